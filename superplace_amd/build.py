"""Builds libsuperplace_hip.so (gfx950) in-tree with hipcc.

Every HIP source is compiled with packed FP32 instructions disabled
(-Xclang -target-feature -Xclang -packed-fp32-ops) and linked into one
shared library next to this file.
"""
from __future__ import annotations

import os
import pathlib
import shutil
import subprocess
import sys

HERE = pathlib.Path(__file__).resolve().parent
CSRC = HERE / "csrc"
INCLUDE = HERE.parent / "include"
LIB = HERE / "libsuperplace_hip.so"
OBJ_DIR = HERE / "_build"
SOURCES = ["poisson.hip", "gp_kernels.hip", "engine.hip", "qp.hip", "bookshelf.cpp", "setup.cpp",
           "metrics.cpp", "legalize.cpp", "macro_lg.cpp",
           "serve.cpp", "noalloc.hip"]
ARCH = os.environ.get("SP_HIP_ARCH", "gfx950")


def hipcc() -> str:
    for cand in (shutil.which("hipcc"), "/opt/rocm/bin/hipcc"):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found")


def flags() -> list[str]:
    return ["-O3", "-std=c++17", f"--offload-arch={ARCH}", "-fPIC", "-Wall",
            "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
            f"-I{INCLUDE}", f"-I{CSRC}"]


def build(verbose: bool = False, jobs: int = 4) -> pathlib.Path:
    """Compiles every source (always) and links the library."""
    OBJ_DIR.mkdir(exist_ok=True)
    cc = hipcc()
    srcs = [s for s in SOURCES if (CSRC / s).exists()]
    procs = []
    objs = []
    for s in srcs:
        obj = OBJ_DIR / (s.rsplit(".", 1)[0] + ".o")
        objs.append(obj)
        cmd = [cc, *flags(), "-c", str(CSRC / s), "-o", str(obj)]
        if verbose:
            print(" ".join(cmd), flush=True)
        procs.append((s, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT,
                                          text=True)))
        if len(procs) >= jobs:
            _drain(procs)
    _drain(procs)
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [cc, "-shared", f"--offload-arch={ARCH}", "-o", str(tmp), *map(str, objs)]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed:\n" + r.stdout + r.stderr)
    os.replace(tmp, LIB)
    return LIB


def _drain(procs):
    while procs:
        name, p = procs.pop(0)
        out, _ = p.communicate()
        noise = [l for l in out.splitlines() if "packed-fp32-ops" not in l
                 and "not a recognized feature" not in l]
        if p.returncode != 0:
            raise RuntimeError(f"hipcc failed on {name}:\n" + "\n".join(noise))
        if noise and any("warning" in l or "error" in l for l in noise):
            print(f"[{name}]\n" + "\n".join(noise), file=sys.stderr)


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
