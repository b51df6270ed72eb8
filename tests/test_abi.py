"""The C-ABI library builds, loads and exports every function that
include/superplace_hip.h declares; struct layouts agree between the header
(compiled with gcc) and the ctypes mirror; without a GPU every entry point
that needs one reports SP_HIP_ENODEV instead of falling back to the CPU."""
import ctypes as C
import os
import subprocess
import tempfile

import numpy as np
import pytest

from superplace_amd import hip

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_exports_every_declared_symbol():
    L = hip.lib()
    names = hip.declared_functions()
    assert len(names) >= 15
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def test_struct_layouts_match_header():
    fields = {
        "sp_hip_design": hip.SpHipDesign, "sp_hip_gp_config": hip.SpHipGpConfig,
        "sp_hip_iter_metrics": hip.SpHipIterMetrics, "sp_hip_gp_result": hip.SpHipGpResult,
        "sp_hip_eval_scalars": hip.SpHipEvalScalars, "sp_hip_eval_debug": hip.SpHipEvalDebug,
        "sp_hip_gp_info": hip.SpHipGpInfo, "sp_hip_job_info": hip.SpHipJobInfo,
        "sp_mlg_config": hip.SpMlgConfig, "sp_mlg_result": hip.SpMlgResult,
        "sp_abacus_result": hip.SpAbacusResult, "sp_qp_config": hip.SpQpConfig,
        "sp_filler_info": hip.SpFillerInfo,
    }
    src = "#include <stdio.h>\n#include <stddef.h>\n#include \"superplace_hip.h\"\nint main(void){\n"
    for cname in fields:
        src += f'printf("{cname} %zu\\n", sizeof({cname}));\n'
    src += "return 0;}\n"
    with tempfile.TemporaryDirectory() as td:
        c = os.path.join(td, "p.c")
        exe = os.path.join(td, "p")
        open(c, "w").write(src)
        subprocess.check_call(["gcc", "-std=c99", "-I", os.path.join(ROOT, "include"), c, "-o", exe])
        out = subprocess.check_output([exe], text=True)
    sizes = dict(line.split() for line in out.strip().splitlines())
    for cname, cls in fields.items():
        assert int(sizes[cname]) == C.sizeof(cls), cname


def test_no_device_paths_report_enodev():
    if hip.device_count() > 0:
        pytest.skip("a GPU is visible; covered by the gpu tests")
    with pytest.raises(hip.NoDevice):
        hip.HipPoissonBackend(hip.BackendConfig(8, 8, 1.0, 1.0))
    from superplace_amd import design as D
    with pytest.raises(hip.NoDevice):
        hip.GpEngine(D.fixture_eplace(20, 0))
    from superplace_amd import qp
    with pytest.raises(hip.NoDevice):
        qp.qp_place(D.fixture_eplace(20, 0))
    from superplace_amd import serve
    with pytest.raises(hip.NoDevice):
        serve.Server(1 << 30)


def test_footprint_estimate_without_device():
    """The admission estimate is host arithmetic: it grows with the design,
    the bin grid and the precision."""
    from superplace_amd import design as D
    from superplace_amd import serve
    small, big = D.fixture_eplace(200, 2), D.fixture_eplace(2000, 8)
    f = lambda d, **kw: serve.footprint(d, hip.GpConfig(**kw))
    assert 0 < f(small) < f(big)
    assert f(small, gp_bin_dim=256) > f(small, gp_bin_dim=64)
    assert f(small, precision=64) > f(small, precision=32)
    with pytest.raises(hip.InvalidArgument):
        serve.footprint(small, hip.GpConfig(precision=16))


def test_invalid_arguments_rejected_before_device():
    L = hip.lib()
    h = hip._P()
    # null output handle
    assert L.sp_hip_poisson_create(8, 8, 1.0, 1.0, 0, None) == hip.SP_HIP_EINVAL
    assert L.sp_hip_gp_create(None, None, C.byref(h)) == hip.SP_HIP_EINVAL
    assert "null" in L.sp_hip_last_error().decode()
