"""Many placement jobs on one GPU through the library's scheduler
(sp_hip_server_*): admission by device footprint, jobs resident on the
device and advanced together, per-job latency.  No counterpart in the
reference, which places one design per process."""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import hip

QUEUED, RUNNING, DONE, FAILED = 0, 1, 2, 3


def footprint(design, cfg: hip.GpConfig | None = None) -> int:
    """Device bytes one instance of the design allocates (an upper bound)."""
    cfg = cfg or hip.GpConfig()
    da = hip._DesignArrays(design)
    cs = cfg.struct()
    out = C.c_int64(0)
    hip.check(hip.lib().sp_hip_gp_footprint(C.byref(da.s), C.byref(cs), C.byref(out)))
    return int(out.value)


class Server:
    def __init__(self, hbm_budget: int, max_active: int = 8):
        h = hip._P()
        hip.check(hip.lib().sp_hip_server_create(int(hbm_budget), int(max_active), C.byref(h)))
        self._h = h
        self._sizes = []

    def submit(self, design, cfg: hip.GpConfig | None = None, arrival_s: float = 0.0) -> int:
        cfg = cfg or hip.GpConfig()
        da = hip._DesignArrays(design)
        cs = cfg.struct()
        job = C.c_int32(-1)
        hip.check(hip.lib().sp_hip_server_submit(self._h, C.byref(da.s), C.byref(cs),
                                                 float(arrival_s), C.byref(job)))
        self._sizes.append(design.total_nodes)
        return job.value

    def pump(self) -> dict:
        now, act, q, fin = C.c_double(), C.c_int32(), C.c_int32(), C.c_int32()
        hip.check(hip.lib().sp_hip_server_pump(self._h, C.byref(now), C.byref(act), C.byref(q),
                                               C.byref(fin)))
        return dict(now=now.value, active=act.value, queued=q.value, finished=fin.value)

    def drain(self, timeout_s: float = 600.0) -> None:
        hip.check(hip.lib().sp_hip_server_drain(self._h, float(timeout_s)))

    def now(self) -> float:
        t = C.c_double()
        hip.check(hip.lib().sp_hip_server_now(self._h, C.byref(t)))
        return t.value

    def job(self, job: int, positions: bool = True) -> dict:
        info, res = hip.SpHipJobInfo(), hip.SpHipGpResult()
        n = self._sizes[job]
        x = np.zeros(n, np.float32) if positions else None
        y = np.zeros(n, np.float32) if positions else None
        hip.check(hip.lib().sp_hip_server_job(self._h, int(job), C.byref(info), C.byref(res),
                                              None if x is None else x.ctypes.data,
                                              None if y is None else y.ctypes.data))
        out = {k: getattr(info, k) for k, _ in hip.SpHipJobInfo._fields_}
        out["result"] = {k: getattr(res, k) for k, _ in hip.SpHipGpResult._fields_}
        if positions and info.state == DONE:
            out["x"], out["y"] = x, y
        return out

    def release(self, job: int) -> None:
        hip.check(hip.lib().sp_hip_server_release(self._h, int(job)))

    def close(self) -> None:
        if getattr(self, "_h", None):
            hip.lib().sp_hip_server_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()
