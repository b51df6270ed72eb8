"""metrics.csv and summary.json writers through the library (sp_metrics_*),
with the reference's MetricsSink interface and formats
(src/util/metrics.h:40-90, src/util/metrics.cpp:19-140)."""
from __future__ import annotations

import ctypes as C

from . import hip


def _rows(rows):
    """A ctypes array of sp_hip_iter_metrics from dicts (the keys of
    GpEngine.run's rows; missing fields are 0, lambda_mu 1)."""
    arr = (hip.SpHipIterMetrics * max(len(rows), 1))()
    for i, r in enumerate(rows):
        arr[i].lambda_mu = 1.0
        for k, v in r.items():
            k = "lambda_" if k == "lambda" else k
            if any(k == f for f, _ in hip.SpHipIterMetrics._fields_):
                setattr(arr[i], k, v)
    return arr


class MetricsSink:
    def __init__(self, out_dir: str | None = None):
        h = hip._P()
        hip.check(hip.lib().sp_metrics_create(None if out_dir is None else str(out_dir).encode(),
                                              C.byref(h)))
        self._h = h

    def begin_stage(self, stage: str) -> None:
        hip.check(hip.lib().sp_metrics_begin_stage(self._h, stage.encode()))

    def push(self, rows, max_displacement=None) -> None:
        """One row (dict) or a list of rows."""
        rows = [rows] if isinstance(rows, dict) else list(rows)
        md = None
        if max_displacement is not None:
            md = (C.c_double * len(rows))(*max_displacement)
        hip.check(hip.lib().sp_metrics_push(self._h, C.cast(_rows(rows), hip._P), len(rows),
                                            None if md is None else C.cast(md, hip._P)))

    def end_stage(self, elapsed_ms: float) -> None:
        hip.check(hip.lib().sp_metrics_end_stage(self._h, float(elapsed_ms)))

    def record_duration(self, label: str, ms: float) -> None:
        hip.check(hip.lib().sp_metrics_record_duration(self._h, label.encode(), float(ms)))

    def set_summary(self, key: str, value) -> None:
        if isinstance(value, str):
            hip.check(hip.lib().sp_metrics_set_string(self._h, key.encode(), value.encode()))
        else:
            hip.check(hip.lib().sp_metrics_set_number(self._h, key.encode(), float(value)))

    def gp_result(self, stage: str, result: dict, hpwl_reference: float,
                  restarts: int = -1) -> None:
        """The summary keys EPlace::run writes for a stage (eplace.cpp:266-270, :509-513)
        from a GpEngine.run result."""
        r = hip.SpHipGpResult(**{k: result[k] for k, _ in hip.SpHipGpResult._fields_})
        hip.check(hip.lib().sp_metrics_gp_result(self._h, stage.encode(), C.byref(r),
                                                 float(hpwl_reference), int(restarts)))

    def write_summary(self, path: str) -> None:
        hip.check(hip.lib().sp_metrics_write_summary(self._h, str(path).encode()))

    def close(self) -> None:
        if getattr(self, "_h", None):
            hip.lib().sp_metrics_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
