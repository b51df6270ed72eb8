"""Many jobs on one GPU (sp_hip_server_*, superplace_amd.serve): every job
served in a mix gives bit for bit the result of running it alone
(sp_hip_gp_run); admission respects arrival times, the job-slot limit and
the device budget; the footprint estimate bounds what an instance
allocates."""
import numpy as np
import pytest

from superplace_amd import design as D
from superplace_amd import hip, serve

pytestmark = pytest.mark.gpu

RESULT_KEYS = ["iterations", "converged", "stagnated", "rollback_iter", "evaluations", "hpwl",
               "overflow", "initial_hpwl", "initial_overflow", "initial_lambda"]


def mix():
    """Uneven synthetic designs (1.5k-12k cells) with their configurations."""
    out = []
    for k, cells in enumerate([1500, 12000, 3000, 8000, 2000, 5000]):
        d = D.synth_design(cells, 4 + k, int(cells * 1.05), 40.0 * np.sqrt(cells), 0.6,
                           seed=100 + k, name=f"mix{k}")
        cfg = hip.GpConfig(gp_bin_dim=0, gp_max_iter=300 + 40 * k,
                           precision=64 if k == 2 else 32)
        out.append((d, cfg))
    return out


def alone(d, cfg):
    eng = hip.GpEngine(d, cfg)
    res, _ = eng.run()
    x, y = eng.positions()
    return res, x, y, eng.device_bytes()


def bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.uint32)


def test_footprint_bounds_instance():
    for d, cfg in mix()[:3]:
        est = serve.footprint(d, cfg)
        _, _, _, used = alone(d, cfg)
        print(d.name, cfg.precision, "estimate", est, "measured", used, flush=True)
        assert used <= est <= 1.5 * used + (1 << 20)


def test_served_equals_alone():
    jobs = mix()
    ref = [alone(d, cfg) for d, cfg in jobs]
    srv = serve.Server(hbm_budget=8 << 30, max_active=4)
    ids = [srv.submit(d, cfg, arrival_s=0.0) for d, cfg in jobs]
    srv.drain(timeout_s=300)
    overlap = 0
    infos = [srv.job(i) for i in ids]
    for (d, _), (res, x, y, used), info in zip(jobs, ref, infos):
        print(d.name, info["state"], info["result"]["iterations"], flush=True)
        assert info["state"] == serve.DONE and info["status"] == 0
        for k in RESULT_KEYS:
            assert info["result"][k] == res[k], (d.name, k)
        assert np.array_equal(bits(info["x"]), bits(x)) and np.array_equal(bits(info["y"]), bits(y))
        assert info["device_bytes"] == used
        assert 0 <= info["admit_s"] <= info["finish_s"]
    for a in infos:
        for b in infos:
            if a is not b and a["admit_s"] < b["admit_s"] < a["finish_s"]:
                overlap += 1
    assert overlap > 0   # jobs did run together
    srv.close()


def test_admission_budget_arrivals_and_failures():
    jobs = mix()[:4]
    est = [serve.footprint(d, cfg) for d, cfg in jobs]
    budget = max(est) + min(est)   # at most two of these fit at once
    srv = serve.Server(hbm_budget=budget, max_active=8)
    arrivals = [0.0, 0.0, 0.05, 0.2]
    ids = [srv.submit(d, cfg, arrival_s=t) for (d, cfg), t in zip(jobs, arrivals)]
    # a job larger than the whole budget fails without running
    big = D.synth_design(30000, 8, 31000, 7000.0, 0.6, seed=5, name="big")
    big_id = srv.submit(big, hip.GpConfig(), arrival_s=0.0)
    srv.drain(timeout_s=300)
    infos = [srv.job(i, positions=False) for i in ids]
    for info, t in zip(infos, arrivals):
        assert info["state"] == serve.DONE
        assert info["admit_s"] >= t
    events = sorted({i["admit_s"] for i in infos})
    for t in events:   # device bytes of the jobs resident at each admission
        live = sum(i["device_bytes"] for i in infos if i["admit_s"] <= t < i["finish_s"])
        assert live <= budget
    b = srv.job(big_id, positions=False)
    assert b["state"] == serve.FAILED and b["status"] == hip.SP_HIP_EOOM
    with pytest.raises(hip.InvalidArgument):
        srv.submit(jobs[0][0], hip.GpConfig(loop_mode=0))
    srv.release(ids[0])
    srv.close()


def test_serving_stream_equals_alone():
    """The serving workload: 12 jobs of 10k-1M cells (log-uniform, the smallest
    10k and the largest 1M cells with 2.1M pins), bins 256-1024 by the
    automatic rule, Poisson arrivals at 20 jobs/s, served together by the
    scheduler (FP32 jobs batched across iterations).  Every job converges and
    its result and positions equal the job run alone; p50/p99 latency (finish
    - arrival) is reported."""
    jobs, arrivals = D.serving_stream(12, 20.0, extremes=True)
    cells = [d.num_movable for d, _ in jobs]
    assert min(cells) == 10000 and max(cells) == 1000000
    assert max(d.num_pins for d, _ in jobs) <= 2_200_000
    assert {dim for _, dim in jobs} >= {256, 1024}
    cfgs = [hip.GpConfig(gp_bin_dim=dim) for _, dim in jobs]
    srv = serve.Server(hbm_budget=200 << 30, max_active=16)
    for (d, _), c, t in zip(jobs, cfgs, arrivals):
        srv.submit(d, c, arrival_s=float(t))
    srv.drain(timeout_s=600)
    infos = [srv.job(i) for i in range(len(jobs))]
    srv.close()
    lat = np.array([i["finish_s"] - i["arrival_s"] for i in infos])
    print(f"latency p50 {np.percentile(lat, 50):.3f} s p99 {np.percentile(lat, 99):.3f} s",
          flush=True)
    overlap = sum(1 for a in infos for b in infos
                  if a is not b and a["admit_s"] < b["admit_s"] < a["finish_s"])
    assert overlap > 0
    for (d, _), c, info in zip(jobs, cfgs, infos):
        assert info["state"] == serve.DONE and info["status"] == 0, d.name
        res, x, y, _ = alone(d, c)
        print(d.name, d.num_movable, c.gp_bin_dim, res["iterations"], res["overflow"], flush=True)
        assert res["converged"] and res["overflow"] <= 0.10
        for k in RESULT_KEYS:
            assert info["result"][k] == res[k], (d.name, k, info["result"][k], res[k])
        assert np.array_equal(bits(info["x"]), bits(x)) and np.array_equal(bits(info["y"]), bits(y))
