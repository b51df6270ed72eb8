"""Many instances advanced together in batched launches
(sp_hip_gp_run_batch): each kernel of an iteration is launched once for all
the instances (the Poisson passes once per grid size), and every
instance's result and positions equal its own run's bit for bit.

The batched workload of the benchmark (64 seeded designs of 5k-80k cells,
bins by the automatic rule within 128-512) is checked in full: every job
through the batch and through the scheduler (sp_hip_server_*) against the
same job run alone."""
import numpy as np
import pytest

from superplace_amd import design as D
from superplace_amd import hip, serve

pytestmark = pytest.mark.gpu

KEYS = ("iterations", "evaluations", "converged", "stagnated", "hpwl", "overflow")


def alone(d, cfg):
    eng = hip.GpEngine(d, cfg)
    res, _ = eng.run(max_rows=1)
    x, y = eng.positions()
    eng.close()
    return res, x, y


def test_batch_equals_alone_mixed_sizes():
    """Designs of 600 to 12k cells on 64^2 to 256^2 grids, one of them
    already below the target overflow (no loop at all)."""
    designs = [D.with_fillers(D.fixture_eplace(600, 2)),
               D.with_fillers(D.fixture_eplace(3000, 12)),
               D.config_small(),
               D.synth_design(12000, 30, 12600, 25.25 * np.sqrt(12000), 0.57, seed=5, name="s12k")]
    converged = D.fixture_eplace(200, 0)
    converged.node_x[:200] = 100 + (np.arange(200) % 40) * 25.0
    converged.node_y[:200] = 100 + (np.arange(200) // 40) * 100.0
    designs.append(converged)
    cfgs = [hip.GpConfig(), hip.GpConfig(), hip.GpConfig(gp_bin_dim=128),
            hip.GpConfig(gp_bin_dim=256), hip.GpConfig()]
    res, xs, ys = hip.run_batch(designs, cfgs)
    for i, (d, c) in enumerate(zip(designs, cfgs)):
        r, x, y = alone(d, c)
        print(i, r["iterations"], res[i]["iterations"], r["hpwl"], res[i]["hpwl"], flush=True)
        for k in KEYS:
            assert res[i][k] == r[k], (i, k, res[i][k], r[k])
        assert np.array_equal(xs[i], x) and np.array_equal(ys[i], y), i
    assert res[4]["iterations"] == 0 and res[4]["converged"]


def test_batch_rejects_fp64():
    d = D.with_fillers(D.fixture_eplace(300, 1))
    with pytest.raises(hip.InvalidArgument):
        hip.run_batch([d], [hip.GpConfig(precision=64)])


def test_config4_batch_and_scheduler_equal_alone():
    """The benchmark's batch: 64 designs of 5k-80k cells (log-uniform, seeded),
    bins 128-512 by the automatic rule.  Every job's result and positions
    through the batched launches and through the scheduler equal the job
    run alone."""
    jobs = D.job_designs(64, 5000, 80000, 0, 128, 512)
    cells = [d.num_movable for d, _ in jobs]
    assert len(jobs) == 64 and min(cells) >= 5000 and max(cells) <= 80000
    cfgs = [hip.GpConfig(gp_bin_dim=dim) for _, dim in jobs]
    designs = [d for d, _ in jobs]
    print("dims", sorted(set(c.gp_bin_dim for c in cfgs)), "cells", min(cells), max(cells),
          flush=True)
    bres, bx, by = hip.run_batch(designs, cfgs)
    srv = serve.Server(hbm_budget=200 << 30, max_active=16)
    for d, c in zip(designs, cfgs):
        srv.submit(d, c, 0.0)
    srv.drain(timeout_s=600)
    sinfo = [srv.job(i, positions=True) for i in range(len(jobs))]
    srv.close()
    for i, (d, c) in enumerate(zip(designs, cfgs)):
        r, x, y = alone(d, c)
        assert r["converged"], i
        for k in KEYS:
            assert bres[i][k] == r[k], ("batch", i, k, bres[i][k], r[k])
            assert sinfo[i]["result"][k] == r[k], ("server", i, k)
        assert np.array_equal(bx[i], x) and np.array_equal(by[i], y), ("batch", i)
        assert np.array_equal(sinfo[i]["x"], x) and np.array_equal(sinfo[i]["y"], y), ("server", i)
