"""Writes tests/golden/ref_schedules.npz: outputs of the reference's own
computeGamma (src/gp/wa_wirelength.cpp:18-29), lambdaMultiplier and
hpwlReference (src/gp/lambda_schedule.h:11-25), compiled from /root/reference
by oracle/Makefile into oracle/_ref (run `make -C oracle ref` first).

Inputs cover every branch: tau on both sides of 0.1 and 1.0 (and exactly at
them), bin steps small enough to hit the 1e-4 floor; HPWL changes that land
inside and on both clamps of the lambda multiplier; the fixed and relative
HPWL references, the FILLERONLY exception, the zero-HPWL fallback and the
inputs the reference rejects.

Run from the repo root: python tests/golden/make_schedule_golden.py
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
OUT = os.path.dirname(os.path.abspath(__file__))


def gamma_inputs():
    taus = [-1.0, 0.0, 0.05, 0.0999999, 0.1, 0.1000001, 0.2956, 0.45, 0.5, 0.75, 0.9999999, 1.0,
            1.0000001, 1.5, 4.0, 1e6]
    taus += list(np.linspace(0.0, 1.2, 97))
    steps = [1e-6, 1e-5, 1.0, 3.3125, 22.65625, 20.0, 156.25]
    t, s = np.meshgrid(np.array(taus, np.float32), np.array(steps, np.float32), indexing="ij")
    return t.ravel(), s.ravel()


def lambda_inputs():
    refs = [3.5e5, 1.0, 1234.5, 6.7e6]
    rng = np.random.default_rng(7)
    dh = [0.0, -1.0, 1.0, 3.5e5, -3.5e5, 1e9, -1e9, 0.5, 2.0]
    dh += list(rng.normal(0.0, 1.0, 40))
    out_dh, out_ref = [], []
    for r in refs:
        for v in dh:
            # scale the unit draws so that they straddle both clamps at every ref
            x = v * r if abs(v) < 10 else v
            out_dh.append(x)
            out_ref.append(r)
        # just inside / outside the clamps: 1.1^(1 - d/r) = 1.05 and 0.95
        for mu in (1.05, 0.95):
            edge = (1.0 - np.log(mu) / np.log(1.1)) * r
            for e in (-1e-9, 0.0, 1e-9):
                out_dh.append(edge * (1 + e))
                out_ref.append(r)
    return np.array(out_dh), np.array(out_ref)


def hpwl_ref_inputs():
    rows = []
    for rel in (0, 1):
        for dref in (3.5e5, 1.0, 0.0, -5.0):
            for drel in (0.01, 0.5, 0.0, -0.1):
                for h0 in (0.0, 1.0, 670765.0, 2.5e8):
                    for fo in (0, 1):
                        rows.append((rel, dref, drel, h0, fo))
    return np.array(rows, dtype=np.float64)


def main():
    import oracle
    if oracle.ref_lib() is None:
        raise SystemExit("oracle/_ref/libsp_ref.so missing: run make -C oracle ref")
    t, s = gamma_inputs()
    gam = np.array([oracle.ref_compute_gamma(a, b) for a, b in zip(t, s)], np.float32)
    dh, ref = lambda_inputs()
    mu = np.array([oracle.ref_lambda_multiplier(a, b) for a, b in zip(dh, ref)], np.float32)
    hin = hpwl_ref_inputs()
    hout = []
    for rel, dref, drel, h0, fo in hin:
        v = oracle.ref_hpwl_reference(int(rel), dref, drel, h0, int(fo))
        hout.append(np.nan if v is None else v)
    np.savez(os.path.join(OUT, "ref_schedules.npz"), gamma_tau=t, gamma_step=s, gamma=gam,
             lam_dh=dh, lam_ref=ref, lam_mu=mu, href_in=hin, href_out=np.array(hout))
    print(f"ref_schedules.npz: {len(gam)} gamma, {len(mu)} multiplier, {len(hout)} reference "
          f"cases ({int(np.isnan(hout).sum())} rejected)")


if __name__ == "__main__":
    main()
