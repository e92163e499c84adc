"""How far the reference dynamics amplify last-ulp differences in exp/log/pow.

The oracle is compared with itself after moving a deterministic 3/8 of its rate functions'
exp/log/pow results by one ulp.  This bounds the per-particle parity any implementation whose
transcendentals are not bit-identical to glibc's (the device library) can reach; the GPU parity
criterion in parity.py is set from it.
"""
import numpy as np

import workloads
from parity import compare, summary


def _run(o, cfg):
    Y = np.ascontiguousarray(cfg["Y"].copy())
    T, dt = cfg["temp"].copy(), cfg["dt"].copy()
    st, sub, new = o.batch(cfg["net"], cfg["rho"], T, Y, dt, cfg["hydro_dt"], eos_kind=cfg["eos"], mu=cfg["mu"],
                           Z=cfg["Z"])
    return Y, T, st, sub, new


def test_ulp_sensitivity_c1(oracle_lib):
    cfg = workloads.config(1)
    base = _run(oracle_lib, cfg)
    oracle_lib.L.oracle_set_ulp_perturbation(1)
    try:
        pert = _run(oracle_lib, cfg)
    finally:
        oracle_lib.L.oracle_set_ulp_perturbation(0)
    A = workloads.mass_numbers(14)
    r = compare(A, pert[0], pert[1], pert[2], pert[3], pert[4], base[0], base[1], base[2], base[3], base[4])
    print("oracle vs 1-ulp-perturbed oracle, C1:", summary(r))
    # the perturbation is visible (the dynamics are sensitive) ...
    assert r["tier1"] < r["n"]
    assert r["max_relX"] > 1e-9
    # ... but bounded: no status change, and within the second tier
    assert r["status_mismatch"] == 0
    assert r["tier2_fail"] == 0
    # unperturbed runs are reproducible bit for bit
    again = _run(oracle_lib, cfg)
    assert np.array_equal(again[0], base[0]) and np.array_equal(again[1], base[1])
