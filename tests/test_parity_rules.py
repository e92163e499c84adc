"""The parity checker itself (tests/parity.py), on the CPU: the oracle compared with itself
passes every bar; mistakes planted in a copy of the oracle's output are caught by the rule meant
for them (a changed status of a capped particle, a non-finite chaotic state, a capped status where
the oracle finished, a regular abundance moved past tier 1)."""
import os

import numpy as np

import workloads
from parity import assert_parity, compare

TABLE = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "sph-exa-nnet_amd", "data",
                     "helm_table.bin")


def _oracle_c2(oracle_lib, n=512):
    oracle_lib.helm_load(TABLE)
    cfg = workloads.config(2, n=n, seed=4711)
    Y = np.ascontiguousarray(cfg["Y"].copy())
    T, dt = cfg["temp"].copy(), cfg["dt"].copy()
    st, sub, new = oracle_lib.batch(14, cfg["rho"].copy(), T, Y, dt, cfg["hydro_dt"], eos_kind=1, Z=cfg["Z"],
                                    max_iter=2000)
    return Y, T, st, sub, new, oracle_lib.last_max_run.copy()


def _cmp(g, o):
    return compare(workloads.mass_numbers(14), g[0], g[1], g[2], g[3], g[4], o[0], o[1], o[2], o[3], o[4],
                   max_run_orc=o[5])


def test_self_comparison_and_planted_mistakes(oracle_lib):
    o = _oracle_c2(oracle_lib)
    r = _cmp(o, o)
    assert r["frac_northstar"] == 1.0 and r["n_chaotic"] > 0
    assert_parity(r)
    capped = np.nonzero(o[2] == 4)[0]
    regular_ok = np.nonzero((o[2] == 0) & (o[5] <= 50))[0]
    assert len(capped) and len(regular_ok)

    g = [x.copy() for x in o]
    g[2][capped[0]] = 0  # a capped particle reported as finished
    assert _cmp(g, o)["chaos_status_mismatch"] == 1

    g = [x.copy() for x in o]
    g[0][0, capped[0]] = np.nan  # non-finite chaotic state
    assert _cmp(g, o)["chaos_bad"] == 1

    g = [x.copy() for x in o]
    g[2][regular_ok[0]] = 4  # the device capped a particle the oracle finished
    assert _cmp(g, o)["status_mismatch"] == 1

    g = [x.copy() for x in o]
    s = int(np.argmax(g[0][:, regular_ok[0]]))
    g[0][s, regular_ok[0]] *= 1 + 1e-8  # past tier 1 (1e-9), inside tier 2 (1e-7)
    r = _cmp(g, o)
    assert r["tier1"] == r["n"] - 1 and r["tier2_fail"] == 0 and r["frac_northstar"] < 1.0
