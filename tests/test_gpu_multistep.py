"""Multi-step parity: K consecutive hydro steps with the particle state kept in device memory
(the device arrays are the next step's input, as in an SPH-EXA run), against the oracle stepping
the same particles the same K steps on the host (nuclear_net.hpp:931-959 per step,
parallel_nuclear_net_impl.cpp:63-103 per batch).  The bar is the calibrated one of
tests/parity.py, evaluated after the last step, with the per-step north-star fractions printed.

Sizes: C2 (net14) 10 steps, C3 (net86) and C4 (net87) 20 steps, 512 particles each.  C4 runs with
a Newton cap of 2,000 per particle and step on both sides (the same cap is a parameter of the
compared semantics; it bounds the oracle's cost on the C4 distribution, where several percent of
the particles cycle to the cap every step)."""
import os

import numpy as np
import pytest

import workloads
from parity import assert_parity, compare, summary

pytestmark = pytest.mark.gpu

TABLE = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "sph-exa-nnet_amd", "data",
                     "helm_table.bin")
NTHREADS = min(16, os.cpu_count() or 1)


@pytest.fixture(scope="module")
def ctx():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    import nnet_mi355x
    c = nnet_mi355x.Context(0, helm_table=TABLE)
    yield c
    c.close()


def _oracle_steps(oracle_lib, cfg, steps, max_iter, perturb):
    oracle_lib.helm_load(TABLE)
    oracle_lib.L.oracle_set_ulp_perturbation(1 if perturb else 0)
    Y = np.ascontiguousarray(cfg["Y"].copy())
    T, dt, rho = cfg["temp"].copy(), cfg["dt"].copy(), cfg["rho"].copy()
    out = []
    for _ in range(steps):
        st, sub, new = oracle_lib.batch(cfg["net"], rho, T, Y, dt, cfg["hydro_dt"], eos_kind=cfg["eos"], mu=cfg["mu"],
                                        Z=cfg["Z"], nthreads=NTHREADS, max_iter=max_iter)
        out.append((Y.copy(), T.copy(), st, sub, new, oracle_lib.last_max_run.copy()))
    oracle_lib.L.oracle_set_ulp_perturbation(0)
    return out


@pytest.mark.parametrize("cid,n,steps,max_iter", [(2, 512, 10, 10000), (3, 512, 20, 10000), (4, 512, 20, 2000)])
def test_device_resident_steps(ctx, oracle_lib, cid, n, steps, max_iter):
    import torch
    import nnet_mi355x as nm
    cfg = workloads.config(cid, n=n, seed=2000 + cid)
    print(f"multistep {cfg['label']} n={n} steps={steps} cap={max_iter}", flush=True)
    ctx.set_max_iterations(max_iter)
    try:
        st = nm.device_state(torch, cfg["net"], cfg["rho"], cfg["temp"], cfg["Y"], cfg["dt"])
        b = ctx.make_batch(cfg["net"], st, eos=cfg["eos"], mu=cfg["mu"], Z=cfg["Z"])
        gpu = []
        for _ in range(steps):
            ctx.compute([b], cfg["hydro_dt"])
            gpu.append(tuple(st[k].cpu().numpy().copy() for k in ("Y", "temp", "status", "substeps", "newton")))
    finally:
        ctx.set_max_iterations(10000)
    orc = _oracle_steps(oracle_lib, cfg, steps, max_iter, False)
    A = workloads.mass_numbers(cfg["net"])
    for k in range(steps):
        g, o = gpu[k], orc[k]
        r = compare(A, g[0], g[1], g[2], g[3].astype(np.uint32), g[4].astype(np.uint32), o[0], o[1], o[2], o[3], o[4],
                    max_run_orc=o[5])
        print(f"  step {k}: north-star fraction {r['frac_northstar']:.4f} chaotic {r['n_chaotic']} "
              f"status mismatch {r['status_mismatch']} max rel X {r['max_relX']:.2e} T {r['max_relT']:.2e}", flush=True)
    pert = _oracle_steps(oracle_lib, cfg, steps, max_iter, True)
    g, o, p = gpu[-1], orc[-1], pert[-1]
    r = compare(A, g[0], g[1], g[2], g[3].astype(np.uint32), g[4].astype(np.uint32), o[0], o[1], o[2], o[3], o[4],
                max_run_orc=o[5])
    rp = compare(A, p[0], p[1], p[2], p[3], p[4], o[0], o[1], o[2], o[3], o[4], max_run_orc=o[5])
    print(cfg["label"], f"after {steps} steps, gpu vs oracle:", summary(r), flush=True)
    print(cfg["label"], f"after {steps} steps, perturbed oracle vs oracle:", summary(rp), flush=True)
    assert_parity(r, rp)
