"""use_drhodt on the GPU.  NNET_DRHODT_CPU (the reference CPU path, parallel_nuclear_net_impl.cpp:
89-92): rho_m1 of every integrated particle becomes (rho - rho_m1)/previous_dt where it was nonzero,
and the integration is exactly the one without drho/dt.  NNET_DRHODT_INTEGRATE (the reference GPU
branch, parallel_nuclear_net.cu:79-81): drho/dt drives rho(t) (nuclear_net.hpp:859) and rho_m1 is
unchanged; the results meet the oracle parity bar (the oracle's single-particle path with drho/dt
is pinned to the reference's own prepareSystemSubstep in tests/test_oracle_reference.py)."""
import os

import numpy as np
import pytest

import workloads
from parity import assert_parity, compare, summary

pytestmark = pytest.mark.gpu

TABLE = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "sph-exa-nnet_amd", "data",
                     "helm_table.bin")
PREV_DT = 2e-4


@pytest.fixture(scope="module")
def ctx():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    import nnet_mi355x
    c = nnet_mi355x.Context(0, helm_table=TABLE)
    yield c
    c.close()


def _cfg(cid, n, seed):
    cfg = workloads.config(cid, n=n, seed=seed)
    rng = np.random.default_rng(seed)
    rm1 = cfg["rho"] * (1 + rng.uniform(-0.3, 0.3, n))
    rm1[rng.random(n) < 0.1] = 0.0  # no previous density: drho/dt = 0 (no division)
    cfg["rho_m1"] = rm1
    return cfg


def _gpu(ctx, cfg, mode, precision=0):
    import torch
    import nnet_mi355x as nm
    st = nm.device_state(torch, cfg["net"], cfg["rho"], cfg["temp"], cfg["Y"], cfg["dt"])
    st["rho_m1"].copy_(torch.from_numpy(cfg["rho_m1"]))
    b = ctx.make_batch(cfg["net"], st, eos=cfg["eos"], mu=cfg["mu"], Z=cfg["Z"], precision=precision)
    ctx.compute([b], cfg["hydro_dt"], previous_dt=PREV_DT, use_drhodt=mode)
    torch.cuda.synchronize()
    out = {k: st[k].cpu().numpy() for k in ("Y", "temp", "dt", "status", "substeps", "newton", "rho_m1")}
    return out


def _oracle(oracle_lib, cfg, mode, perturb=False):
    oracle_lib.helm_load(TABLE)
    oracle_lib.L.oracle_set_ulp_perturbation(1 if perturb else 0)
    Y, T, dt, rm1 = np.ascontiguousarray(cfg["Y"].copy()), cfg["temp"].copy(), cfg["dt"].copy(), cfg["rho_m1"].copy()
    st, sub, new = oracle_lib.batch(cfg["net"], cfg["rho"].copy(), T, Y, dt, cfg["hydro_dt"], previous_dt=PREV_DT,
                                    eos_kind=cfg["eos"], mu=cfg["mu"], Z=cfg["Z"], rho_m1=rm1, use_drhodt=mode)
    oracle_lib.L.oracle_set_ulp_perturbation(0)
    return Y, T, dt, st, sub, new, oracle_lib.last_max_run.copy(), rm1


def test_drhodt_cpu_semantics(ctx, oracle_lib):
    import nnet_mi355x as nm
    cfg = _cfg(2, 2048, 61)
    print("drhodt CPU semantics, net14 C2 n=2048", flush=True)
    a = _gpu(ctx, cfg, nm.DRHODT_CPU)
    b = _gpu(ctx, cfg, nm.DRHODT_OFF)
    for k in ("Y", "temp", "dt", "status", "substeps", "newton"):
        assert np.array_equal(a[k], b[k], equal_nan=True), k
    o = _oracle(oracle_lib, cfg, 1)
    assert np.array_equal(a["rho_m1"], o[7])  # one IEEE division per particle, skipped ones untouched
    assert np.array_equal(b["rho_m1"], cfg["rho_m1"])


@pytest.mark.parametrize("cid,n,precision", [(2, 2048, 0), (3, 384, 0), (3, 384, 1)])
def test_drhodt_integrate(ctx, oracle_lib, cid, n, precision):
    import nnet_mi355x as nm
    cfg = _cfg(cid, n, 70 + cid)
    print(f"drhodt integrate, {cfg['label']} n={n} precision={precision}", flush=True)
    g = _gpu(ctx, cfg, nm.DRHODT_INTEGRATE, precision)
    assert np.array_equal(g["rho_m1"], cfg["rho_m1"])
    off = _gpu(ctx, cfg, nm.DRHODT_OFF, precision)
    moving = (cfg["rho_m1"] != 0) & (g["status"] != 1)
    assert not np.array_equal(g["temp"][moving], off["temp"][moving]), "drho/dt did not enter the integration"
    o = _oracle(oracle_lib, cfg, 2)
    A = workloads.mass_numbers(cfg["net"])
    if precision == 0:
        r = compare(A, g["Y"], g["temp"], g["status"], g["substeps"], g["newton"], o[0], o[1], o[3], o[4], o[5],
                    max_run_orc=o[6])
        p = _oracle(oracle_lib, cfg, 2, perturb=True)
        rp = compare(A, p[0], p[1], p[3], p[4], p[5], o[0], o[1], o[3], o[4], o[5], max_run_orc=o[6])
        print(cfg["label"], "drhodt gpu vs oracle:", summary(r), "perturbed:", summary(rp), flush=True)
        assert_parity(r, rp)
    else:
        from parity import mixed_errors
        r = mixed_errors(A, workloads.binding_energies(cfg["net"]), cfg["Y"], g["Y"], g["temp"], g["status"], o[0],
                         o[1], o[3], max_run_orc=o[6])
        print(cfg["label"], "drhodt mixed vs oracle:", r, flush=True)
        assert r["status_mismatch"] <= max(2, 0.01 * n) and r["within"] >= 0.97 * r["n"], r
