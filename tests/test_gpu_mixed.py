"""Mixed-precision path (FP32 tile LU on MFMA + FP64 iterative refinement; a particle with an
iteration that does not refine is integrated again, from the start of the step, by the FP64
kernel) against the FP64 oracle: mass fractions within 1e-6 relative (floor 1e-14), temperature
within 1e-8, reported per output together with the specific energy release; the FP32
factorization does carry the accepted iterations; and the re-integrated particles are exactly the
FP64 path's.

Where the inputs themselves are ill-conditioned (net87 with the Z = -1 electron entry in the
Helmholtz zbar sum, where a last-ulp change moves Ye at O(1)), the bar is calibrated like the FP64
parity tests: the mixed path may miss the 1e-6 / 1e-8 bar on no more particles than twice the
1-ulp-perturbed oracle does, plus 1%."""
import math
import os

import numpy as np
import pytest

import workloads
from parity import mixed_errors

pytestmark = pytest.mark.gpu

TABLE = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "sph-exa-nnet_amd", "data",
                     "helm_table.bin")
MIN_WITHIN = 0.97


@pytest.fixture(scope="module")
def ctx():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    import nnet_mi355x
    c = nnet_mi355x.Context(0, helm_table=TABLE)
    yield c
    c.close()


@pytest.mark.parametrize("cid,n", [(2, 2048), (3, 512), (4, 64)])
def test_mixed_vs_fp64_oracle(ctx, oracle_lib, cid, n):
    import torch
    import nnet_mi355x as nm
    cfg = workloads.config(cid, n=n, seed=31 + cid)
    print(f"mixed {cfg['label']} n={n}", flush=True)
    st = nm.device_state(torch, cfg["net"], cfg["rho"], cfg["temp"], cfg["Y"], cfg["dt"])
    b = ctx.make_batch(cfg["net"], st, eos=cfg["eos"], mu=cfg["mu"], Z=cfg["Z"], precision=nm.PRECISION_MIXED)
    ctx.compute([b], cfg["hydro_dt"])
    torch.cuda.synchronize()
    Yg, Tg, sg = st["Y"].cpu().numpy(), st["temp"].cpu().numpy(), st["status"].cpu().numpy()
    newton = st["newton"].cpu().numpy().astype(np.int64)
    fb = st["fp64_solves"].cpu().numpy().astype(np.int64)

    oracle_lib.helm_load(TABLE)
    Y = np.ascontiguousarray(cfg["Y"].copy())
    T, dt = cfg["temp"].copy(), cfg["dt"].copy()
    so, _, _ = oracle_lib.batch(cfg["net"], cfg["rho"].copy(), T, Y, dt, cfg["hydro_dt"], eos_kind=cfg["eos"],
                                mu=cfg["mu"], Z=cfg["Z"])
    r = mixed_errors(workloads.mass_numbers(cfg["net"]), workloads.binding_energies(cfg["net"]), cfg["Y"], Yg, Tg, sg,
                     Y, T, so, max_run_orc=oracle_lib.last_max_run)
    print(cfg["label"], "mixed vs FP64 oracle:", r, f"newton {newton.sum()} fp64 re-solves {fb.sum()}", flush=True)
    redo = fb > 0
    assert np.all(fb[redo] == newton[redo])
    if cid != 4:
        # on the net87 C4 distribution (Newton matrices with condition numbers 1e22-1e27) every
        # particle meets an iteration whose FP32 factors do not refine, so the FP64 pass takes all
        assert fb.sum() < newton.sum(), "no Newton iteration was carried by the FP32 factorization"
    # the particles handed to the FP64 pass equal the FP64 path bit for bit
    s64 = nm.device_state(torch, cfg["net"], cfg["rho"], cfg["temp"], cfg["Y"], cfg["dt"])
    ctx.compute([ctx.make_batch(cfg["net"], s64, eos=cfg["eos"], mu=cfg["mu"], Z=cfg["Z"])], cfg["hydro_dt"])
    torch.cuda.synchronize()
    for k in ("Y", "temp", "dt", "status", "substeps", "newton"):
        a, b64 = st[k].cpu().numpy(), s64[k].cpu().numpy()
        assert np.array_equal(a[..., redo], b64[..., redo], equal_nan=True), k
    # a second step with fp64_hint = the first step's fp64_solves: the hinted particles take the
    # FP64 path from the start, so they equal a second FP64 step bit for bit
    st["fp64_hint"] = st["fp64_solves"].clone()
    hinted = st["fp64_hint"].cpu().numpy() > 0
    for k in ("Y", "temp", "dt"):
        st[k][..., torch.from_numpy(hinted).cuda()] = s64[k][..., torch.from_numpy(hinted).cuda()]
    b2 = ctx.make_batch(cfg["net"], st, eos=cfg["eos"], mu=cfg["mu"], Z=cfg["Z"], precision=nm.PRECISION_MIXED)
    ctx.compute([b2], cfg["hydro_dt"])
    ctx.compute([ctx.make_batch(cfg["net"], s64, eos=cfg["eos"], mu=cfg["mu"], Z=cfg["Z"])], cfg["hydro_dt"])
    torch.cuda.synchronize()
    fb2 = st["fp64_solves"].cpu().numpy()
    assert np.all(fb2[hinted] == st["newton"].cpu().numpy()[hinted])
    for k in ("Y", "temp", "dt", "status", "substeps", "newton"):
        a, b64 = st[k].cpu().numpy(), s64[k].cpu().numpy()
        assert np.array_equal(a[..., hinted], b64[..., hinted], equal_nan=True), k
    assert r["status_mismatch"] <= max(2, 0.01 * n), r
    if r["within"] < MIN_WITHIN * r["n"]:
        oracle_lib.L.oracle_set_ulp_perturbation(1)
        Yp = np.ascontiguousarray(cfg["Y"].copy())
        Tp, dtp = cfg["temp"].copy(), cfg["dt"].copy()
        sp, _, _ = oracle_lib.batch(cfg["net"], cfg["rho"].copy(), Tp, Yp, dtp, cfg["hydro_dt"], eos_kind=cfg["eos"],
                                    mu=cfg["mu"], Z=cfg["Z"])
        oracle_lib.L.oracle_set_ulp_perturbation(0)
        rp = mixed_errors(workloads.mass_numbers(cfg["net"]), workloads.binding_energies(cfg["net"]), cfg["Y"], Yp, Tp,
                          sp, Y, T, so, max_run_orc=oracle_lib.last_max_run)
        print(cfg["label"], "perturbed oracle vs oracle:", rp, flush=True)
        assert r["n"] - r["within"] <= 2 * (rp["n"] - rp["within"]) + math.ceil(0.01 * r["n"]), (r, rp)


def test_hint_ignored_in_multi_batch_call(ctx):
    """fp64_hint is honoured in single-batch calls only: in a call with two mixed batches a hint that
    marks every particle changes nothing (same bits and fp64_solves as without a hint)."""
    import torch
    import nnet_mi355x as nm
    cfgs = [workloads.config(2, n=512, seed=61), workloads.config(3, n=256, seed=62)]
    print("mixed two-batch call with and without fp64_hint", flush=True)
    runs = []
    for hint in (False, True):
        sts, bs = [], []
        for cfg in cfgs:
            st = nm.device_state(torch, cfg["net"], cfg["rho"], cfg["temp"], cfg["Y"], cfg["dt"])
            if hint:
                st["fp64_hint"] = torch.ones_like(st["fp64_solves"])
            bs.append(ctx.make_batch(cfg["net"], st, eos=cfg["eos"], mu=cfg["mu"], Z=cfg["Z"],
                                     precision=nm.PRECISION_MIXED))
            sts.append(st)
        ctx.compute(bs, cfgs[0]["hydro_dt"])
        torch.cuda.synchronize()
        runs.append([{k: s[k].cpu().numpy() for k in ("Y", "temp", "dt", "status", "newton", "fp64_solves")}
                     for s in sts])
    assert all(cfg["hydro_dt"] == cfgs[0]["hydro_dt"] for cfg in cfgs)
    for a, b in zip(*runs):
        for k in a:
            assert np.array_equal(a[k], b[k], equal_nan=True), k
    assert runs[0][1]["fp64_solves"].sum() < runs[0][1]["newton"].sum()  # the hint would have changed this
