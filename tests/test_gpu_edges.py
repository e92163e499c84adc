"""Edge cases of nnet_compute on the device, and the full C3 size through size-independent
properties.

- full size: one hydro step of all 2,097,152 C3 particles (FP64 and mixed); a seeded sample of 192
  particles (with the first and the last) integrated again as a batch of their own gives the same
  bits, so nothing depends on the launch size or a particle's position; the sampled particles
  whose run is short also agree with the oracle (|dX| <= 1e-7 |X| + 1e-14 for FP64, 1e-6 for the
  mixed path; |dT| <= 1e-8 |T|).
- a one-particle batch, a batch whose particles are all below the burning thresholds (every
  status SKIPPED, state untouched), a call without batches, and a zero-length batch beside live
  ones."""
import os

import numpy as np
import pytest

import workloads
from parity import MIX_X_RTOL, X_ATOL, X_RTOL2, T_RTOL2

pytestmark = pytest.mark.gpu

TABLE = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "sph-exa-nnet_amd", "data",
                     "helm_table.bin")
KEYS = ("Y", "temp", "dt", "status", "substeps", "newton")


@pytest.fixture(scope="module")
def ctx():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    import nnet_mi355x
    c = nnet_mi355x.Context(0, helm_table=TABLE)
    yield c
    c.close()


def _run(ctx, cfg, precision=0, idx=None):
    import torch
    import nnet_mi355x as nm
    sel = (lambda a: a[..., idx]) if idx is not None else (lambda a: a)
    st = nm.device_state(torch, cfg["net"], sel(cfg["rho"]), sel(cfg["temp"]), sel(cfg["Y"]), sel(cfg["dt"]))
    b = ctx.make_batch(cfg["net"], st, eos=cfg["eos"], mu=cfg["mu"], Z=cfg["Z"], precision=precision)
    ctx.compute([b], cfg["hydro_dt"])
    torch.cuda.synchronize()
    return {k: st[k].cpu().numpy() for k in KEYS}


@pytest.mark.parametrize("precision", [0, 1])
def test_full_size_c3_sample(ctx, oracle_lib, precision):
    cfg = workloads.config(3)
    n = cfg["rho"].size
    print(f"full size C3 n={n} precision={precision}", flush=True)
    full = _run(ctx, cfg, precision)
    idx = np.unique(np.concatenate([[0, n - 1], np.random.default_rng(11).integers(0, n, 190)]))
    small = _run(ctx, cfg, precision, idx)
    for k in KEYS:
        assert np.array_equal(full[k][..., idx], small[k], equal_nan=True), k
    assert np.bincount(full["status"], minlength=5)[1] > 0.2 * n  # the cool set below rho = 1e5
    # the sampled particles that burn briefly, against the oracle (tier-2 bar on the mixed path too)
    short = idx[(full["status"][idx] == 0) & (full["newton"][idx] < 400)]
    assert len(short) >= 20
    oracle_lib.helm_load(TABLE)
    Y, T, dt = np.ascontiguousarray(cfg["Y"][:, short]), cfg["temp"][short].copy(), cfg["dt"][short].copy()
    st, sub, new = oracle_lib.batch(86, cfg["rho"][short].copy(), T, Y, dt, cfg["hydro_dt"], eos_kind=1, Z=cfg["Z"])
    A = workloads.mass_numbers(86)[:, None]
    dx = np.abs(full["Y"][:, short] * A - Y * A)
    xr = X_RTOL2 if precision == 0 else MIX_X_RTOL
    okx = np.all(dx <= xr * np.abs(Y * A) + X_ATOL, axis=0)
    okT = np.abs(full["temp"][short] - T) <= T_RTOL2 * np.abs(T)
    print(f"full size C3 precision={precision}: {len(short)} short particles, within tier 2: "
          f"{int((okx & okT & (st == 0)).sum())}", flush=True)
    assert np.all(st == 0) and np.mean(okx & okT) >= 0.97


def test_one_particle_and_all_skipped(ctx):
    cfg = workloads.config(3, n=64, seed=21)
    one = _run(ctx, cfg, 0, np.array([5]))
    many = _run(ctx, cfg)
    for k in KEYS:
        assert np.array_equal(one[k], many[k][..., [5]], equal_nan=True), k
    cold = workloads.config(3, n=256, seed=22)
    cold["rho"] = np.full(256, 5e4)  # below rho = 1e5: nothing burns
    out = _run(ctx, cold)
    assert np.all(out["status"] == 1) and np.all(out["newton"] == 0) and np.all(out["substeps"] == 0)
    assert np.array_equal(out["Y"], cold["Y"]) and np.array_equal(out["temp"], cold["temp"])
    assert np.array_equal(out["dt"], cold["dt"])


def test_empty_call_and_zero_length_batch(ctx):
    import torch
    import nnet_mi355x as nm
    ctx.compute([], 1e-4)  # no batches: nothing to do
    cfg = workloads.config(2, n=128, seed=23)
    live = nm.device_state(torch, 14, cfg["rho"], cfg["temp"], cfg["Y"], cfg["dt"])
    empty = nm.device_state(torch, 14, cfg["rho"][:0], cfg["temp"][:0], cfg["Y"][:, :0], cfg["dt"][:0])
    ctx.compute([ctx.make_batch(14, empty, eos=1, Z=cfg["Z"]), ctx.make_batch(14, live, eos=1, Z=cfg["Z"])], 1e-4)
    torch.cuda.synchronize()
    ref = _run(ctx, dict(cfg, hydro_dt=1e-4))
    for k in KEYS:
        assert np.array_equal(live[k].cpu().numpy(), ref[k], equal_nan=True), k
