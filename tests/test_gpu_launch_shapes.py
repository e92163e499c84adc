"""Launch shapes of the solve kernels (the autotuner's choices) leave results unchanged.

net14 is one wavefront per particle; 2 or 4 particles per workgroup share the staged assembly
image and use wavefront barriers.  The same seeded particles must come out bit for bit as with one
particle per workgroup, in both EOS and both precisions, and the slot count follows the shape."""
import os

import numpy as np
import pytest

import workloads

pytestmark = pytest.mark.gpu

TABLE = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "sph-exa-nnet_amd", "data",
                     "helm_table.bin")


@pytest.fixture(scope="module")
def ctx():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    import nnet_mi355x
    c = nnet_mi355x.Context(0, helm_table=TABLE)
    yield c
    c.close()


def _run(ctx, cfg, precision, steps=2):
    import torch
    import nnet_mi355x
    st = nnet_mi355x.device_state(torch, cfg["net"], cfg["rho"], cfg["temp"], cfg["Y"], cfg["dt"])
    b = ctx.make_batch(cfg["net"], st, eos=cfg["eos"], mu=cfg["mu"], Z=cfg["Z"], precision=precision)
    for _ in range(steps):
        ctx.compute([b], cfg["hydro_dt"])
    torch.cuda.synchronize()
    return {k: st[k].cpu().numpy() for k in ("Y", "temp", "dt", "status", "substeps", "newton")}


@pytest.mark.parametrize("eos", [0, 1])
@pytest.mark.parametrize("precision", [0, 1])
def test_net14_particles_per_workgroup_bitwise(ctx, eos, precision):
    # C2's net14 distribution (Helmholtz) or C1's (ideal gas), an odd size so the last
    # workgroup of a multi-particle shape has idle groups
    cfg = workloads.config(2 if eos else 1, n=3001, seed=workloads.SEED0 + 77)
    assert cfg["eos"] == eos and cfg["net"] == 14
    out = {}
    try:
        for ppw in (1, 2, 4):
            ctx.set_particles_per_workgroup(14, eos, precision, ppw)
            assert ctx.particles_per_workgroup(14, eos, precision) == ppw
            slots1 = ctx.resident_slots(14, eos, precision)
            out[ppw] = _run(ctx, cfg, precision)
            print(f"net14 eos={eos} precision={precision} ppw={ppw}: slots {slots1}, "
                  f"Newton iterations {int(out[ppw]['newton'].sum())}", flush=True)
    finally:
        ctx.set_particles_per_workgroup(14, eos, precision, 1)
    assert int(out[1]["newton"].sum()) > 0
    for ppw in (2, 4):
        for k, v in out[1].items():
            assert np.array_equal(v.view(np.uint8), out[ppw][k].view(np.uint8)), (ppw, k)


def _multi(ctx, cfgs, precision, fused, hint=False):
    import torch
    import nnet_mi355x
    ctx.set_fused_launch(fused)
    sts = [nnet_mi355x.device_state(torch, c["net"], c["rho"], c["temp"], c["Y"], c["dt"]) for c in cfgs]
    bs = []
    for c, st in zip(cfgs, sts):
        if hint:
            st["cost_hint"] = st["newton"]
        bs.append(ctx.make_batch(c["net"], st, eos=c["eos"], mu=c["mu"], Z=c["Z"], precision=precision))
    for _ in range(2):  # the second step runs with the first step's cost ordering
        ctx.compute(bs, 1e-4)
    torch.cuda.synchronize()
    return [{k: st[k].cpu().numpy() for k in ("Y", "temp", "dt", "status", "substeps", "newton", "fp64_solves")}
            for st in sts]


@pytest.mark.parametrize("nets,precision", [((14, 86, 87), 0), ((14, 86), 1)])
def test_fused_launch_bitwise(ctx, nets, precision):
    """One fused launch for batches of several networks gives the bits of one launch per batch
    (FP64: net14 + net86 + net87, with cost ordering; mixed: net14 + net86)."""
    cid = {14: 2, 86: 3, 87: 4}
    sizes = {14: 700, 86: 300, 87: 90}
    cfgs = [workloads.config(cid[n], n=sizes[n], seed=workloads.SEED0 + 300 + n) for n in nets]
    print(f"fused vs per-batch launches: networks {nets}, precision {precision}", flush=True)
    try:
        a = _multi(ctx, cfgs, precision, True, hint=True)
        b = _multi(ctx, cfgs, precision, False, hint=True)
    finally:
        ctx.set_fused_launch(True)
    for n, x, y in zip(nets, a, b):
        assert int(x["newton"].sum()) > 0
        for k in x:
            assert np.array_equal(x[k].view(np.uint8), y[k].view(np.uint8)), (n, k)


def test_shape_arguments_checked(ctx):
    import nnet_mi355x
    for bad in ((86, 1, 0, 2), (87, 1, 1, 4), (14, 1, 0, 3), (14, 1, 0, 0), (15, 1, 0, 1), (14, 2, 0, 1)):
        with pytest.raises(nnet_mi355x.NnetError):
            ctx.set_particles_per_workgroup(*bad)
    assert ctx.particles_per_workgroup(86, 1, 0) == 1
