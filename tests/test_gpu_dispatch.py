"""Dispatch properties of nnet_compute on the device: a particle's result does not depend on the
order in which the work queue hands particles out (cost hints), nor on sharing a call with
batches of other networks, nor on the range it sits in (firstIndex/lastIndex sub-ranges)."""
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


def _state(cfg):
    import torch
    import nnet_mi355x as nm
    return nm.device_state(torch, cfg["net"], cfg["rho"], cfg["temp"], cfg["Y"], cfg["dt"])


def _host(st):
    return {k: v.cpu().numpy() for k, v in st.items() if k not in ("cost_hint",)}


def _same(a, b):
    for k in a:
        assert np.array_equal(a[k], b[k], equal_nan=True), k


def test_cost_hint_does_not_change_results(ctx):
    import torch
    cfg = workloads.config(3, n=768, seed=91)
    print("dispatch: C3 n=768 with and without cost hints", flush=True)
    s0 = _state(cfg)
    ctx.compute([ctx.make_batch(cfg["net"], s0, eos=cfg["eos"], Z=cfg["Z"])], cfg["hydro_dt"])
    s1 = _state(cfg)
    rng = np.random.default_rng(5)
    s1["cost_hint"] = torch.tensor(rng.integers(0, 20000, cfg["rho"].size), dtype=torch.int32, device="cuda:0")
    ctx.compute([ctx.make_batch(cfg["net"], s1, eos=cfg["eos"], Z=cfg["Z"])], cfg["hydro_dt"])
    # hint = the newton output array itself (previous counts), second step from both states
    s0["cost_hint"] = s0["newton"]
    ctx.compute([ctx.make_batch(cfg["net"], s0, eos=cfg["eos"], Z=cfg["Z"])], cfg["hydro_dt"])
    ctx.compute([ctx.make_batch(cfg["net"], s1, eos=cfg["eos"], Z=cfg["Z"])], cfg["hydro_dt"])
    torch.cuda.synchronize()
    _same(_host(s0), _host(s1))


def test_batches_of_several_networks_in_one_call(ctx):
    import torch
    cfgs = [workloads.config(2, n=512, seed=92), workloads.config(3, n=256, seed=93), workloads.config(4, n=32, seed=94)]
    print("dispatch: net14 + net86 + net87 in one call vs separate calls", flush=True)
    joint = [_state(c) for c in cfgs]
    ctx.compute([ctx.make_batch(c["net"], s, eos=c["eos"], Z=c["Z"]) for c, s in zip(cfgs, joint)], 1e-4)
    for c, sj in zip(cfgs, joint):
        ss = _state(c)
        ctx.compute([ctx.make_batch(c["net"], ss, eos=c["eos"], Z=c["Z"])], 1e-4)
        torch.cuda.synchronize()
        _same(_host(sj), _host(ss))


def test_sub_range_equals_full_range(ctx):
    import torch
    import nnet_mi355x as nm
    cfg = workloads.config(3, n=400, seed=95)
    print("dispatch: particles [100, 300) of a batch vs the whole batch", flush=True)
    full = _state(cfg)
    ctx.compute([ctx.make_batch(cfg["net"], full, eos=cfg["eos"], Z=cfg["Z"])], cfg["hydro_dt"])
    part = _state(cfg)
    view = {k: (v[:, 100:300] if k == "Y" else v[100:300]) for k, v in part.items()}
    ctx.compute([ctx.make_batch(cfg["net"], view, eos=cfg["eos"], Z=cfg["Z"])], cfg["hydro_dt"])
    torch.cuda.synchronize()
    f, p = _host(full), _host(part)
    for k in f:
        a, b = (f[k][:, 100:300], p[k][:, 100:300]) if k == "Y" else (f[k][100:300], p[k][100:300])
        assert np.array_equal(a, b, equal_nan=True), k
