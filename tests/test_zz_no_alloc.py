"""nnet_compute_noalloc: the same computation as nnet_compute with the library's fixed device
scratch instead of its growable buffers -- no allocation in the call.  Two calls in a row, the
second with a smaller input (the fixed scratch keeps the first call's contents, so whatever must
start at zero has to be cleared by the call itself), each against nnet_compute on a copy of the
same input: bit-identical outputs.  A call that needs more than the scratch holds is refused."""
import os

import numpy as np
import pytest

import workloads

pytestmark = pytest.mark.gpu

TABLE = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "sph-exa-nnet_amd", "data",
                     "helm_table.bin")
FIELDS = ("Y", "temp", "dt", "status", "substeps", "newton", "fp64_solves")


def _states(torch, nm, cfgs, hint):
    sts = []
    for c in cfgs:
        st = nm.device_state(torch, c["net"], c["rho"], c["temp"], c["Y"], c["dt"])
        if hint:
            st["cost_hint"] = torch.randint(0, 5000, st["newton"].shape, dtype=torch.int32, device="cuda:0")
        sts.append(st)
    return sts


@pytest.mark.parametrize("precision", [0, 1])
def test_noalloc_equals_compute(precision):
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    import nnet_mi355x as nm
    L = nm.lib()
    ctx = nm.Context(0, helm_table=TABLE)
    # first call: net14 + net86 batches (cost-ordered); second: one smaller net86 batch
    calls = [[workloads.config(2, n=1536, seed=51), workloads.config(3, n=1024, seed=52)],
             [workloads.config(3, n=300, seed=53)]]
    for k, cfgs in enumerate(calls):
        print(f"noalloc call {k}: precision {precision}, batches {[c['rho'].size for c in cfgs]}", flush=True)
        out = []
        for use_noalloc in (False, True):
            sts = _states(torch, nm, cfgs, hint=(k == 0))
            if precision == 1:
                for st in sts:
                    st["fp64_hint"] = None
            bs = [ctx.make_batch(c["net"], st, eos=c["eos"], mu=c["mu"], Z=c["Z"], precision=precision)
                  for c, st in zip(cfgs, sts)]
            arr = (nm.Batch * len(bs))(*bs)
            if use_noalloc:
                rc = L.nnet_compute_noalloc(ctx.h, arr, len(bs), cfgs[0]["hydro_dt"], cfgs[0]["hydro_dt"], 0, None)
                ctx.check(rc, "nnet_compute_noalloc")
            else:
                ctx.compute(bs, cfgs[0]["hydro_dt"], sync=False)
            torch.cuda.synchronize()
            out.append([{f: st[f].cpu().numpy() for f in FIELDS} for st in sts])
        for a, b in zip(*out):
            for f in FIELDS:
                assert np.array_equal(a[f], b[f], equal_nan=True), (k, f)
    # more batches than the fixed counters cover are refused; the batches are empty, so the call
    # would enqueue nothing even if it were accepted
    small = workloads.config(2, n=8, seed=54)
    st = _states(torch, nm, [small], hint=False)[0]
    b = ctx.make_batch(14, st, eos=small["eos"], Z=small["Z"])
    b.n = 0
    arr = (nm.Batch * 120)(*([b] * 120))
    assert L.nnet_compute_noalloc(ctx.h, arr, 120, 1e-4, 1e-4, 0, None) == -1
    assert b"at most" in L.nnet_last_error(ctx.h)
    ctx.close()
