"""Autotuner of the solve launches, run on the device it tunes.

It times every launch choice the library offers on a seeded sample of the workload the choice
matters for, and writes the fastest to the table that nnet_ctx_create loads by default
(data/tuning_gfx950.txt next to lib/; format in include/nnet_mi355x.h):

  ppw 14 <eos> <precision> <1|2|4>   particles per workgroup of net14 (one wavefront per particle):
                                     ideal gas on the net14 ideal-gas sample, Helmholtz on the
                                     net14 Helmholtz sample, FP64 and mixed precision each
  fused <0|1>                        one launch, or one launch per batch, for a call with batches
                                     of net14 + net86 + net87 (the mixed-network hydro-loop shape)

net86 / net87 particles take a whole workgroup of four wavefronts (the matrix tiles of one
particle fill their registers), so their particles per workgroup is 1; their LU block of four
pivots is what two resident workgroups per CU leave room for in LDS.  Both are written as they are.

Each candidate runs from the same seeded state: one warmup step, then --steps timed steps (HIP
events around each call); the candidates alternate over --rounds rounds and each keeps its best
round.  A candidate replaces the default only when it is faster by more than --margin.

usage: python sph-exa-nnet_amd/autotune.py [--out PATH] [--n14 N] [--n86 N] [--steps K]
"""
import argparse
import datetime
import json
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

import numpy as np  # noqa: E402

import workloads  # noqa: E402

TABLE = os.path.join(HERE, "data", "helm_table.bin")
DEFAULT_OUT = os.path.join(HERE, "data", "tuning_gfx950.txt")


def _state(torch, nm, cfg):
    st = nm.device_state(torch, cfg["net"], cfg["rho"], cfg["temp"], cfg["Y"], cfg["dt"])
    st["cost_hint"] = st["newton"]
    return st


def _time_calls(torch, ctx, cfgs, precision, hdts, warmup=1):
    """Seconds per call (mean of the timed calls) of one compute call over the batches of cfgs."""
    import nnet_mi355x as nm
    sts = [_state(torch, nm, c) for c in cfgs]
    bs = [ctx.make_batch(c["net"], st, eos=c["eos"], mu=c["mu"], Z=c["Z"], precision=precision)
          for c, st in zip(cfgs, sts)]
    stream = torch.cuda.current_stream()
    for k in range(warmup):
        ctx.compute(bs, float(hdts[k % len(hdts)]), stream=stream.cuda_stream, sync=False)
    torch.cuda.synchronize()
    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in hdts]
    for (a, b), h in zip(ev, hdts):
        a.record(stream)
        ctx.compute(bs, float(h), stream=stream.cuda_stream, sync=False)
        b.record(stream)
    torch.cuda.synchronize()
    newton = sum(int(st["newton"].sum().item()) for st in sts)
    return float(np.mean([a.elapsed_time(b) for a, b in ev])) / 1e3, newton


def _pick(name, cands, run, rounds, margin, log):
    """Runs every candidate `rounds` times (alternating), keeps each one's best time; returns the
    fastest, or the first (the default) unless another beats it by more than `margin`."""
    best = {c: float("inf") for c in cands}
    its = {}
    for r in range(rounds):
        for c in cands:
            t, newton = run(c)
            best[c] = min(best[c], t)
            its.setdefault(c, newton)
            print(f"  {name} {c}: round {r} {t * 1e3:9.2f} ms per step, Newton iterations {newton}", flush=True)
    default = cands[0]
    win = min(cands, key=lambda c: best[c])
    if best[win] > best[default] * (1 - margin):
        win = default
    log.append({"setting": name, "candidates": {str(c): best[c] for c in cands}, "newton": {str(c): its[c] for c in cands},
                "chosen": win, "gain_over_default": best[default] / best[win] - 1})
    print(f"  {name}: chosen {win} ({best[default] / best[win] - 1:+.1%} over {default})", flush=True)
    return win


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--out", default=DEFAULT_OUT)
    ap.add_argument("--n14", type=int, default=262144, help="particles of the net14 samples")
    ap.add_argument("--n86", type=int, default=65536, help="net86 particles of the mixed-network call "
                    "(as many net14, half as many net87)")
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--rounds", type=int, default=2)
    ap.add_argument("--margin", type=float, default=0.02)
    ap.add_argument("--json", default=None, help="also write the measurements as JSON")
    args = ap.parse_args()

    import torch
    import nnet_mi355x as nm
    if not torch.cuda.is_available():
        sys.exit("autotune.py needs a HIP device")
    os.environ["NNET_MI355X_TUNING"] = ""  # measure from the defaults, not from an earlier table
    ctx = nm.Context(0, helm_table=TABLE)
    ctx.set_max_iterations(10000)
    log = []
    lines = []
    t0 = time.time()
    # net14: particles per workgroup per EOS and precision
    for eos, cid in ((0, 1), (1, 2)):
        cfg = workloads.config(cid, n=args.n14, seed=workloads.SEED0 + cid)
        hdts = [cfg["hydro_dt"]] * args.steps
        for prec in (nm.PRECISION_FP64, nm.PRECISION_MIXED):
            name = f"ppw 14 {eos} {prec}"
            print(f"{name}: {cfg['label']}, {args.n14} particles", flush=True)

            def run(ppw, cfg=cfg, eos=eos, prec=prec, hdts=hdts):
                ctx.set_particles_per_workgroup(14, eos, prec, ppw)
                return _time_calls(torch, ctx, [cfg], prec, hdts)

            win = _pick(name, [1, 2, 4], run, args.rounds, args.margin, log)
            ctx.set_particles_per_workgroup(14, eos, prec, win)
            lines.append(f"ppw 14 {eos} {prec} {win}")
    for net in (86, 87):
        for eos in (0, 1):
            for prec in (0, 1):
                lines.append(f"ppw {net} {eos} {prec} 1")
    # one launch for batches of several networks (FP64; the mixed-network loop's networks and EOS)
    # the mixed-network loop's shape: net14 = net86 = 2 x net87 particles, its per-step dts
    parts, dts = workloads.config5(n=5 * args.n86 // 2, seed=workloads.SEED0 + 5, steps=args.steps)
    print(f"fused: net14 + net86 + net87 ({[p['rho'].size for p in parts]} particles), FP64", flush=True)

    def run_fused(f):
        ctx.set_fused_launch(f)
        return _time_calls(torch, ctx, parts, nm.PRECISION_FP64, list(dts))

    win = _pick("fused", [1, 0], run_fused, args.rounds, args.margin, log)
    lines.append(f"fused {win}")

    dev = torch.cuda.get_device_properties(0)
    head = [f"# launch tuning of libnnet_mi355x, written by sph-exa-nnet_amd/autotune.py",
            f"# device: {dev.name} ({getattr(dev, 'gcnArchName', 'gfx')}), {dev.multi_processor_count} CUs; "
            f"{datetime.date.today().isoformat()}; {time.time() - t0:.0f} s",
            "# ppw <rates kind> <eos 0 ideal, 1 Helmholtz> <precision 0 FP64, 1 mixed> <particles per workgroup>",
            "# fused <one launch for a call with batches of several networks: 1, or one launch per batch: 0>"]
    with open(args.out, "w") as f:
        f.write("\n".join(head + lines) + "\n")
    print(f"wrote {args.out}", flush=True)
    if args.json:
        with open(args.json, "w") as f:
            json.dump({"table": lines, "measurements": log}, f, indent=1)
    ctx.close()


if __name__ == "__main__":
    main()
