#!/usr/bin/env python3
"""Throughput of the batched implicit nuclear-network solve on one MI355X.

Workload (default): net86 + Helmholtz EOS, 2,097,152 particles, 70% cool (T log-U[1e8,1e9] K,
rho log-U[1e4,1e7] g/cm^3) / 30% hot (T log-U[3e9,8e9], rho log-U[1e7,5e9]), X(12C)=0.3,
X(16O)=0.6, X(20Ne)=0.1, hydro dt = 1e-4 s per step, state resident on the device.  A "step" is
one computeNuclearReactions call over all particles (one hydro step).  By default the plain FP64
path and the mixed-precision path (FP32 tile LU on MFMA + FP64 iterative refinement, FP64 re-solve
of iterations that do not refine) each run from the same seeded state; `value` is the mixed path's
rate and config.mixed_over_fp64 the ratio.

Prints one JSON line: particle-steps/s over the timed steps, the dominant kernel's roofline
(algorithmic FLOPs from the device Newton-iteration counters: per iteration of an n = d+1
system, 2n^3/3 + 1.5n^2 - 7n/6 (dense unpivoted LU + solve) + 3d^2 (assembly), divided by the
kernel time measured with HIP events on the stream the kernel runs on), and the CPU baseline:
the FP64 CPU restatement of the reference's OpenMP path timed on this host on a bounded sample.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path[:0] = [os.path.join(ROOT, "sph-exa-nnet_amd")]

import workloads  # noqa: E402

TABLE = os.path.join(ROOT, "sph-exa-nnet_amd", "data", "helm_table.bin")
PEAK_FP64 = 78.6e12  # MI355X dense FP64 (vector and matrix)

# HBM bytes per particle-step of the solve kernels on the net86 C3 distribution, from rocprofv3
# --pmc FETCH_SIZE and --pmc WRITE_SIZE runs, scaled by the same counters on a plain 512 MiB device
# copy (`bench.py --copy-calibration 512`: FETCH_SIZE reads 0.5x and WRITE_SIZE 1.0x of the copied
# bytes).  FP64 (`--precision fp64 --steps 1 --warmup 0 --no-cpu`, one dispatch): 2 x 4,857,641 KiB
# fetched + 4,010,909 KiB written.  Mixed (`--precision mixed --steps 1 --warmup 1 --no-cpu`, the
# hinted step: routed FP64 pass + mixed pass + redo pass): 2 x 6,702,294 KiB fetched + 4,368,279
# KiB written; per 2,097,152 particles.
TRAFFIC_BYTES_PER_PARTICLE_STEP = {(3, "fp64"): (2 * 4857640.94 + 4010909.16) * 1024 / 2097152,
                                   (3, "mixed"): (2 * 6702293.56 + 4368279.28) * 1024 / 2097152}

# The same for the EOS refresh kernel over the C3 particle set (`--workload helm-eos --steps 2
# --warmup 0 --no-cpu`): 2 x 755,378 KiB fetched + 81,920 KiB written per 2,097,152 particles.
EOS_TRAFFIC_BYTES_PER_PARTICLE = {3: (2 * 755378.03 + 81920.0) * 1024 / 2097152}


def flops_per_iter(net):
    d = {14: 14, 86: 86, 87: 87}[net]
    n = d + 1
    return 2 * n ** 3 / 3 + 1.5 * n ** 2 - 7 * n / 6 + 3 * d ** 2


def device_config(torch, cid, n):
    """Upload the seeded particle set of configuration cid (Y built on the device)."""
    return device_config_from(torch, workloads.config(cid, n=n, seed=workloads.SEED0 + cid))


def cpu_baseline(cid, sample, steps_hint):
    """The CPU restatement of the reference OpenMP path (oracle) on a bounded sample."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    from oracle import Oracle
    o = Oracle()
    o.helm_load(TABLE)
    cfg = workloads.config(cid, n=sample, seed=workloads.SEED0 + 100 + cid)
    Y = np.ascontiguousarray(cfg["Y"])
    T, dt, rho = cfg["temp"].copy(), cfg["dt"].copy(), cfg["rho"].copy()
    threads = os.cpu_count() or 1
    t0 = time.perf_counter()
    st, sub, new = o.batch(cfg["net"], rho, T, Y, dt, cfg["hydro_dt"], eos_kind=cfg["eos"], mu=cfg["mu"],
                           Z=cfg["Z"], nthreads=threads)
    el = time.perf_counter() - t0
    return {"value": sample / el, "unit": "particle-steps/s", "cores": threads, "kind": "port",
            "sample": f"first hydro step of {sample} particles of the same distribution (seed +100), "
                      f"{el:.2f} s, statuses {np.bincount(st, minlength=5).tolist()}"}


def run_c5(args, torch, nm):
    """Mixed-network batch (net14 C2 + net86 C3 + net87 C4 distributions, 2:2:1) in one call per
    hydro step; `--steps` steps with per-step dt log-uniform in [1e-6, 1e-3] are captured once in
    a HIP graph; the timed region is `--warmup` untimed replays then one timed replay of the loop
    (a replay continues from the state the previous one left)."""
    parts, dts = workloads.config5(n=args.n, steps=args.steps)
    prec = nm.PRECISION_FP64 if args.precision == "fp64" else nm.PRECISION_MIXED
    ctx = nm.Context(0, helm_table=TABLE)
    sts = []
    for c in parts:
        _, st = device_config_from(torch, c)
        st["cost_hint"] = st["newton"]
        sts.append(st)
    bs = [ctx.make_batch(c["net"], st, eos=c["eos"], mu=c["mu"], Z=c["Z"], precision=prec)
          for c, st in zip(parts, sts)]
    n = sum(c["rho"].size for c in parts)
    stream = torch.cuda.Stream()
    with torch.cuda.stream(stream):
        ctx.compute(bs, float(dts[0]), stream=stream.cuda_stream, sync=False)  # sizes the library's buffers
    stream.synchronize()
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.graph(graph, stream=stream):
        for dt in dts:
            ctx.compute(bs, float(dt), stream=stream.cuda_stream, sync=False)
    print(f"# C5: n={n} ({[c['rho'].size for c in parts]}), {len(dts)} steps captured", file=sys.stderr, flush=True)
    for w in range(args.warmup):
        graph.replay()
        torch.cuda.synchronize()
        print(f"# warmup replay {w} done", file=sys.stderr, flush=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    graph.replay()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    hist = np.zeros(8, np.int64)  # substeps of the last step: 0, 1, 2-3, 4-7, ... >= 64
    for st in sts:
        s_ = st["substeps"].cpu().numpy().astype(np.int64)
        b = np.where(s_ == 0, 0, np.minimum(7, np.floor(np.log2(np.maximum(s_, 1))).astype(int) + 1))
        hist += np.bincount(b, minlength=8)
    out = {
        "metric": "net86 particle-steps/s, mixed vs FP64 on 1 GPU at <=1e-6 rel mass-frac error",
        "value": n * len(dts) / wall,
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": len(dts),
        "warmup": args.warmup,
        "ms_per_step": wall / len(dts) * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if args.precision == "fp64" else "f32+f64",
        "data": "synthetic",
        "config": {"workload": "mixed-network C5 (net14 + net86 + net87, one call per step, HIP graph)",
                   "particles": [int(c["rho"].size) for c in parts], "precision": args.precision,
                   "hydro_dt_range": [float(dts.min()), float(dts.max())],
                   "last_step_substeps_hist_bins": ["0", "1", "2-3", "4-7", "8-15", "16-31", "32-63", ">=64"],
                   "last_step_substeps_hist": hist.tolist()},
        "roofline": None,
        "cpu_baseline": None,
    }
    print(json.dumps(out), flush=True)
    ctx.close()


def run_helm_eos(args, torch, nm):
    """Post-step EOS refresh (computeHelmEOS -> nnet_helm_eos) over the particle set of
    configuration `--config` (default C3: 2,097,152 net86 particles): `--warmup` untimed and
    `--steps` timed launches on one stream, inputs resident in HBM.  Algorithmic HBM bytes per
    particle: Y (d doubles), temp and rho read, u, cv, p, c, dP/dT written."""
    cid = args.config
    n = args.n or {2: 1048576, 3: 2097152, 4: 2097152}[cid]
    cfg, st = device_config(torch, cid, n)
    ctx = nm.Context(0, helm_table=TABLE)
    stream = torch.cuda.current_stream()
    out = {k: torch.empty(n, dtype=torch.float64, device="cuda:0") for k in ("u", "cv", "p", "c", "dpdT")}
    launch = lambda: ctx.helm_eos(cfg["Z"], st["temp"], st["rho"], st["Y"], out, stream=stream.cuda_stream,
                                  sync=False)
    print(f"# helm-eos {cfg['label']}: n={n} warmup={args.warmup} steps={args.steps}", file=sys.stderr, flush=True)
    for _ in range(args.warmup):
        launch()
    torch.cuda.synchronize()
    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(args.steps)]
    t0 = time.perf_counter()
    for k in range(args.steps):
        ev[k][0].record(stream)
        launch()
        ev[k][1].record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    kernel_s = float(np.mean([x.elapsed_time(y) for x, y in ev])) / 1e3
    d = cfg["Y"].shape[0]
    bytes_per_particle = 8 * (d + 2 + 5)
    achieved = bytes_per_particle * n / kernel_s
    res = {
        "metric": f"net{cfg['net']} Helmholtz EOS refresh (computeHelmEOS) particles/s on 1 GPU",
        "value": n * args.steps / wall,
        "unit": "particles/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"EOS refresh over the {cfg['label']} particle set", "particles": n,
                   "kernel_ms": kernel_s * 1e3, "algorithmic_bytes_per_particle": bytes_per_particle},
        "roofline": {"bound": "hbm", "achieved": achieved / 1e9, "peak": 8000.0, "unit": "GB/s",
                     "frac": achieved / 8e12,
                     "traffic": EOS_TRAFFIC_BYTES_PER_PARTICLE[cid] * n if cid in EOS_TRAFFIC_BYTES_PER_PARTICLE
                     else None},
        "cpu_baseline": None,
    }
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        from oracle import Oracle
        o = Oracle()
        o.helm_load(TABLE)
        sample = min(n, 1 << 20)
        c = workloads.config(cid, n=sample, seed=workloads.SEED0 + 100 + cid)
        threads = os.cpu_count() or 1
        t0 = time.perf_counter()
        o.helm_eos_batch(c["Z"], c["temp"], c["rho"], c["Y"], nthreads=threads)
        el = time.perf_counter() - t0
        res["cpu_baseline"] = {"value": sample / el, "unit": "particles/s", "cores": threads, "kind": "port",
                               "sample": f"{sample} particles of the same distribution (seed +100), {el:.2f} s"}
    print(json.dumps(res), flush=True)
    ctx.close()


def device_config_from(torch, cfg):
    """Upload one configuration dict's particle set (Y built on the device)."""
    n = cfg["rho"].size
    dev = "cuda:0"
    st = {
        "rho": torch.tensor(cfg["rho"], dtype=torch.float64, device=dev),
        "rho_m1": torch.tensor(cfg["rho"], dtype=torch.float64, device=dev),
        "temp": torch.tensor(cfg["temp"], dtype=torch.float64, device=dev),
        "dt": torch.tensor(cfg["dt"], dtype=torch.float64, device=dev),
        "status": torch.zeros(n, dtype=torch.uint8, device=dev),
        "substeps": torch.zeros(n, dtype=torch.int32, device=dev),
        "newton": torch.zeros(n, dtype=torch.int32, device=dev),
        "fp64_solves": torch.zeros(n, dtype=torch.int32, device=dev),
    }
    Y = torch.empty(cfg["Y"].shape, dtype=torch.float64, device=dev)
    for s in range(cfg["Y"].shape[0]):
        Y[s].copy_(torch.from_numpy(cfg["Y"][s]))
    st["Y"] = Y
    return cfg, st


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", type=int, default=3, choices=[2, 3, 4, 5],
                    help="3 (default): net86 C3; 5: mixed-network batch, hydro loop captured in a HIP graph")
    ap.add_argument("--n", type=int, default=None, help="particles (default: the configuration's size)")
    ap.add_argument("--precision", default="both", choices=["fp64", "mixed", "both"],
                    help="both: the FP64 path, then the mixed path, each from the same initial state; "
                         "value is the mixed path's rate")
    ap.add_argument("--workload", default="solve", choices=["solve", "helm-eos"],
                    help="solve (default): computeNuclearReactions; helm-eos: the post-step EOS refresh "
                         "computeHelmEOS over the same particle set")
    ap.add_argument("--cpu-sample", type=int, default=524288,
                    help="particles of the CPU-baseline sample (the oracle on all host cores, ~10-30 s)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run a plain device copy of MIB MiB (3 times): the reference kernel whose "
                         "FETCH_SIZE/WRITE_SIZE counters calibrate the HBM traffic counters")
    args = ap.parse_args()
    if args.copy_calibration:
        import torch
        nbytes = args.copy_calibration << 20
        src = torch.ones(nbytes // 8, dtype=torch.float64, device="cuda:0")
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_calibration": {"bytes_read_per_copy": nbytes, "bytes_written_per_copy": nbytes,
                                               "copies": 3}}), flush=True)
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported in this build "
              f"(particles are independent: N GPUs would each run their own share)", file=sys.stderr)
        sys.exit(2)

    import torch
    import nnet_mi355x as nm

    if args.workload == "helm-eos":
        run_helm_eos(args, torch, nm)
        return
    if args.config == 5:
        run_c5(args, torch, nm)
        return
    n = args.n or {2: 1048576, 3: 2097152, 4: 2097152}[args.config]
    ctx = nm.Context(0, helm_table=TABLE)
    stream = torch.cuda.current_stream()

    def run(precision):
        """Warmup + timed hydro steps of one precision from the seeded initial state."""
        cfg, st = device_config(torch, args.config, n)
        prec = nm.PRECISION_FP64 if precision == "fp64" else nm.PRECISION_MIXED
        # each step dispatches the particles in descending order of the previous step's Newton
        # iterations (the newton output array doubles as the cost hint)
        st["cost_hint"] = st["newton"]
        # mixed: particles whose FP32 factors did not refine in the previous step go straight to
        # the FP64 pass (the fp64_solves output doubles as the hint)
        st["fp64_hint"] = st["fp64_solves"]
        batch = ctx.make_batch(cfg["net"], st, eos=cfg["eos"], mu=cfg["mu"], Z=cfg["Z"], precision=prec)
        hdt = cfg["hydro_dt"]
        print(f"# {cfg['label']}: n={n} precision={precision} warmup={args.warmup} steps={args.steps}",
              file=sys.stderr, flush=True)
        for w in range(args.warmup):
            ctx.compute([batch], hdt, stream=stream.cuda_stream, sync=False)
            torch.cuda.synchronize()
            print(f"# {precision} warmup step {w} done", file=sys.stderr, flush=True)
        ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(args.steps)]
        newton_total = fp64_total = 0
        status_hist = np.zeros(5, np.int64)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for k in range(args.steps):
            ev[k][0].record(stream)
            ctx.compute([batch], hdt, stream=stream.cuda_stream, sync=False)
            ev[k][1].record(stream)
            # counters of this step (tiny D2H reads after the kernel on the same stream)
            newton_total += int(st["newton"].sum().item())
            fp64_total += int(st["fp64_solves"].sum().item())
            status_hist += np.bincount(st["status"].cpu().numpy(), minlength=5)
            print(f"# {precision} step {k} done", file=sys.stderr, flush=True)
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        kernel_s = float(np.mean([x.elapsed_time(y) for x, y in ev])) / 1e3
        flops_per_launch = newton_total / args.steps * flops_per_iter(cfg["net"])
        return cfg, {"particle_steps_per_s": n * args.steps / wall, "ms_per_step": wall / args.steps * 1e3,
                     "kernel_ms": kernel_s * 1e3, "achieved_tflops": flops_per_launch / kernel_s / 1e12,
                     "newton_iters_per_step": newton_total / args.steps,
                     "fp64_resolves_per_step": fp64_total / args.steps,
                     "status_per_step": (status_hist / args.steps).tolist()}

    precisions = ["fp64", "mixed"] if args.precision == "both" else [args.precision]
    res = {}
    for p in precisions:
        cfg, res[p] = run(p)
    head = res["mixed"] if "mixed" in res else res["fp64"]
    achieved = head["achieved_tflops"] * 1e12
    head_prec = "mixed" if "mixed" in res else "fp64"
    tb = TRAFFIC_BYTES_PER_PARTICLE_STEP.get((args.config, head_prec))
    traffic = tb * n if tb is not None else None
    conf = {"workload": cfg["label"], "particles": n, "hydro_dt": cfg["hydro_dt"], "precision": args.precision}
    conf.update({p: r for p, r in res.items()})
    if "mixed" in res and "fp64" in res:
        conf["mixed_over_fp64"] = res["mixed"]["particle_steps_per_s"] / res["fp64"]["particle_steps_per_s"]
    out = {
        "metric": "net86 particle-steps/s, mixed vs FP64 on 1 GPU at <=1e-6 rel mass-frac error",
        "value": head["particle_steps_per_s"],
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": head["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if args.precision == "fp64" else "f32+f64",
        "data": "synthetic",
        "config": conf,
        "roofline": {"bound": "mfma", "achieved": achieved / 1e12, "peak": PEAK_FP64 / 1e12, "unit": "TFLOP/s",
                     "frac": achieved / PEAK_FP64, "traffic": traffic},
    }
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(args.config, args.cpu_sample, args.steps)
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out), flush=True)
    ctx.close()


if __name__ == "__main__":
    main()
