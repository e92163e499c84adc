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
# bytes).  FP64 (`--precision fp64 --steps 1 --warmup 0 --no-cpu`, one dispatch): 2 x 4,815,118 KiB
# fetched + 3,973,413 KiB written.  Mixed (`--precision mixed --steps 1 --warmup 1 --no-cpu`, the
# hinted step: routed FP64 pass + mixed pass + redo pass): 2 x 6,702,294 KiB fetched + 4,368,279
# KiB written; per 2,097,152 particles.
TRAFFIC_BYTES_PER_PARTICLE_STEP = {(3, "fp64"): (2 * 4815117.8 + 3973413.2) * 1024 / 2097152,
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


def host_cpus():
    """(threads to use, physical cores available, model name): the physical cores of this process's
    CPU affinity set (lscpu core ids), at most 16 (the host share of one GPU)."""
    import subprocess
    aff = sorted(os.sched_getaffinity(0))
    cores, model = set(), "unknown"
    try:
        for line in subprocess.run(["lscpu", "-p=CPU,CORE,SOCKET"], capture_output=True, text=True,
                                   check=True).stdout.splitlines():
            if line.startswith("#"):
                continue
            cpu, core, sock = (int(x) for x in line.split(",")[:3])
            if cpu in aff:
                cores.add((sock, core))
        for line in subprocess.run(["lscpu"], capture_output=True, text=True, check=True).stdout.splitlines():
            if line.startswith("Model name:"):
                model = line.split(":", 1)[1].strip()
    except (OSError, subprocess.CalledProcessError, ValueError):
        cores = set(aff)
    phys = max(1, len(cores))
    return min(16, phys), phys, model


def stratified_sample(cfg, n, m):
    """m particle indices of an n-particle configuration, taken evenly within each stratum (hot /
    cool for C3, all particles otherwise) in proportion to the stratum's size."""
    strata = [np.nonzero(cfg["hot"])[0], np.nonzero(~cfg["hot"])[0]] if "hot" in cfg else [np.arange(n)]
    idx = []
    for s in strata:
        k = int(round(m * len(s) / n))
        if k:
            idx.append(s[np.linspace(0, len(s) - 1, k).astype(np.int64)])
    return np.sort(np.concatenate(idx))


def cpu_baseline(cfg, snap):
    """The reference CPU path (its compiled substep solver in its OpenMP particle loop,
    parallel_nuclear_net_impl.cpp:63-103, with the replay detection and Newton cap that let it
    return on every particle; oracle/_ref) -- or, where that build is absent, the oracle
    restatement -- on `snap`: a stratified subsample of the device state taken at the start of the
    first timed step, integrated over one hydro step on this host."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as orc
    threads, phys, model = host_cpus()
    Y = np.ascontiguousarray(snap["Y"])
    T, dt, rho = snap["temp"].copy(), snap["dt"].copy(), snap["rho"].copy()
    m = len(rho)
    if orc.reference_available() and orc.Reference().has_helmholtz:
        lib, kind = orc.Reference(), "reference"
        t0 = time.perf_counter()
        st, new = lib.batch_guarded(cfg["net"], rho, T, Y, dt, cfg["hydro_dt"], eos_kind=cfg["eos"], mu=cfg["mu"],
                                    Z=cfg["Z"], nthreads=threads, max_iter=snap["max_iter"])
    else:
        lib, kind = orc.Oracle(), "port"
        lib.helm_load(TABLE)
        t0 = time.perf_counter()
        st, _, new = lib.batch(cfg["net"], rho, T, Y, dt, cfg["hydro_dt"], eos_kind=cfg["eos"], mu=cfg["mu"],
                               Z=cfg["Z"], nthreads=threads, max_iter=snap["max_iter"])
    el = time.perf_counter() - t0
    return {"value": m / el, "unit": "particle-steps/s", "cores": threads, "kind": kind,
            "sample": f"{m} particles, stratified, of the device state at the start of the first timed FP64 step, "
                      f"one hydro step of {cfg['hydro_dt']} s, {el:.2f} s on {threads} threads "
                      f"({phys} physical cores available, {model}); statuses "
                      f"{np.bincount(st, minlength=5).tolist()}, Newton iterations {int(new.sum())}"}


def run_c5(args, torch, nm):
    """Mixed-network batch (net14 C2 + net86 C3 + net87 C4 distributions, 2:2:1) in one call per
    hydro step; `--steps` steps with per-step dt log-uniform in [1e-6, 1e-3] are captured once in
    a HIP graph; the timed region is `--warmup` untimed replays then one timed replay of the loop
    (a replay continues from the state the previous one left)."""
    if args.precision != "fp64":
        # nnet_compute refuses a mixed-precision batch in a captured multi-batch call: a host fault
        # once seen in that combination has no shown cause yet
        print("bench.py --config 5 runs the FP64 path (--precision fp64): mixed precision in a captured "
              "multi-batch call is disabled", file=sys.stderr)
        sys.exit(2)
    parts, dts = workloads.config5(n=args.n, steps=args.steps)
    prec = nm.PRECISION_FP64
    ctx = nm.Context(0, helm_table=TABLE)
    if args.per_batch_launches:
        ctx.set_fused_launch(False)
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
                   "launch": "one fused launch per step" if ctx.fused_launch() else "one launch per batch",
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
    ap.add_argument("--cpu-sample", type=int, default=262144,
                    help="particles of the CPU-baseline sample: a stratified subsample of the device state at the "
                         "start of the first timed step, one hydro step on the host's physical cores (C3: ~5 s "
                         "on 16 cores, ~80 core-seconds)")
    ap.add_argument("--max-iter", type=int, default=10000,
                    help="Newton-iteration cap per particle and step (nnet_ctx_set_max_iterations; the oracle and "
                         "the CPU baseline use the same cap)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--per-batch-launches", action="store_true",
                    help="--config 5: one launch per batch (side streams) instead of the fused launch")
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
    ctx.set_max_iterations(args.max_iter)
    stream = torch.cuda.current_stream()
    slots = ctx.resident_slots({2: 14, 3: 86, 4: 87}[args.config])

    def run(precision, snapshot):
        """Warmup + timed hydro steps of one precision from the seeded initial state.  Returns the
        measurements and the final device state; with snapshot=True also a stratified subsample of
        the state at the start of the first timed step (the CPU baseline's input)."""
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
        print(f"# {cfg['label']}: n={n} precision={precision} warmup={args.warmup} steps={args.steps} "
              f"cap={args.max_iter}", file=sys.stderr, flush=True)
        for w in range(args.warmup):
            ctx.compute([batch], hdt, stream=stream.cuda_stream, sync=False)
            torch.cuda.synchronize()
            print(f"# {precision} warmup step {w} done", file=sys.stderr, flush=True)
        snap = None
        if snapshot:
            idx = torch.from_numpy(stratified_sample(cfg, n, min(args.cpu_sample, n))).cuda()
            snap = {k: st[k][..., idx].cpu().numpy() for k in ("rho", "temp", "dt", "Y")}
            snap["max_iter"] = args.max_iter
        ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(args.steps)]
        newton_total = fp64_total = capped_newton = 0
        status_hist = np.zeros(5, np.int64)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for k in range(args.steps):
            ev[k][0].record(stream)
            ctx.compute([batch], hdt, stream=stream.cuda_stream, sync=False)
            ev[k][1].record(stream)
            # counters of this step (small device reductions and one status copy after the step)
            newton_total += int(st["newton"].sum().item())
            fp64_total += int(st["fp64_solves"].sum().item())
            capped_newton += int(st["newton"][st["status"] == 4].sum().item())
            status_hist += np.bincount(st["status"].cpu().numpy(), minlength=5)
            print(f"# {precision} step {k} done", file=sys.stderr, flush=True)
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        kernel_s = float(np.mean([x.elapsed_time(y) for x, y in ev])) / 1e3
        it_step = newton_total / args.steps
        flops_per_launch = it_step * flops_per_iter(cfg["net"])
        out = {"particle_steps_per_s": n * args.steps / wall, "ms_per_step": wall / args.steps * 1e3,
               "kernel_ms": kernel_s * 1e3, "achieved_tflops": flops_per_launch / kernel_s / 1e12,
               "newton_iters_per_step": it_step, "newton_iters_per_s": it_step / kernel_s,
               "fp64_resolves_per_step": fp64_total / args.steps,
               "status_per_step": (status_hist / args.steps).tolist(),
               "max_iter_particles_per_step": float(status_hist[4]) / args.steps,
               "max_iter_share_of_iterations": capped_newton / max(newton_total, 1),
               # resident workgroups x kernel time / iterations: the mean latency of one Newton
               # iteration of one particle (the work queue keeps every slot busy until the tail)
               "iteration_latency_us": slots * kernel_s / max(it_step, 1) * 1e6}
        final = {k: st[k] for k in ("Y", "temp", "status", "newton")}
        return cfg, out, final, snap

    def precision_errors(ref, mix, cfg):
        """Per-output difference of the mixed path's final state from the FP64 path's (both from the
        same seeded state after the same steps), over the particles with the same status that
        neither path stopped at the Newton cap: mass fractions X = Y A (relative, floor 1e-14),
        temperature (relative) and specific energy release eps = sum_i BE_i (Y_i - Y0_i)."""
        A = torch.tensor(workloads.mass_numbers(cfg["net"]), dtype=torch.float64, device="cuda:0")[:, None]
        BE = torch.tensor(workloads.binding_energies(cfg["net"]), dtype=torch.float64, device="cuda:0")
        Y0 = torch.tensor(cfg["Y"], dtype=torch.float64, device="cuda:0")
        sel = (ref["status"] == mix["status"]) & (ref["status"] != 4) & (ref["status"] != 1)
        Xr, Xm = ref["Y"][:, sel] * A, mix["Y"][:, sel] * A
        relx = ((Xm - Xr).abs() / Xr.abs().clamp_min(1e-14)).amax(0)
        okx = ((Xm - Xr).abs() <= 1e-6 * Xr.abs() + 1e-14).all(0)
        relT = (mix["temp"][sel] - ref["temp"][sel]).abs() / ref["temp"][sel].abs()
        er = BE @ (ref["Y"][:, sel] - Y0[:, sel])
        em = BE @ (mix["Y"][:, sel] - Y0[:, sel])
        nz = er.abs() > 0
        rele = (em - er)[nz].abs() / er[nz].abs()
        m = int(sel.sum().item())
        q = lambda t, f: float(t.max().item()) if f == "max" else float(t.median().item())
        return {"particles_compared": m, "status_mismatch": int((ref["status"] != mix["status"]).sum().item()),
                "X_max_rel": q(relx, "max") if m else 0.0, "X_median_rel": q(relx, "median") if m else 0.0,
                "T_max_rel": q(relT, "max") if m else 0.0, "T_median_rel": q(relT, "median") if m else 0.0,
                "eps_max_rel": q(rele, "max") if rele.numel() else 0.0,
                "eps_median_rel": q(rele, "median") if rele.numel() else 0.0,
                "within_1e-6_X_1e-8_T": float((okx & (relT <= 1e-8)).float().mean().item()) if m else 1.0}

    precisions = ["fp64", "mixed"] if args.precision == "both" else [args.precision]
    res, finals, snap = {}, {}, None
    for p in precisions:
        cfg, res[p], finals[p], sn = run(p, snapshot=(p == precisions[0] and not args.no_cpu))
        snap = snap or sn
    head_prec = "mixed" if "mixed" in res else "fp64"
    head = res[head_prec]
    conf = {"workload": cfg["label"], "particles": n, "hydro_dt": cfg["hydro_dt"], "precision": args.precision,
            "max_iter": args.max_iter, "resident_slots": slots}
    conf.update({p: r for p, r in res.items()})
    if "mixed" in res and "fp64" in res:
        conf["mixed_over_fp64"] = res["mixed"]["particle_steps_per_s"] / res["fp64"]["particle_steps_per_s"]
        conf["mixed_vs_fp64_error"] = precision_errors(finals["fp64"], finals["mixed"], cfg)
    # roofline of the dominant kernel, the FP64 solve (all of the FP64 path's time, and the
    # routed and redo passes that carry most of the mixed path's)
    dom = res["fp64"] if "fp64" in res else head
    achieved = dom["achieved_tflops"] * 1e12
    tb = TRAFFIC_BYTES_PER_PARTICLE_STEP.get((args.config, "fp64" if "fp64" in res else head_prec))
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
                     "frac": achieved / PEAK_FP64, "traffic": tb * n if tb is not None else None,
                     "traffic_source": "constant: calibrated rocprofv3 FETCH_SIZE/WRITE_SIZE of the profile run",
                     "kernel": f"nnet_solve_kernel<{cfg['net']}, HELMHOLTZ, FP64>",
                     # the kernel is latency x slot bound: iterations/s = slots / iteration latency
                     "limiter": "latency",
                     "latency_model": {"resident_slots": slots, "iteration_latency_us": dom["iteration_latency_us"],
                                       "newton_iters_per_s": dom["newton_iters_per_s"]}},
    }
    out["cpu_baseline"] = cpu_baseline(cfg, snap) if snap is not None else None
    print(json.dumps(out), flush=True)
    ctx.close()


if __name__ == "__main__":
    main()
