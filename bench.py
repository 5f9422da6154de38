"""Benchmark of the PIC step on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload lpa_cfg1]
                    [--precision both|fp64|mixed] [--no-cpu-baseline] [--cpu-steps 1]

A step is one ChimeraRun.make_step of the workload (sort/absorb, deposition,
FB transforms, Poisson correction, PSATD, inverse transforms, gather + push,
moving-window actions on their steps), all state resident in HBM.  W untimed
steps, then K timed steps between two device synchronisations.  Prints ONE
JSON line on stdout; progress goes to stderr.

value = pushed-particle steps per second over the timed steps (electrons; the
still ions are not pushed) of the FP64 run.  With --precision both (default)
the same workload then runs in mixed precision (float32 momenta and weights,
complex64 fields, split-BF16 MFMA transforms with FP32 accumulation, FP64
positions) and its figures go in
the "mixed" object of the same line.  Per-kernel-family times come from HIP
events that libchimera_hip records on its own stream around every launch.

The "scan" object (skipped with --no-scan) is the batched scan of 64 uneven
bunches x 300 steps (workloads.beam_scan): particle-steps/s and p50/p99
time-to-result of the plain path (one simulation after another, eager steps)
and of the batched path (all resident, recorded steps replayed on one stream
each, interleaved), in FP64 and mixed precision.  --serve SECONDS runs the
request loop instead (Poisson arrivals, continuous batching).

HIP's GPU_MAX_HW_QUEUES (hardware queues per process) is set to 16 (or to
CHIMERA_HW_QUEUES): the simulations of a batch run on separate streams, and
with HIP's default of 4 queues at most four of them overlap.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

os.environ["GPU_MAX_HW_QUEUES"] = os.environ.get("CHIMERA_HW_QUEUES", "16")

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

METRIC = "particle-steps/s + FB-transform ms/step, FP64 vs mixed; p50/p99 scan latency"
HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFS = 78.6      # dense MFMA
FP32_PEAK_TFS = 157.3
BF16_PEAK_TFS = 2500.0    # dense MFMA
BF16_PRODUCTS = 6         # split-BF16: six BF16 products per algorithmic product
FAMILIES = ["push_coords", "dep_curr", "dep_dens", "gather_push", "dht", "fft", "fb_ops", "psatd", "sort", "other"]


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def prof_read(lib, fam):
    import ctypes
    t, n = ctypes.c_double(0.0), ctypes.c_int64(0)
    lib.chimera_prof_read(fam, ctypes.addressof(t), ctypes.addressof(n))
    return t.value, n.value


def algorithmic(fam, P, Gr, Gk, Nx, Nkr, M, K, prec="fp64"):
    """(amount, unit, bound) per timed region of a family, from the per-unit
    figures: particles P (per launch), real-space nodes Gr = Nx*(Nkr+1)*M,
    spectral elements Gk = Nx*Nkr*M.  s_u = bytes per momentum component or
    weight (8 FP64, 4 mixed), s_f = bytes per complex field value (16, 8);
    positions are FP64 in both."""
    s_u, s_f = (8, 16) if prec == "fp64" else (4, 8)
    if fam == "gather_push":      # x read, u read and written, w read; the 6-component field read once
        return P * (24 + 7 * s_u) + 6 * Gr * s_f, "B", "hbm"
    if fam == "dep_curr":         # fused push + J + rho: x, u, w read, x and x_half written;
        return P * (72 + 4 * s_u) + 4 * Gr * 2 * s_f, "B", "hbm"   # J (3 components) and rho read+written once
    if fam == "psatd":            # E,G,J,grad rho_n, grad rho_n+1 read, E,G written, 10 table reals
        return Gk * (21 * s_f + 10 * s_u), "B", "hbm"
    return None


def run_gpu(args, precision):
    import chimera_amd as G
    from chimera_amd import workloads as W
    from chimera_amd._lib import lib

    spec = W.with_precision(W.WORKLOADS[args.workload](), precision)
    log(f"building {args.workload} ({precision})")
    t0 = time.time()
    opts = {"SortEvery": args.sort_every} if args.sort_every > 0 else {}
    run = W.build(G, spec, seed=args.seed, **opts)
    L = lib()
    L.chimera_synchronize()
    s, e = run.Solvers[0], run.Particles[0]
    A = s.Args
    Nx, Nkr, M = A["Nx"], A["Nkr"], A["Mtot"]
    log(f"built in {time.time() - t0:.1f}s: Nx={Nx} Nkr={Nkr} modes={M} electrons={e.np} ions={run.Particles[1].np}")
    advance = advance_steps(args, spec)
    t0 = time.time()
    for i in range(1, advance + args.warmup + 1):
        run.make_step(i)
    L.chimera_synchronize()
    log(f"advanced {advance} + warmup {args.warmup} steps in {time.time() - t0:.1f}s")
    L.chimera_prof_reset()
    L.chimera_prof_enable(1)
    psteps = 0
    first = advance + args.warmup + 1
    L.chimera_synchronize()
    t0 = time.perf_counter()
    for i in range(first, first + args.steps):
        run.make_step(i)
        psteps += e.np
    L.chimera_synchronize()
    wall = time.perf_counter() - t0
    L.chimera_prof_enable(0)
    fams = {}
    for f, name in enumerate(FAMILIES):
        ms, n = prof_read(L, f)
        fams[name] = (ms, n)
    log(f"timed {args.steps} steps in {wall:.3f}s")
    for name, (ms, n) in fams.items():
        if n:
            log(f"  {name:12s} {ms / args.steps:9.3f} ms/step  ({n} launches)")
    return dict(run=run, wall=wall, psteps=psteps, fams=fams, Nx=Nx, Nkr=Nkr, M=M, P=e.np, workload=args.workload,
                mixed_form=int(L.chimera_get_mixed_gemm()),
                Gr=Nx * A["Nr"] * M, Gk=Nx * Nkr * M, precision=precision, first=first, last=first + args.steps - 1,
                run_steps=spec["steps"])


def advance_steps(args, spec):
    """Untimed steps before the warmup: by default the timed steps are the last
    ones of the workload's run (its own step count), where the window has moved
    many times and the wake has scrambled the particles."""
    if args.advance >= 0:
        return args.advance
    return max(0, spec["steps"] - args.warmup - args.steps)


def family_roofline(res, K, name):
    """Roofline entry of one kernel family over the timed steps: algorithmic
    bytes (HBM-bound families) or flops (the MFMA GEMMs) per launch, divided
    by the family's average launch duration from the library's HIP events."""
    ms, n = res["fams"][name]
    if not n:
        return None
    per_launch = avg_s = None
    if name in ("dht", "fb_ops"):
        Nx, Nkr, M = res["Nx"], res["Nkr"], res["M"]
        flops = 40.0 * M * Nx * Nkr ** 2 if name == "dht" else 4.0 * (18 * M - 3) * Nx * Nkr ** 2
        form = res.get("mixed_form")
        peak = FP64_PEAK_TFS if res["precision"] == "fp64" else BF16_PEAK_TFS if form == 0 else FP32_PEAK_TFS
        per_launch = flops / (n / K)
        avg_s = ms / n / 1e3
        achieved = per_launch / avg_s / 1e12
        out = {"kernel": name, "bound": "mfma", "achieved": round(achieved, 2), "peak": peak, "unit": "TFLOP/s",
               "frac": round(achieved / peak, 4), "traffic": None, "avg_launch_ms": round(avg_s * 1e3, 4),
               "algorithmic_flops_per_launch": int(per_launch), "launches_per_step": n / K,
               "share_of_step": round(ms / sum(v[0] for v in res["fams"].values()), 3)}
        if res["precision"] != "fp64" and form == 0:
            # split-BF16 issues six BF16 products per algorithmic product
            out["issued_bf16_tflops"] = round(BF16_PRODUCTS * achieved, 1)
            out["issued_frac"] = round(BF16_PRODUCTS * achieved / BF16_PEAK_TFS, 4)
        return out
    alg = algorithmic(name, res["P"], res["Gr"], res["Gk"], res["Nx"], res["Nkr"], res["M"], K, res["precision"])
    if alg is None:
        return None
    amount, unit, bound = alg
    avg_s = ms / n / 1e3
    per_launch = amount / (n / K)
    achieved = per_launch / avg_s / 1e9
    return {"kernel": name, "bound": bound, "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
            "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": traffic_for(name, traffic_key(res["workload"], res["precision"])),
            "avg_launch_ms": round(avg_s * 1e3, 4), "algorithmic_bytes_per_launch": int(per_launch),
            "launches_per_step": n / K,
            "share_of_step": round(ms / sum(v[0] for v in res["fams"].values()), 3)}


def roofline(res, K):
    """The dominant kernel family's roofline (the one that takes the most time
    among those with an algorithmic figure), and every family's."""
    every = {}
    for name in res["fams"]:
        r = family_roofline(res, K, name)
        if r is not None:
            every[name] = r
    if not every:
        return None, every
    dom = max(every, key=lambda k: res["fams"][k][0])
    return every[dom], every


def host_info():
    """CPU model, logical CPUs and physical cores of this host (lscpu)."""
    import subprocess
    info = {"nproc": os.cpu_count()}
    try:
        txt = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=20).stdout
        kv = dict(line.split(":", 1) for line in txt.splitlines() if ":" in line)
        kv = {k.strip(): v.strip() for k, v in kv.items()}
        info["model"] = kv.get("Model name")
        cores, sockets = kv.get("Core(s) per socket"), kv.get("Socket(s)")
        if cores and sockets:
            info["physical_cores"] = int(cores) * int(sockets)
        info["threads_per_core"] = kv.get("Thread(s) per core")
    except (OSError, ValueError, subprocess.SubprocessError):
        pass
    try:
        info["affinity_cpus"] = len(os.sched_getaffinity(0))
    except OSError:
        pass
    return info


def cpu_baseline(args):
    """The oracle (the reference's algorithm restated in C + OpenMP, FP64)
    timed on this host's cores at the headline workload, with the x-chunked
    deposition of the reference's OpenMP path (Xchunked = (T, 10), T = the
    thread count: grid_deps_chnk.f90 deposits one chunk per thread).  Bounded
    sample: the build and initial half step are not timed, then args.cpu_steps
    steps are."""
    from chimera_amd import workloads as W
    from oracle import driver as O
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    spec = W.WORKLOADS[args.workload](xchunked=(threads, 10))
    if args.cpu_slab > 1:
        spec = W.x_slab(spec, args.cpu_slab)
    log(f"CPU baseline: building the oracle at {args.workload} (1/{args.cpu_slab} of its x extent) with "
        f"Xchunked=({threads},10), {threads} OpenMP threads")
    t0 = time.time()
    run = W.build(O, spec, seed=args.seed)
    log(f"CPU baseline: built + half step in {time.time() - t0:.1f}s")
    P = run.Particles[0].Data["coords"].shape[1]
    t0 = time.perf_counter()
    for i in range(1, args.cpu_steps + 1):
        run.make_step(i)
    dt = time.perf_counter() - t0
    log(f"CPU baseline: {args.cpu_steps} step(s) in {dt:.1f}s")
    host = host_info()
    return {"value": round(P * args.cpu_steps / dt, 1), "unit": "particle-steps/s", "cores": threads,
            "kind": "port",
            "sample": f"{args.cpu_steps} make_step of {args.workload} on 1/{args.cpu_slab} of its x extent (the "
                      f"same radial grid, modes, plasma and laser: {P} electrons, Nx={run.Solvers[0].Args['Nx']} x "
                      f"Nkr={run.Solvers[0].Args['Nkr']}; per-particle work as in the whole box) after the untimed "
                      f"build and half step, {dt:.2f} s",
            "host": host, "omp_num_threads": threads,
            "code": "oracle/cpu_ref.c driven by oracle/driver.py: FP64; naive O(Nx Nkr^2) Hankel and operator "
                    "loops as fb_io.f90/fb_math.f90; deposition per x-chunk, one OpenMP thread per chunk, as "
                    "grid_deps_chnk.f90 (Xchunked=(T,10), particles binned every 11 steps); own FFT (radix-2 for "
                    "powers of two, long-double direct sum otherwise) in place of FFTW",
            "compiler": "gcc -O2 -fopenmp -ffp-contract=off (oracle/Makefile)"}


def scan_bench(args):
    """The batched scan: args.scan_n bunches of workloads.beam_scan (uneven
    grids and particle counts, modes m=0,1, space charge, undulator, moving
    window), args.scan_steps steps each.  Four runs: one simulation after
    another with eager steps (the plain path) and all at once with recorded,
    interleaved steps (the batched path), each in FP64 and in mixed precision.
    value = particle-steps/s over the stepping (building not timed);
    time-to-result of each simulation from the start of the stepping."""
    from chimera_amd import serve as S
    from chimera_amd import workloads as W
    specs = W.beam_scan(args.scan_n, seed=args.seed % 100000, steps=args.scan_steps)
    out = {"workload": f"beam_scan: {args.scan_n} bunches x {args.scan_steps} steps",
           "grids": f"Nx {min(sp['params']['Nx'] for sp in specs)}-{max(sp['params']['Nx'] for sp in specs)} x "
                    f"Nkr {min(sp['params']['Nkr'] for sp in specs)}-{max(sp['params']['Nkr'] for sp in specs)}, "
                    f"modes m=0,1"}
    runs = [("plain_fp64", "fp64", False), ("plain_mixed", "mixed", False), ("batched_fp64", "fp64", True),
            ("batched_mixed", "mixed", True)]
    for label, prec, batched in runs:
        if args.scan_only and label not in args.scan_only.split(","):
            continue
        reqs = [S.Request(f"beam{k}", sp, 1000 + k, sp["steps"], prec) for k, sp in enumerate(specs)]
        log(f"scan {label}: {len(reqs)} simulations")
        rate, ttr, wall = S.run_scan(reqs, batched=batched, graphs=batched, chunk=args.scan_chunk,
                                     gate=bool(args.scan_gate))
        parts = sum(r.particle_steps for r in reqs) // args.scan_steps
        out[label] = {"value": round(rate, 1), "unit": "particle-steps/s", "stepping_s": round(wall, 3),
                      "p50_s": round(S.percentile(ttr, 50), 4), "p99_s": round(S.percentile(ttr, 99), 4),
                      "particles": int(parts)}
        log(f"scan {label}: {rate:.3g} particle-steps/s, {wall:.2f} s, p50 {S.percentile(ttr, 50):.3f} s "
            f"p99 {S.percentile(ttr, 99):.3f} s ({parts} particles)")
    if "plain_fp64" in out and "batched_mixed" in out:
        out["batched_mixed_vs_plain_fp64"] = round(out["batched_mixed"]["value"] / out["plain_fp64"]["value"], 3)
    return out


def serve_bench(args):
    """The request loop: a Poisson stream of args.serve_rate requests per second
    for args.serve seconds (mix and scan parameters of serve.poisson_requests),
    served with continuous batching until every request has finished; once in
    FP64 and once in mixed precision.  Throughput = particle-steps of the
    finished requests over the wall time from the first arrival to the last
    completion; time-to-result = completion - arrival (queueing, building and
    stepping)."""
    from chimera_amd import serve as S
    out = {"rate_per_s": args.serve_rate, "arrival_window_s": args.serve, "mix": S.MIX}
    for prec in ("fp64", "mixed"):
        reqs = S.poisson_requests(args.serve_rate, args.serve, args.seed % 100000, prec)
        log(f"serve {prec}: {len(reqs)} requests over {args.serve} s")
        server = S.Server(log=log)
        try:
            wall = S.serve(reqs, server, log=log)
        finally:
            server.close()
        fin = server.finished
        ttr = [r.t_done - r.t_arrival_host for r in fin]
        kinds = {}
        for r in fin:
            kinds.setdefault(r.name.split("#")[0], []).append(r.t_done - r.t_arrival_host)
        rec = {"requests": len(reqs), "finished": len(fin), "wall_s": round(wall, 2),
               "value": round(sum(r.particle_steps for r in fin) / wall, 1), "unit": "particle-steps/s",
               "p50_s": round(S.percentile(ttr, 50), 3), "p99_s": round(S.percentile(ttr, 99), 3),
               "by_kind": {k: {"n": len(v), "p50_s": round(S.percentile(v, 50), 3),
                               "p99_s": round(S.percentile(v, 99), 3)} for k, v in sorted(kinds.items())}}
        log(f"serve {prec}: {rec}")
        out[prec] = rec
    return out


def bunch_stages(nx, nkr, precision, particles=3e5, steps=20, reps=50, seed=5):
    """One bunch of workloads.beam (space charge, analytic undulator, window
    moving with the bunch) on an nx x nkr grid, modes 0-1: after three steps,
    each FB stage of the step is called `reps` times and timed (wall clock
    around the calls, us per call), then `steps` whole steps."""
    import chimera_amd as G
    from chimera_amd import workloads as W
    from chimera_amd._lib import lib
    L = lib()
    spec = W.with_precision(W.beam(Nx=nx, Nkr=nkr, gamma=200.0, sig_x=4.0, sig_r=2.0, emittance=0.5,
                                   n_target=particles, K=2.0, lambda_u=300.0), precision)
    run = W.build_beam(G, spec, seed)
    for i in (1, 2, 3):
        run.make_step(i)
    L.chimera_synchronize()
    s = run.Solvers[0]
    print(f"bunch {nx} x {nkr}, modes 0-1, {run.Particles[0].np} particles, {precision}", flush=True)

    def timed(name, fn):
        fn()
        L.chimera_synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            fn()
        L.chimera_synchronize()
        print(f"{name:24s} {(time.perf_counter() - t0) / reps * 1e6:8.1f} us", flush=True)

    timed("fb_curr_in", s.fb_curr_in)
    timed("fb_dens_in", s.fb_dens_in)
    timed("FBGradDens", s.FBGradDens)
    timed("poiss_corr (3 iter)", s.poiss_corr)
    timed("G2B_FBRot", s.G2B_FBRot)
    timed("fb_fld_out", s.fb_fld_out)
    t0 = time.perf_counter()
    for i in range(4, 4 + steps):
        run.make_step(i)
    L.chimera_synchronize()
    print(f"{'step':24s} {(time.perf_counter() - t0) / steps * 1e6:8.1f} us", flush=True)


def copy_calibration(gib):
    """One device-to-device copy of a known size: the FETCH_SIZE / WRITE_SIZE it
    reads under rocprofv3 --pmc calibrate the counters of the PIC kernels."""
    from chimera_amd.device import DeviceArray, call
    n = int(gib * (1 << 30)) // 8
    a, b = DeviceArray((n,), np.float64), DeviceArray((n,), np.float64)
    call("chimera_synchronize")
    call("chimera_memcpy", b.ptr, a.ptr, n * 8)
    call("chimera_synchronize")
    print(json.dumps({"copy_bytes": n * 8}), flush=True)


def traffic_for(kernel, workload):
    """HBM bytes per launch of a kernel family, measured with rocprofv3 --pmc and
    calibrated against a plain copy (profiles/traffic.json), if recorded.
    workload is the traffic_key (workload, or workload:mixed)."""
    path = os.path.join(HERE, "profiles", "traffic.json")
    try:
        rec = json.load(open(path))
    except (OSError, ValueError):
        return None
    e = rec.get(workload, {}).get(kernel)
    return int(e["bytes_per_launch"]) if e else None


# kernels whose HBM traffic makes up a family's figure (the deposit family also
# runs a small bucket sort of the slot descriptors, left out)
TRAFFIC_KERNELS = {"dep_curr": ("k_push_deposit", "k_dep_combine"), "gather_push": ("k_gather_push",),
                   "psatd": ("k_psatd_spchrg",)}


def traffic_key(workload, precision):
    return workload if precision == "fp64" else f"{workload}:{precision}"


def summarize_traffic(root, workload, steps, out_dir, precision="fp64"):
    """Turn four rocprofv3 --pmc runs into HBM bytes per launch.

    root holds pmc_copy_fetch/, pmc_copy_write/ (python bench.py --copy-calibration 1
    under --pmc FETCH_SIZE, then WRITE_SIZE) and pmc_<precision>_fetch/,
    pmc_<precision>_write/ (python bench.py --steps S --warmup 0 --no-cpu-baseline
    --precision P, same counters).  Each counter is scaled by (known copy bytes) /
    (what it read for the copy); a kernel's figure is the mean over its last S
    launches (the launches of the S steps after the initial half step)."""
    import csv

    def rows(tag):
        d = os.path.join(root, f"pmc_{tag}")
        f = [x for x in os.listdir(d) if x.endswith("counter_collection.csv")][0]
        return list(csv.DictReader(open(os.path.join(d, f))))

    copy_bytes = 1 << 30
    cf = [float(r["Counter_Value"]) for r in rows("copy_fetch") if "copyBuffer" in r["Kernel_Name"]][0]
    cw = [float(r["Counter_Value"]) for r in rows("copy_write") if "copyBuffer" in r["Kernel_Name"]][0]
    f_fetch, f_write = copy_bytes / (cf * 1024), copy_bytes / (cw * 1024)
    fetch, write = rows(f"{precision}_fetch"), rows(f"{precision}_write")
    rec = {"_calibration": {"copy_bytes": copy_bytes, "fetch_KiB": cf, "write_KiB": cw,
                            "fetch_factor": round(f_fetch, 4), "write_factor": round(f_write, 4)}}
    lines = []
    for fam, knames in TRAFFIC_KERNELS.items():
        fb = wb = 0.0
        parts = []
        for kname in knames:
            fv = [float(r["Counter_Value"]) for r in fetch if kname in r["Kernel_Name"]]
            wv = [float(r["Counter_Value"]) for r in write if kname in r["Kernel_Name"]]
            if not fv or not wv:
                continue
            n = min(len(fv), steps)            # one launch per step for these kernels
            kf, kw = np.mean(fv[-n:]) * 1024 * f_fetch, np.mean(wv[-n:]) * 1024 * f_write
            fb, wb = fb + kf, wb + kw
            parts.append({"kernel": kname, "launches": n, "fetch_bytes": int(kf), "write_bytes": int(kw)})
            lines.append(f"| {fam} | `{kname}` | {n} | {kf / 1e9:.3f} | {kw / 1e9:.3f} | {(kf + kw) / 1e9:.3f} |")
        if parts:
            rec[fam] = {"kernels": parts, "fetch_bytes": int(fb), "write_bytes": int(wb),
                        "bytes_per_launch": int(fb + wb)}
    os.makedirs(out_dir, exist_ok=True)
    path = os.path.join(out_dir, "traffic.json")
    allrec = {}
    if os.path.exists(path):
        allrec = json.load(open(path))
    allrec[traffic_key(workload, precision)] = rec
    json.dump(allrec, open(path, "w"), indent=1)
    return rec, lines


# MI355X: 256 CUs of 4 SIMDs in 8 XCDs; rocprofv3 sums GRBM_GUI_ACTIVE over the XCDs
N_SIMD, N_XCD = 1024, 8


def summarize_mfma(root, tag):
    """MFMA counters per GEMM kernel instantiation from a rocprofv3 --pmc run
    (pmc_<tag>/: SQ_INSTS_VALU_MFMA_MOPS_F64/_F32/_BF16 -- units of 512 FLOPs per
    the counter definition -- SQ_VALU_MFMA_BUSY_CYCLES, SQ_BUSY_CYCLES,
    GRBM_GUI_ACTIVE), mean per dispatch, and the MFMA busy fraction
    SQ_VALU_MFMA_BUSY_CYCLES / (SIMDs x GPU-active cycles of one XCD) (the
    MfmaUtil expression of the counter list).  Returns ({kernel: {...}}, lines)."""
    import csv
    import re
    d = os.path.join(root, f"pmc_{tag}")
    f = [x for x in os.listdir(d) if x.endswith("counter_collection.csv")][0]
    per = {}
    for r in csv.DictReader(open(os.path.join(d, f))):
        k = r["Kernel_Name"]
        if "k_fb_gemm" not in k:
            continue
        m = re.search(r"(k_fb_gemm\w*<[^>]*>|k_fb_gemm\w*)", k)
        key = m.group(1) if m else k[:60]
        per.setdefault(key, {}).setdefault(r["Counter_Name"], []).append(float(r["Counter_Value"]))
    out, lines = {}, []
    for k, cs in per.items():
        rec = {c: float(np.mean(v)) for c, v in cs.items()}
        rec["dispatches"] = max(len(v) for v in cs.values())
        act = rec.get("GRBM_GUI_ACTIVE", 0.0) / N_XCD
        rec["mfma_busy"] = rec.get("SQ_VALU_MFMA_BUSY_CYCLES", 0.0) / (N_SIMD * act) if act else None
        gflop = {t: rec.get(f"SQ_INSTS_VALU_MFMA_MOPS_{t}", 0.0) * 512 / 1e9 for t in ("F64", "F32", "BF16")}
        rec["issued_gflop"] = gflop
        out[k] = rec
        kind = max(gflop, key=gflop.get)
        busy = "-" if rec["mfma_busy"] is None else f"{rec['mfma_busy']:.3f}"
        lines.append(f"| `{k}` | {rec['dispatches']} | {kind} | {gflop[kind]:.1f} | {act / 1e6:.3f} | {busy} |")
    return out, lines


def summarize_trace(root, steps, step_kernel="k_psatd_spchrg"):
    """Kernel time over the last `steps` steps of a rocprofv3 --kernel-trace run
    (root/*kernel_trace.csv): the window opens at the end of the (steps+1)-th
    last launch of step_kernel (one launch per step).  Returns markdown lines
    (kernel, calls, ms per step, mean us per call, share)."""
    import csv
    import re
    f = [x for x in os.listdir(root) if x.endswith("kernel_trace.csv")][0]
    rows = list(csv.DictReader(open(os.path.join(root, f))))
    marks = sorted(int(r["End_Timestamp"]) for r in rows if step_kernel in r["Kernel_Name"])
    if len(marks) < steps + 1:
        raise SystemExit(f"{root}: {len(marks)} launches of {step_kernel}, need {steps + 1}")
    t0, t1 = marks[-steps - 1], marks[-1]
    agg = {}
    for r in rows:
        a, b = int(r["Start_Timestamp"]), int(r["End_Timestamp"])
        if a < t0 or b > t1:
            continue
        name = r["Kernel_Name"].replace("void ", "").replace("chimera::", "")
        m = re.match(r"([\w:]+(<[^()]*>)?)", name)
        key = m.group(1) if m else name[:60]
        n, t = agg.get(key, (0, 0))
        agg[key] = (n + 1, t + (b - a))
    tot = sum(t for _, t in agg.values())
    lines = [f"| `{k}` | {n} | {t / 1e6 / steps:.3f} | {t / n / 1e3:.1f} | {100 * t / tot:.1f} |"
             for k, (n, t) in sorted(agg.items(), key=lambda kv: -kv[1][1])]
    lines.append(f"| all kernels | {sum(n for n, _ in agg.values())} | {tot / 1e6 / steps:.3f} | | 100 |")
    return lines


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="lpa_cfg2")
    ap.add_argument("--secondary", default="lpa_cfg1", help="comma list of further workloads, each an object "
                                                           "of the line ('' for none)")
    ap.add_argument("--advance", type=int, default=-1, help="untimed steps before the warmup (default: the "
                                                            "timed steps are the last of the workload's run)")
    ap.add_argument("--seed", type=int, default=20261202)
    ap.add_argument("--precision", choices=("both", "fp64", "mixed"), default="both")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--autotune", type=int, default=1, help="1: GEMM tile shapes chosen by timing (first use)")
    ap.add_argument("--sort-every", type=int, default=20, help="sort the electrons by cell every N steps (on the "
                    "window moves of the LPA workloads, where the sort also does the absorption); 0: never")
    ap.add_argument("--cpu-steps", type=int, default=1)
    ap.add_argument("--cpu-slab", type=int, default=8, help="the CPU baseline runs 1/N of the workload's x extent")
    ap.add_argument("--bunch-stages", type=int, nargs=2, metavar=("NX", "NKR"),
                    help="time the FB stages and the step of one bunch on an NX x NKR grid (--precision fp64|mixed)")
    ap.add_argument("--copy-calibration", type=float, default=0.0, metavar="GIB",
                    help="only copy GIB gibibytes device to device once (reference kernel for the HBM counters)")
    ap.add_argument("--no-scan", action="store_true", help="skip the batched bunch scan (scan object)")
    ap.add_argument("--scan-n", type=int, default=64)
    ap.add_argument("--scan-steps", type=int, default=300)
    ap.add_argument("--scan-chunk", type=int, default=4, help="steps issued per simulation per turn")
    ap.add_argument("--scan-gate", type=int, default=1, help="1: a simulation's next chunk waits for the last")
    ap.add_argument("--scan-only", default="", help="comma list of scan runs (plain_fp64, plain_mixed, "
                                                    "batched_fp64, batched_mixed) and no PIC-step bench")
    ap.add_argument("--serve", type=float, default=0.0, metavar="SECONDS",
                    help="only run the request loop for SECONDS of Poisson arrivals (serve object)")
    ap.add_argument("--serve-rate", type=float, default=20.0)
    ap.add_argument("--summarize-traffic", metavar="DIR",
                    help="write profiles/traffic.json from the four rocprofv3 --pmc runs under DIR")
    ap.add_argument("--summarize-mfma", metavar="DIR", help="print the MFMA counters of DIR/pmc_<precision>_mfma")
    ap.add_argument("--summarize-trace", metavar="DIR",
                    help="print the kernel time of the last --steps steps of the rocprofv3 kernel trace in DIR")
    args = ap.parse_args()
    if args.summarize_traffic:
        prec = "fp64" if args.precision == "both" else args.precision
        rec, lines = summarize_traffic(args.summarize_traffic, args.workload, args.steps,
                                       os.path.join(HERE, "profiles"), prec)
        print(json.dumps(rec, indent=1))
        print("\n".join(lines))
        return
    if args.summarize_mfma:
        prec = "fp64" if args.precision == "both" else args.precision
        rec, lines = summarize_mfma(args.summarize_mfma, f"{prec}_mfma")
        print(json.dumps(rec, indent=1))
        print("\n".join(lines))
        return
    if args.summarize_trace:
        print("\n".join(summarize_trace(args.summarize_trace, args.steps)))
        return
    if args.bunch_stages:
        bunch_stages(*args.bunch_stages, "mixed" if args.precision == "both" else args.precision)
        return
    if args.copy_calibration > 0:
        copy_calibration(args.copy_calibration)
        return
    if args.autotune:
        from chimera_amd._lib import lib as _l
        _l().chimera_set_autotune(1, None)          # tile shapes timed on first use (during the warmup)
    if args.gpus != 1:
        print(f"bench.py measures one GPU: the PIC step of one simulation runs on one MI355X and this build "
              f"has no multi-GPU path (asked for --gpus {args.gpus})", file=sys.stderr)
        sys.exit(2)
    if args.scan_only:
        print(json.dumps({"scan": scan_bench(args)}), flush=True)
        return
    if args.serve > 0:
        print(json.dumps({"serve": serve_bench(args)}), flush=True)
        return
    K = args.steps
    precs = ["fp64", "mixed"] if args.precision == "both" else [args.precision]
    summaries = bench_workload(args, args.workload, precs)
    main_prec = precs[0]
    m = summaries[main_prec]
    out = {
        "metric": METRIC, "value": m["value"], "unit": "particle-steps/s", "n_gpus": 1, "steps": K,
        "warmup": args.warmup, "ms_per_step": m["ms_per_step"], "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64" if main_prec == "fp64" else "f32+c64 (f64 positions)",
        "data": "synthetic: seeded FixedCell 2x2x4 plasma (n=0.005 n_c) + Gaussian laser, random draws from NumPy",
        "config": {"workload": args.workload, "Nx": m["Nx"], "Nkr": m["Nkr"], "modes": m["M"],
                   "electrons": m["P"], "precision": main_prec, "moving_window": True,
                   "timed_steps": m["timed_steps"], "sort_every": args.sort_every},
        "fb_transform_ms_per_step": m["fb_transform_ms_per_step"],
        "spectral_operator_ms_per_step": m["spectral_operator_ms_per_step"],
        "kernel_ms_per_step": m["kernel_ms_per_step"],
        "roofline": m["roofline"],
        "families": m["families"],
    }
    if len(precs) == 2:
        out["mixed"] = mixed_object(summaries)
    if args.secondary:
        for wl in args.secondary.split(","):
            sec = bench_workload(args, wl, precs)
            o = {k: sec[main_prec][k] for k in ("value", "ms_per_step", "fb_transform_ms_per_step",
                                                "spectral_operator_ms_per_step", "kernel_ms_per_step", "roofline",
                                                "timed_steps", "Nx", "Nkr", "M", "P")}
            if len(precs) == 2:
                o["mixed"] = mixed_object(sec)
            out[wl] = o
    if not args.no_scan:
        out["scan"] = scan_bench(args)
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(args)
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out), flush=True)


def bench_workload(args, workload, precs):
    K = args.steps
    summaries = {}
    for prec in precs:
        wargs = argparse.Namespace(**vars(args))
        wargs.workload = workload
        res = run_gpu(wargs, prec)
        value = res["psteps"] / res["wall"]
        head, every = roofline(res, K)
        summaries[prec] = {
            "value": round(value, 1), "ms_per_step": round(res["wall"] / K * 1e3, 3),
            "fb_transform_ms_per_step": round((res["fams"]["dht"][0] + res["fams"]["fft"][0]) / K, 3),
            "spectral_operator_ms_per_step": round(res["fams"]["fb_ops"][0] / K, 3),
            "kernel_ms_per_step": {k: round(v[0] / K, 3) for k, v in res["fams"].items() if v[1]},
            "roofline": head, "families": every, "P": res["P"],
            "timed_steps": f"{res['first']}-{res['last']} of {res['run_steps']}",
            "Nx": res["Nx"], "Nkr": res["Nkr"], "M": res["M"]}
        del res["run"]
    return summaries


def mixed_object(summaries):
    mx = summaries["mixed"]
    o = {k: mx[k] for k in ("value", "ms_per_step", "fb_transform_ms_per_step", "spectral_operator_ms_per_step",
                            "kernel_ms_per_step", "roofline", "families", "timed_steps")}
    o["dtype"] = "f32 momenta/weights, c64 fields, split-BF16 MFMA (f32 accumulation); f64 positions and tables"
    o["speedup_vs_fp64"] = round(mx["value"] / summaries["fp64"]["value"], 3)
    return o

if __name__ == "__main__":
    main()
