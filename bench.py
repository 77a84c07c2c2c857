#!/usr/bin/env python3
"""Benchmark: RK3 substeps/s of the InterPartS substep on one MI355X.

A "step" here is one RK3 substep (src/main.f90:268-414): the fused RK stage,
the 3-pass IBM, the Poisson solve, the projection and the FP64 Newton-Euler /
collision update, on the whole domain resident in HBM.  The workload is the
256^3 channel with 1,000 spheres (D = 12h, density ratio 1.5, Re_bulk = 5000),
FP64 path, synthetic seeded inputs.

Prints one JSON line.  `value` is substeps/s with all inputs resident in HBM.
`roofline` is for the dominant kernel: its algorithmic HBM bytes per launch
divided by its average duration, measured here with HIP events on the
library's stream over the K timed substeps.  `cpu_baseline` is the FP64 CPU
oracle timed on this host on a bounded sample of the same workload.
"""
import argparse
import ctypes as C
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def thomas_d_rows(case):
    """Rows of the Thomas multipliers d stored and read back, summed over the (x-mode, |y-mode|)
    columns: rows 1 .. the last row where d changes.  Between the end rows d follows one map
    and reaches a fixed point; the solve sends the later rows' stores and loads to that row,
    which stays in the caches.  The forward recurrence of src/solver.f90:90-135 is repeated here
    with the same tables (src/initsolver.f90:17-56) to find the row per column."""
    import numpy as np
    nx, ny, nz = case.itot, case.jtot, case.ktot
    dxi, dyi, dzi = nx / case.lx, ny / case.ly, nz / case.lz
    f = np.arange(1, nx // 2, dtype=np.float64)
    lx = np.concatenate(([0.0], -4.0 * dxi * dxi * np.sin(f * np.pi / nx) ** 2, [-4.0 * dxi * dxi]))
    f = np.arange(1, ny // 2, dtype=np.float64)
    ly = np.concatenate(([0.0], -4.0 * dyi * dyi * np.sin(f * np.pi / ny) ** 2, [-4.0 * dyi * dyi]))
    xy = (lx[:, None] + ly[None, :]).ravel()
    a = c = dzi * dzi
    b = -(a + c)
    K = nz - 1
    dprev = c * (1.0 / ((b + a) + xy))            # row 1 (Neumann b)
    last = np.ones(xy.shape, dtype=np.int64)
    for k in range(2, K + 1):
        dn = c * (1.0 / ((b + xy) - a * dprev))
        last[dn != dprev] = k
        dprev = dn
    return float(last.sum())


def thomas_words(case, Sw):
    """FP64 words the Thomas solve moves: rows 1 .. K-48 (K = ktot-1) four times, the last 48
    (the kernel's register tail, when K > 48) and row ktot twice, plus the multipliers d up to
    the row where they stop changing, written then read (one word per x-mode, |y-mode| and
    row, shared by the y-modes +-ky)."""
    K = case.ktot - 1
    T = 48 if K > 48 else 0
    return Sw * (4.0 * (K - T) + 2.0 * T + 2.0) / case.ktot + 2.0 * thomas_d_rows(case)


def kernel_bytes_mixed(case, name, rkiter, nref, solve=32):
    """Algorithmic HBM bytes of one launch of a kernel on the mixed path: FP64
    fields (u v w c pnew dudt..dcdt, the right-hand side) at 8 B, the RK history
    at 4 B; with the FP32-stored solve (solve 32) the spectra and the pressure
    pieces at 4 B and the Thomas multipliers d at 8 B, with the FP64 solve
    (solve 64) every Poisson and projection launch as on the FP64 path."""
    N = float(case.itot * case.jtot * case.ktot)
    if solve == 64:
        if name == "rk_stage":
            return 72.0 * N + {1: 16.0, 2: 32.0, 3: 16.0}[rkiter] * N
        return kernel_bytes(case, name, rkiter)
    Sw = 2.0 * (case.itot // 2 + 1) * case.jtot * case.ktot
    d_rows = thomas_d_rows(case)
    b = {
        # u v w c pnew read, dudt..dcdt written (FP64); history 4 written / 4 read + 4 written / 4 read (FP32)
        "rk_stage": 72.0 * N + {1: 16.0, 2: 32.0, 3: 16.0}[rkiter] * N,
        "fft_x_fwd": 32.0 * N + 4.0 * Sw,      # dudt dvdt dwdt read, the right-hand side (FP64) and the spectrum stored
        # refinement r: the right-hand side and r pieces read, the FP32 residual written, then transformed
        "resid": (12.0 + 4.0 * max(1, nref)) * N if nref else None,
        "resid_x_fwd": 4.0 * N + 4.0 * Sw,
        "fft_y_fwd": 8.0 * Sw,
        "thomas": 4.0 * (thomas_words(case, Sw) - 2.0 * d_rows) + 16.0 * d_rows,
        "fft_y_inv": 8.0 * Sw,
        "fft_x_inv": 4.0 * Sw + 4.0 * N,       # the spectrum read, one FP32 piece of p written
        "correc": (64.0 + 4.0 * (nref + 1)) * N,  # dudt dvdt dwdt pnew and the pieces read; u v w pnew written
    }
    return b.get(name)


def mixed_dtype(solve):
    """What the mixed path stores and computes in FP32, by its Poisson solve."""
    if solve == 64:
        return ("mixed: RK history stored f32; u v w p dudt, the Poisson solve, the IBM and the particles f64, "
                "stencil arithmetic f64")
    return ("mixed: RK history, Poisson spectra and the pressure correction's pieces stored f32 (one FP64-residual "
            "refinement); u v w dudt, the IBM and the particles f64, stencil and Poisson arithmetic f64")


def kernel_bytes(case, name, rkiter):
    """Algorithmic HBM bytes of one launch of a kernel (FP64 words x 8).
    N cells; Sw = FP64 words of the half spectrum, 2 (nx/2+1) ny nz."""
    N = float(case.itot * case.jtot * case.ktot)
    Sw = 2.0 * (case.itot // 2 + 1) * case.jtot * case.ktot
    words = {
        "rk_stage": {1: 13.0, 2: 17.0, 3: 13.0}[rkiter] * N,  # reads u v w c p (+4 rk); writes 4 d?dt (+4 rk)
        "fillps": 4.0 * N,
        "fft_x_fwd": 3.0 * N + Sw,    # fillps fused: reads dudt dvdt dwdt, writes the half spectrum
        "fft_y_fwd": 2.0 * Sw,
        # spectrum read + written in each sweep, except the last 48 rows of each column (kept in
        # registers between the sweeps: read once, written once) and the last row; d (one word per
        # mode and plane, shared by the y-modes +-ky) written then read up to its fixed point
        "thomas": thomas_words(case, Sw),
        "fft_y_inv": 2.0 * Sw,
        "fft_x_inv": Sw + N,
        "correc": 9.0 * N,            # dudt dvdt dwdt p pnew -> u v w pnew (cnew = dcdt is a buffer swap)
    }
    w = words.get(name)
    return None if w is None else 8.0 * w


def traffic_from_profile(kernel, workload):
    """HBM bytes per launch of `kernel` measured by the rocprofv3 FETCH_SIZE and
    WRITE_SIZE passes summarised in profiles/hbm_traffic_<kernel>_<workload>.txt
    (null when no such pass exists for this workload)."""
    path = os.path.join(ROOT, "profiles", f"hbm_traffic_{kernel}_{workload}.txt")
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith("traffic_bytes_per_launch:"):
            return float(line.split(":")[1])
    return None


def copy_probe(mib, reps=10):
    """A plain copy kernel of known size: reads and writes mib MiB per launch,
    timed with HIP events.  Runs in a process of its own: torch carries its own
    HIP runtime, which must not share a process with libipg.so's."""
    import torch
    n = mib * (1 << 20) // 8
    x = torch.ones(n, dtype=torch.float64, device="cuda")
    y = torch.empty_like(x)
    y.copy_(x)
    a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    a.record()
    for _ in range(reps):
        y.copy_(x)
    b.record()
    torch.cuda.synchronize()
    ms = a.elapsed_time(b) / reps
    print(json.dumps({"copy_probe_bytes_read": n * 8, "copy_probe_bytes_written": n * 8, "launches": reps + 1,
                      "ms_per_copy": ms, "copy_gbs": 2.0 * n * 8 / (ms * 1e-3) / 1e9}))


def copy_bandwidth(mib=1024):
    """Measured read+write bandwidth of a plain device copy (GB/s), from a
    child process, or None when it cannot run."""
    import subprocess
    r = subprocess.run([sys.executable, os.path.abspath(__file__), "--copy-probe", str(mib)], stdout=subprocess.PIPE,
                       stderr=subprocess.PIPE, text=True, timeout=300)
    if r.returncode != 0:
        print(f"[bench] copy probe failed: {r.stderr[-500:]}", file=sys.stderr)
        return None
    return json.loads(r.stdout.strip().splitlines()[-1])["copy_gbs"]


def run_ensemble(args):
    """Batched ensemble.  Three ways to advance the same cases, timed one after
    another over the same number of substeps:
    - fused: one batch (ipg_batch_*), every kernel of a substep stage merged
      over all cases, one captured graph per substep, launch shapes chosen by
      the batch's tuner (the headline);
    - concurrent: each case its own context, stream and captured graphs, one
      round launching every case's substep before waiting;
    - sequential: the cases one after another."""
    from interparts_mpich_amd.inputs import ensemble_cases, initial_fields, place_spheres, markers_for
    from interparts_mpich_amd.driver import Simulation, Batch
    cases = ensemble_cases(args.ensemble)
    precision = 1 if args.precision == "mixed" else 0
    sims = []
    for c in cases:
        # many contexts overlap one another's streams: no side stream each
        sims.append(Simulation(c, use_graphs=True, side_stream=args.side_stream == "on", precision=precision,
                               solve=args.mixed_solve if precision else None)
                    .initialize(initial_fields(c), place_spheres(c), markers_for(c)))
    print(f"[bench] ensemble of {len(sims)} cases ready", file=sys.stderr, flush=True)
    steps = max(3, args.steps // 3 * 3)
    warm = max(3, args.warmup // 3 * 3)
    for sim in sims:
        for s in range(warm):
            sim.substep(s % 3 + 1)
        sim.sync()
    t0 = time.perf_counter()
    for sim in sims:
        for s in range(steps):
            sim.substep(s % 3 + 1)
        sim.sync()
    t_seq = time.perf_counter() - t0
    print(f"[bench] sequential: {len(sims) * steps / t_seq:.1f} case-substeps/s", file=sys.stderr, flush=True)
    t0 = time.perf_counter()
    for s in range(steps):
        for sim in sims:
            sim.substep(s % 3 + 1)
    for sim in sims:
        sim.sync()
    t_conc = time.perf_counter() - t0
    print(f"[bench] concurrent contexts: {len(sims) * steps / t_conc:.1f} case-substeps/s", file=sys.stderr, flush=True)
    batch = Batch(sims)
    tuned = 0 if args.no_autotune else batch.autotune()
    for s in range(warm):
        batch.substep(s % 3 + 1)
    batch.sync()
    t0 = time.perf_counter()
    for s in range(steps):
        batch.substep(s % 3 + 1)
    batch.sync()
    t_fused = time.perf_counter() - t0
    nl, npers = batch.launches(1)
    ngroups, carry = batch.groups(), batch.carry()
    phases = {}
    for rk in (1, 2, 3):
        for name, ms in batch.phases(rk):
            key = name.split("#")[0]
            phases[key] = phases.get(key, 0.0) + ms / 3.0
    batch.sync()
    print(f"[bench] fused batch: {len(sims) * steps / t_fused:.1f} case-substeps/s, {nl} merged launches per substep "
          f"({npers} persistent) in {ngroups} stream groups, updates carried beside the RK stage: {carry}",
          file=sys.stderr, flush=True)
    print("[bench] fused batch per-substep stage time (ms, eager): " +
          ", ".join(f"{k}={v:.3f}" for k, v in sorted(phases.items(), key=lambda kv: -kv[1])), file=sys.stderr, flush=True)
    n = len(sims) * steps
    cells = sum(c.cells for c in cases) * steps
    out = {
        "metric": f"ensemble case-substeps/s ({len(sims)} cases) " + ("mixed precision" if precision else "FP64"),
        "value": round(n / t_fused, 2),
        "unit": "case-substeps/s", "n_gpus": 1, "steps": steps, "warmup": warm,
        "ms_per_step": round(t_fused * 1e3 / steps, 4), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "mixed" if precision else "f64", "data": "synthetic (seeded cases)",
        "config": {"workload": f"{len(sims)} cases, grids 64^3/96x96x64/128^3/128x128x192, 8-200 spheres, "
                               "D 8-12h, Re_bulk 1000-5000",
                   "gcell_updates_per_s": round(cells / t_fused / 1e9, 4),
                   "merged_launches_per_substep": nl, "persistent_launches": npers, "stream_groups": ngroups,
                   "updates_beside_rk_stage": carry,
                   "tuner_substeps": tuned},
        "concurrent_contexts_case_substeps_per_s": round(n / t_conc, 2),
        "sequential_case_substeps_per_s": round(n / t_seq, 2),
        "speedup_vs_sequential": round(t_seq / t_fused, 3),
        "speedup_vs_concurrent_contexts": round(t_conc / t_fused, 3),
        "fused_stage_ms_per_substep": {k: round(v, 4) for k, v in phases.items()},
    }
    batch.close()
    for sim in sims:
        sim.close()
    print(json.dumps(out))


def run_profile(args):
    """The workload under a profiler: no timing of its own, no graphs unless
    asked for, a few substeps, a line per phase so the run is never silent."""
    from interparts_mpich_amd.inputs import CONFIGS, initial_fields, place_spheres, markers_for
    from interparts_mpich_amd.driver import Simulation
    case = CONFIGS[args.config]
    precision = 1 if args.precision == "mixed" else 0
    print(f"[profile] {case.name} {'mixed' if precision else 'fp64'}: inputs", flush=True)
    inputs = (initial_fields(case), place_spheres(case), markers_for(case))
    sim = Simulation(case, use_graphs=not args.no_graphs, side_stream=args.side_stream == "on",
                     precision=precision, solve=args.mixed_solve if precision else None).initialize(*inputs)
    print("[profile] initialised", flush=True)
    for s in range(args.warmup + args.steps):
        sim.substep(s % 3 + 1)
    sim.sync()
    print(f"[profile] {args.warmup + args.steps} substeps done", flush=True)
    if args.checks:
        # the dt and divergence checks of every ioutchk-th step, repeated
        for name, fn in (("chkdt", sim.chkdt), ("chkdiv", sim.chkdiv)):
            t0 = time.perf_counter()
            for _ in range(args.checks):
                fn()
            print(f"[profile] {name}: {(time.perf_counter() - t0) / args.checks * 1e3:.3f} ms per call "
                  f"(with its synchronisation and copy)", flush=True)
    sim.close()


def measure(case, inputs, args, precision, solve=None):
    """Substeps/s of one path over K timed substeps (after W warmup substeps),
    then per-kernel durations with HIP events over K eager substeps.  solve
    (mixed path): 64 the FP64 Poisson solve, 32 the FP32-stored solve with
    FP64-residual refinement."""
    from interparts_mpich_amd.driver import Simulation
    from interparts_mpich_amd import _lib
    import numpy as np
    fields, spheres, markers = inputs
    name = ("mixed" + ("_fp64_solve" if solve == 64 else "")) if precision else "fp64"
    sim = Simulation(case, use_graphs=not args.no_graphs, side_stream=args.side_stream == "on",
                     precision=precision, solve=solve if precision else None).initialize(fields, spheres, markers)
    rk_kb = None if args.no_autotune else sim.autotune()
    sph_blocks = sim.tuning()[1]
    for s in range(args.warmup):
        sim.substep(s % 3 + 1)
    sim.sync()
    t0 = time.perf_counter()
    for s in range(args.steps):
        sim.substep(s % 3 + 1)
    sim.sync()
    t1 = time.perf_counter()
    res = {"value": args.steps / (t1 - t0), "ms_per_step": (t1 - t0) * 1e3 / args.steps,
           "rk_kb": rk_kb if rk_kb is not None else "default", "nref": None,
           "sphere_int_blocks": sph_blocks if sph_blocks > 0 else "one per sphere"}
    print(f"[bench] {name}: {args.steps} substeps in {t1 - t0:.4f} s ({res['value']:.1f} substeps/s)", file=sys.stderr,
          flush=True)
    if precision:
        res["solve"] = solve
        res["nref"] = int(os.environ.get("IPG_MIXED_REFINE", "1")) if solve == 32 else None
        _, mdiv = sim.chkdiv()
        res["accuracy"] = {"max_div_after_timed_substeps": mdiv}

    # per-kernel durations over the same number of substeps, run as the timed
    # ones were (graph replays beside the carried particle update)
    lib = _lib.load()
    maxn = 64
    ms = (C.c_double * maxn)()
    names = (C.c_char_p * maxn)()
    n = C.c_int()
    acc = {}
    for s in range(args.steps):
        rk = s % 3 + 1
        _lib.check(lib.ipg_substep_phases(sim.ctx, rk, maxn, ms, names, C.byref(n)), "substep_phases")
        for q in range(n.value):
            key = names[q].decode()
            a = acc.setdefault(key, {"ms": 0.0, "launches": 0, "bytes": 0.0})
            a["ms"] += ms[q]
            a["launches"] += 1
            b = kernel_bytes_mixed(case, key, rk, res["nref"], solve) if precision else kernel_bytes(case, key, rk)
            if b is not None:
                a["bytes"] += b
    # the dominant kernel again with no update beside it (the substep's own
    # stream only), so its duration alone is reported beside the overlapped one
    alone = None
    if args.side_stream == "on" and case.np > 0:
        _lib.check(lib.ipg_use_side_stream(sim.ctx, 0), "use_side_stream")
        alone = {}
        nalone = min(args.steps, 60)
        for s in range(nalone):
            rk = s % 3 + 1
            _lib.check(lib.ipg_substep_phases(sim.ctx, rk, maxn, ms, names, C.byref(n)), "substep_phases")
            for q in range(n.value):
                key = names[q].decode()
                b = kernel_bytes_mixed(case, key, rk, res["nref"], solve) if precision else kernel_bytes(case, key, rk)
                if b is not None:
                    a = alone.setdefault(key, [0.0, 0, 0.0])
                    a[0] += ms[q]
                    a[1] += 1
                    a[2] += b
    sim.close()
    per_substep = {k: v["ms"] / args.steps for k, v in acc.items()}
    kernel_gbs = {k: round(v["bytes"] / (v["ms"] * 1e-3) / 1e9, 1) for k, v in acc.items() if v["bytes"] > 0}
    dominant = max((k for k in acc if acc[k]["bytes"] > 0), key=lambda k: acc[k]["ms"])
    d = acc[dominant]
    avg_ms = d["ms"] / d["launches"]
    bytes_per_launch = d["bytes"] / d["launches"]
    achieved = bytes_per_launch / (avg_ms * 1e-3) / 1e9
    print(f"[bench] {name} per-substep kernel time (ms): " +
          ", ".join(f"{k}={v:.4f}" for k, v in sorted(per_substep.items(), key=lambda kv: -kv[1])),
          file=sys.stderr, flush=True)
    res["roofline"] = {"bound": "hbm", "kernel": dominant, "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                       "timed_as": ("graph replays of the timed substeps, wall-clock marks on the kernel's stream"
                                    if not args.no_graphs else "eager launches, HIP events on the kernel's stream"),
                       "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                       "traffic": traffic_from_profile(dominant + ("_mixed" if precision else ""), case.name),
                       "avg_launch_ms": round(avg_ms, 5), "bytes_per_launch": bytes_per_launch}
    if alone and dominant in alone:
        ams, alaunch, abytes = alone[dominant]
        agbs = abytes / alaunch / (ams / alaunch * 1e-3) / 1e9
        res["roofline"]["without_side_stream"] = {
            "avg_launch_ms": round(ams / alaunch, 5), "achieved": round(agbs, 1),
            "frac": round(agbs / HBM_PEAK_GBS, 4),
            "note": "the same kernel over further substeps with the Newton-Euler update in its own substep "
                    "(no second stream), so it has the chip to itself"}
    res["kernel_ms_per_substep"] = {k: round(v, 5) for k, v in per_substep.items()}
    res["kernel_algorithmic_gbs"] = kernel_gbs
    res["event_sum_ms_per_substep"] = round(sum(v["ms"] for v in acc.values()) / args.steps, 4)
    return res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=300, help="timed substeps")
    ap.add_argument("--warmup", type=int, default=30, help="untimed substeps")
    ap.add_argument("--config", default="ch256_1000sph")
    ap.add_argument("--precision", choices=("fp64", "mixed", "both"), default="both",
                    help="path(s) to time: both prints the mixed path's line with the FP64 path's figures beside it "
                         "(with --ensemble: fp64 unless mixed)")
    ap.add_argument("--mixed-solve", type=int, choices=(32, 64), default=32,
                    help="the mixed path's Poisson solve: 32 FP32-stored spectra with FP64-residual refinement (the "
                         "library's default), 64 the FP64 path's solve (timed beside the others with --precision both)")
    ap.add_argument("--profile", action="store_true",
                    help="for a profiler: initialise the case, then run --warmup and --steps substeps of the path "
                         "--precision names (fp64 or mixed) and exit; prints one line per phase")
    ap.add_argument("--checks", type=int, default=0, metavar="N",
                    help="with --profile: then run chkdt and chkdiv N times each")
    ap.add_argument("--no-graphs", action="store_true")
    ap.add_argument("--no-autotune", action="store_true", help="keep the default RK-stage launch shape")
    ap.add_argument("--cpu-substeps", type=int, default=12, help="oracle substeps in the CPU-baseline sample")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="only run a plain device-to-device copy of MIB MiB (calibrates the HBM byte counters of "
                         "a profiler pass; measures the practical copy bandwidth) and exit")
    ap.add_argument("--no-copy-probe", action="store_true", help="skip the copy-bandwidth child process")
    ap.add_argument("--side-stream", choices=("on", "off"), default=None,
                    help="run each substep's Newton-Euler update beside the next RK stage on a second stream "
                         "(default: on for one case, off for an ensemble)")
    ap.add_argument("--ensemble", type=int, default=0, metavar="NCASES",
                    help="instead of the 256^3 workload, advance NCASES independent small cases (uneven grids and "
                         "sphere counts) together in one batch of merged launches and report aggregate "
                         "case-substeps/s against running them as concurrent contexts and one after another")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: every box here has one GPU and this build has no multi-GPU path "
              f"(--gpus {args.gpus} requested)", file=sys.stderr)
        sys.exit(2)

    if args.copy_probe:
        copy_probe(args.copy_probe)
        return
    if args.side_stream is None:
        args.side_stream = "off" if args.ensemble else "on"
    if args.ensemble:
        run_ensemble(args)
        return
    if args.profile:
        run_profile(args)
        return

    from interparts_mpich_amd.inputs import CONFIGS, initial_fields, place_spheres, markers_for

    case = CONFIGS[args.config]
    fields, spheres, markers = initial_fields(case), place_spheres(case), markers_for(case)
    print(f"[bench] {case.name}: {case.itot}x{case.jtot}x{case.ktot}, np={case.np}, nl={case.nl}", file=sys.stderr, flush=True)
    runs = {}
    for prec in (("fp64", "mixed", "mixed_fp64_solve") if args.precision == "both" else (args.precision,)):
        solve = None if prec == "fp64" else 64 if prec == "mixed_fp64_solve" else args.mixed_solve
        runs[prec] = measure(case, (fields, spheres, markers), args, 0 if prec == "fp64" else 1, solve)
    # the line leads with the FP64 path (the workload the metric is quoted
    # on) when it ran; the mixed path (its FP32-stored solve with refinement)
    # and the mixed path with the FP64 solve are reported beside it
    head = runs["fp64"] if "fp64" in runs else runs[args.precision]
    mixed = head is not runs.get("fp64")
    value = head["value"]
    out = {
        "metric": "RK3 substeps/s (256^3, 1k spheres) " + ("mixed precision" if mixed else "FP64"),
        "value": round(value, 3),
        "unit": "substeps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(head["ms_per_step"], 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": (mixed_dtype(head["solve"]) if mixed else "f64"),
        "data": "synthetic (seeded Poiseuille + 0.1 U_bulk noise, seeded sphere placement)",
        "config": {"workload": f"{case.name}: {case.itot}x{case.jtot}x{case.ktot} channel, {case.np} spheres "
                               f"D={case.diameter_h:g}h nl={case.nl}, Re_bulk={case.re_bulk:g}, "
                               + ("mixed-precision path" if mixed else "FP64 path"),
                   "graphs": not args.no_graphs, "side_stream": args.side_stream == "on",
                   "rk_z_planes_per_workgroup": head["rk_kb"],
                   "sphere_integral_blocks": head["sphere_int_blocks"],
                   "gcell_updates_per_s": round(value * case.cells / 1e9, 4)},
        "roofline": head["roofline"],
        "kernel_ms_per_substep": head["kernel_ms_per_substep"],
        "kernel_algorithmic_gbs": head["kernel_algorithmic_gbs"],
        "event_sum_ms_per_substep": head["event_sum_ms_per_substep"],
    }
    if mixed:
        out["config"]["poisson_solve"] = head["solve"]
        out["config"]["poisson_refinement_sweeps"] = head["nref"]
    for key, o in runs.items():
        if o is head:
            continue
        out[key] = {k: o[k] for k in ("value", "ms_per_step", "solve", "nref", "roofline", "kernel_ms_per_substep",
                                      "kernel_algorithmic_gbs", "event_sum_ms_per_substep") if k in o}
        if "solve" in o:
            out[key]["dtype"] = mixed_dtype(o["solve"])
        if "accuracy" in o:
            out[key]["accuracy"] = o["accuracy"]
    if "fp64" in runs:
        for key in ("mixed", "mixed_fp64_solve"):
            if key in runs:
                out[key + "_over_fp64"] = round(runs[key]["value"] / runs["fp64"]["value"], 4)
    if "accuracy" in head:
        out["mixed_accuracy"] = head["accuracy"]
    if not args.no_copy_probe:
        copy_gbs = copy_bandwidth()
        for r in runs.values():
            r["roofline"]["copy_gbs_measured"] = round(copy_gbs, 1) if copy_gbs else None

    if not args.no_cpu_baseline:
        from oracle.pyoracle import Oracle
        orc = Oracle(case).initialize(fields, spheres, markers)
        nsub = args.cpu_substeps
        t0 = time.perf_counter()
        for s in range(nsub):
            orc.substep(s % 3 + 1)
        t1 = time.perf_counter()
        orc.close()
        cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
        out["cpu_baseline"] = {"value": round(nsub / (t1 - t0), 5), "unit": "substeps/s", "cores": cores,
                               "kind": "port",
                               "sample": f"{nsub} FP64 substeps of {case.name} by the C oracle (OpenMP over "
                                         f"field loops, {cores} threads), {t1 - t0:.1f} s"}
    print(json.dumps(out))


if __name__ == "__main__":
    main()
