"""Benchmark: converged patches/s of the MOMFBD patch solver on one GPU.

Workload (default): 256 overlapping 88x88 patches from a seeded 1024x1024 k^-3
field, 2 objects (wideband + narrowband with a 1.0-wave peak-to-valley focus
diversity) x 24 frames sharing wavefronts, 50 KL modes (Kolmogorov, D/r0 = 20),
Poisson noise at 2000 mean counts, pupil 64 px on a 128 x 128 OTF grid; the
reference's solver defaults (FTOL 1e-3, conjugate gradient, MODE_START/STEP 5).

One step = solving every patch of the batch to convergence (Solver::run per
patch: mode stepping, CG + Brent, re-cuts, restoration) with all patches advancing in
lockstep; the batch's metric/gradient evaluations go to the device in alternating
lanes (contiguous shares of the batch) so the host steps one share while the device
evaluates the others.

The timed form is the batched FP32 form (FP32 phase GEMM and transforms, FP64 P/Q,
metric and CG).  A patch counts as converged when its solve exits normally and its
restored objects are within 1e-4 (relative RMS) of the FP64 oracle's solve of the same
patch; the oracle solves a fixed sample of the batch (its timing is the CPU baseline),
and `value` is the FP32 rate times the sample's converged fraction.  Beside it: the
plain FP64 drop-in (one patch at a time, on the sample patches, in the reference's
arithmetic: the mode GEMM per line-search probe, the line search stepped by the host), the
batched FP64 form (line-search probes from base + step * direction, the line search on
the device, as the FP32 form), the reference's arithmetic batched (bit-identical to the
plain drop-in) and the mixed (BF16 mode GEMM) form, each with the same check against the
oracle.

Each form gets a context with its patches loaded (frames resident in HBM); a step is
one mfbd_solve of every loaded patch from zero alphas, so warmup and timed steps repeat
the whole computation on the same context.  Per-kernel HIP-event timings (and the
rooflines) come from one step per form on a separate context that brackets every kernel
with events; the timed steps run without them.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time
from concurrent.futures import ProcessPoolExecutor

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
# Vector (non-matrix) peak for FP64, and for FP32 without packed-FP32 instructions
# (which no kernel here uses): 256 CUs x 64 lanes x 2 flops (FMA) x 2.4 GHz
VECTOR_PEAK_TFLOPS = 78.6


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def _make_patch(args):
    k, seed, cfg, field_size = args
    from redux_amd import synth
    su, fields = _STATE["su"], _STATE["fields"]
    N, s, margin = su["N"], su["s"], cfg.get("margin", 5)
    G = N if N >= s + 2 * margin else int(2 ** math.ceil(math.log2(s + 2 * margin)))
    side = _STATE["side"]
    stride = (field_size - G) // max(side - 1, 1)
    iy, ix = divmod(k, side)
    return synth.build_patch(su, seed * 1000 + k, field=fields, field_offset=(iy * stride, ix * stride))


_STATE = {}


def make_batch(n, seed, workers):
    from redux_amd import synth
    cfg = synth.CFG2_PATCH
    su = synth.build_setup(cfg)
    field_size = 1024
    _STATE.update(su=su, side=int(math.ceil(math.sqrt(n))),
                  fields=[synth.grf_field(field_size, seed + 17 * o, mean=cfg["mean_counts"])
                          for o in range(cfg["n_objects"])])
    jobs = [(k, seed, cfg, field_size) for k in range(n)]
    if workers > 1:
        with ProcessPoolExecutor(max_workers=workers) as ex:
            pbs = list(ex.map(_make_patch, jobs, chunksize=4))
    else:
        pbs = [_make_patch(j) for j in jobs]
    return su, pbs


LINE_GRAPH = True        # device line-search rounds replayed as captured HIP graphs (--line-graph)


def make_solver(pbs, precision, batch, gradient, timing=False, line=None):
    """A context with the patches loaded (frames resident in HBM).  With `timing`,
    every kernel launch is bracketed by HIP events on the pipeline's stream (and the
    line-search rounds are launched kernel by kernel).  `line`: True for line-search probes
    from base + step * direction with the line search on the device, False for the mode
    GEMM per probe stepped by the host (the reference's arithmetic), None for the Solver's
    default of the precision form (FP32 / mixed: the former, FP64: the latter)."""
    from redux_amd.momfbd import Solver
    if timing:
        os.environ["MFBD_KERNEL_TIMING"] = "1"
    lp = {} if line is None else dict(line_phases=bool(line), line_rounds=4 if line else 0)
    try:
        sv = Solver(pbs[0], precision=precision, max_batch=batch, gradient=gradient, line_graph=LINE_GRAPH, **lp)
        sv.load(pbs)
    finally:
        os.environ.pop("MFBD_KERNEL_TIMING", None)
    return sv


def run_step(sv, restore=True):
    """One step: solve every loaded patch from zero alphas (mode stepping, CG + Brent,
    re-cuts) and restore the objects.  Each call repeats the whole computation."""
    t0 = time.perf_counter()
    res = sv.run(restore=restore)
    return res, time.perf_counter() - t0


TRAFFIC_PROFILES = {"fp32": "k_otf_wave_cfg2_traffic.md", "fp64": "k_otf64_cfg2_traffic.md"}


def measured_traffic(form):
    """HBM bytes per image of the hot kernels of one precision form, measured with
    rocprofv3 FETCH_SIZE / WRITE_SIZE (scaled by a plain copy of known size) and
    committed under profiles/."""
    path = os.path.join(ROOT, "profiles", TRAFFIC_PROFILES[form])
    try:
        with open(path) as f:
            for line in f:
                if line.startswith("traffic-json:"):
                    return json.loads(line.split(":", 1)[1])
    except OSError:
        pass
    return {}


SQ_PROFILE = "otf_cfg2_sq_counters.md"
CLOCK_GHZ = 2.4          # MI355X peak engine clock
SIMDS = 1024             # 256 CUs x 4 SIMDs


def measured_valu(name):
    """VALU wave-instructions per image of a transform kernel, counted with rocprofv3
    (SQ_INSTS_VALU) and committed under profiles/."""
    try:
        with open(os.path.join(ROOT, "profiles", SQ_PROFILE)) as f:
            for line in f:
                if line.startswith("sq-json:"):
                    return json.loads(line.split(":", 1)[1]).get(name)
    except OSError:
        pass
    return None


def roofline(stats, N, nPS, nH, form):
    """Dominant kernel by total time; achieved = algorithmic bytes per launch / mean
    launch time, over every launch of one step on a context with kernel events.  S, D, P and
    Q live on the stored (Hermitian) half of the OTF support, nH points; c and r are the
    complex and real element sizes of the form (16 and 8 for FP64, 8 and 4 for FP32).
    Algorithmic bytes per image:
      k_otf_fused, k_otf_wave (FP32): read phi, write S           r nPS + c nH
        (a line-search probe reads the line's base and direction:  2 r nPS + c nH)
      k_otf64_fused (FP64): read phi, write S                     r nPS + c nH
        (a line-search probe reads the line's base and direction:  2 r nPS + c nH)
      k_pq: read S and D                                             2 c nH
      k_phase_mfma: write phi                                        r nPS
      k_vogel_fused: read S, D, P, Q (FP64), phi; write v            2 c nH + 24 nH + 2 r nPS
      k_grad64_fused: read S, D, phi, P and Q (FP64, shared by the
        object's images); write and read back gl (N x N real); write v
                                                                     2 c nH + 24 nH + r nPS + 2 r N^2 + r nPS
      unfused FFT pass: read and write the N x N tile                2 c N^2
    For the fused OTF kernels the FFT arithmetic is reported beside it: two N x N complex
    transforms, 5 N^2 log2(N^2) flops each, against the vector peak (78.6 TFLOP/s for
    FP64, and for FP32 without packed-FP32 instructions, which these kernels do not use;
    the 157.3 TFLOP/s FP32 figure counts packed FP32)."""
    c, r = (16, 8) if form == "fp64" else (8, 4)
    half = N * N // 2
    per_unit = {"fft_rows": 2 * c * N * N, "fft_cols": 2 * c * N * N, "k_field": r * nPS + c * N * N,
                "k_power": 2 * c * N * N, "k_extract": 2 * c * nH, "k_phase_mfma": r * nPS, "k_pq": 2 * c * nH,
                "k_otf_fused": r * nPS + c * nH, "k_otf_wave": r * nPS + c * nH, "k_otf64_fused": r * nPS + c * nH,
                "k_vogel_fused": 2 * c * nH + 24 * nH + 2 * r * nPS,
                "k_grad64_fused": 2 * c * nH + 24 * nH + 2 * r * nPS + 2 * r * N * N}
    kern = {k: v for k, v in stats.items() if not k.startswith("host") and not k.startswith("_")}
    name, s = max(kern.items(), key=lambda kv: kv[1]["ms"])
    out = dict(kernel=name)
    if name in per_unit and s["units"] > 0:
        mean_ms = s["ms"] / s["launches"]
        units_per_launch = s["units"] / s["launches"]
        # a line-search probe's transform reads the line's base and direction (2 r nPS) for phi
        line = stats.get("_line_images", {}).get("units", 0) if name in ("k_otf_fused", "k_otf_wave",
                                                                          "k_otf64_fused") else 0
        bytes_per_launch = (per_unit[name] * s["units"] + r * nPS * line) / s["launches"]
        ach = bytes_per_launch / (mean_ms * 1e-3) / 1e9
        tr = measured_traffic(form).get(name)
        out.update(bound="hbm", achieved=round(ach, 1), peak=HBM_PEAK_GBS, unit="GB/s",
                   frac=round(ach / HBM_PEAK_GBS, 4),
                   traffic=None if tr is None else int(tr * units_per_launch), launches=s["launches"],
                   mean_launch_ms=round(mean_ms, 4), images_per_launch=round(units_per_launch, 1),
                   bytes_per_image=round(bytes_per_launch / units_per_launch, 1), bytes_per_launch=int(bytes_per_launch),
                   line_probe_images=int(line))
        if name in ("k_otf_fused", "k_otf_wave", "k_otf64_fused"):
            flops = 2 * 5 * N * N * math.log2(N * N) * units_per_launch
            tf = flops / (mean_ms * 1e-3) / 1e12
            out.update(fft_tflops=round(tf, 2), fft_peak_tflops=VECTOR_PEAK_TFLOPS,
                       fft_frac=round(tf / VECTOR_PEAK_TFLOPS, 4))
            sq = measured_valu(name)
            if sq:
                # share of the SIMDs' VALU issue cycles: a wave64 VALU instruction takes 4
                # cycles of a 16-lane SIMD; instructions per image counted by rocprofv3
                cyc = sq["valu_per_image"] * units_per_launch * 4
                out.update(valu_issue_frac=round(cyc / (mean_ms * 1e-3 * CLOCK_GHZ * 1e9 * SIMDS), 3),
                           valu_insts_per_image=sq["valu_per_image"])
    return out


COPY_PROBE_BYTES = 256 << 20


def pmc_probe(form):
    """Workload for the rocprofv3 --pmc runs behind `traffic`: three batched metric
    evaluations and three batched Vogel gradients of 32 config-2 patches in the given
    precision form (1536 images per OTF launch), then three device-to-device copies of
    COPY_PROBE_BYTES through the HIP runtime, the plain copy that calibrates FETCH_SIZE /
    WRITE_SIZE."""
    import ctypes as C
    from redux_amd import synth
    su, pbs = synth.build_field_batch(synth.CFG2_PATCH, 32, seed=7)
    from redux_amd.momfbd import Solver
    sv = Solver(pbs[0], precision=form, max_batch=len(pbs))
    sv.load(pbs)
    sv.init()
    a = np.stack([p["truth"]["alpha"] * 0.5 for p in pbs])
    for _ in range(3):
        sv.metric(a)
    for _ in range(3):
        sv.grad(a, method="vogel")
    sv.close()
    hip = C.CDLL("libamdhip64.so")
    src, dst = C.c_void_p(), C.c_void_p()
    if hip.hipMalloc(C.byref(src), C.c_size_t(COPY_PROBE_BYTES)) or hip.hipMalloc(C.byref(dst), C.c_size_t(COPY_PROBE_BYTES)):
        raise RuntimeError("hipMalloc failed")
    hip.hipMemset(src, 1, C.c_size_t(COPY_PROBE_BYTES))
    for _ in range(3):
        if hip.hipMemcpy(dst, src, C.c_size_t(COPY_PROBE_BYTES), 3):      # hipMemcpyDeviceToDevice
            raise RuntimeError("hipMemcpy failed")
    hip.hipDeviceSynchronize()
    hip.hipFree(src)
    hip.hipFree(dst)
    print(json.dumps({"probe": "pmc", "form": form, "images_per_otf_launch": 32 * len(pbs[0]["imgCh"]),
                      "copy_bytes": COPY_PROBE_BYTES}), flush=True)


def uneven(args):
    """The uneven batch: patches of the grids 64 / 128 / 256 with 36-100 KL modes and 12-48
    frames per object drawn per patch from the seed, one context per grid solved at once
    (mfbd_solve_group), FP64; beside it the same patches solved the way the reference runs
    them, one at a time in a context of each patch's own geometry (a bounded sample)."""
    from redux_amd import synth
    from redux_amd.momfbd import Solver, solve_group, uneven_solvers
    workers = min(16, os.cpu_count() or 1)
    t0 = time.perf_counter()
    groups = synth.uneven_batch(synth.UNEVEN_CLASSES, args.patches, seed=args.seed, workers=workers)
    log(f"generated {args.patches} uneven patches in {time.perf_counter() - t0:.1f}s: "
        + ", ".join(f"N={g['setup']['N']}: {len(g['patches'])}" for g in groups))
    groups = [g for g in groups if g["patches"]]
    prec = args.precision
    solvers = uneven_solvers([(g["patches"], g["shapes"]) for g in groups], precision=prec)
    times, res = [], None
    for k in range(args.warmup + args.steps):
        t1 = time.perf_counter()
        out = solve_group(solvers)
        dt = time.perf_counter() - t1
        log(f"uneven {prec} {'warmup' if k < args.warmup else 'step'}: {dt:.2f}s")
        if k >= args.warmup:
            times.append(dt)
            if res is not None:
                assert all(np.array_equal(a["img"], b["img"]) for ra, rb in zip(res, out) for a, b in zip(ra, rb))
            res = out
    lines = [(int(x) for x in sv.debug("fft_lines")) for sv in solvers]
    fft_lines = {f"N={g['setup']['N']}": list(v) for g, v in zip(groups, lines)}
    accuracy = None
    if prec != "fp64":
        # per-result error against the FP64 form of the same batch (itself pinned to the
        # oracle): final metric, alphas and restored objects of the solves, and the metric
        # and Vogel gradient of single evaluations at the FP64 solution
        s64 = uneven_solvers([(g["patches"], g["shapes"]) for g in groups], precision="fp64")
        ref = solve_group(s64)
        acc = {k: [] for k in ("metric_eval", "gradient_eval", "final_metric", "alpha", "object")}
        for g, sv, sr, ra, rr in zip(groups, solvers, s64, res, ref):
            a64 = np.stack([r["alpha"] for r in rr])
            idx = np.arange(len(rr), dtype=np.int32)
            sv.init()          # both contexts back to the frames' first initialisation, so the
            sr.init()          # two evaluate on the same frame state
            L, Lr = sv.metric(a64, patches=idx), sr.metric(a64, patches=idx)
            gv, gr = sv.grad(a64, patches=idx), sr.grad(a64, patches=idx)
            acc["metric_eval"] += list(np.abs(L - Lr) / np.abs(Lr))
            acc["gradient_eval"] += [float(np.linalg.norm(x - y) / np.linalg.norm(y)) for x, y in zip(gv, gr)]
            for a, b in zip(ra, rr):
                acc["final_metric"].append(abs(a["finalMetric"] - b["finalMetric"]) / abs(b["finalMetric"]))
                acc["alpha"].append(float(np.linalg.norm(a["alpha"] - b["alpha"]) / np.linalg.norm(b["alpha"])))
                acc["object"].append(float(np.linalg.norm(a["img"].astype(np.float64) - b["img"])
                                           / np.linalg.norm(b["img"])))
        for sv in s64:
            sv.close()
        accuracy = {k: {"median": float(np.median(v)), "max": float(np.max(v))} for k, v in acc.items()}
        accuracy["objects_within_1e-4"] = int(sum(1 for e in acc["object"] if e <= 1e-4))
    for sv in solvers:
        sv.close()
    dt = max(times)
    n = sum(len(g["patches"]) for g in groups)
    conv = sum(1 for ra in res for r in ra if r["status"] == 0)
    # one patch at a time in its own geometry: a few patches spread evenly over every
    # grid's list, their mean time per patch scaled by that grid's patch count
    plain_est, same, per_grid = 0.0, True, {}
    for g, ra in zip(groups, res):
        ts = []
        picks = sorted(set(np.linspace(0, len(g["own"]) - 1, min(args.plain_patches, len(g["own"]))).round().astype(int)))
        for k in picks:
            own = g["own"][k]
            sv = Solver(own, precision="fp64", max_batch=1)
            sv.load([own])
            t1 = time.perf_counter()
            r = sv.run()[0]
            ts.append(time.perf_counter() - t1)
            sv.close()
            sel = np.flatnonzero(g["shapes"][k]["imgOn"])
            if prec == "fp64":
                same &= bool(np.array_equal(r["img"], ra[k]["img"])
                             and np.array_equal(r["alpha"], ra[k]["alpha"][sel][:, :g["shapes"][k]["nModes"]]))
        plain_est += float(np.mean(ts)) * len(g["patches"])
        per_grid[f"N={g['setup']['N']}"] = round(1.0 / float(np.mean(ts)), 4)
        log(f"plain N={g['setup']['N']}: {np.mean(ts):.3f} s per patch")
    line = {
        "metric": f"converged patches/s, uneven batch ({prec.upper()})",
        "value": round(conv / dt, 4),
        "unit": "patches/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(dt * 1e3, 1),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": {"fp64": "f64", "fp32": "f32", "mixed": "bf16/f32"}[prec], "data": "synthetic",
        "config": {"workload": f"{n} patches, grids 64/128/256 (patches 64/88/128, pupils 32/64/128 px), "
                               "36-100 KL modes and 12-48 frames per object drawn per patch, 2 objects",
                   "patches_per_grid": {f"N={g['setup']['N']}": len(g["patches"]) for g in groups},
                   "form": "one context per grid, ragged shapes inside, contexts solved at once"},
        "converged": conv,
        "plain_one_at_a_time_patches_per_s": round(n / plain_est, 4),
        "plain_per_grid_patches_per_s": per_grid,
        "plain_sample": f"{args.plain_patches} patches spread over every grid, each solved alone in a context of "
                        "its own geometry; the batch's plain time = per grid, mean time per patch x patches",
        "plain_identical_to_batched": same if prec == "fp64" else None,
        "speedup_vs_plain": round(plain_est / dt, 2),
        "fft_lines_autotuned": fft_lines,
    }
    if accuracy:
        line["error_vs_fp64"] = accuracy
    print(json.dumps(line), flush=True)


def _cfg5_patch(k):
    from redux_amd import synth
    return synth.build_patch(_STATE["su5"], _STATE["seed5"] * 1000 + k)


def cfg5(args):
    """Config 5 per wave: patches of 128 x 128 on the 256 x 256 grid, 150 KL modes, 2 objects x
    100 frames, solved with the hand-written adjoint gradient in the FP32 form, `--patches` per
    context (one wave; the full 4,096 are waves of this size); objects compared with the FP64
    solve of the first few patches."""
    from redux_amd import synth
    from redux_amd.momfbd import Solver
    workers = min(16, os.cpu_count() or 1)
    t0 = time.perf_counter()
    _STATE.update(su5=synth.build_setup(synth.CFG5_PATCH), seed5=args.seed)
    if workers > 1:
        with ProcessPoolExecutor(max_workers=workers) as ex:
            pbs = list(ex.map(_cfg5_patch, range(args.patches), chunksize=1))
    else:
        pbs = [_cfg5_patch(k) for k in range(args.patches)]
    log(f"generated {len(pbs)} config-5 patches in {time.perf_counter() - t0:.1f}s")
    sv = make_solver(pbs, "fp32", len(pbs), "adjoint")
    times, res = [], None
    for k in range(args.warmup + args.steps):
        out, dt = run_step(sv)
        log(f"cfg5 fp32 adjoint {'warmup' if k < args.warmup else 'step'}: {dt:.2f}s")
        if k >= args.warmup:
            times.append(dt)
            res = out
    sv.close()
    dt = max(times)
    conv = sum(1 for r in res if r["status"] == 0)
    errs = []
    for k in range(min(args.plain_patches, len(pbs))):
        s64 = Solver(pbs[k], precision="fp64", max_batch=1)
        s64.load([pbs[k]])
        r64 = s64.run()[0]
        s64.close()
        ref = r64["img"].astype(np.float64)
        errs.append(float(np.linalg.norm(res[k]["img"] - ref) / np.linalg.norm(ref)))
    line = {
        "metric": "converged patches/s, config 5 wave (FP32, adjoint gradient)",
        "value": round(conv / dt, 4),
        "unit": "patches/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(dt * 1e3, 1),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": f"{len(pbs)} patches 128x128, N=256 OTF grid, pupil 128 px, 150 KL modes, "
                               "2 objects x 100 frames, D/r0=20, 2000 counts",
                   "form": "FP32 transforms, hand-written adjoint dL/dalpha, one context (a wave of config 5)"},
        "converged": conv,
        "evals_per_patch": {"f": float(np.mean([r["nF"] for r in res])), "grad": float(np.mean([r["nDF"] for r in res]))},
        "obj_rel_err_vs_fp64_solve": errs,
    }
    print(json.dumps(line), flush=True)


def trace_summary(trace_csv, bench_json):
    """Markdown tables of a rocprofv3 kernel trace of this script's default run: per kernel,
    the dispatches of the FP64 and the FP32 statistics steps (the windows from the first to
    the n-th dispatch of each form's roofline kernel, n = the bench line's launch count), and
    the roofline kernels' rocprofv3 mean beside bench.py's HIP-event mean."""
    import csv
    import collections
    with open(bench_json) as f:
        line = json.loads([x for x in f.read().splitlines() if x.startswith("{")][-1])
    with open(trace_csv) as f:
        ks = sorted((int(r["Start_Timestamp"]), int(r["End_Timestamp"]), r["Kernel_Name"]) for r in csv.DictReader(f))

    def short(n):
        return n.split("(")[0].replace("void ", "")

    out = []
    for title, key in (("FP64 statistics step", "fp64_roofline"), ("FP32 statistics step (the timed form)", "roofline")):
        rl = line[key]
        name, n = rl["kernel"], rl["launches"]
        mine = [k for k in ks if short(k[2]).split("<")[0] == "mfbd::" + name][:n]
        t0, t1 = mine[0][0], mine[-1][1]
        tot = collections.defaultdict(lambda: [0, 0])
        for s0, e0, k in ks:
            if s0 >= t0 and e0 <= t1:
                tot[short(k)][0] += e0 - s0
                tot[short(k)][1] += 1
        out.append(f"**{title}** (dispatches from the first `{name}` to its {n}th, "
                   f"{(t1 - t0) / 1e6:.1f} ms):\n")
        out.append("| kernel | dispatches | total ms | mean µs |\n|---|---:|---:|---:|")
        for k, (t, c) in sorted(tot.items(), key=lambda kv: -kv[1][0]):
            if t / 1e6 >= 0.05:
                out.append(f"| `{k}` | {c} | {t / 1e6:.1f} | {t / c / 1e3:.2f} |")
        mean = sum(e0 - s0 for s0, e0, _ in mine) / len(mine) / 1e6
        out.append(f"\n`{name}`: rocprofv3 mean {mean:.4f} ms over {len(mine)} dispatches; bench.py's HIP-event "
                   f"mean {rl['mean_launch_ms']:.4f} ms, {rl['images_per_launch']:.0f} images per launch, "
                   f"{rl['achieved']:.0f} GB/s ({rl['frac']:.3f} of HBM peak), FFT {rl['fft_tflops']:.1f} TFLOP/s "
                   f"({rl['fft_frac']:.3f} of the vector peak)\n")
    print("\n".join(out))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--patches", type=int, default=256)
    ap.add_argument("--plain-patches", type=int, default=16)
    ap.add_argument("--oracle-sample", type=int, default=16,
                    help="patches of the batch solved by the FP64 oracle for the convergence check and the "
                         "CPU baseline")
    ap.add_argument("--gradient", choices=["reference", "adjoint"], default="reference")
    ap.add_argument("--seed", type=int, default=2024)
    ap.add_argument("--workload", choices=["cfg2", "uneven", "cfg5"], default="cfg2",
                    help="cfg2: the headline (256 patches of the benchmark geometry); uneven: the uneven batch "
                         "over three grids; cfg5: a wave of config-5 patches, FP32 adjoint (their own JSON lines)")
    ap.add_argument("--precision", choices=["fp64", "fp32", "mixed"], default="fp64",
                    help="form of the uneven workload (mixed: BF16 mode GEMM, FP32 transforms; errors reported "
                         "against the FP64 form)")
    ap.add_argument("--config3-patches", type=int, default=96,
                    help="uneven patches of the configs 3 / 4 sample in the headline line (0: skip it)")
    ap.add_argument("--config5-patches", type=int, default=8,
                    help="config-5 patches of the wave sample in the headline line")
    ap.add_argument("--pmc-probe", choices=["fp32", "fp64", "mixed"], default=None,
                    help="run the short workload profiled for HBM traffic (FETCH_SIZE / WRITE_SIZE) in one "
                         "precision form and exit")
    ap.add_argument("--trace-summary", nargs=2, metavar=("TRACE_CSV", "BENCH_JSON"), default=None,
                    help="print the statistics steps' kernel tables from a rocprofv3 kernel trace of this "
                         "script's default run and its JSON line, and exit")
    ap.add_argument("--line-graph", type=int, choices=[0, 1], default=1,
                    help="1 (default): the FP32 and mixed forms' device line-search rounds replay captured HIP "
                         "graphs; 0: their kernels launched one by one (the same kernels and results; for "
                         "rocprofv3's kernel trace, which cannot follow graph launches)")
    args = ap.parse_args()
    global LINE_GRAPH
    LINE_GRAPH = bool(args.line_graph)
    if args.trace_summary:
        trace_summary(*args.trace_summary)
        return
    if args.pmc_probe:
        pmc_probe(args.pmc_probe)
        return
    if args.gpus != 1:
        print("bench.py measures one GPU: patches are independent and shard across GPUs with no "
              "collective; run with --gpus 1", file=sys.stderr)
        sys.exit(2)
    if args.workload == "uneven":
        uneven(args)
        return
    if args.workload == "cfg5":
        cfg5(args)
        return
    headline(args)


def other_configs(args):
    """Configs 3, 4 and 5 at their stated sizes on bounded samples, for the headline line:
    an uneven batch (grids 64 / 128 / 256, 36-100 KL modes and 12-48 frames per object drawn
    per patch) solved as one context per grid at once in FP64 (config 3) and in the mixed
    form (config 4, objects against the FP64 solve of the same batch), with one patch per grid
    solved alone in its own geometry beside it; and a wave of config-5 patches (N = 256, 150
    KL modes, 2 x 100 frames) in FP32 with the adjoint gradient (its objects against FP64
    solves are in --workload cfg5).  Each is one warm-up and one timed solve; the full-size
    runs are --workload uneven / cfg5."""
    from redux_amd import synth
    from redux_amd.momfbd import Solver, solve_group, uneven_solvers
    workers = min(16, os.cpu_count() or 1)
    out = {}
    groups = synth.uneven_batch(synth.UNEVEN_CLASSES, args.config3_patches, seed=args.seed, workers=workers)
    groups = [g for g in groups if g["patches"]]
    sets = [(g["patches"], g["shapes"]) for g in groups]
    res = {}
    for prec in ("fp64", "mixed"):
        sv = uneven_solvers(sets, precision=prec)
        solve_group(sv)
        t0 = time.perf_counter()
        res[prec] = solve_group(sv)
        dt = time.perf_counter() - t0
        for x in sv:
            x.close()
        n = sum(len(r) for r in res[prec])
        conv = sum(1 for ra in res[prec] for r in ra if r["status"] == 0)
        out[prec] = dict(patches=n, seconds=round(dt, 3), patches_per_s=round(n / dt, 3), converged=conv)
        log(f"config {3 if prec == 'fp64' else 4}: {n} uneven patches {prec} in {dt:.2f}s")
    errs = [float(np.linalg.norm(a["img"].astype(np.float64) - b["img"]) / np.linalg.norm(b["img"]))
            for ra, rb in zip(res["mixed"], res["fp64"]) for a, b in zip(ra, rb)]
    out["mixed"]["obj_rel_err_vs_fp64"] = {"median": float(np.median(errs)), "max": float(np.max(errs))}
    plain, same = {}, True
    for g, ra in zip(groups, res["fp64"]):
        own = g["own"][0]
        sv = Solver(own, precision="fp64", max_batch=1)
        sv.load([own])
        t0 = time.perf_counter()
        r = sv.run()[0]
        plain[f"N={g['setup']['N']}"] = round(time.perf_counter() - t0, 4)
        sv.close()
        sel = np.flatnonzero(g["shapes"][0]["imgOn"])
        same &= bool(np.array_equal(r["img"], ra[0]["img"])
                     and np.array_equal(r["alpha"], ra[0]["alpha"][sel][:, :g["shapes"][0]["nModes"]]))
    est = sum(plain[f"N={g['setup']['N']}"] * len(g["patches"]) for g in groups)
    config3 = dict(out["fp64"], patches_per_grid={f"N={g['setup']['N']}": len(g["patches"]) for g in groups},
                   plain_seconds_per_patch=plain, plain_identical_to_batched=same,
                   speedup_vs_plain=round(est / out["fp64"]["seconds"], 2))
    # config 5: a wave of patches, FP32 with the adjoint gradient
    _STATE.update(su5=synth.build_setup(synth.CFG5_PATCH), seed5=args.seed)
    with ProcessPoolExecutor(max_workers=workers) as ex:
        pbs = list(ex.map(_cfg5_patch, range(args.config5_patches), chunksize=1))
    sv = make_solver(pbs, "fp32", len(pbs), "adjoint")
    run_step(sv)
    r5, dt5 = run_step(sv)
    sv.close()
    log(f"config 5: {len(pbs)} patches fp32 adjoint in {dt5:.2f}s")
    return {
        "config3_uneven_fp64": config3,
        "config4_uneven_mixed": out["mixed"],
        "config5_wave_fp32_adjoint": dict(patches=len(pbs), seconds=round(dt5, 3),
                                          patches_per_s=round(len(pbs) / dt5, 4),
                                          converged=sum(1 for r in r5 if r["status"] == 0),
                                          evals_per_patch={"f": float(np.mean([r["nF"] for r in r5])),
                                                           "grad": float(np.mean([r["nDF"] for r in r5]))}),
        "note": "one timed solve each after a warm-up, on bounded samples of the stated sizes; full runs: "
                "--workload uneven (1,024 patches), --workload cfg5 (64-patch waves)",
    }


def oracle_sample(pbs, n_sample, threads):
    """Oracle (FP64 C restatement) solves of a fixed sample of the batch, every
    (len/n_sample)-th patch, on `threads` host threads, one patch per thread at a time:
    the reference objects and alphas of the convergence check, and, timed, the CPU
    baseline."""
    from oracle import oracle as O
    O.build()
    idx = list(range(0, len(pbs), max(1, len(pbs) // n_sample)))[:n_sample]
    t0 = time.perf_counter()
    res = O.solve_many([pbs[k] for k in idx], threads=threads)
    wall = time.perf_counter() - t0
    return idx, res, wall


def against_oracle(res, idx, ores):
    """Relative RMS errors of a form's restored objects and alphas against the oracle's solves
    of the sample patches; a patch converges when its solve exits normally and its objects
    are within 1e-4."""
    obj, alp, ok = [], [], 0
    for k, o in zip(idx, ores):
        r = res[k]
        e = float(np.linalg.norm(r["img"].astype(np.float64) - o["img"]) / np.linalg.norm(o["img"]))
        a = float(np.linalg.norm(r["alpha"] - o["alpha"]) / np.linalg.norm(o["alpha"]))
        obj.append(e)
        alp.append(a)
        ok += int(r["status"] == 0 and o["status"] == 0 and e <= 1e-4)
    return {"sample": len(idx), "within_1e-4": ok, "fraction": round(ok / len(idx), 4),
            "obj_rel_err_median": float(np.median(obj)), "obj_rel_err_max": float(np.max(obj)),
            "alpha_rel_err_median": float(np.median(alp)), "alpha_rel_err_max": float(np.max(alp))}


def stage_roofline(stats, N, P, form):
    """The OTF + P/Q + metric stage (pupil field, both transforms, P/Q sums, metric) against the
    stage bytes per image-evaluation of the path's roofline table: read phi on the P x P
    pupil grid and D_j on the N x N grid, r P^2 + c N^2 (c, r the complex and real sizes)."""
    c, r = (16, 8) if form == "fp64" else (8, 4)
    names = [k for k in stats if k in ("k_otf64_fused", "k_otf_fused", "k_otf_wave", "k_pq", "k_field_rows",
                                       "k_cols_power", "k_rows_inv_real", "k_cols_extract")]
    ms = sum(stats[k]["ms"] for k in names)
    imgs = stats.get("k_otf64_fused", stats.get("k_otf_wave", stats.get("k_otf_fused", {"units": 0})))["units"]
    per = r * P * P + c * N * N
    ach = per * imgs / (ms * 1e-3) / 1e9 if ms > 0 else 0.0
    return dict(kernels=names, ms_per_step=round(ms, 3), image_evaluations=int(imgs), bytes_per_image=per,
                achieved=round(ach, 1), peak=HBM_PEAK_GBS, unit="GB/s", frac=round(ach / HBM_PEAK_GBS, 4))


def headline(args):
    """The headline workload: converged patches/s of the batched FP32 form, where a patch
    converges when its solve exits normally and its restored objects are within 1e-4
    (relative RMS) of the FP64 oracle's solve of the same patch, checked on a fixed sample
    of the batch; beside it the plain FP64 drop-in (one patch at a time), the batched FP64
    and mixed forms, all with the same check."""
    workers = min(16, os.cpu_count() or 1)
    t0 = time.perf_counter()
    su, pbs = make_batch(args.patches, args.seed, workers)
    log(f"generated {len(pbs)} patches in {time.perf_counter() - t0:.1f}s")
    settings = pbs[0]["settings"]

    # Oracle solves of the sample (reference objects and the timed CPU baseline)
    threads = min(16, os.cpu_count() or 1)
    idx, ores, owall = oracle_sample(pbs, args.oracle_sample, threads)
    log(f"oracle: {len(idx)} patches on {threads} threads in {owall:.1f}s")
    cpu = dict(value=round(len(idx) / owall, 5), unit="patches/s", cores=threads, kind="port",
               sample=f"oracle FP64 solves (C restatement of Solver::run, to convergence, with restoration) of "
                      f"{len(idx)} of the {len(pbs)} patches (every {max(1, len(pbs) // args.oracle_sample)}th), "
                      f"{threads} host threads, one patch per thread: {owall:.1f}s wall")

    # Per-kernel statistics: one step of each form on a context that brackets every kernel
    # with HIP events and keeps its evaluation lanes on one stream
    def stats_step(prec, line=None):
        sv = make_solver(pbs, prec, len(pbs), args.gradient, timing=True, line=line)
        res, _ = run_step(sv)
        st = sv.kernel_stats()
        sizes = [int(x) for x in sv.debug("sizes")]
        sv.close()
        return res, st, sizes

    ref64, st64, (nps, nos, nh) = stats_step("fp64", line=True)
    ref32, st32, _ = stats_step("fp32")

    # Timed form: batched FP32 (warmup steps, then the timed steps) on a context without events
    sv = make_solver(pbs, "fp32", len(pbs), args.gradient)
    for _ in range(args.warmup):
        _, dt = run_step(sv)
        log(f"fp32 batched warmup: {dt:.2f}s")
    times = []
    for _ in range(args.steps):
        res32, dt = run_step(sv)
        times.append(dt)
        log(f"fp32 batched step: {dt:.2f}s")
    sv.close()
    t32 = max(times)
    same32 = all(np.array_equal(a["img"], b["img"]) for a, b in zip(res32, ref32))
    chk32 = against_oracle(res32, idx, ores)
    rate32 = len(pbs) / t32
    value = rate32 * chk32["fraction"]

    def one_step(prec, line=None):
        sv = make_solver(pbs, prec, len(pbs), args.gradient, line=line)
        run_step(sv)                    # warm-up (first-use costs: graph captures, workspaces)
        res, dt = run_step(sv)
        sv.close()
        log(f"{prec} batched{'' if line is None else ' line phases' if line else ' GEMM per probe'}: {dt:.2f}s")
        return res, dt

    # the batched FP64 form: line phases and the device line search; beside it the
    # reference's arithmetic batched (the mode GEMM per probe, the host's line search)
    res64, t64 = one_step("fp64", line=True)
    res64r, t64r = one_step("fp64", line=False)
    resmx, tmx = one_step("mixed")
    chk64, chkmx = against_oracle(res64, idx, ores), against_oracle(resmx, idx, ores)
    same64 = all(np.array_equal(a["img"], b["img"]) and np.array_equal(a["alpha"], b["alpha"])
                 for a, b in zip(res64, ref64))

    # plain FP64 drop-in: the sample patches one at a time in the reference's arithmetic (the
    # mode GEMM per probe, the host's line search); their results must equal the same
    # arithmetic batched
    plain = [pbs[k] for k in idx[: args.plain_patches]]
    sv = make_solver(plain, "fp64", 1, args.gradient, line=False)
    run_step(sv)                        # warm-up, as for the batched forms
    resp, tp = run_step(sv)
    sv.close()
    plain_rate = len(plain) / tp
    plain_identical = all(np.array_equal(a["img"], res64r[k]["img"]) and np.array_equal(a["alpha"], res64r[k]["alpha"])
                          for a, k in zip(resp, idx))
    log(f"fp64 plain: {len(plain)} patches in {tp:.2f}s, identical to the GEMM-per-probe batch: {plain_identical}")
    chk64r = against_oracle(res64r, idx, ores)
    chk_plain = against_oracle({k: r for k, r in zip(idx, resp)}, idx[: len(plain)], ores[: len(plain)])

    ev32 = {"f": float(np.mean([r["nF"] for r in res32])), "grad": float(np.mean([r["nDF"] for r in res32]))}
    ev64 = {"f": float(np.mean([r["nF"] for r in res64])), "grad": float(np.mean([r["nDF"] for r in res64]))}
    line = {
        "metric": "converged patches/s (rel. obj RMS \u22641e-4 vs FP64) batched FP32 vs plain FP64",
        "value": round(value, 4),
        "unit": "patches/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(t32 * 1e3, 1),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"{len(pbs)} patches 88x88 from a 1024^2 k^-3 field, 2 objects x 24 frames, "
                               f"50 KL modes, N=128 OTF grid, pupil 64 px, D/r0=20, 2000 counts",
                   "form": "batched FP32 (FP32 phase GEMM and transforms, FP64 P/Q/metric reductions and CG)",
                   "solver": {k: settings[k] for k in ("FTOL", "EPS", "MIN_ITER", "MAX_ITER", "N_DONE_ITER",
                                                       "MODE_START", "MODE_STEP", "GRADDIFF_STEP")}
                   | {"GETSTEP": "conjugate", "gradient": args.gradient,
                      "line_phases": "batched forms: probes from base + step * direction (metricAt); plain "
                                     "FP64 drop-in: the mode GEMM per probe (the reference's arithmetic)",
                      "line_search": "batched forms: on the device (bracket + brent state machine per patch, each "
                                     "round of probes " + ("replayed as a captured HIP graph" if LINE_GRAPH else
                                                           "launched kernel by kernel") +
                                     ", read back every 4 rounds); plain FP64 drop-in: stepped by the host"},
                   "converged": "solve exits normally and restored objects within 1e-4 (relative RMS) of the "
                                f"FP64 oracle's solve, on a sample of {len(idx)} patches; value = patches/s x "
                                "sampled fraction",
                   "patches": len(pbs), "parallelism": "batch"},
        "fp32_batched": dict(patches_per_s=round(rate32, 4), vs_oracle=chk32, evals_per_patch=ev32,
                             timed_steps_identical_to_stats_step=bool(same32)),
        "fp64_batched": dict(patches_per_s=round(len(pbs) / t64, 4), vs_oracle=chk64, evals_per_patch=ev64,
                             converged_patches_per_s=round(len(pbs) / t64 * chk64["fraction"], 4),
                             identical_to_stats_step=bool(same64),
                             form="line phases and the device line search"),
        "fp64_batched_gemm_per_probe": dict(patches_per_s=round(len(pbs) / t64r, 4), vs_oracle=chk64r,
                                            converged_patches_per_s=round(len(pbs) / t64r * chk64r["fraction"], 4),
                                            form="the reference's arithmetic (mode GEMM per probe, host line "
                                                 "search), batched"),
        "mixed_bf16_batched": dict(patches_per_s=round(len(pbs) / tmx, 4), vs_oracle=chkmx,
                                   converged_patches_per_s=round(len(pbs) / tmx * chkmx["fraction"], 4)),
        "fp64_plain": dict(patches_per_s=round(plain_rate, 4), patches=len(plain),
                           identical_to_fp64_batched_gemm_per_probe=bool(plain_identical), vs_oracle=chk_plain),
        "ratio_fp32_batched_vs_fp64_plain": {"raw": round(rate32 / plain_rate, 2),
                                             "converged": round(value / (plain_rate * chk_plain["fraction"]), 2)
                                             if chk_plain["fraction"] > 0 else None},
        "ratio_fp64_batched_vs_fp64_plain": round((len(pbs) / t64) / plain_rate, 2),
        "kernels_ms_per_step": {"fp32": {k: round(v["ms"], 3) for k, v in sorted(st32.items(), key=lambda kv: -kv[1]["ms"])
                                         if not k.startswith("_")},
                                "fp64": {k: round(v["ms"], 3) for k, v in sorted(st64.items(), key=lambda kv: -kv[1]["ms"])
                                         if not k.startswith("_")}},
        "kernel_stats_note": "one step per form on a context with HIP events around every kernel and its evaluation "
                             "lanes on one stream, so each launch's duration is its own; the timed steps run "
                             "without events and with the lanes' kernels overlapping on their own streams",
        "roofline": roofline(st32, su["N"], nps, nh, "fp32"),
        "stage_roofline": {"fp32": stage_roofline(st32, su["N"], su["P"], "fp32"),
                           "fp64": stage_roofline(st64, su["N"], su["P"], "fp64")},
        "fp64_roofline": roofline(st64, su["N"], nps, nh, "fp64"),
        "cpu_baseline": cpu,
    }
    if args.config3_patches > 0:
        line["other_configs"] = other_configs(args)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
