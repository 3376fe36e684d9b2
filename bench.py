"""Benchmark: converged patches/s of the batched FP32 solver vs the plain FP64 drop-in.

Workload (default): 256 overlapping 88x88 patches from a seeded 1024x1024 k^-3
field, 2 objects (wideband + narrowband with a 1.0-wave peak-to-valley focus
diversity) x 24 frames sharing wavefronts, 50 KL modes (Kolmogorov, D/r0 = 20),
Poisson noise at 2000 mean counts, pupil 64 px on a 128 x 128 OTF grid.

One step = solving every patch of the batch to convergence (Solver::run per
patch: mode stepping, CG + Brent, restoration) with all patches advancing in
lockstep; the batch's metric/gradient evaluations go to the device in two alternating
passes (halves of the batch) so the host steps one half while the device evaluates the
other.
A patch counts as converged when its solve exits normally and its restored
objects match the FP64 solve of the same patch to relative RMS <= 1e-4.

Each form gets one context with its patches loaded (frames resident in HBM); a
step is one mfbd_solve of every loaded patch from zero alphas, so warmup and timed
steps repeat the whole computation on the same context.

Also reported: the batched FP64 solve (the reference objects; it meets the 1e-4
criterion by construction, being bit-identical to one patch at a time), the plain
FP64 drop-in (one patch at a time, the reference's decomposition) on a slice of the
batch, per-kernel HIP-event timings with the roofline of the dominant kernel, and the
CPU FP64 restatement (oracle) timed on the host cores as a baseline.
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
# FP32 vector peak without packed-FP32 instructions (which no kernel here uses):
# 256 CUs x 64 lanes x 2 flops (FMA) x 2.4 GHz
FP32_VECTOR_PEAK_TFLOPS = 78.6


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


def make_solver(pbs, precision, batch, gradient, timing=False):
    """A context with the patches loaded (frames resident in HBM).  With `timing`,
    every kernel launch is bracketed by HIP events on the pipeline's stream."""
    from redux_amd.momfbd import Solver
    if timing:
        os.environ["MFBD_KERNEL_TIMING"] = "1"
    try:
        sv = Solver(pbs[0], precision=precision, max_batch=batch, gradient=gradient)
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


def measured_traffic():
    """HBM bytes per image of the hot kernels, measured with rocprofv3 FETCH_SIZE /
    WRITE_SIZE (scaled by a plain copy of known size) and committed under profiles/."""
    path = os.path.join(ROOT, "profiles", "k_otf_fused_cfg2_traffic.md")
    try:
        with open(path) as f:
            for line in f:
                if line.startswith("traffic-json:"):
                    return json.loads(line.split(":", 1)[1])
    except OSError:
        pass
    return {}


def roofline(stats, N, nPS, nH):
    """Dominant kernel by total time; achieved = algorithmic bytes per launch / mean
    launch time, over every launch of the batched FP32 context (warmup and timed steps).  S, D, P and Q live on the
    stored (Hermitian) half of the OTF support, nH points.  Algorithmic bytes per image
    (FP32): k_otf_fused reads phi on the pupil support and writes S (4 nPS + 8 nH);
    k_pq reads S and D (16 nH); k_phase_mfma writes phi (4 nPS); the fused Vogel kernel
    reads S, D (8 nH each), P, Q (24 nH, FP64) and phi (4 nPS) and writes v (4 nPS);
    an unfused FFT pass reads and writes the N x N tile (16 N^2).  For k_otf_fused the
    FFT arithmetic is reported beside it: two N x N complex transforms, 5 N^2 log2(N^2)
    flops each, against the FP32 vector peak without packed-FP32 instructions (78.6
    TFLOP/s; the 157.3 TFLOP/s figure counts packed FP32, which these kernels do not use)."""
    per_unit = {"fft_rows": 16 * N * N, "fft_cols": 16 * N * N, "k_field": 4 * nPS + 8 * N * N,
                "k_power": 16 * N * N, "k_extract": 16 * nH, "k_phase_mfma": 4 * nPS, "k_pq": 16 * nH,
                "k_otf_fused": 4 * nPS + 8 * nH,
                "k_vogel_fused": 16 * nH + 24 * nH + 8 * nPS}
    kern = {k: v for k, v in stats.items() if not k.startswith("host") and not k.startswith("_")}
    name, s = max(kern.items(), key=lambda kv: kv[1]["ms"])
    out = dict(kernel=name)
    if name in per_unit and s["units"] > 0:
        mean_ms = s["ms"] / s["launches"]
        units_per_launch = s["units"] / s["launches"]
        bytes_per_launch = per_unit[name] * units_per_launch
        ach = bytes_per_launch / (mean_ms * 1e-3) / 1e9
        tr = measured_traffic().get(name)
        out.update(bound="hbm", achieved=round(ach, 1), peak=HBM_PEAK_GBS, unit="GB/s",
                   frac=round(ach / HBM_PEAK_GBS, 4),
                   traffic=None if tr is None else int(tr * units_per_launch), launches=s["launches"],
                   mean_launch_ms=round(mean_ms, 4), images_per_launch=round(units_per_launch, 1),
                   bytes_per_image=per_unit[name], bytes_per_launch=int(bytes_per_launch))
        if name == "k_otf_fused":
            flops = 2 * 5 * N * N * math.log2(N * N) * units_per_launch
            tf = flops / (mean_ms * 1e-3) / 1e12
            out.update(fft_tflops=round(tf, 2),
                       fft_tflops_frac_of_fp32_vector_peak=round(tf / FP32_VECTOR_PEAK_TFLOPS, 4))
    return out


def cpu_baseline(pbs, evals_f, evals_g, seconds, threads):
    """The FP64 oracle (C restatement) on host threads, one patch per thread, timing
    metric and Vogel-gradient evaluations for about `seconds`; converted to patches/s
    with the mean f/gradient evaluation counts per converged patch of the FP64 solve."""
    from concurrent.futures import ThreadPoolExecutor

    from oracle import oracle as O
    O.build()
    n = min(threads, len(pbs))
    states = [O.State(pbs[k]) for k in range(n)]
    alpha = np.zeros_like(pbs[0]["truth"]["alpha"])

    def work(k):
        st = states[k]
        t0 = time.perf_counter()
        nf = ng = 0
        tf = tg = 0.0
        while time.perf_counter() - t0 < seconds:
            t1 = time.perf_counter()
            st.metric(alpha)
            tf += time.perf_counter() - t1
            nf += 1
            if nf % 8 == 0:
                t1 = time.perf_counter()
                st.grad(alpha, method="vogel")
                tg += time.perf_counter() - t1
                ng += 1
        return nf, ng, tf, tg

    t0 = time.perf_counter()
    with ThreadPoolExecutor(max_workers=n) as ex:
        parts = list(ex.map(work, range(n)))
    wall = time.perf_counter() - t0
    nf = sum(p[0] for p in parts)
    ng = sum(p[1] for p in parts)
    tf = sum(p[2] for p in parts) / max(nf, 1)
    tg = sum(p[3] for p in parts) / max(ng, 1)
    per_patch = evals_f * tf + evals_g * tg            # CPU seconds per converged patch on one core
    rate = n / per_patch                               # n patches in parallel on n cores
    return dict(value=round(rate, 5), unit="patches/s", cores=n, kind="port",
                sample=(f"oracle FP64 on {n} host threads for {wall:.0f}s: {nf} metric evals "
                        f"({tf * 1e3:.1f} ms each), {ng} Vogel gradients ({tg * 1e3:.1f} ms each) on config-2 "
                        f"patches; rate = threads / (mean per-patch evals {evals_f:.0f} f + {evals_g:.0f} grad "
                        f"of the FP64 solve)"))


COPY_PROBE_BYTES = 256 << 20


def pmc_probe():
    """Workload for the rocprofv3 --pmc runs behind `traffic`: three batched FP32 metric
    evaluations of 32 config-2 patches (k_otf_fused at 1536 images per launch), then
    three device-to-device copies of COPY_PROBE_BYTES through the HIP runtime, the
    plain copy that calibrates FETCH_SIZE / WRITE_SIZE."""
    import ctypes as C
    from redux_amd import synth
    su, pbs = synth.build_field_batch(synth.CFG2_PATCH, 32, seed=7)
    from redux_amd.momfbd import Solver
    sv = Solver(pbs[0], precision="fp32", max_batch=len(pbs))
    sv.load(pbs)
    sv.init()
    a = np.stack([p["truth"]["alpha"] * 0.5 for p in pbs])
    for _ in range(3):
        sv.metric(a)
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
    print(json.dumps({"probe": "pmc", "images_per_otf_launch": 32 * len(pbs[0]["imgCh"]),
                      "copy_bytes": COPY_PROBE_BYTES}), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--patches", type=int, default=256)
    ap.add_argument("--plain-patches", type=int, default=4)
    ap.add_argument("--gradient", choices=["reference", "adjoint"], default="reference")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--seed", type=int, default=2024)
    ap.add_argument("--pmc-probe", action="store_true",
                    help="run the short workload profiled for HBM traffic (FETCH_SIZE / WRITE_SIZE) and exit")
    args = ap.parse_args()
    if args.pmc_probe:
        pmc_probe()
        return
    if args.gpus != 1:
        print("bench.py measures one GPU: patches are independent and shard across GPUs with no "
              "collective (see DESIGN.md); run with --gpus 1", file=sys.stderr)
        sys.exit(2)
    workers = min(16, os.cpu_count() or 1)
    t0 = time.perf_counter()
    su, pbs = make_batch(args.patches, args.seed, workers)
    log(f"generated {len(pbs)} patches in {time.perf_counter() - t0:.1f}s")

    # FP64 batched solve: the reference objects (identical to the plain drop-in, batch-invariant)
    sv = make_solver(pbs, "fp64", len(pbs), args.gradient)
    ref, t64b = run_step(sv)
    sv.close()
    log(f"fp64 batched: {t64b:.2f}s")
    ev_f = float(np.mean([r["nF"] for r in ref]))
    ev_g = float(np.mean([r["nDF"] for r in ref]))

    # plain FP64 drop-in: one patch at a time
    plain = pbs[: args.plain_patches]
    sv = make_solver(plain, "fp64", 1, args.gradient)
    _, t64p = run_step(sv)
    sv.close()
    plain_rate = len(plain) / t64p
    log(f"fp64 plain: {len(plain)} patches in {t64p:.2f}s")

    # batched FP32: one context, warmup steps, then the timed steps; kernel statistics
    # cover every launch of the context (warmup and timed steps alike)
    sv = make_solver(pbs, "fp32", len(pbs), args.gradient, timing=True)
    for _ in range(args.warmup):
        run_step(sv)
    times, res32 = [], None
    for _ in range(args.steps):
        res32, dt = run_step(sv)
        times.append(dt)
        log(f"fp32 batched step: {dt:.2f}s")
    stats = sv.kernel_stats()
    nps, nos, nh = (int(x) for x in sv.debug("sizes"))
    sv.close()
    nRuns = args.warmup + args.steps
    t32 = max(times)

    # mixed precision (BF16 mode GEMM): one solve on its own context
    sv = make_solver(pbs, "mixed", len(pbs), args.gradient)
    resmx, tmx = run_step(sv)
    sv.close()
    log(f"mixed batched: {tmx:.2f}s")
    errs_mx = np.array([np.linalg.norm(a["img"].astype(np.float64) - b["img"]) / np.linalg.norm(b["img"])
                        if a["status"] == 0 else np.inf for a, b in zip(resmx, ref)])
    errs = []
    for a, b in zip(res32, ref):
        e = np.linalg.norm(a["img"].astype(np.float64) - b["img"]) / np.linalg.norm(b["img"])
        errs.append(e if a["status"] == 0 else np.inf)
    errs = np.array(errs)
    conv = int(np.sum(errs <= 1e-4))
    value = conv / t32

    roof = roofline(stats, su["N"], nps, nh)
    cpu = cpu_baseline(pbs, ev_f, ev_g, args.cpu_seconds, min(16, os.cpu_count() or 1))
    line = {
        "metric": "converged patches/s (rel. obj RMS <=1e-4 vs FP64) batched FP32 vs plain FP64",
        "value": round(value, 4),
        "unit": "patches/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(t32 * 1e3, 1),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic",
        "config": {"workload": f"{len(pbs)} patches 88x88 from a 1024^2 k^-3 field, 2 objects x 24 frames, "
                               f"50 KL modes, N=128 OTF grid, pupil 64 px, D/r0=20, 2000 counts",
                   "patches": len(pbs), "gradient": args.gradient, "parallelism": "batch"},
        "fp32_batched_patches_per_s": round(len(pbs) / t32, 4),
        "fp32_converged": conv,
        "fp32_obj_rel_err_median": float(np.median(errs[np.isfinite(errs)])) if np.isfinite(errs).any() else None,
        "fp32_obj_rel_err_max": float(np.max(errs)),
        "fp64_batched_patches_per_s": round(len(pbs) / t64b, 4),
        "fp64_batched_converged": len(pbs),
        "mixed_bf16_batched_patches_per_s": round(len(pbs) / tmx, 4),
        "mixed_obj_rel_err_median": float(np.median(errs_mx[np.isfinite(errs_mx)])) if np.isfinite(errs_mx).any()
        else None,
        "mixed_converged": int(np.sum(errs_mx <= 1e-4)),
        "fp64_plain_patches_per_s": round(plain_rate, 4),
        "speedup_fp32_batched_vs_fp64_plain": round((len(pbs) / t32) / plain_rate, 2),
        "evals_per_patch": {"f": ev_f, "grad": ev_g},
        "kernels_ms_per_step": {k: round(v["ms"] / nRuns, 3)
                                for k, v in sorted(stats.items(), key=lambda kv: -kv[1]["ms"])},
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
