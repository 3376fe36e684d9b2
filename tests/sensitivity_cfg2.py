"""Metric-noise sensitivity of the oracle's solve on config-2 patches (CPU; minutes).

Solves the first `n` patches of the benchmark batch (256 patches of 88x88 from the seeded
1024^2 field) with the oracle, then again with every metric value the minimiser sees
multiplied by 1 + noise*u (u uniform in [-1, 1), one seeded stream per patch), and prints
per noise level how many restored objects stay within 1e-4 (relative RMS) of the
noise-free solve and how many patches keep the same integer re-cuts.

Run: python tests/sensitivity_cfg2.py [n_patches] [FTOL] [noise,noise,...] [threads]
"""
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

from oracle import oracle as O  # noqa: E402
from oracle import setup_np as S  # noqa: E402
from redux_amd import synth  # noqa: E402


def main():
    n = int(sys.argv[1]) if len(sys.argv) > 1 else 16
    ftol = float(sys.argv[2]) if len(sys.argv) > 2 else 1e-3
    noises = [float(v) for v in sys.argv[3].split(",")] if len(sys.argv) > 3 else [1e-14, 1e-12, 1e-10]
    threads = int(sys.argv[4]) if len(sys.argv) > 4 else min(16, os.cpu_count() or 1)
    cfg = dict(synth.CFG2_PATCH, settings=dict(FTOL=ftol))
    _, pbs = synth.build_field_batch(cfg, 256, setup=S)
    pbs = pbs[:n]
    t = time.perf_counter()
    base = O.solve_many(pbs, threads=threads)
    print(f"FTOL {ftol:g}: {n} patches in {time.perf_counter() - t:.0f}s, iterations "
          f"{min(r['iterations'] for r in base)}-{max(r['iterations'] for r in base)}", flush=True)
    for noise in noises:
        rs = O.solve_many(pbs, threads=threads, metric_noise=noise, noise_seeds=list(range(1, n + 1)))
        errs = np.array([np.linalg.norm(r["img"] - b["img"]) / np.linalg.norm(b["img"]) for r, b in zip(rs, base)])
        same = sum(int(np.array_equal(r["shifts"], b["shifts"])) for r, b in zip(rs, base))
        print(f"noise {noise:g}: objects within 1e-4 {int((errs <= 1e-4).sum())}/{n}, same re-cuts {same}/{n}, "
              f"median {np.median(errs):.1e}, max {errs.max():.1e}", flush=True)


if __name__ == "__main__":
    main()
