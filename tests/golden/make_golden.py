"""Writes tests/golden/cfg1_golden.npz: one config-1 patch (1 object, 1 channel,
8 frames of 64x64, pupil 32 px on a 64 grid, Zernike Noll 2-21, D/r0 = 10, mean
1000 counts, Poisson noise), built with the oracle's setup, and the oracle's FP64
outputs on it: metric, Vogel and finite-difference gradients at four alpha points,
and the converged alphas, metric history and restored object of a full solve.

Run: python tests/golden/make_golden.py   (CPU, a few seconds)
"""
import json
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from oracle import oracle as O  # noqa: E402
from oracle import setup_np as S  # noqa: E402
from redux_amd import synth  # noqa: E402

OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cfg1_golden.npz")


def pack_problem(pb):
    ob, ch = pb["objects"][0], pb["channels"][0]
    ns = pb["nullspace"]
    return dict(data=pb["data"], pupil=ob["pupil"], modes=ob["modes"], tilt=np.array(ob["tilt"]), J=ob["J"],
                Ji=ob["Ji"], scalars=np.array([pb["s"], pb["P"], pb["N"], pb["K"], pb["start"][0], pb["start"][1]]),
                optics=np.array([ob["weight"], ob["wavelength"], ob["freqCutoff"]]), phiFixed=ch["phiFixed"],
                imgCh=pb["imgCh"], ns_rows=ns["rows"], ns_cols=ns["cols"], ns_vals=ns["vals"],
                ns_sizes=np.array([ns["nPar"], ns["nFree"]]), truth_alpha=pb["truth"]["alpha"],
                settings=np.array(json.dumps(pb["settings"])))


def main():
    pb = synth.build_problem(synth.CFG1, seed=3, setup=S)
    st = O.State(pb)
    rng = np.random.default_rng(11)
    at = pb["truth"]["alpha"]
    alphas = np.stack([np.zeros_like(at), 0.5 * at, at * rng.standard_normal(at.shape) * 0.3,
                       at + 0.05 * rng.standard_normal(at.shape) * np.abs(at).mean()])
    metric = np.array([st.metric(a) for a in alphas])
    gv = np.stack([st.grad(a, method="vogel") for a in alphas])
    gf = np.stack([st.grad(a, method="fd") for a in alphas])
    img = np.stack([st.restore(a) for a in alphas])
    r = O.solve(pb)
    out = pack_problem(pb)
    out.update(alphas=alphas, metric=metric, grad_vogel=gv, grad_fd=gf, restored=img, solve_alpha=r["alpha"],
               solve_img=r["img"], solve_metrics=r["metrics"],
               solve_counts=np.array([r["iterations"], r["nF"], r["nDF"], r["status"]]),
               gamma=np.array([st.scalar("gamma")]), otfnorm=np.array([st.scalar("otfnorm")]))
    np.savez_compressed(OUT, **out)
    print(OUT, os.path.getsize(OUT))


if __name__ == "__main__":
    main()
