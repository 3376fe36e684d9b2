"""Rebuilds the problem dictionary stored in tests/golden/cfg1_golden.npz."""
import json
import os

import numpy as np

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "cfg1_golden.npz")


def load():
    z = np.load(GOLDEN, allow_pickle=False)
    s, P, N, K, sy, sx = (int(v) for v in z["scalars"])
    w, wl, fc = (float(v) for v in z["optics"])
    obj = dict(weight=w, wavelength=wl, pupil=z["pupil"], modes=z["modes"], tilt=tuple(int(t) for t in z["tilt"]),
               J=z["J"], Ji=z["Ji"], freqCutoff=fc)
    ch = dict(obj=0, phiFixed=z["phiFixed"], resOffset=(0.0, 0.0), noiseFudge=1.0)
    nPar, nFree = (int(v) for v in z["ns_sizes"])
    pb = dict(s=s, P=P, N=N, K=K, objects=[obj], channels=[ch], imgCh=z["imgCh"], data=z["data"], start=(sy, sx),
              nullspace=dict(nPar=nPar, nFree=nFree, rows=z["ns_rows"], cols=z["ns_cols"], vals=z["ns_vals"]),
              settings=json.loads(str(z["settings"])), truth=dict(alpha=z["truth_alpha"]))
    return pb, z
