"""The wave-local FP32 OTF kernel (k_otf_wave, N = 128, the default) against the
LDS-staged one (k_otf_fused, MFBD_OTF_WAVE=0): the same butterflies, twiddles and order
except on the OTF columns 0 and 64, which k_otf_wave transforms as one complex line, so
metrics and Vogel gradients on config-2 patches agree to float rounding, and solves with
either kernel converge to objects as close as two FP32 solves get; the statistics show
which kernel ran.  The solve-level parity of the FP32 form against the oracle runs with
k_otf_wave in the other FP32 tests."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _batch(n, seed):
    from redux_amd import synth
    return synth.build_field_batch(synth.CFG2_PATCH, n, seed=seed)[1]


def _run(monkeypatch, pbs, wave, solve):
    from redux_amd.momfbd import Solver
    monkeypatch.setenv("MFBD_OTF_WAVE", "1" if wave else "0")
    monkeypatch.setenv("MFBD_KERNEL_TIMING", "1")
    sv = Solver(pbs[0], precision="fp32", max_batch=len(pbs))
    sv.load(pbs)
    sv.init()
    rng = np.random.default_rng(3)
    a = np.stack([p["truth"]["alpha"] * 0.5 for p in pbs]) + 0.01 * rng.standard_normal((len(pbs), sv.nImg, sv.K))
    out = dict(m0=sv.metric(np.zeros_like(a)), m=sv.metric(a), g=sv.grad(a, method="vogel"))
    if solve:
        out["res"] = sv.run()
    out["stats"] = sv.kernel_stats()
    sv.close()
    return out


def _close(wav, lds):
    """Within the FP32 form's bounds against the oracle (metric 1e-6 relative, Vogel gradient
    1e-4 relative 2-norm per patch), which each kernel meets on its own; measured 2e-8 and
    ~3e-5 apart"""
    for k in ("m0", "m"):
        d = np.max(np.abs(wav[k] - lds[k]) / np.abs(lds[k]))
        print(k, "relative", d, flush=True)
        assert d < 1e-6, k
    d = max(np.linalg.norm(x - y) / np.linalg.norm(y) for x, y in zip(wav["g"], lds["g"]))
    print("g relative 2-norm", d, flush=True)
    assert d < 1e-4


def test_wave_kernel_against_lds_kernel(monkeypatch):
    pbs = _batch(8, 11)
    print("k_otf_wave vs k_otf_fused: 8 config-2 patches", flush=True)
    lds = _run(monkeypatch, pbs, False, True)
    wav = _run(monkeypatch, pbs, True, True)
    assert "k_otf_fused" in lds["stats"] and "k_otf_wave" not in lds["stats"]
    assert "k_otf_wave" in wav["stats"] and "k_otf_fused" not in wav["stats"]
    _close(wav, lds)
    for x, y in zip(wav["res"], lds["res"]):
        assert x["status"] == y["status"] == 0
        d = np.linalg.norm(x["img"] - y["img"]) / np.linalg.norm(y["img"])
        print("solve: objects", d, "iterations", x["iterations"], y["iterations"], flush=True)
        assert d < 1e-2


def test_wave_kernel_small_batch(monkeypatch):
    """Fewer images than persistent workgroups (3 patches, 144 images)."""
    pbs = _batch(3, 5)
    lds = _run(monkeypatch, pbs, False, False)
    wav = _run(monkeypatch, pbs, True, False)
    _close(wav, lds)
