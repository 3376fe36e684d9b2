"""The adjoint gradient kernel behind a torch.autograd.Function (test harness only).

MetricFn.forward evaluates L(alpha) through mfbd_eval_metric, MetricFn.backward
returns dL/dalpha from mfbd_eval_grad(MFBD_GRAD_ADJOINT); torch only carries the
tensors.  Checked on the golden problem (config 1) and on a benchmark-geometry patch
(config 2) at several random alpha points:
  - against central differences of the oracle's FP64 metric, h = 1e-4 lambda
    (FP64: 1e-6 of max|g|, FP32: 1e-3);
  - against the reference's analytic (Vogel) gradient, which is not dL/dalpha (its
    weights carry an extra pupil factor at the rim pixels, subimage.cpp:387) but points
    the same way: cosine similarity > 0.9.
"""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

pytestmark = pytest.mark.gpu


class MetricFn(torch.autograd.Function):
    @staticmethod
    def forward(ctx, alpha, solver):
        a = alpha.detach().cpu().numpy()
        ctx.solver = solver
        ctx.save_for_backward(alpha)
        return torch.tensor(solver.metric(a[None])[0], dtype=torch.float64)

    @staticmethod
    def backward(ctx, gout):
        (alpha,) = ctx.saved_tensors
        g = ctx.solver.grad(alpha.detach().cpu().numpy()[None], method="adjoint")[0]
        return gout * torch.from_numpy(g), None


def _problems():
    from tests.golden_problem import load
    pb, _ = load()
    yield "cfg1", pb
    from oracle import setup_np
    from redux_amd import synth
    yield "cfg2", synth.build_problem(synth.CFG2_PATCH, seed=21, setup=setup_np)


@pytest.fixture(scope="module", params=["cfg1", "cfg2"])
def problem(request):
    return dict(_problems())[request.param]


@pytest.mark.parametrize("prec,tol", [("fp64", 1e-6), ("fp32", 1e-3)])
def test_autograd_adjoint_matches_central_differences(problem, oracle_lib, prec, tol):
    from redux_amd.momfbd import Solver
    sv = Solver(problem, precision=prec, max_batch=1)
    sv.load([problem])
    sv.init()
    st = oracle_lib.State(problem)
    rng = np.random.default_rng(1)
    at = problem["truth"]["alpha"]
    lam = problem["objects"][0]["wavelength"]
    h = 1e-4 * lam
    nImg, K = at.shape
    for trial in range(3):
        a0 = at * rng.uniform(-0.5, 1.0) + 0.1 * lam * rng.standard_normal(at.shape)
        alpha = torch.tensor(a0, dtype=torch.float64, requires_grad=True)
        L = MetricFn.apply(alpha, sv)
        (g,) = torch.autograd.grad(L, alpha)
        g = g.numpy()
        assert abs(L.item() - st.metric(a0)) <= (1e-12 if prec == "fp64" else 1e-6) * abs(L.item())
        picks = [(int(rng.integers(nImg)), int(rng.integers(K))) for _ in range(6)]
        for i, m in picks:
            ap, am = a0.copy(), a0.copy()
            ap[i, m] += h
            am[i, m] -= h
            gc = (st.metric(ap) - st.metric(am)) / (2 * h)
            assert abs(g[i, m] - gc) <= tol * np.abs(g).max(), (trial, i, m, g[i, m], gc)
        gv = st.grad(a0, method="vogel")
        cos = float(np.sum(g * gv) / (np.linalg.norm(g) * np.linalg.norm(gv)))
        assert cos > 0.9, cos
    sv.close()
