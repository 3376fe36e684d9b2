"""The oracle's restatements of GSL 1.16 steepest_descent and vector_bfgs (GETSTEP =
steepest_descent / BFGS, src/lib/momfbd/solver.cpp:355-372), pinned by their mathematics:
the published step rules on a hand-checkable first iteration, monotone descent,
convergence on convex quadratics and on Rosenbrock's function, and the restart of the
BFGS direction along the gradient every n iterations."""
import numpy as np
import pytest

from oracle import oracle as O


def quad(A, b):
    return (lambda x: 0.5 * x @ A @ x - b @ x), (lambda x: A @ x - b)


def test_steepest_descent_step_rules():
    """x1 = x0 - step g/|g|; an accepted first trial doubles the step, a rejected trial
    shrinks it by tol and the accepted one after a shrink shrinks it once more."""
    A = np.diag([1.0, 10.0])
    f, g = quad(A, np.zeros(2))
    x0 = np.array([1.0, 1.0])
    it = O.step_minimize(1, f, g, x0, step=0.1, tol=0.5, iters=1)
    g0 = A @ x0
    np.testing.assert_allclose(it[1][0], x0 - 0.1 * g0 / np.linalg.norm(g0), rtol=0, atol=1e-15)
    it2 = O.step_minimize(1, f, g, x0, step=0.1, tol=0.5, iters=2)
    d1 = it2[1][0] - it2[2][0]
    assert abs(np.linalg.norm(d1) - 0.2) < 1e-14       # the second step is twice the first
    # a step far too long is shrunk by tol while f rises: from x = 1 on f = x^2/2 the trials
    # 8 (x = -7) and 4 (x = -3) raise f, 2 (x = -1) does not (f equal, so accepted)
    it3 = O.step_minimize(1, f, g, np.array([1.0, 0.0]), step=8.0, tol=0.5, iters=1)
    np.testing.assert_allclose(it3[1][0], [-1.0, 0.0], atol=1e-15)
    assert it3[1][4]["f"] == 1 + 3                    # set, then trials at 8, 4, 2
    # after a shrink the accepted step shrinks once more (2 -> 1): the next iterate is 0
    it4 = O.step_minimize(1, f, g, np.array([1.0, 0.0]), step=8.0, tol=0.5, iters=2)
    np.testing.assert_allclose(it4[2][0], [0.0, 0.0], atol=1e-15)


@pytest.mark.parametrize("kind", [1, 3])
def test_descent_and_convergence_on_a_quadratic(kind):
    rng = np.random.default_rng(3)
    M = rng.standard_normal((6, 6))
    A = M @ M.T + 6 * np.eye(6)
    b = rng.standard_normal(6)
    f, g = quad(A, b)
    it = O.step_minimize(kind, f, g, np.zeros(6), step=0.1, tol=0.1 if kind == 3 else 0.5, iters=300)
    fs = [s[1] for s in it]
    assert all(b2 <= a2 + 1e-12 for a2, b2 in zip(fs, fs[1:]))
    xs = np.linalg.solve(A, b)
    best = min(it, key=lambda s: s[1])
    assert np.linalg.norm(best[0] - xs) <= (1e-5 if kind == 3 else 1e-3) * np.linalg.norm(xs)


def test_vector_bfgs_on_rosenbrock():
    f = lambda x: (1 - x[0]) ** 2 + 100 * (x[1] - x[0] ** 2) ** 2
    g = lambda x: np.array([-2 * (1 - x[0]) - 400 * x[0] * (x[1] - x[0] ** 2), 200 * (x[1] - x[0] ** 2)])
    it = O.step_minimize(3, f, g, np.array([-1.2, 1.0]), step=0.01, tol=0.1, iters=2000)
    best = min(it, key=lambda s: s[1])
    assert best[1] < 1e-8 and np.linalg.norm(best[0] - 1.0) < 1e-3


def test_vector_bfgs_restarts_along_the_gradient():
    """On an isotropic quadratic every line search is exact along -g, so x reaches the
    minimum in one line minimisation (the BFGS update then never matters)."""
    f, g = quad(np.eye(3), np.array([1.0, -2.0, 0.5]))
    it = O.step_minimize(3, f, g, np.zeros(3), step=10.0, tol=1e-8, iters=5)
    best = min(it, key=lambda s: s[1])
    np.testing.assert_allclose(best[0], [1.0, -2.0, 0.5], atol=1e-7)
