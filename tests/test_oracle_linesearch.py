"""Pins for the oracle's line search and conjugate-gradient stepper: math::bracket and
math::brent (src/lib/math/helpers.cpp:58-216) on one-dimensional functions whose minimisers
are known in closed form, and conjugate_rdx (src/lib/util/gsl/conjugate_gradient.cpp:128-256)
on convex quadratics, where Polak-Ribiere CG with exact line searches ends in at most n
iterations with mutually conjugate directions.  The solve uses the same bracket, brent and
cg_iterate code the tests drive here."""
import math

import numpy as np
import pytest

CASES = [
    ("parabola", lambda x: (x - 3.0) ** 2 + 1.0, 3.0),
    ("quartic", lambda x: (x + 2.5) ** 4, -2.5),
    ("abs", lambda x: abs(x - 0.7) + 2.0, 0.7),
    ("asymmetric", lambda x: math.exp(x - 1.2) - (x - 1.2), 1.2),      # minimum of e^t - t at t = 0
    ("far", lambda x: (x - 4.0e3) ** 2, 4.0e3),
]


@pytest.mark.parametrize("name,f,c", CASES, ids=[c[0] for c in CASES])
@pytest.mark.parametrize("a,b", [(0.0, 0.1), (0.1, 0.0), (0.0, -0.05), (-7.0, -6.5)])
def test_bracket_contains_minimum(oracle_lib, name, f, c, a, b):
    lo, mid, hi, flo, fmid, fhi = oracle_lib.bracket(f, a, b)
    assert lo <= mid <= hi
    # the returned values belong to the returned points
    for x, fx in ((lo, flo), (mid, fmid), (hi, fhi)):
        assert fx == f(x)
    # a bracket: the middle point is below both ends, and the minimiser lies inside
    assert fmid <= flo and fmid <= fhi
    assert lo <= c <= hi


@pytest.mark.parametrize("name,f,c", CASES, ids=[c[0] for c in CASES])
@pytest.mark.parametrize("tol", [1e-1, 1e-4, 1e-7])
def test_brent_converges_inside_tolerance(oracle_lib, name, f, c, tol):
    lo, mid, hi, *_ = oracle_lib.bracket(f, 0.0, 0.1)
    x, fx = oracle_lib.brent(f, lo, hi, mid, tol)
    assert fx == f(x)
    assert fx <= f(mid)
    # brent stops when [a, b] lies within 2 tol1 of x, tol1 = tol |x| + 1e-32, and the
    # minimiser of a unimodal function stays in [a, b]
    assert abs(x - c) <= 2.0 * (tol * abs(x) + 1e-32) * (1 + 1e-12) + 1e-15 * abs(c)


def test_brent_parabola_is_exact(oracle_lib):
    # on a parabola the parabolic steps land on the vertex, to the sqrt(eps) resolution
    # that rounding of f near a minimum allows
    f = lambda x: 2.0 * (x - 1.75) ** 2 - 4.0   # noqa: E731
    x, fx = oracle_lib.brent(f, 0.0, 5.0, 0.3, 1e-10)
    assert abs(x - 1.75) < 3.0 * math.sqrt(np.finfo(float).eps) * 1.75 and abs(fx + 4.0) < 1e-15


def test_brent_golden_section_on_kink(oracle_lib):
    # |x - c| has no parabolic model at the kink: brent must fall back to golden sections
    # and still shrink onto the kink
    c = 0.123456789
    f = lambda x: abs(x - c)   # noqa: E731
    x, fx = oracle_lib.brent(f, -1.0, 2.0, 1.5, 1e-9)
    assert abs(x - c) < 1e-8


def _quadratic(n, cond, seed, scale):
    rng = np.random.default_rng(seed)
    Qm, _ = np.linalg.qr(rng.standard_normal((n, n)))
    ev = np.geomspace(1.0, cond, n)
    A = scale * (Qm * ev) @ Qm.T
    xs = rng.standard_normal(n)

    def f(x):
        e = x - xs
        return 0.5 * e @ A @ e

    def g(x):
        return A @ (x - xs)

    return A, xs, f, g


# Conditioning is kept moderate: in floating point, conjugacy decays geometrically with
# the iteration count at a rate set by the condition number (the Lanczos loss of
# orthogonality), which is a property of CG, not of this restatement.
@pytest.mark.parametrize("n,cond,seed", [(6, 30.0, 1), (10, 10.0, 3), (12, 5.0, 4)])
def test_cg_terminates_on_quadratic(oracle_lib, n, cond, seed):
    # The first probe of each line search is 1e-21 along the unit direction
    # (conjugate_gradient.cpp:198): the quadratic is scaled so that it changes f.
    A, xs, f, g = _quadratic(n, cond, seed, scale=1e12)
    x0 = np.zeros(n)
    it = oracle_lib.cg_minimize(f, g, x0, step=1e-4, tol=1e-6, iters=n)
    g0 = np.linalg.norm(g(x0))
    fs = [r[1] for r in it]
    assert all(r[3] == 0 for r in it[1:])
    assert all(fs[k + 1] <= fs[k] for k in range(len(fs) - 1))          # monotone descent
    # finite termination: n iterations of PR-CG with exact line searches solve the quadratic
    x = it[-1][0]
    assert np.linalg.norm(g(x)) <= 1e-6 * g0
    assert np.linalg.norm(x - xs) <= 1e-5 * np.linalg.norm(xs)
    # the directions are A-conjugate
    P = [r[5] for r in it[1:]]
    for k in range(1, len(P)):
        for j in range(k):
            num = abs(P[k] @ A @ P[j])
            den = math.sqrt((P[k] @ A @ P[k]) * (P[j] @ A @ P[j]))
            assert num <= 1e-4 * den, (k, j, num / den)


def test_cg_first_direction_is_gradient_and_beta_is_polak_ribiere(oracle_lib):
    A, xs, f, g = _quadratic(8, 50.0, 7, scale=1e12)
    it = oracle_lib.cg_minimize(f, g, np.zeros(8), step=1e-4, tol=1e-6, iters=3)
    # the first direction is the gradient at the start (conjugate_gradient.cpp:155-160)
    assert np.allclose(it[1][5], g(it[0][0]), rtol=0, atol=1e-12 * np.abs(g(it[0][0])).max())
    # the second: p1 = g1 - beta p0 with beta = ((g0 - g1).g1)/|g0|^2 (:165-182); the sign
    # convention (x moves along -lambda p) makes this the Polak-Ribiere update
    g0, g1 = g(it[0][0]), g(it[1][0])
    beta = (g0 - g1) @ g1 / (g0 @ g0)
    assert np.allclose(it[2][5], g1 - beta * it[1][5], rtol=1e-12, atol=1e-9 * np.abs(g1).max())
    # the line search minimises f along the direction: the new gradient is orthogonal to it
    for k in (1, 2):
        gk = g(it[k][0])
        assert abs(gk @ it[k][5]) <= 1e-6 * np.linalg.norm(gk) * np.linalg.norm(it[k][5]) + 1e-9 * np.linalg.norm(g0) * np.linalg.norm(it[k][5])


def test_cg_beats_steepest_descent(oracle_lib):
    # a plausible mistake (a dropped beta term turns PR-CG into steepest descent) leaves
    # the gradient of an ill-conditioned quadratic far from zero after n iterations
    n = 10
    A, xs, f, g = _quadratic(n, 10.0, 3, scale=1e12)
    cg = oracle_lib.cg_minimize(f, g, np.zeros(n), step=1e-4, tol=1e-6, iters=n)
    sd = oracle_lib.step_minimize(1, f, g, np.zeros(n), step=1e-4, tol=1e-6, iters=n)
    g0 = np.linalg.norm(g(np.zeros(n)))
    assert np.linalg.norm(g(cg[-1][0])) <= 1e-6 * g0
    assert np.linalg.norm(g(sd[-1][0])) > 1e-2 * g0
    # steepest descent with exact line searches contracts the A-norm error by at most
    # (k-1)/(k+1) per step (k the condition number): after n steps that is still far above
    # what CG reaches
    assert ((10.0 - 1) / (10.0 + 1)) ** n > 1e3 * np.linalg.norm(g(cg[-1][0])) / g0
