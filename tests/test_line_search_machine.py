"""The device line search's state machine (redux_amd/csrc/line_search.hpp, run by k_ls_step
one round of probes at a time) against the oracle's math::bracket + math::brent as
conjugate_rdx_iterate calls them (conjugate_gradient.cpp:196-231; helpers.cpp:58-216): from
the line's metric fa and lambda, the first probe at 1e-21 / lambda, bracket(0, stepc, limit
100), brent(tol 0.1, limit 100).  The machine is built for the host here (g++, no
contraction) and must ask for the oracle's probes in the oracle's order, bit for bit, and
end at the oracle's point and value: on smooth lines, on lines with rounding-level noise
(ties and near-ties between probes), on flat and rising lines and on one that turns NaN."""
import ctypes as C
import math
import pathlib
import subprocess

import numpy as np
import pytest

ROOT = pathlib.Path(__file__).resolve().parent.parent
CB = C.CFUNCTYPE(C.c_double, C.c_double)


@pytest.fixture(scope="module")
def machine(tmp_path_factory):
    out = tmp_path_factory.mktemp("ls") / "libls_machine.so"
    subprocess.run(["g++", "-O2", "-ffp-contract=off", "-std=c++17", "-shared", "-fPIC", "-D__host__=",
                    "-D__device__=", "-I", str(ROOT / "redux_amd" / "csrc"), str(ROOT / "tests" / "ls_machine_shim.cpp"),
                    "-o", str(out)], check=True)
    lib = C.CDLL(str(out))
    lib.ls_trace.restype = C.c_int
    lib.ls_trace.argtypes = [C.c_double, C.c_double, CB, C.POINTER(C.c_double), C.c_int, C.POINTER(C.c_double)]
    return lib


def device_line(lib, f, fa, lam):
    steps = (C.c_double * 4096)()
    out = (C.c_double * 2)()
    cb = CB(lambda t: float(f(t)))
    n = lib.ls_trace(fa, lam, cb, steps, 4096, out)
    assert n >= 0, "the line did not end within 4096 probes"
    return [steps[k] for k in range(n)], out[0], out[1]


def oracle_line(O, f, lam):
    log = []

    def g(t):
        log.append(t)
        return f(t)
    stepc = 1e-21 / lam
    lo, mid, hi, *_ = O.bracket(g, 0.0, stepc)
    x, fx = O.brent(g, lo, hi, mid, 0.1)
    assert log[0] == 0.0 and log[1] == stepc      # f(0) = fa and the first probe
    return log[1:], x, fx


def _lines():
    rng = np.random.default_rng(11)
    out = []
    for k in range(60):
        lam = (1.0 if k % 2 else -1.0) / 10 ** rng.uniform(-2, 4)
        s0, c2 = rng.uniform(-5, 5) * 10 ** rng.uniform(-3, 1), 10 ** rng.uniform(-2, 3)
        c4, base = rng.uniform(0, 2), rng.uniform(1, 1e4)

        def f(t, lam=lam, s0=s0, c2=c2, c4=c4, base=base):   # a metric along the line, s = -t lambda
            s = -t * lam - s0
            return base + c2 * s * s + c4 * s ** 4
        out.append((f"poly{k}", f, lam))
    for k in range(30):
        lam = -1.0 / 10 ** rng.uniform(-1, 3)
        s0, amp = rng.uniform(-2, 2), 10.0 ** rng.integers(-14, -9)

        def f(t, lam=lam, s0=s0, amp=amp):
            s = -t * lam - s0
            v = 1.0 + 0.5 * s * s
            # rounding-level noise, deterministic in t: probes that tie or nearly tie
            h = (hash(float(t)) & 0xFFFF) / 65535.0 - 0.5
            return v * (1.0 + amp * h)
        out.append((f"noisy{k}", f, lam))
    out.append(("flat", lambda t: 3.0, 0.25))
    out.append(("rising", lambda t: 2.0 + abs(t), -4.0))
    out.append(("kink", lambda t: abs(t * 0.5 + 1.3), 0.5))
    out.append(("nan_far", lambda t: (t - 2.0) ** 2 if abs(t) < 50 else math.nan, 1.0 / 3))
    out.append(("steep_exp", lambda t: math.exp(-t) + 0.01 * t * t, -1.0 / 7))
    return out


LINES = _lines()


@pytest.mark.parametrize("name,f,lam", LINES, ids=[x[0] for x in LINES])
def test_machine_takes_the_oracle_probes(machine, oracle_lib, name, f, lam):
    want, wx, wf = oracle_line(oracle_lib, f, lam)
    got, x, fx = device_line(machine, f, f(0.0), lam)
    assert len(got) == len(want), (len(got), len(want))
    assert all(a == b or (math.isnan(a) and math.isnan(b)) for a, b in zip(got, want))
    assert x == wx and (fx == wf or (math.isnan(fx) and math.isnan(wf)))


def test_machine_probe_counts_span_both_phases(machine):
    """A smooth line whose minimum lies a few first steps out (the first probe is 1e-21 /
    lambda: the metric's scale makes such steps count) takes the first probe, the bracket's
    middle point, brent's re-evaluation of it and brent's own probes, and ends within
    brent's tolerance (0.1 relative) of the minimum."""
    f = lambda t: 1.0 + (t / 1e-21 - 5.0) ** 2   # noqa: E731
    got, x, fx = device_line(machine, f, f(0.0), 1.0)
    assert len(got) >= 4 and got[0] == 1e-21 and abs(x - 5e-21) < 0.2 * 5e-21
