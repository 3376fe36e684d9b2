"""The oracle reproduces the committed golden vectors (regression pin of the checker)."""
import numpy as np

from tests.golden_problem import load


def test_oracle_reproduces_golden(oracle_lib):
    pb, z = load()
    st = oracle_lib.State(pb)
    for k, a in enumerate(z["alphas"]):
        assert abs(st.metric(a) - z["metric"][k]) <= 1e-12 * abs(z["metric"][k])
        g = st.grad(a, method="vogel")
        assert np.abs(g - z["grad_vogel"][k]).max() <= 1e-10 * np.abs(z["grad_vogel"][k]).max()
    assert abs(st.scalar("gamma") - z["gamma"][0]) < 1e-15
    r = oracle_lib.solve(pb)
    assert np.array_equal(r["alpha"], z["solve_alpha"])
    assert r["iterations"] == z["solve_counts"][0] and r["nF"] == z["solve_counts"][1]
