"""Pins the model oracle (Q, P(t)) against closed forms, invariants and scipy
goldens, and checks the library's host model code (C ABI replacements of
lib/mlmodel.c and MlModel helpers) against them."""
import json
import os

import numpy as np
import pytest

from conftest import GOLDEN
from oracle import oracle as O

G = json.load(open(os.path.join(GOLDEN, "model_goldens.json")))


def _jc(n, t):
    P = np.full((n, n), 1.0 / n - np.exp(-n * t / (n - 1)) / n)
    np.fill_diagonal(P, 1.0 / n + (n - 1) / n * np.exp(-n * t / (n - 1)))
    return P


@pytest.mark.parametrize("n", [4, 20])
def test_gtr_q_jc_closed_form(n):
    pi = np.full(n, 1.0 / n)
    Q = O.gtr_q(pi, np.ones((n + 1) * (n - 2) // 2))
    for t in (0.01, 0.3, 2.0):
        assert np.allclose(O.compose(Q, t), _jc(n, t), rtol=0, atol=1e-14)


def test_f81_closed_form(rng):
    pi = rng.dirichlet(np.ones(4) * 3)
    Q = O.gtr_q(pi, np.ones(5))
    beta = 1.0 / (1.0 - (pi ** 2).sum())  # after mean-rate normalisation
    for t in (0.05, 0.5, 4.0):
        P = pi[None, :] + (np.eye(4) - pi[None, :]) * np.exp(-beta * t)
        assert np.allclose(O.compose(Q, t), P, rtol=0, atol=1e-14)


def test_gtr_q_matches_golden_definition():
    for m in G["models"]:
        Q = O.gtr_q(np.array(m["pi"]), np.array(m["co"]))
        assert np.allclose(Q, np.array(m["Q"]), rtol=1e-14, atol=1e-15)


def test_compose_matches_scipy_expm():
    for m in G["models"]:
        Q = np.array(m["Q"])
        for t, P in m["P"].items():
            Pc = O.compose(Q, float(t))
            assert np.allclose(Pc, np.array(P), rtol=1e-12, atol=1e-15), t


def test_compose_special_cases():
    Q = np.array(G["models"][0]["Q"])
    assert np.array_equal(O.compose(Q, -1.0), Q)          # lib/mlmodel.c:332-334
    for t in (0.0, 5e-11, -0.5):
        assert np.array_equal(O.compose(Q, t), np.eye(4))  # lib/mlmodel.c:339-341


def test_p_invariants():
    m = G["models"][2]
    pi, Q = np.array(m["pi"]), np.array(m["Q"])
    Ps, Pt, Pst = O.compose(Q, 0.3), O.compose(Q, 0.45), O.compose(Q, 0.75)
    assert np.allclose(Ps @ Pt, Pst, atol=1e-14)                     # Chapman-Kolmogorov
    assert np.allclose(Pst.sum(1), 1.0, atol=1e-14)                  # rows sum to one
    assert np.allclose(pi[:, None] * Pst, (pi[:, None] * Pst).T, atol=1e-15)  # reversibility
    assert np.allclose(pi @ Q, 0.0, atol=1e-14)                      # stationarity
    assert np.allclose(O.compose(Q, 200.0), np.tile(pi, (20, 1)), atol=1e-12)
    assert abs(-(np.diag(Q) * pi).sum() - 1.0) < 1e-14              # mean rate 1


# ---- the library's host model code against the pinned oracle ----
from phylocaml_amd import mlmodel as MM  # noqa: E402


def test_library_gtr_q_and_compose_match_oracle():
    for m in G["models"]:
        pi, co = np.array(m["pi"]), np.array(m["co"])
        Q = MM.gtr_q(pi, co)
        assert np.allclose(Q, O.gtr_q(pi, co), rtol=1e-15, atol=1e-16)
        U, D, Ui = MM.diagonalize_gtr(Q)
        assert np.allclose(U @ np.diag(np.diag(D)) @ Ui, Q, atol=1e-13)
        model = MM.Model(spec=None, priors=pi, rates=np.ones(1), probs=np.ones(1), q=Q, u=U,
                         d=D, ui=Ui)
        for t in (1e-6, 0.01, 0.1, 0.7, 3.0, 25.0, -1.0, 0.0, 1e-11):
            assert np.allclose(MM.compose(model, t), O.compose(Q, t), rtol=1e-11, atol=1e-14), t


def test_library_rejects_nonreversible_and_nan():
    Q = np.array([[-1.0, 1.0, 0.0], [0.0, -1.0, 1.0], [1.0, 0.0, -1.0]])  # cyclic: complex spectrum
    with pytest.raises(MM.Error):
        MM.diagonalize_gtr(Q)
    Qn = np.array(G["models"][0]["Q"])
    Qn[1, 2] = np.nan
    with pytest.raises(MM.Error):
        MM.diagonalize_gtr(Qn)


def test_library_gamma_rates_match_scipy():
    for g in G["gamma"]:
        a, n = g["alpha"], g["n"]
        ref = MM.gamma_rates(a, a, n, MM.PHY_GAMMA_REF)
        assert np.allclose(ref, g["ref"], rtol=1e-10, atol=1e-300), (a, n)
        mean = MM.gamma_rates(a, a, n, MM.PHY_GAMMA_MEAN)
        assert np.allclose(mean, g["mean"], rtol=1e-9, atol=1e-12), (a, n)
        assert abs(mean.mean() - 1.0) < 1e-9          # test/mlModelTest.ml:44-60 intent
        assert ref[0] == 0.0                          # quantile(0) in the reference


def test_library_symmetric_path_float_t():
    n = 4
    A = O.gtr_q(np.full(n, 0.25), np.ones(5))
    import ctypes as C
    from phylocaml_amd._lib import lib, ptr
    U = A.copy()
    D = np.zeros((n, n))
    assert lib.phy_diagonalize_sym(n, ptr(U, C.c_double), ptr(D, C.c_double)) == 0
    P = np.zeros((n, n))
    t = 0.1234567891234
    assert lib.phy_compose_sym(n, ptr(U, C.c_double), ptr(D, C.c_double), t, ptr(P, C.c_double)) == 0
    # compose_sym rounds t to single precision (lib/mlmodel.c:280)
    assert np.allclose(P, _jc(n, float(np.float32(t))), atol=1e-14)


def test_create_spec_gtr_gamma():
    spec = MM.Spec(4, ("GTR", [1.0, 2.0, 0.5, 0.3, 4.0]), ("DiscreteGamma", 4, 0.5),
                   ("Empirical", [0.1, 0.2, 0.3, 0.4]))
    m = MM.create(spec)
    assert m.rates.size == 4 and abs(m.probs.sum() - 1) < 1e-15
    assert np.allclose(m.priors, [0.1, 0.2, 0.3, 0.4])
    Q = O.gtr_q(m.priors, np.array([1.0, 2.0, 0.5, 0.3, 4.0]))
    assert np.allclose(m.q, Q, atol=1e-15)
    with pytest.raises(MM.Error):
        MM.create(MM.Spec(4, ("GTR", [1.0, 2.0]), ("Constant",), ("Equal",)))
