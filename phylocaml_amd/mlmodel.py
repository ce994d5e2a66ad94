"""MlModel: substitution models for the likelihood path (lib/mlModel.ml).

A host-side mirror of the parts of the reference's model layer that pruning
consumes -- model specification, priors, Gamma rate categories, rate matrix,
eigen-decomposition and P(t) -- with the native stubs (diagonalize_gtr,
compose_gtr, ...) served by libphylo_hip.so instead of LAPACK.
"""
from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass, field

import numpy as np

from ._lib import PHY_GAMMA_MEAN, PHY_GAMMA_REF, check, lib, ptr

MINIMUM = 1e-13  # lib/mlModel.ml:9


class Error(Exception):
    """MlModel.Error (lib/mlModel.ml:3)."""


@dataclass
class Spec:
    """MlModel.s (lib/mlModel.ml:38-44) restricted to what pruning uses.

    substitution: ("JC69",) | ("F81",) | ("GTR", co)  with co the
        (n+1)(n-2)/2 free exchangeabilities (the last one is fixed at 1).
    site_variation: ("Constant",) | ("DiscreteGamma", n, alpha) |
        ("DiscreteCustom", [(rate, prob), ...]).
    base_priors: ("Equal",) | ("Empirical", pi).
    gamma_mode: PHY_GAMMA_REF reproduces gamma_rates exactly
        (lib/mlModel.ml:93-99: quantile(i/n) of Gamma(alpha, scale alpha));
        PHY_GAMMA_MEAN gives Yang's mean-1 categories, the documented intent
        (lib/mlModel.mli:12).
    """

    n_states: int
    substitution: tuple = ("GTR", ())
    site_variation: tuple = ("Constant",)
    base_priors: tuple = ("Equal",)
    gamma_mode: int = PHY_GAMMA_MEAN


@dataclass
class Model:
    """MlModel.t (lib/mlModel.ml:53-63)."""

    spec: Spec
    priors: np.ndarray
    rates: np.ndarray
    probs: np.ndarray
    q: np.ndarray
    u: np.ndarray
    d: np.ndarray
    ui: np.ndarray
    pinvar: float | None = None
    extra: dict = field(default_factory=dict)


def gamma_rates(a: float, b: float, n: int, mode: int = PHY_GAMMA_REF) -> np.ndarray:
    """MlModel.gamma_rates (lib/mlModel.ml:93-99) in PHY_GAMMA_REF mode."""
    out = np.zeros(n)
    check(lib.phy_gamma_rates(int(mode), float(a), float(b), int(n), ptr(out, C.c_double)),
          None, "gamma_rates")
    return out


def gtr_q(pi, co) -> np.ndarray:
    """m_gtr_independent + m_meanrate (lib/mlModel.ml:390-421, 204-213)."""
    pi = np.ascontiguousarray(pi, np.float64)
    n = pi.size
    co = np.ascontiguousarray(co, np.float64)
    if co.size != (n + 1) * (n - 2) // 2:
        raise Error(f"Length of GTR parameters (I) is incorrect for the alphabet. "
                    f"They should be {(n + 1) * (n - 2) // 2}, but are {co.size}.")
    Q = np.zeros((n, n))
    check(lib.phy_gtr_q(n, ptr(pi, C.c_double), ptr(co, C.c_double), ptr(Q, C.c_double)),
          None, "gtr_q")
    return Q


def diagonalize_gtr(q: np.ndarray):
    """likelihood_CAML_diagonalize_gtr replacement: (U, D, Ui), Q = U D Ui."""
    q = np.ascontiguousarray(q, np.float64)
    n = q.shape[0]
    U, Dm, Ui = np.zeros((n, n)), np.zeros((n, n)), np.zeros((n, n))
    rc = lib.phy_diagonalize_gtr(n, ptr(q, C.c_double), ptr(U, C.c_double), ptr(Dm, C.c_double),
                                 ptr(Ui, C.c_double))
    if rc != 0:
        raise Error("Imaginary eigenvalues or invalid rate matrix")
    return U, Dm, Ui


def compose(model: Model, t: float) -> np.ndarray:
    """MlModel.compose (lib/mlModel.ml:585-587): P = exp(Q t)."""
    n = model.q.shape[0]
    P = np.zeros((n, n))
    check(lib.phy_compose_gtr(n, ptr(model.u, C.c_double), ptr(model.d, C.c_double),
                              ptr(model.ui, C.c_double), float(t), ptr(P, C.c_double)),
          None, "compose_gtr")
    return P


def substitution_matrix(model: Model, t: float | None = None) -> np.ndarray:
    """MlModel.substitution_matrix (lib/mlModel.ml:589-620): Q, or Q t."""
    q = model.q.copy()
    return q * t if t is not None else q


def _priors(spec: Spec) -> np.ndarray:
    n = spec.n_states
    if spec.base_priors[0] == "Equal":
        return np.full(n, 1.0 / n)
    p = np.maximum(np.asarray(spec.base_priors[1], np.float64), MINIMUM)
    s = p.sum()
    if p.size != n:
        raise Error(f"Priors (length {p.size}) don't match alphabet (length {n})")
    return p if abs(s - 1.0) <= 2.220446049250313e-16 * 4 else p / s


def _site_rates(spec: Spec):
    sv = spec.site_variation
    if sv[0] == "Constant":
        return np.array([1.0]), np.array([1.0])
    if sv[0] == "DiscreteGamma":
        k, alpha = int(sv[1]), float(sv[2])
        return gamma_rates(alpha, alpha, k, spec.gamma_mode), np.full(k, 1.0 / k)
    if sv[0] == "DiscreteCustom":
        return (np.array([r for r, _ in sv[1]], np.float64),
                np.array([p for _, p in sv[1]], np.float64))
    raise Error(f"unsupported site variation {sv[0]}")


def create(spec: Spec) -> Model:
    """MlModel.create (lib/mlModel.ml:664-746) for JC69 / F81 / GTR."""
    n = spec.n_states
    rates, probs = _site_rates(spec)
    priors = _priors(spec)
    sub = spec.substitution
    if sub[0] == "JC69":
        q = gtr_q(np.full(n, 1.0 / n), np.ones((n + 1) * (n - 2) // 2))
    elif sub[0] == "F81":
        q = gtr_q(priors, np.ones((n + 1) * (n - 2) // 2))
    elif sub[0] == "GTR":
        co = np.asarray(sub[1], np.float64)
        if co.size == 0:
            co = np.ones((n - 2) * (n + 1) // 2)
        q = gtr_q(priors, co)
    else:
        raise Error(f"unsupported substitution model {sub[0]}")
    if np.isnan(q).any():
        raise Error("NaN in rate matrix")
    u, d, ui = diagonalize_gtr(q)
    return Model(spec=spec, priors=priors, rates=rates, probs=probs, q=q, u=u, d=d, ui=ui)


def _same(a, b) -> bool:
    if isinstance(a, (tuple, list)) and isinstance(b, (tuple, list)):
        return len(a) == len(b) and all(_same(x, y) for x, y in zip(a, b))
    if isinstance(a, np.ndarray) or isinstance(b, np.ndarray):
        return np.array_equal(np.asarray(a), np.asarray(b))
    return a == b


def update(model: Model, spec: Spec) -> Model:
    """MlModel.update (lib/mlModel.ml:196, TODO there): the model of spec with
    minimal work -- when only the site-rate variation changed, the rate matrix
    and its eigen-decomposition are kept and only the rate categories are
    recomputed; otherwise the model is created anew."""
    if (spec.n_states == model.spec.n_states and _same(spec.substitution, model.spec.substitution)
            and _same(spec.base_priors, model.spec.base_priors)):
        rates, probs = _site_rates(spec)
        return Model(spec=spec, priors=model.priors, rates=rates, probs=probs, q=model.q,
                     u=model.u, d=model.d, ui=model.ui)
    return create(spec)


def _with(spec: Spec, **kw) -> Spec:
    d = dict(spec.__dict__)
    d.update(kw)
    return Spec(**d)


def gen_subst_opt_func(m: Model):
    """MlModel.gen_subst_opt_func (lib/mlModel.ml:821-822, TODO there): the
    GTR exchangeabilities (the last one fixed at 1) as a log-scale vector, and
    the function from such a vector to the model.  Models without free
    substitution parameters give an empty vector."""
    sub = m.spec.substitution
    if sub[0] != "GTR":
        return np.zeros(0), lambda v: m
    n = m.spec.n_states
    co = np.asarray(sub[1], np.float64)
    if co.size == 0:
        co = np.ones((n + 1) * (n - 2) // 2)
    return np.log(co), lambda v: update(m, _with(m.spec, substitution=("GTR", np.exp(v))))


def gen_rates_opt_func(m: Model):
    """MlModel.gen_rates_opt_func (lib/mlModel.ml:824-825, TODO there): the
    Gamma shape as [log alpha]; other site-rate models give an empty vector."""
    sv = m.spec.site_variation
    if sv[0] != "DiscreteGamma":
        return np.zeros(0), lambda v: m
    k = int(sv[1])
    return (np.array([math.log(float(sv[2]))]),
            lambda v: update(m, _with(m.spec, site_variation=("DiscreteGamma", k,
                                                                 float(np.exp(v[0]))))))


def gen_prior_opt_func(m: Model):
    """MlModel.gen_prior_opt_func (lib/mlModel.ml:827-828, TODO there): the
    empirical base frequencies as n - 1 log ratios log(pi_i / pi_n); equal
    priors give an empty vector."""
    if m.spec.base_priors[0] != "Empirical":
        return np.zeros(0), lambda v: m
    pi = m.priors

    def make(v):
        x = np.append(np.asarray(v, np.float64), 0.0)
        e = np.exp(x - x.max())
        return update(m, _with(m.spec, base_priors=("Empirical", e / e.sum())))

    return np.log(pi[:-1] / pi[-1]), make


def integerized_model(model: Model, t: float, sigma: int = 4) -> np.ndarray:
    """MlModel.integerized_model (lib/mlModel.ml:639-660): the integer cost
    matrix of P(t) scaled by 10^sigma, -int(10^sigma log P_ij) for JC69 and
    -int(10^sigma log(pi_i P_ij)) otherwise (OCaml's int_of_float truncates
    toward zero)."""
    s = 10.0 ** int(sigma)
    p = compose(model, t)
    with np.errstate(divide="ignore"):
        x = np.log(p) if model.spec.substitution[0] == "JC69" else \
            np.log(model.priors[:, None] * p)
    return -np.trunc(s * x).astype(np.int64)


def num_parameters(m: Model) -> int:
    """MlModel.num_parameters (lib/mlModel.ml:193, TODO there): free model
    parameters (substitution, site rates, priors), as used by the
    information criteria; branch lengths are not counted."""
    return sum(g(m)[0].size for g in (gen_subst_opt_func, gen_rates_opt_func, gen_prior_opt_func))


__all__ = ["Spec", "Model", "Error", "create", "compose", "gamma_rates", "gtr_q",
           "diagonalize_gtr", "substitution_matrix", "integerized_model", "update", "gen_subst_opt_func",
           "gen_rates_opt_func", "gen_prior_opt_func", "num_parameters", "PHY_GAMMA_REF",
           "PHY_GAMMA_MEAN"]
