"""ctypes wrapper of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py as the checker.  The product (phylocaml_amd)
never imports it.  Every function restates a reference algorithm; the C
source cites the reference file:line of each.

Parity status per function (the test that pins it):
  fitch_median2 / distance   tests/test_oracle_fitch.py (reference KAT of
                             test/costMatrixTest.ml:83-108, cost = distance
                             invariant of test/bitvectorTest.ml:59-69)
  fitch_downpass / uppass    tests/test_oracle_fitch.py (brute force, N <= 7)
  fitch_median3 / adjust3    tests/test_oracle_fitch.py (brute force per site)
  gtr_q / compose            tests/test_oracle_model.py (closed forms,
                             invariants, scipy goldens)
  lnl                        tests/test_oracle_lnl.py (brute force, 2-taxon
                             closed form, finite differences)
  sankoff                    tests/test_oracle_sankoff.py (brute force over all
                             assignments, unit costs = Fitch, cherry = the
                             reference median search)
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liboracle.so")


def build() -> str:
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return LIB_PATH


if not os.path.exists(LIB_PATH):
    build()

_lib = C.CDLL(LIB_PATH)
_u8 = C.POINTER(C.c_uint8)
_u32 = C.POINTER(C.c_uint32)
_u64 = C.POINTER(C.c_uint64)
_i32 = C.POINTER(C.c_int32)
_f64 = C.POINTER(C.c_double)

_lib.orc_fitch_median2.restype = C.c_double
_lib.orc_fitch_median2.argtypes = [_u8, _u8, _u8, C.c_int64, _f64]
_lib.orc_fitch_distance.restype = C.c_int64
_lib.orc_fitch_distance.argtypes = [_u8, _u8, C.c_int64]
_lib.orc_fitch_median3.restype = C.c_double
_lib.orc_fitch_median3.argtypes = [_u8, _u8, _u8, _u8, C.c_int64, _f64]
_lib.orc_fitch_adjust3.restype = None
_lib.orc_fitch_adjust3.argtypes = [_u8, _u8, _u8, _u8, _u8, C.c_int64]
_lib.orc_fitch_downpass.restype = C.c_uint64
_lib.orc_fitch_downpass.argtypes = [C.c_int32, _i32, C.c_int32, _u8, C.c_int64, _u32, _u32,
                                    _u64, C.c_int32]
_lib.orc_fitch_uppass.restype = None
_lib.orc_fitch_uppass.argtypes = [C.c_int32, _i32, C.c_int32, _u8, _u8, C.c_int64, C.c_int32]
_lib.orc_sankoff.restype = C.c_uint64
_lib.orc_sankoff.argtypes = [C.c_int32, _i32, C.c_int32, C.c_int32, _i32, _u8, C.c_int64, _u32,
                             _u32, C.c_int32]
_lib.orc_fitch_brute.restype = C.c_int64
_lib.orc_fitch_brute.argtypes = [C.c_int32, _i32, C.c_int32, _u8, C.c_int32, _u8]
_lib.orc_gtr_q.restype = None
_lib.orc_gtr_q.argtypes = [C.c_int32, _f64, _f64, _f64]
_lib.orc_compose.restype = None
_lib.orc_compose.argtypes = [C.c_int32, _f64, C.c_double, _f64]
_lib.orc_lnl.restype = C.c_double
_lib.orc_lnl.argtypes = [C.c_int32, _i32, C.c_int32, _f64, C.c_int32, _f64, _f64, C.c_int32,
                         _f64, _f64, _u32, C.c_int64, _u32, _f64, _f64, _f64, _f64, _f64,
                         C.c_int32]
_lib.orc_lnl_brute.restype = C.c_double
_lib.orc_lnl_brute.argtypes = [C.c_int32, _i32, C.c_int32, _f64, C.c_int32, _f64, _f64,
                               C.c_int32, _f64, _f64, _u32, C.c_int64, _u32]


def _p(a, t):
    return None if a is None else a.ctypes.data_as(t)


def _c(a, dt):
    return None if a is None else np.ascontiguousarray(a, dtype=dt)


def fitch_median2(x, y, w=None):
    x, y = _c(x, np.uint8), _c(y, np.uint8)
    out = np.empty_like(x)
    w = _c(w, np.float64)
    cost = _lib.orc_fitch_median2(_p(x, _u8), _p(y, _u8), _p(out, _u8), x.size, _p(w, _f64))
    return out, cost


def fitch_distance(a, b):
    a, b = _c(a, np.uint8), _c(b, np.uint8)
    return _lib.orc_fitch_distance(_p(a, _u8), _p(b, _u8), a.size)


def fitch_median3(a, b, c, w=None):
    a, b, c = (_c(v, np.uint8) for v in (a, b, c))
    out = np.empty_like(a)
    w = _c(w, np.float64)
    cost = _lib.orc_fitch_median3(_p(a, _u8), _p(b, _u8), _p(c, _u8), _p(out, _u8), a.size,
                                  _p(w, _f64))
    return out, cost


def fitch_adjust3(pv, pa, pb, fp):
    pv, pa, pb, fp = (_c(v, np.uint8) for v in (pv, pa, pb, fp))
    out = np.empty_like(pv)
    _lib.orc_fitch_adjust3(_p(pv, _u8), _p(pa, _u8), _p(pb, _u8), _p(fp, _u8), _p(out, _u8),
                           pv.size)
    return out


def fitch_downpass(n_leaves, children, root, leaves, weights=None, threads=1):
    """leaves: [N, S] uint8 sets.  Returns (sets [2N-1, S], site_cost, node_cost, total)."""
    N = int(n_leaves)
    leaves = _c(leaves, np.uint8)
    S = leaves.shape[1]
    sets = np.zeros((2 * N - 1, S), dtype=np.uint8)
    sets[:N] = leaves
    ch = _c(children, np.int32)
    sc = np.zeros(S, dtype=np.uint32)
    nc = np.zeros(N - 1, dtype=np.uint64)
    w = _c(weights, np.uint32)
    tot = _lib.orc_fitch_downpass(N, _p(ch, _i32), int(root), _p(sets, _u8), S, _p(w, _u32),
                                  _p(sc, _u32), _p(nc, _u64), int(threads))
    return sets, sc, nc, int(tot)


def sankoff(n_leaves, children, root, tcm, leaves, weights=None, threads=1):
    """leaves: [N, S] uint8 state-set bitmasks; tcm K x K ints.  Returns
    (site_cost, total)."""
    N = int(n_leaves)
    tcm = _c(tcm, np.int32)
    K = tcm.shape[0]
    leaves = _c(leaves, np.uint8)
    S = leaves.shape[1]
    ch = _c(children, np.int32)
    sc = np.zeros(S, dtype=np.uint32)
    w = _c(weights, np.uint32)
    tot = _lib.orc_sankoff(N, _p(ch, _i32), int(root), K, _p(tcm, _i32), _p(leaves, _u8), S,
                           _p(w, _u32), _p(sc, _u32), int(threads))
    return sc, int(tot)


def sankoff_brute(n_leaves, children, root, tcm, leaves):
    """Per-site minimum over every assignment of states to all nodes (leaf
    states within their sets) of sum over edges tcm[child][parent]; tiny
    trees only."""
    import itertools

    N = int(n_leaves)
    tcm = np.asarray(tcm, np.int64)
    K = tcm.shape[0]
    leaves = np.asarray(leaves, np.uint8)
    ch = np.asarray(children)
    M = 2 * N - 1
    out = np.zeros(leaves.shape[1], np.int64)
    for s in range(leaves.shape[1]):
        opts = []
        for v in range(M):
            if v < N:
                code = int(leaves[v, s]) & ((1 << K) - 1)
                code = code or (1 << K) - 1
                opts.append([k for k in range(K) if (code >> k) & 1])
            else:
                opts.append(list(range(K)))
        best = None
        for a in itertools.product(*opts):
            c = 0
            for v in range(N, M):
                for x in (ch[2 * (v - N)], ch[2 * (v - N) + 1]):
                    c += tcm[a[x], a[v]]
            best = c if best is None or c < best else best
        out[s] = best
    return out


def fitch_uppass(n_leaves, children, root, sets):
    N = int(n_leaves)
    sets = _c(sets, np.uint8)
    fin = np.zeros_like(sets)
    ch = _c(children, np.int32)
    _lib.orc_fitch_uppass(N, _p(ch, _i32), int(root), _p(sets, _u8), _p(fin, _u8),
                          sets.shape[1], 1)
    return fin


def fitch_brute(n_leaves, children, root, leaves, K=4):
    N = int(n_leaves)
    ch = _c(children, np.int32)
    lv = _c(leaves, np.uint8)
    mpr = np.zeros(2 * N - 1, dtype=np.uint8)
    best = _lib.orc_fitch_brute(N, _p(ch, _i32), int(root), _p(lv, _u8), int(K), _p(mpr, _u8))
    return int(best), mpr


def gtr_q(pi, co):
    pi = _c(pi, np.float64)
    co = _c(co, np.float64)
    n = pi.size
    Q = np.zeros((n, n))
    _lib.orc_gtr_q(n, _p(pi, _f64), _p(co, _f64), _p(Q, _f64))
    return Q


def compose(Q, t):
    Q = _c(Q, np.float64)
    n = Q.shape[0]
    P = np.zeros((n, n))
    _lib.orc_compose(n, _p(Q, _f64), float(t), _p(P, _f64))
    return P


def lnl(n_leaves, children, root, blen, pi, Q, rates, probs, tips, weights=None,
        derivs=False, threads=1, scales=False):
    """tips: [N, P] uint32 state bitmasks.  Returns (lnl, site_lnl[, d1, d2
    [, d1abs, d2abs]]): with ``scales``, also sum_p w_p |g_p| and
    sum_p w_p (|L''/L| + g^2) per branch (g = L'/L, h = L''/L - g^2: the
    magnitudes of the terms, for h of its two parts)."""
    N = int(n_leaves)
    ch = _c(children, np.int32)
    bl = _c(blen, np.float64)
    pi = _c(pi, np.float64)
    Q = _c(Q, np.float64)
    rates = _c(rates, np.float64)
    probs = _c(probs, np.float64)
    tips = _c(tips, np.uint32)
    P = tips.shape[1]
    w = _c(weights, np.uint32)
    sl = np.zeros(P)
    d1 = np.zeros(2 * N - 1) if derivs else None
    d2 = np.zeros(2 * N - 1) if derivs else None
    a1 = np.zeros(2 * N - 1) if derivs and scales else None
    a2 = np.zeros(2 * N - 1) if derivs and scales else None
    v = _lib.orc_lnl(N, _p(ch, _i32), int(root), _p(bl, _f64), pi.size, _p(pi, _f64),
                     _p(Q, _f64), rates.size, _p(rates, _f64), _p(probs, _f64), _p(tips, _u32),
                     P, _p(w, _u32), _p(sl, _f64), _p(d1, _f64), _p(d2, _f64), _p(a1, _f64),
                     _p(a2, _f64), int(threads))
    if derivs and scales:
        return v, sl, d1, d2, a1, a2
    if derivs:
        return v, sl, d1, d2
    return v, sl


def lnl_brute(n_leaves, children, root, blen, pi, Q, rates, probs, tips, weights=None):
    N = int(n_leaves)
    ch = _c(children, np.int32)
    bl = _c(blen, np.float64)
    pi = _c(pi, np.float64)
    Q = _c(Q, np.float64)
    rates = _c(rates, np.float64)
    probs = _c(probs, np.float64)
    tips = _c(tips, np.uint32)
    w = _c(weights, np.uint32)
    return _lib.orc_lnl_brute(N, _p(ch, _i32), int(root), _p(bl, _f64), pi.size, _p(pi, _f64),
                              _p(Q, _f64), rates.size, _p(rates, _f64), _p(probs, _f64),
                              _p(tips, _u32), tips.shape[1], _p(w, _u32))
