"""Seeded synthetic workloads (numpy front end of include/phylo_synth.h).

Host-only: nothing here touches the GPU.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from ._lib import lib, ptr
from .device import words


def _ok(rc: int, what: str) -> None:
    if rc != 0:
        raise RuntimeError(f"{what} failed ({rc})")


@dataclass
class Tree:
    n_leaves: int
    children: np.ndarray  # int32 [2(N-1)]
    root: int
    blen: np.ndarray      # float64 [2N-1]; blen[root] = 0


def yule(n_leaves: int, seed: int, mean_bl: float) -> Tree:
    N = int(n_leaves)
    ch = np.zeros(2 * (N - 1), np.int32)
    bl = np.zeros(2 * N - 1, np.float64)
    root = C.c_int32()
    _ok(lib.phy_synth_yule(N, seed, mean_bl, ptr(ch, C.c_int32), C.byref(root),
                           ptr(bl, C.c_double)), "phy_synth_yule")
    return Tree(N, ch, root.value, bl)


def fitch_leaves(n_leaves: int, n_sites: int, seed: int, stride: int | None = None) -> np.ndarray:
    """[N, stride] uint64 nibble blocks: 95% one base, 4% two-base, 1% gap."""
    W = words(n_sites)
    stride = W if stride is None else stride
    out = np.zeros((n_leaves, stride), np.uint64)
    _ok(lib.phy_synth_fitch_leaves(n_leaves, n_sites, seed, ptr(out, C.c_uint64), stride),
        "phy_synth_fitch_leaves")
    return out


def dirichlet(n: int, conc: float, seed: int) -> np.ndarray:
    x = np.zeros(n)
    _ok(lib.phy_synth_dirichlet(n, conc, seed, ptr(x, C.c_double)), "phy_synth_dirichlet")
    return x


def gtr_params(n_states: int, pi_conc: float, seed: int):
    """(pi, co): co = the n(n-1)/2 - 1 free exchangeabilities, last fixed 1."""
    pi = np.zeros(n_states)
    co = np.zeros(n_states * (n_states - 1) // 2 - 1)
    _ok(lib.phy_synth_gtr_params(n_states, pi_conc, seed, ptr(pi, C.c_double),
                                 ptr(co, C.c_double)), "phy_synth_gtr_params")
    return pi, co


def simulate(tree: Tree, pi, Q, rates, n_sites: int, seed: int, unique: bool = False):
    """[N, n_sites] uint8 state indices simulated down the tree."""
    pi = np.ascontiguousarray(pi, np.float64)
    Q = np.ascontiguousarray(Q, np.float64)
    rates = np.ascontiguousarray(rates, np.float64)
    out = np.zeros((tree.n_leaves, n_sites), np.uint8)
    _ok(lib.phy_synth_simulate(tree.n_leaves, ptr(tree.children, C.c_int32), tree.root,
                               ptr(tree.blen, C.c_double), pi.size, ptr(pi, C.c_double),
                               ptr(Q, C.c_double), rates.size, ptr(rates, C.c_double), n_sites,
                               seed, 1 if unique else 0, ptr(out, C.c_uint8)),
        "phy_synth_simulate")
    return out


def compress(states: np.ndarray):
    """Unique columns (first-occurrence order) and their multiplicities."""
    states = np.ascontiguousarray(states, np.uint8)
    N, S = states.shape
    pats = np.zeros((N, S), np.uint8)
    w = np.zeros(S, np.uint32)
    n = C.c_int64()
    _ok(lib.phy_synth_compress(N, S, ptr(states, C.c_uint8), ptr(pats, C.c_uint8),
                               ptr(w, C.c_uint32), C.byref(n)), "phy_synth_compress")
    return np.ascontiguousarray(pats[:, : n.value]), w[: n.value].copy()


def poisson_weights(n: int, lam: float, seed: int) -> np.ndarray:
    w = np.zeros(n, np.uint32)
    _ok(lib.phy_synth_poisson_weights(n, lam, seed, ptr(w, C.c_uint32)), "poisson")
    return w


def pack_dna(states: np.ndarray, stride: int | None = None) -> np.ndarray:
    """state indices (0..3; 255 = missing) -> nibble blocks [rows, stride]."""
    states = np.ascontiguousarray(states, np.uint8)
    rows, S = states.shape
    W = words(S)
    stride = W if stride is None else stride
    out = np.zeros((rows, stride), np.uint64)
    _ok(lib.phy_pack_dna(rows, S, ptr(states, C.c_uint8), S, ptr(out, C.c_uint64), stride),
        "phy_pack_dna")
    return out
