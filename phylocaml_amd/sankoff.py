"""Sankoff parsimony on the GPU over integer transformation cost matrices.

* ``Tcm`` / ``create`` mirror PTcm.X (lib/pTcm.ml:5-50): an int matrix with
  its three properties, ``is_equal`` (every entry equals m[0][1]),
  ``is_symmetric`` and ``is_metric`` (non-negative and the triangle
  inequality), computed over every (i, j, k) as in lib/pTcm.ml:36-48, and the
  cost conventions of lib/pTcm.ml:15-24 (zero 0, addition, ``cost m x y =
  m[x][y]``).
* ``median_cost`` is ``find_median_pair`` over all atomic states
  (lib/costMatrix.ml:52-65): the cheapest k for two child states i, j and
  its cost m[i][k] + m[j][k].
* ``downpass`` runs ``phy_sankoff_downpass``: the whole tree in one launch,
  per-site costs and the weighted total.
* ``integerized_model`` lives in ``mlmodel`` (lib/mlModel.ml:639-660).
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from . import device as D
from ._lib import lib


@dataclass(frozen=True)
class Tcm:
    """PTcm.X.spec (lib/pTcm.ml:7-9); ``properties`` = (equal, symmetric, metric)."""

    matrix: np.ndarray
    properties: tuple

    @property
    def n_states(self) -> int:
        return int(self.matrix.shape[0])

    def cost(self, x: int, y: int) -> int:
        return int(self.matrix[x, y])

    def is_equal(self) -> bool:
        return self.properties[0]

    def is_symmetric(self) -> bool:
        return self.properties[1]

    def is_metric(self) -> bool:
        return self.properties[2]


def create(matrix) -> Tcm:
    """PTcm.X.create (lib/pTcm.ml:34-50)."""
    m = np.array(matrix, dtype=np.int64)
    if m.ndim != 2 or m.shape[0] != m.shape[1] or m.shape[0] < 2:
        raise ValueError("cost matrix must be square, at least 2 x 2")
    # triangle: m[i][j] <= m[i][k] + m[k][j] for every i, j, k; via[i, j, k] = m[i][k] + m[k][j]
    via = m[:, None, :] + m.T[None, :, :]
    tri = bool(np.all(m[:, :, None] <= via))
    positive = bool(np.all(m >= 0))
    symmetric = bool(np.array_equal(m, m.T))
    equal = bool(np.all(m == m[0, 1]))
    return Tcm(m, (equal, symmetric, positive and tri))


def median_cost(t: Tcm, i: int, j: int):
    """find_median_pair over every state (lib/costMatrix.ml:52-65): (cost,
    optimal medians) for child states i and j."""
    costs = t.matrix[i, :] + t.matrix[j, :]
    best = int(costs.min())
    return best, [int(k) for k in np.flatnonzero(costs == best)]


def check_bounds(t: Tcm, n_leaves: int) -> None:
    if np.any(t.matrix < 0):
        raise ValueError("sankoff: negative cost")
    if (2 * n_leaves - 2) * int(t.matrix.max()) >= 2 ** 31:
        raise ValueError("sankoff: costs too large for int32 sums on this tree")


def downpass(ctx: D.Context, topo: D.Topology, t: Tcm, leaves, n_sites: int, stride: int,
             weights=None):
    """Whole-tree Sankoff: leaves is a device uint8 [N, stride] tensor of
    state-set bitmasks (bit k = state k).  Returns (site_cost, total) with
    site_cost a device uint32 [n_sites] and total a device uint64 [1]."""
    torch = __import__("torch")
    K = t.n_states
    if K > 8:
        raise ValueError("sankoff: at most 8 states")
    check_bounds(t, topo.n_leaves)
    dev = f"cuda:{ctx.device}"
    tcm = torch.from_numpy(np.ascontiguousarray(t.matrix, np.int32)).to(dev)
    sc = torch.zeros((max(n_sites, 1),), dtype=torch.uint32, device=dev)
    tot = torch.zeros((1,), dtype=torch.uint64, device=dev)
    ctx.check(lib.phy_sankoff_downpass(ctx.handle, topo.handle, K, D.dptr(tcm), D.dptr(leaves),
                                       int(stride), int(n_sites), D.dptr(weights), D.dptr(sc),
                                       D.dptr(tot)), "sankoff_downpass")
    return sc, tot


def main(argv=None):
    """Time the whole-tree Sankoff pass (default 1,024 taxa x 10^6 sites,
    5 states: DNA with the gap as a fifth state, random integer costs with a
    zero diagonal); prints one JSON line."""
    import argparse
    import json

    from . import synth

    ap = argparse.ArgumentParser()
    ap.add_argument("--taxa", type=int, default=1024)
    ap.add_argument("--sites", type=int, default=1_000_000)
    ap.add_argument("--states", type=int, default=5)
    ap.add_argument("--steps", type=int, default=10)
    args = ap.parse_args(argv)
    torch = __import__("torch")
    N, S, K = args.taxa, args.sites, args.states
    rng = np.random.default_rng(1)
    tree = synth.yule(N, 1, 0.1)
    m = rng.integers(1, 20, (K, K))
    m = np.triu(m, 1) + np.triu(m, 1).T
    t = create(m)
    leaves = (1 << rng.integers(0, K, (N, S))).astype(np.uint8)
    amb = rng.random((N, S)) < 0.05
    leaves[amb] |= (1 << rng.integers(0, K, int(amb.sum()))).astype(np.uint8)
    ctx = D.Context(0)
    topo = D.Topology(N, tree.children, tree.root, ctx)
    dl = D.to_device(leaves, ctx.device)
    for _ in range(3):
        sc, tot = downpass(ctx, topo, t, dl, S, S)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(args.steps):
        sc, tot = downpass(ctx, topo, t, dl, S, S)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / args.steps
    out = {"workload": f"{N}-taxon Yule tree, {S} sites, {K} states, symmetric random integer "
                       "costs 1..19 off the diagonal, 5% two-state ambiguity",
           "ms": ms, "site_node_per_s": S * (N - 1) / (ms / 1e3),
           "leaf_bytes_per_s": N * S / (ms / 1e3), "total_cost": int(tot.cpu()[0]),
           "tcm_properties": list(t.properties)}
    print(json.dumps(out), flush=True)
    return out


__all__ = ["Tcm", "create", "median_cost", "check_bounds", "downpass"]


if __name__ == "__main__":
    main()
