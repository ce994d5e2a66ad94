"""Writes the golden fixtures of tests/golden/ (run once, in a container with
scipy; the outputs are committed).

reference_kats.json -- known-answer data transcribed from the reference's own
    tests and alphabet tables:
      test/alphabetTest.ml:12-18   DNA codes A=1 C=2 G=4 T=8 -=16
      lib/alphabet.ml:301-307      Alphabet.dna states (A C G T - X, BitFlag)
      test/sequenceTest.ml:6-9     "-AGGTCGTCCGTCGGAAT--A" <-> codes
      lib/alphabet.ml:312-322      IUPAC ambiguity codes of Alphabet.nucleotides
      test/costMatrixTest.ml:83-108 the Fitch pair rule checked over every
                                   pair of Alphabet.dna codes (listed here as
                                   explicit (x, y, median, cost) rows, each
                                   computed with Python sets of state names)
      test/treeTest.ml:221-255     Tree.num_unrooted_trees, num_rooted_trees,
                                   num_nodes and num_edges for n = 0 .. 12 / 10
model_goldens.json -- scipy reference values:
      scipy.linalg.expm(Q t) for GTR rate matrices built here from the
      textbook definition (not from the oracle), and Gamma quantiles /
      Yang category means from scipy.stats / scipy.special.
"""
import itertools
import json
import os

import numpy as np
import scipy.linalg
import scipy.special
import scipy.stats

HERE = os.path.dirname(os.path.abspath(__file__))


def kats():
    dna_states = ["A", "C", "G", "T", "-", "X"]       # lib/alphabet.ml:301-307
    code = {s: 1 << i for i, s in enumerate(dna_states)}
    rows = []
    for x, y in itertools.product(dna_states, repeat=2):
        xs, ys = {x}, {y}
        inter = xs & ys
        med, cost = (inter, 0) if inter else (xs | ys, 1)
        rows.append([code[x], code[y], sum(code[s] for s in med), cost])
    return {
        "dna_codes": {"A": 1, "C": 2, "G": 4, "T": 8, "-": 16},
        "dna_states": code,
        "sequence_kat": {
            "string": "-AGGTCGTCCGTCGGAAT--A",
            "codes": [16, 1, 4, 4, 8, 2, 4, 8, 2, 2, 4, 8, 2, 4, 4, 1, 1, 8, 16, 16, 1],
        },
        "iupac": {"M": ["A", "C"], "R": ["A", "G"], "W": ["A", "T"], "S": ["G", "C"],
                  "Y": ["T", "C"], "K": ["G", "T"], "V": ["G", "T", "C"],
                  "H": ["G", "T", "A"], "D": ["C", "T", "A"], "B": ["G", "C", "A"],
                  "N": ["A", "C", "G", "T"]},
        "fitch_pairs": rows,
        "tree_counts": {
            "unrooted": [[0, 0], [1, 1], [2, 1], [3, 1], [4, 3], [5, 15], [6, 105], [7, 945],
                         [8, 10395], [9, 135135], [10, 2027025], [11, 34459425],
                         [12, 654729075]],
            "rooted": [[0, 0], [1, 1], [2, 1], [3, 3], [4, 15], [5, 105], [6, 945], [7, 10395],
                       [8, 135135], [9, 2027025], [10, 34459425]],
            "nodes": [[0, 0], [1, 1], [2, 2], [3, 4], [4, 6], [5, 8], [6, 10], [7, 12], [8, 14],
                      [9, 16], [10, 18]],
            "edges": [[0, 0], [1, 0], [2, 1], [3, 3], [4, 5], [5, 7], [6, 9], [7, 11], [8, 13],
                      [9, 15], [10, 17]],
        },
    }


def gtr_textbook(pi, ex):
    """Q_ij = s_ij pi_j (i != j), rows sum to zero, mean rate 1."""
    n = len(pi)
    S = np.zeros((n, n))
    k = 0
    for i in range(n):
        for j in range(i + 1, n):
            S[i, j] = S[j, i] = ex[k]
            k += 1
    Q = S * pi[None, :]
    np.fill_diagonal(Q, 0.0)
    np.fill_diagonal(Q, -Q.sum(1))
    return Q / (-(np.diag(Q) * pi).sum())


def models():
    rng = np.random.default_rng(20240601)
    out = []
    for n in (4, 4, 20):
        pi = rng.dirichlet(np.full(n, 5.0 if n == 4 else 2.0))
        ex = rng.dirichlet(np.ones(n * (n - 1) // 2))
        ex = ex / ex[-1]
        co = ex[:-1]
        Q = gtr_textbook(pi, ex)
        mats = {}
        for t in (1e-6, 0.01, 0.1, 0.7, 3.0, 25.0):
            mats[repr(t)] = scipy.linalg.expm(Q * t).tolist()
        out.append({"pi": pi.tolist(), "co": co.tolist(), "Q": Q.tolist(), "P": mats})
    gam = []
    for a in (0.1, 0.5, 1.0, 2.7, 20.0):
        for n in (1, 2, 4, 8):
            ref = [float(scipy.stats.gamma.ppf(i / n, a, scale=a)) for i in range(n)]
            x = [scipy.stats.gamma.ppf(i / n, a, scale=1.0 / a) for i in range(1, n)]
            edges = [0.0] + list(x) + [np.inf]
            mean = [float(n * (scipy.special.gammainc(a + 1, a * edges[i + 1])
                               - scipy.special.gammainc(a + 1, a * edges[i])))
                    for i in range(n)]
            gam.append({"alpha": a, "n": n, "ref": ref, "mean": mean})
    return {"models": out, "gamma": gam}


if __name__ == "__main__":
    with open(os.path.join(HERE, "reference_kats.json"), "w") as f:
        json.dump(kats(), f, indent=1)
    with open(os.path.join(HERE, "model_goldens.json"), "w") as f:
        json.dump(models(), f)
    print("wrote fixtures")
