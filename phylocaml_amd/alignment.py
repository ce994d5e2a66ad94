"""Character data in: alphabets, sequence parsing, FASTA / PHYLIP readers,
site-pattern compression, and the leaf formats of the GPU passes.

* Alphabets follow lib/alphabet.ml:301-343.  ``dna`` and ``nucleotides``
  are bit-flag alphabets (A = 1, C = 2, G = 4, T = 8, gap = 16; the code
  table of test/alphabetTest.ml:12-18).  ``dna`` has X as a sixth atomic
  state (32, its "all" and missing symbol) and the equates 0..4 for
  A C G T -; ``nucleotides`` has five atomic states plus the IUPAC and indel
  polymorphism codes of lib/alphabet.ml:313-323, "?" (all, missing) being
  the union of the five.  ``aminoacids`` is sequential: the 20 residues in
  the order of lib/alphabet.ml:329-341 are 0..19, X (all) is 20 and the gap
  21.  Symbols are case-insensitive (``~case:false``).
* Polymorphisms (lib/alphabet.ml:418-466): ``max_polymorphism``,
  ``get_combination``, ``get_state_combination`` (``_exn``),
  ``explode_polymorphisms``, ``compress_polymorphisms``, ``is_atomic``.
* ``of_string`` is Sequence.of_string (lib/sequence.ml:209-227) for these
  one-character alphabets: an unknown symbol raises ``InvalidSequence``.
* ``read_fasta`` / ``read_phylip`` (sequential or interleaved, relaxed
  names) give (names, codes [N, S]).
* ``compress`` collapses identical site columns into patterns with integer
  weights; every score of the GPU passes is invariant under it.
* ``fitch_sets`` / ``tip_sets`` / ``sankoff_sets`` / ``aa_states`` map codes
  to the leaf inputs of the Fitch, pruning, Sankoff and 20-state passes.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

GAP = "-"


class InvalidSequence(ValueError):
    """Sequence.Invalid_Sequence (lib/sequence.ml:209-227)."""


class AlphabetError(ValueError):
    """Alphabet.Error (lib/alphabet.ml): an illegal code, or a set of states
    the alphabet cannot represent as one code."""


@dataclass(frozen=True)
class Alphabet:
    name: str
    codes: dict          # upper-case symbol -> code
    kind: str            # "bitflag" or "sequential"
    n_states: int        # atomic states (the gap included for DNA)

    @property
    def atomic(self) -> list:
        """Codes of the atomic states (t.atomic)."""
        if self.kind == "bitflag":
            return [1 << i for i in range(self.n_states)]
        return list(range(self.n_states))

    def max_polymorphism(self) -> int:
        """lib/alphabet.ml:420-424: all states for a bit-flag alphabet, one
        for a sequential one."""
        return self.n_states if self.kind == "bitflag" else 1

    def is_atomic(self, code: int) -> bool:
        return code in self.atomic

    def get_combination(self, code: int) -> set:
        """The atomic states of one code (lib/alphabet.ml:427-434)."""
        if self.kind == "bitflag":
            return {a for a in self.atomic if code & a}
        if code in self.atomic:
            return {code}
        raise AlphabetError(f"illegal code {code} for alphabet {self.name}")

    def get_state_combination_exn(self, states) -> int:
        """The code of a set of atomic states (lib/alphabet.ml:437-449): their
        sum for a bit-flag alphabet; a sequential alphabet represents single
        states only."""
        states = set(int(x) for x in states)
        if self.kind == "bitflag":
            return sum(states)
        if len(states) != 1:
            raise AlphabetError(f"{self.name}: one state per code, got {sorted(states)}")
        (c,) = states
        if c not in self.atomic:
            raise AlphabetError(f"illegal code {c} for alphabet {self.name}")
        return c

    def get_state_combination(self, states):
        """As ``get_state_combination_exn``, None where the states cannot be
        represented (lib/alphabet.ml:452-456)."""
        try:
            return self.get_state_combination_exn(states)
        except AlphabetError:
            if self.kind == "sequential" and len(set(states)) == 1:
                raise
            return None

    def explode_polymorphisms(self, codes) -> list:
        """Sorted atomic states of the union of codes (lib/alphabet.ml:459-461)."""
        out = set()
        for c in codes:
            out |= self.get_combination(int(c))
        return sorted(out)

    def compress_polymorphisms(self, codes) -> int:
        """The code of the union of codes, truncated to the first
        ``max_polymorphism`` atomic states (lib/alphabet.ml:464-466)."""
        return self.get_state_combination_exn(
            self.explode_polymorphisms(codes)[: self.max_polymorphism()])

    def get_code(self, sym: str) -> int:
        try:
            return self.codes[sym.upper()]
        except KeyError:
            raise InvalidSequence(f"illegal character {sym!r} for alphabet {self.name}") from None

    def get_name(self, code: int) -> str:
        for k, v in self.codes.items():
            if v == code:
                return k
        raise KeyError(code)


def _bitflag(extra: dict) -> dict:
    base = {"A": 1, "C": 2, "G": 4, "T": 8, GAP: 16}
    out = dict(base)
    for sym, members in extra.items():
        out[sym] = sum(base[m] for m in members)
    return out


dna = Alphabet("dna", {"A": 1, "C": 2, "G": 4, "T": 8, GAP: 16, "X": 32,
                      "0": 1, "1": 2, "2": 4, "3": 8, "4": 16}, "bitflag", 6)
nucleotides = Alphabet("nucleotides", _bitflag({
    "M": "AC", "R": "AG", "W": "AT", "S": "GC", "Y": "TC", "K": "GT",
    "V": "GTC", "H": "GTA", "D": "CTA", "B": "GCA", "N": "ACGT", "X": "ACGT",
    "1": "T-", "2": "G-", "3": "GT-", "4": "C-", "5": "TC-", "6": "GC-", "7": "GTC-",
    "8": "A-", "9": "TA-", "0": "GA-", "E": "GTA-", "F": "AC-", "I": "TAC-", "J": "GAC-",
    "P": "GTAC-", "?": "GTAC-"}), "bitflag", 5)
AA_ORDER = "ARNDCQEGHILKMFPSTWYV"
aminoacids = Alphabet("aminoacids", {**{c: i for i, c in enumerate(AA_ORDER)}, "X": 20, GAP: 21},
                      "sequential", 20)


def of_string(s: str, alphabet: Alphabet) -> np.ndarray:
    """Codes of the symbols of s (whitespace ignored)."""
    out = [alphabet.get_code(c) for c in s if not c.isspace()]
    return np.array(out, dtype=np.uint8)


def to_string(codes, alphabet: Alphabet) -> str:
    return "".join(alphabet.get_name(int(c)) for c in codes)


def read_fasta(text: str, alphabet: Alphabet = nucleotides):
    names, seqs, cur = [], [], None
    for line in text.splitlines():
        line = line.strip()
        if not line or line.startswith(";"):
            continue
        if line.startswith(">"):
            names.append(line[1:].strip().split()[0] if line[1:].strip() else "")
            seqs.append([])
            cur = seqs[-1]
        elif cur is None:
            raise InvalidSequence("FASTA: sequence data before the first '>' header")
        else:
            cur.append(line)
    return _matrix(names, ["".join(s) for s in seqs], alphabet)


def read_phylip(text: str, alphabet: Alphabet = nucleotides, interleaved: bool = False):
    """Relaxed PHYLIP: a header 'N S', then N records of a name (no spaces)
    and its sequence (spaces inside ignored).  Sequential: a record runs
    over as many lines as its S symbols take.  Interleaved: the first block
    has the names, later lines continue the taxa in turn."""
    lines = [ln for ln in text.splitlines() if ln.strip()]
    if not lines:
        raise InvalidSequence("PHYLIP: empty input")
    head = lines[0].split()
    N, S = int(head[0]), int(head[1])
    names, seqs, i = [], [], 1

    def body(ln):
        return "".join(ln.split())

    for _ in range(N):
        if i >= len(lines):
            raise InvalidSequence("PHYLIP: fewer records than the header says")
        parts = lines[i].split(None, 1)
        names.append(parts[0])
        seqs.append(body(parts[1]) if len(parts) > 1 else "")
        i += 1
        while not interleaved and len(seqs[-1]) < S and i < len(lines):
            seqs[-1] += body(lines[i])
            i += 1
    k = 0
    while interleaved and i < len(lines):
        seqs[k % N] += body(lines[i])
        k += 1
        i += 1
    if i < len(lines):
        raise InvalidSequence("PHYLIP: text after the last record")
    for nm, sq in zip(names, seqs):
        if len(sq) != S:
            raise InvalidSequence(f"PHYLIP: {nm} has {len(sq)} sites, header says {S}")
    return _matrix(names, seqs, alphabet)


def _matrix(names, seqs, alphabet):
    if not seqs:
        return [], np.zeros((0, 0), np.uint8)
    rows = [of_string(s, alphabet) for s in seqs]
    S = len(rows[0])
    for nm, r in zip(names, rows):
        if len(r) != S:
            raise InvalidSequence(f"{nm}: {len(r)} sites, expected {S}")
    return names, np.stack(rows)


def compress(codes: np.ndarray):
    """(patterns [N, P], weights [P] uint32, index [S]): column s of codes is
    patterns[:, index[s]], weights counts the columns of each pattern."""
    codes = np.asarray(codes)
    if codes.shape[1] == 0:
        return codes.copy(), np.zeros(0, np.uint32), np.zeros(0, np.int64)
    pat, index, counts = np.unique(codes, axis=1, return_inverse=True, return_counts=True)
    return np.ascontiguousarray(pat), counts.astype(np.uint32), index.reshape(-1)


def fitch_sets(codes: np.ndarray) -> np.ndarray:
    """4-bit Fitch / pruning sets of bit-flag DNA codes: the gap, and the
    ``dna`` alphabet's X (32), read as missing data (all four bases); a base
    set with a gap keeps its bases."""
    c = np.asarray(codes, np.uint8) & 15
    return np.where(c == 0, 15, c).astype(np.uint8)


tip_sets = fitch_sets


def sankoff_sets(codes: np.ndarray) -> np.ndarray:
    """5-state Sankoff sets of bit-flag DNA codes, the gap as the fifth state
    (``dna``'s X, 32, becomes the empty set: every state)."""
    return (np.asarray(codes, np.uint8) & 31).astype(np.uint8)


def aa_states(codes: np.ndarray) -> np.ndarray:
    """Tips of the 20-state batch: residue index 0..19, >= 20 missing."""
    return np.asarray(codes, np.uint8).copy()


__all__ = ["Alphabet", "AlphabetError", "InvalidSequence", "dna", "nucleotides", "aminoacids", "of_string",
           "to_string", "read_fasta", "read_phylip", "compress", "fitch_sets", "tip_sets",
           "sankoff_sets", "aa_states"]
