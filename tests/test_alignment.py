"""Character data in: alphabet code tables and sequence parsing pinned by the
reference's own fixtures (test/alphabetTest.ml:12-18, test/sequenceTest.ml:
6-9), FASTA / PHYLIP readers, and pattern compression, whose invariance of
the Fitch cost and of lnL is checked against the oracle."""
import numpy as np
import pytest

from conftest import random_tree
from oracle import oracle as O


def test_reference_code_tables_and_sequence_fixture():
    from phylocaml_amd import alignment as AL

    # test/alphabetTest.ml:12-18
    for name, code in (("A", 1), ("C", 2), ("G", 4), ("T", 8), ("-", 16)):
        assert AL.dna.get_code(name) == code and AL.dna.get_name(code) == name
    # test/sequenceTest.ml:6-9: the same sequence as codes and as a string
    s1_1 = [16, 1, 4, 4, 8, 2, 4, 8, 2, 2, 4, 8, 2, 4, 4, 1, 1, 8, 16, 16, 1]
    s1_2 = "-AGGTCGTCCGTCGGAAT--A"
    assert AL.of_string(s1_2, AL.dna).tolist() == s1_1
    assert AL.to_string(s1_1, AL.dna) == s1_2
    assert AL.of_string(s1_2.lower(), AL.dna).tolist() == s1_1  # ~case:false
    # IUPAC and indel polymorphisms of lib/alphabet.ml:313-323
    nuc = AL.nucleotides
    for sym, members in (("M", "AC"), ("Y", "TC"), ("N", "ACGT"), ("1", "T-"), ("E", "GTA-"),
                         ("P", "GTAC-"), ("?", "GTAC-")):
        assert nuc.get_code(sym) == sum(AL.dna.get_code(m) for m in members)
    # the aminoacid alphabet is sequential in the reference's order
    assert [AL.aminoacids.get_code(c) for c in "ARNDV"] == [0, 1, 2, 3, 19]
    assert AL.aminoacids.get_code("X") == 20 and AL.aminoacids.get_code("-") == 21
    with pytest.raises(AL.InvalidSequence):
        AL.of_string("ACGZ", AL.dna)


def test_fasta_and_phylip_readers():
    from phylocaml_amd import alignment as AL

    fasta = ">t1 first taxon\nACGT-\nNNA\n>t2\nacgtr\nYK?\n"
    names, m = AL.read_fasta(fasta)
    assert names == ["t1", "t2"] and m.shape == (2, 8)
    assert m[0].tolist() == AL.of_string("ACGT-NNA", AL.nucleotides).tolist()
    assert m[1].tolist() == AL.of_string("ACGTRYK?", AL.nucleotides).tolist()
    seq = "2 8\nt1  ACGT -NNA\nt2  acgtr\nYK?\n"
    n2, m2 = AL.read_phylip(seq)
    assert n2 == names and np.array_equal(m2, m)
    inter = "2 8\nt1 ACGT\nt2 acgt\n-NNA\nrYK?\n"
    n3, m3 = AL.read_phylip(inter, interleaved=True)
    assert n3 == names and np.array_equal(m3, m)
    with pytest.raises(AL.InvalidSequence):
        AL.read_phylip("2 9\nt1 ACGT\nt2 ACGT\n")
    with pytest.raises(AL.InvalidSequence):
        AL.read_fasta("ACGT\n>t1\nACGT\n")


def test_compression_keeps_fitch_cost_and_lnl(rng):
    from phylocaml_amd import alignment as AL, mlmodel as MM

    N, S = 12, 3000
    ch, root = random_tree(rng, N)
    alphabet = np.array(list("ACGTACGTACGTRYN-"))
    pool = rng.choice(alphabet, (N, 200))          # 200 distinct-ish columns
    cols = pool[:, rng.integers(0, 200, S)]        # S sites drawn from them
    text = ["".join(row) for row in cols]
    m = np.stack([AL.of_string(t, AL.nucleotides) for t in text])
    pat, w, idx = AL.compress(m)
    assert np.array_equal(pat[:, idx], m) and w.sum() == S and pat.shape[1] < S
    sets = AL.fitch_sets(m)
    psets = AL.fitch_sets(pat)
    assert O.fitch_downpass(N, ch, root, sets)[3] == O.fitch_downpass(N, ch, root, psets, w)[3]
    model = MM.create(MM.Spec(4, ("GTR", [1.0, 2.0, 0.5, 0.7, 3.0]), ("DiscreteGamma", 4, 0.5),
                              ("Empirical", [0.1, 0.2, 0.3, 0.4])))
    bl = rng.exponential(0.1, 2 * N - 1)
    bl[root] = 0.0
    full = O.lnl(N, ch, root, bl, model.priors, model.q, model.rates, model.probs,
                 sets.astype(np.uint32), np.ones(S, np.uint32))[0]
    comp = O.lnl(N, ch, root, bl, model.priors, model.q, model.rates, model.probs,
                 psets.astype(np.uint32), w)[0]
    assert abs(full - comp) <= 1e-10 * abs(full)
    # Sankoff sets keep the gap as the fifth state
    assert AL.sankoff_sets(AL.of_string("A-R1", AL.nucleotides)).tolist() == [1, 16, 5, 24]
    assert AL.fitch_sets(AL.of_string("A-R1", AL.nucleotides)).tolist() == [1, 15, 5, 8]


def test_polymorphisms_follow_the_reference_properties():
    # the properties of test/alphabetTest.ml:120-169 over Alphabet.dna, whose
    # atomic states are A C G T - X (lib/alphabet.ml:301-307)
    import itertools
    import json
    import os

    from conftest import GOLDEN
    from phylocaml_amd import alignment as AL

    kat = json.load(open(os.path.join(GOLDEN, "reference_kats.json")))
    dna = AL.dna
    assert {s: dna.get_code(s) for s in kat["dna_states"]} == kat["dna_states"]
    assert dna.atomic == sorted(kat["dna_states"].values())
    assert [dna.get_code(e) for e in "01234"] == [1, 2, 4, 8, 16]  # the equates
    for x in dna.atomic:
        assert dna.explode_polymorphisms([x, x]) == [x]
    for x, y in itertools.permutations(dna.atomic, 2):
        xy = dna.compress_polymorphisms([x, y])
        assert xy == x + y and dna.explode_polymorphisms([xy]) == sorted([x, y])
    acc, total = set(), 0
    for e in dna.atomic:
        acc.add(e)
        total += e
        assert dna.get_state_combination_exn(acc) == total
    for x, y, z in itertools.permutations(dna.atomic, 3):
        assert dna.explode_polymorphisms([x + y + z]) == sorted([x, y, z])
    # IUPAC codes explode to their members
    nuc = AL.nucleotides
    assert nuc.explode_polymorphisms([nuc.get_code("V")]) == sorted(
        nuc.get_code(c) for c in "GTC")
    assert nuc.compress_polymorphisms([nuc.get_code("M"), nuc.get_code("K")]) == nuc.get_code("N")
    # a sequential alphabet represents one state per code
    aa = AL.aminoacids
    assert aa.max_polymorphism() == 1 and aa.get_state_combination({5}) == 5
    assert aa.get_state_combination({1, 2}) is None
    assert aa.compress_polymorphisms([7, 3]) == 3  # truncated to the first state
    with pytest.raises(AL.AlphabetError):
        aa.get_combination(25)
    # dna's X reads as missing data in the leaf formats
    assert AL.fitch_sets(np.array([32, 16, 5], np.uint8)).tolist() == [15, 15, 5]
    assert AL.sankoff_sets(np.array([32, 17], np.uint8)).tolist() == [0, 17]
