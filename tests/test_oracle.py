"""Pins the oracle (CPU test infrastructure) before it is trusted as the
checker: reference goldens, closed forms, an independent brute-force
enumeration, invariants, and the compiled reference itself."""
import json
import math
import os

import numpy as np
import pytest

from tests._oracle import DBL_MAX, model_from_golden
from whatprot_amd import SequencingModel

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = json.load(open(os.path.join(HERE, "golden", "reference_goldens.json")))


def close(a, b, rel):
    if b == 0.0:
        return a == 0.0
    return abs(a - b) <= rel * abs(b)


@pytest.mark.parametrize("case", GOLD["peptide_hmm"]["cases"], ids=lambda c: c["source"])
def test_peptide_hmm_goldens(oracle, case):
    g = GOLD["peptide_hmm"]
    m = model_from_golden(g["model"], g["num_channels"])
    cut = DBL_MAX if case["cutoff"] is None else case["cutoff"]
    p = oracle.pair_probability(m, g["dye_seq"], np.array(case["rad"]), g["max_num_dyes"], cut)
    assert close(p, case["expected"], g["tolerance"]), (p, case["expected"])


@pytest.mark.parametrize("case", GOLD["channel_pdf"]["cases"], ids=lambda c: c["source"])
def test_channel_pdf_goldens(oracle, case):
    b = GOLD["channel_pdf"]["base"]
    C = case["num_channels"]
    m = SequencingModel.uniform(C, bg_sig=b["bg_sig"], mu=b["mu"], sig=b["sig"])
    for k, v in case["interactions"].items():
        m.channel_models[0].interactions[int(k)] = v
    for k, v in case["flat"].items():
        m.channel_models[0].flat_interactions[int(k)] = v
    p = oracle.pdf(m, 0, case["observed"], case["counts"])
    assert close(p, case["expected"], GOLD["channel_pdf"]["tolerance"]), (p, case["expected"])


def test_binomial_table_closed_form(oracle):
    # BinomialTransition::reserve builds P(i,j) = C(i,j) (1-q)^j q^(i-j) by recurrence.
    for q in (0.07, 0.05, 0.3):
        t = oracle.binomial_table(q, 12)
        for i in range(13):
            for j in range(i + 1):
                exact = math.comb(i, j) * (1 - q) ** j * q ** (i - j)
                assert abs(t[i * (i + 1) // 2 + j] - exact) <= 1e-14 * max(exact, 1e-300) + 1e-300


def _tiny_cases():
    rng = np.random.default_rng(7)
    seqs = ["10.01", "0.1", ".0.0", "1100", "0", "..1.0.1", "00.", "2.1.0"]
    for s in seqs:
        C = 3 if "2" in s else 2
        for T in (1, 2, 3, 4):
            counts = [s.count(str(c)) for c in range(C)]
            rad = np.zeros((T, C))
            for t in range(T):
                for c in range(C):
                    rad[t, c] = max(0, counts[c] - rng.integers(0, 2) * (t > 0)) + rng.normal(0, 0.2)
            yield s, C, rad


@pytest.mark.parametrize("s,C,rad", list(_tiny_cases()))
def test_brute_force_agreement(oracle, s, C, rad):
    """Independent formulation: a forward pass over explicit per-dye fates
    (bitmask of live dyes) following the simulator's event order, with no
    count tensor, no exchangeability argument and no pruning."""
    m = SequencingModel.uniform(C, p_edman_failure=0.08, p_detach=0.04, p_initial_detach=0.03,
                                p_initial_detach_decay=0.2, p_initial_block=0.06, p_cyclic_block=0.03,
                                p_bleach=0.06, p_dud=0.09, bg_sig=0.05, sig=0.2)
    M = max(s.count(str(c)) for c in range(C))
    p = oracle.pair_probability(m, s, rad, M, DBL_MAX)
    bf = oracle.brute_force(m, s, rad)
    assert p > 0
    assert abs(p - bf) <= 1e-12 * bf, (p, bf)


def test_mass_conservation(oracle):
    """With every emission density equal to one the operators only move
    probability mass, so P == 1. sigma = 1/sqrt(2 pi) and intensities equal to
    the mean make pdf == 1 only at the true count; instead use a huge sigma so
    the density is flat and then divide it out."""
    C, T = 2, 5
    big = 1e6
    m = SequencingModel.uniform(C, bg_sig=big, sig=0.0)
    flat = 1.0 / (big * math.sqrt(2 * math.pi))
    rad = np.zeros((T, C))
    p = oracle.pair_probability(m, "10.01.1", rad, 3, DBL_MAX)
    assert abs(p / flat ** (T * C) - 1.0) < 1e-9


def test_emission_range_uncut_is_full(oracle):
    m = SequencingModel.uniform(2)
    lo, hi = oracle.emission_range(m, np.array([[1.0, 2.0]]), 0, 4, DBL_MAX)
    assert list(lo) == [0, 0] and list(hi) == [5, 5]


@pytest.mark.parametrize("cutoff", [DBL_MAX, 5.0, 2.0])
def test_oracle_matches_compiled_reference(oracle, ref, cutoff):
    """The C restatement against the reference's own classes on random pairs."""
    import whatprot_amd as w
    m = SequencingModel.uniform(3)
    seqs = w.generate_dye_seqs(3, 40, 4, 14, 0, 3, 11)
    rads, _ = w.simulate_radiometries(m, 6, seqs, 25, 12)
    M = max(max(s.count(str(c)) for c in range(3)) for s in seqs)
    worst = 0.0
    for r in range(rads.shape[0]):
        for s in seqs[:20]:
            a = oracle.pair_probability(m, s, rads[r], M, cutoff)
            b = ref.pair_probability(m, s, rads[r], M, cutoff)
            if b == 0.0:
                assert a == 0.0
            else:
                worst = max(worst, abs(a - b) / abs(b))
    assert worst <= 1e-12, worst


def test_oracle_classify_matches_compiled_reference(oracle, ref):
    import whatprot_amd as w
    m = SequencingModel.uniform(2)
    seqs = w.generate_dye_seqs(2, 30, 8, 20, 0, 3, 1000)
    rads, _ = w.simulate_radiometries(m, 11, seqs, 40, 2000)
    counts = np.arange(1, 31) % 4 + 1
    ids = np.arange(30) * 3 + 5
    a = oracle.classify(m, seqs, counts, ids, rads, DBL_MAX)
    b = ref.classify(m, seqs, counts, ids, rads, DBL_MAX)
    assert (a[0] == b[0]).all()
    np.testing.assert_allclose(a[1], b[1], rtol=1e-12)
    np.testing.assert_allclose(a[2], b[2], rtol=1e-12)
