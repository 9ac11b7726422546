"""GPU parity of the HMM scoring path against the oracle, through the C ABI.

Bars: FP64 path -- ln P within 1e-12 absolute of the oracle (P within ~1e-12
relative), ids identical, adjusted scores within 1e-10 relative. FP32 path --
|d ln P| / max(1, |ln P|) <= 1e-4 per pair, top-1 agreement >= 99.9%."""
import json
import math
import os

import numpy as np
import pytest

from tests._oracle import DBL_MAX, model_from_golden

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = json.load(open(os.path.join(HERE, "golden", "reference_goldens.json")))
FP32_LOGP_TOL = 1e-4


def make_clf(w, model, T, strings, cutoff, precision, counts=None, ids=None):
    entries = [w.DyeSeqEntry(s, int(counts[i]) if counts is not None else 1,
                             int(ids[i]) if ids is not None else i) for i, s in enumerate(strings)]
    return w.HMMClassifier(T, model.num_channels, model, w.SequencingSettings(cutoff), entries, precision)


def check_empty_pattern(got, want, precision):
    """FP64 must reproduce the oracle's zeros exactly. The FP32 path tracks
    exponents separately, so where the FP64 oracle underflows to 0 it may
    return a finite ln P below the FP64 range; it never returns -inf where
    the oracle is finite."""
    gi, wi = np.isinf(got), np.isinf(want)
    if precision == 0:
        assert np.array_equal(gi, wi), "empty-pair pattern differs"
    else:
        assert not (gi & ~wi).any(), "FP32 lost a pair the FP64 oracle keeps"
        extra = wi & ~gi
        assert (got[extra] < -700.0).all(), "FP32 finite where FP64 is a structural zero"


def log_err(a, b):
    """|a - b| / max(1, |b|) with matching infinities counted as exact."""
    a = np.asarray(a)
    b = np.asarray(b)
    both_inf = np.isinf(a) & np.isinf(b) & (np.sign(a) == np.sign(b))
    d = np.where(both_inf, 0.0, np.abs(a - b) / np.maximum(1.0, np.abs(b)))
    return d


@pytest.mark.parametrize("precision", [0, 1])
def test_peptide_hmm_goldens_on_gpu(gpu, precision):
    w = gpu
    g = GOLD["peptide_hmm"]
    m = model_from_golden(g["model"], g["num_channels"])
    for case in g["cases"]:
        cut = DBL_MAX if case["cutoff"] is None else case["cutoff"]
        clf = make_clf(w, m, g["num_timesteps"], [g["dye_seq"]], cut, precision)
        assert clf.max_num_dyes == g["max_num_dyes"]
        lp = clf.pair_log_probability(np.array([case["rad"]], dtype=np.float64))[0, 0]
        exp = case["expected"]
        print(case["source"], precision, lp, exp, flush=True)
        if exp == 0.0:
            assert lp == -math.inf
        elif precision == 0:
            assert abs(math.exp(lp) - exp) <= 1e-9 * exp
        else:
            assert abs(lp - math.log(exp)) <= FP32_LOGP_TOL * max(1.0, abs(math.log(exp)))


def _workload(w, C, T, S, N, lens, dyes, seed):
    m = w.SequencingModel.uniform(C)
    strings = w.generate_dye_seqs(C, S, lens[0], lens[1], dyes[0], dyes[1], 1000 + seed)
    rads, idx = w.simulate_radiometries(m, T, strings, N, 2000 + seed)
    return m, strings, rads, idx


CASES = {
    # name: C, T, S, N, lens, dyes, cutoff
    "c2t11_uncut": (2, 11, 200, 60, (8, 20), (0, 3), DBL_MAX),
    "c3t12_cut5": (3, 12, 400, 60, (6, 30), (0, 5), 5.0),
    "c1t8_cut3": (1, 8, 50, 80, (3, 12), (0, 4), 3.0),
    "c3t20_cut5": (3, 20, 100, 30, (10, 30), (0, 3), 5.0),
    # boxes of 9..32 states at T <= 16 (split-row kernel in FP32)
    "c3t14_cut6": (3, 14, 200, 120, (6, 30), (0, 5), 6.0),
    "c2t16_uncut": (2, 16, 150, 40, (8, 24), (0, 4), DBL_MAX),
}


@pytest.mark.parametrize("name", list(CASES))
@pytest.mark.parametrize("precision", [0, 1])
def test_pair_log_probability_vs_oracle(gpu, oracle, name, precision):
    C, T, S, N, lens, dyes, cut = CASES[name]
    m, strings, rads, _ = _workload(gpu, C, T, S, N, lens, dyes, seed=len(name))
    print(name, precision, flush=True)
    clf = make_clf(gpu, m, T, strings, cut, precision)
    got = clf.pair_log_probability(rads)
    p = oracle.pair_probabilities(m, strings, rads, cut)
    with np.errstate(divide="ignore"):
        want = np.log(p)
    check_empty_pattern(got, want, precision)
    err = log_err(got, want)
    err[np.isinf(want)] = 0.0
    if precision == 0:
        # Near the bottom of the FP64 range (P < 1e-250) the forward pass of
        # the reference itself goes through subnormal intermediates whose
        # digits depend on the order of the sums; there P is held to 1e-6.
        sub = want < np.log(1e-250)
        assert err[~sub].max(initial=0.0) <= 1e-12, err[~sub].max(initial=0.0)
        assert (np.abs(np.exp(got[sub]) - np.exp(want[sub])) <= 1e-300 + 1e-6 * np.exp(want[sub])).all()
    else:
        assert err.max() <= FP32_LOGP_TOL, err.max()


def tie_aware_agreement(gidx, oidx, p64, rel=1e-5):
    """Share of reads whose pick is, in the FP64 oracle's own numbers, within
    `rel` of the oracle's best (exact FP64 ties are common on small synthetic
    dye-seq sets, and FP32 may break them the other way)."""
    rows = np.arange(len(gidx))
    pg = p64[rows, gidx]
    po = p64[rows, oidx]
    return float(np.mean(pg >= po * (1.0 - rel)))


@pytest.mark.parametrize("name", list(CASES))
@pytest.mark.parametrize("precision", [0, 1])
def test_classify_vs_oracle(gpu, oracle, name, precision):
    C, T, S, N, lens, dyes, cut = CASES[name]
    m, strings, rads, _ = _workload(gpu, C, T, S, N, lens, dyes, seed=3 + len(name))
    rng = np.random.default_rng(5)
    counts = rng.integers(1, 4, size=S)
    ids = rng.permutation(S) + 100
    clf = make_clf(gpu, m, T, strings, cut, precision, counts, ids)
    gid, adj, ls, lt = clf.classify_arrays(rads)
    oid, osc, otot = oracle.classify(m, strings, counts, ids, rads, cut)
    with np.errstate(divide="ignore", invalid="ignore"):
        oadj = np.where(np.isnan(osc / otot), 0.0, osc / otot)
    if precision == 0:
        assert (gid == oid).all()
        np.testing.assert_allclose(adj, oadj, rtol=1e-10, atol=1e-300)
    else:
        inv = np.empty(S + 100, dtype=np.int64)
        inv[ids] = np.arange(S)
        p64 = oracle.pair_probabilities(m, strings, rads, cut)
        raw = float((gid == oid).mean())
        tie = tie_aware_agreement(inv[gid], inv[oid], p64)
        print(f"{name} FP32 top-1 raw {raw:.4f} tie-aware {tie:.4f}", flush=True)
        # near-ties are decided in FP64: the ids are the oracle's
        assert raw == 1.0, (raw, tie)
        np.testing.assert_allclose(adj, oadj, rtol=1e-3, atol=1e-6)


@pytest.mark.parametrize("precision", [0, 1])
def test_candidates_vs_oracle(gpu, oracle, precision):
    C, T, S, N = 3, 12, 300, 40
    m, strings, rads, _ = _workload(gpu, C, T, S, N, (6, 30), (0, 5), seed=21)
    rng = np.random.default_rng(9)
    cands = [list(rng.choice(S, size=int(rng.integers(0, 40)), replace=False)) for _ in range(N)]
    cands[0] = []
    cands[1] = [5, 5, 3]  # repeated candidates are allowed
    for c in cands[2:]:
        if not c:
            c.append(7)
    counts = np.ones(S, dtype=np.int32)
    ids = np.arange(S)
    clf = make_clf(gpu, m, T, strings, 5.0, precision, counts, ids)
    gid, adj, _, _ = clf.classify_arrays(rads, candidates=cands)
    oid, osc, otot = oracle.classify(m, strings, counts, ids, rads, 5.0, candidates=cands)
    with np.errstate(divide="ignore", invalid="ignore"):
        oadj = np.where(np.isnan(osc / otot), 0.0, osc / otot)
    bad = np.nonzero(~np.isclose(adj, np.where(np.isfinite(oadj), oadj, 0.0), rtol=1e-6, atol=1e-300))[0]
    for r in bad[:5]:
        print("mismatch read", r, cands[r][:8], gid[r], oid[r], adj[r], oadj[r], osc[r], otot[r], flush=True)
    # no candidates: the reference indexes dye_seqs[-1]; here id -1, score 0
    assert gid[0] == -1 and oid[0] == -1 and adj[0] == 0.0
    if precision == 0:
        assert (gid == oid).all()
        np.testing.assert_allclose(adj[1:], oadj[1:], rtol=1e-10, atol=1e-300)
    else:
        assert (gid == oid).all()
        np.testing.assert_allclose(adj[1:], oadj[1:], rtol=1e-3, atol=1e-6)


@pytest.mark.parametrize("precision", [0, 1])
def test_edge_cases(gpu, oracle, precision):
    w = gpu
    m = w.SequencingModel.uniform(2)
    strings = ["10.01", "0", "1", "..0.1.1.0", "11"]
    # T = 1; reads that prune everything; an all-zero read
    for T in (1, 2, 5):
        rads = np.zeros((4, T, 2))
        rads[0] = 1.0
        rads[1] = -50.0   # below every count: every pair empty under a cutoff
        rads[2, :, 0] = 7.0
        for cut in (DBL_MAX, 5.0):
            clf = make_clf(w, m, T, strings, cut, precision)
            got = clf.pair_log_probability(rads)
            p = oracle.pair_probabilities(m, strings, rads, cut)
            with np.errstate(divide="ignore"):
                want = np.log(p)
            check_empty_pattern(got, want, precision)
            tol = 1e-12 if precision == 0 else FP32_LOGP_TOL
            err = log_err(got, want)
            err[np.isinf(want)] = 0.0
            assert err.max() <= tol, (T, cut)
            gid, adj, ls, lt = clf.classify_arrays(rads)
            oid, osc, otot = oracle.classify(m, strings, np.ones(5), np.arange(5), rads, cut)
            assert (gid == oid).all()
            if cut == 5.0:
                assert gid[1] == 0 and adj[1] == 0.0  # all-empty: first candidate, NaN guard


@pytest.mark.parametrize("precision", [0, 1])
def test_zero_reads_and_rejected_arguments(gpu, oracle, precision):
    """No reads is a no-op on the host and device entry points; a negative read
    count, decreasing candidate offsets and candidate indices outside the
    dye-seq table return WP_ERR_ARG (-1) before anything is written, and the
    classifier still matches the oracle afterwards."""
    import torch

    w = gpu
    m = w.SequencingModel.uniform(2)
    strings = ["10.01", "0", "1"]
    T = 4
    clf = make_clf(w, m, T, strings, 5.0, precision)
    empty = np.zeros((0, T, 2))
    for out in (clf.classify_arrays(empty), clf.classify_arrays(empty, candidates=[])):
        assert all(len(x) == 0 for x in out)
    assert clf.pair_log_probability(empty).shape == (0, len(strings))
    d_rads = torch.zeros(T * 2, dtype=torch.float64, device="cuda")
    d_id = torch.full((1,), 7, dtype=torch.int32, device="cuda")
    clf.classify_device(0, d_rads.data_ptr(), d_id.data_ptr())
    torch.cuda.synchronize()
    assert int(d_id[0]) == 7

    lib = w.lib()
    rads = np.random.default_rng(11).uniform(0.0, 3.0, size=(2, T, 2))
    ids = np.full(2, 99, dtype=np.int32)
    adj, ls, lt = np.zeros(2), np.zeros(2), np.zeros(2)
    outs = (ids.ctypes.data, adj.ctypes.data, ls.ctypes.data, lt.ctypes.data)
    assert lib.wp_hmm_classify(clf._h, -1, rads.ctypes.data, *outs) == -1
    idx = np.array([0, 1], dtype=np.int32)
    decreasing = np.array([0, 2, 1], dtype=np.int32)
    assert lib.wp_hmm_classify_candidates(clf._h, 2, rads.ctypes.data, decreasing.ctypes.data,
                                          idx.ctypes.data, *outs) == -1
    off = np.array([0, 1, 2], dtype=np.int32)
    for bad in (-1, len(strings)):
        idx = np.array([0, bad], dtype=np.int32)
        assert lib.wp_hmm_classify_candidates(clf._h, 2, rads.ctypes.data, off.ctypes.data,
                                              idx.ctypes.data, *outs) == -1
    assert (ids == 99).all() and not adj.any()

    got = clf.pair_log_probability(rads)
    with np.errstate(divide="ignore"):
        want = np.log(oracle.pair_probabilities(m, strings, rads, 5.0))
    check_empty_pattern(got, want, precision)
    err = log_err(got, want)
    err[np.isinf(want)] = 0.0
    assert err.max() <= (1e-12 if precision == 0 else FP32_LOGP_TOL)


@pytest.mark.parametrize("precision", [0, 1])
def test_host_calls_reuse_staging(gpu, precision):
    """The host-array entry points keep their device staging between calls
    (grown when a call is larger): calls that grow, shrink and switch between
    all-candidate and candidate-list scoring on one classifier give, bit for
    bit, what a fresh classifier gives for each call."""
    C, T, S = 3, 12, 200
    m, strings, rads, _ = _workload(gpu, C, T, S, 120, (6, 30), (0, 5), seed=33)
    rng = np.random.default_rng(4)
    cands = [list(rng.choice(S, size=int(rng.integers(1, 30)), replace=False)) for _ in range(120)]
    clf = make_clf(gpu, m, T, strings, 5.0, precision)
    for lo, hi, with_cands in [(0, 30, False), (0, 120, True), (10, 25, True), (0, 90, False), (50, 51, True)]:
        kw = {"candidates": cands[lo:hi]} if with_cands else {}
        got = clf.classify_arrays(rads[lo:hi], **kw)
        want = make_clf(gpu, m, T, strings, 5.0, precision).classify_arrays(rads[lo:hi], **kw)
        for g, w_ in zip(got, want):
            np.testing.assert_array_equal(g, w_)
