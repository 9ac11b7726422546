"""FP32 path with the FP64 decision of near-ties, through the C ABI.

The FP32 classifier keeps, per read, every candidate whose FP32 ln P lies
within W = 2e-5 + 1e-5 |ln P_best| of its FP32 best and rescores those in FP64
when there are two or more (hmm-classifier.h:59-65 decides: strict '>', the
earlier candidate wins). The bars:

- per pair, |ln P_FP32 - ln P_FP64| stays below W / 4, so the FP64 best is
  always inside the window with room to spare (the window is twice the error
  bound it assumes, and the bound is twice the largest error measured);
- ids are identical to the FP64 path's (which equals the oracle, tested in
  test_gpu_hmm.py) on every read, with and without candidate lists, and when a
  read's window overflows its 32-entry list.
"""
import numpy as np
import pytest

from tests._oracle import DBL_MAX
from tests.test_gpu_hmm import CASES, _workload, make_clf

pytestmark = pytest.mark.gpu

TIE_ABS, TIE_REL = 2e-5, 1e-5

# 2,000 reads of the 10^5-read x 1,000 dye-seq workload (C=3, T=12, lengths
# 6-30, 0-5 dyes per channel, 5-sigma cutoff): same generator seeds
CFG1_SAMPLE = (3, 12, 1000, 2000, (6, 30), (0, 5), 5.0)


def _cfg1_sample(w):
    C, T, S, N, lens, dyes, cut = CFG1_SAMPLE
    m = w.SequencingModel.uniform(C)
    strings = w.generate_dye_seqs(C, S, lens[0], lens[1], dyes[0], dyes[1], 1001)
    rads, _ = w.simulate_radiometries(m, T, strings, N, 2001)
    return m, strings, rads, T, cut


def test_fp32_error_stays_inside_the_window(gpu):
    m, strings, rads, T, cut = _cfg1_sample(gpu)
    p32 = make_clf(gpu, m, T, strings, cut, 1).pair_log_probability(rads)
    p64 = make_clf(gpu, m, T, strings, cut, 0).pair_log_probability(rads)
    fin = np.isfinite(p64)
    assert np.array_equal(fin, np.isfinite(p32))
    d = np.abs(p32[fin] - p64[fin])
    win = TIE_ABS + TIE_REL * np.abs(p64[fin])
    ratio = d / win
    k = int(np.argmax(ratio))
    print(f"pairs {fin.sum()} max |d lnP| {d.max():.3e} max |d|/W {ratio.max():.4f} "
          f"(at ln P {p64[fin][k]:.2f})", flush=True)
    # the best of each read: errors where the decision is made
    best = np.max(np.where(fin, p64, -np.inf), axis=1)
    near = fin & (p64 >= best[:, None] - 1.0)
    dn = np.abs(p32[near] - p64[near]) / (TIE_ABS + TIE_REL * np.abs(p64[near]))
    print(f"near-best pairs {near.sum()} max |d|/W {dn.max():.4f}", flush=True)
    assert ratio.max() <= 0.25


@pytest.mark.parametrize("use_candidates", [False, True])
def test_fp32_ids_equal_fp64_on_cfg1_sample(gpu, use_candidates):
    m, strings, rads, T, cut = _cfg1_sample(gpu)
    S = len(strings)
    rng = np.random.default_rng(17)
    counts = rng.integers(1, 4, size=S)
    ids = rng.permutation(S) + 1000
    cands = None
    if use_candidates:
        # hybrid-like lists: 100-300 candidates per read in a random order,
        # each with its own best
        cands = [list(rng.choice(S, size=int(rng.integers(100, 300)), replace=False))
                 for _ in range(rads.shape[0])]
    c64 = make_clf(gpu, m, T, strings, cut, 0, counts, ids)
    c32 = make_clf(gpu, m, T, strings, cut, 1, counts, ids)
    g64, a64, ls64, lt64 = c64.classify_arrays(rads, candidates=cands)
    c32.set_refine(False)
    raw_id, _, _, _ = c32.classify_arrays(rads, candidates=cands)
    c32.set_refine(True)
    g32, a32, ls32, lt32 = c32.classify_arrays(rads, candidates=cands)
    n_ref, n_all = c32.refine_stats(rads.shape[0])
    print(f"candidates={use_candidates} FP32 argmax agreement {np.mean(raw_id == g64):.4f}; "
          f"with the FP64 decision {np.mean(g32 == g64):.4f}; reads rescored {n_ref} "
          f"(over every candidate {n_all})", flush=True)
    assert (g32 == g64).all()
    # rescored reads carry the FP64 score of their pick, the others the FP32
    # one (within the FP32 error); the normaliser stays FP32
    fin = np.isfinite(ls64)
    assert np.array_equal(fin, np.isfinite(ls32))
    exact = np.mean(np.abs(ls32[fin] - ls64[fin]) <= 1e-12 * np.maximum(1.0, np.abs(ls64[fin])))
    print(f"ln scores equal to FP64: {exact:.4f}", flush=True)
    assert exact >= n_ref / rads.shape[0] - 1e-9
    tol = (TIE_ABS + TIE_REL * np.abs(ls64[fin])) / 4
    assert (np.abs(ls32[fin] - ls64[fin]) <= tol).all()
    np.testing.assert_allclose(lt32[fin], lt64[fin], rtol=1e-4, atol=1e-6)
    np.testing.assert_allclose(a32, a64, rtol=1e-3, atol=1e-9)


@pytest.mark.parametrize("name", list(CASES))
def test_fp32_ids_equal_oracle(gpu, oracle, name):
    C, T, S, N, lens, dyes, cut = CASES[name]
    m, strings, rads, _ = _workload(gpu, C, T, S, N, lens, dyes, seed=3 + len(name))
    rng = np.random.default_rng(5)
    counts = rng.integers(1, 4, size=S)
    ids = rng.permutation(S) + 100
    clf = make_clf(gpu, m, T, strings, cut, 1, counts, ids)
    gid, _, _, _ = clf.classify_arrays(rads)
    oid, _, _ = oracle.classify(m, strings, counts, ids, rads, cut)
    print(name, "rescored", clf.refine_stats(N), flush=True)
    assert (gid == oid).all()


def test_window_overflow_rescores_every_candidate(gpu, oracle):
    """140 copies of one dye-seq tie exactly in FP32 and FP64: the window holds
    more than 128 entries (the near-best list's capacity), the read is
    rescored over every candidate, and the first copy wins as in the
    reference."""
    w = gpu
    C, T, cut = 2, 8, 5.0
    m = w.SequencingModel.uniform(C)
    base = w.generate_dye_seqs(C, 30, 6, 16, 0, 3, 77)
    strings = base[:10] + [base[10]] * 140 + base[11:]
    rads, idx = w.simulate_radiometries(m, T, [base[10]], 12, 78)
    S = len(strings)
    counts = np.ones(S, dtype=np.int32)
    ids = np.arange(S) + 500
    clf = make_clf(w, m, T, strings, cut, 1, counts, ids)
    gid, _, _, _ = clf.classify_arrays(rads)
    n_ref, n_all = clf.refine_stats(rads.shape[0])
    oid, _, _ = oracle.classify(m, strings, counts, ids, rads, cut)
    print("rescored", n_ref, "over every candidate", n_all, flush=True)
    assert n_all >= 1
    assert (gid == oid).all()


def test_reads_below_the_fp64_range_follow_fp64(gpu, oracle):
    """Reads whose best P is below the FP64 range: the reference's FP64 pass
    returns P = 0 (or subnormals) and picks accordingly; the FP32 path defers
    such reads to FP64 over every candidate and returns the same id."""
    w = gpu
    m = w.SequencingModel.uniform(2)
    strings = ["10.01", "0", "1", "..0.1.1.0", "11", "1.0.0.1"]
    T = 12
    rads = np.zeros((4, T, 2))
    rads[0] = 1.0
    rads[1, :, 0] = 7.0        # far from every dye-seq at every step
    rads[2, :, 1] = 3.5
    rads[3, ::2, :] = 2.2
    for cut in (DBL_MAX, 9.0):
        clf = make_clf(w, m, T, strings, cut, 1)
        gid, _, ls, _ = clf.classify_arrays(rads)
        oid, _, _ = oracle.classify(m, strings, np.ones(len(strings)), np.arange(len(strings)), rads, cut)
        print(cut, gid, oid, ls, flush=True)
        assert (gid == oid).all()
