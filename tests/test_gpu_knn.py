"""GPU parity of the kNN and hybrid classifiers through the C ABI: the
kd-tree goldens of kd-tree.test.cc, brute force on tiny inputs, and the
reference's own NNClassifier / HybridClassifier compiled from its sources."""
import numpy as np
import pytest

import whatprot_amd as w

pytestmark = pytest.mark.gpu


def _grid_tracks(hits):
    """The 4 x 3 integer grid of kd-tree.test.cc:48-72 as one-channel dye
    tracks over two timesteps (with mu = 1 and no interactions a track's
    coordinates are its counts)."""
    pts = [(x, y) for x in range(3) for y in range(4)]
    counts = np.array([[[x], [y]] for x, y in pts], dtype=np.int16)
    n = len(pts)
    return pts, w.DyeTracks(2, 1, counts, np.arange(n + 1, dtype=np.int32), np.arange(n, dtype=np.int32),
                            np.ones(n, dtype=np.int32), np.full(n, hits, dtype=np.int32))


def test_kdtree_goldens(gpu):
    model = w.SequencingModel.uniform(1, bg_sig=0.0, sig=0.0)
    q = np.array([[[0.9], [0.8]]])
    # kd-tree.test.cc:47-105: k = 4, farthest first
    pts, tr = _grid_tracks(1)
    cnt, idx, dist = w.NNClassifier(2, 1, model, 4, 0.5, tr).search(q)
    assert cnt[0] == 4
    assert [pts[i] for i in idx[0, :4]] == [(2, 1), (0, 1), (1, 0), (1, 1)]
    want = [(2 - .9) ** 2 + (1 - .8) ** 2, (0 - .9) ** 2 + (1 - .8) ** 2, (1 - .9) ** 2 + (0 - .8) ** 2,
            (1 - .9) ** 2 + (1 - .8) ** 2]
    np.testing.assert_allclose(dist[0, :4], want, rtol=1e-9)
    # kd-tree.test.cc:107-170: two hits per point, k = 5 -> three entries
    pts, tr = _grid_tracks(2)
    cnt, idx, dist = w.NNClassifier(2, 1, model, 5, 0.5, tr).search(q)
    assert cnt[0] == 3
    assert [pts[i] for i in idx[0, :3]] == [(0, 1), (1, 0), (1, 1)]


def test_search_is_brute_force_for_single_hits(gpu):
    """With one hit per point and distinct distances the k-best set is the k
    nearest points, farthest first."""
    rng = np.random.default_rng(3)
    T, C, n, k = 5, 2, 3000, 16
    counts = rng.integers(0, 6, (n, T, C)).astype(np.int16)
    counts = np.unique(counts.reshape(n, -1), axis=0).reshape(-1, T, C)
    n = counts.shape[0]
    tr = w.DyeTracks(T, C, counts, np.arange(n + 1, dtype=np.int32), np.arange(n, dtype=np.int32),
                     np.ones(n, np.int32), np.ones(n, np.int32))
    model = w.SequencingModel.uniform(C, bg_sig=0.0, sig=0.0)
    q = rng.uniform(0, 5, (200, T, C))
    cnt, idx, dist = w.NNClassifier(T, C, model, k, 0.5, tr).search(q)
    X = counts.reshape(n, -1).astype(float)
    for r in range(len(q)):
        d2 = ((X - q[r].ravel()) ** 2).sum(1)
        order = np.argsort(d2, kind="stable")[:k][::-1]
        assert cnt[r] == k
        np.testing.assert_array_equal(np.sort(idx[r]), np.sort(order))
        np.testing.assert_allclose(dist[r], d2[idx[r]], rtol=1e-12)
        assert np.all(np.diff(dist[r]) <= 0)


@pytest.fixture(scope="module")
def workload():
    C, T = 2, 8
    model = w.SequencingModel.uniform(C)
    strings = w.generate_dye_seqs(C, 60, 6, 14, 0, 3, 51)
    seqs = [w.DyeSeqEntry(s, 1 + i % 3, 1000 + i) for i, s in enumerate(strings)]
    tracks = w.simulate_dye_tracks(model, T, seqs, 30, 52)
    rads, _ = w.simulate_radiometries(model, T, strings, 300, 53)
    return C, T, model, seqs, tracks, rads


def test_search_matches_reference(gpu, ref, workload):
    C, T, model, seqs, tracks, rads = workload
    for k in (1, 10, 50):
        got = w.NNClassifier(T, C, model, k, 0.5, tracks).search(rads)
        want = ref.nn_search(model, tracks, k, 0.5, rads)
        np.testing.assert_array_equal(got[0], want[0])
        for r in range(len(rads)):
            n = got[0][r]
            np.testing.assert_array_equal(got[1][r, :n], want[1][r, :n])
            np.testing.assert_array_equal(got[2][r, :n], want[2][r, :n])


def test_nn_classify_matches_reference(gpu, ref, workload):
    C, T, model, seqs, tracks, rads = workload
    nn = w.NNClassifier(T, C, model, 20, 0.5, tracks)
    ids, sc, tot = nn.classify_arrays(rads)
    rid, rsc, rtot = ref.nn_classify(model, tracks, 20, 0.5, rads)
    np.testing.assert_array_equal(ids, rid)
    np.testing.assert_array_equal(sc, rsc)
    np.testing.assert_array_equal(tot, rtot)
    top, ttot = nn.classify_top(rads, 7)
    rtop, rttot = ref.nn_classify_top(model, tracks, 20, 0.5, rads, 7)
    assert top == rtop
    np.testing.assert_array_equal(ttot, rttot)


@pytest.mark.parametrize("cutoff", [w.DBL_MAX, 5.0])
def test_hybrid_matches_reference(gpu, ref, workload, cutoff):
    C, T, model, seqs, tracks, rads = workload
    strings = [s.dye_string for s in seqs]
    hc = w.HybridClassifier(T, C, model, w.SequencingSettings(cutoff), 20, 0.5, tracks, 10, seqs,
                            w.WP_PRECISION_FP64)
    ids, adj = hc.classify_arrays(rads)
    rid, rsc, rtot = ref.hybrid_classify(model, tracks, 20, 0.5, 10, strings, [s.count for s in seqs],
                                         [s.id for s in seqs], rads, cutoff)
    np.testing.assert_array_equal(ids, rid)
    np.testing.assert_allclose(adj, rsc / rtot, rtol=1e-12, atol=1e-300)


def test_search_matches_reference_larger_tree(gpu, ref):
    """Tens of thousands of tracks in 36 dimensions, k = 100: deep trees,
    early exits and many visited leaves."""
    C, T = 3, 12
    model = w.SequencingModel.uniform(C)
    strings = w.generate_dye_seqs(C, 1500, 6, 30, 0, 5, 61)
    seqs = [w.DyeSeqEntry(s, 1, i) for i, s in enumerate(strings)]
    tracks = w.simulate_dye_tracks(model, T, seqs, 20, 62)
    rads, _ = w.simulate_radiometries(model, T, strings, 60, 63)
    got = w.NNClassifier(T, C, model, 100, 0.5, tracks).search(rads)
    want = ref.nn_search(model, tracks, 100, 0.5, rads)
    np.testing.assert_array_equal(got[0], want[0])
    np.testing.assert_array_equal(got[1], want[1])
    np.testing.assert_array_equal(got[2], want[2])


def test_zero_reads_then_reads(gpu, ref, workload):
    """No reads is a no-op for search, NN classify and hybrid classify (the
    hybrid path once rejected it: its empty per-read result vectors reached the
    HMM entry point as null outputs); the same classifiers then still match the
    reference on real reads."""
    C, T, model, seqs, tracks, rads = workload
    empty = np.zeros((0, T, C))
    nn = w.NNClassifier(T, C, model, 20, 0.5, tracks)
    cnt, idx, dist = nn.search(empty)
    assert len(cnt) == 0 and idx.size == 0 and dist.size == 0
    assert all(len(x) == 0 for x in nn.classify_arrays(empty))
    hc = w.HybridClassifier(T, C, model, w.SequencingSettings(5.0), 20, 0.5, tracks, 10, seqs,
                            w.WP_PRECISION_FP64)
    ids, adj = hc.classify_arrays(empty)
    assert len(ids) == 0 and len(adj) == 0
    some = rads[:17]
    ids, adj = hc.classify_arrays(some)
    rid, rsc, rtot = ref.hybrid_classify(model, tracks, 20, 0.5, 10, [s.dye_string for s in seqs],
                                         [s.count for s in seqs], [s.id for s in seqs], some, 5.0)
    np.testing.assert_array_equal(ids, rid)
    np.testing.assert_allclose(adj, rsc / rtot, rtol=1e-12, atol=1e-300)
