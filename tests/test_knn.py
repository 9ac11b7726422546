"""CPU tests of the kNN stage's host parts: the kd-tree construction order
against the reference's KDTree, and the dye-track file format and generator."""
import numpy as np
import pytest

import whatprot_amd as w


def test_kdtree_order_matches_reference(ref):
    rng = np.random.default_rng(0)
    for trial in range(40):
        n = int(rng.integers(1, 3000))
        d = int(rng.integers(1, 12))
        k = int(rng.integers(1, 40))
        if trial % 3 == 0:
            X = rng.normal(size=(n, d))
        else:  # many equal keys: exercises the partition's duplicate handling
            X = rng.integers(0, 4, (n, d)).astype(float) * rng.choice([1.0, 0.37])
        ours, depth = w.kdtree_order(X, k)
        np.testing.assert_array_equal(ours, ref.kdtree_order(X, k))
        assert sorted(ours.tolist()) == list(range(n)) and depth >= 1


def test_kdtree_order_all_equal_points(ref):
    X = np.ones((500, 3))
    ours, depth = w.kdtree_order(X, 7)
    np.testing.assert_array_equal(ours, ref.kdtree_order(X, 7))


def _tracks():
    model = w.SequencingModel.uniform(2)
    seqs = [w.DyeSeqEntry(s, c, i) for i, (s, c) in enumerate([("..0.1..0", 2), (".1.1.0", 1), ("0..0.1", 3)])]
    return model, seqs, w.simulate_dye_tracks(model, 6, seqs, 40, 7)


def test_simulate_dye_tracks_dedup():
    model, seqs, tr = _tracks()
    assert len(tr) > 10
    keys = [tuple(tr.counts[i].ravel()) for i in range(len(tr))]
    assert keys == sorted(keys) and len(set(keys)) == len(keys)
    assert all(tr.counts[i, 0].any() for i in range(len(tr)))  # no invisible tracks
    for i in range(len(tr)):
        sl = slice(tr.offsets[i], tr.offsets[i + 1])
        assert list(tr.ids[sl]) == sorted(tr.ids[sl])
        for sid, cnt in zip(tr.ids[sl], tr.src_counts[sl]):
            assert cnt == seqs[sid].count
    # hits of each dye-seq add up to at most count * per_peptide
    for s in seqs:
        assert tr.hits[tr.ids == s.id].sum() <= s.count * 40


def test_dye_tracks_file_round_trip(tmp_path):
    _, _, tr = _tracks()
    p = tmp_path / "dt.tsv"
    w.write_dye_tracks(str(p), tr)
    back = w.read_dye_tracks(str(p))
    assert (back.num_timesteps, back.num_channels) == (6, 2)
    for a, b in zip(tr._arrays(), back._arrays()):
        np.testing.assert_array_equal(a, b)
    first = p.read_text().split("\n")
    assert first[:3] == ["6", "2", str(len(tr))]
