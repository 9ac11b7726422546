"""The text formats through the library's parallel readers and writers
(tsv_io.cpp), against the reference's own io functions compiled from its
sources: written files identical byte for byte, read values identical bit
for bit, including files whose records are not one per line."""
import numpy as np
import pytest

import whatprot_amd as w


def _model(C):
    m = w.SequencingModel.uniform(C)
    for c, cm in enumerate(m.channel_models):
        cm.mu = 9876.5 + 1234.25 * c   # values are written times mu and read divided by it
    return m


def _rads(N, T, C, seed):
    rng = np.random.default_rng(seed)
    r = rng.normal(1.0, 0.7, size=(N, T, C))
    r.flat[::7] *= 1e-9
    r.flat[::11] *= -3.0
    r.flat[::13] = 0.0
    r.flat[5::17] = rng.uniform(-1e300, 1e300, size=r.flat[5::17].shape)
    return r


@pytest.mark.parametrize("N,T,C", [(0, 3, 2), (1, 1, 1), (2000, 12, 3), (300, 30, 4)])
def test_radiometries_match_reference(ref, tmp_path, N, T, C):
    m = _model(C)
    r = _rads(N, T, C, N + T)
    a, b = tmp_path / "ours.tsv", tmp_path / "ref.tsv"
    w.write_radiometries(a, m, r)
    ref.write_radiometries(b, m, r)
    assert a.read_bytes() == b.read_bytes()
    T2, C2, got = w.read_radiometries(a, m)
    want = ref.read_radiometries(b, m, T, C, max(N, 1))
    assert (T2, C2) == (T, C) and got.shape == want.shape
    assert np.array_equal(got.view(np.uint64), want.view(np.uint64))


def test_radiometries_not_one_per_line(ref, tmp_path):
    """operator>> reads any whitespace: values split over lines, spaces, CRs
    and a '+' sign read the same."""
    m = _model(2)
    r = _rads(40, 5, 2, 3)
    p = tmp_path / "odd.tsv"
    vals = (r * np.array([cm.mu for cm in m.channel_models])).reshape(-1)
    toks = [f"{v:.17g}" for v in vals]
    toks[3] = "+" + toks[3].lstrip("+")
    body = ""
    for i, t in enumerate(toks):
        body += t + ("\n" if i % 7 == 6 else (" \r" if i % 5 == 0 else "\t"))
    p.write_text("5\n2\n40\n" + body + "\n")
    _, _, got = w.read_radiometries(p, m)
    want = ref.read_radiometries(p, m, 5, 2, 40)
    assert np.array_equal(got.view(np.uint64), want.view(np.uint64))


def _tracks(n, T, C, seed):
    rng = np.random.default_rng(seed)
    counts = rng.integers(0, 6, size=(n, T, C)).astype(np.int16)
    ns = rng.integers(0, 4, size=n)
    off = np.concatenate([[0], np.cumsum(ns)]).astype(np.int32)
    k = int(off[-1])
    return w.DyeTracks(T, C, counts, off, rng.integers(-5, 10**6, k).astype(np.int32),
                       rng.integers(1, 9, k).astype(np.int32), rng.integers(1, 10**5, k).astype(np.int32))


@pytest.mark.parametrize("n,T,C", [(0, 4, 2), (1, 1, 1), (5000, 12, 3)])
def test_dye_tracks_match_reference(ref, tmp_path, n, T, C):
    tr = _tracks(n, T, C, n + 1)
    a, b = tmp_path / "ours.tsv", tmp_path / "ref.tsv"
    w.write_dye_tracks(a, tr)
    ref.write_dye_tracks(b, tr)
    assert a.read_bytes() == b.read_bytes()
    got = w.read_dye_tracks(b)
    want = ref.read_dye_tracks(b, T, C, max(n, 1), max(int(tr.offsets[-1]), 1))
    for x, y in zip((got.counts, got.offsets, got.ids, got.src_counts, got.hits), want):
        assert np.array_equal(x, y)


def test_dye_tracks_token_stream(ref, tmp_path):
    tr = _tracks(50, 3, 2, 9)
    a = tmp_path / "one.tsv"
    w.write_dye_tracks(a, tr)
    txt = a.read_text().split()
    b = tmp_path / "stream.tsv"
    b.write_text(" ".join(txt[:3]) + "\n" + "\n".join(" ".join(txt[i:i + 5]) for i in range(3, len(txt), 5)) + "\n")
    got = w.read_dye_tracks(b)
    want = ref.read_dye_tracks(b, 3, 2, 50, int(tr.offsets[-1]))
    for x, y in zip((got.counts, got.offsets, got.ids, got.src_counts, got.hits), want):
        assert np.array_equal(x, y)


def test_results_csv_matches_reference(ref, tmp_path):
    rng = np.random.default_rng(4)
    n = 3000
    ids = rng.integers(-1, 10**6, n)
    sc = rng.uniform(0, 1, n)
    sc[::9] = 0.0
    sc[::13] = 1.0
    sc[5::31] = rng.uniform(0, 1e-300, sc[5::31].shape)
    a, b = tmp_path / "ours.csv", tmp_path / "ref.csv"
    w.write_scored_classifications(a, [w.ScoredClassification(int(i), float(s), 0.0, 0.0) for i, s in zip(ids, sc)])
    ref.write_scored(b, ids, sc)
    assert a.read_bytes() == b.read_bytes()
