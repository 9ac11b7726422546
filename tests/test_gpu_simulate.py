"""GPU simulator against the host simulator: the same per-peptide and
per-attempt random streams, so the integer dye tracks, their deduplication
and the read-to-dye-seq assignment must agree exactly, and the intensities up
to the last float32 bit (the device's log and pow may round differently from
the host's)."""
import numpy as np
import pytest

import whatprot_amd as w

pytestmark = pytest.mark.gpu


def _entries(C, S, seed, lens=(6, 24), dyes=(0, 4), counts=None, ids=None):
    strings = w.generate_dye_seqs(C, S, lens[0], lens[1], dyes[0], dyes[1], seed)
    counts = counts if counts is not None else [1] * S
    ids = ids if ids is not None else list(range(S))
    return [w.DyeSeqEntry(s, int(c), int(i)) for s, c, i in zip(strings, counts, ids)]


def _same_tracks(a, b):
    assert a.counts.shape == b.counts.shape
    np.testing.assert_array_equal(a.counts, b.counts)
    np.testing.assert_array_equal(a.offsets, b.offsets)
    np.testing.assert_array_equal(a.ids, b.ids)
    np.testing.assert_array_equal(a.src_counts, b.src_counts)
    np.testing.assert_array_equal(a.hits, b.hits)


@pytest.mark.parametrize("C,T,S,per,seed", [(1, 8, 30, 50, 1), (2, 10, 60, 20, 7), (3, 12, 200, 10, 11),
                                            (3, 30, 40, 25, 5)])
def test_dye_tracks_identical_to_host(gpu, C, T, S, per, seed):
    rng = np.random.default_rng(seed)
    model = w.SequencingModel.uniform(C)
    # zero counts, repeated counts and ids out of order exercise the grouping
    counts = rng.integers(0, 3, size=S)
    ids = rng.permutation(S) * 3 - 5
    ent = _entries(C, S, 100 + seed, counts=counts, ids=ids)
    host = w.simulate_dye_tracks(model, T, ent, per, seed)
    dev = w.simulate_dye_tracks(model, T, ent, per, seed, gpu=True)
    print(f"C={C} T={T}: {len(host.counts)} distinct tracks from {int(counts.sum()) * per} peptides", flush=True)
    _same_tracks(host, dev)
    # the deduplicated output accounts for every visible peptide
    assert dev.hits.sum() <= counts.sum() * per


def test_dye_tracks_sorted_and_unique(gpu):
    model = w.SequencingModel.uniform(3)
    ent = _entries(3, 100, 3)
    dev = w.simulate_dye_tracks(model, 12, ent, 40, 9, gpu=True)
    flat = dev.counts.reshape(len(dev.counts), -1).astype(np.int64)
    keys = [tuple(r) for r in flat]
    assert keys == sorted(keys) and len(set(keys)) == len(keys)
    for t in range(len(dev.counts)):
        s = dev.ids[dev.offsets[t]:dev.offsets[t + 1]]
        assert (np.diff(s) > 0).all()


@pytest.mark.parametrize("C,T,S,n,seed", [(1, 8, 20, 500, 2), (2, 11, 50, 2000, 4), (3, 12, 300, 5000, 6)])
def test_radiometries_match_host(gpu, C, T, S, n, seed):
    model = w.SequencingModel.uniform(C)
    strings = w.generate_dye_seqs(C, S, 6, 24, 0, 4, 200 + seed)
    weights = np.random.default_rng(seed).integers(1, 5, size=S)
    hr, hi = w.simulate_radiometries(model, T, strings, n, seed, weights=weights)
    dr, di = w.simulate_radiometries(model, T, strings, n, seed, weights=weights, gpu=True)
    np.testing.assert_array_equal(hi, di)
    np.testing.assert_allclose(dr, hr, rtol=2.4e-7, atol=1e-300)
    assert (dr == hr).mean() > 0.999


def test_empty_and_invisible(gpu):
    model = w.SequencingModel.uniform(2)
    ent = [w.DyeSeqEntry("..", 1, 0), w.DyeSeqEntry("0.1", 0, 1)]  # no dyes; no copies
    dev = w.simulate_dye_tracks(model, 6, ent, 10, 1, gpu=True)
    host = w.simulate_dye_tracks(model, 6, ent, 10, 1)
    assert len(dev.counts) == 0
    _same_tracks(host, dev)


@pytest.mark.parametrize("gpu_path", [False, True])
def test_capacity_errors_write_nothing(gpu, gpu_path):
    """Too small max_tracks / max_sources: only the sizes are written and the
    call fails (the host and the GPU entry points alike)."""
    import ctypes
    from whatprot_amd._capi import lib
    model = w.SequencingModel.uniform(2)
    ent = _entries(2, 20, 4)
    full = w.simulate_dye_tracks(model, 8, ent, 30, 2, gpu=gpu_path)
    m = model.to_c()
    arr = (ctypes.c_char_p * len(ent))(*[e.dye_string.encode() for e in ent])
    cnt = np.ones(len(ent), dtype=np.int32)
    ids = np.arange(len(ent), dtype=np.int32)
    nt, ns = ctypes.c_int(-1), ctypes.c_int(-1)
    sentinel = np.full(4, -7, dtype=np.int32)
    counts = np.zeros((len(full.counts), 8, 2), dtype=np.int16)
    off = np.zeros(len(full.counts) + 1, dtype=np.int32)
    fn = lib().wp_simulate_dye_tracks_gpu if gpu_path else lib().wp_simulate_dye_tracks
    rc = fn(ctypes.byref(m), 8, 2, len(ent), arr, cnt.ctypes.data, ids.ctypes.data, 30, 2, len(full.counts), 3,
            ctypes.byref(nt), ctypes.byref(ns), counts.ctypes.data, off.ctypes.data, sentinel.ctypes.data,
            sentinel.ctypes.data, sentinel.ctypes.data)
    assert rc != 0
    assert nt.value == len(full.counts) and ns.value == len(full.ids) > 3
    assert (sentinel == -7).all() and (off == 0).all()
