"""GPU tests at the settings of the hybrid and learned-shortlist workloads,
with dye interactions, through the C ABI.

- Hybrid kNN -> HMM with FP32 rescoring at k = 100, sigma = 0.5, H = 1000,
  5-sigma pruning, C = 3, T = 12 against the reference's own HybridClassifier
  (hybrid-classifier.cc:55-83, compiled from its sources): ids identical,
  adjusted scores within the FP32 bound.
- Learned shortlist with the FP32 classifier at H = 256 and the posterior-mass
  criterion over 5,000 dye-seqs: every read's id is the FP64 pick among its
  shortlist, or over every dye-seq when it falls back.
- Self-quench and FRET interactions (channel-model.h:35-55, the settings of
  channel-model.test.cc:120-243): pair ln P against the oracle (FP64 1e-12,
  FP32 1e-4 relative), the compiled reference on sampled pairs, and ids.
"""
import numpy as np
import pytest

import whatprot_amd as w
from tests._oracle import DBL_MAX, max_dyes_of
from tests.test_gpu_hmm import check_empty_pattern, log_err, make_clf

pytestmark = pytest.mark.gpu


def test_hybrid_fp32_at_workload_settings_matches_reference(gpu, ref):
    C, T, S = 3, 12, 2000
    model = w.SequencingModel.uniform(C)
    strings = w.generate_dye_seqs(C, S, 6, 30, 0, 5, 1102)
    seqs = [w.DyeSeqEntry(s, 1 + i % 2, 7000 + i) for i, s in enumerate(strings)]
    tracks = w.simulate_dye_tracks(model, T, seqs, 20, 1103)
    rads, _ = w.simulate_radiometries(model, T, strings, 48, 1104)
    k, sig, H, cut = 100, 0.5, 1000, 5.0
    print(f"hybrid FP32: {S} dye-seqs, {len(tracks)} tracks, k={k} H={H}", flush=True)
    hc = w.HybridClassifier(T, C, model, w.SequencingSettings(cut), k, sig, tracks, H, seqs, w.WP_PRECISION_FP32)
    ids, adj = hc.classify_arrays(rads)
    rid, rsc, rtot = ref.hybrid_classify(model, tracks, k, sig, H, strings, [s.count for s in seqs],
                                         [s.id for s in seqs], rads, cut)
    with np.errstate(divide="ignore", invalid="ignore"):
        radj = np.where(np.isnan(rsc / rtot), 0.0, rsc / rtot)
    print("ids equal", np.mean(ids == rid), flush=True)
    np.testing.assert_array_equal(ids, rid)
    np.testing.assert_allclose(adj, radj, rtol=1e-3, atol=1e-12)


def test_shortlist_fp32_h256_mass_criterion(gpu):
    C, T, S, H = 3, 12, 5000, 256
    model = w.SequencingModel.uniform(C)
    strings = w.generate_dye_seqs(C, S, 6, 30, 0, 5, 1202)
    entries = [w.DyeSeqEntry(s, 1, i) for i, s in enumerate(strings)]
    tx, ty = w.simulate_radiometries(model, T, strings, 1_000_000, 1203, gpu=True)
    clf32 = w.HMMClassifier(T, C, model, w.SequencingSettings(5.0), entries, w.WP_PRECISION_FP32)
    clf64 = w.HMMClassifier(T, C, model, w.SequencingSettings(5.0), entries, w.WP_PRECISION_FP64)
    # HMM labels: the pick among each read's same-dye-count dye-seqs; checked
    # against the FP64 classifier on a sample
    th = w.hmm_labels(clf32, tx, ty)
    D = w.dye_counts(strings, C)
    for r in range(0, 1_000_000, 50_000):
        same = np.nonzero((D == D[ty[r]]).all(axis=1))[0]
        assert th[r] == clf64.classify_arrays(tx[r:r + 1], candidates=[list(same)])[0][0]
    print(f"labels moved by the HMM: {np.mean(th != ty):.4f}", flush=True)
    net = w.ShortlistMLP(T * C, 512, 3, S, seed=1)
    losses = net.train(tx.reshape(len(tx), -1).astype(np.float32), th, batch=4096, epochs=2, lr=1e-3, seed=9)
    del tx, ty
    qx, _ = w.simulate_radiometries(model, T, strings, 2000, 1204)
    ids, adj, nf = net.classify(clf32, qx, H, 0.5, "mass")
    idx, _, mass = net.shortlist(qx, H, with_mass=True)
    fall = mass < 0.5
    full_ids, full_adj, _, _ = clf64.classify_arrays(qx)
    sl_ids, sl_adj, _, _ = clf64.classify_arrays(qx, candidates=[list(r) for r in idx])
    want_ids = np.where(fall, full_ids, sl_ids)
    want_adj = np.where(fall, full_adj, sl_adj)
    recall = float(np.mean([full_ids[i] in idx[i] for i in range(len(qx))]))
    print(f"losses {losses} recall@{H} {recall:.4f} fallbacks {nf} ids vs full FP64 {np.mean(ids == full_ids):.4f}",
          flush=True)
    assert nf == int(fall.sum())
    np.testing.assert_array_equal(ids, want_ids)
    np.testing.assert_allclose(adj, want_adj, rtol=1e-3, atol=1e-12)
    assert recall > 0.95
    # every read falls back: scored in chunks of 512 dye-seqs, the chunks'
    # picks decided by a second FP32 call over them (near-ties in FP64)
    few = qx[:300]
    ids, adj, nf = net.classify(clf32, few, H, 2.0, "mass")
    assert nf == len(few)
    np.testing.assert_array_equal(ids, full_ids[:300])
    np.testing.assert_allclose(adj, full_adj[:300], rtol=1e-3, atol=1e-12)


def _interacting(C, kind):
    """interactions hold 1 - effect: 0.95 / 0.96 on a channel's own dyes
    (self-quench), 0.97 / 0.98 from the other channels' dyes (FRET), as in
    channel-model.test.cc:120-243."""
    m = w.SequencingModel.uniform(C)
    for c, cm in enumerate(m.channel_models):
        for o in range(C):
            if o == c and kind in ("quench", "both"):
                cm.interactions[o], cm.flat_interactions[o] = 0.95, 0.96
            if o != c and kind in ("fret", "both"):
                cm.interactions[o], cm.flat_interactions[o] = 0.97, 0.98
    return m


@pytest.mark.parametrize("kind", ["quench", "fret", "both"])
@pytest.mark.parametrize("C,T,S,N,cut", [(2, 10, 150, 40, 5.0), (3, 12, 300, 40, 5.0), (2, 8, 60, 30, DBL_MAX)])
def test_interactions_vs_oracle(gpu, oracle, ref, kind, C, T, S, N, cut):
    m = _interacting(C, kind)
    strings = w.generate_dye_seqs(C, S, 6, 24, 0, 4, 1300 + C + T)
    rads, _ = w.simulate_radiometries(m, T, strings, N, 1400 + C + T)
    p = oracle.pair_probabilities(m, strings, rads, cut)
    with np.errstate(divide="ignore"):
        want = np.log(p)
    # the oracle against the compiled reference on sampled pairs
    rng = np.random.default_rng(C * T)
    M = max(max_dyes_of(s, C) for s in strings)
    for r, i in zip(rng.integers(0, N, 12), rng.integers(0, S, 12)):
        pr = ref.pair_probability(m, strings[i], rads[r], M, cut)
        assert p[r, i] == pytest.approx(pr, rel=1e-12, abs=1e-300)
    counts = np.ones(S, dtype=np.int32)
    ids = np.arange(S) + 40
    oid, _, _ = oracle.classify(m, strings, counts, ids, rads, cut)
    for precision in (0, 1):
        print(kind, C, T, "precision", precision, flush=True)
        clf = make_clf(gpu, m, T, strings, cut, precision, counts, ids)
        got = clf.pair_log_probability(rads)
        check_empty_pattern(got, want, precision)
        err = log_err(got, want)
        err[np.isinf(want)] = 0.0
        if precision == 0:
            sub = want < np.log(1e-250)
            assert err[~sub].max(initial=0.0) <= 1e-12
        else:
            assert err.max() <= 1e-4
        gid, _, _, _ = clf.classify_arrays(rads)
        np.testing.assert_array_equal(gid, oid)
