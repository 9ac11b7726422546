"""GPU tests of the learned shortlist: the BF16 MFMA forward and backward
kernels against the NumPy oracle (which rounds to bf16 where the GPU stores
bf16), training, the shortlist, and what the trained network gains over an
untrained one and over the full HMM."""
import numpy as np
import pytest

import whatprot_amd as w
from oracle import mlp as om

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return float(np.linalg.norm(np.asarray(a, np.float64) - b) / max(np.linalg.norm(b), 1e-300))


@pytest.mark.parametrize("dims,B", [((36, 64, 3, 100), 64), ((20, 200, 2, 300), 150), ((36, 512, 3, 1000), 256),
                                    ((36, 512, 3, 20000), 256)])
def test_forward_matches_oracle(gpu, dims, B):
    din, hid, nh, dout = dims
    print(f"mlp forward {dims} B={B}", flush=True)
    net = w.ShortlistMLP(din, hid, nh, dout, seed=3)
    rng = np.random.default_rng(1)
    layers = net.get_layers()
    x = rng.normal(0, 1, (B, din)).astype(np.float32)
    got = net.logits(x)
    _, want = om.forward([(W.astype(np.float64), b.astype(np.float64)) for W, b in layers], x)
    assert _rel(got, want) < 2e-3


@pytest.mark.parametrize("dims,B", [((36, 64, 3, 100), 64), ((20, 200, 2, 300), 150)])
def test_gradients_match_oracle(gpu, dims, B):
    din, hid, nh, dout = dims
    print(f"mlp gradients {dims} B={B}", flush=True)
    net = w.ShortlistMLP(din, hid, nh, dout, seed=5)
    rng = np.random.default_rng(2)
    layers = [(W.astype(np.float64), b.astype(np.float64) + rng.normal(0, 0.05, b.shape)) for W, b in net.get_layers()]
    net.set_layers(layers)
    layers = [(W.astype(np.float32).astype(np.float64), b.astype(np.float32).astype(np.float64)) for W, b in layers]
    x = rng.normal(0, 1, (B, din)).astype(np.float32)
    y = rng.integers(0, dout, B)
    loss, grads = net.gradients(x, y)
    want_loss, want = om.loss_and_gradients(layers, x, y)
    assert loss == pytest.approx(want_loss, rel=2e-3)
    for (gW, gb), (wW, wb) in zip(grads, want):
        assert _rel(gW, wW) < 1e-2 and _rel(gb, wb) < 1e-2


@pytest.fixture(scope="module")
def small_task():
    C, T, S = 3, 12, 40
    model = w.SequencingModel.uniform(C)
    strings = w.generate_dye_seqs(C, S, 6, 20, 0, 3, 81)
    train_x, train_y = w.simulate_radiometries(model, T, strings, 60000, 82)
    test_x, test_y = w.simulate_radiometries(model, T, strings, 2000, 83)
    return C, T, S, model, strings, train_x, train_y, test_x, test_y


def test_training_and_shortlist_gain(gpu, small_task):
    """Training lowers the loss, and the trained network's top-H shortlist
    holds the HMM's FP64 top-1 far more often than the untrained one's."""
    C, T, S, model, strings, train_x, train_y, test_x, test_y = small_task
    net = w.ShortlistMLP(T * C, 128, 3, S, seed=11)
    H = 8
    clf = w.HMMClassifier(T, C, model, w.SequencingSettings(5.0), [w.DyeSeqEntry(s, 1, i) for i, s in
                                                                   enumerate(strings)], w.WP_PRECISION_FP64)
    best = clf.classify_arrays(test_x)[0]

    def recall(n):
        idx, _ = n.shortlist(test_x, H)
        return float(np.mean([best[i] in idx[i] for i in range(len(best))]))

    before = recall(net)
    losses = net.train(train_x.reshape(len(train_x), -1), train_y, batch=512, epochs=3, lr=2e-3)
    after = recall(net)
    print(f"losses {losses}, recall@{H} {before:.3f} -> {after:.3f}")
    assert losses[-1] < losses[0] < np.log(S) + 0.5
    assert after > 0.95 and after > before + 0.3


def test_shortlist_classify(gpu, small_task):
    C, T, S, model, strings, train_x, train_y, test_x, test_y = small_task
    net = w.ShortlistMLP(T * C, 128, 3, S, seed=11)
    net.train(train_x.reshape(len(train_x), -1), train_y, batch=512, epochs=2, lr=2e-3)
    clf = w.HMMClassifier(T, C, model, w.SequencingSettings(5.0), [w.DyeSeqEntry(s, 1, i) for i, s in
                                                                   enumerate(strings)], w.WP_PRECISION_FP64)
    full_ids, full_adj, _, _ = clf.classify_arrays(test_x)
    # every read falls back: the full HMM (scored in chunks of dye-seqs and
    # merged, so the FP64 totals are summed in another order)
    ids, adj, nf = net.classify(clf, test_x, 8, min_posterior=2.0)
    assert nf == len(test_x)
    np.testing.assert_array_equal(ids, full_ids)
    np.testing.assert_allclose(adj, full_adj, rtol=1e-12, atol=0)
    # no fallback: each id is the HMM best among the shortlist
    ids, adj, nf = net.classify(clf, test_x, 8, min_posterior=0.0)
    idx, _ = net.shortlist(test_x, 8)
    assert nf == 0
    assert all(ids[i] in idx[i] for i in range(len(ids)))
    assert np.mean(ids == full_ids) > 0.95


def test_shortlist_classify_device_matches_host(gpu, small_task):
    """wp_shortlist_classify_device on device arrays gives the host entry
    point's ids and adjusted scores bit for bit: every read sure, every read
    falling back, and a threshold that splits them (sure reads gathered and
    scattered back, fallback results scattered into place)."""
    import torch
    C, T, S, model, strings, train_x, train_y, test_x, test_y = small_task
    net = w.ShortlistMLP(T * C, 128, 3, S, seed=11)
    net.train(train_x.reshape(len(train_x), -1), train_y, batch=512, epochs=2, lr=2e-3)
    clf = w.HMMClassifier(T, C, model, w.SequencingSettings(5.0), [w.DyeSeqEntry(s, 1, i) for i, s in
                                                                   enumerate(strings)], w.WP_PRECISION_FP32)
    N = len(test_x)
    dev = torch.device("cuda:0")
    d_x = torch.from_numpy(np.ascontiguousarray(test_x, dtype=np.float64)).to(dev)
    d_id = torch.full((N,), -7, dtype=torch.int32, device=dev)
    d_adj = torch.full((N,), 7.0, dtype=torch.float64, device=dev)
    stream = torch.cuda.current_stream(dev)
    _, post, mass = net.shortlist(test_x, 8, with_mass=True)
    split = float(np.median(mass))
    for thr in (0.0, 2.0, split):
        ids, adj, nf = net.classify(clf, test_x, 8, min_posterior=thr, criterion="mass")
        nfd = net.classify_device(clf, N, d_x.data_ptr(), 8, d_id.data_ptr(), d_adj.data_ptr(), min_posterior=thr,
                                  criterion="mass", stream=stream.cuda_stream)
        assert nfd == nf
        if thr == split:
            assert 0 < nf < N
        np.testing.assert_array_equal(d_id.cpu().numpy(), ids)
        np.testing.assert_array_equal(d_adj.cpu().numpy(), adj)


@pytest.mark.parametrize("S", [40, 700, 5000, 20000, 40000])
def test_shortlist_topk_matches_logits(gpu, S):
    """The shortlist is the top H of the network's own logits (ties to the
    lower index, output in increasing index order) and its mass is their
    softmax mass; covers every register-resident row size and the generic
    kernel above 32768 outputs."""
    rng = np.random.default_rng(S)
    net = w.ShortlistMLP(12, 64, 2, S, seed=5)
    x = rng.normal(0.0, 3.0, size=(300, 12))
    z = net.logits(x.astype(np.float32)).astype(np.float64)
    keys = z.astype(np.float32).view(np.int32).astype(np.int64)
    keys = np.where(keys < 0, -(keys & 0x7FFFFFFF), keys)  # order-preserving integer image
    for H in sorted({1, 37, min(256, S - 1)}):
        idx, post, mass = net.shortlist(x, H, with_mass=True)
        ref = np.sort(np.argsort(-keys, axis=1, kind="stable")[:, :H], axis=1)
        np.testing.assert_array_equal(idx, ref)
        p = np.exp(z - z.max(axis=1, keepdims=True))
        p /= p.sum(axis=1, keepdims=True)
        np.testing.assert_allclose(post, p.max(axis=1), rtol=1e-4)
        np.testing.assert_allclose(mass, np.take_along_axis(p, ref, axis=1).sum(axis=1), rtol=1e-4, atol=1e-6)


@pytest.mark.parametrize("pattern", ["ties", "ascending", "distinct"])
@pytest.mark.parametrize("S", [12000, 20000, 30000])
def test_topk_filter_and_fallback(gpu, S, pattern):
    """Rows whose logits are the last layer's biases (weights zeroed): heavy
    ties and ascending values put more than 4096 logits above the group-maxima
    threshold (the full register search runs), distinct random values keep the
    filtered path; both give the top H of the logits, ties to the lower index,
    in increasing index order."""
    rng = np.random.default_rng(S + len(pattern))
    net = w.ShortlistMLP(12, 64, 2, S, seed=3)
    layers = net.get_layers()
    W, b = layers[-1]
    W[...] = 0.0
    if pattern == "ties":
        b[...] = rng.integers(0, 4, size=b.shape).astype(b.dtype)
    elif pattern == "ascending":
        b[...] = (np.arange(b.size) * 1e-3).reshape(b.shape).astype(b.dtype)
    else:
        b[...] = rng.normal(0.0, 2.0, size=b.shape).astype(b.dtype)
    net.set_layers(layers)
    x = rng.normal(0.0, 3.0, size=(40, 12))
    z = net.logits(x.astype(np.float32)).astype(np.float64)
    keys = z.astype(np.float32).view(np.int32).astype(np.int64)
    keys = np.where(keys < 0, -(keys & 0x7FFFFFFF), keys)
    for H in (1, 256):
        idx, post, mass = net.shortlist(x, H, with_mass=True)
        ref = np.sort(np.argsort(-keys, axis=1, kind="stable")[:, :H], axis=1)
        np.testing.assert_array_equal(idx, ref)
        p = np.exp(z - z.max(axis=1, keepdims=True))
        p /= p.sum(axis=1, keepdims=True)
        np.testing.assert_allclose(mass, np.take_along_axis(p, ref, axis=1).sum(axis=1), rtol=1e-4, atol=1e-6)
