"""Kernel selection and launch scheduling of the classify path, through the C
ABI: the two quad-rows variants (the kernel of runs longer than 12
timesteps), the autotuner's choice and graph replay must all give the same
outputs bit for bit (they change which kernel code runs and how it is
launched, never the arithmetic of a pair)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _setup(w, N=3000, S=400, seed=7, T=12):
    import torch
    C = 3
    m = w.SequencingModel.uniform(C)
    strings = w.generate_dye_seqs(C, S, 6, 30, 0, 5, 1000 + seed)
    rads, _ = w.simulate_radiometries(m, T, strings, N, 2000 + seed)
    entries = [w.DyeSeqEntry(s, 1, i) for i, s in enumerate(strings)]
    clf = w.HMMClassifier(T, C, m, w.SequencingSettings(5.0), entries, w.WP_PRECISION_FP32)
    dev = torch.device("cuda:0")
    return clf, torch.from_numpy(rads).to(dev), rads


def _run(clf, d_rads, N, stream):
    import torch
    dev = d_rads.device
    out = [torch.full((N,), -7, dtype=torch.int32, device=dev)] + \
          [torch.full((N,), 7.0, dtype=torch.float64, device=dev) for _ in range(3)]
    clf.classify_device(N, d_rads.data_ptr(), *[o.data_ptr() for o in out], stream.cuda_stream)
    torch.cuda.synchronize()
    return [o.cpu().numpy() for o in out]


def _same(a, b):
    for x, y in zip(a, b):
        assert np.array_equal(x, y, equal_nan=True), "outputs differ"


def test_variants_identical_and_autotuned(gpu):
    import torch
    w = gpu
    clf, d_rads, rads = _setup(w, T=14)
    N = rads.shape[0]
    stream = torch.cuda.current_stream()
    res = {}
    for mode in (0, 1, -1):
        clf.set_variant(mode)
        res[mode] = _run(clf, d_rads, N, stream)
        v, ms = clf.variants()
        print("mode", mode, "variants", v.tolist(), "ms", ms.tolist(), flush=True)
        assert (v >= 0).all()
        if mode >= 0:
            assert (v == mode).all()
    _same(res[0], res[1])
    _same(res[0], res[-1])
    # a call of 20,000 reads or more times both variants of each quad-rows bin
    big, d_big, rbig = _setup(w, N=20000, S=200, seed=9, T=14)
    big.set_variant(-1)
    _run(big, d_big, rbig.shape[0], stream)
    v, ms = big.variants()
    print("tuned", v.tolist(), ms.tolist(), flush=True)
    assert (v >= 0).all() and (ms[:3] >= 0).all()
    # the host entry point (eager launches) agrees with the device one
    ids, adj, ls, lt = clf.classify_arrays(rads)
    _same(res[0], [ids, adj, ls, lt])


def test_graph_replay_matches_eager(gpu):
    import torch
    w = gpu
    clf, d_rads, rads = _setup(w, N=2500, seed=11)
    N = rads.shape[0]
    stream = torch.cuda.current_stream()
    clf.set_graph(False)
    eager = _run(clf, d_rads, N, stream)
    clf.set_graph(True)
    # same pointers across replays: allocate the outputs once
    dev = d_rads.device
    out = [torch.empty(N, dtype=torch.int32, device=dev)] + [torch.empty(N, dtype=torch.float64, device=dev)
                                                           for _ in range(3)]
    for rep in range(4):
        for o in out:
            o.fill_(-3)  # eager work between the replays
        clf.classify_device(N, d_rads.data_ptr(), *[o.data_ptr() for o in out], stream.cuda_stream)
        scratch = torch.randn(1 << 16, device=dev).sum()  # unrelated work on the same stream
        torch.cuda.synchronize()
        print("replay", rep, float(scratch), flush=True)
        _same(eager, [o.cpu().numpy() for o in out])
    # a different batch size is captured again
    n2 = N // 2
    part = _run(clf, d_rads[:n2].contiguous(), n2, stream)
    _same([e[:n2] for e in eager], part)


def test_outputs_ordered_on_the_callers_stream(gpu):
    """A call's bins run on branches of their own; work enqueued on the
    caller's stream after the call (here clones of the outputs, on a
    non-default stream, with no device synchronisation in between) must see
    every output, in graph replay and in eager launches."""
    import torch
    w = gpu
    clf, d_rads, rads = _setup(w, N=4000, seed=13)
    N = rads.shape[0]
    want = _run(clf, d_rads, N, torch.cuda.current_stream())
    dev = d_rads.device
    side = torch.cuda.Stream(device=dev)
    side.wait_stream(torch.cuda.current_stream())
    out = [torch.empty(N, dtype=torch.int32, device=dev)] + [torch.empty(N, dtype=torch.float64, device=dev)
                                                           for _ in range(3)]
    for graph in (True, False, True):
        clf.set_graph(graph)
        with torch.cuda.stream(side):
            for o in out:
                o.fill_(-5)
            clf.classify_device(N, d_rads.data_ptr(), *[o.data_ptr() for o in out], side.cuda_stream)
            got = [o.clone() for o in out]
            for o in out:
                o.fill_(-9)  # overwritten on the stream after the clones
        side.synchronize()
        print("graph", graph, flush=True)
        _same(want, [g.cpu().numpy() for g in got])
