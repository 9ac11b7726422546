"""wp_hmm_classify_device_noalloc against wp_hmm_classify_device, through the
C ABI: identical outputs bit for bit in both precisions, on two calls in a row
(the second with fewer reads, so the shared scratch holds the first call's
contents), and a call too large for the scratch rejected before anything is
written."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _outputs(torch, n):
    dev = torch.device("cuda:0")
    return (torch.full((n,), -7, dtype=torch.int32, device=dev),
            torch.full((n,), -1.0, dtype=torch.float64, device=dev),
            torch.full((n,), -1.0, dtype=torch.float64, device=dev),
            torch.full((n,), -1.0, dtype=torch.float64, device=dev))


def _call(fn, clf, d_rads, n, outs, stream):
    rc = fn(clf._h, n, d_rads.data_ptr(), *[o.data_ptr() for o in outs], stream)
    assert rc == 0, rc


@pytest.mark.parametrize("precision", [0, 1])
def test_noalloc_matches_device_entry_point(gpu, precision):
    import torch

    w = gpu
    lib = w.lib()
    C, T, S, cut = 3, 12, 1000, 5.0
    m = w.SequencingModel.uniform(C)
    strings = w.generate_dye_seqs(C, S, 6, 30, 0, 5, 1001)
    rads, _ = w.simulate_radiometries(m, T, strings, 3000, 2001)
    entries = [w.DyeSeqEntry(s, 1 + i % 3, 1000 + i) for i, s in enumerate(strings)]
    clf = w.HMMClassifier(T, C, m, w.SequencingSettings(cut), entries, precision)
    stream = torch.cuda.current_stream().cuda_stream
    for n in (3000, 700):  # the second call is smaller
        print("precision", precision, "reads", n, flush=True)
        d_rads = torch.from_numpy(np.ascontiguousarray(rads[:n])).cuda()
        want = _outputs(torch, n)
        got = _outputs(torch, n)
        _call(lib.wp_hmm_classify_device, clf, d_rads, n, want, stream)
        _call(lib.wp_hmm_classify_device_noalloc, clf, d_rads, n, got, stream)
        torch.cuda.synchronize()
        for a, b in zip(got, want):
            assert torch.equal(a, b)
        assert (want[0].cpu().numpy() >= 1000).all()


def test_noalloc_rejects_calls_beyond_its_scratch(gpu):
    import torch

    w = gpu
    lib = w.lib()
    m = w.SequencingModel.uniform(3)
    strings = w.generate_dye_seqs(3, 50, 6, 30, 0, 5, 5)
    clf = w.HMMClassifier(12, 3, m, w.SequencingSettings(5.0), [w.DyeSeqEntry(s, 1, i) for i, s in enumerate(strings)],
                          w.WP_PRECISION_FP32)
    n = 1_000_000  # about 400 MB of workspace at these shapes
    d_rads = torch.zeros(1, dtype=torch.float64, device="cuda")  # never read: rejected on the host
    outs = _outputs(torch, 4)
    rc = lib.wp_hmm_classify_device_noalloc(clf._h, n, d_rads.data_ptr(), *[o.data_ptr() for o in outs],
                                            torch.cuda.current_stream().cuda_stream)
    assert rc == w.WP_ERR_ARG if hasattr(w, "WP_ERR_ARG") else rc == -1
    assert b"scratch" in lib.wp_last_error()
    torch.cuda.synchronize()
    assert (outs[0].cpu().numpy() == -7).all()
