"""The closed form of PeptideHMM's pruned boxes used by the HIP kernels,
checked against the oracle's literal step-by-step prune passes."""
import numpy as np
import pytest

from tests._oracle import DBL_MAX
from whatprot_amd import SequencingModel


def closed_form_boxes(oracle, m, dye_string, rad, M, cutoff):
    T, C = rad.shape
    lo = np.zeros((T, C), dtype=np.int64)
    hi = np.zeros((T, C), dtype=np.int64)
    for t in range(T):
        a, b = oracle.emission_range(m, rad, t, M, cutoff)
        lo[t], hi[t] = a, b
    D = np.array([dye_string.count(str(c)) for c in range(C)])
    L = np.maximum.accumulate(lo[::-1], axis=0)[::-1]
    H = np.minimum.accumulate(np.minimum(hi, M + 1), axis=0)
    H = np.minimum(H, D[None, :] + 1)
    ok = bool((L < H).all())
    return ok, L, H


@pytest.mark.parametrize("cutoff", [5.0, 3.0, 1.5, DBL_MAX])
def test_closed_form_matches_step_algebra(oracle, cutoff):
    import whatprot_amd as w
    m = SequencingModel.uniform(3)
    seqs = w.generate_dye_seqs(3, 60, 6, 30, 0, 5, 1001)
    rads, _ = w.simulate_radiometries(m, 12, seqs, 30, 2001)
    M = max(max(s.count(str(c)) for c in range(3)) for s in seqs)
    nonempty = 0
    for r in range(rads.shape[0]):
        for s in seqs:
            ok1, lo1, hi1 = oracle.emission_boxes(m, s, rads[r], M, cutoff)
            ok2, L, H = closed_form_boxes(oracle, m, s, rads[r], M, cutoff)
            assert ok1 == ok2
            if ok1:
                nonempty += 1
                assert (lo1 == L).all() and (hi1 == H).all()
    if cutoff >= 3.0:  # at 1.5 sigma nearly every pair is empty
        assert nonempty > 0
