"""The simulator's dye-track distribution, on the CPU.

oracle/sim_exact.py gives the exact distribution of generate_dye_track
(simulation/generate-dye-track.cc:29-110) for small dye-seqs. It is pinned
against the reference's own generator, compiled from its sources and run on
a seeded std::default_random_engine, by a chi-square test over 10^6 tracks;
the same test rejects a planted mistake (the detach rate of cycle t read
from p_detach[t + 1]). The host simulator of this library (xoshiro256** streams, same
process) is then held to the exact distribution the same way. Its device
twin is held to it bit for bit (tests/test_gpu_simulate.py)."""
import numpy as np
import pytest
from scipy.stats import chisquare

import whatprot_amd as w
from oracle.sim_exact import track_distribution

N = 1_000_000

# frequent events, so every branch of the process is visited
MODEL = dict(p_edman_failure=0.1, p_detach=0.06, p_initial_detach=0.05, p_initial_detach_decay=0.3,
             p_initial_block=0.1, p_cyclic_block=0.05, p_bleach=0.12, p_dud=0.15)
CASES = [(2, 5, "0.1.0"), (1, 6, "..00."), (3, 4, "120"), (2, 4, "1..0.1")]


def _pvalue(samples, exact):
    """Chi-square p-value of track samples [n, T, C] against an exact
    distribution, pooling tracks whose expected count is below 5."""
    n = samples.shape[0]
    keys, counts = np.unique(samples.reshape(n, -1), axis=0, return_counts=True)
    seen = {tuple(int(v) for v in k): int(c) for k, c in zip(keys, counts)}
    assert set(seen) <= set(exact), "a track the process cannot produce"
    obs, expct = [], []
    pool_o, pool_e = 0, 0.0
    for tr, p in exact.items():
        if n * p >= 5:
            obs.append(seen.get(tr, 0))
            expct.append(n * p)
        else:
            pool_o += seen.get(tr, 0)
            pool_e += n * p
    if pool_e > 0:
        obs.append(pool_o)
        expct.append(pool_e)
    expct = np.array(expct) * (n / np.sum(expct))
    return chisquare(obs, expct).pvalue


def _model(C):
    return w.SequencingModel.uniform(C, **MODEL)


def _host_tracks(model, T, ds, seed):
    tr = w.simulate_dye_tracks(model, T, [w.DyeSeqEntry(ds, 1, 0)], N, seed)
    hits = np.zeros(len(tr), dtype=np.int64)
    for i in range(len(tr)):
        hits[i] = tr.hits[tr.offsets[i]:tr.offsets[i + 1]].sum()
    return np.repeat(tr.counts, hits, axis=0)


@pytest.mark.parametrize("C,T,ds", CASES)
def test_exact_distribution_is_a_distribution(C, T, ds):
    d = track_distribution(_model(C), T, ds)
    assert abs(sum(d.values()) - 1.0) < 1e-12
    assert all(v >= 0 for v in d.values())


@pytest.mark.parametrize("C,T,ds", CASES)
def test_reference_generator_matches_exact(ref, C, T, ds):
    m = _model(C)
    samples = ref.generate_dye_tracks(m, T, ds, N, 12345 + T)
    p = _pvalue(samples, track_distribution(m, T, ds))
    print(f"{ds}: reference generator vs exact, p = {p:.3g}")
    assert p > 1e-4


@pytest.mark.parametrize("C,T,ds", CASES[:2])
def test_planted_mistake_is_rejected(ref, C, T, ds):
    m = _model(C)
    samples = ref.generate_dye_tracks(m, T, ds, N, 999)
    p = _pvalue(samples, track_distribution(m, T, ds, detach_offset=1))
    print(f"{ds}: reference generator vs detach rate off by one cycle, p = {p:.3g}")
    assert p < 1e-8


@pytest.mark.parametrize("C,T,ds", CASES)
def test_host_simulator_matches_exact(C, T, ds):
    """simulate_dye_tracks keeps the tracks with a nonzero first record, as
    generate_dye_tracks does (generate-dye-tracks.cc:53-64): the exact
    distribution is conditioned the same way."""
    m = _model(C)
    samples = _host_tracks(m, T, ds, 77 + T)
    exact = track_distribution(m, T, ds)
    visible = {tr: p for tr, p in exact.items() if any(tr[:C])}
    kept = sum(visible.values())
    # the kept count is binomial(N, kept)
    assert abs(samples.shape[0] - N * kept) <= 5 * np.sqrt(N * kept * (1 - kept)) + 1
    p = _pvalue(samples, {tr: v / kept for tr, v in visible.items()})
    print(f"{ds}: host simulator vs exact, kept {samples.shape[0]} of {N} (expected {N * kept:.0f}), p = {p:.3g}")
    assert p > 1e-4


@pytest.mark.parametrize("kind", ["plain", "fret_and_quench"])
def test_host_radiometries_are_normal_around_adjusted_mu(oracle, kind):
    """generate_radiometry (generate-radiometry.cc:52-60): each intensity is
    Normal(adjusted_mu(counts), sigma(adjusted_mu)). With no duds and two dyes
    of channel 0 and one of channel 1 the first cycle's counts are fixed, so
    the t = 0 intensities of 10^6 reads are two known normals (KS test; the
    library rounds them to float32, far below the test's resolution)."""
    from scipy.stats import kstest

    m = w.SequencingModel.uniform(2, p_dud=0.0, sig=0.16, bg_sig=0.05, mu=1.3)
    if kind == "fret_and_quench":
        for c, cm in enumerate(m.channel_models):
            cm.interactions[c], cm.flat_interactions[c] = 0.9, 0.95
            cm.interactions[1 - c], cm.flat_interactions[1 - c] = 0.97, 0.98
    rads, _ = w.simulate_radiometries(m, 3, ["0.0.1"], 1_000_000, 31)
    counts = np.array([2, 1], dtype=np.uint32)
    for c in range(2):
        mu = oracle.L.wpo_adjusted_mu(oracle._m(m), c, counts.ctypes.data_as(__import__("ctypes").c_void_p))
        sd = oracle.L.wpo_sigma(oracle._m(m), c, mu)
        p = kstest(rads[:, 0, c], "norm", args=(mu, sd)).pvalue
        print(f"{kind} channel {c}: mu {mu:.6f} sigma {sd:.6f} KS p = {p:.3g}")
        assert p > 1e-4
