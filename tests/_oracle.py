"""ctypes face of the oracle (oracle/liboracle.so) and of the compiled
reference (oracle/_ref/libwpref.so) for the tests. Test infrastructure only."""
import ctypes
import os
import sys

import numpy as np

from whatprot_amd import SequencingModel
from whatprot_amd._capi import ModelC

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_SO = os.path.join(ROOT, "oracle", "liboracle.so")
REF_SO = os.path.join(ROOT, "oracle", "_ref", "libwpref.so")
DBL_MAX = sys.float_info.max

P = ctypes.POINTER
c_double, c_int, c_uint, c_void_p = ctypes.c_double, ctypes.c_int, ctypes.c_uint, ctypes.c_void_p


def _ptr(a):
    return a.ctypes.data_as(c_void_p)


def pack_dye_seqs(strings):
    cap = max([len(s) for s in strings] + [1])
    seqs = np.full((len(strings), cap), -1, dtype=np.int16)
    lens = np.zeros(len(strings), dtype=np.int32)
    for i, s in enumerate(strings):
        t = s.rstrip(".")
        lens[i] = len(t)
        for k, ch in enumerate(t):
            seqs[i, k] = int(ch) if ch.isdigit() else -1
    return np.ascontiguousarray(seqs), lens, cap


class Oracle:
    def __init__(self):
        if not os.path.exists(ORACLE_SO):
            import subprocess
            subprocess.run(["make", "-C", os.path.join(ROOT, "oracle")], check=True,
                           stdout=subprocess.DEVNULL)
        L = ctypes.CDLL(ORACLE_SO)
        L.wpo_adjusted_mu.restype = c_double
        L.wpo_adjusted_mu.argtypes = [P(ModelC), c_int, c_void_p]
        L.wpo_sigma.restype = c_double
        L.wpo_sigma.argtypes = [P(ModelC), c_int, c_double]
        L.wpo_pdf.restype = c_double
        L.wpo_pdf.argtypes = [P(ModelC), c_int, c_double, c_void_p]
        L.wpo_p_detach.restype = c_double
        L.wpo_p_detach.argtypes = [P(ModelC), c_uint]
        L.wpo_binomial_table.restype = None
        L.wpo_binomial_table.argtypes = [c_double, c_uint, c_void_p]
        L.wpo_emission_range.restype = None
        L.wpo_emission_range.argtypes = [P(ModelC), c_void_p, c_uint, c_uint, c_uint, c_double, c_void_p, c_void_p]
        L.wpo_emission_boxes.restype = c_int
        L.wpo_emission_boxes.argtypes = [P(ModelC), c_uint, c_uint, c_void_p, c_int, c_void_p, c_uint,
                                         c_double, c_void_p, c_void_p]
        L.wpo_pair_probability.restype = c_double
        L.wpo_pair_probability.argtypes = [P(ModelC), c_uint, c_uint, c_void_p, c_int, c_void_p, c_uint, c_double]
        L.wpo_classify.restype = None
        L.wpo_classify.argtypes = [P(ModelC), c_uint, c_uint, c_int, c_void_p, c_int, c_void_p, c_void_p,
                                   c_void_p, c_int, c_void_p, c_double, c_void_p, c_void_p, c_void_p,
                                   c_void_p, c_void_p]
        L.wpo_pair_probabilities.restype = None
        L.wpo_pair_probabilities.argtypes = [P(ModelC), c_uint, c_uint, c_int, c_void_p, c_int, c_void_p,
                                             c_int, c_void_p, c_double, c_void_p]
        L.wpo_brute_force_probability.restype = c_double
        L.wpo_brute_force_probability.argtypes = [P(ModelC), c_uint, c_uint, c_void_p, c_int, c_void_p]
        L.wpo_num_threads.restype = c_int
        L.wpo_improve_fit.restype = c_double
        L.wpo_improve_fit.argtypes = [P(ModelC), c_uint, c_uint, c_void_p, c_int, c_void_p, c_uint, c_double,
                                      c_void_p]
        L.wpo_fit_estep.restype = None
        L.wpo_fit_estep.argtypes = [P(ModelC), c_uint, c_uint, c_void_p, c_int, c_int, c_void_p, c_uint,
                                    c_void_p, c_void_p]
        self.L = L

    @staticmethod
    def _m(model):
        return model.to_c() if isinstance(model, SequencingModel) else model

    def pdf(self, model, channel, observed, counts):
        m = self._m(model)
        cnt = np.ascontiguousarray(counts, dtype=np.uint32)
        return self.L.wpo_pdf(ctypes.byref(m), channel, observed, _ptr(cnt))

    def binomial_table(self, q, max_n):
        out = np.zeros((max_n + 1) * (max_n + 2) // 2)
        self.L.wpo_binomial_table(q, max_n, _ptr(out))
        return out

    def emission_range(self, model, rad, t, M, cutoff):
        m = self._m(model)
        rad = np.ascontiguousarray(rad, dtype=np.float64)
        C = rad.shape[-1]
        lo = np.zeros(C, dtype=np.uint32)
        hi = np.zeros(C, dtype=np.uint32)
        self.L.wpo_emission_range(ctypes.byref(m), _ptr(rad), C, t, M, cutoff, _ptr(lo), _ptr(hi))
        return lo, hi

    def emission_boxes(self, model, dye_string, rad, M, cutoff):
        m = self._m(model)
        rad = np.ascontiguousarray(rad, dtype=np.float64)
        T, C = rad.shape
        seqs, lens, _ = pack_dye_seqs([dye_string])
        lo = np.zeros(T * C, dtype=np.uint32)
        hi = np.zeros(T * C, dtype=np.uint32)
        ok = self.L.wpo_emission_boxes(ctypes.byref(m), T, C, _ptr(seqs), int(lens[0]), _ptr(rad), M, cutoff,
                                       _ptr(lo), _ptr(hi))
        return bool(ok), lo.reshape(T, C), hi.reshape(T, C)

    def pair_probability(self, model, dye_string, rad, M, cutoff=DBL_MAX):
        m = self._m(model)
        rad = np.ascontiguousarray(rad, dtype=np.float64)
        T, C = rad.shape
        seqs, lens, _ = pack_dye_seqs([dye_string])
        return self.L.wpo_pair_probability(ctypes.byref(m), T, C, _ptr(seqs), int(lens[0]), _ptr(rad), M, cutoff)

    def brute_force(self, model, dye_string, rad):
        m = self._m(model)
        rad = np.ascontiguousarray(rad, dtype=np.float64)
        T, C = rad.shape
        seqs, lens, _ = pack_dye_seqs([dye_string])
        return self.L.wpo_brute_force_probability(ctypes.byref(m), T, C, _ptr(seqs), int(lens[0]), _ptr(rad))

    def improve_fit(self, model, dye_string, rad, M, cutoff=DBL_MAX):
        """(P, statistics) of one pair; layout WP_FIT_STATS_LEN."""
        m = self._m(model)
        rad = np.ascontiguousarray(rad, dtype=np.float64)
        T, C = rad.shape
        seqs, lens, _ = pack_dye_seqs([dye_string])
        st = np.zeros(fit_stats_len(T, C))
        p = self.L.wpo_improve_fit(ctypes.byref(m), T, C, _ptr(seqs), int(lens[0]), _ptr(rad), M, cutoff, _ptr(st))
        return p, st

    def fit_estep(self, model, dye_string, rads, M):
        """(summed statistics, sum of ln P) over reads, uncut."""
        m = self._m(model)
        rads = np.ascontiguousarray(rads, dtype=np.float64)
        N, T, C = rads.shape
        seqs, lens, _ = pack_dye_seqs([dye_string])
        st = np.zeros(fit_stats_len(T, C))
        slp = np.zeros(1)
        self.L.wpo_fit_estep(ctypes.byref(m), T, C, _ptr(seqs), int(lens[0]), N, _ptr(rads), M, _ptr(st),
                             _ptr(slp))
        return st, float(slp[0])

    def pair_probabilities(self, model, dye_strings, rads, cutoff=DBL_MAX):
        m = self._m(model)
        rads = np.ascontiguousarray(rads, dtype=np.float64)
        N, T, C = rads.shape
        seqs, lens, cap = pack_dye_seqs(dye_strings)
        out = np.zeros((N, len(dye_strings)))
        self.L.wpo_pair_probabilities(ctypes.byref(m), T, C, len(dye_strings), _ptr(seqs), cap, _ptr(lens), N,
                                      _ptr(rads), cutoff, _ptr(out))
        return out

    def classify(self, model, dye_strings, counts, ids, rads, cutoff=DBL_MAX, candidates=None):
        m = self._m(model)
        rads = np.ascontiguousarray(rads, dtype=np.float64)
        N, T, C = rads.shape
        seqs, lens, cap = pack_dye_seqs(dye_strings)
        counts = np.ascontiguousarray(counts, dtype=np.int32)
        ids = np.ascontiguousarray(ids, dtype=np.int32)
        off = idx = None
        if candidates is not None:
            off = np.zeros(N + 1, dtype=np.int32)
            off[1:] = np.cumsum([len(c) for c in candidates])
            idx = np.ascontiguousarray(np.concatenate([np.asarray(c, dtype=np.int32) for c in candidates])
                                       if off[-1] else np.zeros(1, np.int32), dtype=np.int32)
        oid = np.zeros(N, dtype=np.int32)
        osc = np.zeros(N)
        otot = np.zeros(N)
        self.L.wpo_classify(ctypes.byref(m), T, C, len(dye_strings), _ptr(seqs), cap, _ptr(lens), _ptr(counts),
                            _ptr(ids), N, _ptr(rads), cutoff, _ptr(off) if off is not None else None,
                            _ptr(idx) if idx is not None else None, _ptr(oid), _ptr(osc), _ptr(otot))
        return oid, osc, otot


class Reference:
    """The reference's own classes, compiled from its sources (oracle/ref.mk)."""

    def __init__(self, L):
        L.wpref_pdf.restype = c_double
        L.wpref_pdf.argtypes = [P(ModelC), c_int, c_double, c_void_p]
        L.wpref_pair_probability.restype = c_double
        L.wpref_pair_probability.argtypes = [P(ModelC), c_uint, c_uint, ctypes.c_char_p, c_void_p, c_uint, c_double]
        L.wpref_improve_fit.restype = c_double
        L.wpref_improve_fit.argtypes = [P(ModelC), c_uint, c_uint, ctypes.c_char_p, c_void_p, c_uint, c_void_p]
        L.wpref_offset_exponential_fit.restype = None
        L.wpref_offset_exponential_fit.argtypes = [c_int, c_void_p, c_void_p, c_void_p, c_void_p]
        L.wpref_mstep.restype = c_double
        L.wpref_mstep.argtypes = [P(ModelC), c_uint, c_uint, ctypes.c_char_p, c_int, c_void_p, c_void_p,
                                  P(ModelC)]
        L.wpref_fit.restype = c_double
        L.wpref_fit.argtypes = [P(ModelC), c_uint, c_uint, ctypes.c_char_p, c_int, c_void_p, c_double, c_double,
                                c_void_p, P(ModelC), c_void_p]
        L.wpref_fit_estep.restype = c_double
        L.wpref_fit_estep.argtypes = [P(ModelC), c_uint, c_uint, ctypes.c_char_p, c_int, c_void_p, c_void_p]
        L.wpref_nn_search.restype = None
        L.wpref_nn_search.argtypes = [P(ModelC), c_uint, c_uint, c_int, c_double, c_int] + [c_void_p] * 5 + \
            [c_int, c_void_p, c_void_p, c_void_p, c_void_p]
        L.wpref_nn_classify.restype = None
        L.wpref_nn_classify.argtypes = [P(ModelC), c_uint, c_uint, c_int, c_double, c_int] + [c_void_p] * 5 + \
            [c_int, c_void_p, c_void_p, c_void_p, c_void_p]
        L.wpref_nn_classify_top.restype = None
        L.wpref_nn_classify_top.argtypes = [P(ModelC), c_uint, c_uint, c_int, c_double, c_int] + [c_void_p] * 5 + \
            [c_int, c_void_p, c_int, c_void_p, c_void_p, c_void_p, c_void_p]
        L.wpref_hybrid_classify.restype = None
        L.wpref_hybrid_classify.argtypes = [P(ModelC), c_uint, c_uint, c_double, c_int, c_double, c_int] + \
            [c_void_p] * 5 + [c_int, c_int, P(ctypes.c_char_p), c_void_p, c_void_p, c_int, c_void_p, c_void_p,
                              c_void_p, c_void_p]
        L.wpref_kdtree_order.restype = None
        L.wpref_kdtree_order.argtypes = [c_int, c_int, c_int, c_void_p, c_void_p]
        L.wpref_generate_dye_tracks.restype = None
        L.wpref_generate_dye_tracks.argtypes = [P(ModelC), c_uint, c_uint, ctypes.c_char_p, c_int, c_uint, c_void_p]
        L.wpref_write_radiometries.restype = None
        L.wpref_write_radiometries.argtypes = [ctypes.c_char_p, P(ModelC), c_uint, c_uint, c_int, c_void_p]
        L.wpref_read_radiometries.restype = c_int
        L.wpref_read_radiometries.argtypes = [ctypes.c_char_p, P(ModelC), c_void_p, c_int]
        L.wpref_write_dye_tracks.restype = None
        L.wpref_write_dye_tracks.argtypes = [ctypes.c_char_p, c_uint, c_uint, c_int] + [c_void_p] * 5
        L.wpref_read_dye_tracks.restype = c_int
        L.wpref_read_dye_tracks.argtypes = [ctypes.c_char_p] + [c_void_p] * 5 + [c_int, c_int, c_void_p]
        L.wpref_write_scored.restype = None
        L.wpref_write_scored.argtypes = [ctypes.c_char_p, c_int, c_void_p, c_void_p]
        L.wpref_classify.restype = None
        L.wpref_classify.argtypes = [P(ModelC), c_uint, c_uint, c_int, P(ctypes.c_char_p), c_void_p, c_void_p,
                                     c_int, c_void_p, c_double, c_void_p, c_void_p, c_void_p]
        self.L = L

    @classmethod
    def load(cls):
        if not os.path.exists(REF_SO):
            return None
        return cls(ctypes.CDLL(REF_SO))

    def write_radiometries(self, path, model, rads):
        r = np.ascontiguousarray(rads, dtype=np.float64)
        N, T, C = r.shape
        self.L.wpref_write_radiometries(str(path).encode(), ctypes.byref(Oracle._m(model)), T, C, N, _ptr(r))

    def read_radiometries(self, path, model, T, C, cap):
        out = np.zeros(cap * T * C)
        n = self.L.wpref_read_radiometries(str(path).encode(), ctypes.byref(Oracle._m(model)), _ptr(out), out.size)
        return out[:n * T * C].reshape(n, T, C)

    def write_dye_tracks(self, path, tracks):
        c, off, ids, cnt, hits = tracks._arrays()
        self.L.wpref_write_dye_tracks(str(path).encode(), tracks.num_timesteps, tracks.num_channels, len(tracks),
                                      _ptr(c), _ptr(off), _ptr(ids), _ptr(cnt), _ptr(hits))

    def read_dye_tracks(self, path, T, C, cap_t, cap_s):
        counts = np.zeros((cap_t, T, C), dtype=np.int16)
        off = np.zeros(cap_t + 1, dtype=np.int32)
        ids, cnt, hits = (np.zeros(cap_s, dtype=np.int32) for _ in range(3))
        ns = ctypes.c_int()
        n = self.L.wpref_read_dye_tracks(str(path).encode(), _ptr(counts), _ptr(off), _ptr(ids), _ptr(cnt),
                                         _ptr(hits), cap_t, cap_s, ctypes.byref(ns))
        k = ns.value
        return counts[:n], off[:n + 1], ids[:k], cnt[:k], hits[:k]

    def write_scored(self, path, ids, scores):
        i = np.ascontiguousarray(ids, dtype=np.int32)
        s = np.ascontiguousarray(scores, dtype=np.float64)
        self.L.wpref_write_scored(str(path).encode(), len(i), _ptr(i), _ptr(s))

    def generate_dye_tracks(self, model, T, dye_string, n, seed):
        """The reference's generate_dye_track n times from one seeded engine: [n, T, C] counts."""
        C = model.num_channels
        out = np.zeros((n, T, C), dtype=np.int16)
        self.L.wpref_generate_dye_tracks(ctypes.byref(Oracle._m(model)), T, C, dye_string.encode(), n, seed,
                                         _ptr(out))
        return out

    def pdf(self, model, channel, observed, counts):
        m = Oracle._m(model)
        cnt = np.ascontiguousarray(counts, dtype=np.uint32)
        return self.L.wpref_pdf(ctypes.byref(m), channel, observed, _ptr(cnt))

    def pair_probability(self, model, dye_string, rad, M, cutoff=DBL_MAX):
        m = Oracle._m(model)
        rad = np.ascontiguousarray(rad, dtype=np.float64)
        T, C = rad.shape
        return self.L.wpref_pair_probability(ctypes.byref(m), T, C, dye_string.encode(), _ptr(rad), M, cutoff)

    def improve_fit(self, model, dye_string, rad, M):
        m = Oracle._m(model)
        rad = np.ascontiguousarray(rad, dtype=np.float64)
        T, C = rad.shape
        st = np.zeros(fit_stats_len(T, C))
        p = self.L.wpref_improve_fit(ctypes.byref(m), T, C, dye_string.encode(), _ptr(rad), M, _ptr(st))
        return p, st

    def offset_exponential_fit(self, y, w, hold, p):
        y = np.ascontiguousarray(y, dtype=np.float64)
        w = np.ascontiguousarray(w, dtype=np.float64)
        h = np.ascontiguousarray(hold, dtype=np.int32)
        out = np.ascontiguousarray(p, dtype=np.float64).copy()
        self.L.wpref_offset_exponential_fit(len(y), _ptr(y), _ptr(w), _ptr(h), _ptr(out))
        return out

    def mstep(self, model, T, C, dye_string, num_reads, stats, holds=None):
        m = Oracle._m(model)
        st = np.ascontiguousarray(stats, dtype=np.float64)
        h = np.ascontiguousarray(holds if holds is not None else np.zeros(6 + 2 * C), dtype=np.int32)
        out = ModelC()
        step = self.L.wpref_mstep(ctypes.byref(m), T, C, dye_string.encode(), num_reads, _ptr(st), _ptr(h),
                                  ctypes.byref(out))
        return out, step

    def fit(self, model, dye_string, rads, stopping_threshold, max_runtime=1e30, holds=None):
        m = Oracle._m(model)
        rads = np.ascontiguousarray(rads, dtype=np.float64)
        N, T, C = rads.shape
        h = np.ascontiguousarray(holds if holds is not None else np.zeros(6 + 2 * C), dtype=np.int32)
        out = ModelC()
        step = np.zeros(1)
        log_l = self.L.wpref_fit(ctypes.byref(m), T, C, dye_string.encode(), N, _ptr(rads), stopping_threshold,
                                 max_runtime, _ptr(h), ctypes.byref(out), _ptr(step))
        return out, float(step[0]), log_l

    def fit_estep(self, model, dye_string, rads):
        """One serial E-step of HMMFitter::fit: (statistics, sum of ln P)."""
        m = Oracle._m(model)
        rads = np.ascontiguousarray(rads, dtype=np.float64)
        N, T, C = rads.shape
        st = np.zeros(fit_stats_len(T, C))
        lp = self.L.wpref_fit_estep(ctypes.byref(m), T, C, dye_string.encode(), N, _ptr(rads), _ptr(st))
        return st, lp

    def kdtree_order(self, points, k):
        X = np.ascontiguousarray(points, dtype=np.float64)
        n, d = X.shape
        out = np.zeros(n, dtype=np.int32)
        self.L.wpref_kdtree_order(k, d, n, _ptr(X), _ptr(out))
        return out

    @staticmethod
    def _tracks(tracks):
        return [np.ascontiguousarray(a) for a in tracks._arrays()]

    def nn_search(self, model, tracks, k, sig, rads):
        m = Oracle._m(model)
        rads = np.ascontiguousarray(rads, dtype=np.float64)
        N, T, C = rads.shape
        ta = self._tracks(tracks)
        cnt = np.zeros(N, dtype=np.int32)
        idx = np.full((N, k), -1, dtype=np.int32)
        dist = np.zeros((N, k))
        self.L.wpref_nn_search(ctypes.byref(m), T, C, k, sig, len(tracks), *[_ptr(a) for a in ta], N, _ptr(rads),
                               _ptr(cnt), _ptr(idx), _ptr(dist))
        return cnt, idx, dist

    def nn_classify(self, model, tracks, k, sig, rads):
        m = Oracle._m(model)
        rads = np.ascontiguousarray(rads, dtype=np.float64)
        N, T, C = rads.shape
        ta = self._tracks(tracks)
        ids = np.zeros(N, dtype=np.int32)
        sc = np.zeros(N)
        tot = np.zeros(N)
        self.L.wpref_nn_classify(ctypes.byref(m), T, C, k, sig, len(tracks), *[_ptr(a) for a in ta], N, _ptr(rads),
                                 _ptr(ids), _ptr(sc), _ptr(tot))
        return ids, sc, tot

    def nn_classify_top(self, model, tracks, k, sig, rads, h):
        m = Oracle._m(model)
        rads = np.ascontiguousarray(rads, dtype=np.float64)
        N, T, C = rads.shape
        ta = self._tracks(tracks)
        cnt = np.zeros(N, dtype=np.int32)
        ids = np.zeros((N, max(h, 1)), dtype=np.int32)
        sc = np.zeros((N, max(h, 1)))
        tot = np.zeros(N)
        self.L.wpref_nn_classify_top(ctypes.byref(m), T, C, k, sig, len(tracks), *[_ptr(a) for a in ta], N,
                                     _ptr(rads), h, _ptr(cnt), _ptr(ids), _ptr(sc), _ptr(tot))
        return [list(zip(ids[i, :cnt[i]].tolist(), sc[i, :cnt[i]].tolist())) for i in range(N)], tot

    def hybrid_classify(self, model, tracks, k, sig, h, dye_strings, counts, ids, rads, cutoff=DBL_MAX):
        m = Oracle._m(model)
        rads = np.ascontiguousarray(rads, dtype=np.float64)
        N, T, C = rads.shape
        ta = self._tracks(tracks)
        arr = (ctypes.c_char_p * len(dye_strings))(*[s.encode() for s in dye_strings])
        counts = np.ascontiguousarray(counts, dtype=np.int32)
        ids = np.ascontiguousarray(ids, dtype=np.int32)
        oid = np.zeros(N, dtype=np.int32)
        osc = np.zeros(N)
        otot = np.zeros(N)
        self.L.wpref_hybrid_classify(ctypes.byref(m), T, C, cutoff, k, sig, len(tracks), *[_ptr(a) for a in ta], h,
                                     len(dye_strings), arr, _ptr(counts), _ptr(ids), N, _ptr(rads), _ptr(oid),
                                     _ptr(osc), _ptr(otot))
        return oid, osc, otot

    def classify(self, model, dye_strings, counts, ids, rads, cutoff=DBL_MAX):
        m = Oracle._m(model)
        rads = np.ascontiguousarray(rads, dtype=np.float64)
        N, T, C = rads.shape
        arr = (ctypes.c_char_p * len(dye_strings))(*[s.encode() for s in dye_strings])
        counts = np.ascontiguousarray(counts, dtype=np.int32)
        ids = np.ascontiguousarray(ids, dtype=np.int32)
        oid = np.zeros(N, dtype=np.int32)
        osc = np.zeros(N)
        otot = np.zeros(N)
        self.L.wpref_classify(ctypes.byref(m), T, C, len(dye_strings), arr, _ptr(counts), _ptr(ids), N,
                              _ptr(rads), cutoff, _ptr(oid), _ptr(osc), _ptr(otot))
        return oid, osc, otot


def fit_stats_len(T, C):
    return 6 + 2 * (T - 1) + 4 * C


def max_dyes_of(dye_string, C):
    """HMMFitter's max_num_dyes (hmm-fitter.cc:44-60)."""
    return max(dye_string.count(str(c)) for c in range(C))


def model_from_golden(g, C):
    return SequencingModel.uniform(C, p_edman_failure=g["p_edman_failure"], p_detach=g["p_detach"],
                                   p_initial_detach=g["p_initial_detach"],
                                   p_initial_detach_decay=g["p_initial_detach_decay"],
                                   p_initial_block=g["p_initial_block"], p_cyclic_block=g["p_cyclic_block"],
                                   p_bleach=g["p_bleach"], p_dud=g["p_dud"], bg_sig=g["bg_sig"], mu=g["mu"],
                                   sig=g["sig"])
