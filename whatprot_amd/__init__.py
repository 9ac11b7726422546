"""whatprot_amd: MI355X (gfx950) implementation of whatprot's tensor-state HMM
scoring path behind the reference's interfaces.

Host-side mirror of the reference classes this path touches (names and
argument meaning follow whatprot's cc_code/src):
  SequencingModel / ChannelModel  parameterization/model/sequencing-model.cc:48-151
  SequencingSettings              parameterization/settings/sequencing-settings.h:14-19
  DyeSeq sourced data             io/dye-seqs-io.cc:31-106
  Radiometry files                io/radiometries-io.cc:33-82
  ScoredClassification CSV        io/scored-classifications-io.cc:30-68
  HMMClassifier                   classifiers/hmm-classifier.h:30-78
  FitSettings                     parameterization/settings/fit-settings.cc:28-100
  Dye-track files                 io/dye-tracks-io.cc:32-103
  NNClassifier, HybridClassifier  classifiers/nn-classifier.h:44-82, hybrid-classifier.h:25-52
  HMMFitter                       fitters/hmm-fitter.h:30-151
All scoring runs in libwhatprot_amd.so on the GPU; there is no CPU path.
"""
import ctypes
import os
import json
import math
import sys
from dataclasses import dataclass, field
from typing import List, Optional, Sequence

import numpy as np

from . import _capi
from ._capi import (WP_MAX_CHANNELS, WP_PRECISION_FP32, WP_PRECISION_FP64, DyeTracksC, FitSettingsC, ModelC,
                    RadiometriesC, WhatprotError, check, fit_stats_len, lib)

DBL_MAX = sys.float_info.max

__all__ = [
    "ChannelModel", "SequencingModel", "SequencingSettings", "DyeSeqEntry", "ScoredClassification",
    "HMMClassifier", "HMMFitter", "FitSettings", "fit_mstep", "offset_exponential_fit", "fit_stats_len",
    "DyeTracks", "read_dye_tracks", "write_dye_tracks", "simulate_dye_tracks", "kdtree_order", "NNClassifier",
    "HybridClassifier", "ShortlistMLP", "read_dye_seqs", "write_dye_seqs", "read_radiometries", "write_radiometries",
    "write_scored_classifications", "generate_dye_seqs", "simulate_radiometries",
    "WP_PRECISION_FP32", "WP_PRECISION_FP64", "WhatprotError", "DBL_MAX", "lib",
]


@dataclass
class ChannelModel:
    """ChannelModel (channel-model.h:20-83); interactions hold 1 - effect."""
    p_bleach: float = 0.0
    p_dud: float = 0.0
    bg_sig: float = 0.0
    mu: float = 1.0
    sig: float = 0.0
    interactions: List[float] = field(default_factory=list)
    flat_interactions: List[float] = field(default_factory=list)


@dataclass
class SequencingModel:
    """SequencingModel (sequencing-model.h:22-49) with DecayingRateModel detach."""
    p_edman_failure: float = 0.0
    p_detach: float = 0.0
    p_initial_detach: float = 0.0
    p_initial_detach_decay: float = 0.0
    p_initial_block: float = 0.0
    p_cyclic_block: float = 0.0
    channel_models: List[ChannelModel] = field(default_factory=list)

    @property
    def num_channels(self):
        return len(self.channel_models)

    @classmethod
    def from_json(cls, path_or_dict):
        """Parses the seqparams JSON exactly as sequencing-model.cc:48-92 reads it
        (p_initial_detach and p_initial_detach_decay are required; optional
        interactions store 1 - effect / 1 - flat_effect)."""
        data = path_or_dict
        if not isinstance(data, dict):
            with open(path_or_dict) as f:
                data = json.load(f)
        chans = data["channel_models"]
        C = len(chans)
        m = cls(p_edman_failure=float(data["p_edman_failure"]), p_detach=float(data["p_detach"]),
                p_initial_detach=float(data["p_initial_detach"]),
                p_initial_detach_decay=float(data["p_initial_detach_decay"]),
                p_initial_block=float(data["p_initial_block"]),
                p_cyclic_block=float(data["p_cyclic_block"]))
        for cd in chans:
            cm = ChannelModel(p_bleach=float(cd["p_bleach"]), p_dud=float(cd["p_dud"]),
                              bg_sig=float(cd["bg_sig"]), mu=float(cd["mu"]), sig=float(cd["sig"]),
                              interactions=[1.0] * C, flat_interactions=[1.0] * C)
            for it in cd.get("interactions", []) or []:
                o = int(it["other_channel"])
                if "effect" in it:
                    cm.interactions[o] = 1.0 - float(it["effect"])
                if "flat_effect" in it:
                    cm.flat_interactions[o] = 1.0 - float(it["flat_effect"])
            m.channel_models.append(cm)
        return m

    def with_mu_as_one(self):
        """sequencing-model.cc:140-151 / channel-model.cc:35-45."""
        out = SequencingModel(self.p_edman_failure, self.p_detach, self.p_initial_detach,
                              self.p_initial_detach_decay, self.p_initial_block, self.p_cyclic_block)
        for cm in self.channel_models:
            out.channel_models.append(ChannelModel(cm.p_bleach, cm.p_dud, cm.bg_sig / cm.mu, 1.0,
                                                   cm.sig / cm.mu, list(cm.interactions),
                                                   list(cm.flat_interactions)))
        return out

    def p_detach_at(self, i):
        """DecayingRateModel::operator[] (decaying-rate-model.cc:47-50)."""
        return self.p_detach + self.p_initial_detach * math.exp(-float(i) * self.p_initial_detach_decay)

    def to_c(self) -> ModelC:
        m = ModelC()
        C = self.num_channels
        if C > WP_MAX_CHANNELS:
            raise WhatprotError(f"at most {WP_MAX_CHANNELS} channels")
        m.num_channels = C
        m.p_edman_failure = self.p_edman_failure
        m.p_detach = self.p_detach
        m.p_initial_detach = self.p_initial_detach
        m.p_initial_detach_decay = self.p_initial_detach_decay
        m.p_initial_block = self.p_initial_block
        m.p_cyclic_block = self.p_cyclic_block
        for c, cm in enumerate(self.channel_models):
            ch = m.channels[c]
            ch.p_bleach, ch.p_dud, ch.bg_sig, ch.mu, ch.sig = cm.p_bleach, cm.p_dud, cm.bg_sig, cm.mu, cm.sig
            for o in range(WP_MAX_CHANNELS):
                ch.interactions[o] = cm.interactions[o] if o < len(cm.interactions) else 1.0
                ch.flat_interactions[o] = cm.flat_interactions[o] if o < len(cm.flat_interactions) else 1.0
        return m

    @classmethod
    def from_c(cls, m: ModelC, like: Optional["SequencingModel"] = None):
        """Inverse of to_c for m.num_channels channels."""
        C = m.num_channels
        out = cls(m.p_edman_failure, m.p_detach, m.p_initial_detach, m.p_initial_detach_decay,
                  m.p_initial_block, m.p_cyclic_block)
        for c in range(C):
            ch = m.channels[c]
            out.channel_models.append(ChannelModel(ch.p_bleach, ch.p_dud, ch.bg_sig, ch.mu, ch.sig,
                                                   [ch.interactions[o] for o in range(C)],
                                                   [ch.flat_interactions[o] for o in range(C)]))
        return out

    def distance(self, other: "SequencingModel"):
        """SequencingModel::distance (sequencing-model.cc:153-166)."""
        d = max(abs(self.p_edman_failure - other.p_edman_failure), abs(self.p_detach - other.p_detach),
                abs(self.p_initial_detach - other.p_initial_detach),
                abs(self.p_initial_detach_decay - other.p_initial_detach_decay),
                abs(self.p_initial_block - other.p_initial_block), abs(self.p_cyclic_block - other.p_cyclic_block))
        for a, b in zip(self.channel_models, other.channel_models):
            d = max(d, abs(a.p_bleach - b.p_bleach), abs(a.p_dud - b.p_dud))
        return d

    @classmethod
    def uniform(cls, C, p_edman_failure=0.06, p_detach=0.05, p_initial_detach=0.04,
                p_initial_detach_decay=0.3, p_initial_block=0.07, p_cyclic_block=0.02,
                p_bleach=0.05, p_dud=0.07, bg_sig=0.00667, mu=1.0, sig=0.16):
        """A model with identical channels and no interactions."""
        m = cls(p_edman_failure, p_detach, p_initial_detach, p_initial_detach_decay,
                p_initial_block, p_cyclic_block)
        for _ in range(C):
            m.channel_models.append(ChannelModel(p_bleach, p_dud, bg_sig, mu, sig, [1.0] * C, [1.0] * C))
        return m


@dataclass
class SequencingSettings:
    """SequencingSettings (sequencing-settings.h:14-19): the -p sigma cutoff."""
    dist_cutoff: float = DBL_MAX


@dataclass
class DyeSeqEntry:
    """SourcedData<DyeSeq, SourceCount<int>> as read by dye-seqs-io.cc:96-103."""
    dye_string: str
    count: int = 1
    id: int = 0


@dataclass
class ScoredClassification:
    """scored-classification.h:16-27; adjusted_score is what the CSV holds."""
    id: int
    adjusted_score: float
    log_score: float
    log_total: float


def read_dye_seqs(path):
    """dye-seqs-io.cc:31-106: C, N, then rows 'dye_string count id'."""
    with open(path) as f:
        tok = f.read().split()
    C, N = int(tok[0]), int(tok[1])
    out = []
    for i in range(N):
        s, cnt, idv = tok[2 + 3 * i], int(tok[3 + 3 * i]), int(tok[4 + 3 * i])
        out.append(DyeSeqEntry(s, cnt, idv))
    return C, out


def write_dye_seqs(path, C, entries: Sequence[DyeSeqEntry]):
    with open(path, "w") as f:
        f.write(f"{C}\n{len(entries)}\n")
        for e in entries:
            f.write(f"{e.dye_string}\t{e.count}\t{e.id}\n")


def read_radiometries(path, seq_model: SequencingModel):
    """read_radiometries (radiometries-io.cc:33-56) through wp_read_radiometries:
    T, C, N then N rows of T*C values (t outer, c inner), each divided by that
    channel's mu. Returns (T, C, array[N,T,C])."""
    out = RadiometriesC()
    m = seq_model.to_c()
    check(lib().wp_read_radiometries(os.fsencode(path), ctypes.byref(m), ctypes.byref(out)))
    try:
        T, C, N = out.num_timesteps, out.num_channels, out.num_reads
        data = np.ctypeslib.as_array(out.values, shape=(max(N * T * C, 1),))[:N * T * C].copy()
    finally:
        lib().wp_free_radiometries(ctypes.byref(out))
    return T, C, data.reshape(N, T, C)


def write_radiometries(path, seq_model: SequencingModel, rads):
    """write_radiometries (radiometries-io.cc:58-82) through
    wp_write_radiometries: values multiplied back by mu, 17 digits."""
    r = np.ascontiguousarray(rads, dtype=np.float64)
    N, T, C = r.shape
    m = seq_model.to_c()
    check(lib().wp_write_radiometries(os.fsencode(path), ctypes.byref(m), T, C, N, r.ctypes.data))


def write_scored_classifications(path, results: Sequence[ScoredClassification]):
    """write_scored_classifications (scored-classifications-io.cc:30-68)."""
    ids = np.ascontiguousarray([r.id for r in results], dtype=np.int32)
    sc = np.ascontiguousarray([r.adjusted_score for r in results], dtype=np.float64)
    check(lib().wp_write_scored_classifications(os.fsencode(path), len(results),
                                                ids.ctypes.data if len(ids) else None,
                                                sc.ctypes.data if len(sc) else None))


def generate_dye_seqs(C, n, len_min, len_max, dyes_min, dyes_max, seed):
    stride = len_max + 1
    buf = ctypes.create_string_buffer(n * stride)
    check(lib().wp_generate_dye_seqs(C, n, len_min, len_max, dyes_min, dyes_max, seed, buf, stride))
    raw = buf.raw
    return [raw[i * stride:(i + 1) * stride].split(b"\0", 1)[0].decode() for i in range(n)]


def _cstrings(strings):
    arr = (ctypes.c_char_p * max(1, len(strings)))()
    enc = [s.encode() for s in strings]
    for i, s in enumerate(enc):
        arr[i] = s
    return arr, enc


def simulate_radiometries(seq_model: SequencingModel, T, dye_strings, n, seed, weights=None, gpu=False):
    """Returns (rads[n,T,C] in the model's units, dye-seq index per read).
    gpu=True runs wp_simulate_radiometries_gpu (the same streams on the
    device: identical dye tracks, intensities equal up to the last float32
    bit)."""
    C = seq_model.num_channels
    m = seq_model.to_c()
    arr, keep = _cstrings(dye_strings)
    out = np.zeros((n, T, C), dtype=np.float64)
    idx = np.zeros(n, dtype=np.int32)
    w = None
    if weights is not None:
        w = np.ascontiguousarray(weights, dtype=np.int32)
    fn = lib().wp_simulate_radiometries_gpu if gpu else lib().wp_simulate_radiometries
    check(fn(ctypes.byref(m), T, C, len(dye_strings), arr, w.ctypes.data if w is not None else None, n, seed,
             out.ctypes.data, idx.ctypes.data))
    return out, idx


class HMMClassifier:
    """HMMClassifier (classifiers/hmm-classifier.h:30-78) on the GPU.

    HMMClassifier(num_timesteps, num_channels, seq_model, seq_settings,
    dye_seqs) as in run-classify-hmm.cc:80-82, plus `precision`
    (WP_PRECISION_FP64 reproduces the reference's arithmetic;
    WP_PRECISION_FP32 is the rescaled batched path)."""

    def __init__(self, num_timesteps, num_channels, seq_model: SequencingModel,
                 seq_settings: SequencingSettings, dye_seqs: Sequence[DyeSeqEntry],
                 precision=WP_PRECISION_FP64):
        self.num_timesteps = int(num_timesteps)
        self.num_channels = int(num_channels)
        self.dye_seqs = list(dye_seqs)
        self._m = seq_model.to_c()
        arr, self._keep = _cstrings([d.dye_string for d in self.dye_seqs])
        S = len(self.dye_seqs)
        counts = (ctypes.c_int * max(1, S))(*[d.count for d in self.dye_seqs])
        ids = (ctypes.c_int * max(1, S))(*[d.id for d in self.dye_seqs])
        h = ctypes.c_void_p()
        check(lib().wp_hmm_classifier_create(self.num_timesteps, self.num_channels, ctypes.byref(self._m),
                                             float(seq_settings.dist_cutoff), S, arr, counts, ids,
                                             int(precision), ctypes.byref(h)))
        self._h = h
        self.max_num_dyes = lib().wp_hmm_classifier_max_num_dyes(h)

    def close(self):
        if getattr(self, "_h", None):
            lib().wp_hmm_classifier_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _rads(self, radiometries):
        r = np.ascontiguousarray(radiometries, dtype=np.float64)
        if r.ndim != 3 or r.shape[1] != self.num_timesteps or r.shape[2] != self.num_channels:
            raise WhatprotError(f"radiometries must be [N, {self.num_timesteps}, {self.num_channels}]")
        return r

    def classify_arrays(self, radiometries, candidates: Optional[Sequence[Sequence[int]]] = None):
        """Returns (ids, adjusted, log_score, log_total) arrays."""
        r = self._rads(radiometries)
        N = r.shape[0]
        ids = np.zeros(N, dtype=np.int32)
        adj = np.zeros(N)
        ls = np.zeros(N)
        lt = np.zeros(N)
        if candidates is None:
            check(lib().wp_hmm_classify(self._h, N, r.ctypes.data, ids.ctypes.data, adj.ctypes.data,
                                        ls.ctypes.data, lt.ctypes.data))
        else:
            off = np.zeros(N + 1, dtype=np.int32)
            off[1:] = np.cumsum([len(c) for c in candidates])
            idx = np.ascontiguousarray(np.concatenate([np.asarray(c, dtype=np.int32) for c in candidates])
                                       if off[-1] else np.zeros(1, dtype=np.int32), dtype=np.int32)
            check(lib().wp_hmm_classify_candidates(self._h, N, r.ctypes.data, off.ctypes.data,
                                                   idx.ctypes.data, ids.ctypes.data, adj.ctypes.data,
                                                   ls.ctypes.data, lt.ctypes.data))
        return ids, adj, ls, lt

    def classify_csr(self, radiometries, offsets, indices):
        """classify_arrays over candidate lists given as CSR arrays (offsets
        [N + 1], indices into the dye-seq table)."""
        r = self._rads(radiometries)
        N = r.shape[0]
        off = np.ascontiguousarray(offsets, dtype=np.int32)
        idx = np.ascontiguousarray(indices if len(indices) else np.zeros(1), dtype=np.int32)
        ids = np.zeros(N, dtype=np.int32)
        adj, ls, lt = np.zeros(N), np.zeros(N), np.zeros(N)
        check(lib().wp_hmm_classify_candidates(self._h, N, r.ctypes.data, off.ctypes.data, idx.ctypes.data,
                                               ids.ctypes.data, adj.ctypes.data, ls.ctypes.data, lt.ctypes.data))
        return ids, adj, ls, lt

    def classify(self, radiometries, candidates=None) -> List[ScoredClassification]:
        """HMMClassifier::classify(vector<Radiometry>) (hmm-classifier.cc:84-93),
        or per-read candidate_indices (hmm-classifier.cc:79-82)."""
        ids, adj, ls, lt = self.classify_arrays(radiometries, candidates)
        return [ScoredClassification(int(i), float(a), float(s), float(t)) for i, a, s, t in zip(ids, adj, ls, lt)]

    def pair_log_probability(self, radiometries):
        """ln PeptideHMM::probability() for every (read, dye-seq) pair."""
        r = self._rads(radiometries)
        out = np.zeros((r.shape[0], len(self.dye_seqs)))
        check(lib().wp_hmm_pair_log_probability(self._h, r.shape[0], r.ctypes.data, out.ctypes.data))
        return out

    def set_timing(self, enable=True):
        check(lib().wp_hmm_set_timing(self._h, 1 if enable else 0))

    def last_timing(self):
        """(prep_ms, classify_ms) of the last call with timing enabled."""
        a, b = ctypes.c_double(), ctypes.c_double()
        check(lib().wp_hmm_last_timing(self._h, ctypes.byref(a), ctypes.byref(b)))
        return a.value, b.value

    def set_variant(self, mode):
        """Quad-rows kernel variant: -1 autotune on the next call (default),
        0 bin-wide boxes, 1 per-timestep box classes (identical outputs)."""
        check(lib().wp_hmm_set_variant(self._h, int(mode)))

    def variants(self):
        """(per-bin variant chosen, or -1 before the first call; per-bin
        [variant 0 ms, variant 1 ms] from the autotuner, -1 where untimed)."""
        n = lib().wp_hmm_get_variants(self._h, None, None, 0)
        check(min(n, 0))
        v = np.zeros(max(n, 1), dtype=np.int32)
        ms = np.zeros((max(n, 1), 2), dtype=np.float64)
        check(min(lib().wp_hmm_get_variants(self._h, v.ctypes.data, ms.ctypes.data, n), 0))
        return v[:n], ms[:n]

    def set_graph(self, enable=True):
        """Graph replay of the device entry points' launch sequence."""
        check(lib().wp_hmm_set_graph(self._h, 1 if enable else 0))

    def set_branches(self, enable=True):
        """Bins as three parallel branches (default) or in order on the
        caller's stream (identical outputs; per-kernel durations for
        profiling)."""
        check(lib().wp_hmm_set_branches(self._h, 1 if enable else 0))

    def set_refine(self, enable=True, abs=2e-5, rel=1e-5):
        """FP32 classifiers: decide the candidates within abs + rel |ln P_best|
        of the FP32 best in FP64 (default on, with these defaults)."""
        check(lib().wp_hmm_set_refine(self._h, 1 if enable else 0, float(abs), float(rel)))

    def refine_stats(self, num_reads):
        """(reads rescored in FP64, of which over every candidate) of the last call."""
        out = np.zeros(2, dtype=np.int32)
        check(lib().wp_hmm_refine_stats(self._h, int(num_reads), out.ctypes.data))
        return int(out[0]), int(out[1])

    def work(self, radiometries):
        """(non-empty pairs, state-timesteps, algorithmic FLOPs) of classifying these reads."""
        r = self._rads(radiometries)
        out = np.zeros(3)
        check(lib().wp_hmm_work(self._h, r.shape[0], r.ctypes.data, out.ctypes.data))
        return tuple(float(x) for x in out)

    def classify_device(self, num_reads, d_rads, d_id, d_adjusted=0, d_log_score=0, d_log_total=0, stream=0):
        """Device-resident classify: integer device addresses, work queued on
        `stream` (a hipStream_t as an integer)."""
        check(lib().wp_hmm_classify_device(self._h, int(num_reads), d_rads, d_id, d_adjusted or None,
                                           d_log_score or None, d_log_total or None, stream or None))


@dataclass
class FitSettings:
    """FitSettings (parameterization/settings/fit-settings.h:20-37): which
    parameters are held at their starting values."""
    hold_p_edman_failure: bool = False
    hold_p_detach: bool = False
    hold_p_initial_detach: bool = False
    hold_p_initial_detach_decay: bool = False
    hold_p_initial_block: bool = False
    hold_p_cyclic_block: bool = False
    hold_p_bleach: List[bool] = field(default_factory=list)
    hold_p_dud: List[bool] = field(default_factory=list)

    @classmethod
    def from_json(cls, num_channels, path_or_dict):
        """Parses a fit settings JSON as fit-settings.cc:44-100 does: every key
        optional (false when absent); per-channel holds under
        "channel_settings" (one entry per channel, in order)."""
        data = path_or_dict
        if not isinstance(data, dict):
            with open(path_or_dict) as f:
                data = json.load(f)
        fs = cls(**{k: bool(data.get(k, False)) for k in (
            "hold_p_edman_failure", "hold_p_detach", "hold_p_initial_detach", "hold_p_initial_detach_decay",
            "hold_p_initial_block", "hold_p_cyclic_block")})
        if "channel_settings" in data:
            for cd in data["channel_settings"]:
                fs.hold_p_bleach.append(bool(cd.get("hold_p_bleach", False)))
                fs.hold_p_dud.append(bool(cd.get("hold_p_dud", False)))
        else:
            fs.hold_p_bleach = [False] * num_channels
            fs.hold_p_dud = [False] * num_channels
        return fs

    def to_c(self) -> FitSettingsC:
        c = FitSettingsC()
        for k in ("hold_p_edman_failure", "hold_p_detach", "hold_p_initial_detach", "hold_p_initial_detach_decay",
                  "hold_p_initial_block", "hold_p_cyclic_block"):
            setattr(c, k, 1 if getattr(self, k) else 0)
        for i, h in enumerate(self.hold_p_bleach[:WP_MAX_CHANNELS]):
            c.hold_p_bleach[i] = 1 if h else 0
        for i, h in enumerate(self.hold_p_dud[:WP_MAX_CHANNELS]):
            c.hold_p_dud[i] = 1 if h else 0
        return c


def offset_exponential_fit(y, w, hold, p):
    """least_squares_fit_of_offset_exponential (util/levmar-helper.cc:44-85);
    returns the fitted (base, initial, initial_decay)."""
    y = np.ascontiguousarray(y, dtype=np.float64)
    w = np.ascontiguousarray(w, dtype=np.float64)
    h = np.ascontiguousarray([1 if x else 0 for x in hold], dtype=np.int32)
    out = np.ascontiguousarray(p, dtype=np.float64).copy()
    check(lib().wp_fit_offset_exponential(len(y), y.ctypes.data, w.ctypes.data, h.ctypes.data, out.ctypes.data))
    return out


def fit_mstep(num_timesteps, num_channels, dye_string, seq_model: SequencingModel, num_reads, stats,
              settings: Optional[FitSettings] = None):
    """M-step of HMMFitter::fit (hmm-fitter.h:80-104): returns (next model, step size)."""
    st = np.ascontiguousarray(stats, dtype=np.float64)
    m = seq_model.to_c()
    out = ModelC()
    step = ctypes.c_double()
    fs = settings.to_c() if settings is not None else None
    check(lib().wp_fit_mstep(num_timesteps, num_channels, dye_string.encode(), ctypes.byref(m), int(num_reads),
                             st.ctypes.data, ctypes.byref(fs) if fs is not None else None, ctypes.byref(out),
                             ctypes.byref(step)))
    return SequencingModel.from_c(out), step.value


class HMMFitter:
    """HMMFitter (fitters/hmm-fitter.h:30-151) on the GPU, holding its reads.

    HMMFitter(num_timesteps, num_channels, dye_string, radiometries) as
    run-fit.cc:94-103 builds it (uncut ranges), plus `precision`. fit()
    mirrors HMMFitter::fit with an optional iteration cap."""

    def __init__(self, num_timesteps, num_channels, dye_string, radiometries, precision=WP_PRECISION_FP64):
        self.num_timesteps = int(num_timesteps)
        self.num_channels = int(num_channels)
        self.dye_string = dye_string
        r = np.ascontiguousarray(radiometries, dtype=np.float64)
        if r.ndim != 3 or r.shape[1] != self.num_timesteps or r.shape[2] != self.num_channels:
            raise WhatprotError(f"radiometries must be [N, {self.num_timesteps}, {self.num_channels}]")
        self.num_reads = r.shape[0]
        h = ctypes.c_void_p()
        check(lib().wp_hmm_fitter_create(self.num_timesteps, self.num_channels, dye_string.encode(),
                                         self.num_reads, r.ctypes.data, int(precision), ctypes.byref(h)))
        self._h = h

    def close(self):
        if getattr(self, "_h", None):
            lib().wp_hmm_fitter_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def estep(self, seq_model: SequencingModel):
        """(statistics [WP_FIT_STATS_LEN], sum of ln P) under seq_model."""
        st = np.zeros(fit_stats_len(self.num_timesteps, self.num_channels))
        slp = ctypes.c_double()
        m = seq_model.to_c()
        check(lib().wp_hmm_fitter_estep(self._h, ctypes.byref(m), st.ctypes.data, ctypes.byref(slp)))
        return st, slp.value

    def log_likelihood(self, seq_model: SequencingModel):
        """HMMFitter::log_likelihood (hmm-fitter.h:120-140)."""
        return self.estep(seq_model)[1]

    def fit(self, seq_model: SequencingModel, stopping_threshold, max_runtime=float("inf"),
            settings: Optional[FitSettings] = None, max_iterations=0):
        """HMMFitter::fit: returns (fitted model, step size, log L, iterations)."""
        m = seq_model.to_c()
        out = ModelC()
        step, log_l, it = ctypes.c_double(), ctypes.c_double(), ctypes.c_int()
        fs = settings.to_c() if settings is not None else None
        check(lib().wp_hmm_fit(self._h, ctypes.byref(m), ctypes.byref(fs) if fs is not None else None,
                               float(stopping_threshold), float(min(max_runtime, 1e300)), int(max_iterations),
                               ctypes.byref(out), ctypes.byref(step), ctypes.byref(log_l), ctypes.byref(it)))
        return SequencingModel.from_c(out), step.value, log_l.value, it.value

    def set_batched(self, enable=True):
        """Bootstrap rounds as one batch (default) or one after another."""
        check(lib().wp_hmm_fitter_set_batched(self._h, 1 if enable else 0))

    def bootstrap_fit(self, seq_model: SequencingModel, stopping_threshold, num_rounds, confidence_interval,
                      seed=1, max_runtime=float("inf"), settings: Optional[FitSettings] = None):
        """bootstrap_fit (fitters/bootstrap-fit.cc:37-173): returns a dict with
        the full fit ('model', 'step', 'log_l'), the rounds ('models',
        'log_ls'), 'ci_min', 'ci_max' and 'worst_step'."""
        m = seq_model.to_c()
        R = int(num_rounds)
        models = (ModelC * max(R, 1))()
        log_ls = np.zeros(max(R, 1))
        best, lo, hi = ModelC(), ModelC(), ModelC()
        step, log_l, worst = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
        fs = settings.to_c() if settings is not None else None
        check(lib().wp_hmm_fit_bootstrap(self._h, ctypes.byref(m), ctypes.byref(fs) if fs is not None else None,
                                         float(stopping_threshold), float(min(max_runtime, 1e300)), R,
                                         float(confidence_interval), int(seed), ctypes.cast(models, ctypes.c_void_p),
                                         log_ls.ctypes.data, ctypes.byref(best), ctypes.byref(step),
                                         ctypes.byref(log_l), ctypes.byref(lo), ctypes.byref(hi),
                                         ctypes.byref(worst)))
        return {"model": SequencingModel.from_c(best), "step": step.value, "log_l": log_l.value,
                "models": [SequencingModel.from_c(models[i]) for i in range(R)], "log_ls": log_ls[:R].copy(),
                "ci_min": SequencingModel.from_c(lo), "ci_max": SequencingModel.from_c(hi),
                "worst_step": worst.value}


def bootstrap_counts(num_reads, seed, round_index):
    """wp_fit_bootstrap_counts: multiplicity of each read in a bootstrap round."""
    out = np.zeros(max(int(num_reads), 1), dtype=np.int32)
    check(lib().wp_fit_bootstrap_counts(int(num_reads), int(seed), int(round_index), out.ctypes.data))
    return out[:int(num_reads)]


def write_params(path, num_channels, models: Sequence[SequencingModel], log_ls):
    """write_params (io/params-io.cc:37-54): one CSV row per bootstrap round,
    17 significant digits."""
    with open(path, "w") as f:
        f.write("p_edman_failure,p_detach,p_initial_detach,p_initial_detach_decay,p_initial_block,p_cyclic_block")
        for c in range(num_channels):
            f.write(f",ch{c}:p_bleach,ch{c}:p_dud")
        f.write(",log(L)\n")
        for m, ll in zip(models, log_ls):
            vals = [m.p_edman_failure, m.p_detach, m.p_initial_detach, m.p_initial_detach_decay,
                    m.p_initial_block, m.p_cyclic_block]
            for c in range(num_channels):
                vals += [m.channel_models[c].p_bleach, m.channel_models[c].p_dud]
            vals.append(ll)
            f.write(",".join(f"{v:.17g}" for v in vals) + "\n")


@dataclass
class DyeTracks:
    """Deduplicated dye-tracks with their sources (SourcedData<DyeTrack,
    SourceCountHitsList<int>>, io/dye-tracks-io.cc:32-62): counts[i, t, c],
    and for track i the sources offsets[i]:offsets[i+1] of ids / src_counts /
    hits."""
    num_timesteps: int
    num_channels: int
    counts: np.ndarray
    offsets: np.ndarray
    ids: np.ndarray
    src_counts: np.ndarray
    hits: np.ndarray

    def __len__(self):
        return self.counts.shape[0]

    def _arrays(self):
        return (np.ascontiguousarray(self.counts, dtype=np.int16), np.ascontiguousarray(self.offsets, dtype=np.int32),
                np.ascontiguousarray(self.ids, dtype=np.int32), np.ascontiguousarray(self.src_counts, dtype=np.int32),
                np.ascontiguousarray(self.hits, dtype=np.int32))


def read_dye_tracks(path) -> DyeTracks:
    """read_dye_tracks (io/dye-tracks-io.cc:32-62) through wp_read_dye_tracks:
    T, C, n, then per track T*C counts, the number of sources and
    (id, count, hits) per source."""
    out = DyeTracksC()
    check(lib().wp_read_dye_tracks(os.fsencode(path), ctypes.byref(out)))
    try:
        T, C, n, ns = out.num_timesteps, out.num_channels, out.num_dye_tracks, out.num_sources

        def arr(p, k, dt):
            return np.ctypeslib.as_array(p, shape=(max(k, 1),))[:k].astype(dt, copy=True)

        tracks = DyeTracks(T, C, arr(out.counts, n * T * C, np.int16).reshape(n, T, C),
                           arr(out.source_offsets, n + 1, np.int32), arr(out.source_ids, ns, np.int32),
                           arr(out.source_counts, ns, np.int32), arr(out.source_hits, ns, np.int32))
    finally:
        lib().wp_free_dye_tracks(ctypes.byref(out))
    return tracks


def write_dye_tracks(path, tracks: DyeTracks):
    """write_dye_tracks (io/dye-tracks-io.cc:64-103) through wp_write_dye_tracks."""
    c, off, ids, cnt, hits = tracks._arrays()
    check(lib().wp_write_dye_tracks(os.fsencode(path), tracks.num_timesteps, tracks.num_channels, len(tracks),
                                    c.ctypes.data, off.ctypes.data, ids.ctypes.data, cnt.ctypes.data,
                                    hits.ctypes.data))


def simulate_dye_tracks(seq_model: SequencingModel, T, dye_seqs: Sequence[DyeSeqEntry], per_peptide, seed,
                        gpu=False, max_tracks=None) -> DyeTracks:
    """generate_dye_tracks + dedup (generate-dye-tracks.cc:28-63,
    dedup-dye-tracks.cc:35-148) with this library's seeded generator;
    gpu=True simulates and deduplicates on the device (identical output)."""
    m = seq_model.to_c()
    C = seq_model.num_channels
    arr, keep = _cstrings([d.dye_string for d in dye_seqs])
    n = len(dye_seqs)
    cnt = np.ascontiguousarray([d.count for d in dye_seqs], dtype=np.int32)
    ids = np.ascontiguousarray([d.id for d in dye_seqs], dtype=np.int32)
    peptides = int(max(1, cnt.clip(min=0).sum() * per_peptide))
    cap = peptides if max_tracks is None else int(max_tracks)
    counts = np.zeros((cap, T, C), dtype=np.int16)
    off = np.zeros(cap + 1, dtype=np.int32)
    sid = np.zeros(peptides, dtype=np.int32)  # one source per (track, id): at most one per peptide
    scnt = np.zeros(peptides, dtype=np.int32)
    shits = np.zeros(peptides, dtype=np.int32)
    nt, ns_ = ctypes.c_int(), ctypes.c_int()
    fn = lib().wp_simulate_dye_tracks_gpu if gpu else lib().wp_simulate_dye_tracks
    check(fn(ctypes.byref(m), T, C, n, arr, cnt.ctypes.data, ids.ctypes.data, int(per_peptide), int(seed), cap,
             peptides, ctypes.byref(nt), ctypes.byref(ns_), counts.ctypes.data, off.ctypes.data, sid.ctypes.data,
             scnt.ctypes.data, shits.ctypes.data))
    k = nt.value
    ns = int(off[k])
    return DyeTracks(T, C, counts[:k].copy(), off[:k + 1].copy(), sid[:ns].copy(), scnt[:ns].copy(),
                     shits[:ns].copy())


def kdtree_order(points, k):
    """The point order the kd-tree construction leaves (make-node.h:28-80):
    (order, depth)."""
    X = np.ascontiguousarray(points, dtype=np.float64)
    n, d = X.shape
    out = np.zeros(max(n, 1), dtype=np.int32)
    depth = ctypes.c_int()
    check(lib().wp_kdtree_order(int(k), d, n, X.ctypes.data, out.ctypes.data, ctypes.byref(depth)))
    return out[:n], depth.value


class NNClassifier:
    """NNClassifier (classifiers/nn-classifier.h:44-82): kd-tree kNN over
    simulated dye-tracks, searched on the GPU."""

    def __init__(self, num_timesteps, num_channels, seq_model: SequencingModel, k, sig, dye_tracks: DyeTracks):
        self.num_timesteps, self.num_channels, self.k = int(num_timesteps), int(num_channels), int(k)
        self._m = seq_model.to_c()
        counts, off, ids, cnts, hits = dye_tracks._arrays()
        h = ctypes.c_void_p()
        check(lib().wp_nn_classifier_create(self.num_timesteps, self.num_channels, ctypes.byref(self._m), self.k,
                                            float(sig), len(dye_tracks), counts.ctypes.data, off.ctypes.data,
                                            ids.ctypes.data, cnts.ctypes.data, hits.ctypes.data, ctypes.byref(h)))
        self._h = h

    def close(self):
        if getattr(self, "_h", None):
            lib().wp_nn_classifier_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _rads(self, radiometries):
        r = np.ascontiguousarray(radiometries, dtype=np.float64)
        if r.ndim != 3 or r.shape[1] != self.num_timesteps or r.shape[2] != self.num_channels:
            raise WhatprotError(f"radiometries must be [N, {self.num_timesteps}, {self.num_channels}]")
        return r

    def search(self, radiometries):
        """KDTree::search per read: (counts [N], tracks [N, k], dists_sq [N, k]),
        farthest first."""
        r = self._rads(radiometries)
        N = r.shape[0]
        cnt = np.zeros(N, dtype=np.int32)
        idx = np.zeros((N, self.k), dtype=np.int32)
        dist = np.zeros((N, self.k))
        check(lib().wp_nn_search(self._h, N, r.ctypes.data, cnt.ctypes.data, idx.ctypes.data, dist.ctypes.data))
        return cnt, idx, dist

    def classify_arrays(self, radiometries):
        """(ids, scores, totals) of NNClassifier::classify."""
        r = self._rads(radiometries)
        N = r.shape[0]
        ids = np.zeros(N, dtype=np.int32)
        sc = np.zeros(N)
        tot = np.zeros(N)
        check(lib().wp_nn_classify(self._h, N, r.ctypes.data, ids.ctypes.data, sc.ctypes.data, tot.ctypes.data))
        return ids, sc, tot

    def classify(self, radiometries) -> List[ScoredClassification]:
        ids, sc, tot = self.classify_arrays(radiometries)
        return [ScoredClassification(int(i), float(s / t) if t else 0.0, math.log(s) if s > 0 else -math.inf,
                                     math.log(t) if t > 0 else -math.inf) for i, s, t in zip(ids, sc, tot)]

    def classify_top(self, radiometries, h):
        """NNClassifier::classify(radiometry, h): per read a list of (id, score)
        in ascending score order, and the totals."""
        r = self._rads(radiometries)
        N = r.shape[0]
        cnt = np.zeros(N, dtype=np.int32)
        ids = np.zeros((N, max(h, 1)), dtype=np.int32)
        sc = np.zeros((N, max(h, 1)))
        tot = np.zeros(N)
        check(lib().wp_nn_classify_top(self._h, N, r.ctypes.data, int(h), cnt.ctypes.data, ids.ctypes.data,
                                       sc.ctypes.data, tot.ctypes.data))
        return [list(zip(ids[i, :cnt[i]].tolist(), sc[i, :cnt[i]].tolist())) for i in range(N)], tot


class HybridClassifier:
    """HybridClassifier (classifiers/hybrid-classifier.h:25-52): kNN top-h
    candidates rescored by the HMM, both on the GPU."""

    def __init__(self, num_timesteps, num_channels, seq_model: SequencingModel, seq_settings: SequencingSettings,
                 k, sig, dye_tracks: DyeTracks, h, dye_seqs: Sequence[DyeSeqEntry], precision=WP_PRECISION_FP64):
        self.num_timesteps, self.num_channels = int(num_timesteps), int(num_channels)
        self._m = seq_model.to_c()
        counts, off, ids, cnts, hits = dye_tracks._arrays()
        arr, self._keep = _cstrings([d.dye_string for d in dye_seqs])
        S = len(dye_seqs)
        dcounts = (ctypes.c_int * max(1, S))(*[d.count for d in dye_seqs])
        dids = (ctypes.c_int * max(1, S))(*[d.id for d in dye_seqs])
        hnd = ctypes.c_void_p()
        check(lib().wp_hybrid_classifier_create(self.num_timesteps, self.num_channels, ctypes.byref(self._m),
                                                float(seq_settings.dist_cutoff), int(k), float(sig), len(dye_tracks),
                                                counts.ctypes.data, off.ctypes.data, ids.ctypes.data,
                                                cnts.ctypes.data, hits.ctypes.data, int(h), S, arr, dcounts, dids,
                                                int(precision), ctypes.byref(hnd)))
        self._h = hnd

    def close(self):
        if getattr(self, "_h", None):
            lib().wp_hybrid_classifier_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def classify_arrays(self, radiometries):
        """(ids, adjusted scores) of HybridClassifier::classify."""
        r = np.ascontiguousarray(radiometries, dtype=np.float64)
        if r.ndim != 3 or r.shape[1] != self.num_timesteps or r.shape[2] != self.num_channels:
            raise WhatprotError(f"radiometries must be [N, {self.num_timesteps}, {self.num_channels}]")
        N = r.shape[0]
        ids = np.zeros(N, dtype=np.int32)
        adj = np.zeros(N)
        check(lib().wp_hybrid_classify(self._h, N, r.ctypes.data, ids.ctypes.data, adj.ctypes.data))
        return ids, adj

    def classify(self, radiometries) -> List[ScoredClassification]:
        ids, adj = self.classify_arrays(radiometries)
        return [ScoredClassification(int(i), float(a), math.nan, math.nan) for i, a in zip(ids, adj)]


def dye_counts(dye_strings, C):
    """Dyes of each channel per dye-seq, [S][C] (DyeSeq digits, dye-seq.cc:23-40)."""
    out = np.zeros((len(dye_strings), C), dtype=np.int32)
    for i, s in enumerate(dye_strings):
        for ch in s:
            if ch.isdigit() and int(ch) < C:
                out[i, int(ch)] += 1
    return out


def hmm_labels(classifier: "HMMClassifier", radiometries, labels, chunk=250000):
    """Training labels from the HMM (no reference counterpart): for each read
    simulated from dye-seq labels[r], the classifier's pick among the dye-seqs
    with the same dye count in every channel (in table order, so exact ties go
    to the lowest index, as HMMClassifier decides them). Reads of a dye-seq
    that a pruned read cannot tell from other dye-seqs (equal P) are labelled
    with the one the HMM returns, which the shortlist then learns to propose.
    Returns dye-seq indices (the classifier's ids mapped back to indices)."""
    strings = [e.dye_string for e in classifier.dye_seqs]
    S = len(strings)
    D = dye_counts(strings, classifier.num_channels)
    _, group = np.unique(D, axis=0, return_inverse=True)
    group = group.reshape(-1)
    order = np.lexsort((np.arange(S), group))       # by group, then index
    gsize = np.bincount(group)
    gstart = np.concatenate([[0], np.cumsum(gsize)[:-1]])
    index_of = {int(e.id): i for i, e in enumerate(classifier.dye_seqs)}
    if len(index_of) != S:
        raise WhatprotError("hmm_labels needs unique dye-seq ids")
    rads = np.asarray(radiometries, dtype=np.float64).reshape(len(labels), classifier.num_timesteps,
                                                               classifier.num_channels)
    labels = np.asarray(labels, dtype=np.int64)
    out = np.empty(len(labels), dtype=np.int32)
    for r0 in range(0, len(labels), chunk):
        lab = labels[r0:r0 + chunk]
        g = group[lab]
        cnt = gsize[g]
        off = np.zeros(len(lab) + 1, dtype=np.int64)
        off[1:] = np.cumsum(cnt)
        pos = np.arange(off[-1]) - np.repeat(off[:-1], cnt) + np.repeat(gstart[g], cnt)
        ids = classifier.classify_csr(rads[r0:r0 + chunk], off, order[pos])[0]
        out[r0:r0 + chunk] = [index_of[int(i)] for i in ids]
    return out


class ShortlistMLP:
    """The learned shortlist's MLP (BF16 on the matrix cores, FP32 master
    weights): in_dim -> hidden x num_hidden -> out_dim, one output per
    dye-seq of the HMM classifier it serves."""

    def __init__(self, in_dim, hidden, num_hidden, out_dim, seed=1):
        h = ctypes.c_void_p()
        check(lib().wp_mlp_create(int(in_dim), int(hidden), int(num_hidden), int(out_dim), int(seed),
                                  ctypes.byref(h)))
        self._h = h
        self.in_dim, self.hidden, self.num_hidden, self.out_dim = in_dim, hidden, num_hidden, out_dim
        self.num_layers = num_hidden + 1

    def close(self):
        if getattr(self, "_h", None):
            lib().wp_mlp_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def layer_shape(self, l):
        r, c = ctypes.c_int(), ctypes.c_int()
        check(lib().wp_mlp_layer_shape(self._h, l, ctypes.byref(r), ctypes.byref(c)))
        return r.value, c.value

    def get_layers(self):
        out = []
        for l in range(self.num_layers):
            r, c = self.layer_shape(l)
            W = np.zeros((r, c), dtype=np.float32)
            b = np.zeros(c, dtype=np.float32)
            check(lib().wp_mlp_get_layer(self._h, l, W.ctypes.data, b.ctypes.data))
            out.append((W, b))
        return out

    def set_layers(self, layers):
        for l, (W, b) in enumerate(layers):
            W = np.ascontiguousarray(W, dtype=np.float32)
            b = np.ascontiguousarray(b, dtype=np.float32)
            if W.shape != self.layer_shape(l) or b.shape != (W.shape[1],):
                raise WhatprotError(f"layer {l} must be {self.layer_shape(l)}")
            check(lib().wp_mlp_set_layer(self._h, l, W.ctypes.data, b.ctypes.data))

    def logits(self, x):
        x = np.ascontiguousarray(x, dtype=np.float32).reshape(-1, self.in_dim)
        out = np.zeros((x.shape[0], self.out_dim), dtype=np.float32)
        check(lib().wp_mlp_logits(self._h, x.shape[0], x.ctypes.data, out.ctypes.data))
        return out

    def gradients(self, x, labels):
        """(mean loss, [(dW, db)...]) without an update."""
        x = np.ascontiguousarray(x, dtype=np.float32).reshape(-1, self.in_dim)
        y = np.ascontiguousarray(labels, dtype=np.int32)
        gW, gb = [], []
        for l in range(self.num_layers):
            r, c = self.layer_shape(l)
            gW.append(np.zeros((r, c), dtype=np.float32))
            gb.append(np.zeros(c, dtype=np.float32))
        pw = (ctypes.c_void_p * self.num_layers)(*[a.ctypes.data for a in gW])
        pb = (ctypes.c_void_p * self.num_layers)(*[a.ctypes.data for a in gb])
        loss = ctypes.c_double()
        check(lib().wp_mlp_gradients(self._h, x.shape[0], x.ctypes.data, y.ctypes.data, ctypes.byref(loss), pw, pb))
        return loss.value, list(zip(gW, gb))

    def train(self, x, labels, batch=4096, epochs=1, lr=1e-3, seed=7):
        """Adam over `epochs` shuffled passes; returns the mean loss per epoch."""
        x = np.ascontiguousarray(x, dtype=np.float32).reshape(-1, self.in_dim)
        y = np.ascontiguousarray(labels, dtype=np.int32)
        losses = np.zeros(max(epochs, 1), dtype=np.float32)
        check(lib().wp_mlp_train(self._h, x.shape[0], x.ctypes.data, y.ctypes.data, int(batch), int(epochs),
                                 float(lr), int(seed), losses.ctypes.data))
        return losses[:epochs]

    def shortlist(self, radiometries, H, with_mass=False):
        """(indices [N, H] in increasing order, max posterior [N][, shortlist mass [N]])."""
        r = np.ascontiguousarray(radiometries, dtype=np.float64).reshape(-1, self.in_dim)
        idx = np.zeros((r.shape[0], H), dtype=np.int32)
        post = np.zeros(r.shape[0], dtype=np.float32)
        mass = np.zeros(r.shape[0], dtype=np.float32)
        check(lib().wp_mlp_shortlist(self._h, r.shape[0], r.ctypes.data, int(H), idx.ctypes.data, post.ctypes.data,
                                     mass.ctypes.data))
        return (idx, post, mass) if with_mass else (idx, post)

    def classify(self, classifier: "HMMClassifier", radiometries, H, min_posterior=0.5, criterion="max"):
        """wp_shortlist_classify: (ids, adjusted, number of full-HMM fallbacks).
        criterion "max" trusts a shortlist when the largest posterior reaches
        min_posterior, "mass" when the shortlist's posterior mass does."""
        r = np.ascontiguousarray(radiometries, dtype=np.float64)
        N = r.shape[0]
        ids = np.zeros(N, dtype=np.int32)
        adj = np.zeros(N)
        nf = ctypes.c_int()
        crit = {"max": 0, "mass": 1}[criterion]
        check(lib().wp_shortlist_classify(classifier._h, self._h, N, r.ctypes.data, int(H), float(min_posterior),
                                          crit, ids.ctypes.data, adj.ctypes.data, ctypes.byref(nf)))
        return ids, adj, nf.value

    def classify_device(self, classifier: "HMMClassifier", num_reads, d_rads, H, d_id, d_adjusted,
                        min_posterior=0.5, criterion="max", stream=0):
        """wp_shortlist_classify_device on device pointers (radiometries
        [N][T][C] FP64 in, ids int32 and adjusted FP64 out); returns the
        number of full-HMM fallbacks."""
        nf = ctypes.c_int()
        crit = {"max": 0, "mass": 1}[criterion]
        check(lib().wp_shortlist_classify_device(classifier._h, self._h, int(num_reads), d_rads, int(H),
                                                 float(min_posterior), crit, d_id, d_adjusted, ctypes.byref(nf),
                                                 stream))
        return nf.value
