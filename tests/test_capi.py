"""The C-ABI library loads and exports every entry point include/*.h
declares; host-only helpers behave. No GPU compute here."""
import ctypes
import os
import re

import numpy as np

import whatprot_amd as w
from whatprot_amd import _capi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    names = set()
    inc = os.path.join(ROOT, "include")
    for fn in os.listdir(inc):
        if not fn.endswith(".h"):
            continue
        text = open(os.path.join(inc, fn)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"\b(wp_[a-z0-9_]+)\s*\(", text):
            names.add(m.group(1))
    return names


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(_capi.LIB_PATH)
    names = declared_symbols()
    assert len(names) >= 10
    missing = [n for n in sorted(names) if not hasattr(lib, n)]
    assert not missing, missing


def test_with_mu_as_one_matches_python_mirror():
    m = w.SequencingModel.uniform(3, mu=10000.0, sig=1600.0, bg_sig=66.7)
    a = m.to_c()
    b = _capi.ModelC()
    w.lib().wp_model_with_mu_as_one(ctypes.byref(a), ctypes.byref(b))
    py = m.with_mu_as_one()
    for c in range(3):
        assert b.channels[c].mu == 1.0
        assert b.channels[c].sig == py.channel_models[c].sig == 0.16
        assert b.channels[c].bg_sig == py.channel_models[c].bg_sig


def test_seqparams_json_parse():
    d = {"p_edman_failure": 0.06, "p_detach": 0.05, "p_initial_detach": 0.04,
         "p_initial_detach_decay": 0.3, "p_initial_block": 0.07, "p_cyclic_block": 0.02,
         "channel_models": [{"p_bleach": 0.05, "p_dud": 0.07, "bg_sig": 66.7, "mu": 10000, "sig": 1600,
                             "interactions": [{"other_channel": 1, "effect": 0.25}]},
                            {"p_bleach": 0.04, "p_dud": 0.08, "bg_sig": 66.7, "mu": 10000, "sig": 1600}]}
    m = w.SequencingModel.from_json(d)
    assert m.num_channels == 2
    assert m.channel_models[0].interactions == [1.0, 0.75]
    assert m.channel_models[0].flat_interactions == [1.0, 1.0]


def test_generators_are_deterministic():
    a = w.generate_dye_seqs(3, 50, 6, 30, 0, 5, 1001)
    b = w.generate_dye_seqs(3, 50, 6, 30, 0, 5, 1001)
    assert a == b
    for s in a:
        assert 6 <= len(s) <= 30
        assert any(ch != "." for ch in s)
    m = w.SequencingModel.uniform(3)
    r1, i1 = w.simulate_radiometries(m, 12, a, 200, 7)
    r2, i2 = w.simulate_radiometries(m, 12, a, 200, 7)
    assert np.array_equal(r1, r2) and np.array_equal(i1, i2)
    assert (r1 == r1.astype(np.float32)).all()


def test_classifier_create_rejects_bad_shapes():
    m = w.SequencingModel.uniform(5)
    try:
        w.HMMClassifier(12, 5, m, w.SequencingSettings(), [w.DyeSeqEntry("01234")])
        raised = False
    except w.WhatprotError:
        raised = True
    assert raised
