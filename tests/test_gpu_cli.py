"""Command line end to end on the GPU: files in the reference's formats in,
results CSV out, against the reference's classifiers and fitter compiled from
its sources (the model read with mu = 10000, radiometries divided by it)."""
import os
import re
import subprocess

import numpy as np
import pytest

import whatprot_amd as w
from tests._files import read_results, write_radiometries_scaled, write_seqparams

pytestmark = pytest.mark.gpu
CLI = os.path.join(os.path.dirname(w.__file__), "whatprot-amd")
MU = 10000.0


def run(*args):
    r = subprocess.run([CLI, *[str(a) for a in args]], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    return r.stdout


@pytest.fixture(scope="module")
def files(tmp_path_factory):
    d = tmp_path_factory.mktemp("cli")
    C, T = 2, 8
    sp = d / "sp.json"
    write_seqparams(sp, C, MU)
    truth = w.SequencingModel.from_json(str(sp))
    model = truth.with_mu_as_one()
    strings = w.generate_dye_seqs(C, 40, 6, 12, 0, 3, 71)
    seqs = [w.DyeSeqEntry(s, 1 + i % 2, 500 + i) for i, s in enumerate(strings)]
    ds = d / "ds.tsv"
    w.write_dye_seqs(str(ds), C, seqs)
    rads, _ = w.simulate_radiometries(model, T, strings, 120, 72)
    rp = d / "r.tsv"
    write_radiometries_scaled(rp, rads, MU)
    _, _, back = w.read_radiometries(str(rp), truth)  # what the reference's reader yields
    tracks = w.simulate_dye_tracks(model, T, seqs, 25, 73)
    tp = d / "dt.tsv"
    w.write_dye_tracks(str(tp), tracks)
    return dict(d=d, sp=sp, ds=ds, rp=rp, tp=tp, model=model, seqs=seqs, rads=back, tracks=tracks, T=T, C=C)


def test_classify_hmm(gpu, ref, files):
    out = files["d"] / "hmm.csv"
    run("classify", "hmm", "-P", files["sp"], "-S", files["ds"], "-R", files["rp"], "-Y", out, "-p", "5")
    ids, scores = read_results(out)
    seqs = files["seqs"]
    rid, rs, rt = ref.classify(files["model"], [s.dye_string for s in seqs], [s.count for s in seqs],
                               [s.id for s in seqs], files["rads"], 5.0)
    np.testing.assert_array_equal(ids, rid)
    np.testing.assert_allclose(scores, rs / rt, rtol=1e-12)


def test_classify_nn(gpu, ref, files):
    out = files["d"] / "nn.csv"
    run("classify", "nn", "-P", files["sp"], "-k", 10, "-s", 0.5, "-T", files["tp"], "-R", files["rp"], "-Y", out)
    ids, scores = read_results(out)
    rid, rs, rt = ref.nn_classify(files["model"], files["tracks"], 10, 0.5, files["rads"])
    np.testing.assert_array_equal(ids, rid)
    np.testing.assert_allclose(scores, rs / rt, rtol=1e-15)


def test_classify_hybrid(gpu, ref, files):
    out = files["d"] / "hyb.csv"
    run("classify", "hybrid", "-P", files["sp"], "-k", 10, "-s", 0.5, "-H", 8, "-S", files["ds"], "-T", files["tp"],
        "-R", files["rp"], "-Y", out, "-p", 5)
    ids, scores = read_results(out)
    seqs = files["seqs"]
    rid, rs, rt = ref.hybrid_classify(files["model"], files["tracks"], 10, 0.5, 8, [s.dye_string for s in seqs],
                                      [s.count for s in seqs], [s.id for s in seqs], files["rads"], 5.0)
    np.testing.assert_array_equal(ids, rid)
    np.testing.assert_allclose(scores, rs / rt, rtol=1e-12)


def test_fit(gpu, ref, files):
    d, C, T = files["d"], files["C"], 10
    ds = ".0.1.1.0"
    model = files["model"]
    rads, _ = w.simulate_radiometries(model, T, [ds], 80, 74)
    rp = d / "fit_r.tsv"
    write_radiometries_scaled(rp, rads, MU)
    _, _, back = w.read_radiometries(str(rp), w.SequencingModel.from_json(str(files["sp"])))
    outp = run("fit", "-P", files["sp"], "-x", ds, "-L", 1e-3, "-R", rp)
    log_l = float(re.search(r"log\(L\): (\S+)", outp).group(1))
    _, _, want = ref.fit(model, ds, back, 1e-3)
    assert log_l == pytest.approx(want, rel=1e-8)
    assert "Edman failure rate: " in outp and "Channel 1 info: (Bleach rate: " in outp


def test_fit_bootstrap(gpu, files):
    """fit -b -c -Y: the full fit as without -b, the bounds printed and one
    write_params row per round (params-io.cc:37-54)."""
    d, C, T = files["d"], files["C"], 10
    ds = ".0.1.1.0"
    rads, _ = w.simulate_radiometries(files["model"], T, [ds], 80, 75)
    rp, yp = d / "boot_r.tsv", d / "boot.csv"
    write_radiometries_scaled(rp, rads, MU)
    plain = run("fit", "-P", files["sp"], "-x", ds, "-L", 1e-3, "-R", rp)
    outp = run("fit", "-P", files["sp"], "-x", ds, "-L", 1e-3, "-R", rp, "-b", 5, "-c", 0.8, "-Y", yp, "--seed", 4)
    assert re.search(r"log\(L\): (\S+)", outp).group(1) == re.search(r"log\(L\): (\S+)", plain).group(1)
    assert "lower bounds: Edman failure rate: " in outp and "upper bounds: " in outp
    assert "worst bootstrap step size: " in outp
    lines = yp.read_text().splitlines()
    assert lines[0] == ("p_edman_failure,p_detach,p_initial_detach,p_initial_detach_decay,p_initial_block,"
                        "p_cyclic_block,ch0:p_bleach,ch0:p_dud,ch1:p_bleach,ch1:p_dud,log(L)")
    assert len(lines) == 6 and all(len(x.split(",")) == 11 for x in lines[1:])


def test_simulate_rad_and_dt(gpu, tmp_path):
    sp = tmp_path / "sp.json"
    write_seqparams(sp, 2)
    ds = tmp_path / "ds.tsv"
    w.write_dye_seqs(str(ds), 2, [w.DyeSeqEntry("..0.1..0", 2, 7), w.DyeSeqEntry(".1.1", 1, 9)])
    rad, ys, dt = tmp_path / "r.tsv", tmp_path / "y.tsv", tmp_path / "dt.tsv"
    run("simulate", "rad", "-P", str(sp), "-t", "6", "-g", "50", "-S", str(ds), "-R", str(rad), "-Y", str(ys),
        "--seed", "3")
    model = w.SequencingModel.from_json(str(sp))
    _, _, rads = w.read_radiometries(str(rad), model)
    assert rads.shape == (50, 6, 2)
    ids = [int(x) for x in ys.read_text().split()]
    assert ids[0] == 50 and set(ids[1:]) <= {7, 9}
    run("simulate", "dt", "-P", str(sp), "-t", "6", "-g", "20", "-S", str(ds), "-T", str(dt), "--seed", "3")
    tr = w.read_dye_tracks(str(dt))
    assert tr.num_timesteps == 6 and tr.num_channels == 2 and len(tr) > 3
    assert set(tr.ids.tolist()) <= {7, 9}
    # the GPU simulator behind `simulate dt` reproduces the host's streams
    ref = w.simulate_dye_tracks(model, 6, [w.DyeSeqEntry("..0.1..0", 2, 7), w.DyeSeqEntry(".1.1", 1, 9)], 20, 3)
    assert (tr.counts == ref.counts).all() and (tr.hits == ref.hits).all() and (tr.ids == ref.ids).all()
