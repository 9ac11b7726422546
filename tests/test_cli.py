"""Command line (whatprot-amd): usage rules and the host-only simulate modes."""
import os
import subprocess

import pytest

import whatprot_amd as w
from tests._files import write_seqparams

CLI = os.path.join(os.path.dirname(w.__file__), "whatprot-amd")


def run(*args):
    return subprocess.run([CLI, *args], capture_output=True, text=True, timeout=120)


@pytest.fixture(scope="module", autouse=True)
def built():
    if not os.path.exists(CLI):
        pytest.skip("whatprot-amd not built")


@pytest.mark.parametrize("args", [[], ["classify"], ["classify", "hmm", "-P", "x.json"],
                                  ["classify", "nn", "-P", "p", "-k", "3", "-T", "t", "-R", "r", "-Y", "y"],
                                  ["fit", "-P", "p", "-x", "..0", "-R", "r"],
                                  ["classify", "hmm", "-P", "p", "-S", "s", "-R", "r", "-Y", "y", "-k", "4"]])
def test_incorrect_usage(args):
    r = run(*args)
    assert r.returncode == 1 and "INCORRECT USAGE" in r.stdout


def test_bootstrap_refused():
    r = run("fit", "-P", "p", "-x", "..0", "-L", "0.01", "-R", "r", "-b", "10", "-c", "0.9")
    assert r.returncode == 1 and "bootstrap" in r.stderr


def test_simulate_rad_and_dt(tmp_path):
    sp = tmp_path / "sp.json"
    write_seqparams(sp, 2)
    ds = tmp_path / "ds.tsv"
    w.write_dye_seqs(str(ds), 2, [w.DyeSeqEntry("..0.1..0", 2, 7), w.DyeSeqEntry(".1.1", 1, 9)])
    rad, ys, dt = tmp_path / "r.tsv", tmp_path / "y.tsv", tmp_path / "dt.tsv"
    r = run("simulate", "rad", "-P", str(sp), "-t", "6", "-g", "50", "-S", str(ds), "-R", str(rad), "-Y", str(ys),
            "--seed", "3")
    assert r.returncode == 0, r.stderr
    model = w.SequencingModel.from_json(str(sp))
    _, _, rads = w.read_radiometries(str(rad), model)
    assert rads.shape == (50, 6, 2)
    ids = [int(x) for x in ys.read_text().split()]
    assert ids[0] == 50 and set(ids[1:]) <= {7, 9}
    r = run("simulate", "dt", "-P", str(sp), "-t", "6", "-g", "20", "-S", str(ds), "-T", str(dt), "--seed", "3")
    assert r.returncode == 0, r.stderr
    tr = w.read_dye_tracks(str(dt))
    assert tr.num_timesteps == 6 and tr.num_channels == 2 and len(tr) > 3
    assert set(tr.ids.tolist()) <= {7, 9}
