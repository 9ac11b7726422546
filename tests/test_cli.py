"""Command line (whatprot-amd): usage rules (no GPU needed)."""
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


@pytest.mark.parametrize("extra", [["-b", "10"], ["-c", "0.9"], ["-Y", "out.csv"]])
def test_bootstrap_flags_need_each_other(extra):
    """-b and -c come together, and -Y only with them (main.cc:413-446)."""
    r = run("fit", "-P", "p", "-x", "..0", "-L", "0.01", "-R", "r", *extra)
    assert r.returncode == 1 and "INCORRECT USAGE" in r.stdout
