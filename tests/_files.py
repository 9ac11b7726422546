"""Writers of the reference's input files for the command-line tests."""
import json

import numpy as np


def write_seqparams(path, C, mu=10000.0):
    """A seqparams JSON in the schema of sequencing-model.cc:51-89 with the
    atto647n-style priors, mu != 1 so the radiometry reader's division shows."""
    data = {"p_edman_failure": 0.06, "p_detach": 0.05, "p_initial_detach": 0.04, "p_initial_detach_decay": 0.3,
            "p_initial_block": 0.07, "p_cyclic_block": 0.02,
            "channel_models": [{"p_bleach": 0.05, "p_dud": 0.07, "bg_sig": 0.00667 * mu, "mu": mu,
                                "sig": 0.16 * mu, "interactions": []} for _ in range(C)]}
    with open(path, "w") as f:
        json.dump(data, f)


def write_radiometries_scaled(path, rads, mu):
    """Radiometries file (radiometries-io.cc:64-81) of intensities times mu."""
    N, T, C = rads.shape
    with open(path, "w") as f:
        f.write(f"{T}\n{C}\n{N}\n")
        for r in rads:
            f.write("\t".join(repr(float(x) * mu) for x in r.ravel()) + "\n")


def read_results(path):
    """(ids, scores) of a results CSV (scored-classifications-io.cc:54-68)."""
    with open(path) as f:
        lines = f.read().strip().split("\n")
    assert lines[0] == "radmat_iz,best_pep_iz,best_pep_score"
    rows = [l.split(",") for l in lines[1:]]
    assert [int(r[0]) for r in rows] == list(range(len(rows)))
    return np.array([int(r[1]) for r in rows]), np.array([float(r[2]) for r in rows])
