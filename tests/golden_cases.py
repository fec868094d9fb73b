"""Loads the golden fixtures written by tests/golden/make_golden.py."""
import glob
import json
import os

import numpy as np

import problems as P

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def names():
    """Contraction cases (the btod_symmetrize2 fixtures symm_* excluded)."""
    return sorted(os.path.basename(p)[:-5] for p in glob.glob(os.path.join(GOLDEN, "*.json"))
                  if not os.path.basename(p).startswith("symm_"))


def symm_names(errors=False):
    """btod_symmetrize2 fixtures: those with a result, or (errors) those
    libtensor rejects."""
    out = []
    for p in sorted(glob.glob(os.path.join(GOLDEN, "symm_*.json"))):
        with open(p) as f:
            meta = json.load(f)
        if bool(meta.get("error")) == errors:
            out.append(os.path.basename(p)[:-5])
    return out


def load_symm(name):
    """(base problem, dense inputs, base golden result, fixture, libtensor's
    symmetrized result or None)."""
    with open(os.path.join(GOLDEN, name + ".json")) as f:
        meta = json.load(f)
    prob, dense, bmeta, bresult = load(meta["base"])
    result = None
    path = os.path.join(GOLDEN, name + ".npz")
    if os.path.exists(path):
        result = np.load(path, allow_pickle=False)["result"]
    return prob, dense, bresult, meta, result


def load(name):
    return P.load_case(os.path.join(GOLDEN, name))
