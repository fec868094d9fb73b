"""Loads the golden fixtures written by tests/golden/make_golden.py."""
import glob
import json
import os

import numpy as np

import problems as P

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def names():
    """btod_contract2 cases (the symmetrization fixtures symm* and the
    btod_contract3 fixtures c3* excluded)."""
    return sorted(os.path.basename(p)[:-5] for p in glob.glob(os.path.join(GOLDEN, "*.json"))
                  if not os.path.basename(p).startswith(("symm", "c3", "add_", "badd_")))


def c3_names():
    return sorted(os.path.basename(p)[:-5] for p in glob.glob(os.path.join(GOLDEN, "c3*.json")))


def load_c3(name):
    """(problem, dense inputs, fixture, libtensor's result)."""
    with open(os.path.join(GOLDEN, name + ".json")) as f:
        meta = json.load(f)
    z = np.load(os.path.join(GOLDEN, name + ".npz"), allow_pickle=False)
    return meta["prob"], {k: z[k] for k in ("A", "B", "C")}, meta, z["result"]


def symm_names(errors=False):
    """btod_symmetrize2 / btod_symmetrize3 fixtures (symm_*, symm3_*): those
    with a result, or (errors) those libtensor rejects."""
    out = []
    for p in sorted(glob.glob(os.path.join(GOLDEN, "symm*.json"))):
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
    path = os.path.join(GOLDEN, name + ".npz")
    z = np.load(path, allow_pickle=False) if os.path.exists(path) else None
    if "base" in meta:
        prob, dense, bmeta, bresult = load(meta["base"])
    else:
        # a problem of its own: inputs stored with the fixture, no base result
        prob, bresult = meta["prob"], None
        dense = {k: z[k] for k in ("A", "B") if z is not None and k in z.files}
    result = z["result"] if z is not None else None
    return prob, dense, bresult, meta, result


def add_names():
    """perform(btc, d) into a target of another symmetry (add_*)."""
    return sorted(os.path.basename(p)[:-5] for p in glob.glob(os.path.join(GOLDEN, "add_*.json")))


def badd_names():
    """btod_add of operands with different symmetries (badd_*)."""
    return sorted(os.path.basename(p)[:-5] for p in glob.glob(os.path.join(GOLDEN, "badd_*.json")))


def load_add(name):
    """(problem with target "C" and "add", dense A, B, C, fixture, libtensor's
    target after perform(btc, d) or btod_add's result)."""
    with open(os.path.join(GOLDEN, name + ".json")) as f:
        meta = json.load(f)
    z = np.load(os.path.join(GOLDEN, name + ".npz"), allow_pickle=False)
    dense = {k: z[k] for k in ("A", "B", "C") if k in z.files}
    return meta["prob"], dense, meta, z["result"]


def load(name):
    return P.load_case(os.path.join(GOLDEN, name))
