"""Loads the golden fixtures written by tests/golden/make_golden.py."""
import glob
import os

import problems as P

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def names():
    return sorted(os.path.basename(p)[:-5] for p in glob.glob(os.path.join(GOLDEN, "*.json")))


def load(name):
    return P.load_case(os.path.join(GOLDEN, name))
