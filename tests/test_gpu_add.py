"""perform(btc, d) into a target of another symmetry, and btod_add of
operands with different symmetries, on the GPU through the C ABI, against
libtensor's results (fixtures add_* / badd_* written by the real reference
build): max|dX| / max|X| <= 1e-12, and the target's symmetry after the call
(canonical block and allowedness of every block) equal to libtensor's."""
import sys

import numpy as np
import pytest

import dense_oracle as O
import golden_cases as G
import libtensor_amd as lt
import problems as P
from libtensor_amd import workloads as W

pytestmark = pytest.mark.gpu

TOL = 1e-12


def _say(*a):
    print(*a)
    sys.stdout.flush()


def _canonical_match(t, symt):
    for aidx, (ac, al, perm, c) in symt.items():
        got = t.orbit(t.index(int(aidx)))
        if (got[0], got[3]) != (ac, bool(al)):
            return int(aidx)
    return None


@pytest.mark.parametrize("mode", [lt.LT_MODE_GROUPED, lt.LT_MODE_PER_BLOCK])
@pytest.mark.parametrize("name", G.add_names())
def test_perform_add_into_target(gpu_ctx, name, mode):
    _say("perform_add", name, mode)
    prob, dense, meta, result = G.load_add(name)
    a, b, op = P.build(gpu_ctx, prob, dense)
    t = W.build_tensor(gpu_ctx, prob["C"])
    t.from_dense(dense["C"])
    op.perform_add(t, prob["add"], mode)
    assert O.rel_err(t.to_dense(), result) <= TOL
    assert _canonical_match(t, meta["symt_add"]) is None
    # a second += reuses the plan and the merged target
    op.perform_add(t, prob["add"], mode)
    r = O.btod_contract2_dense(prob["contr"], dense["A"], dense["B"])
    assert O.rel_err(t.to_dense(), result + prob["add"] * r) <= TOL


@pytest.mark.parametrize("name", G.badd_names())
def test_btod_add_across_symmetries(gpu_ctx, name):
    _say("btod_add", name)
    prob, dense, meta, result = G.load_add(name)
    ba = prob["btod_add"]
    W.register_tables(prob)
    x = W.build_tensor(gpu_ctx, prob["A"])
    x.from_dense(dense["A"])
    y = W.build_tensor(gpu_ctx, prob["B"])
    y.from_dense(dense["B"])
    out = W.build_tensor(gpu_ctx, dict(prob["A"], elems=[]))
    lt.btod_add([(x, None, ba["ca"]), (y, ba["perm"], ba["cb"])], out)
    exp = ba["ca"] * dense["A"] + ba["cb"] * np.transpose(dense["B"], ba["perm"])
    assert O.rel_err(out.to_dense(), exp) <= TOL
    if name != "badd_label01_nopart":
        # libtensor's own result (its schedule misses blocks in that one
        # case, see tests/test_add_symmetry.py)
        assert O.rel_err(out.to_dense(), result) <= TOL
    assert _canonical_match(out, meta["symt"]) is None
    # += into the result: 2 (ca A + cb P(B))
    lt.btod_add([(x, None, ba["ca"]), (y, ba["perm"], ba["cb"])], out, add=True, d=1.0)
    assert O.rel_err(out.to_dense(), 2.0 * exp) <= TOL
