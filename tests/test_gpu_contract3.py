"""btod_contract3 / btod_symcontract3 on the device through the C ABI:
D = kd (A B) C (AB symmetrized first for btod_symcontract3) for every
fixture c3* against libtensor's result, with D := result in both launch
modes, D += k result, repeated performs (the intermediate and both plans
reused), and a symmetrization of the result (btod_symmetrize2 over
btod_contract3) against the closed form D + s P(D)."""
import itertools
import sys

import numpy as np
import pytest

import golden_cases as G
import libtensor_amd as lt
import problems as P

pytestmark = pytest.mark.gpu

TOL = 1e-13


def _say(*a):
    print(*a)
    sys.stdout.flush()


@pytest.mark.parametrize("mode", [lt.LT_MODE_GROUPED, lt.LT_MODE_PER_BLOCK])
@pytest.mark.parametrize("name", G.c3_names())
def test_perform(gpu_ctx, name, mode):
    _say("btod_contract3", name, "mode", mode)
    prob, dense, meta, result = G.load_c3(name)
    ts, op = P.build3(gpu_ctx, prob, dense)
    d = P.result_tensor3(gpu_ctx, prob)
    scale = float(np.max(np.abs(result))) or 1.0
    op.perform(d, mode)
    assert np.max(np.abs(d.to_dense() - result)) <= TOL * scale
    assert set(d.req_nonzero_blocks()) <= set(op.get_schedule())
    op.perform_add(d, 2.0, mode)
    assert np.max(np.abs(d.to_dense() - 3.0 * result)) <= TOL * scale
    op.perform(d, mode)
    assert np.max(np.abs(d.to_dense() - result)) <= TOL * scale
    st = op.stats()
    if meta["schedule"]:
        # AB, written again by the first contraction, keeps its structure:
        # the second contraction's plan for D is reused
        n = op.plans()
        op.perform(d, mode)
        assert op.plans() == n
    assert st["flops"] > 0 or not meta["schedule"]


@pytest.mark.parametrize("name", G.c3_names())
def test_symmetrize_the_result(gpu_ctx, name):
    prob, dense, meta, result = G.load_c3(name)
    ts, op = P.build3(gpu_ctx, prob, dense)
    dd = prob["D"]
    n = len(dd["dims"])

    def splits(i):
        return sorted(p for m, p in dd["splits"] if m[i])
    pairs = [(i, j) for i, j in itertools.combinations(range(n), 2)
             if dd["dims"][i] == dd["dims"][j] and splits(i) == splits(j)]
    if not pairs:
        pytest.skip("no two result indexes of the same block structure")
    i, j = pairs[0]
    perm = list(range(n))
    perm[i], perm[j] = j, i
    scale = float(np.max(np.abs(result))) or 1.0
    ran = 0
    for symm in (True, False):
        try:
            s = lt.btod_symmetrize2(op, perm, symm)
        except lt.LtError:
            continue      # the sign contradicts the result's symmetry
        d = P.result_tensor3(gpu_ctx, prob)
        s.perform(d)
        exp = result + (1.0 if symm else -1.0) * np.transpose(result, perm)
        assert np.max(np.abs(d.to_dense() - exp)) <= TOL * scale, symm
        ran += 1
    if ran == 0:
        pytest.skip("the pair's dimensions carry different labels: no symmetrization")
