"""btod_copy on resident block tensors through the C ABI: B = c P(A) and
B += d c P(A) on the inputs of every golden case (every symmetry kind:
se_perm with signs, se_part, se_label), for several permutations, against
numpy transposes of the dense inputs; B takes the permuted symmetry of A
(its dense expansion must equal the permuted dense A)."""
import sys

import numpy as np
import pytest

import golden_cases as G
import libtensor_amd as lt
import problems as P

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a)
    sys.stdout.flush()


def permuted_tensor(ctx, td, perm):
    bis = lt.block_index_space([td["dims"][p] for p in perm])
    for mask, pos in td.get("splits", []):
        bis.split([mask[p] for p in perm], pos)
    return lt.block_tensor(ctx, bis)


@pytest.mark.parametrize("name", G.names())
def test_btod_copy(gpu_ctx, name):
    _say("btod_copy", name)
    prob, dense, meta, result = G.load(name)
    a, b, op = P.build(gpu_ctx, prob, dense)
    A = dense["A"]
    n = A.ndim
    rng = np.random.default_rng(11)
    perms = [tuple(range(n)), tuple(range(n))[::-1]] + \
        [tuple(int(x) for x in rng.permutation(n)) for _ in range(2)]
    scale = float(np.max(np.abs(A))) or 1.0
    for perm in perms:
        bt = permuted_tensor(gpu_ctx, prob["A"], perm)
        lt.btod_copy(a, bt, perm, c=0.75)
        exp = 0.75 * np.transpose(A, perm)
        assert np.max(np.abs(bt.to_dense() - exp)) <= 1e-15 * scale, perm
        lt.btod_copy(a, bt, perm, c=0.75, add=True, d=2.0)
        assert np.max(np.abs(bt.to_dense() - 3.0 * exp)) <= 1e-14 * scale, perm


def test_btod_copy_add_needs_the_result_symmetry(gpu_ctx):
    prob, dense, meta, result = G.load("ladder_full")
    a, b, op = P.build(gpu_ctx, prob, dense)
    bt = permuted_tensor(gpu_ctx, prob["A"], tuple(range(dense["A"].ndim)))
    with pytest.raises(lt.LtError):
        lt.btod_copy(a, bt, None, add=True)


@pytest.mark.parametrize("name", G.names())
def test_btod_add(gpu_ctx, name):
    """B = 0.5 A + 2 P(A) for P the identity and for a permutation that maps A's
    symmetry onto itself (a transposition of two equal-extent indices when A
    has one, else the identity again), then B += -1 (A - A)."""
    _say("btod_add", name)
    prob, dense, meta, result = G.load(name)
    a, b, op = P.build(gpu_ctx, prob, dense)
    A = dense["A"]
    n = A.ndim
    scale = float(np.max(np.abs(A))) or 1.0
    ident = tuple(range(n))
    perm = ident
    for el in prob["A"].get("elems", []):
        if el[0] == "perm":
            perm = tuple(int(x) for x in el[1])
            break
    bt = permuted_tensor(gpu_ctx, prob["A"], ident)
    lt.btod_add([(a, None, 0.5), (a, perm, 2.0)], bt)
    exp = 0.5 * A + 2.0 * np.transpose(A, perm)
    assert np.max(np.abs(bt.to_dense() - exp)) <= 1e-14 * scale
    lt.btod_add([(a, None, 1.0), (a, None, -1.0)], bt, add=True, d=-1.0)
    assert np.max(np.abs(bt.to_dense() - exp)) <= 1e-14 * scale


def test_btod_add_rejects_different_symmetries(gpu_ctx):
    prob, dense, meta, result = G.load("ladder_full")
    a, b, op = P.build(gpu_ctx, prob, dense)
    plain = permuted_tensor(gpu_ctx, prob["A"], tuple(range(dense["A"].ndim)))
    lt.btod_copy(a, plain)              # same data, A's symmetry
    bare = permuted_tensor(gpu_ctx, prob["A"], tuple(range(dense["A"].ndim)))
    bare.from_dense(dense["A"])         # same data, no symmetry
    out = permuted_tensor(gpu_ctx, prob["A"], tuple(range(dense["A"].ndim)))
    lt.btod_add([(a, None, 1.0), (plain, None, 1.0)], out)
    with pytest.raises(lt.LtError):
        lt.btod_add([(a, None, 1.0), (bare, None, 1.0)], out)
