"""btod_copy on resident block tensors through the C ABI: B = c P(A) and
B += d c P(A) on the inputs of every golden case (every symmetry kind:
se_perm with signs, se_part, se_label), for several permutations, against
numpy transposes of the dense inputs; B takes the permuted symmetry of A
(its dense expansion must equal the permuted dense A).  btod_add and += into
tensors of other symmetries take the symmetry all terms have."""
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


def test_btod_copy_add_into_other_symmetry(gpu_ctx):
    """B += P(A) into a B without symmetry (libtensor's gen_bto_aux_add): B
    keeps no symmetry and holds A's full expansion."""
    prob, dense, meta, result = G.load("ladder_full")
    a, b, op = P.build(gpu_ctx, prob, dense)
    bt = permuted_tensor(gpu_ctx, prob["A"], tuple(range(dense["A"].ndim)))
    lt.btod_copy(a, bt, None, add=True, d=0.5)
    scale = float(np.max(np.abs(dense["A"])))
    assert np.max(np.abs(bt.to_dense() - 0.5 * dense["A"])) <= 1e-15 * scale
    assert len(bt.req_nonzero_blocks()) >= len(a.req_nonzero_blocks())


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


def test_btod_add_different_symmetries(gpu_ctx):
    """Operands of different symmetries (libtensor: so_dirsum + so_merge,
    gen_bto_add_impl.h:238-267): the sum takes the symmetry both have."""
    prob, dense, meta, result = G.load("ladder_full")
    a, b, op = P.build(gpu_ctx, prob, dense)
    plain = permuted_tensor(gpu_ctx, prob["A"], tuple(range(dense["A"].ndim)))
    lt.btod_copy(a, plain)              # same data, A's symmetry
    bare = permuted_tensor(gpu_ctx, prob["A"], tuple(range(dense["A"].ndim)))
    bare.from_dense(dense["A"])         # same data, no symmetry
    out = permuted_tensor(gpu_ctx, prob["A"], tuple(range(dense["A"].ndim)))
    scale = float(np.max(np.abs(dense["A"])))
    lt.btod_add([(a, None, 1.0), (plain, None, 1.0)], out)
    assert np.max(np.abs(out.to_dense() - 2.0 * dense["A"])) <= 1e-15 * scale
    lt.btod_add([(a, None, 1.0), (bare, None, -3.0)], out)
    assert np.max(np.abs(out.to_dense() + 2.0 * dense["A"])) <= 1e-15 * scale
    # every block of the result is its own orbit (bare has no symmetry)
    nblk = int(np.prod(out.bidims))
    assert all(out.orbit(out.index(x))[0] == x for x in range(nblk))


def test_repeated_calls_reuse_descriptors_and_follow_changes(gpu_ctx):
    """btod_copy / btod_add keep their uploaded descriptor lists per operand
    structure version: repeated calls give the same result, a change of the
    source's non-zero blocks or of B's blocks (for +=) gives a fresh list."""
    prob, dense, meta, result = G.load("ladder_full")
    a, b, op = P.build(gpu_ctx, prob, dense)
    A = dense["A"]
    n = A.ndim
    bt = permuted_tensor(gpu_ctx, prob["A"], tuple(range(n)))
    for _ in range(3):
        lt.btod_copy(a, bt, None, c=2.0)
        assert np.array_equal(bt.to_dense(), 2.0 * A)
    for k in range(3):
        lt.btod_copy(a, bt, None, c=1.0, add=True, d=1.0)
        assert np.allclose(bt.to_dense(), (3.0 + k) * A, rtol=0, atol=1e-14 * np.abs(A).max())
    # the source loses a block: the next copy sees it
    nz = a.req_nonzero_blocks()
    a.zero_block(a.index(nz[0]))
    lt.btod_copy(a, bt, None, c=1.0)
    exp = a.to_dense()
    assert np.array_equal(bt.to_dense(), exp)
    out = permuted_tensor(gpu_ctx, prob["A"], tuple(range(n)))
    for _ in range(2):
        lt.btod_add([(a, None, 1.0), (bt, None, -0.5)], out)
        assert np.allclose(out.to_dense(), 0.5 * exp, rtol=0, atol=1e-14 * np.abs(A).max())
