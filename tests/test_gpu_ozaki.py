"""Split-precision mode (LT_MODE_OZAKI): FP64 emulated with int8 slices on
INT8 MFMA, checked against the FP64 reference results.

* exactness where the slices hold the inputs exactly (small integers),
* error decreasing with the slice count and <= 1e-12 (max|dC|/max|C|) at
  7 slices on every golden case,
* per-block errors reported for the pp-ladder shape at reduced size."""
import sys

import numpy as np
import pytest

import dense_oracle as O
import golden_cases as G
import libtensor_amd as lt
import problems as P

pytestmark = pytest.mark.gpu

TOL = 1e-12


def _say(*a):
    print(*a)
    sys.stdout.flush()


def test_small_integers_exact(gpu_ctx):
    """Integers |x| <= 60 are 7-bit multiples of their row's power-of-two
    scale, so one slice holds them exactly: every slice count gives the
    exact integer product (this also pins the MFMA operand and result
    lane layouts)."""
    prob = P.prob_ring(4, 12, so=(2,), sv=(5,))
    rng = np.random.default_rng(3)
    da = tuple(prob["A"]["dims"])
    db = tuple(prob["B"]["dims"])
    dense = {"A": rng.integers(-60, 61, size=da).astype(float),
             "B": rng.integers(-60, 61, size=db).astype(float)}
    a, b, op = P.build(gpu_ctx, prob, dense)
    c = P.make_result_tensor(gpu_ctx, prob)
    op.perform(c, lt.LT_MODE_GROUPED)
    exact = c.to_dense()
    for s in (1, 2, 7):
        op.set_ozaki(s)
        op.perform(c, lt.LT_MODE_OZAKI)
        assert op.stats()["ozaki_slices"] == s
        assert np.array_equal(c.to_dense(), exact), s


@pytest.mark.parametrize("name", G.names())
def test_golden_cases_error_vs_slices(gpu_ctx, name):
    prob, dense, meta, result = G.load(name)
    a, b, op = P.build(gpu_ctx, prob, dense)
    errs = {}
    for s in (3, 5, 7):
        op.set_ozaki(s)
        c = P.make_result_tensor(gpu_ctx, prob)
        op.perform(c, lt.LT_MODE_OZAKI)
        assert c.req_nonzero_blocks() == meta["schedule"]
        errs[s] = O.rel_err(c.to_dense(), result)
    _say("ozaki", name, errs)
    assert errs[7] < TOL
    assert errs[5] <= max(errs[3], 1e-15) and errs[7] <= max(errs[5], 1e-15)


def test_perform_add(gpu_ctx):
    prob, dense, meta, result = G.load("ladder_full")
    a, b, op = P.build(gpu_ctx, prob, dense)
    op.set_ozaki(7)
    c = P.make_result_tensor(gpu_ctx, prob)
    op.perform(c, lt.LT_MODE_OZAKI)
    op.perform_add(c, -0.5, lt.LT_MODE_OZAKI)
    assert O.rel_err(c.to_dense(), 0.5 * result) < TOL


@pytest.mark.parametrize("tile", [0, 1, 2])
def test_every_tile_shape_exact_on_small_integers(gpu_ctx, tile):
    """Each product-kernel tile shape (16x32, 16x64, 32x32), forced, gives the
    exact integer product (C blocks of 60 rows and columns: the last row and
    column panels of every shape are partial)."""
    prob = P.prob_ring(4, 12, so=(2,), sv=(5,))
    rng = np.random.default_rng(11)
    dense = {"A": rng.integers(-60, 61, size=tuple(prob["A"]["dims"])).astype(float),
             "B": rng.integers(-60, 61, size=tuple(prob["B"]["dims"])).astype(float)}
    a, b, op = P.build(gpu_ctx, prob, dense)
    c = P.make_result_tensor(gpu_ctx, prob)
    op.perform(c, lt.LT_MODE_GROUPED)
    exact = c.to_dense()
    op.set_ozaki(3)
    op.set_ozaki_tile(tile)
    op.perform(c, lt.LT_MODE_OZAKI)
    assert np.array_equal(c.to_dense(), exact), tile


@pytest.mark.parametrize("name", ["ladder_full", "ladder_spin", "rand_221"])
def test_tile_shapes_agree(gpu_ctx, name):
    """Forced tile shapes and the measured choice give the same sums (the
    integer slice products are exact, the FP64 recombination per element runs
    in the same order): within 1e-14 of each other and 1e-12 of the FP64
    reference, with 7 slices."""
    prob, dense, meta, result = G.load(name)
    a, b, op = P.build(gpu_ctx, prob, dense)
    op.set_ozaki(7)
    outs = {}
    for tile in (-1, 0, 1, 2):
        op.set_ozaki_tile(tile)
        c = P.make_result_tensor(gpu_ctx, prob)
        op.perform(c, lt.LT_MODE_OZAKI)
        outs[tile] = c.to_dense()
        assert O.rel_err(outs[tile], result) < TOL, tile
    for tile in (0, 1, 2):
        assert O.rel_err(outs[tile], outs[-1]) < 1e-14, tile


def test_slice_count_changes_keep_device_memory(gpu_ctx):
    """Rebuilding the split form for another slice count or tile shape
    frees the previous one: alternating performs leave the context's
    device memory in use unchanged."""
    prob, dense, meta, result = G.load("ladder_full")
    a, b, op = P.build(gpu_ctx, prob, dense)
    c = P.make_result_tensor(gpu_ctx, prob)
    op.set_ozaki(3)
    op.perform(c, lt.LT_MODE_OZAKI)
    op.set_ozaki(7)
    op.perform(c, lt.LT_MODE_OZAKI)
    used = gpu_ctx.mem()[0]
    for s, tile in ((3, -1), (7, -1), (5, 0), (7, 2), (3, 1), (7, -1)):
        op.set_ozaki(s)
        op.set_ozaki_tile(tile)
        op.perform(c, lt.LT_MODE_OZAKI)
    op.set_ozaki(7)
    op.set_ozaki_tile(-1)
    op.perform(c, lt.LT_MODE_OZAKI)
    assert gpu_ctx.mem()[0] == used
    assert O.rel_err(c.to_dense(), result) <= TOL


def _block(dense, t, aidx):
    idx = t.index(aidx)
    st, bd = t.block_start(idx), t.block_dims(idx)
    return dense[tuple(slice(s0, s0 + d) for s0, d in zip(st, bd))]


@pytest.mark.parametrize("name", G.names())
def test_stated_bound_covers_the_error(gpu_ctx, name):
    """The per-block a-priori bound the library states for a split-precision
    perform covers the measured error of every C block (against libtensor's
    FP64 result), and the overall stated relative bound covers the overall
    relative error, at 3, 5 and 7 slices."""
    prob, dense, meta, result = G.load(name)
    a, b, op = P.build(gpu_ctx, prob, dense)
    for s in (3, 5, 7):
        op.set_ozaki(s)
        c = P.make_result_tensor(gpu_ctx, prob)
        op.perform(c, lt.LT_MODE_OZAKI)
        got = c.to_dense()
        bounds = op.ozaki_bounds()
        assert sorted(x[0] for x in bounds) == meta["schedule"]
        for aidx, bound, cmax in bounds:
            err = np.max(np.abs(_block(got, c, aidx) - _block(result, c, aidx)), initial=0.0)
            assert err <= bound, (s, aidx, err, bound)
            assert cmax == np.max(np.abs(_block(got, c, aidx)), initial=0.0)
        assert O.rel_err(got, result) <= op.stats()["ozaki_max_rel_err"], s


def test_immutable_operand_split_once(gpu_ctx):
    """Slices of an operand read in place from an immutable tensor are cut
    once per plan: results equal (bit for bit) those of the same perform with
    a mutable operand, on repeated performs; writes to the immutable tensor
    are refused."""
    prob, dense, meta, result = G.load("ladder_full")
    a, b, op = P.build(gpu_ctx, prob, dense)
    op.set_ozaki(7)
    c = P.make_result_tensor(gpu_ctx, prob)
    op.perform(c, lt.LT_MODE_OZAKI)
    ref = c.to_dense()
    a2, b2, op2 = P.build(gpu_ctx, prob, dense)
    b2.set_immutable()
    assert b2.is_immutable()
    op2.set_ozaki(7)
    c2 = P.make_result_tensor(gpu_ctx, prob)
    for _ in range(3):
        op2.perform(c2, lt.LT_MODE_OZAKI)
        assert np.array_equal(c2.to_dense(), ref)
    with pytest.raises(lt.LtError) as e:
        b2.from_dense(dense["B"])
    assert e.value.code == -6       # LT_ERR_STATE
    with pytest.raises(lt.LtError):
        lt.btod_copy(b2, b2)            # B := P(B) writes into B


@pytest.mark.parametrize("which", ["A", "B"])
@pytest.mark.parametrize("name", G.names())
def test_immutable_operand_copies_prepared_once(gpu_ctx, name, which):
    """Copies of blocks of an immutable operand (permuted or unfolded for the
    contraction) are prepared and split once per plan: repeated
    split-precision performs, with an FP64 perform of the same operation in
    between, equal the perform with mutable operands bit for bit."""
    _say("immutable", name, which)
    prob, dense, meta, result = G.load(name)
    a, b, op = P.build(gpu_ctx, prob, dense)
    op.set_ozaki(5)
    c = P.make_result_tensor(gpu_ctx, prob)
    op.perform(c, lt.LT_MODE_OZAKI)
    ref = c.to_dense()
    a2, b2, op2 = P.build(gpu_ctx, prob, dense)
    (a2 if which == "A" else b2).set_immutable()
    op2.set_ozaki(5)
    c2 = P.make_result_tensor(gpu_ctx, prob)
    for i in range(3):
        op2.perform(c2, lt.LT_MODE_OZAKI)
        assert np.array_equal(c2.to_dense(), ref), i
        if i == 1:
            op2.perform(c2, lt.LT_MODE_GROUPED)
            assert O.rel_err(c2.to_dense(), result) < 1e-12


def test_nonfinite_input_gives_nan(gpu_ctx):
    """A NaN in an operand makes the products of its row NaN (as in FP64)
    and the stated bound infinite, instead of finite wrong values."""
    prob, dense, meta, result = G.load("rand_222")
    A = dense["A"].copy()
    A.flat[np.flatnonzero(A)[0]] = np.nan
    d2 = dict(dense, A=A)
    a, b, op = P.build(gpu_ctx, prob, d2)
    c = P.make_result_tensor(gpu_ctx, prob)
    op.perform(c, lt.LT_MODE_GROUPED)
    fp64 = c.to_dense()
    op.set_ozaki(7)
    op.perform(c, lt.LT_MODE_OZAKI)
    oz = c.to_dense()
    assert np.isnan(fp64).any()
    # every element FP64 makes NaN is NaN in the split form too
    assert np.all(np.isnan(oz[np.isnan(fp64)]))
    assert op.stats()["ozaki_max_rel_err"] == np.inf or np.isnan(op.stats()["ozaki_max_rel_err"])
