"""btod_symmetrize2 / btod_symmetrize3 on the device through the C ABI:
B = X + s P(X), and the sum over the exchanges of three index groups, for
the contraction X of a golden case or of a problem of the fixture's own
(and symmetrizations of symmetrizations) against libtensor's result
(fixtures symm_*, symm3_*), with B := result, B += d result and direct
blocks (compute_block, transformed and accumulated)."""
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


def chain(op, symm):
    ops = [op]
    for entry in symm:
        if len(entry) == 3:
            ops.append(lt.btod_symmetrize3(ops[-1], entry[0], entry[1], entry[2]))
        else:
            ops.append(lt.btod_symmetrize2(ops[-1], entry[0], entry[1]))
    return ops


@pytest.mark.parametrize("mode", [lt.LT_MODE_GROUPED, lt.LT_MODE_PER_BLOCK])
@pytest.mark.parametrize("name", G.symm_names())
def test_perform(gpu_ctx, name, mode):
    _say("btod_symmetrize2", name, "mode", mode)
    prob, dense, bresult, meta, result = G.load_symm(name)
    a, b, op = P.build(gpu_ctx, prob, dense)
    ops = chain(op, meta["symm"])
    s = ops[-1]
    c = P.make_result_tensor(gpu_ctx, prob)
    s.perform(c, mode)
    scale = float(np.max(np.abs(result))) or 1.0
    assert np.max(np.abs(c.to_dense() - result)) <= TOL * scale
    nz = set(c.req_nonzero_blocks())
    assert nz <= set(meta["schedule"])
    st = s.stats()
    assert st["gather_bytes"] > 0 and st["gpu_ms"] >= st["gather_ms"] > 0
    s.perform_add(c, -0.5, mode)
    assert np.max(np.abs(c.to_dense() - 0.5 * result)) <= TOL * scale
    # a second perform reuses the operand's plan and the temporary
    s.perform(c, mode)
    assert np.max(np.abs(c.to_dense() - result)) <= TOL * scale


@pytest.mark.parametrize("name", G.symm_names())
def test_compute_block(gpu_ctx, name):
    _say("btod_symmetrize2 compute_block", name)
    prob, dense, bresult, meta, result = G.load_symm(name)
    a, b, op = P.build(gpu_ctx, prob, dense)
    s = chain(op, meta["symm"])[-1]
    c = P.make_result_tensor(gpu_ctx, prob)
    rng = np.random.default_rng(5)
    nblk = int(np.prod(c.bidims))
    picks = list(meta["schedule"][:6]) + [int(x) for x in rng.integers(0, nblk, size=4)]
    scale = float(np.max(np.abs(result))) or 1.0
    canon = set(meta["schedule"])
    for bi in picks:
        idx = c.index(bi)
        st, dims = c.block_start(idx), c.block_dims(idx)
        ref = result[tuple(slice(s0, s0 + n) for s0, n in zip(st, dims))]
        if bi not in canon and len(meta["symm"][-1]) == 2:
            # btod_symmetrize2 has sources for its scheduled blocks only: the
            # block is written as zero (btod_symmetrize3 gathers the sources
            # of any block, which gives the block's value)
            ref = np.zeros(dims)
        out = lt.device_array(gpu_ctx, tuple(dims), np.full(dims, 7.0))
        s.compute_block(bi, out)
        assert np.max(np.abs(out.download() - ref)) <= TOL * scale, bi
        perm = tuple(int(x) for x in rng.permutation(len(dims)))
        pd = tuple(dims[p] for p in perm)
        out2 = lt.device_array(gpu_ctx, pd, np.ones(pd))
        s.compute_block(bi, out2, zero=False, perm=perm, coeff=-0.5)
        exp = 1.0 - 0.5 * np.transpose(ref, perm)
        assert np.max(np.abs(out2.download() - exp)) <= TOL * max(scale, 1.0), (bi, perm)
        out.free()
        out2.free()


def test_perform_add_needs_the_result_symmetry(gpu_ctx):
    prob, dense, bresult, meta, result = G.load_symm("symm_ring_ijab")
    a, b, op = P.build(gpu_ctx, prob, dense)
    s = chain(op, meta["symm"])[-1]
    c = P.make_result_tensor(gpu_ctx, prob)
    op.perform(c)          # C takes the contraction's symmetry, not the result's
    with pytest.raises(lt.LtError):
        s.perform_add(c, 1.0)
