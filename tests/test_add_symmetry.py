"""Symmetry of a sum, host side (no GPU): perform(btc, d) into a target of
another symmetry lowers the target to the symmetry it shares with the
contraction result (libtensor: so_merge(so_dirsum(...)) in gen_bto_aux_add /
addition_schedule), and btod_add takes the symmetry all its permuted
operands share.  Checked block by block (canonical block, allowedness and
orbit transformation) against libtensor's (fixtures add_* and badd_* written
by the real reference build, tests/golden/make_golden.py); the stored
results are pinned by the closed forms C + d R and ca A + cb P(B) of the
dense inputs and the base case's result."""
import numpy as np
import pytest

import dense_oracle as O
import golden_cases as G
import libtensor_amd as lt
import problems as P
from libtensor_amd import workloads as W


def _orbits_match(t, symt):
    """Number of blocks whose orbit differs from the fixture, and of those
    whose canonical block or allowedness differs."""
    bad, bad_can = [], []
    for aidx, (ac, al, perm, c) in symt.items():
        got = t.orbit(t.index(int(aidx)))
        if (got[0], got[3]) != (ac, bool(al)):
            bad_can.append(int(aidx))
        elif al and (list(got[1]), got[2]) != (perm, c):
            bad.append(int(aidx))
    return bad, bad_can


@pytest.mark.parametrize("name", G.add_names())
def test_target_symmetry_after_add(host_ctx, name):
    prob, dense, meta, result = G.load_add(name)
    a, b, op = P.build(host_ctx, prob, dense)
    r = P.make_result_tensor(host_ctx, prob)
    op.get_symmetry(r)
    t = W.build_tensor(host_ctx, prob["C"])
    t.merge_symmetry_with(r)
    bad, bad_can = _orbits_match(t, meta["symt_add"])
    assert not bad_can and not bad, (bad_can[:5], bad[:5])


@pytest.mark.parametrize("name", G.add_names())
def test_add_fixture_closed_form(name):
    """libtensor's target after perform(btc, d) is C + d R, R the dense
    contraction of the inputs."""
    prob, dense, meta, result = G.load_add(name)
    r = O.btod_contract2_dense(prob["contr"], dense["A"], dense["B"])
    exp = dense["C"] + prob["add"] * r
    assert O.rel_err(result, exp) <= 1e-13


@pytest.mark.parametrize("name", G.badd_names())
def test_btod_add_symmetry(host_ctx, name):
    prob, dense, meta, result = G.load_add(name)
    ba = prob["btod_add"]
    W.register_tables(prob)
    x = W.build_tensor(host_ctx, prob["A"])
    y = W.build_tensor(host_ctx, prob["B"])
    perm = ba["perm"]
    if list(perm) != list(range(len(perm))):
        pytest.skip("host-only check of the identity permutation (the GPU test covers it)")
    x.merge_symmetry_with(y)
    bad, bad_can = _orbits_match(x, meta["symt"])
    assert not bad_can and not bad, (bad_can[:5], bad[:5])


def scheduled_mask(prob, symt, schedule):
    """Dense mask of the blocks whose orbit (under the stored result
    symmetry) has its canonical block in the stored schedule."""
    ctx = lt.Context(-1)
    W.register_tables(prob)
    x = W.build_tensor(ctx, dict(prob["A"], elems=[]))
    mask = np.zeros(tuple(prob["A"]["dims"]), dtype=bool)
    sch = set(schedule)
    for aidx, (ac, al, perm, c) in symt.items():
        if al and ac in sch:
            idx = x.index(int(aidx))
            st, bd = x.block_start(idx), x.block_dims(idx)
            mask[tuple(slice(s0, s0 + d) for s0, d in zip(st, bd))] = True
    ctx.close()
    return mask


@pytest.mark.parametrize("name", G.badd_names())
def test_btod_add_fixture_closed_form(name):
    """libtensor's btod_add result is ca A + cb P(B) on the blocks of its
    schedule and zero elsewhere.  Its schedule (gen_bto_add::make_schedule,
    gen_block_tensor/impl/gen_bto_add_impl.h:290-336) holds, per non-zero
    canonical block of an operand, only the orbit of that block's image: where
    the sum's symmetry is lower than an operand's (badd_label01_nopart: the
    operand's partition maps are lost), other parts of the operand's orbit
    are left out, so the reference result misses those terms.  The library
    computes every block (tests/test_gpu_add.py checks the closed form)."""
    prob, dense, meta, result = G.load_add(name)
    ba = prob["btod_add"]
    exp = ba["ca"] * dense["A"] + ba["cb"] * np.transpose(dense["B"], ba["perm"])
    mask = scheduled_mask(prob, meta["symt"], meta["schedule"])
    assert O.rel_err(np.where(mask, result, 0.0), np.where(mask, exp, 0.0)) <= 1e-14
    assert not np.any(result[~mask])
    missed = np.abs(exp[~mask]).max() if np.any(~mask) else 0.0
    assert (missed > 0) == (name == "badd_label01_nopart")
