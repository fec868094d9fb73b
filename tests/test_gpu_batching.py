"""Memory-bounded performs: when the plan of the whole contraction does not
fit the context's device memory budget (libtensor: allocator<double>::init's
memory limit read by gen_bto_contract2_batching_policy.h:48-70), the
contraction runs in batches of C blocks (gen_bto_contract2_impl.h:122-199)
and gives the same result: = and += on golden cases against libtensor's
results, and on the pp-ladder (o = 40, v = 200) against the one-plan
perform, max|dC|/max|C| <= 1e-12."""
import sys

import numpy as np
import pytest

import dense_oracle as O
import golden_cases as G
import libtensor_amd as lt
import problems as P
from libtensor_amd import _capi
from libtensor_amd import workloads as W

pytestmark = pytest.mark.gpu

TOL = 1e-12

# budgets tried, as fractions of the whole plan above the resident tensors:
# the full plan never fits, and the first fraction at which the batches fit
# is used (a golden case with few C blocks can need most of the plan for one)
FRACS = (0.75, 0.6, 0.5, 0.34)


def _say(*a):
    print(*a)
    sys.stdout.flush()


def _run_capped(ctx, make_op, c, expect, fracs):
    """Runs = and += (d 0.5) of a fresh operation under shrinking budgets;
    returns (fraction, batches) of the first budget the batches fit."""
    op = make_op()
    op.perform(c, lt.LT_MODE_GROUPED)
    assert op.stats()["n_batches"] == 1
    assert O.rel_err(c.to_dense(), expect) <= TOL
    used_plan = ctx.mem()[0]
    op.close()                          # releases the plan
    used = ctx.mem()[0]
    plan_bytes = used_plan - used
    assert plan_bytes > 0
    op2 = make_op()
    cap0 = ctx.mem()[1]
    try:
        for f in fracs:
            cap = used + int(plan_bytes * f)
            ctx.set_cap(cap)
            c.zero()
            try:
                op2.perform(c, lt.LT_MODE_GROUPED)
            except _capi.LtError as e:
                assert e.code == -5, e       # LT_ERR_NOMEM, before device work
                assert np.all(c.to_dense() == 0.0)
                continue
            nb = op2.stats()["n_batches"]
            assert nb > 1, nb
            assert O.rel_err(c.to_dense(), expect) <= TOL
            op2.perform_add(c, 0.5, lt.LT_MODE_GROUPED)
            assert op2.stats()["n_batches"] > 1
            assert O.rel_err(c.to_dense(), 1.5 * expect) <= TOL
            assert ctx.mem()[0] <= cap
            _say("  budget %.2f of the plan (%d bytes): %d batches" % (f, plan_bytes, nb))
            return f, nb
        pytest.fail("no budget in %s fits a batch" % (fracs,))
    finally:
        op2.close()
        ctx.set_cap(cap0)


@pytest.mark.parametrize("name", ["ladder_full", "ladder_spin", "rand_222", "ring_small"])
def test_batches_when_the_plan_does_not_fit(name):
    _say("batching", name)
    ctx = lt.Context(0)
    try:
        prob, dense, meta, result = G.load(name)
        a, b, op = P.build(ctx, prob, dense)
        op.close()
        c = P.make_result_tensor(ctx, prob)
        _run_capped(ctx, lambda: P.build_op(ctx, prob, a, b), c, result, FRACS)
    finally:
        ctx.close()


def test_pp_ladder_in_batches():
    _say("batching pp-ladder o40 v200")
    ctx = lt.Context(0)
    try:
        prob = W.pp_ladder_o40_v200()
        a, b, op = W.build_random(ctx, prob, 1002)
        c = W.result_tensor(ctx, prob)
        op.perform(c, lt.LT_MODE_GROUPED)
        expect = c.to_dense()
        op.close()
        c = W.result_tensor(ctx, prob)

        def make():
            cc = prob["contr"]
            return lt.btod_contract2(W.make_contraction(prob), a, b, cc.get("ka", 1.0),
                                     cc.get("kb", 1.0), cc.get("kc", 1.0))
        f, nb = _run_capped(ctx, make, c, expect, (0.34,))
        assert nb >= 4
    finally:
        ctx.close()


@pytest.mark.parametrize("name", ["ladder_spin", "ring_small"])
def test_split_precision_in_batches(name):
    """A split-precision perform under a budget its plan and slices do not fit
    runs in batches of C blocks: the same bits as the one-plan perform (every
    C element gets the same exact slice products and the same FP64
    recombination), and the stated bounds of every batch's C blocks."""
    _say("batching, split precision", name)
    ctx = lt.Context(0)
    cap0 = ctx.mem()[1]
    try:
        prob, dense, meta, result = G.load(name)
        a, b, op = P.build(ctx, prob, dense)
        op.set_ozaki(7)
        c = P.make_result_tensor(ctx, prob)
        op.perform(c, lt.LT_MODE_OZAKI)
        ref = c.to_dense()
        full = sorted(op.ozaki_bounds())
        stated = op.stats()["ozaki_max_rel_err"]
        assert O.rel_err(ref, result) <= stated
        used_plan = ctx.mem()[0]
        op.close()
        used = ctx.mem()[0]
        plan_bytes = used_plan - used
        op2 = P.build_op(ctx, prob, a, b)
        op2.set_ozaki(7)
        for f in FRACS:
            ctx.set_cap(used + int(plan_bytes * f))
            c.zero()
            try:
                op2.perform(c, lt.LT_MODE_OZAKI)
            except lt.LtError as e:
                assert e.code == -5, e
                assert np.all(c.to_dense() == 0.0)     # refused before device work
                continue
            nb = op2.stats()["n_batches"]
            assert nb > 1, nb
            assert np.array_equal(c.to_dense(), ref)
            assert sorted(op2.ozaki_bounds()) == full
            assert op2.stats()["ozaki_max_rel_err"] == stated
            _say("  budget %.2f of the plan: %d batches" % (f, nb))
            return
        pytest.fail("no budget fits a batch")
    finally:
        ctx.set_cap(cap0)
        ctx.close()
