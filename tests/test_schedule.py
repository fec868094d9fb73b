"""Integer parity of the host schedule with the real libtensor
(golden fixtures of tests/golden/, written by make_golden.py from
oracle/_ref/ref_driver): non-zero schedule, canonical index and allowedness of
every result block, coalesced contraction list of every scheduled block.
Host-side numerics of compute_block against libtensor's result."""
import os
import tempfile

import numpy as np
import pytest

import dense_oracle as O
import golden_cases as G
import problems as P


@pytest.mark.parametrize("name", G.names())
def test_schedule_symmetry_and_lists_match_reference(host_ctx, name):
    prob, dense, meta, result = G.load(name)
    a, b, op = P.build(host_ctx, prob, dense)
    assert op.get_schedule() == meta["schedule"]
    for k, (ac, al) in meta["symc"].items():
        mac, _, _, mal = op.symc_orbit(int(k))
        assert (mac, int(mal)) == (ac, al), (name, k)
    for s in meta["schedule"]:
        mine = P.clst_as_strings(op.clst(s))
        ref = P.normalize_ref_clst(meta["clst"].get(str(s), []))
        assert mine == ref, (name, s)


@pytest.mark.parametrize("name", G.names())
def test_compute_block_host_matches_reference(host_ctx, name):
    prob, dense, meta, result = G.load(name)
    a, b, op = P.build(host_ctx, prob, dense)
    ct = P.make_result_tensor(host_ctx, prob)
    for s in op.get_schedule():
        bd = ct.block_dims(s)
        st = ct.block_start(s)
        blk = op.compute_block_host(s, bd)
        sl = tuple(slice(st[i], st[i] + bd[i]) for i in range(len(bd)))
        assert O.rel_err(blk, result[sl]) < 1e-13 or np.max(np.abs(result[sl])) == 0


@pytest.mark.skipif(not P.ref_available(), reason="reference driver not built")
@pytest.mark.parametrize("seed", [1, 2, 3])
def test_live_reference_ladder_variants(host_ctx, seed):
    """Fresh random data and block splits against the live reference."""
    rng = np.random.default_rng(seed)
    ob = tuple(int(x) for x in rng.integers(1, 4, 2))
    vb = tuple(int(x) for x in rng.integers(2, 5, 3))
    prob = P.prob_ladder(o_blocks=ob, v_blocks=vb)
    dense = P.make_data(host_ctx, prob, 50 + seed)
    a, b, op = P.build(host_ctx, prob, dense)
    with tempfile.TemporaryDirectory() as d:
        ref = P.run_ref(prob, dense, d)
    assert op.get_schedule() == ref["schedule"]
    for s in ref["schedule"]:
        assert P.clst_as_strings(op.clst(s)) == P.normalize_ref_clst(ref["clst"][s])
