"""btod_contract2 on the GPU through the C ABI: every golden case (real
libtensor results) in per-block and grouped modes, the += path, the
config-1 shape at full size and a medium pp-ladder against the dense oracle,
determinism.  Tolerance: max|dC| / max|C| <= 1e-12."""
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


@pytest.mark.parametrize("mode", [lt.LT_MODE_PER_BLOCK, lt.LT_MODE_GROUPED])
@pytest.mark.parametrize("name", G.names())
def test_golden_cases(gpu_ctx, name, mode):
    _say("golden", name, mode)
    prob, dense, meta, result = G.load(name)
    a, b, op = P.build(gpu_ctx, prob, dense)
    c = P.make_result_tensor(gpu_ctx, prob)
    op.perform(c, mode)
    assert c.req_nonzero_blocks() == meta["schedule"]
    assert O.rel_err(c.to_dense(), result) < TOL


@pytest.mark.parametrize("name", ["ring_small", "ladder_full", "rand_221", "randsym_222"])
def test_perform_add(gpu_ctx, name):
    prob, dense, meta, result = G.load(name)
    a, b, op = P.build(gpu_ctx, prob, dense)
    c = P.make_result_tensor(gpu_ctx, prob)
    op.perform(c)
    op.perform_add(c, 0.5)
    assert O.rel_err(c.to_dense(), 1.5 * result) < TOL
    # += into an empty tensor with the result's symmetry
    c2 = P.make_result_tensor(gpu_ctx, prob)
    op.perform(c2)
    c2.zero()
    op.perform_add(c2, -2.0)
    assert O.rel_err(c2.to_dense(), -2.0 * result) < TOL


def test_config1_full_size(gpu_ctx):
    """C[i,j,a,b] = sum_{k,c} A[i,k,a,c] B[k,j,c,b], o=10, v=40, one block
    per dimension, no symmetry, U[-1,1] inputs."""
    prob = P.prob_ring(10, 40)
    rng = np.random.default_rng(1001)
    dense = {"A": rng.uniform(-1, 1, (10, 10, 40, 40)), "B": rng.uniform(-1, 1, (10, 10, 40, 40))}
    a, b, op = P.build(gpu_ctx, prob, dense)
    c = P.make_result_tensor(gpu_ctx, prob)
    for mode in (lt.LT_MODE_PER_BLOCK, lt.LT_MODE_GROUPED):
        op.perform(c, mode)
        ref = O.btod_contract2_dense(prob["contr"], dense["A"], dense["B"])
        assert O.rel_err(c.to_dense(), ref) < TOL
    st = op.stats()
    assert st["flops"] == 2 * 160000 * 400


def test_config3_batched_full_size(gpu_ctx):
    """Ring o=60, v=300, block sizes U{8..96} (1728 block pairs, 1.17e13
    FLOPs): grouped (graph) == eager grouped == per-block launches bit for
    bit with one kernel configuration; the autotuned configuration within
    1e-12; three C blocks against the dense oracle on slices of A and B."""
    from libtensor_amd import workloads as W
    prob = W.ring_batched_o60_v300()
    a, b, op = W.build_random(gpu_ctx, prob, 1003)
    c = W.result_tensor(gpu_ctx, prob)
    _say("config3 built")
    op.set_tile(0)
    op.perform(c, lt.LT_MODE_GROUPED)
    st = op.stats()
    assert st["n_pairs"] == 1728 and st["flops"] == 2.0 * 60**3 * 300**3
    first = c.to_dense()
    op.perform(c, lt.LT_MODE_PER_BLOCK)
    assert np.array_equal(c.to_dense(), first)
    op.set_graph(False)
    op.perform(c, lt.LT_MODE_GROUPED)
    assert np.array_equal(c.to_dense(), first)
    op.set_graph(True)
    op2 = lt.btod_contract2(W.make_contraction(prob), a, b)
    c2 = W.result_tensor(gpu_ctx, prob)
    op2.perform(c2, lt.LT_MODE_GROUPED)
    assert O.rel_err(c2.to_dense(), first) < TOL
    _say("config3 autotuned cfg", op2.stats()["tile_cfg"])
    A, B = a.to_dense(), b.to_dense()
    sched = op.get_schedule()
    for bi in (sched[0], sched[len(sched) // 2], sched[-1]):
        idx = c.index(bi)
        st0, dims = c.block_start(idx), c.block_dims(idx)
        (i0, j0, a0, b0), (ni, nj, na, nb) = st0, dims
        ref = np.einsum("ikac,kjcb->ijab", A[i0:i0 + ni, :, a0:a0 + na, :],
                        B[:, j0:j0 + nj, :, b0:b0 + nb], optimize=True)
        got = first[i0:i0 + ni, j0:j0 + nj, a0:a0 + na, b0:b0 + nb]
        assert O.rel_err(got, ref) < TOL, bi


def test_ladder_medium_against_dense_oracle(gpu_ctx):
    prob = P.prob_ladder(o_blocks=(3, 2, 2, 1), v_blocks=(5, 4, 3, 6, 2),
                         irreps_o=(0, 1, 2, 3), irreps_v=(0, 0, 1, 2, 3))
    dense = P.make_data(gpu_ctx, prob, 77)
    a, b, op = P.build(gpu_ctx, prob, dense)
    c = P.make_result_tensor(gpu_ctx, prob)
    op.perform(c, lt.LT_MODE_GROUPED)
    ref = O.btod_contract2_dense(prob["contr"], dense["A"], dense["B"])
    assert O.rel_err(c.to_dense(), ref) < TOL
    # deterministic: same bits on a second perform and in per-block mode
    first = c.to_dense()
    op.perform(c, lt.LT_MODE_GROUPED)
    assert np.array_equal(c.to_dense(), first)
    op.perform(c, lt.LT_MODE_PER_BLOCK)
    assert np.array_equal(c.to_dense(), first)


@pytest.mark.parametrize("tile", list(range(lt.GEMM_CONFIGS)))
@pytest.mark.parametrize("name", G.names())
def test_kernel_configurations(gpu_ctx, name, tile):
    """Every GEMM kernel configuration (LDS-tiled and direct, all tile
    shapes) on every golden case, in both launch modes."""
    _say("config", name, tile)
    prob, dense, meta, result = G.load(name)
    a, b, op = P.build(gpu_ctx, prob, dense)
    op.set_tile(tile)
    for mode in (lt.LT_MODE_GROUPED, lt.LT_MODE_PER_BLOCK):
        c = P.make_result_tensor(gpu_ctx, prob)
        op.perform(c, mode)
        assert op.stats()["tile_cfg"] == tile
        assert O.rel_err(c.to_dense(), result) < TOL


def test_set_tile_applies_to_the_next_perform(gpu_ctx):
    """A configuration set after a plan was built with another one takes
    effect at the next perform (the stale plan is dropped); results agree."""
    prob, dense, meta, result = G.load("ladder_spin")
    a, b, op = P.build(gpu_ctx, prob, dense)
    c = P.make_result_tensor(gpu_ctx, prob)
    op.set_tile(lt.GEMM_STACKED_FIRST)
    op.perform(c, lt.LT_MODE_GROUPED)
    assert op.stats()["tile_cfg"] == lt.GEMM_STACKED_FIRST
    plans = op.stats()["n_plans"]
    op.set_tile(lt.GEMM_STACKED_FIRST + 3)
    op.perform(c, lt.LT_MODE_GROUPED)
    st = op.stats()
    assert st["tile_cfg"] == lt.GEMM_STACKED_FIRST + 3
    assert st["n_plans"] == plans + 1
    assert O.rel_err(c.to_dense(), result) < TOL
    op.set_tile(lt.GEMM_STACKED_FIRST + 3)          # unchanged: the plan stays
    op.perform(c, lt.LT_MODE_GROUPED)
    assert op.stats()["n_plans"] == plans + 1


def test_autotuner_records_timings(gpu_ctx):
    prob, dense, meta, result = G.load("ladder_full")
    a, b, op = P.build(gpu_ctx, prob, dense)
    c = P.make_result_tensor(gpu_ctx, prob)
    op.perform(c)
    st = op.stats()
    tune = st["tune_ms"][:lt.GEMM_CONFIGS]
    # the stacked configurations are measured, the others are not
    measured = [k for k in range(lt.GEMM_CONFIGS) if tune[k] > 0]
    assert measured == list(range(lt.GEMM_STACKED_FIRST, lt.GEMM_CONFIGS))
    assert tune[st["tile_cfg"]] == min(tune[k] for k in measured)
    assert O.rel_err(c.to_dense(), result) < TOL


@pytest.mark.parametrize("name", ["ladder_full", "rand_221", "randsym_222", "self_ladder"])
def test_graph_replays_match_eager(gpu_ctx, name):
    """The grouped perform runs as a captured HIP graph; every replay, with
    eager work on the same stream in between, gives the eager result bit for
    bit (same kernels, same configuration, fixed summation order)."""
    _say("graph", name)
    prob, dense, meta, result = G.load(name)
    a, b, op = P.build(gpu_ctx, prob, dense)
    c = P.make_result_tensor(gpu_ctx, prob)
    op.set_graph(False)
    op.perform(c, lt.LT_MODE_GROUPED)
    assert op.stats()["graph"] == 0
    eager = c.to_dense()
    assert O.rel_err(eager, result) < TOL
    op.set_graph(True)
    for rep in range(4):
        op.perform(c, lt.LT_MODE_GROUPED)
        assert op.stats()["graph"] == 1
        assert np.array_equal(c.to_dense(), eager), rep
        op.perform(c, lt.LT_MODE_PER_BLOCK)        # eager launches in between
        assert O.rel_err(c.to_dense(), result) < TOL
    # accumulate through the graph: C = r, then C += 0.5 r twice
    op.set_graph(False)
    op.perform(c, lt.LT_MODE_GROUPED)
    op.perform_add(c, 0.5, lt.LT_MODE_GROUPED)
    op.perform_add(c, 0.5, lt.LT_MODE_GROUPED)
    eager_add = c.to_dense()
    op.set_graph(True)
    op.perform(c, lt.LT_MODE_GROUPED)
    op.perform_add(c, 0.5, lt.LT_MODE_GROUPED)
    op.perform_add(c, 0.5, lt.LT_MODE_GROUPED)
    assert op.stats()["graph"] == 1
    assert np.array_equal(c.to_dense(), eager_add)
    assert O.rel_err(eager_add, 2.0 * result) < TOL


def test_plan_not_reused_for_a_new_result_tensor(gpu_ctx):
    """A result tensor created after another was released may occupy the
    same host address; the cached plan must not be taken for it."""
    import gc
    prob, dense, meta, result = G.load("ladder_full")
    a, b, op = P.build(gpu_ctx, prob, dense)
    for rep in range(4):
        c = P.make_result_tensor(gpu_ctx, prob)
        op.perform(c, lt.LT_MODE_GROUPED)
        assert O.rel_err(c.to_dense(), result) < TOL, rep
        del c
        gc.collect()


@pytest.mark.parametrize("mode", [lt.LT_MODE_GROUPED, lt.LT_MODE_PER_BLOCK])
def test_repeated_perform_after_the_result_changed(gpu_ctx, mode):
    """A perform into the tensor the operation wrote last keeps its structure;
    once the tensor was changed in between (a block zeroed, another symmetry
    set) the next perform rebuilds it, and the result is the same."""
    prob, dense, meta, result = G.load("ladder_full")
    a, b, op = P.build(gpu_ctx, prob, dense)
    c = P.make_result_tensor(gpu_ctx, prob)
    op.perform(c, mode)
    nz = c.req_nonzero_blocks()
    op.perform(c, mode)
    assert c.req_nonzero_blocks() == nz
    assert O.rel_err(c.to_dense(), result) < TOL
    c.zero_block(c.index(nz[0]))
    op.perform(c, mode)
    assert c.req_nonzero_blocks() == nz
    assert O.rel_err(c.to_dense(), result) < TOL
    c.zero()
    op.perform(c, mode)
    assert c.req_nonzero_blocks() == nz
    assert O.rel_err(c.to_dense(), result) < TOL
    # another operation writes C and sets its symmetry
    c2 = P.make_result_tensor(gpu_ctx, prob)
    op.perform(c2, mode)
    lt.btod_copy(c2, c, c=2.0)
    assert O.rel_err(c.to_dense(), 2.0 * result) < TOL
    op.perform(c, mode)
    assert O.rel_err(c.to_dense(), result) < TOL


@pytest.mark.parametrize("name", ["ladder_full", "ladder_spin", "rand_222", "randsym_222",
                                  "self_ladder", "ring_small"])
def test_btod_dotprod(gpu_ctx, name):
    """btod_dotprod on resident tensors against the dense oracle: the result
    with itself and with permutations of itself (index pairs exchanged), and
    an input with itself."""
    import itertools
    prob, dense, meta, result = G.load(name)
    a, b, op = P.build(gpu_ctx, prob, dense)
    c = P.make_result_tensor(gpu_ctx, prob)
    op.perform(c)
    R = c.to_dense()
    norm = float(np.sum(R * R))
    for perm in list(itertools.permutations(range(R.ndim)))[:8]:
        if np.transpose(R, perm).shape != R.shape:
            continue
        try:
            got = lt.btod_dotprod(c, c, None, perm, 1.0, -0.5)
        except lt.LtError:
            continue            # block index spaces differ under this permutation
        exp = O.dotprod(R, R, None, perm, -0.5)
        assert abs(got - exp) <= 1e-12 * norm, perm
    A = a.to_dense()
    got = lt.btod_dotprod(a, a)
    assert abs(got - float(np.sum(A * A))) <= 1e-12 * float(np.sum(A * A))


@pytest.mark.parametrize("name", G.names())
def test_compute_block_device(gpu_ctx, name):
    """compute_block on the device (direct tensors): scheduled blocks and
    arbitrary blocks of the result space (canonical or not, allowed or not)
    against the golden dense result, identity and permuted + scaled +
    accumulated, and against the host restatement."""
    _say("compute_block", name)
    prob, dense, meta, result = G.load(name)
    a, b, op = P.build(gpu_ctx, prob, dense)
    c = P.make_result_tensor(gpu_ctx, prob)
    nblk = int(np.prod(c.bidims))
    rng = np.random.default_rng(7)
    picks = list(meta["schedule"][:4]) + [int(x) for x in rng.integers(0, nblk, size=6)]
    scale = float(np.max(np.abs(result))) or 1.0
    for bi in picks:
        idx = c.index(bi)
        st, dims = c.block_start(idx), c.block_dims(idx)
        ref = result[tuple(slice(s0, s0 + n) for s0, n in zip(st, dims))]
        out = lt.device_array(gpu_ctx, tuple(dims))
        op.compute_block(bi, out)
        got = out.download()
        assert np.max(np.abs(got - ref)) <= TOL * scale, bi
        host = op.compute_block_host(bi, tuple(dims))
        assert np.max(np.abs(got - host)) <= TOL * scale, bi
        perm = tuple(int(x) for x in rng.permutation(len(dims)))
        pd = tuple(dims[p] for p in perm)
        out2 = lt.device_array(gpu_ctx, pd, np.ones(pd))
        op.compute_block(bi, out2, zero=False, perm=perm, coeff=-0.5)
        exp = 1.0 - 0.5 * np.transpose(ref, perm)
        assert np.max(np.abs(out2.download() - exp)) <= TOL * max(scale, 1.0), (bi, perm)
        out.free()
        out2.free()


def test_compute_block_sees_structure_changes(gpu_ctx):
    """compute_block keeps each block's contraction terms between calls and
    drops them when an operand's non-zero blocks change: a repeated call
    gives the same bits, the block is zero once A's blocks are zeroed, and
    the first result comes back when they are set again."""
    prob, dense, meta, result = G.load("ladder_spin")
    a, b, op = P.build(gpu_ctx, prob, dense)
    c = P.make_result_tensor(gpu_ctx, prob)
    bi = None
    for x in meta["schedule"]:
        idx = c.index(int(x))
        st, dims = c.block_start(idx), c.block_dims(idx)
        if np.any(result[tuple(slice(s0, s0 + n) for s0, n in zip(st, dims))]):
            bi = int(x)
            break
    assert bi is not None
    dims = tuple(c.block_dims(c.index(bi)))
    out = lt.device_array(gpu_ctx, dims)
    op.compute_block(bi, out)
    first = out.download()
    assert np.any(first)
    op.compute_block(bi, out)
    assert np.array_equal(out.download(), first)
    a.zero()
    op.compute_block(bi, out)
    assert not np.any(out.download())
    a.from_dense(dense["A"])
    op.compute_block(bi, out)
    assert np.array_equal(out.download(), first)
    out.free()


def test_async_performs_are_ordered(gpu_ctx):
    """Performs that return once enqueued (set_async): a tensor upload after
    one waits for it, a second operation's perform sees the new data, reads
    of results wait; results equal those of waited-for performs bit for bit,
    and the device times are available afterwards."""
    from libtensor_amd import workloads as W
    prob = W.pp_ladder_o40_v200()
    a, b, op1 = W.build_random(gpu_ctx, prob, 1002)
    cc = prob["contr"]
    op2 = lt.btod_contract2(W.make_contraction(prob), a, b, cc.get("ka", 1.0),
                            cc.get("kb", 1.0), cc.get("kc", 1.0))
    c1 = W.result_tensor(gpu_ctx, prob)
    c2 = W.result_tensor(gpu_ctx, prob)
    blk = a.req_nonzero_blocks()[0]
    old = a.get_block(blk)
    new = np.random.default_rng(5).standard_normal(old.shape) * 0.01
    # references, one perform at a time
    op1.perform(c1, lt.LT_MODE_GROUPED)
    ref_old = c1.to_dense()
    a.set_block(blk, new)
    op2.perform(c2, lt.LT_MODE_GROUPED)
    ref_new = c2.to_dense()
    a.set_block(blk, old)
    assert not np.array_equal(ref_old, ref_new)
    # the same, enqueued: perform, upload, perform, then read both
    c1.zero()
    c2.zero()
    op1.set_async(True)
    op2.set_async(True)
    for rep in range(3):
        op1.perform(c1, lt.LT_MODE_GROUPED)
        a.set_block(blk, new)
        op2.perform(c2, lt.LT_MODE_GROUPED)
        a.set_block(blk, old)
    assert np.array_equal(c1.to_dense(), ref_old)
    assert np.array_equal(c2.to_dense(), ref_new)
    assert op1.stats()["t_gpu_ms"] > 0.0 and op2.stats()["t_gpu_ms"] > 0.0
