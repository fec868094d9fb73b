"""lt_tod_contract2_noalloc: the tod_contract2 entry point whose temporaries
live in the fixed per-GPU scratch region.  Checked through the C ABI against
lt_tod_contract2 (same kernels: identical bits) and the dense oracle
(max|dC|/max|C| <= 1e-12), on two calls in a row, the second smaller, with
every alignment permutation in use (so the region holds the aligned A, B and
C), plus accumulation and the size refusal."""
import sys

import numpy as np
import pytest

import dense_oracle as O
import libtensor_amd as lt

pytestmark = pytest.mark.gpu

TOL = 1e-12
SCRATCH = 32 << 20


def _say(*a):
    print(*a)
    sys.stdout.flush()


def _ring(o, v):
    """C[i,j,a,b] = sum_{k,c} A[i,k,a,c] B[k,j,c,b]: A, B and C are all
    permuted for the GEMM."""
    c = lt.contraction2(2, 2, 2, permc=[0, 2, 1, 3])
    c.contract(1, 0)
    c.contract(3, 2)
    return c, (o, o, v, v), (o, o, v, v)


def _run(ctx, o, v, seed, zero):
    contr, sa, sb = _ring(o, v)
    rng = np.random.default_rng(seed)
    a = rng.uniform(-1, 1, sa)
    b = rng.uniform(-1, 1, sb)
    c0 = rng.uniform(-1, 1, (o, o, v, v))
    da, db = lt.device_array(ctx, sa, a), lt.device_array(ctx, sb, b)
    dc1 = lt.device_array(ctx, (o, o, v, v), c0)
    dc2 = lt.device_array(ctx, (o, o, v, v), c0)
    lt.tod_contract2(contr, da, db, dc1, d=0.5, zero=zero)
    used = ctx.mem()[0]
    lt.tod_contract2_noalloc(contr, da, db, dc2, d=0.5, zero=zero)
    assert ctx.mem()[0] == used            # no device allocation
    r1, r2 = dc1.download(), dc2.download()
    ref = 0.5 * O.contract2(2, 2, 2, [0, 2, 1, 3], contr.pairs, a, b)
    if not zero:
        ref = ref + c0
    assert np.array_equal(r1, r2)
    assert O.rel_err(r2, ref) <= TOL
    return r2


@pytest.mark.parametrize("zero", [True, False])
def test_noalloc_matches_tod_contract2_twice(gpu_ctx, zero):
    _say("noalloc ring o=10 v=40 then o=6 v=24, zero", zero)
    _run(gpu_ctx, 10, 40, 11, zero)
    _run(gpu_ctx, 6, 24, 12, zero)


def test_noalloc_plain_gemm(gpu_ctx):
    """No alignment permutation: only descriptors and tiles in the region."""
    _say("noalloc plain gemm")
    rng = np.random.default_rng(3)
    a, b = rng.standard_normal((70, 130)), rng.standard_normal((130, 90))
    c = lt.contraction2(1, 1, 1)
    c.contract(1, 0)
    da, db = lt.device_array(gpu_ctx, a.shape, a), lt.device_array(gpu_ctx, b.shape, b)
    dc1, dc2 = lt.device_array(gpu_ctx, (70, 90)), lt.device_array(gpu_ctx, (70, 90))
    lt.tod_contract2(c, da, db, dc1)
    lt.tod_contract2_noalloc(c, da, db, dc2)
    assert np.array_equal(dc1.download(), dc2.download())
    assert O.rel_err(dc2.download(), a @ b) <= TOL


def test_noalloc_refuses_oversized_before_device_work(gpu_ctx):
    """Aligned copies of A, B and C of 16.6 MB each do not fit in the
    32 MiB region together: refused with LT_ERR_PARAM, C untouched."""
    _say("noalloc oversized")
    o, v = 12, 120           # 12*12*120*120*8 = 16.6 MB per tensor
    contr, sa, sb = _ring(o, v)
    assert 3 * o * o * v * v * 8 > SCRATCH
    da = lt.device_array(gpu_ctx, sa, np.ones(sa))
    db = lt.device_array(gpu_ctx, sb, np.ones(sb))
    sentinel = np.full((o, o, v, v), 7.0)
    dc = lt.device_array(gpu_ctx, (o, o, v, v), sentinel)
    with pytest.raises(lt.LtError) as e:
        lt.tod_contract2_noalloc(contr, da, db, dc)
    assert e.value.code == -1 and "scratch" in str(e.value)
    assert np.array_equal(dc.download(), sentinel)
    # the next call after a refusal still works
    _run(gpu_ctx, 4, 16, 13, True)
