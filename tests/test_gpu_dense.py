"""Dense device kernels through the C ABI against the CPU oracle:
grouped FP64 MFMA GEMM (all operand layouts, ragged shapes), N-d permute,
dot product.  Tolerance: max|dC| / max|C| <= 1e-12 (FP64)."""
import itertools
import sys

import numpy as np
import pytest

import dense_oracle as O
import libtensor_amd as lt

pytestmark = pytest.mark.gpu

TOL = 1e-12


def _say(*a):
    print(*a)
    sys.stdout.flush()


@pytest.mark.parametrize("m,n,k", [(1, 1, 1), (16, 16, 4), (17, 33, 5), (64, 64, 64),
                                   (65, 63, 17), (100, 70, 130), (7, 300, 41), (256, 8, 512)])
@pytest.mark.parametrize("layout", ["ik,kj", "ik,jk", "ki,kj", "ki,jk"])
def test_matrix_products(gpu_ctx, m, n, k, layout):
    _say("gemm", m, n, k, layout)
    rng = np.random.default_rng(m * 7 + n * 3 + k)
    la, lb = layout.split(",")
    a = rng.standard_normal((m, k) if la == "ik" else (k, m))
    b = rng.standard_normal((k, n) if lb == "kj" else (n, k))
    c = lt.contraction2(1, 1, 1)
    c.contract(1 if la == "ik" else 0, 0 if lb == "kj" else 1)
    da, db = lt.device_array(gpu_ctx, a.shape, a), lt.device_array(gpu_ctx, b.shape, b)
    dc = lt.device_array(gpu_ctx, (m, n))
    lt.tod_contract2(c, da, db, dc, d=1.0, zero=True)
    ref = O.contract2(1, 1, 1, None, c.pairs, a, b)
    assert O.rel_err(dc.download(), ref) < TOL


def test_mfma_fragment_layout(gpu_ctx):
    """One 16x16x4 tile with structured operands: every output element is
    identified by its value, so a wrong lane mapping cannot pass."""
    a = np.zeros((16, 4))
    b = np.zeros((4, 16))
    a[:, 0] = np.arange(1, 17)
    b[0, :] = 1000.0 * np.arange(1, 17)
    a[:, 3] = 0.001 * np.arange(16)
    b[3, :] = 1.0
    c = lt.contraction2(1, 1, 1)
    c.contract(1, 0)
    da, db = lt.device_array(gpu_ctx, a.shape, a), lt.device_array(gpu_ctx, b.shape, b)
    dc = lt.device_array(gpu_ctx, (16, 16))
    lt.tod_contract2(c, da, db, dc)
    assert np.array_equal(dc.download(), a @ b)


@pytest.mark.parametrize("case", [
    ((2, 2, 2), (1, 0), (3, 2), [0, 2, 1, 3], (5, 7, 6, 9), (7, 4, 9, 8)),
    ((2, 2, 2), (2, 2), (3, 3), None, (6, 5, 8, 9), (7, 4, 8, 9)),
    ((2, 1, 1), (0, 0), None, [2, 0, 1], (13, 5, 11), (13, 17)),
    ((1, 1, 2), (2, 0), (0, 2), [1, 0], (9, 3, 10), (10, 6, 9)),
    ((3, 1, 1), (3, 0), None, [3, 1, 0, 2], (4, 5, 6, 7), (7, 9)),
    ((2, 2, 0), None, None, [3, 0, 2, 1], (4, 5), (6, 7)),
])
def test_tensor_contractions_with_permutations(gpu_ctx, case):
    (N, M, K), p1, p2, permc, da_, db_ = case
    _say("contract", case)
    pairs = [p for p in (p1, p2) if p is not None][:K]
    rng = np.random.default_rng(len(da_) * 10 + len(db_))
    a = rng.standard_normal(da_)
    b = rng.standard_normal(db_)
    c = lt.contraction2(N, M, K, permc)
    for ia, ib in pairs:
        c.contract(ia, ib)
    ref = O.contract2(N, M, K, permc, pairs, a, b, d=-0.75)
    dA, dB = lt.device_array(gpu_ctx, a.shape, a), lt.device_array(gpu_ctx, b.shape, b)
    dC = lt.device_array(gpu_ctx, ref.shape)
    lt.tod_contract2(c, dA, dB, dC, d=-0.75, zero=True)
    assert O.rel_err(dC.download(), ref) < TOL
    # accumulate: C += d contract
    lt.tod_contract2(c, dA, dB, dC, d=-0.75, zero=False)
    assert O.rel_err(dC.download(), 2 * ref) < TOL


def test_permute_all_4d(gpu_ctx):
    rng = np.random.default_rng(5)
    a = rng.standard_normal((3, 5, 7, 4))
    da = lt.device_array(gpu_ctx, a.shape, a)
    for perm in itertools.permutations(range(4)):
        ref = O.copy(a, perm, -2.0)
        db = lt.device_array(gpu_ctx, ref.shape)
        lt.tod_copy(da, db, perm, -2.0, zero=True)
        assert np.array_equal(db.download(), ref), perm
        lt.tod_copy(da, db, perm, 1.0, zero=False)
        assert np.allclose(db.download(), ref + O.copy(a, perm), rtol=0, atol=1e-15)


def test_permute_large_and_dot(gpu_ctx):
    rng = np.random.default_rng(6)
    a = rng.standard_normal((64, 33, 65))
    da = lt.device_array(gpu_ctx, a.shape, a)
    ref = O.copy(a, (2, 0, 1))
    db = lt.device_array(gpu_ctx, ref.shape)
    lt.tod_copy(da, db, (2, 0, 1))
    assert np.array_equal(db.download(), ref)
    b = rng.standard_normal((65, 64, 33))
    dbb = lt.device_array(gpu_ctx, b.shape, b)
    got = lt.tod_dotprod(da, dbb, None, (1, 2, 0), 0.5)
    exp = O.dotprod(a, b, None, (1, 2, 0), 0.5)
    assert abs(got - exp) <= 1e-12 * abs(exp) + 1e-12


def test_set(gpu_ctx):
    d = lt.device_array(gpu_ctx, (1000,))
    lt.tod_set(d, 3.5)
    assert np.all(d.download() == 3.5)


def test_tod_add(gpu_ctx):
    """tod_add: b = (zero ? 0 : b) + sum c_i P_i(a_i), operands in order."""
    rng = np.random.default_rng(8)
    shape = (4, 6, 4, 6)
    arrs = [rng.standard_normal(shape) for _ in range(3)]
    perms = [None, (2, 1, 0, 3), (0, 3, 2, 1)]
    coeffs = [1.0, -0.5, 2.0]
    das = [lt.device_array(gpu_ctx, shape, x) for x in arrs]
    ref = sum(O.copy(x, p, c) for x, p, c in zip(arrs, perms, coeffs))
    db = lt.device_array(gpu_ctx, ref.shape)
    lt.tod_add(list(zip(das, perms, coeffs)), db, zero=True)
    assert np.allclose(db.download(), ref, rtol=0, atol=1e-13)
    lt.tod_add(list(zip(das, perms, coeffs)), db, zero=False)
    assert np.allclose(db.download(), 2 * ref, rtol=0, atol=1e-13)
