"""Pins the CPU oracle (oracle/dense_oracle.py) before it is trusted:
against the real libtensor results of the golden fixtures, closed forms,
brute-force loops and permutation round trips."""
import itertools

import numpy as np
import pytest

import dense_oracle as O
import golden_cases as G


def test_contract2_matches_reference_golden():
    """Dense contraction of the unfolded inputs == libtensor's block result
    unfolded (tests/block_tensor/btod_contract2_test.C:787-817 method)."""
    checked = 0
    for name in G.names():
        prob, dense, meta, result = G.load(name)
        if result is None:
            continue
        b = dense["A"] if prob.get("self") else dense["B"]
        r = O.btod_contract2_dense(prob["contr"], dense["A"], b)
        assert r.shape == result.shape, name
        assert O.rel_err(r, result) < 1e-13, name
        checked += 1
    assert checked >= 20


@pytest.mark.parametrize("N,M,K", [(1, 1, 1), (2, 1, 1), (1, 2, 2), (2, 2, 2), (1, 1, 0),
                                   (0, 2, 1), (2, 0, 1), (3, 1, 2)])
def test_contract2_vs_loops(N, M, K):
    rng = np.random.default_rng(N * 100 + M * 10 + K)
    dims_a = list(rng.integers(2, 4, N + K))
    dims_b = list(rng.integers(2, 4, M + K))
    pa = list(rng.permutation(N + K))
    pb = list(rng.permutation(M + K))
    pairs = [(int(pa[N + k]), int(pb[M + k])) for k in range(K)]
    for ia, ib in pairs:
        dims_b[ib] = dims_a[ia]
    permc = list(rng.permutation(N + M))
    a = rng.standard_normal(dims_a)
    b = rng.standard_normal(dims_b)
    r1 = O.contract2(N, M, K, permc, pairs, a, b)
    r2 = O.contract2_loops(N, M, K, permc, pairs, a, b)
    assert np.allclose(r1, r2, rtol=1e-14, atol=1e-14)


def test_contract2_closed_forms():
    # all-ones inputs: every C element equals the contracted extent
    a = np.ones((3, 4, 5, 6))
    b = np.ones((4, 2, 6, 7))
    r = O.contract2(2, 2, 2, [0, 2, 1, 3], [(1, 0), (3, 2)], a, b)
    assert r.shape == (3, 2, 5, 7)
    assert np.all(r == 24.0)
    # identity matrix contraction reproduces the other operand
    x = np.random.default_rng(0).standard_normal((5, 6))
    r = O.contract2(1, 1, 1, None, [(1, 0)], x, np.eye(6))
    assert np.array_equal(r, x)
    # result permutation: C(j, i) = sum_k A(i, k) B(k, j)
    y = np.random.default_rng(1).standard_normal((6, 3))
    r = O.contract2(1, 1, 1, [1, 0], [(1, 0)], x, y)
    assert np.allclose(r, (x @ y).T, rtol=1e-14)


def test_copy_and_dotprod():
    rng = np.random.default_rng(3)
    a = rng.standard_normal((2, 3, 4, 5))
    for perm in itertools.permutations(range(4)):
        b = O.copy(a, perm, 2.0)
        assert b.shape == tuple(a.shape[p] for p in perm)
        inv = np.argsort(perm)
        assert np.array_equal(O.copy(b, inv, 0.5), a)   # exact round trip
    x = rng.standard_normal((3, 4))
    y = rng.standard_normal((4, 3))
    assert abs(O.dotprod(x, y, None, [1, 0], 2.0) - 2.0 * np.sum(x * y.T)) < 1e-12
    # brute force
    s = 0.0
    for i in range(3):
        for j in range(4):
            s += x[i, j] * y[j, i]
    assert abs(O.dotprod(x, y, None, [1, 0]) - s) < 1e-12


@pytest.mark.parametrize("name", [n for n in G.names() if not n.startswith("self")])
def test_cpu_baseline_restatement(name):
    """The benchmark's host baseline (per-pair GEMMs over libtensor's own
    contraction lists) reproduces libtensor's result on every scheduled
    block."""
    import cpu_contract as CC
    prob, dense, meta, result = G.load(name)
    sched = meta["schedule"]
    clst = {int(k): v for k, v in meta["clst"].items()}
    c = prob["contr"]
    ia, ib = CC.needed_blocks(clst, c["N"] + c["K"])
    ba = CC.blocks_of_dense(dense["A"], prob["A"], ia)
    bb = CC.blocks_of_dense(dense["B"], prob["B"], ib)
    bc = CC.BlockContraction(prob, ba, bb, sched, clst, result.shape)
    blocks = bc.perform()
    assert bc.max_rel_err(blocks, result) <= 1e-13
