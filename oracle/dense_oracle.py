"""Test infrastructure only: the CPU oracle of the dense operations of the
block-contraction path.  Nothing in the product imports this module; tests,
the smoke check and the CPU baseline of the benchmark use it as the checker.

numpy restatement (FP64) of:
  contract2  - tod_contract2<N, M, K> (libtensor/dense_tensor/impl/
               tod_contract2_impl.h:104-240): C (=|+=) d * contraction of A, B
               with the index connections of contraction2
               (libtensor/core/contraction2.h: connect(), contract())
  copy       - tod_copy<N> (libtensor/dense_tensor/impl/tod_copy_impl.h:57-141):
               B (=|+=) c * P(A)
  dotprod    - tod_dotprod<N> (libtensor/dense_tensor/impl/tod_dotprod_impl.h:
               83-153): sum P_a(A) * P_b(B) * c
  btod_contract2_dense - the reference's own block-test method
               (tests/block_tensor/btod_contract2_test.C:787-817): the block
               result unfolded to dense equals the dense contraction of the
               unfolded inputs times ka * kb * kc.
  contract2_loops - brute-force loops over every index (tiny shapes only).

Pinning: test_oracle.py checks these against the reference driver's golden
results (tests/golden/), closed forms (all-ones inputs give K-extent
products, identity matrices), brute-force loops, and permutation
round trips.
"""
import itertools

import numpy as np

_LETTERS = "abcdefghijklmnopqrstuvwxyz"


def result_order(N, M, K, permc, pairs):
    """Sources of the C indexes: list of ("A", i) / ("B", j).  The remaining
    indexes of A then of B, in order, permuted by permc (c[i] = rem[permc[i]])
    as contraction2::connect() (libtensor/core/contraction2.h)."""
    ka = set(p[0] for p in pairs)
    kb = set(p[1] for p in pairs)
    rem = [("A", i) for i in range(N + K) if i not in ka] + \
          [("B", j) for j in range(M + K) if j not in kb]
    if permc is None:
        permc = list(range(N + M))
    return [rem[permc[i]] for i in range(N + M)]


def einsum_spec(N, M, K, permc, pairs):
    la = [None] * (N + K)
    lb = [None] * (M + K)
    it = iter(_LETTERS)
    for ia, ib in pairs:
        l = next(it)
        la[ia] = l
        lb[ib] = l
    for i in range(N + K):
        if la[i] is None:
            la[i] = next(it)
    for j in range(M + K):
        if lb[j] is None:
            lb[j] = next(it)
    lc = [la[i] if t == "A" else lb[i] for t, i in result_order(N, M, K, permc, pairs)]
    return "".join(la) + "," + "".join(lb) + "->" + "".join(lc)


def contract2(N, M, K, permc, pairs, a, b, d=1.0, c=None, zero=True):
    r = d * np.einsum(einsum_spec(N, M, K, permc, pairs), a, b, optimize=True)
    if zero or c is None:
        return r
    return c + r


def contract2_loops(N, M, K, permc, pairs, a, b):
    """Brute-force nested loops (tiny shapes)."""
    order = result_order(N, M, K, permc, pairs)
    dims_c = [a.shape[i] if t == "A" else b.shape[i] for t, i in order]
    kdims = [a.shape[p[0]] for p in pairs]
    c = np.zeros(dims_c)
    for ic in itertools.product(*[range(x) for x in dims_c]):
        s = 0.0
        for ik in itertools.product(*[range(x) for x in kdims]):
            xa = [None] * (N + K)
            xb = [None] * (M + K)
            for pos, (t, i) in enumerate(order):
                if t == "A":
                    xa[i] = ic[pos]
                else:
                    xb[i] = ic[pos]
            for k, (ia, ib) in enumerate(pairs):
                xa[ia] = ik[k]
                xb[ib] = ik[k]
            s += a[tuple(xa)] * b[tuple(xb)]
        c[ic] = s
    return c


def copy(a, perm=None, c=1.0, b=None, zero=True):
    r = c * (a if perm is None else np.transpose(a, perm))
    if zero or b is None:
        return np.ascontiguousarray(r)
    return b + r


def dotprod(a, b, perma=None, permb=None, c=1.0):
    pa = a if perma is None else np.transpose(a, perma)
    pb = b if permb is None else np.transpose(b, permb)
    return float(np.sum(pa * pb) * c)


def btod_contract2_dense(prob_contr, dense_a, dense_b):
    cd = prob_contr
    return contract2(cd["N"], cd["M"], cd["K"], cd.get("permc"), cd["pairs"],
                     dense_a, dense_b,
                     d=cd.get("ka", 1.0) * cd.get("kb", 1.0) * cd.get("kc", 1.0))


def rel_err(x, ref):
    """max |x - ref| / max |ref| (the metric's relative error)."""
    den = np.max(np.abs(ref)) if ref.size else 0.0
    num = np.max(np.abs(x - ref)) if ref.size else 0.0
    return float(num / den) if den > 0 else float(num)
