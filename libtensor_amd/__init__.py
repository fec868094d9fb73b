"""MI355X block-tensor contraction: host-side mirror of libtensor's
btod_contract2 path over the C ABI of liblt_amd.so (include/lt_amd.h).

Names follow libtensor: block_index_space, block_tensor, se_perm, se_part,
se_label, contraction2, btod_contract2, tod_copy, tod_contract2, tod_dotprod.
Block indexes are tuples (or absolute indexes) over the block-index
dimensions; blocks are row-major numpy float64 arrays.
"""
import ctypes

import numpy as np

from . import _capi
from ._capi import (GEMM_CONFIGS, GEMM_STACKED_FIRST, LT_MODE_GROUPED, LT_MODE_OZAKI,
                    LT_MODE_PER_BLOCK, LtError, LtRequestStats, LtStats, call, dptr, ints, sizes)

__all__ = ["Context", "product_table", "block_index_space", "block_tensor",
           "se_perm", "se_part", "se_label", "contraction2", "btod_contract2",
           "device_array", "tod_copy", "tod_add", "tod_set", "tod_contract2", "tod_contract2_noalloc", "tod_dotprod", "btod_dotprod", "btod_copy", "btod_add", "btod_symmetrize2", "btod_symmetrize3", "btod_contract3", "server",
           "LtError", "GEMM_CONFIGS", "GEMM_STACKED_FIRST", "LT_MODE_PER_BLOCK", "LT_MODE_GROUPED", "LT_MODE_OZAKI"]


def _abs_index(idx, dims):
    if isinstance(idx, (int, np.integer)):
        return int(idx)
    a = 0
    for i, d in zip(idx, dims):
        a = a * int(d) + int(i)
    return a


def _index(aidx, dims):
    out = []
    for d in reversed(dims):
        out.append(aidx % d)
        aidx //= d
    return tuple(reversed(out))


class Context(object):
    """Device context (device >= 0) or host-only context (device < 0)."""

    def __init__(self, device=0, hbm_cap=0):
        h = ctypes.c_void_p()
        call("lt_ctx_create", int(device), int(hbm_cap), ctypes.byref(h))
        self._h = h
        self.device = device

    def close(self):
        if self._h:
            call("lt_ctx_destroy", self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def sync(self):
        call("lt_ctx_sync", self._h)

    def mem(self):
        u, c = ctypes.c_size_t(), ctypes.c_size_t()
        call("lt_ctx_mem", self._h, ctypes.byref(u), ctypes.byref(c))
        return u.value, c.value

    def set_cap(self, cap):
        """Device memory budget (a contraction whose plan does not fit runs
        in batches of C blocks)."""
        call("lt_ctx_set_cap", self._h, int(cap))

    def set_threads(self, n):
        call("lt_ctx_set_threads", self._h, int(n))

    def last_op_ms(self):
        """Device time of the kernels of the last btod_copy / btod_add."""
        v = ctypes.c_double()
        call("lt_ctx_last_op_ms", self._h, ctypes.byref(v))
        return v.value


class product_table(object):
    """Point-group product table registered under an id
    (libtensor/symmetry/point_group_table.h)."""

    def __init__(self, table_id, nlabels):
        self.id = table_id
        call("lt_product_table_create", table_id.encode(), int(nlabels))

    def add_product(self, l1, l2, lr):
        call("lt_product_table_add_product", self.id.encode(), int(l1), int(l2), int(lr))


class block_index_space(object):
    """Dimensions plus split points (libtensor/core/block_index_space.h)."""

    def __init__(self, dims):
        self.dims = tuple(int(d) for d in dims)
        self.splits = []

    def split(self, mask, pos):
        self.splits.append((tuple(1 if m else 0 for m in mask), int(pos)))
        return self


class se_perm(object):
    """Permutational symmetry element (libtensor/symmetry/se_perm.h)."""

    def __init__(self, perm, coeff=1.0):
        self.perm = tuple(int(p) for p in perm)
        self.coeff = float(coeff)

    def _make(self, bt):
        h = ctypes.c_void_p()
        call("lt_se_perm_create", len(self.perm), sizes(self.perm), self.coeff,
             ctypes.byref(h))
        return h


class se_part(object):
    """Partition symmetry element (libtensor/symmetry/se_part.h); maps and
    forbidden partitions are applied in call order."""

    def __init__(self, pdims):
        self.pdims = tuple(int(p) for p in pdims)
        self.ops = []

    def add_map(self, idx1, idx2, coeff=1.0):
        self.ops.append(("map", tuple(idx1), tuple(idx2), float(coeff)))
        return self

    def mark_forbidden(self, idx):
        self.ops.append(("forbid", tuple(idx)))
        return self

    def _make(self, bt):
        h = ctypes.c_void_p()
        call("lt_se_part_create", bt._h, sizes(self.pdims), ctypes.byref(h))
        for op in self.ops:
            if op[0] == "map":
                call("lt_se_part_add_map", h, sizes(op[1]), sizes(op[2]), op[3])
            else:
                call("lt_se_part_mark_forbidden", h, sizes(op[1]))
        return h


class se_label(object):
    """Label symmetry element (libtensor/symmetry/se_label.h)."""

    def __init__(self, table_id):
        self.table = table_id
        self.assigns = []
        self.rule = None

    def assign(self, mask, blk, label):
        self.assigns.append((tuple(1 if m else 0 for m in mask), int(blk), int(label)))
        return self

    def set_rule(self, labels):
        if isinstance(labels, int):
            labels = [labels]
        self.rule = [int(l) for l in labels]
        return self

    def _make(self, bt):
        h = ctypes.c_void_p()
        call("lt_se_label_create", bt._h, self.table.encode(), ctypes.byref(h))
        for m, b, l in self.assigns:
            call("lt_se_label_assign", h, ints(m), b, l)
        if self.rule is not None:
            call("lt_se_label_set_rule", h, sizes(self.rule), len(self.rule))
        return h


class block_tensor(object):
    """Block tensor whose canonical non-zero blocks live in HBM
    (libtensor/block_tensor/block_tensor.h)."""

    def __init__(self, ctx, bis):
        if not isinstance(bis, block_index_space):
            bis = block_index_space(bis)
        h = ctypes.c_void_p()
        call("lt_btensor_create", ctx._h, len(bis.dims), sizes(bis.dims), ctypes.byref(h))
        self._h = h
        self.ctx = ctx
        self.dims = bis.dims
        self.n = len(bis.dims)
        for mask, pos in bis.splits:
            call("lt_btensor_split", h, ints(mask), pos)
        self._elems = []

    def __del__(self):
        try:
            if self._h:
                call("lt_btensor_destroy", self._h)
                self._h = None
        except Exception:
            pass

    @property
    def bidims(self):
        out = sizes([0] * self.n)
        call("lt_btensor_bidims", self._h, out)
        return tuple(out[i] for i in range(self.n))

    def abs_index(self, idx):
        return _abs_index(idx, self.bidims)

    def index(self, aidx):
        return _index(int(aidx), self.bidims)

    def insert(self, el):
        h = el._make(self)
        try:
            call("lt_btensor_insert", self._h, h)
        finally:
            call("lt_symel_destroy", h)
        self._elems.append(el)

    def block_dims(self, idx):
        out = sizes([0] * self.n)
        call("lt_btensor_block_dims", self._h, self.abs_index(idx), out)
        return tuple(out[i] for i in range(self.n))

    def block_start(self, idx):
        out = sizes([0] * self.n)
        call("lt_btensor_block_start", self._h, self.abs_index(idx), out)
        return tuple(out[i] for i in range(self.n))

    def req_nonzero_blocks(self):
        cnt = ctypes.c_size_t()
        call("lt_btensor_nonzero", self._h, None, 0, ctypes.byref(cnt))
        out = (ctypes.c_size_t * max(1, cnt.value))()
        call("lt_btensor_nonzero", self._h, out, cnt.value, ctypes.byref(cnt))
        return [out[i] for i in range(cnt.value)]

    def set_block(self, idx, data):
        a = self.abs_index(idx)
        arr = np.ascontiguousarray(data, dtype=np.float64)
        if arr.shape != self.block_dims(a):
            raise ValueError("block shape %s != %s" % (arr.shape, self.block_dims(a)))
        call("lt_btensor_set_block", self._h, a, dptr(arr))

    def get_block(self, idx):
        a = self.abs_index(idx)
        arr = np.empty(self.block_dims(a), dtype=np.float64)
        call("lt_btensor_get_block", self._h, a, dptr(arr))
        return arr

    def zero_block(self, idx):
        call("lt_btensor_zero_block", self._h, self.abs_index(idx))

    def zero(self):
        call("lt_btensor_zero_all", self._h)

    def orbit(self, idx):
        """(canonical abs index, perm canonical->block, coeff, allowed)."""
        ac = ctypes.c_size_t()
        perm = sizes([0] * self.n)
        c = ctypes.c_double()
        al = ctypes.c_int()
        call("lt_btensor_orbit", self._h, self.abs_index(idx), ctypes.byref(ac), perm,
             ctypes.byref(c), ctypes.byref(al))
        return ac.value, tuple(perm[i] for i in range(self.n)), c.value, bool(al.value)

    def random(self, seed, dist="uniform", scale=1.0, scale_dims=None):
        d0, d1 = (-1, -1) if scale_dims is None else scale_dims
        call("lt_btensor_random", self._h, int(seed), 1 if dist == "normal" else 0,
             float(scale), int(d0), int(d1))

    def to_dense(self):
        arr = np.empty(self.dims, dtype=np.float64)
        call("lt_btensor_export", self._h, dptr(arr))
        return arr

    def from_dense(self, arr):
        arr = np.ascontiguousarray(arr, dtype=np.float64)
        if arr.shape != self.dims:
            raise ValueError("shape")
        call("lt_btensor_import", self._h, dptr(arr))

    def set_immutable(self):
        """block_tensor::set_immutable: later writes are refused."""
        call("lt_btensor_set_immutable", self._h)

    def is_immutable(self):
        x = ctypes.c_int()
        call("lt_btensor_is_immutable", self._h, ctypes.byref(x))
        return bool(x.value)

    def copy_symmetry_from(self, other):
        call("lt_btensor_copy_symmetry", self._h, other._h)

    def merge_symmetry_with(self, other):
        """Lowers this tensor's symmetry to the one it shares with `other`
        (the symmetry of a sum; canonical blocks that split off are written
        out on the device)."""
        call("lt_btensor_merge_symmetry", self._h, other._h)


class contraction2(object):
    """Connection of the indexes of A, B and C (libtensor/core/contraction2.h)."""

    def __init__(self, N, M, K, permc=None):
        self.N, self.M, self.K = int(N), int(M), int(K)
        self.permc = tuple(range(N + M)) if permc is None else tuple(int(p) for p in permc)
        self.pairs = []

    def contract(self, ia, ib):
        if len(self.pairs) >= self.K:
            raise LtError(-1, "contraction is complete")
        self.pairs.append((int(ia), int(ib)))
        return self


class btod_contract2(object):
    """C = kc * contract(ka A, kb B) (libtensor/block_tensor/btod_contract2.h)."""

    def __init__(self, contr, bta, btb, ka=1.0, kb=1.0, kc=1.0):
        if len(contr.pairs) != contr.K:
            raise LtError(-1, "contraction is incomplete")
        h = ctypes.c_void_p()
        ka_idx = [p[0] for p in contr.pairs]
        kb_idx = [p[1] for p in contr.pairs]
        call("lt_btod_contract2_create", contr.N, contr.M, contr.K, sizes(contr.permc),
             sizes(ka_idx), sizes(kb_idx), bta._h, float(ka), btb._h, float(kb),
             float(kc), ctypes.byref(h))
        self._h = h
        self.contr = contr
        self.a, self.b = bta, btb

    def close(self):
        """Releases the operation (and its execution plan's device memory)."""
        if self._h:
            call("lt_btod_contract2_destroy", self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def perform(self, btc, mode=LT_MODE_GROUPED):
        call("lt_btod_contract2_perform", self._h, btc._h, int(mode))

    def perform_add(self, btc, d=1.0, mode=LT_MODE_GROUPED):
        call("lt_btod_contract2_perform_add", self._h, btc._h, float(d), int(mode))

    def get_schedule(self):
        cnt = ctypes.c_size_t()
        call("lt_btod_contract2_schedule", self._h, None, 0, ctypes.byref(cnt))
        out = (ctypes.c_size_t * max(1, cnt.value))()
        call("lt_btod_contract2_schedule", self._h, out, cnt.value, ctypes.byref(cnt))
        return [out[i] for i in range(cnt.value)]

    def ozaki_bounds(self):
        """Stated error of the last split-precision perform:
        [(canonical C block, bound of max |error|, max |C|)]."""
        n = ctypes.c_size_t()
        call("lt_btod_contract2_ozaki_bounds", self._h, None, None, None, 0, ctypes.byref(n))
        k = max(1, n.value)
        ai = (ctypes.c_size_t * k)()
        bd = (ctypes.c_double * k)()
        cm = (ctypes.c_double * k)()
        call("lt_btod_contract2_ozaki_bounds", self._h, ai, bd, cm, k, ctypes.byref(n))
        return [(ai[i], bd[i], cm[i]) for i in range(n.value)]

    def get_symmetry(self, dst):
        """dst (C's block index space, no blocks) takes the result symmetry."""
        call("lt_btod_contract2_get_symmetry", self._h, dst._h)

    def symc_orbit(self, aidx):
        nc = self.contr.N + self.contr.M
        ac = ctypes.c_size_t()
        perm = sizes([0] * nc)
        c = ctypes.c_double()
        al = ctypes.c_int()
        call("lt_btod_contract2_symc_orbit", self._h, int(aidx), ctypes.byref(ac), perm,
             ctypes.byref(c), ctypes.byref(al))
        return ac.value, tuple(perm[i] for i in range(nc)), c.value, bool(al.value)

    def clst(self, aidx_c):
        """Coalesced contraction list of a C block (original index space):
        list of (aia, acia, perm_a, coeff_a, aib, acib, perm_b, coeff_b)."""
        na = self.contr.N + self.contr.K
        nb = self.contr.M + self.contr.K
        w = 6 + na + nb
        cnt = ctypes.c_size_t()
        call("lt_btod_contract2_clst", self._h, int(aidx_c), None, 0, ctypes.byref(cnt))
        buf = np.zeros(max(1, cnt.value) * w, dtype=np.float64)
        call("lt_btod_contract2_clst", self._h, int(aidx_c), dptr(buf), cnt.value,
             ctypes.byref(cnt))
        out = []
        for k in range(cnt.value):
            r = buf[k * w:(k + 1) * w]
            j = 0
            aia, acia = int(r[0]), int(r[1])
            pa = tuple(int(x) for x in r[2:2 + na])
            ca = float(r[2 + na])
            j = 3 + na
            aib, acib = int(r[j]), int(r[j + 1])
            pb = tuple(int(x) for x in r[j + 2:j + 2 + nb])
            cb = float(r[j + 2 + nb])
            out.append((aia, acia, pa, ca, aib, acib, pb, cb))
        return out

    def compute_block(self, aidx_c, blk, zero=True, perm=None, coeff=1.0):
        """blk (=, or += unless zero) coeff * P(block aidx_c of the result), computed
        on the device (gen_bto_contract2::compute_block); blk is a device_array
        with the block's dims permuted by perm."""
        p = None if perm is None else sizes(perm)
        call("lt_btod_contract2_compute_block", self._h, 1 if zero else 0, int(aidx_c), p,
             float(coeff), blk.ptr)

    def compute_block_host(self, aidx_c, dims):
        arr = np.empty(dims, dtype=np.float64)
        call("lt_btod_contract2_compute_block_host", self._h, int(aidx_c), dptr(arr))
        return arr

    def stats(self):
        s = LtStats()
        call("lt_btod_contract2_stats", self._h, ctypes.byref(s))
        out = {f: getattr(s, f) for f, _ in LtStats._fields_}
        out["tune_ms"] = list(out["tune_ms"])
        return out

    def set_tile(self, cfg):
        call("lt_btod_contract2_set_tile", self._h, int(cfg))

    def set_graph(self, on):
        call("lt_btod_contract2_set_graph", self._h, 1 if on else 0)

    def set_async(self, on):
        """FP64 performs return once enqueued on the context's stream (later
        operations and reads on the context are ordered after them)."""
        call("lt_btod_contract2_set_async", self._h, 1 if on else 0)

    def set_ozaki(self, slices):
        call("lt_btod_contract2_set_ozaki", self._h, int(slices))

    def set_ozaki_tile(self, tile):
        """Split-precision tile shape: -1 measured (default), 0 16x32,
        1 16x64, 2 32x32."""
        call("lt_btod_contract2_set_ozaki_tile", self._h, int(tile))


class btod_contract3(object):
    """D = kd (A B) C (libtensor/block_tensor/btod_contract3.h): contr1 (a
    contraction2 with N = N1, M = N2 + K2, K = K1) contracts A with B into the
    intermediate AB, contr2 (N = N1 + N2, M = N3, K = K2) contracts AB with C."""

    def __init__(self, contr1, contr2, bta, btb, btc, kd=1.0, sympermab=None, symmab=True):
        """sympermab: btod_symcontract3 (AB replaced by AB + s P(AB), s = +1 if
        symmab, before the second contraction)."""
        for ct in (contr1, contr2):
            if len(ct.pairs) != ct.K:
                raise LtError(-1, "contraction is incomplete")
        n1, k1 = contr1.N, contr1.K
        n3, k2 = contr2.M, contr2.K
        n2 = contr1.M - k2
        if n2 < 0 or contr2.N != n1 + n2:
            raise LtError(-1, "contractions do not chain")
        h = ctypes.c_void_p()
        args = (n1, n2, n3, k1, k2, sizes(contr1.permc), sizes([p[0] for p in contr1.pairs]),
                sizes([p[1] for p in contr1.pairs]), sizes(contr2.permc),
                sizes([p[0] for p in contr2.pairs]), sizes([p[1] for p in contr2.pairs]),
                bta._h, btb._h)
        if sympermab is None:
            call("lt_btod_contract3_create", *args, btc._h, float(kd), ctypes.byref(h))
        else:
            call("lt_btod_symcontract3_create", *args, sizes(sympermab), 1 if symmab else 0,
                 btc._h, float(kd), ctypes.byref(h))
        self._h = h
        self.n = n1 + n2 + n3
        self.a, self.b, self.c = bta, btb, btc

    def __del__(self):
        try:
            if self._h:
                call("lt_btod_contract3_destroy", self._h)
                self._h = None
        except Exception:
            pass

    def perform(self, btd, mode=LT_MODE_GROUPED):
        call("lt_btod_contract3_perform", self._h, btd._h, int(mode))

    def perform_add(self, btd, k=1.0, mode=LT_MODE_GROUPED):
        call("lt_btod_contract3_perform_add", self._h, btd._h, float(k), int(mode))

    def get_schedule(self):
        cnt = ctypes.c_size_t()
        call("lt_btod_contract3_schedule", self._h, None, 0, ctypes.byref(cnt))
        out = (ctypes.c_size_t * max(1, cnt.value))()
        call("lt_btod_contract3_schedule", self._h, out, cnt.value, ctypes.byref(cnt))
        return [out[i] for i in range(cnt.value)]

    def sym_orbit(self, aidx):
        ac = ctypes.c_size_t()
        perm = sizes([0] * self.n)
        c = ctypes.c_double()
        al = ctypes.c_int()
        call("lt_btod_contract3_sym_orbit", self._h, int(aidx), ctypes.byref(ac), perm,
             ctypes.byref(c), ctypes.byref(al))
        return ac.value, tuple(perm[i] for i in range(self.n)), c.value, bool(al.value)

    def stats(self):
        v = [ctypes.c_double() for _ in range(2)]
        call("lt_btod_contract3_stats", self._h, *[ctypes.byref(x) for x in v])
        return {"gpu_ms": v[0].value, "flops": v[1].value}

    def plans(self):
        """Execution plans built so far by the two contractions."""
        n = ctypes.c_uint64()
        call("lt_btod_contract3_plans", self._h, ctypes.byref(n))
        return n.value


class btod_symmetrize2(object):
    """B = X + s P(X) for the result X of a btod_contract2 or of another
    btod_symmetrize2 (libtensor/block_tensor/btod_symmetrize2.h).  perm is a
    pair permutation (P P = 1); btod_symmetrize2(op, i1, i2, symm) of the
    reference is the transposition of i1 and i2 (use ``pair(n, i1, i2)``)."""

    def __init__(self, op, perm, symm):
        h = ctypes.c_void_p()
        if isinstance(op, btod_symmetrize2):
            call("lt_btod_symmetrize2_create_nested", op._h, sizes(perm), 1 if symm else 0,
                 ctypes.byref(h))
        elif isinstance(op, btod_contract3):
            call("lt_btod_symmetrize2_create_c3", op._h, sizes(perm), 1 if symm else 0,
                 ctypes.byref(h))
        else:
            call("lt_btod_symmetrize2_create", op._h, sizes(perm), 1 if symm else 0,
                 ctypes.byref(h))
        self._h = h
        self.op = op          # the operand must outlive this operation
        self.n = len(perm)

    @staticmethod
    def pair(n, i1, i2):
        p = list(range(n))
        p[i1], p[i2] = p[i2], p[i1]
        return p

    def __del__(self):
        try:
            if self._h:
                call("lt_btod_symmetrize2_destroy", self._h)
                self._h = None
        except Exception:
            pass

    def perform(self, btb, mode=LT_MODE_GROUPED):
        call("lt_btod_symmetrize2_perform", self._h, btb._h, int(mode))

    def perform_add(self, btb, d=1.0, mode=LT_MODE_GROUPED):
        call("lt_btod_symmetrize2_perform_add", self._h, btb._h, float(d), int(mode))

    def get_schedule(self):
        cnt = ctypes.c_size_t()
        call("lt_btod_symmetrize2_schedule", self._h, None, 0, ctypes.byref(cnt))
        out = (ctypes.c_size_t * max(1, cnt.value))()
        call("lt_btod_symmetrize2_schedule", self._h, out, cnt.value, ctypes.byref(cnt))
        return [out[i] for i in range(cnt.value)]

    def sym_orbit(self, aidx):
        """(canonical index, orbit transformation perm, coefficient, allowed)
        of a block under the result symmetry."""
        ac = ctypes.c_size_t()
        perm = sizes([0] * self.n)
        c = ctypes.c_double()
        al = ctypes.c_int()
        call("lt_btod_symmetrize2_sym_orbit", self._h, int(aidx), ctypes.byref(ac), perm,
             ctypes.byref(c), ctypes.byref(al))
        return ac.value, tuple(perm[i] for i in range(self.n)), c.value, bool(al.value)

    def stats(self):
        """Last perform: device ms (operand + gather), gather ms, gather bytes,
        and the operand's FLOPs."""
        v = [ctypes.c_double() for _ in range(4)]
        call("lt_btod_symmetrize2_stats", self._h, *[ctypes.byref(x) for x in v])
        return {"gpu_ms": v[0].value, "gather_ms": v[1].value, "gather_bytes": v[2].value,
                "flops": v[3].value}

    def compute_block(self, aidx_b, blk, zero=True, perm=None, coeff=1.0):
        """blk (=, or += unless zero) coeff * P(block aidx_b of the result) on the
        device; blk is a device_array with the block's dims permuted by perm."""
        p = None if perm is None else sizes(perm)
        call("lt_btod_symmetrize2_compute_block", self._h, 1 if zero else 0, int(aidx_b), p,
             float(coeff), blk.ptr)


class btod_symmetrize3(btod_symmetrize2):
    """B = sum over the six exchanges of three index groups generated by the
    pair permutations perm1 and perm2 of s(g) g(X) (libtensor/block_tensor/
    btod_symmetrize3.h; btod_symmetrize3(op, i1, i2, i3, symm) of the
    reference is perm1 = pair(n, i1, i2), perm2 = pair(n, i1, i3)).  X is the
    result of a btod_contract2 or of another symmetrization."""

    def __init__(self, op, perm1, perm2, symm):
        h = ctypes.c_void_p()
        if isinstance(op, btod_symmetrize2):
            call("lt_btod_symmetrize3_create_nested", op._h, sizes(perm1), sizes(perm2),
                 1 if symm else 0, ctypes.byref(h))
        else:
            call("lt_btod_symmetrize3_create", op._h, sizes(perm1), sizes(perm2),
                 1 if symm else 0, ctypes.byref(h))
        self._h = h
        self.op = op
        self.n = len(perm1)


# ---------------------------------------------------------------------------
# Dense tensors in device memory

class device_array(object):
    """Dense float64 array in device memory (row-major)."""

    def __init__(self, ctx, shape, data=None):
        self.ctx = ctx
        self.shape = tuple(int(s) for s in shape)
        self.size = int(np.prod(self.shape)) if self.shape else 1
        p = ctypes.c_void_p()
        call("lt_dev_alloc", ctx._h, max(8, 8 * self.size), ctypes.byref(p))
        self.ptr = p
        self._ctxh = ctx._h
        if data is not None:
            self.upload(data)

    def upload(self, data):
        a = np.ascontiguousarray(data, dtype=np.float64).reshape(self.shape)
        call("lt_dev_upload", self.ctx._h, self.ptr, a.ctypes.data_as(ctypes.c_void_p),
             8 * self.size)

    def download(self):
        a = np.empty(self.shape, dtype=np.float64)
        call("lt_dev_download", self.ctx._h, a.ctypes.data_as(ctypes.c_void_p), self.ptr,
             8 * self.size)
        return a

    def free(self):
        if self.ptr:
            call("lt_dev_free", self._ctxh, self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def tod_copy(a, b, perm=None, c=1.0, zero=True):
    """b = c P(a) (zero) or b += c P(a) (libtensor/dense_tensor/tod_copy.h)."""
    p = None if perm is None else sizes(perm)
    call("lt_tod_copy", a.ctx._h, len(a.shape), sizes(a.shape), a.ptr, p, float(c),
         1 if zero else 0, b.ptr)


def tod_add(ops, b, zero=True):
    """b = (zero ? 0 : b) + sum c_i P_i(a_i) for ops = [(a_i, perm_i, c_i)], all a_i of
    the same dimensions (libtensor/dense_tensor/tod_add.h)."""
    a0 = ops[0][0]
    n = len(a0.shape)
    ptrs = (ctypes.c_void_p * len(ops))(*[a.ptr for a, _, _ in ops])
    perms = []
    for a, p, _ in ops:
        if tuple(a.shape) != tuple(a0.shape):
            raise LtError(-2, "tod_add: operands of different shapes")
        perms += list(range(n)) if p is None else list(p)
    coeffs = (ctypes.c_double * len(ops))(*[float(c) for _, _, c in ops])
    call("lt_tod_add", a0.ctx._h, n, sizes(a0.shape), len(ops), ptrs, sizes(perms), coeffs,
         1 if zero else 0, b.ptr)


def tod_set(a, v=0.0):
    call("lt_tod_set", a.ctx._h, a.size, float(v), a.ptr)


def tod_contract2(contr, a, b, c, d=1.0, zero=True):
    """c (=|+=) d contract(a, b) (libtensor/dense_tensor/tod_contract2.h)."""
    ka = [p[0] for p in contr.pairs]
    kb = [p[1] for p in contr.pairs]
    call("lt_tod_contract2", a.ctx._h, contr.N, contr.M, contr.K, sizes(contr.permc),
         sizes(ka), sizes(kb), sizes(a.shape), a.ptr, sizes(b.shape), b.ptr, float(d),
         1 if zero else 0, c.ptr)


def tod_contract2_noalloc(contr, a, b, c, d=1.0, zero=True):
    """tod_contract2 with every temporary in the fixed per-GPU scratch region
    (no device allocation; LtError LT_ERR_PARAM when it does not fit)."""
    ka = [p[0] for p in contr.pairs]
    kb = [p[1] for p in contr.pairs]
    call("lt_tod_contract2_noalloc", a.ctx._h, contr.N, contr.M, contr.K,
         sizes(contr.permc), sizes(ka), sizes(kb), sizes(a.shape), a.ptr, sizes(b.shape),
         b.ptr, float(d), 1 if zero else 0, c.ptr)


def tod_dotprod(a, b, perma=None, permb=None, c=1.0):
    """sum P_a(a) P_b(b) times c (libtensor/dense_tensor/tod_dotprod.h)."""
    r = ctypes.c_double()
    pa = None if perma is None else sizes(perma)
    pb = None if permb is None else sizes(permb)
    call("lt_tod_dotprod", a.ctx._h, len(a.shape), sizes(a.shape), a.ptr, pa,
         sizes(b.shape), b.ptr, pb, float(c), ctypes.byref(r))
    return r.value


def btod_dotprod(bta, btb, perma=None, permb=None, ca=1.0, cb=1.0):
    """sum of ca P_a(A) * cb P_b(B) over all elements
    (libtensor/block_tensor/btod_dotprod.h)."""
    r = ctypes.c_double()
    call("lt_btod_dotprod", bta._h, None if perma is None else sizes(perma), float(ca),
         btb._h, None if permb is None else sizes(permb), float(cb), ctypes.byref(r))
    return r.value


def btod_copy(bta, btb, perm=None, c=1.0, add=False, d=1.0):
    """btb = c P(bta) (btb takes the permuted symmetry), or btb += d c P(bta)
    (libtensor/block_tensor/btod_copy.h: perform(btb), perform(btb, d))."""
    call("lt_btod_copy", bta._h, None if perm is None else sizes(perm), float(c), btb._h,
         1 if add else 0, float(d))


def btod_add(ops, btb, add=False, d=1.0):
    """btb = sum c_i P_i(bta_i) for ops = [(bta_i, perm_i, c_i)] (btb takes the
    operands' common permuted symmetry), or btb += d sum ... (libtensor/
    block_tensor/btod_add.h: add_op, perform(btb), perform(btb, d))."""
    n = btb.n
    hs = (ctypes.c_void_p * len(ops))(*[x[0]._h.value for x in ops])
    perms = None
    if any(p is not None for _, p, _ in ops):
        flat = []
        for _, p, _ in ops:
            flat += list(range(n)) if p is None else [int(v) for v in p]
        perms = sizes(flat)
    cs = (ctypes.c_double * len(ops))(*[float(x[2]) for x in ops])
    call("lt_btod_add", len(ops), hs, perms, cs, btb._h, 1 if add else 0, float(d))


class server(object):
    """Request loop on one context: submit(op, btc[, d]) returns a ticket,
    wait(ticket) the request's latency figures (continuous batching of
    btod_contract2 performs from many calculations)."""

    def __init__(self, ctx, max_batch=64, hbm_reserve=0):
        h = ctypes.c_void_p()
        call("lt_server_create", ctx._h, int(max_batch), int(hbm_reserve), ctypes.byref(h))
        self._h = h
        self.ctx = ctx

    def close(self):
        if self._h:
            call("lt_server_destroy", self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def submit(self, op, btc, add=False, d=1.0):
        t = ctypes.c_uint64()
        call("lt_server_submit", self._h, op._h, btc._h, 1 if add else 0, float(d),
             ctypes.byref(t))
        return t.value

    def wait(self, ticket):
        st = LtRequestStats()
        call("lt_server_wait", self._h, ctypes.c_uint64(ticket), ctypes.byref(st))
        return {f: getattr(st, f) for f, _ in LtRequestStats._fields_}

    def stats(self):
        b, r, mb = ctypes.c_uint64(), ctypes.c_uint64(), ctypes.c_uint64()
        mean = ctypes.c_double()
        call("lt_server_stats", self._h, ctypes.byref(b), ctypes.byref(r), ctypes.byref(mean),
             ctypes.byref(mb))
        return {"batches": b.value, "requests": r.value, "mean_batch": mean.value,
                "max_batch": mb.value}
