"""Test infrastructure only: the host-CPU baseline of the benchmark (the
oracle's block contraction, timed).  Nothing in the product imports this
module.

A restatement of libtensor's per-block computation over libtensor's own
contraction lists (the schedule and the coalesced list of every scheduled C
block, printed by the reference build, oracle/_ref/ref_driver):
gen_bto_contract2_block::compute_block (libtensor/gen_block_tensor/impl/
gen_bto_contract2_block_impl.h:100-205) forms, for every pair, the A and B
blocks from their canonical blocks (the pair's permutation and scalar), and
tod_contract2 (libtensor/dense_tensor/impl/tod_contract2_impl.h:124-240)
contracts them with one GEMM: here numpy.tensordot, which copies the
operands into matrix layout and calls the BLAS numpy links (OpenBLAS) -- the
reference's linalg_cblas path (linalg/cblas/linalg_cblas_level3.C:10-60).
Blocks run in parallel workers, one BLAS thread each, C blocks dealt out
longest first: libtensor's thread_pool(n, n) model with single-threaded BLAS
(gen_bto_contract2_batch_impl.h:273-276); the workers are forked processes
(numpy's per-call overhead holds the interpreter lock, so threads would not
scale on small blocks).

Pinning: tests/test_oracle.py checks the result of this restatement against
the golden results of the reference build.
"""
import os
import statistics
import time

import numpy as np


def block_bounds(td):
    """Per dimension: list of (start, length) of the blocks of a tensor
    description ({"dims", "splits": [(mask, pos)]})."""
    out = []
    for i, d in enumerate(td["dims"]):
        pts = sorted(set(int(pos) for mask, pos in td.get("splits", []) if mask[i]))
        edges = [0] + pts + [int(d)]
        out.append([(edges[k], edges[k + 1] - edges[k]) for k in range(len(edges) - 1)])
    return out


def block_index(aidx, bounds):
    idx = []
    for b in reversed(bounds):
        idx.append(aidx % len(b))
        aidx //= len(b)
    return tuple(reversed(idx))


def block_view(dense, bounds, aidx):
    idx = block_index(aidx, bounds)
    return dense[tuple(slice(bounds[i][k][0], bounds[i][k][0] + bounds[i][k][1])
                       for i, k in enumerate(idx))]


class BlockContraction:
    """One btod_contract2 of libtensor, restated on the host over its lists.

    prob: problem description (tensors "A", "B", "contr" with N, M, K,
    permc, pairs and the scalars ka, kb, kc); blocks_a / blocks_b: the
    canonical non-zero blocks of the inputs, {absolute index: ndarray} (a
    missing block is zero); schedule: the scheduled canonical C blocks;
    clst: for every scheduled block the rows "aia acia perm_a... coeff_a aib
    acib perm_b... coeff_b" as the reference driver prints them."""

    def __init__(self, prob, blocks_a, blocks_b, schedule, clst, result_shape):
        c = prob["contr"]
        self.N, self.M, self.K = c["N"], c["M"], c["K"]
        self.permc = list(c.get("permc") or range(self.N + self.M))
        self.ka = [p[0] for p in c["pairs"]]
        self.kb = [p[1] for p in c["pairs"]]
        self.scale = c.get("ka", 1.0) * c.get("kb", 1.0) * c.get("kc", 1.0)
        self.a, self.b = blocks_a, blocks_b
        self.ba = block_bounds(prob["A"])
        self.bb = block_bounds(prob["B"] if not prob.get("self") else prob["A"])
        na, nb = self.N + self.K, self.M + self.K
        # the result's blocks: C index i is A's or B's index per the
        # contraction's result order, with that index's splits
        ka, kb = set(self.ka), set(self.kb)
        rem = [self.ba[i] for i in range(na) if i not in ka] + \
              [self.bb[j] for j in range(nb) if j not in kb]
        self.bc = [rem[self.permc[i]] for i in range(self.N + self.M)]
        self.result_shape = tuple(result_shape)
        self.work = []
        for ic in schedule:
            pairs = []
            for row in clst[ic]:
                r = [float(x) if ("." in x or "e" in x) else int(x) for x in row]
                acia, perm_a, ca = int(r[1]), [int(x) for x in r[2:2 + na]], r[2 + na]
                o = 3 + na
                acib, perm_b, cb = int(r[o + 1]), [int(x) for x in r[o + 2:o + 2 + nb]], \
                    r[o + 2 + nb]
                pairs.append((acia, perm_a, acib, perm_b, ca * cb * self.scale))
            self.work.append((ic, pairs))

    def compute_block(self, item):
        ic, pairs = item
        acc = None
        for acia, perm_a, acib, perm_b, d in pairs:
            if d == 0.0 or acia not in self.a or acib not in self.b:
                continue
            blka = np.transpose(self.a[acia], perm_a)
            blkb = np.transpose(self.b[acib], perm_b)
            r = np.tensordot(blka, blkb, axes=(self.ka, self.kb))
            r = np.transpose(r, self.permc)
            if acc is None:
                acc = d * r
            else:
                acc += d * r
        return ic, acc

    def perform(self):
        """Every scheduled block; returns {aidx: block (or None: zero)}."""
        return dict(self.compute_block(w) for w in self.work)

    def max_rel_err(self, blocks, result):
        """max |block - result block| / max |result| over the scheduled
        (canonical) blocks, result the dense reference result."""
        den = float(np.max(np.abs(result))) if result.size else 0.0
        num = 0.0
        for ic, blk in blocks.items():
            ref = block_view(result, self.bc, ic)
            got = np.zeros(ref.shape) if blk is None else blk
            num = max(num, float(np.max(np.abs(got - ref))) if ref.size else 0.0)
        return num / den if den > 0 else num


def blocks_of_dense(dense, td, aidxs):
    """{aidx: block} of the given blocks of an unfolded tensor."""
    bounds = block_bounds(td)
    return {int(a): np.ascontiguousarray(block_view(dense, bounds, int(a))) for a in aidxs}


def needed_blocks(clst, na):
    """Canonical A and B blocks the lists read."""
    ia, ib = set(), set()
    for rows in clst.values():
        for row in rows:
            ia.add(int(row[1]))
            ib.add(int(row[3 + na + 1]))
    return ia, ib


def host_info():
    """CPU model, logical CPUs of the machine and of this process."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        affinity = len(os.sched_getaffinity(0))
    except AttributeError:
        affinity = os.cpu_count() or 1
    return {"cpu_model": model, "nproc": os.cpu_count() or 1, "affinity": affinity}


def blas_info():
    try:
        import threadpoolctl
        for lib in threadpoolctl.threadpool_info():
            if lib.get("user_api") == "blas":
                return "%s %s" % (lib.get("internal_api"), lib.get("version"))
    except Exception:
        pass
    return "numpy BLAS"


# Worker processes of time_perform (forked: they see the contraction and the
# shared result buffer without copying them).
_BC = None
_OUT = None


def _run_chunk(chunk):
    for k in chunk:
        ic, blk = _BC.compute_block(_BC.work[k])
        off, n = _BC.offsets[k]
        if blk is None:
            _OUT[off:off + n] = 0.0
        else:
            _OUT[off:off + n] = blk.reshape(-1)
    return len(chunk)


def time_perform(bc, workers, runs):
    """One warm-up, then `runs` timed performs with `workers` forked worker
    processes (one BLAS thread each; the C blocks dealt out longest first, the
    results written into one shared buffer) or, with workers == 1, in this
    process.  Returns (median ms, all times, {aidx: block})."""
    global _BC, _OUT
    import multiprocessing as mp
    try:
        import threadpoolctl
        limit = threadpoolctl.threadpool_limits(1, user_api="blas")
    except Exception:
        limit = None
    try:
        # result layout: the scheduled blocks back to back
        sizes = []
        for ic, pairs in bc.work:
            idx = block_index(ic, bc.bc)
            sizes.append(int(np.prod([bc.bc[i][k][1] for i, k in enumerate(idx)])))
        bc.offsets = []
        off = 0
        for n in sizes:
            bc.offsets.append((off, n))
            off += n
        shm = mp.RawArray("d", max(1, off))
        _OUT = np.frombuffer(shm, dtype=np.float64)
        _BC = bc
        # cost of a block: sum over its pairs of the block sizes (balanced
        # chunks, longest first, dealt round robin)
        cost = [(sizes[k] * max(1, len(bc.work[k][1])), k) for k in range(len(bc.work))]
        cost.sort(reverse=True)
        chunks = [[] for _ in range(max(1, workers))]
        for j, (cst, k) in enumerate(cost):
            chunks[j % len(chunks)].append(k)
        pool = mp.get_context("fork").Pool(workers) if workers > 1 else None
        try:
            def once():
                if pool is None:
                    _run_chunk(list(range(len(bc.work))))
                else:
                    pool.map(_run_chunk, chunks, chunksize=1)
            once()
            times = []
            for _ in range(runs):
                t0 = time.perf_counter()
                once()
                times.append((time.perf_counter() - t0) * 1e3)
        finally:
            if pool is not None:
                pool.close()
                pool.join()
        blocks = {}
        for k, (ic, pairs) in enumerate(bc.work):
            idx = block_index(ic, bc.bc)
            shape = tuple(bc.bc[i][kk][1] for i, kk in enumerate(idx))
            o, n = bc.offsets[k]
            blocks[ic] = _OUT[o:o + n].reshape(shape).copy()
    finally:
        _BC = None
        _OUT = None
        if limit is not None:
            limit.restore_original_limits()
    return statistics.median(times), times, blocks
