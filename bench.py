"""Benchmark: btod_contract2 on one MI355X for the pp-ladder configuration.

Workload: R[i,j,a,b] = sum_{c,d} T[i,j,c,d] V[a,b,c,d], o = 40, v = 200,
alpha/beta spin partitions, 4 C2v irreps, uneven blocks (o: 7+5+4+4 per spin,
v: 8..32), (ij)/(ab)/(cd) antisymmetry, V symmetric under (ab)<->(cd);
T ~ N(0, 0.01^2)/(1+|c-i|), V ~ N(0,1), symmetrized, fixed seeds.  Inputs are
resident in HBM before timing.  A step is one btod_contract2::perform of the
whole contraction (all non-zero block pairs) with the cached execution plan.

Prints one JSON line: effective FP64-accurate TFLOP/s (schedule FLOPs / wall
time per step), checked against the FP64 reference (max|dR|/max|R| <= 1e-12,
libtensor built from its sources, run on the host cores), the roofline of the
grouped GEMM kernel (HIP events around its launch) and the host-CPU baseline.
"""
import argparse
import hashlib
import json
import os
import random
import statistics
import sys
import tempfile
import threading
import time

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

import numpy as np  # noqa: E402

METRIC = "btod_contract2 eff. TFLOP/s @≤1e-12 rel err vs FP64; p50/p99 request latency"
FP64_MFMA_PEAK = 78.6   # TFLOP/s, MI355X dense FP64 matrix


def log(*a):
    print(*a, file=sys.stderr)
    sys.stderr.flush()


def cpu_baseline(prob, a, b, flops, shape, threads):
    """The host-CPU baseline on this machine's cores, and the FP64 reference
    result.

    libtensor itself (oracle/_ref/ref_driver, built from its sources with its
    default linalg_generic loops: the image has no cblas.h for its BLAS
    backend) runs the contraction on the same inputs: its schedule and
    contraction lists, its result (the parity reference) and its own perform
    time (thread_pool of `threads` workers, median of 5 after a warm-up).  The
    timed baseline is the oracle's restatement of libtensor's per-block work
    over those lists with a BLAS GEMM per block pair (oracle/cpu_contract.py:
    numpy's OpenBLAS, one thread per worker, `threads` workers over C blocks
    as libtensor's thread_pool(n, n)), median of 5 after a warm-up, and on
    one core (median of 3)."""
    sys.path.insert(0, os.path.join(HERE, "oracle"))
    import ref_io
    import cpu_contract as CC
    if not ref_io.ref_available():
        return None, None
    with tempfile.TemporaryDirectory() as d:
        t0 = time.time()
        out = ref_io.run_ref(prob, d, shape, tensors={"A": a, "B": b},
                             extra=["--threads", str(threads), "--repeat", "5", "--no-symc"],
                             timeout=1500)
        log("cpu reference done in %.1fs" % (time.time() - t0))
    ref_ms = out.get("perform_median_ms", out.get("perform_ms"))
    clst = out["clst"]
    na = prob["contr"]["N"] + prob["contr"]["K"]
    ia, ib = CC.needed_blocks(clst, na)
    ba = {x: a.get_block(x) for x in ia if x in set(a.req_nonzero_blocks())}
    nzb = set(b.req_nonzero_blocks())
    bb = {x: b.get_block(x) for x in ib if x in nzb}
    bc = CC.BlockContraction(prob, ba, bb, out["schedule"], clst, shape)
    t0 = time.time()
    med, times, blocks = CC.time_perform(bc, threads, 5)
    med1, times1, _ = CC.time_perform(bc, 1, 3)
    err = bc.max_rel_err(blocks, out["result"]) if "result" in out else None
    log("cpu restatement: %.1f ms (%d threads), %.1f ms (1 thread), err %s, %.1fs" % (
        med, threads, med1, err, time.time() - t0))
    info = CC.host_info()
    base = {"value": flops / (med * 1e-3) / 1e12, "unit": "TFLOP/s", "cores": threads,
            "kind": "port",
            "sample": "full pp-ladder contraction: libtensor's schedule and contraction lists, "
                      "per block pair a %s GEMM (numpy.tensordot, single-threaded BLAS), "
                      "%d worker processes over C blocks; median of 5 after 1 warm-up: %.1f ms; "
                      "1 core: %.1f ms (%.4f TFLOP/s); libtensor itself (linalg_generic "
                      "loops, thread_pool of %d): %.1f ms (%.4f TFLOP/s); host %s, nproc %d, "
                      "affinity %d" % (CC.blas_info(), threads, med, med1,
                                       flops / (med1 * 1e-3) / 1e12, threads, ref_ms,
                                       flops / (ref_ms * 1e-3) / 1e12, info["cpu_model"],
                                       info["nproc"], info["affinity"]),
            "one_core_value": flops / (med1 * 1e-3) / 1e12,
            "reference_value": flops / (ref_ms * 1e-3) / 1e12,
            "reference_ms": ref_ms, "port_ms": med, "port_one_core_ms": med1,
            "port_max_rel_err": err, "cpu_model": info["cpu_model"], "nproc": info["nproc"]}
    return base, out.get("result")


def serving_mix(ctx, lt, W, n_clients, rounds, rate, seed=1005):
    """The serving mix: n_clients calculations (o ~ U{8..48}, v ~ U{40..240},
    ring / ladder), each submitting `rounds` contractions with Poisson
    arrivals (aggregate `rate` per second) to one request loop.  Every result
    is compared bit for bit with the same contraction run alone, and, where the
    dense operands are small enough, with the dense FP64 oracle."""
    probs = W.serving_clients(n_clients, seed)
    clients = []
    solo = {}
    t0 = time.time()
    for i, p in enumerate(probs):
        a, b, op = W.build_random(ctx, p, 2000 + i)
        op.set_tile(-1)      # configuration from the padding estimate
        c = W.result_tensor(ctx, p)
        op.perform(c, lt.LT_MODE_GROUPED)
        if op.stats()["flops"] > 1e11:
            # a large client's kernel is timed on its own work list once,
            # before the mix (as a calculation tunes its first iteration)
            op.set_tile(-2)
            op.perform(c, lt.LT_MODE_GROUPED)
        # time of the same perform alone (plan built: one perform, synchronized)
        ctx.sync()
        ts = time.perf_counter()
        op.perform(c, lt.LT_MODE_GROUPED)
        ctx.sync()
        solo[p["name"]] = (time.perf_counter() - ts) * 1e3
        digest = hashlib.sha1(c.to_dense().tobytes()).hexdigest()
        clients.append((p["name"], a, b, op, c, digest, op.stats()["flops"]))
    log("serving: %d clients ready in %.1fs" % (len(clients), time.time() - t0))
    srv = lt.server(ctx, max_batch=64)
    lat, qlat, bad, reqs = [], [], [], []
    lock = threading.Lock()

    def client(i):
        rng = random.Random(seed * 1000 + i)
        name, a, b, op, c, digest, fl = clients[i]
        for r in range(rounds):
            time.sleep(rng.expovariate(rate / len(clients)))
            st = srv.wait(srv.submit(op, c))
            with lock:
                lat.append(st["t_total_ms"])
                qlat.append(st["t_queue_ms"])
                reqs.append((st["t_total_ms"], st["t_queue_ms"], solo[name], fl / 1e9))

    t1 = time.perf_counter()
    th = [threading.Thread(target=client, args=(i,)) for i in range(len(clients))]
    for t in th:
        t.start()
    for t in th:
        t.join()
    wall = time.perf_counter() - t1
    st = srv.stats()
    srv.close()
    for name, a, b, op, c, digest, fl in clients:
        if hashlib.sha1(c.to_dense().tobytes()).hexdigest() != digest:
            bad.append(name)
    # the results against the FP64 oracle (dense contraction of the unfolded
    # inputs, oracle/dense_oracle.py) for every client whose dense operands
    # and dense contraction stay within a bound (the result in the mix equals
    # the one checked, bit for bit, above)
    sys.path.insert(0, os.path.join(HERE, "oracle"))
    import dense_oracle as O
    checked, worst, t0 = 0, 0.0, time.time()
    for (name, a, b, op, c, digest, fl), p in zip(clients, probs):
        na = int(np.prod(p["A"]["dims"]))
        nb = int(np.prod(p["B"]["dims"]))
        nc = int(np.prod(c.to_dense().shape))
        kd = int(np.prod([p["A"]["dims"][ia] for ia, ib in p["contr"]["pairs"]]))
        if max(na, nb) > (1 << 26) or 2.0 * nc * kd > 4e10:
            continue
        ref = O.btod_contract2_dense(p["contr"], a.to_dense(), b.to_dense())
        worst = max(worst, O.rel_err(c.to_dense(), ref))
        checked += 1
    log("serving: %d of %d results checked against the oracle in %.1fs, max rel err %.2e"
        % (checked, len(clients), time.time() - t0, worst))
    flops = rounds * sum(x[6] for x in clients)
    out = {"clients": len(clients), "requests": int(st["requests"]),
           "batches": int(st["batches"]), "mean_batch": st["mean_batch"],
           "max_batch": int(st["max_batch"]), "arrival_rate_per_s": rate,
           "wall_s": wall, "tflops": flops / wall / 1e12,
           "p50_ms": float(np.percentile(lat, 50)), "p99_ms": float(np.percentile(lat, 99)),
           "queue_p50_ms": float(np.percentile(qlat, 50)),
           "queue_p99_ms": float(np.percentile(qlat, 99)),
           "request_gflop_min": min(x[6] for x in clients) / 1e9,
           "request_gflop_max": max(x[6] for x in clients) / 1e9,
           "largest_solo_ms": max(solo.values()),
           "p99_over_largest_solo": float(np.percentile(lat, 99)) / max(solo.values()),
           "slowest": [{"ms": r[0], "queue_ms": r[1], "solo_ms": r[2], "gflop": r[3]}
                       for r in sorted(reqs, reverse=True)[:6]],
           "results_same_as_alone": not bad, "differing": bad,
           "oracle_checked": checked, "oracle_max_rel_err": worst,
           "oracle_note": "dense FP64 oracle for the clients whose dense operands have at "
                          "most 2^26 elements and whose dense contraction is at most 4e10 "
                          "FLOPs; every other result equals its run alone bit for bit"}
    log("serving: %s" % json.dumps(out))
    return out


def block_ops(ctx, lt, prob, a, b, op, c, steps=10):
    """btod_copy of the two-electron tensor V (every canonical block, one
    batched permute launch; HBM-bound: 16 B per element), compute_block of
    the largest scheduled result block (direct tensors), and btod_symmetrize2
    of the contraction over (ij)(ab) (the contraction, then one gather launch
    per source rank; gather HBM-bound), and btod_contract3 of T T T, timed
    on the host clock around synchronous calls and with HIP events around
    the kernels."""
    perm = (1, 0, 2, 3)
    td = prob["B"]
    bis = lt.block_index_space([td["dims"][p] for p in perm])
    for mask, pos in td.get("splits", []):
        bis.split([mask[p] for p in perm], pos)
    vt = lt.block_tensor(ctx, bis)
    elems = sum(int(np.prod(b.block_dims(b.index(x)))) for x in b.req_nonzero_blocks())
    lt.btod_copy(b, vt, perm, c=-1.0)
    dev_ms = []
    t0 = time.perf_counter()
    for i in range(steps):
        lt.btod_copy(b, vt, perm, c=-1.0)
        dev_ms.append(ctx.last_op_ms())
    ms = (time.perf_counter() - t0) * 1e3 / steps
    kms = statistics.mean(dev_ms)
    gbs = 16.0 * elems / (kms * 1e-3) / 1e9
    sched = op.get_schedule()
    big = max(sched, key=lambda x: int(np.prod(c.block_dims(c.index(x)))))
    dims = tuple(c.block_dims(c.index(big)))
    out = lt.device_array(ctx, dims)
    op.compute_block(big, out)
    t0 = time.perf_counter()
    for i in range(steps):
        op.compute_block(big, out)
    cb_ms = (time.perf_counter() - t0) * 1e3 / steps
    out.free()
    # symm(i|j, a|b, contract(...)): B = R + P(R), P = (ij)(ab) (symmetric)
    from libtensor_amd import workloads as W
    sym = lt.btod_symmetrize2(op, (1, 0, 3, 2), True)
    bs = W.result_tensor(ctx, prob)
    sym.perform(bs)
    sv = []
    t0 = time.perf_counter()
    for i in range(steps):
        sym.perform(bs)
        sv.append(sym.stats())
    sy_ms = (time.perf_counter() - t0) * 1e3 / steps
    g_ms = statistics.mean(x["gather_ms"] for x in sv)
    g_gbs = sv[-1]["gather_bytes"] / (g_ms * 1e-3) / 1e9
    symres = {"perm": [1, 0, 3, 2], "symmetric": True, "blocks": len(sym.get_schedule()),
              "gpu_ms": statistics.mean(x["gpu_ms"] for x in sv), "gather_ms": g_ms,
              "gather_bytes": sv[-1]["gather_bytes"], "gather_GB_per_s": g_gbs,
              "gather_hbm_frac": g_gbs / 8000.0, "call_ms": sy_ms,
              "note": "gpu_ms: contraction (grouped) + gather launches, HIP events; "
                      "call_ms: host clock around the synchronous perform"}
    del sym
    # btod_contract3: D[ijab] = 1/4 sum_{klcd} (T[ijcd] T[klcd]) T[klab] (the
    # quadratic ladder term of CCD), the intermediate X[ijkl] resident
    c1 = lt.contraction2(2, 2, 2)
    c1.contract(2, 2)
    c1.contract(3, 3)
    c2 = lt.contraction2(2, 2, 2)
    c2.contract(2, 0)
    c2.contract(3, 1)
    op3 = lt.btod_contract3(c1, c2, a, a, a, kd=0.25)
    d3 = W.result_tensor(ctx, prob)
    op3.perform(d3)
    op3.perform(d3)
    g3 = []
    t0 = time.perf_counter()
    for i in range(steps):
        op3.perform(d3)
        g3.append(op3.stats()["gpu_ms"])
    c3_ms = (time.perf_counter() - t0) * 1e3 / steps
    f3 = op3.stats()["flops"]
    c3res = {"blocks": len(op3.get_schedule()), "flops": f3, "call_ms": c3_ms,
             "gpu_ms": statistics.mean(g3), "tflops": f3 / (c3_ms * 1e-3) / 1e12,
             "plans_built": op3.plans(),
             "note": "both contractions grouped (HIP graphs); gpu_ms: their device time; "
                     "tflops: schedule FLOPs of both over call_ms"}
    del op3
    res = {"btod_copy": {"elements": elems, "kernel_ms": kms, "GB_per_s": gbs,
                         "hbm_frac": gbs / 8000.0, "call_ms": ms, "perm": list(perm),
                         "note": "kernel_ms: HIP events around the permute launch; call_ms: "
                                 "host clock around the synchronous call (orbits, "
                                 "descriptors, upload, launch)"},
           "compute_block": {"block": int(big), "dims": list(dims),
                             "pairs": len(op.clst(big)), "ms": cb_ms},
           "btod_symmetrize2": symres, "btod_contract3": c3res}
    log("block ops: %s" % json.dumps(res))
    return res


def ozaki_sweep(ctx, lt, W, prob, op, ref, flops, a, b, slices=(3, 4, 5, 6, 7), steps=5):
    """Split-precision mode on the benchmark contraction: error against the
    FP64 reference per block and overall, the error bound the library states,
    and time per perform, for each slice count; then 7 slices with the
    integrals V in an immutable tensor (their slices cut once per plan)."""
    out = []
    c = W.result_tensor(ctx, prob)
    sched = op.get_schedule()
    ref_blocks = {}
    for bi in sched:
        idx = c.index(bi)
        st, dims = c.block_start(idx), c.block_dims(idx)
        sl = tuple(slice(a, a + n) for a, n in zip(st, dims))
        ref_blocks[bi] = (sl, ref[sl])
    den = float(np.max(np.abs(ref)))
    for s in slices:
        op.set_ozaki(s)
        op.perform(c, lt.LT_MODE_OZAKI)
        ctx.sync()
        t0 = time.perf_counter()
        for i in range(steps):
            op.perform(c, lt.LT_MODE_OZAKI)
        ctx.sync()
        ms = (time.perf_counter() - t0) * 1e3 / steps
        got = c.to_dense()
        err = float(np.max(np.abs(got - ref)) / den)
        blk = []
        for bi in sched:
            sl, r = ref_blocks[bi]
            m = float(np.max(np.abs(r)))
            blk.append(float(np.max(np.abs(got[sl] - r)) / m) if m > 0 else 0.0)
        st = op.stats()
        out.append({"slices": s, "ms_per_step": ms, "tflops_equiv": flops / (ms * 1e-3) / 1e12,
                    "max_rel_err": err, "stated_rel_err_bound": st["ozaki_max_rel_err"],
                    "block_rel_err_max": max(blk),
                    "block_rel_err_median": float(np.median(blk)),
                    "gemm_ms": st["t_gemm_ms"]})
        log("ozaki: %s" % json.dumps(out[-1]))
    need = [x["slices"] for x in out if x["max_rel_err"] <= 1e-12]
    res = {"sweep": out, "slices_for_1e-12": min(need) if need else None}
    # V immutable: a copy of the integrals marked immutable, a new operation
    # on it; its slices are cut at the plan's first perform only
    vi = W.build_tensor(ctx, prob["B"])
    lt.btod_copy(b, vi)
    vi.set_immutable()
    cc = prob["contr"]
    opi = lt.btod_contract2(W.make_contraction(prob), a, vi, cc.get("ka", 1.0),
                            cc.get("kb", 1.0), cc.get("kc", 1.0))
    opi.set_ozaki(7)
    opi.perform(c, lt.LT_MODE_OZAKI)
    ctx.sync()
    t0 = time.perf_counter()
    for i in range(steps):
        opi.perform(c, lt.LT_MODE_OZAKI)
    ctx.sync()
    ms = (time.perf_counter() - t0) * 1e3 / steps
    err = float(np.max(np.abs(c.to_dense() - ref)) / den)
    res["immutable_v_7"] = {"ms_per_step": ms, "tflops_equiv": flops / (ms * 1e-3) / 1e12,
                            "max_rel_err": err,
                            "stated_rel_err_bound": opi.stats()["ozaki_max_rel_err"],
                            "gemm_ms": opi.stats()["t_gemm_ms"]}
    log("ozaki, V immutable: %s" % json.dumps(res["immutable_v_7"]))
    opi.close()
    del vi
    return res


# name -> (problem builder, input seed, description)
WORKLOADS = {
    "pp-ladder": (lambda W: W.pp_ladder_o40_v200(), 1002,
                  "pp-ladder R[ijab] = sum_cd T[ijcd] V[abcd], o=40 v=200, alpha/beta se_part, "
                  "C2v se_label, (ij)(ab)(cd) antisymmetric, uneven blocks 4-24"),
    "ring-o10-v40": (lambda W: W.ring(10, 40), 1001,
                     "C[ijab] = sum_kc A[ikac] B[kjcb], o=10 v=40, one block, no symmetry"),
    "ring-o60-v300": (lambda W: W.ring_batched_o60_v300(), 1003,
                      "C[ijab] = sum_kc A[ikac] B[kjcb], o=60 v=300, uneven blocks U{8..96} "
                      "(seed 1003), no symmetry"),
}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--mode", default="grouped", choices=["grouped", "per_block"])
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-per-block", action="store_true",
                    help="skip the per-block comparison run (kernel profiles)")
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--no-serve", action="store_true", help="skip the serving mix")
    ap.add_argument("--no-ozaki", action="store_true", help="skip the split-precision sweep")
    ap.add_argument("--copy-probe", action="store_true",
                    help="only run a plain 256 MiB device copy (tod_copy, identity) three "
                         "times: the reference for scaling HBM counters")
    ap.add_argument("--serve-clients", type=int, default=64)
    ap.add_argument("--serve-rounds", type=int, default=3)
    ap.add_argument("--serve-rate", type=float, default=50.0,
                    help="aggregate Poisson arrival rate of the serving mix (1/s)")
    ap.add_argument("--workload", default="pp-ladder", choices=sorted(WORKLOADS),
                    help="pp-ladder (default, the headline line); ring-o10-v40 (one "
                         "block, permuted operands) and ring-o60-v300 (uneven blocks): "
                         "side measurements without the host baseline, the split-precision "
                         "sweep or the serving mix")
    args = ap.parse_args()
    if args.workload != "pp-ladder":
        args.no_cpu = args.no_serve = args.no_ozaki = True
    if args.gpus != 1:
        print("bench.py runs on one GPU only: it measures one GPU's share of the "
              "contraction; multi-GPU sharding is not built", file=sys.stderr)
        sys.exit(2)

    import libtensor_amd as lt
    from libtensor_amd import workloads as W

    ctx = lt.Context(0)
    if args.copy_probe:
        n = 1 << 25
        src = lt.device_array(ctx, (n,), np.ones(n))
        dst = lt.device_array(ctx, (n,))
        for i in range(3):
            lt.tod_copy(src, dst, (0,), 1.0, zero=True)
        print(json.dumps({"copy_probe_bytes_read": n * 8, "copy_probe_bytes_written": n * 8,
                          "launches": 3}))
        return
    make_prob, seed, wl_desc = WORKLOADS[args.workload]
    prob = make_prob(W)
    t0 = time.time()
    a, b, op = W.build_random(ctx, prob, seed)
    c = W.result_tensor(ctx, prob)
    log("inputs ready in %.1fs: %d + %d canonical blocks" % (
        time.time() - t0, len(a.req_nonzero_blocks()), len(b.req_nonzero_blocks())))
    mode = lt.LT_MODE_GROUPED if args.mode == "grouped" else lt.LT_MODE_PER_BLOCK

    for _ in range(max(1, args.warmup)):
        op.perform(c, mode)
    ctx.sync()
    st = op.stats()
    flops = st["flops"]
    log("plan %.1f ms, schedule %.1f ms, %d C blocks, %d pairs, %d tiles, %.3e flops" % (
        st["t_plan_ms"], st["t_schedule_ms"], st["n_cblocks"], st["n_pairs"],
        st["n_tiles"], flops))

    # timed: K performs back to back, each returning once enqueued on the
    # context's stream (set_async), between two synchronizations
    op.set_async(True)
    ctx.sync()
    t0 = time.perf_counter()
    for i in range(args.steps):
        op.perform(c, mode)
    ctx.sync()
    t1 = time.perf_counter()
    op.set_async(False)
    ms_per_step = (t1 - t0) * 1e3 / args.steps
    value = flops / (ms_per_step * 1e-3) / 1e12
    # the same performs one at a time, each waited for (call latency), and
    # their device time (HIP events around the graph launch)
    gpu_ms = []
    t2 = time.perf_counter()
    for i in range(args.steps):
        op.perform(c, mode)
        gpu_ms.append(op.stats()["t_gpu_ms"])
    ms_per_call_sync = (time.perf_counter() - t2) * 1e3 / args.steps

    # The grouped perform runs as a HIP graph; the GEMM launch alone is timed
    # with HIP events on the context's stream in eager performs of the same
    # plan (same kernel, same work list).
    op.set_graph(False)
    gemm_ms = []
    for i in range(max(5, args.steps // 2)):
        op.perform(c, mode)
        gemm_ms.append(op.stats()["t_gemm_ms"])
    op.set_graph(True)
    gemm_avg = statistics.mean(gemm_ms)
    log("step %.3f ms (device %.3f ms per graph launch, gemm %.3f ms)" % (
        ms_per_step, statistics.mean(gpu_ms), gemm_avg))

    # per-block path on the same operation, for comparison (timed the same
    # way: performs enqueued back to back)
    per_block_ms = None
    if not args.no_per_block:
        op.perform(c, lt.LT_MODE_PER_BLOCK)
        op.set_async(True)
        ctx.sync()
        t2 = time.perf_counter()
        nb = max(3, args.steps // 4)
        for i in range(nb):
            op.perform(c, lt.LT_MODE_PER_BLOCK)
        ctx.sync()
        per_block_ms = (time.perf_counter() - t2) * 1e3 / nb
        op.set_async(False)
        op.perform(c, mode)
    result = c.to_dense()

    base, ref = None, None
    if not args.no_cpu:
        try:
            ncpu = len(os.sched_getaffinity(0))
        except AttributeError:
            ncpu = os.cpu_count() or 1
        threads = args.cpu_threads or min(16, ncpu)
        base, ref = cpu_baseline(prob, a, b, flops, result.shape, threads)
    rel = None
    if ref is not None:
        den = float(np.max(np.abs(ref)))
        rel = float(np.max(np.abs(result - ref)) / den) if den > 0 else None
        log("max rel err vs reference: %.3e" % rel)
    if rel is not None and rel > 1e-12:
        value = 0.0   # not FP64-accurate: no effective throughput

    # split-precision sweep against the FP64 reference (the CPU reference
    # when it ran, else this FP64 result)
    oz = None
    if not args.no_ozaki:
        oz = ozaki_sweep(ctx, lt, W, prob, op, ref if ref is not None else result, flops, a, b)
        if ref is None:
            oz["reference"] = "FP64 grouped result on the GPU"
        else:
            oz["reference"] = "libtensor on the host (oracle/_ref)"

    serve = None
    if not args.no_serve:
        serve = serving_mix(ctx, lt, W, args.serve_clients, args.serve_rounds, args.serve_rate)

    # adjacent block operations on the same resident tensors
    blockops = None
    if args.workload == "pp-ladder":
        blockops = block_ops(ctx, lt, prob, a, b, op, c)

    achieved = flops / (gemm_avg * 1e-3) / 1e12
    # HBM bytes per GEMM launch from the committed rocprofv3 counter runs,
    # looked up for the kernel configuration and tile order chosen here
    traffic = None
    tpath = os.path.join(HERE, "profiles", "pp_ladder_o40_v200_traffic.json")
    if args.workload == "pp-ladder" and os.path.exists(tpath):
        with open(tpath) as f:
            tr = json.load(f)
        for e in tr.get("launches", []):
            if e["cfg"] == int(st["tile_cfg"]) and e["order"] == int(st["tile_order"]):
                traffic = e["traffic_bytes"]
    line = {
        "metric": METRIC, "value": value, "unit": "TFLOP/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms_per_step,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64", "data": "synthetic",
        "config": {"workload": wl_desc, "mode": args.mode,
                   "c_blocks": int(st["n_cblocks"]), "block_pairs": int(st["n_pairs"]),
                   "flops_per_step": flops},
        "roofline": {"bound": "mfma", "achieved": achieved, "peak": FP64_MFMA_PEAK,
                     "unit": "TFLOP/s", "frac": achieved / FP64_MFMA_PEAK, "traffic": traffic},
        "cpu_baseline": base,
        "max_rel_err": rel,
        "timed": "steps performs enqueued back to back (set_async), one synchronization "
                 "before and after; ms_per_call_sync: the same performs each waited for",
        "ms_per_call_sync": ms_per_call_sync,
        "device_ms_per_perform": statistics.mean(gpu_ms),
        "per_block_ms_per_step": per_block_ms,
        "gemm_ms": gemm_avg,
        "plan_ms": st["t_plan_ms"], "schedule_ms": st["t_schedule_ms"],
        "gemm_kernel_cfg": int(st["tile_cfg"]), "gemm_tile_order": int(st["tile_order"]),
        "autotune_ms": st["tune_ms"],
        "p50_ms": serve["p50_ms"] if serve else None,
        "p99_ms": serve["p99_ms"] if serve else None,
        "serving": serve,
        "ozaki": oz,
        "block_ops": blockops,
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
