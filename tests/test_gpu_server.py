"""Request loop (continuous batching): many calculations submit
contractions concurrently; each result equals the result of the same
contraction run alone, bit for bit when the kernel configuration is fixed
(the arithmetic of a C tile does not depend on what else is in the batch),
and within 1e-12 of the FP64 reference with autotuned configurations."""
import random
import sys
import threading
import time

import numpy as np
import pytest

import dense_oracle as O
import golden_cases as G
import libtensor_amd as lt
import problems as P

pytestmark = pytest.mark.gpu

TOL = 1e-12
CASES = ["ladder_full", "ladder_spin", "rand_221", "rand_222", "randsym_222", "rand_112",
         "ring_small", "self_ladder", "rand_311", "randsym_122"]


def _say(*a):
    print(*a)
    sys.stdout.flush()


def _clients(ctx, cfg):
    out = []
    for name in CASES:
        prob, dense, meta, result = G.load(name)
        a, b, op = P.build(ctx, prob, dense)
        if cfg is not None:
            op.set_tile(cfg)
        out.append((name, prob, a, b, op, result))
    return out


@pytest.mark.parametrize("cfg", [5, None])
def test_results_same_alone_and_in_the_mix(gpu_ctx, cfg):
    clients = _clients(gpu_ctx, cfg)
    alone = []
    for name, prob, a, b, op, result in clients:
        c = P.make_result_tensor(gpu_ctx, prob)
        op.perform(c, lt.LT_MODE_GROUPED)
        alone.append(c.to_dense())
        assert O.rel_err(alone[-1], result) < TOL, name
    srv = lt.server(gpu_ctx, max_batch=16)
    outs = [P.make_result_tensor(gpu_ctx, prob) for _, prob, _, _, _, _ in clients]
    lat = []
    errors = []
    rounds = 3

    def client(i):
        rng = random.Random(100 + i)
        name, prob, a, b, op, result = clients[i]
        try:
            for r in range(rounds):
                time.sleep(rng.expovariate(200.0))       # Poisson arrivals
                t = srv.submit(op, outs[i])
                st = srv.wait(t)
                lat.append(st["t_total_ms"])
                got = outs[i].to_dense()
                if cfg is not None and not np.array_equal(got, alone[i]):
                    errors.append((name, r, "differs from alone"))
                if O.rel_err(got, result) >= TOL:
                    errors.append((name, r, O.rel_err(got, result)))
        except Exception as e:       # reported below
            errors.append((name, repr(e)))

    th = [threading.Thread(target=client, args=(i,)) for i in range(len(clients))]
    for t in th:
        t.start()
    for t in th:
        t.join()
    st = srv.stats()
    srv.close()
    _say("server", cfg, st, "p50 %.3f ms p99 %.3f ms" % (np.percentile(lat, 50),
                                                         np.percentile(lat, 99)))
    assert not errors, errors
    assert st["requests"] == rounds * len(clients)


def test_accumulating_requests(gpu_ctx):
    """perform(btc, d) through the loop: C = r, then C += 0.5 r twice."""
    prob, dense, meta, result = G.load("ladder_full")
    a, b, op = P.build(gpu_ctx, prob, dense)
    c = P.make_result_tensor(gpu_ctx, prob)
    srv = lt.server(gpu_ctx)
    srv.wait(srv.submit(op, c))
    srv.wait(srv.submit(op, c, add=True, d=0.5))
    srv.wait(srv.submit(op, c, add=True, d=0.5))
    srv.close()
    assert O.rel_err(c.to_dense(), 2.0 * result) < TOL


def test_one_request_in_flight_per_tensor(gpu_ctx):
    prob, dense, meta, result = G.load("rand_222")
    a, b, op = P.build(gpu_ctx, prob, dense)
    c = P.make_result_tensor(gpu_ctx, prob)
    srv = lt.server(gpu_ctx)
    t = srv.submit(op, c)
    try:
        with pytest.raises(lt.LtError):
            srv.submit(op, c)
    finally:
        srv.wait(t)
        srv.close()
    assert O.rel_err(c.to_dense(), result) < TOL


def test_invalid_request_fails_alone(gpu_ctx):
    """A request with a host-side error (a result tensor of another block
    index space; a += into a tensor of another symmetry, which the request
    loop does not merge) fails with its own error while the requests
    submitted with it complete correctly."""
    clients = _clients(gpu_ctx, None)[:6]
    outs = [P.make_result_tensor(gpu_ctx, prob) for _, prob, _, _, _, _ in clients]
    prob, dense, meta, result = G.load("ladder_full")
    a, b, op = P.build(gpu_ctx, prob, dense)
    wrong = P.make_result_tensor(gpu_ctx, G.load("rand_222")[0])
    from libtensor_amd import workloads as W
    other = W.build_tensor(gpu_ctx, G.load_add("add_ladder_none")[0]["C"])   # no symmetry
    srv = lt.server(gpu_ctx, max_batch=16)
    try:
        tickets = [srv.submit(op_, outs[i]) for i, (_, _, _, _, op_, _) in enumerate(clients)]
        bad = srv.submit(op, wrong)
        for t in tickets:
            srv.wait(t)
        with pytest.raises(lt.LtError) as e:
            srv.wait(bad)
        assert e.value.code == -2          # LT_ERR_DIMS
        c2 = P.make_result_tensor(gpu_ctx, prob)
        srv.wait(srv.submit(op, c2))
        bad2 = srv.submit(op, other, add=True, d=1.0)
        t3 = [srv.submit(op_, outs[i]) for i, (_, _, _, _, op_, _) in enumerate(clients)]
        with pytest.raises(lt.LtError) as e:
            srv.wait(bad2)
        assert e.value.code == -3          # LT_ERR_SYMMETRY
        for t in t3:
            srv.wait(t)
    finally:
        srv.close()
    for i, (name, prob_i, a_i, b_i, op_i, res_i) in enumerate(clients):
        assert O.rel_err(outs[i].to_dense(), res_i) < TOL, name
    assert O.rel_err(c2.to_dense(), result) < TOL


def test_small_requests_run_while_a_large_one_is_on_the_gpu(gpu_ctx):
    """A large request's products run in slices; small requests submitted
    while it runs complete before it (they run between two slices), and all
    results are right."""
    from libtensor_amd import workloads as W
    big = W.ring_batched_o60_v300()
    a, b, op = W.build_random(gpu_ctx, big, 1003)
    cbig = W.result_tensor(gpu_ctx, big)
    op.set_tile(-1)
    op.perform(cbig, lt.LT_MODE_GROUPED)
    ref_big = cbig.to_dense()
    clients = _clients(gpu_ctx, None)[:4]
    outs = [P.make_result_tensor(gpu_ctx, prob) for _, prob, _, _, _, _ in clients]
    srv = lt.server(gpu_ctx, max_batch=16)
    done = {}
    try:
        def run_big():
            t = srv.submit(op, cbig)
            st = srv.wait(t)
            done["big"] = (time.perf_counter(), st)
        th = threading.Thread(target=run_big)
        th.start()
        time.sleep(0.05)
        small_end = []
        for i, (name, prob, _, _, op_i, res) in enumerate(clients):
            st = srv.wait(srv.submit(op_i, outs[i]))
            small_end.append((time.perf_counter(), st))
        th.join()
    finally:
        srv.close()
    t_big, st_big = done["big"]
    _say("large %.1f ms; small %s ms" % (st_big["t_total_ms"],
                                          ["%.2f" % s["t_total_ms"] for _, s in small_end]))
    assert np.array_equal(cbig.to_dense(), ref_big)
    for i, (name, prob, _, _, op_i, res) in enumerate(clients):
        assert O.rel_err(outs[i].to_dense(), res) < TOL, name
    assert all(t < t_big for t, _ in small_end)
    assert max(s["t_total_ms"] for _, s in small_end) < 0.5 * st_big["t_total_ms"]
