// Request loop for many concurrent calculations on one GPU: block tensors stay
// resident in HBM, contractions are admitted by free device memory, planned
// on a host thread while the GPU runs the previous batch, and a batch runs
// ready requests together (batch_run: one grouped GEMM launch over the union
// of their work lists).  A large request's products run in slices, and
// requests that arrive meanwhile run between two slices (continuous
// batching below the perform).
#include "lt_server.h"

#include <algorithm>
#include <chrono>

namespace lta {

namespace {

double now_ms() {
    using namespace std::chrono;
    return duration<double, std::milli>(steady_clock::now().time_since_epoch()).count();
}

} // unnamed namespace

server::server(context &ctx, size_t max_batch, size_t hbm_reserve) :
    m_ctx(ctx), m_max_batch(max_batch ? max_batch : 64), m_reserve(hbm_reserve),
    m_stop(false), m_next(1), m_batches(0), m_done(0), m_batched_requests(0),
    m_max_seen(0), m_rejected(0) {
    m_ctx.require_device("server");
    m_planner = std::thread([this]() { planner_loop(); });
    m_executor = std::thread([this]() { executor_loop(); });
}

server::~server() {
    {
        std::lock_guard<std::mutex> lk(m_mtx);
        m_stop = true;
    }
    m_cv.notify_all();
    if(m_planner.joinable()) m_planner.join();
    if(m_executor.joinable()) m_executor.join();
}

uint64_t server::submit(btod_contract2 *op, btensor *c, bool add, double d) {
    lt_require(op && c, LT_ERR_PARAM, "null request");
    std::lock_guard<std::mutex> lk(m_mtx);
    lt_require(!m_stop, LT_ERR_PARAM, "server is stopping");
    for(std::map<uint64_t, request>::iterator i = m_reqs.begin(); i != m_reqs.end(); ++i) {
        if(i->second.state == DONE) continue;
        const request &q = i->second;
        lt_require(q.op != op && q.c != c, LT_ERR_PARAM,
            "operation or result tensor already has a request in flight");
        // a request may neither read the result of a request in flight nor
        // write an operand one reads (requests of a batch run their prepare
        // steps before any product)
        lt_require(q.c != &op->operand_a() && q.c != &op->operand_b(), LT_ERR_STATE,
            "an operand is the result tensor of a request in flight");
        lt_require(c != &q.op->operand_a() && c != &q.op->operand_b(), LT_ERR_STATE,
            "the result tensor is an operand of a request in flight");
    }
    request r;
    r.op = op; r.c = c; r.add = add; r.d = d;
    r.state = QUEUED;
    r.status = LT_OK;
    r.t_submit = now_ms();
    r.t_start = r.t_end = 0.0;
    r.batch = 0; r.batch_size = 0;
    uint64_t id = m_next++;
    m_reqs[id] = r;
    m_queue.push_back(id);
    m_cv.notify_all();
    return id;
}

void server::wait(uint64_t id, request_stats &out) {
    std::unique_lock<std::mutex> lk(m_mtx);
    std::map<uint64_t, request>::iterator i = m_reqs.find(id);
    lt_require(i != m_reqs.end(), LT_ERR_PARAM, "unknown request");
    m_cv.wait(lk, [&]() { return i->second.state == DONE; });
    const request &r = i->second;
    out.t_queue_ms = r.t_start - r.t_submit;
    out.t_run_ms = r.t_end - r.t_start;
    out.t_total_ms = r.t_end - r.t_submit;
    out.batch = r.batch;
    out.batch_size = r.batch_size;
    out.status = r.status;
    out.message = r.message;
    m_reqs.erase(i);
}

void server::totals(server_stats &out) {
    std::lock_guard<std::mutex> lk(m_mtx);
    out.batches = m_batches;
    out.requests = m_done;
    out.mean_batch = m_batches ? double(m_batched_requests) / double(m_batches) : 0.0;
    out.max_batch = m_max_seen;
    out.rejected = m_rejected;
}

// Planner: admission by device memory, then the execution plan on the host
// (schedule-derived work lists, device upload, kernel autotuning on the
// first use of an operation).
void server::planner_loop() {
    while(true) {
        uint64_t id = 0;
        {
            std::unique_lock<std::mutex> lk(m_mtx);
            m_cv.wait(lk, [&]() { return m_stop || !m_queue.empty(); });
            if(m_queue.empty()) {
                if(m_stop) break;
                continue;
            }
            id = m_queue.front();
            request &r = m_reqs[id];
            // admission: the operation's plan is resident already, or its
            // estimated footprint fits in the free budget; otherwise wait
            // for running batches to finish, and reject what can never fit
            size_t need = r.op->plan_resident(*r.c, r.add, r.d) ? 0 :
                r.op->footprint_estimate();
            size_t used = m_ctx.used(), cap = m_ctx.cap();
            bool fits = used + need + m_reserve <= cap;
            if(!fits && in_flight()) {
                m_cv.wait_for(lk, std::chrono::milliseconds(1));
                continue;
            }
            m_queue.pop_front();
            r.state = PLANNING;
        }
        request snapshot;
        {
            std::lock_guard<std::mutex> lk(m_mtx);
            snapshot = m_reqs[id];
        }
        int status = LT_OK;
        std::string msg;
        try {
            // no timed autotuning here: the executor shares the stream
            snapshot.op->prepare(*snapshot.c, snapshot.add, snapshot.d, false);
        } catch(const lt_error &e) {
            status = e.code;
            msg = e.what();
        } catch(const std::exception &e) {
            status = LT_ERR_INTERNAL;
            msg = e.what();
        }
        {
            std::lock_guard<std::mutex> lk(m_mtx);
            request &r = m_reqs[id];
            if(status != LT_OK) {
                r.status = status;
                r.message = msg;
                r.t_start = r.t_end = now_ms();
                r.state = DONE;
                m_rejected++;
            } else {
                r.state = READY;
                m_ready.push_back(id);
            }
        }
        m_cv.notify_all();
    }
}

// Executor.  Ready requests of at most kSmallFlops form batches of up to
// max_batch (one grouped launch); a medium request (up to kLargeFlops, a
// few milliseconds) runs as a batch of its own; a large one runs alone with
// its GEMM tiles launched in slices of about kSliceFlops.  Between two
// slices the small requests that became ready meanwhile run as one complete
// batch (prepare, products and write-back), then the oldest ready medium
// request, so neither waits for a large request to end.  A request that
// fails its checks fails alone; the others of its batch complete.
namespace {
const double kSmallFlops = 2e10;
const double kLargeFlops = 2e11;
const double kSliceFlops = 4e10;
}

bool server::take_small(std::vector<uint64_t> &ids, std::vector<batch_item> &items,
    bool small_only) {
    // called with m_mtx held
    double t = now_ms();
    for(std::deque<uint64_t>::iterator i = m_ready.begin();
        i != m_ready.end() && ids.size() < m_max_batch;) {
        request &r = m_reqs[*i];
        const bool large = r.op->flops() > kSmallFlops;     // medium or large: alone
        if(large && (small_only || !ids.empty())) { ++i; continue; }
        r.state = RUNNING;
        r.t_start = t;
        ids.push_back(*i);
        batch_item it;
        it.op = r.op; it.c = r.c; it.add = r.add; it.d = r.d;
        items.push_back(it);
        i = m_ready.erase(i);
        if(large) break;         // a large request runs alone
    }
    return !ids.empty();
}

void server::complete(const std::vector<uint64_t> &ids, const std::vector<batch_item> &items,
    int status, const std::string &msg) {
    std::lock_guard<std::mutex> lk(m_mtx);
    double t = now_ms();
    m_batches++;
    m_batched_requests += ids.size();
    m_max_seen = std::max(m_max_seen, ids.size());
    for(size_t i = 0; i < ids.size(); i++) {
        request &r = m_reqs[ids[i]];
        r.t_end = t;
        r.batch = m_batches;
        r.batch_size = ids.size();
        r.status = status != LT_OK ? status : items[i].status;
        r.message = status != LT_OK ? msg : items[i].message;
        if(r.status != LT_OK) m_rejected++;
        r.state = DONE;
        m_done++;
    }
    m_cv.notify_all();
}

bool server::take_medium(std::vector<uint64_t> &ids, std::vector<batch_item> &items) {
    // called with m_mtx held
    for(std::deque<uint64_t>::iterator i = m_ready.begin(); i != m_ready.end(); ++i) {
        request &r = m_reqs[*i];
        const double f = r.op->flops();
        if(f <= kSmallFlops || f > kLargeFlops) continue;
        r.state = RUNNING;
        r.t_start = now_ms();
        ids.push_back(*i);
        batch_item it;
        it.op = r.op; it.c = r.c; it.add = r.add; it.d = r.d;
        items.push_back(it);
        m_ready.erase(i);
        return true;
    }
    return false;
}

void server::run_small(dev_buffer &side) {
    // the ready small requests as one batch, then the oldest ready medium one
    for(int pass = 0; pass < 2; pass++) {
        std::vector<uint64_t> sids;
        std::vector<batch_item> sitems;
        {
            std::lock_guard<std::mutex> lk(m_mtx);
            if(pass == 0) take_small(sids, sitems, true);
            else take_medium(sids, sitems);
        }
        if(sids.empty()) continue;
        int sst = LT_OK;
        std::string smsg;
        try {
            batch_run srun(sitems, side);
            srun.finish();
        } catch(const lt_error &e) {
            sst = e.code;
            smsg = e.what();
        } catch(const std::exception &e) {
            sst = LT_ERR_INTERNAL;
            smsg = e.what();
        }
        complete(sids, sitems, sst, smsg);
    }
}

void server::run_large(batch_item &it, dev_buffer &side) {
    // the large request alone: its own plan and kernel, products in slices;
    // between two slices the small requests that became ready run as a
    // complete batch.  Large requests run in arrival order (first come,
    // first served: the largest requests set the latency tail, and letting
    // smaller large ones overtake them lengthens it)
    auto between = [&]() { run_small(side); };
    it.op->perform_sliced(*it.c, it.add, it.d, kSliceFlops, between);
}

void server::run_batch(std::vector<uint64_t> &ids, std::vector<batch_item> &items,
    dev_buffer &merged, dev_buffer &side) {
    int status = LT_OK;
    std::string msg;
    try {
        if(items.size() == 1 && items[0].op->flops() > kLargeFlops) {
            batch_item &it = items[0];
            try {
                run_large(it, side);
            } catch(const lt_error &e) {
                it.status = e.code;
                it.message = e.what();
            }
        } else {
            batch_run run(items, merged);
            run.finish();
        }
    } catch(const lt_error &e) {
        status = e.code;
        msg = e.what();
    } catch(const std::exception &e) {
        status = LT_ERR_INTERNAL;
        msg = e.what();
    }
    complete(ids, items, status, msg);
}

void server::executor_loop() {
    dev_buffer merged, side;
    while(true) {
        std::vector<uint64_t> ids;
        std::vector<batch_item> items;
        {
            std::unique_lock<std::mutex> lk(m_mtx);
            m_cv.wait(lk, [&]() {
                return !m_ready.empty() || (m_stop && m_queue.empty() && !planning());
            });
            if(m_ready.empty()) break;
            take_small(ids, items, false);
        }
        run_batch(ids, items, merged, side);
    }
}

bool server::in_flight() const {
    if(!m_ready.empty()) return true;
    for(std::map<uint64_t, request>::const_iterator i = m_reqs.begin(); i != m_reqs.end(); ++i)
        if(i->second.state == RUNNING) return true;
    return false;
}

bool server::planning() const {
    for(std::map<uint64_t, request>::const_iterator i = m_reqs.begin(); i != m_reqs.end(); ++i)
        if(i->second.state == PLANNING) return true;
    return false;
}

} // namespace lta
