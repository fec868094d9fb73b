// Request loop of the block-contraction path (continuous batching of
// btod_contract2 performs from many calculations on one GPU).
#pragma once

#include "lt_runtime.h"

#include <condition_variable>
#include <deque>
#include <map>
#include <mutex>
#include <string>
#include <thread>

namespace lta {

struct request_stats {
    double t_queue_ms;     //!< submit -> start of its batch
    double t_run_ms;       //!< start -> end of its batch
    double t_total_ms;     //!< submit -> result ready
    uint64_t batch;        //!< batch number (1-based)
    size_t batch_size;     //!< requests in that batch
    int status;            //!< LT_OK or the error code
    std::string message;
};

struct server_stats {
    uint64_t batches, requests, rejected;
    double mean_batch;
    size_t max_batch;
};

class server {
public:
    //! max_batch: most requests per batch; hbm_reserve: device memory kept
    //! free when admitting a request whose plan is not resident yet.
    server(context &ctx, size_t max_batch, size_t hbm_reserve);
    ~server();   //!< drains every submitted request

    //! C := result (add false) or C += d * result.  One request in flight per
    //! operation and per result tensor; neither may be touched until wait().
    uint64_t submit(btod_contract2 *op, btensor *c, bool add, double d);
    void wait(uint64_t id, request_stats &out);
    void totals(server_stats &out);

private:
    enum state_t { QUEUED, PLANNING, READY, RUNNING, DONE };
    struct request {
        btod_contract2 *op;
        btensor *c;
        bool add;
        double d;
        state_t state;
        int status;
        std::string message;
        double t_submit, t_start, t_end;
        uint64_t batch;
        size_t batch_size;
    };

    context &m_ctx;
    size_t m_max_batch, m_reserve;
    std::mutex m_mtx;
    std::condition_variable m_cv;
    bool m_stop;
    uint64_t m_next;
    std::map<uint64_t, request> m_reqs;
    std::deque<uint64_t> m_queue, m_ready;
    uint64_t m_batches, m_done, m_batched_requests;
    size_t m_max_seen;
    uint64_t m_rejected;
    std::thread m_planner, m_executor;

    void planner_loop();
    void executor_loop();
    //! Moves ready requests into a batch (small ones up to max_batch, or
    //! one medium or large request alone; small_only: small ones only).
    //! m_mtx held.
    bool take_small(std::vector<uint64_t> &ids, std::vector<batch_item> &items,
        bool small_only);
    void run_batch(std::vector<uint64_t> &ids, std::vector<batch_item> &items,
        dev_buffer &merged, dev_buffer &side);
    //! The oldest ready medium request (between the small and the large
    //! ones in size), alone.  m_mtx held.
    bool take_medium(std::vector<uint64_t> &ids, std::vector<batch_item> &items);
    //! The ready small requests as one batch, then the oldest ready medium
    //! request (between two slices of a large one).
    void run_small(dev_buffer &side);
    //! A large request in slices, small batches between the slices.
    void run_large(batch_item &it, dev_buffer &side);
    void complete(const std::vector<uint64_t> &ids, const std::vector<batch_item> &items,
        int status, const std::string &msg);
    bool planning() const;
    bool in_flight() const;
};

} // namespace lta
