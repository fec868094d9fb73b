// Device runtime of the block-contraction path: one context per GPU (stream,
// HBM accounting with a hard cap used for admission), HBM-resident block
// tensors, and the btod_contract2 operation with its cached execution plan.
#pragma once

#include <atomic>
#include <functional>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <vector>
#include <hip/hip_runtime_api.h>
#include "lt_core.h"
#include "lt_sym.h"
#include "lt_sched.h"
#include "../device/lt_device.h"

namespace lta {

void hip_check(hipError_t e, const char *what);

//! Device context: one HIP device, one stream, an HBM budget.
class context {
public:
    //! device < 0 creates a host-only context: blocks live in host memory,
    //! schedules and symmetry can be computed, device operations refuse.
    context(int device, size_t hbm_cap);
    ~context();

    int device() const { return m_device; }
    bool is_device() const { return m_device >= 0; }
    void require_device(const char *what) const;
    hipStream_t stream() const { return m_stream; }

    //! Allocates device memory within the budget; throws LT_ERR_NOMEM when
    //! the request does not fit (callers use this for admission).
    void *alloc(size_t bytes);
    void free(void *p);
    bool fits(size_t bytes) const { return m_used + bytes <= m_cap; }
    size_t used() const { return m_used; }
    size_t cap() const { return m_cap; }
    //! New device memory budget (allocations beyond it fail with
    //! LT_ERR_NOMEM; memory in use is not released).
    void set_cap(size_t cap) { std::lock_guard<std::mutex> lk(m_mtx); m_cap = cap; }

    //! Host-to-device copy, ordered after every operation already enqueued
    //! on the compute stream (which may still read or write dst), complete
    //! on return.
    void upload(void *dst, const void *src, size_t bytes);
    //! Host-to-device copy into memory no enqueued operation uses (allocated
    //! by the caller since): not ordered after the compute stream, so a
    //! host thread that plans does not wait for another thread's kernels.
    void upload_new(void *dst, const void *src, size_t bytes);
    //! Device-to-host copy on the compute stream (after everything enqueued
    //! there), complete on return.
    void download(void *dst, const void *src, size_t bytes);
    void sync();
    //! Makes this context's device current on the calling thread.
    void activate() const;

    size_t nthreads() const { return m_nthreads; }
    void set_nthreads(size_t n) { m_nthreads = n ? n : 1; }

    //! Device time (HIP events) of the last block operation's kernels
    //! (btod_copy, btod_add, btod_symmetrize2) on this context.
    double last_op_ms() const { return m_last_op_ms; }
    void set_last_op_ms(double ms) { m_last_op_ms = ms; }

    //! Device staging area for the descriptor lists of block operations
    //! (grown on demand, released with the context); hold staging_lock()
    //! from the request until the work that reads it has completed.
    void *staging(size_t bytes);
    std::mutex &staging_lock() { return m_stage_mtx; }

    //! Uploaded descriptor lists of block operations, kept per key (the
    //! operation, its operands' identities and structure versions, its
    //! scalars): a repeated call launches without rebuilding them.  At most
    //! kLaunchCache entries, the oldest dropped first.
    struct cached_launch {
        //! one permute launch: descriptors at o_d, work items at o_cd / o_co
        struct segment { size_t o_d, o_cd, o_co, nchunks; };
        void *mem;
        std::vector<segment> segs;       //!< launched in order
        std::vector<size_t> touched;     //!< blocks the launches write
        symmetry sym;                    //!< symmetry the result takes
        double bytes;                    //!< bytes the launches move
        cached_launch() : mem(0), bytes(0.0) { }
    };
    std::shared_ptr<cached_launch> find_launch(const std::string &key);
    void keep_launch(const std::string &key, std::shared_ptr<cached_launch> l);
    static const size_t kLaunchCache = 32;

    //! Side streams (default priority, created on first use) for independent
    //! launches: side_fork makes them wait for the work enqueued on s so
    //! far, side_join makes s wait for the work enqueued on them since.
    //! Hold side_lock() from the fork to the join.
    static const int kSideStreams = 4;
    void side_fork(hipStream_t s);
    void side_join(hipStream_t s);
    hipStream_t side(int i) const { return m_side[i % kSideStreams]; }
    std::mutex &side_lock() { return m_side_mtx; }

private:
    int m_device;
    hipStream_t m_stream;
    hipStream_t m_xfer;   //!< host-to-device copies (default priority)
    hipEvent_t m_xfer_ev; //!< orders ordered uploads after the compute stream
    std::mutex m_xfer_mtx;
    std::mutex m_cache_mtx;
    std::map<std::string, std::shared_ptr<cached_launch> > m_launch_cache;
    std::vector<std::string> m_launch_order;
    size_t m_cap;
    std::atomic<size_t> m_used;
    std::mutex m_mtx;
    std::map<void*, size_t> m_sizes;
    size_t m_nthreads;
    double m_last_op_ms = 0.0;
    void *m_stage = nullptr;
    size_t m_stage_bytes = 0;
    std::mutex m_stage_mtx;
    hipStream_t m_side[kSideStreams] = {};
    hipEvent_t m_side_ev[kSideStreams] = {};
    hipEvent_t m_fork_ev = nullptr;
    std::mutex m_side_mtx;
};

//! Slab allocator for the blocks of one tensor (few hipMallocs, stable
//! block addresses).
class slab {
public:
    explicit slab(context &ctx) : m_ctx(ctx), m_off(0), m_cap(0) { }
    ~slab();
    double *take(size_t n);
    size_t bytes() const;
    void release_all();
    //! True if p lies in one of the slab's chunks.
    bool owns(const void *p) const;
private:
    context &m_ctx;
    std::vector<std::pair<double*, size_t> > m_chunks;
    size_t m_off, m_cap;
};

//! Block tensor with HBM-resident canonical blocks
//! (libtensor/block_tensor/block_tensor.h semantics: only canonical,
//! allowed, non-zero blocks are stored).
class btensor {
public:
    btensor(context &ctx, const bis_t &bis);

    context &ctx() { return m_ctx; }
    const bis_t &bis() const { return m_sym.bis; }
    const symmetry &sym() const { return m_sym; }
    symmetry &sym_rw();
    dims_t bidims() const { return m_bidims; }

    //! Device storage of a block (allocated on first use; stays at the same
    //! address for the lifetime of the tensor).
    double *block_ptr(size_t aidx);
    bool is_nonzero(size_t aidx) const;
    void mark_nonzero(size_t aidx);
    void zero_block(size_t aidx);
    void zero_all();
    std::vector<size_t> nonzero_blocks() const;
    size_t block_size(size_t aidx) const;

    void set_block(size_t aidx, const double *host);
    void get_block(size_t aidx, double *host);

    //! Structure version: changes when the set of non-zero blocks, the
    //! symmetry or the split points change.
    uint64_t version() const { return m_version; }
    void touch() { m_version = s_version++; }
    //! Identity of the tensor object, never reused within the process
    //! (execution plans are keyed on it rather than on the address).
    uint64_t uid() const { return m_uid; }

    void split(const bool *mask, size_t pos);
    //! Sets the symmetry; tag names the operation that set it (0: the caller).
    void set_symmetry(const symmetry &s, uint64_t tag = 0);
    //! The operation that set the symmetry last (0: the caller, or changed
    //! since through sym_rw / split).
    uint64_t sym_tag() const { return m_sym_tag; }

    //! Fills all canonical allowed blocks with random data symmetrized over
    //! each block's stabilizer (btod_random semantics).
    void fill_random(uint64_t seed, int dist, double scale, int sk_d0, int sk_d1);

    //! block_tensor::set_immutable (libtensor/core/immutable.h): from now on
    //! the tensor's structure and data do not change (every write is
    //! refused with LT_ERR_STATE), so derived data may be cached.
    void set_immutable() { m_immutable = true; }
    bool is_immutable() const { return m_immutable; }
    //! Throws LT_ERR_STATE if the tensor is immutable.
    void require_mutable(const char *what) const;
    //! True if p points into the tensor's block storage.
    bool owns(const void *p) const { return m_slab.owns(p); }

private:
    context &m_ctx;
    symmetry m_sym;
    dims_t m_bidims;
    slab m_slab;
    std::map<size_t, double*> m_store;
    std::vector<size_t> m_nonzero;   // sorted
    uint64_t m_version, m_uid;
    uint64_t m_sym_tag = 0;
    bool m_immutable = false;
    static std::atomic<uint64_t> s_version;
};

//! Stabilizer of a block: transformations g (perm, scalar) with g(X) = X.
std::vector<tensor_tr> block_stabilizer(const symmetry &sym, size_t aidx);

//! Applies a tensor transformation to a dense block on the host
//! (dst = c * P(src), dims of dst = P(dims of src)).
void host_transform(const double *src, const dims_t &dsrc, const tensor_tr &tr,
    double *dst, bool accumulate);

//! Work items (chunks, or tiles of transposing descriptors) of permute
//! descriptors for dev::launch_permute.
void make_chunks(const std::vector<dev::perm_desc> &ds, std::vector<uint32_t> &cd,
    std::vector<uint32_t> &co);

//! Device permute descriptor for dst = c * P(src) (dimensions fused).
dev::perm_desc make_perm_desc_public(const double *src, const dims_t &dsrc,
    const tensor_tr &tr, double *dst, bool accumulate);

//! Lowers b's symmetry to the symmetry of a sum of b and a term of symmetry
//! `other` (sum_symmetry): canonical blocks of the new symmetry that b held
//! only through an orbit of the old one are written out on the device.
//! No-op when the orbits stay the same.  ms (optional) accumulates the
//! device time.
void merge_symmetry(btensor &b, const symmetry &other, double *ms = 0);

//! b += d tr(s) for a resident tensor s of any symmetry whose orbits, after
//! tr, contain b's (b already merged): every non-zero canonical block of s
//! goes to each canonical block of b in its orbit (one permute launch).
void add_into(btensor &b, btensor &s, const tensor_tr &tr, double d, double *ms = 0);

//! btod_dotprod(A, tra, B, trb).calculate() on resident block tensors.
double btod_dotprod(btensor &a, const tensor_tr &tra, btensor &b, const tensor_tr &trb);

//! btod_copy(a, tra).perform(b) (add false: B := tra(A), B takes the permuted
//! symmetry of A) or perform(b, d) (add: B += d tra(A), B with that symmetry).
void btod_copy(btensor &a, const tensor_tr &tra, btensor &b, bool add, double d);

//! btod_add: B := sum_i trs[i](ops[i]) (add false; B takes the symmetry all
//! permuted operands have) or B += d sum_i trs[i](ops[i]) (add; B lowered to
//! the symmetry it shares with them first).
void btod_add(const std::vector<btensor*> &ops, const std::vector<tensor_tr> &trs,
    btensor &b, bool add, double d);

//! A new identity for an operation (tags the symmetry it sets on the
//! tensors it writes).
uint64_t next_op_uid();

//! Same canonical block, transformation and allowedness for every block.
bool same_orbits(const symmetry &x, const symmetry &y, const dims_t &bid);

//! Runs a list of permute descriptors as one launch of the batched permute
//! kernel on the context's stream and waits for it; adds the launch's device
//! time (HIP events) to *ms when ms is not null.
void run_permutes(context &ctx, const std::vector<dev::perm_desc> &descs, double *ms = 0);
//! Descriptor lists (one launch each, in order) uploaded into one device
//! buffer, for the context's launch cache.
std::shared_ptr<context::cached_launch> upload_launches(context &ctx,
    const std::vector<std::vector<dev::perm_desc> > &lists);
//! Runs the launches of a cached entry in order and waits; ms (optional)
//! accumulates their device time.
void run_launches(context &ctx, const context::cached_launch &l, double *ms);
//! Launch-cache key parts: a transformation; a tensor's identity and (with
//! version) its structure version.
std::string key_part(const tensor_tr &tr);
std::string key_tensor(const btensor &t, bool version);

enum exec_mode { MODE_PER_BLOCK = 0, MODE_GROUPED = 1, MODE_OZAKI = 2 };

struct exec_stats {
    double t_schedule_ms;   //!< make_schedule (constructor)
    double t_plan_ms;       //!< work-list build on the first perform
    double t_gpu_ms;        //!< device time of the last perform (events)
    double t_gemm_ms;       //!< device time of the GEMM phase of the last perform
    double flops;           //!< 2 |C_blk| K_blk summed over all pairs
    double bytes_perm;      //!< permute traffic of the last perform (bytes)
    size_t n_cblocks, n_pairs, n_tiles, n_unfold, n_launches;
    int tile_cfg;
    double ozaki_max_rel_err; //!< split-precision: stated bound of max |error| / max |C|
    int ozaki_slices;
    double tune_ms[dev::TILE_COUNT]; //!< autotuner timings (0 = not measured)
    int graph;              //!< 1 if the last perform ran as a HIP graph
    int tile_order;         //!< grouped tile order (0 longest first, 1 XCD groups)
    size_t n_plans;         //!< execution plans built so far
    size_t n_batches;       //!< batches of C blocks of the last perform
};

//! Dense helpers on a context's stream (both synchronous): dst = or +=
//! tr(src) (tod_copy), and C = or += d contract(A, B) of device arrays
//! (tod_contract2: align, permute into scratch, one GEMM, permute back).
void dense_permute(context &ctx, const double *src, const dims_t &dsrc,
    const tensor_tr &tr, double *dst, bool accumulate);
void dense_contract2(context &cx, const contraction2 &contr, const dims_t &da, const double *a,
    const dims_t &db, const double *b, double d, bool zero, double *c);
//! One term of dense_contract2_many: C = (zero) or += d contract(A, B).
struct dense_contract_item {
    contraction2 contr;
    dims_t da, db;
    const double *a, *b;
    double d;
    bool zero;
};
//! The terms into one C in order, as that many dense_contract2 calls, with
//! one allocation, one upload and one synchronization for all of them.
void dense_contract2_many(context &cx, const std::vector<dense_contract_item> &items,
    double *c);

struct contract_plan;
class btod_contract2;

//! Device buffer owned by a caller and reused across calls.
struct dev_buffer {
    context *ctx;
    void *mem;
    size_t bytes;
    dev_buffer() : ctx(0), mem(0), bytes(0) { }
    dev_buffer(const dev_buffer&) = delete;
    dev_buffer &operator=(const dev_buffer&) = delete;
    ~dev_buffer() { release(); }
    void release() { if(ctx && mem) ctx->free(mem); mem = 0; bytes = 0; }
};

//! One perform of a batch (btod_contract2::perform_batch, batch_run).
struct batch_item {
    btod_contract2 *op;
    btensor *c;
    bool add;                 //!< false: C := result; true: C += d * result
    double d;
    double batch_gpu_ms;      //!< out: device time of the whole batch
    size_t batch_gemm_launches;
    int status;               //!< out: LT_OK, or the error of this item alone
    std::string message;
    batch_item() : op(0), c(0), add(false), d(1.0), batch_gpu_ms(0.0),
        batch_gemm_launches(0), status(LT_OK) { }
};

struct contract_plan;

//! One batch of performs on the context's stream, run in steps so that
//! other work can be interleaved between its GEMM slices.  The constructor
//! checks and plans every item on its own (an item that fails is left out of
//! the batch with its status and message set; the others go on), launches
//! the result preparation and prepare permutes and uploads the merged work
//! lists; step() launches the next slice of the grouped GEMM tiles (tiles
//! in the order longest first, at least one tile and at most `flop_budget`
//! FLOPs beyond it); finish() launches the write-backs and waits for the
//! stream.  Every C tile is computed whole by one launch, so the results do
//! not depend on how the tiles are sliced.
class batch_run {
public:
    batch_run(std::vector<batch_item> &items, dev_buffer &merged);
    bool step(double flop_budget);
    void finish();
    double flops() const { return m_flops; }
    size_t live() const { return m_live.size(); }

private:
    struct launch { int cfg; size_t ocb, opr, otl; uint32_t nt; std::vector<double> tcost; };
    std::vector<batch_item> &m_items;
    std::vector<size_t> m_live;
    std::vector<contract_plan*> m_plans;
    std::vector<launch> m_launches;
    context *m_ctx;
    char *m_base;
    size_t m_launch, m_tile;
    double m_flops;
    bool m_finished;
};

//! btod_contract2<N, M, K> (libtensor/block_tensor/btod_contract2.h:29):
//! C = kc * contract(ka * A, kb * B).
class btod_contract2 {
public:
    btod_contract2(const contraction2 &c, btensor &a, double ka, btensor &b,
        double kb, double kc);
    ~btod_contract2();

    const symmetry &symc() const { return m_symc; }
    const std::vector<size_t> &schedule() const { return m_sched; }
    const btensor &operand_a() const { return m_a; }
    const btensor &operand_b() const { return m_b; }

    //! C := result (gen_bto_aux_copy): zero C, take the result symmetry,
    //! write every scheduled block.
    void perform(btensor &c, int mode);
    //! C += d * result (gen_bto_aux_add) for C with the result's symmetry.
    void perform_add(btensor &c, double d, int mode);
    //! One block computed directly on the host from canonical blocks
    //! (gen_bto_contract2::compute_block, reference-order numerics on CPU).
    void compute_block_host(const index_t &ic, double *out);
    //! One block computed directly on the device from the resident canonical
    //! blocks (gen_bto_contract2::compute_block, direct block tensors):
    //! blk (=, or += unless zero) trc(C_ic), blk a device array in the layout
    //! of the permuted block.
    void compute_block(bool zero, const index_t &ic, const tensor_tr &trc, double *blk);
    //! Contraction list of a C block in the original index space.
    void clst(const index_t &ic, contr_list &out);
    //! Runs several performs (grouped mode) as one batch on their common
    //! context: every prepare permute, then one grouped GEMM over the union
    //! of the work lists per operand layout and kernel configuration, then
    //! every write-back.  `merged` holds the concatenated lists.
    static void perform_batch(std::vector<batch_item> &items, dev_buffer &merged);
    friend class batch_run;
    //! perform (add false) or perform_add in grouped mode with the plan's own
    //! kernel, the products launched in slices of about slice_flops FLOPs;
    //! `between` runs on the calling thread after each slice has finished
    //! (other work on the same stream goes in there).
    void perform_sliced(btensor &c, bool add, double d, double slice_flops,
        const std::function<void()> &between);
    //! Builds the execution plan for C unless it is cached (device memory
    //! for work lists and operand scratch, kernel autotuning).  measure
    //! false: the kernel configuration is chosen by the padding estimate
    //! instead of timed launches (callers whose stream is shared with other
    //! work, where timings would include that work).
    void prepare(btensor &c, bool add, double d, bool measure = true);
    bool plan_resident(btensor &c, bool add, double d) const {
        return plan_valid(c, add, d, MODE_GROUPED);
    }
    //! Device memory a plan may take: four orientations of every stored A
    //! and B block plus the scheduled C blocks (admission estimate).
    size_t footprint_estimate();

    //! Statistics of the last perform (device times wait for an
    //! asynchronous perform to end).
    const exec_stats &stats() { finish_times(); return m_stats; }
    //! Stated error of the last split-precision perform, per C block.
    struct oz_bound { size_t aidx; double bound, cmax; };
    const std::vector<oz_bound> &ozaki_bounds() const { return m_oz_bounds; }
    //! Kernel configuration of the next plan (a plan built with another one
    //! is dropped, so the next perform uses it).
    void set_tile_cfg(int cfg);
    //! Grouped performs run as a captured HIP graph (default) or eagerly.
    void set_graph(bool on) { m_use_graph = on; }
    //! on: FP64 performs return once their work is enqueued on the context's
    //! stream (later operations on the context are ordered after it; reads
    //! of the result synchronize); off (default): they return when done.
    void set_async(bool on) { m_async = on; }
    void set_ozaki(int slices) { m_oz_slices = slices; }
    //! split-precision tile shape (dev::oz_tile_cfg), -1 = measured per plan
    void set_ozaki_tile(int tile) { m_oz_tile = tile; }
    double flops() const { return m_stats.flops; }

private:
    contraction2 m_contr;
    btensor &m_a, &m_b;
    double m_ka, m_kb, m_kc;
    uint64_t m_uid;         //!< tags the symmetry of the result tensors it writes
    symmetry m_symc;
    std::vector<size_t> m_sched;
    exec_stats m_stats;
    int m_tile_cfg;         //!< >= 0 fixed, -1 estimate, -2 autotune
    int m_tuned_cfg;        //!< configuration found by the autotuner
    int m_tuned_order;      //!< and its grouped-launch tile order
    uint64_t m_tuned_va, m_tuned_vb;   //!< operand versions it was found for
    bool m_use_graph;
    bool m_async;
    //! an asynchronous perform's device times are read from the plan's
    //! events when the statistics are asked for
    bool m_times_pending, m_times_graph;
    void finish_times();
    //! Drops the plan (its device memory; its pending times are lost).
    void drop_plan();
    int m_oz_slices, m_oz_tile;
    std::unique_ptr<contract_plan> m_plan;
    std::vector<oz_bound> m_oz_bounds;
    //! += into a tensor whose orbits differ from the result's: the result
    //! goes to this resident temporary first
    std::unique_ptr<btensor> m_addtmp;
    uint64_t m_add_uid, m_add_ver;   //!< tensor (uid, version) last checked
    bool m_add_direct;               //!< ... and whether its orbits agree
    //! The batch of the schedule the next plan is built for (null: all).
    const std::vector<size_t> *m_sched_subset;
    bool m_use_graph_saved;
    //! compute_block: the contraction terms of each C block asked for (by
    //! absolute block index), valid for the operand structure versions
    //! m_cb_va / m_cb_vb
    std::mutex m_cb_mtx;
    uint64_t m_cb_va = 0, m_cb_vb = 0;
    std::map<size_t, std::vector<dense_contract_item> > m_cb_terms;

    //! Builds the plan; false (and no plan) if it does not fit the device
    //! memory budget.
    bool try_plan(btensor &c, bool add, double d, bool measure = true, int mode = 1);
    //! The split-precision form of the plan for the current slice count and
    //! tile setting (built when missing).
    void ensure_ozaki();
    //! Perform in batches of C blocks whose plans fit the budget.
    void run_batched(btensor &c, bool add, double d, int mode);

    void check_block_sizes() const;
    //! True if c's orbits are the result's (the result adds into c's
    //! canonical blocks directly); cached per (uid, structure version).
    bool add_direct(btensor &c);
    void build_plan(btensor &c, bool add, double d, bool measure = true);
    bool plan_valid(btensor &c, bool add, double d, int mode) const;
    void run(btensor &c, int mode);
    void enqueue(int mode, hipStream_t s, bool timed, size_t &launches);
    void build_ozaki(int slices);
};

//! Operation whose result can be written into a block tensor or computed one
//! block at a time (libtensor's additive_gen_bto, the operand of
//! btod_symmetrize2: gen_block_tensor/additive_gen_bto.h:26-68).
class additive_op {
public:
    virtual ~additive_op() { }
    virtual context &ctx() = 0;
    virtual const symmetry &sym() const = 0;
    virtual const std::vector<size_t> &schedule() const = 0;
    //! T := result (T zeroed, given the result symmetry); returns the
    //! device time of the kernels in ms.
    virtual double perform_into(btensor &t, int mode) = 0;
    //! blk (=, or += unless zero) tr(block idx of the result) on the device.
    virtual void compute_block(bool zero, const index_t &idx, const tensor_tr &tr,
        double *blk) = 0;
    virtual double flops() const = 0;
};

//! btod_contract2 as an additive_op.
class contract2_op : public additive_op {
public:
    explicit contract2_op(btod_contract2 &op, btensor &a) : m_op(op), m_a(a) { }
    context &ctx() { return m_a.ctx(); }
    const symmetry &sym() const { return m_op.symc(); }
    const std::vector<size_t> &schedule() const { return m_op.schedule(); }
    double perform_into(btensor &t, int mode) {
        m_op.perform(t, mode);
        return m_op.stats().t_gpu_ms;
    }
    void compute_block(bool zero, const index_t &idx, const tensor_tr &tr, double *blk) {
        m_op.compute_block(zero, idx, tr, blk);
    }
    double flops() const { return m_op.flops(); }
private:
    btod_contract2 &m_op;
    btensor &m_a;
};

//! btod_contract3<N1, N2, N3, K1, K2> (libtensor/block_tensor/btod_contract3.h;
//! gen_block_tensor/impl/gen_bto_contract3_impl.h): D = kd (A B) C, the
//! intermediate AB = contr1(A, B) and D = contr2(AB, C).  The symmetry and
//! schedule of AB and D are those of the two contractions
//! (gen_bto_contract3::make_schedule, :501-526: the second contraction's
//! non-zero orbits over AB's schedule).  AB is computed whole into a resident
//! temporary instead of in batches of B.
class btod_symmetrize2;
class btod_contract3 : public additive_op {
public:
    //! sympermab: btod_symcontract3 (libtensor/block_tensor/btod_symcontract3.h;
    //! gen_block_tensor/impl/gen_bto_symcontract3_impl.h): AB is replaced by
    //! AB + s P(AB) (s = +1 if symmab) before the second contraction, with
    //! the symmetry and schedule of btod_symmetrize2 (make_schedule :539-590).
    btod_contract3(const contraction2 &c1, const contraction2 &c2, btensor &a, btensor &b,
        btensor &c, double kd, const perm_t *sympermab = 0, bool symmab = true);
    ~btod_contract3();

    context &ctx() { return m_a.ctx(); }
    const symmetry &sym() const { return m_op2->symc(); }
    const std::vector<size_t> &schedule() const { return m_op2->schedule(); }
    double flops() const { return m_op1->flops() + m_op2->flops(); }

    //! D := result (D zeroed, given the result symmetry).
    void perform(btensor &d, int mode);
    //! D += k * result for D with the symmetry of the result.
    void perform_add(btensor &d, double k, int mode);
    double perform_into(btensor &t, int mode);
    //! One block of D on the device (AB computed first; the reference
    //! leaves this unimplemented).
    void compute_block(bool zero, const index_t &id, const tensor_tr &tr, double *blk);
    double last_ms() const { return m_last_ms; }
    btod_contract2 &first() { return *m_op1; }
    btod_contract2 &second() { return *m_op2; }

private:
    btensor &m_a;
    std::unique_ptr<btod_contract2> m_op1;
    std::unique_ptr<additive_op> m_op1x;          //!< m_op1 as an additive_op
    std::unique_ptr<btod_symmetrize2> m_symab;    //!< symmetrization of AB
    std::unique_ptr<btensor> m_ab;
    std::unique_ptr<btod_contract2> m_op2;
    double m_last_ms;
    double make_ab(int mode);
};

//! Sum of transformed copies of another operation's result X
//! (btod_symmetrize2/3): every canonical block of B is a list of sources
//! (canonical block of X, transformation).  perform() computes X once into a
//! resident temporary and gathers every block of B from it with the batched
//! permute kernel, one launch per source rank (the k-th source of every
//! block of B in launch k, so no two descriptors of a launch share a
//! destination).
class symmetrizer : public additive_op {
public:
    context &ctx() { return m_op.ctx(); }
    const symmetry &sym() const { return m_sym; }
    const std::vector<size_t> &schedule() const { return m_sched; }
    double flops() const { return m_op.flops(); }

    //! B := result (B zeroed, given the result symmetry).
    void perform(btensor &b, int mode) { run(b, false, 1.0, mode); }
    //! B += d * result for B with the symmetry of the result.
    void perform_add(btensor &b, double d, int mode) { run(b, true, d, mode); }
    double perform_into(btensor &t, int mode) { return run(t, false, 1.0, mode); }
    //! Device time of the last perform (the operand's kernels and the
    //! gather launches), of the gather launches alone, and the bytes the
    //! gather moved (source read, destination written, and read when added to).
    double last_ms() const { return m_last_ms; }
    double last_gather_ms() const { return m_gather_ms; }
    double last_gather_bytes() const { return m_gather_bytes; }

protected:
    struct rec {
        size_t ai;        //!< canonical block of X (in the operand's schedule)
        tensor_tr tr;     //!< contribution tr(X_ai)
    };
    explicit symmetrizer(additive_op &op) : m_op(op), m_uid(next_op_uid()), m_last_ms(0.0),
        m_gather_ms(0.0), m_gather_bytes(0.0) { }
    const std::vector<rec> *records(size_t aib) const;

    additive_op &m_op;
    symmetry m_sym;
    std::vector<size_t> m_sched;
    std::vector<std::pair<size_t, std::vector<rec> > > m_recs;  //!< by block of B, sorted

private:
    std::unique_ptr<btensor> m_x;
    uint64_t m_uid;
    double m_last_ms, m_gather_ms, m_gather_bytes;
    double run(btensor &b, bool add, double d, int mode);
};

//! btod_symmetrize2<N> (libtensor/block_tensor/btod_symmetrize2.h:20-110;
//! gen_block_tensor/impl/gen_bto_symmetrize2_impl.h): B = X + s P(X) for the
//! result X of another operation and a pair permutation P, s = +1
//! (symmetric) or -1 (antisymmetric).  The result symmetry, the schedule and
//! the sources of every block follow the reference.
class btod_symmetrize2 : public symmetrizer {
public:
    btod_symmetrize2(additive_op &op, const perm_t &perm, bool symm);
    //! gen_bto_symmetrize2::compute_block (impl :66-128) on the device.
    void compute_block(bool zero, const index_t &ib, const tensor_tr &trb, double *blk);
private:
    perm_t m_perm;
    bool m_symm;
    void make_schedule();
};

//! btod_symmetrize3<N> (libtensor/block_tensor/btod_symmetrize3.h;
//! gen_block_tensor/impl/gen_bto_symmetrize3_impl.h): B = sum over the six
//! exchanges g of three index groups generated by the pair permutations p1
//! and p2 of s(g) g(X), s(g) = s for the three pair exchanges and 1 for the
//! identity and the two cycles.
class btod_symmetrize3 : public symmetrizer {
public:
    btod_symmetrize3(additive_op &op, const perm_t &p1, const perm_t &p2, bool symm);
    //! gen_bto_symmetrize3::compute_block (impl :66-128) on the device.
    void compute_block(bool zero, const index_t &ib, const tensor_tr &trb, double *blk);
private:
    perm_t m_p1, m_p2;
    bool m_symm;
    std::vector<size_t> m_sched0;   //!< the operand's schedule, sorted
    void make_schedule();
    void block_records(size_t aib, std::vector<rec> &out) const;
};

} // namespace lta
