// Device context, slab storage and HBM-resident block tensors.
#include "lt_runtime.h"

#include <cmath>
#include <cstdlib>
#include <random>
#include <set>
#include <thread>

namespace lta {

void hip_check(hipError_t e, const char *what) {
    if(e != hipSuccess) {
        throw lt_error(LT_ERR_DEVICE, std::string(what) + ": " + hipGetErrorString(e));
    }
}

// ---------------------------------------------------------------------------

context::context(int device, size_t hbm_cap) :
    m_device(device), m_stream(0), m_xfer(0), m_xfer_ev(0), m_cap(hbm_cap), m_used(0),
    m_nthreads(1) {

    unsigned hc = std::thread::hardware_concurrency();
    m_nthreads = hc ? (hc > 16 ? 16 : hc) : 1;
    if(device < 0) {
        if(m_cap == 0) m_cap = size_t(1) << 40;
        return;
    }
    int n = 0;
    hip_check(hipGetDeviceCount(&n), "hipGetDeviceCount");
    lt_require(device >= 0 && device < n, LT_ERR_DEVICE, "no such HIP device");
    hip_check(hipSetDevice(device), "hipSetDevice");
    hip_check(hipStreamCreateWithFlags(&m_stream, hipStreamNonBlocking),
        "hipStreamCreate");
    // uploads go through a stream of their own, so a host thread that plans
    // (uploads work lists) does not wait for kernels queued on the compute
    // stream by another thread; every upload completes before the host
    // launches anything that reads it
    hip_check(hipStreamCreateWithFlags(&m_xfer, hipStreamNonBlocking), "hipStreamCreate");
    hip_check(hipEventCreateWithFlags(&m_xfer_ev, hipEventDisableTiming), "hipEventCreate");
    if(m_cap == 0) {
        size_t fr = 0, tot = 0;
        hip_check(hipMemGetInfo(&fr, &tot), "hipMemGetInfo");
        m_cap = fr - fr / 16;   // leave headroom for the runtime
    }
}

void context::require_device(const char *what) const {
    if(m_device < 0) {
        throw lt_error(LT_ERR_DEVICE, std::string(what) +
            ": requires a GPU context (this context is host-only)");
    }
}

context::~context() {
    if(m_device < 0) {
        for(std::map<void*, size_t>::iterator i = m_sizes.begin(); i != m_sizes.end(); ++i)
            std::free(i->first);
        return;
    }
    hipSetDevice(m_device);
    if(m_stream) hipStreamSynchronize(m_stream);
    m_launch_cache.clear();      // their memory goes with the loop below
    for(std::map<void*, size_t>::iterator i = m_sizes.begin(); i != m_sizes.end(); ++i)
        hipFree(i->first);
    if(m_xfer_ev) hipEventDestroy(m_xfer_ev);
    for(int i = 0; i < kSideStreams; i++) {
        if(m_side[i]) hipStreamSynchronize(m_side[i]);
        if(m_side_ev[i]) hipEventDestroy(m_side_ev[i]);
        if(m_side[i]) hipStreamDestroy(m_side[i]);
    }
    if(m_fork_ev) hipEventDestroy(m_fork_ev);
    if(m_stream) hipStreamDestroy(m_stream);
    if(m_xfer) hipStreamDestroy(m_xfer);
}

void context::side_fork(hipStream_t s) {
    require_device("side streams");
    if(!m_fork_ev) {
        for(int i = 0; i < kSideStreams; i++) {
            hip_check(hipStreamCreateWithFlags(&m_side[i], hipStreamNonBlocking), "hipStreamCreate");
            hip_check(hipEventCreateWithFlags(&m_side_ev[i], hipEventDisableTiming),
                "hipEventCreate");
        }
        hip_check(hipEventCreateWithFlags(&m_fork_ev, hipEventDisableTiming), "hipEventCreate");
    }
    hip_check(hipEventRecord(m_fork_ev, s), "hipEventRecord");
    for(int i = 0; i < kSideStreams; i++)
        hip_check(hipStreamWaitEvent(m_side[i], m_fork_ev, 0), "hipStreamWaitEvent");
}

void context::side_join(hipStream_t s) {
    for(int i = 0; i < kSideStreams; i++) {
        hip_check(hipEventRecord(m_side_ev[i], m_side[i]), "hipEventRecord");
        hip_check(hipStreamWaitEvent(s, m_side_ev[i], 0), "hipStreamWaitEvent");
    }
}

void context::activate() const {
    if(m_device < 0) return;
    hip_check(hipSetDevice(m_device), "hipSetDevice");
}

std::shared_ptr<context::cached_launch> context::find_launch(const std::string &key) {
    std::lock_guard<std::mutex> lk(m_cache_mtx);
    std::map<std::string, std::shared_ptr<cached_launch> >::iterator i = m_launch_cache.find(key);
    return i == m_launch_cache.end() ? std::shared_ptr<cached_launch>() : i->second;
}

void context::keep_launch(const std::string &key, std::shared_ptr<cached_launch> l) {
    std::vector<std::shared_ptr<cached_launch> > drop;
    {
        std::lock_guard<std::mutex> lk(m_cache_mtx);
        if(!m_launch_cache.count(key)) m_launch_order.push_back(key);
        m_launch_cache[key] = l;
        while(m_launch_order.size() > kLaunchCache) {
            drop.push_back(m_launch_cache[m_launch_order.front()]);
            m_launch_cache.erase(m_launch_order.front());
            m_launch_order.erase(m_launch_order.begin());
        }
    }
    for(size_t i = 0; i < drop.size(); i++) if(drop[i] && drop[i]->mem) free(drop[i]->mem);
}

void *context::staging(size_t bytes) {
    if(bytes > m_stage_bytes) {
        if(m_stage) free(m_stage);
        m_stage = 0;
        m_stage_bytes = 0;
        size_t n = std::max<size_t>(bytes, 65536);
        m_stage = alloc(n);
        m_stage_bytes = n;
    }
    return m_stage;
}

void *context::alloc(size_t bytes) {
    if(bytes == 0) bytes = 8;
    std::lock_guard<std::mutex> lk(m_mtx);
    if(m_used + bytes > m_cap) {
        throw lt_error(LT_ERR_NOMEM, "device memory budget exceeded");
    }
    void *p = 0;
    if(m_device < 0) {
        p = std::aligned_alloc(256, (bytes + 255) & ~size_t(255));
        lt_require(p != 0, LT_ERR_NOMEM, "host allocation failed");
        std::memset(p, 0, bytes);
        m_sizes[p] = bytes;
        m_used += bytes;
        return p;
    }
    activate();
    hipError_t e = hipMalloc(&p, bytes);
    if(e != hipSuccess) {
        hipGetLastError();
        throw lt_error(LT_ERR_NOMEM, std::string("hipMalloc: ") + hipGetErrorString(e));
    }
    m_sizes[p] = bytes;
    m_used += bytes;
    return p;
}

void context::free(void *p) {
    if(!p) return;
    std::lock_guard<std::mutex> lk(m_mtx);
    std::map<void*, size_t>::iterator i = m_sizes.find(p);
    if(i == m_sizes.end()) return;
    if(m_device < 0) {
        std::free(p);
    } else {
        activate();
        hipStreamSynchronize(m_stream);
        hipFree(p);
    }
    m_used -= i->second;
    m_sizes.erase(i);
}

void context::upload(void *dst, const void *src, size_t bytes) {
    if(m_device < 0) { std::memcpy(dst, src, bytes); return; }
    activate();
    std::lock_guard<std::mutex> lk(m_xfer_mtx);
    hip_check(hipEventRecord(m_xfer_ev, m_stream), "hipEventRecord");
    hip_check(hipStreamWaitEvent(m_xfer, m_xfer_ev, 0), "hipStreamWaitEvent");
    hip_check(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, m_xfer),
        "hipMemcpy H2D");
    hip_check(hipStreamSynchronize(m_xfer), "hipStreamSynchronize");
}

void context::upload_new(void *dst, const void *src, size_t bytes) {
    if(m_device < 0) { std::memcpy(dst, src, bytes); return; }
    activate();
    hip_check(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, m_xfer),
        "hipMemcpy H2D");
    hip_check(hipStreamSynchronize(m_xfer), "hipStreamSynchronize");
}

void context::download(void *dst, const void *src, size_t bytes) {
    if(m_device < 0) { std::memcpy(dst, src, bytes); return; }
    activate();
    hip_check(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, m_stream),
        "hipMemcpy D2H");
    hip_check(hipStreamSynchronize(m_stream), "hipStreamSynchronize");
}

void context::sync() {
    if(m_device < 0) return;
    activate();
    hip_check(hipStreamSynchronize(m_stream), "hipStreamSynchronize");
}

// ---------------------------------------------------------------------------

slab::~slab() { release_all(); }

double *slab::take(size_t n) {
    n = (n + 31) & ~size_t(31);      // 256-byte aligned blocks
    if(m_chunks.empty() || m_off + n > m_cap) {
        size_t sz = std::max(n, size_t(64) << 17);   // >= 64 MiB of doubles / 8
        double *p = (double*)m_ctx.alloc(sz * sizeof(double));
        m_chunks.push_back(std::make_pair(p, sz));
        m_off = 0;
        m_cap = sz;
    }
    double *r = m_chunks.back().first + m_off;
    m_off += n;
    return r;
}

bool slab::owns(const void *p) const {
    const double *x = (const double*)p;
    for(size_t i = 0; i < m_chunks.size(); i++)
        if(x >= m_chunks[i].first && x < m_chunks[i].first + m_chunks[i].second) return true;
    return false;
}

size_t slab::bytes() const {
    size_t b = 0;
    for(size_t i = 0; i < m_chunks.size(); i++) b += m_chunks[i].second * sizeof(double);
    return b;
}

void slab::release_all() {
    for(size_t i = 0; i < m_chunks.size(); i++) m_ctx.free(m_chunks[i].first);
    m_chunks.clear();
    m_off = m_cap = 0;
}

// ---------------------------------------------------------------------------

std::atomic<uint64_t> btensor::s_version(1);

btensor::btensor(context &ctx, const bis_t &bis) :
    m_ctx(ctx), m_sym(bis), m_bidims(bis.block_index_dims()), m_slab(ctx),
    m_version(s_version++), m_uid(s_version++) {

    lt_require(bis.order() >= 1 && bis.order() <= LT_MAX_ORDER, LT_ERR_PARAM,
        "block tensor order must be 1..8");
}

void btensor::require_mutable(const char *what) const {
    if(m_immutable)
        throw lt_error(LT_ERR_STATE, std::string(what) + ": the block tensor is immutable");
}

symmetry &btensor::sym_rw() {
    require_mutable("symmetry change");
    touch();
    m_sym_tag = 0;
    return m_sym;
}

void btensor::split(const bool *mask, size_t pos) {
    require_mutable("split");
    lt_require(m_store.empty(), LT_ERR_STATE, "split after blocks were allocated");
    lt_require(m_sym.subsets.empty(), LT_ERR_STATE, "split after symmetry was set");
    for(size_t i = 0; i < m_sym.bis.order(); i++)
        if(mask[i]) m_sym.bis.split(i, pos);
    m_bidims = m_sym.bis.block_index_dims();
    m_sym_tag = 0;
    touch();
}

void btensor::set_symmetry(const symmetry &s, uint64_t tag) {
    require_mutable("set_symmetry");
    lt_require(s.bis == m_sym.bis, LT_ERR_SYMMETRY, "symmetry of a different block index space");
    m_sym = s;
    m_sym_tag = tag;
    touch();
}

size_t btensor::block_size(size_t aidx) const {
    return m_sym.bis.block_dims(get_index(aidx, m_bidims)).size();
}

double *btensor::block_ptr(size_t aidx) {
    std::map<size_t, double*>::iterator i = m_store.find(aidx);
    if(i != m_store.end()) return i->second;
    lt_require(aidx < m_bidims.size(), LT_ERR_PARAM, "block index out of range");
    double *p = m_slab.take(block_size(aidx));
    m_store[aidx] = p;
    return p;
}

bool btensor::is_nonzero(size_t aidx) const {
    return std::binary_search(m_nonzero.begin(), m_nonzero.end(), aidx);
}

void btensor::mark_nonzero(size_t aidx) {
    require_mutable("block write");
    std::vector<size_t>::iterator i = std::lower_bound(m_nonzero.begin(), m_nonzero.end(), aidx);
    if(i != m_nonzero.end() && *i == aidx) return;
    m_nonzero.insert(i, aidx);
    touch();
}

void btensor::zero_block(size_t aidx) {
    require_mutable("zero_block");
    std::vector<size_t>::iterator i = std::lower_bound(m_nonzero.begin(), m_nonzero.end(), aidx);
    if(i != m_nonzero.end() && *i == aidx) {
        m_nonzero.erase(i);
        touch();
    }
}

void btensor::zero_all() {
    require_mutable("zero_all");
    if(!m_nonzero.empty()) touch();
    m_nonzero.clear();
}

std::vector<size_t> btensor::nonzero_blocks() const { return m_nonzero; }

void btensor::set_block(size_t aidx, const double *host) {
    require_mutable("set_block");
    short_orbit so(m_sym, aidx, true);
    lt_require(so.acidx == aidx, LT_ERR_SYMMETRY, "block is not canonical");
    lt_require(so.allowed, LT_ERR_SYMMETRY, "block is forbidden by symmetry");
    double *p = block_ptr(aidx);
    m_ctx.upload(p, host, block_size(aidx) * sizeof(double));
    mark_nonzero(aidx);
}

void btensor::get_block(size_t aidx, double *host) {
    size_t n = block_size(aidx);
    if(!is_nonzero(aidx)) {
        for(size_t i = 0; i < n; i++) host[i] = 0.0;
        return;
    }
    m_ctx.download(host, block_ptr(aidx), n * sizeof(double));
}

// ---------------------------------------------------------------------------

void host_transform(const double *src, const dims_t &dsrc, const tensor_tr &tr,
    double *dst, bool accumulate) {

    const size_t n = dsrc.n;
    dims_t ddst(dsrc);
    ddst.permute(tr.perm);
    perm_t pinv(tr.perm.inverse());
    size_t dstride[kMaxOrder];
    for(size_t j = 0; j < n; j++) dstride[j] = ddst.inc(pinv[j]);
    size_t total = dsrc.size();
    const double c = tr.st.c;
    index_t idx(n);
    size_t doff = 0;
    for(size_t e = 0; e < total; e++) {
        if(accumulate) dst[doff] += c * src[e];
        else dst[doff] = c * src[e];
        // odometer over the source index, tracking the destination offset
        for(size_t i = n; i-- > 0;) {
            if(idx[i] + 1 < dsrc[i]) {
                idx[i]++;
                doff += dstride[i];
                break;
            }
            doff -= idx[i] * dstride[i];
            idx[i] = 0;
        }
    }
}

namespace {
struct tr_key {
    size_t a;
    perm_t p;
    double c;
    bool operator<(const tr_key &o) const {
        if(a != o.a) return a < o.a;
        if(p != o.p) return p < o.p;
        return c < o.c;
    }
};
}

std::vector<tensor_tr> block_stabilizer(const symmetry &sym, size_t aidx) {
    dims_t bid = sym.bis.block_index_dims();
    const size_t n = bid.n;
    std::vector<const sym_element*> elems;
    sym.for_each([&](const sym_element &e) { elems.push_back(&e); });
    std::set<tr_key> seen;
    std::vector<std::pair<index_t, tensor_tr> > q;
    std::vector<tensor_tr> stab;
    tensor_tr t0(n);
    q.push_back(std::make_pair(get_index(aidx, bid), t0));
    tr_key k0 = { aidx, t0.perm, t0.st.c };
    seen.insert(k0);
    stab.push_back(t0);
    while(!q.empty()) {
        std::pair<index_t, tensor_tr> cur = q.back();
        q.pop_back();
        for(size_t i = 0; i < elems.size(); i++) {
            index_t idx(cur.first);
            tensor_tr t(cur.second);
            elems[i]->apply(idx, t);
            tr_key k = { abs_index(idx, bid), t.perm, t.st.c };
            if(seen.count(k)) continue;
            lt_require(seen.size() < 200000, LT_ERR_SYMMETRY, "symmetry group too large");
            seen.insert(k);
            if(k.a == aidx) stab.push_back(t);
            q.push_back(std::make_pair(idx, t));
        }
    }
    return stab;
}

void btensor::fill_random(uint64_t seed, int dist, double scale, int sk_d0, int sk_d1) {
    require_mutable("random fill");
    const size_t n = m_sym.bis.order();
    const size_t nblk = m_bidims.size();
    std::vector<size_t> canon;
    for(size_t a = 0; a < nblk; a++) {
        short_orbit so(m_sym, a, true);
        if(so.allowed && so.acidx == a) canon.push_back(a);
    }
    // Values are drawn in chunks of 2^20 elements, each chunk with its own
    // generator seeded from (seed, block, chunk), so large blocks fill in
    // parallel and the values do not depend on the thread count.
    const size_t kChunk = size_t(1) << 20;
    std::vector<std::vector<double> > data(canon.size());
    std::vector<std::pair<size_t, size_t> > tasks;
    for(size_t t = 0; t < canon.size(); t++) {
        size_t sz = m_sym.bis.block_dims(get_index(canon[t], m_bidims)).size();
        data[t].resize(sz);
        for(size_t q = 0; q * kChunk < sz; q++) tasks.push_back(std::make_pair(t, q));
    }
    parallel_for(tasks.size(), m_ctx.nthreads(), [&](size_t w) {
        size_t t = tasks[w].first, q = tasks[w].second;
        size_t a = canon[t];
        index_t bidx = get_index(a, m_bidims);
        dims_t bd = m_sym.bis.block_dims(bidx);
        index_t start = m_sym.bis.block_start(bidx);
        size_t sz = bd.size(), i0 = q * kChunk, i1 = std::min(sz, i0 + kChunk);
        std::mt19937_64 rng(seed * 0x9E3779B97F4A7C15ull + a * 0xBF58476D1CE4E5B9ull + 1 +
            q * 0x94D049BB133111EBull);
        std::uniform_real_distribution<double> u(-1.0, 1.0);
        std::normal_distribution<double> g(0.0, 1.0);
        index_t e = get_index(i0, bd);
        double *r = data[t].data();
        for(size_t i = i0; i < i1; i++) {
            double v = (dist == 1) ? g(rng) : u(rng);
            v *= scale;
            if(sk_d0 >= 0 && sk_d1 >= 0) {
                double x0 = double(start[sk_d0] + e[sk_d0]);
                double x1 = double(start[sk_d1] + e[sk_d1]);
                v /= 1.0 + std::fabs(x0 - x1);
            }
            r[i] = v;
            inc_index(e, bd);
        }
    });
    // symmetrization over each block's stabilizer
    parallel_for(canon.size(), m_ctx.nthreads(), [&](size_t t) {
        size_t a = canon[t];
        std::vector<tensor_tr> stab = block_stabilizer(m_sym, a);
        if(stab.size() <= 1) return;
        dims_t bd = m_sym.bis.block_dims(get_index(a, m_bidims));
        std::vector<double> sym(data[t].size(), 0.0);
        for(size_t i = 0; i < stab.size(); i++)
            host_transform(data[t].data(), bd, stab[i], sym.data(), true);
        double inv = 1.0 / double(stab.size());
        for(size_t i = 0; i < sym.size(); i++) sym[i] *= inv;
        data[t].swap(sym);
    });
    (void)n;
    for(size_t t = 0; t < canon.size(); t++) {
        double *p = block_ptr(canon[t]);
        m_ctx.upload(p, data[t].data(), data[t].size() * sizeof(double));
        std::vector<double>().swap(data[t]);
        mark_nonzero(canon[t]);
    }
}

} // namespace lta
