// C ABI of the block-contraction path (declarations in include/lt_amd.h).
#include "../../../include/lt_amd.h"
#include "lt_runtime.h"
#include "lt_server.h"

#include <atomic>
#include <cstring>
#include <string>

using namespace lta;

static_assert(LT_NOALLOC_SCRATCH_BYTES == lta::dev::kNoallocScratchBytes,
    "scratch region size in the header and the device source");

// A context stays alive while tensors, device arrays or operations created
// from it exist, whatever order the caller releases handles in.
struct lt_ctx {
    context *c;
    std::atomic<long> refs;
};
struct lt_btensor { btensor *t; lt_ctx *ctx; };
struct lt_symel { std::shared_ptr<sym_element> e; };
struct lt_btod_contract2 { btod_contract2 *op; lt_btensor *a, *b; lt_ctx *ctx; };
struct lt_server { server *s; lt_ctx *ctx; };
struct lt_btod_contract3 { btod_contract3 *op; lt_ctx *ctx; };
struct lt_btod_symmetrize2 {
    symmetrizer *s;     //!< btod_symmetrize2 or btod_symmetrize3
    additive_op *inner;    //!< owned adapter of a btod_contract2 operand, or null
    lt_ctx *ctx;
};

static void ctx_retain(lt_ctx *x) { x->refs++; }
static void ctx_release(lt_ctx *x) {
    if(--x->refs == 0) {
        delete x->c;
        delete x;
    }
}

namespace {

thread_local std::string g_err;

template<typename F>
int guard(F f) {
    try {
        f();
        return LT_OK;
    } catch(const lt_error &e) {
        g_err = e.what();
        return e.code;
    } catch(const std::bad_alloc &) {
        g_err = "host out of memory";
        return LT_ERR_NOMEM;
    } catch(const std::exception &e) {
        g_err = e.what();
        return LT_ERR_INTERNAL;
    } catch(...) {
        g_err = "unknown error";
        return LT_ERR_INTERNAL;
    }
}

void need(const void *p, const char *what) {
    if(!p) throw lt_error(LT_ERR_PARAM, std::string("null argument: ") + what);
}

perm_t perm_or_identity(size_t n, const size_t *p) {
    perm_t r(n);
    if(p) {
        bool seen[kMaxOrder] = {false};
        for(size_t i = 0; i < n; i++) {
            lt_require(p[i] < n && !seen[p[i]], LT_ERR_PARAM, "invalid permutation");
            seen[p[i]] = true;
            r.p[i] = (uint8_t)p[i];
        }
    }
    return r;
}

void fill_transf(const tensor_tr &tr, size_t n, size_t *perm, double *coeff) {
    if(perm) for(size_t i = 0; i < n; i++) perm[i] = tr.perm[i];
    if(coeff) *coeff = tr.st.c;
}

contraction2 make_contraction(size_t N, size_t M, size_t K, const size_t *permc,
    const size_t *ka, const size_t *kb) {
    lt_require(N + K <= LT_MAX_ORDER && M + K <= LT_MAX_ORDER &&
        N + M <= LT_MAX_ORDER && N + M >= 1, LT_ERR_PARAM, "invalid orders");
    perm_t pc = perm_or_identity(N + M, permc);
    contraction2 c(N, M, K, &pc);
    for(size_t i = 0; i < K; i++) c.contract(ka[i], kb[i]);
    return c;
}

//! Device memory for the temporaries of one dense operation: taken from the
//! context's allocator and released when the arena goes out of scope, or
//! carved out of a fixed device region (the no-allocation entry point; no
//! allocation at all, LT_ERR_PARAM when the region is too small).  Callers
//! take every buffer before their first device operation, so a refusal
//! leaves the device untouched.
class scratch_arena {
public:
    explicit scratch_arena(context &ctx, char *region = 0, size_t cap = 0) :
        m_ctx(ctx), m_region(region), m_cap(cap), m_off(0), m_measure(false) { }
    //! A measuring arena: take() hands out placeholder addresses and only
    //! adds up the bytes (used() after the takes of a layout).
    struct measuring { };
    scratch_arena(context &ctx, measuring) :
        m_ctx(ctx), m_region(0), m_cap(0), m_off(0), m_measure(true) { }
    ~scratch_arena() {
        if(!m_region && !m_measure) for(size_t i = 0; i < m_taken.size(); i++) m_ctx.free(m_taken[i]);
    }
    scratch_arena(const scratch_arena&) = delete;
    scratch_arena &operator=(const scratch_arena&) = delete;
    size_t used() const { return m_off; }
    //! A 256-byte aligned buffer of `bytes` bytes.
    char *take(size_t bytes) {
        bytes = (std::max<size_t>(bytes, 8) + 255) & ~size_t(255);
        if(m_measure) {
            char *p = (char*)(uintptr_t(256) + m_off);
            m_off += bytes;
            return p;
        }
        if(!m_region) {
            char *p = (char*)m_ctx.alloc(bytes);
            m_taken.push_back(p);
            return p;
        }
        if(bytes > m_cap - m_off) {
            throw lt_error(LT_ERR_PARAM, "temporaries exceed the fixed scratch region of " +
                std::to_string(m_cap) + " bytes");
        }
        char *p = m_region + m_off;
        m_off += bytes;
        return p;
    }
private:
    context &m_ctx;
    char *m_region;
    size_t m_cap, m_off;
    bool m_measure;
    std::vector<void*> m_taken;
};

//! One permute launch: descriptor and work-item lists laid out for one
//! upload (descriptor at 0, chunk descriptors at ocd, offsets at oco).
struct permute_job {
    std::vector<char> host;
    size_t ocd, oco, nchunks;
    char *mem;
    permute_job() : ocd(0), oco(0), nchunks(0), mem(0) { }
    void prepare(const double *src, const dims_t &dsrc, const tensor_tr &tr, double *dst,
        bool accumulate, scratch_arena &ar) {
        dev::perm_desc d = make_perm_desc_public(src, dsrc, tr, dst, accumulate);
        std::vector<uint32_t> cd, co;
        make_chunks(std::vector<dev::perm_desc>(1, d), cd, co);
        nchunks = cd.size();
        if(!nchunks) return;
        ocd = (sizeof(d) + 255) & ~size_t(255);
        oco = ocd + ((4 * cd.size() + 255) & ~size_t(255));
        host.assign(oco + 4 * co.size(), 0);
        std::memcpy(&host[0], &d, sizeof(d));
        std::memcpy(&host[ocd], cd.data(), 4 * cd.size());
        std::memcpy(&host[oco], co.data(), 4 * co.size());
        mem = ar.take(host.size());
    }
    void upload(context &ctx) { if(nchunks) ctx.upload(mem, host.data(), host.size()); }
    void launch(context &ctx) const {
        if(!nchunks) return;
        hip_check(dev::launch_permute((dev::perm_desc*)mem, (uint32_t*)(mem + ocd),
            (uint32_t*)(mem + oco), (uint32_t)nchunks, ctx.stream()), "permute");
    }
};

void dense_permute_in(context &ctx, const double *src, const dims_t &dsrc,
    const tensor_tr &tr, double *dst, bool accumulate, scratch_arena &ar) {
    permute_job j;
    j.prepare(src, dsrc, tr, dst, accumulate, ar);
    if(!j.nchunks) return;
    j.upload(ctx);
    ctx.activate();
    j.launch(ctx);
    ctx.sync();
}

//! One tod_contract2 laid out on the host: its operand copies and aligned
//! result taken from `tmp`, its descriptor and tile lists from `dsc` (host
//! images kept for the upload), all before any device operation.
struct contract_job {
    bool skip;                  //!< d == 0, C kept
    bool fill;                  //!< d == 0, C := 0
    double *c;
    size_t cn;
    permute_job ja, jb, jc;
    char *mem;                  //!< GEMM block, pair and tile lists
    std::vector<char> host;
    uint32_t ntiles;
    contract_job() : skip(false), fill(false), c(0), cn(0), mem(0), ntiles(0) { }
    //! The uploads of the descriptor lists (each ordered after the stream).
    void upload(context &cx) {
        ja.upload(cx);
        jb.upload(cx);
        jc.upload(cx);
        if(ntiles) cx.upload(mem, host.data(), host.size());
    }
    //! Every blob of the descriptor lists into a host image of the region
    //! at `base` (one upload for many jobs).
    void stage(char *image, const char *base) const {
        const permute_job *js[3] = { &ja, &jb, &jc };
        for(int i = 0; i < 3; i++)
            if(js[i]->nchunks)
                std::memcpy(image + (js[i]->mem - base), js[i]->host.data(), js[i]->host.size());
        if(ntiles) std::memcpy(image + (mem - base), host.data(), host.size());
    }
    //! The launches, in stream order (no synchronization).
    void launch(context &cx) const {
        if(skip) return;
        if(fill) {
            hip_check(dev::launch_fill(c, cn, 0.0, cx.stream()), "fill");
            return;
        }
        ja.launch(cx);
        jb.launch(cx);
        hip_check(dev::launch_gemm(dev::TILE_64x64, (dev::gemm_cblock*)mem,
            (dev::gemm_pair*)(mem + 512), (dev::gemm_tile*)(mem + 1024),
            ntiles, ntiles, cx.stream()), "gemm");
        jc.launch(cx);
    }
};

void contract_prepare(const contraction2 &contr, const dims_t &da, const double *a,
    const dims_t &db, const double *b, double d, bool zero, double *c,
    scratch_arena &tmp, scratch_arena &dsc, contract_job &j) {
    dims_t dc = contract2_dims(contr, da, db);
    // the GEMM kernels address operands with 32-bit byte offsets
    const size_t lim = size_t(1) << 29;
    lt_require(da.size() < lim && db.size() < lim && dc.size() < lim, LT_ERR_PARAM,
        "tod_contract2: a tensor of 2^29 or more elements (32-bit byte offsets in the "
        "GEMM kernels)");
    j.c = c;
    j.cn = dc.size();
    if(d == 0.0) {
        j.fill = zero;
        j.skip = !zero;
        return;
    }
    // align as tod_contract2::perform (tod_contract2_impl.h:124-135),
    // permute A / B into scratch if needed, one GEMM, permute C back
    contraction2_align al(contr);
    contraction2 ca(contr);
    ca.permute_a(al.perma);
    ca.permute_b(al.permb);
    ca.permute_c(al.permc);
    dims_t da2(da), db2(db), dc2(dc);
    da2.permute(al.perma);
    db2.permute(al.permb);
    dc2.permute(al.permc);

    const double *pa = a, *pb = b;
    if(!al.perma.is_identity()) {
        double *t = (double*)tmp.take(da.size() * 8);
        j.ja.prepare(a, da, tensor_tr(al.perma, scalar_tr()), t, false, dsc);
        pa = t;
    }
    if(!al.permb.is_identity()) {
        double *t = (double*)tmp.take(db.size() * 8);
        j.jb.prepare(b, db, tensor_tr(al.permb, scalar_tr()), t, false, dsc);
        pb = t;
    }
    double *pc = c;
    bool cperm = !al.permc.is_identity();
    if(cperm) {
        pc = (double*)tmp.take(dc.size() * 8);
        j.jc.prepare(pc, dc2, tensor_tr(al.permc.inverse(), scalar_tr()), c, !zero, dsc);
    }
    gemm_form g = aligned_gemm_form(ca, da2, db2);
    dev::gemm_cblock cb;
    dev::gemm_pair gp;
    std::memset(&cb, 0, sizeof(cb));
    std::memset(&gp, 0, sizeof(gp));
    size_t lda = g.ta ? g.m : g.k, ldb = g.tb ? g.k : g.n;
    bool ta = g.ta, tb = g.tb;
    if(!g.tc) {
        cb.m = (uint32_t)g.m; cb.n = (uint32_t)g.n;
        gp.a = pa; gp.b = pb; gp.lda = (uint32_t)lda; gp.ldb = (uint32_t)ldb;
    } else {
        cb.m = (uint32_t)g.n; cb.n = (uint32_t)g.m;
        gp.a = pb; gp.b = pa; gp.lda = (uint32_t)ldb; gp.ldb = (uint32_t)lda;
        ta = !g.tb; tb = !g.ta;
    }
    cb.ldc = cb.n;
    cb.c = pc;
    cb.zero = (cperm || zero) ? 1 : 0;
    cb.pair_begin = 0;
    cb.pair_end = 1;
    gp.k = (uint32_t)g.k;
    gp.d = d;
    gp.flags = (ta ? dev::kPairAKMajor : 0) | (tb ? dev::kPairBNMajor : 0);
    uint32_t bm = 64, bn = 64;
    dev::tile_shape(dev::TILE_64x64, bm, bn);
    std::vector<dev::gemm_tile> tiles;
    for(uint32_t i = 0; i < (cb.m + bm - 1) / bm; i++)
        for(uint32_t jj = 0; jj < (cb.n + bn - 1) / bn; jj++) {
            dev::gemm_tile t; t.cblock = 0; t.tm = (uint16_t)i; t.tn = (uint16_t)jj;
            tiles.push_back(t);
        }
    size_t bytes = 512 + 512 + tiles.size() * sizeof(dev::gemm_tile);
    j.mem = dsc.take(bytes);
    j.host.assign(bytes, 0);
    std::memcpy(&j.host[0], &cb, sizeof(cb));
    std::memcpy(&j.host[512], &gp, sizeof(gp));
    std::memcpy(&j.host[1024], tiles.data(), tiles.size() * sizeof(dev::gemm_tile));
    j.ntiles = (uint32_t)tiles.size();
}

//! tod_contract2 with every temporary (aligned operand copies, the aligned
//! result, descriptors and tile list) taken from `ar` before the first
//! device operation.
void dense_contract2_in(context &cx, const contraction2 &contr, const dims_t &da,
    const double *a, const dims_t &db, const double *b, double d, bool zero, double *c,
    scratch_arena &ar) {
    contract_job j;
    contract_prepare(contr, da, a, db, b, d, zero, c, ar, ar, j);
    if(j.skip) return;
    // device side, in stream order
    j.upload(cx);
    cx.activate();
    j.launch(cx);
    cx.sync();
}

} // unnamed namespace

namespace lta {

void dense_permute(context &ctx, const double *src, const dims_t &dsrc,
    const tensor_tr &tr, double *dst, bool accumulate) {
    scratch_arena ar(ctx);
    dense_permute_in(ctx, src, dsrc, tr, dst, accumulate, ar);
}

void dense_contract2(context &cx, const contraction2 &contr, const dims_t &da, const double *a,
    const dims_t &db, const double *b, double d, bool zero, double *c) {
    scratch_arena ar(cx);
    dense_contract2_in(cx, contr, da, a, db, b, d, zero, c, ar);
}

void dense_contract2_many(context &cx, const std::vector<dense_contract_item> &items,
    double *c) {
    if(items.empty()) return;
    // every temporary and descriptor of all items in one allocation: the
    // layout is measured first (same takes, placeholder addresses), then
    // made in the allocation; the descriptors go up in one upload, the
    // launches follow in item order, one synchronization at the end
    size_t tbytes = 0, dbytes = 0;
    {
        scratch_arena mt(cx, scratch_arena::measuring()), md(cx, scratch_arena::measuring());
        for(size_t i = 0; i < items.size(); i++) {
            contract_job j;
            contract_prepare(items[i].contr, items[i].da, items[i].a, items[i].db, items[i].b,
                items[i].d, items[i].zero, c, mt, md, j);
        }
        tbytes = mt.used();
        dbytes = md.used();
    }
    char *region = (char*)cx.alloc(std::max<size_t>(tbytes + dbytes, 256));
    try {
        scratch_arena at(cx, region, tbytes), ad(cx, region + tbytes, dbytes);
        std::vector<contract_job> jobs(items.size());
        for(size_t i = 0; i < items.size(); i++)
            contract_prepare(items[i].contr, items[i].da, items[i].a, items[i].db, items[i].b,
                items[i].d, items[i].zero, c, at, ad, jobs[i]);
        std::vector<char> image(dbytes, 0);
        for(size_t i = 0; i < jobs.size(); i++) jobs[i].stage(image.data(), region + tbytes);
        if(dbytes) cx.upload(region + tbytes, image.data(), dbytes);
        cx.activate();
        for(size_t i = 0; i < jobs.size(); i++) jobs[i].launch(cx);
        cx.sync();
    } catch(...) {
        cx.sync();
        cx.free(region);
        throw;
    }
    cx.free(region);
}
}

extern "C" {

const char *lt_last_error(void) { return g_err.c_str(); }
int lt_version(void) { return 10000; }

// ---- context ----

int lt_ctx_create(int device, size_t cap, lt_ctx **out) {
    return guard([&]() {
        need(out, "out");
        lt_ctx *c = new lt_ctx;
        c->refs = 1;
        try {
            c->c = new context(device, cap);
            if(device >= 0) {
                // resolve this device's fixed scratch region (no-allocation
                // entry points) once, with the device current
                void *p = 0;
                c->c->activate();
                hip_check(dev::noalloc_scratch(device, &p), "scratch region address");
            }
        } catch(...) {
            delete c->c;
            delete c;
            throw;
        }
        *out = c;
    });
}

int lt_ctx_destroy(lt_ctx *ctx) {
    return guard([&]() {
        if(!ctx) return;
        ctx_release(ctx);
    });
}

int lt_ctx_sync(lt_ctx *ctx) {
    return guard([&]() { need(ctx, "ctx"); ctx->c->sync(); });
}

int lt_ctx_mem(lt_ctx *ctx, size_t *used, size_t *cap) {
    return guard([&]() {
        need(ctx, "ctx");
        if(used) *used = ctx->c->used();
        if(cap) *cap = ctx->c->cap();
    });
}

int lt_ctx_set_cap(lt_ctx *ctx, size_t cap) {
    return guard([&]() {
        need(ctx, "ctx");
        lt_require(cap > 0, LT_ERR_PARAM, "device memory budget of 0 bytes");
        ctx->c->set_cap(cap);
    });
}

int lt_ctx_set_threads(lt_ctx *ctx, size_t n) {
    return guard([&]() { need(ctx, "ctx"); ctx->c->set_nthreads(n); });
}

int lt_ctx_last_op_ms(lt_ctx *ctx, double *ms) {
    return guard([&]() { need(ctx, "ctx"); need(ms, "ms"); *ms = ctx->c->last_op_ms(); });
}

int lt_dev_alloc(lt_ctx *ctx, size_t bytes, void **dptr) {
    return guard([&]() {
        need(ctx, "ctx"); need(dptr, "dptr");
        ctx->c->require_device("lt_dev_alloc");
        *dptr = ctx->c->alloc(bytes);
        ctx_retain(ctx);
    });
}

int lt_dev_free(lt_ctx *ctx, void *p) {
    return guard([&]() {
        need(ctx, "ctx");
        if(!p) return;
        ctx->c->free(p);
        ctx_release(ctx);
    });
}

int lt_dev_upload(lt_ctx *ctx, void *dst, const void *host, size_t bytes) {
    return guard([&]() {
        need(ctx, "ctx"); need(dst, "dst"); need(host, "host");
        ctx->c->upload(dst, host, bytes);
    });
}

int lt_dev_download(lt_ctx *ctx, void *host, const void *src, size_t bytes) {
    return guard([&]() {
        need(ctx, "ctx"); need(src, "src"); need(host, "host");
        ctx->c->download(host, src, bytes);
    });
}

// ---- dense ----

int lt_tod_copy(lt_ctx *ctx, size_t n, const size_t *dims_a, const double *a,
    const size_t *perm, double c, int zero, double *b) {
    return guard([&]() {
        need(ctx, "ctx"); need(dims_a, "dims_a"); need(a, "a"); need(b, "b");
        ctx->c->require_device("tod_copy");
        lt_require(n >= 1 && n <= LT_MAX_ORDER, LT_ERR_PARAM, "order");
        tensor_tr tr(perm_or_identity(n, perm), scalar_tr(c));
        dense_permute(*ctx->c, a, dims_t::from(n, dims_a), tr, b, zero == 0);
    });
}

int lt_tod_add(lt_ctx *ctx, size_t n, const size_t *dims, size_t nops,
    const double *const *a, const size_t *perms, const double *c, int zero, double *b) {
    return guard([&]() {
        need(ctx, "ctx"); need(dims, "dims"); need(b, "b");
        ctx->c->require_device("tod_add");
        lt_require(n >= 1 && n <= LT_MAX_ORDER, LT_ERR_PARAM, "order");
        lt_require(nops >= 1, LT_ERR_PARAM, "tod_add needs an operand");
        need(a, "a"); need(c, "c");
        dims_t da = dims_t::from(n, dims);
        // every permuted operand has the dimensions of the first one
        dims_t d0;
        for(size_t i = 0; i < nops; i++) {
            need(a[i], "a[i]");
            perm_t p = perm_or_identity(n, perms ? perms + i * n : 0);
            dims_t di = da;
            di.permute(p);
            if(i == 0) d0 = di;
            lt_require(di == d0, LT_ERR_DIMS, "tod_add: operands of different dimensions");
        }
        for(size_t i = 0; i < nops; i++) {
            tensor_tr tr(perm_or_identity(n, perms ? perms + i * n : 0), scalar_tr(c[i]));
            dense_permute(*ctx->c, a[i], da, tr, b, !(zero && i == 0));
        }
    });
}

int lt_tod_set(lt_ctx *ctx, size_t nelem, double v, double *a) {
    return guard([&]() {
        need(ctx, "ctx"); need(a, "a");
        ctx->c->require_device("tod_set");
        ctx->c->activate();
        hip_check(dev::launch_fill(a, nelem, v, ctx->c->stream()), "fill");
        ctx->c->sync();
    });
}

int lt_tod_contract2(lt_ctx *ctx, size_t N, size_t M, size_t K,
    const size_t *permc, const size_t *ka, const size_t *kb,
    const size_t *dims_a, const double *a, const size_t *dims_b,
    const double *b, double d, int zero, double *c) {
    return guard([&]() {
        need(ctx, "ctx"); need(dims_a, "dims_a"); need(dims_b, "dims_b");
        need(a, "a"); need(b, "b"); need(c, "c");
        if(K) { need(ka, "ka"); need(kb, "kb"); }
        context &cx = *ctx->c;
        cx.require_device("tod_contract2");
        contraction2 contr = make_contraction(N, M, K, permc, ka, kb);
        dense_contract2(cx, contr, dims_t::from(N + K, dims_a), a, dims_t::from(M + K, dims_b),
            b, d, zero != 0, c);
    });
}

int lt_tod_contract2_noalloc(lt_ctx *ctx, size_t N, size_t M, size_t K,
    const size_t *permc, const size_t *ka, const size_t *kb,
    const size_t *dims_a, const double *a, const size_t *dims_b,
    const double *b, double d, int zero, double *c) {
    return guard([&]() {
        need(ctx, "ctx"); need(dims_a, "dims_a"); need(dims_b, "dims_b");
        need(a, "a"); need(b, "b"); need(c, "c");
        if(K) { need(ka, "ka"); need(kb, "kb"); }
        context &cx = *ctx->c;
        cx.require_device("tod_contract2_noalloc");
        contraction2 contr = make_contraction(N, M, K, permc, ka, kb);
        cx.activate();
        void *region = 0;
        hip_check(dev::noalloc_scratch(cx.device(), &region), "scratch region address");
        scratch_arena ar(cx, (char*)region, dev::kNoallocScratchBytes);
        dense_contract2_in(cx, contr, dims_t::from(N + K, dims_a), a,
            dims_t::from(M + K, dims_b), b, d, zero != 0, c, ar);
    });
}

int lt_tod_dotprod(lt_ctx *ctx, size_t n, const size_t *dims_a,
    const double *a, const size_t *perma, const size_t *dims_b,
    const double *b, const size_t *permb, double c, double *result) {
    return guard([&]() {
        need(ctx, "ctx"); need(dims_a, "dims_a"); need(dims_b, "dims_b");
        need(a, "a"); need(b, "b"); need(result, "result");
        context &cx = *ctx->c;
        cx.require_device("tod_dotprod");
        perm_t pa = perm_or_identity(n, perma), pb = perm_or_identity(n, permb);
        dims_t da = dims_t::from(n, dims_a), db = dims_t::from(n, dims_b);
        dims_t da2(da), db2(db);
        da2.permute(pa);
        db2.permute(pb);
        lt_require(da2 == db2, LT_ERR_DIMS, "tod_dotprod: dimensions differ");
        // B seen in the order of A: apply pb, then the inverse of pa
        perm_t q(pb);
        q.permute(pa.inverse());
        dev::perm_desc d = make_perm_desc_public(b, db, tensor_tr(q, scalar_tr(c)),
            const_cast<double*>(a), false);
        std::vector<uint32_t> cd, co;
        for(uint32_t off = 0; off < d.total; off += dev::kPermChunk) {
            cd.push_back(0);
            co.push_back(off);
        }
        size_t nch = cd.size();
        size_t o_cd = 256, o_co = o_cd + ((4 * nch + 255) & ~size_t(255));
        size_t o_part = o_co + ((4 * nch + 255) & ~size_t(255));
        size_t o_out = o_part + ((8 * nch + 255) & ~size_t(255));
        size_t bytes = o_out + 256;
        char *mem = (char*)cx.alloc(bytes);
        try {
            std::vector<char> host(bytes, 0);
            std::memcpy(&host[0], &d, sizeof(d));
            if(nch) {
                std::memcpy(&host[o_cd], cd.data(), 4 * nch);
                std::memcpy(&host[o_co], co.data(), 4 * nch);
            }
            cx.upload(mem, host.data(), bytes);
            cx.activate();
            hip_check(dev::launch_dot((dev::perm_desc*)mem, (uint32_t*)(mem + o_cd),
                (uint32_t*)(mem + o_co), (uint32_t)nch, (double*)(mem + o_part),
                (double*)(mem + o_out), cx.stream()), "dot");
            cx.download(result, mem + o_out, 8);
        } catch(...) {
            cx.free(mem);
            throw;
        }
        cx.free(mem);
    });
}

// ---- product tables ----

int lt_product_table_create(const char *id, size_t nlabels) {
    return guard([&]() {
        need(id, "id");
        product_table_registry::instance().add(product_table(id, nlabels));
    });
}

int lt_product_table_add_product(const char *id, size_t l1, size_t l2, size_t lr) {
    return guard([&]() {
        need(id, "id");
        product_table t = product_table_registry::instance().get(id);
        t.add_product(l1, l2, lr);
        product_table_registry::instance().add(t);
    });
}

// ---- block tensors ----

int lt_btensor_create(lt_ctx *ctx, size_t n, const size_t *dims, lt_btensor **out) {
    return guard([&]() {
        need(ctx, "ctx"); need(dims, "dims"); need(out, "out");
        lt_require(n >= 1 && n <= LT_MAX_ORDER, LT_ERR_PARAM, "order must be 1..8");
        for(size_t i = 0; i < n; i++) lt_require(dims[i] > 0, LT_ERR_PARAM, "zero dimension");
        bis_t bis(dims_t::from(n, dims));
        lt_btensor *b = new lt_btensor;
        b->ctx = ctx;
        try {
            b->t = new btensor(*ctx->c, bis);
        } catch(...) {
            delete b;
            throw;
        }
        ctx_retain(ctx);
        *out = b;
    });
}

int lt_btensor_destroy(lt_btensor *bt) {
    return guard([&]() {
        if(!bt) return;
        lt_ctx *ctx = bt->ctx;
        delete bt->t;
        delete bt;
        ctx_release(ctx);
    });
}

int lt_btensor_split(lt_btensor *bt, const int *mask, size_t pos) {
    return guard([&]() {
        need(bt, "bt"); need(mask, "mask");
        bool m[kMaxOrder] = {false};
        for(size_t i = 0; i < bt->t->bis().order(); i++) m[i] = mask[i] != 0;
        bt->t->split(m, pos);
    });
}

int lt_btensor_order(lt_btensor *bt, size_t *n) {
    return guard([&]() { need(bt, "bt"); need(n, "n"); *n = bt->t->bis().order(); });
}

int lt_btensor_bidims(lt_btensor *bt, size_t *bidims) {
    return guard([&]() {
        need(bt, "bt"); need(bidims, "bidims");
        dims_t d = bt->t->bidims();
        for(size_t i = 0; i < d.n; i++) bidims[i] = d[i];
    });
}

int lt_btensor_block_dims(lt_btensor *bt, size_t aidx, size_t *dims) {
    return guard([&]() {
        need(bt, "bt"); need(dims, "dims");
        lt_require(aidx < bt->t->bidims().size(), LT_ERR_PARAM, "block index out of range");
        dims_t d = bt->t->bis().block_dims(get_index(aidx, bt->t->bidims()));
        for(size_t i = 0; i < d.n; i++) dims[i] = d[i];
    });
}

int lt_btensor_block_start(lt_btensor *bt, size_t aidx, size_t *start) {
    return guard([&]() {
        need(bt, "bt"); need(start, "start");
        lt_require(aidx < bt->t->bidims().size(), LT_ERR_PARAM, "block index out of range");
        index_t s = bt->t->bis().block_start(get_index(aidx, bt->t->bidims()));
        for(size_t i = 0; i < s.n; i++) start[i] = s[i];
    });
}

int lt_btensor_nonzero(lt_btensor *bt, size_t *out, size_t cap, size_t *count) {
    return guard([&]() {
        need(bt, "bt"); need(count, "count");
        std::vector<size_t> v = bt->t->nonzero_blocks();
        *count = v.size();
        if(out) for(size_t i = 0; i < v.size() && i < cap; i++) out[i] = v[i];
    });
}

int lt_btensor_set_block(lt_btensor *bt, size_t aidx, const double *host) {
    return guard([&]() {
        need(bt, "bt"); need(host, "host");
        lt_require(aidx < bt->t->bidims().size(), LT_ERR_PARAM, "block index out of range");
        bt->t->set_block(aidx, host);
    });
}

int lt_btensor_get_block(lt_btensor *bt, size_t aidx, double *host) {
    return guard([&]() {
        need(bt, "bt"); need(host, "host");
        lt_require(aidx < bt->t->bidims().size(), LT_ERR_PARAM, "block index out of range");
        bt->t->get_block(aidx, host);
    });
}

int lt_btensor_zero_block(lt_btensor *bt, size_t aidx) {
    return guard([&]() { need(bt, "bt"); bt->t->zero_block(aidx); });
}

int lt_btensor_zero_all(lt_btensor *bt) {
    return guard([&]() { need(bt, "bt"); bt->t->zero_all(); });
}

int lt_btensor_orbit(lt_btensor *bt, size_t aidx, size_t *acidx, size_t *perm,
    double *coeff, int *allowed) {
    return guard([&]() {
        need(bt, "bt");
        lt_require(aidx < bt->t->bidims().size(), LT_ERR_PARAM, "block index out of range");
        orbit o(bt->t->sym(), aidx, true);
        if(acidx) *acidx = o.acidx;
        fill_transf(o.get_transf(aidx), bt->t->bis().order(), perm, coeff);
        if(allowed) *allowed = o.allowed ? 1 : 0;
    });
}

int lt_btensor_random(lt_btensor *bt, uint64_t seed, int dist, double scale, int d0, int d1) {
    return guard([&]() {
        need(bt, "bt");
        int n = (int)bt->t->bis().order();
        lt_require(d0 < n && d1 < n, LT_ERR_PARAM, "scaling dimension out of range");
        bt->t->fill_random(seed, dist, scale, d0, d1);
    });
}

int lt_btensor_export(lt_btensor *bt, double *host) {
    return guard([&]() {
        need(bt, "bt"); need(host, "host");
        btensor &t = *bt->t;
        const bis_t &bis = t.bis();
        dims_t dims = bis.dims, bid = t.bidims();
        const size_t n = dims.n;
        size_t total = dims.size();
        for(size_t i = 0; i < total; i++) host[i] = 0.0;
        std::map<size_t, std::vector<double> > cache;
        for(size_t a = 0; a < bid.size(); a++) {
            orbit o(t.sym(), a, true);
            if(!o.allowed || !t.is_nonzero(o.acidx)) continue;
            std::vector<double> &src = cache[o.acidx];
            if(src.empty()) {
                src.resize(t.block_size(o.acidx));
                t.get_block(o.acidx, src.data());
            }
            index_t bi = get_index(a, bid);
            dims_t db = bis.block_dims(bi);
            std::vector<double> blk(db.size());
            host_transform(src.data(), bis.block_dims(get_index(o.acidx, bid)),
                o.get_transf(a), blk.data(), false);
            index_t st = bis.block_start(bi);
            index_t e(n);
            for(size_t k = 0; k < blk.size(); k++) {
                size_t off = 0;
                for(size_t i = 0; i < n; i++) off = off * dims[i] + st[i] + e[i];
                host[off] = blk[k];
                inc_index(e, db);
            }
        }
    });
}

int lt_btensor_import(lt_btensor *bt, const double *host) {
    return guard([&]() {
        need(bt, "bt"); need(host, "host");
        btensor &t = *bt->t;
        t.require_mutable("import");
        const bis_t &bis = t.bis();
        dims_t dims = bis.dims, bid = t.bidims();
        const size_t n = dims.n;
        t.zero_all();
        for(size_t a = 0; a < bid.size(); a++) {
            short_orbit so(t.sym(), a, true);
            if(!so.allowed || so.acidx != a) continue;
            index_t bi = get_index(a, bid);
            dims_t db = bis.block_dims(bi);
            index_t st = bis.block_start(bi);
            std::vector<double> blk(db.size());
            index_t e(n);
            bool nz = false;
            for(size_t k = 0; k < blk.size(); k++) {
                size_t off = 0;
                for(size_t i = 0; i < n; i++) off = off * dims[i] + st[i] + e[i];
                blk[k] = host[off];
                if(blk[k] != 0.0) nz = true;
                inc_index(e, db);
            }
            if(nz) t.set_block(a, blk.data());
        }
    });
}

int lt_btensor_set_immutable(lt_btensor *bt) {
    return guard([&]() { need(bt, "bt"); bt->t->set_immutable(); });
}

int lt_btensor_is_immutable(lt_btensor *bt, int *out) {
    return guard([&]() { need(bt, "bt"); need(out, "out"); *out = bt->t->is_immutable() ? 1 : 0; });
}

int lt_btensor_copy_symmetry(lt_btensor *dst, lt_btensor *src) {
    return guard([&]() {
        need(dst, "dst"); need(src, "src");
        dst->t->set_symmetry(src->t->sym());
    });
}

int lt_btensor_merge_symmetry(lt_btensor *bt, lt_btensor *other) {
    return guard([&]() {
        need(bt, "bt"); need(other, "other");
        lt_require(bt->t->bis() == other->t->bis(), LT_ERR_DIMS,
            "merge_symmetry: different block index spaces");
        if(!bt->t->nonzero_blocks().empty())
            bt->ctx->c->require_device("merge_symmetry of a tensor holding blocks");
        merge_symmetry(*bt->t, other->t->sym());
    });
}

// ---- symmetry elements ----

int lt_se_perm_create(size_t n, const size_t *perm, double coeff, lt_symel **out) {
    return guard([&]() {
        need(perm, "perm"); need(out, "out");
        lt_require(n >= 1 && n <= LT_MAX_ORDER, LT_ERR_PARAM, "order");
        lt_symel *e = new lt_symel;
        e->e.reset(new se_perm(perm_or_identity(n, perm), scalar_tr(coeff)));
        *out = e;
    });
}

int lt_se_part_create(lt_btensor *bt, const size_t *pdims, lt_symel **out) {
    return guard([&]() {
        need(bt, "bt"); need(pdims, "pdims"); need(out, "out");
        size_t n = bt->t->bis().order();
        lt_symel *e = new lt_symel;
        e->e.reset(new se_part(bt->t->bis(), dims_t::from(n, pdims)));
        *out = e;
    });
}

int lt_se_part_add_map(lt_symel *e, const size_t *idx1, const size_t *idx2, double coeff) {
    return guard([&]() {
        need(e, "e"); need(idx1, "idx1"); need(idx2, "idx2");
        se_part *p = dynamic_cast<se_part*>(e->e.get());
        lt_require(p != 0, LT_ERR_PARAM, "not an se_part");
        size_t n = p->pdims.n;
        index_t i1(n), i2(n);
        for(size_t i = 0; i < n; i++) {
            lt_require(idx1[i] < p->pdims[i] && idx2[i] < p->pdims[i], LT_ERR_PARAM,
                "partition index out of range");
            i1[i] = idx1[i];
            i2[i] = idx2[i];
        }
        p->add_map(i1, i2, scalar_tr(coeff));
    });
}

int lt_se_part_mark_forbidden(lt_symel *e, const size_t *idx) {
    return guard([&]() {
        need(e, "e"); need(idx, "idx");
        se_part *p = dynamic_cast<se_part*>(e->e.get());
        lt_require(p != 0, LT_ERR_PARAM, "not an se_part");
        size_t n = p->pdims.n;
        index_t i1(n);
        for(size_t i = 0; i < n; i++) {
            lt_require(idx[i] < p->pdims[i], LT_ERR_PARAM, "partition index out of range");
            i1[i] = idx[i];
        }
        p->mark_forbidden(i1);
    });
}

int lt_se_label_create(lt_btensor *bt, const char *table_id, lt_symel **out) {
    return guard([&]() {
        need(bt, "bt"); need(table_id, "table_id"); need(out, "out");
        lt_symel *e = new lt_symel;
        e->e.reset(new se_label(bt->t->bidims(), table_id));
        *out = e;
    });
}

int lt_se_label_assign(lt_symel *e, const int *mask, size_t blk, size_t label) {
    return guard([&]() {
        need(e, "e"); need(mask, "mask");
        se_label *l = dynamic_cast<se_label*>(e->e.get());
        lt_require(l != 0, LT_ERR_PARAM, "not an se_label");
        for(size_t i = 0; i < l->bidims.n; i++)
            if(mask[i]) l->set_label(i, blk, label);
    });
}

int lt_se_label_set_rule(lt_symel *e, const size_t *intr, size_t nintr) {
    return guard([&]() {
        need(e, "e");
        se_label *l = dynamic_cast<se_label*>(e->e.get());
        lt_require(l != 0, LT_ERR_PARAM, "not an se_label");
        std::vector<size_t> v;
        for(size_t i = 0; i < nintr; i++) v.push_back(intr[i]);
        l->set_rule_set(v);
    });
}

int lt_se_label_clear_rule(lt_symel *e) {
    return guard([&]() {
        need(e, "e");
        se_label *l = dynamic_cast<se_label*>(e->e.get());
        lt_require(l != 0, LT_ERR_PARAM, "not an se_label");
        l->rule.clear();
    });
}

int lt_se_label_add_product(lt_symel *e, size_t nterms, const size_t *seqs,
    const size_t *intr) {
    return guard([&]() {
        need(e, "e");
        se_label *l = dynamic_cast<se_label*>(e->e.get());
        lt_require(l != 0, LT_ERR_PARAM, "not an se_label");
        if(nterms) { need(seqs, "seqs"); need(intr, "intr"); }
        const size_t n = l->bidims.n;
        const product_table &pt = product_table_registry::instance().get(l->table_id);
        label_product pr;
        for(size_t t = 0; t < nterms; t++) {
            label_term term;
            term.seq = seq_t(n);
            for(size_t i = 0; i < n; i++) term.seq[i] = seqs[t * n + i];
            term.intrinsic = intr[t];
            lt_require(term.intrinsic == k_label_invalid || term.intrinsic < pt.nlabels,
                LT_ERR_PARAM, "se_label: invalid intrinsic label");
            pr.push_back(term);
        }
        l->rule.push_back(pr);
    });
}

int lt_symel_destroy(lt_symel *e) {
    return guard([&]() { delete e; });
}

int lt_btensor_insert(lt_btensor *bt, const lt_symel *e) {
    return guard([&]() {
        need(bt, "bt"); need(e, "e");
        if(const se_perm *p = dynamic_cast<const se_perm*>(e->e.get())) {
            lt_require(p->tr.perm.n == bt->t->bis().order(), LT_ERR_SYMMETRY, "order");
            bis_t b2(bt->t->bis());
            b2.permute(p->tr.perm);
            lt_require(b2 == bt->t->bis(), LT_ERR_SYMMETRY,
                "se_perm does not preserve the block index space");
        }
        if(const se_part *p = dynamic_cast<const se_part*>(e->e.get())) {
            lt_require(p->bis == bt->t->bis(), LT_ERR_SYMMETRY, "se_part: block index space");
        }
        if(const se_label *l = dynamic_cast<const se_label*>(e->e.get())) {
            lt_require(l->bidims == bt->t->bidims(), LT_ERR_SYMMETRY, "se_label: block index space");
        }
        bt->t->sym_rw().insert(*e->e);
    });
}

// ---- btod_contract2 ----

int lt_btod_contract2_create(size_t N, size_t M, size_t K, const size_t *permc,
    const size_t *ka_idx, const size_t *kb_idx, lt_btensor *a, double ka,
    lt_btensor *b, double kb, double kc, lt_btod_contract2 **out) {
    return guard([&]() {
        need(a, "a"); need(b, "b"); need(out, "out");
        if(K) { need(ka_idx, "ka_idx"); need(kb_idx, "kb_idx"); }
        lt_require(&a->t->ctx() == &b->t->ctx(), LT_ERR_PARAM, "tensors of different contexts");
        contraction2 c = make_contraction(N, M, K, permc, ka_idx, kb_idx);
        lt_btod_contract2 *h = new lt_btod_contract2;
        h->a = a;
        h->b = b;
        h->ctx = a->ctx;
        try {
            h->op = new btod_contract2(c, *a->t, ka, *b->t, kb, kc);
        } catch(...) {
            delete h;
            throw;
        }
        ctx_retain(h->ctx);
        *out = h;
    });
}

int lt_btod_contract2_destroy(lt_btod_contract2 *op) {
    return guard([&]() {
        if(!op) return;
        lt_ctx *ctx = op->ctx;
        delete op->op;
        delete op;
        ctx_release(ctx);
    });
}

int lt_btod_contract2_perform(lt_btod_contract2 *op, lt_btensor *c, int mode) {
    return guard([&]() {
        need(op, "op"); need(c, "c");
        lt_require(mode >= 0 && mode <= 2, LT_ERR_PARAM, "mode");
        op->op->perform(*c->t, mode);
    });
}

int lt_btod_contract2_perform_add(lt_btod_contract2 *op, lt_btensor *c, double d, int mode) {
    return guard([&]() {
        need(op, "op"); need(c, "c");
        lt_require(mode >= 0 && mode <= 2, LT_ERR_PARAM, "mode");
        op->op->perform_add(*c->t, d, mode);
    });
}

int lt_btod_contract2_schedule(lt_btod_contract2 *op, size_t *out, size_t cap, size_t *count) {
    return guard([&]() {
        need(op, "op"); need(count, "count");
        const std::vector<size_t> &s = op->op->schedule();
        *count = s.size();
        if(out) for(size_t i = 0; i < s.size() && i < cap; i++) out[i] = s[i];
    });
}

int lt_btod_contract2_symc_orbit(lt_btod_contract2 *op, size_t aidx, size_t *acidx,
    size_t *perm, double *coeff, int *allowed) {
    return guard([&]() {
        need(op, "op");
        const symmetry &s = op->op->symc();
        lt_require(aidx < s.bis.block_index_dims().size(), LT_ERR_PARAM, "block index");
        orbit o(s, aidx, true);
        if(acidx) *acidx = o.acidx;
        fill_transf(o.get_transf(aidx), s.bis.order(), perm, coeff);
        if(allowed) *allowed = o.allowed ? 1 : 0;
    });
}

int lt_btod_contract2_ozaki_bounds(lt_btod_contract2 *op, size_t *aidx, double *bound,
    double *cmax, size_t cap, size_t *count) {
    return guard([&]() {
        need(op, "op"); need(count, "count");
        const std::vector<btod_contract2::oz_bound> &v = op->op->ozaki_bounds();
        *count = v.size();
        if(cap == 0) return;
        lt_require(cap >= v.size(), LT_ERR_PARAM, "ozaki_bounds: buffer too small");
        for(size_t i = 0; i < v.size(); i++) {
            if(aidx) aidx[i] = v[i].aidx;
            if(bound) bound[i] = v[i].bound;
            if(cmax) cmax[i] = v[i].cmax;
        }
    });
}

int lt_btod_contract2_get_symmetry(lt_btod_contract2 *op, lt_btensor *dst) {
    return guard([&]() {
        need(op, "op"); need(dst, "dst");
        lt_require(dst->t->bis() == op->op->symc().bis, LT_ERR_DIMS,
            "get_symmetry: tensor of another block index space");
        lt_require(dst->t->nonzero_blocks().empty(), LT_ERR_STATE,
            "get_symmetry: the tensor holds blocks");
        dst->t->set_symmetry(op->op->symc());
    });
}

namespace {

void symm_create(lt_ctx *ctx, additive_op *op, additive_op *owned, const size_t *perm,
    const size_t *perm2, int symm, lt_btod_symmetrize2 **out) {
    const size_t n = op->sym().order();
    lt_btod_symmetrize2 *h = new lt_btod_symmetrize2;
    h->inner = owned;
    h->ctx = ctx;
    try {
        if(perm2)
            h->s = new btod_symmetrize3(*op, perm_or_identity(n, perm),
                perm_or_identity(n, perm2), symm != 0);
        else
            h->s = new btod_symmetrize2(*op, perm_or_identity(n, perm), symm != 0);
    } catch(...) {
        delete owned;
        delete h;
        throw;
    }
    ctx_retain(ctx);
    *out = h;
}

} // unnamed namespace

int lt_btod_symmetrize2_create(lt_btod_contract2 *op, const size_t *perm, int symm,
    lt_btod_symmetrize2 **out) {
    return guard([&]() {
        need(op, "op"); need(perm, "perm"); need(out, "out");
        contract2_op *inner = new contract2_op(*op->op, *op->a->t);
        symm_create(op->ctx, inner, inner, perm, 0, symm, out);
    });
}

int lt_btod_symmetrize2_create_c3(lt_btod_contract3 *op, const size_t *perm, int symm,
    lt_btod_symmetrize2 **out) {
    return guard([&]() {
        need(op, "op"); need(perm, "perm"); need(out, "out");
        symm_create(op->ctx, op->op, 0, perm, 0, symm, out);
    });
}

namespace {

void c3_create(size_t N1, size_t N2, size_t N3, size_t K1, size_t K2,
    const size_t *permab, const size_t *ka1, const size_t *kb1, const size_t *permd,
    const size_t *kab2, const size_t *kc2, lt_btensor *a, lt_btensor *b, lt_btensor *c,
    double kd, const size_t *sympermab, int symmab, lt_btod_contract3 **out) {

    need(a, "a"); need(b, "b"); need(c, "c"); need(out, "out");
    if(K1) { need(ka1, "ka1"); need(kb1, "kb1"); }
    if(K2) { need(kab2, "kab2"); need(kc2, "kc2"); }
    lt_require(&a->t->ctx() == &b->t->ctx() && &a->t->ctx() == &c->t->ctx(), LT_ERR_PARAM,
        "tensors of different contexts");
    lt_require(a->t->bis().order() == N1 + K1 && b->t->bis().order() == N2 + K1 + K2 &&
        c->t->bis().order() == N3 + K2, LT_ERR_DIMS, "btod_contract3: tensor orders");
    contraction2 c1 = make_contraction(N1, N2 + K2, K1, permab, ka1, kb1);
    contraction2 c2 = make_contraction(N1 + N2, N3, K2, permd, kab2, kc2);
    perm_t ps;
    if(sympermab) ps = perm_or_identity(N1 + N2 + K2, sympermab);
    lt_btod_contract3 *h = new lt_btod_contract3;
    h->ctx = a->ctx;
    try {
        h->op = new btod_contract3(c1, c2, *a->t, *b->t, *c->t, kd, sympermab ? &ps : 0,
            symmab != 0);
    } catch(...) {
        delete h;
        throw;
    }
    ctx_retain(h->ctx);
    *out = h;
}

} // unnamed namespace

int lt_btod_symcontract3_create(size_t N1, size_t N2, size_t N3, size_t K1, size_t K2,
    const size_t *permab, const size_t *ka1, const size_t *kb1, const size_t *permd,
    const size_t *kab2, const size_t *kc2, lt_btensor *a, lt_btensor *b,
    const size_t *sympermab, int symmab, lt_btensor *c, double kd, lt_btod_contract3 **out) {
    return guard([&]() {
        need(sympermab, "sympermab");
        c3_create(N1, N2, N3, K1, K2, permab, ka1, kb1, permd, kab2, kc2, a, b, c, kd,
            sympermab, symmab, out);
    });
}

int lt_btod_contract3_create(size_t N1, size_t N2, size_t N3, size_t K1, size_t K2,
    const size_t *permab, const size_t *ka1, const size_t *kb1, const size_t *permd,
    const size_t *kab2, const size_t *kc2, lt_btensor *a, lt_btensor *b, lt_btensor *c,
    double kd, lt_btod_contract3 **out) {
    return guard([&]() {
        c3_create(N1, N2, N3, K1, K2, permab, ka1, kb1, permd, kab2, kc2, a, b, c, kd, 0, 1,
            out);
    });
}

int lt_btod_contract3_destroy(lt_btod_contract3 *op) {
    return guard([&]() {
        if(!op) return;
        lt_ctx *ctx = op->ctx;
        delete op->op;
        delete op;
        ctx_release(ctx);
    });
}

int lt_btod_contract3_perform(lt_btod_contract3 *op, lt_btensor *d, int mode) {
    return guard([&]() {
        need(op, "op"); need(d, "d");
        lt_require(mode >= 0 && mode <= 2, LT_ERR_PARAM, "mode");
        op->op->perform(*d->t, mode);
    });
}

int lt_btod_contract3_perform_add(lt_btod_contract3 *op, lt_btensor *d, double k, int mode) {
    return guard([&]() {
        need(op, "op"); need(d, "d");
        lt_require(mode >= 0 && mode <= 2, LT_ERR_PARAM, "mode");
        op->op->perform_add(*d->t, k, mode);
    });
}

int lt_btod_contract3_schedule(lt_btod_contract3 *op, size_t *out, size_t cap, size_t *count) {
    return guard([&]() {
        need(op, "op"); need(count, "count");
        const std::vector<size_t> &v = op->op->schedule();
        *count = v.size();
        if(out) for(size_t i = 0; i < v.size() && i < cap; i++) out[i] = v[i];
    });
}

int lt_btod_contract3_sym_orbit(lt_btod_contract3 *op, size_t aidx, size_t *acidx,
    size_t *perm, double *coeff, int *allowed) {
    return guard([&]() {
        need(op, "op");
        const symmetry &y = op->op->sym();
        lt_require(aidx < y.bis.block_index_dims().size(), LT_ERR_PARAM, "block index");
        orbit o(y, aidx, true);
        if(acidx) *acidx = o.acidx;
        fill_transf(o.get_transf(aidx), y.bis.order(), perm, coeff);
        if(allowed) *allowed = o.allowed ? 1 : 0;
    });
}

int lt_btod_contract3_plans(lt_btod_contract3 *op, uint64_t *n) {
    return guard([&]() {
        need(op, "op"); need(n, "n");
        *n = op->op->first().stats().n_plans + op->op->second().stats().n_plans;
    });
}

int lt_btod_contract3_stats(lt_btod_contract3 *op, double *gpu_ms, double *flops) {
    return guard([&]() {
        need(op, "op");
        if(gpu_ms) *gpu_ms = op->op->last_ms();
        if(flops) *flops = op->op->flops();
    });
}

int lt_btod_symmetrize2_create_nested(lt_btod_symmetrize2 *op, const size_t *perm, int symm,
    lt_btod_symmetrize2 **out) {
    return guard([&]() {
        need(op, "op"); need(perm, "perm"); need(out, "out");
        symm_create(op->ctx, op->s, 0, perm, 0, symm, out);
    });
}

int lt_btod_symmetrize3_create(lt_btod_contract2 *op, const size_t *perm1,
    const size_t *perm2, int symm, lt_btod_symmetrize3 **out) {
    return guard([&]() {
        need(op, "op"); need(perm1, "perm1"); need(perm2, "perm2"); need(out, "out");
        contract2_op *inner = new contract2_op(*op->op, *op->a->t);
        symm_create(op->ctx, inner, inner, perm1, perm2, symm, out);
    });
}

int lt_btod_symmetrize3_create_nested(lt_btod_symmetrize2 *op, const size_t *perm1,
    const size_t *perm2, int symm, lt_btod_symmetrize3 **out) {
    return guard([&]() {
        need(op, "op"); need(perm1, "perm1"); need(perm2, "perm2"); need(out, "out");
        symm_create(op->ctx, op->s, 0, perm1, perm2, symm, out);
    });
}

int lt_btod_symmetrize2_destroy(lt_btod_symmetrize2 *s) {
    return guard([&]() {
        if(!s) return;
        lt_ctx *ctx = s->ctx;
        delete s->s;
        delete s->inner;
        delete s;
        ctx_release(ctx);
    });
}

int lt_btod_symmetrize2_perform(lt_btod_symmetrize2 *s, lt_btensor *b, int mode) {
    return guard([&]() {
        need(s, "op"); need(b, "b");
        lt_require(mode >= 0 && mode <= 2, LT_ERR_PARAM, "mode");
        s->s->perform(*b->t, mode);
    });
}

int lt_btod_symmetrize2_perform_add(lt_btod_symmetrize2 *s, lt_btensor *b, double d,
    int mode) {
    return guard([&]() {
        need(s, "op"); need(b, "b");
        lt_require(mode >= 0 && mode <= 2, LT_ERR_PARAM, "mode");
        s->s->perform_add(*b->t, d, mode);
    });
}

int lt_btod_symmetrize2_schedule(lt_btod_symmetrize2 *s, size_t *out, size_t cap,
    size_t *count) {
    return guard([&]() {
        need(s, "op"); need(count, "count");
        const std::vector<size_t> &v = s->s->schedule();
        *count = v.size();
        if(out) for(size_t i = 0; i < v.size() && i < cap; i++) out[i] = v[i];
    });
}

int lt_btod_symmetrize2_sym_orbit(lt_btod_symmetrize2 *s, size_t aidx, size_t *acidx,
    size_t *perm, double *coeff, int *allowed) {
    return guard([&]() {
        need(s, "op");
        const symmetry &y = s->s->sym();
        lt_require(aidx < y.bis.block_index_dims().size(), LT_ERR_PARAM, "block index");
        orbit o(y, aidx, true);
        if(acidx) *acidx = o.acidx;
        fill_transf(o.get_transf(aidx), y.bis.order(), perm, coeff);
        if(allowed) *allowed = o.allowed ? 1 : 0;
    });
}

int lt_btod_symmetrize2_stats(lt_btod_symmetrize2 *s, double *gpu_ms, double *gather_ms,
    double *gather_bytes, double *flops) {
    return guard([&]() {
        need(s, "op");
        if(gpu_ms) *gpu_ms = s->s->last_ms();
        if(gather_ms) *gather_ms = s->s->last_gather_ms();
        if(gather_bytes) *gather_bytes = s->s->last_gather_bytes();
        if(flops) *flops = s->s->flops();
    });
}

int lt_btod_symmetrize2_compute_block(lt_btod_symmetrize2 *s, int zero, size_t aidx_b,
    const size_t *perm, double coeff, double *blk) {
    return guard([&]() {
        need(s, "op"); need(blk, "blk");
        const symmetry &y = s->s->sym();
        const size_t n = y.bis.order();
        lt_require(aidx_b < y.bis.block_index_dims().size(), LT_ERR_PARAM, "block index");
        tensor_tr tr(perm_or_identity(n, perm), scalar_tr(coeff));
        s->s->compute_block(zero != 0, get_index(aidx_b, y.bis.block_index_dims()), tr, blk);
    });
}

int lt_btod_contract2_clst(lt_btod_contract2 *op, size_t aidx_c, double *out,
    size_t cap, size_t *npairs) {
    return guard([&]() {
        need(op, "op"); need(npairs, "npairs");
        const symmetry &s = op->op->symc();
        index_t ic = get_index(aidx_c, s.bis.block_index_dims());
        contr_list cl;
        op->op->clst(ic, cl);
        *npairs = cl.size();
        if(!out) return;
        size_t na = op->a->t->bis().order(), nb = op->b->t->bis().order();
        size_t w = 6 + na + nb, k = 0;
        for(contr_list::const_iterator p = cl.begin(); p != cl.end() && k < cap; ++p, ++k) {
            double *o = out + k * w;
            size_t j = 0;
            o[j++] = (double)p->aia;
            o[j++] = (double)p->acia;
            for(size_t i = 0; i < na; i++) o[j++] = (double)p->tra.perm[i];
            o[j++] = p->tra.st.c;
            o[j++] = (double)p->aib;
            o[j++] = (double)p->acib;
            for(size_t i = 0; i < nb; i++) o[j++] = (double)p->trb.perm[i];
            o[j++] = p->trb.st.c;
        }
    });
}

int lt_btod_contract2_compute_block_host(lt_btod_contract2 *op, size_t aidx_c, double *host) {
    return guard([&]() {
        need(op, "op"); need(host, "host");
        const symmetry &s = op->op->symc();
        op->op->compute_block_host(get_index(aidx_c, s.bis.block_index_dims()), host);
    });
}

int lt_btod_contract2_compute_block(lt_btod_contract2 *op, int zero, size_t aidx_c,
    const size_t *perm, double coeff, double *blk) {
    return guard([&]() {
        need(op, "op"); need(blk, "blk");
        const symmetry &s = op->op->symc();
        const size_t n = s.bis.order();
        tensor_tr tr(perm_or_identity(n, perm), scalar_tr(coeff));
        op->op->compute_block(zero != 0, get_index(aidx_c, s.bis.block_index_dims()), tr, blk);
    });
}

int lt_btod_contract2_stats(lt_btod_contract2 *op, lt_stats *out) {
    return guard([&]() {
        need(op, "op"); need(out, "out");
        const exec_stats &s = op->op->stats();
        out->t_schedule_ms = s.t_schedule_ms;
        out->t_plan_ms = s.t_plan_ms;
        out->t_gpu_ms = s.t_gpu_ms;
        out->t_gemm_ms = s.t_gemm_ms;
        out->flops = s.flops;
        out->bytes_perm = s.bytes_perm;
        out->n_cblocks = s.n_cblocks;
        out->n_pairs = s.n_pairs;
        out->n_tiles = s.n_tiles;
        out->n_unfold = s.n_unfold;
        out->n_launches = s.n_launches;
        out->tile_cfg = s.tile_cfg;
        out->ozaki_slices = s.ozaki_slices;
        out->ozaki_max_rel_err = s.ozaki_max_rel_err;
        for(int i = 0; i < 32; i++) out->tune_ms[i] = i < dev::TILE_COUNT ? s.tune_ms[i] : 0.0;
        out->graph = s.graph;
        out->tile_order = s.tile_order;
        out->n_plans = s.n_plans;
        out->n_batches = s.n_batches;
    });
}

int lt_btod_contract2_set_tile(lt_btod_contract2 *op, int cfg) {
    return guard([&]() {
        need(op, "op");
        lt_require(cfg >= -2 && cfg < dev::TILE_COUNT, LT_ERR_PARAM, "tile configuration");
        op->op->set_tile_cfg(cfg);
    });
}

int lt_btod_contract2_set_graph(lt_btod_contract2 *op, int on) {
    return guard([&]() {
        need(op, "op");
        op->op->set_graph(on != 0);
    });
}

int lt_btod_contract2_set_async(lt_btod_contract2 *op, int on) {
    return guard([&]() {
        need(op, "op");
        op->op->set_async(on != 0);
    });
}

int lt_btod_contract2_set_ozaki(lt_btod_contract2 *op, int slices) {
    return guard([&]() {
        need(op, "op");
        lt_require(slices >= 0 && slices <= 8, LT_ERR_PARAM, "slices");
        op->op->set_ozaki(slices);
    });
}

int lt_btod_contract2_set_ozaki_tile(lt_btod_contract2 *op, int tile) {
    return guard([&]() {
        need(op, "op");
        lt_require(tile >= -1 && tile < dev::OZ_TILE_COUNT, LT_ERR_PARAM, "tile");
        op->op->set_ozaki_tile(tile);
    });
}

// ---- btod_dotprod ----

int lt_btod_dotprod(lt_btensor *a, const size_t *perma, double ca, lt_btensor *b,
    const size_t *permb, double cb, double *result) {
    return guard([&]() {
        need(a, "a"); need(b, "b"); need(result, "result");
        size_t n = a->t->bis().order();
        lt_require(b->t->bis().order() == n, LT_ERR_DIMS, "btod_dotprod: orders differ");
        tensor_tr tra(perm_or_identity(n, perma), scalar_tr(ca));
        tensor_tr trb(perm_or_identity(n, permb), scalar_tr(cb));
        *result = btod_dotprod(*a->t, tra, *b->t, trb);
    });
}

int lt_btod_copy(lt_btensor *a, const size_t *perm, double c, lt_btensor *b, int add,
    double d) {
    return guard([&]() {
        need(a, "a"); need(b, "b");
        size_t n = a->t->bis().order();
        lt_require(b->t->bis().order() == n, LT_ERR_DIMS, "btod_copy: orders differ");
        btod_copy(*a->t, tensor_tr(perm_or_identity(n, perm), scalar_tr(c)), *b->t, add != 0, d);
    });
}

int lt_btod_add(size_t nops, lt_btensor *const *a, const size_t *perms, const double *c,
    lt_btensor *b, int add, double d) {
    return guard([&]() {
        need(a, "a"); need(c, "c"); need(b, "b");
        lt_require(nops > 0, LT_ERR_PARAM, "btod_add: no operands");
        const size_t n = b->t->bis().order();
        std::vector<btensor*> ops;
        std::vector<tensor_tr> trs;
        for(size_t i = 0; i < nops; i++) {
            need(a[i], "a[i]");
            lt_require(a[i]->t->bis().order() == n, LT_ERR_DIMS, "btod_add: orders differ");
            ops.push_back(a[i]->t);
            trs.push_back(tensor_tr(perm_or_identity(n, perms ? perms + i * n : 0),
                scalar_tr(c[i])));
        }
        btod_add(ops, trs, *b->t, add != 0, d);
    });
}

// ---- request loop ----

int lt_server_create(lt_ctx *ctx, size_t max_batch, size_t hbm_reserve, lt_server **out) {
    return guard([&]() {
        need(ctx, "ctx"); need(out, "out");
        lt_server *x = new lt_server;
        try {
            x->s = new server(*ctx->c, max_batch, hbm_reserve);
        } catch(...) {
            delete x;
            throw;
        }
        x->ctx = ctx;
        ctx_retain(ctx);
        *out = x;
    });
}

int lt_server_destroy(lt_server *srv) {
    return guard([&]() {
        if(!srv) return;
        lt_ctx *ctx = srv->ctx;
        delete srv->s;
        delete srv;
        ctx_release(ctx);
    });
}

int lt_server_submit(lt_server *srv, lt_btod_contract2 *op, lt_btensor *c, int add,
    double d, uint64_t *id) {
    return guard([&]() {
        need(srv, "srv"); need(op, "op"); need(c, "c"); need(id, "id");
        *id = srv->s->submit(op->op, c->t, add != 0, d);
    });
}

int lt_server_wait(lt_server *srv, uint64_t id, lt_request_stats *st) {
    request_stats r;
    int rc = guard([&]() {
        need(srv, "srv");
        srv->s->wait(id, r);
        if(st) {
            st->t_queue_ms = r.t_queue_ms;
            st->t_run_ms = r.t_run_ms;
            st->t_total_ms = r.t_total_ms;
            st->batch = r.batch;
            st->batch_size = r.batch_size;
            st->status = r.status;
            st->pad = 0;
        }
    });
    if(rc != LT_OK) return rc;
    if(r.status != LT_OK) {
        g_err = r.message;
        return r.status;
    }
    return LT_OK;
}

int lt_server_stats(lt_server *srv, uint64_t *batches, uint64_t *requests,
    double *mean_batch, uint64_t *max_batch) {
    return guard([&]() {
        need(srv, "srv");
        server_stats t;
        srv->s->totals(t);
        if(batches) *batches = t.batches;
        if(requests) *requests = t.requests;
        if(mean_batch) *mean_batch = t.mean_batch;
        if(max_batch) *max_batch = t.max_batch;
    });
}

} // extern "C"
