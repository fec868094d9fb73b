// btod_contract2 on the device.
//
// Host side (as libtensor, libtensor/gen_block_tensor/impl/
// gen_bto_contract2_impl.h:42-308 and gen_bto_contract2_batch_impl.h:177-291):
// result symmetry and non-zero schedule in the constructor; on perform the
// contraction is aligned to one matrix product per block pair
// (contraction2_align), A and B symmetries are permuted, per-C-block
// contraction lists are built and coalesced in the aligned space.
//
// Device side, one batch holding all blocks (HBM-resident tensors):
//   1. prepare: every aligned A/B block a pair needs (canonical blocks of the
//      permuted tensors and the non-canonical blocks of unfolded orbits) is
//      produced from the stored canonical block by ONE permute+scale, or used
//      in place when that transformation is the identity;
//   2. contract: grouped FP64 MFMA GEMM, one workgroup per C tile running the
//      whole contraction list of its C block (or one launch per C block);
//   3. write back: aligned C blocks that differ from the stored layout are
//      permuted into the canonical blocks of C.
#include "lt_runtime.h"

#include <functional>

#include <atomic>

#include <map>
#include <tuple>
#include <type_traits>

#include <chrono>
#include <cmath>
#include <numeric>
#include <unordered_map>

namespace lta {

namespace {

double now_ms() {
    return std::chrono::duration<double, std::milli>(
        std::chrono::steady_clock::now().time_since_epoch()).count();
}

//! Device permute descriptor for dst = c * P(src) with dims fusing.
dev::perm_desc make_perm_desc(const double *src, const dims_t &dsrc,
    const tensor_tr &tr, double *dst, bool accumulate) {

    const size_t n = dsrc.n;
    dims_t ddst(dsrc);
    ddst.permute(tr.perm);
    size_t sstr[kMaxOrder];
    for(size_t i = 0; i < n; i++) sstr[i] = dsrc.inc(tr.perm[i]);
    // fuse destination dims that are also adjacent, in order, in the source
    size_t fd[kMaxOrder], fs[kMaxOrder];
    size_t nf = 0;
    for(size_t i = 0; i < n; i++) {
        if(ddst[i] == 1) continue;
        if(nf > 0 && fs[nf - 1] == sstr[i] * ddst[i]) {
            fd[nf - 1] *= ddst[i];
            fs[nf - 1] = sstr[i];
        } else {
            fd[nf] = ddst[i];
            fs[nf] = sstr[i];
            nf++;
        }
    }
    if(nf == 0) { fd[0] = 1; fs[0] = 1; nf = 1; }
    lt_require(nf <= (size_t)dev::kPermMaxOrder, LT_ERR_INTERNAL, "permute order");
    size_t total = dsrc.size();
    lt_require(total < (size_t(1) << 31), LT_ERR_PARAM, "block larger than 2^31 elements");
    dev::perm_desc d;
    std::memset(&d, 0, sizeof(d));
    d.src = src;
    d.dst = dst;
    d.n = (uint32_t)nf;
    d.total = (uint32_t)total;
    for(size_t i = 0; i < (size_t)dev::kPermMaxOrder; i++) {
        d.div[i] = dev::make_fastdiv(i < nf ? (uint32_t)fd[i] : 1);
        d.sstride[i] = i < nf ? (uint32_t)fs[i] : 0;
    }
    d.alpha = tr.st.c;
    d.accumulate = accumulate ? 1 : 0;
    // transposing: the source's unit-stride dimension is not the
    // destination's fastest one
    d.tile_p = -1;
    d.tsq = 0;
    d.dps = d.tdp = d.tdq = dev::make_fastdiv(1);
    if(nf >= 2 && fs[nf - 1] != 1) {
        for(size_t i = 0; i + 1 < nf; i++) {
            if(fs[i] != 1) continue;
            size_t st = 1;
            for(size_t j = i + 1; j < nf; j++) st *= fd[j];
            d.tile_p = (int32_t)i;
            d.tsq = (uint32_t)fs[nf - 1];
            d.dps = dev::make_fastdiv((uint32_t)st);
            d.tdp = dev::make_fastdiv((uint32_t)fd[i]);
            d.tdq = dev::make_fastdiv((uint32_t)fd[nf - 1]);
            break;
        }
    }
    return d;
}

//! Device-side copy of host vectors in one allocation (owned: freed on
//! release, on reassignment and on destruction; not copyable).
struct dev_lists {
    context *ctx;
    void *mem;
    size_t bytes;
    dev_lists() : ctx(0), mem(0), bytes(0) { }
    dev_lists(const dev_lists&) = delete;
    dev_lists &operator=(const dev_lists&) = delete;
    ~dev_lists() { release(); }
    void release() {
        if(ctx && mem) ctx->free(mem);
        mem = 0;
        bytes = 0;
    }
    //! Replaces the allocation with a new one of `n` bytes.
    void *reset(context &c, size_t n) {
        release();
        ctx = &c;
        mem = c.alloc(n);
        bytes = n;
        return mem;
    }
};

size_t align256(size_t x) { return (x + 255) & ~size_t(255); }

//! Largest block of a block index space (product of the largest segment
//! of each dimension).
double max_block_elems(const bis_t &bis) {
    double n = 1.0;
    const dims_t bid = bis.block_index_dims();
    for(size_t i = 0; i < bis.order(); i++) {
        size_t m = 0;
        for(size_t b = 0; b < bid[i]; b++) m = std::max(m, bis.block_len(i, b));
        n *= double(m);
    }
    return n;
}

} // unnamed namespace

//! The GEMM and split-precision kernels address operand and C blocks with
//! 32-bit byte offsets from a block's base address: every block of A, B and
//! C must hold fewer than 2^29 elements (4 GiB).
void btod_contract2::check_block_sizes() const {
    const double lim = double(size_t(1) << 29);
    lt_require(max_block_elems(m_a.bis()) < lim && max_block_elems(m_b.bis()) < lim &&
        max_block_elems(m_symc.bis) < lim, LT_ERR_PARAM,
        "btod_contract2: a block of 2^29 or more elements (the kernels address blocks "
        "with 32-bit byte offsets; split such blocks further)");
}

//! Work items of a list of permute descriptors: chunks of kPermChunk
//! destination elements, or the corners of 32 x 32 tiles for transposing
//! descriptors.
void make_chunks(const std::vector<dev::perm_desc> &ds, std::vector<uint32_t> &cd,
    std::vector<uint32_t> &co) {
    cd.clear();
    co.clear();
    for(size_t i = 0; i < ds.size(); i++) {
        const dev::perm_desc &d = ds[i];
        if(d.tile_p < 0) {
            for(uint32_t off = 0; off < d.total; off += dev::kPermChunk) {
                cd.push_back((uint32_t)i);
                co.push_back(off);
            }
            continue;
        }
        // corners: every index of the other dimensions, p and q in steps of
        // the tile size
        const uint32_t n = d.n, p = (uint32_t)d.tile_p, q = n - 1;
        uint32_t dims[dev::kPermMaxOrder], step[dev::kPermMaxOrder], idx[dev::kPermMaxOrder];
        uint32_t str[dev::kPermMaxOrder];
        uint32_t acc = 1;
        for(int j = (int)n - 1; j >= 0; j--) {
            dims[j] = d.div[j].d;
            str[j] = acc;
            acc *= dims[j];
            step[j] = (uint32_t(j) == p || uint32_t(j) == q) ? dev::kPermTile : 1;
            idx[j] = 0;
        }
        while(true) {
            uint32_t e = 0;
            for(uint32_t j = 0; j < n; j++) e += idx[j] * str[j];
            cd.push_back((uint32_t)i | dev::kPermTileItem);
            co.push_back(e);
            int j = (int)n - 1;
            for(; j >= 0; j--) {
                idx[j] += step[j];
                if(idx[j] < dims[j]) break;
                idx[j] = 0;
            }
            if(j < 0) break;
        }
    }
}

dev::perm_desc make_perm_desc_public(const double *src, const dims_t &dsrc,
    const tensor_tr &tr, double *dst, bool accumulate) {
    return make_perm_desc(src, dsrc, tr, dst, accumulate);
}

//! Execution plan cached by the operation (valid while the structure of A,
//! B and C does not change).
struct contract_plan {
    uint64_t va, vb, cuid;
    btensor *c;
    bool add;
    double d;
    bool empty;

    // host-side work lists
    std::vector<dev::perm_desc> prep, wback;
    std::vector<dev::gemm_cblock> cblocks;
    std::vector<dev::gemm_pair> pairs;
    std::vector<dev::gemm_tile> tiles_sorted, tiles_by_block;
    std::vector<uint32_t> block_tile_begin;
    std::vector<double> cb_ksum;                      // sum of k over a C block's pairs
    std::vector<size_t> c_out;                        // canonical C blocks written
    bool batch;                                       // built for a batch of the schedule
    std::vector<std::pair<double*, size_t> > c_out_mem; // their storage
    bool ta, tb;
    int tile_cfg, tile_order;

    // device copies
    dev_lists lists, tiles;
    // stacks (C blocks sharing their B operands) for the STACK_* kernels
    std::vector<dev::stack_desc> stacks;
    std::vector<dev::stack_pair> spairs;
    std::vector<dev::stack_arow> arows;
    std::vector<dev::stack_crow> crows;
    std::vector<double> stack_ksum;
    std::vector<dev::gemm_tile> stiles;
    std::vector<dev::var_tile> vtiles;
    // offset tables of operands read in place by the stacked kernels
    std::vector<int32_t> otab;
    std::vector<std::pair<int64_t, int64_t> > spair_tab;  // (koff, noff) indexes or -1
    std::vector<dev::perm_desc> prep_st;      // prepare list of the stacked launches
    uint32_t n_prep_st_chunks;
    dev::perm_desc *d_prep_st;
    uint32_t *d_pst_cd, *d_pst_co;
    dev_lists stk, stk_tiles;
    dev::stack_desc *d_stacks;
    dev::stack_pair *d_spairs;
    dev::stack_arow *d_arows;
    dev::stack_crow *d_crows;
    dev::gemm_tile *d_stiles;
    dev::var_tile *d_vtiles;
    // C blocks reading the same B operands (same order and layouts) share a
    // group; n_xcd: XCDs of the device
    std::vector<size_t> cb_group;
    size_t n_groups;
    int n_xcd;
    dev::perm_desc *d_prep, *d_wback;
    uint32_t *d_prep_cd, *d_prep_co, *d_wb_cd, *d_wb_co;
    uint32_t n_prep_chunks, n_wb_chunks;
    dev::gemm_cblock *d_cb;
    dev::gemm_pair *d_pairs;
    dev::gemm_tile *d_tiles_sorted, *d_tiles_block;
    double *scratch;
    size_t scratch_elems;
    context *ctx;

    hipEvent_t ev[4];
    bool events;
    hipGraph_t graph;           // grouped perform, captured on first use
    hipGraphExec_t gexec;
    size_t graph_launches;

    // split-precision form (built on the first perform in that mode)
    int oz_slices;
    int oz_tile;                      //!< dev::oz_tile_cfg of the product launch
    int oz_tile_set;                  //!< the requested shape it was built for (-1 measured)
    dev_lists oz_lists, oz_mem;
    dev::oz_split_desc *d_oz_split;
    uint32_t *d_oz_items;
    uint32_t n_oz_items;
    dev::oz_pair *d_oz_pairs;
    dev::gemm_tile *d_oz_tiles;
    uint32_t n_oz_tiles;
    uint32_t n_oz_items_mut;          //!< split items of operands split every perform
    // prepare of the split-precision performs: the copies of operands of
    // mutable tensors only (the others are prepared once, with their split)
    bool oz_prep_own;
    dev::perm_desc *d_oz_prep;
    uint32_t *d_oz_prep_cd, *d_oz_prep_co;
    uint32_t n_oz_prep_chunks;
    double *d_oz_bound;               //!< stated error per C block (bound, max |C|)
    int32_t *d_oz_opmax;              //!< largest exponent per split operand
    uint32_t n_oz_ops;

    contract_plan() : batch(false), va(0), vb(0), cuid(0), c(0), add(false), d(1.0), empty(false),
        ta(false), tb(false), tile_cfg(0), tile_order(0), d_prep(0), d_wback(0), d_prep_cd(0),
        d_prep_co(0), d_wb_cd(0), d_wb_co(0), n_prep_chunks(0), n_wb_chunks(0),
        d_cb(0), d_pairs(0), d_tiles_sorted(0), d_tiles_block(0), scratch(0),
        scratch_elems(0), ctx(0), events(false), graph(0), gexec(0), graph_launches(0),
        d_stacks(0), d_spairs(0), d_arows(0), d_crows(0), d_stiles(0), d_vtiles(0),
        n_groups(0), n_xcd(8),
        n_prep_st_chunks(0), d_prep_st(0), d_pst_cd(0), d_pst_co(0),
        oz_slices(0), oz_tile(0), oz_tile_set(-1), d_oz_split(0), d_oz_items(0), n_oz_items(0), d_oz_pairs(0),
        d_oz_tiles(0), n_oz_tiles(0), n_oz_items_mut(0), oz_prep_own(false), d_oz_prep(0),
        d_oz_prep_cd(0), d_oz_prep_co(0), n_oz_prep_chunks(0), d_oz_bound(0), d_oz_opmax(0),
        n_oz_ops(0) { }
    ~contract_plan() {
        // an asynchronous perform of this plan may still run
        if(ctx && ctx->is_device()) {
            ctx->activate();
            (void)hipStreamSynchronize(ctx->stream());
        }
        if(gexec) hipGraphExecDestroy(gexec);
        if(graph) hipGraphDestroy(graph);
        if(ctx && scratch) ctx->free(scratch);
        if(events) for(int i = 0; i < 4; i++) hipEventDestroy(ev[i]);
    }
};

//! The grouped GEMM launch of configuration k over a tile list of its own
//! (n tiles at d_list: the autotuner's sample of a large plan).
hipError_t launch_grouped_list(const contract_plan &pl, int k, const void *d_list, uint32_t n,
    hipStream_t s) {
    if(dev::is_var(k))
        return dev::launch_gemm_stacked_var(k, pl.d_stacks, pl.d_spairs, pl.d_arows, pl.d_crows,
            (const dev::var_tile*)d_list, n, s);
    if(dev::is_stacked(k))
        return dev::launch_gemm_stacked(k, pl.d_stacks, pl.d_spairs, pl.d_arows, pl.d_crows,
            (const dev::gemm_tile*)d_list, n, s);
    return dev::launch_gemm(k, pl.d_cb, pl.d_pairs, (const dev::gemm_tile*)d_list, n, n, s);
}

//! The grouped GEMM launch of a plan in configuration k.
hipError_t launch_grouped(const contract_plan &pl, int k, hipStream_t s) {
    if(dev::is_var(k))
        return dev::launch_gemm_stacked_var(k, pl.d_stacks, pl.d_spairs, pl.d_arows, pl.d_crows,
            pl.d_vtiles, (uint32_t)pl.vtiles.size(), s);
    if(dev::is_stacked(k))
        return dev::launch_gemm_stacked(k, pl.d_stacks, pl.d_spairs, pl.d_arows, pl.d_crows,
            pl.d_stiles, (uint32_t)pl.stiles.size(), s);
    uint32_t nt = (uint32_t)pl.tiles_sorted.size();
    return dev::launch_gemm(k, pl.d_cb, pl.d_pairs, pl.d_tiles_sorted, nt, nt, s);
}

// ---------------------------------------------------------------------------

uint64_t next_op_uid() {
    static std::atomic<uint64_t> next(1);
    return next++;
}

btod_contract2::btod_contract2(const contraction2 &c, btensor &a, double ka,
    btensor &b, double kb, double kc) :
    m_contr(c), m_a(a), m_b(b), m_ka(ka), m_kb(kb), m_kc(kc), m_uid(next_op_uid()),
    m_tile_cfg(-2), m_tuned_cfg(-1), m_tuned_order(-1), m_tuned_va(0), m_tuned_vb(0), m_use_graph(true), m_async(false), m_times_pending(false), m_times_graph(false),
    m_oz_slices(0), m_oz_tile(-1),
    m_add_uid(0), m_add_ver(0), m_add_direct(false), m_sched_subset(0),
    m_use_graph_saved(true) {

    lt_require(c.is_complete(), LT_ERR_PARAM, "contraction is incomplete");
    lt_require(a.bis().order() == c.NA() && b.bis().order() == c.NB(),
        LT_ERR_DIMS, "tensor orders do not match the contraction");
    std::memset(&m_stats, 0, sizeof(m_stats));
    double t0 = now_ms();
    bool self = (&a == &b) && (c.N == c.M);
    m_symc = contract2_symmetry(c, a.sym(), b.sym(), self);
    m_sched = contract2_schedule(c, a.sym(), b.sym(), m_symc,
        a.nonzero_blocks(), b.nonzero_blocks(), a.ctx().nthreads());
    m_stats.t_schedule_ms = now_ms() - t0;
}

btod_contract2::~btod_contract2() { }

void btod_contract2::clst(const index_t &ic, contr_list &out) {
    dims_t bida = m_a.bidims(), bidb = m_b.bidims();
    block_list bla(bida, m_a.nonzero_blocks()), blb(bidb, m_b.nonzero_blocks());
    block_list blax(bida), blbx(bidb);
    unfold_block_list(m_a.sym(), bla, blax);
    unfold_block_list(m_b.sym(), blb, blbx);
    contract2_block_list cbl(m_contr, blax, blbx);
    orbit_cache oca(m_a.sym()), ocb(m_b.sym());
    build_clst(m_contr, oca, ocb, blax, blbx, cbl, ic, out);
}

void btod_contract2::set_tile_cfg(int cfg) {
    if(cfg != m_tile_cfg) drop_plan();
    m_tile_cfg = cfg;
}

void btod_contract2::drop_plan() {
    m_times_pending = false;
    m_plan.reset();
}

void btod_contract2::finish_times() {
    if(!m_times_pending) return;
    m_times_pending = false;
    if(!m_plan) return;
    contract_plan &pl = *m_plan;
    m_a.ctx().activate();
    hip_check(hipEventSynchronize(pl.ev[3]), "hipEventSynchronize");
    float ms = 0.f, msg = 0.f;
    hip_check(hipEventElapsedTime(&ms, pl.ev[0], pl.ev[3]), "hipEventElapsedTime");
    if(!m_times_graph)
        hip_check(hipEventElapsedTime(&msg, pl.ev[1], pl.ev[2]), "hipEventElapsedTime");
    m_stats.t_gpu_ms = ms;
    m_stats.t_gemm_ms = msg;
}

bool btod_contract2::plan_valid(btensor &c, bool add, double d, int) const {
    if(m_plan && m_plan->batch) return false;
    if(!m_plan) return false;
    return m_plan->va == m_a.version() && m_plan->vb == m_b.version() &&
        m_plan->cuid == c.uid() && m_plan->add == add && m_plan->d == d;
}

// ---------------------------------------------------------------------------

void btod_contract2::build_plan(btensor &c, bool add, double dscale, bool measure) {
    m_times_pending = false;
    double t0 = now_ms();
    m_stats.n_plans++;
    context &ctx = m_a.ctx();
    m_plan.reset(new contract_plan());
    contract_plan &pl = *m_plan;
    pl.va = m_a.version();
    pl.vb = m_b.version();
    pl.c = &c;
    pl.cuid = c.uid();
    pl.add = add;
    pl.d = dscale;
    pl.ctx = &ctx;

    // the whole schedule, or one batch of it (memory-bounded performs)
    const std::vector<size_t> &sched = m_sched_subset ? *m_sched_subset : m_sched;
    pl.batch = m_sched_subset != 0;
    std::vector<size_t> nza = m_a.nonzero_blocks(), nzb = m_b.nonzero_blocks();
    if(nza.empty() || nzb.empty() || sched.empty()) {
        pl.empty = true;
        return;
    }

    const contraction2_align al(m_contr);
    const perm_t &perma = al.perma, &permb = al.permb, &permc = al.permc;
    perm_t permcinv(permc.inverse());
    contraction2 contr(m_contr);
    contr.permute_a(perma);
    contr.permute_b(permb);
    contr.permute_c(permc);

    symmetry syma2 = so_permute(m_a.sym(), perma);
    symmetry symb2 = so_permute(m_b.sym(), permb);
    symmetry symct = so_permute(m_symc, permc);
    const bis_t &bisa2 = syma2.bis, &bisb2 = symb2.bis, &bisct = symct.bis;
    dims_t bida = m_a.bidims(), bidb = m_b.bidims(), bidc = m_symc.bis.block_index_dims();
    dims_t bida2 = bisa2.block_index_dims(), bidb2 = bisb2.block_index_dims();
    dims_t bidct = bisct.block_index_dims();

    // canonical non-zero blocks of the permuted arguments (the batches)
    std::vector<size_t> bla2, blb2, batchc;
    for(size_t i = 0; i < nza.size(); i++) {
        if(perma.is_identity()) { bla2.push_back(nza[i]); continue; }
        index_t ia = get_index(nza[i], bida);
        perma.apply(ia);
        bla2.push_back(short_orbit(syma2, abs_index(ia, bida2), true).acidx);
    }
    for(size_t i = 0; i < nzb.size(); i++) {
        if(permb.is_identity()) { blb2.push_back(nzb[i]); continue; }
        index_t ib = get_index(nzb[i], bidb);
        permb.apply(ib);
        blb2.push_back(short_orbit(symb2, abs_index(ib, bidb2), true).acidx);
    }
    for(size_t i = 0; i < sched.size(); i++) {
        index_t ic = get_index(sched[i], bidc);
        permc.apply(ic);
        batchc.push_back(short_orbit(symct, abs_index(ic, bidct), true).acidx);
    }
    block_list bla(bida2, bla2), blb(bidb2, blb2);
    block_list blax(bida2), blbx(bidb2);
    unfold_block_list(syma2, bla, blax);
    unfold_block_list(symb2, blb, blbx);
    contract2_block_list cbl(contr, blax, blbx);
    orbit_cache oca(syma2), ocb(symb2);

    // contraction lists of all C blocks
    std::vector<contr_list> clsts(batchc.size());
    parallel_for(batchc.size(), ctx.nthreads(), [&](size_t t) {
        build_clst(contr, oca, ocb, blax, blbx, cbl, get_index(batchc[t], bidct), clsts[t]);
    });

    // aligned operand blocks needed by the pairs
    std::vector<size_t> usea, useb;
    for(size_t t = 0; t < clsts.size(); t++) {
        for(contr_list::const_iterator p = clsts[t].begin(); p != clsts[t].end(); ++p) {
            usea.push_back(p->aia);
            useb.push_back(p->aib);
        }
    }
    std::sort(usea.begin(), usea.end());
    usea.erase(std::unique(usea.begin(), usea.end()), usea.end());
    std::sort(useb.begin(), useb.end());
    useb.erase(std::unique(useb.begin(), useb.end()), useb.end());

    // Where each aligned block comes from: stored canonical block Y and the
    // transformation (A orbit of Y -> idx) then perma then (syma2 orbit).
    struct src_t { size_t y; tensor_tr tr; };
    auto source_of = [&](btensor &t, const symmetry &sym2, const perm_t &perm,
        const dims_t &bid, const dims_t &bid2, size_t z) -> src_t {
        orbit oz(sym2, z, false);
        size_t x = oz.acidx;
        tensor_tr tz = oz.get_transf(z);
        index_t io = get_index(x, bid2);
        perm.inverse().apply(io);
        size_t ao = abs_index(io, bid);
        orbit oa(t.sym(), ao, false);
        src_t s;
        s.y = oa.acidx;
        s.tr = oa.get_transf(ao);
        s.tr.transform(tensor_tr(perm, scalar_tr()));
        s.tr.transform(tz);
        return s;
    };

    // Operand layouts of the aligned product (the same for every pair):
    // A stored m x k (or k x m when ta), B stored k x n (or n x k when tb).
    gemm_form gf0 = aligned_gemm_form(contr,
        bisa2.block_dims(get_index(usea[0], bida2)),
        bisb2.block_dims(get_index(useb[0], bidb2)));

    // How the kernels read an aligned operand block: straight from the
    // stored canonical block when the transformation is a scalar, or a
    // scalar and the exchange of the two matrix index groups (the kernel
    // then reads the operand in the other layout); from a prepared copy in
    // scratch otherwise.
    struct opref { const double *p; double scale; bool flip; uint32_t ldf; };
    auto fold = [&](btensor &t, const dims_t &bid, const src_t &s, const dims_t &dz,
        size_t nfirst, opref &r) -> bool {
        r.p = t.block_ptr(s.y);
        r.scale = 1.0; r.flip = false; r.ldf = 0;
        if(s.tr.is_identity()) return true;
        dims_t dy = t.bis().block_dims(get_index(s.y, bid));
        dev::perm_desc pd = make_perm_desc(r.p, dy, s.tr, 0, false);
        size_t first = 1;
        for(size_t i = 0; i < nfirst; i++) first *= dz[i];
        if(pd.n == 1 && pd.sstride[0] == 1) {
            r.scale = pd.alpha;
            return true;
        }
        if(pd.n == 2 && pd.sstride[0] == 1 && pd.sstride[1] == pd.div[0].d &&
            pd.div[0].d == first) {
            r.scale = pd.alpha;
            r.flip = true;
            r.ldf = (uint32_t)first;
            return true;
        }
        return false;
    };

    // The operand in the kernels' B slot (B, or A when C is computed
    // transposed) can also be read in place through offset tables by the
    // stacked kernels: element (k, n) of the aligned block is
    // Y[koff[k] + noff[n]] of its stored block Y.  The offset of an element
    // is linear in each index, so it splits into a row part and a column
    // part of the aligned matrix.  Where each part is linear in at most two
    // groups of indices, the kernels compute the offsets instead of reading
    // them: off(e) = (e / d) s0 + (e % d) s1.
    const bool tc = gf0.tc;
    struct linform { uint32_t d; int32_t s0, s1; };
    struct tabref {
        const double *y; double alpha; int64_t koff, noff;
        bool arith; linform kf, nf;
    };
    auto make_tables = [&](btensor &t, const dims_t &bid, const src_t &s, const dims_t &dz,
        size_t nfirst, bool first_is_k, tabref &r) {
        dims_t dy = t.bis().block_dims(get_index(s.y, bid));
        size_t n = dy.n;
        std::vector<int64_t> ys(n), w(n);
        int64_t acc = 1;
        for(size_t d = n; d-- > 0;) { ys[d] = acc; acc *= (int64_t)dy[d]; }
        for(size_t i = 0; i < n; i++) w[i] = ys[s.tr.perm[i]];
        auto table = [&](size_t i0, size_t i1) {
            int64_t at = (int64_t)pl.otab.size();
            size_t cnt = 1;
            for(size_t i = i0; i < i1; i++) cnt *= dz[i];
            std::vector<size_t> idx(n, 0);
            for(size_t e = 0; e < cnt; e++) {
                int64_t off = 0;
                for(size_t i = i0; i < i1; i++) off += (int64_t)idx[i] * w[i];
                pl.otab.push_back((int32_t)off);
                for(size_t i = i1; i-- > i0;) {
                    if(++idx[i] < dz[i]) break;
                    idx[i] = 0;
                }
            }
            return at;
        };
        // groups of consecutive indices with one stride (extent-1 indices
        // dropped); at most two groups give the linear form
        auto linear = [&](size_t i0, size_t i1, int64_t at, linform &f) {
            std::vector<std::pair<int64_t, int64_t> > g;    // (extent, stride)
            int64_t cnt = 1;
            for(size_t i = i0; i < i1; i++) {
                cnt *= (int64_t)dz[i];
                if(dz[i] == 1) continue;
                if(!g.empty() && g.back().second == w[i] * (int64_t)dz[i]) {
                    g.back().first *= (int64_t)dz[i];
                    g.back().second = w[i];
                } else g.push_back(std::make_pair((int64_t)dz[i], w[i]));
            }
            int64_t d = 2, s0 = 0, s1 = 0;
            if(g.size() == 1) { d = g[0].first; s1 = g[0].second; }
            else if(g.size() == 2) { d = g[1].first; s0 = g[0].second; s1 = g[1].second; }
            else if(g.size() > 2) return false;
            if(cnt >= (int64_t(1) << 30) || d >= (int64_t(1) << 30)) return false;
            f.d = (uint32_t)d; f.s0 = (int32_t)s0; f.s1 = (int32_t)s1;
            for(int64_t e = 0; e < cnt; e++)
                lt_require((e / d) * s0 + (e % d) * s1 == pl.otab[at + e], LT_ERR_INTERNAL,
                    "offset table is not linear in its index groups");
            return true;
        };
        int64_t rt = table(0, nfirst), ct = table(nfirst, n);
        linform rf, cf;
        bool rl = linear(0, nfirst, rt, rf), cl = linear(nfirst, n, ct, cf);
        r.y = t.block_ptr(s.y);
        r.alpha = s.tr.st.c;
        r.koff = first_is_k ? rt : ct;
        r.noff = first_is_k ? ct : rt;
        r.arith = rl && cl;
        r.kf = first_is_k ? rf : cf;
        r.nf = first_is_k ? cf : rf;
    };
    std::unordered_map<size_t, tabref> taba, tabb;

    size_t scratch = 0;
    std::unordered_map<size_t, size_t> offa, offb;    // aligned block -> scratch offset
    std::unordered_map<size_t, opref> refa, refb;
    std::vector<std::pair<size_t, src_t> > prepa, prepb;
    for(size_t i = 0; i < usea.size(); i++) {
        src_t s = source_of(m_a, syma2, perma, bida, bida2, usea[i]);
        lt_require(m_a.is_nonzero(s.y), LT_ERR_INTERNAL, "pair uses a zero block of A");
        dims_t dz = bisa2.block_dims(get_index(usea[i], bida2));
        opref r;
        if(!fold(m_a, bida, s, dz, gf0.ta ? contr.K : contr.N, r)) {
            offa[usea[i]] = scratch;
            scratch += (dz.size() + 31) & ~size_t(31);
            prepa.push_back(std::make_pair(usea[i], s));
            r.p = 0; r.scale = 1.0; r.flip = false; r.ldf = 0;
            if(tc) make_tables(m_a, bida, s, dz, gf0.ta ? contr.K : contr.N, gf0.ta, taba[usea[i]]);
        }
        refa[usea[i]] = r;
    }
    for(size_t i = 0; i < useb.size(); i++) {
        src_t s = source_of(m_b, symb2, permb, bidb, bidb2, useb[i]);
        lt_require(m_b.is_nonzero(s.y), LT_ERR_INTERNAL, "pair uses a zero block of B");
        dims_t dz = bisb2.block_dims(get_index(useb[i], bidb2));
        opref r;
        if(!fold(m_b, bidb, s, dz, gf0.tb ? contr.M : contr.K, r)) {
            offb[useb[i]] = scratch;
            scratch += (dz.size() + 31) & ~size_t(31);
            prepb.push_back(std::make_pair(useb[i], s));
            r.p = 0; r.scale = 1.0; r.flip = false; r.ldf = 0;
            if(!tc) make_tables(m_b, bidb, s, dz, gf0.tb ? contr.M : contr.K, !gf0.tb, tabb[useb[i]]);
        }
        refb[useb[i]] = r;
    }

    // C blocks: aligned canonical block -> stored canonical block of C
    struct cdst_t { size_t yc; tensor_tr tr; bool direct; size_t off; };
    std::vector<cdst_t> cdst(batchc.size());
    for(size_t t = 0; t < batchc.size(); t++) {
        index_t ib = get_index(batchc[t], bidct);
        permcinv.apply(ib);
        size_t aib = abs_index(ib, bidc);
        orbit ob(m_symc, aib, false);
        tensor_tr tr(permcinv, scalar_tr());
        tr.transform(ob.get_transf(aib).inverse());
        cdst[t].yc = ob.acidx;
        cdst[t].tr = tr;
        cdst[t].direct = tr.is_identity();
        if(!cdst[t].direct) {
            cdst[t].off = scratch;
            scratch += (bisct.block_dims(get_index(batchc[t], bidct)).size() + 31) & ~size_t(31);
        }
    }

    if(scratch > 0) {
        pl.scratch = (double*)ctx.alloc(scratch * sizeof(double));
        pl.scratch_elems = scratch;
    }
    for(std::unordered_map<size_t, size_t>::iterator i = offa.begin(); i != offa.end(); ++i)
        refa[i->first].p = pl.scratch + i->second;
    for(std::unordered_map<size_t, size_t>::iterator i = offb.begin(); i != offb.end(); ++i)
        refb[i->first].p = pl.scratch + i->second;

    // 1. prepare descriptors
    for(size_t i = 0; i < prepa.size(); i++) {
        const src_t &s = prepa[i].second;
        dims_t dy = m_a.bis().block_dims(get_index(s.y, bida));
        pl.prep.push_back(make_perm_desc(m_a.block_ptr(s.y), dy, s.tr,
            pl.scratch + offa[prepa[i].first], false));
    }
    for(size_t i = 0; i < prepb.size(); i++) {
        const src_t &s = prepb[i].second;
        dims_t dy = m_b.bis().block_dims(get_index(s.y, bidb));
        pl.prep.push_back(make_perm_desc(m_b.block_ptr(s.y), dy, s.tr,
            pl.scratch + offb[prepb[i].first], false));
    }
    // the stacked launches read the kernel-B operands through their tables
    if(!tc) pl.prep_st.assign(pl.prep.begin(), pl.prep.begin() + prepa.size());
    else pl.prep_st.assign(pl.prep.begin() + prepa.size(), pl.prep.end());

    // 2. GEMM work lists
    // with C stored as (j, i) the product is computed as C^T = B^T A^T
    pl.ta = tc ? !gf0.tb : gf0.ta;
    pl.tb = tc ? !gf0.ta : gf0.tb;

    double flops = 0.0;
    std::vector<tabref> pair_tab;
    std::vector<std::pair<size_t, size_t> > mn(batchc.size());
    for(size_t t = 0; t < batchc.size(); t++) {
        index_t ic = get_index(batchc[t], bidct);
        dims_t dc = bisct.block_dims(ic);
        size_t m = 1, n = 1;
        // C aligned: (A outer..., B outer...) or swapped when tc
        size_t na = contr.N;
        for(size_t i = 0; i < dc.n; i++) {
            bool froma = (contr.conn[i] - contr.NC()) < contr.NA();
            (void)na;
            if(froma) m *= dc[i]; else n *= dc[i];
        }
        mn[t] = std::make_pair(m, n);
    }

    std::vector<double> cb_ksum;
    for(size_t t = 0; t < batchc.size(); t++) {
        const contr_list &cl = clsts[t];
        dev::gemm_cblock cb;
        std::memset(&cb, 0, sizeof(cb));
        size_t m = mn[t].first, n = mn[t].second;
        size_t rows = tc ? n : m, cols = tc ? m : n;
        cb.m = (uint32_t)rows;
        cb.n = (uint32_t)cols;
        cb.ldc = (uint32_t)cols;
        double *cptr;
        if(cdst[t].direct) {
            cptr = c.block_ptr(cdst[t].yc);
        } else {
            cptr = pl.scratch + cdst[t].off;
        }
        cb.c = cptr;
        cb.zero = (cdst[t].direct && add) ? 0 : 1;
        cb.pair_begin = (uint32_t)pl.pairs.size();
        double ksum = 0.0;
        for(contr_list::const_iterator p = cl.begin(); p != cl.end(); ++p) {
            // scalars as gen_bto_contract2_block::compute_block:
            // inverse orbit transformation of the unfolded block, then the
            // pair's transformation, then ka/kb/kc
            orbit_info oa = oca.get(p->aia), ob = ocb.get(p->aib);
            scalar_tr sa(oa.tr.st); sa.invert(); sa.transform(p->tra.st);
            scalar_tr sb(ob.tr.st); sb.invert(); sb.transform(p->trb.st);
            scalar_tr ka(sa), kb(sb), kc(m_kc);
            ka.transform(scalar_tr(m_ka));
            kb.transform(scalar_tr(m_kb));
            double d = ka.c * kb.c * kc.c;
            if(add) d *= dscale;
            if(d == 0.0) continue;
            dims_t da = bisa2.block_dims(get_index(p->aia, bida2));
            dims_t db = bisb2.block_dims(get_index(p->aib, bidb2));
            gemm_form g = aligned_gemm_form(contr, da, db);
            dev::gemm_pair gp;
            std::memset(&gp, 0, sizeof(gp));
            const opref &ra = refa[p->aia], &rb = refb[p->aib];
            gp.k = (uint32_t)g.k;
            // layout and row stride of A and B as read
            bool ta = g.ta != ra.flip, tb = g.tb != rb.flip;
            size_t lda = ra.flip ? ra.ldf : (g.ta ? g.m : g.k);
            size_t ldb = rb.flip ? rb.ldf : (g.tb ? g.k : g.n);
            if(!tc) {
                gp.a = ra.p; gp.b = rb.p; gp.lda = (uint32_t)lda; gp.ldb = (uint32_t)ldb;
                gp.flags = (ta ? dev::kPairAKMajor : 0) | (tb ? dev::kPairBNMajor : 0);
            } else {
                // C^T = B^T A^T
                gp.a = rb.p; gp.b = ra.p; gp.lda = (uint32_t)ldb; gp.ldb = (uint32_t)lda;
                gp.flags = (!tb ? dev::kPairAKMajor : 0) | (!ta ? dev::kPairBNMajor : 0);
            }
            gp.d = d * ra.scale * rb.scale;
            if(gp.d == 0.0) continue;
            pl.pairs.push_back(gp);
            {
                // kernel-B operand through tables (stacked launches)
                const std::unordered_map<size_t, tabref> &tm = tc ? taba : tabb;
                std::unordered_map<size_t, tabref>::const_iterator ti = tm.find(tc ? p->aia : p->aib);
                if(ti != tm.end()) pair_tab.push_back(ti->second);
                else {
                    tabref z;
                    std::memset(&z, 0, sizeof(z));
                    z.y = 0; z.alpha = 1.0; z.koff = z.noff = -1; z.arith = false;
                    pair_tab.push_back(z);
                }
            }
            flops += 2.0 * double(g.m) * double(g.n) * double(g.k);
            ksum += double(g.k);
        }
        cb.pair_end = (uint32_t)pl.pairs.size();
        pl.cblocks.push_back(cb);
        cb_ksum.push_back(ksum);
        pl.c_out.push_back(cdst[t].yc);
        pl.c_out_mem.push_back(std::make_pair(c.block_ptr(cdst[t].yc), c.block_size(cdst[t].yc)));
    }

    pl.cb_ksum = cb_ksum;

    // Pairs of every C block in a fixed order of their B operands, so that C
    // blocks reading the same B operands list them identically; such blocks
    // form a stack (rows concatenated, each B chunk read once for all).
    typedef std::tuple<uintptr_t, uint32_t, uint32_t, uint32_t> bkey_t;
    auto bkey = [](const dev::gemm_pair &g) {
        return bkey_t((uintptr_t)g.b, g.flags, g.ldb, g.k);
    };
    for(size_t b = 0; b < pl.cblocks.size(); b++) {
        uint32_t p0 = pl.cblocks[b].pair_begin, p1 = pl.cblocks[b].pair_end;
        std::vector<uint32_t> ord(p1 - p0);
        std::iota(ord.begin(), ord.end(), p0);
        std::stable_sort(ord.begin(), ord.end(), [&](uint32_t x, uint32_t y) {
            return bkey(pl.pairs[x]) < bkey(pl.pairs[y]); });
        std::vector<dev::gemm_pair> gs(ord.size());
        std::vector<tabref> ts(ord.size());
        for(size_t i = 0; i < ord.size(); i++) { gs[i] = pl.pairs[ord[i]]; ts[i] = pair_tab[ord[i]]; }
        std::copy(gs.begin(), gs.end(), pl.pairs.begin() + p0);
        std::copy(ts.begin(), ts.end(), pair_tab.begin() + p0);
    }
    {
        std::map<std::pair<uint32_t, std::vector<bkey_t> >, std::vector<size_t> > groups;
        std::vector<std::pair<uint32_t, std::vector<bkey_t> > > order;
        for(size_t b = 0; b < pl.cblocks.size(); b++) {
            const dev::gemm_cblock &cb = pl.cblocks[b];
            std::pair<uint32_t, std::vector<bkey_t> > key;
            key.first = cb.n;
            for(uint32_t p = cb.pair_begin; p < cb.pair_end; p++) key.second.push_back(bkey(pl.pairs[p]));
            std::vector<size_t> &g = groups[key];
            if(g.empty()) order.push_back(key);
            g.push_back(b);
        }
        for(size_t gi = 0; gi < order.size(); gi++) {
            const std::vector<size_t> &bl = groups[order[gi]];
            const dev::gemm_cblock &c0 = pl.cblocks[bl[0]];
            dev::stack_desc sd;
            std::memset(&sd, 0, sizeof(sd));
            sd.n = c0.n;
            sd.npairs = c0.pair_end - c0.pair_begin;
            sd.pair_begin = (uint32_t)pl.spairs.size();
            sd.arow_begin = (uint32_t)pl.arows.size();
            sd.crow_begin = (uint32_t)pl.crows.size();
            size_t rows = 0;
            for(size_t q = 0; q < bl.size(); q++) rows += pl.cblocks[bl[q]].m;
            lt_require(rows < (size_t(1) << 20) && pl.arows.size() + rows * sd.npairs < (size_t(1) << 31),
                LT_ERR_PARAM, "stack too large");
            sd.rows = (uint32_t)rows;
            double ks = 0.0;
            for(uint32_t j = 0; j < sd.npairs; j++) {
                const dev::gemm_pair &g = pl.pairs[c0.pair_begin + j];
                const tabref &tb = pair_tab[c0.pair_begin + j];
                dev::stack_pair sp;
                std::memset(&sp, 0, sizeof(sp));
                sp.b = g.b; sp.k = g.k; sp.ldb = g.ldb; sp.flags = g.flags;
                double bscale = 1.0;
                if(tb.koff >= 0) {
                    sp.b = tb.y;
                    bscale = tb.alpha;    // the copy would have carried it
                    if(tb.arith) {
                        sp.flags |= dev::kPairOffArith;
                        sp.kd = dev::make_fastdiv(tb.kf.d);
                        sp.ks0 = tb.kf.s0; sp.ks1 = tb.kf.s1;
                        sp.nd = dev::make_fastdiv(tb.nf.d);
                        sp.ns0 = tb.nf.s0; sp.ns1 = tb.nf.s1;
                    }
                }
                // two-k loads: k-contiguous rows (A) or columns (B), even
                // extent, 16-byte aligned rows; A and B independently
                if((g.k % 2) == 0) {
                    bool ok = !(g.flags & dev::kPairAKMajor);
                    for(size_t q = 0; q < bl.size() && ok; q++) {
                        const dev::gemm_pair &gq = pl.pairs[pl.cblocks[bl[q]].pair_begin + j];
                        ok = (gq.lda % 2) == 0 && ((uintptr_t)gq.a % 16) == 0;
                    }
                    if(ok) sp.flags |= dev::kPairVecA;
                    if(tb.koff < 0 && (g.flags & dev::kPairBNMajor) && (g.ldb % 2) == 0 &&
                        ((uintptr_t)g.b % 16) == 0) sp.flags |= dev::kPairVecB;
                    if(tb.koff >= 0 && ((uintptr_t)tb.y % 16) == 0) {
                        bool pairs_ok = true;
                        for(uint32_t kk = 0; kk + 1 < g.k && pairs_ok; kk += 2)
                            pairs_ok = pl.otab[tb.koff + kk] % 2 == 0 &&
                                pl.otab[tb.koff + kk + 1] == pl.otab[tb.koff + kk] + 1;
                        for(uint32_t n = 0; n < sd.n && pairs_ok; n++)
                            pairs_ok = pl.otab[tb.noff + n] % 2 == 0;
                        if(pairs_ok) sp.flags |= dev::kPairVecBT;
                    }
                }
                pl.spairs.push_back(sp);
                pl.spair_tab.push_back(std::make_pair(tb.koff, tb.noff));
                ks += double(g.k);
                sd.nchunk += (g.k + 15) / 16;
                for(size_t q = 0; q < bl.size(); q++) {
                    const dev::gemm_cblock &cq = pl.cblocks[bl[q]];
                    const dev::gemm_pair &gq = pl.pairs[cq.pair_begin + j];
                    for(uint32_t r = 0; r < cq.m; r++) {
                        dev::stack_arow ar;
                        std::memset(&ar, 0, sizeof(ar));
                        ar.p = (gq.flags & dev::kPairAKMajor) ? gq.a + r : gq.a + size_t(r) * gq.lda;
                        ar.d = gq.d * bscale;
                        ar.lda = gq.lda;
                        pl.arows.push_back(ar);
                    }
                }
            }
            for(size_t q = 0; q < bl.size(); q++) {
                const dev::gemm_cblock &cq = pl.cblocks[bl[q]];
                for(uint32_t r = 0; r < cq.m; r++) {
                    dev::stack_crow cr;
                    std::memset(&cr, 0, sizeof(cr));
                    cr.c = cq.c + size_t(r) * cq.ldc;
                    cr.zero = (uint32_t)cq.zero;
                    pl.crows.push_back(cr);
                }
            }
            pl.stacks.push_back(sd);
            pl.stack_ksum.push_back(ks);
        }
        dev::stack_desc z;
        std::memset(&z, 0, sizeof(z));
        pl.stacks.push_back(z);      // empty stack of the padding tiles
        pl.stack_ksum.push_back(0.0);
    }

    // Tile lists of one kernel configuration: tiles grouped by C block (one
    // launch per block) and the same tiles longest-first (grouped launch).
    // Grouped launch orders (the autotuner measures both): 0, tiles longest
    // first; 1, XCD-aware.  Workgroups are dealt to the XCDs round-robin
    // (workgroup w runs on XCD w % nx), and each XCD has its own L2: C
    // blocks whose pairs read the same B blocks form one group, groups are
    // assigned to XCDs longest-processing-time first, and each XCD's tiles
    // (groups longest first, a group's tiles together so that they share A
    // rows and B columns in that XCD's L2) go to positions x, x + nx, ...
    // Positions left over at the end hold empty tiles (a C block with no
    // rows).
    int nx = 8;
    {
        int v = 0;
        ctx.activate();
        if(hipDeviceGetAttribute(&v, hipDeviceAttributeNumberOfXccs, ctx.device()) == hipSuccess &&
            v >= 1 && v <= 64) nx = v;
        else hipGetLastError();
    }
    std::vector<size_t> cb_group(pl.cblocks.size());
    size_t ngroups = 0;
    {
        std::map<std::vector<std::pair<const double*, uint32_t> >, size_t> gid;
        for(size_t b = 0; b < pl.cblocks.size(); b++) {
            std::vector<std::pair<const double*, uint32_t> > key;
            for(uint32_t p = pl.cblocks[b].pair_begin; p < pl.cblocks[b].pair_end; p++)
                key.push_back(std::make_pair(pl.pairs[p].b, pl.pairs[p].flags));
            std::sort(key.begin(), key.end());
            std::map<std::vector<std::pair<const double*, uint32_t> >, size_t>::iterator i =
                gid.find(key);
            if(i == gid.end()) i = gid.insert(std::make_pair(key, ngroups++)).first;
            cb_group[b] = i->second;
        }
    }
    pl.cb_group = cb_group;
    pl.n_groups = ngroups;
    pl.n_xcd = nx;
    const uint32_t null_cb = (uint32_t)pl.cblocks.size();
    {
        dev::gemm_cblock z;
        std::memset(&z, 0, sizeof(z));
        z.zero = 1;
        pl.cblocks.push_back(z);     // empty C block of the padding tiles
    }

    // Tile lists of one kernel configuration: tiles grouped by C block (one
    // launch per block) and the grouped-launch order above.
    auto make_tiles = [&](int cfg, int order, std::vector<dev::gemm_tile> &by_block,
        std::vector<uint32_t> &begin, std::vector<dev::gemm_tile> &sorted) {
        // C-block tiles of the kernel that runs them (stacked configurations
        // run C-block lists with their unstacked twin)
        uint32_t bm = 64, bn = 64;
        dev::tile_shape(dev::unstacked(cfg), bm, bn);
        by_block.clear();
        begin.clear();
        std::vector<double> gcost(ngroups, 0.0);
        for(size_t b = 0; b < null_cb; b++) {
            const dev::gemm_cblock &cb = pl.cblocks[b];
            begin.push_back((uint32_t)by_block.size());
            uint32_t tm = (cb.m + bm - 1) / bm, tn = (cb.n + bn - 1) / bn;
            lt_require(tm < 65536 && tn < 65536, LT_ERR_PARAM, "C block too large");
            for(uint32_t i = 0; i < tm; i++) {
                for(uint32_t j = 0; j < tn; j++) {
                    dev::gemm_tile tl;
                    tl.cblock = (uint32_t)b;
                    tl.tm = (uint16_t)i;
                    tl.tn = (uint16_t)j;
                    by_block.push_back(tl);
                    gcost[cb_group[b]] += cb_ksum[b];
                }
            }
        }
        begin.push_back((uint32_t)by_block.size());
        begin.push_back((uint32_t)by_block.size());     // the empty C block
        lt_require(by_block.size() < (size_t(1) << 30), LT_ERR_PARAM, "too many tiles");
        if(order == 0) {
            // longest (largest k sum) first over the whole device
            std::vector<size_t> idx(by_block.size());
            std::iota(idx.begin(), idx.end(), size_t(0));
            std::stable_sort(idx.begin(), idx.end(), [&](size_t x, size_t y) {
                return cb_ksum[by_block[x].cblock] > cb_ksum[by_block[y].cblock]; });
            sorted.resize(idx.size());
            for(size_t i = 0; i < idx.size(); i++) sorted[i] = by_block[idx[i]];
            return;
        }
        // groups to XCDs, longest first onto the least loaded
        std::vector<size_t> gorder(ngroups);
        std::iota(gorder.begin(), gorder.end(), size_t(0));
        std::stable_sort(gorder.begin(), gorder.end(), [&](size_t x, size_t y) {
            return gcost[x] > gcost[y]; });
        std::vector<double> load(nx, 0.0);
        std::vector<std::vector<size_t> > xg(nx);
        for(size_t i = 0; i < gorder.size(); i++) {
            size_t x = std::min_element(load.begin(), load.end()) - load.begin();
            xg[x].push_back(gorder[i]);
            load[x] += gcost[gorder[i]];
        }
        std::vector<std::vector<size_t> > gblocks(ngroups);
        for(size_t b = 0; b < null_cb; b++) gblocks[cb_group[b]].push_back(b);
        std::vector<std::vector<dev::gemm_tile> > q(nx);
        size_t qmax = 0;
        for(int x = 0; x < nx; x++) {
            for(size_t i = 0; i < xg[x].size(); i++) {
                const std::vector<size_t> &bl = gblocks[xg[x][i]];
                for(size_t j = 0; j < bl.size(); j++)
                    for(uint32_t t = begin[bl[j]]; t < begin[bl[j] + 1]; t++)
                        q[x].push_back(by_block[t]);
            }
            qmax = std::max(qmax, q[x].size());
        }
        dev::gemm_tile empty;
        empty.cblock = null_cb; empty.tm = 0; empty.tn = 0;
        sorted.assign(qmax * nx, empty);
        for(int x = 0; x < nx; x++)
            for(size_t l = 0; l < q[x].size(); l++) sorted[l * nx + x] = q[x][l];
    };

    // Kernel configuration: fixed, or the one wasting least padding, or
    // measured on this plan (autotuner, below, after the lists are on the
    // device).
    int cfg = m_tile_cfg;
    std::vector<int> candidates;
    if(cfg >= 0) {
        candidates.push_back(cfg);
    } else {
        std::vector<std::pair<double, int> > est;
        for(int k = 0; k < dev::TILE_COUNT; k++) {
            uint32_t tbm, tbn;
            dev::tile_shape(k, tbm, tbn);
            double waste = 0.0, ntile = 0.0;
            for(size_t b = 0; b < null_cb; b++) {
                size_t tm = (pl.cblocks[b].m + tbm - 1) / tbm;
                size_t tn = (pl.cblocks[b].n + tbn - 1) / tbn;
                waste += double(tm * tbm * tn * tbn) * cb_ksum[b];
                ntile += double(tm * tn);
            }
            // padded work plus a per-tile overhead equivalent to 32x32 MACs
            double kmean = 0.0;
            for(size_t b = 0; b < cb_ksum.size(); b++) kmean += cb_ksum[b];
            if(!cb_ksum.empty()) kmean /= double(cb_ksum.size());
            est.push_back(std::make_pair(waste + ntile * 1024.0 * kmean, k));
        }
        std::sort(est.begin(), est.end());
        // the autotuner measures the stacked configurations only: on every
        // work list measured they beat the per-C-block kernels in grouped
        // launches, and each candidate costs two timed launches
        for(size_t i = 0; i < est.size(); i++)
            if(cfg != -2 || dev::is_stacked(est[i].second)) candidates.push_back(est[i].second);
        if(cfg == -1 || add || !measure) candidates.resize(1);
        // the configuration measured for these operands (same pairs, same
        // work list for any result tensor) is reused without measuring again
        if(cfg == -2 && m_tuned_cfg >= 0 && m_tuned_va == pl.va && m_tuned_vb == pl.vb)
            candidates.assign(1, m_tuned_cfg);
        else if(add && m_tuned_cfg >= 0) candidates.assign(1, m_tuned_cfg);
    }
    cfg = candidates[0];

    // 3. write-back descriptors
    for(size_t t = 0; t < batchc.size(); t++) {
        if(cdst[t].direct) continue;
        dims_t dc = bisct.block_dims(get_index(batchc[t], bidct));
        double *dst = c.block_ptr(cdst[t].yc);
        bool acc = add;
        pl.wback.push_back(make_perm_desc(pl.scratch + cdst[t].off, dc, cdst[t].tr, dst, acc));
    }

    // upload all lists in one allocation
    std::vector<uint32_t> pcd, pco, wcd, wco, scd, sco;
    make_chunks(pl.prep, pcd, pco);
    make_chunks(pl.wback, wcd, wco);
    make_chunks(pl.prep_st, scd, sco);
    pl.n_prep_chunks = (uint32_t)pcd.size();
    pl.n_wb_chunks = (uint32_t)wcd.size();
    pl.n_prep_st_chunks = (uint32_t)scd.size();
    size_t sz = 0;
    size_t o_prep = sz; sz = align256(sz + pl.prep.size() * sizeof(dev::perm_desc));
    size_t o_wb = sz; sz = align256(sz + pl.wback.size() * sizeof(dev::perm_desc));
    size_t o_pcd = sz; sz = align256(sz + pcd.size() * 4);
    size_t o_pco = sz; sz = align256(sz + pco.size() * 4);
    size_t o_wcd = sz; sz = align256(sz + wcd.size() * 4);
    size_t o_wco = sz; sz = align256(sz + wco.size() * 4);
    size_t o_cb = sz; sz = align256(sz + pl.cblocks.size() * sizeof(dev::gemm_cblock));
    size_t o_pr = sz; sz = align256(sz + pl.pairs.size() * sizeof(dev::gemm_pair));
    size_t o_st = sz; sz = align256(sz + pl.stacks.size() * sizeof(dev::stack_desc));
    size_t o_sp = sz; sz = align256(sz + pl.spairs.size() * sizeof(dev::stack_pair));
    size_t o_ar = sz; sz = align256(sz + pl.arows.size() * sizeof(dev::stack_arow));
    size_t o_cr = sz; sz = align256(sz + pl.crows.size() * sizeof(dev::stack_crow));
    size_t o_ot = sz; sz = align256(sz + pl.otab.size() * sizeof(int32_t));
    size_t o_ps = sz; sz = align256(sz + pl.prep_st.size() * sizeof(dev::perm_desc));
    size_t o_scd = sz; sz = align256(sz + scd.size() * 4);
    size_t o_sco = sz; sz = align256(sz + sco.size() * 4);
    std::vector<char> host(sz, 0);
    auto put = [&](size_t off, const void *src, size_t n) { if(n) std::memcpy(&host[off], src, n); };
    put(o_prep, pl.prep.data(), pl.prep.size() * sizeof(dev::perm_desc));
    put(o_wb, pl.wback.data(), pl.wback.size() * sizeof(dev::perm_desc));
    put(o_pcd, pcd.data(), pcd.size() * 4);
    put(o_pco, pco.data(), pco.size() * 4);
    put(o_wcd, wcd.data(), wcd.size() * 4);
    put(o_wco, wco.data(), wco.size() * 4);
    put(o_cb, pl.cblocks.data(), pl.cblocks.size() * sizeof(dev::gemm_cblock));
    put(o_pr, pl.pairs.data(), pl.pairs.size() * sizeof(dev::gemm_pair));
    put(o_st, pl.stacks.data(), pl.stacks.size() * sizeof(dev::stack_desc));
    put(o_ot, pl.otab.data(), pl.otab.size() * sizeof(int32_t));
    put(o_ps, pl.prep_st.data(), pl.prep_st.size() * sizeof(dev::perm_desc));
    put(o_scd, scd.data(), scd.size() * 4);
    put(o_sco, sco.data(), sco.size() * 4);
    put(o_ar, pl.arows.data(), pl.arows.size() * sizeof(dev::stack_arow));
    put(o_cr, pl.crows.data(), pl.crows.size() * sizeof(dev::stack_crow));
    pl.lists.reset(ctx, sz);
    {
        // stack pairs with their offset tables at their device addresses
        const int32_t *ot = (const int32_t*)((char*)pl.lists.mem + o_ot);
        for(size_t j = 0; j < pl.spairs.size(); j++) {
            if(pl.spair_tab[j].first < 0) continue;
            pl.spairs[j].koff = ot + pl.spair_tab[j].first;
            pl.spairs[j].noff = ot + pl.spair_tab[j].second;
        }
        put(o_sp, pl.spairs.data(), pl.spairs.size() * sizeof(dev::stack_pair));
    }
    ctx.upload_new(pl.lists.mem, host.data(), sz);
    char *base = (char*)pl.lists.mem;
    pl.d_prep = (dev::perm_desc*)(base + o_prep);
    pl.d_wback = (dev::perm_desc*)(base + o_wb);
    pl.d_prep_cd = (uint32_t*)(base + o_pcd);
    pl.d_prep_co = (uint32_t*)(base + o_pco);
    pl.d_wb_cd = (uint32_t*)(base + o_wcd);
    pl.d_wb_co = (uint32_t*)(base + o_wco);
    pl.d_cb = (dev::gemm_cblock*)(base + o_cb);
    pl.d_pairs = (dev::gemm_pair*)(base + o_pr);
    pl.d_stacks = (dev::stack_desc*)(base + o_st);
    pl.d_spairs = (dev::stack_pair*)(base + o_sp);
    pl.d_arows = (dev::stack_arow*)(base + o_ar);
    pl.d_crows = (dev::stack_crow*)(base + o_cr);
    pl.d_prep_st = (dev::perm_desc*)(base + o_ps);
    pl.d_pst_cd = (uint32_t*)(base + o_scd);
    pl.d_pst_co = (uint32_t*)(base + o_sco);

    // tile lists live in their own allocation (replaced by the autotuner)
    auto upload_tiles = [&](int k, int order) {
        make_tiles(k, order, pl.tiles_by_block, pl.block_tile_begin, pl.tiles_sorted);
        // the sorted list can be longer than the per-block list (padding
        // tiles of the XCD-aware order)
        size_t nt = pl.tiles_sorted.size() * sizeof(dev::gemm_tile);
        size_t nb = pl.tiles_by_block.size() * sizeof(dev::gemm_tile);
        size_t tsz = align256(nt) + align256(nb) + 256;
        pl.tiles.reset(ctx, tsz);
        std::vector<char> th(tsz, 0);
        if(nt) std::memcpy(&th[0], pl.tiles_sorted.data(), nt);
        if(nb) std::memcpy(&th[align256(nt)], pl.tiles_by_block.data(), nb);
        ctx.upload_new(pl.tiles.mem, th.data(), tsz);
        pl.d_tiles_sorted = (dev::gemm_tile*)pl.tiles.mem;
        pl.d_tiles_block = (dev::gemm_tile*)((char*)pl.tiles.mem + align256(nt));
        pl.tile_cfg = k;
        pl.tile_order = order;
        pl.stiles.clear();
        pl.vtiles.clear();
        if(dev::is_var(k)) {
            // full 64-row tiles and one tile per row / column remainder,
            // largest (k sum x area) first
            const uint32_t fnx = (k == dev::STACK_VAR_64x48_W4) ? 3 : 2;
            std::vector<std::pair<double, dev::var_tile> > vt;
            for(size_t si = 0; si + 1 < pl.stacks.size(); si++) {
                const dev::stack_desc &sd = pl.stacks[si];
                for(uint32_t r0 = 0; r0 < sd.rows; r0 += 64) {
                    uint32_t fm = std::min<uint32_t>(4, (sd.rows - r0 + 15) / 16);
                    for(uint32_t c0 = 0; c0 < sd.n; c0 += 16 * fnx) {
                        uint32_t fn = std::min<uint32_t>(fnx, (sd.n - c0 + 15) / 16);
                        lt_require(r0 / 16 < 65536 && c0 / 16 < 65536, LT_ERR_PARAM, "stack too large");
                        dev::var_tile t;
                        t.stack = (uint32_t)si;
                        t.r16 = (uint16_t)(r0 / 16); t.c16 = (uint16_t)(c0 / 16);
                        t.fm = (uint16_t)fm; t.fn = (uint16_t)fn;
                        vt.push_back(std::make_pair(pl.stack_ksum[si] * fm * fn, t));
                    }
                }
            }
            std::stable_sort(vt.begin(), vt.end(),
                [](const std::pair<double, dev::var_tile> &x, const std::pair<double, dev::var_tile> &y) {
                    return x.first > y.first; });
            for(size_t i = 0; i < vt.size(); i++) pl.vtiles.push_back(vt[i].second);
            size_t nb = pl.vtiles.size() * sizeof(dev::var_tile);
            pl.stk_tiles.reset(ctx, nb ? nb : 256);
            if(nb) ctx.upload_new(pl.stk_tiles.mem, pl.vtiles.data(), nb);
            pl.d_vtiles = (dev::var_tile*)pl.stk_tiles.mem;
        } else if(dev::is_stacked(k)) {
            // stack tiles, longest (largest k sum) first
            uint32_t bm = 64, bn = 64;
            dev::tile_shape(k, bm, bn);
            std::vector<std::pair<double, dev::gemm_tile> > st;
            for(size_t si = 0; si < pl.stacks.size(); si++) {
                uint32_t tm = (pl.stacks[si].rows + bm - 1) / bm, tn = (pl.stacks[si].n + bn - 1) / bn;
                lt_require(tm < 65536 && tn < 65536, LT_ERR_PARAM, "stack too large");
                for(uint32_t i = 0; i < tm; i++) for(uint32_t j = 0; j < tn; j++) {
                    dev::gemm_tile t;
                    t.cblock = (uint32_t)si; t.tm = (uint16_t)i; t.tn = (uint16_t)j;
                    st.push_back(std::make_pair(pl.stack_ksum[si], t));
                }
            }
            std::stable_sort(st.begin(), st.end(),
                [](const std::pair<double, dev::gemm_tile> &x, const std::pair<double, dev::gemm_tile> &y) {
                    return x.first > y.first; });
            if(order == 0) {
                for(size_t i = 0; i < st.size(); i++) pl.stiles.push_back(st[i].second);
            } else {
                // XCD-aware: stacks to XCDs longest-processing-time first, each
                // XCD's tiles longest first at positions x, x + nx, ...; the
                // padding tiles name the empty stack (no rows)
                const size_t ns = pl.stacks.size() - 1;
                std::vector<double> scost(ns, 0.0);
                for(size_t i = 0; i < st.size(); i++) scost[st[i].second.cblock] += st[i].first;
                std::vector<size_t> so(ns);
                std::iota(so.begin(), so.end(), size_t(0));
                std::stable_sort(so.begin(), so.end(), [&](size_t x, size_t y) {
                    return scost[x] > scost[y]; });
                std::vector<double> load(nx, 0.0);
                std::vector<int> sx(ns, 0);
                for(size_t i = 0; i < ns; i++) {
                    int x = int(std::min_element(load.begin(), load.end()) - load.begin());
                    sx[so[i]] = x;
                    load[x] += scost[so[i]];
                }
                std::vector<std::vector<dev::gemm_tile> > q(nx);
                size_t qmax = 0;
                for(size_t i = 0; i < st.size(); i++) {
                    std::vector<dev::gemm_tile> &qq = q[sx[st[i].second.cblock]];
                    qq.push_back(st[i].second);
                    qmax = std::max(qmax, qq.size());
                }
                dev::gemm_tile empty;
                empty.cblock = (uint32_t)ns; empty.tm = 0; empty.tn = 0;
                pl.stiles.assign(qmax * nx, empty);
                for(int x = 0; x < nx; x++)
                    for(size_t l = 0; l < q[x].size(); l++) pl.stiles[l * nx + x] = q[x][l];
            }
            size_t nb = pl.stiles.size() * sizeof(dev::gemm_tile);
            pl.stk_tiles.reset(ctx, nb ? nb : 256);
            if(nb) ctx.upload_new(pl.stk_tiles.mem, pl.stiles.data(), nb);
            pl.d_stiles = (dev::gemm_tile*)pl.stk_tiles.mem;
        }
    };
    int order = (cfg == m_tuned_cfg && m_tuned_order >= 0) ? m_tuned_order : 0;
    upload_tiles(cfg, order);

    ctx.activate();
    for(int i = 0; i < 4; i++) hip_check(hipEventCreate(&pl.ev[i]), "hipEventCreate");
    pl.events = true;

    // Autotuner: time every candidate configuration on this plan's own work
    // list (prepare permute and grouped launch), keep the fastest.  Results of these
    // launches are overwritten by the perform that follows (never run for
    // the accumulating path).
    if(candidates.size() > 1) {
        hipStream_t s = ctx.stream();
        if(pl.n_prep_chunks)
            hip_check(dev::launch_permute(pl.d_prep, pl.d_prep_cd, pl.d_prep_co,
                pl.n_prep_chunks, s), "permute (prepare)");
        // each candidate is timed with the prepare permute it needs (the
        // stacked kernels read B operands in place and prepare less)
        auto prep_for = [&](int k) -> hipError_t {
            if(dev::is_stacked(k))
                return pl.n_prep_st_chunks ? dev::launch_permute(pl.d_prep_st, pl.d_pst_cd,
                    pl.d_pst_co, pl.n_prep_st_chunks, s) : hipSuccess;
            return pl.n_prep_chunks ? dev::launch_permute(pl.d_prep, pl.d_prep_cd,
                pl.d_prep_co, pl.n_prep_chunks, s) : hipSuccess;
        };
        // A large plan is timed on a sample of each candidate's launch list:
        // every stride-th group of nx consecutive tiles (a group keeps its
        // XCD placement), about kTuneFlops of work.  The prepare permute is
        // timed whole, the sampled launch scaled by the list's cost over the
        // sample's (cost of a tile: the k sum of its block or stack times its
        // fragment count).
        const double kTuneFlops = 2e11;
        const size_t stride = flops > 2.0 * kTuneFlops ? size_t(flops / kTuneFlops) : 1;
        dev_lists tune_list;
        auto take = [&](const auto &list, auto cost, double &scale) -> uint32_t {
            typedef typename std::decay<decltype(list[0])>::type tile_t;
            std::vector<tile_t> sub;
            double tot = 0.0, part = 0.0;
            for(size_t i = 0; i < list.size(); i++) {
                const double c = cost(list[i]);
                tot += c;
                if((i / nx) % stride == 0) { sub.push_back(list[i]); part += c; }
            }
            scale = part > 0.0 ? tot / part : 1.0;
            const size_t nb = sub.size() * sizeof(tile_t);
            tune_list.reset(ctx, nb ? nb : 256);
            if(nb) ctx.upload_new(tune_list.mem, sub.data(), nb);
            return (uint32_t)sub.size();
        };
        auto ksum_of = [](const std::vector<double> &ks, uint32_t b) {
            return b < ks.size() ? ks[b] : 0.0; };
        auto sample = [&](int k, double &scale) -> uint32_t {
            if(dev::is_var(k))
                return take(pl.vtiles, [&](const dev::var_tile &t) {
                    return ksum_of(pl.stack_ksum, t.stack) * t.fm * t.fn; }, scale);
            if(dev::is_stacked(k))
                return take(pl.stiles, [&](const dev::gemm_tile &t) {
                    return ksum_of(pl.stack_ksum, t.cblock); }, scale);
            return take(pl.tiles_sorted, [&](const dev::gemm_tile &t) {
                return ksum_of(cb_ksum, t.cblock); }, scale);
        };
        float best = 1e30f;
        int bestk = cfg, besto = 0;
        for(size_t ci = 0; ci < 2 * candidates.size(); ci++) {
            int k = candidates[ci / 2], o = int(ci % 2);
            upload_tiles(k, o);
            double scale = 1.0;
            const uint32_t ntune = stride > 1 ? sample(k, scale) : 0;
            // one timed launch; a short one (first-launch overheads are
            // then not negligible) is timed again
            float ms = 0.f;
            for(int rep = 0; rep < 2; rep++) {
                float mp = 0.f, mg = 0.f;
                hip_check(hipEventRecord(pl.ev[0], s), "hipEventRecord");
                hip_check(prep_for(k), "autotune (prepare)");
                hip_check(hipEventRecord(pl.ev[1], s), "hipEventRecord");
                hip_check(stride > 1 ? launch_grouped_list(pl, k, tune_list.mem, ntune, s) :
                    launch_grouped(pl, k, s), "autotune");
                hip_check(hipEventRecord(pl.ev[2], s), "hipEventRecord");
                hip_check(hipStreamSynchronize(s), "hipStreamSynchronize");
                hip_check(hipEventElapsedTime(&mp, pl.ev[0], pl.ev[1]), "hipEventElapsedTime");
                hip_check(hipEventElapsedTime(&mg, pl.ev[1], pl.ev[2]), "hipEventElapsedTime");
                ms = mp + float(mg * scale);
                if(mp + mg > 20.f) break;
            }
            if(o == 0 || ms < m_stats.tune_ms[k]) m_stats.tune_ms[k] = ms;
            if(ms < best) { best = ms; bestk = k; besto = o; }
        }
        cfg = bestk;
        upload_tiles(cfg, besto);
        m_tuned_cfg = cfg;
        m_tuned_order = besto;
        m_tuned_va = pl.va;
        m_tuned_vb = pl.vb;
    }

    m_stats.n_cblocks = null_cb;
    m_stats.n_pairs = pl.pairs.size();
    m_stats.n_tiles = pl.tiles_by_block.size();
    m_stats.n_unfold = pl.prep.size();
    m_stats.flops = flops;
    m_stats.tile_cfg = cfg;
    m_stats.tile_order = pl.tile_order;
    double bytes = 0.0;
    for(size_t i = 0; i < pl.prep.size(); i++) bytes += 16.0 * pl.prep[i].total;
    for(size_t i = 0; i < pl.wback.size(); i++)
        bytes += (pl.wback[i].accumulate ? 24.0 : 16.0) * pl.wback[i].total;
    m_stats.bytes_perm = bytes;
    m_stats.t_plan_ms = now_ms() - t0;
}

void btod_contract2::build_ozaki(int slices) {
    contract_plan &pl = *m_plan;
    context &ctx = m_a.ctx();
    pl.oz_lists.release();
    pl.oz_mem.release();
    pl.oz_prep_own = false;
    pl.oz_slices = slices;

    // Prepared copies (one descriptor each, each in its own scratch region)
    // of blocks of an immutable tensor never change: they are prepared and
    // split once, with this plan's split form; the performs prepare the
    // others only.
    auto immutable_src = [&](const double *p) {
        return (m_a.is_immutable() && m_a.owns(p)) || (m_b.is_immutable() && m_b.owns(p));
    };
    std::vector<std::pair<const double*, size_t> > inv_copies;    // (dst, elements), sorted
    std::vector<dev::perm_desc> prep_mut, prep_inv;
    for(size_t i = 0; i < pl.prep.size(); i++) {
        if(immutable_src(pl.prep[i].src)) {
            prep_inv.push_back(pl.prep[i]);
            inv_copies.push_back(std::make_pair((const double*)pl.prep[i].dst,
                size_t(pl.prep[i].total)));
        } else prep_mut.push_back(pl.prep[i]);
    }
    std::sort(inv_copies.begin(), inv_copies.end());
    auto in_invariant_copy = [&](const double *p) {
        std::vector<std::pair<const double*, size_t> >::const_iterator i = std::upper_bound(
            inv_copies.begin(), inv_copies.end(), std::make_pair(p, ~size_t(0)));
        if(i == inv_copies.begin()) return false;
        --i;
        return p >= i->first && p < i->first + i->second;
    };

    // one split per distinct operand matrix: A-role operands scaled per row
    // of C, B-role operands per column of C
    struct op_info { const double *src; uint32_t rows, k, kp, rs, ks; size_t off, eoff; bool cached; };
    std::vector<op_info> ops;
    typedef std::map<std::pair<const double*, bool>, size_t> op_index;
    op_index ida, idb;
    std::vector<size_t> pa(pl.pairs.size()), pb(pl.pairs.size());
    size_t bytes = 0, nexp = 0;
    auto add_op = [&](op_index &idx, const double *src,
        uint32_t rows, uint32_t k, uint32_t ld, bool kmajor) {
        op_index::iterator i = idx.find(std::make_pair(src, kmajor));
        if(i != idx.end()) {
            lt_require(ops[i->second].rows == rows && ops[i->second].k == k, LT_ERR_INTERNAL,
                "operand used with two shapes");
            return i->second;
        }
        op_info o;
        o.src = src; o.rows = rows; o.k = k; o.kp = (k + 63) & ~63u;
        o.rs = kmajor ? 1 : ld;
        o.ks = kmajor ? ld : 1;
        o.off = bytes; o.eoff = nexp;
        // an operand read in place from an immutable tensor, or an invariant
        // copy, is split once per plan, the others on every perform
        o.cached = immutable_src(src) || in_invariant_copy(src);
        bytes += size_t(slices) * ((rows + 15) & ~15u) * o.kp;
        bytes = (bytes + 255) & ~size_t(255);
        nexp += rows;
        ops.push_back(o);
        idx[std::make_pair(src, kmajor)] = ops.size() - 1;
        return ops.size() - 1;
    };
    for(size_t b = 0; b < pl.cblocks.size(); b++) {
        const dev::gemm_cblock &cb = pl.cblocks[b];
        for(uint32_t p = cb.pair_begin; p < cb.pair_end; p++) {
            const dev::gemm_pair &gp = pl.pairs[p];
            // int32 accumulation of up to `slices` products of K terms of
            // magnitude <= 127 * 127 must not overflow
            lt_require(double(gp.k) * 16129.0 * slices < 2147483647.0, LT_ERR_PARAM,
                "contraction extent too large for the split-precision mode");
            pa[p] = add_op(ida, gp.a, cb.m, gp.k, gp.lda, (gp.flags & dev::kPairAKMajor) != 0);
            pb[p] = add_op(idb, gp.b, cb.n, gp.k, gp.ldb, (gp.flags & dev::kPairBNMajor) == 0);
        }
    }
    size_t exp_off = bytes;
    bytes += nexp * sizeof(int32_t);
    pl.oz_mem.reset(ctx, bytes ? bytes : 256);
    int8_t *sl = (int8_t*)pl.oz_mem.mem;
    int32_t *ex = (int32_t*)(sl + exp_off);

    std::vector<dev::oz_split_desc> descs(ops.size());
    std::vector<uint32_t> items, items_cached;
    for(size_t i = 0; i < ops.size(); i++) {
        dev::oz_split_desc &d = descs[i];
        std::memset(&d, 0, sizeof(d));
        d.src = ops[i].src;
        d.dst = sl + ops[i].off;
        d.exps = ex + ops[i].eoff;
        d.rows = ops[i].rows; d.k = ops[i].k; d.kp = ops[i].kp;
        d.slices = (uint32_t)slices;
        d.rs = ops[i].rs; d.ks = ops[i].ks;
        const uint32_t prows = (ops[i].rows + 15) & ~15u;   // whole 16-row panels
        d.sstride = prows * ops[i].kp;
        std::vector<uint32_t> &it = ops[i].cached ? items_cached : items;
        for(uint32_t r = 0; r < prows; r++) {
            it.push_back((uint32_t)i);
            it.push_back(r);
        }
    }
    // items of the operands split every perform first, then the cached ones
    const uint32_t n_mut = (uint32_t)(items.size() / 2);
    items.insert(items.end(), items_cached.begin(), items_cached.end());
    std::vector<dev::oz_pair> ozp(pl.pairs.size());
    for(size_t p = 0; p < pl.pairs.size(); p++) {
        const op_info &a = ops[pa[p]], &b = ops[pb[p]];
        dev::oz_pair &o = ozp[p];
        std::memset(&o, 0, sizeof(o));
        o.a = sl + a.off; o.b = sl + b.off;
        o.ea = ex + a.eoff; o.eb = ex + b.eoff;
        o.astride = ((a.rows + 15) & ~15u) * a.kp; o.bstride = ((b.rows + 15) & ~15u) * b.kp;
        o.kp = a.kp;
        o.d = pl.pairs[p].d;
        o.ia = (uint32_t)pa[p]; o.ib = (uint32_t)pb[p];
    }
    // C tiles of every product-kernel shape.  A workgroup runs four tiles
    // (one per wave) and workgroups are dealt to the XCDs round-robin, each
    // XCD with its own L2.  The C blocks of a group (same B operands) go to
    // one XCD, groups longest first onto the least loaded XCD; a group's
    // tiles run column tile by column tile, so the row tiles that read the
    // same B slice panels run together.  Workgroup w takes four tiles of XCD
    // w % nx's queue, or of the longest remaining queue once its own is
    // empty.  The order changes no result: each tile is one wave's sums.
    std::vector<dev::gemm_tile> tl[dev::OZ_TILE_COUNT];
    {
        const size_t nb = pl.cblocks.size();
        std::vector<size_t> grp(nb);
        size_t ng = pl.n_groups;
        if(pl.cb_group.size() == nb) grp = pl.cb_group;
        else {
            std::iota(grp.begin(), grp.end(), size_t(0));
            ng = nb;
        }
        const size_t nx = size_t(std::max(1, pl.n_xcd));
        std::vector<double> cb_cost(nb, 0.0);       // 64-k chunks of a tile
        std::vector<std::vector<size_t> > gblocks(ng);
        for(size_t b = 0; b < nb; b++) {
            for(uint32_t p = pl.cblocks[b].pair_begin; p < pl.cblocks[b].pair_end; p++)
                cb_cost[b] += double((pl.pairs[p].k + 63) / 64);
            gblocks[grp[b]].push_back(b);
        }
        for(int k = 0; k < dev::OZ_TILE_COUNT; k++) {
            uint32_t bm, bn;
            dev::oz_tile_shape(k, bm, bn);
            std::vector<std::vector<dev::gemm_tile> > gt(ng);
            std::vector<double> gc(ng, 0.0);
            for(size_t g = 0; g < ng; g++) {
                uint32_t maxtn = 0;
                for(size_t i = 0; i < gblocks[g].size(); i++) {
                    const dev::gemm_cblock &cb = pl.cblocks[gblocks[g][i]];
                    const uint32_t tm = (cb.m + bm - 1) / bm, tn = (cb.n + bn - 1) / bn;
                    lt_require(tm < 65536 && tn < 65536, LT_ERR_PARAM, "C block too large");
                    maxtn = std::max(maxtn, tn);
                }
                for(uint32_t j = 0; j < maxtn; j++)
                    for(size_t i = 0; i < gblocks[g].size(); i++) {
                        const size_t b = gblocks[g][i];
                        const dev::gemm_cblock &cb = pl.cblocks[b];
                        const uint32_t tm = (cb.m + bm - 1) / bm, tn = (cb.n + bn - 1) / bn;
                        if(j >= tn) continue;
                        for(uint32_t r = 0; r < tm; r++) {
                            dev::gemm_tile t;
                            t.cblock = (uint32_t)b; t.tm = (uint16_t)r; t.tn = (uint16_t)j;
                            gt[g].push_back(t);
                            gc[g] += cb_cost[b];
                        }
                    }
            }
            std::vector<size_t> go(ng);
            std::iota(go.begin(), go.end(), size_t(0));
            std::stable_sort(go.begin(), go.end(), [&](size_t x, size_t y) { return gc[x] > gc[y]; });
            std::vector<double> load(nx, 0.0);
            std::vector<std::vector<dev::gemm_tile> > q(nx);
            size_t total = 0;
            for(size_t i = 0; i < ng; i++) {
                const size_t x = size_t(std::min_element(load.begin(), load.end()) - load.begin());
                q[x].insert(q[x].end(), gt[go[i]].begin(), gt[go[i]].end());
                load[x] += gc[go[i]];
                total += gt[go[i]].size();
            }
            std::vector<size_t> qp(nx, 0);
            tl[k].reserve(total);
            for(size_t w = 0; tl[k].size() < total; w++) {
                for(int j = 0; j < 4 && tl[k].size() < total; j++) {
                    size_t x = w % nx;
                    if(qp[x] >= q[x].size()) {
                        size_t left = 0;
                        for(size_t y = 0; y < nx; y++)
                            if(q[y].size() - qp[y] > left) { left = q[y].size() - qp[y]; x = y; }
                    }
                    tl[k].push_back(q[x][qp[x]++]);
                }
            }
        }
    }
    size_t sz = 0;
    size_t o_d = sz; sz = align256(sz + descs.size() * sizeof(dev::oz_split_desc));
    size_t o_i = sz; sz = align256(sz + items.size() * 4);
    size_t o_p = sz; sz = align256(sz + ozp.size() * sizeof(dev::oz_pair));
    size_t o_bd = sz; sz = align256(sz + 2 * pl.cblocks.size() * sizeof(double));
    size_t o_om = sz; sz = align256(sz + ops.size() * sizeof(int32_t));
    size_t o_t[dev::OZ_TILE_COUNT];
    for(int k = 0; k < dev::OZ_TILE_COUNT; k++) {
        o_t[k] = sz;
        sz = align256(sz + tl[k].size() * sizeof(dev::gemm_tile));
    }
    std::vector<uint32_t> pm_cd, pm_co, pi_cd, pi_co;
    make_chunks(prep_mut, pm_cd, pm_co);
    make_chunks(prep_inv, pi_cd, pi_co);
    size_t o_pm = sz; sz = align256(sz + prep_mut.size() * sizeof(dev::perm_desc));
    size_t o_pmcd = sz; sz = align256(sz + pm_cd.size() * 4);
    size_t o_pmco = sz; sz = align256(sz + pm_co.size() * 4);
    size_t o_pi = sz; sz = align256(sz + prep_inv.size() * sizeof(dev::perm_desc));
    size_t o_picd = sz; sz = align256(sz + pi_cd.size() * 4);
    size_t o_pico = sz; sz = align256(sz + pi_co.size() * 4);
    std::vector<char> host(sz ? sz : 256, 0);
    if(!prep_mut.empty()) std::memcpy(&host[o_pm], prep_mut.data(), prep_mut.size() * sizeof(dev::perm_desc));
    if(!pm_cd.empty()) std::memcpy(&host[o_pmcd], pm_cd.data(), pm_cd.size() * 4);
    if(!pm_co.empty()) std::memcpy(&host[o_pmco], pm_co.data(), pm_co.size() * 4);
    if(!prep_inv.empty()) std::memcpy(&host[o_pi], prep_inv.data(), prep_inv.size() * sizeof(dev::perm_desc));
    if(!pi_cd.empty()) std::memcpy(&host[o_picd], pi_cd.data(), pi_cd.size() * 4);
    if(!pi_co.empty()) std::memcpy(&host[o_pico], pi_co.data(), pi_co.size() * 4);
    if(!descs.empty()) std::memcpy(&host[o_d], descs.data(), descs.size() * sizeof(dev::oz_split_desc));
    if(!items.empty()) std::memcpy(&host[o_i], items.data(), items.size() * 4);
    if(!ozp.empty()) std::memcpy(&host[o_p], ozp.data(), ozp.size() * sizeof(dev::oz_pair));
    for(int k = 0; k < dev::OZ_TILE_COUNT; k++)
        if(!tl[k].empty())
            std::memcpy(&host[o_t[k]], tl[k].data(), tl[k].size() * sizeof(dev::gemm_tile));
    pl.oz_lists.reset(ctx, host.size());
    ctx.upload_new(pl.oz_lists.mem, host.data(), host.size());
    char *base = (char*)pl.oz_lists.mem;
    pl.d_oz_split = (dev::oz_split_desc*)(base + o_d);
    pl.d_oz_items = (uint32_t*)(base + o_i);
    pl.n_oz_items = (uint32_t)(items.size() / 2);
    pl.d_oz_pairs = (dev::oz_pair*)(base + o_p);
    pl.d_oz_bound = (double*)(base + o_bd);
    pl.d_oz_opmax = (int32_t*)(base + o_om);
    pl.n_oz_ops = (uint32_t)ops.size();
    pl.n_oz_items_mut = n_mut;
    pl.d_oz_prep = (dev::perm_desc*)(base + o_pm);
    pl.d_oz_prep_cd = (uint32_t*)(base + o_pmcd);
    pl.d_oz_prep_co = (uint32_t*)(base + o_pmco);
    pl.n_oz_prep_chunks = (uint32_t)pm_cd.size();
    pl.oz_prep_own = true;
    if(pl.n_oz_items > n_mut) {
        // operands of immutable tensors (in place or invariant copies,
        // prepared first): split now, once for this plan
        ctx.activate();
        if(!pi_cd.empty())
            hip_check(dev::launch_permute((const dev::perm_desc*)(base + o_pi),
                (const uint32_t*)(base + o_picd), (const uint32_t*)(base + o_pico),
                (uint32_t)pi_cd.size(), ctx.stream()), "permute (prepare, immutable operands)");
        hip_check(dev::launch_oz_split(pl.d_oz_split, pl.d_oz_items + 2 * n_mut,
            pl.n_oz_items - n_mut, ctx.stream()), "split (ozaki, immutable operands)");
    }

    // Tile shape: each shape's product launch is timed once on this plan's
    // own work list (after one prepare and split) and the fastest kept.
    // Launches that would accumulate into C are never run for this, nor
    // for the plans of a memory-bounded perform (planned before any device
    // work): then the shape reading the fewest operand bytes (tiles x chunks
    // x fragments per chunk) is taken.
    bool all_zero = true;
    for(size_t b = 0; b < pl.cblocks.size(); b++) all_zero = all_zero && pl.cblocks[b].zero;
    int best = dev::OZ_TILE_16x64;
    if(m_oz_tile >= 0) {
        best = m_oz_tile;
    } else if(all_zero && !pl.batch && pl.events && !pl.cblocks.empty()) {
        hipStream_t s = ctx.stream();
        ctx.activate();
        if(pl.n_prep_chunks)
            hip_check(dev::launch_permute(pl.d_prep, pl.d_prep_cd, pl.d_prep_co,
                pl.n_prep_chunks, s), "permute (prepare)");
        hip_check(dev::launch_oz_split(pl.d_oz_split, pl.d_oz_items, pl.n_oz_items_mut, s),
            "split (ozaki)");
        float bestms = 1e30f;
        for(int k = 0; k < dev::OZ_TILE_COUNT; k++) {
            float ms = 0.f;
            hip_check(hipEventRecord(pl.ev[0], s), "hipEventRecord");
            hip_check(dev::launch_oz_gemm(pl.d_cb, pl.d_oz_pairs,
                (const dev::gemm_tile*)(base + o_t[k]), (uint32_t)tl[k].size(), slices, k, s),
                "gemm (ozaki, tile shapes)");
            hip_check(hipEventRecord(pl.ev[1], s), "hipEventRecord");
            hip_check(hipStreamSynchronize(s), "hipStreamSynchronize");
            hip_check(hipEventElapsedTime(&ms, pl.ev[0], pl.ev[1]), "hipEventElapsedTime");
            if(ms < bestms) { bestms = ms; best = k; }
        }
    } else {
        double bestb = 0.0;
        for(int k = 0; k < dev::OZ_TILE_COUNT; k++) {
            uint32_t bm, bn;
            dev::oz_tile_shape(k, bm, bn);
            double bytes = 0.0;
            for(size_t t = 0; t < tl[k].size(); t++) {
                const dev::gemm_cblock &cb = pl.cblocks[tl[k][t].cblock];
                for(uint32_t p = cb.pair_begin; p < cb.pair_end; p++)
                    bytes += double((pl.pairs[p].k + 63) / 64) * double(bm + bn);
            }
            if(k == 0 || bytes < bestb) { bestb = bytes; best = k; }
        }
    }
    pl.oz_tile = best;
    pl.oz_tile_set = m_oz_tile;
    pl.d_oz_tiles = (dev::gemm_tile*)(base + o_t[best]);
    pl.n_oz_tiles = (uint32_t)tl[best].size();
}

void btod_contract2::enqueue(int mode, hipStream_t s, bool timed, size_t &launches) {
    contract_plan &pl = *m_plan;
    if(mode == MODE_GROUPED && dev::is_stacked(pl.tile_cfg)) {
        if(pl.n_prep_st_chunks) {
            hip_check(dev::launch_permute(pl.d_prep_st, pl.d_pst_cd, pl.d_pst_co,
                pl.n_prep_st_chunks, s), "permute (prepare)");
            launches++;
        }
    } else if(mode == MODE_OZAKI && pl.oz_prep_own) {
        if(pl.n_oz_prep_chunks) {
            hip_check(dev::launch_permute(pl.d_oz_prep, pl.d_oz_prep_cd, pl.d_oz_prep_co,
                pl.n_oz_prep_chunks, s), "permute (prepare)");
            launches++;
        }
    } else if(pl.n_prep_chunks) {
        hip_check(dev::launch_permute(pl.d_prep, pl.d_prep_cd, pl.d_prep_co,
            pl.n_prep_chunks, s), "permute (prepare)");
        launches++;
    }
    if(timed) hip_check(hipEventRecord(pl.ev[1], s), "hipEventRecord");
    if(mode == MODE_OZAKI) {
        hip_check(dev::launch_oz_split(pl.d_oz_split, pl.d_oz_items, pl.n_oz_items_mut, s),
            "split (ozaki)");
        hip_check(dev::launch_oz_gemm(pl.d_cb, pl.d_oz_pairs, pl.d_oz_tiles, pl.n_oz_tiles,
            pl.oz_slices, pl.oz_tile, s), "gemm (ozaki)");
        launches += 2;
        if(timed) hip_check(hipEventRecord(pl.ev[2], s), "hipEventRecord");
        hip_check(dev::launch_oz_opmax(pl.d_oz_split, pl.n_oz_ops, pl.d_oz_opmax,
            (uint32_t)pl.c_out.size(), pl.d_oz_bound, s),
            "exponents (ozaki)");
        hip_check(dev::launch_oz_bound(pl.d_cb, pl.d_pairs, pl.d_oz_pairs, pl.d_oz_opmax,
            (uint32_t)pl.c_out.size(), pl.oz_slices, pl.d_oz_bound, s), "bound (ozaki)");
        launches += 2;
    } else if(mode == MODE_PER_BLOCK) {
        // one launch per C block; the blocks are independent, so their
        // launches go round-robin to side streams forked from s and joined
        // back into it (libtensor computes C blocks in parallel too)
        context &ctx = m_a.ctx();
        std::lock_guard<std::mutex> lk(ctx.side_lock());
        ctx.side_fork(s);
        int j = 0;
        for(size_t b = 0; b + 1 < pl.block_tile_begin.size(); b++) {
            uint32_t t0 = pl.block_tile_begin[b], t1 = pl.block_tile_begin[b + 1];
            if(t1 == t0) continue;
            hip_check(dev::launch_gemm(pl.tile_cfg, pl.d_cb, pl.d_pairs,
                pl.d_tiles_block + t0, t1 - t0, t1 - t0, ctx.side(j++)), "gemm (per block)");
            launches++;
        }
        ctx.side_join(s);
    } else {
        hip_check(launch_grouped(pl, pl.tile_cfg, s), "gemm (grouped)");
        launches++;
    }
    if(timed && mode != MODE_OZAKI) hip_check(hipEventRecord(pl.ev[2], s), "hipEventRecord");
    if(pl.n_wb_chunks) {
        hip_check(dev::launch_permute(pl.d_wback, pl.d_wb_cd, pl.d_wb_co,
            pl.n_wb_chunks, s), "permute (write back)");
        launches++;
    }
}

void btod_contract2::run(btensor &c, int mode) {
    contract_plan &pl = *m_plan;
    context &ctx = m_a.ctx();
    ctx.activate();
    hipStream_t s = ctx.stream();
    if(!pl.add && !pl.batch) {
        // A result tensor this operation wrote last keeps its structure, and
        // so its version (plans of operations that read it stay valid):
        // every block of c_out is overwritten below.
        std::vector<size_t> co(pl.c_out);
        std::sort(co.begin(), co.end());
        if(c.sym_tag() != m_uid || c.nonzero_blocks() != co) {
            c.zero_all();
            c.set_symmetry(m_symc, m_uid);
        }
    }
    if(pl.empty) {
        m_stats.t_gpu_ms = m_stats.t_gemm_ms = 0.0;
        m_stats.graph = 0;
        return;
    }
    if(mode == MODE_OZAKI) {
        ensure_ozaki();
        m_stats.ozaki_slices = pl.oz_slices;
    }
    // The grouped perform runs as a HIP graph (prepare permute, grouped
    // GEMM, write-back permute: kernel nodes only), captured once per plan.
    const bool graph = m_use_graph && mode == MODE_GROUPED;
    size_t launches = 0;
    if(graph && !pl.gexec) {
        size_t n = 0;
        hip_check(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal),
            "hipStreamBeginCapture");
        try {
            enqueue(mode, s, false, n);
        } catch(...) {
            hipGraph_t g = 0;
            hipStreamEndCapture(s, &g);
            if(g) hipGraphDestroy(g);
            throw;
        }
        hip_check(hipStreamEndCapture(s, &pl.graph), "hipStreamEndCapture");
        hip_check(hipGraphInstantiate(&pl.gexec, pl.graph, 0, 0, 0), "hipGraphInstantiate");
        pl.graph_launches = n;
    }
    hip_check(hipEventRecord(pl.ev[0], s), "hipEventRecord");
    if(pl.add) {
        // blocks that become non-zero start from zero before accumulation
        for(size_t i = 0; i < pl.c_out.size(); i++) {
            if(c.is_nonzero(pl.c_out[i])) continue;
            hip_check(dev::launch_fill(pl.c_out_mem[i].first, pl.c_out_mem[i].second,
                0.0, s), "fill");
            launches++;
        }
    }
    if(graph) {
        hip_check(hipGraphLaunch(pl.gexec, s), "hipGraphLaunch");
        launches += pl.graph_launches;
    } else {
        enqueue(mode, s, true, launches);
    }
    hip_check(hipEventRecord(pl.ev[3], s), "hipEventRecord");
    m_stats.n_launches = launches;
    m_stats.graph = graph ? 1 : 0;
    m_times_pending = true;
    m_times_graph = graph;
    // an asynchronous perform returns here (its times are read when asked
    // for); a split-precision perform reads its stated bounds, so it waits
    if(!m_async || mode == MODE_OZAKI) finish_times();
    if(mode == MODE_OZAKI) m_oz_bounds.clear();
    if(mode == MODE_OZAKI && !pl.c_out.empty()) {
        // the stated error of this perform: per C block the bound and the
        // largest |C|; overall max bound / max |C| (the relative error as
        // the metric measures it)
        // (the C block list ends with the empty block of the padding tiles)
        std::vector<double> bd(2 * pl.c_out.size());
        ctx.download(bd.data(), pl.d_oz_bound, bd.size() * sizeof(double));
        m_oz_bounds.clear();
        double bmax = 0.0, cmax = 0.0;
        for(size_t b = 0; b < pl.c_out.size(); b++) {
            m_oz_bounds.push_back(oz_bound{pl.c_out[b], bd[2 * b], bd[2 * b + 1]});
            bmax = std::max(bmax, bd[2 * b]);
            cmax = std::max(cmax, bd[2 * b + 1]);
        }
        m_stats.ozaki_max_rel_err = cmax > 0.0 ? bmax / cmax : bmax;
    }
    for(size_t i = 0; i < pl.c_out.size(); i++) c.mark_nonzero(pl.c_out[i]);
}

void btod_contract2::perform_sliced(btensor &c, bool add, double d, double slice_flops,
    const std::function<void()> &between) {
    // run() in grouped mode without the graph, with the products launched in
    // slices of the plan's own tile list and `between` called after each
    // slice has finished
    c.require_mutable("btod_contract2::perform");
    check_block_sizes();
    m_a.ctx().require_device("btod_contract2::perform");
    lt_require(c.bis() == m_symc.bis, LT_ERR_DIMS, "result tensor has a different block index space");
    if(add) {
        lt_require(add_direct(c), LT_ERR_SYMMETRY,
            "perform_add: result symmetry differs from the contraction's (the request "
            "loop adds into tensors of the result's symmetry)");
    }
    if(!plan_valid(c, add, d, MODE_GROUPED)) build_plan(c, add, d, false);
    contract_plan &pl = *m_plan;
    context &ctx = m_a.ctx();
    ctx.activate();
    hipStream_t s = ctx.stream();
    if(!pl.add) {
        std::vector<size_t> co(pl.c_out);
        std::sort(co.begin(), co.end());
        if(c.sym_tag() != m_uid || c.nonzero_blocks() != co) {
            c.zero_all();
            c.set_symmetry(m_symc, m_uid);
        }
    }
    if(pl.empty) return;
    hip_check(hipEventRecord(pl.ev[0], s), "hipEventRecord");
    if(pl.add) {
        for(size_t i = 0; i < pl.c_out.size(); i++) {
            if(c.is_nonzero(pl.c_out[i])) continue;
            hip_check(dev::launch_fill(pl.c_out_mem[i].first, pl.c_out_mem[i].second,
                0.0, s), "fill");
        }
    }
    const int k = pl.tile_cfg;
    if(dev::is_stacked(k)) {
        if(pl.n_prep_st_chunks)
            hip_check(dev::launch_permute(pl.d_prep_st, pl.d_pst_cd, pl.d_pst_co,
                pl.n_prep_st_chunks, s), "permute (prepare)");
    } else if(pl.n_prep_chunks) {
        hip_check(dev::launch_permute(pl.d_prep, pl.d_prep_cd, pl.d_prep_co,
            pl.n_prep_chunks, s), "permute (prepare)");
    }
    hip_check(hipEventRecord(pl.ev[1], s), "hipEventRecord");
    const uint32_t nt = dev::is_var(k) ? (uint32_t)pl.vtiles.size() :
        dev::is_stacked(k) ? (uint32_t)pl.stiles.size() : (uint32_t)pl.tiles_sorted.size();
    const uint32_t nslices = (uint32_t)std::max(1.0, std::min(double(nt),
        std::ceil(m_stats.flops / slice_flops)));
    const uint32_t per = (nt + nslices - 1) / nslices;
    for(uint32_t t0 = 0; t0 < nt; t0 += per) {
        const uint32_t n = std::min(per, nt - t0);
        if(dev::is_var(k))
            hip_check(dev::launch_gemm_stacked_var(k, pl.d_stacks, pl.d_spairs, pl.d_arows,
                pl.d_crows, pl.d_vtiles + t0, n, s), "gemm (slice)");
        else if(dev::is_stacked(k))
            hip_check(dev::launch_gemm_stacked(k, pl.d_stacks, pl.d_spairs, pl.d_arows,
                pl.d_crows, pl.d_stiles + t0, n, s), "gemm (slice)");
        else
            hip_check(dev::launch_gemm(k, pl.d_cb, pl.d_pairs, pl.d_tiles_sorted + t0, n, n, s),
                "gemm (slice)");
        if(t0 + n < nt) {
            hip_check(hipStreamSynchronize(s), "hipStreamSynchronize");
            between();
        }
    }
    hip_check(hipEventRecord(pl.ev[2], s), "hipEventRecord");
    if(pl.n_wb_chunks)
        hip_check(dev::launch_permute(pl.d_wback, pl.d_wb_cd, pl.d_wb_co, pl.n_wb_chunks, s),
            "permute (write back)");
    hip_check(hipEventRecord(pl.ev[3], s), "hipEventRecord");
    hip_check(hipStreamSynchronize(s), "hipStreamSynchronize");
    float ms = 0.f, msg = 0.f;
    hip_check(hipEventElapsedTime(&ms, pl.ev[0], pl.ev[3]), "hipEventElapsedTime");
    hip_check(hipEventElapsedTime(&msg, pl.ev[1], pl.ev[2]), "hipEventElapsedTime");
    m_stats.t_gpu_ms = ms;
    m_stats.t_gemm_ms = msg;
    m_stats.n_launches = nslices + 2;
    m_stats.graph = 0;
    m_stats.n_batches = 1;
    for(size_t i = 0; i < pl.c_out.size(); i++) c.mark_nonzero(pl.c_out[i]);
}

batch_run::batch_run(std::vector<batch_item> &items, dev_buffer &merged) :
    m_items(items), m_ctx(0), m_base(0), m_launch(0), m_tile(0), m_flops(0.0),
    m_finished(false) {

    // every item checked and planned on its own: one that fails keeps its
    // error, the others run
    for(size_t i = 0; i < items.size(); i++) {
        batch_item &it = items[i];
        it.status = LT_OK;
        it.message.clear();
        try {
            lt_require(it.op && it.c, LT_ERR_PARAM, "batch item without operation or tensor");
            it.c->require_mutable("btod_contract2::perform");
            btod_contract2 &op = *it.op;
            if(!m_ctx) m_ctx = &op.m_a.ctx();
            lt_require(&op.m_a.ctx() == m_ctx, LT_ERR_PARAM, "batch spans contexts");
            op.check_block_sizes();
            lt_require(it.c->bis() == op.m_symc.bis, LT_ERR_DIMS,
                "result tensor has a different block index space");
            if(it.add) {
                lt_require(op.add_direct(*it.c), LT_ERR_SYMMETRY,
                    "perform_add: result symmetry differs from the contraction's (the request "
                    "loop adds into tensors of the result's symmetry)");
            }
            // no timed autotuning: the stream may be shared with other work
            if(!op.plan_valid(*it.c, it.add, it.d, MODE_GROUPED))
                op.build_plan(*it.c, it.add, it.d, false);
            m_live.push_back(i);
            m_plans.push_back(op.m_plan.get());
        } catch(const lt_error &e) {
            it.status = e.code;
            it.message = e.what();
        } catch(const std::bad_alloc &) {
            it.status = LT_ERR_NOMEM;
            it.message = "host out of memory";
        }
    }
    if(m_live.empty()) { m_finished = true; return; }
    context &ctx = *m_ctx;
    ctx.require_device("btod_contract2::perform_batch");
    ctx.activate();
    hipStream_t s = ctx.stream();

    // result preparation and prepare permutes
    hip_check(hipEventRecord(m_plans[0]->ev[0], s), "hipEventRecord");
    for(size_t q = 0; q < m_live.size(); q++) {
        contract_plan &pl = *m_plans[q];
        batch_item &it = items[m_live[q]];
        btensor &c = *it.c;
        if(!pl.add) {
            c.zero_all();
            c.set_symmetry(it.op->m_symc);
        } else {
            for(size_t j = 0; j < pl.c_out.size(); j++) {
                if(c.is_nonzero(pl.c_out[j])) continue;
                hip_check(dev::launch_fill(pl.c_out_mem[j].first, pl.c_out_mem[j].second,
                    0.0, s), "fill");
            }
        }
        if(!pl.empty && pl.n_prep_chunks)
            hip_check(dev::launch_permute(pl.d_prep, pl.d_prep_cd, pl.d_prep_co,
                pl.n_prep_chunks, s), "permute (prepare)");
    }

    // one grouped GEMM per kernel configuration: the work lists of the
    // plans concatenated (C block and pair indexes shifted), the tiles
    // ordered longest first across all requests
    std::map<int, std::vector<size_t> > groups;
    for(size_t q = 0; q < m_live.size(); q++) {
        if(m_plans[q]->empty) continue;
        groups[m_plans[q]->tile_cfg].push_back(q);
    }
    std::vector<dev::gemm_cblock> cbs;
    std::vector<dev::gemm_pair> prs;
    std::vector<std::pair<double, dev::gemm_tile> > tl;
    std::vector<char> host;
    size_t sz = 0;
    for(std::map<int, std::vector<size_t> >::iterator g = groups.begin(); g != groups.end(); ++g) {
        cbs.clear(); prs.clear(); tl.clear();
        uint32_t bm = 64, bn = 64;
        dev::tile_shape(g->first, bm, bn);
        for(size_t x = 0; x < g->second.size(); x++) {
            contract_plan &pl = *m_plans[g->second[x]];
            uint32_t cb0 = (uint32_t)cbs.size(), pr0 = (uint32_t)prs.size();
            for(size_t b = 0; b < pl.cblocks.size(); b++) {
                dev::gemm_cblock cb = pl.cblocks[b];
                cb.pair_begin += pr0;
                cb.pair_end += pr0;
                cbs.push_back(cb);
            }
            prs.insert(prs.end(), pl.pairs.begin(), pl.pairs.end());
            for(size_t t = 0; t < pl.tiles_sorted.size(); t++) {
                dev::gemm_tile y = pl.tiles_sorted[t];
                if(y.cblock >= pl.cb_ksum.size()) continue;     // padding tile
                // FLOPs of the tile: its rows x columns (clipped to the
                // block) x 2 x the k extents of the block's pairs
                const dev::gemm_cblock &cb = pl.cblocks[y.cblock];
                const double rows = std::min<double>(bm, double(cb.m) - double(y.tm) * bm);
                const double cols = std::min<double>(bn, double(cb.n) - double(y.tn) * bn);
                const double cost = 2.0 * rows * cols * pl.cb_ksum[y.cblock];
                y.cblock += cb0;
                tl.push_back(std::make_pair(cost, y));
                m_flops += cost;
            }
        }
        std::stable_sort(tl.begin(), tl.end(),
            [](const std::pair<double, dev::gemm_tile> &x, const std::pair<double, dev::gemm_tile> &y) {
                return x.first > y.first; });
        launch L;
        L.cfg = g->first;
        L.ocb = sz; sz = align256(sz + cbs.size() * sizeof(dev::gemm_cblock));
        L.opr = sz; sz = align256(sz + prs.size() * sizeof(dev::gemm_pair));
        L.otl = sz; sz = align256(sz + tl.size() * sizeof(dev::gemm_tile));
        L.nt = (uint32_t)tl.size();
        host.resize(sz, 0);
        if(!cbs.empty()) std::memcpy(&host[L.ocb], cbs.data(), cbs.size() * sizeof(dev::gemm_cblock));
        if(!prs.empty()) std::memcpy(&host[L.opr], prs.data(), prs.size() * sizeof(dev::gemm_pair));
        for(size_t t = 0; t < tl.size(); t++) {
            std::memcpy(&host[L.otl + t * sizeof(dev::gemm_tile)], &tl[t].second,
                sizeof(dev::gemm_tile));
            L.tcost.push_back(tl[t].first);
        }
        m_launches.push_back(L);
    }
    if(sz) {
        if(merged.bytes < sz) {
            merged.release();
            merged.ctx = &ctx;
            merged.mem = ctx.alloc(sz);
            merged.bytes = sz;
        }
        hip_check(hipMemcpyAsync(merged.mem, host.data(), sz, hipMemcpyHostToDevice, s),
            "hipMemcpy H2D");
        // the host buffer goes out of scope: the copy must have read it
        hip_check(hipStreamSynchronize(s), "hipStreamSynchronize");
    }
    m_base = (char*)merged.mem;
    hip_check(hipEventRecord(m_plans[0]->ev[1], s), "hipEventRecord");
}

bool batch_run::step(double flop_budget) {
    if(m_finished) return false;
    hipStream_t s = m_ctx->stream();
    while(m_launch < m_launches.size()) {
        const launch &L = m_launches[m_launch];
        if(m_tile >= L.nt) { m_launch++; m_tile = 0; continue; }
        // at least one tile, then as many as the budget allows
        uint32_t t1 = (uint32_t)m_tile;
        double f = 0.0;
        do { f += L.tcost[t1]; t1++; } while(t1 < L.nt && f + L.tcost[t1] <= flop_budget);
        const uint32_t n = t1 - (uint32_t)m_tile;
        hip_check(dev::launch_gemm(L.cfg, (dev::gemm_cblock*)(m_base + L.ocb),
            (dev::gemm_pair*)(m_base + L.opr),
            (dev::gemm_tile*)(m_base + L.otl) + m_tile, n, n, s), "gemm (batch)");
        m_tile = t1;
        return m_launch < m_launches.size() &&
            (m_tile < L.nt || m_launch + 1 < m_launches.size());
    }
    return false;
}

void batch_run::finish() {
    if(m_finished) return;
    while(step(1e300)) { }
    hipStream_t s = m_ctx->stream();
    hip_check(hipEventRecord(m_plans[0]->ev[2], s), "hipEventRecord");
    for(size_t q = 0; q < m_live.size(); q++) {
        contract_plan &pl = *m_plans[q];
        if(!pl.empty && pl.n_wb_chunks)
            hip_check(dev::launch_permute(pl.d_wback, pl.d_wb_cd, pl.d_wb_co,
                pl.n_wb_chunks, s), "permute (write back)");
    }
    hip_check(hipEventRecord(m_plans[0]->ev[3], s), "hipEventRecord");
    hip_check(hipStreamSynchronize(s), "hipStreamSynchronize");
    float ms = 0.f, msg = 0.f;
    hip_check(hipEventElapsedTime(&ms, m_plans[0]->ev[0], m_plans[0]->ev[3]), "hipEventElapsedTime");
    hip_check(hipEventElapsedTime(&msg, m_plans[0]->ev[1], m_plans[0]->ev[2]), "hipEventElapsedTime");
    for(size_t q = 0; q < m_live.size(); q++) {
        contract_plan &pl = *m_plans[q];
        batch_item &it = m_items[m_live[q]];
        for(size_t j = 0; j < pl.c_out.size(); j++) it.c->mark_nonzero(pl.c_out[j]);
        exec_stats &st = it.op->m_stats;
        st.t_gpu_ms = ms;
        st.t_gemm_ms = msg;
        st.n_launches = m_launches.size();
        st.graph = 0;
        it.batch_gpu_ms = ms;
        it.batch_gemm_launches = m_launches.size();
    }
    m_finished = true;
}

void btod_contract2::perform_batch(std::vector<batch_item> &items, dev_buffer &merged) {
    batch_run run(items, merged);
    run.finish();
}

void btod_contract2::prepare(btensor &c, bool add, double d, bool measure) {
    check_block_sizes();
    m_a.ctx().require_device("btod_contract2::prepare");
    lt_require(c.bis() == m_symc.bis, LT_ERR_DIMS, "result tensor has a different block index space");
    if(add) {
        lt_require(add_direct(c), LT_ERR_SYMMETRY,
            "perform_add: result symmetry differs from the contraction's (the request "
            "loop adds into tensors of the result's symmetry)");
    }
    if(!plan_valid(c, add, d, MODE_GROUPED)) build_plan(c, add, d, measure);
}

size_t btod_contract2::footprint_estimate() {
    size_t n = 0;
    std::vector<size_t> nza = m_a.nonzero_blocks(), nzb = m_b.nonzero_blocks();
    for(size_t i = 0; i < nza.size(); i++) n += 4 * m_a.block_size(nza[i]);
    for(size_t i = 0; i < nzb.size(); i++) n += 4 * m_b.block_size(nzb[i]);
    dims_t bidc = m_symc.bis.block_index_dims();
    for(size_t i = 0; i < m_sched.size(); i++)
        n += m_symc.bis.block_dims(get_index(m_sched[i], bidc)).size();
    return n * sizeof(double) + (size_t(1) << 20);
}

bool btod_contract2::try_plan(btensor &c, bool add, double d, bool measure, int mode) {
    // the plan (and in split-precision mode its split form) within the
    // context's memory budget, or false with nothing allocated
    try {
        if(!plan_valid(c, add, d, mode)) build_plan(c, add, d, measure);
        if(mode == MODE_OZAKI) ensure_ozaki();
        return true;
    } catch(const lt_error &e) {
        if(e.code != LT_ERR_NOMEM) throw;
        drop_plan();           // frees what the plan took
        return false;
    }
}

void btod_contract2::ensure_ozaki() {
    contract_plan &pl = *m_plan;
    const int sl = m_oz_slices > 0 ? m_oz_slices : 7;
    if(pl.oz_slices != sl || pl.oz_tile_set != m_oz_tile) build_ozaki(sl);
}

void btod_contract2::run_batched(btensor &c, bool add, double d, int mode) {
    // gen_bto_contract2::perform (gen_bto_contract2_impl.h:122-199,
    // gen_bto_contract2_batching_policy.h:48-70) runs the contraction in
    // batches of C blocks whose temporaries fit the memory limit.  Here the
    // schedule is cut into 2, 4, 8, ... batches of consecutive blocks until
    // the plan of every batch (prepared operand copies, work lists) fits the
    // context's budget; each batch is planned, run and released in turn.
    // The probe pass plans every batch once before any device work, so an
    // accumulating perform never stops half way.
    const size_t n = m_sched.size();
    std::vector<std::vector<size_t> > parts;
    for(size_t nb = 2; ; nb *= 2) {
        lt_require(nb / 2 < n, LT_ERR_NOMEM,
            "btod_contract2: a single C block's plan exceeds the device memory budget");
        parts.assign(std::min(nb, n), std::vector<size_t>());
        for(size_t i = 0; i < n; i++) parts[i * parts.size() / n].push_back(m_sched[i]);
        bool ok = true;
        for(size_t k = 0; k < parts.size() && ok; k++) {
            m_sched_subset = &parts[k];
            ok = try_plan(c, add, d, false, mode);
            m_sched_subset = 0;
            drop_plan();
        }
        if(ok) break;
    }
    if(!add) {
        c.zero_all();
        c.set_symmetry(m_symc, m_uid);
    }
    // split-precision bounds of every batch's C blocks
    std::vector<oz_bound> bounds;
    for(size_t k = 0; k < parts.size(); k++) {
        m_sched_subset = &parts[k];
        try {
            build_plan(c, add, d, false);
            m_use_graph_saved = m_use_graph;
            m_use_graph = false;        // a batch plan runs once: no capture
            run(c, mode);
            m_use_graph = m_use_graph_saved;
        } catch(...) {
            m_sched_subset = 0;
            m_use_graph = m_use_graph_saved;
            drop_plan();
            throw;
        }
        if(mode == MODE_OZAKI)
            bounds.insert(bounds.end(), m_oz_bounds.begin(), m_oz_bounds.end());
        m_sched_subset = 0;
        drop_plan();
    }
    if(mode == MODE_OZAKI) {
        m_oz_bounds = bounds;
        double bmax = 0.0, cmax = 0.0;
        for(size_t i = 0; i < bounds.size(); i++) {
            bmax = std::max(bmax, bounds[i].bound);
            cmax = std::max(cmax, bounds[i].cmax);
        }
        m_stats.ozaki_max_rel_err = cmax > 0.0 ? bmax / cmax : bmax;
    }
    m_stats.n_batches = parts.size();
}

void btod_contract2::perform(btensor &c, int mode) {
    c.require_mutable("btod_contract2::perform");
    check_block_sizes();
    m_a.ctx().require_device("btod_contract2::perform");
    lt_require(c.bis() == m_symc.bis, LT_ERR_DIMS, "result tensor has a different block index space");
    // (no timed kernel choice for a split-precision perform: its products
    // run in their own kernels, and nothing is launched before it is known
    // whether the plan fits)
    if(!try_plan(c, false, 1.0, mode != MODE_OZAKI, mode)) {
        run_batched(c, false, 1.0, mode);
        return;
    }
    m_stats.n_batches = 1;
    run(c, mode);
}

bool btod_contract2::add_direct(btensor &c) {
    if(c.sym_tag() == m_uid) return true;
    if(c.uid() != m_add_uid || c.version() != m_add_ver) {
        m_add_direct = same_orbits(m_symc, c.sym(), c.bidims());
        m_add_uid = c.uid();
        m_add_ver = c.version();
    }
    return m_add_direct;
}

void btod_contract2::perform_add(btensor &c, double d, int mode) {
    c.require_mutable("btod_contract2::perform");
    check_block_sizes();
    m_a.ctx().require_device("btod_contract2::perform");
    lt_require(c.bis() == m_symc.bis, LT_ERR_DIMS, "result tensor has a different block index space");
    if(!add_direct(c)) {
        // C of another symmetry (btod_contract2_impl.h:70-90: addition_schedule
        // + gen_bto_aux_add): C is lowered to the symmetry it shares with the
        // result, the result is computed into a resident temporary and every
        // canonical block of it added to each canonical block of C in its orbit
        merge_symmetry(c, m_symc);
        if(!m_addtmp) m_addtmp.reset(new btensor(m_a.ctx(), m_symc.bis));
        perform(*m_addtmp, mode);
        tensor_tr id(c.bis().order());
        add_into(c, *m_addtmp, id, d);
        return;
    }
    if(!try_plan(c, true, d, mode != MODE_OZAKI, mode)) {
        run_batched(c, true, d, mode);
        return;
    }
    m_stats.n_batches = 1;
    run(c, mode);
}

// ---------------------------------------------------------------------------

void btod_contract2::compute_block(bool zero, const index_t &ic, const tensor_tr &trc,
    double *blk) {
    // gen_bto_contract2::compute_block (gen_bto_contract2_impl.h:312-345):
    // the full contraction list of the block in the original index space,
    // then gen_bto_contract2_block::compute_block (impl :100-205): every
    // pair contracted from the canonical blocks with the connection permuted
    // by their transformations and the scalars folded into d, pairs with
    // d == 0 skipped (tod_contract2_impl.h:129), the sum transformed by trc
    // into the output block.
    check_block_sizes();
    context &ctx = m_a.ctx();
    ctx.require_device("btod_contract2::compute_block");
    dims_t dc = m_symc.bis.block_dims(ic);
    lt_require(trc.perm.n == dc.n, LT_ERR_PARAM, "block transformation order");
    // the block's contraction terms: built once per operand structure (a
    // direct tensor asks for the same blocks again and again)
    const size_t key = abs_index(ic, m_symc.bis.block_index_dims());
    std::vector<dense_contract_item> items;
    bool known = false;
    {
        std::lock_guard<std::mutex> lk(m_cb_mtx);
        if(m_cb_va != m_a.version() || m_cb_vb != m_b.version()) {
            m_cb_terms.clear();
            m_cb_va = m_a.version();
            m_cb_vb = m_b.version();
        }
        std::map<size_t, std::vector<dense_contract_item> >::const_iterator i =
            m_cb_terms.find(key);
        if(i != m_cb_terms.end()) { items = i->second; known = true; }
    }
    if(!known) {
        const uint64_t va = m_a.version(), vb = m_b.version();
        dims_t bida = m_a.bidims(), bidb = m_b.bidims();
        block_list bla(bida, m_a.nonzero_blocks()), blb(bidb, m_b.nonzero_blocks());
        block_list blax(bida), blbx(bidb);
        unfold_block_list(m_a.sym(), bla, blax);
        unfold_block_list(m_b.sym(), blb, blbx);
        contr_list cl;
        build_clst_full(m_contr, m_a.sym(), m_b.sym(), blax, blbx,
            m_symc.bis.block_index_dims(), ic, false, cl);
        for(contr_list::const_iterator p = cl.begin(); p != cl.end(); ++p) {
            double d = p->tra.st.c * m_ka * p->trb.st.c * m_kb * m_kc;
            if(d == 0.0) continue;
            dense_contract_item it;
            it.da = m_a.bis().block_dims(get_index(p->acia, bida));
            it.db = m_b.bis().block_dims(get_index(p->acib, bidb));
            it.contr = m_contr;
            it.contr.permute_a(p->tra.perm.inverse());
            it.contr.permute_b(p->trb.perm.inverse());
            it.a = m_a.block_ptr(p->acia);
            it.b = m_b.block_ptr(p->acib);
            it.d = d;
            it.zero = items.empty();
            items.push_back(it);
        }
        std::lock_guard<std::mutex> lk(m_cb_mtx);
        if(m_cb_va == va && m_cb_vb == vb) {
            if(m_cb_terms.size() >= 65536) m_cb_terms.clear();
            m_cb_terms[key] = items;
        }
    }
    double *sum = (double*)ctx.alloc(std::max<size_t>(dc.size(), 1) * sizeof(double));
    try {
        // the terms in list order, one allocation, upload and
        // synchronization for all of them
        if(items.empty()) {
            ctx.activate();
            hip_check(dev::launch_fill(sum, dc.size(), 0.0, ctx.stream()), "fill");
        } else dense_contract2_many(ctx, items, sum);
        dense_permute(ctx, sum, dc, trc, blk, !zero);
    } catch(...) {
        ctx.free(sum);
        throw;
    }
    ctx.free(sum);
}

void btod_contract2::compute_block_host(const index_t &ic, double *out) {
    // Direct path of libtensor (gen_bto_contract2::compute_block): full
    // contraction list in the original space, canonical blocks with their
    // transformations, each pair contracted on the host in FP64.
    dims_t bida = m_a.bidims(), bidb = m_b.bidims();
    block_list bla(bida, m_a.nonzero_blocks()), blb(bidb, m_b.nonzero_blocks());
    block_list blax(bida), blbx(bidb);
    unfold_block_list(m_a.sym(), bla, blax);
    unfold_block_list(m_b.sym(), blb, blbx);
    contr_list cl;
    build_clst_full(m_contr, m_a.sym(), m_b.sym(), blax, blbx,
        m_symc.bis.block_index_dims(), ic, false, cl);
    dims_t dc = m_symc.bis.block_dims(ic);
    for(size_t i = 0; i < dc.size(); i++) out[i] = 0.0;
    for(contr_list::const_iterator p = cl.begin(); p != cl.end(); ++p) {
        dims_t da = m_a.bis().block_dims(get_index(p->acia, bida));
        dims_t db = m_b.bis().block_dims(get_index(p->acib, bidb));
        std::vector<double> a(da.size()), b(db.size());
        m_a.get_block(p->acia, a.data());
        m_b.get_block(p->acib, b.data());
        contraction2 cc(m_contr);
        cc.permute_a(p->tra.perm.inverse());
        cc.permute_b(p->trb.perm.inverse());
        double d = p->tra.st.c * m_ka * p->trb.st.c * m_kb * m_kc;
        // brute-force contraction over the connection sequence
        const size_t NC = cc.NC(), NA = cc.NA(), NB = cc.NB();
        dims_t dk(cc.K);
        for(size_t i = 0, k = 0; i < NA; i++)
            if(cc.conn[NC + i] >= NC + NA) dk[k++] = da[i];
        size_t nk = dk.size();
        index_t ie(NC);
        for(size_t e = 0; e < dc.size(); e++) {
            double sum = 0.0;
            index_t ik(cc.K);
            for(size_t kk = 0; kk < nk; kk++) {
                index_t xa(NA), xb(NB);
                for(size_t i = 0; i < NC; i++) {
                    size_t j = cc.conn[i] - NC;
                    if(j < NA) xa[j] = ie[i]; else xb[j - NA] = ie[i];
                }
                for(size_t i = 0, k = 0; i < NA; i++) {
                    if(cc.conn[NC + i] >= NC + NA) {
                        xa[i] = ik[k];
                        xb[cc.conn[NC + i] - NC - NA] = ik[k];
                        k++;
                    }
                }
                sum += a[abs_index(xa, da)] * b[abs_index(xb, db)];
                inc_index(ik, dk);
            }
            out[e] += d * sum;
            inc_index(ie, dc);
        }
    }
}

} // namespace lta
