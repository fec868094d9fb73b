// btod_copy and btod_add on HBM-resident block tensors.
//
// btod_copy: B = c P(A), or B += d c P(A)
// (libtensor/block_tensor/btod_copy.h:20-80; gen_block_tensor/impl/
// gen_bto_copy_impl.h:150-163: the symmetry of B is so_permute(sym A, P);
// :302-330: every non-zero canonical block of A goes to the canonical block
// of the orbit of its image P(ia) in B, transformed by (P, c) and then by the
// inverse of that orbit's transformation).  B += d P(A) into a B of another
// symmetry lowers B to the symmetry both have first, as gen_bto_aux_add
// does.  All blocks are one launch of the batched permute kernel.
//
// btod_add: B = sum_i c_i P_i(A_i), or B += d sum_i ... (libtensor/
// block_tensor/btod_add.h; gen_block_tensor/impl/gen_bto_add_impl.h).  B
// takes the symmetry all permuted operands have (sum_symmetry, the result of
// libtensor's so_dirsum then so_merge, gen_bto_add_impl.h:238-267); each
// non-zero canonical block of an operand goes to every canonical block of B
// in its orbit.  With +=, B's own symmetry takes part as well: B is first
// lowered to the common symmetry (new canonical blocks written out from its
// old ones).  One permute launch per operand (the first write of a block
// overwrites, later ones accumulate), so no two descriptors of a launch
// share a destination.
#include "lt_runtime.h"

#include <algorithm>
#include <cstdio>
#include <set>
#include <unordered_set>

namespace lta {

bool same_orbits(const symmetry &x, const symmetry &y, const dims_t &bid) {
    for(size_t i = 0; i < bid.size(); i++) {
        orbit ox(x, i, true), oy(y, i, true);
        if(ox.acidx != oy.acidx || ox.allowed != oy.allowed) return false;
        if(!(ox.get_transf(i) == oy.get_transf(i))) return false;
    }
    return true;
}

void run_permutes(context &ctx, const std::vector<dev::perm_desc> &descs, double *ms) {
    if(descs.empty()) return;
    std::vector<uint32_t> cd, co;
    make_chunks(descs, cd, co);
    if(cd.empty()) return;
    size_t o_cd = (descs.size() * sizeof(dev::perm_desc) + 255) & ~size_t(255);
    size_t o_co = o_cd + ((4 * cd.size() + 255) & ~size_t(255));
    size_t bytes = o_co + ((4 * co.size() + 255) & ~size_t(255));
    std::vector<char> host(bytes, 0);
    std::memcpy(&host[0], descs.data(), descs.size() * sizeof(dev::perm_desc));
    std::memcpy(&host[o_cd], cd.data(), 4 * cd.size());
    std::memcpy(&host[o_co], co.data(), 4 * co.size());
    std::lock_guard<std::mutex> lk(ctx.staging_lock());
    char *mem = (char*)ctx.staging(bytes);
    hipEvent_t ev[2] = {0, 0};
    try {
        ctx.upload(mem, host.data(), bytes);
        ctx.activate();
        if(ms) {
            hip_check(hipEventCreate(&ev[0]), "hipEventCreate");
            hip_check(hipEventCreate(&ev[1]), "hipEventCreate");
            hip_check(hipEventRecord(ev[0], ctx.stream()), "hipEventRecord");
        }
        hip_check(dev::launch_permute((dev::perm_desc*)mem, (uint32_t*)(mem + o_cd),
            (uint32_t*)(mem + o_co), (uint32_t)cd.size(), ctx.stream()), "permute");
        if(ms) hip_check(hipEventRecord(ev[1], ctx.stream()), "hipEventRecord");
        ctx.sync();
        if(ms) {
            float t = 0.f;
            hip_check(hipEventElapsedTime(&t, ev[0], ev[1]), "hipEventElapsedTime");
            *ms += t;
        }
    } catch(...) {
        for(int i = 0; i < 2; i++) if(ev[i]) (void)hipEventDestroy(ev[i]);
        throw;
    }
    for(int i = 0; i < 2; i++) if(ev[i]) (void)hipEventDestroy(ev[i]);
}

void merge_symmetry(btensor &b, const symmetry &other, double *ms) {
    b.require_mutable("merge_symmetry");
    // the operation's symmetry first, the target's second, as
    // gen_bto_aux_add::open passes them
    const symmetry symc = sum_symmetry(other, b.sym());
    const dims_t bid = b.bidims();
    if(same_orbits(symc, b.sym(), bid)) return;
    // every canonical block of the merged symmetry that b holds only through
    // an orbit of its old symmetry is written out from the old canonical
    // block (gen_bto_aux_add::put / close "touch" the orbit,
    // gen_bto_aux_add_impl.h:115-145,183-208)
    std::vector<dev::perm_desc> descs;
    std::vector<size_t> fresh;
    const std::vector<size_t> nz = b.nonzero_blocks();
    for(size_t i = 0; i < nz.size(); i++) {
        const size_t x = nz[i];
        orbit o(b.sym(), x, false);
        const dims_t dx = b.bis().block_dims(get_index(x, bid));
        for(size_t k = 0; k < o.orb.size(); k++) {
            const size_t y = o.orb[k].first;
            if(y == x) continue;
            short_orbit so(symc, y, true);
            if(so.acidx != y || !so.allowed) continue;
            descs.push_back(make_perm_desc_public(b.block_ptr(x), dx, o.get_transf(y),
                b.block_ptr(y), false));
            fresh.push_back(y);
        }
    }
    run_permutes(b.ctx(), descs, ms);
    b.set_symmetry(symc);
    for(size_t i = 0; i < fresh.size(); i++) b.mark_nonzero(fresh[i]);
}

namespace {

//! Descriptors of b += d tr(s) (see add_into) and the blocks they write.
void add_descs(btensor &b, btensor &s, const tensor_tr &tr, double d,
    std::vector<dev::perm_desc> &descs, std::vector<size_t> &touched) {
    lt_require(&s.ctx() == &b.ctx(), LT_ERR_PARAM, "tensors of different contexts");
    lt_require(tr.perm.n == s.bis().order(), LT_ERR_PARAM, "permutation order");
    bis_t bis = s.bis();
    bis.permute(tr.perm);
    lt_require(bis == b.bis(), LT_ERR_DIMS, "block index space of the added tensor");
    const symmetry syms = so_permute(s.sym(), tr.perm);
    const dims_t bids = s.bidims(), bidb = b.bidims();
    // block x = P(a) of the permuted term, then every canonical block y of b
    // in the term's orbit of x: b[y] += d tr_y inv(tr_x) (c P(s[a]))
    // (gen_bto_aux_add::put, gen_bto_aux_add_impl.h:211-232, with the
    // addition_schedule's nodes of the orbit)
    std::unordered_set<size_t> seen(touched.begin(), touched.end());
    const std::vector<size_t> nz = s.nonzero_blocks();
    for(size_t i = 0; i < nz.size(); i++) {
        index_t ix = get_index(nz[i], bids);
        const dims_t da = s.bis().block_dims(ix);
        tr.perm.apply(ix);
        const size_t x = abs_index(ix, bidb);
        orbit o(syms, x, true);
        if(!o.allowed) continue;
        const tensor_tr inv_x = o.get_transf(x).inverse();
        for(size_t k = 0; k < o.orb.size(); k++) {
            const size_t y = o.orb[k].first;
            short_orbit so(b.sym(), y, true);
            if(so.acidx != y || !so.allowed) continue;
            tensor_tr t(tr);
            t.transform(inv_x);
            t.transform(o.get_transf(y));
            t.st.scale(d);
            const bool acc = b.is_nonzero(y) || seen.count(y) != 0;
            descs.push_back(make_perm_desc_public(s.block_ptr(nz[i]), da, t, b.block_ptr(y),
                acc));
            touched.push_back(y);
            seen.insert(y);
        }
    }
}

} // unnamed namespace

//! Uploads descriptor lists (one launch each) into one device buffer.
std::shared_ptr<context::cached_launch> upload_launches(context &ctx,
    const std::vector<std::vector<dev::perm_desc> > &lists) {
    std::shared_ptr<context::cached_launch> l(new context::cached_launch());
    std::vector<char> host;
    size_t sz = 0;
    for(size_t i = 0; i < lists.size(); i++) {
        if(lists[i].empty()) continue;
        std::vector<uint32_t> cd, co;
        make_chunks(lists[i], cd, co);
        if(cd.empty()) continue;
        context::cached_launch::segment g;
        g.o_d = sz; sz = (sz + lists[i].size() * sizeof(dev::perm_desc) + 255) & ~size_t(255);
        g.o_cd = sz; sz = (sz + 4 * cd.size() + 255) & ~size_t(255);
        g.o_co = sz; sz = (sz + 4 * co.size() + 255) & ~size_t(255);
        g.nchunks = cd.size();
        host.resize(sz, 0);
        std::memcpy(&host[g.o_d], lists[i].data(), lists[i].size() * sizeof(dev::perm_desc));
        std::memcpy(&host[g.o_cd], cd.data(), 4 * cd.size());
        std::memcpy(&host[g.o_co], co.data(), 4 * co.size());
        l->segs.push_back(g);
    }
    if(sz) {
        l->mem = ctx.alloc(sz);
        try {
            ctx.upload(l->mem, host.data(), sz);
        } catch(...) {
            ctx.free(l->mem);
            throw;
        }
    }
    return l;
}

//! Runs the launches of a cached entry (in order) and waits; ms (optional)
//! accumulates their device time.
void run_launches(context &ctx, const context::cached_launch &l, double *ms) {
    if(l.segs.empty()) return;
    ctx.activate();
    hipStream_t s = ctx.stream();
    hipEvent_t ev[2] = {0, 0};
    try {
        if(ms) {
            hip_check(hipEventCreate(&ev[0]), "hipEventCreate");
            hip_check(hipEventCreate(&ev[1]), "hipEventCreate");
            hip_check(hipEventRecord(ev[0], s), "hipEventRecord");
        }
        char *base = (char*)l.mem;
        for(size_t i = 0; i < l.segs.size(); i++) {
            const context::cached_launch::segment &g = l.segs[i];
            hip_check(dev::launch_permute((dev::perm_desc*)(base + g.o_d),
                (uint32_t*)(base + g.o_cd), (uint32_t*)(base + g.o_co),
                (uint32_t)g.nchunks, s), "permute");
        }
        if(ms) hip_check(hipEventRecord(ev[1], s), "hipEventRecord");
        ctx.sync();
        if(ms) {
            float t = 0.f;
            hip_check(hipEventElapsedTime(&t, ev[0], ev[1]), "hipEventElapsedTime");
            *ms += t;
        }
    } catch(...) {
        for(int i = 0; i < 2; i++) if(ev[i]) (void)hipEventDestroy(ev[i]);
        throw;
    }
    for(int i = 0; i < 2; i++) if(ev[i]) (void)hipEventDestroy(ev[i]);
}

std::string key_part(const tensor_tr &tr) {
    std::string k;
    char buf[64];
    for(size_t i = 0; i < tr.perm.n; i++) {
        snprintf(buf, sizeof(buf), "%u,", unsigned(tr.perm.p[i]));
        k += buf;
    }
    snprintf(buf, sizeof(buf), "%a", tr.st.c);
    return k + buf;
}

std::string key_tensor(const btensor &t, bool version) {
    char buf[64];
    snprintf(buf, sizeof(buf), "|%llu:%llu", (unsigned long long)t.uid(),
        version ? (unsigned long long)t.version() : 0ULL);
    return buf;
}

void add_into(btensor &b, btensor &s, const tensor_tr &tr, double d, double *ms) {
    b.require_mutable("add");
    std::vector<dev::perm_desc> descs;
    std::vector<size_t> touched;
    add_descs(b, s, tr, d, descs, touched);
    run_permutes(b.ctx(), descs, ms);
    for(size_t i = 0; i < touched.size(); i++) b.mark_nonzero(touched[i]);
}

void btod_copy(btensor &a, const tensor_tr &tra, btensor &b, bool add, double d) {
    b.require_mutable("btod_copy");
    context &ctx = a.ctx();
    ctx.require_device("btod_copy");
    lt_require(&b.ctx() == &ctx, LT_ERR_PARAM, "tensors of different contexts");
    lt_require(tra.perm.n == a.bis().order(), LT_ERR_PARAM, "btod_copy: permutation order");
    bis_t bisb = a.bis();
    bisb.permute(tra.perm);
    lt_require(bisb == b.bis(), LT_ERR_DIMS, "btod_copy: block index space of B");
    // the descriptor lists are kept per operands' structure versions (B's
    // too for +=, whose blocks decide between writing and adding)
    const std::string key = std::string(add ? "copy+" : "copy") + key_tensor(a, true) +
        key_tensor(b, add) + "|" + key_part(tra) + (add ? "|" + key_part(
        tensor_tr(perm_t(0), scalar_tr(d))) : std::string());
    double ms = 0.0;
    std::shared_ptr<context::cached_launch> l = ctx.find_launch(key);
    if(!l) {
        const symmetry symb = so_permute(a.sym(), tra.perm);
        std::vector<std::vector<dev::perm_desc> > lists(1);
        std::vector<size_t> touched;
        if(add) {
            // B += d P(A): B takes the symmetry both terms have
            // (gen_bto_copy::perform(btb, c) through gen_bto_aux_add)
            merge_symmetry(b, symb, &ms);
        } else {
            b.zero_all();
            b.set_symmetry(symb);
        }
        add_descs(b, a, tra, add ? d : 1.0, lists[0], touched);
        l = upload_launches(ctx, lists);
        l->touched = touched;
        l->sym = add ? b.sym() : symb;
        ctx.keep_launch(key, l);
    } else if(!add) {
        b.zero_all();
        b.set_symmetry(l->sym);
    }
    run_launches(ctx, *l, &ms);
    for(size_t i = 0; i < l->touched.size(); i++) b.mark_nonzero(l->touched[i]);
    ctx.set_last_op_ms(ms);
}

void btod_add(const std::vector<btensor*> &ops, const std::vector<tensor_tr> &trs,
    btensor &b, bool add, double d) {
    lt_require(!ops.empty() && ops.size() == trs.size(), LT_ERR_PARAM, "btod_add: operands");
    b.require_mutable("btod_add");
    context &ctx = b.ctx();
    ctx.require_device("btod_add");
    std::string key = std::string(add ? "add+" : "add") + key_tensor(b, add);
    for(size_t i = 0; i < ops.size(); i++) {
        lt_require(&ops[i]->ctx() == &ctx, LT_ERR_PARAM, "tensors of different contexts");
        lt_require(trs[i].perm.n == b.bis().order(), LT_ERR_PARAM, "btod_add: permutation order");
        bis_t bis = ops[i]->bis();
        bis.permute(trs[i].perm);
        lt_require(bis == b.bis(), LT_ERR_DIMS, "btod_add: block index space of an operand");
        key += key_tensor(*ops[i], true) + "|" + key_part(trs[i]);
    }
    if(add) key += "|" + key_part(tensor_tr(perm_t(0), scalar_tr(d)));
    double ms = 0.0;
    std::shared_ptr<context::cached_launch> l = ctx.find_launch(key);
    if(!l) {
        // the symmetry of the sum: the symmetry every permuted operand has
        // (gen_bto_add_impl.h:238-267, so_dirsum + so_merge per added operand)
        symmetry symr(b.bis());
        for(size_t i = 0; i < ops.size(); i++)
            symr = (i == 0) ? so_permute(ops[i]->sym(), trs[i].perm) :
                sum_symmetry(symr, ops[i]->sym(), &trs[i].perm);
        if(add) {
            merge_symmetry(b, symr, &ms);
        } else {
            b.zero_all();
            b.set_symmetry(symr);
        }
        // one launch per operand: no two descriptors of a launch share a
        // destination (each block of B lies in one orbit of an operand)
        std::vector<std::vector<dev::perm_desc> > lists(ops.size());
        std::vector<size_t> touched;
        for(size_t i = 0; i < ops.size(); i++)
            add_descs(b, *ops[i], trs[i], add ? d : 1.0, lists[i], touched);
        l = upload_launches(ctx, lists);
        l->touched = touched;
        l->sym = b.sym();
        ctx.keep_launch(key, l);
    } else if(!add) {
        b.zero_all();
        b.set_symmetry(l->sym);
    }
    run_launches(ctx, *l, &ms);
    for(size_t i = 0; i < l->touched.size(); i++) b.mark_nonzero(l->touched[i]);
    ctx.set_last_op_ms(ms);
}

} // namespace lta
