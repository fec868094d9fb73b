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

#include <set>

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

void add_into(btensor &b, btensor &s, const tensor_tr &tr, double d, double *ms) {
    context &ctx = b.ctx();
    lt_require(&s.ctx() == &ctx, LT_ERR_PARAM, "tensors of different contexts");
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
    std::vector<dev::perm_desc> descs;
    std::vector<size_t> touched;
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
            descs.push_back(make_perm_desc_public(s.block_ptr(nz[i]), da, t, b.block_ptr(y),
                b.is_nonzero(y)));
            touched.push_back(y);
        }
    }
    run_permutes(ctx, descs, ms);
    for(size_t i = 0; i < touched.size(); i++) b.mark_nonzero(touched[i]);
}

void btod_copy(btensor &a, const tensor_tr &tra, btensor &b, bool add, double d) {
    context &ctx = a.ctx();
    ctx.require_device("btod_copy");
    lt_require(&b.ctx() == &ctx, LT_ERR_PARAM, "tensors of different contexts");
    lt_require(tra.perm.n == a.bis().order(), LT_ERR_PARAM, "btod_copy: permutation order");
    bis_t bisb = a.bis();
    bisb.permute(tra.perm);
    lt_require(bisb == b.bis(), LT_ERR_DIMS, "btod_copy: block index space of B");
    const symmetry symb = so_permute(a.sym(), tra.perm);
    double ms = 0.0;
    if(add) {
        // B += d P(A): B takes the symmetry both terms have
        // (gen_bto_copy::perform(btb, c) through gen_bto_aux_add)
        merge_symmetry(b, symb, &ms);
        add_into(b, a, tra, d, &ms);
    } else {
        b.zero_all();
        b.set_symmetry(symb);
        add_into(b, a, tra, 1.0, &ms);
    }
    ctx.set_last_op_ms(ms);
}

void btod_add(const std::vector<btensor*> &ops, const std::vector<tensor_tr> &trs,
    btensor &b, bool add, double d) {
    lt_require(!ops.empty() && ops.size() == trs.size(), LT_ERR_PARAM, "btod_add: operands");
    context &ctx = b.ctx();
    ctx.require_device("btod_add");
    // the symmetry of the sum: the symmetry every permuted operand has
    // (gen_bto_add_impl.h:238-267, so_dirsum + so_merge per added operand)
    symmetry symr(b.bis());
    for(size_t i = 0; i < ops.size(); i++) {
        lt_require(&ops[i]->ctx() == &ctx, LT_ERR_PARAM, "tensors of different contexts");
        lt_require(trs[i].perm.n == b.bis().order(), LT_ERR_PARAM, "btod_add: permutation order");
        bis_t bis = ops[i]->bis();
        bis.permute(trs[i].perm);
        lt_require(bis == b.bis(), LT_ERR_DIMS, "btod_add: block index space of an operand");
        // the first operand's permuted symmetry, then the sum with each next
        // operand (gen_bto_add::add_operand, gen_bto_add_impl.h:226-270)
        symr = (i == 0) ? so_permute(ops[i]->sym(), trs[i].perm) :
            sum_symmetry(symr, ops[i]->sym(), &trs[i].perm);
    }
    double ms = 0.0;
    if(add) {
        merge_symmetry(b, symr, &ms);
    } else {
        b.zero_all();
        b.set_symmetry(symr);
    }
    // one launch per operand: no two descriptors of a launch share a
    // destination (each block of B lies in one orbit of an operand)
    for(size_t i = 0; i < ops.size(); i++)
        add_into(b, *ops[i], trs[i], add ? d : 1.0, &ms);
    ctx.set_last_op_ms(ms);
}

} // namespace lta
