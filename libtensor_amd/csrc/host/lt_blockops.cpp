// btod_copy and btod_add on HBM-resident block tensors.
//
// btod_copy: B = c P(A), or B += d c P(A)
// (libtensor/block_tensor/btod_copy.h:20-80; gen_block_tensor/impl/
// gen_bto_copy_impl.h:150-163: the symmetry of B is so_permute(sym A, P);
// :302-330: every non-zero canonical block of A goes to the canonical block
// of the orbit of its image P(ia) in B, transformed by (P, c) and then by the
// inverse of that orbit's transformation).  All blocks are one launch of the
// batched permute kernel.
//
// btod_add: B = sum_i c_i P_i(A_i), or B += d sum_i ... (libtensor/
// block_tensor/btod_add.h; gen_block_tensor/impl/gen_bto_add_impl.h).  The
// reference takes the intersection of the operands' permuted symmetries
// (so_dirsum then so_merge, gen_bto_add_impl.h:238-267); this build adds
// operands whose permuted symmetries give every block the same orbit (the
// common case: terms of one equation), and rejects others.  One permute
// launch per operand (the first write of a block overwrites, later ones
// accumulate), so no two descriptors of a launch share a destination.
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

void btod_copy(btensor &a, const tensor_tr &tra, btensor &b, bool add, double d) {
    context &ctx = a.ctx();
    ctx.require_device("btod_copy");
    lt_require(&b.ctx() == &ctx, LT_ERR_PARAM, "tensors of different contexts");
    lt_require(tra.perm.n == a.bis().order(), LT_ERR_PARAM, "btod_copy: permutation order");
    bis_t bisb = a.bis();
    bisb.permute(tra.perm);
    lt_require(bisb == b.bis(), LT_ERR_DIMS, "btod_copy: block index space of B");
    const symmetry symb = so_permute(a.sym(), tra.perm);
    const dims_t bida = a.bidims(), bidb = b.bidims();
    if(add) {
        lt_require(same_orbits(symb, b.sym(), bidb), LT_ERR_SYMMETRY,
            "btod_copy: B += ... needs B with the symmetry of the result");
    } else {
        b.zero_all();
        b.set_symmetry(symb);
    }
    std::vector<dev::perm_desc> descs;
    const std::vector<size_t> nz = a.nonzero_blocks();
    for(size_t i = 0; i < nz.size(); i++) {
        const size_t ia = nz[i];
        index_t ib = get_index(ia, bida);
        tra.perm.apply(ib);
        const size_t xb = abs_index(ib, bidb);
        orbit ob(symb, xb, false);
        tensor_tr tr(tra);
        tr.transform(ob.get_transf(xb).inverse());
        if(add) tr.st.scale(d);
        const bool acc = add && b.is_nonzero(ob.acidx);
        dims_t da = a.bis().block_dims(get_index(ia, bida));
        descs.push_back(make_perm_desc_public(a.block_ptr(ia), da, tr, b.block_ptr(ob.acidx), acc));
        b.mark_nonzero(ob.acidx);
    }
    double ms = 0.0;
    run_permutes(ctx, descs, &ms);
    ctx.set_last_op_ms(ms);
}

void btod_add(const std::vector<btensor*> &ops, const std::vector<tensor_tr> &trs,
    btensor &b, bool add, double d) {
    lt_require(!ops.empty() && ops.size() == trs.size(), LT_ERR_PARAM, "btod_add: operands");
    context &ctx = b.ctx();
    ctx.require_device("btod_add");
    const dims_t bidb = b.bidims();
    std::vector<symmetry> syms;
    for(size_t i = 0; i < ops.size(); i++) {
        lt_require(&ops[i]->ctx() == &ctx, LT_ERR_PARAM, "tensors of different contexts");
        lt_require(trs[i].perm.n == b.bis().order(), LT_ERR_PARAM, "btod_add: permutation order");
        bis_t bis = ops[i]->bis();
        bis.permute(trs[i].perm);
        lt_require(bis == b.bis(), LT_ERR_DIMS, "btod_add: block index space of an operand");
        syms.push_back(so_permute(ops[i]->sym(), trs[i].perm));
        if(i > 0)
            lt_require(same_orbits(syms[0], syms[i], bidb), LT_ERR_SYMMETRY,
                "btod_add: operands whose permuted symmetries differ");
    }
    if(add) {
        lt_require(same_orbits(syms[0], b.sym(), bidb), LT_ERR_SYMMETRY,
            "btod_add: B += ... needs B with the symmetry of the result");
    } else {
        b.zero_all();
        b.set_symmetry(syms[0]);
    }
    double ms = 0.0;
    std::set<size_t> written;
    if(add) {
        const std::vector<size_t> nzb = b.nonzero_blocks();
        written.insert(nzb.begin(), nzb.end());
    }
    for(size_t i = 0; i < ops.size(); i++) {
        btensor &a = *ops[i];
        const dims_t bida = a.bidims();
        std::vector<dev::perm_desc> descs;
        const std::vector<size_t> nz = a.nonzero_blocks();
        for(size_t j = 0; j < nz.size(); j++) {
            index_t ib = get_index(nz[j], bida);
            trs[i].perm.apply(ib);
            const size_t xb = abs_index(ib, bidb);
            orbit ob(syms[0], xb, false);
            tensor_tr tr(trs[i]);
            tr.transform(ob.get_transf(xb).inverse());
            if(add) tr.st.scale(d);
            const bool acc = written.count(ob.acidx) != 0;
            dims_t da = a.bis().block_dims(get_index(nz[j], bida));
            descs.push_back(make_perm_desc_public(a.block_ptr(nz[j]), da, tr,
                b.block_ptr(ob.acidx), acc));
            written.insert(ob.acidx);
            b.mark_nonzero(ob.acidx);
        }
        run_permutes(ctx, descs, &ms);
    }
    ctx.set_last_op_ms(ms);
}

} // namespace lta
