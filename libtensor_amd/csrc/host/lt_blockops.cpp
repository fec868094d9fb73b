// btod_copy on HBM-resident block tensors: B = c P(A), or B += d c P(A)
// (libtensor/block_tensor/btod_copy.h:20-80; gen_block_tensor/impl/
// gen_bto_copy_impl.h:150-163: the symmetry of B is so_permute(sym A, P);
// :302-330: every non-zero canonical block of A goes to the canonical block
// of the orbit of its image P(ia) in B, transformed by (P, c) and then by the
// inverse of that orbit's transformation).  All blocks are one launch of the
// batched permute kernel.
#include "lt_runtime.h"

namespace lta {

namespace {

// Same canonical block, transformation and allowedness for every block.
bool same_orbits(const symmetry &x, const symmetry &y, const dims_t &bid) {
    for(size_t i = 0; i < bid.size(); i++) {
        orbit ox(x, i, true), oy(y, i, true);
        if(ox.acidx != oy.acidx || ox.allowed != oy.allowed) return false;
        if(!(ox.get_transf(i) == oy.get_transf(i))) return false;
    }
    return true;
}

} // unnamed namespace

void run_permutes(context &ctx, const std::vector<dev::perm_desc> &descs) {
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
    char *mem = (char*)ctx.alloc(bytes);
    try {
        ctx.upload(mem, host.data(), bytes);
        ctx.activate();
        hip_check(dev::launch_permute((dev::perm_desc*)mem, (uint32_t*)(mem + o_cd),
            (uint32_t*)(mem + o_co), (uint32_t)cd.size(), ctx.stream()), "permute");
        ctx.sync();
    } catch(...) {
        ctx.free(mem);
        throw;
    }
    ctx.free(mem);
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
    run_permutes(ctx, descs);
}

} // namespace lta
