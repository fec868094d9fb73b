// btod_dotprod on HBM-resident block tensors: sum over all elements of
// tra(A) * trb(B) (libtensor/block_tensor/btod_dotprod.h:24-70,
// gen_block_tensor/impl/gen_bto_dotprod_impl.h:60-95).
//
// Every block X of the (permuted) index space contributes
// dot(A'[X], B'[X]) with A'[X] = tr_a(Ya) and B'[X] = tr_b(Yb) built from
// the stored canonical blocks Ya, Yb.  The block product is evaluated in
// Ya's index space, dot(tr_a(Ya), tr_b(Yb)) = c_a^2 dot(Ya, tr_a^-1 tr_b (Yb))
// (c_a the scalar of tr_a), so each block is one descriptor of the device
// dot kernel reading both canonical blocks in place.  Blocks forbidden by
// either symmetry or zero in either tensor contribute nothing.  All
// partial sums are added in a fixed order.
#include "lt_runtime.h"

namespace lta {

double btod_dotprod(btensor &a, const tensor_tr &tra0, btensor &b, const tensor_tr &trb0) {
    context &ctx = a.ctx();
    ctx.require_device("btod_dotprod");
    lt_require(&b.ctx() == &ctx, LT_ERR_PARAM, "tensors of different contexts");
    bis_t bisa2 = a.bis(), bisb2 = b.bis();
    bisa2.permute(tra0.perm);
    bisb2.permute(trb0.perm);
    lt_require(bisa2 == bisb2, LT_ERR_DIMS, "btod_dotprod: block index spaces differ");
    const dims_t bid2 = bisa2.block_index_dims(), bida = a.bidims(), bidb = b.bidims();
    const perm_t pai = tra0.perm.inverse(), pbi = trb0.perm.inverse();

    std::vector<dev::perm_desc> descs;
    for(size_t x = 0; x < bid2.size(); x++) {
        index_t ix = get_index(x, bid2);
        index_t ia = ix, ib = ix;
        pai.apply(ia);
        pbi.apply(ib);
        size_t xa = abs_index(ia, bida), xb = abs_index(ib, bidb);
        short_orbit sa(a.sym(), xa, true);
        if(!sa.allowed || !a.is_nonzero(sa.acidx)) continue;
        short_orbit sb(b.sym(), xb, true);
        if(!sb.allowed || !b.is_nonzero(sb.acidx)) continue;
        orbit oa(a.sym(), xa, false), ob(b.sym(), xb, false);
        tensor_tr ta = oa.get_transf(xa);
        ta.transform(tensor_tr(tra0.perm, scalar_tr()));
        tensor_tr tb = ob.get_transf(xb);
        tb.transform(tensor_tr(trb0.perm, scalar_tr()));
        tensor_tr t(tb);
        t.transform(ta.inverse());
        dims_t dyb = b.bis().block_dims(get_index(sb.acidx, bidb));
        dev::perm_desc d = make_perm_desc_public(b.block_ptr(sb.acidx), dyb, t,
            a.block_ptr(sa.acidx), false);
        d.alpha *= ta.st.c * ta.st.c;
        descs.push_back(d);
    }
    double scale = tra0.st.c * trb0.st.c;
    if(descs.empty()) return 0.0;

    std::vector<uint32_t> cd, co;
    for(size_t i = 0; i < descs.size(); i++)
        for(uint32_t off = 0; off < descs[i].total; off += dev::kPermChunk) {
            cd.push_back((uint32_t)i);
            co.push_back(off);
        }
    const size_t nch = cd.size();
    size_t o_d = 0;
    size_t o_cd = (descs.size() * sizeof(dev::perm_desc) + 255) & ~size_t(255);
    size_t o_co = o_cd + ((4 * nch + 255) & ~size_t(255));
    size_t o_p = o_co + ((4 * nch + 255) & ~size_t(255));
    size_t o_out = o_p + ((8 * nch + 255) & ~size_t(255));
    size_t bytes = o_out + 256;
    std::vector<char> host(bytes, 0);
    std::memcpy(&host[o_d], descs.data(), descs.size() * sizeof(dev::perm_desc));
    std::memcpy(&host[o_cd], cd.data(), 4 * nch);
    std::memcpy(&host[o_co], co.data(), 4 * nch);
    char *mem = (char*)ctx.alloc(bytes);
    double r = 0.0;
    try {
        ctx.upload(mem, host.data(), bytes);
        ctx.activate();
        hip_check(dev::launch_dot((dev::perm_desc*)(mem + o_d), (uint32_t*)(mem + o_cd),
            (uint32_t*)(mem + o_co), (uint32_t)nch, (double*)(mem + o_p),
            (double*)(mem + o_out), ctx.stream()), "dot");
        ctx.download(&r, mem + o_out, sizeof(double));
    } catch(...) {
        ctx.free(mem);
        throw;
    }
    ctx.free(mem);
    return r * scale;
}

} // namespace lta
