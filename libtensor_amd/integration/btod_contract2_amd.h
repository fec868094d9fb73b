// btod_contract2_amd<N, M, K>: libtensor's btod_contract2 backed by the
// MI355X library (include/lt_amd.h).  Reference-side binding: a header that
// is compiled with libtensor's own headers (C++11) and linked against
// liblt_amd.so; nothing of libtensor is linked into the library itself.
//
// It implements additive_gen_bto<N + M, btod_traits::bti_traits>
// (libtensor/gen_block_tensor/additive_gen_bto.h:26-68,
// direct_gen_bto.h:22-54) with the constructors of btod_contract2
// (libtensor/block_tensor/btod_contract2.h:68-90), so it can stand wherever
// a btod_contract2 is used: directly, or as the operation the expression
// evaluator builds for contract() (the seam that libxm's btod_contract2_xm
// already uses, libtensor/expr/btensor/impl/eval_btensor_double_contract.C:205-221).
//
//   get_bis, get_symmetry   block index space and symmetry of the result as
//                           libtensor objects (gen_bto_contract2_sym, the
//                           host code that computes them today);
//                           verify_structure() checks the library's own
//                           orbits against them block by block
//   get_schedule            the library's non-zero schedule
//                           (lt_btod_contract2_schedule)
//   perform(stream)         one device perform into a resident result mirror,
//                           then every scheduled block put into the stream
//                           (as gen_bto_contract2_task::perform,
//                           gen_bto_contract2_batch_impl.h:339-362)
//   perform(btc)            gen_bto_aux_copy around perform(stream)
//   perform(btc, d)         addition_schedule + gen_bto_aux_add around
//                           perform(stream) (btod_contract2_impl.h:70-90): a
//                           target of any symmetry
//   compute_block           lt_btod_contract2_compute_block on the device
//
// Operands are mirrored into device-resident block tensors when the adapter
// is constructed (block index space, symmetry elements, canonical non-zero
// blocks); the result mirror and its execution plan live as long as the
// adapter, so repeated performs reuse the plan.  The device context is one
// per process (lt_amd_binding::context(); LT_AMD_DEVICE selects the GPU, -1
// a host-only context on which only the structure calls work).
#ifndef LIBTENSOR_BTOD_CONTRACT2_AMD_H
#define LIBTENSOR_BTOD_CONTRACT2_AMD_H

#include <cstdlib>
#include <cstring>
#include <mutex>
#include <set>
#include <string>
#include <vector>

#include <libtensor/core/abs_index.h>
#include <libtensor/core/contraction2.h>
#include <libtensor/core/noncopyable.h>
#include <libtensor/core/orbit.h>
#include <libtensor/core/permutation_builder.h>
#include <libtensor/core/scalar_transf_double.h>
#include <libtensor/core/tensor_transf_double.h>
#include <libtensor/dense_tensor/dense_tensor.h>
#include <libtensor/dense_tensor/dense_tensor_ctrl.h>
#include <libtensor/block_tensor/block_tensor_i.h>
#include <libtensor/block_tensor/btod_traits.h>
#include <libtensor/gen_block_tensor/additive_gen_bto.h>
#include <libtensor/gen_block_tensor/addition_schedule.h>
#include <libtensor/gen_block_tensor/assignment_schedule.h>
#include <libtensor/gen_block_tensor/gen_block_tensor_ctrl.h>
#include <libtensor/gen_block_tensor/gen_bto_aux_add.h>
#include <libtensor/gen_block_tensor/gen_bto_aux_copy.h>
#include <libtensor/gen_block_tensor/impl/gen_bto_contract2_sym.h>
#include <libtensor/symmetry/product_table_container.h>
#include <libtensor/symmetry/se_label.h>
#include <libtensor/symmetry/se_part.h>
#include <libtensor/symmetry/se_perm.h>
#include <libtensor/symmetry/symmetry_element_set_adapter.h>

#include "../../include/lt_amd.h"

namespace libtensor {
namespace lt_amd_binding {

//! Turns a negative status of the C ABI into libtensor's exception of the
//! same meaning (the ABI's codes follow libtensor's exception classes).
inline void check(int rc, const char *method) {
    if(rc == LT_OK) return;
    const char *ns = "libtensor", *cl = "btod_contract2_amd";
    std::string msg = lt_last_error();
    switch(rc) {
    case LT_ERR_PARAM:
        throw bad_parameter(ns, cl, method, __FILE__, __LINE__, msg.c_str());
    case LT_ERR_NOMEM:
        throw out_of_memory(ns, cl, method, __FILE__, __LINE__, msg.c_str());
    default:
        throw generic_exception(ns, cl, method, __FILE__, __LINE__, msg.c_str());
    }
}

//! The process-wide context (LT_AMD_DEVICE, default 0; < 0: host-only).
inline lt_ctx *context() {
    static std::mutex mtx;
    static lt_ctx *ctx = 0;
    std::lock_guard<std::mutex> lk(mtx);
    if(!ctx) {
        const char *e = std::getenv("LT_AMD_DEVICE");
        int dev = e ? std::atoi(e) : 0;
        check(lt_ctx_create(dev, 0, &ctx), "context");
    }
    return ctx;
}

//! Registers libtensor's product table `id` with the library (once).
inline void register_table(const std::string &id) {
    static std::mutex mtx;
    static std::set<std::string> done;
    std::lock_guard<std::mutex> lk(mtx);
    if(done.count(id)) return;
    product_table_container &ptc = product_table_container::get_instance();
    const product_table_i &pt = ptc.req_const_table(id);
    size_t n = pt.get_n_labels();
    try {
        check(lt_product_table_create(id.c_str(), n), "register_table");
        for(size_t l1 = 0; l1 < n; l1++) for(size_t l2 = 0; l2 < n; l2++) {
            product_table_i::label_group_t lg(2);
            lg[0] = l1; lg[1] = l2;
            product_table_i::label_set_t ls;
            pt.product(lg, ls);
            for(product_table_i::label_set_t::const_iterator i = ls.begin();
                i != ls.end(); ++i)
                check(lt_product_table_add_product(id.c_str(), l1, l2, *i),
                    "register_table");
        }
    } catch(...) {
        ptc.ret_table(id);
        throw;
    }
    ptc.ret_table(id);
    done.insert(id);
}

//! Owned symmetry element handle.
struct symel {
    lt_symel *e;
    symel() : e(0) { }
    ~symel() { if(e) lt_symel_destroy(e); }
};

//! Owned block tensor handle.
struct btensor_h {
    lt_btensor *t;
    btensor_h() : t(0) { }
    ~btensor_h() { if(t) lt_btensor_destroy(t); }
    void reset(lt_btensor *x) { if(t) lt_btensor_destroy(t); t = x; }
};

//! Empty block tensor of libtensor's block index space `bis` (dimensions
//! and every split point of every dimension type).
template<size_t N>
lt_btensor *make_btensor(const block_index_space<N> &bis) {
    size_t dims[N];
    const dimensions<N> &d = bis.get_dims();
    for(size_t i = 0; i < N; i++) dims[i] = d[i];
    lt_btensor *t = 0;
    check(lt_btensor_create(context(), N, dims, &t), "mirror");
    std::set<size_t> types;
    for(size_t i = 0; i < N; i++) types.insert(bis.get_type(i));
    for(std::set<size_t>::const_iterator ty = types.begin(); ty != types.end(); ++ty) {
        int mask[N];
        for(size_t i = 0; i < N; i++) mask[i] = bis.get_type(i) == *ty ? 1 : 0;
        const split_points &sp = bis.get_splits(*ty);
        for(size_t k = 0; k < sp.get_num_points(); k++) {
            int rc = lt_btensor_split(t, mask, sp[k]);
            if(rc != LT_OK) { lt_btensor_destroy(t); check(rc, "mirror"); }
        }
    }
    return t;
}

//! Inserts libtensor's symmetry elements (se_perm, se_part, se_label, in
//! the order the symmetry holds them) into the mirror `t`.
template<size_t N>
void mirror_symmetry(const symmetry<N, double> &sym, lt_btensor *t) {
    for(typename symmetry<N, double>::iterator is = sym.begin(); is != sym.end(); ++is) {
        const symmetry_element_set<N, double> &set = sym.get_subset(is);
        const std::string &id = set.get_id();
        if(id == se_perm<N, double>::k_sym_type) {
            symmetry_element_set_adapter<N, double, se_perm<N, double> > ad(set);
            for(typename symmetry_element_set_adapter<N, double,
                    se_perm<N, double> >::iterator i = ad.begin(); i != ad.end(); ++i) {
                const se_perm<N, double> &e = ad.get_elem(i);
                size_t p[N];
                for(size_t k = 0; k < N; k++) p[k] = e.get_perm()[k];
                symel h;
                check(lt_se_perm_create(N, p, e.get_transf().get_coeff(), &h.e), "mirror");
                check(lt_btensor_insert(t, h.e), "mirror");
            }
        } else if(id == se_part<N, double>::k_sym_type) {
            symmetry_element_set_adapter<N, double, se_part<N, double> > ad(set);
            for(typename symmetry_element_set_adapter<N, double,
                    se_part<N, double> >::iterator i = ad.begin(); i != ad.end(); ++i) {
                const se_part<N, double> &e = ad.get_elem(i);
                const dimensions<N> &pd = e.get_pdims();
                size_t pdims[N];
                for(size_t k = 0; k < N; k++) pdims[k] = pd[k];
                symel h;
                check(lt_se_part_create(t, pdims, &h.e), "mirror");
                abs_index<N> ai(pd);
                do {
                    const index<N> &i1 = ai.get_index();
                    size_t x1[N], x2[N];
                    for(size_t k = 0; k < N; k++) x1[k] = i1[k];
                    if(e.is_forbidden(i1)) {
                        check(lt_se_part_mark_forbidden(h.e, x1), "mirror");
                        continue;
                    }
                    const index<N> &i2 = e.get_direct_map(i1);
                    if(i2 == i1) continue;
                    for(size_t k = 0; k < N; k++) x2[k] = i2[k];
                    check(lt_se_part_add_map(h.e, x1, x2,
                        e.get_transf(i1, i2).get_coeff()), "mirror");
                } while(ai.inc());
                check(lt_btensor_insert(t, h.e), "mirror");
            }
        } else if(id == se_label<N, double>::k_sym_type) {
            symmetry_element_set_adapter<N, double, se_label<N, double> > ad(set);
            for(typename symmetry_element_set_adapter<N, double,
                    se_label<N, double> >::iterator i = ad.begin(); i != ad.end(); ++i) {
                const se_label<N, double> &e = ad.get_elem(i);
                register_table(e.get_table_id());
                symel h;
                check(lt_se_label_create(t, e.get_table_id().c_str(), &h.e), "mirror");
                const block_labeling<N> &bl = e.get_labeling();
                const dimensions<N> &bid = bl.get_block_index_dims();
                for(size_t d = 0; d < N; d++) {
                    size_t ty = bl.get_dim_type(d);
                    int mask[N];
                    for(size_t k = 0; k < N; k++) mask[k] = k == d ? 1 : 0;
                    for(size_t b = 0; b < bid[d]; b++) {
                        size_t l = bl.get_label(ty, b);
                        if(l == product_table_i::k_invalid) continue;
                        check(lt_se_label_assign(h.e, mask, b, l), "mirror");
                    }
                }
                // evaluation rule: a disjunction of products of terms
                const evaluation_rule<N> &r = e.get_rule();
                const eval_sequence_list<N> &sl = r.get_sequences();
                check(lt_se_label_clear_rule(h.e), "mirror");
                for(typename evaluation_rule<N>::iterator ip = r.begin(); ip != r.end(); ++ip) {
                    const product_rule<N> &pr = r.get_product(ip);
                    std::vector<size_t> seqs, intr;
                    for(typename product_rule<N>::iterator it = pr.begin(); it != pr.end(); ++it) {
                        const sequence<N, size_t> &s = sl[pr.get_seqno(it)];
                        for(size_t k = 0; k < N; k++) seqs.push_back(s[k]);
                        intr.push_back(pr.get_intrinsic(it));
                    }
                    check(lt_se_label_add_product(h.e, intr.size(),
                        seqs.empty() ? 0 : &seqs[0], intr.empty() ? 0 : &intr[0]), "mirror");
                }
                check(lt_btensor_insert(t, h.e), "mirror");
            }
        } else {
            std::string msg = "symmetry elements of type " + id;
            throw bad_parameter("libtensor", "btod_contract2_amd", "mirror",
                __FILE__, __LINE__, msg.c_str());
        }
    }
}

//! Device-resident mirror of a libtensor block tensor: block index space,
//! symmetry and every canonical non-zero block.
template<size_t N>
lt_btensor *mirror(block_tensor_rd_i<N, double> &bt) {
    typedef block_tensor_i_traits<double> bti_traits;
    gen_block_tensor_rd_ctrl<N, bti_traits> ctrl(bt);
    btensor_h h;
    h.t = make_btensor<N>(bt.get_bis());
    mirror_symmetry<N>(ctrl.req_const_symmetry(), h.t);
    std::vector<size_t> nz;
    ctrl.req_nonzero_blocks(nz);
    dimensions<N> bidims = bt.get_bis().get_block_index_dims();
    for(size_t i = 0; i < nz.size(); i++) {
        index<N> bi;
        abs_index<N>::get_index(nz[i], bidims, bi);
        dense_tensor_rd_i<N, double> &blk = ctrl.req_const_block(bi);
        {
            dense_tensor_rd_ctrl<N, double> dc(blk);
            const double *p = dc.req_const_dataptr();
            int rc = lt_btensor_set_block(h.t, nz[i], p);
            dc.ret_const_dataptr(p);
            if(rc != LT_OK) { ctrl.ret_const_block(bi); check(rc, "mirror"); }
        }
        ctrl.ret_const_block(bi);
    }
    lt_btensor *t = h.t;
    h.t = 0;
    return t;
}

//! (permc, ka, kb) of the C ABI for libtensor's contraction2: ka / kb from
//! the contracted connections (ascending A index), permc the position in
//! contraction2's default result order (uncontracted A indexes, then
//! uncontracted B indexes) of each result index; checked by rebuilding the
//! connections.
template<size_t N, size_t M, size_t K>
void decode_contraction(const contraction2<N, M, K> &contr, std::vector<size_t> &permc,
    std::vector<size_t> &ka, std::vector<size_t> &kb) {
    enum { NA = N + K, NB = M + K, NC = N + M };
    const sequence<2 * (N + M + K), size_t> &conn = contr.get_conn();
    ka.clear(); kb.clear();
    size_t defpos[NA + NB];
    size_t nfree = 0;
    for(size_t ia = 0; ia < NA; ia++) {
        size_t j = conn[NC + ia];
        if(j >= NC + NA) { ka.push_back(ia); kb.push_back(j - NC - NA); }
        else defpos[ia] = nfree++;
    }
    for(size_t ib = 0; ib < NB; ib++) {
        size_t j = conn[NC + NA + ib];
        if(j < NC) defpos[NA + ib] = nfree++;
    }
    permc.assign(NC, 0);
    for(size_t ic = 0; ic < NC; ic++) permc[ic] = defpos[conn[ic] - NC];
    // the C ABI reads permc as libtensor's permutation of the default order
    // (contraction2(perm), permutation_builder of the sequence); rebuild and
    // compare, taking the inverse if the direction differs
    for(int attempt = 0; attempt < 2; attempt++) {
        sequence<NC, size_t> s1(0), s2(0);
        for(size_t i = 0; i < NC; i++) { s1[i] = permc[i]; s2[i] = i; }
        permutation<NC> pc = permutation_builder<NC>(s1, s2).get_perm();
        contraction2<N, M, K> c2(pc);
        for(size_t k = 0; k < K; k++) c2.contract(ka[k], kb[k]);
        bool same = true;
        for(size_t i = 0; i < 2 * (N + M + K); i++)
            if(c2.get_conn()[i] != conn[i]) same = false;
        if(same) return;
        std::vector<size_t> inv(NC);
        for(size_t i = 0; i < NC; i++) inv[permc[i]] = i;
        permc = inv;
    }
    throw generic_exception("libtensor", "btod_contract2_amd", "decode_contraction",
        __FILE__, __LINE__, "contraction could not be expressed in the C ABI");
}

} // namespace lt_amd_binding


template<size_t N, size_t M, size_t K>
class btod_contract2_amd :
    public additive_gen_bto<N + M, typename btod_traits::bti_traits>,
    public noncopyable {
public:
    enum { NA = N + K, NB = M + K, NC = N + M };
    typedef typename btod_traits::bti_traits bti_traits;
    typedef tensor_transf<NC, double> tensor_transf_type;

private:
    gen_bto_contract2_sym<N, M, K, btod_traits> m_symc;
    assignment_schedule<NC, double> m_sch;
    lt_amd_binding::btensor_h m_a, m_b, m_c;
    lt_btod_contract2 *m_op;
    int m_mode;

public:
    btod_contract2_amd(const contraction2<N, M, K> &contr,
        block_tensor_rd_i<NA, double> &bta, block_tensor_rd_i<NB, double> &btb) :
        m_symc(contr, bta, btb), m_sch(m_symc.get_bis().get_block_index_dims()),
        m_op(0), m_mode(LT_MODE_GROUPED) {
        init(contr, bta, 1.0, btb, 1.0, 1.0);
    }

    btod_contract2_amd(const contraction2<N, M, K> &contr,
        block_tensor_rd_i<NA, double> &bta, double ka,
        block_tensor_rd_i<NB, double> &btb, double kb, double kc = 1.0) :
        m_symc(contr, bta, btb), m_sch(m_symc.get_bis().get_block_index_dims()),
        m_op(0), m_mode(LT_MODE_GROUPED) {
        init(contr, bta, ka, btb, kb, kc);
    }

    virtual ~btod_contract2_amd() {
        if(m_op) lt_btod_contract2_destroy(m_op);
    }

    //! Launch mode of the device perform (LT_MODE_GROUPED by default).
    void set_mode(int mode) { m_mode = mode; }

    virtual const block_index_space<NC> &get_bis() const { return m_symc.get_bis(); }
    virtual const symmetry<NC, double> &get_symmetry() const { return m_symc.get_symmetry(); }
    virtual const assignment_schedule<NC, double> &get_schedule() const { return m_sch; }

    //! Every block of C: the library's orbit (canonical index, allowedness,
    //! transformation) equals libtensor's under get_symmetry().  Returns the
    //! number of blocks that differ.
    size_t verify_structure() const {
        dimensions<NC> bid = get_bis().get_block_index_dims();
        size_t bad = 0;
        for(size_t a = 0; a < bid.get_size(); a++) {
            orbit<NC, double> o(get_symmetry(), a);
            size_t ac = 0, perm[NC];
            double coeff = 0.0;
            int allowed = 0;
            lt_amd_binding::check(lt_btod_contract2_symc_orbit(m_op, a, &ac, perm, &coeff,
                &allowed), "verify_structure");
            bool same = (allowed != 0) == o.is_allowed() && ac == o.get_acindex();
            if(same && allowed) {
                const tensor_transf<NC, double> &tr = o.get_transf(a);
                for(size_t i = 0; i < NC; i++) if(tr.get_perm()[i] != perm[i]) same = false;
                if(tr.get_scalar_tr().get_coeff() != coeff) same = false;
            }
            if(!same) bad++;
        }
        return bad;
    }

    virtual void perform(gen_block_stream_i<NC, bti_traits> &out) {
        if(!m_c.t) m_c.t = lt_amd_binding::make_btensor<NC>(get_bis());
        lt_amd_binding::check(lt_btod_contract2_perform(m_op, m_c.t, m_mode), "perform");
        dimensions<NC> bid = get_bis().get_block_index_dims();
        tensor_transf<NC, double> tr0;
        for(typename assignment_schedule<NC, double>::iterator i = m_sch.begin();
            i != m_sch.end(); ++i) {
            size_t a = m_sch.get_abs_index(i);
            index<NC> bi;
            abs_index<NC>::get_index(a, bid, bi);
            dense_tensor<NC, double, allocator<double> > blk(get_bis().get_block_dims(bi));
            {
                dense_tensor_wr_ctrl<NC, double> dc(blk);
                double *p = dc.req_dataptr();
                int rc = lt_btensor_get_block(m_c.t, a, p);
                dc.ret_dataptr(p);
                lt_amd_binding::check(rc, "perform");
            }
            out.put(bi, blk, tr0);
        }
    }

    virtual void perform(gen_block_tensor_i<NC, bti_traits> &btc) {
        gen_bto_aux_copy<NC, btod_traits> out(get_symmetry(), btc);
        out.open();
        perform(out);
        out.close();
    }

    virtual void perform(gen_block_tensor_i<NC, bti_traits> &btc,
        const scalar_transf<double> &d) {
        gen_block_tensor_rd_ctrl<NC, bti_traits> cc(btc);
        std::vector<size_t> nzblkc;
        cc.req_nonzero_blocks(nzblkc);
        addition_schedule<NC, btod_traits> asch(get_symmetry(), cc.req_const_symmetry());
        asch.build(get_schedule(), nzblkc);
        gen_bto_aux_add<NC, btod_traits> out(get_symmetry(), asch, btc, d);
        out.open();
        perform(out);
        out.close();
    }

    void perform(block_tensor_i<NC, double> &btc, double d) {
        perform(btc, scalar_transf<double>(d));
    }

    virtual void compute_block(bool zero, const index<NC> &ic, const tensor_transf_type &trc,
        dense_tensor_wr_i<NC, double> &blkc) {
        dimensions<NC> bid = get_bis().get_block_index_dims();
        size_t a = abs_index<NC>::get_abs_index(ic, bid);
        size_t perm[NC];
        for(size_t i = 0; i < NC; i++) perm[i] = trc.get_perm()[i];
        size_t n = blkc.get_dims().get_size();
        lt_ctx *ctx = lt_amd_binding::context();
        void *d = 0;
        lt_amd_binding::check(lt_dev_alloc(ctx, n * sizeof(double) + 8, &d), "compute_block");
        dense_tensor_wr_ctrl<NC, double> dc(blkc);
        double *p = dc.req_dataptr();
        int rc = LT_OK;
        if(!zero) rc = lt_dev_upload(ctx, d, p, n * sizeof(double));
        if(rc == LT_OK)
            rc = lt_btod_contract2_compute_block(m_op, zero ? 1 : 0, a, perm,
                trc.get_scalar_tr().get_coeff(), (double*)d);
        if(rc == LT_OK) rc = lt_dev_download(ctx, p, d, n * sizeof(double));
        dc.ret_dataptr(p);
        lt_dev_free(ctx, d);
        lt_amd_binding::check(rc, "compute_block");
    }

    using additive_gen_bto<NC, bti_traits>::compute_block;

private:
    void init(const contraction2<N, M, K> &contr, block_tensor_rd_i<NA, double> &bta,
        double ka, block_tensor_rd_i<NB, double> &btb, double kb, double kc) {
        std::vector<size_t> permc, ia, ib;
        lt_amd_binding::decode_contraction<N, M, K>(contr, permc, ia, ib);
        m_a.reset(lt_amd_binding::mirror<NA>(bta));
        lt_btensor *b = m_a.t;
        if((void*)&btb != (void*)&bta) {
            m_b.reset(lt_amd_binding::mirror<NB>(btb));
            b = m_b.t;
        }
        lt_amd_binding::check(lt_btod_contract2_create(N, M, K, &permc[0],
            K ? &ia[0] : 0, K ? &ib[0] : 0, m_a.t, ka, b, kb, kc, &m_op), "init");
        size_t cnt = 0;
        lt_amd_binding::check(lt_btod_contract2_schedule(m_op, 0, 0, &cnt), "init");
        std::vector<size_t> sch(cnt);
        if(cnt) lt_amd_binding::check(lt_btod_contract2_schedule(m_op, &sch[0], cnt, &cnt),
            "init");
        for(size_t i = 0; i < sch.size(); i++) m_sch.insert(sch[i]);
    }
};

} // namespace libtensor

#endif // LIBTENSOR_BTOD_CONTRACT2_AMD_H
