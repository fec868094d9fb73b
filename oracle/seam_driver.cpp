// Test infrastructure (checker only, never the product).
//
// Runs libtensor's own block contraction and, beside it, the MI355X adapter
// btod_contract2_amd (libtensor_amd/integration/btod_contract2_amd.h) on the
// same libtensor block tensors, built from a problem description
// (spec_format.h), and prints how they compare:
//
//   structure (any context, also host-only):
//     bis_match 0|1          block index spaces equal
//     orbit_mismatch n       blocks whose libtensor orbit under the adapter's
//                            get_symmetry() differs from btod_contract2's
//     lib_orbit_mismatch n   blocks whose orbit in the library (canonical
//                            index, allowedness, transformation) differs
//                            from libtensor's (verify_structure)
//     sched n_native n_amd match(0|1)
//   numbers (GPU context; --host-only skips them):
//     perform_err e          perform(btc): adapter vs btod_contract2
//     perform_add_err e      perform(btc, 0.5) into the target tensor
//     compute_block_err e    compute_block(zero, idx, tr, blk) on scheduled
//                            blocks with a transformation, zero and add
//     expr_direct_err e      letter expression vs btod_contract2 (both libtensor)
//     expr_assign_err e      R(...) = contract(...) through the evaluator,
//                            adapter (use_libxm(true) selects it, see the
//                            seam header under oracle/seam/) vs native
//     expr_add_err e         R(...) += contract(...) into the target tensor,
//                            adapter vs native
//
// The target tensor of the additive calls is tensor C of the spec when it
// is declared (its own symmetry and data), else the contraction's result.
// Errors are max|x - ref| / max|ref| over the unfolded tensors.
#include <libtensor/libtensor.h>
#include <libtensor/block_tensor/btod_contract2.h>
#include <libtensor/block_tensor/btod_copy.h>
#include <libtensor/core/batching_policy_base.h>
#include <libtensor/dense_tensor/tod_btconv.h>
#include <libtensor/expr/btensor/eval_btensor.h>
#include <libtensor/expr/operators/contract.h>

#include <cmath>
#include <cstdio>
#include <string>
#include <vector>

#include "spec_format.h"
#include "ref_build.h"
#include "../libtensor_amd/integration/btod_contract2_amd.h"

using namespace libtensor;
using ref_build::make_bis;
using ref_build::fill_tensor;

namespace {

template<size_t N>
std::vector<double> unfold(block_tensor_rd_i<N, double> &bt) {
    dimensions<N> d = bt.get_bis().get_dims();
    dense_tensor<N, double, allocator<double> > t(d);
    tod_btconv<N>(bt).perform(t);
    dense_tensor_rd_ctrl<N, double> c(t);
    const double *p = c.req_const_dataptr();
    std::vector<double> v(p, p + d.get_size());
    c.ret_const_dataptr(p);
    return v;
}

double rel_err(const std::vector<double> &x, const std::vector<double> &ref) {
    double num = 0.0, den = 0.0;
    for(size_t i = 0; i < ref.size(); i++) {
        num = std::max(num, std::fabs(x[i] - ref[i]));
        den = std::max(den, std::fabs(ref[i]));
    }
    return den > 0.0 ? num / den : num;
}

template<size_t N>
std::vector<double> block_data(dense_tensor_i<N, double> &t) {
    dense_tensor_rd_ctrl<N, double> c(t);
    const double *p = c.req_const_dataptr();
    std::vector<double> v(p, p + t.get_dims().get_size());
    c.ret_const_dataptr(p);
    return v;
}

//! Target of the additive calls: a copy of `init` (symmetry and data).
template<size_t N>
void copy_into(block_tensor<N, double, allocator<double> > &init,
    block_tensor<N, double, allocator<double> > &dst) {
    btod_copy<N>(init).perform(dst);
}

//! Letters of the operands and the result of contraction `contr`, as the
//! expression would name them: A index i -> l[i]; a B index contracted
//! with A index i -> l[i], another -> l[NA + j]; a result index -> the
//! letter of the operand index it is connected to.
template<size_t N, size_t M, size_t K>
struct letters {
    enum { NA = N + K, NB = M + K, NC = N + M };
    expr::letter l[NA + NB];
    std::vector<const expr::letter*> la, lb, lc, lk;
    explicit letters(const contraction2<N, M, K> &contr) {
        const sequence<2 * (N + M + K), size_t> &conn = contr.get_conn();
        std::vector<const expr::letter*> byb(NB, 0);
        for(size_t i = 0; i < NA; i++) {
            la.push_back(&l[i]);
            size_t j = conn[NC + i];
            if(j >= NC + NA) { byb[j - NC - NA] = &l[i]; lk.push_back(&l[i]); }
        }
        for(size_t j = 0; j < NB; j++) lb.push_back(byb[j] ? byb[j] : &l[NA + j]);
        for(size_t i = 0; i < NC; i++) {
            size_t j = conn[i] - NC;
            lc.push_back(j < NA ? la[j] : lb[j - NA]);
        }
    }
};

template<size_t N, size_t M, size_t K>
struct expr_runner {
    enum { NA = N + K, NB = M + K, NC = N + M };
    static void run(const contraction2<N, M, K> &contr, expr::btensor<NA> &a,
        expr::btensor<NB> &b, double sa, double sb, double sc, expr::btensor<NC> &r,
        bool add) {
        letters<N, M, K> L(contr);
        expr::label<NA> la(L.la);
        expr::label<NB> lb(L.lb);
        expr::label<NC> lc(L.lc);
        expr::label<K> lk(L.lk);
        if(add) r(lc) += sc * expr::contract(lk, sa * a(la), sb * b(lb));
        else r(lc) = sc * expr::contract(lk, sa * a(la), sb * b(lb));
    }
};

template<size_t N, size_t M>
struct expr_runner<N, M, 0> {
    static void run(const contraction2<N, M, 0>&, expr::btensor<N>&, expr::btensor<M>&,
        double, double, double, expr::btensor<N + M>&, bool) {
        throw std::runtime_error("no contracted index");
    }
};

template<size_t N, size_t M, size_t K>
int run(const spec::problem &p, FILE *f, bool host_only) {
    enum { NA = N + K, NB = M + K, NC = N + M };
    if(p.self) throw std::runtime_error("self-contraction: not driven here");
    block_index_space<NA> bisa = make_bis<NA>(p.a);
    block_index_space<NB> bisb = make_bis<NB>(p.b);
    expr::btensor<NA> bta(bisa);
    expr::btensor<NB> btb(bisb);
    fill_tensor<NA>(p.a, bta);
    fill_tensor<NB>(p.b, btb);
    bta.set_immutable();
    btb.set_immutable();

    sequence<NC, size_t> pcs(0), id(0);
    for(size_t i = 0; i < NC; i++) { pcs[i] = p.permc[i]; id[i] = i; }
    contraction2<N, M, K> contr(permutation_builder<NC>(pcs, id).get_perm());
    for(size_t k = 0; k < K; k++) contr.contract(p.ka[k], p.kb[k]);

    btod_contract2<N, M, K> nat(contr, bta, p.sa, btb, p.sb, p.sc);
    btod_contract2_amd<N, M, K> amd(contr, bta, p.sa, btb, p.sb, p.sc);

    // structure
    fprintf(f, "bis_match %d\n", nat.get_bis().equals(amd.get_bis()) ? 1 : 0);
    dimensions<NC> bid = nat.get_bis().get_block_index_dims();
    size_t bad = 0;
    for(size_t a = 0; a < bid.get_size(); a++) {
        orbit<NC, double> o1(nat.get_symmetry(), a), o2(amd.get_symmetry(), a);
        bool same = o1.is_allowed() == o2.is_allowed() && o1.get_acindex() == o2.get_acindex();
        if(same && o1.is_allowed()) {
            const tensor_transf<NC, double> &t1 = o1.get_transf(a), &t2 = o2.get_transf(a);
            same = t1.get_perm().equals(t2.get_perm()) &&
                t1.get_scalar_tr().get_coeff() == t2.get_scalar_tr().get_coeff();
        }
        if(!same) bad++;
    }
    fprintf(f, "orbit_mismatch %zu\n", bad);
    fprintf(f, "lib_orbit_mismatch %zu\n", amd.verify_structure());
    std::vector<size_t> s1, s2;
    const assignment_schedule<NC, double> &sn = nat.get_schedule(), &sa = amd.get_schedule();
    for(typename assignment_schedule<NC, double>::iterator i = sn.begin(); i != sn.end(); ++i)
        s1.push_back(sn.get_abs_index(i));
    for(typename assignment_schedule<NC, double>::iterator i = sa.begin(); i != sa.end(); ++i)
        s2.push_back(sa.get_abs_index(i));
    std::sort(s1.begin(), s1.end());
    std::sort(s2.begin(), s2.end());
    fprintf(f, "sched %zu %zu %d\n", s1.size(), s2.size(), s1 == s2 ? 1 : 0);
    fflush(f);
    if(host_only) return 0;

    // perform(btc)
    expr::btensor<NC> r1(nat.get_bis()), r2(nat.get_bis());
    nat.perform(r1);
    amd.perform(r2);
    std::vector<double> ref = unfold<NC>(r1);
    fprintf(f, "perform_err %.3e\n", rel_err(unfold<NC>(r2), ref));
    fflush(f);

    // the target of the additive calls
    block_tensor<NC, double, allocator<double> > *tgt0 = 0;
    if(p.c.n == NC) {
        tgt0 = new block_tensor<NC, double, allocator<double> >(make_bis<NC>(p.c));
        fill_tensor<NC>(p.c, *tgt0);
    } else {
        tgt0 = new block_tensor<NC, double, allocator<double> >(nat.get_bis());
        btod_copy<NC>(r1).perform(*tgt0);
    }

    // perform(btc, d)
    {
        expr::btensor<NC> t1(tgt0->get_bis()), t2(tgt0->get_bis());
        copy_into<NC>(*tgt0, t1);
        copy_into<NC>(*tgt0, t2);
        nat.perform(t1, scalar_transf<double>(0.5));
        amd.perform(t2, scalar_transf<double>(0.5));
        fprintf(f, "perform_add_err %.3e\n", rel_err(unfold<NC>(t2), unfold<NC>(t1)));
        fflush(f);
    }

    // compute_block with a scalar and the inverse of a block's orbit
    // permutation (same dimensions), zero and accumulate
    {
        double worst = 0.0;
        size_t nb = 0;
        for(size_t i = 0; i < s1.size() && nb < 8; i++, nb++) {
            libtensor::index<NC> bi;
            abs_index<NC>::get_index(s1[i], bid, bi);
            dimensions<NC> bd = nat.get_bis().get_block_dims(bi);
            tensor_transf<NC, double> tr(permutation<NC>(), scalar_transf<double>(-1.5));
            dense_tensor<NC, double, allocator<double> > b1(bd), b2(bd);
            nat.compute_block(true, bi, tr, b1);
            amd.compute_block(true, bi, tr, b2);
            worst = std::max(worst, rel_err(block_data<NC>(b2), block_data<NC>(b1)));
            nat.compute_block(false, bi, tr, b1);
            amd.compute_block(false, bi, tr, b2);
            worst = std::max(worst, rel_err(block_data<NC>(b2), block_data<NC>(b1)));
        }
        fprintf(f, "compute_block_err %.3e %zu\n", worst, nb);
        fflush(f);
    }

    // letter expressions through the evaluator
    if(K > 0) {
        expr::eval_btensor<double>::use_libxm(false);
        expr::btensor<NC> e1(nat.get_bis());
        expr_runner<N, M, K>::run(contr, bta, btb, p.sa, p.sb, p.sc, e1, false);
        fprintf(f, "expr_direct_err %.3e\n", rel_err(unfold<NC>(e1), ref));
        expr::eval_btensor<double>::use_libxm(true);
        expr::btensor<NC> e2(nat.get_bis());
        expr_runner<N, M, K>::run(contr, bta, btb, p.sa, p.sb, p.sc, e2, false);
        expr::eval_btensor<double>::use_libxm(false);
        fprintf(f, "expr_assign_err %.3e\n", rel_err(unfold<NC>(e2), unfold<NC>(e1)));
        fflush(f);

        expr::btensor<NC> t1(tgt0->get_bis()), t2(tgt0->get_bis());
        copy_into<NC>(*tgt0, t1);
        copy_into<NC>(*tgt0, t2);
        expr_runner<N, M, K>::run(contr, bta, btb, p.sa, p.sb, p.sc, t1, true);
        expr::eval_btensor<double>::use_libxm(true);
        expr_runner<N, M, K>::run(contr, bta, btb, p.sa, p.sb, p.sc, t2, true);
        expr::eval_btensor<double>::use_libxm(false);
        fprintf(f, "expr_add_err %.3e\n", rel_err(unfold<NC>(t2), unfold<NC>(t1)));
        fflush(f);
    }
    delete tgt0;
    return 0;
}

// The orders of ref_driver.cpp.
int dispatch(const spec::problem &p, FILE *f, bool h) {
#define LT_CASE(n, m, k) if(p.N == n && p.M == m && p.K == k) return run<n, m, k>(p, f, h);
    LT_CASE(1, 1, 0) LT_CASE(2, 1, 0) LT_CASE(1, 2, 0) LT_CASE(2, 2, 0)
    LT_CASE(0, 1, 1) LT_CASE(1, 0, 1) LT_CASE(1, 1, 1) LT_CASE(2, 1, 1) LT_CASE(1, 2, 1)
    LT_CASE(2, 2, 1) LT_CASE(3, 1, 1) LT_CASE(1, 3, 1) LT_CASE(3, 3, 1)
    LT_CASE(0, 2, 2) LT_CASE(2, 0, 2) LT_CASE(1, 1, 2) LT_CASE(2, 1, 2) LT_CASE(1, 2, 2)
    LT_CASE(2, 2, 2) LT_CASE(1, 1, 3) LT_CASE(2, 2, 3) LT_CASE(1, 3, 3)
#undef LT_CASE
    fprintf(stderr, "unsupported orders N=%zu M=%zu K=%zu\n", p.N, p.M, p.K);
    return 2;
}

} // unnamed namespace

int main(int argc, char **argv) {
    if(argc < 3) {
        fprintf(stderr, "usage: seam_driver <spec> <out.txt> [--host-only]\n");
        return 1;
    }
    bool host_only = argc > 3 && std::string(argv[3]) == "--host-only";
    try {
        spec::problem p = spec::parse(argv[1]);
        for(size_t t = 0; t < p.tables.size(); t++) ref_build::add_table(p.tables[t]);
        allocator<double>::init(4, 16, 16777216, 16777216);
        batching_policy_base::set_batch_size(size_t(1) << 40);
        FILE *f = fopen(argv[2], "w");
        if(!f) { fprintf(stderr, "cannot write %s\n", argv[2]); return 1; }
        int rc = dispatch(p, f, host_only);
        fclose(f);
        return rc;
    } catch(const std::exception &e) {
        fprintf(stderr, "seam_driver: %s\n", e.what());
        return 1;
    }
}
