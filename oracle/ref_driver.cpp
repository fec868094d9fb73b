// Test infrastructure (checker only, never the product).
//
// Driver of the REAL libtensor block-contraction path, built from the
// reference sources by oracle/Makefile into oracle/_ref/ref_driver.  It reads
// a problem description (spec_format.h), builds the block tensors with
// libtensor's own classes, runs btod_contract2 and writes:
//   <out>.txt         result symmetry (canonical index and allowedness of every
//                     C block), the non-zero schedule, the contraction list of
//                     every scheduled block (gen_bto_contract2_clst_builder,
//                     same inputs as gen_bto_contract2_nzorb), timings
//   <out>.result.bin  the result unfolded to a dense row-major array
//                     (tod_btconv)
// With a contract3 line the problem is btod_contract3 of A, B and C (same
// outputs as with symmetrize2).
// With symmetrize2 lines in the spec the contraction is wrapped in
// btod_symmetrize2 (innermost first); <out>.txt then holds the symmetry of the
// symmetrized result with the orbit transformation of every block ("symt"),
// its schedule, and <out>.result.bin its result.
// Options: --threads T (libutil::thread_pool of T workers), --repeat R
// (timed performs after one warm-up), --no-symc, --no-clst, --no-result.
#include <libtensor/libtensor.h>
#include <libtensor/block_tensor/btod_add.h>
#include <libtensor/block_tensor/btod_contract2.h>
#include <libtensor/block_tensor/btod_symmetrize2.h>
#include <libtensor/block_tensor/btod_symmetrize3.h>
#include <libtensor/block_tensor/btod_contract3.h>
#include <libtensor/block_tensor/btod_symcontract3.h>
#include <libtensor/block_tensor/btod_traits.h>
#include <libtensor/core/orbit.h>
#include <libtensor/core/orbit_list.h>
#include <libtensor/core/batching_policy_base.h>
#include <libtensor/dense_tensor/tod_btconv.h>
#include <libtensor/gen_block_tensor/impl/gen_bto_contract2_block_list.h>
#include <libtensor/gen_block_tensor/impl/gen_bto_contract2_clst_builder.h>
#include <libtensor/gen_block_tensor/impl/gen_bto_unfold_block_list.h>
#include <libtensor/symmetry/point_group_table.h>
#include <libtensor/symmetry/product_table_container.h>
#include <libtensor/symmetry/se_label.h>
#include <libtensor/symmetry/se_part.h>
#include <libtensor/symmetry/se_perm.h>
#include <libutil/thread_pool/thread_pool.h>

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <sys/time.h>
#include <type_traits>
#include <vector>

#include "spec_format.h"
#include "ref_build.h"

using namespace libtensor;
using ref_build::make_perm;
using ref_build::make_index;
using ref_build::make_bis;
using ref_build::fill_tensor;

namespace {

struct options {
    size_t threads, repeat;
    bool symc, clst, result;
    options() : threads(1), repeat(0), symc(true), clst(true), result(true) { }
};

double wall_ms() {
    struct timeval tv;
    gettimeofday(&tv, 0);
    return tv.tv_sec * 1e3 + tv.tv_usec * 1e-3;
}

template<size_t N>
void print_tr(FILE *f, const tensor_transf<N, double> &tr) {
    for(size_t i = 0; i < N; i++) fprintf(f, " %zu", tr.get_perm()[i]);
    fprintf(f, " %.17g", tr.get_scalar_tr().get_coeff());
}

template<size_t NC>
void print_symt(FILE *f, const symmetry<NC, double> &sym) {
    dimensions<NC> bid = sym.get_bis().get_block_index_dims();
    fprintf(f, "symt %zu\n", bid.get_size());
    for(size_t a = 0; a < bid.get_size(); a++) {
        orbit<NC, double> o(sym, a);
        fprintf(f, "%zu %zu %d", a, o.get_acindex(), o.is_allowed() ? 1 : 0);
        print_tr<NC>(f, o.get_transf(a));
        fprintf(f, "\n");
    }
}

template<size_t N>
void write_dense(block_tensor<N, double, allocator<double> > &bt, const std::string &path) {
    dimensions<N> dc = bt.get_bis().get_dims();
    dense_tensor<N, double, allocator<double> > t(dc);
    tod_btconv<N>(bt).perform(t);
    dense_tensor_rd_ctrl<N, double> ctl(t);
    const double *pd = ctl.req_const_dataptr();
    FILE *g = fopen(path.c_str(), "wb");
    fwrite(pd, 8, dc.get_size(), g);
    fclose(g);
    ctl.ret_const_dataptr(pd);
}

// btod_symmetrize3 (libtensor instantiates it for orders 3..8 only).
template<size_t NC>
additive_gen_bto<NC, btod_traits::bti_traits> *make_symm3(
    additive_gen_bto<NC, btod_traits::bti_traits> &op, const spec::problem::symm2 &y,
    std::true_type) {
    return new btod_symmetrize3<NC>(op, make_perm<NC>(y.perm), make_perm<NC>(y.perm2),
        y.symm != 0);
}

template<size_t NC>
additive_gen_bto<NC, btod_traits::bti_traits> *make_symm3(
    additive_gen_bto<NC, btod_traits::bti_traits> &, const spec::problem::symm2 &,
    std::false_type) {
    throw std::runtime_error("symmetrize3 needs a result of order 3 or more");
}

// btod_symmetrize2 chain over the contraction (libtensor instantiates it for
// orders 2..8 only).
template<size_t NC>
int run_symm(const spec::problem &p, additive_gen_bto<NC, btod_traits::bti_traits> &op0,
    FILE *f, const std::string &out, const options &opt, std::true_type) {
    std::vector<additive_gen_bto<NC, btod_traits::bti_traits>*> chain;
    additive_gen_bto<NC, btod_traits::bti_traits> *cur = &op0;
    double t0 = wall_ms();
    for(size_t i = 0; i < p.symm.size(); i++) {
        if(p.symm[i].perm2.empty())
            chain.push_back(new btod_symmetrize2<NC>(*cur, make_perm<NC>(p.symm[i].perm),
                p.symm[i].symm != 0));
        else
            chain.push_back(make_symm3<NC>(*cur, p.symm[i],
                std::integral_constant<bool, (NC >= 3)>()));
        cur = chain.back();
    }
    fprintf(f, "symmetrize_ms %.6f\n", wall_ms() - t0);
    print_symt<NC>(f, cur->get_symmetry());
    const assignment_schedule<NC, double> &sch = cur->get_schedule();
    std::vector<size_t> sched;
    for(typename assignment_schedule<NC, double>::iterator i = sch.begin(); i != sch.end(); ++i)
        sched.push_back(sch.get_abs_index(i));
    std::sort(sched.begin(), sched.end());
    fprintf(f, "schedule %zu\n", sched.size());
    for(size_t i = 0; i < sched.size(); i++) fprintf(f, "%zu\n", sched[i]);
    block_tensor<NC, double, allocator<double> > btc(cur->get_bis());
    t0 = wall_ms();
    chain.back()->perform(btc);
    fprintf(f, "perform_ms %.6f\n", wall_ms() - t0);
    if(opt.result) {
        dimensions<NC> dc = cur->get_bis().get_dims();
        dense_tensor<NC, double, allocator<double> > t(dc);
        tod_btconv<NC>(btc).perform(t);
        dense_tensor_rd_ctrl<NC, double> ctl(t);
        const double *pd = ctl.req_const_dataptr();
        FILE *g = fopen((out + ".result.bin").c_str(), "wb");
        fwrite(pd, 8, dc.get_size(), g);
        fclose(g);
        ctl.ret_const_dataptr(pd);
    }
    for(size_t i = chain.size(); i > 0; i--) delete chain[i - 1];
    return 0;
}

template<size_t NC>
int run_symm(const spec::problem &, additive_gen_bto<NC, btod_traits::bti_traits> &,
    FILE *, const std::string &, const options &, std::false_type) {
    fprintf(stderr, "symmetrize2 needs a result of order 2 or more\n");
    return 3;
}

template<size_t N, size_t M, size_t K>
int run(const spec::problem &p, const std::string &out, const options &opt) {
    enum { NA = N + K, NB = M + K, NC = N + M };

    block_index_space<NA> bisa = make_bis<NA>(p.a);
    block_tensor<NA, double, allocator<double> > bta(bisa);
    fill_tensor<NA>(p.a, bta);
    block_tensor<NB, double, allocator<double> > *btbp = 0;
    if(!p.self) {
        block_index_space<NB> bisb = make_bis<NB>(p.b);
        btbp = new block_tensor<NB, double, allocator<double> >(bisb);
        fill_tensor<NB>(p.b, *btbp);
    }
    block_tensor<NB, double, allocator<double> > &btb = p.self ?
        reinterpret_cast<block_tensor<NB, double, allocator<double> >&>(bta) : *btbp;

    sequence<NC, size_t> pcs(0);
    std::vector<size_t> pcv(p.permc);
    contraction2<N, M, K> contr(make_perm<NC>(pcv));
    for(size_t i = 0; i < K; i++) contr.contract(p.ka[i], p.kb[i]);

    FILE *f = fopen((out + ".txt").c_str(), "w");
    double t0 = wall_ms();
    btod_contract2<N, M, K> op(contr, bta, p.sa, btb, p.sb, p.sc);
    double t_sched = wall_ms() - t0;
    fprintf(f, "schedule_ms %.6f\n", t_sched);
    if(!p.symm.empty()) {
        int rc = run_symm<NC>(p, op, f, out, opt,
            std::integral_constant<bool, (NC >= 2)>());
        fclose(f);
        delete btbp;
        return rc;
    }

    const symmetry<NC, double> &symc = op.get_symmetry();
    dimensions<NC> bidc = symc.get_bis().get_block_index_dims();
    if(opt.symc) {
        fprintf(f, "symc %zu\n", bidc.get_size());
        for(size_t a = 0; a < bidc.get_size(); a++) {
            orbit<NC, double> o(symc, a);
            fprintf(f, "%zu %zu %d\n", a, o.get_acindex(), o.is_allowed() ? 1 : 0);
        }
    }
    const assignment_schedule<NC, double> &sch = op.get_schedule();
    std::vector<size_t> sched;
    for(typename assignment_schedule<NC, double>::iterator i = sch.begin(); i != sch.end(); ++i)
        sched.push_back(sch.get_abs_index(i));
    std::sort(sched.begin(), sched.end());
    fprintf(f, "schedule %zu\n", sched.size());
    for(size_t i = 0; i < sched.size(); i++) fprintf(f, "%zu\n", sched[i]);

    if(opt.clst) {
        block_tensor_ctrl<NA, double> ca(bta);
        block_tensor_ctrl<NB, double> cb(btb);
        const symmetry<NA, double> &syma = ca.req_const_symmetry();
        const symmetry<NB, double> &symb = cb.req_const_symmetry();
        dimensions<NA> bida = bta.get_bis().get_block_index_dims();
        dimensions<NB> bidb = btb.get_bis().get_block_index_dims();
        std::vector<size_t> nza, nzb;
        gen_block_tensor_rd_ctrl<NA, block_tensor_i_traits<double> >(bta).req_nonzero_blocks(nza);
        gen_block_tensor_rd_ctrl<NB, block_tensor_i_traits<double> >(btb).req_nonzero_blocks(nzb);
        block_list<NA> bla(bida, nza), blax(bida);
        block_list<NB> blb(bidb, nzb), blbx(bidb);
        gen_bto_unfold_block_list<NA, btod_traits>(syma, bla).build(blax);
        gen_bto_unfold_block_list<NB, btod_traits>(symb, blb).build(blbx);
        gen_bto_contract2_block_list<N, M, K> cbl(contr, bida, blax, bidb, blbx);
        for(size_t s = 0; s < sched.size(); s++) {
            libtensor::index<NC> ic;
            abs_index<NC>::get_index(sched[s], bidc, ic);
            gen_bto_contract2_clst_builder<N, M, K, btod_traits> bld(contr, syma, symb,
                blax, blbx, bidc, ic);
            bld.build_list(false, cbl);
            const typename gen_bto_contract2_clst<N, M, K, double>::list_type &cl =
                bld.get_clst();
            fprintf(f, "clst %zu %zu\n", sched[s], cl.size());
            for(typename gen_bto_contract2_clst<N, M, K, double>::list_type::const_iterator
                j = cl.begin(); j != cl.end(); ++j) {
                fprintf(f, "%zu %zu", j->get_aindex_a(), j->get_acindex_a());
                print_tr<NA>(f, j->get_transf_a());
                fprintf(f, " %zu %zu", j->get_aindex_b(), j->get_acindex_b());
                print_tr<NB>(f, j->get_transf_b());
                fprintf(f, "\n");
            }
        }
    }

    block_tensor<NC, double, allocator<double> > btc(op.get_bis());
    libutil::thread_pool *tp = 0;
    if(opt.threads > 1) {
        tp = new libutil::thread_pool(opt.threads, opt.threads);
        tp->associate();
    }
    t0 = wall_ms();
    op.perform(btc);
    double t_first = wall_ms() - t0;
    fprintf(f, "perform_ms %.6f\n", t_first);
    std::vector<double> times;
    for(size_t r = 0; r < opt.repeat; r++) {
        t0 = wall_ms();
        op.perform(btc);
        times.push_back(wall_ms() - t0);
    }
    if(!times.empty()) {
        std::sort(times.begin(), times.end());
        fprintf(f, "perform_median_ms %.6f\n", times[times.size() / 2]);
    }
    if(tp) {
        tp->dissociate();
        delete tp;
    }
    fprintf(f, "threads %zu\n", opt.threads);

    if(opt.result) write_dense<NC>(btc, out + ".result.bin");
    if(p.has_add) {
        // perform(btc, d) into tensor C of the spec (its own symmetry and
        // data): gen_bto_aux_add merges the symmetries
        // (gen_bto_aux_add_impl.h:40-66, addition_schedule)
        if(p.c.n != NC) throw std::runtime_error("add: tensor C of order N + M needed");
        block_tensor<NC, double, allocator<double> > btt(make_bis<NC>(p.c));
        fill_tensor<NC>(p.c, btt);
        op.perform(btt, scalar_transf<double>(p.add_d));
        block_tensor_ctrl<NC, double> ct(btt);
        fprintf(f, "add\n");
        print_symt<NC>(f, ct.req_const_symmetry());
        write_dense<NC>(btt, out + ".add.bin");
    }
    fclose(f);
    delete btbp;
    return 0;
}

// btod_add<n>(A, ca) + add_op(B, perm, cb), perform(out): the symmetry of out
// (every orbit: "symt") and its dense result.
template<size_t N>
int run_add(const spec::problem &p, const std::string &out, const options &) {
    block_tensor<N, double, allocator<double> > bta(make_bis<N>(p.a));
    fill_tensor<N>(p.a, bta);
    block_tensor<N, double, allocator<double> > btb(make_bis<N>(p.b));
    fill_tensor<N>(p.b, btb);
    btod_add<N> op(bta, p.add_ca);
    op.add_op(btb, make_perm<N>(p.add_perm), p.add_cb);
    block_tensor<N, double, allocator<double> > bto(op.get_bis());
    op.perform(bto);
    FILE *f = fopen((out + ".txt").c_str(), "w");
    block_tensor_ctrl<N, double> ct(bto);
    print_symt<N>(f, ct.req_const_symmetry());
    const assignment_schedule<N, double> &sch = op.get_schedule();
    std::vector<size_t> sched;
    for(typename assignment_schedule<N, double>::iterator i = sch.begin(); i != sch.end(); ++i)
        sched.push_back(sch.get_abs_index(i));
    std::sort(sched.begin(), sched.end());
    fprintf(f, "schedule %zu\n", sched.size());
    for(size_t i = 0; i < sched.size(); i++) fprintf(f, "%zu\n", sched[i]);
    fclose(f);
    write_dense<N>(bto, out + ".result.bin");
    return 0;
}

// btod_contract3<N1, N2, N3, K1, K2> of A, B and C: the result symmetry with
// every orbit transformation ("symt"), the schedule and the dense result.
template<size_t N1, size_t N2, size_t N3, size_t K1, size_t K2>
int run3(const spec::problem &p, const std::string &out, const options &opt) {
    enum { NA = N1 + K1, NB = N2 + K1 + K2, NC = N3 + K2, NAB = N1 + N2 + K2,
           ND = N1 + N2 + N3 };
    block_tensor<NA, double, allocator<double> > bta(make_bis<NA>(p.a));
    fill_tensor<NA>(p.a, bta);
    block_tensor<NB, double, allocator<double> > btb(make_bis<NB>(p.b));
    fill_tensor<NB>(p.b, btb);
    block_tensor<NC, double, allocator<double> > btc(make_bis<NC>(p.c));
    fill_tensor<NC>(p.c, btc);
    contraction2<N1, N2 + K2, K1> contr1(make_perm<NAB>(p.permab));
    for(size_t i = 0; i < K1; i++) contr1.contract(p.ka1[i], p.kb1[i]);
    contraction2<N1 + N2, N3, K2> contr2(make_perm<ND>(p.permd));
    for(size_t i = 0; i < K2; i++) contr2.contract(p.kab2[i], p.kc2[i]);

    FILE *f = fopen((out + ".txt").c_str(), "w");
    double t0 = wall_ms();
    btod_contract3<N1, N2, N3, K1, K2> op(contr1, contr2, bta, btb, btc, p.kd);
    fprintf(f, "schedule_ms %.6f\n", wall_ms() - t0);
    print_symt<ND>(f, op.get_symmetry());
    const assignment_schedule<ND, double> &sch = op.get_schedule();
    std::vector<size_t> sched;
    for(typename assignment_schedule<ND, double>::iterator i = sch.begin(); i != sch.end(); ++i)
        sched.push_back(sch.get_abs_index(i));
    std::sort(sched.begin(), sched.end());
    fprintf(f, "schedule %zu\n", sched.size());
    for(size_t i = 0; i < sched.size(); i++) fprintf(f, "%zu\n", sched[i]);
    block_tensor<ND, double, allocator<double> > btd(op.get_bis());
    t0 = wall_ms();
    op.perform(btd);
    fprintf(f, "perform_ms %.6f\n", wall_ms() - t0);
    if(opt.result) {
        dimensions<ND> dd = op.get_bis().get_dims();
        dense_tensor<ND, double, allocator<double> > t(dd);
        tod_btconv<ND>(btd).perform(t);
        dense_tensor_rd_ctrl<ND, double> ctl(t);
        const double *pd = ctl.req_const_dataptr();
        FILE *g = fopen((out + ".result.bin").c_str(), "wb");
        fwrite(pd, 8, dd.get_size(), g);
        fclose(g);
        ctl.ret_const_dataptr(pd);
    }
    fclose(f);
    return 0;
}

// btod_symcontract3<N1, N2, N3, K1, K2>: it exposes no symmetry or schedule,
// so the result tensor's symmetry ("symt") and non-zero blocks ("schedule")
// are written after the perform.
template<size_t N1, size_t N2, size_t N3, size_t K1, size_t K2>
int run3s(const spec::problem &p, const std::string &out, const options &opt) {
    enum { NA = N1 + K1, NB = N2 + K1 + K2, NC = N3 + K2, NAB = N1 + N2 + K2,
           ND = N1 + N2 + N3 };
    block_tensor<NA, double, allocator<double> > bta(make_bis<NA>(p.a));
    fill_tensor<NA>(p.a, bta);
    block_tensor<NB, double, allocator<double> > btb(make_bis<NB>(p.b));
    fill_tensor<NB>(p.b, btb);
    block_tensor<NC, double, allocator<double> > btc(make_bis<NC>(p.c));
    fill_tensor<NC>(p.c, btc);
    contraction2<N1, N2 + K2, K1> contr1(make_perm<NAB>(p.permab));
    for(size_t i = 0; i < K1; i++) contr1.contract(p.ka1[i], p.kb1[i]);
    contraction2<N1 + N2, N3, K2> contr2(make_perm<ND>(p.permd));
    for(size_t i = 0; i < K2; i++) contr2.contract(p.kab2[i], p.kc2[i]);
    FILE *f = fopen((out + ".txt").c_str(), "w");
    btod_symcontract3<N1, N2, N3, K1, K2> op(contr1, contr2, bta, btb,
        make_perm<NAB>(p.symab), p.symab_symm != 0, btc, p.kd);
    // the result's block index space is that of the plain contract3
    btod_contract3<N1, N2, N3, K1, K2> op3(contr1, contr2, bta, btb, btc, p.kd);
    block_tensor<ND, double, allocator<double> > btd(op3.get_bis());
    double t0 = wall_ms();
    op.perform(btd);
    fprintf(f, "perform_ms %.6f\n", wall_ms() - t0);
    {
        block_tensor_ctrl<ND, double> cd(btd);
        print_symt<ND>(f, cd.req_const_symmetry());
        std::vector<size_t> nz;
        gen_block_tensor_rd_ctrl<ND, block_tensor_i_traits<double> >(btd).req_nonzero_blocks(nz);
        std::sort(nz.begin(), nz.end());
        fprintf(f, "schedule %zu\n", nz.size());
        for(size_t i = 0; i < nz.size(); i++) fprintf(f, "%zu\n", nz[i]);
    }
    if(opt.result) {
        dimensions<ND> dd = op3.get_bis().get_dims();
        dense_tensor<ND, double, allocator<double> > t(dd);
        tod_btconv<ND>(btd).perform(t);
        dense_tensor_rd_ctrl<ND, double> ctl(t);
        const double *pd = ctl.req_const_dataptr();
        FILE *g = fopen((out + ".result.bin").c_str(), "wb");
        fwrite(pd, 8, dd.get_size(), g);
        fclose(g);
        ctl.ret_const_dataptr(pd);
    }
    fclose(f);
    return 0;
}

typedef int (*runner)(const spec::problem &, const std::string &, const options &);

struct entry { size_t n, m, k; runner r; };

#define E(n, m, k) { n, m, k, &run<n, m, k> }
const entry k_table[] = {
    E(1, 1, 0), E(2, 1, 0), E(1, 2, 0), E(2, 2, 0),
    E(0, 1, 1), E(1, 0, 1), E(1, 1, 1), E(2, 1, 1), E(1, 2, 1), E(2, 2, 1),
    E(3, 1, 1), E(1, 3, 1), E(3, 3, 1),
    E(0, 2, 2), E(2, 0, 2), E(1, 1, 2), E(2, 1, 2), E(1, 2, 2), E(2, 2, 2),
    E(1, 1, 3), E(2, 2, 3), E(1, 3, 3)
};
#undef E

struct entry3 { size_t n1, n2, n3, k1, k2; runner r; };

// The orders libtensor instantiates (block_tensor/impl/btod_contract3.C).
#define E3(a, b, c, d, e) { a, b, c, d, e, &run3<a, b, c, d, e> }
#define E3S(a, b, c, d, e) { a, b, c, d, e, &run3s<a, b, c, d, e> }
const entry3 k_table3[] = {
    E3(1, 0, 1, 1, 1), E3(1, 1, 1, 1, 1), E3(1, 1, 2, 0, 0), E3(1, 1, 2, 1, 2),
    E3(2, 0, 0, 0, 2), E3(2, 0, 1, 1, 2), E3(2, 0, 2, 1, 2)
};

// ... and btod_symcontract3 (block_tensor/impl/btod_symcontract3.C).
const entry3 k_table3s[] = {
    E3S(1, 0, 1, 1, 1), E3S(1, 1, 1, 1, 1), E3S(2, 0, 2, 1, 2)
};
#undef E3
#undef E3S

} // unnamed namespace

int main(int argc, char **argv) {
    if(argc < 3) {
        fprintf(stderr, "usage: %s spec out_prefix [--threads T] [--repeat R] "
            "[--no-symc] [--no-clst] [--no-result]\n", argv[0]);
        return 2;
    }
    options opt;
    for(int i = 3; i < argc; i++) {
        std::string a(argv[i]);
        if(a == "--threads" && i + 1 < argc) opt.threads = atoi(argv[++i]);
        else if(a == "--repeat" && i + 1 < argc) opt.repeat = atoi(argv[++i]);
        else if(a == "--no-symc") opt.symc = false;
        else if(a == "--no-clst") opt.clst = false;
        else if(a == "--no-result") opt.result = false;
    }
    try {
        spec::problem p = spec::parse(argv[1]);
        for(size_t t = 0; t < p.tables.size(); t++) ref_build::add_table(p.tables[t]);
        allocator<double>::init(4, 16, 16777216, 16777216);
        batching_policy_base::set_batch_size(size_t(1) << 40);
        if(p.c3) {
            const entry3 *tb = p.symab.empty() ? k_table3 : k_table3s;
            const size_t nt = p.symab.empty() ? sizeof(k_table3) / sizeof(k_table3[0]) :
                sizeof(k_table3s) / sizeof(k_table3s[0]);
            for(size_t i = 0; i < nt; i++) {
                const entry3 &e = tb[i];
                if(e.n1 == p.N1 && e.n2 == p.N2 && e.n3 == p.N3 && e.k1 == p.K1 &&
                    e.k2 == p.K2) return e.r(p, argv[2], opt);
            }
            fprintf(stderr, "unsupported contract3 orders\n");
            return 3;
        }
        if(p.btod_add_n) {
            switch(p.btod_add_n) {
            case 1: return run_add<1>(p, argv[2], opt);
            case 2: return run_add<2>(p, argv[2], opt);
            case 3: return run_add<3>(p, argv[2], opt);
            case 4: return run_add<4>(p, argv[2], opt);
            default: fprintf(stderr, "btod_add: order 1..4\n"); return 3;
            }
        }
        for(size_t i = 0; i < sizeof(k_table) / sizeof(k_table[0]); i++) {
            const entry &e = k_table[i];
            if(e.n == p.N && e.m == p.M && e.k == p.K) return e.r(p, argv[2], opt);
        }
        fprintf(stderr, "unsupported contraction order (%zu, %zu, %zu)\n", p.N, p.M, p.K);
        return 3;
    } catch(std::exception &e) {
        fprintf(stderr, "error: %s\n", e.what());
        return 1;
    }
}
