// Block schedule of a block-tensor contraction (see lt_sched.h).
#include <functional>
#include "lt_sched.h"

#include <atomic>
#include <thread>

namespace lta {

void parallel_for(size_t n, size_t nthreads, const std::function<void(size_t)> &f) {
    if(nthreads <= 1 || n < 2) {
        for(size_t i = 0; i < n; i++) f(i);
        return;
    }
    if(nthreads > n) nthreads = n;
    std::atomic<size_t> next(0);
    std::vector<std::thread> th;
    std::exception_ptr err;
    std::mutex err_mtx;
    for(size_t t = 0; t < nthreads; t++) {
        th.emplace_back([&]() {
            while(true) {
                size_t i = next.fetch_add(1);
                if(i >= n) break;
                try {
                    f(i);
                } catch(...) {
                    std::lock_guard<std::mutex> lk(err_mtx);
                    if(!err) err = std::current_exception();
                    next = n;
                }
            }
        });
    }
    for(size_t t = 0; t < th.size(); t++) th[t].join();
    if(err) std::rethrow_exception(err);
}

orbit_info orbit_cache::get(size_t aidx) {
    {
        std::lock_guard<std::mutex> lk(m_mtx);
        std::unordered_map<size_t, orbit_info>::const_iterator i = m_map.find(aidx);
        if(i != m_map.end()) return i->second;
    }
    orbit o(m_sym, aidx, false);
    orbit_info inf;
    inf.acidx = o.acidx;
    inf.tr = o.get_transf(aidx);
    std::lock_guard<std::mutex> lk(m_mtx);
    m_map[aidx] = inf;
    return inf;
}

void unfold_block_list(const symmetry &sym, const block_list &bl, block_list &out) {
    for(size_t i = 0; i < bl.blks.size(); i++) {
        orbit o(sym, bl.blks[i], false);
        for(size_t j = 0; j < o.orb.size(); j++) out.add(o.orb[j].first);
    }
    out.sort();
}

// ---------------------------------------------------------------------------

namespace {

struct outer_maps {
    size_t N, M, K;
    size_t mapai[kMaxOrder], mapbj[kMaxOrder], mapak[kMaxOrder], mapbk[kMaxOrder];
    dims_t dimsi, dimsj, dimsk;

    outer_maps(const contraction2 &c, const dims_t &bidimsa, const dims_t &bidimsb) :
        N(c.N), M(c.M), K(c.K), dimsi(c.N), dimsj(c.M), dimsk(c.K) {
        const size_t NC = c.NC(), NA = c.NA(), NB = c.NB();
        for(size_t i = 0, j = 0; i < NA; i++) {
            if(c.conn[NC + i] < NC) { mapai[j] = i; dimsi[j] = bidimsa[i]; j++; }
        }
        for(size_t i = 0, j = 0; i < NB; i++) {
            if(c.conn[NC + NA + i] < NC) { mapbj[j] = i; dimsj[j] = bidimsb[i]; j++; }
        }
        for(size_t i = 0, k = 0; i < NA; i++) {
            if(c.conn[NC + i] >= NC + NA) {
                mapak[k] = i;
                mapbk[k] = c.conn[NC + i] - NC - NA;
                dimsk[k] = bidimsa[i];
                k++;
            }
        }
    }
};

} // unnamed namespace

contract2_block_list::contract2_block_list(const contraction2 &c,
    const block_list &bla, const block_list &blb) {

    if(c.K == 0) return;
    outer_maps om(c, bla.dims, blb.dims);
    bla.sort();
    blb.sort();
    a1.reserve(bla.blks.size());
    b1.reserve(blb.blks.size());
    for(size_t n = 0; n < bla.blks.size(); n++) {
        index_t ba = get_index(bla.blks[n], bla.dims);
        index_t ii(om.N), ik(om.K);
        for(size_t i = 0; i < om.N; i++) ii[i] = ba[om.mapai[i]];
        for(size_t i = 0; i < om.K; i++) ik[i] = ba[om.mapak[i]];
        a1.push_back(std::make_pair(abs_index(ik, om.dimsk), abs_index(ii, om.dimsi)));
    }
    for(size_t n = 0; n < blb.blks.size(); n++) {
        index_t bb = get_index(blb.blks[n], blb.dims);
        index_t ij(om.M), ik(om.K);
        for(size_t i = 0; i < om.M; i++) ij[i] = bb[om.mapbj[i]];
        for(size_t i = 0; i < om.K; i++) ik[i] = bb[om.mapbk[i]];
        b1.push_back(std::make_pair(abs_index(ik, om.dimsk), abs_index(ij, om.dimsj)));
    }
    a2.resize(a1.size());
    b2.resize(b1.size());
    for(size_t i = 0; i < a1.size(); i++) a2[i] = std::make_pair(a1[i].second, a1[i].first);
    for(size_t i = 0; i < b1.size(); i++) b2[i] = std::make_pair(b1[i].second, b1[i].first);
    std::sort(a1.begin(), a1.end());
    std::sort(b1.begin(), b1.end());
    std::sort(a2.begin(), a2.end());
    std::sort(b2.begin(), b2.end());
}

// ---------------------------------------------------------------------------

void coalesce_clst(const contraction2 &c, contr_list &clst) {
    contr_list::iterator j1 = clst.begin();
    while(j1 != clst.end()) {
        contr_list::iterator j2 = j1;
        ++j2;
        bool incj1 = true;
        while(j2 != clst.end()) {
            if(!(j1->acia == j2->acia && j1->acib == j2->acib)) { ++j2; continue; }
            contraction2 c1(c), c2(c);
            c1.permute_a(j1->tra.perm);
            c1.permute_b(j1->trb.perm);
            c2.permute_a(j2->tra.perm);
            c2.permute_b(j2->trb.perm);
            if(!c1.same_conn(c2)) { ++j2; continue; }
            double d1 = j1->tra.st.c * j1->trb.st.c;
            double d2 = j2->tra.st.c * j2->trb.st.c;
            if(d1 + d2 == 0) {
                j1 = clst.erase(j1);
                if(j1 == j2) {
                    j1 = j2 = clst.erase(j2);
                } else {
                    j2 = clst.erase(j2);
                }
                incj1 = false;
                break;
            } else {
                j1->tra.st.reset();
                j1->trb.st.reset();
                j1->tra.st.scale(d1 + d2);
                j2 = clst.erase(j2);
            }
        }
        if(incj1) ++j1;
    }
}

// ---------------------------------------------------------------------------

void build_clst(const contraction2 &c, orbit_cache &oca, orbit_cache &ocb,
    const block_list &blax, const block_list &blbx,
    const contract2_block_list &cbl, const index_t &ic, contr_list &out) {

    const size_t NC = c.NC(), NA = c.NA(), NB = c.NB();
    if(c.K == 0) {
        build_clst_full(c, oca.sym(), ocb.sym(), blax, blbx, dims_t(), ic, false, out);
        return;
    }
    outer_maps om(c, blax.dims, blbx.dims);
    index_t ii(om.N), ij(om.M);
    for(size_t i = 0, j = 0; i < NA; i++) {
        if(c.conn[NC + i] < NC) ii[j++] = ic[c.conn[NC + i]];
    }
    for(size_t i = 0, j = 0; i < NB; i++) {
        if(c.conn[NC + NA + i] < NC) ij[j++] = ic[c.conn[NC + NA + i]];
    }
    size_t aii = abs_index(ii, om.dimsi), aij = abs_index(ij, om.dimsj);

    typedef std::vector<std::pair<size_t, size_t> >::const_iterator it_t;
    std::pair<size_t, size_t> lo(aii, 0), hi(aii + 1, 0);
    it_t ia = std::lower_bound(cbl.a2.begin(), cbl.a2.end(), lo);
    it_t iae = std::lower_bound(ia, cbl.a2.end(), hi);
    std::pair<size_t, size_t> lob(aij, 0), hib(aij + 1, 0);
    it_t ib = std::lower_bound(cbl.b2.begin(), cbl.b2.end(), lob);
    it_t ibe = std::lower_bound(ib, cbl.b2.end(), hib);

    contr_list clst;
    index_t xa(NA), xb(NB);
    while(true) {
        while(ia != iae && ib != ibe && ia->second != ib->second) {
            while(ia != iae && ia->second < ib->second) ++ia;
            while(ia != iae && ib != ibe && ib->second < ia->second) ++ib;
        }
        if(ia == iae || ib == ibe) break;
        index_t ik = get_index(ia->second, om.dimsk);
        for(size_t i = 0; i < om.N; i++) xa[om.mapai[i]] = ii[i];
        for(size_t i = 0; i < om.M; i++) xb[om.mapbj[i]] = ij[i];
        for(size_t i = 0; i < om.K; i++) xa[om.mapak[i]] = xb[om.mapbk[i]] = ik[i];
        size_t aia = abs_index(xa, blax.dims);
        size_t aib = abs_index(xb, blbx.dims);
        orbit_info oa = oca.get(aia), ob = ocb.get(aib);
        contr_pair p;
        p.aia = aia; p.acia = oa.acidx; p.tra = oa.tr;
        p.aib = aib; p.acib = ob.acidx; p.trb = ob.tr;
        clst.push_back(p);
        ++ia;
        ++ib;
    }
    coalesce_clst(c, clst);
    out.splice(out.end(), clst);
}

void build_clst_full(const contraction2 &c, const symmetry &syma,
    const symmetry &symb, const block_list &blax, const block_list &blbx,
    const dims_t &, const index_t &ic, bool testzero, contr_list &out) {

    const size_t N = c.N, M = c.M, K = c.K;
    const size_t NC = c.NC(), NA = c.NA(), NB = c.NB();
    const dims_t &bidimsa = blax.dims, &bidimsb = blbx.dims;

    if(K == 0) {
        index_t ia(NA), ib(NB);
        for(size_t i = 0; i < N; i++) ia[i] = ic[c.conn[NC + i]];
        for(size_t i = 0; i < M; i++) ib[i] = ic[c.conn[NC + NA + i]];
        if(!blax.contains(abs_index(ia, bidimsa)) ||
            !blbx.contains(abs_index(ib, bidimsb))) return;
        orbit oa(syma, abs_index(ia, bidimsa), false);
        orbit ob(symb, abs_index(ib, bidimsb), false);
        contr_list clst;
        for(size_t ja = 0; ja < oa.orb.size(); ja++)
        for(size_t jb = 0; jb < ob.orb.size(); jb++) {
            index_t ia1 = get_index(oa.orb[ja].first, bidimsa);
            index_t ib1 = get_index(ob.orb[jb].first, bidimsb);
            index_t ic1(NC);
            for(size_t i = 0; i < NC; i++) {
                if(c.conn[i] >= 2 * N + M) ic1[i] = ib1[c.conn[i] - 2 * N - M];
                else ic1[i] = ia1[c.conn[i] - N - M];
            }
            if(!(ic1 == ic)) continue;
            contr_pair p;
            p.aia = oa.orb[ja].first; p.acia = oa.acidx; p.tra = oa.tr[oa.orb[ja].second];
            p.aib = ob.orb[jb].first; p.acib = ob.acidx; p.trb = ob.tr[ob.orb[jb].second];
            clst.push_back(p);
        }
        coalesce_clst(c, clst);
        out.splice(out.end(), clst);
        return;
    }

    dims_t bidimsk(K);
    for(size_t i = 0, j = 0; i < NA; i++) {
        if(c.conn[NC + i] > NC) bidimsk[j++] = bidimsa[i];
    }
    size_t nk = bidimsk.size();
    std::vector<char> chk(nk, 1);
    size_t aik = 0;
    while(aik < nk) {
        while(aik < nk && !chk[aik]) aik++;
        if(aik >= nk) break;

        index_t ia(NA), ib(NB);
        index_t ik = get_index(aik, bidimsk);
        size_t ka[kMaxOrder], kb[kMaxOrder];
        for(size_t i = 0, j = 0; i < NA; i++) {
            if(c.conn[NC + i] < NC) {
                ia[i] = ic[c.conn[NC + i]];
            } else {
                ka[j] = i;
                kb[j] = c.conn[NC + i] - 2 * N - M - K;
                ia[ka[j]] = ib[kb[j]] = ik[j];
                j++;
            }
        }
        for(size_t i = 0; i < NB; i++) {
            if(c.conn[2 * N + M + K + i] < N + M) ib[i] = ic[c.conn[2 * N + M + K + i]];
        }
        size_t aia = abs_index(ia, bidimsa), aib = abs_index(ib, bidimsb);
        if(!blax.contains(aia) || !blbx.contains(aib)) {
            chk[aik] = 0;
            continue;
        }
        orbit oa(syma, aia, false);
        orbit ob(symb, aib, false);
        contr_list clst;
        for(size_t ja = 0; ja < oa.orb.size(); ja++)
        for(size_t jb = 0; jb < ob.orb.size(); jb++) {
            index_t ia1 = get_index(oa.orb[ja].first, bidimsa);
            index_t ib1 = get_index(ob.orb[jb].first, bidimsb);
            index_t ika(K), ikb(K), ic1(NC);
            for(size_t i = 0; i < K; i++) { ika[i] = ia1[ka[i]]; ikb[i] = ib1[kb[i]]; }
            if(!(ika == ikb)) continue;
            for(size_t i = 0; i < N + M; i++) {
                if(c.conn[i] >= 2 * N + M + K) ic1[i] = ib1[c.conn[i] - 2 * N - M - K];
                else ic1[i] = ia1[c.conn[i] - N - M];
            }
            if(!(ic1 == ic)) continue;
            contr_pair p;
            p.aia = oa.orb[ja].first; p.acia = oa.acidx; p.tra = oa.tr[oa.orb[ja].second];
            p.aib = ob.orb[jb].first; p.acib = ob.acidx; p.trb = ob.tr[ob.orb[jb].second];
            clst.push_back(p);
            chk[abs_index(ika, bidimsk)] = 0;
        }
        coalesce_clst(c, clst);
        bool empty = clst.empty();
        out.splice(out.end(), clst);
        if(testzero && !empty) break;
    }
}

// ---------------------------------------------------------------------------

std::vector<size_t> contract2_schedule(const contraction2 &c,
    const symmetry &syma, const symmetry &symb, const symmetry &symc,
    const std::vector<size_t> &nza, const std::vector<size_t> &nzb,
    size_t nthreads) {

    const size_t N = c.N, M = c.M, K = c.K;
    const size_t NC = c.NC(), NA = c.NA(), NB = c.NB();
    dims_t bidimsa = syma.bis.block_index_dims();
    dims_t bidimsb = symb.bis.block_index_dims();
    dims_t bidimsc = symc.bis.block_index_dims();

    block_list bla(bidimsa, nza), blb(bidimsb, nzb);
    block_list blax(bidimsa), blbx(bidimsb);
    unfold_block_list(syma, bla, blax);
    unfold_block_list(symb, blb, blbx);

    // permutation taking (A outer ++ B outer) into C order
    size_t seq1[kMaxOrder], seq2[kMaxOrder];
    dims_t dimsci(NC), dimscj(NC);
    for(size_t i = 0, j = 0; i < NA; i++) {
        if(c.conn[NC + i] < NC) {
            dimsci[j] = bidimsa[i];
            seq1[j] = NC + i;
            seq2[j] = c.conn[j];
            j++;
        }
    }
    for(size_t i = 0, j = 0; i < NB; i++) {
        if(c.conn[NC + NA + i] < NC) {
            dimscj[N + j] = bidimsb[i];
            seq1[N + j] = NC + NA + i;
            seq2[N + j] = c.conn[N + j];
            j++;
        }
    }
    perm_t permc = permutation_builder(NC, seq2, seq1);

    std::vector<size_t> candidates;
    if(K == 0) {
        // every pair of unfolded A and B blocks gives a candidate; allowed
        // canonical candidates are non-zero (gen_bto_contract2_nzorb_impl.h,
        // K == 0 task)
        for(size_t a = 0; a < blax.blks.size(); a++) {
            index_t ici = get_index(blax.blks[a], dimsci);
            for(size_t b = 0; b < blbx.blks.size(); b++) {
                index_t icj = get_index(blbx.blks[b], dimscj);
                index_t ic(NC);
                for(size_t i = 0; i < NC; i++) ic[i] = ici[i] + icj[i];
                permc.apply(ic);
                size_t aic = abs_index(ic, bidimsc);
                short_orbit so(symc, aic, true);
                if(so.allowed && so.acidx == aic) candidates.push_back(aic);
            }
        }
        std::sort(candidates.begin(), candidates.end());
        candidates.erase(std::unique(candidates.begin(), candidates.end()),
            candidates.end());
        return candidates;
    }

    contract2_block_list cbl(c, blax, blbx);
    const std::vector<std::pair<size_t, size_t> > &la = cbl.a1, &lb = cbl.b1;

    // contracted block indexes present in both A and B
    std::vector<size_t> ka, kb, ks;
    for(size_t i = 0; i < la.size(); i++)
        if(ka.empty() || ka.back() != la[i].first) ka.push_back(la[i].first);
    for(size_t i = 0; i < lb.size(); i++)
        if(kb.empty() || kb.back() != lb[i].first) kb.push_back(lb[i].first);
    std::set_intersection(ka.begin(), ka.end(), kb.begin(), kb.end(),
        std::back_inserter(ks));

    std::vector<std::vector<size_t> > cand_k(ks.size());
    parallel_for(ks.size(), nthreads, [&](size_t t) {
        size_t k = ks[t];
        std::vector<size_t> &cand = cand_k[t];
        std::pair<size_t, size_t> s(k, 0);
        std::vector<std::pair<size_t, size_t> >::const_iterator ia =
            std::lower_bound(la.begin(), la.end(), s);
        std::vector<std::pair<size_t, size_t> >::const_iterator ib0 =
            std::lower_bound(lb.begin(), lb.end(), s);
        for(; ia != la.end() && ia->first == k; ++ia) {
            index_t ici = get_index(ia->second, dimsci);
            for(std::vector<std::pair<size_t, size_t> >::const_iterator ib = ib0;
                ib != lb.end() && ib->first == k; ++ib) {
                index_t icj = get_index(ib->second, dimscj);
                index_t ic(NC);
                for(size_t i = 0; i < NC; i++) ic[i] = ici[i] + icj[i];
                permc.apply(ic);
                size_t aic = abs_index(ic, bidimsc);
                short_orbit so(symc, aic, true);
                if(so.allowed && so.acidx == aic) cand.push_back(aic);
            }
        }
    });
    for(size_t t = 0; t < ks.size(); t++)
        candidates.insert(candidates.end(), cand_k[t].begin(), cand_k[t].end());
    std::sort(candidates.begin(), candidates.end());
    candidates.erase(std::unique(candidates.begin(), candidates.end()),
        candidates.end());

    orbit_cache oca(syma), ocb(symb);
    std::vector<char> nz(candidates.size(), 0);
    parallel_for(candidates.size(), nthreads, [&](size_t t) {
        index_t ic = get_index(candidates[t], bidimsc);
        contr_list cl;
        build_clst(c, oca, ocb, blax, blbx, cbl, ic, cl);
        nz[t] = cl.empty() ? 0 : 1;
    });
    std::vector<size_t> out;
    for(size_t t = 0; t < candidates.size(); t++) if(nz[t]) out.push_back(candidates[t]);
    (void)M;
    return out;
}

} // namespace lta
