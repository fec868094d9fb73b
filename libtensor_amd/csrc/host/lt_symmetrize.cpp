// btod_symmetrize2 and btod_contract3 on HBM-resident block tensors.
//
// The result symmetry is so_symmetrize of the operand's symmetry
// (gen_block_tensor/impl/gen_bto_symmetrize2_impl.h:131-149) and the schedule
// and per-block source lists follow make_schedule (:152-199): for every
// scheduled block X_ai0 of the operand and every block j of its orbit, the
// block j and its image P(j), where canonical in the new symmetry, receive
// tr_j(X_ai0) and (P, s) tr_j(X_ai0).  perform() computes X once into a
// resident temporary and gathers every block of B from it with the batched
// permute kernel, one launch per source rank (the k-th source of every block
// of B in launch k, so no two descriptors of a launch share a destination).
#include "lt_runtime.h"

#include <algorithm>
#include <list>
#include <map>
#include <set>

namespace lta {

btod_contract3::btod_contract3(const contraction2 &c1, const contraction2 &c2, btensor &a,
    btensor &b, btensor &c, double kd, const perm_t *sympermab, bool symmab) :
    m_a(a), m_last_ms(0.0) {

    m_op1.reset(new btod_contract2(c1, a, 1.0, b, 1.0, 1.0));
    const symmetry *symab = &m_op1->symc();
    const std::vector<size_t> *sab = &m_op1->schedule();
    if(sympermab) {
        m_op1x.reset(new contract2_op(*m_op1, a));
        m_symab.reset(new btod_symmetrize2(*m_op1x, *sympermab, symmab));
        symab = &m_symab->sym();
        sab = &m_symab->schedule();
    }
    // AB carries its symmetry and schedule as its non-zero blocks, so the
    // second contraction's schedule is built over that list as the
    // reference builds it.
    m_ab.reset(new btensor(a.ctx(), symab->bis));
    m_ab->set_symmetry(*symab);
    for(size_t i = 0; i < sab->size(); i++) m_ab->mark_nonzero((*sab)[i]);
    m_op2.reset(new btod_contract2(c2, *m_ab, 1.0, c, 1.0, kd));
}

btod_contract3::~btod_contract3() { }

double btod_contract3::make_ab(int mode) {
    if(m_symab) return m_symab->perform_into(*m_ab, mode);
    m_op1->perform(*m_ab, mode);
    return m_op1->stats().t_gpu_ms;
}

void btod_contract3::perform(btensor &d, int mode) {
    d.require_mutable("btod_contract3::perform");
    const double ms = make_ab(mode);
    m_op2->perform(d, mode);
    m_last_ms = ms + m_op2->stats().t_gpu_ms;
}

void btod_contract3::perform_add(btensor &d, double k, int mode) {
    d.require_mutable("btod_contract3::perform");
    const double ms = make_ab(mode);
    m_op2->perform_add(d, k, mode);
    m_last_ms = ms + m_op2->stats().t_gpu_ms;
}

double btod_contract3::perform_into(btensor &t, int mode) {
    perform(t, mode);
    return m_last_ms;
}

void btod_contract3::compute_block(bool zero, const index_t &id, const tensor_tr &tr,
    double *blk) {
    make_ab(MODE_GROUPED);
    m_op2->compute_block(zero, id, tr, blk);
}

// ---------------------------------------------------------------------------
// Common gather of the symmetrizations.

const std::vector<symmetrizer::rec> *symmetrizer::records(size_t aib) const {
    std::vector<std::pair<size_t, std::vector<rec> > >::const_iterator it =
        std::lower_bound(m_recs.begin(), m_recs.end(), aib,
            [](const std::pair<size_t, std::vector<rec> > &x, size_t v) { return x.first < v; });
    if(it == m_recs.end() || it->first != aib) return 0;
    return &it->second;
}

double symmetrizer::run(btensor &b, bool add, double d, int mode) {
    b.require_mutable("btod_symmetrize::perform");
    context &cx = ctx();
    cx.require_device("btod_symmetrize");
    lt_require(&b.ctx() == &cx, LT_ERR_PARAM, "tensors of different contexts");
    lt_require(b.bis() == m_sym.bis, LT_ERR_DIMS, "btod_symmetrize: block index space of B");
    const dims_t bid = b.bidims();
    if(add) {
        lt_require(same_orbits(m_sym, b.sym(), bid), LT_ERR_SYMMETRY,
            "btod_symmetrize: B += ... needs B with the symmetry of the result");
    }
    if(!m_x) m_x.reset(new btensor(cx, m_sym.bis));
    btensor &x = *m_x;
    double ms = m_op.perform_into(x, mode);
    std::vector<size_t> written;
    for(size_t i = 0; i < m_recs.size(); i++) {
        const std::vector<rec> &rs = m_recs[i].second;
        for(size_t k = 0; k < rs.size(); k++)
            if(x.is_nonzero(rs[k].ai)) { written.push_back(m_recs[i].first); break; }
    }
    if(!add) {
        // A tensor this operation wrote last, with the same blocks, keeps
        // its structure and version (plans reading it stay valid).
        if(b.sym_tag() != m_uid || b.nonzero_blocks() != written) {
            b.zero_all();
            b.set_symmetry(m_sym, m_uid);
        }
    }
    // the gather launches are kept per structure versions of X and B (and
    // the scalar of +=): a repeated perform uploads nothing
    const std::string key = std::string(add ? "symm+" : "symm") + key_tensor(x, true) +
        key_tensor(b, true) + (add ? "|" + key_part(tensor_tr(perm_t(0), scalar_tr(d))) :
        std::string());
    std::shared_ptr<context::cached_launch> l = cx.find_launch(key);
    if(!l) {
        std::vector<std::vector<dev::perm_desc> > rounds;
        std::vector<size_t> touched;
        double bytes = 0.0;
        for(size_t i = 0; i < m_recs.size(); i++) {
            const size_t ib = m_recs[i].first;
            const std::vector<rec> &rs = m_recs[i].second;
            const bool had = add && b.is_nonzero(ib);
            size_t r = 0;
            for(size_t k = 0; k < rs.size(); k++) {
                if(!x.is_nonzero(rs[k].ai)) continue;
                tensor_tr tr(rs[k].tr);
                if(add) tr.st.scale(d);
                if(rounds.size() <= r) rounds.resize(r + 1);
                const dims_t dx = x.bis().block_dims(get_index(rs[k].ai, bid));
                const bool acc = r > 0 || had;
                rounds[r].push_back(make_perm_desc_public(x.block_ptr(rs[k].ai), dx, tr,
                    b.block_ptr(ib), acc));
                bytes += 8.0 * dx.size() * (acc ? 3 : 2);
                r++;
            }
            if(r) touched.push_back(ib);
        }
        l = upload_launches(cx, rounds);
        l->touched = touched;
        l->bytes = bytes;
        cx.keep_launch(key, l);
    }
    double gms = 0.0;
    run_launches(cx, *l, &gms);
    for(size_t i = 0; i < l->touched.size(); i++) b.mark_nonzero(l->touched[i]);
    const double bytes = l->bytes;
    ms += gms;
    m_gather_ms = gms;
    m_gather_bytes = bytes;
    m_last_ms = ms;
    cx.set_last_op_ms(ms);
    return ms;
}

// ---------------------------------------------------------------------------
// btod_symmetrize2

btod_symmetrize2::btod_symmetrize2(additive_op &op, const perm_t &perm, bool symm) :
    symmetrizer(op), m_perm(perm), m_symm(symm) {

    m_sym = so_symmetrize_pair(op.sym(), perm, symm ? 1.0 : -1.0);
    make_schedule();
}

void btod_symmetrize2::make_schedule() {
    // gen_bto_symmetrize2::make_schedule (impl :152-199): for every block j
    // of the orbit of a scheduled X_ai0, the block j and its image P(j),
    // where canonical in the new symmetry, receive tr_j(X_ai0) and
    // (P, s) tr_j(X_ai0).
    const symmetry &s0 = m_op.sym();
    const dims_t bid = m_sym.bis.block_index_dims();
    const double sc = m_symm ? 1.0 : -1.0;
    std::map<size_t, std::vector<rec> > recs;
    std::set<size_t> insch;
    const std::vector<size_t> &sch0 = m_op.schedule();
    for(size_t i = 0; i < sch0.size(); i++) {
        const size_t ai0 = sch0[i];
        orbit o(s0, ai0, true);
        for(size_t j = 0; j < o.orb.size(); j++) {
            const size_t aj1 = o.orb[j].first;
            index_t j2 = get_index(aj1, bid);
            m_perm.apply(j2);
            const size_t aj2 = abs_index(j2, bid);
            short_orbit so1(m_sym, aj1, false), so2(m_sym, aj2, false);
            if(so1.acidx == aj1) {
                if(insch.insert(aj1).second) m_sched.push_back(aj1);
                rec r;
                r.ai = ai0;
                r.tr = o.get_transf(aj1);
                recs[aj1].push_back(r);
            }
            if(so2.acidx == aj2) {
                if(insch.insert(aj2).second) m_sched.push_back(aj2);
                rec r;
                r.ai = ai0;
                r.tr = o.get_transf(aj1);
                r.tr.perm.permute(m_perm);
                r.tr.st.scale(sc);
                recs[aj2].push_back(r);
            }
        }
    }
    m_recs.assign(recs.begin(), recs.end());
}

void btod_symmetrize2::compute_block(bool zero, const index_t &ib, const tensor_tr &trb,
    double *blk) {

    // gen_bto_symmetrize2::compute_block (impl :66-128): sources grouped by
    // operand block; a block used once is computed straight into blk, one
    // used several times once into a temporary that is then added with each
    // record's transformation.
    context &cx = ctx();
    cx.require_device("btod_symmetrize2::compute_block");
    const dims_t bid = m_sym.bis.block_index_dims();
    lt_require(ib.n == bid.n && trb.perm.n == bid.n, LT_ERR_PARAM, "block index order");
    const std::vector<rec> *rp = records(abs_index(ib, bid));
    if(!rp) {
        // No source: the block is zero (the reference leaves blk untouched).
        if(zero) {
            dims_t db = m_sym.bis.block_dims(ib);
            cx.activate();
            hip_check(dev::launch_fill(blk, db.size(), 0.0, cx.stream()), "fill");
            cx.sync();
        }
        return;
    }
    std::list<rec> l(rp->begin(), rp->end());
    bool zero1 = zero;
    while(!l.empty()) {
        const size_t ai = l.front().ai;
        size_t n = 0;
        for(std::list<rec>::const_iterator j = l.begin(); j != l.end(); ++j)
            if(j->ai == ai) n++;
        tensor_tr tri(l.front().tr);
        tri.transform(trb);
        const index_t ia = get_index(ai, bid);
        if(n == 1) {
            m_op.compute_block(zero1, ia, tri, blk);
            zero1 = false;
            l.pop_front();
            continue;
        }
        dims_t dt = m_sym.bis.block_dims(ia);
        dt.permute(tri.perm);
        double *tmp = (double*)cx.alloc(std::max<size_t>(dt.size(), 1) * sizeof(double));
        try {
            m_op.compute_block(true, ia, tri, tmp);
            const tensor_tr tinv(tri.inverse());
            for(std::list<rec>::iterator j = l.begin(); j != l.end();) {
                if(j->ai != ai) { ++j; continue; }
                tensor_tr trj(tinv);
                trj.transform(j->tr).transform(trb);
                dense_permute(cx, tmp, dt, trj, blk, !zero1);
                zero1 = false;
                j = l.erase(j);
            }
        } catch(...) {
            cx.free(tmp);
            throw;
        }
        cx.free(tmp);
    }
}

// ---------------------------------------------------------------------------
// btod_symmetrize3

btod_symmetrize3::btod_symmetrize3(additive_op &op, const perm_t &p1, const perm_t &p2,
    bool symm) : symmetrizer(op), m_p1(p1), m_p2(p2), m_symm(symm) {

    m_sym = so_symmetrize_triple(op.sym(), p1, p2, symm ? 1.0 : -1.0);
    m_sched0 = op.schedule();
    std::sort(m_sched0.begin(), m_sched0.end());
    make_schedule();
}

void btod_symmetrize3::make_schedule() {
    // gen_bto_symmetrize3::make_schedule (impl :162-290): the canonical
    // blocks (new symmetry) of the orbits that the six images of the blocks
    // of every scheduled orbit of X fall in; then the sources of each
    // (make_schedule_blk).
    const symmetry &s0 = m_op.sym();
    const dims_t bid = m_sym.bis.block_index_dims();
    std::set<size_t> sch;
    const std::vector<size_t> &sch0 = m_op.schedule();
    for(size_t i = 0; i < sch0.size(); i++) {
        std::set<size_t> visited;
        orbit o(s0, sch0[i], true);
        for(size_t j = 0; j < o.orb.size(); j++) {
            const index_t i1 = get_index(o.orb[j].first, bid);
            index_t im[6] = { i1, i1, i1, i1, i1, i1 };
            m_p1.apply(im[1]);
            m_p2.apply(im[2]);
            m_p1.apply(im[3]); m_p2.apply(im[3]);
            m_p2.apply(im[4]); m_p1.apply(im[4]);
            m_p1.apply(im[5]); m_p2.apply(im[5]); m_p1.apply(im[5]);
            for(size_t k = 0; k < 6; k++) {
                const size_t a = abs_index(im[k], bid);
                if(visited.count(a)) continue;
                orbit ok(m_sym, a, true);
                sch.insert(ok.acidx);
                for(size_t q = 0; q < ok.orb.size(); q++) visited.insert(ok.orb[q].first);
            }
        }
    }
    m_sched.assign(sch.begin(), sch.end());
    for(size_t i = 0; i < m_sched.size(); i++) {
        std::vector<rec> rs;
        block_records(m_sched[i], rs);
        if(!rs.empty()) m_recs.push_back(std::make_pair(m_sched[i], rs));
    }
}

void btod_symmetrize3::block_records(size_t aib, std::vector<rec> &out) const {
    // gen_bto_symmetrize3::make_schedule_blk (impl :300-380): for each of
    // the six images of the block, the operand's canonical block (when
    // scheduled) with the orbit transformation followed by the exchange back
    // and its sign; sources of one block with one permutation summed.
    const symmetry &s0 = m_op.sym();
    const dims_t bid = m_sym.bis.block_index_dims();
    const double sc = m_symm ? 1.0 : -1.0;
    const index_t i0 = get_index(aib, bid);
    index_t im[6] = { i0, i0, i0, i0, i0, i0 };
    m_p1.apply(im[1]);
    m_p2.apply(im[2]);
    m_p1.apply(im[3]); m_p2.apply(im[3]);
    m_p2.apply(im[4]); m_p1.apply(im[4]);
    m_p1.apply(im[5]); m_p2.apply(im[5]); m_p1.apply(im[5]);
    // the exchanges back, applied to the transformation after the orbit's
    const perm_t *back[6][3] = {
        { 0, 0, 0 }, { &m_p1, 0, 0 }, { &m_p2, 0, 0 }, { &m_p2, &m_p1, 0 },
        { &m_p1, &m_p2, 0 }, { &m_p1, &m_p2, &m_p1 } };
    const bool sign[6] = { false, true, true, false, false, true };
    std::list<rec> l;
    for(size_t k = 0; k < 6; k++) {
        orbit ok(s0, abs_index(im[k], bid), true);
        if(!std::binary_search(m_sched0.begin(), m_sched0.end(), ok.acidx)) continue;
        rec r;
        r.ai = ok.acidx;
        r.tr = ok.get_transf(abs_index(im[k], bid));
        for(size_t q = 0; q < 3; q++) if(back[k][q]) r.tr.perm.permute(*back[k][q]);
        if(sign[k]) r.tr.st.scale(sc);
        l.push_back(r);
    }
    while(!l.empty()) {
        std::list<rec>::iterator i = l.begin();
        const size_t ai = i->ai;
        const perm_t p0 = i->tr.perm;
        double c = 0.0;
        while(i != l.end()) {
            if(i->ai == ai && i->tr.perm == p0) {
                c += i->tr.st.c;
                i = l.erase(i);
            } else {
                ++i;
            }
        }
        if(c != 0.0) {
            rec r;
            r.ai = ai;
            r.tr = tensor_tr(p0, scalar_tr(c));
            out.push_back(r);
        }
    }
}

void btod_symmetrize3::compute_block(bool zero, const index_t &ib, const tensor_tr &trb,
    double *blk) {

    // gen_bto_symmetrize3::compute_block (impl :66-128): as for
    // btod_symmetrize2, with the shared operand block computed untransformed;
    // a block without sources is set to zero.
    context &cx = ctx();
    cx.require_device("btod_symmetrize3::compute_block");
    const dims_t bid = m_sym.bis.block_index_dims();
    lt_require(ib.n == bid.n && trb.perm.n == bid.n, LT_ERR_PARAM, "block index order");
    std::vector<rec> rs;
    block_records(abs_index(ib, bid), rs);
    std::list<rec> l(rs.begin(), rs.end());
    bool zero1 = zero;
    while(!l.empty()) {
        const size_t ai = l.front().ai;
        size_t n = 0;
        for(std::list<rec>::const_iterator j = l.begin(); j != l.end(); ++j)
            if(j->ai == ai) n++;
        const index_t ia = get_index(ai, bid);
        if(n == 1) {
            tensor_tr tri(l.front().tr);
            tri.transform(trb);
            m_op.compute_block(zero1, ia, tri, blk);
            zero1 = false;
            l.pop_front();
            continue;
        }
        const dims_t dt = m_sym.bis.block_dims(ia);
        double *tmp = (double*)cx.alloc(std::max<size_t>(dt.size(), 1) * sizeof(double));
        try {
            m_op.compute_block(true, ia, tensor_tr(bid.n), tmp);
            for(std::list<rec>::iterator j = l.begin(); j != l.end();) {
                if(j->ai != ai) { ++j; continue; }
                tensor_tr trj(j->tr);
                trj.transform(trb);
                dense_permute(cx, tmp, dt, trj, blk, !zero1);
                zero1 = false;
                j = l.erase(j);
            }
        } catch(...) {
            cx.free(tmp);
            throw;
        }
        cx.free(tmp);
    }
    if(zero1 && zero) {
        dims_t db = m_sym.bis.block_dims(ib);
        cx.activate();
        hip_check(dev::launch_fill(blk, db.size(), 0.0, cx.stream()), "fill");
        cx.sync();
    }
}

} // namespace lta
