// Symmetry elements, orbits and symmetry operations of the contraction path.
// See lt_sym.h for the libtensor files each part follows.
#include "lt_sym.h"

#include <algorithm>
#include <list>
#include <set>

namespace lta {

// ---------------------------------------------------------------------------
// Product tables

product_table::product_table(const std::string &id_, size_t n) :
    id(id_), nlabels(n), table(n * n, 0) {
    lt_require(n > 0 && n <= 64, LT_ERR_PARAM, "product_table: 1..64 labels");
    for(size_t i = 0; i < n; i++) {
        table[0 * n + i] = uint64_t(1) << i;
        table[i * n + 0] = uint64_t(1) << i;
    }
}

void product_table::add_product(size_t l1, size_t l2, size_t lr) {
    lt_require(l1 < nlabels && l2 < nlabels && lr < nlabels, LT_ERR_PARAM,
        "product_table: invalid label");
    if(l1 > l2) std::swap(l1, l2);
    if(l1 == 0) {
        lt_require(l2 == lr, LT_ERR_PARAM, "product_table: identity product");
        return;
    }
    table[l1 * nlabels + l2] |= uint64_t(1) << lr;
    table[l2 * nlabels + l1] |= uint64_t(1) << lr;
}

bool product_table::is_in_product(const size_t *lg, size_t nlg, size_t l) const {
    if(nlg == 0) return false;
    uint64_t pr1 = uint64_t(1) << lg[0];
    for(size_t k = 1; k < nlg; k++) {
        uint64_t pr2 = 0;
        for(size_t i = 0; i < nlabels; i++) {
            if(pr1 & (uint64_t(1) << i)) pr2 |= table[i * nlabels + lg[k]];
        }
        pr1 = pr2;
    }
    return (pr1 >> l) & 1;
}

product_table_registry &product_table_registry::instance() {
    static product_table_registry r;
    return r;
}

void product_table_registry::add(const product_table &t) {
    std::lock_guard<std::mutex> lk(m_mtx);
    m_tables[t.id] = std::make_shared<product_table>(t);
}

const product_table &product_table_registry::get(const std::string &id) const {
    std::lock_guard<std::mutex> lk(m_mtx);
    std::map<std::string, std::shared_ptr<product_table> >::const_iterator i =
        m_tables.find(id);
    lt_require(i != m_tables.end(), LT_ERR_PARAM, "unknown product table id");
    return *i->second;
}

bool product_table_registry::has(const std::string &id) const {
    std::lock_guard<std::mutex> lk(m_mtx);
    return m_tables.count(id) != 0;
}

// ---------------------------------------------------------------------------
// se_perm  (libtensor/symmetry/inst/se_perm_impl.h)

se_perm::se_perm(const perm_t &p, const scalar_tr &s) :
    tr(p, s), orderp(1), orderc(1) {

    lt_require(!(p.is_identity() && !s.is_identity()), LT_ERR_SYMMETRY,
        "se_perm: identity permutation with a non-identity scalar");
    perm_t q(p);
    while(!q.is_identity()) { q.permute(p); orderp++; }
    scalar_tr x(s);
    while(!x.is_identity() && orderc < orderp) { x.transform(s); orderc++; }
    lt_require(x.is_identity() && (orderp % orderc) == 0, LT_ERR_SYMMETRY,
        "se_perm: permutation and scalar transformation do not agree");
}

// ---------------------------------------------------------------------------
// se_part  (libtensor/symmetry/inst/se_part_impl.h)

se_part::se_part(const bis_t &b, const dims_t &pd) :
    bis(b), bidims(b.block_index_dims()), pdims(pd), bipdims(pd.n) {

    lt_require(pd.n == b.order(), LT_ERR_SYMMETRY, "se_part: order mismatch");
    for(size_t i = 0; i < pd.n; i++) {
        lt_require(pd[i] >= 1 && bidims[i] % pd[i] == 0, LT_ERR_SYMMETRY,
            "se_part: number of blocks not divisible by number of partitions");
        bipdims[i] = bidims[i] / pd[i];
    }
    init_maps();
}

se_part::se_part(const bis_t &b, const bool *msk, size_t npart) :
    bis(b), bidims(b.block_index_dims()), pdims(b.order()), bipdims(b.order()) {

    lt_require(npart >= 2, LT_ERR_SYMMETRY, "se_part: npart < 2");
    size_t m = 0;
    for(size_t i = 0; i < pdims.n; i++) {
        pdims[i] = msk[i] ? npart : 1;
        if(msk[i]) m++;
    }
    lt_require(m > 0, LT_ERR_SYMMETRY, "se_part: empty mask");
    for(size_t i = 0; i < pdims.n; i++) {
        lt_require(bidims[i] % pdims[i] == 0, LT_ERR_SYMMETRY,
            "se_part: number of blocks not divisible by number of partitions");
        bipdims[i] = bidims[i] / pdims[i];
    }
    init_maps();
}

void se_part::init_maps() {
    size_t sz = pdims.size();
    fmap.assign(sz, 0);
    rmap.assign(sz, 0);
    fmapi.assign(sz, index_t(pdims.n));
    ftr.assign(sz, scalar_tr());
    for(size_t i = 0; i < sz; i++) {
        fmap[i] = rmap[i] = i;
        fmapi[i] = get_index(i, pdims);
    }
}

void se_part::add_map(const index_t &idx1, const index_t &idx2,
    const scalar_tr &tr) {

    size_t a = abs_index(idx1, pdims);
    size_t b = abs_index(idx2, pdims);
    if(a == b) return;
    bool swapped = false;
    if(a > b) { std::swap(a, b); swapped = true; }
    const index_t &idxa = swapped ? idx2 : idx1;
    const index_t &idxb = swapped ? idx1 : idx2;

    // A forbidden partition becomes allowed (a one-loop)
    if(fmap[a] == size_t(-1)) {
        fmap[a] = a; fmapi[a] = idxa; rmap[a] = a; ftr[a].reset();
    }
    if(fmap[b] == size_t(-1)) {
        fmap[b] = b; fmapi[b] = idxb; rmap[b] = b; ftr[b].reset();
    }

    // Is b already in the loop of a?
    size_t ax = a, axf = fmap[ax];
    scalar_tr sx;
    while(ax < axf && ax < b) {
        sx.transform(ftr[ax]);
        ax = axf;
        axf = fmap[ax];
    }
    if(ax == b) {
        if(swapped) sx.invert();
        lt_require(sx == tr, LT_ERR_PARAM,
            "se_part: mapping exists with a different sign");
        return;
    }

    size_t br = rmap[b], bf = fmap[b];
    scalar_tr sab(tr);
    if(swapped) sab.invert();
    while(b != bf) {
        // take b out of its loop
        sx = ftr[b];
        fmap[br] = bf;
        fmapi[br] = get_index(bf, pdims);
        rmap[bf] = br;
        ftr[br].transform(sx);
        // and add it to the loop of a
        add_to_loop(a, b, sab);
        a = b; b = bf; bf = fmap[b];
        sab = sx;
    }
    add_to_loop(a, b, sab);
}

void se_part::add_to_loop(size_t a, size_t b, const scalar_tr &tr) {
    size_t af = fmap[a];
    scalar_tr tx(tr);
    tx.invert();
    if(a < b) {
        while(af < b && a < af) {
            tx.transform(ftr[a]);
            a = af; af = fmap[a];
        }
    } else {
        while((af < b && b < a) || (a < af && af < b) || (b < a && a < af)) {
            tx.transform(ftr[a]);
            a = af; af = fmap[a];
        }
    }
    tx.transform(ftr[a]);
    fmap[a] = b; rmap[b] = a;
    fmapi[a] = get_index(b, pdims);
    fmap[b] = af; rmap[af] = b;
    fmapi[b] = get_index(af, pdims);
    ftr[b] = tx;
    tx.invert();
    ftr[a].transform(tx);
}

void se_part::mark_forbidden(const index_t &idx) {
    size_t a = abs_index(idx, pdims);
    if(fmap[a] == size_t(-1)) return;
    index_t i0(pdims.n);
    size_t af = fmap[a];
    while(af != a) {
        size_t ax = af;
        af = fmap[ax];
        fmap[ax] = size_t(-1); fmapi[ax] = i0; rmap[ax] = size_t(-1);
        ftr[ax].reset();
    }
    fmap[a] = size_t(-1); fmapi[a] = i0; rmap[a] = size_t(-1); ftr[a].reset();
}

const index_t &se_part::get_direct_map(const index_t &idx) const {
    size_t a = abs_index(idx, pdims);
    lt_require(fmap[a] != size_t(-1), LT_ERR_PARAM, "se_part: partition not allowed");
    return fmapi[a];
}

scalar_tr se_part::get_transf(const index_t &from, const index_t &to) const {
    size_t a = abs_index(from, pdims);
    size_t b = abs_index(to, pdims);
    if(a == b) return scalar_tr();
    bool swapped = false;
    if(a > b) { swapped = true; std::swap(a, b); }
    size_t x = fmap[a];
    scalar_tr tr(ftr[a]);
    while(x != b && a < x) {
        tr.transform(ftr[x]);
        x = fmap[x];
    }
    lt_require(x > a, LT_ERR_SYMMETRY, "se_part: no mapping");
    if(swapped) tr.invert();
    return tr;
}

bool se_part::map_exists(const index_t &from, const index_t &to) const {
    size_t a = abs_index(from, pdims);
    size_t b = abs_index(to, pdims);
    if(a > b) std::swap(a, b);
    if(fmap[a] == size_t(-1)) return false;
    if(fmap[b] == size_t(-1)) return false;
    size_t x = fmap[a];
    while(x != b && a < x) x = fmap[x];
    return x == b;
}

bool se_part::is_allowed(const index_t &idx) const {
    index_t pidx(pdims.n);
    for(size_t i = 0; i < pdims.n; i++) pidx[i] = idx[i] / bipdims[i];
    return !is_forbidden(pidx);
}

void se_part::apply(index_t &idx) const {
    index_t p1(pdims.n);
    for(size_t i = 0; i < pdims.n; i++) p1[i] = idx[i] / bipdims[i];
    size_t ap = abs_index(p1, pdims);
    if(fmap[ap] == size_t(-1)) return;
    const index_t &p2 = fmapi[ap];
    for(size_t i = 0; i < pdims.n; i++) idx[i] -= (p1[i] - p2[i]) * bipdims[i];
}

void se_part::apply(index_t &idx, tensor_tr &t) const {
    index_t p1(pdims.n);
    for(size_t i = 0; i < pdims.n; i++) p1[i] = idx[i] / bipdims[i];
    size_t ap = abs_index(p1, pdims);
    if(fmap[ap] == size_t(-1)) return;
    const index_t &p2 = fmapi[ap];
    for(size_t i = 0; i < pdims.n; i++) idx[i] -= (p1[i] - p2[i]) * bipdims[i];
    t.st.transform(ftr[ap]);
}

void se_part::permute(const perm_t &perm) {
    if(perm.is_identity()) return;
    bis.permute(perm);
    bidims.permute(perm);
    bipdims.permute(perm);
    bool affects = false;
    for(size_t i = 0; i < pdims.n; i++) {
        if(pdims[i] != 1 && perm[i] != i) { affects = true; break; }
    }
    if(!affects) return;

    dims_t opd(pdims);
    pdims.permute(perm);
    size_t sz = pdims.size();
    std::vector<size_t> ofmap(fmap);
    std::vector<scalar_tr> oftr(ftr);
    init_maps();
    for(size_t i = 0; i < sz; i++) {
        if(ofmap[i] <= i) continue;
        index_t ia = get_index(i, opd);
        perm.apply(ia);
        size_t naia = abs_index(ia, pdims);
        if(ofmap[i] == size_t(-1)) {
            fmap[naia] = rmap[naia] = size_t(-1);
            continue;
        }
        index_t ib = get_index(ofmap[i], opd);
        perm.apply(ib);
        add_map(ia, ib, oftr[i]);
    }
}

// ---------------------------------------------------------------------------
// se_label  (libtensor/symmetry/inst/se_label_impl.h)

se_label::se_label(const dims_t &bid, const std::string &tid) :
    table_id(tid), bidims(bid), labels(bid.n) {
    lt_require(product_table_registry::instance().has(tid), LT_ERR_PARAM,
        "se_label: unknown product table");
    for(size_t i = 0; i < bid.n; i++) labels[i].assign(bid[i], k_label_invalid);
}

se_label::se_label(const se_label &o) :
    sym_element(o), table_id(o.table_id), bidims(o.bidims), labels(o.labels),
    rule(o.rule), src(o.src), keep(o.keep), rdims(o.rdims), rlabels(o.rlabels),
    alts(o.alts), factors(o.factors), any_of(o.any_of) {
}

void se_label::set_label(size_t dim, size_t blk, size_t label) {
    lt_require(dim < bidims.n && blk < bidims[dim], LT_ERR_PARAM,
        "se_label: block out of range");
    const product_table &pt = product_table_registry::instance().get(table_id);
    lt_require(label == k_label_invalid || label < pt.nlabels, LT_ERR_PARAM,
        "se_label: invalid label");
    labels[dim][blk] = label;
}

void se_label::set_rule(size_t intr) {
    std::vector<size_t> v(1, intr);
    set_rule_set(v);
}

void se_label::set_rule_set(const std::vector<size_t> &intr) {
    rule.clear();
    std::vector<size_t> s(intr);
    std::sort(s.begin(), s.end());
    s.erase(std::unique(s.begin(), s.end()), s.end());
    for(size_t i = 0; i < s.size(); i++) {
        label_term t;
        t.seq = seq_t(bidims.n, 1);
        t.intrinsic = s[i];
        rule.push_back(label_product(1, t));
    }
}

bool se_label::eval_rule(const size_t *lab) const {
    const product_table &pt = product_table_registry::instance().get(table_id);
    size_t lg[64];
    for(size_t ip = 0; ip < rule.size(); ip++) {
        const label_product &pr = rule[ip];
        if(pr.empty()) return false;
        size_t it = 0;
        for(; it < pr.size(); it++) {
            const label_term &t = pr[it];
            if(t.intrinsic == k_label_invalid) continue;
            size_t nlg = 0;
            size_t i = 0;
            for(; i < t.seq.n; i++) {
                if(t.seq[i] == 0) continue;
                size_t l = lab[i];
                if(l == k_label_invalid) break;
                for(size_t r = 0; r < t.seq[i] && nlg < 64; r++) lg[nlg++] = l;
            }
            if(i != t.seq.n) continue;
            if(!pt.is_in_product(lg, nlg, t.intrinsic)) break;
        }
        if(it == pr.size()) return true;
    }
    return false;
}

bool se_label::eval_one(const size_t *lab) const {
    if(!any_of.empty()) {
        for(size_t k = 0; k < any_of.size(); k++) {
            bool all = true;
            for(size_t e = 0; e < any_of[k].size() && all; e++)
                all = any_of[k][e]->eval_labels(lab);
            if(all) return true;
        }
        return false;
    }
    if(!factors.empty()) {
        for(size_t f = 0; f < factors.size(); f++) {
            size_t sub[kMaxOrder];
            const std::vector<size_t> &m = factors[f].map;
            for(size_t k = 0; k < m.size(); k++) sub[k] = lab[m[k]];
            if(!factors[f].el->eval_labels(sub)) return false;
        }
        return true;
    }
    if(src) return eval_reduced(lab);
    return eval_rule(lab);
}

bool se_label::eval_labels(const size_t *lab) const {
    if(eval_one(lab)) return true;
    // Permuted dimensions carry the same labels, so permuting the index
    // permutes the label sequence.
    for(size_t k = 0; k < alts.size(); k++) {
        size_t l2[kMaxOrder];
        for(size_t i = 0; i < bidims.n; i++) l2[i] = lab[alts[k].p[i]];
        if(eval_one(l2)) return true;
    }
    return false;
}

bool se_label::eval_reduced(const size_t *own) const {
    uint64_t key = 0;
    for(size_t i = 0; i < bidims.n; i++) {
        uint64_t l = (own[i] == k_label_invalid) ? 255 : own[i];
        key = (key << 8) | l;
    }
    {
        std::lock_guard<std::mutex> lk(m_memo_mtx);
        std::unordered_map<uint64_t, bool>::const_iterator i = m_memo.find(key);
        if(i != m_memo.end()) return i->second;
    }
    size_t full[4 * kMaxOrder];
    size_t nfull = src->bidims.n;
    for(size_t i = 0; i < nfull; i++) full[i] = k_label_invalid;
    for(size_t i = 0; i < keep.size(); i++) full[keep[i]] = own[i];
    size_t ns = rdims.size();
    std::vector<size_t> pos(ns, 0);
    bool ok = false;
    bool any = true;
    for(size_t s = 0; s < ns; s++) if(rlabels[s].empty()) any = false;
    while(any && !ok) {
        for(size_t s = 0; s < ns; s++)
            for(size_t d = 0; d < rdims[s].size(); d++)
                full[rdims[s][d]] = rlabels[s][pos[s]];
        ok = src->eval_labels(full);
        size_t s = ns;
        while(s > 0) {
            s--;
            if(++pos[s] < rlabels[s].size()) break;
            pos[s] = 0;
            if(s == 0) { any = false; break; }
        }
        if(ns == 0) break;
    }
    std::lock_guard<std::mutex> lk(m_memo_mtx);
    m_memo[key] = ok;
    return ok;
}

bool se_label::is_allowed(const index_t &idx) const {
    size_t lab[kMaxOrder];
    for(size_t i = 0; i < bidims.n; i++) lab[i] = labels[i][idx[i]];
    return eval_labels(lab);
}

void se_label::permute(const perm_t &p) {
    if(p.is_identity()) return;
    bidims.permute(p);
    std::vector<std::vector<size_t> > l(labels);
    for(size_t i = 0; i < bidims.n; i++) labels[i] = l[p.p[i]];
    if(src) {
        std::vector<size_t> k(keep);
        for(size_t i = 0; i < bidims.n; i++) keep[i] = k[p.p[i]];
    } else {
        for(size_t ip = 0; ip < rule.size(); ip++)
            for(size_t it = 0; it < rule[ip].size(); it++)
                p.apply(rule[ip][it].seq);
    }
    for(size_t k = 0; k < any_of.size(); k++)
        for(size_t e = 0; e < any_of[k].size(); e++) {
            std::shared_ptr<se_label> x(new se_label(*any_of[k][e]));
            x->permute(p);
            any_of[k][e] = x;
        }
    // new index y = p(x): alternative q becomes p^-1, then q, then p
    const perm_t pinv(p.inverse());
    for(size_t f = 0; f < factors.size(); f++)
        for(size_t k = 0; k < factors[f].map.size(); k++)
            factors[f].map[k] = pinv[factors[f].map[k]];
    for(size_t k = 0; k < alts.size(); k++) {
        perm_t q(pinv);
        q.permute(alts[k]).permute(p);
        alts[k] = q;
    }
    std::lock_guard<std::mutex> lk(m_memo_mtx);
    m_memo.clear();
}

// ---------------------------------------------------------------------------
// symmetry container

void symmetry::insert(const sym_element &e) {
    insert_shared(std::shared_ptr<sym_element>(e.clone()));
}

void symmetry::insert_shared(std::shared_ptr<sym_element> e) {
    std::string t(e->type());
    for(size_t i = 0; i < subsets.size(); i++) {
        if(subsets[i].first == t) {
            subsets[i].second.push_back(e);
            return;
        }
    }
    subsets.push_back(std::make_pair(t,
        std::vector<std::shared_ptr<sym_element> >(1, e)));
}

// ---------------------------------------------------------------------------
// Orbits  (libtensor/core/impl/orbit_impl.h:209-318)

namespace {
struct pair_less {
    bool operator()(const std::pair<size_t, size_t> &a,
        const std::pair<size_t, size_t> &b) const { return a.first < b.first; }
};
struct pair_eq {
    bool operator()(const std::pair<size_t, size_t> &a,
        const std::pair<size_t, size_t> &b) const { return a.first == b.first; }
};
}

orbit::orbit(const symmetry &sym, size_t aidx, bool compute_allowed) :
    dims(sym.bis.block_index_dims()), acidx(0), allowed(true) {

    const size_t n = dims.n;
    typedef std::pair<size_t, size_t> ip_t;
    pair_less comp;
    pair_eq compeq;

    std::vector<index_t> qi, ti;
    std::vector<tensor_tr> qt, tt;
    std::vector<ip_t> can1, can2;
    qi.reserve(32); qt.reserve(32); ti.reserve(32); tt.reserve(32);

    bool eval = compute_allowed;
    tensor_tr tr1(n);
    tr.reserve(32);
    orb.reserve(32);
    tr.push_back(tr1);
    orb.push_back(ip_t(aidx, 0));
    qi.push_back(get_index(aidx, dims));
    qt.push_back(tr1);

    std::vector<const sym_element*> elems;
    sym.for_each([&](const sym_element &e) { elems.push_back(&e); });

    while(!qi.empty()) {
        index_t idx(qi.back());
        tensor_tr t(qt.back());
        qi.pop_back();
        qt.pop_back();

        for(size_t ie = 0; ie < elems.size(); ie++) {
            const sym_element &e = *elems[ie];
            if(eval && allowed) allowed = e.is_allowed(idx);
            ti.push_back(idx);
            tt.push_back(t);
            e.apply(ti.back(), tt.back());
        }

        for(size_t i = 0; i < ti.size(); i++) {
            size_t a1 = abs_index(ti[i], dims);
            ip_t p(a1, i);
            if(!std::binary_search(orb.begin(), orb.end(), p, comp)) can1.push_back(p);
        }
        std::sort(can1.begin(), can1.end(), comp);
        can1.resize(std::unique(can1.begin(), can1.end(), compeq) - can1.begin());
        for(size_t i = 0; i < can1.size(); i++) {
            size_t a1 = can1[i].first;
            size_t j = tr.size();
            tr.push_back(tt[can1[i].second]);
            can2.push_back(ip_t(a1, j));
            qi.push_back(get_index(a1, dims));
            qt.push_back(tr[j]);
        }
        ti.clear();
        tt.clear();
        can1.resize(orb.size() + can2.size());
        std::merge(orb.begin(), orb.end(), can2.begin(), can2.end(),
            can1.begin(), comp);
        orb.swap(can1);
        can1.clear();
        can2.clear();
    }

    // T(0->i) = [T(s->0) T(s->i)^-1]^-1
    tensor_tr tr0(tr[orb[0].second]);
    for(size_t i = 0; i < orb.size(); i++) {
        tensor_tr &x = tr[orb[i].second];
        x.invert();
        x.transform(tr0);
        x.invert();
    }
    acidx = orb[0].first;
    if(!compute_allowed) allowed = true;
}

const tensor_tr &orbit::get_transf(size_t aidx) const {
    std::pair<size_t, size_t> p(aidx, 0);
    std::vector<std::pair<size_t, size_t> >::const_iterator i =
        std::lower_bound(orb.begin(), orb.end(), p, pair_less());
    lt_require(i != orb.end() && i->first == aidx, LT_ERR_PARAM,
        "orbit: block not in orbit");
    return tr[i->second];
}

bool orbit::contains(size_t aidx) const {
    std::pair<size_t, size_t> p(aidx, 0);
    return std::binary_search(orb.begin(), orb.end(), p, pair_less());
}

short_orbit::short_orbit(const symmetry &sym, size_t aidx, bool compute_allowed) :
    acidx(aidx), allowed(true) {

    dims_t dims(sym.bis.block_index_dims());
    std::vector<size_t> v1, v2, v3, q;
    v1.reserve(32); v2.reserve(32); v3.reserve(32); q.reserve(32);
    q.push_back(aidx);
    v1.push_back(aidx);
    bool eval = compute_allowed;

    std::vector<const sym_element*> elems;
    sym.for_each([&](const sym_element &e) { elems.push_back(&e); });

    while(!q.empty()) {
        v2.clear();
        index_t idx = get_index(q.back(), dims);
        q.pop_back();
        for(size_t ie = 0; ie < elems.size(); ie++) {
            const sym_element &e = *elems[ie];
            if(eval && allowed) allowed = e.is_allowed(idx);
            index_t idx2(idx);
            e.apply(idx2);
            size_t a2 = abs_index(idx2, dims);
            if(!std::binary_search(v1.begin(), v1.end(), a2)) v2.push_back(a2);
        }
        std::sort(v2.begin(), v2.end());
        v2.resize(std::unique(v2.begin(), v2.end()) - v2.begin());
        if(!v2.empty()) {
            q.insert(q.end(), v2.begin(), v2.end());
            v3.resize(v1.size() + v2.size());
            std::merge(v1.begin(), v1.end(), v2.begin(), v2.end(), v3.begin());
            std::swap(v1, v3);
        }
    }
    acidx = v1[0];
    if(!compute_allowed) allowed = true;
}

// ---------------------------------------------------------------------------
// Permutation group  (libtensor/symmetry/inst/permutation_group_impl.h)

permutation_group::branching::branching(size_t n_) : n(n_) {
    for(size_t i = 0; i < kMaxOrder; i++) {
        edges[i] = n;
        sigma[i] = gen_t(perm_t(n), scalar_tr());
        tau[i] = gen_t(perm_t(n), scalar_tr());
    }
}

void permutation_group::branching::reset() {
    for(size_t i = 0; i < kMaxOrder; i++) {
        edges[i] = n;
        sigma[i].first.reset(); tau[i].first.reset();
        sigma[i].second.reset(); tau[i].second.reset();
    }
}

permutation_group::permutation_group(size_t n) : m_n(n), m_br(n) { }

permutation_group::permutation_group(size_t n,
    const std::vector<std::shared_ptr<sym_element> > *set) : m_n(n), m_br(n) {

    std::vector<gen_t> gs1, gs2;
    if(set) {
        for(size_t i = 0; i < set->size(); i++) {
            const se_perm &e = dynamic_cast<const se_perm&>(*(*set)[i]);
            gs1.push_back(gen_t(e.tr.perm, e.tr.st));
        }
    }
    std::vector<gen_t> *p1 = &gs1, *p2 = &gs2;
    for(size_t i = 0; i < m_n; i++) {
        make_branching(m_br, i, *p1, *p2);
        std::swap(p1, p2);
        p2->clear();
    }
}

void permutation_group::add_orbit(const scalar_tr &tr, const perm_t &p) {
    if(p.is_identity()) {
        if(tr.is_identity()) return;
        throw lt_error(LT_ERR_SYMMETRY, "permutation_group: identity with scalar");
    }
    scalar_tr trx(tr);
    if(is_member(m_br, 0, trx, p)) {
        if(trx.is_identity()) return;
        throw lt_error(LT_ERR_SYMMETRY, "permutation_group: inconsistent scalar");
    }
    std::vector<gen_t> gs1, gs2;
    make_genset(m_br, gs1);
    gs1.push_back(gen_t(p, tr));
    m_br.reset();
    std::vector<gen_t> *p1 = &gs1, *p2 = &gs2;
    for(size_t i = 0; i < m_n; i++) {
        make_branching(m_br, i, *p1, *p2);
        std::swap(p1, p2);
        p2->clear();
    }
}

bool permutation_group::is_member(const scalar_tr &tr, const perm_t &p) const {
    if(p.is_identity()) return tr.is_identity();
    scalar_tr trx(tr);
    if(is_member(m_br, 0, trx, p)) return trx.is_identity();
    return false;
}

void permutation_group::convert(std::vector<gen_t> &gs) const {
    make_genset(m_br, gs);
}

void permutation_group::stabilize(const seq_t &seq, permutation_group &g2) const {
    std::vector<gen_t> gs;
    make_setstabilizer(m_br, seq, gs);
    for(size_t i = 0; i < gs.size(); i++) g2.add_orbit(gs[i].second, gs[i].first);
}

void permutation_group::permute(const perm_t &p) {
    if(p.is_identity()) return;
    permute_branching(m_br, p);
}

size_t permutation_group::get_path(const branching &br, size_t i, size_t j,
    size_t *path) const {

    if(j <= i) return 0;
    size_t p[kMaxOrder];
    size_t k = j, len = 0;
    while(k != m_n && k != i) {
        p[len++] = k;
        k = br.edges[k];
    }
    if(k != i) return 0;
    for(k = 0; k < len; k++) path[k] = p[len - k - 1];
    return len;
}

bool permutation_group::is_member(const branching &br, size_t i, scalar_tr &tr,
    const perm_t &perm) const {

    if(perm.is_identity()) return true;
    if(i + 1 >= m_n) return false;
    if(perm[i] == i) return is_member(br, i + 1, tr, perm);
    for(size_t j = i + 1; j < m_n; j++) {
        size_t path[kMaxOrder];
        size_t len = get_path(br, i, j, path);
        if(len == 0) continue;
        perm_t rho(br.tau[i].first), tjinv(br.tau[j].first.inverse());
        rho.permute(tjinv).permute(perm);
        scalar_tr tr_rho(br.tau[i].second), tr_tjinv(br.tau[j].second);
        tr_tjinv.invert();
        tr_rho.transform(tr_tjinv).transform(tr);
        if(rho[i] == i) {
            if(is_member(br, i + 1, tr_rho, rho)) {
                tr = tr_rho;
                return true;
            }
        }
    }
    return false;
}

void permutation_group::make_branching(branching &br, size_t i,
    const std::vector<gen_t> &gs, std::vector<gen_t> &gs2) const {

    if(gs.empty()) return;
    const size_t n = m_n;
    std::vector<gen_t> transv(n, gen_t(perm_t(n), scalar_tr()));
    std::vector<size_t> delta;
    delta.push_back(i);
    std::list<size_t> s;
    s.push_back(i);
    transv[i].first.reset();
    transv[i].second.reset();

    while(!s.empty()) {
        size_t j = s.front();
        s.pop_front();
        for(size_t g = 0; g < gs.size(); g++) {
            size_t k = gs[g].first[j];
            std::vector<size_t>::iterator dd = delta.begin();
            while(dd != delta.end() && *dd != k) ++dd;
            if(dd == delta.end()) {
                gen_t p(gs[g]);
                p.first.permute(transv[j].first);
                p.second.transform(transv[j].second);
                transv[k] = p;
                delta.push_back(k);
                s.push_back(k);
            }
        }
    }

    for(size_t d = 0; d < delta.size(); d++) {
        size_t j = delta[d];
        if(j == i) continue;
        br.edges[j] = i;
        br.sigma[j] = transv[j];
        perm_t tp(br.sigma[j].first);
        tp.permute(br.tau[i].first);
        scalar_tr ts(br.sigma[j].second);
        ts.transform(br.tau[i].second);
        br.tau[j] = gen_t(tp, ts);
    }

    for(size_t g = 0; g < gs.size(); g++) {
        for(size_t d = 0; d < delta.size(); d++) {
            size_t j = delta[d], k = gs[g].first[j];
            gen_t p(transv[k]);
            p.first.invert();
            p.first.permute(gs[g].first).permute(transv[j].first);
            p.second.invert();
            p.second.transform(gs[g].second).transform(transv[j].second);
            if(!p.first.is_identity()) {
                size_t r = 0;
                while(r < gs2.size() && !gs2[r].first.equals(p.first)) r++;
                if(r == gs2.size()) gs2.push_back(p);
            } else if(!p.second.is_identity()) {
                throw lt_error(LT_ERR_SYMMETRY,
                    "permutation_group: illegal permutation");
            }
        }
    }
}

void permutation_group::make_genset(const branching &br,
    std::vector<gen_t> &gs) const {
    for(size_t i = 0; i < m_n; i++) {
        if(br.edges[i] != m_n && !br.sigma[i].first.is_identity())
            gs.push_back(br.sigma[i]);
    }
}

void permutation_group::permute_branching(branching &br, const perm_t &perm) const {
    std::vector<gen_t> gs1, gs2, gs3;
    make_genset(br, gs1);
    for(size_t i = 0; i < gs1.size(); i++) {
        size_t seq1[kMaxOrder], seq2[kMaxOrder];
        for(size_t j = 0; j < m_n; j++) seq1[j] = seq2[j] = j;
        {
            size_t b[kMaxOrder];
            for(size_t j = 0; j < m_n; j++) b[j] = seq2[j];
            for(size_t j = 0; j < m_n; j++) seq2[j] = b[gs1[i].first[j]];
        }
        perm_t pb = permutation_builder(m_n, seq2, seq1, &perm);
        gs2.push_back(gen_t(pb, gs1[i].second));
    }
    br.reset();
    std::vector<gen_t> *p1 = &gs2, *p2 = &gs3;
    for(size_t i = 0; i < m_n; i++) {
        make_branching(br, i, *p1, *p2);
        std::swap(p1, p2);
        p2->clear();
    }
}

void permutation_group::make_setstabilizer(const branching &br,
    const seq_t &msk, std::vector<gen_t> &gs) const {

    const size_t N = m_n;
    size_t m = 0;
    for(size_t i = 0; i < N; i++) if(msk[i] != 0) m++;
    if(m == 0) { make_genset(br, gs); return; }
    if(m == 1) {
        branching brx(N);
        std::vector<gen_t> gsx;
        make_genset(br, gsx);
        size_t i = 0;
        for(; i < N; i++) if(msk[i] != 0) break;
        make_branching(brx, i, gsx, gs);
        return;
    }

    for(size_t i = N - 1, ii = i - 1; i > 0; i--, ii--) {
        std::vector<gen_t> pu(N - i, gen_t(perm_t(N), scalar_tr()));
        std::vector<size_t> ui(N - i);
        for(size_t k = ii; k != N - 1; k++) ui[k - ii] = k;
        ui[0]++;
        for(; ui[0] < N; ui[0]++) {
            size_t k = br.edges[ui[0]];
            while(k != ii && k != N) k = br.edges[k];
            if(k == N) continue;
            pu[0].first.reset();
            pu[0].first.permute(br.tau[ui[0]].first);
            pu[0].first.permute(br.tau[k].first.inverse());
            scalar_tr tkinv(br.tau[k].second); tkinv.invert();
            pu[0].second.reset();
            pu[0].second.transform(br.tau[ui[0]].second);
            pu[0].second.transform(tkinv);
            break;
        }

        while(ui[0] != N) {
            gen_t g{perm_t(N), scalar_tr()};
            for(size_t k = 0; k < ui.size(); k++) {
                if(ui[k] == k + ii) continue;
                g.first.permute(pu[k].first);
                g.second.transform(pu[k].second);
            }
            size_t seq[kMaxOrder];
            for(size_t k = 0; k < N; k++) seq[k] = g.first[k];
            bool done[kMaxOrder] = {false};
            size_t l = 0;
            for(; l < N; l++) {
                if(done[l]) continue;
                if(msk[seq[l]] == 0) {
                    if(msk[l] != 0) break;
                    else continue;
                }
                size_t mm = l + 1;
                for(; mm < N; mm++) {
                    if(msk[mm] != msk[l]) continue;
                    done[mm] = true;
                    if(msk[seq[mm]] != msk[seq[l]]) break;
                }
                if(mm != N) break;
            }
            if(l == N) gs.push_back(g);

            for(size_t k = ui.size(), k1 = k - 1; k > 0; k--, k1--) {
                ui[k1]++;
                for(; ui[k1] < N; ui[k1]++) {
                    size_t mm = br.edges[ui[k1]];
                    while(mm != ii + k1 && mm != N) mm = br.edges[mm];
                    if(mm == N) continue;
                    pu[k1].first.reset();
                    pu[k1].first.permute(br.tau[ui[k1]].first);
                    pu[k1].first.permute(br.tau[mm].first.inverse());
                    scalar_tr tminv(br.tau[mm].second); tminv.invert();
                    pu[k1].second.reset();
                    pu[k1].second.transform(br.tau[ui[k1]].second);
                    pu[k1].second.transform(tminv);
                    break;
                }
                if(ui[k1] != N || k1 == 0) break;
                ui[k1] = ii + k1;
                pu[k1].first.reset();
                pu[k1].second.reset();
            }
        }
    }
}

// ---------------------------------------------------------------------------
// so_permute  (libtensor/symmetry/inst/so_permute_se_*_impl.h)

symmetry so_permute(const symmetry &sym, const perm_t &p) {
    bis_t bis(sym.bis);
    bis.permute(p);
    symmetry r(bis);
    for(size_t s = 0; s < sym.subsets.size(); s++) {
        const std::string &t = sym.subsets[s].first;
        const std::vector<std::shared_ptr<sym_element> > &set = sym.subsets[s].second;
        if(t == "perm") {
            permutation_group g(sym.order(), &set);
            g.permute(p);
            std::vector<permutation_group::gen_t> gs;
            g.convert(gs);
            for(size_t i = 0; i < gs.size(); i++)
                r.insert(se_perm(gs[i].first, gs[i].second));
        } else {
            for(size_t i = 0; i < set.size(); i++) {
                std::shared_ptr<sym_element> e(set[i]->clone());
                e->permute(p);
                r.insert_shared(e);
            }
        }
    }
    return r;
}

// ---------------------------------------------------------------------------
// Direct product and reduction for the result symmetry of a contraction.

namespace {

// so_dirprod for se_perm (inst/so_dirprod_se_perm_impl.h)
void dirprod_perm(size_t na, size_t nb, const perm_t &perm,
    const std::vector<std::shared_ptr<sym_element> > *s1,
    const std::vector<std::shared_ptr<sym_element> > *s2, symmetry &out) {

    size_t n = na + nb;
    perm_t pinv(perm.inverse());
    size_t map[2 * kMaxOrder];
    for(size_t j = 0; j < n; j++) map[j] = pinv[j];
    if(s1) for(size_t e = 0; e < s1->size(); e++) {
        const se_perm &e1 = dynamic_cast<const se_perm&>(*(*s1)[e]);
        size_t a2a[kMaxOrder], a2b[kMaxOrder];
        for(size_t k = 0; k < na; k++) { a2a[map[k]] = k; a2b[map[k]] = e1.tr.perm[k]; }
        for(size_t k = na; k < n; k++) a2a[map[k]] = a2b[map[k]] = k;
        perm_t pb = permutation_builder(n, a2b, a2a);
        out.insert(se_perm(pb, e1.tr.st));
    }
    if(s2) for(size_t e = 0; e < s2->size(); e++) {
        const se_perm &e2 = dynamic_cast<const se_perm&>(*(*s2)[e]);
        size_t a2a[kMaxOrder], a2b[kMaxOrder];
        for(size_t k = 0; k < na; k++) a2a[map[k]] = a2b[map[k]] = k;
        for(size_t k = na, l = 0; k < n; k++, l++) {
            a2a[map[k]] = k; a2b[map[k]] = e2.tr.perm[l] + na;
        }
        perm_t pb = permutation_builder(n, a2b, a2a);
        out.insert(se_perm(pb, e2.tr.st));
    }
}

// so_dirprod for se_part (inst/so_dirprod_se_part_impl.h)
void dirprod_part_one(const se_part &e, const size_t *map, size_t n,
    const bis_t &bis3, symmetry &out) {

    const dims_t &pd = e.pdims;
    index_t i3a(n), i3b(n);
    dims_t pd3(n);
    for(size_t i = 0; i < pd.n; i++) pd3[map[i]] = pd[i];
    se_part e3(bis3, pd3);
    index_t ia(pd.n);
    do {
        for(size_t i = 0; i < pd.n; i++) i3a[map[i]] = ia[i];
        if(e.is_forbidden(ia)) {
            e3.mark_forbidden(i3a);
            continue;
        }
        index_t ib = e.get_direct_map(ia);
        if(ia == ib) continue;
        index_t i3bb(n);
        for(size_t i = 0; i < n; i++) i3bb[i] = 0;
        for(size_t i = 0; i < pd.n; i++) i3bb[map[i]] = ib[i];
        e3.add_map(i3a, i3bb, e.get_transf(ia, ib));
    } while(inc_index(ia, pd));
    out.insert(e3);
}

void dirprod_part(size_t na, size_t nb, const perm_t &perm, const bis_t &bis3,
    const std::vector<std::shared_ptr<sym_element> > *s1,
    const std::vector<std::shared_ptr<sym_element> > *s2, symmetry &out) {

    perm_t pinv(perm.inverse());
    size_t map1[kMaxOrder], map2[kMaxOrder];
    for(size_t i = 0; i < na; i++) map1[i] = pinv[i];
    for(size_t i = 0; i < nb; i++) map2[i] = pinv[i + na];
    if(s1) for(size_t e = 0; e < s1->size(); e++)
        dirprod_part_one(dynamic_cast<const se_part&>(*(*s1)[e]), map1,
            na + nb, bis3, out);
    if(s2) for(size_t e = 0; e < s2->size(); e++)
        dirprod_part_one(dynamic_cast<const se_part&>(*(*s2)[e]), map2,
            na + nb, bis3, out);
}

// Direct product of label elements: conjunction of the rules, grouped by
// product table.
void dirprod_label(size_t na, size_t nb, const perm_t &perm, const bis_t &bis3,
    const std::vector<std::shared_ptr<sym_element> > *s1,
    const std::vector<std::shared_ptr<sym_element> > *s2, symmetry &out) {

    const size_t n = na + nb;
    perm_t pinv(perm.inverse());
    size_t map1[kMaxOrder], map2[kMaxOrder];
    for(size_t i = 0; i < na; i++) map1[i] = pinv[i];
    for(size_t i = 0; i < nb; i++) map2[i] = pinv[i + na];
    dims_t bid3 = bis3.block_index_dims();

    std::vector<std::string> tables;
    if(s1) for(size_t e = 0; e < s1->size(); e++) {
        const se_label &l = dynamic_cast<const se_label&>(*(*s1)[e]);
        if(std::find(tables.begin(), tables.end(), l.table_id) == tables.end())
            tables.push_back(l.table_id);
    }
    if(s2) for(size_t e = 0; e < s2->size(); e++) {
        const se_label &l = dynamic_cast<const se_label&>(*(*s2)[e]);
        if(std::find(tables.begin(), tables.end(), l.table_id) == tables.end())
            tables.push_back(l.table_id);
    }

    for(size_t t = 0; t < tables.size(); t++) {
        se_label e3(bid3, tables[t]);
        // start from the rule "always allowed": one empty-term product is
        // "forbidden", so begin with a product holding one invalid term.
        label_term always;
        always.seq = seq_t(n, 0);
        always.intrinsic = k_label_invalid;
        e3.rule.assign(1, label_product(1, always));

        // Elements without a rule of their own (the result of a contraction,
        // a symmetrization or a product of those) enter as factors: the
        // allowed blocks are the same conjunction of both sides.
        bool product = false;
        for(int side = 0; side < 2; side++) {
            const std::vector<std::shared_ptr<sym_element> > *s = side ? s2 : s1;
            if(s) for(size_t e = 0; e < s->size(); e++) {
                const se_label &l = dynamic_cast<const se_label&>(*(*s)[e]);
                if(l.table_id == tables[t] && (l.src || !l.alts.empty() || !l.factors.empty()))
                    product = true;
            }
        }
        if(product) {
            for(int side = 0; side < 2; side++) {
                const std::vector<std::shared_ptr<sym_element> > *s = side ? s2 : s1;
                const size_t *map = side ? map2 : map1;
                const size_t ns = side ? nb : na;
                if(!s) continue;
                for(size_t e = 0; e < s->size(); e++) {
                    const se_label &l = dynamic_cast<const se_label&>(*(*s)[e]);
                    if(l.table_id != tables[t]) continue;
                    for(size_t i = 0; i < ns; i++) e3.labels[map[i]] = l.labels[i];
                    se_label::factor f;
                    f.el.reset(new se_label(l));
                    f.map.assign(map, map + ns);
                    e3.factors.push_back(f);
                }
            }
            out.insert(e3);
            continue;
        }

        for(int side = 0; side < 2; side++) {
            const std::vector<std::shared_ptr<sym_element> > *s = side ? s2 : s1;
            const size_t *map = side ? map2 : map1;
            size_t ns = side ? nb : na;
            if(!s) continue;
            for(size_t e = 0; e < s->size(); e++) {
                const se_label &l = dynamic_cast<const se_label&>(*(*s)[e]);
                if(l.table_id != tables[t]) continue;
                for(size_t i = 0; i < ns; i++) e3.labels[map[i]] = l.labels[i];
                label_rule nr;
                for(size_t a = 0; a < e3.rule.size(); a++) {
                    for(size_t b = 0; b < l.rule.size(); b++) {
                        label_product p(e3.rule[a]);
                        if(l.rule[b].empty()) {
                            // an empty product forbids everything; keep it
                            // empty so evaluation reports "forbidden"
                            p.clear();
                            nr.push_back(p);
                            continue;
                        }
                        for(size_t k = 0; k < l.rule[b].size(); k++) {
                            label_term tt;
                            tt.seq = seq_t(n, 0);
                            for(size_t i = 0; i < ns; i++)
                                tt.seq[map[i]] = l.rule[b][k].seq[i];
                            tt.intrinsic = l.rule[b][k].intrinsic;
                            p.push_back(tt);
                        }
                        nr.push_back(p);
                    }
                }
                e3.rule.swap(nr);
            }
        }
        out.insert(e3);
    }
}

// combine_part (inst/combine_part_impl.h)
se_part combine_part(const std::vector<std::shared_ptr<sym_element> > &set) {
    lt_require(!set.empty(), LT_ERR_SYMMETRY, "combine_part: empty set");
    const se_part &first = dynamic_cast<const se_part&>(*set[0]);
    size_t n = first.pdims.n;
    dims_t pd(n);
    for(size_t e = 0; e < set.size(); e++) {
        const se_part &el = dynamic_cast<const se_part&>(*set[e]);
        lt_require(el.bis == first.bis, LT_ERR_SYMMETRY, "combine_part: bis");
        for(size_t i = 0; i < n; i++) {
            if(el.pdims[i] != 1) {
                lt_require(pd[i] == 1 || pd[i] == el.pdims[i], LT_ERR_SYMMETRY,
                    "combine_part: pdims");
                pd[i] = el.pdims[i];
            }
        }
    }
    se_part r(first.bis, pd);
    index_t i1(n);
    do {
        for(size_t e = 0; e < set.size(); e++) {
            const se_part &elx = dynamic_cast<const se_part&>(*set[e]);
            const dims_t &pdx = elx.pdims;
            index_t ix1(n);
            for(size_t i = 0; i < n; i++) if(pdx[i] != 1) ix1[i] = i1[i];
            if(elx.is_forbidden(ix1)) {
                r.mark_forbidden(i1);
                continue;
            }
            index_t ix2 = elx.get_direct_map(ix1);
            if(!(ix1 < ix2)) continue;
            scalar_tr trx = elx.get_transf(ix1, ix2);
            for(size_t i = 0; i < n; i++) if(pdx[i] == 1) ix2[i] = i1[i];
            if(i1 == ix2) continue;
            if(r.is_forbidden(i1)) {
                r.mark_forbidden(ix2);
                continue;
            }
            if(r.map_exists(i1, ix2)) {
                scalar_tr tr = r.get_transf(i1, ix2);
                if(trx != tr) r.mark_forbidden(i1);
            } else {
                r.add_map(i1, ix2, trx);
            }
        }
    } while(inc_index(i1, pd));
    return r;
}

struct reduce_params {
    size_t n;                 // order of the input
    bool msk[2 * kMaxOrder];
    size_t rseq[2 * kMaxOrder];
    index_t rbl_begin, rbl_end, ribl_begin, ribl_end;
};

bool part_forbidden(const se_part &el, const index_t &idx, const dims_t &sub) {
    if(!el.is_forbidden(idx)) return false;
    index_t ix(sub.n);
    while(inc_index(ix, sub)) {
        index_t ia(idx.n);
        for(size_t i = 0; i < idx.n; i++) ia[i] = idx[i] + ix[i];
        if(!el.is_forbidden(ia)) return false;
    }
    return true;
}

bool part_map_exists(const se_part &el, const index_t &i1, const index_t &i2,
    const dims_t &sub) {
    if(!el.map_exists(i1, i2)) return false;
    bool exists = true;
    scalar_tr tr = el.get_transf(i1, i2);
    index_t ix(sub.n);
    while(inc_index(ix, sub) && exists) {
        index_t a1(i1.n), a2(i2.n);
        for(size_t i = 0; i < i1.n; i++) { a1[i] = i1[i] + ix[i]; a2[i] = i2[i] + ix[i]; }
        if(!el.map_exists(a1, a2) || tr != el.get_transf(a1, a2)) exists = false;
    }
    return exists;
}

// so_reduce for se_part (inst/so_reduce_se_part_impl.h)
void reduce_part(const reduce_params &rp, size_t nsteps,
    const std::vector<std::shared_ptr<sym_element> > &set, symmetry &out) {

    if(set.empty()) return;
    const size_t NA = rp.n;
    size_t NB = 0;
    for(size_t i = 0; i < NA; i++) if(!rp.msk[i]) NB++;
    const size_t NR = NA - NB;

    size_t map[2 * kMaxOrder];
    bool rsteps[2 * kMaxOrder] = {false};
    for(size_t i = 0, j = 0; i < NA; i++) {
        if(rp.msk[i]) { rsteps[rp.rseq[i]] = true; map[i] = NB + rp.rseq[i]; }
        else map[i] = j++;
    }
    (void)nsteps;

    se_part ela = combine_part(set);
    dims_t bidimsa = ela.bis.block_index_dims();
    const dims_t &pdimsa = ela.pdims;

    index_t rpsz(NR), i1rp(NR), i2rp(NR);
    for(size_t i = 0; i < NR && rsteps[i]; i++) {
        size_t j = 0;
        for(; j < NA; j++) if(rp.msk[j] && rp.rseq[j] == i) break;
        size_t j1 = j++;
        size_t p1 = pdimsa[j1];
        for(; j < NA; j++) {
            if(!rp.msk[j] || rp.rseq[j] != i) continue;
            if(p1 == pdimsa[j]) continue;
            size_t p2 = pdimsa[j];
            if(p2 < p1) std::swap(p1, p2);
            p1 = (p2 % p1 == 0) ? p1 : 1;
        }
        size_t nbpp = bidimsa[j1] / p1;
        rpsz[i] = p1;
        i1rp[i] = rp.rbl_begin[j1] / nbpp;
        i2rp[i] = rp.rbl_end[j1] / nbpp;
    }
    dims_t rpdims(NR);
    for(size_t i = 0; i < NR; i++) rpdims[i] = i2rp[i] - i1rp[i] + 1;

    dims_t pdimsb(NB), pdimsr(NA);
    for(size_t i = 0; i < NA; i++) {
        if(rp.msk[i]) pdimsr[i] = pdimsa[i] / rpsz[rp.rseq[i]];
        else pdimsb[map[i]] = pdimsa[i];
    }

    bis_t bisb{dims_t(NB)};
    for(size_t i = 0, j = 0; i < NA; i++) {
        if(rp.msk[i]) continue;
        bisb.dims[j] = ela.bis.dims[i];
        bisb.splits[j] = ela.bis.splits[i];
        j++;
    }
    se_part elb(bisb, pdimsb);

    typedef std::pair<index_t, scalar_tr> part_t;
    bool empty = true;
    std::vector<bool> done(pdimsb.size(), false);

    auto build_list = [&](const index_t &ib, std::list<part_t> &l) {
        index_t idx(NR);
        do {
            index_t ia(NA);
            for(size_t i = 0, j = 0; i < NA; i++) {
                if(rp.msk[i]) {
                    size_t ir = rp.rseq[i];
                    ia[i] = (i1rp[ir] + idx[ir]) * pdimsr[i];
                } else {
                    ia[i] = ib[j++];
                }
            }
            if(part_forbidden(ela, ia, pdimsr)) continue;
            std::list<part_t>::iterator il = l.begin();
            for(; il != l.end(); ++il)
                if(part_map_exists(ela, il->first, ia, pdimsr)) break;
            if(il == l.end()) {
                l.push_back(part_t(ia, scalar_tr()));
            } else {
                il->second = scalar_tr(il->second.c + ela.get_transf(il->first, ia).c);
            }
        } while(inc_index(idx, rpdims));
        std::list<part_t>::iterator il = l.begin();
        while(il != l.end()) {
            if(il->second.is_zero()) il = l.erase(il);
            else ++il;
        }
    };

    index_t ib1(NB);
    size_t aib1 = 0;
    do {
        aib1 = abs_index(ib1, pdimsb);
        if(done[aib1]) continue;
        std::list<part_t> l1;
        build_list(ib1, l1);
        done[aib1] = true;
        if(l1.empty()) {
            elb.mark_forbidden(ib1);
            empty = false;
            continue;
        }
        index_t ib2(ib1);
        while(inc_index(ib2, pdimsb)) {
            std::list<part_t> l2;
            build_list(ib2, l2);
            if(l2.empty()) {
                elb.mark_forbidden(ib2);
                done[abs_index(ib2, pdimsb)] = true;
                empty = false;
                continue;
            }
            bool found = false;
            scalar_tr trb;
            std::list<part_t>::iterator il1 = l1.begin();
            for(; il1 != l1.end(); ++il1) {
                std::list<part_t>::iterator il2 = l2.begin();
                for(; il2 != l2.end(); ++il2) {
                    if(part_map_exists(ela, il1->first, il2->first, pdimsr) &&
                        il1->second == il2->second) break;
                }
                if(il2 == l2.end()) break;
                scalar_tr tr = ela.get_transf(il1->first, il2->first);
                if(found && trb != tr) break;
                trb = tr;
                found = true;
                l2.erase(il2);
            }
            if(il1 == l1.end()) {
                elb.add_map(ib1, ib2, trb);
                empty = false;
                break;
            }
        }
    } while(inc_index(ib1, pdimsb));

    if(!empty) out.insert(elb);
}

// so_reduce for se_perm (inst/so_reduce_se_perm_impl.h)
void reduce_perm(const reduce_params &rp,
    const std::vector<std::shared_ptr<sym_element> > &set, symmetry &out) {

    const size_t N1 = rp.n;
    permutation_group g1(N1, &set), g2(N1);
    seq_t seq(N1, 0);
    for(size_t i = 0; i < N1; i++) if(rp.msk[i]) seq[i] = rp.rseq[i] + 1;
    g1.stabilize(seq, g2);
    std::vector<permutation_group::gen_t> gs;
    g2.convert(gs);

    for(size_t e = 0; e < gs.size(); e++) {
        const perm_t &p = gs[e].first;
        index_t bia2(rp.rbl_begin), bib2(rp.rbl_end), ia2(rp.ribl_begin), ib2(rp.ribl_end);
        p.apply(bia2); p.apply(bib2); p.apply(ia2); p.apply(ib2);
        size_t j = 0;
        for(; j < N1; j++) {
            if(!rp.msk[j]) continue;
            if(bia2[j] != rp.rbl_begin[j] || bib2[j] != rp.rbl_end[j] ||
                ia2[j] != rp.ribl_begin[j] || ib2[j] != rp.ribl_end[j]) break;
        }
        if(j != N1) continue;

        size_t seq1b[kMaxOrder], seq2b[kMaxOrder];
        size_t k = 0;
        for(j = 0; j < N1; j++) {
            if(rp.msk[j]) continue;
            seq1b[k] = j;
            seq2b[k] = p[j];
            k++;
        }
        perm_t pb = permutation_builder(k, seq2b, seq1b);
        if(pb.is_identity()) {
            lt_require(gs[e].second.is_identity(), LT_ERR_SYMMETRY,
                "anti-symmetric identity permutation");
            continue;
        }
        out.insert(se_perm(pb, gs[e].second));
    }
}

// Reduction of a label element: the result allows a block if some labels of
// the contracted dimensions satisfy the full rule.
void reduce_label(const reduce_params &rp, const bis_t &bisc,
    const std::vector<std::shared_ptr<sym_element> > &set, symmetry &out) {

    const size_t N1 = rp.n;
    dims_t bidc = bisc.block_index_dims();
    for(size_t e = 0; e < set.size(); e++) {
        std::shared_ptr<se_label> full(
            new se_label(dynamic_cast<const se_label&>(*set[e])));
        se_label r(bidc, full->table_id);
        r.src = full;
        size_t nsteps = 0;
        for(size_t i = 0; i < N1; i++)
            if(rp.msk[i] && rp.rseq[i] + 1 > nsteps) nsteps = rp.rseq[i] + 1;
        r.rdims.assign(nsteps, std::vector<size_t>());
        r.rlabels.assign(nsteps, std::vector<size_t>());
        for(size_t i = 0, j = 0; i < N1; i++) {
            if(rp.msk[i]) {
                size_t s = rp.rseq[i];
                r.rdims[s].push_back(i);
                for(size_t b = rp.rbl_begin[i]; b <= rp.rbl_end[i]; b++)
                    r.rlabels[s].push_back(full->labels[i][b]);
            } else {
                r.keep.push_back(i);
                r.labels[j] = full->labels[i];
                j++;
            }
        }
        for(size_t s = 0; s < nsteps; s++) {
            std::vector<size_t> &v = r.rlabels[s];
            std::sort(v.begin(), v.end());
            v.erase(std::unique(v.begin(), v.end()), v.end());
        }
        out.insert(r);
    }
}

} // unnamed namespace

symmetry contract2_symmetry(const contraction2 &c, const symmetry &syma,
    const symmetry &symb, bool self) {

    const size_t N = c.N, M = c.M, K = c.K;
    const size_t NA = N + K, NB = M + K, NC = N + M, NX = NA + NB;

    // Direct-product space: A dims then B dims
    bis_t xbis{dims_t(NX)};
    for(size_t i = 0; i < NA; i++) {
        xbis.dims[i] = syma.bis.dims[i];
        xbis.splits[i] = syma.bis.splits[i];
    }
    for(size_t i = 0; i < NB; i++) {
        xbis.dims[NA + i] = symb.bis.dims[i];
        xbis.splits[NA + i] = symb.bis.splits[i];
    }

    size_t seq1[2 * kMaxOrder], seq2[2 * kMaxOrder];
    reduce_params rp;
    rp.n = NX;
    for(size_t i = 0; i < NX; i++) { rp.msk[i] = false; rp.rseq[i] = 0; }
    for(size_t i = 0, k = 0; i < NX; i++) {
        seq1[i] = i;
        if(c.conn[i + NC] < NC) {
            seq2[c.conn[i + NC]] = i;
        } else if(i < NA) {
            size_t j = N + M + 2 * k;
            rp.msk[j] = rp.msk[j + 1] = true;
            rp.rseq[j] = rp.rseq[j + 1] = k;
            seq2[j] = i;
            seq2[j + 1] = c.conn[i + NC] - NC;
            k++;
        }
    }
    lt_require(NX <= kMaxOrder || true, LT_ERR_PARAM, "order");
    perm_t pb = permutation_builder(NX, seq2, seq1);
    xbis.permute(pb);
    symmetry xsym(xbis);

    // so_dirprod (inst/so_dirprod_impl.h: types of A first, then B-only)
    std::vector<std::string> types;
    for(size_t s = 0; s < syma.subsets.size(); s++) types.push_back(syma.subsets[s].first);
    for(size_t s = 0; s < symb.subsets.size(); s++) {
        const std::string &t = symb.subsets[s].first;
        if(std::find(types.begin(), types.end(), t) == types.end()) types.push_back(t);
    }
    for(size_t t = 0; t < types.size(); t++) {
        const std::vector<std::shared_ptr<sym_element> > *s1 = syma.subset(types[t]);
        const std::vector<std::shared_ptr<sym_element> > *s2 = symb.subset(types[t]);
        if(types[t] == "perm") dirprod_perm(NA, NB, pb, s1, s2, xsym);
        else if(types[t] == "part") dirprod_part(NA, NB, pb, xbis, s1, s2, xsym);
        else if(types[t] == "label") dirprod_label(NA, NB, pb, xbis, s1, s2, xsym);
    }

    if(self) {
        perm_t permab(pb.inverse());
        for(size_t i = 0; i < NB; i++) permab.swap(i, NA + i);
        permab.permute(pb);
        if(!permab.is_identity()) xsym.insert(se_perm(permab, scalar_tr()));
    }

    bis_t bisc = contract2_bis(c, syma.bis, symb.bis);
    symmetry symc(bisc);
    if(K == 0) {
        for(size_t s = 0; s < xsym.subsets.size(); s++)
            for(size_t e = 0; e < xsym.subsets[s].second.size(); e++)
                symc.insert(*xsym.subsets[s].second[e]);
        return symc;
    }

    dims_t bidims = xbis.block_index_dims();
    rp.rbl_begin = index_t(NX); rp.rbl_end = index_t(NX);
    rp.ribl_begin = index_t(NX); rp.ribl_end = index_t(NX);
    index_t bib(NX);
    for(size_t i = 0; i < NX; i++) { bib[i] = bidims[i] - 1; rp.rbl_end[i] = bib[i]; }
    dims_t bdims = xbis.block_dims(bib);
    for(size_t i = 0; i < NX; i++) rp.ribl_end[i] = bdims[i] - 1;

    for(size_t s = 0; s < xsym.subsets.size(); s++) {
        const std::string &t = xsym.subsets[s].first;
        const std::vector<std::shared_ptr<sym_element> > &set = xsym.subsets[s].second;
        if(t == "perm") reduce_perm(rp, set, symc);
        else if(t == "part") reduce_part(rp, K, set, symc);
        else if(t == "label") reduce_label(rp, bisc, set, symc);
    }
    return symc;
}

// ---------------------------------------------------------------------------
// Symmetrization over index groups (so_symmetrize,
// libtensor/symmetry/inst/so_symmetrize_impl.h and the handlers
// so_symmetrize_se_{perm,part,label}_impl.h).  Group g (1..ngrp) holds the
// dimensions with idxgrp == g; symidx orders the dimensions inside a group;
// the symmetrization permutes whole groups.

namespace {

// Permutations of the positions that are not fixed, one exchange of
// neighbouring positions per step: Trotter's algorithm 115 (Commun. ACM 5
// (1962) 434) in the form libtensor enumerates it
// (core/permutation_generator.h); the order of the sequence decides which
// conjugations so_symmetrize composes.
class perm_walk {
public:
    perm_walk(size_t n, const bool *fixed) : m_perm(n), m_m(0), m_done(false) {
        for(size_t i = 0; i < kMaxOrder; i++) { m_up[i] = true; m_c[i] = 0; m_pos[i] = 0; }
        for(size_t i = 0; i < n; i++) if(!fixed[i]) m_pos[m_m++] = i;
        if(m_m < 2) m_done = true;
    }
    const perm_t &perm() const { return m_perm; }
    bool step() {
        if(m_done) return false;
        size_t i = m_m - 1, carried = 0;
        for(; i > 0; i--) {
            if(m_up[i]) m_c[i]++;
            else m_c[i]--;
            if(m_c[i] == i + 1) { m_up[i] = false; continue; }
            if(m_c[i] != 0) break;
            m_up[i] = true;
            carried++;
        }
        if(i == 0) m_done = true;
        const size_t q = m_done ? carried : m_c[i] - 1 + carried;
        m_perm.swap(m_pos[q], m_pos[q + 1]);
        return !m_done;
    }
private:
    perm_t m_perm;
    size_t m_pos[kMaxOrder], m_c[kMaxOrder], m_m;
    bool m_up[kMaxOrder], m_done;
};

struct sym_groups {
    size_t n, ngrp, nidx;
    seq_t idxgrp, symidx;
};

uint64_t perm_code(const perm_t &p) {
    uint64_t c = 0;
    for(size_t i = 0; i < p.n; i++) c = (c << 4) | p.p[i];
    return c;
}

// Position of dimension (group g, index k), both counted from 1.
size_t group_dim(const sym_groups &G, size_t g, size_t k) {
    for(size_t j = 0; j < G.n; j++)
        if(G.idxgrp[j] == g && G.symidx[j] == k) return j;
    return G.n;
}

// se_perm handler (so_symmetrize_se_perm_impl.h): the group elements,
// breadth-first from the distinct generators, that stay in the group under
// every conjugation the walk over group permutations composes, plus the
// cyclic (more than two groups) and pair permutations.
void symmetrize_perm(const sym_groups &G,
    const std::vector<std::shared_ptr<sym_element> > *set,
    const scalar_tr &trp, const scalar_tr &trc, symmetry &out) {

    typedef permutation_group::gen_t gen_t;
    const size_t n = G.n;
    perm_t pp(n), cp(n);
    for(size_t g = 1; g < G.ngrp; g++) {
        for(size_t k = 1; k <= G.nidx; k++) {
            const size_t i1 = group_dim(G, g, k), j = group_dim(G, g + 1, k);
            if(g == 1) pp.swap(i1, j);
            cp.swap(i1, j);
        }
    }
    if(!set || set->empty()) {
        if(G.ngrp > 2) out.insert(se_perm(cp, trc));
        out.insert(se_perm(pp, trp));
        return;
    }
    std::vector<gen_t> l0, l2;
    std::list<gen_t> l1;
    std::set<uint64_t> done;
    done.insert(perm_code(perm_t(n)));
    for(size_t e = 0; e < set->size(); e++) {
        const se_perm &el = dynamic_cast<const se_perm&>(*(*set)[e]);
        uint64_t c = perm_code(el.tr.perm);
        if(done.count(c)) continue;
        done.insert(c);
        l0.push_back(gen_t(el.tr.perm, el.tr.st));
        l1.push_back(l0.back());
    }
    while(!l1.empty()) {
        const gen_t cur = l1.front();
        for(size_t g = 0; g < l0.size(); g++) {
            perm_t p(cur.first);
            p.permute(l0[g].first);
            uint64_t c = perm_code(p);
            if(done.count(c)) continue;
            done.insert(c);
            scalar_tr tr(cur.second);
            tr.transform(l0[g].second);
            l1.push_back(gen_t(p, tr));
        }
        l2.push_back(cur);
        l1.pop_front();
    }
    std::vector<char> keep(l2.size(), 1);
    bool fixed[kMaxOrder];
    for(size_t i = 0; i < n; i++) fixed[i] = G.idxgrp[i] == 0 || G.symidx[i] != 1;
    perm_walk pw(n, fixed);
    scalar_tr trx;
    perm_t px(n), p1(n);
    while(pw.step()) {
        const perm_t &p2 = pw.perm();
        if(G.ngrp == 2) {
            trx.transform(trp);
            px.permute(pp);
        } else {
            size_t i = 0;
            for(; i < n && p2[i] == p1[i]; i++) { trx.transform(trc); px.permute(cp); }
            trx.transform(trp);
            px.permute(pp);
            i++;
            for(; i < n; i++) { trx.transform(trc); px.permute(cp); }
        }
        const perm_t pxinv(px.inverse());
        scalar_tr trxinv(trx);
        trxinv.invert();
        p1 = p2;
        permutation_group gx(n);
        for(size_t g = 0; g < l0.size(); g++) {
            perm_t pi(pxinv);
            pi.permute(l0[g].first).permute(px);
            scalar_tr ti(trxinv);
            ti.transform(l0[g].second).transform(trx);
            gx.add_orbit(ti, pi);
        }
        for(size_t i = 0; i < l2.size(); i++)
            if(keep[i] && !gx.is_member(l2[i].second, l2[i].first)) keep[i] = 0;
    }
    permutation_group g2(n);
    for(size_t i = 0; i < l2.size(); i++)
        if(keep[i]) g2.add_orbit(l2[i].second, l2[i].first);
    if(G.ngrp > 2) g2.add_orbit(trc, cp);
    g2.add_orbit(trp, pp);
    std::vector<gen_t> gs;
    g2.convert(gs);
    for(size_t i = 0; i < gs.size(); i++) out.insert(se_perm(gs[i].first, gs[i].second));
}

// The index i1 with its groups permuted by pn (so_symmetrize_se_part_impl.h:
// map[(symidx - 1) ngrp + idxgrp - 1] is a dimension; pn permutes the
// entries of every run of ngrp).
index_t permute_groups(const index_t &i1, const perm_t &pn, const std::vector<size_t> &map,
    size_t ngrp) {
    index_t ix(i1);
    for(size_t b = 0; b + ngrp <= map.size(); b += ngrp)
        for(size_t j = 0; j < ngrp; j++) ix[map[b + j]] = i1[map[b + pn[j]]];
    return ix;
}

std::vector<perm_t> group_perms(size_t ngrp, size_t n) {
    bool fixed[kMaxOrder];
    for(size_t i = 0; i < n; i++) fixed[i] = i >= ngrp;
    std::vector<perm_t> out;
    perm_walk pw(n, fixed);
    do out.push_back(pw.perm()); while(pw.step());
    return out;
}

// Consistent map i1 -> i2 of the partition for every group permutation of
// the pair (so_symmetrize_se_part_impl.h, map_exists).
bool part_map_all(const se_part &sp, const index_t &i1, const index_t &i2,
    const std::vector<perm_t> &pns, const std::vector<size_t> &map, size_t ngrp) {
    scalar_tr tr;
    size_t k = 0;
    for(; k < pns.size(); k++) {
        index_t j1 = permute_groups(i1, pns[k], map, ngrp);
        index_t j2 = permute_groups(i2, pns[k], map, ngrp);
        if(sp.map_exists(j1, j2)) {
            tr = sp.get_transf(j1, j2);
            break;
        }
        if(!sp.is_forbidden(j1) || !sp.is_forbidden(j2)) return false;
    }
    for(k++; k < pns.size(); k++) {
        index_t j1 = permute_groups(i1, pns[k], map, ngrp);
        index_t j2 = permute_groups(i2, pns[k], map, ngrp);
        if(sp.map_exists(j1, j2)) {
            if(tr != sp.get_transf(j1, j2)) return false;
        } else if(!sp.is_forbidden(j1) || !sp.is_forbidden(j2)) {
            return false;
        }
    }
    return true;
}

// se_part handler (so_symmetrize_se_part_impl.h:20-75).
void symmetrize_part(const sym_groups &G,
    const std::vector<std::shared_ptr<sym_element> > &set, symmetry &out) {

    if(set.empty()) return;
    se_part cp = combine_part(set);
    const dims_t pdims = cp.pdims;
    std::vector<size_t> map(G.ngrp * G.nidx, G.n);
    for(size_t i = 0; i < G.n; i++)
        if(G.idxgrp[i]) map[(G.symidx[i] - 1) * G.ngrp + G.idxgrp[i] - 1] = i;
    // The reference compares these entries of its map (for several indexes
    // per group they are not the dimensions the symmetrization exchanges);
    // kept as it is.
    for(size_t g = 1; g < G.ngrp; g++)
        for(size_t j = 0; j < G.nidx; j++)
            if(pdims[map[g * G.nidx + j]] != pdims[map[j]]) return;
    const std::vector<perm_t> pns = group_perms(G.ngrp, G.n);
    se_part sp2(cp.bis, pdims);
    index_t i1(pdims.n);
    do {
        bool allf = true;
        for(size_t k = 0; k < pns.size() && allf; k++)
            if(!cp.is_forbidden(permute_groups(i1, pns[k], map, G.ngrp))) allf = false;
        if(allf) {
            for(size_t k = 0; k < pns.size(); k++)
                sp2.mark_forbidden(permute_groups(i1, pns[k], map, G.ngrp));
            continue;
        }
        if(cp.is_forbidden(i1)) continue;
        index_t i2 = cp.get_direct_map(i1);
        bool found = false;
        while(!found && i1 < i2) {
            if(part_map_all(cp, i1, i2, pns, map, G.ngrp)) found = true;
            else i2 = cp.get_direct_map(i2);
        }
        if(found) {
            const scalar_tr tr = cp.get_transf(i1, i2);
            for(size_t k = 0; k < pns.size(); k++)
                sp2.add_map(permute_groups(i1, pns[k], map, G.ngrp),
                    permute_groups(i2, pns[k], map, G.ngrp), tr);
        }
    } while(inc_index(i1, pdims));
    out.insert(sp2);
}

// se_label handler (so_symmetrize_se_label_impl.h:20-95): every rule united
// with its copies under all group permutations, kept as the element plus the
// permutations whose images it also allows.
void symmetrize_label(const sym_groups &G,
    const std::vector<std::shared_ptr<sym_element> > &set, symmetry &out) {

    std::vector<size_t> map(G.ngrp * G.nidx, G.n);
    for(size_t i = 0; i < G.n; i++)
        if(G.idxgrp[i]) map[(G.idxgrp[i] - 1) * G.nidx + G.symidx[i] - 1] = i;
    // tensor permutations of the group permutations
    std::vector<perm_t> gp;
    const std::vector<perm_t> pns = group_perms(G.ngrp, G.n);
    for(size_t k = 0; k < pns.size(); k++) {
        perm_t q(G.n);
        for(size_t g = 0; g < G.ngrp; g++)
            for(size_t j = 0; j < G.nidx; j++)
                q.p[map[g * G.nidx + j]] = (uint8_t)map[pns[k][g] * G.nidx + j];
        if(!q.is_identity()) gp.push_back(q);
    }
    for(size_t e = 0; e < set.size(); e++) {
        const se_label &l1 = dynamic_cast<const se_label&>(*set[e]);
        for(size_t j = 0; j < G.nidx; j++) {
            const size_t d0 = map[j];
            for(size_t g = 1; g < G.ngrp; g++) {
                const size_t d = map[g * G.nidx + j];
                lt_require(l1.bidims[d] == l1.bidims[d0] && l1.labels[d] == l1.labels[d0],
                    LT_ERR_SYMMETRY, "so_symmetrize: incompatible dimensions");
            }
        }
        std::vector<perm_t> alts(l1.alts);
        for(size_t k = 0; k < gp.size(); k++) {
            alts.push_back(gp[k]);
            for(size_t a = 0; a < l1.alts.size(); a++) {
                perm_t q(gp[k]);
                q.permute(l1.alts[a]);
                alts.push_back(q);
            }
        }
        se_label l2(l1);
        l2.alts.clear();
        for(size_t k = 0; k < alts.size(); k++) {
            if(alts[k].is_identity()) continue;
            bool seen = false;
            for(size_t j = 0; j < l2.alts.size(); j++) if(l2.alts[j] == alts[k]) seen = true;
            if(!seen) l2.alts.push_back(alts[k]);
        }
        out.insert(l2);
    }
}

symmetry so_symmetrize_groups(const symmetry &sym, const sym_groups &G,
    const scalar_tr &trp, const scalar_tr &trc) {

    symmetry r(sym.bis);
    bool perm_done = false;
    for(size_t k = 0; k < sym.subsets.size(); k++) {
        const std::string &t = sym.subsets[k].first;
        const std::vector<std::shared_ptr<sym_element> > &set = sym.subsets[k].second;
        if(t == "perm") {
            perm_done = true;
            symmetrize_perm(G, &set, trp, trc, r);
        } else if(t == "part") {
            symmetrize_part(G, set, r);
        } else if(t == "label") {
            symmetrize_label(G, set, r);
        } else {
            throw lt_error(LT_ERR_SYMMETRY, "so_symmetrize: unknown element type " + t);
        }
    }
    if(!perm_done) symmetrize_perm(G, 0, trp, trc, r);
    return r;
}

void require_maps_bis(const symmetry &sym, const perm_t &perm, const char *what) {
    bis_t pb(sym.bis);
    pb.permute(perm);
    lt_require(pb == sym.bis, LT_ERR_SYMMETRY, (std::string(what) +
        ": the permutation does not map the block index space onto itself").c_str());
}

bool is_pair_perm(const perm_t &p) {
    perm_t p2(p);
    p2.permute(p);
    return !p.is_identity() && p2.is_identity();
}

} // unnamed namespace

symmetry so_symmetrize_pair(const symmetry &sym, const perm_t &perm, double s) {
    const size_t n = sym.order();
    lt_require(perm.n == n && is_pair_perm(perm), LT_ERR_PARAM,
        "symmetrize2: the permutation must be a non-trivial pair permutation");
    require_maps_bis(sym, perm, "symmetrize2");
    // Index groups as gen_bto_symmetrize2::make_symmetry numbers them: the
    // k-th pair is (i, perm[i]) for the k-th i with perm[i] > i.
    sym_groups G;
    G.n = n;
    G.ngrp = 2;
    G.nidx = 0;
    G.idxgrp = seq_t(n, 0);
    G.symidx = seq_t(n, 0);
    for(size_t i = 0; i < n; i++) {
        if(perm[i] <= i) continue;
        G.nidx++;
        G.idxgrp[i] = 1;
        G.idxgrp[perm[i]] = 2;
        G.symidx[i] = G.symidx[perm[i]] = G.nidx;
    }
    return so_symmetrize_groups(sym, G, scalar_tr(s), scalar_tr(1.0));
}

symmetry so_symmetrize_triple(const symmetry &sym, const perm_t &p1, const perm_t &p2,
    double s) {
    const size_t n = sym.order();
    lt_require(p1.n == n && p2.n == n && is_pair_perm(p1) && is_pair_perm(p2), LT_ERR_PARAM,
        "symmetrize3: each permutation must be a non-trivial pair permutation");
    perm_t p12(p1);
    p12.permute(p2);
    perm_t p123(p12);
    p123.permute(p12).permute(p12);
    lt_require(!p12.is_identity() && p123.is_identity(), LT_ERR_PARAM,
        "symmetrize3: the two permutations must generate the exchanges of three groups");
    require_maps_bis(sym, p1, "symmetrize3");
    require_maps_bis(sym, p2, "symmetrize3");
    // gen_bto_symmetrize3::make_symmetry: dimensions moved by both
    // permutations form group 3, by p1 only group 1, by p2 only group 2,
    // numbered in order within each group.
    sym_groups G;
    G.n = n;
    G.ngrp = 3;
    G.idxgrp = seq_t(n, 0);
    G.symidx = seq_t(n, 0);
    size_t c[4] = { 0, 1, 1, 1 };
    for(size_t i = 0; i < n; i++) {
        const bool m1 = p1[i] != i, m2 = p2[i] != i;
        size_t g = m1 ? (m2 ? 3 : 1) : (m2 ? 2 : 0);
        if(!g) continue;
        G.idxgrp[i] = g;
        G.symidx[i] = c[g]++;
    }
    G.nidx = c[1] - 1;
    lt_require(c[2] - 1 == G.nidx && c[3] - 1 == G.nidx, LT_ERR_PARAM,
        "symmetrize3: groups of different sizes");
    return so_symmetrize_groups(sym, G, scalar_tr(s), scalar_tr(1.0));
}


// ---------------------------------------------------------------------------
// Symmetry of a sum of two tensors.

namespace {

//! The se_perm elements of a subset as (permutation, scalar) pairs.
std::vector<tensor_tr> perm_gens(const std::vector<std::shared_ptr<sym_element> > *set) {
    std::vector<tensor_tr> out;
    if(set) for(size_t i = 0; i < set->size(); i++) {
        const se_perm &e = dynamic_cast<const se_perm&>(*(*set)[i]);
        lt_require(e.tr.st.c == 1.0 || e.tr.st.c == -1.0, LT_ERR_SYMMETRY,
            "sum of tensors: permutational symmetry with a scalar other than +1 or -1");
        out.push_back(e.tr);
    }
    return out;
}

size_t perm_order(const perm_t &p) {
    perm_t q(p);
    size_t k = 1;
    while(!q.is_identity()) { q.permute(p); k++; }
    return k;
}

//! Adds to grp (over na + nb indexes) the elements libtensor's so_dirsum
//! forms from one fixed side (perm `fixed`, scalar t) and the generators of
//! the other side (inst/so_dirsum_se_perm_impl.h, combine()): with t = +1,
//! fixed (+) g for the generators g of scalar +1; with t = -1, fixed (+) g
//! for the generators of scalar -1 when the other side has no generator of
//! scalar +1, else fixed (+) h^j g h^(i-j) for every generator h of scalar
//! +1, 0 <= j < i < order(h), and g of scalar -1.  (For scalars +1 and -1 the
//! reference's search for products of the remaining generators contributes
//! only the identity.)  left: the fixed side is the first na indexes.
void dirsum_combine(size_t na, size_t nb, bool left, const perm_t &fixed, double t,
    const std::vector<tensor_tr> &other, permutation_group &grp) {
    if(other.empty()) return;
    std::vector<perm_t> p1, p2;   // scalar == t, scalar +1 (when t != +1)
    for(size_t i = 0; i < other.size(); i++) {
        if(other[i].st.c == t) p1.push_back(other[i].perm);
        else if(other[i].st.c == 1.0) p2.push_back(other[i].perm);
    }
    const size_t n = na + nb;
    std::vector<perm_t> add;
    if(!p2.empty()) {
        for(size_t a = 0; a < p2.size(); a++) {
            const size_t ord = perm_order(p2[a]);
            for(size_t i = 1; i < ord; i++)
                for(size_t j = 0; j < i; j++)
                    for(size_t b = 0; b < p1.size(); b++) {
                        perm_t px(left ? nb : na);
                        size_t k = 0;
                        for(; k < j; k++) px.permute(p2[a]);
                        px.permute(p1[b]);
                        for(; k < i; k++) px.permute(p2[a]);
                        add.push_back(px);
                    }
        }
    } else {
        add = p1;
    }
    for(size_t x = 0; x < add.size(); x++) {
        // sequence of the combined permutation: the fixed side's images,
        // the other side's images shifted past the first na indexes
        size_t seqa[2 * kMaxOrder], seqb[2 * kMaxOrder];
        for(size_t i = 0; i < n; i++) seqa[i] = i;
        for(size_t i = 0; i < na; i++) seqb[i] = left ? fixed[i] : add[x][i];
        for(size_t i = 0; i < nb; i++) seqb[na + i] = (left ? add[x][i] : fixed[i]) + na;
        grp.add_orbit(scalar_tr(t), permutation_builder(n, seqb, seqa));
    }
}

//! libtensor's so_dirsum of two se_perm sets followed by so_merge of index
//! i with index i + n (gen_bto_aux_add_impl.h:40-66,
//! addition_schedule_impl.h:27-45, gen_bto_add_impl.h:238-267): the
//! combined group is built from the generators as the reference does
//! (inst/so_dirsum_se_perm_impl.h:24-60), permuted by `perm` (the second
//! operand's permutation in btod_add), and its stabilizer of the diagonal
//! pairs mapped back to n indexes (inst/so_merge_se_perm_impl.h:26-97).  The
//! resulting group can be smaller than the intersection of the two groups:
//! the reference combines generators, not group elements.
void sum_perm(const symmetry &a, const symmetry &b, const perm_t &bperm, symmetry &out) {
    const std::vector<tensor_tr> g1 = perm_gens(a.subset("perm")),
        g2 = perm_gens(b.subset("perm"));
    if(g1.empty() && g2.empty()) return;
    const size_t n = a.order(), n2 = 2 * n;
    permutation_group grp(n2);
    const perm_t id(n);
    dirsum_combine(n, n, false, id, 1.0, g1, grp);
    dirsum_combine(n, n, true, id, 1.0, g2, grp);
    for(size_t i = 0; i < g1.size(); i++)
        if(!g1[i].st.is_identity()) dirsum_combine(n, n, true, g1[i].perm, g1[i].st.c, g2, grp);
    for(size_t i = 0; i < g2.size(); i++)
        if(!g2[i].st.is_identity()) dirsum_combine(n, n, false, g2[i].perm, g2[i].st.c, g1, grp);
    // the second operand's permutation acts on the upper n indexes
    perm_t pfull(n2);
    {
        size_t s1[2 * kMaxOrder], s2[2 * kMaxOrder];
        seq_t sa(n);
        for(size_t i = 0; i < n; i++) sa[i] = i + n;
        bperm.apply(sa);
        for(size_t i = 0; i < n; i++) { s1[i] = s2[i] = i; }
        for(size_t i = n; i < n2; i++) { s1[i] = i; s2[i] = sa[i - n]; }
        pfull = permutation_builder(n2, s2, s1);
    }
    grp.permute(pfull);
    std::vector<permutation_group::gen_t> gs;
    grp.convert(gs);
    // so_merge: the group again from those generators, its stabilizer of
    // the index pairs (i, i + n), each element restricted to n indexes
    std::vector<std::shared_ptr<sym_element> > set;
    for(size_t i = 0; i < gs.size(); i++)
        set.push_back(std::shared_ptr<sym_element>(new se_perm(gs[i].first, gs[i].second)));
    permutation_group g1x(n2, &set), g2x(n2);
    seq_t sseq(n2);
    for(size_t i = 0; i < n2; i++) sseq[i] = (i % n) + 1;
    g1x.stabilize(sseq, g2x);
    std::vector<permutation_group::gen_t> st;
    g2x.convert(st);
    for(size_t e = 0; e < st.size(); e++) {
        seq_t img(n2);
        for(size_t i = 0; i < n2; i++) img[i] = i;
        st[e].first.apply(img);
        size_t s1[kMaxOrder], s2[kMaxOrder];
        for(size_t i = 0; i < n; i++) {
            s1[i] = i;
            s2[i] = std::min(img[i], img[i + n]);
        }
        const perm_t pm = permutation_builder(n, s2, s1);
        if(pm.is_identity()) {
            lt_require(st[e].second.is_identity(), LT_ERR_SYMMETRY,
                "sum of tensors: contradictory permutational symmetry");
            continue;
        }
        out.insert(se_perm(pm, st[e].second));
    }
}

void sum_part(const symmetry &a, const symmetry &b, symmetry &out) {
    const std::vector<std::shared_ptr<sym_element> > *sa = a.subset("part"),
        *sb = b.subset("part");
    if(!sa || !sb) return;
    for(size_t i = 0; i < sa->size(); i++) {
        const se_part &pa = dynamic_cast<const se_part&>(*(*sa)[i]);
        for(size_t j = 0; j < sb->size(); j++) {
            const se_part &pb = dynamic_cast<const se_part&>(*(*sb)[j]);
            if(!(pa.pdims == pb.pdims)) continue;
            se_part pc(a.bis, pa.pdims);
            const size_t np = pa.pdims.size();
            // forbidden in both terms, else related to the smallest partition
            // both relate it to with the same scalar
            for(size_t x = 0; x < np; x++) {
                const index_t ix = get_index(x, pa.pdims);
                if(pa.is_forbidden(ix) && pb.is_forbidden(ix)) {
                    pc.mark_forbidden(ix);
                    continue;
                }
                if(pa.is_forbidden(ix) || pb.is_forbidden(ix)) continue;
                for(size_t y = 0; y < x; y++) {
                    const index_t iy = get_index(y, pa.pdims);
                    if(pa.is_forbidden(iy) || pb.is_forbidden(iy)) continue;
                    if(!pa.map_exists(iy, ix) || !pb.map_exists(iy, ix)) continue;
                    const scalar_tr ta = pa.get_transf(iy, ix), tb = pb.get_transf(iy, ix);
                    if(ta != tb) continue;
                    pc.add_map(iy, ix, ta);
                    break;
                }
            }
            out.insert(pc);
        }
    }
}

void sum_label(const symmetry &a, const symmetry &b, symmetry &out) {
    const std::vector<std::shared_ptr<sym_element> > *sa = a.subset("label"),
        *sb = b.subset("label");
    if(!sa || !sb || sa->empty() || sb->empty()) return;
    const se_label &l0 = dynamic_cast<const se_label&>(*(*sa)[0]);
    std::vector<std::shared_ptr<const se_label> > ca, cb;
    for(int t = 0; t < 2; t++) {
        const std::vector<std::shared_ptr<sym_element> > &set = t ? *sb : *sa;
        for(size_t i = 0; i < set.size(); i++) {
            std::shared_ptr<const se_label> e =
                std::dynamic_pointer_cast<const se_label>(set[i]);
            lt_require(e != 0, LT_ERR_INTERNAL, "label subset holds another element");
            // one product table and one block labelling for both terms
            if(e->table_id != l0.table_id || e->labels != l0.labels) return;
            (t ? cb : ca).push_back(e);
        }
    }
    se_label u(l0.bidims, l0.table_id);
    u.labels = l0.labels;
    u.any_of.push_back(ca);
    u.any_of.push_back(cb);
    out.insert(u);
}

} // unnamed namespace

symmetry sum_symmetry(const symmetry &a, const symmetry &b, const perm_t *pb) {
    const perm_t id(a.order());
    const perm_t &p = pb ? *pb : id;
    bis_t bisb(b.bis);
    bisb.permute(p);
    lt_require(a.bis == bisb, LT_ERR_SYMMETRY, "sum of tensors of different block index spaces");
    symmetry out(a.bis);
    sum_perm(a, b, p, out);
    const symmetry bp = p.is_identity() ? b : so_permute(b, p);
    sum_part(a, bp, out);
    sum_label(a, bp, out);
    return out;
}

} // namespace lta
