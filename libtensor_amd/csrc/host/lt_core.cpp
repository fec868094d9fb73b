// Contraction connection algebra and matrix alignment.
// Restates libtensor/core/contraction2.h (connect / permute_a / permute_b /
// permute_c), libtensor/core/contraction2_align.h:159-320 (align) and
// libtensor/dense_tensor/impl/to_contract2_dims_impl.h.  The sequences have to
// be identical to libtensor's because the contraction-list coalescing compares
// them element by element.
#include "lt_core.h"

namespace lta {

contraction2::contraction2(size_t n, size_t m, size_t k, const perm_t *pc) :
    N(n), M(m), K(k), permc(n + m), nk(0) {

    lt_require(n + m + k <= 2 * kMaxOrder && n + k <= kMaxOrder &&
        m + k <= kMaxOrder && n + m <= kMaxOrder, LT_ERR_PARAM,
        "contraction2: order exceeds the maximum");
    for(size_t i = 0; i < 4 * kMaxOrder; i++) conn[i] = k_invalid;
    if(pc) { permc = *pc; permc.n = n + m; }
    if(K == 0) connect();
}

void contraction2::contract(size_t ia, size_t ib) {
    lt_require(!is_complete(), LT_ERR_PARAM, "contraction2: contraction is complete");
    lt_require(ia < NA(), LT_ERR_PARAM, "contraction2: index A out of bounds");
    lt_require(ib < NB(), LT_ERR_PARAM, "contraction2: index B out of bounds");
    size_t ja = NC() + ia, jb = NC() + NA() + ib;
    lt_require(conn[ja] == k_invalid, LT_ERR_PARAM, "contraction2: index ia already contracted");
    lt_require(conn[jb] == k_invalid, LT_ERR_PARAM, "contraction2: index ib already contracted");
    conn[ja] = jb;
    conn[jb] = ja;
    if(++nk == K) connect();
}

void contraction2::connect() {
    size_t nc = NC();
    if(nc == 0) return;
    size_t connc[kMaxOrder];
    size_t j = 0;
    for(size_t i = nc; i < nconn(); i++) {
        if(conn[i] == k_invalid || conn[i] < nc) connc[j++] = i;
    }
    size_t tmp[kMaxOrder];
    for(size_t i = 0; i < nc; i++) tmp[i] = connc[permc.p[i]];
    for(size_t i = 0; i < nc; i++) {
        conn[i] = tmp[i];
        conn[tmp[i]] = i;
    }
}

void contraction2::make_seqc(size_t *seqc) const {
    size_t nc = NC();
    for(size_t i = 0, j = 0; i < NA() + NB(); i++) {
        if(conn[nc + i] < nc) seqc[j++] = conn[nc + i];
    }
}

void contraction2::adjust_permc(const size_t *seqc1, const size_t *seqc2) {
    size_t nc = NC();
    perm_t pb = permutation_builder(nc, seqc1, seqc2);
    perm_t pc(permc);
    // permc := permc . permc^-1 . pb . permc  ==  pb . permc
    perm_t r(pb);
    r.permute(pc);
    permc = r;
}

void contraction2::permute_a(const perm_t &pa) {
    lt_require(is_complete(), LT_ERR_PARAM, "contraction2: incomplete");
    if(pa.is_identity()) return;
    size_t nc = NC(), na = NA();
    size_t seqc1[kMaxOrder], seqc2[kMaxOrder];
    make_seqc(seqc1);
    size_t seqa[kMaxOrder], b[kMaxOrder];
    for(size_t i = 0; i < na; i++) b[i] = conn[nc + i];
    for(size_t i = 0; i < na; i++) seqa[i] = b[pa.p[i]];
    for(size_t i = 0; i < na; i++) {
        conn[nc + i] = seqa[i];
        conn[seqa[i]] = nc + i;
    }
    make_seqc(seqc2);
    adjust_permc(seqc1, seqc2);
}

void contraction2::permute_b(const perm_t &pb) {
    lt_require(is_complete(), LT_ERR_PARAM, "contraction2: incomplete");
    if(pb.is_identity()) return;
    size_t nc = NC(), na = NA(), nb = NB();
    size_t seqc1[kMaxOrder], seqc2[kMaxOrder];
    make_seqc(seqc1);
    size_t seqb[kMaxOrder], b[kMaxOrder];
    for(size_t i = 0; i < nb; i++) b[i] = conn[nc + na + i];
    for(size_t i = 0; i < nb; i++) seqb[i] = b[pb.p[i]];
    for(size_t i = 0; i < nb; i++) {
        conn[nc + na + i] = seqb[i];
        conn[seqb[i]] = nc + na + i;
    }
    make_seqc(seqc2);
    adjust_permc(seqc1, seqc2);
}

void contraction2::permute_c(const perm_t &pc) {
    lt_require(is_complete(), LT_ERR_PARAM, "contraction2: incomplete");
    permc.permute(pc);
    connect();
}

// ---------------------------------------------------------------------------

namespace {

void align_general(const contraction2 &c, perm_t &perma, perm_t &permb,
    perm_t &permc) {

    const size_t N = c.N, M = c.M, K = c.K;
    const size_t NA = N + K, NB = M + K, NC = N + M;
    const size_t *conn = c.conn;

    // Numbering: 0..N-1 outer A, N..N+M-1 outer B, N+M.. inner.
    size_t ioa = 0, iob = N, ii = NC;
    size_t idxa1[kMaxOrder] = {0}, idxa2[kMaxOrder] = {0};
    size_t idxb1[kMaxOrder] = {0}, idxb2[kMaxOrder] = {0};
    size_t idxc1[kMaxOrder] = {0}, idxc2[kMaxOrder] = {0};

    for(size_t i = 0; i < NC; i++) {
        size_t j = conn[i] - NC;
        if(j < NA) {
            idxc1[i] = ioa; idxa1[j] = ioa; ioa++;
        } else {
            j -= NA;
            idxc1[i] = iob; idxb1[j] = iob; iob++;
        }
    }
    for(size_t i = 0; i < NA; i++) {
        if(conn[NC + i] < NC) continue;
        size_t j = conn[NC + i] - NC - NA;
        idxa1[i] = ii; idxb1[j] = ii; ii++;
    }

    size_t iai, iao, ibi, ibo, ica, icb;
    if(idxa1[NA - 1] >= NC) { iai = NA; iao = N; } else { iai = K; iao = NA; }
    if(idxb1[NB - 1] >= NC) { ibi = NB; ibo = M; } else { ibi = K; ibo = NB; }
    if(idxc1[NC - 1] < N) { ica = NC; icb = M; } else { ica = N; icb = NC; }

    for(size_t i = 0; i < NA; i++) {
        size_t v = idxa1[NA - i - 1];
        if(v >= NC) { idxa2[iai - 1] = v; iai--; }
        else { idxa2[iao - 1] = v; iao--; }
    }
    for(size_t i = 0; i < NB; i++) {
        size_t v = idxb1[NB - i - 1];
        if(v >= NC) { idxb2[ibi - 1] = v; ibi--; }
        else { idxb2[ibo - 1] = v; ibo--; }
    }
    for(size_t i = 0; i < NC; i++) {
        size_t v = idxc1[NC - i - 1];
        if(v < N) { idxc2[ica - 1] = v; ica--; }
        else { idxc2[icb - 1] = v; icb--; }
    }

    bool lasta_i = (idxa2[NA - 1] >= NC);
    bool lastb_i = (idxb2[NB - 1] >= NC);
    bool lastc_a = (idxc2[NC - 1] < N);

    if(lastc_a) {
        if(lasta_i) {
            if(lastb_i) {   // C(ji) = A(ik) B(jk)
                for(size_t i = 0; i != N; i++) idxa2[i] = idxc2[M + i];
                for(size_t i = 0; i != M; i++) idxc2[i] = idxb2[i];
                for(size_t i = 0; i != K; i++) idxa2[N + i] = idxb2[M + i];
            } else {        // C(ji) = A(ik) B(kj)
                for(size_t i = 0; i != N; i++) idxa2[i] = idxc2[M + i];
                for(size_t i = 0; i != M; i++) idxc2[i] = idxb2[K + i];
                for(size_t i = 0; i != K; i++) idxb2[i] = idxa2[N + i];
            }
        } else {
            if(lastb_i) {   // C(ji) = A(ki) B(jk)
                for(size_t i = 0; i != N; i++) idxa2[K + i] = idxc2[M + i];
                for(size_t i = 0; i != M; i++) idxc2[i] = idxb2[i];
                for(size_t i = 0; i != K; i++) idxa2[i] = idxb2[M + i];
            } else {        // C(ji) = A(ki) B(kj)
                for(size_t i = 0; i != N; i++) idxa2[K + i] = idxc2[M + i];
                for(size_t i = 0; i != M; i++) idxc2[i] = idxb2[K + i];
                for(size_t i = 0; i != K; i++) idxb2[i] = idxa2[i];
            }
        }
    } else {
        if(lasta_i) {
            if(lastb_i) {   // C(ij) = A(ik) B(jk)
                for(size_t i = 0; i != N; i++) idxa2[i] = idxc2[i];
                for(size_t i = 0; i != M; i++) idxb2[i] = idxc2[N + i];
                for(size_t i = 0; i != K; i++) idxa2[N + i] = idxb2[M + i];
            } else {        // C(ij) = A(ik) B(kj)
                for(size_t i = 0; i != N; i++) idxc2[i] = idxa2[i];
                for(size_t i = 0; i != M; i++) idxb2[K + i] = idxc2[N + i];
                for(size_t i = 0; i != K; i++) idxb2[i] = idxa2[N + i];
            }
        } else {
            if(lastb_i) {   // C(ij) = A(ki) B(jk)
                for(size_t i = 0; i != N; i++) idxc2[i] = idxa2[K + i];
                for(size_t i = 0; i != M; i++) idxb2[i] = idxc2[N + i];
                for(size_t i = 0; i != K; i++) idxa2[i] = idxb2[M + i];
            } else {        // C(ij) = A(ki) B(kj)
                for(size_t i = 0; i != N; i++) idxc2[i] = idxa2[K + i];
                for(size_t i = 0; i != M; i++) idxc2[N + i] = idxb2[K + i];
                for(size_t i = 0; i != K; i++) idxb2[i] = idxa2[i];
            }
        }
    }

    perma = permutation_builder(NA, idxa2, idxa1);
    permb = permutation_builder(NB, idxb2, idxb1);
    permc = permutation_builder(NC, idxc2, idxc1);
}

// N == 0: only B is permuted (contraction2_align.h, <0, M, K> specialization).
void align_n0(const contraction2 &c, perm_t &permb) {
    const size_t M = c.M, K = c.K;
    const size_t NA = K, NB = M + K, NC = M;
    const size_t *conn = c.conn;
    size_t idxa[kMaxOrder] = {0}, idxb1[kMaxOrder] = {0}, idxb2[kMaxOrder] = {0};
    for(size_t i = 0; i < NC; i++) {
        size_t j = conn[i] - NC - NA;
        idxb1[j] = i;
    }
    for(size_t i = 0, ii = NC; i < NA; i++, ii++) {
        size_t j = conn[NC + i] - NC - NA;
        idxa[i] = ii;
        idxb1[j] = ii;
    }
    size_t ibi, ibo;
    if(idxb1[NB - 1] >= NC) { ibi = NB; ibo = M; } else { ibi = K; ibo = NB; }
    for(size_t i = 0; i < NB; i++) {
        size_t v = idxb1[NB - i - 1];
        if(v >= NC) { idxb2[ibi - 1] = v; ibi--; }
        else { idxb2[ibo - 1] = v; ibo--; }
    }
    bool lastb_i = (idxb2[NB - 1] >= NC);
    if(lastb_i) {
        for(size_t i = 0; i != M; i++) idxb2[i] = i;
        for(size_t i = 0; i != K; i++) idxb2[M + i] = idxa[i];
    } else {
        for(size_t i = 0; i != M; i++) idxb2[K + i] = i;
        for(size_t i = 0; i != K; i++) idxb2[i] = idxa[i];
    }
    permb = permutation_builder(NB, idxb2, idxb1);
}

// M == 0: only A is permuted (<N, 0, K> specialization).
void align_m0(const contraction2 &c, perm_t &perma) {
    const size_t N = c.N, K = c.K;
    const size_t NA = N + K, NC = N;
    const size_t *conn = c.conn;
    size_t ii = NC;
    size_t idxa1[kMaxOrder] = {0}, idxa2[kMaxOrder] = {0}, idxb[kMaxOrder] = {0};
    for(size_t i = 0; i < NC; i++) {
        size_t j = conn[i] - NC;
        idxa1[j] = i;
    }
    for(size_t i = 0; i < NA; i++) {
        if(conn[NC + i] < NC) continue;
        size_t j = conn[NC + i] - NC - NA;
        idxa1[i] = ii;
        idxb[j] = ii;
        ii++;
    }
    size_t iai, iao;
    if(idxa1[NA - 1] >= NC) { iai = NA; iao = N; } else { iai = K; iao = NA; }
    for(size_t i = 0; i < NA; i++) {
        size_t v = idxa1[NA - i - 1];
        if(v >= NC) { idxa2[iai - 1] = v; iai--; }
        else { idxa2[iao - 1] = v; iao--; }
    }
    bool lasta_i = (idxa2[NA - 1] >= NC);
    if(lasta_i) {
        for(size_t i = 0; i != N; i++) idxa2[i] = i;
        for(size_t i = 0; i != K; i++) idxa2[N + i] = idxb[i];
    } else {
        for(size_t i = 0; i != N; i++) idxa2[K + i] = i;
        for(size_t i = 0; i != K; i++) idxa2[i] = idxb[i];
    }
    perma = permutation_builder(NA, idxa2, idxa1);
}

} // unnamed namespace

contraction2_align::contraction2_align(const contraction2 &c) :
    perma(c.NA()), permb(c.NB()), permc(c.NC()) {

    lt_require(c.is_complete(), LT_ERR_PARAM, "contraction2_align: incomplete");
    lt_require(c.NC() > 0, LT_ERR_PARAM, "contraction2_align: result of order 0");
    if(c.N == 0) align_n0(c, permb);
    else if(c.M == 0) align_m0(c, perma);
    else align_general(c, perma, permb, permc);
}

dims_t contract2_dims(const contraction2 &c, const dims_t &da, const dims_t &db) {
    lt_require(c.is_complete(), LT_ERR_PARAM, "contract2_dims: incomplete");
    lt_require(da.n == c.NA() && db.n == c.NB(), LT_ERR_DIMS,
        "contract2_dims: operand orders do not match the contraction");
    const size_t NC = c.NC(), NA = c.NA();
    for(size_t i = 0; i < NA; i++) {
        if(c.conn[NC + i] >= NC + NA) {
            size_t j = c.conn[NC + i] - NC - NA;
            lt_require(da[i] == db[j], LT_ERR_DIMS,
                "contract2_dims: contracted dimensions differ");
        }
    }
    dims_t dc(NC);
    for(size_t i = 0; i < NC; i++) {
        size_t j = c.conn[i] - NC;
        dc[i] = (j < NA) ? da[j] : db[j - NA];
    }
    return dc;
}

bis_t contract2_bis(const contraction2 &c, const bis_t &bisa, const bis_t &bisb) {
    dims_t dc = contract2_dims(c, bisa.dims, bisb.dims);
    const size_t NC = c.NC(), NA = c.NA(), NB = c.NB();
    for(size_t i = 0; i < NA; i++) {
        if(c.conn[NC + i] >= NC + NA) {
            size_t j = c.conn[NC + i] - NC - NA;
            lt_require(bisa.splits[i] == bisb.splits[j], LT_ERR_DIMS,
                "contract2_bis: contracted dimensions are split differently");
        }
    }
    (void)NB;
    bis_t bisc(dc);
    for(size_t i = 0; i < NC; i++) {
        size_t j = c.conn[i] - NC;
        bisc.splits[i] = (j < NA) ? bisa.splits[j] : bisb.splits[j - NA];
    }
    return bisc;
}

gemm_form aligned_gemm_form(const contraction2 &c, const dims_t &da,
    const dims_t &db) {

    const size_t N = c.N, M = c.M, K = c.K;
    const size_t NA = N + K, NB = M + K, NC = N + M;
    gemm_form g;
    g.m = 1; g.n = 1; g.k = 1;
    for(size_t i = 0; i < NA; i++) {
        if(c.conn[NC + i] < NC) g.m *= da[i]; else g.k *= da[i];
    }
    for(size_t i = 0; i < NB; i++) {
        if(c.conn[NC + NA + i] < NC) g.n *= db[i];
    }
    // Aligned A is (outer..., inner...) or (inner..., outer...).
    g.ta = (NA > 0 && N > 0 && K > 0 && c.conn[NC + NA - 1] < NC);
    g.tb = (NB > 0 && M > 0 && K > 0 && c.conn[NC + NA + NB - 1] >= NC);
    // C is (A outer..., B outer...) or (B outer..., A outer...).
    g.tc = (N > 0 && M > 0 && (c.conn[NC - 1] - NC) < NA);
    return g;
}

} // namespace lta
