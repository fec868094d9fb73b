// Core index algebra for the block-tensor contraction host side: permutations,
// dimensions, absolute indexes, scalar / tensor transformations, block index
// spaces and contraction connection sequences.  Tensor orders are runtime
// values (at most LT_MAX_ORDER); every convention follows libtensor so that
// schedules built here are bit-identical to libtensor's:
//   * row-major, last index fastest (libtensor/core/impl/dimensions_impl.h:58-67)
//   * permutation p: p[i] is the source position of output position i, and
//     apply(seq): seq'[i] = seq[p[i]] (libtensor/core/permutation.h)
//   * a.permute(b) composes "a then b": (a.b)[i] = a[b[i]]
#pragma once

#include <algorithm>
#include <array>
#include <cstddef>
#include <cstdint>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

#define LT_MAX_ORDER 8

namespace lta {

//! Storage capacity of index-like objects.  Tensors have at most LT_MAX_ORDER
//! dimensions; the direct-product space used to derive the symmetry of a
//! contraction result has up to twice as many.
static const size_t kMaxOrder = 2 * LT_MAX_ORDER;

//! Error raised by the host side; the C ABI converts it into an error code.
struct lt_error : public std::runtime_error {
    int code;
    lt_error(int c, const std::string &m) : std::runtime_error(m), code(c) { }
};

// Error codes shared with the C ABI (include/lt_amd.h).
#ifndef LT_OK
#define LT_OK             0
#define LT_ERR_PARAM     -1   // bad_parameter / out_of_bounds
#define LT_ERR_DIMS      -2   // bad_dimensions
#define LT_ERR_SYMMETRY  -3   // bad_symmetry
#define LT_ERR_DEVICE    -4   // HIP runtime failure
#define LT_ERR_NOMEM     -5   // admission / out of memory
#define LT_ERR_STATE     -6   // call not valid in the current state
#define LT_ERR_INTERNAL  -7
#endif

inline void lt_require(bool ok, int code, const char *msg) {
    if(!ok) throw lt_error(code, msg);
}

//! Fixed-capacity sequence of size_t with a runtime length.
struct seq_t {
    size_t n;
    size_t v[kMaxOrder];
    seq_t() : n(0) { for(size_t i = 0; i < kMaxOrder; i++) v[i] = 0; }
    explicit seq_t(size_t n_, size_t fill = 0) : n(n_) {
        for(size_t i = 0; i < kMaxOrder; i++) v[i] = fill;
    }
    size_t &operator[](size_t i) { return v[i]; }
    const size_t &operator[](size_t i) const { return v[i]; }
    bool operator==(const seq_t &o) const {
        if(n != o.n) return false;
        for(size_t i = 0; i < n; i++) if(v[i] != o.v[i]) return false;
        return true;
    }
    bool operator!=(const seq_t &o) const { return !(*this == o); }
    //! Lexicographic order (libtensor/core/index.h operator<).
    bool operator<(const seq_t &o) const {
        for(size_t i = 0; i < n; i++) {
            if(v[i] < o.v[i]) return true;
            if(v[i] > o.v[i]) return false;
        }
        return false;
    }
};
typedef seq_t index_t;

//! Permutation of tensor indexes (libtensor/core/permutation.h).
struct perm_t {
    size_t n;
    uint8_t p[kMaxOrder];

    perm_t() : n(0) { for(size_t i = 0; i < kMaxOrder; i++) p[i] = (uint8_t)i; }
    explicit perm_t(size_t n_) : n(n_) {
        for(size_t i = 0; i < kMaxOrder; i++) p[i] = (uint8_t)i;
    }
    static perm_t from(size_t n, const size_t *src) {
        perm_t r(n);
        for(size_t i = 0; i < n; i++) r.p[i] = (uint8_t)src[i];
        return r;
    }
    size_t operator[](size_t i) const { return p[i]; }
    bool is_identity() const {
        for(size_t i = 0; i < n; i++) if(p[i] != i) return false;
        return true;
    }
    void reset() { for(size_t i = 0; i < kMaxOrder; i++) p[i] = (uint8_t)i; }
    //! this := this then q.
    perm_t &permute(const perm_t &q) {
        uint8_t c[kMaxOrder];
        std::memcpy(c, p, sizeof(c));
        for(size_t i = 0; i < n; i++) p[i] = c[q.p[i]];
        return *this;
    }
    //! Transposition of positions i and j.
    perm_t &swap(size_t i, size_t j) {
        if(i != j) std::swap(p[i], p[j]);
        return *this;
    }
    perm_t &invert() {
        uint8_t c[kMaxOrder];
        std::memcpy(c, p, sizeof(c));
        for(size_t i = 0; i < n; i++) p[c[i]] = (uint8_t)i;
        return *this;
    }
    perm_t inverse() const { perm_t r(*this); r.invert(); return r; }
    template<typename S>
    void apply(S &s) const {
        S b(s);
        for(size_t i = 0; i < n; i++) s[i] = b[p[i]];
    }
    bool equals(const perm_t &q) const {
        for(size_t i = 0; i < n; i++) if(p[i] != q.p[i]) return false;
        return true;
    }
    bool operator==(const perm_t &q) const { return equals(q); }
    bool operator!=(const perm_t &q) const { return !equals(q); }
    bool operator<(const perm_t &q) const {
        for(size_t i = 0; i < n; i++)
            if(p[i] != q.p[i]) return p[i] < q.p[i];
        return false;
    }
};

//! Builds the permutation that turns seq2 into seq1, applying elementary
//! transpositions through an optional position map in exactly the order of
//! libtensor/core/permutation_builder.h (build()).  With an identity map the
//! result is the unique p with seq1[i] == seq2[p[i]].
template<typename T>
perm_t permutation_builder(size_t n, const T *seq1, const T *seq2,
    const perm_t *pmap = 0) {

    size_t map[kMaxOrder];
    for(size_t i = 0; i < n; i++) map[i] = i;
    if(pmap) {
        perm_t pinv(pmap->inverse());
        size_t tmp[kMaxOrder];
        for(size_t i = 0; i < n; i++) tmp[i] = map[pinv.p[i]];
        for(size_t i = 0; i < n; i++) map[i] = tmp[i];
    }
    size_t idx[kMaxOrder];
    for(size_t i = 0; i < n; i++) {
        for(size_t j = i + 1; j < n; j++)
            lt_require(!(seq1[i] == seq1[j]), LT_ERR_PARAM,
                "permutation_builder: repeated element in seq1");
        size_t j = 0;
        for(; j < n; j++) if(seq2[j] == seq1[i]) { idx[i] = j; break; }
        lt_require(j != n, LT_ERR_PARAM,
            "permutation_builder: sequences differ");
    }
    perm_t perm(n);
    size_t i = 0;
    while(i < n) {
        if(i > idx[i]) {
            perm.swap(map[i], map[idx[i]]);
            size_t j = idx[i];
            idx[i] = idx[j];
            idx[j] = j;
            i = 0;
        } else {
            i++;
        }
    }
    // the transpositions above build the inverse
    perm.invert();
    return perm;
}

//! Tensor dimensions (libtensor/core/dimensions.h).
struct dims_t {
    size_t n;
    size_t d[kMaxOrder];
    dims_t() : n(0) { for(size_t i = 0; i < kMaxOrder; i++) d[i] = 1; }
    explicit dims_t(size_t n_) : n(n_) {
        for(size_t i = 0; i < kMaxOrder; i++) d[i] = 1;
    }
    static dims_t from(size_t n, const size_t *src) {
        dims_t r(n);
        for(size_t i = 0; i < n; i++) r.d[i] = src[i];
        return r;
    }
    size_t operator[](size_t i) const { return d[i]; }
    size_t &operator[](size_t i) { return d[i]; }
    size_t size() const {
        size_t s = 1;
        for(size_t i = 0; i < n; i++) s *= d[i];
        return s;
    }
    //! Increment (stride) of dimension i.
    size_t inc(size_t i) const {
        size_t s = 1;
        for(size_t j = i + 1; j < n; j++) s *= d[j];
        return s;
    }
    void permute(const perm_t &p) {
        size_t b[kMaxOrder];
        for(size_t i = 0; i < n; i++) b[i] = d[i];
        for(size_t i = 0; i < n; i++) d[i] = b[p.p[i]];
    }
    bool operator==(const dims_t &o) const {
        if(n != o.n) return false;
        for(size_t i = 0; i < n; i++) if(d[i] != o.d[i]) return false;
        return true;
    }
    bool operator!=(const dims_t &o) const { return !(*this == o); }
};

inline size_t abs_index(const index_t &idx, const dims_t &dims) {
    size_t a = 0;
    for(size_t i = 0; i < dims.n; i++) a = a * dims.d[i] + idx[i];
    return a;
}

inline index_t get_index(size_t aidx, const dims_t &dims) {
    index_t idx(dims.n);
    if(dims.n == 0) return idx;
    for(size_t i = dims.n; i-- > 0;) {
        idx[i] = aidx % dims.d[i];
        aidx /= dims.d[i];
    }
    return idx;
}

//! Odometer increment; false once the last index is passed.
inline bool inc_index(index_t &idx, const dims_t &dims) {
    for(size_t i = dims.n; i-- > 0;) {
        if(idx[i] + 1 < dims.d[i]) {
            idx[i]++;
            for(size_t j = i + 1; j < dims.n; j++) idx[j] = 0;
            return true;
        }
    }
    return false;
}

//! Scalar transformation of tensor elements (multiplication by a coefficient,
//! libtensor/core/scalar_transf_double.h).
struct scalar_tr {
    double c;
    scalar_tr() : c(1.0) { }
    explicit scalar_tr(double c_) : c(c_) { }
    void reset() { c = 1.0; }
    scalar_tr &transform(const scalar_tr &o) { c *= o.c; return *this; }
    scalar_tr &invert() { c = 1.0 / c; return *this; }
    scalar_tr &scale(double s) { c *= s; return *this; }
    bool is_identity() const { return c == 1.0; }
    bool is_zero() const { return c == 0.0; }
    bool operator==(const scalar_tr &o) const { return c == o.c; }
    bool operator!=(const scalar_tr &o) const { return c != o.c; }
};

//! Permutation plus scalar (libtensor/core/tensor_transf.h).
struct tensor_tr {
    perm_t perm;
    scalar_tr st;
    tensor_tr() { }
    explicit tensor_tr(size_t n) : perm(n) { }
    tensor_tr(const perm_t &p, const scalar_tr &s) : perm(p), st(s) { }
    tensor_tr &transform(const tensor_tr &o) {
        st.transform(o.st);
        perm.permute(o.perm);
        return *this;
    }
    tensor_tr &invert() { st.invert(); perm.invert(); return *this; }
    tensor_tr inverse() const { tensor_tr r(*this); r.invert(); return r; }
    bool is_identity() const { return st.is_identity() && perm.is_identity(); }
    bool operator==(const tensor_tr &o) const {
        return st == o.st && perm == o.perm;
    }
};

//! Block index space: total dimensions plus split points per dimension.
//! Dimensions with identical extent and split points share a type
//! (the result of libtensor's match_splits(), core/block_index_space.h).
struct bis_t {
    dims_t dims;
    std::vector<size_t> splits[kMaxOrder];

    bis_t() { }
    explicit bis_t(const dims_t &d) : dims(d) { }

    size_t order() const { return dims.n; }
    void split(size_t dim, size_t pos) {
        lt_require(dim < dims.n, LT_ERR_PARAM, "split: dimension out of range");
        lt_require(pos < dims.d[dim], LT_ERR_PARAM, "split: position out of range");
        if(pos == 0) return;
        std::vector<size_t> &s = splits[dim];
        std::vector<size_t>::iterator it = std::lower_bound(s.begin(), s.end(), pos);
        if(it == s.end() || *it != pos) s.insert(it, pos);
    }
    dims_t block_index_dims() const {
        dims_t r(dims.n);
        for(size_t i = 0; i < dims.n; i++) r.d[i] = splits[i].size() + 1;
        return r;
    }
    size_t block_start(size_t dim, size_t b) const {
        return b == 0 ? 0 : splits[dim][b - 1];
    }
    size_t block_len(size_t dim, size_t b) const {
        size_t e = (b == splits[dim].size()) ? dims.d[dim] : splits[dim][b];
        return e - block_start(dim, b);
    }
    dims_t block_dims(const index_t &bidx) const {
        dims_t r(dims.n);
        for(size_t i = 0; i < dims.n; i++) r.d[i] = block_len(i, bidx[i]);
        return r;
    }
    index_t block_start(const index_t &bidx) const {
        index_t r(dims.n);
        for(size_t i = 0; i < dims.n; i++) r[i] = block_start(i, bidx[i]);
        return r;
    }
    //! Type of dimension i: index of the first dimension with identical
    //! extent and splits.
    size_t type(size_t i) const {
        for(size_t j = 0; j < i; j++)
            if(dims.d[j] == dims.d[i] && splits[j] == splits[i]) return type(j);
        return i;
    }
    void permute(const perm_t &p) {
        dims.permute(p);
        std::vector<size_t> s[kMaxOrder];
        for(size_t i = 0; i < dims.n; i++) s[i] = splits[p.p[i]];
        for(size_t i = 0; i < dims.n; i++) splits[i].swap(s[i]);
    }
    bool operator==(const bis_t &o) const {
        if(dims != o.dims) return false;
        for(size_t i = 0; i < dims.n; i++) if(splits[i] != o.splits[i]) return false;
        return true;
    }
    bool operator!=(const bis_t &o) const { return !(*this == o); }
};

//! Contraction connection sequence (libtensor/core/contraction2.h).
//! Positions 0..NC-1 are C, NC..NC+NA-1 are A, NC+NA.. are B; conn[x] is
//! the position x is connected to.
struct contraction2 {
    size_t N, M, K;
    perm_t permc;
    size_t nk;          //!< Contracted pairs specified so far
    size_t conn[4 * kMaxOrder];

    static const size_t k_invalid = size_t(-1);

    contraction2() : N(0), M(0), K(0), nk(0) {
        for(size_t i = 0; i < 4 * kMaxOrder; i++) conn[i] = k_invalid;
    }
    contraction2(size_t n, size_t m, size_t k, const perm_t *pc = 0);

    size_t NA() const { return N + K; }
    size_t NB() const { return M + K; }
    size_t NC() const { return N + M; }
    size_t nconn() const { return 2 * (N + M + K); }
    bool is_complete() const { return nk == K; }

    void contract(size_t ia, size_t ib);
    void permute_a(const perm_t &pa);
    void permute_b(const perm_t &pb);
    void permute_c(const perm_t &pc);
    bool same_conn(const contraction2 &o) const {
        for(size_t i = 0; i < nconn(); i++) if(conn[i] != o.conn[i]) return false;
        return true;
    }

private:
    void connect();
    void make_seqc(size_t *seqc) const;
    void adjust_permc(const size_t *seqc1, const size_t *seqc2);
};

//! Permutations of A, B and C that turn a contraction into one matrix
//! product (libtensor/core/contraction2_align.h:159-320).
struct contraction2_align {
    perm_t perma, permb, permc;
    explicit contraction2_align(const contraction2 &c);
};

//! Dimensions of C (libtensor/dense_tensor/to_contract2_dims.h).
dims_t contract2_dims(const contraction2 &c, const dims_t &da, const dims_t &db);

//! Block index space of C (libtensor/gen_block_tensor/impl/
//! gen_bto_contract2_bis_impl.h): every C dimension takes the splits of the
//! A or B dimension it comes from.
bis_t contract2_bis(const contraction2 &c, const bis_t &bisa, const bis_t &bisb);

//! Matrix form of an aligned contraction: C(m,n) [+]= d * op(A) op(B).
struct gemm_form {
    bool ta;        //!< A stored K x M ("ki") instead of M x K ("ik")
    bool tb;        //!< B stored N x K ("jk") instead of K x N ("kj")
    bool tc;        //!< C stored N x M ("ji"): swap operands
    size_t m, n, k;
};

//! For a contraction whose operands are already in matricized order
//! (all three align permutations are identity) computes the GEMM form.
gemm_form aligned_gemm_form(const contraction2 &c, const dims_t &da,
    const dims_t &db);

} // namespace lta
