// Block-tensor symmetry: elements (permutational, partition, label),
// symmetry containers, orbits and the symmetry operations used by the
// contraction path (permute, direct product, reduction).
//
// The orbit traversal, the partition-map bookkeeping and the permutation
// group (Jerrum branching) follow libtensor step by step, because the
// transformations they produce decide which contraction pairs coalesce and
// therefore the integer contraction lists:
//   libtensor/core/impl/orbit_impl.h:209-318, short_orbit_impl.h:93-150
//   libtensor/symmetry/se_perm.h, inst/se_perm_impl.h
//   libtensor/symmetry/se_part.h, inst/se_part_impl.h
//   libtensor/symmetry/se_label.h, inst/se_label_impl.h, point_group_table.C
//   libtensor/symmetry/permutation_group.h, inst/permutation_group_impl.h
//   libtensor/symmetry/inst/so_{permute,dirprod,reduce}_se_{perm,part}_impl.h
//   libtensor/symmetry/inst/combine_part_impl.h
// Label reduction (so_reduce of se_label) is computed by enumerating the
// labels of the contracted dimensions instead of libtensor's rule algebra
// (inst/er_reduce_impl.h); it yields the exact set of allowed blocks.
#pragma once

#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <unordered_map>
#include <vector>
#include "lt_core.h"

namespace lta {

//! Product table of a point group (libtensor/symmetry/point_group_table.C):
//! label 0 is the identity irrep; m_table[pair(l1,l2)] is a bitmask of irreps.
struct product_table {
    std::string id;
    size_t nlabels;
    std::vector<uint64_t> table;   // nlabels*nlabels, symmetric

    product_table() : nlabels(0) { }
    product_table(const std::string &id_, size_t n);
    void add_product(size_t l1, size_t l2, size_t lr);
    //! True if label l is contained in the product of the labels in lg.
    bool is_in_product(const size_t *lg, size_t nlg, size_t l) const;
};

//! Registry of product tables, keyed by id (product_table_container).
class product_table_registry {
public:
    static product_table_registry &instance();
    void add(const product_table &t);
    const product_table &get(const std::string &id) const;
    bool has(const std::string &id) const;
private:
    mutable std::mutex m_mtx;
    std::map<std::string, std::shared_ptr<product_table> > m_tables;
};

static const size_t k_label_invalid = size_t(-1);

//! Base class of symmetry elements (libtensor/core/symmetry_element_i.h).
struct sym_element {
    virtual ~sym_element() { }
    virtual const char *type() const = 0;
    virtual sym_element *clone() const = 0;
    virtual bool is_allowed(const index_t &idx) const = 0;
    virtual void apply(index_t &idx) const = 0;
    virtual void apply(index_t &idx, tensor_tr &tr) const = 0;
    virtual void permute(const perm_t &p) = 0;
};

//! Permutational symmetry element: block B(P idx) = s * P(B(idx)).
struct se_perm : public sym_element {
    tensor_tr tr;
    size_t orderp, orderc;
    se_perm(const perm_t &p, const scalar_tr &s);
    const char *type() const { return "perm"; }
    sym_element *clone() const { return new se_perm(*this); }
    bool is_allowed(const index_t &) const { return true; }
    void apply(index_t &idx) const { tr.perm.apply(idx); }
    void apply(index_t &idx, tensor_tr &t) const {
        tr.perm.apply(idx);
        t.transform(tr);
    }
    void permute(const perm_t &) { throw lt_error(LT_ERR_INTERNAL, "se_perm::permute"); }
};

//! Partition symmetry element: maps between partitions of blocks.
struct se_part : public sym_element {
    bis_t bis;
    dims_t bidims, pdims, bipdims;
    std::vector<size_t> fmap, rmap;
    std::vector<index_t> fmapi;
    std::vector<scalar_tr> ftr;

    se_part(const bis_t &b, const dims_t &pd);
    se_part(const bis_t &b, const bool *msk, size_t npart);
    const char *type() const { return "part"; }
    sym_element *clone() const { return new se_part(*this); }

    void add_map(const index_t &i1, const index_t &i2,
        const scalar_tr &tr = scalar_tr());
    void mark_forbidden(const index_t &i);
    bool is_forbidden(const index_t &i) const {
        return fmap[abs_index(i, pdims)] == size_t(-1);
    }
    const index_t &get_direct_map(const index_t &i) const;
    scalar_tr get_transf(const index_t &from, const index_t &to) const;
    bool map_exists(const index_t &from, const index_t &to) const;

    bool is_allowed(const index_t &idx) const;
    void apply(index_t &idx) const;
    void apply(index_t &idx, tensor_tr &t) const;
    void permute(const perm_t &p);

private:
    void init_maps();
    void add_to_loop(size_t a, size_t b, const scalar_tr &tr);
};

//! One term of a product rule: labels of the dimensions with non-zero
//! multiplicity seq[i] are multiplied; intrinsic must be in the product.
struct label_term {
    seq_t seq;
    size_t intrinsic;
};
typedef std::vector<label_term> label_product;   // conjunction
typedef std::vector<label_product> label_rule;   // disjunction

//! Label symmetry element (libtensor/symmetry/se_label.h): a block is allowed
//! if the evaluation rule holds for the labels of its block indexes.
//! An element may instead be the reduction of a larger element over
//! contracted dimensions; it then allows a block if some assignment of
//! labels to the contracted dimensions satisfies the larger rule.
struct se_label : public sym_element {
    std::string table_id;
    dims_t bidims;
    std::vector<std::vector<size_t> > labels;   //!< per dimension, per block
    label_rule rule;

    // Reduced form
    std::shared_ptr<const se_label> src;   //!< element over the full space
    std::vector<size_t> keep;              //!< dimension of src for each own dim
    std::vector<std::vector<size_t> > rdims;  //!< src dims of each reduction step
    std::vector<std::vector<size_t> > rlabels; //!< candidate labels per step

    //! Symmetrized form: the element also allows a block whose index,
    //! permuted by one of these permutations (idx'[i] = idx[q[i]]), the rule
    //! allows (the union of the rule and its permuted copies that
    //! so_symmetrize builds).
    std::vector<perm_t> alts;

    //! Product form (the direct product of elements that are reduced,
    //! symmetrized or products themselves): a block is allowed if every
    //! factor allows the labels of its dimensions (factor dimension k is
    //! dimension map[k] of this element).
    struct factor {
        std::shared_ptr<const se_label> el;
        std::vector<size_t> map;
    };
    std::vector<factor> factors;

    //! Union form (the symmetry of a sum): a block is allowed if, for one
    //! of the lists, every element of that list allows it.  The elements
    //! share this element's dimensions and block labels.
    std::vector<std::vector<std::shared_ptr<const se_label> > > any_of;

    se_label(const dims_t &bid, const std::string &tid);
    const char *type() const { return "label"; }
    sym_element *clone() const { return new se_label(*this); }
    void set_label(size_t dim, size_t blk, size_t label);
    void set_rule(size_t intrinsic);
    void set_rule_set(const std::vector<size_t> &intr);

    bool eval_labels(const size_t *lab) const;   //!< labels of a full index
    bool is_allowed(const index_t &idx) const;
    void apply(index_t &) const { }
    void apply(index_t &, tensor_tr &) const { }
    void permute(const perm_t &p);

private:
    bool eval_one(const size_t *lab) const;
    bool eval_rule(const size_t *lab) const;
    bool eval_reduced(const size_t *ownlab) const;
    mutable std::mutex m_memo_mtx;
    mutable std::unordered_map<uint64_t, bool> m_memo;
public:
    se_label(const se_label &o);
};

//! Ordered list of symmetry subsets (libtensor/core/symmetry.h): subsets are
//! kept in order of first insertion of their type, elements in insertion
//! order.
struct symmetry {
    bis_t bis;
    std::vector<std::pair<std::string,
        std::vector<std::shared_ptr<sym_element> > > > subsets;

    symmetry() { }
    explicit symmetry(const bis_t &b) : bis(b) { }
    void insert(const sym_element &e);
    void insert_shared(std::shared_ptr<sym_element> e);
    void clear() { subsets.clear(); }
    size_t order() const { return bis.order(); }
    template<typename F> void for_each(F f) const {
        for(size_t i = 0; i < subsets.size(); i++)
            for(size_t j = 0; j < subsets[i].second.size(); j++)
                f(*subsets[i].second[j]);
    }
    const std::vector<std::shared_ptr<sym_element> > *subset(
        const std::string &t) const {
        for(size_t i = 0; i < subsets.size(); i++)
            if(subsets[i].first == t) return &subsets[i].second;
        return 0;
    }
};

//! Orbit of a block with transformations from the canonical block.
struct orbit {
    dims_t dims;
    std::vector<std::pair<size_t, size_t> > orb;   //!< (abs index, tr slot), sorted
    std::vector<tensor_tr> tr;
    size_t acidx;
    bool allowed;

    orbit(const symmetry &sym, size_t aidx, bool compute_allowed = true);
    const tensor_tr &get_transf(size_t aidx) const;
    bool contains(size_t aidx) const;
    size_t size() const { return orb.size(); }
};

//! Canonical index and allowedness only (libtensor/core/short_orbit.h).
struct short_orbit {
    size_t acidx;
    bool allowed;
    short_orbit(const symmetry &sym, size_t aidx, bool compute_allowed = true);
};

//! Permutation group with scalar transformations (Jerrum's branching).
class permutation_group {
public:
    typedef std::pair<perm_t, scalar_tr> gen_t;
    explicit permutation_group(size_t n);
    permutation_group(size_t n, const std::vector<std::shared_ptr<sym_element> > *set);
    void add_orbit(const scalar_tr &tr, const perm_t &p);
    bool is_member(const scalar_tr &tr, const perm_t &p) const;
    void convert(std::vector<gen_t> &gs) const;
    void stabilize(const seq_t &seq, permutation_group &g2) const;
    void permute(const perm_t &p);

private:
    struct branching {
        size_t n;
        gen_t sigma[kMaxOrder];
        gen_t tau[kMaxOrder];
        size_t edges[kMaxOrder];
        explicit branching(size_t n_);
        void reset();
    };
    size_t m_n;
    branching m_br;

    size_t get_path(const branching &br, size_t i, size_t j, size_t *path) const;
    bool is_member(const branching &br, size_t i, scalar_tr &tr, const perm_t &p) const;
    void make_branching(branching &br, size_t i, const std::vector<gen_t> &gs,
        std::vector<gen_t> &gs2) const;
    void make_genset(const branching &br, std::vector<gen_t> &gs) const;
    void permute_branching(branching &br, const perm_t &p) const;
    void make_setstabilizer(const branching &br, const seq_t &msk,
        std::vector<gen_t> &gs) const;
};

//! Symmetry of P(tensor) (libtensor/symmetry/so_permute.h).
symmetry so_permute(const symmetry &sym, const perm_t &p);

//! Symmetry of X + s P(X) for a pair permutation P (P P = 1, P != 1):
//! libtensor/symmetry/so_symmetrize.h with two index groups, as
//! gen_bto_symmetrize2::make_symmetry calls it
//! (gen_block_tensor/impl/gen_bto_symmetrize2_impl.h:131-149).
symmetry so_symmetrize_pair(const symmetry &sym, const perm_t &perm, double s);

//! Symmetry of the sum over the six exchanges of three index groups, the
//! pair exchanges weighted by s (gen_bto_symmetrize3::make_symmetry,
//! gen_block_tensor/impl/gen_bto_symmetrize3_impl.h:132-160): p1 exchanges
//! groups 1 and 2... as libtensor numbers them, p2 the other pair.
symmetry so_symmetrize_triple(const symmetry &sym, const perm_t &p1, const perm_t &p2,
    double s);

//! Symmetry of a sum of two tensors on one block index space, the symmetry
//! both terms have (libtensor forms it as so_merge(so_dirsum(a, b)) over the
//! diagonal: gen_block_tensor/impl/gen_bto_aux_add_impl.h:40-66,
//! addition_schedule_impl.h:27-45, gen_bto_add_impl.h:240-267):
//!  - perm: the group libtensor's so_dirsum builds from the two generating
//!    sets, restricted to the diagonal (it can be smaller than the
//!    intersection of the two groups);
//!  - part: for elements of equal partition dimensions, the maps both have
//!    with the same scalar; a partition is forbidden only if both forbid it;
//!  - label: for elements of one product table and equal block labels, a
//!    block is allowed if either term's label elements allow it.
//! Elements with no counterpart of the same kind in the other term drop out.
//! pb (optional): b enters permuted by pb (btod_add's operand permutation).
//! The order of a and b matters for the generators of the result: a is the
//! operation's result (or the sum so far), b the target (or the next
//! operand), as libtensor passes them.
symmetry sum_symmetry(const symmetry &a, const symmetry &b, const perm_t *pb = 0);

//! Symmetry of the result of a contraction
//! (libtensor/gen_block_tensor/impl/gen_bto_contract2_sym_impl.h:42-86):
//! so_reduce(so_dirprod(symA, symB)) over the contracted index pairs.
symmetry contract2_symmetry(const contraction2 &c, const symmetry &syma,
    const symmetry &symb, bool self);

} // namespace lta
