// Host-side block schedule of a block-tensor contraction: block lists,
// symmetry unfolding of block lists, (k, outer) pair lists, per-C-block
// contraction lists with coalescing, and the list of non-zero canonical C
// blocks.  All integer lists are bit-identical to libtensor's:
//   libtensor/gen_block_tensor/impl/block_list.h
//   libtensor/gen_block_tensor/impl/gen_bto_unfold_block_list_impl.h
//   libtensor/gen_block_tensor/impl/gen_bto_contract2_block_list.h:143-200
//   libtensor/gen_block_tensor/impl/gen_bto_contract2_clst_builder_impl.h
//   libtensor/block_tensor/impl/btod_contract2_clst_optimize_impl.h:11-80
//   libtensor/gen_block_tensor/impl/gen_bto_contract2_nzorb_impl.h:354-476
#pragma once

#include <list>
#include <memory>
#include <mutex>
#include <unordered_map>
#include <vector>
#include "lt_core.h"
#include "lt_sym.h"

namespace lta {

//! Sorted list of absolute block indexes.
struct block_list {
    dims_t dims;
    mutable std::vector<size_t> blks;
    mutable bool sorted;

    block_list() : sorted(true) { }
    explicit block_list(const dims_t &d) : dims(d), sorted(true) { }
    block_list(const dims_t &d, const std::vector<size_t> &b) :
        dims(d), blks(b), sorted(false) { }
    void add(size_t a) {
        blks.push_back(a);
        if(sorted && blks.size() > 1) {
            size_t n = blks.size();
            sorted = blks[n - 2] < blks[n - 1];
        }
    }
    void sort() const {
        if(!sorted) { std::sort(blks.begin(), blks.end()); sorted = true; }
    }
    bool contains(size_t a) const {
        sort();
        return std::binary_search(blks.begin(), blks.end(), a);
    }
    size_t size() const { return blks.size(); }
};

//! Canonical index and transformation of a block, cached per block index.
struct orbit_info {
    size_t acidx;
    tensor_tr tr;   //!< canonical -> block
};

//! Thread-safe cache of orbits started from each block (the starter block
//! determines the transformation chosen, so the cache is keyed by it).
class orbit_cache {
public:
    explicit orbit_cache(const symmetry &sym) : m_sym(sym) { }
    orbit_info get(size_t aidx);
    const symmetry &sym() const { return m_sym; }
private:
    const symmetry &m_sym;
    std::mutex m_mtx;
    std::unordered_map<size_t, orbit_info> m_map;
};

//! Adds every block of the orbits of the listed blocks, then sorts.
void unfold_block_list(const symmetry &sym, const block_list &bl, block_list &out);

//! (k, outer) pair lists of A and B (gen_bto_contract2_block_list):
//! a1/b1 sorted by (k, outer), a2/b2 sorted by (outer, k).
struct contract2_block_list {
    std::vector<std::pair<size_t, size_t> > a1, b1;   // (k, i)
    std::vector<std::pair<size_t, size_t> > a2, b2;   // (i, k)
    contract2_block_list(const contraction2 &c, const block_list &bla,
        const block_list &blb);
};

//! One contraction of a canonical A block with a canonical B block.
struct contr_pair {
    size_t aia, acia;
    tensor_tr tra;
    size_t aib, acib;
    tensor_tr trb;
};
typedef std::list<contr_pair> contr_list;

//! Merges pairs with the same canonical blocks and the same permuted
//! contraction; drops pairs that cancel.
void coalesce_clst(const contraction2 &c, contr_list &clst);

//! Contraction list of one C block from the (k, outer) lists
//! (build_list(false, bl)).
void build_clst(const contraction2 &c, orbit_cache &oca, orbit_cache &ocb,
    const block_list &blax, const block_list &blbx,
    const contract2_block_list &cbl, const index_t &ic, contr_list &out);

//! Contraction list of one C block by enumerating contracted block indexes
//! and orbit pairs (build_list(testzero)).  testzero stops at the first
//! non-empty coalesced sub-list.
void build_clst_full(const contraction2 &c, const symmetry &syma,
    const symmetry &symb, const block_list &blax, const block_list &blbx,
    const dims_t &bidimsc, const index_t &ic, bool testzero, contr_list &out);

//! Non-zero canonical blocks of C (gen_bto_contract2_nzorb::build), sorted.
std::vector<size_t> contract2_schedule(const contraction2 &c,
    const symmetry &syma, const symmetry &symb, const symmetry &symc,
    const std::vector<size_t> &nza, const std::vector<size_t> &nzb,
    size_t nthreads);

//! Runs f(i) for i in [0, n) on up to nthreads host threads.
void parallel_for(size_t n, size_t nthreads, const std::function<void(size_t)> &f);

} // namespace lta
