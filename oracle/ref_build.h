// Test infrastructure (checker only, never the product).
//
// Builds libtensor block tensors from a problem description (spec_format.h):
// block index space with its splits, symmetry elements in file order, and
// the data of the canonical blocks.  Shared by ref_driver.cpp and
// seam_driver.cpp.
#pragma once

#include <libtensor/libtensor.h>
#include <libtensor/symmetry/point_group_table.h>
#include <libtensor/symmetry/product_table_container.h>
#include <libtensor/symmetry/se_label.h>
#include <libtensor/symmetry/se_part.h>
#include <libtensor/symmetry/se_perm.h>

#include <cstdio>
#include <cstring>
#include <stdexcept>
#include <vector>

#include "spec_format.h"

namespace ref_build {

using namespace libtensor;

template<size_t N>
permutation<N> make_perm(const std::vector<size_t> &p) {
    sequence<N, size_t> s1(0), s2(0);
    for(size_t i = 0; i < N; i++) { s1[i] = p[i]; s2[i] = i; }
    return permutation_builder<N>(s1, s2).get_perm();
}

template<size_t N>
libtensor::index<N> make_index(const std::vector<size_t> &v) {
    libtensor::index<N> i;
    for(size_t k = 0; k < N; k++) i[k] = v[k];
    return i;
}

template<size_t N>
block_index_space<N> make_bis(const spec::tensor &t) {
    libtensor::index<N> i1, i2;
    for(size_t i = 0; i < N; i++) i2[i] = t.dims[i] - 1;
    block_index_space<N> bis(dimensions<N>(index_range<N>(i1, i2)));
    for(size_t s = 0; s < t.splits.size(); s++) {
        mask<N> m;
        for(size_t i = 0; i < N; i++) m[i] = t.splits[s].mask[i] != 0;
        bis.split(m, t.splits[s].pos);
    }
    return bis;
}

template<size_t N>
void fill_tensor(const spec::tensor &t, block_tensor<N, double, allocator<double> > &bt) {
    block_tensor_ctrl<N, double> ctrl(bt);
    const block_index_space<N> &bis = bt.get_bis();
    symmetry<N, double> &sym = ctrl.req_symmetry();
    for(size_t e = 0; e < t.elems.size(); e++) {
        const spec::element &el = t.elems[e];
        if(el.kind == 0) {
            sym.insert(se_perm<N, double>(make_perm<N>(el.perm),
                scalar_transf<double>(el.coeff)));
        } else if(el.kind == 1) {
            libtensor::index<N> i1, i2;
            for(size_t i = 0; i < N; i++) i2[i] = el.part.pdims[i] - 1;
            se_part<N, double> sp(bis, dimensions<N>(index_range<N>(i1, i2)));
            size_t im = 0, ifb = 0;
            for(size_t k = 0; k < el.part.order.size(); k++) {
                if(el.part.order[k] == 0) {
                    const spec::part_el::mp &m = el.part.maps[im++];
                    sp.add_map(make_index<N>(m.i1), make_index<N>(m.i2),
                        scalar_transf<double>(m.c));
                } else {
                    sp.mark_forbidden(make_index<N>(el.part.forbid[ifb++]));
                }
            }
            sym.insert(sp);
        } else {
            se_label<N, double> sl(bis.get_block_index_dims(), el.label.table);
            for(size_t a = 0; a < el.label.assigns.size(); a++) {
                mask<N> m;
                for(size_t i = 0; i < N; i++) m[i] = el.label.assigns[a].mask[i] != 0;
                sl.get_labeling().assign(m, el.label.assigns[a].blk,
                    el.label.assigns[a].label);
            }
            if(el.label.has_rule) {
                product_table_i::label_set_t ls;
                for(size_t r = 0; r < el.label.rule.size(); r++) ls.insert(el.label.rule[r]);
                sl.set_rule(ls);
            }
            sym.insert(sl);
        }
    }

    dimensions<N> bidims = bis.get_block_index_dims();
    const dimensions<N> &dims = bis.get_dims();
    if(!t.dense_file.empty()) {
        std::vector<double> dense(dims.get_size());
        FILE *f = fopen(t.dense_file.c_str(), "rb");
        if(!f || fread(&dense[0], 8, dense.size(), f) != dense.size())
            throw std::runtime_error("cannot read " + t.dense_file);
        fclose(f);
        orbit_list<N, double> ol(sym);
        for(typename orbit_list<N, double>::iterator i = ol.begin(); i != ol.end(); ++i) {
            libtensor::index<N> bi;
            ol.get_index(i, bi);
            libtensor::index<N> st = bis.get_block_start(bi);
            dimensions<N> bd = bis.get_block_dims(bi);
            std::vector<double> blk(bd.get_size());
            bool nz = false;
            abs_index<N> ai(bd);
            do {
                const libtensor::index<N> &e = ai.get_index();
                size_t off = 0;
                for(size_t k = 0; k < N; k++) off = off * dims[k] + st[k] + e[k];
                blk[ai.get_abs_index()] = dense[off];
                if(dense[off] != 0.0) nz = true;
            } while(ai.inc());
            if(!nz) continue;
            dense_tensor_wr_i<N, double> &b = ctrl.req_block(bi);
            {
                dense_tensor_wr_ctrl<N, double> dc(b);
                double *p = dc.req_dataptr();
                std::memcpy(p, &blk[0], 8 * blk.size());
                dc.ret_dataptr(p);
            }
            ctrl.ret_block(bi);
        }
    }
    if(!t.blocks_file.empty()) {
        FILE *f = fopen(t.blocks_file.c_str(), "rb");
        if(!f) throw std::runtime_error("cannot read " + t.blocks_file);
        unsigned long long cnt = 0;
        if(fread(&cnt, 8, 1, f) != 1) throw std::runtime_error("blocks: header");
        for(unsigned long long k = 0; k < cnt; k++) {
            unsigned long long aidx = 0, ne = 0;
            if(fread(&aidx, 8, 1, f) != 1 || fread(&ne, 8, 1, f) != 1)
                throw std::runtime_error("blocks: record");
            std::vector<double> blk(ne);
            if(ne && fread(&blk[0], 8, ne, f) != ne) throw std::runtime_error("blocks: data");
            libtensor::index<N> bi;
            abs_index<N>::get_index(aidx, bidims, bi);
            dense_tensor_wr_i<N, double> &b = ctrl.req_block(bi);
            {
                dense_tensor_wr_ctrl<N, double> dc(b);
                double *p = dc.req_dataptr();
                if(ne != b.get_dims().get_size() && ne) throw std::runtime_error("blocks: size");
                std::memcpy(p, &blk[0], 8 * ne);
                dc.ret_dataptr(p);
            }
            ctrl.ret_block(bi);
        }
        fclose(f);
    }
}

//! Registers a product table of the spec with libtensor (point_group_table
//! with labels L0, L1, ...; L0 is the identity).
inline void add_table(const spec::table &tb) {
    std::vector<std::string> irreps;
    char buf[32];
    for(size_t l = 0; l < tb.nlabels; l++) {
        snprintf(buf, sizeof(buf), "L%zu", l);
        irreps.push_back(buf);
    }
    point_group_table pg(tb.id, irreps, irreps[0]);
    for(size_t k = 0; k < tb.products.size(); k++)
        pg.add_product(tb.products[k].l1, tb.products[k].l2, tb.products[k].lr);
    product_table_container::get_instance().add(pg);
}

} // namespace ref_build
