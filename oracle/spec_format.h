// Test infrastructure (checker only, never the product).
//
// Line-based description of a block-tensor contraction, read by the
// reference driver (ref_driver.cpp) and written by the tests:
//
//   table <id> <nlabels>                 point-group product table
//   product <id> <l1> <l2> <lr>
//   tensor <A|B> <n> <dims...>
//   split <A|B> <mask x n> <pos>         block_index_space::split
//   perm <A|B> <perm x n> <coeff>        se_perm
//   part <A|B> <pdims x n>               se_part, followed by map/forbid, end
//     map <idx1 x n> <idx2 x n> <coeff>
//     forbid <idx x n>
//   end
//   label <A|B> <table>                  se_label, followed by assign/rule, end
//     assign <mask x n> <block> <label>
//     rule <count> <intrinsic labels...>
//   end
//   dense <A|B> <file>                   full row-major doubles
//   blocks <A|B> <file>                  u64 count, then per block: u64 aidx,
//                                        u64 nelem, nelem doubles
//   contract <N> <M> <K> <permc x N+M> <ka kb x K> <ka> <kb> <kc>
//   self 1                               B is the same tensor as A
//   contract3 <N1> <N2> <N3> <K1> <K2> <permab x N1+N2+K2> <a b x K1>
//             <permd x N1+N2+N3> <ab c x K2> <kd>
//                                        btod_contract3 of A, B and C
//                                        (tensor C declared like A and B)
//   symab <perm x N1+N2+K2> <symm 0|1>   btod_symcontract3: symmetrize AB
//                                        (after the contract3 line)
//   symmetrize3 <perm1 x N+M> <perm2 x N+M> <symm 0|1>
//                                        wrap it in btod_symmetrize3 (same
//                                        chain as symmetrize2)
//   add <d>                              also perform(btc, d) into tensor C
//                                        (declared like A and B: its own
//                                        symmetry and data) after the plain
//                                        perform
//   btod_add <n> <perm x n> <ca> <cb>    instead of a contraction:
//                                        btod_add(A, ca) + add_op(B, perm, cb),
//                                        perform(out) (A, B of order n)
//   symmetrize2 <perm x N+M> <symm 0|1>  wrap the contraction in
//                                        btod_symmetrize2 (repeatable: each
//                                        line wraps the previous operation)
#pragma once

#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

namespace spec {

struct part_el {
    std::vector<size_t> pdims;
    struct mp { std::vector<size_t> i1, i2; double c; };
    std::vector<mp> maps;           // in file order
    std::vector<std::vector<size_t> > forbid;
    std::vector<int> order;         // 0 = map, 1 = forbid, in file order
};

struct label_el {
    std::string table;
    struct as { std::vector<int> mask; size_t blk, label; };
    std::vector<as> assigns;
    std::vector<size_t> rule;
    bool has_rule;
    label_el() : has_rule(false) { }
};

struct element {
    int kind;   // 0 perm, 1 part, 2 label
    std::vector<size_t> perm;
    double coeff;
    part_el part;
    label_el label;
};

struct tensor {
    size_t n;
    std::vector<size_t> dims;
    struct sp { std::vector<int> mask; size_t pos; };
    std::vector<sp> splits;
    std::vector<element> elems;
    std::string dense_file, blocks_file;
    tensor() : n(0) { }
};

struct table {
    std::string id;
    size_t nlabels;
    struct pr { size_t l1, l2, lr; };
    std::vector<pr> products;
};

struct problem {
    std::vector<table> tables;
    tensor a, b;
    size_t N, M, K;
    std::vector<size_t> permc, ka, kb;
    double sa, sb, sc;
    bool self;
    tensor c;                       // third operand of contract3
    bool c3;                        // contract3 instead of contract
    size_t N1, N2, N3, K1, K2;
    std::vector<size_t> permab, permd, ka1, kb1, kab2, kc2;
    double kd;
    std::vector<size_t> symab;      // btod_symcontract3 when not empty
    int symab_symm;
    struct symm2 { std::vector<size_t> perm, perm2; int symm; };   // perm2: symmetrize3
    std::vector<symm2> symm;        // in file order (innermost first)
    bool has_add;                   // perform(C, add_d) after perform
    double add_d;
    size_t btod_add_n;              // > 0: btod_add of A and P(B) instead
    std::vector<size_t> add_perm;
    double add_ca, add_cb;
    problem() : N(0), M(0), K(0), sa(1), sb(1), sc(1), self(false), c3(false),
        N1(0), N2(0), N3(0), K1(0), K2(0), kd(1), symab_symm(1), has_add(false),
        add_d(1), btod_add_n(0), add_ca(1), add_cb(1) { }
};

inline std::vector<size_t> read_sz(std::istringstream &in, size_t n) {
    std::vector<size_t> v(n);
    for(size_t i = 0; i < n; i++) if(!(in >> v[i])) throw std::runtime_error("spec: short line");
    return v;
}

inline std::vector<int> read_int(std::istringstream &in, size_t n) {
    std::vector<int> v(n);
    for(size_t i = 0; i < n; i++) if(!(in >> v[i])) throw std::runtime_error("spec: short line");
    return v;
}

inline problem parse(const std::string &path) {
    std::ifstream f(path.c_str());
    if(!f) throw std::runtime_error("spec: cannot open " + path);
    problem p;
    std::string line;
    element *cur = 0;
    tensor *curt = 0;
    while(std::getline(f, line)) {
        if(line.empty() || line[0] == '#') continue;
        std::istringstream in(line);
        std::string kw;
        in >> kw;
        if(kw == "table") {
            table t; in >> t.id >> t.nlabels; p.tables.push_back(t);
        } else if(kw == "product") {
            std::string id; table::pr r; in >> id >> r.l1 >> r.l2 >> r.lr;
            for(size_t i = 0; i < p.tables.size(); i++)
                if(p.tables[i].id == id) p.tables[i].products.push_back(r);
        } else if(kw == "tensor" || kw == "split" || kw == "perm" || kw == "part" ||
            kw == "label" || kw == "dense" || kw == "blocks") {
            std::string name; in >> name;
            tensor &t = (name == "A") ? p.a : (name == "C") ? p.c : p.b;
            if(kw == "tensor") {
                in >> t.n; t.dims = read_sz(in, t.n);
            } else if(kw == "split") {
                tensor::sp s; s.mask = read_int(in, t.n); in >> s.pos; t.splits.push_back(s);
            } else if(kw == "perm") {
                element e; e.kind = 0; e.perm = read_sz(in, t.n); in >> e.coeff;
                t.elems.push_back(e);
            } else if(kw == "part") {
                element e; e.kind = 1; e.part.pdims = read_sz(in, t.n);
                t.elems.push_back(e); cur = &t.elems.back(); curt = &t;
            } else if(kw == "label") {
                element e; e.kind = 2; in >> e.label.table;
                t.elems.push_back(e); cur = &t.elems.back(); curt = &t;
            } else if(kw == "dense") {
                in >> t.dense_file;
            } else {
                in >> t.blocks_file;
            }
        } else if(kw == "map") {
            part_el::mp m; m.i1 = read_sz(in, curt->n); m.i2 = read_sz(in, curt->n); in >> m.c;
            cur->part.maps.push_back(m); cur->part.order.push_back(0);
        } else if(kw == "forbid") {
            cur->part.forbid.push_back(read_sz(in, curt->n)); cur->part.order.push_back(1);
        } else if(kw == "assign") {
            label_el::as a; a.mask = read_int(in, curt->n); in >> a.blk >> a.label;
            cur->label.assigns.push_back(a);
        } else if(kw == "rule") {
            size_t k; in >> k; cur->label.rule = read_sz(in, k); cur->label.has_rule = true;
        } else if(kw == "end") {
            cur = 0; curt = 0;
        } else if(kw == "contract") {
            in >> p.N >> p.M >> p.K;
            p.permc = read_sz(in, p.N + p.M);
            for(size_t i = 0; i < p.K; i++) { size_t a, b; in >> a >> b; p.ka.push_back(a); p.kb.push_back(b); }
            in >> p.sa >> p.sb >> p.sc;
        } else if(kw == "contract3") {
            p.c3 = true;
            in >> p.N1 >> p.N2 >> p.N3 >> p.K1 >> p.K2;
            p.permab = read_sz(in, p.N1 + p.N2 + p.K2);
            for(size_t i = 0; i < p.K1; i++) { size_t a, b; in >> a >> b; p.ka1.push_back(a); p.kb1.push_back(b); }
            p.permd = read_sz(in, p.N1 + p.N2 + p.N3);
            for(size_t i = 0; i < p.K2; i++) { size_t a, b; in >> a >> b; p.kab2.push_back(a); p.kc2.push_back(b); }
            in >> p.kd;
        } else if(kw == "symab") {
            p.symab = read_sz(in, p.N1 + p.N2 + p.K2);
            in >> p.symab_symm;
        } else if(kw == "symmetrize3") {
            problem::symm2 y; y.perm = read_sz(in, p.N + p.M); y.perm2 = read_sz(in, p.N + p.M);
            in >> y.symm;
            p.symm.push_back(y);
        } else if(kw == "symmetrize2") {
            problem::symm2 y; y.perm = read_sz(in, p.N + p.M); in >> y.symm;
            p.symm.push_back(y);
        } else if(kw == "add") {
            p.has_add = true; in >> p.add_d;
        } else if(kw == "btod_add") {
            in >> p.btod_add_n; p.add_perm = read_sz(in, p.btod_add_n);
            in >> p.add_ca >> p.add_cb;
        } else if(kw == "self") {
            int s; in >> s; p.self = s != 0;
        } else {
            throw std::runtime_error("spec: unknown keyword " + kw);
        }
    }
    return p;
}

} // namespace spec
