"""Synthetic block-tensor contraction workloads (problem descriptions) and
their construction through the C ABI.

A problem is a plain dict: tensors "A", "B" (dims, split points, symmetry
elements, random-fill parameters) and "contr" (N, M, K, permc, contracted
pairs, ka, kb, kc); "tables" lists point-group product tables.

  ring(o, v, ...)    C[i,j,a,b] = sum_{k,c} A[i,k,a,c] B[k,j,c,b]
  ladder(...)        R[i,j,a,b] = sum_{c,d} T[i,j,c,d] V[a,b,c,d] with
                     alpha/beta spin partitions (se_part), C2v block labels
                     (se_label), (ij)/(ab)/(cd) antisymmetry and the
                     (ab)<->(cd) pair symmetry of V (se_perm)
  pp_ladder_o40_v200()  the benchmark configuration (o = 40, v = 200)
  ring_batched_o60_v300()  ring, o = 60, v = 300, block sizes U{8..96}
  serving_clients()  the serving mix: 64 calculations, o ~ U{8..48},
                     v ~ U{40..240}, ring and ladder patterns
"""
import itertools

import numpy as np

from . import (block_index_space, block_tensor, btod_contract2, contraction2,
               product_table, se_label, se_part, se_perm)

C2V = ("c2v", 4, [(1, 1, 0), (1, 2, 3), (1, 3, 2), (2, 2, 0), (2, 3, 1), (3, 3, 0)])


def ring(o, v, so=(), sv=()):
    A = {"dims": [o, o, v, v], "splits": [([1, 1, 0, 0], int(p)) for p in so] +
         [([0, 0, 1, 1], int(p)) for p in sv]}
    B = {"dims": [o, o, v, v], "splits": [([1, 1, 0, 0], int(p)) for p in so] +
         [([0, 0, 1, 1], int(p)) for p in sv]}
    return {"A": A, "B": B,
            "contr": {"N": 2, "M": 2, "K": 2, "permc": [0, 2, 1, 3],
                      "pairs": [(1, 0), (3, 2)]}}


def spin_part(order, conserve):
    """se_part with two partitions (alpha, beta) on every dimension:
    partitions that do not conserve spin are forbidden; each allowed
    partition maps onto its spin-flipped partner with +1."""
    ops = []
    for p in itertools.product([0, 1], repeat=order):
        if not conserve(p):
            ops.append(("forbid", tuple(p)))
    for p in itertools.product([0, 1], repeat=order):
        if not conserve(p):
            continue
        q = tuple(1 - x for x in p)
        if p < q:
            ops.append(("map", tuple(p), q, 1.0))
    return ("part", tuple([2] * order), ops)


def ladder(o_blocks=(3, 2), v_blocks=(4, 3, 5), irreps_o=(0, 2), irreps_v=(0, 1, 3),
           spin=True, labels=True, antisym=True):
    ob = list(o_blocks) + (list(o_blocks) if spin else [])
    vb = list(v_blocks) + (list(v_blocks) if spin else [])
    o, v = sum(ob), sum(vb)
    so = [int(x) for x in np.cumsum(ob)[:-1]]
    sv = [int(x) for x in np.cumsum(vb)[:-1]]
    lab_o = list(irreps_o) * (2 if spin else 1)
    lab_v = list(irreps_v) * (2 if spin else 1)

    def conserve(p):
        return p[0] + p[1] == p[2] + p[3]

    def mk(kinds, perms):
        dims = [o if k == "o" else v for k in kinds]
        om = [1 if k == "o" else 0 for k in kinds]
        vm = [1 if k == "v" else 0 for k in kinds]
        splits = []
        if any(om):
            splits += [(om, p) for p in so]
        if any(vm):
            splits += [(vm, p) for p in sv]
        elems = [("perm", list(pr), c) for pr, c in perms]
        if spin:
            elems.append(spin_part(4, conserve))
        if labels:
            assigns = []
            if any(om):
                assigns += [(om, b, l) for b, l in enumerate(lab_o)]
            if any(vm):
                assigns += [(vm, b, l) for b, l in enumerate(lab_v)]
            elems.append(("label", "c2v", assigns, [0]))
        return {"dims": dims, "splits": splits, "elems": elems}

    pT = [([1, 0, 2, 3], -1.0), ([0, 1, 3, 2], -1.0)] if antisym else []
    pV = ([([1, 0, 2, 3], -1.0), ([0, 1, 3, 2], -1.0)] if antisym else []) + \
        [([2, 3, 0, 1], 1.0)]
    T = mk("oovv", pT)
    T.update({"dist": "normal", "scale": 0.01, "scale_dims": (2, 0)})
    V = mk("vvvv", pV)
    V["dist"] = "normal"
    prob = {"A": T, "B": V,
            "contr": {"N": 2, "M": 2, "K": 2, "permc": None, "pairs": [(2, 2), (3, 3)]}}
    if labels:
        prob["tables"] = [C2V]
    return prob


def pp_ladder_o40_v200():
    """pp-ladder, o = 40, v = 200: per spin, occupied blocks 7+5+4+4 (one per
    C2v irrep) and virtual blocks 18+16 (A1), 16 (A2), 14+12 (B1), 24 (B2)."""
    return ladder(o_blocks=(7, 5, 4, 4), v_blocks=(18, 16, 16, 14, 12, 24),
                  irreps_o=(0, 1, 2, 3), irreps_v=(0, 0, 1, 2, 2, 3))


def uneven_splits(n, lo, hi, rng):
    """Split points of n into block sizes drawn from U{lo..hi} (the last block
    takes the remainder; a remainder below lo joins the previous block)."""
    sizes = []
    while sum(sizes) < n:
        sizes.append(int(rng.integers(lo, hi + 1)))
    sizes[-1] -= sum(sizes) - n
    if sizes[-1] < lo and len(sizes) > 1:
        last = sizes.pop()
        sizes[-1] += last
    return [int(x) for x in np.cumsum(sizes)[:-1]]


def ring_batched_o60_v300(seed=1003):
    """C[i,j,a,b] = sum_{k,c} A[i,k,a,c] B[k,j,c,b], o = 60, v = 300, no
    symmetry, block sizes drawn from U{8..96} (numpy default_rng(seed))."""
    rng = np.random.default_rng(seed)
    so = uneven_splits(60, 8, 96, rng)
    sv = uneven_splits(300, 8, 96, rng)
    return ring(60, 300, so, sv)


def _parts(n, k, rng):
    """n split into k positive parts of nearly equal size (random order)."""
    base = [n // k + (1 if i < n % k else 0) for i in range(k)]
    rng.shuffle(base)
    return [int(x) for x in base]


def serving_clients(n=64, seed=1005):
    """Problems of the serving mix: client i draws o ~ U{8..48} and
    v ~ U{40..240} (numpy default_rng(seed + i)); even clients run a ring
    (one block per dimension, no symmetry), odd clients a ladder with spin
    and C2v symmetry (occupied: one block per irrep and spin; virtual:
    blocks U{8..32} per spin, irreps assigned in turn)."""
    out = []
    for i in range(n):
        rng = np.random.default_rng(seed + i)
        o = int(rng.integers(8, 49))
        v = int(rng.integers(40, 241))
        if i % 2 == 0:
            p = ring(o, v)
            p["name"] = "ring o=%d v=%d" % (o, v)
        else:
            oh, vh = max(4, o // 2), max(8, v // 2)
            ob = [x for x in _parts(oh, 4, rng) if x > 0]
            vb = []
            while sum(vb) < vh:
                vb.append(int(rng.integers(8, 33)))
            vb[-1] -= sum(vb) - vh
            if vb[-1] < 4 and len(vb) > 1:
                last = vb.pop()
                vb[-1] += last
            p = ladder(o_blocks=tuple(ob), v_blocks=tuple(vb),
                       irreps_o=tuple(range(len(ob))),
                       irreps_v=tuple(j % 4 for j in range(len(vb))))
            p["name"] = "ladder o=%d v=%d" % (2 * oh, 2 * vh)
        out.append(p)
    return out


# ---------------------------------------------------------------------------

def register_tables(prob):
    for tid, n, prods in prob.get("tables", []):
        t = product_table(tid, n)
        for l1, l2, lr in prods:
            t.add_product(l1, l2, lr)


def build_tensor(ctx, td):
    bis = block_index_space(td["dims"])
    for mask, pos in td.get("splits", []):
        bis.split(mask, pos)
    bt = block_tensor(ctx, bis)
    for el in td.get("elems", []):
        if el[0] == "perm":
            bt.insert(se_perm(el[1], el[2]))
        elif el[0] == "part":
            sp = se_part(el[1])
            for op in el[2]:
                if op[0] == "map":
                    sp.add_map(op[1], op[2], op[3])
                else:
                    sp.mark_forbidden(op[1])
            bt.insert(sp)
        else:
            sl = se_label(el[1])
            for mask, blk, lab in el[2]:
                sl.assign(mask, blk, lab)
            if el[3] is not None:
                sl.set_rule(el[3])
            bt.insert(sl)
    return bt


def fill_random(bt, td, seed):
    bt.random(seed, td.get("dist", "uniform"), td.get("scale", 1.0), td.get("scale_dims"))


def make_contraction(prob):
    c = prob["contr"]
    contr = contraction2(c["N"], c["M"], c["K"], c.get("permc"))
    for ia, ib in c["pairs"]:
        contr.contract(ia, ib)
    return contr


def build_random(ctx, prob, seed):
    """A, B filled with symmetric random data (seeded) and the operation."""
    register_tables(prob)
    a = build_tensor(ctx, prob["A"])
    fill_random(a, prob["A"], seed)
    if prob.get("self"):
        b = a
    else:
        b = build_tensor(ctx, prob["B"])
        fill_random(b, prob["B"], seed + 7)
    c = prob["contr"]
    op = btod_contract2(make_contraction(prob), a, b, c.get("ka", 1.0),
                        c.get("kb", 1.0), c.get("kc", 1.0))
    return a, b, op


def result_order(prob):
    c = prob["contr"]
    ka = set(p[0] for p in c["pairs"])
    kb = set(p[1] for p in c["pairs"])
    bdesc = prob["A"] if prob.get("self") else prob["B"]
    rem = [("A", i) for i in range(len(prob["A"]["dims"])) if i not in ka] + \
          [("B", i) for i in range(len(bdesc["dims"])) if i not in kb]
    permc = c.get("permc") or list(range(len(rem)))
    return [rem[permc[i]] for i in range(len(rem))]


def result_tensor(ctx, prob):
    """Empty result block tensor with the splits of the source dimensions."""
    order = result_order(prob)
    dims = []
    splits = []
    for t, i in order:
        td = prob["A"] if (t == "A" or prob.get("self")) else prob["B"]
        dims.append(td["dims"][i])
        splits.append(sorted(set(p for m, p in td.get("splits", []) if m[i])))
    bis = block_index_space(dims)
    for d, pos in enumerate(splits):
        for p in pos:
            m = [0] * len(dims)
            m[d] = 1
            bis.split(m, p)
    return block_tensor(ctx, bis)
