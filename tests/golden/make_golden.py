"""Writes the golden fixtures of tests/golden/ by running the REAL libtensor
block-contraction path (oracle/_ref/ref_driver, built from the reference
sources by oracle/Makefile) on small problems.

Each fixture <name>.json holds the problem description, the non-zero
schedule, the canonical index / allowedness of every result block and the
coalesced contraction list of every scheduled block as printed by libtensor;
<name>.npz holds the dense inputs and libtensor's dense result.

The inputs are generated with libtensor_amd's block_tensor.random on a
host-only context (no GPU needed).  Run:  python tests/golden/make_golden.py
(--symm-only: only the btod_symmetrize2 fixtures symm_*, from the committed
inputs of their base cases; --c3-only: only the btod_contract3 fixtures c3_*;
--add-only: only the fixtures of += into a target of another symmetry, add_*,
and of btod_add across symmetries, badd_*).
"""
import copy
import json
import os
import random
import shutil
import sys
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(HERE)), "oracle"))

import problems as P  # noqa: E402
import libtensor_amd as lt  # noqa: E402
from libtensor_amd import workloads as W  # noqa: E402
import numpy as np  # noqa: E402


def random_problem(rng, N, M, K, sym=False):
    """Random dimensions, splits, contracted pairs and result permutation."""
    def tensor(n, dims_pool):
        return {"dims": list(dims_pool[:n]), "splits": []}
    sizes = [rng.choice([3, 4, 5, 6]) for _ in range(N + M + K)]
    # index kinds: A outer N, B outer M, contracted K
    da = [0] * (N + K)
    db = [0] * (M + K)
    pos_a = list(range(N + K))
    pos_b = list(range(M + K))
    rng.shuffle(pos_a)
    rng.shuffle(pos_b)
    pairs = []
    for k in range(K):
        ia, ib = pos_a[N + k], pos_b[M + k]
        da[ia] = db[ib] = sizes[N + M + k]
        pairs.append((ia, ib))
    for i in range(N):
        da[pos_a[i]] = sizes[i]
    for j in range(M):
        db[pos_b[j]] = sizes[N + j]
    A = {"dims": da, "splits": []}
    B = {"dims": db, "splits": []}

    def add_splits(t, extra):
        for i, d in enumerate(t["dims"]):
            if d >= 4:
                p = rng.randint(1, d - 1)
                m = [1 if x == i else 0 for x in range(len(t["dims"]))]
                t["splits"].append((m, p))
                extra.setdefault(i, []).append(p)
    ea, eb = {}, {}
    add_splits(A, ea)
    # contracted dims must be split identically in A and B
    for ia, ib in pairs:
        for (m, p) in A["splits"]:
            if m[ia]:
                mb = [1 if x == ib else 0 for x in range(len(db))]
                B["splits"].append((mb, p))
    kb = set(p[1] for p in pairs)
    for j, d in enumerate(db):
        if j in kb or d < 4:
            continue
        p = rng.randint(1, d - 1)
        B["splits"].append(([1 if x == j else 0 for x in range(len(db))], p))
    permc = list(range(N + M))
    rng.shuffle(permc)
    prob = {"A": A, "B": B, "contr": {"N": N, "M": M, "K": K, "permc": permc,
                                       "pairs": pairs, "ka": 0.5, "kb": -2.0, "kc": 1.5}}
    if sym:
        # (01)-antisymmetry on A when its first two dims agree
        if len(da) >= 2:
            da[1] = da[0]
            A["splits"] = [(m, p) for (m, p) in A["splits"] if not (m[0] or m[1])]
            if da[0] >= 4:
                A["splits"].append(([1, 1] + [0] * (len(da) - 2), 2))
            # keep contracted partners consistent
            B["splits"] = [s for s in B["splits"] if not any(s[0][ib] for ia, ib in pairs)]
            for ia, ib in pairs:
                db[ib] = da[ia]
                for (m, p) in A["splits"]:
                    if m[ia]:
                        B["splits"].append(([1 if x == ib else 0 for x in range(len(db))], p))
            A["elems"] = [("perm", [1, 0] + list(range(2, len(da))), -1.0)]
    return prob


def cases():
    rng = random.Random(20261202)
    out = [
        ("ring_small", P.prob_ring(4, 6, [2], [2, 4])),
        ("ring_single_block", P.prob_ring(3, 5)),
        ("ladder_nosym", P.prob_ladder(spin=False, labels=False, antisym=False)),
        ("ladder_perm", P.prob_ladder(spin=False, labels=False)),
        ("ladder_spin", P.prob_ladder(labels=False)),
        ("ladder_label", P.prob_ladder(spin=False)),
        ("ladder_full", P.prob_ladder()),
    ]
    for (N, M, K) in [(1, 1, 1), (2, 1, 1), (1, 2, 1), (2, 2, 1), (1, 1, 2), (2, 1, 2),
                      (1, 2, 2), (2, 2, 2), (1, 1, 0), (2, 1, 0), (0, 2, 2), (2, 0, 2),
                      (3, 1, 1), (1, 1, 3)]:
        out.append(("rand_%d%d%d" % (N, M, K), random_problem(rng, N, M, K)))
    for (N, M, K) in [(1, 1, 1), (2, 1, 1), (1, 2, 2), (2, 2, 2)]:
        out.append(("randsym_%d%d%d" % (N, M, K), random_problem(rng, N, M, K, sym=True)))
    self_prob = P.prob_ring(4, 4, [2], [2])
    self_prob["self"] = True
    self_prob["contr"] = {"N": 2, "M": 2, "K": 2, "permc": None, "pairs": [(2, 2), (3, 3)]}
    self_prob["A"]["elems"] = [("perm", [1, 0, 2, 3], -1.0), ("perm", [0, 1, 3, 2], -1.0)]
    out.append(("self_ladder", self_prob))
    return out


# btod_symmetrize2 over the contraction of a golden case (innermost first):
# (name, base case, [(pair permutation, symmetric)], result stored).  The
# inputs are those of the base case.  Cases libtensor rejects (a permutation
# that does not map the block index space onto itself, or a sign that
# contradicts the operand's symmetry) are stored with "error": true.
SYMM_CASES = [
    ("symm_ring_ijab", "ring_small", [([1, 0, 3, 2], True)]),
    ("symm_ring_ab_ij", "ring_small", [([0, 1, 3, 2], False), ([1, 0, 2, 3], False)]),
    ("symm_ladder_full_ijab", "ladder_full", [([1, 0, 3, 2], True)]),
    ("symm_ladder_full_ab", "ladder_full", [([0, 1, 3, 2], False)]),
    ("symm_ladder_full_ab_ij", "ladder_full", [([0, 1, 3, 2], False), ([1, 0, 2, 3], False)]),
    ("symm_ladder_label_ijab", "ladder_label", [([1, 0, 3, 2], True)]),
    ("symm_ladder_nosym_ij", "ladder_nosym", [([1, 0, 2, 3], False)]),
    ("symm_rand_221", "rand_221", [([0, 1, 3, 2], True)]),
    ("symm_rand_311", "rand_311", [([3, 1, 2, 0], False)]),
    ("symm_randsym_211", "randsym_211", [([2, 1, 0], False)]),
    ("symm_self_ladder", "self_ladder", [([2, 1, 0, 3], True)]),
    ("symm_err_sign", "ladder_full", [([0, 1, 3, 2], True)]),
    ("symm_err_bis", "rand_222", [([2, 1, 0, 3], False)]),
]


def swap(n, i, j):
    p = list(range(n))
    p[i], p[j] = p[j], p[i]
    return p


def random_problem3(rng, N1, N2, N3, K1, K2, sym=False, symab=None, labels=False):
    """btod_contract3 problem with random extents, splits and index orders.
    sym: the two K2 indexes (and the two N3 indexes when N3 == 2) get equal
    extents and splits, and B and C antisymmetry in them."""
    letters = iter("abcdefghijklmnopqrstuvw")
    o1 = [next(letters) for _ in range(N1)]
    o2 = [next(letters) for _ in range(N2)]
    o3 = [next(letters) for _ in range(N3)]
    k1 = [next(letters) for _ in range(K1)]
    k2 = [next(letters) for _ in range(K2)]
    size, split = {}, {}
    for i in o1 + o2 + o3 + k1 + k2:
        size[i] = rng.choice([3, 4, 5, 6])
        split[i] = rng.randint(1, size[i] - 1) if size[i] >= 4 else None
    if sym:
        for grp in (k2, o3):
            if len(grp) == 2:
                size[grp[1]], split[grp[1]] = size[grp[0]], split[grp[0]]
    ia = o1 + k1
    ib = o2 + k1 + k2
    ic = o3 + k2
    for x in (ia, ib, ic):
        rng.shuffle(x)
    rem1 = [i for i in ia if i not in k1] + [i for i in ib if i not in k1]
    permab = list(range(len(rem1)))
    rng.shuffle(permab)
    iab = [rem1[permab[i]] for i in range(len(rem1))]
    if symab is not None:
        # btod_symcontract3: the first two indexes of AB get one block
        # structure and AB is (anti)symmetrized in them
        size[rem1[1]], split[rem1[1]] = size[rem1[0]], split[rem1[0]]
    rem2 = [i for i in iab if i not in k2] + [i for i in ic if i not in k2]
    permd = list(range(len(rem2)))
    rng.shuffle(permd)
    idd = [rem2[permd[i]] for i in range(len(rem2))]

    # C2v labels of the blocks of every index (labels: on A, B and C, the
    # totally symmetric product allowed; irreps A1 and A2 only, so that about
    # half the blocks are allowed)
    lab = {}
    if labels:
        lab = {i: [rng.randrange(2) for _ in range(1 if split[i] is None else 2)]
               for i in sorted(size)}
        if sym:
            for grp in (k2, o3):
                if len(grp) == 2:
                    lab[grp[1]] = list(lab[grp[0]])
        if symab is not None:
            lab[rem1[1]] = list(lab[rem1[0]])

    def tensor(ix, with_labels=False):
        t = {"dims": [size[i] for i in ix], "splits": []}
        for p, i in enumerate(ix):
            if split[i] is not None:
                t["splits"].append(([1 if q == p else 0 for q in range(len(ix))], split[i]))
        if with_labels:
            assigns = []
            for p, i in enumerate(ix):
                m = [1 if q == p else 0 for q in range(len(ix))]
                assigns += [(m, b, l) for b, l in enumerate(lab[i])]
            t["elems"] = [("label", "c2v", assigns, [0])]
        return t
    prob = {"A": tensor(ia, labels), "B": tensor(ib, labels), "C": tensor(ic, labels),
            "D": tensor(idd),
            "idx": {"A": ia, "B": ib, "C": ic, "D": idd},
            "contr3": {"N1": N1, "N2": N2, "N3": N3, "K1": K1, "K2": K2, "permab": permab,
                       "pairs1": [(ia.index(k), ib.index(k)) for k in k1],
                       "permd": permd,
                       "pairs2": [(iab.index(k), ic.index(k)) for k in k2], "kd": -1.5}}
    prob["idx"]["AB"] = iab
    if labels:
        prob["tables"] = [W.C2V]
    if symab is not None:
        prob["contr3"]["symab"] = (swap(len(iab), iab.index(rem1[0]), iab.index(rem1[1])),
                                   symab)
    if sym and K2 == 2:
        for name, ix in (("B", ib), ("C", ic)):
            p = list(range(len(ix)))
            a, b = ix.index(k2[0]), ix.index(k2[1])
            p[a], p[b] = b, a
            prob[name]["elems"] = [("perm", p, -1.0)] + prob[name].get("elems", [])
        if N3 == 2:
            p = list(range(len(ic)))
            a, b = ic.index(o3[0]), ic.index(o3[1])
            p[a], p[b] = b, a
            prob["C"]["elems"].insert(1, ("perm", p, -1.0))
    return prob


def c3_cases():
    rng = random.Random(20261203)
    out = []
    for (n1, n2, n3, k1, k2) in [(1, 0, 1, 1, 1), (1, 1, 1, 1, 1), (1, 1, 2, 0, 0),
                                 (1, 1, 2, 1, 2), (2, 0, 0, 0, 2), (2, 0, 1, 1, 2),
                                 (2, 0, 2, 1, 2)]:
        out.append(("c3_%d%d%d%d%d" % (n1, n2, n3, k1, k2),
                    random_problem3(rng, n1, n2, n3, k1, k2)))
    for (n1, n2, n3, k1, k2) in [(1, 1, 2, 1, 2), (2, 0, 2, 1, 2)]:
        out.append(("c3sym_%d%d%d%d%d" % (n1, n2, n3, k1, k2),
                    random_problem3(rng, n1, n2, n3, k1, k2, sym=True)))
    # btod_symcontract3, the orders libtensor instantiates
    rng = random.Random(20261204)
    for (n1, n2, n3, k1, k2), symm in [((1, 0, 1, 1, 1), True), ((1, 1, 1, 1, 1), False),
                                       ((2, 0, 2, 1, 2), True), ((2, 0, 2, 1, 2), False)]:
        out.append(("c3s_%d%d%d%d%d_%s" % (n1, n2, n3, k1, k2, "s" if symm else "a"),
                    random_problem3(rng, n1, n2, n3, k1, k2, sym=(k2 == 2), symab=symm)))
    # with C2v labels on every operand: the second contraction takes the
    # intermediate's reduced labels as an operand's
    rng = random.Random(20261205)
    for (n1, n2, n3, k1, k2), sym in [((1, 1, 1, 1, 1), False), ((1, 1, 2, 1, 2), False),
                                      ((2, 0, 2, 1, 2), True), ((2, 0, 1, 1, 2), False)]:
        out.append(("c3lab_%d%d%d%d%d" % (n1, n2, n3, k1, k2),
                    random_problem3(rng, n1, n2, n3, k1, k2, sym=sym, labels=True)))
    # btod_symcontract3 with labels: the symmetrized intermediate's label
    # element (a reduced element with its permuted copy) feeds the second
    # contraction
    rng = random.Random(20261206)
    for (n1, n2, n3, k1, k2), symm in [((1, 1, 1, 1, 1), True), ((2, 0, 2, 1, 2), False)]:
        out.append(("c3slab_%d%d%d%d%d" % (n1, n2, n3, k1, k2),
                    random_problem3(rng, n1, n2, n3, k1, k2, sym=(k2 == 2), symab=symm,
                                    labels=True)))
    return out


def write_c3(work):
    """Fixtures c3_*.json (problem, result orbits, schedule) and .npz (inputs,
    libtensor's result) for btod_contract3."""
    ctx = lt.Context(-1)
    for i, (name, prob) in enumerate(c3_cases()):
        dense = P.make_data3(ctx, prob, 3000 + 11 * i)
        shape = tuple(prob["D"]["dims"])
        ref = P.ref_io.run_ref(prob, os.path.join(work, name), shape, dense=dense)
        meta = {"prob": prob, "schedule": ref["schedule"],
                "symt": {str(k): list(v) for k, v in ref["symt"].items()}}
        with open(os.path.join(HERE, name + ".json"), "w") as f:
            json.dump(meta, f)
        np.savez_compressed(os.path.join(HERE, name + ".npz"), A=dense["A"], B=dense["B"],
                            C=dense["C"], result=ref["result"])
        print("%-22s schedule %4d" % (name, len(ref["schedule"])))


def sym3_problem(kind, n_o, splits, spin, labels):
    """Contractions whose results have three dimensions (or three pairs) of one
    block structure, with spin partitions and C2v labels on every dimension:
    "k1": C[i,j,k,l] = A[i,j,k,m] B[m,l]; "pairs": C[i,a,j,b,k,c] =
    A[i,a,j,m] B[m,b,k,c] with B symmetric under (ia)(jb)... exchange."""
    blocks = len(splits) + 1
    labs = [(0, 2)[b % 2] for b in range(blocks)]
    labv = [(1, 3)[b % 2] for b in range(blocks)]

    def T(kinds, conserve, perms=()):
        n = len(kinds)
        t = {"dims": [n_o] * n, "splits": [], "elems": [("perm", list(p), c) for p, c in perms]}
        for d in range(n):
            m = [1 if q == d else 0 for q in range(n)]
            t["splits"] += [(m, p) for p in splits]
        if spin:
            t["elems"].append(W.spin_part(n, conserve))
        if labels:
            assigns = []
            for d, k in enumerate(kinds):
                m = [1 if q == d else 0 for q in range(n)]
                assigns += [(m, b, l) for b, l in enumerate(labs if k == "o" else labv)]
            t["elems"].append(("label", "c2v", assigns, [0]))
        return t
    if kind == "k1":
        prob = {"A": T("oooo", lambda p: p[0] + p[1] == p[2] + p[3]),
                "B": T("oo", lambda p: p[0] == p[1]),
                "contr": {"N": 3, "M": 1, "K": 1, "permc": None, "pairs": [(3, 0)]}}
    else:
        prob = {"A": T("ovoo", lambda p: p[0] + p[2] == p[1] + p[3]),
                "B": T("ovov", lambda p: p[0] + p[2] == p[1] + p[3], [([2, 3, 0, 1], 1.0)]),
                "contr": {"N": 3, "M": 3, "K": 1, "permc": None, "pairs": [(3, 0)]}}
    if labels:
        prob["tables"] = [W.C2V]
    return prob


# btod_symmetrize3 (entries (perm1, perm2, symmetric)), over a golden case or
# over a problem of its own (inputs stored with the fixture).
SYMM3_CASES = [
    ("symm3_self_ladder_012", "self_ladder", [(swap(4, 0, 1), swap(4, 0, 2), True)]),
    ("symm3_self_ladder_123", "self_ladder", [(swap(4, 1, 2), swap(4, 1, 3), False)]),
    ("symm3_self_ladder_then_2", "self_ladder", [(swap(4, 1, 2), swap(4, 1, 3), False),
                                                 ([1, 0, 2, 3], True)]),
    ("symm3_k1_spin_label", ("k1", 10, [2, 5, 7], True, True),
     [(swap(4, 0, 1), swap(4, 0, 2), True)]),
    ("symm3_k1_spin", ("k1", 10, [2, 5, 7], True, False),
     [(swap(4, 1, 2), swap(4, 1, 3), False)]),
    ("symm3_pairs_spin", ("pairs", 4, [2], True, False),
     [([2, 3, 0, 1, 4, 5], [4, 5, 2, 3, 0, 1], True)]),
    ("symm3_err_groups", "self_ladder", [(swap(4, 0, 1), swap(4, 2, 3), True)]),
]


def write_symm(work):
    """Fixtures symm_*.json (+ .npz with libtensor's result; with the inputs
    too for a problem of its own)."""
    ctx = lt.Context(-1)
    for name, base, symm in SYMM_CASES + SYMM3_CASES:
        if isinstance(base, tuple):
            prob = sym3_problem(*base)
            dense = P.make_data(ctx, prob, 5000 + len(name))
            out = {"prob": prob, "symm": symm}
        else:
            prob, dense, meta, result = P.load_case(os.path.join(HERE, base))
            out = {"base": base, "symm": symm}
        prob = dict(prob)
        prob["symm"] = symm
        try:
            ref = P.run_ref(prob, dense, os.path.join(work, name))
        except RuntimeError:
            out["error"] = True
            ref = None
        if ref is not None:
            # the schedule as a set: libtensor's btod_symmetrize3 can list a
            # block more than once
            out["schedule"] = sorted(set(ref["schedule"]))
            out["symt"] = {str(k): list(v) for k, v in ref["symt"].items()}
            arrays = {"result": ref["result"]}
            if "prob" in out:
                arrays.update(dense)
            np.savez_compressed(os.path.join(HERE, name + ".npz"), **arrays)
        with open(os.path.join(HERE, name + ".json"), "w") as f:
            json.dump(out, f)
        print("%-26s %s" % (name, "error" if ref is None else
                            "schedule %d" % len(ref["schedule"])))


def target_like(td, kind):
    """A tensor description with td's block index space and the symmetry
    `kind`: "same", "none", "ij" ((01) antisymmetry only), "pair" ((01)(23)
    symmetric only), "nolabel", "nopart", "label01" (label rule {0, 1})."""
    t = copy.deepcopy(td)
    for k in ("scale", "scale_dims", "dist"):
        t.pop(k, None)
    el = t.get("elems", [])
    if kind == "none":
        t["elems"] = []
    elif kind == "ij":
        t["elems"] = [("perm", [1, 0, 2, 3], -1.0)] + [e for e in el if e[0] != "perm"]
    elif kind == "pair":
        t["elems"] = [("perm", [1, 0, 3, 2], 1.0)] + [e for e in el if e[0] != "perm"]
    elif kind == "nolabel":
        t["elems"] = [e for e in el if e[0] != "label"]
    elif kind == "nopart":
        t["elems"] = [e for e in el if e[0] != "part"]
    elif kind == "label01":
        t["elems"] = [e if e[0] != "label" else (e[0], e[1], e[2], [0, 1]) for e in el]
    return t


# perform(btc, d) into a target of another symmetry: (name, base problem,
# target symmetry, d).  The target has the result's block index space (the
# ladder's T and the ring's A share it) and random data of its own symmetry.
ADD_CASES = [
    ("add_ladder_same", "ladder_full", "same", 0.5),
    ("add_ladder_none", "ladder_full", "none", 0.5),
    ("add_ladder_ij", "ladder_full", "ij", -1.5),
    ("add_ladder_pair", "ladder_full", "pair", 0.5),
    ("add_ladder_nolabel", "ladder_full", "nolabel", 2.0),
    ("add_ladder_nopart", "ladder_full", "nopart", 0.5),
    ("add_ladder_label01", "ladder_full", "label01", 1.0),
    ("add_ladder_perm_none", "ladder_perm", "none", 0.5),
    ("add_ring_ij", "ring_small", "ij", 0.25),
]

# btod_add(A, ca) + add_op(B, perm, cb) of two tensors with the ladder T's
# block index space: (name, symmetry of A, symmetry of B, perm, ca, cb).
BADD_CASES = [
    ("badd_full_none", "same", "none", [0, 1, 2, 3], 1.0, 1.0),
    ("badd_ij_pair_perm", "ij", "pair", [1, 0, 2, 3], 0.5, -2.0),
    ("badd_full_nolabel", "same", "nolabel", [0, 1, 3, 2], 1.0, 0.5),
    ("badd_label01_nopart", "label01", "nopart", [0, 1, 2, 3], -1.0, 1.0),
    ("badd_full_full_perm", "same", "same", [1, 0, 3, 2], 1.0, 1.0),
]


def write_add(work):
    """Fixtures add_*.json / .npz (target symmetry after perform(btc, d),
    every orbit; dense inputs, target and result) and badd_*."""
    ctx = lt.Context(-1)
    for i, (name, base, kind, d) in enumerate(ADD_CASES):
        prob, dense, meta, result = P.load_case(os.path.join(HERE, base))
        prob = copy.deepcopy(prob)
        prob["C"] = target_like(prob["A"], kind)
        prob["add"] = d
        W.register_tables(prob)
        bt = W.build_tensor(ctx, prob["C"])
        W.fill_random(bt, prob["C"], 7000 + i)
        dense = dict(dense)
        dense["C"] = bt.to_dense()
        ref = P.run_ref(prob, dense, os.path.join(work, name))
        out = {"prob": prob, "symt_add": {str(k): list(v) for k, v in ref["symt_add"].items()}}
        with open(os.path.join(HERE, name + ".json"), "w") as f:
            json.dump(out, f)
        arrays = dict(dense)
        arrays["result"] = ref["add_result"]
        np.savez_compressed(os.path.join(HERE, name + ".npz"), **arrays)
        print("%-26s canonical allowed %d" % (name, sum(
            1 for k, v in ref["symt_add"].items() if v[0] == k and v[1])))
    base = P.prob_ladder()
    for i, (name, ka, kb, perm, ca, cb) in enumerate(BADD_CASES):
        prob = {"A": target_like(base["A"], ka), "B": target_like(base["A"], kb),
                "tables": base.get("tables", []),
                "btod_add": {"perm": perm, "ca": ca, "cb": cb}}
        W.register_tables(prob)
        dense = {}
        for k, nm in enumerate(("A", "B")):
            bt = W.build_tensor(ctx, prob[nm])
            W.fill_random(bt, prob[nm], 7100 + 10 * i + k)
            dense[nm] = bt.to_dense()
        shape = tuple(prob["A"]["dims"])
        ref = P.ref_io.run_ref(prob, os.path.join(work, name), shape, dense=dense)
        out = {"prob": prob, "schedule": ref["schedule"],
               "symt": {str(k): list(v) for k, v in ref["symt"].items()}}
        with open(os.path.join(HERE, name + ".json"), "w") as f:
            json.dump(out, f)
        arrays = dict(dense)
        arrays["result"] = ref["result"]
        np.savez_compressed(os.path.join(HERE, name + ".npz"), **arrays)
        print("%-26s schedule %d" % (name, len(ref["schedule"])))


def main():
    if not P.ref_available():
        print("oracle/_ref/ref_driver is not built (make -C oracle ref)")
        return 1
    if "--symm-only" in sys.argv or "--c3-only" in sys.argv or "--add-only" in sys.argv:
        work = tempfile.mkdtemp(prefix="lt_golden_")
        try:
            if "--symm-only" in sys.argv:
                write_symm(work)
            if "--c3-only" in sys.argv:
                write_c3(work)
            if "--add-only" in sys.argv:
                write_add(work)
        finally:
            shutil.rmtree(work, ignore_errors=True)
        return 0
    ctx = lt.Context(-1)
    work = tempfile.mkdtemp(prefix="lt_golden_")
    try:
        for i, (name, prob) in enumerate(cases()):
            dense = P.make_data(ctx, prob, 1000 + i)
            ref = P.run_ref(prob, dense, os.path.join(work, name))
            P.save_case(os.path.join(HERE, name), prob, dense, ref)
            print("%-22s schedule %4d" % (name, len(ref["schedule"])))
        write_symm(work)
        write_c3(work)
        write_add(work)
    finally:
        shutil.rmtree(work, ignore_errors=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
