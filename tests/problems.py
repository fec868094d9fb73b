"""Problem descriptions for the parity tests.

A problem is a plain dict (JSON-serialisable apart from the dense input
arrays) describing two block tensors with their splits and symmetry elements,
and a contraction.  The same description is built in libtensor_amd and
written as a spec file for the reference driver (oracle/_ref/ref_driver).
"""
import json
import os
import subprocess

import numpy as np

import libtensor_amd as lt

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF_DRIVER = os.path.join(REPO, "oracle", "_ref", "ref_driver")

C2V = ("c2v", 4, [(1, 1, 0), (1, 2, 3), (1, 3, 2), (2, 2, 0), (2, 3, 1), (3, 3, 0)])


def ref_available():
    return os.path.exists(REF_DRIVER)


# ---------------------------------------------------------------------------
# building

def _tables(prob):
    for tid, n, prods in prob.get("tables", []):
        t = lt.product_table(tid, n)
        for l1, l2, lr in prods:
            t.add_product(l1, l2, lr)


def build_tensor(ctx, td):
    bis = lt.block_index_space(td["dims"])
    for mask, pos in td.get("splits", []):
        bis.split(mask, pos)
    bt = lt.block_tensor(ctx, bis)
    for el in td.get("elems", []):
        if el[0] == "perm":
            bt.insert(lt.se_perm(el[1], el[2]))
        elif el[0] == "part":
            sp = lt.se_part(el[1])
            for op in el[2]:
                if op[0] == "map":
                    sp.add_map(op[1], op[2], op[3])
                else:
                    sp.mark_forbidden(op[1])
            bt.insert(sp)
        else:
            sl = lt.se_label(el[1])
            for mask, blk, lab in el[2]:
                sl.assign(mask, blk, lab)
            if el[3] is not None:
                sl.set_rule(el[3])
            bt.insert(sl)
    return bt


def make_data(ctx, prob, seed):
    """Fills A and B with symmetric random data through the library's
    btod_random analogue and returns the dense inputs (canonical blocks that
    are exactly zero are dropped on both sides by re-importing)."""
    _tables(prob)
    out = {}
    for name in ("A", "B"):
        if name == "B" and prob.get("self"):
            continue
        td = prob[name]
        bt = build_tensor(ctx, td)
        bt.random(seed + (0 if name == "A" else 7), td.get("dist", "uniform"),
                  td.get("scale", 1.0), td.get("scale_dims"))
        out[name] = bt.to_dense()
    return out


def build(ctx, prob, dense):
    """Block tensors with the given dense data and the btod_contract2 op."""
    _tables(prob)
    a = build_tensor(ctx, prob["A"])
    a.from_dense(dense["A"])
    if prob.get("self"):
        b = a
    else:
        b = build_tensor(ctx, prob["B"])
        b.from_dense(dense["B"])
    c = prob["contr"]
    contr = lt.contraction2(c["N"], c["M"], c["K"], c.get("permc"))
    for ia, ib in c["pairs"]:
        contr.contract(ia, ib)
    op = lt.btod_contract2(contr, a, b, c.get("ka", 1.0), c.get("kb", 1.0), c.get("kc", 1.0))
    return a, b, op


def result_dims(prob):
    c = prob["contr"]
    na = len(prob["A"]["dims"])
    conn_src = []
    # C order: remaining A indexes then remaining B indexes, permuted by permc
    ka = set(p[0] for p in c["pairs"])
    kb = set(p[1] for p in c["pairs"])
    bt = prob["A"] if not prob.get("self") else prob["A"]
    bdesc = prob["B"] if not prob.get("self") else prob["A"]
    rem = [("A", i) for i in range(na) if i not in ka] + \
          [("B", i) for i in range(len(bdesc["dims"])) if i not in kb]
    permc = c.get("permc") or list(range(len(rem)))
    seq = [rem[permc[i]] for i in range(len(rem))]
    dims = []
    for t, i in seq:
        dims.append((bt if t == "A" else bdesc)["dims"][i])
    return tuple(dims), seq


def result_splits(prob):
    _, seq = result_dims(prob)
    out = []
    for t, i in seq:
        td = prob["A"] if (t == "A" or prob.get("self")) else prob["B"]
        pos = [pp for mask, pp in td.get("splits", []) if mask[i]]
        out.append(sorted(set(pos)))
    return out


def make_result_tensor(ctx, prob):
    dims, _ = result_dims(prob)
    bis = lt.block_index_space(dims)
    for d, pos in enumerate(result_splits(prob)):
        for p in pos:
            m = [0] * len(dims)
            m[d] = 1
            bis.split(m, p)
    return lt.block_tensor(ctx, bis)


# ---------------------------------------------------------------------------
# reference driver

def write_spec(prob, dense, path):
    d = os.path.dirname(path)
    lines = []
    for tid, n, prods in prob.get("tables", []):
        lines.append("table %s %d" % (tid, n))
        for l1, l2, lr in prods:
            lines.append("product %s %d %d %d" % (tid, l1, l2, lr))
    for name in ("A", "B"):
        if name == "B" and prob.get("self"):
            continue
        td = prob[name]
        n = len(td["dims"])
        lines.append("tensor %s %d %s" % (name, n, " ".join(str(x) for x in td["dims"])))
        for mask, pos in td.get("splits", []):
            lines.append("split %s %s %d" % (name, " ".join(str(int(m)) for m in mask), pos))
        for el in td.get("elems", []):
            if el[0] == "perm":
                lines.append("perm %s %s %.17g" % (name, " ".join(str(p) for p in el[1]), el[2]))
            elif el[0] == "part":
                lines.append("part %s %s" % (name, " ".join(str(p) for p in el[1])))
                for op in el[2]:
                    if op[0] == "map":
                        lines.append("map %s %s %.17g" % (" ".join(str(x) for x in op[1]),
                                                          " ".join(str(x) for x in op[2]), op[3]))
                    else:
                        lines.append("forbid %s" % " ".join(str(x) for x in op[1]))
                lines.append("end")
            else:
                lines.append("label %s %s" % (name, el[1]))
                for mask, blk, lab in el[2]:
                    lines.append("assign %s %d %d" % (" ".join(str(int(m)) for m in mask), blk, lab))
                if el[3] is not None:
                    rl = el[3] if isinstance(el[3], list) else [el[3]]
                    lines.append("rule %d %s" % (len(rl), " ".join(str(x) for x in rl)))
                lines.append("end")
        f = os.path.join(d, "dense_%s.bin" % name)
        np.ascontiguousarray(dense[name], dtype=np.float64).tofile(f)
        lines.append("dense %s %s" % (name, f))
    c = prob["contr"]
    permc = c.get("permc") or list(range(c["N"] + c["M"]))
    lines.append("contract %d %d %d %s %s %.17g %.17g %.17g" % (
        c["N"], c["M"], c["K"], " ".join(str(p) for p in permc),
        " ".join("%d %d" % (a, b) for a, b in c["pairs"]),
        c.get("ka", 1.0), c.get("kb", 1.0), c.get("kc", 1.0)))
    if prob.get("self"):
        lines.append("self 1")
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")


def parse_ref_output(prefix, result_shape=None):
    out = {"symc": {}, "schedule": [], "clst": {}}
    with open(prefix + ".txt") as f:
        lines = f.read().split("\n")
    i = 0
    while i < len(lines):
        l = lines[i].split()
        i += 1
        if not l:
            continue
        if l[0] == "symc":
            n = int(l[1])
            for k in range(n):
                a, ac, al = lines[i + k].split()
                out["symc"][int(a)] = (int(ac), int(al))
            i += n
        elif l[0] == "schedule":
            n = int(l[1])
            out["schedule"] = [int(lines[i + k]) for k in range(n)]
            i += n
        elif l[0] == "clst":
            aidx, n = int(l[1]), int(l[2])
            out["clst"][aidx] = [lines[i + k].split() for k in range(n)]
            i += n
        elif l[0].endswith("_ms") or l[0] == "threads":
            out[l[0]] = float(l[1])
    if result_shape is not None and os.path.exists(prefix + ".result.bin"):
        out["result"] = np.fromfile(prefix + ".result.bin", dtype=np.float64).reshape(result_shape)
    return out


def run_ref(prob, dense, workdir, extra=()):
    os.makedirs(workdir, exist_ok=True)
    spec = os.path.join(workdir, "problem.spec")
    write_spec(prob, dense, spec)
    prefix = os.path.join(workdir, "ref")
    r = subprocess.run([REF_DRIVER, spec, prefix] + list(extra), stdout=subprocess.PIPE,
                       stderr=subprocess.STDOUT, text=True, timeout=600)
    if r.returncode != 0:
        raise RuntimeError("ref_driver failed: %s" % r.stdout)
    return parse_ref_output(prefix, result_dims(prob)[0])


def clst_as_strings(pairs):
    """Contraction list from libtensor_amd in the reference driver's format."""
    out = []
    for aia, acia, pa, ca, aib, acib, pb, cb in pairs:
        row = [str(aia), str(acia)] + [str(x) for x in pa] + [repr_coeff(ca)] + \
              [str(aib), str(acib)] + [str(x) for x in pb] + [repr_coeff(cb)]
        out.append(row)
    return out


def repr_coeff(c):
    return "%.17g" % c


def normalize_ref_clst(rows):
    return [[("%.17g" % float(x)) if ("." in x or "e" in x) else x for x in r] for r in rows]


# ---------------------------------------------------------------------------
# problem catalogue

def prob_ring(o=4, v=6, so=None, sv=None):
    """C[i,j,a,b] = sum_{k,c} A[i,k,a,c] B[k,j,c,b] (ring contraction),
    no symmetry, blocks split at the given points."""
    so = so or []
    sv = sv or []
    A = {"dims": [o, o, v, v], "splits": [([1, 1, 0, 0], p) for p in so] +
         [([0, 0, 1, 1], p) for p in sv]}
    B = {"dims": [o, o, v, v], "splits": [([1, 1, 0, 0], p) for p in so] +
         [([0, 0, 1, 1], p) for p in sv]}
    return {"A": A, "B": B,
            "contr": {"N": 2, "M": 2, "K": 2, "permc": [0, 2, 1, 3],
                      "pairs": [(1, 0), (3, 2)]}}


def spin_part(n_occ_dims, n_vir_dims, order, conserve_pairs):
    """se_part over all dims with 2 partitions (alpha/beta); partitions that
    do not conserve spin between the index pairs are forbidden, and every
    allowed partition maps onto its spin-flipped partner with +1."""
    import itertools
    ops = []
    for p in itertools.product([0, 1], repeat=order):
        if not conserve_pairs(p):
            ops.append(("forbid", tuple(p)))
    for p in itertools.product([0, 1], repeat=order):
        if not conserve_pairs(p):
            continue
        q = tuple(1 - x for x in p)
        if p < q:
            ops.append(("map", tuple(p), q, 1.0))
    return ("part", tuple([2] * order), ops)


def prob_ladder(o_blocks=(3, 2), v_blocks=(4, 3, 5), irreps_o=(0, 2), irreps_v=(0, 1, 3),
                spin=True, labels=True, antisym=True):
    """pp-ladder R[i,j,a,b] = sum_{c,d} T[i,j,c,d] V[a,b,c,d] with alpha/beta
    spin partitions (se_part), C2v block labels (se_label) and (ij), (ab),
    (cd) antisymmetry plus the (ab)<->(cd) pair symmetry of V (se_perm)."""
    ob = list(o_blocks)
    vb = list(v_blocks)
    if spin:
        ob = ob + ob
        vb = vb + vb
    o = sum(ob)
    v = sum(vb)
    so = list(np.cumsum(ob)[:-1])
    sv = list(np.cumsum(vb)[:-1])
    lab_o = list(irreps_o) * (2 if spin else 1)
    lab_v = list(irreps_v) * (2 if spin else 1)

    def conserve_T(p):   # (i j c d): spins of (i,j) equal spins of (c,d) as sets
        return p[0] + p[1] == p[2] + p[3]

    def conserve_V(p):   # (a b c d)
        return p[0] + p[1] == p[2] + p[3]

    def mk(dims_kind, conserve, perms):
        dims = [o if k == "o" else v for k in dims_kind]
        splits = []
        om = [1 if k == "o" else 0 for k in dims_kind]
        vm = [1 if k == "v" else 0 for k in dims_kind]
        if any(om):
            splits += [(om, int(p)) for p in so]
        if any(vm):
            splits += [(vm, int(p)) for p in sv]
        elems = []
        for pr, c in perms:
            elems.append(("perm", pr, c))
        if spin:
            elems.append(spin_part(0, 0, 4, conserve))
        if labels:
            assigns = []
            om_ = [1 if k == "o" else 0 for k in dims_kind]
            vm_ = [1 if k == "v" else 0 for k in dims_kind]
            if any(om_):
                for b, l in enumerate(lab_o):
                    assigns.append((om_, b, l))
            if any(vm_):
                for b, l in enumerate(lab_v):
                    assigns.append((vm_, b, l))
            elems.append(("label", "c2v", assigns, [0]))
        return {"dims": dims, "splits": splits, "elems": elems}

    perms_T = [([1, 0, 2, 3], -1.0), ([0, 1, 3, 2], -1.0)] if antisym else []
    perms_V = [([1, 0, 2, 3], -1.0), ([0, 1, 3, 2], -1.0), ([2, 3, 0, 1], 1.0)] \
        if antisym else [([2, 3, 0, 1], 1.0)]
    T = mk("oovv", conserve_T, perms_T)
    T["dist"] = "normal"
    T["scale"] = 0.01
    T["scale_dims"] = (2, 0)
    V = mk("vvvv", conserve_V, perms_V)
    V["dist"] = "normal"
    prob = {"A": T, "B": V,
            "contr": {"N": 2, "M": 2, "K": 2, "permc": None,
                      "pairs": [(2, 2), (3, 3)]}}
    if labels:
        prob["tables"] = [C2V]
    return prob


def save_case(path, prob, dense, ref):
    """Golden fixture: problem (JSON), dense inputs and reference outputs."""
    meta = {"prob": prob, "schedule": ref["schedule"],
            "symc": {str(k): v for k, v in ref["symc"].items()},
            "clst": {str(k): v for k, v in ref["clst"].items()}}
    with open(path + ".json", "w") as f:
        json.dump(meta, f)
    arrays = {"A": dense["A"]}
    if "B" in dense:
        arrays["B"] = dense["B"]
    if "result" in ref:
        arrays["result"] = ref["result"]
    np.savez_compressed(path + ".npz", **arrays)


def load_case(path):
    with open(path + ".json") as f:
        meta = json.load(f)
    z = np.load(path + ".npz", allow_pickle=False)
    dense = {k: z[k] for k in z.files if k in ("A", "B")}
    result = z["result"] if "result" in z.files else None
    prob = meta["prob"]
    # JSON turns tuples into lists; the builders accept both
    return prob, dense, meta, result
