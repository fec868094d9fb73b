"""Problem helpers of the parity tests: workloads from libtensor_amd.workloads,
reference-driver I/O from oracle/ref_io.py, golden fixture files."""
import json
import os

import numpy as np

import libtensor_amd as lt
from libtensor_amd import workloads as W
import ref_io

REF_DRIVER = ref_io.REF_DRIVER
C2V = W.C2V
ref_available = ref_io.ref_available
prob_ring = W.ring
prob_ladder = W.ladder
build_tensor = W.build_tensor
make_result_tensor = W.result_tensor


def make_data(ctx, prob, seed):
    """Symmetric random inputs from the library's btod_random analogue, as
    dense arrays (re-imported on both sides, so canonical blocks that are
    exactly zero are dropped by both)."""
    W.register_tables(prob)
    out = {}
    for name in ("A", "B"):
        if name == "B" and prob.get("self"):
            continue
        bt = W.build_tensor(ctx, prob[name])
        W.fill_random(bt, prob[name], seed + (0 if name == "A" else 7))
        out[name] = bt.to_dense()
    return out


def build(ctx, prob, dense):
    """Block tensors holding the dense data and the btod_contract2 op."""
    W.register_tables(prob)
    a = W.build_tensor(ctx, prob["A"])
    a.from_dense(dense["A"])
    if prob.get("self"):
        b = a
    else:
        b = W.build_tensor(ctx, prob["B"])
        b.from_dense(dense["B"])
    c = prob["contr"]
    op = lt.btod_contract2(W.make_contraction(prob), a, b, c.get("ka", 1.0),
                           c.get("kb", 1.0), c.get("kc", 1.0))
    return a, b, op


def build_op(ctx, prob, a, b):
    """The btod_contract2 op of a problem on existing block tensors."""
    c = prob["contr"]
    return lt.btod_contract2(W.make_contraction(prob), a, b if not prob.get("self") else a,
                             c.get("ka", 1.0), c.get("kb", 1.0), c.get("kc", 1.0))


def result_dims(prob):
    order = W.result_order(prob)
    dims = []
    for t, i in order:
        td = prob["A"] if (t == "A" or prob.get("self")) else prob["B"]
        dims.append(td["dims"][i])
    return tuple(dims), order


def run_ref(prob, dense, workdir, extra=()):
    return ref_io.run_ref(prob, workdir, result_dims(prob)[0], dense=dense, extra=extra)


def clst_as_strings(pairs):
    """Contraction list from libtensor_amd in the reference driver's format."""
    out = []
    for aia, acia, pa, ca, aib, acib, pb, cb in pairs:
        out.append([str(aia), str(acia)] + [str(x) for x in pa] + ["%.17g" % ca] +
                   [str(aib), str(acib)] + [str(x) for x in pb] + ["%.17g" % cb])
    return out


def normalize_ref_clst(rows):
    return [[("%.17g" % float(x)) if ("." in x or "e" in x) else x for x in r] for r in rows]


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
    return meta["prob"], dense, meta, result


# ---------------------------------------------------------------------------
# btod_contract3 problems: tensors A, B, C and D described by index ids
# (prob["idx"][name] lists the id of every dimension; ids are letters, so the
# dense result is an einsum).

def make_data3(ctx, prob, seed):
    W.register_tables(prob)
    out = {}
    for k, name in enumerate(("A", "B", "C")):
        bt = W.build_tensor(ctx, prob[name])
        W.fill_random(bt, prob[name], seed + 7 * k)
        out[name] = bt.to_dense()
    return out


def build3(ctx, prob, dense):
    """Block tensors A, B, C holding the dense data and the btod_contract3 op."""
    W.register_tables(prob)
    ts = []
    for name in ("A", "B", "C"):
        t = W.build_tensor(ctx, prob[name])
        t.from_dense(dense[name])
        ts.append(t)
    c = prob["contr3"]
    c1 = lt.contraction2(c["N1"], c["N2"] + c["K2"], c["K1"], c["permab"])
    for ia, ib in c["pairs1"]:
        c1.contract(ia, ib)
    c2 = lt.contraction2(c["N1"] + c["N2"], c["N3"], c["K2"], c["permd"])
    for ia, ib in c["pairs2"]:
        c2.contract(ia, ib)
    sym = c.get("symab")
    op = lt.btod_contract3(c1, c2, ts[0], ts[1], ts[2], c.get("kd", 1.0),
                           sympermab=None if sym is None else sym[0],
                           symmab=True if sym is None else sym[1])
    return ts, op


def result_tensor3(ctx, prob):
    return W.build_tensor(ctx, prob["D"])


def einsum3(prob, dense):
    """kd (A B) C of the dense inputs; with "symab" (btod_symcontract3) AB is
    replaced by AB + s P(AB) first."""
    ix = prob["idx"]
    c = prob["contr3"]
    ab = np.einsum("%s,%s->%s" % ("".join(ix["A"]), "".join(ix["B"]), "".join(ix["AB"])),
                   dense["A"], dense["B"], optimize=True)
    if c.get("symab"):
        perm, symm = c["symab"]
        ab = ab + (1.0 if symm else -1.0) * np.transpose(ab, perm)
    return c.get("kd", 1.0) * np.einsum("%s,%s->%s" % ("".join(ix["AB"]), "".join(ix["C"]),
                                                       "".join(ix["D"])),
                                        ab, dense["C"], optimize=True)
