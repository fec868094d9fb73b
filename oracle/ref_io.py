"""Test infrastructure only: input/output of the reference driver
(oracle/_ref/ref_driver, the real libtensor path built from its sources).

write_spec() turns a problem description (libtensor_amd.workloads) plus the
canonical blocks (dense arrays or block lists) into the driver's spec file
(oracle/spec_format.h); run_ref() runs the driver and parses what it prints.
Used by the tests and by the CPU-baseline leg of bench.py only.
"""
import os
import struct
import subprocess

import numpy as np

ORACLE = os.path.dirname(os.path.abspath(__file__))
REF_DRIVER = os.path.join(ORACLE, "_ref", "ref_driver")


def ref_available():
    return os.path.exists(REF_DRIVER)


def write_blocks(bt, path):
    """Canonical non-zero blocks of a libtensor_amd block tensor."""
    blks = bt.req_nonzero_blocks()
    with open(path, "wb") as f:
        f.write(struct.pack("<Q", len(blks)))
        for a in blks:
            x = np.ascontiguousarray(bt.get_block(a), dtype=np.float64)
            f.write(struct.pack("<QQ", a, x.size))
            f.write(x.tobytes())


def write_spec(prob, path, dense=None, tensors=None):
    """dense: {"A": ndarray, "B": ndarray[, "C": ndarray]}, or tensors:
    {"A": bt, "B": bt}.  prob["C"] (with dense["C"]) is the target of
    prob["add"] (perform(btc, d) after perform) or the third operand of a
    btod_contract3; prob["btod_add"] makes the problem btod_add(A) + add_op(B)."""
    d = os.path.dirname(path)
    lines = []
    for tid, n, prods in prob.get("tables", []):
        lines.append("table %s %d" % (tid, n))
        for l1, l2, lr in prods:
            lines.append("product %s %d %d %d" % (tid, l1, l2, lr))
    for name in ("A", "B", "C"):
        if name == "B" and prob.get("self"):
            continue
        if name == "C" and "C" not in prob:
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
        if dense is not None:
            f = os.path.join(d, "dense_%s.bin" % name)
            np.ascontiguousarray(dense[name], dtype=np.float64).tofile(f)
            lines.append("dense %s %s" % (name, f))
        else:
            f = os.path.join(d, "blocks_%s.bin" % name)
            write_blocks(tensors[name], f)
            lines.append("blocks %s %s" % (name, f))
    if "btod_add" in prob:
        ba = prob["btod_add"]
        n = len(prob["A"]["dims"])
        lines.append("btod_add %d %s %.17g %.17g" % (
            n, " ".join(str(x) for x in ba.get("perm") or range(n)), ba.get("ca", 1.0),
            ba.get("cb", 1.0)))
    elif "contr3" in prob:
        c = prob["contr3"]
        lines.append("contract3 %d %d %d %d %d %s %s %s %s %.17g" % (
            c["N1"], c["N2"], c["N3"], c["K1"], c["K2"],
            " ".join(str(p) for p in c["permab"]),
            " ".join("%d %d" % (a, b) for a, b in c["pairs1"]),
            " ".join(str(p) for p in c["permd"]),
            " ".join("%d %d" % (a, b) for a, b in c["pairs2"]), c.get("kd", 1.0)))
        if c.get("symab"):
            lines.append("symab %s %d" % (" ".join(str(x) for x in c["symab"][0]),
                                          1 if c["symab"][1] else 0))
    else:
        c = prob["contr"]
        permc = c.get("permc") or list(range(c["N"] + c["M"]))
        lines.append("contract %d %d %d %s %s %.17g %.17g %.17g" % (
            c["N"], c["M"], c["K"], " ".join(str(p) for p in permc),
            " ".join("%d %d" % (a, b) for a, b in c["pairs"]),
            c.get("ka", 1.0), c.get("kb", 1.0), c.get("kc", 1.0)))
    for entry in prob.get("symm", []):
        if len(entry) == 3:
            perm, perm2, symm = entry
            lines.append("symmetrize3 %s %s %d" % (" ".join(str(x) for x in perm),
                                                  " ".join(str(x) for x in perm2),
                                                  1 if symm else 0))
        else:
            perm, symm = entry
            lines.append("symmetrize2 %s %d" % (" ".join(str(x) for x in perm),
                                                1 if symm else 0))
    if prob.get("self"):
        lines.append("self 1")
    if "add" in prob:
        lines.append("add %.17g" % prob["add"])
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")


def parse_output(prefix, result_shape=None):
    out = {"symc": {}, "symt": {}, "schedule": [], "clst": {}, "symt_add": {}}
    symt_key = "symt"
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
        elif l[0] == "add":
            # the symmetry of the target after perform(btc, d) follows
            symt_key = "symt_add"
        elif l[0] == "symt":
            n = int(l[1])
            for k in range(n):
                r = lines[i + k].split()
                out[symt_key][int(r[0])] = (int(r[1]), int(r[2]), [int(x) for x in r[3:-1]],
                                            float(r[-1]))
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
        out["result"] = np.fromfile(prefix + ".result.bin",
                                    dtype=np.float64).reshape(result_shape)
    if result_shape is not None and os.path.exists(prefix + ".add.bin"):
        out["add_result"] = np.fromfile(prefix + ".add.bin",
                                        dtype=np.float64).reshape(result_shape)
    return out


def run_ref(prob, workdir, result_shape, dense=None, tensors=None, extra=(), timeout=900):
    os.makedirs(workdir, exist_ok=True)
    spec = os.path.join(workdir, "problem.spec")
    write_spec(prob, spec, dense=dense, tensors=tensors)
    prefix = os.path.join(workdir, "ref")
    r = subprocess.run([REF_DRIVER, spec, prefix] + list(extra), stdout=subprocess.PIPE,
                       stderr=subprocess.STDOUT, text=True, timeout=timeout)
    if r.returncode != 0:
        raise RuntimeError("ref_driver failed: %s" % r.stdout)
    return parse_output(prefix, result_shape)
