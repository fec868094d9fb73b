"""libtensor's own btod_contract2 and letter expressions against the MI355X
adapter btod_contract2_amd (libtensor_amd/integration/btod_contract2_amd.h),
both built from the reference sources by `make -C oracle seam` into
oracle/_ref/seam_driver (skipped where it is not built).

Host side (no GPU, host-only context): block index space, the libtensor
orbits of both get_symmetry()s, the library's own orbit of every result block
(verify_structure) and the schedules are identical.  GPU: perform(btc),
perform(btc, 0.5) into a target of another symmetry, compute_block, and
R(...) = contract(...) / R(...) += contract(...) through libtensor's
expression evaluator with the adapter switched in (use_libxm(true), see
oracle/seam/) and out agree within 1e-12 (max|d| / max|ref|)."""
import os
import subprocess
import tempfile

import pytest

import golden_cases as G
import ref_io

SEAM = os.path.join(os.path.dirname(ref_io.REF_DRIVER), "seam_driver")
LIBDIR = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                      "libtensor_amd", "_lib")
TOL = 1e-12

needs_seam = pytest.mark.skipif(not os.path.exists(SEAM),
                                reason="oracle/_ref/seam_driver not built (make -C oracle seam)")


def _cases():
    out = []
    for n in G.names():
        prob = G.load(n)[0]
        if not prob.get("self"):
            out.append(n)
    return out + G.add_names()


def _load(name):
    if name.startswith("add_"):
        prob, dense, meta, result = G.load_add(name)
    else:
        prob, dense, meta, result = G.load(name)
    return prob, dense


def _run(name, host_only):
    prob, dense = _load(name)
    with tempfile.TemporaryDirectory() as d:
        spec = os.path.join(d, "problem.spec")
        ref_io.write_spec(prob, spec, dense=dense)
        out = os.path.join(d, "seam.txt")
        env = dict(os.environ, LT_AMD_DEVICE="-1" if host_only else "0")
        # the library of this tree (ahead of the driver's run path)
        env["LD_LIBRARY_PATH"] = LIBDIR + ":" + env.get("LD_LIBRARY_PATH", "")
        cmd = [SEAM, spec, out] + (["--host-only"] if host_only else [])
        r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True,
                           env=env, timeout=300)
        assert r.returncode == 0, r.stdout
        res = {}
        with open(out) as f:
            for line in f:
                k = line.split()
                res[k[0]] = k[1:]
    return res


@needs_seam
@pytest.mark.parametrize("name", _cases())
def test_structure_host_only(name):
    r = _run(name, True)
    assert r["bis_match"] == ["1"]
    assert r["orbit_mismatch"] == ["0"]
    assert r["lib_orbit_mismatch"] == ["0"]
    assert r["sched"][2] == "1" and r["sched"][0] == r["sched"][1]


@pytest.mark.gpu
@needs_seam
@pytest.mark.parametrize("name", _cases())
def test_results_on_the_gpu(name):
    print("seam", name, flush=True)
    r = _run(name, False)
    assert r["lib_orbit_mismatch"] == ["0"] and r["sched"][2] == "1"
    for key in ("perform_err", "perform_add_err", "compute_block_err"):
        assert float(r[key][0]) <= TOL, (key, r[key])
    if "expr_assign_err" in r:
        assert float(r["expr_direct_err"][0]) <= TOL
        assert float(r["expr_assign_err"][0]) <= TOL
        assert float(r["expr_add_err"][0]) <= TOL
