"""btod_symmetrize2 / btod_symmetrize3 host side (no GPU): the result
symmetry (canonical block, orbit transformation and allowedness of every
block) and the schedule of B = X + s P(X), and of the sum over the six
exchanges of three index groups, against libtensor's (fixtures symm_* and
symm3_* written by the real reference build, tests/golden/make_golden.py),
including symmetrizations of symmetrizations and the cases libtensor
rejects.  The stored results are checked against the closed forms applied
to the contraction result X (the base case's golden result, or the dense
oracle's), which pins what the fixtures mean."""
import numpy as np
import pytest

import dense_oracle as O
import golden_cases as G
import libtensor_amd as lt
import problems as P


def chain(op, symm):
    """The chain of symmetrizations of a fixture: (perm, symmetric) entries
    are btod_symmetrize2, (perm1, perm2, symmetric) btod_symmetrize3."""
    ops = [op]
    for entry in symm:
        if len(entry) == 3:
            ops.append(lt.btod_symmetrize3(ops[-1], entry[0], entry[1], entry[2]))
        else:
            ops.append(lt.btod_symmetrize2(ops[-1], entry[0], entry[1]))
    return ops


@pytest.mark.parametrize("name", G.symm_names())
def test_symmetry_and_schedule(host_ctx, name):
    prob, dense, bresult, meta, result = G.load_symm(name)
    a, b, op = P.build(host_ctx, prob, dense)
    s = chain(op, meta["symm"])[-1]
    assert sorted(s.get_schedule()) == meta["schedule"]
    for aidx, (ac, al, perm, c) in meta["symt"].items():
        got = s.sym_orbit(int(aidx))
        assert (got[0], got[3], list(got[1]), got[2]) == (ac, bool(al), perm, c), aidx


@pytest.mark.parametrize("name", G.symm_names(errors=True))
def test_rejected_like_libtensor(host_ctx, name):
    prob, dense, bresult, meta, result = G.load_symm(name)
    a, b, op = P.build(host_ctx, prob, dense)
    with pytest.raises(lt.LtError):
        chain(op, meta["symm"])


def closed_form(x, symm):
    for entry in symm:
        s = 1.0 if entry[-1] else -1.0
        if len(entry) == 2:
            x = x + s * np.transpose(x, entry[0])
            continue
        p1, p2 = entry[0], entry[1]

        def t(y, *ps):
            for p in ps:
                y = np.transpose(y, p)
            return y
        x = x + s * t(x, p1) + s * t(x, p2) + t(x, p1, p2) + t(x, p2, p1) + \
            s * t(x, p1, p2, p1)
    return x


@pytest.mark.parametrize("name", G.symm_names())
def test_fixture_is_the_closed_form(name):
    prob, dense, bresult, meta, result = G.load_symm(name)
    if bresult is None:
        bresult = O.btod_contract2_dense(prob["contr"], dense["A"],
                                         dense["A"] if prob.get("self") else dense["B"])
    exp = closed_form(bresult, meta["symm"])
    scale = float(np.max(np.abs(exp))) or 1.0
    assert np.max(np.abs(result - exp)) <= 1e-13 * scale


def test_permutation_must_be_a_pair_permutation(host_ctx):
    prob, dense, bresult, meta, result = G.load_symm("symm_ring_ijab")
    a, b, op = P.build(host_ctx, prob, dense)
    for perm in ([0, 1, 2, 3], [1, 2, 0, 3], [1, 0, 3]):
        with pytest.raises(lt.LtError):
            lt.btod_symmetrize2(op, perm, True)
