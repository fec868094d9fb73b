"""btod_symmetrize2 host side (no GPU): the result symmetry (canonical block,
orbit transformation and allowedness of every block) and the schedule of
B = X + s P(X) against libtensor's (fixtures symm_* written by the real
reference build, tests/golden/make_golden.py), including symmetrizations of
symmetrizations and the cases libtensor rejects.  The stored results are
checked against the closed form X + s P(X) of the base case's golden
contraction result, which pins what the fixtures mean."""
import numpy as np
import pytest

import golden_cases as G
import libtensor_amd as lt
import problems as P


def chain(op, symm):
    ops = [op]
    for perm, s in symm:
        ops.append(lt.btod_symmetrize2(ops[-1], perm, s))
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
    for perm, s in symm:
        x = x + (1.0 if s else -1.0) * np.transpose(x, perm)
    return x


@pytest.mark.parametrize("name", G.symm_names())
def test_fixture_is_the_closed_form(name):
    prob, dense, bresult, meta, result = G.load_symm(name)
    exp = closed_form(bresult, meta["symm"])
    scale = float(np.max(np.abs(exp))) or 1.0
    assert np.max(np.abs(result - exp)) <= 1e-13 * scale


def test_permutation_must_be_a_pair_permutation(host_ctx):
    prob, dense, bresult, meta, result = G.load_symm("symm_ring_ijab")
    a, b, op = P.build(host_ctx, prob, dense)
    for perm in ([0, 1, 2, 3], [1, 2, 0, 3], [1, 0, 3]):
        with pytest.raises(lt.LtError):
            lt.btod_symmetrize2(op, perm, True)
