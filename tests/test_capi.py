"""The C ABI library loads and exports every symbol include/lt_amd.h declares;
host-side error behaviour (no GPU needed)."""
import ctypes
import os
import re

import numpy as np
import pytest

import libtensor_amd as lt
from libtensor_amd import _capi

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_symbols():
    with open(os.path.join(REPO, "include", "lt_amd.h")) as f:
        txt = f.read()
    return sorted(set(re.findall(r"\b(lt_[a-z0-9_]+)\s*\(", txt)))


def test_library_exports_all_declared_symbols():
    lib = ctypes.CDLL(_capi.LIB_PATH)
    syms = header_symbols()
    assert len(syms) >= 50
    missing = [s for s in syms if not hasattr(lib, s)]
    assert missing == []
    # every declared symbol has a ctypes signature in the binding
    assert sorted(set(syms) - set(_capi.declared_symbols())) == []


def test_errors_are_codes_not_exceptions(host_ctx):
    bt = lt.block_tensor(host_ctx, [4, 4])
    with pytest.raises(lt.LtError) as e:
        bt.insert(lt.se_perm([0, 0], 1.0))
    assert e.value.code == -1
    # antisymmetric identity is a symmetry error
    with pytest.raises(lt.LtError) as e:
        bt.insert(lt.se_perm([0, 1], -1.0))
    assert e.value.code == -3
    # a permutation that changes the block index space is refused
    bis = lt.block_index_space([4, 6])
    bt2 = lt.block_tensor(host_ctx, bis)
    with pytest.raises(lt.LtError):
        bt2.insert(lt.se_perm([1, 0], 1.0))


def test_bad_dimensions(host_ctx):
    a = lt.block_tensor(host_ctx, [3, 4])
    b = lt.block_tensor(host_ctx, [5, 6])
    c = lt.contraction2(1, 1, 1)
    c.contract(1, 0)
    with pytest.raises(lt.LtError) as e:
        lt.btod_contract2(c, a, b)
    assert e.value.code == -2
    c2 = lt.contraction2(1, 1, 1)
    with pytest.raises(lt.LtError):
        lt.btod_contract2(c2, a, b)   # incomplete contraction


def test_device_ops_refuse_on_host_context(host_ctx):
    a = lt.block_tensor(host_ctx, [4, 4])
    a.from_dense(np.ones((4, 4)))
    c = lt.contraction2(1, 1, 1)
    c.contract(1, 0)
    op = lt.btod_contract2(c, a, a)
    out = lt.block_tensor(host_ctx, [4, 4])
    with pytest.raises(lt.LtError) as e:
        op.perform(out)
    assert e.value.code == -4
    with pytest.raises(lt.LtError):
        lt.device_array(host_ctx, (4,))
    with pytest.raises(lt.LtError) as e:
        lt.server(host_ctx)
    assert e.value.code == -4


def test_block_access_and_orbits(host_ctx):
    bis = lt.block_index_space([5, 5]).split([1, 1], 2)
    bt = lt.block_tensor(host_ctx, bis)
    bt.insert(lt.se_perm([1, 0], -1.0))
    assert bt.bidims == (2, 2)
    assert bt.block_dims((0, 1)) == (2, 3)
    ac, perm, coeff, allowed = bt.orbit((1, 0))
    assert ac == bt.abs_index((0, 1)) and perm == (1, 0) and coeff == -1.0 and allowed
    x = np.arange(6.0).reshape(2, 3)
    bt.set_block((0, 1), x)
    with pytest.raises(lt.LtError):
        bt.set_block((1, 0), x.T)     # not canonical
    d = bt.to_dense()
    assert np.array_equal(d[0:2, 2:5], x)
    assert np.array_equal(d[2:5, 0:2], -x.T)
    assert bt.req_nonzero_blocks() == [1]


def test_large_blocks_refused_before_device_work(host_ctx):
    """The GEMM kernels address blocks with 32-bit byte offsets: a plan
    with a block of 2^29 or more elements is refused with LT_ERR_PARAM (on
    any context, before device work), one element fewer is accepted up to
    the device check."""
    c = lt.contraction2(1, 1, 1)
    c.contract(1, 0)
    big = lt.block_tensor(host_ctx, [1 << 15, 1 << 14])       # one 2^29 block
    b = lt.block_tensor(host_ctx, [1 << 14, 2])
    op = lt.btod_contract2(c, big, b)
    out = lt.block_tensor(host_ctx, [1 << 15, 2])
    with pytest.raises(lt.LtError) as e:
        op.perform(out)
    assert e.value.code == -1 and "2^29" in str(e.value)
    with pytest.raises(lt.LtError) as e:
        op.perform_add(out, 1.0)
    assert e.value.code == -1
    # split in two along the rows: blocks of 2^28 elements pass the guard
    # and stop at the device check
    bis = lt.block_index_space([1 << 15, 1 << 14]).split([1, 0], 1 << 14)
    a2 = lt.block_tensor(host_ctx, bis)
    op2 = lt.btod_contract2(c, a2, b)
    out2 = lt.block_tensor(host_ctx, lt.block_index_space([1 << 15, 2]).split([1, 0], 1 << 14))
    with pytest.raises(lt.LtError) as e:
        op2.perform(out2)
    assert e.value.code == -4
