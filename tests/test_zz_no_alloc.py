"""The allocation-free element insertion (rg_matrix_add_element_blocks_noalloc,
through the C ABI) against rg_matrix_add_element_blocks: identical matrices
on two calls in a row, the second with a smaller input, with no change of
the device's free memory; a staging need beyond the scratch array is
refused with RG_VALUE_ERROR."""
import ctypes

import numpy as np
import pytest
import torch

import ripley_hip as R
from ripley_hip import _lib

pytestmark = pytest.mark.gpu


def _cells(ne):
    nn = [n + 1 for n in ne]
    s2 = nn[0] * nn[1]
    out = []
    for k2 in range(ne[2]):
        for k1 in range(ne[1]):
            for k0 in range(ne[0]):
                b = k0 + nn[0] * k1 + s2 * k2
                out.append([b + dx + nn[0] * dy + s2 * dz for dz in (0, 1) for dy in (0, 1) for dx in (0, 1)])
    return np.asarray(out, dtype=np.int32)


def _add(op, nodes, em, noalloc):
    f = _lib.load().rg_matrix_add_element_blocks_noalloc if noalloc else _lib.load().rg_matrix_add_element_blocks
    op._bind()
    nodes = np.ascontiguousarray(nodes, dtype=np.int32)
    em = np.ascontiguousarray(em, dtype=np.float64)
    return f(op.handle, len(nodes), 8, nodes.ctypes.data_as(ctypes.c_void_p), op.block_size,
             em.ctypes.data_as(ctypes.c_void_p))


@pytest.mark.parametrize("bs,sym", [(1, False), (1, True), (3, False)])
def test_noalloc_matches_the_allocating_entry_point(bs, sym):
    rng = np.random.default_rng(10 * bs + sym)
    calls = [((6, 5, 4), None), ((6, 5, 4), 17)]   # the second call: a smaller input (17 elements)
    ne = calls[0][0]
    dom = R.Brick(*ne)
    nodes = _cells(ne)
    em_all = rng.standard_normal((len(nodes), 8, 8, bs, bs))
    if sym:
        em_all = 0.5 * (em_all + em_all.transpose(0, 2, 1, 4, 3))
    em_all = em_all.reshape(len(nodes), -1)
    ref = dom.newOperator(bs, symmetric=sym)
    new = dom.newOperator(bs, symmetric=sym)
    torch.cuda.synchronize()
    for _, count in calls:
        sel = slice(None) if count is None else slice(3, 3 + count)
        assert _add(ref, nodes[sel], em_all[sel], False) == 0, _lib.load().rg_last_error()
        torch.cuda.synchronize()
        free0 = torch.cuda.mem_get_info()[0]
        rc = _add(new, nodes[sel], em_all[sel], True)
        torch.cuda.synchronize()
        free1 = torch.cuda.mem_get_info()[0]
        assert rc == 0, _lib.load().rg_last_error()
        assert free1 == free0
        # same kernels in the same order: the same bits after every call
        assert np.array_equal(ref.exportCSR()[2], new.exportCSR()[2])


def test_noalloc_refuses_inputs_beyond_the_scratch_array():
    ne = (16, 16, 16)
    dom = R.Brick(*ne)
    op = dom.newOperator(4)
    nodes = _cells(ne)
    em = np.zeros((len(nodes), 8 * 8 * 16))   # 4096 x 8 KiB > 16 MiB
    assert em.nbytes > _lib.RG_SCRATCH_BYTES
    assert _add(op, nodes, em, True) == _lib.RG_VALUE_ERROR
    assert b"scratch" in _lib.load().rg_last_error()
    # within the limit the same call goes through
    assert _add(op, nodes[:512], em[:512], True) == 0
