"""The oracle's stencil (diagonal-offset) layout against the reference's own
known-answer vectors (ripley/test/SystemMatrixTestCase.cpp:41-134, fixture
tests/golden/kat_cds_spmv.json), and the stencil offsets against
getDiagonalIndices."""
import numpy as np
import pytest

import oracle as O
from kat import KAT, kat_elements, kat_expected, kat_x


def test_diagonal_offsets_match_reference_kat():
    # Rectangle(4,3) has nDOF0 = 5: the KAT's offsets are getDiagonalIndices
    off = O.diagonal_indices(2, (5, 4))
    assert list(off) == KAT["diag_off"]
    assert list(O.diagonal_indices(2, (5, 4), upper_only=True)) == KAT["diag_off"][4:]


def test_diagonal_offsets_3d_order():
    off = O.diagonal_indices(3, (4, 5, 6))
    assert len(off) == 27
    assert np.all(np.diff(off) > 0)          # ascending = CSR column order
    assert off[13] == 0 and off[0] == -(4 * 5 + 4 + 1)
    assert len(O.diagonal_indices(3, (4, 5, 6), upper_only=True)) == 14


@pytest.mark.parametrize("bs", [1, 2, 3, 4])
@pytest.mark.parametrize("symmetric", [False, True])
def test_cds_spmv_known_answers(bs, symmetric):
    nodes, em = kat_elements(bs)
    offs = np.asarray(KAT["diag_off"][4:] if symmetric else KAT["diag_off"], dtype=np.int32)
    nd = len(offs)
    values = np.zeros(KAT["rows"] * bs * nd * bs)
    for e in range(8):
        O.cds_add(2, bs, symmetric, nd, values, nodes[e], em[e])
    y = O.cds_spmv(bs, symmetric, KAT["rows"], offs, values, kat_x(bs))
    ref = kat_expected(bs, symmetric)
    assert np.max(np.abs(y - ref)) < KAT["tolerance_lsup"]


def test_cds_kat_detects_transposed_block():
    # planted mistake: swapping the in-block indices must break the known answer
    bs = 2
    nodes, em = kat_elements(bs)
    em_t = em.reshape(8, 16, bs, bs).transpose(0, 1, 3, 2).reshape(8, -1).copy()
    offs = np.asarray(KAT["diag_off"], dtype=np.int32)
    values = np.zeros(KAT["rows"] * bs * 9 * bs)
    for e in range(8):
        O.cds_add(2, bs, False, 9, values, nodes[e], em_t[e])
    y = O.cds_spmv(bs, False, KAT["rows"], offs, values, kat_x(bs))
    assert np.max(np.abs(y - kat_expected(bs, False))) > 1.0
