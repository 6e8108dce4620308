"""Pins the oracle's boundary (FunctionOnBoundary) and Dirac assembly:
closed forms of the face mass matrix and load, the reference's hand-expanded
constants for expanded data (DefaultAssembler2D.cpp:605-663,
DefaultAssembler3D.cpp:2237-2299), the reduced one-point rule (:1018-1059),
the face element numbering, and assemblePDEDirac (RipleyDomain.cpp:
1756-1800)."""
import numpy as np
import pytest

import oracle as O

S3 = np.sqrt(3.0)


def _face_counts(ne):
    dim = len(ne)
    out = []
    for f in range(2 * dim):
        ax = f // 2
        out.append(int(np.prod([ne[a] for a in range(dim) if a != ax])))
    return out


@pytest.mark.parametrize("ne,length", [((3, 2), (1.0, 2.0)), ((2, 3, 2), (1.0, 0.5, 2.0))])
def test_totals_are_boundary_measure(ne, length):
    dim = len(ne)
    ptr, idx, val, rhs = O.assemble_boundary(dim, ne, length, d=2.5, y=1.5)
    if dim == 2:
        meas = 2 * (length[0] + length[1])
    else:
        l0, l1, l2 = length
        meas = 2 * (l0 * l1 + l0 * l2 + l1 * l2)
    assert abs(val.sum() - 2.5 * meas) < 1e-12 * meas
    assert abs(rhs.sum() - 1.5 * meas) < 1e-12 * meas
    D = O.csr_to_dense(1, ptr, idx, val)
    assert np.allclose(D, D.T, atol=1e-15 * abs(D).max(), rtol=0)
    # interior nodes carry nothing
    N = len(ptr) - 1
    nn = [n + 1 for n in ne]
    coords = np.stack(np.unravel_index(np.arange(N), nn[::-1])[::-1], axis=1)
    inner = np.all((coords > 0) & (coords < np.array(nn) - 1), axis=1)
    assert np.all(D[inner] == 0) and np.all(rhs[inner] == 0)


def test_2d_face_mass_closed_form():
    # Rectangle(2, 1), length (1, 1): left face is one edge of length 1
    ptr, idx, val, _ = O.assemble_boundary(2, (2, 1), (1.0, 1.0), d=1.0)
    D = O.csr_to_dense(1, ptr, idx, val)
    # node 0 (corner): left edge h=1 and bottom edge h=0.5: 1/3 + 0.5/3
    assert abs(D[0, 0] - 0.5) < 1e-15
    assert abs(D[0, 3] - 1.0 / 6.0) < 1e-15     # along the left edge
    assert abs(D[0, 1] - 0.5 / 6.0) < 1e-15     # along the bottom edge
    assert D[0, 4] == 0.0                       # diagonal across the element


def _single_face_coef(dim, ne, face, elem, values, npts):
    """Expanded coefficient that is zero except on one face element."""
    F = sum(_face_counts(ne))
    c = np.zeros((F, npts))
    off = sum(_face_counts(ne)[:face])
    c[off + elem] = values
    return c


def test_2d_expanded_constants_of_the_reference():
    d0, d1 = 1.7, 0.4
    d = _single_face_coef(2, (1, 1), 0, 0, [d0, d1], 2)
    ptr, idx, val, _ = O.assemble_boundary(2, (1, 1), (1.0, 1.0), d=d, expanded=("d",))
    D = O.csr_to_dense(1, ptr, idx, val)
    h = 1.0
    w2 = h / 12
    w0, w1 = w2 * (S3 + 2), w2 * (-S3 + 2)
    # DefaultAssembler2D.cpp:640-643 (face 0, local nodes 0 and 2)
    assert abs(D[0, 0] - (d0 * w0 + d1 * w1)) < 1e-15
    assert abs(D[2, 0] - w2 * (d0 + d1)) < 1e-15
    assert abs(D[2, 2] - (d0 * w1 + d1 * w0)) < 1e-15
    y0, y1 = 0.3, -1.1
    y = _single_face_coef(2, (1, 1), 0, 0, [y0, y1], 2)
    _, _, _, rhs = O.assemble_boundary(2, (1, 1), (1.0, 1.0), y=y, expanded=("y",))
    w3, w4 = w2 * (S3 + 3), w2 * (-S3 + 3)
    assert abs(rhs[0] - (w3 * y0 + w4 * y1)) < 1e-15     # :663
    assert abs(rhs[2] - (w3 * y1 + w4 * y0)) < 1e-15


def test_3d_expanded_constants_of_the_reference():
    dq = np.array([1.3, 0.2, 2.1, 0.7])
    ne, ln = (1, 1, 1), (1.0, 2.0, 0.5)
    d = _single_face_coef(3, ne, 0, 0, dq, 4)
    ptr, idx, val, _ = O.assemble_boundary(3, ne, ln, d=d, expanded=("d",))
    D = O.csr_to_dense(1, ptr, idx, val)
    w2 = ln[1] * ln[2] / 144
    w0, w1 = w2 * (-S3 + 2), w2 * (S3 + 2)
    w3, w4 = w2 * (-4 * S3 + 7), w2 * (4 * S3 + 7)
    d_0, d_1, d_2, d_3 = dq
    # DefaultAssembler3D.cpp:2284-2300 (face 0: local nodes 0, 2, 4, 6)
    assert abs(D[0, 0] - (d_0 * w4 + d_3 * w3 + w2 * (d_1 + d_2))) < 1e-15
    assert abs(D[2, 0] - (w0 * (d_2 + d_3) + w1 * (d_0 + d_1))) < 1e-15
    assert abs(D[6, 0] - w2 * dq.sum()) < 1e-15
    assert abs(D[6, 6] - (d_0 * w3 + d_3 * w4 + w2 * (d_1 + d_2))) < 1e-15


def test_expanded_equal_values_match_element_data():
    ne = (3, 2, 2)
    F = sum(_face_counts(ne))
    rng = np.random.default_rng(1)
    dv, yv = rng.uniform(0.5, 2, F), rng.standard_normal(F)
    _, _, v1, r1 = O.assemble_boundary(3, ne, (1, 1, 1), d=dv, y=yv)
    _, _, v2, r2 = O.assemble_boundary(3, ne, (1, 1, 1), d=np.repeat(dv[:, None], 4, 1),
                                       y=np.repeat(yv[:, None], 4, 1), expanded=("d", "y"))
    assert np.allclose(v1, v2, rtol=1e-14, atol=1e-16) and np.allclose(r1, r2, rtol=1e-14, atol=1e-16)


def test_reduced_one_point_rule():
    # DefaultAssembler2D.cpp:1018-1059: EM_S = d h/4 for every face-node pair, EM_F = y h/2
    ptr, idx, val, rhs = O.assemble_boundary(2, (1, 1), (2.0, 1.0), d=_single_face_coef(2, (1, 1), 0, 0, 3.0, 1)[:, 0],
                                             y=_single_face_coef(2, (1, 1), 0, 0, 5.0, 1)[:, 0], reduced=True)
    D = O.csr_to_dense(1, ptr, idx, val)
    h = 1.0    # face 0 spans x1
    assert abs(D[0, 0] - 3.0 * h / 4) < 1e-15 and abs(D[0, 2] - 3.0 * h / 4) < 1e-15
    assert abs(rhs[0] - 5.0 * h / 2) < 1e-15 and abs(rhs[2] - 5.0 * h / 2) < 1e-15


def test_face_element_numbering_3d():
    # face 2 (x1 = 0) numbers its elements INDEX2(k0, k2, NE0)
    # (DefaultAssembler3D.cpp: face 2 loop); load one element and find its nodes
    ne = (3, 2, 4)
    counts = _face_counts(ne)
    k0, k2 = 2, 3
    y = _single_face_coef(3, ne, 2, k0 + ne[0] * k2, 1.0, 1)[:, 0]
    _, _, _, rhs = O.assemble_boundary(3, ne, (1, 1, 1), y=y)
    nn = [n + 1 for n in ne]
    loaded = sorted(np.nonzero(rhs)[0])
    expect = sorted(x + nn[0] * (0 + nn[1] * z) for x in (k0, k0 + 1) for z in (k2, k2 + 1))
    assert loaded == expect
    assert sum(counts) == 2 * (2 * 4 + 3 * 4 + 3 * 2)


def test_dirac_points():
    ptr, idx = O.pattern(2, (2, 2))
    n = 2
    val = np.zeros(len(idx) * n * n)
    rhs = np.zeros(9 * n)
    nodes = [4, 0, 4]
    d = np.array([[[1.0, 2.0], [3.0, 4.0]], [[5.0, 0.0], [0.0, 5.0]], [[0.5, 0.0], [0.0, 0.5]]])
    y = np.array([[1.0, -1.0], [2.0, 0.0], [0.25, 0.25]])
    # d blocks column major per point: d[p][k + n m]
    dcm = np.stack([b.T.reshape(-1) for b in d])
    O.add_dirac(n, ptr, idx, val, nodes, d=dcm, y=y.reshape(-1), rhs=rhs)
    D = O.csr_to_dense(n, ptr, idx, val)
    assert np.allclose(D[8:10, 8:10], d[0] + d[2]) and np.allclose(D[0:2, 0:2], d[1])
    assert np.allclose(rhs[8:10], [1.25, -0.75]) and np.allclose(rhs[0:2], [2.0, 0.0])
    assert abs(D).sum() == pytest.approx(abs(d).sum())
