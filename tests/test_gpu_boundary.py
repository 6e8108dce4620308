"""Boundary (FunctionOnBoundary d, y) and Dirac-point assembly on the GPU
(through the C ABI) against the oracle, plus an exact-solution solve with
Neumann data in the manner of the reference's simple-solve tests
(escriptcore/test/python/test_simplesolve.py:76-160)."""
import numpy as np
import pytest
import torch

import oracle as O
import ripley_hip as R
from gpu_helpers import entrywise_ok, grid, to_dev

pytestmark = pytest.mark.gpu


def _counts(dom):
    return dom.getNumFaceElements(), dom.getNumBoundaryQuadraturePoints()


@pytest.mark.parametrize("dim,ne", [(2, (7, 5)), (3, (5, 4, 3)), (3, (1, 2, 1))])
@pytest.mark.parametrize("n", [1, 2, 3])
@pytest.mark.parametrize("kind", ["element", "expanded", "reduced"])
def test_boundary_assembly_parity(dim, ne, n, kind):
    if n == 3 and dim == 2:
        pytest.skip("3 components on a Rectangle are covered by n=2")
    rng = np.random.default_rng(dim * 100 + n * 10 + len(kind))
    dom, length = grid(dim, ne)
    F, Qb = _counts(dom)
    pd = () if n == 1 else (n, n)
    py = () if n == 1 else (n,)
    lead = (F, Qb) if kind == "expanded" else (F,)
    d = rng.uniform(0.5, 2.0, lead + pd)
    y = rng.standard_normal(lead + py)
    op = dom.newOperator(n)
    rhs = torch.zeros(op.size, dtype=torch.float64, device="cuda")
    cd = R.Coefficient(to_dev(d), pd, F, Qb, dom.device)
    cy = R.Coefficient(to_dev(y), py, F, Qb, dom.device)
    op.assembleBoundary({"d": cd, "y": cy}, rhs, reduced=(kind == "reduced"))
    ptr, idx, val = op.exportCSR()
    optr, oidx, oval, orhs = O.assemble_boundary(dim, ne, length, n=n, d=d, y=y, reduced=(kind == "reduced"),
                                                 expanded=("d", "y") if kind == "expanded" else ())
    assert np.array_equal(ptr, optr) and np.array_equal(idx, oidx)
    ok, worst = entrywise_ok(n, ptr, val, oval)
    assert ok, worst
    scale = np.abs(orhs).max()
    assert np.abs(rhs.cpu().numpy() - orhs).max() <= 1e-13 * scale


def test_boundary_with_volume_terms_and_symmetric_storage():
    rng = np.random.default_rng(5)
    dom, length = grid(3, (6, 5, 4))
    E, Q = dom.getNumElements(), dom.getNumQuadraturePoints()
    F, Qb = _counts(dom)
    kap = rng.uniform(0.5, 2.0, E)
    d = rng.uniform(0.1, 1.0, (F, Qb))
    for sym in (False, True):
        op = dom.newOperator(1, symmetric=sym)
        rhs = torch.zeros(op.size, dtype=torch.float64, device="cuda")
        op.assemble({"A": R.Coefficient(to_dev(kap), (3, 3), E, Q, dom.device, isotropic=True),
                     "Y": R.Coefficient(1.0, (), E, Q, dom.device)}, rhs)
        op.assembleBoundary({"d": R.Coefficient(to_dev(d), (), F, Qb, dom.device)}, rhs)
        ptr, idx, val = op.exportCSR()
        A = kap[:, None, None] * np.eye(3)[None]
        optr, oidx, oval, orhs = O.assemble(3, (6, 5, 4), length, A=A, Y=1.0)
        O.assemble_boundary(3, (6, 5, 4), length, ptr=optr, index=oidx, val=oval, rhs=orhs, d=d, expanded=("d",))
        ok, worst = entrywise_ok(1, ptr, val, oval)
        assert ok, (sym, worst)


@pytest.mark.parametrize("n", [1, 3])
def test_dirac_parity(n):
    dom, length = grid(3, (4, 3, 5))
    rng = np.random.default_rng(n)
    op = dom.newOperator(n)
    rhs = torch.zeros(op.size, dtype=torch.float64, device="cuda")
    nodes = [7, 0, 59, 7]
    d = rng.standard_normal((len(nodes), n, n))
    y = rng.standard_normal((len(nodes), n))
    op.addDirac(nodes, d=d, y=y, rhs=rhs)
    ptr, idx, val = op.exportCSR()
    oval = np.zeros_like(val)
    orhs = np.zeros(op.size)
    dcm = np.stack([b.T.reshape(-1) for b in d])
    O.add_dirac(n, ptr, idx, oval, nodes, d=dcm, y=y.reshape(-1), rhs=orhs)
    assert np.array_equal(val, oval) and np.array_equal(rhs.cpu().numpy(), orhs)
    with pytest.raises(R.RipleyHipError):
        op.addDirac([op.num_rows], y=np.zeros((1, n)), rhs=rhs)


@pytest.mark.parametrize("dim", [2, 3])
def test_neumann_exact_linear_solution(dim):
    # -div grad u = 0, u = 0 on x0 = 0, du/dn = 1 on x0 = L (y = 1 there,
    # 0 elsewhere) => u = x0 exactly for Q1 elements
    ne = (8, 6) if dim == 2 else (6, 5, 4)
    dom = R.Rectangle(*ne, l0=2.0) if dim == 2 else R.Brick(*ne, l0=2.0)
    pde = R.LinearPDE(dom)
    pde.setSymmetryOn()
    X = dom.getBoundaryQuadraturePoints(reduced=True)
    y = (torch.abs(X[:, 0] - 2.0) < 1e-12).to(torch.float64)
    pde.setValue(A=R.kronecker_scaled(1.0), y_reduced=y, q=dom.whereZero(0))
    pde.getSolverOptions().setTolerance(1e-12)
    u = pde.getSolution()
    x0 = dom.getX()[:, 0]
    assert torch.allclose(u, x0, rtol=0, atol=1e-9)


def test_robin_and_dirac_through_linearpde():
    # d u + du/dn = y on the whole boundary with u = 1: y = d; -lap u = 0
    dom = R.Brick(5, 4, 6, diracPoints=[(0.5, 0.5, 0.5)])
    pde = R.LinearPDE(dom)
    pde.setSymmetryOn()
    pde.setValue(A=R.kronecker_scaled(1.0), d=2.0, y=2.0)
    pde.getSolverOptions().setTolerance(1e-12)
    u = pde.getSolution()
    assert torch.allclose(u, torch.ones_like(u), rtol=0, atol=1e-9)
    # a point source adds to the right-hand side at the nearest node
    pde.setValue(y_dirac=[[3.0]])
    op, rhs = pde.getSystem()
    node = dom.getDiracNodes()[0]
    pde2 = R.LinearPDE(dom)
    pde2.setValue(A=R.kronecker_scaled(1.0), d=2.0, y=2.0)
    _, rhs2 = pde2.getSystem()
    diff = (rhs - rhs2).cpu().numpy()
    assert abs(diff[node] - 3.0) < 1e-14 and np.count_nonzero(diff) == 1
