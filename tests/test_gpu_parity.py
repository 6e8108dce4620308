"""Parity of the HIP path (through the C ABI) with the oracle and with the
reference's known answers.  Bars: pattern and index maps bit-exact;
assembled entries |d a_ij| <= 1e-13 max_j |a_ij|; SpMV within
1e-13 (|A||x| + |beta y|); solves converge to the requested tolerance with
iteration counts close to the oracle's PASO restatement."""
import zlib

import numpy as np
import pytest
import torch

import oracle as O
import ripley_hip as R
from gpu_helpers import dirichlet_mask, entrywise_ok, gpu_assemble, grid, to_dev
from kat import KAT, kat_elements, kat_expected, kat_x

pytestmark = pytest.mark.gpu

CASES_1 = [(2, (7, 5)), (3, (5, 4, 3)), (2, (1, 1)), (3, (1, 2, 1))]


@pytest.mark.parametrize("dim,ne", CASES_1)
@pytest.mark.parametrize("bs", [1, 3])
def test_pattern_bit_exact(dim, ne, bs):
    dom, _ = grid(dim, ne)
    op = dom.newOperator(bs)
    ptr, idx, _ = op.exportCSR()
    optr, oidx = O.pattern(dim, ne)
    assert np.array_equal(ptr, optr) and np.array_equal(idx, oidx)
    assert op.nnz == len(oidx)


@pytest.mark.parametrize("bs", [1, 2, 3, 4])
@pytest.mark.parametrize("symmetric", [False, True])
def test_reference_stencil_kat(bs, symmetric):
    # ripley/test/SystemMatrixTestCase.cpp:41-134 on Rectangle(4,3)
    dom = R.Rectangle(4, 3)
    op = dom.newOperator(bs, symmetric=symmetric)
    nodes, em = kat_elements(bs)
    op.addElementBlocks(nodes, em)
    y = op.vectorMultiply(to_dev(kat_x(bs))).cpu().numpy()
    assert np.max(np.abs(y - kat_expected(bs, symmetric))) < KAT["tolerance_lsup"]


def _coefs(dim, E, rng, kinds, expanded=()):
    Q = 1 << dim
    out = {}
    if "A" in kinds:
        A = rng.standard_normal((E, dim, dim)) * 0.3 + 2 * np.eye(dim)
        out["A"] = np.broadcast_to(A[:, None], (E, Q, dim, dim)) * rng.uniform(0.5, 1.5, (E, Q, 1, 1)) \
            if "A" in expanded else A
    if "B" in kinds:
        out["B"] = rng.standard_normal((E, dim))
    if "C" in kinds:
        out["C"] = rng.standard_normal((E, dim))
    if "D" in kinds:
        out["D"] = rng.uniform(0.5, 2, (E, Q)) if "D" in expanded else rng.uniform(0.5, 2, E)
    if "X" in kinds:
        out["X"] = rng.standard_normal((E, dim))
    if "Y" in kinds:
        out["Y"] = rng.standard_normal((E, Q)) if "Y" in expanded else rng.standard_normal(E)
    return out


ASM_CASES = [
    (2, (9, 7), "AY", ()), (2, (9, 7), "ABCDXY", ()), (2, (6, 8), "ADY", ("A", "D", "Y")),
    (3, (6, 5, 4), "AY", ()), (3, (5, 4, 6), "ABCDXY", ()), (3, (4, 5, 3), "ADY", ("A", "D", "Y")),
]


@pytest.mark.parametrize("dim,ne,kinds,expanded", ASM_CASES)
def test_assembly_parity_single(dim, ne, kinds, expanded):
    rng = np.random.default_rng(zlib.crc32(repr((dim, ne, kinds)).encode()))
    dom, length = grid(dim, ne, tuple(0.7 + 0.3 * d for d in range(dim)))
    E = dom.getNumElements()
    c = _coefs(dim, E, rng, kinds, expanded)
    op, rhs = gpu_assemble(dom, 1, **c)
    ptr, idx, val = op.exportCSR()
    _, _, oval, orhs = O.assemble(dim, ne, length, expanded=expanded, **c)
    ok, worst = entrywise_ok(1, ptr, val, oval)
    assert ok, worst
    assert np.max(np.abs(rhs.cpu().numpy() - orhs)) <= 1e-13 * max(np.abs(orhs).max(), 1e-300)


@pytest.mark.parametrize("dim,ne", [(2, (8, 6)), (3, (5, 4, 3))])
def test_assembly_parity_reduced_and_isotropic(dim, ne):
    rng = np.random.default_rng(9)
    dom, length = grid(dim, ne)
    E = dom.getNumElements()
    kap = np.exp(rng.standard_normal(E))
    # isotropic kappa (compact A = kappa I) equals the full tensor form
    op, _ = gpu_assemble(dom, 1, iso=kap)
    ptr, idx, val = op.exportCSR()
    _, _, oval, _ = O.assemble(dim, ne, length, A=kap[:, None, None] * np.eye(dim)[None])
    assert entrywise_ok(1, ptr, val, oval)[0]
    # reduced-order assembly (one-point rule)
    D = rng.uniform(1, 2, E)
    op, rhs = gpu_assemble(dom, 1, reduced=True, A=kap[:, None, None] * np.eye(dim)[None], D=D, Y=D)
    _, _, val = op.exportCSR()
    _, _, oval, orhs = O.assemble(dim, ne, length, reduced=True, A=kap[:, None, None] * np.eye(dim)[None],
                                  D=D, Y=D)
    assert entrywise_ok(1, ptr, val, oval)[0]
    assert np.allclose(rhs.cpu().numpy(), orhs, rtol=1e-13, atol=0)


@pytest.mark.parametrize("dim,ne,sym", [(3, (5, 4, 6), False), (3, (5, 4, 6), True), (2, (9, 6), False),
                                        (2, (9, 6), True)])
def test_assembly_element_constant_isotropic(dim, ne, sym):
    # element-constant kappa I with element-constant Y (the constant-coefficient
    # branch, DefaultAssembler3D.cpp:906-993, :2205-2214): the first assembly
    # writes into the zeroed matrix, the second adds to it; the general
    # quadrature kernel (A as a full tensor) gives the same entries
    rng = np.random.default_rng(31 + dim + 2 * sym)
    length = tuple(0.7 + 0.3 * d for d in range(dim))
    dom, _ = grid(dim, ne, length)
    E = dom.getNumElements()
    kap = np.exp(rng.standard_normal(E))
    Y = rng.standard_normal(E)
    op, rhs = gpu_assemble(dom, 1, symmetric=sym, iso=kap, Y=Y)
    ptr, idx, val = op.exportCSR()
    _, _, oval, orhs = O.assemble(dim, ne, length, A=kap[:, None, None] * np.eye(dim)[None], Y=Y)
    ok, worst = entrywise_ok(1, ptr, val, oval)
    assert ok, worst
    assert np.max(np.abs(rhs.cpu().numpy() - orhs)) <= 1e-13 * np.abs(orhs).max()
    gen, _ = gpu_assemble(dom, 1, symmetric=sym, A=kap[:, None, None] * np.eye(dim)[None], Y=Y)
    ok, worst = entrywise_ok(1, ptr, gen.exportCSR()[2], oval)
    assert ok, worst
    E_, Q = dom.getNumElements(), dom.getNumQuadraturePoints()
    op.assemble({"A": R.Coefficient(to_dev(kap), (dim, dim), E_, Q, dom.device, isotropic=True)}, None)
    ok, worst = entrywise_ok(1, ptr, op.exportCSR()[2], 2 * oval)
    assert ok, worst


@pytest.mark.parametrize("n,dim,ne", [(3, 3, (4, 3, 3)), (2, 2, (6, 5)), (3, 2, (4, 4))])
def test_assembly_parity_system(n, dim, ne):
    rng = np.random.default_rng(n * 10 + dim)
    dom, length = grid(dim, ne)
    E = dom.getNumElements()
    A = rng.standard_normal((E, n, dim, n, dim))
    D = rng.standard_normal((E, n, n))
    X = rng.standard_normal((E, n, dim))
    Y = rng.standard_normal((E, n))
    op, rhs = gpu_assemble(dom, n, A=A, D=D, X=X, Y=Y)
    ptr, idx, val = op.exportCSR()
    _, _, oval, orhs = O.assemble(dim, ne, length, n=n, A=A, D=D, X=X, Y=Y)
    ok, worst = entrywise_ok(n, ptr, val, oval)
    assert ok, worst
    assert np.max(np.abs(rhs.cpu().numpy() - orhs)) <= 1e-13 * np.abs(orhs).max()


@pytest.mark.parametrize("dim,ne,n,sym", [(2, (11, 9), 1, False), (3, (7, 6, 5), 1, False),
                                          (3, (5, 4, 4), 3, False), (3, (7, 6, 5), 1, True),
                                          (3, (5, 4, 4), 3, True)])
def test_spmv_parity(dim, ne, n, sym):
    rng = np.random.default_rng(17)
    dom, length = grid(dim, ne)
    E = dom.getNumElements()
    if n == 1:
        A = np.exp(rng.standard_normal(E))[:, None, None] * np.eye(dim)[None]
        op, _ = gpu_assemble(dom, 1, symmetric=sym, A=A, D=rng.uniform(1, 2, E))
    else:
        lam, mu = rng.uniform(1, 3, E), rng.uniform(1, 3, E)
        eye = np.eye(dim)
        A = (lam[:, None, None, None, None] * np.einsum("ij,kl->ijkl", eye, eye)
             + mu[:, None, None, None, None] * (np.einsum("ik,jl->ijkl", eye, eye)
                                                + np.einsum("il,jk->ijkl", eye, eye)))
        op, _ = gpu_assemble(dom, n, symmetric=sym, A=A)
    ptr, idx, val = op.exportCSR()
    x = rng.standard_normal(op.size)
    y0 = rng.standard_normal(op.size)
    y = to_dev(y0)
    op.MatrixVector(0.75, to_dev(x), -1.5, y)
    ref = O.spmv(n, ptr, idx, val, x, 0.75, -1.5, y0)
    absA = O.spmv(n, ptr, idx, np.abs(val), np.abs(x), 0.75, 1.5, np.abs(y0))
    assert np.all(np.abs(y.cpu().numpy() - ref) <= 1e-13 * absA + 1e-300)


def test_symmetric_storage_equals_full():
    rng = np.random.default_rng(5)
    dom, _ = grid(3, (6, 5, 4))
    E = dom.getNumElements()
    A = np.exp(rng.standard_normal(E))[:, None, None] * np.eye(3)[None]
    full, rf = gpu_assemble(dom, 1, A=A, Y=np.ones(E))
    half, rh = gpu_assemble(dom, 1, symmetric=True, A=A, Y=np.ones(E))
    _, _, vf = full.exportCSR()
    ptr, _, vh = half.exportCSR()
    assert entrywise_ok(1, ptr, vh, vf, rtol=1e-14)[0]
    assert torch.equal(rf, rh)


def test_nullify_parity():
    rng = np.random.default_rng(21)
    for n, dim, ne in [(1, 3, (5, 4, 4)), (3, 3, (3, 3, 3)), (1, 2, (8, 7))]:
        dom, length = grid(dim, ne)
        E = dom.getNumElements()
        A = rng.standard_normal((E,) + ((dim, dim) if n == 1 else (n, dim, n, dim)))
        op, _ = gpu_assemble(dom, n, A=A)
        ptr, idx, val = op.exportCSR()
        qr = (rng.uniform(size=op.size) < 0.2).astype(float)
        qc = (rng.uniform(size=op.size) < 0.2).astype(float)
        op.nullifyRowsAndCols(to_dev(qr), to_dev(qc), 1.0)
        _, _, v2 = op.exportCSR()
        O.nullify(n, ptr, idx, val, qr, qc, 1.0)
        assert np.array_equal(v2, val)


def _poisson_problem(dim, ne, seed=3, sigma=1.0):
    rng = np.random.default_rng(seed)
    dom, length = grid(dim, ne)
    E = dom.getNumElements()
    kap = np.exp(sigma * rng.standard_normal(E))
    A = kap[:, None, None] * np.eye(dim)[None]
    op, rhs = gpu_assemble(dom, 1, A=A, Y=np.ones(E))
    q = dirichlet_mask(dom, dim - 1)
    # insertConstraint with r = 0: rhs[q] = 0, nullify (linearPDEs.py:2084-2112)
    op.nullifyRowsAndCols(to_dev(q), to_dev(q), 1.0)
    rhs = torch.where(to_dev(q) > 0, torch.zeros_like(rhs), rhs)
    ptr, idx, oval, orhs = O.assemble(dim, ne, length, A=A, Y=np.ones(E))
    O.nullify(1, ptr, idx, oval, q, q, 1.0)
    orhs = np.where(q > 0, 0.0, orhs)
    return op, rhs, ptr, idx, oval, orhs


@pytest.mark.parametrize("dim,ne", [(2, (32, 32)), (3, (12, 10, 14))])
@pytest.mark.parametrize("method", ["PCG", "PCG_PLAIN"])
def test_solve_parity(dim, ne, method):
    op, rhs, ptr, idx, oval, orhs = _poisson_problem(dim, ne)
    opts = R.SolverOptions()
    opts.setSymmetryOn()
    opts.setSolverMethod(getattr(R.SolverOptions, method))
    opts.setTolerance(1e-8)
    x = op.solve(rhs, opts).cpu().numpy()
    xo, dg, _, _ = O.solve(1, ptr, idx, oval, orhs, tolerance=1e-8, plain=(method == "PCG_PLAIN"))
    assert opts.hasConverged() and dg["converged"] == 1
    assert abs(opts.getDiagnostics("num_iter") - dg["num_iter"]) <= max(2, 0.02 * dg["num_iter"])
    res = np.linalg.norm(orhs - O.spmv(1, ptr, idx, oval, x)) / np.linalg.norm(orhs)
    assert res <= 1e-7
    assert np.linalg.norm(x - xo) <= 1e-6 * np.linalg.norm(xo)


def test_solve_mixed_precision_reaches_fp64_tolerance():
    op, rhs, ptr, idx, oval, orhs = _poisson_problem(3, (16, 16, 16), sigma=1.5)
    opts = R.SolverOptions()
    opts.setSymmetryOn()
    opts.setSolverMethod(R.SolverOptions.PCG_MIXED)
    x = op.solve(rhs, opts).cpu().numpy()
    assert opts.hasConverged()
    xo, dg, _, _ = O.solve(1, ptr, idx, oval, orhs, tolerance=1e-8)
    res = np.linalg.norm(orhs - O.spmv(1, ptr, idx, oval, x)) / np.linalg.norm(orhs)
    assert res <= 1e-7
    assert np.linalg.norm(x - xo) <= 1e-6 * np.linalg.norm(xo)


@pytest.mark.parametrize("method", ["PCG", "PCG_PLAIN", "PCG_MIXED"])
def test_true_residual_criterion(method):
    # stop on the FP64 unbalanced ||b - A x||_2 <= tol ||b||_2
    op, rhs, ptr, idx, oval, orhs = _poisson_problem(3, (14, 12, 16), sigma=2.0)
    opts = R.SolverOptions()
    opts.setSymmetryOn()
    opts.setSolverMethod(getattr(R.SolverOptions, method))
    opts.setTolerance(1e-9)
    opts.setTrueResidualCriterion(True)
    x = op.solve(rhs, opts).cpu().numpy()
    assert opts.hasConverged()
    res = np.linalg.norm(orhs - O.spmv(1, ptr, idx, oval, x)) / np.linalg.norm(orhs)
    assert res <= 1e-9 * 1.001


def test_solve_elasticity_parity():
    rng = np.random.default_rng(8)
    dom, length = grid(3, (6, 5, 6))
    E = dom.getNumElements()
    lam, mu = rng.uniform(1, 5, E), rng.uniform(1, 5, E)
    eye = np.eye(3)
    A = (lam[:, None, None, None, None] * np.einsum("ij,kl->ijkl", eye, eye)
         + mu[:, None, None, None, None] * (np.einsum("ik,jl->ijkl", eye, eye) + np.einsum("il,jk->ijkl", eye, eye)))
    Y = np.tile([0.0, 0.0, -1.0], (E, 1))
    op, rhs = gpu_assemble(dom, 3, A=A, Y=Y)
    q = dirichlet_mask(dom, 2, 3)
    op.nullifyRowsAndCols(to_dev(q), to_dev(q), 1.0)
    rhs = torch.where(to_dev(q) > 0, torch.zeros_like(rhs), rhs)
    ptr, idx, oval, orhs = O.assemble(3, (6, 5, 6), length, n=3, A=A, Y=Y)
    O.nullify(3, ptr, idx, oval, q, q, 1.0)
    orhs = np.where(q > 0, 0.0, orhs)
    opts = R.SolverOptions()
    opts.setSymmetryOn()
    opts.setSolverMethod(R.SolverOptions.PCG)
    x = op.solve(rhs, opts).cpu().numpy()
    xo, dg, _, _ = O.solve(3, ptr, idx, oval, orhs, tolerance=1e-8)
    assert opts.hasConverged() and dg["converged"]
    assert abs(opts.getDiagnostics("num_iter") - dg["num_iter"]) <= max(2, 0.02 * dg["num_iter"])
    assert np.linalg.norm(x - xo) <= 1e-6 * np.linalg.norm(xo)


def test_solve_is_deterministic_and_graph_replay_stable():
    xs = []
    for _ in range(3):
        op2, rhs2, *_ = _poisson_problem(3, (10, 10, 10))
        opts = R.SolverOptions()
        opts.setSymmetryOn()
        opts.setSolverMethod(R.SolverOptions.PCG)
        opts.setCheckInterval(4)
        xs.append(op2.solve(rhs2, opts).cpu().numpy())
    assert np.array_equal(xs[0], xs[1]) and np.array_equal(xs[1], xs[2])


def test_balanced_state_and_errors():
    op, rhs, *_ = _poisson_problem(2, (8, 8))
    opts = R.SolverOptions()
    opts.setSymmetryOn()
    opts.setSolverMethod(R.SolverOptions.PCG)
    op.solve(rhs, opts)
    with pytest.raises(R.RipleyHipError, match="balanced matrix is not supported"):
        op.vectorMultiply(torch.zeros_like(rhs))
    opts.setTolerance(1e-20)
    with pytest.raises(R.RipleyHipError, match="Tolerance is too small"):
        op.solve(rhs, opts)
    opts.setTolerance(1e-8)
    bad = rhs.clone()
    bad[3] = float("nan")
    with pytest.raises(R.RipleyHipError, match="undefined values"):
        op.solve(bad, opts)
    assert torch.count_nonzero(op.solve(torch.zeros_like(rhs), opts)) == 0
    opts.setIterMax(2)
    with pytest.raises(R.PasoException, match="maximum number of iteration"):
        op.solve(rhs, opts)
    opts.setAcceptanceConvergenceFailureOn()
    op.solve(rhs, opts)
    assert not opts.hasConverged()


def test_singular_diagonal_raises():
    dom = R.Rectangle(3, 3)
    op = dom.newOperator(1)
    opts = R.SolverOptions()
    opts.setSymmetryOn()
    opts.setSolverMethod(R.SolverOptions.PCG)
    with pytest.raises(R.RipleyHipError, match="non-regular main diagonal"):
        op.solve(torch.ones(op.size, dtype=torch.float64, device="cuda"), opts)


def test_linearpde_facade_constant_solution():
    # escriptcore/test/python/test_linearPDEs.py:1741-1748 (PCG + JACOBI): u == 1
    dom = R.Brick(8, 8, 8)
    pde = R.LinearPDE(dom)
    pde.setSymmetryOn()
    pde.setValue(A=R.kronecker_scaled(1.0), D=1.0, Y=1.0)
    pde.getSolverOptions().setSolverMethod(R.SolverOptions.PCG)
    u = pde.getSolution()
    assert torch.max(torch.abs(u - 1.0)).item() <= 1e-6


def test_linearpde_facade_linear_exact_solution():
    # test_simplesolve.py:76-160 (Dirichlet on the whole boundary variant)
    dom = R.Brick(6, 5, 4)
    X = dom.getX()
    u_ex = 1 + 2 * X[:, 0] + 3 * X[:, 1] + 4 * X[:, 2]
    q = ((X < 1e-12) | (X > 1 - 1e-12)).any(dim=1).to(torch.float64)
    pde = R.Poisson(dom)
    pde.setValue(q=q, r=u_ex)
    pde.getSolverOptions().setSolverMethod(R.SolverOptions.PCG)
    u = pde.getSolution()
    assert torch.max(torch.abs(u - u_ex)).item() <= 1e-6 * torch.max(torch.abs(u_ex)).item()


def _advection_problem(dim, ne, seed, n=1):
    """Non-symmetric constrained system on the GPU and in the oracle."""
    rng = np.random.default_rng(seed)
    dom, length = grid(dim, ne)
    E = dom.getNumElements()
    if n == 1:
        A = np.tile(np.eye(dim), (E, 1, 1))
        B = rng.standard_normal((E, dim)) * 3.0
        Y = np.ones(E)
    else:
        A = np.tile(np.einsum("ik,jl->ijkl", np.eye(n), np.eye(dim)), (E, 1, 1, 1, 1))
        B = rng.standard_normal((E, n, dim, n)) * 2.0
        Y = np.ones((E, n))
    op, rhs = gpu_assemble(dom, n, A=A, B=B, Y=Y)
    q = dirichlet_mask(dom, dim - 1, n)
    op.nullifyRowsAndCols(to_dev(q), to_dev(q), 1.0)
    rhs = torch.where(to_dev(q) > 0, torch.zeros_like(rhs), rhs)
    ptr, idx, oval, orhs = O.assemble(dim, ne, length, n=n, A=A, B=B, Y=Y)
    O.nullify(n, ptr, idx, oval, q, q, 1.0)
    orhs = np.where(q > 0, 0.0, orhs)
    return op, rhs, ptr, idx, oval, orhs


@pytest.mark.parametrize("dim,ne,n", [(2, (24, 20), 1), (3, (10, 9, 8), 1), (3, (6, 5, 5), 3), (2, (12, 10), 2)])
def test_bicgstab_parity(dim, ne, n):
    op, rhs, ptr, idx, oval, orhs = _advection_problem(dim, ne, 9, n)
    opts = R.SolverOptions()      # non-symmetric: PASO's default is BiCGStab
    opts.setTolerance(1e-9)
    x = op.solve(rhs, opts).cpu().numpy()
    xo, dg, _, _ = O.solve(n, ptr, idx, oval, orhs, tolerance=1e-9, method="BICGSTAB")
    assert opts.hasConverged() and dg["converged"] == 1
    assert abs(opts.getDiagnostics("num_iter") - dg["num_iter"]) <= max(3, 0.05 * dg["num_iter"])
    res = np.linalg.norm(orhs - O.spmv(n, ptr, idx, oval, x)) / np.linalg.norm(orhs)
    assert res <= 1e-8
    assert np.linalg.norm(x - xo) <= 1e-6 * np.linalg.norm(xo)


@pytest.mark.parametrize("method", ["PCG", "PCG_PLAIN", "PCG_MIXED"])
@pytest.mark.parametrize("dim,ne,n", [(3, (14, 12, 10), 1), (2, (40, 30), 1), (3, (6, 5, 5), 3)])
def test_symmetric_storage_solves(method, dim, ne, n):
    """Solves on upper-half (symmetric) storage against the oracle's full CSR
    solve: the storage the benchmark's fastest method uses."""
    rng = np.random.default_rng(31)
    dom, length = grid(dim, ne)
    E = dom.getNumElements()
    if n == 1:
        A = np.exp(1.5 * rng.standard_normal(E))[:, None, None] * np.eye(dim)[None]
        Y = np.ones(E)
    else:
        lam, mu = rng.uniform(1, 5, E), rng.uniform(1, 5, E)
        eye = np.eye(dim)
        A = (lam[:, None, None, None, None] * np.einsum("ij,kl->ijkl", eye, eye)
             + mu[:, None, None, None, None] * (np.einsum("ik,jl->ijkl", eye, eye)
                                                 + np.einsum("il,jk->ijkl", eye, eye)))
        Y = np.tile([0.0, 0.0, -1.0], (E, 1))
    op, rhs = gpu_assemble(dom, n, symmetric=True, A=A, Y=Y)
    q = dirichlet_mask(dom, dim - 1, n)
    op.nullifyRowsAndCols(to_dev(q), to_dev(q), 1.0)
    rhs = torch.where(to_dev(q) > 0, torch.zeros_like(rhs), rhs)
    ptr, idx, oval, orhs = O.assemble(dim, ne, length, n=n, A=A, Y=Y)
    O.nullify(n, ptr, idx, oval, q, q, 1.0)
    orhs = np.where(q > 0, 0.0, orhs)
    opts = R.SolverOptions()
    opts.setSymmetryOn()
    opts.setSolverMethod(getattr(R.SolverOptions, method))
    opts.setTolerance(1e-9)
    opts.setTrueResidualCriterion(True)
    x = op.solve(rhs, opts).cpu().numpy()
    assert opts.hasConverged()
    res = np.linalg.norm(orhs - O.spmv(n, ptr, idx, oval, x)) / np.linalg.norm(orhs)
    assert res <= 1e-9 * 1.001
    xo, dg, _, _ = O.solve(n, ptr, idx, oval, orhs, tolerance=1e-11, plain=(method == "PCG_PLAIN"))
    assert np.linalg.norm(x - xo) <= 1e-6 * np.linalg.norm(xo)


@pytest.mark.parametrize("method", ["PCG_PLAIN", "PCG"])
def test_chunk_boundaries_and_iteration_limits(method):
    """The textbook PCG folds its update reduction into the direction update
    and commits the scalars at the next stage or at the end of a chunk of
    iterations: the iteration count, the stopping test and the solution must
    not depend on where chunks end (check intervals 1, 7, 32), and an
    iteration limit must stop at the same count as the oracle's."""
    op, rhs, ptr, idx, oval, orhs = _poisson_problem(3, (12, 10, 14))
    xs, its = [], []
    for check in (1, 7, 32):
        opts = R.SolverOptions()
        opts.setSymmetryOn()
        opts.setSolverMethod(getattr(R.SolverOptions, method))
        opts.setTolerance(1e-8)
        opts.setCheckInterval(check)
        xs.append(op.solve(rhs, opts).cpu().numpy())
        its.append(opts.getDiagnostics("num_iter"))
    assert its[0] == its[1] == its[2]
    assert np.array_equal(xs[0], xs[1]) and np.array_equal(xs[0], xs[2])
    for limit in (3, 17):
        for check in (1, 5, 32):
            opts = R.SolverOptions()
            opts.setSymmetryOn()
            opts.setSolverMethod(getattr(R.SolverOptions, method))
            opts.setTolerance(1e-8)
            opts.setCheckInterval(check)
            opts.setIterMax(limit)
            with pytest.raises(R.PasoException, match="maximum number of iteration"):
                op.solve(rhs, opts)
            opts.setAcceptanceConvergenceFailureOn()
            op.solve(rhs, opts)
            _, dg, _, _ = O.solve(1, ptr, idx, oval, orhs, tolerance=1e-8, iter_max=limit,
                                  plain=(method == "PCG_PLAIN"))
            print(method, "limit", limit, "check", check, "iterations gpu", opts.getDiagnostics("num_iter"),
                  "oracle", dg["num_iter"])
            assert dg["rc"] == 1 and not opts.hasConverged()
            assert opts.getDiagnostics("num_iter") == dg["num_iter"]


def test_bicgstab_iteration_limit_ranks_like_paso():
    """Solver_BiCGStab reports a reached iteration limit (and a breakdown)
    ahead of convergence found in the same iteration (BiCGStab.cpp:236-240):
    for every limit up to past the converging count the GPU solve ends with
    the oracle's status and iteration count (a small system, where both take
    the same iterations)."""
    op, rhs, ptr, idx, oval, orhs = _advection_problem(2, (7, 6), 9, 1)
    _, dg0, _, _ = O.solve(1, ptr, idx, oval, orhs, tolerance=1e-9, method="BICGSTAB")
    assert dg0["rc"] == 0
    for limit in range(1, dg0["num_iter"] + 3):
        _, dg, _, _ = O.solve(1, ptr, idx, oval, orhs, tolerance=1e-9, iter_max=limit, method="BICGSTAB")
        opts = R.SolverOptions()
        opts.setTolerance(1e-9)
        opts.setIterMax(limit)
        print("limit", limit, "oracle rc", dg["rc"], "iterations", dg["num_iter"], flush=True)
        if dg["rc"] == 1:
            with pytest.raises(R.PasoException, match="maximum number of iteration"):
                op.solve(rhs, opts)
        else:
            assert dg["rc"] == 0
            op.solve(rhs, opts)
            assert opts.hasConverged()
        assert opts.getDiagnostics("num_iter") == dg["num_iter"]
