"""PASO's MINRES (paso/src/MINRES.cpp:59-191) on the GPU through the C ABI
against the oracle: iteration counts and solutions with every preconditioner
(fused single Jacobi sweep, buffered Jacobi sweeps, Gauss-Seidel, ILU0,
none), full and symmetric storage, block sizes 1 and 3; a symmetric
indefinite system that MINRES solves; the negative-energy-norm error."""
import numpy as np
import pytest
import torch

import oracle as O
import ripley_hip as R
from gpu_helpers import dirichlet_mask, gpu_assemble, grid, to_dev
from test_gpu_precond import _constrained, _options

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("prec,sweeps", [("JACOBI", 1), ("JACOBI", 2), ("GAUSS_SEIDEL", 1), ("ILU0", 1),
                                         ("NONE", 1)])
@pytest.mark.parametrize("dim,ne,n,sym", [(3, (12, 10, 9), 1, False), (2, (30, 26), 1, True),
                                          (3, (5, 5, 4), 3, True), (2, (9, 8), 2, False)])
def test_minres_parity(prec, sweeps, dim, ne, n, sym):
    op, rhs, ptr, idx, oval, orhs = _constrained(dim, ne, n, sym, seed=11 + n)
    opts = _options(prec, sweeps, R.SolverOptions.MINRES)
    opts.setTolerance(1e-9)
    x = op.solve(rhs, opts).cpu().numpy()
    xo, dg, _, _ = O.solve(n, ptr, idx, oval, orhs, tolerance=1e-9, method="MINRES", preconditioner=prec,
                           sweeps=sweeps)
    print(prec, sweeps, dim, ne, n, sym, "iterations gpu", opts.getDiagnostics("num_iter"), "oracle",
          dg["num_iter"])
    assert opts.hasConverged() and dg["converged"] == 1 and dg["rc"] == 0
    assert abs(opts.getDiagnostics("num_iter") - dg["num_iter"]) <= max(3, 0.05 * dg["num_iter"])
    res = np.linalg.norm(orhs - O.spmv(n, ptr, idx, oval, x)) / np.linalg.norm(orhs)
    assert res <= 1e-8
    assert np.linalg.norm(x - xo) <= 1e-6 * np.linalg.norm(xo)


def _reaction(D, ne=(16, 13)):
    dom, length = grid(2, ne)
    E = dom.getNumElements()
    A = np.tile(np.eye(2), (E, 1, 1))
    op, rhs = gpu_assemble(dom, 1, A=A, D=np.full(E, D), Y=np.ones(E))
    ptr, idx, oval, orhs = O.assemble(2, ne, length, A=A, D=np.full(E, D), Y=np.ones(E))
    q = dirichlet_mask(dom, 1)
    op.nullifyRowsAndCols(to_dev(q), to_dev(q), 1.0)
    rhs = torch.where(to_dev(q) > 0, torch.zeros_like(rhs), rhs)
    O.nullify(1, ptr, idx, oval, q, q, 1.0)
    orhs = np.where(q > 0, 0.0, orhs)
    return op, rhs, ptr, idx, oval, orhs


def test_minres_symmetric_indefinite():
    # a negative reaction term: indefinite matrix, positive diagonal
    op, rhs, ptr, idx, oval, orhs = _reaction(-60.0)
    M = O.csr_to_dense(1, ptr, idx, oval)
    ev = np.linalg.eigvalsh(M)
    assert ev.min() < 0 < ev.max() and np.all(np.diag(M) > 0)
    opts = _options("JACOBI", 1, R.SolverOptions.MINRES)
    opts.setTolerance(1e-10)
    x = op.solve(rhs, opts).cpu().numpy()
    xo, dg, _, _ = O.solve(1, ptr, idx, oval, orhs, tolerance=1e-10, method="MINRES")
    xs = np.linalg.solve(M, orhs)
    print("indefinite: iterations gpu", opts.getDiagnostics("num_iter"), "oracle", dg["num_iter"])
    assert dg["rc"] == 0 and opts.hasConverged()
    assert abs(opts.getDiagnostics("num_iter") - dg["num_iter"]) <= max(3, 0.05 * dg["num_iter"])
    assert np.linalg.norm(x - xs) <= 1e-8 * np.linalg.norm(xs)


def test_minres_negative_energy_norm():
    # -lap u - u: negative definite; the Jacobi preconditioner is too, so
    # r.M^-1 r < 0 at the start (MINRES.cpp:88-90, solve.cpp:137-138)
    dom, length = grid(2, (8, 7))
    E = dom.getNumElements()
    op, rhs = gpu_assemble(dom, 1, A=np.tile(-np.eye(2), (E, 1, 1)), D=np.full(E, -1.0), Y=np.ones(E))
    opts = _options("JACOBI", 1, R.SolverOptions.MINRES)
    with pytest.raises(R.RipleyHipError, match="negative energy norm"):
        op.solve(rhs, opts)
    ptr, idx, oval, orhs = O.assemble(2, (8, 7), length, A=np.tile(-np.eye(2), (E, 1, 1)), D=np.full(E, -1.0),
                                      Y=np.ones(E))
    assert O.solve(1, ptr, idx, oval, orhs, method="MINRES")[1]["rc"] == 5
