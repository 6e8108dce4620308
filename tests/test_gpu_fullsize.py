"""Parity at the benchmark's full size: ripley Brick 128^3 Poisson with the
seeded log10-kappa Gaussian random field (sigma 1, correlation 8 cells),
Y = 1, u = 0 on z = 0 (2,146,689 unknowns, 57,066,625 nonzeros).

* assembly: the GPU matrix exported to PASO CSR equals the oracle's entry by
  entry (to 1e-13 of the row scale) and the pattern bit for bit;
* SpMV: full and symmetric storage against the oracle's CSR SpMV;
* solve: PASO's PCG on the GPU against the oracle's PASO PCG on the host
  (the CPU restatement of paso/src/Solver.cpp + PCG.cpp): iteration counts
  within 2 %, solutions within 1e-6; and the plain, FP32-value and
  Gauss-Seidel solves with the true-residual criterion reach an FP64
  ||b - A x|| <= 1e-8 ||b||.
The oracle's solve takes several seconds on the host threads."""
import numpy as np
import pytest
import torch

import oracle as O
import ripley_hip as R
from gpu_helpers import to_dev
from ripley_hip import fields

pytestmark = pytest.mark.gpu

NE = (128, 128, 128)


@pytest.fixture(scope="module")
def c2():
    dom = R.Brick(*NE)
    E, Q = dom.getNumElements(), dom.getNumQuadraturePoints()
    kap = fields.log10_normal_kappa(NE, corr_len=8.0, sigma=1.0, seed=20261202)
    q = dom.whereZero(2)
    ops = {}
    for sym in (False, True):
        op = dom.newOperator(1, symmetric=sym)
        rhs = torch.zeros(op.size, dtype=torch.float64, device="cuda")
        op.assemble({"A": R.Coefficient(to_dev(kap), (3, 3), E, Q, dom.device, isotropic=True),
                     "Y": R.Coefficient(1.0, (), E, Q, dom.device)}, rhs)
        ops[sym] = (op, rhs)
    ptr, idx, oval, orhs = O.assemble(3, NE, (1.0, 1.0, 1.0), A=kap[:, None, None] * np.eye(3)[None], Y=1.0)
    return dict(dom=dom, kap=kap, q=q, ops=ops, ptr=ptr, idx=idx, oval=oval, orhs=orhs, E=E, Q=Q)


def test_full_size_assembly_and_spmv(c2):
    op, rhs = c2["ops"][False]
    ptr, idx, val = op.exportCSR()
    assert np.array_equal(ptr, c2["ptr"]) and np.array_equal(idx, c2["idx"])
    assert len(idx) == 57066625
    # |d a_ij| <= 1e-13 max_j |a_ij| for every entry (vectorised row scale)
    oval = c2["oval"]
    rowmax = np.maximum.reduceat(np.abs(oval), ptr[:-1])
    assert np.all(np.abs(val - oval) <= 1e-13 * np.repeat(rowmax, np.diff(ptr)))
    assert np.abs(rhs.cpu().numpy() - c2["orhs"]).max() <= 1e-13 * np.abs(c2["orhs"]).max()
    x = np.random.default_rng(3).standard_normal(op.size)
    ref = O.spmv(1, c2["ptr"], c2["idx"], c2["oval"], x)
    scale = O.spmv(1, c2["ptr"], c2["idx"], np.abs(c2["oval"]), np.abs(x))
    for sym in (False, True):
        o, _ = c2["ops"][sym]
        y = torch.zeros(o.size, dtype=torch.float64, device="cuda")
        o.MatrixVector(1.0, to_dev(x), 0.0, y)
        assert np.all(np.abs(y.cpu().numpy() - ref) <= 1e-13 * scale + 1e-300), sym


def _constrained(c2, sym):
    """A fresh constrained copy of the GPU system and of the oracle's."""
    dom, E, Q = c2["dom"], c2["E"], c2["Q"]
    op = dom.newOperator(1, symmetric=sym)
    rhs = torch.zeros(op.size, dtype=torch.float64, device="cuda")
    op.assemble({"A": R.Coefficient(to_dev(c2["kap"]), (3, 3), E, Q, dom.device, isotropic=True),
                 "Y": R.Coefficient(1.0, (), E, Q, dom.device)}, rhs)
    q = c2["q"]
    op.nullifyRowsAndCols(q, q, 1.0)
    rhs.masked_fill_(q > 0, 0.0)
    return op, rhs


def test_full_size_paso_pcg_against_oracle(c2):
    qn = c2["q"].cpu().numpy()
    oval = c2["oval"].copy()
    O.nullify(1, c2["ptr"], c2["idx"], oval, qn, qn, 1.0)
    orhs = np.where(qn > 0, 0.0, c2["orhs"])
    xo, dg, _, _ = O.solve(1, c2["ptr"], c2["idx"], oval, orhs, tolerance=1e-8)
    assert dg["rc"] == 0 and dg["converged"] == 1
    for sym in (False, True):
        op, rhs = _constrained(c2, sym)
        opts = R.SolverOptions()
        opts.setSymmetryOn()
        opts.setSolverMethod(R.SolverOptions.PCG)
        x = op.solve(rhs, opts).cpu().numpy()
        print("PASO PCG 128^3", "symmetric" if sym else "full", "iterations gpu", opts.getDiagnostics("num_iter"),
              "oracle", dg["num_iter"])
        assert opts.hasConverged()
        assert abs(opts.getDiagnostics("num_iter") - dg["num_iter"]) <= 0.02 * dg["num_iter"]
        assert np.linalg.norm(x - xo) <= 1e-6 * np.linalg.norm(xo)
        # PASO stops on the balanced residual; the unbalanced one is close to it
        res = np.linalg.norm(orhs - O.spmv(1, c2["ptr"], c2["idx"], oval, x)) / np.linalg.norm(orhs)
        assert res <= 1e-7


@pytest.mark.parametrize("method,prec,sym", [("PCG_PLAIN", "JACOBI", True), ("PCG_MIXED", "JACOBI", True),
                                             ("PCG", "GAUSS_SEIDEL", False)])
def test_full_size_other_methods_reach_the_fp64_residual(c2, method, prec, sym):
    op, rhs = _constrained(c2, sym)
    opts = R.SolverOptions()
    opts.setSymmetryOn()
    opts.setSolverMethod(getattr(R.SolverOptions, method))
    opts.setPreconditioner(getattr(R.SolverOptions, prec))
    opts.setTrueResidualCriterion(True)
    x = op.solve(rhs, opts)
    assert opts.hasConverged()
    fresh, frhs = _constrained(c2, False)
    r = frhs.clone()
    fresh.MatrixVector(-1.0, x, 1.0, r)
    assert (torch.linalg.norm(r) / torch.linalg.norm(frhs)).item() <= 1e-8 * 1.001


@pytest.mark.parametrize("method", ["MINRES", "PCG_PLAIN", "BICGSTAB"])
def test_full_size_other_solvers_against_oracle(c2, method):
    """PASO's MINRES and BiCGStab and the textbook PCG at 128^3 (symmetric
    storage, the folded and non-temporal paths) against the oracle's
    restatements.  MINRES and PCG match the oracle's count to 2 %.
    BiCGStab's residual is not monotone and over ~1200 iterations its count
    moves with the order of the reductions (the oracle's recurrence itself is
    pinned iteration by iteration in tests/test_oracle_recurrences.py): the
    bar is the range the oracle spans over 5, 8 and 16 OpenMP threads (its
    reduction orders), widened by 10 %."""
    qn = c2["q"].cpu().numpy()
    oval = c2["oval"].copy()
    O.nullify(1, c2["ptr"], c2["idx"], oval, qn, qn, 1.0)
    orhs = np.where(qn > 0, 0.0, c2["orhs"])
    xo, dg, _, _ = O.solve(1, c2["ptr"], c2["idx"], oval, orhs, tolerance=1e-8, method=method)
    assert dg["rc"] == 0 and dg["converged"] == 1
    counts = [dg["num_iter"]]
    if method == "BICGSTAB":
        keep = O.num_threads()
        try:
            for nt in (5, 8, 16):
                O.set_num_threads(nt)
                _, dgt, _, _ = O.solve(1, c2["ptr"], c2["idx"], oval, orhs, tolerance=1e-8, method=method)
                assert dgt["rc"] == 0
                counts.append(dgt["num_iter"])
        finally:
            O.set_num_threads(keep)
    op, rhs = _constrained(c2, True)
    opts = R.SolverOptions()
    opts.setSymmetryOn()
    opts.setSolverMethod(getattr(R.SolverOptions, method))
    x = op.solve(rhs, opts).cpu().numpy()
    n_gpu = opts.getDiagnostics("num_iter")
    print(method, "128^3 symmetric iterations gpu", n_gpu, "oracle", counts, flush=True)
    assert opts.hasConverged()
    bar = 0.10 if method == "BICGSTAB" else 0.02
    assert (1 - bar) * min(counts) <= n_gpu <= (1 + bar) * max(counts)
    assert np.linalg.norm(x - xo) <= 1e-6 * np.linalg.norm(xo)
    # PASO's test is on the balanced residual: the unbalanced one of the GPU
    # solution is held to the oracle's own (BiCGStab's recursive residual
    # drifts from the true one over its ~1200 iterations)
    res = np.linalg.norm(orhs - O.spmv(1, c2["ptr"], c2["idx"], oval, x)) / np.linalg.norm(orhs)
    res_o = np.linalg.norm(orhs - O.spmv(1, c2["ptr"], c2["idx"], oval, xo)) / np.linalg.norm(orhs)
    print(method, "unbalanced FP64 residual gpu %.3e oracle %.3e" % (res, res_o), flush=True)
    assert res <= max(1e-7, 5 * res_o)


def _paso_parity(dom, n, coefs_gpu, ptr, idx, oval, orhs, q):
    """PASO PCG on the GPU (full and symmetric storage) against the oracle."""
    E, Q = dom.getNumElements(), dom.getNumQuadraturePoints()
    qn = np.repeat(q, n)
    oval = oval.copy()
    O.nullify(n, ptr, idx, oval, qn, qn, 1.0)
    orhs = np.where(qn > 0, 0.0, orhs)
    xo, dg, _, _ = O.solve(n, ptr, idx, oval, orhs, tolerance=1e-8)
    assert dg["rc"] == 0 and dg["converged"] == 1
    for sym in (False, True):
        op = dom.newOperator(n, symmetric=sym)
        rhs = torch.zeros(op.size, dtype=torch.float64, device="cuda")
        op.assemble({k: R.Coefficient(v, s, E, Q, dom.device, isotropic=iso) for k, (v, s, iso) in coefs_gpu.items()},
                    rhs)
        qd = to_dev(qn)
        op.nullifyRowsAndCols(qd, qd, 1.0)
        rhs.masked_fill_(qd > 0, 0.0)
        opts = R.SolverOptions()
        opts.setSymmetryOn()
        opts.setSolverMethod(R.SolverOptions.PCG)
        x = op.solve(rhs, opts).cpu().numpy()
        print("PASO PCG", dom, "symmetric" if sym else "full", "iterations gpu", opts.getDiagnostics("num_iter"),
              "oracle", dg["num_iter"])
        assert opts.hasConverged()
        assert abs(opts.getDiagnostics("num_iter") - dg["num_iter"]) <= max(2, 0.02 * dg["num_iter"])
        assert np.linalg.norm(x - xo) <= 1e-6 * np.linalg.norm(xo)


def test_full_size_rectangle_256():
    # Rectangle 256^2, A = I, Y = 1, u = 0 on x0 = 0 (the reference's CPU-runnable case)
    dom = R.Rectangle(256, 256)
    ptr, idx, oval, orhs = O.assemble(2, (256, 256), (1.0, 1.0), A=np.eye(2), Y=1.0)
    q = dom.whereZero(0).cpu().numpy()
    _paso_parity(dom, 1, {"A": (to_dev(np.eye(2)), (2, 2), False), "Y": (1.0, (), False)}, ptr, idx, oval, orhs, q)


def test_full_size_brick_96_elasticity():
    # Brick 96^3 Lame elasticity, 3x3 blocks, 8 layers of lambda/mu U[1e9, 5e10] Pa,
    # gravity, u = 0 on z = 0 (2,738,019 unknowns, 217 M scalar entries)
    ne = (96, 96, 96)
    dom = R.Brick(*ne)
    lam, mu = fields.layered_lame(ne)
    eye = np.eye(3)
    A = (lam[:, None, None, None, None] * np.einsum("ij,kl->ijkl", eye, eye)
         + mu[:, None, None, None, None] * (np.einsum("ik,jl->ijkl", eye, eye) + np.einsum("il,jk->ijkl", eye, eye)))
    Y = np.array([0.0, 0.0, -2700.0 * 9.81])
    ptr, idx, oval, orhs = O.assemble(3, ne, (1.0, 1.0, 1.0), n=3, A=A, Y=Y)
    q = dom.whereZero(2).cpu().numpy()
    coefs = {"A": (to_dev(A), (3, 3, 3, 3), False), "Y": (to_dev(Y), (3,), False)}
    del A
    _paso_parity(dom, 3, coefs, ptr, idx, oval, orhs, q)
