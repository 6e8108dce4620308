"""The multilevel preconditioner of SolverOptions.FCG_LEARNED on the GPU
against the oracle's restatement (oracle/multilevel.py):

* Galerkin coarse operators P^T A P (rg_matrix_coarse_operator) entry by
  entry, full and symmetric storage, one and two levels down;
* one V-cycle (solvePreconditioner with FCG_LEARNED): Chebyshev smoothing on
  FP32 values of the balanced level 0 and the physical coarse levels, the
  coarsest level smoothed, or solved by a network through the energy-scaled
  step; the oracle mirrors the FP32 rounding of the values;
* the flexible CG reaches the FP64 residual and the oracle's PASO solution.
Bars: operators 1e-13 of the row scale; V-cycle 1e-9 (max-norm relative)
without a network, 2e-2 (relative L2, the network's BF16 roundings) with
one; solutions 1e-6."""
import numpy as np
import pytest
import scipy.sparse as sp
import torch

import lpc_torch as LT
import multilevel as ML
import oracle as O
import ripley_hip as R
from gpu_helpers import entrywise_ok, to_dev
from ripley_hip import fields
from ripley_hip import learned as L

pytestmark = pytest.mark.gpu


def _system(ne, sym, seed, sigma=1.0, corr=4.0, dirichlet=True):
    dom = R.Brick(*ne)
    E, Q = dom.getNumElements(), dom.getNumQuadraturePoints()
    kap = fields.log10_normal_kappa(ne, corr_len=corr, sigma=sigma, seed=seed)
    op = dom.newOperator(1, symmetric=sym)
    rhs = torch.zeros(op.size, dtype=torch.float64, device="cuda")
    op.assemble({"A": R.Coefficient(to_dev(kap), (3, 3), E, Q, dom.device, isotropic=True),
                 "Y": R.Coefficient(1.0, (), E, Q, dom.device)}, rhs)
    if dirichlet:
        q = dom.whereZero(2)
        op.nullifyRowsAndCols(q, q, 1.0)
        rhs.masked_fill_(q > 0, 0.0)
    return dom, op, rhs


def _csr(op):
    ptr, idx, val = op.exportCSR()
    return ML.csr_from_paso(ptr, idx, val)


@pytest.mark.parametrize("sym", [False, True])
def test_galerkin_coarse_operators(sym):
    ne = (8, 12, 4)
    _, op, _ = _system(ne, sym, 3)
    A = _csr(op)
    nodes = tuple(n + 1 for n in ne)
    for levels in (1, 2):
        C = op.coarseOperator(levels)
        ref, nd = A, nodes
        for _ in range(levels):
            nc = ML.coarsen_nodes(nd)
            P = ML.prolongation(nc)
            ref, nd = (P.T @ ref @ P).tocsr(), nc
        ptr, idx, val = C.exportCSR()
        ref = ref.tocsr()
        ref.sort_indices()
        # the coarse pattern is the 27-point stencil of the coarse Brick
        optr, oidx = O.pattern(3, tuple(n // 2 ** levels for n in ne))
        assert np.array_equal(ptr, optr) and np.array_equal(idx, oidx)
        oval = np.asarray(ref[np.repeat(np.arange(len(ptr) - 1), np.diff(ptr)), idx]).ravel()
        ok, worst = entrywise_ok(1, ptr, val, oval, rtol=1e-13)
        assert ok, (levels, worst)


def _oracle_hierarchy(op_exported, ne, levels, coarse_steps=8, degree=2):
    """The oracle's hierarchy of the balanced form of an exported (unbalanced)
    GPU matrix, FP32 values in smoothing, as the library builds it."""
    ptr, idx, val = op_exported
    bval = np.array(val, copy=True)
    bal = O.balance(1, ptr, idx, bval)
    Ab = ML.csr_from_paso(ptr, idx, bval)
    return ML.Hierarchy(Ab, tuple(n + 1 for n in ne), levels=levels, max_coarse=1, bal=bal, fp32=True,
                        coarse_steps=coarse_steps, degree=degree)


@pytest.mark.parametrize("sym", [False, True])
@pytest.mark.parametrize("levels,degree,coarse_steps", [(3, 2, 8), (2, 3, 4), (1, 2, 6)])
def test_vcycle_against_oracle(sym, levels, degree, coarse_steps):
    ne = (16, 16, 8)
    _, op, rhs = _system(ne, sym, 5)
    exported = op.exportCSR()
    H = _oracle_hierarchy(exported, ne, levels, coarse_steps, degree)
    assert len(H.levels) == levels
    op.setMultilevel(levels, degree, coarse_steps)
    o = R.SolverOptions()
    o.setSolverMethod(R.SolverOptions.FCG_LEARNED)
    r = np.random.default_rng(11).standard_normal(op.size)
    z = op.solvePreconditioner(to_dev(r), o).cpu().numpy()
    zo = H.vcycle(r, coarse="smooth") if levels > 1 else H.smooth(0, r, None, max(degree, coarse_steps))
    err = np.abs(z - zo).max() / np.abs(zo).max()
    print("V-cycle levels %d degree %d coarse %d sym %s: max rel. difference %.2e" % (levels, degree, coarse_steps,
                                                                                     sym, err), flush=True)
    assert err <= 1e-9


@pytest.mark.parametrize("steps", [1, 3, 8])
def test_smooth_against_oracle(steps):
    # the level smoother on an unbalanced matrix (training data of the network)
    ne = (8, 6, 10)
    _, op, _ = _system(ne, True, 9)
    A = _csr(op)
    H = ML.Hierarchy(A, tuple(n + 1 for n in ne), levels=1, fp32=True)
    b = np.random.default_rng(steps).standard_normal(op.size)
    x = op.smooth(to_dev(b), steps).cpu().numpy()
    xo = H.smooth(0, b, None, steps)
    assert np.abs(x - xo).max() <= 1e-9 * np.abs(xo).max()


@pytest.mark.parametrize("levels", [2, 3])
def test_vcycle_with_network_against_oracle(levels):
    # the network corrects level 1 (9^3): the coarsest level (2 levels) or
    # an intermediate one (3 levels, 5^3 below)
    ne = (16, 16, 16)
    _, op, _ = _system(ne, True, 6)
    exported = op.exportCSR()
    H = _oracle_hierarchy(exported, ne, levels)
    p = (np.random.default_rng(3).standard_normal(L.num_params()) * 0.05).astype(np.float32)
    nodes_c = (9, 9, 9)
    net = L.LearnedPreconditioner(nodes_c, params=p)
    op.setMultilevel(levels, 2, 8)
    op.setLearnedPreconditioner(net)
    o = R.SolverOptions()
    o.setSolverMethod(R.SolverOptions.FCG_LEARNED)
    r = np.random.default_rng(12).standard_normal(op.size)
    z = op.solvePreconditioner(to_dev(r), o).cpu().numpy()
    Lc = H.levels[1]
    h = 1.0 / 8
    aref = 8.0 * 3 * h ** 3 / (9.0 * h * h)

    def net_cpu(l, b):
        logk = np.log10(Lc.A.diagonal() / aref).astype(np.float32)
        return LT.apply(p, nodes_c, b, Lc.dinv, logk, bf16=True)
    zo = H.vcycle(r, coarse="smooth", net=net_cpu, net_level=1)
    zs = H.vcycle(r, coarse="smooth")
    err = np.linalg.norm(z - zo) / np.linalg.norm(zo)
    print("V-cycle (%d levels) with network: relative difference %.2e (to the V-cycle without it %.2e)"
          % (levels, err, np.linalg.norm(z - zs) / np.linalg.norm(zs)), flush=True)
    assert err <= 2e-2


def test_vcycle_of_unbalanced_matrix_against_oracle():
    # the V-cycle in physical variables (training data of the network)
    ne = (16, 8, 16)
    _, op, _ = _system(ne, False, 13)
    A = _csr(op)
    H = ML.Hierarchy(A, tuple(n + 1 for n in ne), levels=3, max_coarse=1, fp32=True)
    op.setMultilevel(3, 2, 8)
    b = np.random.default_rng(14).standard_normal(op.size)
    x = op.vcycle(to_dev(b)).cpu().numpy()
    xo = H.vcycle(b, coarse="smooth")
    assert np.abs(x - xo).max() <= 1e-9 * np.abs(xo).max()


def _true_res(op_fresh, x, b):
    r = b.clone()
    op_fresh.MatrixVector(-1.0, x, 1.0, r)
    return (torch.linalg.norm(r) / torch.linalg.norm(b)).item()


@pytest.mark.parametrize("sym", [False, True])
def test_fcg_multilevel_solve(sym):
    ne = (32, 32, 32)
    dom, op, rhs = _system(ne, sym, 7, sigma=1.0, corr=8.0)
    _, fresh, _ = _system(ne, False, 7, sigma=1.0, corr=8.0)
    o = R.SolverOptions()
    o.setSolverMethod(R.SolverOptions.FCG_LEARNED)
    o.setTrueResidualCriterion(True)
    o.setTolerance(1e-10)
    x = op.solve(rhs, o)
    it = o.getDiagnostics("num_iter")
    res = _true_res(fresh, x, rhs)
    ptr, idx, val = fresh.exportCSR()
    xo, dg, _, _ = O.solve(1, ptr, idx, val, rhs.cpu().numpy(), tolerance=1e-10)
    err = np.linalg.norm(x.cpu().numpy() - xo) / np.linalg.norm(xo)
    print("multilevel FCG 32^3 sym %s: %d iterations (PASO PCG %d), FP64 residual %.2e, difference %.2e"
          % (sym, it, dg["num_iter"], res, err), flush=True)
    assert o.hasConverged() and o.getDiagnostics("fallbacks") == 0
    assert res <= 1e-10 * 1.001 and err <= 1e-6
    assert it <= dg["num_iter"] // 5


def test_multilevel_argument_errors():
    ne = (8, 8, 8)
    _, op, rhs = _system(ne, False, 1)
    with pytest.raises(R.RipleyHipError):
        op.setLearnedPreconditioner(L.LearnedPreconditioner((9, 9, 10)))   # not a coarse level
    with pytest.raises(R.RipleyHipError):
        op.setLearnedPreconditioner(L.LearnedPreconditioner((9, 9, 9)))    # the fine grid itself
    with pytest.raises(R.RipleyHipError):
        op.setMultilevel(9, 2, 8)
    op.setMultilevel(5, 2, 8)   # more halvings than 8 elements allow
    o = R.SolverOptions()
    o.setSolverMethod(R.SolverOptions.FCG_LEARNED)
    with pytest.raises(R.RipleyHipError):
        op.solve(rhs, o)
