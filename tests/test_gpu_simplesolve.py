"""The reference's SimpleSolveTestCase (escriptcore/test/python/
test_simplesolve.py:42-160, 267-292) as run on ripley by
ripley/test/python/run_pasoSolversOnRipley.py:40-116 (Rectangle(11, 11),
Brick(11, 11, 7), PASO with PCG, BiCGStab or MINRES and Jacobi), here also
with Gauss-Seidel and ILU0: a linear exact solution u_ex, Dirichlet data on x0 = 0,
Neumann data y = g_ex . n on the rest of the boundary; the system variant
couples the components through D = (1 - f) I + f 1 1^T (f = -0.4) with
Y = D u_ex.  Passes when Lsup(u - u_ex) < 1e-6 Lsup(u_ex) at solver tolerance
1e-8."""
import numpy as np
import pytest
import torch

import ripley_hip as R

pytestmark = pytest.mark.gpu

FAC_DIAG, FAC_OFFDIAG = 1.0, -0.4


def exact(dim, system):
    """(u_ex(x) -> (..., n), gradient g_ex (n, dim)) of SolveTestCaseOrder1."""
    if not system:
        c = np.array([2.0, 3.0, 4.0][:dim])
        return (lambda X: 1.0 + X @ c), c[None, :]
    if dim == 2:
        g = np.array([[2.0, 3.0], [3.0, 2.0]])
        c0 = np.array([1.0, -1.0])
    else:
        g = np.array([[2.0, 3.0, 4.0], [4.0, 1.0, -2.0], [8.0, 4.0, 5.0]])
        c0 = np.array([1.0, -1.0, 5.0])
    return (lambda X: c0 + X @ g.T), g


def face_normals(dom):
    """Outward normal of every face element, in FunctionOnBoundary order."""
    dim = dom.getDim()
    out = []
    for face, axis, side, _, count in dom._faces():
        nvec = np.zeros(dim)
        nvec[axis] = 1.0 if side else -1.0
        out.append(np.repeat(nvec[None, :], count, axis=0))
    return np.concatenate(out)


@pytest.mark.parametrize("dim", [2, 3])
@pytest.mark.parametrize("system", [False, True])
@pytest.mark.parametrize("method,prec", [("PCG", "JACOBI"), ("BICGSTAB", "JACOBI"), ("MINRES", "JACOBI"),
                                         ("PCG", "GAUSS_SEIDEL"), ("PCG", "ILU0")])
def test_simple_solve(dim, system, method, prec):
    dom = R.Rectangle(11, 11) if dim == 2 else R.Brick(11, 11, 7)
    n = dim if system else 1
    u_fn, g = exact(dim, system)
    X = dom.getX().cpu().numpy()
    u_ex = u_fn(X).reshape(-1)                       # node-major, component fastest
    q = (np.abs(X[:, 0]) < 1e-12).astype(float)
    pde = R.LinearPDE(dom, numEquations=n)
    normals = face_normals(dom)
    y = normals @ g.T                                # (F, n): g_ex n, constant per face element
    if not system:
        pde.setValue(A=R.kronecker_scaled(1.0), y=y[:, 0], q=q, r=u_ex)
    else:
        A = np.zeros((n, dim, n, dim))
        for i in range(n):
            A[i, :, i, :] = np.eye(dim)
        D = np.eye(n) * (FAC_DIAG - FAC_OFFDIAG) + np.ones((n, n)) * FAC_OFFDIAG
        XQ = dom.getQuadraturePoints().cpu().numpy()
        Y = u_fn(XQ) @ D.T                           # (E, Q, n): D u_ex at the quadrature points
        pde.setValue(A=A, D=D, Y=torch.as_tensor(Y, device="cuda"), y=y, q=np.repeat(q, n), r=u_ex)
    o = pde.getSolverOptions()
    o.setSolverMethod(getattr(R.SolverOptions, method))
    o.setPreconditioner(getattr(R.SolverOptions, prec))
    o.setTolerance(1e-8)
    u = pde.getSolution().cpu().numpy().reshape(-1)
    err = np.abs(u - u_ex).max()
    print(dim, system, method, prec, "iterations", o.getDiagnostics("num_iter"), "error", err)
    assert err < 1e-6 * np.abs(u_ex).max()
