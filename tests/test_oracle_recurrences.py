"""Pins the recurrences of the oracle's Krylov solvers iteration by iteration,
not only through their converged solutions.

* PASO's residual-smoothed PCG (paso/src/PCG.cpp:62-368): the inner CG
  iterate x2 and its residual r follow the textbook PCG (scipy's ``cg`` with
  the same Jacobi preconditioner, iterate for iterate), r stays the true
  residual of x2, and the smoothed pair (x, rs) of PCG.cpp:262-345 is the
  minimal residual smoothing of the CG sequence: rs = Db - A x, rs is
  orthogonal to r_k - rs_(k-1) (gamma minimises ||rs_(k-1) + gamma (r_k -
  rs_(k-1))||) and so ||rs_k|| <= min(||rs_(k-1)||, ||r_k||).
* Solver_BiCGStab (paso/src/BiCGStab.cpp:77-257): the iterates follow
  scipy's right-preconditioned ``bicgstab`` iterate for iterate, r = Db - A x,
  rtld.s = 0 (the alpha step) and t.r = 0 (the omega step).
* Solver_MINRES (paso/src/MINRES.cpp:59-191): the iterates follow scipy's
  ``minres`` with the same preconditioner, ||Db - A x_k||_(M) never grows,
  and the recurrence's rnorm_prec is that norm.

Each check is also run against a mistake planted in the oracle (the
smoothing switched off, the smoothing coefficient's sign flipped, BiCGStab's
beta without alpha/omega, omega's sign flipped, MINRES's w recurrence without
its a3 term) and must fail there.
"""
import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

import oracle as O


def _heterogeneous_poisson(ne=(14, 11), sigma=1.5, seed=7):
    """-div(kappa grad u) = 1 with a log-normal per-element kappa, u = 0 on
    the x1 = 0 edge: SPD, and rough enough that the CG residual is not
    monotone."""
    rng = np.random.default_rng(seed)
    E = int(np.prod(ne))
    kap = np.exp(sigma * rng.standard_normal(E))
    ptr, idx, val, rhs = O.assemble(2, ne, (1.0, 1.0), A=kap[:, None, None] * np.eye(2)[None], Y=1.0)
    N = len(ptr) - 1
    q = np.zeros(N)
    q[: ne[0] + 1] = 1.0
    O.nullify(1, ptr, idx, val, q, q, 1.0)
    rhs = np.where(q > 0, 0.0, rhs)
    return ptr, idx, val, rhs


def _advection(ne=(12, 10), seed=4):
    """-lap u + B.grad u = 1, u = 0 on x1 = 0: non-symmetric."""
    rng = np.random.default_rng(seed)
    E = int(np.prod(ne))
    ptr, idx, val, rhs = O.assemble(2, ne, (1.0, 1.0), A=np.tile(np.eye(2), (E, 1, 1)),
                                    B=rng.standard_normal((E, 2)) * 3.0, Y=1.0)
    N = len(ptr) - 1
    q = np.zeros(N)
    q[: ne[0] + 1] = 1.0
    O.nullify(1, ptr, idx, val, q, q, 1.0)
    rhs = np.where(q > 0, 0.0, rhs)
    return ptr, idx, val, rhs


def _balanced(ptr, idx, vbal, d, rhs):
    """The system the inner solver sees: balanced matrix, D b, and PASO's
    initial guess x0 = M^-1 D b (Solver.cpp:46-331)."""
    N = len(ptr) - 1
    A = sp.csr_matrix((vbal, idx, ptr), shape=(N, N))
    db = d * rhs
    dia = A.diagonal()
    return A, db, db / dia, dia


def _scipy_iterates(fn, A, b, x0, M, k, **kw):
    xs = []
    fn(A, b, x0=x0.copy(), M=M, maxiter=k, callback=lambda xk: xs.append(np.array(xk, copy=True)), **kw)
    return xs


# ------------------------------------------------------------------ PCG
def pcg_violations(plant="none", ne=(14, 11)):
    """Returns the list of broken properties of the traced PASO PCG."""
    ptr, idx, val, rhs = _heterogeneous_poisson(ne)
    x, dg, vbal, d, vecs, scal = O.traced_solve(1, ptr, idx, val, rhs, method="PCG", cap=400,
                                                tolerance=1e-10, plant=plant)
    A, db, x0, dia = _balanced(ptr, idx, vbal, d, rhs)
    nb = np.linalg.norm(db)
    K = len(vecs)
    bad = []
    # the textbook PCG iterate for iterate (Jacobi preconditioner)
    Minv = sp.diags(1.0 / dia)
    ref = _scipy_iterates(spla.cg, A, db, x0, Minv, min(K, 25), rtol=1e-300, atol=0.0)
    for k in range(min(len(ref), 25)):
        if np.linalg.norm(vecs[k, 2] - ref[k]) > 1e-9 * np.linalg.norm(ref[k]):
            bad.append("x2 differs from the textbook PCG at iteration %d" % (k + 1))
            break
    rs_prev = db - A @ x0
    rough = False
    r_prev_norm = np.linalg.norm(rs_prev)
    for k in range(K):
        xk, rsk, x2k, rk = vecs[k]
        if np.linalg.norm(rk - (db - A @ x2k)) > 1e-10 * nb:
            bad.append("r is not the residual of x2 at iteration %d" % (k + 1)); break
        if np.linalg.norm(rsk - (db - A @ xk)) > 1e-10 * nb:
            bad.append("rs is not the residual of x at iteration %d" % (k + 1)); break
        n_rs, n_r, n_prev = np.linalg.norm(rsk), np.linalg.norm(rk), np.linalg.norm(rs_prev)
        if n_rs > min(n_prev, n_r) * (1 + 1e-10) + 1e-14 * nb:
            bad.append("||rs|| grows at iteration %d" % (k + 1)); break
        dvec = rk - rs_prev
        if abs(rsk @ dvec) > 1e-8 * n_rs * np.linalg.norm(dvec) + 1e-28:
            bad.append("rs is not the minimal residual smoothing at iteration %d" % (k + 1)); break
        if abs(scal[k, 3] - n_rs) > 1e-12 * max(n_rs, 1e-300) + 1e-14 * nb:
            bad.append("reported norm is not ||rs|| at iteration %d" % (k + 1)); break
        rough |= n_r > r_prev_norm
        r_prev_norm = n_r
        rs_prev = rsk
    return bad, rough, dg, K


def test_pcg_paso_recurrence_pinned():
    bad, rough, dg, K = pcg_violations()
    assert dg["rc"] == 0 and K > 30
    assert rough, "test system too smooth: the CG residual never grows, so smoothing is not exercised"
    assert bad == []


@pytest.mark.parametrize("plant", ["pcg_no_smoothing", "pcg_gamma_sign"])
def test_pcg_paso_planted_mistake_is_caught(plant):
    bad, _, _, _ = pcg_violations(plant)
    assert bad, "the planted mistake %s went unnoticed" % plant


# ------------------------------------------------------------- BiCGStab
def bicgstab_violations(plant="none"):
    ptr, idx, val, rhs = _advection()
    x, dg, vbal, d, vecs, scal = O.traced_solve(1, ptr, idx, val, rhs, method="BICGSTAB", cap=400,
                                                tolerance=1e-12, plant=plant)
    A, db, x0, dia = _balanced(ptr, idx, vbal, d, rhs)
    nb = np.linalg.norm(db)
    K = len(vecs)
    bad = []
    Minv = sp.diags(1.0 / dia)
    ref = _scipy_iterates(spla.bicgstab, A, db, x0, Minv, min(K, 12), rtol=1e-300, atol=0.0)
    for k in range(min(len(ref), 12)):
        if np.linalg.norm(vecs[k, 0] - ref[k]) > 1e-8 * np.linalg.norm(ref[k]):
            bad.append("x differs from scipy's bicgstab at iteration %d" % (k + 1))
            break
    rtld = db - A @ x0
    for k in range(K):
        xk, rk, sk, tk = vecs[k]
        if np.linalg.norm(rk - (db - A @ xk)) > 1e-9 * nb:
            bad.append("r is not the residual of x at iteration %d" % (k + 1)); break
        if abs(rtld @ sk) > 1e-9 * np.linalg.norm(rtld) * max(np.linalg.norm(sk), 1e-300):
            bad.append("rtld.s != 0 at iteration %d" % (k + 1)); break
        if scal[k, 2] != 0.0 and abs(tk @ rk) > 1e-9 * np.linalg.norm(tk) * np.linalg.norm(rk):
            bad.append("t.r != 0 at iteration %d" % (k + 1)); break
    return bad, dg, K


def test_bicgstab_recurrence_pinned():
    bad, dg, K = bicgstab_violations()
    assert dg["rc"] == 0 and K > 8
    assert bad == []


@pytest.mark.parametrize("plant", ["bicgstab_beta", "bicgstab_omega_sign"])
def test_bicgstab_planted_mistake_is_caught(plant):
    bad, _, _ = bicgstab_violations(plant)
    assert bad, "the planted mistake %s went unnoticed" % plant


# --------------------------------------------------------------- MINRES
def minres_violations(plant="none", reaction=1.0):
    ptr, idx, val, rhs = O.assemble(2, (12, 10), (1.0, 1.0), A=np.eye(2), D=reaction, Y=1.0)
    N = len(ptr) - 1
    q = np.zeros(N)
    q[:13] = 1.0
    O.nullify(1, ptr, idx, val, q, q, 1.0)
    rhs = np.where(q > 0, 0.0, rhs)
    x, dg, vbal, d, vecs, scal = O.traced_solve(1, ptr, idx, val, rhs, method="MINRES", cap=400,
                                                tolerance=1e-12, plant=plant)
    A, db, x0, dia = _balanced(ptr, idx, vbal, d, rhs)
    K = len(vecs)
    bad = []
    Minv = sp.diags(1.0 / dia)
    ref = _scipy_iterates(spla.minres, A, db, x0, Minv, min(K, 20), rtol=1e-300)
    for k in range(min(len(ref), 20)):
        if np.linalg.norm(vecs[k, 0] - ref[k]) > 1e-8 * np.linalg.norm(ref[k]):
            bad.append("x differs from scipy's minres at iteration %d" % (k + 1))
            break
    r0 = db - A @ x0
    prev = n0 = np.sqrt(r0 @ (r0 / dia))
    for k in range(K):
        rk = db - A @ vecs[k, 0]
        nm = np.sqrt(rk @ (rk / dia))
        if nm > prev * (1 + 1e-9) + 1e-13 * n0:
            bad.append("||r||_M grows at iteration %d" % (k + 1)); break
        if abs(scal[k, 3] - nm) > 1e-6 * nm + 1e-11 * n0:
            bad.append("rnorm_prec is not ||r||_M at iteration %d" % (k + 1)); break
        prev = nm
    return bad, dg, K


@pytest.mark.parametrize("reaction", [1.0, -30.0])
def test_minres_recurrence_pinned(reaction):
    bad, dg, K = minres_violations(reaction=reaction)
    assert dg["rc"] == 0 and K > 10
    assert bad == []


def test_minres_planted_mistake_is_caught():
    bad, _, _ = minres_violations("minres_w_a3")
    assert bad


def test_plant_is_off_after_a_traced_call():
    # the switch applies to one traced call only
    minres_violations("minres_w_a3")
    bad, _, _ = minres_violations()
    assert bad == []
