"""Pins the oracle's PASO solver restatement (paso/src/Solver.cpp:46-331,
PCG.cpp:62-368) with results the mathematics fixes: direct solves of small
SPD systems, the reference's constant-solution and linear-exact-solution
tests (escriptcore/test/python/test_linearPDEs.py:1741-1748,
test_simplesolve.py:76-160), the balance/Jacobi closed forms, and the
error paths of solve.cpp:119-143."""
import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

import oracle as O


def poisson_2d(ne, kappa=None):
    E = int(np.prod(ne))
    A = np.eye(2) if kappa is None else kappa[:, None, None] * np.eye(2)[None]
    ptr, idx, val, rhs = O.assemble(2, ne, (1.0, 1.0), A=A, Y=1.0)
    return ptr, idx, val, rhs


def constrain(ptr, idx, val, rhs, q, r, bs=1):
    """insertConstraint (linearPDEs.py:2084-2112) on the CSR system."""
    u = np.where(q > 0, r, 0.0)
    rhs = O.spmv(bs, ptr, idx, val, u, alpha=-1.0, beta=1.0, y=rhs)
    val = val.copy()
    O.nullify(bs, ptr, idx, val, q, q, 1.0)
    rhs = np.where(q > 0, r, rhs)
    return val, rhs


def test_spmv_matches_dense():
    ptr, idx, val, _ = poisson_2d((5, 4))
    M = O.csr_to_dense(1, ptr, idx, val)
    x = np.random.default_rng(0).standard_normal(M.shape[0])
    y0 = np.random.default_rng(1).standard_normal(M.shape[0])
    assert np.allclose(O.spmv(1, ptr, idx, val, x, 0.5, -2.0, y0), 0.5 * M @ x - 2.0 * y0, atol=1e-14)


def test_nullify_semantics():
    ptr, idx, val, _ = poisson_2d((3, 3))
    N = len(ptr) - 1
    q = np.zeros(N); q[[0, 5]] = 1.0
    v = val.copy()
    O.nullify(1, ptr, idx, v, q, q, 1.0)
    M0, M = O.csr_to_dense(1, ptr, idx, val), O.csr_to_dense(1, ptr, idx, v)
    for i in range(N):
        for j in range(N):
            if q[i] > 0 or q[j] > 0:
                assert M[i, j] == (1.0 if i == j and (idx[ptr[i]:ptr[i + 1]] == j).any() else 0.0)
            else:
                assert M[i, j] == M0[i, j]


def test_balance_closed_form():
    ptr, idx, val, _ = poisson_2d((4, 4), kappa=np.linspace(1, 5, 16))
    M0 = O.csr_to_dense(1, ptr, idx, val)
    v = val.copy()
    d = O.balance(1, ptr, idx, v)
    assert np.allclose(d, 1 / np.sqrt(np.abs(M0).max(axis=1)))
    assert np.allclose(O.csr_to_dense(1, ptr, idx, v), d[:, None] * M0 * d[None, :])


def test_inv_main_blocks():
    rng = np.random.default_rng(4)
    for bs in (1, 2, 3):
        ptr, idx = O.pattern(2, (2, 2))
        val = rng.standard_normal(len(idx) * bs * bs)
        inv = O.inv_main(bs, ptr, idx, val)
        for i in range(len(ptr) - 1):
            k = ptr[i] + np.where(idx[ptr[i]:ptr[i + 1]] == i)[0][0]
            blk = val[k * bs * bs:(k + 1) * bs * bs].reshape(bs, bs, order="F")
            assert np.allclose(inv[i * bs * bs:(i + 1) * bs * bs].reshape(bs, bs, order="F") @ blk, np.eye(bs))


def test_inv_main_singular_raises():
    ptr, idx = O.pattern(2, (1, 1))
    with pytest.raises(RuntimeError, match="non-regular"):
        O.inv_main(1, ptr, idx, np.zeros(len(idx)))


@pytest.mark.parametrize("plain", [False, True])
def test_solve_matches_direct(plain):
    rng = np.random.default_rng(3)
    ne = (12, 9)
    kap = np.exp(rng.standard_normal(int(np.prod(ne))))
    ptr, idx, val, rhs = poisson_2d(ne, kap)
    N = len(ptr) - 1
    q = np.zeros(N); q[: ne[0] + 1] = 1.0                  # y = 0 edge
    val, rhs = constrain(ptr, idx, val, rhs, q, np.zeros(N))
    M = sp.csr_matrix((val, idx, ptr), shape=(N, N))
    x_ref = spla.spsolve(M.tocsc(), rhs)
    x, dg, _, _ = O.solve(1, ptr, idx, val, rhs, tolerance=1e-10, plain=plain)
    assert dg["rc"] == 0 and dg["converged"] == 1
    assert np.linalg.norm(x - x_ref) <= 1e-8 * np.linalg.norm(x_ref)
    res = rhs - M @ x
    assert np.linalg.norm(res) <= 1e-10 * np.linalg.norm(rhs) * 10


def test_constant_solution_pcg_jacobi():
    # test_linearPDEs.py:1741-1748: A = I, D = 1, Y = 1  =>  u == 1
    ne = (8, 8, 8)
    ptr, idx, val, rhs = O.assemble(3, ne, (1.0, 1.0, 1.0), A=np.eye(3), D=1.0, Y=1.0)
    x, dg, _, _ = O.solve(1, ptr, idx, val, rhs, tolerance=1e-8)
    assert dg["rc"] == 0
    assert np.abs(x - 1.0).max() <= 1e-6


def test_linear_exact_solution():
    # test_simplesolve.py:76-160 idea: Q1 reproduces u = 1+2x+3y+4z exactly;
    # Dirichlet data on the whole boundary, A = I, no source.
    ne = (6, 5, 4)
    ptr, idx, val, rhs = O.assemble(3, ne, (1.0, 1.0, 1.0), A=np.eye(3))
    nn = [n + 1 for n in ne]
    g = np.meshgrid(*[np.linspace(0, 1, n) for n in nn[::-1]], indexing="ij")
    X = np.stack([a.reshape(-1) for a in g[::-1]], axis=1)
    u_ex = 1 + 2 * X[:, 0] + 3 * X[:, 1] + 4 * X[:, 2]
    q = np.any((X < 1e-12) | (X > 1 - 1e-12), axis=1).astype(float)
    val, rhs = constrain(ptr, idx, val, rhs, q, u_ex)
    x, dg, _, _ = O.solve(1, ptr, idx, val, rhs, tolerance=1e-8)
    assert np.abs(x - u_ex).max() <= 1e-6 * np.abs(u_ex).max()


def test_zero_rhs_gives_zero():
    ptr, idx, val, _ = poisson_2d((4, 4))
    x, dg, _, _ = O.solve(1, ptr, idx, val, np.zeros(len(ptr) - 1))
    assert dg["rc"] == 0 and np.all(x == 0)


def test_error_paths():
    ptr, idx, val, rhs = poisson_2d((6, 6))
    N = len(ptr) - 1
    q = np.zeros(N); q[0] = 1
    val, rhs = constrain(ptr, idx, val, rhs, q, np.zeros(N))
    assert O.solve(1, ptr, idx, val, rhs, tolerance=1e-16)[1]["rc"] == 100   # too small
    assert O.solve(1, ptr, idx, val, rhs, tolerance=2.0)[1]["rc"] == 100     # > 1
    bad = rhs.copy(); bad[3] = np.nan
    assert O.solve(1, ptr, idx, val, bad)[1]["rc"] == 102
    _, dg, _, _ = O.solve(1, ptr, idx, val, rhs, iter_max=3)
    assert dg["rc"] == 1                                                     # MaxIterReached


def test_sweeps_two_converges():
    ptr, idx, val, rhs = poisson_2d((10, 10))
    N = len(ptr) - 1
    q = np.zeros(N); q[:11] = 1
    val, rhs = constrain(ptr, idx, val, rhs, q, np.zeros(N))
    x1, d1, _, _ = O.solve(1, ptr, idx, val, rhs, sweeps=1)
    x2, d2, _, _ = O.solve(1, ptr, idx, val, rhs, sweeps=2)
    assert d1["rc"] == 0 and d2["rc"] == 0
    assert np.linalg.norm(x1 - x2) <= 1e-6 * np.linalg.norm(x1)


def _advection_system(dim, ne, seed, n=1):
    """Non-symmetric constrained system: -lap u + B . grad u = 1 (n = 1) or
    an elasticity-like system with a skew first-order term (n > 1)."""
    rng = np.random.default_rng(seed)
    E = int(np.prod(ne))
    if n == 1:
        A = np.tile(np.eye(dim), (E, 1, 1))
        B = rng.standard_normal((E, dim)) * 3.0
        ptr, idx, val, rhs = O.assemble(dim, ne, (1.0,) * dim, A=A, B=B, Y=1.0)
    else:
        eye = np.eye(dim)
        A = np.tile(np.einsum("ik,jl->ijkl", np.eye(n), eye), (E, 1, 1, 1, 1))
        B = rng.standard_normal((E, n, dim, n)) * 2.0
        ptr, idx, val, rhs = O.assemble(dim, ne, (1.0,) * dim, n=n, A=A, B=B, Y=np.ones((E, n)))
    N = len(ptr) - 1
    q = np.zeros(N)
    q[: ne[0] + 1] = 1.0
    qn = np.repeat(q, n)
    O.nullify(n, ptr, idx, val, qn, qn, 1.0)
    rhs = np.where(qn > 0, 0.0, rhs)
    return ptr, idx, val, rhs


@pytest.mark.parametrize("dim,ne,n", [(2, (12, 10), 1), (3, (6, 5, 4), 1), (2, (8, 7), 2)])
def test_bicgstab_restatement_solves_nonsymmetric(dim, ne, n):
    import scipy.sparse as sp
    import scipy.sparse.linalg as spl
    ptr, idx, val, rhs = _advection_system(dim, ne, 4, n)
    D = O.csr_to_dense(n, ptr, idx, val)
    assert np.abs(D - D.T).max() > 1e-3          # genuinely non-symmetric
    x, dg, _, _ = O.solve(n, ptr, idx, val, rhs, tolerance=1e-10, method="BICGSTAB")
    assert dg["converged"] == 1 and dg["rc"] == 0
    xs = spl.spsolve(sp.csr_matrix(D), rhs)
    assert np.linalg.norm(x - xs) <= 1e-8 * np.linalg.norm(xs)
    assert np.linalg.norm(rhs - D @ x) <= 1e-9 * np.linalg.norm(rhs)


def test_bicgstab_on_spd_matches_pcg_solution():
    ptr, idx, val, rhs = (lambda r: (r[0], r[1], r[2], r[3]))(O.assemble(2, (10, 9), (1.0, 1.0),
                                                                           A=np.eye(2), Y=1.0))
    N = len(ptr) - 1
    q = np.zeros(N)
    q[:11] = 1.0
    O.nullify(1, ptr, idx, val, q, q, 1.0)
    rhs = np.where(q > 0, 0.0, rhs)
    xb, db, _, _ = O.solve(1, ptr, idx, val, rhs, tolerance=1e-10, method="BICGSTAB")
    xp, dp, _, _ = O.solve(1, ptr, idx, val, rhs, tolerance=1e-10)
    assert np.linalg.norm(xb - xp) <= 1e-8 * np.linalg.norm(xp)


def _reaction_system(D, ne=(10, 9)):
    """-lap u + D u = 1 on Rectangle(ne), u = 0 on the x1 = 0 edge."""
    ptr, idx, val, rhs = O.assemble(2, ne, (1.0, 1.0), A=np.eye(2), D=D, Y=1.0)
    N = len(ptr) - 1
    q = np.zeros(N)
    q[: ne[0] + 1] = 1.0
    val, rhs = constrain(ptr, idx, val, rhs, q, np.zeros(N))
    return ptr, idx, val, rhs


@pytest.mark.parametrize("prec", ["JACOBI", "NONE"])
def test_minres_restatement_spd_matches_direct(prec):
    # Solver_MINRES (paso/src/MINRES.cpp:59-191) on an SPD system
    ptr, idx, val, rhs = _reaction_system(1.0)
    N = len(ptr) - 1
    M = sp.csr_matrix((val, idx, ptr), shape=(N, N))
    x, dg, _, _ = O.solve(1, ptr, idx, val, rhs, tolerance=1e-10, method="MINRES", preconditioner=prec)
    assert dg["rc"] == 0 and dg["converged"] == 1
    xs = spla.spsolve(M.tocsc(), rhs)
    assert np.linalg.norm(x - xs) <= 1e-8 * np.linalg.norm(xs)


def test_minres_restatement_symmetric_indefinite():
    # a negative reaction term makes the matrix indefinite while its diagonal
    # (and so the Jacobi preconditioner) stays positive: MINRES converges
    # where CG has no guarantee
    ptr, idx, val, rhs = _reaction_system(-30.0)
    N = len(ptr) - 1
    M = sp.csr_matrix((val, idx, ptr), shape=(N, N)).toarray()
    ev = np.linalg.eigvalsh(M)
    assert ev.min() < 0 < ev.max()
    assert np.all(np.diag(M) > 0)
    x, dg, _, _ = O.solve(1, ptr, idx, val, rhs, tolerance=1e-10, method="MINRES")
    assert dg["rc"] == 0 and dg["converged"] == 1
    xs = np.linalg.solve(M, rhs)
    assert np.linalg.norm(x - xs) <= 1e-8 * np.linalg.norm(xs)


def test_minres_restatement_negative_norm():
    # MINRES.cpp:88-90: r.M^-1 r < 0 (negative definite system, Jacobi)
    # -> NegativeNormError
    ptr, idx, val, rhs = _reaction_system(1.0)
    x, dg, _, _ = O.solve(1, ptr, idx, -val, rhs, tolerance=1e-8, method="MINRES")
    assert dg["rc"] == 5


def test_minres_restatement_iteration_count_is_minimal_residual():
    # MINRES minimises ||r||_{M^-1} over the Krylov space: without a
    # preconditioner and with the outer loop's first step only it needs no
    # more iterations than CG on an SPD system (same Krylov space)
    ptr, idx, val, rhs = _reaction_system(1.0, ne=(20, 17))
    _, dm, _, _ = O.solve(1, ptr, idx, val, rhs, tolerance=1e-8, method="MINRES", preconditioner="NONE")
    _, dc, _, _ = O.solve(1, ptr, idx, val, rhs, tolerance=1e-8, method="PCG_PLAIN", preconditioner="NONE")
    assert dm["rc"] == 0 and dc["rc"] == 0
    assert dm["num_iter"] <= dc["num_iter"] + 2


def test_minres_restatement_block_system_matches_direct():
    # a coupled 2-component SPD system (block size 2): D = M M^T + 2 I per
    # element, A = I per component
    rng = np.random.default_rng(12)
    ne = (9, 8)
    E = int(np.prod(ne))
    n = 2
    A = np.zeros((E, n, 2, n, 2))
    for k in range(n):
        A[:, k, :, k, :] = np.eye(2) * (1.0 + k)
    M = rng.standard_normal((n, n))
    D = np.tile(M @ M.T + n * np.eye(n), (E, 1, 1))
    ptr, idx, val, rhs = O.assemble(2, ne, (1.0, 1.0), n=n, A=A, D=D, Y=np.ones((E, n)))
    dense = O.csr_to_dense(n, ptr, idx, val)
    x, dg, _, _ = O.solve(n, ptr, idx, val, rhs, tolerance=1e-10, method="MINRES")
    assert dg["rc"] == 0 and dg["converged"] == 1
    xs = np.linalg.solve(dense, rhs)
    assert np.linalg.norm(x - xs) <= 1e-8 * np.linalg.norm(xs)
