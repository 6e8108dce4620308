"""The oracle's multilevel restatement (oracle/multilevel.py), pinned
independently of its own formulas: interpolation exact for trilinear
functions, Galerkin coarse operators equal to the Q1 assembly on the coarse
grid (nested conforming spaces), the Chebyshev smoother against the
closed-form Chebyshev error polynomial from an eigendecomposition, the
symmetry of the V-cycle, the balanced form against the physical one, and the
flexible CG against scipy's CG with the same fixed preconditioner.  Planted
mistakes (a wrong interpolation weight, a flipped Chebyshev coefficient, a
missing balance factor) must fail these checks."""
import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

import multilevel as ML
import oracle as O


def _nodes_coords(nodes):
    n0, n1, n2 = nodes
    z, y, x = np.meshgrid(np.arange(n2), np.arange(n1), np.arange(n0), indexing="ij")
    return x.ravel(), y.ravel(), z.ravel()


def _poisson(ne, kappa=None, dirichlet=True, seed=0):
    """Oracle Q1 assembly of -div(kappa grad u) on Brick(ne) (unit cube),
    u = 0 on z = 0 when dirichlet; returns (A csr, b)."""
    E = ne[0] * ne[1] * ne[2]
    if kappa is None:
        kappa = np.exp(np.random.default_rng(seed).standard_normal(E))
    ptr, idx, val, rhs = O.assemble(3, ne, (1.0, 1.0, 1.0), A=kappa[:, None, None] * np.eye(3)[None], Y=1.0)
    if dirichlet:
        nodes = tuple(n + 1 for n in ne)
        z = np.arange(np.prod(nodes)) // (nodes[0] * nodes[1])
        q = (z == 0).astype(np.float64)
        O.nullify(1, ptr, idx, val, q, q, 1.0)
        rhs = np.where(q > 0, 0.0, rhs)
    return ML.csr_from_paso(ptr, idx, val), rhs


def test_prolongation_reproduces_trilinear_functions():
    nc = (4, 3, 5)
    P = ML.prolongation(nc)
    nf = tuple(2 * n - 1 for n in nc)
    xc, yc, zc = _nodes_coords(nc)
    xf, yf, zf = _nodes_coords(nf)
    for f in (lambda x, y, z: 1 + 0 * x, lambda x, y, z: x - 2 * y + 3 * z,
              lambda x, y, z: x * y * z + x * z - y):
        assert np.allclose(P @ f(xc, yc, zc), f(xf / 2.0, yf / 2.0, zf / 2.0), atol=1e-13)


def test_prolongation_planted_weight_is_caught():
    Pd = ML.prolongation_1d(4).toarray()
    Pd[1, 0] = 0.6   # planted: wrong interpolation weight
    x = np.arange(4.0) + 1.0
    assert np.allclose(ML.prolongation_1d(4) @ x, np.arange(7) / 2.0 + 1.0)
    assert not np.allclose(Pd @ x, np.arange(7) / 2.0 + 1.0)


@pytest.mark.parametrize("ne", [(4, 4, 4), (6, 4, 2)])
def test_galerkin_equals_coarse_assembly_for_constant_coefficient(ne):
    # nested Q1 spaces: the fine stiffness restricted to the coarse space is
    # the coarse stiffness (no Dirichlet rows: a pure Neumann operator)
    E = ne[0] * ne[1] * ne[2]
    Af, _ = _poisson(ne, kappa=np.full(E, 2.5), dirichlet=False)
    nc = tuple(n // 2 + 1 for n in ne)
    P = ML.prolongation(nc)
    Ac_gal = (P.T @ Af @ P).toarray()
    nec = tuple(n // 2 for n in ne)
    Ac, _ = _poisson(nec, kappa=np.full(nec[0] * nec[1] * nec[2], 2.5), dirichlet=False)
    assert np.allclose(Ac_gal, Ac.toarray(), atol=1e-13 * np.abs(Ac_gal).max())


def test_galerkin_of_elementwise_kappa_equals_coarse_assembly_of_block_kappa():
    # kappa constant on each 2x2x2 block of fine elements: the coarse space's
    # functions are then integrated exactly by the coarse assembly with the
    # block's kappa
    ne = (4, 6, 4)
    nec = tuple(n // 2 for n in ne)
    kc = np.exp(np.random.default_rng(3).standard_normal(nec[0] * nec[1] * nec[2]))
    k2, k1, k0 = np.meshgrid(np.arange(ne[2]), np.arange(ne[1]), np.arange(ne[0]), indexing="ij")
    kf = kc[((k2 // 2) * nec[1] + k1 // 2) * nec[0] + k0 // 2].ravel()
    Af, _ = _poisson(ne, kappa=kf, dirichlet=False)
    P = ML.prolongation(tuple(n + 1 for n in nec))
    Ac, _ = _poisson(nec, kappa=kc, dirichlet=False)
    G = (P.T @ Af @ P).toarray()
    assert np.allclose(G, Ac.toarray(), atol=1e-13 * np.abs(G).max())


def _cheb_poly(t, k):
    return np.cos(k * np.arccos(t)) if abs(t) <= 1 else np.cosh(k * np.arccosh(abs(t))) * (np.sign(t) ** k)


@pytest.mark.parametrize("steps", [1, 2, 3, 5])
def test_chebyshev_smoother_against_error_polynomial(steps):
    A, _ = _poisson((3, 3, 3), seed=5)
    H = ML.Hierarchy(A, (4, 4, 4), levels=1, ratio=10.0)
    L = H.levels[0]
    Ad = A.toarray()
    d = np.diag(Ad)
    S = Ad / np.sqrt(np.outer(d, d))           # D^-1/2 A D^-1/2, same spectrum as D^-1 A
    lam, V = np.linalg.eigh(S)
    lmax, lmin = L.lmax, L.lmax / H.ratio
    theta, delta = 0.5 * (lmax + lmin), 0.5 * (lmax - lmin)
    q = np.array([_cheb_poly((theta - l) / delta, steps) / _cheb_poly(theta / delta, steps) for l in lam])
    xs = np.random.default_rng(1).standard_normal(len(d))
    b = Ad @ xs
    x = H.smooth(0, b, None, steps)
    # error e = x* - x_k = D^-1/2 q(S) D^1/2 x*
    e_ref = (V @ (q * (V.T @ (np.sqrt(d) * xs)))) / np.sqrt(d)
    assert np.allclose(xs - x, e_ref, atol=1e-10 * np.abs(xs).max())


def test_chebyshev_planted_coefficient_is_caught():
    A, _ = _poisson((3, 3, 3), seed=5)
    H = ML.Hierarchy(A, (4, 4, 4), levels=1)
    good = H.smooth(0, np.ones(64), None, 3)
    orig = H.cheb_coefficients

    def planted(l, steps):
        c = orig(l, steps)
        return [c[0]] + [(-c1, c2) for c1, c2 in c[1:]]   # planted: sign of the d coefficient
    H.cheb_coefficients = planted
    assert not np.allclose(H.smooth(0, np.ones(64), None, 3), good)


def test_vcycle_is_symmetric_and_converges():
    ne = (8, 8, 8)
    A, b = _poisson(ne, seed=2)
    H = ML.Hierarchy(A, (9, 9, 9), levels=3, max_coarse=1)
    assert len(H.levels) == 3
    rng = np.random.default_rng(4)
    u, v = rng.standard_normal(A.shape[0]), rng.standard_normal(A.shape[0])
    assert abs(u @ H.vcycle(v) - v @ H.vcycle(u)) <= 1e-10 * abs(u @ H.vcycle(v))
    _, it_mg = ML.pcg(A, b, H.vcycle, tol=1e-8, m=1)
    d = 1.0 / A.diagonal()
    _, it_j = ML.pcg(A, b, lambda r: d * r, tol=1e-8, m=1)
    assert it_mg <= 12 and it_j >= 2 * it_mg, (it_mg, it_j)


def test_balanced_form_matches_physical_form():
    # the V-cycle on the balanced system (balanced level 0, physical levels
    # below) equals D^-1 V_phys(D^-1 r) when both use the same level-0 bound
    ne = (8, 8, 4)
    A, _ = _poisson(ne, seed=6)
    ptr, idx, val = A.indptr, A.indices, A.data.copy()
    Ab = A.copy()
    bal = 1.0 / np.sqrt(np.asarray(abs(A).max(axis=1).todense()).ravel())
    Ab = sp.diags(bal) @ A @ sp.diags(bal)
    nodes = (9, 9, 5)
    Hb = ML.Hierarchy(Ab.tocsr(), nodes, levels=3, bal=bal, max_coarse=1)
    assert len(Hb.levels) == 3
    Hp = ML.Hierarchy(A, nodes, levels=3, max_coarse=1)
    Hp.levels[0].lmax = Hb.levels[0].lmax
    r = np.random.default_rng(8).standard_normal(A.shape[0])
    zb = Hb.vcycle(r)
    zp = Hp.vcycle(r / bal) / bal
    assert np.allclose(zb, zp, rtol=0, atol=1e-11 * np.abs(zp).max())
    # planted: the restriction without the balance factor
    Hb.bal = None
    assert not np.allclose(Hb.vcycle(r), zp, rtol=0, atol=1e-6 * np.abs(zp).max())


def test_flexible_cg_with_fixed_preconditioner_matches_cg():
    A, b = _poisson((6, 6, 6), seed=9)
    d = 1.0 / A.diagonal()
    count = [0]

    def cb(xk):
        count[0] += 1
    M = spla.LinearOperator(A.shape, matvec=lambda r: d * r)
    x_ref, info = spla.cg(A, b, rtol=1e-10, atol=0.0, M=M, callback=cb, maxiter=1000)
    for m in (1, 4):
        x, it = ML.pcg(A, b, lambda r: d * r, tol=1e-10, m=m)
        assert abs(it - count[0]) <= 1, (m, it, count[0])
        assert np.allclose(x, x_ref, rtol=0, atol=1e-8 * np.abs(x_ref).max())


def test_energy_scaled_coarse_step_minimises_energy_error():
    A, b = _poisson((4, 4, 4), seed=10)
    H = ML.Hierarchy(A, (5, 5, 5), levels=1)
    y = np.random.default_rng(2).standard_normal(A.shape[0])
    f = H.energy_scaled(lambda l, r: y)
    x = f(0, b)
    xs = spla.spsolve(A.tocsc(), b)

    def energy(a):
        e = xs - a * y
        return e @ (A @ e)
    alpha = x[0] / y[0]
    assert energy(alpha) <= min(energy(alpha * 1.01), energy(alpha * 0.99))
    # after Chebyshev steps the network's correction never raises the error energy
    H.coarse_steps = 4
    x0 = H.smooth(0, b, None, 4)

    def err_energy(v):
        e = xs - v
        return e @ (A @ e)
    assert err_energy(H.net_correct(0, b, x0, lambda l, r: y)) <= err_energy(x0) * (1 + 1e-12)
