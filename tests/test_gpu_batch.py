"""Batched solves on the GPU (through the C ABI): many right-hand sides on
one operator and many independent systems of uneven sizes, against the
oracle's CSR products and solves of each column / system on its own."""
import numpy as np
import pytest
import torch

import oracle as O
import ripley_hip as R
from gpu_helpers import dirichlet_mask, gpu_assemble, grid, to_dev
from ripley_hip import batch as RB

pytestmark = pytest.mark.gpu


def _system(dim, ne, seed, sigma=1.0):
    """Constrained Poisson system with log-normal kappa on the GPU and the
    same system in the oracle (CSR)."""
    rng = np.random.default_rng(seed)
    dom, length = grid(dim, ne)
    E = dom.getNumElements()
    kap = np.exp(sigma * rng.standard_normal(E))
    A = kap[:, None, None] * np.eye(dim)[None]
    Y = rng.uniform(0.5, 1.5, E)
    op, rhs = gpu_assemble(dom, 1, A=A, Y=Y)
    q = dirichlet_mask(dom, 0)
    op.nullifyRowsAndCols(to_dev(q), to_dev(q), 1.0)
    rhs = torch.where(to_dev(q) > 0, torch.zeros_like(rhs), rhs)
    ptr, idx, val, orhs = O.assemble(dim, ne, length, A=A, Y=Y)
    O.nullify(1, ptr, idx, val, q, q, 1.0)
    orhs = np.where(q > 0, 0.0, orhs)
    return op, rhs, (ptr, idx, val, orhs), q


def _opts(method, tol=1e-9):
    o = R.SolverOptions()
    o.setSolverMethod(getattr(R.SolverOptions, method))
    o.setTolerance(tol)
    o.setCheckInterval(16)
    return o


# (33, 9, 40): full and partial 16-node x tiles, partial 4-line y tiles and
# three z segments of the z-marching SpMM
@pytest.mark.parametrize("dim,ne", [(3, (9, 8, 7)), (3, (33, 9, 40)), (2, (30, 21))])
@pytest.mark.parametrize("K", [1, 5, 16, 64])
def test_matrix_vector_many(dim, ne, K):
    op, _, (ptr, idx, val, _), _ = _system(dim, ne, 1)
    rng = np.random.default_rng(K)
    X = rng.standard_normal((op.size, K))
    ref = np.stack([O.spmv(1, ptr, idx, val, X[:, k]) for k in range(K)], axis=1)
    scale = np.stack([O.spmv(1, ptr, idx, np.abs(val), np.abs(X[:, k])) for k in range(K)], axis=1)
    # every SpMM tile shape whose column count divides K (the z-marching
    # shape falls back to the first one on a Rectangle)
    for shape, (_, cpt) in enumerate(RB.SPMM_SHAPES):
        if K % cpt:
            continue
        RB.set_shape_many(op, shape)
        Y = RB.matrix_vector_many(op, to_dev(X)).cpu().numpy()
        assert np.all(np.abs(Y - ref) <= 1e-13 * scale + 1e-300), shape


@pytest.mark.parametrize("method", ["PCG_PLAIN", "PCG_MIXED"])
@pytest.mark.parametrize("K", [1, 7, 64])
@pytest.mark.parametrize("shape", list(range(len(RB.SPMM_SHAPES))))
def test_solve_many_matches_single_solves(method, K, shape):
    print("solve_many", method, K, shape, flush=True)
    op, rhs, (ptr, idx, val, orhs), q = _system(3, (10, 9, 8), 2, sigma=1.5)
    if K % RB.SPMM_SHAPES[shape][1]:
        pytest.skip("shape does not divide the column count")
    RB.set_shape_many(op, shape)
    rng = np.random.default_rng(3)
    B = rng.standard_normal((op.size, K))
    B[q > 0, :] = 0.0
    B[:, 0] = orhs
    X, diags = RB.solve_many(op, to_dev(B), _opts(method))
    X = X.cpu().numpy()
    assert all(d["converged"] for d in diags)
    for k in range(K):
        res = np.linalg.norm(B[:, k] - O.spmv(1, ptr, idx, val, X[:, k])) / np.linalg.norm(B[:, k])
        assert res <= 1e-9 * 1.001, (k, res)
    xo, dg, _, _ = O.solve(1, ptr, idx, val, B[:, K - 1], tolerance=1e-11, plain=True)
    assert np.linalg.norm(X[:, K - 1] - xo) <= 1e-6 * np.linalg.norm(xo)


def _pcg_iterate(ptr, idx, val, b, m):
    """The m-th iterate of textbook Jacobi-PCG from x = 0 (CSR, numpy)."""
    import scipy.sparse as sp
    A = sp.csr_matrix((val, idx, ptr), shape=(len(ptr) - 1, len(ptr) - 1))
    dinv = 1.0 / A.diagonal()
    x = np.zeros_like(b)
    r = b.copy()
    z = dinv * r
    p = z.copy()
    rho = r @ z
    for _ in range(m):
        q = A @ p
        alpha = rho / (p @ q)
        x += alpha * p
        r -= alpha * q
        z = dinv * r
        rho, rho_old = r @ z, rho
        p = z + (rho / rho_old) * p
    return x


# x is updated every second iteration (two directions at once) and a chunk
# ends on the second of a pair: stopping after an odd or even number of
# iterations, inside a chunk or at its end -- also after the first, before
# the second direction block was ever written -- must leave x at that
# iterate
@pytest.mark.parametrize("shape", [0, 6])
@pytest.mark.parametrize("m", [1, 2, 7, 10, 16, 17])
def test_solve_many_iteration_limit_leaves_that_iterate(shape, m):
    print("solve_many iteration limit", shape, m, flush=True)
    op, rhs, (ptr, idx, val, orhs), q = _system(3, (10, 9, 8), 2, sigma=1.5)
    K = 64
    RB.set_shape_many(op, shape)
    rng = np.random.default_rng(5)
    B = rng.standard_normal((op.size, K))
    B[q > 0, :] = 0.0
    o = _opts("PCG_PLAIN", tol=1e-14)
    o.setIterMax(m)
    o.setAcceptanceConvergenceFailureOn()
    X, diags = RB.solve_many(op, to_dev(B), o)
    X = X.cpu().numpy()
    assert all(d["num_iter"] == m and not d["converged"] for d in diags)
    for k in (0, 1, K - 1):
        ref = _pcg_iterate(ptr, idx, val, B[:, k], m)
        assert np.linalg.norm(X[:, k] - ref) <= 1e-10 * np.linalg.norm(ref), k


# columns that start converged (zero right-hand sides) next to running ones
# on the 64-column z-marching shapes, FP32 directions with PCG_MIXED: the
# ended columns stay exactly zero through the alternating direction blocks
@pytest.mark.parametrize("shape", list(RB.ZMARCH_SHAPES))
def test_solve_many_zero_columns_zmarch_fp32_directions(shape):
    op, rhs, (ptr, idx, val, orhs), q = _system(3, (12, 10, 9), 4, sigma=1.5)
    K = 64
    RB.set_shape_many(op, shape)
    rng = np.random.default_rng(7)
    B = rng.standard_normal((op.size, K))
    B[q > 0, :] = 0.0
    zero = [3, 17, 40]
    B[:, zero] = 0.0
    X, diags = RB.solve_many(op, to_dev(B), _opts("PCG_MIXED"))
    X = X.cpu().numpy()
    assert all(d["converged"] for d in diags)
    for k in range(K):
        if k in zero:
            assert np.all(X[:, k] == 0.0) and diags[k]["num_iter"] == 0
        else:
            res = np.linalg.norm(B[:, k] - O.spmv(1, ptr, idx, val, X[:, k])) / np.linalg.norm(B[:, k])
            assert res <= 1e-9 * 1.001, (k, res)


def test_solve_many_zero_column_and_errors():
    op, rhs, (ptr, idx, val, orhs), q = _system(2, (16, 12), 4)
    B = np.stack([orhs, np.zeros_like(orhs), 2.0 * orhs], axis=1)
    X, diags = RB.solve_many(op, to_dev(B), _opts("PCG_PLAIN"))
    X = X.cpu().numpy()
    assert np.all(X[:, 1] == 0.0) and diags[1]["converged"] and diags[1]["num_iter"] == 0
    assert np.allclose(X[:, 2], 2.0 * X[:, 0], rtol=1e-7, atol=1e-14)
    with pytest.raises(ValueError):
        RB.solve_many(op, to_dev(B), _opts("PCG"))
    # a matrix balanced by an ordinary solve is refused until its values are reset
    o = R.SolverOptions()
    o.setSolverMethod(R.SolverOptions.PCG)
    op.solve(rhs, o)
    with pytest.raises(R.RipleyHipError):
        RB.solve_many(op, to_dev(B), _opts("PCG_PLAIN"))


def test_multi_autotune():
    op, *_ = _system(3, (16, 16, 16), 5)
    best, times = RB.autotune_many(op, 32, fp32=True, reps=3)
    # every shape but the 64-column z-marching ones applies to 32 columns
    ok = [j for j in range(len(times)) if j not in RB.ZMARCH_SHAPES]
    assert best in ok and all(times[j] > 0 for j in ok) and all(times[j] == -1 for j in RB.ZMARCH_SHAPES)
    assert times[best] == min(times[j] for j in ok)
    best, times = RB.autotune_many(op, 64, fp32=True, reps=3)
    assert all(t > 0 for t in times) and times[best] == min(times)
    # a column count that only the one-column shape divides
    best, times = RB.autotune_many(op, 7, fp32=False, reps=2)
    assert best == 0 and times[0] > 0 and all(t == -1 for t in times[1:])


def _batch(dim, shapes, seed):
    systems = [_system(dim, ne, seed + j) for j, ne in enumerate(shapes)]
    return systems, RB.BatchSolver([s[0] for s in systems])


def test_batch_vector_uneven_2d():
    shapes = [(3, 4), (17, 9), (40, 33), (64, 64), (100, 7), (1, 1)]
    systems, bs = _batch(2, shapes, 10)
    rng = np.random.default_rng(0)
    xs = [rng.standard_normal(s[0].size) for s in systems]
    y = bs.split(bs.MatrixVector(bs.join(xs)))
    for (op, _, (ptr, idx, val, _), _), x, yj in zip(systems, xs, y):
        ref = O.spmv(1, ptr, idx, val, x)
        scale = O.spmv(1, ptr, idx, np.abs(val), np.abs(x))
        assert np.all(np.abs(yj.cpu().numpy() - ref) <= 1e-13 * scale + 1e-300)


@pytest.mark.parametrize("method", ["PCG_PLAIN", "PCG_MIXED"])
@pytest.mark.parametrize("tile", [0, 1, 8])
def test_batch_solve_uneven_2d(method, tile):
    print("batch", method, tile, flush=True)
    shapes = [(5, 4), (33, 17), (64, 48), (90, 90), (12, 70)]
    systems, bs = _batch(2, shapes, 20)
    bs.set_tile(tile)
    b = bs.join([s[1] for s in systems])
    x, diags = bs.solve(b, _opts(method))
    assert all(d["converged"] for d in diags)
    for (op, _, (ptr, idx, val, orhs), _), xj, d in zip(systems, bs.split(x), diags):
        xj = xj.cpu().numpy()
        res = np.linalg.norm(orhs - O.spmv(1, ptr, idx, val, xj)) / np.linalg.norm(orhs)
        assert res <= 1e-9 * 1.001
        xo, dg, _, _ = O.solve(1, ptr, idx, val, orhs, tolerance=1e-11, plain=True)
        assert np.linalg.norm(xj - xo) <= 1e-6 * np.linalg.norm(xo)
        # per-system iteration counts: each system stops on its own
        assert d["num_iter"] <= 2 * dg["num_iter"] + 32


def test_batch_solve_3d_and_autotune():
    shapes = [(6, 5, 4), (12, 10, 9), (3, 3, 3)]
    systems, bs = _batch(3, shapes, 30)
    best, times = bs.autotune(fp32=False, reps=3)
    assert best in RB.BatchSolver.TILES and times[RB.BatchSolver.TILES.index(best)] == min(times)
    b = bs.join([s[1] for s in systems])
    x, diags = bs.solve(b, _opts("PCG_PLAIN"))
    for (op, _, (ptr, idx, val, orhs), _), xj in zip(systems, bs.split(x)):
        res = np.linalg.norm(orhs - O.spmv(1, ptr, idx, val, xj.cpu().numpy())) / np.linalg.norm(orhs)
        assert res <= 1e-9 * 1.001
