"""Pins the oracle's multi-colour preconditioners:

* the colouring (Pattern::borrowColoringPointer, paso/src/Pattern.cpp:166-195,
  with Pattern::mis, Pattern_mis.cpp:44-110) against an independent
  algorithm: rounds of "every available vertex whose priority is below all
  available neighbours' joins" give the lexicographically-first maximal
  independent set in priority order, i.e. the sequential greedy set, so each
  colour must equal the greedy independent set of the still uncoloured
  vertices taken in increasing fmod(0.4142135623730951 (i+1), 1);
* the coloured Gauss-Seidel sweep (Smoother.cpp:331-507) against dense
  triangular solves: forward (D + L) y = b, backward (D + U) x = D y with
  L / U the couplings to lower / higher colours;
* ILU0 (ILU.cpp:43-462) against the defining property of ILU(0): with
  L = lower blocks + pivots and U = I + scaled upper blocks, (L U)_ij = a_ij
  on every pattern position, and the solve returns (L U)^-1 b;
* PCG with each preconditioner on the reference's own constant-solution
  tests (escriptcore/test/python/test_linearPDEs.py:1741-1768 scalar with
  Gauss-Seidel(80) / ILU0, :2147-2178 the systems with Gauss-Seidel(130) /
  ILU0), u = 1 to Lsup 1e-6."""
import numpy as np
import pytest

import oracle as O

SEED = 0.4142135623730951


def neighbours(ptr, idx, i):
    return [int(j) for j in idx[ptr[i]:ptr[i + 1]] if j != i]


def greedy_colouring(ptr, idx):
    n = len(ptr) - 1
    prio = np.fmod(SEED * np.arange(1, n + 1, dtype=np.float64), 1.0)
    order = np.argsort(prio, kind="stable")
    colour = -np.ones(n, dtype=np.int64)
    c = 0
    while (colour < 0).any():
        blocked = np.zeros(n, dtype=bool)
        for i in order:
            if colour[i] < 0 and not blocked[i]:
                colour[i] = c
                for j in neighbours(ptr, idx, i):
                    blocked[j] = True
        c += 1
    return colour, c


GRIDS = [(4, 3), (7, 5), (1, 6), (9, 9), (3, 2, 4), (5, 4, 3), (1, 1, 3), (6, 6, 6)]


@pytest.mark.parametrize("ne", GRIDS)
def test_colouring_is_greedy_by_priority(ne):
    ptr, idx = O.pattern(len(ne), ne)
    col, nc = O.coloring(ptr, idx)
    ref, rnc = greedy_colouring(ptr, idx)
    assert nc == rnc and np.array_equal(col, ref)
    # a proper colouring: no two neighbours share a colour
    for i in range(len(ptr) - 1):
        assert all(col[j] != col[i] for j in neighbours(ptr, idx, i))


def test_colour_counts_of_the_benchmark_shapes():
    # sizes of the colour classes are a property of the pattern alone; the
    # 3-D stencil needs at least 8 colours (a 2x2x2 block is a clique)
    ptr, idx = O.pattern(3, (16, 16, 16))
    col, nc = O.coloring(ptr, idx)
    assert 8 <= nc <= 27 and np.bincount(col).sum() == 17 ** 3
    ptr, idx = O.pattern(2, (16, 16))
    col, nc = O.coloring(ptr, idx)
    assert 4 <= nc <= 9


def system(dim, ne, n, seed):
    """An assembled SPD system (random kappa Laplacian per component, coupled
    positive definite reaction term)."""
    rng = np.random.default_rng(seed)
    E = int(np.prod(ne))
    length = (1.0,) * dim
    if n == 1:
        A = rng.uniform(0.5, 2.0, E)[:, None, None] * np.eye(dim)[None]
        ptr, idx, val, rhs = O.assemble(dim, ne, length, A=A, D=rng.uniform(0.1, 1.0, E), Y=1.0)
    else:
        A = np.zeros((n, dim, n, dim))
        for k in range(n):
            A[k, :, k, :] = np.eye(dim) * (1.0 + 0.3 * k)
        M = rng.standard_normal((n, n))
        D = M @ M.T + n * np.eye(n)
        ptr, idx, val, rhs = O.assemble(dim, ne, length, n=n, A=A, D=D, Y=np.ones(n))
    return ptr, idx, val, rhs


def block_masks(ptr, idx, col, bs):
    N = len(ptr) - 1
    node_col = np.repeat(col, bs)
    rows = np.arange(N * bs)
    ci = node_col[rows][:, None]
    cj = node_col[rows][None, :]
    same_node = (rows[:, None] // bs) == (rows[None, :] // bs)
    return ci, cj, same_node


@pytest.mark.parametrize("dim,ne,n", [(2, (5, 4), 1), (3, (3, 3, 2), 1), (2, (4, 3), 2), (3, (2, 3, 2), 3)])
def test_gs_sweep_is_symmetric_gauss_seidel(dim, ne, n):
    ptr, idx, val, _ = system(dim, ne, n, seed=dim + n)
    col, nc = O.coloring(ptr, idx)
    inv = O.inv_main(n, ptr, idx, val)
    Mat = O.csr_to_dense(n, ptr, idx, val)
    ci, cj, same = block_masks(ptr, idx, col, n)
    # rows of one colour are never coupled (except inside a node's block)
    assert np.all(Mat[(ci == cj) & ~same] == 0)
    Dm = np.where(same, Mat, 0.0)
    Lm = np.where(cj < ci, Mat, 0.0)
    Um = np.where(cj > ci, Mat, 0.0)
    b = np.random.default_rng(7).standard_normal(Mat.shape[0])
    y = np.linalg.solve(Dm + Lm, b)
    x = np.linalg.solve(Dm + Um, Dm @ y)
    got = O.gs_sweep(n, ptr, idx, val, inv, col, nc, b)
    assert np.allclose(got, x, rtol=1e-12, atol=1e-12 * np.abs(x).max())


def test_gs_sweeps_converge_to_the_solution():
    # Preconditioner_Smoother_solve with many sweeps is a convergent
    # iteration for an SPD matrix: x -> A^-1 b
    ptr, idx, val, _ = system(2, (4, 4), 1, seed=3)
    b = np.random.default_rng(1).standard_normal(len(ptr) - 1)
    Mat = O.csr_to_dense(1, ptr, idx, val)
    col, nc = O.coloring(ptr, idx)
    inv = O.inv_main(1, ptr, idx, val)
    x = O.gs_sweep(1, ptr, idx, val, inv, col, nc, b)
    for _ in range(4000):
        x = x + O.gs_sweep(1, ptr, idx, val, inv, col, nc, b - Mat @ x)
    assert np.allclose(x, np.linalg.solve(Mat, b), rtol=1e-10, atol=1e-12)


def ilu_factors_dense(F, ptr, idx, col, bs):
    N = len(ptr) - 1
    L = np.zeros((N * bs, N * bs))
    U = np.eye(N * bs)
    for i in range(N):
        for k in range(ptr[i], ptr[i + 1]):
            j = idx[k]
            blk = F[k * bs * bs:(k + 1) * bs * bs].reshape(bs, bs, order="F")
            sl_i, sl_j = slice(i * bs, (i + 1) * bs), slice(j * bs, (j + 1) * bs)
            if j == i:
                L[sl_i, sl_j] = np.linalg.inv(blk)    # the factor stores the inverted pivot
            elif col[j] < col[i]:
                L[sl_i, sl_j] = blk
            else:
                U[sl_i, sl_j] = blk
    return L, U


@pytest.mark.parametrize("dim,ne,n", [(2, (5, 4), 1), (3, (3, 2, 3), 1), (2, (3, 4), 2), (3, (2, 2, 3), 3)])
def test_ilu0_matches_matrix_on_pattern(dim, ne, n):
    ptr, idx, val, _ = system(dim, ne, n, seed=10 + dim + n)
    col, nc = O.coloring(ptr, idx)
    F, rc = O.ilu0_factor(n, ptr, idx, val, col, nc)
    assert rc == 0
    L, U = ilu_factors_dense(F, ptr, idx, col, n)
    LU = L @ U
    Mat = O.csr_to_dense(n, ptr, idx, val)
    scale = np.abs(Mat).max()
    for i in range(len(ptr) - 1):
        for k in range(ptr[i], ptr[i + 1]):
            j = idx[k]
            sl_i, sl_j = slice(i * n, (i + 1) * n), slice(j * n, (j + 1) * n)
            assert np.allclose(LU[sl_i, sl_j], Mat[sl_i, sl_j], rtol=0, atol=1e-12 * scale), (i, j)
    b = np.random.default_rng(2).standard_normal(Mat.shape[0])
    z = O.ilu0_solve(n, ptr, idx, F, col, nc, b)
    assert np.allclose(LU @ z, b, rtol=0, atol=1e-11 * np.abs(b).max())


def test_ilu0_singular_pivot_and_block_size_limit():
    ptr, idx = O.pattern(2, (2, 2))
    val = np.zeros(len(idx))
    col, nc = O.coloring(ptr, idx)
    assert O.ilu0_factor(1, ptr, idx, val, col, nc)[1] == 1
    val4 = np.zeros(len(idx) * 16)
    assert O.ilu0_factor(4, ptr, idx, val4, col, nc)[1] == 2


def constant_solution_system(dim, ne, n):
    """test_linearPDEs.py: A = kronecker per component, D = 1 (scalar) or the
    all-ones tensor plus diag(0..dim-1), Y = 1 or dim + i; u = 1."""
    length = (1.0,) * dim
    if n == 1:
        return O.assemble(dim, ne, length, A=np.eye(dim), D=1.0, Y=1.0)
    A = np.zeros((n, dim, n, dim))
    for k in range(n):
        A[k, :, k, :] = np.eye(dim)
    D = np.ones((n, n)) + np.diag(np.arange(n, dtype=float))
    Y = n + np.arange(n, dtype=float)
    return O.assemble(dim, ne, length, n=n, A=A, D=D, Y=Y)


@pytest.mark.parametrize("dim", [2, 3])
@pytest.mark.parametrize("prec,sweeps", [("GAUSS_SEIDEL", 80), ("ILU0", 1), ("NONE", 1), ("JACOBI", 1)])
@pytest.mark.parametrize("system_", [False, True])
def test_reference_constant_solution(dim, prec, sweeps, system_):
    ne = (8, 7) if dim == 2 else (5, 4, 4)
    n = dim if system_ else 1
    if system_ and prec == "GAUSS_SEIDEL":
        sweeps = 130
    ptr, idx, val, rhs = constant_solution_system(dim, ne, n)
    x, dg, _, _ = O.solve(n, ptr, idx, val, rhs, tolerance=1e-10, preconditioner=prec, sweeps=sweeps)
    assert dg["rc"] == 0 and dg["converged"] == 1
    assert np.abs(x - 1.0).max() <= 1e-6


def test_stronger_preconditioners_take_fewer_iterations():
    rng = np.random.default_rng(4)
    ne = (12, 12)
    A = rng.uniform(0.2, 5.0, 144)[:, None, None] * np.eye(2)[None]
    ptr, idx, val, rhs = O.assemble(2, ne, (1.0, 1.0), A=A, D=0.01, Y=1.0)
    its = {}
    for p in ("NONE", "JACOBI", "GAUSS_SEIDEL", "ILU0"):
        x, dg, _, _ = O.solve(1, ptr, idx, val, rhs, tolerance=1e-10, preconditioner=p)
        assert dg["rc"] == 0
        its[p] = dg["num_iter"]
    assert its["GAUSS_SEIDEL"] < its["JACOBI"] and its["ILU0"] < its["JACOBI"], its
