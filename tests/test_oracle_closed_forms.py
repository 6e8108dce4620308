"""Pins the oracle's pattern and element/assembly arithmetic with what the
mathematics and the reference's own constants fix:

* pattern: nnz = (3n-2)^d, ascending rows, in-bounds neighbours only
  (Brick.cpp:3000-3061), brute force on tiny grids;
* element matrices: exact 1-D integrals combined as tensor products (an
  independent derivation of the Q1 integrals, not the oracle's quadrature
  loop), for every coefficient kind, and the reference's constants
  (DefaultAssembler2D.cpp:175-183,310-314, DefaultAssembler3D.cpp:204-231,
  :907-928, :1923, :2186-2214);
* assembly: dense brute-force assembly of tiny grids from those element
  matrices, symmetry, zero row sums, the patch test, kappa linearity;
* quadrature point order of expanded data, pinned by the reference's
  expanded-coefficient constants.
"""
import itertools

import numpy as np
import pytest

import oracle as O

S3 = np.sqrt(3.0)


# ---------------------------------------------------------------- pattern
@pytest.mark.parametrize("dim,ne", [(2, (4, 3)), (2, (1, 1)), (3, (3, 2, 4)), (3, (1, 1, 1))])
def test_pattern_closed_form(dim, ne):
    ptr, idx = O.pattern(dim, ne)
    nn = [n + 1 for n in ne]
    N = int(np.prod(nn))
    assert len(ptr) == N + 1
    assert ptr[-1] == np.prod([3 * n - 2 for n in nn])
    coords = np.array(list(itertools.product(*[range(n) for n in nn[::-1]])))[:, ::-1]
    for i in range(N):
        row = idx[ptr[i]:ptr[i + 1]]
        assert np.all(np.diff(row) > 0)
        expect = [j for j in range(N) if np.max(np.abs(coords[j] - coords[i])) <= 1]
        assert list(row) == expect


# --------------------------------------------------- 1-D exact integrals
def one_d(h):
    """Exact 1-D integrals of the linear hat functions on [0,h]:
    K[a,b] = int phi_a' phi_b', M[a,b] = int phi_a phi_b,
    G[a,b] = int phi_a' phi_b, L[a] = int phi_a, Lp[a] = int phi_a'."""
    K = np.array([[1, -1], [-1, 1]]) / h
    M = h / 6.0 * np.array([[2, 1], [1, 2]])
    G = 0.5 * np.array([[-1, -1], [1, 1]])
    L = np.array([h / 2, h / 2])
    Lp = np.array([-1.0, 1.0])
    return K, M, G, L, Lp


def bits(a, dim):
    return [(a >> d) & 1 for d in range(dim)]


def exact_element(dim, h, A=None, B=None, C=None, D=None, X=None, Y=None):
    """Dense element matrix S[a,b] (a row/test, b column/trial) and load F[a]
    of a single PDE with constant coefficients, from tensor products."""
    tabs = [one_d(hh) for hh in h]
    nN = 1 << dim
    S = np.zeros((nN, nN))
    F = np.zeros(nN)
    for a in range(nN):
        ab = bits(a, dim)
        for b in range(nN):
            bb = bits(b, dim)

            def prod(f):
                return np.prod([f(d) for d in range(dim)])

            s = 0.0
            if A is not None:
                for i in range(dim):
                    for j in range(dim):
                        def f(d):
                            K, M, G, _, _ = tabs[d]
                            if d == i and d == j:
                                return K[ab[d], bb[d]]
                            if d == i:          # d/dx_d on the test function only
                                return G[ab[d], bb[d]]
                            if d == j:          # d/dx_d on the trial function only
                                return G[bb[d], ab[d]]
                            return M[ab[d], bb[d]]
                        s += A[i, j] * prod(f)
            if B is not None:                   # B_i N_b dN_a/dx_i
                for i in range(dim):
                    s += B[i] * prod(lambda d: tabs[d][2][ab[d], bb[d]] if d == i else tabs[d][1][ab[d], bb[d]])
            if C is not None:                   # C_j dN_b/dx_j N_a
                for j in range(dim):
                    s += C[j] * prod(lambda d: tabs[d][2][bb[d], ab[d]] if d == j else tabs[d][1][ab[d], bb[d]])
            if D is not None:
                s += D * prod(lambda d: tabs[d][1][ab[d], bb[d]])
            S[a, b] = s
        f = 0.0
        if X is not None:
            for i in range(dim):
                f += X[i] * prod(lambda d: tabs[d][4][ab[d]] if d == i else tabs[d][3][ab[d]])
        if Y is not None:
            f += Y * prod(lambda d: tabs[d][3][ab[d]])
        F[a] = f
    return S, F


def oracle_S(EM, nN, n=1):
    """EM_S[k + n*(m + n*(a + nN*b))] -> array [a, b, k, m]."""
    return EM.reshape(nN, nN, n, n).transpose(1, 0, 3, 2)


@pytest.mark.parametrize("dim,h", [(2, (0.3, 0.7)), (3, (0.5, 0.25, 0.125)), (3, (1.0, 1.0, 1.0))])
def test_element_matrices_all_coefficients(dim, h):
    rng = np.random.default_rng(7)
    A = rng.standard_normal((dim, dim))
    B, C, X = rng.standard_normal(dim), rng.standard_normal(dim), rng.standard_normal(dim)
    D, Y = 1.7, -0.4
    EM, EF = O.element_matrices(dim, h, A=A, B=B, C=C, D=D, X=X, Y=Y)
    S, F = exact_element(dim, h, A=A, B=B, C=C, D=D, X=X, Y=Y)
    nN = 1 << dim
    assert np.allclose(oracle_S(EM, nN)[:, :, 0, 0], S, rtol=0, atol=1e-14 * np.abs(S).max())
    assert np.allclose(EF, F, rtol=0, atol=1e-14 * np.abs(F).max())


def test_element_planted_transpose_is_caught():
    # a transposed A (row/column swap) must not pass the comparison above
    rng = np.random.default_rng(3)
    A = rng.standard_normal((3, 3))
    EM, _ = O.element_matrices(3, (0.5, 0.5, 0.5), A=A.T)
    S, _ = exact_element(3, (0.5, 0.5, 0.5), A=A)
    assert np.abs(oracle_S(EM, 8)[:, :, 0, 0] - S).max() > 1e-3


def test_reference_constants_2d():
    # DefaultAssembler2D.cpp:175-183, :310, :579 on the unit square, A = I
    EM, EF = O.element_matrices(2, (1.0, 1.0), A=np.eye(2), Y=1.0)
    S = oracle_S(EM, 4)[:, :, 0, 0]
    w3, w6 = 1.0 / 24, -1.0 / 24
    assert np.isclose(S[0, 0], -8 * w6 + 8 * w3, rtol=1e-15)        # 2/3
    assert np.isclose(S[0, 1], -1.0 / 6, rtol=1e-14)
    assert np.isclose(S[0, 3], -1.0 / 3, rtol=1e-14)
    assert np.allclose(EF, 36 * (1.0 / 144))                          # h0 h1 / 4


def test_reference_constants_3d():
    h = 0.5
    EM, EF = O.element_matrices(3, (h, h, h), A=np.eye(3), Y=1.0)
    S = oracle_S(EM, 8)[:, :, 0, 0]
    w27 = -h * h / (288 * h)
    assert np.isclose(S[0, 0], -4 * 3 * (8 * w27), rtol=1e-14)        # h/3  (:907, :928)
    assert abs(S[0, 1]) < 1e-16 and abs(S[0, 2]) < 1e-16             # edge-adjacent: 0
    assert np.isclose(S[0, 3], -h / 12, rtol=1e-13)                   # face diagonal
    assert np.isclose(S[0, 7], -h / 12, rtol=1e-13)                   # body diagonal
    assert np.allclose(EF, 216 * h ** 3 / 1728)                       # h^3/8  (:2206)


def test_reference_reduced_constants_3d():
    # assemblePDESingleReduced: D weight h0h1h2/64, A_00 weight h1h2/(16 h0)
    h = (0.5, 0.25, 0.2)
    EM, _ = O.element_matrices(3, h, A=np.diag([1.0, 0, 0]), reduced=True)
    S = oracle_S(EM, 8)[:, :, 0, 0]
    assert np.isclose(S[0, 0], h[1] * h[2] / (16 * h[0]), rtol=1e-14)
    assert np.isclose(S[1, 0], -h[1] * h[2] / (16 * h[0]), rtol=1e-14)
    EM, _ = O.element_matrices(3, h, D=1.0, reduced=True)
    assert np.allclose(EM, h[0] * h[1] * h[2] / 64)


def test_expanded_quadrature_order_2d():
    # DefaultAssembler2D.cpp:229-232, :289: in EM_S(1,0) the coefficient of
    # A_00 at points 0,1 is w0 = w6 (sqrt3+2) and at points 2,3 w4 = w6 (2-sqrt3)
    h0, h1 = 0.4, 0.9
    w6 = -h1 / (24 * h0)
    expect = [w6 * (S3 + 2), w6 * (S3 + 2), w6 * (2 - S3), w6 * (2 - S3)]
    E = 1
    for q in range(4):
        A = np.zeros((E, 4, 2, 2))
        A[0, q, 0, 0] = 1.0
        EM, _ = O.element_matrices(2, (h0, h1), A=A, expanded=("A",))
        S = oracle_S(EM, 4)[:, :, 0, 0]
        assert np.isclose(S[1, 0], expect[q], rtol=1e-13), q


def test_expanded_quadrature_order_3d():
    # D_0 / D_7 in EM_S(0,0): w59 = w55 (15 sqrt3 + 26), w60 = w55 (26 - 15 sqrt3)
    # (DefaultAssembler3D.cpp:205-209, :1923); Y in EM_F[0]: w70 for Y_0,
    # w71 for Y_1,2,4, w72 for Y_3,5,6, w73 for Y_7 (:210-213, :2186-2200)
    h = (0.3, 0.5, 0.7)
    w55 = h[0] * h[1] * h[2] / 1728
    wD = {0: w55 * (15 * S3 + 26), 7: w55 * (26 - 15 * S3)}
    wY = {0: w55 * 6 * (5 * S3 + 9), 7: w55 * 6 * (9 - 5 * S3)}
    for q in (1, 2, 4):
        wY[q] = w55 * 6 * (S3 + 3)
    for q in (3, 5, 6):
        wY[q] = w55 * 6 * (3 - S3)
    for q in range(8):
        D = np.zeros((1, 8))
        D[0, q] = 1.0
        EM, EF = O.element_matrices(3, h, D=D, Y=D, expanded=("D", "Y"))
        if q in wD:
            assert np.isclose(EM[0], wD[q], rtol=1e-13)
        assert np.isclose(EF[0], wY[q], rtol=1e-13), q


def test_expanded_equal_points_equals_constant_branch():
    rng = np.random.default_rng(11)
    A = rng.standard_normal((3, 3))
    Ae = np.broadcast_to(A, (1, 8, 3, 3)).copy()
    EMc, _ = O.element_matrices(3, (0.2, 0.3, 0.4), A=A)
    EMe, _ = O.element_matrices(3, (0.2, 0.3, 0.4), A=Ae, expanded=("A",))
    assert np.allclose(EMc, EMe, rtol=0, atol=1e-14 * np.abs(EMc).max())


# ----------------------------------------------------------------- systems
def test_system_element_matches_blockwise_single():
    # EM_S[k,m,a,b] for A[k,i,m,j] equals the single-PDE matrix of A[k,:,m,:]
    rng = np.random.default_rng(5)
    n, dim, h = 3, 3, (0.5, 0.4, 0.3)
    A = rng.standard_normal((n, dim, n, dim))
    D = rng.standard_normal((n, n))
    Y = rng.standard_normal(n)
    X = rng.standard_normal((n, dim))
    EM, EF = O.element_matrices(dim, h, n=n, A=A, D=D, X=X, Y=Y)
    Sys = oracle_S(EM, 8, n)
    for k in range(n):
        for m in range(n):
            S, _ = exact_element(dim, h, A=A[k, :, m, :], D=D[k, m])
            assert np.allclose(Sys[:, :, k, m], S, atol=1e-13)
        _, F = exact_element(dim, h, X=X[k], Y=Y[k])
        assert np.allclose(EF.reshape(8, n)[:, k], F, atol=1e-14)


# ---------------------------------------------------------------- assembly
def brute_assemble(dim, ne, length, elem_S, elem_F, n=1):
    """Dense global matrix / rhs by scattering per-element dense blocks."""
    nn = [k + 1 for k in ne]
    N = int(np.prod(nn))
    K = np.zeros((N * n, N * n))
    F = np.zeros(N * n)
    nN = 1 << dim
    for e, kk in enumerate(itertools.product(*[range(k) for k in ne[::-1]])):
        k = kk[::-1]
        first = sum(k[d] * int(np.prod(nn[:d])) for d in range(dim))
        nodes = [first + sum(((a >> d) & 1) * int(np.prod(nn[:d])) for d in range(dim)) for a in range(nN)]
        S, Fe = elem_S(e), elem_F(e)
        for a in range(nN):
            for b in range(nN):
                K[nodes[a] * n:(nodes[a] + 1) * n, nodes[b] * n:(nodes[b] + 1) * n] += S[a, b]
            F[nodes[a] * n:(nodes[a] + 1) * n] += Fe[a]
    return K, F


@pytest.mark.parametrize("dim,ne", [(2, (3, 4)), (3, (3, 2, 2))])
def test_assembly_matches_brute_force(dim, ne):
    rng = np.random.default_rng(1)
    E = int(np.prod(ne))
    length = tuple(0.5 + 0.25 * d for d in range(dim))
    h = [l / k for l, k in zip(length, ne)]
    kap = np.exp(rng.standard_normal(E))
    A = kap[:, None, None] * np.eye(dim)[None]
    A[:, 0, 1] += 0.1 * kap          # non-symmetric part
    Dv = rng.uniform(0.5, 1.5, E)
    Yv = rng.standard_normal(E)
    ptr, idx, val, rhs = O.assemble(dim, ne, length, A=A, D=Dv, Y=Yv)
    M = O.csr_to_dense(1, ptr, idx, val)

    def eS(e):
        return exact_element(dim, h, A=A[e], D=Dv[e])[0][:, :, None, None]

    def eF(e):
        return exact_element(dim, h, Y=Yv[e])[1][:, None]

    K, F = brute_assemble(dim, ne, length, eS, eF)
    scale = np.abs(K).max(axis=1, keepdims=True)
    assert np.all(np.abs(M - K) <= 1e-13 * scale)
    assert np.allclose(rhs, F, rtol=0, atol=1e-14 * np.abs(F).max())


def test_assembly_invariants_3d():
    ne = (4, 3, 5)
    ptr, idx, val, _ = O.assemble(3, ne, (1.0, 1.0, 1.0), A=np.eye(3))
    M = O.csr_to_dense(1, ptr, idx, val)
    assert np.allclose(M, M.T, atol=1e-15)
    assert np.abs(M.sum(axis=1)).max() < 1e-14          # pure A: zero row sums
    # patch test: linear fields are in the kernel on interior rows
    nn = [n + 1 for n in ne]
    X = np.array(list(itertools.product(*[np.linspace(0, 1, n) for n in nn[::-1]])))[:, ::-1]
    u = 1 + 2 * X[:, 0] - 3 * X[:, 1] + 0.5 * X[:, 2]
    r = M @ u
    interior = np.all((X > 1e-12) & (X < 1 - 1e-12), axis=1)
    assert np.abs(r[interior]).max() < 1e-13
    # linear in kappa
    _, _, val2, _ = O.assemble(3, ne, (1.0, 1.0, 1.0), A=2.5 * np.eye(3))
    assert np.allclose(val2, 2.5 * val, rtol=1e-15, atol=1e-16)


def test_system_assembly_matches_brute_force():
    dim, ne, n = 3, (2, 2, 1), 3
    rng = np.random.default_rng(2)
    E = int(np.prod(ne))
    lam, mu = rng.uniform(1, 2, E), rng.uniform(1, 2, E)
    eye = np.eye(dim)
    A = (lam[:, None, None, None, None] * np.einsum("ij,kl->ijkl", eye, eye)[None]
         + mu[:, None, None, None, None] * (np.einsum("ik,jl->ijkl", eye, eye) + np.einsum("il,jk->ijkl", eye, eye))[None])
    Y = np.array([0.0, 0.0, -1.0])
    ptr, idx, val, rhs = O.assemble(dim, ne, (1.0, 1.0, 1.0), n=n, A=A, Y=Y)
    M = O.csr_to_dense(n, ptr, idx, val)
    h = [1.0 / k for k in ne]

    def eS(e):
        S = np.zeros((8, 8, n, n))
        for k in range(n):
            for m in range(n):
                S[:, :, k, m] = exact_element(dim, h, A=A[e, k, :, m, :])[0]
        return S

    def eF(e):
        return np.stack([exact_element(dim, h, Y=Y[k])[1] for k in range(n)], axis=1)

    K, F = brute_assemble(dim, ne, (1.0, 1.0, 1.0), eS, eF, n=n)
    assert np.all(np.abs(M - K) <= 1e-13 * np.abs(K).max(axis=1, keepdims=True))
    assert np.allclose(rhs, F, atol=1e-15)
    assert np.allclose(M, M.T, atol=1e-13 * np.abs(M).max())


def test_add_to_csr_missing_pair_reported():
    ptr, idx = O.pattern(2, (2, 2))
    val = np.zeros(len(idx))
    # nodes 0 and 8 are not neighbours: 2 missing (row,col) pairs
    miss = O.add_to_csr(1, ptr, idx, val, [0, 8], np.ones(4))
    assert miss == 2


def test_reference_constants_b_c_x_3d():
    # constant B / C / X branches of DefaultAssembler3D::assemblePDESingle:
    # EM_S(1,0) += -4 wB0 with wB0 = B0 w53, w53 = -h1 h2/72 (:1353-1357);
    # EM_S(1,0) += 4 wC0 (:1781-1785); EM_F[0] = wX0 + wX1 - wX2 with
    # wX0 = 12 X0 w66, w66 = -h1 h2/48, wX2 = 18 X2 w50, w50 = h0 h1/72 (:2157-2161)
    h = (0.3, 0.5, 0.7)
    w53 = -h[1] * h[2] / 72
    EM, _ = O.element_matrices(3, h, B=np.array([1.0, 0, 0]))
    assert np.isclose(oracle_S(EM, 8)[1, 0, 0, 0], -4 * w53, rtol=1e-13)
    EM, _ = O.element_matrices(3, h, C=np.array([1.0, 0, 0]))
    assert np.isclose(oracle_S(EM, 8)[1, 0, 0, 0], 4 * w53, rtol=1e-13)
    X = np.array([0.7, -0.2, 1.3])
    _, EF = O.element_matrices(3, h, X=X)
    w66, w64, w50 = -h[1] * h[2] / 48, -h[0] * h[2] / 48, h[0] * h[1] / 72
    wX0, wX1, wX2 = 12 * X[0] * w66, 12 * X[1] * w64, 18 * X[2] * w50
    assert np.isclose(EF[0], wX0 + wX1 - wX2, rtol=1e-13)
    assert np.isclose(EF[7], -wX0 - wX1 + wX2, rtol=1e-13)
