"""AbstractSystemMatrix::saveMM (paso/src/SparseMatrix.cpp:182-240) from the
GPU matrix: MatrixMarket coordinate real general, scalar rows/columns
(block entries unrolled), 1-based, entries in PASO's order (block rows, then
the row's columns ascending, then the block's rows and columns), 15
significant digits; the written matrix equals the oracle's assembly."""
import numpy as np
import pytest
import scipy.io

import oracle as O
from gpu_helpers import gpu_assemble, grid

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("dim,ne,n,sym", [(2, (5, 4), 1, False), (3, (3, 3, 2), 1, True), (2, (3, 2), 3, False),
                                          (3, (2, 2, 2), 2, True)])
def test_save_mm(tmp_path, dim, ne, n, sym):
    rng = np.random.default_rng(dim * 10 + n)
    dom, length = grid(dim, ne)
    E = dom.getNumElements()
    if n == 1:
        A = rng.uniform(0.5, 2.0, E)[:, None, None] * np.eye(dim)[None]
        D = rng.uniform(0.1, 1.0, E)
        Y = np.ones(E)
    else:
        A = np.zeros((E, n, dim, n, dim))
        for k in range(n):
            A[:, k, :, k, :] = np.eye(dim) * (1.0 + k)
        M = rng.standard_normal((n, n))
        D = np.tile(M @ M.T + n * np.eye(n), (E, 1, 1))
        Y = np.ones((E, n))
    op, _ = gpu_assemble(dom, n, symmetric=sym, A=A, D=D, Y=Y)
    fn = str(tmp_path / "a.mtx")
    op.saveMM(fn)
    ptr, idx, oval, _ = O.assemble(dim, ne, length, n=n, A=A, D=D, Y=Y)
    ref = O.csr_to_dense(n, ptr, idx, oval)
    lines = open(fn).read().splitlines()
    assert lines[0] == "%%MatrixMarket matrix coordinate real general"
    N = (len(ptr) - 1) * n
    assert lines[1].split() == [str(N), str(N), str(len(idx) * n * n)]
    # PASO's entry order
    order = [(i * n + r + 1, int(idx[k]) * n + c + 1) for i in range(len(ptr) - 1)
             for k in range(ptr[i], ptr[i + 1]) for r in range(n) for c in range(n)]
    got = [tuple(int(t) for t in ln.split()[:2]) for ln in lines[2:]]
    assert got == order
    M = scipy.io.mmread(fn).toarray()
    assert np.allclose(M, ref, rtol=1e-13, atol=1e-14 * np.abs(ref).max())
