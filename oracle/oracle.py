"""ORACLE -- test infrastructure only.

ctypes binding of ``oracle/liboracle.so``, the CPU restatement of the
ripley Q1 assembly + PASO Jacobi-PCG path (see ``ripley_oracle.cpp`` for the
reference file:line each function follows).  Only ``tests/``,
``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py`` may
import this module, and only as the checker or the timed CPU baseline.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

MODE_EMPTY, MODE_ELEMENT, MODE_EXPANDED = 0, 1, 2
SOLVER_RESULT = {0: "NoError", 1: "MaxIterReached", 2: "InputError", 3: "MemoryError",
                 4: "Breakdown", 5: "NegativeNormError", 6: "Divergence"}


class _Coef(ctypes.Structure):
    _fields_ = [("data", ctypes.POINTER(ctypes.c_double)),
                ("mode", ctypes.c_int),
                ("elem_stride", ctypes.c_int)]


class Options(ctypes.Structure):
    _fields_ = [("tolerance", ctypes.c_double), ("iter_max", ctypes.c_int),
                ("sweeps", ctypes.c_int), ("plain", ctypes.c_int), ("verbose", ctypes.c_int),
                ("preconditioner", ctypes.c_int)]

PRECONDITIONERS = {"JACOBI": 0, "GAUSS_SEIDEL": 1, "ILU0": 2, "NONE": 3}


class Diagnostics(ctypes.Structure):
    _fields_ = [("num_iter", ctypes.c_int), ("converged", ctypes.c_int), ("status", ctypes.c_int),
                ("outer_steps", ctypes.c_int), ("residual_norm", ctypes.c_double),
                ("set_up_time", ctypes.c_double), ("net_time", ctypes.c_double),
                ("time", ctypes.c_double)]


def build():
    subprocess.check_call(["make", "-s", "-C", _HERE])


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        P = ctypes.c_void_p
        I, I64, D = ctypes.c_int, ctypes.c_int64, ctypes.c_double
        CP = ctypes.POINTER(_Coef)
        sig = {
            "orc_num_nodes": (I64, [I, P]),
            "orc_diagonal_indices": (I, [I, P, I, P]),
            "orc_pattern": (I64, [I, P, P, P]),
            "orc_element_matrices": (None, [I, P, I, I, CP, CP, CP, CP, CP, CP, P, P]),
            "orc_add_to_csr": (I, [I, I64, P, P, P, I, P, P]),
            "orc_assemble": (None, [I, P, P, I, I, CP, CP, CP, CP, CP, CP, P, P, P, P]),
            "orc_assemble_boundary": (None, [I, P, P, I, I, CP, CP, P, P, P, P]),
            "orc_add_dirac": (None, [I, I64, P, P, P, I64, P, P, P, P]),
            "orc_spmv": (None, [I, I64, P, P, P, D, P, D, P]),
            "orc_nullify": (None, [I, I64, P, P, P, P, P, D]),
            "orc_balance": (None, [I, I64, P, P, P, P]),
            "orc_inv_main": (I, [I, I64, P, P, P, P]),
            "orc_solve": (I, [I, I64, P, P, P, P, P, ctypes.POINTER(Options),
                              ctypes.POINTER(Diagnostics), P]),
            "orc_time_pcg_iterations": (D, [I, I64, P, P, P, P, I, I, P]),
            "orc_coloring": (I, [I64, P, P, P]),
            "orc_gs_sweep": (None, [I, I64, P, P, P, P, P, I, P]),
            "orc_ilu0_factor": (I, [I, I64, P, P, P, P, I, P]),
            "orc_ilu0_solve": (None, [I, I64, P, P, P, P, I, P, P]),
            "orc_gs_apply": (None, [I, I64, P, P, P, P, P, I, I, P, P]),
            "orc_cds_add": (None, [I, I, I, I, P, I, P, P]),
            "orc_cds_spmv": (None, [I, I, I64, I, P, P, P, P]),
            "orc_trace_begin": (None, [P, I, I64]),
            "orc_trace_count": (I, []),
            "orc_set_plant": (None, [I]),
            "orc_num_threads": (I, []),
            "orc_set_num_threads": (None, [I]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p) if a is not None else None


def set_num_threads(n):
    lib().orc_set_num_threads(int(n))


def num_threads():
    return lib().orc_num_threads()


def num_nodes(dim, ne):
    return int(np.prod([n + 1 for n in ne[:dim]]))


def diagonal_indices(dim, ndof, upper_only=False):
    nd = np.asarray(list(ndof) + [1] * (3 - len(ndof)), dtype=np.int32)
    out = np.zeros(27, dtype=np.int32)
    k = lib().orc_diagonal_indices(dim, _p(nd), int(upper_only), _p(out))
    return out[:k]


def pattern(dim, ne):
    """PASO CSR pattern (ptr, index) of a one-rank Rectangle/Brick."""
    nn = np.asarray([n + 1 for n in ne[:dim]] + [1] * (3 - dim), dtype=np.int32)
    N = int(np.prod(nn))
    ptr = np.zeros(N + 1, dtype=np.int32)
    nnz = lib().orc_pattern(dim, _p(nn), _p(ptr), None)
    index = np.zeros(nnz, dtype=np.int32)
    lib().orc_pattern(dim, _p(nn), _p(ptr), _p(index))
    return ptr, index


class Coef:
    """A PDE coefficient in escript layout.

    ``value`` is a data point (same for every element), an array of shape
    (num_elements, *point) for per-element constants, or
    (num_elements, num_quad, *point) with ``expanded=True``.  Points are given
    in logical index order (C order) and stored column-major as escript does.
    """

    def __init__(self, value=None, point_shape=(), expanded=False):
        self.mode = MODE_EMPTY
        self.arr = None
        self.stride = 0
        if value is None:
            return
        v = np.asarray(value, dtype=np.float64)
        npt = len(point_shape)
        if v.ndim == npt:
            self.mode, self.stride = MODE_ELEMENT, 0
            self.arr = (np.ascontiguousarray(v.transpose(tuple(range(npt))[::-1])).ravel()
                        if npt else v.reshape(1).copy())
        elif v.ndim == npt + 1 and not expanded:
            self.mode = MODE_ELEMENT
            t = v.transpose((0,) + tuple(range(1, npt + 1))[::-1]) if npt else v
            self.arr = np.ascontiguousarray(t).ravel()
            self.stride = int(np.prod(point_shape)) if npt else 1
        elif v.ndim == npt + 2:
            self.mode = MODE_EXPANDED
            t = v.transpose((0, 1) + tuple(range(2, npt + 2))[::-1]) if npt else v
            self.arr = np.ascontiguousarray(t).ravel()
            self.stride = int(np.prod(point_shape) if npt else 1) * v.shape[1]
        else:
            raise ValueError("bad coefficient shape %s for point shape %s" % (v.shape, point_shape))

    def c(self):
        c = _Coef()
        c.data = self.arr.ctypes.data_as(ctypes.POINTER(ctypes.c_double)) if self.arr is not None else None
        c.mode = self.mode
        c.elem_stride = self.stride
        return c


def _coefs(dim, n, A=None, B=None, C=None, D=None, X=None, Y=None, expanded=()):
    if n == 1:
        shapes = dict(A=(dim, dim), B=(dim,), C=(dim,), D=(), X=(dim,), Y=())
    else:
        shapes = dict(A=(n, dim, n, dim), B=(n, dim, n), C=(n, n, dim), D=(n, n), X=(n, dim), Y=(n,))
    vals = dict(A=A, B=B, C=C, D=D, X=X, Y=Y)
    return {k: Coef(vals[k], shapes[k], expanded=(k in expanded)) for k in "ABCDXY"}


def element_matrices(dim, h, n=1, expanded=(), reduced=False, **coefs):
    """EM_S[k,m,a,b] (C-order array of shape (nN, nN, n, n) indexed [b,a,m,k])
    and EM_F, as flat arrays in escript layout."""
    cf = _coefs(dim, n, expanded=expanded, **coefs)
    nN = 1 << dim
    EM_S = np.zeros(n * n * nN * nN)
    EM_F = np.zeros(n * nN)
    hh = np.asarray(list(h) + [1.0] * (3 - len(h)), dtype=np.float64)
    cs = {k: cf[k].c() for k in "ABCDXY"}
    lib().orc_element_matrices(dim, _p(hh), n, int(reduced), *[ctypes.byref(cs[k]) for k in "ABCDXY"],
                               _p(EM_S), _p(EM_F))
    return EM_S, EM_F


def assemble(dim, ne, length, n=1, ptr=None, index=None, expanded=(), reduced=False,
             val=None, rhs=None, **coefs):
    """Assemble the PASO CSR matrix values and RHS; returns (ptr, index, val, rhs)."""
    if ptr is None:
        ptr, index = pattern(dim, ne)
    N = len(ptr) - 1
    val = np.zeros(len(index) * n * n) if val is None else val
    rhs = np.zeros(N * n) if rhs is None else rhs
    ne_a = np.asarray(list(ne) + [1] * (3 - len(ne)), dtype=np.int32)
    ln = np.asarray(list(length) + [1.0] * (3 - len(length)), dtype=np.float64)
    cf = _coefs(dim, n, expanded=expanded, **coefs)
    cs = {k: cf[k].c() for k in "ABCDXY"}
    lib().orc_assemble(dim, _p(ne_a), _p(ln), n, int(reduced), *[ctypes.byref(cs[k]) for k in "ABCDXY"],
                       _p(ptr), _p(index), _p(val), _p(rhs))
    return ptr, index, val, rhs


def assemble_boundary(dim, ne, length, n=1, ptr=None, index=None, reduced=False, val=None, rhs=None,
                      d=None, y=None, expanded=()):
    """Boundary terms d (n x n per point) and y (n per point) on the face
    elements; coefficients per face element (reference FunctionOnBoundary
    order), or per face element and point with 'd'/'y' in ``expanded``."""
    if ptr is None:
        ptr, index = pattern(dim, ne)
    N = len(ptr) - 1
    val = np.zeros(len(index) * n * n) if val is None else val
    rhs = np.zeros(N * n) if rhs is None else rhs
    ne_a = np.asarray(list(ne) + [1] * (3 - len(ne)), dtype=np.int32)
    ln = np.asarray(list(length) + [1.0] * (3 - len(length)), dtype=np.float64)
    pd = () if n == 1 else (n, n)
    py = () if n == 1 else (n,)
    cd = Coef(d, pd, expanded="d" in expanded)
    cy = Coef(y, py, expanded="y" in expanded)
    lib().orc_assemble_boundary(dim, _p(ne_a), _p(ln), n, int(reduced), ctypes.byref(cd.c()), ctypes.byref(cy.c()),
                                _p(ptr), _p(index), _p(val), _p(rhs))
    return ptr, index, val, rhs


def add_dirac(n, ptr, index, val, nodes, d=None, y=None, rhs=None):
    nodes = np.ascontiguousarray(nodes, dtype=np.int64)
    dd = None if d is None else np.ascontiguousarray(d, dtype=np.float64)
    yy = None if y is None else np.ascontiguousarray(y, dtype=np.float64)
    lib().orc_add_dirac(n, len(ptr) - 1, _p(ptr), _p(index), _p(val) if val is not None else None, len(nodes),
                        _p(nodes), _p(dd) if dd is not None else None, _p(yy) if yy is not None else None,
                        _p(rhs) if rhs is not None else None)


def add_to_csr(bs, ptr, index, val, nodes, EM):
    nodes = np.ascontiguousarray(nodes, dtype=np.int32)
    EM = np.ascontiguousarray(EM, dtype=np.float64)
    return lib().orc_add_to_csr(bs, len(ptr) - 1, _p(ptr), _p(index), _p(val), len(nodes),
                                _p(nodes), _p(EM))


def spmv(bs, ptr, index, val, x, alpha=1.0, beta=0.0, y=None):
    N = len(ptr) - 1
    x = np.ascontiguousarray(x, dtype=np.float64)
    y = np.zeros(N * bs) if y is None else np.ascontiguousarray(y, dtype=np.float64).copy()
    lib().orc_spmv(bs, N, _p(ptr), _p(index), _p(val), alpha, _p(x), beta, _p(y))
    return y


def nullify(bs, ptr, index, val, mask_row, mask_col, mdv=1.0):
    mask_row = np.ascontiguousarray(mask_row, dtype=np.float64)
    mask_col = np.ascontiguousarray(mask_col, dtype=np.float64)
    lib().orc_nullify(bs, len(ptr) - 1, _p(ptr), _p(index), _p(val), _p(mask_row), _p(mask_col), mdv)


def balance(bs, ptr, index, val):
    d = np.zeros((len(ptr) - 1) * bs)
    lib().orc_balance(bs, len(ptr) - 1, _p(ptr), _p(index), _p(val), _p(d))
    return d


def inv_main(bs, ptr, index, val):
    inv = np.zeros((len(ptr) - 1) * bs * bs)
    rc = lib().orc_inv_main(bs, len(ptr) - 1, _p(ptr), _p(index), _p(val), _p(inv))
    if rc:
        raise RuntimeError("SparseMatrix::invMain: non-regular main diagonal block.")
    return inv


def solve(bs, ptr, index, val, b, tolerance=1e-8, iter_max=100000, sweeps=1, plain=False,
          verbose=False, method=None, preconditioner="JACOBI"):
    """PASO Solver() on a copy of val; returns (x, diagnostics dict, balanced val, d).
    method: "PCG" (PASO's, default), "PCG_PLAIN" (plain=True), "BICGSTAB" or "MINRES";
    preconditioner: "JACOBI", "GAUSS_SEIDEL", "ILU0" or "NONE"."""
    if method is not None:
        plain = {"PCG": 0, "PCG_PLAIN": 1, "BICGSTAB": 2, "MINRES": 3}[method]
    val = np.array(val, dtype=np.float64, copy=True)
    b = np.ascontiguousarray(b, dtype=np.float64)
    N = len(ptr) - 1
    x = np.zeros(N * bs)
    d = np.zeros(N * bs)
    opt = Options(tolerance, iter_max, sweeps, int(plain), int(verbose), PRECONDITIONERS[preconditioner])
    dg = Diagnostics()
    rc = lib().orc_solve(bs, N, _p(ptr), _p(index), _p(val), _p(x), _p(b), ctypes.byref(opt),
                         ctypes.byref(dg), _p(d))
    diag = {f: getattr(dg, f) for f, _ in Diagnostics._fields_}
    diag["rc"] = rc
    return x, diag, val, d


PLANTS = {"none": 0, "pcg_no_smoothing": 1, "pcg_gamma_sign": 2, "bicgstab_beta": 3,
          "bicgstab_omega_sign": 4, "minres_w_a3": 5}


def traced_solve(bs, ptr, index, val, b, method="PCG", cap=200, plant="none", **kw):
    """solve() with the first inner Krylov solve traced.  Returns
    (x, diag, balanced val, d, vecs, scal): vecs[k] holds the four vectors
    recorded after iteration k+1 and scal[k] its four scalars --
    PCG: (x, rs, x2, r), (alpha, tau, gamma_1, ||rs||); BiCGStab: (x, r, s, t),
    (rho, alpha, omega, ||r||); MINRES: (x, r, w, z), (delta, gamma, c,
    rnorm_prec).  plant switches on a known mistake (PLANTS) for this call."""
    n = (len(ptr) - 1) * bs
    buf = np.zeros(cap * (4 * n + 4))
    L = lib()
    L.orc_set_plant(PLANTS[plant])
    L.orc_trace_begin(_p(buf), int(cap), n)
    try:
        x, diag, v, d = solve(bs, ptr, index, val, b, method=method, **kw)
        k = L.orc_trace_count()
    finally:
        L.orc_set_plant(0)
        L.orc_trace_begin(None, 0, 0)
    rec = buf[: k * (4 * n + 4)].reshape(k, 4 * n + 4)
    return x, diag, v, d, rec[:, : 4 * n].reshape(k, 4, n), rec[:, 4 * n:]


def coloring(ptr, index):
    """Pattern::borrowColoringPointer: (colour per row, number of colours)."""
    ptr = np.ascontiguousarray(ptr, dtype=np.int32)
    index = np.ascontiguousarray(index, dtype=np.int32)
    c = np.zeros(len(ptr) - 1, dtype=np.int32)
    nc = lib().orc_coloring(len(ptr) - 1, _p(ptr), _p(index), _p(c))
    return c, nc


def gs_sweep(bs, ptr, index, val, inv, colours, ncolors, b):
    """One coloured symmetric Gauss-Seidel sweep applied to b."""
    x = np.array(b, dtype=np.float64, copy=True)
    lib().orc_gs_sweep(bs, len(ptr) - 1, _p(ptr), _p(index), _p(np.ascontiguousarray(val)),
                       _p(np.ascontiguousarray(inv)), _p(np.ascontiguousarray(colours, dtype=np.int32)),
                       int(ncolors), _p(x))
    return x


def gs_apply(bs, ptr, index, val, inv, colours, ncolors, sweeps, b):
    """Preconditioner_Smoother_solve with `sweeps` coloured Gauss-Seidel sweeps."""
    b = np.ascontiguousarray(b, dtype=np.float64)
    x = np.zeros_like(b)
    lib().orc_gs_apply(bs, len(ptr) - 1, _p(ptr), _p(index), _p(np.ascontiguousarray(val)),
                       _p(np.ascontiguousarray(inv)), _p(np.ascontiguousarray(colours, dtype=np.int32)),
                       int(ncolors), int(sweeps), _p(x), _p(b))
    return x


def ilu0_factor(bs, ptr, index, val, colours, ncolors):
    F = np.zeros(len(index) * bs * bs)
    rc = lib().orc_ilu0_factor(bs, len(ptr) - 1, _p(ptr), _p(index), _p(np.ascontiguousarray(val)),
                               _p(np.ascontiguousarray(colours, dtype=np.int32)), int(ncolors), _p(F))
    return F, rc


def ilu0_solve(bs, ptr, index, F, colours, ncolors, b):
    b = np.ascontiguousarray(b, dtype=np.float64)
    x = np.zeros_like(b)
    lib().orc_ilu0_solve(bs, len(ptr) - 1, _p(ptr), _p(index), _p(F),
                         _p(np.ascontiguousarray(colours, dtype=np.int32)), int(ncolors), _p(x), _p(b))
    return x


def time_pcg_iterations(bs, ptr, index, val, b, iters, plain=False):
    val = np.array(val, dtype=np.float64, copy=True)
    b = np.ascontiguousarray(b, dtype=np.float64)
    d = np.zeros((len(ptr) - 1) * bs)
    return lib().orc_time_pcg_iterations(bs, len(ptr) - 1, _p(ptr), _p(index), _p(val), _p(b),
                                         int(iters), int(plain), _p(d))


def cds_add(dim, bs, symmetric, ndiag, values, nodes, array):
    nodes = np.ascontiguousarray(nodes, dtype=np.int32)
    array = np.ascontiguousarray(array, dtype=np.float64)
    lib().orc_cds_add(dim, bs, int(symmetric), ndiag, _p(values), len(nodes), _p(nodes), _p(array))


def cds_spmv(bs, symmetric, nnodes, offsets, values, x):
    offsets = np.ascontiguousarray(offsets, dtype=np.int32)
    x = np.ascontiguousarray(x, dtype=np.float64)
    y = np.zeros(nnodes * bs)
    lib().orc_cds_spmv(bs, int(symmetric), nnodes, len(offsets), _p(offsets), _p(values), _p(x), _p(y))
    return y


def csr_to_dense(bs, ptr, index, val):
    """Dense scalar matrix from PASO block CSR (for small test cases)."""
    N = len(ptr) - 1
    M = np.zeros((N * bs, N * bs))
    for i in range(N):
        for k in range(ptr[i], ptr[i + 1]):
            j = index[k]
            blk = val[k * bs * bs:(k + 1) * bs * bs].reshape(bs, bs, order="F")
            M[i * bs:(i + 1) * bs, j * bs:(j + 1) * bs] += blk
    return M
