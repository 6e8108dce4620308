"""Batched solves (no counterpart in the reference, whose PASO solves one
right-hand side of one system per call, paso/src/solve.cpp:41-145):

* ``solve_many(op, B, options)`` -- many right-hand sides (columns of a
  row-major N x K device block, K <= 64) on one operator;
* ``BatchSolver(ops)`` -- many independent scalar systems of uneven sizes
  solved together, every kernel launch spanning all of them.

Both run the textbook Jacobi-PCG on the unbalanced systems from x = 0 with
FP64 (SolverOptions.PCG_PLAIN) or FP32 (SolverOptions.PCG_MIXED) matrix
values, FP64 vectors, and stop per column / system on the FP64 residual
||b - A x||_2 <= tol ||b||_2.  Tile shapes are autotuned on first use.
"""
import ctypes

import numpy as np
import torch

from . import _lib
from ._lib import check, load
from .core import SolverOptions, _ptr


def _options(options):
    if options is None:
        options = SolverOptions()
        options.setSolverMethod(SolverOptions.PCG_MIXED)
    m = options.getSolverMethod()
    if m not in (SolverOptions.PCG_PLAIN, SolverOptions.PCG_MIXED):
        raise ValueError("batched solves run PCG_PLAIN or PCG_MIXED")
    return options, options._rg_options()


def _diagnostics(options, dg, n):
    out = []
    for j in range(n):
        d = dg[j]
        out.append({"num_iter": d.num_iter, "converged": bool(d.converged), "status": d.status,
                    "residual_norm": d.residual_norm, "time": d.time})
    options._updateDiagnostics("num_iter", max(d["num_iter"] for d in out))
    options._updateDiagnostics("converged", all(d["converged"] for d in out))
    options._updateDiagnostics("time", out[0]["time"] if out else 0.0)
    return out


def solve_many(op, B, options=None):
    """X (N x K) with A X[:, k] = B[:, k]; B a float64 GPU tensor of shape
    (N, K) (row-major).  Returns (X, per-column diagnostics)."""
    options, o = _options(options)
    if not (isinstance(B, torch.Tensor) and B.is_cuda and B.dtype == torch.float64 and B.dim() == 2
            and B.shape[0] == op.size):
        raise ValueError("B must be a float64 GPU tensor of shape (%d, K)" % op.size)
    B = B.contiguous()
    K = B.shape[1]
    X = torch.empty_like(B)
    dg = (_lib.RgDiagnostics * K)()
    op._bind()
    rc = load().rg_solve_multi(op.handle, _ptr(X), _ptr(B), K, ctypes.byref(o), dg)
    diags = _diagnostics(options, dg, K)
    check(rc)
    return X, diags


def matrix_vector_many(op, X):
    """Y = A X for a float64 GPU block X of shape (N, K)."""
    X = X.contiguous()
    Y = torch.empty_like(X)
    op._bind()
    check(load().rg_matrix_vector_multi(op.handle, _ptr(X), _ptr(Y), X.shape[1]))
    return Y


# (rows, columns) per thread of the row-chunked SpMM tiles 0-3, the column
# entry dividing the column count; 4-6 are the z-marching LDS tiles, which
# need exactly 64 columns (Brick matrices only): (16, 64) and (8, 64) give a
# thread 16 or 8 x-nodes of one column, (1, 64) one x-node of 8 columns
SPMM_SHAPES = ((1, 1), (1, 2), (1, 4), (2, 2), (16, 64), (8, 64), (1, 64))
ZMARCH_SHAPES = (4, 5, 6)


def autotune_many(op, K, fp32=True, reps=5):
    """(index of the chosen SpMM shape in SPMM_SHAPES, seconds per SpMM of
    each shape; -1 for shapes that do not divide K)."""
    r = ctypes.c_int()
    t = (ctypes.c_double * len(SPMM_SHAPES))()
    op._bind()
    check(load().rg_multi_autotune(op.handle, int(K), int(bool(fp32)), int(reps), ctypes.byref(r), t))
    return r.value, list(t)


def set_shape_many(op, shape):
    """Fix the SpMM shape (index into SPMM_SHAPES) or autotune again (-1)."""
    check(load().rg_multi_set_shape(op.handle, int(shape)))


class BatchSolver:
    """Independent scalar systems (SystemMatrix objects of one dimension,
    full storage) solved together.  Vectors are concatenated at
    ``offsets``; ``split``/``join`` convert."""

    TILES = (1, 2, 4, 8)

    def __init__(self, ops):
        self.ops = list(ops)
        if not self.ops:
            raise ValueError("empty batch")
        ctx = self.ops[0].ctx
        arr = (ctypes.c_void_p * len(self.ops))(*[o.handle.value for o in self.ops])
        h = ctypes.c_void_p()
        ctx.bind_stream()
        check(load().rg_batch_create(ctx.handle, len(self.ops), arr, ctypes.byref(h)))
        self.handle = h
        self.ctx = ctx
        off = (ctypes.c_int64 * (len(self.ops) + 1))()
        check(load().rg_batch_offsets(h, off))
        self.offsets = list(off)
        self.device = self.ops[0].device

    def __del__(self):
        try:
            if self.handle:
                load().rg_batch_destroy(self.handle)
                self.handle = None
        except Exception:
            pass

    @property
    def size(self):
        return self.offsets[-1]

    def join(self, vectors):
        out = torch.zeros(self.size, dtype=torch.float64, device=self.device)
        for j, v in enumerate(vectors):
            n = self.ops[j].size
            out[self.offsets[j]:self.offsets[j] + n] = torch.as_tensor(v, device=self.device).reshape(-1)
        return out

    def split(self, v):
        return [v[self.offsets[j]:self.offsets[j] + self.ops[j].size] for j in range(len(self.ops))]

    def MatrixVector(self, x):
        y = torch.zeros(self.size, dtype=torch.float64, device=self.device)
        self.ctx.bind_stream()
        check(load().rg_batch_vector(self.handle, _ptr(x), _ptr(y)))
        return y

    def set_tile(self, mult):
        check(load().rg_batch_set_tile(self.handle, int(mult)))

    def autotune(self, fp32=True, reps=5):
        """(chosen tile multiplier, seconds per batched SpMV of 1, 2, 4, 8)."""
        m = ctypes.c_int()
        t = (ctypes.c_double * len(self.TILES))()
        self.ctx.bind_stream()
        check(load().rg_batch_autotune(self.handle, int(bool(fp32)), int(reps), ctypes.byref(m), t))
        return m.value, list(t)

    def solve(self, b, options=None):
        """x with A_j x_j = b_j for every system (concatenated vectors)."""
        options, o = _options(options)
        b = b.contiguous()
        if b.numel() != self.size:
            raise ValueError("rhs must have %d entries (concatenated at offsets)" % self.size)
        x = torch.zeros_like(b)
        n = len(self.ops)
        dg = (_lib.RgDiagnostics * n)()
        self.ctx.bind_stream()
        rc = load().rg_batch_solve(self.handle, _ptr(x), _ptr(b), ctypes.byref(o), dg)
        diags = _diagnostics(options, dg, n)
        check(rc)
        return x, diags
