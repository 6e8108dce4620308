"""Binary grid I/O of ripley domains (ripley/src/ripleycpp.cpp:63-110 and
:578-590 for the Python signatures; Rectangle/Brick::readBinaryGridImpl and
writeBinaryGridImpl for the file semantics), through the C ABI
(rg_read_binary_grid / rg_write_binary_grid, host code).

Function spaces are given as in escript, e.g. ``ReducedFunction(dom)``; the
data come back as a tensor on the domain's device with one row per sample:
``(samples,) + shape`` when a sample holds one data point (Solution,
ContinuousFunction, ReducedFunction) and ``(samples, 2**dim) + shape`` for
Function.  Data points are stored as escript stores them (the components of
one point contiguous)."""
import ctypes
import gzip
import os
import shutil
import tempfile

import numpy as np
import torch

from . import _lib
from ._lib import check, load

BYTEORDER_LITTLE_ENDIAN = _lib.RG_BYTEORDER_LITTLE_ENDIAN
BYTEORDER_BIG_ENDIAN = _lib.RG_BYTEORDER_BIG_ENDIAN
BYTEORDER_NATIVE = _lib.RG_BYTEORDER_NATIVE
DATATYPE_INT32 = _lib.RG_DATATYPE_INT32
DATATYPE_FLOAT32 = _lib.RG_DATATYPE_FLOAT32
DATATYPE_FLOAT64 = _lib.RG_DATATYPE_FLOAT64


class FunctionSpace:
    """The escript function spaces of a Rectangle/Brick that hold grid data."""

    NODES = ("Solution", "ContinuousFunction")
    ELEMENTS = ("Function", "ReducedFunction")

    def __init__(self, domain, kind):
        if kind not in self.NODES + self.ELEMENTS:
            raise ValueError("unsupported function space %r" % (kind,))
        self._domain = domain
        self.kind = kind

    def getDomain(self):
        return self._domain

    def getDim(self):
        return self._domain.getDim()

    def samplesPerAxis(self):
        d = self._domain
        return d.getNumNodesPerAxis() if self.kind in self.NODES else d.getNumElementsPerAxis()

    def getNumSamples(self):
        return int(np.prod(self.samplesPerAxis()))

    def getNumDataPointsPerSample(self):
        return self._domain.getNumQuadraturePoints() if self.kind == "Function" else 1

    def getX(self):
        d = self._domain
        if self.kind in self.NODES:
            return d.getX()
        return d.getQuadraturePoints() if self.kind == "Function" else d.getElementCenters()

    def __repr__(self):
        return "%s(%r)" % (self.kind, self._domain)


def Solution(domain):
    return FunctionSpace(domain, "Solution")


def ContinuousFunction(domain):
    return FunctionSpace(domain, "ContinuousFunction")


def Function(domain):
    return FunctionSpace(domain, "Function")


def ReducedFunction(domain):
    return FunctionSpace(domain, "ReducedFunction")


def _i64(v, n, name):
    v = list(v)
    if len(v) != n:
        raise ValueError("argument '%s' has wrong length" % name)
    return np.ascontiguousarray(v, dtype=np.int64)


def _i32(v, n, name):
    v = list(v)
    if len(v) != n:
        raise ValueError("argument '%s' has wrong length" % name)
    return np.ascontiguousarray(v, dtype=np.int32)


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def readBinaryGrid(filename, functionspace, shape=(), fill=0.0, byteOrder=BYTEORDER_NATIVE,
                   dataType=DATATYPE_FLOAT64, first=None, numValues=None, multiplier=None, reverse=None):
    """ripley.readBinaryGrid (ripleycpp.cpp:63-86): a Data object of the given
    shape filled with `fill`, then overwritten from the file."""
    fs = functionspace
    dim = fs.getDim()
    grid = np.ascontiguousarray(fs.samplesPerAxis(), dtype=np.int64)
    first = _i64(first if first is not None else [0] * dim, dim, "first")
    numValues = _i64(numValues if numValues is not None else grid, dim, "numValues")
    multiplier = _i32(multiplier if multiplier is not None else [1] * dim, dim, "multiplier")
    reverse = _i32(reverse if reverse is not None else [0] * dim, dim, "reverse")
    shape = tuple(int(s) for s in shape)
    ncomp = int(np.prod(shape)) if shape else 1
    dpp = fs.getNumDataPointsPerSample()
    out = np.full(fs.getNumSamples() * dpp * ncomp, float(fill), dtype=np.float64)
    check(load().rg_read_binary_grid(os.fsencode(filename), dim, _p(grid), ncomp, dpp, _p(first),
                                     _p(numValues), _p(multiplier), _p(reverse), int(byteOrder), int(dataType),
                                     _p(out)))
    lead = (fs.getNumSamples(),) if dpp == 1 else (fs.getNumSamples(), dpp)
    return torch.as_tensor(out.reshape(lead + shape), device=fs.getDomain().device)


def readBinaryGridFromZipped(filename, functionspace, shape=(), fill=0.0, byteOrder=BYTEORDER_NATIVE,
                             dataType=DATATYPE_FLOAT64, first=None, numValues=None, multiplier=None,
                             reverse=None):
    """ripley's _readBinaryGridFromZipped (Brick.cpp:624-646): the gzip
    stream is expanded to a temporary file and read as above."""
    with tempfile.NamedTemporaryFile(suffix=".grid", delete=False) as tmp:
        with gzip.open(filename, "rb") as src:
            shutil.copyfileobj(src, tmp)
        name = tmp.name
    try:
        return readBinaryGrid(name, functionspace, shape, fill, byteOrder, dataType, first, numValues,
                              multiplier, reverse)
    finally:
        os.unlink(name)


def writeBinaryGrid(domain, data, filename, byteOrder=BYTEORDER_NATIVE, dataType=DATATYPE_FLOAT64):
    """RipleyDomain::writeBinaryGrid: scalar data with one value per node or
    per element (samples told apart by their count)."""
    v = data.detach().cpu().numpy() if isinstance(data, torch.Tensor) else np.asarray(data)
    v = np.ascontiguousarray(v, dtype=np.float64).reshape(-1)
    if v.size == domain.getNumNodes():
        grid = domain.getNumNodesPerAxis()
    elif v.size == domain.getNumElements():
        grid = domain.getNumElementsPerAxis()
    else:
        raise ValueError("writeBinaryGrid(): only scalar, single-value data on nodes or elements is supported")
    grid = np.ascontiguousarray(grid, dtype=np.int64)
    check(load().rg_write_binary_grid(os.fsencode(filename), domain.getDim(), _p(grid), _p(v), int(byteOrder),
                                      int(dataType)))
