"""ripley's binary grid reader/writer (rg_read_binary_grid / rg_write_binary_grid,
host code in the C ABI) restating the reference's own tests
(ripley/test/python/run_readWriteOnRipley.py at one rank): writeBinaryGrid of
unique values for every byte order and data type, readBinaryGrid -> write ->
numpy round trips with fill, multipliers and z reversal, and the reference's
fixture files ripley/test/python/ref_data/*.grid(.gz) (copied to
tests/golden/grid/).  No GPU is needed: the domains fall back to the CPU."""
import os

import numpy as np
import pytest
import torch

import ripley_hip as R

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "grid")

COMBOS = [(R.BYTEORDER_LITTLE_ENDIAN, R.DATATYPE_FLOAT32, "<f4"), (R.BYTEORDER_LITTLE_ENDIAN, R.DATATYPE_FLOAT64, "<f8"),
          (R.BYTEORDER_LITTLE_ENDIAN, R.DATATYPE_INT32, "<i4"), (R.BYTEORDER_BIG_ENDIAN, R.DATATYPE_FLOAT32, ">f4"),
          (R.BYTEORDER_BIG_ENDIAN, R.DATATYPE_FLOAT64, ">f8"), (R.BYTEORDER_BIG_ENDIAN, R.DATATYPE_INT32, ">i4")]


def adjust(NE, kind):
    return [n + 1 for n in NE] if kind in ("ContinuousFunction", "Solution") else list(NE)


def domain(NE):
    return R.Rectangle(*NE) if len(NE) == 2 else R.Brick(*NE)


@pytest.mark.parametrize("byteorder,datatype,dtype", COMBOS)
@pytest.mark.parametrize("NE", [[9, 10], [9, 9, 10]])
def test_write_grid(tmp_path, byteorder, datatype, dtype, NE):
    # WriteBinaryGridTestBase: unique values (the sample index; + 0.05 for
    # INT32, which the writer truncates) on ReducedFunction, ContinuousFunction
    # and Solution
    dom = domain(NE)
    for kind in ("ReducedFunction", "ContinuousFunction", "Solution"):
        shape = adjust(NE, kind)
        n = int(np.prod(shape))
        data = torch.arange(n, dtype=torch.float64) + (0.05 if datatype == R.DATATYPE_INT32 else 0.0)
        fn = str(tmp_path / ("w%s" % kind))
        dom.writeBinaryGrid(data, fn, byteorder, datatype)
        result = np.fromfile(fn, dtype=dtype).reshape(tuple(reversed(shape)))
        ref = np.arange(n, dtype=dtype).reshape(tuple(reversed(shape)))
        assert np.abs(result.astype(np.float64) - ref).max() <= 1e-9, kind


READ_CASES = [dict(byteorder=b, datatype=t, dtype=d) for b, t, d in COMBOS] + [
    dict(reverse=[0, 0, 1]), dict(multiplier=[2, 1, 1]), dict(multiplier=[1, 2, 1]), dict(multiplier=[1, 1, 2]),
    dict(multiplier=[2, 3, 4])]


def _read_case(tmp_path, dim, byteorder=R.BYTEORDER_NATIVE, datatype=R.DATATYPE_FLOAT64, dtype="f8",
               multiplier=(1, 1, 1), reverse=(0, 0, 0)):
    """ReadBinaryGridTestBase.test_readGrid2D/3D at one rank."""
    NX, NZ, fill = 10, 8, -42.57
    multiplier, reverse = list(multiplier)[:dim], list(reverse)[:dim]
    if dim == 2:
        NE = [NX - 1 if multiplier[0] == 1 else NX * multiplier[0] - 1, NZ * multiplier[1]]
    else:
        NE = [NX - 1 if multiplier[0] == 1 else NX * multiplier[0] - 1, NX * multiplier[1], NZ * multiplier[2]]
    dom = domain(NE)
    for kind, fs in (("ReducedFunction", R.ReducedFunction(dom)), ("ContinuousFunction", R.ContinuousFunction(dom))):
        Ndata = [NX - 1, NZ] if dim == 2 else [NX - 1, NX, NZ]
        if kind == "ContinuousFunction":
            Ndata[1:] = [n - 1 for n in Ndata[1:]]
        nv = adjust(Ndata, kind)
        ref = np.array(range(int(np.prod(nv))), dtype=dtype).reshape(tuple(reversed(nv)))
        fn = str(tmp_path / ("r%dd%s" % (dim, kind)))
        ref.tofile(fn)
        data = R.readBinaryGrid(fn, fs, shape=(), fill=fill, byteOrder=byteorder, dataType=datatype,
                                first=[0] * dim, numValues=nv, multiplier=multiplier, reverse=reverse)
        dom.writeBinaryGrid(data, fn, byteorder, datatype)
        realNE = adjust(NE, kind)
        result = np.fromfile(fn, dtype=dtype).reshape(tuple(reversed(realNE)))
        for d in range(dim):
            if reverse[d]:
                result = np.flip(result, axis=dim - 1 - d)
        for i in range(dim):
            ref = np.repeat(ref, multiplier[i], axis=dim - 1 - i)
        fillv = np.array(fill, dtype=ref.dtype)
        for d in range(dim):
            excess = realNE[d] - ref.shape[dim - 1 - d]
            if excess > 0:
                shape = list(ref.shape)
                shape[dim - 1 - d] = excess
                extra = fillv * np.ones(shape)
                ref = np.append(extra, ref, axis=dim - 1 - d) if reverse[d] else np.append(ref, extra, axis=dim - 1 - d)
        assert np.abs(ref.astype(np.float64) - result.astype(np.float64)).max() <= 1e-9, (kind, dim)


@pytest.mark.parametrize("case", READ_CASES)
@pytest.mark.parametrize("dim", [2, 3])
def test_read_grid(tmp_path, case, dim):
    _read_case(tmp_path, dim, **case)


@pytest.mark.parametrize("dim", [2, 3])
def test_read_reverse_x_is_refused(tmp_path, dim):
    # reversing is implemented for the z axis of a Brick only (Rectangle.cpp:
    # 598-599, Brick.cpp:697-698)
    with pytest.raises(R.RipleyHipError, match="revers"):
        _read_case(tmp_path, dim, reverse=[1, 0, 0])


@pytest.mark.parametrize("name,NE,kind", [("RectRedF1", [9, 10], "ReducedFunction"),
                                          ("RectConF1", [9, 10], "ContinuousFunction"),
                                          ("BrickRedF1", [9, 9, 10], "ReducedFunction"),
                                          ("BrickConF1", [9, 9, 10], "ContinuousFunction")])
def test_reference_fixtures(name, NE, kind):
    # Test_readBinaryGridZippedRipley: the zipped and plain fixtures read to
    # the same data, which is the file's content in sample order
    dom = domain(NE)
    fs = getattr(R, kind)(dom)
    nv = adjust(NE, kind)
    fn = os.path.join(GOLDEN, name + ".grid")
    plain = R.readBinaryGrid(fn, fs, (), 50000, R.BYTEORDER_NATIVE, R.DATATYPE_FLOAT64, [0] * len(NE), nv,
                             [1] * len(NE), [0] * len(NE))
    zipped = R.readBinaryGridFromZipped(fn + ".gz", fs, (), 50000, R.BYTEORDER_NATIVE, R.DATATYPE_FLOAT64,
                                        [0] * len(NE), nv, [1] * len(NE), [0] * len(NE))
    assert torch.equal(plain, zipped)
    assert np.array_equal(plain.cpu().numpy(), np.fromfile(fn, dtype="<f8"))


def test_function_space_points_components_nan_and_offsets(tmp_path):
    dom = R.Rectangle(4, 3)
    # Function: every quadrature point of an element gets the element's value
    vals = np.arange(12, dtype="<f8")
    fn = str(tmp_path / "f")
    vals.tofile(fn)
    d = R.readBinaryGrid(fn, R.Function(dom), numValues=[4, 3])
    assert d.shape == (12, 4) and np.array_equal(d.cpu().numpy(), np.repeat(vals[:, None], 4, axis=1))
    # two components per value, interleaved in the file
    v2 = np.arange(24, dtype="<f8")
    v2.tofile(fn)
    d = R.readBinaryGrid(fn, R.ReducedFunction(dom), shape=(2,), numValues=[4, 3])
    assert np.array_equal(d.cpu().numpy(), v2.reshape(12, 2))
    # NaN values keep the fill
    v3 = np.arange(12, dtype="<f8")
    v3[5] = np.nan
    v3.tofile(fn)
    d = R.readBinaryGrid(fn, R.ReducedFunction(dom), fill=-1.0, numValues=[4, 3]).cpu().numpy()
    assert d[5] == -1.0 and d[4] == 4.0
    # an offset block: 2 x 2 values placed at first = (1, 1) of the node grid
    np.array([1.0, 2.0, 3.0, 4.0], dtype="<f8").tofile(fn)
    d = R.readBinaryGrid(fn, R.ContinuousFunction(dom), fill=0.0, first=[1, 1], numValues=[2, 2]).cpu().numpy()
    g = d.reshape(4, 5)
    assert np.array_equal(g[1:3, 1:3], [[1, 2], [3, 4]]) and g.sum() == 10.0
    # a negative first skips values of the file
    np.arange(20, dtype="<f8").tofile(fn)
    d = R.readBinaryGrid(fn, R.ContinuousFunction(dom), fill=0.0, first=[-1, 0], numValues=[5, 4]).cpu().numpy()
    assert np.array_equal(d.reshape(4, 5)[:, :4], np.arange(20).reshape(4, 5)[:, 1:])


def test_errors(tmp_path):
    dom = R.Rectangle(4, 3)
    fn = str(tmp_path / "small")
    np.zeros(5).tofile(fn)
    with pytest.raises(R.RipleyHipError, match="not enough data"):
        R.readBinaryGrid(fn, R.ReducedFunction(dom), numValues=[4, 3])
    with pytest.raises(R.RipleyHipError, match="cannot open"):
        R.readBinaryGrid(str(tmp_path / "missing"), R.ReducedFunction(dom), numValues=[4, 3])
    with pytest.raises(R.RipleyHipError, match="multipliers"):
        R.readBinaryGrid(fn, R.ReducedFunction(dom), numValues=[1, 1], multiplier=[0, 1])
    with pytest.raises(ValueError):
        dom.writeBinaryGrid(np.zeros(7), fn, R.BYTEORDER_NATIVE, R.DATATYPE_FLOAT64)
