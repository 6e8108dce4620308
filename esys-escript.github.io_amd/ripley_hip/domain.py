"""Rectangle / Brick domains at one rank (ripley/src/Rectangle.cpp:82-237,
ripley/src/Brick.cpp:89-275; python constructors ripley/src/ripleycpp.cpp:
521-538).  Function spaces are represented by array shapes:

* Solution / ContinuousFunction: one value per node, (N, *shape);
* Function: one value per quadrature point, (E, Q, *shape), Q = 2**dim,
  points x fastest at (1 -+ 1/sqrt3)/2 of each element;
* ReducedFunction: one value per element, (E, *shape);
* FunctionOnBoundary: one value per face quadrature point, (F, 2**(dim-1),
  *shape), face elements in the reference's order (faces x0=0, x0=L, x1=0,
  x1=L[, x2=0, x2=L], within a face by the tangential element indices, lower
  axis fastest; Rectangle.cpp/Brick.cpp m_faceOffset);
* ReducedFunctionOnBoundary: one value per face element, (F, *shape);
* DiracDeltaFunctions: one value per Dirac point (nearest node of each point
  given at construction, RipleyDomain.cpp:1820-1850).

Nodes and elements are numbered lexicographically with x fastest.
"""
import numpy as np
import torch

from .core import SystemMatrix

_G0 = 0.5 - 0.5 / np.sqrt(3.0)
_G1 = 0.5 + 0.5 / np.sqrt(3.0)


class _Grid:
    def __init__(self, ne, length, device=0, diracPoints=()):
        self._ne = tuple(int(n) for n in ne)
        self._len = tuple(float(x) for x in length)
        if any(n < 1 for n in self._ne):
            raise ValueError("need at least one element per axis")
        self._device = torch.device("cuda", device) if torch.cuda.is_available() else torch.device("cpu")
        self._dev_index = device
        self._dirac = [self._nearest_node(p) for p in diracPoints]

    def _nearest_node(self, point):
        h = self.getSpacing()
        nn = self.getNumNodesPerAxis()
        idx, stride = 0, 1
        for d in range(self.getDim()):
            k = int(np.clip(np.floor(point[d] / h[d] + 0.5), 0, nn[d] - 1))
            idx += k * stride
            stride *= nn[d]
        return idx

    def getDiracNodes(self):
        """Node index of each Dirac point (DiracDeltaFunctions samples)."""
        return list(self._dirac)

    # --- boundary (FunctionOnBoundary) ------------------------------------
    def _faces(self):
        """[(face, normal axis, side, tangential axes, element count)]."""
        out = []
        dim = self.getDim()
        for f in range(2 * dim):
            ax, side = f // 2, f % 2
            t = [a for a in range(dim) if a != ax]
            out.append((f, ax, side, t, int(np.prod([self._ne[a] for a in t]))))
        return out

    def getNumFaceElements(self):
        return sum(c for *_, c in self._faces())

    def getNumBoundaryQuadraturePoints(self):
        return 1 << (self.getDim() - 1)

    def getBoundaryQuadraturePoints(self, reduced=False):
        """Coordinates (F, Qb, dim) of the FunctionOnBoundary points (F, dim
        for reduced: the face element centres)."""
        dim = self.getDim()
        h = self.getSpacing()
        pts = []
        for f, ax, side, t, cnt in self._faces():
            ne_t = [self._ne[a] for a in t]
            j = np.arange(cnt)
            jt = [j % ne_t[0]] + ([j // ne_t[0]] if dim == 3 else [])
            nq = 1 if reduced else (1 << (dim - 1))
            P = np.zeros((cnt, nq, dim))
            P[:, :, ax] = self._len[ax] if side else 0.0
            for q in range(nq):
                for i, a in enumerate(t):
                    xi = 0.5 if reduced else (_G1 if (q >> i) & 1 else _G0)
                    P[:, q, a] = (jt[i] + xi) * h[a]
            pts.append(P)
        P = np.concatenate(pts, axis=0)
        P = P[:, 0, :] if reduced else P
        return torch.as_tensor(P, device=self._device)

    # --- sizes (Rectangle.h:352-424, Brick.h:352-416) -------------------
    def getDim(self):
        return len(self._ne)

    def getNumElementsPerAxis(self):
        return self._ne

    def getNumNodesPerAxis(self):
        return tuple(n + 1 for n in self._ne)

    def getLength(self):
        return self._len

    def getSpacing(self):
        return tuple(l / n for l, n in zip(self._len, self._ne))

    def getNumElements(self):
        return int(np.prod(self._ne))

    def getNumNodes(self):
        return int(np.prod(self.getNumNodesPerAxis()))

    getNumDOF = getNumNodes

    def getNumQuadraturePoints(self):
        return 1 << self.getDim()

    @property
    def device(self):
        return self._device

    # --- coordinates ----------------------------------------------------
    def getX(self):
        """Node coordinates (N, dim), x fastest."""
        axes = [torch.linspace(0.0, l, n + 1, dtype=torch.float64, device=self._device)
                for l, n in zip(self._len, self._ne)]
        grids = torch.meshgrid(*axes[::-1], indexing="ij")
        return torch.stack([g.reshape(-1) for g in grids[::-1]], dim=1)

    def getElementCenters(self):
        """Element centres (E, dim) -- ReducedFunction coordinates."""
        h = self.getSpacing()
        axes = [(torch.arange(n, dtype=torch.float64, device=self._device) + 0.5) * hh
                for hh, n in zip(h, self._ne)]
        grids = torch.meshgrid(*axes[::-1], indexing="ij")
        return torch.stack([g.reshape(-1) for g in grids[::-1]], dim=1)

    def getQuadraturePoints(self):
        """Quadrature point coordinates (E, Q, dim) -- Function coordinates."""
        dim = self.getDim()
        h = torch.tensor(self.getSpacing(), dtype=torch.float64, device=self._device)
        c = self.getElementCenters() - 0.5 * h
        q = torch.tensor([[(_G1 if (k >> d) & 1 else _G0) for d in range(dim)] for k in range(1 << dim)],
                         dtype=torch.float64, device=self._device)
        return c[:, None, :] + q[None, :, :] * h

    def whereZero(self, axis, value=0.0, tol=1e-10):
        """Node mask (N,) float64: 1 where x[axis] == value."""
        x = self.getX()[:, axis]
        return (torch.abs(x - value) <= tol * max(self._len)).to(torch.float64)

    # --- binary grid I/O (RipleyDomain::writeBinaryGrid) ------------------
    def writeBinaryGrid(self, data, filename, byteOrder, dataType):
        from .gridio import writeBinaryGrid
        writeBinaryGrid(self, data, filename, byteOrder, dataType)

    # --- operator factory (RipleyDomain::newSystemMatrix) ---------------
    def newOperator(self, block_size=1, symmetric=False):
        return SystemMatrix(self, block_size=block_size, symmetric=symmetric, device=self._dev_index)

    def __repr__(self):
        return "%s(%s)" % (type(self).__name__, ", ".join(str(n) for n in self._ne))


class Rectangle(_Grid):
    """ripley.Rectangle(n0, n1, l0=1, l1=1) at one rank."""

    def __init__(self, n0, n1, l0=1.0, l1=1.0, device=0, diracPoints=()):
        super().__init__((n0, n1), (l0, l1), device, diracPoints)


class Brick(_Grid):
    """ripley.Brick(n0, n1, n2, l0=1, l1=1, l2=1) at one rank."""

    def __init__(self, n0, n1, n2, l0=1.0, l1=1.0, l2=1.0, device=0, diracPoints=()):
        super().__init__((n0, n1, n2), (l0, l1, l2), device, diracPoints)
