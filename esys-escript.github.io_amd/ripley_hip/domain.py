"""Rectangle / Brick domains at one rank (ripley/src/Rectangle.cpp:82-237,
ripley/src/Brick.cpp:89-275; python constructors ripley/src/ripleycpp.cpp:
521-538).  Function spaces are represented by array shapes:

* Solution / ContinuousFunction: one value per node, (N, *shape);
* Function: one value per quadrature point, (E, Q, *shape), Q = 2**dim,
  points x fastest at (1 -+ 1/sqrt3)/2 of each element;
* ReducedFunction: one value per element, (E, *shape).

Nodes and elements are numbered lexicographically with x fastest.
"""
import numpy as np
import torch

from .core import SystemMatrix

_G0 = 0.5 - 0.5 / np.sqrt(3.0)
_G1 = 0.5 + 0.5 / np.sqrt(3.0)


class _Grid:
    def __init__(self, ne, length, device=0):
        self._ne = tuple(int(n) for n in ne)
        self._len = tuple(float(x) for x in length)
        if any(n < 1 for n in self._ne):
            raise ValueError("need at least one element per axis")
        self._device = torch.device("cuda", device) if torch.cuda.is_available() else torch.device("cpu")
        self._dev_index = device

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

    # --- operator factory (RipleyDomain::newSystemMatrix) ---------------
    def newOperator(self, block_size=1, symmetric=False):
        return SystemMatrix(self, block_size=block_size, symmetric=symmetric, device=self._dev_index)

    def __repr__(self):
        return "%s(%s)" % (type(self).__name__, ", ".join(str(n) for n in self._ne))


class Rectangle(_Grid):
    """ripley.Rectangle(n0, n1, l0=1, l1=1) at one rank."""

    def __init__(self, n0, n1, l0=1.0, l1=1.0, device=0):
        super().__init__((n0, n1), (l0, l1), device)


class Brick(_Grid):
    """ripley.Brick(n0, n1, n2, l0=1, l1=1, l2=1) at one rank."""

    def __init__(self, n0, n1, n2, l0=1.0, l1=1.0, l2=1.0, device=0):
        super().__init__((n0, n1, n2), (l0, l1, l2), device)
