"""LinearPDE driver of the ripley->PASO path, mirroring
escriptcore/py_src/linearPDEs.py: setValue / getSystem / insertConstraint /
getSolution (:1861-2112), Poisson and LameEquation (:2594-2706).

The PDE is  -(A_ijkl u_k,l + B_ijk u_k)_,j + C_ikl u_k,l + D_ik u_k
             = -X_ij,j + Y_i   with u_i = r_i where q_i > 0.
"""
import numpy as np
import torch

from .core import Coefficient, SolverOptions


class IsotropicCoefficient:
    """A = value * kronecker(dim) for a single PDE, passed compactly (one
    scalar per data point instead of dim*dim)."""

    def __init__(self, value):
        self.value = value


def kronecker_scaled(value):
    return IsotropicCoefficient(value)


class LinearPDE:
    COEFFICIENTS = ("A", "B", "C", "D", "X", "Y", "A_reduced", "B_reduced", "C_reduced", "D_reduced",
                    "X_reduced", "Y_reduced", "d", "y", "d_reduced", "y_reduced", "d_dirac", "y_dirac",
                    "q", "r")

    def __init__(self, domain, numEquations=1, numSolutions=None):
        if numSolutions is None:
            numSolutions = numEquations
        if numSolutions != numEquations:
            raise ValueError("numEquations and numSolutions must be equal on this path")
        self.domain = domain
        self.n = int(numEquations)
        self.dim = domain.getDim()
        self._coef = {}
        self._options = SolverOptions()
        self._symmetric = False
        self._operator = None
        self._rhs = None
        self._solution = None
        self._operator_valid = False
        self._rhs_valid = False

    # --- coefficient shapes (linearPDEs.py: PDECoef definitions) --------
    def _point_shape(self, name):
        d, n = self.dim, self.n
        base = name.split("_")[0]
        if n == 1:
            return {"A": (d, d), "B": (d,), "C": (d,), "D": (), "X": (d,), "Y": (), "d": (), "y": (), "q": (),
                    "r": ()}[base]
        return {"A": (n, d, n, d), "B": (n, d, n), "C": (n, n, d), "D": (n, n), "X": (n, d), "Y": (n,),
                "d": (n, n), "y": (n,), "q": (n,), "r": (n,)}[base]

    def setValue(self, **coefficients):
        for k, v in coefficients.items():
            if k not in self.COEFFICIENTS:
                raise ValueError("unknown coefficient %s" % k)
            self._coef[k] = v
            if k in ("q", "r"):
                self._operator_valid = self._operator_valid and k == "r"
                self._rhs_valid = False
            elif k[0] in "XYy":
                self._rhs_valid = False
            else:
                self._operator_valid = False
                self._rhs_valid = False
        self._solution = None

    def getCoefficient(self, name):
        return self._coef.get(name)

    def getSolverOptions(self):
        self._options._symmetric = self._options._symmetric or self.isSymmetric()
        return self._options

    def setSolverOptions(self, options=None):
        self._options = options if options is not None else SolverOptions()

    def setSymmetryOn(self):
        self._symmetric = True

    def setSymmetryOff(self):
        self._symmetric = False

    def isSymmetric(self):
        return self._symmetric

    def getNumEquations(self):
        return self.n

    # --- system ---------------------------------------------------------
    def _coefficient(self, name):
        v = self._coef.get(name)
        if v is None:
            return None
        iso = isinstance(v, IsotropicCoefficient)
        if iso:
            if self.n != 1 or name[0] != "A":
                raise ValueError("an isotropic coefficient is defined for A of a single PDE")
            v = v.value
        return Coefficient(v, self._point_shape(name), self.domain.getNumElements(),
                           self.domain.getNumQuadraturePoints(), self.domain.device, isotropic=iso)

    def _boundary_coefficient(self, name):
        v = self._coef.get(name)
        if v is None:
            return None
        return Coefficient(v, self._point_shape(name), self.domain.getNumFaceElements(),
                           self.domain.getNumBoundaryQuadraturePoints(), self.domain.device)

    def _node_vector(self, v):
        N, n = self.domain.getNumNodes(), self.n
        if v is None:
            return torch.zeros(N * n, dtype=torch.float64, device=self.domain.device)
        t = v if isinstance(v, torch.Tensor) else torch.as_tensor(np.asarray(v, dtype=np.float64))
        t = t.to(device=self.domain.device, dtype=torch.float64)
        if t.numel() in (1, n) and t.numel() != N * n:
            t = t.reshape(1, -1).expand(N, n)
        return t.reshape(-1).contiguous()

    def _symmetric_coefficient(self, name):
        """Whether coefficient `name` leaves the operator symmetric: A_ij = A_ji
        (A_ijkl = A_klij for systems), D_km = D_mk, d_km = d_mk."""
        v = self._coef.get(name)
        if v is None or isinstance(v, IsotropicCoefficient):
            return True
        t = torch.as_tensor(v, dtype=torch.float64) if not isinstance(v, torch.Tensor) else v
        shape = self._point_shape(name)
        if len(shape) < 2 or t.dim() < len(shape):
            return True
        t = t.reshape(-1, *shape)
        if len(shape) == 2:
            tt = t.transpose(-1, -2)
        else:
            tt = t.permute(0, 3, 4, 1, 2)
        return bool(torch.equal(t, tt))

    def _use_symmetric_storage(self):
        """Symmetric storage (main + upper stencil diagonals) for a PDE declared
        symmetric -- the flag PASO reads through SolverBuddy::isSymmetric
        (ripley/src/RipleyDomain.cpp:925, paso/src/Options.cpp:39) -- whose
        coefficients give a symmetric matrix: no B or C, symmetric A, D, d."""
        if not self.isSymmetric():
            return False
        if any(self._coef.get(k) is not None for k in ("B", "C", "B_reduced", "C_reduced")):
            return False
        return all(self._symmetric_coefficient(k) for k in ("A", "A_reduced", "D", "D_reduced", "d", "d_reduced",
                                                            "d_dirac"))

    def getSystem(self):
        """(operator, rhs): create, assemble both orders, insert constraints
        (linearPDEs.py:1881-2082).  The operator uses symmetric storage when
        the PDE is declared symmetric and its coefficients keep it so."""
        sym = self._use_symmetric_storage()
        if self._operator is not None and self._operator.symmetric != sym:
            self._operator = None
            self._operator_valid = False
        if self._operator is None:
            self._operator = self.domain.newOperator(self.n, symmetric=sym)
        if not self._operator_valid or not self._rhs_valid:
            op = self._operator
            op.resetValues(preserveSolverData=False)
            rhs = torch.zeros(op.size, dtype=torch.float64, device=self.domain.device)
            full = {k: self._coefficient(k) for k in "ABCDXY"}
            red = {k: self._coefficient(k + "_reduced") for k in "ABCDXY"}
            if any(c is not None for c in full.values()):
                op.assemble(full, rhs, reduced=False)
            if any(c is not None for c in red.values()):
                op.assemble(red, rhs, reduced=True)
            # boundary and Dirac terms (RipleyDomain::addToSystem,
            # RipleyDomain.cpp:1024-1035)
            bnd = {k: self._boundary_coefficient(k) for k in ("d", "y")}
            if any(c is not None for c in bnd.values()):
                op.assembleBoundary(bnd, rhs, reduced=False)
            bred = {k: self._boundary_coefficient(k + "_reduced") for k in ("d", "y")}
            if any(c is not None for c in bred.values()):
                op.assembleBoundary(bred, rhs, reduced=True)
            dd, yd = self._coef.get("d_dirac"), self._coef.get("y_dirac")
            if dd is not None or yd is not None:
                nodes = self.domain.getDiracNodes()
                op.addDirac(nodes, d=None if dd is None else np.broadcast_to(
                    np.asarray(dd, dtype=np.float64), (len(nodes), self.n, self.n)),
                            y=None if yd is None else np.broadcast_to(
                    np.asarray(yd, dtype=np.float64), (len(nodes), self.n)), rhs=rhs)
            self._rhs = rhs
            self.insertConstraint()
            self._operator_valid = True
            self._rhs_valid = True
        return self._operator, self._rhs

    def insertConstraint(self, rhs_only=False):
        """linearPDEs.py:2084-2112: rhs -= A u_r; nullifyRowsAndCols(q, q, 1);
        rhs[q] = r."""
        q = self._coef.get("q")
        if q is None:
            return
        qv = self._node_vector(q)
        rv = self._node_vector(self._coef.get("r"))
        mask = qv > 0
        op = self._operator
        if not rhs_only:
            u = torch.where(mask, rv, torch.zeros_like(rv))
            op.MatrixVector(-1.0, u, 1.0, self._rhs)
            op.nullifyRowsAndCols(qv, qv, 1.0)
        self._rhs = torch.where(mask, rv, self._rhs)

    def getOperator(self):
        return self.getSystem()[0]

    def getRightHandSide(self):
        return self.getSystem()[1]

    def getSolution(self):
        """linearPDEs.py:1861-1879."""
        if self._solution is None:
            op, rhs = self.getSystem()
            opts = self.getSolverOptions()
            x = op.solve(rhs, opts)
            self._solution = x if self.n == 1 else x.reshape(-1, self.n)
        return self._solution


class Poisson(LinearPDE):
    """-grad(grad u) = f, u = 0 where q > 0 (linearPDEs.py Poisson)."""

    def __init__(self, domain):
        super().__init__(domain, 1, 1)
        self.setSymmetryOn()
        super().setValue(A=IsotropicCoefficient(1.0))

    def setValue(self, f=None, q=None, **kw):
        c = {}
        if f is not None:
            c["Y"] = f
        if q is not None:
            c["q"] = q
        c.update(kw)
        super().setValue(**c)


class LameEquation(LinearPDE):
    """-(lambda u_k,k delta_ij + mu (u_i,j + u_j,i))_,j = F_i - sigma_ij,j
    (linearPDEs.py:2594-2706), assembled with the general system assembler
    as the reference does (DefaultAssembler, :2610-2614)."""

    def __init__(self, domain):
        d = domain.getDim()
        super().__init__(domain, d, d)
        self.setSymmetryOn()

    def setValues(self, lame_lambda=None, lame_mu=None, F=None, sigma=None, q=None, r=None):
        d = self.dim
        c = {}
        if lame_lambda is not None or lame_mu is not None:
            lam = torch.as_tensor(lame_lambda if lame_lambda is not None else 0.0, dtype=torch.float64)
            mu = torch.as_tensor(lame_mu if lame_mu is not None else 0.0, dtype=torch.float64)
            lam, mu = torch.broadcast_tensors(lam, mu)
            eye = torch.eye(d, dtype=torch.float64)
            # A_ijkl = lambda d_ij d_kl + mu (d_ik d_jl + d_il d_jk)  (linearPDEs.py:2677-2699)
            base_l = torch.einsum("ij,kl->ijkl", eye, eye)
            base_m = torch.einsum("ik,jl->ijkl", eye, eye) + torch.einsum("il,jk->ijkl", eye, eye)
            A = lam[..., None, None, None, None] * base_l + mu[..., None, None, None, None] * base_m
            c["A"] = A
        if F is not None:
            c["Y"] = F
        if sigma is not None:
            c["X"] = sigma
        if q is not None:
            c["q"] = q
        if r is not None:
            c["r"] = r
        self.setValue(**c)
