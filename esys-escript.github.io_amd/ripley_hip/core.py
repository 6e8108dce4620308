"""Host-side mirror of the reference operator interface for the ripley->PASO
path: ``SystemMatrix`` (escript::AbstractSystemMatrix / paso::SystemMatrix)
and ``SolverOptions`` (escript::SolverBuddy), over the C ABI of
libripley_hip.  Vectors are float64 torch tensors on the GPU (DOF-major, then
component, like escript Data); numpy arrays are accepted by the ``*_host``
style helpers.
"""
import ctypes

import numpy as np
import torch

from . import _lib
from ._lib import check, load

# --------------------------------------------------------------------------
# context


class Context:
    """One device + the stream the caller's torch work runs on."""

    def __init__(self, device=0):
        self.device = device
        h = ctypes.c_void_p()
        check(load().rg_context_create(device, None, ctypes.byref(h)))
        self.handle = h

    def bind_stream(self):
        s = torch.cuda.current_stream(self.device).cuda_stream
        check(load().rg_context_set_stream(self.handle, ctypes.c_void_p(s)))

    def synchronize(self):
        check(load().rg_context_synchronize(self.handle))

    def __del__(self):
        try:
            if self.handle:
                load().rg_context_destroy(self.handle)
        except Exception:
            pass


_contexts = {}


def get_context(device=0):
    if not torch.cuda.is_available():
        raise RuntimeError("ripley_hip needs a ROCm GPU: torch.cuda is not available")
    if device not in _contexts:
        _contexts[device] = Context(device)
    ctx = _contexts[device]
    ctx.bind_stream()
    return ctx


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else None


def _dev_vec(x, n, device):
    """A contiguous float64 device tensor of n entries viewing/copying x."""
    if isinstance(x, torch.Tensor):
        t = x.to(device=device, dtype=torch.float64).contiguous().reshape(-1)
    else:
        t = torch.as_tensor(np.ascontiguousarray(x, dtype=np.float64).reshape(-1), device=device)
    if t.numel() != n:
        raise ValueError("vector has %d entries, expected %d" % (t.numel(), n))
    return t


# --------------------------------------------------------------------------
# solver options


class SolverOptions:
    """Mirror of escript's SolverOptions / SolverBuddy for the settings the
    PASO PCG path reads (escriptcore/src/SolverOptions.h:81-142, defaults
    escriptcore/src/SolverOptions.cpp:40-72)."""

    DEFAULT = 0
    TARGET_CPU, TARGET_GPU = 1, 2
    PASO = 4
    BICGSTAB = 8
    MINRES = 22
    PCG = 24
    GAUSS_SEIDEL = 29
    ILU0 = 30
    JACOBI = 32
    NO_PRECONDITIONER = 33
    # methods of this implementation beyond the reference's enum
    PCG_PLAIN = 1001
    PCG_MIXED = 1002
    FCG_LEARNED = 1003

    def __init__(self):
        self._method = SolverOptions.DEFAULT
        self._preconditioner = SolverOptions.JACOBI
        self._package = SolverOptions.DEFAULT
        self._tolerance = 1e-8
        self._iter_max = 100000
        self._sweeps = 1
        self._verbose = False
        self._symmetric = False
        self._accept = False
        self._inner_tolerance = 1e-4
        self._check_interval = 0
        self._true_residual = False
        self.resetDiagnostics(all=True)

    # setters / getters named as SolverBuddy's python binding
    def setSolverMethod(self, method):
        if method not in (self.DEFAULT, self.PCG, self.PCG_PLAIN, self.PCG_MIXED, self.FCG_LEARNED, self.BICGSTAB,
                          self.MINRES):
            raise ValueError("unknown solver method %r" % (method,))
        self._method = method

    def getSolverMethod(self):
        return self._method

    def setPreconditioner(self, preconditioner):
        if preconditioner not in (self.JACOBI, self.GAUSS_SEIDEL, self.ILU0, self.NO_PRECONDITIONER):
            raise ValueError("preconditioner %r is not provided on this path "
                             "(JACOBI, GAUSS_SEIDEL, ILU0, NO_PRECONDITIONER)" % (preconditioner,))
        self._preconditioner = preconditioner

    def getPreconditioner(self):
        return self._preconditioner

    def setPackage(self, package):
        if package not in (self.DEFAULT, self.PASO):
            raise ValueError("only the PASO package is provided on this path")
        self._package = package

    def getPackage(self):
        return self._package

    def setTolerance(self, rtol=1e-8):
        if rtol < 0 or rtol > 1:
            raise ValueError("tolerance must be between 0 and 1")
        self._tolerance = float(rtol)

    def getTolerance(self):
        return self._tolerance

    def setIterMax(self, iter_max=100000):
        if iter_max < 1:
            raise ValueError("maximum number of iteration steps must be positive")
        self._iter_max = int(iter_max)

    def getIterMax(self):
        return self._iter_max

    def setNumSweeps(self, sweeps=1):
        if sweeps < 1:
            raise ValueError("number of sweeps must be positive")
        self._sweeps = int(sweeps)

    def getNumSweeps(self):
        return self._sweeps

    def setVerbosityOn(self):
        self._verbose = True

    def setVerbosityOff(self):
        self._verbose = False

    def setVerbosity(self, verbose=False):
        self._verbose = bool(verbose)

    def isVerbose(self):
        return self._verbose

    def setSymmetryOn(self):
        self._symmetric = True

    def setSymmetryOff(self):
        self._symmetric = False

    def isSymmetric(self):
        return self._symmetric

    def setAcceptanceConvergenceFailureOn(self):
        self._accept = True

    def setAcceptanceConvergenceFailureOff(self):
        self._accept = False

    def acceptConvergenceFailure(self):
        return self._accept

    # extensions of this implementation
    def setInnerTolerance(self, rtol=1e-4):
        self._inner_tolerance = float(rtol)

    def getInnerTolerance(self):
        return self._inner_tolerance

    def setCheckInterval(self, iters=0):
        self._check_interval = int(iters)

    def setTrueResidualCriterion(self, flag=True):
        """Stop on the FP64 unbalanced ||b - A x||_2 <= tol ||b||_2 instead of
        PASO's balanced two-norm / max-norm test."""
        self._true_residual = bool(flag)

    # diagnostics (paso/src/Options.cpp:441-459)
    _DIAG_KEYS = ("num_iter", "num_level", "num_inner_iter", "time", "set_up_time", "net_time",
                  "residual_norm", "converged", "cum_num_iter", "cum_time", "cum_set_up_time",
                  "cum_net_time", "outer_steps", "fallbacks")

    def resetDiagnostics(self, all=False):
        keep = {} if all else {k: v for k, v in getattr(self, "_diag", {}).items() if k.startswith("cum_")}
        self._diag = {"num_iter": 0, "num_level": 0, "num_inner_iter": 0, "time": 0.0, "set_up_time": 0.0,
                      "net_time": 0.0, "residual_norm": 0.0, "converged": False, "outer_steps": 0, "fallbacks": 0,
                      "cum_num_iter": 0, "cum_time": 0.0, "cum_set_up_time": 0.0, "cum_net_time": 0.0}
        self._diag.update(keep)

    def _updateDiagnostics(self, name, value):
        self._diag[name] = value
        if "cum_" + name in self._diag:
            self._diag["cum_" + name] += value

    def getDiagnostics(self, name):
        if name not in self._diag:
            raise ValueError("unknown diagnostic item %r" % (name,))
        return self._diag[name]

    def hasConverged(self):
        return bool(self._diag["converged"])

    def _rg_options(self):
        o = _lib.RgOptions()
        load().rg_options_default(ctypes.byref(o))
        o.tolerance = self._tolerance
        o.inner_tolerance = self._inner_tolerance
        o.iter_max = self._iter_max
        o.sweeps = self._sweeps
        m = self._method
        if m == self.DEFAULT:
            # PASO's default: PCG for symmetric, BiCGStab otherwise (paso/src/Options.cpp)
            m = self.PCG if self._symmetric else self.BICGSTAB
        o.method = {self.PCG: _lib.RG_METHOD_PCG, self.PCG_PLAIN: _lib.RG_METHOD_PCG_PLAIN,
                    self.PCG_MIXED: _lib.RG_METHOD_PCG_MIXED, self.FCG_LEARNED: _lib.RG_METHOD_FCG_LEARNED,
                    self.BICGSTAB: _lib.RG_METHOD_BICGSTAB, self.MINRES: _lib.RG_METHOD_MINRES}[m]
        o.verbose = int(self._verbose)
        o.accept_failed_convergence = int(self._accept)
        o.check_interval = self._check_interval
        o.criterion = _lib.RG_CRITERION_TRUE_RESIDUAL if self._true_residual else _lib.RG_CRITERION_PASO
        o.preconditioner = {self.JACOBI: _lib.RG_PRECONDITIONER_JACOBI,
                            self.GAUSS_SEIDEL: _lib.RG_PRECONDITIONER_GAUSS_SEIDEL,
                            self.ILU0: _lib.RG_PRECONDITIONER_ILU0,
                            self.NO_PRECONDITIONER: _lib.RG_PRECONDITIONER_NONE}[self._preconditioner]
        return o


# --------------------------------------------------------------------------
# system matrix


def _to_point_major(value, point_shape):
    """escript stores a data point column-major: reverse the point axes."""
    npt = len(point_shape)
    if npt == 0:
        return value
    lead = value.ndim - npt
    perm = tuple(range(lead)) + tuple(range(value.ndim - 1, lead - 1, -1))
    return value.permute(perm).contiguous()


class Coefficient:
    """A PDE coefficient prepared for the C ABI (keeps its device tensor alive)."""

    def __init__(self, value, point_shape, num_elements, num_quad, device, isotropic=False):
        self.tensor = None
        self.mode = _lib.RG_COEF_EMPTY
        self.stride = 0
        self.isotropic = int(isotropic)
        if value is None:
            return
        t = value if isinstance(value, torch.Tensor) else torch.as_tensor(np.asarray(value, dtype=np.float64))
        t = t.to(device=device, dtype=torch.float64)
        ps = () if isotropic else tuple(point_shape)
        npt = len(ps)
        psize = int(np.prod(ps)) if npt else 1
        if t.ndim == npt:
            if tuple(t.shape) != ps:
                raise ValueError("coefficient shape %s, expected %s" % (tuple(t.shape), ps))
            self.mode, self.stride = _lib.RG_COEF_ELEMENT, 0
        elif t.ndim == npt + 1:
            if tuple(t.shape) != (num_elements,) + ps:
                raise ValueError("coefficient shape %s, expected %s" % (tuple(t.shape), (num_elements,) + ps))
            self.mode, self.stride = _lib.RG_COEF_ELEMENT, psize
        elif t.ndim == npt + 2:
            if tuple(t.shape) != (num_elements, num_quad) + ps:
                raise ValueError("coefficient shape %s, expected %s" % (tuple(t.shape),
                                                                        (num_elements, num_quad) + ps))
            self.mode, self.stride = _lib.RG_COEF_EXPANDED, psize * num_quad
        else:
            raise ValueError("bad coefficient rank %d" % t.ndim)
        self.tensor = _to_point_major(t, ps).reshape(-1).contiguous()

    def c(self):
        c = _lib.RgCoef()
        c.data = self.tensor.data_ptr() if self.tensor is not None else None
        c.mode = self.mode
        c.elem_stride = self.stride
        c.isotropic = self.isotropic
        return c


class SystemMatrix:
    """GPU system matrix with the escript::AbstractSystemMatrix /
    paso::SystemMatrix surface (escriptcore/src/AbstractSystemMatrix.h:43-190,
    paso/src/SystemMatrix.h:40-340) for one Rectangle/Brick at one rank."""

    def __init__(self, domain, block_size=1, symmetric=False, device=0):
        self.ctx = get_context(device)
        self.domain = domain
        self.device = torch.device("cuda", device)
        self.block_size = int(block_size)
        self.symmetric = bool(symmetric)
        ne = (ctypes.c_int32 * 3)(*(list(domain.getNumElementsPerAxis()) + [1] * (3 - domain.getDim())))
        ln = (ctypes.c_double * 3)(*(list(domain.getLength()) + [1.0] * (3 - domain.getDim())))
        h = ctypes.c_void_p()
        check(load().rg_matrix_create(self.ctx.handle, domain.getDim(), ne, ln, self.block_size,
                                      _lib.RG_MATRIX_SYMMETRIC if symmetric else 0, ctypes.byref(h)))
        self.handle = h
        nr, nnz, ns, bs = ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int(), ctypes.c_int()
        check(load().rg_matrix_get_info(h, ctypes.byref(nr), ctypes.byref(nnz), ctypes.byref(ns),
                                        ctypes.byref(bs)))
        self.num_rows = nr.value
        self.nnz = nnz.value
        self.num_slots = ns.value
        self.size = self.num_rows * self.block_size

    def __del__(self):
        try:
            if self.handle:
                load().rg_matrix_destroy(self.handle)
                self.handle = None
        except Exception:
            pass

    def _bind(self):
        self.ctx.bind_stream()

    # --- paso::SystemMatrix / AbstractSystemMatrix surface ---------------
    def getRowBlockSize(self):
        return self.block_size

    def getColumnBlockSize(self):
        return self.block_size

    def isEmpty(self):
        return False

    def setValues(self, value):
        self._bind()
        check(load().rg_matrix_set_values(self.handle, float(value)))

    def resetValues(self, preserveSolverData=False):
        self._bind()
        check(load().rg_matrix_reset_values(self.handle, int(bool(preserveSolverData))))

    def MatrixVector(self, alpha, x, beta, y):
        """y = alpha*A*x + beta*y in place (device tensors)."""
        self._bind()
        xd = _dev_vec(x, self.size, self.device)
        if not (isinstance(y, torch.Tensor) and y.is_cuda and y.dtype == torch.float64 and y.is_contiguous()):
            raise ValueError("y must be a contiguous float64 GPU tensor")
        check(load().rg_matrix_vector(self.handle, float(alpha), _ptr(xd), float(beta), _ptr(y)))
        return y

    def ypAx(self, y, x):
        return self.MatrixVector(1.0, x, 1.0, y)

    def vectorMultiply(self, x):
        y = torch.zeros(self.size, dtype=torch.float64, device=self.device)
        return self.ypAx(y, x)

    def __mul__(self, x):
        return self.vectorMultiply(x)

    def nullifyRowsAndCols(self, row_q, col_q, mdv=1.0):
        self._bind()
        r = _dev_vec(row_q, self.size, self.device)
        c = r if col_q is row_q else _dev_vec(col_q, self.size, self.device)
        check(load().rg_matrix_nullify_rows_and_cols(self.handle, _ptr(r), _ptr(c), float(mdv)))

    def saveMM(self, filename):
        check(load().rg_matrix_save_mm(self.handle, filename.encode()))

    def addElementBlocks(self, nodes, em):
        """addToSystemMatrix for a batch of elements (host arrays)."""
        self._bind()
        nodes = np.ascontiguousarray(nodes, dtype=np.int32)
        em = np.ascontiguousarray(em, dtype=np.float64)
        npe = nodes.shape[-1]
        nel = nodes.size // npe
        check(load().rg_matrix_add_element_blocks(self.handle, nel, npe, nodes.ctypes.data_as(ctypes.c_void_p),
                                                  self.block_size, em.ctypes.data_as(ctypes.c_void_p)))

    def assemble(self, coefs, rhs=None, reduced=False):
        """Add the element contributions of the given Coefficient objects
        (keys A,B,C,D,X,Y) to the matrix and to rhs (device tensor or None)."""
        self._bind()
        pc = _lib.RgPdeCoefs()
        for k in "ABCDXY":
            c = coefs.get(k)
            setattr(pc, k, c.c() if c is not None else _lib.RgCoef())
        if rhs is not None and (rhs.numel() != self.size or rhs.dtype != torch.float64 or not rhs.is_cuda):
            raise ValueError("rhs must be a float64 GPU tensor of %d entries" % self.size)
        check(load().rg_matrix_assemble(self.handle, ctypes.byref(pc), int(bool(reduced)), _ptr(rhs)))

    def setLearnedPreconditioner(self, net, logk=None):
        """Attach a LearnedPreconditioner as the coarsest-level solver of
        SolverOptions.FCG_LEARNED's multilevel preconditioner: its node grid
        must be a coarse level of this Brick (every other node per level),
        which fixes the number of levels.  net=None detaches.  The network is
        borrowed: keep it alive while solving.  logk is accepted for
        compatibility and ignored (the network reads its level's operator)."""
        self._bind()
        if net is None:
            check(load().rg_matrix_set_learned_preconditioner(self.handle, None, None))
            self._lpc = None
            return
        check(load().rg_matrix_set_learned_preconditioner(self.handle, net.handle, None))
        self._lpc = net

    def setMultilevel(self, levels=0, degree=2, coarse_steps=8):
        """Hierarchy of FCG_LEARNED without a network: grid levels (0:
        automatic), Chebyshev smoothing steps per side of each coarse
        correction, Chebyshev steps on the coarsest level."""
        self._bind()
        check(load().rg_matrix_set_multilevel(self.handle, int(levels), int(degree), int(coarse_steps)))

    def smooth(self, b, steps):
        """`steps` Chebyshev steps from zero for A x = b (the multilevel
        preconditioner's smoother) on this unbalanced matrix."""
        self._bind()
        bd = _dev_vec(b, self.size, self.device)
        x = torch.empty(self.size, dtype=torch.float64, device=self.device)
        check(load().rg_matrix_smooth(self.handle, _ptr(bd), _ptr(x), int(steps)))
        return x

    def vcycle(self, b):
        """One V-cycle of the multilevel preconditioner for A x = b on this
        unbalanced matrix (its multilevel settings and network)."""
        self._bind()
        bd = _dev_vec(b, self.size, self.device)
        x = torch.empty(self.size, dtype=torch.float64, device=self.device)
        check(load().rg_matrix_vcycle(self.handle, _ptr(bd), _ptr(x)))
        return x

    def coarseOperator(self, levels=1):
        """The Galerkin operator P^T A P `levels` halvings down (P trilinear
        interpolation), as a SystemMatrix of the coarse Brick."""
        from .domain import Brick
        self._bind()
        h = ctypes.c_void_p()
        check(load().rg_matrix_coarse_operator(self.handle, int(levels), ctypes.byref(h)))
        ne = [n // (2 ** levels) for n in self.domain.getNumElementsPerAxis()]
        dom = Brick(*ne, *self.domain.getLength(), device=self.device.index or 0)
        return SystemMatrix._wrap(dom, h, self.symmetric, self.device.index or 0)

    @classmethod
    def _wrap(cls, domain, handle, symmetric, device=0):
        """A SystemMatrix owning an existing library handle."""
        self = cls.__new__(cls)
        self.ctx = get_context(device)
        self.domain = domain
        self.device = torch.device("cuda", device)
        self.block_size = 1
        self.symmetric = bool(symmetric)
        self.handle = handle
        nr, nnz, ns, bs = ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int(), ctypes.c_int()
        check(load().rg_matrix_get_info(handle, ctypes.byref(nr), ctypes.byref(nnz), ctypes.byref(ns),
                                        ctypes.byref(bs)))
        self.num_rows = nr.value
        self.nnz = nnz.value
        self.num_slots = ns.value
        self.size = self.num_rows
        return self

    def assembleBoundary(self, coefs, rhs=None, reduced=False):
        """Face terms d (matrix) and y (rhs) of the Coefficient objects under
        keys "d", "y" (RipleyDomain::assemblePDEBoundary ->
        DefaultAssembler*::assemblePDEBoundarySingle/System[Reduced])."""
        self._bind()
        cd = coefs.get("d")
        cy = coefs.get("y")
        if rhs is not None and (rhs.numel() != self.size or rhs.dtype != torch.float64 or not rhs.is_cuda):
            raise ValueError("rhs must be a float64 GPU tensor of %d entries" % self.size)
        d = cd.c() if cd is not None else _lib.RgCoef()
        y = cy.c() if cy is not None else _lib.RgCoef()
        check(load().rg_matrix_assemble_boundary(self.handle, ctypes.byref(d), ctypes.byref(y), int(bool(reduced)),
                                                 _ptr(rhs)))

    def addDirac(self, nodes, d=None, y=None, rhs=None):
        """Dirac point terms (RipleyDomain::assemblePDEDirac): d (count, n, n)
        added to the diagonal blocks of the nodes, y (count, n) to rhs."""
        self._bind()
        nodes = np.ascontiguousarray(nodes, dtype=np.int64).reshape(-1)
        n = self.block_size
        dd = None if d is None else np.ascontiguousarray(
            _to_point_major(torch.as_tensor(np.asarray(d, dtype=np.float64)).reshape(len(nodes), n, n),
                            (n, n)).numpy())
        yy = None if y is None else np.ascontiguousarray(np.asarray(y, dtype=np.float64).reshape(len(nodes), n))
        check(load().rg_matrix_add_dirac(self.handle, len(nodes), nodes.ctypes.data_as(ctypes.c_void_p),
                                         dd.ctypes.data_as(ctypes.c_void_p) if dd is not None else None,
                                         yy.ctypes.data_as(ctypes.c_void_p) if yy is not None else None,
                                         _ptr(rhs)))

    def exportCSR(self):
        """(ptr, index, val) numpy arrays of the PASO CSR form."""
        ptr = np.zeros(self.num_rows + 1, dtype=np.int32)
        idx = np.zeros(self.nnz, dtype=np.int32)
        val = np.zeros(self.nnz * self.block_size ** 2, dtype=np.float64)
        self._bind()
        check(load().rg_matrix_export_csr(self.handle, ptr.ctypes.data_as(ctypes.c_void_p),
                                          idx.ctypes.data_as(ctypes.c_void_p),
                                          val.ctypes.data_as(ctypes.c_void_p)))
        return ptr, idx, val

    def setToSolution(self, out, in_, options):
        """paso::SystemMatrix::setToSolution (paso/src/SystemMatrix.h:615-644)."""
        self._bind()
        if options is None:
            options = SolverOptions()
            options.setSymmetryOn()
            options.setSolverMethod(SolverOptions.PCG)
        options.resetDiagnostics()
        o = options._rg_options()
        b = _dev_vec(in_, self.size, self.device)
        if not (isinstance(out, torch.Tensor) and out.is_cuda and out.dtype == torch.float64
                and out.is_contiguous() and out.numel() == self.size):
            raise ValueError("out must be a contiguous float64 GPU tensor of %d entries" % self.size)
        dg = _lib.RgDiagnostics()
        rc = load().rg_solve(self.handle, _ptr(out), _ptr(b), ctypes.byref(o), ctypes.byref(dg))
        for k in ("num_iter", "set_up_time", "net_time", "time", "residual_norm", "outer_steps", "fallbacks"):
            options._updateDiagnostics(k, getattr(dg, k))
        options._updateDiagnostics("converged", bool(dg.converged))
        check(rc)
        return out

    def solvePreconditioner(self, in_, options=None):
        """SystemMatrix::setPreconditioner + solvePreconditioner
        (paso/src/SystemMatrix.cpp:42-56): balances the matrix if needed,
        builds the preconditioner the options select and returns M^-1 in_
        (balanced space)."""
        self._bind()
        if options is None:
            options = SolverOptions()
        o = options._rg_options()
        b = _dev_vec(in_, self.size, self.device)
        out = torch.empty(self.size, dtype=torch.float64, device=self.device)
        check(load().rg_matrix_solve_preconditioner(self.handle, ctypes.byref(o), _ptr(out), _ptr(b)))
        return out

    def getColoring(self):
        """Pattern::borrowColoringPointer / getNumColors
        (paso/src/Pattern.cpp:166-195): (colour per node row, number of colours)."""
        self._bind()
        import numpy as np
        c = np.zeros(self.num_rows, dtype=np.int32)
        nc = ctypes.c_int(0)
        check(load().rg_matrix_coloring(self.handle, c.ctypes.data_as(ctypes.c_void_p), ctypes.byref(nc)))
        return c, nc.value

    def solve(self, in_, options=None):
        """AbstractSystemMatrix::solve (escriptcore/src/AbstractSystemMatrix.cpp:89-105)."""
        out = torch.zeros(self.size, dtype=torch.float64, device=self.device)
        return self.setToSolution(out, in_, options)
