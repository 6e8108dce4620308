"""CPU restatement (numpy / scipy.sparse) of the multilevel preconditioner of
the learned flexible CG: test infrastructure only, the checker of the GPU
kernels in esys-escript.github.io_amd/csrc/rg_mg.hip.  The reference has no
counterpart (PASO's preconditioners are Jacobi, Gauss-Seidel, ILU0, RILU and
an algebraic multigrid, paso/src/Preconditioner.cpp:46-137); the method is
standard geometric multigrid on ripley's vertex-centred grid:

* prolongation P: trilinear interpolation from the coarse grid (every other
  fine node, ne/2 elements per axis) -- per axis P[2I, I] = 1,
  P[2I+1, I] = P[2I+1, I+1] = 1/2; 3-D P = Pz (x) Py (x) Px for x-fastest
  node numbering;
* coarse operators A_{l+1} = P^T A_l P (Galerkin), A_0 the unbalanced
  matrix;
* smoother: Chebyshev iteration on D^-1 A (D = diag A) over
  [lmax / ratio, lmax], lmax the Gershgorin bound max_i sum_j |a_ij| / a_ii;
* V-cycle: pre-smoothing from 0, residual, restriction P^T, recursion,
  prolongation, post-smoothing; the coarsest level solved exactly or by
  Chebyshev steps; at one level a given function (the network) may correct
  the iterate: applied to the residual left there, its output low-pass
  filtered by the level's Chebyshev steps on A y = 0 and scaled to minimise
  the level's energy-norm error along it.

Pinned in tests/test_oracle_multilevel.py by: P reproducing trilinear
functions exactly; the Galerkin product equal to the Q1 assembly on the
coarse grid for constant and 2x2x2-blockwise constant kappa (nested
conforming spaces); the Chebyshev smoother against the closed-form
Chebyshev error polynomial from an eigendecomposition; the V-cycle's
symmetry and convergence on a Poisson problem; the balanced form against
the physical one; the flexible CG against scipy's CG with a fixed
preconditioner; the energy-scaled step minimising the energy-norm error.
Planted mistakes (an interpolation weight, a Chebyshev coefficient's sign,
a missing balance factor) fail these checks.
"""
import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spla


def prolongation_1d(nc):
    """(2 nc - 1) x nc linear interpolation."""
    nf = 2 * nc - 1
    rows, cols, vals = [], [], []
    for i in range(nf):
        if i % 2 == 0:
            rows.append(i); cols.append(i // 2); vals.append(1.0)
        else:
            rows += [i, i]; cols += [(i - 1) // 2, (i + 1) // 2]; vals += [0.5, 0.5]
    return sp.csr_matrix((vals, (rows, cols)), shape=(nf, nc))


def prolongation(nodes_coarse):
    """P for node counts (n0, n1[, n2]) of the coarse grid (x fastest)."""
    Ps = [prolongation_1d(n) for n in nodes_coarse]
    P = Ps[0]
    for Pd in Ps[1:]:
        P = sp.kron(Pd, P, format="csr")
    return P


def coarsen_nodes(nodes):
    """Coarse node counts, or None when an axis has an odd number of
    elements or fewer than 2."""
    out = []
    for n in nodes:
        ne = n - 1
        if ne % 2 or ne < 2:
            return None
        out.append(ne // 2 + 1)
    return tuple(out)


def csr_from_paso(ptr, index, val):
    N = len(ptr) - 1
    return sp.csr_matrix((np.asarray(val, dtype=np.float64), np.asarray(index), np.asarray(ptr)), shape=(N, N))


def gershgorin(A):
    d = A.diagonal()
    s = np.asarray(abs(A).sum(axis=1)).ravel()
    return float(np.max(s / d))


class Level:
    def __init__(self, A, nodes, fp32=False):
        self.A = A.tocsr()
        self.nodes = nodes
        self.dinv = 1.0 / self.A.diagonal()
        self.lmax = gershgorin(self.A)
        # the operator of smoothing and residuals: values rounded to FP32 when
        # the kernels read an FP32 copy (arithmetic stays FP64)
        if fp32:
            S = self.A.copy()
            S.data = S.data.astype(np.float32).astype(np.float64)
            self.S = S
        else:
            self.S = self.A
        self.P = None


class Hierarchy:
    """Levels from the fine scalar operator A on `nodes` node counts;
    coarsening stops at max_coarse nodes or when it is impossible; `levels`
    caps the number of levels.  With bal (the balance vector d of a balanced
    matrix, paso/src/SystemMatrix.cpp:262-312) A is the balanced matrix
    D A_phys D: level 0 smooths it (Jacobi-scaled, as the balanced solver
    vectors are), and the levels below use A_phys = D^-1 A D^-1 -- the
    restriction takes P^T D^-1 r, the prolongation adds D^-1 P x_1.  fp32
    rounds the values that smoothing and residuals read."""

    def __init__(self, A, nodes, max_coarse=512, levels=None, degree=2, ratio=10.0, bal=None, fp32=False,
                 coarse_steps=8):
        self.degree = degree
        self.ratio = ratio
        self.coarse_steps = coarse_steps
        self.bal = None if bal is None else np.asarray(bal, dtype=np.float64)
        self.levels = [Level(A, tuple(nodes), fp32)]
        while True:
            L = self.levels[-1]
            if np.prod(L.nodes) <= max_coarse or (levels is not None and len(self.levels) >= levels):
                break
            nc = coarsen_nodes(L.nodes)
            if nc is None:
                break
            L.P = prolongation(nc)
            Af = L.A
            if len(self.levels) == 1 and self.bal is not None:
                Dinv = sp.diags(1.0 / self.bal)
                Af = Dinv @ Af @ Dinv
            Ac = (L.P.T @ Af @ L.P).tocsr()
            self.levels.append(Level(Ac, nc, fp32))
        self.coarse_inv = None

    def cheb_coefficients(self, l, steps):
        """[(c1, c2)] of `steps` Chebyshev steps on level l: d = c1 d + c2 D^-1 r."""
        lmax = self.levels[l].lmax
        lmin = lmax / self.ratio
        theta, delta = 0.5 * (lmax + lmin), 0.5 * (lmax - lmin)
        sigma = theta / delta
        rho_old = 1.0 / sigma
        out = [(0.0, 1.0 / theta)]
        for _ in range(steps - 1):
            rho = 1.0 / (2.0 * sigma - rho_old)
            out.append((rho * rho_old, 2.0 * rho / delta))
            rho_old = rho
        return out

    def smooth(self, l, b, x, steps):
        """`steps` Chebyshev steps on D^-1 A over [lmax/ratio, lmax] from x
        (None: zero)."""
        L = self.levels[l]
        d = None
        for k, (c1, c2) in enumerate(self.cheb_coefficients(l, steps)):
            r = b.copy() if (x is None and k == 0) else b - L.S @ x
            d = c2 * (L.dinv * r) + (c1 * d if c1 != 0.0 else 0.0)
            x = d.copy() if (x is None and k == 0) else x + d
        return x

    def exact_coarse(self, l, b):
        if self.coarse_inv is None:
            self.coarse_inv = np.linalg.inv(self.levels[-1].A.toarray())
        return self.coarse_inv @ b

    def vcycle(self, b, l=0, coarse=None, net=None, net_level=None):
        """x ~ A_l^-1 b.  coarse(l, b) solves the coarsest level (default:
        exactly; "smooth": coarse_steps Chebyshev steps from zero).  With
        net (a function (l, r) -> correction), level net_level ends with
        net_correct (after its post-smoothing, or after the coarse solve when
        it is the coarsest level)."""
        if l == len(self.levels) - 1:
            if coarse is None:
                x = self.exact_coarse(l, b)
            elif coarse == "smooth":
                x = self.smooth(l, b, None, self.coarse_steps)
            else:
                x = coarse(l, b)
            return self.net_correct(l, b, x, net) if (net is not None and l == net_level) else x
        L = self.levels[l]
        x = self.smooth(l, b, None, self.degree)
        r = b - L.S @ x
        if l == 0 and self.bal is not None:
            r = r / self.bal
        xc = self.vcycle(L.P.T @ r, l + 1, coarse, net, net_level)
        c = L.P @ xc
        if l == 0 and self.bal is not None:
            c = c / self.bal
        x = self.smooth(l, b, x + c, self.degree)
        return self.net_correct(l, b, x, net) if (net is not None and l == net_level) else x

    def net_correct(self, l, b, x, net):
        """The library's network correction of level l's iterate x: the
        residual r' = b - A x (FP32 values), y = net(l, r') low-pass filtered
        by coarse_steps Chebyshev steps on A y = 0 from y, then
        x + alpha y with the energy-optimal alpha (FP64 values)."""
        L = self.levels[l]
        r = b - L.S @ x
        y = net(l, r)
        y = self.smooth(l, np.zeros_like(y), y, self.coarse_steps)
        Ay = L.A @ y
        yAy = y @ Ay
        return x + y * ((r @ y) / yAy if yAy > 0 else 0.0)

    def energy_scaled(self, net):
        """coarse(l, b) for vcycle: y = net(l, b) scaled by (b.y)/(y.A y)."""
        def f(l, b):
            y = net(l, b)
            Ay = self.levels[l].A @ y
            yAy = y @ Ay
            return y * ((b @ y) / yAy if yAy > 0 else 0.0)
        return f


def pcg(A, b, prec, tol=1e-8, maxit=1000, m=1):
    """Flexible CG (Notay's FCG(m): each direction A-orthogonalised against
    the last m) with preconditioner prec(r); stops at ||b - A x|| <= tol ||b||.
    Returns (x, iterations)."""
    x = np.zeros_like(b)
    r = b.copy()
    nb = np.linalg.norm(b)
    P, Q, PQ = [], [], []
    for k in range(1, maxit + 1):
        z = prec(r)
        p = z.copy()
        for pj, qj, dj in zip(P, Q, PQ):
            p -= (z @ qj) / dj * pj
        q = A @ p
        pq = p @ q
        alpha = (p @ r) / pq
        x += alpha * p
        r -= alpha * q
        P.append(p); Q.append(q); PQ.append(pq)
        if len(P) > m:
            P.pop(0); Q.pop(0); PQ.pop(0)
        if np.linalg.norm(r) <= tol * nb:
            return x, k
    return x, maxit
