"""Learned coarse-level solver of the multilevel flexible CG
(SolverOptions.FCG_LEARNED).

The network (forward, backward and Adam kernels are the library's own BF16
MFMA kernels, FP32 accumulate) is a 3-level 3-D U-Net on a node grid:

    level 1 (node grid):      2 -> 16 -> 16          skip s1
    level 2 (1/2, avg pool):  16 -> 32 -> 32         skip s2
    level 3 (1/4, avg pool):  32 -> 64 -> 64
    decoder:  up(64) ++ s2 -> 32,  up(32) ++ s1 -> 16,  16 -> 1 (linear)

It corrects one level of the geometric-multigrid V-cycle that preconditions
the flexible CG, after the V-cycle's work there (its coarser levels and
smoothing): on the level's Galerkin operator A_c = P^T A P (trilinear P, the
Brick halved per level), input channels u / rms(u) with u = r / diag(A_c)
for the residual r left there and log10(diag(A_c) / diag of a unit
coefficient), output a correction proportional to the level's error
e = A_c^-1 r (the solver low-pass filters it and scales it to minimise the
energy-norm error).  It is trained with the
energy-norm loss 1 - (r.y)^2 / ((r.e)(y.A_c y)) -- the fraction of the
level's error energy the energy-optimal step along y leaves, which is what
the coarse correction removes; the cosine loss 1 - <y, e>/(|y||e|) ignores
rough output errors that cost much energy -- on residual/error pairs that
the plain Jacobi-PCG produces on A_c: r_k = b - A_c x_k and e_k = x* - x_k
for iterates x_k of the plain solver, with b = A_c x* for smooth random x*,
each after one V-cycle from that level (the solver applies the network to
the residual the V-cycle leaves there).

The reference has no learned preconditioner (PASO offers Jacobi,
Gauss-Seidel, ILU0, RILU and AMG, paso/src/Preconditioner.cpp:46-107); the
plain Jacobi-PCG stays the fallback: the solver switches to it when the
flexible CG stops reducing the residual.
"""
import ctypes

import numpy as np
import torch

from . import _lib
from ._lib import check, load
from .core import SolverOptions, _ptr, get_context

# (cin, cout, relu) of the 9 convolutions, in parameter order
LAYERS = ((16, 16, True), (16, 16, True), (16, 32, True), (32, 32, True), (32, 64, True), (64, 64, True),
          (96, 32, True), (48, 16, True), (16, 16, False))


def num_params():
    return int(load().rg_lpc_num_params())


def split_params(flat):
    """[(W[co][27][ci], b[co]), ...] views of a flat parameter vector."""
    out, o = [], 0
    for cin, cout, _ in LAYERS:
        ws = 27 * cin * cout
        out.append((flat[o:o + ws].reshape(cout, 27, cin), flat[o + ws:o + ws + cout]))
        o += ws + cout
    return out


def init_params(seed=0):
    """He-normal weights over the real fan-in (layer 0 sees 2 channels),
    zero biases."""
    rng = np.random.default_rng(seed)
    parts = []
    for l, (cin, cout, relu) in enumerate(LAYERS):
        fan = 27 * (2 if l == 0 else cin)
        std = np.sqrt((2.0 if relu else 1.0) / fan)
        w = rng.standard_normal((cout, 27, cin)) * std
        if l == 0:
            w[:, :, 2:] = 0.0
        parts.append(w.reshape(-1))
        parts.append(np.zeros(cout))
    return np.concatenate(parts).astype(np.float32)


class LearnedPreconditioner:
    """One network instance for a node grid nodes = (n0, n1, n2)."""

    def __init__(self, nodes, training=False, params=None, seed=0, device=0):
        self.ctx = get_context(device)
        self.device = torch.device("cuda", device)
        self.nodes = tuple(int(n) for n in nodes)
        self.training = bool(training)
        nd = (ctypes.c_int32 * 3)(*self.nodes)
        h = ctypes.c_void_p()
        check(load().rg_lpc_create(self.ctx.handle, nd, int(self.training), ctypes.byref(h)))
        self.handle = h
        self.set_params(init_params(seed) if params is None else params)

    def __del__(self):
        try:
            if self.handle:
                load().rg_lpc_destroy(self.handle)
                self.handle = None
        except Exception:
            pass

    @property
    def num_nodes(self):
        return self.nodes[0] * self.nodes[1] * self.nodes[2]

    def set_params(self, params):
        p = np.ascontiguousarray(params, dtype=np.float32).reshape(-1)
        if p.size != num_params():
            raise ValueError("expected %d parameters, got %d" % (num_params(), p.size))
        self.ctx.bind_stream()
        check(load().rg_lpc_set_params(self.handle, p.ctypes.data_as(ctypes.c_void_p)))

    def get_params(self):
        p = np.zeros(num_params(), dtype=np.float32)
        self.ctx.bind_stream()
        check(load().rg_lpc_get_params(self.handle, p.ctypes.data_as(ctypes.c_void_p)))
        return p

    def save(self, path):
        np.save(path, self.get_params())

    def load(self, path):
        self.set_params(np.load(path, allow_pickle=False))

    def _vec(self, v, dtype=torch.float64):
        if v is None:
            return None
        t = torch.as_tensor(v).to(device=self.device, dtype=dtype).contiguous().reshape(-1)
        if t.numel() != self.num_nodes:
            raise ValueError("vector has %d entries, expected %d" % (t.numel(), self.num_nodes))
        return t

    def apply(self, r, dinv=None, logk=None):
        """z = Net(r) on device vectors over the node grid."""
        r, dinv, lk = self._vec(r), self._vec(dinv), self._vec(logk, torch.float32)
        z = torch.empty(self.num_nodes, dtype=torch.float64, device=self.device)
        self.ctx.bind_stream()
        check(load().rg_lpc_apply(self.handle, _ptr(r), _ptr(dinv), _ptr(lk), _ptr(z)))
        return z

    def set_loss_operator(self, op):
        """The energy-norm loss 1 - (r.z)^2 / ((r.e)(z.A z)) with the level's
        operator op (an unbalanced SystemMatrix on this grid), or, with None,
        the cosine loss."""
        self.ctx.bind_stream()
        check(load().rg_lpc_set_loss_operator(self.handle, op.handle if op is not None else None))
        self._loss_op = op

    def gradients(self, r, dinv, logk, e, op=None):
        """(loss, flat FP32 gradient) of Net(r) against e: the energy-norm
        loss with op, else the cosine loss."""
        r, dinv, lk, e = self._vec(r), self._vec(dinv), self._vec(logk, torch.float32), self._vec(e)
        g = np.zeros(num_params(), dtype=np.float32)
        loss = ctypes.c_double()
        self.set_loss_operator(op)
        check(load().rg_lpc_gradients(self.handle, _ptr(r), _ptr(dinv), _ptr(lk), _ptr(e), ctypes.byref(loss),
                                      g.ctypes.data_as(ctypes.c_void_p)))
        return loss.value, g

    def train_step(self, r, dinv, logk, e, lr=1e-3, op=None):
        """One Adam step on the loss of gradients()."""
        r, dinv, lk, e = self._vec(r), self._vec(dinv), self._vec(logk, torch.float32), self._vec(e)
        loss = ctypes.c_double()
        self.set_loss_operator(op)
        check(load().rg_lpc_train_step(self.handle, _ptr(r), _ptr(dinv), _ptr(lk), _ptr(e), ctypes.c_float(lr),
                                       ctypes.byref(loss)))
        return loss.value


# --------------------------------------------------------------------------
# training data from the plain solver


def operator_diagonal(op):
    """diag(A) of an assembled, not yet balanced SystemMatrix (bs 1), via
    the PASO CSR export."""
    ptr, idx, val = op.exportCSR()
    rows = np.repeat(np.arange(op.num_rows, dtype=np.int64), np.diff(ptr))
    d = np.zeros(op.num_rows)
    m = idx == rows
    d[rows[m]] = val[m]
    return d


def level_inputs(op):
    """(dinv, logk) of the network on a level's operator: 1 / diag(A) and
    log10(diag(A) / diag of a unit coefficient on that grid), the inputs the
    library computes for its coarsest level."""
    d = operator_diagonal(op)
    h = op.domain.getLength()
    ne = op.domain.getNumElementsPerAxis()
    hs = [h[k] / ne[k] for k in range(3)]
    vol = hs[0] * hs[1] * hs[2]
    aref = sum(8.0 * vol / (9.0 * hk * hk) for hk in hs)
    dinv = torch.as_tensor(1.0 / d, device=op.device)
    logk = torch.as_tensor(np.log10(d / aref).astype(np.float32), device=op.device)
    return dinv, logk


def plain_pcg_pairs(make_operator, xstar, iterations, check_interval=32):
    """Residual/error pairs from the plain FP64 Jacobi-PCG on A x = b with
    b = A xstar: for each k in `iterations`, x_k is the plain solver's
    iterate after k iterations and the pair is (r_k = b - A x_k,
    e_k = xstar - x_k).  make_operator() returns a fresh, unbalanced
    operator (called once for products and once per iterate: a solve
    balances its matrix).  Returns [(r_k, e_k), ...] as device tensors."""
    A = make_operator()
    b = torch.zeros(A.size, dtype=torch.float64, device=A.device)
    A.MatrixVector(1.0, xstar, 0.0, b)
    pairs = []
    for k in iterations:
        S = make_operator()
        o = SolverOptions()
        o.setSolverMethod(SolverOptions.PCG_PLAIN)
        o.setTolerance(1e-12)
        o.setIterMax(int(k))
        o.setAcceptanceConvergenceFailureOn()
        o.setCheckInterval(max(1, min(check_interval, int(k))))
        xk = S.solve(b, o)
        del S
        r = b.clone()
        A.MatrixVector(-1.0, xk, 1.0, r)
        pairs.append((r, xstar - xk))
    return pairs


class BrickPoisson:
    """-div(kappa grad u) = f on Brick(ne) (unit cube), u = 0 on z = 0, kappa
    constant per element (isotropic A = kappa I on Function, Y = f) -- the
    problem family of the learned preconditioner."""

    def __init__(self, ne, kappa_elem, f=1.0, device=0, symmetric=False):
        from .core import Coefficient
        from .domain import Brick
        self.ne = tuple(int(n) for n in ne)
        self.dom = Brick(*self.ne, device=device)
        self.device = self.dom.device
        self.symmetric = symmetric
        self.kappa = torch.as_tensor(kappa_elem, dtype=torch.float64).to(self.device).reshape(-1)
        E, Q = self.dom.getNumElements(), self.dom.getNumQuadraturePoints()
        self.cA = Coefficient(self.kappa, (3, 3), E, Q, self.device, isotropic=True)
        self.cY = Coefficient(float(f), (), E, Q, self.device)
        self.q = self.dom.whereZero(2)
        self.nodes = tuple(n + 1 for n in self.ne)
        self.rhs = None

    def operator(self):
        """A freshly assembled, constrained operator (the rhs is kept)."""
        op = self.dom.newOperator(1, symmetric=self.symmetric)
        rhs = torch.zeros(op.size, dtype=torch.float64, device=self.device)
        op.assemble({"A": self.cA, "Y": self.cY}, rhs)
        op.nullifyRowsAndCols(self.q, self.q, 1.0)
        rhs.masked_fill_(self.q > 0, 0.0)
        self.rhs = rhs
        return op


DEFAULT_ITERATES = (1, 2, 4, 8, 16, 32, 64, 128, 256)


def smooth_target(nodes, seed, device):
    """A smooth random x* on a node grid (Gaussian random field, correlation
    2..8 nodes by the seed), zero on the Dirichlet plane z = 0."""
    from . import fields
    rng = np.random.default_rng(seed)
    corr = float(rng.uniform(2.0, 8.0))
    g = fields.gaussian_random_field(nodes, corr, int(rng.integers(1 << 30))).reshape(-1)
    g[: nodes[0] * nodes[1]] = 0.0
    return torch.as_tensor(g, device=device)


def generate_training_set(ne, num_fields, kappa_fn, iterations=DEFAULT_ITERATES, first_seed=1, log=None,
                          per_field=None, levels=1, after_vcycle=True):
    """[(r, e, dinv, logk, A_c), ...] on the device for the network that
    solves the level `levels` halvings below Brick(ne): for each seed s in
    first_seed .. first_seed+num_fields-1, kappa = kappa_fn(ne, s), the
    Galerkin operator A_c of the constrained fine operator, and the plain
    Jacobi-PCG's residual/error pairs on A_c x = A_c x* (x* smooth, random)
    at the given iteration counts (with per_field = m: m of them per field,
    chosen by the seed), each taken after one V-cycle V from that level (the
    default hierarchy below it): r' = r - A_c V r, e' = e - V r, the residual
    and error the network sees and corrects in the solver's V-cycle."""
    data = []
    for s in range(first_seed, first_seed + num_fields):
        prob = BrickPoisson(ne, kappa_fn(ne, s), symmetric=True)
        fine = prob.operator()
        make = lambda: fine.coarseOperator(levels)
        C = make()
        dinv, logk = level_inputs(C)
        nodes = tuple(n + 1 for n in C.domain.getNumElementsPerAxis())
        xstar = smooth_target(nodes, 7919 * s + 1, C.device)
        its = iterations
        if per_field:
            its = [iterations[(7 * s + j) % len(iterations)] for j in range(per_field)]
        for r, e in plain_pcg_pairs(make, xstar, its):
            # what one V-cycle from this level leaves (its levels below, the
            # default smoothing): the network corrects that
            if after_vcycle:
                xs = C.vcycle(r)
                C.MatrixVector(-1.0, xs, 1.0, r)
                e = e - xs
            data.append((r, e, dinv, logk, C))
        n = s - first_seed + 1
        if log and (n % max(1, num_fields // 20) == 0 or n == num_fields):
            log("training set: field %d/%d (%d pairs)" % (n, num_fields, len(data)))
    return data


def train(net, data, steps, lr=1e-3, lr_final=1e-4, seed=0, log=None, log_every=100, energy=True):
    """Adam over random pairs of `data` on the energy-norm loss with each
    pair's level operator (energy=False: the cosine loss); the learning rate
    decays geometrically from lr to lr_final.  Returns the loss history."""
    rng = np.random.default_rng(seed)
    hist = []
    for k in range(steps):
        r, e, dinv, logk, op = data[int(rng.integers(len(data)))]
        a = lr * (lr_final / lr) ** (k / max(1, steps - 1))
        hist.append(net.train_step(r, dinv, logk, e, a, op=op if energy else None))
        if log and (k % log_every == 0 or k == steps - 1):
            log("step %d loss %.4f (mean of last %d: %.4f)" % (k, hist[-1], min(len(hist), log_every),
                                                               float(np.mean(hist[-log_every:]))))
    return hist


def _main():
    import argparse
    import sys
    import time
    from . import fields
    ap = argparse.ArgumentParser(description="train the learned preconditioner on the GPU")
    ap.add_argument("--ne", type=int, default=64, help="elements per axis of the fine training grids")
    ap.add_argument("--levels", type=int, default=1, help="halvings from the fine grid to the network's level")
    ap.add_argument("--fields", type=int, default=32)
    ap.add_argument("--per-field", type=int, default=0,
                    help="residual/error pairs per field (0: one per iteration count 1, 2, 4, ..., 256)")
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--cosine", action="store_true", help="train on the cosine loss instead of the energy loss")
    ap.add_argument("--family", default="log10grf", choices=["log10grf", "lognormal"])
    ap.add_argument("--sigma", type=float, default=1.0)
    ap.add_argument("--corr", type=float, default=8.0)
    ap.add_argument("--init", default=None)
    ap.add_argument("--out", default=None)
    ap.add_argument("--apply-reps", type=int, default=0,
                    help="only time this many applications on a random residual over (ne+1)^3 nodes")
    a = ap.parse_args()
    if a.apply_reps > 0:
        nodes = (a.ne + 1,) * 3
        net = LearnedPreconditioner(nodes)
        if a.init:
            net.load(a.init)
        r = torch.randn(int(np.prod(nodes)), dtype=torch.float64, device="cuda")
        net.apply(r)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(a.apply_reps):
            net.apply(r)
        torch.cuda.synchronize()
        print("apply %s: %.3f ms" % (nodes, 1e3 * (time.perf_counter() - t0) / a.apply_reps), flush=True)
        sys.exit(0)
    if a.out is None:
        ap.error("--out is required for training")

    def kfn(ne, s):
        if a.family == "log10grf":
            return fields.log10_normal_kappa(ne, corr_len=a.corr, sigma=a.sigma, seed=s)
        return fields.lognormal_kappa(ne, corr_len=a.corr, sigma_log=a.sigma, seed=s)

    def log(msg):
        print(msg, flush=True)

    ne = (a.ne,) * 3
    t0 = time.perf_counter()
    data = generate_training_set(ne, a.fields, kfn, log=log, per_field=a.per_field or None, levels=a.levels)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    net = LearnedPreconditioner(tuple(n // 2 ** a.levels + 1 for n in ne), training=True)
    if a.init:
        net.load(a.init)
    hist = train(net, data, a.steps, lr=a.lr, log=log, energy=not a.cosine)
    torch.cuda.synchronize()
    t2 = time.perf_counter()
    net.save(a.out)
    log("data %.1fs, training %.1fs (%.2f ms/step), final loss %.4f" %
        (t1 - t0, t2 - t1, 1e3 * (t2 - t1) / max(1, a.steps), float(np.mean(hist[-100:]))))
    sys.exit(0)


if __name__ == "__main__":
    _main()
