"""Learned preconditioner on the GPU (through the C ABI): forward pass and
gradients of the BF16-MFMA kernels against the CPU restatement of the
network (oracle/lpc_torch.py, which rounds weights and activations to BF16
as the kernels store them), the flexible CG with the network attached
(FP64-verified residual), its fallback to Jacobi-PCG, and what a network
trained on the plain solver's residual/error pairs gains over the untrained
one."""
import numpy as np
import pytest
import torch

import lpc_torch as LT  # noqa: E402
import oracle as O
import ripley_hip as R
from ripley_hip import fields
from ripley_hip import learned as L

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def _random_params(seed):
    rng = np.random.default_rng(seed)
    p = L.init_params(seed)
    for (w, b) in L.split_params(p):
        b[:] = 0.1 * rng.standard_normal(b.shape)
    return p


def _inputs(nodes, seed):
    rng = np.random.default_rng(seed)
    n = nodes[0] * nodes[1] * nodes[2]
    r = rng.standard_normal(n)
    dinv = rng.uniform(0.5, 2.0, n)
    logk = rng.standard_normal(n).astype(np.float32)
    e = rng.standard_normal(n)
    return r, dinv, logk, e


@pytest.mark.parametrize("nodes", [(8, 8, 8), (9, 10, 11), (17, 13, 6)])
def test_forward_matches_oracle(nodes):
    print("forward", nodes, flush=True)
    p = _random_params(1)
    r, dinv, logk, _ = _inputs(nodes, 2)
    net = L.LearnedPreconditioner(nodes, params=p)
    z = net.apply(torch.as_tensor(r, device="cuda"), torch.as_tensor(dinv, device="cuda"),
                  torch.as_tensor(logk, device="cuda")).cpu().numpy()
    zb = LT.apply(p, nodes, r, dinv, logk, bf16=True)
    zf = LT.apply(p, nodes, r, dinv, logk, bf16=False)
    # same BF16 roundings: differences only from FP32 summation order
    assert _rel(z, zb) < 1e-2, _rel(z, zb)
    # against the unrounded FP32 network: BF16 storage error only
    assert _rel(z, zf) < 5e-2, _rel(z, zf)


def test_forward_without_optional_inputs():
    nodes = (8, 9, 10)
    p = _random_params(4)
    r, _, _, _ = _inputs(nodes, 5)
    net = L.LearnedPreconditioner(nodes, params=p)
    z = net.apply(torch.as_tensor(r, device="cuda")).cpu().numpy()
    assert _rel(z, LT.apply(p, nodes, r, None, None, bf16=True)) < 1e-2


def test_params_round_trip_and_errors():
    nodes = (8, 8, 8)
    p = _random_params(6)
    net = L.LearnedPreconditioner(nodes, params=p)
    assert np.array_equal(net.get_params(), p)
    with pytest.raises(R.RipleyHipError):
        net.gradients(torch.zeros(512, device="cuda", dtype=torch.float64), None, None,
                      torch.zeros(512, device="cuda", dtype=torch.float64))   # not created for training
    with pytest.raises(ValueError):
        net.set_params(p[:-1])


@pytest.mark.parametrize("nodes", [(8, 8, 8), (13, 10, 9)])
def test_gradients_match_oracle(nodes):
    print("gradients", nodes, flush=True)
    p = _random_params(7)
    r, dinv, logk, e = _inputs(nodes, 8)
    net = L.LearnedPreconditioner(nodes, training=True, params=p)
    d = lambda a: torch.as_tensor(a, device="cuda")
    loss, g = net.gradients(d(r), d(dinv), d(logk), d(e))
    lref, gref = LT.loss_and_gradients(p, nodes, r, dinv, logk, e, bf16=True)
    assert abs(loss - lref) < 2e-3, (loss, lref)
    # per layer (weights, biases): BF16-stored activations and gradients
    o = 0
    for l, (cin, cout, _) in enumerate(L.LAYERS):
        ws = 27 * cin * cout
        gw, gwr = g[o:o + ws], gref[o:o + ws]
        gb, gbr = g[o + ws:o + ws + cout], gref[o + ws:o + ws + cout]
        if l == 0:   # input channels 2..15 are zero: their gradient is exactly zero
            gw3, gwr3 = gw.reshape(cout, 27, cin), gwr.reshape(cout, 27, cin)
            assert np.all(gw3[:, :, 2:] == 0.0)
            gw, gwr = gw3[:, :, :2], gwr3[:, :, :2]
        assert _rel(gw, gwr) < 5e-2, (l, "W", _rel(gw, gwr))
        assert _rel(gb, gbr) < 5e-2, (l, "b", _rel(gb, gbr))
        o += ws + cout


def test_train_step_is_adam_on_the_gradient():
    nodes = (8, 8, 8)
    p = _random_params(9)
    r, dinv, logk, e = _inputs(nodes, 10)
    d = lambda a: torch.as_tensor(a, device="cuda")
    net = L.LearnedPreconditioner(nodes, training=True, params=p)
    _, g = net.gradients(d(r), d(dinv), d(logk), d(e))
    lr = 1e-3
    net.train_step(d(r), d(dinv), d(logk), d(e), lr)
    p1 = net.get_params()
    # first Adam step: w -= lr * g / (|g| + eps) (bias corrections cancel)
    expect = p - lr * g / (np.abs(g) + 1e-8)
    assert np.allclose(p1, expect, rtol=0, atol=1e-6)


def _problem(ne, seed, sigma=1.0, corr=4.0):
    kap = fields.log10_normal_kappa(ne, corr_len=corr, sigma=sigma, seed=seed)
    return L.BrickPoisson(ne, kap)


def _fcg_solve(prob, net, tol=1e-8, check=8):
    op = prob.operator()
    op.setLearnedPreconditioner(net, prob.logk)
    o = R.SolverOptions()
    o.setSolverMethod(R.SolverOptions.FCG_LEARNED)
    o.setTolerance(tol)
    o.setTrueResidualCriterion(True)
    o.setCheckInterval(check)
    x = op.solve(prob.rhs, o)
    return x, o


def _true_res(prob, x):
    A = prob.operator()
    r = prob.rhs.clone()
    A.MatrixVector(-1.0, x, 1.0, r)
    return (torch.linalg.norm(r) / torch.linalg.norm(prob.rhs)).item()


def test_fcg_converges_fp64_verified():
    ne = (15, 14, 13)
    prob = _problem(ne, 11)
    net = L.LearnedPreconditioner(prob.nodes, params=_random_params(12))
    x, o = _fcg_solve(prob, net, tol=1e-9)
    assert o.hasConverged()
    assert _true_res(prob, x) <= 1e-9 * 1.001
    # same solution as the oracle's PASO restatement
    ptr, idx, val = prob.operator().exportCSR()
    xo, dg, _, _ = O.solve(1, ptr, idx, val, prob.rhs.cpu().numpy(), tolerance=1e-10)
    assert np.linalg.norm(x.cpu().numpy() - xo) <= 1e-6 * np.linalg.norm(xo)


def identity_params():
    """A network whose output channel is its input channel 0 (the
    Jacobi-scaled residual): relu(u) - relu(-u) carried through the level-1
    skip path; everything else zero.  With it the flexible CG is Jacobi-PCG
    up to the BF16 rounding of u."""
    p = np.zeros(L.num_params(), np.float32)
    W = L.split_params(p)
    c = 13   # centre tap (dx, dy, dz) = (0, 0, 0)
    W[0][0][0, c, 0], W[0][0][1, c, 0] = 1.0, -1.0
    W[1][0][0, c, 0], W[1][0][1, c, 1] = 1.0, 1.0
    W[7][0][0, c, 32], W[7][0][1, c, 33] = 1.0, 1.0
    W[8][0][0, c, 0], W[8][0][0, c, 1] = 1.0, -1.0
    return p


def _bf16(v):
    return torch.as_tensor(v.astype(np.float32)).to(torch.bfloat16).to(torch.float64).numpy()


def _cpu_fcg_iterations(M, b, prec, tol):
    """Truncated flexible CG (one previous direction) in FP64 on the CPU:
    the restatement the GPU method is checked against."""
    x = np.zeros(len(b))
    r = b.copy()
    nb = np.linalg.norm(b)
    pp = qp = None
    for k in range(1, 20000):
        z = prec(r)
        p = z if pp is None else z - (z @ qp) / (pp @ qp) * pp
        q = M @ p
        a = (r @ p) / (p @ q)
        x += a * p
        r -= a * q
        pp, qp = p, q
        if np.linalg.norm(r) <= tol * nb:
            return k
    return 20000


def test_fcg_with_identity_network_matches_cpu_restatement():
    """The identity network makes the preconditioner Jacobi applied to the
    BF16-rounded (normalised) residual; the GPU flexible CG must need the
    iterations the CPU restatement of the same method needs (the BF16
    rounding alone roughly doubles them against exact Jacobi-PCG)."""
    import scipy.sparse as sp
    ne = (20, 18, 16)
    prob = _problem(ne, 15, sigma=1.5)
    net = L.LearnedPreconditioner(prob.nodes, params=identity_params())
    r = torch.randn(net.num_nodes, dtype=torch.float64, device="cuda")
    d = torch.rand(net.num_nodes, dtype=torch.float64, device="cuda") + 0.5
    z = net.apply(r, d)
    assert torch.allclose(z, d * r, rtol=2 ** -8, atol=0)
    x, o = _fcg_solve(prob, net, tol=1e-9)
    itf = o.getDiagnostics("num_iter")
    assert o.hasConverged() and o.getDiagnostics("fallbacks") == 0
    assert _true_res(prob, x) <= 1e-9 * 1.001
    ptr, idx, val = prob.operator().exportCSR()
    M = sp.csr_matrix((val, idx, ptr), shape=(len(ptr) - 1,) * 2)
    b = prob.rhs.cpu().numpy()
    dinv = 1.0 / M.diagonal()

    def prec(rv):
        u = dinv * rv
        s = 1.0 / np.sqrt(np.mean(u * u))
        return _bf16(s * u) / s

    itc = _cpu_fcg_iterations(M, b, prec, 1e-9)
    itj = _cpu_fcg_iterations(M, b, lambda rv: dinv * rv, 1e-9)
    print("identity-network FCG: GPU %d, CPU restatement %d, exact Jacobi-PCG %d" % (itf, itc, itj), flush=True)
    assert abs(itf - itc) <= 0.1 * itc + 8


def test_fcg_falls_back_to_jacobi_pcg():
    # an all-zero network gives a zero direction: the solver must switch to
    # the Jacobi-PCG and still converge
    ne = (12, 12, 12)
    prob = _problem(ne, 13)
    net = L.LearnedPreconditioner(prob.nodes, params=np.zeros(L.num_params(), np.float32))
    x, o = _fcg_solve(prob, net)
    assert o.hasConverged()
    assert o.getDiagnostics("fallbacks") >= 1
    assert _true_res(prob, x) <= 1e-8 * 1.001


def test_fcg_needs_matching_network():
    prob = _problem((8, 8, 8), 14)
    op = prob.operator()
    net = L.LearnedPreconditioner((9, 9, 10))
    with pytest.raises(R.RipleyHipError):
        op.setLearnedPreconditioner(net, prob.logk)
    o = R.SolverOptions()
    o.setSolverMethod(R.SolverOptions.FCG_LEARNED)
    with pytest.raises(R.RipleyHipError):
        op.solve(prob.rhs, o)


def _cos(a, b):
    return (torch.dot(a, b) / (torch.linalg.norm(a) * torch.linalg.norm(b))).item()


def test_trained_network_gains_over_untrained():
    """Train on the plain solver's residual/error pairs at 24^3 (the
    library's own forward, backward and Adam kernels) and measure on unseen
    fields what the training gains: the cosine between the network's
    correction and the true error e = A^-1 r, against the untrained network
    and against the Jacobi direction D^-1 r.  The flexible-CG iteration
    counts are reported beside it."""
    ne = (24, 24, 24)
    kfn = lambda n, s: fields.log10_normal_kappa(n, corr_len=4.0, sigma=1.0, seed=s)
    data = L.generate_training_set(ne, 8, kfn, iterations=(1, 2, 4, 8, 16, 32, 64, 128), first_seed=100)
    held = L.generate_training_set(ne, 2, kfn, iterations=(4, 32, 128), first_seed=7)
    nodes = tuple(n + 1 for n in ne)
    net = L.LearnedPreconditioner(nodes, training=True, params=L.init_params(0))
    untrained = L.LearnedPreconditioner(nodes, params=L.init_params(0))

    def mean_cos(pre):
        return float(np.mean([_cos(pre.apply(r, d, lk), e) for (r, e, d, lk) in held]))

    c0 = mean_cos(untrained)
    cj = float(np.mean([_cos(d * r, e) for (r, e, d, lk) in held]))
    hist = L.train(net, data, steps=1500, lr=2e-3, lr_final=2e-4)
    c1 = mean_cos(net)
    test = _problem(ne, 7)
    trained = L.LearnedPreconditioner(test.nodes, params=net.get_params())
    _, o0 = _fcg_solve(test, untrained)
    x1, o1 = _fcg_solve(test, trained)
    op = test.operator()
    po = R.SolverOptions()
    po.setSolverMethod(R.SolverOptions.PCG_PLAIN)
    po.setTrueResidualCriterion(True)
    op.solve(test.rhs, po)
    print("cosine(net(r), e): untrained %.3f, trained %.3f, Jacobi %.3f; loss %.3f -> %.3f; "
          "iterations: untrained FCG %d, trained FCG %d, plain Jacobi-PCG %d" %
          (c0, c1, cj, np.mean(hist[:100]), np.mean(hist[-100:]), o0.getDiagnostics("num_iter"),
           o1.getDiagnostics("num_iter"), po.getDiagnostics("num_iter")), flush=True)
    assert np.mean(hist[-100:]) < 0.5 * np.mean(hist[:100])
    assert c1 > c0 + 0.2 and c1 > cj
    assert o1.hasConverged() and _true_res(test, x1) <= 1e-8 * 1.001
