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

import lpc_torch as LT
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


def test_trained_network_gains_over_untrained():
    """Train on the plain solver's residual/error pairs at 24^3 and compare
    flexible-CG iteration counts on an unseen field: trained vs the same
    network untrained, and vs the plain Jacobi-PCG."""
    torch.manual_seed(0)
    ne = (24, 24, 24)
    kfn = lambda n, s: fields.log10_normal_kappa(n, corr_len=4.0, sigma=1.0, seed=s)
    data = L.generate_training_set(ne, 6, kfn, iterations=(1, 4, 16, 32, 64), first_seed=100)
    net = L.LearnedPreconditioner(tuple(n + 1 for n in ne), training=True, params=L.init_params(0))
    hist = L.train(net, data, steps=400, lr=2e-3, lr_final=2e-4)
    assert np.mean(hist[-50:]) < np.mean(hist[:50])
    test = _problem(ne, 7)
    untrained = L.LearnedPreconditioner(test.nodes, params=L.init_params(0))
    trained = L.LearnedPreconditioner(test.nodes, params=net.get_params())
    x0, o0 = _fcg_solve(test, untrained)
    x1, o1 = _fcg_solve(test, trained)
    op = test.operator()
    po = R.SolverOptions()
    po.setSolverMethod(R.SolverOptions.PCG_PLAIN)
    po.setTrueResidualCriterion(True)
    op.solve(test.rhs, po)
    it0, it1, itp = o0.getDiagnostics("num_iter"), o1.getDiagnostics("num_iter"), po.getDiagnostics("num_iter")
    print("iterations: untrained FCG %d, trained FCG %d, plain Jacobi-PCG %d" % (it0, it1, itp), flush=True)
    assert o1.hasConverged() and _true_res(test, x1) <= 1e-8 * 1.001
    assert it1 < it0
    assert it1 < itp
