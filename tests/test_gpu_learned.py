"""Learned preconditioner on the GPU (through the C ABI): forward pass and
gradients of the BF16-MFMA kernels against the CPU restatement of the
network (oracle/lpc_torch.py, which rounds weights and activations to BF16
as the kernels store them), the multilevel flexible CG with the network as
its coarse-level solver (FP64-verified residual), its fallback to
Jacobi-PCG, and what a network trained on the plain solver's residual/error
pairs gains over the untrained one and over the same hierarchy without a
network."""
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


def test_energy_loss_gradients_match_oracle():
    # the energy-norm loss with a level operator: a Galerkin operator 9^3 of
    # a 16^3 Brick, a pair with r = A e
    print("energy-loss gradients", flush=True)
    prob = _problem((16, 16, 16), 21)
    C = prob.operator().coarseOperator(1)
    nodes = (9, 9, 9)
    p = _random_params(22)
    dinv, logk = L.level_inputs(C)
    e = torch.as_tensor(np.random.default_rng(23).standard_normal(C.size), device="cuda")
    r = torch.zeros_like(e)
    C.MatrixVector(1.0, e, 0.0, r)
    net = L.LearnedPreconditioner(nodes, training=True, params=p)
    loss, g = net.gradients(r, dinv, logk, e, op=C)
    ptr, idx, val = C.exportCSR()
    import scipy.sparse as sp
    A = sp.csr_matrix((val, idx, ptr), shape=(C.size, C.size))
    lref, gref = LT.loss_and_gradients_energy(p, nodes, r.cpu().numpy(), dinv.cpu().numpy(), logk.cpu().numpy(),
                                              e.cpu().numpy(), A, bf16=True)
    print("energy loss gpu %.5f oracle %.5f, gradient rel. difference %.3e" % (loss, lref, _rel(g, gref)), flush=True)
    assert abs(loss - lref) < 2e-3 * max(1.0, abs(lref))
    assert _rel(g, gref) < 5e-2
    # the cosine loss is back without the operator
    l2, _ = net.gradients(r, dinv, logk, e)
    lc, _ = LT.loss_and_gradients(p, nodes, r.cpu().numpy(), dinv.cpu().numpy(), logk.cpu().numpy(),
                                  e.cpu().numpy(), bf16=True)
    assert abs(l2 - lc) < 2e-3


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


def _fcg_solve(prob, net, tol=1e-8, check=8, levels=0, coarse_steps=8):
    """FCG_LEARNED on a fresh operator: with a network (its grid sets the
    levels) or, with net None, the same hierarchy (`levels`) whose coarsest
    level gets `coarse_steps` Chebyshev steps."""
    op = prob.operator()
    if net is not None:
        op.setLearnedPreconditioner(net)
    else:
        op.setMultilevel(levels, 2, coarse_steps)
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
    # an untrained network as the coarse solver of a 2-level hierarchy
    ne = (16, 14, 12)
    prob = _problem(ne, 11)
    net = L.LearnedPreconditioner((9, 8, 7), params=_random_params(12))
    x, o = _fcg_solve(prob, net, tol=1e-9)
    assert o.hasConverged()
    assert _true_res(prob, x) <= 1e-9 * 1.001
    # same solution as the oracle's PASO restatement
    ptr, idx, val = prob.operator().exportCSR()
    xo, dg, _, _ = O.solve(1, ptr, idx, val, prob.rhs.cpu().numpy(), tolerance=1e-10)
    assert np.linalg.norm(x.cpu().numpy() - xo) <= 1e-6 * np.linalg.norm(xo)


@pytest.mark.parametrize("with_net", [True, False])
def test_fcg_true_residual_stop_is_one_inner_solve(with_net):
    # under the true-residual criterion the flexible CG stops on the
    # unbalanced norm of its own residual: the first outer step's inner
    # solve already meets the target and the second step only confirms it
    # (the chunk length left at its default, one iteration)
    ne = (16, 14, 12)
    prob = _problem(ne, 17, sigma=1.5)
    net = L.LearnedPreconditioner((9, 8, 7), params=_random_params(18)) if with_net else None
    x, o = _fcg_solve(prob, net, tol=1e-8, check=0, levels=2)
    assert o.hasConverged()
    assert _true_res(prob, x) <= 1e-8 * 1.001
    print("true-residual stop: network %s, %d iterations, %d outer steps, %d fallbacks"
          % (with_net, o.getDiagnostics("num_iter"), o.getDiagnostics("outer_steps"), o.getDiagnostics("fallbacks")),
          flush=True)
    if not with_net:
        assert o.getDiagnostics("fallbacks") == 0
    if o.getDiagnostics("fallbacks") == 0:   # a fallback continues with the Jacobi-PCG's own stop
        assert o.getDiagnostics("outer_steps") == 2


def test_fcg_falls_back_to_jacobi_pcg():
    # a network whose output is not finite makes the first direction break
    # down: the solver must switch to the Jacobi-PCG and still converge
    ne = (12, 12, 12)
    prob = _problem(ne, 13)
    p = np.full(L.num_params(), np.nan, np.float32)
    net = L.LearnedPreconditioner((7, 7, 7), params=p)
    x, o = _fcg_solve(prob, net)
    assert o.hasConverged()
    assert o.getDiagnostics("fallbacks") >= 1
    assert _true_res(prob, x) <= 1e-8 * 1.001


def test_fcg_needs_a_coarse_level_network():
    prob = _problem((8, 8, 8), 14)
    op = prob.operator()
    with pytest.raises(R.RipleyHipError):
        op.setLearnedPreconditioner(L.LearnedPreconditioner((9, 9, 10)))
    with pytest.raises(R.RipleyHipError):
        op.setLearnedPreconditioner(L.LearnedPreconditioner(prob.nodes))


def _energy_gain(y, r, e, op):
    """(r.y)^2 / ((r.e)(y.A y)): the fraction of the error energy e.A e that
    the energy-optimal step along y removes."""
    ay = torch.zeros_like(y)
    op.MatrixVector(1.0, y, 0.0, ay)
    return (torch.dot(r, y) ** 2 / (torch.dot(r, e) * torch.dot(y, ay))).item()


def test_trained_network_gains_over_untrained_and_over_no_network():
    """Train the network on the plain solver's residual/error pairs on
    Galerkin operators 17^3 of 150 32^3 Bricks, each after one V-cycle from that
    level (the library's own forward, backward and Adam kernels, energy-norm
    loss), then solve unseen fields with the flexible CG and the same
    hierarchy (33^3, 17^3, 9^3) with (a) the trained network correcting the
    17^3 level, (b) the untrained network, (c) no network -- the method
    without it.  The trained network must need fewer iterations than (b)
    and (c), and far fewer than plain Jacobi-PCG; the error energy its step
    removes on held-out pairs is reported beside that of the untrained
    network and of the Jacobi direction."""
    ne = (32, 32, 32)
    kfn = lambda n, s: fields.log10_normal_kappa(n, corr_len=4.0, sigma=1.0, seed=s)
    data = L.generate_training_set(ne, 150, kfn, iterations=(1, 2, 4, 8, 16, 32, 64, 128), first_seed=100)
    held = L.generate_training_set(ne, 2, kfn, iterations=(4, 32, 128), first_seed=7)
    nodes = (17, 17, 17)
    net = L.LearnedPreconditioner(nodes, training=True, params=L.init_params(0))
    untrained = L.LearnedPreconditioner(nodes, params=L.init_params(0))

    def mean_gain(f):
        return float(np.mean([_energy_gain(f(r, d, lk), r, e, op) for (r, e, d, lk, op) in held]))

    g0 = mean_gain(untrained.apply)
    gj = mean_gain(lambda r, d, lk: d * r)
    hist = L.train(net, data, steps=4000, lr=2e-3, lr_final=2e-4)
    g1 = mean_gain(net.apply)
    trained = L.LearnedPreconditioner(nodes, params=net.get_params())
    its = {}
    for seed in (7, 8):
        test = _problem(ne, seed)
        x1, o1 = _fcg_solve(test, trained)
        _, o0 = _fcg_solve(test, untrained)
        _, os_ = _fcg_solve(test, None, levels=0, coarse_steps=8)
        op = test.operator()
        po = R.SolverOptions()
        po.setSolverMethod(R.SolverOptions.PCG_PLAIN)
        po.setTrueResidualCriterion(True)
        op.solve(test.rhs, po)
        its[seed] = [o.getDiagnostics("num_iter") for o in (o1, o0, os_, po)]
        assert o1.hasConverged() and o1.getDiagnostics("fallbacks") == 0
        assert _true_res(test, x1) <= 1e-8 * 1.001
    print("error energy removed by the step along y: untrained %.3f, trained %.3f, Jacobi %.3f; loss %.3f -> %.3f; "
          "iterations (trained, untrained, Chebyshev coarse, plain Jacobi-PCG) %s" %
          (g0, g1, gj, np.mean(hist[:100]), np.mean(hist[-100:]), its), flush=True)
    assert np.mean(hist[-100:]) < 0.7 * np.mean(hist[:100])
    assert g1 > g0 + 0.1 and g1 > gj
    for t, u, c, j in its.values():
        assert t < u and t < c and 5 * t < j
