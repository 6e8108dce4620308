"""Pins the CPU restatement of the learned preconditioner's network
(oracle/lpc_torch.py): tap order and weight layout against a brute-force
convolution, pooling/upsampling/concatenation against explicit loops,
autograd gradients against central finite differences (float64, no BF16
rounding), and the input/output scaling."""
import numpy as np
import pytest
import torch

import lpc_torch as L


def _params(seed, scale=1.0):
    rng = np.random.default_rng(seed)
    parts = []
    for cin, cout, _ in L.LAYERS:
        parts.append(rng.standard_normal(27 * cin * cout) * scale / np.sqrt(27 * cin))
        parts.append(rng.standard_normal(cout) * 0.1)
    return np.concatenate(parts)


def test_param_count():
    assert sum(27 * ci * co + co for ci, co, _ in L.LAYERS) == 332064


@pytest.mark.parametrize("layer", [0, 2, 6, 8])
def test_conv_matches_brute_force(layer):
    cin, cout, _ = L.LAYERS[layer]
    rng = np.random.default_rng(layer)
    flat = _params(11)
    w, b = L.split(flat, torch.float64)[layer]
    x = rng.standard_normal((cin, 4, 5, 6))
    got = torch.nn.functional.conv3d(torch.as_tensor(x)[None], w, b, padding=1)[0].numpy()
    # the library's layout: W[co][tap][ci]
    o = sum(27 * ci * co + co for ci, co, _ in L.LAYERS[:layer])
    wl = flat[o:o + 27 * cin * cout].reshape(cout, 27, cin)
    ref = L.conv_direct(x, wl, flat[o + 27 * cin * cout:o + 27 * cin * cout + cout])
    np.testing.assert_allclose(got, ref, rtol=1e-12, atol=1e-12)


def test_tap_order_single_tap():
    # a weight only on tap (dx, dy, dz) = (+1, 0, -1) reads in[x+1, y, z-1]
    tap = (1 + 1) + 3 * (0 + 1) + 9 * (-1 + 1)
    w = np.zeros((1, 27, 1))
    w[0, tap, 0] = 1.0
    x = np.arange(2 * 3 * 4, dtype=np.float64).reshape(1, 2, 3, 4)   # (C, Z, Y, X)
    out = L.conv_direct(x, w, [0.0])
    assert out[0, 1, 1, 1] == x[0, 0, 1, 2]
    assert out[0, 0, 1, 1] == 0.0          # z-1 outside


def test_flatten_inverts_split():
    flat = _params(3)
    back = L.flatten(L.split(flat, torch.float64)).numpy()
    np.testing.assert_array_equal(back, flat)


def test_pool_up_concat_semantics():
    t = torch.arange(2 * 4 * 4 * 4, dtype=torch.float64).reshape(1, 2, 4, 4, 4)
    p = torch.nn.functional.avg_pool3d(t, 2)
    for c in range(2):
        for Z in range(2):
            for Y in range(2):
                for X in range(2):
                    blk = t[0, c, 2 * Z:2 * Z + 2, 2 * Y:2 * Y + 2, 2 * X:2 * X + 2]
                    assert abs(float(p[0, c, Z, Y, X]) - float(blk.mean())) < 1e-12
    u = torch.nn.functional.interpolate(p, scale_factor=2, mode="nearest")
    for z in range(4):
        for y in range(4):
            for x in range(4):
                assert float(u[0, 1, z, y, x]) == float(p[0, 1, z // 2, y // 2, x // 2])


def test_input_scaling_and_padding():
    nodes = (5, 6, 7)
    rng = np.random.default_rng(1)
    n = 5 * 6 * 7
    r, d, lk = rng.standard_normal(n), rng.uniform(0.5, 2, n), rng.standard_normal(n)
    x, s = L.make_input(nodes, r, d, lk, torch.float64, bf16=False)
    assert x.shape == (1, 16, 8, 8, 8)
    u = r * d
    assert abs(s - 1.0 / np.sqrt(np.mean(u * u))) < 1e-12
    np.testing.assert_allclose(x[0, 0, :7, :6, :5].numpy().reshape(-1), (s * u).astype(np.float32), rtol=0)
    assert float(x[0, :, 7].abs().sum()) == 0.0 and float(x[0, :, :, :, 5:].abs().sum()) == 0.0
    assert float(x[0, 2:].abs().sum()) == 0.0


def test_output_scale_equivariance():
    # the output is linear in the residual's scale (input is normalised)
    nodes = (4, 4, 4)
    rng = np.random.default_rng(2)
    r = rng.standard_normal(64)
    flat = _params(5)
    z1 = L.apply(flat, nodes, r, bf16=False, dtype=torch.float64)
    z2 = L.apply(flat, nodes, 8.0 * r, bf16=False, dtype=torch.float64)
    np.testing.assert_allclose(z2, 8.0 * z1, rtol=1e-12)


def test_gradients_match_finite_differences():
    nodes = (4, 5, 4)
    n = 80
    rng = np.random.default_rng(7)
    r, d, lk, e = rng.standard_normal(n), rng.uniform(0.5, 2, n), rng.standard_normal(n), rng.standard_normal(n)
    flat = _params(9, scale=1.5)
    loss, g = L.loss_and_gradients(flat, nodes, r, d, lk, e, bf16=False, dtype=torch.float64)

    def f(p):
        y = L.apply(p, nodes, r, d, lk, bf16=False, dtype=torch.float64)
        return 1.0 - y @ e / (np.linalg.norm(y) * np.linalg.norm(e))

    assert abs(f(flat) - loss) < 1e-12
    # coordinates in every layer's weights and biases
    offs, o = [], 0
    for cin, cout, _ in L.LAYERS:
        ws = 27 * cin * cout
        offs += [o + int(rng.integers(ws)), o + ws + int(rng.integers(cout))]
        o += ws + cout
    h = 1e-6
    for k in offs:
        p1, p2 = flat.copy(), flat.copy()
        p1[k] += h
        p2[k] -= h
        fd = (f(p1) - f(p2)) / (2 * h)
        assert abs(fd - g[k]) <= 1e-6 + 1e-4 * abs(fd), (k, fd, g[k])
