"""ORACLE (test infrastructure only: imported by tests/ and never by the
product path).  CPU PyTorch restatement of the learned preconditioner's
network, used to check the library's BF16-MFMA forward kernels and its
hand-written backward pass.

The reference has no such network (PASO's preconditioners are Jacobi,
Gauss-Seidel, ILU0, RILU and AMG, paso/src/Preconditioner.cpp:46-107); this
module restates the network the library defines:

  grid: the node grid n0 x n1 x n2 (x fastest) embedded at the origin of a
        grid whose sides are rounded up to multiples of 4; zero padding of
        one voxel around it for every 3x3x3 convolution
  input channels: u * s (s = 1 / rms of u over the nodes, u = dinv * r),
        log10(kappa); zero outside the node grid
  layers (W[co][tap][ci], tap = (dx+1) + 3 (dy+1) + 9 (dz+1), cross-
        correlation out[v] = b + sum in[v + d] W):
        c0 16->16 relu, c1 16->16 relu = s1; avgpool2;
        c2 16->32 relu, c3 32->32 relu = s2; avgpool2;
        c4 32->64 relu, c5 64->64 relu; nearest up x2, concat [up, s2];
        c6 96->32 relu; nearest up x2, concat [up, s1];
        c7 48->16 relu; c8 16->16 linear; output channel 0
  output z = y0 / s on the nodes.
  loss: 1 - <y0, e> / (|y0| |e|) over the nodes, or the energy-norm loss
        1 - (r.z)^2 / ((r.e)(z.A z)) of z with the level's operator A.

With bf16=True every weight and every stored activation (and, in the
library's backward pass, every stored gradient) is rounded to bfloat16 as
the GPU kernels store them; accumulation is FP32 (float64 here when
dtype=float64).
"""
import numpy as np
import torch
import torch.nn.functional as F

LAYERS = ((16, 16, True), (16, 16, True), (16, 32, True), (32, 32, True), (32, 64, True), (64, 64, True),
          (96, 32, True), (48, 16, True), (16, 16, False))


def split(flat, dtype=torch.float32):
    """Flat parameters -> [(weight (co, ci, 3, 3, 3), bias (co,)), ...]."""
    flat = torch.as_tensor(np.asarray(flat), dtype=dtype)
    out, o = [], 0
    for cin, cout, _ in LAYERS:
        ws = 27 * cin * cout
        w = flat[o:o + ws].reshape(cout, 27, cin)
        # tap = kx + 3 ky + 9 kz  ->  torch (co, ci, kz, ky, kx)
        w = w.reshape(cout, 3, 3, 3, cin).permute(0, 4, 1, 2, 3).contiguous()
        out.append((w, flat[o + ws:o + ws + cout]))
        o += ws + cout
    return out


def flatten(grads):
    """Inverse of split for gradients."""
    parts = []
    for (gw, gb) in grads:
        cout, cin = gw.shape[0], gw.shape[1]
        parts.append(gw.permute(0, 2, 3, 4, 1).reshape(cout, 27, cin).reshape(-1))
        parts.append(gb.reshape(-1))
    return torch.cat(parts)


def _rb(t, on):
    return t.to(torch.bfloat16).to(t.dtype) if on else t


def up4(n):
    return (n + 3) // 4 * 4


def make_input(nodes, r, dinv, logk, dtype=torch.float32, bf16=True):
    """(1, 16, Z, Y, X) input tensor and the scale s."""
    n0, n1, n2 = nodes
    r = np.asarray(r, dtype=np.float64)
    u = r * (np.asarray(dinv, dtype=np.float64) if dinv is not None else 1.0)
    rms = np.sqrt(np.mean(u * u))
    s = 1.0 / rms if rms > 0 else 1.0
    x = torch.zeros((1, 16, up4(n2), up4(n1), up4(n0)), dtype=dtype)
    x[0, 0, :n2, :n1, :n0] = torch.as_tensor((s * u).astype(np.float32).reshape(n2, n1, n0), dtype=dtype)
    if logk is not None:
        x[0, 1, :n2, :n1, :n0] = torch.as_tensor(np.asarray(logk, dtype=np.float32).reshape(n2, n1, n0),
                                                 dtype=dtype)
    return _rb(x, bf16), s


def forward(params, x, bf16=True):
    """Network output y (1, 16, Z, Y, X) for input x; params from split()."""
    W = [(_rb(w, bf16), b) for (w, b) in params]

    def conv(l, t):
        w, b = W[l]
        o = F.conv3d(t, w, b, padding=1)
        if LAYERS[l][2]:
            o = torch.relu(o)
        return _rb(o, bf16)

    a1 = conv(0, x)
    s1 = conv(1, a1)
    p1 = _rb(F.avg_pool3d(s1, 2), bf16)
    a2 = conv(2, p1)
    s2 = conv(3, a2)
    p2 = _rb(F.avg_pool3d(s2, 2), bf16)
    a3 = conv(4, p2)
    a3b = conv(5, a3)
    cat2 = torch.cat([F.interpolate(a3b, scale_factor=2, mode="nearest"), s2], dim=1)
    a4 = conv(6, cat2)
    cat1 = torch.cat([F.interpolate(a4, scale_factor=2, mode="nearest"), s1], dim=1)
    a5 = conv(7, cat1)
    return conv(8, a5)


def apply(flat_params, nodes, r, dinv=None, logk=None, bf16=True, dtype=torch.float32):
    """z = Net(r) as numpy float64 over the nodes."""
    n0, n1, n2 = nodes
    x, s = make_input(nodes, r, dinv, logk, dtype, bf16)
    with torch.no_grad():
        y = forward(split(flat_params, dtype), x, bf16)
    return y[0, 0, :n2, :n1, :n0].double().numpy().reshape(-1) / s


def loss_and_gradients(flat_params, nodes, r, dinv, logk, e, bf16=True, dtype=torch.float32):
    """(loss, flat gradient) of the cosine loss by autograd."""
    n0, n1, n2 = nodes
    params = [(w.clone().requires_grad_(True), b.clone().requires_grad_(True)) for (w, b) in split(flat_params, dtype)]
    x, _ = make_input(nodes, r, dinv, logk, dtype, bf16)
    y = forward(params, x, bf16)[0, 0, :n2, :n1, :n0].reshape(-1).to(torch.float64)
    et = torch.as_tensor(np.asarray(e, dtype=np.float64))
    loss = 1.0 - torch.dot(y, et) / (torch.linalg.norm(y) * torch.linalg.norm(et))
    loss.backward()
    return loss.item(), flatten([(w.grad, b.grad) for (w, b) in params]).double().numpy()


def loss_and_gradients_energy(flat_params, nodes, r, dinv, logk, e, A, bf16=True, dtype=torch.float32):
    """(loss, flat gradient) of the energy-norm loss 1 - (r.z)^2/((r.e)(z.A z))
    of the output z = y / s by autograd (A: scipy sparse operator of the
    level)."""
    n0, n1, n2 = nodes
    params = [(w.clone().requires_grad_(True), b.clone().requires_grad_(True)) for (w, b) in split(flat_params, dtype)]
    x, s = make_input(nodes, r, dinv, logk, dtype, bf16)
    y = forward(params, x, bf16)[0, 0, :n2, :n1, :n0].reshape(-1).to(torch.float64)
    z = y / s
    At = torch.sparse_csr_tensor(torch.as_tensor(A.indptr, dtype=torch.int64),
                                 torch.as_tensor(A.indices, dtype=torch.int64),
                                 torch.as_tensor(A.data, dtype=torch.float64), size=A.shape)
    rt = torch.as_tensor(np.asarray(r, dtype=np.float64))
    et = torch.as_tensor(np.asarray(e, dtype=np.float64))
    az = torch.mv(At, z)
    loss = 1.0 - torch.dot(rt, z) ** 2 / (torch.dot(rt, et) * torch.dot(z, az))
    loss.backward()
    return loss.item(), flatten([(w.grad, b.grad) for (w, b) in params]).double().numpy()


def conv_direct(x, w, b):
    """Brute-force 3x3x3 zero-padded cross-correlation, numpy loops over
    taps (x: (C, Z, Y, X); w: (co, 27, ci) in the library's tap order)."""
    C, Z, Y, X = x.shape
    co = w.shape[0]
    xp = np.zeros((C, Z + 2, Y + 2, X + 2))
    xp[:, 1:-1, 1:-1, 1:-1] = x
    out = np.zeros((co, Z, Y, X)) + np.asarray(b).reshape(-1, 1, 1, 1)
    for tap in range(27):
        dx, dy, dz = tap % 3 - 1, (tap // 3) % 3 - 1, tap // 9 - 1
        sh = xp[:, 1 + dz:1 + dz + Z, 1 + dy:1 + dy + Y, 1 + dx:1 + dx + X]
        out += np.einsum("oc,czyx->ozyx", w[:, tap, :], sh)
    return out
