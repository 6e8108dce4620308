"""Shared builders for the GPU parity tests: the same seeded problem on the
HIP path (through the C ABI) and on the oracle."""
import numpy as np
import torch

import oracle as O
import ripley_hip as R


def grid(dim, ne, length=None):
    length = length or tuple(1.0 for _ in range(dim))
    return (R.Rectangle(*ne, *length) if dim == 2 else R.Brick(*ne, *length)), length


def to_dev(a):
    return torch.as_tensor(np.ascontiguousarray(a, dtype=np.float64), device="cuda")


def gpu_assemble(dom, n, symmetric=False, reduced=False, iso=None, **coefs):
    """Assemble on the GPU; coefs are numpy arrays in logical shape."""
    op = dom.newOperator(n, symmetric=symmetric)
    E, Q = dom.getNumElements(), dom.getNumQuadraturePoints()
    dim = dom.getDim()
    if n == 1:
        shapes = dict(A=(dim, dim), B=(dim,), C=(dim,), D=(), X=(dim,), Y=())
    else:
        shapes = dict(A=(n, dim, n, dim), B=(n, dim, n), C=(n, n, dim), D=(n, n), X=(n, dim), Y=(n,))
    cs = {}
    for k, v in coefs.items():
        if v is None:
            continue
        cs[k] = R.Coefficient(to_dev(v), shapes[k], E, Q, dom.device)
    if iso is not None:
        cs["A"] = R.Coefficient(to_dev(iso), shapes["A"], E, Q, dom.device, isotropic=True)
    rhs = torch.zeros(op.size, dtype=torch.float64, device="cuda")
    op.assemble(cs, rhs, reduced=reduced)
    return op, rhs


def row_scale(bs, ptr, val):
    """max |a_ij| of each scalar row of a PASO CSR matrix (per block entry)."""
    N = len(ptr) - 1
    blk = np.abs(val.reshape(-1, bs, bs, order="C"))  # [k, c, r] since val[k*bs*bs + r + bs*c]
    out = np.zeros(N * bs)
    for i in range(N):
        sl = blk[ptr[i]:ptr[i + 1]]
        if len(sl):
            out[i * bs:(i + 1) * bs] = sl.max(axis=(0, 1))
    return out


def entrywise_ok(bs, ptr, val_gpu, val_ref, rtol=1e-13):
    """|d a_ij| <= rtol * max_j |a_ij| for every entry (scalar rows)."""
    scale = row_scale(bs, ptr, val_ref)
    N = len(ptr) - 1
    d = np.abs(val_gpu - val_ref).reshape(-1, bs, bs)
    worst = 0.0
    for i in range(N):
        sl = d[ptr[i]:ptr[i + 1]]
        if len(sl):
            s = scale[i * bs:(i + 1) * bs]
            worst = max(worst, float((sl.max(axis=1) / np.maximum(s, 1e-300)).max()))
    return worst <= rtol, worst


def dirichlet_mask(dom, axis, n=1):
    q = dom.whereZero(axis).cpu().numpy()
    return np.repeat(q, n)
