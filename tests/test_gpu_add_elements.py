"""addToSystemMatrix through the C ABI (rg_matrix_add_element_blocks):
host element matrices of every grid cell, added in one call, in shuffled
order, one colour per call and one element per call, against the oracle's
addToPasoMatrix restatement (ripley/src/RipleyDomain.cpp:1488-1614) in the
reference's element order.  Bars: entries within 1e-13 of the row scale;
the same input gives the same bits on every run."""
import numpy as np
import pytest

import oracle as O
import ripley_hip as R
from gpu_helpers import entrywise_ok, grid

pytestmark = pytest.mark.gpu


def _cells(dim, ne):
    """Node ids of every grid cell in local order (x fastest), cells in the
    reference's order."""
    nn = [n + 1 for n in ne]
    out = []
    if dim == 2:
        for k1 in range(ne[1]):
            for k0 in range(ne[0]):
                b = k0 + nn[0] * k1
                out.append([b, b + 1, b + nn[0], b + nn[0] + 1])
    else:
        s2 = nn[0] * nn[1]
        for k2 in range(ne[2]):
            for k1 in range(ne[1]):
                for k0 in range(ne[0]):
                    b = k0 + nn[0] * k1 + s2 * k2
                    out.append([b + dx + nn[0] * dy + s2 * dz for dz in (0, 1) for dy in (0, 1) for dx in (0, 1)][
                               :1 << dim])
    return np.asarray(out, dtype=np.int32)


def _random_em(rng, nel, npe, bs, symmetric):
    em = rng.standard_normal((nel, npe, npe, bs, bs))   # [e, b, a, m, k] -> EM[k + bs(m + bs(a + npe b))]
    if symmetric:
        # block (a, b) = block (b, a)^T: EM[k, m, a, b] = EM[m, k, b, a]
        em = 0.5 * (em + em.transpose(0, 2, 1, 4, 3))
    return em.reshape(nel, -1)


def _colour(dim, ne, nodes):
    nn0, nn1 = ne[0] + 1, ne[1] + 1
    b = nodes[:, 0].astype(np.int64)
    c = (b % nn0) % 2 + 2 * (((b // nn0) % nn1) % 2)
    if dim == 3:
        c = c + 4 * ((b // (nn0 * nn1)) % 2)
    return c


CASES = [(2, (7, 5), 1), (2, (6, 3), 3), (3, (5, 4, 3), 1), (3, (3, 4, 2), 3)]


@pytest.mark.parametrize("dim,ne,bs", CASES)
@pytest.mark.parametrize("symmetric", [False, True])
def test_add_element_blocks_against_oracle(dim, ne, bs, symmetric):
    rng = np.random.default_rng(1000 * dim + 10 * bs + symmetric)
    dom, _ = grid(dim, ne)
    nodes = _cells(dim, ne)
    npe = nodes.shape[1]
    em = _random_em(rng, len(nodes), npe, bs, symmetric)
    ptr, idx = O.pattern(dim, ne)
    oval = np.zeros(len(idx) * bs * bs)
    for e in range(len(nodes)):
        O.add_to_csr(bs, ptr, idx, oval, nodes[e], em[e])

    def run(order, per_call):
        op = dom.newOperator(bs, symmetric=symmetric)
        for k in range(0, len(order), per_call):
            sel = order[k:k + per_call]
            op.addElementBlocks(nodes[sel], em[sel])
        return op.exportCSR()[2]

    n = len(nodes)
    shuffled = rng.permutation(n)
    col = _colour(dim, ne, nodes)
    by_colour = np.concatenate([np.flatnonzero(col == c) for c in range(npe)])
    v_all = run(np.arange(n), n)
    for name, v in (("all", v_all), ("shuffled", run(shuffled, n)), ("one element per call", run(np.arange(n), 1)),
                    ("by colour", run(by_colour, int(np.bincount(col).max())))):
        ok, worst = entrywise_ok(bs, ptr, v, oval, rtol=1e-13)
        assert ok, (name, worst)
    # the same call twice gives the same bits
    assert np.array_equal(v_all, run(np.arange(n), n))


def test_add_element_blocks_accumulates_over_calls():
    # a call after an assembly adds to the values already there (addToSystemMatrix semantics)
    dim, ne, bs = 3, (4, 3, 3), 1
    dom, _ = grid(dim, ne)
    nodes = _cells(dim, ne)
    em = _random_em(np.random.default_rng(7), len(nodes), 8, bs, False)
    op = dom.newOperator(bs)
    op.addElementBlocks(nodes, em)
    op.addElementBlocks(nodes, em)
    ptr, idx = O.pattern(dim, ne)
    oval = np.zeros(len(idx))
    for e in range(len(nodes)):
        O.add_to_csr(bs, ptr, idx, oval, nodes[e], 2 * em[e])
    ok, worst = entrywise_ok(bs, ptr, op.exportCSR()[2], oval, rtol=1e-13)
    assert ok, worst


def test_add_element_blocks_argument_errors():
    dom, _ = grid(2, (3, 3))
    op = dom.newOperator(1)
    nodes = _cells(2, (3, 3))
    em = np.zeros((len(nodes), 16))
    with pytest.raises(R.RipleyHipError):
        op.addElementBlocks(nodes[:, :3], em)          # 3 nodes per element in 2-D
    bad = nodes.copy()
    bad[2, 1] = 10 ** 6
    with pytest.raises(R.RipleyHipError):
        op.addElementBlocks(bad, em)                   # node id out of range
