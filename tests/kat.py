"""Construction of the reference's stencil-matrix known-answer test
(ripley/test/SystemMatrixTestCase.cpp:169-213) shared by CPU and GPU tests."""
import json
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
KAT = json.load(open(os.path.join(HERE, "golden", "kat_cds_spmv.json")))


def kat_elements(bs):
    """(nodes (8,4) int32, em (8, 4*4*bs*bs) float64) in the element-matrix
    layout EM[k + bs*(m + bs*(a + 4*b))] used by add()."""
    nodes = np.zeros((8, 4), dtype=np.int32)
    em = np.zeros((8, 16 * bs * bs))
    for i in range(8):
        nodes[i] = (2 * i, 2 * i + 1, 2 * i + 4, 2 * i + 5)
        for j in range(16):
            for k in range(bs):
                for l in range(bs):
                    em[i, j * bs * bs + k * bs + l] = (1000 * i + 50 * j + k + l if j % 5 == 0
                                                       else 1000 * i + 50 * j + bs * k + l)
    return nodes, em


def kat_x(bs):
    return np.arange(KAT["rows"] * bs, dtype=np.float64)


def kat_expected(bs, symmetric):
    return np.asarray(KAT["symmetric" if symmetric else "nonsymmetric"][str(bs)])
