"""Writes kat_cds_spmv.json: the known-answer vectors of the reference's
stencil-matrix SpMV test (ripley/test/SystemMatrixTestCase.cpp:30-134) as
data: the diagonal offsets, the element construction parameters and the
expected y = A x for block sizes 1-4, non-symmetric and symmetric storage.

Run in a container that has the reference mounted at /root/reference:
    python tests/golden/extract_kat.py
"""
import json
import os
import re

SRC = "/root/reference/ripley/test/SystemMatrixTestCase.cpp"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "kat_cds_spmv.json")


def arrays(text):
    out = {}
    for m in re.finditer(r"const (?:double|int) (\w+)\[\]\s*=\s*\{([^}]*)\}", text):
        vals = [float(v) for v in re.findall(r"-?\d+(?:\.\d*)?(?:[eE][-+]?\d+)?", m.group(2))]
        out[m.group(1)] = vals
    return out


def main():
    text = open(SRC).read()
    a = arrays(text)
    kat = {
        "source": "ripley/test/SystemMatrixTestCase.cpp:30-134",
        "domain": "Rectangle(4,3): 20 nodes, nDOF0 = 5",
        "rows": 20,
        "diag_off": [int(v) for v in a["diag_off"]],
        "elements": "for i in 0..7: nodes (2i, 2i+1, 2i+4, 2i+5); "
                    "EM[j*bs*bs + k*bs + l] = 1000i + 50j + k + l if j % 5 == 0 "
                    "else 1000i + 50j + bs*k + l  (SystemMatrixTestCase.cpp:184-208)",
        "x": "x[i*bs + j] = i*bs + j",
        "nonsymmetric": {str(bs): a["ref_bs%d" % bs] for bs in range(1, 5)},
        "symmetric": {str(bs): a["ref_symm_bs%d" % bs] for bs in range(1, 5)},
        "tolerance_lsup": 1e-12,
    }
    with open(OUT, "w") as f:
        json.dump(kat, f, indent=1)
    print("wrote", OUT)


if __name__ == "__main__":
    main()
