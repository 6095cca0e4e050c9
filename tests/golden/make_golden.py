"""Writes tests/golden/ntt_vectors.json: small input/output vectors of the
oracle, each cross-checked against a textbook O(n^2) evaluation in Python
integers before it is written.  Run: python tests/golden/make_golden.py"""
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
from oracle import pyoracle as orc  # noqa: E402

BB = 2013265921
P30 = 1056440321
P31 = 2147352577
P62 = 4611686018326724609


def textbook(x, q, nega, inverse):
    n = len(x)
    g = orc.generator(q)
    w = pow(g, (q - 1) // n, q)
    psi = pow(g, (q - 1) // (2 * n), q)
    if not inverse:
        if nega:
            return [sum(x[j] * pow(psi, j * (2 * k + 1), q) for j in range(n)) % q for k in range(n)]
        return [sum(x[j] * pow(w, j * k, q) for j in range(n)) % q for k in range(n)]
    ninv = pow(n, q - 2, q)
    if nega:
        pinv = pow(psi, q - 2, q)
        return [ninv * sum(x[k] * pow(pinv, j * (2 * k + 1), q) for k in range(n)) % q for j in range(n)]
    winv = pow(w, q - 2, q)
    return [ninv * sum(x[k] * pow(winv, j * k, q) for k in range(n)) % q for j in range(n)]


def main():
    cases = []
    specs = [("ramp", BB, 64, 0, 0), ("random", BB, 256, 0, 0), ("random", BB, 256, 0, 1),
             ("random", BB, 256, 1, 0), ("random", BB, 256, 1, 1), ("random", P30, 128, 1, 0),
             ("random", P30, 128, 1, 1), ("random", P31, 32, 1, 0), ("random", P62, 64, 0, 0),
             ("random", P62, 64, 1, 0), ("random", P62, 64, 1, 1), ("random", BB, 2, 0, 0),
             ("random", BB, 1024, 0, 0)]
    for kind, q, n, nega, inv in specs:
        x = list(range(1, n + 1)) if kind == "ramp" else [int(v) for v in orc.random_poly(n, q, 42 + n)]
        import numpy as np
        y = [int(v) for v in orc.transform(np.array(x, dtype=np.uint64), q, bool(nega), bool(inv))]
        if n <= 256:
            assert y == textbook(x, q, nega, inv), (q, n, nega, inv)
        else:
            assert y == [int(v) for v in orc.dft_naive(np.array(x, dtype=np.uint64), q, orc.omega(q, n))]
        cases.append({"q": q, "n": n, "negacyclic": nega, "inverse": inv, "input": x, "output": y})
    # negacyclic product (tests/test_ntt.cu:186-199 semantics)
    import numpy as np
    n = 64
    a = orc.random_poly(n, BB, 42)
    b = orc.random_poly(n, BB, 123)
    r = orc.negacyclic_schoolbook(a, b, BB)
    prods = {"q": BB, "n": n, "a": [int(v) for v in a], "b": [int(v) for v in b], "product": [int(v) for v in r]}
    with open(os.path.join(HERE, "ntt_vectors.json"), "w") as f:
        json.dump({"generator": "oracle/ntt_oracle.c via tests/golden/make_golden.py",
                   "transforms": cases, "negacyclic_products": [prods]}, f)


if __name__ == "__main__":
    main()
