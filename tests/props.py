"""Size-independent checks of a transform's output, for sizes where running
the whole oracle transform would dominate the test time.

Every output of the four transforms is a defining sum
    out[k] = s_k * sum_j x_j * r_k^j  (mod q)
with
    cyclic forward      r_k = w^k,            s_k = 1
    cyclic inverse      r_k = w^-k,           s_k = n^-1
    negacyclic forward  r_k = psi^(2k+1),     s_k = 1
    negacyclic inverse  r_k = w^-k,           s_k = n^-1 psi^-k
(w = g^((q-1)/n), psi = g^((q-1)/2n), g by the reference's generator rule,
include/modular_arith.cuh:130-151; the negacyclic inverse untwists its
output: a_k = psi^-k n^-1 sum_j A_j w^-jk).  `sampled_sums_ok` evaluates a few
of them directly in numpy; q < 2^31 keeps every product below 2^62.
"""
import numpy as np


def powers(r, n, q):
    """r^j mod q for j < n (numpy, by doubling)."""
    p = np.ones(n, dtype=np.uint64)
    m, rm = 1, r % q
    while m < n:
        k = min(m, n - m)
        p[m:m + k] = (p[:k] * np.uint64(rm)) % np.uint64(q)
        rm = rm * rm % q
        m *= 2
    return p


def defining_sum(x, q, g, k, nega, inv):
    n = x.size
    w = pow(g, (q - 1) // n, q)
    psi = pow(g, (q - 1) // (2 * n), q)
    if not inv:
        r, s = (pow(psi, 2 * k + 1, q) if nega else pow(w, k, q)), 1
    else:
        r = pow(pow(w, k, q), q - 2, q)
        s = pow(n, q - 2, q)
        if nega:
            s = s * pow(pow(psi, k, q), q - 2, q) % q
    p = powers(r, n, q)
    t = int(((np.asarray(x, dtype=np.uint64) * p) % np.uint64(q)).sum(dtype=np.uint64) % np.uint64(q))
    return t * s % q


def sampled_sums_ok(got, x, q, g, nega, inv, ks=None):
    """True when got[k] equals the defining sum at the sampled k (q < 2^31)."""
    assert q < (1 << 31)
    n = x.size
    ks = ks if ks is not None else (0, 1, 5, n // 3, n // 2 + 7, n - 1)
    for k in ks:
        if int(got[k]) != defining_sum(x, q, g, k, nega, inv):
            return False
    return True
