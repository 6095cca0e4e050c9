"""K-limb big-prime transforms on the GPU (ntt_c_arb_*), limb for limb
against the oracle's host_ntt_naive restatement on Montgomery data
(reference include/ntt_arbitrary.cuh:2651-2676, run_for_K's check :2870-2897),
for K = 1..12 with FieldConfig's Goldilocks and BN254 fields and random odd
primes; inverse, pointwise, Montgomery conversions and the negacyclic product
(pinned by the schoolbook product mod x^n + 1)."""
import random

import numpy as np
import pytest

import ntt_arb as A
from test_arb_oracle import _is_prime

pytestmark = pytest.mark.gpu


def _prime_with_root(K, two_adicity, seed):
    """A prime of 24K - 1 bits (23 for K=1) with 2^two_adicity | p - 1, and a
    primitive 2^two_adicity-th root of unity."""
    rng = random.Random(seed)
    bits = 24 * K - 1 if K > 1 else 23
    while True:
        c = rng.randrange(1 << (bits - two_adicity - 1), 1 << (bits - two_adicity))
        p = c * (1 << two_adicity) + 1
        if _is_prime(p):
            break
    while True:
        w = pow(rng.randrange(2, p - 1), (p - 1) >> two_adicity, p)
        if pow(w, 1 << (two_adicity - 1), p) == p - 1:
            return p, w


def _field_and_root(K, log_n, seed):
    if K in A.FIELDS and seed % 2 == 0:
        return A.Field(A.FIELDS[K][0], K), A.root_for(K, log_n)
    p, w = _prime_with_root(K, max(log_n, 14), seed)
    w = pow(w, 1 << (max(log_n, 14) - log_n), p)
    return A.Field(p, K), w


def _random_mont(f, n, seed):
    """n canonical values in [0, p) as (n, K) limbs (Montgomery form by convention)."""
    rng = np.random.default_rng(seed)
    a = rng.integers(0, 1 << 24, size=(n, f.K), dtype=np.uint32)
    top = int(f.p_limbs[-1])
    a[:, -1] = rng.integers(0, top, size=n, dtype=np.uint32)  # top limb below p's: value < p
    return a


def _tw_table(f, w, n):
    out, t = [0] * n, f.R % f.p
    for k in range(n):
        out[k] = t
        t = t * w % f.p
    return A.pack(out, f.K)


def _dev(nt, arr):
    return nt.DeviceBuffer.from_array(np.ascontiguousarray(arr))


def _down(buf, shape):
    return buf.download(np.uint32, int(np.prod(shape))).reshape(shape)


CASES = [(1, 1), (1, 10), (2, 12), (2, 13), (3, 1), (3, 2), (3, 11), (3, 12), (4, 12), (5, 11), (6, 12),
         (7, 9), (8, 12), (9, 12), (10, 11), (11, 1), (11, 10), (11, 12), (12, 11), (12, 12)]


@pytest.mark.parametrize("K,log_n", CASES)
@pytest.mark.parametrize("seed", [0, 1])
def test_forward_matches_oracle(nt, orc, K, log_n, seed):
    n = 1 << log_n
    f, w = _field_and_root(K, log_n, 7919 * K + 31 * log_n + seed)
    x = _random_mont(f, n, seed * 1000 + K * 50 + log_n)
    ntt = A.ArbNtt(f, log_n, omega=w)
    buf = _dev(nt, x)
    ntt.forward(buf)
    nt.device_sync()
    ref = orc.arb_ntt_mont(x, _tw_table(f, w, n), f.p_limbs, f.np, K)
    print(f"K={K} log_n={log_n} p={f.p:#x}", flush=True)
    np.testing.assert_array_equal(_down(buf, (n, K)), ref)


@pytest.mark.parametrize("K,log_n", [(2, 12), (3, 12), (11, 11), (12, 10)])
def test_inverse_matches_oracle(nt, orc, K, log_n):
    """inverse = host_ntt_naive with w^-1, then times n^-1 (all in Montgomery form)."""
    n = 1 << log_n
    f, w = _field_and_root(K, log_n, 0)
    x = _random_mont(f, n, 77 + K)
    ntt = A.ArbNtt(f, log_n, omega=w)
    buf = _dev(nt, x)
    ntt.inverse(buf)
    nt.device_sync()
    t = orc.arb_ntt_mont(x, _tw_table(f, pow(w, -1, f.p), n), f.p_limbs, f.np, K)
    ninv = f.limbs(pow(n, -1, f.p) * f.R)
    ref = np.stack([orc.arb_mont_mul(t[i], ninv, f.p_limbs, f.np, K) for i in range(n)])
    np.testing.assert_array_equal(_down(buf, (n, K)), ref)


@pytest.mark.parametrize("K,log_n", [(3, 18), (2, 16), (11, 14)])
def test_forward_large_matches_oracle(nt, orc, K, log_n):
    """Sizes where most stages run as global-stage kernels."""
    n = 1 << log_n
    f, w = _field_and_root(K, log_n, 0)
    x = _random_mont(f, n, 5 + K)
    ntt = A.ArbNtt(f, log_n, omega=w)
    buf = _dev(nt, x)
    ntt.forward(buf)
    nt.device_sync()
    ref = orc.arb_ntt_mont(x, _tw_table(f, w, n), f.p_limbs, f.np, K)
    np.testing.assert_array_equal(_down(buf, (n, K)), ref)


@pytest.mark.parametrize("K,log_n", [(3, 20), (2, 18), (11, 16), (12, 14)])
def test_roundtrip_large(nt, K, log_n):
    n = 1 << log_n
    f, w = _field_and_root(K, log_n, 0)
    x = _random_mont(f, n, 99 + K)
    ntt = A.ArbNtt(f, log_n, omega=w)
    buf = _dev(nt, x)
    ntt.forward(buf)
    ntt.inverse(buf)
    nt.device_sync()
    np.testing.assert_array_equal(_down(buf, (n, K)), x)


@pytest.mark.parametrize("K", [1, 3, 11, 12])
def test_pointwise_scale_and_montgomery_conversions(nt, K):
    rng = random.Random(K)
    f = A.Field(A.FIELDS[K][0], K) if K in A.FIELDS else A.Field(_prime_with_root(K, 8, K)[0], K)
    n = 1000  # not a power of two: pointwise kernels take any length
    xs = [rng.randrange(f.p) for _ in range(n)]
    ys = [rng.randrange(f.p) for _ in range(n)]
    a, b = _dev(nt, A.pack(xs, K)), _dev(nt, A.pack(ys, K))
    A.pointwise_mul(f, a, b, n)
    nt.device_sync()
    ri = pow(f.R, -1, f.p)
    assert A.unpack(_down(a, (n, K))) == [x * y * ri % f.p for x, y in zip(xs, ys)]
    c = _dev(nt, A.pack(xs, K))
    A.to_mont_device(f, c, n)
    nt.device_sync()
    assert A.unpack(_down(c, (n, K))) == f.to_mont(xs)
    A.from_mont_device(f, c, n)
    nt.device_sync()
    assert A.unpack(_down(c, (n, K))) == xs
    s = rng.randrange(f.p)
    A.scale(f, c, n, s)
    nt.device_sync()
    assert A.unpack(_down(c, (n, K))) == [x * s * ri % f.p for x in xs]


def _schoolbook_negacyclic(a, b, p):
    n = len(a)
    out = [0] * n
    for i, x in enumerate(a):
        for j, y in enumerate(b):
            if i + j < n:
                out[i + j] += x * y
            else:
                out[i + j - n] -= x * y
    return [v % p for v in out]


@pytest.mark.parametrize("K,log_n", [(3, 6), (11, 5), (2, 7)])
def test_negacyclic_product_equals_schoolbook(nt, K, log_n):
    n = 1 << log_n
    p = A.FIELDS[K][0]
    f = A.Field(p, K)
    psi = A.root_for(K, log_n + 1)
    ntt = A.ArbNtt(f, log_n, omega=psi * psi % p, psi=psi)
    rng = random.Random(K * 13 + log_n)
    xs = [rng.randrange(p) for _ in range(n)]
    ys = [rng.randrange(p) for _ in range(n)]
    a, b = _dev(nt, A.pack(f.to_mont(xs), K)), _dev(nt, A.pack(f.to_mont(ys), K))
    ntt.forward_negacyclic(a)
    ntt.forward_negacyclic(b)
    A.pointwise_mul(f, a, b, n)
    ntt.inverse_negacyclic(a)
    nt.device_sync()
    got = f.from_mont(A.unpack(_down(a, (n, K))))
    assert got == _schoolbook_negacyclic(xs, ys, p)


def test_forward_on_a_stream(nt, orc):
    K, log_n = 3, 12
    n = 1 << log_n
    f, w = _field_and_root(K, log_n, 0)
    x = _random_mont(f, n, 3)
    ntt = A.ArbNtt(f, log_n, omega=w)
    s = nt.Stream()
    buf = _dev(nt, x)
    ntt.forward(buf, stream=s)
    s.sync()
    np.testing.assert_array_equal(_down(buf, (n, K)), orc.arb_ntt_mont(x, _tw_table(f, w, n), f.p_limbs, f.np, K))
