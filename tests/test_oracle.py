"""Pins the CPU oracle (oracle/ntt_oracle.c) before anything is compared to it.

Each check is independent of the oracle's own text: textbook O(n^2) DFTs in
Python integers, closed forms, the reference's schoolbook product
(tests/test_ntt.cu:186-199), the reference's pow-twiddle DIT
(tests/test_inner_ntt.cu:12-75) and constants of the reference's field rule
(include/modular_arith.cuh:130-151).
"""
import numpy as np
import pytest

BB = 2013265921  # BabyBear 15*2^27+1
P30 = 1056440321  # largest 30-bit prime = 1 mod 2^19
P62 = 4611686018326724609  # largest 62-bit prime = 1 mod 2^23
GOLD = 0xFFFFFFFF00000001  # Goldilocks 2^64 - 2^32 + 1 (sums above 2^64)


def py_dft(x, q, w):
    n = len(x)
    return [sum(int(x[j]) * pow(w, j * k, q) for j in range(n)) % q for k in range(n)]


def py_nega(x, q, psi):
    n = len(x)
    return [sum(int(x[j]) * pow(psi, j * (2 * k + 1), q) for j in range(n)) % q for k in range(n)]


def test_babybear_field_constants(orc):
    # g = 31 is the smallest generator passing the p in {2,3,5,7} checks.
    assert orc.generator(BB) == 31
    # brute force: 31 has order q-1 (q-1 = 2^27 * 3 * 5)
    for p in (2, 3, 5):
        assert pow(31, (BB - 1) // p, BB) != 1
    for g in range(2, 31):  # every smaller g fails some check
        assert any(pow(g, (BB - 1) // p, BB) == 1 for p in (2, 3, 5))
    assert orc.omega(BB, 1 << 12) == 1282623253
    assert orc.omega(BB, 1 << 20) == 195061667
    assert orc.omega(BB, 1 << 24) == 1003846038
    assert orc.psi(BB, 1 << 12) == 298008106
    assert orc.psi(BB, 1 << 20) == 414040701
    assert orc.psi(BB, 1 << 24) == 1149491290


@pytest.mark.parametrize("q", [BB, P30, P62, GOLD])
@pytest.mark.parametrize("log_n", [1, 2, 5, 8])
def test_root_orders(orc, q, log_n):
    n = 1 << log_n
    w, s = orc.omega(q, n), orc.psi(q, n)
    assert pow(w, n, q) == 1 and pow(w, n // 2, q) == q - 1
    assert pow(s, 2, q) == w and pow(s, n, q) == q - 1


def test_goldilocks_generator_and_large_values(orc):
    """The g-rule gives 7 for Goldilocks (q - 1 = 2^32 * 3 * 5 * 17 * 257 * 65537);
    inputs near q make every butterfly sum exceed 2^64."""
    assert orc.generator(GOLD) == 7
    n = 64
    x = np.array([GOLD - 1 - (j % 3) for j in range(n)], dtype=np.uint64)
    w = orc.omega(GOLD, n)
    assert list(map(int, orc.ntt(x, GOLD))) == py_dft(x, GOLD, w)


def test_params_match_create_params(orc):
    p = orc.params(BB, 4096)
    assert p.q == BB and p.q_twice == 2 * BB
    assert p.omega == 1282623253
    assert p.omega * p.omega_inv % BB == 1
    assert 4096 * p.n_inv % BB == 1
    # barrett_k: smallest k >= 64 with 2^(k-64) >= q (capped at 96); m = 2^k // q truncated to 64 bits
    k = 64
    while (1 << (k - 64)) < BB and k < 96:
        k += 1
    assert p.barrett_k == k and p.barrett_m == ((1 << k) // BB) & (2**64 - 1)
    assert orc.lib().orc_mu(BB) == (1 << 64) // BB


@pytest.mark.parametrize("q", [BB, P30, P62, GOLD])
@pytest.mark.parametrize("log_n", [1, 3, 6, 8])
def test_ntt_host_equals_textbook_dft(orc, q, log_n):
    n = 1 << log_n
    x = orc.random_poly(n, q, 100 + log_n)
    w = orc.omega(q, n)
    assert list(map(int, orc.ntt(x, q))) == py_dft(x, q, w)
    winv = pow(w, q - 2, q)
    ninv = pow(n, q - 2, q)
    assert list(map(int, orc.ntt(x, q, inverse=True))) == [v * ninv % q for v in py_dft(x, q, winv)]


@pytest.mark.parametrize("log_n", [9, 10])
def test_ntt_host_equals_c_dft(orc, log_n):
    n = 1 << log_n
    x = orc.random_poly(n, BB, 5)
    assert np.array_equal(orc.ntt(x, BB), orc.dft_naive(x, BB, orc.omega(BB, n)))


@pytest.mark.parametrize("log_n", [1, 4, 6, 11])
def test_ntt_host_equals_pow_twiddle_dit(orc, log_n):
    n = 1 << log_n
    x = orc.random_poly(n, BB, 9)
    assert np.array_equal(orc.ntt(x, BB), orc.ntt_pow_twiddle(x, BB, orc.omega(BB, n)))


@pytest.mark.parametrize("n,head", [
    (64, [2080, 945746745, 44582636, 962587333, 472683146]),
    (4096, [8390656, 552559273, 1526781732, 744143007, 921338756]),
])
def test_kat_ramp(orc, n, head):
    """x_j = j+1 (the reference's debug input, tests/debug_inner.cu:20):
    X_0 = n(n+1)/2 and X_k = n/(w^k - 1)."""
    x = np.arange(1, n + 1, dtype=np.uint64)
    X = orc.ntt(x, BB)
    assert list(map(int, X[:5])) == head
    w = orc.omega(BB, n)
    assert int(X[0]) == n * (n + 1) // 2 % BB
    for k in (1, 2, 3, n // 2, n - 1):
        assert int(X[k]) == n * pow((pow(w, k, BB) - 1) % BB, BB - 2, BB) % BB


@pytest.mark.parametrize("q", [BB, P62])
def test_kat_impulse_constant_and_pure_tone(orc, q):
    n = 256
    d = np.zeros(n, dtype=np.uint64)
    d[0] = 1
    assert np.all(orc.ntt(d, q) == 1)
    c = np.full(n, 7, dtype=np.uint64)
    X = orc.ntt(c, q)
    assert int(X[0]) == 7 * n % q and np.all(X[1:] == 0)
    w = orc.omega(q, n)
    m = 37
    tone = np.array([pow(w, (n - m) * j % n, q) for j in range(n)], dtype=np.uint64)  # w^(-jm)
    X = orc.ntt(tone, q)
    assert int(X[m]) == n % q and int(np.count_nonzero(X)) == 1


@pytest.mark.parametrize("q", [BB, P30, P62, GOLD])
@pytest.mark.parametrize("log_n", [1, 4, 7])
def test_negacyclic_definition(orc, q, log_n):
    n = 1 << log_n
    x = orc.random_poly(n, q, 77)
    psi = orc.psi(q, n)
    assert list(map(int, orc.ntt_negacyclic(x, q))) == py_nega(x, q, psi)
    assert np.array_equal(orc.ntt_negacyclic(orc.ntt_negacyclic(x, q), q, inverse=True), x)


@pytest.mark.parametrize("q", [BB, P62, GOLD])
@pytest.mark.parametrize("n", [8, 64, 256])
def test_negacyclic_product_equals_schoolbook(orc, q, n):
    """INTT_neg(NTT_neg(a) * NTT_neg(b)) == a*b mod (x^n+1) (tests/test_ntt.cu:186-199)."""
    a = orc.random_poly(n, q, 42)
    b = orc.random_poly(n, q, 123)
    ref = orc.negacyclic_schoolbook(a, b, q)
    prod = orc.pointwise_mul(orc.ntt_negacyclic(a, q), orc.ntt_negacyclic(b, q), q)
    assert np.array_equal(orc.ntt_negacyclic(prod, q, inverse=True), ref)
    # and the schoolbook itself against Python integers
    if n <= 64:
        exp = [0] * n
        for i in range(n):
            for j in range(n):
                t = int(a[i]) * int(b[j])
                if i + j >= n:
                    exp[i + j - n] -= t
                else:
                    exp[i + j] += t
        assert list(map(int, ref)) == [v % q for v in exp]


def test_cyclic_convolution(orc):
    n = 128
    a = orc.random_poly(n, BB, 1)
    b = orc.random_poly(n, BB, 2)
    prod = orc.ntt(orc.pointwise_mul(orc.ntt(a, BB), orc.ntt(b, BB), BB), BB, inverse=True)
    exp = [0] * n
    for i in range(n):
        for j in range(n):
            exp[(i + j) % n] += int(a[i]) * int(b[j])
    assert list(map(int, prod)) == [v % BB for v in exp]


def test_random_generator_range_and_determinism(orc):
    a = orc.random_poly(10000, BB, 42)
    b = orc.random_poly(10000, BB, 42)
    assert np.array_equal(a, b) and int(a.max()) < BB
    assert np.array_equal(orc.random_poly_u32(10000, BB, 42).astype(np.uint64), a)
    # first outputs of splitmix64(seed=0) (published test vector of the generator)
    s = 0
    out = []
    for _ in range(3):
        s = (s + 0x9E3779B97F4A7C15) % 2**64
        z = s
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) % 2**64
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) % 2**64
        out.append(z ^ (z >> 31))
    assert out[0] == 0xE220A8397B1DCDAF
    q = 2**61 - 1  # accept-all region covers these draws; value = draw mod q
    assert [int(v) for v in orc.random_poly(3, q, 0)] == [v % q for v in out if v < (2**64 // q) * q][:3]


def test_prime_sets(orc):
    rns = orc.ntt_primes(1 << 31, 17, 24)
    assert len(rns) == 24 and rns[0] == 2147352577 and rns[-1] == 2112225281
    assert all((p - 1) % (1 << 17) == 0 for p in rns)
    p30 = [p for p in orc.ntt_primes(1 << 30, 19, 200) if p > (1 << 29)]
    assert len(p30) == 109 and p30[0] == 1056440321
    p62 = orc.ntt_primes(1 << 62, 23, 6)
    assert p62[0] == P62 and [orc.generator(p) for p in p62] == [3, 5, 5, 3, 14, 3]
    try:
        import sympy
    except ImportError:
        return
    for p in rns[:4] + p30[:4] + p62[:2]:
        assert sympy.isprime(p)
    for c in (2147352577 + 2, 561, 1105, 3215031751):
        assert orc.is_prime(c) == sympy.isprime(c)


# ---- fast base conversion and PMAC (no reference code: pinned by CRT and
# direct Python-integer sums, independent of the oracle's modular loops)

def _rns_primes(orc, count, below=1 << 31, two_adic=17):
    return [int(v) for v in orc.ntt_primes(below, two_adic, count)]


@pytest.mark.parametrize("L,K", [(1, 1), (3, 2), (5, 4), (12, 9)])
def test_bconv_crt_identity(orc, L, K):
    primes = _rns_primes(orc, L + K)
    q, p = primes[:L], primes[L:]
    n = 40
    rng = np.random.default_rng(L * 100 + K)
    x = np.stack([rng.integers(0, qi, size=n, dtype=np.uint64) for qi in q]).astype(np.uint32)
    y = orc.bconv(x, q, p)
    Q = 1
    for qi in q:
        Q *= qi
    for c in range(n):
        xs = [int(x[i, c]) for i in range(L)]
        X = sum(xs[i] * (Q // q[i]) * pow(Q // q[i], -1, q[i]) for i in range(L)) % Q  # CRT
        t = [xs[i] * pow((Q // q[i]) % q[i], -1, q[i]) % q[i] for i in range(L)]
        S = sum(t[i] * (Q // q[i]) for i in range(L))
        assert S % Q == X and 0 <= (S - X) // Q < L
        assert [int(y[j, c]) for j in range(K)] == [S % pj for pj in p]


def test_bconv_small_values_are_exact_up_to_uq(orc):
    # every limb holds the small integer X: the output is X + u Q mod p_j for
    # one u in [0, L) shared by all j (u = 0 for X = 0)
    q = _rns_primes(orc, 4)
    p = _rns_primes(orc, 3, below=q[-1])
    Q = 1
    for qi in q:
        Q *= qi
    for X in (0, 1, 12345):
        x = np.array([[X % qi] for qi in q], dtype=np.uint32)
        y = [int(v) for v in orc.bconv(x, q, p)[:, 0]]
        us = [u for u in range(len(q)) if y == [(X + u * Q) % pj for pj in p]]
        assert len(us) == 1 and (X != 0 or us == [0])


@pytest.mark.parametrize("dnum,L,acc", [(1, 1, False), (3, 4, False), (5, 2, True)])
def test_pmac_direct_sum(orc, dnum, L, acc):
    q = _rns_primes(orc, L)
    n = 33
    rng = np.random.default_rng(dnum * 10 + L)
    a = np.stack([np.stack([rng.integers(0, qi, size=n, dtype=np.uint64) for qi in q]) for _ in range(dnum)]).astype(np.uint32)
    b = np.stack([np.stack([rng.integers(0, qi, size=n, dtype=np.uint64) for qi in q]) for _ in range(dnum)]).astype(np.uint32)
    base = np.stack([rng.integers(0, qi, size=n, dtype=np.uint64) for qi in q]).astype(np.uint32) if acc else None
    r = orc.pmac(a, b, q, base)
    for l in range(L):
        for c in range(n):
            s = (int(base[l, c]) if acc else 0) + sum(int(a[k, l, c]) * int(b[k, l, c]) for k in range(dnum))
            assert int(r[l, c]) == s % q[l]
