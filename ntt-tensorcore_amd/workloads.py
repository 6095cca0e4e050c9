"""Workload definitions and the portable input generator.

The generator is splitmix64 (Steele, Lea, Flood 2014) with rejection into
[0, q): accept x < floor(2^64/q)*q, value x mod q.  It is vectorised here with
numpy and gives the same stream on every host, so GPU inputs and CPU-baseline
inputs are identical for a seed.  (The reference's tests draw through
libstdc++'s uniform_int_distribution, which is implementation-defined.)

Workloads (BabyBear q = 15*2^27 + 1 unless stated):
  * cyclic_4096      : one polynomial, n = 2^12, forward + inverse.
  * negacyclic_large : one polynomial, n = 2^20 or 2^24, forward then inverse.
  * rns              : n = 2^16 negacyclic, 24 limbs, the 24 largest primes
                       q < 2^31 with q = 1 mod 2^17.
  * uneven_batch     : 512 polynomials, log2 n uniform in {10..18}, prime
                       BabyBear or one of the 30-bit primes q = 1 mod 2^19,
                       direction random.
  * serving stream   : Poisson arrivals, 1..32 polynomials per request,
                       log2 n log-uniform in [12, 22], BabyBear or a 62-bit
                       prime q = 1 mod 2^23.
"""
import numpy as np

BABYBEAR = 2013265921
_GOLDEN = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


def _splitmix_block(seed, start, count):
    i = np.arange(start + 1, start + count + 1, dtype=np.uint64)
    with np.errstate(over="ignore"):
        z = np.uint64(seed) + i * _GOLDEN
        z = (z ^ (z >> np.uint64(30))) * _M1
        z = (z ^ (z >> np.uint64(27))) * _M2
    return z ^ (z >> np.uint64(31))


def random_poly(n, q, seed, dtype=np.uint64):
    """n values uniform in [0, q) from splitmix64(seed), rejection sampled."""
    lim = ((1 << 64) // q) * q
    out = np.empty(0, dtype=np.uint64)
    start = 0
    while out.size < n:
        want = n - out.size
        z = _splitmix_block(seed, start, want + 16)
        start += want + 16
        if lim < (1 << 64):
            z = z[z < np.uint64(lim)]
        out = np.concatenate([out, z[:want]])
    return (out % np.uint64(q)).astype(dtype)


def _is_prime(n):
    if n < 2:
        return False
    for p in (2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37):
        if n % p == 0:
            return n == p
    d, r = n - 1, 0
    while d % 2 == 0:
        d //= 2
        r += 1
    for a in (2, 325, 9375, 28178, 450775, 9780504, 1795265022):
        a %= n
        if a == 0:
            continue
        x = pow(a, d, n)
        if x in (1, n - 1):
            continue
        for _ in range(r - 1):
            x = x * x % n
            if x == n - 1:
                break
        else:
            return False
    return True


def ntt_primes(below, two_adic, count, above=0):
    """The `count` largest primes q < below with q = 1 mod 2^two_adic, q > above."""
    step = 1 << two_adic
    q = ((below - 1) // step) * step + 1
    if q >= below:
        q -= step
    out = []
    while len(out) < count and q > max(step, above):
        if _is_prime(q):
            out.append(q)
        q -= step
    return out


def rns_primes():
    return ntt_primes(1 << 31, 17, 24)


def p30_primes():
    return ntt_primes(1 << 30, 19, 1000, above=1 << 29)


def p62_primes(count=6):
    return ntt_primes(1 << 62, 23, count)


def uneven_batch(count=512, seed=4, log_min=10, log_max=18):
    """Per polynomial: log2 n ~ U{log_min..log_max}; prime BabyBear w.p. 1/2
    else uniform over the 30-bit primes q = 1 mod 2^19; direction random;
    negacyclic (both prime sets have 2n | q-1 for n <= 2^18)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    primes = p30_primes()
    jobs = []
    for i in range(count):
        log_n = int(rng.integers(log_min, log_max + 1))
        q = BABYBEAR if rng.random() < 0.5 else int(primes[int(rng.integers(0, len(primes)))])
        inv = bool(rng.random() < 0.5)
        jobs.append({"log_n": log_n, "q": q, "inverse": inv, "nega": True, "seed": seed * 100003 + i})
    return jobs


def serving_requests(count, qps, seed=5, log_min=12, log_max=22, max_polys=32, p62_share=0.5):
    """Poisson arrivals at `qps`: list of (arrival_s, [poly jobs])."""
    rng = np.random.Generator(np.random.PCG64(seed))
    p62 = p62_primes()
    t = 0.0
    reqs = []
    for r in range(count):
        t += float(rng.exponential(1.0 / qps))
        npoly = int(rng.integers(1, max_polys + 1))
        log_n = int(round(float(rng.uniform(log_min, log_max))))
        q = int(p62[int(rng.integers(0, len(p62)))]) if rng.random() < p62_share else BABYBEAR
        inv = bool(rng.random() < 0.5)
        polys = [{"log_n": log_n, "q": q, "inverse": inv, "nega": True, "seed": (seed << 32) + r * 64 + k}
                 for k in range(npoly)]
        reqs.append((t, polys))
    return reqs


def butterflies(log_n):
    return (1 << log_n) // 2 * log_n
