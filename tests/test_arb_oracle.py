"""The K-limb oracle (oracle/ntt_oracle.c orc_arb_*) pinned against Python
integers, which are independent of its limb arithmetic: Montgomery products
(reference include/ntt_arbitrary.cuh:456-474,1218-1298) against a*b*R^-1 mod p,
and host_ntt_naive on Montgomery data (:2651-2676) against the O(n^2) DFT.
The field constants of FieldConfig<3>/<11> (:515-568) are checked by their
root orders.  Also the host side of ntt_arb.py and the argument checks of the
ntt_c_arb_* entry points, which return before any device call."""
import ctypes
import random

import numpy as np
import pytest

import ntt_arb as A


def _is_prime(n, rounds=24):
    if n < 2:
        return False
    for sp in (2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37):
        if n % sp == 0:
            return n == sp
    d, s = n - 1, 0
    while d % 2 == 0:
        d //= 2
        s += 1
    rng = random.Random(n)
    for _ in range(rounds):
        a = rng.randrange(2, n - 1)
        x = pow(a, d, n)
        if x in (1, n - 1):
            continue
        for _ in range(s - 1):
            x = x * x % n
            if x == n - 1:
                break
        else:
            return False
    return True


def test_miller_rabin_helper():
    assert [p for p in range(2, 60) if _is_prime(p)] == [2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37, 41, 43, 47, 53, 59]
    assert not _is_prime(561) and not _is_prime(1543503873) and _is_prime(2013265921)


def test_field_constants():
    # Goldilocks: w_(2^24) of FieldConfig<3> has order exactly 2^24
    p, w, m = A.FIELDS[3]
    assert p == 2**64 - 2**32 + 1 and _is_prime(p)
    assert pow(w, 1 << 23, p) == p - 1 and pow(w, 1 << 24, p) == 1
    # BN254: the 2^28-th root has order exactly 2^28, and p is prime
    assert _is_prime(A.BN254) and (A.BN254 - 1) % (1 << 28) == 0
    assert pow(A.BN254_ROOT_2_28, 1 << 27, A.BN254) == A.BN254 - 1
    p, w, m = A.FIELDS[11]
    assert pow(w, 1 << 23, p) == p - 1
    # K=2: the reference's constant does not fit 48 bits and its truncation is composite
    assert 281476520214529 >> 48 == 1 and 281476520214529 & ((1 << 48) - 1) == 1543503873
    assert not _is_prime(1543503873)
    p, w, m = A.FIELDS[2]
    assert p < 2**48 and _is_prime(p) and (p - 1) % (1 << 24) == 0
    assert pow(w, 1 << 23, p) == p - 1
    for K, (p, w, m) in A.FIELDS.items():
        assert A.limbs_for(p) == K


def _rand_odd_modulus(rng, K):
    return rng.randrange(1 << (24 * K - 1), 1 << (24 * K)) | 1


@pytest.mark.parametrize("K", range(1, 13))
def test_mont_mul_matches_integers(orc, K):
    rng = random.Random(1000 + K)
    for trial in range(40):
        p = _rand_odd_modulus(rng, K) if trial % 2 else A.FIELDS.get(K, (_rand_odd_modulus(rng, K),))[0]
        f = A.Field(p, K)
        assert orc.arb_mont_np(f.p_limbs) == f.np
        assert (f.np * p) % (1 << 24) == (1 << 24) - 1
        ri = pow(f.R, -1, p)
        for a, b in [(rng.randrange(p), rng.randrange(p)), (p - 1, p - 1), (0, rng.randrange(p)), (1, 1)]:
            got = A.from_limbs(orc.arb_mont_mul(f.limbs(a), f.limbs(b), f.p_limbs, f.np, K))
            assert got == a * b * ri % p


def _dft(xs, w, p):
    n = len(xs)
    return [sum(x * pow(w, j * k, p) for j, x in enumerate(xs)) % p for k in range(n)]


@pytest.mark.parametrize("K,log_n", [(2, 3), (2, 6), (3, 1), (3, 5), (3, 7), (11, 4), (11, 6)])
def test_host_ntt_naive_on_montgomery_data_equals_dft(orc, K, log_n):
    p, _, _ = A.FIELDS[K]
    f = A.Field(p, K)
    n = 1 << log_n
    w = A.root_for(K, log_n)
    rng = random.Random(K * 100 + log_n)
    xs = [rng.randrange(p) for _ in range(n)]
    tw = A.pack([pow(w, k, p) * f.R % p for k in range(n)], K)
    got = orc.arb_ntt_mont(A.pack(f.to_mont(xs), K), tw, f.p_limbs, f.np, K)
    assert f.from_mont(A.unpack(got)) == _dft(xs, w, p)


def test_host_ntt_naive_odd_modulus_ramp(orc):
    """x_j = j + 1 over Goldilocks: X_0 = n(n+1)/2 and X_k = n / (w^k - 1)."""
    K, log_n = 3, 6
    p = A.GOLDILOCKS
    f, n = A.Field(p, K), 1 << log_n
    w = A.root_for(K, log_n)
    tw = A.pack([pow(w, k, p) * f.R % p for k in range(n)], K)
    got = f.from_mont(A.unpack(orc.arb_ntt_mont(A.pack(f.to_mont(range(1, n + 1)), K), tw, f.p_limbs, f.np, K)))
    assert got[0] == n * (n + 1) // 2
    for k in range(1, n):
        assert got[k] == n * pow(pow(w, k, p) - 1, -1, p) % p


def test_pack_unpack_roundtrip():
    rng = random.Random(5)
    for K in (1, 3, 11, 12):
        vals = [rng.randrange(1 << (24 * K)) for _ in range(17)]
        arr = A.pack(vals, K)
        assert arr.dtype == np.uint32 and arr.shape == (17, K) and int(arr.max()) < (1 << 24)
        assert A.unpack(arr) == vals
        assert A.from_limbs(A.to_limbs(vals[0], K)) == vals[0]


def _u32(a):
    return np.ascontiguousarray(a, dtype=np.uint32).ctypes.data_as(ctypes.POINTER(ctypes.c_uint32))


def test_arb_entry_points_reject_bad_arguments(nt):
    L = nt.lib()
    f = A.Field(A.GOLDILOCKS, 3)
    pl, ninv = f.p_limbs, f.limbs(1)
    fake = ctypes.c_void_p(0x1000)  # never dereferenced: every call below fails validation first
    bad = [
        L.ntt_c_arb_forward(None, fake, 16, 4, _u32(pl), f.np, 3, None),            # null data
        L.ntt_c_arb_forward(fake, None, 16, 4, _u32(pl), f.np, 3, None),            # null table
        L.ntt_c_arb_forward(fake, fake, 16, 5, _u32(pl), f.np, 3, None),            # n != 2^log_n
        L.ntt_c_arb_forward(fake, fake, 24, 4, _u32(pl), f.np, 3, None),            # not a power of two
        L.ntt_c_arb_forward(fake, fake, 1, 0, _u32(pl), f.np, 3, None),             # log_n = 0
        L.ntt_c_arb_forward(fake, fake, 16, 4, _u32(pl), f.np, 0, None),            # K = 0
        L.ntt_c_arb_forward(fake, fake, 16, 4, _u32(pl), f.np, 13, None),           # K = 13
        L.ntt_c_arb_forward(fake, fake, 16, 4, _u32(pl), f.np ^ 2, 3, None),        # wrong np
        L.ntt_c_arb_forward(fake, fake, 16, 4, _u32([2, 0, 1]), f.np, 3, None),     # even modulus
        L.ntt_c_arb_forward(fake, fake, 16, 4, _u32([1, 1 << 24, 1]), f.np, 3, None),  # limb > 24 bits
        L.ntt_c_arb_forward(fake, fake, 16, 4, _u32([1, 0, 0]), f.np, 3, None),     # zero top limb
        L.ntt_c_arb_inverse(fake, fake, 16, 4, _u32(pl), f.np, None, 3, None),      # null n^-1
        L.ntt_c_arb_pointwise_mul(fake, None, 16, _u32(pl), f.np, 3, None),         # null operand
        L.ntt_c_arb_scale(fake, 16, None, _u32(pl), f.np, 3, None),                 # null scalar
        L.ntt_c_arb_scale(fake, 16, _u32([1 << 24, 0, 0]), _u32(pl), f.np, 3, None),  # scalar limb too wide
    ]
    assert all(s == 1 for s in bad), bad
    assert L.ntt_c_arb_pointwise_mul(fake, fake, 0, _u32(pl), f.np, 3, None) == 0  # empty: no-op
    assert b"arb" in L.ntt_c_last_error()


def test_arbntt_rejects_bad_roots():
    with pytest.raises(ValueError):
        A.root_for(3, 25)
    f = A.Field(A.GOLDILOCKS)
    with pytest.raises(ValueError):
        A.ArbNtt(f, 4, omega=A.root_for(3, 3))  # order 8, not 16
