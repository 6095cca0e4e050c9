"""ctypes view of the C oracle (oracle/ntt_oracle.c).

Test infrastructure: imported only by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, always as the checker.  The library never calls it.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "libntt_oracle.so")
_lib = None

U64P = ctypes.POINTER(ctypes.c_uint64)
U32P = ctypes.POINTER(ctypes.c_uint32)


class Params(ctypes.Structure):
    """ModularParams, include/modular_arith.cuh:11-19."""
    _fields_ = [(n, ctypes.c_uint64) for n in
                ("q", "q_twice", "barrett_k", "barrett_m", "omega", "omega_inv", "n_inv")]


def build():
    subprocess.check_call(["make", "-s", "-C", _HERE])


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        u64 = ctypes.c_uint64
        L.orc_mod_pow.restype = u64
        L.orc_mod_pow.argtypes = [u64, u64, u64]
        L.orc_mod_inverse.restype = u64
        L.orc_mod_inverse.argtypes = [u64, u64]
        L.orc_find_generator.restype = u64
        L.orc_find_generator.argtypes = [u64]
        L.orc_primitive_root.restype = u64
        L.orc_primitive_root.argtypes = [u64, u64]
        L.orc_psi.restype = u64
        L.orc_psi.argtypes = [u64, u64]
        L.orc_mu.restype = u64
        L.orc_mu.argtypes = [u64]
        L.orc_create_params.argtypes = [u64, u64, ctypes.POINTER(Params)]
        L.orc_ntt_host.argtypes = [U64P, u64, ctypes.POINTER(Params), ctypes.c_int]
        L.orc_ntt.argtypes = [U64P, u64, u64, ctypes.c_int]
        L.orc_ntt_negacyclic.argtypes = [U64P, u64, u64, ctypes.c_int]
        L.orc_pointwise_mul.argtypes = [U64P, U64P, U64P, u64, u64]
        L.orc_negacyclic_schoolbook.argtypes = [U64P, U64P, U64P, u64, u64]
        L.orc_bconv.argtypes = [U32P, U32P, u64, U64P, ctypes.c_uint32, U64P, ctypes.c_uint32]
        L.orc_pmac.argtypes = [U32P, U32P, U32P, ctypes.c_uint32, ctypes.c_uint32, u64, U64P, ctypes.c_int]
        L.orc_dft_naive.argtypes = [U64P, U64P, u64, u64, u64]
        L.orc_ntt_pow_twiddle.argtypes = [U64P, u64, u64, u64]
        L.orc_random_poly.argtypes = [U64P, u64, u64, u64]
        L.orc_random_poly_u32.argtypes = [U32P, u64, u64, u64]
        L.orc_is_prime.restype = ctypes.c_int
        L.orc_is_prime.argtypes = [u64]
        L.orc_ntt_primes.restype = ctypes.c_int
        L.orc_ntt_primes.argtypes = [u64, ctypes.c_int, ctypes.c_int, U64P]
        L.orc_arb_mont_np.restype = ctypes.c_uint32
        L.orc_arb_mont_np.argtypes = [U32P]
        L.orc_arb_mont_mul.argtypes = [U32P, U32P, U32P, U32P, ctypes.c_uint32, ctypes.c_int]
        L.orc_arb_ntt_mont.argtypes = [U32P, U32P, ctypes.c_uint32, U32P, ctypes.c_uint32, ctypes.c_int]
        L.orc_time_negacyclic_roundtrip.restype = ctypes.c_double
        L.orc_time_negacyclic_roundtrip.argtypes = [u64, u64, u64, ctypes.c_int]
        L.orc_time_batch.restype = ctypes.c_double
        L.orc_time_batch.argtypes = [U64P, U64P, ctypes.POINTER(ctypes.c_int), ctypes.c_int, ctypes.c_int]
        L.orc_transform_batch.restype = ctypes.c_double
        L.orc_transform_batch.argtypes = [ctypes.POINTER(U64P), U64P, U64P, ctypes.POINTER(ctypes.c_int),
                                          ctypes.c_int, ctypes.c_int]
        _lib = L
    return _lib


def _p64(a):
    return a.ctypes.data_as(U64P)


def _p32(a):
    return a.ctypes.data_as(U32P)


def mod_pow(b, e, q):
    return int(lib().orc_mod_pow(b, e, q))


def generator(q):
    return int(lib().orc_find_generator(q))


def omega(q, n):
    return int(lib().orc_primitive_root(q, n))


def psi(q, n):
    return int(lib().orc_psi(q, n))


def params(q, n):
    p = Params()
    lib().orc_create_params(q, n, ctypes.byref(p))
    return p


def random_poly(n, q, seed):
    out = np.empty(n, dtype=np.uint64)
    lib().orc_random_poly(_p64(out), n, q, seed)
    return out


def random_poly_u32(n, q, seed):
    out = np.empty(n, dtype=np.uint32)
    lib().orc_random_poly_u32(_p32(out), n, q, seed)
    return out


def ntt(x, q, inverse=False):
    """ntt_host on a copy (src/ntt_basic.cu:412-463)."""
    a = np.ascontiguousarray(x, dtype=np.uint64).copy()
    lib().orc_ntt(_p64(a), a.size, q, int(inverse))
    return a


def ntt_negacyclic(x, q, inverse=False):
    a = np.ascontiguousarray(x, dtype=np.uint64).copy()
    lib().orc_ntt_negacyclic(_p64(a), a.size, q, int(inverse))
    return a


def transform(x, q, negacyclic, inverse):
    return ntt_negacyclic(x, q, inverse) if negacyclic else ntt(x, q, inverse)


def pointwise_mul(a, b, q):
    a = np.ascontiguousarray(a, dtype=np.uint64)
    b = np.ascontiguousarray(b, dtype=np.uint64)
    r = np.empty_like(a)
    lib().orc_pointwise_mul(_p64(r), _p64(a), _p64(b), a.size, q)
    return r


def negacyclic_schoolbook(a, b, q):
    a = np.ascontiguousarray(a, dtype=np.uint64)
    b = np.ascontiguousarray(b, dtype=np.uint64)
    r = np.empty_like(a)
    lib().orc_negacyclic_schoolbook(_p64(a), _p64(b), _p64(r), a.size, q)
    return r


def bconv(x, q_from, p_to):
    """Fast base conversion of limb-major x (L x n uint32) to the K moduli p_to."""
    x = np.ascontiguousarray(x, dtype=np.uint32)
    L, n = x.shape
    q = np.ascontiguousarray(q_from, dtype=np.uint64)
    p = np.ascontiguousarray(p_to, dtype=np.uint64)
    y = np.empty((p.size, n), dtype=np.uint32)
    lib().orc_bconv(x.ctypes.data_as(U32P), y.ctypes.data_as(U32P), n, _p64(q), L, _p64(p), p.size)
    return y


def pmac(a, b, q, out=None):
    """(out +) sum_k a[k] * b[k] mod q per limb; a, b: dnum x L x n uint32."""
    a = np.ascontiguousarray(a, dtype=np.uint32)
    b = np.ascontiguousarray(b, dtype=np.uint32)
    dnum, L, n = a.shape
    qq = np.ascontiguousarray(q, dtype=np.uint64)
    acc = out is not None
    r = np.ascontiguousarray(out, dtype=np.uint32).copy() if acc else np.empty((L, n), dtype=np.uint32)
    lib().orc_pmac(r.ctypes.data_as(U32P), a.ctypes.data_as(U32P), b.ctypes.data_as(U32P), dnum, L, n, _p64(qq),
                   int(acc))
    return r


def dft_naive(x, q, w):
    x = np.ascontiguousarray(x, dtype=np.uint64)
    X = np.empty_like(x)
    lib().orc_dft_naive(_p64(x), _p64(X), x.size, q, w)
    return X


def ntt_pow_twiddle(x, q, w):
    a = np.ascontiguousarray(x, dtype=np.uint64).copy()
    lib().orc_ntt_pow_twiddle(_p64(a), a.size, q, w)
    return a


def is_prime(n):
    return bool(lib().orc_is_prime(n))


def ntt_primes(below, two_adic, count):
    out = np.zeros(count, dtype=np.uint64)
    got = lib().orc_ntt_primes(below, two_adic, count, _p64(out))
    return [int(v) for v in out[:got]]


# ---- K-limb (24-bit limbs, Montgomery R = 2^(24K)) ----
LIMB_BITS = 24


def to_limbs(v, K):
    return np.array([(v >> (LIMB_BITS * i)) & ((1 << LIMB_BITS) - 1) for i in range(K)], dtype=np.uint32)


def from_limbs(a):
    return sum(int(x) << (LIMB_BITS * i) for i, x in enumerate(a))


def arb_mont_np(p_limbs):
    return int(lib().orc_arb_mont_np(_p32(np.ascontiguousarray(p_limbs, dtype=np.uint32))))


def arb_mont_mul(a, b, p, np_, K):
    out = np.zeros(K, dtype=np.uint32)
    lib().orc_arb_mont_mul(_p32(out), _p32(np.ascontiguousarray(a, dtype=np.uint32)),
                           _p32(np.ascontiguousarray(b, dtype=np.uint32)),
                           _p32(np.ascontiguousarray(p, dtype=np.uint32)), np_, K)
    return out


def arb_ntt_mont(data_m, tw_m, p_limbs, np_, K):
    """host_ntt_naive on Montgomery-form (n, K) uint32 limbs; returns a copy."""
    a = np.ascontiguousarray(data_m, dtype=np.uint32).copy()
    n = a.shape[0]
    lib().orc_arb_ntt_mont(_p32(a), _p32(np.ascontiguousarray(tw_m, dtype=np.uint32)), n,
                           _p32(np.ascontiguousarray(p_limbs, dtype=np.uint32)), np_, K)
    return a


def time_negacyclic_roundtrip(n, q, seed, reps):
    return float(lib().orc_time_negacyclic_roundtrip(n, q, seed, reps))


def time_batch(ns, qs, dirs, threads):
    ns = np.ascontiguousarray(ns, dtype=np.uint64)
    qs = np.ascontiguousarray(qs, dtype=np.uint64)
    d = (ctypes.c_int * len(dirs))(*[int(x) for x in dirs])
    return float(lib().orc_time_batch(_p64(ns), _p64(qs), d, len(dirs), threads))


def host_threads():
    """Worker threads for the oracle's batch transforms: the cores this
    process may run on, at most 16 (the CPU share a one-GPU host grants)."""
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(16, n))


def transform_batch(polys, qs, negas, invs, threads=None):
    """Oracle transforms of many polynomials (copies), one polynomial per
    worker thread; returns (results, wall seconds)."""
    outs = [np.ascontiguousarray(p, dtype=np.uint64).copy() for p in polys]
    k = len(outs)
    ptrs = (U64P * k)(*[_p64(o) for o in outs])
    ns = np.array([o.size for o in outs], dtype=np.uint64)
    qa = np.array([int(q) for q in qs], dtype=np.uint64)
    fl = (ctypes.c_int * k)(*[(1 if i else 0) | (2 if g else 0) for g, i in zip(negas, invs)])
    t = float(lib().orc_transform_batch(ptrs, _p64(ns), _p64(qa), fl, k, threads or host_threads()))
    return outs, t
