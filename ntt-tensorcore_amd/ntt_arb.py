"""K-limb big-prime NTT: Python host side of include/ntt_arbitrary.cuh.

Values travel as K uint32 limbs of 24 bits, little-endian, in Montgomery form
x*R mod p with R = 2^(24K) (DScalar<K>, reference include/ntt_arbitrary.cuh:51,
573-576).  Host arithmetic uses Python integers; the transforms run in
libntt_mi355x.so (ntt_c_arb_*), never here.

Field constants follow FieldConfig<K> (reference :479-568):
  K=3  Goldilocks p = 2^64 - 2^32 + 1, w_(2^24) = 0x86cdcc31c307e171 (:515-534)
  K=11 BN254 scalar field, the 2^28-th root squared down to 2^24 (:549-568)
  K=2  the reference's 49-bit prime 281476520214529 does not fit two 24-bit
       limbs and is truncated there to 1543503873 = 3*3407*151013 (:488-507);
       this build uses the 48-bit prime 0xffffd8000001 = 16777176*2^24 + 1
       (generator 5) instead.
"""
import ctypes

import numpy as np

import ntt_amd as _n

LIMB_BITS = 24
LIMB_MASK = (1 << LIMB_BITS) - 1
K_MAX = 12

GOLDILOCKS = 0xFFFFFFFF00000001
BN254 = 0x30644E72E131A029B85045B68181585D2833E84879B9709143E1F593F0000001
BN254_ROOT_2_28 = 0x2A3C09F0A58A7E8500E0A7EB8EF62ABC402D111E41112ED49BD61B6E725B19F0
P48 = 0xFFFFD8000001

# K -> (p, primitive 2^max_log_n-th root, max_log_n)
FIELDS = {
    2: (P48, pow(5, (P48 - 1) >> 24, P48), 24),
    3: (GOLDILOCKS, 0x86CDCC31C307E171, 24),
    11: (BN254, pow(BN254_ROOT_2_28, 1 << 4, BN254), 24),
}


def to_limbs(x, K):
    """Integer -> K little-endian 24-bit limbs (uint32)."""
    out = np.empty(K, dtype=np.uint32)
    for j in range(K):
        out[j] = x & LIMB_MASK
        x >>= LIMB_BITS
    if x:
        raise ValueError("value does not fit K limbs")
    return out


def from_limbs(limbs):
    v = 0
    for j in reversed(range(len(limbs))):
        v = (v << LIMB_BITS) | int(limbs[j])
    return v


def pack(values, K):
    """Sequence of integers -> (n, K) uint32 limb array."""
    vals = [int(v) for v in values]
    a = np.empty((len(vals), K), dtype=np.uint32)
    for j in range(K):
        sh = LIMB_BITS * j
        a[:, j] = [(v >> sh) & LIMB_MASK for v in vals]
    return a


def unpack(arr):
    arr = np.asarray(arr).reshape(arr.shape[0], -1)
    K = arr.shape[1]
    out = [0] * arr.shape[0]
    for j in reversed(range(K)):
        col = arr[:, j].tolist()
        out = [(o << LIMB_BITS) | c for o, c in zip(out, col)]
    return out


def mont_np(p):
    """-p^-1 mod 2^24 (mont_np, reference :456-462)."""
    return (-pow(p, -1, 1 << LIMB_BITS)) % (1 << LIMB_BITS)


def limbs_for(p):
    K = max(1, (p.bit_length() + LIMB_BITS - 1) // LIMB_BITS)
    if K > K_MAX:
        raise ValueError("modulus wider than 12 limbs")
    return K


def root_for(K, log_n):
    """Primitive 2^log_n-th root in FieldConfig<K>'s field."""
    p, w, m = FIELDS[K]
    if log_n > m:
        raise ValueError(f"log_n={log_n} exceeds MAX_LOG_N={m} for K={K}")
    for _ in range(m - log_n):
        w = w * w % p
    return w


def _check(st):
    _n.check(st)


def _u32p(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32))


def _sigs(L):
    P = ctypes.POINTER(ctypes.c_uint32)
    vp, u32 = ctypes.c_void_p, ctypes.c_uint32
    for name, args in {
        "ntt_c_arb_forward": [vp, vp, u32, u32, P, u32, u32, vp],
        "ntt_c_arb_inverse": [vp, vp, u32, u32, P, u32, P, u32, vp],
        "ntt_c_arb_pointwise_mul": [vp, vp, u32, P, u32, u32, vp],
        "ntt_c_arb_scale": [vp, u32, P, P, u32, u32, vp],
        "ntt_c_arb_plan_create": [ctypes.POINTER(vp), P, u32, u32, u32, P],
        "ntt_c_arb_plan_forward": [vp, vp, vp],
        "ntt_c_arb_plan_inverse": [vp, vp, vp],
        "ntt_c_arb_plan_destroy": [vp],
    }.items():
        f = getattr(L, name)
        f.argtypes = args
        f.restype = ctypes.c_int


_n.register_sigs(_sigs)


class Field:
    """Modulus p (odd) with its Montgomery constants for K = limbs_for(p) or a given K."""

    def __init__(self, p, K=None):
        self.p = int(p)
        self.K = K or limbs_for(self.p)
        if self.p >> (LIMB_BITS * self.K):
            raise ValueError("modulus does not fit K limbs")
        self.R = 1 << (LIMB_BITS * self.K)
        self.np = mont_np(self.p)
        self.p_limbs = to_limbs(self.p, self.K)

    def to_mont(self, xs):
        return [x * self.R % self.p for x in xs]

    def from_mont(self, xs):
        ri = pow(self.R, -1, self.p)
        return [x * ri % self.p for x in xs]

    def limbs(self, x):
        return to_limbs(int(x) % self.p, self.K)


class ArbNtt:
    """Transforms of length n = 2^log_n over `field` with root w (default: the
    primitive n-th root of FieldConfig<K>).  Holds the device twiddle table
    d_tw[k] = w^k R mod p, and for negacyclic use the psi-twist tables."""

    def __init__(self, field, log_n, omega=None, psi=None):
        self.f = field
        self.log_n = log_n
        self.n = 1 << log_n
        p, n = field.p, self.n
        if omega is None:
            omega = root_for(field.K, log_n)
        if pow(omega, n, p) != 1 or (n > 1 and pow(omega, n // 2, p) == 1):
            raise ValueError("omega is not a primitive n-th root of unity")
        self.omega = omega
        tw = [0] * n
        t = field.R % p
        for k in range(n):
            tw[k] = t
            t = t * omega % p
        self.d_tw = _n.DeviceBuffer.from_array(pack(tw, field.K))
        self.n_inv_mont = field.limbs(pow(n, -1, p) * field.R)
        self.psi = psi
        self.d_psi = self.d_psi_inv = None
        if psi is not None:
            if psi * psi % p != omega or pow(psi, n, p) != p - 1:
                raise ValueError("psi must satisfy psi^2 = omega and psi^n = -1")
            fw, iv = [0] * n, [0] * n
            a, b, pi = field.R % p, field.R % p, pow(psi, -1, p)
            for j in range(n):
                fw[j], iv[j] = a, b
                a, b = a * psi % p, b * pi % p
            self.d_psi = _n.DeviceBuffer.from_array(pack(fw, field.K))
            self.d_psi_inv = _n.DeviceBuffer.from_array(pack(iv, field.K))

    def forward(self, buf, stream=None):
        f = self.f
        _check(_n.lib().ntt_c_arb_forward(buf.ptr, self.d_tw.ptr, self.n, self.log_n, _u32p(f.p_limbs), f.np, f.K,
                                          _n._sp(stream)))

    def inverse(self, buf, stream=None):
        f = self.f
        _check(_n.lib().ntt_c_arb_inverse(buf.ptr, self.d_tw.ptr, self.n, self.log_n, _u32p(f.p_limbs), f.np,
                                          _u32p(self.n_inv_mont), f.K, _n._sp(stream)))

    def forward_negacyclic(self, buf, stream=None):
        if self.d_psi is None:
            raise ValueError("built without psi")
        pointwise_mul(self.f, buf, self.d_psi, self.n, stream)
        self.forward(buf, stream)

    def inverse_negacyclic(self, buf, stream=None):
        if self.d_psi is None:
            raise ValueError("built without psi")
        self.inverse(buf, stream)
        pointwise_mul(self.f, buf, self.d_psi_inv, self.n, stream)


class ArbPlan:
    """The transforms of an ArbNtt on the four-step pass kernels
    (ntt_c_arb_plan_*): moduli that fit 64-bit arithmetic (K <= 2, or K = 3
    below 2^62 or Goldilocks).  Same data and results as ArbNtt.forward /
    inverse.  Raises NttError (status NTT_ERR_UNSUPPORTED) for other moduli."""

    def __init__(self, ntt):
        self.ntt = ntt
        f = ntt.f
        root = f.limbs(ntt.omega * f.R)
        h = ctypes.c_void_p()
        _check(_n.lib().ntt_c_arb_plan_create(ctypes.byref(h), _u32p(f.p_limbs), f.np, f.K, ntt.log_n, _u32p(root)))
        self.h = h.value

    def forward(self, buf, stream=None):
        _check(_n.lib().ntt_c_arb_plan_forward(self.h, buf.ptr, _n._sp(stream)))

    def inverse(self, buf, stream=None):
        _check(_n.lib().ntt_c_arb_plan_inverse(self.h, buf.ptr, _n._sp(stream)))

    def destroy(self):
        if self.h:
            _check(_n.lib().ntt_c_arb_plan_destroy(self.h))
            self.h = None


def pointwise_mul(field, a, b, n, stream=None):
    _check(_n.lib().ntt_c_arb_pointwise_mul(a.ptr, b.ptr, n, _u32p(field.p_limbs), field.np, field.K,
                                            _n._sp(stream)))


def scale(field, a, n, s, stream=None):
    """a[i] <- a[i] * s * R^-1 mod p (s an integer < p)."""
    sl = field.limbs(s)
    _check(_n.lib().ntt_c_arb_scale(a.ptr, n, _u32p(sl), _u32p(field.p_limbs), field.np, field.K, _n._sp(stream)))


def to_mont_device(field, a, n, stream=None):
    scale(field, a, n, field.R * field.R % field.p, stream)


def from_mont_device(field, a, n, stream=None):
    scale(field, a, n, 1, stream)
