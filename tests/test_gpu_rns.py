"""RNS kernels around the transform: fast base conversion and the per-limb
multiply-accumulate, bit-exact against the oracle (oracle/ntt_oracle.c
orc_bconv / orc_pmac, themselves pinned by CRT in tests/test_oracle.py).
The reference has no code for either (src/ntt_optimized.cu:510-513)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

U32 = np.uint32


def _primes(orc, count, below=1 << 31):
    return [int(v) for v in orc.ntt_primes(below, 17, count)]


def _residues(rng, mods, n):
    return np.stack([rng.integers(0, m, size=n, dtype=np.uint64) for m in mods]).astype(U32)


# every (ceil(L/16), ceil(K/16)) instance of the kernel: (1,1) (1,2) (1,3) (1,4) (2,1) (2,2) (2,3) (2,4)
@pytest.mark.parametrize("L,K", [(1, 1), (3, 2), (8, 8), (4, 20), (5, 40), (16, 64), (20, 10), (24, 24), (30, 40),
                                 (32, 64)])
@pytest.mark.parametrize("n", [1, 257, 1028, 4099, 4100, 1 << 16])
def test_bconv_vs_oracle(nt, orc, L, K, n):
    primes = _primes(orc, L + K)
    q, p = primes[:L], primes[L:]
    x = _residues(np.random.default_rng(L * 1000 + K + n), q, n)
    plan = nt.Bconv(q, p)
    din = nt.DeviceBuffer.from_array(x)
    dout = nt.DeviceBuffer(K * n * 4)
    plan.run(dout, din, n)
    nt.device_sync()
    assert np.array_equal(dout.download(U32, K * n).reshape(K, n), orc.bconv(x, q, p))
    plan.destroy()
    din.free()
    dout.free()


@pytest.mark.parametrize("L,K,below,fill,n", [(32, 64, 1 << 31, "max", 4100), (32, 64, 1 << 31, "max", 4099),
                                              (17, 20, 1 << 31, "zero", 4100), (20, 33, 1 << 18, "max", 4099),
                                              (16, 16, 1 << 14, "random", 4100)])
def test_bconv_moduli_sizes_and_extreme_residues(nt, orc, L, K, below, fill, n):
    """Small target moduli (the centred matrix entries are then small) and
    inputs at the ends of the range (every residue q_i - 1, or all zero);
    n = 4100 runs the 16-byte access path with a partial last block, 4099 the
    word path."""
    primes = [int(v) for v in orc.ntt_primes(below, 1, L + K)]
    q, p = primes[:L], primes[L:]
    if fill == "max":
        x = np.stack([np.full(n, m - 1, dtype=np.uint64) for m in q]).astype(U32)
    elif fill == "zero":
        x = np.zeros((L, n), dtype=U32)
    else:
        x = _residues(np.random.default_rng(L + K), q, n)
    plan = nt.Bconv(q, p)
    din = nt.DeviceBuffer.from_array(x)
    dout = nt.DeviceBuffer(K * n * 4)
    plan.run(dout, din, n)
    nt.device_sync()
    assert np.array_equal(dout.download(U32, K * n).reshape(K, n), orc.bconv(x, q, p))
    plan.destroy()
    din.free()
    dout.free()


@pytest.mark.parametrize("L,K,n", [(3, 5, (1 << 21) + 260), (18, 2, (1 << 19) + 131)])
def test_bconv_grid_stride(nt, orc, L, K, n):
    """Sizes beyond one pass of the grid (8192 workgroups): the 16-byte path
    (n % 4 == 0) and the word path each loop over further blocks."""
    primes = _primes(orc, L + K)
    q, p = primes[:L], primes[L:]
    x = _residues(np.random.default_rng(n), q, n)
    plan = nt.Bconv(q, p)
    din = nt.DeviceBuffer.from_array(x)
    dout = nt.DeviceBuffer(K * n * 4)
    plan.run(dout, din, n)
    nt.device_sync()
    assert np.array_equal(dout.download(U32, K * n).reshape(K, n), orc.bconv(x, q, p))
    plan.destroy()
    din.free()
    dout.free()


@pytest.mark.parametrize("dnum,L", [(1, 1), (3, 4), (8, 24), (32, 2)])
@pytest.mark.parametrize("acc", [False, True])
@pytest.mark.parametrize("n", [3000, 3001])
def test_pmac_vs_oracle(nt, orc, dnum, L, acc, n):
    q = _primes(orc, L)
    rng = np.random.default_rng(dnum * 100 + L + n)
    a = np.stack([_residues(rng, q, n) for _ in range(dnum)])
    b = np.stack([_residues(rng, q, n) for _ in range(dnum)])
    base = _residues(rng, q, n)
    plan = nt.Pmac(q)
    da, db = nt.DeviceBuffer.from_array(a), nt.DeviceBuffer.from_array(b)
    dout = nt.DeviceBuffer.from_array(base)
    plan.run(dout, da, db, dnum, n, accumulate=acc)
    nt.device_sync()
    assert np.array_equal(dout.download(U32, L * n).reshape(L, n), orc.pmac(a, b, q, base if acc else None))
    plan.destroy()
    for x in (da, db, dout):
        x.free()


def test_rns_plans_reject_bad_moduli(nt, orc):
    q = _primes(orc, 3)
    with pytest.raises(nt.NttError):
        nt.Bconv([q[0], q[0]], q[1:])               # repeated source modulus
    with pytest.raises(nt.NttError):
        nt.Bconv(q, [q[0] + 2])                     # not a prime (even)
    with pytest.raises(nt.NttError):
        nt.Bconv(_primes(orc, 33), q)               # more than 32 source moduli
    with pytest.raises(nt.NttError):
        nt.Pmac([(1 << 31) + 11])                   # above 2^31
