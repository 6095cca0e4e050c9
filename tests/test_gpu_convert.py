"""Coefficient width converters (ntt_c_u64_to_u32, ntt_c_u32_to_u64): the wire
format between the reference ABI's uint64_t arrays and the native 32-bit
storage.  Bit-exact against numpy, plus the round trip through a 32-bit
descriptor transform checked against the oracle."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

BB = 2013265921
U64, U32 = np.uint64, np.uint32


@pytest.mark.parametrize("n", [1, 3, 4, 5, 1000, (1 << 20) + 3])
@pytest.mark.parametrize("q", [BB, 4294967291, 65537])
def test_narrow_reduces_mod_q(nt, n, q):
    rng = np.random.default_rng(n + q)
    x = rng.integers(0, 1 << 64, size=n, dtype=U64, endpoint=False)
    x[: min(n, 4)] = [0, q - 1, q, (1 << 64) - 1][: min(n, 4)]
    src = nt.DeviceBuffer.from_array(x)
    dst = nt.DeviceBuffer(n * 4 + 16)
    nt.u64_to_u32(dst, src, n, q)
    nt.device_sync()
    assert np.array_equal(dst.download(U32, n), (x % U64(q)).astype(U32))
    src.free()
    dst.free()


@pytest.mark.parametrize("n", [1, 7, 8, 4096, (1 << 18) + 1])
def test_widen_zero_extends(nt, n):
    x = np.random.default_rng(n).integers(0, 1 << 32, size=n, dtype=np.uint64).astype(U32)
    src = nt.DeviceBuffer.from_array(x)
    dst = nt.DeviceBuffer(n * 8)
    nt.u32_to_u64(dst, src, n)
    nt.device_sync()
    assert np.array_equal(dst.download(U64, n), x.astype(U64))
    src.free()
    dst.free()


def test_reference_abi_through_u32_storage(nt, orc):
    """uint64 coefficients -> 32-bit storage -> negacyclic forward (descriptor
    API) -> uint64: the oracle's transform of the original array."""
    log_n = 16
    n = 1 << log_n
    a = orc.random_poly(n, BB, 77)
    d64 = nt.DeviceBuffer.from_array(a)
    d32 = nt.DeviceBuffer(n * 4)
    nt.u64_to_u32(d32, d64, n, BB)
    ctx = nt.Context()
    ctx.transform(d32.ptr, BB, log_n, nt.FLAG_NEGACYCLIC)
    nt.u32_to_u64(d64, d32, n)
    nt.device_sync()
    assert np.array_equal(d64.download(U64, n), orc.ntt_negacyclic(a, BB))
    ctx.destroy()
    d64.free()
    d32.free()


def test_converters_reject_bad_arguments(nt):
    n = 64
    a = nt.DeviceBuffer(n * 8 + 64)
    b = nt.DeviceBuffer(n * 8 + 64)

    class At:
        def __init__(self, buf, off):
            self.ptr = buf.ptr + off
    with pytest.raises(nt.NttError):
        nt.u64_to_u32(b, a, n, 1 << 32)            # q too large for 32-bit storage
    with pytest.raises(nt.NttError):
        nt.u64_to_u32(At(b, 4), a, n, BB)          # misaligned destination
    with pytest.raises(nt.NttError):
        nt.u32_to_u64(At(a, 16), a, n)             # overlapping buffers
    nt.device_sync()
    a.free()
    b.free()
