"""Every plan shape (1..4 passes, uneven radices, narrow and wide tiles,
uint32 and uint64 storage, 31- and 62-bit primes) against the oracle, with
VALU rounds only and with the radix-16 rounds on the matrix cores."""
import numpy as np
import pytest

from props import sampled_sums_ok

pytestmark = pytest.mark.gpu

BB = 2013265921
P62 = 4611686018326724609

SPLITS = [
    (6, [6], 6), (6, [6], 12), (8, [4, 4], 8), (8, [5, 3], 10), (9, [3, 3, 3], 9), (10, [4, 3, 3], 14),
    (10, [2, 2, 3, 3], 11), (12, [12], 14), (13, [7, 6], 11), (14, [14], 14), (16, [8, 8], 12),
    (16, [6, 5, 5], 11), (17, [6, 6, 5], 14), (18, [9, 9], 14), (20, [10, 10], 14), (20, [7, 7, 6], 12),
    (20, [5, 5, 5, 5], 11),
]


@pytest.fixture(scope="module")
def ctx(nt):
    c = nt.Context()
    yield c
    c.destroy()


def _mask(split):
    """Matrix-core rounds in every pass with R >= 16."""
    return sum(1 << i for i, r in enumerate(split) if r >= 4)


@pytest.mark.parametrize("log_n,split,log_e", SPLITS)
@pytest.mark.parametrize("kind", ["cyc", "icyc", "nega", "inega"])
@pytest.mark.parametrize("mf", [False, True])
def test_split_u32(nt, orc, ctx, log_n, split, log_e, kind, mf):
    n = 1 << log_n
    x = orc.random_poly(n, BB, 7 * log_n + len(split))
    flags = {"cyc": 0, "icyc": nt.FLAG_INVERSE, "nega": nt.FLAG_NEGACYCLIC,
             "inega": nt.FLAG_NEGACYCLIC | nt.FLAG_INVERSE}[kind]
    buf = nt.DeviceBuffer.from_array(x.astype(np.uint32))
    ctx.transform_split(buf.ptr, BB, log_n, flags, split, log_e, mfma=_mask(split) if mf else 0)
    nt.device_sync()
    got = buf.download(np.uint32, n).astype(np.uint64)
    assert np.array_equal(got, orc.transform(x, BB, "nega" in kind, kind.startswith("i")))


@pytest.mark.parametrize("log_n,split,log_e", [(8, [4, 4], 8), (12, [6, 6], 10), (13, [5, 4, 4], 13),
                                               (16, [8, 8], 12), (18, [6, 6, 6], 12)])
@pytest.mark.parametrize("q,mf", [(BB, False), (BB, True), (P62, False)])  # 62-bit: VALU rounds only
def test_split_u64(nt, orc, ctx, log_n, split, log_e, q, mf):
    n = 1 << log_n
    x = orc.random_poly(n, q, 3 * log_n)
    for flags, nega, inv in [(0, False, False), (3, True, True), (2, True, False)]:
        buf = nt.DeviceBuffer.from_array(x)
        ctx.transform_split(buf.ptr, q, log_n, flags | nt.FLAG_U64, split, min(log_e, 13 if q == P62 else 14),
                            mfma=_mask(split) if mf else 0)
        nt.device_sync()
        assert np.array_equal(buf.download(np.uint64, n), orc.transform(x, q, nega, inv)), (flags, split)
        buf.free()


def test_split_rejects_bad_shapes(nt, ctx):
    buf = nt.DeviceBuffer(1 << 12)
    with pytest.raises(nt.NttError):
        ctx.transform_split(buf.ptr, BB, 10, 0, [4, 4], 12)      # does not cover n
    with pytest.raises(nt.NttError):
        ctx.transform_split(buf.ptr, BB, 10, 0, [10], 16)        # tile too large


MFMA_SPLITS = [
    (18, [12, 6], 14, 1), (18, [12, 6], 14, 3), (18, [6, 12], 14, 2), (16, [6, 10], 12, 1),
    (24, [12, 12], 14, 3), (24, [12, 12], 13, 3), (24, [8, 8, 8], 13, 7), (24, [8, 8, 8], 12, 5),
    (20, [12, 8], 14, 1), (20, [10, 10], 13, 3), (12, [6, 6], 10, 3), (22, [11, 11], 14, 3),
    (22, [6, 6, 10], 12, 3),
]


@pytest.mark.parametrize("log_n,split,log_e,mask", MFMA_SPLITS)
@pytest.mark.parametrize("kind", ["cyc", "icyc", "nega", "inega"])
@pytest.mark.parametrize("u64", [False, True])
def test_mfma_split(nt, orc, ctx, log_n, split, log_e, mask, kind, u64):
    """Radix-16 rounds on the matrix cores in the passes of `mask`, against the
    oracle, bit-exact; at n >= 2^22, in every direction and on both storage
    widths, through size-independent properties (outputs sampled from the
    defining sums, round trip through the heuristic plan)."""
    q = BB
    n = 1 << log_n
    nega, inv = "nega" in kind, kind.startswith("i")
    flags = {"cyc": 0, "icyc": nt.FLAG_INVERSE, "nega": nt.FLAG_NEGACYCLIC,
             "inega": nt.FLAG_NEGACYCLIC | nt.FLAG_INVERSE}[kind] | (nt.FLAG_U64 if u64 else 0)
    dt = np.uint64 if u64 else np.uint32
    x = orc.random_poly(n, q, 11 * log_n + mask)
    buf = nt.DeviceBuffer.from_array(x.astype(dt))
    ctx.transform_split(buf.ptr, q, log_n, flags, split, log_e, mfma=mask)
    nt.device_sync()
    got = buf.download(dt, n).astype(np.uint64)
    if log_n <= 20:
        assert np.array_equal(got, orc.transform(x, q, nega, inv)), (split, mask)
    else:
        # sampled outputs from the defining sums, then the other direction
        # through the heuristic plan (VALU rounds)
        assert sampled_sums_ok(got, x, q, orc.generator(q), nega, inv, ks=(0, 3, n // 7, n - 1)), (split, mask)
        ctx.transform(buf.ptr, q, log_n, flags ^ nt.FLAG_INVERSE)
        nt.device_sync()
        assert np.array_equal(buf.download(dt, n).astype(np.uint64), x)
    buf.free()


@pytest.mark.parametrize("q,log_n,split,log_e", [(1056440321, 16, [8, 8], 13), (7340033, 15, [7, 8], 12),
                                                 (65537, 14, [6, 8], 14), (2147352577, 16, [12, 4], 13)])
@pytest.mark.parametrize("kind", ["cyc", "inega"])
def test_mfma_primes(nt, orc, ctx, q, log_n, split, log_e, kind):
    """Matrix-core rounds for primes of every size class: below 2^30 the last
    round reduces its outputs by a Shoup step instead of one subtraction."""
    n = 1 << log_n
    x = orc.random_poly(n, q, 5 * log_n)
    flags = 0 if kind == "cyc" else nt.FLAG_NEGACYCLIC | nt.FLAG_INVERSE
    buf = nt.DeviceBuffer.from_array(x.astype(np.uint32))
    ctx.transform_split(buf.ptr, q, log_n, flags, split, log_e, mfma=_mask(split))
    nt.device_sync()
    got = buf.download(np.uint32, n).astype(np.uint64)
    assert np.array_equal(got, orc.transform(x, q, kind == "inega", kind == "inega"))
