"""The shipped plan table (ntt-tensorcore_amd/tuned_mi355x.txt) on the GPU.

bench.py and every caller that loads the table run these exact plans, so each
entry is run here with its own radix split, tile and matrix-core mask, in all
four directions (cyclic / negacyclic, forward / inverse), on 32-bit and
uint64 storage (62-bit and Goldilocks entries: uint64 only), bit-exact
against the oracle up to 2^22.  At 2^23 and 2^24 the output is checked at sampled coefficients
against the defining sums and round-tripped through the same plan's inverse
(the pattern of the reference's tests/test_ntt.cu:288-301: round trip plus
GPU against host at every size).

With the table loaded, the reference's uint64 entry points (ntt_c_forward /
ntt_c_inverse, the negacyclic pair and ntt_c_negacyclic_multiply) also run
once at every shipped single-polynomial size.
"""
import os

import numpy as np
import pytest

from props import sampled_sums_ok
from workloads import p62_primes

pytestmark = pytest.mark.gpu

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
TABLE = os.path.join(REPO, "ntt-tensorcore_amd", "tuned_mi355x.txt")
BB = 2013265921
P62 = p62_primes()[0]
GOLD = 0xFFFFFFFF00000001
Q_OF = {0: BB, 1: P62, 2: GOLD}  # arithmetic kind of an entry -> the prime its cases use
EXACT_MAX = 22  # oracle up to 2^22 (about a second per transform)
KINDS = {"cyc": (False, False), "icyc": (False, True), "nega": (True, False), "inega": (True, True)}


def _entries():
    out = []
    for line in open(TABLE):
        f = line.split()
        if len(f) != 10:
            continue
        ar, log_n, le, npass = (int(v) for v in f[:4])
        split = [int(v) for v in f[4:4 + npass]]
        out.append((ar, log_n, le, split, int(f[8]), int(f[9])))
    return out


ENTRIES = _entries()
CASES = [(e, kind, u64) for e in ENTRIES for kind in KINDS for u64 in ((False, True) if e[0] == 0 else (True,))]


def _id(c):
    (ar, log_n, le, split, log_e, mask), kind, u64 = c
    return f"{('m32', 'm64', 'gold')[ar]}-2^{log_n}@2^{le}-{'x'.join(map(str, split))}-e{log_e}-m{mask}-{kind}-" \
           f"{'u64' if u64 else 'u32'}"


@pytest.fixture(scope="module")
def ctx(nt):
    c = nt.Context()
    yield c
    c.destroy()


_EXPECT = {}


def _inputs(orc, q, log_n, kind):
    """(x, oracle output or None) for a (q, n, kind), computed once per module."""
    key = (q, log_n, kind)
    if key not in _EXPECT:
        x = orc.random_poly(1 << log_n, q, 1000 * log_n + len(_EXPECT) % 7)
        nega, inv = KINDS[kind]
        _EXPECT[key] = (x, orc.transform(x, q, nega, inv) if log_n <= EXACT_MAX else None)
    return _EXPECT[key]


def test_table_is_the_shipped_one():
    """Every entry parses, covers its n, and names a real tile and mask."""
    assert len(ENTRIES) >= 40
    for ar, log_n, le, split, log_e, mask in ENTRIES:
        assert sum(split) == log_n and 1 <= len(split) <= 4
        assert 8 <= log_e <= 14 and mask < (1 << len(split))
        assert ar in (0, 1, 2) and (ar == 0 or mask == 0)


@pytest.mark.parametrize("case", CASES, ids=[_id(c) for c in CASES])
def test_shipped_plan(nt, orc, ctx, case):
    (ar, log_n, le, split, log_e, mask), kind, u64 = case
    q = Q_OF[ar]
    n = 1 << log_n
    nega, inv = KINDS[kind]
    x, expect = _inputs(orc, q, log_n, kind)
    dt = np.uint64 if u64 else np.uint32
    flags = (nt.FLAG_NEGACYCLIC if nega else 0) | (nt.FLAG_INVERSE if inv else 0) | (nt.FLAG_U64 if u64 else 0)
    buf = nt.DeviceBuffer.from_array(x.astype(dt))
    try:
        ctx.transform_split(buf.ptr, q, log_n, flags, split, log_e, mfma=mask)
        nt.device_sync()
        got = buf.download(dt, n).astype(np.uint64)
        if expect is not None:
            assert np.array_equal(got, expect)
        else:
            assert sampled_sums_ok(got, x, q, orc.generator(q), nega, inv, ks=(0, 3, n // 3, n - 1))
        # round trip through the same plan's other direction
        ctx.transform_split(buf.ptr, q, log_n, flags ^ nt.FLAG_INVERSE, split, log_e, mfma=mask)
        nt.device_sync()
        assert np.array_equal(buf.download(dt, n).astype(np.uint64), x)
    finally:
        buf.free()


@pytest.fixture
def shipped_table(nt):
    nt.autotune_clear()
    nt.autotune_load(TABLE)
    yield
    nt.autotune_clear()


SINGLE = sorted({e[1] for e in ENTRIES if e[0] == 0 and e[1] == e[2]})
SINGLE_ALL = [(BB, ln) for ln in SINGLE] + [(GOLD, e[1]) for e in ENTRIES if e[0] == 2 and e[1] == e[2]]


@pytest.mark.parametrize("q,log_n", SINGLE_ALL)
def test_reference_abi_with_shipped_table(nt, orc, shipped_table, q, log_n):
    """ntt_c_forward/inverse and the negacyclic pair on the reference's uint64_t
    data, with the plan each n gets from the shipped table."""
    n = 1 << log_n
    x = orc.random_poly(n, q, 77 + log_n)
    cfg = nt.Config(n, q)
    buf = nt.DeviceBuffer.from_array(x)
    try:
        for fwd, bwd, nega in ((cfg.forward, cfg.inverse, False),
                               (cfg.forward_negacyclic, cfg.inverse_negacyclic, True)):
            fwd(buf)
            nt.device_sync()
            got = buf.download(np.uint64, n)
            if log_n <= EXACT_MAX:
                assert np.array_equal(got, orc.transform(x, q, nega, False)), nega
            else:
                assert sampled_sums_ok(got, x, q, orc.generator(q), nega, False), nega
            bwd(buf)
            nt.device_sync()
            assert np.array_equal(buf.download(np.uint64, n), x), nega
    finally:
        buf.free()
        cfg.cleanup()


@pytest.mark.parametrize("log_n", SINGLE)
def test_negacyclic_multiply_with_shipped_table(nt, orc, shipped_table, log_n):
    """The fused product through the plans of the shipped table: the oracle's
    pipeline (forward, forward, pointwise, inverse) up to 2^20, above that
    multiplication by 3*x^k (a negacyclic shift scaled by 3)."""
    n = 1 << log_n
    q = BB
    a = orc.random_poly(n, q, 5 + log_n)
    if log_n <= 20:
        b = orc.random_poly(n, q, 6 + log_n)
        A, B = orc.ntt_negacyclic(a, q), orc.ntt_negacyclic(b, q)
        expect = orc.ntt_negacyclic(orc.pointwise_mul(A, B, q), q, True)
    else:
        k = n // 2 + 99
        b = np.zeros(n, dtype=np.uint64)
        b[k] = 3
        sa = np.roll(a, k).astype(object) * 3 % q
        sa[:k] = (q - sa[:k]) % q
        expect = sa.astype(np.uint64)
    cfg = nt.Config(n, q)
    da, db, dr = nt.DeviceBuffer.from_array(a), nt.DeviceBuffer.from_array(b), nt.DeviceBuffer(n * 8)
    try:
        cfg.negacyclic_multiply(dr, da, db)
        nt.device_sync()
        assert np.array_equal(dr.download(np.uint64, n), expect)
    finally:
        for d in (da, db, dr):
            d.free()
        cfg.cleanup()
