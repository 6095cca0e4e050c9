"""K-limb transforms on the four-step pass kernels (ntt_c_arb_plan_*) and the
Goldilocks field on the single-word API.

The plan must give the limb kernels' results limb for limb (the same
Montgomery-form data, the same root): checked against the oracle's
host_ntt_naive restatement (reference include/ntt_arbitrary.cuh:2651-2676) at
sizes it finishes in seconds, and against the limb kernels (themselves pinned
by the oracle in test_gpu_arb.py) and by round trip above that.  Fields: K = 1
(23-bit primes), K = 2 (the 48-bit prime of FieldConfig<2> here), K = 3
Goldilocks (FieldConfig<3>, reference :515-534) and a K = 3 prime below 2^62.
Goldilocks through ntt_c_init / descriptors is checked against the same
oracle (the cyclic transform of raw values equals the limb transform of those
values read as Montgomery-form numbers, since x * (w^k R) * R^-1 = x w^k), and
its negacyclic product against the schoolbook product mod x^n + 1."""
import random

import numpy as np
import pytest

import ntt_arb as A
from test_gpu_arb import _dev, _down, _prime_with_root, _random_mont, _tw_table

pytestmark = pytest.mark.gpu

GOLD = A.GOLDILOCKS
P62 = 4611686018326724609  # < 2^62, 2^23 | p - 1
UNSUPPORTED, INVALID = 5, 1


def _root(p, log_n, seed=1):
    rng = random.Random(seed)
    n = 1 << log_n
    while True:
        w = pow(rng.randrange(2, p - 1), (p - 1) >> log_n, p)
        if n == 1 or pow(w, n // 2, p) == p - 1:
            return w


def _field(kind, log_n):
    """(Field, primitive 2^log_n-th root) for a plan-capable field kind."""
    if kind == "k1":
        p, w = _prime_with_root(1, 16, 11)
        return A.Field(p, 1), pow(w, 1 << (16 - log_n), p)
    if kind == "k2":
        return A.Field(A.FIELDS[2][0], 2), A.root_for(2, log_n)
    if kind == "gold":
        return A.Field(GOLD, 3), A.root_for(3, log_n)
    return A.Field(P62, 3), _root(P62, log_n)


KINDS = ["k1", "k2", "gold", "p62k3"]


def _max_log(kind):
    return 16 if kind == "k1" else 22


@pytest.mark.parametrize("kind", KINDS)
@pytest.mark.parametrize("log_n", [1, 2, 5, 10, 12, 13, 16])
def test_plan_forward_inverse_match_oracle(nt, orc, kind, log_n):
    n = 1 << log_n
    f, w = _field(kind, log_n)
    x = _random_mont(f, n, 100 * log_n + len(kind))
    ntt = A.ArbNtt(f, log_n, omega=w)
    plan = A.ArbPlan(ntt)
    buf = _dev(nt, x)
    try:
        print(f"plan {kind} K={f.K} log_n={log_n}", flush=True)
        plan.forward(buf)
        nt.device_sync()
        X = _down(buf, (n, f.K))
        np.testing.assert_array_equal(X, orc.arb_ntt_mont(x, _tw_table(f, w, n), f.p_limbs, f.np, f.K))
        # inverse: the limb kernels' inverse (pinned against the oracle in test_gpu_arb.py)
        ref = _dev(nt, X)
        ntt.inverse(ref)
        plan.inverse(buf)
        nt.device_sync()
        got = _down(buf, (n, f.K))
        np.testing.assert_array_equal(got, _down(ref, (n, f.K)))
        np.testing.assert_array_equal(got, x)
        ref.free()
    finally:
        plan.destroy()
        buf.free()


@pytest.mark.parametrize("kind,log_n", [("k2", 18), ("k2", 20), ("k2", 22), ("gold", 18), ("gold", 20),
                                        ("gold", 22), ("p62k3", 20), ("k1", 16)])
def test_plan_large_equals_limb_kernels(nt, kind, log_n):
    n = 1 << log_n
    f, w = _field(kind, log_n)
    x = _random_mont(f, n, 7 + log_n)
    ntt = A.ArbNtt(f, log_n, omega=w)
    plan = A.ArbPlan(ntt)
    a, b = _dev(nt, x), _dev(nt, x)
    try:
        print(f"plan large {kind} log_n={log_n}", flush=True)
        plan.forward(a)
        ntt.forward(b)
        nt.device_sync()
        np.testing.assert_array_equal(_down(a, (n, f.K)), _down(b, (n, f.K)))
        plan.inverse(a)
        nt.device_sync()
        np.testing.assert_array_equal(_down(a, (n, f.K)), x)
    finally:
        plan.destroy()
        a.free()
        b.free()


def test_goldilocks_edge_values(nt, orc):
    """Inputs that drive every borrow and carry branch of the Goldilocks
    reduction: 0, 1, p - 1, p - 2, 2^32 - 1, 2^32, 2^63, 2^64 - 2^33."""
    log_n = 12
    n = 1 << log_n
    f, w = _field("gold", log_n)
    edge = [0, 1, GOLD - 1, GOLD - 2, (1 << 32) - 1, 1 << 32, 1 << 63, (1 << 64) - (1 << 33)]
    rng = random.Random(5)
    vals = [edge[i % len(edge)] if i % 3 else rng.randrange(GOLD) for i in range(n)]
    x = A.pack(vals, 3)
    ntt = A.ArbNtt(f, log_n, omega=w)
    plan = A.ArbPlan(ntt)
    buf = _dev(nt, x)
    try:
        plan.forward(buf)
        nt.device_sync()
        np.testing.assert_array_equal(_down(buf, (n, 3)), orc.arb_ntt_mont(x, _tw_table(f, w, n), f.p_limbs, f.np, 3))
        # all p - 1: the sum of each output is a multiple of p - 1 at every level
        allm = A.pack([GOLD - 1] * n, 3)
        b2 = _dev(nt, allm)
        plan.forward(b2)
        nt.device_sync()
        np.testing.assert_array_equal(_down(b2, (n, 3)), orc.arb_ntt_mont(allm, _tw_table(f, w, n), f.p_limbs, f.np, 3))
        b2.free()
    finally:
        plan.destroy()
        buf.free()


def test_plan_rejects_other_fields(nt):
    # BN254 (K = 11) and K = 4 need more than 64 bits
    for K, p in [(11, A.FIELDS[11][0]), (4, _prime_with_root(4, 8, 3)[0])]:
        f = A.Field(p, K)
        ntt = A.ArbNtt(f, 4, omega=_root(p, 4))
        with pytest.raises(nt.NttError) as e:
            A.ArbPlan(ntt)
        assert e.value.status == UNSUPPORTED
    # a K = 3 prime between 2^62 and 2^64 other than Goldilocks
    c = (1 << 42) + 1
    while not all(pow(a, ((c << 20) + 1) - 1, (c << 20) + 1) == 1 for a in (2, 3, 5, 7, 11, 13)):
        c += 2
    p = (c << 20) + 1
    assert (1 << 62) <= p < (1 << 64)
    f = A.Field(p, 3)
    with pytest.raises(nt.NttError) as e:
        A.ArbPlan(A.ArbNtt(f, 4, omega=_root(p, 4)))
    assert e.value.status == UNSUPPORTED
    # a root that is not a primitive n-th root
    f, w = _field("gold", 8)
    ntt = A.ArbNtt(f, 8, omega=w)
    ntt.omega = w * w % GOLD
    with pytest.raises(nt.NttError) as e:
        A.ArbPlan(ntt)
    assert e.value.status == INVALID


def _gold_generator_root(log_n):
    """The generator rule's primitive 2^log_n-th root for Goldilocks
    (include/modular_arith.cuh:130-151)."""
    q = GOLD
    g = 2
    while any((q - 1) % p == 0 and pow(g, (q - 1) // p, q) == 1 for p in (2, 3, 5, 7)):
        g += 1
    return pow(g, (q - 1) >> log_n, q), pow(g, (q - 1) >> (log_n + 1), q)


@pytest.mark.parametrize("log_n", [1, 4, 11, 14, 17, 20])
def test_goldilocks_config_transforms(nt, orc, log_n):
    n = 1 << log_n
    w, psi = _gold_generator_root(log_n)
    f = A.Field(GOLD, 3)
    x = np.random.default_rng(log_n).integers(0, GOLD, size=n, dtype=np.uint64)
    xl = A.pack(x.tolist(), 3)
    expect = np.array(A.unpack(orc.arb_ntt_mont(xl, _tw_table(f, w, n), f.p_limbs, f.np, 3)), dtype=np.uint64)
    cfg = nt.Config(n, GOLD)
    buf = nt.DeviceBuffer.from_array(x)
    try:
        cfg.forward(buf)
        nt.device_sync()
        np.testing.assert_array_equal(buf.download(np.uint64, n), expect)
        cfg.inverse(buf)
        nt.device_sync()
        np.testing.assert_array_equal(buf.download(np.uint64, n), x)
        # negacyclic forward = cyclic forward of x_j psi^j
        tw = [int(v) * pow(psi, j, GOLD) % GOLD for j, v in enumerate(x.tolist())] if n <= (1 << 14) else None
        if tw is not None:
            e2 = np.array(A.unpack(orc.arb_ntt_mont(A.pack(tw, 3), _tw_table(f, w, n), f.p_limbs, f.np, 3)),
                          dtype=np.uint64)
            cfg.forward_negacyclic(buf)
            nt.device_sync()
            np.testing.assert_array_equal(buf.download(np.uint64, n), e2)
            cfg.inverse_negacyclic(buf)
            nt.device_sync()
            np.testing.assert_array_equal(buf.download(np.uint64, n), x)
    finally:
        buf.free()
        cfg.cleanup()


def _schoolbook(a, b, p):
    n = len(a)
    out = [0] * n
    for i, u in enumerate(a):
        for j, v in enumerate(b):
            if i + j < n:
                out[i + j] += u * v
            else:
                out[i + j - n] -= u * v
    return [v % p for v in out]


@pytest.mark.parametrize("log_n", [5, 7])
def test_goldilocks_negacyclic_multiply(nt, log_n):
    n = 1 << log_n
    rng = random.Random(log_n)
    a = [rng.randrange(GOLD) for _ in range(n)]
    b = [rng.randrange(GOLD) for _ in range(n)]
    cfg = nt.Config(n, GOLD)
    da, db, dr = (nt.DeviceBuffer.from_array(np.array(v, dtype=np.uint64)) for v in (a, b, [0] * n))
    try:
        cfg.negacyclic_multiply(dr, da, db)
        nt.device_sync()
        assert dr.download(np.uint64, n).tolist() == _schoolbook(a, b, GOLD)
        # the pointwise kernel
        da.upload(np.array(a, dtype=np.uint64))
        db.upload(np.array(b, dtype=np.uint64))
        cfg.poly_multiply(dr, da, db)
        nt.device_sync()
        assert dr.download(np.uint64, n).tolist() == [u * v % GOLD for u, v in zip(a, b)]
    finally:
        for d in (da, db, dr):
            d.free()
        cfg.cleanup()


def test_goldilocks_descriptor_batch(nt, orc):
    """Goldilocks polynomials in a descriptor batch beside 62-bit and 31-bit ones."""
    f = A.Field(GOLD, 3)
    ctx = nt.Context()
    specs = [(GOLD, 12, 0), (GOLD, 16, nt.FLAG_INVERSE), (P62, 12, 0), (2013265921, 14, 0), (GOLD, 10, 0)]
    bufs, xs = [], []
    rng = np.random.default_rng(3)
    for q, ln, fl in specs:
        x = rng.integers(0, q, size=1 << ln, dtype=np.uint64)
        xs.append(x)
        bufs.append(nt.DeviceBuffer.from_array(x))
    try:
        bt = nt.Batch(ctx, [(b.ptr, q, ln, fl | nt.FLAG_U64) for b, (q, ln, fl) in zip(bufs, specs)])
        bt.launch()
        nt.device_sync()
        for b, x, (q, ln, fl) in zip(bufs, xs, specs):
            got = b.download(np.uint64, 1 << ln)
            if q == GOLD and not fl:
                w, _ = _gold_generator_root(ln)
                e = A.unpack(orc.arb_ntt_mont(A.pack(x.tolist(), 3), _tw_table(f, w, 1 << ln), f.p_limbs, f.np, 3))
                np.testing.assert_array_equal(got, np.array(e, dtype=np.uint64))
            elif q == GOLD:
                ctx.transform(b.ptr, q, ln, nt.FLAG_U64)  # forward undoes the inverse
                nt.device_sync()
                np.testing.assert_array_equal(b.download(np.uint64, 1 << ln), x)
            else:
                np.testing.assert_array_equal(got, orc.ntt(x, q))
        bt.destroy()
    finally:
        for b in bufs:
            b.free()
        ctx.destroy()
