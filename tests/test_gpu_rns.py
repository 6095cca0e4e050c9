"""RNS kernels around the transform: fast base conversion and the per-limb
multiply-accumulate, bit-exact against the oracle (oracle/ntt_oracle.c
orc_bconv / orc_pmac, themselves pinned by CRT in tests/test_oracle.py).
The reference has no code for either (src/ntt_optimized.cu:510-513)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

U32 = np.uint32
BB = 2013265921


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


def _ip_expect(orc, a, b, q, nega):
    """Per limb l: the inverse transform of sum_k a[k][l] * b[k][l] mod q_l
    (the oracle's pmac, then its inverse transform)."""
    x = orc.pmac(a.astype(U32), b.astype(U32), q, None)
    return np.stack([np.asarray(orc.transform(x[l].astype(np.uint64), q[l], nega, True), dtype=np.uint64)
                     for l in range(len(q))])


# The key-switching inner product fused into the inverse transforms of a batch
# (ntt_c_batch_prepare_ip): limb l of the output is INTT_{q_l}(sum_k a[k][l] b[k][l]),
# with a and b laid out [k][l][c] as for ntt_c_pmac.  One-pass (2^10, 2^12) and
# multi-pass plans (2^16, 2^18), 32- and 64-bit storage, dnum = 1 (a pointwise
# product) to 5.
@pytest.mark.parametrize("u64", [False, True])
@pytest.mark.parametrize("nega", [True, False])
@pytest.mark.parametrize("dnum,L,log_n", [(1, 2, 10), (3, 4, 12), (3, 6, 16), (5, 2, 18), (2, 3, 4)])
def test_ip_fused_inverse_vs_oracle(nt, orc, dnum, L, log_n, nega, u64):
    n = 1 << log_n
    q = [int(v) for v in orc.ntt_primes(1 << 31, max(17, log_n + 1), L)]
    rng = np.random.default_rng(dnum * 1000 + L * 10 + log_n)
    a = np.stack([_residues(rng, q, n) for _ in range(dnum)])
    b = np.stack([_residues(rng, q, n) for _ in range(dnum)])
    dt, sb = (np.uint64, 8) if u64 else (U32, 4)
    flags = nt.FLAG_INVERSE | (nt.FLAG_NEGACYCLIC if nega else 0) | (nt.FLAG_U64 if u64 else 0)
    da, db = nt.DeviceBuffer.from_array(a.astype(dt)), nt.DeviceBuffer.from_array(b.astype(dt))
    dout = nt.DeviceBuffer(L * n * sb)
    ctx = nt.Context()
    bt = nt.Batch(ctx, [(dout.ptr + l * n * sb, q[l], log_n, flags) for l in range(L)],
                  ips=[(da.ptr + l * n * sb, db.ptr + l * n * sb, dnum, L * n) for l in range(L)])
    expect = _ip_expect(orc, a, b, q, nega)
    for rep in range(2):  # the inputs are read, never overwritten: a second launch gives the same
        bt.launch()
        nt.device_sync()
        got = dout.download(dt, L * n).reshape(L, n).astype(np.uint64)
        for l in range(L):
            assert np.array_equal(got[l], expect[l]), (rep, l)
    bt.destroy()
    ctx.destroy()
    for x in (da, db, dout):
        x.free()


def test_ip_fused_inverse_equals_pmac_then_inverse(nt, orc):
    """The RNS shape (24 limbs x 2^16, dnum = 3, negacyclic): the fused batch,
    replayed from a graph and as one work-queue launch, equals ntt_c_pmac
    followed by a batch of inverse transforms, and both equal the oracle on
    sampled limbs."""
    L, dnum, log_n = 24, 3, 16
    n = 1 << log_n
    q = [int(v) for v in orc.ntt_primes(1 << 31, 17, L)]
    rng = np.random.default_rng(7)
    a = np.stack([_residues(rng, q, n) for _ in range(dnum)])
    b = np.stack([_residues(rng, q, n) for _ in range(dnum)])
    flags = nt.FLAG_INVERSE | nt.FLAG_NEGACYCLIC
    da, db = nt.DeviceBuffer.from_array(a), nt.DeviceBuffer.from_array(b)
    fused, sep = nt.DeviceBuffer(L * n * 4), nt.DeviceBuffer(L * n * 4)
    ctx = nt.Context()
    bf = nt.Batch(ctx, [(fused.ptr + l * n * 4, q[l], log_n, flags) for l in range(L)],
                  ips=[(da.ptr + l * n * 4, db.ptr + l * n * 4, dnum, L * n) for l in range(L)])
    bs = nt.Batch(ctx, [(sep.ptr + l * n * 4, q[l], log_n, flags) for l in range(L)])
    pm = nt.Pmac(q)
    for _ in range(3):
        bf.launch_graph()
    pm.run(sep, da, db, dnum, n, accumulate=False)
    bs.launch()
    nt.device_sync()
    got_f = fused.download(U32, L * n).reshape(L, n)
    got_s = sep.download(U32, L * n).reshape(L, n)
    assert np.array_equal(got_f, got_s)
    expect = _ip_expect(orc, a[:, [0, 11, 23]], b[:, [0, 11, 23]], [q[0], q[11], q[23]], True)
    for i, l in enumerate([0, 11, 23]):
        assert np.array_equal(got_f[l].astype(np.uint64), expect[i]), l
    # the same batch as one work-queue launch
    fused.upload(np.zeros(L * n, dtype=U32), 0)
    bf.set_queue(True)
    bf.launch()
    bf.check()
    assert np.array_equal(fused.download(U32, L * n).reshape(L, n), got_s)
    for x in (bf, bs, pm):
        x.destroy()
    ctx.destroy()
    for x in (da, db, fused, sep):
        x.free()


# The 32-bit product / inner-product instances of the first pass (one per tile
# shape of NTT_IP_SHAPES: (7,5) (8,4) (8,5) (9,4) (10,3)), reached through
# plans with matrix-core rounds installed for the batch's load class.
@pytest.mark.parametrize("log_n,split,log_e", [(12, [7, 5], 12), (16, [8, 8], 12), (16, [8, 8], 13),
                                               (18, [9, 9], 13), (20, [10, 10], 13)])
def test_ip_shape_instances_vs_oracle(nt, orc, tmp_path, log_n, split, log_e):
    L, dnum = 2, 3
    n = 1 << log_n
    q = [int(v) for v in orc.ntt_primes(1 << 31, log_n + 1, L)]
    key = (L * n).bit_length() - 1
    table = tmp_path / "ip_plans.txt"
    table.write_text(f"0 {log_n} {key} 2 {split[0]} {split[1]} 0 0 {log_e} 3\n")
    rng = np.random.default_rng(log_n * 10 + log_e)
    a = np.stack([_residues(rng, q, n) for _ in range(dnum)])
    b = np.stack([_residues(rng, q, n) for _ in range(dnum)])
    flags = nt.FLAG_INVERSE | nt.FLAG_NEGACYCLIC
    da, db = nt.DeviceBuffer.from_array(a), nt.DeviceBuffer.from_array(b)
    dout = nt.DeviceBuffer(L * n * 4)
    ctx = nt.Context()
    nt.autotune_clear()
    try:
        nt.autotune_load(str(table))
        bt = nt.Batch(ctx, [(dout.ptr + l * n * 4, q[l], log_n, flags) for l in range(L)],
                      ips=[(da.ptr + l * n * 4, db.ptr + l * n * 4, dnum, L * n) for l in range(L)])
        bt.launch()
        nt.device_sync()
        got = dout.download(U32, L * n).reshape(L, n).astype(np.uint64)
        bt.destroy()
    finally:
        nt.autotune_clear()
    expect = _ip_expect(orc, a, b, q, True)
    for l in range(L):
        assert np.array_equal(got[l], expect[l]), l
    ctx.destroy()
    for x in (da, db, dout):
        x.free()


def test_ip_and_plain_descriptors_in_one_batch(nt, orc):
    """Inner-product inverses (multi-pass 2^16, one-pass 2^11) and plain
    transforms (forward 2^14, inverse 2^16 of its own data) in one batch:
    each output equals the oracle, and the plain inputs are transformed in
    place as in a batch without inner products."""
    dnum = 2
    q16 = [int(v) for v in orc.ntt_primes(1 << 31, 17, 2)]
    rng = np.random.default_rng(99)
    n16, n11, n14 = 1 << 16, 1 << 11, 1 << 14
    a16 = np.stack([_residues(rng, [q16[0]], n16) for _ in range(dnum)])
    b16 = np.stack([_residues(rng, [q16[0]], n16) for _ in range(dnum)])
    a11 = np.stack([_residues(rng, [q16[1]], n11) for _ in range(dnum)])
    b11 = np.stack([_residues(rng, [q16[1]], n11) for _ in range(dnum)])
    x14 = rng.integers(0, BB, size=n14, dtype=np.uint64).astype(U32)
    x16 = rng.integers(0, q16[1], size=n16, dtype=np.uint64).astype(U32)
    bufs = {k: nt.DeviceBuffer.from_array(v) for k, v in
            dict(a16=a16, b16=b16, a11=a11, b11=b11, x14=x14, x16=x16).items()}
    o16, o11 = nt.DeviceBuffer(n16 * 4), nt.DeviceBuffer(n11 * 4)
    inv, nega = nt.FLAG_INVERSE, nt.FLAG_NEGACYCLIC
    ctx = nt.Context()
    descs = [(o16.ptr, q16[0], 16, inv | nega), (bufs["x14"].ptr, BB, 14, nega), (o11.ptr, q16[1], 11, inv),
             (bufs["x16"].ptr, q16[1], 16, inv)]
    ips = [(bufs["a16"].ptr, bufs["b16"].ptr, dnum, n16), None, (bufs["a11"].ptr, bufs["b11"].ptr, dnum, n11), None]
    bt = nt.Batch(ctx, descs, ips=ips)
    bt.launch()
    nt.device_sync()
    e16 = _ip_expect(orc, a16, b16, [q16[0]], True)[0]
    e11 = _ip_expect(orc, a11, b11, [q16[1]], False)[0]
    assert np.array_equal(o16.download(U32, n16).astype(np.uint64), e16)
    assert np.array_equal(o11.download(U32, n11).astype(np.uint64), e11)
    assert np.array_equal(bufs["x14"].download(U32, n14).astype(np.uint64),
                          np.asarray(orc.transform(x14.astype(np.uint64), BB, True, False), dtype=np.uint64))
    assert np.array_equal(bufs["x16"].download(U32, n16).astype(np.uint64),
                          np.asarray(orc.transform(x16.astype(np.uint64), q16[1], False, True), dtype=np.uint64))
    bt.destroy()
    ctx.destroy()
    for x in list(bufs.values()) + [o16, o11]:
        x.free()
