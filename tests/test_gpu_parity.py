"""GPU parity of the include/ntt.cuh surface against the oracle, bit-exact.

Mirrors the reference's own suite (tests/test_ntt.cu): round trips and
GPU-vs-ntt_host at every size the oracle finishes in seconds, plus the
negacyclic transform it never implemented, 62-bit primes, the batch entry
points, poly multiply, bit reversal and the pre-transposed (RowMaj) layout.
At 2^22..2^24 the checks are size-independent: round trips, linearity, and
direct evaluation of sampled output coefficients."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

BB = 2013265921
P30 = 1056440321
P62 = 4611686018326724609
U64 = np.uint64


def gpu_apply(nt, x, fn):
    buf = nt.DeviceBuffer.from_array(np.ascontiguousarray(x, dtype=U64))
    fn(buf)
    nt.device_sync()
    out = buf.download(U64, x.size)
    buf.free()
    return out


@pytest.fixture(scope="module")
def cfgs(nt):
    cache = {}

    def get(n, q):
        if (n, q) not in cache:
            cache[(n, q)] = nt.Config(n, q)
        return cache[(n, q)]
    yield get
    for c in cache.values():
        c.cleanup()


@pytest.mark.parametrize("variant", ["basic", "four_step", "tlmop", "optimized"])
def test_cfg1_babybear_4096_all_variants(nt, orc, cfgs, variant):
    n = 4096
    x = orc.random_poly(n, BB, 42)
    c = cfgs(n, BB)
    X = gpu_apply(nt, x, lambda b: c.forward(b, variant))
    assert np.array_equal(X, orc.ntt(x, BB))
    y = gpu_apply(nt, X, lambda b: c.inverse(b, variant))
    assert np.array_equal(y, x)


@pytest.mark.parametrize("log_n", list(range(1, 21)))
@pytest.mark.parametrize("q", [BB, P30])
def test_cyclic_vs_oracle(nt, orc, cfgs, log_n, q):
    n = 1 << log_n
    if (q - 1) % n:
        # no n-th root of unity: ntt_init refuses the modulus with a status
        with pytest.raises(nt.NttError) as e:
            nt.Config(n, q)
        assert e.value.status == 2  # NTT_ERR_NO_ROOT
        return
    c = cfgs(n, q)
    x = orc.random_poly(n, q, 1000 + log_n)
    assert np.array_equal(gpu_apply(nt, x, c.forward), orc.ntt(x, q))
    assert np.array_equal(gpu_apply(nt, x, c.inverse), orc.ntt(x, q, inverse=True))


@pytest.mark.parametrize("log_n", list(range(1, 21)))
@pytest.mark.parametrize("q", [BB, P30])
def test_negacyclic_vs_oracle(nt, orc, cfgs, log_n, q):
    n = 1 << log_n
    if (q - 1) % (2 * n):
        # no 2n-th root: the cyclic config may exist, the negacyclic transform
        # is refused with a status and leaves the data untouched
        if (q - 1) % n == 0:
            x = orc.random_poly(n, q, 2000 + log_n)
            buf = nt.DeviceBuffer.from_array(x)
            with pytest.raises(nt.NttError) as e:
                cfgs(n, q).forward_negacyclic(buf)
            assert e.value.status == 2  # NTT_ERR_NO_ROOT
            nt.device_sync()
            assert np.array_equal(buf.download(U64, n), x)
            buf.free()
        return
    c = cfgs(n, q)
    x = orc.random_poly(n, q, 2000 + log_n)
    assert np.array_equal(gpu_apply(nt, x, c.forward_negacyclic), orc.ntt_negacyclic(x, q))
    assert np.array_equal(gpu_apply(nt, x, c.inverse_negacyclic), orc.ntt_negacyclic(x, q, inverse=True))


@pytest.mark.parametrize("log_n", [1, 2, 5, 8, 11, 13, 14, 15, 17, 18])
def test_62bit_vs_oracle(nt, orc, cfgs, log_n):
    n = 1 << log_n
    c = cfgs(n, P62)
    x = orc.random_poly(n, P62, 3000 + log_n)
    assert np.array_equal(gpu_apply(nt, x, c.forward), orc.ntt(x, P62))
    assert np.array_equal(gpu_apply(nt, x, c.inverse), orc.ntt(x, P62, inverse=True))
    assert np.array_equal(gpu_apply(nt, x, c.forward_negacyclic), orc.ntt_negacyclic(x, P62))
    assert np.array_equal(gpu_apply(nt, x, c.inverse_negacyclic), orc.ntt_negacyclic(x, P62, inverse=True))


@pytest.mark.parametrize("n,head", [
    (64, [2080, 945746745, 44582636, 962587333, 472683146]),
    (4096, [8390656, 552559273, 1526781732, 744143007, 921338756]),
])
def test_kat_ramp(nt, cfgs, n, head):
    X = gpu_apply(nt, np.arange(1, n + 1, dtype=U64), cfgs(n, BB).forward)
    assert list(map(int, X[:5])) == head


def test_kat_impulse_and_constant(nt, cfgs):
    n = 1 << 16
    d = np.zeros(n, dtype=U64)
    d[0] = 1
    assert np.all(gpu_apply(nt, d, cfgs(n, BB).forward) == 1)
    X = gpu_apply(nt, np.full(n, 5, dtype=U64), cfgs(n, BB).forward)
    assert int(X[0]) == 5 * n % BB and not X[1:].any()


def _powers(r, n, q):
    """r^j for j < n, vectorised by doubling (uint64 products of 31-bit values)."""
    p = np.ones(n, dtype=U64)
    m, rm = 1, r % q
    while m < n:
        k = min(m, n - m)
        p[m:m + k] = (p[:k] * U64(rm)) % U64(q)
        rm = rm * rm % q
        m *= 2
    return p


def _nega_coeff(a, k, q, psi):
    """A_k = sum_j a_j psi^(j(2k+1)) mod q, directly (O(n))."""
    p = _powers(pow(psi, 2 * k + 1, q), a.size, q)
    return int(((a * p) % U64(q)).sum(dtype=U64) % U64(q)) if a.size <= (1 << 24) else None


@pytest.mark.parametrize("log_n", [22, 24])
def test_large_negacyclic_properties(nt, orc, cfgs, log_n):
    """cfg2 shape: forward then inverse is the identity; sampled outputs equal
    the defining sum; the transform is linear."""
    n = 1 << log_n
    c = cfgs(n, BB)
    a = orc.random_poly(n, BB, 20 + log_n)
    A = gpu_apply(nt, a, c.forward_negacyclic)
    psi = orc.psi(BB, n)
    for k in (0, 1, 12345, n // 2 + 7, n - 1):
        assert int(A[k]) == _nega_coeff(a, k, BB, psi), k
    assert np.array_equal(gpu_apply(nt, A, c.inverse_negacyclic), a)
    b = orc.random_poly(n, BB, 99)
    B = gpu_apply(nt, b, c.forward_negacyclic)
    S = gpu_apply(nt, (a + b) % U64(BB), c.forward_negacyclic)
    assert np.array_equal(S, (A + B) % U64(BB))


def test_large_cyclic_vs_oracle_2_22(nt, orc, cfgs):
    n = 1 << 22
    x = orc.random_poly(n, BB, 22)
    assert np.array_equal(gpu_apply(nt, x, cfgs(n, BB).forward), orc.ntt(x, BB))


@pytest.mark.parametrize("n,batch", [(64, 5), (4096, 3), (1 << 13, 7), (1 << 18, 3)])
def test_batch_entry_points(nt, orc, cfgs, n, batch):
    c = cfgs(n, BB)
    xs = np.concatenate([orc.random_poly(n, BB, 50 + i) for i in range(batch)])
    X = gpu_apply(nt, xs, lambda b: c.forward_batch(b, batch))
    for i in range(batch):
        assert np.array_equal(X[i * n:(i + 1) * n], orc.ntt(xs[i * n:(i + 1) * n], BB)), i
    Y = gpu_apply(nt, X, lambda b: c.inverse_batch(b, batch))
    assert np.array_equal(Y, xs)


@pytest.mark.parametrize("q", [BB, P62])
def test_poly_multiply_pointwise(nt, orc, cfgs, q):
    n = 1 << 12
    c = cfgs(n, q)
    a, b = orc.random_poly(n, q, 1), orc.random_poly(n, q, 2)
    da, db = nt.DeviceBuffer.from_array(a), nt.DeviceBuffer.from_array(b)
    dr = nt.DeviceBuffer(n * 8)
    c.poly_multiply(dr, da, db)
    nt.device_sync()
    assert np.array_equal(dr.download(U64, n), orc.pointwise_mul(a, b, q))


@pytest.mark.parametrize("q,n", [(BB, 256), (BB, 1 << 12), (P62, 512)])
def test_negacyclic_multiply_vs_schoolbook(nt, orc, cfgs, q, n):
    c = cfgs(n, q)
    a, b = orc.random_poly(n, q, 42), orc.random_poly(n, q, 123)
    da, db = nt.DeviceBuffer.from_array(a), nt.DeviceBuffer.from_array(b)
    dr = nt.DeviceBuffer(n * 8)
    c.negacyclic_multiply(dr, da, db)
    nt.device_sync()
    assert np.array_equal(dr.download(U64, n), orc.negacyclic_schoolbook(a, b, q))


# The product runs fwd(a), fwd(b), then an inverse whose first pass loads
# fwd(a)*fwd(b): sizes with one, two and three passes, 32- and 62-bit primes.
@pytest.mark.parametrize("q,log_n", [(BB, 10), (BB, 11), (BB, 12), (BB, 13), (BB, 14), (BB, 16), (BB, 20),
                                     (P30, 18), (P62, 11), (P62, 12), (P62, 15)])
@pytest.mark.parametrize("alias", [None, "a", "b"])
def test_negacyclic_multiply_vs_oracle_pipeline(nt, orc, cfgs, q, log_n, alias):
    n = 1 << log_n
    c = cfgs(n, q)
    a, b = orc.random_poly(n, q, 7 + log_n), orc.random_poly(n, q, 91 + log_n)
    fa, fb = orc.ntt_negacyclic(a, q), orc.ntt_negacyclic(b, q)
    want = orc.ntt_negacyclic(orc.pointwise_mul(fa, fb, q), q, inverse=True)
    da, db = nt.DeviceBuffer.from_array(a), nt.DeviceBuffer.from_array(b)
    dr = {"a": da, "b": db}.get(alias) or nt.DeviceBuffer(n * 8)
    c.negacyclic_multiply(dr, da, db)
    nt.device_sync()
    assert np.array_equal(dr.download(U64, n), want)
    if alias != "a":
        assert np.array_equal(da.download(U64, n), fa)
    if alias != "b":
        assert np.array_equal(db.download(U64, n), fb)
    for buf in {id(x): x for x in (da, db, dr)}.values():
        buf.free()


@pytest.mark.parametrize("q,log_n", [(BB, 10), (BB, 16), (P62, 12)])
def test_negacyclic_square(nt, orc, cfgs, q, log_n):
    # d_a == d_b: one forward transform, then a*a
    n = 1 << log_n
    a = orc.random_poly(n, q, 33 + log_n)
    fa = orc.ntt_negacyclic(a, q)
    want = orc.ntt_negacyclic(orc.pointwise_mul(fa, fa, q), q, inverse=True)
    if log_n <= 12:
        assert np.array_equal(want, orc.negacyclic_schoolbook(a, a, q))
    da, dr = nt.DeviceBuffer.from_array(a), nt.DeviceBuffer(n * 8)
    cfgs(n, q).negacyclic_multiply(dr, da, da)
    nt.device_sync()
    assert np.array_equal(dr.download(U64, n), want)
    assert np.array_equal(da.download(U64, n), fa)
    da.free()
    dr.free()


# a and b inside one allocation: the two forward transforms run as a batch
# whose polynomial stride is b - a (even, odd or negative, in elements).
@pytest.mark.parametrize("q,log_n", [(BB, 12), (BB, 14), (BB, 20), (P62, 12), (P62, 13)])
@pytest.mark.parametrize("gap", [0, 1, 6, -3])
def test_negacyclic_multiply_strides(nt, orc, cfgs, q, log_n, gap):
    n = 1 << log_n
    a, b = orc.random_poly(n, q, 5 + log_n), orc.random_poly(n, q, 8 + log_n)
    fa, fb = orc.ntt_negacyclic(a, q), orc.ntt_negacyclic(b, q)
    want = orc.ntt_negacyclic(orc.pointwise_mul(fa, fb, q), q, inverse=True)
    oa, ob = (0, n + gap) if gap >= 0 else (n - gap, 0)
    host = np.zeros(2 * n + abs(gap), dtype=U64)
    host[oa:oa + n], host[ob:ob + n] = a, b
    buf = nt.DeviceBuffer.from_array(host)
    dr = nt.DeviceBuffer(n * 8)

    class At:
        def __init__(self, off):
            self.ptr = buf.ptr + off * 8
    cfgs(n, q).negacyclic_multiply(dr, At(oa), At(ob))
    nt.device_sync()
    assert np.array_equal(dr.download(U64, n), want)
    got = buf.download(U64, host.size)
    assert np.array_equal(got[oa:oa + n], fa) and np.array_equal(got[ob:ob + n], fb)
    buf.free()
    dr.free()


def test_negacyclic_multiply_rejects_overlap(nt, orc, cfgs):
    n = 1 << 12
    buf = nt.DeviceBuffer.from_array(orc.random_poly(2 * n, BB, 3))
    dr = nt.DeviceBuffer(n * 8)

    class View:  # b = a + 1000 elements: the two inputs overlap
        ptr = buf.ptr + 1000 * 8
    with pytest.raises(nt.NttError):
        cfgs(n, BB).negacyclic_multiply(dr, buf, View())
    nt.device_sync()
    buf.free()
    dr.free()


# Size-independent: multiplying by c*x^k shifts a by k places negacyclically
# (coefficients pushed past x^(n-1) return negated) and scales by c.
@pytest.mark.parametrize("q,log_n,k", [(BB, 22, 1), (BB, 24, (1 << 23) + 12345), (P62, 20, 777)])
def test_negacyclic_multiply_by_monomial(nt, orc, cfgs, q, log_n, k):
    n = 1 << log_n
    a = orc.random_poly(n, q, 5)
    b = np.zeros(n, dtype=U64)
    cst = 3
    b[k] = cst
    da, db = nt.DeviceBuffer.from_array(a), nt.DeviceBuffer.from_array(b)
    cfgs(n, q).negacyclic_multiply(da, da, db)
    nt.device_sync()
    got = da.download(U64, n)
    da.free()
    db.free()
    # r[j] = c*a[j-k] for j >= k, -c*a[j-k+n] for j < k
    sa = np.roll(a, k).astype(object) * cst % q
    sa[:k] = (q - sa[:k]) % q
    assert np.array_equal(got, sa.astype(U64))


@pytest.mark.parametrize("log_n", [1, 6, 13, 20])
def test_bit_reverse(nt, orc, log_n):
    n = 1 << log_n
    x = orc.random_poly(n, BB, 8)
    y = gpu_apply(nt, x, lambda b: nt.bit_reverse(b, n))
    rev = np.array([int(format(i, f"0{log_n}b")[::-1], 2) for i in range(n)]) if log_n <= 13 else None
    if rev is not None:
        assert np.array_equal(y, x[rev])
    assert np.array_equal(gpu_apply(nt, y, lambda b: nt.bit_reverse(b, n)), x)


@pytest.mark.parametrize("log_n", [12, 15, 16])
def test_rowmaj_pretransposed(nt, orc, cfgs, log_n):
    n = 1 << log_n
    n1, n2 = 1 << ((log_n + 1) // 2), 1 << (log_n // 2)
    x = orc.random_poly(n, BB, 4)
    src = nt.DeviceBuffer.from_array(x)
    pre = nt.DeviceBuffer(n * 8)
    nt.rowmaj_to_pretransposed(pre, src, n1, n2)
    nt.device_sync()
    assert np.array_equal(pre.download(U64, n), x.reshape(n1, n2).T.reshape(-1))
    cfgs(n, BB).forward_rowmaj(pre)
    back = nt.DeviceBuffer(n * 8)
    nt.rowmaj_from_pretransposed(back, pre, n1, n2)
    nt.device_sync()
    assert np.array_equal(back.download(U64, n), orc.ntt(x, BB))


def test_config_fields_match_reference_init(nt, orc, cfgs):
    n = 1 << 10
    c = cfgs(n, BB)
    s = c.ptr.contents
    assert s.n == n and s.log_n == 10 and s.q == BB and s.mu == (1 << 64) // BB
    p = orc.params(BB, n)
    assert s.params.omega == p.omega and s.params.n_inv == p.n_inv
    out = np.empty(n, dtype=U64)
    nt.check(nt.lib().ntt_c_memcpy_d2h(out.ctypes.data, s.d_twiddles, n * 8))
    assert list(map(int, out[:4])) == [1, p.omega, p.omega * p.omega % BB, pow(p.omega, 3, BB)]
    assert int(out[n - 1]) == pow(p.omega, n - 1, BB)
    nt.check(nt.lib().ntt_c_memcpy_d2h(out.ctypes.data, s.d_twiddles_inv, n * 8))
    assert int(out[5]) == pow(p.omega_inv, 5, BB)
