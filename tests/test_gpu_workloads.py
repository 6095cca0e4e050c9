"""The benchmark workloads at their stated sizes, with the shipped plan table,
against the oracle (bit-exact):

  * the RNS batch: n = 2^16 negacyclic, 24 limbs (the 24 largest primes
    q < 2^31 with q = 1 mod 2^17), one prepared batch, graph replays;
  * the uneven batch: 512 polynomials, log2 n uniform in {10..18}, BabyBear or
    a 30-bit prime q = 1 mod 2^19, random direction, one prepared batch
    replayed as a graph, and the same polynomials one launch chain each;
  * 62-bit negacyclic transforms at 2^19..2^22 (the serving stream's 62-bit
    sizes above what the other tests cover);
  * the serving stream with its real size law (log2 n in [12, 22], 1..32
    polynomials, half of them 62-bit), including its largest request (31
    polynomials of 2^22 with a 62-bit prime): each result is the same in the
    mix, alone, and in the oracle.
"""
import os

import numpy as np
import pytest

import workloads as W

pytestmark = pytest.mark.gpu

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
TABLE = os.path.join(REPO, "ntt-tensorcore_amd", "tuned_mi355x.txt")


@pytest.fixture
def shipped_table(nt):
    nt.autotune_clear()
    nt.autotune_load(TABLE)
    yield
    nt.autotune_clear()


def _pack(nt, polys, esz):
    offs, tot = [], 0
    for p in polys:
        offs.append(tot)
        tot += (p.size * esz + 255) // 256 * 256
    buf = nt.DeviceBuffer(tot)
    return buf, offs


def _upload(buf, polys, offs, dt):
    for p, o in zip(polys, offs):
        buf.upload(p.astype(dt), o)


def _flags(nt, j, u64=False):
    return ((nt.FLAG_INVERSE if j["inverse"] else 0) | (nt.FLAG_NEGACYCLIC if j["nega"] else 0) |
            (nt.FLAG_U64 if u64 else 0))


def test_rns_24_limbs_2_16_graph(nt, orc, shipped_table):
    primes = W.rns_primes()
    assert len(primes) == 24 and primes[0] == 2147352577
    polys = [orc.random_poly(1 << 16, q, 0x5EED ^ i) for i, q in enumerate(primes)]
    expect, _ = orc.transform_batch(polys, primes, [True] * 24, [False] * 24)
    buf, offs = _pack(nt, polys, 4)
    ctx = nt.Context()
    bt = nt.Batch(ctx, [(buf.ptr + o, q, 16, nt.FLAG_NEGACYCLIC) for q, o in zip(primes, offs)])
    try:
        for rep in range(2):
            _upload(buf, polys, offs, np.uint32)
            bt.launch_graph()
            nt.device_sync()
            for i, (o, e) in enumerate(zip(offs, expect)):
                assert np.array_equal(buf.download(np.uint32, 1 << 16, o).astype(np.uint64), e), (rep, i)
    finally:
        bt.destroy()
        ctx.destroy()
        buf.free()


@pytest.fixture(scope="module")
def uneven(orc):
    jobs = W.uneven_batch()
    assert len(jobs) == 512 and {j["log_n"] for j in jobs} == set(range(10, 19))
    polys = [W.random_poly(1 << j["log_n"], j["q"], j["seed"]) for j in jobs]
    expect, _ = orc.transform_batch(polys, [j["q"] for j in jobs], [j["nega"] for j in jobs],
                                    [j["inverse"] for j in jobs])
    return jobs, polys, expect


def test_uneven_512_graph_matches_oracle(nt, uneven, shipped_table):
    jobs, polys, expect = uneven
    buf, offs = _pack(nt, polys, 4)
    ctx = nt.Context()
    bt = nt.Batch(ctx, [(buf.ptr + o, j["q"], j["log_n"], _flags(nt, j)) for j, o in zip(jobs, offs)])
    try:
        info = bt.info()
        assert info["launches"] < 16
        for rep in range(2):  # two replays of the captured graph, eager uploads between them
            _upload(buf, polys, offs, np.uint32)
            bt.launch_graph()
            nt.device_sync()
            for i, (j, o, e) in enumerate(zip(jobs, offs, expect)):
                got = buf.download(np.uint32, 1 << j["log_n"], o).astype(np.uint64)
                assert np.array_equal(got, e), (rep, i, j)
    finally:
        bt.destroy()
        ctx.destroy()
        buf.free()


def test_uneven_512_one_chain_per_polynomial(nt, uneven, shipped_table):
    """The per-polynomial path the batch is measured against: the same 512
    polynomials, one transform call (one launch chain) each."""
    jobs, polys, expect = uneven
    buf, offs = _pack(nt, polys, 4)
    ctx = nt.Context()
    try:
        _upload(buf, polys, offs, np.uint32)
        for j, o in zip(jobs, offs):
            ctx.transform(buf.ptr + o, j["q"], j["log_n"], _flags(nt, j))
        nt.device_sync()
        for i, (j, o, e) in enumerate(zip(jobs, offs, expect)):
            assert np.array_equal(buf.download(np.uint32, 1 << j["log_n"], o).astype(np.uint64), e), (i, j)
        # the same loop issued from native code (ntt_c_transform_each)
        _upload(buf, polys, offs, np.uint32)
        ctx.transform_each([(buf.ptr + o, j["q"], j["log_n"], _flags(nt, j)) for j, o in zip(jobs, offs)])
        nt.device_sync()
        for i, (j, o, e) in enumerate(zip(jobs, offs, expect)):
            assert np.array_equal(buf.download(np.uint32, 1 << j["log_n"], o).astype(np.uint64), e), (i, j)
    finally:
        ctx.destroy()
        buf.free()


@pytest.mark.parametrize("log_n", [19, 20, 21, 22])
@pytest.mark.parametrize("q", W.p62_primes()[:2])
def test_62bit_negacyclic_large(nt, orc, shipped_table, log_n, q):
    n = 1 << log_n
    x = orc.random_poly(n, q, 62 * log_n + q % 97)
    buf = nt.DeviceBuffer.from_array(x)
    ctx = nt.Context()
    try:
        ctx.transform(buf.ptr, q, log_n, nt.FLAG_NEGACYCLIC | nt.FLAG_U64)
        nt.device_sync()
        X = buf.download(np.uint64, n)
        assert np.array_equal(X, orc.ntt_negacyclic(x, q))
        ctx.transform(buf.ptr, q, log_n, nt.FLAG_NEGACYCLIC | nt.FLAG_INVERSE | nt.FLAG_U64)
        nt.device_sync()
        assert np.array_equal(buf.download(np.uint64, n), x)
    finally:
        ctx.destroy()
        buf.free()


def _serving_set():
    """24 requests of the bench's serving stream (seed 5, real size law) plus
    its largest 62-bit request."""
    base = W.serving_requests(120, 1.0, seed=5)
    big = max(range(len(base)), key=lambda i: sum(1 << p["log_n"] for p in base[i][1]) *
              (2 if base[i][1][0]["q"] > (1 << 31) else 1))
    picked = list(range(24)) + ([big] if big >= 24 else [])
    return [base[i][1] for i in picked]


def test_serving_real_sizes_mix_alone_oracle(nt, orc):
    reqs = _serving_set()
    largest = max(reqs, key=lambda r: len(r) * (1 << r[0]["log_n"]))
    assert largest[0]["log_n"] == 22 and largest[0]["q"] > (1 << 31) and len(largest) >= 30
    flat = [p for r in reqs for p in r]
    polys = [W.random_poly(1 << p["log_n"], p["q"], p["seed"]) for p in flat]
    expect, _ = orc.transform_batch(polys, [p["q"] for p in flat], [True] * len(flat), [p["inverse"] for p in flat])
    u64 = [p["q"] > (1 << 31) for p in flat]
    offs, tot = [], 0
    for x, w in zip(polys, u64):
        offs.append(tot)
        tot += (x.size * (8 if w else 4) + 255) // 256 * 256
    pool = nt.DeviceBuffer(tot)
    descs, k = [], 0
    for r in reqs:
        d = []
        for p in r:
            d.append((pool.ptr + offs[k], p["q"], p["log_n"], nt.FLAG_NEGACYCLIC | (nt.FLAG_INVERSE if p["inverse"] else 0)
                      | (nt.FLAG_U64 if u64[k] else 0)))
            k += 1
        descs.append(d)

    def upload():
        for x, o, w in zip(polys, offs, u64):
            pool.upload(x.astype(np.uint64 if w else np.uint32), o)

    def results():
        return [pool.download(np.uint64 if w else np.uint32, x.size, o).astype(np.uint64)
                for x, o, w in zip(polys, offs, u64)]

    ctx = nt.Context()
    srv = nt.Server(ctx, hbm_budget_bytes=4 << 30, max_inflight=4, max_batch_polys=1024)
    try:
        # the mix: Poisson arrivals at 2000 requests/s, so requests overlap on the device
        upload()
        rng = np.random.default_rng(1)
        arrivals = np.cumsum(rng.exponential(1 / 2000.0, size=len(reqs))).tolist()
        srv.replay(descs, arrivals)
        st = srv.stats()
        assert st["completed"] == len(reqs) and st["failed"] == 0
        mixed = results()
        for i, (g, e) in enumerate(zip(mixed, expect)):
            assert np.array_equal(g, e), (i, flat[i]["log_n"], flat[i]["q"])
        # alone: one request at a time
        upload()
        for d in descs:
            srv.wait(srv.submit(d), 120.0)
        for i, (g, m) in enumerate(zip(results(), mixed)):
            assert np.array_equal(g, m), i
    finally:
        srv.destroy()
        ctx.destroy()
        pool.free()
