"""GPU parity of the descriptor batch API: mixed sizes, primes and
directions in one prepared batch (plain launch and hipGraph replay), RNS
limbs, and the same result for every polynomial alone and in the mix."""
import numpy as np
import pytest

from workloads import rns_primes, uneven_batch

pytestmark = pytest.mark.gpu

BB = 2013265921


def _upload_polys(nt, polys, u64):
    dt = np.uint64 if u64 else np.uint32
    sizes = [p.size * (8 if u64 else 4) for p in polys]
    offs = np.concatenate([[0], np.cumsum([(s + 255) // 256 * 256 for s in sizes])]).astype(np.int64)
    buf = nt.DeviceBuffer(int(offs[-1]))
    for p, o in zip(polys, offs):
        buf.upload(p.astype(dt), int(o))
    return buf, offs


def _download_polys(nt, buf, offs, sizes, u64):
    dt = np.uint64 if u64 else np.uint32
    return [buf.download(dt, n, int(o)).astype(np.uint64) for n, o in zip(sizes, offs)]


@pytest.mark.parametrize("u64", [False, True])
def test_uneven_batch_matches_oracle_and_graph_replays(nt, orc, u64):
    jobs = uneven_batch(count=48, seed=4, log_min=1, log_max=16)
    polys = [orc.random_poly(1 << j["log_n"], j["q"], j["seed"]) for j in jobs]
    expect = [orc.transform(p, j["q"], j["nega"], j["inverse"]) for p, j in zip(polys, jobs)]
    buf, offs = _upload_polys(nt, polys, u64)
    ctx = nt.Context()
    flags = [(nt.FLAG_INVERSE if j["inverse"] else 0) | (nt.FLAG_NEGACYCLIC if j["nega"] else 0) |
             (nt.FLAG_U64 if u64 else 0) for j in jobs]
    descs = [(buf.ptr + int(o), j["q"], j["log_n"], f) for j, o, f in zip(jobs, offs, flags)]
    batch = nt.Batch(ctx, descs)
    info = batch.info()
    # one launch per (pass, storage kind, post-multiply kind): at most 2 passes
    # here (log n <= 16), 3 storage pairs with u64 I/O, 2 post kinds
    assert info["launches"] <= 12 and info["launches"] < len(jobs)
    batch.launch()
    nt.device_sync()
    got = _download_polys(nt, buf, offs, [p.size for p in polys], u64)
    for i, (g, e) in enumerate(zip(got, expect)):
        assert np.array_equal(g, e), (i, jobs[i])
    # graph replays: re-upload inputs between replays (eager work in between)
    for rep in range(3):
        for p, o in zip(polys, offs):
            buf.upload(p.astype(np.uint64 if u64 else np.uint32), int(o))
        batch.launch_graph()
        nt.device_sync()
        got = _download_polys(nt, buf, offs, [p.size for p in polys], u64)
        for i, (g, e) in enumerate(zip(got, expect)):
            assert np.array_equal(g, e), (rep, i)
    batch.destroy()
    ctx.destroy()


def test_alone_equals_in_mix(nt, orc):
    jobs = uneven_batch(count=24, seed=11, log_min=6, log_max=17)
    polys = [orc.random_poly(1 << j["log_n"], j["q"], j["seed"]) for j in jobs]
    ctx = nt.Context()
    flags = [(nt.FLAG_INVERSE if j["inverse"] else 0) | (nt.FLAG_NEGACYCLIC if j["nega"] else 0) for j in jobs]
    buf, offs = _upload_polys(nt, polys, False)
    batch = nt.Batch(ctx, [(buf.ptr + int(o), j["q"], j["log_n"], f) for j, o, f in zip(jobs, offs, flags)])
    batch.launch()
    nt.device_sync()
    mixed = _download_polys(nt, buf, offs, [p.size for p in polys], False)
    for i, (p, j, f) in enumerate(zip(polys, jobs, flags)):
        one = nt.DeviceBuffer.from_array(p.astype(np.uint32))
        ctx.transform(one.ptr, j["q"], j["log_n"], f)
        nt.device_sync()
        alone = one.download(np.uint32, p.size).astype(np.uint64)
        assert np.array_equal(alone, mixed[i]), i
        one.free()
    batch.destroy()
    ctx.destroy()


def test_rns_limbs_one_batch(nt, orc):
    """cfg3 shape at reduced n: 24 limbs (the 24 largest primes < 2^31 with
    q = 1 mod 2^17), negacyclic, one prepared batch."""
    log_n = 12
    primes = rns_primes()
    polys = [orc.random_poly(1 << log_n, q, 0x5EED ^ i) for i, q in enumerate(primes)]
    buf, offs = _upload_polys(nt, polys, False)
    ctx = nt.Context()
    batch = nt.Batch(ctx, [(buf.ptr + int(o), q, log_n, nt.FLAG_NEGACYCLIC) for q, o in zip(primes, offs)])
    batch.launch()
    nt.device_sync()
    got = _download_polys(nt, buf, offs, [p.size for p in polys], False)
    for i, (g, p, q) in enumerate(zip(got, polys, primes)):
        assert np.array_equal(g, orc.ntt_negacyclic(p, q)), i
    batch.destroy()
    ctx.destroy()
