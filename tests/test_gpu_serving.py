"""Continuous-batching server: every request's result is bit-identical alone
and in the mix (and equals the oracle), requests arriving while others run
overlap on the device, and admission respects the HBM budget."""
import numpy as np
import pytest

from workloads import p62_primes, serving_requests

pytestmark = pytest.mark.gpu

BB = 2013265921


def _alloc(nt, reqs, seed_base=0):
    """Device buffers for every polynomial of every request (uint64 storage)."""
    bufs, host = [], []
    from oracle import pyoracle as orc
    for r in reqs:
        rb, rh = [], []
        for p in r:
            x = orc.random_poly(1 << p["log_n"], p["q"], p["seed"] + seed_base)
            rb.append(nt.DeviceBuffer.from_array(x))
            rh.append(x)
        bufs.append(rb)
        host.append(rh)
    return bufs, host


def _descs(nt, r, rb):
    return [(b.ptr, p["q"], p["log_n"], nt.FLAG_U64 | nt.FLAG_NEGACYCLIC | (nt.FLAG_INVERSE if p["inverse"] else 0))
            for p, b in zip(r, rb)]


@pytest.fixture(scope="module")
def reqs():
    # a small version of the serving stream: sizes 2^8..2^14, up to 4 polys each
    return [polys for _, polys in serving_requests(24, 1000.0, seed=9, log_min=8, log_max=14, max_polys=4)]


def test_alone_equals_mix_equals_oracle(nt, orc, reqs):
    ctx = nt.Context()
    srv = nt.Server(ctx, hbm_budget_bytes=1 << 28, max_inflight=3)
    # alone: one request at a time
    bufs_a, host = _alloc(nt, reqs)
    alone = []
    for r, rb in zip(reqs, bufs_a):
        rid = srv.submit(_descs(nt, r, rb))
        srv.wait(rid, 60.0)
        alone.append([b.download(np.uint64, 1 << p["log_n"]) for p, b in zip(r, rb)])
    # mix: all submitted back to back
    bufs_m, _ = _alloc(nt, reqs)
    ids = [srv.submit(_descs(nt, r, rb)) for r, rb in zip(reqs, bufs_m)]
    for rid in ids:
        srv.wait(rid, 60.0)
    st = srv.stats()
    assert st["completed"] == 2 * len(reqs) and st["failed"] == 0
    for i, (r, rb) in enumerate(zip(reqs, bufs_m)):
        for k, (p, b) in enumerate(zip(r, rb)):
            mixed = b.download(np.uint64, 1 << p["log_n"])
            assert np.array_equal(mixed, alone[i][k]), (i, k)
            exp = orc.ntt_negacyclic(host[i][k], p["q"], inverse=p["inverse"])
            assert np.array_equal(mixed, exp), (i, k)
    srv.destroy()
    ctx.destroy()


def test_requests_overlap_and_budget_admission(nt, reqs):
    ctx = nt.Context()
    # budget that fits only a few requests' workspace at a time
    srv = nt.Server(ctx, hbm_budget_bytes=1 << 20, max_inflight=4)
    bufs, _ = _alloc(nt, reqs)
    big = [(b.ptr, p62_primes()[0], 18, nt.FLAG_U64) for b in [nt.DeviceBuffer(8 << 18)]]
    with pytest.raises(nt.NttError) as e:
        srv.submit(big)  # 2 MiB of workspace > 1 MiB budget: rejected at submit
    assert e.value.status == 3
    lat = srv.replay([_descs(nt, r, rb) for r, rb in zip(reqs, bufs)], [0.0005 * i for i in range(len(reqs))])
    st = srv.stats()
    assert st["completed"] == len(reqs) and all(v > 0 for v in lat)
    assert st["arena_free_bytes"] == 1 << 20
    srv.destroy()
    ctx.destroy()


def test_server_queue_under_load(nt):
    """Burst of requests far above one batch: all complete, several batches ran
    concurrently, latency stats are consistent."""
    ctx = nt.Context()
    srv = nt.Server(ctx, hbm_budget_bytes=1 << 30, max_inflight=4, max_batch_polys=64)
    bufs = [nt.DeviceBuffer(4 << 16) for _ in range(96)]
    for b in bufs:
        nt.lib().ntt_c_memset(b.ptr, 1, 4 << 16)
    reqs = [[(b.ptr, BB, 16, nt.FLAG_NEGACYCLIC)] for b in bufs]
    lat = srv.replay(reqs, [0.0] * len(reqs))
    st = srv.stats()
    assert st["completed"] == 96 and st["batches"] >= 2
    assert st["p50_us"] <= st["p99_us"] and max(lat) >= st["p99_us"] * 0.999
    srv.destroy()
    ctx.destroy()
