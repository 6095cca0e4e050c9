"""GPU parity of the descriptor batch API: mixed sizes, primes and
directions in one prepared batch (plain launch and hipGraph replay), RNS
limbs, and the same result for every polynomial alone and in the mix."""
import numpy as np
import pytest

from workloads import rns_primes, uneven_batch

pytestmark = pytest.mark.gpu

BB = 2013265921


@pytest.fixture(autouse=True)
def _plan_table_cleared(nt):
    """Every test starts and ends with an empty plan table, also when it fails
    (a table a test installs must not steer the plans of later tests)."""
    nt.autotune_clear()
    yield
    nt.autotune_clear()


def _mfma_plans(nt, tmp_path, log_ns, elems):
    """Installs a plan table that runs every size of a batch of `elems`
    elements with radix-16 rounds on the matrix cores (the table keys plans
    by log2 n and log2 of the batch's elements): one pass up to 2^14, else
    [8, log n - 8]."""
    key = int(elems).bit_length() - 1
    lines = []
    for ln in sorted(set(log_ns)):
        if ln <= 14:
            lines.append(f"0 {ln} {key} 1 {ln} 0 0 0 {max(ln, 13)} 1")
        else:
            lines.append(f"0 {ln} {key} 2 8 {ln - 8} 0 0 13 3")
    f = tmp_path / "mfma_plans.txt"
    f.write_text("\n".join(lines) + "\n")
    nt.autotune_clear()
    nt.autotune_load(str(f))


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
@pytest.mark.parametrize("mf", [False, True])
def test_uneven_batch_matches_oracle_and_graph_replays(nt, orc, u64, mf, tmp_path):
    jobs = uneven_batch(count=48, seed=4, log_min=1, log_max=16)
    if mf:
        _mfma_plans(nt, tmp_path, [j["log_n"] for j in jobs], sum(1 << j["log_n"] for j in jobs))
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
    nt.autotune_clear()


@pytest.mark.parametrize("mf", [False, True])
def test_alone_equals_in_mix(nt, orc, mf, tmp_path):
    jobs = uneven_batch(count=24, seed=11, log_min=6, log_max=17)
    if mf:
        # the same splits for the batch and for each polynomial alone (a
        # single transform keys its plan by its own size)
        log_ns = [j["log_n"] for j in jobs]
        _mfma_plans(nt, tmp_path, log_ns, sum(1 << ln for ln in log_ns))
        key = sum(1 << ln for ln in log_ns).bit_length() - 1
        extra = tmp_path / "alone.txt"
        extra.write_text("".join(f"0 {ln} {ln} " + ("1 %d 0 0 0 %d 1\n" % (ln, max(ln, 13)) if ln <= 14 else
                                                    "2 8 %d 0 0 13 3\n" % (ln - 8)) for ln in sorted(set(log_ns))
                                 if ln != key))
        nt.autotune_load(str(extra))
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
        assert np.array_equal(alone, orc.transform(p, j["q"], j["nega"], j["inverse"])), i
        one.free()
    batch.destroy()
    ctx.destroy()
    nt.autotune_clear()


@pytest.mark.parametrize("log_n,mf", [(12, False), (12, True), (16, True)])
def test_rns_limbs_one_batch(nt, orc, log_n, mf, tmp_path):
    """cfg3 shape (24 limbs: the 24 largest primes < 2^31 with q = 1 mod
    2^17), negacyclic, one prepared batch: 24 different matrix blocks in one
    launch on the matrix-core path."""
    primes = rns_primes()
    if mf:
        _mfma_plans(nt, tmp_path, [log_n], len(primes) << log_n)
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
    nt.autotune_clear()
