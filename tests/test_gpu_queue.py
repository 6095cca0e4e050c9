"""GPU parity of the single work-queue launch of a descriptor batch
(queue_kernel): every pass of every polynomial in one persistent launch,
a polynomial's next pass gated by device counters.  Each case runs the batch
both ways (one work-queue launch, one launch per phase group) and compares
both with the oracle bit for bit, including hipGraph replays with eager work
between them (the kernel clears its own counters for the next launch)."""
import numpy as np
import pytest

from workloads import rns_primes, uneven_batch

pytestmark = pytest.mark.gpu

BB = 2013265921


@pytest.fixture(autouse=True)
def _plan_table_cleared(nt):
    nt.autotune_clear()
    yield
    nt.autotune_clear()


def _flags(nt, j):
    return (nt.FLAG_INVERSE if j["inverse"] else 0) | (nt.FLAG_NEGACYCLIC if j["nega"] else 0)


def _upload(nt, polys):
    sizes = [p.size * 4 for p in polys]
    offs = np.concatenate([[0], np.cumsum([(s + 255) // 256 * 256 for s in sizes])]).astype(np.int64)
    buf = nt.DeviceBuffer(int(offs[-1]))
    for p, o in zip(polys, offs):
        buf.upload(p.astype(np.uint32), int(o))
    return buf, offs


def _reupload(buf, polys, offs):
    for p, o in zip(polys, offs):
        buf.upload(p.astype(np.uint32), int(o))


def _download(nt, buf, offs, polys):
    return [buf.download(np.uint32, p.size, int(o)).astype(np.uint64) for p, o in zip(polys, offs)]


def _mfma_table(nt, tmp_path, log_ns, elems):
    """Every size of the batch on the matrix cores: one pass up to 2^12,
    else [6 or 8, rest] (the queue needs one round kind for all jobs, and
    matrix-core rounds need tiles of at least 2^8 elements)."""
    key = int(elems).bit_length() - 1
    lines = []
    for ln in sorted(set(log_ns)):
        if ln <= 12:
            lines.append(f"0 {ln} {key} 1 {ln} 0 0 0 {max(ln, 12)} 1")
        elif ln <= 16:
            lines.append(f"0 {ln} {key} 2 6 {ln - 6} 0 0 12 3")
        else:
            lines.append(f"0 {ln} {key} 2 8 {ln - 8} 0 0 13 3")
    f = tmp_path / "queue_plans.txt"
    f.write_text("\n".join(lines) + "\n")
    nt.autotune_load(str(f))


def _run_both(nt, orc, jobs, polys, graph_reps=2):
    expect = [orc.transform(p, j["q"], j["nega"], j["inverse"]) for p, j in zip(polys, jobs)]
    buf, offs = _upload(nt, polys)
    ctx = nt.Context()
    batch = nt.Batch(ctx, [(buf.ptr + int(o), j["q"], j["log_n"], _flags(nt, j)) for j, o in zip(jobs, offs)])
    try:
        assert batch.info()["launches"] >= 2  # the default form: one launch per phase group
        # the work queue with the device's hand-over (one XCD per polynomial
        # where the probe allows it), the phase launches, the work queue with
        # agent-scope fences
        for mode in ("queue", "phases", "agent"):
            batch.set_queue(mode != "phases", agent_fences=mode == "agent")
            info = batch.info()
            assert (info["launches"] == 1) == (mode != "phases")
            _reupload(buf, polys, offs)
            batch.launch()
            batch.check()
            got = _download(nt, buf, offs, polys)
            for i, (g, e) in enumerate(zip(got, expect)):
                assert np.array_equal(g, e), (mode, i, jobs[i])
            for rep in range(graph_reps):
                _reupload(buf, polys, offs)
                batch.launch_graph()
                batch.check()
                got = _download(nt, buf, offs, polys)
                for i, (g, e) in enumerate(zip(got, expect)):
                    assert np.array_equal(g, e), ("graph", mode, rep, i)
    finally:
        batch.destroy()
        ctx.destroy()
        buf.free()


@pytest.mark.parametrize("seed", [4, 9])
def test_uneven_batch_one_launch(nt, orc, seed, tmp_path):
    """Uneven sizes 2^8..2^17, BabyBear and 30-bit primes, both directions,
    cyclic and negacyclic: one- and two-pass plans in one work queue."""
    jobs = uneven_batch(count=40, seed=seed, log_min=8, log_max=17)
    _mfma_table(nt, tmp_path, [j["log_n"] for j in jobs], sum(1 << j["log_n"] for j in jobs))
    polys = [orc.random_poly(1 << j["log_n"], j["q"], j["seed"]) for j in jobs]
    _run_both(nt, orc, jobs, polys)


def test_valu_rounds_one_launch(nt, orc, tmp_path):
    """The same with VALU rounds only (the other queue-kernel instance)."""
    jobs = uneven_batch(count=24, seed=5, log_min=8, log_max=16)
    key = sum(1 << j["log_n"] for j in jobs).bit_length() - 1
    lines = [f"0 {ln} {key} 2 {ln // 2} {ln - ln // 2} 0 0 12 0" for ln in sorted({j["log_n"] for j in jobs})]
    f = tmp_path / "valu.txt"
    f.write_text("\n".join(lines) + "\n")
    nt.autotune_load(str(f))
    polys = [orc.random_poly(1 << j["log_n"], j["q"], j["seed"]) for j in jobs]
    _run_both(nt, orc, jobs, polys, graph_reps=1)


def test_three_pass_plans_one_launch(nt, orc, tmp_path):
    """Three-pass plans: each pass waits for the one before it."""
    jobs = [dict(log_n=ln, q=BB, nega=bool(i & 1), inverse=bool(i & 2), seed=70 + i)
            for i, ln in enumerate([15, 16, 17, 18, 15, 18, 16, 17, 18, 12])]
    key = sum(1 << j["log_n"] for j in jobs).bit_length() - 1
    lines = [f"0 {ln} {key} 3 {ln // 3} {ln // 3} {ln - 2 * (ln // 3)} 0 12 7" for ln in (15, 16, 17, 18)]
    lines.append(f"0 12 {key} 1 12 0 0 0 12 1")
    f = tmp_path / "three.txt"
    f.write_text("\n".join(lines) + "\n")
    nt.autotune_load(str(f))
    polys = [orc.random_poly(1 << j["log_n"], j["q"], j["seed"]) for j in jobs]
    _run_both(nt, orc, jobs, polys, graph_reps=1)


def test_rns_limbs_one_launch(nt, orc, tmp_path):
    """The 24-limb RNS shape (24 primes q = 1 mod 2^17, n = 2^16, negacyclic)."""
    primes = rns_primes()
    log_n = 16
    _mfma_table(nt, tmp_path, [log_n], len(primes) << log_n)
    jobs = [dict(log_n=log_n, q=q, nega=True, inverse=False, seed=0x5EED ^ i) for i, q in enumerate(primes)]
    polys = [orc.random_poly(1 << log_n, q, 0x5EED ^ i) for i, q in enumerate(primes)]
    _run_both(nt, orc, jobs, polys)


def test_mixed_kinds_keep_phase_launches(nt, tmp_path):
    """A mix that needs two kernel instantiations (uint32 and uint64
    storage) has no work-queue form: the batch keeps its phase launches."""
    ctx = nt.Context()
    a = nt.DeviceBuffer(4 << 12)
    b = nt.DeviceBuffer(8 << 12)
    batch = nt.Batch(ctx, [(a.ptr, BB, 12, 0), (b.ptr, BB, 12, nt.FLAG_U64)])
    try:
        assert batch.info()["launches"] >= 2
        with pytest.raises(nt.NttError) as e:
            batch.set_queue(True)
        assert e.value.status == 5
        batch.set_queue(False)
        batch.launch()
        batch.check()
    finally:
        batch.destroy()
        ctx.destroy()
        a.free()
        b.free()


def test_wait_limit_reports_and_recovers(nt, orc, tmp_path):
    """A wait limit of one poll (forced from the host) makes dependent tiles
    give up instead of waiting: every workgroup still leaves, ntt_c_batch_check
    reports it (if any wait gave up), and the next launch with the normal limit
    is exact again (the counters were cleared)."""
    jobs = uneven_batch(count=32, seed=21, log_min=13, log_max=17)
    _mfma_table(nt, tmp_path, [j["log_n"] for j in jobs], sum(1 << j["log_n"] for j in jobs))
    polys = [orc.random_poly(1 << j["log_n"], j["q"], j["seed"]) for j in jobs]
    expect = [orc.transform(p, j["q"], j["nega"], j["inverse"]) for p, j in zip(polys, jobs)]
    buf, offs = _upload(nt, polys)
    ctx = nt.Context()
    batch = nt.Batch(ctx, [(buf.ptr + int(o), j["q"], j["log_n"], _flags(nt, j)) for j, o in zip(jobs, offs)])
    try:
        batch.set_queue(True, spin_limit=1)
        batch.launch()
        try:
            batch.check()
        except nt.NttError as e:
            assert e.status == 4
        batch.check()  # the flag was cleared
        batch.set_queue(True)
        _reupload(buf, polys, offs)
        batch.launch()
        batch.check()
        got = _download(nt, buf, offs, polys)
        for i, (g, e) in enumerate(zip(got, expect)):
            assert np.array_equal(g, e), i
    finally:
        batch.destroy()
        ctx.destroy()
        buf.free()
