"""GPU parity of the pipelined pass kernel (pipe_kernel, NTT_OPT_PIPELINE):
passes whose tiles have a pipelined shape run as workgroups that walk several
tiles and load the next one while transforming the current one.  Every case
compares the pipelined run with the one-tile-per-workgroup run bit for bit,
and with the oracle up to 2^20 (sampled defining sums and the round trip at
2^24).  Option value 2 pipelines every eligible pass (the default, 1, only
passes with two tiles or more per workgroup slot), so that small cases reach
the pipelined kernel.  Batches of equally spaced polynomials that share a plan
run as one job (and so pipelined); other batches keep one job each."""
import numpy as np
import pytest

from props import sampled_sums_ok

pytestmark = pytest.mark.gpu

BB = 2013265921
P30 = 1056440321  # 30-bit prime, q = 1 mod 2^19


@pytest.fixture(autouse=True)
def _options_restored(nt):
    old = nt.set_option(nt.OPT_PIPELINE, 0)
    assert old == 1  # on by default
    nt.autotune_clear()
    yield
    nt.set_option(nt.OPT_PIPELINE, old)
    nt.autotune_clear()


def _run(nt, ctx, x, q, log_n, flags, split, log_e, mask, pipe):
    nt.set_option(nt.OPT_PIPELINE, pipe)
    buf = nt.DeviceBuffer.from_array(x.astype(np.uint32))
    ctx.transform_split(buf.ptr, q, log_n, flags, split, log_e, None, mfma=mask)
    nt.device_sync()
    out = buf.download(np.uint32, x.size).astype(np.uint64)
    buf.free()
    return out


# (log n, split, log2 tile elements): every pass whose shape is (8, 4),
# (8, 5), (10, 3), (9, 5), (10, 4) or (11, 3) takes the pipelined kernel; tile
# counts give 2 to 16 tiles per workgroup range
CASES = [(16, [8, 8], 12), (16, [8, 8], 13), (17, [8, 9], 12), (20, [8, 4, 8], 12), (20, [8, 4, 8], 13),
         (18, [8, 2, 8], 13), (18, [9, 9], 14), (20, [11, 9], 14), (20, [10, 10], 13), (20, [10, 10], 14)]


# the 30-bit prime has 2n | q - 1 only up to n = 2^18
PARAMS = [(ln, sp, le, q) for ln, sp, le in CASES for q in (BB, P30) if q == BB or ln <= 18]


@pytest.mark.parametrize("log_n,split,log_e,q", PARAMS)
@pytest.mark.parametrize("kind", ["fwd", "inv", "nfwd", "ninv"])
def test_pipelined_equals_plain_and_oracle(nt, orc, log_n, split, log_e, q, kind):
    inv, nega = kind in ("inv", "ninv"), kind.startswith("n")
    flags = (nt.FLAG_INVERSE if inv else 0) | (nt.FLAG_NEGACYCLIC if nega else 0)
    x = orc.random_poly(1 << log_n, q, 1000 + log_n)
    ctx = nt.Context()
    mask = (1 << len(split)) - 1
    plain = _run(nt, ctx, x, q, log_n, flags, split, log_e, mask, 0)
    piped = _run(nt, ctx, x, q, log_n, flags, split, log_e, mask, 2)
    ctx.destroy()
    assert np.array_equal(piped, plain)
    assert np.array_equal(piped, orc.transform(x, q, nega, inv))


@pytest.mark.parametrize("nega", [False, True])
def test_pipelined_2e24_headline_plan(nt, orc, nega):
    """The headline's plan [8, 8, 8] with tiles of 2^13 at 2^24: the
    pipelined forward and inverse agree with the plain ones bit for bit,
    sampled outputs match the defining sum, and the round trip is exact."""
    log_n, q = 24, BB
    x = orc.random_poly(1 << log_n, q, 24)
    ctx = nt.Context()
    fwd = nt.FLAG_NEGACYCLIC if nega else 0
    a = _run(nt, ctx, x, q, log_n, fwd, [8, 8, 8], 13, 7, 0)
    b = _run(nt, ctx, x, q, log_n, fwd, [8, 8, 8], 13, 7, 1)
    assert np.array_equal(a, b)
    assert sampled_sums_ok(b, x, q, orc.generator(q), nega, False)
    back = _run(nt, ctx, b, q, log_n, fwd | nt.FLAG_INVERSE, [8, 8, 8], 13, 7, 1)
    ctx.destroy()
    assert np.array_equal(back, x)


def test_option_reports_previous_value(nt):
    # the fixture saved the default (on) and set it off
    assert nt.set_option(nt.OPT_PIPELINE, 1) == 0
    assert nt.set_option(nt.OPT_PIPELINE, 2) == 1
    assert nt.set_option(nt.OPT_PIPELINE, 0) == 2
    with pytest.raises(nt.NttError):
        nt.set_option(99, 1)
    with pytest.raises(nt.NttError):
        nt.set_option(nt.OPT_PIPELINE, 3)
    assert nt.set_option(nt.OPT_PIPELINE, 0) == 0  # a refused value changes nothing


def _install(nt, tmp_path, log_n, count, split, log_e, mask):
    """Plan table entry for `count` polynomials of 2^log_n in one batch."""
    key = (count << log_n).bit_length() - 1
    f = tmp_path / "plans.txt"
    f.write_text(f"0 {log_n} {key} {len(split)} " + " ".join(map(str, split + [0] * (4 - len(split)))) +
                 f" {log_e} {mask}\n")
    nt.autotune_clear()
    nt.autotune_load(str(f))


def _batch_run(nt, ctx, xs, q, log_n, flags, gaps, pipe):
    """One batch launch over polynomials placed with the given byte gaps
    between them (equal gaps: one merged job); returns the outputs."""
    n = 1 << log_n
    offs, o = [], 0
    for g in gaps:
        offs.append(o)
        o += n * 4 + g
    buf = nt.DeviceBuffer(o)
    for x, off in zip(xs, offs):
        buf.upload(x.astype(np.uint32), off)
    nt.set_option(nt.OPT_PIPELINE, pipe)
    bt = nt.Batch(ctx, [(buf.ptr + off, q, log_n, flags) for off in offs])
    bt.launch()
    nt.device_sync()
    bt.destroy()
    out = [buf.download(np.uint32, n, off).astype(np.uint64) for off in offs]
    buf.free()
    return out


# (log n, polynomials, split, log2 tile elements, pipelining option): 16 x
# 2^20 with [11, 9] is the two-pass batch plan of the 2^20 batch workload, with
# exactly two tiles per workgroup slot, so the default pipelines it; 32 x 2^20
# gives four tiles per workgroup; the others need option 2.
BATCHES = [(20, 16, [11, 9], 14, 1), (20, 32, [11, 9], 14, 1), (18, 4, [9, 9], 14, 2), (16, 6, [8, 8], 13, 2)]


@pytest.mark.parametrize("log_n,count,split,log_e,pipe", BATCHES)
@pytest.mark.parametrize("kind", ["nfwd", "ninv", "fwd"])
def test_batch_of_equal_spacing_runs_one_pipelined_job(nt, orc, tmp_path, log_n, count, split, log_e, pipe, kind):
    inv, nega = kind in ("inv", "ninv"), kind.startswith("n")
    flags = (nt.FLAG_INVERSE if inv else 0) | (nt.FLAG_NEGACYCLIC if nega else 0)
    _install(nt, tmp_path, log_n, count, split, log_e, 3)
    xs = [orc.random_poly(1 << log_n, BB, 7000 + 31 * log_n + i) for i in range(count)]
    ctx = nt.Context()
    gaps = [256] * count  # equal spacing, 16-byte aligned
    piped = _batch_run(nt, ctx, xs, BB, log_n, flags, gaps, pipe)
    plain = _batch_run(nt, ctx, xs, BB, log_n, flags, gaps, 0)
    # unequal spacing: only runs of equally spaced neighbours merge, so the
    # group keeps a job table (several jobs, one tile per workgroup)
    uneven = _batch_run(nt, ctx, xs, BB, log_n, flags, [256 * (1 + i % 3) for i in range(count)], pipe)
    ctx.destroy()
    for i, x in enumerate(xs):
        assert np.array_equal(piped[i], plain[i]), i
        assert np.array_equal(uneven[i], plain[i]), i
        if i < 3 or i == count - 1:
            assert np.array_equal(piped[i], orc.transform(x, BB, nega, inv)), i
