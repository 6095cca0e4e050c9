"""The transform path allocates nothing: workspace is reserved at set-up
(ntt_c_init for the NULL stream, ntt_c_config_reserve / ntt_c_context_reserve
for other streams), a call on a stream without a reservation fails with
NTT_ERR_OOM and leaves the data alone, batches larger than the reservation run
in chunks, and the reference entry points can be captured into a hipGraph (in
global capture mode, where an allocation inside the capture would fail it) on a
fresh stream right after init and replayed bit-exactly."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

BB = 2013265921
P62 = 4611686018326724609
OOM = 3


@pytest.mark.parametrize("q,log_n", [(BB, 16), (BB, 20), (P62, 14)])
def test_unreserved_stream_is_refused(nt, orc, q, log_n):
    n = 1 << log_n
    cfg = nt.Config(n, q)
    s = nt.Stream()
    x = orc.random_poly(n, q, 3)
    buf = nt.DeviceBuffer.from_array(x)
    try:
        with pytest.raises(nt.NttError) as e:
            cfg.forward(buf, stream=s)
        assert e.value.status == OOM
        s.sync()
        assert np.array_equal(buf.download(np.uint64, n), x)
        cfg.reserve(s)
        cfg.forward(buf, stream=s)
        s.sync()
        assert np.array_equal(buf.download(np.uint64, n), orc.ntt(x, q))
    finally:
        buf.free()
        s.destroy()
        cfg.cleanup()


def test_context_transform_needs_a_reservation(nt, orc):
    ctx = nt.Context(workspace_bytes=0)
    n, log_n = 1 << 18, 18
    x = orc.random_poly(n, BB, 4)
    buf = nt.DeviceBuffer.from_array(x.astype(np.uint32))
    try:
        with pytest.raises(nt.NttError) as e:
            ctx.transform(buf.ptr, BB, log_n, nt.FLAG_NEGACYCLIC)
        assert e.value.status == OOM
        need = nt.workspace_bytes(BB, log_n, nt.FLAG_NEGACYCLIC)
        assert 0 < need <= 8 * n
        ctx.prepare([(buf.ptr, BB, log_n, nt.FLAG_NEGACYCLIC)])
        ctx.transform(buf.ptr, BB, log_n, nt.FLAG_NEGACYCLIC)
        nt.device_sync()
        assert np.array_equal(buf.download(np.uint32, n).astype(np.uint64), orc.ntt_negacyclic(x, BB))
        assert nt.workspace_bytes(BB, 10, 0) == 0  # one-pass size
    finally:
        buf.free()
        ctx.destroy()


@pytest.mark.parametrize("log_n,batch,reserve", [(16, 7, 2), (14, 5, None), (20, 3, 1)])
def test_batch_larger_than_reservation_runs_in_chunks(nt, orc, log_n, batch, reserve):
    n = 1 << log_n
    cfg = nt.Config(n, BB)
    s = nt.Stream() if reserve else None
    if reserve:
        cfg.reserve(s, reserve)
    xs = [orc.random_poly(n, BB, 50 + i) for i in range(batch)]
    buf = nt.DeviceBuffer.from_array(np.concatenate(xs))
    try:
        cfg.forward_batch(buf, batch, stream=s)
        nt.device_sync()
        got = buf.download(np.uint64, n * batch).reshape(batch, n)
        for i, x in enumerate(xs):
            assert np.array_equal(got[i], orc.ntt(x, BB)), i
        cfg.inverse_batch(buf, batch, stream=s)
        nt.device_sync()
        assert np.array_equal(buf.download(np.uint64, n * batch), np.concatenate(xs))
    finally:
        buf.free()
        if s:
            s.destroy()
        cfg.cleanup()


@pytest.mark.parametrize("q,log_n", [(BB, 12), (BB, 16), (BB, 20), (P62, 15)])
def test_graph_capture_right_after_init(nt, orc, q, log_n):
    """ntt_c_forward and ntt_c_negacyclic_multiply captured on a fresh stream
    (reserved once), replayed three times with eager uploads and syncs between
    the replays; every replay equals the eager result and the oracle."""
    n = 1 << log_n
    cfg = nt.Config(n, q)
    s = nt.Stream()
    cfg.reserve(s)
    x, a, b = (orc.random_poly(n, q, 60 + i) for i in range(3))
    dx, da, db, dr = nt.DeviceBuffer(n * 8), nt.DeviceBuffer(n * 8), nt.DeviceBuffer(n * 8), nt.DeviceBuffer(n * 8)
    expect_x = orc.ntt(x, q)
    A, B = orc.ntt_negacyclic(a, q), orc.ntt_negacyclic(b, q)
    expect_r = orc.ntt_negacyclic(orc.pointwise_mul(A, B, q), q, True)
    graph = None
    try:
        graph = nt.Graph(s, lambda: (cfg.forward(dx, stream=s), cfg.negacyclic_multiply(dr, da, db, stream=s)))
        for rep in range(3):
            dx.upload(x)
            da.upload(a)
            db.upload(b)
            graph.launch(s)
            s.sync()
            assert np.array_equal(dx.download(np.uint64, n), expect_x), rep
            assert np.array_equal(dr.download(np.uint64, n), expect_r), rep
            assert np.array_equal(da.download(np.uint64, n), A), rep
        # the same calls issued eagerly give the same bits
        dx.upload(x)
        cfg.forward(dx, stream=s)
        s.sync()
        assert np.array_equal(dx.download(np.uint64, n), expect_x)
    finally:
        if graph:
            graph.destroy()
        for d in (dx, da, db, dr):
            d.free()
        s.destroy()
        cfg.cleanup()


def test_graph_capture_context_transform(nt, orc):
    """A descriptor transform on a prepared context, captured and replayed."""
    n, log_n = 1 << 22, 22
    s = nt.Stream()
    ctx = nt.Context(workspace_bytes=0)
    x = orc.random_poly(n, BB, 71)
    buf = nt.DeviceBuffer(n * 4)
    desc = (buf.ptr, BB, log_n, nt.FLAG_NEGACYCLIC)
    ctx.prepare([desc], stream=s)
    expect = orc.ntt_negacyclic(x, BB)
    graph = None
    try:
        graph = nt.Graph(s, lambda: ctx.transform(*desc, stream=s))
        for rep in range(3):
            buf.upload(x.astype(np.uint32))
            graph.launch(s)
            s.sync()
            assert np.array_equal(buf.download(np.uint32, n).astype(np.uint64), expect), rep
    finally:
        if graph:
            graph.destroy()
        buf.free()
        ctx.destroy()
        s.destroy()
