"""ntt_c_config_transform_noalloc (workspace in the library's static device
array, no allocation) against the existing entry points it mirrors
(ntt_c_forward / ntt_c_inverse / ntt_c_forward_negacyclic /
ntt_c_inverse_negacyclic), bit-exact, through the C-ABI: two calls in a row on
one stream with no sync between them, the second on a smaller input, in all
four directions; a 62-bit prime; and the size limit."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

BB = 2013265921
P62 = 4611686018326724609
KINDS = [(0, "forward"), (1, "inverse"), (2, "forward_negacyclic"), (3, "inverse_negacyclic")]


def _noalloc(nt, buf, cfg, flags, stream=None):
    nt.check(nt.lib().ntt_c_config_transform_noalloc(ctypes.c_void_p(buf.ptr), cfg.ptr, flags,
                                                      stream.ptr if stream else None))


def _existing(cfg, name, buf):
    getattr(cfg, name)(buf)


@pytest.mark.parametrize("flags,name", KINDS)
@pytest.mark.parametrize("q,big,small", [(BB, 22, 16), (BB, 20, 12), (P62, 21, 14)])
def test_noalloc_matches_existing_entry_point(nt, orc, flags, name, q, big, small):
    s = nt.Stream()
    cfgs = {ln: nt.Config(1 << ln, q) for ln in (big, small)}
    xs = {ln: orc.random_poly(1 << ln, q, 900 + ln + flags) for ln in (big, small)}
    bufs = {ln: nt.DeviceBuffer.from_array(xs[ln]) for ln in (big, small)}
    refs = {ln: nt.DeviceBuffer.from_array(xs[ln]) for ln in (big, small)}
    try:
        # two calls in a row on one stream, the second with a smaller input
        _noalloc(nt, bufs[big], cfgs[big], flags, s)
        _noalloc(nt, bufs[small], cfgs[small], flags, s)
        s.sync()
        for ln in (big, small):
            _existing(cfgs[ln], name, refs[ln])
        nt.device_sync()
        for ln in (big, small):
            got = bufs[ln].download(np.uint64, 1 << ln)
            assert np.array_equal(got, refs[ln].download(np.uint64, 1 << ln)), ln
        small_x = xs[small]
        expect = orc.transform(small_x, q, flags & 2, flags & 1)
        assert np.array_equal(bufs[small].download(np.uint64, 1 << small), expect)
        # and again the other way round: the array keeps the last call's contents
        for ln in (small, big):
            bufs[ln].upload(xs[ln])
            _noalloc(nt, bufs[ln], cfgs[ln], flags)
        nt.device_sync()
        for ln in (big, small):
            assert np.array_equal(bufs[ln].download(np.uint64, 1 << ln), refs[ln].download(np.uint64, 1 << ln)), ln
    finally:
        for d in list(bufs.values()) + list(refs.values()):
            d.free()
        for c in cfgs.values():
            c.cleanup()
        s.destroy()


def test_noalloc_rejects_beyond_the_static_array(nt, orc):
    n = 1 << 24
    cfg = nt.Config(n, BB)
    x = orc.random_poly(n, BB, 5)
    buf = nt.DeviceBuffer.from_array(x)
    try:
        with pytest.raises(nt.NttError) as e:
            _noalloc(nt, buf, cfg, 0)
        assert e.value.status == 5 and "static" in str(e.value)
        with pytest.raises(nt.NttError) as e:
            _noalloc(nt, buf, cfg, 4)  # NTT_FLAG_U64 is not a flag of this call
        assert e.value.status == 1
        nt.device_sync()
        assert np.array_equal(buf.download(np.uint64, n), x)
    finally:
        buf.free()
        cfg.cleanup()
