"""CPU-side checks of the C ABI: the library loads, exports every symbol that
include/ntt_c.h declares, and its host-only functions (params, ntt_host) and
argument validation behave like the reference.  No kernel is launched."""
import ctypes
import subprocess

import numpy as np
import pytest

BB = 2013265921


def test_library_exports_every_declared_symbol(nt):
    syms = nt.header_symbols()
    assert len(syms) >= 40
    L = nt.lib()
    missing = [s for s in syms if not hasattr(L, s)]
    assert not missing, missing


def test_exports_are_c_linkage():
    out = subprocess.check_output(["nm", "-D", "--defined-only", __import__("ntt_amd").LIB_PATH], text=True)
    names = {line.split()[-1] for line in out.splitlines() if line.strip()}
    import ntt_amd
    for s in ntt_amd.header_symbols():
        assert s in names, s  # unmangled


def test_create_params_matches_oracle(nt, orc):
    for q, n in [(BB, 4096), (BB, 1 << 20), (1056440321, 1 << 18), (4611686018326724609, 1 << 22)]:
        a = nt.create_params(q, n)
        b = orc.params(q, n)
        for f in ("q", "q_twice", "barrett_k", "barrett_m", "omega", "omega_inv", "n_inv"):
            assert getattr(a, f) == getattr(b, f), (q, n, f)


@pytest.mark.parametrize("q", [BB, 4611686018326724609])
@pytest.mark.parametrize("log_n", [1, 6, 12])
def test_library_ntt_host_matches_oracle(nt, orc, q, log_n):
    n = 1 << log_n
    x = orc.random_poly(n, q, 3)
    p = nt.create_params(q, n)
    assert np.array_equal(nt.ntt_host(x, p), orc.ntt(x, q))
    assert np.array_equal(nt.ntt_host(x, p, inverse=True), orc.ntt(x, q, inverse=True))


def test_find_primitive_root(nt, orc):
    L = nt.lib()
    for q, n in [(BB, 1 << 12), (BB, 1 << 24), (2147352577, 1 << 16)]:
        assert L.ntt_c_find_primitive_root(q, n) == orc.omega(q, n)


def test_init_rejects_bad_arguments_before_touching_the_device(nt):
    L = nt.lib()
    cfg = ctypes.POINTER(nt.ConfigStruct)()
    assert L.ntt_c_init(3, BB, ctypes.byref(cfg)) == 1            # n not a power of two
    assert L.ntt_c_init(0, BB, ctypes.byref(cfg)) == 1
    assert L.ntt_c_init(1024, BB + 2, ctypes.byref(cfg)) == 2     # q not prime
    assert L.ntt_c_init(1 << 28, BB, ctypes.byref(cfg)) == 2      # 2^28 does not divide q-1
    assert L.ntt_c_init(1024, 1 << 62, ctypes.byref(cfg)) == 5    # q too large
    assert b"prime" in L.ntt_c_last_error() or b"q" in L.ntt_c_last_error()
    assert not cfg


def test_descriptor_validation_is_host_side(nt):
    ctx = nt.Context(workspace_bytes=0)
    with pytest.raises(nt.NttError) as e:
        ctx.transform(0, BB, 10, 0)                                  # null data
    assert e.value.status == 1
    with pytest.raises(nt.NttError) as e:
        ctx.transform(0x1000, 4611686018326724609, 10, 0)            # 62-bit q on u32 storage
    assert e.value.status == 1
    with pytest.raises(nt.NttError) as e:
        ctx.transform(0x1000, BB, 10, 64)                            # unknown flag
    assert e.value.status == 1
    with pytest.raises(nt.NttError) as e:
        ctx.transform(0x1000, BB, 27, 0)                             # n too large
    assert e.value.status == 5
    with pytest.raises(nt.NttError) as e:
        ctx.transform(0x1000, 1073479681, 21, 2)                     # 2n does not divide q-1
    assert e.value.status == 2
    ctx.destroy()


def test_inner_product_batch_checked_on_the_host(nt):
    """ntt_c_batch_prepare_ip refuses a bad inner-product source before any
    device work: a forward transform, a null input, more than 64 pairs,
    slices closer than n elements, data overlapping an input."""
    ctx = nt.Context(workspace_bytes=0)
    n, inv = 1 << 10, nt.FLAG_INVERSE
    a, b, out = 0x200000, 0x400000, 0x100000
    cases = [
        ((out, BB, 10, 0), (a, b, 2, n)),                  # forward transform
        ((out, BB, 10, inv), (0, b, 2, n)),                # null a
        ((out, BB, 10, inv), (a, 0, 1, 0)),                # null b
        ((out, BB, 10, inv), (a, b, 65, n)),               # more than 64 pairs
        ((out, BB, 10, inv), (a, b, 2, n - 1)),            # slices overlap
        ((out, BB, 10, inv), (a, b, 2, (1 << 40) + 1)),    # slices implausibly far apart
        ((a + 4 * n, BB, 10, inv), (a, b, 2, n)),          # data is a's second slice
        ((b - 4, BB, 10, inv), (a, b, 1, 0)),              # data runs into b
        ((a + 8 * n, BB, 10, inv | nt.FLAG_U64), (a, b, 2, n)),  # 8-byte slices: data inside a
    ]
    for d, ip in cases:
        with pytest.raises(nt.NttError) as e:
            nt.Batch(ctx, [d], ips=[ip])
        assert e.value.status == 1, (d, ip)
    with pytest.raises(ValueError):
        nt.Batch(ctx, [(out, BB, 10, inv)], ips=[])
    ctx.destroy()


def test_null_arguments_rejected(nt):
    L = nt.lib()
    assert L.ntt_c_forward(None, None, 3, None) == 1
    assert L.ntt_c_forward(ctypes.c_void_p(16), None, 9, None) == 1   # bad variant
    assert L.ntt_c_bit_reverse(None, 8, None) == 1
    assert L.ntt_c_bit_reverse(ctypes.c_void_p(16), 6, None) == 1
    assert L.ntt_c_rowmaj_to_pretransposed(ctypes.c_void_p(16), ctypes.c_void_p(16), 64, 64, None) == 1


def test_converter_and_rns_arguments_checked_on_the_host(nt):
    """Invalid calls of the width converters and the RNS plans return
    NTT_ERR_INVALID_ARG (1) or NTT_ERR_UNSUPPORTED before any device work."""
    L = nt.lib()
    p16 = ctypes.c_void_p(256)
    assert L.ntt_c_u64_to_u32(None, p16, 8, 2013265921, None) == 1
    assert L.ntt_c_u64_to_u32(p16, ctypes.c_void_p(4096), 8, 1 << 32, None) == 1     # q too large
    assert L.ntt_c_u64_to_u32(ctypes.c_void_p(260), ctypes.c_void_p(4096), 8, 2013265921, None) == 1  # misaligned
    assert L.ntt_c_u32_to_u64(ctypes.c_void_p(4096), ctypes.c_void_p(4096 + 16), 64, None) == 1    # overlap
    q = (ctypes.c_uint64 * 3)(2013265921, 2013265921, 15)
    out = ctypes.c_void_p()
    assert L.ntt_c_bconv_create(q, 2, q, 1, ctypes.byref(out)) == 1          # repeated source modulus
    assert L.ntt_c_bconv_create(q, 1, ctypes.cast(ctypes.byref(q, 16), ctypes.POINTER(ctypes.c_uint64)), 1,
                                ctypes.byref(out)) == 1                       # 15 is not prime
    assert L.ntt_c_bconv_create(q, 0, q, 1, ctypes.byref(out)) != 0          # no source moduli
    assert L.ntt_c_pmac_create(ctypes.cast(ctypes.byref(q, 16), ctypes.POINTER(ctypes.c_uint64)), 1,
                               ctypes.byref(out)) == 1
    assert L.ntt_c_bconv(None, p16, p16, 8, None) == 1
    assert L.ntt_c_pmac(None, p16, p16, p16, 1, 8, 0, None) == 1
    assert out.value is None


def test_every_declared_function_has_a_prototype(nt):
    """A function without argtypes would get 64-bit pointers as C ints."""
    import ntt_arb  # noqa: F401  (registers the K-limb prototypes)
    L = nt.lib()
    missing = [s for s in nt.header_symbols() if getattr(L, s).argtypes is None]
    assert not missing, missing


def test_prototypes_apply_when_a_module_is_imported_after_load(nt):
    import importlib

    import ntt_arb
    L = nt.lib()
    L.ntt_c_arb_forward.argtypes = None
    importlib.reload(ntt_arb)
    assert L.ntt_c_arb_forward.argtypes is not None
