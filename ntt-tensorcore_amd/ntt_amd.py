"""ctypes binding of libntt_mi355x.so (include/ntt_c.h) for Python hosts.

The library is the product; this module only marshals pointers and sizes.
Importing it never falls back to anything: if the shared library is missing
it raises, so no caller can run without the HIP kernels.
"""
import ctypes
import os
import re

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(HERE)
LIB_PATH = os.path.join(HERE, "libntt_mi355x.so")
HEADER = os.path.join(REPO, "include", "ntt_c.h")

NTT_OK = 0
STATUS = {0: "OK", 1: "INVALID_ARG", 2: "NO_ROOT", 3: "OOM", 4: "DEVICE", 5: "UNSUPPORTED", 6: "BUSY"}
VARIANT = {"basic": 0, "four_step": 1, "tlmop": 2, "optimized": 3}
FLAG_INVERSE, FLAG_NEGACYCLIC, FLAG_U64 = 1, 2, 4

BABYBEAR = 2013265921


class NttError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__(f"{STATUS.get(status, status)}: {msg}")
        self.status = status


class ModularParams(ctypes.Structure):
    _fields_ = [(n, ctypes.c_uint64) for n in
                ("q", "q_twice", "barrett_k", "barrett_m", "omega", "omega_inv", "n_inv")]


class ConfigStruct(ctypes.Structure):
    _fields_ = [("n", ctypes.c_uint64), ("log_n", ctypes.c_uint64), ("q", ctypes.c_uint64),
                ("mu", ctypes.c_uint64), ("d_twiddles", ctypes.c_void_p),
                ("d_twiddles_inv", ctypes.c_void_p), ("params", ModularParams),
                ("impl", ctypes.c_void_p)]


class Desc(ctypes.Structure):
    _fields_ = [("data", ctypes.c_void_p), ("q", ctypes.c_uint64), ("log_n", ctypes.c_uint32),
                ("flags", ctypes.c_uint32)]


class IpSrc(ctypes.Structure):
    """ntt_c_ip: the inner product an inverse transform of a batch reads."""
    _fields_ = [("a", ctypes.c_void_p), ("b", ctypes.c_void_p), ("dnum", ctypes.c_uint32),
                ("reserved", ctypes.c_uint32), ("kstride", ctypes.c_int64)]


_lib = None


def header_symbols():
    """Names of every function declared in include/ntt_c.h."""
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(ntt_c_\w+)\s*\(", text)))


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is not built; run `make -C {HERE}` (or __graft_entry__.build())")
    L = ctypes.CDLL(LIB_PATH)
    vp, u64, u32, i32 = ctypes.c_void_p, ctypes.c_uint64, ctypes.c_uint32, ctypes.c_int
    P = ctypes.POINTER
    sigs = {
        "ntt_c_init": [u64, u64, P(P(ConfigStruct))],
        "ntt_c_cleanup": [P(ConfigStruct)],
        "ntt_c_forward": [vp, P(ConfigStruct), i32, vp],
        "ntt_c_inverse": [vp, P(ConfigStruct), i32, vp],
        "ntt_c_forward_negacyclic": [vp, P(ConfigStruct), vp],
        "ntt_c_inverse_negacyclic": [vp, P(ConfigStruct), vp],
        "ntt_c_forward_batch": [vp, u64, P(ConfigStruct), i32, vp],
        "ntt_c_inverse_batch": [vp, u64, P(ConfigStruct), i32, vp],
        "ntt_c_poly_multiply": [vp, vp, vp, P(ConfigStruct), vp],
        "ntt_c_negacyclic_multiply": [vp, vp, vp, P(ConfigStruct), vp],
        "ntt_c_bit_reverse": [vp, u64, vp],
        "ntt_c_u64_to_u32": [vp, vp, u64, u64, vp],
        "ntt_c_u32_to_u64": [vp, vp, u64, vp],
        "ntt_c_forward_rowmaj": [vp, P(ConfigStruct), vp],
        "ntt_c_rowmaj_to_pretransposed": [vp, vp, i32, i32, vp],
        "ntt_c_rowmaj_from_pretransposed": [vp, vp, i32, i32, vp],
        "ntt_c_host": [P(u64), u64, P(ModularParams), i32],
        "ntt_c_create_params": [u64, u64, P(ModularParams)],
        "ntt_c_config_reserve": [P(ConfigStruct), vp, u64],
        "ntt_c_config_transform_noalloc": [vp, P(ConfigStruct), u32, vp],
        "ntt_c_context_reserve": [vp, vp, u64],
        "ntt_c_context_prepare": [vp, P(Desc), ctypes.c_size_t, vp],
        "ntt_c_workspace_bytes": [P(Desc), P(u64)],
        "ntt_c_context_create": [P(vp)],
        "ntt_c_context_destroy": [vp],
        "ntt_c_transform": [vp, P(Desc), vp],
        "ntt_c_transform_split": [vp, P(Desc), u32, P(u32), u32, u32, vp],
        "ntt_c_transform_each": [vp, P(Desc), ctypes.c_size_t, vp],
        "ntt_c_batch_prepare": [vp, P(Desc), ctypes.c_size_t, P(vp)],
        "ntt_c_batch_prepare_ip": [vp, P(Desc), P(IpSrc), ctypes.c_size_t, P(vp)],
        "ntt_c_batch_launch": [vp, vp],
        "ntt_c_batch_launch_graph": [vp, vp],
        "ntt_c_batch_set_queue": [vp, i32, u32],
        "ntt_c_set_option": [u32, ctypes.c_int64, P(ctypes.c_int64)],
        "ntt_c_batch_check": [vp, vp],
        "ntt_c_batch_destroy": [vp],
        "ntt_c_batch_info": [vp, P(u32), P(u64), P(u64)],
        "ntt_c_device_count": [P(i32)],
        "ntt_c_malloc": [P(vp), u64],
        "ntt_c_free": [vp],
        "ntt_c_memcpy_h2d": [vp, vp, u64],
        "ntt_c_memcpy_d2h": [vp, vp, u64],
        "ntt_c_memset": [vp, i32, u64],
        "ntt_c_stream_create": [P(vp)],
        "ntt_c_stream_destroy": [vp],
        "ntt_c_stream_sync": [vp],
        "ntt_c_device_sync": [],
        "ntt_c_event_create": [P(vp)],
        "ntt_c_event_destroy": [vp],
        "ntt_c_event_record": [vp, vp],
        "ntt_c_event_elapsed_ms": [P(ctypes.c_float), vp, vp],
        "ntt_c_graph_begin": [vp],
        "ntt_c_graph_end": [vp, P(vp)],
        "ntt_c_graph_launch": [vp, vp],
        "ntt_c_graph_destroy": [vp],
        "ntt_c_bconv_create": [P(u64), u32, P(u64), u32, P(vp)],
        "ntt_c_bconv": [vp, vp, vp, u64, vp],
        "ntt_c_bconv_destroy": [vp],
        "ntt_c_pmac_create": [P(u64), u32, P(vp)],
        "ntt_c_pmac": [vp, vp, vp, vp, u32, u64, i32, vp],
        "ntt_c_pmac_destroy": [vp],
    }
    for name, args in sigs.items():
        f = getattr(L, name)
        f.argtypes = args
        f.restype = ctypes.c_int
    L.ntt_c_last_error.restype = ctypes.c_char_p
    L.ntt_c_last_error.argtypes = []
    L.ntt_c_find_primitive_root.restype = u64
    L.ntt_c_find_primitive_root.argtypes = [u64, u64]
    L.ntt_c_context_table_bytes.restype = u64
    L.ntt_c_context_table_bytes.argtypes = [vp]
    for extra in _EXTRA_SIGS:
        extra(L)
    _lib = L
    return L


# Later modules of the binding register more prototypes through register_sigs.
_EXTRA_SIGS = []


def register_sigs(fn):
    """Adds a function that sets ctypes prototypes on the library.  It runs at
    once if the library is already loaded, so a module imported after the
    first lib() call still gets its prototypes (without them ctypes would pass
    64-bit pointers as C ints)."""
    _EXTRA_SIGS.append(fn)
    if _lib is not None:
        fn(_lib)


OPT_PIPELINE = 1


def set_option(option, value):
    """Process-wide kernel choice (ntt_c_set_option); returns the old value."""
    old = ctypes.c_int64()
    check(lib().ntt_c_set_option(option, value, ctypes.byref(old)))
    return old.value


def check(status):
    if status != NTT_OK:
        raise NttError(status, lib().ntt_c_last_error().decode(errors="replace"))


# ---------------------------------------------------------------- device memory
class DeviceBuffer:
    def __init__(self, nbytes):
        self.nbytes = int(nbytes)
        p = ctypes.c_void_p()
        check(lib().ntt_c_malloc(ctypes.byref(p), self.nbytes))
        self.ptr = p.value

    @classmethod
    def from_array(cls, a):
        a = np.ascontiguousarray(a)
        b = cls(a.nbytes)
        b.upload(a)
        return b

    def upload(self, a, offset=0):
        a = np.ascontiguousarray(a)
        assert offset + a.nbytes <= self.nbytes
        check(lib().ntt_c_memcpy_h2d(self.ptr + offset, a.ctypes.data, a.nbytes))

    def download(self, dtype, count, offset=0):
        out = np.empty(count, dtype=dtype)
        assert offset + out.nbytes <= self.nbytes
        check(lib().ntt_c_memcpy_d2h(out.ctypes.data, self.ptr + offset, out.nbytes))
        return out

    def free(self):
        if self.ptr:
            check(lib().ntt_c_free(self.ptr))
            self.ptr = None

    def __del__(self):
        try:
            if self.ptr and _lib is not None:
                _lib.ntt_c_free(self.ptr)
        except Exception:
            pass


class Stream:
    def __init__(self):
        p = ctypes.c_void_p()
        check(lib().ntt_c_stream_create(ctypes.byref(p)))
        self.ptr = p.value

    def sync(self):
        check(lib().ntt_c_stream_sync(self.ptr))

    def destroy(self):
        if self.ptr:
            check(lib().ntt_c_stream_destroy(self.ptr))
            self.ptr = None


class Event:
    def __init__(self):
        p = ctypes.c_void_p()
        check(lib().ntt_c_event_create(ctypes.byref(p)))
        self.ptr = p.value

    def record(self, stream=None):
        check(lib().ntt_c_event_record(self.ptr, stream.ptr if stream else None))

    def elapsed_ms(self, end):
        ms = ctypes.c_float()
        check(lib().ntt_c_event_elapsed_ms(ctypes.byref(ms), self.ptr, end.ptr))
        return ms.value

    def destroy(self):
        if self.ptr:
            check(lib().ntt_c_event_destroy(self.ptr))
            self.ptr = None


class Graph:
    """Calls captured on a stream (global capture mode) and replayed as a hipGraph."""

    def __init__(self, stream, record):
        check(lib().ntt_c_graph_begin(stream.ptr))
        try:
            record()
        finally:
            p = ctypes.c_void_p()
            st = lib().ntt_c_graph_end(stream.ptr, ctypes.byref(p))
        check(st)
        self.ptr = p.value

    def launch(self, stream):
        check(lib().ntt_c_graph_launch(self.ptr, stream.ptr))

    def destroy(self):
        if self.ptr:
            check(lib().ntt_c_graph_destroy(self.ptr))
            self.ptr = None


def device_count():
    c = ctypes.c_int()
    check(lib().ntt_c_device_count(ctypes.byref(c)))
    return c.value


def device_sync():
    check(lib().ntt_c_device_sync())


def _sp(stream):
    return stream.ptr if stream is not None else None


# ---------------------------------------------------------------- include/ntt.cuh surface
class Config:
    """NTTConfig from ntt_init (reference src/ntt_basic.cu:217-249)."""

    def __init__(self, n, q=BABYBEAR):
        p = ctypes.POINTER(ConfigStruct)()
        check(lib().ntt_c_init(n, q, ctypes.byref(p)))
        self.ptr = p
        self.n, self.q = n, q

    @property
    def params(self):
        return self.ptr.contents.params

    @property
    def mu(self):
        return self.ptr.contents.mu

    def reserve(self, stream, max_batch=1):
        """Workspace of `stream` for calls of up to max_batch polynomials (set-up)."""
        check(lib().ntt_c_config_reserve(self.ptr, _sp(stream), max_batch))

    def forward(self, buf, variant="optimized", stream=None):
        check(lib().ntt_c_forward(buf.ptr, self.ptr, VARIANT[variant], _sp(stream)))

    def inverse(self, buf, variant="optimized", stream=None):
        check(lib().ntt_c_inverse(buf.ptr, self.ptr, VARIANT[variant], _sp(stream)))

    def forward_negacyclic(self, buf, stream=None):
        check(lib().ntt_c_forward_negacyclic(buf.ptr, self.ptr, _sp(stream)))

    def inverse_negacyclic(self, buf, stream=None):
        check(lib().ntt_c_inverse_negacyclic(buf.ptr, self.ptr, _sp(stream)))

    def forward_batch(self, buf, batch, variant="optimized", stream=None):
        check(lib().ntt_c_forward_batch(buf.ptr, batch, self.ptr, VARIANT[variant], _sp(stream)))

    def inverse_batch(self, buf, batch, variant="optimized", stream=None):
        check(lib().ntt_c_inverse_batch(buf.ptr, batch, self.ptr, VARIANT[variant], _sp(stream)))

    def poly_multiply(self, out, a, b, stream=None):
        check(lib().ntt_c_poly_multiply(out.ptr, a.ptr, b.ptr, self.ptr, _sp(stream)))

    def negacyclic_multiply(self, out, a, b, stream=None):
        check(lib().ntt_c_negacyclic_multiply(out.ptr, a.ptr, b.ptr, self.ptr, _sp(stream)))

    def forward_rowmaj(self, buf, stream=None):
        check(lib().ntt_c_forward_rowmaj(buf.ptr, self.ptr, _sp(stream)))

    def cleanup(self):
        if self.ptr:
            check(lib().ntt_c_cleanup(self.ptr))
            self.ptr = None


def workspace_bytes(q, log_n, flags):
    """Workspace bytes one transform of this kind needs (ntt_c_workspace_bytes)."""
    d = Desc(1, q, log_n, flags)
    out = ctypes.c_uint64()
    check(lib().ntt_c_workspace_bytes(ctypes.byref(d), ctypes.byref(out)))
    return out.value


def create_params(q, n):
    p = ModularParams()
    check(lib().ntt_c_create_params(q, n, ctypes.byref(p)))
    return p


def ntt_host(data, params, inverse=False):
    """The library's host transform (reference ntt_host), on a copy."""
    a = np.ascontiguousarray(data, dtype=np.uint64).copy()
    check(lib().ntt_c_host(a.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)), a.size, ctypes.byref(params),
                           int(inverse)))
    return a


def bit_reverse(buf, n, stream=None):
    check(lib().ntt_c_bit_reverse(buf.ptr, n, _sp(stream)))


def u64_to_u32(dst, src, n, q, stream=None):
    """dst (uint32) = src (uint64) mod q, device to device."""
    check(lib().ntt_c_u64_to_u32(dst.ptr, src.ptr, n, q, _sp(stream)))


def u32_to_u64(dst, src, n, stream=None):
    """dst (uint64) = src (uint32) zero-extended, device to device."""
    check(lib().ntt_c_u32_to_u64(dst.ptr, src.ptr, n, _sp(stream)))


def rowmaj_to_pretransposed(dst, src, n1, n2, stream=None):
    check(lib().ntt_c_rowmaj_to_pretransposed(dst.ptr, src.ptr, n1, n2, _sp(stream)))


def rowmaj_from_pretransposed(dst, src, n1, n2, stream=None):
    check(lib().ntt_c_rowmaj_from_pretransposed(dst.ptr, src.ptr, n1, n2, _sp(stream)))


# ---------------------------------------------------------------- descriptor API
def _u64_array(xs):
    xs = [int(v) for v in xs]
    return (ctypes.c_uint64 * len(xs))(*xs), len(xs)


class Bconv:
    """Fast RNS base conversion plan (include/ntt_c.h, ntt_c_bconv_*): limb-major
    uint32 residues, L source moduli to K target moduli."""

    def __init__(self, q_from, p_to):
        qa, self.L = _u64_array(q_from)
        pa, self.K = _u64_array(p_to)
        p = ctypes.c_void_p()
        check(lib().ntt_c_bconv_create(qa, self.L, pa, self.K, ctypes.byref(p)))
        self.ptr = p.value

    def run(self, out, inp, n, stream=None):
        check(lib().ntt_c_bconv(self.ptr, out.ptr, inp.ptr, n, _sp(stream)))

    def destroy(self):
        if self.ptr:
            check(lib().ntt_c_bconv_destroy(self.ptr))
            self.ptr = None


class Pmac:
    """Per-limb multiply-accumulate plan (ntt_c_pmac_*): out[l] (+)= sum_k a[k][l] * b[k][l] mod q_l."""

    def __init__(self, q):
        qa, self.L = _u64_array(q)
        p = ctypes.c_void_p()
        check(lib().ntt_c_pmac_create(qa, self.L, ctypes.byref(p)))
        self.ptr = p.value

    def run(self, out, a, b, dnum, n, accumulate=False, stream=None):
        check(lib().ntt_c_pmac(self.ptr, out.ptr, a.ptr, b.ptr, dnum, n, int(bool(accumulate)), _sp(stream)))

    def destroy(self):
        if self.ptr:
            check(lib().ntt_c_pmac_destroy(self.ptr))
            self.ptr = None


DEFAULT_WORKSPACE = 256 << 20  # any single transform up to 2^25 x 8 bytes


class Context:
    def __init__(self, workspace_bytes=DEFAULT_WORKSPACE, stream=None):
        """A plan/twiddle cache; reserves `workspace_bytes` of four-step
        workspace for `stream` (the NULL stream by default) at once, since
        transforms never allocate (0: reserve nothing)."""
        p = ctypes.c_void_p()
        check(lib().ntt_c_context_create(ctypes.byref(p)))
        self.ptr = p.value
        if workspace_bytes:
            self.reserve(stream, workspace_bytes)

    def reserve(self, stream, nbytes):
        check(lib().ntt_c_context_reserve(self.ptr, _sp(stream), nbytes))

    def prepare(self, descs, stream=None):
        """Builds the plans of (ptr, q, log_n, flags) descriptors and reserves
        the workspace one of them needs on `stream`."""
        arr = (Desc * len(descs))(*[Desc(*d) for d in descs])
        check(lib().ntt_c_context_prepare(self.ptr, arr, len(descs), _sp(stream)))

    def table_bytes(self):
        return int(lib().ntt_c_context_table_bytes(self.ptr))

    def transform(self, data_ptr, q, log_n, flags, stream=None):
        d = Desc(data_ptr, q, log_n, flags)
        check(lib().ntt_c_transform(self.ptr, ctypes.byref(d), _sp(stream)))

    def transform_each(self, descs, stream=None):
        """One ntt_c_transform per (ptr, q, log_n, flags), in order, from native code."""
        arr = descs if isinstance(descs, ctypes.Array) else (Desc * len(descs))(*[Desc(*d) for d in descs])
        check(lib().ntt_c_transform_each(self.ptr, arr, len(arr), _sp(stream)))

    @staticmethod
    def desc_array(descs):
        return (Desc * len(descs))(*[Desc(*d) for d in descs])

    def transform_split(self, data_ptr, q, log_n, flags, log_r, log_e, stream=None, mfma=0):
        d = Desc(data_ptr, q, log_n, flags)
        arr = (ctypes.c_uint32 * len(log_r))(*log_r)
        check(lib().ntt_c_transform_split(self.ptr, ctypes.byref(d), len(log_r), arr, log_e, mfma, _sp(stream)))

    def destroy(self):
        if self.ptr:
            check(lib().ntt_c_context_destroy(self.ptr))
            self.ptr = None


class Batch:
    def __init__(self, ctx, descs, ips=None):
        """descs: list of (data_ptr, q, log_n, flags).  ips (optional, one per
        descriptor): None, or (a_ptr, b_ptr, dnum, kstride) -- the inverse
        transform of that descriptor reads the inner product
        sum_k a[c + k*kstride] * b[c + k*kstride] mod q instead of its data
        (ntt_c_batch_prepare_ip)."""
        arr = (Desc * len(descs))(*[Desc(*d) for d in descs])
        p = ctypes.c_void_p()
        if ips is None:
            check(lib().ntt_c_batch_prepare(ctx.ptr, arr, len(descs), ctypes.byref(p)))
        else:
            if len(ips) != len(descs):
                raise ValueError("ips: one entry per descriptor")
            src = (IpSrc * len(descs))(*[IpSrc(None, None, 0, 0, 0) if ip is None else
                                         IpSrc(ip[0], ip[1], ip[2], 0, ip[3]) for ip in ips])
            check(lib().ntt_c_batch_prepare_ip(ctx.ptr, arr, src, len(descs), ctypes.byref(p)))
        self.ptr = p.value

    def launch(self, stream=None):
        check(lib().ntt_c_batch_launch(self.ptr, _sp(stream)))

    def launch_graph(self, stream=None):
        check(lib().ntt_c_batch_launch_graph(self.ptr, _sp(stream)))

    def set_queue(self, enable, spin_limit=0, agent_fences=False):
        """enable=True: the whole batch in one work-queue launch (the default
        where the mix allows it), with agent_fences=True forcing the
        agent-scope hand-over; False: one launch per phase group."""
        check(lib().ntt_c_batch_set_queue(self.ptr, (2 if agent_fences else 1) if enable else 0, spin_limit))

    def check(self, stream=None):
        """Synchronises the stream; raises if a work-queue wait gave up."""
        check(lib().ntt_c_batch_check(self.ptr, _sp(stream)))

    def info(self):
        l, w, g = ctypes.c_uint32(), ctypes.c_uint64(), ctypes.c_uint64()
        check(lib().ntt_c_batch_info(self.ptr, ctypes.byref(l), ctypes.byref(w), ctypes.byref(g)))
        return {"launches": l.value, "workspace_bytes": w.value, "workgroups": g.value}

    def destroy(self):
        if self.ptr:
            check(lib().ntt_c_batch_destroy(self.ptr))
            self.ptr = None


# ---------------------------------------------------------------- serving
class ServerOpts(ctypes.Structure):
    _fields_ = [("hbm_budget_bytes", ctypes.c_uint64), ("max_inflight", ctypes.c_uint32),
                ("max_batch_polys", ctypes.c_uint32), ("device", ctypes.c_int)]


class ServerStats(ctypes.Structure):
    _fields_ = [("completed", ctypes.c_uint64), ("failed", ctypes.c_uint64), ("batches", ctypes.c_uint64),
                ("butterflies", ctypes.c_uint64), ("p50_us", ctypes.c_double), ("p99_us", ctypes.c_double),
                ("mean_us", ctypes.c_double), ("wall_s", ctypes.c_double),
                ("max_concurrent_batches", ctypes.c_uint32), ("arena_free_bytes", ctypes.c_uint64)]


def _server_sigs(L):
    vp, u32, u64, P = ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint64, ctypes.POINTER
    for name, args in {
        "ntt_c_server_create": [vp, P(ServerOpts), P(vp)],
        "ntt_c_server_submit": [vp, P(Desc), ctypes.c_size_t, P(u64)],
        "ntt_c_server_wait": [vp, u64, ctypes.c_double],
        "ntt_c_server_poll": [vp, u64, P(ctypes.c_int), P(ctypes.c_double)],
        "ntt_c_server_get_stats": [vp, P(ServerStats)],
        "ntt_c_server_reset_stats": [vp],
        "ntt_c_server_replay": [vp, P(Desc), P(u32), P(ctypes.c_double), ctypes.c_size_t, P(ctypes.c_double)],
        "ntt_c_server_destroy": [vp],
    }.items():
        f = getattr(L, name)
        f.argtypes = args
        f.restype = ctypes.c_int


register_sigs(_server_sigs)


class Server:
    """Continuous-batching server over device-resident polynomials."""

    def __init__(self, ctx, hbm_budget_bytes=1 << 30, max_inflight=4, max_batch_polys=1024):
        o = ServerOpts(hbm_budget_bytes, max_inflight, max_batch_polys, 0)
        p = ctypes.c_void_p()
        check(lib().ntt_c_server_create(ctx.ptr, ctypes.byref(o), ctypes.byref(p)))
        self.ptr = p.value

    def submit(self, polys):
        """polys: list of (data_ptr, q, log_n, flags); returns the request id."""
        arr = (Desc * len(polys))(*[Desc(*d) for d in polys])
        rid = ctypes.c_uint64()
        check(lib().ntt_c_server_submit(self.ptr, arr, len(polys), ctypes.byref(rid)))
        return rid.value

    def wait(self, rid, timeout_s=-1.0):
        check(lib().ntt_c_server_wait(self.ptr, rid, timeout_s))

    def poll(self, rid):
        done, lat = ctypes.c_int(), ctypes.c_double()
        check(lib().ntt_c_server_poll(self.ptr, rid, ctypes.byref(done), ctypes.byref(lat)))
        return bool(done.value), lat.value

    def stats(self):
        s = ServerStats()
        check(lib().ntt_c_server_get_stats(self.ptr, ctypes.byref(s)))
        return {f: getattr(s, f) for f, _ in ServerStats._fields_}

    def reset_stats(self):
        check(lib().ntt_c_server_reset_stats(self.ptr))

    def replay(self, requests, arrivals_s):
        """requests: list of lists of (ptr, q, log_n, flags); returns latencies (us)."""
        flat = [d for r in requests for d in r]
        arr = (Desc * len(flat))(*[Desc(*d) for d in flat])
        counts = (ctypes.c_uint32 * len(requests))(*[len(r) for r in requests])
        arrivals = (ctypes.c_double * len(requests))(*arrivals_s)
        lat = (ctypes.c_double * len(requests))()
        check(lib().ntt_c_server_replay(self.ptr, arr, counts, arrivals, len(requests), lat))
        return list(lat)

    def destroy(self):
        if self.ptr:
            check(lib().ntt_c_server_destroy(self.ptr))
            self.ptr = None


# ---------------------------------------------------------------- autotuner
class TuneResult(ctypes.Structure):
    _fields_ = [("log_n", ctypes.c_uint32), ("npass", ctypes.c_uint32), ("log_r", ctypes.c_uint32 * 4),
                ("log_e", ctypes.c_uint32), ("mfma_mask", ctypes.c_uint32), ("candidates", ctypes.c_uint32),
                ("us_per_transform", ctypes.c_double), ("heuristic_us", ctypes.c_double)]


def _tune_sigs(L):
    P = ctypes.POINTER
    L.ntt_c_autotune.argtypes = [ctypes.c_void_p, ctypes.c_uint64, P(ctypes.c_uint32), ctypes.c_size_t,
                                 ctypes.c_uint64, ctypes.c_uint32, P(TuneResult)]
    L.ntt_c_autotune.restype = ctypes.c_int
    for n in ("ntt_c_autotune_save", "ntt_c_autotune_load"):
        getattr(L, n).argtypes = [ctypes.c_char_p]
        getattr(L, n).restype = ctypes.c_int
    L.ntt_c_autotune_clear.argtypes = []
    L.ntt_c_autotune_clear.restype = ctypes.c_int
    L.ntt_c_split_for.argtypes = [ctypes.c_uint64, ctypes.c_uint32, ctypes.c_uint64, P(ctypes.c_uint32)]
    L.ntt_c_split_for.restype = ctypes.c_int


register_sigs(_tune_sigs)


def autotune(ctx, q, log_ns, elems_in_flight, reps=5):
    arr = (ctypes.c_uint32 * len(log_ns))(*log_ns)
    out = (TuneResult * len(log_ns))()
    check(lib().ntt_c_autotune(ctx.ptr, q, arr, len(log_ns), elems_in_flight, reps, out))
    res = []
    for r in out:
        res.append({"log_n": r.log_n, "split": list(r.log_r)[:r.npass], "log_e": r.log_e, "mfma_mask": r.mfma_mask,
                    "candidates": r.candidates, "us": r.us_per_transform, "heuristic_us": r.heuristic_us})
    return res


def autotune_save(path):
    check(lib().ntt_c_autotune_save(path.encode()))


def autotune_load(path):
    check(lib().ntt_c_autotune_load(path.encode()))


def autotune_clear():
    check(lib().ntt_c_autotune_clear())


def split_for(q, log_n, elems_in_flight):
    o = (ctypes.c_uint32 * 7)()
    check(lib().ntt_c_split_for(q, log_n, elems_in_flight, o))
    return {"split": list(o)[1:1 + o[0]], "log_e": o[5], "mfma_mask": o[6]}
