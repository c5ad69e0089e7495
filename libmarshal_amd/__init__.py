"""libmarshal_amd - in-place layout conversion on AMD Instinct MI355X (gfx950).

Python binding of the C ABI in ``include/hip_marshal.h`` and ``include/marshal.h``. The
compute path is the HIP library ``libmarshal_hip.so`` built next to this file; there is no
fallback: if the library is missing, importing the binding raises.

The functions below mirror the reference's operator API (include/cl_marshal.h:26-49 and the
PyOpenCL layer src/py/transpose.py): same names without the ``cl_`` prefix, same argument
meaning, ``False`` on success and ``True`` on error. Device buffers are torch tensors on a
ROCm device (any dtype; elements are moved as opaque bytes); the stream is torch's current
stream unless one is given.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libmarshal_hip.so")

_lib = None

OP_TRANSPOSE, OP_AOS_ASTA, OP_SOA_ASTA, OP_ASTA_AOS, OP_ASTA_SOA = 0, 1, 2, 3, 4


class Desc(ctypes.Structure):
    """hip_marshal_desc of include/hip_marshal.h."""
    _fields_ = [("ptr", ctypes.c_void_p), ("op", ctypes.c_int32), ("elem_bytes", ctypes.c_int32),
                ("rows", ctypes.c_int64), ("cols", ctypes.c_int64), ("tile", ctypes.c_int64)]


class ServerStats(ctypes.Structure):
    """hip_marshal_server_stats of include/hip_marshal.h."""
    _fields_ = [("completed", ctypes.c_int64), ("failed", ctypes.c_int64),
                ("batches", ctypes.c_int64), ("max_batch", ctypes.c_int64),
                ("p50_ms", ctypes.c_double), ("p99_ms", ctypes.c_double),
                ("mean_ms", ctypes.c_double), ("max_ms", ctypes.c_double),
                ("bytes", ctypes.c_double), ("span_s", ctypes.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class KernelTime(ctypes.Structure):
    """hip_marshal_kernel_time of include/hip_marshal.h."""
    _fields_ = [("name", ctypes.c_char * 48), ("launches", ctypes.c_int64),
                ("ms", ctypes.c_double), ("bytes", ctypes.c_double)]


def lib():
    """Loads the HIP library (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(
            f"{LIB_PATH} is missing: build it with `python -m libmarshal_amd.build` "
            "(or __graft_entry__.build())")
    # One HIP runtime per process: torch ships its own libamdhip64 under the same soname. When
    # torch is loaded first the library binds to it; loaded before torch, the library would
    # pull in the system runtime and torch a second one, and the two do not share the device.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    L = ctypes.CDLL(LIB_PATH)
    vp, i64, u64p, sz, i32, b = (ctypes.c_void_p, ctypes.c_int64,
                                 ctypes.POINTER(ctypes.c_uint64), ctypes.c_size_t,
                                 ctypes.c_int, ctypes.c_bool)
    sig = {
        "hip_aos_asta_bs": (b, [vp, vp, i32, i32, i32]),
        "hip_aos_asta_pttwac": (b, [vp, vp, i32, i32, i32, i32, i32]),
        "hip_aos_asta": (b, [vp, vp, i32, i32, i32, i32]),
        "hip_transpose": (b, [vp, vp, i32, i32, i32, i32, i32, i32, u64p]),
        "hip_transpose_010_bs": (b, [vp, vp, i32, i32, i32, u64p]),
        "hip_transpose_010_pttwac": (b, [vp, vp, i32, i32, i32, u64p, i32, i32]),
        "hip_transpose_100": (b, [vp, vp, i32, i32, i32, u64p]),
        "hip_transpose_0100": (b, [vp, vp, i32, i32, i32, i32, u64p]),
        "hip_padding": (b, [vp, vp, i32, i32, i32, u64p]),
        "hip_unpadding": (b, [vp, vp, i32, i32, i32, u64p]),
        "hip_marshal_finalize": (None, []),
        "hip_transpose_010_bs_es": (b, [vp, vp, sz, i64, i64, i64, u64p]),
        "hip_transpose_010_pttwac_es": (b, [vp, vp, sz, i64, i64, i64, u64p, i32, i32]),
        "hip_transpose_100_es": (b, [vp, vp, sz, i64, i64, i64, u64p]),
        "hip_transpose_0100_es": (b, [vp, vp, sz, i64, i64, i64, i64, u64p]),
        "hip_transpose_0100_noalloc_es": (b, [vp, vp, sz, i64, i64, i64, i64]),
        "hip_transpose_es": (b, [vp, vp, sz, i64, i64, i64, i64, i32, i32, u64p]),
        "hip_transpose_full_es": (b, [vp, vp, sz, i64, i64, u64p]),
        "hip_aos_asta_es": (b, [vp, vp, sz, i64, i64, i64, i32, u64p]),
        "hip_asta_aos_es": (b, [vp, vp, sz, i64, i64, i64, i32, u64p]),
        "hip_soa_asta_es": (b, [vp, vp, sz, i64, i64, i64, u64p]),
        "hip_asta_soa_es": (b, [vp, vp, sz, i64, i64, i64, u64p]),
        "hip_padding_es": (b, [vp, vp, sz, i64, i64, i64, u64p]),
        "hip_unpadding_es": (b, [vp, vp, sz, i64, i64, i64, u64p]),
        "hip_marshal_set_lds_budget": (None, [sz]),
        "hip_marshal_get_lds_budget": (sz, []),
        "hip_marshal_set_profiling": (None, [i32]),
        "hip_marshal_cluster_aborts": (ctypes.c_uint64, []),
        "hip_marshal_last_route": (ctypes.c_char_p, []),
        "hip_marshal_reserve": (b, [sz]),
        "hip_marshal_flag_bytes_0100": (sz, [i64, i64, i64, i64, sz]),
        "hip_marshal_cycle_count": (ctypes.c_uint64, [i64, i64, ctypes.c_uint32,
                                                      ctypes.POINTER(ctypes.c_uint32)]),
        "hip_marshal_version": (ctypes.c_char_p, []),
        "hip_marshal_set_tuning": (b, [ctypes.c_char_p, i64]),
        "hip_marshal_autotune": (b, [vp, vp, sz, i64, i64, ctypes.POINTER(ctypes.c_int)]),
        "hip_marshal_autotune_desc": (b, [vp, ctypes.POINTER(Desc), ctypes.POINTER(ctypes.c_int)]),
        "hip_marshal_plan_transpose": (b, [i64, i64, sz, ctypes.POINTER(ctypes.c_int64),
                                           ctypes.POINTER(ctypes.c_int64),
                                           ctypes.POINTER(ctypes.c_int)]),
        "hip_marshal_batch_create": (vp, [ctypes.POINTER(Desc), i64, i32]),
        "hip_marshal_batch_run": (b, [vp, vp, u64p]),
        "hip_marshal_batch_info": (i64, [vp, ctypes.c_char_p]),
        "hip_marshal_batch_route": (ctypes.c_char_p, [vp, i64]),
        "hip_marshal_batch_destroy": (None, [vp]),
        "hip_marshal_server_create": (vp, [sz, i32, i32]),
        "hip_marshal_server_submit": (i64, [vp, ctypes.POINTER(Desc)]),
        "hip_marshal_server_submit_at": (i64, [vp, ctypes.POINTER(Desc), ctypes.c_double]),
        "hip_marshal_server_now": (ctypes.c_double, []),
        "hip_marshal_server_status": (i32, [vp, i64]),
        "hip_marshal_server_wait": (b, [vp, i64, ctypes.c_double]),
        "hip_marshal_server_drain": (b, [vp, ctypes.c_double]),
        "hip_marshal_server_request_times": (b, [vp, i64, ctypes.POINTER(ctypes.c_double),
                                                 ctypes.POINTER(ctypes.c_double),
                                                 ctypes.POINTER(ctypes.c_double)]),
        "hip_marshal_server_get_stats": (None, [vp, ctypes.POINTER(ServerStats)]),
        "hip_marshal_server_reset_stats": (None, [vp]),
        "hip_marshal_server_destroy": (None, [vp]),
        "hip_marshal_get_tuning": (i64, [ctypes.c_char_p]),
        "hip_marshal_set_kernel_timing": (None, [i32]),
        "hip_marshal_kernel_times": (i32, [ctypes.POINTER(KernelTime), i32]),
        "gpu_aos_asta_bs_float": (b, [vp, i32, i32, i32, vp]),
        "gpu_aos_asta_bs_double": (b, [vp, i32, i32, i32, vp]),
        "gpu_aos_asta_pttwac": (b, [vp, i32, i32, i32, vp]),
        "gpu_soa_asta_pttwac": (b, [vp, i32, i32, i32, vp]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


# Symbols declared by include/hip_marshal.h and include/marshal.h.
EXPORTED = [
    "hip_aos_asta_bs", "hip_aos_asta_pttwac", "hip_aos_asta", "hip_transpose",
    "hip_transpose_010_bs", "hip_transpose_010_pttwac", "hip_transpose_100",
    "hip_transpose_0100", "hip_padding", "hip_unpadding", "hip_marshal_finalize",
    "hip_transpose_010_bs_es", "hip_transpose_010_pttwac_es", "hip_transpose_100_es",
    "hip_transpose_0100_es", "hip_transpose_0100_noalloc_es", "hip_transpose_es", "hip_transpose_full_es", "hip_aos_asta_es",
    "hip_asta_aos_es", "hip_soa_asta_es", "hip_asta_soa_es", "hip_padding_es",
    "hip_unpadding_es", "hip_marshal_set_lds_budget", "hip_marshal_get_lds_budget",
    "hip_marshal_set_profiling", "hip_marshal_cluster_aborts", "hip_marshal_last_route",
    "hip_marshal_reserve",
    "hip_marshal_flag_bytes_0100", "hip_marshal_cycle_count", "hip_marshal_version",
    "hip_marshal_set_tuning",
    "hip_marshal_get_tuning", "hip_marshal_plan_transpose", "hip_marshal_autotune",
    "hip_marshal_autotune_desc", "hip_marshal_batch_create", "hip_marshal_batch_run",
    "hip_marshal_batch_info", "hip_marshal_batch_route", "hip_marshal_batch_destroy",
    "hip_marshal_server_create", "hip_marshal_server_submit", "hip_marshal_server_submit_at",
    "hip_marshal_server_now", "hip_marshal_server_status", "hip_marshal_server_wait",
    "hip_marshal_server_drain", "hip_marshal_server_request_times",
    "hip_marshal_server_get_stats", "hip_marshal_server_reset_stats",
    "hip_marshal_server_destroy", "hip_marshal_set_kernel_timing", "hip_marshal_kernel_times",
    "gpu_aos_asta_bs_float", "gpu_aos_asta_bs_double", "gpu_aos_asta_pttwac",
    "gpu_soa_asta_pttwac",
]


def _ptr(t, need=None):
    """Device address of a contiguous ROCm tensor; with `need`, the tensor must hold at least
    that many bytes (the C ABI trusts the dimensions it is given)."""
    import torch
    if not isinstance(t, torch.Tensor) or t.device.type != "cuda":
        raise TypeError("expected a torch tensor on a ROCm device")
    if not t.is_contiguous():
        raise ValueError("tensor must be contiguous")
    if need is not None and t.numel() * t.element_size() < need:
        raise ValueError(f"tensor holds {t.numel() * t.element_size()} bytes, the call needs {need}")
    return ctypes.c_void_p(t.data_ptr())


def _need(es, *dims):
    n = es
    for d in dims:
        n *= max(int(d), 0)
    return n


def _stream(stream):
    import torch
    s = torch.cuda.current_stream() if stream is None else stream
    return ctypes.c_void_p(s.cuda_stream)


class Elapsed:
    """Accumulating nanosecond counter (the reference's ``cl_ulong *elapsed_time``)."""

    def __init__(self):
        self._v = ctypes.c_uint64(0)

    @property
    def ns(self):
        return self._v.value

    def reset(self):
        self._v.value = 0

    def ref(self):
        return ctypes.byref(self._v)


def _et(elapsed):
    return None if elapsed is None else elapsed.ref()


def _es(t, elem_bytes):
    return t.element_size() if elem_bytes is None else elem_bytes


# ---- mirrors of include/cl_marshal.h (element size taken from the tensor) --------------------

def transpose(buf, A, a, B, b, R=1, stages=3, elapsed=None, elem_bytes=None, stream=None):
    """Full in-place transpose of an (A*a) x (B*b) matrix (cl_transpose)."""
    es = _es(buf, elem_bytes)
    return lib().hip_transpose_es(_stream(stream), _ptr(buf, _need(es, A, a, B, b)), es, A, a, B, b,
                                  R, stages, _et(elapsed))


def transpose_full(buf, M, N, elapsed=None, elem_bytes=None, stream=None):
    """M x N transpose with planner-chosen tile factors and route."""
    es = _es(buf, elem_bytes)
    return lib().hip_transpose_full_es(_stream(stream), _ptr(buf, _need(es, M, N)), es, M, N,
                                       _et(elapsed))


def transpose_010_bs(buf, A, a, B, elapsed=None, elem_bytes=None, stream=None):
    es = _es(buf, elem_bytes)
    return lib().hip_transpose_010_bs_es(_stream(stream), _ptr(buf, _need(es, A, a, B)), es, A,
                                         a, B, _et(elapsed))


def transpose_010_pttwac(buf, A, a, B, elapsed=None, R=1, P=1, elem_bytes=None, stream=None):
    es = _es(buf, elem_bytes)
    return lib().hip_transpose_010_pttwac_es(_stream(stream), _ptr(buf, _need(es, A, a, B)), es,
                                             A, a, B, _et(elapsed), R, P)


def transpose_100(buf, A, B, b, elapsed=None, elem_bytes=None, stream=None):
    es = _es(buf, elem_bytes)
    return lib().hip_transpose_100_es(_stream(stream), _ptr(buf, _need(es, A, B, b)), es, A, B,
                                      b, _et(elapsed))


def transpose_0100(buf, A, a, B, b, elapsed=None, elem_bytes=None, stream=None):
    es = _es(buf, elem_bytes)
    return lib().hip_transpose_0100_es(_stream(stream), _ptr(buf, _need(es, A, a, B, b)), es, A, a,
                                       B, b, _et(elapsed))


def transpose_0100_noalloc(buf, A, a, B, b, elem_bytes=None, stream=None):
    """0100 step that allocates no device memory (hip_transpose_0100_noalloc_es): in-place forms
    only, visited bits in the library's fixed 32 MiB device array; not timed."""
    es = _es(buf, elem_bytes)
    return lib().hip_transpose_0100_noalloc_es(_stream(stream), _ptr(buf, _need(es, A, a, B, b)),
                                               es, A, a, B, b)


def aos_asta(buf, height, width, tile_size, method=0, elapsed=None, elem_bytes=None,
             stream=None):
    """AoS [h][w] -> ASTA [h/t][w][t]; method 0 = BS then PTTWAC, 1 = BS, 2 = PTTWAC."""
    es = _es(buf, elem_bytes)
    return lib().hip_aos_asta_es(_stream(stream), _ptr(buf, _need(es, height, width)), es, height,
                                 width, tile_size, method, _et(elapsed))


def asta_aos(buf, height, width, tile_size, method=0, elapsed=None, elem_bytes=None,
             stream=None):
    es = _es(buf, elem_bytes)
    return lib().hip_asta_aos_es(_stream(stream), _ptr(buf, _need(es, height, width)), es, height,
                                 width, tile_size, method, _et(elapsed))


def soa_asta(buf, height, width, tile_size, elapsed=None, elem_bytes=None, stream=None):
    """SoA [w][h] -> ASTA [h/t][w][t] (100(w, h/t, t))."""
    es = _es(buf, elem_bytes)
    return lib().hip_soa_asta_es(_stream(stream), _ptr(buf, _need(es, height, width)), es, height,
                                 width, tile_size, _et(elapsed))


def asta_soa(buf, height, width, tile_size, elapsed=None, elem_bytes=None, stream=None):
    es = _es(buf, elem_bytes)
    return lib().hip_asta_soa_es(_stream(stream), _ptr(buf, _need(es, height, width)), es, height,
                                 width, tile_size, _et(elapsed))


def padding(buf, x_size, y_size, pad_size, elapsed=None, elem_bytes=None, stream=None):
    es = _es(buf, elem_bytes)
    return lib().hip_padding_es(_stream(stream), _ptr(buf, _need(es, y_size, pad_size)), es, x_size,
                                y_size, pad_size, _et(elapsed))


def unpadding(buf, x_size, y_size, pad_size, elapsed=None, elem_bytes=None, stream=None):
    es = _es(buf, elem_bytes)
    return lib().hip_unpadding_es(_stream(stream), _ptr(buf, _need(es, y_size, pad_size)), es,
                                  x_size, y_size, pad_size, _et(elapsed))


def set_lds_budget(nbytes):
    lib().hip_marshal_set_lds_budget(nbytes)


def lds_budget():
    return lib().hip_marshal_get_lds_budget()


def cluster_aborts():
    """Waits for the device; cluster-staged steps that gave up their barrier wait since the
    last check (their instances left unconverted; 0 normally)."""
    return int(lib().hip_marshal_cluster_aborts())


def set_profiling(on):
    lib().hip_marshal_set_profiling(1 if on else 0)


def set_tuning(key, value):
    if lib().hip_marshal_set_tuning(key.encode(), int(value)):
        raise ValueError(f"bad tuning {key}={value}")


def get_tuning(key):
    return lib().hip_marshal_get_tuning(key.encode())


def autotune(buf, M, N, elem_bytes=None, stream=None):
    """Times the planner's candidate routes for an M x N transpose on `buf` (left unchanged)
    and makes the fastest the plan for this shape. Returns its stage code."""
    st = ctypes.c_int(0)
    es = _es(buf, elem_bytes)
    if lib().hip_marshal_autotune(_stream(stream), _ptr(buf, _need(es, M, N)), es, M, N,
                                  ctypes.byref(st)):
        raise RuntimeError("hip_marshal_autotune failed")
    return st.value


# Form codes hip_marshal_autotune_desc reports for AoS/SoA <-> ASTA classes.
FORM_010 = {1: "010_BS-", 2: "010_OOP-", 3: "010_SPLIT-", 4: "010_RC-", 5: "010_PTTWAC-"}
FORM_0100 = {11: "0100_CYC-", 12: "0100_SEG-", 13: "0100_OOP-", 14: "0100_RC-", 15: "0100_LDS-",
             16: "0100_PTTWAC-", 17: "0100_CL-"}


def autotune_desc(buf, op, elem_bytes, rows, cols, tile=0, stream=None):
    """Measures the forms of the conversion's shape class on `buf` (left unchanged) and makes
    the fastest the choice for the class; returns its form code (a stage code for transposes)."""
    d = Desc(_ptr(buf, _need(elem_bytes, rows, cols)).value, op, elem_bytes, rows, cols, tile)
    f = ctypes.c_int(0)
    if lib().hip_marshal_autotune_desc(_stream(stream), ctypes.byref(d), ctypes.byref(f)):
        raise RuntimeError("hip_marshal_autotune_desc failed")
    return f.value


def cycle_count(R, C, max_order=64):
    """(cycles, order) of the R x C transpose map when its order is at most max_order (host
    only), else (0, 0)."""
    k = ctypes.c_uint32(0)
    n = lib().hip_marshal_cycle_count(R, C, max_order, ctypes.byref(k))
    return (n, k.value) if n else (0, 0)


def plan_transpose(M, N, elem_bytes):
    """(a, b, stages) the planner picks for an M x N transpose (host only)."""
    a, b, st = ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int()
    if lib().hip_marshal_plan_transpose(M, N, elem_bytes, ctypes.byref(a), ctypes.byref(b),
                                        ctypes.byref(st)):
        raise ValueError("bad arguments")
    return a.value, b.value, st.value


def set_kernel_timing(on):
    """Starts (clearing the record) or stops per-launch timing of the library's kernels."""
    lib().hip_marshal_set_kernel_timing(1 if on else 0)


def kernel_times(max_entries=64):
    """Per kernel family since timing was turned on: launches, device ms and algorithmic bytes
    (waits for the recorded launches; clears the record)."""
    arr = (KernelTime * max_entries)()
    n = lib().hip_marshal_kernel_times(arr, max_entries)
    return [{"name": arr[i].name.decode(), "launches": arr[i].launches, "ms": arr[i].ms,
             "bytes": arr[i].bytes} for i in range(n)]


def last_route():
    return lib().hip_marshal_last_route().decode()


def finalize():
    lib().hip_marshal_finalize()


class Batch:
    """Many conversions in one graph launch (hip_marshal_batch_*).

    items: iterable of (tensor, op, elem_bytes, rows, cols, tile) with op one of OP_*.
    The tensors must stay alive while the batch exists.
    """

    def __init__(self, items, use_graph=True):
        items = list(items)
        arr = (Desc * len(items))()
        self._keep = []
        for i, (t, op, es, rows, cols, tile) in enumerate(items):
            arr[i] = Desc(_ptr(t, _need(es, rows, cols)).value, op, es, rows, cols, tile)
            self._keep.append(t)
        self._h = lib().hip_marshal_batch_create(arr, len(items), 1 if use_graph else 0)
        if not self._h:
            raise ValueError("hip_marshal_batch_create failed")

    def run(self, stream=None, elapsed=None):
        return lib().hip_marshal_batch_run(self._h, _stream(stream), _et(elapsed))

    def info(self, key):
        return lib().hip_marshal_batch_info(self._h, key.encode())

    def route(self, i):
        return lib().hip_marshal_batch_route(self._h, i).decode()

    def close(self):
        if self._h:
            lib().hip_marshal_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Server:
    """Continuous-batching request loop (hip_marshal_server_*)."""

    def __init__(self, hbm_budget_bytes, streams=4, max_batch=256):
        self._h = lib().hip_marshal_server_create(hbm_budget_bytes, streams, max_batch)
        if not self._h:
            raise ValueError("hip_marshal_server_create failed")

    @staticmethod
    def now():
        return lib().hip_marshal_server_now()

    def submit(self, tensor, op, elem_bytes, rows, cols, tile=0, arrival=0.0, ptr=None):
        """Submits a conversion of `tensor` (or of the raw device pointer `ptr`); returns its id."""
        p = ptr if ptr is not None else _ptr(tensor, _need(elem_bytes, rows, cols)).value
        d = Desc(p, op, elem_bytes, rows, cols, tile)
        rid = lib().hip_marshal_server_submit_at(self._h, ctypes.byref(d), arrival)
        if rid < 0:
            raise ValueError("request rejected")
        return rid

    def status(self, rid):
        return lib().hip_marshal_server_status(self._h, rid)

    def wait(self, rid, timeout=60.0):
        return lib().hip_marshal_server_wait(self._h, rid, timeout)

    def drain(self, timeout=600.0):
        return lib().hip_marshal_server_drain(self._h, timeout)

    def times(self, rid):
        a, l, d = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
        if lib().hip_marshal_server_request_times(self._h, rid, ctypes.byref(a), ctypes.byref(l),
                                                  ctypes.byref(d)):
            raise KeyError(rid)
        return a.value, l.value, d.value

    def stats(self):
        st = ServerStats()
        lib().hip_marshal_server_get_stats(self._h, ctypes.byref(st))
        return st.as_dict()

    def reset_stats(self):
        lib().hip_marshal_server_reset_stats(self._h)

    def close(self):
        if self._h:
            lib().hip_marshal_server_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def TransposeArrayInPlace(arr, a=1, b=1, stages=None, stream=None):
    """Transposes a 2-D tensor in place and returns the same storage viewed as the transpose.

    Mirrors TransposeArrayInPlace, src/py/transpose.py:108-116: the input's shape is left as it
    is, its data is re-laid out as the transpose, and the returned tensor shares the storage
    with shape (cols, rows). a and b must divide rows and cols. With the default a = b = 1
    the planner picks the tile factors; otherwise the 3-stage route runs with the given ones.
    """
    if arr.dim() != 2:
        raise TypeError("only allows transposing 2D arrays")
    rows, cols = arr.shape
    if rows % a or cols % b:
        raise TypeError("tile size must evenly divide array dimensions")
    if stages is None:
        stages = 0 if (a == 1 and b == 1) else 3
    if transpose(arr, rows // a, a, cols // b, b, 1, stages, stream=stream):
        raise RuntimeError("in-place transpose failed")
    return arr.view(-1)[: rows * cols].view(cols, rows)
