"""MI355X-native arbitrary-precision DECIMAL expression path.

Python host layer over the C ABI of ``libupdec.so`` (include/updec.h).
The compute path is HIP only: when the library is missing, importing this
package raises instead of falling back to anything else.  PyTorch is used
for nothing on the GPU side: device memory and streams come from the
library itself, so one HIP runtime serves the whole process.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libupdec.so")

MODE_AUTO, MODE_FAITHFUL, MODE_EXTENDED = 0, 1, 2
GEN_RANGE, GEN_DIGITS, GEN_BELOW = 0, 1, 2
IMAGE_BYTES = 133  # nullable_type<Decimal>: flag, sign, prec, frac, error, v[32]


class UpdecError(RuntimeError):
    pass


class ColumnDesc(ctypes.Structure):
    _fields_ = [("name", ctypes.c_char_p), ("precision", ctypes.c_int32), ("scale", ctypes.c_int32),
                ("bits", ctypes.c_int32)]


class AggDesc(ctypes.Structure):
    _fields_ = [("cols", ctypes.POINTER(ColumnDesc)), ("ncols", ctypes.c_int32),
                ("exprs", ctypes.POINTER(ctypes.c_char_p)), ("nexprs", ctypes.c_int32),
                ("has_filter", ctypes.c_int32), ("filter_op", ctypes.c_int32), ("filter_value", ctypes.c_int64),
                ("nkeys", ctypes.c_int32), ("mode", ctypes.c_int32), ("avg_mask", ctypes.c_uint32)]


class ColumnBlocks(ctypes.Structure):
    _fields_ = [("d_blocks", ctypes.c_void_p), ("d_starts", ctypes.c_void_p), ("nblocks", ctypes.c_int32),
                ("rows", ctypes.c_int64)]


class BlockColumn:
    """A compact column held as several device blocks (the reference's
    AriesColumnDataIterator data blocks): `blocks` are DeviceBuffers or
    views, `counts` their row counts.  Keeps the device arrays of block
    pointers and start rows alive."""

    def __init__(self, blocks, counts):
        import numpy as np
        if len(blocks) != len(counts) or not blocks:
            raise ValueError("one row count per block")
        self.blocks = list(blocks)
        starts = np.zeros(len(counts), dtype=np.int64)
        starts[1:] = np.cumsum(counts)[:-1]
        self.rows = int(np.sum(counts))
        self._ptrs = DeviceBuffer.from_numpy(np.array([b.ptr for b in blocks], dtype=np.uint64))
        self._starts = DeviceBuffer.from_numpy(starts)
        self.desc = ColumnBlocks(self._ptrs.ptr, self._starts.ptr, len(blocks), self.rows)


def _blocks_array(columns):
    arr = (ColumnBlocks * max(1, len(columns)))()
    for i, c in enumerate(columns):
        arr[i] = c.desc
    return arr


INDEX_NONE, INDEX_INT32, INDEX_NULLABLE_INT32 = 0, 1, 2


class ColumnIterDesc(ctypes.Structure):
    _fields_ = [("d_blocks", ctypes.c_void_p), ("d_starts", ctypes.c_void_p), ("nblocks", ctypes.c_int32),
                ("nullable", ctypes.c_int32), ("positions", ctypes.c_int64), ("index_kind", ctypes.c_int32),
                ("nindex_blocks", ctypes.c_int32), ("d_index", ctypes.c_void_p), ("d_index_blocks", ctypes.c_void_p),
                ("d_index_starts", ctypes.c_void_p)]


class ColumnIter:
    """A column as the reference's AriesColumnDataIterator: data `blocks`
    (DeviceBuffers or views) with `counts` rows each, rows of B(P) bytes or,
    with nullable=True, a flag byte then B(P) bytes; optional index blocks
    (DeviceBuffers of int32 positions, -1 = NULL, or of 5-byte
    nullable_type<int32> entries with index_kind=INDEX_NULLABLE_INT32) with
    `index_counts` entries each.  Keeps the device arrays alive."""

    def __init__(self, blocks, counts, nullable=False, index=None, index_counts=None, index_kind=INDEX_INT32):
        import numpy as np
        if len(blocks) != len(counts) or not blocks:
            raise ValueError("one row count per block")
        self.blocks = list(blocks)
        starts = np.zeros(len(counts), dtype=np.int64)
        starts[1:] = np.cumsum(counts)[:-1]
        self.positions = int(np.sum(counts))
        self._ptrs = DeviceBuffer.from_numpy(np.array([b.ptr for b in blocks], dtype=np.uint64))
        self._starts = DeviceBuffer.from_numpy(starts)
        self.desc = ColumnIterDesc(self._ptrs.ptr, self._starts.ptr, len(blocks), int(bool(nullable)), self.positions,
                                   INDEX_NONE, 0, None, None, None)
        self.index = None
        if index is not None:
            index = list(index)
            self.index = index
            self.desc.index_kind = index_kind
            self.desc.nindex_blocks = len(index)
            if len(index) == 1:
                self.desc.d_index = index[0].ptr
            else:
                if index_counts is None or len(index_counts) != len(index):
                    raise ValueError("one entry count per index block")
                ist = np.zeros(len(index), dtype=np.int64)
                ist[1:] = np.cumsum(index_counts)[:-1]
                self._iptrs = DeviceBuffer.from_numpy(np.array([b.ptr for b in index], dtype=np.uint64))
                self._istarts = DeviceBuffer.from_numpy(ist)
                self.desc.d_index_blocks = self._iptrs.ptr
                self.desc.d_index_starts = self._istarts.ptr


def _iter_array(columns):
    arr = (ColumnIterDesc * max(1, len(columns)))()
    for i, c in enumerate(columns):
        arr[i] = c.desc
    return arr


class GenDesc(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_int32), ("precision", ctypes.c_int32), ("digits", ctypes.c_int32),
                ("neg_permille", ctypes.c_int32), ("lo", ctypes.c_uint64), ("hi", ctypes.c_uint64),
                ("step", ctypes.c_uint64), ("seed", ctypes.c_uint64), ("column_id", ctypes.c_uint32),
                ("bound_words", ctypes.c_int32), ("bound", ctypes.c_uint32 * 16)]


def _load():
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is not built; run __graft_entry__.build() or ultraprecise_amd/build.py")
    lib = ctypes.CDLL(LIB_PATH)
    P, I32, I64, U64, SZ = ctypes.c_void_p, ctypes.c_int32, ctypes.c_int64, ctypes.c_uint64, ctypes.c_size_t
    sig = {
        "updec_last_error": (ctypes.c_char_p, []),
        "updec_version": (ctypes.c_int, []),
        "updec_init": (ctypes.c_int, [ctypes.c_int]),
        "updec_compact_bytes": (ctypes.c_int, [ctypes.c_int]),
        "updec_copy_kernel": (ctypes.c_int, [P, P, SZ, P]),
        "updec_malloc": (ctypes.c_int, [ctypes.POINTER(P), SZ]),
        "updec_free": (ctypes.c_int, [P]),
        "updec_malloc_async": (ctypes.c_int, [ctypes.POINTER(P), SZ, P]),
        "updec_free_async": (ctypes.c_int, [P, P]),
        "updec_memcpy_htod": (ctypes.c_int, [P, P, SZ, P]),
        "updec_memcpy_dtoh": (ctypes.c_int, [P, P, SZ, P]),
        "updec_memset": (ctypes.c_int, [P, ctypes.c_int, SZ, P]),
        "updec_stream_create": (ctypes.c_int, [ctypes.POINTER(P)]),
        "updec_stream_destroy": (ctypes.c_int, [P]),
        "updec_stream_sync": (ctypes.c_int, [P]),
        "updec_device_sync": (ctypes.c_int, []),
        "updec_mem_info": (ctypes.c_int, [ctypes.POINTER(I64), ctypes.POINTER(I64)]),
        "updec_device_info": (ctypes.c_int, [ctypes.POINTER(I32), ctypes.POINTER(I32), ctypes.POINTER(I64),
                                             ctypes.c_char_p, SZ]),
        "updec_event_create": (ctypes.c_int, [ctypes.POINTER(P)]),
        "updec_event_record": (ctypes.c_int, [P, P]),
        "updec_event_elapsed": (ctypes.c_int, [P, P, ctypes.POINTER(ctypes.c_float)]),
        "updec_event_destroy": (ctypes.c_int, [P]),
        "updec_event_query": (ctypes.c_int, [P, ctypes.POINTER(I32)]),
        "updec_event_sync": (ctypes.c_int, [P]),
        "updec_column_maxbits": (ctypes.c_int, [P, ctypes.c_int, I64, ctypes.POINTER(I32), P]),
        "updec_expr_compile": (ctypes.c_int, [ctypes.c_char_p, ctypes.POINTER(ColumnDesc), ctypes.c_int, ctypes.c_int,
                                              ctypes.POINTER(P)]),
        "updec_expr_info": (ctypes.c_int, [P, ctypes.POINTER(I32), ctypes.POINTER(I32), ctypes.POINTER(I32),
                                           ctypes.POINTER(I32)]),
        "updec_expr_work": (ctypes.c_int, [P, ctypes.POINTER(I64)]),
        "updec_mad_peak": (ctypes.c_int, [I32, I32, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double)]),
        "updec_expr_plan": (ctypes.c_int, [P, ctypes.c_char_p, SZ, ctypes.POINTER(SZ)]),
        "updec_expr_source": (ctypes.c_int, [P, ctypes.c_char_p, SZ, ctypes.POINTER(SZ)]),
        "updec_expr_source_iter": (ctypes.c_int, [P, ctypes.c_char_p, SZ, ctypes.POINTER(SZ)]),
        "updec_expr_lanes": (ctypes.c_int, [P, I32, ctypes.POINTER(I32)]),
        "updec_expr_source_ml": (ctypes.c_int, [P, I32, ctypes.c_char_p, SZ, ctypes.POINTER(SZ)]),
        "updec_expr_calc": (ctypes.c_int, [P, ctypes.POINTER(P), I64, P, P]),
        "updec_expr_free": (None, [P]),
        "updec_agg_compile": (ctypes.c_int, [ctypes.POINTER(AggDesc), ctypes.POINTER(P)]),
        "updec_agg_run": (ctypes.c_int, [P, ctypes.POINTER(P), P, P, P, I64, P, P, I32, ctypes.POINTER(I32),
                                         ctypes.POINTER(ctypes.c_uint16), P]),
        "updec_agg_last_kernel_ms": (ctypes.c_int, [P, ctypes.POINTER(ctypes.c_float)]),
        "updec_agg_source": (ctypes.c_int, [P, I32, ctypes.c_char_p, SZ, ctypes.POINTER(SZ)]),
        "updec_agg_source_iter": (ctypes.c_int, [P, I32, ctypes.c_char_p, SZ, ctypes.POINTER(SZ)]),
        "updec_agg_free": (None, [P]),
        "updec_div_int64": (ctypes.c_int, [P, I64, P, P, I64, ctypes.c_int, P]),
        "updec_segsum_compact": (ctypes.c_int, [P, ctypes.c_int, ctypes.c_int, I64, P, P, I32, P, ctypes.c_int, P]),
        "updec_segsum_compact_noalloc": (ctypes.c_int, [P, ctypes.c_int, ctypes.c_int, I64, P, P, I32, P, ctypes.c_int,
                                                        P]),
        "updec_segsum_noalloc_max_segments": (ctypes.c_int, [ctypes.c_int, ctypes.POINTER(I32)]),
        "updec_compile": (ctypes.c_int, [ctypes.c_char_p, SZ, ctypes.POINTER(P)]),
        "updec_launch": (ctypes.c_int, [P, ctypes.c_char_p, ctypes.c_uint32, ctypes.c_uint32, P, SZ, P]),
        "updec_module_free": (None, [P]),
        "updec_check_compile": (ctypes.c_int, [ctypes.c_char_p, SZ]),
        "updec_module_cache_stats": (ctypes.c_int, [ctypes.POINTER(I64)] * 3),
        "updec_gen_compact": (ctypes.c_int, [ctypes.POINTER(GenDesc), I64, I64, P, ctypes.c_int, P]),
        "updec_gen_int32": (ctypes.c_int, [U64, ctypes.c_uint32, I32, I32, I64, I64, P, ctypes.c_int, P]),
        "updec_gen_byte": (ctypes.c_int, [U64, ctypes.c_uint32, P, I32, I64, I64, P, ctypes.c_int, P]),
        "updec_vm_compile": (ctypes.c_int, [ctypes.c_char_p, ctypes.POINTER(ColumnDesc), ctypes.c_int, ctypes.c_int,
                                            ctypes.POINTER(P)]),
        "updec_vm_info": (ctypes.c_int, [P] + [ctypes.POINTER(I32)] * 6),
        "updec_vm_listing": (ctypes.c_int, [P, ctypes.c_char_p, SZ, ctypes.POINTER(SZ)]),
        "updec_vm_free": (None, [P]),
        "updec_batch_create": (ctypes.c_int, [I32, I32, ctypes.POINTER(P)]),
        "updec_batch_tuning": (ctypes.c_int, [P, I32, I64, I32]),
        "updec_batch_lanes": (ctypes.c_int, [P, I32, I32]),
        "updec_batch_add": (ctypes.c_int, [P, P, ctypes.POINTER(P), I64, P, P, ctypes.POINTER(I32)]),
        "updec_batch_pending": (ctypes.c_int, [P, ctypes.POINTER(I32)]),
        "updec_batch_launch": (ctypes.c_int, [P, ctypes.c_int, P]),
        "updec_batch_free": (None, [P]),
        "updec_parse_compact": (ctypes.c_int, [P, P, I64, I32, I32, P, P]),
        "updec_expr_calc_blocks": (ctypes.c_int, [P, ctypes.POINTER(ColumnBlocks), P, I64, P, P, P]),
        "updec_agg_run_blocks": (ctypes.c_int, [P, ctypes.POINTER(ColumnBlocks), P, I64, P, P, P]),
        "updec_expr_calc_iter": (ctypes.c_int, [P, ctypes.POINTER(ColumnIterDesc), I64, P, I32, P, P]),
        "updec_agg_run_iter": (ctypes.c_int, [P, ctypes.POINTER(ColumnIterDesc), I64, P, P, P, P, P, I32,
                                              ctypes.POINTER(I32), ctypes.POINTER(ctypes.c_uint16), P]),
        "updec_format_stride": (ctypes.c_int, [I32, I32, I32, ctypes.POINTER(I64)]),
        "updec_format_compact": (ctypes.c_int, [P, I64, I32, I32, I32, P, I64, P]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib


_lib = _load()

# C ABI symbols declared in include/updec.h (checked by the CPU tests)
EXPORTED = ["updec_last_error", "updec_version", "updec_init", "updec_compact_bytes", "updec_column_maxbits",
            "updec_malloc", "updec_free", "updec_malloc_async", "updec_free_async", "updec_memcpy_htod", "updec_memcpy_dtoh", "updec_memset",
            "updec_stream_create", "updec_stream_destroy", "updec_stream_sync", "updec_device_sync",
            "updec_device_info", "updec_mem_info", "updec_event_create", "updec_event_record", "updec_event_elapsed",
            "updec_event_destroy", "updec_event_query", "updec_event_sync", "updec_copy_kernel",
            "updec_expr_compile", "updec_expr_info", "updec_expr_plan", "updec_expr_work", "updec_mad_peak", "updec_expr_source", "updec_expr_calc",
            "updec_expr_free", "updec_agg_compile", "updec_agg_run", "updec_agg_last_kernel_ms", "updec_agg_source",
            "updec_agg_free", "updec_div_int64", "updec_segsum_compact", "updec_segsum_compact_noalloc",
            "updec_segsum_noalloc_max_segments", "updec_compile", "updec_launch",
            "updec_module_free", "updec_check_compile", "updec_module_cache_stats", "updec_gen_compact", "updec_gen_int32", "updec_gen_byte",
            "updec_vm_compile", "updec_vm_info", "updec_vm_listing", "updec_vm_free", "updec_batch_create",
            "updec_batch_tuning", "updec_batch_lanes", "updec_batch_add", "updec_batch_pending", "updec_batch_launch", "updec_batch_free",
            "updec_parse_compact", "updec_format_stride", "updec_format_compact", "updec_expr_calc_blocks",
            "updec_agg_run_blocks", "updec_expr_calc_iter", "updec_agg_run_iter", "updec_expr_source_iter",
            "updec_agg_source_iter", "updec_expr_lanes", "updec_expr_source_ml"]


def lib():
    return _lib


def _check(rc):
    if rc != 0:
        raise UpdecError(f"updec error {rc}: {_lib.updec_last_error().decode(errors='replace')}")


def _text(fn, h, *extra):
    need = ctypes.c_size_t(0)
    _check(fn(h, *extra, None, 0, ctypes.byref(need)))
    buf = ctypes.create_string_buffer(need.value)
    _check(fn(h, *extra, buf, need.value, ctypes.byref(need)))
    return buf.value.decode()


def compact_bytes(precision):
    return _lib.updec_compact_bytes(precision)


def _ptr(t):
    return ctypes.c_void_p(0 if t is None else t.ptr)


def _stream(stream):
    return ctypes.c_void_p(0 if stream is None else stream.handle)


class Stream:
    """A HIP stream created by the library (default priority)."""

    def __init__(self):
        h = ctypes.c_void_p()
        _check(_lib.updec_stream_create(ctypes.byref(h)))
        self.handle = h.value

    def sync(self):
        _check(_lib.updec_stream_sync(ctypes.c_void_p(self.handle)))

    def __del__(self):
        if getattr(self, "handle", None):
            _lib.updec_stream_destroy(ctypes.c_void_p(self.handle))
            self.handle = None


class DeviceBuffer:
    """Device memory owned by this object (hipMalloc through the library;
    with ordered=True stream-ordered hipMallocAsync / hipFreeAsync on
    `stream`, which frees without a device-wide synchronization)."""

    def __init__(self, nbytes, ordered=False, stream=None):
        self.nbytes = int(nbytes)
        self.ordered = ordered
        self.stream = stream
        p = ctypes.c_void_p()
        if ordered:
            _check(_lib.updec_malloc_async(ctypes.byref(p), max(4, self.nbytes), _stream(stream)))
        else:
            _check(_lib.updec_malloc(ctypes.byref(p), max(4, self.nbytes)))
        self.ptr = p.value

    @classmethod
    def from_numpy(cls, arr, stream=None):
        import numpy as np
        a = np.ascontiguousarray(arr)
        b = cls(a.nbytes)
        _check(_lib.updec_memcpy_htod(ctypes.c_void_p(b.ptr), a.ctypes.data_as(ctypes.c_void_p), a.nbytes,
                                      _stream(stream)))
        return b

    def numpy(self, dtype="uint8", count=None, stream=None):
        import numpy as np
        dt = np.dtype(dtype)
        n = self.nbytes // dt.itemsize if count is None else int(count)
        out = np.empty(n, dtype=dt)
        _check(_lib.updec_memcpy_dtoh(out.ctypes.data_as(ctypes.c_void_p), ctypes.c_void_p(self.ptr), out.nbytes,
                                      _stream(stream)))
        return out

    def zero(self, stream=None):
        _check(_lib.updec_memset(ctypes.c_void_p(self.ptr), 0, self.nbytes, _stream(stream)))
        return self

    def __del__(self):
        if getattr(self, "ptr", None):
            if getattr(self, "ordered", False):
                _lib.updec_free_async(ctypes.c_void_p(self.ptr), _stream(self.stream))
            else:
                _lib.updec_free(ctypes.c_void_p(self.ptr))
            self.ptr = None


class Event:
    def __init__(self):
        h = ctypes.c_void_p()
        _check(_lib.updec_event_create(ctypes.byref(h)))
        self.handle = h.value

    def record(self, stream=None):
        _check(_lib.updec_event_record(ctypes.c_void_p(self.handle), _stream(stream)))

    def done(self):
        d = ctypes.c_int32()
        _check(_lib.updec_event_query(ctypes.c_void_p(self.handle), ctypes.byref(d)))
        return bool(d.value)

    def sync(self):
        _check(_lib.updec_event_sync(ctypes.c_void_p(self.handle)))

    def elapsed_ms(self, end):
        ms = ctypes.c_float(0)
        _check(_lib.updec_event_elapsed(ctypes.c_void_p(self.handle), ctypes.c_void_p(end.handle), ctypes.byref(ms)))
        return ms.value

    def __del__(self):
        if getattr(self, "handle", None):
            _lib.updec_event_destroy(ctypes.c_void_p(self.handle))
            self.handle = None


def copy_kernel(dst, src, nbytes, stream=None):
    _check(_lib.updec_copy_kernel(_ptr(dst), _ptr(src), nbytes, _stream(stream)))


def device_sync():
    _check(_lib.updec_device_sync())


def mad_peak(blocks=None, iters=20000):
    """measured issue rate of 32 x 32 -> 64 multiply-adds: (ms, MACs, MACs/s)"""
    if blocks is None:
        blocks = device_info()["cus"] * 8
    ms, macs = ctypes.c_double(), ctypes.c_double()
    _check(_lib.updec_mad_peak(blocks, iters, ctypes.byref(ms), ctypes.byref(macs)))
    return ms.value, macs.value, macs.value / (ms.value / 1000.0)


def mem_info():
    """(free, total) device memory in bytes (hipMemGetInfo)"""
    f, t = ctypes.c_int64(), ctypes.c_int64()
    _check(_lib.updec_mem_info(ctypes.byref(f), ctypes.byref(t)))
    return f.value, t.value


def device_info():
    cus, clk, mem = ctypes.c_int32(), ctypes.c_int32(), ctypes.c_int64()
    name = ctypes.create_string_buffer(256)
    _check(_lib.updec_device_info(ctypes.byref(cus), ctypes.byref(clk), ctypes.byref(mem), name, 256))
    return {"cus": cus.value, "clock_khz": clk.value, "mem_bytes": mem.value, "name": name.value.decode()}


def _col_array(cols):
    arr = (ColumnDesc * len(cols))()
    keep = []
    for i, c in enumerate(cols):
        name, p, s = c[0], c[1], c[2]
        bits = c[3] if len(c) > 3 else 0
        b = name.encode()
        keep.append(b)
        arr[i] = ColumnDesc(b, p, s, bits)
    return arr, keep


def init(device=0):
    _check(_lib.updec_init(device))


class Expr:
    """A compiled decimal expression: parse, the reference's alignment
    rewrite, typing, HIP code generation and (lazily) hipRTC compilation.
    ``cols`` is a list of (name, precision, scale[, bits])."""

    def __init__(self, text, cols, mode=MODE_AUTO):
        self.cols = [tuple(c) for c in cols]
        arr, keep = _col_array(self.cols)
        h = ctypes.c_void_p()
        _check(_lib.updec_expr_compile(text.encode(), arr, len(self.cols), mode, ctypes.byref(h)))
        self._h = h
        P, S, B, F = ctypes.c_int32(), ctypes.c_int32(), ctypes.c_int32(), ctypes.c_int32()
        _check(_lib.updec_expr_info(h, ctypes.byref(P), ctypes.byref(S), ctypes.byref(B), ctypes.byref(F)))
        self.precision, self.scale, self.out_bytes, self.faithful = P.value, S.value, B.value, bool(F.value)

    @property
    def plan(self):
        return _text(_lib.updec_expr_plan, self._h)

    @property
    def limb_products(self):
        """algorithmic 32 x 32 -> 64 limb products per row"""
        v = ctypes.c_int64()
        _check(_lib.updec_expr_work(self._h, ctypes.byref(v)))
        return v.value

    @property
    def source(self):
        return _text(_lib.updec_expr_source, self._h)

    @property
    def multi_lane_width(self):
        """limbs W of the multi-lane form (0: one row per lane only)"""
        w = ctypes.c_int32()
        _check(_lib.updec_expr_lanes(self._h, 0, ctypes.byref(w)))
        return w.value

    def set_lanes(self, lanes_per_row):
        """lanes per row of calc(): 1, or W/4 or W/8 (at most 64) for an
        expression with a multi-lane form"""
        _check(_lib.updec_expr_lanes(self._h, int(lanes_per_row), None))

    def ml_source(self, lanes_per_row):
        return _text(_lib.updec_expr_source_ml, self._h, ctypes.c_int32(int(lanes_per_row)))

    def iter_source(self):
        """source of the column-iterator variant (calc_iter)"""
        return _text(_lib.updec_expr_source_iter, self._h)

    def calc_blocks(self, columns, rows, out, indices=None, valid=None, stream=None):
        """columns: BlockColumns in descriptor order; output row i is the
        expression at position indices[i] (int32 DeviceBuffer, -1 = NULL)
        or i; valid (optional DeviceBuffer, one byte per row) gets 1 / 0"""
        _check(_lib.updec_expr_calc_blocks(self._h, _blocks_array(columns), _ptr(indices), rows, _ptr(out), _ptr(valid),
                                           _stream(stream)))

    def calc_iter(self, columns, rows, out, out_nullable=False, valid=None, stream=None):
        """columns: ColumnIters in descriptor order; output row i of out_bytes
        (or 1 + out_bytes with out_nullable: flag byte first); valid
        (optional DeviceBuffer, one byte per row) gets 1 / 0 (NULL)"""
        _check(_lib.updec_expr_calc_iter(self._h, _iter_array(columns), rows, _ptr(out), int(bool(out_nullable)),
                                         _ptr(valid), _stream(stream)))

    def calc(self, columns, rows, out, stream=None):
        """columns: DeviceBuffers in descriptor order; out: DeviceBuffer of
        at least rows * out_bytes bytes, 4-byte aligned."""
        ptrs = (ctypes.c_void_p * max(1, len(columns)))(*[c.ptr for c in columns])
        _check(_lib.updec_expr_calc(self._h, ptrs, rows, _ptr(out), _stream(stream)))

    def __del__(self):
        if getattr(self, "_h", None):
            _lib.updec_expr_free(self._h)
            self._h = None


FILTER_OPS = {"<=": 0, "<": 1, ">=": 2, ">": 3, "==": 4, "!=": 5}


class Agg:
    """Fused SUM(expr_k) / COUNT(*) with an optional int32 filter and up to
    two one-byte group keys.  avg: indices k of the expressions whose
    AVG = SUM / COUNT(*) is also produced, after the SUM images."""

    def __init__(self, exprs, cols, filter=None, nkeys=0, mode=MODE_AUTO, avg=()):
        self.cols = [tuple(c) for c in cols]
        self.exprs = list(exprs)
        self.avg = sorted(set(int(k) for k in avg))
        mask = 0
        for k in self.avg:
            if not 0 <= k < min(32, len(self.exprs)):
                raise ValueError("avg names an expression past the list")
            mask |= 1 << k
        arr, keep = _col_array(self.cols)
        ex = (ctypes.c_char_p * len(self.exprs))(*[e.encode() for e in self.exprs])
        d = AggDesc(arr, len(self.cols), ex, len(self.exprs), 0, 0, 0, nkeys, mode, mask)
        if filter is not None:
            d.has_filter, d.filter_op, d.filter_value = 1, FILTER_OPS[filter[0]], int(filter[1])
        h = ctypes.c_void_p()
        _check(_lib.updec_agg_compile(ctypes.byref(d), ctypes.byref(h)))
        self._h = h
        self.nkeys = nkeys

    def source(self, ngroups=1, iterators=False):
        fn = _lib.updec_agg_source_iter if iterators else _lib.updec_agg_source
        return _text(fn, self._h, ctypes.c_int32(ngroups))

    def run_blocks(self, columns, rows, images, counts, indices=None, stream=None):
        """SUM of each expression over BlockColumns, at positions indices[0..rows)
        (or 0..rows); NULL positions are skipped.  Ungrouped, unfiltered."""
        _check(_lib.updec_agg_run_blocks(self._h, _blocks_array(columns), _ptr(indices), rows, _ptr(images),
                                         _ptr(counts), _stream(stream)))

    def run_iter(self, columns, rows, images, counts, filter_col=None, key0=None, key1=None, max_groups=255,
                 stream=None):
        """SUM / COUNT(*) over ColumnIters with the filter and keys given per
        launch row; NULL values add to no SUM.  Returns (ngroups, keys)."""
        ng = ctypes.c_int32(0)
        keys = (ctypes.c_uint16 * max(1, max_groups))()
        _check(_lib.updec_agg_run_iter(self._h, _iter_array(columns), rows, _ptr(filter_col), _ptr(key0), _ptr(key1),
                                       _ptr(images), _ptr(counts), max_groups, ctypes.byref(ng), keys,
                                       _stream(stream)))
        return ng.value, [keys[i] for i in range(ng.value)]

    def run(self, columns, rows, images, counts, filter_col=None, key0=None, key1=None, max_groups=255, stream=None):
        """Returns (ngroups, keys).  images: DeviceBuffer of at least
        max_groups * (nexprs + len(avg)) * 133 bytes, SUM image of expression k
        in group g at g * nexprs + k, AVG j at max_groups * nexprs + g * len(avg)
        + j; counts: DeviceBuffer of max_groups int64."""
        # the argument arrays are kept between calls with the same columns
        # and group bound (a serving or benchmark loop repeats them)
        cols = tuple(c.ptr for c in columns)
        cached = getattr(self, "_run_args", None)
        if cached is None or cached[0] != cols or cached[1] != max_groups:
            cached = (cols, max_groups, (ctypes.c_void_p * max(1, len(cols)))(*cols), ctypes.c_int32(0),
                      (ctypes.c_uint16 * max(1, max_groups))())
            self._run_args = cached
        ptrs, ng, keys = cached[2], cached[3], cached[4]
        _check(_lib.updec_agg_run(self._h, ptrs, _ptr(filter_col), _ptr(key0), _ptr(key1), rows, _ptr(images),
                                  _ptr(counts), max_groups, ctypes.byref(ng), keys, _stream(stream)))
        n = ng.value
        return n, list(keys[:n])

    def last_kernel_ms(self):
        ms = ctypes.c_float(0)
        _check(_lib.updec_agg_last_kernel_ms(self._h, ctypes.byref(ms)))
        return ms.value

    def __del__(self):
        if getattr(self, "_h", None):
            _lib.updec_agg_free(self._h)
            self._h = None


def div_int64(sums, stride, counts, out, n, mode=MODE_AUTO, stream=None, sums_offset=0, out_offset=0):
    """AVG images out[i] = sums[i] / counts[i]; offsets are byte offsets
    into the two buffers (e.g. one expression column of a [group][expr]
    image array)."""
    _check(_lib.updec_div_int64(ctypes.c_void_p(sums.ptr + sums_offset), stride, _ptr(counts),
                                ctypes.c_void_p(out.ptr + out_offset), n, mode, _stream(stream)))


def segsum_compact(col, precision, scale, rows, associated, groups, ngroups, out, mode=MODE_AUTO, stream=None):
    _check(_lib.updec_segsum_compact(_ptr(col), precision, scale, rows, _ptr(associated), _ptr(groups), ngroups,
                                     _ptr(out), mode, _stream(stream)))


FORMAT_HEX, FORMAT_TEXT = 0, 1


def parse_compact(chars, offsets, rows, precision, scale, out, stream=None):
    """Decimal(P, S, text) -> compact rows for a text column on the device:
    row r is chars[offsets[r] .. offsets[r+1])"""
    _check(_lib.updec_parse_compact(_ptr(chars), _ptr(offsets), rows, precision, scale, _ptr(out), _stream(stream)))


def format_stride(precision, scale, form):
    """bytes per output slot of format_compact"""
    n = ctypes.c_int64(0)
    _check(_lib.updec_format_stride(precision, scale, form, ctypes.byref(n)))
    return n.value


def format_compact(col, rows, precision, scale, form, out, stride, stream=None):
    """GetDecimal hex (FORMAT_HEX) or decimal text (FORMAT_TEXT) of every
    compact row, into stride-byte NUL-padded slots"""
    _check(_lib.updec_format_compact(_ptr(col), rows, precision, scale, form, _ptr(out), stride, _stream(stream)))


def column_maxbits(col, precision, rows, stream=None):
    b = ctypes.c_int32(0)
    _check(_lib.updec_column_maxbits(_ptr(col), precision, rows, ctypes.byref(b), _stream(stream)))
    return b.value


def module_cache_stats():
    """(hipRTC compiles, in-memory hits, on-disk hits) since the process started"""
    c, m, d = ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int64()
    _check(_lib.updec_module_cache_stats(ctypes.byref(c), ctypes.byref(m), ctypes.byref(d)))
    return c.value, m.value, d.value


def check_compile(src):
    b = src.encode()
    _check(_lib.updec_check_compile(b, len(b)))


def gen_desc(kind, precision, seed, column_id, digits=0, lo=0, hi=0, step=1, neg_permille=0, bound=0):
    """`bound` (GEN_BELOW): magnitudes are drawn uniformly from [0, bound)."""
    d = GenDesc(kind, precision, digits, neg_permille, lo, hi, step, seed, column_id)
    if kind == GEN_BELOW:
        words = []
        while bound:
            words.append(bound & 0xFFFFFFFF)
            bound >>= 32
        if not 1 <= len(words) <= 16:
            raise ValueError("bound must be a positive integer below 2^512")
        d.bound_words = len(words)
        for i, w in enumerate(words):
            d.bound[i] = w
    return d


def gen_compact_host(desc, rows, row0=0):
    """Generate a compact column on the host; returns a numpy uint8 array."""
    import numpy as np
    nb = compact_bytes(desc.precision)
    out = np.empty(rows * nb, dtype=np.uint8)
    _check(_lib.updec_gen_compact(ctypes.byref(desc), row0, rows, out.ctypes.data_as(ctypes.c_void_p), 0, None))
    return out


class DeviceView:
    """A range of a DeviceBuffer (no ownership): usable wherever a buffer's
    .ptr is read."""

    def __init__(self, base, offset, nbytes):
        self.base = base
        self.ptr = base.ptr + offset
        self.nbytes = nbytes


def gen_compact_into(desc, rows, dst, row0=0, stream=None):
    """generate a compact column into an existing device buffer or view"""
    if dst.nbytes < rows * compact_bytes(desc.precision):
        raise ValueError("destination too small")
    _check(_lib.updec_gen_compact(ctypes.byref(desc), row0, rows, _ptr(dst), 1, _stream(stream)))


def gen_compact_device(desc, rows, row0=0, stream=None):
    nb = compact_bytes(desc.precision)
    out = DeviceBuffer(max(4, rows * nb + 16))
    _check(_lib.updec_gen_compact(ctypes.byref(desc), row0, rows, _ptr(out), 1, _stream(stream)))
    return out


def gen_int32_host(seed, column_id, lo, hi, rows, row0=0):
    import numpy as np
    out = np.empty(rows, dtype=np.int32)
    _check(_lib.updec_gen_int32(seed, column_id, lo, hi, row0, rows, out.ctypes.data_as(ctypes.c_void_p), 0, None))
    return out


def gen_int32_device(seed, column_id, lo, hi, rows, row0=0, stream=None):
    out = DeviceBuffer(max(4, rows * 4))
    _check(_lib.updec_gen_int32(seed, column_id, lo, hi, row0, rows, _ptr(out), 1, _stream(stream)))
    return out


def gen_byte_host(seed, column_id, choices, rows, row0=0):
    import numpy as np
    ch = (ctypes.c_uint8 * len(choices))(*choices)
    out = np.empty(rows, dtype=np.uint8)
    _check(_lib.updec_gen_byte(seed, column_id, ch, len(choices), row0, rows, out.ctypes.data_as(ctypes.c_void_p), 0,
                               None))
    return out


def gen_byte_device(seed, column_id, choices, rows, row0=0, stream=None):
    ch = (ctypes.c_uint8 * len(choices))(*choices)
    out = DeviceBuffer(max(4, rows))
    _check(_lib.updec_gen_byte(seed, column_id, ch, len(choices), row0, rows, _ptr(out), 1, _stream(stream)))
    return out


VM_CLASS_WIDTHS = (2, 4, 8, 16, 32, 64, 128, 256, 512)
VM_FIRST_MULTI_LANE = 5   # classes from here spread a value over W / 4 or W / 8 lanes
EFULL = 5   # UPDEC_EFULL: the batch has no room for the query


class VmProgram:
    """An expression compiled for the precompiled interpreter (no hipRTC):
    the same parse / rewrite / typing as Expr, run as a register program."""

    def __init__(self, text, cols, mode=MODE_AUTO):
        self.text = text
        self.cols = [tuple(c) for c in cols]
        arr, keep = _col_array(self.cols)
        h = ctypes.c_void_p()
        _check(_lib.updec_vm_compile(text.encode(), arr, len(self.cols), mode, ctypes.byref(h)))
        self._h = h
        v = [ctypes.c_int32() for _ in range(6)]
        _check(_lib.updec_vm_info(self._h, *[ctypes.byref(x) for x in v]))
        self.width, self.nins, self.nregs, self.precision, self.scale, f = [x.value for x in v]
        self.faithful = bool(f)
        self.width_class = VM_CLASS_WIDTHS.index(self.width)

    @property
    def listing(self):
        n = ctypes.c_size_t()
        _check(_lib.updec_vm_listing(self._h, None, 0, ctypes.byref(n)))
        buf = ctypes.create_string_buffer(n.value)
        _check(_lib.updec_vm_listing(self._h, buf, n.value, ctypes.byref(n)))
        return buf.value.decode()

    def __del__(self):
        if getattr(self, "_h", None):
            _lib.updec_vm_free(self._h)
            self._h = None


class Batch:
    """Queries collected for one set of interpreter launches.  add() stages
    a query (None when the batch is full: launch it first); launch() runs
    every staged query (as one HIP graph when graph=True) and writes each
    query's SUM image and COUNT."""

    def __init__(self, max_queries=1024, max_items=65536):
        h = ctypes.c_void_p()
        _check(_lib.updec_batch_create(max_queries, max_items, ctypes.byref(h)))
        self._h = h
        self._keep = []

    def tune(self, width_class, rows_per_item, blocks_per_cu, lanes=None):
        _check(_lib.updec_batch_tuning(self._h, width_class, rows_per_item, blocks_per_cu))
        if lanes is not None and width_class >= VM_FIRST_MULTI_LANE:
            _check(_lib.updec_batch_lanes(self._h, width_class, lanes))

    def add(self, prog, columns, rows, image, count, image_offset=0, count_offset=0):
        """columns: DeviceBuffers in descriptor order; image: 133 bytes at
        image_offset; count: one int64 at count_offset."""
        ptrs = (ctypes.c_void_p * max(1, len(columns)))(*[c.ptr for c in columns])
        idx = ctypes.c_int32()
        rc = _lib.updec_batch_add(self._h, prog._h, ptrs, rows, ctypes.c_void_p(image.ptr + image_offset),
                                  ctypes.c_void_p(count.ptr + count_offset), ctypes.byref(idx))
        if rc == EFULL:
            return None
        _check(rc)
        self._keep.append(prog)
        return idx.value

    @property
    def pending(self):
        n = ctypes.c_int32()
        _check(_lib.updec_batch_pending(self._h, ctypes.byref(n)))
        return n.value

    def launch(self, graph=True, stream=None):
        _check(_lib.updec_batch_launch(self._h, 1 if graph else 0, _stream(stream)))
        self._keep = []

    def __del__(self):
        if getattr(self, "_h", None):
            _lib.updec_batch_free(self._h)
            self._h = None
