"""TEST INFRASTRUCTURE ONLY: ctypes binding of the C oracle (updec_oracle.c).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
import this module, and only as the checker / the timed CPU baseline.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liboracle.so")

FAITHFUL, EXTENDED = 0, 1
OD_MAXW = 192


class OdDec(ctypes.Structure):
    _fields_ = [("sign", ctypes.c_int32), ("prec", ctypes.c_int32), ("frac", ctypes.c_int32),
                ("error", ctypes.c_int32), ("v", ctypes.c_uint32 * OD_MAXW)]

    @property
    def magnitude(self):
        m = 0
        for i in reversed(range(OD_MAXW)):
            m = (m << 32) | self.v[i]
        return m

    def set_magnitude(self, m):
        for i in range(OD_MAXW):
            self.v[i] = m & 0xffffffff
            m >>= 32


def build():
    import subprocess
    subprocess.check_call(["make", "-s", "-C", _HERE])


def _load(path=LIB_PATH):
    if not os.path.exists(path):
        build()
    lib = ctypes.CDLL(path)
    D = ctypes.POINTER(OdDec)
    P, I64 = ctypes.c_void_p, ctypes.c_int64
    sig = {
        "od_compact_bytes": (ctypes.c_int, [ctypes.c_int]),
        "od_zero": (None, [D]),
        "od_from_compact": (None, [D, P, ctypes.c_int, ctypes.c_int]),
        "od_to_compact": (None, [D, P, ctypes.c_int]),
        "od_from_string": (ctypes.c_int, [D, ctypes.c_char_p, ctypes.c_int]),
        "od_cast": (None, [D, ctypes.c_int, ctypes.c_int, ctypes.c_int]),
        "od_truncate": (None, [D, ctypes.c_int, ctypes.c_int]),
        "od_from_int32_scale": (None, [D, ctypes.c_int32, ctypes.c_int32, ctypes.c_int]),
        "od_from_int64": (None, [D, I64]),
        "od_rescale_up": (None, [D, ctypes.c_int, ctypes.c_int]),
        "od_shl10": (None, [D, ctypes.c_int, ctypes.c_int]),
        "od_shr10": (None, [D, ctypes.c_int, ctypes.c_int]),
        "od_neg": (None, [D]),
        "od_add": (None, [D, D, ctypes.c_int]),
        "od_sub": (None, [D, D, ctypes.c_int]),
        "od_mul": (None, [D, D, ctypes.c_int]),
        "od_div": (None, [D, D, ctypes.c_int]),
        "od_mod": (None, [D, D, ctypes.c_int]),
        "od_cmp": (ctypes.c_int, [D, D, ctypes.c_int]),
        "od_is_zero": (ctypes.c_int, [D]),
        "od_bitlen": (ctypes.c_int, [D]),
        "od_to_nullable_image": (None, [D, P]),
        "od_to_string": (ctypes.c_int, [D, ctypes.c_char_p, ctypes.c_int]),
        "od_overflow_flag": (ctypes.c_int, []),
        "od_clear_overflow": (None, []),
        "od_plan_parse": (P, [ctypes.c_char_p]),
        "od_plan_free": (None, [P]),
        "od_plan_result": (ctypes.c_int, [P, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int)]),
        "od_plan_eval": (ctypes.c_int, [P, P, P, P, I64, P, ctypes.c_int, ctypes.c_int, ctypes.c_int]),
        "od_plan_sum": (I64, [P, P, P, P, I64, P, P, ctypes.c_int, ctypes.c_int]),
        "od_plan_group_sum": (ctypes.c_int, [P, P, P, P, I64, P, ctypes.c_int, P, P, ctypes.c_int, ctypes.c_int]),
        "od_segsum_compact": (ctypes.c_int, [P, ctypes.c_int, ctypes.c_int, I64, P, P, ctypes.c_int, P, ctypes.c_int,
                                             ctypes.c_int]),
        "od_avg_from_sum": (None, [P, I64, P, ctypes.c_int]),
        "od_format_hex": (ctypes.c_int, [D, ctypes.c_char_p]),
        "od_parse_compact": (None, [P, P, I64, ctypes.c_int, ctypes.c_int, P]),
        "od_format_compact": (None, [P, I64, ctypes.c_int, ctypes.c_int, ctypes.c_int, P, I64]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib


lib = _load()


_wide_lib = None


def plan_sum_wide(plan_text, cols, precisions, scales, rows, mode=EXTENDED, threads=8):
    """plan_sum with 640-limb extended values (liboracle_wide.so, built with
    OD_MAXW = 640): for plans whose exact intermediates pass 192 limbs"""
    import numpy as np
    global _wide_lib
    if _wide_lib is None:
        _wide_lib = _load(os.path.join(_HERE, "liboracle_wide.so"))
    parts = plan_text.split(" ", 4)
    if parts[0] in ("F", "X") and len(parts) == 5:
        plan_text = parts[4]
    h = _wide_lib.od_plan_parse(plan_text.encode())
    if not h:
        raise ValueError("unparsable plan: " + plan_text)
    try:
        arrs, ptrs, P, S = _colargs(cols, precisions, scales)
        out = ctypes.create_string_buffer(133)
        n = _wide_lib.od_plan_sum(h, ptrs, P, S, rows, None, out, mode, threads)
    finally:
        _wide_lib.od_plan_free(h)
    return out.raw, n


def plan_eval_wide(plan_text, cols, precisions, scales, rows, out_bytes, mode=EXTENDED, threads=8):
    """plan_eval with 640-limb extended values (liboracle_wide.so)"""
    import numpy as np
    global _wide_lib
    if _wide_lib is None:
        _wide_lib = _load(os.path.join(_HERE, "liboracle_wide.so"))
    parts = plan_text.split(" ", 4)
    if parts[0] in ("F", "X") and len(parts) == 5:
        plan_text = parts[4]
    h = _wide_lib.od_plan_parse(plan_text.encode())
    if not h:
        raise ValueError("unparsable plan: " + plan_text)
    try:
        arrs, ptrs, P, S = _colargs(cols, precisions, scales)
        out = np.zeros(rows * out_bytes, dtype=np.uint8)
        _wide_lib.od_plan_eval(h, ptrs, P, S, rows, out.ctypes.data, out_bytes, mode, threads)
    finally:
        _wide_lib.od_plan_free(h)
    return out


def use_native():
    """Rebuild the oracle with -O3 -march=native for the host this process
    runs on (make native) and route every call below through it.  Used by
    the CPU baseline leg only, before its timed region; returns the library
    path, or None when the native build fails (the portable build stays)."""
    import subprocess
    global lib
    path = os.path.join(_HERE, "liboracle_native.so")
    try:
        subprocess.check_call(["make", "-s", "-B", "-C", _HERE, "native"], stdout=subprocess.DEVNULL)
        lib = _load(path)
    except (OSError, subprocess.CalledProcessError):
        return None
    return path


def host_cpu():
    """(model name, logical CPUs visible, CPUs this process may run on)"""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        allowed = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        allowed = os.cpu_count() or 1
    return model, os.cpu_count() or 1, allowed


def dec(s, mode=FAITHFUL):
    d = OdDec()
    lib.od_from_string(ctypes.byref(d), s.encode(), mode)
    return d


def dec_cast(s, P, S, mode=FAITHFUL):
    """Decimal(P, S, "s") (decimal.cu:154-165): parse then cast to scale S."""
    d = dec(s, mode)
    lib.od_cast(ctypes.byref(d), P, S, mode)
    return d


def to_str(d):
    buf = ctypes.create_string_buffer(4096)
    lib.od_to_string(ctypes.byref(d), buf, 4096)
    return buf.value.decode()


def copy(d):
    e = OdDec()
    ctypes.memmove(ctypes.byref(e), ctypes.byref(d), ctypes.sizeof(OdDec))
    return e


def binop(op, a, b, mode=FAITHFUL):
    r = copy(a)
    getattr(lib, "od_" + op)(ctypes.byref(r), ctypes.byref(b), mode)
    return r


def from_compact(buf, P, S):
    d = OdDec()
    b = bytes(buf)
    lib.od_from_compact(ctypes.byref(d), b, P, S)
    return d


def to_compact(d, nbytes):
    out = ctypes.create_string_buffer(nbytes)
    lib.od_to_compact(ctypes.byref(d), out, nbytes)
    return out.raw


def image(d):
    out = ctypes.create_string_buffer(133)
    lib.od_to_nullable_image(ctypes.byref(d), out)
    return out.raw


class Plan:
    def __init__(self, text):
        # accept the product's "F|X P S B <plan>" header form too
        parts = text.split(" ", 4)
        if parts[0] in ("F", "X") and len(parts) == 5:
            text = parts[4]
        self._h = lib.od_plan_parse(text.encode())
        if not self._h:
            raise ValueError("unparsable plan: " + text)
        P, S = ctypes.c_int(), ctypes.c_int()
        lib.od_plan_result(self._h, ctypes.byref(P), ctypes.byref(S))
        self.P, self.S = P.value, S.value

    def __del__(self):
        if getattr(self, "_h", None):
            lib.od_plan_free(self._h)
            self._h = None


def _colargs(cols, precisions, scales):
    import numpy as np
    n = len(cols)
    arrs = [np.ascontiguousarray(c, dtype=np.uint8) for c in cols]
    ptrs = (ctypes.c_void_p * max(1, n))(*[a.ctypes.data for a in arrs])
    P = (ctypes.c_int * max(1, n))(*precisions)
    S = (ctypes.c_int * max(1, n))(*scales)
    return arrs, ptrs, P, S


def plan_eval(plan, cols, precisions, scales, rows, out_bytes, mode=FAITHFUL, threads=8):
    import numpy as np
    arrs, ptrs, P, S = _colargs(cols, precisions, scales)
    out = np.zeros(rows * out_bytes, dtype=np.uint8)
    lib.od_plan_eval(plan._h, ptrs, P, S, rows, out.ctypes.data, out_bytes, mode, threads)
    return out


def plan_sum(plan, cols, precisions, scales, rows, mask=None, mode=FAITHFUL, threads=8):
    import numpy as np
    arrs, ptrs, P, S = _colargs(cols, precisions, scales)
    out = ctypes.create_string_buffer(133)
    m = None
    if mask is not None:
        m = np.ascontiguousarray(mask, dtype=np.uint8)
    n = lib.od_plan_sum(plan._h, ptrs, P, S, rows, None if m is None else m.ctypes.data, out, mode, threads)
    return out.raw, n


def plan_group_sum(plan, cols, precisions, scales, rows, gid, ngroups, mode=FAITHFUL, threads=8):
    import numpy as np
    arrs, ptrs, P, S = _colargs(cols, precisions, scales)
    g = np.ascontiguousarray(gid, dtype=np.int32)
    out = np.zeros(ngroups * 133, dtype=np.uint8)
    cnt = np.zeros(ngroups, dtype=np.int64)
    lib.od_plan_group_sum(plan._h, ptrs, P, S, rows, g.ctypes.data, ngroups, out.ctypes.data, cnt.ctypes.data, mode,
                          threads)
    return out.reshape(ngroups, 133), cnt


def segsum_compact(col, P, S, rows, associated, groups, mode=FAITHFUL, threads=8):
    import numpy as np
    c = np.ascontiguousarray(col, dtype=np.uint8)
    a = None if associated is None else np.ascontiguousarray(associated, dtype=np.int32)
    g = np.ascontiguousarray(groups, dtype=np.int32)
    out = np.zeros(len(g) * 133, dtype=np.uint8)
    lib.od_segsum_compact(c.ctypes.data, P, S, rows, None if a is None else a.ctypes.data, g.ctypes.data, len(g),
                          out.ctypes.data, mode, threads)
    return out.reshape(len(g), 133)


def avg_from_sum(img133, count, mode=FAITHFUL):
    out = ctypes.create_string_buffer(133)
    lib.od_avg_from_sum(bytes(img133), count, out, mode)
    return out.raw


def image_value(img):
    """(sign, prec, frac, error, magnitude) of a 133-byte image."""
    img = bytes(img)
    m = int.from_bytes(img[5:133], "little")
    return img[1], img[2], img[3], img[4], m


def format_hex(d):
    """GetDecimal (decimal.cu:362-397) of an oracle value"""
    buf = ctypes.create_string_buffer(8 * 32 + 8)
    lib.od_format_hex(ctypes.byref(d), buf)
    return buf.value.decode()


def parse_compact(texts, P, S):
    """Decimal(P, S, text) -> ToCompactDecimal for each text (bytes or str):
    returns (rows * B(P) bytes as numpy uint8)"""
    import numpy as np
    enc = [t.encode("latin-1") if isinstance(t, str) else bytes(t) for t in texts]
    chars = b"".join(enc)
    offs = np.zeros(len(enc) + 1, dtype=np.int64)
    offs[1:] = np.cumsum([len(t) for t in enc])
    B = lib.od_compact_bytes(P)
    out = np.zeros(len(enc) * B, dtype=np.uint8)
    cbuf = ctypes.create_string_buffer(chars, len(chars) + 1)
    lib.od_parse_compact(cbuf, offs.ctypes.data, len(enc), P, S, out.ctypes.data)
    return out


def format_compact(col, rows, P, S, hex_form, stride):
    """stride-byte NUL-padded slots of each compact row, GetDecimal hex
    (hex_form true) or decimal text"""
    import numpy as np
    out = np.zeros(rows * stride, dtype=np.uint8)
    src = np.ascontiguousarray(col, dtype=np.uint8)
    lib.od_format_compact(src.ctypes.data, rows, P, S, 1 if hex_form else 0, out.ctypes.data, stride)
    return out
