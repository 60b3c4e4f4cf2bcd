"""ctypes binding of the CPU oracle (oracle/_build/libbpc_oracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, where it is the checker.  The product path
(cuda_bpc-paco_amd/) never imports it.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "libbpc_oracle.so")
_lib = None

OK = 0
ERR_ARG = -1
ERR_OVERFLOW = -3
ERR_RANGE = -4
ERR_MALFORMED = -5


class Params(ctypes.Structure):
    _fields_ = [
        ("cpass3", ctypes.c_int32),
        ("dwt_levels", ctypes.c_int32),
        ("cblock_length", ctypes.c_int32),
        ("lut_n_bitplanes", ctypes.c_int32),
        ("cb_stride_words", ctypes.c_int32),
    ]


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


NATIVE_FLAGS = "-O3 -march=native -fopenmp"


def use_native():
    """Builds the oracle for the host it runs on (make native: -O3 -march=native
    -fopenmp) and loads that build instead of the portable one; for the CPU
    baseline.  Must be called before the first use.  Returns the flags used."""
    global _LIB_PATH
    if _lib is not None:
        raise RuntimeError("oracle already loaded")
    subprocess.run(["make", "-s", "-C", _HERE, "native"], check=True)
    _LIB_PATH = os.path.join(_HERE, "_build", "native", "libbpc_oracle.so")
    return NATIVE_FLAGS


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            raise RuntimeError("oracle library not built: run `make -C oracle`")
        L = ctypes.CDLL(_LIB_PATH)
        P = ctypes.c_void_p
        i32 = ctypes.c_int
        PP = ctypes.POINTER(Params)
        L.orc_lut_len.argtypes = [PP]
        L.orc_encode.argtypes = [P, i32, i32, P, i32, PP, P, P, i32]
        L.orc_decode.argtypes = [P, i32, i32, P, i32, PP, P, i32]
        L.orc_encode_list.argtypes = [P, i32, i32, P, i32, P, i32, PP, P, P, i32]
        L.orc_decode_list.argtypes = [P, i32, i32, P, i32, P, i32, PP, P, i32]
        L.orc_encode_cb.argtypes = [P, i32, i32, i32, P, i32, PP, P, P]
        L.orc_decode_cb.argtypes = [P, i32, i32, i32, P, i32, PP, P]
        L.orc_symbol_counts.argtypes = [P, i32, i32, PP, P, P, i32, i32]
        L.orc_cb_file_order.argtypes = [i32, i32, i32, i32, P, i32]
        L.orc_load_cstream.argtypes = [P, ctypes.c_size_t, P, i32]
        L.orc_sig_context.argtypes = [P]
        L.orc_sign_context_hv.argtypes = [i32, i32]
        L.orc_sign_context.argtypes = [ctypes.c_uint32] * 4
        L.orc_find_subband.argtypes = [i32, i32, i32, i32, i32, ctypes.POINTER(i32), ctypes.POINTER(i32)]
        L.orc_codeblock_msb.argtypes = [P, i32, i32, i32, i32]
        L.orc_ac_encode_seq.argtypes = [P, P, i32, P, i32]
        L.orc_ac_decode_seq.argtypes = [P, i32, P, i32, P]
        L.orc_lut_load_text.argtypes = [ctypes.c_char_p, i32, i32, P, i32, ctypes.POINTER(i32), ctypes.POINTER(i32)]
        L.orc_encode_cb_trace.argtypes = [P, i32, i32, i32, P, i32, PP, P, P, P, i32, ctypes.POINTER(i32)]
        L.orc_decode_cb_trace.argtypes = [P, i32, i32, i32, P, i32, PP, P, P, i32, ctypes.POINTER(i32)]
        _lib = L
    return _lib


def params(cpass3=0, dwt_levels=5, cblock_length=64, lut_n_bitplanes=15, cb_stride_words=0):
    return Params(cpass3, dwt_levels, cblock_length, lut_n_bitplanes, cb_stride_words)


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def lut_len(p):
    return lib().orc_lut_len(ctypes.byref(p))


def stride(p):
    return p.cb_stride_words if p.cb_stride_words > 0 else 64 * p.cblock_length


def n_codeblocks(size_x, size_y, p):
    return (size_x // 64) * (size_y // p.cblock_length)


def encode(img, lut, p, nthreads=0):
    """Encode a (Y, X) int32 image.  Returns (status, stream[nCB*stride], ncw[nCB])."""
    img = np.ascontiguousarray(img, dtype=np.int32)
    lut = np.ascontiguousarray(lut, dtype=np.int32)
    Y, X = img.shape
    n = n_codeblocks(X, Y, p)
    stream = np.empty(n * stride(p), dtype=np.int32)
    ncw = np.zeros(n, dtype=np.int32)
    st = lib().orc_encode(_ptr(img), X, Y, _ptr(lut), lut.size, ctypes.byref(p), _ptr(stream), _ptr(ncw), nthreads)
    return st, stream, ncw


def decode(stream, size_x, size_y, lut, p, nthreads=0):
    stream = np.ascontiguousarray(stream, dtype=np.int32)
    lut = np.ascontiguousarray(lut, dtype=np.int32)
    img = np.zeros((size_y, size_x), dtype=np.int32)
    st = lib().orc_decode(_ptr(stream), size_x, size_y, _ptr(lut), lut.size, ctypes.byref(p), _ptr(img), nthreads)
    return st, img


# orc_trace_rec (bpc_oracle.h): one record per coded symbol
TRACE_DTYPE = np.dtype([("plane", "<i2"), ("pass", "i1"), ("site", "i1"), ("row", "<i2"), ("column", "i1"),
                        ("lane", "i1"), ("lut_index", "<i4"), ("prob", "<i4"), ("L", "<u4"), ("S", "<u4"),
                        ("slot", "<u4"), ("sym", "<u4"), ("opened", "<u4")])
PASS_NAMES = ("SPP", "MRP", "CP")
SITE_NAMES = ("significance", "sign", "refinement")


def _trace_call(fn, n_hint):
    cap = n_hint
    while True:
        rec = np.zeros(max(cap, 1), dtype=TRACE_DTYPE)
        n = ctypes.c_int()
        st = fn(rec, cap, n)
        if n.value <= cap:
            return st, rec[: n.value]
        cap = n.value


def encode_cb_trace(img, cb, lut, p):
    """Encode codeblock cb of a (Y, X) image and return (status, region, ncw, trace):
    every coded symbol of every lane as a TRACE_DTYPE record, in coding order."""
    img = np.ascontiguousarray(img, dtype=np.int32)
    lut = np.ascontiguousarray(lut, dtype=np.int32)
    Y, X = img.shape
    region = np.full(stride(p), -1, dtype=np.int32)
    ncw = np.zeros(1, dtype=np.int32)
    st, rec = _trace_call(lambda rec, cap, n: lib().orc_encode_cb_trace(
        _ptr(img), X, Y, cb, _ptr(lut), lut.size, ctypes.byref(p), _ptr(region), _ptr(ncw), _ptr(rec), cap,
        ctypes.byref(n)), 4096)
    return st, region, int(ncw[0]), rec


def decode_cb_trace(region, size_x, size_y, cb, lut, p):
    """Decode one codeblock region; returns (status, image with that codeblock, trace)."""
    region = np.ascontiguousarray(region, dtype=np.int32)
    lut = np.ascontiguousarray(lut, dtype=np.int32)
    img = np.zeros((size_y, size_x), dtype=np.int32)
    st, rec = _trace_call(lambda rec, cap, n: lib().orc_decode_cb_trace(
        _ptr(region), size_x, size_y, cb, _ptr(lut), lut.size, ctypes.byref(p), _ptr(img), _ptr(rec), cap,
        ctypes.byref(n)), 4096)
    return st, img, rec


def format_trace(rec, lane=None, limit=None):
    """Text lines of a trace (optionally one lane), one per symbol."""
    out = []
    for r in rec:
        if lane is not None and r["lane"] != lane:
            continue
        out.append("bp=%d %s row=%d col=%d lane=%d %s idx=%d P=%d L=%d S=%d slot=%d sym=%d%s" % (
            r["plane"], PASS_NAMES[r["pass"]], r["row"], r["column"], r["lane"], SITE_NAMES[r["site"]],
            r["lut_index"], r["prob"], r["L"], r["S"], r["slot"], r["sym"], " open" if r["opened"] else ""))
        if limit is not None and len(out) >= limit:
            break
    return out


def encode_list(img, cbs, lut, p, nthreads=0):
    """Encode only the codeblocks in cbs; untouched regions stay -1."""
    img = np.ascontiguousarray(img, dtype=np.int32)
    lut = np.ascontiguousarray(lut, dtype=np.int32)
    cbs = np.ascontiguousarray(cbs, dtype=np.int32)
    Y, X = img.shape
    n = n_codeblocks(X, Y, p)
    stream = np.full(n * stride(p), -1, dtype=np.int32)
    ncw = np.zeros(n, dtype=np.int32)
    st = lib().orc_encode_list(_ptr(img), X, Y, _ptr(cbs), cbs.size, _ptr(lut), lut.size, ctypes.byref(p),
                               _ptr(stream), _ptr(ncw), nthreads)
    return st, stream, ncw


def decode_list(stream, size_x, size_y, cbs, lut, p, nthreads=0):
    stream = np.ascontiguousarray(stream, dtype=np.int32)
    lut = np.ascontiguousarray(lut, dtype=np.int32)
    cbs = np.ascontiguousarray(cbs, dtype=np.int32)
    img = np.zeros((size_y, size_x), dtype=np.int32)
    st = lib().orc_decode_list(_ptr(stream), size_x, size_y, _ptr(cbs), cbs.size, _ptr(lut), lut.size,
                               ctypes.byref(p), _ptr(img), nthreads)
    return st, img


def symbol_counts(img, p, nthreads=0):
    """Per LUT entry: (status, n0, n) = symbols 0 and all symbols the encoder codes
    with that entry's probability (planes >= lut_n_bitplanes not counted)."""
    img = np.ascontiguousarray(img, dtype=np.int32)
    Y, X = img.shape
    n = lut_len(p)
    n0 = np.zeros(n, dtype=np.uint32)
    nn = np.zeros(n, dtype=np.uint32)
    st = lib().orc_symbol_counts(_ptr(img), X, Y, ctypes.byref(p), _ptr(n0), _ptr(nn), n, nthreads)
    return st, n0, nn


def cb_file_order(size_x, size_y, dwt_levels, cblock_length=64):
    """Tile index of each codeblock file k (load_Data's order); None if not a bijection."""
    n = (size_x // 64) * (size_y // cblock_length)
    out = np.zeros(n + 1, dtype=np.int32)
    k = lib().orc_cb_file_order(size_x, size_y, dwt_levels, cblock_length, _ptr(out), out.size)
    return None if k < 0 else out[:k].copy()


def load_cstream(data, words):
    """load_CStream of one file's bytes into a region of `words` int32."""
    buf = np.frombuffer(bytes(data), dtype=np.uint8) if len(data) else np.zeros(1, np.uint8)
    out = np.zeros(words, dtype=np.int32)
    lib().orc_load_cstream(_ptr(buf), len(data), _ptr(out), words)
    return out


def sig_context(nb):
    a = np.ascontiguousarray(nb, dtype=np.uint32)
    return lib().orc_sig_context(_ptr(a))


def sign_context_hv(h, v):
    return lib().orc_sign_context_hv(h, v)


def sign_context(u, l, r, d):
    return lib().orc_sign_context(u, l, r, d)


def find_subband(x, y, size_x, size_y, levels):
    lv, sb = ctypes.c_int(), ctypes.c_int()
    lib().orc_find_subband(x, y, size_x, size_y, levels, ctypes.byref(lv), ctypes.byref(sb))
    return lv.value, sb.value


def codeblock_msb(img, cb, cblock_length=64):
    img = np.ascontiguousarray(img, dtype=np.int32)
    Y, X = img.shape
    return lib().orc_codeblock_msb(_ptr(img), X, Y, cb, cblock_length)


def ac_encode_seq(syms, probs):
    syms = np.ascontiguousarray(syms, dtype=np.uint8)
    probs = np.ascontiguousarray(probs, dtype=np.int32)
    cw = np.zeros(max(1, syms.size + 1), dtype=np.uint32)
    n = lib().orc_ac_encode_seq(_ptr(syms), _ptr(probs), syms.size, _ptr(cw), cw.size)
    return cw[:n].copy()


def ac_decode_seq(cw, probs):
    cw = np.ascontiguousarray(cw, dtype=np.uint32)
    probs = np.ascontiguousarray(probs, dtype=np.int32)
    out = np.zeros(probs.size, dtype=np.uint8)
    lib().orc_ac_decode_seq(_ptr(cw), cw.size, _ptr(probs), probs.size, _ptr(out))
    return out


def lut_load_text(folder, cpass3, dwt_levels=5):
    """Returns (lut int32 array, nbp)."""
    if not folder.endswith("/"):
        folder += "/"
    cap = 1 << 16
    out = np.zeros(cap, dtype=np.int32)
    nbp, ln = ctypes.c_int(), ctypes.c_int()
    st = lib().orc_lut_load_text(folder.encode(), cpass3, dwt_levels, _ptr(out), cap, ctypes.byref(nbp), ctypes.byref(ln))
    if st:
        raise RuntimeError("orc_lut_load_text failed: %d" % st)
    return out[: ln.value].copy(), nbp.value
