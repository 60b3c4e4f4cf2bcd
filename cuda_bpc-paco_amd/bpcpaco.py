"""Python mirror of the BPC-PaCo C-ABI (libbpc_paco.so, declared in include/bpc_paco.h).

The reference's host interface is

    void CUDA_BPC(int Direction, DATATYPE* HData, int HDSize_X, int HDSize_Y, int* LUT)
    void load_LUTs(int** LUT, const char* folder)

(source/BPC_PaCo.cu:1305, source/main_example.cu:219).  ``run`` and
``Lut.from_text`` mirror them: same direction codes (CODE=0, DECODE=1), same
in-place buffer semantics, same LUT layout.  Errors raise ``BPCError`` with the
library's status code instead of exiting the process (common.cu:3-8).

Everything here calls the HIP library; there is no CPU fallback.  Loading fails
loudly when libbpc_paco.so has not been built.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libbpc_paco.so")

CODE = 0
DECODE = 1
LAYOUT_RASTER = 0
LAYOUT_TILED = 1

OK = 0
ERR_ARG = -1
ERR_HIP = -2
ERR_OVERFLOW = -3
ERR_RANGE = -4
ERR_MALFORMED = -5
ERR_WIDE = -6
ERR_NODEVICE = -7
ERR_TIMEOUT = -8


class BPCError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("BPC status %d: %s" % (code, msg))
        self.code = code


class Params(ctypes.Structure):
    """bpc_params: runtime form of CPASS_3, DWT_LEVELS, CBLOCK_LENGTH, LUT_N_BITPLANES (common.h)."""
    _fields_ = [
        ("cpass3", ctypes.c_int32),
        ("dwt_levels", ctypes.c_int32),
        ("cblock_length", ctypes.c_int32),
        ("lut_n_bitplanes", ctypes.c_int32),
        ("cb_stride_words", ctypes.c_int32),
        ("layout", ctypes.c_int32),
    ]

    def stride(self):
        return self.cb_stride_words if self.cb_stride_words > 0 else 64 * self.cblock_length

    def n_codeblocks(self, size_x, size_y):
        return (size_x // 64) * (size_y // self.cblock_length)


class Image(ctypes.Structure):
    """bpc_image: one image of a batch (device pointers)."""
    _fields_ = [
        ("src", ctypes.c_void_p),
        ("dst", ctypes.c_void_p),
        ("lut", ctypes.c_void_p),
        ("lut_len", ctypes.c_int32),
        ("ncw", ctypes.c_void_p),
        ("size_x", ctypes.c_int32),
        ("size_y", ctypes.c_int32),
        ("params", Params),
        ("fill_unused", ctypes.c_int32),
        ("int16_coeffs", ctypes.c_int32),
    ]


MEM_INFO_FN = ctypes.CFUNCTYPE(ctypes.c_int, ctypes.c_void_p, ctypes.POINTER(ctypes.c_size_t),
                               ctypes.POINTER(ctypes.c_size_t))


class ServerConfig(ctypes.Structure):
    """bpc_server_config."""
    _fields_ = [("max_inflight", ctypes.c_int32), ("max_batch_images", ctypes.c_int32),
                ("max_batch_codeblocks", ctypes.c_longlong), ("hbm_budget_bytes", ctypes.c_size_t),
                ("hbm_reserve_bytes", ctypes.c_size_t), ("pool_limit_bytes", ctypes.c_size_t),
                ("waves_per_block", ctypes.c_int32), ("vgpr_cap_waves", ctypes.c_int32),
                ("mem_info", MEM_INFO_FN), ("mem_info_ctx", ctypes.c_void_p)]


class Request(ctypes.Structure):
    _fields_ = [("direction", ctypes.c_int32), ("src", ctypes.c_void_p), ("lut", ctypes.c_void_p),
                ("lut_len", ctypes.c_int32), ("size_x", ctypes.c_int32), ("size_y", ctypes.c_int32),
                ("params", Params), ("int16_coeffs", ctypes.c_int32)]


class Result(ctypes.Structure):
    _fields_ = [("status", ctypes.c_int32), ("output", ctypes.c_void_p), ("ncw", ctypes.c_void_p),
                ("queue_us", ctypes.c_double), ("latency_us", ctypes.c_double)]


class ServerStats(ctypes.Structure):
    _fields_ = [("batches", ctypes.c_ulonglong), ("completed", ctypes.c_ulonglong), ("pending", ctypes.c_ulonglong),
                ("committed_bytes", ctypes.c_size_t), ("max_batch_images", ctypes.c_size_t)]


def params(cpass3=0, dwt_levels=5, cblock_length=64, lut_n_bitplanes=15, cb_stride_words=0, layout=LAYOUT_RASTER):
    return Params(cpass3, dwt_levels, cblock_length, lut_n_bitplanes, cb_stride_words, layout)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("libbpc_paco.so is not built (run __graft_entry__.build() or make -C cuda_bpc-paco_amd)")
        L = ctypes.CDLL(LIB_PATH)
        P = ctypes.c_void_p
        i32 = ctypes.c_int
        PP = ctypes.POINTER(Params)
        sz = ctypes.c_size_t
        sig = {
            "bpc_default_params": (None, [PP]),
            "bpc_lut_len": (i32, [PP]),
            "bpc_stream_words": (ctypes.c_int64, [i32, i32, PP]),
            "bpc_last_error": (ctypes.c_char_p, []),
            "bpc_device_count": (i32, []),
            "bpc_run": (i32, [i32, P, i32, i32, P]),
            "bpc_set_config": (i32, [PP]),
            "bpc_get_config": (i32, [PP]),
            "bpc_run_ex": (i32, [i32, P, i32, i32, P, i32, PP]),
            "bpc_lut_load_text": (i32, [ctypes.c_char_p, i32, i32, ctypes.POINTER(P), ctypes.POINTER(i32),
                                        ctypes.POINTER(i32)]),
            "bpc_lut_parse_text": (i32, [ctypes.c_char_p, i32, i32, P, i32, ctypes.POINTER(i32), ctypes.POINTER(i32)]),
            "bpc_lut_upload": (i32, [P, i32, ctypes.POINTER(P)]),
            "bpc_lut_free": (i32, [P]),
            "bpc_lut_train": (i32, [P, i32, i32, PP, P, i32, i32]),
            "bpc_lut_train_device": (i32, [P, i32, i32, PP, P, P, P, P]),
            "bpc_compact_index_device": (i32, [P, ctypes.c_int64, i32, P, P, P, P, P]),
            "bpc_compact_pack_device": (i32, [P, ctypes.c_int64, i32, P, P, ctypes.c_uint64, P, P]),
            "bpc_compact_unpack_device": (i32, [P, P, P, ctypes.c_uint64, ctypes.c_int64, i32, P, P, P]),
            "bpc_cb_file_order": (i32, [i32, i32, i32, i32, P, i32]),
            "bpc_dwt53_forward_device": (i32, [P, P, i32, i32, i32, P]),
            "bpc_dwt53_inverse_device": (i32, [P, P, i32, i32, i32, P]),
            "bpc_bin_export": (i32, [ctypes.c_char_p, P, P, P, i32, i32, PP]),
            "bpc_bin_import": (i32, [ctypes.c_char_p, i32, i32, PP, P]),
            "bpc_encode_device": (i32, [P, P, i32, i32, P, i32, PP, i32, P, P, P]),
            "bpc_decode_device": (i32, [P, P, i32, i32, P, i32, PP, P, P]),
            "bpc_encode_device_i16": (i32, [P, P, i32, i32, P, i32, PP, i32, P, P, P]),
            "bpc_decode_device_i16": (i32, [P, P, i32, i32, P, i32, PP, P, P]),
            "bpc_decode_compact_device": (i32, [P, P, P, ctypes.c_uint64, P, i32, i32, P, i32, PP, P, P]),
            "bpc_status_take": (i32, [P, P]),
            "bpc_codec_device_noalloc": (i32, [i32, P, P, i32, i32, P, i32, PP, i32, P, P, P]),
            "bpc_dev_malloc": (i32, [ctypes.POINTER(P), sz]),
            "bpc_dev_free": (i32, [P]),
            "bpc_dev_memset": (i32, [P, i32, sz, P]),
            "bpc_copy_h2d": (i32, [P, P, sz, P]),
            "bpc_copy_d2h": (i32, [P, P, sz, P]),
            "bpc_copy_d2d": (i32, [P, P, sz, P]),
            "bpc_stream_create": (i32, [ctypes.POINTER(P)]),
            "bpc_stream_destroy": (i32, [P]),
            "bpc_stream_sync": (i32, [P]),
            "bpc_device_sync": (i32, []),
            "bpc_event_create": (i32, [ctypes.POINTER(P)]),
            "bpc_event_destroy": (i32, [P]),
            "bpc_event_record": (i32, [P, P]),
            "bpc_event_elapsed_ms": (i32, [P, P, ctypes.POINTER(ctypes.c_float)]),
            "bpc_mem_info": (i32, [ctypes.POINTER(sz), ctypes.POINTER(sz)]),
            "bpc_server_default_config": (None, [ctypes.POINTER(ServerConfig)]),
            "bpc_server_create": (i32, [ctypes.POINTER(P), ctypes.POINTER(ServerConfig)]),
            "bpc_server_submit": (i32, [P, ctypes.POINTER(Request), ctypes.POINTER(ctypes.c_uint64)]),
            "bpc_server_wait": (i32, [P, ctypes.c_uint64, i32, ctypes.POINTER(Result)]),
            "bpc_server_release": (i32, [P, ctypes.c_uint64]),
            "bpc_server_stats": (i32, [P, ctypes.POINTER(ServerStats)]),
            "bpc_server_destroy": (i32, [P]),
            "bpc_batch_create": (i32, [ctypes.POINTER(P), i32, ctypes.c_longlong]),
            "bpc_batch_destroy": (i32, [P]),
            "bpc_batch_set_tuning": (i32, [P, i32, i32]),
            "bpc_batch_set_vgpr_cap": (i32, [P, i32]),
            "bpc_batch_run": (i32, [P, i32, ctypes.POINTER(Image), i32, i32, P, P]),
            "bpc_batch_graph_count": (i32, [P]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _check(code):
    if code != OK:
        raise BPCError(code, lib().bpc_last_error().decode(errors="replace"))
    return code


def _ptr(a):
    return ctypes.c_void_p(a.ctypes.data)


def lut_len(p):
    return lib().bpc_lut_len(ctypes.byref(p))


def stream_words(size_x, size_y, p):
    return lib().bpc_stream_words(size_x, size_y, ctypes.byref(p))


def device_count():
    return lib().bpc_device_count()


# ---- device memory -----------------------------------------------------------------

class DeviceBuffer:
    """A hipMalloc'ed buffer of n elements of dtype."""

    def __init__(self, n, dtype=np.int32):
        self.dtype = np.dtype(dtype)
        self.n = int(n)
        self.nbytes = self.n * self.dtype.itemsize
        p = ctypes.c_void_p()
        _check(lib().bpc_dev_malloc(ctypes.byref(p), max(1, self.nbytes)))
        self.ptr = p

    @classmethod
    def from_array(cls, arr, stream=None):
        arr = np.ascontiguousarray(arr)
        b = cls(arr.size, arr.dtype)
        b.upload(arr, stream)
        return b

    def upload(self, arr, stream=None):
        arr = np.ascontiguousarray(arr, dtype=self.dtype)
        assert arr.size <= self.n
        _check(lib().bpc_copy_h2d(self.ptr, _ptr(arr), arr.nbytes, _sp(stream)))

    def download(self, n=None, stream=None):
        n = self.n if n is None else n
        out = np.empty(n, dtype=self.dtype)
        _check(lib().bpc_copy_d2h(_ptr(out), self.ptr, out.nbytes, _sp(stream)))
        return out

    def copy_from(self, other, nbytes=None, stream=None):
        """Asynchronous device-to-device copy from another DeviceBuffer."""
        nb = other.nbytes if nbytes is None else nbytes
        _check(lib().bpc_copy_d2d(self.ptr, other.ptr, nb, _sp(stream)))

    def fill_bytes(self, value, stream=None):
        _check(lib().bpc_dev_memset(self.ptr, value, self.nbytes, _sp(stream)))

    def free(self):
        if self.ptr is not None and self.ptr.value:
            lib().bpc_dev_free(self.ptr)
        self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Stream:
    def __init__(self):
        p = ctypes.c_void_p()
        _check(lib().bpc_stream_create(ctypes.byref(p)))
        self.ptr = p

    def sync(self):
        _check(lib().bpc_stream_sync(self.ptr))

    def destroy(self):
        if self.ptr is not None:
            lib().bpc_stream_destroy(self.ptr)
            self.ptr = None


class Event:
    def __init__(self):
        p = ctypes.c_void_p()
        _check(lib().bpc_event_create(ctypes.byref(p)))
        self.ptr = p

    def record(self, stream=None):
        _check(lib().bpc_event_record(self.ptr, _sp(stream)))

    def elapsed_ms(self, later):
        ms = ctypes.c_float()
        _check(lib().bpc_event_elapsed_ms(self.ptr, later.ptr, ctypes.byref(ms)))
        return ms.value

    def destroy(self):
        if self.ptr is not None:
            lib().bpc_event_destroy(self.ptr)
            self.ptr = None


def _sp(stream):
    if stream is None:
        return None
    return stream.ptr if isinstance(stream, Stream) else stream


def device_sync():
    _check(lib().bpc_device_sync())


def mem_info():
    f, t = ctypes.c_size_t(), ctypes.c_size_t()
    _check(lib().bpc_mem_info(ctypes.byref(f), ctypes.byref(t)))
    return f.value, t.value


# ---- probability LUTs --------------------------------------------------------------

class Lut:
    """A device-resident probability LUT in the reference layout (main_example.cu:219-448)."""

    def __init__(self, dev_ptr, length, nbp=None, host=None):
        self.ptr = dev_ptr
        self.len = length
        self.nbp = nbp
        self.host = host

    @classmethod
    def from_text(cls, folder, cpass3, dwt_levels=5):
        """load_LUTs (main_example.cu:219): parse ref/sig/sign(/cp_sig/cp_sign).txt and upload."""
        p = ctypes.c_void_p()
        ln, nbp = ctypes.c_int(), ctypes.c_int()
        _check(lib().bpc_lut_load_text(folder.encode(), cpass3, dwt_levels, ctypes.byref(p), ctypes.byref(ln),
                                       ctypes.byref(nbp)))
        return cls(p, ln.value, nbp.value)

    @classmethod
    def from_array(cls, arr):
        arr = np.ascontiguousarray(arr, dtype=np.int32)
        p = ctypes.c_void_p()
        _check(lib().bpc_lut_upload(_ptr(arr), arr.size, ctypes.byref(p)))
        return cls(p, arr.size, host=arr)

    @classmethod
    def train_device(cls, d_img, size_x, size_y, p, stream=None):
        """LUT trained on the device from a device-resident training image (p.layout)."""
        ptr = ctypes.c_void_p()
        n = lut_len(p)
        _check(lib().bpc_dev_malloc(ctypes.byref(ptr), n * 4))
        try:
            train_lut_device(d_img, size_x, size_y, p, d_lut=ptr, stream=stream)
        except BPCError:
            lib().bpc_lut_free(ptr)
            raise
        return cls(ptr, n, p.lut_n_bitplanes)

    def free(self):
        if self.ptr is not None:
            lib().bpc_lut_free(self.ptr)
            self.ptr = None


def parse_lut_text(folder, cpass3, dwt_levels=5):
    """Host-side parse of the LUT text files; returns (int32 array, nbp)."""
    cap = 1 << 16
    out = np.zeros(cap, dtype=np.int32)
    ln, nbp = ctypes.c_int(), ctypes.c_int()
    _check(lib().bpc_lut_parse_text(folder.encode(), cpass3, dwt_levels, _ptr(out), cap, ctypes.byref(ln),
                                    ctypes.byref(nbp)))
    return out[: ln.value].copy(), nbp.value


def train_lut(img, p, nthreads=0):
    """Probability LUT from symbol frequencies of a raster training image."""
    img = np.ascontiguousarray(img, dtype=np.int32)
    Y, X = img.shape
    n = lut_len(p)
    out = np.zeros(n, dtype=np.int32)
    _check(lib().bpc_lut_train(_ptr(img), X, Y, ctypes.byref(p), _ptr(out), n, nthreads))
    return out


def _dptr(b):
    if b is None:
        return None
    return b.ptr if hasattr(b, "ptr") else b


def train_lut_device(d_img, size_x, size_y, p, d_lut=None, d_n0=None, d_n=None, stream=None):
    """Device trainer (asynchronous): LUT and/or per-entry counts from a device image."""
    _check(lib().bpc_lut_train_device(_dptr(d_img), size_x, size_y, ctypes.byref(p), _dptr(d_lut), _dptr(d_n0),
                                      _dptr(d_n), _sp(stream)))


def train_lut_gpu(img, p, counts=False):
    """Device trainer on a host image (2-D raster, or flattened codeblock-contiguous data
    with p.layout tiled and shape (size_y, size_x)).  Returns the LUT, or (lut, n0, n)."""
    img = np.ascontiguousarray(img, dtype=np.int32)
    Y, X = img.shape
    n = lut_len(p)
    d_img = DeviceBuffer.from_array(img)
    d_lut = DeviceBuffer(n)
    d_n0 = DeviceBuffer(n, np.uint32) if counts else None
    d_n = DeviceBuffer(n, np.uint32) if counts else None
    try:
        train_lut_device(d_img, X, Y, p, d_lut, d_n0, d_n)
        lut = d_lut.download()
        if counts:
            return lut, d_n0.download(), d_n.download()
        return lut
    finally:
        for b in (d_img, d_lut, d_n0, d_n):
            if b is not None:
                b.free()


# ---- codec -------------------------------------------------------------------------

def run(direction, hdata, size_x, size_y, lut, p=None):
    """CUDA_BPC (BPC_PaCo.cu:1305): hdata (host int32 numpy array) is overwritten in place.

    CODE: coefficients -> codestream regions; DECODE: regions -> coefficients.
    With p=None the library's current configuration (bpc_set_config) is used."""
    assert hdata.dtype == np.int32 and hdata.flags["C_CONTIGUOUS"]
    if p is None:
        _check(lib().bpc_run(direction, _ptr(hdata), size_x, size_y, lut.ptr))
    else:
        need = max(size_x * size_y, stream_words(size_x, size_y, p))
        assert hdata.size >= need, "hdata must hold %d int32" % need
        _check(lib().bpc_run_ex(direction, _ptr(hdata), size_x, size_y, lut.ptr, lut.len, ctypes.byref(p)))
    return hdata


def set_config(p):
    _check(lib().bpc_set_config(ctypes.byref(p)))


def get_config():
    p = Params()
    _check(lib().bpc_get_config(ctypes.byref(p)))
    return p


def encode_device(d_coeffs, d_stream, size_x, size_y, lut, p, fill_unused=1, d_ncw=None, d_status=None, stream=None):
    _check(lib().bpc_encode_device(d_coeffs.ptr, d_stream.ptr, size_x, size_y, lut.ptr, lut.len, ctypes.byref(p),
                                   fill_unused, d_ncw.ptr if d_ncw is not None else None,
                                   d_status.ptr if d_status is not None else None, _sp(stream)))


def decode_device(d_stream, d_coeffs, size_x, size_y, lut, p, d_status=None, stream=None):
    _check(lib().bpc_decode_device(d_stream.ptr, d_coeffs.ptr, size_x, size_y, lut.ptr, lut.len, ctypes.byref(p),
                                   d_status.ptr if d_status is not None else None, _sp(stream)))


def encode_device_i16(d_coeffs, d_stream, size_x, size_y, lut, p, fill_unused=1, d_ncw=None, d_status=None,
                      stream=None):
    """encode_device with int16 coefficients (d_coeffs holds size_x*size_y int16)."""
    _check(lib().bpc_encode_device_i16(_dptr(d_coeffs), _dptr(d_stream), size_x, size_y, lut.ptr, lut.len,
                                       ctypes.byref(p), fill_unused, _dptr(d_ncw), _dptr(d_status), _sp(stream)))


def decode_device_i16(d_stream, d_coeffs, size_x, size_y, lut, p, d_status=None, stream=None):
    """decode_device into int16 coefficients (saturated; out-of-range values are reported)."""
    _check(lib().bpc_decode_device_i16(_dptr(d_stream), _dptr(d_coeffs), size_x, size_y, lut.ptr, lut.len,
                                       ctypes.byref(p), _dptr(d_status), _sp(stream)))


def codec_device_noalloc(direction, d_src, d_dst, size_x, size_y, lut, p, d_status, fill_unused=1, d_ncw=None,
                         stream=None):
    """encode_device / decode_device without any device allocation (fixed per-device
    scratch; d_status required; calls overlapping on one GPU race)."""
    _check(lib().bpc_codec_device_noalloc(direction, _dptr(d_src), _dptr(d_dst), size_x, size_y, lut.ptr, lut.len,
                                          ctypes.byref(p), fill_unused, _dptr(d_ncw), _dptr(d_status), _sp(stream)))


def status_take(d_status=None, stream=None):
    """Synchronizes and returns the BPC status code of the launches since the last call (no raise)."""
    return lib().bpc_status_take(d_status.ptr if d_status is not None else None, _sp(stream))


def encode(img, lut, p, fill_unused=1):
    """Convenience: upload, encode on the device, download.  Returns (stream words, ncw per codeblock)."""
    Y, X = img.shape if p.layout == LAYOUT_RASTER else _tiled_shape(img, p)
    d_in = DeviceBuffer.from_array(np.ascontiguousarray(img, dtype=np.int32).ravel())
    nw = stream_words(X, Y, p)
    d_out = DeviceBuffer(nw)
    ncb = p.n_codeblocks(X, Y)
    d_ncw = DeviceBuffer(ncb)
    try:
        encode_device(d_in, d_out, X, Y, lut, p, fill_unused, d_ncw)
        _check(status_take())
        return d_out.download(), d_ncw.download()
    finally:
        d_in.free()
        d_out.free()
        d_ncw.free()


def decode(stream_words_arr, size_x, size_y, lut, p):
    d_in = DeviceBuffer.from_array(np.ascontiguousarray(stream_words_arr, dtype=np.int32))
    d_out = DeviceBuffer(size_x * size_y)
    try:
        decode_device(d_in, d_out, size_x, size_y, lut, p)
        _check(status_take())
        out = d_out.download()
    finally:
        d_in.free()
        d_out.free()
    return out.reshape(size_y, size_x) if p.layout == LAYOUT_RASTER else out


def _tiled_shape(img, p):
    raise ValueError("tiled input needs explicit sizes; use encode_device")


# ---- batched codec -------------------------------------------------------------------

class Batch:
    """A device work queue coding many images of uneven size in one persistent launch."""

    def __init__(self, max_images, max_codeblocks):
        p = ctypes.c_void_p()
        _check(lib().bpc_batch_create(ctypes.byref(p), max_images, max_codeblocks))
        self.ptr = p

    def set_tuning(self, waves_per_block=4, waves_per_simd=0):
        _check(lib().bpc_batch_set_tuning(self.ptr, waves_per_block, waves_per_simd))

    def set_vgpr_cap(self, waves_per_simd=0):
        """0: the compiler's register budget; 5..8: registers capped for that many waves per SIMD."""
        _check(lib().bpc_batch_set_vgpr_cap(self.ptr, waves_per_simd))

    def run(self, direction, images, use_graph=False, d_status=None, stream=None):
        """images: list of Image (or of dicts with the Image fields)."""
        arr = (Image * len(images))()
        for k, im in enumerate(images):
            arr[k] = im if isinstance(im, Image) else make_image(**im)
        _check(lib().bpc_batch_run(self.ptr, direction, arr, len(images), 1 if use_graph else 0,
                                   d_status.ptr if d_status is not None else None, _sp(stream)))

    def graph_count(self):
        n = lib().bpc_batch_graph_count(self.ptr)
        if n < 0:
            _check(n)
        return n

    def destroy(self):
        if self.ptr is not None:
            lib().bpc_batch_destroy(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass


def make_image(src, dst, lut, size_x, size_y, p, ncw=None, fill_unused=1, int16=False):
    """Image descriptor from DeviceBuffers and a Lut; int16: the coefficients (encoder
    src, decoder dst) are int16."""
    return Image(src.ptr, dst.ptr, lut.ptr, lut.len, ncw.ptr if ncw is not None else None, size_x, size_y, p,
                 fill_unused, 1 if int16 else 0)


# ---- continuous-batching server ------------------------------------------------------

class Server:
    """Request server: submit encode/decode requests on device buffers, wait for results.

    mem_info: optional callable returning (free_bytes, total_bytes), used by admission
    instead of the device's figures (to exercise admission under memory pressure)."""

    def __init__(self, max_inflight=2, max_batch_images=64, max_batch_codeblocks=1 << 18, hbm_budget_bytes=0,
                 hbm_reserve_bytes=None, pool_limit_bytes=None, waves_per_block=4, vgpr_cap_waves=8, mem_info=None):
        cfg = ServerConfig()
        lib().bpc_server_default_config(ctypes.byref(cfg))
        cfg.max_inflight, cfg.max_batch_images = max_inflight, max_batch_images
        cfg.max_batch_codeblocks, cfg.hbm_budget_bytes = max_batch_codeblocks, hbm_budget_bytes
        if hbm_reserve_bytes is not None:
            cfg.hbm_reserve_bytes = hbm_reserve_bytes
        if pool_limit_bytes is not None:
            cfg.pool_limit_bytes = pool_limit_bytes
        cfg.waves_per_block = waves_per_block
        cfg.vgpr_cap_waves = vgpr_cap_waves
        self._probe = None
        if mem_info is not None:
            def probe(ctx, fr, tot):
                try:
                    f, t = mem_info()
                    fr[0], tot[0] = int(f), int(t)
                    return 0
                except Exception:
                    return 1
            self._probe = MEM_INFO_FN(probe)  # kept alive as long as the server
            cfg.mem_info = self._probe
        p = ctypes.c_void_p()
        _check(lib().bpc_server_create(ctypes.byref(p), ctypes.byref(cfg)))
        self.ptr = p

    def submit(self, direction, src, lut, size_x, size_y, p, int16=False):
        req = Request(direction, src.ptr if hasattr(src, "ptr") else src, lut.ptr, lut.len, size_x, size_y, p,
                      1 if int16 else 0)
        rid = ctypes.c_uint64()
        _check(lib().bpc_server_submit(self.ptr, ctypes.byref(req), ctypes.byref(rid)))
        return rid.value

    def wait(self, rid, timeout_ms=-1):
        res = Result()
        _check(lib().bpc_server_wait(self.ptr, rid, timeout_ms, ctypes.byref(res)))
        return res

    def release(self, rid):
        _check(lib().bpc_server_release(self.ptr, rid))

    def stats(self):
        st = ServerStats()
        _check(lib().bpc_server_stats(self.ptr, ctypes.byref(st)))
        return st

    def destroy(self):
        if self.ptr is not None:
            lib().bpc_server_destroy(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass


def download_ptr(ptr, n, dtype=np.int32, stream=None):
    """Copies n elements from a raw device pointer to a new host array."""
    out = np.empty(n, dtype=dtype)
    _check(lib().bpc_copy_d2h(_ptr(out), ctypes.c_void_p(ptr), out.nbytes, _sp(stream)))
    return out


# ---- compact codestream container / per-codeblock files -------------------------------

class Compact:
    """A compact codestream on the device: msb (uint8[ncb]), offsets (uint64[ncb+1],
    in codewords) and payload (uint16[offsets[ncb]])."""

    def __init__(self, msb, offsets, payload, ncb, total):
        self.msb, self.offsets, self.payload = msb, offsets, payload
        self.ncb, self.total = ncb, total

    @property
    def nbytes(self):
        return self.ncb + 8 * (self.ncb + 1) + 2 * self.total

    def download(self, stream=None):
        return (self.msb.download(stream=stream), self.offsets.download(stream=stream),
                self.payload.download(self.total, stream=stream))

    def free(self):
        for b in (self.msb, self.offsets, self.payload):
            b.free()


def compact_index(d_stream, ncb, stride, d_msb, d_offsets, d_ncw=None, d_status=None, stream=None):
    _check(lib().bpc_compact_index_device(_dptr(d_stream), ncb, stride, _dptr(d_ncw), _dptr(d_msb), _dptr(d_offsets),
                                          _dptr(d_status), _sp(stream)))


def compact_pack_into(d_stream, ncb, stride, d_offsets, d_payload, cap, d_status=None, stream=None):
    _check(lib().bpc_compact_pack_device(_dptr(d_stream), ncb, stride, _dptr(d_offsets), _dptr(d_payload), cap,
                                         _dptr(d_status), _sp(stream)))


def compact_pack(d_stream, ncb, stride, d_ncw=None, d_status=None, stream=None):
    """Reference-layout regions on the device -> Compact (synchronizes once to size
    the payload)."""
    msb = DeviceBuffer(ncb, np.uint8)
    off = DeviceBuffer(ncb + 1, np.uint64)
    compact_index(d_stream, ncb, stride, msb, off, d_ncw, d_status, stream)
    total = int(download_ptr(off.ptr.value + 8 * ncb, 1, np.uint64, stream)[0])
    payload = DeviceBuffer(max(total, 1), np.uint16)
    compact_pack_into(d_stream, ncb, stride, off, payload, total, d_status, stream)
    return Compact(msb, off, payload, ncb, total)


def compact_unpack(c, d_stream, stride, d_status=None, stream=None):
    """Compact -> reference-layout regions (MSB, codewords, -1 fill) in d_stream."""
    _check(lib().bpc_compact_unpack_device(_dptr(c.msb), _dptr(c.offsets), _dptr(c.payload), c.total, c.ncb, stride,
                                           _dptr(d_stream), _dptr(d_status), _sp(stream)))


def decode_compact_device(c, d_coeffs, size_x, size_y, lut, p, d_status=None, stream=None):
    """Compact -> int32 coefficients in p.layout, with no regions in between."""
    _check(lib().bpc_decode_compact_device(_dptr(c.msb), _dptr(c.offsets), _dptr(c.payload), c.total,
                                           _dptr(d_coeffs), size_x, size_y, lut.ptr, lut.len, ctypes.byref(p),
                                           _dptr(d_status), _sp(stream)))


def cb_file_order(size_x, size_y, dwt_levels=5, cblock_length=64):
    """Tile index of every codeblock file k (main_example.cu:71-179's order)."""
    n = (size_x // 64) * (size_y // cblock_length)
    out = np.zeros(max(n, 1), dtype=np.int32)
    k = lib().bpc_cb_file_order(size_x, size_y, dwt_levels, cblock_length, _ptr(out), out.size)
    if k < 0:
        raise BPCError(k, lib().bpc_last_error().decode(errors="replace"))
    return out[:k].copy()


def bin_export(folder, msb, offsets, payload, size_x, size_y, p):
    msb = np.ascontiguousarray(msb, dtype=np.uint8)
    offsets = np.ascontiguousarray(offsets, dtype=np.uint64)
    payload = np.ascontiguousarray(payload, dtype=np.uint16)
    if payload.size == 0:
        payload = np.zeros(1, np.uint16)
    _check(lib().bpc_bin_export(folder.encode(), _ptr(msb), _ptr(offsets), _ptr(payload), size_x, size_y,
                                ctypes.byref(p)))


def bin_import(folder, size_x, size_y, p):
    out = np.zeros(stream_words(size_x, size_y, p), dtype=np.int32)
    _check(lib().bpc_bin_import(folder.encode(), size_x, size_y, ctypes.byref(p), _ptr(out)))
    return out


# ---- reversible 5/3 wavelet transform ----------------------------------------------

def dwt53_forward_device(d_img, d_coeffs, size_x, size_y, levels, stream=None):
    _check(lib().bpc_dwt53_forward_device(_dptr(d_img), _dptr(d_coeffs), size_x, size_y, levels, _sp(stream)))


def dwt53_inverse_device(d_coeffs, d_img, size_x, size_y, levels, stream=None):
    _check(lib().bpc_dwt53_inverse_device(_dptr(d_coeffs), _dptr(d_img), size_x, size_y, levels, _sp(stream)))


def dwt53(img, levels, inverse=False):
    """Host convenience: one transform of a (Y, X) int32 image on the device."""
    img = np.ascontiguousarray(img, dtype=np.int32)
    Y, X = img.shape
    a = DeviceBuffer.from_array(img)
    b = DeviceBuffer(X * Y)
    try:
        (dwt53_inverse_device if inverse else dwt53_forward_device)(a, b, X, Y, levels)
        return b.download().reshape(Y, X)
    finally:
        a.free()
        b.free()
