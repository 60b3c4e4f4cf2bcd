"""CPU-side checks of the C-ABI library: it loads, exports every declared symbol,
validates arguments, and its host-only parts (LUT text parser, LUT trainer) agree
with the oracle.  No GPU compute here."""
import ctypes
import os
import re

import numpy as np
import pytest

import synth
from conftest import ROOT, lut_folder


def declared_functions():
    text = open(os.path.join(ROOT, "include", "bpc_paco.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(bpc_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol(bpc):
    names = declared_functions()
    assert len(names) >= 30
    for n in names:
        assert hasattr(bpc.lib(), n), n


def test_default_params_are_reference_defaults(bpc):
    p = bpc.Params()
    bpc.lib().bpc_default_params(ctypes.byref(p))
    assert (p.cpass3, p.dwt_levels, p.cblock_length, p.lut_n_bitplanes, p.cb_stride_words, p.layout) == \
        (0, 5, 64, 15, 0, 0)
    assert bpc.lut_len(p) == 3360
    p.cpass3 = 1
    assert bpc.lut_len(p) == 6480
    assert bpc.stream_words(2048, 2048, p) == 2048 * 2048
    p.cb_stride_words = 8192
    assert bpc.stream_words(2048, 2048, p) == 1024 * 8192


@pytest.mark.parametrize("cpass3", [0, 1])
def test_lut_parser_matches_oracle_loader(bpc, O, cpass3):
    a, nbp_a = bpc.parse_lut_text(lut_folder(cpass3), cpass3)
    b, nbp_b = O.lut_load_text(lut_folder(cpass3), cpass3)
    assert nbp_a == nbp_b == 15
    assert np.array_equal(a, b)


def test_invalid_arguments_are_rejected(bpc):
    lib = bpc.lib()
    p = bpc.params()
    buf = np.zeros(64 * 64, np.int32)
    bad = bpc.params(cblock_length=1)
    assert lib.bpc_set_config(ctypes.byref(bad)) == bpc.ERR_ARG
    assert b"cblock_length" in lib.bpc_last_error()
    # codeblock regions: 2 .. 2^26 words (the per-image kernel addresses a pair's regions with 32-bit offsets)
    for stride in (1, (1 << 26) + 1):
        bad = bpc.params(cb_stride_words=stride)
        assert lib.bpc_set_config(ctypes.byref(bad)) == bpc.ERR_ARG
        assert b"cb_stride_words" in lib.bpc_last_error()
    ok = bpc.params(cb_stride_words=1 << 26)
    assert lib.bpc_set_config(ctypes.byref(ok)) == bpc.OK
    assert lib.bpc_set_config(ctypes.byref(p)) == bpc.OK
    # image not a whole number of codeblocks
    assert lib.bpc_run_ex(0, buf.ctypes.data, 100, 64, ctypes.c_void_p(8), 3360, ctypes.byref(p)) == bpc.ERR_ARG
    # LUT too short
    assert lib.bpc_run_ex(0, buf.ctypes.data, 64, 64, ctypes.c_void_p(8), 100, ctypes.byref(p)) == bpc.ERR_ARG
    # bad direction
    assert lib.bpc_run_ex(5, buf.ctypes.data, 64, 64, ctypes.c_void_p(8), 3360, ctypes.byref(p)) == bpc.ERR_ARG
    # LUT entries must fit the coder's 7-bit probabilities
    lut = np.full(3360, 64, np.int32)
    lut[7] = 128
    out = ctypes.c_void_p()
    assert lib.bpc_lut_upload(lut.ctypes.data, lut.size, ctypes.byref(out)) == bpc.ERR_RANGE


def test_trainer_counts_symbol_frequencies(bpc, O):
    """The trained table codes the image it was trained on better than a flat table,
    stays in [1,127], and round-trips through the oracle."""
    img = synth.synth_image(512, 512, levels=3, seed=11)
    for cp3 in (0, 1):
        p = bpc.params(cpass3=cp3, dwt_levels=3, lut_n_bitplanes=15)
        lut = bpc.train_lut(img, p)
        assert lut.size == bpc.lut_len(p)
        assert lut.min() >= 1 and lut.max() <= 127
        assert np.array_equal(lut, bpc.train_lut(img, p, nthreads=1))  # thread count does not matter
        op = O.params(cpass3=cp3, dwt_levels=3, lut_n_bitplanes=15)
        st, stream, ncw = O.encode(img, lut, op)
        assert st == 0
        st, rec = O.decode(stream, 512, 512, lut, op)
        assert st == 0 and np.array_equal(rec, img)
        st, _, ncw_flat = O.encode(img, np.full(lut.size, 64, np.int32), op)
        assert ncw.sum() < 0.9 * ncw_flat.sum()


def test_trainer_single_context_frequency(bpc):
    """A codeblock whose only nonzero coefficient is 1 at the top-left corner, 2-pass,
    MSB 0: plane 0 codes 4096 significance symbols, all in context 0 except the three
    neighbours coded after the corner became significant... counted by hand below."""
    img = np.zeros((64, 64), np.int32)
    img[0, 0] = 1
    p = bpc.params(cpass3=0, dwt_levels=0, lut_n_bitplanes=1)
    lut = bpc.train_lut(img, p)
    # group 0 (LL with 0 levels), plane 0.  Scan: row 0 even columns, then row 0 odd
    # columns, then row 1 ...  (0,0) is coded first in context 0 (symbol 1).  Its
    # neighbours (0,1), (1,0), (1,1) are coded later with context 1 (symbol 0).
    # All other 4092 coefficients: context 0, symbol 0.
    sig = lut[1:1 + 9]  # ref region is 1 entry (1 group x 1 plane)
    assert sig[0] == 127  # 4092 zeros of 4093 -> round(127.97) = 128 -> clamped to 127
    assert sig[1] == 127  # 3 of 3 zeros -> 128 -> clamp
    assert all(v == 64 for v in sig[2:])  # never visited
    sign = lut[10:14]
    # the sign of (0,0): no significant neighbours -> context 0, predicted +, positive
    # -> symbol 0 once -> 128 -> 127
    assert sign[0] == 127 and all(v == 64 for v in sign[1:])
