"""Pins the oracle's primitives against closed forms, exhaustive tables and hand traces.

Each check is an independent characterisation (arithmetic worked by hand, a
mathematical property, or the Mallat geometry), not a retyping of the oracle's code.
"""
import itertools

import numpy as np
import pytest

import synth

S = 0x80000000


# ---- arithmetic coder (BPC_PaCo.cu:261-339) -----------------------------------------

def test_ac_halving_interval(O):
    # P = 64 halves the interval: 65535 -> 32767 -> ... -> 1 -> 0 is 16 symbols, and
    # all-zero symbols never raise the lower bound, so every codeword is 0.
    cw = O.ac_encode_seq(np.zeros(32, np.uint8), np.full(32, 64))
    assert cw.tolist() == [0, 0]
    cw = O.ac_encode_seq(np.zeros(33, np.uint8), np.full(33, 64))
    assert cw.tolist() == [0, 0, 0]  # third codeword opened by the 33rd symbol, flushed


def test_ac_zero_probability(O):
    # P = 0: t = 0.  A 1 moves L up by one and S down by one; a 0 sets S = 0 and
    # emits L.  [1,1,1,0] -> codeword 3, [1,0] -> codeword 1.
    cw = O.ac_encode_seq(np.array([1, 1, 1, 0, 1, 0], np.uint8), np.zeros(6))
    assert cw.tolist() == [3, 1]


def test_ac_hand_trace_p127(O):
    # S=65535, P=127: t = 8322945 >> 7 = 65023; symbol 1 -> S = 65535-65024 = 511, L = 65024
    # S=511: t = 64897 >> 7 = 507; -> S = 3, L = 65532
    # S=3:   t = 381 >> 7 = 2;     -> S = 0, L = 65535 -> codeword 65535
    cw = O.ac_encode_seq(np.array([1, 1, 1], np.uint8), np.full(3, 127))
    assert cw.tolist() == [65535]


def test_ac_hand_trace_mixed(O):
    # P=100: t = 6553500 >> 7 = 51199 ; symbol 0 -> S = 51199
    #        t = 5119900 >> 7 = 39999 ; symbol 1 -> S = 11199, L = 40000 ; flush 40000
    cw = O.ac_encode_seq(np.array([0, 1], np.uint8), np.full(2, 100))
    assert cw.tolist() == [40000]


def test_ac_decoder_inverts_encoder(O):
    rng = np.random.default_rng(7)
    for trial in range(50):
        n = int(rng.integers(1, 3000))
        probs = rng.integers(0, 128, size=n)
        p1 = rng.random()
        syms = (rng.random(n) < p1).astype(np.uint8)
        cw = O.ac_encode_seq(syms, probs)
        assert cw.max(initial=0) <= 0xFFFF
        out = O.ac_decode_seq(cw, probs)
        assert np.array_equal(out, syms), trial


# ---- contexts -------------------------------------------------------------------------

def test_significance_context_is_popcount_of_S(O):
    rng = np.random.default_rng(1)
    for pattern in range(256):
        nb = np.array([(S if (pattern >> k) & 1 else 0) | int(rng.integers(0, 1 << 29)) for k in range(8)],
                      dtype=np.uint32)
        assert O.sig_context(nb) == bin(pattern).count("1")


def _word(state):
    # state: 0 insignificant, +1 significant positive, -1 significant negative
    if state == 0:
        return 1  # sign bit set but not significant: must read as 0
    return S | (1 if state < 0 else 0)


def test_sign_context_characterisation(O):
    """Exhaustive over the 3^4 (U, L, R, D) states.  Properties of the reference table
    (BPC_PaCo.cu:178-211): the predicted-sign bit is the sign of h, or of v when h = 0;
    the LUT index (ctx >> 1) is 0 for (0,0) and opposite signs, 1 for h = 0, 2 for
    v = 0, 3 for equal signs."""
    for u, l, r, d in itertools.product((-1, 0, 1), repeat=4):
        h, v = l + r, u + d
        ctx = O.sign_context(_word(u), _word(l), _word(r), _word(d))
        pred_negative = h < 0 or (h == 0 and v < 0)
        assert (ctx & 1) == int(pred_negative), (u, l, r, d)
        if (h == 0 and v == 0) or h * v < 0:
            idx = 0
        elif h == 0:
            idx = 1
        elif v == 0:
            idx = 2
        else:
            idx = 3
        assert ctx >> 1 == idx, (u, l, r, d)
        assert O.sign_context_hv(h, v) == ctx
    # sign-flip symmetry: negating every neighbour flips only the predicted-sign bit
    for h, v in itertools.product(range(-2, 3), repeat=2):
        if (h, v) != (0, 0):
            assert O.sign_context_hv(-h, -v) == O.sign_context_hv(h, v) ^ 1


# ---- geometry -------------------------------------------------------------------------

@pytest.mark.parametrize("size,levels", [(2048, 5), (4096, 5), (1024, 3), (512, 1), (256, 0)])
def test_find_subband_matches_mallat_geometry(O, size, levels):
    code = {"HL": 0, "LH": 1, "HH": 2}
    for name, d, (sy, sx) in synth.band_slices(size, size, levels):
        for y in range(sy.start, sy.stop, 64):
            for x in range(sx.start, sx.stop, 16):
                lv, sb = O.find_subband(x, y, size, size, levels)
                if name == "LL":
                    assert (lv, sb) == (levels, 0)
                else:
                    assert (lv, sb) == (d - 1, code[name]), (name, d, x, y)


def test_msb_is_bit_length_minus_one(O):
    rng = np.random.default_rng(3)
    img = np.zeros((128, 256), np.int32)
    for cb in range(8):
        ty, tx = divmod(cb, 4)
        bits = int(rng.integers(1, 28))
        blk = rng.integers(-(1 << bits) + 1, 1 << bits, size=(64, 64))
        img[ty * 64:(ty + 1) * 64, tx * 64:(tx + 1) * 64] = blk
        m = int(np.bitwise_or.reduce(np.abs(blk).ravel()))
        assert O.codeblock_msb(img, cb) == m.bit_length() - 1
    img[:64, :64] = 0
    assert O.codeblock_msb(img, 0) == 32  # all zero: 32 - ffs(brev(0)) = 32


# ---- LUT loader (main_example.cu:219-448) --------------------------------------------

def _naive_table(folder, cpass3, levels=5):
    names = ["ref.txt", "sig.txt", "sign.txt"] + (["cp_sig.txt", "cp_sign.txt"] if cpass3 else [])
    nvals = {"ref.txt": 1, "sig.txt": 9, "sign.txt": 4, "cp_sig.txt": 9, "cp_sign.txt": 4}
    rows = {}
    maxbp = 0
    for nm in names:
        rows[nm] = []
        for line in open(folder + nm):
            a, b = line.split(":")
            lv, sb, bp = map(int, a.split())
            vals = list(map(int, b.split()))
            assert len(vals) == nvals[nm]
            rows[nm].append((lv, sb, bp, vals))
            if nm in ("ref.txt", "sig.txt", "sign.txt"):
                maxbp = max(maxbp, bp + 1)
    G = 3 * levels + 1
    out = []
    for nm in names:
        N = nvals[nm]
        reg = np.full(G * maxbp * N, 64, np.int64)  # rows a file never lists read 64
        for lv, sb, bp, vals in rows[nm]:
            g = lv * 3 + sb
            reg[g * maxbp * N + bp * N: g * maxbp * N + bp * N + N] = vals
        out.append(reg)
    return np.concatenate(out), maxbp


@pytest.mark.parametrize("cpass3", [0, 1])
def test_lut_loader_layout(O, cpass3):
    from conftest import lut_folder
    lut, nbp = O.lut_load_text(lut_folder(cpass3), cpass3)
    assert nbp == 15
    assert lut.size == (6480 if cpass3 else 3360)
    naive, maxbp = _naive_table(lut_folder(cpass3), cpass3)
    assert maxbp == 15
    # the loader's padding rule (64 past each group's last listed row, starting one
    # entry early) leaves every listed value intact on the shipped files
    assert np.array_equal(lut, naive)
    assert lut.min() >= 1 and lut.max() <= 127
