"""The 5/3 transform oracle against hand-derived values, closed forms and a
sample-by-sample restatement of the same definitions."""
import ctypes

import numpy as np
import pytest

import dwt53


def test_hand_traced_1d():
    # x = 1..8: d[n] = x[2n+1] - floor((x[2n] + x[2n+2]) / 2) = 0, except d[3] = 8 - floor((7 + x[6]) / 2) = 1
    # (x[8] = x[6]); s[n] = x[2n] + floor((d[n-1] + d[n] + 2) / 4) = x[2n] (s[3]: floor(3/4) = 0)
    assert dwt53.forward_1d_loop(range(1, 9)) == [1, 3, 5, 7, 0, 0, 0, 1]
    # floor, not truncation, for negative sums: x = [-2, 0, -1, 1]
    # d0 = 0 - floor(-3/2) = 2, d1 = 1 - floor((-1 + -1)/2) = 2; s0 = -2 + floor(6/4) = -1, s1 = -1 + 1 = 0
    assert dwt53.forward_1d_loop([-2, 0, -1, 1]) == [-1, 0, 2, 2]
    # two samples: d0 = 8 - 5 = 3 (x[2] = x[0]); s0 = 5 + floor((3 + 3 + 2) / 4) = 7
    assert dwt53.forward_1d_loop([5, 8]) == [7, 3]
    for x in ([1, 2, 3, 4, 5, 6, 7, 8], [-2, 0, -1, 1], [5, 8]):
        a = np.array([x], np.int32)
        assert list(dwt53.forward(a.T, 0)[:, 0]) == x  # zero levels: identity
        assert list(dwt53._fwd1(np.array(x), 0)) == dwt53.forward_1d_loop(x)


def test_constant_and_ramp_closed_forms():
    c = np.full((64, 128), 77, np.int32)
    f = dwt53.forward(c, 3)
    assert (f[:8, :16] == 77).all()  # LL stays the constant (s = c + floor(2/4))
    f[:8, :16] = 0
    assert not f.any()  # every detail band is zero
    ramp = np.tile(np.arange(16, dtype=np.int32), (8, 1))  # f(x, y) = x
    f = dwt53.forward(ramp, 1)
    # columns are constant: vertical H = 0, L = x; rows: s[n] = 2n, d = 0 but d[last] = 15 - 14 = 1
    assert (f[:4, :8] == np.arange(0, 16, 2)).all()
    assert (f[:4, 8:15] == 0).all() and (f[:4, 15] == 1).all()
    assert not f[4:].any()


@pytest.mark.parametrize("shape,levels", [((8, 8), 1), ((8, 8), 3), ((16, 4), 2), ((2, 2), 1), ((4, 8), 2),
                                          ((32, 64), 5), ((12, 20), 2), ((64, 64), 6)])
def test_vectorized_equals_loop_form(shape, levels):
    rng = np.random.default_rng(shape[0] * 100 + shape[1] + levels)
    a = rng.integers(-(1 << 20), 1 << 20, shape).astype(np.int32)
    f = dwt53.forward(a, levels)
    assert np.array_equal(f, dwt53.forward_loop(a, levels))
    assert np.array_equal(dwt53.inverse(f, levels), a)


def test_band_layout_matches_find_subband(O):
    """Alternating columns are horizontal detail of the finest level: the transform
    puts it where Find_Subband reads HL of level 0 (the rest is a constant LL);
    alternating rows go to LH.  Closed form: rows [0, 100, 0, 100, ...] give
    d = 100 everywhere (x[N] = x[N-2] = 0) and s = floor((100 + 100 + 2) / 4) = 50."""
    X = Y = 256
    for pattern, detail in ((np.s_[:, 1::2], (0, 0)), (np.s_[1::2, :], (0, 1))):
        img = np.zeros((Y, X), np.int32)
        img[pattern] = 100
        f = dwt53.forward(img, 3)
        ys, xs = np.nonzero(f)
        bands = {O.find_subband(int(x) & ~1, int(y) & ~63, X, Y, 3) for y, x in zip(ys, xs)}
        assert bands == {detail, (3, 0)}
        assert set(np.unique(f).tolist()) == {0, 50, 100}


def test_device_entry_checks_sizes(bpc):
    lib = bpc.lib()
    a, b = ctypes.c_void_p(4096), ctypes.c_void_p(1 << 40)
    assert lib.bpc_dwt53_forward_device(a, b, 96, 64, 6, None) == bpc.ERR_ARG  # 96 not divisible by 64
    assert lib.bpc_dwt53_inverse_device(a, a, 64, 64, 2, None) == bpc.ERR_ARG  # in place
    assert lib.bpc_dwt53_forward_device(None, b, 64, 64, 1, None) == bpc.ERR_ARG
