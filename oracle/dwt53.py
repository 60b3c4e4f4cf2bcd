"""Reversible integer 5/3 (LeGall) wavelet transform -- CPU oracle.

TEST INFRASTRUCTURE ONLY: imported by tests/ and bench.py's checks, never by the
product path (cuda_bpc-paco_amd/).

The reference codes coefficients that were transformed before they reach it (its
sample image ships transformed, main_example.cu:71-179) and holds no transform
code, so this restates the published algorithm: the reversible 5/3 lifting of
ISO/IEC 15444-1 Annex F with whole-sample symmetric extension.  For a signal
x[0..2M) (even length, M >= 1):

    d[n] = x[2n+1] - floor((x[2n] + x[2n+2]) / 2)        x[2M] := x[2M-2]
    s[n] = x[2n]   + floor((d[n-1] + d[n] + 2) / 4)      d[-1] := d[0]

and the inverse undoes the two steps in reverse order.  In 2-D each level
transforms the columns (vertical) and then the rows (horizontal) of the current
LL band and stores the bands in the Mallat layout that Find_Subband
(BPC_PaCo.cu:84-111) assumes: LL top-left, HL (horizontal high) top-right, LH
(vertical high) bottom-left, HH bottom-right.
"""
import numpy as np


def _fwd1(x, axis):
    x = np.moveaxis(x, axis, 0).astype(np.int64)
    ev, od = x[0::2], x[1::2]
    nxt = np.concatenate([ev[1:], ev[-1:]], axis=0) if ev.shape[0] > 1 else ev  # x[2n+2], x[2M] = x[2M-2]
    d = od - ((ev + nxt) >> 1)
    prv = np.concatenate([d[:1], d[:-1]], axis=0)  # d[n-1], d[-1] = d[0]
    s = ev + ((prv + d + 2) >> 2)
    return np.moveaxis(np.concatenate([s, d], axis=0), 0, axis)


def _inv1(y, axis):
    y = np.moveaxis(y, axis, 0).astype(np.int64)
    m = y.shape[0] // 2
    s, d = y[:m], y[m:]
    prv = np.concatenate([d[:1], d[:-1]], axis=0)
    ev = s - ((prv + d + 2) >> 2)
    nxt = np.concatenate([ev[1:], ev[-1:]], axis=0) if m > 1 else ev
    od = d + ((ev + nxt) >> 1)
    x = np.empty_like(y)
    x[0::2], x[1::2] = ev, od
    return np.moveaxis(x, 0, axis)


def check_shape(shape, levels):
    Y, X = shape
    if levels < 0 or X < 1 or Y < 1 or X % (1 << levels) or Y % (1 << levels):
        raise ValueError("%dx%d cannot take %d levels" % (X, Y, levels))


def forward(img, levels):
    """levels-level 2-D forward transform of a (Y, X) integer image (int32 result)."""
    check_shape(img.shape, levels)
    out = np.array(img, dtype=np.int64)
    Y, X = out.shape
    for lv in range(levels):
        h, w = Y >> lv, X >> lv
        band = out[:h, :w]
        band = _fwd1(band, 0)
        band = _fwd1(band, 1)
        out[:h, :w] = band
    return out.astype(np.int32)


def inverse(coeffs, levels):
    check_shape(coeffs.shape, levels)
    out = np.array(coeffs, dtype=np.int64)
    Y, X = out.shape
    for lv in range(levels - 1, -1, -1):
        h, w = Y >> lv, X >> lv
        band = out[:h, :w]
        band = _inv1(band, 1)
        band = _inv1(band, 0)
        out[:h, :w] = band
    return out.astype(np.int32)


# ---- plain-loop form (small inputs): the same definitions, sample by sample ----

def _ext(x, i):
    """Whole-sample symmetric extension: x[-i] = x[i], x[N-1+i] = x[N-1-i]."""
    n = len(x)
    if n == 1:
        return x[0]
    p = 2 * (n - 1)
    i %= p
    return x[i] if i < n else x[p - i]


def forward_1d_loop(x):
    x = [int(v) for v in x]
    m = len(x) // 2
    d = [x[2 * k + 1] - (x[2 * k] + _ext(x, 2 * k + 2)) // 2 for k in range(m)]
    dd = lambda k: d[0] if k < 0 else d[k]  # noqa: E731  (d[-1] = x[-1] - (x[-2] + x[0]) // 2 = d[0])
    s = [x[2 * k] + (dd(k - 1) + d[k] + 2) // 4 for k in range(m)]
    return s + d


def forward_loop(img, levels):
    out = [[int(v) for v in row] for row in np.asarray(img)]
    Y, X = len(out), len(out[0])
    for lv in range(levels):
        h, w = Y >> lv, X >> lv
        for c in range(w):
            col = forward_1d_loop([out[r][c] for r in range(h)])
            for r in range(h):
                out[r][c] = col[r]
        for r in range(h):
            out[r][:w] = forward_1d_loop(out[r][:w])
    return np.array(out, dtype=np.int32)
