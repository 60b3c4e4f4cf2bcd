"""Symbol statistics of the oracle (the checker of the device LUT trainer).

The per-entry counts are pinned three ways:
- closed form: summed over contexts, the symbols of each (group, plane, pass kind)
  follow from the magnitudes alone.  A coefficient of magnitude m (top bit t,
  t = -1 for m = 0) in a codeblock of MSB B has a significance symbol at every
  plane B..t (symbol 1 at t), a sign symbol at plane t and a refinement symbol
  at every plane below t (the bit of m).  In 3-pass mode the significance symbol
  of a plane falls in SPP or CP, never both.
- a hand count on a one-coefficient codeblock;
- agreement with the library's host trainer (an independent implementation
  over packed words).
"""
import numpy as np
import pytest

import synth


def groups_of(size_x, size_y, levels, rows):
    """LUT group (level*3 + subband) of every column pair's codeblock row, per sample."""
    gx = np.zeros((size_y, size_x), np.int64)
    for y0 in range(0, size_y, rows):
        for x in range(0, size_x, 2):
            g = levels * 3
            for aux in range(1, levels + 1):
                bx, by = x >= (size_x >> aux), y0 >= (size_y >> aux)
                if bx or by:
                    g = (aux - 1) * 3 + ((2 if by else 0) if bx else 1)
                    break
            gx[y0:y0 + rows, x:x + 2] = g
    return gx


def closed_form(img, levels, rows, nbp):
    """(sig, sig_zero, sign, ref, ref_zero) totals, each [group, plane]."""
    Y, X = img.shape
    G = 3 * levels + 1
    mag = np.abs(img.astype(np.int64))
    top = np.where(mag > 0, np.floor(np.log2(np.maximum(mag, 1))).astype(np.int64), -1)
    grp = groups_of(X, Y, levels, rows)
    out = [np.zeros((G, nbp), np.int64) for _ in range(5)]
    for ty in range(Y // rows):
        for tx in range(X // 64):
            sl = (slice(ty * rows, ty * rows + rows), slice(tx * 64, tx * 64 + 64))
            m, t, g = mag[sl], top[sl], grp[sl]
            if m.max() == 0 or m.max() >= 1 << 28:
                continue
            B = int(t.max())
            for bp in range(min(B, nbp - 1) + 1):
                coded = t <= bp
                np.add.at(out[0][:, bp], g[coded], 1)
                np.add.at(out[1][:, bp], g[t < bp], 1)
                np.add.at(out[2][:, bp], g[t == bp], 1)
                np.add.at(out[3][:, bp], g[t > bp], 1)
                np.add.at(out[4][:, bp], g[(t > bp) & (((m >> bp) & 1) == 0)], 1)
    return out


def region_sums(n, levels, nbp, cp3):
    """Oracle/trainer counts folded to [group, plane] per pass kind (SPP+CP summed)."""
    G = 3 * levels + 1
    ref = n[:G * nbp].reshape(G, nbp)
    sig = n[G * nbp:G * nbp * 10].reshape(G, nbp, 9).sum(-1)
    sign = n[G * nbp * 10:G * nbp * 14].reshape(G, nbp, 4).sum(-1)
    if cp3:
        off = G * nbp * 13
        sig = sig + n[off + G * nbp:off + G * nbp * 10].reshape(G, nbp, 9).sum(-1)
        sign = sign + n[off + G * nbp * 10:off + G * nbp * 14].reshape(G, nbp, 4).sum(-1)
    return sig.astype(np.int64), sign.astype(np.int64), ref.astype(np.int64)


def lut_from_counts(n0, n):
    n0 = n0.astype(np.int64)
    n = n.astype(np.int64)
    v = (n0 * 128 + n // 2) // np.maximum(n, 1)
    return np.where(n > 0, np.clip(v, 1, 127), 64).astype(np.int32)


def counts_image(levels, rows, seed, size=256):
    img = synth.synth_image(size, size, levels, seed, ll_bits=12, detail_exp=7, mag_bits=13)
    img[0:rows, 64:128] = 0          # an all-zero codeblock: counts nothing
    img[rows:2 * rows, 0:64][5, 7] = 1 << 28  # a rejected codeblock: counts nothing
    return img


@pytest.mark.parametrize("cp3,levels,rows,nbp", [(0, 2, 64, 8), (1, 2, 64, 8), (1, 3, 32, 14), (0, 0, 64, 15)])
def test_counts_closed_form(O, cp3, levels, rows, nbp):
    img = counts_image(levels, rows, seed=50 + cp3 + levels)
    op = O.params(cpass3=cp3, dwt_levels=levels, cblock_length=rows, lut_n_bitplanes=nbp)
    st, n0, n = O.symbol_counts(img, op)
    assert st == 0
    sig, sig0, sign, ref, ref0 = closed_form(img, levels, rows, nbp)
    osig, osign, oref = region_sums(n, levels, nbp, cp3)
    osig0, _, oref0 = region_sums(n0, levels, nbp, cp3)
    assert np.array_equal(osig, sig)
    assert np.array_equal(osig0, sig0)
    assert np.array_equal(osign, sign)
    assert np.array_equal(oref, ref)
    assert np.array_equal(oref0, ref0)
    assert sig.sum() > 0 and ref.sum() > 0
    if cp3:  # both passes take part
        G = 3 * levels + 1
        off = G * nbp * 13
        assert n[G * nbp:G * nbp * 10].sum() > 0 and n[off + G * nbp:off + G * nbp * 10].sum() > 0


def test_counts_hand_traced(O):
    """One coefficient = 1 at the top-left corner, 2-pass, MSB 0: 4093 significance
    symbols 0 ... see the host-trainer test of the same case."""
    img = np.zeros((64, 64), np.int32)
    img[0, 0] = 1
    st, n0, n = O.symbol_counts(img, O.params(cpass3=0, dwt_levels=0, lut_n_bitplanes=1))
    assert st == 0
    sig_n, sig_n0 = n[1:10], n0[1:10]
    # (0,0) coded first in context 0 (symbol 1); its 3 neighbours later in context 1
    assert list(sig_n[:2]) == [4093, 3] and list(sig_n0[:2]) == [4092, 3] and sig_n[2:].sum() == 0
    assert n[10] == 1 and n0[10] == 1 and n[11:14].sum() == 0  # one sign, context 0, predicted
    assert n[0] == 0  # no refinement at plane 0


@pytest.mark.parametrize("cp3,levels,rows", [(0, 3, 64), (1, 3, 64), (1, 5, 16)])
def test_counts_match_host_trainer(bpc, O, cp3, levels, rows):
    img = synth.synth_image(512, 512, levels, 60 + cp3, mag_bits=14)
    op = O.params(cpass3=cp3, dwt_levels=levels, cblock_length=rows, lut_n_bitplanes=15)
    st, n0, n = O.symbol_counts(img, op)
    assert st == 0
    lut = bpc.train_lut(img, bpc.params(cpass3=cp3, dwt_levels=levels, cblock_length=rows, lut_n_bitplanes=15))
    assert np.array_equal(lut_from_counts(n0, n), lut)
