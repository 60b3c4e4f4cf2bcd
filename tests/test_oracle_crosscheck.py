"""Pins the oracle's codeword slot order with an independent formulation.

The oracle interleaves 32 lanes step by step and hands out slots at each coder call
site in lane order.  Here each lane's coder runs on its own over the lane's symbol
sequence, recording when (plane, pass, row, column parity, call site) it opens a
codeword; slots are then assigned by sorting all openings by (time, lane).  A
mistake in either (lane order, site order, pass order, scan order, the reservation
rule) makes the two codestreams differ.  Pure Python, small codeblocks only.
"""
import numpy as np
import pytest

S_BIT, C_BIT, R_BIT = 1 << 31, 1 << 30, 1 << 29


def chi(w):
    return 0 if not (w & S_BIT) else (-1 if w & 1 else 1)


def sign_ctx(h, v):
    table = {(0, 0): 0, (0, 1): 2, (0, -1): 3, (1, 0): 4, (1, 1): 6, (1, -1): 0,
             (-1, 0): 5, (-1, 1): 1, (-1, -1): 7}
    sg = lambda t: (t > 0) - (t < 0)
    return table[(sg(h), sg(v))]


def group(x, y, X, Y, L):
    for aux in range(1, L + 1):
        if x >= X >> aux or y >= Y >> aux:
            return (aux - 1) * 3 + (2 if (x >= X >> aux and y >= Y >> aux) else (0 if x >= X >> aux else 1))
    return 3 * L


def encode_events(block, X, Y, tx, ty, lut, cp3, L, nbp):
    """Symbol events of one codeblock: list of (time, lane, symbol, prob)."""
    rows = block.shape[0]
    mag = np.abs(block.astype(np.int64))
    m = int(np.bitwise_or.reduce(mag.ravel()))
    msb = m.bit_length() - 1
    w = [[(int(mag[r, c]) << 1) | int(block[r, c] < 0) | (C_BIT if cp3 else 0) for c in range(64)] for r in range(rows)]
    G = 3 * L + 1
    off_sig, off_sign, cpoff = G * nbp, G * nbp * 10, G * nbp * 13
    grp = [group(tx * 64 + (c & ~1), ty * rows, X, Y, L) for c in range(64)]

    def at(r, c):
        return w[r][c] if 0 <= r < rows and 0 <= c < 64 else 0

    events = []
    planes = list(range(msb, -1, -1))
    for pi, bp in enumerate(planes):
        passes = ["spp", "mrp"] + (["cp"] if cp3 else [])
        if cp3 and bp == msb:
            passes = ["cp"]
        for pass_name in passes:
            pidx = {"spp": 0, "mrp": 1, "cp": 2}[pass_name]
            for r in range(rows):
                for par in (0, 1):
                    for c in range(par, 64, 2):
                        lane = c // 2
                        g = grp[c]
                        x = w[r][c]
                        t1 = (pi, pidx, r, par, 0)
                        t2 = (pi, pidx, r, par, 1)
                        if pass_name == "mrp":
                            if x & R_BIT:
                                events.append((t1, lane, (x >> (bp + 1)) & 1, lut[g * nbp + bp]))
                            elif x & S_BIT:
                                w[r][c] = x | R_BIT
                            continue
                        nb = [at(r - 1, c - 1), at(r - 1, c), at(r - 1, c + 1), at(r, c - 1), at(r, c + 1),
                              at(r + 1, c - 1), at(r + 1, c), at(r + 1, c + 1)]
                        ctx = sum(1 for n in nb if n & S_BIT)
                        clean = pass_name == "cp"
                        if not clean:
                            if x & S_BIT:
                                continue
                            if cp3 and ctx == 0:
                                w[r][c] = x | C_BIT
                                continue
                        else:
                            if not x & C_BIT:
                                continue
                            x &= ~C_BIT
                        base = off_sig + (cpoff if clean else 0) + g * nbp * 9 + bp * 9
                        sym = (x >> (bp + 1)) & 1
                        events.append((t1, lane, sym, lut[base + ctx]))
                        if sym:
                            x |= S_BIT | (R_BIT if clean else 0)
                            sc = sign_ctx(chi(nb[3]) + chi(nb[4]), chi(nb[1]) + chi(nb[6]))
                            sbase = off_sign + (cpoff if clean else 0) + g * nbp * 4 + bp * 4
                            events.append((t2, lane, (x & 1) ^ (sc & 1), lut[sbase + (sc >> 1)]))
                        w[r][c] = x
    return msb, events


def decoupled_region(block, X, Y, tx, ty, lut, cp3, L, nbp, stride):
    msb, events = encode_events(block, X, Y, tx, ty, lut, cp3, L, nbp)
    openings = []  # (time, lane, codeword value)
    per_lane = {}
    for t, lane, sym, prob in events:
        per_lane.setdefault(lane, []).append((t, sym, prob))
    for lane, evs in per_lane.items():
        Lo, Sz = 0, 0
        cur = None
        for t, sym, prob in evs:
            if Sz == 0:
                Lo, Sz = 0, 0xFFFF
                cur = [t, lane, None]
                openings.append(cur)
            aux = ((Sz * prob) >> 7) + sym
            if sym:
                Sz -= aux
                Lo += aux
            else:
                Sz = aux
            cur[2] = Lo
    openings.sort(key=lambda o: (o[0], o[1]))
    region = np.full(stride, -1, np.int64)
    region[0] = msb
    for k, (_, _, val) in enumerate(openings):
        region[1 + k] = val
    return region.astype(np.int32), len(openings)


@pytest.mark.parametrize("cp3", [0, 1])
@pytest.mark.parametrize("seed", [0, 1, 2])
def test_decoupled_coders_match_lockstep_oracle(O, cp3, seed):
    rng = np.random.default_rng(seed)
    rows, L = 8, 1
    X, Y = 128, 16  # 4 codeblocks of 64x8, level-0 bands and LL
    img = np.floor(rng.exponential(6.0, size=(Y, X))).astype(np.int32)
    img *= np.where(rng.random((Y, X)) < 0.5, -1, 1).astype(np.int32)
    img[0:8, 0:64] = 0
    img[2, 10] = 5  # one isolated coefficient in codeblock 0
    nbp = 4 + int(np.abs(img).max()).bit_length()
    p = O.params(cpass3=cp3, dwt_levels=L, cblock_length=rows, lut_n_bitplanes=nbp)
    lut = rng.integers(1, 128, size=O.lut_len(p)).astype(np.int32)
    st, stream, ncw = O.encode(img, lut, p)
    assert st == 0
    stride = 64 * rows
    for cb in range(4):
        ty, tx = divmod(cb, X // 64)
        block = img[ty * rows:(ty + 1) * rows, tx * 64:(tx + 1) * 64]
        region, n = decoupled_region(block, X, Y, tx, ty, lut, cp3, L, nbp, stride)
        assert n == ncw[cb]
        assert np.array_equal(region, stream[cb * stride:(cb + 1) * stride]), cb
