"""Synthetic DWT-coefficient images for the BPC-PaCo workloads (numpy only).

An image of size X*Y in the Mallat layout the codec classifies (BPC_PaCo.cu:84-111):
the LL band (coarsest) at the top-left (X>>L)*(Y>>L), and for each level d = 1
(finest) .. L (coarsest) the HL band right of, the LH band below and the HH band
diagonal to the level's low-pass square.

Magnitudes: LL uniform on [0, 2^ll_bits); HL/LH floor(|Laplace(0, b)|) with
b = 2^(detail_exp - d); HH with b/2; all capped at 2^mag_bits - 1.  Signs uniform.
Stored as two's-complement int32 (the codec's input format).
"""
import numpy as np


def band_slices(size_x, size_y, levels):
    """Yields (name, d, (slice_y, slice_x)) for every subband; d = 0 for LL."""
    lx, ly = size_x >> levels, size_y >> levels
    yield "LL", 0, (slice(0, ly), slice(0, lx))
    for d in range(1, levels + 1):
        sx, sy = size_x >> d, size_y >> d
        yield "HL", d, (slice(0, sy), slice(sx, 2 * sx))
        yield "LH", d, (slice(sy, 2 * sy), slice(0, sx))
        yield "HH", d, (slice(sy, 2 * sy), slice(sx, 2 * sx))


def synth_image(size_x, size_y, levels=5, seed=0, ll_bits=13, detail_exp=9, mag_bits=14):
    rng = np.random.default_rng(seed)
    img = np.empty((size_y, size_x), dtype=np.int32)
    cap = (1 << mag_bits) - 1
    for name, d, (sy, sx) in band_slices(size_x, size_y, levels):
        shape = (sy.stop - sy.start, sx.stop - sx.start)
        if name == "LL":
            mag = rng.integers(0, 1 << ll_bits, size=shape, dtype=np.int64)
        else:
            b = float(2 ** (detail_exp - d))
            if name == "HH":
                b *= 0.5
            mag = np.floor(rng.exponential(b, size=shape)).astype(np.int64)
        np.minimum(mag, cap, out=mag)
        neg = rng.integers(0, 2, size=shape, dtype=np.int64).astype(bool)
        img[sy, sx] = np.where(neg, -mag, mag).astype(np.int32)
    return img


def to_tiled(img, rows=64):
    """Raster (Y, X) -> codeblock-contiguous flat array (tile t at t*64*rows, row-major inside)."""
    Y, X = img.shape
    return np.ascontiguousarray(img.reshape(Y // rows, rows, X // 64, 64).transpose(0, 2, 1, 3)).ravel()


def from_tiled(flat, size_x, size_y, rows=64):
    t = flat.reshape(size_y // rows, size_x // 64, rows, 64).transpose(0, 2, 1, 3)
    return np.ascontiguousarray(t).reshape(size_y, size_x)


# Workloads: configuration of each named case (image sizes, bitplanes, passes).
WORKLOADS = {
    "plain_2048_2pass": dict(sizes=[2048], cpass3=[0], nbp=15, mag_bits=14, ll_bits=13, detail_exp=9),
    "plain_4096_3pass": dict(sizes=[4096], cpass3=[1], nbp=16, mag_bits=16, ll_bits=13, detail_exp=9),
    "batched_96": dict(sizes=[2048] * 32 + [3072] * 32 + [4096] * 32, cpass3=[0, 1], nbp=15, mag_bits=14,
                       ll_bits=13, detail_exp=9),
    "large_16384": dict(sizes=[16384], cpass3=[1], nbp=20, mag_bits=20, ll_bits=19, detail_exp=14),
}
