"""Per-codeblock codestream files and the compact container, host side.

The file order restates load_Data (main_example.cu:71-179); the oracle's copy is
pinned by the reference's own sample image, whose codeblock file k carries the
level, subband and MSB of the tile it is placed at.  Files are checked byte for
byte and through the oracle's load_CStream restatement (main_example.cu:51-68).
"""
import os

import numpy as np
import pytest

import synth


def compact_of(stream, ncw, stride):
    """Compact form of reference-layout regions (MSB byte, offsets, uint16 codewords)."""
    regions = stream.reshape(-1, stride)
    msb = regions[:, 0].astype(np.uint8)
    offsets = np.concatenate([[0], np.cumsum(ncw.astype(np.uint64))]).astype(np.uint64)
    parts = [regions[k, 1:1 + int(n)] for k, n in enumerate(ncw)]
    payload = np.concatenate(parts).astype(np.uint16) if parts else np.zeros(0, np.uint16)
    return msb, offsets, payload


def test_file_order_pinned_by_sample_headers(O, sample_headers):
    """File k of the reference's sample image sits at the tile whose subband and MSB
    its header names (the placement is checked against the MSBs in test_oracle_sample)."""
    X = sample_headers["size_x"]
    order = O.cb_file_order(X, sample_headers["size_y"], sample_headers["dwt_levels"])
    assert order is not None and len(order) == len(sample_headers["codeblocks"])
    for h in sample_headers["codeblocks"]:
        t = int(order[h["k"]])
        assert (t % (X // 64), t // (X // 64)) == (h["x"] // 64, h["y"] // 64)
        assert O.find_subband(h["x"], h["y"], X, X, sample_headers["dwt_levels"]) == (h["level"], h["subband"])


@pytest.mark.parametrize("size_x,size_y,levels,rows", [
    (2048, 2048, 5, 64), (4096, 4096, 5, 64), (16384, 16384, 5, 64), (1024, 512, 3, 32), (512, 512, 0, 64),
    (4096, 2048, 5, 16), (3072, 3072, 5, 64), (2048, 2048, 6, 64)])
def test_file_order_matches_oracle(bpc, O, size_x, size_y, levels, rows):
    want = O.cb_file_order(size_x, size_y, levels, rows)
    if want is None:
        with pytest.raises(bpc.BPCError):
            bpc.cb_file_order(size_x, size_y, levels, rows)
        return
    got = bpc.cb_file_order(size_x, size_y, levels, rows)
    assert np.array_equal(got, want)
    assert np.array_equal(np.sort(got), np.arange(got.size))  # every tile exactly once


def test_load_cstream_restatement(O):
    # big-endian 16-bit words, every word past the end of the file is -1 (fgetc EOF)
    assert list(O.load_cstream(b"\x00\x0c\x12\x34", 4)) == [12, 0x1234, -1, -1]
    assert list(O.load_cstream(b"\xff\xff\x00", 3)) == [0xFFFF, -1, -1]  # odd length: lone byte | EOF
    assert list(O.load_cstream(b"", 2)) == [-1, -1]


@pytest.mark.parametrize("cp3", [0, 1])
def test_bin_export_import_round_trip(bpc, O, tmp_path, cp3):
    size, levels = 512, 3
    img = synth.synth_image(size, size, levels, 120 + cp3)
    img[0:64, 0:64] = 0  # an all-zero codeblock: its file holds the MSB word only
    op = O.params(cpass3=cp3, dwt_levels=levels)
    lut = bpc.train_lut(img, bpc.params(cpass3=cp3, dwt_levels=levels))
    st, stream, ncw = O.encode(img, lut, op)
    assert st == 0
    p = bpc.params(cpass3=cp3, dwt_levels=levels)
    msb, offsets, payload = compact_of(stream, ncw, p.stride())
    folder = str(tmp_path) + "/"
    bpc.bin_export(folder, msb, offsets, payload, size, size, p)
    order = O.cb_file_order(size, size, levels)
    regions = stream.reshape(-1, p.stride())
    for k, t in enumerate(order):
        data = open(os.path.join(folder, "%d.bin" % k), "rb").read()
        assert len(data) == 2 * (1 + int(ncw[t]))
        assert np.array_equal(O.load_cstream(data, p.stride()), regions[t])
    assert np.array_equal(bpc.bin_import(folder, size, size, p), stream)
    st, rec = O.decode(bpc.bin_import(folder, size, size, p), size, size, lut, op)
    assert st == 0 and np.array_equal(rec, img)


def test_bin_full_region(bpc, O, tmp_path):
    """A codeblock whose codewords fill its region: the file is the whole region."""
    p = bpc.params(dwt_levels=0, cb_stride_words=8)
    msb = np.array([3, 32], np.uint8)
    offsets = np.array([0, 7, 7], np.uint64)
    payload = np.array([1, 2, 65535, 4, 5, 6, 7], np.uint16)
    bpc.bin_export(str(tmp_path), msb, offsets, payload, 128, 64, p)
    got = bpc.bin_import(str(tmp_path), 128, 64, p)
    assert list(got) == [3, 1, 2, 65535, 4, 5, 6, 7] + [32] + [-1] * 7
    assert os.path.getsize(os.path.join(str(tmp_path), "0.bin")) == 16
