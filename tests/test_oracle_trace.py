"""The oracle's per-symbol trace (the reference's THREAD_DEBUG_ID printfs,
BPC_PaCo.cu:266,317,434,445, kept as records): consistent with the codestream it
produced, with the decoder, with the symbol statistics and with a hand-derived
tiny case, so that it can be trusted when a parity failure is chased down."""
import numpy as np
import pytest

import synth


@pytest.mark.parametrize("cp3,rows", [(0, 64), (1, 64), (1, 32), (0, 16)])
def test_trace_encoder_equals_decoder(O, cp3, rows):
    img = synth.synth_image(256, 256, 3, 40 + cp3 + rows)
    p = O.params(cpass3=cp3, dwt_levels=3, cblock_length=rows)
    train = synth.synth_image(256, 256, 3, 7)  # a table trained from the oracle's own counts
    st, n0, n = O.symbol_counts(train, p)
    assert st == 0
    n0, n = n0.astype(np.int64), n.astype(np.int64)
    table = np.where(n > 0, np.clip((n0 * 128 + n // 2) // np.maximum(n, 1), 1, 127), 64).astype(np.int32)
    for cb in (0, 5, O.n_codeblocks(256, 256, p) - 1):
        st, region, ncw, tr = O.encode_cb_trace(img, cb, table, p)
        assert st == 0
        st2, rec, tr2 = O.decode_cb_trace(region, 256, 256, cb, table, p)
        assert st2 == 0
        assert np.array_equal(tr, tr2)
        # the decoded codeblock is the original one
        x0, y0 = (cb % 4) * 64, (cb // 4) * rows
        assert np.array_equal(rec[y0:y0 + rows, x0:x0 + 64], img[y0:y0 + rows, x0:x0 + 64])


def test_trace_slots_and_codewords(O):
    """Openings take slots 1, 2, 3, ... in trace order (lock-step order with lanes
    ascending), and replaying each lane's (symbol, probability) sequence through the
    single-lane coder gives exactly the codewords found at that lane's slots."""
    img = synth.synth_image(512, 512, 4, 77)
    p = O.params(cpass3=1, dwt_levels=4)
    lut = np.full(O.lut_len(p), 64, np.int32)
    lut[::3] = 100
    lut[1::7] = 20
    for cb in (0, 9, 40):
        st, region, ncw, tr = O.encode_cb_trace(img, cb, lut, p)
        assert st == 0 and ncw > 0
        opened = tr[tr["opened"] == 1]
        assert np.array_equal(opened["slot"], np.arange(1, len(opened) + 1))
        assert len(opened) == ncw
        assert np.array_equal(tr["prob"], lut[tr["lut_index"]])
        for lane in range(32):
            t = tr[tr["lane"] == lane]
            cw = O.ac_encode_seq(t["sym"].astype(np.uint8), t["prob"].astype(np.int32))
            slots = t[t["opened"] == 1]["slot"]
            assert np.array_equal(region[slots].astype(np.uint32), np.asarray(cw, dtype=np.uint32)), lane


def test_trace_counts_equal_symbol_statistics(O):
    """Summed over every codeblock of an image, the trace's symbols per LUT entry are
    the oracle's symbol counts (an independent pass of the same modelling)."""
    img = synth.synth_image(256, 192, 2, 91)
    for cp3 in (0, 1):
        p = O.params(cpass3=cp3, dwt_levels=2, lut_n_bitplanes=15)
        lut = np.full(O.lut_len(p), 64, np.int32)
        n_tr = np.zeros(lut.size, np.int64)
        n0_tr = np.zeros(lut.size, np.int64)
        for cb in range(O.n_codeblocks(256, 192, p)):
            st, _, _, tr = O.encode_cb_trace(img, cb, lut, p)
            assert st == 0
            np.add.at(n_tr, tr["lut_index"], 1)
            np.add.at(n0_tr, tr["lut_index"][tr["sym"] == 0], 1)
        st, n0, n = O.symbol_counts(img, p)
        assert st == 0 and np.array_equal(n_tr, n) and np.array_equal(n0_tr, n0)


def test_trace_one_coefficient_hand_case(O):
    """A 64x64 image (one codeblock, no transform levels) whose only nonzero
    coefficient is -1 at row 0, column 0, 2 passes, every probability 64:
    plane 0 only.  The significance pass codes every coefficient (2-pass mode has
    no cleanup pass): at the even step of row 0, lanes 0..31 code their
    significance symbols in lane order and open slots 1..32, then lane 0 codes the
    sign (no significant neighbour: context 0, predicted positive, so the symbol is
    1 for a negative coefficient); the odd step follows, and so on.  4096
    significance symbols, one sign, no refinement symbol (the coefficient only
    becomes significant in this plane)."""
    img = np.zeros((64, 64), np.int32)
    img[0, 0] = -1
    p = O.params(cpass3=0, dwt_levels=0, lut_n_bitplanes=1)
    lut = np.full(O.lut_len(p), 64, np.int32)
    st, region, ncw, tr = O.encode_cb_trace(img, 0, lut, p)
    assert st == 0 and region[0] == 0
    assert len(tr) == 4097
    first = tr[:32]
    assert np.array_equal(first["lane"], np.arange(32)) and (first["site"] == 0).all()
    assert np.array_equal(first["column"], 2 * np.arange(32)) and (first["row"] == 0).all()
    assert np.array_equal(first["slot"], np.arange(1, 33)) and (first["opened"] == 1).all()
    assert first["sym"][0] == 1 and (first["sym"][1:] == 0).all()
    sign = tr[32]
    assert sign["site"] == 1 and sign["lane"] == 0 and sign["sym"] == 1 and sign["opened"] == 0
    # after the 1 (aux = 32767 + 1) the interval is [32768, 32768 + 32767)
    assert sign["L"] == 32768 and sign["S"] == 32767
    odd = tr[33:65]
    assert (odd["column"] == 2 * np.arange(32) + 1).all() and (odd["sym"] == 0).all()
    assert (tr["site"] != 2).sum() == 4097 and (tr["pass"] == 0).all()
    # P = 64 halves the interval: S = 2^16 - 1 reaches 0 after 16 zeros, so lanes
    # 1..31 (128 zeros each) close exactly 8 codewords.  Lane 0 codes a 1 (S becomes
    # 32767), the sign 1 (aux = 16383 + 1, S = 16383), then 127 zeros: 14 close its
    # first codeword, 112 more close 7, and the last zero opens a ninth.
    assert ncw == 31 * 8 + 9
    lane0 = tr[tr["lane"] == 0]
    assert lane0[2]["S"] == 16383 and lane0[2]["L"] == 32768 + 16384
    assert np.array_equal(np.nonzero(lane0["opened"])[0], [0, 16] + [16 + 16 * k for k in range(1, 8)])
