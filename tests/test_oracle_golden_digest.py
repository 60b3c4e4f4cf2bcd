"""The oracle's codestream for the reference's sample image with its shipped
probability tables (the reference's own example, main_example.cu:452-548) is
pinned by committed digests (tests/golden/make_digests.py), so that any change of
the oracle's output shows here, on the CPU."""
import hashlib
import json
import os

import pytest

from conftest import GOLDEN, lut_folder


@pytest.mark.parametrize("cpass3", [0, 1])
def test_oracle_sample_codestream_digest(O, sample_image, cpass3):
    with open(os.path.join(GOLDEN, "sample_codestream.json")) as f:
        golden = json.load(f)["%d_coding_pass" % (3 if cpass3 else 2)]
    lut, nbp = O.lut_load_text(lut_folder(cpass3), cpass3)
    assert nbp == golden["lut_n_bitplanes"]
    st, cs, ncw = O.encode(sample_image, lut, O.params(cpass3=cpass3, lut_n_bitplanes=nbp))
    assert st == 0 and int(ncw.sum()) == golden["codewords"]
    assert hashlib.sha256(cs.astype("<i4").tobytes()).hexdigest() == golden["sha256"]
