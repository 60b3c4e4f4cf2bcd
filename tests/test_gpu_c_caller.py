"""The drop-in boundary called from C: tests/c/bpc_main (built against
include/bpc_paco.h and linked to libbpc_paco.so) reproduces the reference's
example main (main_example.cu:452-548) on its sample image with its shipped
probability tables, in both pass modes.  Its codestream must equal the oracle's,
whose digest is committed (tests/golden/sample_codestream.json), and its decode
must give back the image."""
import hashlib
import json
import os
import subprocess

import numpy as np
import pytest

from conftest import GOLDEN, ROOT, lut_folder

pytestmark = pytest.mark.gpu

EXE = os.path.join(ROOT, "tests", "c", "bpc_main")


@pytest.mark.parametrize("cpass3", [0, 1])
def test_c_main_round_trip_and_golden(O, sample_image, tmp_path, cpass3):
    if not os.path.exists(EXE):
        raise RuntimeError("tests/c/bpc_main is not built (make -C tests/c)")
    with open(os.path.join(GOLDEN, "sample_codestream.json")) as f:
        golden = json.load(f)["%d_coding_pass" % (3 if cpass3 else 2)]
    olut, nbp = O.lut_load_text(lut_folder(cpass3), cpass3)
    st, ostream, ncw = O.encode(sample_image, olut, O.params(cpass3=cpass3, lut_n_bitplanes=nbp))
    assert st == 0
    assert hashlib.sha256(ostream.astype("<i4").tobytes()).hexdigest() == golden["sha256"]
    img_path, want_path, got_path = (str(tmp_path / n) for n in ("image.i32", "expected.i32", "got.i32"))
    sample_image.astype("<i4").tofile(img_path)
    ostream.astype("<i4").tofile(want_path)
    r = subprocess.run([EXE, img_path, "2048", lut_folder(cpass3), str(cpass3), want_path, got_path],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "TEST ENCODER OK" in r.stdout and "TEST DECODER 2048x2048 OK" in r.stdout
    got = np.fromfile(got_path, dtype="<i4")
    assert hashlib.sha256(got.tobytes()).hexdigest() == golden["sha256"]
