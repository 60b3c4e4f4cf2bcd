import json
import lzma
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "cuda_bpc-paco_amd")
ORACLE = os.path.join(ROOT, "oracle")
GOLDEN = os.path.join(ROOT, "tests", "golden")
for p in (PKG, ORACLE):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")


def _load_oracle():
    import oracle
    if not os.path.exists(oracle._LIB_PATH):
        oracle.build()
    return oracle


@pytest.fixture(scope="session")
def O():
    return _load_oracle()


@pytest.fixture(scope="session")
def sample_image():
    with open(os.path.join(GOLDEN, "sample_image.i16.xz"), "rb") as f:
        raw = lzma.decompress(f.read())
    return np.frombuffer(raw, dtype="<i2").reshape(2048, 2048).astype(np.int32)


@pytest.fixture(scope="session")
def sample_headers():
    with open(os.path.join(GOLDEN, "sample_headers.json")) as f:
        return json.load(f)


def lut_folder(cpass3):
    return os.path.join(GOLDEN, "luts", "%d_coding_pass" % (3 if cpass3 else 2)) + "/"


@pytest.fixture(scope="session")
def bpc():
    import bpcpaco
    bpcpaco.lib()
    return bpcpaco
