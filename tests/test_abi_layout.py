"""The Python mirror's ctypes structures (cuda_bpc-paco_amd/bpcpaco.py) have the
C compiler's layout of include/bpc_paco.h: every size and every field offset,
printed by tests/c/abi_layout.c.  CPU only (no library call)."""
import ctypes
import json
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def c_layout(tmp_path_factory):
    if not shutil.which("gcc"):
        pytest.skip("no C compiler")
    exe = str(tmp_path_factory.mktemp("abi") / "abi_layout")
    subprocess.run(["gcc", "-std=c99", "-Wall", "-Wextra", "-Werror", "-I", os.path.join(ROOT, "include"),
                    os.path.join(ROOT, "tests", "c", "abi_layout.c"), "-o", exe], check=True)
    return json.loads(subprocess.run([exe], capture_output=True, text=True, check=True).stdout)


def test_ctypes_structs_match_header(c_layout, bpc):
    mirror = {"bpc_params": bpc.Params, "bpc_image": bpc.Image, "bpc_server_config": bpc.ServerConfig,
              "bpc_request": bpc.Request, "bpc_result": bpc.Result, "bpc_server_stats_t": bpc.ServerStats}
    assert set(mirror) == set(c_layout)
    for name, cls in mirror.items():
        want = c_layout[name]
        assert ctypes.sizeof(cls) == want["sizeof"], name
        names = [f[0] for f in cls._fields_]
        assert names == list(want["fields"]), name
        for field, off in want["fields"].items():
            assert getattr(cls, field).offset == off, (name, field)
