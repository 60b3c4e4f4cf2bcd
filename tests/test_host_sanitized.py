"""The host-only library code (argument checks, batch queue entries, server
admission rule, LUT text parser, host LUT trainer, codeblock files) and the CPU
oracle, built with g++/gcc under AddressSanitizer and UndefinedBehaviorSanitizer
and exercised by tests/cpp/host_selftest.cpp.  CPU only: the device code is not
part of this build.  The reference's own undefined behaviour sites (shift by 32 at
BPC_PaCo.cu:275,311; reads outside TData at :644,654,688,698) have no counterpart
in the oracle, and this build would report one if it had."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SAN = ["-fsanitize=address,undefined", "-fno-sanitize-recover=all", "-fno-omit-frame-pointer", "-g", "-O1"]


@pytest.fixture(scope="module")
def selftest(tmp_path_factory):
    if not shutil.which("g++") or not shutil.which("gcc"):
        pytest.skip("no host compiler")
    d = tmp_path_factory.mktemp("san")
    oracle_o = str(d / "bpc_oracle.o")
    subprocess.run(["gcc", "-std=c99", "-Wall", "-Wextra", "-Werror", "-Wno-unknown-pragmas"] + SAN +
                   ["-c", os.path.join(ROOT, "oracle", "bpc_oracle.c"), "-o", oracle_o], check=True)
    exe = str(d / "host_selftest")
    subprocess.run(["g++", "-std=c++17", "-Wall", "-Wextra", "-Werror"] + SAN +
                   [os.path.join(ROOT, "tests", "cpp", "host_selftest.cpp"),
                    os.path.join(ROOT, "cuda_bpc-paco_amd", "csrc", "bpc_host.cpp"), oracle_o, "-o", exe, "-lpthread"],
                   check=True)
    return exe, d


def test_host_code_and_oracle_under_sanitizers(selftest):
    exe, d = selftest
    files = d / "bin"
    files.mkdir()
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0:halt_on_error=1",
               UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1")
    r = subprocess.run([exe, os.path.join(ROOT, "tests", "golden", "luts"), str(files)], capture_output=True,
                       text=True, env=env, timeout=600)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "0 failed" in r.stdout
    assert "runtime error" not in r.stderr and "AddressSanitizer" not in r.stderr, r.stderr
