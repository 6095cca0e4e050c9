"""The C++ drop-in header (include/ntt.cuh) builds against the library with a
plain host compiler, and a caller written like the reference's
tests/test_ntt.cu passes on the GPU."""
import os
import subprocess

import pytest

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cpp")
BIN = os.path.join(HERE, "test_ntt_dropin")


def test_dropin_header_compiles_and_links():
    subprocess.check_call(["make", "-s", "-C", HERE])
    assert os.path.exists(BIN)


@pytest.mark.gpu
def test_dropin_cpp_caller_passes():
    if not os.path.exists(BIN):
        subprocess.check_call(["make", "-s", "-C", HERE])
    r = subprocess.run([BIN], capture_output=True, text=True, timeout=600)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "FAILED" not in r.stdout
