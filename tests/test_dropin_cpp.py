"""The C++ drop-in headers (include/ntt.cuh, include/ntt_arbitrary.cuh) build
against the library with a plain host compiler, and callers written like the
reference's tests/test_ntt.cu and run_for_K (include/ntt_arbitrary.cuh:2808-2906)
pass on the GPU."""
import os
import subprocess

import pytest

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cpp")
BIN = os.path.join(HERE, "test_ntt_dropin")
ARB = os.path.join(HERE, "test_arb_dropin")


def _build():
    subprocess.check_call(["make", "-s", "-C", HERE])


def test_dropin_headers_compile_and_link():
    _build()
    assert os.path.exists(BIN) and os.path.exists(ARB)


def test_arb_header_host_arithmetic():
    """BigInt, mod_*, mont_np/mont_R2/mont_mul_big and FieldConfig root orders (no GPU)."""
    if not os.path.exists(ARB):
        _build()
    r = subprocess.run([ARB, "--host"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0 and "ALL PASSED" in r.stdout, r.stdout + r.stderr


@pytest.mark.gpu
def test_dropin_cpp_caller_passes():
    if not os.path.exists(BIN):
        _build()
    r = subprocess.run([BIN], capture_output=True, text=True, timeout=600)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "FAILED" not in r.stdout


@pytest.mark.gpu
def test_arb_dropin_caller_passes():
    if not os.path.exists(ARB):
        _build()
    r = subprocess.run([ARB], capture_output=True, text=True, timeout=600)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "ALL PASSED" in r.stdout and "FAILED" not in r.stdout
