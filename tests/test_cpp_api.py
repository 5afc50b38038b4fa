"""C++ plan-API programs (tests/cpp, built by __graft_entry__.build()) and the reference's
own SYCL-free test programs built from its sources (oracle/_ref, built only where the
reference is present).

* span / plugin-ABI / host tests run on the CPU (host::plan through the FFTW3-API host
  provider module, backed by MKL's FFTW3 interface in this image);
* the reference's own tests/host_tests.cpp, compiled against the reference's loader and
  host front-end, loads this repository's syclfft_provider_fftw.so: the host ABI v1 is
  bit-compatible and the module passes the reference's tolerances;
* device_tests (GPU) drive syclfft::plan over the HIP provider module.
"""
import os
import subprocess

import pytest

from conftest import ROOT

BIN = os.path.join(ROOT, "tests", "cpp", "bin")
REF = os.path.join(ROOT, "oracle", "_ref")
PROVIDERS = os.path.join(ROOT, "sycl-fft_amd", "lib", "syclfft", "providers")


def run(exe, env_extra=None, timeout=300):
    if not os.path.exists(exe):
        pytest.fail(f"{exe} not built (run __graft_entry__.build())")
    env = dict(os.environ)
    env.update(env_extra or {})
    r = subprocess.run([exe], capture_output=True, text=True, timeout=timeout, env=env)
    assert r.returncode == 0, r.stdout + r.stderr
    return r.stdout


def fftw_backend_available():
    import ctypes
    so = os.path.join(PROVIDERS, "syclfft_provider_fftw.so")
    if not os.path.exists(so):
        return False
    lib = ctypes.CDLL(so)
    lib.syclfft_fftw_provider_backend.restype = ctypes.c_char_p
    return bool(lib.syclfft_fftw_provider_backend())


def test_span():
    assert "passed" in run(os.path.join(BIN, "span_tests"))


def test_plugin_abi_mismatch_rejected():
    assert "passed" in run(os.path.join(BIN, "plugin_tests"), {"SYCLFFT_PLUGIN_PATH": os.path.join(BIN, "bad")})


def test_host_plans():
    if not fftw_backend_available():
        pytest.skip("no FFTW3-API library (libfftw3 or libmkl_rt) on this host")
    assert "passed" in run(os.path.join(BIN, "host_tests"))


@pytest.mark.skipif(not os.path.exists(os.path.join(REF, "ref_host_tests")),
                    reason="reference sources not present (oracle/_ref is built only where they are)")
def test_reference_host_tests_against_our_host_provider():
    if not fftw_backend_available():
        pytest.skip("no FFTW3-API library on this host")
    assert "passed" in run(os.path.join(REF, "ref_host_tests"), {"SYCLFFT_PLUGIN_PATH": PROVIDERS})


@pytest.mark.skipif(not os.path.exists(os.path.join(REF, "ref_plugin_tests")),
                    reason="reference sources not present")
def test_reference_plugin_test_against_bad_module():
    run(os.path.join(REF, "ref_plugin_tests"), {"SYCLFFT_PLUGIN_PATH": os.path.join(REF, "bad")})


@pytest.mark.skipif(not os.path.exists(os.path.join(REF, "ours_host_tests")),
                    reason="reference sources not present")
def test_reference_host_tests_against_our_front_end():
    """The reference's tests/host_tests.cpp compiled against this repository's host::plan
    front-end and provider loader (include/, sycl-fft_amd/lib)."""
    if not fftw_backend_available():
        pytest.skip("no FFTW3-API library on this host")
    assert "passed" in run(os.path.join(REF, "ours_host_tests"), {"SYCLFFT_PLUGIN_PATH": PROVIDERS})


@pytest.mark.skipif(not os.path.exists(os.path.join(REF, "ours_plugin_tests")),
                    reason="reference sources not present")
def test_reference_plugin_test_against_our_loader():
    run(os.path.join(REF, "ours_plugin_tests"), {"SYCLFFT_PLUGIN_PATH": os.path.join(REF, "bad")})


@pytest.mark.gpu
def test_device_plan_api():
    assert "passed" in run(os.path.join(BIN, "device_tests"), timeout=120)
