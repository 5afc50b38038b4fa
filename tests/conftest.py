"""Shared pytest setup: markers, paths to the provider binding and the CPU oracle."""
import ctypes
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "sycl-fft_amd")
ORACLE_SO = os.path.join(ROOT, "oracle", "liboracle.so")
if PKG not in sys.path:
    sys.path.insert(0, PKG)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


_oracle = None


def oracle_lib():
    """The CPU oracle (test infrastructure, oracle/oracle.c)."""
    global _oracle
    if _oracle is None:
        if not os.path.exists(ORACLE_SO):
            raise OSError(f"oracle not built: {ORACLE_SO} (run make -C oracle)")
        lib = ctypes.CDLL(ORACLE_SO)
        u64p = ctypes.POINTER(ctypes.c_uint64)
        dp = ctypes.c_void_p
        lib.oracle_dft_ld.argtypes = [ctypes.c_uint64, dp, dp, ctypes.c_int]
        lib.oracle_fft_ld.argtypes = [ctypes.c_uint64, dp, dp, ctypes.c_int]
        lib.oracle_fftn_ld.argtypes = [ctypes.c_int, u64p, ctypes.c_uint64, dp, dp, ctypes.c_int, ctypes.c_int]
        lib.oracle_norm_scale.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_uint64]
        lib.oracle_norm_scale.restype = ctypes.c_double
        lib.ref_portable_exec_f32.argtypes = [ctypes.c_int, u64p, ctypes.c_uint64, ctypes.c_int, ctypes.c_int, dp, dp]
        lib.ref_portable_exec_f64.argtypes = [ctypes.c_int, u64p, ctypes.c_uint64, ctypes.c_int, ctypes.c_int, dp, dp]
        lib.oracle_max_threads.restype = ctypes.c_int
        lib.oracle_set_threads.argtypes = [ctypes.c_int]
        _oracle = lib
    return _oracle


NORM_ID = {"none": 0, "forward": 1, "backward": 2, "orthogonal": 3}


def oracle_fftn(x: np.ndarray, lengths, batch, direction=-1, norm="none") -> np.ndarray:
    """FP64 oracle of a batched C-order transform; x is complex, any shape of the right size."""
    lib = oracle_lib()
    xin = np.ascontiguousarray(x, dtype=np.complex128).reshape(-1)
    out = np.empty_like(xin)
    lens = (ctypes.c_uint64 * 3)(*[int(n) for n in lengths], *([1] * (3 - len(lengths))))
    lib.oracle_fftn_ld(len(lengths), lens, int(batch), xin.ctypes.data, out.ctypes.data,
                       int(direction), NORM_ID[norm])
    return out


def oracle_dft(x: np.ndarray, direction=-1) -> np.ndarray:
    lib = oracle_lib()
    xin = np.ascontiguousarray(x, dtype=np.complex128).reshape(-1)
    out = np.empty_like(xin)
    lib.oracle_dft_ld(xin.size, xin.ctypes.data, out.ctypes.data, int(direction))
    return out


def ref_portable(x: np.ndarray, lengths, batch, direction=-1, norm="none", double=False) -> np.ndarray:
    """Restatement of the reference portable provider in its own Scalar precision."""
    lib = oracle_lib()
    dt = np.complex128 if double else np.complex64
    xin = np.ascontiguousarray(x, dtype=dt).reshape(-1)
    out = np.empty_like(xin)
    lens = (ctypes.c_uint64 * 3)(*[int(n) for n in lengths], *([1] * (3 - len(lengths))))
    fn = lib.ref_portable_exec_f64 if double else lib.ref_portable_exec_f32
    rc = fn(len(lengths), lens, int(batch), int(direction), NORM_ID[norm], xin.ctypes.data, out.ctypes.data)
    assert rc == 0
    return out


def rel_l2(a, b) -> float:
    a = np.asarray(a).reshape(-1).astype(np.complex128)
    b = np.asarray(b).reshape(-1).astype(np.complex128)
    d = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / (d if d > 0 else 1.0))


def ref_close(actual, expected, tol) -> bool:
    """The reference tests' element check |a - e| <= tol (1 + |e|) (tests/sycl_tests.cpp:186-191)."""
    a = np.asarray(actual).reshape(-1).astype(np.complex128)
    e = np.asarray(expected).reshape(-1).astype(np.complex128)
    return bool(np.all(np.abs(a - e) <= tol * (1.0 + np.abs(e))))


def gpu_available() -> bool:
    try:
        import torch
        return torch.cuda.is_available()
    except Exception:
        return False


@pytest.fixture(scope="session")
def oracle():
    return oracle_lib()
