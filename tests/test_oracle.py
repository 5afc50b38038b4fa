"""CPU tests pinning the oracle (oracle/oracle.c) before it is trusted as the checker.

Each function of the oracle is pinned by something independent of its own formula:
known answers (impulse, shifted impulse, tone, constant), the textbook O(n^2) DFT at
exact integer exponents, numpy's pocketfft (committed golden fixtures, written with
scipy in long double where available), and invariants (linearity, Parseval, round trip,
separability).  A planted mistake in the oracle (wrong sign, dropped twiddle, transposed
axis) fails at least one of these.
"""
import os

import numpy as np
import pytest

from conftest import ROOT, oracle_dft, oracle_fftn, ref_close, ref_portable, rel_l2

GOLDEN = os.path.join(ROOT, "tests", "golden", "fft_golden.npz")


def rnd(n, seed=0):
    g = np.random.default_rng(seed)
    return g.uniform(-1, 1, n) + 1j * g.uniform(-1, 1, n)


# ---------------------------------------------------------------- known answers

@pytest.mark.parametrize("n", [1, 2, 3, 7, 8, 11, 13, 60, 64])
def test_dft_impulse_is_constant(n):
    x = np.zeros(n, complex)
    x[0] = 1
    y = oracle_dft(x)
    assert np.allclose(y, np.ones(n), atol=1e-15)


@pytest.mark.parametrize("n,s", [(8, 1), (12, 5), (13, 3)])
def test_dft_shifted_impulse_is_twiddle_row(n, s):
    x = np.zeros(n, complex)
    x[s] = 1
    k = np.arange(n)
    for sign in (-1, 1):
        y = oracle_dft(x, sign)
        assert np.allclose(y, np.exp(sign * 2j * np.pi * s * k / n), atol=1e-14)


@pytest.mark.parametrize("n,f", [(16, 3), (15, 4), (17, 5)])
def test_dft_tone_is_delta(n, f):
    j = np.arange(n)
    x = np.exp(2j * np.pi * f * j / n)
    y = oracle_dft(x, -1)
    e = np.zeros(n, complex)
    e[f] = n
    assert np.allclose(y, e, atol=1e-12)


def test_dft_matches_numpy_small():
    for n in (5, 6, 9, 16, 31, 100):
        x = rnd(n, n)
        assert rel_l2(oracle_dft(x, -1), np.fft.fft(x)) < 1e-14
        assert rel_l2(oracle_dft(x, 1), np.fft.ifft(x) * n) < 1e-14


# ---------------------------------------------------------------- fast oracle

@pytest.mark.parametrize("n", [1, 2, 8, 12, 60, 64, 97, 128, 210, 251, 1000, 1009, 2048, 2520])
def test_fast_matches_bruteforce(n):
    x = rnd(n, n + 1)
    for sign in (-1, 1):
        a = oracle_fftn(x, [n], 1, sign)
        b = oracle_dft(x, sign)
        assert rel_l2(a, b) < 1e-15 * max(10, np.log2(max(n, 2)) * 4)


@pytest.mark.parametrize("n", [4096, 6561, 10007, 65537, 2 ** 16])
def test_fast_matches_numpy_large(n):
    x = rnd(n, 7)
    assert rel_l2(oracle_fftn(x, [n], 1, -1), np.fft.fft(x)) < 1e-13


def test_linearity_and_parseval():
    n = 1013
    x, y = rnd(n, 1), rnd(n, 2)
    a, b = 0.3 - 0.2j, -1.7 + 0.5j
    fx, fy = oracle_fftn(x, [n], 1), oracle_fftn(y, [n], 1)
    assert rel_l2(oracle_fftn(a * x + b * y, [n], 1), a * fx + b * fy) < 1e-15
    assert abs(np.sum(abs(fx) ** 2) - n * np.sum(abs(x) ** 2)) / (n * np.sum(abs(x) ** 2)) < 1e-14


def test_roundtrip_identity():
    for n in (13, 360, 4099):
        x = rnd(n, n)
        y = oracle_fftn(oracle_fftn(x, [n], 1, -1), [n], 1, 1, "backward")
        assert rel_l2(y, x) < 1e-15


def test_multid_matches_numpy_and_separability():
    for shape, batch in (((3, 4), 2), ((2, 3, 5), 1), ((11, 3), 1), ((16, 12, 10), 2)):
        n = int(np.prod(shape)) * batch
        x = rnd(n, 3).reshape((batch,) + shape)
        ax = tuple(range(1, len(shape) + 1))
        assert rel_l2(oracle_fftn(x, shape, batch, -1), np.fft.fftn(x, axes=ax)) < 1e-14
        assert rel_l2(oracle_fftn(x, shape, batch, 1), np.fft.ifftn(x, axes=ax) * np.prod(shape)) < 1e-14


@pytest.mark.parametrize("norm", ["none", "forward", "backward", "orthogonal"])
@pytest.mark.parametrize("sign", [-1, 1])
def test_normalization_scale(norm, sign):
    n = 12
    x = rnd(n, 4)
    raw = oracle_fftn(x, [n], 1, sign)
    s = {"none": 1.0, "forward": 1 / n if sign == -1 else 1.0,
         "backward": 1 / n if sign == 1 else 1.0, "orthogonal": 1 / np.sqrt(n)}[norm]
    assert rel_l2(oracle_fftn(x, [n], 1, sign, norm), raw * s) < 1e-15


def test_golden_fixtures():
    """Vectors written by tests/golden/make_golden.py (numpy pocketfft, float64)."""
    g = np.load(GOLDEN, allow_pickle=False)
    names = sorted({k.rsplit("_", 1)[0] for k in g.files})
    assert names
    for name in names:
        x, y, meta = g[name + "_in"], g[name + "_out"], g[name + "_meta"]
        rank = int(meta[0])
        shape = [int(v) for v in meta[1:1 + rank]]
        batch, sign, normid = int(meta[4]), int(meta[5]), int(meta[6])
        norm = ["none", "forward", "backward", "orthogonal"][normid]
        got = oracle_fftn(x, shape, batch, sign, norm)
        assert rel_l2(got, y) < 1e-13, name


# ---------------------------------------------------------------- reference restatement

REF_TOL_1D = {False: 4e-4, True: 2e-11}      # tests/sycl_tests.cpp:242
REF_TOL_MANY = {False: 8e-4, True: 5e-11}    # tests/sycl_tests.cpp:186


def _ref_input_1d(n, double):
    i = np.arange(n)
    x = ((i * 5 + 1) % 9) / 4 + 1j * (((i * 3 + 2) % 7) / 5)   # tests/sycl_tests.cpp:200-205
    return x.astype(np.complex128 if double else np.complex64)


def _ref_input_many(count, double):
    i = np.arange(count)
    x = ((i * 11 + 2) % 17) / 8 + 1j * (((i * 3 + 1) % 13) / 7)  # tests/sycl_tests.cpp:160-165
    return x.astype(np.complex128 if double else np.complex64)


@pytest.mark.parametrize("n,sign,norm,double", [
    (8, -1, "none", False), (8, -1, "forward", False), (8, -1, "backward", False),
    (8, -1, "orthogonal", False), (7, 1, "none", False), (6, -1, "none", False),
    (15, 1, "none", False), (35, -1, "none", False), (11, 1, "none", False),
    (11, -1, "orthogonal", True)])
def test_restatement_matches_reference_cases(n, sign, norm, double):
    """The reference's own 1-D cases (tests/sycl_tests.cpp:318-370), same inputs/tolerances."""
    x = _ref_input_1d(n, double)
    got = ref_portable(x, [n], 1, sign, norm, double)
    exp = oracle_fftn(x, [n], 1, sign, norm)
    assert ref_close(got, exp, REF_TOL_1D[double])


@pytest.mark.parametrize("shape,batch", [((3, 4), 2), ((2, 3, 5), 1), ((11, 3), 1)])
def test_restatement_matches_reference_many(shape, batch):
    count = int(np.prod(shape)) * batch
    x = _ref_input_many(count, False)
    got = ref_portable(x, shape, batch, -1, "none", False)
    exp = oracle_fftn(x, shape, batch, -1, "none")
    assert ref_close(got, exp, REF_TOL_MANY[False])


@pytest.mark.parametrize("n,double", [(4096, False), (2520, True), (6561, True), (1024, True), (360, False)])
def test_restatement_paths_larger(n, double):
    x = np.tile(rnd(n, 11), 2).astype(np.complex128 if double else np.complex64)
    err = rel_l2(ref_portable(x, [n], 2, -1, "none", double), oracle_fftn(x, [n], 2, -1))
    assert err < (1e-14 if double else 2e-6) * np.log2(n)


def test_restatement_bluestein_fp32_error_is_the_references():
    """The reference's fp32 chirp uses an unreduced k^2 in float (src/sycl/sycl.cpp:548-553),
    so its Bluestein error grows with n: rel-L2 ~1e-3 at n = 10007, far above the FP64
    oracle's; this is why the FP64 oracle, not the reference, is the accuracy bar there."""
    n = 10007
    x = rnd(n, 5).astype(np.complex64)
    err = rel_l2(ref_portable(x, [n], 1, -1, "none", False), oracle_fftn(x, [n], 1, -1))
    assert 1e-5 < err < 1e-1
    err64 = rel_l2(ref_portable(x.astype(np.complex128), [n], 1, -1, "none", True), oracle_fftn(x, [n], 1, -1))
    assert err64 < 1e-10


def test_restatement_roundtrip_13():
    """Round trip n = 13 with backward normalization (tests/sycl_tests.cpp:372-399)."""
    i = np.arange(13)
    x = (i / 7 + 1j * ((i % 3) / 5)).astype(np.complex64)
    f = ref_portable(x, [13], 1, -1, "none", False)
    b = ref_portable(f, [13], 1, 1, "backward", False)
    assert np.all(np.abs(b - x) < 2e-3)
