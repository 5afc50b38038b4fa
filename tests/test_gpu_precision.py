"""GPU parity of the split-precision MFMA plans (fp32, precision split_bf16 / split_fp16)
against the FP64 oracle, through the provider's C ABI.

The radix-16 passes of these plans run as 16-point DFT matrix products on MFMA
(v_mfma_f32_16x16x32_bf16 / _f16) with hi + lo operands and FP32 accumulation; every other
pass is the VALU kernel's.  Stated errors (relative L2 against the FP64 oracle):
  split_fp16 (scaled FP16 hi/lo, 22 operand bits): rel-L2 <= 1e-6, the fp32 accuracy target
  split_bf16 (BF16 hi/lo, 16 operand bits):        rel-L2 <= 2e-5
on uniform[-1,1) inputs and on inputs with log-uniform magnitudes 1e-4..1e4.
"""
import sys

import numpy as np
import pytest

from conftest import oracle_fftn, rel_l2

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
import syclfft_hip as sh  # noqa: E402

TOL = {"split_fp16": 1e-6, "split_bf16": 2e-5}


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def uniform(n, seed):
    g = np.random.default_rng(seed)
    return g.uniform(-1, 1, n) + 1j * g.uniform(-1, 1, n)


def log_uniform(n, seed):
    g = np.random.default_rng(seed)
    mag = 10.0 ** g.uniform(-4, 4, n)
    return mag * np.exp(2j * np.pi * g.uniform(0, 1, n))


def run(x, lengths, batch, direction, norm, precision, in_place=False):
    say(f"case lengths={list(lengths)} batch={batch} dir={direction} norm={norm} "
        f"precision={precision} in_place={in_place}")
    p = sh.Plan(lengths, batch, direction, in_place=in_place, normalization=norm, precision=precision)
    desc = p.describe()
    xd = torch.from_numpy(np.ascontiguousarray(x).astype(np.complex64)).to("cuda")
    if in_place:
        p.execute(xd)
        out = xd
    else:
        out = torch.empty_like(xd)
        p.execute(xd, out)
    torch.cuda.synchronize()
    r = out.cpu().numpy()
    p.close()
    return r, desc


CASES = [
    ([16], 64, -1, "none"),
    ([256], 32, 1, "orthogonal"),
    ([4096], 8, -1, "none"),
    ([4096], 4, 1, "backward"),
    ([16384], 4, -1, "none"),
    ([16384], 2, 1, "forward"),
    ([65536], 2, -1, "none"),
    ([1 << 20], 1, -1, "orthogonal"),
    ([256, 256], 1, 1, "none"),
    ([64, 4096], 1, -1, "forward"),
]


@pytest.mark.parametrize("precision", ["split_fp16", "split_bf16"])
@pytest.mark.parametrize("lengths,batch,direction,norm", CASES)
def test_split_precision_vs_fp64_oracle(lengths, batch, direction, norm, precision):
    n = int(np.prod(lengths))
    x = uniform(n * batch, 20260300 + n + batch).astype(np.complex64)
    y, desc = run(x, lengths, batch, direction, norm, precision)
    ref = oracle_fftn(x.astype(np.complex128), lengths, batch, direction, norm)
    err = rel_l2(y, ref)
    say(f"  rel-L2 {err:.3e} (tol {TOL[precision]:.0e})")
    assert "split=" in desc, desc  # the radix-16 passes ran on MFMA
    assert err <= TOL[precision], (err, desc)


@pytest.mark.parametrize("precision", ["split_fp16", "split_bf16"])
@pytest.mark.parametrize("n,batch", [(16384, 4), (1 << 20, 1)])
def test_split_precision_log_uniform_magnitudes(n, batch, precision):
    """Inputs spanning eight decades of magnitude: the per-butterfly FP16 scaling keeps the
    error relative to each butterfly's largest component."""
    x = log_uniform(n * batch, 20260400 + n).astype(np.complex64)
    y, desc = run(x, [n], batch, -1, "none", precision)
    ref = oracle_fftn(x.astype(np.complex128), [n], batch, -1, "none")
    err = rel_l2(y, ref)
    say(f"  rel-L2 {err:.3e} (tol {TOL[precision]:.0e})")
    assert err <= TOL[precision], (err, desc)


@pytest.mark.parametrize("precision", ["split_fp16", "split_bf16"])
def test_split_precision_in_place_round_trip(precision):
    n, batch = 4096, 16
    x = uniform(n * batch, 20260500).astype(np.complex64)
    xd = torch.from_numpy(x).to("cuda")
    fwd = sh.Plan([n], batch, sh.FORWARD, in_place=True, normalization="orthogonal", precision=precision)
    bwd = sh.Plan([n], batch, sh.BACKWARD, in_place=True, normalization="orthogonal", precision=precision)
    fwd.execute(xd)
    bwd.execute(xd)
    torch.cuda.synchronize()
    err = rel_l2(xd.cpu().numpy(), x)
    fwd.close()
    bwd.close()
    say(f"round trip rel-L2 {err:.3e}")
    assert err <= 2 * TOL[precision]


def test_split_precision_is_deterministic_and_matches_valu_closely():
    """Two executions give identical bits; the fp16 split result differs from the plain
    FP32 VALU result by no more than both differ from the FP64 oracle."""
    n, batch = 16384, 4
    x = uniform(n * batch, 20260600).astype(np.complex64)
    a, _ = run(x, [n], batch, -1, "none", "split_fp16")
    b, _ = run(x, [n], batch, -1, "none", "split_fp16")
    v, _ = run(x, [n], batch, -1, "none", "valu")
    assert np.array_equal(a, b)
    ref = oracle_fftn(x.astype(np.complex128), [n], batch, -1, "none")
    assert rel_l2(a, v) <= rel_l2(a, ref) + rel_l2(v, ref)


@pytest.mark.parametrize("precision", ["split_fp16", "split_bf16"])
def test_split_precision_measure_mode_keeps_mfma(precision):
    """planning_mode::measure may pick another tile shape or variant, but never one that
    drops the MFMA passes (e.g. a larger variant without radix-16 passes)."""
    n, batch = 16384, 256
    say(f"case measure-mode split plan n={n} batch={batch} precision={precision}")
    p = sh.Plan([n], batch, sh.FORWARD, precision=precision, measure=True)
    desc = p.describe()
    x = uniform(n * batch, 20260700).astype(np.complex64)
    xd = torch.from_numpy(x).to("cuda")
    out = torch.empty_like(xd)
    p.execute(xd, out)
    torch.cuda.synchronize()
    p.close()
    assert "split=" in desc, desc
    ref = oracle_fftn(x[: 2 * n].astype(np.complex128), [n], 2, -1, "none")
    assert rel_l2(out[: 2 * n].cpu().numpy(), ref) <= TOL[precision]
