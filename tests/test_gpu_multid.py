"""Full-size multidimensional transforms of the multi-D configuration (3-D 256^3
complex<float> and 2-D 4096^2 complex<double>, in place, forward normalization, C order),
checked through size-independent properties: the input is a sum of 8 seeded complex
exponentials with integer frequency vectors plus N(0, 0.01) noise, so the normalized
spectrum must show exactly those 8 peaks with the tones' amplitudes; a backward transform
(normalization none) must return the input; Parseval must hold.
"""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
import syclfft_hip as sh  # noqa: E402

U = {False: 2.0 ** -24, True: 2.0 ** -53}


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def tones_input(shape, double, seed):
    """x = sum_t a_t exp(2 pi i <k_t, j / shape>) + noise; returns (x, freqs, amps)."""
    g = np.random.default_rng(seed)
    freqs = []
    while len(freqs) < 8:
        k = tuple(int(g.integers(n)) for n in shape)
        if k not in freqs:
            freqs.append(k)
    amps = np.exp(2j * np.pi * g.random(8))  # unit modulus, random phase
    rdt = torch.float64 if double else torch.float32
    cdt = torch.complex128 if double else torch.complex64
    grids = torch.meshgrid(*[torch.arange(n, dtype=torch.float64, device="cuda") / n for n in shape],
                           indexing="ij")
    x = torch.zeros(shape, dtype=torch.complex128, device="cuda")
    for k, a in zip(freqs, amps):
        ph = sum(2 * np.pi * kk * gg for kk, gg in zip(k, grids))
        x += complex(a) * torch.exp(1j * ph)
    gen = torch.Generator(device="cuda")
    gen.manual_seed(seed)
    noise = torch.randn(shape + (2,), generator=gen, dtype=torch.float64, device="cuda") * 0.01
    x += torch.view_as_complex(noise)
    del grids, noise
    return x.to(cdt).contiguous(), freqs, amps


@pytest.mark.parametrize("measure", [False, True], ids=["estimate", "measure"])
@pytest.mark.parametrize("shape,double", [((256, 256, 256), False), ((4096, 4096), True)],
                         ids=["256^3_f32", "4096^2_f64"])
def test_multid_tone_peaks_roundtrip(shape, double, measure):
    """measure=True builds the plans bench.py times (planning_mode::measure)."""
    n = int(np.prod(shape))
    say(f"case multi-D {shape} double={double} measure={measure}")
    x, freqs, amps = tones_input(shape, double, 20260003)
    x0 = x.clone()
    with sh.Plan(list(shape), 1, sh.FORWARD, double=double, in_place=True, normalization="forward",
                 measure=measure) as fwd:
        say("plan " + fwd.describe().strip().replace("\n", " | "))
        fwd.execute(x)
    torch.cuda.synchronize()
    # peaks: the 8 tone bins carry the amplitudes, every other bin only noise
    mag = x.abs().reshape(-1)
    idx = [int(np.ravel_multi_index(k, shape)) for k in freqs]
    got = x.reshape(-1)[idx].cpu().numpy().astype(np.complex128)
    assert np.max(np.abs(got - amps)) < 1e-3, np.abs(got - amps)
    mag[idx] = 0
    assert float(mag.max()) < 1e-3
    # Parseval with the forward normalization: sum |X|^2 = sum |x|^2 / N
    ex = float((x0.abs() ** 2).sum()) / n
    eX = float((x.abs() ** 2).sum())
    assert abs(eX - ex) / ex < 1e-5
    # round trip
    with sh.Plan(list(shape), 1, sh.BACKWARD, double=double, in_place=True, normalization="none") as bwd:
        bwd.execute(x)
    torch.cuda.synchronize()
    err = float(torch.linalg.vector_norm(x - x0) / torch.linalg.vector_norm(x0))
    tol = 3 * 2 * 2 * np.log2(n) * U[double]
    say(f"round-trip rel-L2 {err:.3e} (tol {tol:.1e})")
    assert err <= tol
