"""Real 1-D plans (syclfft_hip_real_plan_create) on the GPU, through the C ABI.

R2C: batch x N real values -> batch x (N/2 + 1) complex values, the first N/2 + 1 bins of
the complex DFT of the real input.  C2R: the inverse, from a Hermitian half spectrum.  Both
run the complex plan of length N/2 over the real data viewed as complex values plus one
untangle / tangle pass.  Checked against the complex FP64 oracle within c log2(N) u
(c = 4: the extra pass adds two roundings), and through R2C -> C2R round trips."""
import sys

import numpy as np
import pytest

from conftest import oracle_fftn, rel_l2

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
import syclfft_hip as sh  # noqa: E402

U = {False: 2.0 ** -24, True: 2.0 ** -53}


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def tol(n, double):
    return 4 * max(1.0, np.log2(n)) * U[double]


CASES = [(2, 5), (16, 7), (30, 4), (256, 16), (1000, 3), (4096, 8), (8192, 300), (16384, 2), (20014, 2),
         (65536, 2), (1 << 20, 1), (2520, 64)]


@pytest.mark.parametrize("double", [False, True])
@pytest.mark.parametrize("norm", ["none", "forward", "orthogonal"])
@pytest.mark.parametrize("n,batch", CASES)
def test_r2c_vs_oracle(n, batch, norm, double):
    say(f"case r2c n={n} batch={batch} norm={norm} double={double}")
    g = np.random.default_rng(20261200 + n + batch)
    x = g.uniform(-1, 1, n * batch).astype(np.float64 if double else np.float32)
    m = n // 2
    p = sh.Plan([n], batch, sh.FORWARD, double=double, normalization=norm, real=True)
    xd = torch.from_numpy(x).to("cuda")
    yd = torch.empty(batch * (m + 1), dtype=torch.complex128 if double else torch.complex64, device="cuda")
    p.execute(xd, yd)
    torch.cuda.synchronize()
    p.close()
    ref = oracle_fftn(x.astype(np.complex128), [n], batch, -1, norm).reshape(batch, n)[:, : m + 1]
    err = rel_l2(yd.cpu().numpy(), ref)
    say(f"  rel-L2 {err:.3e} (tol {tol(n, double):.1e})")
    assert err <= tol(n, double)


@pytest.mark.parametrize("double", [False, True])
@pytest.mark.parametrize("norm", ["none", "backward", "orthogonal"])
@pytest.mark.parametrize("n,batch", CASES)
def test_c2r_vs_oracle(n, batch, norm, double):
    say(f"case c2r n={n} batch={batch} norm={norm} double={double}")
    g = np.random.default_rng(20261300 + n + batch)
    r = g.uniform(-1, 1, n * batch)
    m = n // 2
    full = oracle_fftn(r.astype(np.complex128), [n], batch, -1, "none").reshape(batch, n)
    half = full[:, : m + 1].copy()
    half[:, 0] = half[:, 0].real
    half[:, m] = half[:, m].real
    half = half.astype(np.complex128 if double else np.complex64)
    herm = np.zeros((batch, n), dtype=np.complex128)
    herm[:, : m + 1] = half
    herm[:, m + 1:] = np.conj(half[:, 1:m][:, ::-1])
    expect = oracle_fftn(herm, [n], batch, 1, norm).real
    p = sh.Plan([n], batch, sh.BACKWARD, double=double, normalization=norm, real=True)
    hd = torch.from_numpy(half.reshape(-1)).to("cuda")
    yd = torch.empty(n * batch, dtype=torch.float64 if double else torch.float32, device="cuda")
    p.execute(hd, yd)
    torch.cuda.synchronize()
    p.close()
    err = rel_l2(yd.cpu().numpy(), expect)
    say(f"  rel-L2 {err:.3e} (tol {tol(n, double):.1e})")
    assert err <= tol(n, double)


@pytest.mark.parametrize("double", [False, True])
@pytest.mark.parametrize("n,batch", [(4096, 32), (1 << 20, 2), (10008, 4)])
def test_real_round_trip(n, batch, double):
    """R2C then C2R with orthogonal normalization returns the input."""
    say(f"case real round trip n={n} batch={batch} double={double}")
    g = np.random.default_rng(20261400 + n)
    x = g.uniform(-1, 1, n * batch).astype(np.float64 if double else np.float32)
    fwd = sh.Plan([n], batch, sh.FORWARD, double=double, normalization="orthogonal", real=True)
    bwd = sh.Plan([n], batch, sh.BACKWARD, double=double, normalization="orthogonal", real=True)
    xd = torch.from_numpy(x).to("cuda")
    sd = torch.empty(batch * (n // 2 + 1), dtype=torch.complex128 if double else torch.complex64, device="cuda")
    zd = torch.empty_like(xd)
    fwd.execute(xd, sd)
    bwd.execute(sd, zd)
    torch.cuda.synchronize()
    err = rel_l2(zd.cpu().numpy(), x)
    fwd.close()
    bwd.close()
    say(f"  rel-L2 {err:.3e}")
    assert err <= 2 * tol(n, double)


def test_real_plan_buffer_checks():
    p = sh.Plan([256], 2, sh.FORWARD, real=True)
    x = torch.zeros(512, dtype=torch.float32, device="cuda")
    host = np.zeros(2 * 129, dtype=np.complex64)
    with pytest.raises(sh.SyclfftError) as e:
        p.execute_ptr(x.data_ptr(), host.ctypes.data, 0, [], 0)
    assert e.value.code == "invalid_pointer"
    with pytest.raises(sh.SyclfftError) as e:
        p.execute_ptr(x.data_ptr(), x.data_ptr(), 0, [], 0)
    assert e.value.code == "invalid_argument"
    assert p.launch_count() == 2
    p.close()


# multi-D real plans: the last axis as 1-D real rows, then the leading axes of the half spectrum
MD_CASES = [([8, 6], 2), ([32, 64], 3), ([12, 10, 8], 1), ([5, 14], 3), ([64, 2048], 2), ([256, 256], 2),
            ([7, 9, 10], 2), ([1000, 30], 1), ([2, 2], 4), ([17, 4096], 1), ([4, 8, 10008], 1),
            ([2048, 256], 2), ([2048, 2048], 1)]


def hermitian(S, lengths):
    """S symmetrized to exact Hermitian symmetry over the transform axes (all but the first)."""
    axes = tuple(range(1, 1 + len(lengths)))
    M = np.conj(np.roll(np.flip(S, axes), 1, axes))
    return 0.5 * (S + M)


@pytest.mark.parametrize("double", [False, True])
@pytest.mark.parametrize("lengths,batch", MD_CASES)
def test_real_nd_vs_oracle(lengths, batch, double):
    """R2C of a real [batch][lengths] array against the first N/2 + 1 bins of the last axis of
    the complex FP64 oracle; C2R of that half spectrum (made exactly Hermitian, rounded to the
    plan's type) against the oracle's inverse of the full spectrum it defines."""
    T = int(np.prod(lengths))
    n = lengths[-1]
    h = n // 2 + 1
    norm = ["none", "orthogonal", "forward", "backward"][(T + batch) % 4]
    say(f"case real nd lengths={lengths} batch={batch} norm={norm} double={double}")
    g = np.random.default_rng(20261500 + T + batch)
    rt, ct = (np.float64, np.complex128) if double else (np.float32, np.complex64)
    x = g.uniform(-1, 1, batch * T).astype(rt)
    shape = (batch, *lengths)
    # R2C
    fwd_norm = norm if norm != "backward" else "none"
    p = sh.Plan(lengths, batch, sh.FORWARD, double=double, normalization=fwd_norm, real=True)
    xd = torch.from_numpy(x).to("cuda")
    yd = torch.empty(batch * T // n * h, dtype=torch.complex128 if double else torch.complex64, device="cuda")
    p.execute(xd, yd)
    torch.cuda.synchronize()
    say(f"  r2c {p.describe().strip()}")
    p.close()
    ref = oracle_fftn(x.astype(np.complex128), lengths, batch, -1, fwd_norm).reshape(shape)[..., :h]
    err = rel_l2(yd.cpu().numpy(), ref)
    say(f"  r2c rel-L2 {err:.3e} (tol {tol(T, double):.1e})")
    assert err <= tol(T, double)
    # C2R
    S = oracle_fftn(g.uniform(-1, 1, batch * T).astype(np.complex128), lengths, batch, -1, "none").reshape(shape)
    F = hermitian(S, lengths)
    F = F.astype(ct).astype(np.complex128)  # rounded: conjugate pairs stay exact conjugates
    half = np.ascontiguousarray(F[..., :h]).astype(ct)
    bwd_norm = norm if norm != "forward" else "none"
    expect = oracle_fftn(F, lengths, batch, 1, bwd_norm).real
    q = sh.Plan(lengths, batch, sh.BACKWARD, double=double, normalization=bwd_norm, real=True)
    hd = torch.from_numpy(half.reshape(-1)).to("cuda")
    zd = torch.empty(batch * T, dtype=torch.float64 if double else torch.float32, device="cuda")
    q.execute(hd, zd)
    torch.cuda.synchronize()
    say(f"  c2r {q.describe().strip()}")
    q.close()
    err = rel_l2(zd.cpu().numpy(), expect)
    say(f"  c2r rel-L2 {err:.3e} (tol {tol(T, double):.1e})")
    assert err <= tol(T, double)


@pytest.mark.parametrize("lengths,batch", [([64, 2048], 2), ([32, 32, 30], 3)])
def test_real_nd_round_trip(lengths, batch):
    """Multi-D R2C then C2R with orthogonal normalization returns the input (fp32)."""
    say(f"case real nd round trip lengths={lengths} batch={batch}")
    T = int(np.prod(lengths))
    h = lengths[-1] // 2 + 1
    x = np.random.default_rng(20261600 + T).uniform(-1, 1, batch * T).astype(np.float32)
    fwd = sh.Plan(lengths, batch, sh.FORWARD, normalization="orthogonal", real=True)
    bwd = sh.Plan(lengths, batch, sh.BACKWARD, normalization="orthogonal", real=True)
    xd = torch.from_numpy(x).to("cuda")
    sd = torch.empty(batch * T // lengths[-1] * h, dtype=torch.complex64, device="cuda")
    zd = torch.empty_like(xd)
    fwd.execute(xd, sd)
    bwd.execute(sd, zd)
    torch.cuda.synchronize()
    fwd.close()
    bwd.close()
    err = rel_l2(zd.cpu().numpy(), x)
    say(f"  rel-L2 {err:.3e}")
    assert err <= 2 * tol(T, False)


def test_real_nd_round_trip_full_size():
    """At the measured size (fp32 2048 x 2048 batch 16, 256 MiB of real values): 2-D R2C then
    C2R with orthogonal normalization returns the input."""
    lengths, batch = [2048, 2048], 16
    say(f"case real nd full-size round trip lengths={lengths} batch={batch}")
    T = 2048 * 2048
    gen = torch.Generator(device="cuda").manual_seed(20261800)
    xd = torch.rand(T * batch, generator=gen, device="cuda") * 2 - 1
    sd = torch.empty(batch * 2048 * 1025, dtype=torch.complex64, device="cuda")
    zd = torch.empty_like(xd)
    with sh.Plan(lengths, batch, sh.FORWARD, normalization="orthogonal", real=True) as f, \
            sh.Plan(lengths, batch, sh.BACKWARD, normalization="orthogonal", real=True) as b:
        f.execute(xd, sd)
        b.execute(sd, zd)
        torch.cuda.synchronize()
    err = float(torch.linalg.vector_norm((zd - xd).double()) / torch.linalg.vector_norm(xd.double()))
    say(f"  rel-L2 {err:.3e}")
    assert err <= 2 * tol(T, False)
