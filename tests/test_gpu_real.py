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


CASES = [(2, 5), (16, 7), (30, 4), (256, 16), (1000, 3), (4096, 8), (16384, 2), (20014, 2), (65536, 2),
         (1 << 20, 1), (2520, 64)]


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
