"""GPU parity of the HIP provider (through its C ABI) against the FP64 oracle and the
restatement of the reference portable provider, on seeded inputs.

Tolerances: per-element |a - e| <= tol (1 + |e|) with the reference's own tolerances
(tests/sycl_tests.cpp:186, 242) for the reference's cases, and a relative L2 bound
c * log2(N) * u (u = 2^-24 for float, 2^-53 for double, c = 2) against the FP64 oracle
for everything else.
"""
import sys

import numpy as np
import pytest

from conftest import oracle_fftn, ref_close, ref_portable, rel_l2

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
import syclfft_hip as sh  # noqa: E402

U = {False: 2.0 ** -24, True: 2.0 ** -53}


def tol_l2(n, double, c=2.0):
    return c * max(1.0, np.log2(max(n, 2))) * U[double]


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def dev_tensor(x, double):
    dt = torch.complex128 if double else torch.complex64
    return torch.from_numpy(np.ascontiguousarray(x).astype(np.complex128 if double else np.complex64)).to(
        "cuda", dtype=dt)


def run_plan(x, lengths, batch, direction=-1, norm="none", double=False, in_place=False):
    """Execute one plan on x (numpy, flat) and return the result as numpy."""
    say(f"case lengths={list(lengths)} batch={batch} dir={direction} norm={norm} "
        f"double={double} in_place={in_place}")
    p = sh.Plan(lengths, batch, direction, double=double, in_place=in_place, normalization=norm)
    xd = dev_tensor(x, double)
    if in_place:
        p.execute(xd)
        out = xd
    else:
        out = torch.empty_like(xd)
        p.execute(xd, out)
    torch.cuda.synchronize()
    r = out.cpu().numpy()
    p.close()
    return r


def rnd(n, seed):
    g = np.random.default_rng(seed)
    return g.uniform(-1, 1, n) + 1j * g.uniform(-1, 1, n)


# ------------------------------------------------------------------ reference's own cases

def _in_1d(n, double):
    i = np.arange(n)
    x = ((i * 5 + 1) % 9) / 4 + 1j * (((i * 3 + 2) % 7) / 5)      # tests/sycl_tests.cpp:200-205
    return x.astype(np.complex128 if double else np.complex64)


@pytest.mark.parametrize("n,direction,norm,placement,double", [
    (8, -1, "none", "oop", False), (8, -1, "forward", "oop", False),
    (8, -1, "backward", "oop", False), (8, -1, "orthogonal", "oop", False),
    (7, 1, "none", "inplace", False), (6, -1, "none", "oop", False),
    (15, 1, "none", "oop", False), (35, -1, "none", "oop", False),
    (11, 1, "none", "inplace", False), (11, -1, "orthogonal", "oop", True)])
def test_reference_1d_cases(n, direction, norm, placement, double):
    """tests/sycl_tests.cpp:318-370 (run_1d): same inputs, same tolerance, plus the
    reference restatement at the same precision."""
    x = _in_1d(n, double)
    got = run_plan(x, [n], 1, direction, norm, double, placement == "inplace")
    exp = oracle_fftn(x, [n], 1, direction, norm)
    tol = 2e-11 if double else 4e-4
    assert ref_close(got, exp, tol)
    assert ref_close(got, ref_portable(x, [n], 1, direction, norm, double), 2 * tol)


@pytest.mark.parametrize("shape,batch", [((3, 4), 2), ((2, 3, 5), 1), ((11, 3), 1)])
def test_reference_many_cases(shape, batch):
    """tests/sycl_tests.cpp:153-193 (run_many)."""
    count = int(np.prod(shape)) * batch
    i = np.arange(count)
    x = (((i * 11 + 2) % 17) / 8 + 1j * (((i * 3 + 1) % 13) / 7)).astype(np.complex64)
    got = run_plan(x, shape, batch)
    assert ref_close(got, oracle_fftn(x, shape, batch), 8e-4)


def test_reference_roundtrip_13():
    """tests/sycl_tests.cpp:372-399: forward then backward(normalization backward)."""
    i = np.arange(13)
    x = (i / 7 + 1j * ((i % 3) / 5)).astype(np.complex64)
    say("case roundtrip 13")
    fwd = sh.Plan([13], 1, -1)
    bwd = sh.Plan([13], 1, 1, normalization="backward")
    xd = dev_tensor(x, False)
    spec = torch.empty_like(xd)
    back = torch.empty_like(xd)
    ev = torch.cuda.Event()
    fwd.execute(xd, spec, done=ev)
    bwd.execute(spec, back, deps=[ev])
    torch.cuda.synchronize()
    assert np.all(np.abs(back.cpu().numpy() - x) < 2e-3)


def test_repeated_execution_serializes():
    """A second execute on the same plan (sharing scratch) must give the same result
    (tests/sycl_tests.cpp:227-256)."""
    n, b = 2 ** 18, 4   # two-pass path with plan scratch
    x = rnd(n * b, 5)
    say("case repeated execution 2^18 x 4")
    p = sh.Plan([n], b)
    xd = dev_tensor(x, False)
    o1, o2 = torch.empty_like(xd), torch.empty_like(xd)
    s2 = torch.cuda.Stream()
    p.execute(xd, o1)
    with torch.cuda.stream(s2):
        p.execute(xd, o2, stream=s2)
    torch.cuda.synchronize()
    a, c = o1.cpu().numpy(), o2.cpu().numpy()
    assert np.array_equal(a, c)
    assert rel_l2(a, oracle_fftn(x, [n], b)) < tol_l2(n, False)


@pytest.mark.parametrize("in_place", [False, True])
def test_bluestein_pair_repeats(in_place, monkeypatch):
    """The Bluestein form over two workgroups per transform (n = 10007, half length 10240):
    the second workgroup of each pair to finish adds the halves, counting arrivals on a
    per-transform counter that every execution advances by two.  Repeated executions must
    give bit-identical results, within the oracle bound."""
    monkeypatch.setenv("SYCLFFT_HIP_BLUE_HALVES", "1")
    n, b = 10007, 48
    x = rnd(n * b, 17)
    say(f"case bluestein pair repeats in_place={in_place}")
    p = sh.Plan([n], b, 1, in_place=in_place, normalization="orthogonal")
    assert "pair" in p.describe()
    outs = []
    for _ in range(3):
        xd = dev_tensor(x, False)
        if in_place:
            p.execute(xd)
            outs.append(xd.cpu().numpy())
        else:
            yd = torch.empty_like(xd)
            p.execute(xd, yd)
            outs.append(yd.cpu().numpy())
    p.close()
    assert np.array_equal(outs[0], outs[1]) and np.array_equal(outs[0], outs[2])
    assert rel_l2(outs[0], oracle_fftn(x, [n], b, 1, "orthogonal")) < tol_l2(n, False, 4.0)


@pytest.mark.parametrize("n,b,direction,norm", [
    (10007, 300, -1, "orthogonal"), (10007, 257, 1, "none"), (9973, 200, -1, "forward"), (7919, 160, 1, "backward"),
])
def test_bluestein_one_workgroup_large_batch(n, b, direction, norm):
    """Bluestein with one workgroup per transform (batches above half the CU count): n = 10007
    runs the length-specialized 20160-point convolution, 9973 and 7919 the generic kernel.
    Against the oracle."""
    x = rnd(n * b, 47)
    say(f"case bluestein one workgroup n={n} b={b}")
    p = sh.Plan([n], b, direction, normalization=norm)
    desc = p.describe()
    assert desc.startswith("bluestein-fused"), desc
    xd = dev_tensor(x, False)
    yd = torch.empty_like(xd)
    p.execute(xd, yd)
    torch.cuda.synchronize()
    p.close()
    assert rel_l2(yd.cpu().numpy(), oracle_fftn(x, [n], b, direction, norm)) < tol_l2(n, False, 4.0)


def test_concurrent_plans_and_impulse():
    """Two plans on one stream, impulse inputs (tests/sycl_tests.cpp:410-435)."""
    say("case concurrent plans")
    a = np.zeros(8, complex)
    a[0] = 1
    b = np.zeros(8, complex)
    b[1] = 1
    p1, p2 = sh.Plan([8]), sh.Plan([8])
    da, db = dev_tensor(a, False), dev_tensor(b, False)
    oa, ob = torch.empty_like(da), torch.empty_like(db)
    p1.execute(da, oa)
    p2.execute(db, ob)
    torch.cuda.synchronize()
    assert abs(oa.cpu().numpy()[0].real - 1) < 1e-5 and abs(ob.cpu().numpy()[0].real - 1) < 1e-5
    assert np.allclose(oa.cpu().numpy(), 1, atol=1e-6)


def test_dependency_events_span():
    """Two producer kernels on another stream; execute waits on both events
    (tests/sycl_tests.cpp:483-506)."""
    say("case dependency span")
    p = sh.Plan([8])
    x = torch.empty(8, dtype=torch.complex64, device="cuda")
    out = torch.empty_like(x)
    prod = torch.cuda.Stream()
    e1, e2 = torch.cuda.Event(), torch.cuda.Event()
    with torch.cuda.stream(prod):
        torch.cuda._sleep(2_000_000)
        x[:4] = torch.tensor([1, 0, 0, 0], dtype=torch.complex64, device="cuda")
        e1.record(prod)
        x[4:] = 0
        e2.record(prod)
    p.execute(x, out, deps=[e1, e2])
    torch.cuda.synchronize()
    assert np.allclose(out.cpu().numpy(), 1, atol=1e-5)


def test_error_cases():
    """tests/sycl_tests.cpp:508-538: null / wrong placement / closed plan / host memory."""
    say("case errors")
    p = sh.Plan([8])
    with pytest.raises(sh.SyclfftError) as e:
        p.execute_ptr(0, 0)
    assert e.value.code == "invalid_pointer"
    x = torch.zeros(8, dtype=torch.complex64, device="cuda")
    with pytest.raises(sh.SyclfftError) as e:
        p.execute(x, x)
    assert e.value.code == "invalid_argument"
    host = np.zeros(8, np.complex64)
    with pytest.raises(sh.SyclfftError) as e:
        p.execute_ptr(host.ctypes.data, x.data_ptr())
    assert e.value.code == "invalid_pointer"
    # (a buffer shorter than the transform is caught against its HIP allocation; torch's
    # caching allocator sub-allocates, so that case is tested with raw hipMalloc buffers in
    # tests/cpp/device_tests.cpp instead)
    ip = sh.Plan([8], in_place=True)
    y = torch.zeros(8, dtype=torch.complex64, device="cuda")
    with pytest.raises(sh.SyclfftError) as e:
        ip.execute(x, y)
    assert e.value.code == "invalid_argument"
    p.close()
    with pytest.raises(sh.SyclfftError) as e:
        p.execute(x, y)
    assert e.value.code == "invalid_state"


# ------------------------------------------------------------------ FP64-oracle parity

POW2 = [2, 4, 8, 16, 32, 64, 128, 256, 512, 1024, 2048, 4096, 8192, 16384, 32768, 65536,
        2 ** 17, 2 ** 18, 2 ** 20]
SMOOTH = [6, 12, 15, 35, 49, 60, 81, 120, 240, 343, 360, 625, 1000, 2187, 2520, 3125, 6561,
          10000, 16807, 30030 // 11 * 10, 40320, 100000, 362880]
PRIMES = [11, 13, 17, 97, 257, 1009, 4099, 10007, 65537]


def _batch_for(n, budget=1 << 20):
    return max(1, min(64, budget // n))


@pytest.mark.parametrize("double", [False, True])
@pytest.mark.parametrize("n", POW2)
def test_pow2_forward(n, double):
    b = _batch_for(n)
    x = rnd(n * b, n)
    got = run_plan(x, [n], b, -1, "none", double)
    assert rel_l2(got, oracle_fftn(x, [n], b, -1)) < tol_l2(n, double)


@pytest.mark.parametrize("double", [False, True])
@pytest.mark.parametrize("n", [m for m in SMOOTH if m > 1])
def test_smooth_backward(n, double):
    b = _batch_for(n)
    x = rnd(n * b, n + 1)
    got = run_plan(x, [n], b, 1, "none", double)
    assert rel_l2(got, oracle_fftn(x, [n], b, 1)) < tol_l2(n, double)


@pytest.mark.parametrize("double", [False, True])
@pytest.mark.parametrize("n", PRIMES)
def test_bluestein(n, double):
    b = _batch_for(n, 1 << 18)
    x = rnd(n * b, n + 2)
    for direction in (-1, 1):
        got = run_plan(x, [n], b, direction, "orthogonal", double)
        assert rel_l2(got, oracle_fftn(x, [n], b, direction, "orthogonal")) < tol_l2(n, double, 4.0)


@pytest.mark.parametrize("shape,batch,double", [
    ((16, 16), 3, False), ((32, 32, 32), 2, False), ((64, 48, 10), 1, True),
    ((128, 128), 4, True), ((7, 9, 11), 2, False), ((13, 256), 2, False),
    ((256, 13), 2, True), ((1, 64), 2, False), ((64, 1, 8), 1, True),
    ((2048, 96), 1, False), ((1024, 1024), 1, False), ((4100, 6), 1, True),
    # f64 4096-point strided axes on the column-tile kernel (4 columns per tile)
    ((4096, 8), 2, True), ((4096, 4, 3), 1, True), ((4096, 12), 1, True),
    # column tiles with a ragged last tile per row (33 = 8 x 4 + 1, 65 = 8 x 8 + 1 columns),
    # and fp32 2048-point strided axes (8 columns per tile)
    ((4096, 33), 1, True), ((2048, 65), 2, False), ((2048, 8, 9), 1, False), ((2048, 2048), 1, False)])
def test_multid(shape, batch, double):
    n = int(np.prod(shape))
    x = rnd(n * batch, n)
    for direction, norm in ((-1, "forward"), (1, "none")):
        got = run_plan(x, shape, batch, direction, norm, double)
        assert rel_l2(got, oracle_fftn(x, shape, batch, direction, norm)) < tol_l2(n, double, 3.0)


@pytest.mark.parametrize("norm", ["none", "forward", "backward", "orthogonal"])
def test_in_place_norms(norm):
    n, b = 4096, 8
    x = rnd(n * b, 3)
    for direction in (-1, 1):
        got = run_plan(x, [n], b, direction, norm, False, in_place=True)
        assert rel_l2(got, oracle_fftn(x, [n], b, direction, norm)) < tol_l2(n, False)


def test_all_lengths_one():
    x = rnd(5, 1)
    got = run_plan(x, [1, 1], 5, -1, "orthogonal", True)
    assert np.allclose(got, x)


def test_restatement_parity_pow2_and_smooth():
    """Same-precision parity with the reference portable provider (pow2 and 7-smooth)."""
    for n, double in ((4096, False), (2520, True), (6561, True), (1000, False)):
        b = 4
        x = rnd(n * b, 9).astype(np.complex128 if double else np.complex64)
        got = run_plan(x, [n], b, -1, "orthogonal", double)
        ref = ref_portable(x, [n], b, -1, "orthogonal", double)
        tol = 2e-11 if double else 4e-4
        assert ref_close(got, ref, tol), n
