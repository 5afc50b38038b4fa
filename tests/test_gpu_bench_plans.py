"""GPU parity of the exact plans bench.py times, at their full batch counts.

bench.py builds every timed plan with planning_mode::measure, and planning depends on the
batch count (column-tile width, global-table twin, persistent forms), so these tests
build the same plans the same way: same lengths, batch, precision, direction,
normalization and placement, measure=True.  Each check runs the plan on seeded
uniform[-1,1) inputs and compares the first, middle and last transform of the batch
with the FP64 oracle; round-trip configurations also run the backward plan on the
forward output and compare with the input.

The autotuner can select any of its kernel forms on another box or batch: the
multi-launch form (s1), the column-tile width halved or doubled (w-1 / w1), the
persistent prefetching kernel (p1) and the next larger kernel variant (g1).  Each form
is forced through SYCLFFT_HIP_TUNE on the bench's shapes and checked the same way.

Tolerance: rel-L2 <= c log2(N) u against the FP64 oracle (u = 2^-24 for float, 2^-53
for double; c = 2, 4 for Bluestein).  Split BF16 operands carry ~2^-16 relative error
per product, so that mode is checked at its own bound (2e-5) and is below the fp32
accuracy target of 1e-6; split FP16 is checked at 1e-6.

Reference: src/sycl/sycl.cpp:688-775 (execute), tests/sycl_tests.cpp:318-370 (cases).
"""
import sys

import numpy as np
import pytest

from conftest import oracle_fftn, rel_l2

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
import syclfft_hip as sh  # noqa: E402

U = {False: 2.0 ** -24, True: 2.0 ** -53}
SWEEP_POINTS = (512 << 20) // 8  # fp32 sweep: 512 MiB per buffer


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def is_prime(n):
    return n > 1 and all(n % f for f in range(2, int(n ** 0.5) + 1))


def bound(n, double, precision="valu"):
    if precision == "split_bf16":
        return 2e-5
    if precision == "split_fp16":
        return 1e-6
    c = 4.0 if is_prime(n) and n > 7 else 2.0
    return c * max(1.0, np.log2(n)) * U[double]


def device_input(n, batch, double, seed):
    g = torch.Generator(device="cuda")
    g.manual_seed(seed)
    rdt = torch.float64 if double else torch.float32
    x = torch.rand((n * batch, 2), generator=g, device="cuda", dtype=rdt) * 2 - 1
    return torch.view_as_complex(x.contiguous())


def sample_lines(t, n, batch):
    """First, middle and last transform of a batch (host copies)."""
    idx = sorted({0, batch // 2, batch - 1})
    return idx, [t[i * n:(i + 1) * n].cpu().numpy().astype(np.complex128) for i in idx]


def check_plan(n, batch, double, direction, norm, precision="valu", measure=True, roundtrip=False,
               seed=20260101, expect=None):
    say(f"case n={n} batch={batch} double={double} dir={direction} norm={norm} precision={precision} "
        f"measure={measure} roundtrip={roundtrip}")
    x = device_input(n, batch, double, seed)
    y = torch.empty_like(x)
    with sh.Plan([n], batch, direction, double=double, normalization=norm, precision=precision,
                 measure=measure) as p:
        desc = p.describe().strip()
        say("plan " + desc.replace("\n", " | "))
        if expect:
            assert expect in desc, (expect, desc)
        p.execute(x, y)
        torch.cuda.synchronize()
    idx, xs = sample_lines(x, n, batch)
    _, ys = sample_lines(y, n, batch)
    tol = bound(n, double, precision)
    for i, xi, yi in zip(idx, xs, ys):
        e = rel_l2(yi, oracle_fftn(xi, [n], 1, direction, norm))
        say(f"  line {i}: rel-L2 {e:.3e} (bound {tol:.1e})")
        assert e <= tol, (i, e)
    if roundtrip:
        z = torch.empty_like(x)
        with sh.Plan([n], batch, -direction, double=double, normalization=norm, precision=precision,
                     measure=measure) as q:
            say("plan " + q.describe().strip().replace("\n", " | "))
            q.execute(y, z)
            torch.cuda.synchronize()
        _, zs = sample_lines(z, n, batch)
        for i, xi, zi in zip(idx, xs, zs):
            e = rel_l2(zi, xi)
            say(f"  round trip line {i}: rel-L2 {e:.3e}")
            assert e <= 2 * tol, (i, e)
    del x, y
    torch.cuda.empty_cache()


# (id, n, batch, double, direction, norm, precision, roundtrip): the bench's timed plans
BENCH_CASES = [
    ("c1_f32_4096x256", 4096, 256, False, -1, "none", "valu", False),
    ("c2_f64_2520x1024", 2520, 1024, True, -1, "orthogonal", "valu", True),
    ("c2_f64_6561x512", 6561, 512, True, -1, "orthogonal", "valu", True),
    ("c2_f32_10007x128", 10007, 128, False, -1, "orthogonal", "valu", True),
] + [
    (f"c4_f32_{n}x{b}_{prec}", n, b, False, -1, "none", prec, False)
    for n, b in ((16384, 4096), (1 << 20, 16)) for prec in ("valu", "split_bf16", "split_fp16")
] + [
    (f"sweep_f32_{1 << lg}", 1 << lg, SWEEP_POINTS >> lg, False, -1, "none", "valu", False)
    for lg in range(12, 21)
]


@pytest.mark.parametrize("case", BENCH_CASES, ids=[c[0] for c in BENCH_CASES])
def test_bench_plan_full_batch(case):
    _, n, batch, double, direction, norm, precision, roundtrip = case
    check_plan(n, batch, double, direction, norm, precision, measure=True, roundtrip=roundtrip)


# Forced autotuner forms on the bench's shapes.  `expect` names the describe() marker of
# forms that change the launch itself where the shape admits them.  The length-specialized
# kernels (4096 ... 32768, f64 2520 and 6561, the Bluestein convolution of 10007) have no
# persistent or column-width forms: there those forms leave the launch as it is, and the
# last shape (the generic kernel forced with SYCLFFT_HIP_FIXED=0) is where the persistent
# form runs.
FORMS = ["s1w0p0g0", "s0w-1p0g0", "s0w1p0g0", "s0w0p1g0", "s0w0p0g1"]
FORM_SHAPES = [
    (4096, 256, False, "none", True), (2520, 1024, True, "orthogonal", True),
    (6561, 512, True, "orthogonal", True), (10007, 128, False, "orthogonal", True),
    (16384, 4096, False, "none", True), (1 << 20, 16, False, "none", True),
    (4096, SWEEP_POINTS >> 12, False, "none", True), (8192, SWEEP_POINTS >> 13, False, "none", True),
    (1 << 15, SWEEP_POINTS >> 15, False, "none", True), (1 << 18, SWEEP_POINTS >> 18, False, "none", True),
    (4096, SWEEP_POINTS >> 12, False, "none", False),
]


def _form_id(shape, form):
    n, b, d, _, fixed = shape
    return f"{form}_{n}x{b}{'_d' if d else '_f'}{'' if fixed else '_generic'}"


@pytest.mark.parametrize("shape,form", [(s, f) for s in FORM_SHAPES for f in FORMS],
                         ids=[_form_id(s, f) for s in FORM_SHAPES for f in FORMS])
def test_forced_tuning_form(shape, form, monkeypatch):
    n, batch, double, norm, fixed = shape
    monkeypatch.setenv("SYCLFFT_HIP_TUNE", form)
    if not fixed:
        monkeypatch.setenv("SYCLFFT_HIP_FIXED", "0")
    expect = None
    if form == "s0w0p1g0" and not fixed:
        expect = "persistent"  # many more tiles than resident workgroups
    check_plan(n, batch, double, -1, norm, "valu", measure=False, roundtrip=(norm == "orthogonal"),
               seed=20260102, expect=expect)


def test_forced_forms_reach_the_kernels(monkeypatch):
    """The forced forms are not no-ops: the persistent form (generic kernel) and the
    multi-launch form each change the launch sequence of a shape where they apply."""
    import os

    monkeypatch.setenv("SYCLFFT_HIP_FIXED", "0")
    descs = {}
    for form in ("s0w0p0g0", "s0w0p1g0", "s1w0p0g0"):
        os.environ["SYCLFFT_HIP_TUNE"] = form
        try:
            with sh.Plan([4096], SWEEP_POINTS >> 12) as p:
                descs[form] = (p.describe(), p.launch_count())
        finally:
            del os.environ["SYCLFFT_HIP_TUNE"]
    say(f"descs {descs}")
    assert "persistent" in descs["s0w0p1g0"][0] and "persistent" not in descs["s0w0p0g0"][0]
    assert descs["s1w0p0g0"][1] == 2 and descs["s0w0p0g0"][1] == 1
