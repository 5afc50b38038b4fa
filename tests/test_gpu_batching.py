"""GPU tests of execution over pointer arrays (syclfft_hip_plan_execute_many): each
transform in its own allocation, fewer transforms than the plan's batch count, and the
error paths.  The results must be bit-identical to the plan's contiguous execution (the
same kernels run on the same values; only the addressing differs) and within the
oracle bound c * log2(N) * u.
"""
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


def rnd(n, seed):
    g = np.random.default_rng(seed)
    return g.uniform(-1, 1, n) + 1j * g.uniform(-1, 1, n)


# (lengths, batch, double, direction, norm, in_place, forced tuning)
CASES = [
    ([4096], 6, False, -1, "none", False, None),
    ([300], 9, False, 1, "orthogonal", False, None),
    ([2520], 5, True, 1, "orthogonal", True, None),
    ([10007], 4, False, -1, "forward", False, None),
    ([10007], 4, False, 1, "none", False, "s1w0p0g0"),
    ([32768], 3, False, -1, "none", False, None),
    ([2 ** 20], 2, False, 1, "backward", False, None),
    ([1009], 5, True, -1, "orthogonal", True, "s1w0p0g0"),
    ([13, 256], 3, False, -1, "none", False, None),
    ([256, 13], 3, True, 1, "none", False, None),
    ([32, 32, 32], 3, False, 1, "none", True, None),
    ([64, 48, 10], 2, True, -1, "forward", False, None),
    ([10007], 3, False, -1, "orthogonal", False, "halves"),
    ([1009], 4, True, 1, "none", True, "halves"),
]


@pytest.mark.parametrize("case", CASES, ids=lambda c: "x".join(map(str, c[0])) + f"_b{c[1]}" + ("_d" if c[2] else "_f") + (f"_{c[6]}" if c[6] else ""))
def test_execute_many_matches_contiguous(case, monkeypatch):
    lengths, batch, double, direction, norm, in_place, tune = case
    if tune == "halves":  # Bluestein over two workgroups per transform
        monkeypatch.setenv("SYCLFFT_HIP_BLUE_HALVES", "1")
    elif tune:
        monkeypatch.setenv("SYCLFFT_HIP_TUNE", tune)
    n = int(np.prod(lengths))
    say(f"case lengths={lengths} batch={batch} double={double} dir={direction} norm={norm} "
        f"in_place={in_place} tune={tune}")
    dt = torch.complex128 if double else torch.complex64
    x = torch.from_numpy(rnd(n * batch, 4000 + n).astype(np.complex128 if double else np.complex64)).to("cuda")
    with sh.Plan(lengths, batch, direction, double=double, in_place=in_place, normalization=norm) as p:
        # contiguous execution
        if in_place:
            ref = x.clone()
            p.execute(ref)
        else:
            ref = torch.empty_like(x)
            p.execute(x, ref)
        # one allocation per transform, allocated in reverse order; all, then the first half
        for count in (batch, (batch + 1) // 2):
            ins = [None] * count
            for k in reversed(range(count)):
                ins[k] = x[k * n:(k + 1) * n].clone()
            outs = ins if in_place else [torch.empty(n, dtype=dt, device="cuda") for _ in range(count)]
            p.execute_many(ins, outs)
            torch.cuda.synchronize()
            got = torch.cat(outs).cpu().numpy()
            want = ref[: count * n].cpu().numpy()
            assert np.array_equal(got, want), f"count={count}: pointer-array result differs from contiguous"
        torch.cuda.synchronize()
    xo = x.cpu().numpy().astype(np.complex128)
    sign = -1 if direction < 0 else 1
    expect = oracle_fftn(xo, lengths, batch, sign, norm)
    err = rel_l2(ref.cpu().numpy().astype(np.complex128), expect)
    tol = (12.0 if len(lengths) > 1 else 4.0) * max(1.0, np.log2(n)) * U[double]
    assert err <= tol, (err, tol)


def test_execute_many_errors():
    say("case execute_many error paths")
    n = 256
    with sh.Plan([n], 4, sh.FORWARD) as p:
        xs = [torch.zeros(n, dtype=torch.complex64, device="cuda") for _ in range(5)]
        ys = [torch.zeros(n, dtype=torch.complex64, device="cuda") for _ in range(5)]
        with pytest.raises(sh.SyclfftError) as e:
            p.execute_many(xs, ys)  # 5 > batch count 4
        assert e.value.code == "invalid_argument"
        with pytest.raises(sh.SyclfftError) as e:
            p.execute_many_ptr([], [])
        assert e.value.code == "invalid_argument"
        host = np.zeros(n, dtype=np.complex64)
        with pytest.raises(sh.SyclfftError) as e:
            p.execute_many_ptr([host.ctypes.data], [ys[0].data_ptr()])
        assert e.value.code == "invalid_pointer"
        with pytest.raises(sh.SyclfftError) as e:
            p.execute_many(xs[:2], xs[:2])  # out-of-place plan, in == out
        assert e.value.code == "invalid_argument"
        # a valid call after the rejected ones still works
        p.execute_many(xs[:2], ys[:2])
        torch.cuda.synchronize()
