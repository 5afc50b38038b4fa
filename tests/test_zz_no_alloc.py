"""syclfft_hip_plan_execute_noalloc, called through the C ABI, against
syclfft_hip_plan_execute on the same plan and input: launches that pass intermediates
through scratch (four-step axes, the three-launch Bluestein form) run with the library's
file-scope device array instead of the plan's own scratch, with the same kernels in the
same order, so the outputs must be bit-identical.  Two calls in a row, the second with a
smaller input, then the error paths: a plan whose scratch exceeds the array and a real plan
are rejected with invalid_argument.
"""
import ctypes
import sys

import numpy as np
import pytest

from conftest import oracle_fftn, rel_l2

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
import syclfft_hip as sh  # noqa: E402


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def noalloc(plan, x, y):
    lib = sh.load()
    err = ctypes.create_string_buffer(1024)
    s = torch.cuda.current_stream(x.device).cuda_stream
    st = lib.syclfft_hip_plan_execute_noalloc(plan.handle, ctypes.c_void_p(x.data_ptr()),
                                              ctypes.c_void_p(y.data_ptr()), ctypes.c_void_p(s), None, 0,
                                              None, err, 1024)
    return st, err.value.decode()


def device_input(n, double, seed):
    g = torch.Generator(device="cuda")
    g.manual_seed(seed)
    rdt = torch.float64 if double else torch.float32
    return torch.view_as_complex((torch.rand((n, 2), generator=g, device="cuda", dtype=rdt) * 2 - 1).contiguous())


# (lengths, batch, double, direction, norm, in_place, forced tuning): the second case is the
# smaller one that follows the first in the same process
CASES = [
    ([1 << 18], 8, False, -1, "none", False, None),        # four-step, 16 MiB of scratch
    ([1 << 16], 4, False, 1, "backward", False, None),     # four-step, smaller
    ([10007], 4, False, -1, "orthogonal", False, "s1w0p0g0"),  # three-launch Bluestein
    ([16384], 8, True, 1, "none", True, None),            # f64 four-step, in place
    ([64, 4096], 2, False, -1, "forward", False, None),    # 2-D with a four-step-free plan
]


def test_noalloc_matches_execute(monkeypatch):
    lib = sh.load()
    limit = int(lib.syclfft_hip_noalloc_scratch_limit())
    say(f"no-allocation scratch limit {limit} bytes")
    assert limit == 64 << 20
    for lengths, batch, double, direction, norm, in_place, tune in CASES:
        if tune:
            monkeypatch.setenv("SYCLFFT_HIP_TUNE", tune)
        else:
            monkeypatch.delenv("SYCLFFT_HIP_TUNE", raising=False)
        n = int(np.prod(lengths))
        say(f"case lengths={lengths} batch={batch} double={double} dir={direction} norm={norm} "
            f"in_place={in_place} tune={tune}")
        x = device_input(n * batch, double, 20260200 + n)
        with sh.Plan(lengths, batch, direction, double=double, in_place=in_place, normalization=norm) as p:
            say(f"  plan {p.describe().strip()} scratch={p.scratch_size_bytes()}")
            assert p.scratch_size_bytes() <= limit
            if in_place:
                y1 = x.clone()
                y2 = x.clone()
                p.execute(y1)
                st, msg = noalloc(p, y2, y2)
            else:
                y1 = torch.empty_like(x)
                y2 = torch.full_like(x, float("nan"))
                p.execute(x, y1)
                st, msg = noalloc(p, x, y2)
            assert st == 0, msg
            torch.cuda.synchronize()
        a, b = y1.cpu().numpy(), y2.cpu().numpy()
        assert np.array_equal(a, b), "no-allocation execution differs from execute"
        xs = x[:n].cpu().numpy().astype(np.complex128)
        e = rel_l2(b[:n].astype(np.complex128), oracle_fftn(xs, lengths, 1, direction, norm))
        say(f"  bit-identical; first transform rel-L2 vs oracle {e:.3e}")
        assert e < 4 * np.log2(n) * (2.0 ** -53 if double else 2.0 ** -24)
        del x, y1, y2


def test_noalloc_rejects_what_it_cannot_run():
    lib = sh.load()
    limit = int(lib.syclfft_hip_noalloc_scratch_limit())
    say("case over-limit plan")
    n, batch = 1 << 20, 16  # four-step, 128 MiB of scratch
    x = torch.zeros(n * batch, dtype=torch.complex64, device="cuda")
    y = torch.empty_like(x)
    with sh.Plan([n], batch) as p:
        assert p.scratch_size_bytes() > limit
        st, msg = noalloc(p, x, y)
    say(f"  status {st}: {msg}")
    assert st == 1 and "scratch" in msg
    say("case real plan")
    r = torch.zeros(4096 * 2, dtype=torch.float32, device="cuda")
    c = torch.zeros(2049 * 2, dtype=torch.complex64, device="cuda")
    with sh.Plan([4096], 2, sh.FORWARD, real=True) as p:
        st, msg = noalloc(p, r, c)
    say(f"  status {st}: {msg}")
    assert st == 1
    torch.cuda.synchronize()
