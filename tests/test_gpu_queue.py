"""Work-queue execution (syclfft_hip_execute_queue) on the GPU, through the C ABI.

Plans whose execution is one plain tile launch and share a kernel variant run in one
persistent launch whose workgroups take tiles from a device counter; other plans execute
first on the same stream.  Each queued execution must give exactly the bits of the same
plan executed alone (the tile code and its arguments are the same), and match the FP64
oracle within c log2(N) u (c = 2)."""
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


def make_case(lengths, batch, direction, norm, double, in_place, seed):
    n = int(np.prod(lengths))
    g = np.random.default_rng(seed)
    x = (g.uniform(-1, 1, n * batch) + 1j * g.uniform(-1, 1, n * batch)).astype(
        np.complex128 if double else np.complex64)
    p = sh.Plan(lengths, batch, direction, double=double, in_place=in_place, normalization=norm)
    return dict(lengths=lengths, batch=batch, direction=direction, norm=norm, double=double, in_place=in_place,
                x=x, plan=p)


def run_alone(c):
    xd = torch.from_numpy(c["x"].copy()).to("cuda")
    if c["in_place"]:
        c["plan"].execute(xd)
        out = xd
    else:
        out = torch.empty_like(xd)
        c["plan"].execute(xd, out)
    torch.cuda.synchronize()
    return out.cpu().numpy()


def run_queue(cases):
    xs = [torch.from_numpy(c["x"].copy()).to("cuda") for c in cases]
    ys = [x if c["in_place"] else torch.empty_like(x) for c, x in zip(cases, xs)]
    sh.execute_queue([c["plan"] for c in cases], xs, ys)
    torch.cuda.synchronize()
    return [y.cpu().numpy() for y in ys]


MIXED = [
    ([256], 5, -1, "none", False, False),
    ([240], 3, 1, "orthogonal", False, False),
    ([4096], 2, -1, "forward", False, True),
    ([2520], 8, 1, "orthogonal", True, False),
    ([16], 33, -1, "none", False, False),
    ([64], 7, -1, "none", True, True),
    ([512], 4, 1, "backward", False, False),
    ([1000], 3, -1, "none", False, False),
    ([6561], 2, -1, "orthogonal", True, False),
    ([16384], 3, 1, "none", False, False),
    ([10007], 2, -1, "none", False, False),      # Bluestein: executes directly
    ([32, 32], 2, 1, "backward", False, False),  # two launches: executes directly
    ([65536], 1, -1, "none", False, False),      # four-step: executes directly
]


def test_queue_with_real_plan_entry():
    """A real (R2C) plan in a queue call executes directly with its own buffer sizes (N real
    values in, N/2 + 1 complex values out) beside queued complex plans."""
    n, batch = 512, 3
    g = np.random.default_rng(20261500)
    xr = g.uniform(-1, 1, n * batch).astype(np.float32)
    rp = sh.Plan([n], batch, sh.FORWARD, real=True)
    c = make_case([256], 4, -1, "none", False, False, seed=20261501)
    xd = torch.from_numpy(xr).to("cuda")
    yd = torch.empty(batch * (n // 2 + 1), dtype=torch.complex64, device="cuda")
    cx = torch.from_numpy(c["x"].copy()).to("cuda")
    cy = torch.empty_like(cx)
    say("case: work queue with a real plan entry")
    sh.execute_queue_ptr([rp, c["plan"]], [xd.data_ptr(), cx.data_ptr()], [yd.data_ptr(), cy.data_ptr()])
    torch.cuda.synchronize()
    ref = oracle_fftn(xr.astype(np.complex128), [n], batch, -1, "none").reshape(batch, n)[:, : n // 2 + 1]
    assert rel_l2(yd.cpu().numpy(), ref) <= 4 * np.log2(n) * U[False]
    assert np.array_equal(cy.cpu().numpy(), run_alone(c))
    rp.close()
    c["plan"].close()


def test_queue_mixed_plans_bit_identical_to_alone_and_oracle():
    cases = [make_case(*m, seed=20260800 + k) for k, m in enumerate(MIXED)]
    say("case: work queue over " + ", ".join(f"{c['lengths']}x{c['batch']}{'d' if c['double'] else 'f'}"
                                              for c in cases))
    got = run_queue(cases)
    for c, y in zip(cases, got):
        alone = run_alone(c)
        n = int(np.prod(c["lengths"]))
        ref = oracle_fftn(c["x"].astype(np.complex128), c["lengths"], c["batch"], c["direction"], c["norm"])
        err = rel_l2(y, ref)
        tol = 2 * max(1.0, np.log2(n)) * U[c["double"]] * (4 if n == 10007 else 1)
        say(f"  {c['lengths']} batch {c['batch']} rel-L2 {err:.3e} (tol {tol:.1e})")
        assert np.array_equal(y, alone), c["lengths"]
        assert err <= tol
        c["plan"].close()


def test_queue_many_small_plans():
    """200 plans drawn from the serving length mix (7-smooth, one tile), one call."""
    g = np.random.default_rng(20260900)
    smooth = [n for n in range(16, 4097) if _smooth7(n)]
    cases = []
    for k in range(200):
        n = int(smooth[int(g.integers(len(smooth)))])
        batch = int(np.exp(g.uniform(0, np.log(64))))
        double = bool(g.random() < 0.2)
        cases.append(make_case([n], batch, int(g.choice([-1, 1])), ["none", "forward", "backward", "orthogonal"][k % 4],
                               double, False, 20261000 + k))
    say(f"case: work queue over {len(cases)} small plans")
    got = run_queue(cases)
    bad = 0
    for c, y in zip(cases[:40], got[:40]):
        bad += not np.array_equal(y, run_alone(c))
    for c, y in zip(cases, got):
        n = c["lengths"][0]
        ref = oracle_fftn(c["x"].astype(np.complex128), [n], c["batch"], c["direction"], c["norm"])
        assert rel_l2(y, ref) <= 2 * max(1.0, np.log2(n)) * U[c["double"]]
    for c in cases:
        c["plan"].close()
    assert bad == 0


def _smooth7(n):
    for p in (2, 3, 5, 7):
        while n % p == 0:
            n //= p
    return n == 1


def test_queue_errors():
    p = sh.Plan([256], 2, sh.FORWARD)
    x = torch.zeros(512, dtype=torch.complex64, device="cuda")
    host = np.zeros(512, dtype=np.complex64)  # not device memory: rejected before any launch
    with pytest.raises(sh.SyclfftError) as e:
        sh.execute_queue_ptr([p], [x.data_ptr()], [host.ctypes.data])
    assert e.value.code == "invalid_pointer"
    with pytest.raises(sh.SyclfftError) as e:
        sh.execute_queue_ptr([p], [x.data_ptr()], [x.data_ptr()])  # out-of-place plan, in == out
    assert e.value.code == "invalid_argument"
    sh.execute_queue_ptr([], [], [])  # nothing to do
    p.close()
