"""GPU tests of plan execution captured into a HIP graph (torch.cuda.graph): the captured
launches must replay correctly several times, with new input copied in eagerly and the
stream synchronized between replays, and every replay must be bit-identical to an eager
execution of the same plan on the same input (the same kernels run on the same values)
and within the oracle bound c * log2(N) * u.  Pointer-array executions whose pointers are
contiguous (transform k at base + k * transform size) need no unit tables and can be captured; the others are
rejected with invalid_state before anything is enqueued.
"""
import sys

import numpy as np
import pytest

from conftest import oracle_fftn, rel_l2

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
import syclfft_hip as sh  # noqa: E402

U = {False: 2.0 ** -24, True: 2.0 ** -53}
REPLAYS = 4


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def rnd(n, seed, double):
    g = np.random.default_rng(seed)
    x = g.uniform(-1, 1, n) + 1j * g.uniform(-1, 1, n)
    return x.astype(np.complex128 if double else np.complex64)


def tol(lengths, double):
    n = int(np.prod(lengths))
    return (12.0 if len(lengths) > 1 else 4.0) * max(1.0, np.log2(n)) * U[double]


# (lengths, batch, double, direction, norm, in_place)
CASES = [
    ([4096], 8, False, -1, "none", False),
    ([2520], 4, True, 1, "orthogonal", True),
    ([10007], 3, False, -1, "forward", False),   # Bluestein (chirp tables)
    ([65536], 2, False, 1, "backward", False),   # two launches through scratch
    ([1009], 3, True, -1, "none", False),
    ([64, 48], 3, False, 1, "none", True),
    ([32, 32, 32], 2, True, -1, "forward", False),
]


def warm(p, x, y, in_place):
    """One eager execution on a side stream (torch's recommended warm-up before capture)."""
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(s):
        p.execute(x) if in_place else p.execute(x, y)
    torch.cuda.current_stream().wait_stream(s)
    torch.cuda.synchronize()


@pytest.mark.parametrize("case", CASES, ids=lambda c: "x".join(map(str, c[0])) + f"_b{c[1]}" + ("_d" if c[2] else "_f"))
def test_captured_execute_replays(case):
    lengths, batch, double, direction, norm, in_place = case
    n = int(np.prod(lengths))
    say(f"case graph lengths={lengths} batch={batch} double={double} dir={direction} norm={norm} "
        f"in_place={in_place}")
    dt = torch.complex128 if double else torch.complex64
    with sh.Plan(lengths, batch, direction, double=double, in_place=in_place, normalization=norm) as p:
        xs = torch.zeros(n * batch, dtype=dt, device="cuda")
        ys = xs if in_place else torch.zeros_like(xs)
        warm(p, xs, ys, in_place)
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g):
            p.execute(xs) if in_place else p.execute(xs, ys)
        for r in range(REPLAYS):
            host = rnd(n * batch, 7000 + 31 * r + n, double)
            src = torch.from_numpy(host).to("cuda")
            xs.copy_(src)
            torch.cuda.synchronize()
            g.replay()
            torch.cuda.synchronize()
            got = ys.cpu().numpy()
            # the same plan, eagerly, between replays
            xe = src.clone()
            ye = xe if in_place else torch.empty_like(xe)
            p.execute(xe) if in_place else p.execute(xe, ye)
            torch.cuda.synchronize()
            assert np.array_equal(got, ye.cpu().numpy()), f"replay {r} differs from the eager execution"
            want = oracle_fftn(host.astype(np.complex128), lengths, batch, direction, norm)
            e = rel_l2(got.astype(np.complex128), want)
            assert e <= tol(lengths, double), (r, e)
        say(f"{REPLAYS} replays bit-identical to eager")
        del g


def test_captured_execute_many_affine_and_rejects_tables():
    n, batch = 4096, 6
    say(f"case graph execute_many n={n} batch={batch}")
    with sh.Plan([n], batch, sh.FORWARD) as p:
        x = torch.zeros(n * batch, dtype=torch.complex64, device="cuda")
        y = torch.zeros_like(x)
        # the first half of the batch as views of one buffer: contiguous pointers
        used = batch // 2
        xv = [x[k * n:(k + 1) * n] for k in range(used)]
        yv = [y[k * n:(k + 1) * n] for k in range(used)]
        scattered_in = [torch.zeros(n, dtype=torch.complex64, device="cuda") for _ in range(3)]
        scattered_out = [torch.zeros_like(t) for t in scattered_in]
        scattered_in.reverse()  # descending addresses: not contiguous
        warm(p, x, y, False)
        p.execute_many(xv, yv)
        torch.cuda.synchronize()
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g):
            p.execute_many(xv, yv)
            with pytest.raises(sh.SyclfftError) as e:
                p.execute_many(scattered_in, scattered_out)
        assert e.value.code == "invalid_state"
        for r in range(REPLAYS):
            host = rnd(n * batch, 9100 + r, False)
            x.copy_(torch.from_numpy(host).to("cuda"))
            y.zero_()
            torch.cuda.synchronize()
            g.replay()
            torch.cuda.synchronize()
            got = y.cpu().numpy().reshape(batch, n)
            want = oracle_fftn(host.astype(np.complex128), [n], batch, -1, "none").reshape(batch, n)
            for k in range(batch):
                if k >= used:
                    assert not got[k].any(), f"replay {r} wrote transform {k}"
                else:
                    assert rel_l2(got[k].astype(np.complex128), want[k]) <= tol([n], False), (r, k)
        del g


# layout plans: fused padded rows, interleaved batches (strided-axis plan), planar gather / scatter
LAYOUT_CASES = [
    ([4096], 4, (1, 4096 + 5, 1, 4096 + 5, 0, 0)),
    ([1000], 8, (8, 1, 8, 1, 0, 0)),
    ([64, 48], 3, (1, 64 * 48, 1, 64 * 48, 1, 1)),
]


@pytest.mark.parametrize("lengths,batch,lay", LAYOUT_CASES, ids=lambda c: str(c))
def test_captured_layout_replays(lengths, batch, lay):
    """A captured layout execution (fused, strided-axis, or gather / core / scatter) replays
    bit-identically to an eager execution of the same plan."""
    say(f"case graph layout lengths={lengths} batch={batch} layout={lay}")
    T = int(np.prod(lengths))
    istride, idist, ostride, odist, ip, op = lay
    iext = (batch - 1) * idist + (T - 1) * istride + 1
    oext = (batch - 1) * odist + (T - 1) * ostride + 1
    with sh.Plan(lengths, batch, sh.FORWARD, layout=sh.Layout(*lay)) as p:
        def side(ext, planar):
            if planar:
                return (torch.zeros(ext, device="cuda"), torch.zeros(ext, device="cuda"))
            return torch.zeros(ext, dtype=torch.complex64, device="cuda")
        xs, ys = side(iext, ip), side(oext, op)
        s = torch.cuda.Stream()
        s.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(s):
            p.execute_layout(xs, ys)
        torch.cuda.current_stream().wait_stream(s)
        torch.cuda.synchronize()
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g):
            p.execute_layout(xs, ys)
        flat = lambda v: torch.cat([v[0], v[1]]) if isinstance(v, tuple) else v  # noqa: E731
        for r in range(REPLAYS):
            gen = torch.Generator(device="cuda").manual_seed(8100 + r)
            for t in (xs if isinstance(xs, tuple) else (torch.view_as_real(xs),)):
                t.copy_(torch.rand(t.shape, generator=gen, device="cuda") * 2 - 1)
            torch.cuda.synchronize()
            g.replay()
            torch.cuda.synchronize()
            got = flat(ys).cpu().numpy().copy()
            ye = side(oext, op)
            p.execute_layout(xs, ye)
            torch.cuda.synchronize()
            assert np.array_equal(got, flat(ye).cpu().numpy()), f"replay {r} differs from the eager execution"
        say(f"{REPLAYS} replays bit-identical to eager [{p.describe().strip()}]")
        del g


@pytest.mark.parametrize("lengths,batch,direction", [([4096], 8, -1), ([64, 48], 3, -1), ([64, 48], 3, 1),
                                                     ([16, 10, 12], 2, 1)])
def test_captured_real_replays(lengths, batch, direction):
    """Captured real plans (1-D, and multi-D with the leading axes over the half spectrum)
    replay bit-identically to eager executions; R2C replays also meet the oracle bound."""
    say(f"case graph real lengths={lengths} batch={batch} dir={direction}")
    T = int(np.prod(lengths))
    H = T // lengths[-1] * (lengths[-1] // 2 + 1)
    with sh.Plan(lengths, batch, direction, real=True) as p:
        if direction == -1:
            xs = torch.zeros(T * batch, device="cuda")
            ys = torch.zeros(H * batch, dtype=torch.complex64, device="cuda")
        else:
            xs = torch.zeros(H * batch, dtype=torch.complex64, device="cuda")
            ys = torch.zeros(T * batch, device="cuda")
        warm(p, xs, ys, False)
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g):
            p.execute(xs, ys)
        for r in range(REPLAYS):
            gen = torch.Generator(device="cuda").manual_seed(8200 + r)
            src = xs if direction == -1 else torch.view_as_real(xs)
            src.copy_(torch.rand(src.shape, generator=gen, device="cuda") * 2 - 1)
            torch.cuda.synchronize()
            g.replay()
            torch.cuda.synchronize()
            got = ys.cpu().numpy().copy()
            ye = torch.empty_like(ys)
            p.execute(xs, ye)
            torch.cuda.synchronize()
            assert np.array_equal(got, ye.cpu().numpy()), f"replay {r} differs from the eager execution"
            if direction == -1:
                x = xs.cpu().numpy().astype(np.complex128)
                want = oracle_fftn(x, lengths, batch, -1, "none").reshape(batch, T // lengths[-1], lengths[-1])
                want = want[..., : lengths[-1] // 2 + 1]
                assert rel_l2(got.astype(np.complex128), want) <= tol(lengths, False), r
        say(f"{REPLAYS} replays bit-identical to eager [{p.describe().strip()}]")
        del g
