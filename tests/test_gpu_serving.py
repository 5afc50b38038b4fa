"""GPU tests of the serving loop (include/syclfft_hip_server.h): requests of mixed shapes,
types, directions and normalizations submitted together are merged into shared launches
and overlap on several streams; every result must match the FP64 oracle within
c * log2(N) * u, and must be the same (within that bound) when the request runs alone.
Requests that arrive while others are on the GPU are admitted as slots free up.
"""
import sys
import threading
import time

import numpy as np
import pytest

from conftest import oracle_fftn, rel_l2

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
import syclfft_hip as sh  # noqa: E402

U = {False: 2.0 ** -24, True: 2.0 ** -53}
NORMS = ["none", "forward", "backward", "orthogonal"]


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def tol(lengths, double):
    n = int(np.prod(lengths))
    return (12.0 if len(lengths) > 1 else 4.0) * max(1.0, np.log2(n)) * U[double]


def make_requests(count, seed):
    g = np.random.default_rng(seed)
    shapes = [[256], [4096], [2520], [65536], [1009], [10007], [64, 64], [13, 32]]
    reqs = []
    for k in range(count):
        lengths = shapes[int(g.integers(len(shapes)))]
        n = int(np.prod(lengths))
        batch = int(g.integers(1, 5)) if n <= 16384 else 1
        double = bool(g.random() < 0.25)
        direction = int(g.choice([-1, 1]))
        norm = NORMS[int(g.integers(4))]
        in_place = bool(g.random() < 0.2)
        x = (g.uniform(-1, 1, n * batch) + 1j * g.uniform(-1, 1, n * batch))
        xd = torch.from_numpy(x.astype(np.complex128 if double else np.complex64)).to("cuda")
        reqs.append(dict(lengths=lengths, batch=batch, double=double, direction=direction, norm=norm,
                         in_place=in_place, x=x, xd=xd))
    return reqs


def run_one(srv, r, tag):
    """Submit request r (fresh copies of its input); returns (ticket, output tensor)."""
    xin = r["xd"].clone()
    out = xin if r["in_place"] else torch.empty_like(xin)
    t = srv.submit(xin, out, r["lengths"], r["batch"], r["direction"], r["norm"], tag=tag)
    return t, xin, out


def test_serving_mix_matches_alone_and_oracle():
    reqs = make_requests(48, 20260006)
    say(f"case serving mix of {len(reqs)} requests")
    torch.cuda.synchronize()
    with sh.Server(budget_bytes=8 << 30, streams=4, max_batch=64) as srv:
        subs = [run_one(srv, r, k) for k, r in enumerate(reqs)]
        done = [srv.wait(t) for t, _, _ in subs]
        mix = [o.cpu().numpy() for _, _, o in subs]
        st = srv.stats()
        say(f"stats {st}")
        assert st["completed"] >= len(reqs) and st["failed"] == 0
        assert st["launches"] < len(reqs), "waiting requests of one shape were not merged"
        assert max(d["group_size"] for d in done) > 1
        assert st["peak_footprint_bytes"] <= st["budget_bytes"]
        for d in done:
            assert d["latency_s"] >= d["run_s"] >= 0 and d["queue_s"] >= 0
        for k, r in enumerate(reqs):
            xo = r["xd"].cpu().numpy().astype(np.complex128)
            want = oracle_fftn(xo, r["lengths"], r["batch"], r["direction"], r["norm"])
            e = rel_l2(mix[k].astype(np.complex128), want)
            assert e <= tol(r["lengths"], r["double"]), (k, r["lengths"], r["double"], e)
            t, _, o = run_one(srv, r, 1000 + k)
            d = srv.wait(t)
            assert d["group_size"] == 1
            alone = o.cpu().numpy()
            ea = rel_l2(alone.astype(np.complex128), mix[k].astype(np.complex128))
            assert ea <= tol(r["lengths"], r["double"]), (k, r["lengths"], ea)


def test_serving_continuous_arrivals():
    """Requests arrive while earlier ones run (a submitting thread and a waiting thread);
    every one completes correctly, and later arrivals are admitted before earlier
    launches finish (their work overlaps)."""
    reqs = make_requests(64, 20260007)
    say(f"case serving continuous arrivals of {len(reqs)} requests")
    torch.cuda.synchronize()
    results = {}
    with sh.Server(budget_bytes=4 << 30, streams=4, max_batch=32) as srv:
        tickets = []
        lock = threading.Lock()

        def waiter():
            k = 0
            while k < len(reqs):
                with lock:
                    if k >= len(tickets):
                        t = None
                    else:
                        t = tickets[k]
                if t is None:
                    time.sleep(0.0005)
                    continue
                results[k] = srv.wait(t[0])
                k += 1

        th = threading.Thread(target=waiter)
        th.start()
        for k, r in enumerate(reqs):
            sub = run_one(srv, r, k)
            with lock:
                tickets.append(sub)
            time.sleep(0.0002 * (k % 4))
        th.join(timeout=120)
        assert not th.is_alive(), "waiter did not finish"
        st = srv.stats()
        say(f"stats {st}")
        assert st["completed"] >= len(reqs) and st["failed"] == 0
        for k, r in enumerate(reqs):
            xo = r["xd"].cpu().numpy().astype(np.complex128)
            want = oracle_fftn(xo, r["lengths"], r["batch"], r["direction"], r["norm"])
            got = tickets[k][2].cpu().numpy().astype(np.complex128)
            assert rel_l2(got, want) <= tol(r["lengths"], r["double"]), (k, r["lengths"])
        # overlap: some request was launched before an earlier-submitted one had completed
        overlapped = sum(1 for j in range(1, len(reqs))
                         if results[j]["launch_at"] < max(results[i]["done_at"] for i in range(j)))
        say(f"requests launched while an earlier one was still running: {overlapped}")
        assert overlapped > 0


def test_serving_rejects_bad_requests():
    say("case serving errors")
    with sh.Server(budget_bytes=1 << 20, streams=1) as srv:
        x = torch.zeros(1 << 17, dtype=torch.complex64, device="cuda")
        y = torch.zeros_like(x)
        with pytest.raises(sh.SyclfftError) as e:
            srv.submit(x, y, [1 << 17], 1)  # 1 MiB in + 1 MiB out > 1 MiB budget
        assert e.value.code == "invalid_argument"
        host = np.zeros(16, dtype=np.complex64)
        with pytest.raises(sh.SyclfftError) as e:
            srv.submit_ptr(host.ctypes.data, y.data_ptr(), [16], 1)
        assert e.value.code == "invalid_pointer"
        with pytest.raises(sh.SyclfftError) as e:
            srv.wait(987654)
        assert e.value.code == "invalid_argument"
        small = torch.zeros(256, dtype=torch.complex64, device="cuda")
        out = torch.zeros_like(small)
        t = srv.submit(small, out, [256], 1)
        srv.wait(t)
