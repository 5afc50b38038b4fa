"""The serving loop on the serving configuration's own request mix.

Requests are drawn by bench.py's `serving_requests` (the generator of the measured
stream): lengths 50% powers of two 2^8..2^20, 30% 7-smooth 240..500000, 20% primes
257..65537 (Bluestein), one in ten 2-D; batch count log-uniform 1..512 capped at 2^22
points; 80% complex<float>, 20% complex<double>; random direction and normalization.  The
server runs with the measured configuration's 64 GiB admission budget and batch limit
512.  Every request gets its own seeded input and output buffer.

Checks:
  * every request completes, none fails, the footprint stays within the budget, and
    waiting requests are merged into shared launches;
  * each request's first and last transform match the FP64 oracle within c log2(N) u
    (c = 4 for 1-D, 12 for 2-D: Bluestein and two axes);
  * each request run alone afterwards (a group of one) gives the same result as in the mix,
    within the same bound, over the whole batch.

Reference concurrency model: README.md:45-47, tests/sycl_tests.cpp:410-435.
"""
import os
import sys
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

from conftest import oracle_fftn, rel_l2

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
import syclfft_hip as sh  # noqa: E402
from bench import serving_requests  # noqa: E402

U = {False: 2.0 ** -24, True: 2.0 ** -53}
NREQ = 240


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def tol(lengths, double):
    n = int(np.prod(lengths))
    return (12.0 if len(lengths) > 1 else 4.0) * max(1.0, np.log2(n)) * U[double]


@pytest.mark.timeout(900)
def test_serving_configuration_mix():
    reqs = serving_requests(NREQ, 20260005)
    say(f"case serving configuration mix: {len(reqs)} requests, "
        f"{sum(r['n'] * r['batch'] for r in reqs)} points")
    gen = torch.Generator(device="cuda")
    gen.manual_seed(20260008)
    for r in reqs:
        rdt = torch.float64 if r["double"] else torch.float32
        x = torch.rand((r["n"] * r["batch"], 2), generator=gen, device="cuda", dtype=rdt) * 2 - 1
        r["x"] = torch.view_as_complex(x.contiguous())
        r["y"] = torch.empty_like(r["x"])
    torch.cuda.synchronize()
    with sh.Server(budget_bytes=64 << 30, streams=4, max_batch=512) as srv:
        tickets = [srv.submit(r["x"], r["y"], r["lengths"], r["batch"], r["direction"], r["norm"], tag=k)
                   for k, r in enumerate(reqs)]
        done = [srv.wait(t) for t in tickets]
        torch.cuda.synchronize()
        st = srv.stats()
        say(f"stats {st}")
        assert st["completed"] >= len(reqs) and st["failed"] == 0
        assert st["peak_footprint_bytes"] <= st["budget_bytes"]
        assert st["launches"] < len(reqs), "waiting requests were not merged"
        say(f"largest group {max(d['group_size'] for d in done)}")

        # oracle on the first and last transform of every request (the C oracle releases
        # the GIL, so lines run on several threads)
        jobs = []
        for k, r in enumerate(reqs):
            n, b = r["n"], r["batch"]
            for i in sorted({0, b - 1}):
                xi = r["x"][i * n:(i + 1) * n].cpu().numpy().astype(np.complex128)
                yi = r["y"][i * n:(i + 1) * n].cpu().numpy().astype(np.complex128)
                jobs.append((k, i, xi, yi))

        def check(job):
            k, i, xi, yi = job
            r = reqs[k]
            return k, i, rel_l2(yi, oracle_fftn(xi, r["lengths"], 1, r["direction"], r["norm"]))

        with ThreadPoolExecutor(max_workers=8) as pool:
            errs = list(pool.map(check, jobs))
        worst = max(errs, key=lambda e: e[2] / tol(reqs[e[0]]["lengths"], reqs[e[0]]["double"]))
        say(f"oracle lines checked {len(errs)}; worst request {worst[0]} line {worst[1]} "
            f"{reqs[worst[0]]['lengths']} rel-L2 {worst[2]:.3e}")
        for k, i, e in errs:
            r = reqs[k]
            assert e <= tol(r["lengths"], r["double"]), (k, i, r["lengths"], r["batch"], r["double"], e)

        # each request alone (a group of one) against its result in the mix
        for k, r in enumerate(reqs):
            alone = torch.empty_like(r["x"])
            t = srv.submit(r["x"], alone, r["lengths"], r["batch"], r["direction"], r["norm"], tag=10000 + k)
            d = srv.wait(t)
            assert d["group_size"] == 1
            torch.cuda.synchronize()
            e = rel_l2(alone.cpu().numpy(), r["y"].cpu().numpy())
            assert e <= tol(r["lengths"], r["double"]), (k, r["lengths"], r["batch"], e)
            del alone
