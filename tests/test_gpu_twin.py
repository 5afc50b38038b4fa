"""GPU tests of the f64 global-table twin variants: batched f64 lengths whose tiles fit more
workgroups per CU only without twiddle tables in LDS and at 128 VGPRs (2049..2560 points:
four instead of three; 4097..5120 points: two instead of one) run the twin once their
tiles outnumber the LDS-table variant's resident slots.  Results are checked
against the FP64 oracle (relative L2 <= 2 log2(N) u) and against the same plan built with
the twin switched off (SYCLFFT_HIP_TW_GLOBAL=0): the twin reads the same table values and
runs the same arithmetic, so the outputs are bit-identical.  The length-specialized kernels
are switched off here (SYCLFFT_HIP_FIXED=0): they take some of these lengths (2520) first,
and the twin is a form of the generic kernel.
"""
import os
import re
import sys

import numpy as np
import pytest

from conftest import oracle_fftn, rel_l2

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

import syclfft_hip as sh  # noqa: E402

U64 = 2.0 ** -53


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def run(n, batch, direction, norm, x, twin_env):
    old = os.environ.get("SYCLFFT_HIP_TW_GLOBAL")
    old_fixed = os.environ.get("SYCLFFT_HIP_FIXED")
    if twin_env is None:
        os.environ.pop("SYCLFFT_HIP_TW_GLOBAL", None)
    else:
        os.environ["SYCLFFT_HIP_TW_GLOBAL"] = twin_env
    os.environ["SYCLFFT_HIP_FIXED"] = "0"
    try:
        p = sh.Plan([n], batch, direction, double=True, normalization=norm)
    finally:
        if old is None:
            os.environ.pop("SYCLFFT_HIP_TW_GLOBAL", None)
        else:
            os.environ["SYCLFFT_HIP_TW_GLOBAL"] = old
        if old_fixed is None:
            os.environ.pop("SYCLFFT_HIP_FIXED", None)
        else:
            os.environ["SYCLFFT_HIP_FIXED"] = old_fixed
    desc = p.describe().strip()
    xd = torch.from_numpy(x).to("cuda")
    yd = torch.empty_like(xd)
    p.execute(xd, yd)
    torch.cuda.synchronize()
    y = yd.cpu().numpy()
    p.close()
    return y, desc


@pytest.mark.parametrize("n,batch,direction,norm", [
    (2520, 1024, -1, "orthogonal"),
    (2187, 1024, 1, "none"),
    (2400, 2048, -1, "forward"),
    (2560, 1024, 1, "backward"),
    (4608, 1024, -1, "none"),
    (5040, 512, 1, "orthogonal"),
    (4374, 1024, -1, "forward"),
])
def test_twin_matches_oracle_and_lds_tables(n, batch, direction, norm):
    say(f"case twin n={n} batch={batch} dir={direction} norm={norm}")
    g = np.random.default_rng(20260100 + n)
    x = g.uniform(-1, 1, n * batch) + 1j * g.uniform(-1, 1, n * batch)
    y_auto, d_auto = run(n, batch, direction, norm, x, None)
    y_off, d_off = run(n, batch, direction, norm, x, "0")
    say(f"  auto: {d_auto}")
    say(f"  off:  {d_off}")
    assert "tw=global" not in d_off
    lp = int(re.search(r"lp=(\d+)", d_off).group(1))  # LDS slots of the tile (pads included)
    nt = int(re.search(r"nt=(\d+)", d_off).group(1))
    if (nt == 256 and lp * 16 <= 40960) or (nt == 512 and lp * 16 <= 81920):
        # 256 threads, data within 40 KiB: four per CU without tables (three with them);
        # 512 threads, data within 80 KiB: two per CU (one at the LDS-table variant's registers)
        assert "tw=global" in d_auto
    want = oracle_fftn(x, [n], batch, direction, norm)
    e = rel_l2(y_auto, want)
    say(f"  rel-L2 vs oracle {e:.3e}")
    assert e <= 2 * np.log2(n) * U64
    assert np.array_equal(y_auto, y_off), "twin and LDS-table variant differ"
