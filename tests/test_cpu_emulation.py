"""The provider's real plan compiler and kernel templates, compiled for the host against an
emulated HIP runtime (tests/cpuemu) with AddressSanitizer and UBSan: every global and LDS
access of every launch shape is bounds-checked and every result is compared with the FP64
oracle, without a GPU.  Cases cover each launch kind (single tile, W > 1 tiles, four-step
A/B, Bluestein fused and A/B/C, strided axes, in-place, all normalizations, both
directions, both precisions); SYCLFFT_HIP_MAX_TILE forces the multi-launch paths at small
lengths."""
import os
import subprocess

import pytest

from conftest import ROOT

EMU = os.path.join(ROOT, "tests", "cpuemu", "emu_driver")

CASES = [
    "f:8:1:-1:0:0", "f:8:1:-1:1:0", "f:8:1:-1:2:0", "f:8:1:-1:3:0", "f:7:1:1:0:1",
    "f:6:1:-1:0:0", "f:15:1:1:0:0", "f:35:1:-1:0:0", "f:11:1:1:0:1", "d:11:1:-1:3:0",
    "f:3x4:2:-1:0:0", "f:2x3x5:1:-1:0:0", "f:11x3:1:-1:0:0", "f:13:1:1:2:0",
    "f:4096:2:-1:0:0", "d:2520:2:1:3:0", "d:6561:1:-1:3:0", "f:16384:1:1:1:1",
    "d:10000:1:1:3:0", "d:8192:1:-1:0:1", "f:32x32x32:1:1:0:1", "d:64x48x10:1:-1:1:0",
    "f:7x9x11:2:1:0:0", "f:13x256:2:-1:0:0", "d:256x13:2:1:0:0", "d:64x1x8:1:1:0:0",
    "f:257:3:1:2:0", "d:1009:2:-1:3:1", "f:1:5:-1:3:0", "f:262144:1:-1:0:0",
    "d:65536:1:1:2:0", "f:10007:1:-1:3:0",
]
CAPPED = [
    "f:4096:2:-1:0:0", "d:4096:1:1:3:1", "f:3000:2:1:1:0", "f:1009:2:-1:0:0",
    "d:1009:1:1:3:1", "f:4099:1:-1:0:0", "f:2048x96:1:-1:1:0", "d:96x2048:1:1:0:1",
    "d:1031x3:2:-1:0:0", "f:3x1031:2:1:3:0",
]


def _run(cases, env_extra=None):
    if not os.path.exists(EMU):
        pytest.fail(f"{EMU} is not built (run make -C tests/cpuemu, or __graft_entry__.build())")
    env = dict(os.environ)
    env["ASAN_OPTIONS"] = "detect_leaks=0:abort_on_error=0"
    env.update(env_extra or {})
    r = subprocess.run([EMU] + cases, capture_output=True, text=True, timeout=900, env=env)
    lines = r.stdout.splitlines()
    bad = [l for l in lines if not l.startswith("ok ")]
    assert r.returncode == 0 and not bad and len(lines) == len(cases), (r.stdout[-4000:], r.stderr[-4000:])


def test_emulated_launch_kinds():
    _run(CASES)


# split-precision MFMA plans (b: BF16 hi/lo, h: scaled FP16 hi/lo): the MFMA products and
# the lane exchange are emulated with the v_mfma_f32_16x16x32 lane layout; shapes cover
# W = 1, W > 1 with row-major and column-fast lanes, four-step A/B, multi-D
SPLIT = ["b:16:3:-1:0:0", "h:16:3:1:3:0", "h:256:5:1:3:1", "b:4096:1:1:1:0", "h:4096:1:-1:0:0",
         "h:64x1x8:3:1:0:0", "b:80x48:2:1:0:0", "h:256x16:3:1:1:0"]


def test_emulated_split_precision():
    _run(SPLIT)
    _run(["h:4096:2:-1:0:0", "b:2048x96:1:-1:1:0"], {"SYCLFFT_HIP_MAX_TILE": "1024"})
    _run(["h:4096:3:-1:0:0", "b:256x16:3:1:1:0"], {"EMU_MANY": "1"})


def test_emulated_multi_launch_paths():
    _run(CAPPED, {"SYCLFFT_HIP_MAX_TILE": "1024"})


CHUNKED = ["f:65536:9:-1:0:0", "d:32768:5:1:3:1", "f:5x65536:1:1:1:0", "f:65536x3:2:-1:0:0"]
CHUNKED_CAPPED = ["f:4099:40:1:3:0", "d:1031x3:40:1:1:1"]


def test_emulated_chunked_launch_groups():
    """Multi-launch axes split into launch groups of ~1 MiB (SYCLFFT_HIP_CHUNK_MB): the
    direction conjugations and the scale must reach every group."""
    _run(CHUNKED, {"SYCLFFT_HIP_CHUNK_MB": "1"})
    _run(CHUNKED_CAPPED, {"SYCLFFT_HIP_CHUNK_MB": "1", "SYCLFFT_HIP_MAX_TILE": "1024"})


TUNED = ["f:4096:9:-1:3:0", "d:2520:5:1:3:1", "f:10007:3:1:1:0", "f:32x64x16:2:-1:0:1", "d:16384:2:1:2:0"]


@pytest.mark.parametrize("tune", ["s1w0p0g0", "s0w-1p0g0", "s0w1p0g0", "s0w0p1g0", "s0w0p0g1", "s1w1p1g1"])
def test_emulated_tuning_choices(tune):
    """Every planner choice the autotuner can make gives a correct, in-bounds plan."""
    _run(TUNED, {"SYCLFFT_HIP_TUNE": tune})


def test_emulated_measure_mode_writes_cache(tmp_path):
    cache = tmp_path / "tune.txt"
    _run(["f:4096:9:-1:3:0"], {"EMU_MEASURE": "1", "SYCLFFT_HIP_TUNE_CACHE": str(cache)})
    text = cache.read_text()
    assert "f32" in text and "4096x1x1|9" in text
    key, value = text.strip().split("\t")
    assert value.startswith("s") and "w" in value


MANY = ["f:4096:3:-1:0:0", "f:300:5:1:3:0", "d:2520:3:1:3:1", "f:10007:3:-1:3:0", "f:32768:3:-1:0:0",
        "d:1009:3:-1:3:1", "f:13x256:3:-1:0:0", "d:256x13:3:1:0:0", "f:32x32x32:3:1:0:1",
        "d:64x48x10:3:-1:1:0", "f:16384:1:1:1:1", "d:4096x8:2:1:0:1"]

# column-tile length-specialized kernel (f64 4096-point strided axes, 4 columns per tile):
# 2-D and 3-D, in and out of place, every normalization
COLS = ["d:4096x4:2:-1:0:0", "d:4096x8:1:1:3:1", "d:4096x12:1:-1:1:0", "d:4096x2x4:1:1:2:0", "d:4096x4:3:1:3:0",
        # ragged last tiles (33 and 65 columns), fp32 2048 x 8 columns, multi-D real plans over them
        "d:4096x33:1:-1:0:0", "c:2048x65:1:-1:1:0", "c:2048x8:2:1:3:1", "R:2048x128:2:-1:0:0",
        "R:2048x128:2:1:2:0",
        # 64 and 32 tiles: every tile goes through the XCD group remap (8 and 4 tiles per group)
        "c:2048x512:1:-1:0:0", "d:4096x128:1:1:3:0"]


def test_emulated_column_tiles():
    _run(COLS)


@pytest.mark.parametrize("mode", ["1", "2", "3"])
def test_emulated_pointer_array_execution(mode):
    """syclfft_hip_plan_execute_many: every transform in its own allocation (mode 1), a
    count below the plan's batch count (mode 2), and pointers into one contiguous batch
    (mode 3, no unit tables), across single, four-step, fused and three-launch Bluestein
    and multi-D plans."""
    _run(MANY, {"EMU_MANY": mode})
    _run(["f:10007:3:-1:3:0", "d:1009:5:1:2:1", "f:4096:3:1:0:0", "d:1031x3:3:-1:0:0"],
         {"EMU_MANY": mode, "SYCLFFT_HIP_TUNE": "s1w0p0g0"})


EMU_SERVER = os.path.join(os.path.dirname(EMU), "emu_server")


@pytest.mark.parametrize("env", [{}, {"SYCLFFT_HIP_SERVE_SJF": "1", "SYCLFFT_HIP_SERVE_SJF_AGE_US": "200"},
                                 {"SYCLFFT_HIP_SERVE_QUEUE": "0"}])
def test_emulated_serving_loop(env):
    """The serving loop on the emulated device: a mix of 36 requests (6 shapes, both
    types, directions, normalizations, in-place and out-of-place) submitted at once is
    merged into shared launches, each result matches the FP64 oracle and the same request
    run alone; budget, ticket and host-buffer errors are rejected."""
    if not os.path.exists(EMU_SERVER):
        pytest.fail(f"{EMU_SERVER} is not built (run make -C tests/cpuemu)")
    run_env = dict(os.environ)
    run_env["ASAN_OPTIONS"] = "detect_leaks=0:abort_on_error=0"
    run_env.update(env)  # default dispatch, size-first dispatch, no work-queue merging
    r = subprocess.run([EMU_SERVER], capture_output=True, text=True, timeout=900, env=run_env)
    lines = r.stdout.splitlines()
    bad = [l for l in lines if not l.startswith("ok ")]
    assert r.returncode == 0 and not bad and len(lines) > 100, (r.stdout[-4000:], r.stderr[-4000:])


HALVES = ["f:10007:2:-1:0:0", "f:10007:3:1:3:1", "d:1009:3:-1:1:0", "d:1031:2:1:2:1", "f:13:5:1:0:0",
          "f:257:4:-1:3:0", "f:11x13:2:1:0:0", "d:17x2:3:-1:3:0"]


def test_emulated_bluestein_two_workgroups():
    """Bluestein over two workgroups per transform (the small-batch form), forced on for
    every transform size, in both precisions and directions, every normalization,
    in place and out of place, and through pointer arrays."""
    _run(HALVES, {"SYCLFFT_HIP_BLUE_HALVES": "1"})
    _run(["f:10007:3:-1:3:0", "d:1009:3:1:1:1"], {"SYCLFFT_HIP_BLUE_HALVES": "1", "EMU_MANY": "1"})
    _run(["f:1009:5:-1:3:0"], {"SYCLFFT_HIP_BLUE_HALVES": "1", "EMU_MANY": "2"})


def test_emulated_work_queue_execution():
    """One syclfft_hip_execute_queue call over plans of mixed lengths, batch counts, precisions,
    directions, normalizations and placements: single-launch plain plans share persistent
    work-queue launches per kernel variant (the emulated device counter hands out their
    tiles), the others (Bluestein, four-step, multi-D) execute first on the same stream."""
    _run(["f:256:5:-1:0:0", "f:240:3:1:3:0", "f:4096:2:-1:1:1", "d:2520:2:1:3:0", "f:16:33:-1:0:0",
          "f:10007:2:-1:0:0", "d:64:7:-1:0:1", "f:512:4:1:0:0", "f:1000:3:-1:0:0", "f:32x32:2:1:2:0",
          "f:65536:1:-1:0:0"], {"EMU_QUEUE": "1"})


def test_emulated_bluestein_pair():
    """The length-specialized Bluestein kernel over two workgroups per transform (forced with
    SYCLFFT_HIP_BLUE_HALVES=1 where the emulation reports no CU count): every direction and
    normalization, in place, lengths whose half length is 10240, contiguous and over pointer
    arrays (the second half of each pair to finish adds the two partial results)."""
    env = {"SYCLFFT_HIP_BLUE_HALVES": "1"}
    cases = ["f:10007:3:-1:3:0", "f:10007:2:1:0:1", "f:9973:2:-1:1:0", "f:10079:1:1:2:0"]
    _run(cases, env)
    _run(cases[:2], dict(env, EMU_MANY="2"))


def test_emulated_global_table_twin():
    """The f64 variant twins that read their twiddle tables from global memory (no tables in
    LDS), forced on every launch that has a twin: single tiles of 2049..5120 points in every
    normalization and direction, in place, a multi-D axis, and a work-queue launch of twin
    plans."""
    env = {"SYCLFFT_HIP_TW_GLOBAL": "1"}
    _run(["d:2520:3:-1:3:0", "d:2520:2:1:0:1", "d:2187:2:-1:1:0", "d:4096:2:-1:0:0", "d:3000:2:1:2:0",
          "d:48x2520:1:-1:0:0", "d:2560:3:1:3:0", "d:4608:2:-1:0:0", "d:5040:2:1:3:0", "d:4097:1:-1:1:1"], env)
    _run(["d:2520:2:1:3:0", "d:3000:3:-1:0:0", "f:256:5:-1:0:0"], dict(env, EMU_QUEUE="1"))


def test_emulated_real_plans():
    """R2C (N real -> N/2 + 1 complex) and C2R plans ('R' float, 'Q' double) against the
    complex FP64 oracle: the half-length complex plan over the real data viewed as complex,
    single tile, Bluestein (N/2 prime) and four-step half plans, every normalization; and
    2-D / 3-D real plans (C2R inputs made exactly Hermitian)."""
    _run(["R:2:3:-1:0:0", "R:16:4:-1:1:0", "R:4096:2:-1:0:0", "Q:2520:2:-1:3:0", "R:30:3:1:0:0",
          "R:4096:2:1:2:0", "Q:1000:2:1:3:0", "R:20014:1:-1:0:0", "Q:2:2:1:0:0", "R:65536:1:1:1:0",
          # multi-D: the last axis as 1-D real rows, the leading axes over the half spectrum
          "R:8x6:2:-1:0:0", "Q:8x6:2:1:3:0", "R:12x10x8:1:-1:2:0", "Q:5x14:3:1:1:0", "R:32x64:2:1:0:0",
          "Q:7x9x10:1:-1:3:0", "R:3x1000:2:1:2:0",
          # fused: the untangle in the last pass / the tangle in the first pass of the
          # length-specialized half-length FFT (M = 1024, 2048, 4096, 8192)
          "R:2048:3:-1:0:0", "R:2048:3:1:1:0", "Q:4096:3:-1:3:0", "Q:4096:2:1:3:0", "R:16384:1:-1:1:0",
          "Q:8192:2:-1:2:0"])


def test_emulated_layout_plans():
    """Layout plans ('L' float, 'M' double; ...:istride:idist:ostride:odist:iplanar:oplanar)
    against the FP64 oracle over the gathered data, with sentinels in the layouts' gaps:
    rank-1 interleaved layouts addressed by the one tile launch (single tile, in place,
    Bluestein fused, N = 16384), interleaved batches planned as strided axes (four-step,
    Bluestein fused and three-launch, 2-D, 3-D), and the gather / in-place core / scatter
    path (planar sides, 2-D, a three-launch Bluestein core)."""
    _run(["L:256:4:-1:0:0:2:600:3:900:0:0", "L:64:8:1:1:0:8:1:8:1:0:0", "M:120:3:-1:3:1:2:300:2:300:0:0",
          "L:257:3:1:2:0:3:800:1:300:0:0", "L:16384:1:1:1:0:2:40000:1:16384:0:0",
          "L:4096:2:-1:0:0:1:4096:1:4096:1:0", "M:2520:2:1:3:0:1:2600:2:5100:0:1", "L:1000:3:-1:2:1:1:1000:1:1000:1:1",
          "M:12x10:2:-1:0:0:3:400:1:120:0:0", "L:64:4:1:1:0:5:400:1:64:1:1", "M:10007:1:-1:3:1:2:20014:2:20014:0:0",
          # interleaved batches (stride = batch, distance 1): strided-axis plans over the caller's buffers
          "M:12x10:4:-1:0:1:4:1:4:1:0:0", "L:16384:4:1:0:0:4:1:4:1:0:0", "L:257:3:-1:2:0:3:1:3:1:0:0",
          "M:10007:2:1:3:0:2:1:2:1:0:0", "L:5x7x9:3:1:1:1:3:1:3:1:0:0"])
