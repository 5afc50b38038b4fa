"""CPU tests of the C-ABI boundary: the library loads, exports every declared symbol, and
validates plan configurations the way the reference does before any device work
(checked_element_count, src/core/provider_loader.cpp:112-132; direction check,
src/sycl/sycl.cpp:435-438).  No compute call is made here."""
import ctypes
import glob
import os
import re

import pytest

from conftest import ROOT

import syclfft_hip as sh

HEADERS = sorted(glob.glob(os.path.join(ROOT, "include", "*.h")))


def declared_symbols():
    names = set()
    for h in HEADERS:
        text = "\n".join(l for l in open(h).read().splitlines() if not l.lstrip().startswith("#"))
        names.update(re.findall(r"SYCLFFT_HIP_API[^;(]*?\b(\w+)\s*\(", text))
    return sorted(names)


def test_header_symbols_match_binding_list():
    assert declared_symbols() == sorted(sh.EXPORTED_SYMBOLS)


def test_library_exports_every_declared_symbol():
    lib = sh.load()
    for name in declared_symbols():
        assert hasattr(lib, name), name


def test_provider_table():
    t = sh.provider_table()
    assert t.abi_version == sh.ABI_VERSION
    assert t.name == b"hip_mi355x"
    for f in ("supports", "create", "destroy", "scratch_bytes", "execute"):
        assert ctypes.cast(getattr(t, f), ctypes.c_void_p).value


def test_config_struct_layout():
    # fixed C layout: 4 x u32, 3 x u64 lengths, u64 batch, 5 x u32/i32, 3 x u32 reserved
    assert ctypes.sizeof(sh.PlanConfig) == 16 + 24 + 8 + 20 + 12
    assert sh.PlanConfig.lengths.offset == 16
    assert sh.PlanConfig.batch_count.offset == 40


def _create_status(cfg):
    lib = sh.load()
    h = ctypes.c_void_p()
    err = ctypes.create_string_buffer(256)
    st = lib.syclfft_hip_plan_create_ex(ctypes.byref(cfg), ctypes.byref(h), err, 256)
    if h.value:
        lib.syclfft_hip_plan_destroy(h)
    return st, err.value.decode()


@pytest.mark.parametrize("mutate,status,text", [
    (lambda c: setattr(c, "rank", 0), 1, "rank"),
    (lambda c: setattr(c, "rank", 4), 1, "rank"),
    (lambda c: setattr(c, "batch_count", 0), 1, "batch"),
    (lambda c: c.lengths.__setitem__(0, 0), 1, "non-zero"),
    (lambda c: (setattr(c, "rank", 2), c.lengths.__setitem__(0, 2 ** 63), c.lengths.__setitem__(1, 4)), 1, "overflow"),
    (lambda c: setattr(c, "direction", 0), 1, "direction"),
    (lambda c: setattr(c, "normalization", 9), 1, "normalization"),
    (lambda c: setattr(c, "scalar", 3), 4, "scalar"),
    (lambda c: setattr(c, "abi_version", 2), 8, "ABI"),
    (lambda c: (setattr(c, "scalar", 2), setattr(c, "precision", 1)), 4, "split"),
])
def test_invalid_configs_rejected(mutate, status, text):
    cfg = sh.make_config([8], 1)
    mutate(cfg)
    st, msg = _create_status(cfg)
    assert st == status, msg
    assert text.lower() in msg.lower()


@pytest.mark.parametrize("mutate,status,text", [
    (lambda c: c.lengths.__setitem__(0, 7), 1, "even"),
    (lambda c: (setattr(c, "rank", 2), c.lengths.__setitem__(1, 5)), 1, "even last length"),
    (lambda c: setattr(c, "in_place", 1), 1, "out of place"),
    (lambda c: setattr(c, "precision", 2), 4, "VALU"),
    (lambda c: setattr(c, "batch_count", 0), 1, "batch"),
])
def test_invalid_real_configs_rejected(mutate, status, text):
    """Real (R2C/C2R) plans: even last length, out of place, VALU precision; checked before
    any device work."""
    lib = sh.load()
    cfg = sh.make_config([8], 1)
    mutate(cfg)
    h = ctypes.c_void_p()
    err = ctypes.create_string_buffer(256)
    st = lib.syclfft_hip_real_plan_create(ctypes.byref(cfg), ctypes.byref(h), err, 256)
    if h.value:
        lib.syclfft_hip_plan_destroy(h)
    assert st == status, err.value
    assert text.lower() in err.value.decode().lower()


@pytest.mark.parametrize("field,value,text", [
    ("in_stride", 0, "strides"),
    ("out_distance", 0, "strides"),
    ("in_distance", 1 << 40, "strides"),
    ("in_planar", 2, "planar"),
])
def test_invalid_layouts_rejected(field, value, text):
    """Layout plans: strides and distances in [1, 2^40), planar flags 0 / 1; checked before
    any device work (INVALID_ARGUMENT)."""
    lib = sh.load()
    cfg = sh.make_config([8], 2)
    lay = sh.Layout(1, 8, 1, 8, 0, 0)
    setattr(lay, field, value)
    h = ctypes.c_void_p()
    err = ctypes.create_string_buffer(256)
    st = lib.syclfft_hip_layout_plan_create(ctypes.byref(cfg), ctypes.byref(lay), ctypes.byref(h), err, 256)
    if h.value:
        lib.syclfft_hip_plan_destroy(h)
    assert st == 1, err.value
    assert text in err.value.decode().lower()


def test_layout_null_arguments():
    lib = sh.load()
    err = ctypes.create_string_buffer(64)
    cfg = sh.make_config([8], 1)
    h = ctypes.c_void_p()
    assert lib.syclfft_hip_layout_plan_create(ctypes.byref(cfg), None, ctypes.byref(h), err, 64) == 1
    assert b"layout is null" in err.value and not h.value
    assert lib.syclfft_hip_layout_execute(None, None, None, None, None, None, None, 0, None, err, 64) == 2


def test_null_arguments():
    lib = sh.load()
    err = ctypes.create_string_buffer(64)
    assert lib.syclfft_hip_plan_create(None, err, 64) is None
    assert b"non-null" in err.value
    # null plan handle
    st = lib.syclfft_hip_plan_execute(None, None, None, None, None, 0, None, err, 64)
    assert st == 2
    lib.syclfft_hip_plan_destroy(None)
    assert lib.syclfft_hip_plan_scratch_bytes(None) == 0


def test_error_buffer_truncation():
    lib = sh.load()
    cfg = sh.make_config([8], 1)
    cfg.rank = 0
    err = ctypes.create_string_buffer(b"\xff" * 8, 8)
    h = ctypes.c_void_p()
    st = lib.syclfft_hip_plan_create_ex(ctypes.byref(cfg), ctypes.byref(h), err, 8)
    assert st == 1
    assert err.raw[7] == 0 and len(err.value) == 7


def test_no_gpu_reports_unavailable():
    import torch
    if torch.cuda.is_available():
        pytest.skip("GPU present")
    ok, reason = sh.supports(-1)
    assert not ok and reason
    with pytest.raises(sh.SyclfftError) as e:
        sh.Plan([8])
    assert e.value.code == "provider_unavailable"


def test_server_struct_layout():
    # syclfft_hip_request_v1: 4 x u32, 3 x u64 lengths, u64 batch, 2 x u32, 2 pointers, u64 tag
    assert ctypes.sizeof(sh.Request) == 16 + 24 + 8 + 8 + 16 + 8
    assert sh.Request.lengths.offset == 16 and sh.Request.inp.offset == 56 and sh.Request.tag.offset == 72
    assert ctypes.sizeof(sh.Completion) == 8 + 8 + 4 + 4 + 6 * 8
    assert ctypes.sizeof(sh.ServerStats) == 9 * 8


def test_server_without_gpu_reports_unavailable():
    import torch

    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    with pytest.raises(sh.SyclfftError) as e:
        sh.Server(1 << 30)
    assert e.value.code == "provider_unavailable"
