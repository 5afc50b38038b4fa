"""ctypes binding of the MI355X syclfft provider's C ABI (include/syclfft_hip_provider.h).

Mirrors the reference plan API's meaning (include/syclfft/syclfft.hpp:45-129,
include/syclfft/common.hpp:34-110): a plan is built once for (lengths, batch, direction,
placement, normalization) and executed on device buffers; errors raise `SyclfftError`
carrying the reference's `error_code` name.  Device memory and streams come from PyTorch
(tensors' data_ptr() and torch.cuda streams); no computation happens in Python.

The provider library is found in ./lib/syclfft/providers next to this file (built by
`make` in this directory), or at $SYCLFFT_HIP_PROVIDER.  Loading fails loudly if it is
missing: there is no fallback path.
"""
from __future__ import annotations

import ctypes
import os
from typing import Iterable, Optional, Sequence

_HERE = os.path.dirname(os.path.abspath(__file__))
PROVIDER_PATH = os.path.join(_HERE, "lib", "syclfft", "providers", "syclfft_provider_hip.so")

ABI_VERSION = 1
SCALAR_FLOAT = 1
SCALAR_DOUBLE = 2
FORWARD = -1
BACKWARD = 1
NORM = {"none": 0, "forward": 1, "backward": 2, "orthogonal": 3}
PRECISION = {"valu": 0, "split_bf16": 1, "split_fp16": 2}

# status = 1 + syclfft::error_code (include/syclfft/common.hpp:83-93)
ERROR_CODES = {
    1: "invalid_argument",
    2: "invalid_state",
    3: "invalid_pointer",
    4: "unsupported_precision",
    5: "provider_unavailable",
    6: "planning_failed",
    7: "execution_failed",
    8: "plugin_error",
}

# every symbol include/syclfft_hip_provider.h declares
EXPORTED_SYMBOLS = (
    "syclfft_get_device_provider_v1",
    "syclfft_hip_supports",
    "syclfft_hip_plan_create",
    "syclfft_hip_plan_create_ex",
    "syclfft_hip_real_plan_create",
    "syclfft_hip_layout_plan_create",
    "syclfft_hip_layout_execute",
    "syclfft_hip_plan_destroy",
    "syclfft_hip_plan_scratch_bytes",
    "syclfft_hip_plan_execute",
    "syclfft_hip_plan_execute_noalloc",
    "syclfft_hip_noalloc_scratch_limit",
    "syclfft_hip_plan_execute_many",
    "syclfft_hip_execute_queue",
    "syclfft_hip_plan_describe",
    "syclfft_hip_plan_launch_count",
    "syclfft_hip_server_create",
    "syclfft_hip_server_submit",
    "syclfft_hip_server_wait",
    "syclfft_hip_server_drain",
    "syclfft_hip_server_stats",
    "syclfft_hip_server_replay",
    "syclfft_hip_server_destroy",
)


class SyclfftError(RuntimeError):
    """Error raised by the provider; `code` is the syclfft::error_code name."""

    def __init__(self, code: str, message: str):
        super().__init__(f"{code}: {message}")
        self.code = code
        self.message = message


class Layout(ctypes.Structure):
    """syclfft_hip_layout_v1: element x of transform b at b * distance + x * stride (complex
    elements for interleaved sides, real elements of each array for planar sides)."""
    _fields_ = [
        ("in_stride", ctypes.c_uint64),
        ("in_distance", ctypes.c_uint64),
        ("out_stride", ctypes.c_uint64),
        ("out_distance", ctypes.c_uint64),
        ("in_planar", ctypes.c_uint32),
        ("out_planar", ctypes.c_uint32),
    ]


class PlanConfig(ctypes.Structure):
    _fields_ = [
        ("abi_version", ctypes.c_uint32),
        ("scalar", ctypes.c_uint32),
        ("direction", ctypes.c_int32),
        ("rank", ctypes.c_uint32),
        ("lengths", ctypes.c_uint64 * 3),
        ("batch_count", ctypes.c_uint64),
        ("in_place", ctypes.c_uint32),
        ("normalization", ctypes.c_uint32),
        ("measure", ctypes.c_uint32),
        ("precision", ctypes.c_uint32),
        ("device", ctypes.c_int32),
        ("reserved", ctypes.c_uint32 * 3),
    ]


_SUPPORTS = ctypes.CFUNCTYPE(ctypes.c_int, ctypes.c_int, ctypes.c_char_p, ctypes.c_size_t)
_CREATE = ctypes.CFUNCTYPE(ctypes.c_void_p, ctypes.POINTER(PlanConfig), ctypes.c_char_p, ctypes.c_size_t)
_DESTROY = ctypes.CFUNCTYPE(None, ctypes.c_void_p)
_SCRATCH = ctypes.CFUNCTYPE(ctypes.c_uint64, ctypes.c_void_p)
_EXECUTE = ctypes.CFUNCTYPE(
    ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
    ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t, ctypes.c_void_p, ctypes.c_char_p,
    ctypes.c_size_t)


class ProviderTable(ctypes.Structure):
    _fields_ = [
        ("abi_version", ctypes.c_uint32),
        ("name", ctypes.c_char_p),
        ("supports", _SUPPORTS),
        ("create", _CREATE),
        ("destroy", _DESTROY),
        ("scratch_bytes", _SCRATCH),
        ("execute", _EXECUTE),
    ]


_lib = None


def load(path: Optional[str] = None) -> ctypes.CDLL:
    """Load the provider library (raises OSError if it is not built)."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = path or os.environ.get("SYCLFFT_HIP_PROVIDER", PROVIDER_PATH)
    if not os.path.exists(p):
        raise OSError(f"syclfft HIP provider not built: {p} is missing (run make in {_HERE})")
    lib = ctypes.CDLL(p, mode=ctypes.RTLD_LOCAL)
    lib.syclfft_hip_supports.argtypes = [ctypes.c_int, ctypes.c_char_p, ctypes.c_size_t]
    lib.syclfft_hip_supports.restype = ctypes.c_int
    lib.syclfft_hip_plan_create.argtypes = [ctypes.POINTER(PlanConfig), ctypes.c_char_p, ctypes.c_size_t]
    lib.syclfft_hip_plan_create.restype = ctypes.c_void_p
    lib.syclfft_hip_plan_create_ex.argtypes = [
        ctypes.POINTER(PlanConfig), ctypes.POINTER(ctypes.c_void_p), ctypes.c_char_p, ctypes.c_size_t]
    lib.syclfft_hip_plan_create_ex.restype = ctypes.c_int
    lib.syclfft_hip_real_plan_create.argtypes = [ctypes.POINTER(PlanConfig), ctypes.POINTER(ctypes.c_void_p),
                                                 ctypes.c_char_p, ctypes.c_size_t]
    lib.syclfft_hip_real_plan_create.restype = ctypes.c_int
    lib.syclfft_hip_layout_plan_create.argtypes = [ctypes.POINTER(PlanConfig), ctypes.POINTER(Layout),
                                                   ctypes.POINTER(ctypes.c_void_p), ctypes.c_char_p, ctypes.c_size_t]
    lib.syclfft_hip_layout_plan_create.restype = ctypes.c_int
    lib.syclfft_hip_layout_execute.argtypes = [
        ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
        ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t, ctypes.c_void_p, ctypes.c_char_p, ctypes.c_size_t]
    lib.syclfft_hip_layout_execute.restype = ctypes.c_int
    lib.syclfft_hip_plan_destroy.argtypes = [ctypes.c_void_p]
    lib.syclfft_hip_plan_destroy.restype = None
    lib.syclfft_hip_plan_scratch_bytes.argtypes = [ctypes.c_void_p]
    lib.syclfft_hip_plan_scratch_bytes.restype = ctypes.c_uint64
    lib.syclfft_hip_plan_execute.argtypes = [
        ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
        ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t, ctypes.c_void_p, ctypes.c_char_p,
        ctypes.c_size_t]
    lib.syclfft_hip_plan_execute.restype = ctypes.c_int
    lib.syclfft_hip_plan_execute_noalloc.argtypes = lib.syclfft_hip_plan_execute.argtypes
    lib.syclfft_hip_plan_execute_noalloc.restype = ctypes.c_int
    lib.syclfft_hip_noalloc_scratch_limit.argtypes = []
    lib.syclfft_hip_noalloc_scratch_limit.restype = ctypes.c_uint64
    lib.syclfft_hip_plan_execute_many.argtypes = [
        ctypes.c_void_p, ctypes.c_uint64, ctypes.POINTER(ctypes.c_void_p), ctypes.POINTER(ctypes.c_void_p),
        ctypes.c_void_p, ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t, ctypes.c_void_p, ctypes.c_char_p,
        ctypes.c_size_t]
    lib.syclfft_hip_plan_execute_many.restype = ctypes.c_int
    lib.syclfft_hip_execute_queue.argtypes = [
        ctypes.POINTER(ctypes.c_void_p), ctypes.POINTER(ctypes.c_void_p), ctypes.POINTER(ctypes.c_void_p),
        ctypes.c_size_t, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_char_p, ctypes.c_size_t]
    lib.syclfft_hip_execute_queue.restype = ctypes.c_int
    lib.syclfft_hip_plan_describe.argtypes = [ctypes.c_void_p, ctypes.c_char_p, ctypes.c_size_t]
    lib.syclfft_hip_plan_describe.restype = ctypes.c_size_t
    lib.syclfft_hip_plan_launch_count.argtypes = [ctypes.c_void_p]
    lib.syclfft_hip_plan_launch_count.restype = ctypes.c_uint32
    lib.syclfft_hip_server_create.argtypes = [ctypes.c_int, ctypes.c_uint64, ctypes.c_uint32, ctypes.c_uint64,
                                              ctypes.c_char_p, ctypes.c_size_t]
    lib.syclfft_hip_server_create.restype = ctypes.c_void_p
    lib.syclfft_hip_server_submit.argtypes = [ctypes.c_void_p, ctypes.POINTER(Request),
                                              ctypes.POINTER(ctypes.c_uint64), ctypes.c_char_p, ctypes.c_size_t]
    lib.syclfft_hip_server_submit.restype = ctypes.c_int
    lib.syclfft_hip_server_wait.argtypes = [ctypes.c_void_p, ctypes.c_uint64, ctypes.POINTER(Completion),
                                            ctypes.c_char_p, ctypes.c_size_t]
    lib.syclfft_hip_server_wait.restype = ctypes.c_int
    lib.syclfft_hip_server_drain.argtypes = [ctypes.c_void_p, ctypes.c_char_p, ctypes.c_size_t]
    lib.syclfft_hip_server_drain.restype = ctypes.c_int
    lib.syclfft_hip_server_stats.argtypes = [ctypes.c_void_p, ctypes.POINTER(ServerStats)]
    lib.syclfft_hip_server_stats.restype = None
    lib.syclfft_hip_server_replay.argtypes = [ctypes.c_void_p, ctypes.POINTER(Request),
                                              ctypes.POINTER(ctypes.c_double), ctypes.c_size_t, ctypes.c_int,
                                              ctypes.POINTER(Completion), ctypes.c_char_p, ctypes.c_size_t]
    lib.syclfft_hip_server_replay.restype = ctypes.c_int
    lib.syclfft_hip_server_destroy.argtypes = [ctypes.c_void_p]
    lib.syclfft_hip_server_destroy.restype = None
    lib.syclfft_get_device_provider_v1.argtypes = []
    lib.syclfft_get_device_provider_v1.restype = ctypes.POINTER(ProviderTable)
    if path is None:
        _lib = lib
    return lib


SERVER_ABI_VERSION = 1


class Request(ctypes.Structure):
    """syclfft_hip_request_v1 (include/syclfft_hip_server.h)."""
    _fields_ = [
        ("abi_version", ctypes.c_uint32),
        ("scalar", ctypes.c_uint32),
        ("direction", ctypes.c_int32),
        ("rank", ctypes.c_uint32),
        ("lengths", ctypes.c_uint64 * 3),
        ("batch_count", ctypes.c_uint64),
        ("normalization", ctypes.c_uint32),
        ("in_place", ctypes.c_uint32),
        ("inp", ctypes.c_void_p),
        ("out", ctypes.c_void_p),
        ("tag", ctypes.c_uint64),
    ]


class Completion(ctypes.Structure):
    """syclfft_hip_completion_v1."""
    _fields_ = [
        ("ticket", ctypes.c_uint64),
        ("tag", ctypes.c_uint64),
        ("status", ctypes.c_int32),
        ("group_size", ctypes.c_uint32),
        ("queue_seconds", ctypes.c_double),
        ("run_seconds", ctypes.c_double),
        ("latency_seconds", ctypes.c_double),
        ("submit_at", ctypes.c_double),
        ("launch_at", ctypes.c_double),
        ("done_at", ctypes.c_double),
    ]


class ServerStats(ctypes.Structure):
    """syclfft_hip_server_stats_v1."""
    _fields_ = [(n, ctypes.c_uint64) for n in (
        "submitted", "completed", "failed", "launches", "plans_created", "plans_resident", "plans_evicted",
        "peak_footprint_bytes", "budget_bytes")]


class Server:
    """The provider's serving loop: continuous batching of transform requests over
    resident plans (include/syclfft_hip_server.h)."""

    def __init__(self, budget_bytes: int, streams: int = 4, max_batch: int = 512, device: int = -1):
        self._lib = load()
        err = ctypes.create_string_buffer(1024)
        h = self._lib.syclfft_hip_server_create(int(device), int(budget_bytes), int(streams), int(max_batch),
                                                err, 1024)
        if not h:
            raise SyclfftError("provider_unavailable", err.value.decode())
        self._h = ctypes.c_void_p(h)

    def submit_ptr(self, in_ptr: int, out_ptr: int, lengths: Sequence[int], batch: int = 1,
                   direction: int = FORWARD, double: bool = False, normalization: str = "none",
                   tag: int = 0) -> int:
        r = Request()
        r.abi_version = SERVER_ABI_VERSION
        r.scalar = SCALAR_DOUBLE if double else SCALAR_FLOAT
        r.direction = int(direction)
        r.rank = len(lengths)
        for i, n in enumerate(list(lengths)[:3]):
            r.lengths[i] = int(n)
        r.batch_count = int(batch)
        r.normalization = NORM[normalization] if isinstance(normalization, str) else int(normalization)
        r.in_place = 1 if in_ptr == out_ptr else 0
        r.inp = in_ptr
        r.out = out_ptr
        r.tag = int(tag)
        t = ctypes.c_uint64()
        err = ctypes.create_string_buffer(1024)
        st = self._lib.syclfft_hip_server_submit(self._h, ctypes.byref(r), ctypes.byref(t), err, 1024)
        if st != 0:
            _raise(st, err)
        return int(t.value)

    def submit(self, x, y=None, lengths=None, batch: int = 1, direction: int = FORWARD,
               normalization: str = "none", tag: int = 0) -> int:
        """Queue a transform of torch tensor x into y (y=None: in place)."""
        import torch

        y = x if y is None else y
        return self.submit_ptr(x.data_ptr(), y.data_ptr(), lengths, batch, direction,
                               x.dtype == torch.complex128, normalization, tag)

    def wait(self, ticket: int) -> dict:
        c = Completion()
        err = ctypes.create_string_buffer(1024)
        st = self._lib.syclfft_hip_server_wait(self._h, int(ticket), ctypes.byref(c), err, 1024)
        if st != 0:
            _raise(st, err)
        return {"ticket": c.ticket, "tag": c.tag, "group_size": c.group_size, "queue_s": c.queue_seconds,
                "run_s": c.run_seconds, "latency_s": c.latency_seconds, "submit_at": c.submit_at,
                "launch_at": c.launch_at, "done_at": c.done_at}

    def drain(self) -> None:
        err = ctypes.create_string_buffer(1024)
        st = self._lib.syclfft_hip_server_drain(self._h, err, 1024)
        if st != 0:
            _raise(st, err)

    @staticmethod
    def make_request(in_ptr: int, out_ptr: int, lengths: Sequence[int], batch: int = 1,
                     direction: int = FORWARD, double: bool = False, normalization: str = "none",
                     tag: int = 0) -> "Request":
        r = Request()
        r.abi_version = SERVER_ABI_VERSION
        r.scalar = SCALAR_DOUBLE if double else SCALAR_FLOAT
        r.direction = int(direction)
        r.rank = len(lengths)
        for i, n in enumerate(list(lengths)[:3]):
            r.lengths[i] = int(n)
        r.batch_count = int(batch)
        r.normalization = NORM[normalization] if isinstance(normalization, str) else int(normalization)
        r.in_place = 1 if in_ptr == out_ptr else 0
        r.inp = in_ptr
        r.out = out_ptr
        r.tag = int(tag)
        return r

    def replay(self, requests: Sequence["Request"], arrivals=None, sequential: bool = False) -> list:
        """Submit the requests (at `arrivals` seconds, back to back, or one at a time) from
        native threads and return their completion records in order."""
        n = len(requests)
        arr = (Request * max(1, n))(*requests)
        comp = (Completion * max(1, n))()
        times = (ctypes.c_double * n)(*[float(t) for t in arrivals]) if arrivals is not None else None
        err = ctypes.create_string_buffer(1024)
        st = self._lib.syclfft_hip_server_replay(self._h, arr, times, n, 1 if sequential else 0, comp, err, 1024)
        if st != 0:
            _raise(st, err)
        return [{"ticket": c.ticket, "tag": c.tag, "group_size": c.group_size, "queue_s": c.queue_seconds,
                 "run_s": c.run_seconds, "latency_s": c.latency_seconds, "submit_at": c.submit_at,
                 "launch_at": c.launch_at, "done_at": c.done_at} for c in comp[:n]]

    def stats(self) -> dict:
        s = ServerStats()
        self._lib.syclfft_hip_server_stats(self._h, ctypes.byref(s))
        return {n: int(getattr(s, n)) for n, _ in ServerStats._fields_}

    def close(self) -> None:
        if getattr(self, "_h", None) is not None:
            self._lib.syclfft_hip_server_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def execute_queue_ptr(plans: Sequence["Plan"], ins: Sequence[int], outs: Sequence[int], stream: int = 0,
                      done: int = 0) -> None:
    """Execute plans[k] on raw device pointers ins[k] -> outs[k], k = 0..n-1, with one
    persistent work-queue launch per kernel variant (syclfft_hip_execute_queue)."""
    if not (len(plans) == len(ins) == len(outs)):
        raise ValueError("plans, ins and outs must have the same length")
    n = len(plans)
    lib = load()
    pa = (ctypes.c_void_p * max(1, n))(*[p.handle.value for p in plans])
    ia = (ctypes.c_void_p * max(1, n))(*[int(v) for v in ins])
    oa = (ctypes.c_void_p * max(1, n))(*[int(v) for v in outs])
    err = ctypes.create_string_buffer(1024)
    st = lib.syclfft_hip_execute_queue(pa, ia, oa, n, ctypes.c_void_p(stream), ctypes.c_void_p(done), err, 1024)
    if st != 0:
        _raise(st, err)


def execute_queue(plans: Sequence["Plan"], xs, ys=None, stream=None) -> None:
    """execute_queue_ptr over torch tensors (ys defaults to xs: in-place plans)."""
    import torch

    ys = xs if ys is None else ys
    s = stream if stream is not None else torch.cuda.current_stream(xs[0].device)
    execute_queue_ptr(plans, [x.data_ptr() for x in xs], [y.data_ptr() for y in ys], s.cuda_stream)


def provider_table(lib: Optional[ctypes.CDLL] = None) -> ProviderTable:
    lib = lib or load()
    return lib.syclfft_get_device_provider_v1().contents


def supports(device: int = -1) -> tuple[bool, str]:
    err = ctypes.create_string_buffer(1024)
    ok = load().syclfft_hip_supports(device, err, 1024)
    return bool(ok), err.value.decode()


def make_config(lengths: Sequence[int], batch: int = 1, direction: int = FORWARD,
                double: bool = False, in_place: bool = False, normalization: str = "none",
                precision: str = "valu", measure: bool = False, device: int = -1) -> PlanConfig:
    cfg = PlanConfig()
    cfg.abi_version = ABI_VERSION
    cfg.scalar = SCALAR_DOUBLE if double else SCALAR_FLOAT
    cfg.direction = int(direction)
    cfg.rank = len(lengths)
    for i, n in enumerate(list(lengths)[:3]):
        cfg.lengths[i] = int(n)
    cfg.batch_count = int(batch)
    cfg.in_place = 1 if in_place else 0
    cfg.normalization = NORM[normalization] if isinstance(normalization, str) else int(normalization)
    cfg.precision = PRECISION[precision] if isinstance(precision, str) else int(precision)
    cfg.measure = 1 if measure else 0
    cfg.device = int(device)
    return cfg


def _raise(status: int, err: ctypes.Array) -> None:
    raise SyclfftError(ERROR_CODES.get(status, f"status{status}"), err.value.decode())


class Plan:
    """A provider plan (RAII over syclfft_hip_plan_create / _destroy)."""

    def __init__(self, lengths: Sequence[int], batch: int = 1, direction: int = FORWARD,
                 double: bool = False, in_place: bool = False, normalization: str = "none",
                 precision: str = "valu", measure: bool = False, device: int = -1, real: bool = False,
                 layout: "Layout | None" = None):
        """real=True: a real 1-D plan (syclfft_hip_real_plan_create): FORWARD = R2C (N real ->
        N/2 + 1 complex per transform), BACKWARD = C2R; execute on float/double tensors for
        the real side and complex tensors for the spectrum side.

        layout: a layout plan (syclfft_hip_layout_plan_create) over strided and / or planar
        buffers; execute it with execute_layout."""
        self._lib = load()
        self.real = bool(real)
        self.lengths = tuple(int(n) for n in lengths)
        self.batch = int(batch)
        self.direction = int(direction)
        self.double = bool(double)
        self.in_place = bool(in_place)
        self.normalization = normalization
        self.precision = precision
        self.layout = layout
        cfg = make_config(lengths, batch, direction, double, in_place, normalization, precision,
                          measure, device)
        err = ctypes.create_string_buffer(1024)
        h = ctypes.c_void_p()
        self._h = None
        if layout is not None:
            st = self._lib.syclfft_hip_layout_plan_create(ctypes.byref(cfg), ctypes.byref(layout), ctypes.byref(h),
                                                          err, 1024)
        else:
            create = self._lib.syclfft_hip_real_plan_create if real else self._lib.syclfft_hip_plan_create_ex
            st = create(ctypes.byref(cfg), ctypes.byref(h), err, 1024)
        if st != 0:
            _raise(st, err)
        self._h = h

    @property
    def handle(self) -> ctypes.c_void_p:
        if self._h is None:
            raise SyclfftError("invalid_state", "Cannot execute a moved-from plan")
        return self._h

    def close(self) -> None:
        if getattr(self, "_h", None) is not None:
            self._lib.syclfft_hip_plan_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    @property
    def element_count(self) -> int:
        n = self.batch
        for x in self.lengths:
            n *= x
        return n

    def scratch_size_bytes(self) -> int:
        return int(self._lib.syclfft_hip_plan_scratch_bytes(self.handle))

    def describe(self) -> str:
        n = self._lib.syclfft_hip_plan_describe(self.handle, None, 0)
        buf = ctypes.create_string_buffer(n + 1)
        self._lib.syclfft_hip_plan_describe(self.handle, buf, n + 1)
        return buf.value.decode()

    def launch_count(self) -> int:
        return int(self._lib.syclfft_hip_plan_launch_count(self.handle))

    def execute_ptr(self, in_ptr: int, out_ptr: int, stream: int = 0,
                    deps: Iterable[int] = (), done: int = 0) -> None:
        """Enqueue on a raw hipStream_t with raw hipEvent_t dependencies."""
        dl = [int(d) for d in deps]
        arr = (ctypes.c_void_p * max(1, len(dl)))(*dl) if dl else None
        err = ctypes.create_string_buffer(1024)
        st = self._lib.syclfft_hip_plan_execute(
            self.handle, ctypes.c_void_p(in_ptr), ctypes.c_void_p(out_ptr),
            ctypes.c_void_p(stream), arr, len(dl), ctypes.c_void_p(done), err, 1024)
        if st != 0:
            _raise(st, err)

    def execute_many_ptr(self, ins: Sequence[int], outs: Sequence[int], stream: int = 0,
                         deps: Iterable[int] = (), done: int = 0) -> None:
        """Enqueue len(ins) transforms of the plan's shape over raw device pointers
        (transform k reads ins[k], writes outs[k]); at most the plan's batch count."""
        if len(ins) != len(outs):
            raise ValueError("ins and outs must have the same length")
        n = len(ins)
        ia = (ctypes.c_void_p * max(1, n))(*[int(v) for v in ins])
        oa = (ctypes.c_void_p * max(1, n))(*[int(v) for v in outs])
        dl = [int(d) for d in deps]
        arr = (ctypes.c_void_p * max(1, len(dl)))(*dl) if dl else None
        err = ctypes.create_string_buffer(1024)
        st = self._lib.syclfft_hip_plan_execute_many(
            self.handle, n, ia, oa, ctypes.c_void_p(stream), arr, len(dl), ctypes.c_void_p(done), err, 1024)
        if st != 0:
            _raise(st, err)

    def execute_layout_ptr(self, x, x_im, y, y_im, stream: int = 0, deps=(), done: int = 0) -> None:
        """Execute a layout plan over raw device pointers (x_im / y_im: imaginary-part arrays
        of planar sides, 0 for interleaved sides)."""
        dl = [int(d) for d in deps]
        arr = (ctypes.c_void_p * max(1, len(dl)))(*dl) if dl else None
        err = ctypes.create_string_buffer(1024)
        st = self._lib.syclfft_hip_layout_execute(
            self.handle, ctypes.c_void_p(int(x)), ctypes.c_void_p(int(x_im)), ctypes.c_void_p(int(y)),
            ctypes.c_void_p(int(y_im)), ctypes.c_void_p(stream), arr, len(dl), ctypes.c_void_p(done), err, 1024)
        if st != 0:
            _raise(st, err)

    def execute_layout(self, x, y=None, stream=None, deps=(), done=None) -> None:
        """Execute a layout plan on torch tensors: an interleaved side is one complex tensor, a
        planar side a (real, imaginary) pair of float / double tensors."""
        import torch

        y = x if y is None else y

        def parts(v):
            if isinstance(v, (tuple, list)):
                return v[0].data_ptr(), v[1].data_ptr(), v[0].device
            return v.data_ptr(), 0, v.device

        xr, xi, dev = parts(x)
        yr, yi, _ = parts(y)
        s = stream if stream is not None else torch.cuda.current_stream(dev)
        self.execute_layout_ptr(xr, xi, yr, yi, s.cuda_stream, [e.cuda_event for e in deps], 0)
        if done is not None:
            done.record(s)

    def execute_many(self, xs, ys=None, stream=None, deps=(), done=None) -> None:
        """Execute on lists of torch tensors, one transform of the plan's shape each."""
        import torch

        ys = xs if ys is None else ys
        dev = xs[0].device
        s = stream if stream is not None else torch.cuda.current_stream(dev)
        self.execute_many_ptr([x.data_ptr() for x in xs], [y.data_ptr() for y in ys], s.cuda_stream,
                              [e.cuda_event for e in deps], 0)
        if done is not None:
            done.record(s)

    def execute(self, x, y=None, stream=None, deps=(), done=None) -> None:
        """Execute on torch tensors (complex64 / complex128 on a HIP device).

        execute(inout) for in-place plans, execute(input, output) for out-of-place plans;
        `stream` is a torch.cuda.Stream (default: the current stream), `deps` torch.cuda.Events
        to wait for, `done` a torch.cuda.Event recorded at completion."""
        import torch

        y = x if y is None else y
        s = stream if stream is not None else torch.cuda.current_stream(x.device)
        dep_handles = [e.cuda_event for e in deps]
        self.execute_ptr(x.data_ptr(), y.data_ptr(), s.cuda_stream, dep_handles, 0)
        if done is not None:
            done.record(s)
