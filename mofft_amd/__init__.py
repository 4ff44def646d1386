"""MI355X-native MoFFT: Python bindings over the C ABI of libmofft.so.

The product is the C library (``include/mofft.h`` drop-in plus
``include/mofft_hip.h`` extensions) built by ``mofft_amd.build`` into
``mofft_amd/lib/libmofft.so``.  This module binds it with ctypes so tests,
the benchmark and Python users call exactly the entry points a C caller
would (the reference's API, include/mofft.h:107-231).  There is no Python or
CPU compute path: if the library is missing, importing the bindings raises.
"""
from __future__ import annotations

import ctypes
import os
from typing import Optional, Sequence

_HERE = os.path.dirname(os.path.abspath(__file__))
# MOFFT_LIBRARY selects another build of the same library (e.g. one made
# with `python -m mofft_amd.build --out-dir DIR --define ...`).
LIB_PATH = os.environ.get("MOFFT_LIBRARY") or os.path.join(_HERE, "lib", "libmofft.so")

# include/mofft.h:14-35
SUCCESS = 0
INVALID_ARGUMENT = 1
UNSUPPORTED_SIZE = 2
ALLOCATION_FAILURE = 3
UNSUPPORTED_HARDWARE = 4
INTERNAL_ERROR = 5
WISDOM_NOT_FOUND = 6
INCOMPATIBLE_RECIPE = 7
BUFFER_TOO_SMALL = 8
FORWARD = -1
BACKWARD = 1
OUT_OF_PLACE = 0
IN_PLACE = 1
PRETUNED = 0
MEASURE = 1
STRICT_WISDOM = 1
FLAG_MEASURE = 1
FLAG_WISDOM_ONLY = 2
MAX_PLAN_DEPTH = 64
PLAN_MFMA = 0x100          # mofft_hip.h: radix-16 stages on the matrix cores
MFMA_BANK = 3
MAX_PASSES = 16


class MofftError(RuntimeError):
    def __init__(self, status: int, what: str = ""):
        self.status = status
        super().__init__(f"{what}: {status_string(status)} ({status})")


class PlanOptions(ctypes.Structure):
    _fields_ = [("mode", ctypes.c_int), ("flags", ctypes.c_uint32),
                ("candidate_limit", ctypes.c_size_t),
                ("timing_samples", ctypes.c_uint32)]


class PlanDiagnostics(ctypes.Structure):
    _fields_ = [("measured", ctypes.c_int), ("used_wisdom", ctypes.c_int),
                ("used_model_fallback", ctypes.c_int),
                ("candidates_measured", ctypes.c_size_t),
                ("radix_count", ctypes.c_size_t),
                ("radices", ctypes.c_int * MAX_PLAN_DEPTH),
                ("kernel_bank", ctypes.c_uint),
                ("kernel_bank_count", ctypes.c_uint),
                ("blocked_transpose_mask", ctypes.c_uint64),
                ("direct_input", ctypes.c_int),
                ("stage_layout", ctypes.c_int),
                ("planning_seconds", ctypes.c_double),
                ("selected_median_seconds", ctypes.c_double),
                ("length", ctypes.c_size_t),
                ("precision_bits", ctypes.c_int),
                ("direction", ctypes.c_int),
                ("placement", ctypes.c_int)]

    @property
    def chain(self) -> list[int]:
        return list(self.radices[: self.radix_count])


class CostBreakdown(ctypes.Structure):
    _fields_ = [("kernel_cost", ctypes.c_double), ("memory_cost", ctypes.c_double),
                ("layout_cost", ctypes.c_double), ("total_cost", ctypes.c_double),
                ("logical_read_bytes", ctypes.c_size_t),
                ("logical_write_bytes", ctypes.c_size_t),
                ("transferred_read_bytes", ctypes.c_size_t),
                ("transferred_write_bytes", ctypes.c_size_t),
                ("peak_working_set_bytes", ctypes.c_size_t)]


class PlanDescription(ctypes.Structure):
    _fields_ = [("length", ctypes.c_size_t), ("batch", ctypes.c_size_t),
                ("precision_bits", ctypes.c_int),
                ("radix_count", ctypes.c_size_t),
                ("radices", ctypes.c_int * MAX_PLAN_DEPTH),
                ("pass_mask", ctypes.c_uint64),
                ("pass_count", ctypes.c_int),
                ("pass_length", ctypes.c_size_t * MAX_PASSES),
                ("pass_stride", ctypes.c_size_t * MAX_PASSES),
                ("pass_span", ctypes.c_size_t * MAX_PASSES),
                ("pass_kernel", ctypes.c_int * MAX_PASSES),
                ("chunk_transforms", ctypes.c_size_t),
                ("workspace_bytes", ctypes.c_size_t),
                ("twiddle_bytes", ctypes.c_size_t)]

    @property
    def chain(self) -> list[int]:
        return list(self.radices[: self.radix_count])

    @property
    def passes(self) -> list[int]:
        return list(self.pass_length[: self.pass_count])


class KernelInfo(ctypes.Structure):
    _fields_ = [("precision_bits", ctypes.c_int), ("length", ctypes.c_int),
                ("rows", ctypes.c_int), ("threads_per_column", ctypes.c_int),
                ("columns_per_block", ctypes.c_int), ("stage_count", ctypes.c_int),
                ("stages", ctypes.c_int * 8), ("variant", ctypes.c_int)]


class BatchOptions(ctypes.Structure):
    _fields_ = [("flags", ctypes.c_uint32), ("grid", ctypes.c_uint32),
                ("spin_limit", ctypes.c_uint64), ("owned_tiles", ctypes.c_uint32),
                ("item_tiles", ctypes.c_uint32)]


class BatchDescription(ctypes.Structure):
    _fields_ = [("job_count", ctypes.c_size_t), ("distinct_lengths", ctypes.c_size_t),
                ("multi_pass_jobs", ctypes.c_size_t), ("shared_jobs", ctypes.c_size_t),
                ("owned_tiles", ctypes.c_size_t), ("item_tiles", ctypes.c_size_t),
                ("wisdom_lengths", ctypes.c_size_t),
                ("segment_count", ctypes.c_size_t), ("item_count", ctypes.c_size_t),
                ("grid", ctypes.c_size_t), ("blocks_per_cu", ctypes.c_size_t),
                ("lds_bytes", ctypes.c_size_t), ("workspace_bytes", ctypes.c_size_t),
                ("twiddle_bytes", ctypes.c_size_t), ("flops", ctypes.c_double),
                ("algorithmic_bytes", ctypes.c_double), ("precision_bits", ctypes.c_int),
                ("graph", ctypes.c_int)]


class ServerOptions(ctypes.Structure):
    _fields_ = [("budget_bytes", ctypes.c_size_t), ("streams", ctypes.c_uint32),
                ("max_inflight", ctypes.c_uint32), ("max_batch", ctypes.c_uint32),
                ("flags", ctypes.c_uint32)]


class Request(ctypes.Structure):
    _fields_ = [("length", ctypes.c_size_t), ("batch", ctypes.c_size_t),
                ("precision_bits", ctypes.c_int), ("direction", ctypes.c_int),
                ("input", ctypes.c_void_p), ("seed", ctypes.c_uint64),
                ("tag", ctypes.c_uint64)]


class Completion(ctypes.Structure):
    _fields_ = [("ticket", ctypes.c_uint64), ("tag", ctypes.c_uint64),
                ("status", ctypes.c_int), ("submit_seconds", ctypes.c_double),
                ("start_seconds", ctypes.c_double), ("done_seconds", ctypes.c_double),
                ("output", ctypes.c_void_p), ("bytes", ctypes.c_size_t)]


class ServerStats(ctypes.Structure):
    _fields_ = [("submitted", ctypes.c_size_t), ("completed", ctypes.c_size_t),
                ("pending", ctypes.c_size_t), ("inflight", ctypes.c_size_t),
                ("resident_plans", ctypes.c_size_t), ("bytes_in_use", ctypes.c_size_t),
                ("peak_bytes_in_use", ctypes.c_size_t), ("budget_bytes", ctypes.c_size_t),
                ("peak_inflight", ctypes.c_size_t), ("flops_completed", ctypes.c_double),
                ("bytes_completed", ctypes.c_double), ("executions", ctypes.c_size_t)]


SERVER_NO_COALESCE = 1


BATCH_NO_GRAPH = 1
BATCH_SHARE_ALL = 2

_lib = None

# (name, restype, argtypes); P = void pointer
_P = ctypes.c_void_p
_SZ = ctypes.c_size_t
_I = ctypes.c_int
_U64 = ctypes.c_uint64
_SIGS = []
for _s in ("f32", "f64"):
    _SIGS += [
        (f"mofft_plan_create_{_s}", _I, [ctypes.POINTER(_P), _SZ, _I, _I]),
        (f"mofft_plan_create_with_options_{_s}", _I,
         [ctypes.POINTER(_P), _SZ, _I, _I, ctypes.POINTER(PlanOptions),
          ctypes.POINTER(PlanDiagnostics)]),
        (f"mofft_plan_dft_1d_{_s}", _P, [_SZ, _P, _P, _I, ctypes.c_uint32]),
        (f"mofft_execute_plan_{_s}", _I, [_P]),
        (f"mofft_plan_create_with_radices_{_s}", _I,
         [ctypes.POINTER(_P), _SZ, _I, _I, ctypes.POINTER(_I), _SZ]),
        (f"mofft_plan_set_transpose_strategy_{_s}", _I, [_P, _I]),
        (f"mofft_plan_set_blocked_transpose_mask_{_s}", _I, [_P, _U64]),
        (f"mofft_plan_set_direct_input_{_s}", _I, [_P, _I]),
        (f"mofft_plan_set_stage_layout_{_s}", _I, [_P, _I]),
        (f"mofft_plan_set_kernel_bank_{_s}", _I, [_P, ctypes.c_uint]),
        (f"mofft_plan_get_kernel_bank_{_s}", _I, [_P, ctypes.POINTER(ctypes.c_uint)]),
        (f"mofft_plan_export_{_s}", _I, [_P, ctypes.c_char_p, _SZ, ctypes.POINTER(_SZ)]),
        (f"mofft_plan_import_{_s}", _I,
         [ctypes.POINTER(_P), ctypes.c_char_p, _SZ, ctypes.POINTER(PlanDiagnostics)]),
        (f"mofft_execute_{_s}", _I, [_P, _P, _P]),
        (f"mofft_plan_destroy_{_s}", None, [_P]),
        (f"mofft_plan_create_batched_{_s}", _I,
         [ctypes.POINTER(_P), _SZ, _SZ, _I, _I, ctypes.POINTER(PlanOptions),
          ctypes.POINTER(PlanDiagnostics)]),
        (f"mofft_plan_create_batched_with_radices_{_s}", _I,
         [ctypes.POINTER(_P), _SZ, _SZ, _I, _I, ctypes.POINTER(_I), _SZ]),
        (f"mofft_plan_get_batch_{_s}", _I, [_P, ctypes.POINTER(_SZ)]),
        (f"mofft_execute_async_{_s}", _I, [_P, _P, _P, _P]),
        (f"mofft_plan_workspace_size_{_s}", _I, [_P, ctypes.POINTER(_SZ)]),
        (f"mofft_execute_async_workspace_{_s}", _I, [_P, _P, _P, _P, _SZ, _P]),
        (f"mofft_plan_get_description_{_s}", _I, [_P, ctypes.POINTER(PlanDescription)]),
        (f"mofft_plan_graph_create_{_s}", _I, [_P, _P, _P, ctypes.POINTER(_P)]),
        (f"mofft_execute_timed_{_s}", _I,
         [_P, _P, _P, _P, ctypes.POINTER(ctypes.c_double), _SZ]),
    ]
_SIGS += [
    ("mofft_plan_options_init", None, [ctypes.POINTER(PlanOptions)]),
    ("mofft_build_id", ctypes.c_char_p, []),
    ("mofft_estimated_radix_cost", ctypes.c_double, [_I, _I, _I]),
    ("mofft_estimated_stage_cost", ctypes.c_double, [_I, _I, _I, _SZ, _SZ]),
    ("mofft_estimated_plan_cost", ctypes.c_double,
     [_SZ, _I, ctypes.POINTER(_I), _SZ, _U64, _I]),
    ("mofft_estimated_plan_cost_with_layout", ctypes.c_double,
     [_SZ, _I, ctypes.POINTER(_I), _SZ, _U64, _I, _I]),
    ("mofft_estimated_plan_cost_breakdown_with_layout", _I,
     [_SZ, _I, ctypes.POINTER(_I), _SZ, _U64, _I, _I, ctypes.POINTER(CostBreakdown)]),
    ("mofft_wisdom_is_formal", _I, []),
    ("mofft_wisdom_source", ctypes.c_char_p, []),
    ("mofft_status_string", ctypes.c_char_p, [_I]),
    ("mofft_version", ctypes.c_char_p, []),
    ("mofft_plan_describe", _I,
     [_SZ, _SZ, _I, ctypes.POINTER(_I), _SZ, _U64, ctypes.POINTER(PlanDescription)]),
    ("mofft_estimated_schedule_seconds", ctypes.c_double,
     [_SZ, _SZ, _I, ctypes.POINTER(_I), _SZ, _U64, ctypes.c_uint]),
    ("mofft_estimated_pass_costs_measured", _I, []),
    ("mofft_kernel_count", _I, []),
    ("mofft_kernel_get_info", _I, [_I, ctypes.POINTER(KernelInfo)]),
    ("mofft_device_available", _I, []),
    ("mofft_graph_launch", _I, [_P, _P]),
    ("mofft_graph_destroy", None, [_P]),
    ("mofft_batch_options_init", None, [ctypes.POINTER(BatchOptions)]),
    ("mofft_batch_plan_create", _I,
     [ctypes.POINTER(_P), _I, _SZ, ctypes.POINTER(_SZ), ctypes.POINTER(_I),
      ctypes.POINTER(_P), ctypes.POINTER(_P), ctypes.POINTER(BatchOptions)]),
    ("mofft_batch_set_buffers", _I, [_P, ctypes.POINTER(_P), ctypes.POINTER(_P)]),
    ("mofft_batch_execute_async", _I, [_P, _P]),
    ("mofft_batch_check", _I, [_P]),
    ("mofft_batch_execute", _I, [_P]),
    ("mofft_batch_describe", _I, [_P, ctypes.POINTER(BatchDescription)]),
    ("mofft_batch_describe_length", _I, [_I, _SZ, ctypes.POINTER(PlanDescription)]),
    ("mofft_batch_plan_destroy", None, [_P]),
    ("mofft_server_options_init", None, [ctypes.POINTER(ServerOptions)]),
    ("mofft_server_create", _I, [ctypes.POINTER(_P), ctypes.POINTER(ServerOptions)]),
    ("mofft_server_submit", _I, [_P, ctypes.POINTER(Request), ctypes.POINTER(ctypes.c_uint64)]),
    ("mofft_server_poll", _I, [_P, ctypes.POINTER(Completion), _SZ, ctypes.POINTER(_SZ), _I]),
    ("mofft_server_release", _I, [_P, ctypes.c_uint64]),
    ("mofft_server_get_stats", _I, [_P, ctypes.POINTER(ServerStats)]),
    ("mofft_server_run_trace", _I,
     [_P, ctypes.POINTER(Request), ctypes.POINTER(ctypes.c_double), _SZ,
      ctypes.POINTER(Completion)]),
    ("mofft_server_generate", _I, [_P, _I, _SZ, ctypes.c_uint64, _P]),
    ("mofft_server_destroy", None, [_P]),
]


def lib() -> ctypes.CDLL:
    """Load libmofft.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is missing: run `python -m mofft_amd.build`")
    handle = ctypes.CDLL(LIB_PATH)
    for name, res, args in _SIGS:
        fn = getattr(handle, name)
        fn.restype = res
        fn.argtypes = args
    _lib = handle
    return handle


def exported_symbols() -> list[str]:
    return [name for name, _, _ in _SIGS]


def status_string(status: int) -> str:
    return lib().mofft_status_string(int(status)).decode()


def build_id() -> str:
    return lib().mofft_build_id().decode()


def device_available() -> bool:
    return bool(lib().mofft_device_available())


def _check(status: int, what: str) -> None:
    if status != SUCCESS:
        raise MofftError(status, what)


def _int_array(values: Sequence[int]):
    arr = (ctypes.c_int * max(1, len(values)))(*values)
    return arr, len(values)


def describe(length: int, batch: int = 1, precision_bits: int = 32,
             radices: Optional[Sequence[int]] = None, pass_mask: int = 0) -> PlanDescription:
    """Host-only schedule of a plan (no device needed)."""
    d = PlanDescription()
    if radices is None:
        st = lib().mofft_plan_describe(length, batch, precision_bits, None, 0, 0,
                                       ctypes.byref(d))
    else:
        arr, n = _int_array(radices)
        st = lib().mofft_plan_describe(length, batch, precision_bits, arr, n,
                                       pass_mask, ctypes.byref(d))
    _check(st, "mofft_plan_describe")
    return d


def kernels() -> list[KernelInfo]:
    out = []
    for i in range(lib().mofft_kernel_count()):
        k = KernelInfo()
        _check(lib().mofft_kernel_get_info(i, ctypes.byref(k)), "kernel_get_info")
        out.append(k)
    return out


def _ptr(x) -> int:
    """Data pointer of a torch tensor (device or host) or a numpy array."""
    if hasattr(x, "data_ptr"):
        return int(x.data_ptr())
    return int(x.ctypes.data)


class Plan:
    """A batched 1-D C2C plan (mofft_plan_create_batched_*)."""

    def __init__(self, length: int, batch: int = 1, precision_bits: int = 32,
                 direction: int = FORWARD, in_place: bool = False,
                 radices: Optional[Sequence[int]] = None,
                 mode: int = PRETUNED, flags: int = 0, candidate_limit: int = 0,
                 timing_samples: int = 0, pass_mask: Optional[int] = None,
                 mfma: bool = False):
        if precision_bits not in (32, 64):
            raise ValueError("precision_bits must be 32 or 64")
        self.length = int(length)
        self.batch = int(batch)
        self.precision_bits = precision_bits
        self.direction = direction
        self.in_place = bool(in_place)
        self._s = "f32" if precision_bits == 32 else "f64"
        self._handle = ctypes.c_void_p()
        self.diagnostics = PlanDiagnostics()
        placement = IN_PLACE if in_place else OUT_OF_PLACE
        L = lib()
        if radices is None:
            opts = PlanOptions()
            L.mofft_plan_options_init(ctypes.byref(opts))
            opts.mode = mode
            opts.flags = flags | (PLAN_MFMA if mfma else 0)
            opts.candidate_limit = candidate_limit
            opts.timing_samples = timing_samples
            st = getattr(L, f"mofft_plan_create_batched_{self._s}")(
                ctypes.byref(self._handle), self.length, self.batch, direction,
                placement, ctypes.byref(opts), ctypes.byref(self.diagnostics))
        else:
            arr, n = _int_array(radices)
            st = getattr(L, f"mofft_plan_create_batched_with_radices_{self._s}")(
                ctypes.byref(self._handle), self.length, self.batch, direction,
                placement, arr, n)
        _check(st, f"plan N={length} batch={batch} f{precision_bits}")
        if pass_mask is not None:
            self.set_blocked_transpose_mask(pass_mask)

    @property
    def handle(self) -> int:
        return self._handle.value

    def _fn(self, name):
        return getattr(lib(), f"{name}_{self._s}")

    def description(self) -> PlanDescription:
        d = PlanDescription()
        _check(self._fn("mofft_plan_get_description")(self._handle, ctypes.byref(d)),
               "get_description")
        return d

    def workspace_size(self) -> int:
        n = ctypes.c_size_t()
        _check(self._fn("mofft_plan_workspace_size")(self._handle, ctypes.byref(n)),
               "workspace_size")
        return n.value

    def set_blocked_transpose_mask(self, mask: int) -> None:
        _check(self._fn("mofft_plan_set_blocked_transpose_mask")(self._handle, mask),
               "set_blocked_transpose_mask")

    def execute(self, x_in, x_out) -> None:
        """Synchronous mofft_execute_* (host or device buffers)."""
        _check(self._fn("mofft_execute")(self._handle, _ptr(x_in), _ptr(x_out)),
               "execute")

    def execute_async(self, x_in, x_out, stream: int = 0) -> None:
        """Stream-ordered execution on device buffers."""
        _check(self._fn("mofft_execute_async")(self._handle, _ptr(x_in), _ptr(x_out),
                                               stream or None), "execute_async")

    def execute_async_workspace(self, x_in, x_out, workspace, nbytes: int,
                                stream: int = 0) -> None:
        _check(self._fn("mofft_execute_async_workspace")(
            self._handle, _ptr(x_in), _ptr(x_out),
            _ptr(workspace) if workspace is not None else None, nbytes,
            stream or None), "execute_async_workspace")

    def capture(self, x_in, x_out) -> "Graph":
        """The plan's launches on these device buffers as a HIP graph."""
        g = ctypes.c_void_p()
        _check(self._fn("mofft_plan_graph_create")(self._handle, _ptr(x_in), _ptr(x_out),
                                                   ctypes.byref(g)), "plan_graph_create")
        return Graph(g)

    def execute_timed(self, x_in, x_out, stream: int = 0) -> list[float]:
        """One execution with device timestamps between pass launches;
        returns seconds per pass (summed over batch chunks)."""
        secs = (ctypes.c_double * MAX_PASSES)()
        _check(self._fn("mofft_execute_timed")(self._handle, _ptr(x_in), _ptr(x_out),
                                               stream or None, secs, MAX_PASSES),
               "execute_timed")
        return list(secs[: self.description().pass_count])

    def export(self) -> str:
        need = ctypes.c_size_t()
        _check(self._fn("mofft_plan_export")(self._handle, None, 0, ctypes.byref(need)),
               "export size")
        buf = ctypes.create_string_buffer(need.value)
        _check(self._fn("mofft_plan_export")(self._handle, buf, need.value,
                                             ctypes.byref(need)), "export")
        return buf.value.decode()

    def destroy(self) -> None:
        if self._handle and self._handle.value:
            self._fn("mofft_plan_destroy")(self._handle)
            self._handle = ctypes.c_void_p()

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.destroy()


class Graph:
    """A captured plan execution (mofft_plan_graph_create_*)."""

    def __init__(self, handle):
        self._h = handle

    def launch(self, stream: int = 0) -> None:
        _check(lib().mofft_graph_launch(self._h, stream or None), "graph_launch")

    def destroy(self) -> None:
        if self._h and self._h.value:
            lib().mofft_graph_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.destroy()


def batch_describe_length(length: int, precision_bits: int = 32) -> PlanDescription:
    """Host-only schedule a batch plan uses for one length."""
    d = PlanDescription()
    _check(lib().mofft_batch_describe_length(precision_bits, length, ctypes.byref(d)),
           f"batch_describe_length N={length}")
    return d


class BatchPlan:
    """Many transforms of uneven length in one persistent launch
    (mofft_batch_plan_create).  `inputs`/`outputs` are device buffers
    (torch tensors or raw pointers); a transform is in place when its input
    and output are the same buffer."""

    def __init__(self, lengths: Sequence[int], directions: Sequence[int],
                 inputs, outputs, precision_bits: int = 32, graph: bool = True,
                 grid: int = 0, spin_limit: int = 0, share_all: bool = False,
                 owned_tiles: int = 0, item_tiles: int = 0):
        n = len(lengths)
        if not (len(directions) == len(inputs) == len(outputs) == n):
            raise ValueError("lengths, directions, inputs and outputs differ in size")
        self.precision_bits = precision_bits
        self.count = n
        self._lengths = (ctypes.c_size_t * n)(*[int(x) for x in lengths])
        self._dirs = (ctypes.c_int * n)(*[int(d) for d in directions])
        ins, outs = self._pointers(inputs, outputs)
        opts = BatchOptions()
        lib().mofft_batch_options_init(ctypes.byref(opts))
        opts.flags = (0 if graph else BATCH_NO_GRAPH) | (BATCH_SHARE_ALL if share_all else 0)
        opts.grid = grid
        opts.spin_limit = spin_limit
        opts.owned_tiles = owned_tiles
        opts.item_tiles = item_tiles
        self._handle = ctypes.c_void_p()
        _check(lib().mofft_batch_plan_create(ctypes.byref(self._handle), precision_bits, n,
                                             self._lengths, self._dirs, ins, outs,
                                             ctypes.byref(opts)),
               f"batch plan of {n} transforms f{precision_bits}")

    def _pointers(self, inputs, outputs):
        n = self.count
        ins = (ctypes.c_void_p * n)(*[x if isinstance(x, int) else _ptr(x) for x in inputs])
        outs = (ctypes.c_void_p * n)(*[x if isinstance(x, int) else _ptr(x) for x in outputs])
        return ins, outs

    def set_buffers(self, inputs, outputs) -> None:
        ins, outs = self._pointers(inputs, outputs)
        _check(lib().mofft_batch_set_buffers(self._handle, ins, outs), "batch_set_buffers")

    def execute_async(self, stream: int = 0) -> None:
        _check(lib().mofft_batch_execute_async(self._handle, stream or None),
               "batch_execute_async")

    def execute(self) -> None:
        _check(lib().mofft_batch_execute(self._handle), "batch_execute")

    def check(self) -> None:
        _check(lib().mofft_batch_check(self._handle), "batch_check")

    def description(self) -> BatchDescription:
        d = BatchDescription()
        _check(lib().mofft_batch_describe(self._handle, ctypes.byref(d)), "batch_describe")
        return d

    def destroy(self) -> None:
        if self._handle and self._handle.value:
            lib().mofft_batch_plan_destroy(self._handle)
            self._handle = ctypes.c_void_p()

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.destroy()


class Server:
    """Continuous-batching request loop (mofft_server_*): admission against
    a device-memory budget, resident plans, results in server buffers."""

    def __init__(self, budget_bytes: int, streams: int = 0, max_inflight: int = 0,
                 max_batch: int = 0, coalesce: bool = True):
        o = ServerOptions()
        lib().mofft_server_options_init(ctypes.byref(o))
        o.budget_bytes = int(budget_bytes)
        o.streams = streams
        o.max_inflight = max_inflight
        o.max_batch = max_batch
        o.flags = 0 if coalesce else SERVER_NO_COALESCE
        self._h = ctypes.c_void_p()
        _check(lib().mofft_server_create(ctypes.byref(self._h), ctypes.byref(o)),
               "server_create")

    def submit(self, length: int, batch: int = 1, precision_bits: int = 32,
               direction: int = FORWARD, seed: int = 0, input=None, tag: int = 0) -> int:
        r = Request(length, batch, precision_bits, direction,
                    None if input is None else (input if isinstance(input, int) else _ptr(input)),
                    seed, tag)
        t = ctypes.c_uint64()
        _check(lib().mofft_server_submit(self._h, ctypes.byref(r), ctypes.byref(t)),
               f"server_submit N={length} batch={batch}")
        return t.value

    def poll(self, wait: bool = False, capacity: int = 64) -> list:
        buf = (Completion * capacity)()
        n = ctypes.c_size_t()
        _check(lib().mofft_server_poll(self._h, buf, capacity, ctypes.byref(n), int(wait)),
               "server_poll")
        return [buf[i] for i in range(n.value)]

    def release(self, ticket: int) -> None:
        _check(lib().mofft_server_release(self._h, ticket), "server_release")

    def stats(self) -> ServerStats:
        st = ServerStats()
        _check(lib().mofft_server_get_stats(self._h, ctypes.byref(st)), "server_stats")
        return st

    def run_trace(self, requests: Sequence[Request], arrivals: Sequence[float]):
        n = len(requests)
        reqs = (Request * n)(*requests)
        arr = (ctypes.c_double * n)(*arrivals)
        out = (Completion * n)()
        _check(lib().mofft_server_run_trace(self._h, reqs, arr, n, out), "server_run_trace")
        return out

    def destroy(self) -> None:
        if self._h and self._h.value:
            lib().mofft_server_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.destroy()


def server_generate(out, precision_bits: int, count: int, seed: int, stream: int = 0) -> None:
    """The server's seeded device input generator on a caller buffer."""
    _check(lib().mofft_server_generate(_ptr(out), precision_bits, count, seed, stream or None),
           "server_generate")


def estimated_schedule_seconds(length: int, batch: int, precision_bits: int, radices,
                               mask: int = 0, bank: int = 0) -> float:
    """The planner's predicted device seconds of a schedule
    (mofft_estimated_schedule_seconds); inf for an invalid plan."""
    arr, n = _int_array(radices)
    v = lib().mofft_estimated_schedule_seconds(length, batch, precision_bits, arr, n,
                                               mask, bank)
    return float("inf") if v > 1e300 else v


def estimated_plan_breakdown(length: int, precision_bits: int, radices,
                             mask: int = 0, direct_input: int = 0,
                             section: bool = False) -> CostBreakdown:
    arr, n = _int_array(radices)
    b = CostBreakdown()
    _check(lib().mofft_estimated_plan_cost_breakdown_with_layout(
        length, precision_bits, arr, n, mask, direct_input, int(section),
        ctypes.byref(b)), "estimated_plan_cost_breakdown")
    return b
