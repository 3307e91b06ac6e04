"""clbellard_amd -- MI355X-native Bellard-formula hex digit extraction of pi.

Python mirror of the C ABI in include/clbellard.h (libclbellard.so, built in this directory by
``python -m clbellard_amd.build``).  The names follow the reference's host surface:

* ``normalize`` / ``formula``          -- BuildOptions (Bellard_192.hsp:468-474), numer_denom_init
                                          (KernelFuncs.as:3-13)
* ``Context.set0mem / batch_calc /
  add_ans / read_partial / read_ans``  -- Set0mem*, BatchCalc, AddAns, HCLReadBuffer
                                          (KernelFuncs.as:15-74, Bellard_192.hsp:264-296)
* ``Context.sglobal / sglobal_after /
  ssum``                               -- the Sglobal* / Sglobal_after* / Ssum kernels
* ``Context.normal``                   -- Allloop (Bellard_192.hsp:264-296)
* ``Context.run_worklist``, ``work_sum`` -- work mode (Bellard_192.hsp:88-340, 514-533) and
                                          workmodeSum0_6.hsp
* ``hex48`` / ``hex_dd``               -- hexadecimal_i32 / hexadecimal_dd (hexsademical.as)

The library has no CPU fallback: every term is evaluated by the HIP kernels.  Loading fails
loudly if the shared library has not been built.
"""
from __future__ import annotations

import ctypes
import os
from typing import Iterable, List, Optional, Sequence, Tuple

__all__ = [
    "ClbError", "lib", "LIB_PATH", "normalize", "formula", "hex48", "hex_dd", "u192_to_dd",
    "work_sum", "parse_worklist", "write_work_file", "device_count", "device_name", "Context",
    "u192_to_int", "int_to_u192", "issue_rate", "PEAK_OPS", "read_schedule", "write_schedule",
    "SCHEDULE_KEYS",
    "OK", "EINVAL", "EDEVICE", "ERANGE", "EIO", "ENOMEM", "EFORMAT",
    "PREC_DD106", "PREC_U192", "PREC_U192_BETA", "KERNEL_AUTO", "KERNEL_PLAIN", "KERNEL_SPLIT",
    "NUMER", "NUMESIGN", "DEN0", "DEN1", "REF_GLOBAL_WORK_SIZE", "NOALLOC_MAX_JOBS",
]

OK, EINVAL, EDEVICE, ERANGE, EIO, ENOMEM, EFORMAT = 0, -1, -2, -3, -4, -5, -6
PREC_DD106, PREC_U192, PREC_U192_BETA = 0, 1, 2
KERNEL_AUTO, KERNEL_PLAIN, KERNEL_SPLIT = 0, 1, 2
REF_GLOBAL_WORK_SIZE = 65536 * 256  # the reference's work-items per launch, Bellard_192.hsp:5-7
CLOCK_XCDS = 8  # CLB_CLOCK_XCDS: head-kernel clock samples, one per XCD
NOALLOC_MAX_JOBS = 4096  # CLB_NOALLOC_MAX_JOBS

# KernelFuncs.as:9-12
NUMER = (32, 1, 256, 64, 4, 4, 1)
NUMESIGN = (1, 1, 0, 1, 1, 1, 0)
DEN0 = (4, 4, 10, 10, 10, 10, 10)
DEN1 = (1, 3, 1, 3, 5, 7, 9)

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libclbellard.so")

_ERRORS = {
    -1: "CLB_EINVAL", -2: "CLB_EDEVICE", -3: "CLB_ERANGE", -4: "CLB_EIO", -5: "CLB_ENOMEM",
    -6: "CLB_EFORMAT",
}


class ClbError(RuntimeError):
    """A negative status from libclbellard.so, with the library's message."""

    def __init__(self, code: int, message: str):
        super().__init__(f"{_ERRORS.get(code, code)}: {message}")
        self.code = code


class _Schedule(ctypes.Structure):
    _fields_ = [("device", ctypes.c_char * 256)] + [
        (k, ctypes.c_int) for k in ("terms_per_lane", "min_tiles_per_wave", "blocks_per_cu",
                                    "waves_per_group", "min_chunk", "chunks_per_wave",
                                    "max_replicas")]


SCHEDULE_KEYS = ("terms_per_lane", "min_tiles_per_wave", "blocks_per_cu", "waves_per_group",
                 "min_chunk", "chunks_per_wave", "max_replicas")


class _Job(ctypes.Structure):
    _fields_ = [("d", ctypes.c_uint64), ("k_begin", ctypes.c_uint64), ("k_end", ctypes.c_uint64),
                ("formula", ctypes.c_int32), ("reserved", ctypes.c_int32)]


_U64 = ctypes.c_uint64
_P64 = ctypes.POINTER(ctypes.c_uint64)
_CTX = ctypes.c_void_p

_SIGNATURES = {
    "clb_last_error": ([], ctypes.c_char_p),
    "clb_version": ([], ctypes.c_char_p),
    "clb_get_dev_count": ([ctypes.POINTER(ctypes.c_int)], ctypes.c_int),
    "clb_get_device_name": ([ctypes.c_int, ctypes.c_char_p, ctypes.c_size_t], ctypes.c_int),
    "clb_ctx_create": ([ctypes.c_int, ctypes.c_int, ctypes.POINTER(_CTX)], ctypes.c_int),
    "clb_ctx_destroy": ([_CTX], ctypes.c_int),
    "clb_set_kernel": ([_CTX, ctypes.c_int], ctypes.c_int),
    "clb_set_launch": ([_CTX, ctypes.c_int], ctypes.c_int),
    "clb_set_graph": ([_CTX, ctypes.c_int], ctypes.c_int),
    "clb_set_replay": ([_CTX, _U64], ctypes.c_int),
    "clb_set_tiling": ([_CTX, ctypes.c_int, ctypes.c_int], ctypes.c_int),
    "clb_set_schedule": ([_CTX, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int],
                         ctypes.c_int),
    "clb_get_schedule": ([_CTX] + [ctypes.POINTER(ctypes.c_int)] * 6, ctypes.c_int),
    "clb_autotune": ([_CTX, ctypes.c_uint64, ctypes.c_double, ctypes.POINTER(ctypes.c_int),
                      ctypes.POINTER(ctypes.c_int)], ctypes.c_int),
    "clb_read_schedule": ([ctypes.c_char_p, ctypes.POINTER(_Schedule)], ctypes.c_int),
    "clb_write_schedule": ([ctypes.c_char_p, ctypes.POINTER(_Schedule)], ctypes.c_int),
    "clb_save_schedule": ([_CTX, ctypes.c_char_p], ctypes.c_int),
    "clb_load_schedule": ([_CTX, ctypes.c_char_p], ctypes.c_int),
    "clb_normalize": ([_U64, _P64, _P64, ctypes.POINTER(ctypes.c_int)], ctypes.c_int),
    "clb_formula": ([ctypes.c_int, _P64, ctypes.POINTER(ctypes.c_int64), _P64, _P64],
                    ctypes.c_int),
    "clb_set0mem": ([_CTX], ctypes.c_int),
    "clb_batch_calc": ([_CTX, _U64, _U64, _U64, ctypes.c_int], ctypes.c_int),
    "clb_add_ans": ([_CTX, ctypes.c_int], ctypes.c_int),
    "clb_read_partial": ([_CTX, _P64], ctypes.c_int),
    "clb_read_ans": ([_CTX, _P64], ctypes.c_int),
    "clb_clear_ans": ([_CTX], ctypes.c_int),
    "clb_sglobal": ([_CTX, _U64, _U64, _U64, ctypes.c_int64, _U64, _U64], ctypes.c_int),
    "clb_sglobal_after": ([_CTX, _U64, ctypes.c_int64, _U64, _U64], ctypes.c_int),
    "clb_ssum": ([_CTX], ctypes.c_int),
    "clb_normal": ([_CTX, _U64, _P64, ctypes.c_char_p, _P64], ctypes.c_int),
    "clb_batch_positions": ([_CTX, _P64, ctypes.c_size_t, _P64], ctypes.c_int),
    "clb_run_jobs": ([_CTX, ctypes.POINTER(_Job), ctypes.c_size_t, _P64], ctypes.c_int),
    "clb_run_jobs_noalloc": ([_CTX, ctypes.POINTER(_Job), ctypes.c_size_t, _P64], ctypes.c_int),
    "clb_hex48": ([_P64, ctypes.c_char_p], ctypes.c_int),
    "clb_hex_dd": ([ctypes.POINTER(ctypes.c_double), ctypes.c_char_p], ctypes.c_int),
    "clb_u192_to_dd": ([_P64, ctypes.POINTER(ctypes.c_double)], ctypes.c_int),
    "clb_run_worklist": ([_CTX, ctypes.c_char_p, ctypes.c_char_p, _U64, ctypes.c_char_p,
                          ctypes.POINTER(ctypes.c_int)], ctypes.c_int),
    "clb_set_worklist_launch": ([_CTX, ctypes.c_uint32, ctypes.c_double, ctypes.c_uint32],
                                ctypes.c_int),
    "clb_work_sum": ([ctypes.c_char_p, ctypes.c_char_p, ctypes.POINTER(ctypes.c_int)],
                     ctypes.c_int),
    "clb_write_work_file": ([ctypes.c_char_p, ctypes.c_int, ctypes.c_int64, _P64, ctypes.c_int,
                             ctypes.c_char_p, ctypes.c_int64], ctypes.c_int),
    "clb_parse_worklist": ([ctypes.c_char_p, _P64, ctypes.POINTER(ctypes.c_int),
                            ctypes.POINTER(ctypes.c_int64), ctypes.c_size_t,
                            ctypes.POINTER(ctypes.c_size_t)], ctypes.c_int),
    "clb_issue_rate": ([ctypes.c_int, ctypes.c_int, ctypes.c_double,
                        ctypes.POINTER(ctypes.c_double)], ctypes.c_int),
    "clb_last_timing": ([_CTX, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double)],
                        ctypes.c_int),
    "clb_last_clock": ([_CTX, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double)],
                       ctypes.c_int),
}


def _load() -> ctypes.CDLL:
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            f"{LIB_PATH} is missing: build it with `python -m clbellard_amd.build` "
            "(there is no CPU fallback)")
    handle = ctypes.CDLL(LIB_PATH)
    for name, (args, res) in _SIGNATURES.items():
        fn = getattr(handle, name)
        fn.argtypes = args
        fn.restype = res
    return handle


class _Lib:
    """libclbellard.so, loaded on first use so that the package (and its build module) import
    before the library exists; any call without the built library raises ImportError."""

    _handle = None

    def __getattr__(self, name):
        if _Lib._handle is None:
            _Lib._handle = _load()
        return getattr(_Lib._handle, name)


lib = _Lib()


def _check(rc: int) -> None:
    if rc != 0:
        raise ClbError(rc, lib.clb_last_error().decode(errors="replace"))


def u192_to_int(w: Sequence[int]) -> int:
    """bigSum word order (w[0] most significant) -> integer."""
    return (int(w[0]) << 128) | (int(w[1]) << 64) | int(w[2])


def int_to_u192(v: int) -> Tuple[int, int, int]:
    v %= 1 << 192
    return (v >> 128, (v >> 64) & (2**64 - 1), v & (2**64 - 1))


def _arr3(w: Optional[Sequence[int]] = None):
    a = (ctypes.c_uint64 * 3)()
    if w is not None:
        a[0], a[1], a[2] = (int(x) for x in w)
    return a


def normalize(n: int) -> Tuple[int, int, bool]:
    """Digits N -> (N', d, overdouble), Bellard_192.hsp:468-474."""
    n_out, d_out, od = _U64(), _U64(), ctypes.c_int()
    _check(lib.clb_normalize(n, ctypes.byref(n_out), ctypes.byref(d_out), ctypes.byref(od)))
    return n_out.value, d_out.value, bool(od.value)


def formula(j: int) -> Tuple[int, int, int, int]:
    """(c_j, numesign_j, den0_j, den1_j), KernelFuncs.as:3-13."""
    c, s, a, b = _U64(), ctypes.c_int64(), _U64(), _U64()
    _check(lib.clb_formula(j, ctypes.byref(c), ctypes.byref(s), ctypes.byref(a), ctypes.byref(b)))
    return c.value, s.value, a.value, b.value


def hex48(w: Sequence[int]) -> str:
    buf = ctypes.create_string_buffer(49)
    _check(lib.clb_hex48(_arr3(w), buf))
    return buf.value.decode()


def hex_dd(hi: float, lo: float) -> str:
    dd = (ctypes.c_double * 2)(hi, lo)
    buf = ctypes.create_string_buffer(49)
    _check(lib.clb_hex_dd(dd, buf))
    return buf.value.decode()


def u192_to_dd(w: Sequence[int]) -> Tuple[float, float]:
    dd = (ctypes.c_double * 2)()
    _check(lib.clb_u192_to_dd(_arr3(w), dd))
    return dd[0], dd[1]


def work_sum(work_dir: str) -> Tuple[str, List[int]]:
    """workmodeSum0_6: (48 hex digits, files used per formula)."""
    buf = ctypes.create_string_buffer(49)
    counts = (ctypes.c_int * 7)()
    _check(lib.clb_work_sum(os.fsencode(work_dir), buf, counts))
    return buf.value.decode(), list(counts)


def write_work_file(work_dir: str, j: int, b: int, w: Sequence[int], precision: int = PREC_U192,
                    device_label: str = "", ms: int = 0) -> None:
    """fsave: <work_dir>/<j>/<j>_<b>.txt with the formatted value, a device line and the time."""
    _check(lib.clb_write_work_file(os.fsencode(work_dir), j, b, _arr3(w), precision,
                                   device_label.encode(), ms))


def read_schedule(path: str) -> dict:
    """A launch-settings file (clb_read_schedule): {"device": name, "terms_per_lane": ..., ...}."""
    s = _Schedule()
    _check(lib.clb_read_schedule(os.fsencode(path), ctypes.byref(s)))
    out = {"device": s.device.decode(errors="replace")}
    out.update({k: getattr(s, k) for k in SCHEDULE_KEYS})
    return out


def write_schedule(path: str, settings: dict) -> None:
    """Writes a launch-settings file (clb_write_schedule) from read_schedule's dict form."""
    s = _Schedule()
    s.device = settings["device"].encode()
    for k in SCHEDULE_KEYS:
        setattr(s, k, settings[k])
    _check(lib.clb_write_schedule(os.fsencode(path), ctypes.byref(s)))


def parse_worklist(path: str) -> Tuple[int, int, List[Tuple[int, int]]]:
    """Worklist -> (digits, precision (-1 if invalid), [(formula, batch), ...])."""
    digits, prec, n = _U64(), ctypes.c_int(), ctypes.c_size_t()
    _check(lib.clb_parse_worklist(os.fsencode(path), ctypes.byref(digits), ctypes.byref(prec),
                                  None, 0, ctypes.byref(n)))
    arr = (ctypes.c_int64 * (2 * max(1, n.value)))()
    _check(lib.clb_parse_worklist(os.fsencode(path), ctypes.byref(digits), ctypes.byref(prec),
                                  arr, n.value, ctypes.byref(n)))
    return digits.value, prec.value, [(arr[2 * i], arr[2 * i + 1]) for i in range(n.value)]


# Instructions (and whole modular squares) of the issue-rate microbenchmark, CLB_PEAK_* order.
PEAK_OPS = ("v_fma_f64", "v_mul_f64", "v_add_f64", "v_rndne_f64", "v_ldexp_f64", "v_mad_u64_u32",
            "v_mul_hi_u32", "v_mul_lo_u32", "v_add_u32", "v_fma_f32", "fp64 modular square",
            "int64 Montgomery square")


def issue_rate(op: int, target_ms: float = 100.0, device: int = 0) -> dict:
    """One microbenchmark run: lane-operations (or modular squares) per second, the shader clock
    measured during the run (GHz), the run's ms and the CU count."""
    out = (ctypes.c_double * 4)()
    _check(lib.clb_issue_rate(device, op, target_ms, out))
    return {"op": PEAK_OPS[op], "per_s": out[0], "sclk_ghz": out[1], "ms": out[2],
            "cus": int(out[3])}


def device_count() -> int:
    n = ctypes.c_int()
    _check(lib.clb_get_dev_count(ctypes.byref(n)))
    return n.value


def device_name(device: int = 0) -> str:
    buf = ctypes.create_string_buffer(256)
    _check(lib.clb_get_device_name(device, buf, 256))
    return buf.value.decode()


class Context:
    """A GPU context: HCLSetDev + the bigSumMem / dev_ans buffers of Bellard_192.hsp:408-420."""

    def __init__(self, device: int = 0, precision: int = PREC_U192, kernel: int = KERNEL_AUTO):
        self._ctx = _CTX()
        _check(lib.clb_ctx_create(device, precision, ctypes.byref(self._ctx)))
        self.precision = precision
        if kernel != KERNEL_AUTO:
            self.set_kernel(kernel)

    def close(self) -> None:
        if self._ctx:
            lib.clb_ctx_destroy(self._ctx)
            self._ctx = _CTX()

    def __enter__(self) -> "Context":
        return self

    def __exit__(self, *exc) -> None:
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_kernel(self, kernel: int) -> None:
        _check(lib.clb_set_kernel(self._ctx, kernel))

    def set_launch(self, blocks_per_cu: int) -> None:
        _check(lib.clb_set_launch(self._ctx, blocks_per_cu))

    def set_tiling(self, terms_per_lane: int = 0, min_tiles_per_wave: int = 0) -> None:
        """Terms per lane per tile (16, 4, 1; 0 = chosen per launch: the longest giving every
        resident wave min_tiles_per_wave tiles, or with 0 one tile to every 8th wave)."""
        _check(lib.clb_set_tiling(self._ctx, terms_per_lane, min_tiles_per_wave))

    def set_schedule(self, waves_per_group: int = 0, min_chunk: int = 0, chunks_per_wave: int = 0,
                     max_replicas: int = 0) -> None:
        """Workgroup size (waves), queue chunk (the guided queue's last), chunks per wave per phase
        of a guided queue (0: fixed chunks), most copies per accumulator slot; 0 = default."""
        _check(lib.clb_set_schedule(self._ctx, waves_per_group, min_chunk, chunks_per_wave,
                                    max_replicas))

    def get_schedule(self) -> dict:
        """The launch settings in force (clb_get_schedule)."""
        v = [ctypes.c_int() for _ in range(6)]
        _check(lib.clb_get_schedule(self._ctx, *[ctypes.byref(x) for x in v]))
        keys = ("terms_per_lane", "blocks_per_cu", "waves_per_group", "min_chunk",
                "chunks_per_wave", "max_replicas")
        return dict(zip(keys, (x.value for x in v)))

    def save_schedule(self, path: str) -> None:
        """Writes the context's launch settings with its device's name (clb_save_schedule)."""
        _check(lib.clb_save_schedule(self._ctx, os.fsencode(path)))

    def load_schedule(self, path: str) -> None:
        """Applies a launch-settings file written for this device (clb_load_schedule)."""
        _check(lib.clb_load_schedule(self._ctx, os.fsencode(path)))

    def autotune(self, n: int, budget_ms: float = 2000.0) -> Tuple[int, int]:
        """Time tile lengths and grid sizes on a sample of position n; keep and return the
        fastest (terms per lane, workgroups per CU)."""
        tpl, bpc = ctypes.c_int(), ctypes.c_int()
        _check(lib.clb_autotune(self._ctx, n, budget_ms, ctypes.byref(tpl), ctypes.byref(bpc)))
        return tpl.value, bpc.value

    def set_graph(self, enable: bool) -> None:
        """Replay each call's kernel sequence as a captured HIP graph (identical results)."""
        _check(lib.clb_set_graph(self._ctx, 1 if enable else 0))

    def set_replay(self, global_work_size: int = 0) -> None:
        """Reproduce the reference's precision-1/2 kernels bit for bit, double-double rounding
        included, as launched with global_work_size work-items (REF_GLOBAL_WORK_SIZE is the
        reference's); 0 = exact evaluation (the default)."""
        _check(lib.clb_set_replay(self._ctx, global_work_size))

    # --- KernelFuncs.as surface
    def set0mem(self) -> None:
        _check(lib.clb_set0mem(self._ctx))

    def clear_ans(self) -> None:
        _check(lib.clb_clear_ans(self._ctx))

    def batch_calc(self, d: int, k_start: int, k_end: int, j: int) -> None:
        _check(lib.clb_batch_calc(self._ctx, d, k_start, k_end, j))

    def add_ans(self, j: int) -> None:
        _check(lib.clb_add_ans(self._ctx, j))

    def read_partial(self) -> Tuple[int, int, int]:
        a = _arr3()
        _check(lib.clb_read_partial(self._ctx, a))
        return tuple(a)

    def read_ans(self) -> Tuple[int, int, int]:
        a = _arr3()
        _check(lib.clb_read_ans(self._ctx, a))
        return tuple(a)

    # --- kernel operators
    def sglobal(self, offset: int, d: int, k_max: int, numesign: int, den0: int, den1: int) -> None:
        _check(lib.clb_sglobal(self._ctx, offset, d, k_max, numesign, den0, den1))

    def sglobal_after(self, d: int, numesign: int, den0: int, den1: int) -> None:
        _check(lib.clb_sglobal_after(self._ctx, d, numesign, den0, den1))

    def ssum(self) -> None:
        _check(lib.clb_ssum(self._ctx))

    # --- whole computations
    def normal(self, n: int) -> Tuple[Tuple[int, int, int], str, int]:
        """Allloop: (frac words, display hex, N')."""
        a, buf, n_out = _arr3(), ctypes.create_string_buffer(49), _U64()
        _check(lib.clb_normal(self._ctx, n, a, buf, ctypes.byref(n_out)))
        return tuple(a), buf.value.decode(), n_out.value

    def batch_positions(self, positions: Iterable[int]) -> List[Tuple[int, int, int]]:
        pos = [int(p) for p in positions]
        n = len(pos)
        if n == 0:
            return []
        arr = (ctypes.c_uint64 * n)(*pos)
        out = (ctypes.c_uint64 * (3 * n))()
        _check(lib.clb_batch_positions(self._ctx, arr, n, out))
        return [tuple(out[3 * i:3 * i + 3]) for i in range(n)]

    def run_jobs(self, jobs: Sequence[Tuple[int, int, int, int]]) -> List[Tuple[int, int, int]]:
        """jobs: (d, k_begin, k_end, formula) -> c_j * partial mod 1 per job."""
        n = len(jobs)
        if n == 0:
            return []
        arr = (_Job * n)()
        for i, (d, kb, ke, j) in enumerate(jobs):
            arr[i].d, arr[i].k_begin, arr[i].k_end, arr[i].formula = d, kb, ke, j
        out = (ctypes.c_uint64 * (3 * n))()
        _check(lib.clb_run_jobs(self._ctx, arr, n, out))
        return [tuple(out[3 * i:3 * i + 3]) for i in range(n)]

    def run_jobs_noalloc(self, jobs: Sequence[Tuple[int, int, int, int]]
                         ) -> List[Tuple[int, int, int]]:
        """run_jobs through clb_run_jobs_noalloc: no allocation in the call (at most
        NOALLOC_MAX_JOBS jobs)."""
        n = len(jobs)
        if n == 0:
            return []
        arr = (_Job * n)()
        for i, (d, kb, ke, j) in enumerate(jobs):
            arr[i].d, arr[i].k_begin, arr[i].k_end, arr[i].formula = d, kb, ke, j
        out = (ctypes.c_uint64 * (3 * n))()
        _check(lib.clb_run_jobs_noalloc(self._ctx, arr, n, out))
        return [tuple(out[3 * i:3 * i + 3]) for i in range(n)]

    def run_worklist(self, worklist: str, work_dir: str, batch_size: int = 0,
                     device_label: Optional[str] = None) -> int:
        ran = ctypes.c_int()
        label = device_label.encode() if device_label is not None else None
        _check(lib.clb_run_worklist(self._ctx, os.fsencode(worklist), os.fsencode(work_dir),
                                    batch_size, label, ctypes.byref(ran)))
        return ran.value

    def set_worklist_launch(self, max_jobs: int = 0, max_seconds: float = 60.0,
                            max_launches: int = 0) -> None:
        """Work mode: jobs per launch (0 = any), estimated seconds per launch (0 = any) and
        launches per run_worklist call (0 = until done); files are written launch by launch."""
        _check(lib.clb_set_worklist_launch(self._ctx, max_jobs, max_seconds, max_launches))

    def last_timing(self) -> Tuple[float, float]:
        """(head-term kernel ms, whole call ms) of the last computation, from HIP events."""
        h, t = ctypes.c_double(), ctypes.c_double()
        _check(lib.clb_last_timing(self._ctx, ctypes.byref(h), ctypes.byref(t)))
        return h.value, t.value

    def last_clock(self) -> Tuple[float, List[float]]:
        """(mean GHz, GHz per XCD) of the last call's head-term kernel, from the shader-clock and
        real-time counters its first wave on each XCD read at the start and at the end."""
        per = (ctypes.c_double * CLOCK_XCDS)()
        mean = ctypes.c_double()
        _check(lib.clb_last_clock(self._ctx, per, ctypes.byref(mean)))
        return mean.value, list(per)
