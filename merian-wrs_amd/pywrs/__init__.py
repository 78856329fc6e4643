"""ctypes binding of libwrs_amd.so (the C ABI in include/wrs.h and
include/wrs_runtime.h) with numpy host arrays.

Device memory is owned by :class:`DeviceArray` objects allocated through the
library's own runtime entry points, so nothing here needs torch or HIP headers.
Loading fails loudly when the shared library is missing: there is no CPU
fallback for any operation.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(os.path.dirname(_HERE), "libwrs_amd.so")

WRS_OK, WRS_EINVAL, WRS_ENOMEM, WRS_EHIP, WRS_ENOTSUP, WRS_EWAIT = range(6)
PREC_FP32, PREC_FP64_CARRY, PREC_COMPENSATED, PREC_FP64 = 0, 1, 2, 3
GEN_UNIFORM01, GEN_LOGNORMAL, GEN_ZIPF, GEN_SPIKE = 0, 1, 2, 3
PACK_DEFAULT, PACK_DIRECT, PACK_CLASSIC, PACK_SERIAL, PACK_INLINE, PACK_SUBGROUP, PACK_STREAM = (
    0, 1, 2, 3, 4, 5, 6)  # wrs_tuning.direct

ALIAS_DTYPE = np.dtype([("p", np.float32), ("a", np.uint32)])
SPLIT_DTYPE = np.dtype([("i", np.uint32), ("j", np.uint32), ("spill", np.float64)])

# every symbol the headers declare (checked by the CPU test suite)
EXPORTS = [
    "wrs_plan_create", "wrs_plan_destroy", "wrs_psa_build", "wrs_psa_build_debug", "wrs_mean",
    "wrs_prefix_partition", "wrs_scan_inclusive", "wrs_its_build", "wrs_alias_sample",
    "wrs_its_sample", "wrs_its_sample_noalloc", "wrs_alias_sample_window", "wrs_its_sample_window", "wrs_cutpoint_build", "wrs_cutpoint_sample", "wrs_philox", "wrs_generate",
    "wrs_batch_plan_create",
    "wrs_batch_plan_destroy", "wrs_psa_build_batched", "wrs_alias_sample_batched",
    "wrs_autotune_psa", "wrs_autotune_batched", "wrs_histogram", "wrs_sample_stats",
    "wrs_last_error", "wrs_version", "wrs_plan_set_timing", "wrs_plan_set_tuning", "wrs_plan_phase_times",
    "wrs_rt_device_count", "wrs_rt_device_info", "wrs_rt_malloc", "wrs_rt_free",
    "wrs_rt_memcpy_htod", "wrs_rt_memcpy_dtoh", "wrs_rt_memcpy_dtod", "wrs_rt_memset",
    "wrs_rt_stream_create", "wrs_rt_stream_destroy", "wrs_rt_stream_sync", "wrs_rt_device_sync",
    "wrs_rt_event_create", "wrs_rt_event_destroy", "wrs_rt_event_record", "wrs_rt_event_elapsed",
    "wrs_rt_copy_f32", "wrs_rt_psa_build_sample_each", "wrs_rt_graph_begin", "wrs_rt_graph_end",
    "wrs_rt_graph_launch", "wrs_rt_graph_destroy", "wrs_plan_status", "wrs_batch_plan_status",
    "wrs_plan_set_test_hooks",
]


class WrsError(RuntimeError):
    def __init__(self, status: int, msg: str):
        super().__init__(f"wrs status {status}: {msg}")
        self.status = status


class Tuning(C.Structure):
    """wrs_tuning: launch geometry (zero fields take the defaults)."""
    _fields_ = [("merge_tile", C.c_uint32), ("merge_per", C.c_uint32),
                ("partition", C.c_uint32), ("direct", C.c_uint32)]

    def as_dict(self) -> dict:
        return dict(merge_tile=self.merge_tile, merge_per=self.merge_per,
                    partition=self.partition, direct=self.direct)


def tuning(merge_tile: int = 0, merge_per: int = 0, partition: int = 0,
           direct: int = 0) -> Tuning:
    return Tuning(int(merge_tile), int(merge_per), int(partition), int(direct))


class Config(C.Structure):
    _fields_ = [("split_size", C.c_uint32), ("precision", C.c_int), ("tuning", Tuning)]


class TuneEntry(C.Structure):
    _fields_ = [("tuning", Tuning), ("ms", C.c_float)]


TUNE_NO_CACHE = 1


def _config(split_size: int, precision: int, tune) -> Config:
    if tune is None:
        tune = Tuning()
    elif isinstance(tune, dict):
        tune = tuning(**tune)
    return Config(int(split_size), int(precision), tune)


class PsaDebug(C.Structure):
    _fields_ = [("mean", C.c_void_p), ("heavy_count", C.c_void_p), ("indices", C.c_void_p),
                ("prefix_hi", C.c_void_p), ("prefix_lo", C.c_void_p), ("splits", C.c_void_p)]


_lib = None


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is not built: run __graft_entry__.build() "
                          "(python merian-wrs_amd/build.py)")
    L = C.CDLL(LIB_PATH)
    vp, u32, u64, i32, f64 = C.c_void_p, C.c_uint32, C.c_uint64, C.c_int, C.c_double
    sig = {
        "wrs_plan_create": [vp, u64, C.POINTER(vp)],
        "wrs_plan_destroy": [vp],
        "wrs_psa_build": [vp, vp, u64, vp, vp],
        "wrs_psa_build_debug": [vp, vp, u64, vp, C.POINTER(PsaDebug), vp],
        "wrs_mean": [vp, vp, u64, vp, vp],
        "wrs_prefix_partition": [vp, vp, u64, f64, vp, vp, vp, vp, vp],
        "wrs_scan_inclusive": [vp, vp, u64, vp, vp, vp],
        "wrs_its_build": [vp, vp, u64, vp, vp, vp],
        "wrs_alias_sample": [vp, u64, u64, u64, u64, vp, vp],
        "wrs_its_sample": [vp, vp, u64, u64, u64, u64, vp, vp],
        "wrs_its_sample_noalloc": [vp, vp, u64, u64, u64, u64, vp, vp],
        "wrs_alias_sample_window": [vp, u64, u64, u64, u64, u32, vp, vp],
        "wrs_its_sample_window": [vp, vp, u64, u64, u64, u64, u32, i32, vp, vp],
        "wrs_cutpoint_build": [vp, vp, u64, vp, u64, vp],
        "wrs_cutpoint_sample": [vp, vp, u64, vp, u64, u64, u64, u64, vp, vp],
        "wrs_philox": [u64, u32, u64, u64, vp, vp],
        "wrs_generate": [i32, f64, f64, u64, u64, vp, vp],
        "wrs_batch_plan_create": [vp, vp, u32, C.POINTER(vp)],
        "wrs_batch_plan_destroy": [vp],
        "wrs_psa_build_batched": [vp, vp, vp, vp],
        "wrs_alias_sample_batched": [vp, vp, u32, u32, u64, u64, vp, vp],
        "wrs_autotune_psa": [C.POINTER(Config), vp, u64, u32, C.POINTER(Tuning),
                             C.POINTER(TuneEntry), u32, C.POINTER(u32), vp],
        "wrs_autotune_batched": [C.POINTER(Config), vp, u32, vp, u32, C.POINTER(Tuning),
                                 C.POINTER(TuneEntry), u32, C.POINTER(u32), vp],
        "wrs_histogram": [vp, u64, u64, vp, i32, vp],
        "wrs_sample_stats": [vp, vp, u64, vp, u64, f64, vp, vp],
        "wrs_plan_set_timing": [vp, i32],
        "wrs_plan_set_tuning": [vp, vp],
        "wrs_plan_phase_times": [vp, vp, u32],
        "wrs_plan_status": [vp, vp],
        "wrs_batch_plan_status": [vp, vp],
        "wrs_plan_set_test_hooks": [vp, u32, u32],
        "wrs_rt_device_count": [C.POINTER(C.c_int)],
        "wrs_rt_device_info": [C.c_char_p, C.c_size_t, C.c_char_p, C.c_size_t,
                               C.POINTER(C.c_int), C.POINTER(C.c_uint64)],
        "wrs_rt_malloc": [C.POINTER(vp), C.c_size_t],
        "wrs_rt_free": [vp],
        "wrs_rt_memcpy_htod": [vp, vp, C.c_size_t],
        "wrs_rt_memcpy_dtoh": [vp, vp, C.c_size_t],
        "wrs_rt_memcpy_dtod": [vp, vp, C.c_size_t, vp],
        "wrs_rt_memset": [vp, i32, C.c_size_t, vp],
        "wrs_rt_stream_create": [C.POINTER(vp)],
        "wrs_rt_stream_destroy": [vp],
        "wrs_rt_stream_sync": [vp],
        "wrs_rt_device_sync": [],
        "wrs_rt_event_create": [C.POINTER(vp)],
        "wrs_rt_event_destroy": [vp],
        "wrs_rt_event_record": [vp, vp],
        "wrs_rt_event_elapsed": [C.POINTER(C.c_float), vp, vp],
        "wrs_rt_copy_f32": [vp, vp, u64, vp],
        "wrs_rt_psa_build_sample_each": [vp, vp, vp, u32, vp, u32, u64, u64, vp, vp],
        "wrs_rt_graph_begin": [vp],
        "wrs_rt_graph_end": [vp, C.POINTER(vp)],
        "wrs_rt_graph_launch": [vp, vp],
        "wrs_rt_graph_destroy": [vp],
    }
    for name, args in sig.items():
        f = getattr(L, name)
        f.argtypes = args
        f.restype = C.c_int
    L.wrs_last_error.restype = C.c_char_p
    L.wrs_last_error.argtypes = []
    L.wrs_version.restype = C.c_char_p
    L.wrs_version.argtypes = []
    _lib = L
    return L


def check(status: int) -> None:
    if status != WRS_OK:
        raise WrsError(status, lib().wrs_last_error().decode(errors="replace"))


def version() -> str:
    return lib().wrs_version().decode()


def device_count() -> int:
    n = C.c_int(0)
    check(lib().wrs_rt_device_count(C.byref(n)))
    return n.value


def device_info() -> dict:
    name = C.create_string_buffer(256)
    arch = C.create_string_buffer(256)
    cus = C.c_int(0)
    mem = C.c_uint64(0)
    check(lib().wrs_rt_device_info(name, 256, arch, 256, C.byref(cus), C.byref(mem)))
    return dict(name=name.value.decode(), arch=arch.value.decode(), cu_count=cus.value,
                hbm_bytes=mem.value)


def synchronize() -> None:
    check(lib().wrs_rt_device_sync())


# ------------------------------------------------------------ device arrays --
class DeviceArray:
    """A device allocation holding `count` elements of numpy dtype `dtype`."""

    def __init__(self, count: int, dtype):
        self.dtype = np.dtype(dtype)
        self.count = int(count)
        self.nbytes = self.count * self.dtype.itemsize
        p = C.c_void_p()
        check(lib().wrs_rt_malloc(C.byref(p), self.nbytes))
        self.ptr = p.value

    @classmethod
    def from_numpy(cls, arr: np.ndarray) -> "DeviceArray":
        arr = np.ascontiguousarray(arr)
        d = cls(arr.size, arr.dtype)
        if arr.size:
            check(lib().wrs_rt_memcpy_htod(d.ptr, arr.ctypes.data, d.nbytes))
        return d

    def numpy(self, count: int | None = None) -> np.ndarray:
        count = self.count if count is None else int(count)
        out = np.empty(count, self.dtype)
        if count:
            check(lib().wrs_rt_memcpy_dtoh(out.ctypes.data, self.ptr, count * self.dtype.itemsize))
        return out

    def offset(self, elements: int) -> int:
        return self.ptr + int(elements) * self.dtype.itemsize

    def fill_bytes(self, value: int = 0, stream=None) -> None:
        check(lib().wrs_rt_memset(self.ptr, value, self.nbytes, stream))

    def free(self) -> None:
        if getattr(self, "ptr", None):
            lib().wrs_rt_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def empty(count: int, dtype) -> DeviceArray:
    return DeviceArray(count, dtype)


def to_device(arr) -> DeviceArray:
    return DeviceArray.from_numpy(np.asarray(arr))


# ------------------------------------------------------------- streams/events --
class Stream:
    def __init__(self):
        s = C.c_void_p()
        check(lib().wrs_rt_stream_create(C.byref(s)))
        self.handle = s.value

    def synchronize(self) -> None:
        check(lib().wrs_rt_stream_sync(self.handle))

    def __del__(self):
        try:
            if self.handle:
                lib().wrs_rt_stream_destroy(self.handle)
                self.handle = None
        except Exception:
            pass


class Event:
    def __init__(self):
        e = C.c_void_p()
        check(lib().wrs_rt_event_create(C.byref(e)))
        self.handle = e.value

    def record(self, stream=None) -> None:
        check(lib().wrs_rt_event_record(self.handle, _h(stream)))

    def elapsed_ms(self, end: "Event") -> float:
        ms = C.c_float(0.0)
        check(lib().wrs_rt_event_elapsed(C.byref(ms), self.handle, end.handle))
        return ms.value

    def __del__(self):
        try:
            if self.handle:
                lib().wrs_rt_event_destroy(self.handle)
                self.handle = None
        except Exception:
            pass


def _h(stream):
    if stream is None:
        return None
    return stream.handle if isinstance(stream, Stream) else stream


def _p(x):
    if x is None:
        return None
    return x.ptr if isinstance(x, DeviceArray) else int(x)


# ------------------------------------------------------------------- plans --
class Plan:
    """wrs_plan: build workspace for distributions of up to max_n weights."""

    def __init__(self, max_n: int, split_size: int = 32, precision: int = PREC_COMPENSATED,
                 tune=None):
        self.max_n = int(max_n)
        self.split_size = int(split_size)
        self.precision = int(precision)
        cfg = _config(self.split_size, self.precision, tune)
        p = C.c_void_p()
        check(lib().wrs_plan_create(C.byref(cfg), self.max_n, C.byref(p)))
        self.handle = p.value

    def split_count(self, n: int) -> int:
        return max(1, -(-int(n) // self.split_size))

    def psa_build(self, weights, n: int, table, stream=None) -> None:
        check(lib().wrs_psa_build(self.handle, _p(weights), n, _p(table), _h(stream)))

    def psa_build_debug(self, weights, n: int, table, dbg: dict, stream=None) -> None:
        d = PsaDebug(*[_p(dbg.get(k)) for k in
                       ("mean", "heavy_count", "indices", "prefix_hi", "prefix_lo", "splits")])
        check(lib().wrs_psa_build_debug(self.handle, _p(weights), n, _p(table), C.byref(d),
                                        _h(stream)))

    def mean(self, weights, n: int, out2, stream=None) -> None:
        check(lib().wrs_mean(self.handle, _p(weights), n, _p(out2), _h(stream)))

    def prefix_partition(self, weights, n: int, pivot: float, indices, prefix_hi, prefix_lo,
                         heavy_count, stream=None) -> None:
        check(lib().wrs_prefix_partition(self.handle, _p(weights), n, float(pivot), _p(indices),
                                         _p(prefix_hi), _p(prefix_lo), _p(heavy_count),
                                         _h(stream)))

    def scan_inclusive(self, x, n: int, out_hi, out_lo=None, stream=None) -> None:
        check(lib().wrs_scan_inclusive(self.handle, _p(x), n, _p(out_hi), _p(out_lo), _h(stream)))

    def its_build(self, weights, n: int, cdf_hi, cdf_lo=None, stream=None) -> None:
        check(lib().wrs_its_build(self.handle, _p(weights), n, _p(cdf_hi), _p(cdf_lo),
                                  _h(stream)))

    def sample_stats(self, weights, n: int, hist, s: int, min_expected: float = 32.0,
                     stream=None) -> dict:
        """chi-square (pooled), squared error and zero-mass hits of a histogram
        (wrs.histogram) of s samples against weights, computed on the device."""
        out = empty(6, np.float64)
        check(lib().wrs_sample_stats(self.handle, _p(weights), n, _p(hist), s,
                                     float(min_expected), _p(out), _h(stream)))
        synchronize()
        v = out.numpy()
        chi2, dof, sse, esse, zh, pools = v.tolist()
        return dict(chi2=chi2, df=int(dof), sse=sse, expected_sse=esse, zero_mass_hits=int(zh),
                    pools=int(pools), rmse=float(np.sqrt(sse / n)),
                    expected_rmse=float(np.sqrt(esse / n)))

    def set_timing(self, enable: bool) -> None:
        check(lib().wrs_plan_set_timing(self.handle, int(bool(enable))))

    def set_tuning(self, tune) -> None:
        """Launch geometry of later builds (a dict as `tune` of the constructor)."""
        t = tuning(**tune) if isinstance(tune, dict) else (tune if tune is not None else Tuning())
        check(lib().wrs_plan_set_tuning(self.handle, C.byref(t)))

    PHASES = ("reduce", "finalize", "tile_aggregate", "tile_scan", "partition_emit",
              "boundary_splits", "merge_pack")

    def phase_times(self) -> list:
        """ms of the kernels of the last timed build, in the order of PHASES."""
        ms = (C.c_float * len(self.PHASES))()
        check(lib().wrs_plan_phase_times(self.handle, ms, len(self.PHASES)))
        return list(ms)

    def status(self, stream=None) -> None:
        """Raises WrsError(WRS_EWAIT) if a device-side wait of the most recent
        build on this plan gave up (its outputs are then invalid)."""
        check(lib().wrs_plan_status(self.handle, _h(stream)))

    def set_test_hooks(self, wait_polls: int = 0, hooks: int = 0) -> None:
        """Failure injection for error-path tests (0, 0: normal builds)."""
        check(lib().wrs_plan_set_test_hooks(self.handle, int(wait_polls), int(hooks)))

    def close(self) -> None:
        if getattr(self, "handle", None):
            lib().wrs_plan_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class BatchPlan:
    """wrs_batch_plan: segmented build of many distributions (CSR offsets)."""

    def __init__(self, offsets: np.ndarray, split_size: int = 32,
                 precision: int = PREC_COMPENSATED, tune=None):
        self.offsets = np.ascontiguousarray(offsets, np.uint64)
        self.nseg = self.offsets.size - 1
        cfg = _config(int(split_size), int(precision), tune)
        p = C.c_void_p()
        check(lib().wrs_batch_plan_create(C.byref(cfg), self.offsets.ctypes.data, self.nseg,
                                          C.byref(p)))
        self.handle = p.value

    def build(self, weights, tables, stream=None) -> None:
        check(lib().wrs_psa_build_batched(self.handle, _p(weights), _p(tables), _h(stream)))

    def status(self, stream=None) -> None:
        """Raises WrsError(WRS_EWAIT) if a device-side wait of the last batch gave up."""
        check(lib().wrs_batch_plan_status(self.handle, _h(stream)))

    def close(self) -> None:
        if getattr(self, "handle", None):
            lib().wrs_batch_plan_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# -------------------------------------------------------------- autotuner --
def _tune_result(best: Tuning, log, count: int) -> tuple:
    entries = [dict(log[i].tuning.as_dict(), ms=log[i].ms) for i in range(min(count, len(log)))]
    return best.as_dict(), entries


def autotune_psa(weights, n: int, precision: int = PREC_COMPENSATED, split_size: int = 32,
                 no_cache: bool = False, stream=None) -> tuple:
    """(best tuning dict, [candidate dicts with 'ms']) for a PSA build of n weights."""
    cfg = _config(split_size, precision, None)
    best = Tuning()
    log = (TuneEntry * 32)()
    cnt = C.c_uint32(0)
    check(lib().wrs_autotune_psa(C.byref(cfg), _p(weights), int(n),
                                 TUNE_NO_CACHE if no_cache else 0, C.byref(best), log, 32,
                                 C.byref(cnt), _h(stream)))
    return _tune_result(best, log, cnt.value)


def autotune_batched(offsets: np.ndarray, weights, precision: int = PREC_COMPENSATED,
                     split_size: int = 32, no_cache: bool = False, stream=None) -> tuple:
    """(best tuning dict, [candidate dicts with 'ms']) for a batched build."""
    offsets = np.ascontiguousarray(offsets, np.uint64)
    cfg = _config(split_size, precision, None)
    best = Tuning()
    log = (TuneEntry * 32)()
    cnt = C.c_uint32(0)
    check(lib().wrs_autotune_batched(C.byref(cfg), offsets.ctypes.data, offsets.size - 1,
                                     _p(weights), TUNE_NO_CACHE if no_cache else 0,
                                     C.byref(best), log, 32, C.byref(cnt), _h(stream)))
    return _tune_result(best, log, cnt.value)


# --------------------------------------------------------------- sampling --
def alias_sample(table, n: int, s: int, seed: int, out, counter_base: int = 0,
                 stream=None) -> None:
    check(lib().wrs_alias_sample(_p(table), n, s, seed, counter_base, _p(out), _h(stream)))


def its_sample(cdf_hi, cdf_lo, n: int, s: int, seed: int, out, counter_base: int = 0,
               stream=None) -> None:
    check(lib().wrs_its_sample(_p(cdf_hi), _p(cdf_lo), n, s, seed, counter_base, _p(out),
                               _h(stream)))


ITS_NOALLOC_MAX_N = 1 << 25


def its_sample_noalloc(cdf_hi, cdf_lo, n: int, s: int, seed: int, out, counter_base: int = 0,
                       stream=None) -> None:
    """its_sample's samples with the search index in the library's static
    device array (no allocation; n <= ITS_NOALLOC_MAX_N; calls overlapping on
    one GPU race)."""
    check(lib().wrs_its_sample_noalloc(_p(cdf_hi), _p(cdf_lo), n, s, seed, counter_base,
                                       _p(out), _h(stream)))


def alias_sample_window(table, n: int, s: int, seed: int, window: int, out,
                        counter_base: int = 0, stream=None) -> None:
    """Windowed alias sampling: groups of 64 counters share a window of buckets
    (exact marginals, samples of a group not independent)."""
    check(lib().wrs_alias_sample_window(_p(table), n, s, seed, counter_base, int(window), _p(out),
                                        _h(stream)))


def its_sample_window(cdf_hi, cdf_lo, n: int, s: int, seed: int, window: int, out,
                      parts: bool = True, counter_base: int = 0, stream=None) -> None:
    """Windowed ITS sampling: groups of 64 counters share a window of <= window
    CDF entries (exact marginals, samples of a group not independent)."""
    check(lib().wrs_its_sample_window(_p(cdf_hi), _p(cdf_lo), n, s, seed, counter_base,
                                      int(window), int(bool(parts)), _p(out), _h(stream)))


def histogram(samples, s: int, n: int, hist, clear: bool = True, stream=None) -> None:
    """hist[0..n] (n + 1 uint64; hist[n] counts samples >= n)."""
    check(lib().wrs_histogram(_p(samples), s, n, _p(hist), int(bool(clear)), _h(stream)))


def cutpoint_build(cdf_hi, cdf_lo, n: int, guide, m: int, stream=None) -> None:
    """guide[0..m] over a normalised CDF (guide holds m + 1 uint32)."""
    check(lib().wrs_cutpoint_build(_p(cdf_hi), _p(cdf_lo), n, _p(guide), m, _h(stream)))


def cutpoint_sample(cdf_hi, cdf_lo, n: int, guide, m: int, s: int, seed: int, out,
                    counter_base: int = 0, stream=None) -> None:
    check(lib().wrs_cutpoint_sample(_p(cdf_hi), _p(cdf_lo), n, _p(guide), m, s, seed,
                                    counter_base, _p(out), _h(stream)))


def alias_sample_batched(tables, offsets_dev, nseg: int, per_segment: int, seed: int, out,
                         counter_base: int = 0, stream=None) -> None:
    check(lib().wrs_alias_sample_batched(_p(tables), _p(offsets_dev), nseg, per_segment, seed,
                                         counter_base, _p(out), _h(stream)))


def build_sample_each(plan: Plan, weights, offsets: np.ndarray, tables, per_segment: int,
                      seed: int, samples, counter_base: int = 0, stream=None) -> None:
    offsets = np.ascontiguousarray(offsets, np.uint64)
    check(lib().wrs_rt_psa_build_sample_each(plan.handle, _p(weights), offsets.ctypes.data,
                                             offsets.size - 1, _p(tables), per_segment, seed,
                                             counter_base, _p(samples), _h(stream)))


def philox(seed: int, n: int, out, base: int = 0, stream_word: int = 0, stream=None) -> None:
    check(lib().wrs_philox(seed, stream_word, base, n, _p(out), _h(stream)))


def generate(kind: int, n: int, seed: int, out, a: float = 0.0, b: float = 0.0,
             stream=None) -> None:
    check(lib().wrs_generate(kind, a, b, seed, n, _p(out), _h(stream)))


def copy_f32(dst, src, n: int, stream=None) -> None:
    check(lib().wrs_rt_copy_f32(_p(dst), _p(src), n, _h(stream)))


class Graph:
    """Everything issued on `stream` inside the `with` block, captured once."""

    def __init__(self, stream: Stream):
        self.stream = stream
        self.handle = None

    def __enter__(self):
        check(lib().wrs_rt_graph_begin(self.stream.handle))
        return self

    def __exit__(self, exc_type, exc, tb):
        g = C.c_void_p()
        st = lib().wrs_rt_graph_end(self.stream.handle, C.byref(g))
        if exc_type is None:
            check(st)
            self.handle = g.value
        return False

    def launch(self) -> None:
        check(lib().wrs_rt_graph_launch(self.handle, self.stream.handle))

    def __del__(self):
        try:
            if self.handle:
                lib().wrs_rt_graph_destroy(self.handle)
                self.handle = None
        except Exception:
            pass


# -------------------------------------------------------------- conveniences --
def psa_alias_table(weights: np.ndarray, split_size: int = 32,
                    precision: int = PREC_COMPENSATED) -> np.ndarray:
    """Host convenience: upload, build, download (structured ALIAS_DTYPE array)."""
    w = np.ascontiguousarray(weights, np.float32)
    n = w.size
    plan = Plan(n, split_size, precision)
    dw = to_device(w)
    dt = empty(n, ALIAS_DTYPE)
    plan.psa_build(dw, n, dt)
    synchronize()
    return dt.numpy()


def segment_offsets(sizes) -> np.ndarray:
    sizes = np.asarray(sizes, np.uint64)
    off = np.zeros(sizes.size + 1, np.uint64)
    np.cumsum(sizes, out=off[1:])
    return off


def segment_sizes(seed: int, nseg: int, lo: int = 256, hi: int = 1 << 18) -> np.ndarray:
    """Log-uniform segment sizes in [lo, hi]: size_b = floor(lo * (hi/lo)^u_b) with
    u_b the 53-bit uniform of Philox4x32-10 at counter (b, 0, "SEG1", 0), key = seed."""
    d = empty(4 * nseg, np.uint32)
    philox(seed, nseg, d, base=0, stream_word=0x53454731)
    synchronize()
    r = d.numpy().reshape(nseg, 4).astype(np.uint64)
    u = (((r[:, 0] << np.uint64(32)) | r[:, 1]) >> np.uint64(11)).astype(np.float64) * 2.0 ** -53
    v = np.floor(float(lo) * np.power(float(hi) / float(lo), u)).astype(np.uint64)
    return np.clip(v, lo, hi).astype(np.uint64)
