"""Python (ctypes + numpy) front end of the CPU oracle in oracle/wrs_oracle.c.

TEST INFRASTRUCTURE ONLY: imported by tests/, by __graft_entry__.smoke() and by
bench.py's cpu_baseline leg, always as the checker / the CPU baseline, never as
part of the product path.  Every function restates the sequential CPU reference
of kistenklaus/merian-wrs; the C source names the reference file:line of each.
"""
from __future__ import annotations

import ctypes as C
import math
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

u32p = np.ctypeslib.ndpointer(dtype=np.uint32, flags="C_CONTIGUOUS")
u64p = np.ctypeslib.ndpointer(dtype=np.uint64, flags="C_CONTIGUOUS")
f32p = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")
f64p = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")


class Split32(C.Structure):
    _fields_ = [("i", C.c_uint32), ("j", C.c_uint32), ("spill", C.c_float)]


class Split64(C.Structure):
    _fields_ = [("i", C.c_uint32), ("j", C.c_uint32), ("spill", C.c_double)]


class PsaDebug(C.Structure):
    _fields_ = [
        ("mean", C.c_double),
        ("heavy_count", C.c_uint32),
        ("idx_cpu", C.c_void_p),
        ("heavy_prefix", C.c_void_p),
        ("light_prefix", C.c_void_p),
        ("splits", C.c_void_p),
    ]


SPLIT_REFERENCE_CPU = 0  # src/host/reference/split.hpp:14-57
SPLIT_LOWER_BOUND = 1    # src/device/wrs/alias/psa/split/scalar/shader.comp:72-109

PACK_SCALAR = 0   # pack.hpp:14-83 pack2 (ScalarPack, pack/scalar/float.comp:60-158)
PACK_GROUPED = 1  # SubgroupPack, pack/subgroup/shader.comp:68-246

SPLIT64_DTYPE = np.dtype([("i", np.uint32), ("j", np.uint32), ("spill", np.float64)])
SPLIT32_DTYPE = np.dtype([("i", np.uint32), ("j", np.uint32), ("spill", np.float32)])


def build() -> None:
    subprocess.run(["bash", os.path.join(_HERE, "build_oracle.sh")], check=True)


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(_LIB_PATH):
        build()
    L = C.CDLL(_LIB_PATH)
    L.oro_philox4x32_10.argtypes = [u32p, u32p, u32p]
    L.oro_philox_block.argtypes = [C.c_uint64, C.c_uint32, C.c_uint64, C.c_uint64, u32p]
    L.oro_gen_uniform01.argtypes = [C.c_uint64, C.c_uint64, f32p]
    L.oro_gen_lognormal.argtypes = [C.c_uint64, C.c_double, C.c_double, C.c_uint64, f32p]
    L.oro_feistel_perm.argtypes = [C.c_uint64, C.c_uint64, C.c_uint64]
    L.oro_feistel_perm.restype = C.c_uint64
    L.oro_gen_zipf.argtypes = [C.c_uint64, C.c_double, C.c_uint64, f32p]
    L.oro_gen_spike.argtypes = [C.c_uint64, C.c_uint64, f32p]
    L.oro_segment_sizes.argtypes = [C.c_uint64, C.c_uint32, C.c_uint64, C.c_uint64, u64p]
    L.oro_kahan_reduce_f32.argtypes = [f32p, C.c_uint64]
    L.oro_kahan_reduce_f32.restype = C.c_float
    L.oro_kahan_reduce_f64.argtypes = [f64p, C.c_uint64]
    L.oro_kahan_reduce_f64.restype = C.c_double
    L.oro_stable_partition_indices_f32.argtypes = [f32p, C.c_uint32, C.c_float, u32p]
    L.oro_stable_partition_indices_f32.restype = C.c_uint32
    L.oro_stable_partition_indices_f64.argtypes = [f64p, C.c_uint32, C.c_double, u32p]
    L.oro_stable_partition_indices_f64.restype = C.c_uint32
    L.oro_prefix_sum_f32.argtypes = [f32p, C.c_uint64, f32p]
    L.oro_prefix_sum_f64.argtypes = [f64p, C.c_uint64, f64p]
    L.oro_ceil_mul_div.argtypes = [C.c_uint64, C.c_uint64, C.c_uint64]
    L.oro_ceil_mul_div.restype = C.c_uint64
    L.oro_split_f32.argtypes = [f32p, C.c_uint32, f32p, C.c_uint32, C.c_float, C.c_uint32]
    L.oro_split_f32.restype = Split32
    L.oro_split_f64.argtypes = [f64p, C.c_uint32, f64p, C.c_uint32, C.c_double, C.c_uint32]
    L.oro_split_f64.restype = Split64
    L.oro_split_lb_f64.argtypes = [f64p, C.c_uint32, f64p, C.c_uint32, C.c_double, C.c_uint32]
    L.oro_split_lb_f64.restype = Split64
    L.oro_splitK_f64.argtypes = [f64p, C.c_uint32, f64p, C.c_uint32, C.c_double,
                                 C.c_uint32, C.c_uint32, C.c_int, C.c_void_p]
    L.oro_splitK_f32.argtypes = [f32p, C.c_uint32, f32p, C.c_uint32, C.c_float,
                                 C.c_uint32, C.c_uint32, C.c_int, C.c_void_p]
    L.oro_sweep_states_f64.argtypes = [u32p, C.c_uint32, u32p, C.c_uint32, f64p, C.c_double,
                                       u32p, u32p]
    L.oro_pack_splits_f64.argtypes = [u32p, C.c_uint32, u32p, C.c_uint32, f64p, C.c_uint32,
                                      C.c_double, C.c_void_p, C.c_uint32, f64p, u32p]
    L.oro_pack_grouped_f64.argtypes = [u32p, C.c_uint32, u32p, C.c_uint32, f64p, C.c_uint32,
                                       C.c_double, C.c_void_p, C.c_uint32, f64p, u32p]
    L.oro_pack_grouped_f32.argtypes = [u32p, C.c_uint32, u32p, C.c_uint32, f32p, C.c_uint32,
                                       C.c_float, C.c_void_p, C.c_uint32, f32p, u32p]
    L.oro_psa_alias_table_f32.argtypes = [f32p, C.c_uint32, C.c_uint32, C.c_int, C.c_int, f32p,
                                          u32p, C.c_void_p]
    L.oro_psa_alias_table_f64.argtypes = [f32p, C.c_uint32, C.c_uint32, C.c_int, C.c_int,
                                          C.c_double, f64p, u32p, C.c_void_p]
    L.oro_psa_stream_table.argtypes = [f32p, C.c_uint32, C.c_double, C.c_int, f32p, u32p]
    L.oro_sample_alias.argtypes = [f32p, u32p, C.c_uint64, C.c_uint64, C.c_uint64, C.c_uint64,
                                   u32p]
    L.oro_its_sample.argtypes = [f32p, C.c_void_p, C.c_uint64, C.c_uint64, C.c_uint64,
                                 C.c_uint64, u32p]
    L.oro_implied_distribution_f32.argtypes = [f32p, u32p, C.c_uint64, f64p]
    L.oro_implied_distribution_f64.argtypes = [f64p, u32p, C.c_uint64, f64p]
    _lib = L
    return L


# ---------------------------------------------------------------- Philox ----
def philox4x32_10(ctr, key) -> np.ndarray:
    out = np.zeros(4, np.uint32)
    lib().oro_philox4x32_10(np.asarray(ctr, np.uint32), np.asarray(key, np.uint32), out)
    return out


def philox_block(seed: int, n: int, base: int = 0, stream: int = 0) -> np.ndarray:
    out = np.zeros((n, 4), np.uint32)
    lib().oro_philox_block(seed, stream, base, n, out.reshape(-1))
    return out


# ------------------------------------------------------------ generators ----
def gen_uniform01(seed: int, n: int) -> np.ndarray:
    out = np.empty(n, np.float32)
    lib().oro_gen_uniform01(seed, n, out)
    return out


def gen_lognormal(seed: int, mu: float, sigma: float, n: int) -> np.ndarray:
    out = np.empty(n, np.float32)
    lib().oro_gen_lognormal(seed, mu, sigma, n, out)
    return out


def feistel_perm(seed: int, n: int, i: int) -> int:
    return int(lib().oro_feistel_perm(seed, n, i))


def gen_zipf(seed: int, s: float, n: int) -> np.ndarray:
    out = np.empty(n, np.float32)
    lib().oro_gen_zipf(seed, s, n, out)
    return out


def gen_spike(seed: int, n: int) -> np.ndarray:
    out = np.empty(n, np.float32)
    lib().oro_gen_spike(seed, n, out)
    return out


def segment_sizes(seed: int, b: int, lo: int = 256, hi: int = 1 << 18) -> np.ndarray:
    out = np.empty(b, np.uint64)
    lib().oro_segment_sizes(seed, b, lo, hi, out)
    return out


# ---------------------------------------------------------- PSA pieces ----
def kahan_reduce_f32(x) -> float:
    x = np.ascontiguousarray(x, np.float32)
    return float(lib().oro_kahan_reduce_f32(x, x.size))


def kahan_reduce_f64(x) -> float:
    x = np.ascontiguousarray(x, np.float64)
    return float(lib().oro_kahan_reduce_f64(x, x.size))


def stable_partition_indices(w, pivot, f64: bool = True):
    """Returns (idx_cpu_layout, H)."""
    n = len(w)
    idx = np.empty(n, np.uint32)
    if f64:
        H = lib().oro_stable_partition_indices_f64(np.ascontiguousarray(w, np.float64), n,
                                                    float(pivot), idx)
    else:
        H = lib().oro_stable_partition_indices_f32(np.ascontiguousarray(w, np.float32), n,
                                                    np.float32(pivot), idx)
    return idx, int(H)


def cpu_to_gpu_layout(idx_cpu: np.ndarray, H: int) -> np.ndarray:
    """F11 layout: heavies ascending at [0,H), light rank r at N-1-r."""
    out = idx_cpu.copy()
    out[H:] = idx_cpu[H:][::-1]
    return out


def prefix_sum_f64(x) -> np.ndarray:
    x = np.ascontiguousarray(x, np.float64)
    out = np.empty_like(x)
    lib().oro_prefix_sum_f64(x, x.size, out)
    return out


def prefix_sum_f32(x) -> np.ndarray:
    x = np.ascontiguousarray(x, np.float32)
    out = np.empty_like(x)
    lib().oro_prefix_sum_f32(x, x.size, out)
    return out


def ceil_mul_div(a: int, b: int, c: int) -> int:
    return int(lib().oro_ceil_mul_div(a, b, c))


def split_f64(hp, lp, mean: float, n: int):
    hp = np.ascontiguousarray(hp, np.float64)
    lp = np.ascontiguousarray(lp, np.float64)
    s = lib().oro_split_f64(hp if hp.size else np.zeros(1), hp.size,
                            lp if lp.size else np.zeros(1), lp.size, mean, n)
    return int(s.i), int(s.j), float(s.spill)


def split_f32(hp, lp, mean: float, n: int):
    hp = np.ascontiguousarray(hp, np.float32)
    lp = np.ascontiguousarray(lp, np.float32)
    s = lib().oro_split_f32(hp if hp.size else np.zeros(1, np.float32), hp.size,
                            lp if lp.size else np.zeros(1, np.float32), lp.size,
                            np.float32(mean), n)
    return int(s.i), int(s.j), float(s.spill)


def split_lb_f64(hp, lp, mean: float, n: int):
    hp = np.ascontiguousarray(hp, np.float64)
    lp = np.ascontiguousarray(lp, np.float64)
    s = lib().oro_split_lb_f64(hp if hp.size else np.zeros(1), hp.size,
                               lp if lp.size else np.zeros(1), lp.size, mean, n)
    return int(s.i), int(s.j), float(s.spill)


def splitK_f64(hp, lp, mean: float, N: int, K: int, rule: int = SPLIT_LOWER_BOUND) -> np.ndarray:
    hp = np.ascontiguousarray(hp, np.float64)
    lp = np.ascontiguousarray(lp, np.float64)
    out = np.zeros(K, SPLIT64_DTYPE)
    lib().oro_splitK_f64(hp, hp.size, lp, lp.size, mean, N, K, rule, out.ctypes.data)
    return out


def splitK_f32(hp, lp, mean: float, N: int, K: int, rule: int = SPLIT_REFERENCE_CPU) -> np.ndarray:
    hp = np.ascontiguousarray(hp, np.float32)
    lp = np.ascontiguousarray(lp, np.float32)
    out = np.zeros(K, SPLIT32_DTYPE)
    lib().oro_splitK_f32(hp, hp.size, lp, lp.size, np.float32(mean), N, K, rule, out.ctypes.data)
    return out


def sweep_states(w, mean: float):
    """(i, j) after n buckets of the split-free sequential sweep, n = 0..N-1."""
    idx, H = stable_partition_indices(w, mean)
    n = len(w)
    si = np.zeros(n, np.uint32)
    sj = np.zeros(n, np.uint32)
    lib().oro_sweep_states_f64(idx[:H].copy(), H, idx[H:].copy(), n - H,
                               np.ascontiguousarray(w, np.float64), float(mean), si, sj)
    return si, sj, H


def psa_alias_table_f32(w, K: int, reference_scale: bool = False,
                        split_rule: int = SPLIT_REFERENCE_CPU, pack: int = PACK_SCALAR):
    """Reference CPU path with T = P = float.  Returns (p float32, a uint32, debug dict).
    `pack` = PACK_GROUPED re-packs the same split records with SubgroupPack's sweep."""
    w = np.ascontiguousarray(w, np.float32)
    n = w.size
    p = np.empty(n, np.float32)
    a = np.empty(n, np.uint32)
    idx = np.empty(n, np.uint32)
    hp = np.zeros(n, np.float64)
    lp = np.zeros(n, np.float64)
    sp = np.zeros(K, SPLIT64_DTYPE)
    dbg = PsaDebug(0.0, 0, idx.ctypes.data, hp.ctypes.data, lp.ctypes.data, sp.ctypes.data)
    lib().oro_psa_alias_table_f32(w, n, K, int(reference_scale), int(split_rule), p, a,
                                  C.byref(dbg))
    H = int(dbg.heavy_count)
    if pack == PACK_GROUPED and 0 < H < n:
        # the weights the reference path packs (`weights[i] *= 0.0001` with
        # T = float is evaluated in double and rounded back to float)
        ws = (w.astype(np.float64) * 0.0001).astype(np.float32) if reference_scale else w
        sp32 = np.zeros(K, SPLIT32_DTYPE)
        sp32["i"], sp32["j"], sp32["spill"] = sp["i"], sp["j"], sp["spill"].astype(np.float32)
        lib().oro_pack_grouped_f32(idx[:H].copy(), H, idx[H:].copy(), n - H, ws, n,
                                   np.float32(dbg.mean), sp32.ctypes.data, K, p, a)
    return p, a, dict(mean=dbg.mean, H=H, idx_cpu=idx, heavy_prefix=hp[:H],
                      light_prefix=lp[:n - H], splits=sp)


def psa_alias_table_f64(w, K: int, reference_scale: bool = False, mean=None,
                        split_rule: int = SPLIT_LOWER_BOUND, pack: int = PACK_SCALAR):
    """Reference CPU path with T = P = double (tests-of-tests instantiation).
    `mean` (optional) pins the pivot/bucket size ("given the pivot");
    `split_rule` selects the reference CPU split or the lower-bound split of
    the reference GPU kernels (the default: it always matches the sweep);
    `pack` = PACK_GROUPED re-packs the same split records with SubgroupPack's
    sweep."""
    w = np.ascontiguousarray(w, np.float32)
    n = w.size
    p = np.empty(n, np.float64)
    a = np.empty(n, np.uint32)
    idx = np.empty(n, np.uint32)
    hp = np.zeros(n, np.float64)
    lp = np.zeros(n, np.float64)
    sp = np.zeros(K, SPLIT64_DTYPE)
    dbg = PsaDebug(0.0, 0, idx.ctypes.data, hp.ctypes.data, lp.ctypes.data, sp.ctypes.data)
    m = float("nan") if mean is None else float(mean)
    lib().oro_psa_alias_table_f64(w, n, K, int(reference_scale), int(split_rule), m, p, a,
                                  C.byref(dbg))
    H = int(dbg.heavy_count)
    if pack == PACK_GROUPED and 0 < H < n:
        w64 = w.astype(np.float64) * 0.0001 if reference_scale else w.astype(np.float64)
        lib().oro_pack_grouped_f64(idx[:H].copy(), H, idx[H:].copy(), n - H, w64, n,
                                   float(dbg.mean), sp.ctypes.data, K, p, a)
    return p, a, dict(mean=dbg.mean, H=H, idx_cpu=idx, heavy_prefix=hp[:H],
                      light_prefix=lp[:n - H], splits=sp)


def psa_stream_table(w, mean: float, f32: bool = False):
    """The reference sweep (pack.hpp:14-83) with exact integer keys, as the
    streamed pack form computes it: (p float32, a uint32) for the given mean."""
    w = np.ascontiguousarray(w, np.float32)
    n = w.size
    p = np.empty(n, np.float32)
    a = np.empty(n, np.uint32)
    lib().oro_psa_stream_table(w, n, float(mean), int(bool(f32)), p, a)
    return p, a


# ------------------------------------------------------------ sampling ----
def sample_alias(p, a, s: int, seed: int, counter_base: int = 0) -> np.ndarray:
    p = np.ascontiguousarray(p, np.float32)
    a = np.ascontiguousarray(a, np.uint32)
    out = np.empty(s, np.uint32)
    lib().oro_sample_alias(p, a, p.size, s, seed, counter_base, out)
    return out


def its_sample(cdf_hi, cdf_lo, s: int, seed: int, counter_base: int = 0) -> np.ndarray:
    cdf_hi = np.ascontiguousarray(cdf_hi, np.float32)
    lo_ptr = None
    if cdf_lo is not None:
        cdf_lo = np.ascontiguousarray(cdf_lo, np.float32)
        lo_ptr = cdf_lo.ctypes.data
    out = np.empty(s, np.uint32)
    lib().oro_its_sample(cdf_hi, lo_ptr, cdf_hi.size, s, seed, counter_base, out)
    return out


def implied_distribution(p, a) -> np.ndarray:
    a = np.ascontiguousarray(a, np.uint32)
    q = np.empty(a.size, np.float64)
    if np.asarray(p).dtype == np.float64:
        lib().oro_implied_distribution_f64(np.ascontiguousarray(p, np.float64), a, a.size, q)
    else:
        lib().oro_implied_distribution_f32(np.ascontiguousarray(p, np.float32), a, a.size, q)
    return q


# ------------------------------------------------- FP64 truth and stats ----
def exact_total(w) -> float:
    """Correctly rounded sum of the (float) weights (math.fsum)."""
    return math.fsum(np.asarray(w, np.float64).tolist())


def target_distribution(w) -> np.ndarray:
    w64 = np.asarray(w, np.float64)
    return w64 / exact_total(w64)


def distribution_error(q: np.ndarray, w) -> dict:
    """Per-element error of an implied distribution q against w/sum(w) in FP64."""
    t = target_distribution(w)
    d = np.abs(q - t)
    nz = t > 0
    rel = np.zeros_like(d)
    rel[nz] = d[nz] / t[nz]
    return dict(max_abs=float(d.max(initial=0.0)), max_rel=float(rel.max(initial=0.0)),
                zero_weight_mass=float(q[~nz].sum()) if (~nz).any() else 0.0)
