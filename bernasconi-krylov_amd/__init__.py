"""MI355X-native Trotter engine for H = E_LABS(s)/N - g sum_i sigma^x_i.

Python mirror of the reference's sampler and subspace interfaces over the C
ABI of ``libbk.so`` (include/bk_capi.h):

* ``SamplerConfig`` / ``SamplingResult`` / ``ISampler`` / ``create_sampler``
  mirror sampler.cuh:235-310, 737-743 (same field names, defaults and
  meaning); ``create_sampler(n).sample(start_state, cfg)`` runs the whole
  ISampler::sample (sampler.cuh:412-707) on the GPU.
* ``Engine`` exposes the lower-level entry points: Trotter evolution with
  fused P_< / <d_H> reductions, projective sampling, the matrix-free H.v
  (diagonalizer.cuh:60-77), GPU Lanczos (diagonalizer.cuh:420-554) and the
  Krylov propagator.
* ``energy`` / ``d_min`` / ``known_optimal_energy`` / ``make_trap`` mirror
  labs_energy.hpp:121-207 and classical.cuh:602-713.

There is no CPU fallback: every call goes through the HIP library and fails
loudly when it is missing or no GPU is present.
"""
from __future__ import annotations

import ctypes as C
import dataclasses
import os
import subprocess
from typing import List, Optional, Sequence

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libbk.so")

BK_OK = 0
BK_C64 = 1
BK_C128 = 2
BK_LIE = 1
BK_STRANG = 2
BK_MAX_N = 32
BK_ENERGY_ONTHEFLY = 0
BK_ENERGY_CACHE16 = 1


class BKError(RuntimeError):
    """Error code returned by libbk (message from bk_last_error())."""

    def __init__(self, code: int, msg: str):
        super().__init__(f"libbk error {code}: {msg}")
        self.code = code


# ---------------------------------------------------------------- ctypes mirror
class EvolveParams(C.Structure):
    _fields_ = [("n", C.c_int), ("precision", C.c_int), ("order", C.c_int), ("dt", C.c_double),
                ("steps", C.c_int), ("g", C.c_double), ("g_steps", C.POINTER(C.c_double)),
                ("start_bits", C.c_uint64), ("e0", C.c_int64), ("init_from_start", C.c_int),
                ("measure_every", C.c_int)]


class AdjointParams(C.Structure):
    _fields_ = [("ev", EvolveParams), ("checkpoint_every", C.c_int)]


class Instance(C.Structure):
    _fields_ = [("n", C.c_int), ("start_bits", C.c_uint64), ("e0", C.c_int64), ("g", C.c_double)]


class BatchParams(C.Structure):
    _fields_ = [("precision", C.c_int), ("order", C.c_int), ("dt", C.c_double), ("steps", C.c_int),
                ("use_graph", C.c_int), ("autotune", C.c_int), ("block_bytes", C.c_int64),
                ("out_of_place", C.c_int)]


class BatchResult(C.Structure):
    _fields_ = [("p_less", C.c_double), ("dh", C.c_double), ("norm", C.c_double)]


class BatchStats(C.Structure):
    _fields_ = [("device_ms", C.c_double), ("grid", C.c_int), ("row_bits", C.c_int),
                ("block_bytes", C.c_int64), ("launches", C.c_int64), ("tuned", C.c_int),
                ("out_of_place", C.c_int)]


class Series(C.Structure):
    _fields_ = [("p_less", C.POINTER(C.c_double)), ("dh", C.POINTER(C.c_double)),
                ("norm", C.POINTER(C.c_double)), ("capacity", C.c_int64), ("count", C.c_int64),
                ("e0", C.c_int64), ("p_blind", C.c_double)]


class _SamplerConfigC(C.Structure):
    _fields_ = [("n", C.c_int), ("g", C.c_double), ("dt", C.c_double), ("t_max", C.c_double),
                ("num_sample_points", C.c_int), ("shots_per_point", C.c_int),
                ("seed", C.c_uint64), ("num_sampling_seeds", C.c_int),
                ("seed_stride", C.c_uint64), ("t_start", C.c_double),
                ("dt_sample", C.c_double), ("track_p_less", C.c_int), ("verbose", C.c_int),
                ("record_trajectory", C.c_int), ("dt_trajectory", C.c_double),
                ("precision", C.c_int), ("order", C.c_int)]


class _SeedStatsC(C.Structure):
    _fields_ = [("seed", C.c_uint64), ("total_shots_collected", C.c_uint64),
                ("improved_shots", C.c_uint64), ("success_rate", C.c_double),
                ("per_eval_rate", C.c_double), ("mean_hamming_flips", C.c_double),
                ("best_energy", C.c_int64)]


class _SamplingResultC(C.Structure):
    _fields_ = [("unique_basis", C.POINTER(C.c_uint64)), ("unique_basis_len", C.c_int64),
                ("total_shots_collected", C.c_uint64), ("improved_shots", C.c_uint64),
                ("success_rate", C.c_double), ("per_eval_rate", C.c_double),
                ("total_evals", C.c_uint64), ("mean_hamming_flips", C.c_double),
                ("seed_stats", C.POINTER(_SeedStatsC)), ("seed_stats_len", C.c_int64),
                ("time_points", C.POINTER(C.c_double)),
                ("p_less_history", C.POINTER(C.c_double)),
                ("dh_history", C.POINTER(C.c_double)), ("history_len", C.c_int64),
                ("window_avg_dh", C.c_double), ("initial_rise_rate_avg", C.c_double),
                ("initial_rise_rate_max", C.c_double), ("d_min", C.c_int),
                ("start_energy", C.c_int64), ("p_blind", C.c_double),
                ("peak_p_less", C.c_double), ("peak_time", C.c_double),
                ("samples", C.POINTER(C.c_int64)), ("samples_len", C.c_int64)]


class _MinimumFamilyC(C.Structure):
    _fields_ = [("id", C.c_int64), ("canonical_state", C.c_uint64), ("sample_state", C.c_uint64),
                ("energy", C.c_int64), ("times_visited", C.c_uint64), ("evals_to_first_find", C.c_uint64)]


BK_MINIMA_MAX_LEVELS = 8


class _MinimaOrbitC(C.Structure):
    _fields_ = [("canonical_state", C.c_uint64), ("energy", C.c_int64), ("level", C.c_int), ("d_min", C.c_int),
                ("sd_hit_fraction", C.c_double)]


class _ExhaustiveResultC(C.Structure):
    _fields_ = [("n", C.c_int), ("e_star", C.c_int64), ("total_minima_found", C.c_int64),
                ("total_orbits_found", C.c_int64), ("num_energies", C.c_int64), ("num_levels", C.c_int),
                ("level_energy", C.c_int64 * BK_MINIMA_MAX_LEVELS), ("level_orbits", C.c_int64 * BK_MINIMA_MAX_LEVELS),
                ("orbits_written", C.c_int64), ("search_time_sec", C.c_double)]


class _TrapRecordC(C.Structure):
    _fields_ = [("n", C.c_int), ("g", C.c_double), ("trap_rank", C.c_int64), ("trap_bits", C.c_uint64),
                ("canonical_bits", C.c_uint64), ("trap_energy", C.c_int64), ("d_min", C.c_int),
                ("quantum_time_sec", C.c_double), ("window_t_start", C.c_double), ("window_points", C.c_int),
                ("window_dt", C.c_double), ("t_max", C.c_double), ("result", C.POINTER(_SamplingResultC))]


class KernelTiming(C.Structure):
    _fields_ = [("name", C.c_char * 48), ("launches", C.c_int64), ("total_ms", C.c_double),
                ("bytes_per_launch", C.c_double)]


# names and argtypes of every exported entry point (checked by the CPU tests)
_SIGS = {
    "bk_last_error": (C.c_char_p, []),
    "bk_version": (C.c_int, []),
    "bk_device_count": (C.c_int, []),
    "bk_set_device_memory_limit": (C.c_int, [C.c_uint64]),
    "bk_device_memory_held": (C.c_uint64, []),
    "bk_ctx_create": (C.c_int, [C.c_int, C.POINTER(C.c_void_p)]),
    "bk_ctx_destroy": (None, [C.c_void_p]),
    "bk_energy": (C.c_int64, [C.c_int, C.c_uint64]),
    "bk_d_min": (C.c_int, [C.c_int, C.c_uint64, C.c_int64, C.c_int]),
    "bk_known_optimal_energy": (C.c_int64, [C.c_int]),
    "bk_make_trap": (C.c_int, [C.c_int, C.c_uint64, C.POINTER(C.c_uint64),
                               C.POINTER(C.c_uint64), C.POINTER(C.c_int64)]),
    "bk_state_init": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_uint64]),
    "bk_state_upload": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_void_p]),
    "bk_state_download": (C.c_int, [C.c_void_p, C.c_void_p]),
    "bk_state_device": (C.c_int, [C.c_void_p, C.POINTER(C.c_void_p), C.POINTER(C.c_void_p),
                                  C.POINTER(C.c_int), C.POINTER(C.c_int)]),
    "bk_evolve": (C.c_int, [C.c_void_p, C.POINTER(EvolveParams), C.POINTER(Series)]),
    "bk_adjoint_forward": (C.c_int, [C.c_void_p, C.POINTER(AdjointParams), C.POINTER(C.c_double)]),
    "bk_adjoint_backward": (C.c_int, [C.c_void_p, C.c_double, C.c_double, C.POINTER(C.c_double),
                                      C.POINTER(C.c_double)]),
    "bk_adjoint_grad": (C.c_int, [C.c_void_p, C.POINTER(AdjointParams), C.c_double, C.c_double,
                                  C.POINTER(C.c_double), C.POINTER(C.c_double),
                                  C.POINTER(C.c_double)]),
    "bk_batch_run": (C.c_int, [C.c_void_p, C.POINTER(Instance), C.c_int, C.POINTER(BatchParams),
                               C.POINTER(BatchResult), C.POINTER(BatchStats)]),
    "bk_observables": (C.c_int, [C.c_void_p, C.c_uint64, C.c_int64, C.POINTER(C.c_double)]),
    "bk_state_overlap": (C.c_int, [C.c_void_p, C.c_void_p, C.POINTER(C.c_double)]),
    "bk_count_below": (C.c_int, [C.c_void_p, C.c_int, C.c_int64, C.POINTER(C.c_uint64)]),
    "bk_sample": (C.c_int, [C.c_void_p, C.POINTER(C.c_double), C.c_int64,
                            C.POINTER(C.c_int64)]),
    "bk_sample_noalloc": (C.c_int, [C.c_int, C.c_int, C.c_void_p, C.c_void_p, C.c_int64, C.c_void_p,
                                    C.c_void_p]),
    "bk_hv": (C.c_int, [C.c_void_p, C.c_int, C.c_double, C.c_int, C.c_int, C.c_void_p,
                        C.c_void_p, C.c_void_p]),
    "bk_lanczos": (C.c_int, [C.c_void_p, C.c_int, C.c_double, C.c_int, C.c_int, C.c_double,
                             C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_double)]),
    "bk_krylov_propagate": (C.c_int, [C.c_void_p, C.c_double, C.c_int, C.c_double, C.c_double,
                                      C.c_int]),
    "bk_sampler_config_defaults": (None, [C.POINTER(_SamplerConfigC)]),
    "bk_sampler_run": (C.c_int, [C.c_void_p, C.POINTER(_SamplerConfigC), C.c_uint64,
                                 C.POINTER(C.POINTER(_SamplingResultC))]),
    "bk_result_free": (None, [C.POINTER(_SamplingResultC)]),
    "bk_set_energy_mode": (C.c_int, [C.c_void_p, C.c_int]),
    "bk_set_tuning": (C.c_int, [C.c_void_p, C.c_char_p, C.c_int]),
    "bk_sweep_plan": (C.c_int, [C.c_int, C.c_int, C.c_int, C.c_int, C.POINTER(C.c_int), C.c_int]),
    "bk_timing_enable": (C.c_int, [C.c_void_p, C.c_int]),
    "bk_copy_probe": (C.c_int, [C.c_void_p, C.c_uint64, C.c_int, C.POINTER(C.c_double)]),
    "bk_copy_probe_pattern": (C.c_int, [C.c_void_p, C.c_uint64, C.c_int, C.c_int, C.POINTER(C.c_double)]),
    "bk_subspace_lanczos": (C.c_int, [C.c_void_p, C.c_int, C.c_double, C.c_int, C.POINTER(C.c_uint64), C.c_int64,
                                      C.c_int, C.c_double, C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_double)]),
    "bk_subspace_dense": (C.c_int, [C.c_void_p, C.c_int, C.c_double, C.c_int, C.POINTER(C.c_uint64), C.c_int64,
                                    C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_double)]),
    "bk_fullspace_dense": (C.c_int, [C.c_void_p, C.c_int, C.c_double, C.c_int, C.c_int, C.POINTER(C.c_double),
                                     C.POINTER(C.c_double)]),
    "bk_canonical": (C.c_uint64, [C.c_int, C.c_uint64]),
    "bk_descent_batch": (C.c_int, [C.c_void_p, C.c_int, C.POINTER(C.c_uint64), C.c_int64, C.c_uint64,
                                   C.POINTER(C.c_uint64), C.POINTER(C.c_int64), C.POINTER(C.c_int64)]),
    "bk_collect_minima": (C.c_int, [C.c_void_p, C.c_int, C.c_uint64, C.c_int64, C.c_int64, C.c_int64,
                                    C.POINTER(_MinimumFamilyC), C.c_int64, C.POINTER(C.c_int64),
                                    C.POINTER(C.c_int64), C.POINTER(C.c_int64)]),
    "bk_d_min_batch": (C.c_int, [C.c_void_p, C.c_int, C.POINTER(C.c_uint64), C.POINTER(C.c_int64), C.c_int64,
                                 C.c_int, C.POINTER(C.c_int)]),
    "bk_exhaustive_minima": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_uint64, C.c_uint64,
                                       C.POINTER(_ExhaustiveResultC), C.POINTER(_MinimaOrbitC), C.c_int64]),
    "bk_export_exhaustive_json": (C.c_int, [C.c_char_p, C.POINTER(_ExhaustiveResultC), C.POINTER(_MinimaOrbitC),
                                            C.c_int64]),
    "bk_export_results_csv": (C.c_int, [C.c_char_p, C.POINTER(_TrapRecordC), C.c_int]),
    "bk_export_time_series_csv": (C.c_int, [C.c_char_p, C.POINTER(_TrapRecordC), C.c_int]),
    "bk_export_structured_json": (C.c_int, [C.c_char_p, C.POINTER(_TrapRecordC), C.c_int, C.c_int]),
    "bk_timing_read": (C.c_int, [C.c_void_p, C.POINTER(KernelTiming), C.c_int,
                                 C.POINTER(C.c_int)]),
}

_lib = None


def build() -> None:
    """Compile libbk.so in-tree (hipcc, gfx950)."""
    subprocess.run(["make", "-s", "-C", HERE, "-j8"], check=True)


def lib() -> C.CDLL:
    """The loaded libbk.so; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is missing: build it with "
                              f"`make -C {HERE}` (there is no CPU fallback)")
        # PyTorch-ROCm ships its own HIP runtime with the same SONAME
        # (libamdhip64.so.7).  Loading torch first makes libbk bind to that
        # copy, so one process holds one HIP runtime and torch tensors and
        # libbk pointers share a device context.
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def exported_symbols() -> List[str]:
    return list(_SIGS)


def _check(rc: int) -> int:
    if rc < 0:
        raise BKError(rc, lib().bk_last_error().decode())
    return rc


def _dp(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_double))


# ---------------------------------------------------------------- L0 core
def energy(n: int, s: int) -> int:
    """LabsEvaluator::calcEnergy (labs_energy.hpp:121-132)."""
    return lib().bk_energy(n, s)


def d_min(n: int, s0: int, target: int = -1, max_d: int = -1) -> int:
    """LabsEvaluator::compute_d_min (labs_energy.hpp:175-207)."""
    return lib().bk_d_min(n, s0, target, max_d)


def known_optimal_energy(n: int) -> int:
    return lib().bk_known_optimal_energy(n)


def canonical(n: int, s: int) -> int:
    """SpinState::canonical (labs_energy.hpp:51-78)."""
    return lib().bk_canonical(n, s)


def make_trap(n: int, seed: int):
    """seed -> random_state -> steepest_descent (classical.cuh:602-713): (start, s0, E0)."""
    a, b, e = C.c_uint64(), C.c_uint64(), C.c_int64()
    _check(lib().bk_make_trap(n, seed, C.byref(a), C.byref(b), C.byref(e)))
    return a.value, b.value, e.value


class MT19937_64:
    """std::mt19937_64 (its output sequence is fixed by the C++ standard), for
    seed-drawn synthetic inputs on the host."""

    _M = (1 << 64) - 1

    def __init__(self, seed: int = 5489):
        self.mt = [0] * 312
        self.mt[0] = seed & self._M
        for i in range(1, 312):
            self.mt[i] = (6364136223846793005 * (self.mt[i - 1] ^ (self.mt[i - 1] >> 62)) + i) & self._M
        self.idx = 312

    def __call__(self) -> int:
        if self.idx >= 312:
            mt = self.mt
            for i in range(312):
                y = (mt[i] & 0xFFFFFFFF80000000) | (mt[(i + 1) % 312] & 0x7FFFFFFF)
                v = mt[(i + 156) % 312] ^ (y >> 1)
                if y & 1:
                    v ^= 0xB5026F5AA96619E9
                mt[i] = v
            self.idx = 0
        x = self.mt[self.idx]
        self.idx += 1
        x ^= (x >> 29) & 0x5555555555555555
        x ^= (x << 17) & 0x71D67FFFEDA60000
        x ^= (x << 37) & 0xFFF7EEE000000000
        x ^= x >> 43
        return x & self._M


def field_schedule(seed: int, n_segments: int, center: float = 0.35, spread: float = 0.05) -> np.ndarray:
    """Piecewise-constant field g_j = center + spread * (2u - 1), u = (x >> 11) * 2^-53
    with x drawn from mt19937_64(seed): a portable uniform (unlike
    std::uniform_real_distribution, whose algorithm the standard leaves open)."""
    rng = MT19937_64(seed)
    u = np.array([(rng() >> 11) * 2.0 ** -53 for _ in range(n_segments)])
    return center + spread * (2.0 * u - 1.0)


def batch_workload(count: int = 512, seed: int = 1, trap_base: int = 1000, sizes=(16, 18, 20, 22),
                   g_lo: float = 0.10, g_hi: float = 0.60, n_g: int = 16):
    """Seed-drawn batch of traps: instance i has n_i = sizes[x >> 62] for successive
    draws x of mt19937_64(seed) (a portable uniform choice among four sizes),
    g_i = g_lo + (g_hi - g_lo) * (i mod n_g) / (n_g - 1), and its trap from seed
    trap_base + i.  Returns a list of (n, s0, E0, g)."""
    if len(sizes) != 4:
        raise ValueError("four sizes")
    rng = MT19937_64(seed)
    out = []
    for i in range(count):
        n = sizes[rng() >> 62]
        _, s0, e0 = make_trap(n, trap_base + i)
        out.append((n, s0, e0, g_lo + (g_hi - g_lo) * (i % n_g) / (n_g - 1)))
    return out


def sweep_plan(n: int, precision: int = BK_C64, row_bits: int = 5, page_bits: int = 0):
    """HIGH sweeps of one Trotter step: list of (h1, a1, h2, a2, R)."""
    buf = (C.c_int * 40)()
    k = _check(lib().bk_sweep_plan(n, precision, row_bits, page_bits, buf, 8))
    return [tuple(buf[5 * i:5 * i + 5]) for i in range(k)]


def make_trap_min_dmin(n: int, seed: int, min_dmin: int, max_tries: int = 100000):
    """Iterate seed, seed+1, ... until the trap's exact d_min >= min_dmin (and
    E0 > E*(N)).  Returns (seed, start, s0, E0, d_min)."""
    opt = known_optimal_energy(n)
    for k in range(max_tries):
        sd = seed + k
        start, s0, e0 = make_trap(n, sd)
        if opt >= 0 and e0 <= opt:
            continue
        dm = d_min(n, s0, e0)
        if dm >= min_dmin:
            return sd, start, s0, e0, dm
    raise RuntimeError("no trap found")


# ---------------------------------------------------------------- engine
class Engine:
    """One libbk context: a device, a HIP stream and one statevector."""

    def __init__(self, device: int = 0):
        h = C.c_void_p()
        _check(lib().bk_ctx_create(device, C.byref(h)))
        self._h = h
        self.device = device

    def close(self):
        if getattr(self, "_h", None):
            lib().bk_ctx_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- state
    def state_init(self, n: int, precision: int = BK_C128, start_bits: int = 0):
        _check(lib().bk_state_init(self._h, n, precision, start_bits))

    def upload(self, psi: np.ndarray, precision: Optional[int] = None):
        psi = np.asarray(psi)
        if precision is None:
            precision = BK_C64 if psi.dtype == np.complex64 else BK_C128
        dt = np.complex64 if precision == BK_C64 else np.complex128
        a = np.ascontiguousarray(psi, dtype=dt)
        n = int(a.shape[0]).bit_length() - 1
        if (1 << n) != a.shape[0]:
            raise ValueError("state length must be a power of two")
        _check(lib().bk_state_upload(self._h, n, precision, a.ctypes.data_as(C.c_void_p)))

    def device_state(self):
        p, s, n, prec = C.c_void_p(), C.c_void_p(), C.c_int(), C.c_int()
        _check(lib().bk_state_device(self._h, C.byref(p), C.byref(s), C.byref(n), C.byref(prec)))
        return p.value, s.value, n.value, prec.value

    def download(self) -> np.ndarray:
        _, _, n, prec = self.device_state()
        out = np.empty(1 << n, dtype=np.complex64 if prec == BK_C64 else np.complex128)
        _check(lib().bk_state_download(self._h, out.ctypes.data_as(C.c_void_p)))
        return out

    # -- evolution
    def evolve(self, n: int, start_bits: int, g: float, dt: float, steps: int,
               order: int = BK_LIE, precision: int = BK_C128, e0: int = -1,
               g_steps: Optional[Sequence[float]] = None, init: bool = True,
               measure_every: int = 1) -> dict:
        """Trotter steps with P_< / <d_H> / norm recorded every `measure_every` steps."""
        gs = None
        if g_steps is not None:
            gs = np.ascontiguousarray(g_steps, dtype=np.float64)
            if gs.shape != (steps,):
                raise ValueError("g_steps must have length steps")
        p = EvolveParams(n=n, precision=precision, order=order, dt=dt, steps=steps, g=g,
                         g_steps=_dp(gs) if gs is not None else None, start_bits=start_bits,
                         e0=e0, init_from_start=1 if init else 0, measure_every=measure_every)
        cap = 1 + ((steps + measure_every - 1) // measure_every if measure_every > 0 else 0)
        pl, dh, nm = np.zeros(cap), np.zeros(cap), np.zeros(cap)
        s = Series(p_less=_dp(pl), dh=_dp(dh), norm=_dp(nm), capacity=cap, count=0, e0=0,
                   p_blind=0.0)
        _check(lib().bk_evolve(self._h, C.byref(p), C.byref(s)))
        k = s.count
        return {"p_less": pl[:k], "dh": dh[:k], "norm": nm[:k], "e0": s.e0,
                "p_blind": s.p_blind, "_keep": gs}

    # ---- batched evolution ----
    def batch_run(self, instances, dt: float, steps: int, order: int = BK_STRANG,
                  precision: int = BK_C64, use_graph: bool = True, autotune: bool = False,
                  block_bytes: int = 0, out_of_place: bool = False):
        """Evolve many traps at once.  instances: iterable of (n, start_bits, e0, g)
        (e0 < 0 -> E(start_bits)).  Without autotune, block_bytes > 0 runs the
        states in cache blocks of that size (each block through all steps), out of
        place with out_of_place.  Returns (array count x 3 of P_<, <d_H>, norm after
        the last step, stats dict)."""
        inst = list(instances)
        arr = (Instance * len(inst))(*[Instance(int(n), int(s), int(e), float(g)) for n, s, e, g in inst])
        res = (BatchResult * len(inst))()
        st = BatchStats()
        p = BatchParams(precision=precision, order=order, dt=dt, steps=steps,
                        use_graph=1 if use_graph else 0, autotune=1 if autotune else 0,
                        block_bytes=int(block_bytes), out_of_place=1 if out_of_place else 0)
        _check(lib().bk_batch_run(self._h, arr, len(inst), C.byref(p), res, C.byref(st)))
        out = np.array([[r.p_less, r.dh, r.norm] for r in res])
        stats = {"device_ms": st.device_ms, "grid": st.grid, "row_bits": st.row_bits,
                 "block_bytes": st.block_bytes, "launches": st.launches, "tuned": bool(st.tuned),
                 "out_of_place": bool(st.out_of_place)}
        return out, stats

    # ---- adjoint gradient ----
    def adjoint_forward(self, n: int, start_bits: int, g_steps: Sequence[float], dt: float,
                        order: int = BK_STRANG, precision: int = BK_C64, e0: int = -1,
                        checkpoint_every: int = 64) -> tuple:
        """Forward pass that records the checkpoint tape.  Returns (P_<, <d_H>, norm)
        of the final state."""
        gs = np.ascontiguousarray(g_steps, dtype=np.float64)
        steps = int(gs.shape[0])
        ev = EvolveParams(n=n, precision=precision, order=order, dt=dt, steps=steps, g=0.0,
                          g_steps=_dp(gs), start_bits=start_bits, e0=e0, init_from_start=1,
                          measure_every=0)
        p = AdjointParams(ev=ev, checkpoint_every=checkpoint_every)
        obs = np.zeros(3)
        _check(lib().bk_adjoint_forward(self._h, C.byref(p), _dp(obs)))
        self._adj_steps = steps
        return float(obs[0]), float(obs[1]), float(obs[2])

    def adjoint_backward(self, w_pless: float = 1.0, w_dh: float = 0.0):
        """dL/dg_k per step for L = w_pless P_<(T) + w_dh <d_H>(T) over the tape of the
        last adjoint_forward.  Returns (grad, drift)."""
        grad = np.zeros(getattr(self, "_adj_steps", 0))
        drift = C.c_double()
        _check(lib().bk_adjoint_backward(self._h, float(w_pless), float(w_dh), _dp(grad),
                                         C.byref(drift)))
        return grad, drift.value

    def observables(self, s0: int, e0: int):
        out = np.zeros(3)
        _check(lib().bk_observables(self._h, s0, e0, _dp(out)))
        return float(out[0]), float(out[1]), float(out[2])

    def overlap(self, other: "Engine") -> complex:
        """<psi_self|psi_other> of the two held states (same n, same device)."""
        out = np.zeros(2)
        _check(lib().bk_state_overlap(self._h, other._h, _dp(out)))
        return complex(out[0], out[1])

    def count_below(self, n: int, e0: int) -> int:
        c = C.c_uint64()
        _check(lib().bk_count_below(self._h, n, e0, C.byref(c)))
        return c.value

    def sample(self, randnums: Sequence[float]) -> np.ndarray:
        r = np.ascontiguousarray(randnums, dtype=np.float64)
        out = np.empty(r.shape[0], dtype=np.int64)
        _check(lib().bk_sample(self._h, _dp(r), r.shape[0],
                               out.ctypes.data_as(C.POINTER(C.c_int64))))
        return out

    def set_energy_mode(self, mode: int):
        """BK_ENERGY_CACHE16 (exact E(s) cache, 12-bit codes, 1.5 B per state) or
        BK_ENERGY_ONTHEFLY (popcounts)."""
        _check(lib().bk_set_energy_mode(self._h, mode))

    def set_tuning(self, key: str, value: int):
        """Sweep knobs: 'grid', 'row_bits', 'page_bits', 'pingpong', 'prefetch_low',
        'prefetch_high', 'fused', 'wide_low', 'narrow_phase', 'lds_tab13', 'nt_plain', 'grid_bphase', 'grid_lphase', 'rev_grid_low',
        'rev_grid_bphase', 'fold_finish'
        (see bk_set_tuning)."""
        _check(lib().bk_set_tuning(self._h, key.encode(), int(value)))

    # -- instrumentation
    def timing_enable(self, on: bool = True):
        _check(lib().bk_timing_enable(self._h, 1 if on else 0))

    def timing_read(self) -> List[dict]:
        buf = (KernelTiming * 64)()
        cnt = C.c_int()
        _check(lib().bk_timing_read(self._h, buf, 64, C.byref(cnt)))
        return [{"name": buf[i].name.decode(), "launches": buf[i].launches,
                 "total_ms": buf[i].total_ms, "bytes_per_launch": buf[i].bytes_per_launch}
                for i in range(cnt.value)]

    def copy_probe(self, nbytes: int = 1 << 33, reps: int = 5, pattern: int = 0) -> float:
        """GB/s (read+write) of a device copy of nbytes.  pattern 0: grid-stride
        copy; 1: 4 KB per workgroup, nontemporal (the part's fastest stream);
        2: 128 KB per workgroup (a complex64 sweep tile's footprint)."""
        g = C.c_double()
        if pattern == 0:
            _check(lib().bk_copy_probe(self._h, nbytes, reps, C.byref(g)))
        else:
            _check(lib().bk_copy_probe_pattern(self._h, nbytes, reps, pattern, C.byref(g)))
        return g.value

    # -- subspace
    def hv(self, n: int, g: float, x, y, normalize_by_n: bool = True, stream=None):
        """y = H x on device tensors (torch, real float64 or complex128)."""
        import torch  # device memory plumbing only
        if not (x.is_cuda and y.is_cuda):
            raise ValueError("hv takes device tensors")
        if x.dtype not in (torch.float64, torch.complex128) or y.dtype != x.dtype:
            raise ValueError("hv takes float64 or complex128 tensors")
        if x.numel() != (1 << n) or y.numel() != (1 << n) or not x.is_contiguous() \
                or not y.is_contiguous():
            raise ValueError("x, y must be contiguous with 2^n elements")
        torch.cuda.current_stream().synchronize()
        _check(lib().bk_hv(self._h, n, g, 1 if normalize_by_n else 0,
                           1 if x.dtype == torch.complex128 else 0, C.c_void_p(x.data_ptr()),
                           C.c_void_p(y.data_ptr()), stream))
        return y

    def lanczos(self, n: int, g: float, max_krylov_dim: int = 100, tol: float = 1e-12,
                k_pairs: int = 1, normalize_by_n: bool = True, vectors: bool = True):
        ev = np.zeros(k_pairs)
        vec = np.zeros((k_pairs, 1 << n)) if vectors else None
        k = _check(lib().bk_lanczos(self._h, n, g, 1 if normalize_by_n else 0, max_krylov_dim,
                                    tol, k_pairs, _dp(ev), _dp(vec) if vectors else None))
        return ev[:k], (vec[:k] if vectors else None)

    def descent_batch(self, n: int, starts, max_evals: int = 0):
        """Steepest descent of every start on the GPU (classical.cuh:459-535):
        (final states, final energies, evaluation counts)."""
        st = np.ascontiguousarray(np.asarray(starts, dtype=np.uint64))
        k = st.shape[0]
        fin = np.zeros(k, np.uint64)
        fe = np.zeros(k, np.int64)
        ev = np.zeros(k, np.int64)
        u64 = C.POINTER(C.c_uint64)
        i64 = C.POINTER(C.c_int64)
        _check(lib().bk_descent_batch(self._h, n, st.ctypes.data_as(u64), k, max_evals, fin.ctypes.data_as(u64),
                                      fe.ctypes.data_as(i64), ev.ctypes.data_as(i64)))
        return fin, fe, ev

    def collect_minima(self, n: int, seed: int, target_families: int, batch_size: int = 4096,
                       max_starts: int = 100000) -> dict:
        """collect_distinct_minima_gpu (classical.cuh:861-950): distinct symmetry
        families of local minima, deepest first."""
        out = (_MinimumFamilyC * target_families)()
        cnt, starts, evals = C.c_int64(), C.c_int64(), C.c_int64()
        _check(lib().bk_collect_minima(self._h, n, seed, target_families, batch_size, max_starts, out,
                                       target_families, C.byref(cnt), C.byref(starts), C.byref(evals)))
        fams = [{f: getattr(out[i], f) for f, _ in _MinimumFamilyC._fields_} for i in range(cnt.value)]
        return {"families": fams, "total_starts": starts.value, "total_evals": evals.value}

    def d_min_batch(self, n: int, states, targets=None, max_d: int = -1) -> np.ndarray:
        """LabsEvaluator::compute_d_min (labs_energy.hpp:175-207) of many states on the
        GPU; targets None (or entries < 0): each state's own energy."""
        st = np.ascontiguousarray(np.asarray(states, dtype=np.uint64))
        k = st.shape[0]
        tg = None if targets is None else np.ascontiguousarray(np.asarray(targets, dtype=np.int64))
        out = np.zeros(k, np.int32)
        tp = tg.ctypes.data_as(C.POINTER(C.c_int64)) if tg is not None else None
        _check(lib().bk_d_min_batch(self._h, n, st.ctypes.data_as(C.POINTER(C.c_uint64)), tp, k, max_d,
                                    out.ctypes.data_as(C.POINTER(C.c_int))))
        return out

    def exhaustive_minima(self, n: int, levels: int = 2, num_sd_starts: int = 0, sd_seed: int = 42,
                          cap: int = 1 << 16) -> dict:
        """run_exploration (exhaustive_minima.cpp:109-216) on the GPU: E*, the minima
        and orbit totals, and the trap levels with every orbit's canonical state,
        exact d_min and (with num_sd_starts > 0) sd_hit_fraction."""
        res = _ExhaustiveResultC()
        orb = (_MinimaOrbitC * cap)()
        _check(lib().bk_exhaustive_minima(self._h, n, levels, num_sd_starts, sd_seed, C.byref(res), orb, cap))
        lv = []
        for i in range(res.num_levels):
            lv.append({"E0": res.level_energy[i], "num_orbits": res.level_orbits[i], "orbits": []})
        for i in range(res.orbits_written):
            o = orb[i]
            lv[o.level - 1]["orbits"].append({"canonical": o.canonical_state, "d_min": o.d_min,
                                              "sd_hit_fraction": o.sd_hit_fraction})
        return {"n": res.n, "e_star": res.e_star, "total_minima_found": res.total_minima_found,
                "total_orbits_found": res.total_orbits_found, "num_energies": res.num_energies,
                "levels": lv, "search_time_sec": res.search_time_sec}

    def _basis(self, basis) -> np.ndarray:
        return np.ascontiguousarray(np.asarray(basis, dtype=np.uint64))

    def subspace_lanczos(self, n: int, g: float, basis, max_krylov_dim: int = 100, tol: float = 1e-12,
                         k_pairs: int = 1, normalize_by_n: bool = True, vectors: bool = True):
        """Diagonalizer::solve(basis, g, n, cfg) with DiagMethod::Lanczos
        (diagonalizer.cuh:299-372); eigenvectors in the order of basis."""
        b = self._basis(basis)
        ev = np.zeros(k_pairs)
        vec = np.zeros((k_pairs, b.shape[0])) if vectors else None
        k = _check(lib().bk_subspace_lanczos(self._h, n, g, 1 if normalize_by_n else 0,
                                             b.ctypes.data_as(C.POINTER(C.c_uint64)), b.shape[0],
                                             max_krylov_dim, tol, k_pairs, _dp(ev),
                                             _dp(vec) if vectors else None))
        return ev[:k], (vec[:k] if vectors else None)

    def subspace_dense(self, n: int, g: float, basis, k_pairs: int = 1, normalize_by_n: bool = True,
                       vectors: bool = True):
        """DiagMethod::ExactDense on a basis (diagonalizer.cuh:188-245), rocSOLVER on the device."""
        b = self._basis(basis)
        ev = np.zeros(k_pairs)
        vec = np.zeros((k_pairs, b.shape[0])) if vectors else None
        k = _check(lib().bk_subspace_dense(self._h, n, g, 1 if normalize_by_n else 0,
                                           b.ctypes.data_as(C.POINTER(C.c_uint64)), b.shape[0], k_pairs,
                                           _dp(ev), _dp(vec) if vectors else None))
        return ev[:k], (vec[:k] if vectors else None)

    def fullspace_dense(self, n: int, g: float, k_pairs: int = 1, normalize_by_n: bool = True,
                        vectors: bool = True):
        """DiagMethod::ExactDense on the full space (diagonalizer.cuh:250-294), n <= 14."""
        ev = np.zeros(k_pairs)
        vec = np.zeros((k_pairs, 1 << n)) if vectors else None
        k = _check(lib().bk_fullspace_dense(self._h, n, g, 1 if normalize_by_n else 0, k_pairs, _dp(ev),
                                            _dp(vec) if vectors else None))
        return ev[:k], (vec[:k] if vectors else None)

    def krylov_propagate(self, g: float, tau: float, tau_restart: float = 0.5, m: int = 30,
                         normalize_by_n: bool = True):
        _check(lib().bk_krylov_propagate(self._h, g, 1 if normalize_by_n else 0, tau,
                                         tau_restart, m))


# ---------------------------------------------------------------- sampler API
@dataclasses.dataclass
class SamplerConfig:
    """SamplerConfig (sampler.cuh:235-260); precision/order are engine extensions
    whose defaults (complex128, Lie phase-first) reproduce the reference."""
    n: int = 16
    g: float = 0.5
    dt: float = 0.05
    t_max: float = 60.0
    num_sample_points: int = 20
    shots_per_point: int = 100
    seed: int = 42
    num_sampling_seeds: int = 3
    seed_stride: int = 10007
    t_start: float = 0.0
    dt_sample: float = 0.0
    track_p_less: bool = True
    verbose: bool = True
    record_trajectory: bool = False
    dt_trajectory: float = 0.0
    precision: int = BK_C128
    order: int = BK_LIE


@dataclasses.dataclass
class SeedSamplingStats:
    seed: int = 0
    total_shots_collected: int = 0
    improved_shots: int = 0
    success_rate: float = 0.0
    per_eval_rate: float = 0.0
    mean_hamming_flips: float = 0.0
    best_energy: int = 0


@dataclasses.dataclass
class SamplingResult:
    """SamplingResult (sampler.cuh:272-297); unique_basis is sorted ascending."""
    unique_basis: List[int] = dataclasses.field(default_factory=list)
    total_shots_collected: int = 0
    improved_shots: int = 0
    success_rate: float = 0.0
    per_eval_rate: float = 0.0
    total_evals: int = 0
    mean_hamming_flips: float = 0.0
    seed_stats: List[SeedSamplingStats] = dataclasses.field(default_factory=list)
    time_points: np.ndarray = dataclasses.field(default_factory=lambda: np.zeros(0))
    p_less_history: np.ndarray = dataclasses.field(default_factory=lambda: np.zeros(0))
    dh_history: np.ndarray = dataclasses.field(default_factory=lambda: np.zeros(0))
    window_avg_dh: float = 0.0
    initial_rise_rate_avg: float = 0.0
    initial_rise_rate_max: float = 0.0
    d_min: int = 0
    start_energy: int = 0
    p_blind: float = 0.0
    peak_p_less: float = 0.0
    peak_time: float = 0.0
    samples: np.ndarray = dataclasses.field(default_factory=lambda: np.zeros(0, np.int64))


class ISampler:
    """ISampler (sampler.cuh:303-310)."""

    def sample(self, start_state: int, config: SamplerConfig) -> SamplingResult:
        raise NotImplementedError


class HipSampler(ISampler):
    """The GPU backend behind create_sampler (replaces CuStateVecSampler)."""

    def __init__(self, engine: Optional[Engine] = None):
        self.engine = engine or Engine(0)

    def sample(self, start_state: int, config: SamplerConfig) -> SamplingResult:
        cfg = _SamplerConfigC()
        lib().bk_sampler_config_defaults(C.byref(cfg))
        for f in dataclasses.fields(SamplerConfig):
            v = getattr(config, f.name)
            setattr(cfg, f.name, int(v) if isinstance(v, bool) else v)
        res = C.POINTER(_SamplingResultC)()
        _check(lib().bk_sampler_run(self.engine._h, C.byref(cfg), int(start_state), C.byref(res)))
        try:
            r = res.contents
            hl = r.history_len

            def arr(ptr, k, dt):
                return np.ctypeslib.as_array(ptr, shape=(k,)).astype(dt).copy() if k > 0 \
                    else np.zeros(0, dt)

            out = SamplingResult(
                unique_basis=[int(x) for x in arr(r.unique_basis, r.unique_basis_len, np.uint64)],
                total_shots_collected=r.total_shots_collected, improved_shots=r.improved_shots,
                success_rate=r.success_rate, per_eval_rate=r.per_eval_rate,
                total_evals=r.total_evals, mean_hamming_flips=r.mean_hamming_flips,
                seed_stats=[SeedSamplingStats(s.seed, s.total_shots_collected, s.improved_shots,
                                              s.success_rate, s.per_eval_rate,
                                              s.mean_hamming_flips, s.best_energy)
                            for s in (r.seed_stats[i] for i in range(r.seed_stats_len))],
                time_points=arr(r.time_points, hl, np.float64),
                p_less_history=arr(r.p_less_history, hl, np.float64),
                dh_history=arr(r.dh_history, hl, np.float64),
                window_avg_dh=r.window_avg_dh, initial_rise_rate_avg=r.initial_rise_rate_avg,
                initial_rise_rate_max=r.initial_rise_rate_max, d_min=r.d_min,
                start_energy=r.start_energy, p_blind=r.p_blind, peak_p_less=r.peak_p_less,
                peak_time=r.peak_time, samples=arr(r.samples, r.samples_len, np.int64))
        finally:
            lib().bk_result_free(res)
        return out


class _TooLargeSampler(ISampler):
    def __init__(self, n: int):
        self.n = n

    def sample(self, start_state: int, config: SamplerConfig) -> SamplingResult:
        raise RuntimeError(f"statevector backend supports N <= {BK_MAX_N} (got {self.n})")


def create_sampler(n: int, engine: Optional[Engine] = None) -> ISampler:
    """create_sampler (sampler.cuh:737-743): the HIP backend for n <= 32, a
    throwing placeholder above (the reference's cut is n <= 30)."""
    if n <= BK_MAX_N:
        return HipSampler(engine)
    return _TooLargeSampler(n)
