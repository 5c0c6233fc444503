"""ctypes front-end of the CPU oracle (oracle/libbk_oracle.so).

TEST INFRASTRUCTURE ONLY.  The checker for the HIP product path: imported by
tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg, never by the
product package.  Every function cites, through bk_oracle.h, the reference
file:line whose algorithm it restates.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libbk_oracle.so")
REF_PATH = os.path.join(HERE, "_ref", "libbk_ref.so")

_lib = None
_ref = None

_i64 = C.c_int64
_u64 = C.c_uint64
_dp = C.POINTER(C.c_double)
_i64p = C.POINTER(C.c_int64)


def build(with_ref: bool | None = None) -> None:
    """Compile the oracle (and oracle/_ref when the reference sources exist)."""
    targets = ["all"]
    if with_ref is None:
        with_ref = os.path.isdir("/root/reference/src")
    if with_ref:
        targets.append("ref")
    subprocess.run(["make", "-s", "-C", HERE, *targets], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build(with_ref=False)
        L = C.CDLL(LIB_PATH)
        L.bko_energy.restype = _i64
        L.bko_energy.argtypes = [C.c_int, _u64]
        L.bko_ck.argtypes = [C.c_int, _u64, _i64p]
        L.bko_canonical.restype = _u64
        L.bko_canonical.argtypes = [C.c_int, _u64]
        L.bko_known_optimal_energy.restype = _i64
        L.bko_known_optimal_energy.argtypes = [C.c_int]
        L.bko_d_min.restype = C.c_int
        L.bko_d_min.argtypes = [C.c_int, _u64, _i64, C.c_int]
        L.bko_all_energies.argtypes = [C.c_int, C.POINTER(C.c_int32)]
        L.bko_count_below.restype = _i64
        L.bko_count_below.argtypes = [C.c_int, _i64]
        L.bko_random_state.restype = _u64
        L.bko_random_state.argtypes = [C.c_int, _u64]
        L.bko_steepest_descent.argtypes = [C.c_int, _u64, C.POINTER(_u64), _i64p,
                                           C.POINTER(_u64), C.POINTER(_u64)]
        L.bko_mt64_seed.argtypes = [C.c_void_p, _u64]
        L.bko_mt64_next.restype = _u64
        L.bko_mt64_next.argtypes = [C.c_void_p]
        L.bko_mt64_uniform.restype = C.c_double
        L.bko_mt64_uniform.argtypes = [C.c_void_p, C.c_double, C.c_double]
        L.bko_trotter.restype = C.c_int
        L.bko_trotter.argtypes = [C.c_int, _u64, _i64, C.c_double, _dp, C.c_double, C.c_int,
                                  C.c_int, C.c_int, _dp, _dp, _dp, _dp, C.c_int]
        L.bko_trotter_bench.restype = C.c_int
        L.bko_adjoint_grad.argtypes = [C.c_int, _u64, _i64, _dp, C.c_double, C.c_int, C.c_int,
                                       C.c_double, C.c_double, _dp, _dp, C.c_int]
        L.bko_adjoint_grad.restype = C.c_int
        L.bko_trotter_bench.argtypes = [C.c_int, _u64, C.c_double, C.c_double, C.c_int, C.c_int,
                                        C.c_int, _dp]
        L.bko_observables.argtypes = [C.c_int, _dp, _u64, _i64, _dp, _dp, _dp]
        L.bko_hv_real.argtypes = [C.c_int, C.c_double, C.c_int, _dp, _dp]
        L.bko_hv_complex.argtypes = [C.c_int, C.c_double, C.c_int, _dp, _dp]
        L.bko_dense_h.argtypes = [C.c_int, C.c_double, C.c_int, _dp]
        L.bko_syev.restype = C.c_int
        L.bko_syev.argtypes = [C.c_int, _dp, _dp]
        L.bko_tridiag_eig.restype = C.c_int
        L.bko_tridiag_eig.argtypes = [C.c_int, _dp, _dp, _dp]
        L.bko_ed_evolve.restype = C.c_int
        L.bko_ed_evolve.argtypes = [C.c_int, C.c_double, _u64, _i64, _dp, C.c_int, _dp, _dp,
                                    C.c_int, C.c_int]
        L.bko_lanczos.restype = C.c_int
        L.bko_lanczos.argtypes = [C.c_int, C.c_double, C.c_int, C.c_int, C.c_double, C.c_int,
                                  _dp, _dp]
        L.bko_krylov_propagate.restype = C.c_int
        L.bko_krylov_propagate.argtypes = [C.c_int, C.c_double, C.c_int, _dp, C.c_double,
                                           C.c_int]
        L.bko_sample_inverse_cdf.argtypes = [C.c_int, _dp, _dp, C.c_int, _i64p]
        L.bko_derived_metrics.argtypes = [_dp, _dp, _dp, C.c_int, C.c_double, C.c_double,
                                          _dp, _dp, _dp, _dp, _dp, _dp]
        L.bko_schedule.restype = C.c_int
        L.bko_schedule.argtypes = [C.c_double, C.c_double, C.c_int, C.c_int, C.c_double,
                                   C.c_double, C.c_int, C.c_int, C.c_double,
                                   C.POINTER(C.c_ubyte), C.c_int]
        L.bko_exhaustive_minima.restype = C.c_int
        L.bko_exhaustive_minima.argtypes = [C.c_int, C.c_int, _i64p, _i64p, _i64p, _i64p, _i64p, _i64p,
                                            C.POINTER(_u64), _i64]
        _lib = L
    return _lib


def ref():
    """The reference's own CPU code (oracle/_ref), or None when not built."""
    global _ref
    if _ref is None and os.path.exists(REF_PATH):
        R = C.CDLL(REF_PATH)
        R.bkref_energy.restype = _i64
        R.bkref_energy.argtypes = [C.c_int, _u64]
        R.bkref_ck.argtypes = [C.c_int, _u64, _i64p]
        R.bkref_canonical.restype = _u64
        R.bkref_canonical.argtypes = [C.c_int, _u64]
        R.bkref_known_optimal_energy.restype = _i64
        R.bkref_known_optimal_energy.argtypes = [C.c_int]
        R.bkref_d_min.restype = C.c_int
        R.bkref_d_min.argtypes = [C.c_int, _u64, _i64, C.c_int]
        R.bkref_random_state.restype = _u64
        R.bkref_random_state.argtypes = [C.c_int, _u64]
        R.bkref_steepest_descent.argtypes = [C.c_int, _u64, C.POINTER(_u64), _i64p,
                                             C.POINTER(_u64), C.POINTER(_u64)]
        R.bkref_exhaustive_minima.restype = C.c_int
        R.bkref_exhaustive_minima.argtypes = [C.c_int, C.c_int, _u64, _i64p, _i64p, _i64p, _i64p, _i64p,
                                              C.POINTER(_u64), C.POINTER(C.c_int), _dp, _i64, C.c_char_p]
        _ref = R
    return _ref


def _p(a):
    return a.ctypes.data_as(_dp)


# ---------------------------------------------------------------- L0 core
def energy(n: int, s: int) -> int:
    return lib().bko_energy(n, s)


def ck(n: int, s: int) -> np.ndarray:
    out = np.zeros(n, dtype=np.int64)
    lib().bko_ck(n, s, out.ctypes.data_as(_i64p))
    return out


def canonical(n: int, s: int) -> int:
    return lib().bko_canonical(n, s)


def known_optimal_energy(n: int) -> int:
    return lib().bko_known_optimal_energy(n)


def d_min(n: int, s0: int, target: int = -1, max_d: int = -1) -> int:
    return lib().bko_d_min(n, s0, target, max_d)


def all_energies(n: int) -> np.ndarray:
    out = np.empty(1 << n, dtype=np.int32)
    lib().bko_all_energies(n, out.ctypes.data_as(C.POINTER(C.c_int32)))
    return out


def count_below(n: int, e0: int) -> int:
    return lib().bko_count_below(n, e0)


class MT64:
    """std::mt19937_64 (+ libstdc++ uniform_real_distribution<double>)."""

    def __init__(self, seed: int):
        self._buf = C.create_string_buffer(312 * 8 + 16)
        lib().bko_mt64_seed(self._buf, seed)

    def next(self) -> int:
        return lib().bko_mt64_next(self._buf)

    def uniform(self, a: float = 0.0, b: float = 1.0) -> float:
        return lib().bko_mt64_uniform(self._buf, a, b)


def random_state(n: int, seed: int) -> int:
    return lib().bko_random_state(n, seed)


def steepest_descent(n: int, start: int):
    fs, fe, ev, st = _u64(), _i64(), _u64(), _u64()
    lib().bko_steepest_descent(n, start, C.byref(fs), C.byref(fe), C.byref(ev), C.byref(st))
    return fs.value, fe.value, ev.value, st.value


def make_trap(n: int, seed: int):
    """seed -> random_state -> steepest_descent (classical.cuh:602-713)."""
    start = random_state(n, seed)
    s0, e0, evals, steps = steepest_descent(n, start)
    return start, s0, e0


# ---------------------------------------------------------------- Trotter
def trotter(n, s0, g, dt, steps, order=1, e0=None, g_steps=None, psi=None, threads=0,
            want_state=False):
    """FP64 Trotter evolution from |s0> (or from psi).  Returns dict of series."""
    dim = 1 << n
    if e0 is None:
        e0 = energy(n, s0)
    init = psi is None
    state = np.zeros(2 * dim, dtype=np.float64) if init else \
        np.ascontiguousarray(np.asarray(psi, dtype=np.complex128)).view(np.float64).copy()
    pl = np.zeros(steps + 1)
    dh = np.zeros(steps + 1)
    nm = np.zeros(steps + 1)
    gs = None
    if g_steps is not None:
        gs = np.ascontiguousarray(g_steps, dtype=np.float64)
        assert gs.shape == (steps,)
    rc = lib().bko_trotter(n, s0, e0, float(g), _p(gs) if gs is not None else None, float(dt),
                           int(steps), int(order), 1 if init else 0, _p(state), _p(pl), _p(dh),
                           _p(nm), int(threads))
    if rc != 0:
        raise RuntimeError(f"bko_trotter failed rc={rc}")
    out = {"p_less": pl, "dh": dh, "norm": nm}
    if want_state:
        out["psi"] = state.view(np.complex128)
    return out


def adjoint_grad(n, s0, g_steps, dt, order=2, e0=None, wp=1.0, wd=0.0, threads=0):
    """dL/dg_k for L = wp*P_<(T) + wd*<d_H>(T) (FP64 reverse mode).  Returns
    (grad per step, (P_<(T), <d_H>(T)))."""
    if e0 is None:
        e0 = energy(n, s0)
    gs = np.ascontiguousarray(g_steps, dtype=np.float64)
    grad = np.zeros(gs.shape[0])
    obs = np.zeros(2)
    rc = lib().bko_adjoint_grad(n, s0, e0, _p(gs), float(dt), int(gs.shape[0]), int(order), float(wp),
                                float(wd), _p(grad), _p(obs), int(threads))
    if rc != 0:
        raise RuntimeError(f"bko_adjoint_grad failed rc={rc}")
    return grad, (obs[0], obs[1])


def trotter_bench(n, s0, g, dt, steps, order=2, threads=0):
    """Seconds spent in `steps` FP64 Trotter steps (setup excluded)."""
    sec = C.c_double()
    rc = lib().bko_trotter_bench(n, s0, g, dt, steps, order, threads, C.byref(sec))
    if rc != 0:
        raise RuntimeError(f"bko_trotter_bench failed rc={rc}")
    return sec.value


def observables(n, psi, s0, e0):
    st = np.ascontiguousarray(np.asarray(psi, dtype=np.complex128)).view(np.float64)
    a, b, c = C.c_double(), C.c_double(), C.c_double()
    lib().bko_observables(n, _p(st), s0, e0, C.byref(a), C.byref(b), C.byref(c))
    return a.value, b.value, c.value


# ---------------------------------------------------------------- H.v / ED
def hv_real(n, g, x, normalize_by_n=True):
    x = np.ascontiguousarray(x, dtype=np.float64)
    y = np.empty_like(x)
    lib().bko_hv_real(n, g, 1 if normalize_by_n else 0, _p(x), _p(y))
    return y


def hv_complex(n, g, x, normalize_by_n=True):
    xv = np.ascontiguousarray(x, dtype=np.complex128).view(np.float64)
    y = np.empty_like(xv)
    lib().bko_hv_complex(n, g, 1 if normalize_by_n else 0, _p(xv), _p(y))
    return y.view(np.complex128)


def dense_h(n, g, normalize_by_n=True):
    dim = 1 << n
    h = np.empty((dim, dim), dtype=np.float64)
    lib().bko_dense_h(n, g, 1 if normalize_by_n else 0, _p(h))
    return h


def syev(a):
    a = np.array(a, dtype=np.float64, order="C", copy=True)
    dim = a.shape[0]
    w = np.empty(dim)
    rc = lib().bko_syev(dim, _p(a), _p(w))
    if rc != 0:
        raise RuntimeError(f"bko_syev failed rc={rc}")
    return w, a  # eigenvectors as rows


def tridiag_eig(d, e, vectors=True):
    d = np.array(d, dtype=np.float64, copy=True)
    e = np.array(e, dtype=np.float64, copy=True)
    m = d.shape[0]
    z = np.eye(m) if vectors else None
    rc = lib().bko_tridiag_eig(m, _p(d), _p(e), _p(z) if vectors else None)
    if rc != 0:
        raise RuntimeError(f"bko_tridiag_eig failed rc={rc}")
    return d, z


def ed_evolve(n, g, s0, times, e0=None, inner_loop_trig=False, threads=0):
    if e0 is None:
        e0 = energy(n, s0)
    t = np.ascontiguousarray(times, dtype=np.float64)
    pl = np.empty_like(t)
    dh = np.empty_like(t)
    rc = lib().bko_ed_evolve(n, g, s0, e0, _p(t), t.shape[0], _p(pl), _p(dh),
                             1 if inner_loop_trig else 0, threads)
    if rc != 0:
        raise RuntimeError(f"bko_ed_evolve failed rc={rc}")
    return pl, dh


def lanczos(n, g, max_krylov_dim=100, tol=1e-12, k_pairs=1, normalize_by_n=True,
            vectors=True):
    dim = 1 << n
    ev = np.empty(k_pairs)
    vec = np.empty((k_pairs, dim)) if vectors else None
    rc = lib().bko_lanczos(n, g, 1 if normalize_by_n else 0, max_krylov_dim, tol, k_pairs,
                           _p(ev), _p(vec) if vectors else None)
    if rc < 0:
        raise RuntimeError(f"bko_lanczos failed rc={rc}")
    return ev[:rc], (vec[:rc] if vectors else None)


def krylov_propagate(n, g, psi, tau, m=30, normalize_by_n=True):
    st = np.array(psi, dtype=np.complex128, copy=True).view(np.float64)
    rc = lib().bko_krylov_propagate(n, g, 1 if normalize_by_n else 0, _p(st), float(tau), m)
    if rc != 0:
        raise RuntimeError(f"bko_krylov_propagate failed rc={rc}")
    return st.view(np.complex128)


def sample_inverse_cdf(n, psi, randnums):
    st = np.ascontiguousarray(psi, dtype=np.complex128).view(np.float64)
    r = np.ascontiguousarray(randnums, dtype=np.float64)
    out = np.empty(r.shape[0], dtype=np.int64)
    lib().bko_sample_inverse_cdf(n, _p(st), _p(r), r.shape[0], out.ctypes.data_as(_i64p))
    return out


def derived_metrics(t, p, dh, avg_start, avg_end):
    t = np.ascontiguousarray(t, dtype=np.float64)
    p = np.ascontiguousarray(p, dtype=np.float64)
    dh = np.ascontiguousarray(dh, dtype=np.float64)
    outs = [C.c_double() for _ in range(6)]
    lib().bko_derived_metrics(_p(t), _p(p), _p(dh), t.shape[0], avg_start, avg_end,
                              *[C.byref(o) for o in outs])
    keys = ["rise_avg", "rise_max", "peak_p", "peak_t", "window_dh", "window_p"]
    return {k: o.value for k, o in zip(keys, outs)}


def schedule(t_max, dt, num_sample_points=20, shots_per_point=100, t_start=0.0,
             dt_sample=0.0, track_p_less=True, record_trajectory=False, dt_trajectory=0.0):
    total = int(round(t_max / dt))
    flags = np.zeros(total + 1, dtype=np.uint8)
    lib().bko_schedule(t_max, dt, num_sample_points, shots_per_point, t_start, dt_sample,
                       1 if track_p_less else 0, 1 if record_trajectory else 0, dt_trajectory,
                       flags.ctypes.data_as(C.POINTER(C.c_ubyte)), flags.shape[0])
    return flags


# ---------------------------------------------------------------- exhaustive minima
def exhaustive_minima(n: int, levels: int = 2, cap: int = 1 << 16) -> dict:
    """run_exploration (exhaustive_minima.cpp:109-208) without d_min: E*, the
    reported minima and orbit totals, and the trap levels' energies and
    canonical orbits (ascending)."""
    es, tm, to, ne = _i64(), _i64(), _i64(), _i64()
    le = np.zeros(max(levels, 1), dtype=np.int64)
    lo = np.zeros(max(levels, 1), dtype=np.int64)
    orb = np.zeros(cap, dtype=np.uint64)
    k = lib().bko_exhaustive_minima(n, levels, C.byref(es), C.byref(tm), C.byref(to), C.byref(ne),
                                    le.ctypes.data_as(_i64p), lo.ctypes.data_as(_i64p),
                                    orb.ctypes.data_as(C.POINTER(_u64)), cap)
    if k < 0:
        raise RuntimeError("bko_exhaustive_minima failed")
    out, pos = [], 0
    for lv in range(k):
        cnt = int(lo[lv])
        out.append({"E0": int(le[lv]), "orbits": [int(x) for x in orb[pos:pos + cnt]]})
        pos += cnt
    return {"n": n, "e_star": es.value, "total_minima": tm.value, "total_orbits": to.value,
            "num_energies": ne.value, "levels": out}


def ref_exhaustive_minima(n: int, with_sd: bool = False, num_sd: int = 100000, json_path=None,
                          cap: int = 1 << 16):
    """The reference's own run_exploration (oracle/_ref), or None when not built."""
    R = ref()
    if R is None:
        return None
    es, tm, to = _i64(), _i64(), _i64()
    le, lo = np.zeros(2, dtype=np.int64), np.zeros(2, dtype=np.int64)
    canon = np.zeros(cap, dtype=np.uint64)
    dmin = np.zeros(cap, dtype=np.int32)
    sd = np.zeros(cap)
    k = R.bkref_exhaustive_minima(n, 1 if with_sd else 0, num_sd, C.byref(es), C.byref(tm), C.byref(to),
                                  le.ctypes.data_as(_i64p), lo.ctypes.data_as(_i64p),
                                  canon.ctypes.data_as(C.POINTER(_u64)),
                                  dmin.ctypes.data_as(C.POINTER(C.c_int)), _p(sd), cap,
                                  json_path.encode() if json_path else None)
    out, pos = [], 0
    for lv in range(k):
        cnt = int(lo[lv])
        out.append({"E0": int(le[lv]),
                    "orbits": [{"canonical": int(canon[i]), "d_min": int(dmin[i]), "sd_hit_fraction": float(sd[i])}
                               for i in range(pos, pos + cnt)]})
        pos += cnt
    return {"n": n, "e_star": es.value, "total_minima": tm.value, "total_orbits": to.value, "levels": out}
