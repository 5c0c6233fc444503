"""The benchmark connector's output files for sampler results (Python side).

`TrapRecord` is what BenchmarkConnector::run keeps of one trap at one field g
(SingleTrapBenchmark, connector.cuh:26-117, quantum part).  The three export
functions write the results CSV, the time-series CSV and the structured JSON
through the library's C++ writers (include/bk/wire.hpp, connector.cuh:839-1155),
byte-for-byte as the reference writes them for the same values.

    rec = wire.TrapRecord(n=16, g=0.3, trap_rank=1, trap_bits=s0, canonical_bits=c,
                          trap_energy=e0, result=bk.create_sampler(16).sample(s0, cfg),
                          window_t_start=cfg.t_start, window_points=cfg.num_sample_points,
                          window_dt=cfg.dt_sample, t_max=cfg.t_max)
    wire.export_time_series_csv("out/probability_dynamics_n16.csv", [rec])
"""
from __future__ import annotations

import ctypes as C
import dataclasses
import importlib
from typing import List, Optional, Sequence

import numpy as np

_bk = importlib.import_module("bernasconi-krylov_amd")


@dataclasses.dataclass
class TrapRecord:
    n: int
    g: float
    trap_rank: int
    trap_bits: int
    canonical_bits: int
    trap_energy: int
    d_min: int = 0                       # 0: taken from the result
    quantum_time_sec: float = 0.0
    window_t_start: float = 0.0          # window of window_avg_p_less (connector.cuh:816-830)
    window_points: int = 0
    window_dt: float = 0.0
    t_max: float = 0.0
    result: Optional["_bk.SamplingResult"] = None


def _result_c(r, keep: list):
    """A C view of a SamplingResult; the arrays it points at go into `keep`."""
    def darr(v):
        a = np.ascontiguousarray(v, dtype=np.float64)
        keep.append(a)
        return a.ctypes.data_as(C.POINTER(C.c_double)), a.shape[0]

    def uarr(v):
        a = np.ascontiguousarray(np.asarray(v, dtype=np.uint64))
        keep.append(a)
        return a.ctypes.data_as(C.POINTER(C.c_uint64)), a.shape[0]

    c = _bk._SamplingResultC()
    c.unique_basis, c.unique_basis_len = uarr(r.unique_basis)
    c.total_shots_collected = int(r.total_shots_collected)
    c.improved_shots = int(r.improved_shots)
    c.success_rate = float(r.success_rate)
    c.per_eval_rate = float(r.per_eval_rate)
    c.total_evals = int(r.total_evals)
    c.mean_hamming_flips = float(r.mean_hamming_flips)
    stats = (_bk._SeedStatsC * max(1, len(r.seed_stats)))()
    for i, s in enumerate(r.seed_stats):
        stats[i] = _bk._SeedStatsC(int(s.seed), int(s.total_shots_collected), int(s.improved_shots),
                                   float(s.success_rate), float(s.per_eval_rate),
                                   float(s.mean_hamming_flips), int(s.best_energy))
    keep.append(stats)
    c.seed_stats = C.cast(stats, C.POINTER(_bk._SeedStatsC))
    c.seed_stats_len = len(r.seed_stats)
    k = min(len(r.time_points), len(r.p_less_history), len(r.dh_history))
    c.time_points, _ = darr(np.asarray(r.time_points)[:k])
    c.p_less_history, _ = darr(np.asarray(r.p_less_history)[:k])
    c.dh_history, _ = darr(np.asarray(r.dh_history)[:k])
    c.history_len = k
    c.window_avg_dh = float(r.window_avg_dh)
    c.initial_rise_rate_avg = float(r.initial_rise_rate_avg)
    c.initial_rise_rate_max = float(r.initial_rise_rate_max)
    c.d_min = int(r.d_min)
    c.start_energy = int(r.start_energy)
    c.p_blind = float(r.p_blind)
    c.peak_p_less = float(r.peak_p_less)
    c.peak_time = float(r.peak_time)
    keep.append(c)
    return C.pointer(c)


def _records(recs: Sequence[TrapRecord]):
    keep: list = []
    arr = (_bk._TrapRecordC * max(1, len(recs)))()
    for i, r in enumerate(recs):
        arr[i] = _bk._TrapRecordC(int(r.n), float(r.g), int(r.trap_rank), int(r.trap_bits),
                                  int(r.canonical_bits), int(r.trap_energy), int(r.d_min),
                                  float(r.quantum_time_sec), float(r.window_t_start),
                                  int(r.window_points), float(r.window_dt), float(r.t_max),
                                  _result_c(r.result, keep) if r.result is not None else None)
    return arr, keep


def export_results_csv(path: str, records: List[TrapRecord]) -> None:
    """export_results_csv (connector.cuh:839-904)."""
    arr, _keep = _records(records)
    _bk._check(_bk.lib().bk_export_results_csv(str(path).encode(), arr, len(records)))


def export_time_series_csv(path: str, records: List[TrapRecord]) -> None:
    """export_time_series_csv (connector.cuh:906-933)."""
    arr, _keep = _records(records)
    _bk._check(_bk.lib().bk_export_time_series_csv(str(path).encode(), arr, len(records)))


def export_structured_json(path: str, records: List[TrapRecord], n_override: int = 0) -> None:
    """export_structured_json (connector.cuh:935-1155)."""
    arr, _keep = _records(records)
    _bk._check(_bk.lib().bk_export_structured_json(str(path).encode(), arr, len(records), int(n_override)))


def export_exhaustive_json(path: str, result: dict) -> None:
    """save_to_json (exhaustive_minima.cpp:328-372) of a result shaped like
    Engine.exhaustive_minima's: {"n", "e_star", "levels": [{"E0", "orbits":
    [{"canonical", "d_min", "sd_hit_fraction"}]}]}."""
    levels = result["levels"]
    if len(levels) > _bk.BK_MINIMA_MAX_LEVELS:
        raise ValueError("too many levels")
    res = _bk._ExhaustiveResultC()
    res.n = int(result["n"])
    res.e_star = int(result["e_star"])
    res.num_levels = len(levels)
    orbs = []
    for i, lv in enumerate(levels):
        res.level_energy[i] = int(lv["E0"])
        res.level_orbits[i] = len(lv["orbits"])
        for o in lv["orbits"]:
            orbs.append(_bk._MinimaOrbitC(int(o["canonical"]), int(lv["E0"]), i + 1, int(o["d_min"]),
                                          float(o.get("sd_hit_fraction", 0.0))))
    arr = (_bk._MinimaOrbitC * max(1, len(orbs)))(*orbs)
    _bk._check(_bk.lib().bk_export_exhaustive_json(path.encode(), C.byref(res), arr, len(orbs)))
