"""The reference's benchmark configuration files (.cfg / .ini) for this engine.

Python twin of include/bk/config.hpp: the sections, fields, defaults, keys,
aliases and parsing rules of the reference's src/config/config.hpp
(config.hpp:82-389), so a reference config drives the engine unchanged, and
`to_sampler_config`, the mapping of BenchmarkConnector::run_quantum_simulation
(connector.cuh:628-643).

    cfg = config.load_from_file("configs/probability_scan_n30.cfg")
    scfg = config.to_sampler_config(cfg.quantum, n=30, g=0.35, seed=cfg.system.seed_base)
"""
from __future__ import annotations

import dataclasses
import importlib
import io
import os
import sys
from typing import Callable, Dict, List

_bk = importlib.import_module("bernasconi-krylov_amd")


def _trim(s: str) -> str:
    return s.strip(" \t\r\n")


def _uncomment(line: str) -> str:
    quoted = False
    for i, ch in enumerate(line):
        if ch in "\"'":
            quoted = not quoted
        elif not quoted and ch in "#;":
            return line[:i]
    return line


def _boolean(v: str) -> bool:
    s = _trim(v).lower()
    if s in ("1", "true", "yes", "on"):
        return True
    if s in ("0", "false", "no", "off"):
        return False
    raise ValueError(f"Cannot parse boolean value: '{v}'")


def _leading(item: str, typ):
    """The longest leading number of item, as operator>> reads it; None if none."""
    best = None
    for end in range(1, len(item) + 1):
        try:
            best = typ(item[:end])
        except ValueError:
            if best is not None and typ is int:
                break
    return best


def _list(v: str, typ) -> list:
    out = []
    for item in v.split(","):
        item = _trim(item)
        if item:
            x = _leading(item, typ)
            if x is not None:
                out.append(x)
    return out


def _stoi(v: str) -> int:
    """std::stoi / stoul / stoull: leading integer after blanks, else error."""
    x = _leading(v.strip(), int)
    if x is None:
        raise ValueError(f"invalid integer '{v}'")
    return x


def _stod(v: str) -> float:
    x = _leading(v.strip(), float)
    if x is None:
        raise ValueError(f"invalid number '{v}'")
    return x


@dataclasses.dataclass
class SystemConfig:
    n_values: List[int] = dataclasses.field(default_factory=lambda: [16, 20, 24])
    seed_base: int = 42
    device: str = "auto"


@dataclasses.dataclass
class TrapsConfig:
    pool_minima_count: int = 100
    benchmark_trap_count: int = 5
    max_energy_ranks: int = 2
    minima_cache_dir: str = "seeds"
    force_regenerate_traps: bool = False


@dataclasses.dataclass
class QuantumConfig:
    g_values: List[float] = dataclasses.field(default_factory=lambda: [0.25, 0.5, 0.75])
    dt: float = 0.05
    t_start: float = 10.0
    dt_sample: float = 1.5
    k_sample_points: int = 20
    m_shots_per_point: int = 100
    num_sampling_seeds: int = 3
    seed_stride: int = 10007
    track_p_less: bool = True
    verbose: bool = True
    record_trajectory: bool = False
    dt_trajectory: float = 0.0
    t_max_val: float = 0.0
    run_classical: bool = True

    def t_max(self) -> float:
        if self.t_max_val > 0.0:
            return self.t_max_val
        if self.k_sample_points <= 1:
            return self.t_start
        return self.t_start + (self.k_sample_points - 1) * (self.dt_sample if self.dt_sample > 0.0 else self.dt)

    def sample_times(self) -> List[float]:
        if self.record_trajectory:
            step = self.dt_trajectory if self.dt_trajectory > 0.0 else self.dt
            if step <= 0.0:
                step = 0.05
            return [k * step for k in range(int(round(self.t_max() / step)) + 1)]
        step = self.dt_sample if self.dt_sample > 0.0 else self.dt
        return [self.t_start + k * step for k in range(self.k_sample_points)]

    def trotter_steps_per_sample(self) -> int:
        if self.dt <= 0.0:
            return 1
        step = self.dt_sample if self.dt_sample > 0.0 else self.dt
        return max(1, int(round(step / self.dt)))


@dataclasses.dataclass
class ClassicalConfig:
    kick_trials_per_d: int = 50000
    kick_d_values: List[int] = dataclasses.field(default_factory=lambda: [1, 2, 3, 4, 5, 6])
    restart_trials: int = 50000
    max_evals_per_start: int = 0
    batch_size: int = 10000
    enable_tabu: bool = True
    tabu_trials: int = 50000
    tabu_max_steps: int = 100
    tabu_kick_d_values: List[int] = dataclasses.field(default_factory=lambda: [2, 4, 6, 8, 10, 12])
    tabu_tenure: int = 0


@dataclasses.dataclass
class OutputConfig:
    output_dir: str = "data_raw"
    results_csv: str = "benchmark_results_n{N}.csv"
    time_series_csv: str = "p_less_dynamics_n{N}.csv"
    summary_csv: str = "benchmark_summary.csv"
    structured_json: str = "structured_benchmark_n{N}.json"
    all_structured_json: str = "all_seeds_benchmark.json"
    verbose: bool = True

    def resolve_path(self, pattern: str, n: int = 0) -> str:
        s = pattern.replace("{N}", str(n)).replace("{n}", str(n))
        if not os.path.isabs(s) and self.output_dir and os.path.dirname(s) == "":
            return os.path.join(self.output_dir, s)
        return s

    def get_results_path(self, n: int) -> str:
        return self.resolve_path(self.results_csv, n)

    def get_time_series_path(self, n: int) -> str:
        return self.resolve_path(self.time_series_csv, n)

    def get_summary_path(self) -> str:
        return self.resolve_path(self.summary_csv, 0)

    def get_structured_json_path(self, n: int) -> str:
        return self.resolve_path(self.structured_json, n)

    def get_all_structured_json_path(self) -> str:
        return self.resolve_path(self.all_structured_json, 0)


@dataclasses.dataclass
class BenchmarkConfig:
    system: SystemConfig = dataclasses.field(default_factory=SystemConfig)
    traps: TrapsConfig = dataclasses.field(default_factory=TrapsConfig)
    quantum: QuantumConfig = dataclasses.field(default_factory=QuantumConfig)
    classical: ClassicalConfig = dataclasses.field(default_factory=ClassicalConfig)
    output: OutputConfig = dataclasses.field(default_factory=OutputConfig)

    def apply_key_value(self, section: str, key: str, val: str, line_no: int) -> None:
        f = _SETTERS.get(key)
        if f is None:
            print(f"[Config Warning] Unknown parameter on line {line_no}: '{key}'", file=sys.stderr)
            return
        f(self, val)


def _kick(c: BenchmarkConfig, v: str) -> None:
    c.classical.kick_trials_per_d = _stoi(v)
    c.classical.restart_trials = c.classical.kick_trials_per_d


def _verbose(c: BenchmarkConfig, v: str) -> None:
    c.output.verbose = c.quantum.verbose = _boolean(v)


def _set(path: str, conv: Callable):
    part, field = path.split(".")

    def f(c: BenchmarkConfig, v: str) -> None:
        setattr(getattr(c, part), field, conv(v))
    return f


_TABLE = [
    (("n_values", "n", "n_spins"), _set("system.n_values", lambda v: _list(v, int))),
    (("seed_base", "seed", "master_seed"), _set("system.seed_base", _stoi)),
    (("device", "backend"), _set("system.device", str.lower)),
    (("pool_minima_count", "target_minima_count", "minima_pool_size"), _set("traps.pool_minima_count", _stoi)),
    (("benchmark_trap_count", "top_minima_count", "num_traps"), _set("traps.benchmark_trap_count", _stoi)),
    (("max_energy_ranks", "max_ranks", "energy_ranks", "num_ranks"), _set("traps.max_energy_ranks", _stoi)),
    (("minima_cache_dir", "seeds_dir", "seed_dir"), _set("traps.minima_cache_dir", str)),
    (("force_regenerate_traps", "force_regenerate_seeds", "regenerate_seeds"),
     _set("traps.force_regenerate_traps", _boolean)),
    (("g_values",), _set("quantum.g_values", lambda v: _list(v, float))),
    (("g",), _set("quantum.g_values", lambda v: [_stod(v)])),
    (("dt", "simulation_dt"), _set("quantum.dt", _stod)),
    (("t_start", "sampling_start_time", "burn_in_time"), _set("quantum.t_start", _stod)),
    (("dt_sample", "sample_dt", "sampling_delta_t"), _set("quantum.dt_sample", _stod)),
    (("k_sample_points", "k_steps", "num_sample_points", "k"), _set("quantum.k_sample_points", _stoi)),
    (("m_shots_per_point", "m_samples", "shots_per_point", "m"), _set("quantum.m_shots_per_point", _stoi)),
    (("num_sampling_seeds", "sampling_seeds", "num_seeds"), _set("quantum.num_sampling_seeds", _stoi)),
    (("seed_stride", "sampling_seed_stride"), _set("quantum.seed_stride", _stoi)),
    (("track_p_less",), _set("quantum.track_p_less", _boolean)),
    (("record_trajectory", "track_trajectory", "continuous_trajectory"), _set("quantum.record_trajectory", _boolean)),
    (("dt_trajectory", "dt_log", "trajectory_dt"), _set("quantum.dt_trajectory", _stod)),
    (("t_max", "simulation_t_max"), _set("quantum.t_max_val", _stod)),
    (("run_classical", "enable_classical"), _set("quantum.run_classical", _boolean)),
    (("compute_transverse", "eval_transverse"), lambda c, v: None),
    (("kick_trials_per_d", "classical_trials_per_d", "num_classical_trials"), _kick),
    (("kick_d_values", "kick_scan_d", "d_scan", "kick_distances"), _set("classical.kick_d_values", lambda v: _list(v, int))),
    (("restart_trials", "num_restart_trials"), _set("classical.restart_trials", _stoi)),
    (("max_evals_per_start", "evaluation_budget"), _set("classical.max_evals_per_start", _stoi)),
    (("batch_size",), _set("classical.batch_size", _stoi)),
    (("enable_tabu", "run_tabu"), _set("classical.enable_tabu", _boolean)),
    (("tabu_trials", "num_tabu_trials"), _set("classical.tabu_trials", _stoi)),
    (("tabu_max_steps", "tabu_steps"), _set("classical.tabu_max_steps", _stoi)),
    (("tabu_kick_d_values", "tabu_kick_d", "tabu_d_values"), _set("classical.tabu_kick_d_values", lambda v: _list(v, int))),
    (("tabu_tenure", "tenure"), _set("classical.tabu_tenure", _stoi)),
    (("output_dir",), _set("output.output_dir", str)),
    (("results_csv", "csv_results", "csv_output"), _set("output.results_csv", str)),
    (("time_series_csv", "csv_time_series", "csv_trajectory"), _set("output.time_series_csv", str)),
    (("structured_json", "json_output", "structured_output", "json_benchmark"), _set("output.structured_json", str)),
    (("all_structured_json", "all_json_output", "overall_json"), _set("output.all_structured_json", str)),
    (("summary_csv", "csv_summary"), _set("output.summary_csv", str)),
    (("verbose",), _verbose),
]
_SETTERS: Dict[str, Callable] = {k: f for keys, f in _TABLE for k in keys}


def load_from_string(content: str) -> BenchmarkConfig:
    cfg = BenchmarkConfig()
    section = ""
    for line_no, raw in enumerate(io.StringIO(content).read().split("\n"), 1):
        line = _trim(_uncomment(raw))
        if not line:
            continue
        if line[0] == "[" and line[-1] == "]":
            section = _trim(line[1:-1]).lower()
            continue
        eq = line.find("=")
        if eq < 0:
            eq = line.find(":")
        if eq < 0:
            print(f"[Config Warning] Malformed line {line_no}: {raw}", file=sys.stderr)
            continue
        cfg.apply_key_value(section, _trim(line[:eq]).lower(), _trim(line[eq + 1:]), line_no)
    return cfg


def load_from_file(path: str) -> BenchmarkConfig:
    try:
        with open(path) as f:
            return load_from_string(f.read())
    except OSError:
        raise RuntimeError(f"BenchmarkConfig: Cannot open file: {path}") from None


def to_sampler_config(q: QuantumConfig, n: int, g: float, seed: int) -> "_bk.SamplerConfig":
    """connector.cuh:628-643: the sampler run of one (n, g) of a benchmark config."""
    return _bk.SamplerConfig(n=n, g=g, dt=q.dt, t_max=q.t_max(), num_sample_points=q.k_sample_points,
                             shots_per_point=q.m_shots_per_point, seed=seed, t_start=q.t_start,
                             dt_sample=q.dt_sample, track_p_less=q.track_p_less, verbose=q.verbose,
                             record_trajectory=q.record_trajectory, dt_trajectory=q.dt_trajectory,
                             num_sampling_seeds=q.num_sampling_seeds, seed_stride=q.seed_stride)
