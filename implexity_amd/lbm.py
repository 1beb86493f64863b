"""Python binding of ``libimplexity_lbm.so`` (the C ABI in include/implexity_lbm.h).

The names mirror the reference's LBM solver and adjoint interfaces
(crates/implexity-physics-lbm/src/solver.rs:359-947 and crates/implexity-ad/src/
scan.rs, revolve.rs) so that host code reads like the reference:

    ``Lattice.initial_populations``  LbmProblem::initial_populations (solver.rs:592)
    ``Lattice.interval``             LbmProblem::interval (solver.rs:604)
    ``Lattice.interval_vjp``         LbmProblem::interval_vjp (solver.rs:642)
    ``Lattice.history``              LbmProblem::history + objective (solver.rs:677)
    ``Lattice.gradient``             history_adjoint behind scan_adjoint with
                                     Checkpointing::Binomial (solver.rs:798, scan.rs:140)
    ``Lattice.fields``               LbmProblem::fields in lattice units (solver.rs:690)
    ``BinomialSchedule``             implexity-ad BinomialSchedule (revolve.rs:128)

The library is the only compute path: there is no CPU fallback, and importing this
module raises if the HIP library has not been built.
"""
from __future__ import annotations

import ctypes
import math
import os
from dataclasses import dataclass, field

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libimplexity_lbm.so")

PRECISION = {"f64": 0, "f32": 1, "f16s": 2, "bf16s": 3}
ADJOINT_STORAGE = {"auto": 0, "f64": 1, "f32": 2, "f16": 3}
COLLISION = {"bgk": 0, "trt": 1, "mrt": 2, "cumulant": 3, "bgk_guo": 4, "regularized": 5}
FACE = {"xmin": 0, "xmax": 1, "ymin": 2, "ymax": 3, "zmin": 4, "zmax": 5}
COUPLING = {"brinkman": 0, "psm": 1, "psm_superposition": 2, "brinkman_occupancy": 3}
OBJECTIVE = {"none": 0, "dissipation": 1, "pressure_drop": 2, "vector": 3, "thermal": 4, "solid_force": 5}
MAGIC_DEFAULT = 3.0 / 16.0  # moving/collision.rs:36


class LbmError(RuntimeError):
    """Base error; ``code`` is the lbm_status value."""

    code = -1


class ContractError(LbmError):
    """CaeError::Contract (crates/implexity-core/src/error.rs:34)."""

    code = 1


class ConvergenceError(LbmError):
    code = 2


class ShapeError(LbmError):
    code = 3


class InvalidError(LbmError):
    code = 4


class DeviceError(LbmError):
    code = 5


class MemoryError_(LbmError):
    code = 6


_ERRORS = {c.code: c for c in (ContractError, ConvergenceError, ShapeError, InvalidError, DeviceError, MemoryError_)}


class _Port(ctypes.Structure):
    _fields_ = [
        ("face", ctypes.c_uint32), ("kind", ctypes.c_uint32), ("value", ctypes.c_double * 3),
        ("profile", ctypes.c_void_p), ("profile_len", ctypes.c_uint64),
        ("mask", ctypes.c_void_p), ("mask_len", ctypes.c_uint64),
    ]


class _Sponge(ctypes.Structure):
    _fields_ = [
        ("face", ctypes.c_uint32), ("thickness", ctypes.c_uint32), ("strength", ctypes.c_double),
        ("density", ctypes.c_double), ("velocity", ctypes.c_double * 3), ("wave", ctypes.c_uint32),
        ("reserved", ctypes.c_uint32), ("wave_amplitude", ctypes.c_double * 3),
        ("wave_direction", ctypes.c_double * 3), ("wave_speed", ctypes.c_double),
        ("wave_frequency", ctypes.c_double), ("wave_phase", ctypes.c_double), ("wave_ramp", ctypes.c_double),
        ("wave_origin", ctypes.c_double * 3),
    ]


class _Config(ctypes.Structure):
    _fields_ = [
        ("struct_size", ctypes.c_uint32), ("lattice", ctypes.c_uint32), ("collision", ctypes.c_uint32),
        ("precision", ctypes.c_uint32), ("shape", ctypes.c_int32 * 3), ("periodic", ctypes.c_uint8 * 3),
        ("reserved0", ctypes.c_uint8 * 5), ("omega", ctypes.c_double), ("magic", ctypes.c_double),
        ("bulk_rate", ctypes.c_double), ("even_rate", ctypes.c_double), ("accel", ctypes.c_double * 3),
        ("n_ports", ctypes.c_uint32), ("reserved1", ctypes.c_uint32), ("ports", ctypes.c_void_p),
        ("checkpoint_budget_bytes", ctypes.c_uint64), ("adjoint_storage", ctypes.c_uint32),
        ("checkpoint_every", ctypes.c_uint32),
        ("thermal", ctypes.c_uint32), ("reserved2", ctypes.c_uint32), ("thermal_kappa", ctypes.c_double),
        ("thermal_ratio", ctypes.c_double), ("thermal_source", ctypes.c_double),
        ("thermal_inlet", ctypes.c_double), ("thermal_range", ctypes.c_double),
        ("les_constant", ctypes.c_double), ("les_norm_floor", ctypes.c_double),
        ("wale_constant", ctypes.c_double), ("wale_floor", ctypes.c_double),
        ("n_sponges", ctypes.c_uint32), ("reserved3", ctypes.c_uint32), ("sponges", ctypes.c_void_p),
        ("coupling", ctypes.c_uint32), ("reserved4", ctypes.c_uint32), ("psm_kappa", ctypes.c_double),
        ("psm_c_alpha", ctypes.c_double), ("psm_drag_shape", ctypes.c_double),
        ("mach_limit", ctypes.c_double), ("drift_limit", ctypes.c_double), ("tau", ctypes.c_double),
    ]


class _Scales(ctypes.Structure):
    _fields_ = [("velocity", ctypes.c_double), ("density", ctypes.c_double), ("step_s", ctypes.c_double),
                ("cell_volume", ctypes.c_double)]


def _scales(sc):
    v = _Scales()
    v.velocity, v.density, v.step_s, v.cell_volume = (float(x) for x in sc)
    return v


class _Objective(ctypes.Structure):
    _fields_ = [
        ("kind", ctypes.c_uint32), ("axis", ctypes.c_uint32), ("plane_a", ctypes.c_int32),
        ("plane_b", ctypes.c_int32), ("t_begin", ctypes.c_uint64), ("t_end", ctypes.c_uint64),
        ("weight", ctypes.c_double), ("terminal_weight", ctypes.c_double),
        ("vector_weights", ctypes.c_double * 6), ("scales", _Scales),
    ]


class Stats(ctypes.Structure):
    _fields_ = [
        ("forward_steps", ctypes.c_uint64), ("reverse_steps", ctypes.c_uint64),
        ("snapshot_slots", ctypes.c_uint64), ("peak_snapshots", ctypes.c_uint64),
        ("snapshot_copies", ctypes.c_uint64), ("forward_ms", ctypes.c_double), ("adjoint_ms", ctypes.c_double),
        ("total_ms", ctypes.c_double), ("max_mach", ctypes.c_double), ("mass_drift", ctypes.c_double),
        ("bad_cells", ctypes.c_uint64), ("forward_kernel_avg_ms", ctypes.c_double),
        ("adjoint_kernel_avg_ms", ctypes.c_double), ("kernel_samples", ctypes.c_uint64),
        ("cotangent_max", ctypes.c_double), ("cotangent_exponent", ctypes.c_int64),
        ("min_population", ctypes.c_double),
    ]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


_lib = None

# Every symbol include/implexity_lbm.h declares, with its argument and result types.
_SIGS = {
    "lbm_abi_version": ([], ctypes.c_int),
    "lbm_last_error": ([ctypes.c_char_p, ctypes.c_uint64], ctypes.c_int),
    "lbm_create": ([ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint64, ctypes.c_void_p], ctypes.c_int),
    "lbm_destroy": ([ctypes.c_void_p], ctypes.c_int),
    "lbm_set_design": ([ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint64], ctypes.c_int),
    "lbm_set_solid_momentum": ([ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint64], ctypes.c_int),
    "lbm_design_map_create": ([ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint64,
                               ctypes.c_void_p], ctypes.c_int),
    "lbm_design_map_destroy": ([ctypes.c_void_p], ctypes.c_int),
    "lbm_pushforward": ([ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint64,
                         ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint64], ctypes.c_int),
    "lbm_pushforward_vjp": ([ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint64,
                             ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint64, ctypes.c_void_p, ctypes.c_void_p,
                             ctypes.c_void_p], ctypes.c_int),
    "lbm_design_map_apply": ([ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint64], ctypes.c_int),
    "lbm_design_map_vjp": ([ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint64],
                           ctypes.c_int),
    "lbm_solid_momentum_bar": ([ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint64], ctypes.c_int),
    "lbm_set_occupancy_end": ([ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint64, ctypes.c_uint64,
                               ctypes.c_uint64], ctypes.c_int),
    "lbm_occupancy_end_bar": ([ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint64], ctypes.c_int),
    "lbm_subcycle": ([ctypes.c_void_p, ctypes.c_uint32, ctypes.c_void_p, ctypes.c_uint64, ctypes.c_void_p],
                     ctypes.c_int),
    "lbm_subcycle_adjoint": ([ctypes.c_void_p, ctypes.c_uint32, ctypes.c_void_p, ctypes.c_uint64, ctypes.c_void_p,
                              ctypes.c_uint64, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint64, ctypes.c_void_p],
                             ctypes.c_int),
    "lbm_reset_initial": ([ctypes.c_void_p], ctypes.c_int),
    "lbm_set_state": ([ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_uint64], ctypes.c_int),
    "lbm_get_state": ([ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_uint64], ctypes.c_int),
    "lbm_step_count": ([ctypes.c_void_p], ctypes.c_uint64),
    "lbm_step": ([ctypes.c_void_p, ctypes.c_uint64], ctypes.c_int),
    "lbm_forward": ([ctypes.c_void_p, ctypes.c_uint64, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p],
                    ctypes.c_int),
    "lbm_gradient": ([ctypes.c_void_p, ctypes.c_uint64, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                      ctypes.c_uint64, ctypes.c_void_p, ctypes.c_uint64, ctypes.c_void_p], ctypes.c_int),
    "lbm_gradient_seeded": ([ctypes.c_void_p, ctypes.c_uint64, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint64,
                             ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint64, ctypes.c_void_p, ctypes.c_uint64,
                             ctypes.c_void_p], ctypes.c_int),
    "lbm_vector": ([ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint64, ctypes.c_void_p, ctypes.c_void_p], ctypes.c_int),
    "lbm_vector_vjp": ([ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint64, ctypes.c_void_p, ctypes.c_void_p,
                        ctypes.c_void_p, ctypes.c_uint64, ctypes.c_void_p, ctypes.c_void_p], ctypes.c_int),
    "lbm_step_vjp": ([ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint64],
                     ctypes.c_int),
    "lbm_reserve": ([ctypes.c_void_p, ctypes.c_uint64, ctypes.c_void_p], ctypes.c_int),
    "lbm_alpha_bar": ([ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint64], ctypes.c_int),
    "lbm_step_jvp": ([ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint64],
                     ctypes.c_int),
    "lbm_vector_noalloc": ([ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint64, ctypes.c_void_p, ctypes.c_void_p],
                           ctypes.c_int),
    "lbm_get_fields": ([ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint64], ctypes.c_int),
    "lbm_debug_link_table": ([ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint64], ctypes.c_int),
    "lbm_port_cells": ([ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint64, ctypes.c_void_p],
                       ctypes.c_int),
    "lbm_revolve_schedule": ([ctypes.c_uint64, ctypes.c_uint64, ctypes.c_void_p, ctypes.c_uint64, ctypes.c_void_p,
                              ctypes.c_void_p, ctypes.c_void_p], ctypes.c_int),
    "lbm_every_schedule": ([ctypes.c_uint64, ctypes.c_uint64, ctypes.c_void_p, ctypes.c_uint64, ctypes.c_void_p,
                            ctypes.c_void_p, ctypes.c_void_p], ctypes.c_int),
    "lbm_revolve_optimal_forward_steps": ([ctypes.c_uint64, ctypes.c_uint64], ctypes.c_uint64),
    "lbm_time_kernel": ([ctypes.c_void_p, ctypes.c_int, ctypes.c_uint64, ctypes.c_void_p], ctypes.c_int),
    "lbm_copy_bandwidth": ([ctypes.c_uint64, ctypes.c_uint64, ctypes.c_void_p, ctypes.c_void_p], ctypes.c_int),
    "lbm_state_bytes": ([ctypes.c_void_p], ctypes.c_uint64),
    "lbm_set_thermal_design": ([ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint64], ctypes.c_int),
    "lbm_get_thermal_state": ([ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint64], ctypes.c_int),
    "lbm_set_thermal_state": ([ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint64], ctypes.c_int),
    "lbm_get_temperature": ([ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint64], ctypes.c_int),
    "lbm_thermal_design_bar": ([ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint64], ctypes.c_int),
    "lbm_get_thermal_lambda": ([ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint64], ctypes.c_int),
}


def lib():
    """Load the HIP library (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(
                f"{LIB_PATH} is missing: build it with `python -m implexity_amd.build` "
                "(there is no CPU fallback for the LBM hot path)")
        _lib = ctypes.CDLL(LIB_PATH)
        for name, (args, res) in _SIGS.items():
            fn = getattr(_lib, name)
            fn.argtypes = args
            fn.restype = res
    return _lib


def _check(code: int) -> None:
    if code == 0:
        return
    buf = ctypes.create_string_buffer(1024)
    lib().lbm_last_error(buf, len(buf))
    raise _ERRORS.get(code, LbmError)(buf.value.decode(errors="replace"))


def _p(a):
    return ctypes.c_void_p(a.ctypes.data) if a is not None else None


# lbm_port_kind: the moving solver's NEE rule (moving/boundary.rs:445-506), and the
# stationary solver's reconstruct_neighbor (ports.rs:484-497), moment reconstruction
# ("legacy_moment", ports.rs:343-406) and regularized reconstruction
# ("regularized_known_stress", ports.rs:409-482), D3Q19.
PORT_KIND = {"velocity": 0, "pressure": 1, "pressure_neighbor": 2, "velocity_moment": 3, "pressure_moment": 4,
             "velocity_regularized": 5, "pressure_regularized": 6}


@dataclass
class Port:
    """Inlet/outlet on a face (moving/boundary.rs:126-143 PortSpec; ports.rs:56-71 Port).

    value: velocity (3,) in lattice units for the velocity kinds; (rho,) for "pressure";
    (rho, u_t0, u_t1) for the stationary pressure kinds, the authored tangential velocity
    in ascending axis order.  profile: one factor per port cell (velocity kinds).  mask:
    lattice-shaped booleans selecting the port's cells on the face (default: the whole
    face)."""

    face: str
    kind: str  # a key of PORT_KIND
    value: tuple
    profile: np.ndarray | None = None
    mask: np.ndarray | None = None


@dataclass
class Wave:
    """Travelling wave on a sponge's far-field velocity (moving/sponge.rs:22-29), lattice
    units: amplitude (velocity), unit direction, speed (cells per step), frequency
    (cycles per step), phase (rad), ramp (steps), origin (cells)."""

    amplitude: tuple
    direction: tuple
    speed: float
    frequency: float
    phase: float = 0.0
    ramp: float = 0.0
    origin: tuple = (0.0, 0.0, 0.0)


@dataclass
class Sponge:
    """Sponge layer (moving/sponge.rs:12-20 SpongeSpec) in lattice units: the far field
    is a density and a velocity."""

    face: str
    thickness: int
    strength: float
    density: float = 1.0
    velocity: tuple = (0.0, 0.0, 0.0)
    wave: Wave | None = None


@dataclass
class LatticeConfig:
    shape: tuple
    periodic: tuple = (False, False, False)
    lattice: int = 19
    collision: str = "trt"
    precision: str = "f32"
    omega: float = 1.0
    magic: float = MAGIC_DEFAULT
    bulk_rate: float = float("nan")
    even_rate: float = float("nan")
    accel: tuple = (0.0, 0.0, 0.0)
    ports: list = field(default_factory=list)
    checkpoint_budget_bytes: int = 0
    solid: np.ndarray | None = None
    # cotangent storage: "auto" (FP64 / FP32 / scaled FP16 for f64 / f32, bf16s / f16s),
    # "f64", "f32" or "f16" (scaled FP16, f16s storage only)
    adjoint_storage: str = "auto"
    # 0: binomial checkpointing inside the budget; k: Checkpointing::Every(k) (scan.rs:109-246)
    checkpoint_every: int = 0
    # Coupled D3Q7 temperature field (no counterpart in the reference): diffusivity
    # kappa (1 + (ratio - 1) phi), source * phi per step, T = inlet at velocity ports,
    # zero-gradient at pressure ports; `thermal_range` bounds |T| for FP16 storage.
    thermal: bool = False
    thermal_kappa: float = 0.0
    thermal_ratio: float = 1.0
    thermal_source: float = 0.0
    thermal_inlet: float = 0.0
    thermal_range: float = 0.0
    # Smagorinsky closure on BGK/TRT (moving/psm.rs:45-48,161-197); 0 = off
    les_constant: float = 0.0
    les_norm_floor: float = 0.0
    # WALE closure on BGK/TRT (moving/wale.rs); 0 = off
    wale_constant: float = 0.0
    wale_floor: float = 1e-12
    # Sponge layers blended after collision (moving/sponge.rs, field.rs:1461-1475)
    sponges: list = field(default_factory=list)
    # Fluid-solid coupling (moving/psm.rs:198-267): "brinkman" (the design is alpha), or
    # "psm" / "psm_superposition" / "brinkman_occupancy" (the design is the occupancy d,
    # with a solid momentum per cell and the saturation width psm_kappa; the last with
    # alpha = resistance(d; psm_c_alpha, psm_drag_shape))
    coupling: str = "brinkman"
    psm_kappa: float = 0.0
    psm_c_alpha: float = 0.0
    psm_drag_shape: float = 0.1
    # Admission of the first sweep's trajectory (LbmProblem::admit, solver.rs:707-793):
    # 0 = none; otherwise history/gradient raise ConvergenceError on a non-finite or
    # non-positive fluid population, Mach above mach_limit (<= 0.2) or a relative mass
    # drift above drift_limit (0: 1e-6)
    mach_limit: float = 0.0
    drift_limit: float = 0.0
    # the relaxation time itself (0: 1/omega); > 0 replaces omega by 1/tau
    tau: float = 0.0


@dataclass
class Objective:
    """J = sum_{s in [t_begin, t_end]} weight * phi(state s) + terminal_weight * phi(state n)."""

    kind: str = "none"
    t_begin: int = 1
    t_end: int = 1
    weight: float = 0.0
    terminal_weight: float = 0.0
    axis: int = 0
    plane_a: int = 1
    plane_b: int = -2
    # "vector": phi = w . vector(state) (LbmProblem::vector, solver.rs:874-894) with
    # scales = (velocity, density, step_s, cell_volume); w[5] (design volume) is not in J
    vector_weights: tuple = (0.0,) * 6
    scales: tuple = (1.0, 1.0, 1.0, 1.0)

    def _c(self):
        o = _Objective()
        o.vector_weights = (ctypes.c_double * 6)(*[float(x) for x in self.vector_weights])
        o.scales = _scales(self.scales)
        o.kind = OBJECTIVE[self.kind]
        o.axis = self.axis
        o.plane_a = self.plane_a
        o.plane_b = self.plane_b
        o.t_begin = self.t_begin
        o.t_end = self.t_end
        o.weight = self.weight
        o.terminal_weight = self.terminal_weight
        return o


class Lattice:
    """One lattice problem resident on the GPU (an lbm_handle)."""

    def __init__(self, cfg: LatticeConfig):
        self.cfg = cfg
        self.shape = tuple(int(s) for s in cfg.shape)
        self.n = int(np.prod(self.shape))
        self.Q = int(cfg.lattice)
        c = _Config()
        c.struct_size = ctypes.sizeof(_Config)
        c.lattice = self.Q
        c.collision = COLLISION[cfg.collision]
        c.precision = PRECISION[cfg.precision]
        for d in range(3):
            c.shape[d] = self.shape[d]
            c.periodic[d] = 1 if cfg.periodic[d] else 0
            c.accel[d] = float(cfg.accel[d])
        c.omega = float(cfg.omega)
        c.magic = float(cfg.magic)
        c.bulk_rate = float(cfg.bulk_rate)
        c.even_rate = float(cfg.even_rate)
        self._profiles = []
        ports = (_Port * max(1, len(cfg.ports)))()
        for i, port in enumerate(cfg.ports):
            ports[i].face = FACE[port.face]
            if port.kind not in PORT_KIND:
                raise ContractError(f"unknown port kind {port.kind!r}")
            ports[i].kind = PORT_KIND[port.kind]
            vals = list(port.value) + [0.0] * (3 - len(port.value))
            for d in range(3):
                ports[i].value[d] = float(vals[d])
            if port.profile is not None:
                prof = np.ascontiguousarray(port.profile, np.float64).reshape(-1)
                self._profiles.append(prof)
                ports[i].profile = prof.ctypes.data
                ports[i].profile_len = prof.size
            if port.mask is not None:
                pm = np.ascontiguousarray(np.asarray(port.mask, bool).reshape(-1), np.uint8)
                self._profiles.append(pm)
                ports[i].mask = pm.ctypes.data
                ports[i].mask_len = pm.size
        c.n_ports = len(cfg.ports)
        c.ports = ctypes.cast(ports, ctypes.c_void_p)
        c.checkpoint_budget_bytes = int(cfg.checkpoint_budget_bytes)
        if cfg.adjoint_storage not in ADJOINT_STORAGE:
            raise ValueError(f"adjoint_storage must be one of {sorted(ADJOINT_STORAGE)}")
        c.adjoint_storage = ADJOINT_STORAGE[cfg.adjoint_storage]
        c.checkpoint_every = int(cfg.checkpoint_every)
        c.thermal = 1 if cfg.thermal else 0
        c.thermal_kappa = float(cfg.thermal_kappa)
        c.thermal_ratio = float(cfg.thermal_ratio)
        c.thermal_source = float(cfg.thermal_source)
        c.thermal_inlet = float(cfg.thermal_inlet)
        c.thermal_range = float(cfg.thermal_range)
        c.les_constant = float(cfg.les_constant)
        c.les_norm_floor = float(cfg.les_norm_floor)
        c.wale_constant = float(cfg.wale_constant)
        c.wale_floor = float(cfg.wale_floor)
        sponges = (_Sponge * max(1, len(cfg.sponges)))()
        for i, sp in enumerate(cfg.sponges):
            sponges[i].face = FACE[sp.face]
            sponges[i].thickness = int(sp.thickness)
            sponges[i].strength = float(sp.strength)
            sponges[i].density = float(sp.density)
            for d in range(3):
                sponges[i].velocity[d] = float(sp.velocity[d])
            if sp.wave is not None:
                w = sp.wave
                sponges[i].wave = 1
                for d in range(3):
                    sponges[i].wave_amplitude[d] = float(w.amplitude[d])
                    sponges[i].wave_direction[d] = float(w.direction[d])
                    sponges[i].wave_origin[d] = float(w.origin[d])
                sponges[i].wave_speed = float(w.speed)
                sponges[i].wave_frequency = float(w.frequency)
                sponges[i].wave_phase = float(w.phase)
                sponges[i].wave_ramp = float(w.ramp)
        if cfg.coupling not in COUPLING:
            raise ValueError(f"coupling must be one of {sorted(COUPLING)}")
        c.coupling = COUPLING[cfg.coupling]
        c.psm_kappa = float(cfg.psm_kappa)
        c.psm_c_alpha = float(cfg.psm_c_alpha)
        c.psm_drag_shape = float(cfg.psm_drag_shape)
        c.mach_limit = float(cfg.mach_limit)
        c.drift_limit = float(cfg.drift_limit)
        c.tau = float(cfg.tau)
        c.n_sponges = len(cfg.sponges)
        c.sponges = ctypes.cast(sponges, ctypes.c_void_p)
        solid = None
        if cfg.solid is not None:
            solid = np.ascontiguousarray(np.asarray(cfg.solid, bool).reshape(-1), np.uint8)
        h = ctypes.c_void_p()
        _check(lib().lbm_create(ctypes.byref(c), _p(solid), 0 if solid is None else solid.size, ctypes.byref(h)))
        self._h = h
        self._ports_c = ports
        self._sponges_c = sponges

    def close(self):
        if getattr(self, "_h", None):
            lib().lbm_destroy(self._h)
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

    # --- design and state ----------------------------------------------------------
    def set_design(self, alpha):
        a = np.ascontiguousarray(alpha, np.float64).reshape(-1)
        _check(lib().lbm_set_design(self._h, _p(a), a.size))

    def initial_populations(self):
        _check(lib().lbm_reset_initial(self._h))

    # --- occupancy coupling ---------------------------------------------------------------
    def set_solid_momentum(self, m):
        """Solid momentum per cell, (cells, 3) (PSM couplings; the design is the occupancy)."""
        a = np.ascontiguousarray(m, np.float64).reshape(-1)
        _check(lib().lbm_set_solid_momentum(self._h, _p(a), a.size))

    def solid_momentum_bar(self):
        out = np.zeros(self.n * 3)
        _check(lib().lbm_solid_momentum_bar(self._h, _p(out), out.size))
        return out.reshape(self.n, 3)

    def set_occupancy_end(self, d1, m1, t0, t1):
        """Second occupancy endpoint blended in over the steps t0..t1."""
        d = np.ascontiguousarray(d1, np.float64).reshape(-1)
        m = None if m1 is None else np.ascontiguousarray(m1, np.float64).reshape(-1)
        _check(lib().lbm_set_occupancy_end(self._h, _p(d), _p(m), d.size, int(t0), int(t1)))

    def occupancy_end_bar(self):
        d = np.zeros(self.n)
        m = np.zeros(self.n * 3)
        _check(lib().lbm_occupancy_end_bar(self._h, _p(d), _p(m), self.n))
        return d, m.reshape(self.n, 3)

    def subcycle(self, substeps):
        """One macro step of the subcycled moving field (SubcycledField::subcycle,
        moving/field.rs:1870-1919) from the held state between the occupancy endpoints
        (design, solid momentum) and set_occupancy_end; returns (g, stats), g the
        endpoints' impulse shares (cells, 2, 3)."""
        g = np.zeros(self.n * 6)
        st = Stats()
        _check(lib().lbm_subcycle(self._h, int(substeps), _p(g), g.size, ctypes.byref(st)))
        return g.reshape(self.n, 2, 3), st

    def subcycle_adjoint(self, substeps, state_bar, g_bar=None):
        """Transpose of subcycle from the held (start) state: returns (previous_bar, d0_bar,
        m0_bar, d1_bar, m1_bar, stats) for the cotangents of the end state (AoS) and of g."""
        sb = np.ascontiguousarray(state_bar, np.float64).reshape(-1)
        gb = None if g_bar is None else np.ascontiguousarray(g_bar, np.float64).reshape(-1)
        pb = np.zeros(sb.size)
        d0 = np.zeros(self.n)
        st = Stats()
        _check(lib().lbm_subcycle_adjoint(self._h, int(substeps), _p(sb), sb.size, _p(gb), 0 if gb is None else gb.size,
                                          _p(pb), _p(d0), self.n, ctypes.byref(st)))
        d1, m1 = self.occupancy_end_bar()
        return pb, d0, self.solid_momentum_bar(), d1, m1, st

    # --- coupled temperature field ------------------------------------------------------
    def set_thermal_design(self, phi):
        a = np.ascontiguousarray(phi, np.float64).reshape(-1)
        _check(lib().lbm_set_thermal_design(self._h, _p(a), a.size))

    def thermal_state(self):
        out = np.empty(self.n * 7)
        _check(lib().lbm_get_thermal_state(self._h, _p(out), out.size))
        return out

    def set_thermal_state(self, g):
        g = np.ascontiguousarray(g, np.float64).reshape(-1)
        _check(lib().lbm_set_thermal_state(self._h, _p(g), g.size))

    def temperature(self):
        out = np.empty(self.n)
        _check(lib().lbm_get_temperature(self._h, _p(out), out.size))
        return out

    def thermal_design_bar(self):
        """dJ/dphi through the temperature field of the last gradient call."""
        out = np.empty(self.n)
        _check(lib().lbm_thermal_design_bar(self._h, _p(out), out.size))
        return out

    def thermal_lambda(self):
        """Cotangent of the temperature populations of state 0 of the last gradient call."""
        out = np.empty(self.n * 7)
        _check(lib().lbm_get_thermal_lambda(self._h, _p(out), out.size))
        return out

    def set_state(self, f, layout="aos"):
        f = np.ascontiguousarray(f, np.float64).reshape(-1)
        _check(lib().lbm_set_state(self._h, 0 if layout == "aos" else 1, _p(f), f.size))

    def state(self, layout="aos"):
        out = np.empty(self.n * self.Q)
        _check(lib().lbm_get_state(self._h, 0 if layout == "aos" else 1, _p(out), out.size))
        return out

    @property
    def step_count(self):
        return int(lib().lbm_step_count(self._h))

    @property
    def state_bytes(self):
        return int(lib().lbm_state_bytes(self._h))

    # --- dynamics -------------------------------------------------------------------
    def interval(self, steps=1):
        _check(lib().lbm_step(self._h, steps))

    def history(self, steps, objective: Objective | None = None):
        """Forward sweep; returns (J, stats)."""
        J = ctypes.c_double(0.0)
        st = Stats()
        o = objective._c() if objective else None
        _check(lib().lbm_forward(self._h, steps, ctypes.byref(o) if o else None, ctypes.byref(J), ctypes.byref(st)))
        return J.value, st

    def reserve(self, steps):
        """Allocate the snapshot pool for a horizon ahead of time; returns the slot count."""
        slots = ctypes.c_uint64(0)
        _check(lib().lbm_reserve(self._h, steps, ctypes.byref(slots)))
        return slots.value

    def gradient(self, steps, objective: Objective | None, want_lambda0=False, lambda_final=None,
                 alpha_bar_on_device=False):
        """Forward + discrete adjoint from the current state (binomial checkpointing).
        lambda_final (reference AoS, optional) is added to the final state's cotangent, as
        history_adjoint's lambda_final (solver.rs:798-815).  With alpha_bar_on_device the
        gradient stays in device memory (alpha_bar is returned as None; fetch it with
        ``alpha_bar()``).
        Returns (J, alpha_bar, lambda0 or None, stats)."""
        J = ctypes.c_double(0.0)
        st = Stats()
        ab = None if alpha_bar_on_device else np.empty(self.n)
        lam = np.zeros(self.n * self.Q) if want_lambda0 else None
        o = objective._c() if objective is not None else None
        lf = None if lambda_final is None else np.ascontiguousarray(lambda_final, np.float64).reshape(-1)
        _check(lib().lbm_gradient_seeded(self._h, steps, ctypes.byref(o) if o is not None else None, _p(lf),
                                         0 if lf is None else lf.size, ctypes.byref(J), _p(ab), self.n, _p(lam),
                                         0 if lam is None else lam.size, ctypes.byref(st)))
        return J.value, ab, lam, st

    def alpha_bar(self):
        """dJ/dalpha of the last gradient or step VJP, copied out of device memory."""
        ab = np.empty(self.n)
        _check(lib().lbm_alpha_bar(self._h, _p(ab), self.n))
        return ab

    def vector(self, scales, phi=None):
        """Terminal responses of the current state (solver.rs:874-894)."""
        out = np.zeros(6)
        ph = None if phi is None else np.ascontiguousarray(phi, np.float64).reshape(-1)
        sc = _scales(scales)
        _check(lib().lbm_vector(self._h, _p(ph), 0 if ph is None else ph.size, ctypes.byref(sc), _p(out)))
        return out

    def vector_noalloc(self, scales, out_dev, phi_dev=None):
        """lbm_vector into device memory without allocating: out_dev (6 doubles) and phi_dev
        (cells doubles, or None) are device addresses (ints) on the handle's device."""
        sc = _scales(scales)
        _check(lib().lbm_vector_noalloc(self._h, None if phi_dev is None else ctypes.c_void_p(phi_dev),
                                        0 if phi_dev is None else self.n, ctypes.byref(sc), ctypes.c_void_p(out_dev)))

    def vector_vjp(self, scales, w, phi=None):
        """(f_bar, alpha_bar, phi_bar) of w . vector at the current state (solver.rs:896-947)."""
        ph = None if phi is None else np.ascontiguousarray(phi, np.float64).reshape(-1)
        ww = np.ascontiguousarray(w, np.float64)
        fb = np.zeros(self.n * self.Q)
        ab = np.zeros(self.n)
        pb = np.zeros(self.n)
        sc = _scales(scales)
        _check(lib().lbm_vector_vjp(self._h, _p(ph), 0 if ph is None else ph.size, ctypes.byref(sc), _p(ww), _p(fb),
                                    fb.size, _p(ab), _p(pb)))
        return fb, ab, pb

    def interval_vjp(self, lambda_next):
        lam = np.ascontiguousarray(lambda_next, np.float64).reshape(-1)
        out = np.empty_like(lam)
        ab = np.empty(self.n)
        _check(lib().lbm_step_vjp(self._h, _p(lam), _p(out), _p(ab), lam.size))
        return out, ab

    def interval_jvp(self, d_state, d_alpha=None):
        """Tangent of one step at the current state: J_state d_state + J_alpha d_alpha
        (ScanStep::jvp, scan.rs:46-80); the state is not advanced."""
        dx = np.ascontiguousarray(d_state, np.float64).reshape(-1)
        da = None if d_alpha is None else np.ascontiguousarray(d_alpha, np.float64).reshape(-1)
        if da is not None and da.size != self.n:
            raise ValueError("d_alpha needs one entry per cell")
        out = np.empty_like(dx)
        _check(lib().lbm_step_jvp(self._h, _p(dx), _p(da), _p(out), dx.size))
        return out

    def fields(self):
        rho = np.empty(self.n)
        u = np.empty(3 * self.n)
        _check(lib().lbm_get_fields(self._h, _p(rho), _p(u), self.n))
        return rho, u.reshape(self.n, 3)

    def link_table(self):
        out = np.empty(self.n * self.Q, np.uint32)
        _check(lib().lbm_debug_link_table(self._h, _p(out), out.size))
        return out

    def port_cells(self):
        cnt = ctypes.c_uint64(0)
        _check(lib().lbm_port_cells(self._h, None, None, 0, ctypes.byref(cnt)))
        cells = np.empty(cnt.value, np.int32)
        nbs = np.empty(cnt.value, np.int32)
        _check(lib().lbm_port_cells(self._h, _p(cells), _p(nbs), cnt.value, ctypes.byref(cnt)))
        return cells, nbs

    def time_kernel(self, which="forward", launches=10):
        ms = ctypes.c_double(0.0)
        code = {"forward": 0, "adjoint": 1, "stream": 2, "coupled_forward": 3, "thermal_adjoint": 4, "sponge": 5,
                "sponge_transpose": 6, "forward_diagnostics": 7}[which]
        _check(lib().lbm_time_kernel(self._h, code, launches, ctypes.byref(ms)))
        return ms.value


class _DesignMapConfig(ctypes.Structure):
    _fields_ = [("struct_size", ctypes.c_uint32), ("shape", ctypes.c_int32 * 3), ("periodic", ctypes.c_uint8 * 3),
                ("reserved", ctypes.c_uint8 * 5), ("radius", ctypes.c_double), ("beta", ctypes.c_double),
                ("eta", ctypes.c_double)]


class DesignMap:
    """Design map raw -> phi on the device (DesignMap, design.rs:95-280): cone filter of
    `radius` cells over the editable cells (design region, not solid), then the tanh
    projection (beta, eta); solid cells map to 0 and the rest of the lattice to `fixed`."""

    def __init__(self, shape, periodic=(False, False, False), radius=0.0, beta=0.0, eta=0.5, solid=None,
                 region=None, fixed=None):
        self.shape = tuple(int(s) for s in shape)
        self.n = int(np.prod(self.shape))
        c = _DesignMapConfig()
        c.struct_size = ctypes.sizeof(_DesignMapConfig)
        for d in range(3):
            c.shape[d] = self.shape[d]
            c.periodic[d] = 1 if periodic[d] else 0
        c.radius, c.beta, c.eta = float(radius), float(beta), float(eta)
        sol = None if solid is None else np.ascontiguousarray(np.asarray(solid, bool).reshape(-1), np.uint8)
        reg = None if region is None else np.ascontiguousarray(np.asarray(region, bool).reshape(-1), np.uint8)
        fix = None if fixed is None else np.ascontiguousarray(fixed, np.float64).reshape(-1)
        for arr in (sol, reg, fix):
            if arr is not None and arr.size != self.n:
                raise ShapeError("design map masks need one entry per cell")
        h = ctypes.c_void_p()
        _check(lib().lbm_design_map_create(ctypes.byref(c), _p(sol), _p(reg), _p(fix), self.n, ctypes.byref(h)))
        self._h = h

    def close(self):
        if getattr(self, "_h", None):
            lib().lbm_design_map_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def map(self, raw):
        r = np.ascontiguousarray(raw, np.float64).reshape(-1)
        out = np.empty(self.n)
        _check(lib().lbm_design_map_apply(self._h, _p(r), _p(out), r.size))
        return out

    def vjp(self, raw, phi_bar):
        r = np.ascontiguousarray(raw, np.float64).reshape(-1)
        pb = np.ascontiguousarray(phi_bar, np.float64).reshape(-1)
        if pb.size != r.size:
            raise ShapeError("phi_bar needs one value per cell")
        out = np.empty(self.n)
        _check(lib().lbm_design_map_vjp(self._h, _p(r), _p(pb), _p(out), r.size))
        return out


class _PushConfig(ctypes.Structure):
    _fields_ = [("struct_size", ctypes.c_uint32), ("shape", ctypes.c_int32 * 3), ("periodic", ctypes.c_uint8 * 3),
                ("mirrored", ctypes.c_uint8 * 3), ("reserved", ctypes.c_uint8 * 2), ("width_cells", ctypes.c_uint32),
                ("kernel", ctypes.c_uint32)]


PUSH_KERNEL = {"cubic": 0, "peskin4": 1}


def _push_config(shape, periodic, width, kernel, mirrored=None):
    if kernel not in PUSH_KERNEL:
        raise ContractError("unknown interpolation kernel")
    c = _PushConfig()
    c.struct_size = ctypes.sizeof(_PushConfig)
    for d in range(3):
        c.shape[d] = int(shape[d])
        c.periodic[d] = 1 if periodic[d] else 0
        if mirrored is not None:
            lo, hi = mirrored[d]
            c.mirrored[d] = (1 if lo else 0) | (2 if hi else 0)
    c.width_cells, c.kernel = int(width), PUSH_KERNEL[kernel]
    return c


def pushforward(shape, xi, a, v=None, periodic=(False, False, False), width=1, kernel="cubic", mirrored=None):
    """Occupancy d (cells) and solid momentum m (cells x 3) of body points xi (lattice
    coordinates) with weights a and velocities v (moving/pushforward.rs scatter);
    mirrored[axis] = (low face, high face) marks symmetry faces (Frame::with_mirrors)."""
    n = int(np.prod(shape))
    x = np.ascontiguousarray(xi, np.float64).reshape(-1)
    aa = np.ascontiguousarray(a, np.float64).reshape(-1)
    vv = None if v is None else np.ascontiguousarray(v, np.float64).reshape(-1)
    if x.size != 3 * aa.size or (vv is not None and vv.size != x.size):
        raise ShapeError("points need three coordinates, one weight and three velocity components each")
    d = np.empty(n)
    m = np.empty(3 * n)
    c = _push_config(shape, periodic, width, kernel, mirrored)
    _check(lib().lbm_pushforward(ctypes.byref(c), _p(x), _p(aa), _p(vv), aa.size, _p(d), _p(m), n))
    return d, m.reshape(n, 3)


def pushforward_vjp(shape, xi, a, v, d_bar, m_bar, periodic=(False, False, False), width=1, kernel="cubic",
                    mirrored=None):
    """(xi_bar, a_bar, v_bar) of <d_bar, d> + <m_bar, m> (endpoint_vjp)."""
    n = int(np.prod(shape))
    x = np.ascontiguousarray(xi, np.float64).reshape(-1)
    aa = np.ascontiguousarray(a, np.float64).reshape(-1)
    vv = np.ascontiguousarray(v, np.float64).reshape(-1)
    db = np.ascontiguousarray(d_bar, np.float64).reshape(-1)
    mb = np.ascontiguousarray(m_bar, np.float64).reshape(-1)
    if x.size != 3 * aa.size or vv.size != x.size or db.size != n or mb.size != 3 * n:
        raise ShapeError("push-forward cotangent shapes")
    xb, ab, vb = np.empty(x.size), np.empty(aa.size), np.empty(x.size)
    c = _push_config(shape, periodic, width, kernel, mirrored)
    _check(lib().lbm_pushforward_vjp(ctypes.byref(c), _p(x), _p(aa), _p(vv), aa.size, _p(db), _p(mb), n, _p(xb),
                                     _p(ab), _p(vb)))
    return xb.reshape(-1, 3), ab, vb.reshape(-1, 3)


class BinomialSchedule:
    """Revolve schedule built by the library (revolve.rs:128-329)."""

    KINDS = ("advance", "snapshot", "restore", "reverse", "release")

    _EXPORT = "lbm_revolve_schedule"

    def __init__(self, steps: int, snapshots: int):
        cnt = ctypes.c_uint64(0)
        fwd = ctypes.c_uint64(0)
        rs = ctypes.c_uint64(0)
        export = getattr(lib(), self._EXPORT)
        _check(export(steps, snapshots, None, 0, ctypes.byref(cnt), None, None))
        buf = np.zeros(5 * cnt.value, np.int64)
        _check(export(steps, snapshots, _p(buf), cnt.value, ctypes.byref(cnt), ctypes.byref(fwd), ctypes.byref(rs)))
        self.steps = steps
        self.forward_steps = fwd.value
        self.reverse_start = rs.value
        acts = []
        for k, a, b, _, _ in buf.reshape(-1, 5):
            name = self.KINDS[k]
            if name in ("advance", "snapshot", "restore"):
                acts.append((name, int(a), int(b)))
            elif name == "reverse":
                acts.append((name, int(a)))
            else:
                acts.append((name, int(a)))
        self.actions = acts


class EverySchedule(BinomialSchedule):
    """Checkpointing::Every(k) action list (scan.rs:109-246) as the gradient driver runs it."""

    _EXPORT = "lbm_every_schedule"


def optimal_forward_steps(steps: int, snapshots: int) -> int:
    return int(lib().lbm_revolve_optimal_forward_steps(steps, snapshots))


def copy_bandwidth(nbytes: int = 1 << 30, iters: int = 20):
    g = ctypes.c_double(0.0)
    ms = ctypes.c_double(0.0)
    _check(lib().lbm_copy_bandwidth(nbytes, iters, ctypes.byref(g), ctypes.byref(ms)))
    return g.value, ms.value


def odd_rate(omega: float, magic: float) -> float:
    """TRT odd relaxation rate (moving/collision.rs:302-305)."""
    tp = 1.0 / omega - 0.5
    return 1.0 / (0.5 + magic / tp)


def declared_symbols(header: str) -> list[str]:
    """Function names declared in include/implexity_lbm.h."""
    import re

    text = open(header).read()
    return sorted(set(re.findall(r"^\s*(?:int|uint64_t)\s+(lbm_\w+)\s*\(", text, flags=re.M)))
