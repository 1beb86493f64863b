"""Problem normalisation: the reference's JSON LBM problem to the library's lattice.

``normalize`` follows LbmProblem::normalize (crates/implexity-physics-lbm/src/
solver.rs:414-550): the same fifteen required keys and five optional ones, the same
checks in the same order and the same messages, with the solver's derived quantities
in lattice units:

    tau      = 1/2 + 3 nu dt / dx^2, 0.51 <= tau <= 2          (solver.rs:487-493)
    a_lattice = a dt^2 / dx                                    (solver.rs:575-578)
    alpha(phi) = dt drag_max q (1 - phi) / (q + phi)           (design.rs:282-294)
    vector scales (dx/dt, density, dt, dx^3)                   (solver.rs:862-871)
    topology map radius in cells = filter_radius_m / dx        (design.rs:51-80)

The reference's stationary step is BGK with Guo forcing and halfway bounce-back
(METHOD_VERSION "d3q19-bgk-guo-bounceback-porous-v1"), so the lattice configuration is
D3Q19 BGK-Guo with the relaxation time tau handed over as tau (not as its reciprocal),
so the library divides by the reference's tau bit for bit; admission (LbmProblem::admit)
uses the problem's Mach limit.  Ports map onto the library's stationary port kinds with
their masks: "legacy_moment" to the moment reconstruction (ports.rs:343-406),
"regularized_known_stress" to the regularized reconstruction (:409-482),
"neighbor_non_equilibrium" to the neighbour rule (:484-497).  The implicit-geometry
source and field registration (geometry::masks_from_document, check_registration) are
refused with a contract error: they stay with the caller's host pipeline.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

from . import design
from .lbm import ContractError, DesignMap, LatticeConfig, Port

PROBLEM_KEYS = ("acceleration_m_s2", "density_kg_m3", "design_region", "drag_max_per_s", "drag_shape",
                "fixed_design", "history_byte_budget", "kinematic_viscosity_m2_s", "mach_limit", "periodic_axes",
                "shape", "solid_mask", "spacing_m", "step_count", "step_s")  # solver.rs:46-62
OPTIONAL_KEYS = ("topology_map", "ports", "origin_m", "field_registration", "geometry_source")  # solver.rs:64-65
FACES = ("xmin", "xmax", "ymin", "ymax", "zmin", "zmax")


def _err(msg):
    return ContractError(msg)


def _is_int(v):
    return isinstance(v, int) and not isinstance(v, bool)


def _real_scalar(v):
    if isinstance(v, bool) or not isinstance(v, (int, float)):
        return None
    v = float(v)
    return v if math.isfinite(v) else None


def _positive_scalar(p, key):
    v = _real_scalar(p.get(key))
    if v is None or not v > 0.0:
        raise _err(f"{key} must be a finite positive scalar")
    return v


def _shape_of(value):
    if not isinstance(value, (list, tuple)) or len(value) != 3 or not all(_is_int(v) and v >= 2 for v in value):
        return None
    return tuple(int(v) for v in value)


def _bool_array(value, shape):
    try:
        a = np.asarray(value)
    except (ValueError, TypeError):
        return None
    if a.shape != shape or a.dtype != np.bool_:
        return None
    return a


@dataclass
class TopologyMap:
    """design.rs:37-80 in metres."""

    filter_radius_m: float = 0.0
    projection_beta: float = 0.0
    projection_eta: float = 0.5


@dataclass
class LbmProblem:
    """The normalised problem: physical inputs, masks and the lattice-unit quantities the
    library takes."""

    shape: tuple
    spacing_m: float
    step_s: float
    step_count: int
    density_kg_m3: float
    kinematic_viscosity_m2_s: float
    acceleration_m_s2: tuple
    periodic_axes: tuple
    solid_mask: np.ndarray
    design_region: np.ndarray
    fixed_design: np.ndarray
    drag_max_per_s: float
    drag_shape: float
    mach_limit: float
    history_byte_budget: int
    tau: float
    topology_map: TopologyMap
    ports: list = field(default_factory=list)
    origin_m: tuple = (0.0, 0.0, 0.0)

    @property
    def acceleration_lattice(self):
        """solver.rs:575-578."""
        return tuple(v * self.step_s * self.step_s / self.spacing_m for v in self.acceleration_m_s2)

    @property
    def alpha_max(self):
        """step_s * drag_max_per_s: the resistance's lattice-unit scale (design.rs:282-286)."""
        return self.step_s * self.drag_max_per_s

    def resistance(self, phi):
        """LbmProblem::resistance (solver.rs:580-584, design.rs:282-286)."""
        return design.resistance(phi, self.alpha_max, self.drag_shape)

    def resistance_derivative(self, phi):
        """LbmProblem::resistance_derivative (solver.rs:586-590, design.rs:288-294)."""
        return design.resistance_derivative(phi, self.alpha_max, self.drag_shape)

    def vector_scales(self):
        """LbmProblem::vector_scales (solver.rs:862-871): (velocity, density, step, cell volume)."""
        return (self.spacing_m / self.step_s, self.density_kg_m3, self.step_s, self.spacing_m ** 3)

    def lattice_config(self, precision="f64", **kw):
        """The library's configuration of the reference's stationary step (D3Q19 BGK-Guo,
        halfway bounce-back, these ports) with admission at the problem's Mach limit and the
        history budget as the checkpoint budget."""
        return LatticeConfig(shape=self.shape, periodic=self.periodic_axes, lattice=19, collision="bgk_guo",
                             precision=precision, omega=1.0 / self.tau, tau=self.tau,
                             accel=self.acceleration_lattice,
                             ports=list(self.ports), solid=self.solid_mask if self.solid_mask.any() else None,
                             checkpoint_budget_bytes=int(self.history_byte_budget), mach_limit=self.mach_limit, **kw)

    def design_map(self):
        """DesignMap::new (design.rs:108-135) on the device: radius filter_radius_m /
        spacing_m cells over the editable cells (design region, not solid)."""
        t = self.topology_map
        return DesignMap(self.shape, self.periodic_axes, radius=t.filter_radius_m / self.spacing_m,
                         beta=t.projection_beta, eta=t.projection_eta, solid=self.solid_mask,
                         region=self.design_region, fixed=self.fixed_design)


def _topology_map(value, spacing):
    """TopologyMap::normalize (design.rs:51-80)."""
    keys = ("filter_radius_m", "projection_beta", "projection_eta")
    if not isinstance(value, dict) or len(value) != 3 or not all(k in value for k in keys):
        raise _err("topology_map requires filter_radius_m, projection_beta and projection_eta")
    out = {}
    for k, v in value.items():
        r = _real_scalar(v)
        if r is None:
            raise _err(f"topology_map.{k} must be finite real scalar")
        out[k] = r
    t = TopologyMap(out["filter_radius_m"], out["projection_beta"], out["projection_eta"])
    if not (0.0 <= t.filter_radius_m <= 4.0 * spacing):
        raise _err("filter_radius_m must lie in [0,4*spacing_m]")
    if not (0.0 <= t.projection_beta <= 32.0) or not (0.0 < t.projection_eta < 1.0):
        raise _err("Require beta in [0,32] and eta in (0,1)")
    return t


def _ports(value, shape, periodic, solid, design_region, fixed, accel, step_s, spacing, mach_limit, density):
    """normalize_ports (ports.rs:107-270), mapped onto the library's stationary port kinds."""
    if not isinstance(value, list):
        raise _err("ports must be a list")
    accelerated = any(a != 0.0 for a in accel)
    used = np.zeros(shape, bool)
    ids = set()
    out, sources = [], []
    for port in value:
        kind = port.get("kind") if isinstance(port, dict) else None
        if kind not in ("velocity", "pressure"):
            raise _err("Port kind must be velocity or pressure")
        pressure = kind == "pressure"
        required = ["id", "kind", "face", "mask", "velocity_m_s"] + (["gauge_pressure_Pa"] if pressure else [])
        keys = [k for k in port if k != "reconstruction"]
        pid = port.get("id")
        if (len(keys) != len(required) or not all(r in keys for r in required) or not isinstance(pid, str)
                or not pid.strip() or pid in ids):
            raise _err("Unique port id and exact port fields required")
        rec = port.get("reconstruction", "legacy_moment")
        if rec not in ("legacy_moment", "regularized_known_stress", "neighbor_non_equilibrium"):
            raise _err("Port reconstruction must be legacy_moment, regularized_known_stress or "
                       "neighbor_non_equilibrium")
        if rec == "neighbor_non_equilibrium" and (not pressure or accelerated):
            raise _err("neighbor_non_equilibrium currently requires a pressure port and zero acceleration")
        if rec == "regularized_known_stress" and accelerated:
            raise _err("regularized_known_stress ports require zero acceleration; forcing-consistent stress "
                       "reconstruction is not implemented")
        ids.add(pid)
        face = port.get("face")
        if face not in FACES:
            raise _err("Port must identify one Cartesian face")
        axis = "xyz".index(face[0])
        low = face.endswith("min")
        if periodic[axis]:
            raise _err("A periodic face cannot be a port")
        mask = _bool_array(port.get("mask"), shape)
        if mask is None or not mask.any():
            raise _err("Nonempty shape-matched boolean port mask required")
        face_index = 0 if low else shape[axis] - 1
        coord = np.indices(shape)
        if np.any(mask & ((coord[axis] != face_index) | used)):
            raise _err("Port mask must lie on its face and cannot overlap another port")
        for a in range(3):
            if a != axis and not periodic[a] and np.any(mask & ((coord[a] == 0) | (coord[a] == shape[a] - 1))):
                raise _err("Port corners/edges with other closed faces require an unsupported corner closure")
        if np.any(mask & (solid | design_region | (fixed != 1.0))):
            raise _err("Port cells must be protected, unobstructed, fixed-design fluid")
        if rec == "neighbor_non_equilibrium":
            step = 1 if low else -1
            if not (0 <= face_index + step < shape[axis]):
                raise _err("Neighbor pressure reconstruction requires an inward fluid plane")
            nb = np.roll(mask, step, axis=axis)
            if np.any(nb & (solid | design_region | (fixed != 1.0))):
                raise _err("Neighbor pressure reconstruction requires protected, unobstructed, fixed-design "
                           "inward fluid cells")
        vel = port.get("velocity_m_s")
        vel_ok = isinstance(vel, (list, tuple)) and len(vel) == 3 and all(_real_scalar(v) is not None for v in vel)
        if not vel_ok:
            raise _err("Port velocity requires three finite components")
        u = [float(v) * step_s / spacing for v in vel]
        if pressure and float(vel[axis]) != 0.0:
            raise _err("Pressure port velocity specifies tangential components only; normal component must be zero")
        if math.sqrt(sum(x * x for x in u)) * math.sqrt(3.0) > mach_limit:
            raise _err("Authored port velocity exceeds Mach limit")
        rho = None
        if pressure:
            gauge = _real_scalar(port.get("gauge_pressure_Pa"))
            if gauge is None:
                raise _err("Finite scalar gauge pressure required")
            unit = spacing / step_s
            rho = 1.0 + gauge * 3.0 / (density * (unit * unit))
            if not (0.9 <= rho <= 1.1):
                raise _err("Pressure port density must remain within 10% of reference")
        used |= mask
        tang = [u[a] for a in range(3) if a != axis]
        rule = "regularized" if rec == "regularized_known_stress" else "moment"
        if rec == "neighbor_non_equilibrium":
            out.append(Port(face, "pressure_neighbor", (rho, tang[0], tang[1]), mask=mask))
        elif pressure:
            out.append(Port(face, f"pressure_{rule}", (rho, tang[0], tang[1]), mask=mask))
        else:
            out.append(Port(face, f"velocity_{rule}", tuple(u), mask=mask))
        sources.append(np.roll(mask, 1 if low else -1, axis=axis) if rec == "neighbor_non_equilibrium" else None)
    for src in sources:  # ports.rs:264-270
        if src is not None and np.any(src & used):
            raise _err("Neighbor pressure reconstruction cannot sample another port")
    return out


def normalize(problem) -> LbmProblem:
    """LbmProblem::normalize (solver.rs:414-550)."""
    if not isinstance(problem, dict) or sorted(k for k in problem if k not in OPTIONAL_KEYS) != sorted(PROBLEM_KEYS):
        raise _err("LBM requires exactly these problem keys: [" + ", ".join(f"'{k}'" for k in PROBLEM_KEYS) + "]")
    p = problem
    shape = _shape_of(p["shape"])
    if shape is None:
        raise _err("LBM shape must contain three integers >= 2")
    spacing = _positive_scalar(p, "spacing_m")
    step_s = _positive_scalar(p, "step_s")
    density = _positive_scalar(p, "density_kg_m3")
    nu = _positive_scalar(p, "kinematic_viscosity_m2_s")
    drag_max = _positive_scalar(p, "drag_max_per_s")
    drag_shape = _positive_scalar(p, "drag_shape")
    mach_limit = _positive_scalar(p, "mach_limit")
    if mach_limit > 0.2:
        raise _err("LBM mach_limit must be <= 0.2; this is not a compressible gas solver")
    count = p["step_count"]
    if not (_is_int(count) and 1 <= count <= 1_000_000):
        raise _err("step_count must be an integer in [1,1000000]")
    cells = shape[0] * shape[1] * shape[2]
    required = (count + 1) * cells * 19 * 8
    budget = p["history_byte_budget"]
    if not (_is_int(budget) and budget >= required):
        raise _err("Population history exceeds explicit byte budget; AD working memory is additional")
    if "geometry_source" in p:
        raise _err("geometry_source: the implicit-geometry pipeline stays on the host; pass solid_mask, "
                   "design_region and fixed_design")
    masks = []
    for key in ("solid_mask", "design_region"):
        a = _bool_array(p[key], shape)
        if a is None:
            raise _err(f"{key} must be a shape-matched boolean array")
        masks.append(a)
    solid, design_region = masks
    try:
        fixed = np.asarray(p["fixed_design"])
    except (ValueError, TypeError):
        fixed = None
    if (fixed is None or fixed.shape != shape or fixed.dtype.kind not in "iuf" or not np.all(np.isfinite(fixed))
            or np.any((fixed < 0.0) | (fixed > 1.0))):
        raise _err("fixed_design must be finite and in [0,1]")
    fixed = fixed.astype(np.float64)
    if solid.all() or np.any(solid & design_region):
        raise _err("Need fluid cells; fixed solid cells cannot be designable")
    acc = p["acceleration_m_s2"]
    if not (isinstance(acc, (list, tuple)) and len(acc) == 3 and all(_real_scalar(v) is not None for v in acc)):
        raise _err("acceleration_m_s2 requires three finite components")
    acc = tuple(float(v) for v in acc)
    per = p["periodic_axes"]
    if not (isinstance(per, (list, tuple)) and len(per) == 3 and all(isinstance(v, bool) for v in per)):
        raise _err("periodic_axes requires three booleans; false means closed no-slip walls")
    per = tuple(per)
    tau = 0.5 + 3.0 * nu * step_s / (spacing * spacing)
    if not (0.51 <= tau <= 2.0):
        raise _err("Choose spacing/time step such that 0.51 <= relaxation time <= 2")
    tmap = _topology_map(p.get("topology_map", {"filter_radius_m": 0.0, "projection_beta": 0.0,
                                                "projection_eta": 0.5}), spacing)
    ports = _ports(p.get("ports", []), shape, per, solid, design_region, fixed, acc, step_s, spacing, mach_limit,
                   density)
    origin = p.get("origin_m", [0.0, 0.0, 0.0])
    if not (isinstance(origin, (list, tuple)) and len(origin) == 3 and all(_real_scalar(v) is not None for v in origin)):
        raise _err("origin_m requires three finite components")
    if "field_registration" in p:
        raise _err("field_registration: registered field checks stay with the caller's host pipeline")
    return LbmProblem(shape=shape, spacing_m=spacing, step_s=step_s, step_count=int(count), density_kg_m3=density,
                      kinematic_viscosity_m2_s=nu, acceleration_m_s2=acc, periodic_axes=per, solid_mask=solid,
                      design_region=design_region, fixed_design=fixed, drag_max_per_s=drag_max,
                      drag_shape=drag_shape, mach_limit=mach_limit, history_byte_budget=int(budget), tau=tau,
                      topology_map=tmap, ports=ports, origin_m=tuple(float(v) for v in origin))
