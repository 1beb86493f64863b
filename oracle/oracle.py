"""CPU FP64 oracle for the differentiable LBM hot path -- TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may import this
module, and only as the checker (or the timed CPU baseline).  The product
(``implexity_amd``) never imports it.

It wraps ``oracle/lbm_oracle.c`` (a C restatement of the reference's arithmetic, built
by ``oracle/Makefile`` with -ffp-contract=off) and adds the reference's driver logic:

* port enumeration       -- moving/boundary.rs:318-380 (Topology::with_symmetry ports)
* ``interval``           -- solver.rs:604-614 (collide -> stream -> ports)
* ``interval_vjp``       -- solver.rs:642-674
* ``history``            -- solver.rs:677-687 (all states kept)
* ``scan_adjoint``       -- crates/implexity-ad/src/scan.rs:140-246 (Every(k) policy)
* ``binomial_scan``      -- scan.rs:284-390 driven by oracle/revolve.py
* ``fd_vs_ad_probe``     -- crates/implexity-ad/src/gradcheck.rs:50-123

Parity pins (each checked in tests/test_oracle_*.py):

* tables                 -- closed forms (weights sum to 1, isotropy, opposite pairs)
* stream table           -- brute-force geometric enumeration of halfway bounce-back
* TRT/BGK collide        -- equilibrium moments, BGK == TRT at omega- == omega+,
                            exact mass conservation, body-force Poiseuille closed form
* MRT collide            -- MRT with default rates == TRT (options.rs:56-59)
* cumulant collide       -- conservation of rho and (with force) of momentum update,
                            omega = bulk = 1 Gaussian (Wick) closure
* every VJP              -- central finite differences and dot-product identities
* NEE ports              -- equilibrium reproduction on uniform flow, FD of the VJP
* objectives             -- FD of their gradients, closed form on uniform states
* revolve                -- optimal_forward_steps identity and exhaustive replay
* Smagorinsky / WALE     -- Hou et al.'s and Nicoud & Ducros' closed forms, pure-shear
                            invariance, linear-field stencil, central differences, the
                            dot-product identity of the step (tests/test_oracle_les.py)
* occupancy coupling     -- Noble & Torczynski's weighting B = d t / (1 - d + t), the
                            saturation's C1 joins, full-solid cells moving with the solid
                            (mass kept, momentum rho (u_s + a/2)), the momentum exchange
                            as the change of momentum, d -> 0 continuity, finite
                            differences of the cell and step transposes
                            (tests/test_oracle_psm.py)
* sponge layers          -- the ramp's Hermite conditions, mirror symmetry of opposite
                            faces, the far-field equilibrium as a fixed point, the blend's
                            mass and momentum, travelling-wave invariance, the step's
                            dot-product identity and finite differences
                            (tests/test_oracle_sponge.py)
"""
from __future__ import annotations

import ctypes
import math
import os
import subprocess
from dataclasses import dataclass, field

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "liblbm_oracle.so")

KIND = {"bgk": 0, "trt": 1, "mrt": 2, "cumulant": 3, "bgk_guo": 4, "regularized": 5}
FACES = {"xmin": (0, 1), "xmax": (0, -1), "ymin": (1, 1), "ymax": (1, -1), "zmin": (2, 1), "zmax": (2, -1)}
NONE_U32 = 0xFFFFFFFF

_lib = None


def build() -> None:
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        _lib = ctypes.CDLL(LIB_PATH)
        d = ctypes.c_double
        pd = ctypes.POINTER(ctypes.c_double)
        _lib.orc_odd_rate.restype = d
        _lib.orc_odd_rate.argtypes = [d, d]
        _lib.orc_odd_rate_derivative.restype = d
        _lib.orc_odd_rate_derivative.argtypes = [d, d]
        _lib.orc_obj_dissipation.restype = d
        _lib.orc_obj_pressure_drop.restype = d
        _lib.orc_psm_saturate.restype = d
        _lib.orc_obj_solid_force.restype = d
        _lib.orc_psm_saturate.argtypes = [d, d]
        _lib.orc_stationary_stress_basis.restype = d
        del pd
    return _lib


class _Problem(ctypes.Structure):
    _fields_ = [
        ("Q", ctypes.c_int32), ("kind", ctypes.c_int32),
        ("nx", ctypes.c_int32), ("ny", ctypes.c_int32), ("nz", ctypes.c_int32),
        ("periodic", ctypes.c_int32 * 3),
        ("omega", ctypes.c_double), ("magic", ctypes.c_double),
        ("bulk_rate", ctypes.c_double), ("even_rate", ctypes.c_double),
        ("accel", ctypes.c_double * 3),
        ("n_port_cells", ctypes.c_int32), ("pad_", ctypes.c_int32),
        ("port_cell", ctypes.c_void_p), ("port_nb", ctypes.c_void_p),
        ("port_kind", ctypes.c_void_p), ("port_value", ctypes.c_void_p),
        ("solid", ctypes.c_void_p),
        ("mrt_rows", ctypes.c_void_p), ("mrt_norms", ctypes.c_void_p), ("mrt_groups", ctypes.c_void_p),
        ("les_constant", ctypes.c_double), ("les_floor", ctypes.c_double),
        ("wale_constant", ctypes.c_double), ("wale_floor", ctypes.c_double),
        ("n_sponge_cells", ctypes.c_int32), ("n_sponges", ctypes.c_int32),
        ("sponge_cell", ctypes.c_void_p), ("sponge_sigma", ctypes.c_void_p), ("sponge_off", ctypes.c_void_p),
        ("sponge_part", ctypes.c_void_p), ("sponge_weight", ctypes.c_void_p), ("sponge_phase", ctypes.c_void_p),
        ("sponge_target", ctypes.c_void_p),
        ("psm_law", ctypes.c_int32), ("pad2_", ctypes.c_int32), ("psm_kappa", ctypes.c_double),
        ("occ_m", ctypes.c_void_p), ("occ_m_bar", ctypes.c_void_p),
        ("psm_c_alpha", ctypes.c_double), ("psm_drag_shape", ctypes.c_double),
        ("occ_d1", ctypes.c_void_p), ("occ_m1", ctypes.c_void_p), ("theta", ctypes.c_double),
        ("occ_d1_bar", ctypes.c_void_p), ("occ_m1_bar", ctypes.c_void_p), ("tau", ctypes.c_double),
        ("sub_g", ctypes.c_void_p), ("sub_gbar", ctypes.c_void_p),
    ]

PSM_LAW = {"brinkman": -1, "psm": 0, "psm_superposition": 1, "brinkman_occupancy": 2}


def _ptr(a):
    return ctypes.c_void_p(a.ctypes.data) if a is not None else None


def lattice(Q: int):
    c = np.zeros((Q, 3), np.int32)
    w = np.zeros(Q, np.float64)
    opp = np.zeros(Q, np.int32)
    lib().orc_lattice(ctypes.c_int(Q), _ptr(c), _ptr(w), _ptr(opp))
    return c, w, opp


def mrt_basis(Q: int):
    rows = np.zeros((Q, Q), np.float64)
    norms = np.zeros(Q, np.float64)
    groups = np.zeros(Q, np.int32)
    rc = lib().orc_mrt_basis(ctypes.c_int(Q), _ptr(rows), _ptr(norms), _ptr(groups))
    if rc != 0:
        raise RuntimeError("MRT basis does not span the lattice")
    return rows, norms, groups


def odd_rate(omega: float, magic: float) -> float:
    return lib().orc_odd_rate(omega, magic)


def psm_beta(d: float, t: float, kappa: float):
    """beta of psm.rs:123-140: (B/d, d(B/d)/dd, d(B/d)/dt)."""
    out = np.zeros(3)
    lib().orc_psm_beta(ctypes.c_double(d), ctypes.c_double(t), ctypes.c_double(kappa), _ptr(out))
    return tuple(out)


def psm_resistance(d: float, c_alpha: float, q: float, kappa: float):
    """resistance of psm.rs:143-160: (alpha, gamma, [dalpha/dd, dalpha/dc, dgamma/dd, dgamma/dc])."""
    out = np.zeros(6)
    lib().orc_psm_resistance(ctypes.c_double(d), ctypes.c_double(c_alpha), ctypes.c_double(q),
                             ctypes.c_double(kappa), _ptr(out))
    return out[0], out[1], out[2:]


def psm_saturate(d: float, kappa: float) -> float:
    return lib().orc_psm_saturate(ctypes.c_double(d), ctypes.c_double(kappa))


# Port kinds: the moving solver's NEE rule (boundary.rs:445-506) and the stationary
# solver's rules (ports.rs): "pressure_neighbor" = reconstruct_neighbor (:484-497),
# "velocity_moment" / "pressure_moment" = reconstruct (:343-406, "legacy_moment"),
# "velocity_regularized" / "pressure_regularized" = reconstruct_regularized (:455-482,
# "regularized_known_stress").
PORT_KINDS = ("velocity", "pressure", "pressure_neighbor", "velocity_moment", "pressure_moment",
              "velocity_regularized", "pressure_regularized")
_SELF_SOURCE = ("velocity_moment", "pressure_moment", "velocity_regularized", "pressure_regularized")
ST_PORT_BASE = 8  # row kind of a stationary rule: ST_PORT_BASE + 6 * rule + face
_REGULARIZED_ACCEL = ("regularized_known_stress ports require zero acceleration; forcing-consistent stress "
                      "reconstruction is not implemented")  # ports.rs:154-158


@dataclass
class Port:
    face: str
    kind: str  # one of PORT_KINDS
    # velocity kinds: velocity (3,) lattice units; "pressure": (rho,); the stationary
    # pressure kinds: (rho, u_t0, u_t1), the authored tangential velocity in ascending axes
    value: tuple
    profile: np.ndarray | None = None  # one factor per port cell, ascending cell order
    mask: np.ndarray | None = None  # lattice-shaped boolean subset of the face, or all of it


@dataclass
class Wave:
    """Travelling wave on a sponge's far-field velocity (moving/sponge.rs:22-29), in
    lattice units: amplitude (velocity), unit direction, speed (cells per step),
    frequency (cycles per step), phase (rad), ramp (steps), origin (cells)."""
    amplitude: tuple
    direction: tuple
    speed: float
    frequency: float
    phase: float = 0.0
    ramp: float = 0.0
    origin: tuple = (0.0, 0.0, 0.0)


@dataclass
class Sponge:
    """SpongeSpec (moving/sponge.rs:12-20) in lattice units: the far-field state is a
    density and a velocity instead of a pressure in Pa and a velocity in m/s."""
    face: str
    thickness: int
    strength: float
    density: float = 1.0
    velocity: tuple = (0.0, 0.0, 0.0)
    wave: Wave | None = None


def sponge_ramp(s: float) -> float:
    """ramp of moving/sponge.rs:72-81 (the quintic smoothstep, in its operation order)."""
    if s <= 0.0:
        return 0.0
    if s >= 1.0:
        return 1.0
    return s * s * s * (10.0 - 15.0 * s + 6.0 * s * s)


def signal_factors(ramp_steps: float, frequency: float, phase: float, t: float):
    """Signal::factors (moving/boundary.rs:110-125) in steps: (ramp, wave)."""
    if ramp_steps > 0.0 and t < ramp_steps:
        s = t / ramp_steps
        s = 0.0 if s < 0.0 else s
        ramp = s * s * s * (10.0 - s * 15.0 + s * s * 6.0)
    else:
        ramp = 1.0
    wave = 0.0 if frequency == 0.0 else math.sin(t * (2.0 * math.pi * frequency) + phase)
    return ramp, wave


def _wave_ok(w: Wave) -> bool:
    """TravellingWave::validate (moving/sponge.rs:33-60); the signal's frequency and phase
    are the wave's own."""
    vals = list(w.amplitude) + list(w.direction) + list(w.origin)
    norm = math.sqrt(sum(v * v for v in w.direction))
    return (all(math.isfinite(v) for v in vals) and math.isfinite(w.speed) and w.speed > 0.0
            and math.isfinite(w.frequency) and w.frequency > 0.0 and math.isfinite(w.phase)
            and math.isfinite(w.ramp) and w.ramp >= 0.0 and abs(norm - 1.0) <= 1e-9)


def sponge_field(shape, periodic, sponges, solid=None):
    """SpongeField::new (moving/sponge.rs:112-181) on the stationary grid with unit
    spacing and the lattice origin at 0 (cell centres at ijk + 1/2).  Returns
    (cells, sigma, off, part, weight, phase) of the fluid cells some sponge covers, in
    ascending cell order; raises ValueError with the reference's contract messages."""
    nx, ny, nz = shape
    n = nx * ny * nz
    empty = (np.zeros(0, np.int32), np.zeros(0), np.zeros(1, np.int32), np.zeros(0, np.int32), np.zeros(0),
             np.zeros(0))
    if not sponges:
        return empty
    solid = np.zeros(n, bool) if solid is None else np.asarray(solid, bool).reshape(-1)
    sigma = np.zeros(n)
    parts = {}
    idx = np.arange(n)
    co = np.stack([idx // (ny * nz), (idx // nz) % ny, idx % nz])
    for l, sp in enumerate(sponges):
        if sp.wave is not None and not _wave_ok(sp.wave):
            raise ValueError("sponge wave needs finite amplitude and origin, a unit direction, a positive speed "
                             "and a positive frequency")
        axis, inward = FACES[sp.face]
        if periodic[axis]:
            raise ValueError(f"sponge on face {sp.face} lies on a periodic axis")
        if sp.thickness == 0 or 2 * sp.thickness > shape[axis]:
            raise ValueError(f"sponge on face {sp.face} must be between 1 cell and half the lattice thick")
        if not (math.isfinite(sp.strength) and 0.0 < sp.strength < 1.0):
            raise ValueError("sponge strength must lie in (0, 1)")
        if not math.isfinite(sp.density) or not all(math.isfinite(v) for v in sp.velocity):
            raise ValueError("sponge far-field state must be finite")
        length = float(sp.thickness)
        c = co[axis].astype(np.float64)
        depth = (float(sp.thickness) - c) if inward > 0 else (c - float(shape[axis] - sp.thickness) + 1.0)
        for x in np.nonzero(depth > 0.0)[0]:
            sv = sp.strength * sponge_ramp((float(depth[x]) - 0.5) / length)
            if sv <= 0.0:
                continue
            sigma[x] += sv
            if sp.wave is None:
                ph = (1.0, 0.0)
            else:
                w = sp.wave
                pos = [float(co[a, x]) + 0.5 for a in range(3)]
                s = 0.0
                for a in range(3):
                    s += (pos[a] - w.origin[a]) * w.direction[a]
                theta = 2.0 * math.pi * w.frequency * s / w.speed
                ph = (math.cos(theta), math.sin(theta))
            parts.setdefault(int(x), []).append((l, sv, ph))
    if np.any(sigma >= 1.0):
        raise ValueError("summed sponge strength must stay below one")
    cells = [x for x in sorted(parts) if not solid[x]]
    off = [0]
    part, weight, phase = [], [], []
    for x in cells:
        for l, sv, ph in parts[x]:
            part.append(l)
            weight.append(sv)
            phase.extend(ph)
        off.append(len(part))
    return (np.asarray(cells, np.int32), np.asarray([sigma[x] for x in cells], np.float64),
            np.asarray(off, np.int32), np.asarray(part, np.int32), np.asarray(weight, np.float64),
            np.asarray(phase, np.float64))


def sponge_targets(sponges, t: float):
    """Far-field targets of every sponge at step time t (field.rs:1015-1036 with c = 1):
    rows [rho, mean[3], sin[3], cos[3]]."""
    out = np.zeros((max(1, len(sponges)), 10))
    for l, sp in enumerate(sponges):
        out[l, 0] = sp.density
        out[l, 1:4] = sp.velocity
        if sp.wave is not None:
            w = sp.wave
            ramp, _ = signal_factors(w.ramp, w.frequency, w.phase, t)
            arg = t * (2.0 * math.pi * w.frequency) + w.phase
            sw, cw = ramp * math.sin(arg), ramp * math.cos(arg)
            out[l, 4:7] = [sw * a for a in w.amplitude]
            out[l, 7:10] = [cw * a for a in w.amplitude]
    return out


@dataclass
class Problem:
    """Plain description of one lattice problem in lattice units."""

    shape: tuple
    periodic: tuple = (False, False, False)
    Q: int = 19
    kind: str = "trt"
    omega: float = 1.0
    tau: float = 0.0  # BGK-Guo relaxation time as given; 0: 1/omega
    magic: float = 3.0 / 16.0
    bulk_rate: float = float("nan")
    even_rate: float = float("nan")
    accel: tuple = (0.0, 0.0, 0.0)
    solid: np.ndarray | None = None  # bool, shape
    ports: list = field(default_factory=list)
    # Smagorinsky LES on BGK/TRT (psm.rs:45-48,161-197): constant (0 = off), norm floor
    les_constant: float = 0.0
    les_floor: float = 0.0
    # WALE closure on BGK/TRT (moving/wale.rs): constant (0 = off), denominator floor
    wale_constant: float = 0.0
    wale_floor: float = 1e-12
    # Sponge layers (moving/sponge.rs) blended after collision
    sponges: list = field(default_factory=list)
    # Occupancy coupling (moving/psm.rs:198-267): "brinkman" (the design is alpha), or
    # "psm" / "psm_superposition" / "brinkman_occupancy" (the design is the occupancy d;
    # solid momentum m per cell; the last with alpha = resistance(d; c_alpha, drag_shape))
    coupling: str = "brinkman"
    psm_kappa: float = 0.0
    psm_c_alpha: float = 0.0
    psm_drag_shape: float = 0.1
    # second occupancy endpoint (d1, m1 (cells x 3), t0, t1): the step reading state t
    # collides with d = d0 (1 - theta) + d1 theta, theta = clamp((t - t0) / (t1 - t0), 0, 1)
    occupancy_end: tuple | None = None
    solid_momentum: np.ndarray | None = None  # (cells, 3)


class Oracle:
    """FP64 CPU reference implementation (AoS state, index x*Q+q)."""

    def __init__(self, prob: Problem):
        self.prob = prob
        nx, ny, nz = prob.shape
        self.n = nx * ny * nz
        self.Q = prob.Q
        self.solid = (np.zeros(self.n, np.uint8) if prob.solid is None
                      else np.ascontiguousarray(np.asarray(prob.solid, bool).reshape(-1), np.uint8))
        cells, nbs, kinds, values = self.port_cells()
        self.port_cell = np.asarray(cells, np.int32)
        self.port_nb = np.asarray(nbs, np.int32)
        self.port_kind = np.asarray(kinds, np.int32)
        self.port_value = np.ascontiguousarray(np.asarray(values, np.float64).reshape(-1, 3))
        self._mrt = mrt_basis(self.Q) if prob.kind in ("mrt", "regularized") else None
        p = _Problem()
        p.Q, p.kind = self.Q, KIND[prob.kind]
        p.nx, p.ny, p.nz = nx, ny, nz
        for a in range(3):
            p.periodic[a] = int(bool(prob.periodic[a]))
            p.accel[a] = float(prob.accel[a])
        p.omega, p.magic = float(prob.omega), float(prob.magic)
        p.tau = float(prob.tau)
        p.bulk_rate, p.even_rate = float(prob.bulk_rate), float(prob.even_rate)
        p.n_port_cells = len(self.port_cell)
        p.port_cell = _ptr(self.port_cell)
        p.port_nb = _ptr(self.port_nb)
        p.port_kind = _ptr(self.port_kind)
        p.port_value = _ptr(self.port_value)
        p.solid = _ptr(self.solid)
        p.les_constant, p.les_floor = float(prob.les_constant), float(prob.les_floor)
        p.wale_constant, p.wale_floor = float(prob.wale_constant), float(prob.wale_floor)
        if self._mrt is not None:
            p.mrt_rows, p.mrt_norms, p.mrt_groups = (_ptr(self._mrt[0]), _ptr(self._mrt[1]), _ptr(self._mrt[2]))
        self._sp = sponge_field(prob.shape, prob.periodic, prob.sponges, self.solid.astype(bool))
        cells, sigma, off, part, weight, phase = self._sp
        p.n_sponge_cells, p.n_sponges = len(cells), len(prob.sponges)
        p.sponge_cell, p.sponge_sigma, p.sponge_off = _ptr(cells), _ptr(sigma), _ptr(off)
        p.sponge_part, p.sponge_weight, p.sponge_phase = _ptr(part), _ptr(weight), _ptr(phase)
        self._target = sponge_targets(prob.sponges, 0.0)
        p.sponge_target = _ptr(self._target)
        p.psm_law, p.psm_kappa = PSM_LAW[prob.coupling], float(prob.psm_kappa)
        p.psm_c_alpha, p.psm_drag_shape = float(prob.psm_c_alpha), float(prob.psm_drag_shape)
        self.psm = prob.coupling != "brinkman"
        self._m = None
        self.m_bar = np.zeros((self.n, 3)) if self.psm else None
        if self.psm:
            self._m = (np.zeros((self.n, 3)) if prob.solid_momentum is None
                       else np.ascontiguousarray(np.asarray(prob.solid_momentum, np.float64).reshape(self.n, 3)))
            p.occ_m, p.occ_m_bar = _ptr(self._m), _ptr(self.m_bar)
        self.d1_bar = self.m1_bar = None
        if self.psm and prob.occupancy_end is not None:
            d1, m1, _, _ = prob.occupancy_end
            self._d1 = np.ascontiguousarray(d1, np.float64).reshape(-1)
            self._m1 = (np.zeros((self.n, 3)) if m1 is None
                        else np.ascontiguousarray(np.asarray(m1, np.float64).reshape(self.n, 3)))
            self.d1_bar, self.m1_bar = np.zeros(self.n), np.zeros((self.n, 3))
            p.occ_d1, p.occ_m1 = _ptr(self._d1), _ptr(self._m1)
            p.occ_d1_bar, p.occ_m1_bar = _ptr(self.d1_bar), _ptr(self.m1_bar)
        self._p = p
        self._pp = ctypes.byref(p)
        self.source = np.zeros(self.n * self.Q, np.uint32)
        lib().orc_stream_table(self._pp, _ptr(self.source))

    # --- grid helpers (d3q19.rs:70-128) -------------------------------------------
    def coords(self, x):
        nx, ny, nz = self.prob.shape
        return x // (ny * nz), (x // nz) % ny, x % nz

    def index(self, i, j, k):
        nx, ny, nz = self.prob.shape
        return (i * ny + j) * nz + k

    def port_cells(self):
        """Port cells in the order of moving/boundary.rs:318-380: ports in list order,
        cells of a face in ascending cell index among non-wall cells (those of the
        port's mask when it has one, ports.rs:200); the inward neighbour is one cell
        along the face normal, or the cell itself for the moment rules, whose source is
        the port cell (Port::sources, ports.rs:74-82); the value is the port value times
        the cell's profile factor (field.rs:1242-1248) and, for the moment rules, less
        half the lattice acceleration (v0, ports.rs:352)."""
        nx, ny, nz = self.prob.shape
        shape = (nx, ny, nz)
        solid = (np.zeros(self.n, bool) if self.prob.solid is None
                 else np.asarray(self.prob.solid, bool).reshape(-1))
        cells, nbs, kinds, values = [], [], [], []
        for port in self.prob.ports:
            axis, inward = FACES[port.face]
            if self.prob.periodic[axis]:
                raise ValueError("port on periodic axis")
            layer = 0 if inward > 0 else shape[axis] - 1
            idx = np.arange(self.n)
            co = np.stack(self.coords(idx))
            on_face = (co[axis] == layer) & ~solid
            if port.mask is not None:
                on_face &= np.asarray(port.mask, bool).reshape(-1)
            face = idx[on_face]
            prof = port.profile
            if prof is not None and len(prof) != len(face):
                raise ValueError("profile needs one factor per port cell")
            fi = 2 * axis + (0 if inward > 0 else 1)
            moment = port.kind in _SELF_SOURCE
            acc = [float(a) for a in self.prob.accel]
            for r, x in enumerate(face):
                c = list(self.coords(int(x)))
                c[axis] += inward
                nb = int(x) if moment else self.index(*c)
                if solid[nb]:
                    raise ValueError("inward neighbour of a port cell is a wall")
                factor = 1.0 if prof is None else float(prof[r])
                cells.append(int(x))
                nbs.append(nb)
                if port.kind == "velocity":
                    kinds.append(0)
                    values.append([float(v) * factor for v in port.value])
                elif port.kind == "pressure":
                    kinds.append(1)
                    values.append([float(port.value[0]), 0.0, 0.0])
                elif port.kind == "pressure_neighbor":
                    kinds.append(ST_PORT_BASE + fi)
                    values.append([float(v) for v in port.value])
                elif port.kind in ("velocity_moment", "velocity_regularized"):
                    if port.kind == "velocity_regularized" and any(a != 0.0 for a in acc):
                        raise ValueError(_REGULARIZED_ACCEL)
                    kinds.append(ST_PORT_BASE + (6 if port.kind == "velocity_moment" else 18) + fi)
                    values.append([float(port.value[d]) * factor - 0.5 * acc[d] for d in range(3)])
                elif port.kind in ("pressure_moment", "pressure_regularized"):
                    if port.kind == "pressure_regularized" and any(a != 0.0 for a in acc):
                        raise ValueError(_REGULARIZED_ACCEL)
                    kinds.append(ST_PORT_BASE + (12 if port.kind == "pressure_moment" else 24) + fi)
                    tang = [d for d in range(3) if d != axis]
                    values.append([float(port.value[0])] + [float(port.value[1 + t]) - 0.5 * acc[d]
                                                            for t, d in enumerate(tang)])
                else:
                    raise ValueError(f"unknown port kind {port.kind!r}")
        return cells, nbs, kinds, values

    # --- cell level ---------------------------------------------------------------------
    def cell_collide(self, f, alpha):
        f = np.ascontiguousarray(f, np.float64)
        out = np.zeros(self.Q)
        lib().orc_cell_collide(self._pp, _ptr(f), ctypes.c_double(alpha), _ptr(out))
        return out

    def cell_vjp(self, f, alpha, g):
        f = np.ascontiguousarray(f, np.float64)
        g = np.ascontiguousarray(g, np.float64)
        fb = np.zeros(self.Q)
        ab = ctypes.c_double(0.0)
        lib().orc_cell_vjp(self._pp, _ptr(f), ctypes.c_double(alpha), _ptr(g), _ptr(fb), ctypes.byref(ab))
        return fb, ab.value

    def port_populations(self, nb, kind, value):
        nb = np.ascontiguousarray(nb, np.float64)
        v = np.ascontiguousarray(np.asarray(value, np.float64).reshape(3))
        out = np.zeros(self.Q)
        lib().orc_port_populations(ctypes.c_int(self.Q), _ptr(nb), ctypes.c_int(kind), _ptr(v), _ptr(out))
        return out

    @staticmethod
    def stationary_port(src, kind, value):
        """ports.rs rule of row kind `kind` on the source cell's 19 populations."""
        src = np.ascontiguousarray(src, np.float64)
        v = np.ascontiguousarray(np.asarray(value, np.float64).reshape(3))
        out = np.zeros(19)
        lib().orc_stationary_port(_ptr(src), ctypes.c_int(kind), _ptr(v), _ptr(out))
        return out

    @staticmethod
    def stationary_port_jacobian(src, kind, value):
        """cell_jacobian (ports.rs:532-557): (19, 19), [r, k] = d out_r / d src_k."""
        src = np.ascontiguousarray(src, np.float64)
        v = np.ascontiguousarray(np.asarray(value, np.float64).reshape(3))
        jac = np.zeros(19 * 19)
        lib().orc_stationary_port_jacobian(_ptr(src), ctypes.c_int(kind), _ptr(v), _ptr(jac))
        return jac.reshape(19, 19)

    @staticmethod
    def stationary_stress_operator(face):
        """known_stress_operator (ports.rs:409-440) of a face: (known q (14,), operator
        (6, 14)), the operator the exact rational weighted least-squares left inverse
        rounded once per entry."""
        op = np.zeros(84)
        known = np.zeros(19, np.int32)
        nk = lib().orc_stationary_stress_operator(ctypes.c_int(face), _ptr(op),
                                                  known.ctypes.data_as(ctypes.POINTER(ctypes.c_int)))
        if nk != 14:
            raise ValueError("stress fit of a face without rank 6")
        return known[:nk].copy(), op.reshape(6, 14)

    @staticmethod
    def stationary_stress_basis():
        """basis (ports.rs:413-417): (19, 6), 4.5 w_q (Q00, Q11, Q22, 2 Q01, 2 Q02, 2 Q12)."""
        return np.array([[lib().orc_stationary_stress_basis(ctypes.c_int(q), ctypes.c_int(r)) for r in range(6)]
                         for q in range(19)])

    @staticmethod
    def stationary_right_inverse(face):
        """right_inverse (ports.rs:325-341) of a face's unknown set: (n_unknown, 3)."""
        ri = np.zeros(27)
        nu = lib().orc_stationary_right_inverse(ctypes.c_int(face), _ptr(ri))
        if nu < 0:
            raise ValueError("right inverse of a non-diagonal moment matrix")
        return ri[:3 * nu].reshape(nu, 3)

    def port_populations_vjp(self, nb, kind, value, g):
        nb = np.ascontiguousarray(nb, np.float64)
        v = np.ascontiguousarray(np.asarray(value, np.float64).reshape(3))
        g = np.ascontiguousarray(g, np.float64)
        out = np.zeros(self.Q)
        lib().orc_port_populations_vjp(ctypes.c_int(self.Q), _ptr(nb), ctypes.c_int(kind), _ptr(v), _ptr(g),
                                       _ptr(out))
        return out

    # --- occupancy coupling (moving/psm.rs:96-152, 228-266, 331-510) -------------------------
    def psm_cell(self, f, d, m):
        post, dp = np.zeros(self.Q), np.zeros(3)
        lib().orc_psm_cell(self._pp, _ptr(np.ascontiguousarray(f, np.float64)), ctypes.c_double(d),
                           _ptr(np.ascontiguousarray(m, np.float64)), _ptr(post), _ptr(dp))
        return post, dp

    def psm_cell_vjp(self, f, d, m, g, dp_bar=(0.0, 0.0, 0.0)):
        fb, mb = np.zeros(self.Q), np.zeros(3)
        db = ctypes.c_double(0.0)
        lib().orc_psm_cell_vjp(self._pp, _ptr(np.ascontiguousarray(f, np.float64)), ctypes.c_double(d),
                               _ptr(np.ascontiguousarray(m, np.float64)), _ptr(np.ascontiguousarray(g, np.float64)),
                               _ptr(np.ascontiguousarray(dp_bar, np.float64)), _ptr(fb), ctypes.byref(db), _ptr(mb))
        return fb, db.value, mb

    # --- WALE closure (moving/wale.rs) -------------------------------------------------------
    def wale_rates(self, state):
        rate = np.zeros(self.n)
        lib().orc_wale_rates(self._pp, _ptr(np.ascontiguousarray(state, np.float64)), _ptr(rate))
        return rate

    def wale_omega(self, g, tau0):
        out = ctypes.c_double(0.0)
        lib().orc_wale_omega(self._pp, _ptr(np.ascontiguousarray(g, np.float64).reshape(9)), ctypes.c_double(tau0),
                             ctypes.byref(out))
        return out.value

    def wale_omega_vjp(self, g, tau0, w):
        gb = np.zeros(9)
        lib().orc_wale_omega_vjp(self._pp, _ptr(np.ascontiguousarray(g, np.float64).reshape(9)),
                                 ctypes.c_double(tau0), ctypes.c_double(w), _ptr(gb))
        return gb.reshape(3, 3)

    # --- problem level --------------------------------------------------------------------
    def initial(self):
        f = np.zeros(self.n * self.Q)
        lib().orc_initial(self._pp, _ptr(f))
        return f

    def theta(self, t):
        """Interpolation weight of the second occupancy endpoint for the step reading state t."""
        if self.prob.occupancy_end is None:
            return 0.0
        _, _, t0, t1 = self.prob.occupancy_end
        return 0.0 if t <= t0 else (1.0 if t >= t1 else (t - t0) / (t1 - t0))

    def set_time(self, t):
        """Sponge targets and occupancy weight of the step that reads state t."""
        self._target[:] = sponge_targets(self.prob.sponges, float(t))
        self._p.theta = float(self.theta(t))

    def interval(self, alpha, before, want_post=False, t=0):
        self.set_time(t)
        alpha = np.ascontiguousarray(alpha, np.float64)
        before = np.ascontiguousarray(before, np.float64)
        post = np.empty_like(before)
        after = np.empty_like(before)
        ex = np.zeros(max(1, len(self.port_cell)))
        lib().orc_interval(self._pp, _ptr(self.source), _ptr(alpha), _ptr(before), _ptr(post), _ptr(after),
                           _ptr(ex))
        if want_post:
            return after, ex[: len(self.port_cell)], post
        return after

    def interval_vjp(self, alpha, before, after_bar, after=None, t=0):
        alpha = np.ascontiguousarray(alpha, np.float64)
        before = np.ascontiguousarray(before, np.float64)
        after_bar = np.ascontiguousarray(after_bar, np.float64)
        if after is None:
            after = self.interval(alpha, before, t=t)
        self.set_time(t)
        work = np.empty_like(before)
        f_bar = np.empty_like(before)
        a_bar = np.empty(self.n)
        lib().orc_interval_vjp(self._pp, _ptr(self.source), _ptr(alpha), _ptr(before), _ptr(after),
                               _ptr(after_bar), _ptr(work), _ptr(f_bar), _ptr(a_bar))
        return f_bar, a_bar

    # --- subcycled moving field (moving/field.rs:1870-1919, 2055-2392) -----------------------
    def _substep(self, alpha, before, t, theta):
        self.set_time(t)
        self._p.theta = float(theta)
        after = np.empty_like(before)
        post = np.empty_like(before)
        ex = np.zeros(max(1, len(self.port_cell)))
        lib().orc_interval(self._pp, _ptr(self.source), _ptr(alpha), _ptr(before), _ptr(post), _ptr(after), _ptr(ex))
        return after

    def subcycle(self, d0, f_prev, substeps, t0=0):
        """SubcycledField::subcycle on the occupancy coupling: `substeps` steps from f_prev,
        step k = 1..m blending the endpoints (d0 with the problem's solid momentum, and the
        problem's occupancy_end) at theta = (k - 1/2) / m (Field::params, field.rs:1049),
        with the impulse shares of Field::accumulate (:1499-1541).  Returns (states 0..m,
        g (cells, 2, 3))."""
        d0 = np.ascontiguousarray(d0, np.float64)
        g = np.zeros(self.n * 6)
        states = [np.ascontiguousarray(f_prev, np.float64).copy()]
        self._p.sub_g = _ptr(g)
        try:
            for k in range(1, substeps + 1):
                states.append(self._substep(d0, states[-1], t0 + k - 1, (k - 0.5) / substeps))
        finally:
            self._p.sub_g = None
        return states, g.reshape(self.n, 2, 3)

    def subcycle_adjoint(self, d0, states, substeps, state_bar, g_bar=None, t0=0):
        """Transpose of subcycle (Field::reverse, field.rs:2055-2300, reverse_substep
        :2482-2848) over its stored states: (previous_bar, d0_bar, m0_bar, d1_bar, m1_bar)."""
        d0 = np.ascontiguousarray(d0, np.float64)
        lam = np.ascontiguousarray(state_bar, np.float64).copy()
        gb = None if g_bar is None else np.ascontiguousarray(g_bar, np.float64).reshape(-1)
        d0b, d1b = np.zeros(self.n), np.zeros(self.n)
        m0b, m1b = np.zeros((self.n, 3)), np.zeros((self.n, 3))
        self._p.sub_gbar = None if gb is None else _ptr(gb)
        try:
            for k in range(substeps, 0, -1):
                theta = (k - 0.5) / substeps
                before = states[k - 1]
                after = self._substep(d0, before, t0 + k - 1, theta)
                work = np.empty_like(before)
                f_bar = np.empty_like(before)
                a_bar = np.empty(self.n)
                lib().orc_interval_vjp(self._pp, _ptr(self.source), _ptr(d0), _ptr(before), _ptr(after), _ptr(lam),
                                       _ptr(work), _ptr(f_bar), _ptr(a_bar))
                lam = f_bar
                d0b += a_bar
                m0b += self.m_bar
                if self.d1_bar is not None:
                    d1b += self.d1_bar
                    m1b += self.m1_bar
        finally:
            self._p.sub_gbar = None
        return lam, d0b, m0b, d1b, m1b

    def history(self, alpha, steps, f0=None):
        """solver.rs:677-687 -- every state kept (use small problems)."""
        states = [self.initial() if f0 is None else np.array(f0, np.float64)]
        for t in range(steps):
            states.append(self.interval(alpha, states[-1], t=t))
        return states

    def run(self, alpha, steps, f0=None, t0=0):
        f = self.initial() if f0 is None else np.array(f0, np.float64)
        for t in range(steps):
            f = self.interval(alpha, f, t=t0 + t)
        return f

    def macroscopic(self, alpha, state):
        rho = np.zeros(self.n)
        u = np.zeros(3 * self.n)
        lib().orc_macroscopic(self._pp, _ptr(np.ascontiguousarray(alpha, np.float64)),
                              _ptr(np.ascontiguousarray(state, np.float64)), _ptr(rho), _ptr(u))
        return rho, u.reshape(self.n, 3)

    # --- objectives -------------------------------------------------------------------------
    def dissipation(self, alpha, state):
        return lib().orc_obj_dissipation(self._pp, _ptr(np.ascontiguousarray(alpha, np.float64)),
                                         _ptr(np.ascontiguousarray(state, np.float64)))

    def dissipation_grad(self, alpha, state, weight=1.0):
        fb = np.zeros(self.n * self.Q)
        ab = np.zeros(self.n)
        lib().orc_obj_dissipation_grad(self._pp, _ptr(np.ascontiguousarray(alpha, np.float64)),
                                       _ptr(np.ascontiguousarray(state, np.float64)), ctypes.c_double(weight),
                                       _ptr(fb), _ptr(ab))
        return fb, ab

    def pressure_drop(self, state, axis, plane_a, plane_b):
        return lib().orc_obj_pressure_drop(self._pp, _ptr(np.ascontiguousarray(state, np.float64)),
                                           ctypes.c_int(axis), ctypes.c_int(plane_a), ctypes.c_int(plane_b))

    def vector(self, phi, alpha, state, scales):
        """solver.rs:874-894: [mean u (3), kinetic, porous dissipation, volume]."""
        out = np.zeros(6)
        sc = np.asarray(scales, np.float64)
        lib().orc_vector(self._pp, _ptr(None if phi is None else np.ascontiguousarray(phi, np.float64)),
                         _ptr(np.ascontiguousarray(alpha, np.float64)), _ptr(np.ascontiguousarray(state, np.float64)),
                         _ptr(sc), _ptr(out))
        return out

    def vector_vjp(self, phi, alpha, state, scales, w):
        """solver.rs:896-947 (forward-mode gradient of w . vector): (f_bar, alpha_bar, phi_bar)."""
        fb = np.zeros(self.n * self.Q)
        ab = np.zeros(self.n)
        pb = np.zeros(self.n)
        sc = np.asarray(scales, np.float64)
        ww = np.asarray(w, np.float64)
        lib().orc_vector_vjp(self._pp, _ptr(None if phi is None else np.ascontiguousarray(phi, np.float64)),
                             _ptr(np.ascontiguousarray(alpha, np.float64)),
                             _ptr(np.ascontiguousarray(state, np.float64)), _ptr(sc), _ptr(ww), _ptr(fb), _ptr(ab),
                             _ptr(pb))
        return fb, ab, pb

    def solid_force(self, alpha, state, w):
        """w . (momentum exchanged with the solid by the collision of state)."""
        return lib().orc_obj_solid_force(self._pp, _ptr(np.ascontiguousarray(alpha, np.float64)),
                                         _ptr(np.ascontiguousarray(state, np.float64)),
                                         _ptr(np.ascontiguousarray(w, np.float64)))

    def solid_force_grad(self, alpha, state, w, weight=1.0):
        """(f_bar, d_bar, m_bar) of weight * solid_force."""
        fb, db, mb = np.zeros(self.n * self.Q), np.zeros(self.n), np.zeros(3 * self.n)
        lib().orc_obj_solid_force_grad(self._pp, _ptr(np.ascontiguousarray(alpha, np.float64)),
                                       _ptr(np.ascontiguousarray(state, np.float64)),
                                       _ptr(np.ascontiguousarray(w, np.float64)), ctypes.c_double(weight), _ptr(fb),
                                       _ptr(db), _ptr(mb))
        return fb, db, mb.reshape(self.n, 3)

    def pressure_drop_grad(self, axis, plane_a, plane_b, weight=1.0):
        fb = np.zeros(self.n * self.Q)
        lib().orc_obj_pressure_drop_grad(self._pp, ctypes.c_int(axis), ctypes.c_int(plane_a), ctypes.c_int(plane_b),
                                         ctypes.c_double(weight), _ptr(fb))
        return fb


# ----------------------------------------------------------------------------------------
# Objectives in the ScanObjective form (scan.rs:96-107): stage(t, x_{t+1}) for t in
# [0, n), terminal(x_n).  ``window`` lists the state indices s = t+1 that carry a stage
# cost; the cost of a state is weight * phi(state).
# ----------------------------------------------------------------------------------------
@dataclass
class Objective:
    kind: str  # "dissipation" | "pressure_drop" | "vector" | "solid_force" (w = vector_weights[:3])
    window: tuple = (1, 1)  # first and last state index (inclusive) carrying stage cost
    weight: float = 1.0
    terminal_weight: float = 0.0
    axis: int = 0
    plane_a: int = 1
    plane_b: int = -2
    vector_weights: tuple = (0.0,) * 6  # "vector": phi = w . vector(state) (phi input not used)
    scales: tuple = (1.0, 1.0, 1.0, 1.0)  # velocity, density, step_s, cell volume

    def planes(self, shape):
        n = shape[self.axis]
        return self.plane_a % n, self.plane_b % n

    def value(self, orc: Oracle, alpha, state):
        if self.kind == "solid_force":
            return orc.solid_force(alpha, state, self.vector_weights[:3])
        if self.kind == "dissipation":
            return orc.dissipation(alpha, state)
        if self.kind == "vector":
            return float(np.dot(self.vector_weights, orc.vector(None, alpha, state, self.scales)))
        a, b = self.planes(orc.prob.shape)
        return orc.pressure_drop(state, self.axis, a, b)

    def grad(self, orc: Oracle, alpha, state, weight):
        """(f_bar, alpha_bar); the solid force also leaves its m_bar in self.last_m_bar."""
        if self.kind == "solid_force":
            fb, db, self.last_m_bar = orc.solid_force_grad(alpha, state, self.vector_weights[:3], weight)
            return fb, db
        if self.kind == "dissipation":
            return orc.dissipation_grad(alpha, state, weight)
        if self.kind == "vector":
            fb, ab, _ = orc.vector_vjp(None, alpha, state, self.scales, np.asarray(self.vector_weights) * weight)
            return fb, ab
        a, b = self.planes(orc.prob.shape)
        return orc.pressure_drop_grad(self.axis, a, b, weight), np.zeros(orc.n)

    def stage_weight(self, s):
        lo, hi = self.window
        return self.weight if lo <= s <= hi else 0.0


def scan_adjoint(orc: Oracle, alpha, steps, objective: Objective, every=None, f0=None, m_bar=None, end_bar=None):
    """scan_adjoint with Checkpointing::Every(k) / All (scan.rs:140-246) over the oracle
    step.  Returns (J, alpha_bar, lambda_0, final_state); with occupancy coupling the
    solid-momentum cotangent is summed into m_bar (cells x 3) when given, and with a second
    occupancy endpoint its cotangents into end_bar = (d1_bar, m1_bar)."""
    alpha = np.asarray(alpha, np.float64)
    seg = steps if every is None else every
    x = orc.initial() if f0 is None else np.array(f0, np.float64)
    J = 0.0
    checkpoints = [(0, x.copy())]
    for t in range(steps):
        x = orc.interval(alpha, x, t=t)
        w = objective.stage_weight(t + 1)
        if w:
            orc.set_time(t + 1)
            J += w * objective.value(orc, alpha, x)
        if (t + 1) % seg == 0 and t + 1 < steps:
            checkpoints.append((t + 1, x.copy()))
    final = x
    abar = np.zeros(orc.n)
    lam = np.zeros(orc.n * orc.Q)
    if objective.terminal_weight:
        J += objective.terminal_weight * objective.value(orc, alpha, final)
        g, ga = objective.grad(orc, alpha, final, objective.terminal_weight)
        lam += g
        abar += ga
    after = final
    while checkpoints:
        start, xs = checkpoints.pop()
        end = min(start + seg, steps)
        states = [xs]
        for t in range(start, end - 1):
            states.append(orc.interval(alpha, states[-1], t=t))
        for t in range(end - 1, start - 1, -1):
            x_next = states[t + 1 - start] if t + 1 < end else after
            w = objective.stage_weight(t + 1)
            if w:
                orc.set_time(t + 1)  # the objective on state t + 1 reads its occupancy
                g, ga = objective.grad(orc, alpha, x_next, w)
                lam = lam + g
                th = orc.theta(t + 1) if orc.d1_bar is not None else 0.0
                abar += ga * (1.0 - th) if objective.kind == "solid_force" and orc.d1_bar is not None else ga
                if objective.kind == "solid_force":
                    if m_bar is not None:
                        m_bar += objective.last_m_bar * (1.0 - th)
                    if end_bar is not None and orc.d1_bar is not None:
                        end_bar[0][...] += ga * th
                        end_bar[1][...] += objective.last_m_bar * th
            fb, ab = orc.interval_vjp(alpha, states[t - start], lam, t=t)
            abar += ab
            if m_bar is not None:
                m_bar += orc.m_bar
            if end_bar is not None and orc.d1_bar is not None:
                end_bar[0][...] += orc.d1_bar
                end_bar[1][...] += orc.m1_bar
            lam = fb
        after = states[0]
    return J, abar, lam, final


def trajectory(orc: Oracle, alpha, steps, f0=None):
    """history (solver.rs:676-687) with the port mass exchange of every step, as the
    reference's History carries it (boundary_mass_changes_lattice): (states, exchanges)."""
    states = [orc.initial() if f0 is None else np.array(f0, np.float64)]
    changes = []
    for t in range(steps):
        after, ex, _ = orc.interval(alpha, states[-1], want_post=True, t=t)
        states.append(after)
        changes.append(float(np.sum(ex)))
    return states, changes


def admit(orc: Oracle, alpha, states, changes, mach_limit, drift_limit=1e-6):
    """LbmProblem::admit (solver.rs:707-793) over a kept trajectory.  Returns
    (refusal or None, max Mach, relative mass drift, min fluid population); the refusal is
    the reference's message category: "populations" (a non-finite value anywhere, or a
    fluid population <= 0, :716-728) or "rejected" (Mach above mach_limit or drift above
    drift_limit, :762-771).  Mach is sqrt(3) times the largest Brinkman velocity |u| over
    the fluid cells of every state (:730-745); the drift is max_t |M_t - (M_0 + sum of the
    port exchanges before t)| / M_0 with M_t the fluid mass (:746-761)."""
    Q = orc.Q
    fluid = orc.solid == 0
    minimum = np.inf
    for st in states:
        if not np.all(np.isfinite(st)):
            return "populations", np.nan, np.nan, np.nan
        minimum = min(minimum, float(st.reshape(-1, Q)[fluid].min()))
    mach, mass = 0.0, []
    for st in states:
        rho, u = orc.macroscopic(alpha, st)
        mach = max(mach, float(np.sqrt((u[fluid] ** 2).sum(axis=1)).max()))
        mass.append(float(rho[fluid].sum()))
    mach *= np.sqrt(3.0)
    err, cum = 0.0, 0.0
    for t, m in enumerate(mass):
        if t > 0 and changes:
            cum += changes[t - 1]
        err = max(err, abs(m - (mass[0] + cum)))
    drift = err / mass[0]
    if minimum <= 0.0:
        return "populations", mach, drift, minimum
    if mach > mach_limit or drift > drift_limit:
        return "rejected", mach, drift, minimum
    return None, mach, drift, minimum


def objective_value(orc: Oracle, alpha, steps, objective: Objective, f0=None):
    """Forward-only objective (for finite differences)."""
    x = orc.initial() if f0 is None else np.array(f0, np.float64)
    J = 0.0
    for t in range(steps):
        x = orc.interval(alpha, x, t=t)
        w = objective.stage_weight(t + 1)
        if w:
            orc.set_time(t + 1)
            J += w * objective.value(orc, alpha, x)
    if objective.terminal_weight:
        J += objective.terminal_weight * objective.value(orc, alpha, x)
    return J


def fd_vs_ad_probe(f, x, index, value, ad, step=None):
    """gradcheck.rs:50-123: central difference with h = step*max(|x_i|, 1) and
    rel_err = |ad/s - fd/s|, s = max(|ad|, |fd|)."""
    step = np.finfo(float).eps ** (1.0 / 3.0) if step is None else step
    x = np.asarray(x, np.float64)
    h = step * max(abs(x[index]), 1.0)
    xp = x.copy()
    xp[index] += h
    xm = x.copy()
    xm[index] -= h
    fd = (f(xp) - f(xm)) / (2.0 * h)
    scale = max(abs(ad), abs(fd))
    rel = abs(ad / scale - fd / scale) if scale > 0 else 0.0
    return fd, rel
