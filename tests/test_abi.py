"""The C ABI library, without a GPU: it loads, exports every symbol the header
declares, builds revolve schedules on the host, and validates problems before any
device call."""
from __future__ import annotations

import ctypes
import os

import numpy as np
import pytest

from implexity_amd import lbm
from oracle import revolve as rev

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "implexity_lbm.h")


def test_library_exports_every_declared_symbol():
    names = lbm.declared_symbols(HEADER)
    assert len(names) >= 20
    lib = lbm.lib()
    for name in names:
        assert hasattr(lib, name), name
    assert set(names) == set(lbm._SIGS)


def test_abi_version():
    assert lbm.lib().lbm_abi_version() == 6


@pytest.mark.parametrize("n,s", [(1, 1), (3, 1), (10, 2), (100, 10), (1000, 30), (8000, 156)])
def test_revolve_schedule_host(n, s):
    sch = lbm.BinomialSchedule(n, s)
    acts, rstart, fwd, _ = rev.binomial_schedule(n, s)
    assert sch.actions == acts and sch.reverse_start == rstart and sch.forward_steps == fwd


@pytest.mark.parametrize("n,k", [(1, 1), (10, 1), (10, 2), (7, 3), (24, 5), (100, 10), (3000, 100)])
def test_every_schedule_replays(n, k):
    """Checkpointing::Every(k) (scan.rs:109-246) as an action list: every reverse holds
    state k-1 at the cursor and k ahead, the recomputation is n + sum(len - 1) steps, and
    at most ceil(n/k) + max(k - 2, 0) slots are live."""
    sch = lbm.EverySchedule(n, k)
    evals = rev.replay_actions(n, sch.actions)
    nseg = -(-n // k)
    lens = [min(k, n - g * k) for g in range(nseg)]
    assert evals == sch.forward_steps == n + sum(length - 1 for length in lens)
    used = {a[2] for a in sch.actions if a[0] == "snapshot"}
    assert max(used) + 1 <= nseg + max(k - 2, 0)


def test_revolve_refuses_no_slots():
    with pytest.raises(lbm.LbmError):
        lbm.BinomialSchedule(10, 0)


def _create(**kw):
    cfg = lbm.LatticeConfig(shape=kw.pop("shape", (4, 4, 4)), **kw)
    return lbm.Lattice(cfg)


@pytest.mark.parametrize("kw,msg", [
    (dict(shape=(1, 4, 4)), "shape"),
    (dict(omega=2.5), "relaxation time"),
    (dict(collision="cumulant"), "D3Q27"),
    (dict(lattice=27, collision="bgk_guo"), "D3Q19"),
    (dict(magic=-1.0), "magic"),
    (dict(precision="bf16s", les_constant=0.1), "BF16"),
    (dict(precision="f32", adjoint_storage="f16"), "FP16 cotangents"),
    (dict(precision="f64", adjoint_storage="f32"), "FP64 storage"),
    (dict(periodic=(True, False, False), ports=[lbm.Port("xmin", "pressure", (1.0,))]), "periodic"),
    (dict(shape=(3, 4, 4), ports=[lbm.Port("xmin", "velocity", (0.01, 0.0, 0.0)), lbm.Port("xmax", "pressure", (1.0,))]),
     "share an inward neighbour"),
    (dict(thermal=True, thermal_kappa=0.0), "thermal_kappa"),
    (dict(thermal=True, thermal_kappa=0.02, thermal_ratio=-1.0), "thermal_ratio"),
    (dict(thermal=True, thermal_kappa=0.02, collision="bgk_guo"), "BGK-Guo"),
    (dict(thermal=True, thermal_kappa=0.02, precision="bf16s"), "FP64, FP32 and FP16"),
    (dict(thermal=True, thermal_kappa=0.02, solid=np.eye(4, dtype=bool)[None].repeat(4, 0)), "no solid cells"),
    (dict(les_constant=0.1, collision="mrt"), "Smagorinsky closure applies"),
    (dict(mach_limit=0.3), "mach_limit"),
    (dict(mach_limit=-0.1), "mach_limit"),
    (dict(mach_limit=0.1, drift_limit=-1.0), "drift_limit"),
    (dict(les_constant=-0.1), "non-negative"),
    (dict(wale_constant=0.5, collision="mrt"), "WALE closure applies"),
    (dict(wale_constant=0.5, les_constant=0.1), "choose one closure"),
    (dict(wale_constant=0.5, wale_floor=0.0), "denominator_floor"),
    (dict(wale_constant=0.5, precision="f16s", adjoint_storage="f16"), "FP32 cotangents"),
    (dict(sponges=[lbm.Sponge("xmin", 2, 0.5)], periodic=(True, False, False)), "periodic axis"),
    (dict(sponges=[lbm.Sponge("ymax", 3, 0.5)]), "half the lattice"),
    (dict(sponges=[lbm.Sponge("ymax", 0, 0.5)]), "half the lattice"),
    (dict(sponges=[lbm.Sponge("zmin", 1, 1.0)]), r"\(0, 1\)"),
    (dict(sponges=[lbm.Sponge("zmin", 1, 0.5, density=float("inf"))]), "far-field state must be finite"),
    (dict(sponges=[lbm.Sponge("zmin", 1, 0.5, wave=lbm.Wave((0.01, 0, 0), (1.0, 1.0, 0.0), 0.1, 0.01))]),
     "unit direction"),
    (dict(sponges=[lbm.Sponge("zmin", 1, 0.5, wave=lbm.Wave((0.01, 0, 0), (1.0, 0.0, 0.0), 0.1, -0.01))]),
     "positive frequency"),
    (dict(sponges=[lbm.Sponge("xmin", 2, 0.6), lbm.Sponge("xmin", 2, 0.6)]), "below one"),
    (dict(sponges=[lbm.Sponge("xmin", 1, 0.1)] * 9), "at most 8"),
    (dict(coupling="psm", collision="mrt"), "plain BGK and TRT"),
    (dict(coupling="psm_superposition", les_constant=0.1), "plain BGK and TRT"),
    (dict(coupling="psm", psm_kappa=1.0), "kappa"),
    (dict(coupling="psm", precision="bf16s"), "FP64, FP32 and FP16"),
    (dict(coupling="psm", thermal=True, thermal_kappa=0.02), "temperature field"),
    (dict(coupling="brinkman_occupancy", psm_c_alpha=-1.0), "c_alpha"),
    (dict(coupling="brinkman_occupancy", psm_c_alpha=0.1, psm_drag_shape=0.0), "drag shape"),
    (dict(lattice=27, ports=[lbm.Port("xmin", "velocity_moment", (0.01, 0, 0))]), "D3Q19"),
    (dict(ports=[lbm.Port("xmin", "velocity_moment", (0.01, 0, 0), mask=np.eye(4, dtype=bool)[None].repeat(4, 0))]),
     "lie on its face"),
    (dict(ports=[lbm.Port("xmin", "pressure_moment", (0.0, 0, 0))]), "density must be positive"),
    (dict(ports=[lbm.Port("xmin", "pressure_neighbor", (1.0, 0, 0), profile=np.ones(16))]), "profile"),
    (dict(ports=[lbm.Port("xmin", "velocity_moment", (float("nan"), 0, 0))]), "finite"),
    (dict(ports=[lbm.Port("xmin", "velocity_moment", (0.01, 0, 0), mask=np.zeros((4, 4, 4), bool))]), "no cells"),
    (dict(thermal=True, thermal_kappa=0.02, ports=[lbm.Port("xmin", "velocity_moment", (0.01, 0, 0))]),
     "temperature field"),
])
def test_contract_errors_before_device(kw, msg):
    """CaeError::Contract conditions (solver.rs:414-550, collision.rs:244-292) are
    refused with status 1 before anything touches the GPU."""
    with pytest.raises(lbm.ContractError, match=msg):
        _create(**kw)


def test_solid_mask_length_is_checked():
    with pytest.raises(lbm.ShapeError):
        _create(solid=np.zeros(5, bool))


@pytest.mark.parametrize("kw,msg", [
    (dict(radius=4.5), "radius"),
    (dict(radius=-1.0), "radius"),
    (dict(beta=40.0), "beta"),
    (dict(eta=1.0), "eta"),
])
def test_design_map_contract_errors_before_device(kw, msg):
    """TopologyMap::normalize (design.rs:51-80) conditions, refused before any device call."""
    with pytest.raises(lbm.ContractError, match=msg):
        lbm.DesignMap((4, 4, 4), **kw)


def test_design_map_mask_length_is_checked():
    with pytest.raises(lbm.ShapeError):
        lbm.DesignMap((4, 4, 4), solid=np.zeros(5, bool))


@pytest.mark.parametrize("kw,msg", [
    (dict(width=3), "half-width"),
    (dict(xi=[[0.5, 4.0, 4.0]]), "pushforward_support"),
    (dict(xi=[[4.0, float("nan"), 4.0]]), "not finite"),
    (dict(xi=[[4.0, 4.0, -3.0e9]], periodic=(True, True, True)), "too far"),
])
def test_pushforward_contract_errors_before_device(kw, msg):
    args = dict(xi=[[4.0, 4.0, 4.0]], width=1, periodic=(False, False, False))
    args.update(kw)
    with pytest.raises(lbm.ContractError, match=msg):
        lbm.pushforward((8, 8, 8), args["xi"], [1.0], periodic=args["periodic"], width=args["width"])


NO_FACES = ((False, False), (False, False), (False, False))


@pytest.mark.parametrize("shape,periodic,mirrored,xi,msg", [
    ((3, 8, 8), (True, False, False), NO_FACES, [4.0, 4.0, 4.0], "shorter than the push-forward kernel support"),
    ((8, 1, 8), (False, False, False), NO_FACES, [4.0, 0.0, 4.0], "must be periodic"),
    ((8, 8, 8), (True, False, False), ((True, False), (False, False), (False, False)), [4.0, 4.0, 4.0],
     "mirror face"),
    ((8, 8, 8), (False, False, False), ((True, False), (False, False), (False, False)), [-0.51, 4.0, 4.0],
     "beyond a symmetry face"),
    ((8, 8, 8), (False, False, False), ((True, False), (False, False), (False, False)), [6.2, 4.0, 4.0],
     "outside the lattice"),
])
def test_pushforward_frame_errors_before_device(shape, periodic, mirrored, xi, msg):
    with pytest.raises(lbm.ContractError, match=msg):
        lbm.pushforward(shape, [xi], [1.0], periodic=periodic, width=1, mirrored=mirrored)
