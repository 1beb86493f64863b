"""Benchmark: D3Q19 TRT forward+adjoint design gradient on one MI355X.

Workload (one "step" = one lattice time step of the design-gradient horizon): the
256^3 D3Q19 TRT Brinkman channel -- velocity inlet u_x = 0.03 with a parabolic duct
profile on x = 0, zero-gradient pressure outlet rho = 1 on x = nx-1, halfway
bounce-back walls on y and z, tau+ = 0.55, magic 3/16, a design of 64 random
Gaussian blobs (radius U(6,20)) smoothed and sigmoided, alpha = 0.1*0.1*(1-phi)/
(0.1+phi); objective = pressure drop between planes x = 1 and x = nx-2 averaged over
the last 500 states.  The gradient over K steps runs forward + adjoint with binomial
checkpointing inside a 200 GB snapshot budget, in FP64 (reference), FP32 and
FP16-shifted storage.  Synthetic data: the design is generated from a seed.

value = lattice updates of the horizon per second of the whole FP16-shifted gradient
(cells * K / wall time / 1e6, recomputation and snapshot copies included), with the
design and state 0 in HBM when the timed region starts and dJ/dalpha left in HBM when it
ends; the copy of dJ/dalpha to the host is timed separately (alpha_bar_download_ms).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

HBM_PEAK_GBS = 8000.0
# Variants: name -> (state storage, cotangent storage, state bytes/value, cotangent
# bytes/value, alpha bytes).  f16s keeps FP16 cotangents (scaled), f16s_l32 FP32 ones.
VARIANTS = {
    "f64": ("f64", "auto", 8, 8, 8),
    "f32": ("f32", "auto", 4, 4, 4),
    "f16s": ("f16s", "f16", 2, 2, 4),
    "f16s_l32": ("f16s", "f32", 2, 4, 4),
}


# HBM traffic per lattice update of the step kernels at 256^3 (D3Q19 TRT, no solids),
# from rocprofv3 --pmc FETCH_SIZE and WRITE_SIZE passes over `bench.py --mode kernels`,
# each counter scaled by its reading for the 1 GiB copy kernel of the same run
# (FETCH_SIZE x 2.0, WRITE_SIZE x 1.0); the passes and the scaling are recorded in
# profiles/pmc_traffic_d3q19_trt_256cubed.txt.
PMC_TRAFFIC_BYTES_PER_LU = {
    ("forward", "f16s"): 81.7,
    ("adjoint", "f16s"): 127.8,
    ("forward", "f32"): 159.0,
    ("adjoint", "f32"): 243.2,
}


ABAR_RUN = 32  # reverse steps whose alpha_bar terms one FP32 partial gathers (FP32 arithmetic)


def abar_bytes(variant, steps=None):
    """alpha_bar bytes per lattice update of an adjoint step.  FP64: the FP64 read and
    write (16 B).  FP32 arithmetic: runs of ABAR_RUN reverse steps gather it in an FP32
    partial -- the first step of a run writes it (4 B), the middle ones read and write it
    (8 B), the last adds it and its own term to the FP64 alpha_bar (4 + 8 + 8 B); a run of
    one step updates alpha_bar directly (16 B).  steps None: a middle step (isolated
    launches); else the average over a sweep of that many reverse steps."""
    if variant == "f64":
        return 16.0
    if steps is None:
        return 8.0

    def run(n):
        return 16 if n == 1 else 4 + 8 * (n - 2) + 20
    full, rest = divmod(steps, ABAR_RUN)
    return (full * run(ABAR_RUN) + (run(rest) if rest else 0)) / steps


def algorithmic_bytes(which, variant, Q=19, steps=None):
    """Bytes per lattice update: the forward step reads and writes every population once
    and reads alpha; the adjoint step reads the state, reads and writes every cotangent,
    reads alpha and updates alpha_bar (abar_bytes)."""
    _, _, s, l, a = VARIANTS[variant]
    if which == "forward":
        return Q * s * 2 + a
    return Q * s + Q * l * 2 + a + abar_bytes(variant, steps)


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def build_problem(shape, precision, budget, seed=3):
    from implexity_amd import design

    cfg = _cfg_only(shape, precision, budget)
    phi = design.blob_design(shape, 64, seed, radius=(6.0, 20.0))
    alpha = design.resistance(phi, 0.1, 0.1).reshape(-1)
    return cfg, alpha


def objective(shape, steps):
    from implexity_amd.lbm import Objective

    window = min(500, steps)
    return Objective(kind="pressure_drop", axis=0, plane_a=1, plane_b=shape[0] - 2, t_begin=steps - window + 1,
                     t_end=steps, weight=1.0 / window)


def run_precision(shape, precision, steps, warmup, budget, alpha_cfg, timing_launches=20):
    from implexity_amd.lbm import Lattice

    cfg, alpha = alpha_cfg(precision)
    lat = Lattice(cfg)
    try:
        lat.set_design(alpha)
        lat.reserve(steps)  # snapshot pool allocated outside the timed region
        ob_w = objective(shape, max(warmup, 2))
        if warmup:
            lat.initial_populations()
            lat.gradient(warmup, ob_w)
        lat.initial_populations()
        ob = objective(shape, steps)
        # Timed: the whole gradient with the design and state 0 resident in HBM; dJ/dalpha
        # stays in HBM (where the design map's transpose consumes it) and its copy to the
        # host is timed on its own below.
        t0 = time.perf_counter()
        J, _, _, st = lat.gradient(steps, ob, alpha_bar_on_device=True)
        wall = time.perf_counter() - t0
        t1 = time.perf_counter()
        ab = lat.alpha_bar()
        t_dl = time.perf_counter() - t1
        # per-kernel device time (HIP events on the library's stream)
        lat.initial_populations()
        t_fwd = lat.time_kernel("forward", timing_launches)
        lat.initial_populations()
        t_adj = lat.time_kernel("adjoint", timing_launches)
        return dict(J=J, alpha_bar=ab, wall=wall, download_s=t_dl, stats=st.as_dict(), t_fwd_ms=t_fwd, t_adj_ms=t_adj,
                    state_bytes=lat.state_bytes, run_fwd_ms=st.forward_kernel_avg_ms,
                    run_adj_ms=st.adjoint_kernel_avg_ms)
    finally:
        lat.close()


def cpu_threads():
    """Host threads of the CPU baseline: the cores this process may run on, within the
    OMP_NUM_THREADS share the environment grants (when it sets one)."""
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    env = os.environ.get("OMP_NUM_THREADS", "")
    if env.isdigit() and int(env) > 0:
        n = min(n, int(env))
    return n


def cpu_baseline(shape, window, alpha_cfg, threads, runs=3):
    """The oracle (C restatement of the reference CPU algorithm: AoS FP64, collide then
    table stream, all states kept, -O3 -fopenmp -ffp-contract=off) timed on a bounded
    window of the same workload: one warm-up run, then the median of `runs`.  The
    oracle's tables are built before the timed runs."""
    os.environ["OMP_NUM_THREADS"] = str(threads)
    from oracle import oracle as orc

    cfg, alpha = alpha_cfg("f64")
    prob = orc.Problem(shape=shape, periodic=(False, False, False), Q=19, kind="trt", omega=cfg.omega,
                       magic=cfg.magic, ports=[orc.Port(p.face, p.kind, p.value, p.profile) for p in cfg.ports])
    o = orc.Oracle(prob)
    ob = orc.Objective(kind="pressure_drop", window=(1, window), weight=1.0 / window, axis=0, plane_a=1,
                       plane_b=shape[0] - 2)
    times = []
    for r in range(runs + 1):
        t0 = time.perf_counter()
        orc.scan_adjoint(o, alpha, window, ob)
        if r:  # the first run warms up
            times.append(time.perf_counter() - t0)
    dt = float(np.median(times))
    return o.n * window / dt / 1e6, dt, times


# Algorithmic bytes per lattice update of the coupled kernels (D3Q19 flow, D3Q7
# temperature): the forward pass reads and writes the 19 flow and 7 temperature
# populations and reads alpha and phi; the thermal adjoint reads the flow and
# temperature state, alpha and phi, reads and writes the temperature cotangent, writes
# the velocity cotangent (3 values, read back by the flow adjoint) and updates the FP64
# phi_bar.
def thermal_bytes(which, precision):
    s, a = {"f64": (8, 8), "f32": (4, 4), "f16s": (2, 4)}[precision]
    l = 8 if precision == "f64" else 4
    if which == "forward":
        return 38 * s + 14 * s + 2 * a
    return 19 * s + 7 * s + 2 * a + 14 * l + 3 * a + 16


def thermofluid(args):
    """The 384 x 192 x 192 flow + temperature design gradient (6000 steps by default):
    D3Q19 TRT flow (tau+ = 0.55, inlet u_x = 0.02, density-1 outlet) coupled to the D3Q7
    temperature (inlet T = 0, zero-gradient outlet, source 1e-5 per step and 10x
    diffusivity in the solid phi = 1 - rho), J = mean solid T + 0.1 pressure drop over the
    last 500 steps.  Prints one JSON line (not the headline metric)."""
    from implexity_amd import design
    from implexity_amd.lbm import Lattice, LatticeConfig, Objective, Port

    shape = (384, 192, 192) if tuple(args.shape) == (256, 256, 256) else tuple(args.shape)
    steps = args.steps if args.steps != 8000 else 6000
    n = int(np.prod(shape))
    rho = design.blob_design(shape, 64, 4, radius=(6.0, 20.0)).reshape(-1)
    prof = design.parabolic_profile(shape[1], shape[2]).reshape(-1)
    ports = [Port("xmin", "velocity", (0.02, 0.0, 0.0), prof), Port("xmax", "pressure", (1.0,))]
    window = min(500, steps)
    ob = Objective(kind="thermal", axis=0, plane_a=1, plane_b=shape[0] - 2, t_begin=steps - window + 1,
                   t_end=steps, weight=1.0 / window, vector_weights=(1.0, 0.1, 0.0, 0.0, 0.0, 0.0))
    res = {}
    for prec in args.precisions.split(","):
        cfg = LatticeConfig(shape=shape, periodic=(False, False, False), lattice=19, collision="trt",
                            precision=prec, omega=1.0 / 0.55, ports=ports,
                            checkpoint_budget_bytes=int(args.budget_gb * 1e9), thermal=True, thermal_kappa=0.02,
                            thermal_ratio=10.0, thermal_source=1e-5, thermal_inlet=0.0, thermal_range=0.25)
        with Lattice(cfg) as lat:
            lat.set_design(design.resistance(rho))
            lat.set_thermal_design(1.0 - rho)
            lat.reserve(steps)
            lat.initial_populations()
            lat.gradient(min(steps, 20), Objective(kind="thermal", axis=0, plane_a=1, plane_b=shape[0] - 2,
                                                   t_begin=1, t_end=min(steps, 20), weight=1.0,
                                                   vector_weights=(1.0, 0.1, 0, 0, 0, 0)))
            lat.initial_populations()
            t0 = time.perf_counter()
            J, ab, _, st = lat.gradient(steps, ob)
            wall = time.perf_counter() - t0
            g = ab * design.resistance_derivative(rho) - lat.thermal_design_bar()
            lat.initial_populations()
            tf = lat.time_kernel("coupled_forward", args.launches)
            ta = lat.time_kernel("thermal_adjoint", args.launches)
        res[prec] = {"J": J, "wall_s": wall, "mlups_fwd_adj": n * steps / wall / 1e6, "grad": g,
                     "flow_forward_kernel_ms": st.forward_kernel_avg_ms,
                     "flow_adjoint_kernel_ms": st.adjoint_kernel_avg_ms, "forward_steps": st.forward_steps,
                     "snapshot_slots": st.snapshot_slots, "coupled_forward_ms": tf, "thermal_adjoint_ms": ta,
                     "coupled_forward_GBs": thermal_bytes("forward", prec) * n / (tf * 1e-3) / 1e9,
                     "thermal_adjoint_GBs": thermal_bytes("adjoint", prec) * n / (ta * 1e-3) / 1e9}
        log(f"[thermofluid] {prec}: wall {wall:.2f} s, J {J:.6e}, coupled forward {tf:.3f} ms, thermal adjoint "
            f"{ta:.3f} ms")
    out = {"workload": f"D3Q19 TRT + D3Q7 temperature {shape[0]}x{shape[1]}x{shape[2]}, {steps} steps",
           "per_precision": {p: {k: v for k, v in r.items() if k != "grad"} for p, r in res.items()}}
    if "f64" in res:
        g64 = res["f64"]["grad"]
        out["error_vs_f64"] = {
            p: {"grad_rel_l2": float(np.linalg.norm(r["grad"] - g64) / np.linalg.norm(g64)),
                "grad_cosine": float(np.dot(r["grad"], g64) / np.linalg.norm(r["grad"]) / np.linalg.norm(g64)),
                "J_rel": abs(r["J"] - res["f64"]["J"]) / abs(res["f64"]["J"])}
            for p, r in res.items() if p != "f64"}
    if "f32" in res and "f16s" in res:
        out["speedup_f16s_vs_f32"] = res["f32"]["wall_s"] / res["f16s"]["wall_s"]
    print(json.dumps(out))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=8000)
    ap.add_argument("--warmup", type=int, default=100)
    ap.add_argument("--shape", type=int, nargs=3, default=(256, 256, 256))
    ap.add_argument("--budget-gb", type=float, default=200.0)
    ap.add_argument("--precisions", default="f64,f32,f16s")
    ap.add_argument("--cpu-window", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--mode", default="gradient", choices=["gradient", "kernels", "thermofluid"])
    ap.add_argument("--ports", default="nee", choices=["nee", "moment", "regularized"],
                    help="the channel's ports: NEE (the benchmark), or the stationary moment / regularized rules")
    ap.add_argument("--kernel-precision", default="f32")
    ap.add_argument("--launches", type=int, default=20)
    ap.add_argument("--les", type=float, default=0.0,
                    help="kernels mode: Smagorinsky constant of the timed operator (0: plain TRT)")
    ap.add_argument("--wale", type=float, default=0.0,
                    help="kernels mode: WALE constant of the timed operator (0: plain TRT)")
    ap.add_argument("--coupling", default="brinkman",
                    choices=["brinkman", "psm", "psm_superposition", "brinkman_occupancy"],
                    help="kernels mode: fluid-solid coupling (PSM laws take the blob design's solid fraction as "
                         "occupancy)")
    ap.add_argument("--design-map", type=float, default=0.0,
                    help="kernels mode: also time the design map and its transpose with this filter radius "
                         "(cells, beta 8)")
    ap.add_argument("--pushforward", type=int, default=0,
                    help="kernels mode: also time the push-forward of this many body points (cubic, width 1) "
                         "and its transpose")
    ap.add_argument("--sponge", type=int, default=0,
                    help="kernels mode: thickness of sponge layers on the outlet and on one side wall (0: none); "
                         "also times the sponge pass and its transpose")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: --gpus {args.gpus} is not supported (the lattice step runs on one "
              "GPU; slab decomposition along x with a per-step halo exchange is designed, not built)",
              file=sys.stderr)
        sys.exit(2)
    shape = tuple(args.shape)
    budget = int(args.budget_gb * 1e9)
    cache = {}

    def alpha_cfg(precision):
        if "alpha" not in cache:
            cache["alpha"] = build_problem(shape, precision, budget)[1]
        return _cfg_only(shape, precision, budget, args.ports), cache["alpha"]

    if args.mode == "thermofluid":
        thermofluid(args)
        return

    if args.mode == "kernels":
        # Per-kernel timing only (used under rocprofv3): forward and adjoint launches.
        from implexity_amd.lbm import Lattice

        from implexity_amd.lbm import Sponge, Wave

        cfg, alpha = alpha_cfg(args.kernel_precision)
        cfg.les_constant = args.les
        cfg.wale_constant = args.wale
        if args.wale and cfg.adjoint_storage == "f16":
            cfg.adjoint_storage = "auto"  # WALE keeps FP32 cotangents
        design_in = alpha
        if args.coupling != "brinkman":
            from implexity_amd import design

            cfg.coupling, cfg.psm_kappa, cfg.psm_c_alpha, cfg.psm_drag_shape = args.coupling, 0.1, 0.1, 0.1
            design_in = 1.0 - design.blob_design(shape, 64, 3, radius=(6.0, 20.0)).reshape(-1)  # solid fraction
        if args.sponge:
            cfg.sponges = [Sponge("xmax", args.sponge, 0.5, density=1.0, velocity=(0.03, 0.0, 0.0)),
                           Sponge("ymin", args.sponge, 0.3, wave=Wave((0.003, 0.0, 0.0), (0.0, 1.0, 0.0), 0.1, 0.002))]
        sp = {}
        with Lattice(cfg) as lat:
            lat.set_design(design_in)
            lat.initial_populations()
            tf = lat.time_kernel("forward", args.launches)
            ta = lat.time_kernel("adjoint", args.launches)
            ts = lat.time_kernel("stream", args.launches)
            tfd = lat.time_kernel("forward_diagnostics", args.launches)
            if args.sponge:
                t_sp = lat.time_kernel("sponge", args.launches)
                t_spv = lat.time_kernel("sponge_transpose", args.launches)
                T = args.sponge
                nx, ny, nz = shape
                cells = 2 * T * ny * nz - T * T * nz  # the two layers share a T x T x nz corner
                parts = cells + T * T * nz
                store = {"f64": 8, "f32": 4, "f16s": 2, "bf16s": 2}[args.kernel_precision]
                lam = {"f64": 8, "f32": 4, "f16s": 2 if cfg.adjoint_storage in ("auto", "f16") else 4}.get(
                    args.kernel_precision, 4)
                meta = cells * (4 + 8 + 4) + parts * (4 + 8 + 16)
                sp = {"sponge_thickness": T, "sponge_cells": cells, "sponge_ms": t_sp, "sponge_transpose_ms": t_spv,
                      "sponge_GBs": (2 * 19 * store * cells + meta) / (t_sp * 1e-3) / 1e9,
                      "sponge_transpose_GBs": (2 * 19 * lam * cells + cells * (4 + 8 + 4)) / (t_spv * 1e-3) / 1e9}
        if args.design_map:
            from implexity_amd.lbm import DesignMap

            raw = np.random.default_rng(0).random(int(np.prod(shape)))
            with DesignMap(shape, (False, False, False), args.design_map, 8.0, 0.5) as dm:
                dm.map(raw)
                dm.vjp(raw, raw)
                t0 = time.perf_counter()
                for _ in range(args.launches):
                    phi = dm.map(raw)
                t1 = time.perf_counter()
                for _ in range(args.launches):
                    dm.vjp(raw, phi)
                t2 = time.perf_counter()
            sp["design_map_radius"] = args.design_map
            sp["design_map_ms_with_copies"] = (t1 - t0) * 1e3 / args.launches
            sp["design_map_vjp_ms_with_copies"] = (t2 - t1) * 1e3 / args.launches
        if args.pushforward:
            from implexity_amd.lbm import pushforward, pushforward_vjp

            rng = np.random.default_rng(0)
            npt = args.pushforward
            xi = 3.0 + (np.array(shape, float) - 7.0) * rng.random((npt, 3))
            pa, pv = rng.random(npt) * 1e-3, rng.standard_normal((npt, 3)) * 0.01
            n = int(np.prod(shape))
            db, mb = rng.standard_normal(n), rng.standard_normal((n, 3))
            pushforward(shape, xi, pa, pv)
            t0 = time.perf_counter()
            for _ in range(args.launches):
                pushforward(shape, xi, pa, pv)
            t1 = time.perf_counter()
            for _ in range(args.launches):
                pushforward_vjp(shape, xi, pa, pv, db, mb)
            t2 = time.perf_counter()
            sp["pushforward_points"] = npt
            sp["pushforward_ms_with_copies"] = (t1 - t0) * 1e3 / args.launches
            sp["pushforward_vjp_ms_with_copies"] = (t2 - t1) * 1e3 / args.launches
        from implexity_amd.lbm import copy_bandwidth

        copy_gbs, copy_ms = copy_bandwidth(1 << 30, args.launches)  # 1 GiB read + 1 GiB written
        n = int(np.prod(shape))
        out = {"precision": args.kernel_precision, "les_constant": args.les, "wale_constant": args.wale,
               "coupling": args.coupling, **sp, "forward_ms": tf, "adjoint_ms": ta,
               "stream_only_ms": ts, "forward_diagnostics_ms": tfd,
               "forward_GBs": algorithmic_bytes("forward", args.kernel_precision) * n / (tf * 1e-3) / 1e9,
               "adjoint_GBs": algorithmic_bytes("adjoint", args.kernel_precision) * n / (ta * 1e-3) / 1e9,
               "stream_only_GBs": algorithmic_bytes("forward", args.kernel_precision) * n / (ts * 1e-3) / 1e9,
               "copy_GBs": copy_gbs, "copy_ms": copy_ms}
        print(json.dumps(out))
        return

    n = int(np.prod(shape))
    res = {}
    for prec in args.precisions.split(","):
        log(f"[bench] {prec}: gradient over {args.steps} steps on {shape} (warmup {args.warmup})")
        res[prec] = run_precision(shape, prec, args.steps, args.warmup, budget, alpha_cfg)
        r = res[prec]
        log(f"[bench] {prec}: wall {r['wall']:.2f} s, J {r['J']:.6e}, fwd kernel {r['t_fwd_ms']:.3f} ms, "
            f"adj kernel {r['t_adj_ms']:.3f} ms, stats {r['stats']}")

    from implexity_amd.lbm import copy_bandwidth

    copy_gbs, _ = copy_bandwidth(1 << 31, 20)
    main_prec = "f16s" if "f16s" in res else list(res)[-1]
    r = res[main_prec]
    value = n * args.steps / r["wall"] / 1e6

    def kern(prec, which, live=True):
        """Average kernel time (ms) and algorithmic GB/s: live = HIP events around every
        64th step kernel inside the timed gradient; otherwise back-to-back launches."""
        if live:
            t = res[prec]["run_fwd_ms" if which == "forward" else "run_adj_ms"]
        else:
            t = res[prec]["t_fwd_ms" if which == "forward" else "t_adj_ms"]
        gbs = algorithmic_bytes(which, prec, steps=args.steps) * n / (t * 1e-3) / 1e9
        return t, gbs

    # dominant kernel of the timed run: forward (incl. recomputation) or adjoint
    st = r["stats"]
    fwd_total = st["forward_steps"] * r["run_fwd_ms"]
    adj_total = st["reverse_steps"] * r["run_adj_ms"]
    dom = "forward" if fwd_total >= adj_total else "adjoint"
    t_dom, gbs_dom = kern(main_prec, dom)
    traffic = PMC_TRAFFIC_BYTES_PER_LU.get((dom, main_prec))
    if traffic is not None and shape == (256, 256, 256):
        traffic = traffic * n
    else:
        traffic = None
    out = {
        "metric": "D3Q19 fwd+adjoint MLUPS, % HBM roofline; FP16-store vs FP32 at grad err<=1e-3",
        "value": value,
        "unit": "MLUPS (fwd+adj)",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": r["wall"] * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f16-shifted storage, f32 arithmetic, f64 sensitivities" if main_prec == "f16s" else main_prec,
        "data": "synthetic (seeded 64-blob Brinkman design)",
        "config": {"workload": f"D3Q19 TRT channel {shape[0]}x{shape[1]}x{shape[2]}, tau+=0.55, u_in=0.03, "
                               f"forward+adjoint over {args.steps} steps, binomial checkpointing in "
                               f"{args.budget_gb:g} GB"
                               + ("" if args.ports == "nee" else f", stationary {args.ports} ports"),
                   "precision": main_prec},
        "roofline": {"bound": "hbm", "achieved": gbs_dom, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": gbs_dom / HBM_PEAK_GBS, "traffic": traffic, "kernel": f"{dom} ({main_prec})",
                     "algorithmic_bytes_per_launch": algorithmic_bytes(dom, main_prec, steps=args.steps) * n,
                     "algorithmic_bytes_per_lu": algorithmic_bytes(dom, main_prec, steps=args.steps),
                     "avg_kernel_ms": t_dom,
                     "avg_kernel_ms_source": "HIP events on the library stream around every min(64, K/4)-th step kernel of the timed gradient",
                     "traffic_source": "rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE, copy-kernel scaled"},
    }
    per = {}
    for prec, rr in res.items():
        tf, gf = kern(prec, "forward")
        ta, ga = kern(prec, "adjoint")
        tf_iso, _ = kern(prec, "forward", live=False)
        ta_iso, _ = kern(prec, "adjoint", live=False)
        per[prec] = {"gradient_wall_s": rr["wall"], "mlups_fwd_adj": n * args.steps / rr["wall"] / 1e6,
                     "alpha_bar_download_ms": rr["download_s"] * 1e3,
                     "mlups_fwd_adj_with_download": n * args.steps / (rr["wall"] + rr["download_s"]) / 1e6,
                     "kernel_samples": rr["stats"]["kernel_samples"],
                     "min_population": rr["stats"]["min_population"],
                     "J": rr["J"], "forward_kernel_ms": tf, "forward_GBs": gf, "forward_frac": gf / HBM_PEAK_GBS,
                     "forward_mlups": n / (tf * 1e-3) / 1e6, "adjoint_kernel_ms": ta, "adjoint_GBs": ga,
                     "adjoint_frac": ga / HBM_PEAK_GBS, "adjoint_mlups": n / (ta * 1e-3) / 1e6,
                     "forward_steps": rr["stats"]["forward_steps"], "snapshot_slots": rr["stats"]["snapshot_slots"],
                     "max_mach": rr["stats"]["max_mach"], "mass_drift": rr["stats"]["mass_drift"],
                     "forward_kernel_ms_back_to_back": tf_iso, "adjoint_kernel_ms_back_to_back": ta_iso}
    out["per_precision"] = per
    if "f32" in res and "f16s" in res:
        out["speedup_f16s_vs_f32"] = res["f32"]["wall"] / res["f16s"]["wall"]
        out["speedup_f16s_vs_f32_with_download"] = ((res["f32"]["wall"] + res["f32"]["download_s"]) /
                                                    (res["f16s"]["wall"] + res["f16s"]["download_s"]))
    if "f32" in res and "f16s_l32" in res:
        out["speedup_f16s_l32_vs_f32"] = res["f32"]["wall"] / res["f16s_l32"]["wall"]
    if "f64" in res:
        g64 = res["f64"]["alpha_bar"]
        errs = {}
        for prec in res:
            if prec == "f64":
                continue
            g = res[prec]["alpha_bar"]
            # max relative error over the voxels that carry the gradient: |g64| >= 1 % of max
            big = np.abs(g64) >= 0.01 * np.abs(g64).max()
            errs[prec] = {
                "grad_rel_l2": float(np.linalg.norm(g - g64) / np.linalg.norm(g64)),
                "grad_cosine": float(g @ g64 / (np.linalg.norm(g) * np.linalg.norm(g64))),
                "grad_max_rel_error": float(np.max(np.abs(g[big] - g64[big]) / np.abs(g64[big]))),
                "grad_max_rel_error_voxels": int(big.sum()),
                "grad_max_abs_error_over_max": float(np.abs(g - g64).max() / np.abs(g64).max()),
                "J_rel": abs(res[prec]["J"] - res["f64"]["J"]) / abs(res["f64"]["J"]),
            }
        out["error_vs_f64"] = errs
    out["copy_bandwidth_GBs"] = copy_gbs
    out["roofline"]["frac_of_copy"] = gbs_dom / copy_gbs
    if not args.no_cpu:
        threads = cpu_threads()
        log(f"[bench] cpu baseline: oracle, {args.cpu_window}-step forward+adjoint window, {threads} threads")
        mlups, dt, times = cpu_baseline(shape, args.cpu_window, alpha_cfg, threads)
        try:
            affinity = len(os.sched_getaffinity(0))
        except AttributeError:
            affinity = os.cpu_count() or 1
        out["cpu_baseline"] = {"value": mlups, "unit": "MLUPS (fwd+adj)", "cores": threads, "kind": "port",
                               "sample": f"{args.cpu_window}-step forward+adjoint window (all states kept) of the "
                                         f"same {shape[0]}x{shape[1]}x{shape[2]} problem; median of "
                                         f"{len(times)} runs after one warm-up ({', '.join(f'{t:.2f}' for t in times)}"
                                         f" s); {threads} OpenMP threads of {affinity} cores in the process's "
                                         f"affinity, OMP_NUM_THREADS={os.environ.get('OMP_NUM_THREADS')}"}
    print(json.dumps(out), flush=True)


def _cfg_only(shape, variant, budget, port_rule="nee"):
    """The channel: NEE ports (the benchmark's), or the stationary solver's moment or
    regularized rules on the x faces without their y/z edges (ports.rs refuses port cells
    on the edges of closed faces)."""
    from implexity_amd import design
    from implexity_amd.lbm import LatticeConfig, Port

    precision, adjoint_storage = VARIANTS[variant][:2]
    nx, ny, nz = shape
    if port_rule == "nee":
        prof = design.parabolic_profile(ny, nz).reshape(-1)
        ports = [Port("xmin", "velocity", (0.03, 0.0, 0.0), prof), Port("xmax", "pressure", (1.0,))]
    else:
        mi = np.zeros(shape, bool)
        mo = np.zeros(shape, bool)
        mi[0, 1:-1, 1:-1] = True
        mo[-1, 1:-1, 1:-1] = True
        ports = [Port("xmin", f"velocity_{port_rule}", (0.03, 0.0, 0.0), mask=mi),
                 Port("xmax", f"pressure_{port_rule}", (1.0, 0.0, 0.0), mask=mo)]
    return LatticeConfig(shape=shape, periodic=(False, False, False), lattice=19, collision="trt",
                         precision=precision, omega=1.0 / 0.55, magic=3.0 / 16.0, ports=ports,
                         checkpoint_budget_bytes=budget, adjoint_storage=adjoint_storage)


if __name__ == "__main__":
    main()
