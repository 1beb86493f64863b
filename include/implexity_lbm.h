/*
 * implexity_lbm.h -- C ABI of the MI355X differentiable lattice-Boltzmann library.
 *
 * Drop-in boundary for Implexity's LBM hot path.  The GPU library replaces the
 * reference's in-process CPU loops below these interfaces (paths relative to
 * crates/implexity-physics-lbm/src/ unless noted):
 *
 *   lbm_create            LbmProblem::normalize + Streaming::new      solver.rs:414-550, 83-120
 *                         (grid, periodic axes, solid mask, halfway bounce-back table) and,
 *                         for the collision operator, Collision::new   moving/collision.rs:244-292
 *   lbm_set_design        the Brinkman coefficient field alpha = LbmProblem::resistance(phi)
 *                                                                       solver.rs:580-590, design.rs:282-294
 *   lbm_reset_initial     LbmProblem::initial_populations             solver.rs:592-601
 *   lbm_step              LbmProblem::interval (collide->stream->ports) solver.rs:603-614
 *   lbm_forward           LbmProblem::history + objective             solver.rs:676-687
 *   lbm_gradient          history + history_adjoint behind scan_adjoint with binomial
 *                         checkpointing                                solver.rs:797-815,
 *                         crates/implexity-ad/src/scan.rs:140-390, revolve.rs:128-388
 *   lbm_gradient_seeded   history_adjoint(history, alpha, tau, lambda_final) solver.rs:798-815
 *                         (both with the admission of LbmProblem::admit when mach_limit > 0,
 *                         solver.rs:707-793)
 *   lbm_step_vjp          LbmProblem::interval_vjp                    solver.rs:641-674
 *   lbm_step_jvp          ScanStep::jvp of the step                   crates/implexity-ad/src/scan.rs:46-80
 *   lbm_vector            LbmProblem::vector (terminal responses)      solver.rs:874-894
 *   lbm_vector_vjp        LbmProblem::vector_vjp                      solver.rs:896-947
 *   lbm_get_state         the reference's AoS state vector (x*Q+q)    solver.rs:90,325
 *   lbm_get_fields        LbmProblem::fields (lattice units)          solver.rs:689-705
 *   lbm_debug_link_table  Streaming::source_of for every (cell, q)    solver.rs:160-168
 *   lbm_revolve_schedule  BinomialSchedule::new / actions             revolve.rs:261-329
 *   lbm_every_schedule    Checkpointing::Every(k) of scan_adjoint     scan.rs:109-246
 *
 * Conventions
 *   - Plain pointers and explicit lengths; host buffers are caller-owned (like the
 *     reference's borrowed &[f64] inputs and owned Vec<f64> outputs).
 *   - Every call returns an lbm_status; on failure lbm_last_error() gives the message
 *     (thread-local).  Nothing unwinds across the ABI.
 *   - A handle owns its device memory and one HIP stream.  A handle is not safe for
 *     concurrent calls from several threads (wrap it in a mutex, as the reference
 *     wraps its adjoint cache, moving/field.rs:542).
 *   - All quantities are in lattice units.  Cell index x = (i*ny + j)*nz + k.
 */
#ifndef IMPLEXITY_LBM_H
#define IMPLEXITY_LBM_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define LBM_ABI_VERSION 6

/* Status codes, one per reference error category (crates/implexity-core/src/error.rs:34-60,
 * crates/implexity-ad/src/error.rs). */
typedef enum {
    LBM_OK = 0,
    LBM_ERR_CONTRACT = 1,    /* CaeError::Contract: invalid problem or argument     */
    LBM_ERR_CONVERGENCE = 2, /* CaeError::Convergence: admission refused (NaN, f<=0, Mach) */
    LBM_ERR_SHAPE = 3,       /* AdError::Shape: buffer length mismatch              */
    LBM_ERR_INVALID = 4,     /* AdError::Invalid: schedule or state inconsistency   */
    LBM_ERR_DEVICE = 5,      /* HIP runtime failure                                 */
    LBM_ERR_MEMORY = 6       /* device allocation failed or budget too small        */
} lbm_status;

typedef enum { LBM_D3Q19 = 19, LBM_D3Q27 = 27 } lbm_lattice;

/* Collision operators.  BGK/TRT: moving/collision.rs:457-488; MRT: :79-197,349-375;
 * CUMULANT (D3Q27 only): :636-752; BGK_GUO: the stationary solver's collide_cell
 * (solver.rs:298-322, D3Q19 only).  All use the Brinkman branch of psm.rs:213-226. */
typedef enum {
    LBM_BGK = 0,
    LBM_TRT = 1,
    LBM_MRT = 2,
    LBM_CUMULANT = 3,
    LBM_BGK_GUO = 4,
    LBM_REGULARIZED = 5  /* CollisionKind::Regularized (collision.rs:376-394,419-431) */
} lbm_collision;

/* Population storage.  Arithmetic is FP64 for LBM_F64 and FP32 otherwise. */
typedef enum {
    LBM_F64 = 0,
    LBM_F32 = 1,
    LBM_F16_SHIFTED = 2, /* FP16 of (f_q - w_q) times an exact power of two */
    LBM_BF16_SHIFTED = 3 /* BF16 of (f_q - w_q); no LES closures            */
} lbm_precision;

typedef enum { LBM_FACE_XMIN = 0, LBM_FACE_XMAX, LBM_FACE_YMIN, LBM_FACE_YMAX, LBM_FACE_ZMIN, LBM_FACE_ZMAX } lbm_face;

typedef enum {
    LBM_PORT_VELOCITY = 0,          /* NEE velocity port (moving/boundary.rs:445-506)          */
    LBM_PORT_PRESSURE = 1,          /* NEE pressure port, u_b = u_nb (boundary.rs:445-506)     */
    LBM_PORT_PRESSURE_NEIGHBOR = 2, /* ports.rs reconstruct_neighbor (:484-497), D3Q19         */
    LBM_PORT_VELOCITY_MOMENT = 3,   /* ports.rs reconstruct, "legacy_moment" (:343-406), D3Q19 */
    LBM_PORT_PRESSURE_MOMENT = 4,   /* the same with a given density                          */
    LBM_PORT_VELOCITY_REGULARIZED = 5, /* ports.rs reconstruct_regularized (:455-482), D3Q19  */
    LBM_PORT_PRESSURE_REGULARIZED = 6  /* the same with a given density                       */
} lbm_port_kind;

/* One inlet/outlet on a face, applied after streaming.  Port cells are the non-solid
 * cells of the face (of `mask` when given) in ascending cell order (boundary.rs:331-333,
 * ports.rs:200).
 *   VELOCITY, PRESSURE: the moving solver's NEE rule from the inward neighbour.
 *     Velocity: rho_b = rho_nb, u_b = value * profile[r].  Pressure: rho_b = value[0],
 *     u_b = u_nb (zero-gradient outlet).
 *   The stationary solver's rules (ports.rs, D3Q19 only, no temperature field):
 *   PRESSURE_NEIGHBOR (ports.rs:484-497): from the inward neighbour,
 *     f = f^eq(value[0], t) + f_nb - f^eq(rho_nb, u_nb), t = u_nb along the normal and
 *     (value[1], value[2]) across it (the tangential axes in ascending order).
 *   VELOCITY_MOMENT (ports.rs:343-406): from the port cell's own populations, the
 *     unknown populations from bounce-back of the non-equilibrium part corrected to
 *     density rho = (zero + 2 outgoing sums) / (1 - sign v0_n) and momentum rho v0,
 *     v0 = value * profile[r] - a/2 with a the lattice acceleration.
 *   PRESSURE_MOMENT: the same with rho = value[0] given, the normal velocity from it and
 *     tangential v0 = (value[1], value[2]) - a/2.
 *   VELOCITY_REGULARIZED, PRESSURE_REGULARIZED ("regularized_known_stress",
 *     ports.rs:409-482): the density and momentum of the MOMENT rule, the populations
 *     rebuilt as f^eq + the second-order stress basis times the six stresses fitted (least
 *     squares weighted by 1/w) to the non-equilibrium parts of the 14 known populations.
 *     Refused with a nonzero acceleration (ports.rs:154-158).
 * A neighbour rule's inward neighbour must be fluid, not a port cell, and not shared with
 * another port cell.  The reference's authoring checks (ports.rs:107-262: no corners on
 * closed faces, protected fluid, Mach and density limits) are the caller's. */
typedef struct {
    uint32_t face;          /* lbm_face */
    uint32_t kind;          /* lbm_port_kind */
    double value[3];        /* see lbm_port_kind */
    const double* profile;  /* one factor per port cell (velocity kinds), or NULL for 1 */
    uint64_t profile_len;
    const uint8_t* mask;    /* cells x (lattice order) nonzero on this port's cells, all on
                               the face and fluid; NULL: every non-solid cell of the face */
    uint64_t mask_len;      /* cells, or 0 with mask NULL */
} lbm_port;

/* Storage of the adjoint cotangent field (lambda).  AUTO: FP64 for FP64 storage, FP16
 * for FP16-shifted storage, FP32 otherwise.  F16_SCALED stores lambda * 2^e with one
 * exponent for the whole field, chosen every adjoint step from the measured max |lambda|
 * (no counterpart in the reference, whose cotangents are FP64 Vec<f64>); a step that
 * would leave the FP16 range makes lbm_gradient / lbm_step_vjp fail with
 * LBM_ERR_CONVERGENCE. */
/* One sponge layer (moving/sponge.rs:12-29, SpongeField::new :112-181), in lattice units:
 * the fluid cells within `thickness` layers of `face` are blended after collision,
 *   post := post (1 - sigma) + sum_l s_l f^eq(rho_l, u_l)       (field.rs:1461-1475),
 * s_l = strength * ramp((depth - 1/2) / thickness) with the quintic smoothstep ramp and
 * depth 1 .. thickness counted inward from the face's innermost layer (thickness at the
 * face), sigma = sum_l s_l < 1.  The far-field target of the step that reads state t is
 *   rho_l = density,  u_l = velocity + A r(t) sin(2 pi f t + phase - theta(x)),
 *   theta(x) = 2 pi f ((x - origin) . direction) / speed,  x = cell centre (ijk + 1/2),
 * r(t) the smoothstep over wave_ramp steps (field.rs:1015-1036, boundary.rs:110-125). */
typedef struct {
    uint32_t face;              /* lbm_face; the axis must not be periodic */
    uint32_t thickness;         /* 1 .. half the lattice along the face's axis */
    double strength;            /* in (0, 1) */
    double density;             /* far-field density */
    double velocity[3];         /* far-field mean velocity */
    uint32_t wave;              /* 0: steady far field; 1: travelling wave below */
    uint32_t reserved;
    double wave_amplitude[3];   /* A (velocity) */
    double wave_direction[3];   /* unit vector */
    double wave_speed;          /* cells per step, > 0 */
    double wave_frequency;      /* cycles per step, > 0 */
    double wave_phase;          /* rad */
    double wave_ramp;           /* steps over which the wave ramps in, >= 0 */
    double wave_origin[3];      /* cells */
} lbm_sponge;

#define LBM_MAX_SPONGES 8

/* Fluid-solid coupling of a cell.  BRINKMAN: the design array is the Brinkman
 * coefficient alpha (psm.rs:213-226 with m = 0).  PSM / PSM_SUPERPOSITION: the design
 * array is the occupancy d (solid fraction; d = 0 is fluid), with the solid momentum m
 * of lbm_set_solid_momentum (default 0), the Psm and PsmSuperposition branches of
 * cell_update (moving/psm.rs:228-266): a cell with d > 0 leaves the fluid collision
 * (u = j/rho + a/2, force rho a) as
 *   PSM:                f + (1 - B)(fluid - f) + B (f_opp - eq_opp - f) + s_eq
 *   PSM_SUPERPOSITION:  fluid + s_eq - B eq
 * with s_eq = rho w (B + (B/d)(3 c.m + 4.5 (c.m)^2/d - 1.5 |m|^2/d)) and Noble &
 * Torczynski's B = d t / (1 - d + t), t = tau+ - 1/2, saturated smoothly above
 * 1 - psm_kappa (psm.rs:96-152).  BRINKMAN_OCCUPANCY: the Brinkman branch of
 * cell_update with a moving solid (psm.rs:143-160, 213-226): alpha = psm_c_alpha e /
 * ((psm_drag_shape + 1) - e) of the saturated occupancy e, gamma = alpha / d,
 * u = (j/rho + a/2 + gamma m/2) / (1 + alpha/2) and the force rho (a + gamma m - alpha u).
 * The gradient's alpha_bar is then d_bar (0 at d = 0), and lbm_solid_momentum_bar
 * returns m_bar.  BGK and TRT only; the objective is the pressure drop. */
typedef enum {
    LBM_COUPLING_BRINKMAN = 0,
    LBM_COUPLING_PSM = 1,
    LBM_COUPLING_PSM_SUPERPOSITION = 2,
    LBM_COUPLING_BRINKMAN_OCCUPANCY = 3
} lbm_coupling;

typedef enum {
    LBM_ADJ_AUTO = 0,
    LBM_ADJ_F64 = 1,
    LBM_ADJ_F32 = 2,
    LBM_ADJ_F16_SCALED = 3
} lbm_adjoint_storage;

typedef struct {
    uint32_t struct_size;         /* sizeof(lbm_config) */
    uint32_t lattice;             /* lbm_lattice */
    uint32_t collision;           /* lbm_collision */
    uint32_t precision;           /* lbm_precision */
    int32_t shape[3];             /* nx, ny, nz (each >= 2) */
    uint8_t periodic[3];          /* 0: closed halfway-bounce-back walls */
    uint8_t reserved0[5];
    double omega;                 /* 1/tau+ ; tau+ in [0.51, 2] */
    double magic;                 /* TRT magic Lambda (default 3/16) / MRT odd_magic */
    double bulk_rate;             /* MRT bulk rate or cumulant bulk rate; NaN = default */
    double even_rate;             /* MRT even higher-moment rate; NaN = default */
    double accel[3];              /* body acceleration (lattice) */
    uint32_t n_ports;
    uint32_t reserved1;
    const lbm_port* ports;
    uint64_t checkpoint_budget_bytes; /* device bytes for revolve snapshots (0: 1/2 of free) */
    uint32_t adjoint_storage;     /* lbm_adjoint_storage */
    uint32_t checkpoint_every;    /* 0: binomial (revolve) in the budget; k > 0: Checkpointing::Every(k)
                                     (scan.rs:109-246), needing ceil(n/k) + max(k-2, 0) snapshots */
    /* Coupled D3Q7 temperature field (double distribution).  No counterpart in the
     * reference, whose thermal coupling is finite-volume transport on D3Q19 links
     * (lbm/thermal.rs:91-134).  T = sum g_i, g_i^eq = w_i T (1 + 4 c_i.u) with the flow's
     * Brinkman velocity u, diffusivity kappa(phi) = thermal_kappa (1 + (thermal_ratio - 1)
     * phi), tau_T = 1/2 + 4 kappa, source thermal_source * phi per step, adiabatic
     * (bounce-back) walls; at the flow's ports: velocity ports hold T = thermal_inlet,
     * pressure ports are zero-gradient.  phi is the thermal design (lbm_set_thermal_design,
     * default 0).  Needs a fluid-only lattice (no solid mask), a non-BGK-Guo operator and
     * FP64/FP32/FP16 storage; cotangents stay FP64 (FP64 storage) or FP32. */
    uint32_t thermal;             /* 0: flow only; 1: flow + temperature */
    uint32_t reserved2;
    double thermal_kappa;         /* fluid diffusivity (lattice units), > 0 */
    double thermal_ratio;         /* solid:fluid diffusivity ratio, > 0 */
    double thermal_source;        /* heat source per step at phi = 1 */
    double thermal_inlet;         /* temperature held at velocity ports */
    double thermal_range;         /* FP16 storage: bound on |T| that sets the storage scale (0: 1) */
    /* Smagorinsky closure on BGK/TRT (moving/psm.rs:45-48,161-197): the effective
     * omega+ of a cell is 1/tau_e, tau_e = (tau0 + sqrt(tau0^2 + 18 sqrt(2) C^2 P / rho)) / 2,
     * P = sqrt(|Pi_neq|^2 + floor^2) - floor; TRT's omega- follows through the magic
     * parameter.  les_constant = 0: no closure. */
    double les_constant;          /* C >= 0 */
    double les_norm_floor;        /* floor >= 0 */
    /* WALE closure on BGK/TRT (moving/wale.rs), exclusive with Smagorinsky: the rate of a
     * fluid cell is 1 / (tau0 + 3 nu_t), nu_t = C^2 A^(3/2) / (B^(5/2) + A^(5/4) + floor)
     * from the invariants A, B of the velocity gradient of u = j / rho of the state the
     * step collides (central differences, one-sided next to walls and closed faces). */
    double wale_constant;         /* C >= 0; 0: no closure */
    double wale_floor;            /* > 0 when wale_constant > 0 */
    /* Sponge layers, at most LBM_MAX_SPONGES (lbm_sponge). */
    uint32_t n_sponges;
    uint32_t reserved3;
    const lbm_sponge* sponges;
    uint32_t coupling;            /* lbm_coupling */
    uint32_t reserved4;
    double psm_kappa;             /* saturation width in [0, 1) (occupancy couplings) */
    double psm_c_alpha;           /* BRINKMAN_OCCUPANCY: drag coefficient c_alpha >= 0 */
    double psm_drag_shape;        /* BRINKMAN_OCCUPANCY: drag shape q > 0 */
    /* Admission of the trajectory of a sweep (LbmProblem::admit, solver.rs:707-793; the
     * moving solver's check_state, moving/field.rs:1735-1775).  With mach_limit > 0,
     * lbm_forward and lbm_gradient(_seeded) refuse the states of their first sweep
     * (states 0..n) with LBM_ERR_CONVERGENCE when a fluid population is non-finite or
     * <= 0, when the Mach number sqrt(3) max |u| exceeds mach_limit (at most 0.2,
     * solver.rs:434-437), or when the relative fluid-mass drift, port and sponge exchange
     * included, exceeds drift_limit (0: 1e-6, the reference's bound, solver.rs:765).  The
     * gradient refuses before its reverse sweep.  mach_limit = 0: no refusal (lbm_stats
     * still reports the quantities). */
    double mach_limit;            /* 0, or in (0, 0.2] */
    double drift_limit;           /* >= 0; 0: 1e-6 */
    /* The relaxation time itself (solver.rs:487, tau = 1/2 + 3 nu dt / dx^2), for callers
     * that hold tau rather than omega: > 0 replaces omega by 1/tau and BGK-Guo divides by
     * this tau as given; 0: tau = 1/omega. */
    double tau;
} lbm_config;

/* Objective J = sum over states s in [t_begin, t_end] of weight * phi(s)
 *             + terminal_weight * phi(final state), states indexed 1..n (scan.rs:96-107). */
typedef enum {
    LBM_OBJ_NONE = 0,
    LBM_OBJ_DISSIPATION = 1,   /* sum over fluid cells of rho alpha |u|^2 (lattice units) */
    LBM_OBJ_PRESSURE_DROP = 2, /* mean rho/3 on plane a minus mean rho/3 on plane b */
    LBM_OBJ_VECTOR = 3,        /* w . vector(state), the terminal responses of solver.rs:874-894
                                  (the design-volume term w[5] is not state-dependent and is
                                  left out of J; lbm_vector reports it) */
    LBM_OBJ_THERMAL = 4,       /* coupled temperature field: w[0] * (mean T over the cells with
                                  thermal design phi > 1/2) + w[1] * (pressure drop between
                                  plane_a and plane_b); w = vector_weights (no counterpart in
                                  the reference) */
    LBM_OBJ_SOLID_FORCE = 5    /* occupancy couplings: w . (momentum the coupled cells exchange
                                  with the solid in the collision of the state, CellOutput::dp,
                                  psm.rs:240-262, summed over the cells), w = vector_weights[0..2];
                                  a stage cost with t_end < nsteps and no terminal weight */
} lbm_objective_kind;

/* Physical scales of the terminal responses (LbmProblem::vector_scales, solver.rs:862-871). */
typedef struct {
    double velocity;    /* spacing / step (m/s per lattice unit) */
    double density;     /* kg/m^3 per lattice unit */
    double step_s;      /* time step (s) */
    double cell_volume; /* spacing^3 (m^3) */
} lbm_vector_scales;

typedef struct {
    uint32_t kind;        /* lbm_objective_kind */
    uint32_t axis;        /* pressure drop: axis normal to the planes */
    int32_t plane_a;      /* pressure drop: mean rho/3 on plane a minus on plane b */
    int32_t plane_b;
    uint64_t t_begin;     /* first state carrying the stage cost (>= 1) */
    uint64_t t_end;       /* last state carrying the stage cost (<= n) */
    double weight;        /* stage weight */
    double terminal_weight;
    double vector_weights[6];     /* LBM_OBJ_VECTOR and LBM_OBJ_THERMAL: w */
    lbm_vector_scales scales;     /* LBM_OBJ_VECTOR */
} lbm_objective;

typedef struct {
    uint64_t forward_steps;      /* step evaluations, recomputation included */
    uint64_t reverse_steps;
    uint64_t snapshot_slots;     /* slots the budget allowed */
    uint64_t peak_snapshots;     /* BinomialSchedule::peak_snapshots */
    uint64_t snapshot_copies;    /* snapshot + restore copies */
    double forward_ms;           /* device time in forward-step kernels */
    double adjoint_ms;           /* device time in adjoint-step kernels */
    double total_ms;             /* wall time of the call */
    double max_mach;             /* sqrt(3) max |u| over the first sweep (admit, solver.rs:742-751) */
    double mass_drift;           /* max relative drift of fluid mass incl. port exchange */
    uint64_t bad_cells;          /* cells with a non-finite or non-positive population */
    double forward_kernel_avg_ms; /* forward step kernel alone, HIP events on every 64th step
                                     (every n/4-th, so at least four, on horizons below 256) */
    double adjoint_kernel_avg_ms; /* adjoint step kernel alone, sampled the same way */
    uint64_t kernel_samples;      /* steps timed for the two averages (gradient only) */
    double cotangent_max;         /* FP16 cotangents: max |lambda| of the last adjoint step (else NaN) */
    int64_t cotangent_exponent;   /* FP16 cotangents: exponent e of the stored lambda * 2^e (else 0) */
    double min_population;        /* min over the fluid populations f (not f - w) of the states of
                                     the first sweep (admit, solver.rs:716-727); NaN if none */
} lbm_stats;

typedef struct lbm_handle lbm_handle;

int lbm_abi_version(void);
int lbm_last_error(char* buf, uint64_t len);

int lbm_create(const lbm_config* cfg, const uint8_t* solid_mask, uint64_t mask_len, lbm_handle** out);
int lbm_destroy(lbm_handle* h);

/* alpha: one Brinkman coefficient per cell (solid cells ignored); with a PSM coupling,
 * the occupancy d of every cell. */
int lbm_set_design(lbm_handle* h, const double* alpha, uint64_t n);
int lbm_reset_initial(lbm_handle* h);

/* PSM couplings: the solid momentum m of every cell (SolidInput::m, psm.rs:56-59; the
 * solid velocity times d), AoS [cell][3], len = 3 * cells, and its cotangent summed over
 * the last lbm_gradient / lbm_step_vjp (CellBar::m). */
int lbm_set_solid_momentum(lbm_handle* h, const double* m, uint64_t len);
int lbm_solid_momentum_bar(lbm_handle* h, double* out, uint64_t len);

/* Second occupancy endpoint (Occupancy::{d, m}[1], a body moving within an interval):
 * the step that reads state t collides with d = d0 (1 - theta) + d1 theta and m likewise
 * (Field::solid_at, field.rs:1266-1275), theta = 0 up to t0, (t - t0) / (t1 - t0) between,
 * 1 from t1; d0, m0 are the design and lbm_set_solid_momentum.  m1 is [cell][3] or NULL
 * (0).  The gradient's alpha_bar and lbm_solid_momentum_bar then hold the first
 * endpoint's cotangents and lbm_occupancy_end_bar the second's. */
int lbm_set_occupancy_end(lbm_handle* h, const double* d1, const double* m1, uint64_t n, uint64_t t0, uint64_t t1);
int lbm_occupancy_end_bar(lbm_handle* h, double* d1_bar, double* m1_bar, uint64_t n);

/* One macro step of the subcycled moving field (SubcycledField::subcycle and
 * subcycle_adjoint, crates/implexity-solve/src/multirate_coupling/mod.rs:86-166, on
 * MovingField, moving/field.rs:1870-1919 and :2055-2392).  The handle's occupancy coupling
 * holds the interval's endpoints: d0 = the design (lbm_set_design), m0 =
 * lbm_set_solid_momentum, and d1, m1 = lbm_set_occupancy_end (whose t0, t1 these calls
 * replace).  Substep k = 1 .. m collides the state with the endpoints blended at theta =
 * (k - 1/2) / m (Field::params, field.rs:1049) and adds each occupied cell's impulse
 * shares (Field::accumulate, :1499-1541)
 *   g0 += (-dp / d)(1 - theta),   g1 += (-dp / d) theta,
 * dp the momentum the cell exchanges with the solid, d the blended occupancy.
 *
 * lbm_subcycle advances the handle's state by m substeps and returns g as [cell][6]
 * (g0 x, y, z, g1 x, y, z; 0 where no solid), NULL to skip.  With a Mach limit it refuses
 * (LBM_ERR_CONVERGENCE) as Field::check_state (:1735-1775): a non-finite or non-positive
 * population of the end state, or a substep's Mach number above the limit.
 *
 * lbm_subcycle_adjoint, from the state the subcycle started at (the handle's state), forms
 * the cotangents of the end state (state_bar, AoS) and of g (g_bar, [cell][6], or NULL for
 * 0) back to the start: previous_bar (AoS), d0_bar (one per cell, or NULL), and through
 * lbm_solid_momentum_bar and lbm_occupancy_end_bar m0_bar, d1_bar and m1_bar.  The
 * substeps are recomputed under the handle's checkpoint budget (the inner binomial
 * schedule of Field::prepare, :2312-2385); the handle keeps the start state.  D3Q19
 * only: the D3Q27 transpose is refused (LBM_ERR_CONTRACT).  The
 * reference's observables, entrainment compensation, carrier flux and time-scale
 * derivatives stay with the caller. */
int lbm_subcycle(lbm_handle* h, uint32_t substeps, double* g, uint64_t g_len, lbm_stats* stats);
int lbm_subcycle_adjoint(lbm_handle* h, uint32_t substeps, const double* state_bar, uint64_t state_len,
                         const double* g_bar, uint64_t g_len, double* previous_bar, double* d0_bar, uint64_t n,
                         lbm_stats* stats);

/* State exchange in the reference layout: layout 0 = AoS (x*Q+q), 1 = SoA (q*N+x). */
int lbm_set_state(lbm_handle* h, int layout, const double* f, uint64_t len);
int lbm_get_state(lbm_handle* h, int layout, double* f, uint64_t len);
uint64_t lbm_step_count(const lbm_handle* h);

int lbm_step(lbm_handle* h, uint64_t nsteps);
int lbm_forward(lbm_handle* h, uint64_t nsteps, const lbm_objective* obj, double* J, lbm_stats* stats);

/* Forward + discrete adjoint from the current state (taken as state 0) over nsteps with
 * binomial checkpointing inside the handle's snapshot budget.  alpha_bar (n cells,
 * FP64, or NULL to leave it in device memory for lbm_alpha_bar) receives dJ/dalpha;
 * lambda0 (optional, AoS N*Q) receives dJ/d state0.  The handle's state is left at
 * state 0. */
int lbm_gradient(lbm_handle* h, uint64_t nsteps, const lbm_objective* obj, double* J, double* alpha_bar,
                 uint64_t n_alpha, double* lambda0, uint64_t lambda_len, lbm_stats* stats);

/* dJ/dalpha of the last lbm_gradient / lbm_gradient_seeded / lbm_step_vjp call (n cells,
 * FP64).  A gradient called with alpha_bar = NULL leaves it in device memory, so that a
 * caller that does not need it on the host at once (or a timed loop) skips the copy; this
 * copies it out afterwards through the handle's pinned staging buffer.  No counterpart in
 * the reference, whose history_adjoint returns it as Vec<f64> (solver.rs:797-815). */
int lbm_alpha_bar(lbm_handle* h, double* alpha_bar, uint64_t n);

/* lbm_gradient with an additional cotangent lambda_final (AoS N*Q, in the reference
 * layout) added to the seed of the final state: the drop-in for
 * LbmProblem::history_adjoint(history, alpha, tau, lambda_final) (solver.rs:798-815),
 * where the caller computed lambda_final (e.g. with lbm_vector_vjp).  obj may be NULL. */
int lbm_gradient_seeded(lbm_handle* h, uint64_t nsteps, const lbm_objective* obj, const double* lambda_final,
                        uint64_t lambda_final_len, double* J, double* alpha_bar, uint64_t n_alpha, double* lambda0,
                        uint64_t lambda_len, lbm_stats* stats);

/* Terminal responses of the current state (LbmProblem::vector, solver.rs:874-894):
 * out = [mean u_x, mean u_y, mean u_z over fluid cells, 0.5 sum rho |u|^2 dV,
 * sum rho alpha/dt |u|^2 dV, sum phi dV], u and rho scaled to physical units.  phi
 * (n = cells) may be NULL (volume 0).  Sums are FP64. */
int lbm_vector(lbm_handle* h, const double* phi, uint64_t n, const lbm_vector_scales* scales, double* out);

/* lbm_vector without any device allocation, for callers that cannot allocate (code
 * around HIP graph capture, tight serving loops, harnesses that forbid hipMalloc); no
 * counterpart in the reference.  phi_dev (n = cells doubles, or NULL) and out_dev (6
 * doubles) are device pointers on the handle's device, which must be current; the
 * responses are those of lbm_vector (sums folded on the device in a fixed order instead
 * of on the host).  The per-block partial sums live in a fixed 4 MiB device workspace
 * of the library, so lattices needing more than (2^19 - 256) / 5 step blocks are
 * refused with LBM_ERR_MEMORY (a 256^3 lattice needs 65536).  Calls on one GPU share the
 * workspace: overlapping calls on one GPU race.  Returns after the stream has completed. */
int lbm_vector_noalloc(lbm_handle* h, const double* phi_dev, uint64_t n, const lbm_vector_scales* scales,
                       double* out_dev);

/* VJP of w . vector at the current state (LbmProblem::vector_vjp, solver.rs:896-947):
 * f_bar (AoS N*Q), alpha_bar (N), phi_bar (N, may be NULL). */
int lbm_vector_vjp(lbm_handle* h, const double* phi, uint64_t n, const lbm_vector_scales* scales, const double* w,
                   double* f_bar, uint64_t f_len, double* alpha_bar, double* phi_bar);

/* Allocate the adjoint buffers and the revolve snapshot pool for a horizon of nsteps
 * (min(budget / state bytes, nsteps) slots) ahead of lbm_gradient, so that the
 * allocation is not part of a timed gradient.  lbm_gradient does this itself when needed. */
int lbm_reserve(lbm_handle* h, uint64_t nsteps, uint64_t* slots);

/* VJP of one step at the current state: lambda_prev = J^T lambda_next (AoS N*Q), and
 * alpha_bar (N) = d(lambda_next . step)/d alpha. */
int lbm_step_vjp(lbm_handle* h, const double* lambda_next, double* lambda_prev, double* alpha_bar, uint64_t len);

/* Tangent of one step at the current state (ScanStep::jvp, crates/implexity-ad/src/
 * scan.rs:46-80, whose default makes one vjp call per state entry): d_next = J_state d_state
 * + J_alpha d_alpha (AoS N*Q; d_alpha has one entry per cell, or NULL for 0), computed in
 * FP64 forward mode through the step's own collision, streaming and port code.  The
 * handle's state is not advanced.  Covers the BGK, TRT, MRT, regularized, cumulant and
 * BGK-Guo Brinkman operators, the Smagorinsky and WALE closures on BGK/TRT, sponge
 * layers and the occupancy couplings (d_alpha is then the occupancy's tangent; the Psm
 * law on D3Q19 only), not the temperature field. */
int lbm_step_jvp(lbm_handle* h, const double* d_state, const double* d_alpha, double* d_next, uint64_t len);

int lbm_get_fields(lbm_handle* h, double* rho, double* u, uint64_t n);

/* Source index of every (cell, q) exactly as Streaming::new builds it (NONE = 0xFFFFFFFF
 * for solid cells), derived by running the device's own streaming addressing once on a
 * state of unique identifiers. */
int lbm_debug_link_table(lbm_handle* h, uint32_t* source, uint64_t len);

/* Port cells (cell, neighbour) in application order. */
int lbm_port_cells(const lbm_handle* h, int32_t* cells, int32_t* nbs, uint64_t cap, uint64_t* count);

/* Binomial schedule as int64 quintuples (kind, a, b, 0, 0): kind 0 advance(from,to),
 * 1 snapshot(step,slot), 2 restore(step,slot), 3 reverse(step), 4 release(slot). */
int lbm_revolve_schedule(uint64_t steps, uint64_t snapshots, int64_t* actions, uint64_t cap, uint64_t* count,
                         uint64_t* forward_steps, uint64_t* reverse_start);
uint64_t lbm_revolve_optimal_forward_steps(uint64_t steps, uint64_t snapshots);
/* The Checkpointing::Every(k) action list the gradient driver runs when
 * checkpoint_every = k (same encoding as lbm_revolve_schedule). */
int lbm_every_schedule(uint64_t steps, uint64_t every, int64_t* actions, uint64_t cap, uint64_t* count,
                       uint64_t* forward_steps, uint64_t* reverse_start);

/* Coupled temperature field (lbm_config.thermal = 1); no counterpart in the reference.
 * phi: thermal design per cell in [0, 1] (diffusivity and source interpolate with it;
 * the cells with phi > 1/2 are the "solid" cells of LBM_OBJ_THERMAL).  States are AoS
 * x*7+q over the D3Q7 lattice (rest, then the axis vectors in lexicographic order).
 * lbm_thermal_design_bar returns dJ/dphi through the temperature field of the last
 * lbm_gradient / lbm_gradient_seeded call (dJ/dalpha of the flow, coupling included,
 * is that call's alpha_bar); lbm_get_thermal_lambda returns the cotangent of the
 * temperature populations of state 0 of that call.  lbm_step_vjp covers flow-only
 * handles. */
int lbm_set_thermal_design(lbm_handle* h, const double* phi, uint64_t n);
int lbm_get_thermal_state(lbm_handle* h, double* g, uint64_t len);
int lbm_set_thermal_state(lbm_handle* h, const double* g, uint64_t len);
int lbm_get_temperature(lbm_handle* h, double* T, uint64_t n);
int lbm_thermal_design_bar(lbm_handle* h, double* phi_bar, uint64_t n);
int lbm_get_thermal_lambda(lbm_handle* h, double* lambda, uint64_t len);

/* Design map raw -> phi (DesignMap, design.rs:95-280; TopologyMap with the radius in
 * cells, design.rs:37-80): phi = 0 on solid cells, the fixed value outside the design
 * region, and on editable cells (design region, not solid) the cone filter
 * sum_o w_o raw[x - o] / sum_o w_o over editable sources with w_o = radius - |o| > 0,
 * followed by the projection (tanh(beta eta) + tanh(beta (f - eta))) /
 * (tanh(beta eta) + tanh(beta (1 - eta))) when beta > 0.  The device computes the map
 * (DesignMap::map) and its transpose (DesignMap::vjp) as gathers. */
typedef struct lbm_design_map lbm_design_map;

typedef struct {
    uint32_t struct_size;   /* sizeof(lbm_design_map_config) */
    int32_t shape[3];
    uint8_t periodic[3];
    uint8_t reserved[5];
    double radius;          /* filter radius in cells (filter_radius_m / spacing_m), in [0, 4]; 0: none */
    double beta;            /* projection sharpness in [0, 32]; 0: none */
    double eta;             /* projection threshold in (0, 1) */
} lbm_design_map_config;

/* solid, region: one byte per cell or NULL (no solid cells / every cell editable);
 * fixed: the design outside the region, or NULL (0). */
int lbm_design_map_create(const lbm_design_map_config* cfg, const uint8_t* solid, const uint8_t* region,
                          const double* fixed, uint64_t n, lbm_design_map** out);
int lbm_design_map_destroy(lbm_design_map* m);
int lbm_design_map_apply(lbm_design_map* m, const double* raw, double* phi, uint64_t n);
int lbm_design_map_vjp(lbm_design_map* m, const double* raw, const double* phi_bar, double* raw_bar, uint64_t n);

/* Push-forward of body points into the occupancy (moving/pushforward.rs):
 * d[c] = sum_q a_q W(c - xi_q) and m[c] = sum_q a_q v_q W(c - xi_q), W
 * the product of per-axis weights of the cubic B-spline (kernel 0) or Peskin's 4-point
 * kernel (1) of half-width width_cells (support 4 width_cells cells per axis), positions
 * in lattice coordinates (cell c at coordinate c; Frame::to_lattice maps metres to them).
 * d and m feed lbm_set_design / lbm_set_solid_momentum of an occupancy-coupled handle;
 * m is [cell][3].  The transpose gives the cotangents of the points, their weights and
 * their velocities (endpoint_vjp).  Frame::new / with_mirrors (:113-165): a periodic
 * axis is at least 4 width_cells long unless it has one cell (planar: weight 1 on it), a
 * closed axis has more than one cell, and a symmetry face (mirrored bit 0 at the low end,
 * bit 1 at the high end) sits on a closed axis of at least 4 width_cells; the support
 * is folded back into the lattice at a face it crosses (fold_axis, :344-362).  A point
 * whose support leaves a closed axis without a face, or that lies beyond a face, is
 * refused (pushforward_support). */
typedef struct {
    uint32_t struct_size;   /* sizeof(lbm_pushforward_config) */
    int32_t shape[3];
    uint8_t periodic[3];
    uint8_t mirrored[3];    /* per axis: bit 0 symmetry face at cell -1/2, bit 1 at n - 1/2 */
    uint8_t reserved[2];
    uint32_t width_cells;   /* 1 or 2 */
    uint32_t kernel;        /* 0 cubic B-spline (m4), 1 Peskin 4-point */
} lbm_pushforward_config;

int lbm_pushforward(const lbm_pushforward_config* cfg, const double* xi, const double* a, const double* v,
                    uint64_t npoints, double* d, double* m, uint64_t ncells);
int lbm_pushforward_vjp(const lbm_pushforward_config* cfg, const double* xi, const double* a, const double* v,
                        uint64_t npoints, const double* d_bar, const double* m_bar, uint64_t ncells, double* xi_bar,
                        double* a_bar, double* v_bar);

/* Timing support: average device time of `launches` back-to-back launches of one kernel
 * (0 forward step, 1 adjoint step, 2 forward addressing with the collision replaced by
 * a copy, 3 forward step of the flow and the coupled temperature field, 4 thermal
 * adjoint step, 5 sponge blend pass, 6 sponge transpose, 7 forward step as a gradient's
 * first sweep runs it: admission diagnostics on and the per-block partials folded) on
 * the handle's stream, measured with HIP events. */
int lbm_time_kernel(lbm_handle* h, int which, uint64_t launches, double* avg_ms);
/* Device copy bandwidth (GB/s, read+write bytes) of a plain copy kernel. */
int lbm_copy_bandwidth(uint64_t bytes, uint64_t iters, double* gbps, double* avg_ms);
/* Bytes of one population buffer on the device (storage padding included, and with
 * regularized ports the pre-port populations of their cells). */
uint64_t lbm_state_bytes(const lbm_handle* h);

#ifdef __cplusplus
}
#endif

#endif /* IMPLEXITY_LBM_H */
