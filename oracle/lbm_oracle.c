/*
 * lbm_oracle.c -- CPU FP64 restatement of the reference's differentiable LBM path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product (implexity_amd/, include/) links,
 * loads or calls this file.  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg use it, and only as the checker / the timed CPU baseline.
 *
 * Every function restates one reference function operation for operation, in the
 * reference's evaluation order, so that an FP64 build with -ffp-contract=off
 * reproduces the reference's IEEE results (Rust never contracts a*b+c into an FMA).
 * Reference paths are relative to crates/implexity-physics-lbm/src/ unless noted.
 *
 * Layout: AoS, index x*Q+q, cell x = (i*ny + j)*nz + k (z fastest), as the
 * stationary solver (solver.rs:90,325; d3q19.rs:86-89).
 *
 * Parity status of each function (how it is pinned) is listed in the module
 * docstring of oracle/oracle.py and checked by tests/test_oracle_*.py.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define MAXQ 27
#define NONE_U32 0xFFFFFFFFu

/* ------------------------------------------------------------------------- */
/* Lattices: d3q19.rs:29-61 (D3Q19), moving/lattice.rs:93-129,149-156 (D3Q27), */
/* opposite_of moving/lattice.rs:49-63.                                         */
/* ------------------------------------------------------------------------- */
typedef struct {
    int Q;
    int c[MAXQ][3];
    double cf[MAXQ][3];
    double w[MAXQ];
    int opp[MAXQ];
} lattice_t;

static lattice_t L19, L27, L7;
static int lattices_ready = 0;

static const int C19[19][3] = {
    {0, 0, 0},  {-1, -1, 0}, {-1, 0, -1}, {-1, 0, 0}, {-1, 0, 1}, {-1, 1, 0}, {0, -1, -1},
    {0, -1, 0}, {0, -1, 1},  {0, 0, -1},  {0, 0, 1},  {0, 1, -1}, {0, 1, 0},  {0, 1, 1},
    {1, -1, 0}, {1, 0, -1},  {1, 0, 0},   {1, 0, 1},  {1, 1, 0}};

static void init_lattices(void) {
    if (lattices_ready) return;
    /* D3Q19: W0=1/3, W1=1/18 (axis), W2=1/36 (edge); OPPOSITE[i] = 19-i for i >= 1. */
    L19.Q = 19;
    for (int q = 0; q < 19; q++) {
        int n = 0;
        for (int d = 0; d < 3; d++) {
            L19.c[q][d] = C19[q][d];
            L19.cf[q][d] = (double)C19[q][d];
            n += abs(C19[q][d]);
        }
        L19.w[q] = n == 0 ? 1.0 / 3.0 : (n == 1 ? 1.0 / 18.0 : 1.0 / 36.0);
        L19.opp[q] = q == 0 ? 0 : 19 - q;
    }
    /* D3Q27: rest first, then (a,b,c) in {-1,0,1}^3 lexicographic without origin. */
    L27.Q = 27;
    L27.c[0][0] = L27.c[0][1] = L27.c[0][2] = 0;
    int n = 1;
    for (int a = -1; a <= 1; a++)
        for (int b = -1; b <= 1; b++)
            for (int c = -1; c <= 1; c++) {
                if (a == 0 && b == 0 && c == 0) continue;
                L27.c[n][0] = a;
                L27.c[n][1] = b;
                L27.c[n][2] = c;
                n++;
            }
    for (int q = 0; q < 27; q++) {
        int m = abs(L27.c[q][0]) + abs(L27.c[q][1]) + abs(L27.c[q][2]);
        L27.w[q] = m == 0 ? 8.0 / 27.0 : (m == 1 ? 2.0 / 27.0 : (m == 2 ? 1.0 / 54.0 : 1.0 / 216.0));
        for (int d = 0; d < 3; d++) L27.cf[q][d] = (double)L27.c[q][d];
    }
    for (int i = 0; i < 27; i++)
        for (int j = 0; j < 27; j++)
            if (L27.c[j][0] == -L27.c[i][0] && L27.c[j][1] == -L27.c[i][1] && L27.c[j][2] == -L27.c[i][2])
                L27.opp[i] = j;
    /* D3Q7 (the thermal lattice; no counterpart in the reference): rest first, then the
     * six axis vectors in lexicographic order, so OPPOSITE[i] = 7-i; weights 1/4, 1/8,
     * sound speed c_s^2 = 1/4. */
    L7.Q = 7;
    n = 0;
    for (int a = -1; a <= 1; a++)
        for (int b = -1; b <= 1; b++)
            for (int c = -1; c <= 1; c++) {
                if (abs(a) + abs(b) + abs(c) > 1) continue;
                if (a == 0 && b == 0 && c == 0) continue;
                n++;
                L7.c[n][0] = a;
                L7.c[n][1] = b;
                L7.c[n][2] = c;
            }
    L7.c[0][0] = L7.c[0][1] = L7.c[0][2] = 0;
    for (int q = 0; q < 7; q++) {
        L7.w[q] = q == 0 ? 0.25 : 0.125;
        L7.opp[q] = q == 0 ? 0 : 7 - q;
        for (int d = 0; d < 3; d++) L7.cf[q][d] = (double)L7.c[q][d];
    }
    lattices_ready = 1;
}

static const lattice_t *lat(int Q) {
    init_lattices();
    return Q == 27 ? &L27 : (Q == 7 ? &L7 : &L19);
}

int orc_lattice(int Q, int *c_out, double *w_out, int *opp_out) {
    const lattice_t *L = lat(Q);
    if (Q != 19 && Q != 27 && Q != 7) return -1;
    for (int q = 0; q < Q; q++) {
        for (int d = 0; d < 3; d++) c_out[3 * q + d] = L->c[q][d];
        w_out[q] = L->w[q];
        opp_out[q] = L->opp[q];
    }
    return 0;
}

/* ------------------------------------------------------------------------- */
/* Problem description (plain data filled by oracle/oracle.py).               */
/* ------------------------------------------------------------------------- */
enum { K_BGK = 0, K_TRT = 1, K_MRT = 2, K_CUMULANT = 3, K_BGK_GUO = 4, K_REGULARIZED = 5 };

typedef struct {
    int32_t Q, kind;
    int32_t nx, ny, nz;
    int32_t periodic[3];
    double omega;     /* omega+ = 1/tau+ */
    double magic;     /* TRT magic / MRT odd_magic */
    double bulk_rate; /* MRT bulk rate (NaN = none) / cumulant bulk rate */
    double even_rate; /* MRT even rate (NaN = none) */
    double accel[3];  /* lattice body acceleration */
    int32_t n_port_cells;
    int32_t pad_;
    const int32_t *port_cell;
    const int32_t *port_nb;
    const int32_t *port_kind;   /* 0 velocity, 1 pressure */
    const double *port_value;   /* 3 per port cell: u_b, or rho_b in [0] */
    const uint8_t *solid;       /* N entries, may be NULL (all fluid) */
    const double *mrt_rows;     /* Q*Q, row-major rows[k][i] (from orc_mrt_basis) */
    const double *mrt_norms;    /* Q */
    const int32_t *mrt_groups;  /* Q */
    double les_constant;        /* Smagorinsky constant (psm.rs:45-48); <= 0 or NaN: no LES */
    double les_floor;           /* Smagorinsky norm_floor */
    double wale_constant;       /* WALE constant (moving/wale.rs); <= 0 or NaN: none */
    double wale_floor;          /* WALE denominator_floor */
    /* Sponge layers (moving/sponge.rs, blended after collision as field.rs:1461-1475):
     * the listed fluid cells take post (1 - sigma) + sum_parts s_l f^eq(rho_l, u_l) with
     * u_l = mean_l + sin_l cos(theta) - cos_l sin(theta).  Built by oracle/oracle.py
     * (SpongeField::new, moving/sponge.rs:112-181; targets field.rs:1015-1036). */
    int32_t n_sponge_cells;
    int32_t n_sponges;
    const int32_t *sponge_cell;    /* ascending cell indices */
    const double *sponge_sigma;    /* summed strength of each listed cell */
    const int32_t *sponge_off;     /* parts of listed cell r: [off[r], off[r+1]) */
    const int32_t *sponge_part;    /* sponge index l of each part */
    const double *sponge_weight;   /* s_l of each part */
    const double *sponge_phase;    /* [cos theta, sin theta] of each part */
    const double *sponge_target;   /* per sponge at the step's time: rho, mean[3], sin[3], cos[3] */
    /* Occupancy coupling (moving/psm.rs:198-267, cell_update without Brinkman): with
     * psm_law >= 0 the per-cell design array is the occupancy d, cells with d > 0 couple
     * to a solid of momentum occ_m (3 per cell, NULL: at rest); orc_interval_vjp writes
     * the momentum cotangent into occ_m_bar (3 per cell) when it is set. */
    int32_t psm_law;               /* -1 none, 0 Psm, 1 PsmSuperposition, 2 Brinkman */
    int32_t pad2_;
    double psm_kappa;              /* saturation width */
    const double *occ_m;
    double *occ_m_bar;
    double psm_c_alpha;            /* Brinkman law: c_alpha (CellParameters) */
    double psm_drag_shape;         /* Brinkman law: drag_shape q */
    /* Second occupancy endpoint (Occupancy::{d, m}[1]): with occ_d1 set, a step collides
     * with d = d0 (1 - theta) + d1 theta and m likewise (field.rs:1266-1275); the
     * transpose splits the cotangents between the endpoints (occ_d1_bar, occ_m1_bar). */
    const double *occ_d1;
    const double *occ_m1;
    double theta;
    double *occ_d1_bar;
    double *occ_m1_bar;
    double tau; /* relaxation time tau+ as given (solver.rs:487); 0: 1/omega */
    /* Subcycle (moving/field.rs:1499-1541, 2682-2690): with sub_g set, each coupled cell
     * adds its impulse shares g0 += (-dp/d)(1 - theta), g1 += (-dp/d) theta to sub_g
     * ([cell][6]: g0 x, y, z, g1 x, y, z); with sub_gbar set, the step transpose takes their
     * cotangents ([cell][6]). */
    double *sub_g;
    const double *sub_gbar;
} orc_problem;

/* tau+ of the problem: as given, else 1/omega (solver.rs:487 states tau, the moving solver omega). */
static double relax_tau(const orc_problem *p) { return p->tau > 0.0 ? p->tau : 1.0 / p->omega; }

static inline int is_solid(const orc_problem *p, long x) { return p->solid ? p->solid[x] != 0 : 0; }

/* ------------------------------------------------------------------------- */
/* Grid: d3q19.rs:70-128 (index, coords, wrap with rem_euclid, inside).        */
/* ------------------------------------------------------------------------- */
static inline long g_index(const orc_problem *p, long i, long j, long k) { return (i * p->ny + j) * p->nz + k; }

static inline void g_coords(const orc_problem *p, long x, long out[3]) {
    long k = x % p->nz;
    long ij = x / p->nz;
    out[0] = ij / p->ny;
    out[1] = ij % p->ny;
    out[2] = k;
}

static inline long rem_euclid(long v, long n) {
    long r = v % n;
    return r < 0 ? r + n : r;
}

static long g_wrap(const orc_problem *p, long x, const long off[3]) {
    long c[3], n[3] = {p->nx, p->ny, p->nz};
    g_coords(p, x, c);
    for (int a = 0; a < 3; a++) c[a] = rem_euclid(c[a] + off[a], n[a]);
    return g_index(p, c[0], c[1], c[2]);
}

static int g_inside(const orc_problem *p, long x, const long off[3]) {
    long c[3], n[3] = {p->nx, p->ny, p->nz};
    g_coords(p, x, c);
    for (int a = 0; a < 3; a++) {
        long v = c[a] + off[a];
        if (!(p->periodic[a] || (v >= 0 && v < n[a]))) return 0;
    }
    return 1;
}

/* Streaming::new (solver.rs:83-120): halfway bounce-back link table.
 * blocked[x,q] = solid[wrap(x-c_q)] || !inside(x-c_q); source = x*Q+opp(q) if blocked,
 * else wrap(x-c_q)*Q+q.  Solid cells have no source (NONE): they keep their previous
 * state (solver.rs:131-136). */
void orc_stream_table(const orc_problem *p, uint32_t *source) {
    const lattice_t *L = lat(p->Q);
    long n = (long)p->nx * p->ny * p->nz;
    int Q = p->Q;
    for (long x = 0; x < n; x++) {
        for (int q = 0; q < Q; q++) {
            if (is_solid(p, x)) {
                source[x * Q + q] = NONE_U32;
                continue;
            }
            long back[3] = {-L->c[q][0], -L->c[q][1], -L->c[q][2]};
            long y = g_wrap(p, x, back);
            int blocked = is_solid(p, y) || !g_inside(p, x, back);
            source[x * Q + q] = blocked ? (uint32_t)(x * Q + L->opp[q]) : (uint32_t)(y * Q + q);
        }
    }
}

/* ------------------------------------------------------------------------- */
/* Moving-lattice helpers: moving/lattice.rs:160-222.                          */
/* ------------------------------------------------------------------------- */
static inline double dot_c(const lattice_t *L, int i, const double v[3]) {
    /* v0*c0 + v1*c1 + v2*c2, left to right (lattice.rs:160-165). */
    return v[0] * L->cf[i][0] + v[1] * L->cf[i][1] + v[2] * L->cf[i][2];
}

static void moments(const lattice_t *L, const double *f, double *rho, double j[3]) {
    /* lattice.rs:168-180: rho = f0 + f1 + ...; j[d] += f_i*c_id for i >= 1. */
    double r = f[0];
    double m[3] = {0.0, 0.0, 0.0};
    for (int i = 1; i < L->Q; i++) {
        r += f[i];
        for (int d = 0; d < 3; d++) m[d] += f[i] * L->cf[i][d];
    }
    *rho = r;
    j[0] = m[0];
    j[1] = m[1];
    j[2] = m[2];
}

static void equilibrium(const lattice_t *L, double rho, const double u[3], double *eq) {
    /* lattice.rs:183-191 */
    double usq = u[0] * u[0] + u[1] * u[1] + u[2] * u[2];
    double base = 1.0 - usq * 1.5;
    for (int i = 0; i < L->Q; i++) {
        double cu = dot_c(L, i, u);
        eq[i] = rho * L->w[i] * (base + cu * 3.0 + cu * cu * 4.5);
    }
}

static void equilibrium_vjp(const lattice_t *L, double rho, const double u[3], const double *g, double *rho_bar,
                            double u_bar[3]) {
    /* lattice.rs:193-222 */
    double usq = u[0] * u[0] + u[1] * u[1] + u[2] * u[2];
    double base = 1.0 - usq * 1.5;
    double usq_bar = 0.0;
    for (int i = 0; i < L->Q; i++) {
        double cu = dot_c(L, i, u);
        double poly = base + cu * 3.0 + cu * cu * 4.5;
        double gw = g[i] * L->w[i];
        *rho_bar += gw * poly;
        double e = gw * rho;
        double cu_bar = e * (cu * 9.0 + 3.0);
        for (int d = 0; d < 3; d++)
            if (L->cf[i][d] != 0.0) u_bar[d] += cu_bar * L->cf[i][d];
        usq_bar -= e * 1.5;
    }
    for (int d = 0; d < 3; d++) u_bar[d] += usq_bar * u[d] * 2.0;
}

static void source_term(const lattice_t *L, const double u[3], const double F[3], double *src) {
    /* collision.rs:316-323 */
    double uf = u[0] * F[0] + u[1] * F[1] + u[2] * F[2];
    for (int i = 0; i < L->Q; i++) {
        double cu = dot_c(L, i, u);
        double cf = dot_c(L, i, F);
        src[i] = (cf * 3.0 + cu * cf * 9.0 - uf * 3.0) * L->w[i];
    }
}

/* odd_rate / odd_rate_derivative: collision.rs:302-313. */
double orc_odd_rate(double omega, double magic) {
    double tp = 1.0 / omega - 0.5;
    return 1.0 / (0.5 + magic / tp);
}

double orc_odd_rate_derivative(double omega, double magic) {
    double tau = 1.0 / omega;
    double tp = tau - 0.5;
    double tm = 0.5 + magic / tp;
    return -(tau * tau * magic) / (tm * tm * tp * tp);
}

/* collide_pairs: collision.rs:457-488 (TRT; BGK when om == omega). */
static void collide_pairs(const lattice_t *L, const double *f, double rho, const double u[3], const double F[3],
                          double omega, double om, double *out) {
    double usq = u[0] * u[0] + u[1] * u[1] + u[2] * u[2];
    double uf = u[0] * F[0] + u[1] * F[1] + u[2] * F[2];
    double base = 1.0 - usq * 1.5;
    double kp = 1.0 - omega * 0.5;
    double km = 1.0 - om * 0.5;
    double w0 = L->w[0];
    double eq0 = rho * w0 * base;
    out[0] = f[0] - omega * (f[0] - eq0) - kp * uf * (3.0 * w0);
    for (int i = 1; i < L->Q; i++) {
        int o = L->opp[i];
        if (o < i) continue;
        double w = L->w[i];
        double cu = dot_c(L, i, u);
        double cf = dot_c(L, i, F);
        double wr = rho * w;
        double eqp = wr * (base + cu * cu * 4.5);
        double eqm = wr * cu * 3.0;
        double sp = (cu * cf * 9.0 - uf * 3.0) * w;
        double sm = cf * (3.0 * w);
        double fp = (f[i] + f[o]) * 0.5;
        double fm = (f[i] - f[o]) * 0.5;
        double a = kp * sp - omega * (fp - eqp);
        double b = km * sm - om * (fm - eqm);
        out[i] = f[i] + a + b;
        out[o] = f[o] + a - b;
    }
}

/* collide_pairs_vjp: collision.rs:492-570.  Returns f_bar, rho_bar, u_bar, force_bar. */
static void collide_pairs_vjp(const lattice_t *L, const double *f, double rho, const double u[3], const double F[3],
                              double omega, double om, const double *g, double *f_bar, double *rho_bar_out,
                              double u_bar_out[3], double force_bar_out[3], double *omega_bar_out,
                              double *om_bar_out) {
    double usq = u[0] * u[0] + u[1] * u[1] + u[2] * u[2];
    double base = 1.0 - usq * 1.5;
    double kp = 1.0 - omega * 0.5;
    double km = 1.0 - om * 0.5;
    double rho_bar = 0.0, usq_bar = 0.0, uf_bar = 0.0;
    double u_bar[3] = {0, 0, 0}, force_bar[3] = {0, 0, 0};
    double w0 = L->w[0];
    double g0 = g[0];
    double uf = u[0] * F[0] + u[1] * F[1] + u[2] * F[2];
    double eq0 = rho * w0 * base;
    double s0 = -uf * (3.0 * w0);
    double omega_bar = 0.0, om_bar = 0.0;
    for (int i = 0; i < L->Q; i++) f_bar[i] = 0.0;
    f_bar[0] = g0 * (1.0 - omega);
    rho_bar += g0 * omega * w0 * base;
    usq_bar -= g0 * omega * rho * (1.5 * w0);
    uf_bar -= g0 * kp * (3.0 * w0);
    omega_bar -= g0 * ((f[0] - eq0) + s0 * 0.5);
    for (int i = 1; i < L->Q; i++) {
        int o = L->opp[i];
        if (o < i) continue;
        double w = L->w[i];
        double cu = dot_c(L, i, u);
        double cf = dot_c(L, i, F);
        double wr = rho * w;
        double a_bar = g[i] + g[o];
        double b_bar = g[i] - g[o];
        {
            double eqp = wr * (base + cu * cu * 4.5);
            double eqm = wr * cu * 3.0;
            double sp = (cu * cf * 9.0 - uf * 3.0) * w;
            double sm = cf * (3.0 * w);
            double fp = (f[i] + f[o]) * 0.5;
            double fm = (f[i] - f[o]) * 0.5;
            omega_bar -= a_bar * ((fp - eqp) + sp * 0.5);
            om_bar -= b_bar * ((fm - eqm) + sm * 0.5);
        }
        double fp_bar = -omega * a_bar;
        double fm_bar = -om * b_bar;
        double eqp_bar = omega * a_bar;
        double eqm_bar = om * b_bar;
        double sp_bar = kp * a_bar;
        double sm_bar = km * b_bar;
        f_bar[i] = g[i] + (fp_bar + fm_bar) * 0.5;
        f_bar[o] = g[o] + (fp_bar - fm_bar) * 0.5;
        rho_bar += eqp_bar * w * (base + cu * cu * 4.5) + eqm_bar * cu * (3.0 * w);
        usq_bar -= eqp_bar * wr * 1.5;
        double cu_bar = eqp_bar * wr * cu * 9.0 + eqm_bar * wr * 3.0 + sp_bar * cf * (9.0 * w);
        double cf_bar = sp_bar * cu * (9.0 * w) + sm_bar * (3.0 * w);
        uf_bar -= sp_bar * (3.0 * w);
        for (int d = 0; d < 3; d++) {
            if (L->cf[i][d] != 0.0) {
                u_bar[d] += cu_bar * L->cf[i][d];
                force_bar[d] += cf_bar * L->cf[i][d];
            }
        }
    }
    for (int d = 0; d < 3; d++) {
        u_bar[d] += usq_bar * u[d] * 2.0 + uf_bar * F[d];
        force_bar[d] += uf_bar * u[d];
    }
    *rho_bar_out = rho_bar;
    for (int d = 0; d < 3; d++) {
        u_bar_out[d] = u_bar[d];
        force_bar_out[d] = force_bar[d];
    }    if (omega_bar_out) *omega_bar_out = omega_bar;
    if (om_bar_out) *om_bar_out = om_bar;
}

/* ------------------------------------------------------------------------- */
/* Smagorinsky LES (moving/psm.rs:161-197): effective omega from the norm of    */
/* the non-equilibrium momentum flux, applied to BGK/TRT (omega- follows from  */
/* the effective omega+ through the magic parameter).                          */
/* ------------------------------------------------------------------------- */
static int les_on(const orc_problem *p) {
    return p->les_constant > 0.0 && (p->kind == K_BGK || p->kind == K_TRT);
}

/* pi[x][y] = sum_i (f_i - eq_i) c_ix c_iy and its squared norm, psm.rs:169-189. */
static double les_flux(const lattice_t *L, const double *f, double rho, const double u[3], double pi[3][3]) {
    double eq[MAXQ];
    equilibrium(L, rho, u, eq);
    for (int x = 0; x < 3; x++)
        for (int y = 0; y < 3; y++) pi[x][y] = 0.0;
    for (int i = 0; i < L->Q; i++) {
        double n = f[i] - eq[i];
        for (int x = 0; x < 3; x++) {
            if (L->cf[i][x] == 0.0) continue;
            for (int y = 0; y < 3; y++)
                if (L->cf[i][y] != 0.0) pi[x][y] += n * (L->cf[i][x] * L->cf[i][y]);
        }
    }
    double qn = 0.0;
    for (int x = 0; x < 3; x++)
        for (int y = 0; y < 3; y++) qn += pi[x][y] * pi[x][y];
    return qn;
}

static double smagorinsky(const orc_problem *p, const lattice_t *L, const double *f, double rho, const double u[3],
                          double omega) {
    double pi[3][3];
    double qn = les_flux(L, f, rho, u, pi);
    double delta = p->les_floor;
    double pn = sqrt(qn + delta * delta) - delta;
    double k = 18.0 * 1.4142135623730951 * p->les_constant * p->les_constant;
    double tau0 = 1.0 / omega;
    double r = sqrt(tau0 * tau0 + pn * k / rho);
    return 1.0 / ((tau0 + r) * 0.5);
}

/* Transpose of smagorinsky (psm.rs:415-473) given the cotangent of the effective
 * omega: adds to f_bar, rho_bar, u_bar (the equilibrium's inputs). */
static void smagorinsky_vjp(const orc_problem *p, const lattice_t *L, const double *f, double rho, const double u[3],
                            double omega, double omega_eff_bar, double *f_bar, double *rho_bar, double u_bar[3]) {
    double pi[3][3];
    double qn = les_flux(L, f, rho, u, pi);
    double delta = p->les_floor;
    double root = sqrt(qn + delta * delta);
    double pn = root - delta;
    double k = 18.0 * 1.4142135623730951 * p->les_constant * p->les_constant;
    double tau0 = 1.0 / omega;
    double r = sqrt(tau0 * tau0 + pn * k / rho);
    double tau_e = (tau0 + r) * 0.5;
    double tau_e_bar = -omega_eff_bar / (tau_e * tau_e);
    double r_bar = tau_e_bar * 0.5;
    double inner_bar = r_bar / (r * 2.0);
    double p_bar = inner_bar * k / rho;
    *rho_bar -= inner_bar * pn * k / (rho * rho);
    double qn_bar = p_bar / (root * 2.0);
    double n_bar[MAXQ], neg[MAXQ];
    for (int i = 0; i < L->Q; i++) {
        double acc = 0.0;
        for (int x = 0; x < 3; x++) {
            if (L->cf[i][x] == 0.0) continue;
            for (int y = 0; y < 3; y++)
                if (L->cf[i][y] != 0.0) acc += pi[x][y] * (L->cf[i][x] * L->cf[i][y]);
        }
        n_bar[i] = acc * qn_bar * 2.0;
        f_bar[i] += n_bar[i];
        neg[i] = -n_bar[i];
    }
    equilibrium_vjp(L, rho, u, neg, rho_bar, u_bar);
}

/* ------------------------------------------------------------------------- */
/* MRT: MomentBasis::new (collision.rs:87-153), forward/inverse/transpose       */
/* (:156-197), relax MRT arm (:349-375), relax_transpose MRT arm (:403-418).    */
/* Groups: 0 Mass, 1 Momentum, 2 Bulk, 3 Shear, 4 Odd, 5 EvenHigh.              */
/* ------------------------------------------------------------------------- */
static double ipow(double v, int e) {
    /* f64::powi with small non-negative integer exponents: repeated products. */
    double r = 1.0;
    for (int k = 0; k < e; k++) r *= v;
    return r;
}

int orc_mrt_basis(int Q, double *rows_out, double *norms_out, int32_t *groups_out) {
    const lattice_t *L = lat(Q);
    double cand[64][MAXQ];
    int cgroup[64];
    int nc = 0;
    for (int i = 0; i < Q; i++) cand[nc][i] = 1.0;
    cgroup[nc++] = 0;
    for (int d = 0; d < 3; d++) {
        for (int i = 0; i < Q; i++) cand[nc][i] = L->cf[i][d];
        cgroup[nc++] = 1;
    }
    for (int i = 0; i < Q; i++) {
        const double *v = L->cf[i];
        cand[nc][i] = v[0] * v[0] + v[1] * v[1] + v[2] * v[2];
    }
    cgroup[nc++] = 2;
    for (int i = 0; i < Q; i++) {
        const double *v = L->cf[i];
        cand[nc][i] = 2.0 * v[0] * v[0] - v[1] * v[1] - v[2] * v[2];
    }
    cgroup[nc++] = 3;
    for (int i = 0; i < Q; i++) {
        const double *v = L->cf[i];
        cand[nc][i] = v[1] * v[1] - v[2] * v[2];
    }
    cgroup[nc++] = 3;
    for (int i = 0; i < Q; i++) cand[nc][i] = L->cf[i][0] * L->cf[i][1];
    cgroup[nc++] = 3;
    for (int i = 0; i < Q; i++) cand[nc][i] = L->cf[i][0] * L->cf[i][2];
    cgroup[nc++] = 3;
    for (int i = 0; i < Q; i++) cand[nc][i] = L->cf[i][1] * L->cf[i][2];
    cgroup[nc++] = 3;
    for (int degree = 3; degree <= 6; degree++)
        for (int a = 0; a <= 2; a++)
            for (int b = 0; b <= 2; b++) {
                int cc = degree - a - b;
                if (cc < 0 || cc > 2) continue;
                for (int i = 0; i < Q; i++)
                    cand[nc][i] = ipow(L->cf[i][0], a) * ipow(L->cf[i][1], b) * ipow(L->cf[i][2], cc);
                cgroup[nc++] = (degree % 2 == 1) ? 4 : 5;
            }
    int nrows = 0;
    double rows[MAXQ][MAXQ];
    double norms[MAXQ];
    for (int ci = 0; ci < nc; ci++) {
        if (nrows == Q) break;
        const double *pv = cand[ci];
        double scale = 0.0;
        for (int i = 0; i < Q; i++) scale += L->w[i] * pv[i] * pv[i];
        if (scale <= 0.0) continue;
        double v[MAXQ];
        for (int i = 0; i < Q; i++) v[i] = pv[i];
        for (int k = 0; k < nrows; k++) {
            double ip = 0.0;
            for (int i = 0; i < Q; i++) ip += L->w[i] * pv[i] * rows[k][i];
            double coef = ip / norms[k];
            for (int i = 0; i < Q; i++) v[i] -= coef * rows[k][i];
        }
        double nn = 0.0;
        for (int i = 0; i < Q; i++) nn += L->w[i] * v[i] * v[i];
        if (nn <= 1e-10 * scale) continue;
        norms[nrows] = nn;
        groups_out[nrows] = cgroup[ci];
        for (int i = 0; i < Q; i++) rows[nrows][i] = v[i];
        nrows++;
    }
    if (nrows != Q) return -1;
    for (int k = 0; k < Q; k++) {
        norms_out[k] = norms[k];
        for (int i = 0; i < Q; i++) rows_out[k * Q + i] = rows[k][i];
    }
    return 0;
}

/* MRT arm of relax (collision.rs:349-375) and the Regularized arm (:376-394): bulk and
 * shear moments relax with omega, every other moment is removed (rate 1). */
static void mrt_rates(const orc_problem *p, double *rate) {
    double om = orc_odd_rate(p->omega, p->magic);
    if (p->kind == K_REGULARIZED) {
        for (int k = 0; k < p->Q; k++) rate[k] = (p->mrt_groups[k] == 2 || p->mrt_groups[k] == 3) ? p->omega : 1.0;
        return;
    }
    for (int k = 0; k < p->Q; k++) {
        switch (p->mrt_groups[k]) {
        case 0:
        case 3: rate[k] = p->omega; break;
        case 1:
        case 4: rate[k] = om; break;
        case 2: rate[k] = isnan(p->bulk_rate) ? p->omega : p->bulk_rate; break;
        default: rate[k] = isnan(p->even_rate) ? p->omega : p->even_rate; break;
        }
    }
}

static void mrt_forward(const orc_problem *p, const double *v, double *m) {
    int Q = p->Q;
    for (int k = 0; k < Q; k++) {
        double acc = 0.0;
        for (int i = 0; i < Q; i++) {
            double r = p->mrt_rows[k * Q + i];
            if (r != 0.0) acc += v[i] * r;
        }
        m[k] = acc;
    }
}

static void mrt_inverse(const orc_problem *p, const lattice_t *L, const double *m, double *out) {
    int Q = p->Q;
    double scaled[MAXQ];
    for (int k = 0; k < Q; k++) scaled[k] = m[k] / p->mrt_norms[k];
    for (int i = 0; i < Q; i++) {
        double acc = 0.0;
        for (int k = 0; k < Q; k++) {
            double r = p->mrt_rows[k * Q + i];
            if (r != 0.0) acc += scaled[k] * r;
        }
        out[i] = acc * L->w[i];
    }
}

static void mrt_transpose(const orc_problem *p, const double *m, double *out) {
    int Q = p->Q;
    for (int i = 0; i < Q; i++) {
        double acc = 0.0;
        for (int k = 0; k < Q; k++) {
            double r = p->mrt_rows[k * Q + i];
            if (r != 0.0) acc += m[k] * r;
        }
        out[i] = acc;
    }
}

/* Generic (MRT) collide: collision.rs:449-453 with relax MRT arm :349-375. */
static void collide_mrt(const orc_problem *p, const lattice_t *L, const double *f, double rho, const double u[3],
                        const double F[3], double *out) {
    int Q = p->Q;
    double eq[MAXQ], src[MAXQ], y[MAXQ], m[MAXQ], rm[MAXQ], r[MAXQ], rate[MAXQ];
    equilibrium(L, rho, u, eq);
    source_term(L, u, F, src);
    for (int i = 0; i < Q; i++) y[i] = f[i] - eq[i] + src[i] * 0.5;
    mrt_forward(p, y, m);
    mrt_rates(p, rate);
    for (int k = 0; k < Q; k++) rm[k] = m[k] * rate[k];
    mrt_inverse(p, L, rm, r);
    for (int i = 0; i < Q; i++) out[i] = f[i] - r[i] + src[i];
}

/* Generic collide_vjp (MRT arm): collision.rs:597-633, relax_transpose :403-418. */
static void collide_mrt_vjp(const orc_problem *p, const lattice_t *L, const double *f, double rho, const double u[3],
                            const double F[3], const double *g, double *f_bar, double *rho_bar_out,
                            double u_bar_out[3], double force_bar_out[3]) {
    (void)f;
    int Q = p->Q;
    double wg[MAXQ], m[MAXQ], scaled[MAXQ], rt[MAXQ], rate[MAXQ];
    for (int i = 0; i < Q; i++) wg[i] = g[i] * L->w[i];
    mrt_forward(p, wg, m);
    mrt_rates(p, rate);
    for (int k = 0; k < Q; k++) scaled[k] = m[k] * rate[k] / p->mrt_norms[k];
    mrt_transpose(p, scaled, rt);
    double src_bar[MAXQ];
    for (int i = 0; i < Q; i++) {
        f_bar[i] = g[i] - rt[i];
        src_bar[i] = g[i] - rt[i] * 0.5;
    }
    double rho_bar = 0.0, u_bar[3] = {0, 0, 0};
    equilibrium_vjp(L, rho, u, rt, &rho_bar, u_bar);
    double force_bar[3] = {0, 0, 0}, uf_bar = 0.0;
    for (int i = 0; i < Q; i++) {
        double h = src_bar[i] * L->w[i];
        double cu = dot_c(L, i, u);
        double cf = dot_c(L, i, F);
        double cf_bar = h * (cu * 9.0 + 3.0);
        double cu_bar = h * cf * 9.0;
        uf_bar -= h * 3.0;
        for (int d = 0; d < 3; d++) {
            if (L->cf[i][d] != 0.0) {
                force_bar[d] += cf_bar * L->cf[i][d];
                u_bar[d] += cu_bar * L->cf[i][d];
            }
        }
    }
    for (int d = 0; d < 3; d++) {
        u_bar[d] += uf_bar * F[d];
        force_bar[d] += uf_bar * u[d];
    }
    *rho_bar_out = rho_bar;
    for (int d = 0; d < 3; d++) {
        u_bar_out[d] = u_bar[d];
        force_bar_out[d] = force_bar[d];
    }
}

/* ------------------------------------------------------------------------- */
/* Cumulant (D3Q27): collision.rs:636-752, with forward-mode tangents for the  */
/* VJP as collision.rs:755-788 does (one tangent direction per pass).          */
/* ------------------------------------------------------------------------- */
typedef struct {
    double v, d;
} dual;

static inline dual dv(double v) { dual r = {v, 0.0}; return r; }
static inline dual dadd(dual a, dual b) { dual r = {a.v + b.v, a.d + b.d}; return r; }
static inline dual dsub(dual a, dual b) { dual r = {a.v - b.v, a.d - b.d}; return r; }
static inline dual dmul(dual a, dual b) { dual r = {a.v * b.v, a.d * b.v + a.v * b.d}; return r; }
static inline dual dmuls(dual a, double s) { dual r = {a.v * s, a.d * s}; return r; }
static inline dual ddiv(dual a, dual b) {
    /* implexity-ad dual.rs:136-146: eps = a'/b - b'*q*inv with inv = 1/b, q = a*inv. */
    double inv = 1.0 / b.v;
    double q = a.v * inv;
    dual r = {a.v / b.v, a.d / b.v - b.d * q * inv};
    return r;
}
static inline dual ddivs(dual a, double s) { dual r = {a.v / s, a.d / s}; return r; }

/* Wick pairings per cube index (collision.rs:212-239,276-285). */
static int wick_n[27];
static int wick_pairs[27][16][3][2];
static int wick_ready = 0;

static int matchings_rec(const int *items, int n, int out[][3][2], int depth_base, int cap) {
    /* Returns matchings of items[0..n) in the order of collision.rs:216-239. */
    if (n == 0) return 1; /* one empty matching; caller handles */
    int count = 0;
    int first = items[0];
    for (int j = 1; j < n; j++) {
        int rest[8], nr = 0;
        for (int t = 1; t < n; t++)
            if (t != j) rest[nr++] = items[t];
        int partner = items[j];
        int lo = first < partner ? first : partner, hi = first < partner ? partner : first;
        int sub[16][3][2];
        int ns;
        if (nr == 0) {
            ns = 1;
        } else {
            ns = matchings_rec(rest, nr, sub, depth_base + 1, 16);
        }
        for (int s = 0; s < ns; s++) {
            if (count >= cap) return count;
            out[count][0][0] = lo;
            out[count][0][1] = hi;
            for (int pidx = 0; pidx < nr / 2; pidx++) {
                out[count][pidx + 1][0] = sub[s][pidx][0];
                out[count][pidx + 1][1] = sub[s][pidx][1];
            }
            count++;
        }
    }
    return count;
}

static void init_wick(void) {
    if (wick_ready) return;
    for (int k = 0; k < 27; k++) {
        int e[3] = {k / 9, (k / 3) % 3, k % 3};
        int axes[8], na = 0;
        for (int d = 0; d < 3; d++)
            for (int t = 0; t < e[d]; t++) axes[na++] = d;
        if (na % 2 == 1 || na < 4) {
            wick_n[k] = 0;
        } else {
            wick_n[k] = matchings_rec(axes, na, wick_pairs[k], 0, 16);
        }
    }
    wick_ready = 1;
}

int orc_wick_table(int32_t *counts, int32_t *pairs) {
    /* counts[27]; pairs[27][16][3][2] (unused entries untouched). */
    init_wick();
    for (int k = 0; k < 27; k++) {
        counts[k] = wick_n[k];
        for (int m = 0; m < wick_n[k]; m++) {
            int np = 0;
            int e[3] = {k / 9, (k / 3) % 3, k % 3};
            np = (e[0] + e[1] + e[2]) / 2;
            for (int t = 0; t < np; t++) {
                pairs[((k * 16 + m) * 3 + t) * 2 + 0] = wick_pairs[k][m][t][0];
                pairs[((k * 16 + m) * 3 + t) * 2 + 1] = wick_pairs[k][m][t][1];
            }
        }
    }
    return 0;
}

static void cumulant_dual(const dual *f, dual rho, const dual u[3], const dual F[3], dual omega, double bulk,
                          dual *out) {
    const lattice_t *L = lat(27);
    init_wick();
    dual cube[27];
    for (int i = 0; i < 27; i++) {
        int k = (L->c[i][0] + 1) * 9 + (L->c[i][1] + 1) * 3 + (L->c[i][2] + 1);
        cube[k] = f[i];
    }
    const int stride[3] = {9, 3, 1};
    const int fwd_axes[3] = {2, 1, 0};
    for (int ai = 0; ai < 3; ai++) {
        int axis = fwd_axes[ai];
        int s = stride[axis];
        for (int base = 0; base < 27; base++) {
            if ((base / s) % 3 != 0) continue;
            dual fm = cube[base], f0 = cube[base + s], fp = cube[base + 2 * s];
            dual k0 = dadd(dadd(fm, f0), fp);
            dual diff = dsub(fp, fm);
            dual ua = u[axis];
            dual k1 = dsub(diff, dmul(ua, k0));
            dual k2 = dadd(dsub(dadd(fp, fm), dmuls(dmul(ua, diff), 2.0)), dmul(dmul(ua, ua), k0));
            cube[base] = k0;
            cube[base + s] = k1;
            cube[base + 2 * s] = k2;
        }
    }
#define IDX(a, b, c) ((a) * 9 + (b) * 3 + (c))
    dual kxx = cube[IDX(2, 0, 0)], kyy = cube[IDX(0, 2, 0)], kzz = cube[IDX(0, 0, 2)];
    dual one_minus = dsub(dv(1.0), omega);
    dual dxy = dmul(dsub(kxx, kyy), one_minus);
    dual dxz = dmul(dsub(kxx, kzz), one_minus);
    dual trace = dadd(dadd(kxx, kyy), kzz);
    trace = dadd(trace, dmuls(dsub(rho, trace), bulk));
    dual post_xx = ddivs(dadd(dadd(trace, dxy), dxz), 3.0);
    dual post_yy = ddivs(dadd(dsub(trace, dmuls(dxy, 2.0)), dxz), 3.0);
    dual post_zz = ddivs(dsub(dadd(trace, dxy), dmuls(dxz, 2.0)), 3.0);
    dual post_xy = dmul(cube[IDX(1, 1, 0)], one_minus);
    dual post_xz = dmul(cube[IDX(1, 0, 1)], one_minus);
    dual post_yz = dmul(cube[IDX(0, 1, 1)], one_minus);
    dual sigma[3][3] = {{ddiv(post_xx, rho), ddiv(post_xy, rho), ddiv(post_xz, rho)},
                        {ddiv(post_xy, rho), ddiv(post_yy, rho), ddiv(post_yz, rho)},
                        {ddiv(post_xz, rho), ddiv(post_yz, rho), ddiv(post_zz, rho)}};
    dual post[27];
    for (int k = 0; k < 27; k++) {
        int e[3] = {k / 9, (k / 3) % 3, k % 3};
        int order = e[0] + e[1] + e[2];
        if (order == 0) {
            post[k] = rho;
        } else if (order == 1) {
            int axis = e[0] == 1 ? 0 : (e[1] == 1 ? 1 : 2);
            post[k] = dmuls(F[axis], 0.5);
        } else if (order == 2) {
            if (e[0] == 2) post[k] = post_xx;
            else if (e[1] == 2) post[k] = post_yy;
            else if (e[2] == 2) post[k] = post_zz;
            else if (e[0] == 1 && e[1] == 1) post[k] = post_xy;
            else if (e[0] == 1 && e[2] == 1) post[k] = post_xz;
            else post[k] = post_yz;
        } else {
            dual acc = dv(0.0);
            int np = order / 2;
            for (int m = 0; m < wick_n[k]; m++) {
                dual pr = dv(1.0);
                for (int t = 0; t < np; t++) pr = dmul(pr, sigma[wick_pairs[k][m][t][0]][wick_pairs[k][m][t][1]]);
                acc = dadd(acc, pr);
            }
            post[k] = dmul(acc, rho);
        }
    }
#undef IDX
    for (int axis = 0; axis < 3; axis++) {
        int s = stride[axis];
        for (int base = 0; base < 27; base++) {
            if ((base / s) % 3 != 0) continue;
            dual k0 = post[base], k1 = post[base + s], k2 = post[base + 2 * s];
            dual ua = u[axis];
            dual m1 = dadd(k1, dmul(ua, k0));
            dual m2 = dadd(dadd(k2, dmuls(dmul(ua, k1), 2.0)), dmul(dmul(ua, ua), k0));
            post[base] = dmuls(dsub(m2, m1), 0.5);
            post[base + s] = dsub(k0, m2);
            post[base + 2 * s] = dmuls(dadd(m2, m1), 0.5);
        }
    }
    for (int i = 0; i < 27; i++) {
        int k = (L->c[i][0] + 1) * 9 + (L->c[i][1] + 1) * 3 + (L->c[i][2] + 1);
        out[i] = post[k];
    }
}

static void collide_cumulant(const double *f, double rho, const double u[3], const double F[3], double omega,
                             double bulk, double *out) {
    dual fd[27], ud[3], Fd[3], od[27];
    for (int i = 0; i < 27; i++) fd[i] = dv(f[i]);
    for (int d = 0; d < 3; d++) {
        ud[d] = dv(u[d]);
        Fd[d] = dv(F[d]);
    }
    cumulant_dual(fd, dv(rho), ud, Fd, dv(omega), bulk, od);
    for (int i = 0; i < 27; i++) out[i] = od[i].v;
}

/* cumulant_vjp (collision.rs:755-788): tangent k of the 34 inputs (27 f, rho, 3 u,
 * 3 F) contracted with g.  The omega tangent (index 34) is not needed here. */
static void collide_cumulant_vjp(const double *f, double rho, const double u[3], const double F[3], double omega,
                                 double bulk, const double *g, double *f_bar, double *rho_bar, double u_bar[3],
                                 double force_bar[3]) {
    double bar[34];
    for (int t = 0; t < 34; t++) {
        dual fd[27], ud[3], Fd[3], od[27];
        for (int i = 0; i < 27; i++) fd[i] = (dual){f[i], t == i ? 1.0 : 0.0};
        dual rd = {rho, t == 27 ? 1.0 : 0.0};
        for (int d = 0; d < 3; d++) {
            ud[d] = (dual){u[d], t == 28 + d ? 1.0 : 0.0};
            Fd[d] = (dual){F[d], t == 31 + d ? 1.0 : 0.0};
        }
        cumulant_dual(fd, rd, ud, Fd, dv(omega), bulk, od);
        double acc = 0.0;
        for (int i = 0; i < 27; i++)
            if (g[i] != 0.0) acc += g[i] * od[i].d;
        bar[t] = acc;
    }
    for (int i = 0; i < 27; i++) f_bar[i] = bar[i];
    *rho_bar = bar[27];
    for (int d = 0; d < 3; d++) {
        u_bar[d] = bar[28 + d];
        force_bar[d] = bar[31 + d];
    }
}

/* ------------------------------------------------------------------------- */
/* Stationary BGK-Guo (solver.rs:189-196,214-285,298-322).                      */
/* ------------------------------------------------------------------------- */
static void collide_bgk_guo(const double *f, double alpha, const double a[3], double tau, double *out) {
    const lattice_t *L = lat(19);
    double rho = f[0];
    for (int q = 1; q < 19; q++) rho += f[q];
    double mx = 0.0, my = 0.0, mz = 0.0;
    for (int q = 0; q < 19; q++) {
        mx += f[q] * L->cf[q][0];
        my += f[q] * L->cf[q][1];
        mz += f[q] * L->cf[q][2];
    }
    double s = alpha * 0.5 + 1.0;
    double u[3] = {(mx / rho + 0.5 * a[0]) / s, (my / rho + 0.5 * a[1]) / s, (mz / rho + 0.5 * a[2]) / s};
    double force[3] = {rho * (a[0] - alpha * u[0]), rho * (a[1] - alpha * u[1]), rho * (a[2] - alpha * u[2])};
    double uf = u[0] * force[0] + u[1] * force[1] + u[2] * force[2];
    double usq = u[0] * u[0] + u[1] * u[1] + u[2] * u[2];
    double k = 1.0 - 0.5 / tau;
    for (int q = 0; q < 19; q++) {
        double cu = 0.0 + u[0] * L->cf[q][0] + u[1] * L->cf[q][1] + u[2] * L->cf[q][2];
        double cf = 0.0 + force[0] * L->cf[q][0] + force[1] * L->cf[q][1] + force[2] * L->cf[q][2];
        double eq = rho * L->w[q] * (1.0 + cu * 3.0 + cu * cu * 4.5 - usq * 1.5);
        double source = k * L->w[q] * (cf * 3.0 - uf * 3.0 + cu * cf * 9.0);
        out[q] = f[q] - (f[q] - eq) / tau + source;
    }
}

/* ports.rs dot_c / moments (used by collide_cell_vjp via macroscopic). */
static double st_dot_c(int q, const double u[3]) {
    double acc = 0.0;
    int first = 1;
    for (int d = 0; d < 3; d++) {
        double term;
        if (C19[q][d] == 0) continue;
        term = C19[q][d] == 1 ? u[d] : -u[d];
        acc = first ? term : acc + term;
        first = 0;
    }
    return first ? 0.0 : acc;
}

static void st_moments(const double *f, double *rho, double m[3]) {
    double r = f[0];
    for (int q = 1; q < 19; q++) r += f[q];
    for (int d = 0; d < 3; d++) {
        double acc = 0.0;
        for (int q = 0; q < 19; q++) {
            if (C19[q][d] == 1) acc += f[q];
            else if (C19[q][d] == -1) acc -= f[q];
        }
        m[d] = acc;
    }
    *rho = r;
}

/* collide_cell_vjp: solver.rs:214-285 (returns f_bar, alpha_bar; tau_bar dropped). */
static void collide_bgk_guo_vjp(const double *f, double alpha, const double a[3], double tau, const double *g,
                                double *f_bar, double *alpha_bar_out) {
    const lattice_t *L = lat(19);
    double rho, m[3];
    st_moments(f, &rho, m);
    double s = 1.0 + 0.5 * alpha;
    double u[3], force[3];
    for (int d = 0; d < 3; d++) u[d] = (m[d] / rho + 0.5 * a[d]) / s;
    for (int d = 0; d < 3; d++) force[d] = rho * (a[d] - alpha * u[d]);
    double usq = u[0] * u[0] + u[1] * u[1] + u[2] * u[2];
    double k = 1.0 - 0.5 / tau;
    double rho_bar = 0.0, uf_bar = 0.0, usq_bar = 0.0;
    double u_bar[3] = {0, 0, 0}, force_bar[3] = {0, 0, 0};
    for (int q = 0; q < 19; q++) f_bar[q] = 0.0;
    for (int q = 0; q < 19; q++) {
        double gq = g[q];
        if (gq == 0.0) continue;
        double cu = st_dot_c(q, u);
        double cf = st_dot_c(q, force);
        double poly = 1.0 + 3.0 * cu + 4.5 * cu * cu - 1.5 * usq;
        f_bar[q] += gq * (1.0 - 1.0 / tau);
        double eq_bar = gq / tau;
        double h = gq * L->w[q] * k;
        double cf_bar = h * (3.0 + 9.0 * cu);
        double cu_bar = h * 9.0 * cf;
        uf_bar -= 3.0 * h;
        rho_bar += eq_bar * L->w[q] * poly;
        double e = eq_bar * rho * L->w[q];
        cu_bar += e * (3.0 + 9.0 * cu);
        usq_bar -= 1.5 * e;
        for (int d = 0; d < 3; d++) {
            double cd = L->cf[q][d];
            u_bar[d] += cu_bar * cd;
            force_bar[d] += cf_bar * cd;
        }
    }
    for (int d = 0; d < 3; d++) {
        u_bar[d] += usq_bar * 2.0 * u[d] + uf_bar * force[d];
        force_bar[d] += uf_bar * u[d];
    }
    double alpha_bar = 0.0;
    for (int d = 0; d < 3; d++) {
        rho_bar += force_bar[d] * (a[d] - alpha * u[d]);
        alpha_bar -= force_bar[d] * rho * u[d];
        u_bar[d] -= force_bar[d] * rho * alpha;
    }
    double m_bar[3], s_bar = 0.0;
    for (int d = 0; d < 3; d++) {
        m_bar[d] = u_bar[d] / (s * rho);
        rho_bar -= u_bar[d] * m[d] / (s * rho * rho);
        s_bar -= u_bar[d] * u[d] / s;
    }
    alpha_bar += 0.5 * s_bar;
    for (int q = 0; q < 19; q++) {
        double v = f_bar[q] + rho_bar;
        for (int d = 0; d < 3; d++) v += m_bar[d] * L->cf[q][d];
        f_bar[q] = v;
    }
    *alpha_bar_out = alpha_bar;
}

/* ------------------------------------------------------------------------- */
/* One cell: the Brinkman branch of psm.rs::cell_update (psm.rs:213-226) with   */
/* zero solid momentum (static design), then Collision::collide.               */
/* ------------------------------------------------------------------------- */
static void brinkman_macro(const lattice_t *L, const double *f, double alpha, const double a[3], double *rho,
                           double j[3], double *den, double sv[3], double u[3], double F[3]) {
    moments(L, f, rho, j);
    *den = 1.0 + alpha * 0.5;
    for (int d = 0; d < 3; d++) {
        sv[d] = j[d] / *rho + a[d] * 0.5; /* + gamma*m*0.5 with m = 0 */
        u[d] = sv[d] / *den;
    }
    for (int d = 0; d < 3; d++) {
        double z = -(alpha * u[d]); /* gamma*m - alpha*u with m = 0 */
        F[d] = *rho * (a[d] + z);
    }
}

/* One cell with omega+ given (rate < 0: the handle's omega, through the Smagorinsky
 * closure when it is on). */
static void cell_collide_rate(const orc_problem *p, const double *f, double alpha, double rate, double *out) {
    const lattice_t *L = lat(p->Q);
    if (p->kind == K_BGK_GUO) {
        collide_bgk_guo(f, alpha, p->accel, relax_tau(p), out);
        return;
    }
    double rho, j[3], den, sv[3], u[3], F[3];
    brinkman_macro(L, f, alpha, p->accel, &rho, j, &den, sv, u, F);
    const double om_eff = rate >= 0.0 ? rate : (les_on(p) ? smagorinsky(p, L, f, rho, u, p->omega) : p->omega);
    switch (p->kind) {
    case K_BGK: collide_pairs(L, f, rho, u, F, om_eff, om_eff, out); break;
    case K_TRT: collide_pairs(L, f, rho, u, F, om_eff, orc_odd_rate(om_eff, p->magic), out); break;
    case K_MRT:
    case K_REGULARIZED: collide_mrt(p, L, f, rho, u, F, out); break;
    case K_CUMULANT: collide_cumulant(f, rho, u, F, p->omega, isnan(p->bulk_rate) ? 1.0 : p->bulk_rate, out); break;
    }
}

void orc_cell_collide(const orc_problem *p, const double *f, double alpha, double *out) {
    cell_collide_rate(p, f, alpha, -1.0, out);
}

/* VJP of one cell w.r.t. (f, alpha): psm.rs::cell_update_vjp Brinkman branch
 * (psm.rs:301-330 and :455-469) with m = 0, dp_bar = 0; rate as in cell_collide_rate,
 * and with rate >= 0 also the cotangent of omega+ in *rate_bar (BGK/TRT). */
static void cell_vjp_rate(const orc_problem *p, const double *f, double alpha, double rate, const double *g,
                          double *f_bar, double *alpha_bar_out, double *rate_bar);

void orc_cell_vjp(const orc_problem *p, const double *f, double alpha, const double *g, double *f_bar,
                  double *alpha_bar_out) {
    cell_vjp_rate(p, f, alpha, -1.0, g, f_bar, alpha_bar_out, NULL);
}

static void cell_vjp_rate(const orc_problem *p, const double *f, double alpha, double rate, const double *g,
                          double *f_bar, double *alpha_bar_out, double *rate_bar) {
    const lattice_t *L = lat(p->Q);
    if (p->kind == K_BGK_GUO) {
        collide_bgk_guo_vjp(f, alpha, p->accel, relax_tau(p), g, f_bar, alpha_bar_out);
        return;
    }
    int Q = p->Q;
    const double *a = p->accel;
    double rho, j[3], den, sv[3], u[3], F[3];
    brinkman_macro(L, f, alpha, a, &rho, j, &den, sv, u, F);
    double cb_f[MAXQ], cb_rho, cb_u[3], cb_F[3];
    const double om_eff = rate >= 0.0 ? rate : (les_on(p) ? smagorinsky(p, L, f, rho, u, p->omega) : p->omega);
    double ob = 0.0, omb = 0.0, dom = 1.0;
    switch (p->kind) {
    case K_BGK: collide_pairs_vjp(L, f, rho, u, F, om_eff, om_eff, g, cb_f, &cb_rho, cb_u, cb_F, &ob, &omb); break;
    case K_TRT:
        collide_pairs_vjp(L, f, rho, u, F, om_eff, orc_odd_rate(om_eff, p->magic), g, cb_f, &cb_rho, cb_u, cb_F,
                          &ob, &omb);
        dom = orc_odd_rate_derivative(om_eff, p->magic);
        break;
    case K_MRT:
    case K_REGULARIZED: collide_mrt_vjp(p, L, f, rho, u, F, g, cb_f, &cb_rho, cb_u, cb_F); break;
    default:
        collide_cumulant_vjp(f, rho, u, F, p->omega, isnan(p->bulk_rate) ? 1.0 : p->bulk_rate, g, cb_f, &cb_rho,
                             cb_u, cb_F);
        break;
    }
    double rho_bar = 0.0 + cb_rho;
    double u_bar[3] = {0.0 + cb_u[0], 0.0 + cb_u[1], 0.0 + cb_u[2]};
    double alpha_bar = 0.0;
    for (int d = 0; d < 3; d++) {
        double z = -(alpha * u[d]);
        rho_bar += cb_F[d] * (a[d] + z) + 0.0 * z;
        double zb = (cb_F[d] + 0.0) * rho;
        alpha_bar -= zb * u[d];
        u_bar[d] -= zb * alpha;
    }
    if (rate_bar) *rate_bar = ob + omb * dom;
    if (rate < 0.0 && les_on(p))  /* psm.rs:415-473: the effective omega's cotangent into f, rho, u */
        smagorinsky_vjp(p, L, f, rho, u, p->omega, ob + omb * dom, cb_f, &rho_bar, u_bar);
    double j_bar[3];
    for (int d = 0; d < 3; d++) {
        double s_bar = u_bar[d] / den;
        alpha_bar -= u_bar[d] * sv[d] / (den * den) * 0.5;
        j_bar[d] = s_bar / rho;
        rho_bar -= s_bar * j[d] / (rho * rho);
    }
    for (int i = 0; i < Q; i++) {
        double v = cb_f[i] + rho_bar;
        for (int d = 0; d < 3; d++)
            if (L->cf[i][d] != 0.0) v += j_bar[d] * L->cf[i][d];
        f_bar[i] = v;
    }
    *alpha_bar_out = alpha_bar;
}

/* ------------------------------------------------------------------------- */
/* NEE ports: moving/boundary.rs:445-506 (port_populations and its VJP).       */
/* ------------------------------------------------------------------------- */
/* ------------------------------------------------------------------------- */
/* Occupancy coupling: saturate / saturate_derivative / beta (psm.rs:96-152) and the   */
/* Psm / PsmSuperposition branches of cell_update (psm.rs:228-266) and of          */
/* cell_update_vjp (psm.rs:331-415, 460-510).                                       */
/* ------------------------------------------------------------------------- */
static double psm_saturate(double d, double kappa) {
    double lo = 1.0 - kappa;
    if (d <= lo) return d;
    if (d >= 1.0 + kappa) return 1.0;
    double t = (d - lo) / (2.0 * kappa);
    double t2 = t * t;
    double h = t - t2 * t + t2 * t2 * 0.5;
    return lo + h * (2.0 * kappa);
}

static double psm_saturate_derivative(double d, double kappa) {
    double lo = 1.0 - kappa;
    if (d <= lo) return 1.0;
    if (d >= 1.0 + kappa) return 0.0;
    double t = (d - lo) / (2.0 * kappa);
    double s = 1.0 - t;
    return s * s * (t * 2.0 + 1.0);
}

/* (B / d, dB/dd / ..., dB/dt / d) as beta() returns them: B / d, d(B/d)/dd, d(B/d)/dt. */
static void psm_beta(double d, double t, double kappa, double *b_over_d, double *beta_d, double *beta_t) {
    if (d <= 1.0 - kappa) {
        double den = 1.0 - d + t;
        double den2 = den * den;
        *b_over_d = t / den;
        *beta_d = t / den2;
        *beta_t = (1.0 - d) / den2;
    } else {
        double e = psm_saturate(d, kappa);
        double de = psm_saturate_derivative(d, kappa);
        double den = 1.0 - e + t;
        double bf = e * t / den;
        double den2 = den * den;
        double bf_e = t * (t + 1.0) / den2;
        double bf_t = e * (1.0 - e) / den2;
        *b_over_d = bf / d;
        *beta_d = bf_e * de / d - bf / (d * d);
        *beta_t = bf_t / d;
    }
}

void orc_psm_beta(double d, double t, double kappa, double *out3) { psm_beta(d, t, kappa, out3, out3 + 1, out3 + 2); }

/* resistance (psm.rs:143-160): alpha = c_alpha e / ((q + 1) - e), gamma (= alpha / d) and
 * the partials [d alpha/d d, d alpha/d c_alpha, d gamma/d d, d gamma/d c_alpha]. */
static void psm_resistance(double d, double c_alpha, double q, double kappa, double *alpha, double *gamma,
                           double partials[4]) {
    double e = psm_saturate(d, kappa);
    double de = psm_saturate_derivative(d, kappa);
    double den = 1.0 * (q + 1.0) - e;
    double a = c_alpha * e / den;
    double a_e = c_alpha * (q + 1.0) / (den * den);
    double a_c = e / den;
    *alpha = a;
    partials[0] = a_e * de;
    partials[1] = a_c;
    if (d <= 1.0 - kappa) {
        *gamma = c_alpha / den;
        partials[2] = c_alpha / (den * den);
        partials[3] = 1.0 / den;
    } else {
        *gamma = a / d;
        partials[2] = a_e * de / d - a / (d * d);
        partials[3] = a_c / d;
    }
}

void orc_psm_resistance(double d, double c_alpha, double q, double kappa, double *out6) {
    psm_resistance(d, c_alpha, q, kappa, out6, out6 + 1, out6 + 2);
}

/* The Brinkman branch of cell_update with a moving solid (psm.rs:213-226):
 * u = (j/rho + a/2 + gamma m/2) / (1 + alpha/2), z = gamma m - alpha u, force rho (a + z),
 * dp = rho z. */
static void brinkman_moving_macro(const lattice_t *L, const double *f, double alpha, double gamma, const double a[3],
                                  const double m[3], double *rho, double j[3], double *den, double sv[3], double u[3],
                                  double F[3], double z[3]) {
    moments(L, f, rho, j);
    *den = 1.0 + alpha * 0.5;
    for (int d = 0; d < 3; d++) {
        sv[d] = j[d] / *rho + a[d] * 0.5 + gamma * m[d] * 0.5;
        u[d] = sv[d] / *den;
    }
    for (int d = 0; d < 3; d++) {
        z[d] = gamma * m[d] - alpha * u[d];
        F[d] = *rho * (a[d] + z[d]);
    }
}

static void brinkman_moving_cell(const orc_problem *p, const double *f, double d, const double m[3], double *post,
                                 double dp[3]) {
    const lattice_t *L = lat(p->Q);
    double alpha, gamma, part[4], rho, j[3], den, sv[3], u[3], F[3], z[3];
    psm_resistance(d, p->psm_c_alpha, p->psm_drag_shape, p->psm_kappa, &alpha, &gamma, part);
    brinkman_moving_macro(L, f, alpha, gamma, p->accel, m, &rho, j, &den, sv, u, F, z);
    double om = p->kind == K_TRT ? orc_odd_rate(p->omega, p->magic) : p->omega;
    collide_pairs(L, f, rho, u, F, p->omega, om, post);
    for (int a = 0; a < 3; a++) dp[a] = rho * z[a];
}

/* cell_update_vjp, Brinkman branch (psm.rs:301-330, 478-492) without LES and omega. */
static void brinkman_moving_vjp(const orc_problem *p, const double *f, double d, const double m[3], const double *g,
                                const double dp_bar[3], double *f_bar, double *d_bar, double m_bar[3]) {
    const lattice_t *L = lat(p->Q);
    int Q = p->Q;
    double alpha, gamma, part[4], rho, j[3], den, sv[3], u[3], F[3], z[3];
    psm_resistance(d, p->psm_c_alpha, p->psm_drag_shape, p->psm_kappa, &alpha, &gamma, part);
    brinkman_moving_macro(L, f, alpha, gamma, p->accel, m, &rho, j, &den, sv, u, F, z);
    double om = p->kind == K_TRT ? orc_odd_rate(p->omega, p->magic) : p->omega;
    double cb_rho, cb_u[3], cb_F[3], ob, omb;
    collide_pairs_vjp(L, f, rho, u, F, p->omega, om, g, f_bar, &cb_rho, cb_u, cb_F, &ob, &omb);
    double rho_bar = cb_rho, u_bar[3] = {cb_u[0], cb_u[1], cb_u[2]};
    double alpha_bar = 0.0, gamma_bar = 0.0;
    for (int a = 0; a < 3; a++) {
        m_bar[a] = 0.0;
        rho_bar += cb_F[a] * (p->accel[a] + z[a]) + dp_bar[a] * z[a];
        double zb = (cb_F[a] + dp_bar[a]) * rho;
        gamma_bar += zb * m[a];
        m_bar[a] += zb * gamma;
        alpha_bar -= zb * u[a];
        u_bar[a] -= zb * alpha;
    }
    double j_bar[3];
    for (int a = 0; a < 3; a++) {
        double s_bar = u_bar[a] / den;
        alpha_bar -= u_bar[a] * sv[a] / (den * den) * 0.5;
        j_bar[a] = s_bar / rho;
        rho_bar -= s_bar * j[a] / (rho * rho);
        gamma_bar += s_bar * m[a] * 0.5;
        m_bar[a] += s_bar * gamma * 0.5;
    }
    *d_bar = alpha_bar * part[0] + gamma_bar * part[2];
    for (int i = 0; i < Q; i++) {
        double v = f_bar[i] + rho_bar;
        for (int a = 0; a < 3; a++)
            if (L->cf[i][a] != 0.0) v += j_bar[a] * L->cf[i][a];
        f_bar[i] = v;
    }
}
double orc_psm_saturate(double d, double kappa) { return psm_saturate(d, kappa); }

/* One coupled cell (d > 0): post-collision populations and the momentum exchange dp. */
static void psm_cell(const orc_problem *p, const double *f, double d, const double m[3], double *post, double dp[3]) {
    if (p->psm_law == 2) {
        brinkman_moving_cell(p, f, d, m, post, dp);
        return;
    }
    const lattice_t *L = lat(p->Q);
    int Q = p->Q;
    double rho, j[3], den, sv[3], u[3], F[3];
    brinkman_macro(L, f, 0.0, p->accel, &rho, j, &den, sv, u, F); /* u = j/rho + a/2, F = rho a */
    double fluid[MAXQ], eq[MAXQ];
    cell_collide_rate(p, f, 0.0, -1.0, fluid);
    double t = relax_tau(p) - 0.5;
    double b_over_d, beta_d, beta_t;
    psm_beta(d, t, p->psm_kappa, &b_over_d, &beta_d, &beta_t);
    double bb = b_over_d * d;
    equilibrium(L, rho, u, eq);
    double m2 = m[0] * m[0] + m[1] * m[1] + m[2] * m[2];
    dp[0] = dp[1] = dp[2] = 0.0;
    for (int i = 0; i < Q; i++) {
        int o = L->opp[i];
        double cm = dot_c(L, i, m);
        double solid_eq =
            rho * L->w[i] * (bb + b_over_d * cm * 3.0 + b_over_d * cm * cm * 4.5 / d - b_over_d * m2 * 1.5 / d);
        double omega_s = p->psm_law == 1 ? solid_eq - bb * eq[i] : bb * (f[o] - eq[o] - f[i]) + solid_eq;
        for (int a = 0; a < 3; a++)
            if (L->cf[i][a] != 0.0) dp[a] += omega_s * L->cf[i][a];
        post[i] = p->psm_law == 1 ? fluid[i] + omega_s : f[i] + (1.0 - bb) * (fluid[i] - f[i]) + omega_s;
    }
}

/* Cotangents (f_bar, d_bar, m_bar) of a coupled cell given the cotangent g of post and
 * dp_bar of dp.  The fluid collision's transpose is the Brinkman one at alpha = 0 (the
 * same u and force); the equilibrium and solid-equilibrium terms are chained through
 * rho and u = j/rho + a/2 separately (the reference sums both chains first). */
static void psm_cell_vjp(const orc_problem *p, const double *f, double d, const double m[3], const double *g,
                         const double dp_bar[3], double *f_bar, double *d_bar_out, double m_bar[3]) {
    if (p->psm_law == 2) {
        brinkman_moving_vjp(p, f, d, m, g, dp_bar, f_bar, d_bar_out, m_bar);
        return;
    }
    const lattice_t *L = lat(p->Q);
    int Q = p->Q;
    double rho, j[3], den, sv[3], u[3], F[3];
    brinkman_macro(L, f, 0.0, p->accel, &rho, j, &den, sv, u, F);
    double fluid[MAXQ], eq[MAXQ], eq_bar[MAXQ], fluid_bar[MAXQ];
    cell_collide_rate(p, f, 0.0, -1.0, fluid);
    double t = relax_tau(p) - 0.5;
    double b_over_d, beta_d, beta_t;
    psm_beta(d, t, p->psm_kappa, &b_over_d, &beta_d, &beta_t);
    double bb = b_over_d * d;
    equilibrium(L, rho, u, eq);
    double m2 = m[0] * m[0] + m[1] * m[1] + m[2] * m[2];
    double b_bar = 0.0, beta_bar = 0.0, wsum = 0.0, rho_bar = 0.0, u_bar[3] = {0.0, 0.0, 0.0}, d_bar = 0.0;
    for (int i = 0; i < Q; i++) {
        f_bar[i] = 0.0;
        eq_bar[i] = 0.0;
    }
    m_bar[0] = m_bar[1] = m_bar[2] = 0.0;
    const int sup = p->psm_law == 1;
    for (int i = 0; i < Q; i++) {
        int o = L->opp[i];
        double h = g[i];
        for (int a = 0; a < 3; a++)
            if (L->cf[i][a] != 0.0) h += dp_bar[a] * L->cf[i][a];
        if (sup) {
            fluid_bar[i] = g[i];
            b_bar += h * (rho * L->w[i] - eq[i]);
            eq_bar[i] -= h * bb;
        } else {
            fluid_bar[i] = g[i] * (1.0 - bb);
            f_bar[i] += g[i] * bb;
            b_bar += g[i] * (f[i] - fluid[i]);
            b_bar += h * (f[o] - eq[o] - f[i] + rho * L->w[i]);
            f_bar[o] += h * bb;
            f_bar[i] -= h * bb;
            eq_bar[o] -= h * bb;
        }
        double cm = dot_c(L, i, m);
        double w = h * L->w[i];
        rho_bar += w * (bb + b_over_d * cm * 3.0 + b_over_d * cm * cm * 4.5 / d - b_over_d * m2 * 1.5 / d);
        double wr = w * rho;
        beta_bar += wr * (cm * 3.0 + cm * cm * 4.5 / d - m2 * 1.5 / d);
        double cm_bar = wr * b_over_d * (cm * 9.0 / d + 3.0);
        for (int a = 0; a < 3; a++)
            if (L->cf[i][a] != 0.0) m_bar[a] += cm_bar * L->cf[i][a];
        wsum += wr;
        d_bar += wr * b_over_d * (m2 * 1.5 - cm * cm * 4.5) / (d * d);
    }
    for (int a = 0; a < 3; a++) m_bar[a] -= wsum * b_over_d * m[a] * 3.0 / d;
    equilibrium_vjp(L, rho, u, eq_bar, &rho_bar, u_bar);
    beta_bar += b_bar * d;
    d_bar += b_bar * b_over_d + beta_bar * beta_d;
    double fb[MAXQ], ab;
    cell_vjp_rate(p, f, 0.0, -1.0, fluid_bar, fb, &ab, NULL);
    double j_bar[3];
    for (int a = 0; a < 3; a++) {
        j_bar[a] = u_bar[a] / rho;
        rho_bar -= u_bar[a] * j[a] / (rho * rho);
    }
    for (int i = 0; i < Q; i++) {
        double v = f_bar[i] + fb[i] + rho_bar;
        for (int a = 0; a < 3; a++)
            if (L->cf[i][a] != 0.0) v += j_bar[a] * L->cf[i][a];
        f_bar[i] = v;
    }
    *d_bar_out = d_bar;
}

void orc_psm_cell(const orc_problem *p, const double *f, double d, const double *m, double *post, double *dp) {
    psm_cell(p, f, d, m, post, dp);
}

void orc_psm_cell_vjp(const orc_problem *p, const double *f, double d, const double *m, const double *g,
                      const double *dp_bar, double *f_bar, double *d_bar, double *m_bar) {
    psm_cell_vjp(p, f, d, m, g, dp_bar, f_bar, d_bar, m_bar);
}

/* The occupancy and solid momentum a step collides with (Field::solid_at, field.rs:1266-1275). */
static inline double psm_d(const orc_problem *p, const double *alpha, long x) {
    return p->occ_d1 ? alpha[x] * (1.0 - p->theta) + p->occ_d1[x] * p->theta : alpha[x];
}
static inline void psm_mv(const orc_problem *p, long x, double m[3]) {
    for (int a = 0; a < 3; a++) {
        double m0 = p->occ_m ? p->occ_m[3 * x + a] : 0.0;
        m[a] = p->occ_d1 ? m0 * (1.0 - p->theta) + (p->occ_m1 ? p->occ_m1[3 * x + a] : 0.0) * p->theta : m0;
    }
}

/* Stage cost on the momentum the coupled cells exchange with the solid in the collision
 * of `state` (CellOutput::dp, psm.rs:240-262): phi = sum over coupled cells of w . dp. */
double orc_obj_solid_force(const orc_problem *p, const double *alpha, const double *state, const double w[3]) {
    long n = (long)p->nx * p->ny * p->nz;
    int Q = p->Q;
    double acc = 0.0, post[MAXQ], dp[3];
    for (long x = 0; x < n; x++) {
        double d = psm_d(p, alpha, x), m[3];
        if (is_solid(p, x) || p->psm_law < 0 || !(d > 0.0)) continue;
        psm_mv(p, x, m);
        psm_cell(p, state + x * Q, d, m, post, dp);
        acc += w[0] * dp[0] + w[1] * dp[1] + w[2] * dp[2];
    }
    return acc;
}

/* weight * d phi / d (state, d, m): f_bar (AoS), d_bar and m_bar (3 per cell), overwritten. */
void orc_obj_solid_force_grad(const orc_problem *p, const double *alpha, const double *state, const double w[3],
                              double weight, double *f_bar, double *d_bar, double *m_bar) {
    long n = (long)p->nx * p->ny * p->nz;
    int Q = p->Q;
    double g[MAXQ], fb[MAXQ], db, mb[3];
    const double dpb[3] = {weight * w[0], weight * w[1], weight * w[2]};
    for (int q = 0; q < Q; q++) g[q] = 0.0;
    for (long x = 0; x < n; x++) {
        for (int q = 0; q < Q; q++) f_bar[x * Q + q] = 0.0;
        d_bar[x] = 0.0;
        for (int a = 0; a < 3; a++) m_bar[3 * x + a] = 0.0;
        double d = psm_d(p, alpha, x), m[3];
        if (is_solid(p, x) || p->psm_law < 0 || !(d > 0.0)) continue;
        psm_mv(p, x, m);
        psm_cell_vjp(p, state + x * Q, d, m, g, dpb, fb, &db, mb);
        for (int q = 0; q < Q; q++) f_bar[x * Q + q] = fb[q];
        d_bar[x] = db;
        for (int a = 0; a < 3; a++) m_bar[3 * x + a] = mb[a];
    }
}

void orc_port_populations(int Q, const double *nb, int kind, const double value[3], double *out) {
    const lattice_t *L = lat(Q);
    double rho_n, j[3], u_n[3], eb[MAXQ], en[MAXQ];
    moments(L, nb, &rho_n, j);
    for (int d = 0; d < 3; d++) u_n[d] = j[d] / rho_n;
    if (kind == 0) {
        equilibrium(L, rho_n, value, eb);
    } else {
        equilibrium(L, value[0], u_n, eb);
    }
    equilibrium(L, rho_n, u_n, en);
    for (int i = 0; i < Q; i++) out[i] = eb[i] + nb[i] - en[i];
}

void orc_port_populations_vjp(int Q, const double *nb, int kind, const double value[3], const double *g,
                              double *nb_bar) {
    const lattice_t *L = lat(Q);
    double rho_n, j[3], u_n[3];
    moments(L, nb, &rho_n, j);
    for (int d = 0; d < 3; d++) u_n[d] = j[d] / rho_n;
    double rho_b = kind == 0 ? rho_n : value[0];
    const double *u_b = kind == 0 ? value : u_n;
    double rb_bar = 0.0, ub_bar[3] = {0, 0, 0};
    equilibrium_vjp(L, rho_b, u_b, g, &rb_bar, ub_bar);
    double neg[MAXQ];
    for (int i = 0; i < Q; i++) neg[i] = -g[i];
    double rn_bar = 0.0, un_bar[3] = {0, 0, 0};
    equilibrium_vjp(L, rho_n, u_n, neg, &rn_bar, un_bar);
    if (kind == 0) {
        rn_bar += rb_bar;
    } else {
        for (int d = 0; d < 3; d++) un_bar[d] += ub_bar[d];
    }
    double j_bar[3];
    for (int d = 0; d < 3; d++) {
        j_bar[d] = un_bar[d] / rho_n;
        rn_bar -= un_bar[d] * j[d] / (rho_n * rho_n);
    }
    for (int i = 0; i < Q; i++) {
        double v = g[i] + rn_bar;
        for (int d = 0; d < 3; d++)
            if (L->cf[i][d] != 0.0) v += j_bar[d] * L->cf[i][d];
        nb_bar[i] = v;
    }
}

/* ------------------------------------------------------------------------- */
/* Stationary D3Q19 ports (ports.rs): the neighbour non-equilibrium pressure     */
/* rule reconstruct_neighbor (:484-497) and the moment reconstruction           */
/* `reconstruct` of "legacy_moment" ports (:317-406).  Each rule is written once */
/* on dual numbers: the value part is the FP64 rule in the reference's operation */
/* order, and the tangent of input direction k is column k of the cell Jacobian, */
/* the way CompiledPort::cell_jacobian (:532-557) forms it with Dual<19>.        */
/*                                                                               */
/* A port row's kind is ST_PORT_BASE + 6 * rule + face (face 0..5 = xmin, xmax,  */
/* ymin, ymax, zmin, zmax), rule 0 = neighbour pressure, 1 = moment velocity,    */
/* 2 = moment pressure, 3 and 4 = the regularized rules of velocity and pressure */
/* ports (reconstruct_regularized, :455-482, "regularized_known_stress"); its    */
/* three values are                                                              */
/*   rule 0: rho_b and the two tangential velocities (ascending axes),          */
/*   rules 1, 3: v0 = u - a/2 (ports.rs:352), three components,                   */
/*   rules 2, 4: rho_b and the two tangential components of v0.                  */
/* (The regularized rules are refused with an acceleration, ports.rs:154-158, so */
/* their v0 is u.)                                                               */
/* ------------------------------------------------------------------------- */
#define ST_PORT_BASE 8

static dual st_dot_c_d(int q, const dual u[3]) {
    /* ports.rs:285-297: the first nonzero term starts the sum */
    dual acc = dv(0.0);
    int first = 1;
    for (int d = 0; d < 3; d++) {
        if (C19[q][d] == 0) continue;
        dual term = C19[q][d] == 1 ? u[d] : dmuls(u[d], -1.0);
        acc = first ? term : dadd(acc, term);
        first = 0;
    }
    return acc;
}

static void st_moments_d(const dual *f, dual *rho, dual m[3]) {
    /* ports.rs:299-316 */
    dual r = f[0];
    for (int q = 1; q < 19; q++) r = dadd(r, f[q]);
    for (int d = 0; d < 3; d++) {
        dual acc = dv(0.0);
        for (int q = 0; q < 19; q++) {
            if (C19[q][d] == 1) acc = dadd(acc, f[q]);
            else if (C19[q][d] == -1) acc = dsub(acc, f[q]);
        }
        m[d] = acc;
    }
    *rho = r;
}

static void st_equilibrium_d(dual rho, const dual u[3], dual *eq) {
    /* ports.rs:276-282: rho W (1 + 3 cu + 4.5 cu^2 - 1.5 u^2) */
    const dual usq = dadd(dadd(dmul(u[0], u[0]), dmul(u[1], u[1])), dmul(u[2], u[2]));
    for (int q = 0; q < 19; q++) {
        const dual cu = st_dot_c_d(q, u);
        dual b = dadd(dadd(dv(1.0), dmuls(cu, 3.0)), dmuls(dmul(cu, cu), 4.5));
        b = dsub(b, dmuls(usq, 1.5));
        eq[q] = dmul(dmuls(rho, L19.w[q]), b);
    }
}

/* right_inverse (ports.rs:325-341) of the unknown set: rows = (1, c_t0, c_t1) over
 * the unknowns, m = rows rows^T, ri[k][j] = sum_i rows[i][k] inv(m)[i][j].  For D3Q19
 * m is diagonal (the unknowns of a face are symmetric in each tangential axis), so the
 * LU solve of linalg3::inv returns inv = diag(1 / m_ii), which is what is formed here;
 * a non-diagonal m is refused. */
static int st_right_inverse(const int *unknown, int nu, const int tangent[2], double ri[][3]) {
    double rows[3][9], m[3][3], inv[3][3];
    for (int k = 0; k < nu; k++) {
        rows[0][k] = 1.0;
        rows[1][k] = (double)C19[unknown[k]][tangent[0]];
        rows[2][k] = (double)C19[unknown[k]][tangent[1]];
    }
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) {
            double acc = 0.0;
            for (int k = 0; k < nu; k++) acc += rows[i][k] * rows[j][k];
            m[i][j] = acc;
        }
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) {
            if (i != j && m[i][j] != 0.0) return -1;
            inv[i][j] = i == j ? 1.0 / m[i][i] : 0.0;
        }
    for (int k = 0; k < nu; k++)
        for (int j = 0; j < 3; j++) {
            double acc = 0.0;
            for (int i = 0; i < 3; i++) acc += rows[i][k] * inv[i][j];
            ri[k][j] = acc;
        }
    return 0;
}

/* reconstruct (ports.rs:343-406) of the cell's own populations f on face `face`;
 * density_lattice given (pressure, rule 2) or not (velocity, rule 1). */
static void st_reconstruct_d(const dual *f, int face, const double v0in[3], int pressure, double density,
                             dual *out) {
    const int axis = face / 2, sign = face % 2 == 0 ? 1 : -1;
    int unknown[9], outgoing[9], zero[9], nu = 0, no = 0, nz = 0;
    for (int q = 0; q < 19; q++) {
        if (C19[q][axis] == sign) unknown[nu++] = q;
        else if (C19[q][axis] == -sign) outgoing[no++] = q;
        else zero[nz++] = q;
    }
    int tangent[2], nt = 0;
    for (int a = 0; a < 3; a++)
        if (a != axis) tangent[nt++] = a;
    dual compat = f[zero[0]];
    for (int k = 1; k < nz; k++) compat = dadd(compat, f[zero[k]]);
    dual out_sum = f[outgoing[0]];
    for (int k = 1; k < no; k++) out_sum = dadd(out_sum, f[outgoing[k]]);
    compat = dadd(compat, dmuls(out_sum, 2.0));
    const double sgn = (double)sign;
    dual rho, v[3];
    for (int d = 0; d < 3; d++) v[d] = dv(v0in[d]);
    if (!pressure) {
        rho = ddivs(compat, 1.0 - sgn * v0in[axis]);
    } else {
        rho = dv(density);
        v[axis] = dmuls(dsub(dv(1.0), ddiv(compat, rho)), sgn);
    }
    dual proposal[9];
    for (int k = 0; k < nu; k++) {
        const int q = unknown[k];
        proposal[k] = dadd(f[L19.opp[q]], dmul(dmuls(rho, 6.0 * L19.w[q]), st_dot_c_d(q, v)));
    }
    dual known[19], km[3], kmass;
    for (int q = 0; q < 19; q++) known[q] = f[q];
    for (int k = 0; k < nu; k++) known[unknown[k]] = dv(0.0);
    st_moments_d(known, &kmass, km);
    const dual target[3] = {dsub(rho, kmass), dsub(dmul(rho, v[tangent[0]]), km[tangent[0]]),
                            dsub(dmul(rho, v[tangent[1]]), km[tangent[1]])};
    dual pa[3] = {dv(0.0), dv(0.0), dv(0.0)};
    for (int k = 0; k < nu; k++) {
        const int q = unknown[k];
        pa[0] = dadd(pa[0], proposal[k]);
        pa[1] = dadd(pa[1], dmuls(proposal[k], (double)C19[q][tangent[0]]));
        pa[2] = dadd(pa[2], dmuls(proposal[k], (double)C19[q][tangent[1]]));
    }
    const dual residual[3] = {dsub(target[0], pa[0]), dsub(target[1], pa[1]), dsub(target[2], pa[2])};
    double ri[9][3];
    st_right_inverse(unknown, nu, tangent, ri);
    for (int q = 0; q < 19; q++) out[q] = f[q];
    for (int k = 0; k < nu; k++) {
        dual corr = dadd(dadd(dmuls(residual[0], ri[k][0]), dmuls(residual[1], ri[k][1])),
                         dmuls(residual[2], ri[k][2]));
        out[unknown[k]] = dadd(proposal[k], corr);
    }
}

/* reconstruct_neighbor (ports.rs:484-497) from the inward neighbour's populations. */
static void st_neighbor_d(const dual *nb, int face, double density, const double tang[2], dual *out) {
    const int axis = face / 2;
    dual rho, m[3], vel[3], target[3], eb[19], el[19];
    st_moments_d(nb, &rho, m);
    for (int d = 0; d < 3; d++) vel[d] = ddiv(m[d], rho);
    for (int d = 0, t = 0; d < 3; d++) target[d] = d == axis ? vel[d] : dv(tang[t++]);
    st_equilibrium_d(dv(density), target, eb);
    st_equilibrium_d(rho, vel, el);
    for (int q = 0; q < 19; q++) out[q] = dsub(dadd(eb[q], nb[q]), el[q]);
}

/* ------------------------------------------------------------------------- */
/* The known-stress fit of the regularized rules (known_stress_operator,          */
/* ports.rs:409-440).  basis[q] = 4.5 w_q (Q00, Q11, Q22, 2 Q01, 2 Q02, 2 Q12),   */
/* Q_ab = c_a c_b - delta_ab / 3, in the reference's floating-point order; the    */
/* known set is every q with c_axis != sign (14 directions, ascending); the       */
/* operator is the pseudo-inverse of the known rows of the basis weighted by       */
/* 1 / sqrt(w), times 1 / sqrt(w) on its columns.  With B the 14 x 6 known rows    */
/* (rank 6) that is the weighted least-squares left inverse                        */
/*     op = (B^T W^-1 B)^-1 B^T W^-1,                                             */
/* a matrix of rationals (the square roots cancel).  It is formed here in exact    */
/* rational arithmetic and each entry rounded once to the nearest double; the      */
/* reference's SVD route (faer) lands within a few units in the last place of it.  */
/* ------------------------------------------------------------------------- */
typedef struct { long long n, d; } st_rat;  /* d > 0, lowest terms */

static int st_rat_norm(__int128 n, __int128 d, st_rat *out) {
    if (d == 0) return -1;
    if (d < 0) { n = -n; d = -d; }
    __int128 a = n < 0 ? -n : n, b = d;
    while (b) { __int128 t = a % b; a = b; b = t; }
    if (a > 1) { n /= a; d /= a; }
    const __int128 lim = (__int128)1 << 62;
    if (n >= lim || n <= -lim || d >= lim) return -1;
    out->n = (long long)n;
    out->d = (long long)d;
    return 0;
}

static int st_rat_add(st_rat a, st_rat b, st_rat *out) {
    return st_rat_norm((__int128)a.n * b.d + (__int128)b.n * a.d, (__int128)a.d * b.d, out);
}

static int st_rat_mul(st_rat a, st_rat b, st_rat *out) {
    return st_rat_norm((__int128)a.n * b.n, (__int128)a.d * b.d, out);
}

/* 4.5 w_q e as an exact rational: 4.5 w_q = 3/2, 1/4, 1/8 for |c|^2 = 0, 1, 2, and
 * e = c_a^2 - 1/3 on the diagonal entries, 2 c_a c_b off it. */
static st_rat st_basis_rat(int q, int r) {
    static const int A[6] = {0, 1, 2, 0, 0, 1}, B[6] = {0, 1, 2, 1, 2, 2};
    const int a = A[r], b = B[r];
    int n2 = 0;
    for (int d = 0; d < 3; d++) n2 += C19[q][d] * C19[q][d];
    const long long sd = n2 == 0 ? 2 : (n2 == 1 ? 4 : 8), sn = n2 == 0 ? 3 : 1;
    st_rat e, out = {0, 0};
    if (a == b) {
        e.n = 3LL * C19[q][a] * C19[q][a] - 1;
        e.d = 3;
    } else {
        e.n = 2LL * C19[q][a] * C19[q][b];
        e.d = 1;
    }
    st_rat scale = {sn, sd};
    if (st_rat_mul(scale, e, &out) != 0) out.d = 0;
    return out;
}

/* basis[q][r] in the reference's floating-point order (ports.rs:413-417). */
static double st_basis(int q, int r) {
    static const int A[6] = {0, 1, 2, 0, 0, 1}, B[6] = {0, 1, 2, 1, 2, 2};
    const int a = A[r], b = B[r];
    const double qq = (double)C19[q][a] * (double)C19[q][b] - (a == b ? 1.0 / 3.0 : 0.0);
    const double e = r < 3 ? qq : 2.0 * qq;
    return 4.5 * L19.w[q] * e;
}

static int st_stress_known(int face, int known[19]) {
    const int axis = face / 2, sign = face % 2 == 0 ? 1 : -1;
    int nk = 0;
    for (int q = 0; q < 19; q++)
        if (C19[q][axis] != sign) known[nk++] = q;
    return nk;
}

/* op[r * 14 + k], r over the six stress components, k over the known set. */
static int st_stress_operator_exact(int face, double *op) {
    int known[19];
    const int nk = st_stress_known(face, known);
    if (nk != 14) return -1;
    st_rat bk[14][6], inv_w[14], n[6][12];
    for (int k = 0; k < nk; k++) {
        const int q = known[k];
        int n2 = 0;
        for (int d = 0; d < 3; d++) n2 += C19[q][d] * C19[q][d];
        inv_w[k].n = n2 == 0 ? 3 : (n2 == 1 ? 18 : 36);
        inv_w[k].d = 1;
        for (int r = 0; r < 6; r++) {
            bk[k][r] = st_basis_rat(q, r);
            if (bk[k][r].d == 0) return -1;
        }
    }
    /* [N | I], N = B^T W^-1 B, reduced to [I | N^-1] by Gauss-Jordan with the first
     * nonzero pivot of each column (exact arithmetic needs no pivot search). */
    for (int r = 0; r < 6; r++)
        for (int s = 0; s < 12; s++) {
            st_rat acc = {s == r + 6 ? 1 : 0, 1};
            if (s < 6) {
                acc.n = 0;
                for (int k = 0; k < nk; k++) {
                    st_rat t;
                    if (st_rat_mul(bk[k][r], bk[k][s], &t) || st_rat_mul(t, inv_w[k], &t) ||
                        st_rat_add(acc, t, &acc))
                        return -1;
                }
            }
            n[r][s] = acc;
        }
    for (int c = 0; c < 6; c++) {
        int piv = -1;
        for (int r = c; r < 6 && piv < 0; r++)
            if (n[r][c].n != 0) piv = r;
        if (piv < 0) return -1; /* rank below 6 */
        for (int s = 0; s < 12; s++) {
            st_rat t = n[c][s];
            n[c][s] = n[piv][s];
            n[piv][s] = t;
        }
        const st_rat p = n[c][c], pinv = {p.n < 0 ? -p.d : p.d, p.n < 0 ? -p.n : p.n};
        for (int s = 0; s < 12; s++)
            if (st_rat_mul(n[c][s], pinv, &n[c][s])) return -1;
        for (int r = 0; r < 6; r++) {
            if (r == c || n[r][c].n == 0) continue;
            const st_rat f = {-n[r][c].n, n[r][c].d};
            for (int s = 0; s < 12; s++) {
                st_rat t;
                if (st_rat_mul(f, n[c][s], &t) || st_rat_add(n[r][s], t, &n[r][s])) return -1;
            }
        }
    }
    for (int r = 0; r < 6; r++)
        for (int k = 0; k < nk; k++) {
            st_rat acc = {0, 1};
            for (int s = 0; s < 6; s++) {
                st_rat t;
                if (st_rat_mul(n[r][6 + s], bk[k][s], &t) || st_rat_add(acc, t, &acc)) return -1;
            }
            if (st_rat_mul(acc, inv_w[k], &acc)) return -1;
            op[r * 14 + k] = (double)acc.n / (double)acc.d; /* both exact: |n|, d < 2^53 */
            if (acc.n >= (1LL << 53) || acc.n <= -(1LL << 53) || acc.d >= (1LL << 53)) return -1;
        }
    return 0;
}

static double g_stress_op[6][84];
static int g_stress_ready[6];

static const double *st_stress_operator(int face) {
    if (!g_stress_ready[face]) {
        if (st_stress_operator_exact(face, g_stress_op[face]) != 0) return NULL;
        g_stress_ready[face] = 1;
    }
    return g_stress_op[face];
}

int orc_stationary_stress_operator(int face, double *op, int *known) {
    init_lattices();
    if (face < 0 || face > 5) return -1;
    const double *src = st_stress_operator(face);
    if (!src) return -1;
    for (int i = 0; i < 84; i++) op[i] = src[i];
    return st_stress_known(face, known);
}

double orc_stationary_stress_basis(int q, int r) {
    init_lattices();
    return st_basis(q, r);
}

/* reconstruct_regularized (ports.rs:455-482): the moment reconstruction `target` of
 * the cell's own populations f, its equilibrium, the six stresses fitted to the known
 * populations' non-equilibrium parts, and eq + basis . stress. */
static void st_regularized_d(const dual *f, int face, const dual *target, dual *out) {
    int known[19];
    const int nk = st_stress_known(face, known);
    const double *op = st_stress_operator(face);
    dual rho, m[3], vel[3], eq[19], stress[6];
    if (!op) { /* not reached for D3Q19: every face's fit has rank 6 */
        for (int q = 0; q < 19; q++) out[q] = dv(NAN);
        return;
    }
    st_moments_d(target, &rho, m);
    for (int d = 0; d < 3; d++) vel[d] = ddiv(m[d], rho);
    st_equilibrium_d(rho, vel, eq);
    for (int r = 0; r < 6; r++) {
        dual acc = dv(0.0);
        for (int k = 0; k < nk; k++)
            acc = dadd(acc, dmuls(dsub(f[known[k]], eq[known[k]]), op[r * 14 + k]));
        stress[r] = acc;
    }
    for (int q = 0; q < 19; q++) {
        dual acc = dv(0.0);
        for (int r = 0; r < 6; r++) acc = dadd(acc, dmuls(stress[r], st_basis(q, r)));
        out[q] = dadd(eq[q], acc);
    }
}

static void st_port_d(const dual *src, int kind, const double value[3], dual *out) {
    const int rule = (kind - ST_PORT_BASE) / 6, face = (kind - ST_PORT_BASE) % 6;
    const int axis = face / 2;
    if (rule == 0) {
        st_neighbor_d(src, face, value[0], value + 1, out);
        return;
    }
    if (rule >= 3) {
        dual target[19];
        st_port_d(src, kind - 12, value, target);
        st_regularized_d(src, face, target, out);
        return;
    }
    double v0[3];
    if (rule == 1) {
        for (int d = 0; d < 3; d++) v0[d] = value[d];
    } else {
        for (int d = 0, t = 1; d < 3; d++) v0[d] = d == axis ? 0.0 : value[t++];
    }
    st_reconstruct_d(src, face, v0, rule == 2, value[0], out);
}

/* The stationary rule of row kind `kind` on its source cell (the neighbour for rule 0,
 * the port cell itself for the moment rules). */
void orc_stationary_port(const double *src, int kind, const double value[3], double *out) {
    init_lattices();
    dual x[19], y[19];
    for (int q = 0; q < 19; q++) x[q] = dv(src[q]);
    st_port_d(x, kind, value, y);
    for (int q = 0; q < 19; q++) out[q] = y[q].v;
}

/* Cell Jacobian jac[r * 19 + k] = d out_r / d src_k (cell_jacobian, ports.rs:532-557).
 * The "legacy_moment" rules (1, 2) are affine, so their Jacobian is taken once at the
 * lattice weights, as CompiledPort::new does (:514-520); the neighbour and regularized
 * rules at the source itself. */
void orc_stationary_port_jacobian(const double *src, int kind, const double value[3], double *jac) {
    init_lattices();
    const int rule = (kind - ST_PORT_BASE) / 6;
    const int at_src = rule == 0 || rule >= 3;
    dual x[19], y[19];
    for (int k = 0; k < 19; k++) {
        for (int q = 0; q < 19; q++) {
            x[q].v = at_src ? src[q] : L19.w[q];
            x[q].d = q == k ? 1.0 : 0.0;
        }
        st_port_d(x, kind, value, y);
        for (int r = 0; r < 19; r++) jac[r * 19 + k] = y[r].d;
    }
}

int orc_stationary_right_inverse(int face, double *ri_out) {
    const int axis = face / 2, sign = face % 2 == 0 ? 1 : -1;
    int unknown[9], nu = 0, tangent[2], nt = 0;
    for (int q = 0; q < 19; q++)
        if (C19[q][axis] == sign) unknown[nu++] = q;
    for (int a = 0; a < 3; a++)
        if (a != axis) tangent[nt++] = a;
    double ri[9][3];
    if (st_right_inverse(unknown, nu, tangent, ri) != 0) return -1;
    for (int k = 0; k < nu; k++)
        for (int j = 0; j < 3; j++) ri_out[k * 3 + j] = ri[k][j];
    return nu;
}

/* ------------------------------------------------------------------------- */
/* Problem level: initial_populations (solver.rs:593-601), interval             */
/* (solver.rs:604-614: collide -> stream -> ports), interval_vjp (:642-674).    */
/* ------------------------------------------------------------------------- */

/* ------------------------------------------------------------------------- */
/* WALE closure (moving/wale.rs) on BGK/TRT, restated for the stationary step:  */
/* the rate of a fluid cell is 1 / (tau0 + 3 nu_t) with nu_t from the velocity  */
/* gradient of u = j / rho of the state the step collides (field.rs:1294-1320), */
/* by central differences, or one-sided ones next to walls and closed faces     */
/* (Stencil::new / gradient).                                                   */
/* ------------------------------------------------------------------------- */
static int wale_on(const orc_problem *p) {
    return p->wale_constant > 0.0 && (p->kind == K_BGK || p->kind == K_TRT);
}

/* u = j / rho per fluid cell (0 at solid cells) and rho. */
static void wale_velocity(const orc_problem *p, const double *state, double *u, double *rho_out) {
    const lattice_t *L = lat(p->Q);
    long n = (long)p->nx * p->ny * p->nz;
#pragma omp parallel for schedule(static)
    for (long x = 0; x < n; x++) {
        if (is_solid(p, x)) {
            u[3 * x] = u[3 * x + 1] = u[3 * x + 2] = 0.0;
            rho_out[x] = 1.0;
            continue;
        }
        double rho, j[3];
        moments(L, state + x * p->Q, &rho, j);
        for (int d = 0; d < 3; d++) u[3 * x + d] = j[d] / rho;
        rho_out[x] = rho;
    }
}

/* Neighbour of x one cell along +-axis that the stencil may use (inside or across a
 * periodic axis, not solid); -1 if none. */
static long wale_nb(const orc_problem *p, long x, int axis, int dir) {
    long off[3] = {0, 0, 0};
    off[axis] = dir;
    if (!g_inside(p, x, off)) return -1;
    long y = g_wrap(p, x, off);
    return is_solid(p, y) ? -1 : y;
}

/* Difference kind along axis b at x: 0 zero, 1 central (pp, mm), 2 forward (pp), 3 backward (mm). */
static int wale_diff(const orc_problem *p, long x, int b, long *pp, long *mm) {
    const int n[3] = {p->nx, p->ny, p->nz};
    *pp = *mm = -1;
    if (is_solid(p, x) || n[b] == 1) return 0;
    *pp = wale_nb(p, x, b, 1);
    *mm = wale_nb(p, x, b, -1);
    if (*pp >= 0 && *mm >= 0) return 1;
    if (*pp >= 0) return 2;
    if (*mm >= 0) return 3;
    return 0;
}

/* g[a][b] = d u_a / d x_b (Stencil::gradient). */
static void wale_gradient(const orc_problem *p, const double *u, long x, double g[3][3]) {
    for (int b = 0; b < 3; b++) {
        long pp, mm;
        int kind = wale_diff(p, x, b, &pp, &mm);
        for (int a = 0; a < 3; a++) {
            switch (kind) {
            case 1: g[a][b] = (u[3 * pp + a] - u[3 * mm + a]) * 0.5; break;
            case 2: g[a][b] = u[3 * pp + a] - u[3 * x + a]; break;
            case 3: g[a][b] = u[3 * x + a] - u[3 * mm + a]; break;
            default: g[a][b] = 0.0; break;
            }
        }
    }
}

static void wale_invariants(const double g[3][3], double sd[3][3], double s[3][3], double *a_inv, double *b_inv) {
    double g2[3][3];
    for (int a = 0; a < 3; a++)
        for (int b = 0; b < 3; b++) {
            g2[a][b] = 0.0;
            for (int c = 0; c < 3; c++) g2[a][b] += g[a][c] * g[c][b];
        }
    double tr = (g2[0][0] + g2[1][1] + g2[2][2]) / 3.0;
    *a_inv = 0.0;
    *b_inv = 0.0;
    for (int a = 0; a < 3; a++)
        for (int b = 0; b < 3; b++) {
            sd[a][b] = (g2[a][b] + g2[b][a]) * 0.5 - (a == b ? tr : 0.0);
            s[a][b] = (g[a][b] + g[b][a]) * 0.5;
            *a_inv += sd[a][b] * sd[a][b];
            *b_inv += s[a][b] * s[a][b];
        }
}

/* omega = 1 / (tau0 + 3 C^2 A^(3/2) / (B^(5/2) + A^(5/4) + floor)) (wale.rs omega). */
static double wale_omega(const orc_problem *p, const double g[3][3], double tau0) {
    double sd[3][3], s[3][3], a, b;
    wale_invariants(g, sd, s, &a, &b);
    double num = pow(a, 1.5);
    double den = pow(b, 2.5) + pow(a, 1.25) + p->wale_floor;
    double nu = num / den * (p->wale_constant * p->wale_constant);
    return 1.0 / (tau0 + nu * 3.0);
}

/* wale.rs omega_vjp: cotangent of the gradient given the cotangent w of omega. */
static void wale_omega_vjp(const orc_problem *p, const double g[3][3], double tau0, double w, double g_bar[3][3]) {
    double sd[3][3], s[3][3], a, b;
    wale_invariants(g, sd, s, &a, &b);
    double c2 = p->wale_constant * p->wale_constant;
    double num = pow(a, 1.5);
    double den = pow(b, 2.5) + pow(a, 1.25) + p->wale_floor;
    double nu = num / den * c2;
    double tau = tau0 + nu * 3.0;
    double tau_bar = -w / (tau * tau);
    double nu_bar = tau_bar * 3.0;
    double a_bar = nu_bar * c2 * (pow(a, 0.5) * 1.5 / den - num * pow(a, 0.25) * 1.25 / (den * den));
    double b_bar = -nu_bar * c2 * num * pow(b, 1.5) * 2.5 / (den * den);
    for (int x = 0; x < 3; x++)
        for (int y = 0; y < 3; y++) {
            double da = 0.0;
            for (int k = 0; k < 3; k++) da += sd[x][k] * g[y][k] + g[k][x] * sd[k][y];
            g_bar[x][y] = a_bar * da * 2.0 + b_bar * s[x][y] * 2.0;
        }
}

/* Rate of every cell of a state (p->omega at solid cells). */
static void wale_rates(const orc_problem *p, const double *state, double *rate, double *u, double *rho) {
    long n = (long)p->nx * p->ny * p->nz;
    wale_velocity(p, state, u, rho);
    const double tau0 = relax_tau(p);
#pragma omp parallel for schedule(static)
    for (long x = 0; x < n; x++) {
        if (is_solid(p, x)) {
            rate[x] = p->omega;
            continue;
        }
        double g[3][3];
        wale_gradient(p, u, x, g);
        rate[x] = wale_omega(p, g, tau0);
    }
}

void orc_wale_rates(const orc_problem *p, const double *state, double *rate) {
    long n = (long)p->nx * p->ny * p->nz;
    double *u = (double *)malloc(sizeof(double) * 3 * n), *rho = (double *)malloc(sizeof(double) * n);
    wale_rates(p, state, rate, u, rho);
    free(u);
    free(rho);
}

void orc_wale_omega(const orc_problem *p, const double *g9, double tau0, double *omega_out) {
    double g[3][3];
    for (int a = 0; a < 9; a++) g[a / 3][a % 3] = g9[a];
    *omega_out = wale_omega(p, g, tau0);
}

void orc_wale_omega_vjp(const orc_problem *p, const double *g9, double tau0, double w, double *g_bar9) {
    double g[3][3], gb[3][3];
    for (int a = 0; a < 9; a++) g[a / 3][a % 3] = g9[a];
    wale_omega_vjp(p, g, tau0, w, gb);
    for (int a = 0; a < 9; a++) g_bar9[a] = gb[a / 3][a % 3];
}

void orc_initial(const orc_problem *p, double *f0) {
    const lattice_t *L = lat(p->Q);
    long n = (long)p->nx * p->ny * p->nz;
    int Q = p->Q;
    double fluid[MAXQ];
    for (int q = 0; q < Q; q++) fluid[q] = L->w[q] - 1.5 * L->w[q] * dot_c(L, q, p->accel);
    for (long x = 0; x < n; x++)
        for (int q = 0; q < Q; q++) f0[x * Q + q] = is_solid(p, x) ? L->w[q] : fluid[q];
}

/* Stationary ports applied in place after streaming; the rule is the moving NEE
 * port_populations from the pulled neighbour (field.rs:1251-1263), i.e. from the
 * neighbour's streamed populations.  exchange[r] = sum_q (new - old) at port cell r
 * (ports.rs:560-580). */
static void apply_ports(const orc_problem *p, double *after, double *exchange) {
    int Q = p->Q;
    double vals[MAXQ];
    for (int r = 0; r < p->n_port_cells; r++) {
        long x = p->port_cell[r], nb = p->port_nb[r];
        if (p->port_kind[r] >= ST_PORT_BASE)
            orc_stationary_port(after + nb * Q, p->port_kind[r], p->port_value + 3 * r, vals);
        else
            orc_port_populations(Q, after + nb * Q, p->port_kind[r], p->port_value + 3 * r, vals);
        double delta = 0.0;
        for (int q = 0; q < Q; q++) {
            delta += vals[q] - after[x * Q + q];
            after[x * Q + q] = vals[q];
        }
        if (exchange) exchange[r] = delta;
    }
}

/* The sponge blend of field.rs:1461-1475 on the post-collision populations:
 * target = sum over parts (in sponge order) of f^eq(rho_l, velocity(phase)) * s_l
 * (Target::velocity, field.rs:373-376), then post = post * (1 - sigma) + target. */
static void apply_sponges(const orc_problem *p, double *post) {
    const lattice_t *L = lat(p->Q);
    int Q = p->Q;
    for (int r = 0; r < p->n_sponge_cells; r++) {
        long x = p->sponge_cell[r];
        double target[MAXQ], e[MAXQ];
        for (int q = 0; q < Q; q++) target[q] = 0.0;
        for (int k = p->sponge_off[r]; k < p->sponge_off[r + 1]; k++) {
            const double *t = p->sponge_target + 10 * p->sponge_part[k];
            const double *ph = p->sponge_phase + 2 * k;
            double u[3];
            for (int d = 0; d < 3; d++) u[d] = t[1 + d] + t[4 + d] * ph[0] - t[7 + d] * ph[1];
            equilibrium(L, t[0], u, e);
            for (int q = 0; q < Q; q++) target[q] += e[q] * p->sponge_weight[k];
        }
        double keep = 1.0 - p->sponge_sigma[r];
        for (int q = 0; q < Q; q++) post[x * Q + q] = post[x * Q + q] * keep + target[q];
    }
}

void orc_interval(const orc_problem *p, const uint32_t *source, const double *alpha, const double *before,
                  double *post, double *after, double *exchange) {
    long n = (long)p->nx * p->ny * p->nz;
    int Q = p->Q;
    double *rate = NULL, *wu = NULL, *wr = NULL;
    if (wale_on(p)) {
        rate = (double *)malloc(sizeof(double) * n);
        wu = (double *)malloc(sizeof(double) * 3 * n);
        wr = (double *)malloc(sizeof(double) * n);
        wale_rates(p, before, rate, wu, wr);
    }
#pragma omp parallel for schedule(static)
    for (long x = 0; x < n; x++) {
        if (is_solid(p, x)) {
            for (int q = 0; q < Q; q++) post[x * Q + q] = before[x * Q + q];
            continue;
        }
        if (p->psm_law >= 0) {
            double dp[3];
            double d = psm_d(p, alpha, x), m[3];
            psm_mv(p, x, m);
            if (d > 0.0) {
                psm_cell(p, before + x * Q, d, m, post + x * Q, dp);
                if (p->sub_g) { /* Field::accumulate, field.rs:1518-1526 */
                    for (int e = 0; e < 3; e++) {
                        double gv = -dp[e] / d;
                        p->sub_g[6 * x + e] += gv * (1.0 - p->theta);
                        p->sub_g[6 * x + 3 + e] += gv * p->theta;
                    }
                }
            } else {
                cell_collide_rate(p, before + x * Q, 0.0, -1.0, post + x * Q);
            }
            continue;
        }
        cell_collide_rate(p, before + x * Q, alpha[x], rate ? rate[x] : -1.0, post + x * Q);
    }
    free(rate);
    free(wu);
    free(wr);
    apply_sponges(p, post);
#pragma omp parallel for schedule(static)
    for (long i = 0; i < n * Q; i++) {
        uint32_t s = source[i];
        after[i] = s == NONE_U32 ? before[i] : post[s];
    }
    apply_ports(p, after, exchange);
}

/* interval_vjp (solver.rs:642-674): ports VJP, stream VJP (each fluid post value has
 * exactly one reader), collide VJP, plus previous_bar at solid cells.
 * `after_pre_port` holds stream(collide(before)) at the port neighbours (the
 * reference's reflected_cells, solver.rs:616-639). */
void orc_interval_vjp(const orc_problem *p, const uint32_t *source, const double *alpha, const double *before,
                      const double *after_pre_port, const double *after_bar, double *work, double *f_bar,
                      double *alpha_bar) {
    long n = (long)p->nx * p->ny * p->nz;
    int Q = p->Q;
    /* work := reflected_bar */
    memcpy(work, after_bar, sizeof(double) * n * Q);
    for (int r = 0; r < p->n_port_cells; r++) {
        long x = p->port_cell[r];
        for (int q = 0; q < Q; q++) work[x * Q + q] = 0.0;
    }
    double nb_bar[MAXQ];
    for (int r = 0; r < p->n_port_cells; r++) {
        long x = p->port_cell[r], nb = p->port_nb[r];
        const double *g = after_bar + x * Q;
        int nz = 0;
        for (int q = 0; q < Q; q++) nz |= g[q] != 0.0;
        if (!nz) continue;
        if (p->port_kind[r] >= ST_PORT_BASE) {  /* apply_ports_vjp, ports.rs:582-621 */
            double jac[19 * 19];
            orc_stationary_port_jacobian(after_pre_port + nb * Q, p->port_kind[r], p->port_value + 3 * r, jac);
            for (int i = 0; i < Q; i++) {
                if (g[i] == 0.0) continue;
                for (int k = 0; k < Q; k++) work[nb * Q + k] += jac[i * Q + k] * g[i];
            }
            continue;
        }
        orc_port_populations_vjp(Q, after_pre_port + nb * Q, p->port_kind[r], p->port_value + 3 * r, g, nb_bar);
        for (int k = 0; k < Q; k++) work[nb * Q + k] += nb_bar[k];
    }
    /* stream_vjp: post_bar[s] += reflected_bar[i] for every i whose source is s;
     * previous_bar = reflected_bar at solid cells. */
    double *post_bar = f_bar; /* reuse: filled then consumed cell by cell below */
    memset(post_bar, 0, sizeof(double) * n * Q);
    for (long i = 0; i < n * Q; i++) {
        uint32_t s = source[i];
        if (s != NONE_U32) post_bar[s] += work[i];
    }
    /* the sponge blend is affine in post: post_bar *= 1 - sigma at the listed cells */
    for (int r = 0; r < p->n_sponge_cells; r++) {
        long x = p->sponge_cell[r];
        double keep = 1.0 - p->sponge_sigma[r];
        for (int q = 0; q < Q; q++) post_bar[x * Q + q] *= keep;
    }
    const int wale = wale_on(p);
    double *rate = NULL, *wu = NULL, *wr = NULL, *gbar = NULL;
    if (wale) {
        rate = (double *)malloc(sizeof(double) * n);
        wu = (double *)malloc(sizeof(double) * 3 * n);
        wr = (double *)malloc(sizeof(double) * n);
        gbar = (double *)calloc((size_t)9 * n, sizeof(double));
        wale_rates(p, before, rate, wu, wr);
    }
    const double tau0 = relax_tau(p);
    if (p->psm_law >= 0 && p->occ_m_bar) memset(p->occ_m_bar, 0, sizeof(double) * 3 * n);
    if (p->psm_law >= 0 && p->occ_d1 && p->occ_d1_bar) memset(p->occ_d1_bar, 0, sizeof(double) * n);
    if (p->psm_law >= 0 && p->occ_d1 && p->occ_m1_bar) memset(p->occ_m1_bar, 0, sizeof(double) * 3 * n);
    double g[MAXQ], fb[MAXQ];
#pragma omp parallel for schedule(static) private(g, fb)
    for (long x = 0; x < n; x++) {
        alpha_bar[x] = 0.0;
        if (is_solid(p, x)) {
            for (int q = 0; q < Q; q++) f_bar[x * Q + q] = work[x * Q + q];
            continue;
        }
        int nz = 0;
        for (int q = 0; q < Q; q++) {
            g[q] = post_bar[x * Q + q];
            nz |= g[q] != 0.0;
        }
        if (p->sub_gbar && p->psm_law >= 0) /* the impulse shares' cotangents reach d even at g = 0 */
            for (int e = 0; e < 6; e++) nz |= p->sub_gbar[6 * x + e] != 0.0;
        if (!nz) {
            for (int q = 0; q < Q; q++) f_bar[x * Q + q] = 0.0;
            continue;
        }
        double ab, rb = 0.0;
        if (p->psm_law >= 0) {
            static const double no_dp[3] = {0.0, 0.0, 0.0};
            double mb[3] = {0.0, 0.0, 0.0};
            double d = psm_d(p, alpha, x), m[3];
            psm_mv(p, x, m);
            if (d > 0.0) {
                if (p->sub_gbar) { /* reverse_substep, field.rs:2682-2690 */
                    double post_s[MAXQ], dp[3], dpb[3], extra = 0.0;
                    psm_cell(p, before + x * Q, d, m, post_s, dp);
                    for (int e = 0; e < 3; e++) {
                        double gt = p->sub_gbar[6 * x + e] * (1.0 - p->theta) + p->sub_gbar[6 * x + 3 + e] * p->theta;
                        dpb[e] = 0.0 - gt / d;
                        extra += gt * dp[e] / (d * d);
                    }
                    psm_cell_vjp(p, before + x * Q, d, m, g, dpb, fb, &ab, mb);
                    ab += extra;
                } else {
                    psm_cell_vjp(p, before + x * Q, d, m, g, no_dp, fb, &ab, mb);
                }
            } else {
                cell_vjp_rate(p, before + x * Q, 0.0, -1.0, g, fb, &ab, NULL);
                ab = 0.0;
            }
            for (int q = 0; q < Q; q++) f_bar[x * Q + q] = fb[q];
            if (p->occ_d1) {  /* d = d0 (1 - theta) + d1 theta */
                alpha_bar[x] = ab * (1.0 - p->theta);
                if (p->occ_d1_bar) p->occ_d1_bar[x] = ab * p->theta;
                for (int a = 0; a < 3; a++) {
                    if (p->occ_m_bar) p->occ_m_bar[3 * x + a] = mb[a] * (1.0 - p->theta);
                    if (p->occ_m1_bar) p->occ_m1_bar[3 * x + a] = mb[a] * p->theta;
                }
                continue;
            }
            alpha_bar[x] = ab;
            if (p->occ_m_bar)
                for (int a = 0; a < 3; a++) p->occ_m_bar[3 * x + a] = mb[a];
            continue;
        }
        cell_vjp_rate(p, before + x * Q, alpha[x], wale ? rate[x] : -1.0, g, fb, &ab, wale ? &rb : NULL);
        for (int q = 0; q < Q; q++) f_bar[x * Q + q] = fb[q];
        alpha_bar[x] = ab;
        if (wale) {  /* cotangent of the cell's velocity gradient */
            double gr[3][3], gb[3][3];
            wale_gradient(p, wu, x, gr);
            wale_omega_vjp(p, gr, tau0, rb, gb);
            for (int a = 0; a < 9; a++) gbar[9 * x + a] = gb[a / 3][a % 3];
        }
    }
    if (wale) {
        /* Stencil::gradient_transpose by scatter: u_bar of every cell a gradient used */
        double *ub = (double *)calloc((size_t)3 * n, sizeof(double));
        for (long x = 0; x < n; x++) {
            if (is_solid(p, x)) continue;
            for (int b = 0; b < 3; b++) {
                long pp, mm;
                int kind = wale_diff(p, x, b, &pp, &mm);
                for (int a = 0; a < 3; a++) {
                    double v = gbar[9 * x + 3 * a + b];
                    switch (kind) {
                    case 1: ub[3 * pp + a] += v * 0.5; ub[3 * mm + a] -= v * 0.5; break;
                    case 2: ub[3 * pp + a] += v; ub[3 * x + a] -= v; break;
                    case 3: ub[3 * x + a] += v; ub[3 * mm + a] -= v; break;
                    default: break;
                    }
                }
            }
        }
        /* u = j / rho:  d u_a / d f_q = (c_qa - u_a) / rho */
        const lattice_t *L = lat(Q);
        for (long y = 0; y < n; y++) {
            if (is_solid(p, y)) continue;
            for (int q = 0; q < Q; q++) {
                double acc = 0.0;
                for (int a = 0; a < 3; a++) acc += ub[3 * y + a] * (L->cf[q][a] - wu[3 * y + a]);
                f_bar[y * Q + q] += acc / wr[y];
            }
        }
        free(ub);
    }
    free(rate);
    free(wu);
    free(wr);
    free(gbar);
}

/* ------------------------------------------------------------------------- */
/* Macroscopic fields (solver.rs:690-705 in lattice units) and objectives.      */
/* ------------------------------------------------------------------------- */
void orc_macroscopic(const orc_problem *p, const double *alpha, const double *state, double *rho_out,
                     double *u_out) {
    const lattice_t *L = lat(p->Q);
    long n = (long)p->nx * p->ny * p->nz;
    int Q = p->Q;
    for (long x = 0; x < n; x++) {
        double rho, j[3], den, sv[3], u[3], F[3];
        brinkman_macro(L, state + x * Q, alpha[x], p->accel, &rho, j, &den, sv, u, F);
        rho_out[x] = rho;
        for (int d = 0; d < 3; d++) u_out[3 * x + d] = is_solid(p, x) ? 0.0 : u[d];
    }
}

/* Porous dissipation sum_{fluid} rho*alpha*|u|^2: vector_cell's drag term
 * (solver.rs:835-850) in lattice units (density, velocity, step and dV scales 1). */
double orc_obj_dissipation(const orc_problem *p, const double *alpha, const double *state) {
    const lattice_t *L = lat(p->Q);
    long n = (long)p->nx * p->ny * p->nz;
    int Q = p->Q;
    double acc = 0.0;
#pragma omp parallel for schedule(static) reduction(+ : acc)
    for (long x = 0; x < n; x++) {
        if (is_solid(p, x)) continue;
        double rho, j[3], den, sv[3], u[3], F[3];
        brinkman_macro(L, state + x * Q, alpha[x], p->accel, &rho, j, &den, sv, u, F);
        double usq = u[0] * u[0] + u[1] * u[1] + u[2] * u[2];
        acc += rho * alpha[x] * usq;
    }
    return acc;
}

/* Gradient of weight * sum rho*alpha*|u|^2, accumulated into f_bar / alpha_bar. */
void orc_obj_dissipation_grad(const orc_problem *p, const double *alpha, const double *state, double weight,
                              double *f_bar, double *alpha_bar) {
    const lattice_t *L = lat(p->Q);
    long n = (long)p->nx * p->ny * p->nz;
    int Q = p->Q;
    for (long x = 0; x < n; x++) {
        if (is_solid(p, x)) continue;
        double rho, j[3], den, sv[3], u[3], F[3];
        double al = alpha[x];
        brinkman_macro(L, state + x * Q, al, p->accel, &rho, j, &den, sv, u, F);
        double usq = u[0] * u[0] + u[1] * u[1] + u[2] * u[2];
        /* d = rho*al*usq; u = (j/rho + a/2)/den */
        double rho_bar = weight * al * usq;
        double usq_bar = weight * rho * al;
        double al_bar = weight * rho * usq;
        double jb[3];
        for (int d = 0; d < 3; d++) {
            double ub = usq_bar * 2.0 * u[d];
            al_bar -= ub * sv[d] / (den * den) * 0.5;
            double sb = ub / den;
            jb[d] = sb / rho;
            rho_bar -= sb * j[d] / (rho * rho);
        }
        for (int i = 0; i < Q; i++) {
            double v = rho_bar;
            for (int d = 0; d < 3; d++) v += jb[d] * L->cf[i][d];
            f_bar[x * Q + i] += v;
        }
        alpha_bar[x] += al_bar;
    }
}

/* Pressure drop between two planes normal to `axis`: mean(rho/3) over fluid cells
 * of plane a minus that of plane b (lattice units, cs^2 = 1/3). */
double orc_obj_pressure_drop(const orc_problem *p, const double *state, int axis, int plane_a, int plane_b) {
    long n = (long)p->nx * p->ny * p->nz;
    int Q = p->Q;
    double sa = 0.0, sb = 0.0;
    long na = 0, nb = 0;
    for (long x = 0; x < n; x++) {
        if (is_solid(p, x)) continue;
        long c[3];
        g_coords(p, x, c);
        if (c[axis] != plane_a && c[axis] != plane_b) continue;
        double rho = state[x * Q];
        for (int q = 1; q < Q; q++) rho += state[x * Q + q];
        if (c[axis] == plane_a) {
            sa += rho;
            na++;
        } else {
            sb += rho;
            nb++;
        }
    }
    return (sa / (double)na - sb / (double)nb) / 3.0;
}

void orc_obj_pressure_drop_grad(const orc_problem *p, int axis, int plane_a, int plane_b, double weight,
                                double *f_bar) {
    long n = (long)p->nx * p->ny * p->nz;
    int Q = p->Q;
    long na = 0, nb = 0;
    for (long x = 0; x < n; x++) {
        if (is_solid(p, x)) continue;
        long c[3];
        g_coords(p, x, c);
        if (c[axis] == plane_a) na++;
        else if (c[axis] == plane_b) nb++;
    }
    double ga = weight / (3.0 * (double)na), gb = -weight / (3.0 * (double)nb);
    for (long x = 0; x < n; x++) {
        if (is_solid(p, x)) continue;
        long c[3];
        g_coords(p, x, c);
        double gv = c[axis] == plane_a ? ga : (c[axis] == plane_b ? gb : 0.0);
        if (gv == 0.0) continue;
        for (int q = 0; q < Q; q++) f_bar[x * Q + q] += gv;
    }
}

/* ------------------------------------------------------------------------- */
/* Terminal responses: vector_cell / vector / vector_vjp (solver.rs:835-947).  */
/* scales = {velocity, density, step_s, dv}; out = [mean u_x, u_y, u_z over   */
/* fluid cells, 0.5 sum rho|u|^2 dV, sum rho alpha/dt |u|^2 dV, sum phi dV].   */
/* The VJP is the forward-mode gradient the reference takes (one tangent per   */
/* pass over the Q + 2 inputs f, alpha, phi).                                  */
/* ------------------------------------------------------------------------- */
static void vec_moments(const orc_problem *p, const double *f, double *rho, double m[3]) {
    if (p->Q == 19) {
        st_moments(f, rho, m);
    } else {
        moments(lat(p->Q), f, rho, m);
    }
}

static void vector_cell(const orc_problem *p, const double *f, double alpha, double phi, int solid,
                        const double sc[4], double out[6]) {
    double rho, m[3];
    vec_moments(p, f, &rho, m);
    double s = alpha * 0.5 + 1.0;
    double u[3];
    for (int d = 0; d < 3; d++) u[d] = (solid ? 0.0 : (m[d] / rho + 0.5 * p->accel[d]) / s) * sc[0];
    double r = rho * sc[1];
    double usq = u[0] * u[0] + u[1] * u[1] + u[2] * u[2];
    out[0] = u[0];
    out[1] = u[1];
    out[2] = u[2];
    out[3] = solid ? 0.0 : r * usq;
    out[4] = solid ? 0.0 : r * alpha / sc[2] * usq;
    out[5] = phi;
}

void orc_vector(const orc_problem *p, const double *phi, const double *alpha, const double *state, const double sc[4],
                double out[6]) {
    long n = (long)p->nx * p->ny * p->nz;
    double sums[6] = {0, 0, 0, 0, 0, 0};
    double fluid = 0.0;
    for (long x = 0; x < n; x++) fluid += is_solid(p, x) ? 0.0 : 1.0;
    for (long x = 0; x < n; x++) {
        double c[6];
        vector_cell(p, state + x * p->Q, alpha[x], phi ? phi[x] : 0.0, is_solid(p, x), sc, c);
        for (int k = 0; k < 6; k++) sums[k] += c[k];
    }
    out[0] = sums[0] / fluid;
    out[1] = sums[1] / fluid;
    out[2] = sums[2] / fluid;
    out[3] = 0.5 * sums[3] * sc[3];
    out[4] = sums[4] * sc[3];
    out[5] = sums[5] * sc[3];
}

static dual dual_sum_c(const orc_problem *p, const dual *f, int d) {
    /* stationary ports.rs moments order: acc starts at 0, q ascending */
    const lattice_t *L = lat(p->Q);
    dual acc = dv(0.0);
    for (int q = 0; q < p->Q; q++) {
        if (L->c[q][d] == 1) acc = dadd(acc, f[q]);
        else if (L->c[q][d] == -1) acc = dsub(acc, f[q]);
    }
    return acc;
}

static dual vector_cell_weighted_dual(const orc_problem *p, const dual *f, dual alpha, dual phi, int solid,
                                      const double sc[4], const double wts[6]) {
    dual rho = f[0];
    for (int q = 1; q < p->Q; q++) rho = dadd(rho, f[q]);
    dual m[3];
    for (int d = 0; d < 3; d++) m[d] = dual_sum_c(p, f, d);
    dual s = dadd(dmuls(alpha, 0.5), dv(1.0));
    dual u[3];
    for (int d = 0; d < 3; d++) {
        dual ud = ddiv(dadd(ddiv(m[d], rho), dv(0.5 * p->accel[d])), s);
        u[d] = dmuls(solid ? dv(0.0) : ud, sc[0]);
    }
    dual r = dmuls(rho, sc[1]);
    dual usq = dadd(dadd(dmul(u[0], u[0]), dmul(u[1], u[1])), dmul(u[2], u[2]));
    dual c[6];
    c[0] = u[0];
    c[1] = u[1];
    c[2] = u[2];
    c[3] = solid ? dv(0.0) : dmul(r, usq);
    c[4] = solid ? dv(0.0) : dmul(ddivs(dmul(r, alpha), sc[2]), usq);
    c[5] = phi;
    dual acc = dmuls(c[0], wts[0]);
    for (int k = 1; k < 6; k++) acc = dadd(acc, dmuls(c[k], wts[k]));
    return acc;
}

void orc_vector_vjp(const orc_problem *p, const double *phi, const double *alpha, const double *state,
                    const double sc[4], const double w[6], double *f_bar, double *alpha_bar, double *phi_bar) {
    long n = (long)p->nx * p->ny * p->nz;
    int Q = p->Q;
    double fluid = 0.0;
    for (long x = 0; x < n; x++) fluid += is_solid(p, x) ? 0.0 : 1.0;
    double wts[6] = {w[0] / fluid, w[1] / fluid, w[2] / fluid, 0.5 * w[3] * sc[3], w[4] * sc[3], w[5] * sc[3]};
#pragma omp parallel for schedule(static)
    for (long x = 0; x < n; x++) {
        dual in[MAXQ + 2];
        for (int t = 0; t < Q + 2; t++) {
            for (int q = 0; q < Q; q++) in[q] = dv(state[x * Q + q]);
            in[Q] = dv(alpha[x]);
            in[Q + 1] = dv(phi ? phi[x] : 0.0);
            in[t].d = 1.0;
            dual acc = vector_cell_weighted_dual(p, in, in[Q], in[Q + 1], is_solid(p, x), sc, wts);
            if (t < Q) f_bar[x * Q + t] = acc.d;
            else if (t == Q) alpha_bar[x] = acc.d;
            else phi_bar[x] = acc.d;
        }
    }
}

/* ------------------------------------------------------------------------- */
/* Coupled D3Q7 advection-diffusion of a temperature T (double distribution).  */
/* Not in the reference (its thermal coupling is finite-volume transport on    */
/* D3Q19 links, lbm/thermal.rs:91-134); the model is defined here:             */
/*   T = sum_i g_i,  g_i^eq = w_i T (1 + 4 c_i.u),  u = the Brinkman velocity   */
/*   of the flow state the same step collides (brinkman_macro),                */
/*   kappa(phi) = kappa_f (1 + (ratio - 1) phi),  tau_T = 1/2 + 4 kappa,       */
/*   S(phi) = source * phi,                                                    */
/*   g_i* = g_i - omega_T (g_i - g_i^eq) + w_i S,                              */
/* streamed with halfway bounce-back (adiabatic walls) and, at the flow's port */
/* cells after streaming: velocity ports hold T = T_in by non-equilibrium      */
/* extrapolation g_b = w T_in + (g_nb - w T_nb); pressure ports are            */
/* zero-gradient, g_b = g_nb.  Hard solid cells are not part of this model.    */
/* ------------------------------------------------------------------------- */
typedef struct {
    double kappa;  /* fluid diffusivity (lattice units) */
    double ratio;  /* solid:fluid diffusivity ratio */
    double source; /* heat source per step at phi = 1 */
    double inlet;  /* T held at velocity ports */
} orc_thermal;

static inline double th_omega(const orc_thermal *th, double phi) {
    return 1.0 / (0.5 + 4.0 * (th->kappa * (1.0 + (th->ratio - 1.0) * phi)));
}

/* One thermal cell: g (7) and the flow velocity u -> post-collision g*. */
static void thermal_cell(const double *g, const double u[3], double om, double S, double *out) {
    const lattice_t *L = lat(7);
    double T = 0.0;
    for (int q = 0; q < 7; q++) T += g[q];
    for (int q = 0; q < 7; q++) {
        double cu = dot_c(L, q, u);
        double eq = L->w[q] * T * (1.0 + 4.0 * cu);
        out[q] = g[q] - om * (g[q] - eq) + L->w[q] * S;
    }
}

/* VJP of thermal_cell: gb = cotangent of g*; returns g_bar (7), u_bar (3), and the
 * cotangents of omega and S. */
static void thermal_cell_vjp(const double *g, const double u[3], double om, const double *gb, double *g_bar,
                             double u_bar[3], double *om_bar, double *S_bar) {
    const lattice_t *L = lat(7);
    double T = 0.0;
    for (int q = 0; q < 7; q++) T += g[q];
    double t_bar = 0.0, ob = 0.0, sb = 0.0;
    u_bar[0] = u_bar[1] = u_bar[2] = 0.0;
    for (int q = 0; q < 7; q++) {
        double cu = dot_c(L, q, u);
        double eq = L->w[q] * T * (1.0 + 4.0 * cu);
        double eq_bar = om * gb[q];
        g_bar[q] = (1.0 - om) * gb[q];
        ob += gb[q] * (eq - g[q]);
        sb += gb[q] * L->w[q];
        t_bar += eq_bar * L->w[q] * (1.0 + 4.0 * cu);
        double cu_bar = eq_bar * L->w[q] * T * 4.0;
        for (int d = 0; d < 3; d++) u_bar[d] += cu_bar * L->cf[q][d];
    }
    for (int q = 0; q < 7; q++) g_bar[q] += t_bar;
    *om_bar = ob;
    *S_bar = sb;
}

void orc_thermal_cell(const double *g, const double u[3], double om, double S, double *out) {
    thermal_cell(g, u, om, S, out);
}

void orc_thermal_cell_vjp(const double *g, const double u[3], double om, const double *gb, double *g_bar,
                          double *u_bar, double *om_bar, double *S_bar) {
    thermal_cell_vjp(g, u, om, gb, g_bar, u_bar, om_bar, S_bar);
}

void orc_thermal_stream_table(const orc_problem *p, uint32_t *source) {
    orc_problem q7 = *p;
    q7.Q = 7;
    q7.solid = NULL;
    orc_stream_table(&q7, source);
}

void orc_thermal_initial(const orc_problem *p, double *g0) {
    long n = (long)p->nx * p->ny * p->nz;
    for (long i = 0; i < n * 7; i++) g0[i] = 0.0;
}

static void thermal_ports(const orc_problem *p, const orc_thermal *th, double *after) {
    const lattice_t *L = lat(7);
    for (int r = 0; r < p->n_port_cells; r++) {
        long x = p->port_cell[r], nb = p->port_nb[r];
        const double *gn = after + nb * 7;
        if (p->port_kind[r] == 0) {
            double Tn = 0.0;
            for (int q = 0; q < 7; q++) Tn += gn[q];
            for (int q = 0; q < 7; q++) after[x * 7 + q] = L->w[q] * th->inlet + (gn[q] - L->w[q] * Tn);
        } else {
            for (int q = 0; q < 7; q++) after[x * 7 + q] = gn[q];
        }
    }
}

/* One thermal step on the flow state f_before (pre-collision, the state the flow step of
 * the same interval reads): collide, stream with source7 (orc_thermal_stream_table),
 * ports.  post is scratch (n*7). */
void orc_thermal_interval(const orc_problem *p, const orc_thermal *th, const uint32_t *source7, const double *alpha,
                          const double *phi, const double *f_before, const double *g_before, double *post,
                          double *g_after) {
    const lattice_t *L = lat(p->Q);
    long n = (long)p->nx * p->ny * p->nz;
    int Q = p->Q;
#pragma omp parallel for schedule(static)
    for (long x = 0; x < n; x++) {
        double rho, j[3], den, sv[3], u[3], F[3];
        brinkman_macro(L, f_before + x * Q, alpha[x], p->accel, &rho, j, &den, sv, u, F);
        thermal_cell(g_before + x * 7, u, th_omega(th, phi[x]), th->source * phi[x], post + x * 7);
    }
#pragma omp parallel for schedule(static)
    for (long i = 0; i < n * 7; i++) g_after[i] = post[source7[i]];
    thermal_ports(p, th, g_after);
}

/* VJP of orc_thermal_interval.  g_after_bar (n*7) is consumed (the ports' transpose
 * works on it in place).  Outputs: g_bar (n*7, overwritten); f_bar, alpha_bar, phi_bar
 * accumulated (+=). */
void orc_thermal_interval_vjp(const orc_problem *p, const orc_thermal *th, const uint32_t *source7,
                              const double *alpha, const double *phi, const double *f_before,
                              const double *g_before, double *g_after_bar, double *post_bar, double *g_bar,
                              double *f_bar, double *alpha_bar, double *phi_bar) {
    const lattice_t *L = lat(p->Q);
    const lattice_t *L7p = lat(7);
    long n = (long)p->nx * p->ny * p->nz;
    int Q = p->Q;
    /* ports, in reverse order of application */
    for (int r = p->n_port_cells - 1; r >= 0; r--) {
        long x = p->port_cell[r], nb = p->port_nb[r];
        double *bx = g_after_bar + x * 7, *bn = g_after_bar + nb * 7;
        if (p->port_kind[r] == 0) {
            double s = 0.0;
            for (int q = 0; q < 7; q++) s += L7p->w[q] * bx[q];
            for (int q = 0; q < 7; q++) bn[q] += bx[q] - s;
        } else {
            for (int q = 0; q < 7; q++) bn[q] += bx[q];
        }
        for (int q = 0; q < 7; q++) bx[q] = 0.0;
    }
    /* stream: every post slot has exactly one reader */
    for (long i = 0; i < n * 7; i++) post_bar[source7[i]] = g_after_bar[i];
#pragma omp parallel for schedule(static)
    for (long x = 0; x < n; x++) {
        double rho, j[3], den, sv[3], u[3], F[3];
        const double *f = f_before + x * Q;
        brinkman_macro(L, f, alpha[x], p->accel, &rho, j, &den, sv, u, F);
        double om = th_omega(th, phi[x]);
        double ub[3], omb, sb;
        thermal_cell_vjp(g_before + x * 7, u, om, post_bar + x * 7, g_bar + x * 7, ub, &omb, &sb);
        /* u = (j/rho + a/2) / (1 + alpha/2) */
        double v[3], s[3];
        for (int d = 0; d < 3; d++) {
            v[d] = j[d] / rho;
            s[d] = ub[d] / den / rho;
        }
        for (int q = 0; q < Q; q++) {
            double acc = 0.0;
            for (int d = 0; d < 3; d++) acc += s[d] * (L->cf[q][d] - v[d]);
            f_bar[x * Q + q] += acc;
        }
        double ab = 0.0;
        for (int d = 0; d < 3; d++) ab += ub[d] * u[d];
        alpha_bar[x] += -0.5 * ab / den;
        /* omega = 1 / (1/2 + 4 kappa_f (1 + (ratio - 1) phi)) */
        double dom = -om * om * 4.0 * th->kappa * (th->ratio - 1.0);
        phi_bar[x] += omb * dom + sb * th->source;
    }
}

/* Mean temperature over the cells with phi > 1/2 ("solid" cells of the design). */
double orc_obj_solid_temperature(const orc_problem *p, const double *phi, const double *g) {
    long n = (long)p->nx * p->ny * p->nz;
    double acc = 0.0;
    long cnt = 0;
    for (long x = 0; x < n; x++) {
        if (!(phi[x] > 0.5)) continue;
        double T = 0.0;
        for (int q = 0; q < 7; q++) T += g[x * 7 + q];
        acc += T;
        cnt++;
    }
    return cnt ? acc / (double)cnt : 0.0;
}

void orc_obj_solid_temperature_grad(const orc_problem *p, const double *phi, double weight, double *g_bar) {
    long n = (long)p->nx * p->ny * p->nz;
    long cnt = 0;
    for (long x = 0; x < n; x++) cnt += phi[x] > 0.5;
    if (!cnt) return;
    double c = weight / (double)cnt;
    for (long x = 0; x < n; x++)
        if (phi[x] > 0.5)
            for (int q = 0; q < 7; q++) g_bar[x * 7 + q] += c;
}

void orc_temperature(const orc_problem *p, const double *g, double *T) {
    long n = (long)p->nx * p->ny * p->nz;
    for (long x = 0; x < n; x++) {
        double t = 0.0;
        for (int q = 0; q < 7; q++) t += g[x * 7 + q];
        T[x] = t;
    }
}
