// Fused stream-collide kernels for gfx950 (forward, discrete adjoint, ports,
// objectives, layout conversion).  Included by the per-precision translation units.
//
// Data layout in HBM
//   Populations are structure-of-arrays: array q holds one value per storage cell,
//   storage cell (i,j,k) at i*sxy + j*SZ + k (z fastest, as the reference's Grid,
//   d3q19.rs:86-89).  Non-periodic axes carry one extra storage layer at index n
//   (the esoteric-twist ghost layer); the z pitch SZ is nz + 1.
//
// In-place esoteric-twist streaming
//   With a_q = max(c_q, 0) and a_opp(q) = max(-c_q, 0) (so a_q - a_opp(q) = c_q), a
//   step on a state of parity p (p = t mod 2) at cell x does
//     read   f_q  from P[(p xor blocked_q) ? q : opp(q)] at x + a_opp(q)
//     write  f*_q to   P[p ? opp(q) : q]                 at x + a_q
//   where blocked_q means x - c_q is solid or outside a non-periodic face (halfway
//   bounce-back, solver.rs:84-105).  Every storage location belongs to exactly one
//   (cell, link), so one buffer is updated in place without races, each population
//   is read and written once per step, and the state after the step is the
//   reference's post-stream state (solver.rs:604-614) in the layout of parity 1-p.
//   Solid cells are never updated; their slots carry the reflected populations of
//   their fluid neighbours.  The same step may also run out of place (read buffer A,
//   write buffer B): the locations it writes are exactly those the next step reads,
//   so B needs no other content -- this lets the checkpoint driver start a step from
//   a snapshot slot, or finish one into a slot, without copying whole states.
//
// Discrete adjoint in place
//   The cotangent of f*_q(x) is the cotangent of whoever reads that value next step,
//   and that reader stores its cotangent exactly where f*_q(x) was written.  So the
//   adjoint of step t reads post_bar_q(x) at the forward write location of (x,q),
//   applies the transposed collision Jacobian, and writes f_bar_q(x) at the forward
//   read location of (x,q): solver.rs:642-674 (stream_vjp + collide_vjp) in one pass.
#pragma once
#include <type_traits>
#include <hip/hip_fp16.h>
#include <hip/hip_runtime.h>

#include <cstdint>

#include "cell.h"
#include "cumulant.h"
#include "dual.h"
#include "les.h"

namespace lbm {

enum Prec : int { P_F64 = 0, P_F32 = 1, P_F16S = 2, P_BF16S = 3 };
// K_TRT_LES / K_BGK_LES: TRT / BGK with the Smagorinsky closure (les.h); their constants
// [magic, k, floor] travel in the collision-constant array (`mrt`).  K_TRT_WALE /
// K_BGK_WALE: TRT / BGK with the WALE closure (wale.h), constants [magic, C^2, floor].
enum Kind : int {
    K_BGK = 0, K_TRT = 1, K_MRT = 2, K_CUMULANT = 3, K_BGK_GUO = 4, K_STREAM_ONLY = 5, K_TRT_LES = 6, K_BGK_LES = 7,
    K_TRT_WALE = 8, K_BGK_WALE = 9, K_TRT_PSM = 10, K_BGK_PSM = 11, K_TRT_PSMS = 12, K_BGK_PSMS = 13,
    K_TRT_BRO = 14, K_BGK_BRO = 15
};
constexpr bool is_wale(int kind) { return kind == K_TRT_WALE || kind == K_BGK_WALE; }
// K_TRT_PSM / K_BGK_PSM (Psm law), K_TRT_PSMS / K_BGK_PSMS (PsmSuperposition) and
// K_TRT_BRO / K_BGK_BRO (Brinkman law with a moving solid): TRT / BGK with the occupancy
// coupling (psm.h); the design array holds the occupancy d, constants
// [magic, kappa, law, c_alpha, drag shape].
constexpr bool is_psm(int kind) { return kind >= K_TRT_PSM && kind <= K_BGK_BRO; }
constexpr bool psm_sup(int kind) { return kind == K_TRT_PSMS || kind == K_BGK_PSMS; }
constexpr bool is_bro(int kind) { return kind == K_TRT_BRO || kind == K_BGK_BRO; }
constexpr int psm_base(int kind) {
    return kind == K_TRT_PSM || kind == K_TRT_PSMS || kind == K_TRT_BRO ? K_TRT : K_BGK;
}
// OBJ_SOLID_FORCE: w . (momentum the coupled cells exchange with the solid in the
// collision of the state read), w = vw[0..2] (occupancy couplings only).
enum ObjKind : int { OBJ_NONE = 0, OBJ_DISSIPATION = 1, OBJ_PRESSURE_DROP = 2, OBJ_VECTOR = 3, OBJ_SOLID_FORCE = 5 };

constexpr uint32_t SOLID_BIT = 0x80000000u;

// Storage geometry (the row-blocked layout is described with the addressing below).
struct Geom {
    int nx, ny, nz;
    int SX, SY, SZ;  // storage planes, rows per plane (ghost layers included) and row length
    int GY;          // +z ghost rows at the end of each plane (0 for a periodic z axis)
    int px, py, pz;
};

// Objective evaluated on one state: weight * phi(state), phi per cell as
//   dissipation:    rho * alpha * |u|^2 over fluid cells
//   pressure drop:  (rho - 1)/3 * (1/na on plane a, -1/nb on plane b)
struct ObjArgs {
    int kind;
    int axis, plane_a, plane_b;
    double weight;
    double inv_na, inv_nb;
    // OBJ_VECTOR: per-cell weights of the terminal responses (solver.rs:896-915):
    // [w0/fluid, w1/fluid, w2/fluid, 0.5 w3 dV, w4 dV] on [u_x, u_y, u_z, rho|u|^2,
    // rho alpha/dt |u|^2] in physical units (u * vel, rho * dens).
    double vw[5];
    double vel, dens, inv_dt;
    // Coupled temperature field (host side only): 1 / (cells with phi > 1/2) in the
    // template, and the stage weight per such cell on T of the state.
    double inv_ns, tw;
};

struct StepArgs {
    Geom g;
    void* pop;        // state read (parity odd)
    void* pop_out;    // buffer written (== pop for an in-place step)
    const void* alpha;
    const uint32_t* mask;
    double omega, om, tau;
    double a[3];
    const void* mrt;
    int odd;
    int diag;         // 1: accumulate diagnostics
    uint32_t step;    // index of the interval (rounding key of stochastically rounded stores)
    // Coupled temperature field (k_forward with TH): temperature populations read and
    // written in the same pass, the thermal design, its constants, the FP16 storage scale
    // and the stage weight on T per cell with phi > 1/2.
    const void* tpop;
    void* tpop_out;
    const void* phi;
    double kappa, ratio_m1, source, tw;
    float tenc, tdec;
    const void* wu;   // WALE: u (3 x cells) and rho (cells) of the state read, arithmetic precision
    const void* sm;   // occupancy coupling: solid momentum m (3 x cells), arithmetic precision
    const void* d1;   // occupancy coupling: second endpoint d1 (cells) and m1 (3 x cells), or null
    const void* sm1;
    double theta;     // weight of the second endpoint in this step
    double* gacc;     // subcycle: impulse shares g0, g1 of the endpoints (6 x cells, SoA, FP64), or null
    ObjArgs obj;      // objective of the state being read (kind 0: none)
    double* part;     // per-block partials [J, mass, max |u|^2, bad cells, max -f], PART_STRIDE per block
    long long npart;  // capacity of part, in blocks
    double* stage;    // FOLD_MAX_BLOCKS x FOLD_VALUES fold stage and its ticket counter (k_fold)
    unsigned* ticket;
    int* lsc;         // scale record of FP16 cotangents (seed writes), else null
    double* jacc;     // objective accumulator      (written by k_fold)
    double* macc;     // fluid mass of the state    (written by k_fold)
    unsigned* dstat;  // [0] max |u|^2 (float bits), [1] bad cells (non-finite or f <= 0),
                      // [2] float_key of max -f over fluid populations (0: none seen)
};

struct AdjArgs {
    Geom g;
    const void* pop;
    void* lam;
    const void* alpha;
    double* alpha_bar;
    const uint32_t* mask;
    double omega, om, tau;
    double a[3];
    const void* mrt;
    int odd;
    ObjArgs obj;  // stage cost of the state being read (added to its cotangent)
    // Scaled FP16 cotangents only: per-block max |lambda| partials, the scale record and
    // the bound on the stage-cost seed of the following reverse step.
    double* part;
    long long npart;
    double* stage;
    unsigned* ticket;
    int* lsc;
    double s_next;
    // Coupled temperature field: cotangent of the Brinkman velocity from the thermal
    // step (SoA, 3 x cells, arithmetic precision), added through u(f, alpha); else null.
    const void* ubar;
    // WALE: u and rho of the state read (as StepArgs::wu) and the cotangent of every
    // cell's velocity gradient written for k_wale_transpose (9 x cells, G[a][b] at 3a+b).
    const void* wu;
    void* wgb;
    // Occupancy coupling: solid momentum (as StepArgs::sm) and its cotangent, FP64
    // (3 x cells), accumulated over the reverse sweep.
    const void* sm;
    double* sm_bar;
    // Second occupancy endpoint (as StepArgs) and its FP64 cotangents.
    const void* d1;
    const void* sm1;
    double theta;
    double* d1_bar;
    double* sm1_bar;
    // Subcycle: cotangents of the impulse shares g0, g1 (as StepArgs::gacc), or null.
    const double* gbar;
    // Runs of reverse steps may gather alpha_bar in an FP32 partial per cell (abar32) and
    // fold it into the FP64 alpha_bar once per run: ABAR_FIRST writes the partial,
    // ABAR_MID adds to it, ABAR_FOLD adds partial + this step's term to alpha_bar.
    // abar32 null: every step adds to alpha_bar directly.
    float* abar32;
    int abar_mode;
};

enum { ABAR_FIRST = 0, ABAR_MID = 1, ABAR_FOLD = 2 };

// 32-bit integer hash (xorshift-multiply, constants of the "lowbias32" finaliser).
LBM_DEV uint32_t mix32(uint32_t x) {
    x ^= x >> 16;
    x *= 0x7feb352du;
    x ^= x >> 15;
    x *= 0x846ca68bu;
    x ^= x >> 16;
    return x;
}

// FP16 with stochastic rounding: the 13 float significand bits below the FP16 ones are
// replaced by a carry from r's low 13 bits, so a value rounds up with probability equal
// to its distance from the lower FP16 neighbour (E[result] = x).  A field that moves by
// small increments per step next to an 11-bit stored value turns nearest rounding into a
// bias that accumulates over thousands of steps; stochastic rounding makes it zero-mean
// noise.  Magnitudes below the FP16 normal range round to nearest.
LBM_DEV _Float16 f16_stochastic(float x, uint32_t r) {
    const uint32_t b = __float_as_uint(x);
    if ((b & 0x7fffffffu) < 0x38800000u) return _Float16(x);
    return _Float16(__uint_as_float((b + (r & 0x1fffu)) & ~0x1fffu));
}

// ---------------------------------------------------------------------------------
// Storage codecs.  C is the arithmetic type; T the stored type.
// ---------------------------------------------------------------------------------
template <int P>
struct Store;

template <>
struct Store<P_F64> {
    using T = double;
    using C = double;
    static constexpr bool shifted = false;  // holds f itself
    template <int Q, int q>
    static LBM_DEV C dec(T v) {
        return v;
    }
    template <int Q, int q>
    static LBM_DEV T enc(C v) {
        return v;
    }
    template <int Q, int q>
    static LBM_DEV C ld(const T* p) {
        return *p;
    }
    template <int Q, int q>
    static LBM_DEV void st(T* p, C v) {
        *p = v;
    }
};

// FP32 holding d = f - w (same bytes as f, more significant bits of the deviation).
template <>
struct Store<P_F32> {
    using T = float;
    using C = float;
    static constexpr bool shifted = true;
    template <int Q, int q>
    static LBM_DEV C dec(T v) {
        return v;
    }
    template <int Q, int q>
    static LBM_DEV T enc(C v) {
        return v;
    }
    template <int Q, int q>
    static LBM_DEV C ld(const T* p) {
        return *p;
    }
    template <int Q, int q>
    static LBM_DEV void st(T* p, C v) {
        *p = v;
    }
};

// FP16 holding d_q = f_q - w_q times 2^k_q; the arithmetic value is d (FP32).
template <>
struct Store<P_F16S> {
    using T = _Float16;
    using C = float;
    static constexpr bool shifted = true;
    template <int Q, int q>
    static LBM_DEV C dec(T v) {
        constexpr float inv = 1.0f / float(1u << fp16_exp<Q>(q));
        return float(v) * inv;
    }
    template <int Q, int q>
    static LBM_DEV T enc(C v) {
        constexpr float s = float(1u << fp16_exp<Q>(q));
        return _Float16(v * s);
    }
    template <int Q, int q>
    static LBM_DEV C ld(const T* p) {
        return dec<Q, q>(*p);
    }
    template <int Q, int q>
    static LBM_DEV void st(T* p, C v) {
        *p = enc<Q, q>(v);
    }
};

LBM_DEV float bf16_to_f32(uint16_t h) { return __uint_as_float(uint32_t(h) << 16); }
LBM_DEV uint16_t f32_to_bf16(float f) {
    uint32_t u = __float_as_uint(f);
    u += 0x7fffu + ((u >> 16) & 1u);  // round to nearest even
    return uint16_t(u >> 16);
}

// BF16 holding d_q = f_q - w_q; the arithmetic value is d (FP32).
template <>
struct Store<P_BF16S> {
    using T = uint16_t;
    using C = float;
    static constexpr bool shifted = true;
    template <int Q, int q>
    static LBM_DEV C dec(T v) {
        return bf16_to_f32(v);
    }
    template <int Q, int q>
    static LBM_DEV T enc(C v) {
        return f32_to_bf16(v);
    }
    template <int Q, int q>
    static LBM_DEV C ld(const T* p) {
        return bf16_to_f32(*p);
    }
    template <int Q, int q>
    static LBM_DEV void st(T* p, C v) {
        *p = f32_to_bf16(v);
    }
};

// Full population f from the arithmetic value of a codec (d + w when shifted).
template <bool SH, int Q, class C>
LBM_DEV void to_full(const C* v, C* f) {
#pragma unroll
    for (int q = 0; q < Q; q++) f[q] = SH ? v[q] + C(weight<Q>(q)) : v[q];
}
template <bool SH, int Q, class C>
LBM_DEV void from_full(const C* f, C* v) {
#pragma unroll
    for (int q = 0; q < Q; q++) v[q] = SH ? f[q] - C(weight<Q>(q)) : f[q];
}

// Cotangent (adjoint) storage: plain, no shift.
template <int P>
struct LamStore;
// Scale of FP16 cotangents: lambda is stored as lambda * 2^e with one exponent for the
// whole field.  The record lsc[LAM_REC_INTS] on the device holds
//   lsc[0]  e of the field as stored now,
//   lsc[1]  e for the next whole-field write (an adjoint step),
//   lsc[2]  max |lambda| of the last adjoint step (float bits),
//   lsc[3]  count of adjoint steps whose scaled values left the FP16 range,
//   lsc[LAM_SLOT0 ..]  LAM_SLOTS max slots the blocks of an adjoint step raise.
// Readers decode with lsc[0]; an adjoint step writes with lsc[1], after which
// k_lam_update folds the slots, moves lsc[1] into lsc[0] and picks the next exponent
// from the measured max.  Writers that do not replace the whole field (ports, seeds)
// keep lsc[0].
constexpr int LAM_SLOT0 = 8, LAM_SLOTS = 64, LAM_REC_INTS = LAM_SLOT0 + LAM_SLOTS;
struct LScale {
    float dec, enc;
};
template <int LP>
LBM_DEV LScale lam_scale(const int* lsc, bool whole_write) {
    if constexpr (LP != P_F16S) {
        return LScale{1.0f, 1.0f};
    } else {
        const int ec = lsc[0];
        return LScale{ldexpf(1.0f, -ec), ldexpf(1.0f, whole_write ? lsc[1] : ec)};
    }
}

// Growth allowed between two measured maxima (one port pass plus one transposed step),
// and the exponent that maps a bound b on |lambda| to at most 2^14 (FP16 overflows at
// 65504 ~ 2^16; values down to 2^-14 stay normal, 28 binades below the bound).
constexpr double LAM_GROWTH = 256.0;
LBM_HD inline int lam_exponent(double b) {
    if (!(b > 0.0) || !(b < 1e300)) return 0;
    int ex = 0;
    (void)frexp(b, &ex);  // b < 2^ex
    const int e = 14 - ex;
    return e < -100 ? -100 : (e > 100 ? 100 : e);
}

template <>
struct LamStore<P_F64> {
    using T = double;
    template <class C>
    static LBM_DEV C ld(const T* p, LScale = LScale{1.0f, 1.0f}) {
        return C(*p);
    }
    template <class C>
    static LBM_DEV void st(T* p, C v, LScale = LScale{1.0f, 1.0f}) {
        *p = double(v);
    }
    template <class C>
    static LBM_DEV T enc(C v, LScale = LScale{1.0f, 1.0f}) {
        return double(v);
    }
};
template <>
struct LamStore<P_F32> {
    using T = float;
    template <class C>
    static LBM_DEV C ld(const T* p, LScale = LScale{1.0f, 1.0f}) {
        return C(*p);
    }
    template <class C>
    static LBM_DEV void st(T* p, C v, LScale = LScale{1.0f, 1.0f}) {
        *p = float(v);
    }
    template <class C>
    static LBM_DEV T enc(C v, LScale = LScale{1.0f, 1.0f}) {
        return float(v);
    }
};
// FP16 cotangent times 2^e (LScale carries 2^-e for loads and the write exponent).
template <>
struct LamStore<P_F16S> {
    using T = _Float16;
    template <class C>
    static LBM_DEV C ld(const T* p, LScale s) {
        return C(float(*p) * s.dec);
    }
    template <class C>
    static LBM_DEV void st(T* p, C v, LScale s) {
        *p = _Float16(float(v) * s.enc);
    }
    template <class C>
    static LBM_DEV T enc(C v, LScale s) {
        return _Float16(float(v) * s.enc);
    }
};

// ---------------------------------------------------------------------------------
// Addressing.
// ---------------------------------------------------------------------------------
// Row-blocked layout of a field of Qn population arrays.  A storage row (i, j) holds the
// Qn arrays of its z row one after the other, SZ elements each (SZ: nz rounded up to 64,
// so every array row starts on a 128-byte line):
//   element (arr, i, j, k) = i * PP + j * RP + apos(arr) * SZ + k,
//   RP = Qn * SZ (row pitch),  PP = (SY + GY) * RP (plane pitch).
// A non-periodic x or y axis carries one extra storage plane / row (the esoteric-twist
// ghost layer).  On a non-periodic z axis the +z ghost slot of row (i, j) sits in GY =
// ceil(SY / SZ) ghost rows at the end of plane i: element
//   i * PP + (SY + j / SZ) * RP + apos(arr) * SZ + j % SZ,
// so a row of one array is fetched as whole lines, and the ghost slots of 64 rows share
// a line.  Compared with one array after the other (structure of arrays), the 2 x Qn
// streams of a step then fall in a few kB per row instead of Qn regions a field apart,
// which measured 6-7 % more bandwidth for the twisted access pattern at 256^3.
// An address is a row index (64-bit, uniform over a block of a step launch, where the row
// is (blockIdx.z, blockIdx.y)) plus a 32-bit lane offset within the plane: the global
// accesses take the row part in SGPRs and the lane part in one VGPR (an unsigned 32-bit
// byte offset).  The host refuses a lattice whose x plane of populations reaches 2 GiB.
template <int Qn>
struct Cursor {
    struct Loc {
        long long row;  // element index of the row block
        uint32_t lane;  // element offset from it (array 0)
    };
    long long R[4];  // row blocks of the cell's row and its +x, +y, +x+y neighbours (wrapped)
    uint32_t k, kz;  // the cell's z position and that of its +z neighbour (wrapped)
    uint32_t G[4];   // +z ghost slot of each of the four rows, relative to its row block
    bool zghost;     // the +z neighbour is the ghost slot (k = nz-1, z not periodic)
    LBM_DEV Loc at(int corner) const {
        const long long r = R[corner & 3];
        if (!(corner & 4)) return Loc{r, k};
        return Loc{r, zghost ? G[corner & 3] : kz};
    }
};

template <int Qn>
LBM_HD inline long long row_pitch(const Geom& g) {
    return (long long)Qn * g.SZ;
}
template <int Qn>
LBM_HD inline long long plane_pitch(const Geom& g) {
    return (long long)(g.SY + g.GY) * row_pitch<Qn>(g);
}

template <int Qn>
LBM_DEV Cursor<Qn> make_cursor(const Geom& g, int i, int j, int k) {
    Cursor<Qn> c;
    const long long RP = row_pitch<Qn>(g), PP = plane_pitch<Qn>(g);
    const int i1 = (g.px && i + 1 == g.nx) ? 0 : i + 1;
    const int j1 = (g.py && j + 1 == g.ny) ? 0 : j + 1;
    c.R[0] = i * PP + j * RP;
    c.R[1] = i1 * PP + j * RP;
    c.R[2] = i * PP + j1 * RP;
    c.R[3] = i1 * PP + j1 * RP;
    c.k = uint32_t(k);
    c.kz = (g.pz && k + 1 == g.nz) ? 0u : uint32_t(k + 1);
    c.zghost = !g.pz && k + 1 == g.nz;
    const uint32_t sz = uint32_t(g.SZ), sy = uint32_t(g.SY), rp = uint32_t(RP);
    const uint32_t g0 = (sy + uint32_t(j) / sz - uint32_t(j)) * rp + uint32_t(j) % sz;
    const uint32_t g1 = (sy + uint32_t(j1) / sz - uint32_t(j1)) * rp + uint32_t(j1) % sz;
    c.G[0] = c.G[1] = g0;
    c.G[2] = c.G[3] = g1;
    return c;
}

// Element index of (array arr, location) from the field base.
template <int Q>
LBM_DEV long long field_index(const Geom& g, int arr, const typename Cursor<Q>::Loc& L) {
    return L.row + (long long)apos<Q>(arr) * g.SZ + L.lane;
}

// Directions q whose velocity component along `axis` has sign `sgn`, as a bit mask.
template <int Q>
constexpr uint32_t dir_bits(int axis, int sgn) {
    uint32_t b = 0;
    for (int q = 1; q < Q; q++)
        if (cvec<Q>(q, axis) * sgn > 0) b |= 1u << q;
    return b;
}

// Blocked-link mask of a cell from the domain faces alone (no solid cells): on the low
// face of an axis every link with a positive component there is blocked (its upstream
// neighbour is outside), on the high face every link with a negative one.  The x and y
// faces are uniform over a block's z row, so only the z terms are per lane.
template <int Q>
LBM_DEV uint32_t face_mask(const Geom& g, int i, int j, int k) {
    const bool lx = !g.px && i == 0, hx = !g.px && i == g.nx - 1;
    const bool ly = !g.py && j == 0, hy = !g.py && j == g.ny - 1;
    const bool lz = !g.pz && k == 0, hz = !g.pz && k == g.nz - 1;
    return (lx ? dir_bits<Q>(0, 1) : 0u) | (hx ? dir_bits<Q>(0, -1) : 0u) | (ly ? dir_bits<Q>(1, 1) : 0u) |
           (hy ? dir_bits<Q>(1, -1) : 0u) | (lz ? dir_bits<Q>(2, 1) : 0u) | (hz ? dir_bits<Q>(2, -1) : 0u);
}

template <int Q, bool MASK>
LBM_DEV uint32_t cell_mask(const Geom& g, const uint32_t* mask, long long cell, int i, int j, int k) {
    if constexpr (MASK)
        return mask[cell];
    else
        return face_mask<Q>(g, i, j, k);
}

// Read location (element index from the field base) of f_q at a cell for a state of
// parity odd.
template <int Q, int q>
LBM_DEV long long read_loc(const Geom& g, const Cursor<Q>& cu, uint32_t m, int odd) {
    const bool bl = (m >> q) & 1u;
    const int arr = ((odd != 0) != bl) ? q : opp<Q>(q);
    return field_index<Q>(g, arr, cu.at(corner_neg<Q>(q)));
}
// Write location of f*_q when stepping a state of parity odd.
template <int Q, int q>
LBM_DEV long long write_loc(const Geom& g, const Cursor<Q>& cu, int odd) {
    const int arr = odd ? opp<Q>(q) : q;
    return field_index<Q>(g, arr, cu.at(corner_pos<Q>(q)));
}

// Element `lane` (+ one array when hi) of the array pair starting at array `lo` of the
// row block at element `row`: a uniform base (row and pair) plus a 32-bit byte offset.
// OPAQUE: pass the 32-bit byte offset through an empty asm statement first.  A kernel
// that reads a field at these locations and writes another one there later (the adjoint
// reads f and writes lambda at the read locations) would otherwise reuse the read's
// offset across basic blocks as a 64-bit value and address the store with 64-bit VGPR
// pairs; the barrier keeps the 32-bit offset, and the store takes the uniform base from
// SGPRs with the offset in one VGPR.
template <int Q, class T, bool OPAQUE = false>
LBM_DEV T& pair_elem(T* pop, const Geom& g, long long row, int lo, uint32_t lane, bool hi) {
    using B = typename std::conditional<std::is_const<T>::value, const char, char>::type;
    B* base = reinterpret_cast<B*>(pop + (row + (long long)apos<Q>(lo) * g.SZ));
    uint32_t off = (lane + (hi ? uint32_t(g.SZ) : 0u)) * uint32_t(sizeof(T));
    if constexpr (OPAQUE) asm("" : "+v"(off));
    return *reinterpret_cast<T*>(base + size_t(off));
}

// Element at the read location of f_q (see read_loc): the slot at x + a_opp(q) in the
// wave-uniform array of an open link, or in the opposite array for a blocked link
// (halfway bounce-back).  The array choice is a select of the lane offset: the opposite
// arrays of a pair are adjacent in the row block.
template <int Q, int q, class T, bool OPAQUE = false>
LBM_DEV T& at_read(T* pop, const Geom& g, const Cursor<Q>& cu, uint32_t m, int odd) {
    const auto L = cu.at(corner_neg<Q>(q));
    if constexpr (q == 0) {
        return pair_elem<Q, T, OPAQUE>(pop, g, L.row, 0, L.lane, false);
    } else {
        // Array read: (odd xor blocked) ? q : opp(q); the pair's higher index sits one
        // array after the lower one.
        constexpr int lo = q < opp<Q>(q) ? q : opp<Q>(q);
        constexpr bool q_is_hi = q != lo;
        const bool pick_q = (odd != 0) != (((m >> q) & 1u) != 0);
        return pair_elem<Q, T, OPAQUE>(pop, g, L.row, lo, L.lane, pick_q == q_is_hi);
    }
}
template <int Q, int q, class T>
LBM_DEV T load_read(const T* pop, const Geom& g, const Cursor<Q>& cu, uint32_t m, int odd) {
    return at_read<Q, q>(pop, g, cu, m, odd);
}
template <int Q, int q, class T>
LBM_DEV void store_read(T* pop, const Geom& g, const Cursor<Q>& cu, uint32_t m, int odd, T v) {
    at_read<Q, q, T, true>(pop, g, cu, m, odd) = v;
}
// Element at the write location of f*_q (see write_loc): array odd ? opp(q) : q.
template <int Q, int q, class T>
LBM_DEV T& at_write(T* pop, const Geom& g, const Cursor<Q>& cu, int odd) {
    const auto L = cu.at(corner_pos<Q>(q));
    if constexpr (q == 0) {
        return pair_elem<Q, T>(pop, g, L.row, 0, L.lane, false);
    } else {
        constexpr int lo = q < opp<Q>(q) ? q : opp<Q>(q);
        constexpr bool q_is_hi = q != lo;
        return pair_elem<Q, T>(pop, g, L.row, lo, L.lane, (odd == 0) == q_is_hi);
    }
}

template <int P, int Q>
LBM_DEV void load_cell(const Geom& g, const typename Store<P>::T* pop, const Cursor<Q>& cu, uint32_t m, int odd,
                       typename Store<P>::C* f) {
    static_for<0, Q>([&](auto I) {
        constexpr int q = decltype(I)::value;
        f[q] = Store<P>::template dec<Q, q>(load_read<Q, q>(pop, g, cu, m, odd));
    });
}

template <int P, int Q>
LBM_DEV void store_cell_read(const Geom& g, typename Store<P>::T* pop, const Cursor<Q>& cu, uint32_t m, int odd,
                             const typename Store<P>::C* f) {
    static_for<0, Q>([&](auto I) {
        constexpr int q = decltype(I)::value;
        store_read<Q, q>(pop, g, cu, m, odd, Store<P>::template enc<Q, q>(f[q]));
    });
}

#include "thermal_cell.h"
#include "wale.h"
#include "psm.h"

// ---------------------------------------------------------------------------------
// Reductions.
// ---------------------------------------------------------------------------------
template <class T>
LBM_DEV T wave_sum(T v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    return v;
}

// Wave reductions with DPP row shifts and row broadcasts (an inclusive scan whose lane 63
// holds the result): row_shr 1, 2, 4, 8 within rows of 16 lanes, then row_bcast 15 and
// 31 across rows; a lane whose source is outside its row, and every lane of a row the
// broadcast leaves out, reads 0, the neutral value of both reductions below.
// Max of unsigned values: six VALU operations.
LBM_DEV unsigned wave_max_u32(unsigned x) {
    x = max(x, unsigned(__builtin_amdgcn_update_dpp(0, int(x), 0x111, 0xf, 0xf, false)));  // row_shr:1
    x = max(x, unsigned(__builtin_amdgcn_update_dpp(0, int(x), 0x112, 0xf, 0xf, false)));  // row_shr:2
    x = max(x, unsigned(__builtin_amdgcn_update_dpp(0, int(x), 0x114, 0xf, 0xf, false)));  // row_shr:4
    x = max(x, unsigned(__builtin_amdgcn_update_dpp(0, int(x), 0x118, 0xf, 0xf, false)));  // row_shr:8
    x = max(x, unsigned(__builtin_amdgcn_update_dpp(0, int(x), 0x142, 0xa, 0xf, false)));  // row_bcast:15
    x = max(x, unsigned(__builtin_amdgcn_update_dpp(0, int(x), 0x143, 0xc, 0xf, false)));  // row_bcast:31
    return unsigned(__builtin_amdgcn_readlane(int(x), 63));
}
// Max over the 64 lanes of non-negative floats (compared as their bit patterns).
LBM_DEV float wave_max_nonneg(float v) { return __uint_as_float(wave_max_u32(__float_as_uint(v))); }

// Order-preserving key of a float as an unsigned integer (0 sorts below every float),
// so that a running max of signed values can be held in a zero-initialised word.
LBM_DEV unsigned float_key(float v) {
    const unsigned u = __float_as_uint(v);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
LBM_DEV float key_float(unsigned x) { return __uint_as_float((x & 0x80000000u) ? (x & 0x7fffffffu) : ~x); }

// A DPP move of a double (its two halves); 0 where the source is outside the row or the
// row is masked off.
template <int CTRL, int ROWS>
LBM_DEV double dpp_f64(double v) {
    const unsigned long long b = static_cast<unsigned long long>(__double_as_longlong(v));
    const unsigned lo = unsigned(__builtin_amdgcn_update_dpp(0, int(unsigned(b)), CTRL, ROWS, 0xf, false));
    const unsigned hi = unsigned(__builtin_amdgcn_update_dpp(0, int(unsigned(b >> 32)), CTRL, ROWS, 0xf, false));
    return __longlong_as_double(static_cast<long long>((static_cast<unsigned long long>(hi) << 32) | lo));
}
// Sum over the 64 lanes in double precision, in a fixed order (the scan above).
LBM_DEV double wave_sum_f64(double v) {
    v += dpp_f64<0x111, 0xf>(v);
    v += dpp_f64<0x112, 0xf>(v);
    v += dpp_f64<0x114, 0xf>(v);
    v += dpp_f64<0x118, 0xf>(v);
    v += dpp_f64<0x142, 0xa>(v);
    v += dpp_f64<0x143, 0xc>(v);
    const unsigned long long b = static_cast<unsigned long long>(__double_as_longlong(v));
    const unsigned lo = unsigned(__builtin_amdgcn_readlane(int(unsigned(b)), 63));
    const unsigned hi = unsigned(__builtin_amdgcn_readlane(int(unsigned(b >> 32)), 63));
    return __longlong_as_double(static_cast<long long>((static_cast<unsigned long long>(hi) << 32) | lo));
}

LBM_DEV long long block_linear() {
    return (long long)blockIdx.x + (long long)gridDim.x * ((long long)blockIdx.y + (long long)gridDim.y * blockIdx.z);
}

// Per-block partials of the objective, the fluid mass, max |u|^2, the count of bad
// cells and the largest negated population (-min f over the block's fluid cells),
// PART_STRIDE doubles per block, written without atomics (a single address hit by every
// block serialises in L2); k_fold combines them, in a fixed order, after the step.
constexpr int PART_STRIDE = 8;

LBM_DEV void block_partials(double jv, double mv, float umax, bool bad, float fneg, double* part) {
    __shared__ double red[5][16];
    jv = wave_sum_f64(jv);
    mv = wave_sum_f64(mv);
    umax = wave_max_nonneg(umax);
    fneg = key_float(wave_max_u32(float_key(fneg)));
    const double nbad = double(__popcll(__ballot(bad)));
    const int tid = threadIdx.x + blockDim.x * threadIdx.y;  // blockDim.x is a multiple of 64
    const int lane = tid & 63, wid = tid >> 6;
    if (lane == 0) {
        red[0][wid] = jv;
        red[1][wid] = mv;
        red[2][wid] = double(umax);
        red[3][wid] = nbad;
        red[4][wid] = double(fneg);
    }
    __syncthreads();
    if (tid == 0) {
        const int nw = int(blockDim.x * blockDim.y + 63) >> 6;
        double a = 0.0, b = 0.0, c = 0.0, d = 0.0, e = -INFINITY;
        for (int w = 0; w < nw; w++) {
            a += red[0][w];
            b += red[1][w];
            c = fmax(c, red[2][w]);
            d += red[3][w];
            e = fmax(e, red[4][w]);
        }
        double* o = part + PART_STRIDE * block_linear();
        o[0] = a;
        o[1] = b;
        o[2] = c;
        o[3] = d;
        o[4] = e;
    }
}



// Block max of |lambda| (>= 0) raised into one of the LAM_SLOTS slots of the scale
// record (float bits order like the values; one atomic per block).
LBM_DEV void block_raise_max(float v, int* lsc) {
    __shared__ float red[16];
    v = wave_max_nonneg(v);
    const int tid = threadIdx.x + blockDim.x * threadIdx.y;
    const int lane = tid & 63, wid = tid >> 6;
    if (lane == 0) red[wid] = v;
    __syncthreads();
    if (tid == 0) {
        const int nw = int(blockDim.x * blockDim.y + 63) >> 6;
        float m = 0.0f;
        for (int w = 0; w < nw; w++) m = fmaxf(m, red[w]);
        atomicMax(reinterpret_cast<unsigned*>(lsc + LAM_SLOT0 + int(block_linear() % LAM_SLOTS)), __float_as_uint(m));
    }
}

// After an FP16 cotangent step (one wave of LAM_SLOTS lanes): the field now has exponent
// lsc[1]; record the max of the slots, count an overflow, choose the next exponent from
// LAM_GROWTH * max + s_next (s_next bounds the stage seed of the following step) and
// clear the slots.
template <int UNUSED = 0>
__global__ __launch_bounds__(64) void k_lam_update(int* lsc, double s_next) {
    const int t = threadIdx.x;
    const float m = wave_max_nonneg(__uint_as_float(unsigned(lsc[LAM_SLOT0 + t])));
    lsc[LAM_SLOT0 + t] = 0;
    if (t == 0) {
        const int ew = lsc[1];
        if (!(double(m) * ldexp(1.0, ew) <= 65504.0)) lsc[3] = lsc[3] + 1;
        lsc[0] = ew;
        lsc[2] = __float_as_int(m);
        lsc[1] = lam_exponent(LAM_GROWTH * double(m) + s_next);
    }
}

// Folds the per-block partials of one launch (part[PART_STRIDE b + 0..4] = objective,
// mass, max |u|^2, bad-cell count, max -f) in two fixed-order stages: each fold block
// reduces 256 partials per pass into stage[FOLD_VALUES block ..]; the last fold block to
// finish (a ticket counter, released with device-scope fences) loads the stage values
// one per thread, reduces them in the same fixed tree order, writes the results and
// resets the ticket.  Deterministic for a given launch shape.
//   FOLD_STEP:   jacc += objective; *macc = mass (if set); dstat[0] = max(dstat[0],
//                max |u|^2), dstat[1] += bad and dstat[2] = max(dstat[2], float_key(max -f))
//                (if set).
constexpr int FOLD_MAX_BLOCKS = 256;
constexpr int FOLD_VALUES = 5;
enum FoldMode { FOLD_STEP = 0 };

LBM_HD inline int fold_blocks(long long nb) {
    const long long g = (nb + 255) / 256;
    return int(g < 1 ? 1 : (g > FOLD_MAX_BLOCKS ? FOLD_MAX_BLOCKS : g));
}

// Block-wide fixed-order reduction of (sum a, sum b, max c, sum d, max e); result in
// thread 0.
LBM_DEV void block_fold5(double& a, double& b, double& c, double& d, double& e) {
    __shared__ double red[5][4];
    a = wave_sum(a);
    b = wave_sum(b);
    d = wave_sum(d);
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        c = fmax(c, __shfl_down(c, off, 64));
        e = fmax(e, __shfl_down(e, off, 64));
    }
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (lane == 0) {
        red[0][wid] = a;
        red[1][wid] = b;
        red[2][wid] = c;
        red[3][wid] = d;
        red[4][wid] = e;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        a = b = c = d = 0.0;
        e = -INFINITY;
        for (int w = 0; w < int(blockDim.x >> 6); w++) {
            a += red[0][w];
            b += red[1][w];
            c = fmax(c, red[2][w]);
            d += red[3][w];
            e = fmax(e, red[4][w]);
        }
    }
}

template <int MODE>
__global__ __launch_bounds__(256) void k_fold(const double* part, long long nb, double* stage, unsigned* ticket,
                                             double* jacc, double* macc, unsigned* dstat) {
    __shared__ bool last;
    double a = 0.0, b = 0.0, c = 0.0, d = 0.0, e = -INFINITY;
    const long long step = (long long)gridDim.x * blockDim.x;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < nb; i += step) {
        const double4 v = *reinterpret_cast<const double4*>(part + PART_STRIDE * i);
        a += v.x;
        b += v.y;
        c = fmax(c, v.z);
        d += v.w;
        e = fmax(e, part[PART_STRIDE * i + 4]);
    }
    block_fold5(a, b, c, d, e);
    if (threadIdx.x == 0) {
        double* so = stage + FOLD_VALUES * blockIdx.x;
        so[0] = a;
        so[1] = b;
        so[2] = c;
        so[3] = d;
        so[4] = e;
        __threadfence();
        last = atomicAdd(ticket, 1u) == gridDim.x - 1;
    }
    __syncthreads();
    if (!last) return;
    __threadfence();
    a = b = c = d = 0.0;
    e = -INFINITY;
    if (threadIdx.x < gridDim.x) {
        const double* sp = stage + FOLD_VALUES * threadIdx.x;
        a = __hip_atomic_load(sp, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        b = __hip_atomic_load(sp + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        c = __hip_atomic_load(sp + 2, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        d = __hip_atomic_load(sp + 3, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        e = __hip_atomic_load(sp + 4, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    block_fold5(a, b, c, d, e);
    if (threadIdx.x != 0) return;
    if (MODE == FOLD_STEP) {
        if (jacc) *jacc += a;
        if (macc) *macc = b;
        if (dstat) {
            const unsigned ub = __float_as_uint(float(c));
            if (ub > dstat[0]) dstat[0] = ub;
            dstat[1] += unsigned(d);
            if (e > -INFINITY) {
                const unsigned ke = float_key(float(e));
                if (ke > dstat[2]) dstat[2] = ke;
            }
        }
    }
    *ticket = 0u;
}

// ---------------------------------------------------------------------------------
// Objective per cell.
// ---------------------------------------------------------------------------------
template <bool SH, int Q, class C>
LBM_DEV double objective_cell(const ObjArgs& ob, const C* f, const Macro<Q, C>& m, C alpha, int coord_axis) {
    if (ob.kind == OBJ_DISSIPATION) {
        const C usq = m.u[0] * m.u[0] + m.u[1] * m.u[1] + m.u[2] * m.u[2];
        return ob.weight * double(m.rho * alpha * usq);
    }
    if (ob.kind == OBJ_VECTOR) {
        const double u0 = double(m.u[0]) * ob.vel, u1 = double(m.u[1]) * ob.vel, u2 = double(m.u[2]) * ob.vel;
        const double r = double(m.rho) * ob.dens;
        const double usq = u0 * u0 + u1 * u1 + u2 * u2;
        const double v = ob.vw[0] * u0 + ob.vw[1] * u1 + ob.vw[2] * u2 + ob.vw[3] * (r * usq) +
                         ob.vw[4] * (r * double(alpha) * ob.inv_dt * usq);
        return ob.weight * v;
    }
    if (ob.kind == OBJ_PRESSURE_DROP) {
        const double s = coord_axis == ob.plane_a ? ob.inv_na : (coord_axis == ob.plane_b ? -ob.inv_nb : 0.0);
        if (s == 0.0) return 0.0;
        // rho - 1 summed from the deviations avoids cancellation against 1.
        C dev = SH ? f[0] : f[0] - C(weight<Q>(0));
        static_for<1, Q>([&](auto I) {
            constexpr int q = decltype(I)::value;
            dev += SH ? f[q] : f[q] - C(weight<Q>(q));
        });
        return ob.weight * s * (double(dev) / 3.0);
    }
    return 0.0;
}

// Gradient of weight * phi(state) as cotangents of the density and momentum of the
// cell: d(weight phi)/d f_q = R + c_q . J (every objective depends on f through rho and
// j only), and d(weight phi)/d alpha added to ab.
template <int Q, class C>
LBM_DEV void objective_grad_rj(const ObjArgs& ob, const Macro<Q, C>& m, C alpha, int coord_axis, C& R, C J[3],
                               C& ab) {
    R = C(0);
    J[0] = J[1] = J[2] = C(0);
    if (ob.kind == OBJ_DISSIPATION) {
        const C w = C(ob.weight);
        const C usq = m.u[0] * m.u[0] + m.u[1] * m.u[1] + m.u[2] * m.u[2];
        C rho_bar = w * alpha * usq;
        const C usq_bar = w * m.rho * alpha;
        C al_bar = w * m.rho * usq;
#pragma unroll
        for (int d = 0; d < 3; d++) {
            const C ub = usq_bar * C(2.0) * m.u[d];
            al_bar -= qdiv(ub * m.sv[d], m.den * m.den) * C(0.5);
            const C sb = qdiv(ub, m.den);
            J[d] = qdiv(sb, m.rho);
            rho_bar -= qdiv(sb * m.j[d], m.rho * m.rho);
        }
        R = rho_bar;
        ab += al_bar;
    } else if (ob.kind == OBJ_VECTOR) {
        // phi = sum_d A_d u_d + B rho |u|^2 with A_d = w vw_d vel and
        // B = w dens vel^2 (vw3 + vw4 alpha / dt); then the Brinkman chain as above.
        const C w = C(ob.weight);
        const C V = C(ob.vel);
        const C bk = w * C(ob.dens) * V * V * C(ob.vw[3]);
        const C bd = w * C(ob.dens) * V * V * C(ob.vw[4]) * C(ob.inv_dt);
        const C B = bk + bd * alpha;
        const C usq = m.u[0] * m.u[0] + m.u[1] * m.u[1] + m.u[2] * m.u[2];
        C rho_bar = B * usq;
        C al_bar = bd * m.rho * usq;
#pragma unroll
        for (int d = 0; d < 3; d++) {
            const C ub = w * C(ob.vw[d]) * V + C(2) * B * m.rho * m.u[d];
            al_bar -= qdiv(ub * m.sv[d], m.den * m.den) * C(0.5);
            const C sb = qdiv(ub, m.den);
            J[d] = qdiv(sb, m.rho);
            rho_bar -= qdiv(sb * m.j[d], m.rho * m.rho);
        }
        R = rho_bar;
        ab += al_bar;
    } else if (ob.kind == OBJ_PRESSURE_DROP) {
        const double s = coord_axis == ob.plane_a ? ob.inv_na : (coord_axis == ob.plane_b ? -ob.inv_nb : 0.0);
        R = C(ob.weight * s / 3.0);
    }
}

// Adds weight * d phi / d f and d phi / d alpha to fb / ab.
template <int Q, class C>
LBM_DEV void objective_grad(const ObjArgs& ob, const Macro<Q, C>& m, C alpha, int coord_axis, C* fb, C& ab) {
    if (ob.kind != OBJ_DISSIPATION && ob.kind != OBJ_VECTOR && ob.kind != OBJ_PRESSURE_DROP) return;
    C R, J[3];
    objective_grad_rj<Q, C>(ob, m, alpha, coord_axis, R, J, ab);
    if (ob.kind == OBJ_PRESSURE_DROP) {
        if (R != C(0)) {
#pragma unroll
            for (int q = 0; q < Q; q++) fb[q] += R;
        }
        return;
    }
    static_for<0, Q>([&](auto I) {
        constexpr int i = decltype(I)::value;
        C v = R;
        if constexpr (cvec<Q>(i, 0) != 0) v += J[0] * C(cvec<Q>(i, 0));
        if constexpr (cvec<Q>(i, 1) != 0) v += J[1] * C(cvec<Q>(i, 1));
        if constexpr (cvec<Q>(i, 2) != 0) v += J[2] * C(cvec<Q>(i, 2));
        fb[i] += v;
    });
}

// ---------------------------------------------------------------------------------
// Collision dispatch (forward and transposed).
// ---------------------------------------------------------------------------------
template <int Q, class C, bool SH>
LBM_DEV void macro_of(const C* v, C alpha, const C a[3], Macro<Q, C>& m) {
    if constexpr (SH)
        brinkman_macro_shifted<Q>(v, alpha, a, m);
    else
        brinkman_macro<Q>(v, alpha, a, m);
}

// v holds f (double precision, reference operation order) or d = f - w (single
// precision, shifted formulation); out is returned in the same representation.
template <int Q, int KIND, class C, bool SH>
LBM_DEV void collide(const C* v, C alpha, const C a[3], C omega, C om, C tau, const C* mrt, C* out,
                     Macro<Q, C>& m) {
    if constexpr (KIND == K_STREAM_ONLY) {
#pragma unroll
        for (int q = 0; q < Q; q++) out[q] = v[q];
        m.rho = C(1);
    } else if constexpr (KIND == K_BGK_GUO) {
        static_assert(Q == 19, "BGK-Guo is the D3Q19 stationary operator");
        macro_of<Q, C, SH>(v, alpha, a, m);
        if constexpr (SH)
            collide_bgk_guo_shifted(v, m, tau, out);
        else
            collide_bgk_guo(v, alpha, a, tau, out);
    } else {
        macro_of<Q, C, SH>(v, alpha, a, m);
        if constexpr (KIND == K_TRT_LES || KIND == K_BGK_LES) {
            LesEval<Q, C> e;
            les_eval<Q, C, SH>(v, m, omega, mrt, e);
            const C oe = e.oe;
            const C oo = KIND == K_TRT_LES ? odd_rate_of(oe, mrt[0]) : oe;
            if constexpr (SH)
                collide_pairs_shifted<Q>(v, m, oe, oo, out);
            else
                collide_pairs<Q>(v, m.rho, m.u, m.F, oe, oo, out);
        } else if constexpr (is_wale(KIND)) {  // omega is the cell's WALE rate
            const C oo = KIND == K_TRT_WALE ? odd_rate_of(omega, mrt[0]) : omega;
            if constexpr (SH)
                collide_pairs_shifted<Q>(v, m, omega, oo, out);
            else
                collide_pairs<Q>(v, m.rho, m.u, m.F, omega, oo, out);
        } else if constexpr (KIND == K_TRT || KIND == K_BGK) {
            if constexpr (SH)
                collide_pairs_shifted<Q>(v, m, omega, KIND == K_BGK ? omega : om, out);
            else
                collide_pairs<Q>(v, m.rho, m.u, m.F, omega, KIND == K_BGK ? omega : om, out);
        } else if constexpr (KIND == K_MRT) {
            if constexpr (SH)
                collide_mrt_shifted<Q>(v, m, mrt, out);
            else
                collide_mrt<Q>(v, m.rho, m.u, m.F, mrt, out);
        } else {
            if constexpr (SH) {
                collide_cumulant_shifted(v, m.drho, m.rho, m.u, m.F, omega, mrt[0], out);
            } else {
                C f[Q], o[Q];
                to_full<SH, Q>(v, f);
                collide_cumulant(f, m.rho, m.u, m.F, omega, mrt[0], o);
                from_full<SH, Q>(o, out);
            }
        }
    }
}

// WALE kinds: omega is the cell's rate and *rate_bar receives its cotangent.
template <int Q, int KIND, class C, bool SH>
LBM_DEV void collide_vjp(const C* v, C alpha, const C a[3], C omega, C om, C tau, const C* mrt, const C* g,
                         Macro<Q, C>& m, C* fb, C& ab, C* rate_bar = nullptr) {
    macro_of<Q, C, SH>(v, alpha, a, m);
    if constexpr (KIND == K_BGK_GUO) {
        collide_bgk_guo_vjp(m, alpha, a, tau, g, fb, ab);
    } else {
        C cb_rho, cb_u[3], cb_F[3];
        if constexpr (KIND == K_TRT_LES || KIND == K_BGK_LES) {
            LesEval<Q, C> e;
            les_eval<Q, C, SH>(v, m, omega, mrt, e);
            const C oe = e.oe;
            const C oo = KIND == K_TRT_LES ? odd_rate_of(oe, mrt[0]) : oe;
            const C dom = KIND == K_TRT_LES ? odd_rate_derivative_of(oe, mrt[0]) : C(1);
            // the rate cotangents first: the non-equilibrium part is not needed after them
            C ob, omb;
            pairs_omega_bar<Q, C>(e.n, m, g, ob, omb);
            if constexpr (std::is_same<C, float>::value)
                collide_pairs_vjp_fast<Q>(m.rho, m.u, m.F, oe, oo, g, fb, cb_rho, cb_u, cb_F);
            else
                collide_pairs_vjp<Q>(m.rho, m.u, m.F, oe, oo, g, fb, cb_rho, cb_u, cb_F);
            smagorinsky_vjp<Q, C>(e, m, mrt, ob + omb * dom, fb, cb_rho, cb_u);
        } else if constexpr (is_wale(KIND)) {
            const C oo = KIND == K_TRT_WALE ? odd_rate_of(omega, mrt[0]) : omega;
            const C dom = KIND == K_TRT_WALE ? odd_rate_derivative_of(omega, mrt[0]) : C(1);
            C n[Q], ob, omb;
            nonequilibrium<Q, C, SH>(v, m, n);
            pairs_omega_bar<Q, C>(n, m, g, ob, omb);
            *rate_bar = ob + omb * dom;
            if constexpr (std::is_same<C, float>::value)
                collide_pairs_vjp_fast<Q>(m.rho, m.u, m.F, omega, oo, g, fb, cb_rho, cb_u, cb_F);
            else
                collide_pairs_vjp<Q>(m.rho, m.u, m.F, omega, oo, g, fb, cb_rho, cb_u, cb_F);
        } else if constexpr (KIND == K_TRT || KIND == K_BGK) {
            if constexpr (SH && std::is_same<C, float>::value) {  // the moment-form transpose, chain included
                const C zero3[3] = {C(0), C(0), C(0)};
                ab = trt_adjoint_shifted<Q>(m, alpha, a, omega, KIND == K_BGK ? omega : om, g, C(0), zero3, fb);
                return;
            }
            if constexpr (std::is_same<C, float>::value)
                collide_pairs_vjp_fast<Q>(m.rho, m.u, m.F, omega, KIND == K_BGK ? omega : om, g, fb, cb_rho, cb_u,
                                          cb_F);
            else
                collide_pairs_vjp<Q>(m.rho, m.u, m.F, omega, KIND == K_BGK ? omega : om, g, fb, cb_rho, cb_u, cb_F);
        } else if constexpr (KIND == K_MRT) {
            collide_mrt_vjp<Q>(m.rho, m.u, m.F, mrt, g, fb, cb_rho, cb_u, cb_F);
        } else {
            C f[Q];
            to_full<SH, Q>(v, f);
            collide_cumulant_vjp(f, m.rho, m.u, m.F, omega, mrt[0], g, fb, cb_rho, cb_u, cb_F);
        }
        brinkman_chain<Q>(m, alpha, a, cb_rho, cb_u, cb_F, fb, ab);
    }
}

// ---------------------------------------------------------------------------------
// Forward step: one thread per cell.
// ---------------------------------------------------------------------------------
// TH: also advance the coupled temperature field of the cell (thermal_cell.h) with the
// Brinkman velocity the flow collision computes.
template <int P, int Q, int KIND, bool MASK, bool TH = false>
__global__ __launch_bounds__(256) void k_forward(StepArgs A) {
    using S = Store<P>;
    using C = typename S::C;
    const Geom& g = A.g;
    // step launches: a block is a run of blockDim.x cells of one z row (blockDim.y = 1),
    // so the row (i, j) is uniform over the block
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y, i = blockIdx.z;
    const bool active = k < g.nz && j < g.ny;
    double jsum = 0.0, msum = 0.0;
    float umax = 0.0f, fneg = -INFINITY;
    unsigned bad = 0;
    if (active) {
        const long long cell = ((long long)i * g.ny + j) * g.nz + k;
        const uint32_t m = cell_mask<Q, MASK>(g, A.mask, cell, i, j, k);
        if (!(m & SOLID_BIT)) {
            const Cursor<Q> cu = make_cursor<Q>(g, i, j, k);
            typename S::T* pop = static_cast<typename S::T*>(A.pop);
            C f[Q], fs[Q];
            load_cell<P, Q>(g, pop, cu, m, A.odd, f);
            const C alpha = static_cast<const C*>(A.alpha)[cell];
            const C a[3] = {C(A.a[0]), C(A.a[1]), C(A.a[2])};
            Macro<Q, C> mac;
            C rate = C(A.omega);
            if constexpr (is_wale(KIND)) {
                C G[3][3];
                wale_gradient<C, MASK>(g, A.mask, static_cast<const C*>(A.wu), (long long)g.nx * g.ny * g.nz, i, j,
                                       k, cell, G);
                rate = wale_omega<C>(G, C(1) / C(A.omega), static_cast<const C*>(A.mrt));
            }
            if constexpr (is_bro(KIND)) {  // the design is the occupancy d
                const C* pk = static_cast<const C*>(A.mrt);
                C dc, ms[3];
                psm_input<C>(alpha, A.sm, A.d1, A.sm1, A.theta, (long long)g.nx * g.ny * g.nz, cell, dc, ms);
                if (dc > C(0)) {
                    C al, ga, part[4], z[3];
                    psm_resistance<C>(dc, pk[3], pk[4], pk[1], al, ga, part);
                    brinkman_moving_macro<Q, C, S::shifted>(f, al, ga, a, ms, mac, z);
                    const C om = psm_base(KIND) == K_TRT ? C(A.om) : rate;
                    if constexpr (S::shifted)
                        collide_pairs_shifted<Q>(f, mac, rate, om, fs);
                    else
                        collide_pairs<Q>(f, mac.rho, mac.u, mac.F, rate, om, fs);
                    if (A.obj.kind == OBJ_SOLID_FORCE)  // dp = rho z
                        jsum += A.obj.weight * (A.obj.vw[0] * double(mac.rho * z[0]) +
                                                A.obj.vw[1] * double(mac.rho * z[1]) +
                                                A.obj.vw[2] * double(mac.rho * z[2]));
                    if (A.gacc) {
                        const C dp[3] = {mac.rho * z[0], mac.rho * z[1], mac.rho * z[2]};
                        sub_accumulate<C>(A.gacc, (long long)g.nx * g.ny * g.nz, cell, dp, dc, A.theta);
                    }
                } else {
                    collide<Q, psm_base(KIND), C, S::shifted>(f, C(0), a, rate, C(A.om), C(A.tau), pk, fs, mac);
                }
            } else if constexpr (is_psm(KIND)) {  // the design is the occupancy d
                const C* pk = static_cast<const C*>(A.mrt);
                collide<Q, psm_base(KIND), C, S::shifted>(f, C(0), a, rate, C(A.om), C(A.tau), pk, fs, mac);
                C dc, ms[3];
                psm_input<C>(alpha, A.sm, A.d1, A.sm1, A.theta, (long long)g.nx * g.ny * g.nz, cell, dc, ms);
                if (dc > C(0)) {
                    C dp[3];
                    psm_blend<Q, C, S::shifted, psm_sup(KIND)>(f, mac, dc, ms, pk, C(A.omega), fs, dp);
                    if (A.obj.kind == OBJ_SOLID_FORCE)
                        jsum += A.obj.weight * (A.obj.vw[0] * double(dp[0]) + A.obj.vw[1] * double(dp[1]) +
                                                A.obj.vw[2] * double(dp[2]));
                    if (A.gacc) sub_accumulate<C>(A.gacc, (long long)g.nx * g.ny * g.nz, cell, dp, dc, A.theta);
                }
            } else {
                collide<Q, KIND, C, S::shifted>(f, alpha, a, rate, C(A.om), C(A.tau), static_cast<const C*>(A.mrt),
                                                fs, mac);
            }
            if (A.obj.kind != OBJ_NONE) {  // added to the solid-force share the coupling formed
                const int coord = A.obj.axis == 0 ? i : (A.obj.axis == 1 ? j : k);
                jsum += objective_cell<S::shifted, Q, C>(A.obj, f, mac, alpha, coord);
            }
            if (A.diag) {
                msum = double(mac.rho);
                umax = float(mac.u[0] * mac.u[0] + mac.u[1] * mac.u[1] + mac.u[2] * mac.u[2]);
                if constexpr (std::is_same<C, float>::value) {
                    // a population is bad when it is not positive or not finite: its bits,
                    // less one, as unsigned, reach 0x7f7fffff exactly then (zero, sign bit,
                    // infinity, NaN); one unsigned max and one min per population
                    float mn = INFINITY;
                    unsigned key = 0u;
#pragma unroll
                    for (int q = 0; q < Q; q++) {
                        const float fq = S::shifted ? f[q] + float(weight<Q>(q)) : f[q];
                        mn = fminf(mn, fq);
                        key = max(key, __float_as_uint(fq) - 1u);
                    }
                    bad = key >= 0x7f7fffffu;
                    fneg = -mn;
                } else {
#pragma unroll
                    for (int q = 0; q < Q; q++) {
                        const C fq = S::shifted ? f[q] + C(weight<Q>(q)) : f[q];
                        bad |= !(fq > C(0)) || !isfinite(float(fq));
                        fneg = fmaxf(fneg, -float(fq));
                    }
                }
            }
            if constexpr (TH) {
                using TS = TStore<P>;
                const uint32_t mt = face_mask<TQ>(g, i, j, k);
                const Cursor<TQ> ct = make_cursor<TQ>(g, i, j, k);  // the temperature field's rows
                C gv[TQ], go[TQ], T;
                load_tcell<P>(g, static_cast<const typename TS::T*>(A.tpop), ct, mt, A.odd, A.tdec, gv);
                const C phi = static_cast<const C*>(A.phi)[cell];
                thermal_collide<C>(gv, mac.u, thermal_omega<C>(A.kappa, A.ratio_m1, phi), C(A.source) * phi, go, T);
                if (A.tw != 0.0 && phi > C(0.5)) jsum += A.tw * double(T);
                typename TS::T* to = static_cast<typename TS::T*>(A.tpop_out);
                const uint32_t key = (uint32_t(cell) * uint32_t(TQ)) ^ (A.step * 0x9e3779b9u);
                static_for<0, TQ>([&](auto I) {
                    constexpr int q = decltype(I)::value;
                    at_write<TQ, q>(to, g, ct, A.odd) = tstore_step<P>(go[q], A.tenc, key + uint32_t(q));
                });
            }
            typename S::T* out = static_cast<typename S::T*>(A.pop_out);
            static_for<0, Q>([&](auto I) {
                constexpr int q = decltype(I)::value;
                at_write<Q, q>(out, g, cu, A.odd) = S::template enc<Q, q>(fs[q]);
            });
        }
    }
    if (A.part) block_partials(jsum, msum, umax, bad != 0, fneg, A.part);
}

// ---------------------------------------------------------------------------------
// Adjoint step, in place on the cotangent buffer.
// ---------------------------------------------------------------------------------
// u = (j / rho + a / 2) / (1 + alpha / 2): adds (du/df)^T ub to fb and (du/dalpha) . ub
// to ab, with d u_d / d f_q = (c_qd - j_d / rho) / (rho den) and d u / d alpha = -u / (2 den).
template <int Q, class C>
LBM_DEV void brinkman_u_vjp(const Macro<Q, C>& m, const C ub[3], C* fb, C& ab) {
    C v[3], s[3];
#pragma unroll
    for (int d = 0; d < 3; d++) {
        v[d] = m.j[d] / m.rho;
        s[d] = ub[d] / m.den / m.rho;
    }
    static_for<0, Q>([&](auto I) {
        constexpr int q = decltype(I)::value;
        C acc = C(0);
#pragma unroll
        for (int d = 0; d < 3; d++) acc += s[d] * (C(cvec<Q>(q, d)) - v[d]);
        fb[q] += acc;
    });
    C au = C(0);
#pragma unroll
    for (int d = 0; d < 3; d++) au += ub[d] * m.u[d];
    ab += C(-0.5) * au / m.den;
}

// UB: add the velocity cotangent A.ubar of the coupled temperature field.
template <int P, int LP, int Q, int KIND, bool MASK, bool UB = false>
__global__ __launch_bounds__(256) void k_adjoint(AdjArgs A) {
    using S = Store<P>;
    using L = LamStore<LP>;
    using C = typename S::C;
    // TRT / BGK in single precision: the moment-form transpose (trt_adjoint_shifted)
    constexpr bool FUSED = S::shifted && (KIND == K_TRT || KIND == K_BGK);
    const Geom& g = A.g;
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y, i = blockIdx.z;  // one z row per block, as k_forward
    float lmax = 0.0f;  // max |lambda| written (FP16 cotangents)
    if (k < g.nz && j < g.ny) {
        const long long cell = ((long long)i * g.ny + j) * g.nz + k;
        const uint32_t m = cell_mask<Q, MASK>(g, A.mask, cell, i, j, k);
        if (!(m & SOLID_BIT)) {
            const LScale sc = lam_scale<LP>(A.lsc, true);
            const Cursor<Q> cu = make_cursor<Q>(g, i, j, k);
            const typename S::T* pop = static_cast<const typename S::T*>(A.pop);
            typename L::T* lam = static_cast<typename L::T*>(A.lam);
            C f[Q], gb[Q], fb[Q];
            load_cell<P, Q>(g, pop, cu, m, A.odd, f);
            static_for<0, Q>([&](auto I) {
                constexpr int q = decltype(I)::value;
                gb[q] = L::template ld<C>(&at_write<Q, q>(lam, g, cu, A.odd), sc);
            });
            const C alpha = static_cast<const C*>(A.alpha)[cell];
            const C a[3] = {C(A.a[0]), C(A.a[1]), C(A.a[2])};
            Macro<Q, C> mac;
            C ab;
            if constexpr (is_wale(KIND)) {
                const long long n = (long long)g.nx * g.ny * g.nz;
                const C* wk = static_cast<const C*>(A.mrt);
                const C tau0 = C(1) / C(A.omega);
                C G[3][3], Gb[3][3], rb;
                wale_gradient<C, MASK>(g, A.mask, static_cast<const C*>(A.wu), n, i, j, k, cell, G);
                const C rate = wale_omega<C>(G, tau0, wk);
                collide_vjp<Q, KIND, C, S::shifted>(f, alpha, a, rate, C(A.om), C(A.tau), wk, gb, mac, fb, ab, &rb);
                wale_omega_vjp<C>(G, tau0, wk, rb, Gb);
                C* W = static_cast<C*>(A.wgb);
#pragma unroll
                for (int t = 0; t < 9; t++) W[t * n + cell] = Gb[t / 3][t % 3];
            } else if constexpr (is_bro(KIND)) {  // the design is the occupancy d; ab := d_bar
                constexpr int BK = psm_base(KIND);
                const C* pk = static_cast<const C*>(A.mrt);
                const long long n = (long long)g.nx * g.ny * g.nz;
                C dc, ms[3];
                psm_input<C>(alpha, A.sm, A.d1, A.sm1, A.theta, n, cell, dc, ms);
                if (dc > C(0)) {
                    C al, ga, part[4], z[3], cb_rho, cb_u[3], cb_F[3];
                    psm_resistance<C>(dc, pk[3], pk[4], pk[1], al, ga, part);
                    brinkman_moving_macro<Q, C, S::shifted>(f, al, ga, a, ms, mac, z);
                    const C om = BK == K_TRT ? C(A.om) : C(A.omega);
                    if constexpr (std::is_same<C, float>::value)
                        collide_pairs_vjp_fast<Q>(mac.rho, mac.u, mac.F, C(A.omega), om, gb, fb, cb_rho, cb_u, cb_F);
                    else
                        collide_pairs_vjp<Q>(mac.rho, mac.u, mac.F, C(A.omega), om, gb, fb, cb_rho, cb_u, cb_F);
                    C abr, gbr, mb[3], extra = C(0);
                    C dpb[3] = {psm_dp_bar<C>(A.obj, 0), psm_dp_bar<C>(A.obj, 1), psm_dp_bar<C>(A.obj, 2)};
                    if constexpr (Q == 19) {  // lbm_subcycle_adjoint is D3Q19 only
                        if (A.gbar) {
                            const C dp[3] = {mac.rho * z[0], mac.rho * z[1], mac.rho * z[2]};
                            sub_cotangent<C>(A.gbar, n, cell, dp, dc, A.theta, dpb, extra);
                        }
                    }
                    brinkman_moving_chain<Q, C>(mac, al, ga, a, ms, z, dpb, cb_rho, cb_u, cb_F, fb, abr, gbr, mb);
                    C db = abr * part[0] + gbr * part[2];
                    if constexpr (Q == 19) db += extra;
                    ab = psm_accumulate<C>(A, n, cell, db, mb);
                } else {
                    collide_vjp<Q, BK, C, S::shifted>(f, C(0), a, C(A.omega), C(A.om), C(A.tau), pk, gb, mac, fb, ab);
                    ab = C(0);
                }
            } else if constexpr (is_psm(KIND)) {  // the design is the occupancy d; ab := d_bar
                constexpr int BK = psm_base(KIND);
                const C* pk = static_cast<const C*>(A.mrt);
                const long long n = (long long)g.nx * g.ny * g.nz;
                C dc, ms[3];
                psm_input<C>(alpha, A.sm, A.d1, A.sm1, A.theta, n, cell, dc, ms);
                if (dc > C(0)) {
                    constexpr bool SUP = psm_sup(KIND);
                    C gfl = C(0);  // <g, fluid collision> (Psm law)
                    if constexpr (!SUP) {
                        C fl[Q];
                        collide<Q, BK, C, S::shifted>(f, C(0), a, C(A.omega), C(A.om), C(A.tau), pk, fl, mac);
#pragma unroll
                        for (int q = 0; q < Q; q++) gfl += gb[q] * fl[q];
                    }
                    collide_vjp<Q, BK, C, S::shifted>(f, C(0), a, C(A.omega), C(A.om), C(A.tau), pk, gb, mac, fb, ab);
                    C db, mb[3], extra = C(0);
                    C dpb[3] = {psm_dp_bar<C>(A.obj, 0), psm_dp_bar<C>(A.obj, 1), psm_dp_bar<C>(A.obj, 2)};
                    if constexpr (Q == 19) {  // lbm_subcycle_adjoint is D3Q19 only
                        if (A.gbar) {
                            C dp[3];
                            psm_exchange<Q, C, S::shifted, SUP>(f, mac, dc, ms, pk, C(A.omega), dp);
                            sub_cotangent<C>(A.gbar, n, cell, dp, dc, A.theta, dpb, extra);
                        }
                    }
                    psm_blend_vjp<Q, C, S::shifted, SUP>(f, mac, dc, ms, pk, C(A.omega), gb, gfl, dpb, fb, db, mb);
                    if constexpr (Q == 19) db += extra;
                    ab = psm_accumulate<C>(A, n, cell, db, mb);
                } else {
                    collide_vjp<Q, BK, C, S::shifted>(f, C(0), a, C(A.omega), C(A.om), C(A.tau), pk, gb, mac, fb, ab);
                    ab = C(0);
                }
            } else if constexpr (FUSED) {
                // objective and temperature-coupling terms enter as density and momentum
                // cotangents, added in the transpose's single per-direction pass
                macro_of<Q, C, true>(f, alpha, a, mac);
                C R0 = C(0), J0[3] = {C(0), C(0), C(0)};
                ab = C(0);
                if (A.obj.kind != OBJ_NONE) {
                    const int coord = A.obj.axis == 0 ? i : (A.obj.axis == 1 ? j : k);
                    objective_grad_rj<Q, C>(A.obj, mac, alpha, coord, R0, J0, ab);
                }
                if constexpr (UB) {
                    const C* U = static_cast<const C*>(A.ubar);
                    const long long n = (long long)g.nx * g.ny * g.nz;
                    const C ub[3] = {U[cell], U[n + cell], U[2 * n + cell]};
                    C au = C(0);
#pragma unroll
                    for (int d = 0; d < 3; d++) {
                        const C s = ub[d] / mac.den / mac.rho;
                        R0 -= s * (mac.j[d] / mac.rho);
                        J0[d] += s;
                        au += ub[d] * mac.u[d];
                    }
                    ab += C(-0.5) * au / mac.den;
                }
                ab += trt_adjoint_shifted<Q>(mac, alpha, a, C(A.omega), KIND == K_BGK ? C(A.omega) : C(A.om), gb, R0,
                                             J0, fb);
            } else {
                collide_vjp<Q, KIND, C, S::shifted>(f, alpha, a, C(A.omega), C(A.om), C(A.tau),
                                                    static_cast<const C*>(A.mrt), gb, mac, fb, ab);
            }
            if constexpr (!FUSED) {
                if (A.obj.kind != OBJ_NONE) {
                    const int coord = A.obj.axis == 0 ? i : (A.obj.axis == 1 ? j : k);
                    objective_grad<Q, C>(A.obj, mac, alpha, coord, fb, ab);
                }
                if constexpr (UB) {
                    const C* U = static_cast<const C*>(A.ubar);
                    const long long n = (long long)g.nx * g.ny * g.nz;
                    const C ub[3] = {U[cell], U[n + cell], U[2 * n + cell]};
                    brinkman_u_vjp<Q, C>(mac, ub, fb, ab);
                }
            }
            if (!A.abar32)
                A.alpha_bar[cell] += double(ab);
            else if (A.abar_mode == ABAR_FIRST)
                A.abar32[cell] = float(ab);
            else if (A.abar_mode == ABAR_MID)
                A.abar32[cell] += float(ab);
            else
                A.alpha_bar[cell] += double(A.abar32[cell]) + double(ab);
            static_for<0, Q>([&](auto I) {
                constexpr int q = decltype(I)::value;
                store_read<Q, q>(lam, g, cu, m, A.odd, L::template enc<C>(fb[q], sc));
            });
            if constexpr (LP == P_F16S) {
                // The first value to leave the FP16 range shows as a large or infinite max
                // (a NaN can only follow an earlier overflow or a non-finite state).
                float m0 = fabsf(float(fb[0]));
#pragma unroll
                for (int q = 1; q + 1 < Q; q += 2) m0 = fmaxf(m0, fmaxf(fabsf(float(fb[q])), fabsf(float(fb[q + 1]))));
                lmax = m0;
            }
        }
    }
    if constexpr (LP == P_F16S) block_raise_max(lmax, A.lsc);
}


// ---------------------------------------------------------------------------------
// WALE pre-pass and transpose (wale.h).
// ---------------------------------------------------------------------------------
// U[a * n + cell] = u_a = j_a / rho and U[3n + cell] = rho of the state of parity odd
// (u = 0, rho = 1 at solid cells); field.rs:1298-1308.
template <int P, int Q, bool MASK>
__global__ __launch_bounds__(256) void k_wale_u(Geom g, const void* pop_, const uint32_t* mask, int odd, void* U_) {
    using S = Store<P>;
    using C = typename S::C;
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y * blockDim.y + threadIdx.y, i = blockIdx.z;
    if (k >= g.nz || j >= g.ny) return;
    const long long n = (long long)g.nx * g.ny * g.nz;
    const long long cell = ((long long)i * g.ny + j) * g.nz + k;
    C* U = static_cast<C*>(U_);
    const uint32_t m = cell_mask<Q, MASK>(g, mask, cell, i, j, k);
    if (m & SOLID_BIT) {
        U[cell] = U[n + cell] = U[2 * n + cell] = C(0);
        U[3 * n + cell] = C(1);
        return;
    }
    const Cursor<Q> cu = make_cursor<Q>(g, i, j, k);
    C f[Q], rho, jv[3];
    load_cell<P, Q>(g, static_cast<const typename S::T*>(pop_), cu, m, odd, f);
    moments<Q>(f, rho, jv);
    if constexpr (S::shifted) rho = C(1) + rho;
#pragma unroll
    for (int d = 0; d < 3; d++) U[d * n + cell] = jv[d] / rho;
    U[3 * n + cell] = rho;
}

// Transpose of the gradients (Stencil::gradient_transpose): every fluid cell y gathers
// the gradient cotangents GB of the cells whose stencil used u_y (itself and its
// neighbours along each axis) into u_bar, then adds (du/df)^T u_bar,
// d u_a / d f_q = (c_qa - u_a) / rho, to its cotangent at its read locations (state of
// parity odd, after the adjoint kernel wrote it).
template <int P, int LP, int Q, bool MASK>
__global__ __launch_bounds__(256) void k_wale_transpose(Geom g, const uint32_t* mask, const void* U_, const void* GB_,
                                                        void* lam_, int odd) {
    using C = typename Store<P>::C;
    using L = LamStore<LP>;
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y * blockDim.y + threadIdx.y, i = blockIdx.z;
    if (k >= g.nz || j >= g.ny) return;
    const long long n = (long long)g.nx * g.ny * g.nz;
    const long long y = ((long long)i * g.ny + j) * g.nz + k;
    const uint32_t m = cell_mask<Q, MASK>(g, mask, y, i, j, k);
    if (m & SOLID_BIT) return;
    const C* U = static_cast<const C*>(U_);
    const C* GB = static_cast<const C*>(GB_);
    C ub[3] = {C(0), C(0), C(0)};
    auto add = [&](long long x, int b, C coef) {
#pragma unroll
        for (int a = 0; a < 3; a++) ub[a] += GB[(3 * a + b) * n + x] * coef;
    };
    const int ijk[3] = {i, j, k};
#pragma unroll
    for (int b = 0; b < 3; b++) {
        // candidates: y itself and its neighbours along b (each cell once)
        long long cand[3];
        int nc = 0;
        cand[nc++] = y;
        const long long xp = wale_nb<MASK>(g, mask, i, j, k, b, 1);
        const long long xm = wale_nb<MASK>(g, mask, i, j, k, b, -1);
        if (xp >= 0 && xp != y) cand[nc++] = xp;
        if (xm >= 0 && xm != y && xm != xp) cand[nc++] = xm;
        for (int t = 0; t < nc; t++) {
            const long long x = cand[t];
            int c3[3] = {ijk[0], ijk[1], ijk[2]};
            if (x != y) {  // coordinates of the neighbour
                const int nb3[3] = {g.nx, g.ny, g.nz};
                c3[b] += (x == xp) ? 1 : -1;
                if (c3[b] < 0) c3[b] += nb3[b];
                if (c3[b] >= nb3[b]) c3[b] -= nb3[b];
            }
            long long pp, mm;
            const int kind = wale_diff<MASK>(g, mask, c3[0], c3[1], c3[2], b, pp, mm);
            if (kind == 1) {
                if (pp == y) add(x, b, C(0.5));
                if (mm == y) add(x, b, C(-0.5));
            } else if (kind == 2) {
                if (pp == y) add(x, b, C(1));
                if (x == y) add(x, b, C(-1));
            } else if (kind == 3) {
                if (x == y) add(x, b, C(1));
                if (mm == y) add(x, b, C(-1));
            }
        }
    }
    if (ub[0] == C(0) && ub[1] == C(0) && ub[2] == C(0)) return;
    const C rho = U[3 * n + y];
    const C u[3] = {U[y], U[n + y], U[2 * n + y]};
    const Cursor<Q> cu = make_cursor<Q>(g, i, j, k);
    typename L::T* lam = static_cast<typename L::T*>(lam_);
    static_for<0, Q>([&](auto I) {
        constexpr int q = decltype(I)::value;
        C acc = C(0);
#pragma unroll
        for (int a = 0; a < 3; a++) acc += ub[a] * (C(cvec<Q>(q, a)) - u[a]);
        typename L::T& slot = at_read<Q, q>(lam, g, cu, m, odd);
        slot = L::template enc<C>(L::template ld<C>(&slot) + acc / rho);
    });
}

// ---------------------------------------------------------------------------------
// Objective on one state: accumulate its value (eval) and/or write the cotangent seed
// lambda = d J / d state at every fluid cell and alpha_bar = d J / d alpha (seed).
// ---------------------------------------------------------------------------------
template <int P, int LP, int Q, bool MASK>
__global__ __launch_bounds__(256) void k_objective(StepArgs A, ObjArgs seed, void* lam, double* alpha_bar,
                                                   int do_seed) {
    using S = Store<P>;
    using L = LamStore<LP>;
    using C = typename S::C;
    const Geom& g = A.g;
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y, i = blockIdx.z;  // one z row per block, as k_forward
    double jsum = 0.0, msum = 0.0;
    float umax = 0.0f, fneg = -INFINITY;
    bool bad = false;
    if (k < g.nz && j < g.ny) {
        const long long cell = ((long long)i * g.ny + j) * g.nz + k;
        const uint32_t m = cell_mask<Q, MASK>(g, A.mask, cell, i, j, k);
        if (!(m & SOLID_BIT)) {
            const Cursor<Q> cu = make_cursor<Q>(g, i, j, k);
            C f[Q];
            load_cell<P, Q>(g, static_cast<const typename S::T*>(A.pop), cu, m, A.odd, f);
            const C alpha = static_cast<const C*>(A.alpha)[cell];
            const C a[3] = {C(A.a[0]), C(A.a[1]), C(A.a[2])};
            Macro<Q, C> mac;
            macro_of<Q, C, S::shifted>(f, alpha, a, mac);
            if (A.obj.kind != OBJ_NONE) {
                const int coord = A.obj.axis == 0 ? i : (A.obj.axis == 1 ? j : k);
                jsum = objective_cell<S::shifted, Q, C>(A.obj, f, mac, alpha, coord);
            }
            msum = double(mac.rho);
            if (A.diag) {  // the admission quantities of this state (as the step kernel's)
                umax = float(mac.u[0] * mac.u[0] + mac.u[1] * mac.u[1] + mac.u[2] * mac.u[2]);
#pragma unroll
                for (int q = 0; q < Q; q++) {
                    const C fq = S::shifted ? f[q] + C(weight<Q>(q)) : f[q];
                    bad |= !(fq > C(0)) || !isfinite(float(fq));
                    fneg = fmaxf(fneg, -float(fq));
                }
            }
            if (do_seed) {
                C fb[Q];
                C ab = C(0);
#pragma unroll
                for (int q = 0; q < Q; q++) fb[q] = C(0);
                if (seed.kind != OBJ_NONE) {
                    const int coord = seed.axis == 0 ? i : (seed.axis == 1 ? j : k);
                    objective_grad<Q, C>(seed, mac, alpha, coord, fb, ab);
                }
                alpha_bar[cell] = double(ab);
                typename L::T* lp = static_cast<typename L::T*>(lam);
                const LScale sc = lam_scale<LP>(A.lsc, false);
                static_for<0, Q>([&](auto I) {
                    constexpr int q = decltype(I)::value;
                    L::template st<C>(lp + read_loc<Q, q>(g, cu, m, A.odd), fb[q], sc);
                });
            }
        } else if (do_seed) {
            alpha_bar[cell] = 0.0;
        }
    }
    if (A.part) block_partials(jsum, msum, umax, bad, fneg, A.part);
}

// ---------------------------------------------------------------------------------
// Ports: moving/boundary.rs:445-506 (NEE port_populations and its VJP) applied in the
// stationary order: after streaming, replace every population of a port cell.
// ---------------------------------------------------------------------------------
constexpr int PORT_BLOCK = 128;  // threads per block of the port kernels

struct PortArgs {
    Geom g;
    int np;
    const int* cell;      // port cell (reference index)
    const int* nb;        // inward neighbour
    const int* kind;      // 0 velocity, 1 pressure (NEE), ST_PORT_BASE + 6 rule + face
    const double* value;  // 3 per port cell
    const uint32_t* mask;
    int odd;              // parity of the state the ports act on
    int st;               // 1: some rows carry a stationary rule (the kernels with them compiled in)
    const double* stress; // regularized rules: 6 faces x (6 x 14) known-stress operators, else null
    long long pre_off;    // regularized rules: byte offset from a state buffer to the pre-port
                          // populations of its regularized rows (pre[q * np + r], storage type);
                          // 0: not kept (buffers without the region)
};

// ---------------------------------------------------------------------------------
// Stationary D3Q19 port rules (ports.rs), on a row kind ST_PORT_BASE + 6 rule + face
// (face 0..5 = xmin, xmax, ymin, ymax, zmin, zmax; sign +1 on the min faces):
//   rule 0  reconstruct_neighbor (:484-497) from the inward neighbour nb:
//           f = f^eq(rho_b, t) + nb - f^eq(rho_n, u_n), t = u_n along the normal and
//           the authored tangential velocity (value[1], value[2]) across it;
//   rule 1  reconstruct (:343-406, "legacy_moment") of a velocity port from the cell's
//           own populations, value = v0 = u - a/2;
//   rule 2  the same for a pressure port, value = (rho_b, tangential v0);
//   rules 3, 4  reconstruct_regularized (:455-482, "regularized_known_stress") of a
//           velocity / pressure port: the moment rule 1 / 2 gives the density and
//           momentum, and the populations are rebuilt as eq(rho, u) + basis . s with the
//           six stresses s fitted to the non-equilibrium parts of the 14 known
//           populations (c_axis != sign) by the operator of known_stress_operator
//           (:409-440); basis[q] = 4.5 w_q (Q00, Q11, Q22, 2 Q01, 2 Q02, 2 Q12),
//           Q_ab = c_a c_b - delta_ab / 3.  The rule reads the port cell's own pre-port
//           populations and is not affine, so its transpose needs them: the forward pass
//           keeps them beside the state (PortArgs::pre_off).
// The moment rule replaces the unknown populations (c_axis = sign) by the bounce-back
// proposal f_opp + 6 w rho (c.v) corrected through the right inverse of (1, c_t0, c_t1)
// over the unknowns, which for D3Q19 is (1/5, c_t0 / 2, c_t1 / 2).  Sums run in
// ascending q, as the reference's partitions list them, and the equilibrium is ports.rs's
// rho w (1 + 3 cu + 4.5 cu^2 - 1.5 u^2), so the FP64 result is the reference's.
// ---------------------------------------------------------------------------------
constexpr int ST_PORT_BASE = 8;

template <class C>
LBM_DEV C st_dot_c(int q, const C u[3]) {
    C acc = C(0);
    bool first = true;
#pragma unroll
    for (int d = 0; d < 3; d++) {
        const int c = cvec<19>(q, d);
        if (c == 0) continue;
        const C term = c > 0 ? u[d] : -u[d];
        acc = first ? term : acc + term;
        first = false;
    }
    return acc;
}

template <class C>
LBM_DEV void st_moments(const C* f, C& rho, C m[3]) {
    C r = f[0];
#pragma unroll
    for (int q = 1; q < 19; q++) r += f[q];
#pragma unroll
    for (int d = 0; d < 3; d++) {
        C acc = C(0);
#pragma unroll
        for (int q = 0; q < 19; q++) {
            if (cvec<19>(q, d) == 1) acc += f[q];
            if (cvec<19>(q, d) == -1) acc -= f[q];
        }
        m[d] = acc;
    }
    rho = r;
}

template <class C>
LBM_DEV void st_equilibrium(C rho, const C u[3], C* eq) {
    const C usq = u[0] * u[0] + u[1] * u[1] + u[2] * u[2];
#pragma unroll
    for (int q = 0; q < 19; q++) {
        const C cu = st_dot_c(q, u);
        eq[q] = rho * C(weight<19>(q)) * (C(1) + cu * C(3.0) + cu * cu * C(4.5) - usq * C(1.5));
    }
}

// Entry r of basis[q] in the reference's floating-point order (ports.rs:413-417);
// r = 0..5 <-> (00, 11, 22, 01, 02, 12).
LBM_HD constexpr double st_basis(int q, int r) {
    const int a = (r == 0 || r == 3 || r == 4) ? 0 : ((r == 1 || r == 5) ? 1 : 2);
    const int b = r == 0 ? 0 : ((r == 1 || r == 3) ? 1 : 2);
    const double qq = double(cvec<19>(q, a)) * double(cvec<19>(q, b)) - (a == b ? 1.0 / 3.0 : 0.0);
    return 4.5 * weight<19>(q) * (r < 3 ? qq : 2.0 * qq);
}

template <class C>
LBM_DEV void st_port_moment(const C* f, int rule, int face, const double* value, C* out);

template <class C>
LBM_DEV void st_port(const C* f, int kind, const double* value, C* out, const double* stress) {
    const int rule = (kind - ST_PORT_BASE) / 6, face = (kind - ST_PORT_BASE) % 6;
    const int axis = face >> 1, sign = (face & 1) ? -1 : 1;
    const int t0 = axis == 0 ? 1 : 0;  // the first tangential axis
    if (rule == 0) {
        C rho, m[3], vel[3], target[3], eb[19], el[19];
        st_moments(f, rho, m);
#pragma unroll
        for (int d = 0; d < 3; d++) vel[d] = qdiv(m[d], rho);
#pragma unroll
        for (int d = 0; d < 3; d++) target[d] = d == axis ? vel[d] : C(d == t0 ? value[1] : value[2]);
        st_equilibrium(C(value[0]), target, eb);
        st_equilibrium(rho, vel, el);
#pragma unroll
        for (int q = 0; q < 19; q++) out[q] = eb[q] + f[q] - el[q];
        return;
    }
    if (rule <= 2) {
        st_port_moment(f, rule, face, value, out);
        return;
    }
    // the regularized rules: target = the moment rule, then eq + basis . (op . neq_known)
    C target[19];
    st_port_moment(f, rule - 2, face, value, target);
    C rho, m[3], vel[3], eq[19];
    st_moments(target, rho, m);
#pragma unroll
    for (int d = 0; d < 3; d++) vel[d] = qdiv(m[d], rho);
    st_equilibrium(rho, vel, eq);
    const double* op = stress + face * 84;
    C s[6];
#pragma unroll
    for (int r = 0; r < 6; r++) {
        C acc = C(0);
        int k = 0;
#pragma unroll
        for (int q = 0; q < 19; q++) {
            const int c = cvec<19>(q, 0) * (axis == 0) + cvec<19>(q, 1) * (axis == 1) + cvec<19>(q, 2) * (axis == 2);
            if (c == sign) continue;
            acc += (f[q] - eq[q]) * C(op[r * 14 + k]);
            k++;
        }
        s[r] = acc;
    }
#pragma unroll
    for (int q = 0; q < 19; q++) {
        C acc = C(0);
#pragma unroll
        for (int r = 0; r < 6; r++) acc += s[r] * C(st_basis(q, r));
        out[q] = eq[q] + acc;
    }
}

// The moment reconstruction (rule 1: velocity, 2: pressure) of the cell's own populations.
template <class C>
LBM_DEV void st_port_moment(const C* f, int rule, int face, const double* value, C* out) {
    const int axis = face >> 1, sign = (face & 1) ? -1 : 1;
    const int t0 = axis == 0 ? 1 : 0, t1 = axis == 2 ? 1 : 2;
    // compatibility: zero set (the rest population first) plus twice the outgoing set
    C compat = f[0], out_sum = C(0);
    bool first = true;
#pragma unroll
    for (int q = 1; q < 19; q++) {
        const int c = cvec<19>(q, 0) * (axis == 0) + cvec<19>(q, 1) * (axis == 1) + cvec<19>(q, 2) * (axis == 2);
        if (c == 0) compat += f[q];
        if (c == -sign) {
            out_sum = first ? f[q] : out_sum + f[q];
            first = false;
        }
    }
    compat += out_sum * C(2.0);
    const double sgn = double(sign);
    C rho, v[3];
    if (rule == 1) {
#pragma unroll
        for (int d = 0; d < 3; d++) v[d] = C(value[d]);
        rho = compat / C(1.0 - sgn * value[axis]);
    } else {
        rho = C(value[0]);
#pragma unroll
        for (int d = 0; d < 3; d++) v[d] = C(d == t0 ? value[1] : value[2]);
        v[axis] = (C(1) - compat / rho) * C(sgn);
    }
    C known[19], km[3], kmass, prop[19];
#pragma unroll
    for (int q = 0; q < 19; q++) {
        const int c = cvec<19>(q, 0) * (axis == 0) + cvec<19>(q, 1) * (axis == 1) + cvec<19>(q, 2) * (axis == 2);
        const bool unknown = c == sign;
        known[q] = unknown ? C(0) : f[q];
        prop[q] = unknown ? f[opp<19>(q)] + rho * C(6.0 * weight<19>(q)) * st_dot_c(q, v) : C(0);
    }
    st_moments(known, kmass, km);
    const C target0 = rho - kmass;
    const C target1 = rho * v[t0] - km[t0], target2 = rho * v[t1] - km[t1];
    C pa0 = C(0), pa1 = C(0), pa2 = C(0);
#pragma unroll
    for (int q = 0; q < 19; q++) {
        const int c = cvec<19>(q, 0) * (axis == 0) + cvec<19>(q, 1) * (axis == 1) + cvec<19>(q, 2) * (axis == 2);
        if (c != sign) continue;
        pa0 += prop[q];
        pa1 += prop[q] * C(double(cvec<19>(q, t0)));
        pa2 += prop[q] * C(double(cvec<19>(q, t1)));
    }
    const C r0 = target0 - pa0, r1 = target1 - pa1, r2 = target2 - pa2;
#pragma unroll
    for (int q = 0; q < 19; q++) {
        const int c = cvec<19>(q, 0) * (axis == 0) + cvec<19>(q, 1) * (axis == 1) + cvec<19>(q, 2) * (axis == 2);
        if (c != sign) {
            out[q] = f[q];
            continue;
        }
        const C corr = r0 * C(0.2) + r1 * C(0.5 * cvec<19>(q, t0)) + r2 * C(0.5 * cvec<19>(q, t1));
        out[q] = prop[q] + corr;
    }
}

// Analytic transposes of the stationary rules, in FP64, equal to the dual-number cell
// Jacobian's J^T g (cell_jacobian, ports.rs:532-557) up to rounding.
//
// Moment rule (rules 1, 2; affine in f): known populations pass through; an unknown q
// gets prop_q + r . (1/5, c_t0 / 2, c_t1 / 2) with prop_q = f_opp(q) + 6 w_q rho (c_q . v)
// and r = (rho - kmass, rho v_t0 - km_t0, rho v_t1 - km_t1) - sum_unknown prop (1, c_t0,
// c_t1); rho = compat / (1 - sign v0_n) for a velocity port, or given with
// v_n = (1 - compat / rho) sign for a pressure port; compat = zero set + 2 outgoing set.
// The transpose runs these steps backwards and adds to fb.
LBM_DEV void st_moment_vjp(const double* f, int rule, int face, const double* value, const double* g, double* fb) {
    const int axis = face >> 1, sign = (face & 1) ? -1 : 1;
    const int t0 = axis == 0 ? 1 : 0, t1 = axis == 2 ? 1 : 2;
    const double sgn = double(sign);
    // forward quantities the transpose needs: rho and v
    double compat = 0.0;
#pragma unroll
    for (int q = 0; q < 19; q++) {
        const int c = cvec<19>(q, 0) * (axis == 0) + cvec<19>(q, 1) * (axis == 1) + cvec<19>(q, 2) * (axis == 2);
        if (c == 0) compat += f[q];
        if (c == -sign) compat += 2.0 * f[q];
    }
    double rho, v[3];
    if (rule == 1) {
        for (int d = 0; d < 3; d++) v[d] = value[d];
        rho = compat / (1.0 - sgn * value[axis]);
    } else {
        rho = value[0];
        v[t0] = value[1];
        v[t1] = value[2];
        v[axis] = (1.0 - compat / rho) * sgn;
    }
    // out[q] = prop_q + 0.2 r0 + 0.5 c_t0 r1 + 0.5 c_t1 r2 (unknown), f_q (known)
    double r0b = 0.0, r1b = 0.0, r2b = 0.0;
#pragma unroll
    for (int q = 0; q < 19; q++) {
        const int c = cvec<19>(q, 0) * (axis == 0) + cvec<19>(q, 1) * (axis == 1) + cvec<19>(q, 2) * (axis == 2);
        if (c != sign) {
            fb[q] += g[q];
            continue;
        }
        r0b += 0.2 * g[q];
        r1b += 0.5 * double(cvec<19>(q, t0)) * g[q];
        r2b += 0.5 * double(cvec<19>(q, t1)) * g[q];
    }
    // r = target - pa: target_bar = r_bar, pa_bar = -r_bar; target0 = rho - kmass,
    // target1/2 = rho v_t - km_t: the known populations' moments get -r_bar
    double rhob = r0b + r1b * v[t0] + r2b * v[t1], vnb = 0.0;
#pragma unroll
    for (int q = 0; q < 19; q++) {
        const int c = cvec<19>(q, 0) * (axis == 0) + cvec<19>(q, 1) * (axis == 1) + cvec<19>(q, 2) * (axis == 2);
        if (c == sign) {
            // prop_bar = g_q + pa_bar . (1, c_t0, c_t1)
            const double pb = g[q] - r0b - r1b * double(cvec<19>(q, t0)) - r2b * double(cvec<19>(q, t1));
            fb[opp<19>(q)] += pb;
            const double s6 = 6.0 * weight<19>(q);
            double cv = 0.0;
#pragma unroll
            for (int d = 0; d < 3; d++) cv += double(cvec<19>(q, d)) * v[d];
            rhob += pb * s6 * cv;
            vnb += pb * rho * s6 * double(cvec<19>(q, axis));
        } else {
            fb[q] -= r0b + r1b * double(cvec<19>(q, t0)) + r2b * double(cvec<19>(q, t1));
        }
    }
    // compat's cotangent: through rho (velocity port) or v_n (pressure port)
    const double cb = rule == 1 ? rhob / (1.0 - sgn * value[axis]) : -vnb * sgn / rho;
#pragma unroll
    for (int q = 0; q < 19; q++) {
        const int c = cvec<19>(q, 0) * (axis == 0) + cvec<19>(q, 1) * (axis == 1) + cvec<19>(q, 2) * (axis == 2);
        if (c == 0) fb[q] += cb;
        if (c == -sign) fb[q] += 2.0 * cb;
    }
}

// Transpose of eq_q = rho w_q (1 + 3 cu + 4.5 cu^2 - 1.5 u^2) (ports.rs:276-282) with the
// populations' cotangent e: rho_bar and u_bar added to rb, ub.
LBM_DEV void st_equilibrium_vjp(double rho, const double u[3], const double* e, double& rb, double ub[3]) {
    const double usq = u[0] * u[0] + u[1] * u[1] + u[2] * u[2];
#pragma unroll
    for (int q = 0; q < 19; q++) {
        const double cu = st_dot_c(q, u);
        const double w = weight<19>(q);
        rb += e[q] * w * (1.0 + 3.0 * cu + 4.5 * cu * cu - 1.5 * usq);
        const double sq = e[q] * rho * w;
#pragma unroll
        for (int d = 0; d < 3; d++) ub[d] += sq * ((3.0 + 9.0 * cu) * double(cvec<19>(q, d)) - 3.0 * u[d]);
    }
}

// Transpose of (rho, m) = moments(f), u = m / rho, given rho_bar and u_bar: f_bar.
LBM_DEV void st_velocity_vjp(double rho, const double u[3], double rb, const double ub[3], double* fb) {
    double mb[3];
#pragma unroll
    for (int d = 0; d < 3; d++) {
        mb[d] = ub[d] / rho;
        rb -= ub[d] * u[d] / rho;
    }
#pragma unroll
    for (int q = 0; q < 19; q++)
        fb[q] += rb + mb[0] * double(cvec<19>(q, 0)) + mb[1] * double(cvec<19>(q, 1)) + mb[2] * double(cvec<19>(q, 2));
}

// J^T g of a stationary row's rule at its source populations src, added to acc.
//   rule 0 (neighbour): out = eq(rho_b, t) + nb - eq(rho_n, u_n), t = (u_n along the
//     normal, the authored tangential velocity across it);
//   rules 1, 2: the moment rule above;
//   rules 3, 4 (regularized): out = eq(rho, m / rho) + basis . s, s = op . (f_known -
//     eq_known), (rho, m) the moments of target = the moment rule of f.
template <class C>
LBM_DEV void st_port_transpose(const C* src, int kind, const double* value, const C* gc, C* acc,
                               const double* stress) {
    const int rule = (kind - ST_PORT_BASE) / 6, face = (kind - ST_PORT_BASE) % 6;
    const int axis = face >> 1, sign = (face & 1) ? -1 : 1;
    double f[19], g[19], fb[19];
#pragma unroll
    for (int q = 0; q < 19; q++) {
        f[q] = double(src[q]);
        g[q] = double(gc[q]);
        fb[q] = 0.0;
    }
    if (rule == 0) {
        double rho, m[3], u[3], t[3];
        st_moments(f, rho, m);
#pragma unroll
        for (int d = 0; d < 3; d++) u[d] = m[d] / rho;
        const int t0 = axis == 0 ? 1 : 0;
#pragma unroll
        for (int d = 0; d < 3; d++) t[d] = d == axis ? u[d] : (d == t0 ? value[1] : value[2]);
        double rbb = 0.0, tb[3] = {0.0, 0.0, 0.0}, rnb = 0.0, unb[3] = {0.0, 0.0, 0.0}, ng[19];
        st_equilibrium_vjp(value[0], t, g, rbb, tb);  // rho_b is given: rbb unused
#pragma unroll
        for (int q = 0; q < 19; q++) {
            fb[q] = g[q];
            ng[q] = -g[q];
        }
        st_equilibrium_vjp(rho, u, ng, rnb, unb);
        unb[axis] += tb[axis];  // only the normal component of t follows the neighbour
        st_velocity_vjp(rho, u, rnb, unb, fb);
    } else if (rule <= 2) {
        st_moment_vjp(f, rule, face, value, g, fb);
    } else {
        double target[19], rho, m[3], vel[3];
        st_port_moment(f, rule - 2, face, value, target);
        st_moments(target, rho, m);
#pragma unroll
        for (int d = 0; d < 3; d++) vel[d] = m[d] / rho;
        const double* op = stress + face * 84;
        double sb[6];
#pragma unroll
        for (int r = 0; r < 6; r++) {
            double t = 0.0;
#pragma unroll
            for (int q = 0; q < 19; q++) t += st_basis(q, r) * g[q];
            sb[r] = t;
        }
        double eqb[19];
        int k = 0;
#pragma unroll
        for (int q = 0; q < 19; q++) {
            eqb[q] = g[q];
            const int c = cvec<19>(q, 0) * (axis == 0) + cvec<19>(q, 1) * (axis == 1) + cvec<19>(q, 2) * (axis == 2);
            if (c == sign) continue;
            double t = 0.0;
#pragma unroll
            for (int r = 0; r < 6; r++) t += op[r * 14 + k] * sb[r];
            eqb[q] -= t;
            fb[q] += t;
            k++;
        }
        double rb = 0.0, vb[3] = {0.0, 0.0, 0.0}, tgb[19];
        st_equilibrium_vjp(rho, vel, eqb, rb, vb);
#pragma unroll
        for (int q = 0; q < 19; q++) tgb[q] = 0.0;
        st_velocity_vjp(rho, vel, rb, vb, tgb);
        st_moment_vjp(f, rule - 2, face, value, tgb, fb);
    }
#pragma unroll
    for (int q = 0; q < 19; q++) acc[q] = C(double(acc[q]) + fb[q]);
}

template <int Q, class C, bool ST = false>
LBM_DEV void port_populations(const C* nb, int kind, const double* value, C* out, const double* stress) {
    if constexpr (Q == 19 && ST) {
        if (kind >= ST_PORT_BASE) {
            st_port(nb, kind, value, out, stress);
            return;
        }
    }
    C rho_n, j[3], u_n[3];
    moments<Q>(nb, rho_n, j);
#pragma unroll
    for (int d = 0; d < 3; d++) u_n[d] = qdiv(j[d], rho_n);
    C eb[Q], en[Q];
    if (kind == 0) {
        const C ub[3] = {C(value[0]), C(value[1]), C(value[2])};
        equilibrium<Q>(rho_n, ub, eb);
    } else {
        equilibrium<Q>(C(value[0]), u_n, eb);
    }
    equilibrium<Q>(rho_n, u_n, en);
#pragma unroll
    for (int q = 0; q < Q; q++) out[q] = eb[q] + nb[q] - en[q];
}

template <int Q, class C>
LBM_DEV void port_populations_vjp(const C* nb, int kind, const double* value, const C* g, C* nb_bar) {
    C rho_n, j[3], u_n[3];
    moments<Q>(nb, rho_n, j);
#pragma unroll
    for (int d = 0; d < 3; d++) u_n[d] = qdiv(j[d], rho_n);
    C rb_bar = C(0), ub_bar[3] = {C(0), C(0), C(0)};
    if (kind == 0) {
        const C ub[3] = {C(value[0]), C(value[1]), C(value[2])};
        equilibrium_vjp<Q>(rho_n, ub, g, rb_bar, ub_bar);
    } else {
        equilibrium_vjp<Q>(C(value[0]), u_n, g, rb_bar, ub_bar);
    }
    C neg[Q];
#pragma unroll
    for (int q = 0; q < Q; q++) neg[q] = -g[q];
    C rn_bar = C(0), un_bar[3] = {C(0), C(0), C(0)};
    equilibrium_vjp<Q>(rho_n, u_n, neg, rn_bar, un_bar);
    if (kind == 0) {
        rn_bar += rb_bar;
    } else {
#pragma unroll
        for (int d = 0; d < 3; d++) un_bar[d] += ub_bar[d];
    }
    C j_bar[3];
#pragma unroll
    for (int d = 0; d < 3; d++) {
        j_bar[d] = qdiv(un_bar[d], rho_n);
        rn_bar -= qdiv(un_bar[d] * j[d], rho_n * rho_n);
    }
    static_for<0, Q>([&](auto I) {
        constexpr int i = decltype(I)::value;
        C v = g[i] + rn_bar;
        if constexpr (cvec<Q>(i, 0) != 0) v += j_bar[0] * C(cvec<Q>(i, 0));
        if constexpr (cvec<Q>(i, 1) != 0) v += j_bar[1] * C(cvec<Q>(i, 1));
        if constexpr (cvec<Q>(i, 2) != 0) v += j_bar[2] * C(cvec<Q>(i, 2));
        nb_bar[i] = v;
    });
}

LBM_DEV void decode_cell(const Geom& g, long long cell, int& i, int& j, int& k) {
    const uint32_t c = uint32_t(cell);  // cell < 2^31
    const uint32_t ij = c / uint32_t(g.nz);
    k = int(c - ij * uint32_t(g.nz));
    i = int(ij / uint32_t(g.ny));
    j = int(ij - uint32_t(i) * uint32_t(g.ny));
}

template <int Q, bool MASK>
LBM_DEV uint32_t mask_of(const PortArgs& A, long long cell, Cursor<Q>& cu) {
    int i, j, k;
    decode_cell(A.g, cell, i, j, k);
    cu = make_cursor<Q>(A.g, i, j, k);
    return cell_mask<Q, MASK>(A.g, A.mask, cell, i, j, k);
}

// f(port cell) := NEE(f(neighbour)); exch_part[block] gets the block's mass change
// sum_q (new - old).
// ST: the stationary rules compiled in (rows of kind >= ST_PORT_BASE); the NEE-only
// variant keeps the register budget of the plain rule.
template <int P, int Q, bool MASK, bool ST = false>
__global__ __launch_bounds__(128) void k_port_apply(PortArgs A, void* pop_, double* exch_part) {
    using S = Store<P>;
    using C = typename S::C;
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    double delta = 0.0;
    if (r < A.np) {
        typename S::T* pop = static_cast<typename S::T*>(pop_);
        Cursor<Q> cx, cn;
        const uint32_t mx = mask_of<Q, MASK>(A, A.cell[r], cx);
        const uint32_t mn = mask_of<Q, MASK>(A, A.nb[r], cn);
        C vnb[Q], old[Q], fnb[Q], fout[Q], out[Q];
        load_cell<P, Q>(A.g, pop, cn, mn, A.odd, vnb);
        load_cell<P, Q>(A.g, pop, cx, mx, A.odd, old);
        to_full<S::shifted, Q>(vnb, fnb);
        const int kind = A.kind[r];
        port_populations<Q, C, ST>(fnb, kind, A.value + 3 * r, fout, A.stress);
        from_full<S::shifted, Q>(fout, out);
        store_cell_read<P, Q>(A.g, pop, cx, mx, A.odd, out);
        if constexpr (Q == 19 && ST) {  // the regularized rows keep their pre-port populations
            if (A.pre_off && kind >= ST_PORT_BASE + 18) {
                typename S::T* pre = reinterpret_cast<typename S::T*>(reinterpret_cast<char*>(pop_) + A.pre_off);
                static_for<0, Q>([&](auto I) {
                    constexpr int q = decltype(I)::value;
                    pre[(long long)q * A.np + r] = S::template enc<Q, q>(old[q]);
                });
            }
        }
        // Mass exchange of the stored (encoded) values, summed in order of q.
        static_for<0, Q>([&](auto I) {
            constexpr int q = decltype(I)::value;
            delta += double(S::template dec<Q, q>(S::template enc<Q, q>(out[q]))) - double(old[q]);
        });
    }
    if (exch_part) {  // per-block mass exchange, summed on the host in block order
        __shared__ double red[2];
        delta = wave_sum(delta);
        if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = delta;
        __syncthreads();
        if (threadIdx.x == 0) exch_part[blockIdx.x] = blockDim.x > 64 ? red[0] + red[1] : red[0];
    }
}

// Save the populations of every port neighbour (state of parity A.odd) to buf[q*np+r];
// for a regularized row, the port cell's pre-port populations kept beside the state.
template <int P, int Q, bool MASK>
__global__ __launch_bounds__(128) void k_port_save(PortArgs A, const void* pop_, typename Store<P>::C* buf) {
    using S = Store<P>;
    using C = typename S::C;
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= A.np) return;
    C vnb[Q], fnb[Q];
    if (Q == 19 && A.pre_off && A.kind[r] >= ST_PORT_BASE + 18) {
        const typename S::T* pre =
            reinterpret_cast<const typename S::T*>(static_cast<const char*>(pop_) + A.pre_off);
        static_for<0, Q>([&](auto I) {
            constexpr int q = decltype(I)::value;
            vnb[q] = S::template dec<Q, q>(pre[(long long)q * A.np + r]);
        });
    } else {
        Cursor<Q> cn;
        const uint32_t mn = mask_of<Q, MASK>(A, A.nb[r], cn);
        load_cell<P, Q>(A.g, static_cast<const typename S::T*>(pop_), cn, mn, A.odd, vnb);
    }
    to_full<S::shifted, Q>(vnb, fnb);
#pragma unroll
    for (int q = 0; q < Q; q++) buf[(long long)q * A.np + r] = fnb[q];
}

// Adjoint of the port pass on the cotangent of a state of parity A.odd:
// lam(nb) += J_NEE(f_nb)^T lam(port); lam(port) := 0   (solver.rs:642-656).
template <int P, int LP, int Q, bool MASK, bool ST = false>
__global__ __launch_bounds__(128) void k_port_vjp(PortArgs A, void* lam_, const typename Store<P>::C* buf, const int* lsc) {
    using L = LamStore<LP>;
    using C = typename Store<P>::C;
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= A.np) return;
    const LScale sc = lam_scale<LP>(lsc, false);
    typename L::T* lam = static_cast<typename L::T*>(lam_);
    Cursor<Q> cx, cn;
    const uint32_t mx = mask_of<Q, MASK>(A, A.cell[r], cx);
    const uint32_t mn = mask_of<Q, MASK>(A, A.nb[r], cn);
    C gx[Q], gn[Q], fnb[Q], nbb[Q];
    static_for<0, Q>([&](auto I) {
        constexpr int q = decltype(I)::value;
        gx[q] = L::template ld<C>(lam + read_loc<Q, q>(A.g, cx, mx, A.odd), sc);
        gn[q] = L::template ld<C>(lam + read_loc<Q, q>(A.g, cn, mn, A.odd), sc);
    });
#pragma unroll
    for (int q = 0; q < Q; q++) fnb[q] = buf[(long long)q * A.np + r];
    const int kind = A.kind[r];
    if constexpr (Q == 19 && ST) {
        if (kind >= ST_PORT_BASE) {
            // the moment rules read the port cell itself: its cotangent is J^T g alone
            const bool self = A.nb[r] == A.cell[r];
#pragma unroll
            for (int q = 0; q < Q; q++) nbb[q] = self ? C(0) : gn[q];
            st_port_transpose(fnb, kind, A.value + 3 * r, gx, nbb, A.stress);
            static_for<0, Q>([&](auto I) {
                constexpr int q = decltype(I)::value;
                if (!self) L::template st<C>(lam + read_loc<Q, q>(A.g, cx, mx, A.odd), C(0), sc);
                L::template st<C>(lam + read_loc<Q, q>(A.g, cn, mn, A.odd), nbb[q], sc);
            });
            return;
        }
    }
    port_populations_vjp<Q, C>(fnb, kind, A.value + 3 * r, gx, nbb);
    static_for<0, Q>([&](auto I) {
        constexpr int q = decltype(I)::value;
        L::template st<C>(lam + read_loc<Q, q>(A.g, cn, mn, A.odd), gn[q] + nbb[q], sc);
        L::template st<C>(lam + read_loc<Q, q>(A.g, cx, mx, A.odd), C(0), sc);
    });
}

// ---------------------------------------------------------------------------------
// Sponge layers (moving/sponge.rs; the blend of field.rs:1461-1475):
//   post := post (1 - sigma) + sum_l s_l f^eq(rho_l, u_l),
//   u_l = mean_l + sin_l cos(theta) - cos_l sin(theta)   (Target::velocity, field.rs:373-376).
// The step kernel leaves the post-collision value of (cell, q) in the cell's write slot,
// which no other cell touches in that step, so the blend is a pass over the sponge cells'
// write slots after the step and before the ports, and its transpose scales the
// cotangents in the same slots before the transposed step reads them.  The pass touches
// only the listed cells (the layers), so the step kernels keep their register budget.
// ---------------------------------------------------------------------------------
constexpr int SPONGE_BLOCK = 128;
constexpr int MAX_SPONGES = 8;

struct SpongeArgs {
    Geom g;
    int n;                // listed cells
    int ns;               // sponges
    int odd;              // parity of the step that wrote the slots
    const int* cell;      // ascending fluid cells covered by some sponge
    const double* sigma;  // summed strength of each listed cell
    const int* off;       // parts of listed cell r: [off[r], off[r + 1]), in sponge order
    const int* part;      // sponge index of each part
    const double* weight; // s_l of each part
    const double* phase;  // [cos theta, sin theta] of each part
    double target[MAX_SPONGES][10];  // at the step's time: rho, mean[3], sin[3], cos[3]
};

// The far-field target sum_l s_l f^eq(rho_l, u_l) of listed cell r (f^eq - w in the
// shifted storages), accumulated in sponge order.  The sponge index of a part selects
// its target row through a loop over compile-time rows, so the rows stay kernel-argument
// (scalar) loads.
template <int Q, class C, bool SH>
LBM_DEV void sponge_target(const SpongeArgs& A, int r, C* tgt) {
#pragma unroll
    for (int q = 0; q < Q; q++) tgt[q] = C(0);
    int p = A.off[r];
    const int end = A.off[r + 1];
#pragma unroll
    for (int l = 0; l < MAX_SPONGES; l++) {
        if (l < A.ns && p < end && A.part[p] == l) {
            const double* t = A.target[l];
            const C c0 = C(A.phase[2 * p]), c1 = C(A.phase[2 * p + 1]);
            Macro<Q, C> m;
            m.rho = C(t[0]);
            m.drho = C(t[0] - 1.0);
#pragma unroll
            for (int d = 0; d < 3; d++) m.u[d] = C(t[1 + d]) + C(t[4 + d]) * c0 - C(t[7 + d]) * c1;
            C e[Q];
            if constexpr (SH)
                equilibrium_shifted<Q>(m, e);
            else
                equilibrium<Q>(m.rho, m.u, e);
            const C s = C(A.weight[p]);
#pragma unroll
            for (int q = 0; q < Q; q++) tgt[q] += e[q] * s;
            p++;
        }
    }
}

// Blend of the slots written by the step of parity A.odd; exch_part[block] gets the
// block's mass change sum_q (new - old) of the stored values.
template <int P, int Q>
__global__ __launch_bounds__(SPONGE_BLOCK) void k_sponge(SpongeArgs A, void* pop_, double* exch_part) {
    using S = Store<P>;
    using C = typename S::C;
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    double delta = 0.0;
    if (r < A.n) {
        typename S::T* pop = static_cast<typename S::T*>(pop_);
        int i, j, k;
        decode_cell(A.g, A.cell[r], i, j, k);
        const Cursor<Q> cu = make_cursor<Q>(A.g, i, j, k);
        C tgt[Q];
        sponge_target<Q, C, S::shifted>(A, r, tgt);
        // in the shifted storages d' = d (1 - sigma) + sum_l s_l (f^eq_l - w), as sigma = sum_l s_l
        const C keep = C(1.0 - A.sigma[r]);
        static_for<0, Q>([&](auto I) {
            constexpr int q = decltype(I)::value;
            typename S::T& slot = at_write<Q, q>(pop, A.g, cu, A.odd);
            const C old = S::template dec<Q, q>(slot);
            const typename S::T nv = S::template enc<Q, q>(old * keep + tgt[q]);
            slot = nv;
            delta += double(S::template dec<Q, q>(nv)) - double(old);
        });
    }
    if (exch_part) {  // per-block mass exchange, summed on the host in block order
        __shared__ double red[SPONGE_BLOCK / 64];
        delta = wave_sum(delta);
        if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = delta;
        __syncthreads();
        if (threadIdx.x == 0) exch_part[blockIdx.x] = red[0] + red[1];
    }
}

// Transpose of the blend on the cotangent of the post-collision values (the write slots
// of parity A.odd): lambda *= 1 - sigma.  The far field does not depend on the state.
template <int P, int LP, int Q>
__global__ __launch_bounds__(SPONGE_BLOCK) void k_sponge_vjp(SpongeArgs A, void* lam_, const int* lsc) {
    using L = LamStore<LP>;
    using C = typename Store<P>::C;
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= A.n) return;
    const LScale sc = lam_scale<LP>(lsc, false);
    typename L::T* lam = static_cast<typename L::T*>(lam_);
    int i, j, k;
    decode_cell(A.g, A.cell[r], i, j, k);
    const Cursor<Q> cu = make_cursor<Q>(A.g, i, j, k);
    const C keep = C(1.0 - A.sigma[r]);
    static_for<0, Q>([&](auto I) {
        constexpr int q = decltype(I)::value;
        typename L::T* slot = &at_write<Q, q>(lam, A.g, cu, A.odd);
        L::template st<C>(slot, L::template ld<C>(slot, sc) * keep, sc);
    });
}

// ---------------------------------------------------------------------------------
// Layout conversion between the twisted device layout and the reference AoS layout.
// ---------------------------------------------------------------------------------
template <int P, int Q, bool MASK>
__global__ __launch_bounds__(256) void k_get_state(Geom g, const void* pop_, const uint32_t* mask, int odd, double* aos) {
    using S = Store<P>;
    using C = typename S::C;
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y * blockDim.y + threadIdx.y, i = blockIdx.z;
    if (k >= g.nz || j >= g.ny) return;
    const long long cell = ((long long)i * g.ny + j) * g.nz + k;
    const uint32_t m = cell_mask<Q, MASK>(g, mask, cell, i, j, k);
    if (m & SOLID_BIT) {
#pragma unroll
        for (int q = 0; q < Q; q++) aos[cell * Q + q] = weight<Q>(q);
        return;
    }
    const Cursor<Q> cu = make_cursor<Q>(g, i, j, k);
    C f[Q];
    load_cell<P, Q>(g, static_cast<const typename S::T*>(pop_), cu, m, odd, f);
#pragma unroll
    for (int q = 0; q < Q; q++) aos[cell * Q + q] = S::shifted ? weight<Q>(q) + double(f[q]) : double(f[q]);
}

template <int P, int Q, bool MASK>
__global__ __launch_bounds__(256) void k_set_state(Geom g, void* pop_, const uint32_t* mask, int odd, const double* aos) {
    using S = Store<P>;
    using C = typename S::C;
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y * blockDim.y + threadIdx.y, i = blockIdx.z;
    if (k >= g.nz || j >= g.ny) return;
    const long long cell = ((long long)i * g.ny + j) * g.nz + k;
    const uint32_t m = cell_mask<Q, MASK>(g, mask, cell, i, j, k);
    if (m & SOLID_BIT) return;
    const Cursor<Q> cu = make_cursor<Q>(g, i, j, k);
    C f[Q];
#pragma unroll
    for (int q = 0; q < Q; q++) f[q] = S::shifted ? C(aos[cell * Q + q] - weight<Q>(q)) : C(aos[cell * Q + q]);
    store_cell_read<P, Q>(g, static_cast<typename S::T*>(pop_), cu, m, odd, f);
}

template <int LP, int Q, bool MASK>
__global__ __launch_bounds__(256) void k_get_lam(Geom g, const void* lam_, const uint32_t* mask, int odd, double* aos, const int* lsc) {
    using L = LamStore<LP>;
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y * blockDim.y + threadIdx.y, i = blockIdx.z;
    if (k >= g.nz || j >= g.ny) return;
    const long long cell = ((long long)i * g.ny + j) * g.nz + k;
    const uint32_t m = cell_mask<Q, MASK>(g, mask, cell, i, j, k);
    if (m & SOLID_BIT) {
#pragma unroll
        for (int q = 0; q < Q; q++) aos[cell * Q + q] = 0.0;
        return;
    }
    const Cursor<Q> cu = make_cursor<Q>(g, i, j, k);
    const typename L::T* lam = static_cast<const typename L::T*>(lam_);
    const LScale sc = lam_scale<LP>(lsc, false);
    static_for<0, Q>([&](auto I) {
        constexpr int q = decltype(I)::value;
        aos[cell * Q + q] = L::template ld<double>(lam + read_loc<Q, q>(g, cu, m, odd), sc);
    });
}

template <int LP, int Q, bool MASK>
__global__ __launch_bounds__(256) void k_set_lam(Geom g, void* lam_, const uint32_t* mask, int odd, const double* aos,
                                                const int* lsc, int accumulate) {
    using L = LamStore<LP>;
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y * blockDim.y + threadIdx.y, i = blockIdx.z;
    if (k >= g.nz || j >= g.ny) return;
    const long long cell = ((long long)i * g.ny + j) * g.nz + k;
    const uint32_t m = cell_mask<Q, MASK>(g, mask, cell, i, j, k);
    if (m & SOLID_BIT) return;
    const Cursor<Q> cu = make_cursor<Q>(g, i, j, k);
    typename L::T* lam = static_cast<typename L::T*>(lam_);
    const LScale sc = lam_scale<LP>(lsc, false);
    static_for<0, Q>([&](auto I) {
        constexpr int q = decltype(I)::value;
        typename L::T* p = lam + read_loc<Q, q>(g, cu, m, odd);
        const double prev = accumulate ? L::template ld<double>(p, sc) : 0.0;
        L::template st<double>(p, prev + aos[cell * Q + q], sc);
    });
}

// Macroscopic fields (solver.rs:690-705 in lattice units): rho, and u (0 in solids).
// Terminal responses of a state (solver.rs:874-894): per block the sums of
// [u_x, u_y, u_z, rho|u|^2, rho alpha/dt |u|^2] (physical units, fluid cells) into
// sums[5 * block + 0..4], folded on the host in block order.
template <int P, int Q, bool MASK>
__global__ __launch_bounds__(256) void k_vector(Geom g, const void* pop_, const uint32_t* mask, int odd,
                                               const void* alpha_, double a0, double a1, double a2, double vel,
                                               double dens, double inv_dt, double* sums) {
    using S = Store<P>;
    using C = typename S::C;
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y * blockDim.y + threadIdx.y, i = blockIdx.z;
    double v[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
    if (k < g.nz && j < g.ny) {
        const long long cell = ((long long)i * g.ny + j) * g.nz + k;
        const uint32_t m = cell_mask<Q, MASK>(g, mask, cell, i, j, k);
        if (!(m & SOLID_BIT)) {
            const Cursor<Q> cu = make_cursor<Q>(g, i, j, k);
            C f[Q];
            load_cell<P, Q>(g, static_cast<const typename S::T*>(pop_), cu, m, odd, f);
            const C alpha = static_cast<const C*>(alpha_)[cell];
            const C a[3] = {C(a0), C(a1), C(a2)};
            Macro<Q, C> mac;
            macro_of<Q, C, S::shifted>(f, alpha, a, mac);
            const double u0 = double(mac.u[0]) * vel, u1 = double(mac.u[1]) * vel, u2 = double(mac.u[2]) * vel;
            const double r = double(mac.rho) * dens;
            const double usq = u0 * u0 + u1 * u1 + u2 * u2;
            v[0] = u0;
            v[1] = u1;
            v[2] = u2;
            v[3] = r * usq;
            v[4] = r * double(alpha) * inv_dt * usq;
        }
    }
    __shared__ double red[5][4];
    const int tid = threadIdx.x + blockDim.x * threadIdx.y;
#pragma unroll
    for (int c = 0; c < 5; c++) {
        const double t = wave_sum(v[c]);
        if ((tid & 63) == 0) red[c][tid >> 6] = t;
    }
    __syncthreads();
    if (tid == 0) {
        const int nw = int(blockDim.x * blockDim.y + 63) >> 6;
        for (int c = 0; c < 5; c++) {
            double acc = 0.0;
            for (int w = 0; w < nw; w++) acc += red[c][w];
            sums[5 * block_linear() + c] = acc;
        }
    }
}

// vector_vjp (solver.rs:896-947) without the design-volume term: f_bar (reference AoS,
// w.r.t. f) and alpha_bar of ob (an OBJ_VECTOR objective of weight 1) at every cell;
// zero at solid cells.
template <int P, int Q, bool MASK>
__global__ __launch_bounds__(256) void k_vector_vjp(Geom g, const void* pop_, const uint32_t* mask, int odd,
                                                   const void* alpha_, double a0, double a1, double a2, ObjArgs ob,
                                                   double* f_bar, double* alpha_bar) {
    using S = Store<P>;
    using C = typename S::C;
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y * blockDim.y + threadIdx.y, i = blockIdx.z;
    if (k >= g.nz || j >= g.ny) return;
    const long long cell = ((long long)i * g.ny + j) * g.nz + k;
    const uint32_t m = cell_mask<Q, MASK>(g, mask, cell, i, j, k);
    if (m & SOLID_BIT) {
#pragma unroll
        for (int q = 0; q < Q; q++) f_bar[cell * Q + q] = 0.0;
        alpha_bar[cell] = 0.0;
        return;
    }
    const Cursor<Q> cu = make_cursor<Q>(g, i, j, k);
    C f[Q], fb[Q];
    load_cell<P, Q>(g, static_cast<const typename S::T*>(pop_), cu, m, odd, f);
    const C alpha = static_cast<const C*>(alpha_)[cell];
    const C a[3] = {C(a0), C(a1), C(a2)};
    Macro<Q, C> mac;
    macro_of<Q, C, S::shifted>(f, alpha, a, mac);
#pragma unroll
    for (int q = 0; q < Q; q++) fb[q] = C(0);
    C ab = C(0);
    objective_grad<Q, C>(ob, mac, alpha, 0, fb, ab);
#pragma unroll
    for (int q = 0; q < Q; q++) f_bar[cell * Q + q] = double(fb[q]);
    alpha_bar[cell] = double(ab);
}

template <int P, int Q, bool MASK>
__global__ __launch_bounds__(256) void k_fields(Geom g, const void* pop_, const uint32_t* mask, int odd, const void* alpha_, double a0,
                         double a1, double a2, double* rho, double* u) {
    using S = Store<P>;
    using C = typename S::C;
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y * blockDim.y + threadIdx.y, i = blockIdx.z;
    if (k >= g.nz || j >= g.ny) return;
    const long long cell = ((long long)i * g.ny + j) * g.nz + k;
    const uint32_t m = cell_mask<Q, MASK>(g, mask, cell, i, j, k);
    C f[Q];
    if (m & SOLID_BIT) {
#pragma unroll
        for (int q = 0; q < Q; q++) f[q] = S::shifted ? C(0) : C(weight<Q>(q));
    } else {
        const Cursor<Q> cu = make_cursor<Q>(g, i, j, k);
        load_cell<P, Q>(g, static_cast<const typename S::T*>(pop_), cu, m, odd, f);
    }
    const C a[3] = {C(a0), C(a1), C(a2)};
    Macro<Q, C> mac;
    macro_of<Q, C, S::shifted>(f, static_cast<const C*>(alpha_)[cell], a, mac);
    rho[cell] = double(mac.rho);
    for (int d = 0; d < 3; d++) u[3 * cell + d] = (m & SOLID_BIT) ? 0.0 : double(mac.u[d]);
}

}  // namespace lbm
