// Per-cell collision arithmetic and its transposed Jacobian (device code).
//
// Every function mirrors the reference's operation order so that, compiled with
// -ffp-contract=off in double precision, it reproduces the reference's IEEE results:
//   moments / equilibrium / equilibrium_vjp   moving/lattice.rs:160-222
//   collide_pairs / collide_pairs_vjp (TRT,BGK) moving/collision.rs:457-570
//   MRT collide / collide_vjp                 moving/collision.rs:326-453,574-633
//   Brinkman branch of cell_update(_vjp)      moving/psm.rs:213-226,301-330,455-469
//   stationary BGK-Guo collide_cell(_vjp)     solver.rs:189-196,214-285,298-322
// In single precision the same order is used with float arithmetic.
#pragma once
#include <type_traits>

#include "lattice.h"

namespace lbm {

#define LBM_DEV __device__ __forceinline__

// v . c_q summing only the non-zero components, in x, y, z order.  Equal in value to
// the reference's v0*c0 + v1*c1 + v2*c2 with c in {-1,0,1}.
template <int Q, int q, class C>
LBM_DEV C dotc(const C v[3]) {
    constexpr int c0 = cvec<Q>(q, 0), c1 = cvec<Q>(q, 1), c2 = cvec<Q>(q, 2);
    C r = C(0);
    bool first = true;
    if constexpr (c0 != 0) {
        r = c0 > 0 ? v[0] : -v[0];
        first = false;
    }
    if constexpr (c1 != 0) {
        const C t = c1 > 0 ? v[1] : -v[1];
        r = first ? t : r + t;
        first = false;
    }
    if constexpr (c2 != 0) {
        const C t = c2 > 0 ? v[2] : -v[2];
        r = first ? t : r + t;
    }
    return r;
}

// Compile-time loop helper: calls f(std::integral_constant<int, I>) for I in [B, E).
template <int B, int E, class F>
LBM_DEV void static_for(F&& f) {
    if constexpr (B < E) {
        f(std::integral_constant<int, B>{});
        static_for<B + 1, E>(f);
    }
}

// rho = f0 + f1 + ...; j_d = sum over i >= 1 of f_i c_id (lattice.rs:168-180).
template <int Q, class C>
LBM_DEV void moments(const C* f, C& rho, C j[3]) {
    C r = f[0];
    C m0 = C(0), m1 = C(0), m2 = C(0);
    bool s0 = false, s1 = false, s2 = false;
    static_for<1, Q>([&](auto I) {
        constexpr int i = decltype(I)::value;
        r += f[i];
        if constexpr (cvec<Q>(i, 0) != 0) {
            const C t = cvec<Q>(i, 0) > 0 ? f[i] : -f[i];
            m0 = s0 ? m0 + t : t;
            s0 = true;
        }
        if constexpr (cvec<Q>(i, 1) != 0) {
            const C t = cvec<Q>(i, 1) > 0 ? f[i] : -f[i];
            m1 = s1 ? m1 + t : t;
            s1 = true;
        }
        if constexpr (cvec<Q>(i, 2) != 0) {
            const C t = cvec<Q>(i, 2) > 0 ? f[i] : -f[i];
            m2 = s2 ? m2 + t : t;
            s2 = true;
        }
    });
    rho = r;
    j[0] = m0;
    j[1] = m1;
    j[2] = m2;
}

// Brinkman macroscopic quantities (psm.rs:213-222 with zero solid momentum, equal in
// value to solver.rs:190-196): u = (j/rho + a/2)/(1 + alpha/2), F = rho (a - alpha u).
// a / b: an IEEE division in FP64, keeping the reference's operations bit for bit; in
// FP32 a multiply by the hardware reciprocal (1 ulp), shared by the compiler between
// divisions by the same value (an FP32 IEEE division is about ten instructions).
template <class C>
LBM_DEV C qdiv(C a, C b) {
    return a / b;
}
template <>
LBM_DEV float qdiv<float>(float a, float b) {
    return a * __builtin_amdgcn_rcpf(b);
}

template <int Q, class C>
struct Macro {
    C rho, drho, j[3], den, sv[3], u[3], F[3];  // drho = rho - 1
};

template <int Q, class C>
LBM_DEV void brinkman_macro(const C* f, C alpha, const C a[3], Macro<Q, C>& m) {
    moments<Q>(f, m.rho, m.j);
    m.drho = m.rho - C(1);
    m.den = C(1) + alpha * C(0.5);
#pragma unroll
    for (int d = 0; d < 3; d++) {
        m.sv[d] = qdiv(m.j[d], m.rho) + a[d] * C(0.5);
        m.u[d] = qdiv(m.sv[d], m.den);
    }
#pragma unroll
    for (int d = 0; d < 3; d++) m.F[d] = m.rho * (a[d] + (-(alpha * m.u[d])));
}

// TRT / BGK collide_pairs (collision.rs:457-488).
template <int Q, class C>
LBM_DEV void collide_pairs(const C* f, C rho, const C u[3], const C F[3], C omega, C om, C* out) {
    const C usq = u[0] * u[0] + u[1] * u[1] + u[2] * u[2];
    const C uf = u[0] * F[0] + u[1] * F[1] + u[2] * F[2];
    const C base = C(1) - usq * C(1.5);
    const C kp = C(1) - omega * C(0.5);
    const C km = C(1) - om * C(0.5);
    constexpr double w0 = weight<Q>(0);
    const C eq0 = rho * C(w0) * base;
    out[0] = f[0] - omega * (f[0] - eq0) - kp * uf * C(3.0 * w0);
    static_for<1, Q>([&](auto I) {
        constexpr int i = decltype(I)::value;
        constexpr int o = opp<Q>(i);
        if constexpr (o > i) {
            constexpr double w = weight<Q>(i);
            const C cu = dotc<Q, i>(u);
            const C cf = dotc<Q, i>(F);
            const C wr = rho * C(w);
            const C eqp = wr * (base + cu * cu * C(4.5));
            const C eqm = wr * cu * C(3.0);
            const C sp = (cu * cf * C(9.0) - uf * C(3.0)) * C(w);
            const C sm = cf * C(3.0 * w);
            const C fp = (f[i] + f[o]) * C(0.5);
            const C fm = (f[i] - f[o]) * C(0.5);
            const C aa = kp * sp - omega * (fp - eqp);
            const C bb = km * sm - om * (fm - eqm);
            out[i] = f[i] + aa + bb;
            out[o] = f[o] + aa - bb;
        }
    });
}

// collide_pairs_vjp (collision.rs:492-570) without the omega cotangent.
template <int Q, class C>
LBM_DEV void collide_pairs_vjp(const C rho, const C u[3], const C F[3], C omega, C om, const C* g, C* fb,
                               C& rho_bar, C u_bar[3], C F_bar[3]) {
    const C usq = u[0] * u[0] + u[1] * u[1] + u[2] * u[2];
    const C base = C(1) - usq * C(1.5);
    const C kp = C(1) - omega * C(0.5);
    const C km = C(1) - om * C(0.5);
    C rb = C(0), usq_bar = C(0), uf_bar = C(0);
    C ub0 = C(0), ub1 = C(0), ub2 = C(0), fb0 = C(0), fb1 = C(0), fb2 = C(0);
    constexpr double w0 = weight<Q>(0);
    const C g0 = g[0];
    fb[0] = g0 * (C(1) - omega);
    rb += g0 * omega * C(w0) * base;
    usq_bar -= g0 * omega * rho * C(1.5 * w0);
    uf_bar -= g0 * kp * C(3.0 * w0);
    static_for<1, Q>([&](auto I) {
        constexpr int i = decltype(I)::value;
        constexpr int o = opp<Q>(i);
        if constexpr (o > i) {
            constexpr double w = weight<Q>(i);
            const C cu = dotc<Q, i>(u);
            const C cf = dotc<Q, i>(F);
            const C wr = rho * C(w);
            const C a_bar = g[i] + g[o];
            const C b_bar = g[i] - g[o];
            const C fp_bar = -omega * a_bar;
            const C fm_bar = -om * b_bar;
            const C eqp_bar = omega * a_bar;
            const C eqm_bar = om * b_bar;
            const C sp_bar = kp * a_bar;
            const C sm_bar = km * b_bar;
            fb[i] = g[i] + (fp_bar + fm_bar) * C(0.5);
            fb[o] = g[o] + (fp_bar - fm_bar) * C(0.5);
            rb += eqp_bar * C(w) * (base + cu * cu * C(4.5)) + eqm_bar * cu * C(3.0 * w);
            usq_bar -= eqp_bar * wr * C(1.5);
            const C cu_bar = eqp_bar * wr * cu * C(9.0) + eqm_bar * wr * C(3.0) + sp_bar * cf * C(9.0 * w);
            const C cf_bar = sp_bar * cu * C(9.0 * w) + sm_bar * C(3.0 * w);
            uf_bar -= sp_bar * C(3.0 * w);
            constexpr int c0 = cvec<Q>(i, 0), c1 = cvec<Q>(i, 1), c2 = cvec<Q>(i, 2);
            if constexpr (c0 != 0) {
                ub0 += cu_bar * C(c0);
                fb0 += cf_bar * C(c0);
            }
            if constexpr (c1 != 0) {
                ub1 += cu_bar * C(c1);
                fb1 += cf_bar * C(c1);
            }
            if constexpr (c2 != 0) {
                ub2 += cu_bar * C(c2);
                fb2 += cf_bar * C(c2);
            }
        }
    });
    u_bar[0] = ub0 + (usq_bar * u[0] * C(2.0) + uf_bar * F[0]);
    u_bar[1] = ub1 + (usq_bar * u[1] * C(2.0) + uf_bar * F[1]);
    u_bar[2] = ub2 + (usq_bar * u[2] * C(2.0) + uf_bar * F[2]);
    F_bar[0] = fb0 + uf_bar * u[0];
    F_bar[1] = fb1 + uf_bar * u[1];
    F_bar[2] = fb2 + uf_bar * u[2];
    rho_bar = rb;
}

// FP32 form of collide_pairs_vjp (same cotangents, regrouped per pair with the
// per-weight coefficients formed once per cell); the FP64 path keeps the reference's
// operation order above.
template <int Q>
LBM_DEV void collide_pairs_vjp_fast(const float rho, const float u[3], const float F[3], float omega, float om,
                                    const float* g, float* fb, float& rho_bar, float u_bar[3], float F_bar[3]) {
    const float usq = u[0] * u[0] + u[1] * u[1] + u[2] * u[2];
    const float base = 1.0f - usq * 1.5f;
    const float kp = 1.0f - omega * 0.5f;
    const float km = 1.0f - om * 0.5f;
    const float hp = omega * 0.5f, hm = om * 0.5f;
    const float r1 = omega * base;        // times w: d rho_bar / d (g_i + g_o), cu-free part
    const float r2 = 4.5f * omega;        // times w cu^2
    const float r3 = 3.0f * om;           // times w cu (g_i - g_o)
    const float c1 = 9.0f * omega * rho;  // cu_bar: times w cu (g_i + g_o)
    const float c2 = 9.0f * kp;           // cu_bar: times w cf (g_i + g_o); cf_bar: times w cu (g_i + g_o)
    const float c3 = 3.0f * om * rho;     // cu_bar: times w (g_i - g_o)
    const float c4 = 3.0f * km;           // cf_bar: times w (g_i - g_o)
    float rb = 0.0f, wa = 0.0f;           // wa = sum_pairs w (g_i + g_o) (+ w0 g0)
    float ub0 = 0.0f, ub1 = 0.0f, ub2 = 0.0f, fb0 = 0.0f, fb1 = 0.0f, fb2 = 0.0f;
    constexpr float w0 = float(weight<Q>(0));
    const float g0 = g[0];
    fb[0] = g0 - omega * g0;
    rb += w0 * g0 * r1;
    wa += w0 * g0;
    static_for<1, Q>([&](auto I) {
        constexpr int i = decltype(I)::value;
        constexpr int o = opp<Q>(i);
        if constexpr (o > i) {
            constexpr float w = float(weight<Q>(i));
            const float cu = dotc<Q, i>(u);
            const float cf = dotc<Q, i>(F);
            const float a = g[i] + g[o];
            const float b = g[i] - g[o];
            const float hb = hm * b;
            fb[i] = g[i] - (hp * a + hb);
            fb[o] = g[o] - (hp * a - hb);
            const float wa_i = w * a, wb_i = w * b;
            rb += wa_i * (r1 + r2 * cu * cu) + wb_i * (r3 * cu);
            wa += wa_i;
            const float cu_bar = wa_i * (c1 * cu + c2 * cf) + wb_i * c3;
            const float cf_bar = wa_i * (c2 * cu) + wb_i * c4;
            constexpr int e0 = cvec<Q>(i, 0), e1 = cvec<Q>(i, 1), e2 = cvec<Q>(i, 2);
            if constexpr (e0 != 0) {
                ub0 += e0 > 0 ? cu_bar : -cu_bar;
                fb0 += e0 > 0 ? cf_bar : -cf_bar;
            }
            if constexpr (e1 != 0) {
                ub1 += e1 > 0 ? cu_bar : -cu_bar;
                fb1 += e1 > 0 ? cf_bar : -cf_bar;
            }
            if constexpr (e2 != 0) {
                ub2 += e2 > 0 ? cu_bar : -cu_bar;
                fb2 += e2 > 0 ? cf_bar : -cf_bar;
            }
        }
    });
    // usq_bar = -1.5 omega rho sum w a; uf_bar = -3 kp sum w a
    const float usq_bar = -1.5f * omega * rho * wa;
    const float uf_bar = -3.0f * kp * wa;
    u_bar[0] = ub0 + (usq_bar * u[0] * 2.0f + uf_bar * F[0]);
    u_bar[1] = ub1 + (usq_bar * u[1] * 2.0f + uf_bar * F[1]);
    u_bar[2] = ub2 + (usq_bar * u[2] * 2.0f + uf_bar * F[2]);
    F_bar[0] = fb0 + uf_bar * u[0];
    F_bar[1] = fb1 + uf_bar * u[1];
    F_bar[2] = fb2 + uf_bar * u[2];
    rho_bar = rb;
}

// Transposed TRT / BGK collision and Brinkman map of one cell in the shifted single
// precision form: collide_pairs_vjp (collision.rs:492-570) followed by brinkman_chain
// below (psm.rs:317-330, 455-469), evaluated through the moments of the cotangent.
// The equilibrium and source terms of direction q are w_q P(c_q), P of degree two in c_q
// with coefficients in (rho, u, F):
//   w_q [omega A - 3 kp uf + 4.5 omega rho (c.u)^2 + 9 kp (c.u)(c.F) + 3 om rho c.u + 3 km c.F],
//   A = drho - 1.5 rho |u|^2,
// so their pairing with g needs only M0 = sum w g, M1 = sum w c g and M2 = sum w c c^T g,
// and the cotangents of rho, u and F follow from these ten numbers.  The Brinkman chain
// maps them to cotangents R of the density and J of the momentum, and one pass forms
//   fb_q = g_q - omega g+_q - om g-_q + R + c_q . J
// (g+-: even/odd parts over the pair q, opp(q)).  R0, J0 are density and momentum
// cotangents the caller adds (objective terms).  Returns the cotangent of alpha.
template <int Q>
LBM_DEV float trt_adjoint_shifted(const Macro<Q, float>& m, float alpha, const float a[3], float omega, float om,
                                  const float* g, float R0, const float J0[3], float* fb) {
    // moments of g; M2 as xx, yy, zz, xy, xz, yz
    float M0 = float(weight<Q>(0)) * g[0];
    float M1[3] = {0.0f, 0.0f, 0.0f}, M2[6] = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
    static_for<1, Q>([&](auto I) {
        constexpr int i = decltype(I)::value;
        constexpr int o = opp<Q>(i);
        if constexpr (o > i) {
            constexpr float w = float(weight<Q>(i));
            constexpr int c0 = cvec<Q>(i, 0), c1 = cvec<Q>(i, 1), c2 = cvec<Q>(i, 2);
            const float pa = g[i] + g[o], pb = g[i] - g[o];
            M0 += w * pa;
            if constexpr (c0 != 0) M1[0] += float(c0) * w * pb;
            if constexpr (c1 != 0) M1[1] += float(c1) * w * pb;
            if constexpr (c2 != 0) M1[2] += float(c2) * w * pb;
            if constexpr (c0 != 0) M2[0] += w * pa;
            if constexpr (c1 != 0) M2[1] += w * pa;
            if constexpr (c2 != 0) M2[2] += w * pa;
            if constexpr (c0 * c1 != 0) M2[3] += float(c0 * c1) * w * pa;
            if constexpr (c0 * c2 != 0) M2[4] += float(c0 * c2) * w * pa;
            if constexpr (c1 * c2 != 0) M2[5] += float(c1 * c2) * w * pa;
        }
    });
    const float* u = m.u;
    const float* F = m.F;
    const float Mu[3] = {M2[0] * u[0] + M2[3] * u[1] + M2[4] * u[2], M2[3] * u[0] + M2[1] * u[1] + M2[5] * u[2],
                         M2[4] * u[0] + M2[5] * u[1] + M2[2] * u[2]};
    const float MF[3] = {M2[0] * F[0] + M2[3] * F[1] + M2[4] * F[2], M2[3] * F[0] + M2[1] * F[1] + M2[5] * F[2],
                         M2[4] * F[0] + M2[5] * F[1] + M2[2] * F[2]};
    const float usq = u[0] * u[0] + u[1] * u[1] + u[2] * u[2];
    const float uMu = u[0] * Mu[0] + u[1] * Mu[1] + u[2] * Mu[2];
    const float uM1 = u[0] * M1[0] + u[1] * M1[1] + u[2] * M1[2];
    const float rho = m.rho;
    const float kp = 1.0f - omega * 0.5f, km = 1.0f - om * 0.5f;
    // cotangents of rho (through drho and rho), u and F of the collision
    float R = omega * M0 * (1.0f - 1.5f * usq) + 4.5f * omega * uMu + 3.0f * om * uM1;
    const float s1 = -3.0f * omega * rho * M0, s2 = -3.0f * kp * M0;
    const float s3 = 9.0f * omega * rho, s4 = 9.0f * kp, s5 = 3.0f * om * rho, s6 = 3.0f * km;
    float ub[3], Fb[3];
#pragma unroll
    for (int d = 0; d < 3; d++) {
        ub[d] = s1 * u[d] + s2 * F[d] + s3 * Mu[d] + s4 * MF[d] + s5 * M1[d];
        Fb[d] = s2 * u[d] + s4 * Mu[d] + s6 * M1[d];
    }
    // F = rho (a - alpha u)
    float ab = 0.0f;
#pragma unroll
    for (int d = 0; d < 3; d++) {
        R += Fb[d] * (a[d] - alpha * u[d]);
        ab -= rho * Fb[d] * u[d];
        ub[d] -= alpha * rho * Fb[d];
    }
    // u = (j / rho + a / 2) / den
    const float irho = __builtin_amdgcn_rcpf(rho);
    const float iden = __builtin_amdgcn_rcpf(m.den);
    const float ir = irho * iden;
    float J[3];
#pragma unroll
    for (int d = 0; d < 3; d++) {
        ab -= 0.5f * ub[d] * u[d] * iden;
        J[d] = ub[d] * ir;
        R -= J[d] * (m.j[d] * irho);
    }
    R += R0;
#pragma unroll
    for (int d = 0; d < 3; d++) J[d] += J0[d];
    // per-direction pass: fb_i = k1 g_i + k2 g_o + R + c_i.J, fb_o = k1 g_o + k2 g_i + R - c_i.J
    const float k1 = 1.0f - 0.5f * (omega + om), k2 = 0.5f * (om - omega);
    fb[0] = (1.0f - omega) * g[0] + R;
    static_for<1, Q>([&](auto I) {
        constexpr int i = decltype(I)::value;
        constexpr int o = opp<Q>(i);
        if constexpr (o > i) {
            const float cJ = dotc<Q, i>(J);
            fb[i] = k1 * g[i] + (k2 * g[o] + (R + cJ));
            fb[o] = k1 * g[o] + (k2 * g[i] + (R - cJ));
        }
    });
    return ab;
}

// Chain the collision cotangents through the Brinkman macroscopic map back to f and
// alpha: psm.rs:317-330 and :455-469 with m = 0, dp_bar = 0.
template <int Q, class C>
LBM_DEV void brinkman_chain(const Macro<Q, C>& m, C alpha, const C a[3], C cb_rho, const C cb_u[3], const C cb_F[3],
                            C* fb, C& alpha_bar) {
    C rho_bar = C(0) + cb_rho;
    C u_bar[3] = {C(0) + cb_u[0], C(0) + cb_u[1], C(0) + cb_u[2]};
    C ab = C(0);
#pragma unroll
    for (int d = 0; d < 3; d++) {
        const C z = -(alpha * m.u[d]);
        rho_bar += cb_F[d] * (a[d] + z) + C(0) * z;
        const C zb = (cb_F[d] + C(0)) * m.rho;
        ab -= zb * m.u[d];
        u_bar[d] -= zb * alpha;
    }
    C j_bar[3];
#pragma unroll
    for (int d = 0; d < 3; d++) {
        const C s_bar = qdiv(u_bar[d], m.den);
        ab -= qdiv(u_bar[d] * m.sv[d], m.den * m.den) * C(0.5);
        j_bar[d] = qdiv(s_bar, m.rho);
        rho_bar -= qdiv(s_bar * m.j[d], m.rho * m.rho);
    }
    static_for<0, Q>([&](auto I) {
        constexpr int i = decltype(I)::value;
        C v = fb[i] + rho_bar;
        if constexpr (cvec<Q>(i, 0) != 0) v += j_bar[0] * C(cvec<Q>(i, 0));
        if constexpr (cvec<Q>(i, 1) != 0) v += j_bar[1] * C(cvec<Q>(i, 1));
        if constexpr (cvec<Q>(i, 2) != 0) v += j_bar[2] * C(cvec<Q>(i, 2));
        fb[i] = v;
    });
    alpha_bar = ab;
}

// ---------------------------------------------------------------------------------
// Stationary BGK-Guo, D3Q19 (solver.rs:298-322 and 214-285).
// ---------------------------------------------------------------------------------
template <class C>
LBM_DEV C st_dot(int c0, int c1, int c2, const C v[3]) {
    // 0.0 + v0*c0 + v1*c1 + v2*c2 in value (solver.rs:314-315), zero terms skipped.
    C r = C(0);
    if (c0 != 0) r = r + (c0 > 0 ? v[0] : -v[0]);
    if (c1 != 0) r = r + (c1 > 0 ? v[1] : -v[1]);
    if (c2 != 0) r = r + (c2 > 0 ? v[2] : -v[2]);
    return r;
}

template <class C>
LBM_DEV void collide_bgk_guo(const C* f, C alpha, const C a[3], C tau, C* out) {
    constexpr int Q = 19;
    C rho, m[3];
    moments<Q>(f, rho, m);
    const C s = alpha * C(0.5) + C(1.0);
    C u[3], F[3];
#pragma unroll
    for (int d = 0; d < 3; d++) u[d] = (m[d] / rho + C(0.5) * a[d]) / s;
#pragma unroll
    for (int d = 0; d < 3; d++) F[d] = rho * (a[d] - alpha * u[d]);
    const C uf = u[0] * F[0] + u[1] * F[1] + u[2] * F[2];
    const C usq = u[0] * u[0] + u[1] * u[1] + u[2] * u[2];
    const C k = C(1.0) - C(0.5) / tau;
    static_for<0, Q>([&](auto I) {
        constexpr int q = decltype(I)::value;
        constexpr double w = weight<Q>(q);
        const C cu = st_dot(cvec<Q>(q, 0), cvec<Q>(q, 1), cvec<Q>(q, 2), u);
        const C cf = st_dot(cvec<Q>(q, 0), cvec<Q>(q, 1), cvec<Q>(q, 2), F);
        const C eq = rho * C(w) * (C(1.0) + cu * C(3.0) + cu * cu * C(4.5) - usq * C(1.5));
        const C src = k * C(w) * (cf * C(3.0) - uf * C(3.0) + cu * cf * C(9.0));
        out[q] = f[q] - (f[q] - eq) / tau + src;
    });
}

// Takes the macroscopic quantities from brinkman_macro, equal in value to the
// reference's own recomputation (ports.rs moments; solver.rs:222-226).
template <class C>
LBM_DEV void collide_bgk_guo_vjp(const Macro<19, C>& mc, C alpha, const C a[3], C tau, const C* g, C* fb,
                                 C& alpha_bar) {
    constexpr int Q = 19;
    const C rho = mc.rho;
    const C m[3] = {mc.j[0], mc.j[1], mc.j[2]};
    const C s = mc.den;
    const C u[3] = {mc.u[0], mc.u[1], mc.u[2]};
    const C F[3] = {mc.F[0], mc.F[1], mc.F[2]};
    const C usq = u[0] * u[0] + u[1] * u[1] + u[2] * u[2];
    const C k = C(1.0) - C(0.5) / tau;
    C rho_bar = C(0), uf_bar = C(0), usq_bar = C(0);
    C u_bar[3] = {C(0), C(0), C(0)}, F_bar[3] = {C(0), C(0), C(0)};
    static_for<0, Q>([&](auto I) {
        constexpr int q = decltype(I)::value;
        constexpr double w = weight<Q>(q);
        const C gq = g[q];
        fb[q] = C(0);
        if (gq != C(0)) {
            const C cu = st_dot(cvec<Q>(q, 0), cvec<Q>(q, 1), cvec<Q>(q, 2), u);
            const C cf = st_dot(cvec<Q>(q, 0), cvec<Q>(q, 1), cvec<Q>(q, 2), F);
            const C poly = C(1.0) + C(3.0) * cu + C(4.5) * cu * cu - C(1.5) * usq;
            fb[q] = fb[q] + gq * (C(1.0) - C(1.0) / tau);
            const C eq_bar = gq / tau;
            const C h = gq * C(w) * k;
            const C cf_bar = h * (C(3.0) + C(9.0) * cu);
            C cu_bar = h * C(9.0) * cf;
            uf_bar -= C(3.0) * h;
            rho_bar += eq_bar * C(w) * poly;
            const C e = eq_bar * rho * C(w);
            cu_bar += e * (C(3.0) + C(9.0) * cu);
            usq_bar -= C(1.5) * e;
#pragma unroll
            for (int d = 0; d < 3; d++) {
                const C cd = C(cvec<Q>(q, d));
                u_bar[d] += cu_bar * cd;
                F_bar[d] += cf_bar * cd;
            }
        }
    });
#pragma unroll
    for (int d = 0; d < 3; d++) {
        u_bar[d] += usq_bar * C(2.0) * u[d] + uf_bar * F[d];
        F_bar[d] += uf_bar * u[d];
    }
    C ab = C(0);
#pragma unroll
    for (int d = 0; d < 3; d++) {
        rho_bar += F_bar[d] * (a[d] - alpha * u[d]);
        ab -= F_bar[d] * rho * u[d];
        u_bar[d] -= F_bar[d] * rho * alpha;
    }
    C m_bar[3], s_bar = C(0);
#pragma unroll
    for (int d = 0; d < 3; d++) {
        m_bar[d] = u_bar[d] / (s * rho);
        rho_bar -= u_bar[d] * m[d] / (s * rho * rho);
        s_bar -= u_bar[d] * u[d] / s;
    }
    ab += C(0.5) * s_bar;
    static_for<0, Q>([&](auto I) {
        constexpr int q = decltype(I)::value;
        C v = fb[q] + rho_bar;
#pragma unroll
        for (int d = 0; d < 3; d++) v += m_bar[d] * C(cvec<Q>(q, d));
        fb[q] = v;
    });
    alpha_bar = ab;
}

// ---------------------------------------------------------------------------------
// MRT (collision.rs:326-453): y = f - eq + src/2; m = M y; f' = f - M^-1 R m + src.
// `mrt` holds rows[Q*Q] (row-major, rows[k*Q+i]), inv_norms[Q], rates[Q] in C.
// ---------------------------------------------------------------------------------
template <int Q, class C>
LBM_DEV void equilibrium(C rho, const C u[3], C* eq) {
    const C usq = u[0] * u[0] + u[1] * u[1] + u[2] * u[2];
    const C base = C(1) - usq * C(1.5);
    static_for<0, Q>([&](auto I) {
        constexpr int i = decltype(I)::value;
        const C cu = dotc<Q, i>(u);
        eq[i] = rho * C(weight<Q>(i)) * (base + cu * C(3.0) + cu * cu * C(4.5));
    });
}

template <int Q, class C>
LBM_DEV void source_term(const C u[3], const C F[3], C* src) {
    const C uf = u[0] * F[0] + u[1] * F[1] + u[2] * F[2];
    static_for<0, Q>([&](auto I) {
        constexpr int i = decltype(I)::value;
        const C cu = dotc<Q, i>(u);
        const C cf = dotc<Q, i>(F);
        src[i] = (cf * C(3.0) + cu * cf * C(9.0) - uf * C(3.0)) * C(weight<Q>(i));
    });
}

template <int Q, class C>
LBM_DEV void collide_mrt(const C* f, C rho, const C u[3], const C F[3], const C* __restrict__ mrt, C* out) {
    const C* rows = mrt;
    const C* norms = mrt + Q * Q;  // norms (the reference divides by them)
    const C* rates = mrt + Q * Q + Q;
    C eq[Q], src[Q], y[Q];
    equilibrium<Q>(rho, u, eq);
    source_term<Q>(u, F, src);
#pragma unroll
    for (int i = 0; i < Q; i++) y[i] = f[i] - eq[i] + src[i] * C(0.5);
    C sc[Q];
#pragma unroll
    for (int k = 0; k < Q; k++) {
        C acc = C(0);
#pragma unroll
        for (int i = 0; i < Q; i++) acc += y[i] * rows[k * Q + i];
        sc[k] = acc * rates[k] / norms[k];
    }
    static_for<0, Q>([&](auto I) {
        constexpr int i = decltype(I)::value;
        C acc = C(0);
#pragma unroll
        for (int k = 0; k < Q; k++) acc += sc[k] * rows[k * Q + i];
        out[i] = f[i] - acc * C(weight<Q>(i)) + src[i];
    });
}

template <int Q, class C>
LBM_DEV void equilibrium_vjp(C rho, const C u[3], const C* g, C& rho_bar, C u_bar[3]) {
    const C usq = u[0] * u[0] + u[1] * u[1] + u[2] * u[2];
    const C base = C(1) - usq * C(1.5);
    C usq_bar = C(0);
    static_for<0, Q>([&](auto I) {
        constexpr int i = decltype(I)::value;
        const C cu = dotc<Q, i>(u);
        const C poly = base + cu * C(3.0) + cu * cu * C(4.5);
        const C gw = g[i] * C(weight<Q>(i));
        rho_bar += gw * poly;
        const C e = gw * rho;
        const C cu_bar = e * (cu * C(9.0) + C(3.0));
        if constexpr (cvec<Q>(i, 0) != 0) u_bar[0] += cu_bar * C(cvec<Q>(i, 0));
        if constexpr (cvec<Q>(i, 1) != 0) u_bar[1] += cu_bar * C(cvec<Q>(i, 1));
        if constexpr (cvec<Q>(i, 2) != 0) u_bar[2] += cu_bar * C(cvec<Q>(i, 2));
        usq_bar -= e * C(1.5);
    });
#pragma unroll
    for (int d = 0; d < 3; d++) u_bar[d] += usq_bar * u[d] * C(2.0);
}

// Generic collide_vjp MRT arm (collision.rs:597-633) with relax_transpose (:403-418).
template <int Q, class C>
LBM_DEV void collide_mrt_vjp(C rho, const C u[3], const C F[3], const C* __restrict__ mrt, const C* g, C* fb,
                             C& rho_bar, C u_bar[3], C F_bar[3]) {
    const C* rows = mrt;
    const C* norms = mrt + Q * Q;
    const C* rates = mrt + Q * Q + Q;
    C sc[Q];
#pragma unroll
    for (int k = 0; k < Q; k++) {
        C acc = C(0);
#pragma unroll
        for (int i = 0; i < Q; i++) acc += (g[i] * C(weight<Q>(i))) * rows[k * Q + i];
        sc[k] = acc * rates[k] / norms[k];
    }
    C rt[Q];
#pragma unroll
    for (int i = 0; i < Q; i++) {
        C acc = C(0);
#pragma unroll
        for (int k = 0; k < Q; k++) acc += sc[k] * rows[k * Q + i];
        rt[i] = acc;
    }
    C src_bar[Q];
#pragma unroll
    for (int i = 0; i < Q; i++) {
        fb[i] = g[i] - rt[i];
        src_bar[i] = g[i] - rt[i] * C(0.5);
    }
    C rb = C(0), ub[3] = {C(0), C(0), C(0)};
    equilibrium_vjp<Q>(rho, u, rt, rb, ub);
    C Fb[3] = {C(0), C(0), C(0)}, uf_bar = C(0);
    static_for<0, Q>([&](auto I) {
        constexpr int i = decltype(I)::value;
        const C h = src_bar[i] * C(weight<Q>(i));
        const C cu = dotc<Q, i>(u);
        const C cf = dotc<Q, i>(F);
        const C cf_bar = h * (cu * C(9.0) + C(3.0));
        const C cu_bar = h * cf * C(9.0);
        uf_bar -= h * C(3.0);
#pragma unroll
        for (int d = 0; d < 3; d++) {
            if (cvec<Q>(i, d) != 0) {
                Fb[d] += cf_bar * C(cvec<Q>(i, d));
                ub[d] += cu_bar * C(cvec<Q>(i, d));
            }
        }
    });
#pragma unroll
    for (int d = 0; d < 3; d++) {
        u_bar[d] = ub[d] + uf_bar * F[d];
        F_bar[d] = Fb[d] + uf_bar * u[d];
    }
    rho_bar = rb;
}


// ---------------------------------------------------------------------------------
// Shifted formulation for single precision: the cell holds d_q = f_q - w_q and every
// quantity near its equilibrium value is carried as a deviation, so FP32 rounding acts
// on the deviations rather than on f ~ w (rho - 1 = sum d, j = sum c d, and
// eq_q - w_q = w_q (drho + rho (3 cu + 4.5 cu^2 - 1.5 u^2))).  Algebraically identical
// to the reference operators; used only where C = float.
// ---------------------------------------------------------------------------------
template <int Q, class C>
LBM_DEV void brinkman_macro_shifted(const C* d, C alpha, const C a[3], Macro<Q, C>& m) {
    C dr;
    moments<Q>(d, dr, m.j);  // sum of deviations, sum of c d (sum_q c_q w_q = 0)
    m.drho = dr;
    m.rho = C(1) + dr;
    m.den = C(1) + alpha * C(0.5);
#pragma unroll
    for (int k = 0; k < 3; k++) {
        m.sv[k] = qdiv(m.j[k], m.rho) + a[k] * C(0.5);
        m.u[k] = qdiv(m.sv[k], m.den);
    }
#pragma unroll
    for (int k = 0; k < 3; k++) m.F[k] = m.rho * (a[k] - alpha * m.u[k]);
}

template <int Q, class C>
LBM_DEV void collide_pairs_shifted(const C* d, const Macro<Q, C>& m, C omega, C om, C* out) {
    // Pair (i, o = opp(i)) with s = d_i + d_o, t = d_i - d_o, cu = c_i.u, cf = c_i.F:
    //   aa = kp s+ - omega (s/2 - e+) = w (omega A - 3 kp uf) + cu (4.5 omega w rho cu + 9 kp w cf) - omega/2 s
    //   bb = km s- - om (t/2 - e-)    = 3 w om rho cu + 3 w km cf - om/2 t
    // with A = drho - 1.5 rho |u|^2 (so e+ - w = w (A + 4.5 rho cu^2)); out_i = d_i + aa + bb,
    // out_o = d_o + aa - bb.  The per-weight coefficients are formed once per cell.
    const C* u = m.u;
    const C* F = m.F;
    const C usq = u[0] * u[0] + u[1] * u[1] + u[2] * u[2];
    const C uf = u[0] * F[0] + u[1] * F[1] + u[2] * F[2];
    const C kp = C(1) - omega * C(0.5);
    const C km = C(1) - om * C(0.5);
    const C A = m.drho - C(1.5) * m.rho * usq;
    const C hp = omega * C(0.5), hm = om * C(0.5);
    const C p0 = omega * A - C(3) * kp * uf;  // times w
    const C pc = C(4.5) * omega * m.rho;      // times w
    const C pf = C(9) * kp;                   // times w
    const C mc = C(3) * om * m.rho;           // times w
    const C mf = C(3) * km;                   // times w
    constexpr double w0 = weight<Q>(0);
    out[0] = d[0] - omega * d[0] + C(w0) * p0;
    static_for<1, Q>([&](auto I) {
        constexpr int i = decltype(I)::value;
        constexpr int o = opp<Q>(i);
        if constexpr (o > i) {
            const C w = C(weight<Q>(i));
            const C cu = dotc<Q, i>(u);
            const C cf = dotc<Q, i>(F);
            const C sp = d[i] + d[o];
            const C tm = d[i] - d[o];
            const C aa = w * (p0 + cu * (pc * cu + pf * cf)) - hp * sp;
            const C bb = w * (mc * cu + mf * cf) - hm * tm;
            out[i] = d[i] + aa + bb;
            out[o] = d[o] + aa - bb;
        }
    });
}

// eq_q - w_q for all q.
template <int Q, class C>
LBM_DEV void equilibrium_shifted(const Macro<Q, C>& m, C* e) {
    const C* u = m.u;
    const C usq = u[0] * u[0] + u[1] * u[1] + u[2] * u[2];
    static_for<0, Q>([&](auto I) {
        constexpr int i = decltype(I)::value;
        const C cu = dotc<Q, i>(u);
        e[i] = C(weight<Q>(i)) * (m.drho + m.rho * (cu * C(3.0) + cu * cu * C(4.5) - usq * C(1.5)));
    });
}

template <int Q, class C>
LBM_DEV void collide_mrt_shifted(const C* d, const Macro<Q, C>& m, const C* __restrict__ mrt, C* out) {
    const C* rows = mrt;
    const C* norms = mrt + Q * Q;
    const C* rates = mrt + Q * Q + Q;
    C e[Q], src[Q], y[Q];
    equilibrium_shifted<Q>(m, e);
    source_term<Q>(m.u, m.F, src);
#pragma unroll
    for (int i = 0; i < Q; i++) y[i] = d[i] - e[i] + src[i] * C(0.5);
    C sc[Q];
#pragma unroll
    for (int k = 0; k < Q; k++) {
        C acc = C(0);
#pragma unroll
        for (int i = 0; i < Q; i++) acc += y[i] * rows[k * Q + i];
        sc[k] = acc * rates[k] / norms[k];
    }
    static_for<0, Q>([&](auto I) {
        constexpr int i = decltype(I)::value;
        C acc = C(0);
#pragma unroll
        for (int k = 0; k < Q; k++) acc += sc[k] * rows[k * Q + i];
        out[i] = d[i] - acc * C(weight<Q>(i)) + src[i];
    });
}

// Stationary BGK-Guo (solver.rs:298-322) on deviations.
template <class C>
LBM_DEV void collide_bgk_guo_shifted(const C* d, const Macro<19, C>& m, C tau, C* out) {
    constexpr int Q = 19;
    const C* u = m.u;
    const C* F = m.F;
    const C uf = u[0] * F[0] + u[1] * F[1] + u[2] * F[2];
    const C k = C(1.0) - C(0.5) / tau;
    C e[Q];
    equilibrium_shifted<Q>(m, e);
    static_for<0, Q>([&](auto I) {
        constexpr int q = decltype(I)::value;
        constexpr double w = weight<Q>(q);
        const C cu = dotc<Q, q>(u);
        const C cf = dotc<Q, q>(F);
        const C src = k * C(w) * (cf * C(3.0) - uf * C(3.0) + cu * cf * C(9.0));
        out[q] = d[q] - (d[q] - e[q]) / tau + src;
    });
}

}  // namespace lbm
