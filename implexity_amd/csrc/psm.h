// Occupancy coupling of the partially saturated method (moving/psm.rs): a cell with
// occupancy d > 0 and solid momentum m (u_s = m / d) leaves the fluid collision as
//   Psm:               post = f + (1 - B)(fluid - f) + Omega_s,  Omega_s = B (f_o - eq_o - f) + s_eq
//   PsmSuperposition:  post = fluid + Omega_s,                   Omega_s = s_eq - B eq
// with s_eq = rho w (B + (B/d)(3 c.m + 4.5 (c.m)^2 / d - 1.5 |m|^2 / d)), B = beta(d, tau - 1/2)
// (Noble & Torczynski's weighting with a smooth saturation above 1 - kappa), and the
// fluid collision evaluated with u = j/rho + a/2 and force rho a (psm.rs:228-266, the
// Brinkman macroscopic step at alpha = 0).  In the shifted storages every population
// term is a deviation from w: s_eq - B w = w (B (rho - 1) + rho (B/d)(...)), and the
// w terms of the bounce-back part cancel.
// Constants of a handle with the coupling: pk[0] = magic, pk[1] = kappa, pk[2] = law
// (0 Psm, 1 PsmSuperposition; the kernels take the law as a template argument).
// Included inside namespace lbm by kernels.h.
#pragma once

template <class C>
LBM_DEV C psm_saturate(C d, C kappa) {
    const C lo = C(1) - kappa;
    if (d <= lo) return d;
    if (d >= C(1) + kappa) return C(1);
    const C t = (d - lo) / (C(2.0) * kappa);
    const C t2 = t * t;
    const C h = t - t2 * t + t2 * t2 * C(0.5);
    return lo + h * (C(2.0) * kappa);
}

template <class C>
LBM_DEV C psm_saturate_derivative(C d, C kappa) {
    const C lo = C(1) - kappa;
    if (d <= lo) return C(1);
    if (d >= C(1) + kappa) return C(0);
    const C t = (d - lo) / (C(2.0) * kappa);
    const C s = C(1) - t;
    return s * s * (t * C(2.0) + C(1));
}

// B / d and its partials in d and t (psm.rs:123-140).
template <class C>
LBM_DEV void psm_beta(C d, C t, C kappa, C& bod, C& bd, C& bt) {
    if (d <= C(1) - kappa) {
        const C den = C(1) - d + t;
        const C den2 = den * den;
        bod = t / den;
        bd = t / den2;
        bt = (C(1) - d) / den2;
    } else {
        const C e = psm_saturate(d, kappa);
        const C de = psm_saturate_derivative(d, kappa);
        const C den = C(1) - e + t;
        const C bf = e * t / den;
        const C den2 = den * den;
        const C bf_e = t * (t + C(1)) / den2;
        const C bf_t = e * (C(1) - e) / den2;
        bod = bf / d;
        bd = bf_e * de / d - bf / (d * d);
        bt = bf_t / d;
    }
}

// x / d: a division in double precision (the reference's operation order), a product
// with the reciprocal in single precision (the shifted storages).
template <class C, bool SH>
LBM_DEV C psm_over(C x, C d, C inv_d) {
    if constexpr (SH)
        return x * inv_d;
    else
        return x / d;
}

// The solid-equilibrium term of direction i: s_eq (f storage) or s_eq - B w (shifted).
template <int Q, int i, class C, bool SH>
LBM_DEV C psm_solid_eq(const Macro<Q, C>& m, C bb, C bod, C cm, C m2, C d, C inv_d) {
    constexpr double w = weight<Q>(i);
    if constexpr (SH)
        return C(w) * (bb * m.drho + m.rho * (bod * cm * C(3.0) + bod * cm * cm * C(4.5) * inv_d -
                                              bod * m2 * C(1.5) * inv_d));
    else
        return m.rho * C(w) * (bb + bod * cm * C(3.0) + bod * cm * cm * C(4.5) / d - bod * m2 * C(1.5) / d);
}

// eq_k (f storage) or eq_k - w_k (shifted) of the cell's macroscopic state.
template <int Q, int k, class C, bool SH>
LBM_DEV C psm_eq(const Macro<Q, C>& m, C usq) {
    const C cu = dotc<Q, k>(m.u);
    if constexpr (SH)
        return C(weight<Q>(k)) * (m.drho + m.rho * (cu * C(3.0) + cu * cu * C(4.5) - usq * C(1.5)));
    else
        return m.rho * C(weight<Q>(k)) * ((C(1) - usq * C(1.5)) + cu * C(3.0) + cu * cu * C(4.5));
}

// out holds the fluid collision of v on entry and the coupled post-collision values on
// return (same representation as v); m is the alpha = 0 macroscopic state of v; dp gets
// the momentum exchanged with the solid, sum_i Omega_s,i c_i.  SUP selects
// PsmSuperposition.  The equilibrium is formed per direction, not held.
template <int Q, class C, bool SH, bool SUP>
LBM_DEV void psm_blend(const C* v, const Macro<Q, C>& m, C d, const C ms[3], const C* pk, C omega, C* out, C dp[3]) {
    const C t = C(1) / omega - C(0.5);
    C bod, bd, bt;
    psm_beta<C>(d, t, pk[1], bod, bd, bt);
    const C bb = bod * d;
    const C m2 = ms[0] * ms[0] + ms[1] * ms[1] + ms[2] * ms[2];
    const C usq = m.u[0] * m.u[0] + m.u[1] * m.u[1] + m.u[2] * m.u[2];
    const C inv_d = SH ? C(1) / d : C(0);
    dp[0] = dp[1] = dp[2] = C(0);
    auto exch = [&](auto I, C os) {
        constexpr int i = decltype(I)::value;
        if constexpr (cvec<Q>(i, 0) != 0) dp[0] += os * C(cvec<Q>(i, 0));
        if constexpr (cvec<Q>(i, 1) != 0) dp[1] += os * C(cvec<Q>(i, 1));
        if constexpr (cvec<Q>(i, 2) != 0) dp[2] += os * C(cvec<Q>(i, 2));
    };
    if constexpr (!SH) {
        // the reference's equilibrium (lattice.rs), bit for bit in double precision
        C eq[Q];
        equilibrium<Q>(m.rho, m.u, eq);
        static_for<0, Q>([&](auto I) {
            constexpr int i = decltype(I)::value;
            constexpr int o = opp<Q>(i);
            const C cm = dotc<Q, i>(ms);
            const C se = psm_solid_eq<Q, i, C, SH>(m, bb, bod, cm, m2, d, inv_d);
            const C os = SUP ? se - bb * eq[i] : bb * (v[o] - eq[o] - v[i]) + se;
            out[i] = SUP ? out[i] + os : v[i] + (C(1) - bb) * (out[i] - v[i]) + os;
            exch(I, os);
        });
    } else {
        static_for<0, Q>([&](auto I) {
            constexpr int i = decltype(I)::value;
            constexpr int o = opp<Q>(i);
            const C cm = dotc<Q, i>(ms);
            const C se = psm_solid_eq<Q, i, C, SH>(m, bb, bod, cm, m2, d, inv_d);
            const C os = SUP ? se - bb * psm_eq<Q, i, C, SH>(m, usq)
                             : bb * (v[o] - psm_eq<Q, o, C, SH>(m, usq) - v[i]) + se;
            out[i] = SUP ? out[i] + os : v[i] + (C(1) - bb) * (out[i] - v[i]) + os;
            exch(I, os);
        });
    }
}

// The momentum exchanged with the solid alone, dp = sum_i Omega_s,i c_i of psm_blend (the
// same expressions in the same order, so the same bits), without the blended values.
template <int Q, class C, bool SH, bool SUP>
LBM_DEV void psm_exchange(const C* v, const Macro<Q, C>& m, C d, const C ms[3], const C* pk, C omega, C dp[3]) {
    const C t = C(1) / omega - C(0.5);
    C bod, bd, bt;
    psm_beta<C>(d, t, pk[1], bod, bd, bt);
    const C bb = bod * d;
    const C m2 = ms[0] * ms[0] + ms[1] * ms[1] + ms[2] * ms[2];
    const C usq = m.u[0] * m.u[0] + m.u[1] * m.u[1] + m.u[2] * m.u[2];
    const C inv_d = SH ? C(1) / d : C(0);
    dp[0] = dp[1] = dp[2] = C(0);
    C eq[Q];
    if constexpr (!SH) equilibrium<Q>(m.rho, m.u, eq);
    static_for<0, Q>([&](auto I) {
        constexpr int i = decltype(I)::value;
        constexpr int o = opp<Q>(i);
        const C cm = dotc<Q, i>(ms);
        const C se = psm_solid_eq<Q, i, C, SH>(m, bb, bod, cm, m2, d, inv_d);
        C os;
        if constexpr (!SH)
            os = SUP ? se - bb * eq[i] : bb * (v[o] - eq[o] - v[i]) + se;
        else
            os = SUP ? se - bb * psm_eq<Q, i, C, SH>(m, usq) : bb * (v[o] - psm_eq<Q, o, C, SH>(m, usq) - v[i]) + se;
        if constexpr (cvec<Q>(i, 0) != 0) dp[0] += os * C(cvec<Q>(i, 0));
        if constexpr (cvec<Q>(i, 1) != 0) dp[1] += os * C(cvec<Q>(i, 1));
        if constexpr (cvec<Q>(i, 2) != 0) dp[2] += os * C(cvec<Q>(i, 2));
    });
}

// Subcycled occupancy (Field::accumulate, moving/field.rs:1499-1541): the impulse shares of
// the two endpoints, g0 += (-dp / d)(1 - theta), g1 += (-dp / d) theta, in G (6 x cells,
// SoA: g0 x, y, z then g1 x, y, z; FP64).  Each cell's entries belong to its own thread.
template <class C>
LBM_DEV void sub_accumulate(double* G, long long n, long long cell, const C dp[3], C d, double theta) {
    const C keep = C(1.0 - theta), th = C(theta);
#pragma unroll
    for (int e = 0; e < 3; e++) {
        const C gv = -dp[e] / d;
        G[e * n + cell] += double(gv * keep);
        G[(3 + e) * n + cell] += double(gv * th);
    }
}

// Transpose of sub_accumulate at one cell (reverse_substep, field.rs:2682-2690): with the
// blended cotangent gt = gbar0 (1 - theta) + gbar1 theta, the exchange's cotangent takes
// -gt / d and the occupancy's gains sum_e gt dp_e / d^2, returned in extra.
template <class C>
LBM_DEV void sub_cotangent(const double* GB, long long n, long long cell, const C dp[3], C d, double theta,
                           C dpb[3], C& extra) {
    extra = C(0);
#pragma unroll
    for (int e = 0; e < 3; e++) {
        const C gt = C(GB[e * n + cell]) * C(1.0 - theta) + C(GB[(3 + e) * n + cell]) * C(theta);
        dpb[e] -= gt / d;
        extra += gt * dp[e] / (d * d);
    }
}

// One direction's share of the equilibrium transpose (cell.h equilibrium_vjp) for the
// cotangent e of eq_k; usq_bar is folded into u_bar by psm_eq_vjp_end.
template <int Q, int k, class C>
LBM_DEV void psm_eq_vjp(const Macro<Q, C>& m, C usq, C e, C& rho_bar, C u_bar[3], C& usq_bar) {
    const C cu = dotc<Q, k>(m.u);
    const C gw = e * C(weight<Q>(k));
    rho_bar += gw * ((C(1) - usq * C(1.5)) + cu * C(3.0) + cu * cu * C(4.5));
    const C E = gw * m.rho;
    const C cu_bar = E * (cu * C(9.0) + C(3.0));
    if constexpr (cvec<Q>(k, 0) != 0) u_bar[0] += cu_bar * C(cvec<Q>(k, 0));
    if constexpr (cvec<Q>(k, 1) != 0) u_bar[1] += cu_bar * C(cvec<Q>(k, 1));
    if constexpr (cvec<Q>(k, 2) != 0) u_bar[2] += cu_bar * C(cvec<Q>(k, 2));
    usq_bar -= E * C(1.5);
}

// Transpose of psm_blend (psm.rs:331-415) given the cotangent g of post and dpb of the
// exchange dp.  On entry fb holds the transpose of the fluid collision applied to g;
// gfl = <g, fluid> (Psm only).
// Adds the blend's own terms to fb (the fluid part scaled by 1 - B for Psm, the
// bounce-back and equilibrium terms, and rho, u = j/rho + a/2 carried to the
// populations) and returns the cotangents of d and m.
template <int Q, class C, bool SH, bool SUP>
LBM_DEV void psm_blend_vjp(const C* v, const Macro<Q, C>& m, C d, const C ms[3], const C* pk, C omega, const C* g,
                           C gfl, const C dpb[3], C* fb, C& d_bar, C m_bar[3]) {
    const C t = C(1) / omega - C(0.5);
    C bod, bd, bt;
    psm_beta<C>(d, t, pk[1], bod, bd, bt);
    const C bb = bod * d;
    const C m2 = ms[0] * ms[0] + ms[1] * ms[1] + ms[2] * ms[2];
    const C usq = m.u[0] * m.u[0] + m.u[1] * m.u[1] + m.u[2] * m.u[2];
    // rho w - eq of the bounce-back and superposition terms: w (rho - 1) - (eq - w) when shifted
    const C rw = SH ? m.drho : m.rho;
    const C inv_d = SH ? C(1) / d : C(0);
    C b_bar = C(0), beta_bar = C(0), wsum = C(0), rho_bar = C(0), usq_bar = C(0);
    C u_bar[3] = {C(0), C(0), C(0)};
    d_bar = C(0);
#pragma unroll
    for (int a = 0; a < 3; a++) m_bar[a] = C(0);
    if constexpr (!SUP) {
        // post = f + (1 - B)(fluid - f) + ...: the fluid transpose scales by 1 - B, f gets B g,
        // and B gets <g, f - fluid>
        C gf = C(0);
#pragma unroll
        for (int q = 0; q < Q; q++) {
            fb[q] = fb[q] * (C(1) - bb) + g[q] * bb;
            gf += g[q] * v[q];
        }
        b_bar = gf - gfl;
    }
    static_for<0, Q>([&](auto I) {
        constexpr int i = decltype(I)::value;
        constexpr int o = opp<Q>(i);
        constexpr double w = weight<Q>(i);
        C h = g[i];  // cotangent of Omega_s,i: through post and through the exchange dp
        if constexpr (cvec<Q>(i, 0) != 0) h += dpb[0] * C(cvec<Q>(i, 0));
        if constexpr (cvec<Q>(i, 1) != 0) h += dpb[1] * C(cvec<Q>(i, 1));
        if constexpr (cvec<Q>(i, 2) != 0) h += dpb[2] * C(cvec<Q>(i, 2));
        if constexpr (SUP) {
            b_bar += h * (rw * C(w) - psm_eq<Q, i, C, SH>(m, usq));
            psm_eq_vjp<Q, i, C>(m, usq, -h * bb, rho_bar, u_bar, usq_bar);
        } else {
            b_bar += h * (v[o] - psm_eq<Q, o, C, SH>(m, usq) - v[i] + rw * C(w));
            fb[o] += h * bb;
            fb[i] -= h * bb;
            psm_eq_vjp<Q, o, C>(m, usq, -h * bb, rho_bar, u_bar, usq_bar);
        }
        const C cm = dotc<Q, i>(ms);
        const C hw = h * C(w);
        rho_bar += hw * (bb + bod * cm * C(3.0) + psm_over<C, SH>(bod * cm * cm * C(4.5), d, inv_d) -
                         psm_over<C, SH>(bod * m2 * C(1.5), d, inv_d));
        const C wr = hw * m.rho;
        beta_bar += wr * (cm * C(3.0) + psm_over<C, SH>(cm * cm * C(4.5), d, inv_d) -
                          psm_over<C, SH>(m2 * C(1.5), d, inv_d));
        const C cm_bar = wr * bod * (psm_over<C, SH>(cm * C(9.0), d, inv_d) + C(3.0));
        if constexpr (cvec<Q>(i, 0) != 0) m_bar[0] += cm_bar * C(cvec<Q>(i, 0));
        if constexpr (cvec<Q>(i, 1) != 0) m_bar[1] += cm_bar * C(cvec<Q>(i, 1));
        if constexpr (cvec<Q>(i, 2) != 0) m_bar[2] += cm_bar * C(cvec<Q>(i, 2));
        wsum += wr;
        d_bar += psm_over<C, SH>(psm_over<C, SH>(wr * bod * (m2 * C(1.5) - cm * cm * C(4.5)), d, inv_d), d, inv_d);
    });
#pragma unroll
    for (int a = 0; a < 3; a++) {
        m_bar[a] -= psm_over<C, SH>(wsum * bod * ms[a] * C(3.0), d, inv_d);
        u_bar[a] += usq_bar * m.u[a] * C(2.0);
    }
    beta_bar += b_bar * d;
    d_bar += b_bar * bod + beta_bar * bd;
    // rho and u = j/rho + a/2 carried to the populations
    C jb[3];
#pragma unroll
    for (int a = 0; a < 3; a++) {
        jb[a] = u_bar[a] / m.rho;
        rho_bar -= u_bar[a] * m.j[a] / (m.rho * m.rho);
    }
    static_for<0, Q>([&](auto I) {
        constexpr int q = decltype(I)::value;
        C x = fb[q] + rho_bar;
        if constexpr (cvec<Q>(q, 0) != 0) x += jb[0] * C(cvec<Q>(q, 0));
        if constexpr (cvec<Q>(q, 1) != 0) x += jb[1] * C(cvec<Q>(q, 1));
        if constexpr (cvec<Q>(q, 2) != 0) x += jb[2] * C(cvec<Q>(q, 2));
        fb[q] = x;
    });
}

// ---------------------------------------------------------------------------------
// Brinkman law with a moving solid (psm.rs:143-160, 213-226, 301-330, 478-492): the
// occupancy sets alpha = c_alpha e / ((q + 1) - e) and gamma (= alpha / d), the
// macroscopic step is u = (j/rho + a/2 + gamma m/2) / (1 + alpha/2) with the force
// rho (a + z), z = gamma m - alpha u, and the fluid collision runs on them.  At m = 0 it
// is the stationary Brinkman operator with alpha = resistance(d).
// Constants: pk[3] = c_alpha, pk[4] = drag shape q.
// ---------------------------------------------------------------------------------
template <class C>
LBM_DEV void psm_resistance(C d, C c_alpha, C q, C kappa, C& alpha, C& gamma, C part[4]) {
    const C e = psm_saturate(d, kappa);
    const C de = psm_saturate_derivative(d, kappa);
    const C den = C(1) * (q + C(1)) - e;
    alpha = c_alpha * e / den;
    const C a_e = c_alpha * (q + C(1)) / (den * den);
    const C a_c = e / den;
    part[0] = a_e * de;
    part[1] = a_c;
    if (d <= C(1) - kappa) {
        gamma = c_alpha / den;
        part[2] = c_alpha / (den * den);
        part[3] = C(1) / den;
    } else {
        gamma = alpha / d;
        part[2] = a_e * de / d - alpha / (d * d);
        part[3] = a_c / d;
    }
}

template <int Q, class C, bool SH>
LBM_DEV void brinkman_moving_macro(const C* v, C alpha, C gamma, const C a[3], const C ms[3], Macro<Q, C>& m, C z[3]) {
    if constexpr (SH) {
        C dr;
        moments<Q>(v, dr, m.j);
        m.drho = dr;
        m.rho = C(1) + dr;
    } else {
        moments<Q>(v, m.rho, m.j);
        m.drho = m.rho - C(1);
    }
    m.den = C(1) + alpha * C(0.5);
#pragma unroll
    for (int k = 0; k < 3; k++) {
        m.sv[k] = qdiv(m.j[k], m.rho) + a[k] * C(0.5) + gamma * ms[k] * C(0.5);
        m.u[k] = qdiv(m.sv[k], m.den);
    }
#pragma unroll
    for (int k = 0; k < 3; k++) {
        z[k] = gamma * ms[k] - alpha * m.u[k];
        m.F[k] = m.rho * (a[k] + z[k]);
    }
}

// Cotangents of (f, alpha, gamma, m) from those of the collision's (rho, u, F) and of the
// exchange dp (dpb); fb holds the collision's direct cotangent on entry and the total on
// return.
template <int Q, class C>
LBM_DEV void brinkman_moving_chain(const Macro<Q, C>& m, C alpha, C gamma, const C a[3], const C ms[3], const C z[3],
                                   const C dpb[3], C cb_rho, const C cb_u[3], const C cb_F[3], C* fb, C& alpha_bar,
                                   C& gamma_bar, C m_bar[3]) {
    C rho_bar = cb_rho;
    C u_bar[3] = {cb_u[0], cb_u[1], cb_u[2]};
    alpha_bar = C(0);
    gamma_bar = C(0);
#pragma unroll
    for (int k = 0; k < 3; k++) {  // force rho (a + z) and exchange dp = rho z
        rho_bar += cb_F[k] * (a[k] + z[k]) + dpb[k] * z[k];
        const C zb = (cb_F[k] + dpb[k]) * m.rho;
        gamma_bar += zb * ms[k];
        m_bar[k] = zb * gamma;
        alpha_bar -= zb * m.u[k];
        u_bar[k] -= zb * alpha;
    }
    C jb[3];
#pragma unroll
    for (int k = 0; k < 3; k++) {
        const C s_bar = qdiv(u_bar[k], m.den);
        alpha_bar -= qdiv(u_bar[k] * m.sv[k], m.den * m.den) * C(0.5);
        jb[k] = qdiv(s_bar, m.rho);
        rho_bar -= qdiv(s_bar * m.j[k], m.rho * m.rho);
        gamma_bar += s_bar * ms[k] * C(0.5);
        m_bar[k] += s_bar * gamma * C(0.5);
    }
    static_for<0, Q>([&](auto I) {
        constexpr int q = decltype(I)::value;
        C x = fb[q] + rho_bar;
        if constexpr (cvec<Q>(q, 0) != 0) x += jb[0] * C(cvec<Q>(q, 0));
        if constexpr (cvec<Q>(q, 1) != 0) x += jb[1] * C(cvec<Q>(q, 1));
        if constexpr (cvec<Q>(q, 2) != 0) x += jb[2] * C(cvec<Q>(q, 2));
        fb[q] = x;
    });
}

// Cotangent of the exchange dp from a solid-force stage cost (0 for other objectives).
template <class C>
LBM_DEV C psm_dp_bar(const ObjArgs& ob, int k) {
    return ob.kind == OBJ_SOLID_FORCE ? C(ob.weight * ob.vw[k]) : C(0);
}

// The occupancy d and solid momentum m a step collides with: the design value d0 and
// m0, blended with the second endpoint when one is set as d0 (1 - theta) + d1 theta
// (Field::solid_at, field.rs:1266-1275).
template <class C>
LBM_DEV void psm_input(C d0, const void* M0, const void* D1, const void* M1, double theta, long long n,
                       long long cell, C& d, C ms[3]) {
    const C* M = static_cast<const C*>(M0);
    if (D1) {
        const C th = C(theta), keep = C(1.0 - theta);
        const C* Db = static_cast<const C*>(D1);
        const C* Mb = static_cast<const C*>(M1);
        d = d0 * keep + Db[cell] * th;
#pragma unroll
        for (int e = 0; e < 3; e++) ms[e] = M[e * n + cell] * keep + Mb[e * n + cell] * th;
    } else {
        d = d0;
#pragma unroll
        for (int e = 0; e < 3; e++) ms[e] = M[e * n + cell];
    }
}

// Adds the solid-momentum cotangent mb to the endpoints' buffers and returns the share
// of the occupancy cotangent db that goes where alpha_bar goes (the first endpoint's);
// the second endpoint's shares are theta db and theta mb.
template <class C, class Args>
LBM_DEV C psm_accumulate(const Args& A, long long n, long long cell, C db, const C mb[3]) {
    if (A.d1) {
        const C th = C(A.theta), keep = C(1.0 - A.theta);
        A.d1_bar[cell] += double(db * th);
#pragma unroll
        for (int e = 0; e < 3; e++) {
            A.sm_bar[e * n + cell] += double(mb[e] * keep);
            A.sm1_bar[e * n + cell] += double(mb[e] * th);
        }
        return db * keep;
    }
#pragma unroll
    for (int e = 0; e < 3; e++) A.sm_bar[e * n + cell] += double(mb[e]);
    return db;
}
