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
// (0 Psm, 1 PsmSuperposition).  Included inside namespace lbm by kernels.h.
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

// The solid-equilibrium term of direction i: s_eq (f storage) or s_eq - B w (shifted).
template <int Q, int i, class C, bool SH>
LBM_DEV C psm_solid_eq(const Macro<Q, C>& m, C bb, C bod, C cm, C m2, C d) {
    constexpr double w = weight<Q>(i);
    if constexpr (SH)
        return C(w) * (bb * m.drho + m.rho * (bod * cm * C(3.0) + bod * cm * cm * C(4.5) / d - bod * m2 * C(1.5) / d));
    else
        return m.rho * C(w) * (bb + bod * cm * C(3.0) + bod * cm * cm * C(4.5) / d - bod * m2 * C(1.5) / d);
}

// out holds the fluid collision of v on entry and the coupled post-collision values on
// return (same representation as v); m is the alpha = 0 macroscopic state of v.
template <int Q, class C, bool SH>
LBM_DEV void psm_blend(const C* v, const Macro<Q, C>& m, C d, const C ms[3], const C* pk, C omega, C* out) {
    const C t = C(1) / omega - C(0.5);
    C bod, bd, bt;
    psm_beta<C>(d, t, pk[1], bod, bd, bt);
    const C bb = bod * d;
    C eq[Q];
    if constexpr (SH)
        equilibrium_shifted<Q>(m, eq);
    else
        equilibrium<Q>(m.rho, m.u, eq);
    const C m2 = ms[0] * ms[0] + ms[1] * ms[1] + ms[2] * ms[2];
    const bool sup = pk[2] != C(0);
    static_for<0, Q>([&](auto I) {
        constexpr int i = decltype(I)::value;
        constexpr int o = opp<Q>(i);
        const C cm = dotc<Q, i>(ms);
        const C se = psm_solid_eq<Q, i, C, SH>(m, bb, bod, cm, m2, d);
        const C os = sup ? se - bb * eq[i] : bb * (v[o] - eq[o] - v[i]) + se;
        out[i] = sup ? out[i] + os : v[i] + (C(1) - bb) * (out[i] - v[i]) + os;
    });
}

// Transpose of psm_blend given the cotangent g of post: the cotangent of the fluid
// collision's output (fluid_bar), the direct cotangent of v (fb, overwritten), the
// cotangents of rho and u = j/rho + a/2 from the equilibrium and solid-equilibrium terms
// (rho_bar, u_bar, overwritten), of d and of m (psm.rs:331-415).  fluid is the fluid
// collision of v; fluid_bar may be the same array.
template <int Q, class C, bool SH>
LBM_DEV void psm_blend_vjp(const C* v, const C* fluid, const Macro<Q, C>& m, C d, const C ms[3], const C* pk, C omega,
                           const C* g, C* fluid_bar, C* fb, C& rho_bar, C u_bar[3], C& d_bar, C m_bar[3]) {
    const C t = C(1) / omega - C(0.5);
    C bod, bd, bt;
    psm_beta<C>(d, t, pk[1], bod, bd, bt);
    const C bb = bod * d;
    C eq[Q], eq_bar[Q];
    if constexpr (SH)
        equilibrium_shifted<Q>(m, eq);
    else
        equilibrium<Q>(m.rho, m.u, eq);
    const C m2 = ms[0] * ms[0] + ms[1] * ms[1] + ms[2] * ms[2];
    const bool sup = pk[2] != C(0);
    // rho w - eq of the bounce-back and superposition terms: w (rho - 1) - (eq - w) when shifted
    const C rw = SH ? m.drho : m.rho;
    C b_bar = C(0), beta_bar = C(0), wsum = C(0);
    rho_bar = C(0);
    d_bar = C(0);
#pragma unroll
    for (int a = 0; a < 3; a++) {
        u_bar[a] = C(0);
        m_bar[a] = C(0);
    }
#pragma unroll
    for (int q = 0; q < Q; q++) {
        fb[q] = C(0);
        eq_bar[q] = C(0);
    }
    static_for<0, Q>([&](auto I) {
        constexpr int i = decltype(I)::value;
        constexpr int o = opp<Q>(i);
        constexpr double w = weight<Q>(i);
        const C h = g[i];
        const C fl = fluid[i];  // read before fluid_bar[i] is written: the two may alias
        if (sup) {
            fluid_bar[i] = g[i];
            b_bar += h * (rw * C(w) - eq[i]);
            eq_bar[i] -= h * bb;
        } else {
            fluid_bar[i] = g[i] * (C(1) - bb);
            fb[i] += g[i] * bb;
            b_bar += g[i] * (v[i] - fl);
            b_bar += h * (v[o] - eq[o] - v[i] + rw * C(w));
            fb[o] += h * bb;
            fb[i] -= h * bb;
            eq_bar[o] -= h * bb;
        }
        const C cm = dotc<Q, i>(ms);
        const C hw = h * C(w);
        rho_bar += hw * (bb + bod * cm * C(3.0) + bod * cm * cm * C(4.5) / d - bod * m2 * C(1.5) / d);
        const C wr = hw * m.rho;
        beta_bar += wr * (cm * C(3.0) + cm * cm * C(4.5) / d - m2 * C(1.5) / d);
        const C cm_bar = wr * bod * (cm * C(9.0) / d + C(3.0));
        if constexpr (cvec<Q>(i, 0) != 0) m_bar[0] += cm_bar * C(cvec<Q>(i, 0));
        if constexpr (cvec<Q>(i, 1) != 0) m_bar[1] += cm_bar * C(cvec<Q>(i, 1));
        if constexpr (cvec<Q>(i, 2) != 0) m_bar[2] += cm_bar * C(cvec<Q>(i, 2));
        wsum += wr;
        d_bar += wr * bod * (m2 * C(1.5) - cm * cm * C(4.5)) / (d * d);
    });
#pragma unroll
    for (int a = 0; a < 3; a++) m_bar[a] -= wsum * bod * ms[a] * C(3.0) / d;
    equilibrium_vjp<Q>(m.rho, m.u, eq_bar, rho_bar, u_bar);
    beta_bar += b_bar * d;
    d_bar += b_bar * bod + beta_bar * bd;
}

// fb[q] += rho_bar + c_q . (u_bar / rho) - u_bar . j / rho^2: the cotangents of rho and
// u = j/rho + a/2 carried to the populations.
template <int Q, class C>
LBM_DEV void psm_macro_chain(const Macro<Q, C>& m, C rho_bar, const C u_bar[3], C* fb) {
    C jb[3];
#pragma unroll
    for (int a = 0; a < 3; a++) {
        jb[a] = u_bar[a] / m.rho;
        rho_bar -= u_bar[a] * m.j[a] / (m.rho * m.rho);
    }
    static_for<0, Q>([&](auto I) {
        constexpr int q = decltype(I)::value;
        C v = fb[q] + rho_bar;
        if constexpr (cvec<Q>(q, 0) != 0) v += jb[0] * C(cvec<Q>(q, 0));
        if constexpr (cvec<Q>(q, 1) != 0) v += jb[1] * C(cvec<Q>(q, 1));
        if constexpr (cvec<Q>(q, 2) != 0) v += jb[2] * C(cvec<Q>(q, 2));
        fb[q] = v;
    });
}
