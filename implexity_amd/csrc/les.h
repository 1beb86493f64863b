// Smagorinsky large-eddy closure on the BGK/TRT collision (moving/psm.rs:161-197) and
// its transpose (psm.rs:415-473), in the operation order of the reference so that the
// double-precision path reproduces it bit for bit:
//   Pi_xy = sum_i (f_i - eq_i) c_ix c_iy,  |Pi|^2 = sum over all nine (x, y) of Pi_xy^2,
//   P = sqrt(|Pi|^2 + floor^2) - floor,   tau0 = 1 / omega,
//   tau_e = (tau0 + sqrt(tau0^2 + P k / rho)) / 2 with k = 18 sqrt(2) C^2,
//   omega_eff = 1 / tau_e,  and for TRT omega- = odd_rate(omega_eff, magic).
// Constants of a handle with the closure: les[0] = magic, les[1] = k, les[2] = floor.
#pragma once
#include "cell.h"

namespace lbm {

// Non-equilibrium part n_i = f_i - eq_i: from f (double precision) or, in the shifted
// form, n_i = d_i - w_i (drho + rho (3 cu + 4.5 cu^2 - 1.5 |u|^2)).
template <int Q, class C, bool SH>
LBM_DEV void nonequilibrium(const C* v, const Macro<Q, C>& m, C* n) {
    if constexpr (SH) {
        const C usq = m.u[0] * m.u[0] + m.u[1] * m.u[1] + m.u[2] * m.u[2];
        static_for<0, Q>([&](auto I) {
            constexpr int i = decltype(I)::value;
            const C cu = dotc<Q, i>(m.u);
            n[i] = v[i] - C(weight<Q>(i)) * (m.drho + m.rho * (cu * C(3.0) + cu * cu * C(4.5) - usq * C(1.5)));
        });
    } else {
        C eq[Q];
        equilibrium<Q>(m.rho, m.u, eq);
#pragma unroll
        for (int i = 0; i < Q; i++) n[i] = v[i] - eq[i];
    }
}

// The six distinct components (xx, xy, xz, yy, yz, zz) of Pi, each summed over i in
// order; returns |Pi|^2 summed over the nine (x, y) in row order.
template <int Q, class C>
LBM_DEV C les_flux(const C* n, C pi[6]) {
#pragma unroll
    for (int t = 0; t < 6; t++) pi[t] = C(0);
    static_for<0, Q>([&](auto I) {
        constexpr int i = decltype(I)::value;
        constexpr int cx = cvec<Q>(i, 0), cy = cvec<Q>(i, 1), cz = cvec<Q>(i, 2);
        if constexpr (cx != 0) pi[0] += n[i] * C(cx * cx);
        if constexpr (cx != 0 && cy != 0) pi[1] += n[i] * C(cx * cy);
        if constexpr (cx != 0 && cz != 0) pi[2] += n[i] * C(cx * cz);
        if constexpr (cy != 0) pi[3] += n[i] * C(cy * cy);
        if constexpr (cy != 0 && cz != 0) pi[4] += n[i] * C(cy * cz);
        if constexpr (cz != 0) pi[5] += n[i] * C(cz * cz);
    });
    C q = C(0);
    q += pi[0] * pi[0];
    q += pi[1] * pi[1];
    q += pi[2] * pi[2];
    q += pi[1] * pi[1];
    q += pi[3] * pi[3];
    q += pi[4] * pi[4];
    q += pi[2] * pi[2];
    q += pi[4] * pi[4];
    q += pi[5] * pi[5];
    return q;
}

// The closure of one cell, evaluated once and shared by the forward rate, the omega
// cotangents and the transpose.
template <int Q, class C>
struct LesEval {
    C n[Q];    // non-equilibrium populations
    C pi[6];   // Pi_xx, Pi_xy, Pi_xz, Pi_yy, Pi_yz, Pi_zz
    C root, pn, tau0, r, oe;  // sqrt(|Pi|^2 + floor^2), P, 1/omega, sqrt(tau0^2 + P k / rho), omega_eff
};

template <int Q, class C, bool SH>
LBM_DEV void les_eval(const C* v, const Macro<Q, C>& m, C omega, const C* les, LesEval<Q, C>& e) {
    nonequilibrium<Q, C, SH>(v, m, e.n);
    const C qn = les_flux<Q, C>(e.n, e.pi);
    const C delta = les[2];
    e.root = sqrt(qn + delta * delta);
    e.pn = e.root - delta;
    e.tau0 = C(1) / omega;
    e.r = sqrt(e.tau0 * e.tau0 + e.pn * les[1] / m.rho);
    e.oe = C(1) / ((e.tau0 + e.r) * C(0.5));
}

// omega- = 1 / (1/2 + magic / (1 / omega - 1/2)) and its derivative (collision.rs:302-313).
template <class C>
LBM_DEV C odd_rate_of(C omega, C magic) {
    const C tp = C(1) / omega - C(0.5);
    return C(1) / (C(0.5) + magic / tp);
}
template <class C>
LBM_DEV C odd_rate_derivative_of(C omega, C magic) {
    const C tau = C(1) / omega;
    const C tp = tau - C(0.5);
    const C tm = C(0.5) + magic / tp;
    return -(tau * tau * magic) / (tm * tm * tp * tp);
}

// Cotangents of omega+ and omega- of collide_pairs (collision.rs:513-569):
//   ob  = -g_0 ((f_0 - eq_0) + s_0 / 2) - sum_pairs (g_i + g_o) ((fp - eqp) + sp / 2),
//   omb = -sum_pairs (g_i - g_o) ((fm - eqm) + sm / 2).
template <int Q, class C>
LBM_DEV void pairs_omega_bar(const C* n, const Macro<Q, C>& m, const C* g, C& ob, C& omb) {
    const C* u = m.u;
    const C* F = m.F;
    const C uf = u[0] * F[0] + u[1] * F[1] + u[2] * F[2];
    constexpr double w0 = weight<Q>(0);
    C o = -g[0] * (n[0] + (-uf * C(3.0 * w0)) * C(0.5));
    C om = C(0);
    static_for<1, Q>([&](auto I) {
        constexpr int i = decltype(I)::value;
        constexpr int op = opp<Q>(i);
        if constexpr (op > i) {
            constexpr double w = weight<Q>(i);
            const C cu = dotc<Q, i>(u);
            const C cf = dotc<Q, i>(F);
            const C sp = (cu * cf * C(9.0) - uf * C(3.0)) * C(w);
            const C sm = cf * C(3.0 * w);
            // fp - eqp = (n_i + n_o) / 2 and fm - eqm = (n_i - n_o) / 2 (eq = eqp +- eqm)
            o -= (g[i] + g[op]) * ((n[i] + n[op]) * C(0.5) + sp * C(0.5));
            om -= (g[i] - g[op]) * ((n[i] - n[op]) * C(0.5) + sm * C(0.5));
        }
    });
    ob = o;
    omb = om;
}

// Transpose of the closure given the cotangent of omega_eff: adds to fb (through the
// non-equilibrium part) and to the equilibrium's inputs rho_bar, u_bar.
template <int Q, class C>
LBM_DEV void smagorinsky_vjp(const LesEval<Q, C>& e, const Macro<Q, C>& m, const C* les, C oe_bar, C* fb,
                             C& rho_bar, C u_bar[3]) {
    const C k = les[1];
    const C tau_e = (e.tau0 + e.r) * C(0.5);
    const C tau_e_bar = -oe_bar / (tau_e * tau_e);
    const C r_bar = tau_e_bar * C(0.5);
    const C inner_bar = r_bar / (e.r * C(2.0));
    const C p_bar = inner_bar * k / m.rho;
    rho_bar -= inner_bar * e.pn * k / (m.rho * m.rho);
    const C qn_bar2 = p_bar / (e.root * C(2.0)) * C(2.0);
    C nb[Q];
    static_for<0, Q>([&](auto I) {
        constexpr int i = decltype(I)::value;
        constexpr int cx = cvec<Q>(i, 0), cy = cvec<Q>(i, 1), cz = cvec<Q>(i, 2);
        C acc = C(0);
        if constexpr (cx != 0) acc += e.pi[0] * C(cx * cx);
        if constexpr (cx != 0 && cy != 0) acc += C(2) * e.pi[1] * C(cx * cy);
        if constexpr (cx != 0 && cz != 0) acc += C(2) * e.pi[2] * C(cx * cz);
        if constexpr (cy != 0) acc += e.pi[3] * C(cy * cy);
        if constexpr (cy != 0 && cz != 0) acc += C(2) * e.pi[4] * C(cy * cz);
        if constexpr (cz != 0) acc += e.pi[5] * C(cz * cz);
        nb[i] = acc * qn_bar2;
        fb[i] += nb[i];
        nb[i] = -nb[i];
    });
    equilibrium_vjp<Q>(m.rho, m.u, nb, rho_bar, u_bar);
}

}  // namespace lbm
