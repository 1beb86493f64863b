// D3Q27 cumulant collision (moving/collision.rs:636-752) and its transposed Jacobian.
//
// Forward: chimera transform of f into central moments about u along z, y, x; the
// six second-order moments are relaxed (shear with omega, trace towards rho with the
// bulk rate); order 0 is rho, order 1 is F/2; orders >= 3 are rebuilt from Wick
// pairings of sigma = M2/rho (higher cumulants relaxed to zero; odd orders vanish);
// then the inverse transform along x, y, z.  The operation order follows the
// reference so the double-precision build reproduces it bit for bit.
//
// Reverse (our own): the reference differentiates this with 35 forward-mode tangents
// per cell (collision.rs:755-788).  Here the transpose is analytic: the inverse
// transform is a tensor product of 3x3 maps T(c, e; u) per axis, so its transpose is
// the same tensor product transposed, and its u-derivative replaces one factor by
// dT/du; only six central moments of the forward transform survive the relaxation,
// with d kappa_ab / d f_i = (c_ia - u_a)(c_ib - u_b) and
// d kappa_ab / d u_d = -(delta_ad kappa_b + delta_bd kappa_a).
#pragma once
#include "lattice.h"

namespace lbm {

// Cube index of a D3Q27 direction: (cx+1)*9 + (cy+1)*3 + (cz+1).
template <int q>
LBM_HD constexpr int cube_of() {
    return (cvec<27>(q, 0) + 1) * 9 + (cvec<27>(q, 1) + 1) * 3 + (cvec<27>(q, 2) + 1);
}

template <class C>
__device__ __forceinline__ void chimera_forward(C* cube, const C u[3]) {
    // axis 2 (stride 1), then axis 1 (stride 3), then axis 0 (stride 9)
#pragma unroll
    for (int ai = 0; ai < 3; ai++) {
        const int axis = 2 - ai;
        const int s = axis == 2 ? 1 : (axis == 1 ? 3 : 9);
        const C ua = u[axis];
#pragma unroll
        for (int base = 0; base < 27; base++) {
            if ((base / s) % 3 != 0) continue;
            const C fm = cube[base], f0 = cube[base + s], fp = cube[base + 2 * s];
            const C k0 = fm + f0 + fp;
            const C diff = fp - fm;
            const C k1 = diff - ua * k0;
            const C k2 = (fp + fm) - ua * diff * C(2.0) + ua * ua * k0;
            cube[base] = k0;
            cube[base + s] = k1;
            cube[base + 2 * s] = k2;
        }
    }
}

template <class C>
__device__ __forceinline__ void chimera_inverse(C* post, const C u[3]) {
#pragma unroll
    for (int axis = 0; axis < 3; axis++) {
        const int s = axis == 2 ? 1 : (axis == 1 ? 3 : 9);
        const C ua = u[axis];
#pragma unroll
        for (int base = 0; base < 27; base++) {
            if ((base / s) % 3 != 0) continue;
            const C k0 = post[base], k1 = post[base + s], k2 = post[base + 2 * s];
            const C m1 = k1 + ua * k0;
            const C m2 = k2 + ua * k1 * C(2.0) + ua * ua * k0;
            post[base] = (m2 - m1) * C(0.5);
            post[base + s] = k0 - m2;
            post[base + 2 * s] = (m2 + m1) * C(0.5);
        }
    }
}

// Post-collision central moments (collision.rs:677-729).
template <class C>
struct CumPost {
    C pxx, pyy, pzz, pxy, pxz, pyz;
    C s00, s11, s22, s01, s02, s12;
};

template <class C>
__device__ __forceinline__ void cumulant_relax(const C* cube, C rho, C omega, C bulk, CumPost<C>& p) {
    const C kxx = cube[18], kyy = cube[6], kzz = cube[2];
    const C one_minus = C(1) - omega;
    const C dxy = (kxx - kyy) * one_minus;
    const C dxz = (kxx - kzz) * one_minus;
    C trace = kxx + kyy + kzz;
    trace = trace + (rho - trace) * bulk;
    p.pxx = (trace + dxy + dxz) / C(3.0);
    p.pyy = (trace - dxy * C(2.0) + dxz) / C(3.0);
    p.pzz = (trace + dxy - dxz * C(2.0)) / C(3.0);
    p.pxy = cube[12] * one_minus;
    p.pxz = cube[10] * one_minus;
    p.pyz = cube[4] * one_minus;
    p.s00 = p.pxx / rho;
    p.s01 = p.pxy / rho;
    p.s02 = p.pxz / rho;
    p.s11 = p.pyy / rho;
    p.s12 = p.pyz / rho;
    p.s22 = p.pzz / rho;
}

// Wick (Gaussian) closures of the even orders >= 4, in the pairing order the
// reference's matchings() produces (collision.rs:216-239).
template <class C>
__device__ __forceinline__ C wick4(C a1, C a2, C b1, C b2, C c1, C c2) {
    C acc = C(0);
    acc += C(1) * a1 * a2;
    acc += C(1) * b1 * b2;
    acc += C(1) * c1 * c2;
    return acc;
}

template <class C>
__device__ __forceinline__ void cumulant_fill(C* post, const CumPost<C>& p, C rho, const C F[3]) {
#pragma unroll
    for (int k = 0; k < 27; k++) post[k] = C(0) * rho;  // odd and unpaired orders: 0 * rho
    post[0] = rho;
    post[9] = F[0] * C(0.5);
    post[3] = F[1] * C(0.5);
    post[1] = F[2] * C(0.5);
    post[18] = p.pxx;
    post[6] = p.pyy;
    post[2] = p.pzz;
    post[12] = p.pxy;
    post[10] = p.pxz;
    post[4] = p.pyz;
    // order 4
    post[24] = wick4(p.s00, p.s11, p.s01, p.s01, p.s01, p.s01) * rho;  // (2,2,0)
    post[20] = wick4(p.s00, p.s22, p.s02, p.s02, p.s02, p.s02) * rho;  // (2,0,2)
    post[8] = wick4(p.s11, p.s22, p.s12, p.s12, p.s12, p.s12) * rho;   // (0,2,2)
    post[22] = wick4(p.s00, p.s12, p.s01, p.s02, p.s02, p.s01) * rho;  // (2,1,1)
    post[16] = wick4(p.s01, p.s12, p.s01, p.s12, p.s02, p.s11) * rho;  // (1,2,1)
    post[14] = wick4(p.s01, p.s22, p.s02, p.s12, p.s02, p.s12) * rho;  // (1,1,2)
    // order 6: (2,2,2), 15 pairings in the reference's order
    C acc = C(0);
    acc += C(1) * p.s00 * p.s11 * p.s22;
    acc += C(1) * p.s00 * p.s12 * p.s12;
    acc += C(1) * p.s00 * p.s12 * p.s12;
    acc += C(1) * p.s01 * p.s01 * p.s22;
    acc += C(1) * p.s01 * p.s02 * p.s12;
    acc += C(1) * p.s01 * p.s02 * p.s12;
    acc += C(1) * p.s01 * p.s01 * p.s22;
    acc += C(1) * p.s01 * p.s02 * p.s12;
    acc += C(1) * p.s01 * p.s02 * p.s12;
    acc += C(1) * p.s02 * p.s01 * p.s12;
    acc += C(1) * p.s02 * p.s01 * p.s12;
    acc += C(1) * p.s02 * p.s02 * p.s11;
    acc += C(1) * p.s02 * p.s01 * p.s12;
    acc += C(1) * p.s02 * p.s01 * p.s12;
    acc += C(1) * p.s02 * p.s02 * p.s11;
    post[26] = acc * rho;
}

template <class C>
__device__ __forceinline__ void collide_cumulant(const C* f, C rho, const C u[3], const C F[3], C omega, C bulk,
                                                 C* out) {
    C cube[27];
    static_for<0, 27>([&](auto I) {
        constexpr int q = decltype(I)::value;
        cube[cube_of<q>()] = f[q];
    });
    chimera_forward(cube, u);
    CumPost<C> p;
    cumulant_relax(cube, rho, omega, bulk, p);
    C post[27];
    cumulant_fill(post, p, rho, F);
    chimera_inverse(post, u);
    static_for<0, 27>([&](auto I) {
        constexpr int q = decltype(I)::value;
        out[q] = post[cube_of<q>()];
    });
}

// The same collision on deviations d = f - w (single-precision shifted storage).  The
// chimera transform is linear in f for a fixed u, and the D3Q27 weights are the product
// w1(cx) w1(cy) w1(cz) with w1(0) = 2/3, w1(+-1) = 1/6, whose central moments about u are
// per axis (1, -u, 1/3 + u^2).  So the transform of d gives kappa(f) - kappa(w); the
// post-collision moments are formed as deviations Delta = post - kappa(w) in which no two
// O(1) quantities are subtracted (rho = 1 + drho, M2 = I/3 + E, the Wick products
// expanded about I/3); and the inverse transform of Delta is d' = f' - w.  In exact
// arithmetic this equals collide_cumulant; in FP32 it keeps the deviations' significant
// bits instead of rounding at ulp(w).
template <class C>
__device__ __forceinline__ C mw(int e, C u) {  // central moment e of w1 about u
    return e == 0 ? C(1) : (e == 1 ? -u : C(1.0 / 3.0) + u * u);
}

template <class C>
__device__ __forceinline__ void collide_cumulant_shifted(const C* d, C drho, C rho, const C u[3], const C F[3],
                                                         C omega, C bulk, C* out) {
    C cube[27];
    static_for<0, 27>([&](auto I) {
        constexpr int q = decltype(I)::value;
        cube[cube_of<q>()] = d[q];
    });
    chimera_forward(cube, u);  // kappa(d) = kappa(f) - kappa(w)
    const C u0 = u[0], u1 = u[1], u2 = u[2];
    const C q0 = u0 * u0, q1 = u1 * u1, q2 = u2 * u2;
    // second order of f as deviations from I/3: kappa_aa(f) - 1/3 = u_a^2 + kappa_aa(d)
    const C dxx = q0 + cube[18], dyy = q1 + cube[6], dzz = q2 + cube[2];
    const C one_minus = C(1) - omega;
    const C sxy = (dxx - dyy) * one_minus;
    const C sxz = (dxx - dzz) * one_minus;
    C dtr = dxx + dyy + dzz;        // trace - 1
    dtr = dtr + (drho - dtr) * bulk;  // relaxed towards rho
    const C ex = (dtr + sxy + sxz) * C(1.0 / 3.0);  // post M2 - I/3
    const C ey = (dtr - sxy * C(2.0) + sxz) * C(1.0 / 3.0);
    const C ez = (dtr + sxy - sxz * C(2.0)) * C(1.0 / 3.0);
    const C pxy = (u0 * u1 + cube[12]) * one_minus;
    const C pxz = (u0 * u2 + cube[10]) * one_minus;
    const C pyz = (u1 * u2 + cube[4]) * one_minus;
    const C irho = C(1) / rho;
    const C r1 = -drho * irho;                          // 1/rho - 1
    const C r2 = -drho * (C(2) + drho) * irho * irho;  // 1/rho^2 - 1
    C post[27];
    // odd orders 3 and 5 of the post state vanish: Delta = -kappa(w)
#pragma unroll
    for (int k = 0; k < 27; k++) {
        const int a = k / 9, b = (k / 3) % 3, c = k % 3;
        post[k] = -(mw(a, u0) * mw(b, u1) * mw(c, u2));
    }
    post[0] = drho;
    post[9] = F[0] * C(0.5) + u0;
    post[3] = F[1] * C(0.5) + u1;
    post[1] = F[2] * C(0.5) + u2;
    post[18] = ex - q0;
    post[6] = ey - q1;
    post[2] = ez - q2;
    post[12] = pxy - u0 * u1;
    post[10] = pxz - u0 * u2;
    post[4] = pyz - u1 * u2;
    const C third = C(1.0 / 3.0), ninth = C(1.0 / 9.0);
    // order 4, (a,a,b,b): W = M_aa M_bb + 2 M_ab^2 = 1/9 + A; kappa(w) = 1/9 + B
    auto pair4 = [&](C ea, C eb, C pab, C qa, C qb) {
        const C A = (ea + eb) * third + ea * eb + C(2) * pab * pab;
        const C B = (qa + qb) * third + qa * qb;
        return (A - B) + r1 * (ninth + A);
    };
    post[24] = pair4(ex, ey, pxy, q0, q1);  // (2,2,0)
    post[20] = pair4(ex, ez, pxz, q0, q2);  // (2,0,2)
    post[8] = pair4(ey, ez, pyz, q1, q2);   // (0,2,2)
    // order 4, (a,a,b,c): W = M_aa M_bc + 2 M_ab M_ac; kappa(w) = (1/3 + u_a^2) u_b u_c
    auto cross4 = [&](C ea, C pbc, C pab, C pac, C qa, C ub, C uc) {
        const C W = (third + ea) * pbc + C(2) * pab * pac;
        return (W - (third + qa) * ub * uc) + r1 * W;
    };
    post[22] = cross4(ex, pyz, pxy, pxz, q0, u1, u2);  // (2,1,1)
    post[16] = cross4(ey, pxz, pxy, pyz, q1, u0, u2);  // (1,2,1)
    post[14] = cross4(ez, pxy, pxz, pyz, q2, u0, u1);  // (1,1,2)
    // order 6: W = Mxx Myy Mzz + 2 (Mxx Myz^2 + Myy Mxz^2 + Mzz Mxy^2) + 8 Mxy Mxz Myz
    {
        const C A6 = (ex + ey + ez) * ninth + (ex * ey + ex * ez + ey * ez) * third + ex * ey * ez +
                     C(2) * ((third + ex) * pyz * pyz + (third + ey) * pxz * pxz + (third + ez) * pxy * pxy) +
                     C(8) * pxy * pxz * pyz;
        const C B6 = (q0 + q1 + q2) * ninth + (q0 * q1 + q0 * q2 + q1 * q2) * third + q0 * q1 * q2;
        post[26] = (A6 - B6) + r2 * (C(1.0 / 27.0) + A6);
    }
    chimera_inverse(post, u);  // d' = f' - w
    static_for<0, 27>([&](auto I) {
        constexpr int q = decltype(I)::value;
        out[q] = post[cube_of<q>()];
    });
}

// 1D inverse-transform kernel T(c, e; u): raw value at velocity c in {-1,0,1} (index
// c+1) produced by central moment e in {0,1,2}, and its derivative in u.
template <class C>
__device__ __forceinline__ C tker(int c, int e, C u) {
    if (e == 2) return c == 0 ? C(-1) : C(0.5);
    if (e == 1) return c == 0 ? C(-2) * u : (c < 0 ? u - C(0.5) : u + C(0.5));
    return c == 0 ? C(1) - u * u : (c < 0 ? (u * u - u) * C(0.5) : (u * u + u) * C(0.5));
}
template <class C>
__device__ __forceinline__ C dtker(int c, int e, C u) {
    if (e == 2) return C(0);
    if (e == 1) return c == 0 ? C(-2) : C(1);
    return c == 0 ? C(-2) * u : (c < 0 ? u - C(0.5) : u + C(0.5));
}

// out[.., e, ..] = sum_c K(c, e) in[.., c, ..] along one axis (stride s): the
// transpose of the inverse transform's 1D step (D = true uses dT/du).
template <class C, bool D>
__device__ __forceinline__ void tmap_t(const C* in, C* out, int s, C u) {
#pragma unroll
    for (int base = 0; base < 27; base++) {
        if ((base / s) % 3 != 0) continue;
        const C gm = in[base], g0 = in[base + s], gp = in[base + 2 * s];
#pragma unroll
        for (int e = 0; e < 3; e++) {
            const C km = D ? dtker<C>(-1, e, u) : tker<C>(-1, e, u);
            const C k0 = D ? dtker<C>(0, e, u) : tker<C>(0, e, u);
            const C kp = D ? dtker<C>(1, e, u) : tker<C>(1, e, u);
            out[base + e * s] = km * gm + k0 * g0 + kp * gp;
        }
    }
}

// Cotangents of (f, rho, u, F) given the post-collision cotangent g.
template <class C>
__device__ __forceinline__ void collide_cumulant_vjp(const C* f, C rho, const C u[3], const C F[3], C omega,
                                                     C bulk, const C* g, C* fb, C& rho_bar, C u_bar[3],
                                                     C F_bar[3]) {
    (void)F;
    // Forward quantities.
    C cube[27];
    static_for<0, 27>([&](auto I) {
        constexpr int q = decltype(I)::value;
        cube[cube_of<q>()] = f[q];
    });
    chimera_forward(cube, u);
    CumPost<C> p;
    cumulant_relax(cube, rho, omega, bulk, p);
    C post[27];
    cumulant_fill(post, p, rho, F);

    // Transposed inverse transform: axes were applied x (s=9), y (s=3), z (s=1), so the
    // transpose applies z, y, x.  Cotangent cube G indexed by velocity.
    C G[27], H1[27], H2[27], PB[27], T1[27], T2[27];
    static_for<0, 27>([&](auto I) {
        constexpr int q = decltype(I)::value;
        G[cube_of<q>()] = g[q];
    });
    tmap_t<C, false>(G, H1, 1, u[2]);
    tmap_t<C, false>(H1, H2, 3, u[1]);
    tmap_t<C, false>(H2, PB, 9, u[0]);  // PB = d out / d post, transposed
    C ub[3];
    // u_x: dT along x on H2
    tmap_t<C, true>(H2, T1, 9, u[0]);
    {
        C s = C(0);
#pragma unroll
        for (int k = 0; k < 27; k++) s += T1[k] * post[k];
        ub[0] = s;
    }
    // u_y: dT along y on H1, then T along x
    tmap_t<C, true>(H1, T1, 3, u[1]);
    tmap_t<C, false>(T1, T2, 9, u[0]);
    {
        C s = C(0);
#pragma unroll
        for (int k = 0; k < 27; k++) s += T2[k] * post[k];
        ub[1] = s;
    }
    // u_z: dT along z on G, then T along y and x
    tmap_t<C, true>(G, T1, 1, u[2]);
    tmap_t<C, false>(T1, T2, 3, u[1]);
    tmap_t<C, false>(T2, T1, 9, u[0]);
    {
        C s = C(0);
#pragma unroll
        for (int k = 0; k < 27; k++) s += T1[k] * post[k];
        ub[2] = s;
    }

    // Relaxation and closure.
    C rb = PB[0];
    F_bar[0] = PB[9] * C(0.5);
    F_bar[1] = PB[3] * C(0.5);
    F_bar[2] = PB[1] * C(0.5);
    // d(sigma) from the Wick terms, and rho from post = W * rho.
    C sb00 = C(0), sb11 = C(0), sb22 = C(0), sb01 = C(0), sb02 = C(0), sb12 = C(0);
    {
        const C r = rho;
        // (2,2,0): s00 s11 + 2 s01^2
        C w = PB[24];
        rb += w * (p.s00 * p.s11 + C(2) * p.s01 * p.s01);
        sb00 += w * r * p.s11;
        sb11 += w * r * p.s00;
        sb01 += w * r * C(4) * p.s01;
        // (2,0,2): s00 s22 + 2 s02^2
        w = PB[20];
        rb += w * (p.s00 * p.s22 + C(2) * p.s02 * p.s02);
        sb00 += w * r * p.s22;
        sb22 += w * r * p.s00;
        sb02 += w * r * C(4) * p.s02;
        // (0,2,2): s11 s22 + 2 s12^2
        w = PB[8];
        rb += w * (p.s11 * p.s22 + C(2) * p.s12 * p.s12);
        sb11 += w * r * p.s22;
        sb22 += w * r * p.s11;
        sb12 += w * r * C(4) * p.s12;
        // (2,1,1): s00 s12 + 2 s01 s02
        w = PB[22];
        rb += w * (p.s00 * p.s12 + C(2) * p.s01 * p.s02);
        sb00 += w * r * p.s12;
        sb12 += w * r * p.s00;
        sb01 += w * r * C(2) * p.s02;
        sb02 += w * r * C(2) * p.s01;
        // (1,2,1): 2 s01 s12 + s02 s11
        w = PB[16];
        rb += w * (C(2) * p.s01 * p.s12 + p.s02 * p.s11);
        sb01 += w * r * C(2) * p.s12;
        sb12 += w * r * C(2) * p.s01;
        sb02 += w * r * p.s11;
        sb11 += w * r * p.s02;
        // (1,1,2): s01 s22 + 2 s02 s12
        w = PB[14];
        rb += w * (p.s01 * p.s22 + C(2) * p.s02 * p.s12);
        sb01 += w * r * p.s22;
        sb22 += w * r * p.s01;
        sb02 += w * r * C(2) * p.s12;
        sb12 += w * r * C(2) * p.s02;
        // (2,2,2): s00 s11 s22 + 2 s00 s12^2 + 2 s11 s02^2 + 2 s22 s01^2 + 8 s01 s02 s12
        w = PB[26];
        rb += w * (p.s00 * p.s11 * p.s22 + C(2) * p.s00 * p.s12 * p.s12 + C(2) * p.s11 * p.s02 * p.s02 +
                   C(2) * p.s22 * p.s01 * p.s01 + C(8) * p.s01 * p.s02 * p.s12);
        sb00 += w * r * (p.s11 * p.s22 + C(2) * p.s12 * p.s12);
        sb11 += w * r * (p.s00 * p.s22 + C(2) * p.s02 * p.s02);
        sb22 += w * r * (p.s00 * p.s11 + C(2) * p.s01 * p.s01);
        sb01 += w * r * (C(4) * p.s22 * p.s01 + C(8) * p.s02 * p.s12);
        sb02 += w * r * (C(4) * p.s11 * p.s02 + C(8) * p.s01 * p.s12);
        sb12 += w * r * (C(4) * p.s00 * p.s12 + C(8) * p.s01 * p.s02);
    }
    // sigma = post2 / rho
    const C ir = C(1) / rho;
    C pxx_b = PB[18] + sb00 * ir, pyy_b = PB[6] + sb11 * ir, pzz_b = PB[2] + sb22 * ir;
    C pxy_b = PB[12] + sb01 * ir, pxz_b = PB[10] + sb02 * ir, pyz_b = PB[4] + sb12 * ir;
    rb -= (sb00 * p.pxx + sb11 * p.pyy + sb22 * p.pzz + sb01 * p.pxy + sb02 * p.pxz + sb12 * p.pyz) * ir * ir;
    // second-order relaxation
    const C one_minus = C(1) - omega;
    const C tr_b = (pxx_b + pyy_b + pzz_b) / C(3);
    const C dxy_b = (pxx_b - C(2) * pyy_b + pzz_b) / C(3);
    const C dxz_b = (pxx_b + pyy_b - C(2) * pzz_b) / C(3);
    const C kxx_b = tr_b * (C(1) - bulk) + (dxy_b + dxz_b) * one_minus;
    const C kyy_b = tr_b * (C(1) - bulk) - dxy_b * one_minus;
    const C kzz_b = tr_b * (C(1) - bulk) - dxz_b * one_minus;
    rb += tr_b * bulk;
    const C kxy_b = pxy_b * one_minus, kxz_b = pxz_b * one_minus, kyz_b = pyz_b * one_minus;
    // central moments of the forward transform
    C kap[3] = {C(0), C(0), C(0)};  // first-order central moments sum f (c - u)
    static_for<0, 27>([&](auto I) {
        constexpr int q = decltype(I)::value;
        const C dx = C(cvec<27>(q, 0)) - u[0], dy = C(cvec<27>(q, 1)) - u[1], dz = C(cvec<27>(q, 2)) - u[2];
        kap[0] += f[q] * dx;
        kap[1] += f[q] * dy;
        kap[2] += f[q] * dz;
        fb[q] = kxx_b * dx * dx + kyy_b * dy * dy + kzz_b * dz * dz + kxy_b * dx * dy + kxz_b * dx * dz +
                kyz_b * dy * dz;
        // plus the transposed inverse transform: g -> post_bar -> (already folded above)
    });
    u_bar[0] = ub[0] - (C(2) * kxx_b * kap[0] + kxy_b * kap[1] + kxz_b * kap[2]);
    u_bar[1] = ub[1] - (C(2) * kyy_b * kap[1] + kxy_b * kap[0] + kyz_b * kap[2]);
    u_bar[2] = ub[2] - (C(2) * kzz_b * kap[2] + kxz_b * kap[0] + kyz_b * kap[1]);
    rho_bar = rb;
}

}  // namespace lbm
