// WALE closure on the BGK/TRT collision (moving/wale.rs), restated for the stationary
// step: the rate of a fluid cell is 1 / (tau0 + 3 nu_t) with
//   nu_t = C^2 A^(3/2) / (B^(5/2) + A^(5/4) + floor),
// A = Sd:Sd (Sd the traceless symmetric part of g^2), B = S:S (S the symmetric part of g),
// g the velocity gradient of u = j / rho of the state the step collides
// (field.rs:1294-1320), by central differences, one-sided where a neighbour along the
// axis is outside a closed face or solid, zero where both are (Stencil::new / gradient).
// The gradient needs the velocities of the neighbours at the same time level, which the
// in-place step overwrites, so a pre-pass (k_wale_u) writes u and rho of every cell; the
// transpose scatters the gradient cotangent of each cell back to the velocities its
// stencil used (Stencil::gradient_transpose, as a gather in k_wale_transpose).
// Constants of a handle with the closure: wk[0] = magic, wk[1] = C^2, wk[2] = floor.
// Included inside namespace lbm by kernels.h, after the storage geometry.
#pragma once

// Storage index of the neighbour of (i, j, k) one cell along +-axis that a stencil may
// use (inside, or across a periodic axis, and not solid); -1 if there is none.
template <bool MASK>
LBM_DEV long long wale_nb(const Geom& g, const uint32_t* mask, int i, int j, int k, int axis, int dir) {
    int c[3] = {i, j, k};
    const int n[3] = {g.nx, g.ny, g.nz};
    const int per[3] = {g.px, g.py, g.pz};
    c[axis] += dir;
    if (c[axis] < 0 || c[axis] >= n[axis]) {
        if (!per[axis]) return -1;
        c[axis] = c[axis] < 0 ? c[axis] + n[axis] : c[axis] - n[axis];
    }
    const long long y = ((long long)c[0] * g.ny + c[1]) * g.nz + c[2];
    if constexpr (MASK) {
        if (mask[y] & SOLID_BIT) return -1;
    }
    return y;
}

// Difference kind of a fluid cell along axis b: 0 zero, 1 central, 2 forward, 3 backward.
template <bool MASK>
LBM_DEV int wale_diff(const Geom& g, const uint32_t* mask, int i, int j, int k, int b, long long& pp,
                      long long& mm) {
    const int n[3] = {g.nx, g.ny, g.nz};
    pp = mm = -1;
    if (n[b] == 1) return 0;
    pp = wale_nb<MASK>(g, mask, i, j, k, b, 1);
    mm = wale_nb<MASK>(g, mask, i, j, k, b, -1);
    if (pp >= 0 && mm >= 0) return 1;
    if (pp >= 0) return 2;
    if (mm >= 0) return 3;
    return 0;
}

// G[a][b] = d u_a / d x_b from the pre-pass velocities U[a * n + cell].
template <class C, bool MASK>
LBM_DEV void wale_gradient(const Geom& g, const uint32_t* mask, const C* U, long long n, int i, int j, int k,
                           long long cell, C G[3][3]) {
#pragma unroll
    for (int b = 0; b < 3; b++) {
        long long pp, mm;
        const int kind = wale_diff<MASK>(g, mask, i, j, k, b, pp, mm);
        // absent neighbours read the cell itself, so that every load stays inside U
        const long long ip = pp >= 0 ? pp : cell, im = mm >= 0 ? mm : cell;
#pragma unroll
        for (int a = 0; a < 3; a++) {
            const C* ua = U + (long long)a * n;
            const C up = ua[ip], um = ua[im], uc = ua[cell];
            G[a][b] = kind == 1 ? (up - um) * C(0.5) : (kind == 2 ? up - uc : (kind == 3 ? uc - um : C(0)));
        }
    }
}

template <class C>
LBM_DEV void wale_invariants(const C G[3][3], C sd[3][3], C s[3][3], C& ai, C& bi) {
    C g2[3][3];
#pragma unroll
    for (int a = 0; a < 3; a++)
#pragma unroll
        for (int b = 0; b < 3; b++) {
            C v = C(0);
#pragma unroll
            for (int c = 0; c < 3; c++) v += G[a][c] * G[c][b];
            g2[a][b] = v;
        }
    const C tr = (g2[0][0] + g2[1][1] + g2[2][2]) / C(3.0);
    ai = C(0);
    bi = C(0);
#pragma unroll
    for (int a = 0; a < 3; a++)
#pragma unroll
        for (int b = 0; b < 3; b++) {
            sd[a][b] = (g2[a][b] + g2[b][a]) * C(0.5) - (a == b ? tr : C(0));
            s[a][b] = (G[a][b] + G[b][a]) * C(0.5);
            ai += sd[a][b] * sd[a][b];
            bi += s[a][b] * s[a][b];
        }
}

template <class C>
LBM_DEV C wale_omega(const C G[3][3], C tau0, const C* wk) {
    C sd[3][3], s[3][3], a, b;
    wale_invariants<C>(G, sd, s, a, b);
    const C num = pow(a, C(1.5));
    const C den = pow(b, C(2.5)) + pow(a, C(1.25)) + wk[2];
    const C nu = num / den * wk[1];
    return C(1) / (tau0 + nu * C(3.0));
}

// wale.rs omega_vjp: cotangent Gb of the gradient given the cotangent w of the rate.
template <class C>
LBM_DEV void wale_omega_vjp(const C G[3][3], C tau0, const C* wk, C w, C Gb[3][3]) {
    C sd[3][3], s[3][3], a, b;
    wale_invariants<C>(G, sd, s, a, b);
    const C c2 = wk[1];
    const C num = pow(a, C(1.5));
    const C den = pow(b, C(2.5)) + pow(a, C(1.25)) + wk[2];
    const C nu = num / den * c2;
    const C tau = tau0 + nu * C(3.0);
    const C tau_bar = -w / (tau * tau);
    const C nu_bar = tau_bar * C(3.0);
    const C a_bar = nu_bar * c2 * (pow(a, C(0.5)) * C(1.5) / den - num * pow(a, C(0.25)) * C(1.25) / (den * den));
    const C b_bar = -nu_bar * c2 * num * pow(b, C(1.5)) * C(2.5) / (den * den);
#pragma unroll
    for (int x = 0; x < 3; x++)
#pragma unroll
        for (int y = 0; y < 3; y++) {
            C da = C(0);
#pragma unroll
            for (int k = 0; k < 3; k++) da += sd[x][k] * G[y][k] + G[k][x] * sd[k][y];
            Gb[x][y] = a_bar * da * C(2.0) + b_bar * s[x][y] * C(2.0);
        }
}
