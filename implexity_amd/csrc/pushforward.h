// Push-forward of body points into the lattice occupancy (moving/pushforward.rs):
//   d[c] = sum_q a_q W(c - xi_q),  m[c] = sum_q a_q v_q W(c - xi_q),
// W the tensor product of per-axis weights of the cubic B-spline m4 or Peskin's 4-point
// kernel of half-width h (support 4h cells from floor(x) - 2h + 1; the closed-form cubic
// for h = 1), positions in lattice coordinates (cell c at coordinate c).  A periodic
// one-cell axis is planar (weight 1); at a symmetry face the support is folded back into
// the lattice (Stencils::new_with_kernel, :205-318).  The scatter
// runs as a gather: the points are sorted by the cell their stencil starts at (a stable
// radix sort), and every cell sums the points of the 64 h^3 start cells whose stencils
// cover it, in a fixed order, so d and m are the same bits run to run and agree with the
// reference's point-order sums to rounding.  The transpose (endpoint_vjp, :946-967) is a
// per-point gather.  Neither uses atomics.
#pragma once
#include <hip/hip_runtime.h>

namespace lbm {

constexpr int PF_MAX_H = 2;

struct PushArgs {
    int nx, ny, nz;
    int px, py, pz;
    int mirror[3];  // per axis: bit 0 symmetry face at the low end, bit 1 at the high end
    int h;        // kernel half-width in cells (1 or 2)
    int peskin;   // 0 cubic B-spline, 1 Peskin 4-point
    long long np; // points
    const double* xi;  // 3 per point
    const double* a;   // weight per point
    const double* v;   // 3 per point, or null
};

__device__ __forceinline__ double pf_m4(double s) {
    const double a = s < 0.0 ? -s : s;
    if (a < 1.0) return (4.0 - a * a * 6.0 + a * a * a * 3.0) / 6.0;
    if (a < 2.0) {
        const double t = 2.0 - a;
        return t * t * t / 6.0;
    }
    return 0.0;
}
__device__ __forceinline__ double pf_m4_derivative(double s) {
    const bool neg = s < 0.0;
    const double a = neg ? -s : s;
    double d = 0.0;
    if (a < 1.0) {
        d = -a * 2.0 + a * a * 1.5;
    } else if (a < 2.0) {
        const double t = 2.0 - a;
        d = -(t * t) * 0.5;
    }
    return neg ? -d : d;
}
__device__ __forceinline__ double pf_peskin4(double s) {
    const double r = s < 0.0 ? -s : s;
    if (r < 1.0) return (3.0 - r * 2.0 + sqrt(1.0 + r * 4.0 - r * r * 4.0)) / 8.0;
    if (r < 2.0) return (5.0 - r * 2.0 - sqrt(-7.0 + r * 12.0 - r * r * 4.0)) / 8.0;
    return 0.0;
}
__device__ __forceinline__ double pf_peskin4_derivative(double s) {
    const bool neg = s < 0.0;
    const double r = neg ? -s : s;
    double d = 0.0;
    if (r < 1.0)
        d = (-2.0 + (2.0 - r * 4.0) / sqrt(1.0 + r * 4.0 - r * r * 4.0)) / 8.0;
    else if (r < 2.0)
        d = (-2.0 - (6.0 - r * 4.0) / sqrt(-7.0 + r * 12.0 - r * r * 4.0)) / 8.0;
    return neg ? -d : d;
}

// Weights (and d/dx) of one coordinate over its 4h cells; returns the first cell.
__device__ __forceinline__ int pf_axis(const PushArgs& A, double x, double* w, double* dw) {
    const double bf = floor(x);
    const int b = int(bf);
    const int h = A.h;
    const int first = b - 2 * h + 1;
    if (h == 1 && !A.peskin) {
        const double phi = x - bf;
        const double r = 1.0 - phi;
        const double p2 = phi * phi, r2 = r * r;
        w[0] = r2 * r / 6.0;
        w[1] = (p2 * (phi * 3.0 - 6.0) + 4.0) / 6.0;
        w[2] = (((-phi * 3.0 + 3.0) * phi + 3.0) * phi + 1.0) / 6.0;
        w[3] = p2 * phi / 6.0;
        dw[0] = -r2 * 0.5;
        dw[1] = phi * (phi * 1.5 - 2.0);
        dw[2] = (-phi * 1.5 + 1.0) * phi + 0.5;
        dw[3] = p2 * 0.5;
        return first;
    }
    const double hf = double(h);
    for (int t = 0; t < 4 * h; t++) {
        const double r = (double(first + t) - x) / hf;
        w[t] = (A.peskin ? pf_peskin4(r) : pf_m4(r)) / hf;
        dw[t] = -(A.peskin ? pf_peskin4_derivative(r) : pf_m4_derivative(r)) / (hf * hf);
    }
    return first;
}

__device__ __forceinline__ int pf_wrap(int c, int n) {
    c %= n;
    return c < 0 ? c + n : c;
}

// A periodic one-cell axis is planar: weight 1 on its cell, support 1.
__device__ __forceinline__ bool pf_planar(const PushArgs& A, int e) {
    const int n = e == 0 ? A.nx : (e == 1 ? A.ny : A.nz);
    const int p = e == 0 ? A.px : (e == 1 ? A.py : A.pz);
    return p && n == 1;
}
__device__ __forceinline__ int pf_support(const PushArgs& A, int e) { return pf_planar(A, e) ? 1 : 4 * A.h; }

// The cell the support of coordinate x on axis e starts at: on a closed axis whose
// symmetry face the support crosses, the start of the folded window (0 or n - 4h);
// on a periodic axis, wrapped into [0, n).
__device__ __forceinline__ int pf_start(const PushArgs& A, int e, double x) {
    if (pf_planar(A, e)) return 0;
    const int n = e == 0 ? A.nx : (e == 1 ? A.ny : A.nz);
    const int p = e == 0 ? A.px : (e == 1 ? A.py : A.pz);
    const int span = 4 * A.h;
    const int first = int(floor(x)) - 2 * A.h + 1;
    if ((A.mirror[e] & 1) && first < 0) return 0;
    if ((A.mirror[e] & 2) && first + span > n) return n - span;
    return p ? pf_wrap(first, n) : first;
}

// Weights (and d/dx) of coordinate x on axis e over its support, folded into the lattice
// at a symmetry face it crosses (a cell c < 0 reflects to -1 - c, c >= n to 2n - 1 - c;
// fold_axis, moving/pushforward.rs:344-362); returns pf_start.
__device__ __forceinline__ int pf_axis_at(const PushArgs& A, int e, double x, double* w, double* dw) {
    if (pf_planar(A, e)) {
        w[0] = 1.0;
        dw[0] = 0.0;
        return 0;
    }
    const int n = e == 0 ? A.nx : (e == 1 ? A.ny : A.nz);
    const int span = 4 * A.h;
    const int first = pf_axis(A, x, w, dw);
    int start = -1;
    if ((A.mirror[e] & 1) && first < 0) start = 0;
    else if ((A.mirror[e] & 2) && first + span > n) start = n - span;
    if (start < 0) return pf_start(A, e, x);
    double fw[4 * PF_MAX_H], fd[4 * PF_MAX_H];
    for (int t = 0; t < span; t++) fw[t] = fd[t] = 0.0;
    for (int t = 0; t < span; t++) {
        const int c = first + t;
        const int m = c < 0 ? -1 - c : (c >= n ? 2 * n - 1 - c : c);
        fw[m - start] += w[t];
        fd[m - start] += dw[t];
    }
    for (int t = 0; t < span; t++) {
        w[t] = fw[t];
        dw[t] = fd[t];
    }
    return start;
}

// Sort key of every point: the cell its (folded, wrapped) support starts at.
__global__ __launch_bounds__(256) void k_pf_keys(PushArgs A, int* key, int* idx) {
    const long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= A.np) return;
    int first[3];
    for (int e = 0; e < 3; e++) first[e] = pf_start(A, e, A.xi[3 * q + e]);
    key[q] = int(((long long)first[0] * A.ny + first[1]) * A.nz + first[2]);
    idx[q] = int(q);
}

// Per sorted point p: the bounds [start, end) of its start cell's bin, and its weights in
// sorted order, wts[p * 3s + axis * s + t] (s = 4h, t below the axis's support, the x
// axis times a_q), vel[3p + e].
__global__ __launch_bounds__(256) void k_pf_bins(PushArgs A, const int* skey, const int* sidx, int* start, int* end,
                                                 double* wts, double* vel) {
    const long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= A.np) return;
    const int k = skey[p];
    if (p == 0 || skey[p - 1] != k) start[k] = int(p);
    if (p == A.np - 1 || skey[p + 1] != k) end[k] = int(p + 1);
    const long long q = sidx[p];
    const int s = 4 * A.h;
    const double aq = A.a[q];
    for (int e = 0; e < 3; e++) {
        double w[4 * PF_MAX_H], dw[4 * PF_MAX_H];
        pf_axis_at(A, e, A.xi[3 * q + e], w, dw);
        for (int t = 0; t < pf_support(A, e); t++) wts[p * 3 * s + e * s + t] = e == 0 ? aq * w[t] : w[t];
        if (vel) vel[3 * p + e] = A.v ? A.v[3 * q + e] : 0.0;
    }
}

// d[c] (and m[c]): over the start cells c - t, t below the support on each axis (wrapped
// on periodic axes), and their points in point order, a_q w_x w_y w_z (times v_q).
__global__ __launch_bounds__(256) void k_pf_gather(PushArgs A, const int* start, const int* end, const double* wts,
                                                   const double* vel, double* d, double* m, long long ncell) {
    const long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= ncell) return;
    const int ck = int(c % A.nz), cj = int((c / A.nz) % A.ny), ci = int(c / ((long long)A.nz * A.ny));
    const int s = 4 * A.h;
    const int sx = pf_support(A, 0), sy = pf_support(A, 1), sz = pf_support(A, 2);
    double acc = 0.0, am[3] = {0.0, 0.0, 0.0};
    for (int tx = 0; tx < sx; tx++) {
        int bx = ci - tx;
        if (bx < 0) {
            if (!A.px) break;
            bx = pf_wrap(bx, A.nx);
        }
        for (int ty = 0; ty < sy; ty++) {
            int by = cj - ty;
            if (by < 0) {
                if (!A.py) break;
                by = pf_wrap(by, A.ny);
            }
            for (int tz = 0; tz < sz; tz++) {
                int bz = ck - tz;
                if (bz < 0) {
                    if (!A.pz) break;
                    bz = pf_wrap(bz, A.nz);
                }
                const long long b = ((long long)bx * A.ny + by) * A.nz + bz;
                const int p1 = end[b];
                for (int p = start[b]; p < p1; p++) {
                    const double* w = wts + (long long)p * 3 * s;
                    const double ww = w[tx] * w[s + ty] * w[2 * s + tz];
                    acc += ww;
                    if (m) {
                        am[0] += ww * vel[3 * (long long)p];
                        am[1] += ww * vel[3 * (long long)p + 1];
                        am[2] += ww * vel[3 * (long long)p + 2];
                    }
                }
            }
        }
    }
    d[c] = acc;
    if (m) {
        m[3 * c] = am[0];
        m[3 * c + 1] = am[1];
        m[3 * c + 2] = am[2];
    }
}

// xi_bar = a (sum grad W d_bar + sum grad W (m_bar . v)), a_bar = sum W (d_bar + m_bar . v),
// v_bar = a sum W m_bar (endpoint_vjp).
__global__ __launch_bounds__(128) void k_pushforward_vjp(PushArgs A, const double* d_bar, const double* m_bar,
                                                         double* xi_bar, double* a_bar, double* v_bar) {
    const long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= A.np) return;
    double w[3][4 * PF_MAX_H], dw[3][4 * PF_MAX_H];
    int first[3], sup[3];
    for (int e = 0; e < 3; e++) {
        first[e] = pf_axis_at(A, e, A.xi[3 * q + e], w[e], dw[e]);
        sup[e] = pf_support(A, e);
    }
    double vq[3] = {0.0, 0.0, 0.0};
    if (A.v)
        for (int e = 0; e < 3; e++) vq[e] = A.v[3 * q + e];
    double ks = 0.0, kv[3] = {0.0, 0.0, 0.0}, gs[3] = {0.0, 0.0, 0.0}, gv[3][3] = {};
    for (int tx = 0; tx < sup[0]; tx++) {
        const int cx = pf_wrap(first[0] + tx, A.nx);
        for (int ty = 0; ty < sup[1]; ty++) {
            const int cy = pf_wrap(first[1] + ty, A.ny);
            for (int tz = 0; tz < sup[2]; tz++) {
                const long long cell = ((long long)cx * A.ny + cy) * A.nz + pf_wrap(first[2] + tz, A.nz);
                const double W = w[0][tx] * w[1][ty] * w[2][tz];
                const double dW[3] = {dw[0][tx] * w[1][ty] * w[2][tz], w[0][tx] * dw[1][ty] * w[2][tz],
                                      w[0][tx] * w[1][ty] * dw[2][tz]};
                const double db = d_bar ? d_bar[cell] : 0.0;
                double mb[3] = {0.0, 0.0, 0.0};
                if (m_bar)
                    for (int e = 0; e < 3; e++) mb[e] = m_bar[3 * cell + e];
                ks += W * db;
                for (int e = 0; e < 3; e++) {
                    kv[e] += W * mb[e];
                    gs[e] += dW[e] * db;
                    for (int f = 0; f < 3; f++) gv[e][f] += dW[e] * mb[f];
                }
            }
        }
    }
    const double aq = A.a[q];
    a_bar[q] = ks + kv[0] * vq[0] + kv[1] * vq[1] + kv[2] * vq[2];
    for (int e = 0; e < 3; e++) {
        xi_bar[3 * q + e] = aq * (gs[e] + gv[e][0] * vq[0] + gv[e][1] * vq[1] + gv[e][2] * vq[2]);
        if (v_bar) v_bar[3 * q + e] = aq * kv[e];
    }
}

}  // namespace lbm
