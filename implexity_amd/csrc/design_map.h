// The design map raw -> phi on the device (design.rs:95-280, DesignMap::{map, vjp}): a
// cone filter over the editable cells within `radius` cells, then the normalised tanh
// projection; solid cells map to 0 and fixed cells to their fixed value.  The filter
// is a gather with the offsets and weights in the reference's (a, b, c) loop order, so
// the double-precision map reproduces the reference's sums; the transpose is a second
// gather (out[s] = sum_o w_o scale[s + o]), in place of the reference's scatter.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace lbm {

constexpr uint8_t DM_SOLID = 1, DM_EDIT = 2;

struct DesignMapArgs {
    int nx, ny, nz;
    int px, py, pz;
    int noff;             // offsets with a non-zero weight (0: no filter)
    const int* off;       // (a, b, c) per offset
    const double* w;      // weight per offset
    const uint8_t* flags; // DM_SOLID / DM_EDIT per cell
    const double* fixed;  // value of the cells outside the design region
    double beta, eta;     // projection (beta = 0: none)
    double pa, pden;      // tanh(beta eta) and pa + tanh(beta (1 - eta)), formed on the host
};

// Cell at (i + a, j + b, k + c) with periodic wrap, or -1 outside a closed axis
// (Grid::inside / Grid::wrap, d3q19.rs:101-128).
__device__ __forceinline__ long long dm_shift(const DesignMapArgs& A, int i, int j, int k, int a, int b, int c) {
    int v[3] = {i + a, j + b, k + c};
    const int n[3] = {A.nx, A.ny, A.nz};
    const int p[3] = {A.px, A.py, A.pz};
#pragma unroll
    for (int t = 0; t < 3; t++) {
        if (v[t] < 0 || v[t] >= n[t]) {
            if (!p[t]) return -1;
            v[t] %= n[t];
            if (v[t] < 0) v[t] += n[t];
        }
    }
    return ((long long)v[0] * A.ny + v[1]) * A.nz + v[2];
}

__device__ __forceinline__ void dm_cell(const DesignMapArgs& A, long long x, int& i, int& j, int& k) {
    k = int(x % A.nz);
    j = int((x / A.nz) % A.ny);
    i = int(x / ((long long)A.nz * A.ny));
}

// DesignMap::filtered and the row's denominator (1 without a filter).
__device__ __forceinline__ void dm_filter(const DesignMapArgs& A, const double* raw, long long x, double& f,
                                          double& den) {
    if (A.noff == 0) {
        f = raw[x];
        den = 1.0;
        return;
    }
    int i, j, k;
    dm_cell(A, x, i, j, k);
    double num = 0.0, d = 0.0;
    for (int o = 0; o < A.noff; o++) {
        // the row of x holds the editable source x - offset
        const long long s = dm_shift(A, i, j, k, -A.off[3 * o], -A.off[3 * o + 1], -A.off[3 * o + 2]);
        if (s < 0 || !(A.flags[s] & DM_EDIT)) continue;
        num += A.w[o] * raw[s];
        d += A.w[o];
    }
    den = d > 0.0 ? d : 1.0;
    f = num / den;
}

__global__ __launch_bounds__(256) void k_design_map(DesignMapArgs A, const double* raw, double* phi, long long n) {
    const long long x = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (x >= n) return;
    const uint8_t fl = A.flags[x];
    double v;
    if (fl & DM_SOLID) {
        v = 0.0;
    } else if (fl & DM_EDIT) {
        double f, den;
        dm_filter(A, raw, x, f, den);
        v = A.beta > 0.0 ? (A.pa + tanh(A.beta * (f - A.eta))) / A.pden : f;
    } else {
        v = A.fixed[x];
    }
    phi[x] = v;
}

// scale[x] = phi_bar[x] * (projection slope) / (row denominator) on editable cells, else 0.
__global__ __launch_bounds__(256) void k_design_scale(DesignMapArgs A, const double* raw, const double* phi_bar,
                                                      double* scale, long long n) {
    const long long x = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (x >= n) return;
    double s = 0.0;
    if ((A.flags[x] & DM_EDIT) && phi_bar[x] != 0.0) {
        double f, den;
        dm_filter(A, raw, x, f, den);
        double g = phi_bar[x];
        if (A.beta > 0.0) {
            const double t = tanh(A.beta * (f - A.eta));
            g *= A.beta * (1.0 - t * t) / A.pden;
        }
        s = A.noff == 0 ? g : g / den;
    }
    scale[x] = s;
}

// raw_bar[s] = sum over the rows x that hold s (x = s + offset) of w * scale[x].
__global__ __launch_bounds__(256) void k_design_gather(DesignMapArgs A, const double* scale, double* out, long long n) {
    const long long s = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    double v = 0.0;
    if (A.flags[s] & DM_EDIT) {
        if (A.noff == 0) {
            v = scale[s];
        } else {
            int i, j, k;
            dm_cell(A, s, i, j, k);
            for (int o = 0; o < A.noff; o++) {
                const long long x = dm_shift(A, i, j, k, A.off[3 * o], A.off[3 * o + 1], A.off[3 * o + 2]);
                if (x < 0) continue;
                v += A.w[o] * scale[x];
            }
        }
    }
    out[s] = v;
}

}  // namespace lbm
