// Tangent (forward-mode) evaluation of one step: the step and port kernels of kernels.h
// instantiated on a dual number (value, tangent), so that the Jacobian-vector product
// J_state dx + J_alpha dalpha of the step (ScanStep::jvp, crates/implexity-ad/src/
// scan.rs:46-80) runs through the same collision, streaming and port code as the step
// itself.  The value part is FP64 arithmetic in the kernels' operation order (the
// translation unit is built with -ffp-contract=off), so it reproduces the FP64 step.
#pragma once
#include <hip/hip_runtime.h>

#include "kernels.h"

namespace lbm {

// Storage/arithmetic tag of the dual field (after the Prec values of kernels.h).
constexpr int P_JVP = 4;

// The dual field holds full populations f (no shift), 16 bytes per value.
template <>
struct Store<P_JVP> {
    using T = Dual;
    using C = Dual;
    static constexpr bool shifted = false;
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

// AoS values and tangents (x*Q + q) into a dual field of parity odd (fluid cells).
template <int Q, bool MASK>
__global__ __launch_bounds__(256) void k_set_dual(Geom g, Dual* pop, const uint32_t* mask, int odd, const double* val,
                                                  const double* tan) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y * blockDim.y + threadIdx.y, i = blockIdx.z;
    if (k >= g.nz || j >= g.ny) return;
    const long long cell = ((long long)i * g.ny + j) * g.nz + k;
    const uint32_t m = cell_mask<Q, MASK>(g, mask, cell, i, j, k);
    if (m & SOLID_BIT) return;
    const Cursor<Q> cu = make_cursor<Q>(g, i, j, k);
    static_for<0, Q>([&](auto I) {
        constexpr int q = decltype(I)::value;
        store_read<Q, q>(pop, g, cu, m, odd, Dual(val[cell * Q + q], tan[cell * Q + q]));
    });
}

// Tangents of a dual field of parity odd into AoS (fluid cells; solid cells untouched).
template <int Q, bool MASK>
__global__ __launch_bounds__(256) void k_get_tangent(Geom g, const Dual* pop, const uint32_t* mask, int odd,
                                                     double* tan) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y * blockDim.y + threadIdx.y, i = blockIdx.z;
    if (k >= g.nz || j >= g.ny) return;
    const long long cell = ((long long)i * g.ny + j) * g.nz + k;
    const uint32_t m = cell_mask<Q, MASK>(g, mask, cell, i, j, k);
    if (m & SOLID_BIT) return;
    const Cursor<Q> cu = make_cursor<Q>(g, i, j, k);
    static_for<0, Q>([&](auto I) {
        constexpr int q = decltype(I)::value;
        tan[cell * Q + q] = load_read<Q, q>(pop, g, cu, m, odd).d;
    });
}

// Launchers of the tangent step (kern_jvp.hip).  The step and port arguments are those of
// the handle's step with the state, the design and the collision constants replaced by
// dual arrays.
struct JvpRunner {
    hipError_t step(int Q, int kind, bool mask, const StepArgs& a, dim3 grid, dim3 block, hipStream_t s) const;
    hipError_t ports(int Q, bool mask, const PortArgs& a, void* pop, hipStream_t s) const;
    hipError_t sponge(int Q, const SpongeArgs& a, void* pop, hipStream_t s) const;
    hipError_t wale_u(int Q, bool mask, const Geom& g, const void* pop, const uint32_t* m, int odd, void* U,
                      dim3 grid, dim3 block, hipStream_t s) const;
    hipError_t set(int Q, bool mask, const Geom& g, void* pop, const uint32_t* m, int odd, const double* val,
                   const double* tan, dim3 grid, dim3 block, hipStream_t s) const;
    hipError_t get(int Q, bool mask, const Geom& g, const void* pop, const uint32_t* m, int odd, double* tan,
                   dim3 grid, dim3 block, hipStream_t s) const;
};
const JvpRunner* jvp_runner();

}  // namespace lbm
