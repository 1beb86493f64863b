// Host-side launchers for the kernels of one storage precision.  Each precision is a
// separate translation unit (kern_f64.hip is built with -ffp-contract=off so the
// double-precision path keeps the reference's IEEE operation order; the others allow
// FMA contraction).  api.cpp dispatches on the precision at run time.
#pragma once
#include <hip/hip_runtime.h>

#include "kernels.h"

namespace lbm {

struct LaunchCfg {
    dim3 grid, block;
    hipStream_t stream;
};

struct Runner {
    virtual ~Runner() = default;
    virtual hipError_t forward(int Q, int kind, bool mask, const StepArgs& a, const LaunchCfg& c) const = 0;
    virtual hipError_t adjoint(int Q, int kind, bool mask, const AdjArgs& a, const LaunchCfg& c) const = 0;
    virtual hipError_t objective(int Q, bool mask, const StepArgs& a, const ObjArgs& seed, void* lam, double* abar,
                                 int do_seed, const LaunchCfg& c) const = 0;
    virtual hipError_t port_apply(int Q, bool mask, const PortArgs& a, void* pop, double* exch, hipStream_t s) const = 0;
    virtual hipError_t port_save(int Q, bool mask, const PortArgs& a, const void* pop, void* buf,
                                 hipStream_t s) const = 0;
    virtual hipError_t port_vjp(int Q, bool mask, const PortArgs& a, void* lam, const void* buf,
                                hipStream_t s) const = 0;
    virtual hipError_t get_state(int Q, bool mask, const Geom& g, const void* pop, const uint32_t* m, int odd,
                                 double* aos, const LaunchCfg& c) const = 0;
    virtual hipError_t set_state(int Q, bool mask, const Geom& g, void* pop, const uint32_t* m, int odd,
                                 const double* aos, const LaunchCfg& c) const = 0;
    virtual hipError_t get_lam(int Q, bool mask, const Geom& g, const void* lam, const uint32_t* m, int odd,
                               double* aos, const LaunchCfg& c) const = 0;
    virtual hipError_t set_lam(int Q, bool mask, const Geom& g, void* lam, const uint32_t* m, int odd,
                               const double* aos, const LaunchCfg& c) const = 0;
    virtual hipError_t fields(int Q, bool mask, const Geom& g, const void* pop, const uint32_t* m, int odd,
                              const void* alpha, const double a[3], double* rho, double* u,
                              const LaunchCfg& c) const = 0;
};

const Runner* runner_f64();
const Runner* runner_f32();
const Runner* runner_f16s();
const Runner* runner_bf16s();

#define LBM_LAUNCH(kern, cfg, ...)                                                     \
    do {                                                                               \
        hipLaunchKernelGGL((kern), (cfg).grid, (cfg).block, 0, (cfg).stream, __VA_ARGS__); \
        return hipGetLastError();                                                      \
    } while (0)

// Supported (lattice, collision) pairs: D3Q19 with BGK, TRT, MRT, stationary BGK-Guo
// and the streaming-only debug operator; D3Q27 with BGK, TRT, MRT and cumulant.
template <int P, int LP, bool ALL_KINDS>
struct RunnerImpl final : Runner {
    template <int Q, bool M>
    static hipError_t fwd_q(int kind, const StepArgs& a, const LaunchCfg& c) {
        switch (kind) {
        case K_TRT: LBM_LAUNCH((k_forward<P, Q, K_TRT, M>), c, a);
        case K_BGK: LBM_LAUNCH((k_forward<P, Q, K_BGK, M>), c, a);
        default: break;
        }
        if constexpr (ALL_KINDS) {
            if (kind == K_MRT) LBM_LAUNCH((k_forward<P, Q, K_MRT, M>), c, a);
            if constexpr (Q == 19) {
                if (kind == K_BGK_GUO) LBM_LAUNCH((k_forward<P, 19, K_BGK_GUO, M>), c, a);
                if (kind == K_STREAM_ONLY) LBM_LAUNCH((k_forward<P, 19, K_STREAM_ONLY, M>), c, a);
            } else {
                if (kind == K_CUMULANT) LBM_LAUNCH((k_forward<P, 27, K_CUMULANT, M>), c, a);
                if (kind == K_STREAM_ONLY) LBM_LAUNCH((k_forward<P, 27, K_STREAM_ONLY, M>), c, a);
            }
        }
        return hipErrorInvalidValue;
    }
    template <int Q, bool M>
    static hipError_t adj_q(int kind, const AdjArgs& a, const LaunchCfg& c) {
        switch (kind) {
        case K_TRT: LBM_LAUNCH((k_adjoint<P, LP, Q, K_TRT, M>), c, a);
        case K_BGK: LBM_LAUNCH((k_adjoint<P, LP, Q, K_BGK, M>), c, a);
        default: break;
        }
        if constexpr (ALL_KINDS) {
            if (kind == K_MRT) LBM_LAUNCH((k_adjoint<P, LP, Q, K_MRT, M>), c, a);
            if constexpr (Q == 19) {
                if (kind == K_BGK_GUO) LBM_LAUNCH((k_adjoint<P, LP, 19, K_BGK_GUO, M>), c, a);
            } else {
                if (kind == K_CUMULANT) LBM_LAUNCH((k_adjoint<P, LP, 27, K_CUMULANT, M>), c, a);
            }
        }
        return hipErrorInvalidValue;
    }

    hipError_t forward(int Q, int kind, bool mask, const StepArgs& a, const LaunchCfg& c) const override {
        if (Q == 19) return mask ? fwd_q<19, true>(kind, a, c) : fwd_q<19, false>(kind, a, c);
        return mask ? fwd_q<27, true>(kind, a, c) : fwd_q<27, false>(kind, a, c);
    }
    hipError_t adjoint(int Q, int kind, bool mask, const AdjArgs& a, const LaunchCfg& c) const override {
        if (Q == 19) return mask ? adj_q<19, true>(kind, a, c) : adj_q<19, false>(kind, a, c);
        return mask ? adj_q<27, true>(kind, a, c) : adj_q<27, false>(kind, a, c);
    }
    hipError_t objective(int Q, bool mask, const StepArgs& a, const ObjArgs& seed, void* lam, double* abar,
                         int do_seed, const LaunchCfg& c) const override {
        if (Q == 19) {
            if (mask) LBM_LAUNCH((k_objective<P, LP, 19, true>), c, a, seed, lam, abar, do_seed);
            LBM_LAUNCH((k_objective<P, LP, 19, false>), c, a, seed, lam, abar, do_seed);
        }
        if (mask) LBM_LAUNCH((k_objective<P, LP, 27, true>), c, a, seed, lam, abar, do_seed);
        LBM_LAUNCH((k_objective<P, LP, 27, false>), c, a, seed, lam, abar, do_seed);
    }
    static LaunchCfg port_cfg(const PortArgs& a, hipStream_t s) {
        return LaunchCfg{dim3((a.np + 63) / 64), dim3(64), s};
    }
    hipError_t port_apply(int Q, bool mask, const PortArgs& a, void* pop, double* exch, hipStream_t s) const override {
        const LaunchCfg c = port_cfg(a, s);
        if (Q == 19) {
            if (mask) LBM_LAUNCH((k_port_apply<P, 19, true>), c, a, pop, exch);
            LBM_LAUNCH((k_port_apply<P, 19, false>), c, a, pop, exch);
        }
        if (mask) LBM_LAUNCH((k_port_apply<P, 27, true>), c, a, pop, exch);
        LBM_LAUNCH((k_port_apply<P, 27, false>), c, a, pop, exch);
    }
    hipError_t port_save(int Q, bool mask, const PortArgs& a, const void* pop, void* buf,
                         hipStream_t s) const override {
        using C = typename Store<P>::C;
        const LaunchCfg c = port_cfg(a, s);
        C* b = static_cast<C*>(buf);
        if (Q == 19) {
            if (mask) LBM_LAUNCH((k_port_save<P, 19, true>), c, a, pop, b);
            LBM_LAUNCH((k_port_save<P, 19, false>), c, a, pop, b);
        }
        if (mask) LBM_LAUNCH((k_port_save<P, 27, true>), c, a, pop, b);
        LBM_LAUNCH((k_port_save<P, 27, false>), c, a, pop, b);
    }
    hipError_t port_vjp(int Q, bool mask, const PortArgs& a, void* lam, const void* buf,
                        hipStream_t s) const override {
        using C = typename Store<P>::C;
        const LaunchCfg c = port_cfg(a, s);
        const C* b = static_cast<const C*>(buf);
        if (Q == 19) {
            if (mask) LBM_LAUNCH((k_port_vjp<P, LP, 19, true>), c, a, lam, b);
            LBM_LAUNCH((k_port_vjp<P, LP, 19, false>), c, a, lam, b);
        }
        if (mask) LBM_LAUNCH((k_port_vjp<P, LP, 27, true>), c, a, lam, b);
        LBM_LAUNCH((k_port_vjp<P, LP, 27, false>), c, a, lam, b);
    }
    hipError_t get_state(int Q, bool mask, const Geom& g, const void* pop, const uint32_t* m, int odd, double* aos,
                         const LaunchCfg& c) const override {
        if (Q == 19) {
            if (mask) LBM_LAUNCH((k_get_state<P, 19, true>), c, g, pop, m, odd, aos);
            LBM_LAUNCH((k_get_state<P, 19, false>), c, g, pop, m, odd, aos);
        }
        if (mask) LBM_LAUNCH((k_get_state<P, 27, true>), c, g, pop, m, odd, aos);
        LBM_LAUNCH((k_get_state<P, 27, false>), c, g, pop, m, odd, aos);
    }
    hipError_t set_state(int Q, bool mask, const Geom& g, void* pop, const uint32_t* m, int odd, const double* aos,
                         const LaunchCfg& c) const override {
        if (Q == 19) {
            if (mask) LBM_LAUNCH((k_set_state<P, 19, true>), c, g, pop, m, odd, aos);
            LBM_LAUNCH((k_set_state<P, 19, false>), c, g, pop, m, odd, aos);
        }
        if (mask) LBM_LAUNCH((k_set_state<P, 27, true>), c, g, pop, m, odd, aos);
        LBM_LAUNCH((k_set_state<P, 27, false>), c, g, pop, m, odd, aos);
    }
    hipError_t get_lam(int Q, bool mask, const Geom& g, const void* lam, const uint32_t* m, int odd, double* aos,
                       const LaunchCfg& c) const override {
        if (Q == 19) {
            if (mask) LBM_LAUNCH((k_get_lam<LP, 19, true>), c, g, lam, m, odd, aos);
            LBM_LAUNCH((k_get_lam<LP, 19, false>), c, g, lam, m, odd, aos);
        }
        if (mask) LBM_LAUNCH((k_get_lam<LP, 27, true>), c, g, lam, m, odd, aos);
        LBM_LAUNCH((k_get_lam<LP, 27, false>), c, g, lam, m, odd, aos);
    }
    hipError_t set_lam(int Q, bool mask, const Geom& g, void* lam, const uint32_t* m, int odd, const double* aos,
                       const LaunchCfg& c) const override {
        if (Q == 19) {
            if (mask) LBM_LAUNCH((k_set_lam<LP, 19, true>), c, g, lam, m, odd, aos);
            LBM_LAUNCH((k_set_lam<LP, 19, false>), c, g, lam, m, odd, aos);
        }
        if (mask) LBM_LAUNCH((k_set_lam<LP, 27, true>), c, g, lam, m, odd, aos);
        LBM_LAUNCH((k_set_lam<LP, 27, false>), c, g, lam, m, odd, aos);
    }
    hipError_t fields(int Q, bool mask, const Geom& g, const void* pop, const uint32_t* m, int odd, const void* alpha,
                      const double a[3], double* rho, double* u, const LaunchCfg& c) const override {
        if (Q == 19) {
            if (mask) LBM_LAUNCH((k_fields<P, 19, true>), c, g, pop, m, odd, alpha, a[0], a[1], a[2], rho, u);
            LBM_LAUNCH((k_fields<P, 19, false>), c, g, pop, m, odd, alpha, a[0], a[1], a[2], rho, u);
        }
        if (mask) LBM_LAUNCH((k_fields<P, 27, true>), c, g, pop, m, odd, alpha, a[0], a[1], a[2], rho, u);
        LBM_LAUNCH((k_fields<P, 27, false>), c, g, pop, m, odd, alpha, a[0], a[1], a[2], rho, u);
    }
};

}  // namespace lbm
