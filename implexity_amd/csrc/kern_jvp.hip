// Tangent-step kernels on dual numbers (jvp.h); built with -ffp-contract=off so that the
// value part follows the FP64 step's operation order.
#include "jvp.h"

namespace lbm {

namespace {
template <int Q, bool M>
hipError_t step_q(int kind, const StepArgs& a, dim3 grid, dim3 block, hipStream_t s) {
    switch (kind) {
    case K_BGK: hipLaunchKernelGGL((k_forward<P_JVP, Q, K_BGK, M>), grid, block, 0, s, a); break;
    case K_TRT: hipLaunchKernelGGL((k_forward<P_JVP, Q, K_TRT, M>), grid, block, 0, s, a); break;
    case K_MRT: hipLaunchKernelGGL((k_forward<P_JVP, Q, K_MRT, M>), grid, block, 0, s, a); break;
    case K_TRT_LES: hipLaunchKernelGGL((k_forward<P_JVP, Q, K_TRT_LES, M>), grid, block, 0, s, a); break;
    case K_BGK_LES: hipLaunchKernelGGL((k_forward<P_JVP, Q, K_BGK_LES, M>), grid, block, 0, s, a); break;
    case K_TRT_WALE: hipLaunchKernelGGL((k_forward<P_JVP, Q, K_TRT_WALE, M>), grid, block, 0, s, a); break;
    case K_BGK_WALE: hipLaunchKernelGGL((k_forward<P_JVP, Q, K_BGK_WALE, M>), grid, block, 0, s, a); break;
    case K_TRT_PSMS: hipLaunchKernelGGL((k_forward<P_JVP, Q, K_TRT_PSMS, M>), grid, block, 0, s, a); break;
    case K_BGK_PSMS: hipLaunchKernelGGL((k_forward<P_JVP, Q, K_BGK_PSMS, M>), grid, block, 0, s, a); break;
    case K_TRT_BRO: hipLaunchKernelGGL((k_forward<P_JVP, Q, K_TRT_BRO, M>), grid, block, 0, s, a); break;
    case K_BGK_BRO: hipLaunchKernelGGL((k_forward<P_JVP, Q, K_BGK_BRO, M>), grid, block, 0, s, a); break;
    default:
        if constexpr (Q == 19) {
            // the Psm law on dual numbers is D3Q19 only (on D3Q27 it spills to scratch heavily)
            if (kind == K_TRT_PSM) {
                hipLaunchKernelGGL((k_forward<P_JVP, 19, K_TRT_PSM, M>), grid, block, 0, s, a);
                break;
            }
            if (kind == K_BGK_PSM) {
                hipLaunchKernelGGL((k_forward<P_JVP, 19, K_BGK_PSM, M>), grid, block, 0, s, a);
                break;
            }
            if (kind == K_BGK_GUO) {
                hipLaunchKernelGGL((k_forward<P_JVP, 19, K_BGK_GUO, M>), grid, block, 0, s, a);
                break;
            }
        } else {
            if (kind == K_CUMULANT) {
                hipLaunchKernelGGL((k_forward<P_JVP, 27, K_CUMULANT, M>), grid, block, 0, s, a);
                break;
            }
        }
        return hipErrorInvalidValue;
    }
    return hipGetLastError();
}
}  // namespace

hipError_t JvpRunner::step(int Q, int kind, bool mask, const StepArgs& a, dim3 grid, dim3 block,
                           hipStream_t s) const {
    if (!a.pop || !a.pop_out || !a.alpha || (mask && !a.mask) || a.part || a.tpop || a.gacc) return hipErrorInvalidValue;
    if ((kind == K_MRT || kind == K_CUMULANT || kind == K_TRT_LES || kind == K_BGK_LES || is_wale(kind)) && !a.mrt)
        return hipErrorInvalidValue;
    if (is_wale(kind) && !a.wu) return hipErrorInvalidValue;
    if (is_psm(kind) && (!a.sm || !a.mrt)) return hipErrorInvalidValue;
    if (Q == 19) return mask ? step_q<19, true>(kind, a, grid, block, s) : step_q<19, false>(kind, a, grid, block, s);
    if (Q == 27) return mask ? step_q<27, true>(kind, a, grid, block, s) : step_q<27, false>(kind, a, grid, block, s);
    return hipErrorInvalidValue;
}

// The WALE pre-pass on the dual state: u and rho of every cell with their tangents.
hipError_t JvpRunner::wale_u(int Q, bool mask, const Geom& g, const void* pop, const uint32_t* m, int odd, void* U,
                             dim3 grid, dim3 block, hipStream_t s) const {
    if (!pop || !U || (mask && !m)) return hipErrorInvalidValue;
    if (Q == 19) {
        if (mask)
            hipLaunchKernelGGL((k_wale_u<P_JVP, 19, true>), grid, block, 0, s, g, pop, m, odd, U);
        else
            hipLaunchKernelGGL((k_wale_u<P_JVP, 19, false>), grid, block, 0, s, g, pop, m, odd, U);
    } else if (Q == 27) {
        if (mask)
            hipLaunchKernelGGL((k_wale_u<P_JVP, 27, true>), grid, block, 0, s, g, pop, m, odd, U);
        else
            hipLaunchKernelGGL((k_wale_u<P_JVP, 27, false>), grid, block, 0, s, g, pop, m, odd, U);
    } else {
        return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

// The sponge blend on dual numbers: tangent (1 - sigma), the far field a constant.
hipError_t JvpRunner::sponge(int Q, const SpongeArgs& a, void* pop, hipStream_t s) const {
    if (a.n <= 0) return hipSuccess;
    if (!pop || a.ns <= 0 || a.ns > MAX_SPONGES || !a.cell || !a.sigma || !a.off || !a.part || !a.weight || !a.phase)
        return hipErrorInvalidValue;
    const dim3 grid((a.n + SPONGE_BLOCK - 1) / SPONGE_BLOCK), block(SPONGE_BLOCK);
    if (Q == 19)
        hipLaunchKernelGGL((k_sponge<P_JVP, 19>), grid, block, 0, s, a, pop, nullptr);
    else if (Q == 27)
        hipLaunchKernelGGL((k_sponge<P_JVP, 27>), grid, block, 0, s, a, pop, nullptr);
    else
        return hipErrorInvalidValue;
    return hipGetLastError();
}

hipError_t JvpRunner::ports(int Q, bool mask, const PortArgs& a, void* pop, hipStream_t s) const {
    if (a.np <= 0) return hipSuccess;
    if (!a.cell || !a.nb || !a.kind || !a.value || (mask && !a.mask)) return hipErrorInvalidValue;
    const dim3 grid((a.np + PORT_BLOCK - 1) / PORT_BLOCK), block(PORT_BLOCK);
    if (Q == 19 && a.st) {
        if (mask)
            hipLaunchKernelGGL((k_port_apply<P_JVP, 19, true, true>), grid, block, 0, s, a, pop, nullptr);
        else
            hipLaunchKernelGGL((k_port_apply<P_JVP, 19, false, true>), grid, block, 0, s, a, pop, nullptr);
    } else if (Q == 19) {
        if (mask)
            hipLaunchKernelGGL((k_port_apply<P_JVP, 19, true>), grid, block, 0, s, a, pop, nullptr);
        else
            hipLaunchKernelGGL((k_port_apply<P_JVP, 19, false>), grid, block, 0, s, a, pop, nullptr);
    } else if (Q == 27) {
        if (mask)
            hipLaunchKernelGGL((k_port_apply<P_JVP, 27, true>), grid, block, 0, s, a, pop, nullptr);
        else
            hipLaunchKernelGGL((k_port_apply<P_JVP, 27, false>), grid, block, 0, s, a, pop, nullptr);
    } else {
        return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t JvpRunner::set(int Q, bool mask, const Geom& g, void* pop, const uint32_t* m, int odd, const double* val,
                          const double* tan, dim3 grid, dim3 block, hipStream_t s) const {
    Dual* p = static_cast<Dual*>(pop);
    if (Q == 19) {
        if (mask)
            hipLaunchKernelGGL((k_set_dual<19, true>), grid, block, 0, s, g, p, m, odd, val, tan);
        else
            hipLaunchKernelGGL((k_set_dual<19, false>), grid, block, 0, s, g, p, m, odd, val, tan);
    } else {
        if (mask)
            hipLaunchKernelGGL((k_set_dual<27, true>), grid, block, 0, s, g, p, m, odd, val, tan);
        else
            hipLaunchKernelGGL((k_set_dual<27, false>), grid, block, 0, s, g, p, m, odd, val, tan);
    }
    return hipGetLastError();
}

hipError_t JvpRunner::get(int Q, bool mask, const Geom& g, const void* pop, const uint32_t* m, int odd, double* tan,
                          dim3 grid, dim3 block, hipStream_t s) const {
    const Dual* p = static_cast<const Dual*>(pop);
    if (Q == 19) {
        if (mask)
            hipLaunchKernelGGL((k_get_tangent<19, true>), grid, block, 0, s, g, p, m, odd, tan);
        else
            hipLaunchKernelGGL((k_get_tangent<19, false>), grid, block, 0, s, g, p, m, odd, tan);
    } else {
        if (mask)
            hipLaunchKernelGGL((k_get_tangent<27, true>), grid, block, 0, s, g, p, m, odd, tan);
        else
            hipLaunchKernelGGL((k_get_tangent<27, false>), grid, block, 0, s, g, p, m, odd, tan);
    }
    return hipGetLastError();
}

const JvpRunner* jvp_runner() {
    static const JvpRunner r;
    return &r;
}

}  // namespace lbm
