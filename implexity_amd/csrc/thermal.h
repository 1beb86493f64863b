// Coupled D3Q7 temperature field (a double-distribution advection-diffusion lattice)
// on the storage geometry of the flow.  No counterpart in the reference, whose thermal
// coupling is finite-volume energy transport on D3Q19 links
// (crates/implexity-physics-lbm/src/lbm/thermal.rs:91-134); the model:
//   T = sum_i g_i,  g_i^eq = w_i T (1 + 4 c_i.u)   (w = 1/4, 1/8; c_s^2 = 1/4),
//   u = the Brinkman velocity of the flow state the same interval collides,
//   kappa(phi) = kappa_f (1 + (ratio - 1) phi),  tau_T = 1/2 + 4 kappa,  S = source phi,
//   g_i* = g_i - omega_T (g_i - g_i^eq) + w_i S,
// streamed in place with the esoteric twist and halfway bounce-back (adiabatic walls),
// then at the flow's port cells: velocity ports hold T = T_in (g_b = w T_in + g_nb -
// w T_nb), pressure ports are zero-gradient (g_b = g_nb).
//
// The temperature populations sit in the same state buffer as the flow, after its Q
// arrays, so snapshots and restores move both fields together.  The flow's forward
// kernel advances both fields of a cell in one pass (k_forward with TH, using the
// velocity its collision computes; thermal_cell.h); the ports of both fields follow.
// A reverse step runs k_thermal_adj, then the flow adjoint kernel, which adds the
// velocity cotangent k_thermal_adj hands over.  Hard solid cells are not part of the
// model (the host refuses them), so every link mask is the face mask.
#pragma once
#include "kernels.h"

namespace lbm {

struct ThermalArgs {
    Geom g;
    const void* pop;    // flow state read (parity odd)
    const void* tpop;   // temperature populations read
    void* tpop_out;     // temperature populations written (== tpop in place)
    void* tlam;         // cotangent of the temperature populations (adjoint)
    void* ubar;         // adjoint: cotangent of the flow's Brinkman velocity (SoA 3 x cells)
    const void* alpha;  // flow design (arithmetic precision)
    const void* phi;    // thermal design (arithmetic precision)
    double* alpha_bar;
    double* phi_bar;
    double a[3];
    double kappa, ratio_m1, source;  // ratio_m1 = ratio - 1
    float tenc, tdec;
    int odd;
    uint32_t step;  // index of the interval (rounding key of FP16 stores)
    double obj_w;  // stage cost of the state read: obj_w * T on every cell with phi > 1/2
    double* part;  // per-block partials (J in slot 0), or null
    double* stage;
    unsigned* ticket;
    double* jacc;
};


// Adjoint, before the flow adjoint kernel of the same reverse step: lambda_g in place
// (write locations -> read locations), the design cotangent phi_bar, and the cotangent
// of the flow velocity the step advected with, which the flow adjoint kernel adds
// through u(f, alpha) (brinkman_u_vjp).
template <int P, int LP, int Q>
__global__ __launch_bounds__(256) void k_thermal_adj(ThermalArgs A) {
    using S = Store<P>;
    using L = LamStore<LP>;
    using C = typename S::C;
    const Geom& g = A.g;
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y * blockDim.y + threadIdx.y, i = blockIdx.z;
    if (k >= g.nz || j >= g.ny) return;
    const long long cell = ((long long)i * g.ny + j) * g.nz + k;
    const Cursor<Q> cf = make_cursor<Q>(g, i, j, k);
    const Cursor<TQ> cu = make_cursor<TQ>(g, i, j, k);
    const uint32_t m = face_mask<Q>(g, i, j, k);
    C f[Q];
    load_cell<P, Q>(g, static_cast<const typename S::T*>(A.pop), cf, m, A.odd, f);
    const C alpha = static_cast<const C*>(A.alpha)[cell];
    const C a[3] = {C(A.a[0]), C(A.a[1]), C(A.a[2])};
    Macro<Q, C> mac;
    macro_of<Q, C, S::shifted>(f, alpha, a, mac);
    const uint32_t mt = face_mask<TQ>(g, i, j, k);
    C gv[TQ], gb[TQ], g_bar[TQ], ub[3], ob, sb;
    load_tcell<P>(g, static_cast<const typename TStore<P>::T*>(A.tpop), cu, mt, A.odd, A.tdec, gv);
    typename L::T* tl = static_cast<typename L::T*>(A.tlam);
    static_for<0, TQ>([&](auto I) {
        constexpr int q = decltype(I)::value;
        gb[q] = L::template ld<C>(&at_write<TQ, q>(tl, g, cu, A.odd));
    });
    const C phi = static_cast<const C*>(A.phi)[cell];
    const C om = thermal_omega<C>(A.kappa, A.ratio_m1, phi);
    thermal_collide_vjp<C>(gv, mac.u, om, gb, g_bar, ub, ob, sb);
    if (A.obj_w != 0.0 && phi > C(0.5)) {
#pragma unroll
        for (int q = 0; q < TQ; q++) g_bar[q] += C(A.obj_w);
    }
    static_for<0, TQ>([&](auto I) {
        constexpr int q = decltype(I)::value;
        store_read<TQ, q>(tl, g, cu, mt, A.odd, L::template enc<C>(g_bar[q]));
    });
    // the velocity cotangent goes to the flow adjoint kernel of the same reverse step
    C* U = static_cast<C*>(A.ubar);
    const long long n = (long long)g.nx * g.ny * g.nz;
    U[cell] = ub[0];
    U[n + cell] = ub[1];
    U[2 * n + cell] = ub[2];
    const C dom = -om * om * C(4) * C(A.kappa) * C(A.ratio_m1);
    A.phi_bar[cell] += double(ob * dom + sb * C(A.source));
}

// Final state of a sweep: its stage/terminal cost into the partials and, with do_seed,
// lambda_g of the state := seed_w on the cells with phi > 1/2 (0 elsewhere).
template <int P, int LP>
__global__ __launch_bounds__(256) void k_thermal_seed(ThermalArgs A, double seed_w, int do_seed) {
    using TS = TStore<P>;
    using L = LamStore<LP>;
    using C = typename TS::C;
    const Geom& g = A.g;
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y * blockDim.y + threadIdx.y, i = blockIdx.z;
    double jsum = 0.0;
    if (k < g.nz && j < g.ny) {
        const long long cell = ((long long)i * g.ny + j) * g.nz + k;
        const Cursor<TQ> cu = make_cursor<TQ>(g, i, j, k);
        const uint32_t mt = face_mask<TQ>(g, i, j, k);
        const C phi = static_cast<const C*>(A.phi)[cell];
        const bool solid = phi > C(0.5);
        if (A.obj_w != 0.0 && solid) {
            C gv[TQ];
            load_tcell<P>(g, static_cast<const typename TS::T*>(A.tpop), cu, mt, A.odd, A.tdec, gv);
            C T = C(0);
#pragma unroll
            for (int q = 0; q < TQ; q++) T += gv[q];
            jsum = A.obj_w * double(T);
        }
        if (do_seed) {
            typename L::T* tl = static_cast<typename L::T*>(A.tlam);
            const C s = solid ? C(seed_w) : C(0);
            static_for<0, TQ>([&](auto I) {
                constexpr int q = decltype(I)::value;
                store_read<TQ, q>(tl, g, cu, mt, A.odd, L::template enc<C>(s));
            });
        }
    }
    if (A.part) block_partials(jsum, 0.0, 0.0f, false, -INFINITY, A.part);
}

// Port cells after streaming (state of parity A.odd).
template <int P>
__global__ __launch_bounds__(128) void k_thermal_port(PortArgs A, void* tpop_, double t_in, float tenc, float tdec,
                                                       uint32_t step) {
    using TS = TStore<P>;
    using C = typename TS::C;
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= A.np) return;
    typename TS::T* tp = static_cast<typename TS::T*>(tpop_);
    Cursor<TQ> cx, cn;
    const uint32_t mx = mask_of<TQ, false>(A, A.cell[r], cx);
    const uint32_t mn = mask_of<TQ, false>(A, A.nb[r], cn);
    C gn[TQ], out[TQ];
    load_tcell<P>(A.g, tp, cn, mn, A.odd, tdec, gn);
    if (A.kind[r] == 0) {
        C Tn = C(0);
#pragma unroll
        for (int q = 0; q < TQ; q++) Tn += gn[q];
#pragma unroll
        for (int q = 0; q < TQ; q++) {
            const C w = C(weight<TQ>(q));
            out[q] = w * C(t_in) + (gn[q] - w * Tn);
        }
    } else {
#pragma unroll
        for (int q = 0; q < TQ; q++) out[q] = gn[q];
    }
    const uint32_t key = (uint32_t(A.cell[r]) * uint32_t(TQ)) ^ (step * 0x85ebca6bu);
    static_for<0, TQ>([&](auto I) {
        constexpr int q = decltype(I)::value;
        store_read<TQ, q>(tp, A.g, cx, mx, A.odd, tstore_step<P>(out[q], tenc, key + uint32_t(q)));
    });
}

// Transpose of the thermal port pass on lambda_g of a state of parity A.odd.
template <int LP, class C>
__global__ __launch_bounds__(128) void k_thermal_port_vjp(PortArgs A, void* tlam_) {
    using L = LamStore<LP>;
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= A.np) return;
    typename L::T* tl = static_cast<typename L::T*>(tlam_);
    Cursor<TQ> cx, cn;
    const uint32_t mx = mask_of<TQ, false>(A, A.cell[r], cx);
    const uint32_t mn = mask_of<TQ, false>(A, A.nb[r], cn);
    C gx[TQ], gn[TQ];
    static_for<0, TQ>([&](auto I) {
        constexpr int q = decltype(I)::value;
        gx[q] = L::template ld<C>(&at_read<TQ, q>(tl, A.g, cx, mx, A.odd));
        gn[q] = L::template ld<C>(&at_read<TQ, q>(tl, A.g, cn, mn, A.odd));
    });
    C s = C(0);
    if (A.kind[r] == 0) {
#pragma unroll
        for (int q = 0; q < TQ; q++) s += C(weight<TQ>(q)) * gx[q];
    }
    static_for<0, TQ>([&](auto I) {
        constexpr int q = decltype(I)::value;
        store_read<TQ, q>(tl, A.g, cn, mn, A.odd, L::template enc<C>(gn[q] + (gx[q] - s)));
        store_read<TQ, q>(tl, A.g, cx, mx, A.odd, L::template enc<C>(C(0)));
    });
}

// Layout conversion of the temperature populations / their cotangent (AoS x*7+q).
template <int P>
__global__ __launch_bounds__(256) void k_get_tstate(Geom g, const void* tpop, int odd, double* aos, float tdec) {
    using TS = TStore<P>;
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y * blockDim.y + threadIdx.y, i = blockIdx.z;
    if (k >= g.nz || j >= g.ny) return;
    const long long cell = ((long long)i * g.ny + j) * g.nz + k;
    const Cursor<TQ> cu = make_cursor<TQ>(g, i, j, k);
    typename TS::C gv[TQ];
    load_tcell<P>(g, static_cast<const typename TS::T*>(tpop), cu, face_mask<TQ>(g, i, j, k), odd, tdec, gv);
#pragma unroll
    for (int q = 0; q < TQ; q++) aos[cell * TQ + q] = double(gv[q]);
}

template <int P>
__global__ __launch_bounds__(256) void k_set_tstate(Geom g, void* tpop, int odd, const double* aos, float tenc) {
    using TS = TStore<P>;
    using C = typename TS::C;
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y * blockDim.y + threadIdx.y, i = blockIdx.z;
    if (k >= g.nz || j >= g.ny) return;
    const long long cell = ((long long)i * g.ny + j) * g.nz + k;
    const Cursor<TQ> cu = make_cursor<TQ>(g, i, j, k);
    const uint32_t mt = face_mask<TQ>(g, i, j, k);
    typename TS::T* tp = static_cast<typename TS::T*>(tpop);
    static_for<0, TQ>([&](auto I) {
        constexpr int q = decltype(I)::value;
        store_read<TQ, q>(tp, g, cu, mt, odd, TS::enc(C(aos[cell * TQ + q]), tenc));
    });
}

template <int LP>
__global__ __launch_bounds__(256) void k_get_tlam(Geom g, const void* tlam, int odd, double* aos) {
    using L = LamStore<LP>;
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y * blockDim.y + threadIdx.y, i = blockIdx.z;
    if (k >= g.nz || j >= g.ny) return;
    const long long cell = ((long long)i * g.ny + j) * g.nz + k;
    const Cursor<TQ> cu = make_cursor<TQ>(g, i, j, k);
    const uint32_t mt = face_mask<TQ>(g, i, j, k);
    const typename L::T* tl = static_cast<const typename L::T*>(tlam);
    static_for<0, TQ>([&](auto I) {
        constexpr int q = decltype(I)::value;
        aos[cell * TQ + q] = L::template ld<double>(&at_read<TQ, q>(tl, g, cu, mt, odd));
    });
}

template <int LP>
__global__ __launch_bounds__(256) void k_set_tlam(Geom g, void* tlam, int odd, const double* aos) {
    using L = LamStore<LP>;
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y * blockDim.y + threadIdx.y, i = blockIdx.z;
    if (k >= g.nz || j >= g.ny) return;
    const long long cell = ((long long)i * g.ny + j) * g.nz + k;
    const Cursor<TQ> cu = make_cursor<TQ>(g, i, j, k);
    const uint32_t mt = face_mask<TQ>(g, i, j, k);
    typename L::T* tl = static_cast<typename L::T*>(tlam);
    static_for<0, TQ>([&](auto I) {
        constexpr int q = decltype(I)::value;
        store_read<TQ, q>(tl, g, cu, mt, odd, L::template enc<double>(aos[cell * TQ + q]));
    });
}

}  // namespace lbm
