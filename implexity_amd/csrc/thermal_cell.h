// One cell of the coupled D3Q7 temperature field and the storage of its populations
// (the model is described in thermal.h).  Included by kernels.h so that the flow's
// forward kernel can advance the temperature of the same cell in the same pass; it is
// included inside namespace lbm, after the addressing helpers.
#pragma once

constexpr int TQ = 7;

// Temperature population storage: FP64 / FP32 as is; FP16 times a power of two chosen
// on the host from the expected temperature range (tenc = 2^e, tdec = 2^-e).
template <int P>
struct TStore;
template <>
struct TStore<P_F64> {
    using T = double;
    using C = double;
    static LBM_DEV C dec(T v, float) { return v; }
    static LBM_DEV T enc(C v, float) { return v; }
};
template <>
struct TStore<P_F32> {
    using T = float;
    using C = float;
    static LBM_DEV C dec(T v, float) { return v; }
    static LBM_DEV T enc(C v, float) { return v; }
};
template <>
struct TStore<P_F16S> {
    using T = _Float16;
    using C = float;
    static LBM_DEV C dec(T v, float s) { return float(v) * s; }
    static LBM_DEV T enc(C v, float s) { return _Float16(v * s); }
};

// Store of a temperature population written by a step: stochastic rounding for FP16
// (bits from the cell, the direction and the step, so recomputed steps reproduce it).
template <int P>
LBM_DEV typename TStore<P>::T tstore_step(typename TStore<P>::C v, float s, uint32_t key) {
    if constexpr (P == P_F16S)
        return f16_stochastic(v * s, mix32(key));
    else
        return TStore<P>::enc(v, s);
}

// omega_T(phi) = 1 / (1/2 + 4 kappa_f (1 + (ratio - 1) phi))
template <class C>
LBM_DEV C thermal_omega(double kappa, double ratio_m1, C phi) {
    return C(1) / (C(0.5) + C(4) * (C(kappa) * (C(1) + C(ratio_m1) * phi)));
}

template <class C>
LBM_DEV C thermal_cu(const C u[3], int q) {
    return u[0] * C(cvec<TQ>(q, 0)) + u[1] * C(cvec<TQ>(q, 1)) + u[2] * C(cvec<TQ>(q, 2));
}

// One cell: populations g, flow velocity u -> post-collision populations; T = sum g.
template <class C>
LBM_DEV void thermal_collide(const C* gv, const C u[3], C om, C S, C* out, C& T) {
    T = C(0);
#pragma unroll
    for (int q = 0; q < TQ; q++) T += gv[q];
#pragma unroll
    for (int q = 0; q < TQ; q++) {
        const C w = C(weight<TQ>(q));
        const C eq = w * T * (C(1) + C(4) * thermal_cu(u, q));
        out[q] = gv[q] - om * (gv[q] - eq) + w * S;
    }
}

// Transpose of thermal_collide: gb = cotangent of the output.
template <class C>
LBM_DEV void thermal_collide_vjp(const C* gv, const C u[3], C om, const C* gb, C* g_bar, C u_bar[3], C& om_bar,
                                 C& S_bar) {
    C T = C(0);
#pragma unroll
    for (int q = 0; q < TQ; q++) T += gv[q];
    C t_bar = C(0), ob = C(0), sb = C(0);
    u_bar[0] = u_bar[1] = u_bar[2] = C(0);
    static_for<0, TQ>([&](auto I) {
        constexpr int q = decltype(I)::value;
        const C w = C(weight<TQ>(q));
        const C cu = thermal_cu(u, q);
        const C eq = w * T * (C(1) + C(4) * cu);
        const C eq_bar = om * gb[q];
        g_bar[q] = (C(1) - om) * gb[q];
        ob += gb[q] * (eq - gv[q]);
        sb += gb[q] * w;
        t_bar += eq_bar * w * (C(1) + C(4) * cu);
        const C cu_bar = eq_bar * w * T * C(4);
        if constexpr (cvec<TQ>(q, 0) != 0) u_bar[0] += cu_bar * C(cvec<TQ>(q, 0));
        if constexpr (cvec<TQ>(q, 1) != 0) u_bar[1] += cu_bar * C(cvec<TQ>(q, 1));
        if constexpr (cvec<TQ>(q, 2) != 0) u_bar[2] += cu_bar * C(cvec<TQ>(q, 2));
    });
#pragma unroll
    for (int q = 0; q < TQ; q++) g_bar[q] += t_bar;
    om_bar = ob;
    S_bar = sb;
}

template <int P>
LBM_DEV void load_tcell(const Geom& g, const typename TStore<P>::T* tp, const Cursor<TQ>& cu, uint32_t mt, int odd,
                        float tdec, typename TStore<P>::C* gv) {
    static_for<0, TQ>([&](auto I) {
        constexpr int q = decltype(I)::value;
        gv[q] = TStore<P>::dec(load_read<TQ, q>(tp, g, cu, mt, odd), tdec);
    });
}
