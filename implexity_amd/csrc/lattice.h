// Lattice velocity sets, weights and the esoteric-twist storage offsets.
//
// D3Q19 order and weights follow crates/implexity-physics-lbm/src/d3q19.rs:29-61
// (rest first, then the lexicographic velocity list; OPPOSITE[i] = 19 - i).
// D3Q27 follows moving/lattice.rs:93-129 (rest first, then (a,b,c) in {-1,0,1}^3
// lexicographic without the origin; the opposite of i >= 1 is 27 - i).  D3Q7 is the
// lattice of the coupled temperature field (no counterpart in the reference), ordered
// the same way.
// Both orders must match bit for bit: the oracle and the reference index populations
// this way.  Everything here is a constexpr function of the direction index (no
// tables), so unrolled device loops fold to immediates.
#pragma once
#include <cstdint>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define LBM_HD __host__ __device__
#else
#define LBM_HD
#endif

namespace lbm {

// Velocity component d of direction q.
template <int Q>
LBM_HD constexpr int cvec(int q, int d) {
    if constexpr (Q == 7) {
        // The thermal D3Q7 lattice: rest, then the six axis vectors in lexicographic
        // order (opposite of i >= 1 is 7 - i), packed like D3Q19 below.
        constexpr int code[7] = {13, 4, 10, 12, 14, 16, 22};
        const int k = code[q];
        return d == 0 ? k / 9 - 1 : (d == 1 ? (k / 3) % 3 - 1 : k % 3 - 1);
    } else if constexpr (Q == 19) {
        // d3q19.rs:29-49, packed as (cx+1)*9 + (cy+1)*3 + (cz+1).
        constexpr int code[19] = {13, 1, 3, 4, 5, 7, 9, 10, 11, 12, 14, 15, 16, 17, 19, 21, 22, 23, 25};
        const int k = code[q];
        return d == 0 ? k / 9 - 1 : (d == 1 ? (k / 3) % 3 - 1 : k % 3 - 1);
    } else {
        if (q == 0) return 0;
        const int k = q <= 13 ? q - 1 : q;  // lexicographic index of (a,b,c) skipping the origin
        return d == 0 ? k / 9 - 1 : (d == 1 ? (k / 3) % 3 - 1 : k % 3 - 1);
    }
}

template <int Q>
LBM_HD constexpr int opp(int q) {
    return q == 0 ? 0 : Q - q;
}

// Storage position of population array q: the rest population first, then every
// opposite pair (q, Q - q), q < Q - q, in two adjacent arrays.  A bounce-back read
// chooses between the two arrays of one pair, so it is a select of a byte offset from
// one wave-uniform base instead of a select of 64-bit pointers.
template <int Q>
LBM_HD constexpr int apos(int q) {
    return q == 0 ? 0 : (q < Q - q ? 2 * q - 1 : 2 * (Q - q));
}

template <int Q>
LBM_HD constexpr int vorder(int q) {
    return (cvec<Q>(q, 0) != 0) + (cvec<Q>(q, 1) != 0) + (cvec<Q>(q, 2) != 0);
}

template <int Q>
LBM_HD constexpr double weight(int q) {
    if constexpr (Q == 7)
        return vorder<Q>(q) == 0 ? 0.25 : 0.125;  // c_s^2 = 1/4
    else if constexpr (Q == 19)
        return vorder<Q>(q) == 0 ? 1.0 / 3.0 : (vorder<Q>(q) == 1 ? 1.0 / 18.0 : 1.0 / 36.0);
    else
        return vorder<Q>(q) == 0 ? 8.0 / 27.0
                                 : (vorder<Q>(q) == 1 ? 2.0 / 27.0 : (vorder<Q>(q) == 2 ? 1.0 / 54.0 : 1.0 / 216.0));
}

// Exponent k of the power-of-two factor 2^k applied to (f_q - w_q) in FP16 storage.
// It keeps w_q * 2^k near 1e4, so |f_q - w_q| up to 6 w_q stays far inside the FP16
// normal range and deviations down to ~1e-9 stay normal.  An exact power of two
// changes only the exponent, never the stored significand.
template <int Q>
LBM_HD constexpr int fp16_exp(int q) {
    if constexpr (Q == 19)
        return vorder<Q>(q) == 0 ? 15 : (vorder<Q>(q) == 1 ? 18 : 19);
    else
        return vorder<Q>(q) == 0 ? 15 : (vorder<Q>(q) == 1 ? 17 : (vorder<Q>(q) == 2 ? 19 : 21));
}

// Esoteric-twist storage offsets, as a 3-bit corner of the unit cube at the cell
// (bit 0 = +x, bit 1 = +y, bit 2 = +z):  a_q = max(c_q, 0), a_opp(q) = max(-c_q, 0).
template <int Q>
LBM_HD constexpr int corner_pos(int q) {
    return (cvec<Q>(q, 0) > 0 ? 1 : 0) | (cvec<Q>(q, 1) > 0 ? 2 : 0) | (cvec<Q>(q, 2) > 0 ? 4 : 0);
}
template <int Q>
LBM_HD constexpr int corner_neg(int q) {
    return (cvec<Q>(q, 0) < 0 ? 1 : 0) | (cvec<Q>(q, 1) < 0 ? 2 : 0) | (cvec<Q>(q, 2) < 0 ? 4 : 0);
}

}  // namespace lbm
