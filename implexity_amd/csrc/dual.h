// Dual numbers for forward-mode derivatives on the device: the tangent step (jvp.h),
// which runs the collision, closure, sponge and port code on them.
#pragma once
#include <hip/hip_runtime.h>

#include "cell.h"

namespace lbm {

// a + eps b with eps^2 = 0: every operation propagates the derivative b of its value a.
struct Dual {
    double v, d;
    __host__ __device__ constexpr Dual() : v(0.0), d(0.0) {}
    __host__ __device__ constexpr Dual(double x) : v(x), d(0.0) {}
    __host__ __device__ constexpr Dual(double x, double t) : v(x), d(t) {}
    __host__ __device__ explicit constexpr operator double() const { return v; }
    // d sqrt(x) = dx / (2 sqrt(x)), the value the FP64 square root (the Smagorinsky
    // closure).  A hidden friend: found for Dual arguments only, so it does not hide
    // ::sqrt for the other arithmetic types.
    friend LBM_DEV Dual sqrt(const Dual& a) {
        const double r = ::sqrt(a.v);
        return Dual(r, a.d / (2.0 * r));
    }
    // a^p for a constant exponent p (the WALE closure's powers 1.25, 1.5, 2.5 of
    // non-negative invariants): d(a^p) = p a^(p-1) da; the tangent of p is not used.
    friend LBM_DEV Dual pow(const Dual& a, const Dual& p) {
        return Dual(::pow(a.v, p.v), p.v * ::pow(a.v, p.v - 1.0) * a.d);
    }
    __host__ __device__ explicit constexpr operator float() const { return float(v); }
    LBM_DEV Dual& operator+=(const Dual& b) {
        v += b.v;
        d += b.d;
        return *this;
    }
    LBM_DEV Dual& operator-=(const Dual& b) {
        v -= b.v;
        d -= b.d;
        return *this;
    }
    LBM_DEV Dual& operator*=(const Dual& b) {
        const double bv = b.v, bd = b.d;
        d = d * bv + v * bd;
        v *= bv;
        return *this;
    }
    LBM_DEV Dual& operator/=(const Dual& b) {
        const double bv = b.v, bd = b.d;
        v /= bv;
        d = (d - v * bd) / bv;
        return *this;
    }
};

LBM_DEV Dual operator+(Dual a, const Dual& b) { return a += b; }
LBM_DEV Dual operator-(Dual a, const Dual& b) { return a -= b; }
LBM_DEV Dual operator*(Dual a, const Dual& b) { return a *= b; }
LBM_DEV Dual operator/(Dual a, const Dual& b) { return a /= b; }
LBM_DEV Dual operator-(const Dual& a) { return Dual(-a.v, -a.d); }
LBM_DEV bool operator<(const Dual& a, const Dual& b) { return a.v < b.v; }
LBM_DEV bool operator>(const Dual& a, const Dual& b) { return a.v > b.v; }
LBM_DEV bool operator<=(const Dual& a, const Dual& b) { return a.v <= b.v; }
LBM_DEV bool operator>=(const Dual& a, const Dual& b) { return a.v >= b.v; }
LBM_DEV bool operator==(const Dual& a, const Dual& b) { return a.v == b.v; }
LBM_DEV bool operator!=(const Dual& a, const Dual& b) { return a.v != b.v; }

}  // namespace lbm
