// Device-side lattice tables and per-cell LBM algebra for gfx950.
//
// Directions, weights and the moment algebra follow the reference velocity
// sets: src/velocitySet/D3Q19.cuh:55-187, src/velocitySet/D3Q27.cuh:55-206,
// src/velocitySet/D3Q7.cuh:55-153.  Every loop over directions is expanded at
// compile time (the role of device::constexpr_for, src/functions/constexprFor.cuh),
// so each direction's velocity components are immediate constants and zero
// terms disappear.  The order of floating-point additions matches the
// reference's written sums, so results differ from a straight evaluation
// only by FMA contraction.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <utility>

namespace pf {

// Direction tables in the reference's ordering (D3Q7 and D3Q19 are prefixes
// of D3Q27).  Opposite directions are the pairs (2j-1, 2j).
constexpr int kCX[27] = {0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 0, 0, 1, -1,
                                             1, -1, 0, 0, 1, -1, 1, -1, 1, -1, -1, 1};
constexpr int kCY[27] = {0, 0, 0, 1, -1, 0, 0, 1, -1, 0, 0, 1, -1, -1, 1,
                                             0, 0, 1, -1, 1, -1, 1, -1, -1, 1, 1, -1};
constexpr int kCZ[27] = {0, 0, 0, 0, 0, 1, -1, 0, 0, 1, -1, 1, -1, 0, 0,
                                             -1, 1, -1, 1, 1, -1, -1, 1, 1, -1, 1, -1};

__host__ __device__ constexpr int opp(int q) { return q == 0 ? 0 : ((q & 1) ? q + 1 : q - 1); }
__host__ __device__ constexpr int wclass(int q) { return q == 0 ? 0 : (q <= 6 ? 1 : (q <= 18 ? 2 : 3)); }
__host__ __device__ constexpr int comp(int axis, int q) {
    return axis == 0 ? kCX[q] : (axis == 1 ? kCY[q] : kCZ[q]);
}

// Weights: the reference casts double quotients to float.
template <int Q>
__host__ __device__ constexpr float wk(int k) {
    if constexpr (Q == 19) {
        return k == 0 ? (float)(1.0 / 3.0) : k == 1 ? (float)(1.0 / 18.0) : k == 2 ? (float)(1.0 / 36.0) : 0.0f;
    } else if constexpr (Q == 27) {
        return k == 0 ? (float)(8.0 / 27.0)
                      : k == 1 ? (float)(2.0 / 27.0) : k == 2 ? (float)(1.0 / 54.0) : (float)(1.0 / 216.0);
    } else {
        return k == 0 ? 0.25f : k == 1 ? 0.125f : 0.0f;
    }
}
template <int Q>
__host__ __device__ constexpr float wq(int q) { return wk<Q>(wclass(q)); }

constexpr float kCs2 = (float)(1.0 / 3.0);  // D3Q19/D3Q27
constexpr float kAs2 = 3.0f;
constexpr float kAs2g = 4.0f;               // D3Q7

// Compile-time loop.
template <class F, int... I>
__device__ __forceinline__ void static_for_impl(F&& f, std::integer_sequence<int, I...>) {
    (f(std::integral_constant<int, I>{}), ...);
}
template <int N, class F>
__device__ __forceinline__ void static_for(F&& f) {
    static_for_impl(f, std::make_integer_sequence<int, N>{});
}

// First opposite pair (odd member) with c_axis != 0 in weight class K
// (K < 0: any class), or 0 when none.
template <int Q, int AXIS, int K>
__host__ __device__ constexpr int first_pair() {
    for (int p = 1; p < Q; p += 2)
        if ((K < 0 || wclass(p) == K) && comp(AXIS, p) != 0) return p;
    return 0;
}

// Pair-ordered first moment: for each opposite pair in increasing order with
// c_axis != 0, add the sample of the +1 member and subtract the sample of the
// -1 member, left to right.  This is the order of the reference's hand-written
// sums (src/lbm.cu:72-80, src/phaseField.cuh:82-122 gradient, :168-210 divergence).
template <int Q, int AXIS, int K, class V>
__device__ __forceinline__ float pair_chain(V&& val) {
    constexpr int first = first_pair<Q, AXIS, K>();
    float acc = 0.0f;
    static_for<(Q - 1) / 2>([&](auto jc) {
        constexpr int p = 2 * decltype(jc)::value + 1;
        if constexpr ((K < 0 || wclass(p) == K) && comp(AXIS, p) != 0) {
            constexpr int pos = comp(AXIS, p) > 0 ? p : p + 1;
            constexpr int neg = comp(AXIS, p) > 0 ? p + 1 : p;
            if constexpr (p == first) {
                acc = val(std::integral_constant<int, pos>{}) - val(std::integral_constant<int, neg>{});
            } else {
                acc = acc + val(std::integral_constant<int, pos>{});
                acc = acc - val(std::integral_constant<int, neg>{});
            }
        }
    });
    return acc;
}

// Isotropic stencil sum along AXIS (src/phaseField.cuh:82-122, :168-210): weight class
// 1, then 2, then (D3Q27) 3.  The caller multiplies by as2.
template <int Q, int AXIS, class V>
__device__ __forceinline__ float iso_sum(V&& val) {
    float s = wk<Q>(1) * pair_chain<Q, AXIS, 1>(val) + wk<Q>(2) * pair_chain<Q, AXIS, 2>(val);
    if constexpr (Q == 27) s = s + wk<Q>(3) * pair_chain<Q, AXIS, 3>(val);
    return s;
}

template <int Q, int q>
__device__ __forceinline__ float cu_of(float ux, float uy, float uz) {
    constexpr float cx = (float)kCX[q], cy = (float)kCY[q], cz = (float)kCZ[q];
    float s = 0.0f;
    // c.u with zero components skipped (adding an exact +0 is an identity)
    bool any = false;
    if constexpr (kCX[q] != 0) { s = cx * ux; any = true; }
    if constexpr (kCY[q] != 0) { s = any ? s + cy * uy : cy * uy; any = true; }
    if constexpr (kCZ[q] != 0) { s = any ? s + cz * uz : cz * uz; }
    return kAs2 * s;
}

// Shifted equilibrium f_eq - w (D3Q19.cuh:143-150, D3Q27.cuh:152-159).
template <int Q, int q>
__device__ __forceinline__ float f_eq(float rho, float uu, float cu) {
    constexpr float w = wq<Q>(q);
    if constexpr (Q == 19) {
        return w * rho * (1.0f - uu + cu + 0.5f * cu * cu) - w;
    } else {
        constexpr float oos = (float)(1.0 / 6.0);
        return w * rho * (1.0f - uu + cu + 0.5f * cu * cu + oos * cu * cu * cu - uu * cu) - w;
    }
}

// Guo forcing (D3Q19.cuh:173-187).
template <int Q, int q>
__device__ __forceinline__ float force_of(float cu, float ux, float uy, float uz, float fx, float fy, float fz) {
    constexpr float w = wq<Q>(q);
    constexpr float cx = (float)kCX[q], cy = (float)kCY[q], cz = (float)kCZ[q];
    return w * ((kAs2 * (cx - ux) + kAs2 * cu * cx) * fx + (kAs2 * (cy - uy) + kAs2 * cu * cy) * fy +
                (kAs2 * (cz - uz) + kAs2 * cu * cz) * fz);
}

// Regularized non-equilibrium population from the deviatoric stress
// (D3Q19.cuh:152-171 second order; D3Q27.cuh:161-190 adds the third-order
// Hermite terms).  Coefficients are folded at compile time; the terms with a
// zero coefficient contribute exact zeros in the reference and are dropped.
template <int Q, int q>
__device__ __forceinline__ float f_neq(float pxx, float pyy, float pzz, float pxy, float pxz, float pyz, float ux,
                                       float uy, float uz) {
    constexpr int cx = kCX[q], cy = kCY[q], cz = kCZ[q];
    constexpr float A = (float)(cx * cx) - kCs2, B = (float)(cy * cy) - kCs2, C = (float)(cz * cz) - kCs2;
    float s = A * pxx + B * pyy + C * pzz;
    {
        float t = 0.0f;
        bool any = false;
        if constexpr (cx * cy != 0) { t = (float)(cx * cy) * pxy; any = true; }
        if constexpr (cx * cz != 0) { t = any ? t + (float)(cx * cz) * pxz : (float)(cx * cz) * pxz; any = true; }
        if constexpr (cy * cz != 0) { t = any ? t + (float)(cy * cz) * pyz : (float)(cy * cz) * pyz; any = true; }
        if (any) s = s + 2.0f * t;
    }
    if constexpr (Q == 27) {
        constexpr float k1 = (float)(cx * cx * cy) - kCs2 * (float)cy;
        constexpr float k2 = (float)(cx * cx * cz) - kCs2 * (float)cz;
        constexpr float k3 = (float)(cx * cy * cy) - kCs2 * (float)cx;
        constexpr float k4 = (float)(cy * cy * cz) - kCs2 * (float)cz;
        constexpr float k5 = (float)(cx * cz * cz) - kCs2 * (float)cx;
        constexpr float k6 = (float)(cy * cz * cz) - kCs2 * (float)cy;
        float t = 0.0f;
        bool any = false;
        auto acc = [&](float k, float v) {
            t = any ? t + k * v : k * v;
            any = true;
        };
        if constexpr (k1 != 0.0f) acc(k1, pxx * uy + 2.0f * ux * pxy);
        if constexpr (k2 != 0.0f) acc(k2, pxx * uz + 2.0f * ux * pxz);
        if constexpr (k3 != 0.0f) acc(k3, pyy * ux + 2.0f * uy * pxy);
        if constexpr (k4 != 0.0f) acc(k4, pyy * uz + 2.0f * uy * pyz);
        if constexpr (k5 != 0.0f) acc(k5, pzz * ux + 2.0f * uz * pxz);
        if constexpr (k6 != 0.0f) acc(k6, pzz * uy + 2.0f * uz * pyz);
        if (any) s = s + kAs2 * t;
        if constexpr (cx * cy * cz != 0) s = s + (float)(6 * cx * cy * cz) * (ux * pyz + uy * pxz + uz * pxy);
    }
    return (wq<Q>(q) * 4.5f) * s;
}

// D3Q7 equilibrium plus sharpening flux (D3Q7.cuh:129-153).
template <int q>
__device__ __forceinline__ float g_post(float phi, float ux, float uy, float uz, float sharp, float nx, float ny,
                                        float nz) {
    constexpr float w = wq<7>(q);
    float cu = 0.0f, cn = 0.0f;
    if constexpr (kCX[q] != 0) { cu = (float)kCX[q] * ux; cn = (float)kCX[q] * nx; }
    if constexpr (kCY[q] != 0) { cu = (float)kCY[q] * uy; cn = (float)kCY[q] * ny; }
    if constexpr (kCZ[q] != 0) { cu = (float)kCZ[q] * uz; cn = (float)kCZ[q] * nz; }
    const float geq = w * phi * (1.0f + kAs2g * cu);
    if constexpr (q == 0) {
        return geq;  // the rest direction carries no sharpening flux
    } else {
        return geq + w * sharp * cn;
    }
}

}  // namespace pf
