// Moment storage form (PF_STORE_MOM): each cell stores, in FP32, the
// coefficients of the polynomial in c that its regularized collision
// output is, and the next step rebuilds every arriving population from the
// coefficients of the cell that sent it.
//
// The reference's post-collision population (src/lbm.cu:206, with
// D3Q19.cuh:143-187 / D3Q27.cuh:152-206)
//     f~*_q = f~eq_q(rho, u) + (1 - omega) f_neq_q(P, u) + F_q / 2
// is w_q times a polynomial of degree <= 3 in c_q whose coefficients depend on
// the sending cell alone.  With c_i in {-1, 0, 1} (so c_i^3 = c_i) it has 10
// monomials on D3Q19 and 17 on D3Q27:
//     A + sum_i B_i c_i + sum_i G_i c_i^2 + sum_{i<j} D_ij c_i c_j
//       [D3Q27: + sum_{i != j} E_iij c_i^2 c_j + Z c_x c_y c_z]
// With rho, u, the Guo force F, the deviatoric stress P and omega of the
// sending cell (t = c.u, s = c.F, V = 1 - 1.5|u|^2, O = 4.5 (1 - omega)):
//     A     = rho V - 1 - 1.5 u.F
//     G_i   = 4.5 rho u_i^2 + 4.5 u_i F_i + O P_ii
//     D_ij  = 9 rho u_i u_j + 4.5 (u_i F_j + u_j F_i) + 2 O P_ij
//     B_i   = 3 rho u_i + 1.5 F_i                                   (D3Q19)
//     B_i   = 3 rho (V u_i + 1.5 u_i^3) + 1.5 F_i - O (X_a + X_b)    (D3Q27)
//     E_iij = 13.5 rho u_i^2 u_j + 3 O X_iij,  Z = 27 rho u_x u_y u_z + O X7s
// where X_iij = P_ii u_j + 2 u_i P_ij are the third-order Hermite blocks of
// D3Q27.cuh:161-190 and X7s = 6 (u_x P_yz + u_y P_xz + u_z P_xy).
//
// A step moves 2 x 17 x 4 = 136 B of coefficients per D3Q27 cell instead of
// 2 x 27 x 4 = 216 B of populations, with FP32 arithmetic and FP32 storage
// throughout: the rebuilt populations equal the stored ones of the
// population form up to the rounding of a different summation order.
#pragma once

#include "kparams.hpp"
#include "lattice.hpp"

namespace pf {

// Coefficient slots.
enum : int {
    kMoA = 0,
    kMoBx, kMoBy, kMoBz,
    kMoGx, kMoGy, kMoGz,
    kMoDxy, kMoDxz, kMoDyz,
    kMoExxy, kMoExxz, kMoEyyx, kMoEyyz, kMoEzzx, kMoEzzy, kMoZ,
};
template <int Q>
__host__ __device__ constexpr int mom_nc() { return Q == 27 ? 17 : 10; }

// Coefficient arrays are interleaved per (y, z) row, as the population
// arrays: coefficient k of cell (x, y, z) sits at
// z*NC*nx*ny + y*NC*nx + k*nx + x, so the NC coefficients of a row of cells
// are NC contiguous runs.  Accessed from a per-plane buffer resource with
// y*NC*nx + x as the 32-bit vector offset and k*nx as the scalar offset.
template <int Q>
__device__ __forceinline__ uint64_t rec_elem(uint32_t k, uint32_t c, uint32_t nx, uint32_t ny) {
    const uint32_t plane = nx * ny, z = c / plane, rem = c - z * plane, y = rem / nx, x = rem - y * nx;
    return ((uint64_t)z * ny + y) * ((uint64_t)mom_nc<Q>() * nx) + (uint64_t)k * nx + x;
}
template <int Q>
__device__ __forceinline__ uint32_t rec_off(uint32_t x, uint32_t y, uint32_t nx) {
    return y * (uint32_t)mom_nc<Q>() * nx + x;
}
template <int Q>
__device__ __forceinline__ __amdgpu_buffer_rsrc_t rec_rsrc(const float* base, uint32_t plane, uint32_t z) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<float*>(base) + (uint64_t)z * (uint32_t)mom_nc<Q>() * plane,
                                             (short)0, (int)((uint32_t)mom_nc<Q>() * plane * 4u), 0x00020000);
}
// Element `off` of run `slot` (runs `stride` elements apart) of a buffer.
__device__ __forceinline__ float slot_ld(__amdgpu_buffer_rsrc_t r, uint32_t off, uint32_t slot, uint32_t stride) {
    return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, off * 4u, slot * stride * 4u, 0));
}

// Raw moments of the populations arriving at a cell: sum f~, sum f~ c,
// sum f~ c c (src/lbm.cu:49-115).
struct MomAcc {
    float m0, jx, jy, jz, xx, yy, zz, xy, xz, yz;
};
__device__ __forceinline__ MomAcc mom_zero() { return {0, 0, 0, 0, 0, 0, 0, 0, 0, 0}; }

// In-plane partial sums of the (up to) 9 populations of one c_z layer:
// S = sum f, X = sum c_x f, Y = sum c_y f, XX = sum c_x^2 f, YY, XY.
struct MomLayer {
    float S, X, Y, XX, YY, XY;
    bool any;
};
template <int cx, int cy>
__device__ __forceinline__ void layer_add(MomLayer& g, float f) {
    if (!g.any) {
        g.S = f;
        g.X = cx == 0 ? 0.0f : (cx > 0 ? f : -f);
        g.Y = cy == 0 ? 0.0f : (cy > 0 ? f : -f);
        g.XX = cx == 0 ? 0.0f : f;
        g.YY = cy == 0 ? 0.0f : f;
        g.XY = (cx * cy == 0) ? 0.0f : (cx * cy > 0 ? f : -f);
        g.any = true;
        return;
    }
    g.S += f;
    if constexpr (cx != 0) {
        g.X = cx > 0 ? g.X + f : g.X - f;
        g.XX += f;
    }
    if constexpr (cy != 0) {
        g.Y = cy > 0 ? g.Y + f : g.Y - f;
        g.YY += f;
    }
    if constexpr (cx * cy != 0) g.XY = cx * cy > 0 ? g.XY + f : g.XY - f;
}
// Adds a finished c_z layer to the raw moments of its reader.
template <int cz>
__device__ __forceinline__ void layer_to(MomAcc& a, const MomLayer& g) {
    a.m0 += g.S;
    a.jx += g.X;
    a.jy += g.Y;
    a.xx += g.XX;
    a.yy += g.YY;
    a.xy += g.XY;
    if constexpr (cz != 0) {
        a.jz = cz > 0 ? a.jz + g.S : a.jz - g.S;
        a.zz += g.S;
        a.xz = cz > 0 ? a.xz + g.X : a.xz - g.X;
        a.yz = cz > 0 ? a.yz + g.Y : a.yz - g.Y;
    }
}

// Whether direction (cx, cy, cz) belongs to the velocity set.
template <int Q>
__host__ __device__ constexpr bool mom_has(int cx, int cy, int cz) {
    const int n = (cx != 0) + (cy != 0) + (cz != 0);
    return Q == 27 || n <= 2;
}

template <int s>
__device__ __forceinline__ float sgn_add(float a, float b) {
    return s > 0 ? a + b : a - b;
}

// The three populations (c_x, c_y, c_z = -1, 0, +1) sent by one cell, from its
// coefficients R(k): the c_z-free part P0, the parts odd (A1) and even (A2)
// in c_z, then f(c_z) = w (P0 + c_z A1 + c_z^2 A2).  WM / W0 / WP select the
// outputs wanted.
template <int Q, int cx, int cy, bool WM, bool W0, bool WP, class R>
__device__ __forceinline__ void mom_column(R&& r, float& fm, float& f0, float& fp) {
    constexpr int n0 = (cx != 0) + (cy != 0);
    float p0 = r(kMoA);
    if constexpr (cx != 0) p0 = sgn_add<cx>(p0, r(kMoBx));
    if constexpr (cy != 0) p0 = sgn_add<cy>(p0, r(kMoBy));
    if constexpr (cx != 0) p0 = p0 + r(kMoGx);
    if constexpr (cy != 0) p0 = p0 + r(kMoGy);
    if constexpr (cx * cy != 0) p0 = sgn_add<cx * cy>(p0, r(kMoDxy));
    if constexpr (Q == 27) {
        if constexpr (cx != 0 && cy != 0) {
            p0 = sgn_add<cy>(p0, r(kMoExxy));
            p0 = sgn_add<cx>(p0, r(kMoEyyx));
        }
    }
    if constexpr (W0) f0 = wk<Q>(n0) * p0;
    if constexpr ((WM || WP) && mom_has<Q>(cx, cy, 1)) {
        float a1 = r(kMoBz);
        if constexpr (cx != 0) a1 = sgn_add<cx>(a1, r(kMoDxz));
        if constexpr (cy != 0) a1 = sgn_add<cy>(a1, r(kMoDyz));
        float a2 = r(kMoGz);
        if constexpr (Q == 27) {
            if constexpr (cx != 0) a1 = a1 + r(kMoExxz);
            if constexpr (cy != 0) a1 = a1 + r(kMoEyyz);
            if constexpr (cx * cy != 0) a1 = sgn_add<cx * cy>(a1, r(kMoZ));
            if constexpr (cx != 0) a2 = sgn_add<cx>(a2, r(kMoEzzx));
            if constexpr (cy != 0) a2 = sgn_add<cy>(a2, r(kMoEzzy));
        }
        const float e = p0 + a2;
        constexpr float w1 = wk<Q>(n0 + 1);
        if constexpr (WP) fp = w1 * (e + a1);
        if constexpr (WM) fm = w1 * (e - a1);
    }
}

// Population of direction q sent by the cell with coefficients R(k), by
// the same operations as mom_column.
template <int Q, int q, class R>
__device__ __forceinline__ float mom_pop(R&& r) {
    constexpr int cx = kCX[q], cy = kCY[q], cz = kCZ[q];
    float fm = 0.0f, f0 = 0.0f, fp = 0.0f;
    mom_column<Q, cx, cy, (cz < 0), (cz == 0), (cz > 0)>(r, fm, f0, fp);
    return cz < 0 ? fm : (cz == 0 ? f0 : fp);
}

// Coefficients of one cell's collision output from its density (with
// m0 = rho - 1 as summed), velocity, force, deviatoric stress and 1 - omega.
template <int Q>
__device__ __forceinline__ void mom_coeffs(float m0, float rho, float ux, float uy, float uz, float fx, float fy,
                                           float fz, float pxx, float pyy, float pzz, float pxy, float pxz,
                                           float pyz, float omco, float (&c)[mom_nc<Q>()]) {
    const float usq = ux * ux + uy * uy + uz * uz;
    const float uF = ux * fx + uy * fy + uz * fz;
    const float R3 = 3.0f * rho, R45 = 4.5f * rho, O45 = 4.5f * omco;
    const float R9 = 2.0f * R45;
    c[kMoA] = m0 - 1.5f * (rho * usq + uF);
    c[kMoGx] = R45 * (ux * ux) + 4.5f * (ux * fx) + O45 * pxx;
    c[kMoGy] = R45 * (uy * uy) + 4.5f * (uy * fy) + O45 * pyy;
    c[kMoGz] = R45 * (uz * uz) + 4.5f * (uz * fz) + O45 * pzz;
    c[kMoDxy] = R9 * (ux * uy) + 4.5f * (ux * fy + uy * fx) + O45 * (2.0f * pxy);
    c[kMoDxz] = R9 * (ux * uz) + 4.5f * (ux * fz + uz * fx) + O45 * (2.0f * pxz);
    c[kMoDyz] = R9 * (uy * uz) + 4.5f * (uy * fz + uz * fy) + O45 * (2.0f * pyz);
    if constexpr (Q == 19) {
        c[kMoBx] = R3 * ux + 1.5f * fx;
        c[kMoBy] = R3 * uy + 1.5f * fy;
        c[kMoBz] = R3 * uz + 1.5f * fz;
    } else {
        const float V = 1.0f - 1.5f * usq;
        const float X1 = pxx * uy + 2.0f * ux * pxy, X2 = pxx * uz + 2.0f * ux * pxz;
        const float X3 = pyy * ux + 2.0f * uy * pxy, X4 = pyy * uz + 2.0f * uy * pyz;
        const float X5 = pzz * ux + 2.0f * uz * pxz, X6 = pzz * uy + 2.0f * uz * pyz;
        const float X7s = 6.0f * (ux * pyz + uy * pxz + uz * pxy);
        c[kMoBx] = R3 * (V * ux + 1.5f * (ux * ux * ux)) + 1.5f * fx - O45 * (X3 + X5);
        c[kMoBy] = R3 * (V * uy + 1.5f * (uy * uy * uy)) + 1.5f * fy - O45 * (X1 + X6);
        c[kMoBz] = R3 * (V * uz + 1.5f * (uz * uz * uz)) + 1.5f * fz - O45 * (X2 + X4);
        const float R135 = 4.5f * R3;
        c[kMoExxy] = R135 * (ux * ux * uy) + 3.0f * O45 * X1;
        c[kMoExxz] = R135 * (ux * ux * uz) + 3.0f * O45 * X2;
        c[kMoEyyx] = R135 * (uy * uy * ux) + 3.0f * O45 * X3;
        c[kMoEyyz] = R135 * (uy * uy * uz) + 3.0f * O45 * X4;
        c[kMoEzzx] = R135 * (uz * uz * ux) + 3.0f * O45 * X5;
        c[kMoEzzy] = R135 * (uz * uz * uy) + 3.0f * O45 * X6;
        c[kMoZ] = 9.0f * R3 * (ux * uy * uz) + O45 * X7s;
    }
}

// Raw moments of the Q populations of one cell's own coefficients (the
// first step after initialisation reads each cell's own equilibrium, as the
// reference's setDistros leaves it, src/initialConditions.cu:101-138).
template <int Q, class R>
__device__ __forceinline__ MomAcc mom_local(R&& r) {
    MomLayer gm{0, 0, 0, 0, 0, 0, false}, g0 = gm, gp = gm;
    static_for<9>([&](auto oc) {
        constexpr int o = decltype(oc)::value, cx = o % 3 - 1, cy = o / 3 - 1;
        float fm = 0.0f, f0 = 0.0f, fp = 0.0f;
        constexpr bool hz = mom_has<Q>(cx, cy, 1);
        mom_column<Q, cx, cy, hz, true, hz>(r, fm, f0, fp);
        if constexpr (hz) {
            layer_add<cx, cy>(gm, fm);
            layer_add<cx, cy>(gp, fp);
        }
        layer_add<cx, cy>(g0, f0);
    });
    MomAcc a = mom_zero();
    layer_to<1>(a, gp);
    layer_to<0>(a, g0);
    layer_to<-1>(a, gm);
    return a;
}

}  // namespace pf
