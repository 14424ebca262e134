// Binary16 non-equilibrium storage of the populations.
//
// The binary16 form stores f~ = f - w (the reference's shifted population,
// src/cuda/precision.cuh:29-42) rounded to binary16.  In a jet f~ carries
// 3 w c.u (0.15 w at u = 0.05), so each store rounds away ~4e-5 w, and the
// shear-layer instability amplifies that seed until the flow parts from the
// FP64 one within a few thousand steps.  This form stores instead
//     v_q = K (f~_q - S_q(M)),   S_q(M) = w_q (A + b (1 + b / 2)),  b = c_q . B,
// with M = (A, B) = (rho (1 - 1.5 |u|^2) - 1, 3 rho u) of the cell that
// wrote the population: S_q is the second-order equilibrium of that cell
// (f_eq - w of D3Q19.cuh:143-150 up to O(u^2 (rho - 1)) and the third-order
// terms of D3Q27.cuh:152-159), so v_q is the small non-equilibrium
// remainder.  K = 2^12 lifts the remainder out of binary16's subnormal range
// (below 6.1e-5, where the absolute spacing 6e-8 is as coarse as the
// rounding of f~ itself).  Each cell keeps its M of the step, rounded to
// binary16 (8 B), and the reader of a population adds S_q of its writer
// back:
//     f~_q(x) = v_q / K + S_q(M(x - c_q)).
// The shift needs only to be the same on both sides, not exact: reader and
// writer evaluate it from the same stored M in the same operation order
// (ne_shift per direction; the writer's ne_sub_pair, two directions of an
// opposite pair at a time, gives the same bits).
#pragma once

#include <hip/hip_runtime.h>

#include "kparams.hpp"
#include "lattice.hpp"

namespace pf {

// M of a cell with density rho and velocity u, rounded to binary16 (the
// stored form: 8 B per cell; the shift needs only be the same for writer and
// reader, and the rounding moves it by ~1e-4 of the equilibrium, which the
// stored remainder absorbs).  `packed` is what the cell stores.
struct NeM {
    float4 m;
    uint2 packed;
};
__device__ __forceinline__ NeM ne_mom(float rho, float ux, float uy, float uz) {
#pragma clang fp contract(off)
    const float v = 1.0f - 1.5f * (ux * ux + uy * uy + uz * uz);
    const float r3 = 3.0f * rho;
    const _Float16 a = (_Float16)(rho * v - 1.0f), bx = (_Float16)(r3 * ux), by = (_Float16)(r3 * uy),
                   bz = (_Float16)(r3 * uz);
    NeM r;
    r.m = make_float4((float)a, (float)bx, (float)by, (float)bz);
    r.packed = make_uint2((uint32_t)__builtin_bit_cast(uint16_t, a) | ((uint32_t)__builtin_bit_cast(uint16_t, bx) << 16),
                          (uint32_t)__builtin_bit_cast(uint16_t, by) | ((uint32_t)__builtin_bit_cast(uint16_t, bz) << 16));
    return r;
}
__device__ __forceinline__ float4 ne_unpack(uint2 v) {
    return make_float4((float)__builtin_bit_cast(_Float16, (uint16_t)(v.x & 0xffffu)),
                       (float)__builtin_bit_cast(_Float16, (uint16_t)(v.x >> 16)),
                       (float)__builtin_bit_cast(_Float16, (uint16_t)(v.y & 0xffffu)),
                       (float)__builtin_bit_cast(_Float16, (uint16_t)(v.y >> 16)));
}

// b = c.B for a compile-time direction (signs folded, zero terms dropped).
template <int cx, int cy, int cz>
__device__ __forceinline__ float ne_cdot(const float4& m) {
    float s = 0.0f;
    bool any = false;
    if constexpr (cx != 0) {
        s = cx > 0 ? m.y : -m.y;
        any = true;
    }
    if constexpr (cy != 0) {
        s = any ? (cy > 0 ? s + m.z : s - m.z) : (cy > 0 ? m.z : -m.z);
        any = true;
    }
    if constexpr (cz != 0) s = any ? (cz > 0 ? s + m.w : s - m.w) : (cz > 0 ? m.w : -m.w);
    return s;
}

// S_q(M) for direction q: w (A + b (1 + b/2)) as w fma(b, fma(0.5, b, 1), A),
// with every rounding spelled out (contraction off), so that it has the
// same bits wherever it is evaluated.
template <int Q, int q>
__device__ __forceinline__ float ne_shift(const float4& m) {
#pragma clang fp contract(off)
    constexpr float w = wq<Q>(q);
    if constexpr (q == 0) {
        return w * m.x;
    } else {
        const float b = ne_cdot<kCX[q], kCY[q], kCZ[q]>(m);
        return w * fmaf(b, fmaf(0.5f, b, 1.0f), m.x);
    }
}

// The writer's side for an opposite pair: a = direction q (c) and b = its
// opposite (-c), with op and om the pair's outputs scaled by K:
//     op -= K S_a(M),   om -= K S_b(M).
// b of the opposite direction is -b exactly, and K = 2^12 scales without
// rounding, so the results are K (out - ne_shift) bit for bit, for one
// c . B per pair.
template <int Q, int q>
__device__ __forceinline__ void ne_sub_pair(const float4& m, float& op, float& om) {
#pragma clang fp contract(off)
    constexpr float kw = kNeK * wq<Q>(q);
    const float b = ne_cdot<kCX[q], kCY[q], kCZ[q]>(m);
    op = op - kw * fmaf(b, fmaf(0.5f, b, 1.0f), m.x);
    om = om - kw * fmaf(-b, fmaf(-0.5f, b, 1.0f), m.x);
}

}  // namespace pf
