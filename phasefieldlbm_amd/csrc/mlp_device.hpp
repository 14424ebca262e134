// Learned curvature closure: a 27 -> 64 -> 64 -> 1 MLP with GELU, BF16
// weights and FP32 accumulation on the gfx950 matrix cores
// (v_mfma_f32_16x16x32_bf16).  No counterpart in the reference, whose
// curvature is always the finite-difference div(n) (src/phaseField.cuh:142-219).
//
// MFMA 16x16x32 lane layout (lane = i + 16 g, i in 0..15, g in 0..3):
//   A (16 x 32): lane holds A[i][8g .. 8g+7]
//   B (32 x 16): lane holds B[8g .. 8g+7][i]
//   D (16 x 16): lane holds D[4g + r][i], r = 0..3
// Activations are kept transposed (hidden unit x sample), so the D tiles of
// one layer are the B operand of the next without leaving registers: for
// the K-block s of the second layer, lane group g supplies hidden units
// {32s + 4g + r, 32s + 16 + 4g + r}, and the W2 fragment is gathered in that
// same order.
#pragma once

#include <hip/hip_runtime.h>

#include "kparams.hpp"

namespace pf {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int kMlpInPad = 32;                                   // 27 features + zero padding
constexpr int kMlpW1 = kMlpHidden * kMlpInPad;                  // bf16 W1 [64][32]
constexpr int kMlpW2 = kMlpHidden * kMlpHidden;                 // bf16 W2 [64][64]
constexpr int kMlpWords = kMlpW1 + kMlpW2 + kMlpHidden;         // + w3 [64]
constexpr int kMlpBias = 2 * kMlpHidden + 1;                    // b1, b2, b3 (fp32)
constexpr int kMlpParams = kMlpHidden * kMlpIn + kMlpHidden + kMlpW2 + kMlpHidden + kMlpHidden + 1;  // 6017

__device__ __forceinline__ f32x4 mfma32(bf16x8 a, bf16x8 b, f32x4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

// tanh from the identity tanh(u) = 1 - 2 / (1 + e^(2u)), with the hardware
// base-2 exponential (e^(2u) = 2^(2u log2 e)) and reciprocal: a few
// instructions, relative error ~1e-6, far below the BF16 rounding of the
// network's operands.  |u| is clamped at 15, where tanh is 1 to FP32.
__device__ __forceinline__ float fast_tanh(float u) {
    const float a = fminf(fmaxf(u, -15.0f), 15.0f);
    const float e = __builtin_amdgcn_exp2f(a * 2.8853900817779268f);  // 2 log2(e)
    return 1.0f - 2.0f * __builtin_amdgcn_rcpf(1.0f + e);
}

// GELU, tanh form, and its derivative.  The value is evaluated through the
// identity 0.5 x (1 + tanh u) = x / (1 + e^(-2u)), u = sqrt(2/pi) (x +
// 0.044715 x^3): one exponential and one reciprocal.  For large negative x,
// e^(-2u) overflows to +inf and the result is x * 0, the limit; for large
// positive x it is x.
__device__ __forceinline__ float gelu(float x) {
    constexpr float k1 = -2.0f * 0.7978845608028654f * 1.4426950408889634f;  // -2 sqrt(2/pi) log2(e)
    constexpr float k3 = k1 * 0.044715f;
    const float v = x * (k1 + k3 * (x * x));
    return x * __builtin_amdgcn_rcpf(1.0f + __builtin_amdgcn_exp2f(v));
}
__device__ __forceinline__ float gelu_grad(float x) {
    const float u = 0.7978845608028654f * (x + 0.044715f * x * x * x);
    const float t = fast_tanh(u);
    return 0.5f * (1.0f + t) + 0.5f * x * (1.0f - t * t) * 0.7978845608028654f * (1.0f + 3.0f * 0.044715f * x * x);
}

// Feature f of the 3x3x3 stencil sits at offset (f%3-1, (f/3)%3-1, f/9-1).
// Normalised stencil: feature 13 (the centre) is 2*phi_c - 1, the others are
// (phi_f - phi_c) / (|g| + 1e-4) with g the central-difference gradient of
// phi at the centre, so the inputs are O(1) and the planar part of the
// profile only sets their scale.  Inputs enter the matrix cores as a
// hi + lo pair of BF16 numbers (about 16 significant bits): the curvature is
// a percent-level second-order feature of the stencil that one BF16 rounding
// would erase.
__device__ __forceinline__ int feat_dx(int f) { return f % 3 - 1; }
__device__ __forceinline__ int feat_dy(int f) { return (f / 3) % 3 - 1; }
__device__ __forceinline__ int feat_dz(int f) { return f / 9 - 1; }
constexpr int kCentre = 13;

// inv = 1 / (|g| + 1e-4), with g from the six face neighbours
// (features 12/14, 10/16, 4/22).
__device__ __forceinline__ float feat_scale(float xm, float xp, float ym, float yp, float zm, float zp) {
    const float gx = 0.5f * (xp - xm), gy = 0.5f * (yp - ym), gz = 0.5f * (zp - zm);
    return 1.0f / (sqrtf(gx * gx + gy * gy + gz * gz) + 1e-4f);
}

// Canonical frame of a stencil under the 48 symmetries of the cube.  The
// curvature is invariant under them, and the closure evaluates the network on
// one representative of each orbit: every axis is reflected so that the
// difference across the centre along it, d = phi(+1) - phi(-1), is >= 0, and
// the axes are ordered by |d|, largest first (the earlier axis first at a
// tie).  Canonical offset (u0, u1, u2) is then original offset
// (sx u[rx], sy u[ry], sz u[rz]).  A cube symmetry permutes and negates the
// differences exactly, so a stencil and its images pick the same canonical
// stencil (up to exact ties of |d|), and the forces on symmetric parts of an
// interface are equal by construction with one network evaluation per cell.
struct CanonFrame {
    int sx, sy, sz;  // reflection of each original axis
    int rx, ry, rz;  // canonical position of each original axis
};
__device__ __forceinline__ CanonFrame canon_frame(float dx, float dy, float dz) {
    const float a = fabsf(dx), b = fabsf(dy), c = fabsf(dz);
    CanonFrame fr;
    fr.sx = dx < 0.0f ? -1 : 1;
    fr.sy = dy < 0.0f ? -1 : 1;
    fr.sz = dz < 0.0f ? -1 : 1;
    fr.rx = (int)(b > a) + (int)(c > a);
    fr.ry = (int)(a >= b) + (int)(c > b);
    fr.rz = (int)(a >= c) + (int)(b >= c);
    return fr;
}
__device__ __forceinline__ int canon_pick(int r, int u0, int u1, int u2) { return r == 0 ? u0 : (r == 1 ? u1 : u2); }

// ---------------------------------------------------------------------------
// inference: 16 cells per wavefront job
// ---------------------------------------------------------------------------
struct MlpWeightsLds {
    __bf16 w[kMlpWords];
    float b[kMlpBias];
};

// Network output for the 16 cells of one job: lane i + 16g contributes
// features 8g..8g+7 of cell i as xh + xl.  Returns the prediction in every
// lane of column i (lanes i, i+16, i+32, i+48).
//
// The weight fragments are read from LDS on every call.  They are the same
// on every call of a kernel, so the compiler would hoist them out of the
// caller's job loop.  The offset `zero` passes through an empty asm
// statement, an optimisation barrier that emits no instruction, which
// prevents that.  Hoisted, they held ~90 registers and spilled to scratch at
// the closure kernel's 128-VGPR bound (94 spilled VGPRs against 8); kept in
// the loop, the closure took 0.21 instead of 0.33 ms per step of the
// developed 256x256x1024 jet, with identical results.
__device__ __forceinline__ float mlp_eval16(const MlpWeightsLds& s, bf16x8 xh, bf16x8 xl) {
    const int lane = threadIdx.x & 63, i = lane & 15, g = lane >> 4;
    int zero = 0;
    asm volatile("" : "+v"(zero));
    const __bf16* W1 = s.w + zero;
    const __bf16* W2 = s.w + kMlpW1 + zero;
    const __bf16* w3 = s.w + kMlpW1 + kMlpW2 + zero;
    f32x4 d1[4];
#pragma unroll
    for (int m = 0; m < 4; ++m) {
        bf16x8 a;
#pragma unroll
        for (int r = 0; r < 8; ++r) a[r] = W1[(16 * m + i) * kMlpInPad + 8 * g + r];
        // the bias seeds the accumulator (row 16m + 4g + r of this lane)
        const f32x4 b1 = {s.b[16 * m + 4 * g], s.b[16 * m + 4 * g + 1], s.b[16 * m + 4 * g + 2],
                          s.b[16 * m + 4 * g + 3]};
        d1[m] = mfma32(a, xl, mfma32(a, xh, b1));
#pragma unroll
        for (int r = 0; r < 4; ++r) d1[m][r] = gelu(d1[m][r]);
    }
    float y = 0.0f;
#pragma unroll
    for (int m = 0; m < 4; ++m) {
        const int hb = kMlpHidden + 16 * m + 4 * g;
        f32x4 acc = {s.b[hb], s.b[hb + 1], s.b[hb + 2], s.b[hb + 3]};
#pragma unroll
        for (int ks = 0; ks < 2; ++ks) {
            bf16x8 a, b;
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                a[r] = W2[(16 * m + i) * kMlpHidden + 32 * ks + 4 * g + r];
                a[4 + r] = W2[(16 * m + i) * kMlpHidden + 32 * ks + 16 + 4 * g + r];
                b[r] = (__bf16)d1[2 * ks][r];
                b[4 + r] = (__bf16)d1[2 * ks + 1][r];
            }
            acc = mfma32(a, b, acc);
        }
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int h = 16 * m + 4 * g + r;
            y += (float)w3[h] * gelu(acc[r]);
        }
    }
    y += __shfl_xor(y, 16, 64);
    y += __shfl_xor(y, 32, 64);
    return y + s.b[2 * kMlpHidden];
}

}  // namespace pf
