// Training and batched inference of the learned curvature closure on the
// matrix cores (BF16 operands, FP32 accumulation).
//
//   k_mlp_grad      forward + backward of 64 samples per workgroup; writes
//                   the workgroup's partial gradient of all 6017 parameters
//   k_mlp_adam      sums the partial gradients of a minibatch and applies one
//                   Adam step to the FP32 master parameters, refreshing the
//                   BF16 copies the kernels read
//   k_mlp_forward   network output for a batch of samples
//   k_mfma_layout   one MFMA tile from global operands, to verify the lane
//                   layout the kernels assume
//
// Parameter vector (FP32, 6017): W1[64][27] | b1[64] | W2[64][64] | b2[64] |
// w3[64] | b3.  Samples are 32 FP32 features (27 used; the normalised stencil,
// see mlp_device.hpp), entering the matrix cores as hi + lo BF16 pairs, and
// one FP32 target (4 x curvature).
#include <hip/hip_runtime.h>

#include "mlp_device.hpp"

namespace pf {

constexpr int kOffB1 = kMlpHidden * kMlpIn;
constexpr int kOffW2 = kOffB1 + kMlpHidden;
constexpr int kOffB2 = kOffW2 + kMlpW2;
constexpr int kOffW3 = kOffB2 + kMlpHidden;
constexpr int kOffB3 = kOffW3 + kMlpHidden;
static_assert(kOffB3 + 1 == kMlpParams, "parameter layout");

constexpr int kMB = 64;  // samples per workgroup

// One 16x16 tile of C = A B (A: M x K, B: K x N), operands in LDS with
// arbitrary strides; returns lane (i + 16g) = C[m0 + 4g + r][n0 + i].
__device__ __forceinline__ f32x4 tile_gemm(const __bf16* A, int sam, int sak, const __bf16* B, int sbk, int sbn,
                                           int m0, int n0, int K) {
    const int lane = threadIdx.x & 63, i = lane & 15, g = lane >> 4;
    f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
    for (int k0 = 0; k0 < K; k0 += 32) {
        bf16x8 a, b;
#pragma unroll
        for (int r = 0; r < 8; ++r) {
            a[r] = A[(m0 + i) * sam + (k0 + 8 * g + r) * sak];
            b[r] = B[(k0 + 8 * g + r) * sbk + (n0 + i) * sbn];
        }
        acc = mfma32(a, b, acc);
    }
    return acc;
}

struct MlpLds {
    __bf16 W1[kMlpHidden * kMlpInPad];
    __bf16 W2[kMlpHidden * kMlpHidden];
    float w3[kMlpHidden];
    float b1[kMlpHidden], b2[kMlpHidden];
    float b3;
    __bf16 X[kMB * kMlpInPad];           // [sample][feature], high BF16 part
    __bf16 XL[kMB * kMlpInPad];          // low BF16 part
    float Z1[kMlpHidden * kMB];          // [hidden][sample] pre-activations
    float Z2[kMlpHidden * kMB];
    __bf16 H1[kMlpHidden * kMB];         // GELU(Z1), transposed
    float H2[kMlpHidden * kMB];          // GELU(Z2), transposed
    __bf16 D2[kMlpHidden * kMB];         // dL/dZ2, transposed
    __bf16 D1[kMlpHidden * kMB];         // dL/dZ1, transposed
    float dy[kMB];
    float y[kMB];
};

// Load the BF16 weights / FP32 biases and kMB samples (with the cube-symmetry
// permutation `sym` of the stencil, if given).
__device__ void mlp_load(MlpLds& L, const uint16_t* w16, const float* bias, const float* xs, const int* idx,
                         const uint8_t* symperm, int nsym, uint32_t sym_seed, size_t n, size_t s0) {
    const int tid = threadIdx.x;
    const __bf16* w = reinterpret_cast<const __bf16*>(w16);
    for (int k = tid; k < kMlpW1; k += 256) L.W1[k] = w[k];
    for (int k = tid; k < kMlpW2; k += 256) L.W2[k] = w[kMlpW1 + k];
    for (int k = tid; k < kMlpHidden; k += 256) {
        L.w3[k] = (float)w[kMlpW1 + kMlpW2 + k];
        L.b1[k] = bias[k];
        L.b2[k] = bias[kMlpHidden + k];
    }
    if (tid == 0) L.b3 = bias[2 * kMlpHidden];
    for (int k = tid; k < kMB * kMlpInPad; k += 256) {
        const int b = k / kMlpInPad, f = k % kMlpInPad;
        const size_t s = s0 + b;
        float v = 0.0f;
        if (s < n && f < kMlpIn) {
            const size_t row = idx ? (size_t)idx[s] : s;
            int fs = f;
            if (symperm) {
                uint32_t h = (uint32_t)row * 0x9E3779B1u ^ sym_seed;
                h ^= h >> 15;
                h *= 0x2C1B3C6Du;
                h ^= h >> 12;
                fs = symperm[(h % (uint32_t)nsym) * kMlpIn + f];
            }
            v = xs[row * kMlpInPad + fs];
        }
        const __bf16 hi = (__bf16)v;
        L.X[k] = hi;
        L.XL[k] = (__bf16)(v - (float)hi);
    }
}

// Forward pass of the workgroup's kMB samples into L.y.
__device__ void mlp_forward_lds(MlpLds& L) {
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63, i = lane & 15, g = lane >> 4;
    __syncthreads();
    // Z1^T = W1 X^T : 4 x 4 tiles of 16 (hidden x sample), 4 per wave
    for (int t = wave; t < 16; t += 4) {
        const int m0 = 16 * (t >> 2), n0 = 16 * (t & 3);
        const f32x4 d = tile_gemm(L.W1, kMlpInPad, 1, L.X, 1, kMlpInPad, m0, n0, kMlpInPad) +
                        tile_gemm(L.W1, kMlpInPad, 1, L.XL, 1, kMlpInPad, m0, n0, kMlpInPad);
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int h = m0 + 4 * g + r, b = n0 + i;
            const float z = d[r] + L.b1[h];
            L.Z1[h * kMB + b] = z;
            L.H1[h * kMB + b] = (__bf16)gelu(z);
        }
    }
    __syncthreads();
    // Z2^T = W2 H1^T
    for (int t = wave; t < 16; t += 4) {
        const int m0 = 16 * (t >> 2), n0 = 16 * (t & 3);
        const f32x4 d = tile_gemm(L.W2, kMlpHidden, 1, L.H1, kMB, 1, m0, n0, kMlpHidden);
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int h = m0 + 4 * g + r, b = n0 + i;
            const float z = d[r] + L.b2[h];
            L.Z2[h * kMB + b] = z;
            L.H2[h * kMB + b] = gelu(z);
        }
    }
    __syncthreads();
    if (tid < kMB) {
        float y = 0.0f;
        for (int h = 0; h < kMlpHidden; ++h) y += L.w3[h] * L.H2[h * kMB + tid];
        L.y[tid] = y + L.b3;
    }
    __syncthreads();
}

// grad_out[blockIdx.x][0..6016]: gradient of sum over samples of
// scale * (y - t)^2; grad_out[..][6017] = sum of (y - t)^2.
__global__ __launch_bounds__(256) void k_mlp_grad(const uint16_t* w16, const float* bias, const float* xs,
                                                  const float* ts, const int* idx, const uint8_t* symperm, int nsym,
                                                  uint32_t sym_seed, size_t n, size_t s_begin, float scale,
                                                  float* grad_out) {
    __shared__ MlpLds L;
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63, i = lane & 15, g = lane >> 4;
    const size_t s0 = s_begin + (size_t)blockIdx.x * kMB;
    mlp_load(L, w16, bias, xs, idx, symperm, nsym, sym_seed, n, s0);
    mlp_forward_lds(L);
    float* G = grad_out + (size_t)blockIdx.x * (kMlpParams + 1);
    // dL/dy
    if (tid < kMB) {
        const size_t s = s0 + tid;
        float d = 0.0f, e2 = 0.0f;
        if (s < n) {
            const float t = ts[idx ? (size_t)idx[s] : s];
            const float e = L.y[tid] - t;
            d = 2.0f * scale * e;
            e2 = e * e;
        }
        L.dy[tid] = d;
        L.y[tid] = e2;
    }
    __syncthreads();
    // dZ2 = dy w3 gelu'(Z2); bias and w3 gradients
    for (int k = tid; k < kMlpHidden * kMB; k += 256) {
        const int h = k / kMB, b = k % kMB;
        L.D2[k] = (__bf16)(L.dy[b] * L.w3[h] * gelu_grad(L.Z2[k]));
    }
    if (tid < kMlpHidden) {
        float gb2 = 0.0f, gw3 = 0.0f;
        for (int b = 0; b < kMB; ++b) {
            gb2 += L.dy[b] * L.w3[tid] * gelu_grad(L.Z2[tid * kMB + b]);
            gw3 += L.dy[b] * L.H2[tid * kMB + b];
        }
        G[kOffB2 + tid] = gb2;
        G[kOffW3 + tid] = gw3;
    } else if (tid == kMlpHidden) {
        float gb3 = 0.0f, l2 = 0.0f;
        for (int b = 0; b < kMB; ++b) {
            gb3 += L.dy[b];
            l2 += L.y[b];
        }
        G[kOffB3] = gb3;
        G[kMlpParams] = l2;
    }
    __syncthreads();
    // dW2 = dZ2^T-blocks times H1 (K = samples); dH1^T = W2^T dZ2^T
    for (int t = wave; t < 16; t += 4) {
        const int m0 = 16 * (t >> 2), n0 = 16 * (t & 3);
        const f32x4 d = tile_gemm(L.D2, kMB, 1, L.H1, 1, kMB, m0, n0, kMB);
#pragma unroll
        for (int r = 0; r < 4; ++r) G[kOffW2 + (m0 + 4 * g + r) * kMlpHidden + n0 + i] = d[r];
    }
    for (int t = wave; t < 16; t += 4) {
        const int m0 = 16 * (t >> 2), n0 = 16 * (t & 3);
        const f32x4 d = tile_gemm(L.W2, 1, kMlpHidden, L.D2, kMB, 1, m0, n0, kMlpHidden);
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int h = m0 + 4 * g + r, b = n0 + i;
            L.D1[h * kMB + b] = (__bf16)(d[r] * gelu_grad(L.Z1[h * kMB + b]));
            L.Z2[h * kMB + b] = d[r] * gelu_grad(L.Z1[h * kMB + b]);  // FP32 copy for the bias sum
        }
    }
    __syncthreads();
    if (tid < kMlpHidden) {
        float gb1 = 0.0f;
        for (int b = 0; b < kMB; ++b) gb1 += L.Z2[tid * kMB + b];
        G[kOffB1 + tid] = gb1;
    }
    // dW1 = dZ1^T X (K = samples): 4 x 2 tiles (hidden x feature)
    for (int t = wave; t < 8; t += 4) {
        const int m0 = 16 * (t >> 1), n0 = 16 * (t & 1);
        const f32x4 d = tile_gemm(L.D1, kMB, 1, L.X, kMlpInPad, 1, m0, n0, kMB) +
                        tile_gemm(L.D1, kMB, 1, L.XL, kMlpInPad, 1, m0, n0, kMB);
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int f = n0 + i;
            if (f < kMlpIn) G[(m0 + 4 * g + r) * kMlpIn + f] = d[r];
        }
    }
}

// Sum nparts partial gradients, Adam step on the FP32 master parameters,
// refresh BF16 weights and FP32 biases.  loss_acc += sum of squared errors.
__global__ __launch_bounds__(256) void k_mlp_adam(const float* parts, int nparts, float* master, float* m1, float* m2,
                                                  uint16_t* w16, float* bias, float lr, float beta1, float beta2,
                                                  float eps, int step, int apply, float* grad_sum, double* loss_acc) {
    __bf16* w = reinterpret_cast<__bf16*>(w16);
    const float c1 = 1.0f / (1.0f - powf(beta1, (float)step));
    const float c2 = 1.0f / (1.0f - powf(beta2, (float)step));
    for (int p = blockIdx.x * 256 + threadIdx.x; p <= kMlpParams; p += gridDim.x * 256) {
        float gsum = 0.0f;
        for (int k = 0; k < nparts; ++k) gsum += parts[(size_t)k * (kMlpParams + 1) + p];
        if (p == kMlpParams) {
            if (loss_acc) atomicAdd(loss_acc, (double)gsum);
            continue;
        }
        if (grad_sum) grad_sum[p] = gsum;
        if (!apply) continue;
        const float a = beta1 * m1[p] + (1.0f - beta1) * gsum;
        const float b = beta2 * m2[p] + (1.0f - beta2) * gsum * gsum;
        m1[p] = a;
        m2[p] = b;
        const float v = master[p] - lr * (a * c1) / (sqrtf(b * c2) + eps);
        master[p] = v;
        // refresh the copy the kernels read
        if (p < kOffB1) {
            w[(p / kMlpIn) * kMlpInPad + (p % kMlpIn)] = (__bf16)v;
        } else if (p < kOffW2) {
            bias[p - kOffB1] = v;
        } else if (p < kOffB2) {
            w[kMlpW1 + (p - kOffW2)] = (__bf16)v;
        } else if (p < kOffW3) {
            bias[kMlpHidden + (p - kOffB2)] = v;
        } else if (p < kOffB3) {
            w[kMlpW1 + kMlpW2 + (p - kOffW3)] = (__bf16)v;
        } else {
            bias[2 * kMlpHidden] = v;
        }
    }
}

__global__ __launch_bounds__(256) void k_mlp_forward(const uint16_t* w16, const float* bias, const float* xs,
                                                     size_t n, float* y) {
    __shared__ MlpLds L;
    const size_t s0 = (size_t)blockIdx.x * kMB;
    mlp_load(L, w16, bias, xs, nullptr, nullptr, 1, 0, n, s0);
    mlp_forward_lds(L);
    if (threadIdx.x < kMB && s0 + threadIdx.x < n) y[s0 + threadIdx.x] = L.y[threadIdx.x];
}

// A: 16 x 32 row-major bf16, B: 32 x 16 row-major bf16, D: 16 x 16 fp32.
__global__ __launch_bounds__(64) void k_mfma_layout(const uint16_t* a16, const uint16_t* b16, float* d) {
    const __bf16* A = reinterpret_cast<const __bf16*>(a16);
    const __bf16* B = reinterpret_cast<const __bf16*>(b16);
    const int lane = threadIdx.x, i = lane & 15, g = lane >> 4;
    bf16x8 a, b;
#pragma unroll
    for (int r = 0; r < 8; ++r) {
        a[r] = A[i * 32 + 8 * g + r];
        b[r] = B[(8 * g + r) * 16 + i];
    }
    const f32x4 acc = mfma32(a, b, f32x4{0.0f, 0.0f, 0.0f, 0.0f});
#pragma unroll
    for (int r = 0; r < 4; ++r) d[(4 * g + r) * 16 + i] = acc[r];
}

// ---------------------------------------------------------------------------
// k_closure: network curvature at the interface cells of the current phi.
// A workgroup covers 64 x 4 cells over kClZ planes.  It reads the phi of all
// its planes at once and leaves if none of its cells has
// mlp_lo < phi < mlp_hi (most workgroups).  Otherwise it marches its planes
// with phi of planes z-1..z+1 (1-cell halo) staged in LDS, compacts the
// interface cells of each plane, loads the weights once, and runs the cells
// through the network 16 per wavefront job, features read from LDS.
// ---------------------------------------------------------------------------
// Planes per workgroup.  Measured on the D3Q27 256x256x1024 jet after 2000
// steps (phase sum + closure per step): 0.51 ms at 8, 0.65 at 16, 0.71 at
// 32 -- longer chunks serialise more network work per workgroup.
#ifndef PF_CLOSURE_ZCHUNK
#define PF_CLOSURE_ZCHUNK 8
#endif
constexpr int kClZ = PF_CLOSURE_ZCHUNK;
static_assert(kClZ >= 1 && kClZ <= 32, "one bit per plane in the 32-bit interface mask");

// Launch bound: 4 waves per SIMD, 128 VGPRs.  With the weight fragments
// read inside the job loop (mlp_eval16) only 13 VGPRs spill.  When they were
// hoisted (94 spilled), this bound was still 5 % faster than 2 waves per SIMD
// with 256 VGPRs and no spills.
__global__ __launch_bounds__(256, 4) void k_closure(KParams p) {
    __shared__ MlpWeightsLds W;
    constexpr int kPX = 66, kPY = 6;  // 64 x 4 tile with a 1-cell halo
    __shared__ float s_ph[3][kPY * kPX];
    __shared__ int s_list[256];
    __shared__ int s_count[4];
    __shared__ int s_loaded;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int x = blockIdx.x * 64 + lane, y = blockIdx.y * 4 + wave;
    const int nx = p.nx, ny = p.ny, nz = p.nz;
    const float* phi_g = p.fld[10];
    if (tid == 0) s_loaded = 0;
    // Every plane's load is issued before any is tested: the addresses are
    // clamped into the interior (grids are at least 4 cells on a side) and
    // the conditions applied afterwards.  Loads under per-plane conditions
    // compiled to one load and one wait per plane, and the scan of a
    // workgroup without interface cells, most of them, cost kClZ memory
    // latencies instead of one.
    const int xc = min(max(x, 1), nx - 2), yc = min(max(y, 1), ny - 2);
    const bool inner = x >= 1 && x <= nx - 2 && y >= 1 && y <= ny - 2;
    float v[kClZ];
#pragma unroll
    for (int k = 0; k < kClZ; ++k) {
        const int z = min(1 + (int)blockIdx.z * kClZ + k, nz - 2);
        v[k] = phi_g[(uint32_t)xc + (uint32_t)nx * ((uint32_t)yc + (uint32_t)ny * (uint32_t)z)];
    }
    uint32_t wants = 0;  // bit k: the cell of plane k is an interface cell
#pragma unroll
    for (int k = 0; k < kClZ; ++k) {
        const int z = 1 + (int)blockIdx.z * kClZ + k;
        if (inner && z <= nz - 2 && v[k] > p.mlp_lo && v[k] < p.mlp_hi) wants |= 1u << k;
    }
    if (!__syncthreads_or(wants != 0)) return;  // uniform: no interface cell in this workgroup
    const int x0 = blockIdx.x * 64, y0 = blockIdx.y * 4;
    auto stage = [&](int pz) {  // phi of plane pz, tile + halo, into ring slot pz % 3
        float* dst = s_ph[pz % 3];
        constexpr int kIt = (kPY * kPX + 255) / 256;
        float w[kIt];
#pragma unroll
        for (int j = 0; j < kIt; ++j) {  // loads first, at clamped addresses, as in the scan
            const int i = min(tid + 256 * j, kPY * kPX - 1);
            const int gx = min(max(x0 - 1 + i % kPX, 0), nx - 1), gy = min(max(y0 - 1 + i / kPX, 0), ny - 1);
            w[j] = phi_g[(uint32_t)gx + (uint32_t)nx * ((uint32_t)gy + (uint32_t)ny * (uint32_t)pz)];
        }
#pragma unroll
        for (int j = 0; j < kIt; ++j) {
            const int i = tid + 256 * j, gx = x0 - 1 + i % kPX, gy = y0 - 1 + i / kPX;
            if (i < kPY * kPX) dst[i] = (gx >= 0 && gx < nx && gy >= 0 && gy < ny) ? w[j] : 0.0f;
        }
    };
    const int zfirst = 1 + blockIdx.z * kClZ;
    stage(zfirst - 1);
    stage(zfirst);
    for (int k = 0; k < kClZ; ++k) {
        const int z = zfirst + k;
        if (z > nz - 2) break;  // uniform over the workgroup
        stage(z + 1);  // its slot held plane z - 2, last read before the previous barrier
        const bool want = (wants >> k) & 1u;
        const unsigned long long mask = __ballot(want);
        const int rank = __builtin_amdgcn_mbcnt_hi((unsigned)(mask >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)mask, 0));
        if (lane == 0) s_count[wave] = __popcll(mask);
        __syncthreads();
        int base = 0, total = 0;
        for (int w = 0; w < 4; ++w) {
            base += (w < wave) ? s_count[w] : 0;
            total += s_count[w];
        }
        if (want) s_list[base + rank] = tid;
        if (total > 0 && !s_loaded) {  // first plane with interface cells: fetch the weights
            const __bf16* w = reinterpret_cast<const __bf16*>(p.mlp_w);
            for (int i = tid; i < kMlpWords; i += 256) W.w[i] = w[i];
            for (int i = tid; i < kMlpBias; i += 256) W.b[i] = p.mlp_b[i];
        }
        __syncthreads();
        if (tid == 0 && total > 0) s_loaded = 1;
        const int i = lane & 15, g = lane >> 4;
        for (int job = wave; job < (total + 15) / 16; job += 4) {
            const int slot = 16 * job + i;
            const bool valid = slot < total;
            const int cell = valid ? s_list[slot] : 0;
            const int cx = x0 + (cell & 63), cy = y0 + (cell >> 6);
            const int o = ((cell >> 6) + 1) * kPX + (cell & 63) + 1;
            auto ph = [&](int dx, int dy, int dz) { return s_ph[(z + dz) % 3][o + dy * kPX + dx]; };
            float pc = 0.0f, inv = 0.0f;
            CanonFrame fr{1, 1, 1, 0, 1, 2};
            if (valid) {
                pc = ph(0, 0, 0);
                const float xm = ph(-1, 0, 0), xp = ph(1, 0, 0), ym = ph(0, -1, 0), yp = ph(0, 1, 0);
                const float zm = ph(0, 0, -1), zp = ph(0, 0, 1);
                inv = feat_scale(xm, xp, ym, yp, zm, zp);
                // the stencil in its canonical frame under the cube
                // symmetries: one network evaluation per cell
                fr = canon_frame(xp - xm, yp - ym, zp - zm);
            }
            bf16x8 xh, xl;
#pragma unroll
            for (int r = 0; r < 8; ++r) {
                const int f = 8 * g + r;
                float v = 0.0f;
                if (valid && f < kMlpIn) {
                    const int u0 = feat_dx(f), u1 = feat_dy(f), u2 = feat_dz(f);
                    v = (f == kCentre) ? 2.0f * pc - 1.0f
                                       : (ph(fr.sx * canon_pick(fr.rx, u0, u1, u2), fr.sy * canon_pick(fr.ry, u0, u1, u2),
                                             fr.sz * canon_pick(fr.rz, u0, u1, u2)) -
                                          pc) *
                                             inv;
                }
                xh[r] = (__bf16)v;
                xl[r] = (__bf16)(v - (float)xh[r]);
            }
            const float yv = mlp_eval16(W, xh, xl);
            if (g == 0 && valid)
                p.kappa[(uint32_t)cx + (uint32_t)nx * ((uint32_t)cy + (uint32_t)ny * (uint32_t)z)] = 0.25f * yv;
        }
        __syncthreads();
    }
}

hipError_t launch_closure(const KParams& p, hipStream_t s) {
    const dim3 grid((p.nx + 63) / 64, (p.ny + 3) / 4, (p.nz - 2 + kClZ - 1) / kClZ);
    hipLaunchKernelGGL(k_closure, grid, dim3(256), 0, s, p);
    return hipGetLastError();
}

hipError_t launch_mlp_grad(const uint16_t* w16, const float* bias, const float* xs, const float* ts, const int* idx,
                           const uint8_t* symperm, int nsym, uint32_t sym_seed, size_t n, size_t s_begin,
                           size_t count, float scale, float* grad_out, hipStream_t s) {
    const unsigned blocks = (unsigned)((count + kMB - 1) / kMB);
    hipLaunchKernelGGL(k_mlp_grad, dim3(blocks), dim3(256), 0, s, w16, bias, xs, ts, idx, symperm, nsym, sym_seed,
                       n, s_begin, scale, grad_out);
    return hipGetLastError();
}

hipError_t launch_mlp_adam(const float* parts, int nparts, float* master, float* m1, float* m2, uint16_t* w16,
                           float* bias, float lr, int step, int apply, float* grad_sum, double* loss_acc,
                           hipStream_t s) {
    hipLaunchKernelGGL(k_mlp_adam, dim3(8), dim3(256), 0, s, parts, nparts, master, m1, m2, w16, bias, lr, 0.9f,
                       0.999f, 1e-8f, step, apply, grad_sum, loss_acc);
    return hipGetLastError();
}

hipError_t launch_mlp_forward(const uint16_t* w16, const float* bias, const float* xs, size_t n, float* y,
                              hipStream_t s) {
    hipLaunchKernelGGL(k_mlp_forward, dim3((unsigned)((n + kMB - 1) / kMB)), dim3(256), 0, s, w16, bias, xs, n, y);
    return hipGetLastError();
}

hipError_t launch_mfma_layout(const uint16_t* a, const uint16_t* b, float* d, hipStream_t s) {
    hipLaunchKernelGGL(k_mfma_layout, dim3(1), dim3(64), 0, s, a, b, d);
    return hipGetLastError();
}

}  // namespace pf
