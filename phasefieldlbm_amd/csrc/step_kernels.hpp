// One timestep of the phase-field LBM on gfx950: three kernels.
//
//   k_phase_sum  phi = sum_q g_q, the g rebuilt from the neighbours' phi and
//                phase flux (phase_flux.hpp)      (src/phaseField.cuh:31-52)
//   k_step       fused normals + curvature + CSF force + moments +
//                regularized collision + f streaming + phase flux
//                (src/phaseField.cuh:54-219, src/lbm.cu:33-239)
//   k_jet_bc     inflow and outflow planes of the jet
//                (src/BoundaryConditions.cuh:36-169)
//
// f streams in place with the A-A pattern: one population buffer, two
// alternating step kinds.  With S[slot][cell]:
//   even step  in_q  = S[q][x]              out_q -> S[opp q][x]
//   odd  step  in_q  = S[opp q][x - c_q]    out_q -> S[q][x + c_q]
// Each cell reads and writes the same set of locations, so no other cell
// touches them within a step.  Neighbour coordinates wrap with the
// reference's ghost-shell periodicity (0 -> n-2, n-1 -> 1,
// src/functions/deviceFunctions.cuh:63-76): x and y for the jet, x, y and z
// for the droplet.  The jet's z = 0 and z = nz-1 ghost planes carry the
// populations the boundary kernel supplies for the next step.
//
// k_step marches a 128 x 4 column of cells (64 x 8 on grids at most 64
// wide) through z.  A ring of the phi tile (5 planes in FP32, 4 in
// binary16) with a 2-cell halo and a 3-plane ring of normals with a 1-cell
// halo live in LDS, so phi is read from HBM about once per cell and the
// normals never leave the chip (the reference stores normx/y/z, ind and the
// force in HBM between three kernels).
#pragma once

#include <hip/hip_runtime.h>

#include <type_traits>

#include "kparams.hpp"
#include "lattice.hpp"
#include "ne.hpp"
#include "phase_flux.hpp"

namespace pf {

// Workgroup barrier for LDS only: the LDS writes before it are complete and
// visible after it, while global loads and stores stay in flight.  k_step
// passes nothing between threads through global memory (phi is read-only in
// it, and under A-A addressing each location a cell reads is written only by
// that cell), so its barriers need not drain the population stores of the
// previous plane, which __syncthreads' full fence does.
__device__ __forceinline__ void lds_barrier() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// ---------------------------------------------------------------------------
// population storage
// ---------------------------------------------------------------------------
// Population stores are consumed only by the next step, so they carry the
// non-temporal (streaming) hint (measured ~7% faster on MI355X).  Loads keep
// the default policy: in the odd step neighbouring tiles re-read the lines at
// their common edge, and streaming loads measured slower.
template <bool HALF>
struct Pop;

// Location of a cell in a population array: plane index z and the in-plane
// element offset x + y*pitch.  z is wave-uniform in every kernel that
// marches planes.
struct Loc {
    uint32_t z;
    uint32_t xy;
};

// Raw buffer resource of plane z of a population array (gfx9 descriptor
// word 3: 32-bit data format).  Every access is then one 32-bit vector
// offset (x + y*pitch) plus a scalar offset (slot*nx) from a scalar base.
// num_records stops at the end of the array: an offset beyond it would read
// zero and drop the write rather than touch other memory.
__device__ __forceinline__ __amdgpu_buffer_rsrc_t plane_rsrc(const void* base, uint64_t plane_bytes,
                                                           uint64_t total_bytes, uint32_t z) {
    const uint64_t off = (uint64_t)z * plane_bytes;
    const uint64_t rem = total_bytes > off ? total_bytes - off : 0;
    const uint32_t nrec = rem > 0xffffffffull ? 0xffffffffu : (uint32_t)rem;
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<char*>(reinterpret_cast<const char*>(base)) + off, (short)0,
                                             (int)nrec, 0x00020000);
}

#ifdef PF_CACHED_STORE
constexpr int kStoreAux = 0;
#else
constexpr int kStoreAux = 2;  // nt: streaming store
#endif

template <>
struct Pop<false> {
    using T = float;
    static __device__ __forceinline__ float ld(const void* base, uint64_t plane, uint64_t total, uint32_t nx, int slot,
                                               Loc l) {
        const auto r = plane_rsrc(base, plane * 4u, total * 4u, l.z);
        return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, l.xy << 2, (uint32_t)slot * nx * 4u, 0));
    }
    static __device__ __forceinline__ void st(void* base, uint64_t plane, uint64_t total, uint32_t nx, int slot, Loc l,
                                              float v) {
        const auto r = plane_rsrc(base, plane * 4u, total * 4u, l.z);
        __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(uint32_t, v), r, l.xy << 2, (uint32_t)slot * nx * 4u,
                                              kStoreAux);
    }
};
template <>
struct Pop<true> {
    using T = _Float16;
    static __device__ __forceinline__ float ld(const void* base, uint64_t plane, uint64_t total, uint32_t nx, int slot,
                                               Loc l) {
        const auto r = plane_rsrc(base, plane * 2u, total * 2u, l.z);
        const unsigned short h = __builtin_amdgcn_raw_buffer_load_b16(r, l.xy << 1, (uint32_t)slot * nx * 2u, 0);
        return (float)__builtin_bit_cast(_Float16, h);
    }
    static __device__ __forceinline__ void st(void* base, uint64_t plane, uint64_t total, uint32_t nx, int slot, Loc l,
                                              float v) {
        const auto r = plane_rsrc(base, plane * 2u, total * 2u, l.z);
        __builtin_amdgcn_raw_buffer_store_b16(__builtin_bit_cast(unsigned short, (_Float16)v), r, l.xy << 1,
                                              (uint32_t)slot * nx * 2u, kStoreAux);
    }
};

// Accessors of the two population arrays of a KParams.
template <bool HALF>
__device__ __forceinline__ float ldf(const KParams& p, int slot, Loc l) {
    return Pop<HALF>::ld(p.f, p.plane_f, p.elems_f, (uint32_t)p.nx, slot, l);
}
template <bool HALF>
__device__ __forceinline__ void stf(const KParams& p, int slot, Loc l, float v) {
    Pop<HALF>::st(p.f, p.plane_f, p.elems_f, (uint32_t)p.nx, slot, l, v);
}

// Reciprocal and square root: IEEE division and the correctly rounded square
// root by default; the hardware reciprocal and square root (1 ulp) with FAST.
// The non-equilibrium binary16 step takes FAST: it is VALU-bound, and the
// hardware forms took it from 2.88 to 2.74 ms on the D3Q27 256x256x1024 jet
// (MI355X, alternated builds); the plain binary16 step measured no faster
// with them (1.70 ms either way) and keeps the IEEE forms, as does FP32.
template <bool FAST = false>
__device__ __forceinline__ float recip(float x) {
    if constexpr (FAST) return __builtin_amdgcn_rcpf(x);
    return 1.0f / x;
}
template <bool FAST>
__device__ __forceinline__ float root(float x) {
    if constexpr (FAST) return __builtin_amdgcn_sqrtf(x);
    return sqrtf(x);
}

// Neighbours of one cell (ghost-shell periodicity applied): x, y*pitch of
// the population array and z for offsets -1, 0, +1 along each axis.
struct Nbr {
    uint32_t xw[3];
    uint32_t yf[3];
    uint32_t zz[3];
};

template <bool WRAPZ>
__device__ __forceinline__ Nbr make_nbr(const KParams& p, int x, int y, int z) {
    const int nx = p.nx, ny = p.ny, nz = p.nz;
    Nbr n;
    const int xm = (x == 1) ? nx - 2 : x - 1, xp = (x == nx - 2) ? 1 : x + 1;
    const int ym = (y == 1) ? ny - 2 : y - 1, yp = (y == ny - 2) ? 1 : y + 1;
    int zm = z - 1, zp = z + 1;
    if constexpr (WRAPZ) {
        zm = (z == 1) ? nz - 2 : z - 1;
        zp = (z == nz - 2) ? 1 : z + 1;
    }
    n.xw[0] = (uint32_t)xm;
    n.xw[1] = (uint32_t)x;
    n.xw[2] = (uint32_t)xp;
    const int yy[3] = {ym, y, yp};
#pragma unroll
    for (int a = 0; a < 3; ++a) n.yf[a] = p.pitch_f * (uint32_t)yy[a];
    n.zz[0] = (uint32_t)zm;
    n.zz[1] = (uint32_t)z;
    n.zz[2] = (uint32_t)zp;
    return n;
}

// The cell itself and its neighbour x + SIGN*c_q in the population array.
__device__ __forceinline__ Loc self_f(const Nbr& n) { return {n.zz[1], n.xw[1] + n.yf[1]}; }
template <int q, int SIGN>
__device__ __forceinline__ Loc nb_f(const Nbr& n) {
    return {n.zz[SIGN * kCZ[q] + 1], n.xw[SIGN * kCX[q] + 1] + n.yf[SIGN * kCY[q] + 1]};
}

// Where the next step of parity PAR reads f_in_q of the cell with
// neighbour table n.
template <int q, int PAR>
__device__ __forceinline__ void in_slot(const Nbr& n, int& slot, Loc& l) {
    if constexpr (PAR == 0) {
        slot = q;
        l = self_f(n);
    } else {
        slot = opp(q);
        l = nb_f<q, -1>(n);
    }
}

// ---------------------------------------------------------------------------
// k_phase_sum: phi = sum_q g_q over interior cells (src/phaseField.cuh:31-52),
// with the seven g rebuilt from the neighbours' phi and phase flux
// (phase_flux.hpp), written to the second phi buffer.  Each thread handles a
// column of ps_z(HALF) cells in z.  FP32: the cells' loads are issued
// together (arriving_g per cell).  Binary16: the thread marches the column
// and carries phi and h_z of the planes below and above in registers, so each
// plane is read once per column.  Measured on MI355X (D3Q27 256x256x1024):
// FP32 0.325 ms with 2 stacked cells (the march: 0.327-0.354 ms and a 1 %
// slower k_step after it), binary16 0.21 ms with a 4-cell march (stacked:
// 0.26 ms).  Both form the same sums.
// ---------------------------------------------------------------------------
constexpr int ps_z(bool half) { return half ? 4 : 2; }
#ifndef PF_PS_TY
#define PF_PS_TY 4
#endif
// Rows of a phase-sum workgroup (kTX wide).  Measured on MI355X (FP32,
// 256x256x1024): 4 rows 0.325 ms; 8 rows fetch 22.2 instead of 24.2 B/cell
// (fewer halo rows) but take 0.336 ms; 2 rows 0.338 ms.
constexpr int kPsTY = PF_PS_TY;
constexpr int kPsThreads = kTX * kPsTY;

#ifndef PF_PS_BANDS
#define PF_PS_BANDS 0
#endif
// Tile of a phase-sum workgroup.  PF_PS_BANDS = 0: the hardware order of the
// 3-D grid (x, then y, then z blocks).  PF_PS_BANDS = 1 (1-D grid): workgroup
// b runs on XCD k = b % 8, which takes the k-th band of ceil(ty / 8) tile
// rows and sweeps it x fastest, then through the band, then in z, so the
// y-neighbour rows of a tile are fetched into the same L2.  Measured on
// MI355X for the D3Q27 256x256x1024 jet: the bands cut the FP32 phase sum's
// HBM fetch from 24.2 to 16.6 B/cell (binary16: 17.3 to 10.5) but not its
// time (FP32 0.334-0.341 ms against 0.325-0.328 ms, binary16 0.235 against
// 0.234 ms), so the default keeps the hardware order.
__device__ __forceinline__ bool ps_tile(int ntx, int nty, int& bx, int& by, int& bz) {
#if PF_PS_BANDS
    const int band = (nty + 7) / 8;
    const int k = (int)(blockIdx.x & 7u), j = (int)(blockIdx.x >> 3);
    bx = j % ntx;
    by = k * band + (j / ntx) % band;
    bz = j / (ntx * band);
    return by < nty;
#else
    (void)ntx;
    (void)nty;
    bx = (int)blockIdx.x;
    by = (int)blockIdx.y;
    bz = (int)blockIdx.z;
    return true;
#endif
}

template <bool JET, bool LOCAL, bool HALF>
__global__ __launch_bounds__(kPsThreads) void k_phase_sum(KParams p) {
    constexpr int kPsZ = ps_z(HALF);
    const int nx = p.nx, ny = p.ny, nz = p.nz;
    int bx, by, bz;
    if (!ps_tile((nx + kTX - 1) / kTX, (ny + kPsTY - 1) / kPsTY, bx, by, bz)) return;
    const int x = bx * kTX + (threadIdx.x & (kTX - 1));
    const int y = by * kPsTY + (int)(threadIdx.x / kTX);
    if (x < 1 || x > nx - 2 || y < 1 || y > ny - 2) return;
    const int z0 = 1 + bz * kPsZ;
    const int z1 = min(z0 + kPsZ, nz - 1);  // one past the last plane of the column
    const uint32_t row = (uint32_t)nx, plane = (uint32_t)nx * (uint32_t)ny;
    const uint32_t col = (uint32_t)x + row * (uint32_t)y;
    if constexpr (!HALF && !LOCAL) {
        float g[kPsZ][7];
#pragma unroll
        for (int k = 0; k < kPsZ; ++k) arriving_g<JET, false, HALF>(p, x, y, min(z0 + k, nz - 2), g[k]);
#pragma unroll
        for (int k = 0; k < kPsZ; ++k) {
            const int z = z0 + k;
            if (z >= z1) break;  // clamped duplicates are not stored
            float phi = 0.0f;
#pragma unroll
            for (int q = 0; q < 7; ++q) phi += g[k][q];
            p.phi_next[col + plane * (uint32_t)z] = phi;
        }
        return;
    }
    if constexpr (LOCAL) {
        for (int z = z0; z < z1; ++z) {
            float g[7];
            arriving_g<JET, true, HALF>(p, x, y, z, g);
            float phi = 0.0f;
#pragma unroll
            for (int q = 0; q < 7; ++q) phi += g[q];
            p.phi_next[col + plane * (uint32_t)z] = phi;
        }
        return;
    }
    using HT = typename std::conditional<HALF, _Float16, float>::type;
    const HT* __restrict__ h0 = static_cast<const HT*>(p.h[0]);
    const HT* __restrict__ h1 = static_cast<const HT*>(p.h[1]);
    const HT* __restrict__ h2 = static_cast<const HT*>(p.h[2]);
    const float* __restrict__ phi = p.fld[10];
    // lateral neighbours with the ghost-shell periodicity, as offsets
    const int xm = (x == 1) ? nx - 2 : x - 1, xp = (x == nx - 2) ? 1 : x + 1;
    const int ym = (y == 1) ? ny - 2 : y - 1, yp = (y == ny - 2) ? 1 : y + 1;
    const uint32_t dxm = (uint32_t)(xm - x), dxp = (uint32_t)(xp - x);
    const uint32_t dym = row * (uint32_t)(ym - y), dyp = row * (uint32_t)(yp - y);
    auto zwrap = [&](int z) {
        if constexpr (!JET) {
            if (z < 1) return nz - 2;
            if (z > nz - 2) return 1;
        }
        return z;
    };
    uint32_t cc = col + plane * (uint32_t)z0;
    const uint32_t cm = col + plane * (uint32_t)zwrap(z0 - 1);
    float phi_m = phi[cm], phi_c = phi[cc];
    float h2_m = (float)h2[cm], h2_c = (float)h2[cc];
    for (int z = z0; z < z1; ++z) {
        const uint32_t cp = col + plane * (uint32_t)zwrap(z + 1);
        const float phi_p = phi[cp], h2_p = (float)h2[cp];
        // the arriving g in the reference's D3Q7 order, as arriving_g forms them
        float g[7];
        g[0] = 0.25f * phi_c;
        g[1] = 0.125f * phi[cc + dxm] + (float)h0[cc + dxm];
        g[2] = 0.125f * phi[cc + dxp] - (float)h0[cc + dxp];
        g[3] = 0.125f * phi[cc + dym] + (float)h1[cc + dym];
        g[4] = 0.125f * phi[cc + dyp] - (float)h1[cc + dyp];
        g[5] = 0.125f * phi_m + h2_m;
        g[6] = 0.125f * phi_p - h2_p;
        if constexpr (JET) {
            if (z == 1) g[5] = p.gbc[col];
            if (z == nz - 2) g[6] = p.gbc[plane + col];
        }
        float sum = 0.0f;
#pragma unroll
        for (int q = 0; q < 7; ++q) sum += g[q];
        p.phi_next[cc] = sum;
        phi_m = phi_c;
        phi_c = phi_p;
        h2_m = h2_c;
        h2_c = h2_p;
        cc += plane;  // the next plane of the column (no wrap inside it)
    }
}

// ---------------------------------------------------------------------------
// k_step: fused phase-field forcing + regularized collide-stream
// ---------------------------------------------------------------------------
// Workgroup b runs on XCD b % 8 (the hardware deals workgroups round-robin
// over the 8 XCDs); map it to a tile (x fastest, then y, then z).
//  - Rows of `row` >= 2 tiles: each XCD takes whole rows, so x-adjacent
//    tiles, which share cache lines at their common edge in the odd step,
//    meet in one L2, and consecutive rows go to consecutive XCDs, so all
//    XCDs work on neighbouring rows at a time.  Any last incomplete round
//    of 8 rows keeps the identity.  Measured on MI355X (D3Q27 256x256x1024,
//    2 tiles per row, one process, alternated): k_step FP32 3.12-3.15 ms
//    against 3.25 ms for the contiguous ranges below, FP16 1.74 against
//    1.80 ms; chunks of half a row (single tiles) 3.44 ms, of 2 rows
//    3.18-3.19 ms, of 4 rows 3.02-3.18 ms.
//  - Rows of one tile: each XCD takes a contiguous range of tiles.  Dealing
//    single rows was 1-4 % slower there (64x64x128 and 128x128x512 jets).
__device__ __forceinline__ uint32_t xcd_remap(uint32_t b, uint32_t nblk, uint32_t row) {
#ifdef PF_NO_XCD_REMAP
    (void)nblk;
    (void)row;
    return b;
#endif
    const uint32_t k = b & 7u, j = b >> 3;
    if (row < 2) {
        const uint32_t q = nblk >> 3, r = nblk & 7u;
        return k * q + min(k, r) + j;
    }
    const uint32_t m = nblk / (8u * row) * (8u * row);
    if (b >= m) return b;
    return ((j / row) * 8u + k) * row + (j % row);
}

// c.v for a compile-time direction (zero components skipped, signs folded).
template <int cx, int cy, int cz>
__device__ __forceinline__ float cdot(float x, float y, float z) {
    float s = 0.0f;
    bool any = false;
    if constexpr (cx != 0) {
        s = cx > 0 ? x : -x;
        any = true;
    }
    if constexpr (cy != 0) {
        s = any ? (cy > 0 ? s + y : s - y) : (cy > 0 ? y : -y);
        any = true;
    }
    if constexpr (cz != 0) s = any ? (cz > 0 ? s + z : s - z) : (cz > 0 ? z : -z);
    return s;
}

// Per-cell quantities shared by all pairs of the regularized collision.
struct Coll {
    float B0;               // rho (1 - 1.5|u|^2) - 1 - 1.5 u.F : even part common to all directions
    float R3, R45, V, O45;  // 3 rho, 4.5 rho, 1 - 1.5|u|^2, 4.5 (1 - omega)
    float Pxx, Pyy, Pzz, Sxy, Sxz, Syz, S3, Txy, Txz, Tyz;  // stress sums, T = 2 P_ab
    float X1, X2, X3, X4, X5, X6, X7s;  // third-order Hermite blocks (D3Q27), X7s = 6 X7
    float Bx, By, Bz;                    // X3 + X5, X1 + X6, X2 + X4
};

// out(+c) and out(-c) of one opposite pair (c = direction of the odd member):
//   even = w (B0 + 4.5 rho t^2 + 4.5 t s + 4.5 (1-omega) c.P.c)
//   odd  = w (3 rho t [1 + 1.5 t^2 - 1.5|u|^2]_{D3Q27} + 1.5 s + 4.5 (1-omega) N3_{D3Q27})
// with t = c.u and s = c.F.  This is f_eq + (1-omega) f_neq(P) + F_q/2 of the
// reference (D3Q19.cuh:143-187, D3Q27.cuh:152-206, lbm.cu:206) split into
// its even and odd parts in c, each evaluated once per pair.  KS: both
// outputs scaled by kNeK (folded into w; a power of two, so exactly).
template <int Q, int cx, int cy, int cz, bool KS = false>
__device__ __forceinline__ void pair_out(const Coll& k, float ux, float uy, float uz, float fx, float fy, float fz,
                                         float& out_p, float& out_m) {
    constexpr float w = (KS ? kNeK : 1.0f) * wk<Q>((cx != 0) + (cy != 0) + (cz != 0));
    constexpr int nz_ = (cx != 0) + (cy != 0) + (cz != 0);
    const float t = cdot<cx, cy, cz>(ux, uy, uz);
    const float s = cdot<cx, cy, cz>(fx, fy, fz);
    const float t2 = t * t;
    float cpc;
    if constexpr (nz_ == 1) {
        cpc = cx != 0 ? k.Pxx : (cy != 0 ? k.Pyy : k.Pzz);
    } else if constexpr (nz_ == 2) {
        const float d = (cz == 0) ? k.Sxy : ((cy == 0) ? k.Sxz : k.Syz);
        const float o = (cz == 0) ? k.Txy : ((cy == 0) ? k.Txz : k.Tyz);
        cpc = ((cz == 0 ? cx * cy : (cy == 0 ? cx * cz : cy * cz)) > 0) ? d + o : d - o;
    } else {
        float o = cx * cy > 0 ? k.Txy : -k.Txy;
        o = cx * cz > 0 ? o + k.Txz : o - k.Txz;
        o = cy * cz > 0 ? o + k.Tyz : o - k.Tyz;
        cpc = k.S3 + o;
    }
    const float even = w * (k.B0 + k.R45 * t2 + 4.5f * t * s + k.O45 * cpc);
    float odd;
    if constexpr (Q == 19) {
        odd = w * (k.R3 * t + 1.5f * s);
    } else {
        // third-order block sum_i K_i X_i + 6 cx cy cz X7, K_i = 3 k_i of D3Q27.cuh:180-189
        float nb;
        if constexpr (nz_ == 3) {
            nb = 2.0f * cdot<cx, cy, cz>(k.Bx, k.By, k.Bz);
            nb = (cx * cy * cz > 0) ? nb + k.X7s : nb - k.X7s;
        } else {
            constexpr int K1 = cy * (3 * cx * cx - 1), K2 = cz * (3 * cx * cx - 1);
            constexpr int K3 = cx * (3 * cy * cy - 1), K4 = cz * (3 * cy * cy - 1);
            constexpr int K5 = cx * (3 * cz * cz - 1), K6 = cy * (3 * cz * cz - 1);
            nb = 0.0f;
            if constexpr (K1 != 0) nb = nb + (float)K1 * k.X1;
            if constexpr (K2 != 0) nb = nb + (float)K2 * k.X2;
            if constexpr (K3 != 0) nb = nb + (float)K3 * k.X3;
            if constexpr (K4 != 0) nb = nb + (float)K4 * k.X4;
            if constexpr (K5 != 0) nb = nb + (float)K5 * k.X5;
            if constexpr (K6 != 0) nb = nb + (float)K6 * k.X6;
        }
        odd = w * (k.R3 * t * (k.V + 1.5f * t2) + 1.5f * s + k.O45 * nb);
    }
    out_p = even + odd;
    out_m = even - odd;
}

// LDS of k_step.  Full: a 5-plane phi ring and 3 planes of (n, |grad phi|)
// with a 1-cell halo, 58.9 KB, 2 workgroups per CU.  Slim: a 4-plane phi ring,
// 3 planes of n and 2 planes of |grad phi| at the tile cells only (it is read
// at the centre only), 49.1 KB, 3 workgroups per CU.
#ifndef PF_SLIM_LDS_FP32
#define PF_SLIM_LDS_FP32 0
#endif
#ifndef PF_SLIM_LDS_FP16
#define PF_SLIM_LDS_FP16 1
#endif

template <int Q, bool JET, bool HALF, int PAR, bool MLP, bool LIT, bool SAVE, int TX, bool NE = false>
__global__ __launch_bounds__(kThreads) void k_step(KParams p) {
    static_assert(!NE || (HALF && !LIT), "the non-equilibrium form is binary16 with the pair collision");
    constexpr int TY = kThreads / TX;                  // rows of the TX x TY tile
    constexpr int HX = TX + 4, HY = TY + 4;            // phi tile with 2-cell halo
    constexpr int NX1 = TX + 2, NY1 = TY + 2;          // normal tile with 1-cell halo
    constexpr bool SLIM = HALF ? (PF_SLIM_LDS_FP16 != 0) : (PF_SLIM_LDS_FP32 != 0);
    constexpr int kRing = SLIM ? 4 : 5;  // phi planes
    constexpr int kNC = SLIM ? 3 : 4;    // components per normal plane
    __shared__ float s_phi[kRing][HY * HX];
    __shared__ float s_n[3][kNC][NY1 * NX1];
    __shared__ float s_ind[SLIM ? 2 : 1][SLIM ? kThreads : 1];
    // non-equilibrium form: M of planes z-1, z, z+1 over the tile and a
    // 1-cell halo, taken at the wrapped coordinates of the cells that wrote
    // the populations each cell reads (ne.hpp)
    __shared__ uint2 s_m[NE ? 3 : 1][NE ? NX1 * NY1 : 1];
    constexpr int NP = (Q - 1) / 2;

    const int tid = threadIdx.x;
    const int tx = tid & (TX - 1), ty = tid / TX;
    const int nx = p.nx, ny = p.ny, nz = p.nz;
    const uint32_t lb = xcd_remap(blockIdx.x, gridDim.x, (uint32_t)p.ntx);
    const int bx = (int)(lb % (uint32_t)p.ntx);
    const uint32_t rem = lb / (uint32_t)p.ntx;
    const int by = (int)(rem % (uint32_t)p.nty);
    const int bz = (int)(rem / (uint32_t)p.nty);
    const int x0 = bx * TX, y0 = by * TY;
    const int zs = 1 + bz * p.zchunk;
    const int ze = min(zs + p.zchunk, nz - 1);
    const float* __restrict__ phi_g = p.fld[10];

    // phi of plane pz, tile + 2-cell halo, zero outside the grid, into ring
    // slot pz % kRing.  fetch_phi issues all of a thread's loads at once, at
    // addresses clamped into the grid; put_phi applies the bounds as it
    // stores to LDS.  (Loads under per-element bounds tests compiled to a
    // load and a wait each.)
    constexpr int kPhiIt = (HX * HY + kThreads - 1) / kThreads;
    auto fetch_phi = [&](int pz, float (&v)[kPhiIt]) {
        const int cz = min(max(pz, 0), nz - 1);
#pragma unroll
        for (int j = 0; j < kPhiIt; ++j) {
            const int i = min(tid + j * kThreads, HX * HY - 1);
            const int gx = min(max(x0 - 2 + i % HX, 0), nx - 1), gy = min(max(y0 - 2 + i / HX, 0), ny - 1);
            v[j] = phi_g[(uint32_t)gx + (uint32_t)nx * ((uint32_t)gy + (uint32_t)ny * (uint32_t)cz)];
        }
    };
    auto put_phi = [&](int pz, const float (&v)[kPhiIt]) {
        float* dst = s_phi[(pz + kRing) % kRing];
        const bool zin = pz >= 0 && pz < nz;
#pragma unroll
        for (int j = 0; j < kPhiIt; ++j) {
            const int i = tid + j * kThreads;
            const int gx = x0 - 2 + i % HX, gy = y0 - 2 + i / HX;
            if (i < HX * HY) dst[i] = (zin && gx >= 0 && gx < nx && gy >= 0 && gy < ny) ? v[j] : 0.0f;
        }
    };

    // M of plane pz over the 1-halo tile (non-equilibrium form), as fetch_phi
    // and put_phi: x and y with the ghost-shell wrap (a halo cell outside
    // the grid belongs to a ghost column that reads nothing), z wrapped for
    // the droplet
    constexpr int kMIt = NE ? (NX1 * NY1 + kThreads - 1) / kThreads : 1;
    auto fetch_m = [&](int pz, uint2 (&v)[kMIt]) {
        if constexpr (NE) {
            int cz = pz;
            if constexpr (!JET) cz = (pz < 1) ? nz - 2 : (pz > nz - 2 ? 1 : pz);
            cz = min(max(cz, 0), nz - 1);
#pragma unroll
            for (int j = 0; j < kMIt; ++j) {
                const int i = min(tid + j * kThreads, NX1 * NY1 - 1);
                const int gx = x0 - 1 + i % NX1, gy = y0 - 1 + i / NX1;
                const int wx = gx <= 0 ? nx - 2 : (gx >= nx - 1 ? 1 : gx);
                const int wy = gy <= 0 ? ny - 2 : (gy >= ny - 1 ? 1 : gy);
                v[j] = p.mom_prev[(uint32_t)wx + (uint32_t)nx * ((uint32_t)wy + (uint32_t)ny * (uint32_t)cz)];
            }
        }
    };
    auto put_m = [&](int pz, const uint2 (&v)[kMIt]) {
        if constexpr (NE) {
            uint2* dst = s_m[pz % 3];
#pragma unroll
            for (int j = 0; j < kMIt; ++j) {
                const int i = tid + j * kThreads;
                if (i < NX1 * NY1) dst[i] = v[j];
            }
        }
    };

    // normals and |grad phi| on the 1-halo tile of plane pz
    // (src/phaseField.cuh:54-140); zero at ghost cells, which the reference
    // never writes.
    auto comp_normals = [&](int pz) {
        const int r = pz % 3;
        const float* pm = s_phi[(pz + kRing - 1) % kRing];
        const float* p0 = s_phi[pz % kRing];
        const float* pp = s_phi[(pz + 1) % kRing];
        for (int i = tid; i < NX1 * NY1; i += kThreads) {
            const int lx = i % NX1, ly = i / NX1;
            const int gx = x0 - 1 + lx, gy = y0 - 1 + ly;
            float nrx = 0.0f, nry = 0.0f, nrz = 0.0f, ind = 0.0f;
            if (gx >= 1 && gx <= nx - 2 && gy >= 1 && gy <= ny - 2 && pz >= 1 && pz <= nz - 2) {
                const int o = (ly + 1) * HX + (lx + 1);
                auto val = [&](auto qc) {
                    constexpr int q = decltype(qc)::value;
                    constexpr int off = kCY[q] * HX + kCX[q];
                    const float* pl = kCZ[q] < 0 ? pm : (kCZ[q] > 0 ? pp : p0);
                    return pl[o + off];
                };
                const float gx_ = kAs2 * iso_sum<Q, 0>(val);
                const float gy_ = kAs2 * iso_sum<Q, 1>(val);
                const float gz_ = kAs2 * iso_sum<Q, 2>(val);
                ind = root<NE>(gx_ * gx_ + gy_ * gy_ + gz_ * gz_);
                const float inv = recip<NE>(ind + 1e-9f);
                nrx = gx_ * inv;
                nry = gy_ * inv;
                nrz = gz_ * inv;
            }
            s_n[r][0][i] = nrx;
            s_n[r][1][i] = nry;
            s_n[r][2][i] = nrz;
            if constexpr (SLIM) {
                if (lx >= 1 && lx <= TX && ly >= 1 && ly <= TY) s_ind[pz & 1][(ly - 1) * TX + (lx - 1)] = ind;
            } else {
                s_n[r][3][i] = ind;
            }
        }
    };

    // prologue: the four planes' loads all issued before any is stored
    {
        float v[4][kPhiIt];
#pragma unroll
        for (int k = 0; k < 4; ++k) fetch_phi(zs - 2 + k, v[k]);
        uint2 mv[2][kMIt];
        fetch_m(zs - 1, mv[0]);
        fetch_m(zs, mv[1]);
#pragma unroll
        for (int k = 0; k < 4; ++k) put_phi(zs - 2 + k, v[k]);
        put_m(zs - 1, mv[0]);
        put_m(zs, mv[1]);
    }
    lds_barrier();
    comp_normals(zs - 1);
    comp_normals(zs);
    // the 4-plane ring reuses the slot of phi(zs - 2) for the first load below
    if constexpr (SLIM) lds_barrier();

    const int x = x0 + tx, y = y0 + ty;
    const bool active = (x >= 1 && x <= nx - 2 && y >= 1 && y <= ny - 2);

    // Binary16: phi of plane z + 2 is fetched one iteration ahead, so its
    // loads are in flight during the collision of plane z - 1.  Measured on
    // MI355X (D3Q27 256x256x1024, alternated builds): binary16 k_step 1.695
    // against 1.710 ms fetched in place; FP32 3.065 against 3.030 ms, so FP32
    // fetches each plane at the top of its iteration.
    constexpr bool kPrefetch = HALF;
    float phi_next[kPhiIt];
    if constexpr (kPrefetch) fetch_phi(zs + 2, phi_next);
    for (int z = zs; z < ze; ++z) {
        // M of plane z + 1 replaces that of z - 2, which the previous
        // iteration read before the first barrier below
        uint2 m_next[kMIt];
        fetch_m(z + 1, m_next);
        // The closure's neighbour mean read phi of plane z - 2 after the
        // previous iteration's last barrier; the 4-plane ring puts plane
        // z + 2 in that slot below.
        if constexpr (SLIM && MLP) lds_barrier();
        if constexpr (kPrefetch) {
            put_phi(z + 2, phi_next);
            if (z + 1 < ze) fetch_phi(z + 3, phi_next);
        } else {
            fetch_phi(z + 2, phi_next);
            put_phi(z + 2, phi_next);
        }
        lds_barrier();
        put_m(z + 1, m_next);
        comp_normals(z + 1);
        lds_barrier();

        // ---- curvature and continuum surface force (src/phaseField.cuh:142-219)
        const int o = (ty + 1) * NX1 + (tx + 1);
        const float* nm[3] = {s_n[(z + 2) % 3][0], s_n[z % 3][0], s_n[(z + 1) % 3][0]};
        auto nval = [&](int comp_, auto qc) {
            constexpr int q = decltype(qc)::value;
            constexpr int off = kCY[q] * NX1 + kCX[q];
            return nm[kCZ[q] + 1][comp_ * (NY1 * NX1) + o + off];
        };
        const float phi = s_phi[z % kRing][(ty + 2) * HX + (tx + 2)];
        float curvature = 0.0f;
        if (active) {
            const float scx = iso_sum<Q, 0>([&](auto qc) { return nval(0, qc); });
            const float scy = iso_sum<Q, 1>([&](auto qc) { return nval(1, qc); });
            const float scz = iso_sum<Q, 2>([&](auto qc) { return nval(2, qc); });
            curvature = kAs2 * (scx + scy + scz);
        }
        if (!active) continue;
        const uint32_t c = (uint32_t)x + (uint32_t)nx * ((uint32_t)y + (uint32_t)ny * (uint32_t)z);
        if constexpr (MLP) {
            // Learned closure at interface cells whose FD curvature lies in
            // the trained range: the mean of the network curvature (written
            // by k_closure for this step at every cell with
            // mlp_lo < phi < mlp_hi) over the cell and those of its 26
            // neighbours that are interior interface cells.  The network is
            // evaluated on each stencil in its canonical frame under the cube
            // symmetries; the mean spreads what remains of its pointwise
            // error along the interface, where an error that is not a function
            // of phi drives spurious currents.
            if (phi > p.mlp_lo && phi < p.mlp_hi && curvature >= p.mlp_kmin && curvature <= p.mlp_kmax) {
                const int oc = (ty + 2) * HX + (tx + 2);
                float ks = p.kappa[c], ws = 1.0f;
                static_for<27>([&](auto qc) {
                    constexpr int q = decltype(qc)::value;
                    constexpr int dx = q % 3 - 1, dy = (q / 3) % 3 - 1, dz = q / 9 - 1;
                    if constexpr (q != 13) {
                        const float pn = s_phi[(z + kRing + dz) % kRing][oc + dy * HX + dx];
                        const int xn = x + dx, yn = y + dy, zn = z + dz;
                        if (pn > p.mlp_lo && pn < p.mlp_hi && xn >= 1 && xn <= nx - 2 && yn >= 1 && yn <= ny - 2 &&
                            zn >= 1 && zn <= p.nz - 2) {
                            ks += p.kappa[(uint32_t)xn + (uint32_t)nx * ((uint32_t)yn + (uint32_t)ny * (uint32_t)zn)];
                            ws += 1.0f;
                        }
                    }
                });
                curvature = ks / ws;
            }
        }

        float ind;
        if constexpr (SLIM)
            ind = s_ind[z & 1][tid];
        else
            ind = s_n[z % 3][3][o];
        const float nrx = s_n[z % 3][0][o], nry = s_n[z % 3][1][o], nrz = s_n[z % 3][2][o];
        const float stCurv = -p.sigma * curvature * ind;
        const float fsx = stCurv * nrx, fsy = stCurv * nry, fsz = stCurv * nrz;

        const Nbr nb = make_nbr<!JET>(p, x, y, z);
        const Loc sf = self_f(nb);

        // ---- relaxation (src/lbm.cu:158-172)
        float omega;
        if constexpr (JET) {
            const float tau_phi = (1.0f - phi) * p.tau_water + phi * p.tau_oil;
            const float zn = (float)z * p.sp_inv_nz_m1;
            float s = (zn - p.sp_z_start) * p.sp_inv_sponge;
            s = fminf(fmaxf(s, 0.0f), 1.0f);
            const float r = s * s * (3.0f - 2.0f * s);
            const float visc_local = (1.0f - phi) * p.visc_water + phi * p.visc_oil;
            const float tau_zmax = 0.5f + kAs2 * (visc_local * (1.0f + 100.0f));
            const float tau_eff = tau_phi + r * (tau_zmax - tau_phi);
            omega = recip<NE>(tau_eff);
        } else {
            omega = p.omega_ref;
        }
        const float omco = 1.0f - omega;

        float rho, ux, uy, uz, pxx, pyy, pzz, pxy, pxz, pyz;
        if constexpr (LIT) {
            // ---- literal evaluation, direction by direction (src/lbm.cu:33-207)
            float pop[Q];
            static_for<Q>([&](auto qc) {
                constexpr int q = decltype(qc)::value;
                if constexpr (PAR == 0) {
                    pop[q] = ldf<HALF>(p, q, sf);
                } else {
                    pop[q] = ldf<HALF>(p, opp(q), nb_f<q, -1>(nb));
                }
            });
            rho = 0.0f;
#pragma unroll
            for (int q = 0; q < Q; ++q) rho += pop[q];
            rho += 1.0f;
            const float invRho = 1.0f / rho;
            auto pv = [&](auto qc) { return pop[decltype(qc)::value]; };
            ux = invRho * pair_chain<Q, 0, -1>(pv);
            uy = invRho * pair_chain<Q, 1, -1>(pv);
            uz = invRho * pair_chain<Q, 2, -1>(pv);
            ux += fsx * 0.5f * invRho;
            uy += fsy * 0.5f * invRho;
            uz += fsz * 0.5f * invRho;
            const float uu = 1.5f * (ux * ux + uy * uy + uz * uz);
            pxx = pyy = pzz = pxy = pxz = pyz = 0.0f;
            static_for<Q>([&](auto qc) {
                constexpr int q = decltype(qc)::value;
                const float cu = cu_of<Q, q>(ux, uy, uz);
                const float feq = f_eq<Q, q>(rho, uu, cu);
                const float frc = force_of<Q, q>(cu, ux, uy, uz, fsx, fsy, fsz);
                const float fneq = pop[q] - feq + 0.5f * frc;
                if constexpr (kCX[q] != 0) pxx += fneq;
                if constexpr (kCY[q] != 0) pyy += fneq;
                if constexpr (kCZ[q] != 0) pzz += fneq;
                if constexpr (kCX[q] * kCY[q] > 0) pxy += fneq;
                if constexpr (kCX[q] * kCY[q] < 0) pxy -= fneq;
                if constexpr (kCX[q] * kCZ[q] > 0) pxz += fneq;
                if constexpr (kCX[q] * kCZ[q] < 0) pxz -= fneq;
                if constexpr (kCY[q] * kCZ[q] > 0) pyz += fneq;
                if constexpr (kCY[q] * kCZ[q] < 0) pyz -= fneq;
            });
            const float trace = pxx + pyy + pzz;
            pxx -= kCs2 * trace;
            pyy -= kCs2 * trace;
            pzz -= kCs2 * trace;
            static_for<Q>([&](auto qc) {
                constexpr int q = decltype(qc)::value;
                const float cu = cu_of<Q, q>(ux, uy, uz);
                const float feq = f_eq<Q, q>(rho, uu, cu);
                const float frc = force_of<Q, q>(cu, ux, uy, uz, fsx, fsy, fsz);
                const float fneq = f_neq<Q, q>(pxx, pyy, pzz, pxy, pxz, pyz, ux, uy, uz);
                const float fo = feq + omco * fneq + 0.5f * frc;
                if constexpr (PAR == 0) {
                    stf<HALF>(p, opp(q), sf, fo);
                } else {
                    stf<HALF>(p, q, nb_f<q, 1>(nb), fo);
                }
            });
        } else {
            // ---- pair-symmetric evaluation
            // Loads: pair j holds directions a = 2j+1 (c) and b = 2j+2 (-c).
            float fa[NP], fb[NP];
            const float f0 = ldf<HALF>(p, 0, sf);
            static_for<NP>([&](auto jc) {
                constexpr int j = decltype(jc)::value, a = 2 * j + 1, b = 2 * j + 2;
                if constexpr (PAR == 0) {
                    fa[j] = ldf<HALF>(p, a, sf);
                    fb[j] = ldf<HALF>(p, b, sf);
                } else {
                    fa[j] = ldf<HALF>(p, b, nb_f<a, -1>(nb));
                    fb[j] = ldf<HALF>(p, a, nb_f<a, 1>(nb));
                }
            });
            float f0r = f0;
            if constexpr (NE) {
                // In both parities the population of direction +c that
                // reaches x was written by x - c and that of -c by x + c:
                // add back their shifts (ne.hpp).
                const uint2* sm[3] = {s_m[(z + 2) % 3], s_m[z % 3], s_m[(z + 1) % 3]};
                const int o1 = (ty + 1) * NX1 + (tx + 1);
                auto m_at = [&](auto qc, auto sgn) {
                    constexpr int q = decltype(qc)::value, sg = decltype(sgn)::value;
                    return ne_unpack(sm[sg * kCZ[q] + 1][o1 + sg * (kCY[q] * NX1 + kCX[q])]);
                };
                const float kinv = p.ne_kinv;
                f0r = fmaf(f0, kinv, ne_shift<Q, 0>(ne_unpack(sm[1][o1])));
                static_for<NP>([&](auto jc) {
                    constexpr int j = decltype(jc)::value, a = 2 * j + 1, b = 2 * j + 2;
                    const float4 ma = m_at(std::integral_constant<int, a>{}, std::integral_constant<int, -1>{});
                    const float4 mb = m_at(std::integral_constant<int, a>{}, std::integral_constant<int, 1>{});
                    fa[j] = fmaf(fa[j], kinv, ne_shift<Q, a>(ma));
                    fb[j] = fmaf(fb[j], kinv, ne_shift<Q, b>(mb));
                });
            }
            // Moments from pair sums and differences: rho = 1 + sum f,
            // j = sum f c, Pi = sum f c c (src/lbm.cu:49-115).
            float sp[NP], dp[NP];
#pragma unroll
            for (int j = 0; j < NP; ++j) {
                sp[j] = fa[j] + fb[j];
                dp[j] = fa[j] - fb[j];
            }
            float m0 = f0r;
#pragma unroll
            for (int j = 0; j < NP; ++j) m0 += sp[j];
            rho = m0 + 1.0f;
            float jx = 0.0f, jy = 0.0f, jz = 0.0f, Pi[6] = {0, 0, 0, 0, 0, 0};
            static_for<NP>([&](auto jc) {
                constexpr int j = decltype(jc)::value, a = 2 * j + 1;
                constexpr int cx = kCX[a], cy = kCY[a], cz = kCZ[a];
                if constexpr (cx != 0) { jx = cx > 0 ? jx + dp[j] : jx - dp[j]; Pi[0] += sp[j]; }
                if constexpr (cy != 0) { jy = cy > 0 ? jy + dp[j] : jy - dp[j]; Pi[1] += sp[j]; }
                if constexpr (cz != 0) { jz = cz > 0 ? jz + dp[j] : jz - dp[j]; Pi[2] += sp[j]; }
                if constexpr (cx * cy != 0) Pi[3] = cx * cy > 0 ? Pi[3] + sp[j] : Pi[3] - sp[j];
                if constexpr (cx * cz != 0) Pi[4] = cx * cz > 0 ? Pi[4] + sp[j] : Pi[4] - sp[j];
                if constexpr (cy * cz != 0) Pi[5] = cy * cz > 0 ? Pi[5] + sp[j] : Pi[5] - sp[j];
            });
            const float invRho = recip<NE>(rho);
            ux = (jx + 0.5f * fsx) * invRho;
            uy = (jy + 0.5f * fsy) * invRho;
            uz = (jz + 0.5f * fsz) * invRho;
            // Deviatoric non-equilibrium stress: Pi minus the equilibrium
            // second moment rho u u (+ isotropic part) plus half the Guo
            // force moment F u + u F.
            const float axx = Pi[0] - rho * ux * ux + fsx * ux;
            const float ayy = Pi[1] - rho * uy * uy + fsy * uy;
            const float azz = Pi[2] - rho * uz * uz + fsz * uz;
            pxy = Pi[3] - rho * ux * uy + 0.5f * (fsx * uy + fsy * ux);
            pxz = Pi[4] - rho * ux * uz + 0.5f * (fsx * uz + fsz * ux);
            pyz = Pi[5] - rho * uy * uz + 0.5f * (fsy * uz + fsz * uy);
            const float tr3 = (axx + ayy + azz) * kCs2;
            pxx = axx - tr3;
            pyy = ayy - tr3;
            pzz = azz - tr3;

            const float usq = ux * ux + uy * uy + uz * uz;
            const float uF = ux * fsx + uy * fsy + uz * fsz;
            Coll k;
            k.V = 1.0f - 1.5f * usq;
            // rho V - 1 - 1.5 u.F with rho - 1 taken as the population sum m0:
            // rho = m0 + 1 rounds the density deviation to a multiple of
            // ulp(1) ~ 1.2e-7, the same way in every cell of a uniform
            // region, which would bias the mass by that much per step.
            k.B0 = m0 - 1.5f * (rho * usq + uF);
            k.R3 = 3.0f * rho;
            k.R45 = 4.5f * rho;
            k.O45 = 4.5f * omco;
            k.Pxx = pxx; k.Pyy = pyy; k.Pzz = pzz;
            k.Sxy = pxx + pyy; k.Sxz = pxx + pzz; k.Syz = pyy + pzz;
            k.S3 = k.Sxy + pzz;
            k.Txy = 2.0f * pxy; k.Txz = 2.0f * pxz; k.Tyz = 2.0f * pyz;
            if constexpr (Q == 27) {
                k.X1 = pxx * uy + 2.0f * ux * pxy;
                k.X2 = pxx * uz + 2.0f * ux * pxz;
                k.X3 = pyy * ux + 2.0f * uy * pxy;
                k.X4 = pyy * uz + 2.0f * uy * pyz;
                k.X5 = pzz * ux + 2.0f * uz * pxz;
                k.X6 = pzz * uy + 2.0f * uz * pyz;
                k.Bx = k.X3 + k.X5;
                k.By = k.X1 + k.X6;
                k.Bz = k.X2 + k.X4;
                k.X7s = 6.0f * (ux * pyz + uy * pxz + uz * pxy);
            }
            // non-equilibrium form: this cell's M, with which it stores its
            // populations and its readers rebuild them next step
            float4 mw = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            if constexpr (NE) {
                const NeM me = ne_mom(rho, ux, uy, uz);
                mw = me.m;
                p.mom_cur[c] = me.packed;
            }
            // rest direction: f_eq + F_0/2 (its regularized part is -cs^2 tr P = 0)
            {
                float o0 = wk<Q>(0) * k.B0;
                if constexpr (NE) o0 = kNeK * (o0 - ne_shift<Q, 0>(mw));
                stf<HALF>(p, 0, sf, o0);
            }
            static_for<NP>([&](auto jc) {
                constexpr int j = decltype(jc)::value, a = 2 * j + 1, b = 2 * j + 2;
                float op, om;
                // non-equilibrium form: outputs scaled by K, less K S(M) (ne.hpp)
                pair_out<Q, kCX[a], kCY[a], kCZ[a], NE>(k, ux, uy, uz, fsx, fsy, fsz, op, om);
                if constexpr (NE) ne_sub_pair<Q, a>(mw, op, om);
                if constexpr (PAR == 0) {
                    stf<HALF>(p, b, sf, op);
                    stf<HALF>(p, a, sf, om);
                } else {
                    stf<HALF>(p, a, nb_f<a, 1>(nb), op);
                    stf<HALF>(p, b, nb_f<a, -1>(nb), om);
                }
            });
        }

        // ---- phase populations (src/lbm.cu:209-238): g_eq + sharpening flux,
        // kept as the cell's phase flux h (phase_flux.hpp), from which the
        // next phase sum forms g_0 = phi/4 and g_(+-k) = phi/8 +- h_k
        {
            using HT = typename std::conditional<HALF, _Float16, float>::type;
            const float sharp = p.gamma * phi * (1.0f - phi);
            const float p4 = 4.0f * phi;
            __builtin_nontemporal_store((HT)(0.125f * (p4 * ux + sharp * nrx)), static_cast<HT*>(p.h[0]) + c);
            __builtin_nontemporal_store((HT)(0.125f * (p4 * uy + sharp * nry)), static_cast<HT*>(p.h[1]) + c);
            __builtin_nontemporal_store((HT)(0.125f * (p4 * uz + sharp * nrz)), static_cast<HT*>(p.h[2]) + c);
        }

        // ---- macroscopic fields for output and for the jet boundary planes
        if constexpr (SAVE) {
            float* const* F = p.fld;
            F[0][c] = rho; F[1][c] = ux; F[2][c] = uy; F[3][c] = uz;
            F[4][c] = pxx; F[5][c] = pyy; F[6][c] = pzz; F[7][c] = pxy; F[8][c] = pxz; F[9][c] = pyz;
            F[11][c] = nrx; F[12][c] = nry; F[13][c] = nrz; F[14][c] = ind;
            F[15][c] = fsx; F[16][c] = fsy; F[17][c] = fsz;
        }
        if constexpr (JET) {
            if (z == 1 || z == nz - 2) {
                const uint32_t plane = (uint32_t)nx * (uint32_t)ny;
                float* b = p.bcp + (z == 1 ? 0u : (uint32_t)kPlaneVals * plane) + (uint32_t)x + (uint32_t)nx * (uint32_t)y;
                const float v[kPlaneVals] = {rho, ux, uy, uz, pxx, pyy, pzz, pxy, pxz, pyz, fsx, fsy, fsz};
#pragma unroll
                for (int kk = 0; kk < kPlaneVals; ++kk) b[(uint32_t)kk * plane] = v[kk];
            }
        }
    }
}

// ---------------------------------------------------------------------------
// k_jet_bc: nozzle inflow at z = 1 and convective outflow at z = nz-2,
// written where the next step reads them.
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint32_t hash32(uint32_t x) {  // src/BoundaryConditions.cuh:172-181
    x ^= x >> 16;
    x *= 0x7FEB352Du;
    x ^= x >> 15;
    x *= 0x846CA68Bu;
    x ^= x >> 16;
    return x;
}

__device__ __forceinline__ float uniform01(uint32_t s) {  // :183-188
#pragma clang fp contract(off)
    return ((float)s + 0.5f) * 2.3283064365386963e-10f;
}

__device__ __forceinline__ uint32_t noise_base(const KParams& p, uint32_t salt, uint32_t x, uint32_t y) {
    return (p.seed ^ salt) ^ x ^ (y * 0x85EBCA6Bu) ^ (p.t * 0xC2B2AE35u);
}

// Nozzle velocity at (x, y) for this step (src/BoundaryConditions.cuh:58-69).
__device__ __forceinline__ void nozzle_u(const KParams& p, int x, int y, float& ux, float& uy, float& uz) {
#pragma clang fp contract(off)
    ux = 0.0f;
    uy = 0.0f;
    uz = p.u_inf;
    if (p.noise == 0) {  // Gaussian on u_x, u_y (Box-Muller, :190-213)
        const float sigma_u = 0.08f * p.u_inf;
        const uint32_t salts[2] = {0xA341316Cu, 0xC8013EA4u};
        float z2[2];
        for (int k = 0; k < 2; ++k) {
            const uint32_t base = noise_base(p, salts[k], (uint32_t)x, (uint32_t)y);
            float rrx = uniform01(hash32(base));
            const float rry = uniform01(hash32(base ^ 0x68BC21EBu));
            rrx = fmaxf(rrx, 1e-12f);
            const float r = sqrtf(-2.0f * logf(rrx));
            const float theta = (2.0f * 3.141592654f) * rry;
            z2[k] = r * cosf(theta);
        }
        ux = sigma_u * z2[0];
        uy = sigma_u * z2[1];
    } else if (p.noise == 1) {  // uniform +-a on u_z
        const float U = uniform01(hash32(noise_base(p, 0xA341316Cu, (uint32_t)x, (uint32_t)y)));
        uz = p.u_inf * (1.0f + p.noise_amp * (2.0f * U - 1.0f));
    }
}

__device__ __forceinline__ bool in_nozzle(const KParams& p, int x, int y) {
#pragma clang fp contract(off)
    const float dx = (float)x - p.cxg, dy = (float)y - p.cyg;
    return dx * dx + dy * dy <= p.R2;
}

// MOM: the moment form, which rebuilds populations from coefficients: the
// boundary populations go to p.bcf (kparams.hpp) instead of f.
template <int Q, bool HALF, int NPAR, bool NE = false, bool MOM = false>
__global__ __launch_bounds__(kThreads) void k_jet_bc(KParams p) {
#pragma clang fp contract(off)
    const int x = 1 + blockIdx.x * kTX + (threadIdx.x & (kTX - 1));
    const int y = 1 + blockIdx.y * kTY + (int)(threadIdx.x / kTX);
    const int nx = p.nx, ny = p.ny, nz = p.nz;
    if (x > nx - 2 || y > ny - 2) return;
    const uint32_t plane = (uint32_t)nx * (uint32_t)ny;
    const uint32_t pi = (uint32_t)x + (uint32_t)nx * (uint32_t)y;
    float* const* F = p.fld;
    // wrapped lateral neighbours: the source columns of the populations that
    // enter this column (ghost-shell periodicity)
    const int xw3[3] = {(x == 1) ? nx - 2 : x - 1, x, (x == nx - 2) ? 1 : x + 1};
    const int yw3[3] = {(y == 1) ? ny - 2 : y - 1, y, (y == ny - 2) ? 1 : y + 1};
    // Non-equilibrium form: the populations entering z = 1 are written by the
    // ghost cells of z = 0 and those entering z = nz-2 by the ghost cells of
    // z = nz-1; their M (ne.hpp) goes into the ghost planes of mom_cur, and
    // each stored value is taken relative to its source cell's shift.  The
    // nozzle lies inside the lateral ghost shell, so its source columns need
    // no wrap; the frozen wall has rho = 1, u = 0, so M = 0 and S = 0 there.
    auto inflow_m = [&](int xs, int ys) {
        NeM m{make_float4(0.0f, 0.0f, 0.0f, 0.0f), make_uint2(0u, 0u)};
        if (in_nozzle(p, xs, ys)) {
            float ux, uy, uz;
            nozzle_u(p, xs, ys, ux, uy, uz);
            m = ne_mom(1.0f, ux, uy, uz);
        }
        return m;
    };
    if constexpr (NE) p.mom_cur[pi] = inflow_m(x, y).packed;

    // ---------------- inflow, target cell (x, y, 1) ----------------
    {
        const float* b = p.bcp;
        const float tux = b[1 * plane + pi], tuy = b[2 * plane + pi], tuz = b[3 * plane + pi];
        const float pxx = b[4 * plane + pi], pyy = b[5 * plane + pi], pzz = b[6 * plane + pi];
        const float pxy = b[7 * plane + pi], pxz = b[8 * plane + pi], pyz = b[9 * plane + pi];
        const Nbr nb = make_nbr<false>(p, x, y, 1);
        static_for<Q>([&](auto qc) {
            constexpr int q = decltype(qc)::value;
            if constexpr (kCZ[q] == 1) {
                const int xs = x - kCX[q], ys = y - kCY[q];
                float val = 0.0f;  // frozen equilibrium f_eq(1, 0) - w = 0 off the nozzle
                if (in_nozzle(p, xs, ys)) {
                    float ux, uy, uz;
                    nozzle_u(p, xs, ys, ux, uy, uz);
                    const float uu = 1.5f * (ux * ux + uy * uy + uz * uz);
                    const float cu = cu_of<Q, q>(ux, uy, uz);
                    val = f_eq<Q, q>(1.0f, uu, cu) + p.omco_zmin * f_neq<Q, q>(pxx, pyy, pzz, pxy, pxz, pyz, tux, tuy, tuz);
                }
                if constexpr (NE)
                    val = p.ne_k * (val - ne_shift<Q, q>(inflow_m(xw3[1 - kCX[q]], yw3[1 - kCY[q]]).m));
                if constexpr (MOM) {
                    p.bcf[(uint32_t)((kCY[q] + 1) * 3 + kCX[q] + 1) * plane + pi] = val;
                } else {
                    int slot;
                    Loc l;
                    in_slot<q, NPAR>(nb, slot, l);
                    stf<HALF>(p, slot, l, val);
                }
            }
        });
        float g5 = 0.0f;
        if (in_nozzle(p, x, y)) {
            float ux, uy, uz;
            nozzle_u(p, x, y, ux, uy, uz);
            g5 = wq<7>(5) * 1.0f * (1.0f + kAs2g * uz);
            F[10][pi] = 1.0f;  // ghost phi at z = 0, in both phi buffers
            p.phi_next[pi] = 1.0f;
            if (p.save) {
                F[0][pi] = 1.0f;
                F[1][pi] = ux;
                F[2][pi] = uy;
                F[3][pi] = uz;
            }
        }
        p.gbc[pi] = g5;  // off the nozzle: the frozen wall equilibrium of phi = 0
    }

    // ---------------- outflow, target cell (x, y, nz-2) ----------------
    {
        const float* b = p.bcp + (uint32_t)kPlaneVals * plane;
        const float tux = b[1 * plane + pi], tuy = b[2 * plane + pi], tuz = b[3 * plane + pi];
        const float pxx = b[4 * plane + pi], pyy = b[5 * plane + pi], pzz = b[6 * plane + pi];
        const float pxy = b[7 * plane + pi], pxz = b[8 * plane + pi], pyz = b[9 * plane + pi];
        const float tfx = b[10 * plane + pi], tfy = b[11 * plane + pi], tfz = b[12 * plane + pi];
        const uint32_t zoff = plane * (uint32_t)(nz - 2);
        const uint32_t c = pi + zoff;
        const Nbr nb = make_nbr<false>(p, x, y, nz - 2);
        const float* phi_g = F[10];
        static_for<Q>([&](auto qc) {
            constexpr int q = decltype(qc)::value;
            if constexpr (kCZ[q] == -1) {
                // source column (wrapped x - c_x, y - c_y) whose boundary
                // values feed this target (src/BoundaryConditions.cuh:139-147)
                const uint32_t si = (uint32_t)xw3[1 - kCX[q]] + (uint32_t)nx * (uint32_t)yw3[1 - kCY[q]];
                const float rho_b = b[0 * plane + si];
                const float ux = b[1 * plane + si], uy = b[2 * plane + si], uz = b[3 * plane + si];
                const float phi_b = phi_g[si + zoff];
                const float uu = 1.5f * (ux * ux + uy * uy + uz * uz);
                const float cu = cu_of<Q, q>(ux, uy, uz);
                const float visc_local = (1.0f - phi_b) * p.visc_water + phi_b * p.visc_oil;
                const float omega_zmax =
                    (float)(1.0 / (0.5 + (double)kAs2 * (double)(visc_local * (1.0f + 100.0f))));
                const float omco_zmax = 1.0f - omega_zmax;
                float val = f_eq<Q, q>(rho_b, uu, cu) +
                            omco_zmax * f_neq<Q, q>(pxx, pyy, pzz, pxy, pxz, pyz, tux, tuy, tuz) +
                            0.5f * force_of<Q, q>(cu, ux, uy, uz, tfx, tfy, tfz);
                if constexpr (NE) val = p.ne_k * (val - ne_shift<Q, q>(ne_mom(rho_b, ux, uy, uz).m));
                if constexpr (MOM) {
                    p.bcf[(uint32_t)(9 + (kCY[q] + 1) * 3 + kCX[q] + 1) * plane + pi] = val;
                } else {
                    int slot;
                    Loc l;
                    in_slot<q, NPAR>(nb, slot, l);
                    stf<HALF>(p, slot, l, val);
                }
            }
        });
        const float phi_b = phi_g[c];
        const float uz_b = b[3 * plane + pi];
        p.gbc[plane + pi] = wq<7>(6) * phi_b * (1.0f - kAs2g * uz_b);
        // zero-gradient ghost plane z = nz-1, in both phi buffers
        const uint32_t cg = c + plane;
        F[10][cg] = phi_b;
        p.phi_next[cg] = phi_b;
        if constexpr (NE) p.mom_cur[cg] = ne_mom(b[0 * plane + pi], b[1 * plane + pi], b[2 * plane + pi], uz_b).packed;
        if (p.save) {
            F[0][cg] = b[0 * plane + pi];
            F[1][cg] = b[1 * plane + pi];
            F[2][cg] = b[2 * plane + pi];
            F[3][cg] = uz_b;
        }
    }
}

// ---------------------------------------------------------------------------
// launchers (instantiated per velocity set in step_q19.hip / step_q27.hip)
// ---------------------------------------------------------------------------
template <bool JET, bool LOCAL, bool HALF>
static hipError_t phase_sum_t(const KParams& p, hipStream_t s) {
    constexpr int kPsZ = ps_z(HALF);
    const int ntx = (p.nx + kTX - 1) / kTX, nty = (p.ny + kPsTY - 1) / kPsTY, nzb = (p.nz - 2 + kPsZ - 1) / kPsZ;
#if PF_PS_BANDS
    const dim3 grid((unsigned)(8 * ntx * ((nty + 7) / 8) * nzb));
#else
    const dim3 grid((unsigned)ntx, (unsigned)nty, (unsigned)nzb);
#endif
    hipLaunchKernelGGL((k_phase_sum<JET, LOCAL, HALF>), grid, dim3(kPsThreads), 0, s, p);
    return hipGetLastError();
}

template <int Q, bool JET, bool HALF, bool MLP, bool LIT, bool SAVE, bool NE>
static hipError_t step_t(const KParams& p, int parity, hipStream_t s) {
    const uint32_t nblk = (uint32_t)p.ntx * (uint32_t)p.nty * (uint32_t)p.ntz;
    if (p.tile_x == kTXNarrow && kTXNarrow != kTX) {
        if (parity)
            hipLaunchKernelGGL((k_step<Q, JET, HALF, 1, MLP, LIT, SAVE, kTXNarrow, NE>), dim3(nblk), dim3(kThreads), 0, s, p);
        else
            hipLaunchKernelGGL((k_step<Q, JET, HALF, 0, MLP, LIT, SAVE, kTXNarrow, NE>), dim3(nblk), dim3(kThreads), 0, s, p);
    } else {
        if (parity)
            hipLaunchKernelGGL((k_step<Q, JET, HALF, 1, MLP, LIT, SAVE, kTX, NE>), dim3(nblk), dim3(kThreads), 0, s, p);
        else
            hipLaunchKernelGGL((k_step<Q, JET, HALF, 0, MLP, LIT, SAVE, kTX, NE>), dim3(nblk), dim3(kThreads), 0, s, p);
    }
    return hipGetLastError();
}

template <int Q, bool JET, bool HALF>
static hipError_t step_v(const KParams& p, int parity, int variant, bool save, hipStream_t s) {
    if constexpr (HALF) {
        if (p.ne) {  // non-equilibrium form: pair collision, with or without the closure
            if (variant == 1) return hipErrorInvalidValue;
            if (variant == 2)
                return save ? step_t<Q, JET, true, true, false, true, true>(p, parity, s)
                            : step_t<Q, JET, true, true, false, false, true>(p, parity, s);
            return save ? step_t<Q, JET, true, false, false, true, true>(p, parity, s)
                        : step_t<Q, JET, true, false, false, false, true>(p, parity, s);
        }
    }
    if (variant == 1) {  // literal evaluation (FP32 and FP16 storage)
        return save ? step_t<Q, JET, HALF, false, true, true, false>(p, parity, s)
                    : step_t<Q, JET, HALF, false, true, false, false>(p, parity, s);
    }
    if (variant == 2) {
        return save ? step_t<Q, JET, HALF, true, false, true, false>(p, parity, s)
                    : step_t<Q, JET, HALF, true, false, false, false>(p, parity, s);
    }
    return save ? step_t<Q, JET, HALF, false, false, true, false>(p, parity, s)
                : step_t<Q, JET, HALF, false, false, false, false>(p, parity, s);
}

template <int Q>
static hipError_t step_q(const KParams& p, bool jet, bool half, int parity, int variant, hipStream_t s) {
    const bool save = p.save != 0;
    if (jet) return half ? step_v<Q, true, true>(p, parity, variant, save, s) : step_v<Q, true, false>(p, parity, variant, save, s);
    return half ? step_v<Q, false, true>(p, parity, variant, save, s) : step_v<Q, false, false>(p, parity, variant, save, s);
}

template <int Q, bool HALF>
static hipError_t jet_bc_t(const KParams& p, int npar, hipStream_t s) {
    dim3 grid((p.nx - 2 + kTX - 1) / kTX, (p.ny - 2 + kTY - 1) / kTY, 1);
    if (HALF && p.ne) {
        if (npar)
            hipLaunchKernelGGL((k_jet_bc<Q, true, 1, true>), grid, dim3(kThreads), 0, s, p);
        else
            hipLaunchKernelGGL((k_jet_bc<Q, true, 0, true>), grid, dim3(kThreads), 0, s, p);
    } else if (npar) {
        hipLaunchKernelGGL((k_jet_bc<Q, HALF, 1>), grid, dim3(kThreads), 0, s, p);
    } else {
        hipLaunchKernelGGL((k_jet_bc<Q, HALF, 0>), grid, dim3(kThreads), 0, s, p);
    }
    return hipGetLastError();
}

}  // namespace pf
