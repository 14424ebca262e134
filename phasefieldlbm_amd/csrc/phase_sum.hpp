// k_phase_row: phi = sum_q g_q over interior cells (src/phaseField.cuh:31-52)
// with the seven g rebuilt from the neighbours' phi and FP32 phase flux
// (phase_flux.hpp), for grids whose rows are a multiple of 4 cells.
//
// Each thread owns 4 consecutive cells of a row and marches them through a
// z-chunk.  Per plane it issues one 16-byte load per array and row it needs: phi and h_x of its own row, phi and h_y of the
// rows above and below, phi and h_z of the next plane (the planes below and
// at z are carried in registers).  The x-neighbours at the ends of its 4
// cells come from the adjacent lanes of the row (cross-lane moves); only the
// lanes at the ends of a tile row load them, and the lanes holding x = 1 and
// x = nx-2 load the ghost-shell-wrapped neighbours (period nx-2,
// src/lbm.cu:194-198).  The per-cell arithmetic and summation order are those
// of arriving_g + k_phase_sum, so phi is bit-for-bit the same.
//
// Tile: TXC lanes x 4 cells in x, 256 / TXC rows, one z-chunk.  Workgroups
// are dealt to the XCDs in bands of tile rows (workgroup b runs on XCD
// b % 8, which takes the (b % 8)-th band of ceil(rows / 8) tile rows and
// sweeps it x fastest, then through the band, then in z), so the rows above
// and below a tile are fetched into the same L2 as the tile.
#pragma once

#include <hip/hip_runtime.h>

#include "kparams.hpp"

namespace pf {

constexpr int kPrThreads = 256;

__device__ __forceinline__ float f4_at(const float4& v, int e) {
    return e == 0 ? v.x : e == 1 ? v.y : e == 2 ? v.z : v.w;
}

// Value of lane (lane - 1) / (lane + 1) within a group of TXC lanes; the
// lanes at the group's ends get their own value back (they load instead).
template <int TXC>
__device__ __forceinline__ float from_left(float v) {
    return __shfl_up(v, 1, TXC);
}
template <int TXC>
__device__ __forceinline__ float from_right(float v) {
    return __shfl_down(v, 1, TXC);
}

template <bool JET, int TXC>
__global__ __launch_bounds__(kPrThreads) void k_phase_row(KParams p, int ZC) {
    constexpr int TR = kPrThreads / TXC;  // rows per tile
    constexpr int TW = 4 * TXC;           // cells per tile row
    const int nx = p.nx, ny = p.ny, nz = p.nz;
    const int ntx = (nx + TW - 1) / TW, nty = (ny + TR - 1) / TR;
    // band dealing of the tiles over the XCDs
    const int band = (nty + 7) / 8;
    const int k = (int)(blockIdx.x & 7u), j = (int)(blockIdx.x >> 3);
    const int bx = j % ntx;
    const int by = k * band + (j / ntx) % band;
    const int bz = j / (ntx * band);
    if (by >= nty) return;  // uniform: this XCD's band ends inside the grid
    const int lc = (int)threadIdx.x % TXC, r = (int)threadIdx.x / TXC;
    const int x0 = bx * TW + 4 * lc;
    const int y = by * TR + r;
    const int z0 = 1 + bz * ZC;
    const int z1 = min(z0 + ZC, nz - 1);
    // Lanes outside the grid still take part in the cross-lane moves: their
    // addresses are clamped into the grid and they store nothing.
    const bool live = x0 < nx && y >= 1 && y <= ny - 2;
    const int xc = min(x0, nx - 4), ycl = min(max(y, 1), ny - 2);
    const uint32_t row = (uint32_t)nx, plane = (uint32_t)nx * (uint32_t)ny;
    const int ym = (ycl == 1) ? ny - 2 : ycl - 1, yp = (ycl == ny - 2) ? 1 : ycl + 1;
    const uint32_t col = (uint32_t)xc + row * (uint32_t)ycl;
    const uint32_t colm = (uint32_t)xc + row * (uint32_t)ym, colp = (uint32_t)xc + row * (uint32_t)yp;
    // the ends of this thread's cells: neighbours loaded rather than moved
    const bool left_load = (lc == 0) && xc >= 1;
    const bool right_load = (lc == TXC - 1) && xc + 4 <= nx - 1;
    const bool wrap_l = (xc == 0), wrap_r = (xc + 4 == nx);  // x = 1 / x = nx-2 in these cells
    const float* __restrict__ h0 = static_cast<const float*>(p.h[0]);
    const float* __restrict__ h1 = static_cast<const float*>(p.h[1]);
    const float* __restrict__ h2 = static_cast<const float*>(p.h[2]);
    const float* __restrict__ phi = p.fld[10];
    auto zwrap = [&](int z) {
        if constexpr (!JET) {
            if (z < 1) return nz - 2;
            if (z > nz - 2) return 1;
        }
        return z;
    };
    auto ld4 = [&](const float* a, uint32_t i) { return *reinterpret_cast<const float4*>(a + i); };
    uint32_t cm = col + plane * (uint32_t)zwrap(z0 - 1);
    uint32_t cc = col + plane * (uint32_t)z0;
    float4 phi_m = ld4(phi, cm), phi_c = ld4(phi, cc);
    float4 hz_m = ld4(h2, cm), hz_c = ld4(h2, cc);
    for (int z = z0; z < z1; ++z) {
        const uint32_t cp = col + plane * (uint32_t)zwrap(z + 1);
        const uint32_t dz = plane * (uint32_t)z;
        // every load of the plane first
        const float4 phi_p = ld4(phi, cp), hz_p = ld4(h2, cp);
        const float4 phi_ym = ld4(phi, colm + dz), phi_yp = ld4(phi, colp + dz);
        const float4 hy_m = ld4(h1, colm + dz), hy_p = ld4(h1, colp + dz);
        const float4 hx = ld4(h0, cc);
        float phi_l = 0.0f, hx_l = 0.0f, phi_r = 0.0f, hx_r = 0.0f;
        if (left_load) {
            phi_l = phi[cc - 1];
            hx_l = h0[cc - 1];
        }
        if (right_load) {
            phi_r = phi[cc + 4];
            hx_r = h0[cc + 4];
        }
        float phi_wl = 0.0f, hx_wl = 0.0f, phi_wr = 0.0f, hx_wr = 0.0f;
        if (wrap_l) {  // x = 1: its -x neighbour is x = nx-2
            phi_wl = phi[cc + (uint32_t)(nx - 2)];
            hx_wl = h0[cc + (uint32_t)(nx - 2)];
        }
        if (wrap_r) {  // x = nx-2: its +x neighbour is x = 1
            phi_wr = phi[cc + 1u - (uint32_t)(nx - 4)];
            hx_wr = h0[cc + 1u - (uint32_t)(nx - 4)];
        }
        // neighbours across the ends of the 4 cells from the adjacent lanes
        const float mv_phi_l = from_left<TXC>(phi_c.w), mv_hx_l = from_left<TXC>(f4_at(hx, 3));
        const float mv_phi_r = from_right<TXC>(phi_c.x), mv_hx_r = from_right<TXC>(f4_at(hx, 0));
        if (lc != 0) {
            phi_l = mv_phi_l;
            hx_l = mv_hx_l;
        }
        if (lc != TXC - 1) {
            phi_r = mv_phi_r;
            hx_r = mv_hx_r;
        }
        float out[4];
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            float pl = e > 0 ? f4_at(phi_c, e - 1) : phi_l, hl = e > 0 ? f4_at(hx, e - 1) : hx_l;
            float pr = e < 3 ? f4_at(phi_c, e + 1) : phi_r, hr = e < 3 ? f4_at(hx, e + 1) : hx_r;
            if (e == 1 && wrap_l) {
                pl = phi_wl;
                hl = hx_wl;
            }
            if (e == 2 && wrap_r) {
                pr = phi_wr;
                hr = hx_wr;
            }
            float g[7];
            g[0] = 0.25f * f4_at(phi_c, e);
            g[1] = 0.125f * pl + hl;
            g[2] = 0.125f * pr - hr;
            g[3] = 0.125f * f4_at(phi_ym, e) + f4_at(hy_m, e);
            g[4] = 0.125f * f4_at(phi_yp, e) - f4_at(hy_p, e);
            g[5] = 0.125f * f4_at(phi_m, e) + f4_at(hz_m, e);
            g[6] = 0.125f * f4_at(phi_p, e) - f4_at(hz_p, e);
            if constexpr (JET) {
                if (z == 1) g[5] = p.gbc[col + (uint32_t)e];
                if (z == nz - 2) g[6] = p.gbc[plane + col + (uint32_t)e];
            }
            float sum = 0.0f;
#pragma unroll
            for (int q = 0; q < 7; ++q) sum += g[q];
            out[e] = sum;
        }
        if (live) {
            if (xc >= 1 && xc + 3 <= nx - 2) {
                *reinterpret_cast<float4*>(p.phi_next + cc) = make_float4(out[0], out[1], out[2], out[3]);
            } else {
#pragma unroll
                for (int e = 0; e < 4; ++e)
                    if (xc + e >= 1 && xc + e <= nx - 2) p.phi_next[cc + (uint32_t)e] = out[e];
            }
        }
        phi_m = phi_c;
        phi_c = phi_p;
        hz_m = hz_c;
        hz_c = hz_p;
        cc += plane;  // the next plane of the column (no wrap inside it)
    }
}

}  // namespace pf
