// The D3Q7 phase populations arriving at a cell, rebuilt from phi and the
// phase flux h (kparams.hpp).
//
// The reference stores the seven post-collision g of every cell, streams them
// and sums them into phi (src/lbm.cu:209-238, src/phaseField.cuh:31-52).  The
// g carry no non-equilibrium part (src/velocitySet/D3Q7.cuh:139-143): a cell
// with phase phi, velocity u and interface normal n emits
//   g_0 = phi/4,   g_(+-k) = phi/8 +- h_k,   h = (4 phi u + gamma phi (1-phi) n) / 8
// (D3Q7.cuh:129-153 with w = 1/4, 1/8 and c_s^2 = 1/4).  The timestep stores h
// at the cell (3 values) and the phase kernel forms the g arriving at x from
// phi and h of its six neighbours, in the same FP32 operations the
// post-collision g would have been formed with, so phi is unchanged.  In the
// binary16 form h is stored in binary16; the h of each neighbour enters once
// with + and once with - across the grid, so sum phi stays conserved.
//
// Jet planes: the g_5 entering z = 1 and the g_6 entering z = nz-2 are set by
// the boundary kernel (src/BoundaryConditions.cuh:104,168) and kept in gbc.
#pragma once

#include <hip/hip_runtime.h>

#include <type_traits>

#include "kparams.hpp"
#include "lattice.hpp"

namespace pf {

// g[q] arriving at interior cell (x, y, z), q in the reference's D3Q7 order
// (rest, +x, -x, +y, -y, +z, -z).  LOCAL: the first step after the
// populations were set, when every cell holds its own equilibrium
// (setDistros, src/initialConditions.cu:101-138) and sums that.
template <bool JET, bool LOCAL, bool HALF>
__device__ __forceinline__ void arriving_g(const KParams& p, int x, int y, int z, float (&g)[7]) {
    using HT = typename std::conditional<HALF, _Float16, float>::type;
    const HT* __restrict__ h0 = static_cast<const HT*>(p.h[0]);
    const HT* __restrict__ h1 = static_cast<const HT*>(p.h[1]);
    const HT* __restrict__ h2 = static_cast<const HT*>(p.h[2]);
    const int nx = p.nx, ny = p.ny, nz = p.nz;
    const uint32_t row = (uint32_t)nx, plane = (uint32_t)nx * (uint32_t)ny;
    const uint32_t c = (uint32_t)x + row * (uint32_t)y + plane * (uint32_t)z;
    const float* __restrict__ phi = p.fld[10];
    if constexpr (LOCAL) {
        const float ph = phi[c], ux = p.fld[1][c], uy = p.fld[2][c], uz = p.fld[3][c];
        static_for<7>([&](auto qc) {
            constexpr int q = decltype(qc)::value;
            g[q] = g_post<q>(ph, ux, uy, uz, 0.0f, 0.0f, 0.0f, 0.0f);
        });
        return;
    }
    const int xm = (x == 1) ? nx - 2 : x - 1, xp = (x == nx - 2) ? 1 : x + 1;
    const int ym = (y == 1) ? ny - 2 : y - 1, yp = (y == ny - 2) ? 1 : y + 1;
    int zm = z - 1, zp = z + 1;
    if constexpr (!JET) {
        zm = (z == 1) ? nz - 2 : z - 1;
        zp = (z == nz - 2) ? 1 : z + 1;
    }
    const uint32_t cxm = c - (uint32_t)x + (uint32_t)xm, cxp = c - (uint32_t)x + (uint32_t)xp;
    const uint32_t cym = c + row * (uint32_t)(ym - y), cyp = c + row * (uint32_t)(yp - y);
    const uint32_t czm = c + plane * (uint32_t)(zm - z), czp = c + plane * (uint32_t)(zp - z);
    g[0] = 0.25f * phi[c];
    g[1] = 0.125f * phi[cxm] + (float)h0[cxm];
    g[2] = 0.125f * phi[cxp] - (float)h0[cxp];
    g[3] = 0.125f * phi[cym] + (float)h1[cym];
    g[4] = 0.125f * phi[cyp] - (float)h1[cyp];
    if constexpr (JET) {
        const uint32_t pi = (uint32_t)x + row * (uint32_t)y;
        g[5] = (z == 1) ? p.gbc[pi] : 0.125f * phi[czm] + (float)h2[czm];
        g[6] = (z == nz - 2) ? p.gbc[plane + pi] : 0.125f * phi[czp] - (float)h2[czp];
    } else {
        g[5] = 0.125f * phi[czm] + (float)h2[czm];
        g[6] = 0.125f * phi[czp] - (float)h2[czp];
    }
}

}  // namespace pf
