// k_step_mom: one timestep of the moment storage form (mom.hpp).
//
// The same step as k_step (src/phaseField.cuh:54-219, src/lbm.cu:33-239),
// with the populations kept as the coefficients of their senders' collision
// output instead of stored one by one.  A workgroup marches a TX x TY column
// of cells through its z-chunk.  Per plane:
//   - a phi ring and the normals of plane w, as in k_step; each thread adds
//     their divergence terms to the curvature sums of its column's cells in
//     planes w - 1, w and w + 1;
//   - the coefficients of one sending plane w, with a 1-cell halo, are staged
//     in LDS, and each thread rebuilds from them the populations its column
//     receives from plane w: those with c_z = +1 belong to its cell in plane
//     w + 1, c_z = 0 to plane w, c_z = -1 to plane w - 1.  It sums each c_z
//     layer into the raw moments (sum f, sum f c, sum f c c) of the cell it
//     arrives at, so it carries the partial moments of three cells;
//   - the cell of plane w - 1 then has all its arrivals and curvature terms
//     (from planes w - 2, w - 1, w: always in that order, so the sums do not
//     depend on the tile or the z-chunk): its collision writes its own
//     coefficients, the phase flux, and on save steps the macroscopic
//     fields.
// The jet's ghost planes send the boundary populations of k_jet_bc (p.bcf)
// instead.  The first step after initialisation (local) sums each cell's
// own coefficients, the equilibrium setDistros leaves in every slot
// (src/initialConditions.cu:101-138).
#pragma once

#include "mom.hpp"
#include "step_kernels.hpp"

namespace pf {

// Tiles of TX x 4 cells (TX * 4 threads): 64 x 4 by default, whose LDS
// (40.3 KB) and registers (at most 128) let four workgroups share a CU.
// Measured on MI355X (D3Q27 256x256x1024 jet, k_step per step):
//   64 x 4, 4 workgroups per CU, the curvature marched in z (one plane of
//   normals in LDS) and a 3-plane phi ring, 16-plane z-chunks
//                                                    2.37 ms
//   the same with 24-plane z-chunks 2.39 ms, 32-plane 2.40-2.42, 64-plane
//   2.41, 128-plane 2.46, 256-plane (one round of 1024 workgroups) 2.61; at 3
//   workgroups per CU (134 registers) 2.42-2.45 ms; tile rows dealt
//   round-robin (PF_MOM_BANDS=2) 2.37 ms, the hardware order 3.17 ms;
//   every eighth z-chunk layer to one XCD (all tile rows of a layer on one
//   XCD) 2.61-2.66 ms
//   64 x 8 (66.7 KB, 2 workgroups of 8 waves per CU, a 1.29x halo tile
//   against 1.55x): 2.50-2.53 ms; 128 x 4 (79.2 KB, 2 of 8 waves per CU at
//   a 128-register cap): 2.57-2.62 ms; 32 x 4 (20.8 KB, 7 workgroups of 2
//   waves per CU): 2.69 ms
//   (timing copies, results discarded: without the coefficient staging
//   1.60 ms, without the scatter 2.08, without the normals 2.08 against
//   2.34; the fetched traffic runs at 0.91 of the copy rate)
// and with the curvature summed over three planes of normals in LDS and a
// 4-plane phi ring (51.9 KB, 3 workgroups per CU):
//   64 x 4, coefficients interleaved per row, 32-plane z-chunks
//                                                    2.49 ms
//   64 x 6 (PF_MOM_ROWS=6, 68.9 KB, 2 per CU), 43 tile rows dealt
//   round-robin or in hardware order                 4.45-4.62 ms
//   32 x 8 (44.3 KB, 125 VGPRs, 3 per CU; a 34 x 10 halo tile, 1.33x its
//   cells against 1.55x)                             2.56 ms
//   the next sending plane's coefficients loaded straight into LDS (buffer
//   loads with the lds modifier) right after the scatter, and phi of the
//   plane after it into registers, both in flight during the collision:
//   2.47 ms, no gain, so the step does not wait on its loads' latency
// and, with the earlier layout of one plane-sized array per coefficient:
//   64 x 4, 3 workgroups per CU                      2.82 ms
//   64 x 4, 2 per CU (no register cap)               2.85 ms (2.92 before
//           the hardware reciprocal); 3.42 with a 5-plane ring and full
//           normal planes (56.8 KB)
//   128 x 4, 1 per CU (102 KB)                       3.75-3.95 ms
//   32 x 4                                           3.09 ms
//   PF_MOM_PREFETCH (next plane's loads issued a plane ahead: after the
//   first barrier 3.6 ms with spills at 3 per CU, 3.13 ms at 2 per CU; after
//   the scatter 2.89 ms)
//   every lane issuing the plane's stores (those outside the interior to a
//   dropped offset), so that loads issued before them need not wait for
//   them: 2.85 ms, with the loads a plane ahead 2.92 ms
//   a scheduling fence after each sender in the scatter (80 instead of
//   168 registers): 3.07-3.13 ms
//   the ghost shell holding copies of the cells it images, so that the
//   coefficients are staged with 16-byte loads and no wrap: 2.93 ms (2.67
//   ms against 2.49 with the row-interleaved coefficients)
//   (timing copy with the coefficient staging left out: 1.93 ms; with the
//   scatter left out 2.73, with the normals and curvature left out 2.68)
// z-chunks of 16, 32 and 64 planes are within 2 %.  The droplet's step
// (periodic in z) guards its scatter with a branch as the jet's boundary
// planes do; without it the compiler hoisted all nine columns' LDS reads
// together and spilled (D3Q27 128^3 droplet: 0.104 ms against 0.170 at a
// 168-register cap and 0.196 at 128).
constexpr int kMomTY = kMomRows;
#ifndef PF_MOM_PREFETCH
#define PF_MOM_PREFETCH 0
#endif
#ifndef PF_MOM_WAVES
#define PF_MOM_WAVES 4
#endif
// PF_MOM_FAST: the hardware reciprocal and square root (1 ulp) for |grad
// phi|, the normals, omega and 1/rho (recip / root of step_kernels.hpp);
// 2.85 -> 2.82 ms.
#ifndef PF_MOM_FAST
#define PF_MOM_FAST 1
#endif

// Workgroup b runs on XCD b % 8.  PF_MOM_BANDS = 1: XCD k takes the k-th band
// of nty / 8 tile rows and sweeps it x fastest, then row by row, then
// z-chunk by z-chunk, so a tile's halo rows (coefficients and phi, 1.5x
// its own cells on 4-row tiles) were fetched by the tiles beside it into
// the same L2.  Grids whose tile rows do not divide into 8 bands keep the
// hardware order.  PF_MOM_BANDS = 2 deals whole tile rows round-robin as
// k_step does: the same time with the row-interleaved coefficients (2.52 ms
// at 16-plane chunks), the hardware order 2.98 ms with the earlier
// per-plane layout.
#ifndef PF_MOM_BANDS
#define PF_MOM_BANDS 1
#endif
__device__ __forceinline__ uint32_t xcd_band_remap(uint32_t b, uint32_t nblk, uint32_t ntx, uint32_t nty) {
    if (PF_MOM_BANDS == 2) return xcd_remap(b, nblk, ntx);  // whole tile rows dealt round-robin (k_step)
    if (PF_MOM_BANDS == 0 || nty % 8u != 0u) return b;
    const uint32_t k = b & 7u, j = b >> 3;
    const uint32_t rows = nty / 8u, layer = ntx * rows;
    const uint32_t zc = j / layer, r = j - zc * layer;
    return (zc * nty + k * rows + r / ntx) * ntx + r % ntx;
}

template <int Q, bool JET, bool SAVE, bool LOCAL, int TX>
__global__ __launch_bounds__(TX * kMomTY, TX == 64 ? PF_MOM_WAVES : 2) void k_step_mom(KParams p) {
    constexpr int NC = mom_nc<Q>();
    constexpr bool kFast = PF_MOM_FAST != 0;
    constexpr int TY = kMomTY;
    constexpr int kThreads = TX * TY;
    constexpr int HX = TX + 4, HY = TY + 4;    // phi tile with 2-cell halo
    constexpr int NX1 = TX + 2, NY1 = TY + 2;  // normal and coefficient tiles with 1-cell halo
    // LDS (40.3 KB at 64 x 4, so that four workgroups share a CU): a 3-plane
    // phi ring, the normals of one plane with a 1-cell halo, |grad phi| of 2
    // planes at the tile cells (it is read at the centre only), and the
    // coefficients of one sending plane with a 1-cell halo
    constexpr int kRing = 3;
    constexpr int kRec = NX1 * NY1;
    __shared__ float s_phi[kRing][HY * HX];
    __shared__ float s_n[3][NY1 * NX1];
    __shared__ float s_ind[2][kThreads];
    __shared__ float s_rec[NC][kRec];

    const int tid = threadIdx.x;
    const int tx = tid & (TX - 1), ty = tid / TX;
    const int nx = p.nx, ny = p.ny, nz = p.nz;
    const uint32_t lb = xcd_band_remap(blockIdx.x, gridDim.x, (uint32_t)p.ntx, (uint32_t)p.nty);
    const int bx = (int)(lb % (uint32_t)p.ntx);
    const uint32_t rem = lb / (uint32_t)p.ntx;
    const int by = (int)(rem % (uint32_t)p.nty);
    const int bz = (int)(rem / (uint32_t)p.nty);
    const int x0 = bx * TX, y0 = by * TY;
    const int zs = 1 + bz * p.zchunk;
    const int ze = min(zs + p.zchunk, nz - 1);
    const float* __restrict__ phi_g = p.fld[10];
    const uint32_t plane = (uint32_t)nx * (uint32_t)ny;

    // phi of plane pz, tile + 2-cell halo, zero outside the grid (as k_step)
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

    // Coefficients of sending plane pw over the 1-halo tile, taken at the
    // ghost-shell-wrapped coordinates of the cells that sent them (x and y;
    // z for the droplet).  The jet's ghost planes are never read from here
    // (their populations come from p.bcf); the load is clamped into the
    // interior.
    constexpr int kRIt = (kRec + kThreads - 1) / kThreads;
    auto fetch_rec = [&](int pw, float (&v)[kRIt][NC]) {
        int cz = pw;
        if constexpr (!JET) cz = (pw < 1) ? nz - 2 : (pw > nz - 2 ? 1 : pw);
        cz = min(max(cz, 1), nz - 2);
        const auto rs = rec_rsrc<Q>(p.rec_prev, plane, (uint32_t)cz);
#pragma unroll
        for (int j = 0; j < kRIt; ++j) {
            const int i = min(tid + j * kThreads, kRec - 1);
            const int gx = x0 - 1 + i % NX1, gy = y0 - 1 + i / NX1;
            const int wx = gx <= 0 ? nx - 2 : (gx >= nx - 1 ? 1 : gx);
            const int wy = gy <= 0 ? ny - 2 : (gy >= ny - 1 ? 1 : gy);
            const uint32_t ro = rec_off<Q>((uint32_t)wx, (uint32_t)wy, (uint32_t)nx);
#pragma unroll
            for (int k = 0; k < NC; ++k) v[j][k] = slot_ld(rs, ro, (uint32_t)k, (uint32_t)nx);
        }
    };
    auto put_rec = [&](const float (&v)[kRIt][NC]) {
#pragma unroll
        for (int j = 0; j < kRIt; ++j) {
            const int i = tid + j * kThreads;
            if (i < kRec) {
#pragma unroll
                for (int k = 0; k < NC; ++k) s_rec[k][i] = v[j][k];
            }
        }
    };

    // normals and |grad phi| on the 1-halo tile of plane pz
    // (src/phaseField.cuh:54-140), as k_step
    auto comp_normals = [&](int pz) {
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
                ind = root<kFast>(gx_ * gx_ + gy_ * gy_ + gz_ * gz_);
                const float inv = recip<kFast>(ind + 1e-9f);
                nrx = gx_ * inv;
                nry = gy_ * inv;
                nrz = gz_ * inv;
            }
            s_n[0][i] = nrx;
            s_n[1][i] = nry;
            s_n[2][i] = nrz;
            if (lx >= 1 && lx <= TX && ly >= 1 && ly <= TY) s_ind[pz & 1][(ly - 1) * TX + (lx - 1)] = ind;
        }
    };

    // Divergence of n (src/phaseField.cuh:168-213) marched in z as the
    // moments are: from the normals of plane pw in s_n, the terms
    // w_q c_q . n(x + c_q) of this thread's column whose neighbour lies in
    // pw, summed per c_z: tm (c_z = -1) for the cell at pw + 1, t0 for the
    // cell at pw, tp (c_z = +1) for the cell at pw - 1.  Every cell adds
    // those of planes z - 1, z, z + 1 in that order.  nself, phself: n and
    // phi of the column's own cell in pw.
    const int o = (ty + 1) * NX1 + (tx + 1);
    auto div_terms = [&](int pw, float& tm, float& t0, float& tp, float (&nself)[3], float& phself) {
        phself = s_phi[pw % kRing][(ty + 2) * HX + (tx + 2)];
        float s0 = 0.0f, s1 = 0.0f, b1 = 0.0f;
        static_for<9>([&](auto oc) {
            constexpr int oo = decltype(oc)::value, cx = oo % 3 - 1, cy = oo / 3 - 1;
            const int src = o + cy * NX1 + cx;
            const float nz_ = s_n[2][src];
            float sxy = 0.0f;
            if constexpr (cx != 0) sxy = (float)cx * s_n[0][src];
            if constexpr (cy != 0) sxy = sxy + (float)cy * s_n[1][src];
            if constexpr (cx == 0 && cy == 0) {
                nself[0] = s_n[0][src];
                nself[1] = s_n[1][src];
                nself[2] = nz_;
            }
            if constexpr (cx != 0 || cy != 0) s0 = s0 + wk<Q>((cx != 0) + (cy != 0)) * sxy;
            if constexpr (mom_has<Q>(cx, cy, 1)) {
                constexpr float w1 = wk<Q>((cx != 0) + (cy != 0) + 1);
                if constexpr (cx != 0 || cy != 0) s1 = s1 + w1 * sxy;
                b1 = b1 + w1 * nz_;
            }
        });
        t0 = s0;
        tp = s1 + b1;
        tm = s1 - b1;
    };

    // Rebuild the populations sent by plane pw to this thread's column and
    // add each c_z layer to the raw moments of the cell it reaches: am (plane
    // pw - 1), a0 (pw), ap (pw + 1).  DM / D0 / DP: which layers are wanted.
    const int o1 = (ty + 1) * NX1 + (tx + 1);
    const uint32_t pcol = (uint32_t)min(x0 + tx, nx - 1) + (uint32_t)nx * (uint32_t)min(y0 + ty, ny - 1);
    auto bc_layer = [&](int base) {
        const auto rb = __builtin_amdgcn_make_buffer_rsrc(p.bcf, (short)0, (int)(18u * plane * 4u), 0x00020000);
        float v[9];
        static_for<9>([&](auto oc) {
            constexpr int o = decltype(oc)::value, cx = o % 3 - 1, cy = o / 3 - 1;
            if constexpr (mom_has<Q>(cx, cy, 1)) v[o] = slot_ld(rb, pcol, (uint32_t)(base + o), plane);
        });
        MomLayer g{0, 0, 0, 0, 0, 0, false};
        static_for<9>([&](auto oc) {
            constexpr int o = decltype(oc)::value, cx = o % 3 - 1, cy = o / 3 - 1;
            if constexpr (mom_has<Q>(cx, cy, 1)) layer_add<cx, cy>(g, v[o]);
        });
        return g;
    };
    auto scatter = [&](auto dm_, auto d0_, auto dp_, int pw, MomAcc& am, MomAcc& a0, MomAcc& ap) {
        constexpr bool DM = decltype(dm_)::value, D0 = decltype(d0_)::value, DP = decltype(dp_)::value;
        if constexpr (JET) {
            if (pw == 0) {  // nozzle inflow and frozen wall entering z = 1
                if constexpr (DP) layer_to<1>(ap, bc_layer(0));
                return;
            }
            if (pw == nz - 1) {  // outflow entering z = nz-2
                if constexpr (DM) layer_to<-1>(am, bc_layer(9));
                return;
            }
        }
        if constexpr (!JET) {
            // a plane outside the grid sends nothing (the chunking never asks
            // for one; as a branch it also keeps the compiler from hoisting
            // all nine columns' reads together, which spilled registers)
            if (pw < 0 || pw > nz - 1) return;
        }
        MomLayer gm{0, 0, 0, 0, 0, 0, false}, g0 = gm, gp = gm;
        static_for<9>([&](auto oc) {
            constexpr int o = decltype(oc)::value, cx = o % 3 - 1, cy = o / 3 - 1;
            constexpr bool hz = mom_has<Q>(cx, cy, 1);
            const int src = o1 - cy * NX1 - cx;
            auto r = [&](int k) { return s_rec[k][src]; };
            float fm = 0.0f, f0 = 0.0f, fp = 0.0f;
            mom_column<Q, cx, cy, DM && hz, D0, DP && hz>(r, fm, f0, fp);
            if constexpr (DM && hz) layer_add<cx, cy>(gm, fm);
            if constexpr (D0) layer_add<cx, cy>(g0, f0);
            if constexpr (DP && hz) layer_add<cx, cy>(gp, fp);
        });
        if constexpr (DP) layer_to<1>(ap, gp);
        if constexpr (D0) layer_to<0>(a0, g0);
        if constexpr (DM) layer_to<-1>(am, gm);
    };
    using T = std::true_type;
    using F = std::false_type;

    // prologue: phi planes zs-2 .. zs+1, normals of zs-1 and zs, and the
    // arrivals of plane zs from zs-1 and zs (and of zs+1 from zs)
    MomAcc a_lo = mom_zero(), a_mid = mom_zero(), a_hi = mom_zero(), a_off = mom_zero();
    // curvature sums of the cells at z (k_lo) and z + 1 (k_mid); n and phi
    // of the cell at z (n_cur, ph_cur)
    float k_lo, k_mid, k_off, n_cur[3], n_next[3], ph_cur, ph_next;
    {
        float v[4][kPhiIt];
#pragma unroll
        for (int k = 0; k < 4; ++k) fetch_phi(zs - 2 + k, v[k]);
        float rv[kRIt][NC];
        if constexpr (!LOCAL) fetch_rec(zs - 1, rv);
#pragma unroll
        for (int k = 0; k < 3; ++k) put_phi(zs - 2 + k, v[k]);
        if constexpr (!LOCAL) put_rec(rv);
        lds_barrier();
        comp_normals(zs - 1);
        if constexpr (!LOCAL) scatter(F{}, F{}, T{}, zs - 1, a_off, a_off, a_lo);
        lds_barrier();
        div_terms(zs - 1, k_lo, k_off, k_off, n_next, ph_next);
        if constexpr (!LOCAL) fetch_rec(zs, rv);
        lds_barrier();
        put_phi(zs + 1, v[3]);  // into the slot of zs - 2
        if constexpr (!LOCAL) put_rec(rv);
        lds_barrier();
        comp_normals(zs);
        lds_barrier();
        if constexpr (!LOCAL) scatter(F{}, T{}, T{}, zs, a_off, a_lo, a_mid);
        {
            float t0;
            div_terms(zs, k_mid, t0, k_off, n_cur, ph_cur);
            k_lo = k_lo + t0;
        }
    }
    lds_barrier();

    const int x = x0 + tx, y = y0 + ty;
    const bool active = (x >= 1 && x <= nx - 2 && y >= 1 && y <= ny - 2);

    // PF_MOM_PREFETCH: the coefficients of sending plane z + 1 and phi of
    // plane z + 2 are loaded during the previous plane's collision, so their
    // loads are in flight while it computes (34 more registers there).
    constexpr bool kAhead = PF_MOM_PREFETCH != 0;
    float rv[kRIt][NC];
    float phv[kPhiIt];
    if constexpr (kAhead) {
        if constexpr (!LOCAL) fetch_rec(zs + 1, rv);
        fetch_phi(zs + 2, phv);
    }
    for (int z = zs; z < ze; ++z) {
        if constexpr (!kAhead) {
            if constexpr (!LOCAL) fetch_rec(z + 1, rv);
            fetch_phi(z + 2, phv);
        }
        put_phi(z + 2, phv);
        if constexpr (!LOCAL) put_rec(rv);
        lds_barrier();
        comp_normals(z + 1);
        if constexpr (!LOCAL) scatter(T{}, T{}, T{}, z + 1, a_lo, a_mid, a_hi);
        lds_barrier();
        if (kAhead && z + 1 < ze) {
            if constexpr (!LOCAL) fetch_rec(z + 2, rv);
            fetch_phi(z + 3, phv);
        }

        // the arrivals of plane z are complete, and with the normals of
        // plane z + 1 its curvature sum
        MomAcc a = a_lo;
        a_lo = a_mid;
        a_mid = a_hi;
        a_hi = mom_zero();
        float ksum;
        {
            float tm, t0, tp;
            div_terms(z + 1, tm, t0, tp, n_next, ph_next);
            ksum = k_lo + tp;
            k_lo = k_mid + t0;
            k_mid = tm;
        }
        const float nrx = n_cur[0], nry = n_cur[1], nrz = n_cur[2];
        n_cur[0] = n_next[0];
        n_cur[1] = n_next[1];
        n_cur[2] = n_next[2];
        const float phi = ph_cur;
        ph_cur = ph_next;

        // ---- curvature and continuum surface force (src/phaseField.cuh:142-219)
        if (!active) continue;
        const float curvature = kAs2 * ksum;
        const uint32_t c = (uint32_t)x + (uint32_t)nx * ((uint32_t)y + (uint32_t)ny * (uint32_t)z);
        const float ind = s_ind[z & 1][tid];
        const float stCurv = -p.sigma * curvature * ind;
        const float fsx = stCurv * nrx, fsy = stCurv * nry, fsz = stCurv * nrz;

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
            omega = recip<kFast>(tau_eff);
        } else {
            omega = p.omega_ref;
        }
        const float omco = 1.0f - omega;

        if constexpr (LOCAL) {
            const auto rs = rec_rsrc<Q>(p.rec_prev, plane, (uint32_t)z);
            float own[NC];
#pragma unroll
            for (int k = 0; k < NC; ++k)
                own[k] = slot_ld(rs, rec_off<Q>((uint32_t)x, (uint32_t)y, (uint32_t)nx), (uint32_t)k, (uint32_t)nx);
            a = mom_local<Q>([&](int k) { return own[k]; });
        }

        // ---- moments (src/lbm.cu:49-120): rho = 1 + sum f~, u with half the
        // force, deviatoric non-equilibrium stress
        const float m0 = a.m0;
        const float rho = m0 + 1.0f;
        const float invRho = recip<kFast>(rho);
        const float ux = (a.jx + 0.5f * fsx) * invRho;
        const float uy = (a.jy + 0.5f * fsy) * invRho;
        const float uz = (a.jz + 0.5f * fsz) * invRho;
        const float axx = a.xx - rho * ux * ux + fsx * ux;
        const float ayy = a.yy - rho * uy * uy + fsy * uy;
        const float azz = a.zz - rho * uz * uz + fsz * uz;
        const float pxy = a.xy - rho * ux * uy + 0.5f * (fsx * uy + fsy * ux);
        const float pxz = a.xz - rho * ux * uz + 0.5f * (fsx * uz + fsz * ux);
        const float pyz = a.yz - rho * uy * uz + 0.5f * (fsy * uz + fsz * uy);
        const float tr3 = (axx + ayy + azz) * kCs2;
        const float pxx = axx - tr3, pyy = ayy - tr3, pzz = azz - tr3;

        // ---- collision: this cell's coefficients (mom.hpp)
        {
            float cf[NC];
            mom_coeffs<Q>(m0, rho, ux, uy, uz, fsx, fsy, fsz, pxx, pyy, pzz, pxy, pxz, pyz, omco, cf);
            const auto rs = rec_rsrc<Q>(p.rec_cur, plane, (uint32_t)z);
#pragma unroll
            for (int k = 0; k < NC; ++k)
                __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(uint32_t, cf[k]), rs,
                                                      rec_off<Q>((uint32_t)x, (uint32_t)y, (uint32_t)nx) * 4u,
                                                      (uint32_t)k * (uint32_t)nx * 4u, kStoreAux);
        }

        // ---- phase flux (src/lbm.cu:209-238, phase_flux.hpp)
        {
            const float sharp = p.gamma * phi * (1.0f - phi);
            const float p4 = 4.0f * phi;
            __builtin_nontemporal_store(0.125f * (p4 * ux + sharp * nrx), static_cast<float*>(p.h[0]) + c);
            __builtin_nontemporal_store(0.125f * (p4 * uy + sharp * nry), static_cast<float*>(p.h[1]) + c);
            __builtin_nontemporal_store(0.125f * (p4 * uz + sharp * nrz), static_cast<float*>(p.h[2]) + c);
        }

        // ---- macroscopic fields for output and for the jet boundary planes
        if constexpr (SAVE) {
            float* const* Fd = p.fld;
            Fd[0][c] = rho; Fd[1][c] = ux; Fd[2][c] = uy; Fd[3][c] = uz;
            Fd[4][c] = pxx; Fd[5][c] = pyy; Fd[6][c] = pzz; Fd[7][c] = pxy; Fd[8][c] = pxz; Fd[9][c] = pyz;
            Fd[11][c] = nrx; Fd[12][c] = nry; Fd[13][c] = nrz; Fd[14][c] = ind;
            Fd[15][c] = fsx; Fd[16][c] = fsy; Fd[17][c] = fsz;
        }
        if constexpr (JET) {
            if (z == 1 || z == nz - 2) {
                float* b = p.bcp + (z == 1 ? 0u : (uint32_t)kPlaneVals * plane) + (uint32_t)x + (uint32_t)nx * (uint32_t)y;
                const float v[kPlaneVals] = {rho, ux, uy, uz, pxx, pyy, pzz, pxy, pxz, pyz, fsx, fsy, fsz};
#pragma unroll
                for (int kk = 0; kk < kPlaneVals; ++kk) b[(uint32_t)kk * plane] = v[kk];
            }
        }
    }
}

template <int Q, bool JET, bool SAVE, bool LOCAL>
static hipError_t step_mom_t(const KParams& p, hipStream_t s) {
    const uint32_t nblk = (uint32_t)p.ntx * (uint32_t)p.nty * (uint32_t)p.ntz;
    if (p.tile_x == 128)
        hipLaunchKernelGGL((k_step_mom<Q, JET, SAVE, LOCAL, 128>), dim3(nblk), dim3(128 * kMomTY), 0, s, p);
    else if (p.tile_x == 64)
        hipLaunchKernelGGL((k_step_mom<Q, JET, SAVE, LOCAL, 64>), dim3(nblk), dim3(64 * kMomTY), 0, s, p);
    else
        return hipErrorInvalidValue;
    return hipGetLastError();
}

template <int Q, bool JET>
static hipError_t step_mom_j(const KParams& p, bool local, hipStream_t s) {
    if (p.save != 0) return local ? step_mom_t<Q, JET, true, true>(p, s) : step_mom_t<Q, JET, true, false>(p, s);
    return local ? step_mom_t<Q, JET, false, true>(p, s) : step_mom_t<Q, JET, false, false>(p, s);
}

template <int Q>
static hipError_t step_mom_q(const KParams& p, bool jet, bool local, hipStream_t s) {
    return jet ? step_mom_j<Q, true>(p, local, s) : step_mom_j<Q, false>(p, local, s);
}

template <int Q>
static hipError_t jet_bc_mom_q(const KParams& p, hipStream_t s) {
    dim3 grid((p.nx - 2 + kTX - 1) / kTX, (p.ny - 2 + kTY - 1) / kTY, 1);
    hipLaunchKernelGGL((k_jet_bc<Q, false, 0, false, true>), grid, dim3(kThreads), 0, s, p);
    return hipGetLastError();
}

}  // namespace pf
