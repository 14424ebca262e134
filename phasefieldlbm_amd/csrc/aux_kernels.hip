// Initial conditions, population gather and FP64 reductions.
//
//   k_init_fields    setInitialDensity + setJet (src/initialConditions.cu:33-99);
//                    the droplet profile is computed on the host
//   k_init_distros   setDistros (src/initialConditions.cu:101-138) for f,
//                    written in the layout the first (even) A-A step reads;
//                    the first phase sum forms the equilibrium g itself
//   k_gather         post-streaming populations in the reference layout
//   k_diagnostics    FP64 sums of rho, rho*u, phi and max|u| (no reference
//                    counterpart; the reference prints no diagnostics)
//   k_phase_mass     FP64 sum of phi = sum_q g_q over the arriving populations
#include <hip/hip_runtime.h>

#include "kparams.hpp"
#include "lattice.hpp"
#include "mom.hpp"
#include "ne.hpp"
#include "phase_flux.hpp"

namespace pf {

// Element index of population `slot` of cell (x, y, z) in an array with row
// pitch `pitch` and plane pitch `plane` (the layout of kparams.hpp).
__device__ __forceinline__ uint64_t pop_elem(uint64_t plane, uint32_t pitch, int nx, int slot, int x, int y, int z) {
    return (uint64_t)z * plane + (uint64_t)y * pitch + (uint64_t)slot * (uint32_t)nx + (uint32_t)x;
}
template <bool HALF>
__device__ __forceinline__ void st_pop(void* base, uint64_t e, float v) {
    if constexpr (HALF)
        reinterpret_cast<_Float16*>(base)[e] = (_Float16)v;
    else
        reinterpret_cast<float*>(base)[e] = v;
}
template <bool HALF>
__device__ __forceinline__ float ld_pop(const void* base, uint64_t e) {
    if constexpr (HALF)
        return (float)reinterpret_cast<const _Float16*>(base)[e];
    else
        return reinterpret_cast<const float*>(base)[e];
}

__global__ __launch_bounds__(256) void k_init_fields(KParams p, int jet) {
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i >= p.N) return;
    p.fld[0][i] = 1.0f;
    if (jet) {
        const uint32_t plane = (uint32_t)p.nx * (uint32_t)p.ny;
        if (i < plane) {
#pragma clang fp contract(off)
            const int x = (int)(i % (uint32_t)p.nx), y = (int)(i / (uint32_t)p.nx);
            const float dx = (float)x - p.cxg, dy = (float)y - p.cyg;
            if (dx * dx + dy * dy <= p.R2) {
                p.fld[10][i] = 1.0f;
                p.fld[3][i] = p.u_inf;
            }
        }
    }
}

// Cell indices of the 3x3x3 neighbourhood with the ghost-shell wrap of the
// timestep (x, y always; z for the droplet), as make_nbr.
template <bool WRAPZ>
__device__ __forceinline__ void cell_nbr(const KParams& p, int x, int y, int z, uint32_t (&cx)[3], uint32_t (&cy)[3],
                                         uint32_t (&cz)[3]) {
    const int nx = p.nx, ny = p.ny, nz = p.nz;
    const int xm = (x == 1) ? nx - 2 : x - 1, xp = (x == nx - 2) ? 1 : x + 1;
    const int ym = (y == 1) ? ny - 2 : y - 1, yp = (y == ny - 2) ? 1 : y + 1;
    int zm = z - 1, zp = z + 1;
    if constexpr (WRAPZ) {
        zm = (z == 1) ? nz - 2 : z - 1;
        zp = (z == nz - 2) ? 1 : z + 1;
    }
    const uint32_t row = (uint32_t)nx, plane = (uint32_t)nx * (uint32_t)ny;
    cx[0] = (uint32_t)xm; cx[1] = (uint32_t)x; cx[2] = (uint32_t)xp;
    cy[0] = row * (uint32_t)ym; cy[1] = row * (uint32_t)y; cy[2] = row * (uint32_t)yp;
    cz[0] = plane * (uint32_t)zm; cz[1] = plane * (uint32_t)z; cz[2] = plane * (uint32_t)zp;
}

// Non-equilibrium form: the population of direction q that cell (x, y, z)
// holds for the first (even) step counts as written by x - c_q, so it is
// stored relative to that cell's shift; every cell's M goes into both
// M buffers.  Ghost cells hold zeros (no step reads them before the boundary
// kernel rewrites them).
template <int Q, bool HALF, bool NE, bool WRAPZ>
__global__ __launch_bounds__(256) void k_init_distros(KParams p) {
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i >= p.N) return;
    const float rho = p.fld[0][i], ux = p.fld[1][i], uy = p.fld[2][i], uz = p.fld[3][i];
    const float uu = 1.5f * (ux * ux + uy * uy + uz * uz);
    const int x = (int)(i % (uint32_t)p.nx);
    const int y = (int)((i / (uint32_t)p.nx) % (uint32_t)p.ny);
    const int z = (int)(i / ((uint32_t)p.nx * (uint32_t)p.ny));
    if constexpr (NE) {
        const NeM me = ne_mom(rho, ux, uy, uz);
        p.mom_cur[i] = me.packed;
        const_cast<uint2*>(p.mom_prev)[i] = me.packed;
        const bool interior = x >= 1 && x <= p.nx - 2 && y >= 1 && y <= p.ny - 2 && z >= 1 && z <= p.nz - 2;
        uint32_t cx[3] = {0, 0, 0}, cy[3] = {0, 0, 0}, cz[3] = {0, 0, 0};
        if (interior) cell_nbr<WRAPZ>(p, x, y, z, cx, cy, cz);
        static_for<Q>([&](auto qc) {
            constexpr int q = decltype(qc)::value;
            float v = 0.0f;
            if (interior) {
                const uint32_t w = cx[1 - kCX[q]] + cy[1 - kCY[q]] + cz[1 - kCZ[q]];
                const float4 mw = ne_mom(p.fld[0][w], p.fld[1][w], p.fld[2][w], p.fld[3][w]).m;
                v = p.ne_k * (f_eq<Q, q>(rho, uu, cu_of<Q, q>(ux, uy, uz)) - ne_shift<Q, q>(mw));
            }
            st_pop<HALF>(p.f, pop_elem(p.plane_f, p.pitch_f, p.nx, q, x, y, z), v);
        });
        return;
    }
    static_for<Q>([&](auto qc) {
        constexpr int q = decltype(qc)::value;
        st_pop<HALF>(p.f, pop_elem(p.plane_f, p.pitch_f, p.nx, q, x, y, z), f_eq<Q, q>(rho, uu, cu_of<Q, q>(ux, uy, uz)));
    });
}

// Moment form (mom.hpp): every cell's coefficients of its own equilibrium
// f~eq(rho, u) (no stress, no force), in both coefficient buffers; the
// first step reads each cell's own (setDistros, src/initialConditions.cu:101-138).
template <int Q>
__global__ __launch_bounds__(256) void k_init_mom(KParams p) {
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i >= p.N) return;
    const float rho = p.fld[0][i], ux = p.fld[1][i], uy = p.fld[2][i], uz = p.fld[3][i];
    float cf[mom_nc<Q>()];
    mom_coeffs<Q>(rho - 1.0f, rho, ux, uy, uz, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, cf);
    for (int k = 0; k < mom_nc<Q>(); ++k) {
        const uint64_t e = rec_elem<Q>((uint32_t)k, i, (uint32_t)p.nx, (uint32_t)p.ny);
        p.rec_cur[e] = cf[k];
        const_cast<float*>(p.rec_prev)[e] = cf[k];
    }
}

// Moment form: the populations the next step reads, rebuilt from the
// coefficients of the cells that sent them (wrapped x - c_q), or from the
// jet's boundary populations at z = 1 (c_z = +1) and z = nz-2 (c_z = -1), by
// the same operations as the step (mom_pop); LOCAL: each cell's own.
template <int Q, bool JET, bool LOCAL>
__global__ __launch_bounds__(256) void k_gather_mom(KParams p, float* out) {
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i >= p.N) return;
    const int nx = p.nx, ny = p.ny, nz = p.nz;
    const int x = (int)(i % (uint32_t)nx);
    const int y = (int)((i / (uint32_t)nx) % (uint32_t)ny);
    const int z = (int)(i / ((uint32_t)nx * (uint32_t)ny));
    const bool interior = x >= 1 && x <= nx - 2 && y >= 1 && y <= ny - 2 && z >= 1 && z <= nz - 2;
    uint32_t cx[3] = {0, 0, 0}, cy[3] = {0, 0, 0}, cz[3] = {0, 0, 0};
    if (interior) cell_nbr<!JET>(p, x, y, z, cx, cy, cz);
    const uint32_t plane = (uint32_t)nx * (uint32_t)ny, pi = (uint32_t)x + (uint32_t)nx * (uint32_t)y;
    static_for<Q>([&](auto qc) {
        constexpr int q = decltype(qc)::value;
        float v = 0.0f;
        if (interior) {
            const uint32_t w = LOCAL ? i : cx[1 - kCX[q]] + cy[1 - kCY[q]] + cz[1 - kCZ[q]];
            if (JET && !LOCAL && kCZ[q] == 1 && z == 1) {
                v = p.bcf[(uint32_t)((kCY[q] + 1) * 3 + kCX[q] + 1) * plane + pi];
            } else if (JET && !LOCAL && kCZ[q] == -1 && z == nz - 2) {
                v = p.bcf[(uint32_t)(9 + (kCY[q] + 1) * 3 + kCX[q] + 1) * plane + pi];
            } else {
                v = mom_pop<Q, q>([&](int k) { return p.rec_prev[rec_elem<Q>((uint32_t)k, w, (uint32_t)nx, (uint32_t)ny)]; });
            }
        }
        out[(size_t)q * p.N + i] = v;
    });
}

// which = 0: f (Q slots, stored shift -w kept); which = 1: g (7 slots, the
// populations the next phase sum adds)
template <int Q, bool JET, bool HALF, int PAR, bool LOCAL, bool NE = false>
__global__ __launch_bounds__(256) void k_gather(KParams p, int which, float* out) {
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i >= p.N) return;
    const int nx = p.nx, ny = p.ny, nz = p.nz;
    const int x = (int)(i % (uint32_t)nx);
    const int y = (int)((i / (uint32_t)nx) % (uint32_t)ny);
    const int z = (int)(i / ((uint32_t)nx * (uint32_t)ny));
    const bool interior = x >= 1 && x <= nx - 2 && y >= 1 && y <= ny - 2 && z >= 1 && z <= nz - 2;
    const int xm = (x == 1) ? nx - 2 : x - 1, xp = (x == nx - 2) ? 1 : x + 1;
    const int ym = (y == 1) ? ny - 2 : y - 1, yp = (y == ny - 2) ? 1 : y + 1;
    int zm = z - 1, zp = z + 1;
    if constexpr (!JET) {
        zm = (z == 1) ? nz - 2 : z - 1;
        zp = (z == nz - 2) ? 1 : z + 1;
    }
    const int xs[3] = {xm, x, xp}, ys[3] = {ym, y, yp}, zs[3] = {zm, z, zp};
    if (which == 1) {
        float g[7] = {0, 0, 0, 0, 0, 0, 0};
        if (interior) arriving_g<JET, LOCAL, HALF>(p, x, y, z, g);
        for (int q = 0; q < 7; ++q) out[(size_t)q * p.N + i] = g[q];
        return;
    }
    if constexpr (NE) {
        // stored value plus the shift of the cell that wrote it, x - c_q (ne.hpp)
        uint32_t cx[3] = {0, 0, 0}, cy[3] = {0, 0, 0}, cz[3] = {0, 0, 0};
        if (interior) cell_nbr<!JET>(p, x, y, z, cx, cy, cz);
        static_for<Q>([&](auto qc) {
            constexpr int q = decltype(qc)::value;
            float v = 0.0f;
            if (interior) {
                int sx = x, sy = y, sz = z, slot = q;
                if constexpr (PAR == 1) {
                    slot = opp(q);
                    sx = xs[1 - kCX[q]];
                    sy = ys[1 - kCY[q]];
                    sz = zs[1 - kCZ[q]];
                }
                const uint32_t w = cx[1 - kCX[q]] + cy[1 - kCY[q]] + cz[1 - kCZ[q]];
                v = fmaf(ld_pop<HALF>(p.f, pop_elem(p.plane_f, p.pitch_f, nx, slot, sx, sy, sz)), p.ne_kinv,
                         ne_shift<Q, q>(ne_unpack(p.mom_prev[w])));
            }
            out[(size_t)q * p.N + i] = v;
        });
        return;
    }
    for (int q = 0; q < Q; ++q) {
        float v = 0.0f;
        if (interior) {
            int sx = x, sy = y, sz = z, slot = q;
            if constexpr (PAR == 1) {
                slot = opp(q);
                sx = xs[1 - kCX[q]];
                sy = ys[1 - kCY[q]];
                sz = zs[1 - kCZ[q]];
            }
            v = ld_pop<HALF>(p.f, pop_elem(p.plane_f, p.pitch_f, nx, slot, sx, sy, sz));
        }
        out[(size_t)q * p.N + i] = v;
    }
}

// Maximum that propagates NaN: fmax(NaN, b) returns b, which would let a
// non-finite velocity vanish from max|u|.
__device__ __forceinline__ double nan_max(double a, double b) {
    return (a != a || b != b) ? __builtin_nan("") : fmax(a, b);
}

// Block reduction of NV doubles (sum, except the last, which is max when
// LASTMAX).  256 threads = 4 wavefronts of 64.
template <int NV, bool LASTMAX>
__device__ __forceinline__ void block_reduce(double (&v)[NV], double* out) {
    __shared__ double s[4][NV];
#pragma unroll
    for (int k = 0; k < NV; ++k) {
        double a = v[k];
#pragma unroll
        for (int off = 32; off >= 1; off >>= 1) {
            const double b = __shfl_xor(a, off, 64);
            a = (LASTMAX && k == NV - 1) ? nan_max(a, b) : a + b;
        }
        v[k] = a;
    }
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (lane == 0)
#pragma unroll
        for (int k = 0; k < NV; ++k) s[wid][k] = v[k];
    __syncthreads();
    if (threadIdx.x == 0) {
#pragma unroll
        for (int k = 0; k < NV; ++k) {
            double a = s[0][k];
            for (int w = 1; w < 4; ++w) a = (LASTMAX && k == NV - 1) ? nan_max(a, s[w][k]) : a + s[w][k];
            out[k] = a;
        }
    }
}

__global__ __launch_bounds__(256) void k_diagnostics(KParams p, double* partials) {
    double v[6] = {0, 0, 0, 0, 0, 0};
    const int nx = p.nx, ny = p.ny, nz = p.nz;
    for (uint32_t i = blockIdx.x * 256u + threadIdx.x; i < p.N; i += gridDim.x * 256u) {
        const int x = (int)(i % (uint32_t)nx);
        const int y = (int)((i / (uint32_t)nx) % (uint32_t)ny);
        const int z = (int)(i / ((uint32_t)nx * (uint32_t)ny));
        if (x < 1 || x > nx - 2 || y < 1 || y > ny - 2 || z < 1 || z > nz - 2) continue;
        const double rho = p.fld[0][i], ux = p.fld[1][i], uy = p.fld[2][i], uz = p.fld[3][i];
        v[0] += rho;
        v[1] += rho * ux;
        v[2] += rho * uy;
        v[3] += rho * uz;
        v[4] += (double)p.fld[10][i];
        const double um = sqrt(ux * ux + uy * uy + uz * uz);
        v[5] = nan_max(v[5], um);  // a NaN, once seen, stays
    }
    block_reduce<6, true>(v, partials + (size_t)blockIdx.x * 6);
}

template <bool JET, bool LOCAL, bool HALF>
__global__ __launch_bounds__(256) void k_phase_mass(KParams p, double* partials) {
    double v[1] = {0.0};
    const int nx = p.nx, ny = p.ny, nz = p.nz;
    for (uint32_t i = blockIdx.x * 256u + threadIdx.x; i < p.N; i += gridDim.x * 256u) {
        const int x = (int)(i % (uint32_t)nx);
        const int y = (int)((i / (uint32_t)nx) % (uint32_t)ny);
        const int z = (int)(i / ((uint32_t)nx * (uint32_t)ny));
        if (x < 1 || x > nx - 2 || y < 1 || y > ny - 2 || z < 1 || z > nz - 2) continue;
        float g[7];
        arriving_g<JET, LOCAL, HALF>(p, x, y, z, g);
        double phi = 0.0;
        for (int q = 0; q < 7; ++q) phi += (double)g[q];
        v[0] += phi;
    }
    block_reduce<1, false>(v, partials + blockIdx.x);
}

// Final pass over per-block partials: sums, except value 5 of a 6-wide
// diagnostics record, which is a max.
__global__ __launch_bounds__(256) void k_reduce_partials(const double* partials, int nblocks, int nvals,
                                                          double* out) {
    for (int k = 0; k < nvals; ++k) {
        const bool is_max = (nvals == 6 && k == 5);
        double a = 0.0;
        for (int b = threadIdx.x; b < nblocks; b += 256) {
            const double x = partials[(size_t)b * nvals + k];
            a = is_max ? nan_max(a, x) : a + x;
        }
        double v[1] = {a};
        if (is_max)
            block_reduce<1, true>(v, out + k);
        else
            block_reduce<1, false>(v, out + k);
        __syncthreads();
    }
}

// Streaming kernels that measure the achievable HBM bandwidth (and
// calibrate memory-traffic counters against a known byte count): copy,
// read-only and write-only, four 16-byte accesses per thread, one pass over
// the array (a grid of n/1024 workgroups of 256 threads).  Measured on
// MI355X, 4 GiB: copy 5.52-5.57 TB/s, read-only 6.15-6.17, write-only
// 5.82; a grid-stride copy on 1024-65536 workgroups, 2-16 accesses per
// thread and non-temporal loads and stores all stayed at 4.6-5.54 TB/s.
__global__ __launch_bounds__(256) void k_copy(const float4* __restrict__ src, float4* __restrict__ dst, size_t n4) {
    const size_t stride = (size_t)gridDim.x * 1024u;
    size_t i = (size_t)blockIdx.x * 1024u + threadIdx.x;
    for (; i + 768 < n4; i += stride) {
        const float4 a = src[i], b = src[i + 256], c = src[i + 512], d = src[i + 768];
        dst[i] = a;
        dst[i + 256] = b;
        dst[i + 512] = c;
        dst[i + 768] = d;
    }
    for (; i < n4; i += 256) dst[i] = src[i];
}

// Reads every element; the sum is stored only if it equals a value a zeroed
// array cannot produce, so the loads stay live.
__global__ __launch_bounds__(256) void k_read_only(const float4* __restrict__ src, size_t n4, float* sink) {
    const size_t stride = (size_t)gridDim.x * 1024u;
    size_t i = (size_t)blockIdx.x * 1024u + threadIdx.x;
    float acc = 0.0f;
    for (; i + 768 < n4; i += stride) {
        const float4 a = src[i], b = src[i + 256], c = src[i + 512], d = src[i + 768];
        acc += ((a.x + a.y) + (a.z + a.w)) + ((b.x + b.y) + (b.z + b.w)) + ((c.x + c.y) + (c.z + c.w)) +
               ((d.x + d.y) + (d.z + d.w));
    }
    for (; i < n4; i += 256) acc += src[i].x;
    if (acc == 1234.5f) sink[0] = acc;
}

__global__ __launch_bounds__(256) void k_write_only(float4* __restrict__ dst, size_t n4) {
    const size_t stride = (size_t)gridDim.x * 1024u;
    size_t i = (size_t)blockIdx.x * 1024u + threadIdx.x;
    const float4 z = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    for (; i + 768 < n4; i += stride) {
        dst[i] = z;
        dst[i + 256] = z;
        dst[i + 512] = z;
        dst[i + 768] = z;
    }
    for (; i < n4; i += 256) dst[i] = z;
}

// ---------------------------------------------------------------------------
static dim3 stream_grid(size_t n4) {
    const size_t blocks = (n4 + 1023) / 1024;
    return dim3((unsigned)(blocks < 0x7fffffffu ? blocks : 0x7fffffffu));
}

hipError_t launch_copy(const float4* src, float4* dst, size_t n4, hipStream_t s) {
    hipLaunchKernelGGL(k_copy, stream_grid(n4), dim3(256), 0, s, src, dst, n4);
    return hipGetLastError();
}

hipError_t launch_read_only(const float4* src, size_t n4, float* sink, hipStream_t s) {
    hipLaunchKernelGGL(k_read_only, stream_grid(n4), dim3(256), 0, s, src, n4, sink);
    return hipGetLastError();
}

hipError_t launch_write_only(float4* dst, size_t n4, hipStream_t s) {
    hipLaunchKernelGGL(k_write_only, stream_grid(n4), dim3(256), 0, s, dst, n4);
    return hipGetLastError();
}

hipError_t launch_init_fields(const KParams& p, bool jet, hipStream_t s) {
    hipLaunchKernelGGL(k_init_fields, dim3((p.N + 255) / 256), dim3(256), 0, s, p, jet ? 1 : 0);
    return hipGetLastError();
}

template <int Q>
static hipError_t init_distros_q(const KParams& p, bool jet, bool half, hipStream_t s) {
    const dim3 grid((p.N + 255) / 256);
    if (half && p.ne) {
        if (jet) hipLaunchKernelGGL((k_init_distros<Q, true, true, false>), grid, dim3(256), 0, s, p);
        else hipLaunchKernelGGL((k_init_distros<Q, true, true, true>), grid, dim3(256), 0, s, p);
    } else if (half) {
        hipLaunchKernelGGL((k_init_distros<Q, true, false, false>), grid, dim3(256), 0, s, p);
    } else {
        hipLaunchKernelGGL((k_init_distros<Q, false, false, false>), grid, dim3(256), 0, s, p);
    }
    return hipGetLastError();
}

hipError_t launch_init_distros(const KParams& p, int Q, bool jet, bool half, hipStream_t s) {
    if (p.mom) {
        const dim3 grid((p.N + 255) / 256);
        if (Q == 27) hipLaunchKernelGGL((k_init_mom<27>), grid, dim3(256), 0, s, p);
        else hipLaunchKernelGGL((k_init_mom<19>), grid, dim3(256), 0, s, p);
        return hipGetLastError();
    }
    return Q == 27 ? init_distros_q<27>(p, jet, half, s) : init_distros_q<19>(p, jet, half, s);
}

template <int Q, bool JET>
static hipError_t gather_mom(const KParams& p, bool local, float* out, hipStream_t s) {
    const dim3 grid((p.N + 255) / 256);
    if (local) hipLaunchKernelGGL((k_gather_mom<Q, JET, true>), grid, dim3(256), 0, s, p, out);
    else hipLaunchKernelGGL((k_gather_mom<Q, JET, false>), grid, dim3(256), 0, s, p, out);
    return hipGetLastError();
}

template <int Q, bool JET, bool HALF, bool NE>
static hipError_t gather_n(const KParams& p, int parity, bool local, int which, float* out, hipStream_t s) {
    const dim3 grid((p.N + 255) / 256);
    if (local) hipLaunchKernelGGL((k_gather<Q, JET, HALF, 0, true, NE>), grid, dim3(256), 0, s, p, which, out);
    else if (parity) hipLaunchKernelGGL((k_gather<Q, JET, HALF, 1, false, NE>), grid, dim3(256), 0, s, p, which, out);
    else hipLaunchKernelGGL((k_gather<Q, JET, HALF, 0, false, NE>), grid, dim3(256), 0, s, p, which, out);
    return hipGetLastError();
}

template <int Q, bool JET, bool HALF>
static hipError_t gather_t(const KParams& p, int parity, bool local, int which, float* out, hipStream_t s) {
    if (HALF && p.ne) return gather_n<Q, JET, true, true>(p, parity, local, which, out, s);
    return gather_n<Q, JET, HALF, false>(p, parity, local, which, out, s);
}

template <int Q>
static hipError_t gather_q(const KParams& p, bool jet, bool half, int parity, bool local, int which, float* out,
                           hipStream_t s) {
    if (jet)
        return half ? gather_t<Q, true, true>(p, parity, local, which, out, s)
                    : gather_t<Q, true, false>(p, parity, local, which, out, s);
    return half ? gather_t<Q, false, true>(p, parity, local, which, out, s)
                : gather_t<Q, false, false>(p, parity, local, which, out, s);
}

hipError_t launch_gather(const KParams& p, int Q, bool jet, bool half, int parity, bool local, int which,
                         float* out, hipStream_t s) {
    if (p.mom && which == 0) {
        if (Q == 27) return jet ? gather_mom<27, true>(p, local, out, s) : gather_mom<27, false>(p, local, out, s);
        return jet ? gather_mom<19, true>(p, local, out, s) : gather_mom<19, false>(p, local, out, s);
    }
    return Q == 27 ? gather_q<27>(p, jet, half, parity, local, which, out, s)
                   : gather_q<19>(p, jet, half, parity, local, which, out, s);
}

// Running means of the derived-field set over every cell, ghosts included
// (timeAverage and reynoldsMomentsAverage, src/derivedFields/operators/
// timeAverage.cuh:34-88, reynoldsMoments.cuh:34-104): a += (v - a) / t.
// Each thread handles 4 consecutive cells.
__global__ __launch_bounds__(256) void k_average(KParams p, DerivedPtrs d, float t) {
    const uint32_t i0 = (blockIdx.x * 256u + threadIdx.x) * 4u;
#pragma unroll
    for (uint32_t k = 0; k < 4; ++k) {
        const uint32_t i = i0 + k;
        if (i >= p.N) return;
        const float phi = p.fld[10][i], ux = p.fld[1][i], uy = p.fld[2][i], uz = p.fld[3][i];
        const float v[10] = {phi, ux, uy, uz, ux * ux, uy * uy, uz * uz, ux * uy, ux * uz, uy * uz};
#pragma unroll
        for (int a = 0; a < 10; ++a) {
            float* f = d.avg[a];
            if (f) f[i] = f[i] + (v[a] - f[i]) / t;
        }
    }
}

// Vorticity by central differences at interior cells, 0 on the ghost shell
// (vorticityCompute, src/derivedFields/operators/vorticityFields.cuh:34-100).
__global__ __launch_bounds__(256) void k_vorticity(KParams p, DerivedPtrs d) {
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i >= p.N) return;
    const uint32_t nx = (uint32_t)p.nx, ny = (uint32_t)p.ny, nz = (uint32_t)p.nz;
    const uint32_t x = i % nx, y = (i / nx) % ny, z = i / (nx * ny);
    float wx = 0.0f, wy = 0.0f, wz = 0.0f;
    if (x > 0 && x < nx - 1 && y > 0 && y < ny - 1 && z > 0 && z < nz - 1) {
        const float* ux = p.fld[1];
        const float* uy = p.fld[2];
        const float* uz = p.fld[3];
        const uint32_t sy = nx, sz = nx * ny;
        const float duz_dy = 0.5f * (uz[i + sy] - uz[i - sy]);
        const float duy_dz = 0.5f * (uy[i + sz] - uy[i - sz]);
        const float dux_dz = 0.5f * (ux[i + sz] - ux[i - sz]);
        const float duz_dx = 0.5f * (uz[i + 1] - uz[i - 1]);
        const float duy_dx = 0.5f * (uy[i + 1] - uy[i - 1]);
        const float dux_dy = 0.5f * (ux[i + sy] - ux[i - sy]);
        wx = duz_dy - duy_dz;
        wy = dux_dz - duz_dx;
        wz = duy_dx - dux_dy;
    }
    if (d.vort[0]) d.vort[0][i] = wx;
    if (d.vort[1]) d.vort[1][i] = wy;
    if (d.vort[2]) d.vort[2][i] = wz;
    if (d.vort[3]) d.vort[3][i] = sqrtf(wx * wx + wy * wy + wz * wz);
}

hipError_t launch_average(const KParams& p, const DerivedPtrs& d, float t, hipStream_t s) {
    hipLaunchKernelGGL(k_average, dim3((p.N + 1023u) / 1024u), dim3(256), 0, s, p, d, t);
    return hipGetLastError();
}

hipError_t launch_vorticity(const KParams& p, const DerivedPtrs& d, hipStream_t s) {
    hipLaunchKernelGGL(k_vorticity, dim3((p.N + 255u) / 256u), dim3(256), 0, s, p, d);
    return hipGetLastError();
}

hipError_t launch_diagnostics(const KParams& p, double* partials, int nblocks, hipStream_t s) {
    hipLaunchKernelGGL(k_diagnostics, dim3(nblocks), dim3(256), 0, s, p, partials);
    return hipGetLastError();
}

template <bool JET, bool HALF>
static hipError_t mass_t(const KParams& p, bool local, double* partials, int nblocks, hipStream_t s) {
    if (local) hipLaunchKernelGGL((k_phase_mass<JET, true, HALF>), dim3(nblocks), dim3(256), 0, s, p, partials);
    else hipLaunchKernelGGL((k_phase_mass<JET, false, HALF>), dim3(nblocks), dim3(256), 0, s, p, partials);
    return hipGetLastError();
}

hipError_t launch_phase_mass(const KParams& p, bool jet, bool local, bool half, double* partials, int nblocks,
                             hipStream_t s) {
    if (jet) return half ? mass_t<true, true>(p, local, partials, nblocks, s) : mass_t<true, false>(p, local, partials, nblocks, s);
    return half ? mass_t<false, true>(p, local, partials, nblocks, s) : mass_t<false, false>(p, local, partials, nblocks, s);
}

hipError_t launch_reduce_partials(const double* partials, int nblocks, int nvals, double* out, hipStream_t s) {
    hipLaunchKernelGGL(k_reduce_partials, dim3(1), dim3(256), 0, s, partials, nblocks, nvals, out);
    return hipGetLastError();
}

}  // namespace pf
