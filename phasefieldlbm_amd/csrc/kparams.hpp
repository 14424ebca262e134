// Kernel parameter block and launch entry points shared by the HIP kernels
// and the C-ABI layer.
//
// HBM layout (one context):
//   f   : Q population slots of every cell, FP32 or binary16, interleaved
//         per (y, z) row: element (slot, x, y, z) sits at
//         z*plane_f + y*pitch_f + slot*nx + x with pitch_f >= Q*nx and
//         plane_f = pitch_f*ny.  One row holds the Q populations of its nx
//         cells as Q contiguous runs, so the populations a workgroup touches
//         together are one contiguous block per row instead of Q streams a
//         whole grid apart (the reference's slot-major f[q*N + idx],
//         src/functions/deviceFunctions.cuh:33-48, made timings depend on
//         where the driver placed the physical pages).  Addressed in place
//         with the A-A pattern (see the step kernels).
//   h   : 3 arrays [N] (FP32, binary16 in the binary16 form), the phase
//         flux of every cell,
//         h = (4 phi u + gamma phi (1 - phi) n) / 8.  The D3Q7 phase
//         populations have no non-equilibrium part (D3Q7.cuh:139-143), so the
//         post-collision g of a cell is fixed by its phi and h:
//         g_0 = phi/4, g_(+-k) = phi/8 +- h_k.  The phase kernel rebuilds the
//         seven g arriving at each cell from its neighbours' phi and h instead
//         of storing and streaming the g (12 B written per cell instead of
//         7 populations).
//   fld : 18 FP32 macroscopic fields [N] in host::FieldID order; phi is
//         live every step, the others are written on save steps only
//   phi_next : the second phi buffer: the phase kernel reads phi of the
//         previous step from fld[10] and writes the new phi here; the host
//         then swaps the two
//   bcp : jet boundary planes, 2 planes x 13 values x nx*ny (rho, u, P, F at
//         z = 1 and z = nz-2), written by the fused step for the boundary
//         kernel
//   mom : binary16 non-equilibrium form only, 2 x N x 4 binary16 (ne.hpp)
//   rec : moment form only, 2 x NC x N FP32 (NC = 10 / 17 for D3Q19 / D3Q27),
//         the coefficients of each cell's collision output (mom.hpp), in
//         place of f
//   gbc : jet boundary phase populations, 2 planes x nx*ny: the g_5 entering
//         z = 1 and the g_6 entering z = nz-2 (src/BoundaryConditions.cuh:104,
//         168)
// idx = x + nx*(y + ny*z), x fastest; the outer 1-cell shell is a ghost layer.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace pf {

// Workgroup tile of the timestep kernels: kTX cells in x (a multiple of the
// 64-lane wavefront, a power of two) by kTY rows in y.  128 x 4 measured best
// on MI355X for the D3Q27 jet: a wider x tile halves the cache lines that two
// tiles share in the in-place odd step (each x-shifted population of a tile
// row touches one line of the neighbouring tile).
#ifndef PF_TILE_X
#define PF_TILE_X 128
#endif
#ifndef PF_TILE_Y
#define PF_TILE_Y 4
#endif
constexpr int kTX = PF_TILE_X;
constexpr int kTY = PF_TILE_Y;
constexpr int kThreads = kTX * kTY;
static_assert((kTX & (kTX - 1)) == 0 && kTX >= 64, "tile width must be a power of two >= 64");
static_assert(kThreads <= 1024, "workgroup too large");
// Narrow tile of the timestep kernel, with the same thread count
// (64 x 8 for the default), for grids with at most 64 cells in x: a 128-wide
// tile leaves half its lanes idle there.  Measured on MI355X: 64^3 droplet
// 14.7k against 12.5k MLUPS at the best z-chunk of each; the reference's
// 64x64x128 jet 14.0k against 12.8k (D3Q19) and 10.7k against 9.9k (D3Q27).
constexpr int kTXNarrow = 64;
static_assert(kThreads % kTXNarrow == 0, "narrow tile");
constexpr int tile_width(int nx) { return nx <= kTXNarrow ? kTXNarrow : kTX; }
// The moment form's step tiles are tile_x x kMomRows cells (step_mom.hpp).
#ifndef PF_MOM_ROWS
#define PF_MOM_ROWS 4
#endif
constexpr int kMomRows = PF_MOM_ROWS;

// Scale K of the binary16 non-equilibrium form (ne.hpp).
constexpr float kNeK = 4096.0f;

constexpr int kNumFields = 18;
constexpr int kPlaneVals = 13;  // rho, ux, uy, uz, pxx, pyy, pzz, pxy, pxz, pyz, fsx, fsy, fsz

// MLP closure dimensions (27 -> 64 -> 64 -> 1).
constexpr int kMlpIn = 27;
constexpr int kMlpHidden = 64;

struct KParams {
    void* f;
    void* h[3];  // float, or _Float16 in the binary16 form
    float* fld[kNumFields];
    float* phi_next;
    float* bcp;
    float* gbc;
    int nx, ny, nz;
    uint32_t N;  // cells
    uint32_t pitch_f;  // population row pitch in elements (y stride); slot stride is nx
    uint64_t plane_f;  // population plane pitch in elements (z stride) = pitch * ny
    uint64_t elems_f;  // population array size in elements = plane * nz
    int zchunk;
    int tile_x;         // k_step tile width: kTX, or kTXNarrow (tile_width)
    int ntx, nty, ntz;  // tile counts
    int save;           // write every macroscopic field this step
    // physics
    float sigma, gamma;
    float u_inf;
    float tau_water, tau_oil, visc_water, visc_oil, omco_zmin;
    float omega_ref;
    float sp_z_start, sp_inv_nz_m1, sp_inv_sponge;
    float cxg, cyg, R2;  // nozzle centre and radius^2
    // inflow noise
    int noise;
    float noise_amp;
    uint32_t seed;
    uint32_t t;  // time index of this step (inflow)
    // learned closure
    const uint16_t* mlp_w;  // bf16: W1[64x32] (27 used) | W2[64x64] | w3[64]
    const float* mlp_b;     // fp32: b1[64] | b2[64] | b3[1]
    float mlp_lo, mlp_hi;   // phi window of the closure
    float mlp_kmin, mlp_kmax;  // FD-curvature window of the closure
    float* kappa;              // closure curvature at interface cells [N]
    // binary16 non-equilibrium form (ne != 0; f and h in binary16): a
    // population is stored as ne_k (f~ - S_q(M)), S_q the shift of the cell
    // that wrote it, and M = (A, Bx, By, Bz) of every cell is kept as four
    // binary16 (ne.hpp).  mom_prev holds the M the populations now stored
    // were written with, mom_cur receives this step's; the host swaps them.
    int ne;
    float ne_k, ne_kinv;
    const uint2* mom_prev;
    uint2* mom_cur;
    // moment form (mom != 0, mom.hpp): the FP32 coefficients of every
    // cell's collision output, interleaved per (y, z) row (rec_elem); rec_prev holds those the
    // arriving populations are rebuilt from, rec_cur receives this step's
    // (the host swaps them).  bcf: the jet's boundary populations for the
    // next step, 18 x nx*ny: the 9 c_z = +1 directions entering z = 1, then
    // the 9 c_z = -1 entering z = nz-2, slot (c_y + 1) * 3 + (c_x + 1), by
    // target column.
    int mom;
    const float* rec_prev;
    float* rec_cur;
    float* bcf;
};

// Launchers; all return hipError_t of the launch.
// local: the first step after the populations were set, when each cell's
// own equilibrium g (not its neighbours' post-collision g) is summed
hipError_t launch_phase_sum(const KParams& p, bool jet, bool local, bool half, hipStream_t s);
// variant: 0 pair-symmetric collision, 1 literal per-direction evaluation,
// 2 pair-symmetric with the learned curvature closure
hipError_t launch_step(const KParams& p, int Q, bool jet, bool half, int parity, int variant, hipStream_t s);
hipError_t launch_jet_bc(const KParams& p, int Q, bool half, int next_parity, hipStream_t s);
// moment form: the step (local: the first step after initialisation, which
// reads every cell's own coefficients) and its boundary populations
hipError_t launch_step_mom(const KParams& p, int Q, bool jet, bool local, hipStream_t s);
hipError_t launch_jet_bc_mom(const KParams& p, int Q, hipStream_t s);
// rho = 1, the jet's nozzle plane (phi, u_z); the droplet profile is
// uploaded by the host
hipError_t launch_init_fields(const KParams& p, bool jet, hipStream_t s);
hipError_t launch_init_distros(const KParams& p, int Q, bool jet, bool half, hipStream_t s);
hipError_t launch_gather(const KParams& p, int Q, bool jet, bool half, int parity, bool local, int which,
                         float* out, hipStream_t s);
hipError_t launch_diagnostics(const KParams& p, double* partials, int nblocks, hipStream_t s);
hipError_t launch_phase_mass(const KParams& p, bool jet, bool local, bool half, double* partials, int nblocks,
                             hipStream_t s);
hipError_t launch_reduce_partials(const double* partials, int nblocks, int nvals, double* out, hipStream_t s);
hipError_t launch_copy(const float4* src, float4* dst, size_t n4, hipStream_t s);
hipError_t launch_read_only(const float4* src, size_t n4, float* sink, hipStream_t s);
hipError_t launch_write_only(float4* dst, size_t n4, hipStream_t s);

// Derived fields (src/derivedFields/): avg[0..3] running means of phi, ux,
// uy, uz; avg[4..9] of ux*ux, uy*uy, uz*uz, ux*uy, ux*uz, uy*uz; vort[0..3]
// curl u and its magnitude.  Null pointers are skipped.
struct DerivedPtrs {
    float* avg[10];
    float* vort[4];
};
hipError_t launch_average(const KParams& p, const DerivedPtrs& d, float t, hipStream_t s);
hipError_t launch_vorticity(const KParams& p, const DerivedPtrs& d, hipStream_t s);
hipError_t launch_mlp_grad(const uint16_t* w16, const float* bias, const float* xs, const float* ts, const int* idx,
                           const uint8_t* symperm, int nsym, uint32_t sym_seed, size_t n, size_t s_begin,
                           size_t count, float scale, float* grad_out, hipStream_t s);
hipError_t launch_mlp_adam(const float* parts, int nparts, float* master, float* m1, float* m2, uint16_t* w16,
                           float* bias, float lr, int step, int apply, float* grad_sum, double* loss_acc,
                           hipStream_t s);
hipError_t launch_mlp_forward(const uint16_t* w16, const float* bias, const float* xs, size_t n, float* y,
                              hipStream_t s);
hipError_t launch_mfma_layout(const uint16_t* a, const uint16_t* b, float* d, hipStream_t s);
hipError_t launch_closure(const KParams& p, hipStream_t s);

}  // namespace pf
