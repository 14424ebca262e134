// Kernel parameter block and launch entry points shared by the HIP kernels
// and the C-ABI layer.
//
// HBM layout (one context):
//   f   : Q slots x N cells, FP32 or binary16, slot-major (SoA), the
//         reference's f[q*N + idx] (src/functions/deviceFunctions.cuh:33-48)
//         but addressed in place with the A-A pattern (see step kernels)
//   g   : 7 slots x N cells, same storage type as f
//   fld : 18 FP32 macroscopic fields [N] in host::FieldID order; phi is
//         live every step, the others are written on save steps only
//   bcp : jet boundary planes, 2 planes x 13 values x nx*ny (rho, u, P, F at
//         z = 1 and z = nz-2), written by the fused step for the boundary
//         kernel
// idx = x + nx*(y + ny*z), x fastest; the outer 1-cell shell is a ghost layer.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace pf {

constexpr int kTX = 64;  // tile width in x (one wavefront)
constexpr int kTY = 4;   // tile rows in y (4 wavefronts per workgroup)
constexpr int kThreads = kTX * kTY;

constexpr int kNumFields = 18;
constexpr int kPlaneVals = 13;  // rho, ux, uy, uz, pxx, pyy, pzz, pxy, pxz, pyz, fsx, fsy, fsz

// MLP closure dimensions (27 -> 64 -> 64 -> 1).
constexpr int kMlpIn = 27;
constexpr int kMlpHidden = 64;

struct KParams {
    void* f;
    void* g;
    float* fld[kNumFields];
    float* bcp;
    int nx, ny, nz;
    uint32_t N;
    int zchunk;
    int ntx, nty, ntz;  // tile counts
    int save;           // write every macroscopic field this step
    // physics
    float sigma, gamma;
    float u_inf;
    float tau_water, tau_oil, visc_water, visc_oil, omco_zmin;
    float omega_ref;
    float sp_z_start, sp_inv_nz_m1, sp_inv_sponge;
    float cxg, cyg, R2;  // nozzle centre and radius^2
    float wphi0[7];      // FP16 storage shift of g
    // inflow noise
    int noise;
    float noise_amp;
    uint32_t seed;
    uint32_t t;  // time index of this step (inflow)
    // learned closure
    const uint16_t* mlp_w;  // bf16: W1[64x27] | W2[64x64] | W3[64]
    const float* mlp_b;     // fp32: b1[64] | b2[64] | b3[1]
    float mlp_lo, mlp_hi;
};

// Launchers; all return hipError_t of the launch.
hipError_t launch_phase_sum(const KParams& p, int Q, bool jet, bool half, int parity, hipStream_t s);
hipError_t launch_step(const KParams& p, int Q, bool jet, bool half, int parity, bool mlp, hipStream_t s);
hipError_t launch_jet_bc(const KParams& p, int Q, bool half, int next_parity, hipStream_t s);
hipError_t launch_init_fields(const KParams& p, bool jet, float radius, float ax, float ay, float az, float W,
                              hipStream_t s);
hipError_t launch_init_distros(const KParams& p, int Q, bool half, hipStream_t s);
hipError_t launch_gather(const KParams& p, int Q, bool jet, bool half, int parity, int which, float* out,
                         hipStream_t s);
hipError_t launch_diagnostics(const KParams& p, double* partials, int nblocks, hipStream_t s);
hipError_t launch_phase_mass(const KParams& p, bool jet, bool half, int parity, double* partials, int nblocks,
                             hipStream_t s);
hipError_t launch_reduce_partials(const double* partials, int nblocks, int nvals, double* out, hipStream_t s);

}  // namespace pf
