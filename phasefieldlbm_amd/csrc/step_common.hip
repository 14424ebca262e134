// Velocity-set independent launches: phase sum and dispatch to the per-Q units.
#include "step_kernels.hpp"
#include "phase_sum.hpp"

namespace pf {

hipError_t launch_step_q19(const KParams& p, bool jet, bool half, int parity, int variant, hipStream_t s);
hipError_t launch_step_q27(const KParams& p, bool jet, bool half, int parity, int variant, hipStream_t s);
hipError_t launch_jet_bc_q19(const KParams& p, bool half, int next_parity, hipStream_t s);
hipError_t launch_jet_bc_q27(const KParams& p, bool half, int next_parity, hipStream_t s);
hipError_t launch_step_mom_q19(const KParams& p, bool jet, bool local, hipStream_t s);
hipError_t launch_step_mom_q27(const KParams& p, bool jet, bool local, hipStream_t s);
hipError_t launch_jet_bc_mom_q19(const KParams& p, hipStream_t s);
hipError_t launch_jet_bc_mom_q27(const KParams& p, hipStream_t s);

template <bool HALF>
static hipError_t phase_sum_h(const KParams& p, bool jet, bool local, hipStream_t s) {
    if (jet) return local ? phase_sum_t<true, true, HALF>(p, s) : phase_sum_t<true, false, HALF>(p, s);
    return local ? phase_sum_t<false, true, HALF>(p, s) : phase_sum_t<false, false, HALF>(p, s);
}

// The row kernel (phase_sum.hpp) on FP32 grids wider than 128 cells, with
// 2-plane z-chunks.  Measured on MI355X (D3Q27 256x256x1024 jet, 50 steps
// per setting, one process): 0.302 ms against 0.326 ms for k_phase_sum;
// 0.308 / 0.312 / 0.333 ms with 4 / 8 / 16-plane chunks.  It was slower in
// binary16 (0.240 against 0.233 ms) and on 128-wide rows (D3Q19
// 128x128x512: 0.059 against 0.058 ms FP32, 0.042 against 0.038 ms
// binary16), which keep k_phase_sum.  Both give the same phi bit for bit.
template <bool JET>
static hipError_t phase_row_t(const KParams& p, hipStream_t s) {
    constexpr int TXC = 64, TR = kPrThreads / TXC, TW = 4 * TXC, ZC = 2;
    const int ntx = (p.nx + TW - 1) / TW, nty = (p.ny + TR - 1) / TR, band = (nty + 7) / 8;
    const int nzb = (p.nz - 2 + ZC - 1) / ZC;
    const dim3 grid((unsigned)(8 * ntx * band * nzb));
    hipLaunchKernelGGL((k_phase_row<JET, TXC>), grid, dim3(kPrThreads), 0, s, p, ZC);
    return hipGetLastError();
}

// Rows of a multiple of 4 cells, wider than 128, 16-byte aligned arrays.
static bool row_kernel_fits(const KParams& p, bool half) {
    if (half || p.nx <= 128 || p.nx % 4 != 0) return false;
    for (int i = 0; i < 3; ++i)
        if ((uintptr_t)p.h[i] % 16) return false;
    return (uintptr_t)p.fld[10] % 16 == 0 && (uintptr_t)p.phi_next % 16 == 0 && (uintptr_t)p.gbc % 16 == 0;
}

hipError_t launch_phase_sum(const KParams& p, bool jet, bool local, bool half, hipStream_t s) {
    if (!local && row_kernel_fits(p, half)) return jet ? phase_row_t<true>(p, s) : phase_row_t<false>(p, s);
    return half ? phase_sum_h<true>(p, jet, local, s) : phase_sum_h<false>(p, jet, local, s);
}

hipError_t launch_step(const KParams& p, int Q, bool jet, bool half, int parity, int variant, hipStream_t s) {
    return Q == 27 ? launch_step_q27(p, jet, half, parity, variant, s) : launch_step_q19(p, jet, half, parity, variant, s);
}

hipError_t launch_jet_bc(const KParams& p, int Q, bool half, int next_parity, hipStream_t s) {
    return Q == 27 ? launch_jet_bc_q27(p, half, next_parity, s) : launch_jet_bc_q19(p, half, next_parity, s);
}

hipError_t launch_step_mom(const KParams& p, int Q, bool jet, bool local, hipStream_t s) {
    return Q == 27 ? launch_step_mom_q27(p, jet, local, s) : launch_step_mom_q19(p, jet, local, s);
}

hipError_t launch_jet_bc_mom(const KParams& p, int Q, hipStream_t s) {
    return Q == 27 ? launch_jet_bc_mom_q27(p, s) : launch_jet_bc_mom_q19(p, s);
}

}  // namespace pf
