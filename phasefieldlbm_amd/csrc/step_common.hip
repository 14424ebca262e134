// Velocity-set independent launches: phase sum and dispatch to the per-Q units.
#include "step_kernels.hpp"

namespace pf {

hipError_t launch_step_q19(const KParams& p, bool jet, bool half, int parity, int variant, hipStream_t s);
hipError_t launch_step_q27(const KParams& p, bool jet, bool half, int parity, int variant, hipStream_t s);
hipError_t launch_jet_bc_q19(const KParams& p, bool half, int next_parity, hipStream_t s);
hipError_t launch_jet_bc_q27(const KParams& p, bool half, int next_parity, hipStream_t s);

template <bool HALF>
static hipError_t phase_sum_h(const KParams& p, bool jet, bool local, hipStream_t s) {
    if (jet) return local ? phase_sum_t<true, true, HALF>(p, s) : phase_sum_t<true, false, HALF>(p, s);
    return local ? phase_sum_t<false, true, HALF>(p, s) : phase_sum_t<false, false, HALF>(p, s);
}

hipError_t launch_phase_sum(const KParams& p, bool jet, bool local, bool half, hipStream_t s) {
    return half ? phase_sum_h<true>(p, jet, local, s) : phase_sum_h<false>(p, jet, local, s);
}

hipError_t launch_step(const KParams& p, int Q, bool jet, bool half, int parity, int variant, hipStream_t s) {
    return Q == 27 ? launch_step_q27(p, jet, half, parity, variant, s) : launch_step_q19(p, jet, half, parity, variant, s);
}

hipError_t launch_jet_bc(const KParams& p, int Q, bool half, int next_parity, hipStream_t s) {
    return Q == 27 ? launch_jet_bc_q27(p, half, next_parity, s) : launch_jet_bc_q19(p, half, next_parity, s);
}

}  // namespace pf
