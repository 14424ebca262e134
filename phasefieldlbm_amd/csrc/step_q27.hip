// D3Q27 instantiations of the timestep kernels (templates in step_kernels.hpp).
#include "step_kernels.hpp"

namespace pf {

hipError_t launch_step_q27(const KParams& p, bool jet, bool half, int parity, int variant, hipStream_t s) {
    return step_q<27>(p, jet, half, parity, variant, s);
}

hipError_t launch_jet_bc_q27(const KParams& p, bool half, int next_parity, hipStream_t s) {
    return half ? jet_bc_t<27, true>(p, next_parity, s) : jet_bc_t<27, false>(p, next_parity, s);
}

}  // namespace pf
