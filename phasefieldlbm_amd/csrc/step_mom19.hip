// D3Q19 instantiations of the moment-form timestep (templates in step_mom.hpp).
#include "step_mom.hpp"

namespace pf {

hipError_t launch_step_mom_q19(const KParams& p, bool jet, bool local, hipStream_t s) {
    return step_mom_q<19>(p, jet, local, s);
}

hipError_t launch_jet_bc_mom_q19(const KParams& p, hipStream_t s) { return jet_bc_mom_q<19>(p, s); }

}  // namespace pf
