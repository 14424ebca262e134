// pf_mlp_predict_noalloc: network output for host samples without any
// allocation.  The features and outputs are staged in `scratch`, a
// file-scope device array of fixed size, instead of buffers allocated per
// call as pf_mlp_predict does.
//
// In host code `scratch` names a host-side stand-in, not device memory: its
// device address is resolved once per device with hipGetSymbolAddress
// (mlp_scratch_setup, called by pf_mlp_create with the device current) and
// only that address is passed to kernels and copies.  Each device holds its
// own copy of the array until the process ends.  Calls on one device share
// it: calls that overlap in time on the same device race.
#include <hip/hip_runtime.h>

#include <mutex>

#include "../../include/phasefieldlbm.h"
#include "kparams.hpp"
#include "pf_mlp.hpp"

namespace pf {
namespace {

constexpr size_t kMax = PF_MLP_PREDICT_MAX;
constexpr size_t kStride = 32;  // padded features per sample (27 used)
// layout (floats): raw features [kMax * 27] | padded features [kMax * 32] | outputs [kMax]
constexpr size_t kOffPadded = kMax * kMlpIn;
constexpr size_t kOffOut = kOffPadded + kMax * kStride;
constexpr size_t kScratchFloats = kOffOut + kMax;

static __device__ __align__(256) float scratch[kScratchFloats];

constexpr int kMaxDevices = 64;
std::mutex g_mu;
float* g_ptr[kMaxDevices] = {};

// padded[s][f] = raw[s][f] for f < 27, 0 for 27 <= f < 32 (every padded
// value of the first n samples is written, so no clearing is needed)
__global__ __launch_bounds__(256) void k_pad_features(const float* raw, float* padded, uint32_t n) {
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i >= n * (uint32_t)kStride) return;
    const uint32_t s = i / (uint32_t)kStride, f = i % (uint32_t)kStride;
    padded[i] = f < (uint32_t)kMlpIn ? raw[s * (uint32_t)kMlpIn + f] : 0.0f;
}

}  // namespace

int mlp_scratch_setup(int device) {
    if (device < 0 || device >= kMaxDevices) return PF_ERR_DEVICE;
    std::lock_guard<std::mutex> lock(g_mu);
    if (g_ptr[device]) return PF_OK;
    void* ptr = nullptr;
    if (hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)) != hipSuccess || !ptr) return PF_ERR_LAUNCH;
    g_ptr[device] = static_cast<float*>(ptr);
    return PF_OK;
}

}  // namespace pf

extern "C" int pf_mlp_predict_noalloc(pf_mlp* m, const float* x, size_t n, float* y) {
    if (!m || !x || !y || n == 0 || n > pf::kMax) return PF_ERR_CONFIG;
    if (m->dev < 0 || m->dev >= pf::kMaxDevices) return PF_ERR_DEVICE;
    float* base = nullptr;
    {
        std::lock_guard<std::mutex> lock(pf::g_mu);
        base = pf::g_ptr[m->dev];
    }
    if (!base) return PF_ERR_STATE;
    if (hipSetDevice(m->dev) != hipSuccess) return PF_ERR_DEVICE;
    float* raw = base;
    float* padded = base + pf::kOffPadded;
    float* out = base + pf::kOffOut;
    const uint32_t nn = (uint32_t)n;
    if (hipMemcpyAsync(raw, x, n * pf::kMlpIn * sizeof(float), hipMemcpyHostToDevice, m->stream) != hipSuccess)
        return PF_ERR_LAUNCH;
    hipLaunchKernelGGL(pf::k_pad_features, dim3((nn * (uint32_t)pf::kStride + 255u) / 256u), dim3(256), 0, m->stream,
                       raw, padded, nn);
    if (hipGetLastError() != hipSuccess) return PF_ERR_LAUNCH;
    if (pf::launch_mlp_forward(m->w16, m->bias, padded, n, out, m->stream) != hipSuccess ||
        hipMemcpyAsync(y, out, n * sizeof(float), hipMemcpyDeviceToHost, m->stream) != hipSuccess ||
        hipStreamSynchronize(m->stream) != hipSuccess)
        return PF_ERR_LAUNCH;
    return PF_OK;
}
