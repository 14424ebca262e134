// The learned closure's network object on one device (pf_mlp of the C-ABI),
// shared by the C-ABI layer and the allocation-free prediction entry point.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

struct pf_mlp {
    int dev = 0;
    hipStream_t stream = nullptr;
    float *master = nullptr, *m1 = nullptr, *m2 = nullptr, *bias = nullptr, *parts = nullptr, *gsum = nullptr;
    uint16_t* w16 = nullptr;
    float* xs = nullptr;
    float* ts = nullptr;
    int* idx = nullptr;
    uint8_t* sym = nullptr;
    double* loss = nullptr;
    size_t n = 0, cap_parts = 0;
    int step = 0;
};

namespace pf {
// Resolves, with `device` current, the device address of the library's fixed
// prediction staging buffer on that device (once per device).  PF_OK or a
// pf_status code.
int mlp_scratch_setup(int device);
}  // namespace pf
