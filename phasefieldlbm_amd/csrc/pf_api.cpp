// C-ABI of the phase-field LBM timestep library (declared in
// include/phasefieldlbm.h).  Owns every device buffer, the stream and the
// A-A step parity of one simulation; launches the kernels of
// step_kernels.hip and aux_kernels.hip.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "../../include/phasefieldlbm.h"
#include "kparams.hpp"

struct pf_ctx {
    pf_config cfg{};
    int dev = 0;
    hipStream_t stream = nullptr;
    size_t N = 0;
    size_t sN = 0;  // population slot stride (elements)
    int Q = 19;
    bool jet = true, half = false;
    void* f = nullptr;
    void* g = nullptr;
    float* fld[pf::kNumFields] = {};
    float* bcp = nullptr;
    double* partials = nullptr;
    double* dout = nullptr;
    int nred = 0;
    float* gather_buf = nullptr;
    uint16_t* mlp_w = nullptr;
    float* mlp_b = nullptr;
    bool has_mlp = false;
    pf::KParams kp{};
    int parity = 0;
    bool inited = false;
    size_t bytes = 0;
    std::string err;
};

namespace {

int fail(pf_ctx* c, int code, const std::string& msg) {
    if (c) c->err = msg;
    return code;
}

int hip_fail(pf_ctx* c, hipError_t e, const char* what) {
    return fail(c, PF_ERR_LAUNCH, std::string(what) + ": " + hipGetErrorString(e));
}

#define PF_HIP(ctx, call)                                        \
    do {                                                         \
        hipError_t e_ = (call);                                  \
        if (e_ != hipSuccess) return hip_fail((ctx), e_, #call); \
    } while (0)

size_t pop_bytes(const pf_ctx* c) { return c->half ? 2 : 4; }

// Round to nearest even float -> bfloat16.
uint16_t f2bf(float v) {
    uint32_t x;
    std::memcpy(&x, &v, 4);
    if ((x & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((x >> 16) | 0x40u);
    const uint32_t lsb = (x >> 16) & 1u;
    x += 0x7fffu + lsb;
    return (uint16_t)(x >> 16);
}

// Physical and relaxation constants, evaluated exactly as the reference's
// constexpr helpers (src/constants.cuh:92-141,
// src/functions/globalFunctions.cuh:104-462): FP32 arithmetic, with the
// double-precision quotients where the reference uses them.
void fill_constants(pf_ctx* c) {
    const pf_config& g = c->cfg;
    pf::KParams& k = c->kp;
    const float cs2g = 0.25f;
    const float diff_int = cs2g * (g.tau_g - 0.5f);
    const float kappa = 4.0f * diff_int / g.interface_width;
    k.gamma = kappa / cs2g;
    const float as2 = 3.0f;
    auto omega_from_nu = [&](float nu) { return (float)(1.0 / (0.5 + (double)as2 * (double)nu)); };
    auto tau_from_nu = [&](float nu) { return 0.5f + as2 * nu; };
    if (c->jet) {
        k.u_inf = g.u_inf;
        k.sigma = (g.u_inf * g.u_inf * (float)g.diam) / g.weber;
        k.visc_water = (float)(((double)g.u_inf * (double)g.diam) / (double)g.reynolds_water);
        k.visc_oil = (float)(((double)g.u_inf * (double)g.diam) / (double)g.reynolds_oil);
        k.tau_water = tau_from_nu(k.visc_water);
        k.tau_oil = tau_from_nu(k.visc_oil);
        k.omco_zmin = 1.0f - omega_from_nu(k.visc_oil);
        const int cells = g.nz / 12;
        const float sponge = (float)((double)cells / (double)(g.nz - 1));
        k.sp_z_start = (float)((double)(g.nz - 1 - cells) / (double)(g.nz - 1));
        k.sp_inv_nz_m1 = (float)(1.0 / (double)(g.nz - 1));
        k.sp_inv_sponge = (float)(1.0 / (double)sponge);
        const int radius = g.diam / 2;
        k.R2 = (float)(radius * radius);
        k.omega_ref = omega_from_nu(k.visc_water);
    } else {
        k.u_inf = 0.0f;
        k.sigma = g.sigma;
        const float visc_ref = (g.tau - 0.5f) / as2;
        k.omega_ref = omega_from_nu(visc_ref);
    }
    k.cxg = (float)(g.nx - 1) * 0.5f;
    k.cyg = (float)(g.ny - 1) * 0.5f;
    const float w7[7] = {0.25f, 0.125f, 0.125f, 0.125f, 0.125f, 0.125f, 0.125f};
    for (int q = 0; q < 7; ++q) k.wphi0[q] = c->half ? w7[q] * g.phi0 : 0.0f;
    k.noise = g.noise;
    k.noise_amp = g.noise_amp;
    k.seed = g.seed;
    k.mlp_lo = g.mlp_lo;
    k.mlp_hi = g.mlp_hi;
}

// z cells marched per workgroup.  Measured on MI355X (D3Q27 jet
// 256x256x1024): FP32 is fastest with ~16k workgroups (z-chunk 8), the
// binary16 form with ~4k (z-chunk 32); short chunks re-load more halo planes,
// long ones leave the last wave of workgroups unbalanced.
int choose_zchunk(const pf_config& g) {
    if (g.zchunk > 0) return g.zchunk;
    const int ntx = (g.nx + pf::kTX - 1) / pf::kTX, nty = (g.ny + pf::kTY - 1) / pf::kTY;
    const long cols = (long)ntx * nty;
    const long target = g.storage == PF_STORE_FP16 ? 4096 : 16384;
    long zc = ((long)(g.nz - 2) * cols + target / 2) / target;
    if (zc < 4) zc = 4;
    if (zc > 64) zc = 64;
    return (int)zc;
}

}  // namespace

extern "C" {

const char* pf_status_string(int s) {
    switch (s) {
        case PF_OK: return "ok";
        case PF_ERR_CONFIG: return "invalid configuration or argument";
        case PF_ERR_DEVICE: return "device unavailable";
        case PF_ERR_OOM: return "device allocation failed";
        case PF_ERR_LAUNCH: return "kernel launch or runtime error";
        case PF_ERR_IO: return "file could not be written";
        case PF_ERR_NONFINITE: return "non-finite value";
        case PF_ERR_STATE: return "call out of order";
        default: return "unknown status";
    }
}

const char* pf_last_error(const pf_ctx* c) { return c ? c->err.c_str() : "null context"; }

int pf_config_default(pf_config* cfg, int32_t flow_case, int32_t velocity_set) {
    if (!cfg) return PF_ERR_CONFIG;
    if (flow_case != PF_JET && flow_case != PF_DROPLET) return PF_ERR_CONFIG;
    if (velocity_set != PF_D3Q19 && velocity_set != PF_D3Q27) return PF_ERR_CONFIG;
    std::memset(cfg, 0, sizeof(*cfg));
    cfg->velocity_set = velocity_set;
    cfg->flow_case = flow_case;
    cfg->storage = PF_STORE_FP32;
    cfg->closure = PF_CLOSURE_FD;
    cfg->interface_width = 4.0f;
    cfg->tau_g = 1.0f;
    cfg->phi0 = 0.0f;
    cfg->device = -1;
    cfg->seed = 0x9E3779B9u;
    cfg->noise = PF_NOISE_GAUSSIAN_XY;
    cfg->noise_amp = 0.05f;
    cfg->mlp_lo = 0.05f;
    cfg->mlp_hi = 0.95f;
    cfg->slot_pad = -1;
    cfg->axis_x = 2.0f;  // src/initialConditions.cu:74 halves the x offset
    cfg->axis_y = 1.0f;
    cfg->axis_z = 1.0f;
    if (flow_case == PF_JET) {  // src/constants.cuh:82-109
        cfg->nx = 64; cfg->ny = 64; cfg->nz = 128;
        cfg->diam = 10;
        cfg->u_inf = 0.05f;
        cfg->reynolds_water = 5000.0f;
        cfg->reynolds_oil = 5000.0f;
        cfg->weber = 500.0f;
        cfg->tau = 0.0f;
        cfg->sigma = 0.0f;
    } else {  // src/constants.cuh:113-141
        cfg->nx = 128; cfg->ny = 128; cfg->nz = 128;
        cfg->radius = 35.0f;
        cfg->sigma = 0.03f;
        cfg->tau = 0.55f;
    }
    return PF_OK;
}

int pf_create(const pf_config* cfg, pf_ctx** out) {
    if (!cfg || !out) return PF_ERR_CONFIG;
    *out = nullptr;
    const pf_config& g = *cfg;
    if (g.velocity_set != PF_D3Q19 && g.velocity_set != PF_D3Q27) return PF_ERR_CONFIG;
    if (g.flow_case != PF_JET && g.flow_case != PF_DROPLET) return PF_ERR_CONFIG;
    if (g.storage != PF_STORE_FP32 && g.storage != PF_STORE_FP16) return PF_ERR_CONFIG;
    if (g.closure != PF_CLOSURE_FD && g.closure != PF_CLOSURE_MLP) return PF_ERR_CONFIG;
    if (g.nx < 4 || g.ny < 4 || g.nz < 4) return PF_ERR_CONFIG;
    if (g.flow_case == PF_JET && (g.nz < 12 || g.diam < 1 || !(g.weber > 0) || !(g.reynolds_water > 0) ||
                                  !(g.reynolds_oil > 0)))
        return PF_ERR_CONFIG;
    if (g.flow_case == PF_DROPLET && !(g.tau > 0.5f)) return PF_ERR_CONFIG;
    if (!(g.interface_width > 0) || g.axis_x == 0 || g.axis_y == 0 || g.axis_z == 0) return PF_ERR_CONFIG;
    if (g.noise < 0 || g.noise > 2) return PF_ERR_CONFIG;
    if (g.kernel != PF_KERNEL_PAIR && g.kernel != PF_KERNEL_LITERAL) return PF_ERR_CONFIG;
    const size_t N = (size_t)g.nx * g.ny * g.nz;
    if (N >= (size_t)1 << 30) return PF_ERR_CONFIG;  // byte offsets of one slot stay below 2^32

    pf_ctx* c = new (std::nothrow) pf_ctx();
    if (!c) return PF_ERR_OOM;
    c->cfg = g;
    c->N = N;
    c->Q = g.velocity_set;
    c->jet = g.flow_case == PF_JET;
    c->half = g.storage == PF_STORE_FP16;
    // Slot stride: N rounded up to 256 cells plus `slot_pad` cells (default
    // 1 KiB-worth of cells times an odd factor) so that successive slots of
    // the same cell start at different HBM channel offsets.
    const long pad = g.slot_pad >= 0 ? g.slot_pad : (c->half ? 1536 : 768);
    c->sN = ((N + 255) / 256) * 256 + (size_t)pad;
    if ((c->sN * 27 * (c->half ? 2 : 4)) >> 36) {  // keep every byte offset representable
        delete c;
        return PF_ERR_CONFIG;
    }

    // device selection (src/functions/hostFunctions.cuh:117-150)
    int dev = g.device;
    if (dev < 0) {
        dev = 0;
        if (const char* env = std::getenv("GPU_INDEX")) {
            char* end = nullptr;
            const long v = std::strtol(env, &end, 10);
            if (end != env && v >= 0) dev = (int)v;
        }
    }
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || dev >= ndev) {
        delete c;
        return PF_ERR_DEVICE;
    }
    if (hipSetDevice(dev) != hipSuccess) {
        delete c;
        return PF_ERR_DEVICE;
    }
    c->dev = dev;
    if (hipStreamCreate(&c->stream) != hipSuccess) {
        delete c;
        return PF_ERR_DEVICE;
    }

    auto alloc = [&](void** p, size_t bytes) -> bool {
        if (hipMalloc(p, bytes) != hipSuccess) {
            *p = nullptr;
            return false;
        }
        c->bytes += bytes;
        return true;
    };
    bool ok = alloc(&c->f, (size_t)c->Q * c->sN * pop_bytes(c)) && alloc(&c->g, (size_t)7 * c->sN * pop_bytes(c));
    for (int i = 0; ok && i < pf::kNumFields; ++i) ok = alloc((void**)&c->fld[i], N * sizeof(float));
    const size_t plane = (size_t)g.nx * g.ny;
    ok = ok && alloc((void**)&c->bcp, 2 * pf::kPlaneVals * plane * sizeof(float));
    c->nred = 1024;
    ok = ok && alloc((void**)&c->partials, (size_t)c->nred * 6 * sizeof(double));
    ok = ok && alloc((void**)&c->dout, 8 * sizeof(double));
    if (!ok) {
        pf_destroy(c);
        return PF_ERR_OOM;
    }

    pf::KParams& k = c->kp;
    k.f = c->f;
    k.g = c->g;
    for (int i = 0; i < pf::kNumFields; ++i) k.fld[i] = c->fld[i];
    k.bcp = c->bcp;
    k.nx = g.nx;
    k.ny = g.ny;
    k.nz = g.nz;
    k.N = (uint32_t)N;
    k.sN = (uint32_t)c->sN;
    k.zchunk = choose_zchunk(g);
    k.ntx = (g.nx + pf::kTX - 1) / pf::kTX;
    k.nty = (g.ny + pf::kTY - 1) / pf::kTY;
    k.ntz = (g.nz - 2 + k.zchunk - 1) / k.zchunk;
    fill_constants(c);
    *out = c;
    return PF_OK;
}

void pf_destroy(pf_ctx* c) {
    if (!c) return;
    (void)hipSetDevice(c->dev);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    (void)hipFree(c->f);
    (void)hipFree(c->g);
    for (auto* p : c->fld) (void)hipFree(p);
    (void)hipFree(c->bcp);
    (void)hipFree(c->partials);
    (void)hipFree(c->dout);
    (void)hipFree(c->gather_buf);
    (void)hipFree(c->mlp_w);
    (void)hipFree(c->mlp_b);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
}

size_t pf_device_bytes(const pf_ctx* c) { return c ? c->bytes : 0; }

int pf_init_distros(pf_ctx* c) {
    if (!c) return PF_ERR_CONFIG;
    PF_HIP(c, hipSetDevice(c->dev));
    PF_HIP(c, pf::launch_init_distros(c->kp, c->Q, c->half, c->stream));
    const size_t plane = (size_t)c->cfg.nx * c->cfg.ny;
    PF_HIP(c, hipMemsetAsync(c->bcp, 0, 2 * pf::kPlaneVals * plane * sizeof(float), c->stream));
    c->parity = 0;
    c->inited = true;
    return PF_OK;
}

int pf_init(pf_ctx* c) {
    if (!c) return PF_ERR_CONFIG;
    PF_HIP(c, hipSetDevice(c->dev));
    const pf_config& g = c->cfg;
    for (int i = 0; i < pf::kNumFields; ++i)
        PF_HIP(c, hipMemsetAsync(c->fld[i], 0, c->N * sizeof(float), c->stream));
    PF_HIP(c, hipMemsetAsync(c->f, 0, (size_t)c->Q * c->sN * pop_bytes(c), c->stream));
    PF_HIP(c, hipMemsetAsync(c->g, 0, (size_t)7 * c->sN * pop_bytes(c), c->stream));
    PF_HIP(c, pf::launch_init_fields(c->kp, c->jet, g.radius, g.axis_x, g.axis_y, g.axis_z, g.interface_width,
                                     c->stream));
    if (!c->jet) {
        // setDroplet (src/initialConditions.cu:58-83), evaluated on the host
        // with the C library's tanhf so the profile is the same everywhere.
        std::vector<float> phi(c->N, 0.0f);
        const float cx = (float)(g.nx - 1) * 0.5f, cy = (float)(g.ny - 1) * 0.5f, cz = (float)(g.nz - 1) * 0.5f;
        for (int z = 1; z < g.nz - 1; ++z)
            for (int y = 1; y < g.ny - 1; ++y)
                for (int x = 1; x < g.nx - 1; ++x) {
                    volatile float dx = ((float)x - cx) / g.axis_x;
                    volatile float dy = ((float)y - cy) / g.axis_y;
                    volatile float dz = ((float)z - cz) / g.axis_z;
                    volatile float dx2 = dx * dx, dy2 = dy * dy, dz2 = dz * dz;
                    volatile float s2 = dx2 + dy2;
                    s2 = s2 + dz2;
                    const float L2 = std::sqrt((float)s2);
                    const float arg = (float)(((double)g.radius - (double)L2) / (double)g.interface_width);
                    volatile float t = std::tanh(2.0f * arg);
                    volatile float h = 0.5f * t;
                    phi[(size_t)x + (size_t)g.nx * ((size_t)y + (size_t)g.ny * z)] = 0.5f + h;
                }
        PF_HIP(c, hipMemcpyAsync(c->fld[PF_FIELD_PHI], phi.data(), c->N * sizeof(float), hipMemcpyHostToDevice,
                                 c->stream));
        PF_HIP(c, hipStreamSynchronize(c->stream));
    }
    return pf_init_distros(c);
}

int pf_set_field(pf_ctx* c, int32_t id, const float* host, size_t n) {
    if (!c || !host) return PF_ERR_CONFIG;
    if (id < 0 || id >= PF_NUM_FIELDS) return fail(c, PF_ERR_CONFIG, "bad field id");
    if (n != c->N) return fail(c, PF_ERR_CONFIG, "size mismatch");
    PF_HIP(c, hipSetDevice(c->dev));
    PF_HIP(c, hipMemcpyAsync(c->fld[id], host, n * sizeof(float), hipMemcpyHostToDevice, c->stream));
    PF_HIP(c, hipStreamSynchronize(c->stream));
    return PF_OK;
}

static int run_steps(pf_ctx* c, uint64_t t0, uint32_t nsteps, uint32_t flags, hipEvent_t* ev) {
    for (uint32_t s = 0; s < nsteps; ++s) {
        c->kp.t = (uint32_t)(t0 + s);
        c->kp.save = ((flags & PF_STEP_SAVE_LAST) && s + 1 == nsteps) ? 1 : 0;
        const int variant = c->cfg.closure == PF_CLOSURE_MLP ? 2 : c->cfg.kernel;
        if (ev) PF_HIP(c, hipEventRecord(ev[0], c->stream));
        PF_HIP(c, pf::launch_phase_sum(c->kp, c->Q, c->jet, c->half, c->parity, c->stream));
        if (ev) PF_HIP(c, hipEventRecord(ev[1], c->stream));
        PF_HIP(c, pf::launch_step(c->kp, c->Q, c->jet, c->half, c->parity, variant, c->stream));
        if (ev) PF_HIP(c, hipEventRecord(ev[2], c->stream));
        if (c->jet) PF_HIP(c, pf::launch_jet_bc(c->kp, c->Q, c->half, c->parity ^ 1, c->stream));
        if (ev) PF_HIP(c, hipEventRecord(ev[3], c->stream));
        c->parity ^= 1;
        if (ev) ev += 4;
    }
    return PF_OK;
}

int pf_step(pf_ctx* c, uint64_t t0, uint32_t nsteps, uint32_t flags) {
    if (!c) return PF_ERR_CONFIG;
    if (!c->inited) return fail(c, PF_ERR_STATE, "pf_step before pf_init");
    if (c->cfg.closure == PF_CLOSURE_MLP && !c->has_mlp) return fail(c, PF_ERR_STATE, "MLP closure without weights");
    PF_HIP(c, hipSetDevice(c->dev));
    return run_steps(c, t0, nsteps, flags, nullptr);
}

int pf_sync(pf_ctx* c) {
    if (!c) return PF_ERR_CONFIG;
    PF_HIP(c, hipSetDevice(c->dev));
    PF_HIP(c, hipStreamSynchronize(c->stream));
    return PF_OK;
}

int pf_get_field(pf_ctx* c, int32_t id, float* host, size_t n) {
    if (!c || !host) return PF_ERR_CONFIG;
    if (id < 0 || id >= PF_NUM_FIELDS) return fail(c, PF_ERR_CONFIG, "bad field id");
    if (n != c->N) return fail(c, PF_ERR_CONFIG, "size mismatch");
    PF_HIP(c, hipSetDevice(c->dev));
    PF_HIP(c, hipMemcpyAsync(host, c->fld[id], n * sizeof(float), hipMemcpyDeviceToHost, c->stream));
    PF_HIP(c, hipStreamSynchronize(c->stream));
    return PF_OK;
}

int pf_get_populations(pf_ctx* c, int32_t which, float* host, size_t n) {
    if (!c || !host || (which != 0 && which != 1)) return PF_ERR_CONFIG;
    const size_t nq = which == 0 ? (size_t)c->Q : 7;
    if (n != nq * c->N) return fail(c, PF_ERR_CONFIG, "size mismatch");
    if (!c->inited) return fail(c, PF_ERR_STATE, "populations before pf_init");
    PF_HIP(c, hipSetDevice(c->dev));
    if (!c->gather_buf) {
        if (hipMalloc((void**)&c->gather_buf, (size_t)c->Q * c->N * sizeof(float)) != hipSuccess) {
            c->gather_buf = nullptr;
            return fail(c, PF_ERR_OOM, "gather buffer");
        }
        c->bytes += (size_t)c->Q * c->N * sizeof(float);
    }
    PF_HIP(c, pf::launch_gather(c->kp, c->Q, c->jet, c->half, c->parity, which, c->gather_buf, c->stream));
    PF_HIP(c, hipMemcpyAsync(host, c->gather_buf, n * sizeof(float), hipMemcpyDeviceToHost, c->stream));
    PF_HIP(c, hipStreamSynchronize(c->stream));
    return PF_OK;
}

int pf_diagnostics(pf_ctx* c, double out[6]) {
    if (!c || !out) return PF_ERR_CONFIG;
    PF_HIP(c, hipSetDevice(c->dev));
    PF_HIP(c, pf::launch_diagnostics(c->kp, c->partials, c->nred, c->stream));
    PF_HIP(c, pf::launch_reduce_partials(c->partials, c->nred, 6, c->dout, c->stream));
    PF_HIP(c, hipMemcpyAsync(out, c->dout, 6 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    PF_HIP(c, hipStreamSynchronize(c->stream));
    for (int i = 0; i < 6; ++i)
        if (!std::isfinite(out[i])) return fail(c, PF_ERR_NONFINITE, "non-finite diagnostics");
    return PF_OK;
}

int pf_phase_mass(pf_ctx* c, double* mass) {
    if (!c || !mass) return PF_ERR_CONFIG;
    if (!c->inited) return fail(c, PF_ERR_STATE, "phase mass before pf_init");
    PF_HIP(c, hipSetDevice(c->dev));
    PF_HIP(c, pf::launch_phase_mass(c->kp, c->jet, c->half, c->parity, c->partials, c->nred, c->stream));
    PF_HIP(c, pf::launch_reduce_partials(c->partials, c->nred, 1, c->dout, c->stream));
    PF_HIP(c, hipMemcpyAsync(mass, c->dout, sizeof(double), hipMemcpyDeviceToHost, c->stream));
    PF_HIP(c, hipStreamSynchronize(c->stream));
    if (!std::isfinite(*mass)) return fail(c, PF_ERR_NONFINITE, "non-finite phase mass");
    return PF_OK;
}

int pf_set_closure(pf_ctx* c, const float* W1, const float* b1, const float* W2, const float* b2, const float* W3,
                   const float* b3) {
    if (!c || !W1 || !b1 || !W2 || !b2 || !W3 || !b3) return PF_ERR_CONFIG;
    PF_HIP(c, hipSetDevice(c->dev));
    const int H = pf::kMlpHidden, I = pf::kMlpIn;
    std::vector<uint16_t> w((size_t)H * I + (size_t)H * H + H);
    size_t o = 0;
    for (int i = 0; i < H * I; ++i) w[o++] = f2bf(W1[i]);
    for (int i = 0; i < H * H; ++i) w[o++] = f2bf(W2[i]);
    for (int i = 0; i < H; ++i) w[o++] = f2bf(W3[i]);
    std::vector<float> b(2 * H + 1);
    for (int i = 0; i < H; ++i) b[i] = b1[i];
    for (int i = 0; i < H; ++i) b[H + i] = b2[i];
    b[2 * H] = b3[0];
    if (!c->mlp_w) {
        if (hipMalloc((void**)&c->mlp_w, w.size() * sizeof(uint16_t)) != hipSuccess ||
            hipMalloc((void**)&c->mlp_b, b.size() * sizeof(float)) != hipSuccess)
            return fail(c, PF_ERR_OOM, "closure weights");
        c->bytes += w.size() * sizeof(uint16_t) + b.size() * sizeof(float);
    }
    PF_HIP(c, hipMemcpy(c->mlp_w, w.data(), w.size() * sizeof(uint16_t), hipMemcpyHostToDevice));
    PF_HIP(c, hipMemcpy(c->mlp_b, b.data(), b.size() * sizeof(float), hipMemcpyHostToDevice));
    c->kp.mlp_w = c->mlp_w;
    c->kp.mlp_b = c->mlp_b;
    c->has_mlp = true;
    return PF_OK;
}

int pf_time_steps(pf_ctx* c, uint64_t t0, uint32_t nsteps, float out_ms[4]) {
    if (!c || !out_ms || nsteps == 0) return PF_ERR_CONFIG;
    if (!c->inited) return fail(c, PF_ERR_STATE, "timing before pf_init");
    if (c->cfg.closure == PF_CLOSURE_MLP && !c->has_mlp) return fail(c, PF_ERR_STATE, "MLP closure without weights");
    PF_HIP(c, hipSetDevice(c->dev));
    hipEvent_t a, b;
    PF_HIP(c, hipEventCreate(&a));
    PF_HIP(c, hipEventCreate(&b));
    PF_HIP(c, hipStreamSynchronize(c->stream));
    PF_HIP(c, hipEventRecord(a, c->stream));
    int rc = run_steps(c, t0, nsteps, 0, nullptr);
    if (rc) return rc;
    PF_HIP(c, hipEventRecord(b, c->stream));
    PF_HIP(c, hipEventSynchronize(b));
    PF_HIP(c, hipEventElapsedTime(&out_ms[0], a, b));
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    // second pass: events around every kernel of every step
    std::vector<hipEvent_t> ev((size_t)4 * nsteps);
    for (auto& e : ev) PF_HIP(c, hipEventCreate(&e));
    rc = run_steps(c, t0 + nsteps, nsteps, 0, ev.data());
    if (rc) return rc;
    PF_HIP(c, hipStreamSynchronize(c->stream));
    out_ms[1] = out_ms[2] = out_ms[3] = 0.0f;
    for (uint32_t s = 0; s < nsteps; ++s) {
        float t;
        for (int k = 0; k < 3; ++k) {
            PF_HIP(c, hipEventElapsedTime(&t, ev[4 * s + k], ev[4 * s + k + 1]));
            out_ms[1 + k] += t;
        }
    }
    for (auto& e : ev) (void)hipEventDestroy(e);
    return PF_OK;
}

int pf_copy_bandwidth(int32_t device, size_t bytes, uint32_t iters, double* gbs) {
    if (!gbs || bytes < 1024 || iters == 0) return PF_ERR_CONFIG;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || device < 0 || device >= ndev) return PF_ERR_DEVICE;
    if (hipSetDevice(device) != hipSuccess) return PF_ERR_DEVICE;
    const size_t n4 = bytes / 16;
    float4 *a = nullptr, *b = nullptr;
    if (hipMalloc((void**)&a, n4 * 16) != hipSuccess) return PF_ERR_OOM;
    if (hipMalloc((void**)&b, n4 * 16) != hipSuccess) {
        (void)hipFree(a);
        return PF_ERR_OOM;
    }
    hipStream_t s;
    hipEvent_t e0, e1;
    int rc = PF_OK;
    if (hipStreamCreate(&s) != hipSuccess) rc = PF_ERR_DEVICE;
    if (!rc && (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess)) rc = PF_ERR_DEVICE;
    if (!rc) {
        (void)hipMemsetAsync(a, 0, n4 * 16, s);
        (void)pf::launch_copy(a, b, n4, s);  // warm-up
        (void)hipEventRecord(e0, s);
        for (uint32_t i = 0; i < iters; ++i) (void)pf::launch_copy(a, b, n4, s);
        (void)hipEventRecord(e1, s);
        if (hipEventSynchronize(e1) != hipSuccess) rc = PF_ERR_LAUNCH;
        float ms = 0.0f;
        if (!rc && hipEventElapsedTime(&ms, e0, e1) != hipSuccess) rc = PF_ERR_LAUNCH;
        if (!rc) *gbs = 2.0 * (double)(n4 * 16) * iters / (ms * 1e-3) / 1e9;
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        (void)hipStreamDestroy(s);
    }
    (void)hipFree(a);
    (void)hipFree(b);
    return rc;
}

}  // extern "C"
