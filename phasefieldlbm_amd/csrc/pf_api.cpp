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
#include <random>
#include <algorithm>
#include <numeric>
#include <string>
#include <vector>

#include "../../include/phasefieldlbm.h"
#include "kparams.hpp"
#include "pf_io.hpp"
#include "pf_mlp.hpp"

struct pf_ctx {
    pf_config cfg{};
    int dev = 0;
    hipStream_t stream = nullptr;
    size_t N = 0;
    size_t f_elems = 0;  // population array size (elements, row-interleaved layout)
    int Q = 19;
    bool jet = true, half = false, ne = false, moments = false;
    void* f = nullptr;  // populations (not allocated in the moment form)
    float* rec[2] = {nullptr, nullptr};  // moment form: coefficients, previous / current step
    float* bcf = nullptr;                // moment form: jet boundary populations
    uint2* mom[2] = {nullptr, nullptr};  // non-equilibrium form: M of every cell, previous / current step
    void* h = nullptr;  // phase flux, 3 x N (FP32, or binary16 in the binary16 form)
    float* fld[pf::kNumFields] = {};
    float* phi_next = nullptr;  // second phi buffer, swapped with fld[PF_FIELD_PHI] every step
    float* bcp = nullptr;
    float* gbc = nullptr;  // jet inflow g_5 / outflow g_6 planes
    double* partials = nullptr;
    double* dout = nullptr;
    int nred = 0;
    float* gather_buf = nullptr;
    float* kappa = nullptr;
    float* der[PF_NUM_FIELDS - PF_NUM_BASE_FIELDS] = {};  // derived fields, by id - PF_NUM_BASE_FIELDS
    uint16_t* mlp_w = nullptr;
    float* mlp_b = nullptr;
    bool has_mlp = false;
    pf::KParams kp{};
    int parity = 0;
    bool fresh = false;  // populations just set: the next phase sum is the local one
    bool inited = false;
    // PF_STEP_GRAPH: one captured graph per A-A parity (phase sum, closure,
    // step), valid while the kernel parameters it was captured with hold.
    hipGraphExec_t gexec[2] = {nullptr, nullptr};
    pf::KParams gkey[2]{};
    int gvariant[2] = {-1, -1};
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

// Device copy of the closure parameters: BF16 W1 padded to [64][32], W2
// [64][64], w3 [64]; FP32 b1, b2, b3.
void pack_mlp(const float* params, std::vector<uint16_t>& w, std::vector<float>& b) {
    w.assign(64 * 32 + 64 * 64 + 64, 0);
    b.assign(2 * 64 + 1, 0.0f);
    const float* W1 = params;
    const float* B1 = W1 + 64 * 27;
    const float* W2 = B1 + 64;
    const float* B2 = W2 + 64 * 64;
    const float* W3 = B2 + 64;
    const float* B3 = W3 + 64;
    for (int o = 0; o < 64; ++o)
        for (int i = 0; i < 27; ++i) w[o * 32 + i] = f2bf(W1[o * 27 + i]);
    for (int i = 0; i < 64 * 64; ++i) w[64 * 32 + i] = f2bf(W2[i]);
    for (int i = 0; i < 64; ++i) w[64 * 32 + 64 * 64 + i] = f2bf(W3[i]);
    for (int i = 0; i < 64; ++i) {
        b[i] = B1[i];
        b[64 + i] = B2[i];
    }
    b[128] = B3[0];
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
    k.noise = g.noise;
    k.noise_amp = g.noise_amp;
    k.seed = g.seed;
    k.mlp_lo = g.mlp_lo;
    k.mlp_hi = g.mlp_hi;
    k.mlp_kmin = g.mlp_kmin;
    k.mlp_kmax = g.mlp_kmax;
}

// z cells marched per workgroup.  Measured on MI355X (D3Q27 jet
// 256x256x1024, tile rows dealt to the XCDs, fresh process per setting, two
// runs each), with each thread's phi-tile loads issued together: FP32 k_step
// 2.95-2.97 ms at z-chunk 2 (65k workgroups) against 3.03 at 4, 3.05 at 3
// and 3.12-3.13 at 1; the binary16 form (3 workgroups per CU) within 0.3 %
// from z-chunk 4 to 16 (~8k workgroups at 16).  Before those loads were
// issued together, z-chunk 4 was the FP32 optimum (3.05 ms, 3.12 at 2).
// The 128x128x512 D3Q19 jet takes 0.294 ms at z-chunk 2 or 4.  Short chunks re-load
// more halo planes, long ones leave the last wave of workgroups unbalanced.
// Small grids get the longest chunk, down to 1, that still gives every one
// of the 256 CUs a workgroup.  Measured on 64 x 8 tiles:
// 64^3 droplet 14.7k MLUPS at z-chunk 1 (496 workgroups), 13.8k at 2 (248),
// 10.9k at 4 (128); the reference's 64x64x128 D3Q19 jet 14.2k at 4 (256),
// 14.0k at 2 (504), 12.9k at 1.
// The moment form takes 64 x 4 tiles unless told otherwise (step_mom.hpp).
int tile_x(const pf_config& g) {
    if (g.tile_x > 0) return g.tile_x;
    return g.storage == PF_STORE_MOM ? 64 : pf::tile_width(g.nx);
}
int tile_rows(const pf_config& g, int tw) { return g.storage == PF_STORE_MOM ? pf::kMomRows : pf::kThreads / tw; }

int choose_zchunk(const pf_config& g) {
    if (g.zchunk > 0) return g.zchunk;
    const int tw = tile_x(g), th = tile_rows(g, tw);
    const long cols = (long)((g.nx + tw - 1) / tw) * ((g.ny + th - 1) / th);
    long zc;
    if (g.storage == PF_STORE_MOM) {
        // The moment form stages two sending planes more than it updates per
        // chunk: 16 planes where the grid fills the GPU with them.  Its step
        // took 2.37 ms at 16 planes and 2.40-2.42 at 32 on the 256x256x1024
        // D3Q27 jet, and 0.220 ms at 8 or 16 planes against 0.232 at 4 on
        // the 128x128x512 D3Q19 jet.
        zc = 16;
    } else {
        // binary16 marches long chunks
        const bool fp16 = g.storage == PF_STORE_FP16;
        const long target = fp16 ? 8192 : 32768;
        zc = ((long)(g.nz - 2) * cols + target / 2) / target;
        if (zc < 4) zc = 4;
        if (zc > 64) zc = 64;
        // FP32 grids that reach the target with 4-plane chunks take half the
        // chunk, at least 2 planes (see the measurements above).
        if (!fp16 && ((long)(g.nz - 2) + 3) / 4 * cols >= target) zc = std::max(2L, zc / 2);
    }
    constexpr long kCUs = 256;
    while (zc > 1 && ((long)(g.nz - 2) + zc - 1) / zc * cols < kCUs) --zc;
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
    cfg->reserved0 = 0.0f;
    cfg->device = -1;
    cfg->seed = 0x9E3779B9u;
    cfg->noise = PF_NOISE_GAUSSIAN_XY;
    cfg->noise_amp = 0.05f;
    // the learned closure over the whole diffuse interface, on convex parts
    // with radii of curvature from 1 to ~67 cells
    cfg->mlp_lo = 0.001f;
    cfg->mlp_hi = 0.999f;
    cfg->slot_pad = -1;
    cfg->mlp_kmin = -2.0f;
    cfg->mlp_kmax = -0.03f;
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
    if (g.storage != PF_STORE_FP32 && g.storage != PF_STORE_FP16 && g.storage != PF_STORE_FP16_NE &&
        g.storage != PF_STORE_MOM)
        return PF_ERR_CONFIG;
    // the moment form: pair collision and FD curvature, and one plane of
    // coefficients addressable by a 32-bit buffer range (step_mom.hpp)
    if (g.storage == PF_STORE_MOM && (g.kernel == PF_KERNEL_LITERAL || g.closure == PF_CLOSURE_MLP ||
                                      (uint64_t)g.nx * (uint64_t)g.ny * 17u * 4u >= ((uint64_t)1 << 32)))
        return PF_ERR_CONFIG;
    if (g.storage == PF_STORE_FP16_NE) {
        // pair collision only; the nozzle must lie inside the lateral ghost
        // shell, where the boundary kernel's source columns need no wrap
        if (g.kernel == PF_KERNEL_LITERAL) return PF_ERR_CONFIG;
        if (g.flow_case == PF_JET && ((g.nx - 1) * 0.5 - g.diam / 2 < 1.0 || (g.ny - 1) * 0.5 - g.diam / 2 < 1.0))
            return PF_ERR_CONFIG;
    }
    if (g.closure != PF_CLOSURE_FD && g.closure != PF_CLOSURE_MLP) return PF_ERR_CONFIG;
    if (g.nx < 4 || g.ny < 4 || g.nz < 4) return PF_ERR_CONFIG;
    if (g.flow_case == PF_JET && (g.nz < 12 || g.diam < 1 || !(g.weber > 0) || !(g.reynolds_water > 0) ||
                                  !(g.reynolds_oil > 0)))
        return PF_ERR_CONFIG;
    if (g.flow_case == PF_DROPLET && !(g.tau > 0.5f)) return PF_ERR_CONFIG;
    if (!(g.interface_width > 0) || g.axis_x == 0 || g.axis_y == 0 || g.axis_z == 0) return PF_ERR_CONFIG;
    if (g.noise < 0 || g.noise > 2) return PF_ERR_CONFIG;
    if (g.kernel != PF_KERNEL_PAIR && g.kernel != PF_KERNEL_LITERAL) return PF_ERR_CONFIG;
    if (g.derived & ~(PF_DERIVED_AVERAGE | PF_DERIVED_REYNOLDS | PF_DERIVED_VORTICITY)) return PF_ERR_CONFIG;
    if (g.tile_x != 0 && g.tile_x != pf::kTX && g.tile_x != pf::kTXNarrow) return PF_ERR_CONFIG;
    if (g.reserved0 != 0.0f) return PF_ERR_CONFIG;
    const size_t N = (size_t)g.nx * g.ny * g.nz;
    if (N >= (size_t)1 << 30) return PF_ERR_CONFIG;  // byte offsets of one slot stay below 2^32

    pf_ctx* c = new (std::nothrow) pf_ctx();
    if (!c) return PF_ERR_OOM;
    c->cfg = g;
    c->N = N;
    c->Q = g.velocity_set;
    c->jet = g.flow_case == PF_JET;
    c->half = g.storage == PF_STORE_FP16 || g.storage == PF_STORE_FP16_NE;
    c->ne = g.storage == PF_STORE_FP16_NE;
    c->moments = g.storage == PF_STORE_MOM;
    // Row-interleaved population layout: row pitch S*nx plus `slot_pad`
    // elements (default none), plane pitch = row pitch * ny.  In-plane byte
    // offsets are 32-bit in the kernels.
    const uint64_t pad = g.slot_pad >= 0 ? (uint64_t)g.slot_pad : 0;
    const uint64_t pitch_f = (uint64_t)c->Q * g.nx + pad;
    const uint64_t esize = c->half ? 2 : 4;
    if (pitch_f * g.ny * esize >= ((uint64_t)1 << 32)) {
        delete c;
        return PF_ERR_CONFIG;
    }
    c->f_elems = (size_t)(pitch_f * g.ny * g.nz);

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
    bool ok = (c->moments || alloc(&c->f, c->f_elems * pop_bytes(c))) && alloc(&c->h, 3 * N * pop_bytes(c));
    if (c->moments) {
        const size_t nc = c->Q == 27 ? 17 : 10;
        ok = ok && alloc((void**)&c->rec[0], nc * N * sizeof(float)) && alloc((void**)&c->rec[1], nc * N * sizeof(float));
        ok = ok && alloc((void**)&c->bcf, 18 * (size_t)g.nx * g.ny * sizeof(float));
    }
    for (int i = 0; ok && i < pf::kNumFields; ++i) ok = alloc((void**)&c->fld[i], N * sizeof(float));
    ok = ok && alloc((void**)&c->phi_next, N * sizeof(float));
    const size_t plane = (size_t)g.nx * g.ny;
    ok = ok && alloc((void**)&c->bcp, 2 * pf::kPlaneVals * plane * sizeof(float));
    ok = ok && alloc((void**)&c->gbc, 2 * plane * sizeof(float));
    c->nred = 1024;
    ok = ok && alloc((void**)&c->partials, (size_t)c->nred * 6 * sizeof(double));
    ok = ok && alloc((void**)&c->dout, 8 * sizeof(double));
    if (g.closure == PF_CLOSURE_MLP) ok = ok && alloc((void**)&c->kappa, N * sizeof(float));
    if (c->ne) ok = ok && alloc((void**)&c->mom[0], N * sizeof(uint2)) && alloc((void**)&c->mom[1], N * sizeof(uint2));
    for (int i = 0; ok && i < PF_NUM_FIELDS - PF_NUM_BASE_FIELDS; ++i) {
        const int id = PF_NUM_BASE_FIELDS + i;
        const bool want = (id <= PF_FIELD_AVG_UZ && (g.derived & PF_DERIVED_AVERAGE)) ||
                          (id >= PF_FIELD_AVG_UXUX && id <= PF_FIELD_AVG_UYUZ && (g.derived & PF_DERIVED_REYNOLDS)) ||
                          (id >= PF_FIELD_VORT_X && (g.derived & PF_DERIVED_VORTICITY));
        if (want) ok = alloc((void**)&c->der[i], N * sizeof(float));
    }
    if (!ok) {
        pf_destroy(c);
        return PF_ERR_OOM;
    }

    pf::KParams& k = c->kp;
    k.f = c->f;
    for (int i = 0; i < 3; ++i) k.h[i] = static_cast<char*>(c->h) + (size_t)i * N * pop_bytes(c);
    for (int i = 0; i < pf::kNumFields; ++i) k.fld[i] = c->fld[i];
    k.phi_next = c->phi_next;
    k.bcp = c->bcp;
    k.gbc = c->gbc;
    k.kappa = c->kappa;
    k.ne = c->ne ? 1 : 0;
    k.ne_k = pf::kNeK;
    k.ne_kinv = 1.0f / pf::kNeK;
    k.mom_prev = c->mom[0];
    k.mom_cur = c->mom[1];
    k.mom = c->moments ? 1 : 0;
    k.rec_prev = c->rec[0];
    k.rec_cur = c->rec[1];
    k.bcf = c->bcf;
    k.nx = g.nx;
    k.ny = g.ny;
    k.nz = g.nz;
    k.N = (uint32_t)N;
    k.pitch_f = (uint32_t)pitch_f;
    k.plane_f = pitch_f * (uint64_t)g.ny;
    k.elems_f = c->f_elems;
    k.zchunk = choose_zchunk(g);
    k.tile_x = tile_x(g);
    k.ntx = (g.nx + k.tile_x - 1) / k.tile_x;
    k.nty = (g.ny + tile_rows(g, k.tile_x) - 1) / tile_rows(g, k.tile_x);
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
    (void)hipFree(c->h);
    for (auto* p : c->fld) (void)hipFree(p);
    (void)hipFree(c->phi_next);
    (void)hipFree(c->bcp);
    (void)hipFree(c->gbc);
    (void)hipFree(c->partials);
    (void)hipFree(c->dout);
    (void)hipFree(c->gather_buf);
    (void)hipFree(c->kappa);
    for (auto* p : c->mom) (void)hipFree(p);
    for (auto* p : c->rec) (void)hipFree(p);
    (void)hipFree(c->bcf);
    for (auto* p : c->der) (void)hipFree(p);
    (void)hipFree(c->mlp_w);
    (void)hipFree(c->mlp_b);
    for (auto& g : c->gexec)
        if (g) (void)hipGraphExecDestroy(g);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
}

size_t pf_device_bytes(const pf_ctx* c) { return c ? c->bytes : 0; }

int pf_init_distros(pf_ctx* c) {
    if (!c) return PF_ERR_CONFIG;
    PF_HIP(c, hipSetDevice(c->dev));
    PF_HIP(c, pf::launch_init_distros(c->kp, c->Q, c->jet, c->half, c->stream));
    const size_t plane = (size_t)c->cfg.nx * c->cfg.ny;
    PF_HIP(c, hipMemsetAsync(c->bcp, 0, 2 * pf::kPlaneVals * plane * sizeof(float), c->stream));
    PF_HIP(c, hipMemsetAsync(c->gbc, 0, 2 * plane * sizeof(float), c->stream));
    if (c->bcf) PF_HIP(c, hipMemsetAsync(c->bcf, 0, 18 * plane * sizeof(float), c->stream));
    PF_HIP(c, hipMemsetAsync(c->h, 0, 3 * c->N * pop_bytes(c), c->stream));
    // the ghost shell of phi is read from whichever buffer is current
    PF_HIP(c, hipMemcpyAsync(c->phi_next, c->fld[PF_FIELD_PHI], c->N * sizeof(float), hipMemcpyDeviceToDevice,
                             c->stream));
    c->parity = 0;
    c->fresh = true;
    c->inited = true;
    return PF_OK;
}

int pf_init(pf_ctx* c) {
    if (!c) return PF_ERR_CONFIG;
    PF_HIP(c, hipSetDevice(c->dev));
    const pf_config& g = c->cfg;
    for (int i = 0; i < pf::kNumFields; ++i)
        PF_HIP(c, hipMemsetAsync(c->fld[i], 0, c->N * sizeof(float), c->stream));
    for (auto* d : c->der)
        if (d) PF_HIP(c, hipMemsetAsync(d, 0, c->N * sizeof(float), c->stream));
    if (c->f) PF_HIP(c, hipMemsetAsync(c->f, 0, c->f_elems * pop_bytes(c), c->stream));
    PF_HIP(c, pf::launch_init_fields(c->kp, c->jet, c->stream));
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
    if (id < 0 || id >= PF_NUM_BASE_FIELDS) return fail(c, PF_ERR_CONFIG, "bad field id");
    if (n != c->N) return fail(c, PF_ERR_CONFIG, "size mismatch");
    PF_HIP(c, hipSetDevice(c->dev));
    PF_HIP(c, hipMemcpyAsync(c->fld[id], host, n * sizeof(float), hipMemcpyHostToDevice, c->stream));
    PF_HIP(c, hipStreamSynchronize(c->stream));
    return PF_OK;
}

static pf::DerivedPtrs derived_ptrs(const pf_ctx* c) {
    pf::DerivedPtrs d{};
    for (int i = 0; i < 10; ++i) d.avg[i] = c->der[i];
    for (int i = 0; i < 4; ++i) d.vort[i] = c->der[10 + i];
    return d;
}

// The phase sum wrote the new phi into phi_next: make it the current field.
static void swap_phi(pf_ctx* c) {
    std::swap(c->fld[PF_FIELD_PHI], c->phi_next);
    c->kp.fld[PF_FIELD_PHI] = c->fld[PF_FIELD_PHI];
    c->kp.phi_next = c->phi_next;
}

// Non-equilibrium form: the M this step wrote become those the next step
// reads.
static void swap_mom(pf_ctx* c) {
    if (c->moments) {  // moment form: likewise the coefficients
        std::swap(c->rec[0], c->rec[1]);
        c->kp.rec_prev = c->rec[0];
        c->kp.rec_cur = c->rec[1];
    }
    if (!c->ne) return;
    std::swap(c->mom[0], c->mom[1]);
    c->kp.mom_prev = c->mom[0];
    c->kp.mom_cur = c->mom[1];
}

// The kernels of one step before the boundary conditions: the phase sum, the
// learned closure when on, and the collide-and-stream of this parity.
static int launch_core(pf_ctx* c, int variant, hipEvent_t* ev) {
    if (ev) PF_HIP(c, hipEventRecord(ev[0], c->stream));
    const bool local = c->fresh;
    PF_HIP(c, pf::launch_phase_sum(c->kp, c->jet, c->fresh, c->half, c->stream));
    swap_phi(c);
    c->fresh = false;
    if (variant == 2) PF_HIP(c, pf::launch_closure(c->kp, c->stream));
    if (ev) PF_HIP(c, hipEventRecord(ev[1], c->stream));
    if (c->moments)
        PF_HIP(c, pf::launch_step_mom(c->kp, c->Q, c->jet, local, c->stream));
    else
        PF_HIP(c, pf::launch_step(c->kp, c->Q, c->jet, c->half, c->parity, variant, c->stream));
    if (ev) PF_HIP(c, hipEventRecord(ev[2], c->stream));
    return PF_OK;
}

// launch_core through a graph, as the reference replays its captured graph
// every step (src/cuda/CUDAGraph.cuh:34-58, src/main.cu:166).  The graph of a
// parity is captured on first use and kept while the kernel parameters other
// than the time index and the save flag (which the graph never sees: it is
// captured with save = 0 and the inflow kernel runs outside it) equal those
// it was captured with; phi's two buffers swap every step, so each parity
// always sees the same pair.  Otherwise it is captured again.
static int launch_core_graph(pf_ctx* c, int variant) {
    const int par = c->parity;
    pf::KParams key;
    std::memcpy(&key, &c->kp, sizeof key);  // bytewise, padding included, for the memcmp below
    key.t = 0;
    if (!c->gexec[par] || c->gvariant[par] != variant || std::memcmp(&key, &c->gkey[par], sizeof key) != 0) {
        if (c->gexec[par]) {
            PF_HIP(c, hipGraphExecDestroy(c->gexec[par]));
            c->gexec[par] = nullptr;
        }
        PF_HIP(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
        const int rc = launch_core(c, variant, nullptr);
        hipGraph_t graph = nullptr;
        const hipError_t e = hipStreamEndCapture(c->stream, &graph);
        if (rc != PF_OK) {
            if (graph) (void)hipGraphDestroy(graph);
            return rc;
        }
        if (e != hipSuccess) return hip_fail(c, e, "hipStreamEndCapture");
        const hipError_t ei = hipGraphInstantiate(&c->gexec[par], graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        if (ei != hipSuccess) {
            c->gexec[par] = nullptr;
            return hip_fail(c, ei, "hipGraphInstantiate");
        }
        std::memcpy(&c->gkey[par], &key, sizeof key);
        c->gvariant[par] = variant;
        swap_phi(c);  // capture ran nothing: undo its swap, the launch below redoes it
    }
    PF_HIP(c, hipGraphLaunch(c->gexec[par], c->stream));
    swap_phi(c);
    return PF_OK;
}

static int run_steps(pf_ctx* c, uint64_t t0, uint32_t nsteps, uint32_t flags, hipEvent_t* ev) {
    // Running means need the macroscopic fields of every step (the reference
    // computes them every step and averages after the boundary conditions).
    const bool avg = (c->cfg.derived & (PF_DERIVED_AVERAGE | PF_DERIVED_REYNOLDS)) != 0;
    const pf::DerivedPtrs dp = derived_ptrs(c);
    for (uint32_t s = 0; s < nsteps; ++s) {
        c->kp.t = (uint32_t)(t0 + s);
        c->kp.save = (avg || ((flags & PF_STEP_SAVE_LAST) && s + 1 == nsteps)) ? 1 : 0;
        const int variant = c->cfg.closure == PF_CLOSURE_MLP ? 2 : c->cfg.kernel;
        // The first step after the populations are set runs the local phase
        // sum once, and save steps run the materialising kernel variant:
        // both are launched directly.
        if ((flags & PF_STEP_GRAPH) && !ev && !c->fresh && !c->kp.save) {
            const int rc = launch_core_graph(c, variant);
            if (rc != PF_OK) return rc;
        } else {
            const int rc = launch_core(c, variant, ev);
            if (rc != PF_OK) return rc;
        }
        if (c->jet && c->moments)
            PF_HIP(c, pf::launch_jet_bc_mom(c->kp, c->Q, c->stream));
        else if (c->jet)
            PF_HIP(c, pf::launch_jet_bc(c->kp, c->Q, c->half, c->parity ^ 1, c->stream));
        swap_mom(c);
        if (avg) PF_HIP(c, pf::launch_average(c->kp, dp, (float)(t0 + s + 1), c->stream));
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
    const float* src = c->fld[0];
    if (id < PF_NUM_BASE_FIELDS) {
        src = c->fld[id];
    } else {
        src = c->der[id - PF_NUM_BASE_FIELDS];
        if (!src) return fail(c, PF_ERR_CONFIG, "derived field not enabled in pf_config.derived");
        if (id >= PF_FIELD_VORT_X) PF_HIP(c, pf::launch_vorticity(c->kp, derived_ptrs(c), c->stream));
    }
    PF_HIP(c, hipMemcpyAsync(host, src, n * sizeof(float), hipMemcpyDeviceToHost, c->stream));
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
    PF_HIP(c, pf::launch_gather(c->kp, c->Q, c->jet, c->half, c->parity, c->fresh, which, c->gather_buf,
                                c->stream));
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
    PF_HIP(c, pf::launch_phase_mass(c->kp, c->jet, c->fresh, c->half, c->partials, c->nred, c->stream));
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
    std::vector<float> params(PF_MLP_PARAMS);
    size_t o = 0;
    for (int i = 0; i < 64 * 27; ++i) params[o++] = W1[i];
    for (int i = 0; i < 64; ++i) params[o++] = b1[i];
    for (int i = 0; i < 64 * 64; ++i) params[o++] = W2[i];
    for (int i = 0; i < 64; ++i) params[o++] = b2[i];
    for (int i = 0; i < 64; ++i) params[o++] = W3[i];
    params[o++] = b3[0];
    std::vector<uint16_t> w;
    std::vector<float> b;
    pack_mlp(params.data(), w, b);
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

namespace {
const char* const kFieldNames[PF_NUM_FIELDS] = {
    "rho", "ux", "uy", "uz", "pxx", "pyy", "pzz", "pxy", "pxz", "pyz", "phi", "normx", "normy", "normz", "ind",
    "fsx", "fsy", "fsz", "avg_phi", "avg_ux", "avg_uy", "avg_uz", "avg_uxux", "avg_uyuy", "avg_uzuz", "avg_uxuy",
    "avg_uxuz", "avg_uyuz", "vort_x", "vort_y", "vort_z", "vort_mag"};
}

const char* pf_field_name(int32_t id) { return id >= 0 && id < PF_NUM_FIELDS ? kFieldNames[id] : nullptr; }

int pf_get_config(const pf_ctx* c, pf_config* out) {
    if (!c || !out) return PF_ERR_CONFIG;
    *out = c->cfg;
    out->device = c->dev;
    return PF_OK;
}

int pf_device_name(int32_t device, char* buf, size_t n) {
    if (!buf || n == 0) return PF_ERR_CONFIG;
    hipDeviceProp_t prop{};
    if (hipGetDeviceProperties(&prop, device) != hipSuccess) return PF_ERR_DEVICE;
    std::snprintf(buf, n, "%s (%s)", prop.name, prop.gcnArchName);
    return PF_OK;
}

int pf_write_fields(pf_ctx* c, const char* dir, uint64_t step, const int32_t* fields, int32_t nfields,
                    int32_t formats) {
    if (!c || !dir || !fields || nfields < 1) return PF_ERR_CONFIG;
    if (formats <= 0 || (formats & ~(PF_OUT_BIN | PF_OUT_VTI | PF_OUT_VTR))) return fail(c, PF_ERR_CONFIG, "bad formats");
    std::vector<std::string> names;
    for (int32_t i = 0; i < nfields; ++i) {
        if (!pf_field_name(fields[i])) return fail(c, PF_ERR_CONFIG, "bad field id");
        names.emplace_back(pf_field_name(fields[i]));
    }
    const std::string sfx = pfio::step_suffix(step);
    pfio::VtkWriter vti, vtr;
    const int nx = c->cfg.nx, ny = c->cfg.ny, nz = c->cfg.nz;
    if ((formats & PF_OUT_VTI) && !vti.open(pfio::join(dir, "step_" + sfx + ".vti"), pfio::kVti, nx, ny, nz, names))
        return fail(c, PF_ERR_IO, std::string("cannot write ") + pfio::join(dir, "step_" + sfx + ".vti"));
    if ((formats & PF_OUT_VTR) && !vtr.open(pfio::join(dir, "step_" + sfx + ".vtr"), pfio::kVtr, nx, ny, nz, names))
        return fail(c, PF_ERR_IO, std::string("cannot write ") + pfio::join(dir, "step_" + sfx + ".vtr"));
    std::vector<float> h(c->N);
    for (int32_t i = 0; i < nfields; ++i) {
        const int rc = pf_get_field(c, fields[i], h.data(), c->N);
        if (rc) return rc;
        if (formats & PF_OUT_BIN) {
            const std::string path = pfio::join(dir, names[(size_t)i] + sfx + ".bin");
            if (!pfio::write_bin(path, h.data(), c->N)) return fail(c, PF_ERR_IO, "cannot write " + path);
        }
        if ((formats & PF_OUT_VTI) && !vti.put(h.data())) return fail(c, PF_ERR_IO, "vti write failed");
        if ((formats & PF_OUT_VTR) && !vtr.put(h.data())) return fail(c, PF_ERR_IO, "vtr write failed");
    }
    if ((formats & PF_OUT_VTI) && !vti.close()) return fail(c, PF_ERR_IO, "vti write failed");
    if ((formats & PF_OUT_VTR) && !vtr.close()) return fail(c, PF_ERR_IO, "vtr write failed");
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

// ---------------------------------------------------------------------------
// learned closure: training on the device
// ---------------------------------------------------------------------------

extern "C++" {
namespace {
int mlp_fail(hipError_t e) { return e == hipSuccess ? PF_OK : PF_ERR_LAUNCH; }

// The 48 symmetries of the cube (axis permutations x sign flips) acting on
// the 3x3x3 stencil: perm[s][f] = source feature of output feature f.
std::vector<uint8_t> cube_symmetries() {
    std::vector<uint8_t> t;
    int ax[3] = {0, 1, 2};
    do {
        for (int sg = 0; sg < 8; ++sg) {
            for (int f = 0; f < 27; ++f) {
                const int d[3] = {f % 3 - 1, (f / 3) % 3 - 1, f / 9 - 1};
                int e[3];
                for (int a = 0; a < 3; ++a) e[a] = ((sg >> a) & 1 ? -1 : 1) * d[ax[a]];
                t.push_back((uint8_t)((e[2] + 1) * 9 + (e[1] + 1) * 3 + (e[0] + 1)));
            }
        }
    } while (std::next_permutation(ax, ax + 3));
    return t;
}

int mlp_upload_params(pf_mlp* m, const float* params) {
    std::vector<uint16_t> w;
    std::vector<float> b;
    pack_mlp(params, w, b);
    if (hipMemcpy(m->master, params, PF_MLP_PARAMS * sizeof(float), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(m->w16, w.data(), w.size() * 2, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(m->bias, b.data(), b.size() * 4, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemset(m->m1, 0, PF_MLP_PARAMS * 4) != hipSuccess || hipMemset(m->m2, 0, PF_MLP_PARAMS * 4) != hipSuccess)
        return PF_ERR_LAUNCH;
    m->step = 0;
    return PF_OK;
}

int mlp_ensure_parts(pf_mlp* m, size_t blocks) {
    if (blocks <= m->cap_parts) return PF_OK;
    (void)hipFree(m->parts);
    m->parts = nullptr;
    if (hipMalloc((void**)&m->parts, blocks * (PF_MLP_PARAMS + 1) * sizeof(float)) != hipSuccess) return PF_ERR_OOM;
    m->cap_parts = blocks;
    return PF_OK;
}

std::vector<float> pad_features(const float* x, size_t n) {
    std::vector<float> out(n * 32, 0.0f);
    for (size_t s = 0; s < n; ++s)
        for (int f = 0; f < 27; ++f) out[s * 32 + f] = x[s * 27 + f];
    return out;
}
}  // namespace
}  // extern "C++"

int pf_mlp_create(int32_t device, pf_mlp** out) {
    if (!out) return PF_ERR_CONFIG;
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || device < 0 || device >= ndev) return PF_ERR_DEVICE;
    if (hipSetDevice(device) != hipSuccess) return PF_ERR_DEVICE;
    pf_mlp* m = new (std::nothrow) pf_mlp();
    if (!m) return PF_ERR_OOM;
    m->dev = device;
    bool ok = hipStreamCreate(&m->stream) == hipSuccess;
    ok = ok && hipMalloc((void**)&m->master, PF_MLP_PARAMS * 4) == hipSuccess;
    ok = ok && hipMalloc((void**)&m->m1, PF_MLP_PARAMS * 4) == hipSuccess;
    ok = ok && hipMalloc((void**)&m->m2, PF_MLP_PARAMS * 4) == hipSuccess;
    ok = ok && hipMalloc((void**)&m->gsum, PF_MLP_PARAMS * 4) == hipSuccess;
    ok = ok && hipMalloc((void**)&m->w16, (64 * 32 + 64 * 64 + 64) * 2) == hipSuccess;
    ok = ok && hipMalloc((void**)&m->bias, 129 * 4) == hipSuccess;
    ok = ok && hipMalloc((void**)&m->loss, sizeof(double)) == hipSuccess;
    const std::vector<uint8_t> sym = cube_symmetries();
    ok = ok && hipMalloc((void**)&m->sym, sym.size()) == hipSuccess;
    // the fixed staging buffer of pf_mlp_predict_noalloc on this device
    ok = ok && pf::mlp_scratch_setup(device) == PF_OK;
    ok = ok && hipMemcpy(m->sym, sym.data(), sym.size(), hipMemcpyHostToDevice) == hipSuccess;
    if (!ok) {
        pf_mlp_destroy(m);
        return PF_ERR_OOM;
    }
    // default initialisation: uniform(-1/sqrt(fan_in), 1/sqrt(fan_in))
    std::vector<float> p(PF_MLP_PARAMS, 0.0f);
    std::mt19937 rng(1234);
    auto fill = [&](size_t off, size_t count, int fan_in) {
        std::uniform_real_distribution<float> u(-1.0f / std::sqrt((float)fan_in), 1.0f / std::sqrt((float)fan_in));
        for (size_t i = 0; i < count; ++i) p[off + i] = u(rng);
    };
    fill(0, 64 * 27, 27);
    fill(64 * 27 + 64, 64 * 64, 64);
    fill(64 * 27 + 64 + 64 * 64 + 64, 64, 64);
    if (mlp_upload_params(m, p.data())) {
        pf_mlp_destroy(m);
        return PF_ERR_LAUNCH;
    }
    *out = m;
    return PF_OK;
}

void pf_mlp_destroy(pf_mlp* m) {
    if (!m) return;
    (void)hipSetDevice(m->dev);
    if (m->stream) (void)hipStreamSynchronize(m->stream);
    for (void* p : {(void*)m->master, (void*)m->m1, (void*)m->m2, (void*)m->bias, (void*)m->parts, (void*)m->gsum,
                    (void*)m->w16, (void*)m->xs, (void*)m->ts, (void*)m->idx, (void*)m->sym, (void*)m->loss})
        (void)hipFree(p);
    if (m->stream) (void)hipStreamDestroy(m->stream);
    delete m;
}

int pf_mlp_set_params(pf_mlp* m, const float* params) {
    if (!m || !params) return PF_ERR_CONFIG;
    if (hipSetDevice(m->dev) != hipSuccess) return PF_ERR_DEVICE;
    return mlp_upload_params(m, params);
}

int pf_mlp_get_params(pf_mlp* m, float* params) {
    if (!m || !params) return PF_ERR_CONFIG;
    if (hipSetDevice(m->dev) != hipSuccess) return PF_ERR_DEVICE;
    if (hipStreamSynchronize(m->stream) != hipSuccess) return PF_ERR_LAUNCH;
    return mlp_fail(hipMemcpy(params, m->master, PF_MLP_PARAMS * 4, hipMemcpyDeviceToHost));
}

int pf_mlp_set_data(pf_mlp* m, const float* x, const float* t, size_t n) {
    if (!m || !x || !t || n == 0 || n > ((size_t)1 << 31)) return PF_ERR_CONFIG;
    if (hipSetDevice(m->dev) != hipSuccess) return PF_ERR_DEVICE;
    (void)hipFree(m->xs);
    (void)hipFree(m->ts);
    (void)hipFree(m->idx);
    m->xs = nullptr;
    m->ts = nullptr;
    m->idx = nullptr;
    const std::vector<float> xb = pad_features(x, n);
    if (hipMalloc((void**)&m->xs, xb.size() * 4) != hipSuccess || hipMalloc((void**)&m->ts, n * 4) != hipSuccess ||
        hipMalloc((void**)&m->idx, n * 4) != hipSuccess)
        return PF_ERR_OOM;
    if (hipMemcpy(m->xs, xb.data(), xb.size() * 4, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(m->ts, t, n * 4, hipMemcpyHostToDevice) != hipSuccess)
        return PF_ERR_LAUNCH;
    m->n = n;
    return PF_OK;
}

int pf_mlp_train(pf_mlp* m, uint32_t epochs, uint32_t batch, float lr, uint32_t seed, int32_t augment,
                 float* epoch_loss) {
    if (!m || m->n == 0 || batch == 0 || batch % 64 != 0 || !(lr > 0)) return PF_ERR_CONFIG;
    if (hipSetDevice(m->dev) != hipSuccess) return PF_ERR_DEVICE;
    if (int rc = mlp_ensure_parts(m, batch / 64)) return rc;
    std::vector<int> perm(m->n);
    std::iota(perm.begin(), perm.end(), 0);
    std::mt19937 rng(seed);
    for (uint32_t e = 0; e < epochs; ++e) {
        std::shuffle(perm.begin(), perm.end(), rng);
        if (hipMemcpyAsync(m->idx, perm.data(), m->n * 4, hipMemcpyHostToDevice, m->stream) != hipSuccess ||
            hipMemsetAsync(m->loss, 0, sizeof(double), m->stream) != hipSuccess)
            return PF_ERR_LAUNCH;
        const uint32_t sym_seed = seed * 0x85EBCA6Bu ^ (e + 1) * 0xC2B2AE35u;
        for (size_t s0 = 0; s0 < m->n; s0 += batch) {
            const size_t count = std::min<size_t>(batch, m->n - s0);
            const size_t blocks = (count + 63) / 64;
            hipError_t err = pf::launch_mlp_grad(m->w16, m->bias, m->xs, m->ts, m->idx, augment ? m->sym : nullptr, 48,
                                                 sym_seed, m->n, s0, count, 1.0f / (float)count, m->parts, m->stream);
            if (err != hipSuccess) return PF_ERR_LAUNCH;
            err = pf::launch_mlp_adam(m->parts, (int)blocks, m->master, m->m1, m->m2, m->w16, m->bias, lr, ++m->step,
                                      1, nullptr, m->loss, m->stream);
            if (err != hipSuccess) return PF_ERR_LAUNCH;
        }
        double l = 0.0;
        if (hipMemcpyAsync(&l, m->loss, sizeof(double), hipMemcpyDeviceToHost, m->stream) != hipSuccess ||
            hipStreamSynchronize(m->stream) != hipSuccess)
            return PF_ERR_LAUNCH;
        if (epoch_loss) epoch_loss[e] = (float)(l / (double)m->n);
    }
    return PF_OK;
}

int pf_mlp_gradients(pf_mlp* m, size_t n, float* grad, double* mse) {
    if (!m || !grad || n == 0 || n > m->n) return PF_ERR_CONFIG;
    if (hipSetDevice(m->dev) != hipSuccess) return PF_ERR_DEVICE;
    const size_t blocks = (n + 63) / 64;
    if (int rc = mlp_ensure_parts(m, blocks)) return rc;
    if (hipMemsetAsync(m->loss, 0, sizeof(double), m->stream) != hipSuccess) return PF_ERR_LAUNCH;
    if (pf::launch_mlp_grad(m->w16, m->bias, m->xs, m->ts, nullptr, nullptr, 1, 0, n, 0, n, 1.0f / (float)n,
                            m->parts, m->stream) != hipSuccess)
        return PF_ERR_LAUNCH;
    if (pf::launch_mlp_adam(m->parts, (int)blocks, m->master, m->m1, m->m2, m->w16, m->bias, 0.0f, 1, 0, m->gsum,
                            m->loss, m->stream) != hipSuccess)
        return PF_ERR_LAUNCH;
    double l = 0.0;
    if (hipMemcpyAsync(grad, m->gsum, PF_MLP_PARAMS * 4, hipMemcpyDeviceToHost, m->stream) != hipSuccess ||
        hipMemcpyAsync(&l, m->loss, sizeof(double), hipMemcpyDeviceToHost, m->stream) != hipSuccess ||
        hipStreamSynchronize(m->stream) != hipSuccess)
        return PF_ERR_LAUNCH;
    if (mse) *mse = l / (double)n;
    return PF_OK;
}

int pf_mlp_predict(pf_mlp* m, const float* x, size_t n, float* y) {
    if (!m || !x || !y || n == 0) return PF_ERR_CONFIG;
    if (hipSetDevice(m->dev) != hipSuccess) return PF_ERR_DEVICE;
    const std::vector<float> xb = pad_features(x, n);
    float* dx = nullptr;
    float* dy = nullptr;
    int rc = PF_OK;
    if (hipMalloc((void**)&dx, xb.size() * 4) != hipSuccess || hipMalloc((void**)&dy, n * 4) != hipSuccess) {
        rc = PF_ERR_OOM;
    } else if (hipMemcpyAsync(dx, xb.data(), xb.size() * 4, hipMemcpyHostToDevice, m->stream) != hipSuccess ||
               pf::launch_mlp_forward(m->w16, m->bias, dx, n, dy, m->stream) != hipSuccess ||
               hipMemcpyAsync(y, dy, n * 4, hipMemcpyDeviceToHost, m->stream) != hipSuccess ||
               hipStreamSynchronize(m->stream) != hipSuccess) {
        rc = PF_ERR_LAUNCH;
    }
    (void)hipFree(dx);
    (void)hipFree(dy);
    return rc;
}

int pf_mfma_check(int32_t device, const float* a, const float* b, float* c) {
    if (!a || !b || !c) return PF_ERR_CONFIG;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || device < 0 || device >= ndev) return PF_ERR_DEVICE;
    if (hipSetDevice(device) != hipSuccess) return PF_ERR_DEVICE;
    std::vector<uint16_t> ab(16 * 32), bb(32 * 16);
    for (int i = 0; i < 16 * 32; ++i) {
        ab[i] = f2bf(a[i]);
        bb[i] = f2bf(b[i]);
    }
    uint16_t *da = nullptr, *db = nullptr;
    float* dc = nullptr;
    int rc = PF_OK;
    if (hipMalloc((void**)&da, 1024) != hipSuccess || hipMalloc((void**)&db, 1024) != hipSuccess ||
        hipMalloc((void**)&dc, 1024) != hipSuccess) {
        rc = PF_ERR_OOM;
    } else if (hipMemcpy(da, ab.data(), 1024, hipMemcpyHostToDevice) != hipSuccess ||
               hipMemcpy(db, bb.data(), 1024, hipMemcpyHostToDevice) != hipSuccess ||
               pf::launch_mfma_layout(da, db, dc, nullptr) != hipSuccess ||
               hipMemcpy(c, dc, 1024, hipMemcpyDeviceToHost) != hipSuccess) {
        rc = PF_ERR_LAUNCH;
    }
    (void)hipFree(da);
    (void)hipFree(db);
    (void)hipFree(dc);
    return rc;
}

// Times `iters` launches of each streaming kernel over two arrays of
// `bytes` bytes with HIP events: out[0] read-only, out[1] write-only, out[2]
// copy (read + write bytes counted), in GB/s.
static int stream_bandwidth(int32_t device, size_t bytes, uint32_t iters, double out[3]) {
    if (!out || bytes < 1024 || iters == 0) return PF_ERR_CONFIG;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || device < 0 || device >= ndev) return PF_ERR_DEVICE;
    if (hipSetDevice(device) != hipSuccess) return PF_ERR_DEVICE;
    const size_t n4 = bytes / 16;
    float4 *a = nullptr, *b = nullptr;
    float* sink = nullptr;
    if (hipMalloc((void**)&a, n4 * 16) != hipSuccess) return PF_ERR_OOM;
    if (hipMalloc((void**)&b, n4 * 16) != hipSuccess || hipMalloc((void**)&sink, 16) != hipSuccess) {
        (void)hipFree(a);
        (void)hipFree(b);
        return PF_ERR_OOM;
    }
    hipStream_t s;
    hipEvent_t e0, e1;
    int rc = PF_OK;
    if (hipStreamCreate(&s) != hipSuccess) rc = PF_ERR_DEVICE;
    if (!rc && (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess)) rc = PF_ERR_DEVICE;
    if (!rc) {
        (void)hipMemsetAsync(a, 0, n4 * 16, s);
        const double moved[3] = {(double)(n4 * 16), (double)(n4 * 16), 2.0 * (double)(n4 * 16)};
        for (int k = 0; k < 3 && !rc; ++k) {
            auto launch = [&]() {
                if (k == 0) return pf::launch_read_only(a, n4, sink, s);
                if (k == 1) return pf::launch_write_only(b, n4, s);
                return pf::launch_copy(a, b, n4, s);
            };
            if (launch() != hipSuccess) rc = PF_ERR_LAUNCH;  // warm-up
            (void)hipEventRecord(e0, s);
            for (uint32_t i = 0; i < iters && !rc; ++i)
                if (launch() != hipSuccess) rc = PF_ERR_LAUNCH;
            (void)hipEventRecord(e1, s);
            if (!rc && hipEventSynchronize(e1) != hipSuccess) rc = PF_ERR_LAUNCH;
            float ms = 0.0f;
            if (!rc && hipEventElapsedTime(&ms, e0, e1) != hipSuccess) rc = PF_ERR_LAUNCH;
            if (!rc) out[k] = moved[k] * iters / (ms * 1e-3) / 1e9;
        }
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        (void)hipStreamDestroy(s);
    }
    (void)hipFree(a);
    (void)hipFree(b);
    (void)hipFree(sink);
    return rc;
}

int pf_copy_bandwidth(int32_t device, size_t bytes, uint32_t iters, double* gbs) {
    if (!gbs) return PF_ERR_CONFIG;
    double out[3] = {0, 0, 0};
    const int rc = stream_bandwidth(device, bytes, iters, out);
    if (!rc) *gbs = out[2];
    return rc;
}

int pf_stream_bandwidth(int32_t device, size_t bytes, uint32_t iters, double gbs[3]) {
    return stream_bandwidth(device, bytes, iters, gbs);
}

}  // extern "C"
