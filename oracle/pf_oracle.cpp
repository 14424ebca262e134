// ORACLE -- TEST INFRASTRUCTURE ONLY.
//
// A plain C++ (OpenMP) restatement of the reference phaseFieldLBM timestep,
// kernel by kernel, in the reference's order and index semantics.  It is the
// checker for the HIP library: only tests/, __graft_entry__.smoke() and
// bench.py's cpu_baseline leg may load it.  Nothing in the product links it.
//
// Instantiations: FP32 (parity oracle), FP32 arithmetic with binary16 f
// storage (the reference's FP16 mode; g stays FP32 as in precision.cuh),
// FP32 arithmetic with binary16 f and a binary16 phase flux (the library's
// binary16 form, below), and FP64 (accuracy oracle).
//
// The library's binary16 form stores, instead of the seven post-collision g
// of a cell, its phase flux h = (4 phi u + gamma phi (1 - phi) n) / 8 rounded
// to binary16; the g arriving at a neighbour are phi/8 +- h_k and the rest
// population phi/4.  The reference's g have no non-equilibrium part
// (src/velocitySet/D3Q7.cuh:139-143), so this is the same scheme with one
// more rounding; the mode streams g formed that way, so its fields can be
// held to FP32 round-off against the library.
//
// Reference files restated here (paths relative to the reference root):
//   src/phaseField.cuh:31-219        computePhase / computeNormals / computeForces
//   src/lbm.cu:33-239                computeMoments / streamCollide
//   src/BoundaryConditions.cuh:36-213 applyInflow / applyOutflow / noise
//   src/initialConditions.cu:33-138  setJet / setDroplet / setInitialDensity / setDistros
//   src/velocitySet/D3Q7.cuh, D3Q19.cuh, D3Q27.cuh   lattice tables and moments
//   src/functions/globalFunctions.cuh:104-462        geometry, sponge, relaxation
//   src/functions/deviceFunctions.cuh:33-101         indexing, guard, wrap, ramp
//   src/cuda/CUDAGraph.cuh:44-50 + src/main.cu:163-172  step order
//
// Pinning (what checks each part, all under tests/):
//   lattice tables, f_eq, f_neq, force, g_eq   closed-form moment identities
//   gradient / divergence stencils            exact on linear fields, sphere div n
//   computePhase / streamCollide              sum-phi and mass conservation, quiescent state;
//                                             every streamed g equals phi/8 +- h_k of its
//                                             source cell (the phase-flux identity)
//   binary16 conversion                       exhaustive round trip + numpy float16
//   inflow hash / uniform noise               known values of lowbias32 and hand-computed samples
//   Box-Muller inflow noise                   numpy sqrt(-2 ln u1) cos(2 pi u2) of the same uniforms
//   jet relaxation, sponge, sigma, gamma      closed forms per configuration (nu = uD/Re,
//                                             tau = 1/2 + 3 nu, sigma = u^2 D/We, nz/12 sponge
//                                             cells, cubic smoothstep, K = 100), pfo_consts/pfo_omega
//   droplet initial condition                 numpy FP64 tanh profile with the axis divisors
//   full step                                 Laplace law on the FP64 droplet, FP32 vs FP64
// The reference ships no golden vectors and cannot be built here (it needs
// nvcc), so no function is pinned against reference outputs directly.

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <new>
#include <type_traits>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

#include "../include/phasefieldlbm.h"

namespace {

// ---------------------------------------------------------------------------
// binary16 conversion, round to nearest even (the behaviour of __float2half,
// src/cuda/precision.cuh:36)
// ---------------------------------------------------------------------------
inline uint16_t f2h(float f) {
    uint32_t x;
    std::memcpy(&x, &f, 4);
    const uint32_t sign = (x >> 16) & 0x8000u;
    const uint32_t ax = x & 0x7fffffffu;
    if (ax >= 0x7f800000u) return (uint16_t)(sign | 0x7c00u | (ax > 0x7f800000u ? 0x200u : 0u));
    if (ax >= 0x477ff000u) return (uint16_t)(sign | 0x7c00u);  // >= 65520 -> inf
    if (ax >= 0x38800000u) {                                     // normal half
        const uint32_t mant = ax & 0x7fffffu;
        const uint32_t e = (ax >> 23) - 127u + 15u;
        uint32_t h = (e << 10) | (mant >> 13);
        const uint32_t rem = mant & 0x1fffu;
        if (rem > 0x1000u || (rem == 0x1000u && (h & 1u))) h++;
        return (uint16_t)(sign | h);
    }
    if (ax < 0x33000000u) return (uint16_t)sign;  // below 2^-25 -> signed zero
    const uint32_t m = (ax & 0x7fffffu) | 0x800000u;
    const int e = (int)(ax >> 23) - 127;  // -25 .. -15
    const int s = -(e + 1);               // 14 .. 24
    uint32_t h = m >> s;
    const uint32_t rem = m & ((1u << s) - 1u);
    const uint32_t half = 1u << (s - 1);
    if (rem > half || (rem == half && (h & 1u))) h++;
    return (uint16_t)(sign | h);
}

inline float h2f(uint16_t h) {
    const uint32_t sign = (uint32_t)(h & 0x8000u) << 16;
    const uint32_t e = (h >> 10) & 0x1fu;
    const uint32_t m = h & 0x3ffu;
    uint32_t x;
    if (e == 0) {
        if (m == 0) {
            x = sign;
        } else {  // subnormal: m * 2^-24
            float v = (float)m * 5.9604644775390625e-08f;
            std::memcpy(&x, &v, 4);
            x |= sign;
        }
    } else if (e == 31) {
        x = sign | 0x7f800000u | (m << 13);
    } else {
        x = sign | ((e - 15u + 127u) << 23) | (m << 13);
    }
    float f;
    std::memcpy(&f, &x, 4);
    return f;
}

// ---------------------------------------------------------------------------
// Lattices (src/velocitySet/D3Q19.cuh:92-141, D3Q27.cuh:96-145, D3Q7.cuh:80-127)
// ---------------------------------------------------------------------------
const int CX27[27] = {0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 0, 0, 1, -1, 1, -1, 0, 0, 1, -1, 1, -1, 1, -1, -1, 1};
const int CY27[27] = {0, 0, 0, 1, -1, 0, 0, 1, -1, 0, 0, 1, -1, -1, 1, 0, 0, 1, -1, 1, -1, 1, -1, -1, 1, 1, -1};
const int CZ27[27] = {0, 0, 0, 0, 0, 1, -1, 0, 0, 1, -1, 1, -1, 0, 0, -1, 1, -1, 1, 1, -1, -1, 1, 1, -1, 1, -1};

inline int wclass(int q) { return q == 0 ? 0 : (q <= 6 ? 1 : (q <= 18 ? 2 : 3)); }

template <class T>
struct Lat {
    int Q;
    T w[27];
    T wk[4];   // weight per class
    T cs2, as2;
};

template <class T>
Lat<T> make_lat(int Q) {
    Lat<T> L{};
    L.Q = Q;
    if (Q == 19) {
        L.wk[0] = (T)(1.0 / 3.0); L.wk[1] = (T)(1.0 / 18.0); L.wk[2] = (T)(1.0 / 36.0); L.wk[3] = (T)0;
    } else if (Q == 27) {
        L.wk[0] = (T)(8.0 / 27.0); L.wk[1] = (T)(2.0 / 27.0); L.wk[2] = (T)(1.0 / 54.0); L.wk[3] = (T)(1.0 / 216.0);
    } else {  // D3Q7
        L.wk[0] = (T)(1.0 / 4.0); L.wk[1] = (T)(1.0 / 8.0); L.wk[2] = (T)0; L.wk[3] = (T)0;
    }
    for (int q = 0; q < Q; ++q) L.w[q] = L.wk[wclass(q)];
    L.cs2 = (Q == 7) ? (T)(1.0 / 4.0) : (T)(1.0 / 3.0);
    L.as2 = (Q == 7) ? (T)4 : (T)3;
    return L;
}

// f_eq, shifted by -w (D3Q19.cuh:143-150, D3Q27.cuh:152-159)
template <class T>
inline T f_eq(const Lat<T>& L, int q, T rho, T uu, T cu) {
    const T w = L.w[q];
    if (L.Q == 19) return w * rho * ((T)1 - uu + cu + (T)0.5 * cu * cu) - w;
    const T oos = (T)(1.0 / 6.0);
    return w * rho * ((T)1 - uu + cu + (T)0.5 * cu * cu + oos * cu * cu * cu - uu * cu) - w;
}

// Regularized non-equilibrium part (D3Q19.cuh:152-171, D3Q27.cuh:161-190)
template <class T>
inline T f_neq(const Lat<T>& L, int q, T pxx, T pyy, T pzz, T pxy, T pxz, T pyz, T ux, T uy, T uz) {
    const int cx = CX27[q], cy = CY27[q], cz = CZ27[q];
    const T cs2 = L.cs2;
    T s = ((T)(cx * cx) - cs2) * pxx + ((T)(cy * cy) - cs2) * pyy + ((T)(cz * cz) - cs2) * pzz +
          (T)2 * ((T)(cx * cy) * pxy + (T)(cx * cz) * pxz + (T)(cy * cz) * pyz);
    if (L.Q == 27) {
        const T as2 = L.as2;
        s = s + (T)(cx * cx * cx - cx) * (as2 * ux * pxx) + (T)(cy * cy * cy - cy) * (as2 * uy * pyy) +
            (T)(cz * cz * cz - cz) * (as2 * uz * pzz) +
            as2 * (((T)(cx * cx * cy) - cs2 * (T)cy) * (pxx * uy + (T)2 * ux * pxy) +
                   ((T)(cx * cx * cz) - cs2 * (T)cz) * (pxx * uz + (T)2 * ux * pxz) +
                   ((T)(cx * cy * cy) - cs2 * (T)cx) * (pyy * ux + (T)2 * uy * pxy) +
                   ((T)(cy * cy * cz) - cs2 * (T)cz) * (pyy * uz + (T)2 * uy * pyz) +
                   ((T)(cx * cz * cz) - cs2 * (T)cx) * (pzz * ux + (T)2 * uz * pxz) +
                   ((T)(cy * cz * cz) - cs2 * (T)cy) * (pzz * uy + (T)2 * uz * pyz)) +
            (T)6 * (T)cx * (T)cy * (T)cz * (ux * pyz + uy * pxz + uz * pxy);
    }
    return (L.w[q] * (T)4.5) * s;
}

// Guo forcing term (D3Q19.cuh:173-187, D3Q27.cuh:192-206)
template <class T>
inline T force_q(const Lat<T>& L, int q, T cu, T ux, T uy, T uz, T fx, T fy, T fz) {
    const T cx = (T)CX27[q], cy = (T)CY27[q], cz = (T)CZ27[q];
    const T as2 = L.as2;
    return L.w[q] * ((as2 * (cx - ux) + as2 * cu * cx) * fx + (as2 * (cy - uy) + as2 * cu * cy) * fy +
                     (as2 * (cz - uz) + as2 * cu * cz) * fz);
}

template <class T>
inline T cu_q(const Lat<T>& L, int q, T ux, T uy, T uz) {
    return L.as2 * ((T)CX27[q] * ux + (T)CY27[q] * uy + (T)CZ27[q] * uz);
}

// D3Q7 equilibrium and sharpening flux (D3Q7.cuh:129-153)
template <class T>
inline T g_eq(const Lat<T>& G, int q, T phi, T ux, T uy, T uz) {
    return G.w[q] * phi * ((T)1 + G.as2 * ((T)CX27[q] * ux + (T)CY27[q] * uy + (T)CZ27[q] * uz));
}
template <class T>
inline T anti_diffusion(const Lat<T>& G, int q, T sharp, T nx, T ny, T nz) {
    return G.w[q] * sharp * ((T)CX27[q] * nx + (T)CY27[q] * ny + (T)CZ27[q] * nz);
}

// Pair-ordered first moment along `axis` of a value sampled at lattice
// directions, restricted to weight class k.  The reference writes these sums
// out by hand (src/phaseField.cuh:82-122, :168-210, src/lbm.cu:72-80): for each
// opposite pair (2j-1, 2j) in increasing j with c_axis != 0, it adds the
// value at the +1 member and subtracts the one at the -1 member, left to
// right.  `val(q)` returns the sample for direction q.
template <class T, class V>
inline T pair_chain(int Q, int axis, int k, V val) {
    T acc = (T)0;
    bool first = true;
    for (int p = 1; p + 1 < Q + 1; p += 2) {
        if (k >= 0 && wclass(p) != k) continue;
        const int* C = axis == 0 ? CX27 : (axis == 1 ? CY27 : CZ27);
        if (C[p] == 0) continue;
        const int pos = C[p] > 0 ? p : p + 1;
        const int neg = C[p] > 0 ? p + 1 : p;
        if (first) {
            acc = val(pos) - val(neg);
            first = false;
        } else {
            acc = acc + val(pos);
            acc = acc - val(neg);
        }
    }
    return acc;
}

// Isotropic gradient component (src/phaseField.cuh:82-128, :168-213): classes 1, 2
// (and 3 for D3Q27) each weighted, summed in that order.
template <class T, class V>
inline T iso_grad(const Lat<T>& L, int axis, V val) {
    T s = L.wk[1] * pair_chain<T>(L.Q, axis, 1, val) + L.wk[2] * pair_chain<T>(L.Q, axis, 2, val);
    if (L.Q == 27) s = s + L.wk[3] * pair_chain<T>(L.Q, axis, 3, val);
    return s;
}

// ---------------------------------------------------------------------------
// Simulation state
// ---------------------------------------------------------------------------
template <class T>
struct Consts {
    T sigma, gamma;
    T u_inf;
    // jet relaxation (globalFunctions.cuh:298-393)
    T visc_water, visc_oil, tau_water, tau_oil, omco_zmin;
    // droplet relaxation (globalFunctions.cuh:396-460)
    T omega_ref;
    // sponge (globalFunctions.cuh:260-292)
    T sp_z_start, sp_inv_nz_m1, sp_inv_sponge;
    // geometry (globalFunctions.cuh:104-125)
    T cx, cy, cz, R2;
};

template <class T>
inline T omega_from_nu(T nu, T as2) {
    return (T)(1.0 / (0.5 + (double)as2 * (double)nu));
}
template <class T>
inline T tau_from_nu(T nu, T as2) {
    return (T)0.5 + as2 * nu;
}

struct Base {
    virtual ~Base() {}
    virtual void consts(double* out) const = 0;
    virtual double omega_at(double phi, int z) const = 0;
    virtual double omco_zmax_at(double phi) const = 0;
    virtual double ramp_at(int z) const = 0;
    virtual int init() = 0;
    virtual int init_distros() = 0;
    virtual int step(uint64_t t0, uint32_t n) = 0;
    virtual int get_field(int id, double* out) = 0;
    virtual int set_field(int id, const double* in) = 0;
    virtual int get_pop(int which, double* out) = 0;
    virtual int set_pop(int which, const double* in) = 0;
};

// Shift of the library's binary16 non-equilibrium form: M = (A, B) =
// (rho (1 - 1.5|u|^2) - 1, 3 rho u), rounded to binary16, of the cell that
// writes a population,
// S_q(M) = w_q fma(b, fma(0.5, b, 1), A) = w_q (A + b (1 + b/2)), b = c_q . B
// (float fma, one rounding each); the population is stored as
// K (f~ - S_q(M)) in binary16 with K = 2^12 and read back as
// stored / K + S_q(M).  Evaluated in the library's operation order.
struct NeMom {
    float A, Bx, By, Bz;
};
// M is kept rounded to binary16, as the library stores it
inline NeMom ne_mom(float rho, float ux, float uy, float uz) {
    const float v = 1.0f - 1.5f * (ux * ux + uy * uy + uz * uz);
    const float r3 = 3.0f * rho;
    auto r16 = [](float a) { return h2f(f2h(a)); };
    return {r16(rho * v - 1.0f), r16(r3 * ux), r16(r3 * uy), r16(r3 * uz)};
}
inline float ne_shift(const Lat<float>& L, int q, const NeMom& m) {
    const float w = L.w[q];
    if (q == 0) return w * m.A;
    float b = 0.0f;
    bool any = false;
    const float comp[3] = {m.Bx, m.By, m.Bz};
    const int c[3] = {CX27[q], CY27[q], CZ27[q]};
    for (int a = 0; a < 3; ++a) {
        if (c[a] == 0) continue;
        b = any ? (c[a] > 0 ? b + comp[a] : b - comp[a]) : (c[a] > 0 ? comp[a] : -comp[a]);
        any = true;
    }
    return w * std::fma(b, std::fma(0.5f, b, 1.0f), m.A);
}
constexpr float kNeK = 4096.0f;

// HALF: f stored in binary16.  HALF_H: the phase populations are formed from
// a binary16 phase flux, as the library's binary16 form does.  NE: f stored
// as the library's binary16 non-equilibrium form does (kept here as the FP32
// value that form reads back).
template <class T, bool HALF, bool HALF_H = false, bool NE = false>
struct Sim final : Base {
    using S = typename std::conditional<HALF, uint16_t, T>::type;
    pf_config c;
    Lat<T> L, G;
    Lat<float> Lf;  // the lattice in FP32, for the non-equilibrium shift
    Consts<T> k;
    int nx, ny, nz;
    size_t N;
    std::vector<T> fld[PF_NUM_FIELDS];
    std::vector<S> f;
    std::vector<T> g;
    bool jet;

    explicit Sim(const pf_config& cfg) : c(cfg) {
        nx = c.nx; ny = c.ny; nz = c.nz;
        N = (size_t)nx * ny * nz;
        L = make_lat<T>(c.velocity_set);
        G = make_lat<T>(7);
        Lf = make_lat<float>(c.velocity_set);
        jet = c.flow_case == PF_JET;
        for (auto& v : fld) v.assign(N, (T)0);
        f.assign((size_t)L.Q * N, S{});
        g.assign((size_t)7 * N, S{});
        // physics (src/constants.cuh:92-141)
        const T W = (T)c.interface_width;
        const T tau_g = (T)c.tau_g;
        const T diff_int = G.cs2 * (tau_g - (T)0.5);
        const T kappa = (T)4 * diff_int / W;
        k.gamma = kappa / G.cs2;
        if (jet) {
            k.u_inf = (T)c.u_inf;
            k.sigma = (k.u_inf * k.u_inf * (T)c.diam) / (T)c.weber;
            k.visc_water = (T)(((double)k.u_inf * (double)c.diam) / (double)(T)c.reynolds_water);
            k.visc_oil = (T)(((double)k.u_inf * (double)c.diam) / (double)(T)c.reynolds_oil);
            k.tau_water = tau_from_nu(k.visc_water, L.as2);
            k.tau_oil = tau_from_nu(k.visc_oil, L.as2);
            k.omco_zmin = (T)1 - omega_from_nu(k.visc_oil, L.as2);
            const int cells = nz / 12;
            const T sponge = (T)((double)cells / (double)(nz - 1));
            k.sp_z_start = (T)((double)(nz - 1 - cells) / (double)(nz - 1));
            k.sp_inv_nz_m1 = (T)(1.0 / (double)(nz - 1));
            k.sp_inv_sponge = (T)(1.0 / (double)sponge);
            const int radius = c.diam / 2;
            k.R2 = (T)(radius * radius);
        } else {
            k.u_inf = (T)0;
            k.sigma = (T)c.sigma;
            const T tau = (T)c.tau;
            const T visc_ref = (tau - (T)0.5) / L.as2;
            k.omega_ref = omega_from_nu(visc_ref, L.as2);
        }
        k.cx = (T)(nx - 1) * (T)0.5;
        k.cy = (T)(ny - 1) * (T)0.5;
        k.cz = (T)(nz - 1) * (T)0.5;
    }

    size_t id3(int x, int y, int z) const { return (size_t)x + (size_t)nx * ((size_t)y + (size_t)ny * (size_t)z); }
    bool guard(int x, int y, int z) const {
        return x < 1 || y < 1 || z < 1 || x >= nx - 1 || y >= ny - 1 || z >= nz - 1;
    }
    static int wrap(int x, int n) { return x < 1 ? n - 2 : (x > n - 2 ? 1 : x); }

    T ld_f(int q, size_t i) const {
        if constexpr (HALF) return (T)h2f(f[(size_t)q * N + i]);
        else return f[(size_t)q * N + i];
    }
    void st_f(int q, size_t i, T v) {
        if constexpr (HALF) f[(size_t)q * N + i] = f2h((float)v);
        else f[(size_t)q * N + i] = v;
    }
    // store as the non-equilibrium form does, written by a cell with moments m
    void st_f_ne(int q, size_t i, T v, const NeMom& m) {
        if constexpr (NE) {
            const float sh = ne_shift(Lf, q, m);
            const float stored = h2f(f2h(kNeK * ((float)v - sh)));
            f[(size_t)q * N + i] = (T)(sh + stored / kNeK);
        } else {
            (void)m;
            st_f(q, i, v);
        }
    }
    T ld_g(int q, size_t i) const { return g[(size_t)q * N + i]; }
    void st_g(int q, size_t i, T v) { g[(size_t)q * N + i] = v; }

    T& F(int id, size_t i) { return fld[id][i]; }

    // -- initial conditions (src/initialConditions.cu) ----------------------
    int init() override {
        for (auto& v : fld) std::fill(v.begin(), v.end(), (T)0);
        // setInitialDensity (:85-99)
        std::fill(fld[PF_FIELD_RHO].begin(), fld[PF_FIELD_RHO].end(), (T)1);
        if (jet) {
            // setJet (:33-56)
            for (int y = 0; y < ny; ++y)
                for (int x = 0; x < nx; ++x) {
                    const T dx = (T)x - k.cx, dy = (T)y - k.cy;
                    if (dx * dx + dy * dy > k.R2) continue;
                    F(PF_FIELD_PHI, id3(x, y, 0)) = (T)1;
                    F(PF_FIELD_UZ, id3(x, y, 0)) = k.u_inf;
                }
        } else {
            // setDroplet (:58-83), axis divisors generalise the reference's dx/2
#pragma omp parallel for schedule(static)
            for (int z = 1; z < nz - 1; ++z)
                for (int y = 1; y < ny - 1; ++y)
                    for (int x = 1; x < nx - 1; ++x) {
                        const T dx = ((T)x - k.cx) / (T)c.axis_x;
                        const T dy = ((T)y - k.cy) / (T)c.axis_y;
                        const T dz = ((T)z - k.cz) / (T)c.axis_z;
                        const T L2 = std::sqrt(dx * dx + dy * dy + dz * dz);
                        const T arg = (T)(((double)(T)c.radius - (double)L2) / (double)(T)c.interface_width);
                        F(PF_FIELD_PHI, id3(x, y, z)) = (T)0.5 + (T)0.5 * std::tanh((T)2 * arg);
                    }
        }
        return init_distros();
    }

    // setDistros (src/initialConditions.cu:101-138), every cell incl. ghosts
    int init_distros() override {
#pragma omp parallel for schedule(static)
        for (int z = 0; z < nz; ++z)
            for (int y = 0; y < ny; ++y)
                for (int x = 0; x < nx; ++x) {
                    const size_t i = id3(x, y, z);
                    const T ux = F(PF_FIELD_UX, i), uy = F(PF_FIELD_UY, i), uz = F(PF_FIELD_UZ, i);
                    const T uu = (T)1.5 * (ux * ux + uy * uy + uz * uz);
                    const bool interior = !guard(x, y, z);
                    for (int q = 0; q < L.Q; ++q) {
                        const T feq = f_eq(L, q, F(PF_FIELD_RHO, i), uu, cu_q(L, q, ux, uy, uz));
                        if (NE && interior) {
                            // held for the first step as written by x - c_q
                            const int wx = wrap(x - CX27[q], nx), wy = wrap(y - CY27[q], ny);
                            const int wz = jet ? z - CZ27[q] : wrap(z - CZ27[q], nz);
                            const size_t w = id3(wx, wy, wz);
                            st_f_ne(q, i, feq,
                                    ne_mom((float)F(PF_FIELD_RHO, w), (float)F(PF_FIELD_UX, w), (float)F(PF_FIELD_UY, w),
                                           (float)F(PF_FIELD_UZ, w)));
                        } else {
                            st_f(q, i, feq);
                        }
                    }
                    for (int q = 0; q < 7; ++q) st_g(q, i, g_eq(G, q, F(PF_FIELD_PHI, i), ux, uy, uz));
                }
        return PF_OK;
    }

    // -- phase field (src/phaseField.cuh) -----------------------------------
    void computePhase() {  // :31-52
#pragma omp parallel for schedule(static)
        for (int z = 1; z < nz - 1; ++z)
            for (int y = 1; y < ny - 1; ++y)
                for (int x = 1; x < nx - 1; ++x) {
                    const size_t i = id3(x, y, z);
                    T phi = (T)0;
                    for (int q = 0; q < 7; ++q) phi += ld_g(q, i);
                    F(PF_FIELD_PHI, i) = phi;
                }
    }

    void computeNormals() {  // :54-140
#pragma omp parallel for schedule(static)
        for (int z = 1; z < nz - 1; ++z)
            for (int y = 1; y < ny - 1; ++y)
                for (int x = 1; x < nx - 1; ++x) {
                    const std::vector<T>& phi = fld[PF_FIELD_PHI];
                    auto val = [&](int q) { return phi[id3(x + CX27[q], y + CY27[q], z + CZ27[q])]; };
                    const T gx = L.as2 * iso_grad(L, 0, val);
                    const T gy = L.as2 * iso_grad(L, 1, val);
                    const T gz = L.as2 * iso_grad(L, 2, val);
                    const T ind = std::sqrt(gx * gx + gy * gy + gz * gz);
                    const T invInd = (T)1 / (ind + (T)1e-9);
                    const size_t i = id3(x, y, z);
                    F(PF_FIELD_IND, i) = ind;
                    F(PF_FIELD_NORMX, i) = gx * invInd;
                    F(PF_FIELD_NORMY, i) = gy * invInd;
                    F(PF_FIELD_NORMZ, i) = gz * invInd;
                }
    }

    void computeForces() {  // :142-219
#pragma omp parallel for schedule(static)
        for (int z = 1; z < nz - 1; ++z)
            for (int y = 1; y < ny - 1; ++y)
                for (int x = 1; x < nx - 1; ++x) {
                    auto at = [&](int id) {
                        return [&, id](int q) { return fld[id][id3(x + CX27[q], y + CY27[q], z + CZ27[q])]; };
                    };
                    const T scx = iso_grad(L, 0, at(PF_FIELD_NORMX));
                    const T scy = iso_grad(L, 1, at(PF_FIELD_NORMY));
                    const T scz = iso_grad(L, 2, at(PF_FIELD_NORMZ));
                    const T curvature = L.as2 * (scx + scy + scz);
                    const size_t i = id3(x, y, z);
                    const T stCurv = -k.sigma * curvature * F(PF_FIELD_IND, i);
                    F(PF_FIELD_FSX, i) = stCurv * F(PF_FIELD_NORMX, i);
                    F(PF_FIELD_FSY, i) = stCurv * F(PF_FIELD_NORMY, i);
                    F(PF_FIELD_FSZ, i) = stCurv * F(PF_FIELD_NORMZ, i);
                }
    }

    // -- hydrodynamics (src/lbm.cu) -----------------------------------------
    void computeMoments() {  // :33-128
#pragma omp parallel for schedule(static)
        for (int z = 1; z < nz - 1; ++z)
            for (int y = 1; y < ny - 1; ++y)
                for (int x = 1; x < nx - 1; ++x) {
                    const size_t i = id3(x, y, z);
                    T pop[27];
                    T rho = (T)0;
                    for (int q = 0; q < L.Q; ++q) {
                        pop[q] = ld_f(q, i);
                        rho += pop[q];
                    }
                    rho += (T)1;
                    F(PF_FIELD_RHO, i) = rho;
                    const T fsx = F(PF_FIELD_FSX, i), fsy = F(PF_FIELD_FSY, i), fsz = F(PF_FIELD_FSZ, i);
                    const T invRho = (T)1 / rho;
                    auto pv = [&](int q) { return pop[q]; };
                    T ux = invRho * pair_chain<T>(L.Q, 0, -1, pv);
                    T uy = invRho * pair_chain<T>(L.Q, 1, -1, pv);
                    T uz = invRho * pair_chain<T>(L.Q, 2, -1, pv);
                    ux += fsx * (T)0.5 * invRho;
                    uy += fsy * (T)0.5 * invRho;
                    uz += fsz * (T)0.5 * invRho;
                    F(PF_FIELD_UX, i) = ux;
                    F(PF_FIELD_UY, i) = uy;
                    F(PF_FIELD_UZ, i) = uz;
                    T pxx = 0, pyy = 0, pzz = 0, pxy = 0, pxz = 0, pyz = 0;
                    const T uu = (T)1.5 * (ux * ux + uy * uy + uz * uz);
                    for (int q = 0; q < L.Q; ++q) {
                        const T cx = (T)CX27[q], cy = (T)CY27[q], cz = (T)CZ27[q];
                        const T cu = cu_q(L, q, ux, uy, uz);
                        const T feq = f_eq(L, q, rho, uu, cu);
                        const T frc = force_q(L, q, cu, ux, uy, uz, fsx, fsy, fsz);
                        const T fneq = pop[q] - feq + (T)0.5 * frc;
                        pxx += fneq * cx * cx;
                        pyy += fneq * cy * cy;
                        pzz += fneq * cz * cz;
                        pxy += fneq * cx * cy;
                        pxz += fneq * cx * cz;
                        pyz += fneq * cy * cz;
                    }
                    const T trace = pxx + pyy + pzz;
                    pxx -= L.cs2 * trace;
                    pyy -= L.cs2 * trace;
                    pzz -= L.cs2 * trace;
                    F(PF_FIELD_PXX, i) = pxx;
                    F(PF_FIELD_PYY, i) = pyy;
                    F(PF_FIELD_PZZ, i) = pzz;
                    F(PF_FIELD_PXY, i) = pxy;
                    F(PF_FIELD_PXZ, i) = pxz;
                    F(PF_FIELD_PYZ, i) = pyz;
                }
    }

    // sponge_ramp (src/functions/deviceFunctions.cuh:94-101)
    T sponge_ramp(int z) const {
        const T zn = (T)z * k.sp_inv_nz_m1;
        T s = (zn - k.sp_z_start) * k.sp_inv_sponge;
        s = std::fmin(std::fmax(s, (T)0), (T)1);
        return s * s * ((T)3 - (T)2 * s);
    }
    T visc_local(T phi) const { return ((T)1 - phi) * k.visc_water + phi * k.visc_oil; }
    T tau_zmax(T phi) const { return tau_from_nu(visc_local(phi) * ((T)1 + (T)100), L.as2); }
    T omco_zmax(T phi) const { return (T)1 - omega_from_nu(visc_local(phi) * ((T)1 + (T)100), L.as2); }
    // collision frequency of a cell (src/lbm.cu:158-170)
    T omega_cell(T phi, int z) const {
        if (!jet) return k.omega_ref;
        const T tau_phi = ((T)1 - phi) * k.tau_water + phi * k.tau_oil;
        const T r = sponge_ramp(z);
        const T tau_eff = tau_phi + r * (tau_zmax(phi) - tau_phi);
        return (T)1 / tau_eff;
    }

    // order of pfo_consts
    void consts(double* o) const override {
        const double v[] = {(double)k.sigma,     (double)k.gamma,       (double)k.u_inf,
                            (double)k.visc_water, (double)k.visc_oil,   (double)k.tau_water,
                            (double)k.tau_oil,   (double)k.omco_zmin,   (double)k.omega_ref,
                            (double)k.sp_z_start, (double)k.sp_inv_nz_m1, (double)k.sp_inv_sponge,
                            (double)k.cx,        (double)k.cy,          (double)k.cz,
                            (double)k.R2};
        for (size_t i = 0; i < sizeof(v) / sizeof(v[0]); ++i) o[i] = v[i];
    }
    double omega_at(double phi, int z) const override { return (double)omega_cell((T)phi, z); }
    double omco_zmax_at(double phi) const override { return (double)omco_zmax((T)phi); }
    double ramp_at(int z) const override { return (double)sponge_ramp(z); }

    void streamCollide() {  // :130-239
#pragma omp parallel for schedule(static)
        for (int z = 1; z < nz - 1; ++z)
            for (int y = 1; y < ny - 1; ++y)
                for (int x = 1; x < nx - 1; ++x) {
                    const size_t i = id3(x, y, z);
                    const T rho = F(PF_FIELD_RHO, i), ux = F(PF_FIELD_UX, i), uy = F(PF_FIELD_UY, i),
                            uz = F(PF_FIELD_UZ, i);
                    const T pxx = F(PF_FIELD_PXX, i), pyy = F(PF_FIELD_PYY, i), pzz = F(PF_FIELD_PZZ, i);
                    const T pxy = F(PF_FIELD_PXY, i), pxz = F(PF_FIELD_PXZ, i), pyz = F(PF_FIELD_PYZ, i);
                    const T fsx = F(PF_FIELD_FSX, i), fsy = F(PF_FIELD_FSY, i), fsz = F(PF_FIELD_FSZ, i);
                    const T phi = F(PF_FIELD_PHI, i);
                    const T omega = omega_cell(phi, z);
                    const T omco = (T)1 - omega;
                    const T uu = (T)1.5 * (ux * ux + uy * uy + uz * uz);
                    const NeMom mw = ne_mom((float)rho, (float)ux, (float)uy, (float)uz);
                    for (int q = 0; q < L.Q; ++q) {
                        const T cu = cu_q(L, q, ux, uy, uz);
                        const T feq = f_eq(L, q, rho, uu, cu);
                        const T frc = force_q(L, q, cu, ux, uy, uz, fsx, fsy, fsz);
                        const T fneq = f_neq(L, q, pxx, pyy, pzz, pxy, pxz, pyz, ux, uy, uz);
                        int xx = x + CX27[q], yy = y + CY27[q], zz = z + CZ27[q];
                        xx = wrap(xx, nx);
                        yy = wrap(yy, ny);
                        if (!jet) zz = wrap(zz, nz);
                        st_f_ne(q, id3(xx, yy, zz), feq + omco * fneq + (T)0.5 * frc, mw);
                    }
                    const T nrx = F(PF_FIELD_NORMX, i), nry = F(PF_FIELD_NORMY, i), nrz = F(PF_FIELD_NORMZ, i);
                    const T sharp = k.gamma * phi * ((T)1 - phi);
                    T hflux[3] = {0, 0, 0};
                    if constexpr (HALF_H) {
                        // the binary16 phase flux and the g formed from it
                        const T p4 = (T)4 * phi;
                        hflux[0] = (T)h2f(f2h((float)((T)0.125 * (p4 * ux + sharp * nrx))));
                        hflux[1] = (T)h2f(f2h((float)((T)0.125 * (p4 * uy + sharp * nry))));
                        hflux[2] = (T)h2f(f2h((float)((T)0.125 * (p4 * uz + sharp * nrz))));
                    }
                    for (int q = 0; q < 7; ++q) {
                        T gq;
                        if constexpr (HALF_H) {
                            if (q == 0) gq = (T)0.25 * phi;
                            else if (q & 1) gq = (T)0.125 * phi + hflux[(q - 1) / 2];
                            else gq = (T)0.125 * phi - hflux[(q - 1) / 2];
                        } else {
                            gq = g_eq(G, q, phi, ux, uy, uz) + anti_diffusion(G, q, sharp, nrx, nry, nrz);
                        }
                        int xx = x + CX27[q], yy = y + CY27[q], zz = z + CZ27[q];
                        xx = wrap(xx, nx);
                        yy = wrap(yy, ny);
                        if (!jet) zz = wrap(zz, nz);
                        st_g(q, id3(xx, yy, zz), gq);
                    }
                }
    }

    // -- boundary conditions (src/BoundaryConditions.cuh) -------------------
    static uint32_t hash32(uint32_t x) {  // :172-181 (lowbias32)
        x ^= x >> 16;
        x *= 0x7FEB352Du;
        x ^= x >> 15;
        x *= 0x846CA68Bu;
        x ^= x >> 16;
        return x;
    }
    static T uniform01(uint32_t s) {  // :183-188
        const T inv2_32 = (T)2.3283064365386963e-10;
        return ((T)s + (T)0.5) * inv2_32;
    }
    static T box_muller(T rrx, T rry) {  // :190-198
        rrx = std::fmax(rrx, (T)1e-12);
        const T r = std::sqrt(-(T)2 * std::log(rrx));
        const T two_pi = (T)2 * (T)3.141592654f;
        const T theta = two_pi * rry;
        return r * std::cos(theta);
    }
    uint32_t noise_base(uint32_t salt, uint32_t x, uint32_t y, uint32_t t) const {  // :200-213
        return (c.seed ^ salt) ^ x ^ (y * 0x85EBCA6Bu) ^ (t * 0xC2B2AE35u);
    }
    T white_noise(uint32_t salt, uint32_t x, uint32_t y, uint32_t t) const {
        const uint32_t base = noise_base(salt, x, y, t);
        return box_muller(uniform01(hash32(base)), uniform01(hash32(base ^ 0x68BC21EBu)));
    }

    void applyInflow(uint32_t t) {  // :36-105
        const T sigma_u = (T)0.08 * k.u_inf;
        for (int y = 0; y < ny; ++y)
            for (int x = 0; x < nx; ++x) {
                const T dx = (T)x - k.cx, dy = (T)y - k.cy;
                // Frozen wall: no population reaches z = 1 from z = 0 off the
                // nozzle, so g_5 there keeps its initial value 0.
                if (dx * dx + dy * dy > k.R2) continue;
                const size_t ib = id3(x, y, 0);
                T ux = 0, uy = 0, uz = k.u_inf;
                if (c.noise == PF_NOISE_GAUSSIAN_XY) {
                    ux = sigma_u * white_noise(0xA341316Cu, x, y, t);
                    uy = sigma_u * white_noise(0xC8013EA4u, x, y, t);
                } else if (c.noise == PF_NOISE_UNIFORM_Z) {
                    const T U = uniform01(hash32(noise_base(0xA341316Cu, x, y, t)));
                    uz = k.u_inf * ((T)1 + (T)c.noise_amp * ((T)2 * U - (T)1));
                }
                const T rho = (T)1, phi = (T)1;
                F(PF_FIELD_RHO, ib) = rho;
                F(PF_FIELD_PHI, ib) = phi;
                F(PF_FIELD_UX, ib) = ux;
                F(PF_FIELD_UY, ib) = uy;
                F(PF_FIELD_UZ, ib) = uz;
                const T uu = (T)1.5 * (ux * ux + uy * uy + uz * uz);
                for (int q = 0; q < L.Q; ++q) {
                    if (CZ27[q] != 1) continue;
                    const size_t fn = id3(x + CX27[q], y + CY27[q], 1);
                    const T cu = cu_q(L, q, ux, uy, uz);
                    const T feq = f_eq(L, q, rho, uu, cu);
                    const T fneq = f_neq(L, q, F(PF_FIELD_PXX, fn), F(PF_FIELD_PYY, fn), F(PF_FIELD_PZZ, fn),
                                         F(PF_FIELD_PXY, fn), F(PF_FIELD_PXZ, fn), F(PF_FIELD_PYZ, fn),
                                         F(PF_FIELD_UX, fn), F(PF_FIELD_UY, fn), F(PF_FIELD_UZ, fn));
                    st_f_ne(q, fn, feq + k.omco_zmin * fneq, ne_mom(1.0f, (float)ux, (float)uy, (float)uz));
                }
                st_g(5, id3(x, y, 1), G.w[5] * phi * ((T)1 + G.as2 * uz));
            }
    }

    void applyOutflow() {  // :107-169
        for (int y = 1; y < ny - 1; ++y)
            for (int x = 1; x < nx - 1; ++x) {
                const size_t ib = id3(x, y, nz - 1), im = id3(x, y, nz - 2);
                for (int id : {PF_FIELD_RHO, PF_FIELD_PHI, PF_FIELD_UX, PF_FIELD_UY, PF_FIELD_UZ})
                    F(id, ib) = F(id, im);
                const T rho = F(PF_FIELD_RHO, ib), phi = F(PF_FIELD_PHI, ib);
                const T ux = F(PF_FIELD_UX, ib), uy = F(PF_FIELD_UY, ib), uz = F(PF_FIELD_UZ, ib);
                const T uu = (T)1.5 * (ux * ux + uy * uy + uz * uz);
                for (int q = 0; q < L.Q; ++q) {
                    if (CZ27[q] != -1) continue;
                    const int xx = wrap(x + CX27[q], nx), yy = wrap(y + CY27[q], ny);
                    const size_t fn = id3(xx, yy, nz - 2);
                    const T cu = cu_q(L, q, ux, uy, uz);
                    const T feq = f_eq(L, q, rho, uu, cu);
                    const T fneq = f_neq(L, q, F(PF_FIELD_PXX, fn), F(PF_FIELD_PYY, fn), F(PF_FIELD_PZZ, fn),
                                         F(PF_FIELD_PXY, fn), F(PF_FIELD_PXZ, fn), F(PF_FIELD_PYZ, fn),
                                         F(PF_FIELD_UX, fn), F(PF_FIELD_UY, fn), F(PF_FIELD_UZ, fn));
                    const T frc = force_q(L, q, cu, ux, uy, uz, F(PF_FIELD_FSX, fn), F(PF_FIELD_FSY, fn),
                                          F(PF_FIELD_FSZ, fn));
                    st_f_ne(q, fn, feq + omco_zmax(phi) * fneq + (T)0.5 * frc,
                            ne_mom((float)rho, (float)ux, (float)uy, (float)uz));
                }
                st_g(6, im, G.w[6] * phi * ((T)1 - G.as2 * uz));
            }
    }

    int step(uint64_t t0, uint32_t n) override {
        for (uint32_t s = 0; s < n; ++s) {
            computePhase();
            computeNormals();
            computeForces();
            computeMoments();
            streamCollide();
            if (jet) {
                applyInflow((uint32_t)(t0 + s));
                applyOutflow();
            }
        }
        return PF_OK;
    }

    int get_field(int id, double* out) override {
        if (id < 0 || id >= PF_NUM_FIELDS) return PF_ERR_CONFIG;
        for (size_t i = 0; i < N; ++i) out[i] = (double)fld[id][i];
        return PF_OK;
    }
    int set_field(int id, const double* in) override {
        if (id < 0 || id >= PF_NUM_FIELDS) return PF_ERR_CONFIG;
        for (size_t i = 0; i < N; ++i) fld[id][i] = (T)in[i];
        return PF_OK;
    }
    // Replace the post-streaming populations (the state the next step reads),
    // e.g. with a GPU state, to take one step from it.
    int set_pop(int which, const double* in) override {
        if (which == 0) {
            for (size_t q = 0; q < (size_t)L.Q; ++q)
                for (size_t i = 0; i < N; ++i) st_f((int)q, i, (T)in[q * N + i]);
        } else {
            for (size_t q = 0; q < 7; ++q)
                for (size_t i = 0; i < N; ++i) st_g((int)q, i, (T)in[q * N + i]);
        }
        return PF_OK;
    }
    int get_pop(int which, double* out) override {
        if (which == 0) {
            for (size_t q = 0; q < (size_t)L.Q; ++q)
                for (size_t i = 0; i < N; ++i) out[q * N + i] = (double)ld_f((int)q, i);
        } else {
            for (size_t q = 0; q < 7; ++q)
                for (size_t i = 0; i < N; ++i) out[q * N + i] = (double)ld_g((int)q, i);
        }
        return PF_OK;
    }
};

}  // namespace

extern "C" {

// precision: 32 = FP32, 16 = FP32 math with binary16 f (the reference's FP16
// mode), 17 = FP32 math with binary16 f and binary16 phase flux (the
// library's binary16 form), 18 = the library's binary16 non-equilibrium
// form (f as it reads it back, binary16 phase flux), 64 = FP64
void* pfo_create(const pf_config* cfg, int precision) {
    if (!cfg) return nullptr;
    if (cfg->velocity_set != 19 && cfg->velocity_set != 27) return nullptr;
    if (cfg->nx < 4 || cfg->ny < 4 || cfg->nz < (cfg->flow_case == PF_JET ? 12 : 4)) return nullptr;
    try {
        if (precision == 32) return static_cast<Base*>(new Sim<float, false>(*cfg));
        if (precision == 16) return static_cast<Base*>(new Sim<float, true>(*cfg));
        if (precision == 17) return static_cast<Base*>(new Sim<float, true, true>(*cfg));
        if (precision == 18) return static_cast<Base*>(new Sim<float, false, true, true>(*cfg));
        if (precision == 64) return static_cast<Base*>(new Sim<double, false>(*cfg));
    } catch (const std::bad_alloc&) {
        return nullptr;
    }
    return nullptr;
}
void pfo_destroy(void* h) { delete static_cast<Base*>(h); }
int pfo_init(void* h) { return static_cast<Base*>(h)->init(); }
int pfo_init_distros(void* h) { return static_cast<Base*>(h)->init_distros(); }
int pfo_step(void* h, uint64_t t0, uint32_t n) { return static_cast<Base*>(h)->step(t0, n); }
int pfo_get_field(void* h, int id, double* out) { return static_cast<Base*>(h)->get_field(id, out); }
int pfo_set_field(void* h, int id, const double* in) { return static_cast<Base*>(h)->set_field(id, in); }
int pfo_get_pop(void* h, int which, double* out) { return static_cast<Base*>(h)->get_pop(which, out); }
int pfo_set_pop(void* h, int which, const double* in) { return static_cast<Base*>(h)->set_pop(which, in); }

// Derived constants of a configuration, in the oracle's precision:
// out[0..15] = sigma, gamma, u_inf, visc_water, visc_oil, tau_water, tau_oil,
// omco_zmin, omega_ref, sponge z_start, 1/(nz-1), 1/sponge, centre x, y, z,
// nozzle radius^2 (src/constants.cuh:92-141,
// src/functions/globalFunctions.cuh:104-125,260-462).
int pfo_consts(void* h, double* out) {
    static_cast<Base*>(h)->consts(out);
    return PF_OK;
}
// Collision frequency streamCollide uses at phase phi in plane z
// (src/lbm.cu:158-170), 1 - omega_zmax(phi) of the outflow
// (src/BoundaryConditions.cuh:164) and the sponge ramp of plane z
// (src/functions/deviceFunctions.cuh:94-101).
double pfo_omega(void* h, double phi, int z) { return static_cast<Base*>(h)->omega_at(phi, z); }
double pfo_omco_zmax(void* h, double phi) { return static_cast<Base*>(h)->omco_zmax_at(phi); }
double pfo_sponge_ramp(void* h, int z) { return static_cast<Base*>(h)->ramp_at(z); }

int pfo_num_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}
void pfo_set_num_threads(int n) {
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
#else
    (void)n;
#endif
}

// Per-cell lattice primitives in FP64 for the closed-form tests.
int pfo_lattice(int Q, int* cx, int* cy, int* cz, double* w) {
    if (Q != 7 && Q != 19 && Q != 27) return PF_ERR_CONFIG;
    const Lat<double> L = make_lat<double>(Q);
    for (int q = 0; q < Q; ++q) {
        cx[q] = CX27[q]; cy[q] = CY27[q]; cz[q] = CZ27[q]; w[q] = L.w[q];
    }
    return PF_OK;
}
int pfo_cell_feq(int Q, double rho, const double* u, double* out) {
    const Lat<double> L = make_lat<double>(Q);
    const double uu = 1.5 * (u[0] * u[0] + u[1] * u[1] + u[2] * u[2]);
    for (int q = 0; q < Q; ++q) out[q] = f_eq(L, q, rho, uu, cu_q(L, q, u[0], u[1], u[2]));
    return PF_OK;
}
int pfo_cell_fneq(int Q, const double* P, const double* u, double* out) {
    const Lat<double> L = make_lat<double>(Q);
    for (int q = 0; q < Q; ++q) out[q] = f_neq(L, q, P[0], P[1], P[2], P[3], P[4], P[5], u[0], u[1], u[2]);
    return PF_OK;
}
int pfo_cell_force(int Q, const double* u, const double* F, double* out) {
    const Lat<double> L = make_lat<double>(Q);
    for (int q = 0; q < Q; ++q)
        out[q] = force_q(L, q, cu_q(L, q, u[0], u[1], u[2]), u[0], u[1], u[2], F[0], F[1], F[2]);
    return PF_OK;
}
int pfo_cell_geq(double phi, const double* u, const double* n, double sharp, double* geq, double* h) {
    const Lat<double> G = make_lat<double>(7);
    for (int q = 0; q < 7; ++q) {
        geq[q] = g_eq(G, q, phi, u[0], u[1], u[2]);
        h[q] = anti_diffusion(G, q, sharp, n[0], n[1], n[2]);
    }
    return PF_OK;
}
// Isotropic gradient of samples v[q] at the Q directions (FP64).
int pfo_iso_grad(int Q, const double* v, double* out3) {
    const Lat<double> L = make_lat<double>(Q);
    auto val = [&](int q) { return v[q]; };
    for (int a = 0; a < 3; ++a) out3[a] = L.as2 * iso_grad(L, a, val);
    return PF_OK;
}
uint16_t pfo_f2h(float x) { return f2h(x); }
float pfo_h2f(uint16_t h) { return h2f(h); }
uint32_t pfo_hash32(uint32_t x) { return Sim<float, false>::hash32(x); }

}  // extern "C"
