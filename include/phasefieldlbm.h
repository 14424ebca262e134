/*
 * phasefieldlbm.h -- C-ABI of the MI355X phase-field LBM timestep library.
 *
 * The reference (brenogemelgo/phaseFieldLBM) has no FFI: its host driver
 * (src/main.cu:35-215) calls compile-time C++ kernels that take an
 * `LBMFields` POD of device pointers (src/LBMIncludes.cuh:58-136).  This
 * header is the boundary that replaces those calls: a host driver in C,
 * C++ or Python (ctypes) links only against these symbols.  No torch or HIP
 * types cross it; every pointer argument is caller-owned host memory.
 *
 * Every entry point returns PF_OK (0) or a negative pf_status code; the
 * message of the last failure on a context is available from
 * pf_last_error().  Contexts are independent; one context is not
 * thread-safe.
 */
#ifndef PHASEFIELDLBM_H
#define PHASEFIELDLBM_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Compile-time velocity set of the reference (-DVS_D3Q19 | -DVS_D3Q27,
 * compile.sh:30, src/constants.cuh:39-43).  Each value selects a separate
 * template instantiation of the kernels. */
enum pf_velocity_set { PF_D3Q19 = 19, PF_D3Q27 = 27 };

/* Compile-time flow case of the reference (-DJET | -DDROPLET,
 * src/constants.cuh:45-49, src/flowCase/{Jet,Droplet}.cuh). */
enum pf_flow_case { PF_JET = 0, PF_DROPLET = 1 };

/* Population storage.  FP32 is the plain path.  FP16 stores f - w*rho0 and
 * the phase flux (the 3-vector from which the D3Q7 populations are rebuilt)
 * in binary16 (round to nearest even) with FP32 arithmetic, the role of
 * ENABLE_FP16 (src/cuda/precision.cuh:29-42), which in the reference stores
 * f in FP16 and g in FP32.  FP16_NE stores in binary16 2^12 times the
 * population's departure from the second-order equilibrium of the cell that
 * wrote it, w (A + b + b^2/2) with b = c.B, (A, B) = (rho (1 - 1.5|u|^2) - 1,
 * 3 rho u) of that cell, kept per cell in FP32 (16 B more per cell and step
 * written and read); the phase flux as in FP16.  Pair collision only
 * (PF_KERNEL_PAIR); no counterpart in the reference.  MOM stores, in FP32,
 * the coefficients of the polynomial in c that each cell's regularized
 * collision output is (10 per D3Q19 cell, 17 per D3Q27 cell, instead of Q
 * populations), and rebuilds every arriving population from the
 * coefficients of the cell that sent it: the reference's scheme with FP32
 * storage and arithmetic, 136 B of f traffic per D3Q27 update instead of
 * 216 B.  Pair collision with the FD closure only (PF_KERNEL_PAIR,
 * PF_CLOSURE_FD); no counterpart in the reference. */
enum pf_storage { PF_STORE_FP32 = 0, PF_STORE_FP16 = 1, PF_STORE_FP16_NE = 2, PF_STORE_MOM = 3 };

/* Inflow perturbation.  GAUSSIAN_XY is the reference's Box-Muller noise with
 * sigma = 0.08*u_inf on u_x, u_y (src/BoundaryConditions.cuh:60-69).
 * UNIFORM_Z is u_z = u_inf*(1 + a*(2U-1)) with U drawn from the same integer
 * hash; it is bit-reproducible between CPU and GPU. */
enum pf_noise { PF_NOISE_GAUSSIAN_XY = 0, PF_NOISE_UNIFORM_Z = 1, PF_NOISE_NONE = 2 };

/* Curvature closure: finite-difference div(n) (src/phaseField.cuh:142-219)
 * or the learned MLP at interface cells with the FD value as fallback. */
enum pf_closure { PF_CLOSURE_FD = 0, PF_CLOSURE_MLP = 1 };

/* Collision evaluation.  PAIR evaluates each opposite pair of directions
 * once (even and odd parts in c) with closed-form equilibrium and forcing
 * moments; LITERAL evaluates the reference's per-direction formulas
 * (src/lbm.cu:33-207) term by term.  Both are the same scheme; they differ by
 * floating-point rounding only. */
enum pf_kernel { PF_KERNEL_PAIR = 0, PF_KERNEL_LITERAL = 1 };

/* Field identifiers in host::FieldID order (src/fileIO/fields.cuh:33-68):
 * 0..17 the timestep's fields, 18..31 the derived fields of
 * src/derivedFields/ (available when enabled in pf_config.derived).  The
 * reference's passive scalar "c" is a stub that traps
 * (src/derivedFields/operators/passiveScalar.cuh:34-38) and is not offered. */
enum pf_field {
    PF_FIELD_RHO = 0, PF_FIELD_UX, PF_FIELD_UY, PF_FIELD_UZ,
    PF_FIELD_PXX, PF_FIELD_PYY, PF_FIELD_PZZ, PF_FIELD_PXY, PF_FIELD_PXZ, PF_FIELD_PYZ,
    PF_FIELD_PHI, PF_FIELD_NORMX, PF_FIELD_NORMY, PF_FIELD_NORMZ, PF_FIELD_IND,
    PF_FIELD_FSX, PF_FIELD_FSY, PF_FIELD_FSZ,
    PF_NUM_BASE_FIELDS,
    PF_FIELD_AVG_PHI = PF_NUM_BASE_FIELDS, PF_FIELD_AVG_UX, PF_FIELD_AVG_UY, PF_FIELD_AVG_UZ,
    PF_FIELD_AVG_UXUX, PF_FIELD_AVG_UYUY, PF_FIELD_AVG_UZUZ, PF_FIELD_AVG_UXUY, PF_FIELD_AVG_UXUZ,
    PF_FIELD_AVG_UYUZ,
    PF_FIELD_VORT_X, PF_FIELD_VORT_Y, PF_FIELD_VORT_Z, PF_FIELD_VORT_MAG,
    PF_NUM_FIELDS
};

/* Derived fields (bit mask, pf_config.derived), the reference's
 * TIME_AVERAGE, REYNOLDS_MOMENTS and VORTICITY_FIELDS builds
 * (src/derivedFields/operators/timeAverage.cuh, reynoldsMoments.cuh,
 * vorticityFields.cuh):
 *   AVERAGE    running means of phi, ux, uy, uz over every step, updated
 *              a += (v - a) / (t + 1) after step t, ghost cells included
 *   REYNOLDS   running means of ux*ux, uy*uy, uz*uz, ux*uy, ux*uz, uy*uz
 *   VORTICITY  curl u by central differences at interior cells (0 on the
 *              ghost shell), evaluated from the last saved velocity when
 *              read or written
 * With AVERAGE or REYNOLDS every step materialises the fields (as if run
 * with PF_STEP_SAVE_LAST), as the reference's per-step update needs. */
enum pf_derived { PF_DERIVED_AVERAGE = 1, PF_DERIVED_REYNOLDS = 2, PF_DERIVED_VORTICITY = 4 };

enum pf_status {
    PF_OK = 0,
    PF_ERR_CONFIG = -1,    /* invalid configuration or argument */
    PF_ERR_DEVICE = -2,    /* bad device index / no device */
    PF_ERR_OOM = -3,       /* device allocation failed */
    PF_ERR_LAUNCH = -4,    /* kernel launch or runtime error */
    PF_ERR_IO = -5,        /* file could not be written */
    PF_ERR_NONFINITE = -6, /* diagnostics found NaN/Inf */
    PF_ERR_STATE = -7      /* call out of order (e.g. step before init) */
};

/* pf_step flags */
#define PF_STEP_SAVE_LAST 1u /* materialise every pf_field on the last step */
/* Replay the step's kernels before the boundary conditions (phase sum,
 * closure, collide-and-stream) as a captured hipGraph, one per A-A parity,
 * as the reference replays its CUDA graph every step
 * (src/cuda/CUDAGraph.cuh:34-58, src/main.cu:163-172).  The boundary
 * kernels, which take the time index, are launched after it as in the
 * reference.  Results are bit-identical to direct launches.  A graph is
 * captured on first use and again whenever a kernel parameter it holds
 * changes; the first step after the populations are set and steps that
 * materialise the fields launch directly. */
#define PF_STEP_GRAPH 2u

/*
 * Run configuration.  The reference keeps these as constexpr values
 * (src/constants.cuh:82-141, src/functions/globalFunctions.cuh:260-462);
 * here they are runtime values so one library serves every grid.
 * pf_config_default() fills the reference's own JET or DROPLET values.
 */
typedef struct pf_config {
    int32_t velocity_set;  /* pf_velocity_set */
    int32_t flow_case;     /* pf_flow_case */
    int32_t storage;       /* pf_storage */
    int32_t closure;       /* pf_closure */
    int32_t nx, ny, nz;    /* grid including the 1-cell ghost shell */
    /* jet (src/constants.cuh:84-109) */
    int32_t diam;          /* nozzle diameter; radius = diam/2 (integer) */
    float u_inf;
    float reynolds_water;
    float reynolds_oil;
    float weber;
    int32_t noise;         /* pf_noise */
    float noise_amp;       /* UNIFORM_Z relative amplitude a */
    uint32_t seed;         /* hash base, reference 0x9E3779B9 */
    /* droplet (src/constants.cuh:113-141, src/initialConditions.cu:58-83) */
    float radius;
    float axis_x, axis_y, axis_z; /* coordinate divisors; reference 2,1,1 */
    float tau;             /* droplet relaxation time */
    float sigma;           /* droplet surface tension (jet: u^2 D / We) */
    /* phase field */
    float interface_width;
    float tau_g;           /* only sets the sharpening gamma, as in the reference */
    float reserved0;       /* reserved: must be 0 (pf_create returns PF_ERR_CONFIG otherwise) */
    /* execution */
    int32_t device;        /* HIP device index; <0 reads GPU_INDEX like setDeviceFromEnv */
    int32_t zchunk;        /* z cells marched per workgroup; 0 = automatic */
    float mlp_lo, mlp_hi;  /* MLP applies where mlp_lo < phi < mlp_hi (default 0.001, 0.999) */
    int32_t kernel;        /* pf_kernel */
    int32_t slot_pad;      /* elements of padding per population row (beyond Q*nx); <0 = automatic (0) */
    float mlp_kmin, mlp_kmax; /* MLP applies where the FD curvature lies in [kmin, kmax] (default -2, -0.03) */
    int32_t derived;       /* pf_derived bit mask; 0 = none (the reference's default build) */
    int32_t tile_x;        /* timestep tile width: 0 = automatic (64 when nx <= 64, else 128), 64 or 128 */
} pf_config;

typedef struct pf_ctx pf_ctx;

/* Fill `cfg` with the reference defaults of a flow case and velocity set. */
int pf_config_default(pf_config *cfg, int32_t flow_case, int32_t velocity_set);

/* Allocate every device buffer and the stream.  Replaces setDeviceFromEnv
 * (src/functions/hostFunctions.cuh:117-150), FieldAllocate
 * (src/fieldAllocate/FieldAllocate.cuh:47-122) and cudaStreamCreate
 * (src/main.cu:110-111). */
int pf_create(const pf_config *cfg, pf_ctx **out);
void pf_destroy(pf_ctx *ctx);

/* Initial conditions: setInitialDensity, setJet|setDroplet, setDistros
 * (src/main.cu:114-116, src/initialConditions.cu:33-138). */
int pf_init(pf_ctx *ctx);

/* Replace field `id` (one of the first PF_NUM_BASE_FIELDS) with host data of n =
 * nx*ny*nz floats; follow with pf_init_distros.  Used to start from an
 * arbitrary state (host-computed ICs). */
int pf_set_field(pf_ctx *ctx, int32_t id, const float *host, size_t n);
/* setDistros only (src/initialConditions.cu:101-138): populations at
 * equilibrium with the current rho, u, phi fields. */
int pf_init_distros(pf_ctx *ctx);

/* Advance nsteps timesteps; step s uses time index t0+s for the inflow
 * noise.  One step = graph (computePhase, computeNormals, computeForces,
 * computeMoments, streamCollide) + flow-case boundary conditions, as in
 * src/main.cu:163-172 and src/cuda/CUDAGraph.cuh:44-50.  Asynchronous. */
int pf_step(pf_ctx *ctx, uint64_t t0, uint32_t nsteps, uint32_t flags);
int pf_sync(pf_ctx *ctx);

/* Copy field `id` (n = nx*ny*nz floats) to host memory.  Fields other than
 * PHI hold the values of the last step run with PF_STEP_SAVE_LAST (or the
 * initial conditions), as the reference's arrays do after that step
 * (PostProcess::bin, src/postProcess/PostProcess.cuh:43-97).  Derived fields
 * must be enabled in pf_config.derived (else PF_ERR_CONFIG). */
int pf_get_field(pf_ctx *ctx, int32_t id, float *host, size_t n);

/* Copy the post-streaming populations f (which=0, Q*N) or g (which=1, 7*N)
 * in the reference's layout f[q*N + idx], unshifted for g and shifted by -w
 * for f (the reference's stored value), as FP32.  g is rebuilt from phi and
 * the stored phase flux: the populations the next phase sum adds. */
int pf_get_populations(pf_ctx *ctx, int32_t which, float *host, size_t n);

/* FP64 reductions over interior cells of the last saved fields:
 * out[0]=sum rho, out[1..3]=sum rho*u, out[4]=sum phi, out[5]=max|u|.
 * Returns PF_ERR_NONFINITE if any value is not finite. */
int pf_diagnostics(pf_ctx *ctx, double out[6]);

/* FP64 sum of phi over interior cells, from the current populations (no
 * save needed): the phase mass. */
int pf_phase_mass(pf_ctx *ctx, double *mass);

/* Learned-closure weights, fp32 host arrays in row-major [out][in]:
 * W1[64*27], b1[64], W2[64*64], b2[64], W3[64], b3[1].  Stored on the
 * device as BF16 weights and FP32 biases.  The network maps the normalised
 * 3x3x3 phi stencil to 4 x curvature: feature (dz+1)*9 + (dy+1)*3 + (dx+1)
 * is (phi(x+d) - phi(x)) / (|g| + 1e-4), g the central-difference gradient
 * of phi at x, except the centre (feature 13), which is 2*phi(x) - 1.  The
 * features enter the matrix cores as hi + lo BF16 pairs.  The timestep
 * evaluates the network once per cell with mlp_lo < phi < mlp_hi, on the
 * stencil in its canonical frame under the 48 symmetries of the cube (each
 * axis reflected so that phi(+1) - phi(-1) >= 0 along it, the axes ordered
 * by that difference's magnitude, largest first, the earlier axis first at
 * a tie), and applies, where the FD curvature lies in
 * [mlp_kmin, mlp_kmax], the mean over the cell and its interior neighbours
 * in that phi window.  No counterpart in the reference. */
int pf_set_closure(pf_ctx *ctx, const float *W1, const float *b1, const float *W2,
                   const float *b2, const float *W3, const float *b3);

/*
 * Training of the learned closure on the matrix cores (no counterpart in the
 * reference).  Parameters are one FP32 vector of PF_MLP_PARAMS values:
 * W1[64][27] | b1[64] | W2[64][64] | b2[64] | w3[64] | b3.  Samples are the
 * 27 stencil features described at pf_set_closure and a target
 * (4 x curvature).
 */
#define PF_MLP_PARAMS 6017
typedef struct pf_mlp pf_mlp;
int pf_mlp_create(int32_t device, pf_mlp **out);
void pf_mlp_destroy(pf_mlp *m);
int pf_mlp_set_params(pf_mlp *m, const float *params);
int pf_mlp_get_params(pf_mlp *m, float *params);
/* Copy n samples (x: n*27 floats, kept in FP32; t: n floats) to the device. */
int pf_mlp_set_data(pf_mlp *m, const float *x, const float *t, size_t n);
/* Adam (beta 0.9/0.999) on the mean squared error over minibatches of
 * `batch` samples (a multiple of 64), shuffled each epoch from `seed`; with
 * augment != 0 every sample is seen under a random one of the 48 symmetries
 * of the cube.  epoch_loss[e] = mean squared error seen during epoch e. */
int pf_mlp_train(pf_mlp *m, uint32_t epochs, uint32_t batch, float lr, uint32_t seed, int32_t augment,
                 float *epoch_loss);
/* Gradient of the mean squared error over the first n samples (one forward
 * and backward pass, no update): grad[PF_MLP_PARAMS], *mse. */
int pf_mlp_gradients(pf_mlp *m, size_t n, float *grad, double *mse);
/* Network output for n samples x[n*27]. */
int pf_mlp_predict(pf_mlp *m, const float *x, size_t n, float *y);
/* As pf_mlp_predict, but allocating nothing: no hipMalloc, hipFree or host
 * allocation.  No counterpart in the reference; for callers that cannot
 * allocate (code under HIP graph capture, tight serving loops, harnesses that
 * forbid it).  The samples and outputs are staged in a fixed device buffer of
 * the library (about 2 MB per device, held until the process ends) that
 * pf_mlp_create sets up on its device; n may be at most
 * PF_MLP_PREDICT_MAX (PF_ERR_CONFIG beyond).  Every call on one device shares
 * that buffer: calls that overlap in time on the same device race.  Results
 * equal pf_mlp_predict's bit for bit. */
#define PF_MLP_PREDICT_MAX 8192
int pf_mlp_predict_noalloc(pf_mlp *m, const float *x, size_t n, float *y);
/* One 16x16x32 BF16 MFMA tile, C = A B with A 16x32 and B 32x16 row-major
 * (rounded to BF16), as the kernels lay the operands out; for verifying the
 * lane layout. */
int pf_mfma_check(int32_t device, const float *a, const float *b, float *c);

/* Time `nsteps` steps (flags as pf_step) with HIP events on the context's
 * stream.  out_ms[0] = total, out_ms[1] = phase-sum kernel total (with the
 * learned-closure pass when it is enabled), out_ms[2] = fused
 * collide-stream kernel total, out_ms[3] = boundary total.
 * Per-kernel times are measured in a second, separately-evented pass. */
int pf_time_steps(pf_ctx *ctx, uint64_t t0, uint32_t nsteps, float out_ms[4]);

/* Achievable HBM bandwidth of `device`: a streaming copy of `bytes` bytes
 * repeated `iters` times, timed with HIP events; *gbs = (read + write bytes)
 * / time.  No counterpart in the reference; used to state the measured peak
 * beside the roofline and to calibrate memory-traffic counters. */
int pf_copy_bandwidth(int32_t device, size_t bytes, uint32_t iters, double *gbs);
/* The same measurement for three streaming kernels over `bytes`-byte
 * arrays: gbs[0] read-only, gbs[1] write-only, gbs[2] copy (read + write
 * bytes, the value pf_copy_bandwidth returns).  No counterpart in the
 * reference. */
int pf_stream_bandwidth(int32_t device, size_t bytes, uint32_t iters, double gbs[3]);

/*
 * Output.  Formats are a bit mask: PF_OUT_BIN writes <dir>/<name><step6>.bin,
 * raw FP32 of nx*ny*nz values (PostProcess::bin,
 * src/postProcess/PostProcess.cuh:43-97); PF_OUT_VTI writes
 * <dir>/step_<step6>.vti, VTK XML ImageData with raw appended FP32 point
 * arrays, origin 0 and unit spacing (PostProcess::vti, :99-261); PF_OUT_VTR
 * writes <dir>/step_<step6>.vtr, the same arrays on a VTK XML
 * RectilinearGrid with coordinates 0..n-1.  <step6> is the step number
 * zero-padded to six digits.  Names are those of pf_field_name.  The
 * directory must exist.
 */
enum pf_output { PF_OUT_BIN = 1, PF_OUT_VTI = 2, PF_OUT_VTR = 4 };
/* Fields `fields[0..nfields)` of the context as pf_get_field returns them;
 * the main loop's output step (src/main.cu:178-185). */
int pf_write_fields(pf_ctx *ctx, const char *dir, uint64_t step, const int32_t *fields, int32_t nfields,
                    int32_t formats);
/* One .vti (format PF_OUT_VTI) or .vtr (PF_OUT_VTR) file at `path` from
 * host arrays of nx*ny*nz floats; needs no device. */
int pf_write_vtk(const char *path, int32_t format, int32_t nx, int32_t ny, int32_t nz, const char *const *names,
                 const float *const *arrays, int32_t narrays);
/* Name of field `id` as the reference's output names it ("rho", "ux", ...,
 * src/main.cu:57-76); NULL for an unknown id. */
const char *pf_field_name(int32_t id);
/* The configuration a context was created with (device resolved). */
int pf_get_config(const pf_ctx *ctx, pf_config *out);
/* Marketing name and gfx architecture of a device, "name (gfxNNN)", for the
 * driver's banner (setDeviceFromEnv, src/functions/hostFunctions.cuh:117-150). */
int pf_device_name(int32_t device, char *buf, size_t n);

/* Bytes of device memory held by the context. */
size_t pf_device_bytes(const pf_ctx *ctx);

const char *pf_last_error(const pf_ctx *ctx);
const char *pf_status_string(int status);

#ifdef __cplusplus
}
#endif

#endif /* PHASEFIELDLBM_H */
