/* The smallest host program on the C-ABI: plain C (C99), no HIP headers.
 *
 *   minimal [DIR] [NSTEPS]
 *
 * Runs the reference's default D3Q27 jet (64x64x128, D = 10) for NSTEPS
 * steps (default 100), writes phi and u_z as step_NNNNNN.vtr into DIR
 * (default ".") every 50 steps and at the end, and prints the FP64
 * diagnostics.  It is the C counterpart of src/main.cu:114-188: pf_create +
 * pf_init replace the allocation and initial conditions, pf_step the graph
 * launch and boundary conditions, pf_write_fields the PostProcess calls. */
#include <inttypes.h>
#include <stdio.h>
#include <stdlib.h>

#include "phasefieldlbm.h"

static int check(pf_ctx *ctx, int rc, const char *what) {
    if (rc != PF_OK)
        fprintf(stderr, "%s: %s (%s)\n", what, pf_status_string(rc), ctx ? pf_last_error(ctx) : "");
    return rc != PF_OK;
}

int main(int argc, char **argv) {
    const char *dir = argc > 1 ? argv[1] : ".";
    const uint64_t nsteps = argc > 2 ? strtoull(argv[2], NULL, 10) : 100;
    const uint64_t save = 50;

    pf_config cfg;
    if (check(NULL, pf_config_default(&cfg, PF_JET, PF_D3Q27), "pf_config_default")) return 1;
    pf_ctx *ctx = NULL;
    if (check(NULL, pf_create(&cfg, &ctx), "pf_create")) return 1;
    int bad = check(ctx, pf_init(ctx), "pf_init");
    const int32_t ids[2] = {PF_FIELD_PHI, PF_FIELD_UZ};
    for (uint64_t t = 0; !bad && t < nsteps; t += save) {
        const uint32_t n = (uint32_t)(t + save <= nsteps ? save : nsteps - t);
        bad = check(ctx, pf_step(ctx, t, n, PF_STEP_SAVE_LAST), "pf_step") ||
              check(ctx, pf_write_fields(ctx, dir, t + n - 1, ids, 2, PF_OUT_VTR), "pf_write_fields");
    }
    double d[6];
    if (!bad && !check(ctx, pf_diagnostics(ctx, d), "pf_diagnostics"))
        printf("steps %" PRIu64 ": sum rho %.9g, sum phi %.9g, max |u| %.6g\n", nsteps, d[0], d[4], d[5]);
    pf_destroy(ctx);
    return bad;
}
