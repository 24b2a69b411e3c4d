/* A C11 program that drives the engine through the C-ABI alone (include/rkdg.h, librkdg.so):
 * generates a 20x20-quad mesh with land walls over seeded Fourier bathymetry, creates a context
 * per precision, steps a lake at rest (ze = 0.5, q = 0; p = 2, SSP-RK3) and checks that it stays
 * at rest, as test_lake_at_rest checks the oracle.  Exit status: 0 at rest in every precision,
 * 2 when no HIP device is visible (rkdg_create says so), 1 on any other failure. */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "rkdg.h"

static int run(const rkdg_mesh_desc* desc, rkdg_precision precision, const char* name, double tol_ze, double tol_q) {
    rkdg_params prm;
    memset(&prm, 0, sizeof prm);
    prm.p = 2, prm.nstages = 3, prm.order = 3, prm.dt = 1.0, prm.g = 9.81;
    prm.problem = RKDG_PROBLEM_NONE, prm.ic_constant = 1, prm.ze0 = 0.5;
    prm.h_o = 0.1, prm.sl_M = 1e-8, prm.sl_nu = 1.5;
    rkdg_ctx* ctx = NULL;
    rkdg_status s = rkdg_create(desc, &prm, precision, &ctx);
    if (s == RKDG_ERR_NO_DEVICE) {
        printf("%s: no HIP device (%s)\n", name, rkdg_create_error());
        return 2;
    }
    if (s != RKDG_OK) {
        printf("%s: rkdg_create failed %d: %s\n", name, (int)s, rkdg_create_error());
        return 1;
    }
    const int ne = desc->n_elements, nd = 6;
    double* q0 = malloc(sizeof(double) * 3 * nd * ne);
    double* q = malloc(sizeof(double) * 3 * nd * ne);
    int rc = 1;
    if (!q0 || !q) goto done;
    if (rkdg_get_state(ctx, q0, NULL, NULL, NULL) != RKDG_OK) goto done;
    if ((s = rkdg_step(ctx, 20)) != RKDG_OK) {
        printf("%s: rkdg_step failed %d: %s\n", name, (int)s, rkdg_last_error(ctx));
        goto done;
    }
    if (rkdg_get_state(ctx, q, NULL, NULL, NULL) != RKDG_OK) goto done;
    double dze = 0.0, mq = 0.0, t = 0.0;
    for (int i = 0; i < nd * ne; ++i) dze = fmax(dze, fabs(q[i] - q0[i]));       /* ze modes */
    for (int i = nd * ne; i < 3 * nd * ne; ++i) mq = fmax(mq, fabs(q[i]));       /* qx, qy modes */
    int32_t step = 0, stage = 0;
    rkdg_get_time(ctx, &t, &step, &stage);
    printf("%s: %d triangles, t = %g after %d steps, max |dze| = %.3e, max |q| = %.3e\n", name, ne, t, step, dze, mq);
    rc = (dze < tol_ze && mq < tol_q && step == 20) ? 0 : 1;
done:
    free(q0), free(q);
    rkdg_destroy(ctx);
    return rc;
}

int main(void) {
    rkdg_quad_mesh_spec spec;
    memset(&spec, 0, sizeof spec);
    spec.x2 = 20000.0, spec.y2 = 20000.0, spec.m = 20, spec.n = 20;
    for (int k = 0; k < 4; ++k) spec.bc[k] = RKDG_BC_LAND;
    spec.bathy_kind = 2, spec.bmin = 5.0, spec.bmax = 400.0, spec.n_modes = 8, spec.jitter = 0.2, spec.seed = 7;
    rkdg_mesh* mesh = NULL;
    rkdg_mesh_desc desc;
    if (rkdg_mesh_generate(&spec, &mesh, &desc) != RKDG_OK) {
        printf("mesh generation failed: %s\n", rkdg_create_error());
        return 1;
    }
    int rc = run(&desc, RKDG_FP64_PLAIN, "fp64 plain", 1e-12, 1e-9);
    if (rc == 0) rc = run(&desc, RKDG_FP64_FAST, "fp64 fast", 1e-12, 1e-9);
    if (rc == 0) rc = run(&desc, RKDG_FP32_BULK, "fp32 bulk", 1e-5, 1e-3);
    rkdg_mesh_free(mesh);
    return rc;
}
