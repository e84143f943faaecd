/*
 * Oracle command-line driver: the reference's q3d() set-up and solve
 * (src/heat3ds.jl:169-232 and main() 107-160) on the C restatement.
 * TEST INFRASTRUCTURE ONLY (CPU baseline and fixture generation).
 *
 *   orc_q3d NX NY NZ solver smoother eps [steady=1] [threads=0] [out.bin]
 *   (environment ORC_MAXITR caps the iteration count, default 20000;
 *    ORC_SUM_MODE selects the grid sums: 0 the reference's sequential order,
 *    1 exactly rounded, 2 OpenMP partial sums as the threaded path)
 *
 * Prints the reference's log lines ("Initial residual", "%10d %24.14E",
 * "Converged at"), theta min/max over the interior and the solve time of
 * calRHS! + the Krylov solve (what tm = @elapsed main(...) covers minus the
 * log/plot I/O).  With out.bin, writes theta (MX*MY*MZ doubles).
 */
#include "h3d_oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

static double now(void)
{
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return ts.tv_sec + 1e-9 * ts.tv_nsec;
}

int main(int argc, char **argv)
{
    if (argc < 7) {
        fprintf(stderr, "usage: %s NX NY NZ solver smoother eps [steady] [threads] [out.bin]\n", argv[0]);
        return 2;
    }
    int NX = atoi(argv[1]), NY = atoi(argv[2]), NZ = atoi(argv[3]);
    const char *solver = argv[4], *smoother = argv[5];
    double eps = atof(argv[6]);
    int steady = argc > 7 ? atoi(argv[7]) : 1;
    int threads = argc > 8 ? atoi(argv[8]) : 0;
    const char *out = argc > 9 ? argv[9] : NULL;
    orc_set_threads(threads);
    const char *sum_mode = getenv("ORC_SUM_MODE");  /* 0 sequential (default), 1 exact, 2 threaded */
    if (sum_mode) orc_set_sum_mode(atoi(sum_mode));

    int MX = NX + 2, MY = NY + 2, MZ = NZ + 2;
    size_t n = (size_t)MX * MY * MZ;
    double dx = orc_pitch(NX), dy = orc_pitch(NY);
    double *Z = calloc(MZ + 1, sizeof(double)), *ZC = calloc(MZ, sizeof(double)), *dZ = calloc(MZ, sizeof(double));
    if (orc_genZ(NZ, Z, ZC, dZ)) { fprintf(stderr, "genZ: NZ must be a multiple of 30\n"); return 1; }
    uint8_t *ID = calloc(n, 1);
    double *arr[15];
    for (int a = 0; a < 15; ++a) arr[a] = calloc(n, sizeof(double));
    orc_work_t wk = {arr[0], arr[1], arr[2], arr[3], arr[4], arr[5], arr[6], arr[7],
                     arr[8], arr[9], arr[10], arr[11], arr[12], arr[13], arr[14]};
    for (size_t c = 0; c < n; ++c) { wk.mask[c] = 1.0; wk.lam[c] = 1.0; wk.cp[c] = 1.0; }

    double t0 = now();
    if (orc_fillID(ID, MX, MY, MZ, dx, dy, Z)) { fprintf(stderr, "fillID failed\n"); return 1; }
    orc_set_properties(wk.lam, wk.rho, wk.cp, ID, n);
    double t_geom = now() - t0;
    orc_bc_t bc[6];
    orc_mode3_bc(bc);
    for (size_t c = 0; c < n; ++c) wk.theta[c] = 300.0;
    orc_apply_bc(wk.theta, wk.lam, wk.rho, wk.cp, wk.mask, MX, MY, MZ, bc);

    orc_heat_src(wk.hsrc, ID, MX, MY, MZ);
    orc_grid_t g = {MX, MY, MZ, dx, dy, 10000.0, ZC, dZ, {0}, steady};
    orc_set_bc_coef(bc, g.HC);

    int sm = strcmp(smoother, "gs") == 0 ? 1 : 0;
    const char *mi = getenv("ORC_MAXITR");   /* bounded runs for the CPU baseline */
    int maxItr = mi ? atoi(mi) : 20000;
    if (maxItr < 1) maxItr = 1;
    int cap = maxItr, itr = 0;
    double res0 = 0.0, *hist = calloc(cap, sizeof(double));
    double t1 = now();
    orc_cal_rhs(&wk, &g, bc);
    int conv;
    if (strcmp(solver, "cg") == 0)
        conv = orc_cg(&wk, &g, eps, maxItr, sm, 1, &itr, &res0, hist, cap);
    else
        conv = orc_pbicgstab(&wk, &g, eps, maxItr, sm, 1, &itr, &res0, hist, cap);
    double t_solve = now() - t1;

    double mn = 1e300, mx = -1e300, l2 = 0.0;
    for (int k = 1; k < MZ - 1; ++k)
        for (int j = 1; j < MY - 1; ++j)
            for (int i = 1; i < MX - 1; ++i) {
                double v = wk.theta[i + (size_t)MX * (j + (size_t)MY * k)];
                mn = v < mn ? v : mn;
                mx = v > mx ? v : mx;
                l2 += v * v;
            }
    printf("converged=%d itr=%d res0=%.17g\n", conv, itr, res0);
    printf("theta_min=%.17g theta_max=%.17g l2=%.17g\n", mn, mx, sqrt(l2));
    printf("geometry_s=%.3f solve_s=%.3f threads=%d\n", t_geom, t_solve, orc_get_threads());
    if (out) {
        FILE *f = fopen(out, "wb");
        fwrite(wk.theta, sizeof(double), n, f);
        fclose(f);
    }
    return 0;
}
