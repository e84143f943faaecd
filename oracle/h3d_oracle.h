/*
 * CPU oracle for the Heat3ds linear-solve path.  TEST INFRASTRUCTURE ONLY.
 *
 * This is a plain-C, FP64, op-for-op restatement of the reference Julia code
 * (kenoogl/H2).  It exists so that tests, the smoke check and the CPU-baseline
 * leg of the benchmark have something to compare the HIP library against.
 * Nothing in the product (h2_amd/) may link, load or call it.
 *
 * Conventions
 *   - Arrays are the reference's Julia column-major (MX,MY,MZ) arrays with a
 *     one-cell halo, i fastest.  0-based offset = i + MX*(j + MY*k), where
 *     0-based (i,j,k) = 1-based (i-1,j-1,k-1).
 *   - Built with -ffp-contract=off: Julia never contracts a*b+c to an FMA, so
 *     every operation here rounds exactly where the reference's does.
 *   - Reductions run in the reference's sequential order (k outer, j, i inner,
 *     left to right), which is what FLoops' SequentialEx produces.
 *   - Element-wise loops may run under OpenMP: they have no cross-cell
 *     dependencies (red-black colouring guarantees that for the smoother), so
 *     results are bitwise independent of the thread count.
 *
 * Parity status of each part (how it is pinned) is recorded in the tests
 * that exercise it; see the per-function comments in h3d_oracle.c.
 */
#ifndef H3D_ORACLE_H
#define H3D_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* boundary types, src/common.jl:20-24 */
enum { ORC_ISOTHERMAL = 0, ORC_HEAT_FLUX = 1, ORC_CONVECTION = 2 };

/* src/boundary_conditions.jl:15-21 */
typedef struct {
    int type;
    double temperature;
    double heat_flux;
    double htc;       /* heat_transfer_coefficient */
    double ambient;   /* ambient_temperature */
} orc_bc_t;

/* Geometry/metrics shared by the operator kernels (Δh, Δt, ZC, ΔZ, HC). */
typedef struct {
    int MX, MY, MZ;
    double dx, dy, dt;
    const double *ZC;   /* length MZ */
    const double *dZ;   /* length MZ */
    double HC[6];
    int steady;
} orc_grid_t;

/* WorkBuffers, src/common.jl:49-89 (15 arrays of MX*MY*MZ doubles) */
typedef struct {
    double *theta, *b, *mask, *rho, *lam, *cp;
    double *p, *p_, *r, *r0, *q, *s, *s_, *t_;
    double *hsrc;
} orc_work_t;

/* ---- host set-up (src/heat3ds.jl, src/Zcoord.jl, src/modelA.jl) ---- */
double orc_pitch(int n);                                   /* heat3ds.jl:197-200 */
int orc_genZ(int nk, double *Z, double *ZC, double *dZ);   /* Zcoord.jl:88-166 */
int orc_fillID(uint8_t *ID, int MX, int MY, int MZ, double dx, double dy,
               const double *Z);                           /* modelA.jl:545-556 */
void orc_set_properties(double *lam, double *rho, double *cp,
                        const uint8_t *ID, size_t n);      /* modelA.jl:499-542 */
void orc_heat_src(double *hsrc, const uint8_t *ID, int MX, int MY, int MZ); /* heat3ds.jl:55-64 */
void orc_mode3_bc(orc_bc_t bc[6]);                         /* heat3ds.jl:27-46 */
void orc_apply_bc(double *theta, double *lam, double *rho, double *cp, double *mask,
                  int MX, int MY, int MZ, const orc_bc_t bc[6]); /* boundary_conditions.jl:94-313 */
void orc_set_bc_coef(const orc_bc_t bc[6], double HC[6]); /* boundary_conditions.jl:360-394 */

/* ---- the solve path (src/RHS.jl, src/NonUniform.jl, src/common.jl) ---- */
double orc_lamf(double a, double b, double ma, double mb); /* common.jl:35 */
void orc_bc2rhs(double *b, const double *theta, const orc_grid_t *g,
                const orc_bc_t bc[6]);                     /* RHS.jl:23-192 */
void orc_cal_rhs(orc_work_t *wk, const orc_grid_t *g, const orc_bc_t bc[6]); /* NonUniform.jl:57-91 */
void orc_calc_ax(double *ax, const double *x, const orc_work_t *wk, const orc_grid_t *g);
double orc_calc_rk(double *r, const double *theta, const double *b,
                   const orc_work_t *wk, const orc_grid_t *g);
double orc_rbsor_core(double *x, const double *b, const orc_work_t *wk,
                      const orc_grid_t *g, double omega, int color);
double orc_res_sor(const double *x, const double *b, const orc_work_t *wk, const orc_grid_t *g, double omega);
double orc_rbsor(double *x, const double *b, const orc_work_t *wk, const orc_grid_t *g, double omega);
/* sor! (NonUniform.jl:762-826): one lexicographic SOR sweep in place */
double orc_sor(double *x, const double *b, const orc_work_t *wk, const orc_grid_t *g, double omega);
int orc_solve_sor(orc_work_t *wk, const orc_grid_t *g, double omega, double tol, int maxItr, int *itr,
                  double *res0_out, double *hist, int hist_cap);
void orc_precond(double *xx, const double *bb, const orc_work_t *wk,
                 const orc_grid_t *g, int smoother);      /* 0 none, 1 gs */
double orc_fdot1(const double *x, int MX, int MY, int MZ);
double orc_fdot2(const double *x, const double *y, int MX, int MY, int MZ);
void orc_bicg1(double *p, const double *r, const double *q, double beta, double omg,
               int MX, int MY, int MZ);
void orc_triad(double *z, const double *x, const double *y, double a, int MX, int MY, int MZ);
void orc_bicg2(double *z, const double *x, const double *y, double a, double b,
               int MX, int MY, int MZ);
void orc_fill(double *a, double v, size_t n);
void orc_copy(double *dst, const double *src, size_t n);

/* Krylov drivers.  Return isconverged (0/1); *itr, *res0 as in the reference;
 * hist[k-1] = relative residual printed at iteration k (cap entries). */
int orc_pbicgstab(orc_work_t *wk, const orc_grid_t *g, double tol, int maxItr,
                  int smoother, int verbose, int *itr, double *res0,
                  double *hist, int cap);                  /* NonUniform.jl:248-350 */
int orc_cg(orc_work_t *wk, const orc_grid_t *g, double tol, int maxItr,
           int smoother, int verbose, int *itr, double *res0,
           double *hist, int cap);                         /* NonUniform.jl:115-226 */

/* Grid-sum mode: 0 = the reference's sequential (k,j,i) running sum
 * (default), 1 = exactly rounded sum of the same terms (order-independent),
 * 2 = per-thread partial sums combined by OpenMP, the threaded reference
 * path's shape (order unpinned; timing only). */
void orc_set_sum_mode(int mode);
int orc_get_sum_mode(void);
double orc_exact_sum(const double *t, size_t n);

/* Number of OpenMP threads the element-wise loops use (0 = runtime default). */
void orc_set_threads(int n);
int orc_get_threads(void);

#ifdef __cplusplus
}
#endif
#endif
