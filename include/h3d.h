/*
 * libh3d — MI355X (gfx950) native linear-solve path of Heat3ds.
 *
 * C ABI: plain pointers and sizes, every function returns an int status
 * (H3D_OK == 0, negative on error; h3d_last_error() gives the message).
 * Nothing here ever exits or aborts.  A context owns device-resident copies
 * of the reference's WorkBuffers arrays and runs all work on its own HIP
 * stream.  One context per host thread; a context is not re-entrant.
 *
 * Array layout everywhere is the reference's Julia Array{Float64,3} of size
 * (MX,MY,MZ) = (NX+2,NY+2,NZ+2), column-major, i fastest, with a one-cell
 * halo: 0-based offset = i + MX*(j + MY*k).  A Julia caller passes
 * pointer(A) directly.
 *
 * Reference interfaces replaced (paths are into the kenoogl/H2 tree):
 *   WorkBuffers            src/common.jl:49-89        -> h3d_create/upload/download
 *   myfill!/mycopy!        src/common.jl:95-113       -> h3d_fill/h3d_copy
 *   apply_boundary_conditions!  src/boundary_conditions.jl:94-119 -> h3d_apply_bc
 *   set_BC_coef            src/boundary_conditions.jl:360-394 -> h3d_set_bc
 *   calRHS!/apply_bc2RHS!  src/NonUniform.jl:57-91, src/RHS.jl:23-192 -> h3d_calc_rhs
 *   PBiCGSTAB!             src/NonUniform.jl:248-350  -> h3d_pbicgstab
 *   CG!                    src/NonUniform.jl:115-226  -> h3d_cg
 *   CalcAX!                src/NonUniform.jl:454-511  -> h3d_calc_ax
 *   CalcRK!                src/NonUniform.jl:372-434  -> h3d_calc_rk
 *   Preconditioner!        src/NonUniform.jl:531-559  -> h3d_precond
 *   rbsor_core!            src/NonUniform.jl:848-915  -> h3d_rbsor_color
 *   Fdot1/Fdot2            src/NonUniform.jl:567-592  -> h3d_fdot1/h3d_fdot2
 *   BiCG1!/Triad!/BICG2!   src/NonUniform.jl:603-659  -> h3d_bicg1/h3d_triad/h3d_bicg2
 */
#ifndef H3D_H
#define H3D_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* status codes */
#define H3D_OK 0
#define H3D_ERR_ARG (-1)         /* bad argument (size, field id, enum) */
#define H3D_ERR_HIP (-2)         /* a HIP runtime call failed */
#define H3D_ERR_NAN (-3)         /* NaN/Inf residual during a solve */
#define H3D_ERR_UNSUPPORTED (-4) /* input the device path does not support */
#define H3D_ERR_STATE (-5)       /* call order (e.g. solve before set_grid) */

/* boundary types: src/common.jl:20-24 */
#define H3D_ISOTHERMAL 0
#define H3D_HEAT_FLUX 1
#define H3D_CONVECTION 2

#define H3D_SOLVER_PBICGSTAB 0
#define H3D_SOLVER_CG 1
#define H3D_SMOOTHER_NONE 0 /* Preconditioner! :none (copy) */
#define H3D_SMOOTHER_GS 1   /* Preconditioner! :gs (5 red-black sweeps) */

/* arithmetic of the Krylov iteration */
#define H3D_FP64 0          /* FP64 throughout (the reference) */
#define H3D_FP32_REFINED 1  /* FP32 inner PBiCGSTAB/CG inside FP64 refinement */

/* fields: the WorkBuffers members, src/common.jl:49-65, in that order */
#define H3D_F_THETA 0
#define H3D_F_B 1
#define H3D_F_MASK 2
#define H3D_F_RHO 3
#define H3D_F_LAMBDA 4
#define H3D_F_CP 5
#define H3D_F_PCG_P 6
#define H3D_F_PCG_P_ 7
#define H3D_F_PCG_R 8
#define H3D_F_PCG_R0 9
#define H3D_F_PCG_Q 10
#define H3D_F_PCG_S 11
#define H3D_F_PCG_S_ 12
#define H3D_F_PCG_T_ 13
#define H3D_F_HSRC 14
#define H3D_NFIELDS 15

/* options (h3d_set_option / h3d_get_option) */
#define H3D_OPT_COEF_FORM 0    /* how the operator kernels obtain the face conductivities */
#define H3D_COEF_AUTO 0        /* class tables when λ has at most 32 distinct values, else direct */
#define H3D_COEF_DIRECT 1      /* always the reference's per-face harmonic mean of the λ field */
#define H3D_OPT_NUM_CLASSES 1  /* read-only: distinct λ values in use by the table form (0 = direct) */
#define H3D_OPT_LAYOUT 2       /* device layout of the solver's Krylov vectors */
#define H3D_LAYOUT_AUTO 0      /* colour-split (each red-black colour contiguous) when the table form is active */
#define H3D_LAYOUT_STANDARD 1  /* the reference's column-major layout */
#define H3D_OPT_OPERATOR_FORM 3 /* how the FP64 operator evaluates A·x */
#define H3D_FORM_REFERENCE 0   /* ss - dd·x0, the reference's expression (bitwise the reference) */
#define H3D_FORM_FLUX 1        /* Σ c_n (x_n - x0) - (Σconv + ap0)·x0: the same operator with rounding
                                  relative to temperature differences instead of temperatures (used for
                                  finite-difference checks of the adjoint) */
#define H3D_OPT_SWEEP_CHUNK 4  /* Gauss-Seidel sweeps of the colour-split layout: 0 = one launch per colour,
                                  C >= 1 = the two colours of a sweep in one launch marching chunks of C
                                  z-planes (C capped at NZ and 125; operators with more than 16 classes
                                  keep one launch per colour), -1 = chosen from the grid (default; one
                                  launch per colour on every grid measured); h3d_cg
                                  runs its last sweep per colour either way (that pass also forms
                                  rho = (r, z)); bitwise identical */
#define H3D_OPT_RESIDENT 6     /* Gauss-Seidel preconditioner of the colour-split layout on grids whose
                                  iterate fits the CUs' LDS (240x240x30 in FP64 and below), used when
                                  H3D_OPT_SWEEP_CHUNK is -1 and the grid fits: 2 (default) = as 1, and
                                  h3d_pbicgstab and h3d_cg run every iteration of a solve in ONE launch
                                  of the resident grid (in-launch grid barriers for their sums; not
                                  with H3D_SUM_SEQUENTIAL); 1 = all five sweeps in one launch of
                                  co-resident workgroups that keep the iterate in LDS (three launches
                                  per PBiCGSTAB iteration); 0 = one launch per colour pass; bitwise
                                  identical */
#define H3D_OPT_RESIDENT_WGS 7 /* read-only: workgroups of the last solve's resident preconditioner
                                  launches (0 when it did not use one) */
#define H3D_OPT_SUM_ORDER 5    /* order of the Krylov solvers' grid sums (dots, residual norms) */
#define H3D_SUM_EXACT 0        /* exactly rounded, independent of order (default; parallel, fast) */
#define H3D_SUM_SEQUENTIAL 1   /* the reference's sequential order: one running FP64 sum over the
                                  interior cells, k outer, then j, then i (Fdot1/Fdot2/CalcRK! under
                                  FLoops' SequentialEx, src/NonUniform.jl:372-434, 567-592, with
                                  src/common.jl:42), so a solve reproduces the reference's sequential
                                  run bit for bit; each sum runs on one workgroup (milliseconds per
                                  sum at 240x240x30), for parity runs, not speed.  The RB-SOR
                                  solver's and the sensitivity extension's sums stay exact. */

typedef struct h3d_ctx h3d_ctx;

/* WorkBuffers(mx,my,mz) (src/common.jl:71-89): all fields zero except
 * mask = lambda = cp = 1.  device = HIP device ordinal. */
int h3d_create(h3d_ctx **ctx, int MX, int MY, int MZ, int device);
int h3d_destroy(h3d_ctx *ctx);
const char *h3d_last_error(const h3d_ctx *ctx);
/* message of the last h3d_create failure (no context exists then) */
const char *h3d_create_error(void);

/* Both coefficient forms give bitwise identical operators; the table form
 * (a one-byte class per cell and a λf table per class pair) moves 1 byte
 * per cell instead of 8 and drops the per-face division.  Both layouts give
 * bitwise identical solves: the colour-split layout only changes where the
 * Krylov vectors' elements live (fields keep the reference's layout).  No
 * reference counterpart: the reference has one form and one layout. */
int h3d_set_option(h3d_ctx *ctx, int option, int value);
int h3d_get_option(const h3d_ctx *ctx, int option, int *value);

/* Δh[1], Δh[2] and the Z metrics ZC, ΔZ (both MZ entries, genZ! layout,
 * src/Zcoord.jl:133-166). */
int h3d_set_grid(h3d_ctx *ctx, double dx, double dy, const double *ZC, const double *dZ, int mz);

/* The six faces x-,x+,y-,y+,z-,z+ (BoundaryConditionSet,
 * src/boundary_conditions.jl:24-31).  Also forms HC[6] as set_BC_coef does. */
int h3d_set_bc(h3d_ctx *ctx, const int type[6], const double temperature[6],
               const double heat_flux[6], const double htc[6], const double ambient[6]);
int h3d_get_hc(const h3d_ctx *ctx, double HC[6]);

/* Host <-> device copies of a whole field (n must equal MX*MY*MZ).
 * The mask is not stored: the device path requires the mask the
 * reference's apply_boundary_conditions! leaves (0 on every halo cell, 1
 * inside); uploading any other mask returns H3D_ERR_UNSUPPORTED.  Before
 * h3d_apply_bc the reference mask is 1 everywhere; the operator kernels
 * always use the post-boundary-condition mask. */
int h3d_upload(h3d_ctx *ctx, int field, const double *host, size_t n);
int h3d_download(h3d_ctx *ctx, int field, double *host, size_t n);
int h3d_fill(h3d_ctx *ctx, int field, double value);           /* myfill! */
int h3d_copy(h3d_ctx *ctx, int dst, int src);                   /* mycopy! */

/* apply_boundary_conditions! on the device copies of θ, λ, ρ, cp. */
int h3d_apply_bc(h3d_ctx *ctx);

/* calRHS!(wk, Δh, Δt, ΔZ, bc_set, qsrf, par; is_steady): adds the heat-flux
 * and convection face terms into b, then b = -(ap0*θ + hsrc*dx*dy*dz + b).
 * Like the reference it does not clear b first (h3d_fill(ctx,H3D_F_B,0)). */
int h3d_calc_rhs(h3d_ctx *ctx, double dt, int is_steady);

/* PBiCGSTAB!/CG!(wk, Δh, Δt, ZC, ΔZ, HC; tol, maxItr, smoother, is_steady).
 * Outputs mirror the reference's (isconverged, itr, res0) tuple.  Non-
 * convergence and Krylov breakdown are not errors (isconverged = 0);
 * a NaN/Inf residual returns H3D_ERR_NAN.  precision = H3D_FP64 runs the
 * reference iteration; H3D_FP32_REFINED runs FP32 inner solves (relative
 * tolerance inner_tol, 1e-4 if <= 0) inside an FP64 refinement loop that
 * stops when the true FP64 ||b - A θ|| / res0 < tol; then *itr counts
 * inner iterations and the history holds the outer true residuals.  The
 * refinement also reports convergence when the true residual is within
 * 9u||s|| (u = 2^-53, s = |b| + Σ|terms of Aθ| per cell), the error bound
 * of its own FP64 evaluation (h3d_last_solve_floor then gives 1), and it
 * stops unconverged after two corrections in a row that did not halve it,
 * or after a correction that raised it, which it takes back (θ is then the
 * iterate whose residual is the history's last entry).
 * No counterpart in the reference: its solvers test the recursive residual. */
int h3d_pbicgstab(h3d_ctx *ctx, double dt, double tol, int maxItr, int smoother,
                  int is_steady, int precision, double inner_tol,
                  int *isconverged, int *itr, double *res0);
int h3d_cg(h3d_ctx *ctx, double dt, double tol, int maxItr, int smoother,
           int is_steady, int precision, double inner_tol,
           int *isconverged, int *itr, double *res0);

/* Relative residuals of the last solve (what the reference prints as
 * "%10d %24.14E" per iteration).  *n = number available. */
int h3d_get_history(const h3d_ctx *ctx, double *res, int cap, int *n);
/* Device time of the last h3d_pbicgstab/h3d_cg call (HIP events on the
 * context stream, initial residual to final residual). */
int h3d_last_solve_ms(const h3d_ctx *ctx, double *ms);
/* 1 when the last H3D_FP32_REFINED solve converged by reaching the rounding
 * floor of its FP64 residual rather than tol, else 0.  No counterpart in
 * the reference. */
int h3d_last_solve_floor(const h3d_ctx *ctx, int *at_floor);

/* Single operators on fields, for unit tests and custom drivers. */
int h3d_calc_ax(h3d_ctx *ctx, int ax, int x, double dt, int is_steady);
int h3d_calc_rk(h3d_ctx *ctx, int r, int theta, int b, double dt, int is_steady, double *res);
/* ||s||_2 with s = |b| + Σ|terms of Aθ| per interior cell, the terms in the
 * current operator form: 2^-53·||s|| is the scale of the rounding error of
 * the FP64 residual b - Aθ (the refinement's floor test).  No counterpart
 * in the reference. */
int h3d_residual_scale(h3d_ctx *ctx, int theta, int b, double dt, int is_steady, double *scale);
int h3d_rbsor_color(h3d_ctx *ctx, int x, int b, int color, double dt, int is_steady);
int h3d_precond(h3d_ctx *ctx, int xx, int bb, int smoother, double dt, int is_steady);
int h3d_fdot1(h3d_ctx *ctx, int x, double *out);
int h3d_fdot2(h3d_ctx *ctx, int x, int y, double *out);
int h3d_bicg1(h3d_ctx *ctx, int p, int r, int q, double beta, double omg);
int h3d_triad(h3d_ctx *ctx, int z, int x, int y, double a);
int h3d_bicg2(h3d_ctx *ctx, int z, int x, int y, double a, double b);

/* Sensitivity extension (no counterpart in the reference): a smooth maximum
 * of θ, its adjoint, and the gradients the adjoint gives.
 *
 * h3d_objective: J = m + log(Σ_R exp(β(θ_c - m)))/β over the interior cells
 * whose region byte is non-zero (region: MX*MY*MZ bytes, reference layout),
 * m = max_R θ; keeps the seed ∂J/∂θ (the softmax weights) on the device.
 * H3D_ERR_ARG when the region has no interior cell. */
int h3d_objective(h3d_ctx *ctx, const uint8_t *region, size_t n, double beta, double *J);
/* Solves Aᵀ μ = ∂J/∂θ for the adjoint μ (A is the operator of the last
 * forward solve's dt and is_steady; it is symmetric, so the same PBiCGSTAB!
 * or CG! runs on it with a zero halo).  θ and b are left as they were.
 * Outputs as h3d_pbicgstab. */
int h3d_adjoint_solve(h3d_ctx *ctx, int solver, double dt, double tol, int maxItr, int smoother, int is_steady,
                      int *isconverged, int *itr, double *res0);
/* Per-cell gradients of J for the current θ and μ (MX*MY*MZ each, zero on
 * the halo): dJ/dλ (including the isothermal halo copy of λ) and dJ/dhsrc.
 * Either pointer may be NULL. */
int h3d_gradients(h3d_ctx *ctx, double *dJ_dlambda, double *dJ_dhsrc, size_t n);
/* The adjoint μ and the seed ∂J/∂θ (either may be NULL). */
int h3d_get_adjoint(h3d_ctx *ctx, double *mu, double *seed, size_t n);

/* One step of the adjoint of an implicit-Euler run (no reference
 * counterpart; the forward step is calRHS! then PBiCGSTAB!/CG! with
 * is_steady = false, src/heat3ds.jl:124-139).  For
 * J_total = Σ_n w_n J(θⁿ) and θⁿ solving A θⁿ = b(θⁿ⁻¹), the adjoint runs
 * backwards from μ_{N+1} = 0:
 *     Aᵀ μ_n = w_n ∂J/∂θ(θⁿ) + D μ_{n+1},
 *     D_c = ∂b_c/∂θ_c^old = -ap0_c + Σ_{convective faces f of c} ±h_f·area_f
 * (+ on x+/y+/z+ faces, - on x-/y-/z- faces: calRHS! subtracts the minus
 * faces' h·area·(θ∞ - θ) and adds the plus faces', src/RHS.jl).
 * Call h3d_objective on θⁿ first (it sets the seed); chain = 0 for the last
 * step n = N, 1 for the others (μ_{n+1} is the adjoint the previous call
 * left).  precision as h3d_pbicgstab.  θ and b are left as they were. */
int h3d_adjoint_step(h3d_ctx *ctx, int solver, double dt, double tol, int maxItr, int smoother, int precision,
                     double weight, int chain, int *isconverged, int *itr, double *res0);
/* Adds w_lambda·dJ/dλ and w_hsrc·dJ/dhsrc of the current θ and μ (the
 * per-cell gradients of h3d_gradients) to device accumulators; with
 * hsrc_n = f_n·hsrc_base, w_hsrc = f_n gives dJ_total/dhsrc_base. */
int h3d_gradients_accumulate(h3d_ctx *ctx, double w_lambda, double w_hsrc);
/* Copies the accumulators out (either pointer may be NULL; zeros before
 * the first accumulation) and, when reset != 0, clears them. */
int h3d_gradients_accumulated(h3d_ctx *ctx, double *dJ_dlambda, double *dJ_dhsrc, size_t n, int reset);

/* Device-resident field copies (checkpoints; no reference counterpart).
 * h3d_store_field: slot := field (allocated on first use, 0 <= slot <
 * H3D_MAX_SLOTS).  h3d_load_field: field := scale·slot (scale 1 copies
 * bit for bit; the time-varying heat source is a scaled base field).
 * h3d_release_slots frees them all.  The mask field has no storage. */
#define H3D_MAX_SLOTS 4096
int h3d_store_field(h3d_ctx *ctx, int field, int slot);
int h3d_load_field(h3d_ctx *ctx, int field, int slot, double scale);
int h3d_release_slots(h3d_ctx *ctx);
/* Device memory in use (all processes) and the device's total, in bytes
 * (hipMemGetInfo on the context's device). */
int h3d_mem_info(h3d_ctx *ctx, size_t *used, size_t *total);

/* Stand-alone red-black SOR (FP64, the reference's operator form).
 * h3d_res_sor: resSOR (src/NonUniform.jl:680-744), sqrt(Σ r²) over every
 *   interior cell of r = (dd + ω(cond_xm + cond_ym + cond_zm))·dp/ω,
 *   dp = (ss - b)/dd - x, without updating x.
 * h3d_rbsor: rbsor! (src/NonUniform.jl:936-957), colour 0 then colour 1
 *   of rbsor_core! with relaxation ω in place on x; *res = sqrt of the
 *   summed residual squares of both colours.
 * h3d_rbsor_solve: solveSOR! (src/NonUniform.jl:1017-1050) on θ and b:
 *   res0 = resSOR (1 when it is 0), then rbsor! until its residual / res0
 *   < tol or maxItr iterations; outputs and history as h3d_pbicgstab.
 * The sums are exactly rounded (the reference's are sequential or
 * threaded). */
int h3d_res_sor(h3d_ctx *ctx, int x, int b, double omega, double dt, int is_steady, double *res);
int h3d_rbsor(h3d_ctx *ctx, int x, int b, double omega, double dt, int is_steady, double *res);
int h3d_rbsor_solve(h3d_ctx *ctx, double dt, double omega, double tol, int maxItr, int is_steady,
                    int *isconverged, int *itr, double *res0);
/* sor! (src/NonUniform.jl:762-826): one lexicographic point-SOR sweep in
 * place on x (cells in the loop order k, j, i, each seeing the new values
 * of the cells before it), computed as a wavefront over the planes
 * i + j + k = t (one launch per plane, NX+NY+NZ-2 launches), so bitwise the
 * sequential sweep; *res = sqrt(Σ r²), r = (dd + ω(cond_xm + cond_ym +
 * cond_zm))·dp/ω, exactly rounded (or in the reference's order with
 * H3D_SUM_SEQUENTIAL). */
int h3d_sor(h3d_ctx *ctx, int x, int b, double omega, double dt, int is_steady, double *res);

/* Timing of one kernel family on the context stream (HIP events around
 * reps back-to-back launches after one warm-up launch; *ms = mean per
 * launch).  which: 0 = one full red-black GS sweep (both colours),
 * 1 = CalcAX!, 2 = CalcAX! with its fused dot; 3 and 4 = the FP32 operator's
 * sweep and CalcAX! (the kernels of the H3D_FP32_REFINED inner solve);
 * 5 and 6 = one full sweep as one two-colour launch, FP64 and FP32, with
 * the chunk H3D_OPT_SWEEP_CHUNK sets (H3D_ERR_ARG when it sets none, as the
 * default does, or the operator has none); 7 and 8 =
 * the whole preconditioner (five sweeps from zero) as one resident launch,
 * FP64 and FP32 (H3D_OPT_RESIDENT; H3D_ERR_ARG when the grid does not fit);
 * 9 and 10 = PBiCGSTAB!'s first resident launch of an iteration (BiCG1!,
 * the preconditioner, CalcAX! and its (q, r0) partials), FP64 and FP32;
 * 11 = PBiCGSTAB!'s one-launch loop (FP64) run for `reps` iterations with
 * tolerance 0 from θ = 0 on the current b (a scratch θ: the fields are left
 * as they are), *ms = the launch's time per iteration run. */
int h3d_time_kernel(h3d_ctx *ctx, int which, int reps, double dt, int is_steady, double *ms);

/* Fdot1/Fdot2 on caller-owned device arrays without a context and without
 * any device allocation.  No counterpart in the reference: it is for
 * callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).
 * x, y: device arrays of MX*MY*MZ doubles in the layout above (y NULL:
 * Σ x², as Fdot1); *out (a device double) receives Σ x·y over the interior,
 * exactly rounded, bitwise what h3d_fdot1/h3d_fdot2 return for the same
 * arrays.  stream: a hipStream_t (NULL: the null stream); the call only
 * enqueues work, so *out is ready when the stream reaches it.  The scratch
 * memory is one fixed device array per GPU (787 KB, held until the process
 * ends; its address is looked up on the first call on each device, so make
 * that call outside graph capture); it serves interiors with
 * ceil(NX/64)*ceil(NY/4) <= 16384 (H3D_ERR_UNSUPPORTED beyond).  Calls
 * that overlap on one GPU (several streams or host threads) share it and
 * race; calls on one stream are ordered.  Returns H3D_ERR_ARG for null
 * pointers or extents below 3, H3D_ERR_HIP when a launch fails. */
int h3d_fdot2_noalloc(const double *x, const double *y, int MX, int MY, int MZ, double *out, void *stream);

#ifdef __cplusplus
}
#endif
#endif
