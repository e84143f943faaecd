// Internal declarations shared by the HIP kernels (h3d_kernels.hip) and the
// host-side C ABI (h3d_api.hip).  Not part of the public interface.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

namespace h3d {

// Status of a device-resident Krylov iteration (KState::status).
enum : int {
    ST_RUNNING = 0,
    ST_CONVERGED = 1,   // res < tol, or res0 == 0
    ST_BREAKDOWN = 2,   // |ρ|, |(t,t)|, |(p,q)| below FloatMin (src/common.jl:12)
    ST_MAXITER = 3,     // loop ran maxItr iterations without converging
    ST_NAN = 4,         // NaN/Inf residual (guard the reference lacks)
};

// Finalisers run by the last workgroup of a reducing kernel on the summed
// partials s[0..2].  Each one restates one scalar step of PBiCGSTAB!
// (src/NonUniform.jl:287-347) or CG! (src/NonUniform.jl:139-224).
enum : int {
    FIN_STORE = 0,       // out[0..ns-1] = s
    FIN_RES0,            // res0 = sqrt(s0); converged at 0 if res0 == 0
    FIN_RES0_BICG,       // FIN_RES0, then FIN_BICG_RHO with ρ = s0 (r0 == r at iteration 1)
    FIN_BICG_RHO,        // itr += 1; rho = s0; breakdown; beta
    FIN_BICG_ALPHA,      // alpha = rho / s0
    FIN_BICG_OMEGA,      // breakdown on s0 = (t,t); omega = s1 / s0
    FIN_BICG_RES,        // res = sqrt(s0)/res0; history; tol / maxItr; rho_old = rho
    FIN_BICG_RES_RHO,    // FIN_BICG_RES on s0, then FIN_BICG_RHO on s1 for the next iteration
    FIN_CG_RHO0,         // rho_old = s0
    FIN_CG_ALPHA,        // itr += 1; breakdown on s0 = (p,q); alpha = rho_old / s0
    FIN_CG_RES,          // res = sqrt(s0)/res0; history; tol
    FIN_CG_BETA,         // breakdown on s0 = rho_new; beta; rho_old = rho_new; maxItr
};

struct KState {
    double rho, rho_old, alpha, omega, beta, res0, res, tol;
    double out[4];          // FIN_STORE destination
    int itr, status, maxItr, pad;
};

// Scratch for cross-workgroup reductions: per-block partial sums (up to 3
// per block, each a double-double hi/lo pair) and the arrival counter of the
// last-block pattern.
struct RedBuf {
    double *partials;       // [nblocks_max][3][2]
    unsigned int *counter;  // reset to 0 by the last block
    int nblocks_max;
};

// Per-solve operator constants.  T is the storage/arithmetic type of the
// stencil (double for the reference iteration, float for the FP32 inner
// solve of the refinement loop).
template <typename T>
struct Geo {
    int MX, MY, MZ;         // stored extents (with halo)
    int NX, NY, NZ;         // interior extents
    T dx0, dy0, ddx, ddy;
    T HC[6];
    const T *ZC;            // device, MZ entries
    const T *dZ;            // device, MZ entries
    const T *ap0;           // device, per-cell ρ·cp·dx·dy·dz/Δt (transient); nullptr when steady
};

// Launch helpers implemented in h3d_kernels.hip.  All take the stream and
// return hipError_t of the launch.
template <typename T> hipError_t launch_fill(T *a, T v, size_t n, hipStream_t s);
template <typename T> hipError_t launch_copy_interior(T *dst, const T *src, const Geo<T> &g, const KState *st, hipStream_t s);
template <typename T>
hipError_t launch_calc_ax(T *ax, const T *x, const T *lam, const Geo<T> &g,
                          const T *dot_y, const T *dot_z, int fin, KState *st, double *hist,
                          const RedBuf &rb, hipStream_t s);
template <typename T>
hipError_t launch_calc_rk(T *r, T *r0copy, const T *th, const T *b, const T *lam, const Geo<T> &g,
                          int fin, KState *st, double *hist, const RedBuf &rb, hipStream_t s);
template <typename T>
hipError_t launch_rbsor(T *x, const T *b, const T *lam, const Geo<T> &g, int color,
                        const KState *st, hipStream_t s);
template <typename T>
hipError_t launch_dot(const T *x, const T *y, const T *z, const Geo<T> &g, int fin, KState *st,
                      double *hist, const RedBuf &rb, hipStream_t s);
template <typename T>
hipError_t launch_axpby(int op, T *z, const T *x, const T *y, const T *w, double a, double b,
                        const Geo<T> &g, const KState *st, hipStream_t s);
template <typename T>
hipError_t launch_bicg_p(T *p, const T *r, const T *q, const Geo<T> &g, const KState *st, hipStream_t s);
template <typename T>
hipError_t launch_bicg_s(T *s_, const T *q, const T *r, const Geo<T> &g, const KState *st, hipStream_t s);
template <typename T>
hipError_t launch_bicg_update(T *th, T *r, const T *p_, const T *s_, const T *t, const T *sv, const T *r0,
                              const Geo<T> &g, KState *st, double *hist, const RedBuf &rb, hipStream_t s);
template <typename T>
hipError_t launch_cg_update(T *th, T *r, const T *p, const T *q, const Geo<T> &g, KState *st,
                            double *hist, const RedBuf &rb, hipStream_t s);
template <typename T>
hipError_t launch_cg_p(T *p, const T *z, const Geo<T> &g, const KState *st, hipStream_t s);

// axpby ops (interior only), src/NonUniform.jl:603-659
enum : int {
    AX_BICG1 = 0,   // z = x + a*(z - b*y)          BiCG1!(p=z, r=x, q=y, beta=a, omg=b)
    AX_TRIAD = 1,   // z = a*x + y                  Triad!
    AX_BICG2 = 2,   // z = z + (a*x + b*y)          BICG2!
};

// boundary-condition set-up and RHS (h3d_kernels.hip)
hipError_t launch_apply_bc_face(int face, int type, double temperature, double ambient,
                                double *th, double *lam, double *rho, double *cp,
                                int MX, int MY, int MZ, hipStream_t s);
hipError_t launch_bc2rhs_face(int face, int type, double q, double h, double ta, double *b,
                              const double *th, const Geo<double> &g, hipStream_t s);
hipError_t launch_rhs_final(double *b, const double *th, const double *hsrc, const Geo<double> &g,
                            hipStream_t s);
hipError_t launch_ap0(double *ap0, const double *rho, const double *cp, const Geo<double> &g,
                      double ddt, hipStream_t s);
hipError_t launch_check_mask(const double *mask, int MX, int MY, int MZ, unsigned int *bad, hipStream_t s);
template <typename D, typename S>
hipError_t launch_convert(D *dst, const S *src, size_t n, hipStream_t s);
// θ += d (FP64 += FP32) on the interior, refinement update
hipError_t launch_add_correction(double *th, const float *d, const Geo<double> &g, hipStream_t s);
// b32 = (float) r64 on the whole array
}  // namespace h3d
