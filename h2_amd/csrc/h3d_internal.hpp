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
    ST_DIVERGED = 5,    // guarded solve only: residual far above its minimum, NaN, or no new minimum for long
};

// Guarded solves (the FP32 inner solves of the refinement loop) stop when the
// relative residual exceeds kDivergeFactor times the smallest one seen, or
// when kStallIters iterations pass without a new minimum; the iterate with
// the smallest residual is kept (see KState::improved).
constexpr double kDivergeFactor = 1.0e4;
constexpr int kStallIters = 500;

// Scalar steps run by the finalize kernel once the grid sums s[0..2] of the
// preceding reducing kernel are known.  Each one restates one scalar step of
// PBiCGSTAB! (src/NonUniform.jl:287-347) or CG! (src/NonUniform.jl:139-224).
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
    FIN_SOR_RES0,        // res0 = sqrt(s0), 1 when 0 (solveSOR!); itr = 0
    FIN_SOR_RES,         // itr += 1; res = sqrt(out[0] + s0)/res0; history; tol / maxItr
};

struct KState {
    double rho, rho_old, alpha, omega, beta, res0, res, tol;
    double res_min;         // guarded solves: smallest relative residual so far (1 = the start)
    double out[4];          // FIN_STORE destination
    int itr, status, maxItr;
    int guard;              // 1: divergence guard and best-iterate tracking on
    int improved;           // guarded: the last residual was a new minimum, so the next update
                            // kernel copies the iterate it is about to overwrite into the best-iterate buffer
    int since_best;         // guarded: iterations since the last new minimum
};

// Per-workgroup partial sums of a reducing kernel (up to three sums per
// workgroup, each a double-double hi/lo pair), read by the finalize kernel.
// More than kFinalizeMax partials are first pre-reduced 1024 to one into
// partials2 by a many-workgroup kernel.
struct RedBuf {
    double *partials;       // [nblocks_max][3][2]
    double *partials2;      // [nblocks_max/1024 + 1][3][2]
    int nblocks_max;
    int seq;                // 1: the finalize steps use reference-order sums (launch_finalize_sums)
};
constexpr unsigned kFinalizeMax = 4096;

// Terms of up to two grid sums for the reference-order mode: sum s adds
// x[s][c]·(y[s] ? y[s][c] : x[s][c]) over the interior cells c in the
// reference's loop order (k outer, then j, then i), one running FP64 sum
// from 0.0, as Fdot1/Fdot2 and CalcRK! sum under FLoops' SequentialEx
// (src/NonUniform.jl:372-434, 567-592; src/common.jl:42).
template <typename T>
struct SeqTerms {
    const T *x[2], *y[2];
    int NX, NY, NZ, MX, MY, HX;
    size_t HS;              // 0: the reference's layout; else the colour-split one with this half size
};

// Per-solve operator constants.  T is the storage/arithmetic type of the
// stencil (double for the reference iteration, float for the FP32 inner
// solve of the refinement loop).
//
// Conductivities reach the kernels in one of two equivalent forms:
//   cls == nullptr : the λ field, harmonic means formed per face
//                    (the reference's on-the-fly form);
//   cls != nullptr : a one-byte class per cell (its index among the K
//                    distinct λ values) and tab[m][a][b] = λf(v_a, v_b, m, 1)
//                    for neighbour class a, centre class b, neighbour mask m.
//                    Bitwise the same values, without the per-face division.
// flux selects how A·x is formed: 0 = ss - dd·x0 (the reference's
// expression, FP64); 1 = Σ c_n (x_n - x0) - (Σ conv + ap0)·x0, the same
// operator without the cancellation that makes the reference form useless
// in FP32 for smooth x (used by the FP32 inner solve only).
template <typename T>
struct Geo {
    int MX, MY, MZ;         // stored extents (with halo)
    int NX, NY, NZ;         // interior extents
    T dx0, dy0, ddx, ddy;
    T HC[6];
    const T *ZC;            // device, MZ entries
    // device, MZ-1 z-face distances ZC[k+1] - ZC[k] formed in FP64 and
    // rounded to T (the FP32 operator: the difference of two rounded ZC
    // would carry their absolute rounding error); nullptr for FP64, which
    // forms the differences of ZC as the reference does
    const T *DZC;
    const T *dZ;            // device, MZ entries
    const T *ap0;           // device, per-cell ρ·cp·dx·dy·dz/Δt (transient); nullptr when steady
    const T *lam;           // device λ field (direct form)
    const uint8_t *cls;     // device class field (table form) or nullptr
    const T *tab;           // device [2][K][K] λf table (table form)
    int K;                  // number of classes (table form)
    int flux;               // 1: flux form of A·x (FP32 inner operator)
};

constexpr int kMaxClasses = 32;

// Colour-split ("RB") layout of the solver's Krylov vectors.  Cell
// (i,j,k) of the (MX,MY,MZ) array lives in half h = (i+j+k) & 1 at
//   h*HS + (i >> 1) + HX*(j + MY*k),   HX = (MX+1)/2,  HS = HX*MY*MZ,
// so the cells one Gauss-Seidel colour updates are contiguous and every
// neighbour of a cell is in the other half at the same (j,k) row: the
// i-neighbours at m-1+s and m+s (m = i>>1, s = i&1), the others at m.
// Colour c of rbsor_core! updates half 1-c.  Entries with i >= MX are
// padding and never read.  The operator takes its coefficients from the
// class table form only: the class field in RB layout, the λf table, and a
// per-face table of the z conductances,
//   ZF[kf][a][b] = ((λf(v_a, v_b, m) * dx0) * dy0) / (ZC[kf+1] - ZC[kf]),
// for the face between planes kf and kf+1 (m = 0 on the two boundary faces,
// 1 inside), which is the value CalcAX! forms for the zm term of plane
// kf+1 and the zp term of plane kf (λf is symmetric in its two values);
// in FP32 the distance is the FP64 difference rounded (Geo::DZC).
template <typename T>
struct RbGeo {
    int MX, MY, MZ, NX, NY, NZ;
    int HX;                 // half-row length
    size_t HS;              // half size
    T dx0, dy0, ddx, ddy;
    T HC[6];
    const T *dZ;            // device, MZ entries
    const T *ZF;            // device, (NZ+1)*K*K face conductances
    // transient term ap0 = ((((ρ·cp)·dx)·dy)·ΔZ[k])·ddt formed per class and
    // plane from the FP64 values (the products k_ap0 forms per cell)
    int ap;                 // 1 for a transient operator
    const double *capc;     // device, ρ·cp per class
    const double *dZd;      // device, FP64 ΔZ
    double dx0d, dy0d, ddt;
    const uint8_t *cls;     // device, RB layout class field
    const T *tab;           // device [2][K][K] λf table
    int K;
    int flux;               // as Geo::flux
    // launch grid (rb_shape): one plane per thread, gx*gy*NZ workgroups
    // rounded up to a multiple of 8
    int gx, gy;
    unsigned nb1;
};

// fills the launch grid fields of g
template <typename T> void rb_shape(RbGeo<T> &g);

// Tile plan of the resident Gauss-Seidel preconditioner (h3d_resident.hip):
// the interior rows (j, k) cut into tiles of TJ x TK rows, ntj x ntk of
// them, one workgroup per tile, all co-resident (G <= the CU count).
// G = 0: the grid does not fit and the per-colour launches run.
struct RbRes {
    int TJ = 0, TK = 0, ntj = 0, ntk = 0;
    int G = 0;
    size_t lds = 0;             // dynamic LDS bytes per workgroup
    unsigned *bar = nullptr;    // device synchronisation words (kResBarWords)
    // per-cell coefficients (launch_rb_res_coef), 2*HS entries each: the
    // diagonal and the flux form's sink (elements of the solve's type), and
    // three index words per cell (ix + 2*HS*w, w = 0..2)
    void *dgl = nullptr, *snk = nullptr;
    unsigned *ix = nullptr;
    double *wpart = nullptr;    // partials of sums reduced inside a launch (3 x 4 per workgroup)
};
template <typename T>
struct ResCoef {
    T *dgl, *snk;
    unsigned *ix;
};
// The update fused into PBiCGSTAB!'s second resident launch
template <typename T>
struct ResUpd {
    double *th = nullptr, *xbest = nullptr;  // θ (standard layout) and the guarded solve's best iterate
    T *r = nullptr;                          // r, written
    const T *p_ = nullptr, *r0 = nullptr;
    double *wpart = nullptr;                 // the launch's own partials of (t, t), (t, s)
    const double *apart = nullptr;           // the first launch's partials of (q, r0)
};
// The vectors of the one-launch PBiCGSTAB! loop (launch_rb_bicg_solve):
// the Krylov vectors in the colour-split layout, θ (and the guarded solve's
// best iterate) in the standard one.
template <typename T>
struct ResBicg {
    T *p, *p_, *r, *q, *s_, *t;
    const T *r0;
    double *th, *xbest;
};
// The vectors of the one-launch CG! loop (launch_rb_cg_solve): p, r, q
// and z = M⁻¹r in the colour-split layout, θ (and the guarded solve's best
// iterate) in the standard one.
template <typename T>
struct ResCg {
    T *p, *r, *q, *z;
    double *th, *xbest;
};
constexpr size_t kResLdsMax = 150 * 1024;
constexpr int kResMaxWG = 1024;
// synchronisation words: an error word, then a pass counter per workgroup
// (64 bytes apart), then the grid barrier's; the host zeroes them whenever
// a solve selects a plan
constexpr int kResBarWords = 64 + 16 * kResMaxWG + 16 * 10;
constexpr int kResBarErrWord = 0;

// A grid sum formed inside a launch and its finalize step
struct DotFin {
    int fin;
    KState *st;
    double *hist;
    const RedBuf *rb;
};

// ---- launchers (h3d_kernels.hip); all return hipError_t of the launch -----
// Sum nb workgroup partials of ns sums each and run finalizer fin.
hipError_t launch_finalize_nb(int ns, unsigned nb, int fin, KState *st, double *hist, const RedBuf &rb,
                              hipStream_t s);
// The ns sums of t in the reference's order (one workgroup), then finalizer fin.
template <typename T>
hipError_t launch_finalize_seq(int ns, const SeqTerms<T> &t, int fin, KState *st, double *hist, hipStream_t s);

template <typename T>
inline SeqTerms<T> seq_terms(const Geo<T> &g, const T *x0, const T *y0, const T *x1 = nullptr,
                             const T *y1 = nullptr)
{
    SeqTerms<T> t{};
    t.x[0] = x0; t.y[0] = y0; t.x[1] = x1; t.y[1] = y1;
    t.NX = g.NX; t.NY = g.NY; t.NZ = g.NZ; t.MX = g.MX; t.MY = g.MY;
    t.HX = 0; t.HS = 0;
    return t;
}

template <typename T>
inline SeqTerms<T> seq_terms(const RbGeo<T> &g, const T *x0, const T *y0, const T *x1 = nullptr,
                             const T *y1 = nullptr)
{
    SeqTerms<T> t{};
    t.x[0] = x0; t.y[0] = y0; t.x[1] = x1; t.y[1] = y1;
    t.NX = g.NX; t.NY = g.NY; t.NZ = g.NZ; t.MX = g.MX; t.MY = g.MY;
    t.HX = g.HX; t.HS = g.HS;
    return t;
}

// Finalize step of a reducing launch of nb workgroups: the exactly rounded
// sums of its partials, or, in the reference-order mode (rb.seq), the
// sequential sums of the terms t, re-read from the vectors the launch wrote.
template <typename T>
inline hipError_t launch_finalize_sums(int ns, unsigned nb, int fin, KState *st, double *hist, const RedBuf &rb,
                                       hipStream_t s, const SeqTerms<T> &t)
{
    if (rb.seq) return launch_finalize_seq<T>(ns, t, fin, st, hist, s);
    return launch_finalize_nb(ns, nb, fin, st, hist, rb, s);
}
template <typename T> hipError_t launch_fill(T *a, T v, size_t n, hipStream_t s);
template <typename T>
hipError_t launch_copy_interior(T *dst, const T *src, const Geo<T> &g, const KState *st, hipStream_t s);

// Reducing kernels write per-workgroup partials into rb; a second,
// single-workgroup kernel sums them in a fixed order and runs finalizer fin.
// CalcAX! (src/NonUniform.jl:454-511): ax = A x; with fin >= 0 also
// s0 = Σ ax·(dot_y ? dot_y : ax), s1 = Σ ax·dot_z.
template <typename T>
hipError_t launch_calc_ax(T *ax, const T *x, const Geo<T> &g, const T *dot_y, const T *dot_z, int fin, KState *st,
                          double *hist, const RedBuf &rb, hipStream_t s);
// CalcRK! (src/NonUniform.jl:372-434): r = (b - Aθ)·m, s0 = Σ r², optional r0 copy
template <typename T>
hipError_t launch_calc_rk(T *r, T *r0copy, const T *th, const T *b, const Geo<T> &g, int fin, KState *st,
                          double *hist, const RedBuf &rb, hipStream_t s);
// st->out[0] = Σ s², s = |b| + Σ|terms of Aθ| per cell: ε·sqrt(out[0]) is the
// rounding scale of the computed residual b - Aθ (double only)
template <typename T>
hipError_t launch_rk_scale(const T *th, const T *b, const Geo<T> &g, KState *st, const RedBuf &rb, hipStream_t s);
// rbsor_core! (src/NonUniform.jl:848-915), one colour, ω = 1
template <typename T>
hipError_t launch_rbsor(T *x, const T *b, const Geo<T> &g, int color, const KState *st, hipStream_t s);
// Colour 0 of the first sweep of Preconditioner! from x = 0, fused with the
// myfill!(x, 0) before it and with forming its right-hand side bb:
//   FIRST_COPY   bb given (not written)
//   FIRST_BICG_P bb = p: iteration 1 p = r, later p = r + β(p - ω q)   (BiCG1!)
//   FIRST_BICG_S bb = s = (-α) q + r                                   (Triad!)
enum : int { FIRST_COPY = 0, FIRST_BICG_P = 1, FIRST_BICG_S = 2 };
template <typename T>
hipError_t launch_rbsor_first(int mode, T *x, T *bb, const T *r, const T *q, const Geo<T> &g, const KState *st,
                              hipStream_t s);
// The right-hand side alone (mode FIRST_BICG_P or FIRST_BICG_S), for the
// solve without a smoother
template <typename T>
hipError_t launch_form_rhs(int mode, T *bb, const T *r, const T *q, const Geo<T> &g, const KState *st,
                           hipStream_t s);
// workgroups (and so partials) of the reference-layout interior kernels
// (launch_dot, launch_calc_ax, ...) on an NX x NY x NZ interior
unsigned interior_launch_blocks(int NX, int NY, int NZ);
// Fdot1/Fdot2 (src/NonUniform.jl:567-592): s0 = Σ x·(y ? y : x), s1 = Σ x·z
template <typename T>
hipError_t launch_dot(const T *x, const T *y, const T *z, const Geo<T> &g, int fin, KState *st, double *hist,
                      const RedBuf &rb, hipStream_t s);
// BiCG1!/Triad!/BICG2! with host scalars (src/NonUniform.jl:603-659)
enum : int {
    AX_BICG1 = 0,   // z = x + a*(z - b*y)          BiCG1!(p=z, r=x, q=y, beta=a, omg=b)
    AX_TRIAD = 1,   // z = a*x + y                  Triad!
    AX_BICG2 = 2,   // z = z + (a*x + b*y)          BICG2!
};
template <typename T>
hipError_t launch_axpby(int op, T *z, const T *x, const T *y, double a, double b, const Geo<T> &g,
                        const KState *st, hipStream_t s);
// BICG2!(θ,p_,s_,α,ω); Triad!(r,t,s,-ω); Σ r²; with r0 also Σ r·r0.
// xbest (guarded solves, else nullptr) receives θ before the update when
// the previous residual was a new minimum.
template <typename T>
hipError_t launch_bicg_update(double *th, double *xbest, T *r, const T *p_, const T *s_, const T *t, const T *sv,
                              const T *r0, const Geo<T> &g, KState *st, double *hist, const RedBuf &rb,
                              hipStream_t s);
// CG!: θ += α p, r -= α q, Σ r²
template <typename T>
hipError_t launch_cg_update(double *th, double *xbest, T *r, const T *p, const T *q, const Geo<T> &g, KState *st,
                            double *hist, const RedBuf &rb, hipStream_t s);
// CG!: p = z + β p
template <typename T>
hipError_t launch_cg_p(T *p, const T *z, const Geo<T> &g, const KState *st, hipStream_t s);

// boundary conditions, RHS and set-up (FP64)
hipError_t launch_apply_bc_face(int face, int type, double temperature, double ambient, double *th, double *lam,
                                double *rho, double *cp, int MX, int MY, int MZ, hipStream_t s);
hipError_t launch_bc2rhs_face(int face, int type, double q, double h, double ta, double *b, const double *th,
                              const Geo<double> &g, hipStream_t s);
hipError_t launch_rhs_final(double *b, const double *th, const double *hsrc, const Geo<double> &g, hipStream_t s);
hipError_t launch_ap0(double *ap0, const double *rho, const double *cp, const Geo<double> &g, double ddt,
                      hipStream_t s);
// class of every cell: index of λ's bit pattern in the sorted list vals[0..K-1]
hipError_t launch_classify(uint8_t *cls, const double *lam, const double *rho, const double *cp, const uint64_t *vals,
                           int K, size_t n, hipStream_t s);
template <typename D, typename S> hipError_t launch_convert(D *dst, const S *src, size_t n, hipStream_t s);
// θ += d on the interior: the refinement update
hipError_t launch_add_correction(double *th, const double *d, const Geo<double> &g, hipStream_t s);

// ---- objective and adjoint gradients (h3d_adjoint.hip) ---------------------
// out[0] = max of θ over the region's interior cells (-DBL_MAX if none);
// bmax holds one value per workgroup of the cell grid
hipError_t launch_objective_max(const double *th, const uint8_t *region, const Geo<double> &g, double *bmax,
                                double *out, hipStream_t s);
// w = exp(β(θ - *mptr)) on the region (0 elsewhere); st->out[0] = Σ w
hipError_t launch_objective_exp(const double *th, const uint8_t *region, const Geo<double> &g, double beta,
                                const double *mptr, double *w, KState *st, const RedBuf &rb, hipStream_t s);
hipError_t launch_scale_interior(double *w, const Geo<double> &g, double sc, hipStream_t s);
// per-cell dJ/dλ and dJ/dhsrc from θ and the adjoint μ; iso: isothermal faces (bit f)
hipError_t launch_gradients(const double *th, const double *mu, const Geo<double> &g, int iso, double *glam,
                            double *ghs, hipStream_t s);

// Stand-alone RB-SOR (FP64, the reference's operator form): one colour of
// rbsor_core! with relaxation omega and the sum of its residual squares,
// and resSOR's sum over every cell; each followed by finalizer fin.
hipError_t launch_sor_color(double *x, const double *b, const Geo<double> &g, int color, double omega, int fin,
                            KState *st, double *hist, const RedBuf &rb, hipStream_t s);
hipError_t launch_res_sor(const double *x, const double *b, const Geo<double> &g, double omega, int fin, KState *st,
                          double *hist, const RedBuf &rb, hipStream_t s);

// sor! (lexicographic point SOR, FP64, the reference's operator form): one
// sweep in place on x as a hyperplane wavefront; rout receives each interior
// cell's residual term r (its sum of squares is sor!'s result)
hipError_t launch_sor_lex(double *x, const double *b, const Geo<double> &g, double omega, double *rout,
                          hipStream_t s);

// rhs = w·seed + D∘μ_next (μ_next may be nullptr): one step of the transient adjoint
hipError_t launch_adjoint_rhs(double *rhs, const double *seed, const double *mu_next, const Geo<double> &g, double w,
                              const double hconv[6], hipStream_t s);
hipError_t launch_accumulate(double *acc_l, double *acc_h, const double *gl, const double *gh, double wl, double wh,
                             size_t n, hipStream_t s);
hipError_t launch_scaled_copy(double *dst, const double *src, double f, size_t n, hipStream_t s);

// ---- RB-layout launchers (h3d_rb.hip) ---------------------------------------
// dst (RB) = src (standard layout), every cell including the halo
template <typename D> hipError_t launch_to_rb(D *dst, const D *src, int MX, int MY, int MZ, hipStream_t s);
// dst (standard) = src (RB), every cell including the halo
template <typename D> hipError_t launch_from_rb(D *dst, const D *src, int MX, int MY, int MZ, hipStream_t s);
template <typename T>
hipError_t launch_rb_copy_interior(T *dst, const T *src, const RbGeo<T> &g, const KState *st, hipStream_t s);
template <typename T>
hipError_t launch_rb_calc_ax(T *ax, const T *x, const RbGeo<T> &g, const T *dot_y, const T *dot_z, int fin,
                             KState *st, double *hist, const RedBuf &rb, hipStream_t s);
// r (RB) = b - A θ with θ and b in standard layout
template <typename T>
hipError_t launch_rb_calc_rk(T *r, T *r0copy, const T *th, const T *b, const RbGeo<T> &g, int fin, KState *st,
                             double *hist, const RedBuf &rb, hipStream_t s);
template <typename T>
hipError_t launch_rb_sweep(T *x, const T *b, const RbGeo<T> &g, int color, const KState *st, hipStream_t s,
                           bool from_zero = false);
template <typename T>
hipError_t launch_rb_cg_first(double *th, double *xbest, T *r, const T *p, const T *q, T *z, const RbGeo<T> &g,
                              KState *st, double *hist, const RedBuf &rb, hipStream_t s);
template <typename T>
hipError_t launch_rb_sweep_dot(T *x, const T *b, const RbGeo<T> &g, int fin, KState *st, double *hist,
                               const RedBuf &rb, hipStream_t s);
// Colour passes of halves hA then 1 - hA in one launch, xin -> xout (both
// whole RB arrays with equal halos), z-chunks of C planes; the result is
// bitwise that of the two in-place passes.  Only for operators for which
// rb_sweep2_fits (the class count bounds its LDS tables).
template <typename T>
hipError_t launch_rb_sweep2(const T *xin, T *xout, const T *b, const RbGeo<T> &g, int hA, int C, const KState *st,
                            hipStream_t s);
template <typename T>
bool rb_sweep2_fits(const RbGeo<T> &g);
// LDS budget of the two-colour sweep (above it the per-colour launches run)
constexpr size_t kSweep2LdsMax = 64 * 1024;
// longest z-chunk of the two-colour sweep
constexpr int kSweep2ChunkMax = 125;
// z-chunk the two-colour sweep uses on an HX x NY x NZ colour-split grid of
// elem_bytes-sized values when none is set (0: the per-colour launches)
int sweep_chunk_auto(int HX, int NY, int NZ, int elem_bytes);
template <typename T>
hipError_t launch_rb_first(int mode, T *x, T *bb, const T *r, const T *q, const RbGeo<T> &g, const KState *st,
                           hipStream_t s);
template <typename T>
hipError_t launch_rb_form_rhs(int mode, T *bb, const T *r, const T *q, const RbGeo<T> &g, const KState *st,
                              hipStream_t s);
template <typename T>
hipError_t launch_rb_dot(const T *x, const T *y, const T *z, const RbGeo<T> &g, int fin, KState *st,
                         double *hist, const RedBuf &rb, hipStream_t s);
// θ and xbest in standard layout (FP64), the rest RB
template <typename T>
hipError_t launch_rb_bicg_update(double *th, double *xbest, T *r, const T *p_, const T *s_, const T *t,
                                 const T *sv, const T *r0, const RbGeo<T> &g, KState *st, double *hist,
                                 const RedBuf &rb, hipStream_t s);
template <typename T>
hipError_t launch_rb_cg_update(double *th, double *xbest, T *r, const T *p, const T *q, const RbGeo<T> &g,
                               KState *st, double *hist, const RedBuf &rb, hipStream_t s);
template <typename T>
hipError_t launch_rb_cg_p(T *p, const T *z, const RbGeo<T> &g, const KState *st, hipStream_t s);

// ---- resident preconditioner (h3d_resident.hip) ------------------------------
// Tile plan for ncu CUs (G = 0 when the grid does not fit).
template <typename T> RbRes rb_res_plan(const RbGeo<T> &g, int ncu);
// Raise the kernels' dynamic LDS limit to kResLdsMax (once, outside capture).
template <typename T> hipError_t rb_res_prepare();
// The per-cell coefficients of the operator g into cf (once per solve).
template <typename T> hipError_t launch_rb_res_coef(const RbGeo<T> &g, const ResCoef<T> &cf, hipStream_t s);
// PBiCGSTAB!'s second half-iteration as one resident launch (see
// k_rb_resident, MODE 4) and its finalize (FIN_BICG_RES_RHO): alpha from the
// partials of (q, r0) the first launch left (launched with fin < 0),
// s = r - alpha q into sv, s_ = M⁻¹s, t = A s_, omega, θ += alpha p_ +
// omega s_, r = s - omega t, residual and next ρ.  Not in the reference-
// order sum mode.
template <typename T>
hipError_t launch_rb_resident_bicg_s(T *s_, T *sv, const T *r, const T *q, T *t, const T *p_, double *th,
                                     double *xbest, T *rw, const T *r0, const RbGeo<T> &g, const RbRes &pl, KState *st,
                                     double *hist, const RedBuf &rb, hipStream_t s);
// Every PBiCGSTAB! iteration of a solve (after CalcRK! and the first ρ,
// FIN_RES0_BICG) in one launch of the resident grid, until the state leaves
// ST_RUNNING (converged, breakdown, NaN, maxItr, guard); the state and the
// residual history are written as the per-iteration launches write them.
// Not in the reference-order sum mode.
template <typename T>
hipError_t launch_rb_bicg_solve(const ResBicg<T> &v, const RbGeo<T> &g, const RbRes &pl, KState *st, double *hist,
                                hipStream_t s);
// Every CG! iteration of a solve (after CalcRK!, z = M⁻¹r, p = z and
// ρ_old = (r, z), FIN_CG_RHO0) in one launch of the resident grid, until
// the state leaves ST_RUNNING; the state and the residual history are
// written as the per-iteration launches write them.  Not in the
// reference-order sum mode.
template <typename T>
hipError_t launch_rb_cg_solve(const ResCg<T> &v, const RbGeo<T> &g, const RbRes &pl, KState *st, double *hist,
                              hipStream_t s);
// CalcAX! fused after the resident preconditioner: ax = A x of its result,
// s0 = Σ ax·(dy ? dy : ax), s1 = Σ ax·dz (when dz), finalized with fin.
template <typename T>
struct ResAx {
    T *ax;
    const T *dy, *dz;
    int fin;
    KState *st;
    double *hist;
    const RedBuf *rb;
};
// Preconditioner! with :gs (five red-black sweeps) from x = 0 on the
// right-hand side bb formed per mode (as launch_rb_first), in one launch;
// with dot also Σ bb·x over the interior, finalized with dot->fin; with
// axf also CalcAX! of the result and its sums (not with dot).
template <typename T>
hipError_t launch_rb_resident(int mode, T *x, T *bb, const T *r, const T *q, const RbGeo<T> &g, const RbRes &pl,
                              const KState *st, const DotFin *dot, hipStream_t s, const ResAx<T> *axf = nullptr);

}  // namespace h3d
