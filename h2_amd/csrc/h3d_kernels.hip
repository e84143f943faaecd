// HIP kernels for the Heat3ds linear-solve path on gfx950 (MI355X).
//
// Every floating-point expression here performs the same IEEE operations,
// in the same order, as the reference Julia code it cites (the library is
// built with -ffp-contract=off, and FP64 division on gfx950 is correctly
// rounded), so element-wise results are bitwise equal to the reference's.
// Only sums over the grid differ: they run as fixed-order block trees (wave
// butterfly, then LDS, then a last-workgroup pass over per-block partials),
// which are deterministic run to run but not the reference's sequential
// order.
//
// The mask array of the reference is implicit: after
// apply_boundary_conditions! it is 0 on every halo cell and 1 inside, so for
// an interior cell the neighbour mask is (neighbour is interior ? 1 : 0).
//
// Work decomposition: a workgroup is 64 (x) by 4 (y) threads; each thread
// owns one cell column (i,j) of the interior and marches over KC planes in
// z.  Red-black sweeps give each thread the cells of one colour in a row
// pair i = i0 + 2t.

#include "h3d_internal.hpp"

#include <cmath>

namespace h3d {

constexpr int BX = 64;
constexpr int BY = 4;
constexpr int NT = BX * BY;
constexpr int NWAVE = NT / 64;
constexpr double kFloatMin = 1.0e-37;  // src/common.jl:12

// ---------------------------------------------------------------------------
// stencil (src/NonUniform.jl:477-507 and the identical blocks in CalcRK! and
// rbsor_core!).  (i,j,k) are 0-based interior indices, 1..N.
// ---------------------------------------------------------------------------

// λf(a, b, ma, mb) (src/common.jl:35) with mb = m0 = 1: the factor
// 2 - div(ma+1, 2) is 1 for an interior neighbour and 2 for a halo one.
template <typename T>
__device__ __forceinline__ T lamf(T a, T b, T ma)
{
    const T f = (ma == T(1)) ? T(1) : T(2);
    return T(2) * a * b / (a + b) * f;
}

template <typename T>
struct Sten {
    T dd, ss;
};

template <typename T>
__device__ __forceinline__ Sten<T> stencil(const T *__restrict__ x, const T *__restrict__ lam,
                                           const Geo<T> &g, int i, int j, int k, size_t c)
{
    const size_t sy = (size_t)g.MX, sz = (size_t)g.MX * (size_t)g.MY;
    const T dz_k = g.dZ[k];
    const T l0 = lam[c];
    const T mxm = (i > 1) ? T(1) : T(0), mxp = (i < g.NX) ? T(1) : T(0);
    const T mym = (j > 1) ? T(1) : T(0), myp = (j < g.NY) ? T(1) : T(0);
    const T mzm = (k > 1) ? T(1) : T(0), mzp = (k < g.NZ) ? T(1) : T(0);

    const T cond_xm = lamf(lam[c - 1], l0, mxm) * g.dy0 * dz_k * g.ddx;
    const T cond_xp = lamf(lam[c + 1], l0, mxp) * g.dy0 * dz_k * g.ddx;
    const T cond_ym = lamf(lam[c - sy], l0, mym) * g.dx0 * dz_k * g.ddy;
    const T cond_yp = lamf(lam[c + sy], l0, myp) * g.dx0 * dz_k * g.ddy;
    const T cond_zm = lamf(lam[c - sz], l0, mzm) * g.dx0 * g.dy0 / (g.ZC[k] - g.ZC[k - 1]);
    const T cond_zp = lamf(lam[c + sz], l0, mzp) * g.dx0 * g.dy0 / (g.ZC[k + 1] - g.ZC[k]);

    const T conv_xm = g.HC[0] * g.dy0 * dz_k * (T(1) - mxm);
    const T conv_xp = g.HC[1] * g.dy0 * dz_k * (T(1) - mxp);
    const T conv_ym = g.HC[2] * g.dx0 * dz_k * (T(1) - mym);
    const T conv_yp = g.HC[3] * g.dx0 * dz_k * (T(1) - myp);
    const T conv_zm = g.HC[4] * g.dx0 * g.dy0 * (T(1) - mzm);
    const T conv_zp = g.HC[5] * g.dx0 * g.dy0 * (T(1) - mzp);

    const T a_p_0 = g.ap0 ? g.ap0[c] : T(0);

    Sten<T> s;
    s.dd = (T(1) - T(1)) + (cond_xp + cond_xm + cond_yp + cond_ym + cond_zp + cond_zm + conv_xp +
                            conv_xm + conv_yp + conv_ym + conv_zp + conv_zm + a_p_0) *
                               T(1);
    s.ss = (cond_xp * x[c + 1] + cond_xm * x[c - 1] + cond_yp * x[c + sy] + cond_ym * x[c - sy] +
            cond_zp * x[c + sz] + cond_zm * x[c - sz]);
    return s;
}

// ---------------------------------------------------------------------------
// reductions: fixed-order block tree + last-workgroup finaliser
// ---------------------------------------------------------------------------

// Grid sums are accumulated in double-double (an unevaluated hi + lo pair,
// error-free TwoSum additions) through every level of the reduction tree,
// then rounded once.  With at most a few dozen terms per thread and a tree
// above that, the pair carries the sum to far below one ulp of the result,
// so the rounded value is the exactly rounded sum of the terms independent
// of how the grid is split into threads and workgroups; it is not the
// reference's sequential running sum, whose rounding depends on the order.
struct dd {
    double hi, lo;
};

__device__ __forceinline__ dd two_sum(double a, double b)
{
    const double s = a + b;
    const double bb = s - a;
    const double e = (a - (s - bb)) + (b - bb);
    return {s, e};
}

__device__ __forceinline__ dd quick_two_sum(double a, double b)
{
    const double s = a + b;
    return {s, b - (s - a)};
}

__device__ __forceinline__ dd dd_add(dd x, double t)
{
    const dd s = two_sum(x.hi, t);
    return quick_two_sum(s.hi, s.lo + x.lo);
}

__device__ __forceinline__ dd dd_add(dd x, dd y)
{
    const dd s = two_sum(x.hi, y.hi);
    const dd t = two_sum(x.lo, y.lo);
    dd u = quick_two_sum(s.hi, s.lo + t.hi);
    return quick_two_sum(u.hi, u.lo + t.lo);
}

__device__ __forceinline__ dd wave_sum(dd v)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        dd o;
        o.hi = __shfl_xor(v.hi, off, 64);
        o.lo = __shfl_xor(v.lo, off, 64);
        v = dd_add(v, o);
    }
    return v;
}

__device__ __forceinline__ double ld_coherent(const double *p)
{
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Scalar steps of the Krylov recurrences, run by a single thread once the
// sums s[] over the whole interior are known.
__device__ void fin_res0(double s0, KState *st)  // CalcRK! result; NonUniform.jl:263-272
{
    const double r0 = sqrt(s0);
    st->res0 = r0;
    st->itr = 0;
    if (r0 == 0.0) st->status = ST_CONVERGED;
    else if (!isfinite(r0)) st->status = ST_NAN;
}

__device__ void fin_bicg_rho(double rho, KState *st)  // NonUniform.jl:287-300
{
    st->itr += 1;
    st->rho = rho;
    if (fabs(rho) < kFloatMin) {
        st->status = ST_BREAKDOWN;
        return;
    }
    if (st->itr > 1) st->beta = st->rho / st->rho_old * st->alpha / st->omega;
}

__device__ void fin_bicg_res(double s0, KState *st, double *hist)  // NonUniform.jl:331-346
{
    double res = sqrt(s0);
    res /= st->res0;
    st->res = res;
    if (hist) hist[st->itr - 1] = res;
    if (!isfinite(res)) st->status = ST_NAN;
    else if (res < st->tol) st->status = ST_CONVERGED;
    else {
        st->rho_old = st->rho;
        if (st->itr >= st->maxItr) st->status = ST_MAXITER;
    }
}

__device__ void finalize(int fin, const double *s, KState *st, double *hist)
{
    switch (fin) {
    case FIN_STORE:
        st->out[0] = s[0];
        st->out[1] = s[1];
        st->out[2] = s[2];
        break;
    case FIN_RES0:
        fin_res0(s[0], st);
        break;
    case FIN_RES0_BICG:  // res0, then ρ of iteration 1 = (r, r0) with r0 == r: the same sum
        fin_res0(s[0], st);
        if (st->status == ST_RUNNING) fin_bicg_rho(s[0], st);
        break;
    case FIN_BICG_RHO:
        fin_bicg_rho(s[0], st);
        break;
    case FIN_BICG_ALPHA:  // NonUniform.jl:309
        st->alpha = st->rho / s[0];
        break;
    case FIN_BICG_OMEGA:  // NonUniform.jl:320-326
        if (fabs(s[0]) < kFloatMin) {
            st->status = ST_BREAKDOWN;
            break;
        }
        st->omega = s[1] / s[0];
        break;
    case FIN_BICG_RES:
        fin_bicg_res(s[0], st, hist);
        break;
    case FIN_BICG_RES_RHO:  // this iteration's residual, then the next iteration's ρ
        fin_bicg_res(s[0], st, hist);
        if (st->status == ST_RUNNING) fin_bicg_rho(s[1], st);
        break;
    case FIN_CG_RHO0:  // NonUniform.jl:155
        st->rho_old = s[0];
        break;
    case FIN_CG_ALPHA:  // NonUniform.jl:160-176
        st->itr += 1;
        if (fabs(s[0]) < kFloatMin) {
            st->status = ST_BREAKDOWN;
            break;
        }
        st->alpha = st->rho_old / s[0];
        break;
    case FIN_CG_RES: {  // NonUniform.jl:184-197
        double res = sqrt(s[0]);
        res /= st->res0;
        st->res = res;
        if (hist) hist[st->itr - 1] = res;
        if (!isfinite(res)) st->status = ST_NAN;
        else if (res < st->tol) st->status = ST_CONVERGED;
        break;
    }
    case FIN_CG_BETA: {  // NonUniform.jl:205-222
        const double rho_new = s[0];
        if (fabs(rho_new) < kFloatMin) {
            st->status = ST_BREAKDOWN;
            break;
        }
        st->beta = rho_new / st->rho_old;
        st->rho_old = rho_new;
        if (st->itr >= st->maxItr) st->status = ST_MAXITER;
        break;
    }
    default:
        break;
    }
}

// Block-reduce up to three per-thread sums, publish the block's partials and
// let the last block to arrive reduce all partials and run the finaliser.
// Every thread of the block must call this.
template <int NS>
__device__ void block_finish(const dd (&v)[3], int fin, KState *st, double *hist, const RedBuf &rb)
{
    __shared__ dd sh[3][NWAVE];
    __shared__ int am_last;
    const int tid = threadIdx.x + BX * threadIdx.y;
    const int lane = tid & 63, wave = tid >> 6;
#pragma unroll
    for (int s = 0; s < NS; ++s) {
        const dd w = wave_sum(v[s]);
        if (lane == 0) sh[s][wave] = w;
    }
    __syncthreads();
    const unsigned nb = gridDim.x * gridDim.y * gridDim.z;
    const unsigned bid = blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z);
    if (tid == 0) {
#pragma unroll
        for (int s = 0; s < NS; ++s) {
            dd b = sh[s][0];
            for (int w = 1; w < NWAVE; ++w) b = dd_add(b, sh[s][w]);
            rb.partials[6 * (size_t)bid + 2 * s] = b.hi;
            rb.partials[6 * (size_t)bid + 2 * s + 1] = b.lo;
        }
        __threadfence();
        const unsigned prev = atomicAdd(rb.counter, 1u);
        am_last = (prev == nb - 1);
    }
    __syncthreads();
    if (!am_last) return;
    __threadfence();
    dd acc[3] = {{0.0, 0.0}, {0.0, 0.0}, {0.0, 0.0}};
    for (unsigned b = tid; b < nb; b += NT) {
#pragma unroll
        for (int s = 0; s < NS; ++s) {
            dd p;
            p.hi = ld_coherent(&rb.partials[6 * (size_t)b + 2 * s]);
            p.lo = ld_coherent(&rb.partials[6 * (size_t)b + 2 * s + 1]);
            acc[s] = dd_add(acc[s], p);
        }
    }
    __syncthreads();
#pragma unroll
    for (int s = 0; s < NS; ++s) {
        const dd w = wave_sum(acc[s]);
        if (lane == 0) sh[s][wave] = w;
    }
    __syncthreads();
    if (tid == 0) {
        double tot[3] = {0.0, 0.0, 0.0};
#pragma unroll
        for (int s = 0; s < NS; ++s) {
            dd t = sh[s][0];
            for (int w = 1; w < NWAVE; ++w) t = dd_add(t, sh[s][w]);
            tot[s] = t.hi + t.lo;
        }
        finalize(fin, tot, st, hist);
        __hip_atomic_store(rb.counter, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
}

__device__ __forceinline__ bool halted(const KState *st) { return st != nullptr && st->status != ST_RUNNING; }

// ---------------------------------------------------------------------------
// launch geometry
// ---------------------------------------------------------------------------
struct Shape {
    dim3 grid, block;
    int KC;
};

static Shape interior_shape(int nx_threads, int NY, int NZ)
{
    Shape sh;
    sh.block = dim3(BX, BY, 1);
    const int gx = (nx_threads + BX - 1) / BX, gy = (NY + BY - 1) / BY;
    const long nxy = (long)gx * gy;
    // enough planes per thread that the grid stays near 2k workgroups
    int kc = (int)((nxy * NZ + 2047) / 2048);
    if (kc < 1) kc = 1;
    if (kc > NZ) kc = NZ;
    const int gz = (NZ + kc - 1) / kc;
    sh.grid = dim3(gx, gy, gz);
    sh.KC = kc;
    return sh;
}

// ---------------------------------------------------------------------------
// element-wise kernels
// ---------------------------------------------------------------------------
template <typename T>
__global__ void __launch_bounds__(256) k_fill(T *a, T v, size_t n)
{
    for (size_t c = blockIdx.x * (size_t)blockDim.x + threadIdx.x; c < n; c += (size_t)gridDim.x * blockDim.x)
        a[c] = v;
}

template <typename T>
hipError_t launch_fill(T *a, T v, size_t n, hipStream_t s)
{
    const int nb = (int)((n + 255) / 256 < 8192 ? (n + 255) / 256 : 8192);
    hipLaunchKernelGGL(k_fill<T>, dim3(nb > 0 ? nb : 1), dim3(256), 0, s, a, v, n);
    return hipGetLastError();
}

#define INTERIOR_COLUMN(g, KC)                                              \
    const int i = blockIdx.x * BX + threadIdx.x + 1;                        \
    const int j = blockIdx.y * BY + threadIdx.y + 1;                        \
    const bool active = (i <= (g).NX) && (j <= (g).NY);                     \
    const int kb = 1 + (int)blockIdx.z * (KC);                              \
    const int ke = min(kb + (KC), (g).NZ + 1);                              \
    const size_t sz_ = (size_t)(g).MX * (size_t)(g).MY;                     \
    const size_t c0_ = (size_t)i + (size_t)(g).MX * (size_t)j;

template <typename T>
__global__ void __launch_bounds__(256) k_copy_interior(T *dst, const T *src, Geo<T> g, const KState *st, int KC)
{
    if (halted(st)) return;
    INTERIOR_COLUMN(g, KC)
    if (!active) return;
    for (int k = kb; k < ke; ++k) dst[c0_ + sz_ * k] = src[c0_ + sz_ * k];
}

template <typename T>
hipError_t launch_copy_interior(T *dst, const T *src, const Geo<T> &g, const KState *st, hipStream_t s)
{
    Shape sh = interior_shape(g.NX, g.NY, g.NZ);
    hipLaunchKernelGGL(k_copy_interior<T>, sh.grid, sh.block, 0, s, dst, src, g, st, sh.KC);
    return hipGetLastError();
}

// CalcAX! (src/NonUniform.jl:454-511), optionally fused with one or two dot
// products of the result: s0 = Σ ax·(y ? y : ax), s1 = Σ ax·z.
template <typename T, int NS>
__global__ void __launch_bounds__(256) k_calc_ax(T *ax, const T *x, const T *lam, Geo<T> g, const T *dy,
                                                 const T *dz, int fin, KState *st, double *hist, RedBuf rb, int KC)
{
    if (halted(st)) return;
    INTERIOR_COLUMN(g, KC)
    dd v[3] = {{0.0, 0.0}, {0.0, 0.0}, {0.0, 0.0}};
    if (active) {
        for (int k = kb; k < ke; ++k) {
            const size_t c = c0_ + sz_ * k;
            const Sten<T> s = stencil(x, lam, g, i, j, k, c);
            const T a = (s.ss - s.dd * x[c]) * T(1);
            ax[c] = a;
            if (NS >= 1) v[0] = dd_add(v[0], (double)a * (double)(dy ? dy[c] : a));
            if (NS >= 2) v[1] = dd_add(v[1], (double)a * (double)dz[c]);
        }
    }
    if (NS > 0) block_finish<NS>(v, fin, st, hist, rb);
}

template <typename T>
hipError_t launch_calc_ax(T *ax, const T *x, const T *lam, const Geo<T> &g, const T *dot_y, const T *dot_z,
                          int fin, KState *st, double *hist, const RedBuf &rb, hipStream_t s)
{
    Shape sh = interior_shape(g.NX, g.NY, g.NZ);
    if (fin < 0)
        hipLaunchKernelGGL((k_calc_ax<T, 0>), sh.grid, sh.block, 0, s, ax, x, lam, g, dot_y, dot_z, fin, st, hist, rb, sh.KC);
    else if (dot_z == nullptr)
        hipLaunchKernelGGL((k_calc_ax<T, 1>), sh.grid, sh.block, 0, s, ax, x, lam, g, dot_y, dot_z, fin, st, hist, rb, sh.KC);
    else
        hipLaunchKernelGGL((k_calc_ax<T, 2>), sh.grid, sh.block, 0, s, ax, x, lam, g, dot_y, dot_z, fin, st, hist, rb, sh.KC);
    return hipGetLastError();
}

// CalcRK! (src/NonUniform.jl:372-434): r = (b - Aθ)·m and Σ r².  With
// r0copy it also performs PBiCGSTAB!'s mycopy!(r0, r) (NonUniform.jl:274).
template <typename T>
__global__ void __launch_bounds__(256) k_calc_rk(T *r, T *r0copy, const T *th, const T *b, const T *lam, Geo<T> g,
                                                 int fin, KState *st, double *hist, RedBuf rb, int KC)
{
    if (halted(st)) return;
    INTERIOR_COLUMN(g, KC)
    dd v[3] = {{0.0, 0.0}, {0.0, 0.0}, {0.0, 0.0}};
    if (active) {
        for (int k = kb; k < ke; ++k) {
            const size_t c = c0_ + sz_ * k;
            const Sten<T> s = stencil(th, lam, g, i, j, k, c);
            const T rs = (b[c] - (s.ss - s.dd * th[c])) * T(1);
            r[c] = rs;
            if (r0copy) r0copy[c] = rs;
            v[0] = dd_add(v[0], (double)rs * (double)rs);
        }
    }
    block_finish<1>(v, fin, st, hist, rb);
}

template <typename T>
hipError_t launch_calc_rk(T *r, T *r0copy, const T *th, const T *b, const T *lam, const Geo<T> &g, int fin,
                          KState *st, double *hist, const RedBuf &rb, hipStream_t s)
{
    Shape sh = interior_shape(g.NX, g.NY, g.NZ);
    hipLaunchKernelGGL(k_calc_rk<T>, sh.grid, sh.block, 0, s, r, r0copy, th, b, lam, g, fin, st, hist, rb, sh.KC);
    return hipGetLastError();
}

// rbsor_core! (src/NonUniform.jl:848-915) with ω = 1: one colour of
// Gauss-Seidel in place.  Colour c updates 0-based (i+j+k+c) odd, i.e. the
// reference's i in 2+mod(k+j+c,2):2:MX-1.  The residual sum the reference
// returns is discarded by Preconditioner! and is not formed.
template <typename T>
__global__ void __launch_bounds__(256) k_rbsor(T *x, const T *b, const T *lam, Geo<T> g, int color,
                                               const KState *st, int KC)
{
    if (halted(st)) return;
    const int t = blockIdx.x * BX + threadIdx.x;
    const int j = blockIdx.y * BY + threadIdx.y + 1;
    if (j > g.NY) return;
    const int kb = 1 + (int)blockIdx.z * KC;
    const int ke = min(kb + KC, g.NZ + 1);
    const size_t sz = (size_t)g.MX * (size_t)g.MY;
    for (int k = kb; k < ke; ++k) {
        const int i = (((j + k + color) & 1) ? 2 : 1) + 2 * t;
        if (i > g.NX) continue;
        const size_t c = (size_t)i + (size_t)g.MX * (size_t)j + sz * (size_t)k;
        const T pp = x[c];
        const Sten<T> s = stencil(x, lam, g, i, j, k, c);
        const T dp = (((s.ss - b[c]) / s.dd - pp)) * T(1);
        x[c] = pp + T(1) * dp;
    }
}

template <typename T>
hipError_t launch_rbsor(T *x, const T *b, const T *lam, const Geo<T> &g, int color, const KState *st, hipStream_t s)
{
    Shape sh = interior_shape((g.NX + 1) / 2, g.NY, g.NZ);
    hipLaunchKernelGGL(k_rbsor<T>, sh.grid, sh.block, 0, s, x, b, lam, g, color, st, sh.KC);
    return hipGetLastError();
}

// Fdot1/Fdot2 (src/NonUniform.jl:567-592), up to three dots in one pass:
// s0 = Σ x·y (y == nullptr: Σ x²), s1 = Σ x·z.
template <typename T, int NS>
__global__ void __launch_bounds__(256) k_dot(const T *x, const T *y, const T *z, Geo<T> g, int fin, KState *st,
                                             double *hist, RedBuf rb, int KC)
{
    if (halted(st)) return;
    INTERIOR_COLUMN(g, KC)
    dd v[3] = {{0.0, 0.0}, {0.0, 0.0}, {0.0, 0.0}};
    if (active) {
        for (int k = kb; k < ke; ++k) {
            const size_t c = c0_ + sz_ * k;
            const double xv = (double)x[c];
            v[0] = dd_add(v[0], xv * (double)(y ? y[c] : x[c]));
            if (NS >= 2) v[1] = dd_add(v[1], xv * (double)z[c]);
        }
    }
    block_finish<NS>(v, fin, st, hist, rb);
}

template <typename T>
hipError_t launch_dot(const T *x, const T *y, const T *z, const Geo<T> &g, int fin, KState *st, double *hist,
                      const RedBuf &rb, hipStream_t s)
{
    Shape sh = interior_shape(g.NX, g.NY, g.NZ);
    if (z == nullptr)
        hipLaunchKernelGGL((k_dot<T, 1>), sh.grid, sh.block, 0, s, x, y, z, g, fin, st, hist, rb, sh.KC);
    else
        hipLaunchKernelGGL((k_dot<T, 2>), sh.grid, sh.block, 0, s, x, y, z, g, fin, st, hist, rb, sh.KC);
    return hipGetLastError();
}

// BiCG1! / Triad! / BICG2! (src/NonUniform.jl:603-659) with host scalars.
template <typename T>
__global__ void __launch_bounds__(256) k_axpby(int op, T *z, const T *x, const T *y, T a, T b, Geo<T> g,
                                               const KState *st, int KC)
{
    if (halted(st)) return;
    INTERIOR_COLUMN(g, KC)
    if (!active) return;
    for (int k = kb; k < ke; ++k) {
        const size_t c = c0_ + sz_ * k;
        if (op == AX_BICG1) z[c] = x[c] + a * (z[c] - b * y[c]);
        else if (op == AX_TRIAD) z[c] = a * x[c] + y[c];
        else z[c] = z[c] + (a * x[c] + b * y[c]);
    }
}

template <typename T>
hipError_t launch_axpby(int op, T *z, const T *x, const T *y, const T *w, double a, double b, const Geo<T> &g,
                        const KState *st, hipStream_t s)
{
    (void)w;
    Shape sh = interior_shape(g.NX, g.NY, g.NZ);
    hipLaunchKernelGGL(k_axpby<T>, sh.grid, sh.block, 0, s, op, z, x, y, (T)a, (T)b, g, st, sh.KC);
    return hipGetLastError();
}

// PBiCGSTAB! p update with device scalars: iteration 1 mycopy!(p, r)
// (NonUniform.jl:297), later BiCG1!(p, r, q, beta, omega) (:299-300).
template <typename T>
__global__ void __launch_bounds__(256) k_bicg_p(T *p, const T *r, const T *q, Geo<T> g, const KState *st, int KC)
{
    if (halted(st)) return;
    INTERIOR_COLUMN(g, KC)
    if (!active) return;
    const bool first = (st->itr == 1);
    const T beta = (T)st->beta, omg = (T)st->omega;
    for (int k = kb; k < ke; ++k) {
        const size_t c = c0_ + sz_ * k;
        p[c] = first ? r[c] : r[c] + beta * (p[c] - omg * q[c]);
    }
}

template <typename T>
hipError_t launch_bicg_p(T *p, const T *r, const T *q, const Geo<T> &g, const KState *st, hipStream_t s)
{
    Shape sh = interior_shape(g.NX, g.NY, g.NZ);
    hipLaunchKernelGGL(k_bicg_p<T>, sh.grid, sh.block, 0, s, p, r, q, g, st, sh.KC);
    return hipGetLastError();
}

// s = (-alpha) q + r  (Triad!, NonUniform.jl:310-311)
template <typename T>
__global__ void __launch_bounds__(256) k_bicg_s(T *sv, const T *q, const T *r, Geo<T> g, const KState *st, int KC)
{
    if (halted(st)) return;
    INTERIOR_COLUMN(g, KC)
    if (!active) return;
    const T r_alpha = (T)(-st->alpha);
    for (int k = kb; k < ke; ++k) {
        const size_t c = c0_ + sz_ * k;
        sv[c] = r_alpha * q[c] + r[c];
    }
}

template <typename T>
hipError_t launch_bicg_s(T *s_, const T *q, const T *r, const Geo<T> &g, const KState *st, hipStream_t s)
{
    Shape sh = interior_shape(g.NX, g.NY, g.NZ);
    hipLaunchKernelGGL(k_bicg_s<T>, sh.grid, sh.block, 0, s, s_, q, r, g, st, sh.KC);
    return hipGetLastError();
}

// BICG2!(θ, p_, s_, alpha, omega); Triad!(r, t_, s, -omega); Σ r², and with
// r0 also Σ r·r0 for the next iteration's ρ (NonUniform.jl:328-332, 289).
template <typename T, int NS>
__global__ void __launch_bounds__(256) k_bicg_update(T *th, T *r, const T *p_, const T *s_, const T *t, const T *sv,
                                                     const T *r0, Geo<T> g, KState *st, double *hist, RedBuf rb,
                                                     int fin, int KC)
{
    if (halted(st)) return;
    INTERIOR_COLUMN(g, KC)
    const T alpha = (T)st->alpha, omega = (T)st->omega, r_omega = (T)(-st->omega);
    dd v[3] = {{0.0, 0.0}, {0.0, 0.0}, {0.0, 0.0}};
    if (active) {
        for (int k = kb; k < ke; ++k) {
            const size_t c = c0_ + sz_ * k;
            th[c] = th[c] + (alpha * p_[c] + omega * s_[c]);
            const T rv = r_omega * t[c] + sv[c];
            r[c] = rv;
            v[0] = dd_add(v[0], (double)rv * (double)rv);
            if (NS >= 2) v[1] = dd_add(v[1], (double)rv * (double)r0[c]);
        }
    }
    block_finish<NS>(v, fin, st, hist, rb);
}

template <typename T>
hipError_t launch_bicg_update(T *th, T *r, const T *p_, const T *s_, const T *t, const T *sv, const T *r0,
                              const Geo<T> &g, KState *st, double *hist, const RedBuf &rb, hipStream_t s)
{
    Shape sh = interior_shape(g.NX, g.NY, g.NZ);
    if (r0 == nullptr)
        hipLaunchKernelGGL((k_bicg_update<T, 1>), sh.grid, sh.block, 0, s, th, r, p_, s_, t, sv, r0, g, st, hist, rb,
                           (int)FIN_BICG_RES, sh.KC);
    else
        hipLaunchKernelGGL((k_bicg_update<T, 2>), sh.grid, sh.block, 0, s, th, r, p_, s_, t, sv, r0, g, st, hist, rb,
                           (int)FIN_BICG_RES_RHO, sh.KC);
    return hipGetLastError();
}

// CG!: θ += α p, r -= α q, Σ r² (NonUniform.jl:179-186)
template <typename T>
__global__ void __launch_bounds__(256) k_cg_update(T *th, T *r, const T *p, const T *q, Geo<T> g, KState *st,
                                                   double *hist, RedBuf rb, int KC)
{
    if (halted(st)) return;
    INTERIOR_COLUMN(g, KC)
    const T alpha = (T)st->alpha;
    dd v[3] = {{0.0, 0.0}, {0.0, 0.0}, {0.0, 0.0}};
    if (active) {
        for (int k = kb; k < ke; ++k) {
            const size_t c = c0_ + sz_ * k;
            th[c] = th[c] + alpha * p[c];
            const T rv = r[c] - alpha * q[c];
            r[c] = rv;
            v[0] = dd_add(v[0], (double)rv * (double)rv);
        }
    }
    block_finish<1>(v, FIN_CG_RES, st, hist, rb);
}

template <typename T>
hipError_t launch_cg_update(T *th, T *r, const T *p, const T *q, const Geo<T> &g, KState *st, double *hist,
                            const RedBuf &rb, hipStream_t s)
{
    Shape sh = interior_shape(g.NX, g.NY, g.NZ);
    hipLaunchKernelGGL(k_cg_update<T>, sh.grid, sh.block, 0, s, th, r, p, q, g, st, hist, rb, sh.KC);
    return hipGetLastError();
}

// CG!: p = z + β p (NonUniform.jl:218-220)
template <typename T>
__global__ void __launch_bounds__(256) k_cg_p(T *p, const T *z, Geo<T> g, const KState *st, int KC)
{
    if (halted(st)) return;
    INTERIOR_COLUMN(g, KC)
    if (!active) return;
    const T beta = (T)st->beta;
    for (int k = kb; k < ke; ++k) {
        const size_t c = c0_ + sz_ * k;
        p[c] = z[c] + beta * p[c];
    }
}

template <typename T>
hipError_t launch_cg_p(T *p, const T *z, const Geo<T> &g, const KState *st, hipStream_t s)
{
    Shape sh = interior_shape(g.NX, g.NY, g.NZ);
    hipLaunchKernelGGL(k_cg_p<T>, sh.grid, sh.block, 0, s, p, z, g, st, sh.KC);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// boundary conditions and RHS (FP64 only)
// ---------------------------------------------------------------------------

// apply_isothermal!/apply_heat_flux!/apply_convection! for one face over its
// full extent 1:SZ (src/boundary_conditions.jl:156-313).  Faces must be
// launched in the reference order x-,x+,y-,y+,z-,z+ because edge cells are
// written by more than one face.
__global__ void __launch_bounds__(256) k_apply_bc_face(int face, int type, double temperature, double ambient,
                                                       double *th, double *lam, double *rho, double *cp,
                                                       int MX, int MY, int MZ)
{
    const int axis = face >> 1, plus = face & 1;
    const int n1 = axis == 0 ? MY : MX;
    const int n2 = axis == 2 ? MY : MZ;
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    const int bq = blockIdx.y;
    if (a >= n1 || bq >= n2) return;
    const int S = axis == 0 ? MX : (axis == 1 ? MY : MZ);
    const int h = plus ? S - 1 : 0, in = plus ? S - 2 : 1;
    int ih, jh, kh, ii, ji, ki;
    if (axis == 0) { ih = h; ii = in; jh = ji = a; kh = ki = bq; }
    else if (axis == 1) { jh = h; ji = in; ih = ii = a; kh = ki = bq; }
    else { kh = h; ki = in; ih = ii = a; jh = ji = bq; }
    const size_t c = (size_t)ih + (size_t)MX * ((size_t)jh + (size_t)MY * kh);
    const size_t ci = (size_t)ii + (size_t)MX * ((size_t)ji + (size_t)MY * ki);
    if (type == 0) {          // ISOTHERMAL
        lam[c] = lam[ci];
        rho[c] = rho[ci];
        cp[c] = cp[ci];
        th[c] = temperature;
    } else if (type == 1) {   // HEAT_FLUX
        lam[c] = 0.0;
    } else if (type == 2) {   // CONVECTION
        lam[c] = 0.0;
        th[c] = ambient;
    }
}

hipError_t launch_apply_bc_face(int face, int type, double temperature, double ambient, double *th, double *lam,
                                double *rho, double *cp, int MX, int MY, int MZ, hipStream_t s)
{
    const int axis = face >> 1;
    const int n1 = axis == 0 ? MY : MX;
    const int n2 = axis == 2 ? MY : MZ;
    hipLaunchKernelGGL(k_apply_bc_face, dim3((n1 + 255) / 256, n2), dim3(256), 0, s, face, type, temperature,
                       ambient, th, lam, rho, cp, MX, MY, MZ);
    return hipGetLastError();
}

// RHS_heat_flux!/RHS_convection! for one face (src/RHS.jl:83-192): the first
// interior layer gets b ± q·area (flux) or b ∓ h·area·(θ∞ − θ) (convection).
__global__ void __launch_bounds__(256) k_bc2rhs_face(int face, int type, double q, double hc, double ta, double *b,
                                                     const double *th, Geo<double> g)
{
    const int axis = face >> 1, plus = face & 1;
    const int n1 = axis == 0 ? g.NY : g.NX;
    const int n2 = axis == 2 ? g.NY : g.NZ;
    const int a = blockIdx.x * blockDim.x + threadIdx.x + 1;
    const int bq = blockIdx.y + 1;
    if (a > n1 || bq > n2) return;
    int i, j, k;
    double area;
    if (axis == 0) { i = plus ? g.NX : 1; j = a; k = bq; area = g.dy0 * g.dZ[k]; }
    else if (axis == 1) { j = plus ? g.NY : 1; i = a; k = bq; area = g.dx0 * g.dZ[k]; }
    else { k = plus ? g.NZ : 1; i = a; j = bq; area = g.dx0 * g.dy0; }
    const size_t c = (size_t)i + (size_t)g.MX * ((size_t)j + (size_t)g.MY * k);
    if (type == 1) {
        b[c] = plus ? b[c] - q * area : b[c] + q * area;
    } else if (type == 2) {
        b[c] = plus ? b[c] + hc * area * (ta - th[c]) : b[c] - hc * area * (ta - th[c]);
    }
}

hipError_t launch_bc2rhs_face(int face, int type, double q, double h, double ta, double *b, const double *th,
                              const Geo<double> &g, hipStream_t s)
{
    const int axis = face >> 1;
    const int n1 = axis == 0 ? g.NY : g.NX;
    const int n2 = axis == 2 ? g.NY : g.NZ;
    hipLaunchKernelGGL(k_bc2rhs_face, dim3((n1 + 255) / 256, n2), dim3(256), 0, s, face, type, q, h, ta, b, th, g);
    return hipGetLastError();
}

// calRHS! final loop (src/NonUniform.jl:85-90):
// b = -(ap0·θ + hsrc·dx·dy·dz + b), ap0 = 0 when steady.
__global__ void __launch_bounds__(256) k_rhs_final(double *b, const double *th, const double *hsrc, Geo<double> g,
                                                   int KC)
{
    INTERIOR_COLUMN(g, KC)
    if (!active) return;
    for (int k = kb; k < ke; ++k) {
        const size_t c = c0_ + sz_ * k;
        const double dz_k = g.dZ[k];
        const double a_p_0 = g.ap0 ? g.ap0[c] : 0.0;
        b[c] = -(a_p_0 * th[c] + hsrc[c] * g.dx0 * g.dy0 * dz_k + b[c]);
    }
}

hipError_t launch_rhs_final(double *b, const double *th, const double *hsrc, const Geo<double> &g, hipStream_t s)
{
    Shape sh = interior_shape(g.NX, g.NY, g.NZ);
    hipLaunchKernelGGL(k_rhs_final, sh.grid, sh.block, 0, s, b, th, hsrc, g, sh.KC);
    return hipGetLastError();
}

// a_p_0 = ρ·cp·dx·dy·dz_k·(1/Δt) per interior cell (src/NonUniform.jl:88,420)
__global__ void __launch_bounds__(256) k_ap0(double *ap0, const double *rho, const double *cp, Geo<double> g,
                                             double ddt, int KC)
{
    INTERIOR_COLUMN(g, KC)
    if (!active) return;
    for (int k = kb; k < ke; ++k) {
        const size_t c = c0_ + sz_ * k;
        ap0[c] = rho[c] * cp[c] * g.dx0 * g.dy0 * g.dZ[k] * ddt;
    }
}

hipError_t launch_ap0(double *ap0, const double *rho, const double *cp, const Geo<double> &g, double ddt,
                      hipStream_t s)
{
    Shape sh = interior_shape(g.NX, g.NY, g.NZ);
    hipLaunchKernelGGL(k_ap0, sh.grid, sh.block, 0, s, ap0, rho, cp, g, ddt, sh.KC);
    return hipGetLastError();
}

template <typename D, typename S>
__global__ void __launch_bounds__(256) k_convert(D *dst, const S *src, size_t n)
{
    for (size_t c = blockIdx.x * (size_t)blockDim.x + threadIdx.x; c < n; c += (size_t)gridDim.x * blockDim.x)
        dst[c] = (D)src[c];
}

template <typename D, typename S>
hipError_t launch_convert(D *dst, const S *src, size_t n, hipStream_t s)
{
    const int nb = (int)((n + 255) / 256 < 8192 ? (n + 255) / 256 : 8192);
    hipLaunchKernelGGL((k_convert<D, S>), dim3(nb > 0 ? nb : 1), dim3(256), 0, s, dst, src, n);
    return hipGetLastError();
}

__global__ void __launch_bounds__(256) k_add_correction(double *th, const float *d, Geo<double> g, int KC)
{
    INTERIOR_COLUMN(g, KC)
    if (!active) return;
    for (int k = kb; k < ke; ++k) {
        const size_t c = c0_ + sz_ * k;
        th[c] = th[c] + (double)d[c];
    }
}

hipError_t launch_add_correction(double *th, const float *d, const Geo<double> &g, hipStream_t s)
{
    Shape sh = interior_shape(g.NX, g.NY, g.NZ);
    hipLaunchKernelGGL(k_add_correction, sh.grid, sh.block, 0, s, th, d, g, sh.KC);
    return hipGetLastError();
}

// explicit instantiations
#define INST(T)                                                                                                   \
    template hipError_t launch_fill<T>(T *, T, size_t, hipStream_t);                                              \
    template hipError_t launch_copy_interior<T>(T *, const T *, const Geo<T> &, const KState *, hipStream_t);     \
    template hipError_t launch_calc_ax<T>(T *, const T *, const T *, const Geo<T> &, const T *, const T *, int,   \
                                          KState *, double *, const RedBuf &, hipStream_t);                      \
    template hipError_t launch_calc_rk<T>(T *, T *, const T *, const T *, const T *, const Geo<T> &, int,        \
                                          KState *, double *, const RedBuf &, hipStream_t);                      \
    template hipError_t launch_rbsor<T>(T *, const T *, const T *, const Geo<T> &, int, const KState *,          \
                                        hipStream_t);                                                            \
    template hipError_t launch_dot<T>(const T *, const T *, const T *, const Geo<T> &, int, KState *, double *,  \
                                      const RedBuf &, hipStream_t);                                              \
    template hipError_t launch_axpby<T>(int, T *, const T *, const T *, const T *, double, double,               \
                                        const Geo<T> &, const KState *, hipStream_t);                            \
    template hipError_t launch_bicg_p<T>(T *, const T *, const T *, const Geo<T> &, const KState *, hipStream_t); \
    template hipError_t launch_bicg_s<T>(T *, const T *, const T *, const Geo<T> &, const KState *, hipStream_t); \
    template hipError_t launch_bicg_update<T>(T *, T *, const T *, const T *, const T *, const T *, const T *,   \
                                              const Geo<T> &, KState *, double *, const RedBuf &, hipStream_t);  \
    template hipError_t launch_cg_update<T>(T *, T *, const T *, const T *, const Geo<T> &, KState *, double *,  \
                                            const RedBuf &, hipStream_t);                                        \
    template hipError_t launch_cg_p<T>(T *, const T *, const Geo<T> &, const KState *, hipStream_t);
INST(double)
INST(float)
template hipError_t launch_convert<float, double>(float *, const double *, size_t, hipStream_t);
template hipError_t launch_convert<double, float>(double *, const float *, size_t, hipStream_t);

}  // namespace h3d
