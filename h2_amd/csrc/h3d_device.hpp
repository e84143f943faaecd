// Device helpers shared by the kernel files: launch constants, the
// double-double grid sums and the halted() test of the solver state.
#pragma once

#include "h3d_internal.hpp"

namespace h3d {

constexpr int BX = 64;
constexpr int BY = 4;
constexpr int NT = BX * BY;
constexpr int NWAVE = NT / 64;
constexpr double kFloatMin = 1.0e-37;  // src/common.jl:12

// ---------------------------------------------------------------------------
// grid sums: double-double through a fixed reduction tree
// ---------------------------------------------------------------------------
// An unevaluated hi + lo pair with error-free TwoSum additions.  Each thread
// adds at most a few dozen terms, the tree above adds pairs (one TwoSum of
// the high parts, the low parts added in FP64); the pair keeps the sum
// within about 2^-100 of the sum of the terms' magnitudes, so the final
// rounding gives the exactly rounded sum of the terms whatever the
// decomposition, unless the exact sum lies that close to a rounding
// boundary.
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
    const dd u = quick_two_sum(s.hi, s.lo + t.hi);
    return quick_two_sum(u.hi, u.lo + t.lo);
}

// v of lane l-N of the same 16-lane row, 0 where there is none (DPP
// row_shr, a VALU operation, on both 32-bit halves).
template <int N>
__device__ __forceinline__ double row_shr(double v)
{
    const unsigned long long u = (unsigned long long)__double_as_longlong(v);
    const int lo = __builtin_amdgcn_update_dpp(0, (int)(unsigned)(u & 0xffffffffull), 0x110 + N, 0xf, 0xf, true);
    const int hi = __builtin_amdgcn_update_dpp(0, (int)(unsigned)(u >> 32), 0x110 + N, 0xf, 0xf, true);
    return __longlong_as_double((long long)(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo));
}

// Inclusive scan within each 16-lane row: lane 15 of a row ends with the
// row's sum.  All lanes of the wave must be active.
// A thread's two terms as one exact pair (an invalid term counts as 0):
// one TwoSum instead of two dd_add steps from zero.
__device__ __forceinline__ dd pair_sum(bool v0, double t0, bool v1, double t1)
{
    return two_sum(v0 ? t0 : 0.0, v1 ? t1 : 0.0);
}

// Pair addition of the reduction tree: 11 FP64 operations against the 20
// of dd_add(dd, dd), whose second TwoSum only matters when the high parts
// cancel to far below the low parts' scale.  Error at most about 2u²
// (|x| + |y|), u = 2^-53: over the tree's levels it stays within the
// bound above (measured: whole solves bitwise unchanged, 1.9 % faster at
// 240x240x30).
__device__ __forceinline__ dd dd_add_tree(dd x, dd y)
{
    const dd s = two_sum(x.hi, y.hi);
    return quick_two_sum(s.hi, s.lo + (x.lo + y.lo));
}

__device__ __forceinline__ dd row_sum(dd v)
{
    v = dd_add_tree(v, dd{row_shr<1>(v.hi), row_shr<1>(v.lo)});
    v = dd_add_tree(v, dd{row_shr<2>(v.hi), row_shr<2>(v.lo)});
    v = dd_add_tree(v, dd{row_shr<4>(v.hi), row_shr<4>(v.lo)});
    v = dd_add_tree(v, dd{row_shr<8>(v.hi), row_shr<8>(v.lo)});
    return v;
}

// Workgroup barrier that orders LDS only.  __syncthreads() also orders
// global memory, which makes every wave wait for its outstanding global
// stores first; the reductions exchange data through LDS alone.
__device__ __forceinline__ void lds_barrier()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// Sum up to three per-thread values over the NT = 256-thread workgroup:
// row sums (16 of them) through LDS into the first row of wave 0, whose
// lane 15 ends with the totals (the return value is meaningful there only:
// leader() is true).  Every thread must call it.
template <int NS>
__device__ __forceinline__ void block_sum(const dd (&v)[3], dd (&out)[3])
{
    static_assert(NT == 256, "16 row sums fill one row");
    __shared__ dd sh[3][16];
    const int tid = threadIdx.x + BX * threadIdx.y;
    const int lane = tid & 63, wave = tid >> 6;
#pragma unroll
    for (int s = 0; s < NS; ++s) {
        const dd w = row_sum(v[s]);
        if ((lane & 15) == 15) sh[s][wave * 4 + (lane >> 4)] = w;
    }
    lds_barrier();
    if (wave == 0) {
#pragma unroll
        for (int s = 0; s < NS; ++s) out[s] = row_sum(lane < 16 ? sh[s][lane] : dd{0.0, 0.0});
    }
}

__device__ __forceinline__ bool block_leader() { return threadIdx.x + BX * threadIdx.y == 15; }

// Reduce up to three per-thread sums over the workgroup and store the
// workgroup's partials.  Every thread must call it.
template <int NS>
__device__ __forceinline__ void block_partials(const dd (&v)[3], const RedBuf &rb)
{
    dd t[3];
    block_sum<NS>(v, t);
    if (block_leader()) {
        const unsigned bid = blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z);
#pragma unroll
        for (int s = 0; s < NS; ++s) {
            rb.partials[6 * (size_t)bid + 2 * s] = t[s].hi;
            rb.partials[6 * (size_t)bid + 2 * s + 1] = t[s].lo;
        }
    }
}

__device__ __forceinline__ bool halted(const KState *st) { return st != nullptr && st->status != ST_RUNNING; }

// ---------------------------------------------------------------------------
// conductances (src/NonUniform.jl:477-503 and the identical blocks of
// CalcRK! and rbsor_core!)
// ---------------------------------------------------------------------------
// λf(a, b, ma, mb) (src/common.jl:35) with mb = m0 = 1: the factor
// 2 - div(ma+1, 2) is 1 for an interior neighbour and 2 for a halo one.
template <typename T>
__host__ __device__ __forceinline__ T lamf(T a, T b, T ma)
{
    const T f = (ma == T(1)) ? T(1) : T(2);
    return T(2) * a * b / (a + b) * f;
}

template <typename T>
struct Cond {
    T xm, xp, ym, yp, zm, zp;            // cond_* (conduction, area/distance included)
    T cxm, cxp, cym, cyp, czm, czp;      // conv_* (convective boundary terms)
    T ap0;                               // a_p_0 (transient)
};

// dd = (1-m0) + (Σcond + Σconv + a_p_0)·m0 in the reference's order
template <typename T>
__device__ __forceinline__ T diag(const Cond<T> &o)
{
    return (T(1) - T(1)) + (o.xp + o.xm + o.yp + o.ym + o.zp + o.zm + o.cxp + o.cxm + o.cyp + o.cym + o.czp +
                            o.czm + o.ap0) *
                               T(1);
}



// ---------------------------------------------------------------------------
// scalar steps (one thread, after the sums are known)
// ---------------------------------------------------------------------------
__device__ __forceinline__ void fin_res0(double s0, KState *st)  // CalcRK! result; NonUniform.jl:263-272
{
    const double r0 = sqrt(s0);
    st->res0 = r0;
    st->itr = 0;
    if (r0 == 0.0) st->status = ST_CONVERGED;
    else if (!isfinite(r0)) st->status = ST_NAN;
}

__device__ __forceinline__ void fin_bicg_rho(double rho, KState *st)  // NonUniform.jl:287-300
{
    st->itr += 1;
    st->rho = rho;
    if (fabs(rho) < kFloatMin) {
        st->status = ST_BREAKDOWN;
        return;
    }
    if (st->itr > 1) st->beta = st->rho / st->rho_old * st->alpha / st->omega;
}

// Divergence guard of a guarded solve (no reference counterpart): returns
// false after stopping the solve.
__device__ __forceinline__ bool guard_step(double res, KState *st)
{
    st->improved = 0;
    if (!isfinite(res) || res > kDivergeFactor * st->res_min || st->since_best >= kStallIters) {
        st->status = ST_DIVERGED;
        return false;
    }
    if (res < st->res_min) {
        st->res_min = res;
        st->improved = 1;
        st->since_best = 0;
    } else {
        st->since_best += 1;
    }
    return true;
}

__device__ __forceinline__ void fin_bicg_res(double s0, KState *st, double *hist)  // NonUniform.jl:331-346
{
    double res = sqrt(s0);
    res /= st->res0;
    st->res = res;
    if (hist) hist[st->itr - 1] = res;
    if (st->guard && !guard_step(res, st)) return;
    if (!isfinite(res)) st->status = ST_NAN;
    else if (res < st->tol) st->status = ST_CONVERGED;
    else {
        st->rho_old = st->rho;
        if (st->itr >= st->maxItr) st->status = ST_MAXITER;
    }
}

__device__ __forceinline__ void finalize(int fin, const double *s, KState *st, double *hist)
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
        if (st->guard && !guard_step(res, st)) break;
        if (!isfinite(res)) st->status = ST_NAN;
        else if (res < st->tol) st->status = ST_CONVERGED;
        break;
    }
    case FIN_SOR_RES0: {  // solveSOR!, NonUniform.jl:1035-1038
        double r0 = sqrt(s[0]);
        if (r0 == 0.0) r0 = 1.0;
        st->res0 = r0;
        st->itr = 0;
        if (!isfinite(r0)) st->status = ST_NAN;
        break;
    }
    case FIN_SOR_RES: {  // rbsor!'s sqrt((0 + res_c0) + res_c1) / res0, NonUniform.jl:950-956, 1042-1049
        st->itr += 1;
        double res = sqrt(st->out[0] + s[0]);
        res /= st->res0;
        st->res = res;
        if (hist) hist[st->itr - 1] = res;
        if (!isfinite(res)) st->status = ST_NAN;
        else if (res < st->tol) st->status = ST_CONVERGED;
        else if (st->itr >= st->maxItr) st->status = ST_MAXITER;
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


}  // namespace h3d
