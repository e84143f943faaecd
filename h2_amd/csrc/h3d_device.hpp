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
// adds at most a few dozen terms, the tree above adds pairs; the pair keeps
// the sum to far below one ulp of the result, so the final rounding gives
// the exactly rounded sum of the terms whatever the decomposition.
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

// Reduce up to three per-thread sums over the workgroup and store the
// workgroup's partials.  Every thread must call it.
template <int NS>
__device__ __forceinline__ void block_partials(const dd (&v)[3], const RedBuf &rb)
{
    __shared__ dd sh[3][NWAVE];
    const int tid = threadIdx.x + BX * threadIdx.y;
    const int lane = tid & 63, wave = tid >> 6;
#pragma unroll
    for (int s = 0; s < NS; ++s) {
        const dd w = wave_sum(v[s]);
        if (lane == 0) sh[s][wave] = w;
    }
    __syncthreads();
    if (tid == 0) {
        const unsigned bid = blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z);
#pragma unroll
        for (int s = 0; s < NS; ++s) {
            dd b = sh[s][0];
#pragma unroll
            for (int w = 1; w < NWAVE; ++w) b = dd_add(b, sh[s][w]);
            rb.partials[6 * (size_t)bid + 2 * s] = b.hi;
            rb.partials[6 * (size_t)bid + 2 * s + 1] = b.lo;
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



}  // namespace h3d
