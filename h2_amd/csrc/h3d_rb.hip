// Kernels of the solver's inner loop on Krylov vectors in the colour-split
// ("RB") layout (layout and coefficient tables: RbGeo in h3d_internal.hpp).
//
// In the reference's column-major layout one red-black colour is every
// second double of a row, so a half-sweep moves whole cache lines of x and
// of the right-hand side for half of their bytes and writes back lines of
// which half the bytes changed.  In the RB layout the cells a colour
// updates are contiguous and their six neighbours are contiguous runs of
// the other half: every load and store of a wave is a dense, aligned run.
//
// Every cell performs exactly the operations of the standard-layout kernels
// (h3d_kernels.hip), in the same order, so the results are bitwise those of
// the reference's expressions; grid sums are exactly rounded and so do not
// depend on the layout either.  Terms the reference adds that are exact
// zeros (the convective terms of cells away from the boundary) are skipped,
// which does not change a sum's bits.
//
// Work decomposition: a workgroup is 64 (half-row index m) by 4 (rows j)
// threads of one plane (the two-colour sweep, k_rb_sweep2, marches a 64 x 8
// footprint through a chunk of planes instead).  "Pair" kernels handle the
// two cells i = 2m and 2m+1 of a thread (one in each half); the sweep
// handles the one cell of the colour being updated.  On large grids the
// sweep, CalcAX! and their sums take NC neighbouring cells of a half per
// thread instead, one row per wave (k_rb_sweep_w, k_rb_sweep_w2,
// k_rb_calc_ax_w).  Every thread computes from clamped, always
// in-bounds indices and only its stores and sums are predicated on the
// cell being an interior cell, so no branch separates a thread's loads and
// they are all in flight together.  Workgroups are renumbered so that each
// XCD (workgroups are dealt to the eight XCDs round-robin) receives a
// contiguous range of rows and planes, keeping the rows and planes
// neighbouring workgroups share in one L2.

#include "h3d_rb.hpp"

namespace h3d {

template <typename T>
void rb_shape(RbGeo<T> &g)
{
    g.gx = (g.HX + BX - 1) / BX;
    g.gy = (g.NY + BY - 1) / BY;
    const unsigned nb1 = (unsigned)((long)g.gx * g.gy * g.NZ);
    g.nb1 = (nb1 + 7u) & ~7u;
}
template void rb_shape<double>(RbGeo<double> &);
template void rb_shape<float>(RbGeo<float> &);

// ---------------------------------------------------------------------------
// layout conversion (set-up)
// ---------------------------------------------------------------------------
template <typename D>
__global__ void __launch_bounds__(256) k_to_rb(D *dst, const D *src, int MX, int MY, int MZ, int dir)
{
    const size_t H = (size_t)MX * MY * MZ;
    const size_t HX = (size_t)(MX + 1) / 2, HS = HX * MY * MZ;
    for (size_t c = blockIdx.x * (size_t)blockDim.x + threadIdx.x; c < H; c += (size_t)gridDim.x * blockDim.x) {
        const size_t i = c % MX, t = c / MX, j = t % MY, k = t / MY;
        const size_t h = (i + j + k) & 1;
        const size_t o = h * HS + (i >> 1) + HX * (j + MY * k);
        if (dir == 0) dst[o] = src[c];
        else dst[c] = src[o];
    }
}

template <typename D>
hipError_t launch_to_rb(D *dst, const D *src, int MX, int MY, int MZ, hipStream_t s)
{
    const size_t H = (size_t)MX * MY * MZ;
    const int nb = (int)std::min<size_t>((H + 255) / 256, 8192);
    hipLaunchKernelGGL(k_to_rb<D>, dim3(nb > 0 ? nb : 1), dim3(256), 0, s, dst, src, MX, MY, MZ, 0);
    return hipGetLastError();
}

template <typename D>
hipError_t launch_from_rb(D *dst, const D *src, int MX, int MY, int MZ, hipStream_t s)
{
    const size_t H = (size_t)MX * MY * MZ;
    const int nb = (int)std::min<size_t>((H + 255) / 256, 8192);
    hipLaunchKernelGGL(k_to_rb<D>, dim3(nb > 0 ? nb : 1), dim3(256), 0, s, dst, src, MX, MY, MZ, 1);
    return hipGetLastError();
}

// Pointers of the two halves of an RB array
template <typename T>
struct Halves {
    T *h[2];
};

template <typename T>
static Halves<T> halves(T *a, size_t HS)
{
    return Halves<T>{{a, a ? a + HS : nullptr}};
}

// ---------------------------------------------------------------------------
// red-black Gauss-Seidel
// ---------------------------------------------------------------------------
// Workgroup decode for kernels whose threads handle R rows each (rows j0,
// j0 + BY, ..., j0 + (R-1)·BY of one plane): gx x ceil(NY / (BY·R)) x NZ
// workgroups rounded up to a multiple of 8, renumbered per XCD as in
// rb_thread.  live = the workgroup and m are real; rows are checked per row.
struct RbRows {
    int m, j0, k;
    bool live;
};

template <int R, typename T>
__device__ __forceinline__ RbRows rb_rows(const RbGeo<T> &g)
{
    const unsigned gy = ((unsigned)g.NY + BY * R - 1) / (BY * R);
    const unsigned gz = (unsigned)g.NZ;
    const unsigned nreal = (unsigned)g.gx * gy * gz;
    const unsigned nb = (nreal + 7u) & ~7u;
    const unsigned p = blockIdx.x;
    const unsigned L = (nb >> 3) * (p & 7u) + (p >> 3);
    const unsigned bx = L % (unsigned)g.gx, t = L / (unsigned)g.gx;
    const unsigned bz = t % gz, by = t / gz;
    const int m = (int)bx * BX + threadIdx.x;
    RbRows r;
    r.live = L < nreal && m < g.HX;
    r.m = min(m, g.HX - 1);
    r.j0 = (int)by * BY * R + __builtin_amdgcn_readfirstlane((int)threadIdx.y) + 1;  // wave-uniform
    r.k = (int)bz + 1;
    return r;
}

template <int R, typename T>
static unsigned rb_rows_nb(const RbGeo<T> &g)
{
    const unsigned gy = ((unsigned)g.NY + BY * R - 1) / (BY * R);
    return ((unsigned)g.gx * gy * (unsigned)g.NZ + 7u) & ~7u;
}

// One colour of rbsor_core! (src/NonUniform.jl:848-915), ω = 1, in place:
// colour c updates half h = 1-c.  The kernel gets the two halves of x as
// separate (non-aliasing) pointers: xo = the half it updates, xt = the half
// it reads the neighbours from; bo = b's half h; clo/clt = the class
// field's halves.  Same arithmetic as gs_update().  Each thread updates one cell in each of R rows; all their
// loads are issued before the table staging.
template <typename T, int R, bool ZERO, bool DOT>
__global__ void __launch_bounds__(256) k_rb_sweep(T *__restrict__ xo, const T *__restrict__ xt,
                                                  const T *__restrict__ bo, const uint8_t *__restrict__ clo,
                                                  const uint8_t *__restrict__ clt, RbGeo<T> g, int h, const KState *st,
                                                  const T *__restrict__ bt, RedBuf rbuf)
{
    // small grids (the rows per thread they use): the loads go out before
    // the status test, whose latency they then hide
    constexpr bool early = R == (sizeof(T) == 8 ? 1 : 2);
    if (!early && halted(st)) return;
    const RbRows t = rb_rows<R>(g);
    const unsigned HX = (unsigned)g.HX, HXY = HX * (unsigned)g.MY;
    Nb<T> n[R];
    Cls q[R];
    T pp[R], bc[R], bp[R];
    int s[R], jj[R];
    unsigned rr[R];
#pragma unroll
    for (int a = 0; a < R; ++a) {
        jj[a] = min(t.j0 + a * BY, g.NY);
        s[a] = (h + jj[a] + t.k) & 1;
        rr[a] = (unsigned)t.m + HX * (unsigned)jj[a] + HXY * (unsigned)t.k;
        const unsigned rb = rr[a] * (unsigned)sizeof(T);
        n[a] = rb_nb(xt, rr[a], s[a], HX, HXY);
        pp[a] = ZERO ? T(0) : ldb(xo, rb);
        bc[a] = ldb(bo, rb);
        bp[a] = DOT ? ldb(bt, rb) : T(0);
        q[a] = rb_cls(clo, clt, rr[a], rr[a], s[a], HX, HXY);
    }
    if (early && halted(st)) return;
    const T *stab = rb_stage(g, t.k);
    T xv[R];
    bool valid[R];
    dd v[3] = {{0.0, 0.0}, {0.0, 0.0}, {0.0, 0.0}};
#pragma unroll
    for (int a = 0; a < R; ++a) {
        const int i = 2 * t.m + s[a];
        const bool rowok = t.live && t.j0 + a * BY <= g.NY;
        valid[a] = rowok && i >= 1 && i <= g.NX;
        const Cond<T> o = rb_cond(g, stab, q[a], i, jj[a], t.k);
        const T dd = diag(o);
        T xn;
        if (g.flux) {
            xn = pp[a] + (rb_flux(o, n[a], pp[a]) - bc[a]) / dd;
        } else {
            const T dp = (((rb_offdiag(o, n[a]) - bc[a]) / dd - pp[a])) * T(1);
            xn = pp[a] + T(1) * dp;
        }
        xv[a] = xn;
        if (DOT) {
            // this cell, and the other half's cell at the same offset
            // (i' = 2m + 1 - s, already loaded as the x-neighbour on that side)
            const int ip = 2 * t.m + 1 - s[a];
            const T xp = s[a] ? n[a].xm : n[a].xp;
            if (valid[a]) v[0] = dd_add(v[0], (double)bc[a] * (double)xn);
            if (rowok && ip >= 1 && ip <= g.NX) v[0] = dd_add(v[0], (double)bp[a] * (double)xp);
        }
    }
    // the workgroup sum before the stores (a store ahead of the reduction's
    // barrier would hold the wave until it completes)
    if (DOT) block_partials<1>(v, rbuf);
#pragma unroll
    for (int a = 0; a < R; ++a)
        if (valid[a]) stb(xo, rr[a] * (unsigned)sizeof(T), xv[a]);
}

// The per-colour sweep with two neighbouring cells of a row per thread
// (m = mb + 2l and mb + 2l + 1 of a 128-wide tile, one wave per row): each
// load of x, b and the class bytes covers both cells in one instruction
// (the cells are contiguous in every array), and the x neighbour that lies
// in the next thread's pair comes from that thread by a DPP lane shift, or
// for the wave's end thread from one more load.  Every cell's update is
// k_rb_sweep's, so the results are bitwise the same.
constexpr int kDppWaveShr1 = 0x138;  // lane l reads lane l - 1 (lane 0: old)
constexpr int kDppWaveShl1 = 0x130;  // lane l reads lane l + 1 (lane 63: old)

template <int CTRL, typename T>
__device__ __forceinline__ T dpp_shift(T v, T old)
{
    if constexpr (sizeof(T) == 8) {
        const unsigned long long u = __builtin_bit_cast(unsigned long long, v);
        const unsigned long long o = __builtin_bit_cast(unsigned long long, old);
        const unsigned lo = (unsigned)__builtin_amdgcn_update_dpp((int)(unsigned)o, (int)(unsigned)u, CTRL, 0xf, 0xf,
                                                                  false);
        const unsigned hi = (unsigned)__builtin_amdgcn_update_dpp((int)(unsigned)(o >> 32), (int)(unsigned)(u >> 32),
                                                                  CTRL, 0xf, 0xf, false);
        return __builtin_bit_cast(T, ((unsigned long long)hi << 32) | lo);
    } else {
        return __builtin_bit_cast(T, __builtin_amdgcn_update_dpp(__builtin_bit_cast(int, old),
                                                                 __builtin_bit_cast(int, v), CTRL, 0xf, 0xf, false));
    }
}

// elements e and e + 1 (alignment: one element)
template <typename T>
__device__ __forceinline__ void ld_pair(const T *__restrict__ p, unsigned e, T &a, T &b)
{
    typedef T v2 __attribute__((ext_vector_type(2), aligned(sizeof(T))));
    const v2 v = *reinterpret_cast<const v2 *>(p + e);
    a = v.x;
    b = v.y;
}

template <typename T>
__device__ __forceinline__ void st_pair(T *__restrict__ p, unsigned e, T a, T b)
{
    typedef T v2 __attribute__((ext_vector_type(2), aligned(sizeof(T))));
    v2 v;
    v.x = a;
    v.y = b;
    *reinterpret_cast<v2 *>(p + e) = v;
}

__device__ __forceinline__ void ld_cls_pair(const uint8_t *__restrict__ p, unsigned e, int &a, int &b)
{
    uint16_t v;
    __builtin_memcpy(&v, p + e, 2);
    a = v & 255;
    b = v >> 8;
}

template <typename T>
__device__ __forceinline__ T gs_cell(const RbGeo<T> &g, const T *stab, const Cls &q, const Nb<T> &n, T pp, T bc,
                                     int i, int j, int k)
{
    const Cond<T> o = rb_cond(g, stab, q, i, j, k);
    const T dd = diag(o);
    if (g.flux) return pp + (rb_flux(o, n, pp) - bc) / dd;
    const T dp = (((rb_offdiag(o, n) - bc) / dd - pp)) * T(1);
    return pp + T(1) * dp;
}

// NC elements from e on (alignment: one element)
template <int NC, typename T>
__device__ __forceinline__ void ld_vec(const T *__restrict__ p, unsigned e, T (&o)[NC])
{
    typedef T vn __attribute__((ext_vector_type(NC), aligned(sizeof(T))));
    const vn v = *reinterpret_cast<const vn *>(p + e);
#pragma unroll
    for (int c = 0; c < NC; ++c) o[c] = v[c];
}

template <int NC, typename T>
__device__ __forceinline__ void st_vec(T *__restrict__ p, unsigned e, const T (&a)[NC])
{
    typedef T vn __attribute__((ext_vector_type(NC), aligned(sizeof(T))));
    vn v;
#pragma unroll
    for (int c = 0; c < NC; ++c) v[c] = a[c];
    *reinterpret_cast<vn *>(p + e) = v;
}

template <int NC>
__device__ __forceinline__ void ld_cls_vec(const uint8_t *__restrict__ p, unsigned e, int (&o)[NC])
{
    if constexpr (NC == 4) {
        uint32_t v;
        __builtin_memcpy(&v, p + e, 4);
#pragma unroll
        for (int c = 0; c < 4; ++c) o[c] = (int)((v >> (8 * c)) & 255u);
    } else {
        uint16_t v;
        __builtin_memcpy(&v, p + e, 2);
        o[0] = v & 255;
        o[1] = v >> 8;
    }
}

// The per-colour sweep with NC neighbouring cells of a row per thread
// (m = mb + NC·l ... of an NC·64-wide tile, one wave per row): each load of
// x, b and the class bytes covers the NC cells in one instruction (they are
// contiguous in every array), and the x neighbour that lies in the next
// thread's cells comes from that thread by a DPP lane shift, or for the
// wave's end thread from one more load.  Every cell's update is
// k_rb_sweep's, so the results are bitwise the same.  DOT: also Σ b·x over
// both halves (the pass of launch_rb_sweep_dot): the thread's new cells
// and the other half's cells at the same offsets (its x-neighbour values),
// with bt = b's other half.
template <typename T, int NC, bool ZERO, bool DOT = false>
__global__ void __launch_bounds__(256) k_rb_sweep_w(T *__restrict__ xo, const T *__restrict__ xt,
                                                    const T *__restrict__ bo, const uint8_t *__restrict__ clo,
                                                    const uint8_t *__restrict__ clt, RbGeo<T> g, int h,
                                                    const KState *st, const T *__restrict__ bt = nullptr,
                                                    RedBuf rbuf = RedBuf{})
{
    if (halted(st)) return;
    constexpr int TW = NC * BX;  // the tile's half-row indices
    const unsigned gx = ((unsigned)g.HX + TW - 1) / TW, gy = ((unsigned)g.NY + BY - 1) / BY;
    const unsigned gz = (unsigned)g.NZ, nreal = gx * gy * gz, nb = (nreal + 7u) & ~7u;
    const unsigned p = blockIdx.x;
    const unsigned L = (nb >> 3) * (p & 7u) + (p >> 3);
    const unsigned bx = L % gx, t1 = L / gx, bz = t1 % gz, by = t1 / gz;
    const int HXi = g.HX;
    const int mb = (int)bx * TW, m0 = mb + NC * (int)threadIdx.x;
    // one row per wave: its row index is wave-uniform (scalar registers)
    const int jr = (int)by * BY + __builtin_amdgcn_readfirstlane((int)threadIdx.y) + 1;
    const int k = (int)min(bz, gz - 1u) + 1;
    const bool rowok = L < nreal && jr <= g.NY;
    const int j = min(jr, g.NY);
    // the cells mc .. mc + NC - 1: past the row's end they run into the next
    // row, which is in the array for every row and plane the loads touch
    const int mc = min(m0, HXi - 1);
    const unsigned HX = (unsigned)HXi, HXY = HX * (unsigned)g.MY;
    const unsigned e = (unsigned)mc + HX * (unsigned)j + HXY * (unsigned)k;
    const int s = (h + j + k) & 1;
    // the x neighbour outside the wave's cells: m = mb - 1 for the first
    // thread (s = 0), mb + TW for the last (s = 1), clamped into the row
    // (a clamped one belongs to a cell outside the interior)
    const int me = s ? min(mb + TW, HXi - 1) : max(mb - 1, 0);
    const unsigned ee = (unsigned)me + HX * (unsigned)j + HXY * (unsigned)k;
    T P[NC], ym[NC], yp[NC], zm[NC], zp[NC], pp[NC], bv[NC], bp[NC];
    ld_vec<NC>(xt, e, P);
    ld_vec<NC>(xt, e - HX, ym);
    ld_vec<NC>(xt, e + HX, yp);
    ld_vec<NC>(xt, e - HXY, zm);
    ld_vec<NC>(xt, e + HXY, zp);
    if (ZERO) {
#pragma unroll
        for (int c = 0; c < NC; ++c) pp[c] = T(0);
    } else {
        ld_vec<NC>(xo, e, pp);
    }
    ld_vec<NC>(bo, e, bv);
    if (DOT) ld_vec<NC>(bt, e, bp);
    const T xe = xt[ee];
    int co[NC], cn[NC], cym[NC], cyp[NC], czm[NC], czp[NC];
    ld_cls_vec<NC>(clo, e, co);
    ld_cls_vec<NC>(clt, e, cn);
    ld_cls_vec<NC>(clt, e - HX, cym);
    ld_cls_vec<NC>(clt, e + HX, cyp);
    ld_cls_vec<NC>(clt, e - HXY, czm);
    ld_cls_vec<NC>(clt, e + HXY, czp);
    const int ce = clt[ee];
    const T *stab = rb_stage(g, k);
    // x neighbours of cell c: other[m0 + c - 1 + s] and other[m0 + c + s];
    // the ones past the thread's cells: other[m0 - 1] from the thread
    // before, other[m0 + NC] from the thread after
    const T sr = dpp_shift<kDppWaveShr1>(P[NC - 1], xe);
    const T sl = dpp_shift<kDppWaveShl1>(P[0], xe);
    const int csr = dpp_shift<kDppWaveShr1>(cn[NC - 1], ce), csl = dpp_shift<kDppWaveShl1>(cn[0], ce);
    Nb<T> n[NC];
    Cls q[NC];
#pragma unroll
    for (int c = 0; c < NC; ++c) {
        const T lo = c > 0 ? P[c - 1] : sr, hi = c < NC - 1 ? P[c + 1] : sl;
        const int clo_ = c > 0 ? cn[c - 1] : csr, chi = c < NC - 1 ? cn[c + 1] : csl;
        n[c].xm = s ? P[c] : lo;
        n[c].xp = s ? hi : P[c];
        n[c].ym = ym[c];
        n[c].yp = yp[c];
        n[c].zm = zm[c];
        n[c].zp = zp[c];
        q[c].c = co[c];
        q[c].xm = s ? cn[c] : clo_;
        q[c].xp = s ? chi : cn[c];
        q[c].ym = cym[c];
        q[c].yp = cyp[c];
        q[c].zm = czm[c];
        q[c].zp = czp[c];
    }
    T xn[NC];
    bool valid[NC];
#pragma unroll
    for (int c = 0; c < NC; ++c) {
        const int i = 2 * (m0 + c) + s;
        valid[c] = rowok && m0 + c < HXi && i >= 1 && i <= g.NX;
    }
#pragma unroll
    for (int c = 0; c < NC; ++c) xn[c] = gs_cell(g, stab, q[c], n[c], pp[c], bv[c], 2 * (m0 + c) + s, j, k);
    if (DOT) {
        // the other half's cell at offset m0 + c is i = 2(m0 + c) + 1 - s
        dd v[3] = {{0.0, 0.0}, {0.0, 0.0}, {0.0, 0.0}};
#pragma unroll
        for (int c = 0; c < NC; ++c) {
            const int io = 2 * (m0 + c) + 1 - s;
            if (valid[c]) v[0] = dd_add(v[0], (double)bv[c] * (double)xn[c]);
            if (rowok && m0 + c < HXi && io >= 1 && io <= g.NX) v[0] = dd_add(v[0], (double)bp[c] * (double)P[c]);
        }
        // the workgroup sum before the stores, as in k_rb_sweep
        block_partials<1>(v, rbuf);
    }
    bool all = true;
#pragma unroll
    for (int c = 0; c < NC; ++c) all = all && valid[c];
    if (all) {
        st_vec<NC>(xo, e, xn);
    } else {
#pragma unroll
        for (int c = 0; c < NC; ++c)
            if (valid[c]) xo[e + (unsigned)c] = xn[c];
    }
}

// The FP64 form of k_rb_sweep_w (NC = 2) written out for the two cells: the
// same loads, lane shifts and updates, but the compiler allocates 80 VGPRs
// for it against 88 for the generic form with NC = 2 (six waves per SIMD
// against five), which measured 2-3 % slower (77.8 against 75.4 us per
// full sweep at 480x480x60, 890 against 881 us at 960x960x120).
template <typename T, bool ZERO, bool DOT = false>
__global__ void __launch_bounds__(256) k_rb_sweep_w2(T *__restrict__ xo, const T *__restrict__ xt,
                                                     const T *__restrict__ bo, const uint8_t *__restrict__ clo,
                                                     const uint8_t *__restrict__ clt, RbGeo<T> g, int h,
                                                     const KState *st, const T *__restrict__ bt = nullptr,
                                                     RedBuf rbuf = RedBuf{})
{
    if (halted(st)) return;
    constexpr int TW = 2 * BX;  // the tile's half-row indices
    const unsigned gx = ((unsigned)g.HX + TW - 1) / TW, gy = ((unsigned)g.NY + BY - 1) / BY;
    const unsigned gz = (unsigned)g.NZ, nreal = gx * gy * gz, nb = (nreal + 7u) & ~7u;
    const unsigned p = blockIdx.x;
    const unsigned L = (nb >> 3) * (p & 7u) + (p >> 3);
    const unsigned bx = L % gx, t1 = L / gx, bz = t1 % gz, by = t1 / gz;
    const int HXi = g.HX;
    const int mb = (int)bx * TW, m0 = mb + 2 * (int)threadIdx.x;
    // one row per wave: its row index is wave-uniform (scalar registers)
    const int jr = (int)by * BY + __builtin_amdgcn_readfirstlane((int)threadIdx.y) + 1;
    const int k = (int)min(bz, gz - 1u) + 1;
    const bool rowok = L < nreal && jr <= g.NY;
    const int j = min(jr, g.NY);
    // the pair (mc, mc + 1): mc + 1 may be the next row's first element,
    // which is in the array for every row and plane the loads touch
    const int mc = min(m0, HXi - 1);
    const unsigned HX = (unsigned)HXi, HXY = HX * (unsigned)g.MY;
    const unsigned e = (unsigned)mc + HX * (unsigned)j + HXY * (unsigned)k;
    const int s = (h + j + k) & 1;
    // the x neighbour outside the wave's pairs: m = mb - 1 for the first
    // thread (s = 0), mb + 128 for the last (s = 1), clamped into the row
    // (a clamped one belongs to a cell outside the interior)
    const int me = s ? min(mb + TW, HXi - 1) : max(mb - 1, 0);
    const unsigned ee = (unsigned)me + HX * (unsigned)j + HXY * (unsigned)k;
    T p0, p1, ym0, ym1, yp0, yp1, zm0, zm1, zp0, zp1, pp0 = T(0), pp1 = T(0), b0, b1;
    ld_pair(xt, e, p0, p1);
    ld_pair(xt, e - HX, ym0, ym1);
    ld_pair(xt, e + HX, yp0, yp1);
    ld_pair(xt, e - HXY, zm0, zm1);
    ld_pair(xt, e + HXY, zp0, zp1);
    if (!ZERO) ld_pair(xo, e, pp0, pp1);
    ld_pair(bo, e, b0, b1);
    T bp0 = T(0), bp1 = T(0);
    if (DOT) ld_pair(bt, e, bp0, bp1);
    const T xe = xt[ee];
    int c0, c1, n0, n1, cym0, cym1, cyp0, cyp1, czm0, czm1, czp0, czp1;
    ld_cls_pair(clo, e, c0, c1);
    ld_cls_pair(clt, e, n0, n1);
    ld_cls_pair(clt, e - HX, cym0, cym1);
    ld_cls_pair(clt, e + HX, cyp0, cyp1);
    ld_cls_pair(clt, e - HXY, czm0, czm1);
    ld_cls_pair(clt, e + HXY, czp0, czp1);
    const int ce = clt[ee];
    const T *stab = rb_stage(g, k);
    // x neighbours: cell A = m0 has other[m0 - 1 + s], other[m0 + s], cell
    // B = m0 + 1 has other[m0 + s], other[m0 + 1 + s]
    const T sr = dpp_shift<kDppWaveShr1>(p1, xe);  // other[m0 - 1]
    const T sl = dpp_shift<kDppWaveShl1>(p0, xe);  // other[m0 + 2]
    const int csr = dpp_shift<kDppWaveShr1>(n1, ce), csl = dpp_shift<kDppWaveShl1>(n0, ce);
    Nb<T> na, nbb;
    Cls qa, qb;
    na.xm = s ? p0 : sr;
    na.xp = s ? p1 : p0;
    nbb.xm = s ? p1 : p0;
    nbb.xp = s ? sl : p1;
    na.ym = ym0;
    nbb.ym = ym1;
    na.yp = yp0;
    nbb.yp = yp1;
    na.zm = zm0;
    nbb.zm = zm1;
    na.zp = zp0;
    nbb.zp = zp1;
    qa.c = c0;
    qb.c = c1;
    qa.xm = s ? n0 : csr;
    qa.xp = s ? n1 : n0;
    qb.xm = s ? n1 : n0;
    qb.xp = s ? csl : n1;
    qa.ym = cym0;
    qb.ym = cym1;
    qa.yp = cyp0;
    qb.yp = cyp1;
    qa.zm = czm0;
    qb.zm = czm1;
    qa.zp = czp0;
    qb.zp = czp1;
    const int ia = 2 * m0 + s, ib = ia + 2;
    const bool va = rowok && m0 < HXi && ia >= 1 && ia <= g.NX;
    const bool vb = rowok && m0 + 1 < HXi && ib >= 1 && ib <= g.NX;
    const T xa = gs_cell(g, stab, qa, na, pp0, b0, ia, j, k);
    const T xb = gs_cell(g, stab, qb, nbb, pp1, b1, ib, j, k);
    if (DOT) {
        // the other half's cells at offsets m0, m0 + 1: i = 2m0 + 1 - s, + 2
        const int ja = 2 * m0 + 1 - s, jb2 = ja + 2;
        dd v[3] = {{0.0, 0.0}, {0.0, 0.0}, {0.0, 0.0}};
        if (va) v[0] = dd_add(v[0], (double)b0 * (double)xa);
        if (rowok && m0 < HXi && ja >= 1 && ja <= g.NX) v[0] = dd_add(v[0], (double)bp0 * (double)p0);
        if (vb) v[0] = dd_add(v[0], (double)b1 * (double)xb);
        if (rowok && m0 + 1 < HXi && jb2 >= 1 && jb2 <= g.NX) v[0] = dd_add(v[0], (double)bp1 * (double)p1);
        // the workgroup sum before the stores, as in k_rb_sweep
        block_partials<1>(v, rbuf);
    }
    if (va && vb) {
        st_pair(xo, e, xa, xb);
    } else {
        if (va) xo[e] = xa;
        if (vb) xo[e + 1] = xb;
    }
}

// The large grids' sweep and CalcAX!: the NC-cell form (kSweepWide), NC
// cells of a half per thread (one 16-byte load: two FP64 or four FP32
// values; FP32 with two cells measured slower than three rows per thread,
// 58.1 against 51.5 us at 480x480x60), or the rows form.
template <typename T>
constexpr bool kSweepWide = true;
template <typename T>
constexpr int kSweepCells = sizeof(T) == 8 ? 2 : 4;

template <typename T, int NC = 2>
static unsigned rb_wide_nb(const RbGeo<T> &g)
{
    const unsigned gx = ((unsigned)g.HX + NC * BX - 1) / (NC * BX), gy = ((unsigned)g.NY + BY - 1) / BY;
    return (gx * gy * (unsigned)g.NZ + 7u) & ~7u;
}

// Rows per thread of the sweep, by precision and grid size (measured, HIP
// events, one full sweep): FP64 one row on small grids (17.7 against
// 18.4 us at 240x240x30), two on large ones (91.0 against 95.8 us at
// 480x480x60, 1025 against 1078 us at 960x960x120; three rows, at 124
// VGPRs, 86 against 82 us at 480x480x60); FP32 two rows on small grids
// (12.2 against 13.4 us with four), three on large ones (51.9 against
// 58.4 us with four at 480x480x60, 548 against 615 us at 960x960x120:
// 73-77 VGPRs keep six waves per SIMD, four rows' 97 only four).  Large:
// at least 2^21 cells per half.  On large grids the sweep runs in the
// NC-cell form (k_rb_sweep_w) while kSweepWide holds; the rows form stays
// for the small grids and as the large grids' alternative.
template <typename T>
constexpr int rb_rows_large()
{
    return sizeof(T) == 4 ? 3 : 2;
}

template <typename T, int R>
static void rb_sweep_rows(T *x, const T *b, const RbGeo<T> &g, int h, bool from_zero, const KState *st,
                          hipStream_t s)
{
    const size_t own = h ? g.HS : 0, oth = h ? 0 : g.HS;
    const RedBuf none{};
    if (from_zero)
        hipLaunchKernelGGL((k_rb_sweep<T, R, true, false>), dim3(rb_rows_nb<R>(g)), dim3(BX, BY), rb_lds(g), s,
                           x + own, x + oth, b + own, g.cls + own, g.cls + oth, g, h, st, nullptr, none);
    else
        hipLaunchKernelGGL((k_rb_sweep<T, R, false, false>), dim3(rb_rows_nb<R>(g)), dim3(BX, BY), rb_lds(g), s,
                           x + own, x + oth, b + own, g.cls + own, g.cls + oth, g, h, st, nullptr, none);
}

// from_zero: the colour's own cells of x are zero (the first colour-1 pass
// of a preconditioner that starts from x = 0); the kernel takes them as 0
// instead of reading them, so they need not have been written.
template <typename T>
hipError_t launch_rb_sweep(T *x, const T *b, const RbGeo<T> &g, int color, const KState *st, hipStream_t s,
                           bool from_zero)
{
    constexpr int R = sizeof(T) == 4 ? 2 : 1;
    const int h = 1 - color;
    // large grids: the NC-cell form (measured, HIP events, both colours:
    // FP64 two cells 75-78 against 84.0 us at 480x480x60 and 881 against
    // 1019 us at 960x960x120 for two rows per thread; FP32 four cells 47.0
    // against 51.5 us and 467 against 536 us for three rows)
    if (kSweepWide<T> && g.HS >= ((size_t)1 << 21)) {
        constexpr int NC = kSweepCells<T>;
        const size_t own = h ? g.HS : 0, oth = h ? 0 : g.HS;
        const unsigned nb = rb_wide_nb<T, NC>(g);
        if constexpr (NC == 2) {
            if (from_zero)
                hipLaunchKernelGGL((k_rb_sweep_w2<T, true>), dim3(nb), dim3(BX, BY), rb_lds(g), s, x + own, x + oth,
                                   b + own, g.cls + own, g.cls + oth, g, h, st);
            else
                hipLaunchKernelGGL((k_rb_sweep_w2<T, false>), dim3(nb), dim3(BX, BY), rb_lds(g), s, x + own,
                                   x + oth, b + own, g.cls + own, g.cls + oth, g, h, st);
        } else {
            if (from_zero)
                hipLaunchKernelGGL((k_rb_sweep_w<T, NC, true>), dim3(nb), dim3(BX, BY), rb_lds(g), s, x + own,
                                   x + oth, b + own, g.cls + own, g.cls + oth, g, h, st);
            else
                hipLaunchKernelGGL((k_rb_sweep_w<T, NC, false>), dim3(nb), dim3(BX, BY), rb_lds(g), s, x + own,
                                   x + oth, b + own, g.cls + own, g.cls + oth, g, h, st);
        }
    } else if (g.HS >= ((size_t)1 << 21)) {
        rb_sweep_rows<T, rb_rows_large<T>()>(x, b, g, h, from_zero, st, s);
    } else {
        rb_sweep_rows<T, R>(x, b, g, h, from_zero, st, s);
    }
    return hipGetLastError();
}

// Colour 1 of the preconditioner's last sweep (not from zero) with
// Σ b·x over every interior cell of both halves, finalized with fin: CG!'s
// ρ = (r, z) when b is r (NonUniform.jl:205).  The colour-0 half of x is
// final before this pass, so each thread adds its own cell (new value) and
// the colour-0 cell at the same offset in the other half, which it has
// already loaded as a neighbour.  The sum is exactly rounded, so it is
// bitwise the separate dot kernel's.
template <typename T>
hipError_t launch_rb_sweep_dot(T *x, const T *b, const RbGeo<T> &g, int fin, KState *st, double *hist,
                               const RedBuf &rb, hipStream_t s)
{
    constexpr int R0 = sizeof(T) == 4 ? 2 : 1;
    const int h = 0;  // colour 1
    const size_t own = 0, oth = g.HS;
    unsigned nb;
    if (kSweepWide<T> && g.HS >= ((size_t)1 << 21)) {
        constexpr int NC = kSweepCells<T>;
        nb = rb_wide_nb<T, NC>(g);
        if constexpr (NC == 2)
            hipLaunchKernelGGL((k_rb_sweep_w2<T, false, true>), dim3(nb), dim3(BX, BY), rb_lds(g), s, x + own,
                               x + oth, b + own, g.cls + own, g.cls + oth, g, h, st, b + oth, rb);
        else
            hipLaunchKernelGGL((k_rb_sweep_w<T, NC, false, true>), dim3(nb), dim3(BX, BY), rb_lds(g), s, x + own,
                               x + oth, b + own, g.cls + own, g.cls + oth, g, h, st, b + oth, rb);
    } else if (g.HS >= ((size_t)1 << 21)) {
        constexpr int RL = rb_rows_large<T>();
        nb = rb_rows_nb<RL>(g);
        hipLaunchKernelGGL((k_rb_sweep<T, RL, false, true>), dim3(nb), dim3(BX, BY), rb_lds(g), s, x + own,
                           x + oth, b + own, g.cls + own, g.cls + oth, g, h, st, b + oth, rb);
    } else {
        nb = rb_rows_nb<R0>(g);
        hipLaunchKernelGGL((k_rb_sweep<T, R0, false, true>), dim3(nb), dim3(BX, BY), rb_lds(g), s, x + own,
                           x + oth, b + own, g.cls + own, g.cls + oth, g, h, st, b + oth, rb);
    }
    hipError_t e = hipGetLastError();
    return e != hipSuccess ? e : launch_finalize_sums<T>(1, nb, fin, st, hist, rb, s, seq_terms<T>(g, b, x));
}

// ---------------------------------------------------------------------------
// Two colours of rbsor_core! in one pass: half hA, then half hB = 1 - hA,
// from xin into xout (whole RB arrays, both halves; their halos must be
// equal, as the zero halos of the Krylov vectors are).
//
// A workgroup of S2X x S2Y (64 x 8) threads covers a footprint of S2X
// half-row indices by S2Y rows and owns its inner (S2X-2) x (S2Y-2) cells
// over a chunk of planes [kb, ke], which it marches in z.  At step ka (from kb-1
// to ke+1) every thread updates half A at its (m, j, ka) into the LDS plane
// slot of ka; after the barrier the owned threads update half B at plane
// ka-1 from the A planes ka-2, ka-1, ka in LDS, and store both halves of
// plane ka-1.  The footprint's ring of A cells (one cell wide in m, j and,
// through the chunk's first and last steps, k) belongs to neighbouring
// workgroups, which compute the same values with the same operations.
//
// Every cell's update is k_rb_sweep's, so the result is bitwise that of the
// two in-place colour passes.  The pass is out of place because a ring cell
// must be computed from the old values its owner overwrites.  Compared with
// the two passes it reads x and the class field once instead of twice and
// launches once; it recomputes the ring, about (S2X·S2Y)/((S2X-2)(S2Y-2))
// × (C+2)/C of the A updates for a chunk of C planes.  A 64 x 16 footprint
// (fewer ring cells, one 16-wave workgroup per CU) measured slower than
// 64 x 8 (two 8-wave workgroups per CU, whose lockstep steps interleave).
//
// Four LDS slots (plane index & 3) for the A planes, their class bytes and
// the conductance tables let one barrier per step order everything: a slot
// written at step ka was last read at step ka-2, before the barrier of step
// ka-1.
constexpr int S2X = 64, S2Y = 8, S2OX = S2X - 2, S2OY = S2Y - 2;
// longest z-chunk; its planes kb-1..ke+1 keep their ΔZ in LDS, so the loop
// issues no global load besides the prefetched ones (a vector load's wait
// would also wait for them)
constexpr int S2CMAX = kSweep2ChunkMax, S2ZMAX = S2CMAX + 3;

template <typename T>
struct S2Loads {
    Nb<T> n;
    T ppA, bcA, ppB, bcB;
    T zf;  // this thread's z-face table entry of the plane after
    Cls q;
    int sA, clA, clB;
};

// rb_stage_into for a workgroup of at least 2K² threads, each holding its
// λf table entry f (plane-independent), its z-face entry zf of plane k and,
// for a transient operator, capc = ρ·cp of class tid (tid < K)
// dz_k = ΔZ[k] and dzd_k = its FP64 value are read by the caller outside
// any divergent branch (scalar loads, which do not wait for the vector
// loads in flight).
template <typename T>
__device__ __forceinline__ void rb_stage_regs(const RbGeo<T> &g, T *t, int tid, T f, T zf, double capc, T dz_k,
                                              double dzd_k)
{
    const int n = 2 * g.K * g.K;
    if (tid < n) {
        t[tid] = f * g.dy0 * dz_k * g.ddx;
        t[n + tid] = f * g.dx0 * dz_k * g.ddy;
        t[2 * n + tid] = zf;
    }
    if (g.ap && tid < g.K) t[3 * n + tid] = (T)(capc * g.dx0d * g.dy0d * dzd_k * g.ddt);
}

template <typename T>
static size_t rb_sweep2_lds(const RbGeo<T> &g)
{
    return S2ZMAX * (sizeof(double) + sizeof(T)) + 4 * rb_table_entries(g) * sizeof(T) +
           4 * (size_t)(S2X * S2Y) * (sizeof(T) + 1);
}

template <typename T>
__global__ void __launch_bounds__(S2X *S2Y) k_rb_sweep2(const T *__restrict__ xin, T *__restrict__ xout,
                                                         const T *__restrict__ b, const uint8_t *__restrict__ cls,
                                                         RbGeo<T> g, int hA, int C, unsigned nbx, unsigned nby,
                                                         unsigned nch, const KState *st)
{
    if (halted(st)) return;
    const unsigned nreal = nbx * nby * nch;
    const unsigned nb = (nreal + 7u) & ~7u;
    const unsigned p = blockIdx.x;
    const unsigned L = (nb >> 3) * (p & 7u) + (p >> 3);
    if (L >= nreal) return;  // the whole workgroup
    // m-tile fastest, then the row tile, then the chunk: an XCD's workgroups
    // in flight are neighbours in m and j and share their rings in its L2
    const unsigned tm = L % nbx, t1 = L / nbx, tj = t1 % nby, ch = t1 / nby;

    extern __shared__ __align__(16) unsigned char h3d_dyn_smem[];
    const int TB = 6 * g.K * g.K + g.K;
    double *zd = reinterpret_cast<double *>(h3d_dyn_smem);  // [S2ZMAX] FP64 ΔZ of planes kb-1..
    T *zt = reinterpret_cast<T *>(zd + S2ZMAX);              // [S2ZMAX] ΔZ in T
    T *tabs = zt + S2ZMAX;                                   // [4][TB]
    T *as = tabs + 4 * TB;                                   // [4][S2Y][S2X]
    uint8_t *acl = reinterpret_cast<uint8_t *>(as + 4 * S2X * S2Y);

    const int tx = threadIdx.x, ty = threadIdx.y, tid = tx + S2X * ty;
    const int HX = g.HX, MY = g.MY, NX = g.NX, NY = g.NY, NZ = g.NZ;
    const int mr = (int)tm * S2OX - 1 + tx, jr = (int)tj * S2OY + ty;
    const int mc = min(max(mr, 0), HX - 1), jc = min(jr, MY - 1);
    const bool col = (mr == mc) && (jr == jc);
    const bool own = col && tx >= 1 && tx <= S2OX && ty >= 1 && ty <= S2OY;
    // rows of the loads: a neighbour of any row in 1..NY is in the array
    const int jl = min(max(jc, 1), NY);
    const int kb = 1 + (int)ch * C, ke = min(kb + C - 1, NZ);
    const int hB = 1 - hA;
    const unsigned HXu = (unsigned)HX, HXY = HXu * (unsigned)MY;
    const size_t oA = hA ? g.HS : 0, oB = hB ? g.HS : 0;
    const T *xA = xin + oA, *xB = xin + oB, *bA = b + oA, *bB = b + oB;
    T *yA = xout + oA, *yB = xout + oB;
    const uint8_t *cA = g.cls + oA, *cB = g.cls + oB;
    const int slot0 = tid;  // (ty, tx) within a plane slot

    // table entries this thread stages (2K² <= S2X·S2Y: the host keeps K
    // within the LDS budget, which is far smaller)
    const int KK = g.K * g.K, ntab = min(tid, 2 * KK - 1);
    // the global loads of step ka: the A cell at plane ka (at its load
    // position (mc, jl, kl), whose neighbours are all in the array, and at
    // its true position for the value kept where it is not updated) and
    // the B cell at plane ka-1; issued one step ahead of their use
    auto load = [&](int ka, S2Loads<T> &v) {
        const int kl = min(max(ka, 1), NZ);
        v.sA = (hA + jl + kl) & 1;
        // (the loop's last steps load planes past the chunk, clamped into
        // the array, and use nothing of them)
        const unsigned ra = (unsigned)mc + HXu * (unsigned)jc + HXY * (unsigned)min(ka, g.MZ - 1);
        const unsigned rl = (unsigned)mc + HXu * (unsigned)jl + HXY * (unsigned)kl;
        v.n = rb_nb(xB, rl, v.sA, HXu, HXY);
        v.ppA = xA[ra];
        v.bcA = bA[rl];
        v.q = rb_cls(cA, cB, rl, rl, v.sA, HXu, HXY);
        v.clA = cA[ra];
        const unsigned rb_ = (unsigned)mc + HXu * (unsigned)jl + HXY * (unsigned)min(max(ka - 1, 1), NZ);
        v.ppB = xB[rb_];
        v.bcB = bB[rb_];
        v.clB = cB[rb_];
        v.zf = g.ZF[(size_t)(min(ka + 1, NZ) - 1) * KK + ntab];
    };
    const T ftab = g.tab[ntab];
    const double capc = g.ap ? g.capc[min(tid, g.K - 1)] : 0.0;
    if (tid <= ke - kb + 2) {
        zd[tid] = g.dZd[kb - 1 + tid];
        zt[tid] = g.dZ[kb - 1 + tid];
    }
    S2Loads<T> v0, v1;
    load(kb - 1, v0);
    lds_barrier();
    if (kb - 1 >= 1)
        rb_stage_regs(g, tabs + ((kb - 1) & 3) * TB, tid, ftab, g.ZF[(size_t)(kb - 2) * KK + ntab], capc, zt[0],
                      zd[0]);
    lds_barrier();
    // step ka uses cur (loaded one step earlier) and loads step ka+1 into
    // nxt; the loop alternates the two register sets instead of copying
    // (a copy would wait for the loads in flight)
    auto step = [&](int ka, const S2Loads<T> &cur, S2Loads<T> &nxt) {
        const int kt = min(ka + 1, min(NZ, ke + 1));  // the plane whose table this step stages
        const T dz_t = zt[kt - (kb - 1)];
        const double dzd_t = zd[kt - (kb - 1)];
        load(ka + 1, nxt);  // unconditional: a skipped load would make later waits drain everything
        const int kl = min(max(ka, 1), NZ);
        const int sA = cur.sA;
        const Nb<T> n = cur.n;
        const T ppA = cur.ppA, bcA = cur.bcA, ppB = cur.ppB, bcB = cur.bcB;
        const Cls q = cur.q;
        const uint8_t clA = (uint8_t)cur.clA;
        const int clB = cur.clB;
        const int kB = ka - 1;
        const bool doB = own && kB >= kb && kB <= ke;
        const unsigned rb_ = (unsigned)mc + HXu * (unsigned)jl + HXY * (unsigned)min(max(kB, 1), NZ);
        // ---- the table of plane ka+1, used from the next step on
        if (ka + 1 >= 1 && ka + 1 <= NZ && ka + 1 <= ke + 1)
            rb_stage_regs(g, tabs + ((ka + 1) & 3) * TB, tid, ftab, cur.zf, capc, dz_t, dzd_t);
        // ---- A at (mc, jc, ka) into its slot (the old value where the cell
        // is not an interior cell of this footprint position)
        const int iA = 2 * mc + sA;
        const bool validA = col && jc == jl && ka == kl && ka <= ke + 1 && iA >= 1 && iA <= NX;
        T anew = ppA;
        if (validA) {
            const T *stab = tabs + (ka & 3) * TB;
            const Cond<T> o = rb_cond_dz(g, stab, q, iA, jc, ka, [&] { return zt[ka - (kb - 1)]; });
            const T dd = diag(o);
            if (g.flux) {
                anew = ppA + (rb_flux(o, n, ppA) - bcA) / dd;
            } else {
                const T dp = (((rb_offdiag(o, n) - bcA) / dd - ppA)) * T(1);
                anew = ppA + T(1) * dp;
            }
        }
        as[(ka & 3) * (S2X * S2Y) + slot0] = anew;
        acl[(ka & 3) * (S2X * S2Y) + slot0] = clA;
        lds_barrier();
        // ---- B at (mc, jc, kB) from the A planes kB-1..kB+1
        if (doB) {
            const int sB = (hB + jc + kB) & 1;
            const int iB = 2 * mc + sB;
            const T *a0 = as + (kB & 3) * (S2X * S2Y);
            const uint8_t *c0 = acl + (kB & 3) * (S2X * S2Y);
            const int zm = ((kB - 1) & 3) * (S2X * S2Y) + slot0, zp = ((kB + 1) & 3) * (S2X * S2Y) + slot0;
            Nb<T> nb2;
            nb2.xm = a0[slot0 - 1 + sB];
            nb2.xp = a0[slot0 + sB];
            nb2.ym = a0[slot0 - S2X];
            nb2.yp = a0[slot0 + S2X];
            nb2.zm = as[zm];
            nb2.zp = as[zp];
            Cls qb;
            qb.c = clB;
            qb.xm = c0[slot0 - 1 + sB];
            qb.xp = c0[slot0 + sB];
            qb.ym = c0[slot0 - S2X];
            qb.yp = c0[slot0 + S2X];
            qb.zm = acl[zm];
            qb.zp = acl[zp];
            if (jc == jl && iB >= 1 && iB <= NX) {
                const T *stab = tabs + (kB & 3) * TB;
                const Cond<T> o = rb_cond_dz(g, stab, qb, iB, jc, kB, [&] { return zt[kB - (kb - 1)]; });
                const T dd = diag(o);
                T xn;
                if (g.flux) {
                    xn = ppB + (rb_flux(o, nb2, ppB) - bcB) / dd;
                } else {
                    const T dp = (((rb_offdiag(o, nb2) - bcB) / dd - ppB)) * T(1);
                    xn = ppB + T(1) * dp;
                }
                yB[rb_] = xn;
            }
            // the owned A cell of plane kB (interior cells only: halos stay)
            const int sA2 = (hA + jc + kB) & 1;
            const int iA2 = 2 * mc + sA2;
            if (jc == jl && iA2 >= 1 && iA2 <= NX) yA[rb_] = a0[slot0];
        }
    };
    // steps kb-1 .. ke+1 in pairs; an odd count ends with a step that
    // updates and stores nothing
    for (int ka = kb - 1; ka <= ke + 1; ka += 2) {
        step(ka, v0, v1);
        step(ka + 1, v1, v0);
    }
}

template <typename T>
hipError_t launch_rb_sweep2(const T *xin, T *xout, const T *b, const RbGeo<T> &g, int hA, int C, const KState *st,
                            hipStream_t s)
{
    if (C < 1) return hipErrorInvalidValue;
    const unsigned nbx = (unsigned)((g.HX + S2OX - 1) / S2OX), nby = (unsigned)((g.NY + S2OY - 1) / S2OY);
    const unsigned nch = (unsigned)((g.NZ + C - 1) / C);
    const unsigned nb = (nbx * nby * nch + 7u) & ~7u;
    hipLaunchKernelGGL(k_rb_sweep2<T>, dim3(nb), dim3(S2X, S2Y), rb_sweep2_lds(g), s, xin, xout, b, g.cls, g, hA, C,
                       nbx, nby, nch, st);
    return hipGetLastError();
}

template <typename T>
bool rb_sweep2_fits(const RbGeo<T> &g)
{
    // the LDS budget, and one staged table entry per thread (rb_stage_regs)
    return g.K >= 1 && rb_sweep2_lds(g) <= kSweep2LdsMax && 2 * g.K * g.K <= S2X * S2Y;
}

// Measured on one MI355X (HIP events, one full sweep, 64 x 8 footprint):
// at 960x960x120 (transient) the two-colour launches with 60-plane chunks
// take 933 us in FP64 and 520 us in FP32 (40-plane chunks: 951 and 522 us;
// 120: 929 and 527 us), against 881-892 us and 467 us for the per-colour
// launches of the NC-cell form; at 480x480x60 and 240x240x30 every chunk
// length was slower than the per-colour launches of any form (best 132
// against 96 us and 69 against 58 us at 480x480x60).  So every operator
// takes one launch per colour unless the caller sets a chunk.
int sweep_chunk_auto(int HX, int NY, int NZ, int elem_bytes)
{
    (void)HX;
    (void)NY;
    (void)NZ;
    (void)elem_bytes;
    return 0;
}

// The Gauss-Seidel value of a cell whose old value and neighbours are all
// zero (the first colour of a preconditioner that starts from x = 0): the
// general update's operations with those zeros, so bitwise the same.
template <typename T>
__device__ __forceinline__ T rb_gs_from_zero(const RbGeo<T> &g, const Cond<T> &cd, T bv)
{
    const T pp = T(0);
    const T dd = diag(cd);
    if (g.flux) return pp + (T(0) - bv) / dd;
    const T ss = (cd.xp * T(0) + cd.xm * T(0) + cd.yp * T(0) + cd.ym * T(0) + cd.zp * T(0) + cd.zm * T(0));
    const T dp = (((ss - bv) / dd - pp)) * T(1);
    return pp + T(1) * dp;
}

// Colour 0 of the first sweep from x = 0, fused with forming its right-hand
// side bb (BiCG1! / Triad! / given): the colour-0 cell of the pair (half 1)
// gets the Gauss-Seidel value from all-zero neighbours.  Same arithmetic as
// k_rbsor_first.  The colour-1 cell of x is not written: the colour-1 pass
// that follows (launch_rb_sweep with from_zero) takes it as zero without
// reading it and writes every interior one.
template <typename T>
__global__ void __launch_bounds__(256) k_rb_first(int mode, T *__restrict__ x1,
                                                  T *__restrict__ b0, T *__restrict__ b1, const T *__restrict__ r0,
                                                  const T *__restrict__ r1, const T *__restrict__ q0,
                                                  const T *__restrict__ q1, const uint8_t *__restrict__ c0,
                                                  const uint8_t *__restrict__ c1, RbGeo<T> g, const KState *st)
{
    if (halted(st)) return;
    const RbThread t = rb_thread(g);
    const unsigned HX = (unsigned)g.HX, HXY = HX * (unsigned)g.MY;
    T beta = T(0), omg = T(0), r_alpha = T(0);
    bool first_itr = false;
    if (mode == FIRST_BICG_P) {
        first_itr = (st->itr == 1);
        beta = (T)st->beta;
        omg = (T)st->omega;
    } else if (mode == FIRST_BICG_S) {
        r_alpha = (T)(-st->alpha);
    }
    RbCell c[2];
    int h[2];
    rb_pair(g, t, c, h);
    // right-hand side of both cells: half 0 and half 1 entries at the same offset
    const unsigned o = c[0].co;
    const unsigned ob = o * (unsigned)sizeof(T);
    T bv[2];
    const T *rr[2] = {r0, r1}, *qq[2] = {q0, q1};
    T *bb[2] = {b0, b1};
#pragma unroll
    for (int hh = 0; hh < 2; ++hh) {
        if (mode == FIRST_BICG_P)
            bv[hh] = first_itr ? ldb(rr[hh], ob) : ldb(rr[hh], ob) + beta * (ldb(bb[hh], ob) - omg * ldb(qq[hh], ob));
        else if (mode == FIRST_BICG_S) bv[hh] = r_alpha * ldb(qq[hh], ob) + ldb(rr[hh], ob);
        else bv[hh] = ldb(bb[hh], ob);
    }
    // the half-1 cell (colour 0) of the pair
    const int a1 = h[0] == 1 ? 0 : 1;
    const RbCell &cr = c[a1];
    const Cls q = rb_cls(c1, c0, o, o, cr.s, HX, HXY);
    const T *stab = rb_stage(g, t.k);
    const Cond<T> cd = rb_cond(g, stab, q, cr.i, t.j, t.k);
    const T xn = rb_gs_from_zero(g, cd, bv[1]);
    const bool v1 = c[a1].valid, v0 = c[1 - a1].valid;
    if (mode != FIRST_COPY) {
        if (v0) stb(b0, ob, bv[0]);
        if (v1) stb(b1, ob, bv[1]);
    }
    if (v1) stb(x1, ob, xn);
}

template <typename T>
hipError_t launch_rb_first(int mode, T *x, T *bb, const T *r, const T *q, const RbGeo<T> &g, const KState *st,
                           hipStream_t s)
{
    const size_t HS = g.HS;
    // q is not read in FIRST_COPY mode (nullptr); r neither
    const T *r1 = r ? r + HS : nullptr, *q1 = q ? q + HS : nullptr;
    hipLaunchKernelGGL(k_rb_first<T>, dim3(g.nb1), dim3(BX, BY), rb_lds(g), s, mode, x + HS, bb, bb + HS, r, r1, q,
                       q1, g.cls, g.cls + HS, g, st);
    return hipGetLastError();
}

// The preconditioner's right-hand side alone (no smoother)
template <typename T>
__global__ void __launch_bounds__(256) k_rb_form_rhs(int mode, T *__restrict__ bb, const T *__restrict__ r,
                                                     const T *__restrict__ q, RbGeo<T> g, const KState *st)
{
    if (halted(st)) return;
    const RbThread t = rb_thread(g);
    const bool first_itr = (st->itr == 1);
    const T beta = (T)st->beta, omg = (T)st->omega, r_alpha = (T)(-st->alpha);
    RbCell c[2];
    int h[2];
    rb_pair(g, t, c, h);
#pragma unroll
    for (int a = 0; a < 2; ++a) {
        const size_t o = (h[a] ? g.HS : 0) + c[a].co;
        const T v = mode == FIRST_BICG_P ? (first_itr ? r[o] : r[o] + beta * (bb[o] - omg * q[o])) : r_alpha * q[o] + r[o];
        if (c[a].valid) bb[o] = v;
    }
}

template <typename T>
hipError_t launch_rb_form_rhs(int mode, T *bb, const T *r, const T *q, const RbGeo<T> &g, const KState *st,
                              hipStream_t s)
{
    hipLaunchKernelGGL(k_rb_form_rhs<T>, dim3(g.nb1), dim3(BX, BY), 0, s, mode, bb, r, q, g, st);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// operator, residual, BLAS-1 with fused sums
// ---------------------------------------------------------------------------
// CalcAX! (src/NonUniform.jl:454-511): ax = A x; with NS >= 1 also
// s0 = Σ ax·(dy ? dy : ax), s1 = Σ ax·dz.  Pair mapping (cells 2m and
// 2m+1 of a row) over R rows per thread (rb_rows).  Offsets are into the
// whole arrays (both halves).
template <typename T, int NS, int R>
__global__ void __launch_bounds__(256) k_rb_calc_ax(T *__restrict__ ax, const T *__restrict__ x,
                                                    const T *__restrict__ dy, const T *__restrict__ dz, RbGeo<T> g,
                                                    KState *st, RedBuf rb)
{
    // small grids (fewer than 2^21 cells per half): the loads go out
    // before the status test, whose latency they then hide
    const bool early = g.HS < ((size_t)1 << 21);
    if (!early && halted(st)) return;
    const RbRows t = rb_rows<R>(g);
    const unsigned HX = (unsigned)g.HX, HXY = HX * (unsigned)g.MY, HS = (unsigned)g.HS;
    const unsigned sz = (unsigned)sizeof(T);
    Nb<T> n[R][2];
    Cls q[R][2];
    T xv[R][2], av[R][2], dyv[R][2], dzv[R][2];
    unsigned co[R][2];
    bool valid[R][2];
    int jj[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
        jj[r] = min(t.j0 + r * BY, g.NY);
        const bool rowok = t.live && t.j0 + r * BY <= g.NY;
        const unsigned base = (unsigned)t.m + HX * (unsigned)jj[r] + HXY * (unsigned)t.k;
        const int h0 = (jj[r] + t.k) & 1;
#pragma unroll
        for (int a = 0; a < 2; ++a) {
            const int h = h0 ^ a, i = 2 * t.m + a;
            co[r][a] = (h ? HS : 0u) + base;
            const unsigned no = (h ? 0u : HS) + base;
            valid[r][a] = rowok && i >= 1 && i <= g.NX;
            n[r][a] = rb_nb<T>(x, no, a, HX, HXY);
            xv[r][a] = ldb(x, co[r][a] * sz);
            q[r][a] = rb_cls(g.cls, g.cls, co[r][a], no, a, HX, HXY);
            dyv[r][a] = (NS >= 1 && dy) ? ldb(dy, co[r][a] * sz) : T(0);
            dzv[r][a] = NS >= 2 ? ldb(dz, co[r][a] * sz) : T(0);
        }
        // the pair's cells are each other's x-neighbours (cell 2m+1 sits at
        // the other half's offset of cell 2m and the reverse): reuse their
        // values and classes instead of loading them again
        n[r][0].xp = xv[r][1];
        n[r][1].xm = xv[r][0];
        q[r][0].xp = q[r][1].c;
        q[r][1].xm = q[r][0].c;
    }
    if (early && halted(st)) return;
    const T *stab = rb_stage(g, t.k);
    dd v[3] = {{0.0, 0.0}, {0.0, 0.0}, {0.0, 0.0}};
#pragma unroll
    for (int r = 0; r < R; ++r)
#pragma unroll
        for (int a = 0; a < 2; ++a) {
            const Cond<T> cd = rb_cond(g, stab, q[r][a], 2 * t.m + a, jj[r], t.k);
            av[r][a] = rb_apply(g, cd, n[r][a], xv[r][a]);
            if (!(NS >= 1 && dy)) dyv[r][a] = av[r][a];
            if (R > 1 && valid[r][a]) {
                if (NS >= 1) v[0] = dd_add(v[0], (double)av[r][a] * (double)dyv[r][a]);
                if (NS >= 2) v[1] = dd_add(v[1], (double)av[r][a] * (double)dzv[r][a]);
            }
        }
    if (R == 1) {
        if (NS >= 1)
            v[0] = pair_sum(valid[0][0], (double)av[0][0] * (double)dyv[0][0], valid[0][1],
                            (double)av[0][1] * (double)dyv[0][1]);
        if (NS >= 2)
            v[1] = pair_sum(valid[0][0], (double)av[0][0] * (double)dzv[0][0], valid[0][1],
                            (double)av[0][1] * (double)dzv[0][1]);
    }
    // the workgroup sum first, the stores last: a store before the
    // reduction's barrier would hold the wave until it completes
    if (NS > 0) block_partials<NS>(v, rb);
#pragma unroll
    for (int r = 0; r < R; ++r)
#pragma unroll
        for (int a = 0; a < 2; ++a)
            if (valid[r][a]) stb(ax, co[r][a] * sz, av[r][a]);
}

// CalcAX! in the form of k_rb_sweep_w: a thread takes the 2·NC consecutive
// cells i = 2m0 .. 2m0 + 2NC - 1 of a row (m0 = mb + NC·l of an NC·64-wide
// tile, one wave per row), NC in each half, so that each load of x, dy,
// dz and the class bytes covers NC cells and the x neighbours at the ends
// come from the next threads by DPP lane shifts (or, at the wave's ends,
// from one more load each).  Every cell's value is k_rb_calc_ax's and the
// sums are exactly rounded, so the results are bitwise the same.
template <typename T, int NC, int NS>
__global__ void __launch_bounds__(256) k_rb_calc_ax_w(T *__restrict__ ax, const T *__restrict__ x,
                                                      const T *__restrict__ dy, const T *__restrict__ dz,
                                                      RbGeo<T> g, KState *st, RedBuf rb)
{
    if (halted(st)) return;
    constexpr int TW = NC * BX;
    const unsigned gx = ((unsigned)g.HX + TW - 1) / TW, gy = ((unsigned)g.NY + BY - 1) / BY;
    const unsigned gz = (unsigned)g.NZ, nreal = gx * gy * gz, nb = (nreal + 7u) & ~7u;
    const unsigned p = blockIdx.x;
    const unsigned L = (nb >> 3) * (p & 7u) + (p >> 3);
    const unsigned bx = L % gx, t1 = L / gx, bz = t1 % gz, by = t1 / gz;
    const int HXi = g.HX;
    const int mb = (int)bx * TW, m0 = mb + NC * (int)threadIdx.x;
    // one row per wave: its row index is wave-uniform (scalar registers)
    const int jr = (int)by * BY + __builtin_amdgcn_readfirstlane((int)threadIdx.y) + 1;
    const int k = (int)min(bz, gz - 1u) + 1;
    const bool rowok = L < nreal && jr <= g.NY;
    const int j = min(jr, g.NY);
    const int mc = min(m0, HXi - 1);  // as in k_rb_sweep_w
    const unsigned HX = (unsigned)HXi, HXY = HX * (unsigned)g.MY;
    const unsigned e = (unsigned)mc + HX * (unsigned)j + HXY * (unsigned)k;
    const int h0 = (j + k) & 1;
    // half A holds the cells i = 2(m0 + c), half B i = 2(m0 + c) + 1
    const size_t oA = h0 ? g.HS : 0, oB = h0 ? 0 : g.HS;
    const T *xA = x + oA, *xB = x + oB;
    const uint8_t *cA = g.cls + oA, *cB = g.cls + oB;
    // the x neighbours outside the wave's cells: B at mb - 1 (first thread),
    // A at mb + TW (last thread), clamped into the row (a clamped one
    // belongs to a cell outside the interior)
    const unsigned eB = (unsigned)max(mb - 1, 0) + HX * (unsigned)j + HXY * (unsigned)k;
    const unsigned eA = (unsigned)min(mb + TW, HXi - 1) + HX * (unsigned)j + HXY * (unsigned)k;
    // [0]: half A's cells (neighbours in B), [1]: half B's (neighbours in A)
    T own[2][NC], ym[2][NC], yp[2][NC], zm[2][NC], zp[2][NC];
    ld_vec<NC>(xA, e, own[0]);
    ld_vec<NC>(xB, e, own[1]);
    ld_vec<NC>(xB, e - HX, ym[0]);
    ld_vec<NC>(xB, e + HX, yp[0]);
    ld_vec<NC>(xB, e - HXY, zm[0]);
    ld_vec<NC>(xB, e + HXY, zp[0]);
    ld_vec<NC>(xA, e - HX, ym[1]);
    ld_vec<NC>(xA, e + HX, yp[1]);
    ld_vec<NC>(xA, e - HXY, zm[1]);
    ld_vec<NC>(xA, e + HXY, zp[1]);
    const T xeB = xB[eB], xeA = xA[eA];
    T dyv[2][NC] = {}, dzv[2][NC] = {};
    if (NS >= 1 && dy) {
        ld_vec<NC>(dy + oA, e, dyv[0]);
        ld_vec<NC>(dy + oB, e, dyv[1]);
    }
    if (NS >= 2) {
        ld_vec<NC>(dz + oA, e, dzv[0]);
        ld_vec<NC>(dz + oB, e, dzv[1]);
    }
    int cc[2][NC], cym[2][NC], cyp[2][NC], czm[2][NC], czp[2][NC];
    ld_cls_vec<NC>(cA, e, cc[0]);
    ld_cls_vec<NC>(cB, e, cc[1]);
    ld_cls_vec<NC>(cB, e - HX, cym[0]);
    ld_cls_vec<NC>(cB, e + HX, cyp[0]);
    ld_cls_vec<NC>(cB, e - HXY, czm[0]);
    ld_cls_vec<NC>(cB, e + HXY, czp[0]);
    ld_cls_vec<NC>(cA, e - HX, cym[1]);
    ld_cls_vec<NC>(cA, e + HX, cyp[1]);
    ld_cls_vec<NC>(cA, e - HXY, czm[1]);
    ld_cls_vec<NC>(cA, e + HXY, czp[1]);
    const int ceB = cB[eB], ceA = cA[eA];
    const T *stab = rb_stage(g, k);
    // x neighbours: A's cell c has B[m0 + c - 1] and B[m0 + c], B's cell c
    // has A[m0 + c] and A[m0 + c + 1]
    const T bsr = dpp_shift<kDppWaveShr1>(own[1][NC - 1], xeB);   // B[m0 - 1]
    const T asl = dpp_shift<kDppWaveShl1>(own[0][0], xeA);        // A[m0 + NC]
    const int cbsr = dpp_shift<kDppWaveShr1>(cc[1][NC - 1], ceB);
    const int casl = dpp_shift<kDppWaveShl1>(cc[0][0], ceA);
    T av[2][NC];
    bool valid[2][NC];
    dd v[3] = {{0.0, 0.0}, {0.0, 0.0}, {0.0, 0.0}};
#pragma unroll
    for (int c = 0; c < NC; ++c)
#pragma unroll
        for (int a = 0; a < 2; ++a) {
            Nb<T> n;
            Cls q;
            if (a == 0) {
                n.xm = c > 0 ? own[1][c - 1] : bsr;
                n.xp = own[1][c];
                q.xm = c > 0 ? cc[1][c - 1] : cbsr;
                q.xp = cc[1][c];
            } else {
                n.xm = own[0][c];
                n.xp = c < NC - 1 ? own[0][c + 1] : asl;
                q.xm = cc[0][c];
                q.xp = c < NC - 1 ? cc[0][c + 1] : casl;
            }
            n.ym = ym[a][c];
            n.yp = yp[a][c];
            n.zm = zm[a][c];
            n.zp = zp[a][c];
            q.c = cc[a][c];
            q.ym = cym[a][c];
            q.yp = cyp[a][c];
            q.zm = czm[a][c];
            q.zp = czp[a][c];
            const int i = 2 * (m0 + c) + a;
            valid[a][c] = rowok && m0 + c < HXi && i >= 1 && i <= g.NX;
            const Cond<T> cd = rb_cond(g, stab, q, i, j, k);
            av[a][c] = rb_apply(g, cd, n, own[a][c]);
            const T yv = (NS >= 1 && dy) ? dyv[a][c] : av[a][c];
            if (valid[a][c]) {
                if (NS >= 1) v[0] = dd_add(v[0], (double)av[a][c] * (double)yv);
                if (NS >= 2) v[1] = dd_add(v[1], (double)av[a][c] * (double)dzv[a][c]);
            }
        }
    if (NS > 0) block_partials<NS>(v, rb);
#pragma unroll
    for (int a = 0; a < 2; ++a) {
        T *ah = ax + (a ? oB : oA);
        bool all = true;
#pragma unroll
        for (int c = 0; c < NC; ++c) all = all && valid[a][c];
        if (all) {
            st_vec<NC>(ah, e, av[a]);
        } else {
#pragma unroll
            for (int c = 0; c < NC; ++c)
                if (valid[a][c]) ah[e + (unsigned)c] = av[a][c];
        }
    }
}

// One row per thread: two rows measured slower at 960x960x120 (888 against
// 838 µs for CalcAX!, 1013 against 975 µs with its dot) and equal at
// 240x240x30.
constexpr int kAxRows = 1;

template <typename T>
hipError_t launch_rb_calc_ax(T *ax, const T *x, const RbGeo<T> &g, const T *dot_y, const T *dot_z, int fin,
                             KState *st, double *hist, const RedBuf &rb, hipStream_t s)
{
    constexpr int R = kAxRows;
    const int ns = fin < 0 ? 0 : (dot_z ? 2 : 1);
    // large grids: the NC-cell form (measured, HIP events: FP64 53.5
    // against 66.7 us at 480x480x60 and 677 against 822 us at 960x960x120,
    // FP32 32.1 against 40.5 us and 334 against 432 us)
    const bool wide = g.HS >= ((size_t)1 << 21);
    constexpr int NC = kSweepCells<T>;
    const unsigned nb = wide ? rb_wide_nb<T, NC>(g) : rb_rows_nb<R>(g);
    const dim3 grid(nb), block(BX, BY);
    if (wide && ns == 0)
        hipLaunchKernelGGL((k_rb_calc_ax_w<T, NC, 0>), grid, block, rb_lds(g), s, ax, x, dot_y, dot_z, g, st, rb);
    else if (wide && ns == 1)
        hipLaunchKernelGGL((k_rb_calc_ax_w<T, NC, 1>), grid, block, rb_lds(g), s, ax, x, dot_y, dot_z, g, st, rb);
    else if (wide)
        hipLaunchKernelGGL((k_rb_calc_ax_w<T, NC, 2>), grid, block, rb_lds(g), s, ax, x, dot_y, dot_z, g, st, rb);
    else if (ns == 0)
        hipLaunchKernelGGL((k_rb_calc_ax<T, 0, R>), grid, block, rb_lds(g), s, ax, x, dot_y, dot_z, g, st, rb);
    else if (ns == 1)
        hipLaunchKernelGGL((k_rb_calc_ax<T, 1, R>), grid, block, rb_lds(g), s, ax, x, dot_y, dot_z, g, st, rb);
    else hipLaunchKernelGGL((k_rb_calc_ax<T, 2, R>), grid, block, rb_lds(g), s, ax, x, dot_y, dot_z, g, st, rb);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess || ns == 0) return e;
    return launch_finalize_sums<T>(ns, nb, fin, st, hist, rb, s, seq_terms<T>(g, ax, dot_y, ax, dot_z));
}

// CalcRK! (src/NonUniform.jl:372-434) with θ, b in standard layout; r (and
// r0copy) in RB layout; s0 = Σ r².
template <typename T>
__global__ void __launch_bounds__(256) k_rb_calc_rk(T *__restrict__ r, T *__restrict__ r0copy,
                                                    const T *__restrict__ th, const T *__restrict__ b, RbGeo<T> g,
                                                    KState *st, RedBuf rb)
{
    if (halted(st)) return;
    const RbThread t = rb_thread(g);
    const unsigned HX = (unsigned)g.HX, HXY = HX * (unsigned)g.MY;
    const size_t sy = (size_t)g.MX, sz = (size_t)g.MX * (size_t)g.MY;
    RbCell c[2];
    int h[2];
    rb_pair(g, t, c, h);
    const unsigned o = c[0].co;
    const uint8_t *cl = g.cls;
    const size_t HS = g.HS;
    Nb<T> n[2];
    Cls q[2];
    T x0[2], bv[2], rs[2];
#pragma unroll
    for (int a = 0; a < 2; ++a) {
        // standard-layout offset of the cell (i is at most MX-1 for a loadable m)
        const size_t cs = (size_t)min(c[a].i, g.MX - 2 + a) + sy * (size_t)t.j + sz * (size_t)t.k;
        n[a].xm = th[cs - 1];
        n[a].xp = th[cs + 1];
        n[a].ym = th[cs - sy];
        n[a].yp = th[cs + sy];
        n[a].zm = th[cs - sz];
        n[a].zp = th[cs + sz];
        x0[a] = th[cs];
        bv[a] = b[cs];
        const int hh = h[a];
        q[a] = rb_cls(cl + (hh ? HS : 0), cl + (hh ? 0 : HS), o, o, c[a].s, HX, HXY);
    }
    const T *stab = rb_stage(g, t.k);
#pragma unroll
    for (int a = 0; a < 2; ++a) {
        const Cond<T> cd = rb_cond(g, stab, q[a], c[a].i, t.j, t.k);
        if (g.flux) rs[a] = bv[a] - rb_flux(cd, n[a], x0[a]);
        else rs[a] = (bv[a] - (rb_offdiag(cd, n[a]) - diag(cd) * x0[a])) * T(1);
    }
    dd v[3] = {{0.0, 0.0}, {0.0, 0.0}, {0.0, 0.0}};
#pragma unroll
    for (int a = 0; a < 2; ++a)
        if (c[a].valid) v[0] = dd_add(v[0], (double)rs[a] * (double)rs[a]);
    block_partials<1>(v, rb);
#pragma unroll
    for (int a = 0; a < 2; ++a) {
        if (c[a].valid) {
            const size_t ro = (h[a] ? HS : 0) + o;
            r[ro] = rs[a];
            if (r0copy) r0copy[ro] = rs[a];
        }
    }
}

template <typename T>
hipError_t launch_rb_calc_rk(T *r, T *r0copy, const T *th, const T *b, const RbGeo<T> &g, int fin, KState *st,
                             double *hist, const RedBuf &rb, hipStream_t s)
{
    hipLaunchKernelGGL(k_rb_calc_rk<T>, dim3(g.nb1), dim3(BX, BY), rb_lds(g), s, r, r0copy, th, b, g, st, rb);
    hipError_t e = hipGetLastError();
    return e != hipSuccess ? e : launch_finalize_sums<T>(1, g.nb1, fin, st, hist, rb, s, seq_terms<T>(g, r, nullptr));
}

// Fdot1/Fdot2 (src/NonUniform.jl:567-592): s0 = Σ x·(y ? y : x), s1 = Σ x·z
template <typename T, int NS>
__global__ void __launch_bounds__(256) k_rb_dot(const T *__restrict__ x, const T *__restrict__ y,
                                                const T *__restrict__ z, RbGeo<T> g, KState *st, RedBuf rb)
{
    if (halted(st)) return;
    const RbThread t = rb_thread(g);
    RbCell c[2];
    int h[2];
    rb_pair(g, t, c, h);
    double xv[2], yv[2], zv[2];
#pragma unroll
    for (int a = 0; a < 2; ++a) {
        const size_t o = (h[a] ? g.HS : 0) + c[a].co;
        xv[a] = (double)x[o];
        yv[a] = (double)(y ? y[o] : x[o]);
        zv[a] = NS >= 2 ? (double)z[o] : 0.0;
    }
    dd v[3] = {{0.0, 0.0}, {0.0, 0.0}, {0.0, 0.0}};
#pragma unroll
    for (int a = 0; a < 2; ++a) {
        if (c[a].valid) {
            v[0] = dd_add(v[0], xv[a] * yv[a]);
            if (NS >= 2) v[1] = dd_add(v[1], xv[a] * zv[a]);
        }
    }
    block_partials<NS>(v, rb);
}

template <typename T>
hipError_t launch_rb_dot(const T *x, const T *y, const T *z, const RbGeo<T> &g, int fin, KState *st, double *hist,
                         const RedBuf &rb, hipStream_t s)
{
    const dim3 grid(g.nb1), block(BX, BY);
    if (z) hipLaunchKernelGGL((k_rb_dot<T, 2>), grid, block, 0, s, x, y, z, g, st, rb);
    else hipLaunchKernelGGL((k_rb_dot<T, 1>), grid, block, 0, s, x, y, z, g, st, rb);
    hipError_t e = hipGetLastError();
    return e != hipSuccess ? e
                           : launch_finalize_sums<T>(z ? 2 : 1, g.nb1, fin, st, hist, rb, s, seq_terms<T>(g, x, y, x, z));
}

// BICG2!(θ, p_, s_, alpha, omega); Triad!(r, t_, s, -omega); Σ r², and with
// r0 also Σ r·r0 (NonUniform.jl:328-332, 289).  θ (and xbest) in standard
// layout, FP64.
template <typename T, int NS>
__global__ void __launch_bounds__(256) k_rb_bicg_update(double *__restrict__ th, double *__restrict__ xbest,
                                                        T *__restrict__ r, const T *__restrict__ p_,
                                                        const T *__restrict__ s_, const T *__restrict__ t,
                                                        const T *__restrict__ sv, const T *__restrict__ r0,
                                                        RbGeo<T> g, KState *st, RedBuf rb)
{
    if (halted(st)) return;
    const RbThread tt = rb_thread(g);
    const double alpha = st->alpha, omega = st->omega;
    const T r_omega = (T)(-st->omega);
    const bool snap = xbest != nullptr && st->improved != 0;
    const size_t sy = (size_t)g.MX, sz = (size_t)g.MX * (size_t)g.MY;
    RbCell c[2];
    int h[2];
    rb_pair(g, tt, c, h);
    // the pair's two standard-layout cells are adjacent: i = 2m, 2m+1
    const size_t cs0 = (size_t)(2 * tt.m) + sy * (size_t)tt.j + sz * (size_t)tt.k;
    double xo[2], xn[2];
    T rv[2], r0v[2];
    const unsigned HS = (unsigned)g.HS, esz = (unsigned)sizeof(T);
#pragma unroll
    for (int a = 0; a < 2; ++a) {
        const unsigned ob = ((h[a] ? HS : 0u) + c[a].co) * esz;
        const size_t cs = min(cs0 + a, cs0 + (size_t)(g.MX - 1 - 2 * tt.m));
        xo[a] = th[cs];
        xn[a] = xo[a] + (alpha * (double)ldb(p_, ob) + omega * (double)ldb(s_, ob));
        rv[a] = r_omega * ldb(t, ob) + ldb(sv, ob);
        r0v[a] = NS >= 2 ? ldb(r0, ob) : T(0);
    }
    dd v[3] = {{0.0, 0.0}, {0.0, 0.0}, {0.0, 0.0}};
    v[0] = pair_sum(c[0].valid, (double)rv[0] * (double)rv[0], c[1].valid, (double)rv[1] * (double)rv[1]);
    if (NS >= 2)
        v[1] = pair_sum(c[0].valid, (double)rv[0] * (double)r0v[0], c[1].valid, (double)rv[1] * (double)r0v[1]);
    block_partials<NS>(v, rb);
#pragma unroll
    for (int a = 0; a < 2; ++a) {
        if (c[a].valid) {
            const size_t cs = cs0 + a;
            if (snap) xbest[cs] = xo[a];
            th[cs] = xn[a];
            stb(r, ((h[a] ? HS : 0u) + c[a].co) * esz, rv[a]);
        }
    }
}

template <typename T>
hipError_t launch_rb_bicg_update(double *th, double *xbest, T *r, const T *p_, const T *s_, const T *t,
                                 const T *sv, const T *r0, const RbGeo<T> &g, KState *st, double *hist,
                                 const RedBuf &rb, hipStream_t s)
{
    const dim3 grid(g.nb1), block(BX, BY);
    if (r0 == nullptr) {
        hipLaunchKernelGGL((k_rb_bicg_update<T, 1>), grid, block, 0, s, th, xbest, r, p_, s_, t, sv, r0, g, st, rb);
        hipError_t e = hipGetLastError();
        return e != hipSuccess ? e
                               : launch_finalize_sums<T>(1, g.nb1, FIN_BICG_RES, st, hist, rb, s,
                                                         seq_terms<T>(g, r, nullptr));
    }
    hipLaunchKernelGGL((k_rb_bicg_update<T, 2>), grid, block, 0, s, th, xbest, r, p_, s_, t, sv, r0, g, st, rb);
    hipError_t e = hipGetLastError();
    return e != hipSuccess ? e
                           : launch_finalize_sums<T>(2, g.nb1, FIN_BICG_RES_RHO, st, hist, rb, s,
                                                     seq_terms<T>(g, r, nullptr, r, r0));
}

// CG!: θ += α p, r -= α q, Σ r² (NonUniform.jl:179-186)
template <typename T>
__global__ void __launch_bounds__(256) k_rb_cg_update(double *__restrict__ th, double *__restrict__ xbest,
                                                      T *__restrict__ r, const T *__restrict__ p,
                                                      const T *__restrict__ q, RbGeo<T> g, KState *st, RedBuf rb)
{
    if (halted(st)) return;
    const RbThread tt = rb_thread(g);
    const double alpha_d = st->alpha;
    const T alpha = (T)st->alpha;
    const bool snap = xbest != nullptr && st->improved != 0;
    const size_t sy = (size_t)g.MX, sz = (size_t)g.MX * (size_t)g.MY;
    RbCell c[2];
    int h[2];
    rb_pair(g, tt, c, h);
    const size_t cs0 = (size_t)(2 * tt.m) + sy * (size_t)tt.j + sz * (size_t)tt.k;
    double xo[2], xn[2];
    T rv[2];
#pragma unroll
    for (int a = 0; a < 2; ++a) {
        const size_t o = (h[a] ? g.HS : 0) + c[a].co;
        const size_t cs = min(cs0 + a, cs0 + (size_t)(g.MX - 1 - 2 * tt.m));
        xo[a] = th[cs];
        xn[a] = xo[a] + alpha_d * (double)p[o];
        rv[a] = r[o] - alpha * q[o];
    }
    dd v[3] = {{0.0, 0.0}, {0.0, 0.0}, {0.0, 0.0}};
    v[0] = pair_sum(c[0].valid, (double)rv[0] * (double)rv[0], c[1].valid, (double)rv[1] * (double)rv[1]);
    block_partials<1>(v, rb);
#pragma unroll
    for (int a = 0; a < 2; ++a) {
        if (c[a].valid) {
            const size_t o = (h[a] ? g.HS : 0) + c[a].co;
            const size_t cs = cs0 + a;
            if (snap) xbest[cs] = xo[a];
            th[cs] = xn[a];
            r[o] = rv[a];
        }
    }
}

template <typename T>
hipError_t launch_rb_cg_update(double *th, double *xbest, T *r, const T *p, const T *q, const RbGeo<T> &g,
                               KState *st, double *hist, const RedBuf &rb, hipStream_t s)
{
    hipLaunchKernelGGL(k_rb_cg_update<T>, dim3(g.nb1), dim3(BX, BY), 0, s, th, xbest, r, p, q, g, st, rb);
    hipError_t e = hipGetLastError();
    return e != hipSuccess ? e
                           : launch_finalize_sums<T>(1, g.nb1, FIN_CG_RES, st, hist, rb, s,
                                                     seq_terms<T>(g, r, nullptr));
}

// CG!'s update followed by the first colour of its preconditioner, in one
// pass: θ += α p, r -= α q and Σ r² (NonUniform.jl:179-186, as
// k_rb_cg_update), then colour 0 of z = M⁻¹ r from z = 0 with the new r as
// its right-hand side (NonUniform.jl:201-203, as k_rb_first in FIRST_COPY
// mode).  The residual's finalize follows this kernel, so a converged
// iteration has also written colour 0 of z, which nothing reads; the
// sweeps after the finalize see the halted state and return.
template <typename T>
__global__ void __launch_bounds__(256) k_rb_cg_first(double *__restrict__ th, double *__restrict__ xbest,
                                                     T *__restrict__ r, const T *__restrict__ p,
                                                     const T *__restrict__ q, T *__restrict__ z1, RbGeo<T> g,
                                                     KState *st, RedBuf rb)
{
    if (halted(st)) return;
    const RbThread tt = rb_thread(g);
    const unsigned HX = (unsigned)g.HX, HXY = HX * (unsigned)g.MY;
    const double alpha_d = st->alpha;
    const T alpha = (T)st->alpha;
    const bool snap = xbest != nullptr && st->improved != 0;
    const size_t sy = (size_t)g.MX, sz = (size_t)g.MX * (size_t)g.MY;
    RbCell c[2];
    int h[2];
    rb_pair(g, tt, c, h);
    const size_t cs0 = (size_t)(2 * tt.m) + sy * (size_t)tt.j + sz * (size_t)tt.k;
    double xo[2], xn[2];
    T rv[2];
#pragma unroll
    for (int a = 0; a < 2; ++a) {
        const size_t o = (h[a] ? g.HS : 0) + c[a].co;
        const size_t cs = min(cs0 + a, cs0 + (size_t)(g.MX - 1 - 2 * tt.m));
        xo[a] = th[cs];
        xn[a] = xo[a] + alpha_d * (double)p[o];
        rv[a] = r[o] - alpha * q[o];
    }
    // the half-1 cell (colour 0) of the pair
    const int a1 = h[0] == 1 ? 0 : 1;
    const RbCell &cr = c[a1];
    const unsigned o = c[0].co;
    const Cls cq = rb_cls(g.cls + g.HS, g.cls, o, o, cr.s, HX, HXY);
    const T *stab = rb_stage(g, tt.k);
    const Cond<T> cd = rb_cond(g, stab, cq, cr.i, tt.j, tt.k);
    const T zn = rb_gs_from_zero(g, cd, rv[a1]);
    dd v[3] = {{0.0, 0.0}, {0.0, 0.0}, {0.0, 0.0}};
    v[0] = pair_sum(c[0].valid, (double)rv[0] * (double)rv[0], c[1].valid, (double)rv[1] * (double)rv[1]);
    block_partials<1>(v, rb);
#pragma unroll
    for (int a = 0; a < 2; ++a) {
        if (c[a].valid) {
            const size_t oa = (h[a] ? g.HS : 0) + c[a].co;
            const size_t cs = cs0 + a;
            if (snap) xbest[cs] = xo[a];
            th[cs] = xn[a];
            r[oa] = rv[a];
        }
    }
    if (cr.valid) stb(z1, o * (unsigned)sizeof(T), zn);
}

template <typename T>
hipError_t launch_rb_cg_first(double *th, double *xbest, T *r, const T *p, const T *q, T *z, const RbGeo<T> &g,
                              KState *st, double *hist, const RedBuf &rb, hipStream_t s)
{
    hipLaunchKernelGGL(k_rb_cg_first<T>, dim3(g.nb1), dim3(BX, BY), rb_lds(g), s, th, xbest, r, p, q, z + g.HS, g, st,
                       rb);
    hipError_t e = hipGetLastError();
    return e != hipSuccess ? e
                           : launch_finalize_sums<T>(1, g.nb1, FIN_CG_RES, st, hist, rb, s,
                                                     seq_terms<T>(g, r, nullptr));
}

// CG!: p = z + β p (NonUniform.jl:218-220)
template <typename T>
__global__ void __launch_bounds__(256) k_rb_cg_p(T *__restrict__ p, const T *__restrict__ z, RbGeo<T> g,
                                                 const KState *st)
{
    if (halted(st)) return;
    const RbThread t = rb_thread(g);
    const T beta = (T)st->beta;
    RbCell c[2];
    int h[2];
    rb_pair(g, t, c, h);
#pragma unroll
    for (int a = 0; a < 2; ++a) {
        const size_t o = (h[a] ? g.HS : 0) + c[a].co;
        const T v = z[o] + beta * p[o];
        if (c[a].valid) p[o] = v;
    }
}

template <typename T>
hipError_t launch_rb_cg_p(T *p, const T *z, const RbGeo<T> &g, const KState *st, hipStream_t s)
{
    hipLaunchKernelGGL(k_rb_cg_p<T>, dim3(g.nb1), dim3(BX, BY), 0, s, p, z, g, st);
    return hipGetLastError();
}

template <typename T>
__global__ void __launch_bounds__(256) k_rb_copy_interior(T *__restrict__ dst, const T *__restrict__ src, RbGeo<T> g,
                                                          const KState *st)
{
    if (halted(st)) return;
    const RbThread t = rb_thread(g);
    RbCell c[2];
    int h[2];
    rb_pair(g, t, c, h);
#pragma unroll
    for (int a = 0; a < 2; ++a) {
        const size_t o = (h[a] ? g.HS : 0) + c[a].co;
        const T v = src[o];
        if (c[a].valid) dst[o] = v;
    }
}

template <typename T>
hipError_t launch_rb_copy_interior(T *dst, const T *src, const RbGeo<T> &g, const KState *st, hipStream_t s)
{
    hipLaunchKernelGGL(k_rb_copy_interior<T>, dim3(g.nb1), dim3(BX, BY), 0, s, dst, src, g, st);
    return hipGetLastError();
}

#define RB_INST(T)                                                                                              \
    template hipError_t launch_rb_copy_interior<T>(T *, const T *, const RbGeo<T> &, const KState *, hipStream_t); \
    template hipError_t launch_rb_calc_ax<T>(T *, const T *, const RbGeo<T> &, const T *, const T *, int, KState *, \
                                             double *, const RedBuf &, hipStream_t);                            \
    template hipError_t launch_rb_calc_rk<T>(T *, T *, const T *, const T *, const RbGeo<T> &, int, KState *,    \
                                             double *, const RedBuf &, hipStream_t);                            \
    template hipError_t launch_rb_sweep<T>(T *, const T *, const RbGeo<T> &, int, const KState *, hipStream_t,    \
                                           bool);                                                                  \
    template hipError_t launch_rb_cg_first<T>(double *, double *, T *, const T *, const T *, T *,                \
                                              const RbGeo<T> &, KState *, double *, const RedBuf &, hipStream_t); \
    template hipError_t launch_rb_sweep_dot<T>(T *, const T *, const RbGeo<T> &, int, KState *, double *,       \
                                               const RedBuf &, hipStream_t);                                    \
    template hipError_t launch_rb_sweep2<T>(const T *, T *, const T *, const RbGeo<T> &, int, int, const KState *, \
                                            hipStream_t);                                                       \
    template bool rb_sweep2_fits<T>(const RbGeo<T> &);                                                         \
    template hipError_t launch_rb_first<T>(int, T *, T *, const T *, const T *, const RbGeo<T> &, const KState *, \
                                           hipStream_t);                                                        \
    template hipError_t launch_rb_form_rhs<T>(int, T *, const T *, const T *, const RbGeo<T> &, const KState *,  \
                                              hipStream_t);                                                     \
    template hipError_t launch_rb_dot<T>(const T *, const T *, const T *, const RbGeo<T> &, int, KState *,      \
                                         double *, const RedBuf &, hipStream_t);                                \
    template hipError_t launch_rb_bicg_update<T>(double *, double *, T *, const T *, const T *, const T *,      \
                                                 const T *, const T *, const RbGeo<T> &, KState *, double *,    \
                                                 const RedBuf &, hipStream_t);                                  \
    template hipError_t launch_rb_cg_update<T>(double *, double *, T *, const T *, const T *, const RbGeo<T> &, \
                                               KState *, double *, const RedBuf &, hipStream_t);                \
    template hipError_t launch_rb_cg_p<T>(T *, const T *, const RbGeo<T> &, const KState *, hipStream_t);

RB_INST(double)
RB_INST(float)
#undef RB_INST

template hipError_t launch_to_rb<uint8_t>(uint8_t *, const uint8_t *, int, int, int, hipStream_t);
template hipError_t launch_to_rb<double>(double *, const double *, int, int, int, hipStream_t);
template hipError_t launch_to_rb<float>(float *, const float *, int, int, int, hipStream_t);
template hipError_t launch_from_rb<double>(double *, const double *, int, int, int, hipStream_t);
template hipError_t launch_from_rb<float>(float *, const float *, int, int, int, hipStream_t);

}  // namespace h3d
