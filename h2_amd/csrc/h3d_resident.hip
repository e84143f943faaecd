// Preconditioner! (src/NonUniform.jl:531-559) with the :gs smoother on
// small grids as ONE launch: the five red-black sweeps (ten colour passes,
// rbsor_core! at :848-915 with ω = 1) run inside a grid of co-resident
// workgroups that keep the iterate in LDS from the first pass to the last.
// The same machinery runs larger units as one launch: the preconditioner
// with CalcAX! and its sums fused (k_rb_resident), and the whole iteration
// loops of PBiCGSTAB! and CG! (k_rb_bicg_solve, k_rb_cg_solve).
//
// Why: on a grid whose vectors sit in the Infinity Cache (240x240x30 and
// below) a colour pass launched on its own is latency bound.  Every launch
// pays the dispatch of two rounds of workgroups, one dependent round trip
// for x, b and the class bytes, the table staging, and the drain; ten of
// them per preconditioner are most of a BiCGSTAB iteration.  Here each
// cell's diagonal and conductance table indices (formed once per solve by
// k_rb_res_coef) and its right-hand side are held in registers, x lives
// in LDS, and between passes only the rows a neighbouring tile needs cross
// HBM.
//
// Decomposition: the interior rows (j, k) are cut into tiles of TJ x TK
// rows (RbRes, planned on the host), one 1024-thread workgroup per tile,
// all resident at once (at most one tile per CU).  LDS holds both colour
// halves of the tile's rows plus a one-row ring in j and k (the rows the
// stencil reads from the neighbouring tiles), the staged conductance
// tables of the tile's planes and a table of the ring cells' addresses.
// Thread t owns the slots t, t+1024, ... of the tile's (row, m) index
// space in both halves.
//
// A pass updates the cells of one half from LDS into LDS, and stores the
// rows on the tile's edge into x in HBM (and, after the last sweep, the
// other rows when a caller reads x back from HBM); the workgroup then
// publishes its pass counter.  A pass first updates the rows inside the
// tile, whose neighbours are all in LDS, while the neighbouring tiles
// finish the previous pass; then it waits for the counters of the (at
// most four) tiles whose edge rows it reads, loads those ring rows of the
// half the previous pass updated, and updates its own edge rows.  The
// passes need no grid-wide barrier: each tile waits only for its
// neighbours (the sums of the fused launches and of the loops do, through
// res_grid_sync).  Cross-workgroup data moves with device-scope (sc1) loads
// and stores, which the L2s of the eight XCDs keep coherent without the
// write-back and invalidate a device-scope release/acquire fence costs;
// each wave waits for its stores before the counter is published.  Every
// wait is bounded: a timed-out wait sets an error word the host checks
// after the solve, and later waits then return at once.  The counters grow
// by one per pass across launches and start equal (the host zeroes them
// whenever a solve selects a plan).
//
// Every cell performs the per-colour kernels' operations in the same
// order (rb_offdiag / rb_flux / diag, the from-zero start as explicit
// zeros), so x is bitwise the result of the ten separate launches.

#include "h3d_rb.hpp"

namespace h3d {

constexpr int kResThreads = 1024;
constexpr int kResSlots = 4;                    // slots per thread and half
constexpr unsigned kResSpin = 1u << 16;         // polls before a wait gives up

// Words of the synchronisation buffer (RbRes::bar, kResBarWords, zeroed
// by the host whenever a solve selects a tile plan): the error word, then
// one pass counter per workgroup, 64 bytes apart.
constexpr int kBarErr = kResBarErrWord, kBarFlags = 64;
constexpr int kBarFlagStride = 16;
// the grid barrier's words after the pass counters: one arrival counter
// per XCD group, the counter of the groups, the generation
constexpr int kBarGrid = kBarFlags + kBarFlagStride * kResMaxWG;
constexpr int kBarTop = kBarGrid + 16 * 8, kBarGen = kBarGrid + 16 * 9;

__device__ __forceinline__ unsigned bar_load(const unsigned *p)
{
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void bar_store(unsigned *p, unsigned v)
{
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <typename T>
__device__ __forceinline__ void sc_store(T *p, T v)
{
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <typename T>
__device__ __forceinline__ T sc_load(const T *p)
{
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// x, opaque to the compiler: an address formed from it is formed where it
// is used, not hoisted out of the loops around the use into registers.
__device__ __forceinline__ unsigned opq(unsigned x)
{
    asm volatile("" : "+v"(x));
    return x;
}

// The tile of workgroup bid = a + ntj*b: first row (j0, k0), extents
// tj x tk, and its up to four neighbours in (j, k) (-1: none).
struct ResTile {
    int j0, k0, tj, tk;
    int nb[4];
};

__device__ __forceinline__ ResTile res_tile(const RbRes &pl, int NY, int NZ)
{
    ResTile t;
    const int a = (int)blockIdx.x % pl.ntj, b = (int)blockIdx.x / pl.ntj;
    t.j0 = 1 + a * pl.TJ;
    t.k0 = 1 + b * pl.TK;
    t.tj = min(pl.TJ, NY - t.j0 + 1);
    t.tk = min(pl.TK, NZ - t.k0 + 1);
    t.nb[0] = a > 0 ? (int)blockIdx.x - 1 : -1;
    t.nb[1] = a + 1 < pl.ntj ? (int)blockIdx.x + 1 : -1;
    t.nb[2] = b > 0 ? (int)blockIdx.x - pl.ntj : -1;
    t.nb[3] = b + 1 < pl.ntk ? (int)blockIdx.x + pl.ntj : -1;
    return t;
}

// Grid-wide barrier (every workgroup of the launch): two levels, one
// counter per XCD group (workgroups are dealt to the eight XCDs round-
// robin) and one for the groups; the last arriver of each level resets its
// counter, and the last of all bumps the generation the others poll.
// Device-scope stores issued before it are visible to every workgroup's
// device-scope loads after it.  Bounded like every wait here.  (A flat
// variant, one arrival word per workgroup polled by every workgroup,
// measured slower: 132.8 against 130.7 µs per PBiCGSTAB iteration at
// 240x240x30.)
__device__ __forceinline__ void res_grid_sync(unsigned *bar)
{
    asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
    __syncthreads();
    if (threadIdx.x == 0) {
        const unsigned G = gridDim.x, grp = blockIdx.x & 7u;
        const unsigned ng = G / 8u + (grp < (G & 7u) ? 1u : 0u);
        const unsigned ngroups = G < 8u ? G : 8u;
        // the generation cannot move before this workgroup arrives
        const unsigned gen0 = bar_load(bar + kBarGen);
        if (__hip_atomic_fetch_add(bar + kBarGrid + 16 * grp, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) ==
            ng - 1u) {
            bar_store(bar + kBarGrid + 16 * grp, 0u);
            asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
            if (__hip_atomic_fetch_add(bar + kBarTop, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) ==
                ngroups - 1u) {
                bar_store(bar + kBarTop, 0u);
                asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
                __hip_atomic_fetch_add(bar + kBarGen, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            }
        }
        unsigned n = 0;
        while (bar_load(bar + kBarGen) == gen0) {
            __builtin_amdgcn_s_sleep(1);
            if ((++n & 255u) == 0u && (n >= kResSpin || bar_load(bar + kBarErr) != 0u)) {
                bar_store(bar + kBarErr, 1u);
                break;
            }
        }
    }
    __syncthreads();
}

// A timed-out wait leaves the iterates wrong: the one-launch loops stop
// at the end of the iteration (the host reads the error word and reports
// the timeout, whatever the state says).  Called by thread 0 before the
// barrier after which every thread tests the state, so that every wave of
// the workgroup leaves the loop together.
__device__ __forceinline__ void res_stop_on_error(unsigned *bar, KState &ks)
{
    if (ks.status == ST_RUNNING && bar_load(bar + kBarErr) != 0u) ks.status = ST_BREAKDOWN;
}

// Publish that this workgroup has finished pass count `cnt`: its stores
// of the rows its neighbours read are complete (visible to their
// device-scope loads) before the counter moves.  Every thread must call it.
__device__ __forceinline__ void res_publish(unsigned *bar, unsigned cnt)
{
    asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
    __syncthreads();
    if (threadIdx.x == 0) bar_store(bar + kBarFlags + kBarFlagStride * blockIdx.x, cnt);
}

// Wait until the neighbouring tiles have published pass count `want`.
// Every thread must call it; ends with a barrier.
__device__ __forceinline__ void res_wait(unsigned *bar, const ResTile &tl, unsigned want)
{
    if (threadIdx.x < 64) {
        const int lane = threadIdx.x;
        // selects instead of an indexed read, which would put the tile in
        // scratch memory
        const int nb = lane == 0 ? tl.nb[0] : lane == 1 ? tl.nb[1] : lane == 2 ? tl.nb[2] : lane == 3 ? tl.nb[3] : -1;
        const unsigned *f = bar + kBarFlags + kBarFlagStride * (nb < 0 ? 0 : nb);
        unsigned n = 0;
        while (true) {
            const bool ok = nb < 0 || (int)(bar_load(f) - want) >= 0;
            if (__all(ok)) break;
            __builtin_amdgcn_s_sleep(1);
            if ((++n & 255u) == 0u && (n >= kResSpin || bar_load(bar + kBarErr) != 0u)) {
                bar_store(bar + kBarErr, 1u);
                break;
            }
        }
    }
    __syncthreads();
}

// The tile's ring (ring row hr = 0..2tj+2tk-1: planes k0-1 and k0+tk, then
// rows j0-1 and j0+tj of each plane, HX cells each) as a table in LDS, one
// entry per ring cell: the cell's offset in an HBM half, and its offset in
// an LDS half with bit 16 + h set when the cell is an interior cell in half
// h.  Only interior cells are ever read: the others (rows outside the
// interior, the cells i = 0 and i = NX+1 and the padding of each row) are
// the domain's halo, zero in LDS.  Formed once per launch, so that a ring
// load is one round of independent loads with no index arithmetic.
__device__ __forceinline__ void res_ring_table(uint2 *rt, const ResTile &tl, int RJ, int HX, unsigned HXY, int NX,
                                               int NY, int NZ)
{
    const unsigned nring = (unsigned)(2 * tl.tj + 2 * tl.tk) * (unsigned)HX;
    for (unsigned e = threadIdx.x; e < nring; e += kResThreads) {
        const unsigned hr = e / (unsigned)HX, m = e - hr * (unsigned)HX;
        int jj, kk;
        if (hr < (unsigned)tl.tj) {
            jj = (int)hr + 1;
            kk = 0;
        } else if (hr < 2u * tl.tj) {
            jj = (int)hr - tl.tj + 1;
            kk = tl.tk + 1;
        } else if (hr < 2u * tl.tj + tl.tk) {
            jj = 0;
            kk = (int)hr - 2 * tl.tj + 1;
        } else {
            jj = tl.tj + 1;
            kk = (int)hr - 2 * tl.tj - tl.tk + 1;
        }
        const int j = tl.j0 + jj - 1, k = tl.k0 + kk - 1;
        unsigned go = 0, bits = 0;
        if (j >= 1 && j <= NY && k >= 1 && k <= NZ) {
            go = m + (unsigned)HX * (unsigned)j + HXY * (unsigned)k;
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                const int i = 2 * (int)m + ((h + j + k) & 1);
                if (i >= 1 && i <= NX) bits |= 1u << (16 + h);
            }
        }
        rt[e] = make_uint2(go, (unsigned)((kk * RJ + jj) * HX + (int)m) | bits);
    }
}

constexpr int kRingBatch = 4;  // ring loads a thread has in flight at once

// The ring cells of half h from x in HBM (xg: that half) into LDS xo, by the
// table of res_ring_table.  Every thread must call it; ends with a barrier.
template <typename T>
__device__ __forceinline__ void res_load_ring(T *xo, const T *xg, const uint2 *rt, unsigned nring, int h)
{
    const unsigned sh = 16u + (unsigned)h;
#pragma unroll 1
    for (unsigned e0 = 0; e0 < nring; e0 += kRingBatch * kResThreads) {
        T v[kRingBatch];
        unsigned lo[kRingBatch];
        bool ok[kRingBatch];
#pragma unroll
        for (int q = 0; q < kRingBatch; ++q) {
            const unsigned e = e0 + threadIdx.x + (unsigned)(q * kResThreads);
            const uint2 w = e < nring ? rt[e] : make_uint2(0u, 0u);
            ok[q] = ((w.y >> sh) & 1u) != 0u;
            lo[q] = w.y & 0xffffu;
            v[q] = ok[q] ? sc_load(xg + w.x) : T(0);
        }
#pragma unroll
        for (int q = 0; q < kRingBatch; ++q)
            if (ok[q]) xo[lo[q]] = v[q];
    }
    __syncthreads();
}

// Row r of a tile's tj x tk rows in slot order: the rows on the tile's edge
// (the ones neighbouring tiles read) first, so that the cells of a pass's
// edge phase, which is on the path from tile to tile, take the low slots of
// every thread and spread over all of them.
__device__ __forceinline__ void res_row(int r, int tj, int tk, int &jj, int &kk)
{
    const int ij = max(tj - 2, 0), ik = max(tk - 2, 0);
    const int nedge = tj * tk - ij * ik;
    if (r >= nedge) {
        const int q = r - nedge;
        jj = 2 + q % ij;
        kk = 2 + q / ij;
        return;
    }
    if (r < tj) {  // first plane of the tile
        jj = r + 1;
        kk = 1;
        return;
    }
    r -= tj;
    if (tk > 1 && r < tj) {  // last plane
        jj = r + 1;
        kk = tk;
        return;
    }
    if (tk > 1) r -= tj;
    // planes in between: first and last row of each
    if (tj > 1) {
        kk = 2 + r / 2;
        jj = (r & 1) ? tj : 1;
    } else {
        kk = 2 + r;
        jj = 1;
    }
}

// One slot of a thread: cell (m, j, k) of the tile in both halves.  e = its
// element offset in an LDS half, base = its offset in an HBM half; edge =
// its row is on the tile's edge (read by a neighbouring tile).
struct ResSlot {
    unsigned e, base;
    bool own, edge;
};

// Per-cell coefficients of the resident launches, formed once per solve
// (they depend on the operator only): for every interior cell of both
// halves the diagonal dd = diag(Cond), the sink Σconv + ap0 of the flux
// form, and the plane-relative indices of its six conductances in the
// staged tables (rb_stage_into's layout), two 16-bit indices to a word.
// The same rb_cond the per-colour kernels evaluate, from the same staged
// tables, so the values are bitwise theirs.
template <typename T>
__global__ void __launch_bounds__(256) k_rb_res_coef(RbGeo<T> g, ResCoef<T> cf)
{
    const RbThread t = rb_thread(g);
    const T *stab = rb_stage(g, t.k);
    if (!t.live) return;
    const unsigned HX = (unsigned)g.HX, HXY = HX * (unsigned)g.MY, HS = (unsigned)g.HS;
    const unsigned base = (unsigned)t.m + HX * (unsigned)t.j + HXY * (unsigned)t.k;
    const int K = g.K, n2 = 2 * K * K;
    const int h0 = (t.j + t.k) & 1;
#pragma unroll
    for (int a = 0; a < 2; ++a) {
        const int h = h0 ^ a, i = 2 * t.m + a;
        if (i < 1 || i > g.NX) continue;
        const unsigned co = (h ? HS : 0u) + base, no = (h ? 0u : HS) + base;
        const Cls q = rb_cls(g.cls, g.cls, co, no, a, HX, HXY);
        const Cond<T> o = rb_cond(g, stab, q, i, t.j, t.k);
        cf.dgl[co] = diag(o);
        if (cf.snk) cf.snk[co] = o.cxp + o.cxm + o.cyp + o.cym + o.czp + o.czm + o.ap0;
        const int b = q.c;
        const unsigned ixm = ((i > 1 ? K : 0) + q.xm) * K + b;
        const unsigned ixp = ((i < g.NX ? K : 0) + q.xp) * K + b;
        const unsigned iym = n2 + ((t.j > 1 ? K : 0) + q.ym) * K + b;
        const unsigned iyp = n2 + ((t.j < g.NY ? K : 0) + q.yp) * K + b;
        const unsigned izm = 2 * n2 + q.zm * K + b;
        const unsigned izp = 2 * n2 + K * K + q.zp * K + b;
        cf.ix[co] = ixm | (ixp << 16);
        cf.ix[2 * (size_t)HS + co] = iym | (iyp << 16);
        cf.ix[4 * (size_t)HS + co] = izm | (izp << 16);
    }
}

template <typename T>
hipError_t launch_rb_res_coef(const RbGeo<T> &g, const ResCoef<T> &cf, hipStream_t s)
{
    hipLaunchKernelGGL(k_rb_res_coef<T>, dim3(g.nb1), dim3(BX, BY), rb_lds(g), s, g, cf);
    return hipGetLastError();
}

// Up to two per-thread sums over the 1024-thread workgroup into the
// workgroup's partials (rbuf layout).  Every thread must call it.
template <int NS>
__device__ __forceinline__ void res_partials(const dd (&v)[2], dd *red, const RedBuf &rbuf)
{
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
#pragma unroll
    for (int s = 0; s < NS; ++s) {
        const dd w = row_sum(v[s]);
        if ((lane & 15) == 15) red[s * 68 + wave * 4 + (lane >> 4)] = w;
    }
    __syncthreads();
    if (wave == 0) {
#pragma unroll
        for (int s = 0; s < NS; ++s) {
            const dd z = row_sum(red[s * 68 + lane]);
            if ((lane & 15) == 15) red[s * 68 + 64 + (lane >> 4)] = z;
        }
    }
    __syncthreads();
    if (tid == 0) {
#pragma unroll
        for (int s = 0; s < NS; ++s) {
            const dd *q = red + s * 68 + 64;
            const dd t = dd_add_tree(dd_add_tree(q[0], q[1]), dd_add_tree(q[2], q[3]));
            rbuf.partials[6 * (size_t)blockIdx.x + 2 * s] = t.hi;
            rbuf.partials[6 * (size_t)blockIdx.x + 2 * s + 1] = t.lo;
        }
    }
}

// The exactly rounded totals (hi + lo, as the finalize kernel forms them)
// of NS sums given as the double-double partials of G workgroups, in
// every thread.  Every thread must call it; G <= 1024.
template <int NS, bool SC>
__device__ __forceinline__ void res_total(const double *part, int stride, unsigned G, dd *red, double (&tot)[2])
{
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    dd v[2] = {{0.0, 0.0}, {0.0, 0.0}};
    if ((unsigned)tid < G) {
#pragma unroll
        for (int q = 0; q < NS; ++q) {
            const double *w = part + (size_t)stride * tid + 2 * q;
            v[q] = SC ? dd{sc_load(w), sc_load(w + 1)} : dd{w[0], w[1]};
        }
    }
#pragma unroll
    for (int q = 0; q < NS; ++q) {
        const dd w = row_sum(v[q]);
        if ((lane & 15) == 15) red[q * 68 + wave * 4 + (lane >> 4)] = w;
    }
    __syncthreads();
    if (wave == 0) {
#pragma unroll
        for (int q = 0; q < NS; ++q) {
            const dd z = row_sum(red[q * 68 + lane]);
            if ((lane & 15) == 15) red[q * 68 + 64 + (lane >> 4)] = z;
        }
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < NS; ++q) {
        const dd *r = red + q * 68 + 64;
        const dd t = dd_add_tree(dd_add_tree(r[0], r[1]), dd_add_tree(r[2], r[3]));
        tot[q] = t.hi + t.lo;
    }
    __syncthreads();  // red is reused
}

// What a tile's passes use besides its per-slot registers.
template <typename T>
struct ResView {
    T *xs0, *xs1;      // the tile's two colour halves in LDS (rows and ring)
    const T *tabs;     // staged conductance tables of the tile's planes
    unsigned *bar;
    ResTile tl;
    int RJ, HX, NX, NY, NZ;
    unsigned HXY, zs;  // zs: the LDS offset of the next plane
    size_t HS;
    const uint2 *rt;   // the ring table (res_ring_table)
    unsigned nring;    // its entries
};

// The tile's view of the launch's dynamic LDS (rb_res_lds gives its size):
// both colour halves of its rows and ring, the staged tables of its planes,
// the sums' scratch, then the ring table.
template <typename T>
__device__ __forceinline__ ResView<T> res_view(const RbGeo<T> &g, const RbRes &pl, unsigned char *smem)
{
    const int HX = g.HX, RJ = pl.TJ + 2;
    const unsigned hsz = (unsigned)(RJ * (pl.TK + 2) * HX + 1);  // one LDS half (+1 pad)
    const int ntab = 6 * g.K * g.K + g.K;
    ResView<T> w;
    w.xs0 = reinterpret_cast<T *>(smem);
    w.xs1 = w.xs0 + hsz;
    w.tabs = w.xs1 + hsz;
    w.bar = pl.bar;
    w.tl = res_tile(pl, g.NY, g.NZ);
    w.RJ = RJ;
    w.HX = HX;
    w.NX = g.NX;
    w.NY = g.NY;
    w.NZ = g.NZ;
    w.HXY = (unsigned)HX * (unsigned)g.MY;
    w.zs = (unsigned)(RJ * HX);
    w.HS = g.HS;
    const size_t tend = ((2 * (size_t)hsz + (size_t)pl.TK * ntab) * sizeof(T) + 15) & ~(size_t)15;
    w.rt = reinterpret_cast<const uint2 *>(smem + tend + 2 * 68 * sizeof(dd));
    w.nring = (unsigned)(2 * w.tl.tj + 2 * w.tl.tk) * (unsigned)HX;
    return w;
}

// The sums' scratch of the view (2 x 68 double-double words, before the
// ring table).
template <typename T>
__device__ __forceinline__ dd *res_red(const ResView<T> &w)
{
    return reinterpret_cast<dd *>(const_cast<uint2 *>(w.rt)) - 2 * 68;
}

// The per-slot registers of a tile's cells (see k_rb_resident).
template <typename T>
struct ResRegs {
    ResSlot sl[kResSlots];
    unsigned par[kResSlots];  // (j + k) & 1 of the slot
    bool valid[kResSlots][2];
    T dgl[kResSlots][2], snk[kResSlots][2];
    unsigned ix[kResSlots][2][3];
};

// Preconditioner! :gs, the five red-black sweeps from x = 0 (LDS zeroed
// by the caller) on the right-hand side bv, into LDS and into x in HBM (the
// rows on the tile's edge every pass, for the neighbours; the others after
// the last sweep when all_rows, else left to the caller's LDS); publishes
// pass counts f0+1 .. f0+10.  Every thread must call it.
template <typename T, bool FLUX>
__device__ __forceinline__ void res_gs5(const ResView<T> &w, T *x, const ResRegs<T> &rg,
                                        const T (&bv)[kResSlots][2], unsigned f0, bool all_rows = true)
{
    const unsigned HX = (unsigned)w.HX, zs = w.zs;
#pragma unroll 1
    for (int sweep = 0; sweep < 5; ++sweep) {
#pragma unroll
        for (int half = 0; half < 2; ++half) {
            // one colour pass: colour 0 updates half 1, colour 1 half 0
            const int h = 1 - half;
            const int pass = 2 * sweep + half;
            T *xo = h ? w.xs1 : w.xs0;
            T *xt = h ? w.xs0 : w.xs1;
            T *xg = x + (h ? w.HS : 0);
            // the tile's inner rows, then (after the ring rows of the half
            // the previous pass updated have arrived) its edge rows
#pragma unroll
            for (int phase = 0; phase < 2; ++phase) {
                if (phase == 1 && pass > 0) {
                    res_wait(w.bar, w.tl, f0 + (unsigned)pass);
                    res_load_ring(xt, (const T *)(x + (h ? 0 : w.HS)), w.rt, w.nring, 1 - h);
                }
#pragma unroll
                for (int u = 0; u < kResSlots; ++u) {
                    if (!rg.sl[u].own || rg.sl[u].edge != (phase == 1)) continue;
                    // the offsets and index words are opaque to the
                    // compiler here, so that it forms the addresses inside
                    // the loop instead of hoisting some hundred of them
                    // into registers
                    unsigned e = rg.sl[u].e, base = rg.sl[u].base, w0 = rg.ix[u][h][0], w1 = rg.ix[u][h][1],
                             w2 = rg.ix[u][h][2];
                    asm volatile("" : "+v"(e), "+v"(base), "+v"(w0), "+v"(w1), "+v"(w2));
                    const unsigned sx = (unsigned)h ^ rg.par[u];
                    Nb<T> n;
                    n.xm = xt[e - 1u + sx];
                    n.xp = xt[e + sx];
                    n.ym = xt[e - HX];
                    n.yp = xt[e + HX];
                    n.zm = xt[e - zs];
                    n.zp = xt[e + zs];
                    const T pp = xo[e];
                    Cond<T> o;
                    o.xm = w.tabs[w0 & 0xffffu];
                    o.xp = w.tabs[w0 >> 16];
                    o.ym = w.tabs[w1 & 0xffffu];
                    o.yp = w.tabs[w1 >> 16];
                    o.zm = w.tabs[w2 & 0xffffu];
                    o.zp = w.tabs[w2 >> 16];
                    T xn;
                    if (FLUX) {
                        const T f = o.xp * (n.xp - pp) + o.xm * (n.xm - pp) + o.yp * (n.yp - pp) +
                                    o.ym * (n.ym - pp) + o.zp * (n.zp - pp) + o.zm * (n.zm - pp);
                        xn = pp + ((f - rg.snk[u][h] * pp) - bv[u][h]) / rg.dgl[u][h];
                    } else {
                        const T dp = (((rb_offdiag(o, n) - bv[u][h]) / rg.dgl[u][h] - pp)) * T(1);
                        xn = pp + T(1) * dp;
                    }
                    if (rg.valid[u][h]) {
                        xo[e] = xn;
                        // HBM receives the rows neighbours read, and in the
                        // last sweep (final values) every row if asked
                        if (phase == 1) sc_store(xg + base, xn);
                        else if (sweep == 4 && all_rows) xg[base] = xn;
                    }
                    // one slot's loads in flight at a time: sixteen waves
                    // per CU hide the LDS latency, and the registers stay
                    // within 128
                    __builtin_amdgcn_sched_barrier(0);
                }
            }
            res_publish(w.bar, f0 + (unsigned)pass + 1u);
        }
    }
}

// CalcAX! (src/NonUniform.jl:454-511) of x, read from LDS (both halves and
// their rings in place): ax = A x on the tile's cells (not stored when ax
// is null), tav = the same values for the caller, v[0] += Σ ax·y with y = x
// (DYX), dy, or ax itself, v[1] += Σ ax·(dz_bb ? bv : dz) when NS = 2.
// Every thread must call it.
template <typename T, bool FLUX, int NS, bool DYX = false>
__device__ __forceinline__ void res_calc_ax_core(const ResView<T> &w, const ResRegs<T> &rg,
                                                 const T (&bv)[kResSlots][2], T *ax, const T *dy, const T *dz,
                                                 bool dz_bb, dd (&v)[2], T (&tav)[kResSlots][2])
{
    const unsigned HX = (unsigned)w.HX, zs = w.zs;
#pragma unroll
    for (int u = 0; u < kResSlots; ++u) {
        if (!rg.sl[u].own) continue;
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            unsigned e = rg.sl[u].e, base = rg.sl[u].base, w0 = rg.ix[u][h][0], w1 = rg.ix[u][h][1],
                     w2 = rg.ix[u][h][2];
            asm volatile("" : "+v"(e), "+v"(base), "+v"(w0), "+v"(w1), "+v"(w2));
            const T *xo = h ? w.xs1 : w.xs0, *xt = h ? w.xs0 : w.xs1;
            const unsigned sx = (unsigned)h ^ rg.par[u];
            Nb<T> n;
            n.xm = xt[e - 1u + sx];
            n.xp = xt[e + sx];
            n.ym = xt[e - HX];
            n.yp = xt[e + HX];
            n.zm = xt[e - zs];
            n.zp = xt[e + zs];
            const T x0 = xo[e];
            Cond<T> o;
            o.xm = w.tabs[w0 & 0xffffu];
            o.xp = w.tabs[w0 >> 16];
            o.ym = w.tabs[w1 & 0xffffu];
            o.yp = w.tabs[w1 >> 16];
            o.zm = w.tabs[w2 & 0xffffu];
            o.zp = w.tabs[w2 >> 16];
            // rb_apply with the diagonal (and sink) formed once per solve
            T av;
            if (FLUX) {
                const T f = o.xp * (n.xp - x0) + o.xm * (n.xm - x0) + o.yp * (n.yp - x0) + o.ym * (n.ym - x0) +
                            o.zp * (n.zp - x0) + o.zm * (n.zm - x0);
                av = f - rg.snk[u][h] * x0;
            } else {
                av = (rb_offdiag(o, n) - rg.dgl[u][h] * x0) * T(1);
            }
            tav[u][h] = av;
            if (rg.valid[u][h]) {
                const size_t c = (h ? w.HS : 0) + base;
                if (ax) ax[c] = av;
                const T yv = DYX ? x0 : dy ? dy[c] : av;
                v[0] = dd_add(v[0], (double)av * (double)yv);
                if (NS >= 2) {
                    const T zv = dz_bb ? bv[u][h] : dz[c];
                    v[1] = dd_add(v[1], (double)av * (double)zv);
                }
            }
            __builtin_amdgcn_sched_barrier(0);
        }
    }
}

// res_calc_ax_core after the ring rows of half 0 (updated by the last
// pass, published as count want) have arrived.
template <typename T, bool FLUX, int NS>
__device__ __forceinline__ void res_calc_ax(const ResView<T> &w, const T *x, const ResRegs<T> &rg,
                                            const T (&bv)[kResSlots][2], unsigned want, T *ax, const T *dy,
                                            const T *dz, bool dz_bb, dd (&v)[2], T (&tav)[kResSlots][2])
{
    res_wait(w.bar, w.tl, want);
    res_load_ring(w.xs0, x, w.rt, w.nring, 0);
    res_calc_ax_core<T, FLUX, NS>(w, rg, bv, ax, dy, dz, dz_bb, v, tav);
}

// The tables of the tile's planes in one round of loads (the entries
// rb_stage_into forms, per plane, with the same operations), and the
// tile's slots with their per-cell constants (cf, formed once per solve).
// Ends without a barrier.
template <typename T, bool FLUX>
__device__ __forceinline__ void res_setup(const RbGeo<T> &g, const ResCoef<T> &cf, const ResView<T> &w, int ntab,
                                          T *tabs, uint2 *rt, ResRegs<T> &rg)
{
    const int tid = threadIdx.x, K = g.K;
    const ResTile &tl = w.tl;
    res_ring_table(rt, tl, w.RJ, w.HX, w.HXY, g.NX, g.NY, g.NZ);
    {
        const int n = 2 * K * K, per = n + K;
        for (int idx = tid; idx < tl.tk * per; idx += kResThreads) {
            const int kk = idx / per, a = idx - kk * per;
            const int k = tl.k0 + kk;
            T *t = tabs + (size_t)kk * ntab;
            const T dz_k = g.dZ[k];
            if (a < n) {
                const T f = g.tab[a];
                t[a] = f * g.dy0 * dz_k * g.ddx;
                t[n + a] = f * g.dx0 * dz_k * g.ddy;
                t[2 * n + a] = g.ZF[(size_t)(k - 1) * K * K + a];
            } else if (g.ap) {
                const int c = a - n;
                t[3 * n + c] = (T)(g.capc[c] * g.dx0d * g.dy0d * g.dZd[k] * g.ddt);
            }
        }
    }
    const unsigned HX = (unsigned)w.HX;
    const unsigned nown = (unsigned)(tl.tj * tl.tk) * HX;
#pragma unroll
    for (int u = 0; u < kResSlots; ++u) {
        const unsigned qi = (unsigned)tid + (unsigned)kResThreads * u;
        rg.sl[u].own = qi < nown;
        const unsigned qq = rg.sl[u].own ? qi : 0u;
        const int row = (int)(qq / HX);
        const int m = (int)(qq - (unsigned)row * HX);
        int jj, kk;
        res_row(row, tl.tj, tl.tk, jj, kk);
        const int j = tl.j0 + jj - 1, k = tl.k0 + kk - 1;
        rg.sl[u].edge = row < tl.tj * tl.tk - max(tl.tj - 2, 0) * max(tl.tk - 2, 0);
        rg.par[u] = (unsigned)((j + k) & 1);
        rg.sl[u].e = (unsigned)((kk * w.RJ + jj) * w.HX + m);
        rg.sl[u].base = (unsigned)m + HX * (unsigned)j + w.HXY * (unsigned)k;
        // the tile's plane offset in the staged tables, in both 16-bit halves
        const unsigned tb = (unsigned)((kk - 1) * ntab) * 0x10001u;
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            const size_t oh = (h ? w.HS : 0) + rg.sl[u].base;
            const int i = 2 * m + (int)((unsigned)h ^ rg.par[u]);
            rg.valid[u][h] = rg.sl[u].own && i >= 1 && i <= g.NX;
            // cells outside the interior have no coefficients (never stored)
            rg.dgl[u][h] = rg.valid[u][h] ? cf.dgl[oh] : T(1);
            rg.snk[u][h] = FLUX && rg.valid[u][h] ? cf.snk[oh] : T(0);
#pragma unroll
            for (int q = 0; q < 3; ++q) rg.ix[u][h][q] = rg.valid[u][h] ? cf.ix[2 * q * w.HS + oh] + tb : tb;
        }
    }
}

// Up to two per-thread sums over the workgroup into its device-scope
// partials wp[4 * blockIdx.x + 2(s0 + s) + {0, 1}] (hi, lo), for a grid
// reduction inside the launch.  Every thread must call it.
template <int NS>
__device__ __forceinline__ void res_wpartials(const dd (&v)[2], dd *red, double *wp, int s0 = 0)
{
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
#pragma unroll
    for (int q = 0; q < NS; ++q) {
        const dd w = row_sum(v[q]);
        if ((lane & 15) == 15) red[q * 68 + wave * 4 + (lane >> 4)] = w;
    }
    __syncthreads();
    if (wave == 0) {
#pragma unroll
        for (int q = 0; q < NS; ++q) {
            const dd z = row_sum(red[q * 68 + lane]);
            if ((lane & 15) == 15) red[q * 68 + 64 + (lane >> 4)] = z;
        }
    }
    __syncthreads();
    if (tid == 0) {
#pragma unroll
        for (int q = 0; q < NS; ++q) {
            const dd *rq = red + q * 68 + 64;
            const dd t = dd_add_tree(dd_add_tree(rq[0], rq[1]), dd_add_tree(rq[2], rq[3]));
            sc_store(wp + 4 * (size_t)blockIdx.x + 2 * (s0 + q), t.hi);
            sc_store(wp + 4 * (size_t)blockIdx.x + 2 * (s0 + q) + 1, t.lo);
        }
    }
    __syncthreads();
}

// BICG2!(θ, p_, s_, alpha, omega); Triad!(r, t, s, -omega); Σ r², Σ r·r0
// (src/NonUniform.jl:328-332, 289) on the tile's cells, as k_rb_bicg_update
// forms them: s_ from LDS, s = bv, t = tav; θ in the standard layout; with
// xbest the iterate is copied there before it is overwritten.
template <typename T>
__device__ __forceinline__ void res_bicg_update(const RbGeo<T> &g, const ResView<T> &w, const ResRegs<T> &rg,
                                                const T (&bv)[kResSlots][2], const T (&tav)[kResSlots][2],
                                                double *th, double *xbest, T *r, const T *p_, const T *r0,
                                                double alpha, double omega, dd (&v)[2])
{
    const T r_omega = (T)(-omega);
    const unsigned HX = (unsigned)w.HX, HXY = w.HXY;
    const size_t sy = (size_t)g.MX, sz = (size_t)g.MX * (size_t)g.MY;
    v[0] = dd{0.0, 0.0};
    v[1] = dd{0.0, 0.0};
#pragma unroll
    for (int u = 0; u < kResSlots; ++u) {
        const unsigned b0 = opq(rg.sl[u].base);
        const unsigned k = b0 / HXY, jr = b0 - k * HXY, j = jr / HX, m = jr - j * HX;
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            if (!rg.valid[u][h]) continue;
            const size_t c = (h ? w.HS : 0) + b0;
            const size_t cs = (size_t)(2 * m + ((unsigned)h ^ rg.par[u])) + sy * j + sz * k;
            const T sh = (h ? w.xs1 : w.xs0)[opq(rg.sl[u].e)];
            const double xo = th[cs];
            const double xn = xo + (alpha * (double)p_[c] + omega * (double)sh);
            const T rv = r_omega * tav[u][h] + bv[u][h];
            const T r0v = r0[c];
            v[0] = dd_add(v[0], (double)rv * (double)rv);
            v[1] = dd_add(v[1], (double)rv * (double)r0v);
            if (xbest) xbest[cs] = xo;
            th[cs] = xn;
            r[c] = rv;
        }
    }
}

// MODE: 0 = the preconditioner alone; 1 = and Σ bb·x (CG!'s ρ); 2, 3 =
// and CalcAX! ax = A x of its result with one or two fused sums,
// s0 = Σ ax·(dy ? dy : ax), s1 = Σ ax·dz (dz_is_bb: dz is the right-hand
// side bb the launch formed); 4 = PBiCGSTAB!'s second half-iteration
// (mode FIRST_BICG_S) whole: alpha from the previous launch's partials of
// (q, r0), s = r - alpha q, s_ = M⁻¹s, t = A s_, omega from (t, t) and
// (t, s) reduced over the grid inside the launch, BICG2!/Triad! on θ and r
// and the partials of (r, r) and (r, r0) for the finalize launch.
template <typename T, int MODE, bool FLUX>
__global__ void __launch_bounds__(kResThreads) k_rb_resident(int mode, T *__restrict__ x, T *__restrict__ bb,
                                                             const T *__restrict__ r, const T *__restrict__ q,
                                                             RbGeo<T> g, RbRes pl, ResCoef<T> cf, const KState *st,
                                                             RedBuf rbuf, T *__restrict__ ax, const T *__restrict__ dy,
                                                             const T *__restrict__ dz, int dz_is_bb, ResUpd<T> up)
{
    if (halted(st)) return;  // uniform: every workgroup reads the same state
    extern __shared__ __align__(16) unsigned char h3d_dyn_smem[];
    const int tid = threadIdx.x;
    const int ntab = 6 * g.K * g.K + g.K;
    const ResView<T> w = res_view(g, pl, h3d_dyn_smem);
    const unsigned hsz = (unsigned)(w.xs1 - w.xs0);  // one LDS half
    T *tabs = const_cast<T *>(w.tabs);
    dd *red = res_red(w);
    uint2 *rt = const_cast<uint2 *>(w.rt);
    const size_t HS = g.HS;

    // tables, slots and per-cell constants
    ResRegs<T> rg;
    res_setup<T, FLUX>(g, cf, w, ntab, tabs, rt, rg);

    // the right-hand side (formed as k_rb_first forms it)
    T beta = T(0), omg = T(0), r_alpha = T(0);
    bool first_itr = false;
    double alpha = 0.0;
    if (MODE == 4) {
        // FIN_BICG_ALPHA on the previous launch's partials of (q, r0), in
        // every workgroup: alpha = rho / (q, r0)
        double tq[2];
        res_total<1, false>(up.apart, 6, gridDim.x, red, tq);
        alpha = st->rho / tq[0];
        r_alpha = (T)(-alpha);
    } else if (mode == FIRST_BICG_P) {
        first_itr = (st->itr == 1);
        beta = (T)st->beta;
        omg = (T)st->omega;
    } else if (mode == FIRST_BICG_S) {
        r_alpha = (T)(-st->alpha);
    }
    T bv[kResSlots][2];
#pragma unroll
    for (int u = 0; u < kResSlots; ++u)
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            const size_t oh = (h ? HS : 0) + rg.sl[u].base;
            T v;
            if (mode == FIRST_BICG_P) v = first_itr ? r[oh] : r[oh] + beta * (bb[oh] - omg * q[oh]);
            else if (mode == FIRST_BICG_S) v = r_alpha * q[oh] + r[oh];
            else v = bb[oh];
            bv[u][h] = v;
        }
    __syncthreads();  // tables staged

    // pass counter at the start of this launch (equal in every workgroup)
    const unsigned f0 = bar_load(pl.bar + kBarFlags + kBarFlagStride * blockIdx.x);

    // x = 0 everywhere in LDS (the preconditioner starts from zero; the
    // domain's halo rows and columns stay zero)
    for (unsigned e = tid; e < 2 * hsz; e += kResThreads) w.xs0[e] = T(0);
    // the bb stores of modes that form it (interior cells only)
    if (mode != FIRST_COPY) {
#pragma unroll
        for (int u = 0; u < kResSlots; ++u)
#pragma unroll
            for (int h = 0; h < 2; ++h)
                if (rg.valid[u][h]) bb[(h ? HS : 0) + rg.sl[u].base] = bv[u][h];
    }
    __syncthreads();  // x zeroed

    res_gs5<T, FLUX>(w, x, rg, bv, f0);

    if (MODE == 1) {
        // Σ bb·x over the tile's interior cells of both halves (x final:
        // half 0 from the last pass, half 1 from the one before; each
        // thread reads back its own cells)
        dd v[2] = {{0.0, 0.0}, {0.0, 0.0}};
#pragma unroll
        for (int u = 0; u < kResSlots; ++u) {
            if (rg.valid[u][0]) v[0] = dd_add(v[0], (double)bv[u][0] * (double)w.xs0[rg.sl[u].e]);
            if (rg.valid[u][1]) v[0] = dd_add(v[0], (double)bv[u][1] * (double)w.xs1[rg.sl[u].e]);
        }
        res_partials<1>(v, red, rbuf);
    }
    if (MODE >= 2) {
        dd v[2] = {{0.0, 0.0}, {0.0, 0.0}};
        T tav[kResSlots][2];  // mode 4: t kept for the update
        res_calc_ax<T, FLUX, (MODE >= 3 ? 2 : 1)>(w, x, rg, bv, f0 + 10u, ax, dy, dz, dz_is_bb != 0, v, tav);
        if (MODE == 2) {
            res_partials<1>(v, red, rbuf);
        } else if (MODE == 3) {
            res_partials<2>(v, red, rbuf);
        } else {
            // omega: the grid's (t, t) and (t, s) in every workgroup
            res_wpartials<2>(v, red, up.wpart);
            res_grid_sync(pl.bar);
            double tw[2];
            res_total<2, true>(up.wpart, 4, gridDim.x, red, tw);
            // FIN_BICG_OMEGA (NonUniform.jl:320-326)
            const bool breakdown = fabs(tw[0]) < kFloatMin;
            const double omega = breakdown ? 0.0 : tw[1] / tw[0];
            // every workgroup has read the state by now: publish the scalars
            if (tid == 0) {
                KState *sw = const_cast<KState *>(st);
                sw->alpha = alpha;
                if (breakdown) sw->status = ST_BREAKDOWN;
                else sw->omega = omega;
            }
            if (!breakdown) {
                dd wv[2];
                res_bicg_update<T>(g, w, rg, bv, tav, up.th, up.xbest != nullptr && st->improved != 0 ? up.xbest
                                                                                                        : nullptr,
                                   up.r, up.p_, up.r0, alpha, omega, wv);
                res_partials<2>(wv, red, rbuf);
            }
        }
    }
}

// The whole PBiCGSTAB! loop (src/NonUniform.jl:283-349, the iterations
// after CalcRK! and the first ρ) with the :gs preconditioner in ONE launch
// of the resident grid.  Each iteration is the two halves the launches
// above run (k_rb_resident modes FIRST_BICG_P with CalcAX! and 4), with
// the separate launches' boundaries replaced by in-launch grid barriers:
//   p = r (first iteration) or BiCG1! p = r + beta (p - omega q);
//   p_ = M⁻¹p; q = A p_; (q, r0) -> barrier -> alpha = rho / (q, r0);
//   s = r - alpha q; s_ = M⁻¹s; t = A s_; (t, t), (t, s) -> barrier ->
//   omega; BICG2!/Triad! on θ and r; (r, r), (r, r0) -> barrier ->
//   FIN_BICG_RES_RHO (residual, history, convergence, next rho and beta).
// The scalar steps run redundantly in every workgroup on the same exactly
// rounded totals, so every workgroup holds the same solver state (in LDS);
// workgroup 0 writes the history and, at the end, the state.  The tables,
// slots and per-cell constants are set up once per solve instead of twice
// per iteration.  s and t live in registers (neither is stored), s_ goes
// to HBM only on the tiles' edge rows (the update reads it from LDS), and
// q is reused from CalcAX!'s registers.  The three partial buffers of the in-launch
// sums (4 doubles per workgroup each) are each reused once per iteration:
// every workgroup has read one before any can pass the next barrier.
template <typename T, bool FLUX>
__global__ void __launch_bounds__(kResThreads) k_rb_bicg_solve(RbGeo<T> g, RbRes pl, ResCoef<T> cf, KState *st,
                                                               double *hist, ResBicg<T> v)
{
    __shared__ KState ks;
    const int tid = threadIdx.x;
    if (tid == 0) ks = *st;
    __syncthreads();
    if (ks.status != ST_RUNNING) return;  // uniform
    extern __shared__ __align__(16) unsigned char h3d_dyn_smem[];
    const int ntab = 6 * g.K * g.K + g.K;
    const ResView<T> w = res_view(g, pl, h3d_dyn_smem);
    const unsigned hsz = (unsigned)(w.xs1 - w.xs0);  // one LDS half
    T *tabs = const_cast<T *>(w.tabs);
    dd *red = res_red(w);
    uint2 *rt = const_cast<uint2 *>(w.rt);
    const size_t HS = g.HS;
    ResRegs<T> rg;
    res_setup<T, FLUX>(g, cf, w, ntab, tabs, rt, rg);
    __syncthreads();  // tables staged
    double *wq = pl.wpart, *wt = pl.wpart + 4 * kResMaxWG, *wr = pl.wpart + 8 * kResMaxWG;
    unsigned fc = bar_load(pl.bar + kBarFlags + kBarFlagStride * blockIdx.x);

#pragma unroll 1
    for (;;) {
        // ---- p formed; p_ = M⁻¹p; q = A p_; Σ q·r0
        T bv[kResSlots][2];
        {
            const bool first = ks.itr == 1;
            const T beta = (T)ks.beta, omg = (T)ks.omega;
#pragma unroll
            for (int u = 0; u < kResSlots; ++u)
#pragma unroll
                for (int h = 0; h < 2; ++h) {
                    const size_t oh = (h ? HS : 0) + opq(rg.sl[u].base);
                    bv[u][h] = first ? v.r[oh] : v.r[oh] + beta * (v.p[oh] - omg * v.q[oh]);
                }
        }
        for (unsigned e = tid; e < 2 * hsz; e += kResThreads) w.xs0[e] = T(0);
#pragma unroll
        for (int u = 0; u < kResSlots; ++u)
#pragma unroll
            for (int h = 0; h < 2; ++h)
                if (rg.valid[u][h]) v.p[(h ? HS : 0) + opq(rg.sl[u].base)] = bv[u][h];
        __syncthreads();
        res_gs5<T, FLUX>(w, v.p_, rg, bv, fc);
        dd sv[2] = {{0.0, 0.0}, {0.0, 0.0}};
        T tav[kResSlots][2] = {};
        res_calc_ax<T, FLUX, 1>(w, v.p_, rg, bv, fc + 10u, v.q, v.r0, nullptr, false, sv, tav);
        fc += 10u;
        res_wpartials<1>(sv, red, wq);
        res_grid_sync(pl.bar);
        double tq[2];
        res_total<1, true>(wq, 4, gridDim.x, red, tq);
        const double alpha = ks.rho / tq[0];  // FIN_BICG_ALPHA (NonUniform.jl:309)

        // ---- s = r - alpha q; s_ = M⁻¹s; t = A s_; Σ t·t, Σ t·s
        {
            const T r_alpha = (T)(-alpha);
#pragma unroll
            for (int u = 0; u < kResSlots; ++u)
#pragma unroll
                for (int h = 0; h < 2; ++h) bv[u][h] = r_alpha * tav[u][h] + v.r[(h ? HS : 0) + opq(rg.sl[u].base)];
        }
        for (unsigned e = tid; e < 2 * hsz; e += kResThreads) w.xs0[e] = T(0);
        __syncthreads();
        // (the update reads s_ from LDS: only its edge rows go to HBM)
        res_gs5<T, FLUX>(w, v.s_, rg, bv, fc, false);
        sv[0] = dd{0.0, 0.0};
        // (t is used from registers only: not stored)
        res_calc_ax<T, FLUX, 2>(w, v.s_, rg, bv, fc + 10u, nullptr, nullptr, nullptr, true, sv, tav);
        fc += 10u;
        res_wpartials<2>(sv, red, wt);
        res_grid_sync(pl.bar);
        double tw[2];
        res_total<2, true>(wt, 4, gridDim.x, red, tw);
        // FIN_BICG_OMEGA (NonUniform.jl:320-326)
        if (fabs(tw[0]) < kFloatMin) {
            if (tid == 0) {
                ks.alpha = alpha;
                ks.status = ST_BREAKDOWN;
            }
            break;
        }
        const double omega = tw[1] / tw[0];

        // ---- θ += alpha p_ + omega s_; r = s - omega t; Σ r·r, Σ r·r0
        dd wv[2];
        res_bicg_update<T>(g, w, rg, bv, tav, v.th, v.xbest != nullptr && ks.improved != 0 ? v.xbest : nullptr, v.r,
                           v.p_, v.r0, alpha, omega, wv);
        res_wpartials<2>(wv, red, wr);
        res_grid_sync(pl.bar);
        double tr[2];
        res_total<2, true>(wr, 4, gridDim.x, red, tr);
        if (tid == 0) {
            ks.alpha = alpha;
            ks.omega = omega;
            finalize(FIN_BICG_RES_RHO, tr, &ks, blockIdx.x == 0 ? hist : nullptr);
            res_stop_on_error(pl.bar, ks);
        }
        __syncthreads();
        if (ks.status != ST_RUNNING) break;
    }
    __syncthreads();
    if (blockIdx.x == 0 && tid == 0) *st = ks;
}

// The whole CG! loop (src/NonUniform.jl:160-223, the iterations after
// CalcRK!, the first preconditioner application, p = z and ρ_old) with the
// :gs preconditioner in ONE launch of the resident grid.  An iteration:
//   q = A p from LDS (p's ring rows exchanged once), Σ p·q -> barrier ->
//   FIN_CG_ALPHA; θ += α p, r -= α q, Σ r²; z = M⁻¹ r (ten passes in LDS),
//   Σ r·z; both sums -> one barrier -> FIN_CG_RES, then FIN_CG_BETA while
//   the solve runs; p = z + β p, stored and published for the next CalcAX!.
// Sharing one barrier between Σ r² and Σ r·z runs the preconditioner of
// the iteration that converges as well; its z is workspace nothing reads,
// and the scalar steps run in the reference's order, so θ, r, the history
// and the state are those of the per-iteration launches.  As in
// k_rb_bicg_solve the scalar steps run in every workgroup on the same
// exactly rounded totals; workgroup 0 writes the history and the state.
template <typename T, bool FLUX>
__global__ void __launch_bounds__(kResThreads) k_rb_cg_solve(RbGeo<T> g, RbRes pl, ResCoef<T> cf, KState *st,
                                                             double *hist, ResCg<T> v)
{
    __shared__ KState ks;
    const int tid = threadIdx.x;
    if (tid == 0) ks = *st;
    __syncthreads();
    if (ks.status != ST_RUNNING) return;  // uniform
    extern __shared__ __align__(16) unsigned char h3d_dyn_smem[];
    const int HX = g.HX, ntab = 6 * g.K * g.K + g.K;
    const ResView<T> w = res_view(g, pl, h3d_dyn_smem);
    const unsigned hsz = (unsigned)(w.xs1 - w.xs0);  // one LDS half
    T *tabs = const_cast<T *>(w.tabs);
    dd *red = res_red(w);
    uint2 *rt = const_cast<uint2 *>(w.rt);
    const size_t HS = g.HS;
    ResRegs<T> rg;
    res_setup<T, FLUX>(g, cf, w, ntab, tabs, rt, rg);
    // p's cells of the tile into LDS, whose other entries (the domain's
    // halo and the ring) start at zero
    for (unsigned e = tid; e < 2 * hsz; e += kResThreads) w.xs0[e] = T(0);
    __syncthreads();  // tables staged, LDS zeroed
#pragma unroll
    for (int u = 0; u < kResSlots; ++u)
#pragma unroll
        for (int h = 0; h < 2; ++h)
            if (rg.valid[u][h]) (h ? w.xs1 : w.xs0)[rg.sl[u].e] = v.p[(h ? HS : 0) + opq(rg.sl[u].base)];
    double *wq = pl.wpart, *wr = pl.wpart + 4 * kResMaxWG;
    unsigned fc = bar_load(pl.bar + kBarFlags + kBarFlagStride * blockIdx.x);
    const unsigned HXY = w.HXY;
    const size_t sy = (size_t)g.MX, sz = (size_t)g.MX * (size_t)g.MY;

#pragma unroll 1
    for (;;) {
        // ---- q = A p; Σ p·q (the neighbours' p published as count fc)
        res_wait(pl.bar, w.tl, fc);
        res_load_ring(w.xs0, (const T *)v.p, w.rt, w.nring, 0);
        res_load_ring(w.xs1, (const T *)(v.p + HS), w.rt, w.nring, 1);
        dd sv[2] = {{0.0, 0.0}, {0.0, 0.0}};
        T tav[kResSlots][2] = {};
        // (q is used from registers only, not stored; the right-hand side
        // argument is read only with two sums)
        res_calc_ax_core<T, FLUX, 1, true>(w, rg, tav, nullptr, nullptr, nullptr, false, sv, tav);
        res_wpartials<1>(sv, red, wq);
        res_grid_sync(pl.bar);
        double tq[2];
        res_total<1, true>(wq, 4, gridDim.x, red, tq);
        if (tid == 0) finalize(FIN_CG_ALPHA, tq, &ks, nullptr);  // NonUniform.jl:160-176
        __syncthreads();
        if (ks.status != ST_RUNNING) break;

        // ---- θ += α p; r -= α q; Σ r² (as k_rb_cg_update forms them)
        T bv[kResSlots][2];  // the new r: the preconditioner's right-hand side
        {
            const double alpha_d = ks.alpha;
            const T alpha = (T)ks.alpha;
            double *xbest = v.xbest != nullptr && ks.improved != 0 ? v.xbest : nullptr;
            dd rr[2] = {{0.0, 0.0}, {0.0, 0.0}};
#pragma unroll
            for (int u = 0; u < kResSlots; ++u) {
                const unsigned b0 = opq(rg.sl[u].base);
                const unsigned k = b0 / HXY, jr = b0 - k * HXY, j = jr / (unsigned)HX, m = jr - j * (unsigned)HX;
#pragma unroll
                for (int h = 0; h < 2; ++h) {
                    const size_t c = (h ? HS : 0) + b0;
                    const T rv = v.r[c] - alpha * tav[u][h];
                    bv[u][h] = rv;
                    if (!rg.valid[u][h]) continue;
                    const size_t cs = (size_t)(2 * m + ((unsigned)h ^ rg.par[u])) + sy * j + sz * k;
                    const T pv = (h ? w.xs1 : w.xs0)[opq(rg.sl[u].e)];
                    const double xo = v.th[cs];
                    rr[0] = dd_add(rr[0], (double)rv * (double)rv);
                    if (xbest) xbest[cs] = xo;
                    v.th[cs] = xo + alpha_d * (double)pv;
                    v.r[c] = rv;
                }
            }
            res_wpartials<1>(rr, red, wr, 0);
        }

        // ---- z = M⁻¹ r from z = 0; Σ r·z
        for (unsigned e = tid; e < 2 * hsz; e += kResThreads) w.xs0[e] = T(0);
        __syncthreads();
        // (z is read from LDS only: its edge rows alone go to HBM)
        res_gs5<T, FLUX>(w, v.z, rg, bv, fc, false);
        fc += 10u;
        {
            // each thread reads back its own cells (half 0 from the last
            // pass, half 1 from the one before)
            dd rz[2] = {{0.0, 0.0}, {0.0, 0.0}};
#pragma unroll
            for (int u = 0; u < kResSlots; ++u) {
                if (rg.valid[u][0]) rz[0] = dd_add(rz[0], (double)bv[u][0] * (double)w.xs0[rg.sl[u].e]);
                if (rg.valid[u][1]) rz[0] = dd_add(rz[0], (double)bv[u][1] * (double)w.xs1[rg.sl[u].e]);
            }
            res_wpartials<1>(rz, red, wr, 1);
        }
        res_grid_sync(pl.bar);
        double tr[2];
        res_total<2, true>(wr, 4, gridDim.x, red, tr);
        if (tid == 0) {
            finalize(FIN_CG_RES, &tr[0], &ks, blockIdx.x == 0 ? hist : nullptr);  // NonUniform.jl:184-197
            if (ks.status == ST_RUNNING) finalize(FIN_CG_BETA, &tr[1], &ks, nullptr);  // :205-222
            res_stop_on_error(pl.bar, ks);
        }
        __syncthreads();
        if (ks.status != ST_RUNNING) break;

        // ---- p = z + β p (NonUniform.jl:218-220): into LDS and HBM, the
        // rows on the tile's edge with device-scope stores, then published
        {
            const T beta = (T)ks.beta;
#pragma unroll
            for (int u = 0; u < kResSlots; ++u)
#pragma unroll
                for (int h = 0; h < 2; ++h) {
                    if (!rg.valid[u][h]) continue;
                    const unsigned e = opq(rg.sl[u].e);
                    const size_t c = (h ? HS : 0) + opq(rg.sl[u].base);
                    T *xh = h ? w.xs1 : w.xs0;
                    const T pn = xh[e] + beta * v.p[c];
                    xh[e] = pn;
                    if (rg.sl[u].edge) sc_store(v.p + c, pn);
                    else v.p[c] = pn;
                }
        }
        res_publish(pl.bar, fc + 1u);
        fc += 1u;
    }
    __syncthreads();
    if (blockIdx.x == 0 && tid == 0) *st = ks;
}

template <typename T>
size_t rb_res_lds(int TJ, int TK, int HX, int K)
{
    const size_t hsz = (size_t)(TJ + 2) * (TK + 2) * HX + 1;
    const size_t ntab = (size_t)6 * K * K + K;
    size_t n = (2 * hsz + (size_t)TK * ntab) * sizeof(T);
    n = (n + 15) & ~(size_t)15;
    // the sums' scratch, then the ring table
    return n + 2 * 68 * sizeof(dd) + (size_t)(2 * TJ + 2 * TK) * HX * sizeof(uint2);
}

// Tile plan: the tiling with the fewest slot rounds per thread (compute per
// pass), then the fewest ring elements (HBM reads per pass), then the
// fewest workgroups (barrier arrivals), within ncu workgroups, the LDS
// budget and 16-bit table indices.  G = 0 when none fits.
template <typename T>
RbRes rb_res_plan(const RbGeo<T> &g, int ncu)
{
    RbRes best{};
    if (g.K < 1 || ncu < 1) return best;
    const long ntab = 6L * g.K * g.K + g.K;
    long bc = -1, bh = 0, bg = 0;
    for (int TK = 1; TK <= g.NZ; ++TK)
        for (int TJ = 1; TJ <= g.NY; ++TJ) {
            const long rows = (long)TJ * TK;
            if (rows * g.HX > (long)kResThreads * kResSlots) break;
            const int ntj = (g.NY + TJ - 1) / TJ, ntk = (g.NZ + TK - 1) / TK;
            const long G = (long)ntj * ntk;
            // 16-bit table indices and ring-table LDS offsets
            const long hsz = (long)(TJ + 2) * (TK + 2) * g.HX + 1;
            if (G > ncu || G > kResMaxWG || TK * ntab >= 65536 || hsz >= 65536) continue;
            const size_t lds = rb_res_lds<T>(TJ, TK, g.HX, g.K);
            if (lds > kResLdsMax) continue;
            const long rounds = (rows * g.HX + kResThreads - 1) / kResThreads;
            const long halo = (2L * TJ + 2L * TK) * g.HX;
            if (bc < 0 || rounds < bc || (rounds == bc && (halo < bh || (halo == bh && G < bg)))) {
                bc = rounds;
                bh = halo;
                bg = G;
                best.TJ = TJ;
                best.TK = TK;
                best.ntj = ntj;
                best.ntk = ntk;
                best.G = (int)G;
                best.lds = lds;
            }
        }
    return best;
}

template <typename T, int MODE, bool FLUX>
static hipError_t res_attr()
{
    return hipFuncSetAttribute((const void *)k_rb_resident<T, MODE, FLUX>, hipFuncAttributeMaxDynamicSharedMemorySize,
                               (int)kResLdsMax);
}

template <typename T, int MODE>
static hipError_t res_attr_mode()
{
    hipError_t e = res_attr<T, MODE, false>();
    return e != hipSuccess ? e : res_attr<T, MODE, true>();
}

// Allow the resident kernels the LDS budget (above the default dynamic
// limit); called once per context, outside any stream capture.
template <typename T>
hipError_t rb_res_prepare()
{
    hipError_t e;
    if ((e = hipFuncSetAttribute((const void *)k_rb_bicg_solve<T, false>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                 (int)kResLdsMax)) != hipSuccess)
        return e;
    if ((e = hipFuncSetAttribute((const void *)k_rb_bicg_solve<T, true>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                 (int)kResLdsMax)) != hipSuccess)
        return e;
    if ((e = hipFuncSetAttribute((const void *)k_rb_cg_solve<T, false>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                 (int)kResLdsMax)) != hipSuccess)
        return e;
    if ((e = hipFuncSetAttribute((const void *)k_rb_cg_solve<T, true>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                 (int)kResLdsMax)) != hipSuccess)
        return e;
    if ((e = res_attr_mode<T, 0>()) != hipSuccess) return e;
    if ((e = res_attr_mode<T, 1>()) != hipSuccess) return e;
    if ((e = res_attr_mode<T, 2>()) != hipSuccess) return e;
    if ((e = res_attr_mode<T, 3>()) != hipSuccess) return e;
    return res_attr_mode<T, 4>();
}

template <typename T, int MODE>
static void res_launch(int mode, T *x, T *bb, const T *r, const T *q, const RbGeo<T> &g, const RbRes &pl,
                       const ResCoef<T> &cf, const KState *st, const RedBuf &rb, T *ax, const T *dy, const T *dz,
                       int dz_is_bb, hipStream_t s, const ResUpd<T> &up = ResUpd<T>{})
{
    const dim3 grid(pl.G), block(kResThreads);
    if (g.flux)
        hipLaunchKernelGGL((k_rb_resident<T, MODE, true>), grid, block, pl.lds, s, mode, x, bb, r, q, g, pl, cf, st, rb,
                           ax, dy, dz, dz_is_bb, up);
    else
        hipLaunchKernelGGL((k_rb_resident<T, MODE, false>), grid, block, pl.lds, s, mode, x, bb, r, q, g, pl, cf, st,
                           rb, ax, dy, dz, dz_is_bb, up);
}

template <typename T>
hipError_t launch_rb_resident(int mode, T *x, T *bb, const T *r, const T *q, const RbGeo<T> &g, const RbRes &pl,
                              const KState *st, const DotFin *dot, hipStream_t s, const ResAx<T> *axf)
{
    const ResCoef<T> cf{reinterpret_cast<T *>(pl.dgl), reinterpret_cast<T *>(pl.snk), pl.ix};
    if (!cf.dgl || !cf.ix || (g.flux && !cf.snk)) return hipErrorInvalidValue;
    if (pl.G <= 0 || !pl.bar || pl.lds > kResLdsMax || (dot && axf)) return hipErrorInvalidValue;
    const RedBuf *rbp = dot ? dot->rb : axf ? axf->rb : nullptr;
    if (rbp && pl.G > rbp->nblocks_max) return hipErrorInvalidValue;
    const RedBuf none{};
    const RedBuf &rb = rbp ? *rbp : none;
    if (dot) {
        res_launch<T, 1>(mode, x, bb, r, q, g, pl, cf, st, rb, nullptr, nullptr, nullptr, 0, s);
    } else if (axf) {
        const int zb = axf->dz != nullptr && axf->dz == bb;
        if (axf->dz) res_launch<T, 3>(mode, x, bb, r, q, g, pl, cf, st, rb, axf->ax, axf->dy, axf->dz, zb, s);
        else res_launch<T, 2>(mode, x, bb, r, q, g, pl, cf, st, rb, axf->ax, axf->dy, nullptr, 0, s);
    } else {
        res_launch<T, 0>(mode, x, bb, r, q, g, pl, cf, st, rb, nullptr, nullptr, nullptr, 0, s);
    }
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    if (dot)
        return launch_finalize_sums<T>(1, (unsigned)pl.G, dot->fin, dot->st, dot->hist, *dot->rb, s,
                                       seq_terms<T>(g, bb, x));
    if (axf && axf->fin >= 0)
        return launch_finalize_sums<T>(axf->dz ? 2 : 1, (unsigned)pl.G, axf->fin, axf->st, axf->hist, *axf->rb, s,
                                       seq_terms<T>(g, axf->ax, axf->dy, axf->ax, axf->dz));
    return hipSuccess;
}

template <typename T>
hipError_t launch_rb_resident_bicg_s(T *s_, T *sv, const T *r, const T *q, T *t, const T *p_, double *th,
                                     double *xbest, T *rw, const T *r0, const RbGeo<T> &g, const RbRes &pl, KState *st,
                                     double *hist, const RedBuf &rb, hipStream_t s)
{
    if (pl.G <= 0 || !pl.bar || !pl.wpart || pl.lds > kResLdsMax || pl.G > rb.nblocks_max || rb.seq)
        return hipErrorInvalidValue;
    const ResCoef<T> cf{reinterpret_cast<T *>(pl.dgl), reinterpret_cast<T *>(pl.snk), pl.ix};
    if (!cf.dgl || !cf.ix || (g.flux && !cf.snk)) return hipErrorInvalidValue;
    ResUpd<T> up;
    up.th = th;
    up.xbest = xbest;
    up.r = rw;
    up.p_ = p_;
    up.r0 = r0;
    up.wpart = pl.wpart;
    up.apart = rb.partials;
    res_launch<T, 4>(FIRST_BICG_S, s_, sv, r, q, g, pl, cf, st, rb, t, nullptr, sv, 1, s, up);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    return launch_finalize_sums<T>(2, (unsigned)pl.G, FIN_BICG_RES_RHO, st, hist, rb, s,
                                   seq_terms<T>(g, rw, nullptr, rw, r0));
}


template <typename T>
hipError_t launch_rb_bicg_solve(const ResBicg<T> &v, const RbGeo<T> &g, const RbRes &pl, KState *st, double *hist,
                                hipStream_t s)
{
    if (pl.G <= 0 || !pl.bar || !pl.wpart || pl.lds > kResLdsMax) return hipErrorInvalidValue;
    const ResCoef<T> cf{reinterpret_cast<T *>(pl.dgl), reinterpret_cast<T *>(pl.snk), pl.ix};
    if (!cf.dgl || !cf.ix || (g.flux && !cf.snk)) return hipErrorInvalidValue;
    if (!v.p || !v.p_ || !v.r || !v.q || !v.s_ || !v.t || !v.r0 || !v.th) return hipErrorInvalidValue;
    const dim3 grid(pl.G), block(kResThreads);
    if (g.flux)
        hipLaunchKernelGGL((k_rb_bicg_solve<T, true>), grid, block, pl.lds, s, g, pl, cf, st, hist, v);
    else
        hipLaunchKernelGGL((k_rb_bicg_solve<T, false>), grid, block, pl.lds, s, g, pl, cf, st, hist, v);
    return hipGetLastError();
}


template <typename T>
hipError_t launch_rb_cg_solve(const ResCg<T> &v, const RbGeo<T> &g, const RbRes &pl, KState *st, double *hist,
                              hipStream_t s)
{
    if (pl.G <= 0 || !pl.bar || !pl.wpart || pl.lds > kResLdsMax) return hipErrorInvalidValue;
    const ResCoef<T> cf{reinterpret_cast<T *>(pl.dgl), reinterpret_cast<T *>(pl.snk), pl.ix};
    if (!cf.dgl || !cf.ix || (g.flux && !cf.snk)) return hipErrorInvalidValue;
    if (!v.p || !v.r || !v.q || !v.z || !v.th) return hipErrorInvalidValue;
    const dim3 grid(pl.G), block(kResThreads);
    if (g.flux)
        hipLaunchKernelGGL((k_rb_cg_solve<T, true>), grid, block, pl.lds, s, g, pl, cf, st, hist, v);
    else
        hipLaunchKernelGGL((k_rb_cg_solve<T, false>), grid, block, pl.lds, s, g, pl, cf, st, hist, v);
    return hipGetLastError();
}

#define RES_INST(T)                                                                                              \
    template RbRes rb_res_plan<T>(const RbGeo<T> &, int);                                                       \
    template hipError_t rb_res_prepare<T>();                                                                    \
    template hipError_t launch_rb_res_coef<T>(const RbGeo<T> &, const ResCoef<T> &, hipStream_t);              \
    template hipError_t launch_rb_resident<T>(int, T *, T *, const T *, const T *, const RbGeo<T> &,          \
                                              const RbRes &, const KState *, const DotFin *, hipStream_t,     \
                                              const ResAx<T> *);                                              \
    template hipError_t launch_rb_resident_bicg_s<T>(T *, T *, const T *, const T *, T *, const T *, double *,   \
                                                     double *, T *, const T *, const RbGeo<T> &, const RbRes &,  \
                                                     KState *, double *, const RedBuf &, hipStream_t);      \
    template hipError_t launch_rb_bicg_solve<T>(const ResBicg<T> &, const RbGeo<T> &, const RbRes &, KState *,   \
                                                double *, hipStream_t);                                         \
    template hipError_t launch_rb_cg_solve<T>(const ResCg<T> &, const RbGeo<T> &, const RbRes &, KState *,       \
                                              double *, hipStream_t);

RES_INST(double)
RES_INST(float)
#undef RES_INST

}  // namespace h3d
