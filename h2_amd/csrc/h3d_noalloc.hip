// h3d_fdot2_noalloc: Fdot1/Fdot2 on caller-owned device arrays with no
// context and no device allocation (declared in include/h3d.h).
//
// The reduction's scratch memory (per-workgroup double-double partials, the
// pre-reduction area and the finalize state) is the file-scope device array
// `scratch` below, of a fixed size.  Each GPU holds its own copy for the
// life of the process; its device address is looked up once per device
// with hipGetSymbolAddress (the host-side name of the array is not a device
// address) and kept in g_scratch_ptr.  Every call on a GPU shares that
// GPU's copy, so calls that overlap on one GPU (two streams, two host
// threads) race; calls on one stream are ordered.

#include "h3d.h"
#include "h3d_internal.hpp"

#include <cstdint>
#include <mutex>

using namespace h3d;

namespace {

// Largest launch served: kNoallocMaxBlocks workgroups of the dot kernel,
// i.e. ceil(NX/64)·ceil(NY/4) plane columns at most 16384 (for example
// 960 x 960 interior cells: 3600).
constexpr unsigned kNoallocMaxBlocks = 16384;
constexpr size_t kPartialWords = (size_t)kNoallocMaxBlocks * 6;             // [nb][3][2]
constexpr size_t kPrereduceWords = ((size_t)kNoallocMaxBlocks / 1024 + 2) * 6;
constexpr size_t kStateWords = (sizeof(KState) + 7) / 8;
constexpr size_t kScratchWords = kPartialWords + kPrereduceWords + kStateWords;
static_assert(sizeof(KState) % 8 == 0, "the state is cleared as whole doubles");

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mutex;
double *g_scratch_ptr[kMaxDevices] = {};

}  // namespace

// 787 KB per GPU: partials, pre-reduction area, finalize state
static __device__ __align__(256) double scratch[kScratchWords];

namespace {

// Device address of this GPU's scratch, looked up on the first call on the
// device (with that device current).
int scratch_for(int device, double **out)
{
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (!g_scratch_ptr[device]) {
        void *p = nullptr;
        if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return H3D_ERR_HIP;
        g_scratch_ptr[device] = static_cast<double *>(p);
    }
    *out = g_scratch_ptr[device];
    return H3D_OK;
}

// copies the finalize result out[0] of the state into the caller's double
__global__ void k_result_out(const KState *st, double *out) { *out = st->out[0]; }

}  // namespace

extern "C" int h3d_fdot2_noalloc(const double *x, const double *y, int MX, int MY, int MZ, double *out,
                                 void *stream)
{
    if (!x || !out || MX < 3 || MY < 3 || MZ < 3) return H3D_ERR_ARG;
    const int NX = MX - 2, NY = MY - 2, NZ = MZ - 2;
    if (interior_launch_blocks(NX, NY, NZ) > kNoallocMaxBlocks) return H3D_ERR_UNSUPPORTED;
    int device = 0;
    if (hipGetDevice(&device) != hipSuccess || device < 0 || device >= kMaxDevices) return H3D_ERR_HIP;
    double *base = nullptr;
    int rc = scratch_for(device, &base);
    if (rc != H3D_OK) return rc;
    hipStream_t s = static_cast<hipStream_t>(stream);
    RedBuf rb{};
    rb.partials = base;
    rb.partials2 = base + kPartialWords;
    rb.nblocks_max = (int)kNoallocMaxBlocks;
    rb.seq = 0;
    KState *st = reinterpret_cast<KState *>(base + kPartialWords + kPrereduceWords);
    Geo<double> g{};
    g.MX = MX; g.MY = MY; g.MZ = MZ;
    g.NX = NX; g.NY = NY; g.NZ = NZ;
    // the state's status must read "running" (0) for the reduction to run:
    // clear it with a fill kernel (no memset node, so the call can be captured)
    if (launch_fill<double>(reinterpret_cast<double *>(st), 0.0, kStateWords, s) != hipSuccess) return H3D_ERR_HIP;
    if (launch_dot<double>(x, y, nullptr, g, FIN_STORE, st, nullptr, rb, s) != hipSuccess) return H3D_ERR_HIP;
    hipLaunchKernelGGL(k_result_out, dim3(1), dim3(1), 0, s, st, out);
    return hipGetLastError() == hipSuccess ? H3D_OK : H3D_ERR_HIP;
}
