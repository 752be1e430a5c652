// nbg_noalloc.hip -- the direct sum for callers that cannot allocate device memory (code
// under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
//
// nbg_accel_direct_noalloc is the direct-sum evaluation of nbg_accel (NBG_SOLVER_DIRECT)
// on caller-owned device arrays, with the same kernels.  Its only device memory beyond
// the caller's inputs and outputs is the staging of the packed source records
// (x, y, z, G m), which lives in `scratch` below: a file-scope device array of
// NBG_NOALLOC_MAX_N records (16 MiB) that each GPU the library runs on holds until the
// process ends.  The entry point calls no hipMalloc, hipFree or allocating routine.
//
// Every call on one device shares that device's array: calls that overlap on one device
// (two host threads, or two streams whose work overlaps) race on it.  A call overwrites
// the first n records before its force kernel reads them, so nothing needs clearing.
#include <mutex>

#include "nbg_internal.hpp"

static __device__ __align__(256) float4 scratch[NBG_NOALLOC_MAX_N];

namespace {

constexpr int kMaxDevices = 64;
// Device address of `scratch` on each device, looked up once per device with that device
// current.  In host code the name `scratch` is a host-side stand-in, not a device address:
// only these addresses are passed to kernels.
float4* g_scratch_dev[kMaxDevices] = {};
std::mutex g_scratch_mu;

// Called with `device` current (the first call on a device looks the address up).
int scratch_on(int device, float4** out) {
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_dev[device]) {
        void* ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !ptr) return NBG_E_HIP;
        g_scratch_dev[device] = static_cast<float4*>(ptr);
    }
    *out = g_scratch_dev[device];
    return NBG_OK;
}

}  // namespace

extern "C" {

int nbg_accel_direct_noalloc(int device, void* stream, const nbg_params* p, int64_t n, const float* x,
                             const float* y, const float* z, const float* m, float* ax, float* ay, float* az) {
    if (!p || !x || !y || !z || !m || !ax || !ay || !az) return NBG_E_INVALID;
    if (n < 1 || n > NBG_NOALLOC_MAX_N) return NBG_E_INVALID;
    if (p->solver != NBG_SOLVER_DIRECT || !(p->eps2 >= 0.0) || !(p->G == p->G) ||
        (p->softening != NBG_SOFT_PLUMMER && p->softening != NBG_SOFT_SPLINE) ||
        (p->softening == NBG_SOFT_SPLINE && !(p->eps2 > 0.0)))
        return NBG_E_INVALID;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess) return NBG_E_HIP;
    if (device < 0 || device >= ndev || device >= kMaxDevices) return NBG_E_INVALID;
    if (hipSetDevice(device) != hipSuccess) return NBG_E_HIP;
    float4* stage = nullptr;
    if (int r = scratch_on(device, &stage)) return r;
    (void)hipGetLastError();  // report this call's launches only
    nbg::launch_direct_raw(static_cast<hipStream_t>(stream), *p, static_cast<int>(n), x, y, z, m, stage, ax, ay,
                           az);
    return hipGetLastError() == hipSuccess ? NBG_OK : NBG_E_HIP;
}

}  // extern "C"
