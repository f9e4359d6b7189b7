// nuosc_noalloc.hip -- nuosc_p3_slice_noalloc, the variant of nuosc_p3_slice
// that allocates nothing.
//
// nuosc_p3_slice stages the slice's index lists and its output in the
// handle's staging buffer, which it (re)allocates with hipMalloc when a
// slice is larger than any before.  Callers that may not allocate device
// memory at that point (inside a serving loop, between graph replays, or in
// harnesses that forbid it) use this variant instead: the index lists and the
// output are staged in `scratch`, a fixed array in device memory that every
// GPU the library runs on holds once for the whole process.
//
// Limits: the index lists hold at most kIdxBytes / 4 = 16384 values
// (sx + sv), the output at most (kScratchBytes - kIdxBytes) / 8 = 2,088,960
// values (sx * nz * sv); larger slices are refused with NUOSC_EINVAL.
// Nothing in the array must start at zero: every call writes the indices and
// the output values it reads back.  The array is shared by every call on its
// GPU; calls of this process on one GPU take a per-device lock for their
// whole (synchronous) duration, so they do not overlap.  Anything else that
// used the same array concurrently would race with them.
#include <hip/hip_runtime.h>

#include <mutex>

#include "../../include/nuosc_hip.h"
#include "nuosc_dev.h"

namespace {

constexpr size_t kScratchBytes = size_t(16) << 20;   // per GPU, held until the process ends
constexpr size_t kIdxBytes = size_t(64) << 10;       // index region (x indices, then v indices)
constexpr int kMaxDevices = 64;

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

std::mutex g_init_mu;
void* g_scratch_ptr[kMaxDevices] = {};   // device address of `scratch` on each GPU
std::mutex g_call_mu[kMaxDevices];

// Device address of this GPU's copy of `scratch`, looked up once per device
// with that device current.  (In host code the name `scratch` is only a
// stand-in for the symbol: its device address must come from
// hipGetSymbolAddress, and that address is what kernels and copies take.)
hipError_t scratch_address(int dev, void** out) {
    std::lock_guard<std::mutex> g(g_init_mu);
    if (!g_scratch_ptr[dev]) {
        int cur = -1;
        hipError_t e = hipGetDevice(&cur);
        if (e != hipSuccess) return e;
        if (cur != dev && (e = hipSetDevice(dev)) != hipSuccess) return e;
        void* p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (cur != dev) (void)hipSetDevice(cur);
        if (e != hipSuccess) return e;
        g_scratch_ptr[dev] = p;
    }
    *out = g_scratch_ptr[dev];
    return hipSuccess;
}

}  // namespace

extern "C" int nuosc_p3_slice_noalloc(nuosc_sim* s, const int* xidx, int sx, const int* vidx, int sv,
                                      double* dst) {
    using namespace nuosc;
    if (!s || sx < 0 || sv < 0 || (sx && !xidx) || (sv && !vidx) || !dst) return NUOSC_EINVAL;
    const SimView v = sim_view(s);
    for (int a = 0; a < sx; ++a)
        if (xidx[a] < 0 || xidx[a] >= v.geo.nx) return sim_fail(s, NUOSC_EINVAL, "x index out of range");
    for (int b = 0; b < sv; ++b)
        if (vidx[b] < 0 || vidx[b] >= v.geo.nv) return sim_fail(s, NUOSC_EINVAL, "v index out of range");
    const size_t n = (size_t)sx * v.geo.nz * sv;
    if (n == 0) return NUOSC_OK;
    if (((size_t)sx + (size_t)sv) * sizeof(int) > kIdxBytes || n * sizeof(double) > kScratchBytes - kIdxBytes)
        return sim_fail(s, NUOSC_EINVAL,
                        "slice too large for nuosc_p3_slice_noalloc: at most 16384 indices (sx + sv) and "
                        "2088960 values (sx * nz * sv); use nuosc_p3_slice");
    if (v.device < 0 || v.device >= kMaxDevices) return sim_fail(s, NUOSC_EINVAL, "device ordinal out of range");
    std::lock_guard<std::mutex> call(g_call_mu[v.device]);
    void* base = nullptr;
    hipError_t e = scratch_address(v.device, &base);
    if (e != hipSuccess)
        return sim_fail(s, NUOSC_EHIP, "hipGetSymbolAddress of the scratch array failed");
    int* di = static_cast<int*>(base);
    double* dd = reinterpret_cast<double*>(static_cast<unsigned char*>(base) + kIdxBytes);
    if (sx) e = hipMemcpyAsync(di, xidx, sizeof(int) * sx, hipMemcpyHostToDevice, v.stream);
    if (e == hipSuccess && sv) e = hipMemcpyAsync(di + sx, vidx, sizeof(int) * sv, hipMemcpyHostToDevice, v.stream);
    if (e == hipSuccess)
        e = v.esize == 8 ? launch_p3_slice<double>(v.geo, static_cast<const double*>(v.state), v.g0, di, sx, di + sx,
                                                   sv, dd, v.stream)
                         : launch_p3_slice<float>(v.geo, static_cast<const float*>(v.state), v.g0, di, sx, di + sx,
                                                  sv, dd, v.stream);
    if (e == hipSuccess) e = hipMemcpyAsync(dst, dd, sizeof(double) * n, hipMemcpyDeviceToHost, v.stream);
    if (e == hipSuccess) e = hipStreamSynchronize(v.stream);
    if (e != hipSuccess)
        return sim_fail(s, NUOSC_EHIP, "p3 slice (no allocation): a HIP call failed");
    return NUOSC_OK;
}
