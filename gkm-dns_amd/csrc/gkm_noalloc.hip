// gkm_noalloc.hip — gkm_flux_points_noalloc: the BGK flux on sampled interface
// records (the kernel of gkm_flux_points) with no device allocation per call,
// for callers that cannot allocate (code under HIP graph capture, tight
// serving loops, harnesses that forbid hipMalloc).
//
// The records and the fluxes are staged through `scratch`, a static device
// array of fixed size that every GPU the library runs on holds until the
// process ends.  Host code never passes the array's name to a kernel or a
// copy (on the host it is a stand-in symbol, not a device address): the
// device address is looked up once per device with hipGetSymbolAddress, with
// that device current, on the entry point's first call there, and kept in
// scratch_ptr[].  A lock per device serialises the calls on it, so calls from
// several host threads on one GPU do not overlap in the array; each call
// overwrites the part of it that it reads back.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <mutex>

#include "gkm_internal.h"

namespace {

constexpr int kMaxDevices = 64;
// records per call: 52 input doubles and 5 output doubles each (3.7 MB in all)
constexpr int kMaxRecords = 8192;
constexpr size_t kInDoubles = (size_t)kMaxRecords * 52;
constexpr size_t kScratchDoubles = kInDoubles + (size_t)kMaxRecords * 5;

__device__ __align__(256) double scratch[kScratchDoubles];

std::mutex g_lock[kMaxDevices];
double* scratch_ptr[kMaxDevices] = {};

int fail(int code, const char* msg, hipError_t e) {
    char buf[256];
    snprintf(buf, sizeof buf, "%s: %s", msg, hipGetErrorString(e));
    return gkm::report_error(code, buf);
}

}  // namespace

extern "C" int gkm_flux_points_noalloc(const gkm_params* p, int n, const double* in, double dx, double* F) {
    if (!p || !in || !F || n < 0) return gkm::report_error(GKM_E_ARG, "null argument or negative count");
    if (n > kMaxRecords) {
        char buf[160];
        snprintf(buf, sizeof buf, "%d records exceed the %d that the static scratch array holds", n, kMaxRecords);
        return gkm::report_error(GKM_E_ARG, buf);
    }
    if (n == 0) return GKM_OK;
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return fail(GKM_E_HIP, "hipGetDevice", e);
    if (dev < 0 || dev >= kMaxDevices) return gkm::report_error(GKM_E_ARG, "device index beyond the supported 64");
    std::lock_guard<std::mutex> guard(g_lock[dev]);
    if (!scratch_ptr[dev]) {
        void* ptr = nullptr;
        e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return fail(GKM_E_HIP, "hipGetSymbolAddress(scratch)", e);
        scratch_ptr[dev] = (double*)ptr;
    }
    double* din = scratch_ptr[dev];
    double* dout = din + kInDoubles;
    // the null stream: no stream is created or destroyed per call
    const hipStream_t s = nullptr;
    e = hipMemcpyAsync(din, in, sizeof(double) * 52 * (size_t)n, hipMemcpyHostToDevice, s);
    if (e != hipSuccess) return fail(GKM_E_HIP, "copy of the records", e);
    gkm::launch_flux_points(s, p->precision, n, din, dx, p->gamma, 1.0 / p->Pr, p->K, dout);
    e = hipGetLastError();
    if (e == hipSuccess) e = hipMemcpyAsync(F, dout, sizeof(double) * 5 * (size_t)n, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) return fail(GKM_E_HIP, "flux points (no-alloc)", e);
    return GKM_OK;
}
