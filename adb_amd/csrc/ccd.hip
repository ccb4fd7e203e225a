// ccd.hip - batched collision-time and distance kernels of the kernel-level API.
//
// pt_collision_time / ee_collision_time (model/collision.cpp:299-348, exposed by
// python/bindings.cpp:81-82) and the distance type + squared distance pair
// (model/geometry.cpp:101-112, model/abd.cpp:353-354), one entry per lane.
#include <mutex>
#include "dmath.h"
#include "world.h"

namespace adb {

template <int IS_EE>
__global__ void k_toi_api(int n, const double* __restrict__ x, double* __restrict__ toi) {
    int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n) return;
    const double* p = x + 24 * (size_t)e;
    d3 P0[4], P1[4];
    for (int k = 0; k < 4; k++) { P0[k] = ld3(p + 3 * k); P1[k] = ld3(p + 12 + 3 * k); }
    toi[e] = IS_EE ? ee_toi(P0, P1) : pt_toi(P0, P1);
}
template <int IS_EE>
__global__ void k_dist_api(int n, const double* __restrict__ x, int* __restrict__ type, double* __restrict__ dist) {
    int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n) return;
    const double* p = x + 12 * (size_t)e;
    d3 a = ld3(p), b = ld3(p + 3), c = ld3(p + 6), d = ld3(p + 9);
    int t = IS_EE ? ee_type(a, b, c, d) : pt_type(a, b, c, d);
    type[e] = t;
    dist[e] = IS_EE ? ee_dist(a, b, c, d, t) : pt_dist(a, b, c, d, t);
}

} // namespace adb

using namespace adb;

extern "C" int adb_pt_collision_time_batched(int n, const double* x, double* toi, void* stream) {
    if (n < 0) return ADB_E_ARG;
    if (n == 0) return ADB_OK;
    hipLaunchKernelGGL(k_toi_api<0>, dim3((n + 127) / 128), dim3(128), 0, (hipStream_t)stream, n, x, toi);
    return hipGetLastError() == hipSuccess ? ADB_OK : ADB_E_HIP;
}
extern "C" int adb_ee_collision_time_batched(int n, const double* x, double* toi, void* stream) {
    if (n < 0) return ADB_E_ARG;
    if (n == 0) return ADB_OK;
    hipLaunchKernelGGL(k_toi_api<1>, dim3((n + 127) / 128), dim3(128), 0, (hipStream_t)stream, n, x, toi);
    return hipGetLastError() == hipSuccess ? ADB_OK : ADB_E_HIP;
}
extern "C" int adb_pt_distance_batched(int n, const double* x, int* type, double* dist, void* stream) {
    if (n < 0) return ADB_E_ARG;
    if (n == 0) return ADB_OK;
    hipLaunchKernelGGL(k_dist_api<0>, dim3((n + 127) / 128), dim3(128), 0, (hipStream_t)stream, n, x, type, dist);
    return hipGetLastError() == hipSuccess ? ADB_OK : ADB_E_HIP;
}
extern "C" int adb_ee_distance_batched(int n, const double* x, int* type, double* dist, void* stream) {
    if (n < 0) return ADB_E_ARG;
    if (n == 0) return ADB_OK;
    hipLaunchKernelGGL(k_dist_api<1>, dim3((n + 127) / 128), dim3(128), 0, (hipStream_t)stream, n, x, type, dist);
    return hipGetLastError() == hipSuccess ? ADB_OK : ADB_E_HIP;
}

static double single_toi(int is_ee, const double x[24]) {
    double *dx = nullptr, *dt = nullptr;
    double out = -1.0;
    if (hipMalloc((void**)&dx, 24 * sizeof(double)) == hipSuccess && hipMalloc((void**)&dt, sizeof(double)) == hipSuccess) {
        hipMemcpy(dx, x, 24 * sizeof(double), hipMemcpyHostToDevice);
        int rc = is_ee ? adb_ee_collision_time_batched(1, dx, dt, nullptr) : adb_pt_collision_time_batched(1, dx, dt, nullptr);
        if (rc == ADB_OK && hipDeviceSynchronize() == hipSuccess) hipMemcpy(&out, dt, sizeof(double), hipMemcpyDeviceToHost);
    }
    hipFree(dx); hipFree(dt);
    return out;
}
extern "C" double adb_pt_collision_time(const double x[24]) { return single_toi(0, x); }
extern "C" double adb_ee_collision_time(const double x[24]) { return single_toi(1, x); }

// ---------------------------------------------------------------------------
// No-allocation variant of adb_pt/ee_collision_time for callers that cannot
// call hipMalloc.  The device staging of the inputs (24 doubles per entry)
// and outputs (1 double per entry) lives in a file-scope device array of
// NOALLOC_MAX entries (800 KB on each device that runs it, kept until the
// process ends).  Every entry of the staging area that a call reads is
// written by that call first, so nothing needs clearing between calls.
// ---------------------------------------------------------------------------
namespace adb {
constexpr int NOALLOC_MAX = 4096;
static __device__ __align__(256) double scratch[NOALLOC_MAX * 25];
constexpr int NOALLOC_MAX_DEVICES = 64;
static std::mutex g_noalloc_mu;                       // one call at a time in this process
static double* g_scratch_dev[NOALLOC_MAX_DEVICES];    // device address of `scratch`, per device
}

extern "C" int adb_collision_time_noalloc(int is_ee, int n, const double* x, double* toi) {
    if (n < 0 || (n > 0 && (!x || !toi)) || (is_ee != 0 && is_ee != 1)) return ADB_E_ARG;
    if (n > NOALLOC_MAX) return ADB_E_CAPACITY;
    if (n == 0) return ADB_OK;
    std::lock_guard<std::mutex> lock(g_noalloc_mu);
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return ADB_E_HIP;
    if (dev < 0 || dev >= NOALLOC_MAX_DEVICES) return ADB_E_NO_DEVICE;
    if (!g_scratch_dev[dev]) {
        // first call on this device, with it current: the array's address on it
        void* p = nullptr;
        if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return ADB_E_HIP;
        g_scratch_dev[dev] = static_cast<double*>(p);
    }
    double* dx = g_scratch_dev[dev];
    double* dt = dx + 24 * (size_t)NOALLOC_MAX;
    if (hipMemcpy(dx, x, 24 * sizeof(double) * (size_t)n, hipMemcpyHostToDevice) != hipSuccess) return ADB_E_HIP;
    if (is_ee) hipLaunchKernelGGL(k_toi_api<1>, dim3((n + 127) / 128), dim3(128), 0, nullptr, n, dx, dt);
    else hipLaunchKernelGGL(k_toi_api<0>, dim3((n + 127) / 128), dim3(128), 0, nullptr, n, dx, dt);
    if (hipGetLastError() != hipSuccess) return ADB_E_HIP;
    // the copy on the null stream waits for the kernel
    if (hipMemcpy(toi, dt, sizeof(double) * (size_t)n, hipMemcpyDeviceToHost) != hipSuccess) return ADB_E_HIP;
    return ADB_OK;
}
