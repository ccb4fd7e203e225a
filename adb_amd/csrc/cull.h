// cull.h - device helpers shared by the two culling methods (iAABB body
// overlaps in cull.hip, spatial hashing in hash.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <cstring>

namespace adb {

// Monotone map of doubles to unsigned 64-bit keys (negative values included),
// so an integer atomicMin is an exact double min.
__device__ __forceinline__ unsigned long long order_key(double t) {
    unsigned long long b = (unsigned long long)__double_as_longlong(t);
    return (b >> 63) ? ~b : (b | 0x8000000000000000ull);
}
inline double order_key_inv(unsigned long long k) {
    unsigned long long b = (k >> 63) ? (k & 0x7fffffffffffffffull) : ~k;
    double t;
    std::memcpy(&t, &b, 8);
    return t;
}
inline unsigned long long order_key_host(double t) {
    unsigned long long b;
    std::memcpy(&b, &t, 8);
    return (b >> 63) ? ~b : (b | 0x8000000000000000ull);
}

__device__ __forceinline__ bool ovl(const double* alo, const double* ahi, const double* blo, const double* bhi) {
    return alo[0] <= bhi[0] && ahi[0] >= blo[0] && alo[1] <= bhi[1] && ahi[1] >= blo[1] && alo[2] <= bhi[2] && ahi[2] >= blo[2];
}

// Append `key` for the lanes with `hit`: one atomic per wavefront, lanes in
// order.  Entries past `cap` are counted but not written (the caller grows
// the buffer and repeats the pass).
__device__ __forceinline__ void append_key(bool hit, unsigned long long key, unsigned long long* out, int cap, int* counter) {
    unsigned long long m = __ballot(hit);
    if (m == 0) return;
    const int lane = __lane_id();
    const int leader = __ffsll((long long)m) - 1;
    int base = 0;
    if (lane == leader) base = atomicAdd(counter, __popcll(m));
    base = __shfl(base, leader);
    if (hit) {
        int off = __popcll(m & ((1ull << lane) - 1));
        if (base + off < cap) out[base + off] = key;
    }
}

// Contact-set keys are appended to NBK buckets (one counter and one region
// of `sub` keys each; a workgroup uses bucket blockIdx.x % NBK): device-scope
// atomics on one address serialize (~11 ns each on gfx950), on 64 addresses
// they overlap.  The buckets are packed before the sort (cull.hip).
constexpr int NBK = 64;
struct BucketList { int cnt[NBK]; int off[NBK]; };

} // namespace adb
