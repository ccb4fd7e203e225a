// cull.hip - broad phase, narrow-phase contact sets and the CCD step bound.
//
// Replaces IAABB::iaabb_brute_force (model/iaabb.cpp:23-568) and its pair
// tests (model/thread_local_culling.cpp:28-386): vtn = 1, 2 contact sets and
// the vtn = 3 trajectory step bound.
//
// Pipeline per pass (all on the device, one host read of the counters):
//   1. per-body AABB of its vertices (contact sets: inflated by r >= sqrt(d_hat);
//      CCD: union of the start and end positions), one workgroup per body;
//   2. body pairs whose boxes overlap (tiled all-pairs test, wave-aggregated
//      compaction);
//   3. six workgroups per body pair: primitives of both bodies that touch the
//      pair's overlap box are compacted into LDS lists, then every list pair is
//      tested (conservative box prefilter, exact FP64 predicate); hits gather
//      in an LDS buffer that goes out with one atomic on one of 64 bucket
//      counters per flush;
//   4. vertex-ground candidates by a vertex-parallel scan;
//   5. the keys are grouped by their high bits (counting sort), each key is
//      placed by its rank inside its group (the canonical sorted order), then
//      decoded.
// The box prefilters are conservative, so the sets equal the exact predicate
// over all primitive pairs (d_PT < d_hat for I != J, d_EE < d_hat for I < J).
#include <hipcub/hipcub.hpp>
#include <algorithm>
#include <vector>
#include <cstdio>
#include "dmath.h"
#include "world.h"
#include "cull.h"

namespace adb {

constexpr int NB_THREADS = 256;


// ------------------------------------------------------------------ transforms
__global__ void k_transform(int V, const int* __restrict__ vbody, const double* __restrict__ xbar,
                            const double* __restrict__ q, double* __restrict__ X) {
    int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= V) return;
    const double* Q = q + 12 * (size_t)vbody[v];
    double x0 = xbar[3 * (size_t)v], x1 = xbar[3 * (size_t)v + 1], x2 = xbar[3 * (size_t)v + 2];
#pragma unroll
    for (int i = 0; i < 3; i++) X[3 * (size_t)v + i] = ((Q[3 + i] * x0 + Q[6 + i] * x1) + Q[9 + i] * x2) + Q[i];
}
__global__ void k_axpy(int n, const double* __restrict__ q, const double* __restrict__ dq, double alpha, double* __restrict__ out) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = q[i] + (dq ? dq[i] * alpha : 0.0);
}

// q' = q + alpha dq (dq null: q' = q) written to qout (thread i < 12 B writes
// entry i, as k_axpy computes it), and the vertex positions at q' to X (and
// X2 if given).  A workgroup's 256 vertices belong to consecutive bodies: when
// there are at most TQ_BODIES of them, their q' is formed once into LDS,
// else every vertex thread forms its body's q'.
constexpr int TQ_BODIES = 32;
__global__ __launch_bounds__(256) void k_transform_q(int V, int n12, const int* __restrict__ vbody, const double* __restrict__ xbar,
                                                     const double* __restrict__ q, const double* __restrict__ dq, double alpha,
                                                     double* __restrict__ qout, double* __restrict__ X, double* __restrict__ X2) {
    __shared__ double sq[12 * TQ_BODIES];
    const int v0 = blockIdx.x * 256, v = v0 + threadIdx.x;
    if (v < n12 && qout) qout[v] = dq ? q[v] + dq[v] * alpha : q[v];
    if (v0 >= V) return;
    const int b0 = vbody[v0], nb = vbody[min(v0 + 255, V - 1)] - b0 + 1;
    const bool staged = nb <= TQ_BODIES;
    if (staged) {
        for (int t = threadIdx.x; t < 12 * nb; t += 256) {
            const size_t k = 12 * (size_t)b0 + t;
            sq[t] = dq ? q[k] + dq[k] * alpha : q[k];
        }
        __syncthreads();
    }
    if (v >= V) return;
    const int bv = vbody[v];
    const size_t b = 12 * (size_t)bv;
    double Q[12];
    if (staged) {
#pragma unroll
        for (int k = 0; k < 12; k++) Q[k] = sq[12 * (bv - b0) + k];
    } else {
#pragma unroll
        for (int k = 0; k < 12; k++) Q[k] = dq ? q[b + k] + dq[b + k] * alpha : q[b + k];
    }
    const double x0 = xbar[3 * (size_t)v], x1 = xbar[3 * (size_t)v + 1], x2 = xbar[3 * (size_t)v + 2];
#pragma unroll
    for (int i = 0; i < 3; i++) {
        const double x = ((Q[3 + i] * x0 + Q[6 + i] * x1) + Q[9 + i] * x2) + Q[i];
        X[3 * (size_t)v + i] = x;
        if (X2) X2[3 * (size_t)v + i] = x;
    }
}
void transform_step(World& w, const double* q, const double* dq, double alpha, double* qout, double* X, double* X2) {
    ProfScope ps(w, PK_TRANSFORM, (X2 ? 76.0 : 52.0) * w.V + (dq ? 288.0 : 192.0) * w.B);
    const int n = std::max(w.V, 12 * w.B);
    hipLaunchKernelGGL(k_transform_q, dim3((n + 255) / 256), dim3(256), 0, w.stream, w.V, 12 * w.B, w.vbody.p, w.xbar.p, q, dq, alpha,
                       qout, X, X2);
    ADB_HIP(hipGetLastError());
}
void transform_positions(World& w, const double* q, double* X) {
    ProfScope ps(w, PK_TRANSFORM, 52.0 * w.V + 96.0 * w.B);   // xbar + vbody in, X out, q
    hipLaunchKernelGGL(k_transform, dim3((w.V + 255) / 256), dim3(256), 0, w.stream, w.V, w.vbody.p, w.xbar.p, q, X);
    ADB_HIP(hipGetLastError());
}
void q_axpy(World& w, const double* q, const double* dq, double alpha, double* out) {
    int n = 12 * w.B;
    hipLaunchKernelGGL(k_axpy, dim3((n + 255) / 256), dim3(256), 0, w.stream, n, q, dq, alpha, out);
    ADB_HIP(hipGetLastError());
}
// per-scene factors: out = q + f[scene] dq; v *= f[scene]
__global__ void k_axpy_scenes(int n, const int* __restrict__ scene_of, const double* __restrict__ q,
                              const double* __restrict__ dq, const double* __restrict__ f, double* __restrict__ out) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = q[i] + dq[i] * f[scene_of[i / 12]];
}
__global__ void k_scale_scenes(int n, const int* __restrict__ scene_of, double* __restrict__ v, const double* __restrict__ f) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) v[i] = v[i] * f[scene_of[i / 12]];
}
void q_axpy_scenes(World& w, const double* q, const double* dq, const double* f, double* out) {
    int n = 12 * w.B;
    hipLaunchKernelGGL(k_axpy_scenes, dim3((n + 255) / 256), dim3(256), 0, w.stream, n, w.scene_of.p, q, dq, f, out);
    ADB_HIP(hipGetLastError());
}
void scale_scenes(World& w, double* v, const double* f) {
    int n = 12 * w.B;
    hipLaunchKernelGGL(k_scale_scenes, dim3((n + 255) / 256), dim3(256), 0, w.stream, n, w.scene_of.p, v, f);
    ADB_HIP(hipGetLastError());
}

// ------------------------------------------------------------------ body boxes
// one workgroup per body; X1 != null: union of both position sets
// Body AABB (inflated by `infl`) and bounding sphere about the box centre,
// both from the vertex positions at X0 (and X1: the swept set, as the
// segments x0 -> x1 stay inside the ball of their end points).  Workgroup 0
// also clears the pass's counters zero[0, nzero) and, with toi, sets the two
// step bounds toi[0, 2) to toi_one (the kernels after it in the stream
// count and take minima there).
__global__ void k_body_box(const int* __restrict__ voff, const double* __restrict__ X0, const double* __restrict__ X1,
                           double infl, double* __restrict__ box, double* __restrict__ sph, int* __restrict__ zero, int nzero,
                           unsigned long long* __restrict__ toi, unsigned long long toi_one) {
    const int b = blockIdx.x;
    if (b == 0) {
        for (int i = threadIdx.x; i < nzero; i += blockDim.x) zero[i] = 0;
        if (toi && threadIdx.x < 2) toi[threadIdx.x] = toi_one;
    }
    double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
    for (int v = voff[b] + threadIdx.x; v < voff[b + 1]; v += blockDim.x) {
        for (int k = 0; k < 3; k++) {
            double x = X0[3 * (size_t)v + k];
            lo[k] = fmin(lo[k], x); hi[k] = fmax(hi[k], x);
            if (X1) { double y = X1[3 * (size_t)v + k]; lo[k] = fmin(lo[k], y); hi[k] = fmax(hi[k], y); }
        }
    }
    __shared__ double s[6][NB_THREADS / 64];
    for (int k = 0; k < 3; k++) {
        for (int m = 32; m >= 1; m >>= 1) { lo[k] = fmin(lo[k], __shfl_xor(lo[k], m)); hi[k] = fmax(hi[k], __shfl_xor(hi[k], m)); }
    }
    const int wv = threadIdx.x / 64, ln = threadIdx.x % 64;
    if (ln == 0) for (int k = 0; k < 3; k++) { s[k][wv] = lo[k]; s[3 + k][wv] = hi[k]; }
    __syncthreads();
    __shared__ double cen[3];
    if (threadIdx.x == 0) {
        for (int k = 0; k < 3; k++) {
            double l = s[k][0], h = s[3 + k][0];
            for (int i = 1; i < (int)(blockDim.x / 64); i++) { l = fmin(l, s[k][i]); h = fmax(h, s[3 + k][i]); }
            box[6 * (size_t)b + k] = l - infl;
            box[6 * (size_t)b + 3 + k] = h + infl;
            cen[k] = 0.5 * (l + h);
        }
    }
    __syncthreads();
    double r2 = 0.0;
    for (int v = voff[b] + threadIdx.x; v < voff[b + 1]; v += blockDim.x) {
        const double dx = X0[3 * (size_t)v] - cen[0], dy = X0[3 * (size_t)v + 1] - cen[1], dz = X0[3 * (size_t)v + 2] - cen[2];
        r2 = fmax(r2, dx * dx + dy * dy + dz * dz);
        if (X1) {
            const double ex = X1[3 * (size_t)v] - cen[0], ey = X1[3 * (size_t)v + 1] - cen[1], ez = X1[3 * (size_t)v + 2] - cen[2];
            r2 = fmax(r2, ex * ex + ey * ey + ez * ez);
        }
    }
    for (int m = 32; m >= 1; m >>= 1) r2 = fmax(r2, __shfl_xor(r2, m));
    __syncthreads();
    if (ln == 0) s[0][wv] = r2;
    __syncthreads();
    if (threadIdx.x == 0) {
        double rr = s[0][0];
        for (int i = 1; i < (int)(blockDim.x / 64); i++) rr = fmax(rr, s[0][i]);
        sph[4 * (size_t)b] = cen[0]; sph[4 * (size_t)b + 1] = cen[1]; sph[4 * (size_t)b + 2] = cen[2];
        sph[4 * (size_t)b + 3] = sqrt(rr) * (1.0 + 1e-12) + 1e-300;
    }
}

// The same boxes and spheres with one wavefront per body (four bodies per
// workgroup), for bodies of at most 64 * BOX_VC vertices: each lane keeps its
// positions in registers between the box pass and the sphere pass, and both
// reductions stay inside the wavefront (no LDS, no workgroup barrier).  The
// results equal k_body_box's bit for bit: the minima, maxima and squared
// distances come from the same expressions, and a minimum or maximum does not
// depend on the order it is taken in.
constexpr int BOX_VC = 16;
template <bool SWEPT>
__global__ __launch_bounds__(256) void k_body_box_w(int B, const int* __restrict__ voff, const double* __restrict__ X0,
                                                    const double* __restrict__ X1, double infl, double* __restrict__ box,
                                                    double* __restrict__ sph, int* __restrict__ zero, int nzero,
                                                    unsigned long long* __restrict__ toi, unsigned long long toi_one) {
    if (blockIdx.x == 0) {
        for (int i = threadIdx.x; i < nzero; i += blockDim.x) zero[i] = 0;
        if (toi && threadIdx.x < 2) toi[threadIdx.x] = toi_one;
    }
    const int b = blockIdx.x * 4 + threadIdx.x / 64, ln = threadIdx.x % 64;
    if (b >= B) return;
    const int v0 = voff[b], nv = voff[b + 1] - v0;
    double x[BOX_VC][3], y[SWEPT ? BOX_VC : 1][3];
    double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
#pragma unroll
    for (int i = 0; i < BOX_VC; i++) {
        const int t = ln + 64 * i;
        if (t < nv) {
            const size_t o = 3 * (size_t)(v0 + t);
#pragma unroll
            for (int k = 0; k < 3; k++) {
                x[i][k] = X0[o + k];
                lo[k] = fmin(lo[k], x[i][k]); hi[k] = fmax(hi[k], x[i][k]);
                if (SWEPT) {
                    y[i][k] = X1[o + k];
                    lo[k] = fmin(lo[k], y[i][k]); hi[k] = fmax(hi[k], y[i][k]);
                }
            }
        }
    }
    double cen[3];
#pragma unroll
    for (int k = 0; k < 3; k++) {
        for (int m = 32; m >= 1; m >>= 1) { lo[k] = fmin(lo[k], __shfl_xor(lo[k], m)); hi[k] = fmax(hi[k], __shfl_xor(hi[k], m)); }
        cen[k] = 0.5 * (lo[k] + hi[k]);
    }
    if (ln < 3) {
        const double l = ln == 0 ? lo[0] : ln == 1 ? lo[1] : lo[2], h = ln == 0 ? hi[0] : ln == 1 ? hi[1] : hi[2];
        box[6 * (size_t)b + ln] = l - infl;
        box[6 * (size_t)b + 3 + ln] = h + infl;
    }
    double r2 = 0.0;
#pragma unroll
    for (int i = 0; i < BOX_VC; i++) {
        if (ln + 64 * i < nv) {
            const double dx = x[i][0] - cen[0], dy = x[i][1] - cen[1], dz = x[i][2] - cen[2];
            r2 = fmax(r2, dx * dx + dy * dy + dz * dz);
            if (SWEPT) {
                const double ex = y[i][0] - cen[0], ey = y[i][1] - cen[1], ez = y[i][2] - cen[2];
                r2 = fmax(r2, ex * ex + ey * ey + ez * ez);
            }
        }
    }
    for (int m = 32; m >= 1; m >>= 1) r2 = fmax(r2, __shfl_xor(r2, m));
    if (ln < 4) sph[4 * (size_t)b + ln] = ln < 3 ? cen[ln == 0 ? 0 : ln == 1 ? 1 : 2] : sqrt(r2) * (1.0 + 1e-12) + 1e-300;
}

// ------------------------------------------------------------------ body pairs
// tiled all-pairs: thread (I, J) over a 2-D grid, only I < J, wave-aggregated append
// body pairs whose (inflated) boxes overlap and whose bounding spheres are
// within `infl` (1e-9 relative slack covers the rounding of the test)
__global__ void k_body_pairs(int B, const double* __restrict__ box, const double* __restrict__ sph, double infl,
                             int2* __restrict__ pairs, int cap, int* __restrict__ counter) {
    const int I = blockIdx.y * 16 + threadIdx.y;
    const int J = blockIdx.x * 16 + threadIdx.x;
    bool hit = false;
    if (I < B && J < B && I < J) {
        const double* a = box + 6 * (size_t)I;
        const double* b = box + 6 * (size_t)J;
        hit = ovl(a, a + 3, b, b + 3);
        if (hit) {
            const double* sa = sph + 4 * (size_t)I;
            const double* sb = sph + 4 * (size_t)J;
            const double dx = sa[0] - sb[0], dy = sa[1] - sb[1], dz = sa[2] - sb[2];
            const double lim = (sa[3] + sb[3] + infl) * (1.0 + 1e-9);
            hit = dx * dx + dy * dy + dz * dz <= lim * lim;
        }
    }
    unsigned long long m = __ballot(hit);
    if (m == 0) return;
    const int lane = __lane_id();
    int base = 0;
    if (lane == __ffsll((long long)m) - 1) base = atomicAdd(counter, __popcll(m));
    base = __shfl(base, __ffsll((long long)m) - 1);
    if (hit) {
        int off = __popcll(m & ((1ull << lane) - 1));
        if (base + off < cap) pairs[base + off] = make_int2(I, J);
    }
}

// outward rounding of doubles to FP32
__device__ __forceinline__ float round_down(double x) { float f = (float)x; return (double)f > x ? nextafterf(f, -INFINITY) : f; }
__device__ __forceinline__ float round_up(double x) { float f = (float)x; return (double)f < x ? nextafterf(f, INFINITY) : f; }

// ------------------------------------------------------------------ sort and sweep
// Bodies sorted by (scene, box min x rounded down to FP32); each sorted body
// scans the following ones of its scene while their min x does not pass its
// max x (rounded up), and tests those with the same box and sphere tests as
// k_body_pairs: the same pairs, without testing all B^2 / 2.
__device__ __forceinline__ unsigned int f32_order(float x) {
    unsigned int b = __float_as_uint(x);
    return (b >> 31) ? ~b : (b | 0x80000000u);
}
__device__ __forceinline__ float f32_from_order(unsigned int k) {
    unsigned int b = (k >> 31) ? (k & 0x7fffffffu) : ~k;
    return __uint_as_float(b);
}
__global__ void k_sweep_keys(int B, const int* __restrict__ scene_of, const double* __restrict__ box,
                             unsigned long long* __restrict__ keys, int* __restrict__ ids) {
    int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= B) return;
    const float lo = round_down(box[6 * (size_t)b]);
    keys[b] = ((unsigned long long)(scene_of ? scene_of[b] : 0) << 32) | f32_order(lo);
    ids[b] = b;
}
__global__ void k_sweep_pairs(int B, const unsigned long long* __restrict__ keys, const int* __restrict__ ids,
                              const double* __restrict__ box, const double* __restrict__ sph, double infl,
                              int2* __restrict__ pairs, int cap, int* __restrict__ counter) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= B) return;
    const int I = ids[i];
    const unsigned int scene = (unsigned int)(keys[i] >> 32);
    const double* a = box + 6 * (size_t)I;
    const float hx = round_up(a[3]);
    const double* sa = sph + 4 * (size_t)I;
    for (int j = i + 1; j < B; j++) {
        const unsigned long long kj = keys[j];
        if ((unsigned int)(kj >> 32) != scene || f32_from_order((unsigned int)kj) > hx) break;
        const int J = ids[j];
        const double* bb = box + 6 * (size_t)J;
        if (!ovl(a, a + 3, bb, bb + 3)) continue;
        const double* sb = sph + 4 * (size_t)J;
        const double dx = sa[0] - sb[0], dy = sa[1] - sb[1], dz = sa[2] - sb[2];
        const double lim = (sa[3] + sb[3] + infl) * (1.0 + 1e-9);
        if (dx * dx + dy * dy + dz * dz > lim * lim) continue;
        const int k = atomicAdd(counter, 1);
        if (k < cap) pairs[k] = make_int2(min(I, J), max(I, J));
    }
}

// ------------------------------------------------------------------ narrow phase
// One workgroup per overlapping body pair, three passes (vertices of I vs
// faces of J, vertices of J vs faces of I, edges of I vs edges of J):
//   1. stage: the primitives of each side whose box meets the overlap region
//      of the two body boxes go to LDS with their boxes (BT = float: rounded
//      outward, the mixed-precision broad phase; BT = double: exact), at most
//      CAP per side per round (longer lists are streamed in rounds);
//   2. box tests: all staged pairs, PAIRS_PT per thread per round, LDS reads
//      only; overlapping pairs go to an LDS candidate queue;
//   3. exact tests (FP64, from the positions in HBM) of the queued candidates:
//      distance type and value against d_hat (sets, CCD = 0), the CCD time of
//      impact (CCD = 1), or, over the segment X0 -> X1 (CCD = 2, swept boxes),
//      the distance at X0 against the reach sqrt(d_hat) plus the largest
//      displacement of each side: a superset of the set at every point of the
//      segment, from which the line search's trial sets are filtered.
// Boxes only cull; the FP64 predicate decides, so the sets equal those of the
// all-pairs tests (model/thread_local_culling.cpp:209-386 pt/ee_col_set_task).
constexpr int PAIRS_PT = 4;                    // box tests per thread per round
// workgroups per body pair: 2 PT passes + the EE pass in ee_slices slices (tuning)
constexpr int QCAP = NB_THREADS * PAIRS_PT * 3 / 2; // candidate queue (i << 16 | j: staged indices < 2^16)

struct NarrowArgs {
    const int* n_pairs; int pair_cap; const int2* pairs;   // pair count on the device (may exceed pair_cap)
    const int* voff; const int* foff; const int* eoff; const int* face; const int* edge;
    const double* X0; const double* X1;   // X1: CCD and swept candidate sets
    const double* box;                    // body boxes
    const double* sph;                    // body bounding spheres (centre, radius)
    double r; double d_hat;
    int enable_pt, enable_ee;
    unsigned long long* pt_keys; unsigned long long* ee_keys;   // NBK buckets of `sub` keys each
    int sub;
    int* bcnt;                            // bucket counters: [0, NBK) pt, [NBK, 2 NBK) ee
    unsigned long long* toi_bits;         // CCD: [1] pt/ee min; per scene [2 s + 1] when scene_of
    const int* scene_of;                  // batched worlds: body -> scene (CCD minimum per scene)
    int ee_slices;                        // workgroups per body pair of the EE pass
};

template <int CCD>
__device__ __forceinline__ void prim_box(const double* X0, const double* X1, const int* ids, int n, double lo[3], double hi[3]) {
    for (int k = 0; k < 3; k++) { lo[k] = 1e300; hi[k] = -1e300; }
    for (int c = 0; c < n; c++) {
        const double* p = X0 + 3 * (size_t)ids[c];
        for (int k = 0; k < 3; k++) { lo[k] = fmin(lo[k], p[k]); hi[k] = fmax(hi[k], p[k]); }
        if (CCD) {
            const double* p1 = X1 + 3 * (size_t)ids[c];
            for (int k = 0; k < 3; k++) { lo[k] = fmin(lo[k], p1[k]); hi[k] = fmax(hi[k], p1[k]); }
        }
    }
}
// outward rounding to the box type: round_down / round_up (above)
template <class BT> __device__ __forceinline__ BT lo_of(double x);
template <class BT> __device__ __forceinline__ BT hi_of(double x);
template <> __device__ __forceinline__ float lo_of<float>(double x) { return round_down(x); }
template <> __device__ __forceinline__ float hi_of<float>(double x) { return round_up(x); }
template <> __device__ __forceinline__ double lo_of<double>(double x) { return x; }
template <> __device__ __forceinline__ double hi_of<double>(double x) { return x; }

// primitive `id` of kind 0 (vertex), 1 (face), 2 (edge): its vertex ids
__device__ __forceinline__ int prim_verts(const NarrowArgs& a, int kind, int id, int v[3]) {
    if (kind == 0) { v[0] = id; return 1; }
    if (kind == 1) { v[0] = a.face[3 * (size_t)id]; v[1] = a.face[3 * (size_t)id + 1]; v[2] = a.face[3 * (size_t)id + 2]; return 3; }
    v[0] = a.edge[2 * (size_t)id]; v[1] = a.edge[2 * (size_t)id + 1]; return 2;
}

// Stage primitives [*cursor, p1) of one kind into LDS until CAP entries are
// filled (whole rounds of NB_THREADS); boxes are inflated by `infl`.  A
// primitive is staged when its box meets the overlap box [xlo, xhi] and comes
// within sc[3] of the other body's bounding-sphere centre sc (that body's
// primitives all lie in its ball, swept positions included, so nothing within
// reach is dropped).  Returns the staged count; *cursor advances past the
// scanned primitives.
template <int CCD, class BT, int CAP>
__device__ int stage_prims(const NarrowArgs& a, int kind, int* cursor, int p1, const double* xlo, const double* xhi,
                           const double* sc, double infl, BT (*bx)[6], int* ids, int* cnt) {
    __syncthreads();
    if (threadIdx.x == 0) *cnt = 0;
    __syncthreads();
    int c = *cursor;
    while (true) {
        // every thread reads *cnt before any thread adds to it again
        const bool go = c < p1 && *cnt + NB_THREADS <= CAP;
        __syncthreads();
        if (!go) break;
        const int p = c + threadIdx.x;
        if (p < p1) {
            int v[3];
            const int nv = prim_verts(a, kind, p, v);
            double lo[3], hi[3];
            prim_box<CCD>(a.X0, a.X1, v, nv, lo, hi);
            double d2 = 0.0;
            for (int i = 0; i < 3; i++) {
                const double g = fmax(fmax(lo[i] - sc[i], sc[i] - hi[i]), 0.0);
                d2 += g * g;
            }
            if (ovl(lo, hi, xlo, xhi) && d2 <= sc[3] * sc[3]) {
                const int k = atomicAdd(cnt, 1);
                ids[k] = p;
                for (int i = 0; i < 3; i++) { bx[k][i] = lo_of<BT>(lo[i] - infl); bx[k][3 + i] = hi_of<BT>(hi[i] + infl); }
            }
        }
        c += NB_THREADS;
        __syncthreads();
    }
    *cursor = c;
    return *cnt;
}

// Keys found by a workgroup gather in LDS (KB of them) and go out in one
// piece: one atomic on the workgroup's bucket counter per flush.
constexpr int KB = 512;
struct KeyBuf { unsigned long long k[KB]; int n; int base; };
// all threads; kind 0 = PT, 1 = EE
__device__ void flush_keys(const NarrowArgs& a, int kind, KeyBuf& kb) {
    __syncthreads();
    const int n = kb.n;
    if (n > 0) {
        const int b = blockIdx.x % NBK;
        if (threadIdx.x == 0) kb.base = atomicAdd(&a.bcnt[kind * NBK + b], n);
        __syncthreads();
        unsigned long long* out = (kind == 0 ? a.pt_keys : a.ee_keys) + (size_t)b * a.sub;
        const int base = kb.base;
        for (int i = threadIdx.x; i < n; i += NB_THREADS)
            if (base + i < a.sub) out[base + i] = kb.k[i];
        __syncthreads();
        if (threadIdx.x == 0) kb.n = 0;
    }
    __syncthreads();
}

// exact tests of the queued candidates (LDS index pairs into the staged lists)
template <int CCD> struct KeySpace { KeyBuf kb; };   // LDS key buffer of the contact-set pass only
template <> struct KeySpace<1> {};

template <int CCD>
__device__ void flush_candidates(const NarrowArgs& a, int kind, const int* idA, const int* idB, const int* q, int* qn,
                                 double& tmin, KeyBuf* kb) {
    __syncthreads();
    const int n = *qn;
    const int nround = (n + NB_THREADS - 1) / NB_THREADS;
    for (int it = 0; it < nround; it++) {
        if constexpr (CCD != 1) {
            // all threads read kb->n before any thread appends to it again
            __syncthreads();
            const bool full = kb->n + NB_THREADS > KB;
            __syncthreads();
            if (full) flush_keys(a, kind, *kb);
        }
        const int t = it * NB_THREADS + threadIdx.x;
        bool hit = false;
        unsigned long long key = 0;
        if (t < n) {
            const int2 c = make_int2(q[t] >> 16, q[t] & 0xffff);
            const int pa = idA[c.x], pb = idB[c.y];
            if (kind == 0) {
                const int v = pa, f = pb;
                const int* fv = a.face + 3 * (size_t)f;
                const d3 p = ld3(a.X0 + 3 * (size_t)v), t0 = ld3(a.X0 + 3 * (size_t)fv[0]), t1 = ld3(a.X0 + 3 * (size_t)fv[1]),
                         t2 = ld3(a.X0 + 3 * (size_t)fv[2]);
                if (CCD == 0) {
                    const int ty = pt_type(p, t0, t1, t2);
                    hit = pt_dist(p, t0, t1, t2, ty) < a.d_hat;
                    key = ((unsigned long long)v << 32) | (unsigned long long)f;
                } else if (CCD == 2) {
                    // the distance moves by at most the largest displacement of
                    // each side along X0 -> X1
                    const double da = sqrt(sqn(ld3(a.X1 + 3 * (size_t)v) - p));
                    const double db = fmax(fmax(sqrt(sqn(ld3(a.X1 + 3 * (size_t)fv[0]) - t0)), sqrt(sqn(ld3(a.X1 + 3 * (size_t)fv[1]) - t1))),
                                           sqrt(sqn(ld3(a.X1 + 3 * (size_t)fv[2]) - t2)));
                    const double reach = (a.r + da + db) * (1.0 + 1e-6) + 1e-9;
                    const int ty = pt_type(p, t0, t1, t2);
                    hit = pt_dist(p, t0, t1, t2, ty) < reach * reach;
                    key = ((unsigned long long)v << 32) | (unsigned long long)f;
                } else {
                    d3 P0[4] = {p, t0, t1, t2};
                    d3 P1[4] = {ld3(a.X1 + 3 * (size_t)v), ld3(a.X1 + 3 * (size_t)fv[0]), ld3(a.X1 + 3 * (size_t)fv[1]),
                                ld3(a.X1 + 3 * (size_t)fv[2])};
                    tmin = fmin(tmin, pt_toi(P0, P1));
                }
            } else {
                const int ea = pa, eb = pb;
                const int* va = a.edge + 2 * (size_t)ea;
                const int* vb = a.edge + 2 * (size_t)eb;
                const d3 a0 = ld3(a.X0 + 3 * (size_t)va[0]), a1 = ld3(a.X0 + 3 * (size_t)va[1]);
                const d3 b0 = ld3(a.X0 + 3 * (size_t)vb[0]), b1 = ld3(a.X0 + 3 * (size_t)vb[1]);
                if (CCD == 0) {
                    const int ty = ee_type(a0, a1, b0, b1);
                    hit = ee_dist(a0, a1, b0, b1, ty) < a.d_hat;
                    key = ((unsigned long long)ea << 32) | (unsigned long long)eb;
                } else if (CCD == 2) {
                    const double da = fmax(sqrt(sqn(ld3(a.X1 + 3 * (size_t)va[0]) - a0)), sqrt(sqn(ld3(a.X1 + 3 * (size_t)va[1]) - a1)));
                    const double db = fmax(sqrt(sqn(ld3(a.X1 + 3 * (size_t)vb[0]) - b0)), sqrt(sqn(ld3(a.X1 + 3 * (size_t)vb[1]) - b1)));
                    const double reach = (a.r + da + db) * (1.0 + 1e-6) + 1e-9;
                    const int ty = ee_type(a0, a1, b0, b1);
                    hit = ee_dist(a0, a1, b0, b1, ty) < reach * reach;
                    key = ((unsigned long long)ea << 32) | (unsigned long long)eb;
                } else {
                    d3 P0[4] = {a0, a1, b0, b1};
                    d3 P1[4] = {ld3(a.X1 + 3 * (size_t)va[0]), ld3(a.X1 + 3 * (size_t)va[1]), ld3(a.X1 + 3 * (size_t)vb[0]),
                                ld3(a.X1 + 3 * (size_t)vb[1])};
                    tmin = fmin(tmin, ee_toi(P0, P1));
                }
            }
        }
        if constexpr (CCD != 1) {
            const unsigned long long m = __ballot(hit);
            if (m) {
                const int lane = __lane_id(), leader = __ffsll((long long)m) - 1;
                int b = 0;
                if (lane == leader) b = atomicAdd(&kb->n, __popcll(m));
                b = __shfl(b, leader);
                if (hit) kb->k[b + __popcll(m & ((1ull << lane) - 1ull))] = key;
            }
        }
    }
    __syncthreads();
    if (threadIdx.x == 0) *qn = 0;
    __syncthreads();
}

template <class BT> struct NarrowCap;
template <> struct NarrowCap<float> { static constexpr int A = 512, B = 512; };
template <> struct NarrowCap<double> { static constexpr int A = 256, B = 256; };

template <int CCD, class BT>
__global__ __launch_bounds__(NB_THREADS) void k_pair_narrow(NarrowArgs a) {
    constexpr int CA = NarrowCap<BT>::A, CB = NarrowCap<BT>::B;
    __shared__ BT bxA[CA][6], bxB[CB][6];
    __shared__ int idA[CA], idB[CB];
    __shared__ int q[QCAP];
    __shared__ int cntA, cntB, qn;
    __shared__ double bmin[NB_THREADS / 64];
    __shared__ KeySpace<CCD> ks;
    KeyBuf* kb = nullptr;
    if constexpr (CCD != 1) {
        kb = &ks.kb;
        if (threadIdx.x == 0) kb->n = 0;
    }
    if (threadIdx.x == 0) qn = 0;
    double tmin = 1.0;
    const double infl = CCD == 1 ? 0.0 : a.r;
    // task (blockIdx.y): 0 = vertices of I vs faces of J, 1 = vertices of J vs
    // faces of I, 2 + q = edges of I in slice q (of a.ee_slices) vs edges of J
    const int task = blockIdx.y;
    const int pass = task < 2 ? task : 2;
    const int kind = pass < 2 ? 0 : 1;                            // 0 = PT, 1 = EE
    const int np = min(*a.n_pairs, a.pair_cap);
    for (int pidx = blockIdx.x; pidx < np; pidx += gridDim.x) {
        const int2 pr = a.pairs[pidx];
        double xlo[3], xhi[3];
        for (int k = 0; k < 3; k++) {
            xlo[k] = fmax(a.box[6 * (size_t)pr.x + k], a.box[6 * (size_t)pr.y + k]);
            xhi[k] = fmin(a.box[6 * (size_t)pr.x + 3 + k], a.box[6 * (size_t)pr.y + 3 + k]);
        }
        // bounding spheres of both bodies, radii grown by the reach of a contact
        // (1e-9 relative slack covers the rounding of the test)
        double sx[4], sy[4];
        for (int k = 0; k < 4; k++) { sx[k] = a.sph[4 * (size_t)pr.x + k]; sy[k] = a.sph[4 * (size_t)pr.y + k]; }
        sx[3] = (sx[3] + infl) * (1.0 + 1e-9); sy[3] = (sy[3] + infl) * (1.0 + 1e-9);
        if (pass < 2 ? a.enable_pt : a.enable_ee) {
            const int bA = pass == 1 ? pr.y : pr.x, bB = pass == 1 ? pr.x : pr.y;
            const double* sA = pass == 1 ? sy : sx;   // sphere of body A: filters B's primitives
            const double* sB = pass == 1 ? sx : sy;
            const int kA = pass < 2 ? 0 : 2, kB = pass < 2 ? 1 : 2;   // primitive kinds
            const int* offA = kA == 0 ? a.voff : a.eoff;
            const int* offB = kB == 1 ? a.foff : a.eoff;
            int curA = offA[bA];
            int endA = offA[bA + 1];
            if (pass == 2) {
                const int len = endA - curA, q = task - 2;
                const int s0 = curA + (int)((long long)len * q / a.ee_slices), s1 = curA + (int)((long long)len * (q + 1) / a.ee_slices);
                curA = s0; endA = s1;
            }
            while (curA < endA) {
                // the A side is inflated by r (vertex or edge of body I), as the all-pairs test
                const int nA = stage_prims<CCD, BT, CA>(a, kA, &curA, endA, xlo, xhi, sB, infl, bxA, idA, &cntA);
                if (nA == 0) continue;
                int curB = offB[bB];
                const int endB = offB[bB + 1];
                while (curB < endB) {
                    const int nB = stage_prims<CCD, BT, CB>(a, kB, &curB, endB, xlo, xhi, sA, 0.0, bxB, idB, &cntB);
                    const int total = nA * nB;
                    for (int t0 = 0; t0 < total; t0 += NB_THREADS * PAIRS_PT) {
                        // all threads read qn before any thread appends to the queue again
                        __syncthreads();
                        const bool full = qn + NB_THREADS * PAIRS_PT > QCAP;
                        __syncthreads();
                        if (full) flush_candidates<CCD>(a, kind, idA, idB, q, &qn, tmin, kb);
    #pragma unroll
                        for (int u = 0; u < PAIRS_PT; u++) {
                            const int t = t0 + u * NB_THREADS + threadIdx.x;
                            if (t < total) {
                                const int i = t / nB, j = t % nB;
                                const bool o = bxA[i][0] <= bxB[j][3] && bxB[j][0] <= bxA[i][3] &&
                                               bxA[i][1] <= bxB[j][4] && bxB[j][1] <= bxA[i][4] &&
                                               bxA[i][2] <= bxB[j][5] && bxB[j][2] <= bxA[i][5];
                                if (o) { const int s = atomicAdd(&qn, 1); q[s] = (i << 16) | j; }
                            }
                        }
                    }
                    flush_candidates<CCD>(a, kind, idA, idB, q, &qn, tmin, kb);
                }
            }
        }
        if (CCD == 1 && a.scene_of) {
            // batched world: this pair's minimum goes to its scene's slot
            for (int m = 32; m >= 1; m >>= 1) tmin = fmin(tmin, __shfl_xor(tmin, m));
            __syncthreads();
            if (__lane_id() == 0) bmin[threadIdx.x / 64] = tmin;
            __syncthreads();
            if (threadIdx.x == 0) {
                double t = bmin[0];
                for (int i = 1; i < NB_THREADS / 64; i++) t = fmin(t, bmin[i]);
                if (t < 1.0) atomicMin(&a.toi_bits[2 * a.scene_of[pr.x] + 1], order_key(t));
            }
            tmin = 1.0;
        }
    }   // pairs
    if constexpr (CCD != 1) flush_keys(a, kind, *kb);
    if (CCD == 1 && !a.scene_of) {
        for (int m = 32; m >= 1; m >>= 1) tmin = fmin(tmin, __shfl_xor(tmin, m));
        if (__lane_id() == 0) bmin[threadIdx.x / 64] = tmin;
        __syncthreads();
        if (threadIdx.x == 0) {
            double t = bmin[0];
            for (int i = 1; i < NB_THREADS / 64; i++) t = fmin(t, bmin[i]);
            if (t < 1.0) atomicMin(&a.toi_bits[1], order_key(t));
        }
    }
}

// ------------------------------------------------------------------ vertex-ground
__global__ void k_vg_scan(int V, const double* __restrict__ X, double ground_y, double d_hat,
                          unsigned long long* keys, int cap, int* counter) {
    int v = blockIdx.x * blockDim.x + threadIdx.x;
    bool hit = false;
    if (v < V) {
        double dd = X[3 * (size_t)v + 1] - ground_y;
        hit = dd * dd < d_hat;
    }
    append_key(hit, (unsigned long long)v, keys, cap, counter);
}
// vertex-ground CCD (model/collision.cpp:83-105)
__global__ void k_vg_toi(int V, const double* __restrict__ X0, const double* __restrict__ X1, double ground_y,
                         unsigned long long* toi_bits, const int* __restrict__ vbody, const int* __restrict__ scene_of) {
    int v = blockIdx.x * blockDim.x + threadIdx.x;
    double t = 1.0;
    if (v < V) {
        double d2 = X1[3 * (size_t)v + 1] - ground_y;
        if (d2 < 0) {
            double d1 = X0[3 * (size_t)v + 1] - ground_y;
            t = fmin(1.0, d1 / (d1 - d2));
        }
    }
    if (scene_of) {   // batched world: each vertex's scene slot (rare: only vertices that reach the ground)
        if (t < 1.0) atomicMin(&toi_bits[2 * scene_of[vbody[v]]], order_key(t));
        return;
    }
    for (int m = 32; m >= 1; m >>= 1) t = fmin(t, __shfl_xor(t, m));
    if (__lane_id() == 0 && t < 1.0) atomicMin(&toi_bits[0], order_key(t));
}

// ------------------------------------------------------------------ decode
// Contact-set keys are unique and sorted by value.  The high part of a key,
// h = key >> shift, is a vertex or edge id (< nh).  The keys are grouped by
// g = h >> gs (2^gs consecutive ids) with a counting sort; then, up to
// RANK_MAX keys, every key goes to its group's start plus the number of
// smaller keys in its group (k_group_rank: one thread per key, a loop over
// its group, groups of 4 ids), above it hipCUB's segmented radix sort orders
// groups of 16 ids (ranking is quadratic in the group size, and the contact
// patches of a dense pile make groups of hundreds of keys; the radix sort
// slows down with more, smaller segments).
// The narrow phase flushes the keys of one body pair and task row together,
// so consecutive keys mostly share a group: the counting passes take one
// atomic per run of equal groups in a wavefront, not one per key
// (device-scope atomics cost ~10 ns each).  Sorting all 43-54 key bits with
// one radix pass per 8-bit digit took 1.5x longer.
constexpr int GS_RANK = 2, GS_RADIX = 4;
constexpr int RANK_MAX = 1 << 20;

// Lanes [0, nv) of the wavefront hold keys of group g.  Each run of equal g
// (in lane order) updates ctr[g] by its length with one atomic: added
// (!FILL), or taken from the run's end (FILL: the run then owns
// [ctr[g] after, ctr[g] before)).  Returns the lane's slot (FILL).
template <bool FILL>
__device__ __forceinline__ int run_atomic(bool valid, unsigned g, int* ctr) {
    const int lane = __lane_id();
    const unsigned gp = __shfl_up(g, 1);
    const bool head = valid && (lane == 0 || gp != g);
    const unsigned long long mh = __ballot(head);
    const int nv = __popcll(__ballot(valid));
    const unsigned long long above = lane == 63 ? 0ull : (mh & (~0ull << (lane + 1)));
    const int next = above ? __ffsll((long long)above) - 1 : nv;
    int base = 0;
    if (head) {
        if (FILL) base = atomicSub(&ctr[g], next - lane) - (next - lane);
        else atomicAdd(&ctr[g], next - lane);
    }
    const unsigned long long upto = mh & (lane == 63 ? ~0ull : ((1ull << (lane + 1)) - 1ull));
    const int myhead = upto ? 63 - __clzll(upto) : 0;
    base = __shfl(base, myhead);
    return base + lane - myhead;
}
// The three contact sets are sorted together: PT keys (v << 32 | f, buckets
// y < NBK), EE keys (ea << 32 | eb, buckets NBK <= y < 2 NBK; bucket y at
// buckets + y * sub) and VG keys (v, one flat list, y = 2 NBK).  Their groups
// follow each other (PT, EE, VG), so the sorted keys are the sorted PT set,
// then the sorted EE set, then the sorted VG set.
struct KeySets {
    BucketList L[2];
    int sub, nvg, V, E, gs, ng_pt, ng_ee;
    const unsigned long long* buckets; const unsigned long long* vg;
};
// the loop runs whole wavefronts.  The fill pass also clears the counts of
// the count pass (clear[0, nclear), read by the scan before it) for the next
// sort.
template <bool FILL>
__global__ void k_group_keys(KeySets ks, int* __restrict__ ctr, unsigned long long* __restrict__ dst, int* __restrict__ clear,
                             int nclear) {
    const int y = blockIdx.y;
    if (FILL) {
        const int nt = gridDim.x * gridDim.y * blockDim.x;
        for (int i = (blockIdx.y * gridDim.x + blockIdx.x) * blockDim.x + threadIdx.x; i < nclear; i += nt) clear[i] = 0;
    }
    int n, shift, nh, gbase;
    const unsigned long long* src;
    if (y < 2 * NBK) {
        const int set = y / NBK;
        n = ks.L[set].cnt[y % NBK];
        src = ks.buckets + (size_t)y * ks.sub;
        shift = 32; nh = set ? ks.E : ks.V; gbase = set ? ks.ng_pt : 0;
    } else {
        n = ks.nvg; src = ks.vg; shift = 0; nh = ks.V; gbase = ks.ng_pt + ks.ng_ee;
    }
    const int lane = __lane_id();
    for (int base = blockIdx.x * blockDim.x + (threadIdx.x & ~63); base < n; base += gridDim.x * blockDim.x) {
        const int i = base + lane;
        const bool valid = i < n;
        unsigned long long key = 0;
        unsigned g = 0;
        if (valid) {
            key = src[i];
            g = (unsigned)gbase + (unsigned)(std::min(key >> shift, (unsigned long long)(nh - 1)) >> ks.gs);
        }
        const int pos = run_atomic<FILL>(valid, g, ctr);
        if (FILL && valid) dst[pos] = key;
    }
}

// Keys [0, npt) of tmp are PT keys, [npt, npt + nee) EE keys and the rest VG
// keys, grouped as k_group_keys<true> placed them (group g in [start[g],
// start[g + 1]), in no order inside).  The keys of a set are unique, so a
// key's rank among the keys of its group is its offset in sorted order; the
// key is decoded there into the contact set (body, local index) entries.
struct SetOut {
    const int* vbody; const int* voff; const int* fbody; const int* foff; const int* ebody; const int* eoff;
    int4* pt; int4* ee; int2* vg;
};
__global__ __launch_bounds__(256) void k_group_rank(KeySets ks, int npt, int nee, int n, const int* __restrict__ start,
                                                    const unsigned long long* __restrict__ tmp, SetOut o) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const unsigned long long key = tmp[i];
    int shift, nh, gbase;
    if (i < npt) { shift = 32; nh = ks.V; gbase = 0; }
    else if (i < npt + nee) { shift = 32; nh = ks.E; gbase = ks.ng_pt; }
    else { shift = 0; nh = ks.V; gbase = ks.ng_pt + ks.ng_ee; }
    const unsigned g = (unsigned)gbase + (unsigned)(std::min(key >> shift, (unsigned long long)(nh - 1)) >> ks.gs);
    const int s = start[g], e = start[g + 1];
    int r = 0;
#pragma unroll 4
    for (int j = s; j < e; j++) r += tmp[j] < key ? 1 : 0;
    const int pos = s + r;
    const int a = (int)(key >> 32), b = (int)(key & 0xffffffffull);
    if (i < npt) {
        const int I = o.vbody[a], J = o.fbody[b];
        o.pt[pos] = make_int4(I, a - o.voff[I], J, b - o.foff[J]);
    } else if (i < npt + nee) {
        const int I = o.ebody[a], J = o.ebody[b];
        o.ee[pos - npt] = make_int4(I, a - o.eoff[I], J, b - o.eoff[J]);
    } else {
        const int I = o.vbody[b];
        o.vg[pos - npt - nee] = make_int2(I, b - o.voff[I]);
    }
}

// Sorted keys [0, npt + nee + nvg) decoded into the three lists.
__global__ void k_decode_all(int npt, int nee, int n, const unsigned long long* __restrict__ keys, SetOut o) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const unsigned long long key = keys[i];
    const int a = (int)(key >> 32), b = (int)(key & 0xffffffffull);
    if (i < npt) {
        const int I = o.vbody[a], J = o.fbody[b];
        o.pt[i] = make_int4(I, a - o.voff[I], J, b - o.foff[J]);
    } else if (i < npt + nee) {
        const int I = o.ebody[a], J = o.ebody[b];
        o.ee[i - npt] = make_int4(I, a - o.eoff[I], J, b - o.eoff[J]);
    } else {
        const int I = o.vbody[b];
        o.vg[i - npt - nee] = make_int2(I, b - o.voff[I]);
    }
}
// Sorts the n keys of ks (npt PT, nee EE, then VG; at most `most` per bucket
// or list) and decodes them into out's lists; tmp and sorted (radix path)
// hold n keys each.
static void sort_keys(World& w, const KeySets& ks, int npt, int nee, int n, int most, unsigned long long* tmp,
                      unsigned long long* sorted, ContactSet& out) {
    ProfScope ps(w, PK_SORT, 40.0 * n);
    const int ng = ks.ng_pt + ks.ng_ee + (w.V >> ks.gs) + 1;
    // the counts are zero between sorts (cleared by the fill pass), once set here
    if (w.hcnt.cap < (size_t)ng + 1) {
        w.hcnt.reserve((size_t)ng + 1);
        ADB_HIP(hipMemsetAsync(w.hcnt.p, 0, w.hcnt.cap * sizeof(int), w.stream));
    }
    w.hend.reserve((size_t)ng + 1);
    const dim3 grid(std::max(1, std::min(64, (most + 255) / 256)), 2 * NBK + 1);
    hipLaunchKernelGGL(k_group_keys<false>, grid, dim3(256), 0, w.stream, ks, w.hcnt.p, nullptr, nullptr, 0);
    ADB_HIP(hipGetLastError());
    size_t tb = 0;
    ADB_HIP(hipcub::DeviceScan::InclusiveSum(nullptr, tb, w.hcnt.p, w.hend.p, ng + 1, w.stream));
    w.sort_tmp.reserve(tb + 16);
    ADB_HIP(hipcub::DeviceScan::InclusiveSum(w.sort_tmp.p, tb, w.hcnt.p, w.hend.p, ng + 1, w.stream));
    // hend[g] = end of group g; after the fill its start, and hend[ng] = n
    hipLaunchKernelGGL(k_group_keys<true>, grid, dim3(256), 0, w.stream, ks, w.hend.p, tmp, w.hcnt.p, ng + 1);
    ADB_HIP(hipGetLastError());
    SetOut o;
    o.vbody = w.vbody.p; o.voff = w.voff.p; o.fbody = w.fbody.p; o.foff = w.foff.p; o.ebody = w.ebody.p; o.eoff = w.eoff.p;
    o.pt = out.pt.p; o.ee = out.ee.p; o.vg = out.vg.p;
    if (ks.gs == GS_RANK) {
        hipLaunchKernelGGL(k_group_rank, dim3((n + 255) / 256), dim3(256), 0, w.stream, ks, npt, nee, n, w.hend.p, tmp, o);
        ADB_HIP(hipGetLastError());
        return;
    }
    // within a group the keys differ in their low 32 + gs bits (VG keys: gs)
    tb = 0;
    ADB_HIP(hipcub::DeviceSegmentedRadixSort::SortKeys(nullptr, tb, tmp, sorted, n, ng, w.hend.p, w.hend.p + 1, 0, 32 + ks.gs,
                                                       w.stream));
    w.sort_tmp.reserve(tb + 16);
    ADB_HIP(hipcub::DeviceSegmentedRadixSort::SortKeys(w.sort_tmp.p, tb, tmp, sorted, n, ng, w.hend.p, w.hend.p + 1, 0, 32 + ks.gs,
                                                       w.stream));
    hipLaunchKernelGGL(k_decode_all, dim3((n + 255) / 256), dim3(256), 0, w.stream, npt, nee, n, sorted, o);
    ADB_HIP(hipGetLastError());
}

// body boxes + pairs; the pair count stays on the device (*cnt, which may
// exceed the capacity: the caller checks it with its other counters).  The
// box kernel clears zero[0, nzero) (which holds *cnt) and sets toi (if given).
static void body_pairs(World& w, const double* X0, const double* X1, double infl, int* cnt, int* zero, int nzero,
                       unsigned long long* toi = nullptr) {
    w.bbox.reserve(6 * (size_t)w.B);
    w.bsph.reserve(4 * (size_t)w.B);
    {
        ProfScope ps(w, PK_BODY_BOX, (X1 ? 48.0 : 24.0) * w.V + 52.0 * w.B);
        if (w.max_vpb <= 64 * BOX_VC && !w.general_paths) {
            auto kern = X1 ? k_body_box_w<true> : k_body_box_w<false>;
            hipLaunchKernelGGL(kern, dim3((w.B + 3) / 4), dim3(256), 0, w.stream, w.B, w.voff.p, X0, X1, infl, w.bbox.p,
                               w.bsph.p, zero, nzero, toi, order_key_host(1.0));
        } else {
            hipLaunchKernelGGL(k_body_box, dim3(w.B), dim3(NB_THREADS), 0, w.stream, w.voff.p, X0, X1, infl, w.bbox.p,
                               w.bsph.p, zero, nzero, toi, order_key_host(1.0));
        }
    }
    ADB_HIP(hipGetLastError());
    if (cnt < zero || cnt >= zero + nzero) throw AdbError(ADB_E_ARG, "body-pair counter outside the cleared range");
    w.pairs.reserve(std::max<size_t>(w.pairs.cap, std::max(4096, 16 * w.B)));
    if (w.sweep) {
        const int B = w.B;
        w.swk.reserve(B); w.swk_s.reserve(B); w.swi.reserve(B); w.swi_s.reserve(B);
        ProfScope ps(w, PK_BODY_PAIRS, 48.0 * w.B);
        hipLaunchKernelGGL(k_sweep_keys, dim3((B + 255) / 256), dim3(256), 0, w.stream, B,
                           w.S > 1 ? (const int*)w.scene_of.p : (const int*)nullptr, w.bbox.p, w.swk.p, w.swi.p);
        ADB_HIP(hipGetLastError());
        int hi = 32;
        while (hi < 64 && (1ll << (hi - 32)) <= w.S) hi++;
        size_t tb = 0;
        ADB_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, tb, w.swk.p, w.swk_s.p, w.swi.p, w.swi_s.p, B, 0, hi, w.stream));
        w.sort_tmp.reserve(tb + 16);
        ADB_HIP(hipcub::DeviceRadixSort::SortPairs(w.sort_tmp.p, tb, w.swk.p, w.swk_s.p, w.swi.p, w.swi_s.p, B, 0, hi, w.stream));
        hipLaunchKernelGGL(k_sweep_pairs, dim3((B + 127) / 128), dim3(128), 0, w.stream, B, w.swk_s.p, w.swi_s.p, w.bbox.p,
                           w.bsph.p, infl, w.pairs.p, (int)w.pairs.cap, cnt);
        ADB_HIP(hipGetLastError());
        return;
    }
    dim3 grid((w.B + 15) / 16, (w.B + 15) / 16), blk(16, 16);
    {
        ProfScope ps(w, PK_BODY_PAIRS, 48.0 * w.B);
        hipLaunchKernelGGL(k_body_pairs, grid, blk, 0, w.stream, w.B, w.bbox.p, w.bsph.p, infl, w.pairs.p, (int)w.pairs.cap, cnt);
    }
    ADB_HIP(hipGetLastError());
}

// workgroups per task of a narrow-phase launch (each loops over the pairs):
// about the pair count of the previous pass of the same kind (contact set,
// CCD, segment candidates: their inflations and sweeps give very different
// counts), so most take one pair
static int narrow_grid(World& w, int kind) {
    const double g = w.pair_hint[kind] * w.tune.narrow_per_pair + 64.0;
    return std::max(1, (int)std::min<double>((double)w.pairs.cap, g));
}

static NarrowArgs narrow_args(World& w, const int* npairs, const double* X0, const double* X1, double r) {
    NarrowArgs a;
    a.n_pairs = npairs; a.pair_cap = (int)w.pairs.cap; a.pairs = w.pairs.p;
    a.voff = w.voff.p; a.foff = w.foff.p; a.eoff = w.eoff.p; a.face = w.face.p; a.edge = w.edge.p;
    a.X0 = X0; a.X1 = X1; a.box = w.bbox.p; a.sph = w.bsph.p; a.r = r; a.d_hat = w.P.d_hat;
    a.enable_pt = w.P.enable_pt && w.F > 0;
    a.enable_ee = w.P.enable_ee && w.E > 0;
    a.pt_keys = nullptr; a.ee_keys = nullptr; a.sub = 0; a.bcnt = nullptr;
    a.toi_bits = nullptr;
    a.scene_of = nullptr;
    a.ee_slices = w.tune.ee_slices;
    return a;
}

// The contact sets at X (X1 == nullptr), or with X1 the PT / EE candidates of
// the segment X -> X1 (no VG), sorted and decoded into out.
static int* small_host(World& w) {
    if (!w.h_small) ADB_HIP(hipHostMalloc((void**)&w.h_small, 256 * sizeof(int), hipHostMallocDefault));
    return w.h_small;
}
static void set_pass(World& w, const double* X, const double* X1, ContactSet& out) {
    const double r = sqrt(w.P.d_hat) * (1.0 + 1e-6) + 1e-12;
    const bool swept = X1 != nullptr;
    const bool hashed = !swept && !(w.P.iaabb & 1);    // spatial hashing (hash.hip) instead of body overlaps
    const bool ground = w.P.enable_ground && !swept;
    w.bcnt.reserve(2 * NBK + 2);             // PT buckets, EE buckets, VG counter, body-pair counter
    int* pcnt = w.bcnt.p + 2 * NBK + 1;
    for (int attempt = 0; attempt < 4; attempt++) {
        const size_t sub = std::max<size_t>(w.keys_a.cap / (2 * NBK), 2048);
        const size_t pcap = sub * NBK;       // keys_a: PT buckets in [0, pcap), EE buckets in [pcap, 2 pcap)
        const size_t vcap = w.V + 64;
        w.keys_a.reserve(2 * pcap);
        w.keys_b.reserve(2 * (2 * pcap + vcap));      // sorted PT, EE, VG; sort scratch
        w.vg_keys.reserve(vcap);
        if (hashed) ADB_HIP(hipMemsetAsync(w.bcnt.p, 0, (2 * NBK + 2) * sizeof(int), w.stream));
        else body_pairs(w, X, X1, r, pcnt, w.bcnt.p, 2 * NBK + 2);   // also clears the bucket counters
        NarrowArgs a = narrow_args(w, pcnt, X, X1, r);
        a.pt_keys = w.keys_a.p; a.ee_keys = w.keys_a.p + pcap; a.sub = (int)sub; a.bcnt = w.bcnt.p;
        if (hashed) {
            hash_sets(w, X, r, a.pt_keys, a.ee_keys, (int)sub, w.bcnt.p);
        } else {
            ProfScope ps(w, PK_NARROW_SET, 0.0);
            const dim3 grid(narrow_grid(w, swept ? 2 : 0), 2 + w.tune.ee_slices);
            if (swept) {
                if (w.P.broad_phase_fp32) hipLaunchKernelGGL((k_pair_narrow<2, float>), grid, dim3(NB_THREADS), 0, w.stream, a);
                else hipLaunchKernelGGL((k_pair_narrow<2, double>), grid, dim3(NB_THREADS), 0, w.stream, a);
            } else {
                if (w.P.broad_phase_fp32) hipLaunchKernelGGL((k_pair_narrow<0, float>), grid, dim3(NB_THREADS), 0, w.stream, a);
                else hipLaunchKernelGGL((k_pair_narrow<0, double>), grid, dim3(NB_THREADS), 0, w.stream, a);
            }
            ADB_HIP(hipGetLastError());
        }
        if (ground) {
            ProfScope ps(w, PK_VG_SCAN, 8.0 * w.V);
            hipLaunchKernelGGL(k_vg_scan, dim3((w.V + 255) / 256), dim3(256), 0, w.stream, w.V, X, w.P.ground_y, w.P.d_hat,
                               w.vg_keys.p, (int)vcap, w.bcnt.p + 2 * NBK);
            ADB_HIP(hipGetLastError());
        }
        static_assert(2 * NBK + 2 <= 256, "bucket counters fit the pinned scratch");
        int* h = small_host(w);
        ADB_HIP(hipMemcpyAsync(h, w.bcnt.p, (2 * NBK + 2) * sizeof(int), hipMemcpyDeviceToHost, w.stream));
        wait_stream(w);
        if (!hashed) w.pair_hint[swept ? 2 : 0] = h[2 * NBK + 1];
        if (!hashed && (size_t)h[2 * NBK + 1] > w.pairs.cap) {   // more overlapping body pairs than slots
            w.pairs.reserve(h[2 * NBK + 1]);
            continue;
        }
        const int nvg = ground ? h[2 * NBK] : 0;
        int most = 0;
        for (int i = 0; i < 2 * NBK; i++) most = std::max(most, h[i]);
        if ((size_t)most > sub) {
            const size_t need = (size_t)most + most / 4 + 64;
            w.keys_a.reserve(2 * NBK * need);
            continue;
        }
        BucketList Lp, Le;
        int npt = 0, nee = 0;
        for (int b = 0; b < NBK; b++) {
            Lp.cnt[b] = h[b]; Lp.off[b] = npt; npt += h[b];
            Le.cnt[b] = h[NBK + b]; Le.off[b] = nee; nee += h[NBK + b];
        }
        out.pt.reserve(std::max(npt, 1)); out.ee.reserve(std::max(nee, 1)); out.vg.reserve(std::max(nvg, 1));
        // sort all three sets (grouped in keys_b's second half, sorted in its
        // first half when radix-sorted) and decode them into out
        const size_t half = 2 * pcap + vcap;
        const int nall = npt + nee + nvg;
        if (nall > 0) {
            KeySets ks;
            ks.L[0] = Lp; ks.L[1] = Le;
            ks.sub = (int)sub; ks.nvg = nvg; ks.V = w.V; ks.E = std::max(w.E, 1);
            ks.gs = nall <= RANK_MAX && !w.general_paths ? GS_RANK : GS_RADIX;
            ks.ng_pt = (w.V >> ks.gs) + 1; ks.ng_ee = (ks.E >> ks.gs) + 1;
            ks.buckets = w.keys_a.p; ks.vg = w.vg_keys.p;
            sort_keys(w, ks, npt, nee, nall, std::max(most, nvg), w.keys_b.p + half, w.keys_b.p, out);
        }
        out.n_pt = npt; out.n_ee = nee; out.n_vg = nvg;
        // PT / EE contacts only come from the overlapping body pairs of the pass
        out.np_bound = hashed ? -1 : h[2 * NBK + 1];
        out.serial = ++w.cs_serial;
        return;
    }
    throw AdbError(ADB_E_CAPACITY, "contact-set capacity could not be grown");
}

void collision_set(World& w, const double* X, ContactSet& out) { set_pass(w, X, nullptr, out); }

void segment_candidates(World& w, const double* X0, const double* X1, ContactSet& cand) { set_pass(w, X0, X1, cand); }

// ---- trial sets from the segment's candidates
// One launch over three ranges of 256-item blocks: PT candidates, EE
// candidates, vertices (VG).  Each item is tested as the narrow phase / the
// VG scan tests it (same distance type and value against d_hat, so the kept
// lists equal the full pass's sets); pass 1 counts the kept items per block,
// pass 2 writes them at the block's scanned offset + their rank in the block,
// which keeps the (sorted) candidate order.
struct TrialArgs {
    int n_pt, n_ee, V, nb_pt, nb_ee, nb_v, ground;
    const int4* pt; const int4* ee;
    const int* voff; const int* foff; const int* eoff; const int* face; const int* edge; const int* vbody;
    const double* X; double d_hat, ground_y;
    int* bcount;                       // kept items per block (pass 1), their exclusive scan (pass 2)
    unsigned long long* keep;          // pass 1's test results, one bit per item (a word per wavefront)
    int4* out_pt; int4* out_ee; int2* out_vg;
};
__device__ __forceinline__ bool trial_keep(const TrialArgs& a, int range, int i) {
    if (range == 0) {
        const int4 c = a.pt[i];
        const int v = a.voff[c.x] + c.y, f = a.foff[c.z] + c.w;
        const int* fv = a.face + 3 * (size_t)f;
        const d3 p = ld3(a.X + 3 * (size_t)v), t0 = ld3(a.X + 3 * (size_t)fv[0]), t1 = ld3(a.X + 3 * (size_t)fv[1]),
                 t2 = ld3(a.X + 3 * (size_t)fv[2]);
        return pt_dist(p, t0, t1, t2, pt_type(p, t0, t1, t2)) < a.d_hat;
    }
    if (range == 1) {
        const int4 c = a.ee[i];
        const int ea = a.eoff[c.x] + c.y, eb = a.eoff[c.z] + c.w;
        const int* va = a.edge + 2 * (size_t)ea;
        const int* vb = a.edge + 2 * (size_t)eb;
        const d3 a0 = ld3(a.X + 3 * (size_t)va[0]), a1 = ld3(a.X + 3 * (size_t)va[1]);
        const d3 b0 = ld3(a.X + 3 * (size_t)vb[0]), b1 = ld3(a.X + 3 * (size_t)vb[1]);
        return ee_dist(a0, a1, b0, b1, ee_type(a0, a1, b0, b1)) < a.d_hat;
    }
    const double dd = a.X[3 * (size_t)i + 1] - a.ground_y;
    return dd * dd < a.d_hat;
}
template <bool WRITE>
__global__ __launch_bounds__(256) void k_trial_set(TrialArgs a) {
    __shared__ int wcnt[4];
    int b = blockIdx.x, range = 0, n = a.n_pt;
    if (b >= a.nb_pt) { b -= a.nb_pt; range = 1; n = a.n_ee; if (b >= a.nb_ee) { b -= a.nb_ee; range = 2; n = a.V; } }
    const int i = b * 256 + threadIdx.x;
    const int wv = threadIdx.x / 64, lane = __lane_id();
    unsigned long long m;
    if (!WRITE) {
        const bool keep = i < n && (range < 2 || a.ground) && trial_keep(a, range, i);
        m = __ballot(keep);
        if (lane == 0) { wcnt[wv] = __popcll(m); a.keep[4 * (size_t)blockIdx.x + wv] = m; }
        __syncthreads();
        if (threadIdx.x == 0) a.bcount[blockIdx.x] = wcnt[0] + wcnt[1] + wcnt[2] + wcnt[3];
        return;
    }
    m = a.keep[4 * (size_t)blockIdx.x + wv];
    if (lane == 0) wcnt[wv] = __popcll(m);
    __syncthreads();
    const bool keep = (m >> lane) & 1ull;
    if (!keep) return;
    int pos = a.bcount[blockIdx.x] + __popcll(m & ((1ull << lane) - 1ull));
    for (int k = 0; k < wv; k++) pos += wcnt[k];
    if (range == 0) a.out_pt[pos] = a.pt[i];
    else if (range == 1) a.out_ee[pos - a.bcount[a.nb_pt]] = a.ee[i];
    else {
        const int I = a.vbody[i];
        a.out_vg[pos - a.bcount[a.nb_pt + a.nb_ee]] = make_int2(I, i - a.voff[I]);
    }
}

// The sets at X from the candidates of a segment through X (their sorted
// order is kept), VG by a scan of all vertices.
__global__ void k_counts3(const int* __restrict__ off, int i0, int i1, int i2, int* __restrict__ dst) {
    if (threadIdx.x == 0) { dst[0] = off[i0]; dst[1] = off[i1]; dst[2] = off[i2]; }
}
void trial_set(World& w, const double* X, const ContactSet& cand, ContactSet& out) {
    w.tcount.reserve(4);
    if (!w.h_tcount) ADB_HIP(hipHostMalloc((void**)&w.h_tcount, 4 * sizeof(int), hipHostMallocDefault));
    TrialArgs a;
    a.n_pt = cand.n_pt; a.n_ee = cand.n_ee; a.V = w.V; a.ground = w.P.enable_ground ? 1 : 0;
    a.nb_pt = (a.n_pt + 255) / 256; a.nb_ee = (a.n_ee + 255) / 256; a.nb_v = a.ground ? (w.V + 255) / 256 : 0;
    const int nb = a.nb_pt + a.nb_ee + a.nb_v;
    a.pt = cand.pt.p; a.ee = cand.ee.p;
    a.voff = w.voff.p; a.foff = w.foff.p; a.eoff = w.eoff.p; a.face = w.face.p; a.edge = w.edge.p; a.vbody = w.vbody.p;
    a.X = X; a.d_hat = w.P.d_hat; a.ground_y = w.P.ground_y;
    // the candidate lists bound the kept PT / EE lists; VG at most one per vertex
    out.pt.reserve(std::max(a.n_pt, 1)); out.ee.reserve(std::max(a.n_ee, 1)); out.vg.reserve(a.ground ? w.V : 1);
    a.out_pt = out.pt.p; a.out_ee = out.ee.p; a.out_vg = out.vg.p;
    w.vsel.reserve(2 * (size_t)nb + 2);
    int* cnt = w.vsel.p;                      // [0, nb]: block counts; [nb + 1, 2 nb + 2): their exclusive scan
    int* off = w.vsel.p + nb + 1;
    w.keepbits.reserve(4 * (size_t)std::max(nb, 1));
    a.bcount = cnt; a.keep = w.keepbits.p;
    ProfScope ps(w, PK_NARROW_SET, 0.0);
    if (nb > 0) {
        hipLaunchKernelGGL(k_trial_set<false>, dim3(nb), dim3(256), 0, w.stream, a);
        ADB_HIP(hipGetLastError());
    }
    ADB_HIP(hipMemsetAsync(cnt + nb, 0, sizeof(int), w.stream));
    size_t tb = 0;
    ADB_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, tb, cnt, off, nb + 1, w.stream));
    w.sort_tmp.reserve(tb + 16);
    ADB_HIP(hipcub::DeviceScan::ExclusiveSum(w.sort_tmp.p, tb, cnt, off, nb + 1, w.stream));
    a.bcount = off;
    if (nb > 0) {
        hipLaunchKernelGGL(k_trial_set<true>, dim3(nb), dim3(256), 0, w.stream, a);
        ADB_HIP(hipGetLastError());
    }
    // the three counts gathered on the device, one read into pinned memory
    hipLaunchKernelGGL(k_counts3, dim3(1), dim3(64), 0, w.stream, off, a.nb_pt, a.nb_pt + a.nb_ee, nb, w.tcount.p);
    ADB_HIP(hipGetLastError());
    ADB_HIP(hipMemcpyAsync(w.h_tcount, w.tcount.p, 3 * sizeof(int), hipMemcpyDeviceToHost, w.stream));
    wait_stream(w);
    const int* h = w.h_tcount;
    out.n_pt = h[0]; out.n_ee = h[1] - h[0]; out.n_vg = h[2] - h[1];
    out.np_bound = cand.np_bound;   // a subset of the candidates
    out.serial = ++w.cs_serial;
}

// The CCD step bound of X0 -> X1 (model/abd.cpp:142-196) in two parts: the
// launches and the copy of the bounds and the body-pair count to pinned
// memory (ccd_queue, no wait), then the wait, a repeat with grown pair slots
// when they overflowed, and the stall test (ccd_finish).  toi_bits holds the
// two bounds, then the body-pair counter (one read for all).
static bool ccd_hashed(const World& w) { return !(w.P.iaabb & 2); }   // spatial hashing (hash.hip) instead of body overlaps
void ccd_queue(World& w, const double* X0, const double* X1) {
    w.toi_bits.reserve(3);
    const bool hashed = ccd_hashed(w);
    int* pcnt = reinterpret_cast<int*>(w.toi_bits.p + 2);
    int* hs = small_host(w);                  // [2, 6): the two 64-bit bounds; [6]: pair count
    unsigned long long* bits = reinterpret_cast<unsigned long long*>(hs + 2);
    // the bounds start at 1 (set by the box kernel, or copied when hashing)
    if (!hashed) {
        body_pairs(w, X0, X1, 0.0, pcnt, pcnt, 1, w.toi_bits.p);
    } else {
        const unsigned long long one = order_key_host(1.0);
        const unsigned long long init[2] = {one, one};
        ADB_HIP(hipMemcpyAsync(w.toi_bits.p, init, sizeof init, hipMemcpyHostToDevice, w.stream));
    }
    if (w.P.enable_ground) {
        ProfScope ps(w, PK_VG_TOI, 16.0 * w.V);
        hipLaunchKernelGGL(k_vg_toi, dim3((w.V + 255) / 256), dim3(256), 0, w.stream, w.V, X0, X1, w.P.ground_y, w.toi_bits.p,
                           w.vbody.p, (const int*)nullptr);
        ADB_HIP(hipGetLastError());
    }
    NarrowArgs a = narrow_args(w, pcnt, X0, X1, 0.0);
    a.toi_bits = w.toi_bits.p;
    if (hashed) {
        hash_ccd(w, X0, X1, w.toi_bits.p);
    } else {
        ProfScope ps(w, PK_NARROW_CCD, 0.0);
        const dim3 grid(narrow_grid(w, 1), 2 + w.tune.ee_slices);
        if (w.P.broad_phase_fp32) hipLaunchKernelGGL((k_pair_narrow<1, float>), grid, dim3(NB_THREADS), 0, w.stream, a);
        else hipLaunchKernelGGL((k_pair_narrow<1, double>), grid, dim3(NB_THREADS), 0, w.stream, a);
        ADB_HIP(hipGetLastError());
    }
    ADB_HIP(hipMemcpyAsync(bits, w.toi_bits.p, (hashed ? 2 : 3) * sizeof(unsigned long long), hipMemcpyDeviceToHost, w.stream));
}
double ccd_finish(World& w, const double* X0, const double* X1, int* status) {
    *status = ADB_OK;
    const bool hashed = ccd_hashed(w);
    int* hs = small_host(w);
    const unsigned long long* bits = reinterpret_cast<const unsigned long long*>(hs + 2);
    for (int attempt = 0;; attempt++) {
        wait_stream(w);
        const int np = hashed ? 0 : hs[6];
        if (!hashed) w.pair_hint[1] = np;
        if ((size_t)np <= w.pairs.cap || attempt == 1) break;
        w.pairs.reserve(np);               // more overlapping body pairs than slots: grow, test again
        ccd_queue(w, X0, X1);
    }
    const double tg = order_key_inv(bits[0]), tp = order_key_inv(bits[1]);
    // toi < 1e-6 twice in a row is a stall (model/iaabb.cpp:305-311,533-537)
    if (tg < 1e-6) { if (++w.stall_vg > 1) *status = ADB_E_TOI_STALL; } else w.stall_vg = 0;
    if (tp < 1e-6) { if (++w.stall_pt > 1) *status = ADB_E_TOI_STALL; } else w.stall_pt = 0;
    return tg < tp ? tg : tp;
}
double ccd_toi(World& w, const double* X0, const double* X1, int* status) {
    ccd_queue(w, X0, X1);
    return ccd_finish(w, X0, X1, status);
}

// Per-scene CCD bounds of a batched world (slots 2 s: vertex-ground, 2 s + 1:
// PT / EE), stalls counted per scene as ccd_toi counts them for one scene.
void ccd_toi_scenes(World& w, const double* X0, const double* X1, double* toi, int* status) {
    const int S = w.S;
    w.toi_bits.reserve(2 * (size_t)S);
    std::vector<unsigned long long> bits(2 * (size_t)S, order_key_host(1.0));
    ADB_HIP(hipMemcpyAsync(w.toi_bits.p, bits.data(), bits.size() * sizeof(unsigned long long), hipMemcpyHostToDevice, w.stream));
    if (w.P.enable_ground) {
        ProfScope ps(w, PK_VG_TOI, 16.0 * w.V);
        hipLaunchKernelGGL(k_vg_toi, dim3((w.V + 255) / 256), dim3(256), 0, w.stream, w.V, X0, X1, w.P.ground_y, w.toi_bits.p,
                           w.vbody.p, w.scene_of.p);
        ADB_HIP(hipGetLastError());
    }
    w.bcnt.reserve(2 * NBK + 2);
    int* pcnt = w.bcnt.p + 2 * NBK + 1;
    for (int attempt = 0; attempt < 2; attempt++) {
        int np = 0;
        body_pairs(w, X0, X1, 0.0, pcnt, pcnt, 1);
        NarrowArgs a = narrow_args(w, pcnt, X0, X1, 0.0);
        a.toi_bits = w.toi_bits.p;
        a.scene_of = w.scene_of.p;
        {
            ProfScope ps(w, PK_NARROW_CCD, 0.0);
            const dim3 grid(narrow_grid(w, 1), 2 + w.tune.ee_slices);
            if (w.P.broad_phase_fp32) hipLaunchKernelGGL((k_pair_narrow<1, float>), grid, dim3(NB_THREADS), 0, w.stream, a);
            else hipLaunchKernelGGL((k_pair_narrow<1, double>), grid, dim3(NB_THREADS), 0, w.stream, a);
            ADB_HIP(hipGetLastError());
        }
        ADB_HIP(hipMemcpyAsync(&np, pcnt, sizeof(int), hipMemcpyDeviceToHost, w.stream));
        ADB_HIP(hipMemcpyAsync(bits.data(), w.toi_bits.p, bits.size() * sizeof(unsigned long long), hipMemcpyDeviceToHost, w.stream));
        wait_stream(w);
        w.pair_hint[1] = np;
        if ((size_t)np <= w.pairs.cap) break;
        w.pairs.reserve(np);
        std::fill(bits.begin(), bits.end(), order_key_host(1.0));
        ADB_HIP(hipMemcpyAsync(w.toi_bits.p, bits.data(), bits.size() * sizeof(unsigned long long), hipMemcpyHostToDevice, w.stream));
        if (w.P.enable_ground)
            hipLaunchKernelGGL(k_vg_toi, dim3((w.V + 255) / 256), dim3(256), 0, w.stream, w.V, X0, X1, w.P.ground_y, w.toi_bits.p,
                               w.vbody.p, w.scene_of.p);
    }
    w.h_stall_pt.resize(S, 0); w.h_stall_vg.resize(S, 0);
    for (int s = 0; s < S; s++) {
        const double tg = order_key_inv(bits[2 * s]), tp = order_key_inv(bits[2 * s + 1]);
        status[s] = ADB_OK;
        if (tg < 1e-6) { if (++w.h_stall_vg[s] > 1) status[s] = ADB_E_TOI_STALL; } else w.h_stall_vg[s] = 0;
        if (tp < 1e-6) { if (++w.h_stall_pt[s] > 1) status[s] = ADB_E_TOI_STALL; } else w.h_stall_pt[s] = 0;
        toi[s] = tg < tp ? tg : tp;
    }
}

} // namespace adb
