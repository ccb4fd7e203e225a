// hash.hip - spatial-hash culling: the reference's alternative to iAABB for
// the contact sets (model/col_set.cpp:20-339, `iaabb` bit 0 clear) and the
// CCD step bound (model/upper_bound.cpp:16-228, `iaabb` bit 1 clear), on the
// grid of model/spatial_hashing.cpp:11-211.
//
// Grid: cell coordinate floor(max(x - min_xyz, 0) / dx) per axis; the hash
// concatenates the low `bits` bits of the three coordinates, so cells 2^bits
// apart share a slot (spatial_hashing.cpp:11-27).  The reference's table has
// n_buffer slots per cell and 8 overflow slots, and loses entries beyond
// them; here the table is a counting sort, exact at any occupancy:
//   1. register: every registered primitive (vertex, or edge) writes its
//      cell range (lower corner + lengths, each at most 2^bits: farther cells
//      alias ones already covered) and counts itself into each slot;
//   2. exclusive scan of the slot counts;
//   3. fill: each primitive writes its id into its slots (atomic cursors; the
//      order inside a slot is irrelevant, the sets are sorted afterwards);
//   4. query: one thread per query primitive (triangle, or edge) walks the
//      slots of its range.  A registered primitive found in several of those
//      slots is taken once, in the first cell of the walk whose slot holds
//      it: per axis, the first coordinate of the query range whose residue
//      mod 2^bits lies in the entry's cyclic residue interval (the
//      reference sorts and de-duplicates instead, spatial_hashing.cpp:111-112).
//      Own-body entries are dropped, EE pairs keep I < J, candidates pass the
//      same FP64 primitive box test as the iAABB path, and the exact FP64
//      predicate decides (distance < d_hat, or the time of impact).
// Query boxes of contact sets are inflated by 2 r (r = the culling radius),
// so every pair whose boxes overlap meets in some cell; the sets and bounds
// therefore equal those of the iAABB path bit for bit.  The edge's CCD range
// spans its end points at t0 and t1 (the reference's
// register_edge_trajectory reads the t0 end points twice,
// spatial_hashing.cpp:150-151).
#include <hipcub/hipcub.hpp>
#include <climits>
#include "dmath.h"
#include "world.h"
#include "cull.h"

namespace adb {

constexpr int HB_THREADS = 256;

struct GridP { double mn, dx; int bits; };

__device__ __forceinline__ int cell_of(double x, const GridP& g) {
    const double u = fmax(x - g.mn, 0.0) / g.dx;
    return u < 1073741824.0 ? (int)u : 1073741824;
}
// lower corner (x, y, z) and lengths - 1 packed 10 bits each (w)
__device__ __forceinline__ int4 cell_range(const double lo[3], const double hi[3], const GridP& g) {
    const int M = 1 << g.bits;
    int c[3], L[3];
    for (int k = 0; k < 3; k++) {
        c[k] = cell_of(lo[k], g);
        L[k] = min(max(cell_of(hi[k], g) - c[k] + 1, 1), M);
    }
    return make_int4(c[0], c[1], c[2], (L[0] - 1) | ((L[1] - 1) << 10) | ((L[2] - 1) << 20));
}
__device__ __forceinline__ int slot_of(int i, int j, int k, int bits) {
    const int m = (1 << bits) - 1;
    return ((i & m) << (2 * bits)) | ((j & m) << bits) | (k & m);
}
// first coordinate of [q, ...) whose residue lies in the entry's residue
// interval [le, le + L) mod M
__device__ __forceinline__ int first_hit(int q, int le, int L, int M) {
    if (L >= M) return q;
    const int d = ((q - le) % M + M) % M;
    return d < L ? q : q + (M - d);
}

// box of the primitive with vertices v[0..nv); CCD: over the positions at t0 and t1
template <int CCD>
__device__ __forceinline__ void prim_bounds(const double* X0, const double* X1, const int* v, int nv, double lo[3], double hi[3]) {
    for (int k = 0; k < 3; k++) { lo[k] = 1e300; hi[k] = -1e300; }
    for (int c = 0; c < nv; c++) {
        const double* p = X0 + 3 * (size_t)v[c];
        for (int k = 0; k < 3; k++) { lo[k] = fmin(lo[k], p[k]); hi[k] = fmax(hi[k], p[k]); }
        if (CCD) {
            const double* p1 = X1 + 3 * (size_t)v[c];
            for (int k = 0; k < 3; k++) { lo[k] = fmin(lo[k], p1[k]); hi[k] = fmax(hi[k], p1[k]); }
        }
    }
}

// 1. ranges and slot counts of the registered primitives (vertices, or edges
// when `edge` is non-null), and the 64-bit total of entries (checked on the
// host before the int scan)
template <int CCD>
__global__ void k_hash_count(int n, const int* __restrict__ edge, const double* __restrict__ X0, const double* __restrict__ X1,
                             GridP g, int4* __restrict__ rng, int* __restrict__ cnt, unsigned long long* __restrict__ total) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    unsigned long long mine = 0;
    if (p < n) {
        int v[2] = {p, p};
        int nv = 1;
        if (edge) { v[0] = edge[2 * (size_t)p]; v[1] = edge[2 * (size_t)p + 1]; nv = 2; }
        double lo[3], hi[3];
        prim_bounds<CCD>(X0, X1, v, nv, lo, hi);
        const int4 r = cell_range(lo, hi, g);
        rng[p] = r;
        const int Lx = (r.w & 1023) + 1, Ly = ((r.w >> 10) & 1023) + 1, Lz = ((r.w >> 20) & 1023) + 1;
        mine = (unsigned long long)Lx * Ly * Lz;
        for (int i = 0; i < Lx; i++)
            for (int j = 0; j < Ly; j++)
                for (int k = 0; k < Lz; k++) atomicAdd(&cnt[slot_of(r.x + i, r.y + j, r.z + k, g.bits)], 1);
    }
    for (int m = 32; m >= 1; m >>= 1) mine += __shfl_xor(mine, m);
    if (__lane_id() == 0 && mine) atomicAdd(total, mine);
}

// 3. slot entries; the counts are consumed (left at zero)
__global__ void k_hash_fill(int n, const int4* __restrict__ rng, GridP g, const int* __restrict__ start, int* __restrict__ cnt,
                            int* __restrict__ ent) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    const int4 r = rng[p];
    const int Lx = (r.w & 1023) + 1, Ly = ((r.w >> 10) & 1023) + 1, Lz = ((r.w >> 20) & 1023) + 1;
    for (int i = 0; i < Lx; i++)
        for (int j = 0; j < Ly; j++)
            for (int k = 0; k < Lz; k++) {
                const int h = slot_of(r.x + i, r.y + j, r.z + k, g.bits);
                ent[start[h] + atomicSub(&cnt[h], 1) - 1] = p;
            }
}

struct HashQueryArgs {
    GridP g;
    int nq;                                  // query primitives (faces, or edges)
    const int* face; const int* edge;        // global vertex ids
    const int* vbody; const int* fbody; const int* ebody;
    const double* X0; const double* X1;      // X1 only for CCD
    const int4* rng; const int* start; const int* ent;
    double r; double d_hat;
    unsigned long long* keys; int sub; int* bcnt;      // contact sets: NBK buckets of `sub` keys, their counters
    unsigned long long* toi_bits;                       // CCD: [1] pair minimum
};

// 4. queries.  PT: registered vertices vs query faces (pairs (v, f) of
// different bodies); EE: registered edges vs query edges (pairs (ea, eb) with
// body(ea) < body(eb)).  Keys v << 32 | f and ea << 32 | eb, as the iAABB path.
template <int EE, int CCD>
__global__ __launch_bounds__(HB_THREADS) void k_hash_query(HashQueryArgs a) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    double tmin = 1.0;
    if (q < a.nq) {
        const GridP g = a.g;
        const int M = 1 << g.bits;
        int qv[3];
        const int nqv = EE ? 2 : 3;
        for (int c = 0; c < nqv; c++) qv[c] = EE ? a.edge[2 * (size_t)q + c] : a.face[3 * (size_t)q + c];
        const int J = EE ? a.ebody[q] : a.fbody[q];
        double blo[3], bhi[3];                 // the query primitive's box (not inflated)
        prim_bounds<CCD>(a.X0, a.X1, qv, nqv, blo, bhi);
        double slo[3], shi[3];                 // its search box
        const double infl = CCD ? 0.0 : 2.0 * a.r;
        for (int k = 0; k < 3; k++) { slo[k] = blo[k] - infl; shi[k] = bhi[k] + infl; }
        const int4 rq = cell_range(slo, shi, g);
        const int Lx = (rq.w & 1023) + 1, Ly = ((rq.w >> 10) & 1023) + 1, Lz = ((rq.w >> 20) & 1023) + 1;
        const d3 Q0 = ld3(a.X0 + 3 * (size_t)qv[0]), Q1 = ld3(a.X0 + 3 * (size_t)qv[1]);
        const d3 Q2 = EE ? Q1 : ld3(a.X0 + 3 * (size_t)qv[2]);
        for (int i = 0; i < Lx; i++)
            for (int j = 0; j < Ly; j++)
                for (int k = 0; k < Lz; k++) {
                    const int cx = rq.x + i, cy = rq.y + j, cz = rq.z + k;
                    const int h = slot_of(cx, cy, cz, g.bits);
                    const int s1 = a.start[h + 1];
                    for (int s = a.start[h]; s < s1; s++) {
                        const int e = a.ent[s];
                        const int bE = EE ? a.ebody[e] : a.vbody[e];
                        if (EE ? bE >= J : bE == J) continue;
                        const int4 re = a.rng[e];
                        if (cx != first_hit(rq.x, re.x, (re.w & 1023) + 1, M) ||
                            cy != first_hit(rq.y, re.y, ((re.w >> 10) & 1023) + 1, M) ||
                            cz != first_hit(rq.z, re.z, ((re.w >> 20) & 1023) + 1, M))
                            continue;
                        int ev[2] = {e, e};
                        if (EE) { ev[0] = a.edge[2 * (size_t)e]; ev[1] = a.edge[2 * (size_t)e + 1]; }
                        double elo[3], ehi[3];
                        prim_bounds<CCD>(a.X0, a.X1, ev, EE ? 2 : 1, elo, ehi);
                        if (!CCD)
                            for (int c = 0; c < 3; c++) { elo[c] -= a.r; ehi[c] += a.r; }
                        if (!ovl(elo, ehi, blo, bhi)) continue;
                        if (!CCD) {
                            bool hit;
                            if (EE) {
                                const d3 a0 = ld3(a.X0 + 3 * (size_t)ev[0]), a1 = ld3(a.X0 + 3 * (size_t)ev[1]);
                                hit = ee_dist(a0, a1, Q0, Q1, ee_type(a0, a1, Q0, Q1)) < a.d_hat;
                            } else {
                                const d3 p = ld3(a.X0 + 3 * (size_t)e);
                                hit = pt_dist(p, Q0, Q1, Q2, pt_type(p, Q0, Q1, Q2)) < a.d_hat;
                            }
                            if (hit) {
                                const int b = blockIdx.x % NBK;
                                const int at = atomicAdd(&a.bcnt[b], 1);
                                if (at < a.sub) a.keys[(size_t)b * a.sub + at] = ((unsigned long long)e << 32) | (unsigned long long)q;
                            }
                        } else if (EE) {
                            d3 P0[4] = {ld3(a.X0 + 3 * (size_t)ev[0]), ld3(a.X0 + 3 * (size_t)ev[1]), Q0, Q1};
                            d3 P1[4] = {ld3(a.X1 + 3 * (size_t)ev[0]), ld3(a.X1 + 3 * (size_t)ev[1]),
                                        ld3(a.X1 + 3 * (size_t)qv[0]), ld3(a.X1 + 3 * (size_t)qv[1])};
                            tmin = fmin(tmin, ee_toi(P0, P1));
                        } else {
                            d3 P0[4] = {ld3(a.X0 + 3 * (size_t)e), Q0, Q1, Q2};
                            d3 P1[4] = {ld3(a.X1 + 3 * (size_t)e), ld3(a.X1 + 3 * (size_t)qv[0]), ld3(a.X1 + 3 * (size_t)qv[1]),
                                        ld3(a.X1 + 3 * (size_t)qv[2])};
                            tmin = fmin(tmin, pt_toi(P0, P1));
                        }
                    }
                }
    }
    if (CCD) {
        for (int m = 32; m >= 1; m >>= 1) tmin = fmin(tmin, __shfl_xor(tmin, m));
        if (__lane_id() == 0 && tmin < 1.0) atomicMin(&a.toi_bits[1], order_key(tmin));
    }
}

static GridP grid_params(const World& w) { return GridP{w.P.sh_min_xyz, w.P.sh_dx, w.P.sh_bits}; }

// Builds the table of the vertices (edges = false) or edges of the scene;
// CCD: ranges over the positions at X0 and X1.
static void hash_build(World& w, bool edges, const double* X0, const double* X1) {
    const GridP g = grid_params(w);
    const int n = edges ? w.E : w.V;
    const size_t nslot = (size_t)1 << (3 * g.bits);
    w.sh_cnt.reserve(nslot + 1);
    w.sh_start.reserve(nslot + 1);
    w.sh_rng.reserve(std::max(n, 1));
    ProfScope ps(w, PK_HASH_BUILD, 0.0);
    w.sh_total.reserve(1);
    ADB_HIP(hipMemsetAsync(w.sh_cnt.p, 0, (nslot + 1) * sizeof(int), w.stream));
    ADB_HIP(hipMemsetAsync(w.sh_total.p, 0, sizeof(unsigned long long), w.stream));
    const dim3 grid((n + HB_THREADS - 1) / HB_THREADS);
    if (n > 0) {
        if (X1) hipLaunchKernelGGL(k_hash_count<1>, grid, dim3(HB_THREADS), 0, w.stream, n, edges ? w.edge.p : nullptr, X0, X1, g,
                                   w.sh_rng.p, w.sh_cnt.p, w.sh_total.p);
        else hipLaunchKernelGGL(k_hash_count<0>, grid, dim3(HB_THREADS), 0, w.stream, n, edges ? w.edge.p : nullptr, X0, X1, g,
                                w.sh_rng.p, w.sh_cnt.p, w.sh_total.p);
        ADB_HIP(hipGetLastError());
    }
    unsigned long long total = 0;
    ADB_HIP(hipMemcpyAsync(&total, w.sh_total.p, sizeof total, hipMemcpyDeviceToHost, w.stream));
    wait_stream(w);
    // the slot counts are left set: a failed build must not leave them for the next
    if (total > (unsigned long long)INT_MAX - 1) {
        ADB_HIP(hipMemsetAsync(w.sh_cnt.p, 0, (nslot + 1) * sizeof(int), w.stream));
        throw AdbError(ADB_E_CAPACITY, "spatial hash: more than 2^31 cell entries (cell size too small for the scene)");
    }
    size_t tmp = 0;
    ADB_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, tmp, w.sh_cnt.p, w.sh_start.p, (int)(nslot + 1), w.stream));
    w.sort_tmp.reserve(tmp + 16);
    ADB_HIP(hipcub::DeviceScan::ExclusiveSum(w.sort_tmp.p, tmp, w.sh_cnt.p, w.sh_start.p, (int)(nslot + 1), w.stream));
    w.sh_ent.reserve(std::max<size_t>(total, 1));
    if (n > 0) {
        hipLaunchKernelGGL(k_hash_fill, grid, dim3(HB_THREADS), 0, w.stream, n, w.sh_rng.p, g, w.sh_start.p, w.sh_cnt.p, w.sh_ent.p);
        ADB_HIP(hipGetLastError());
    }
}

static HashQueryArgs query_args(World& w, int nq, const double* X0, const double* X1, double r) {
    HashQueryArgs a;
    a.g = grid_params(w); a.nq = nq;
    a.face = w.face.p; a.edge = w.edge.p; a.vbody = w.vbody.p; a.fbody = w.fbody.p; a.ebody = w.ebody.p;
    a.X0 = X0; a.X1 = X1;
    a.rng = w.sh_rng.p; a.start = w.sh_start.p; a.ent = w.sh_ent.p;
    a.r = r; a.d_hat = w.P.d_hat;
    a.keys = nullptr; a.sub = 0; a.bcnt = nullptr; a.toi_bits = nullptr;
    return a;
}

void hash_sets(World& w, const double* X, double r, unsigned long long* pt_keys, unsigned long long* ee_keys, int sub,
               int* bcnt) {
    if (w.P.enable_pt && w.F > 0) {
        hash_build(w, false, X, nullptr);
        HashQueryArgs a = query_args(w, w.F, X, nullptr, r);
        a.keys = pt_keys; a.sub = sub; a.bcnt = bcnt;
        ProfScope ps(w, PK_HASH_QUERY, 0.0);
        hipLaunchKernelGGL((k_hash_query<0, 0>), dim3((w.F + HB_THREADS - 1) / HB_THREADS), dim3(HB_THREADS), 0, w.stream, a);
        ADB_HIP(hipGetLastError());
    }
    if (w.P.enable_ee && w.E > 0) {
        hash_build(w, true, X, nullptr);
        HashQueryArgs a = query_args(w, w.E, X, nullptr, r);
        a.keys = ee_keys; a.sub = sub; a.bcnt = bcnt + NBK;
        ProfScope ps(w, PK_HASH_QUERY, 0.0);
        hipLaunchKernelGGL((k_hash_query<1, 0>), dim3((w.E + HB_THREADS - 1) / HB_THREADS), dim3(HB_THREADS), 0, w.stream, a);
        ADB_HIP(hipGetLastError());
    }
}

void hash_ccd(World& w, const double* X0, const double* X1, unsigned long long* toi_bits) {
    if (w.P.enable_pt && w.F > 0) {
        hash_build(w, false, X0, X1);
        HashQueryArgs a = query_args(w, w.F, X0, X1, 0.0);
        a.toi_bits = toi_bits;
        ProfScope ps(w, PK_HASH_QUERY, 0.0);
        hipLaunchKernelGGL((k_hash_query<0, 1>), dim3((w.F + HB_THREADS - 1) / HB_THREADS), dim3(HB_THREADS), 0, w.stream, a);
        ADB_HIP(hipGetLastError());
    }
    if (w.P.enable_ee && w.E > 0) {
        hash_build(w, true, X0, X1);
        HashQueryArgs a = query_args(w, w.E, X0, X1, 0.0);
        a.toi_bits = toi_bits;
        ProfScope ps(w, PK_HASH_QUERY, 0.0);
        hipLaunchKernelGGL((k_hash_query<1, 1>), dim3((w.E + HB_THREADS - 1) / HB_THREADS), dim3(HB_THREADS), 0, w.stream, a);
        ADB_HIP(hipGetLastError());
    }
}

} // namespace adb
