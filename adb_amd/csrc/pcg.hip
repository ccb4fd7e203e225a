// pcg.hip - block-Jacobi preconditioned conjugate gradients on the BSR
// Hessian, FP64 throughout.  Replaces the direct PardisoLLT / SimplicialLLT
// solve of ABD::solve (model/abd.cpp:263-291): dq = -H^{-1} g.
//
// Two launches per iteration, no atomics and no host round trip:
//   spmv(k)  : every wave first sums the update partials of iteration k - 1
//              (r.z, r.r per update workgroup) in a fixed order; from them the
//              stop test (r.r <= rtol^2 b.b) and beta = rz_k / rz_{k-1}; then
//              p = z + beta p_old (on the fly for every column block, written
//              for the workgroup's own rows into the other p buffer), y = H p,
//              and one p.y partial per workgroup
//   update(k): every wave sums the p.y partials in a fixed order, alpha =
//              rz / pAp, x += alpha p, r -= alpha y, z = M^-1 r (explicit
//              inverse of the diagonal block, 12 lanes per body), and one
//              (r.z, r.r) partial per workgroup
// Every workgroup evaluates the same scalars from the same partials in the
// same order, so they agree bit for bit; once the stop test holds every kernel
// returns at once, and the host queues iterations in batches, checking the
// flag of the previous batch while the current one runs.
#include <cmath>
#include <functional>
#include "dmath.h"
#include "world.h"

namespace adb {

// scalars
enum { S_STOP = 0, S_ITERS, S_BB, S_RZ0, S_RZ1, S_COUNT };
// stop flag values
enum { RUN = 0, CONVERGED = 1, BREAKDOWN = 2, NONFINITE = 3 };

constexpr int SP_ROWS = 4;       // block rows per spmv workgroup (one 64-lane team each)
// iterations queued between host checks: w.tune.pcg_batch (even)

// fixed-order wave sum, broadcast from lane 0
__device__ __forceinline__ double wave_sum(double v) {
    for (int m = 32; m >= 1; m >>= 1) v += __shfl_down(v, m);
    return __shfl(v, 0);
}
// Fixed-order sums of n values (lane l adds l, l + 64, ... in order, then
// the wave sum).  The loads are issued 8 per lane at a time: a loop of
// dependent adds would wait for one L2 round trip per value, which for ~1000
// partials made up most of a PCG iteration.
constexpr int SUM_U = 8;
__device__ __forceinline__ double wave_sum_array(const double* v, int n) {
    const int l = threadIdx.x % 64;
    double s = 0.0;
    for (int i0 = l; i0 < n; i0 += SUM_U * 64) {
        double t[SUM_U];
#pragma unroll
        for (int u = 0; u < SUM_U; u++) t[u] = i0 + 64 * u < n ? v[i0 + 64 * u] : 0.0;
#pragma unroll
        for (int u = 0; u < SUM_U; u++)
            if (i0 + 64 * u < n) s += t[u];
    }
    return wave_sum(s);
}
__device__ __forceinline__ void wave_sum_array2(const double* a, const double* b, int n, double& ra, double& rb) {
    const int l = threadIdx.x % 64;
    double sa = 0.0, sb = 0.0;
    for (int i0 = l; i0 < n; i0 += SUM_U * 64) {
        double ta[SUM_U], tb[SUM_U];
#pragma unroll
        for (int u = 0; u < SUM_U; u++) {
            const bool in = i0 + 64 * u < n;
            ta[u] = in ? a[i0 + 64 * u] : 0.0;
            tb[u] = in ? b[i0 + 64 * u] : 0.0;
        }
#pragma unroll
        for (int u = 0; u < SUM_U; u++)
            if (i0 + 64 * u < n) { sa += ta[u]; sb += tb[u]; }
    }
    ra = wave_sum(sa);
    rb = wave_sum(sb);
}
// fixed-order workgroup sum (one value per wave combined in wave order)
__device__ double wg_sum(double v, double* sh) {
    v = wave_sum(v);
    const int w = threadIdx.x / 64, nw = (blockDim.x + 63) / 64;
    __syncthreads();
    if (threadIdx.x % 64 == 0) sh[w] = v;
    __syncthreads();
    double r = 0.0;
    for (int i = 0; i < nw; i++) r += sh[i];
    __syncthreads();
    return r;
}

// Preconditioner application for one body group (one wave per group):
// z = M_g^-1 r on the group's rows (n = 12 bodies <= GN), rows i = lane, lane + 64.
constexpr int GN = 96;
struct Groups { const int* off; const int* body; const long long* moff; const double* Minv; const int* count; };

// initial: x = 0, r = -g, z = M^-1 r, p_old = 0, (r.z, r.r) partials per group
__global__ __launch_bounds__(64) void k_pcg_init(int B, Groups G, const double* __restrict__ g, double* __restrict__ x,
                                                 double* __restrict__ r, double* __restrict__ z, double* __restrict__ p0,
                                                 double* __restrict__ partB, double* __restrict__ scal, int* __restrict__ kbase) {
    __shared__ double rs[GN];
    __shared__ int bod[GN / 12];
    const int gi = blockIdx.x, l = threadIdx.x;
    if (gi == 0 && l == 0) { scal[S_STOP] = RUN; scal[S_ITERS] = 0.0; *kbase = 0; }
    if (gi >= *G.count) return;
    const int m = G.off[gi + 1] - G.off[gi], n = 12 * m;
    if (l < m) bod[l] = G.body[G.off[gi] + l];
    __syncthreads();
    for (int i = l; i < n; i += 64) {
        const size_t k = 12 * (size_t)bod[i / 12] + i % 12;
        const double rk = -g[k];
        x[k] = 0.0; r[k] = rk; p0[k] = 0.0;
        rs[i] = rk;
    }
    __syncthreads();
    const double* M = G.Minv + G.moff[gi];
    double rz = 0.0, rr = 0.0;
    for (int i = l; i < n; i += 64) {
        double zi = 0.0;
        for (int j = 0; j < n; j++) zi += M[(size_t)j * n + i] * rs[j];
        z[12 * (size_t)bod[i / 12] + i % 12] = zi;
        rz += rs[i] * zi; rr += rs[i] * rs[i];
    }
    rz = wave_sum(rz);
    rr = wave_sum(rr);
    if (l == 0) { partB[gi] = rz; partB[B + gi] = rr; }
}
// Batched worlds: each scene's residual test, |r_s|^2 <= rtol^2 |b_s|^2, from
// the group partials (groups are numbered scene by scene); first = the
// partials of iteration 0 (|b_s|^2 = |r_s|^2 then).  One workgroup; brun = 1
// while some scene is above its tolerance.
__global__ __launch_bounds__(1024) void k_pcg_scene_stop(int S, int B, const int* __restrict__ sc_goff, double rtol2, int first,
                                                         const double* __restrict__ partB, double* __restrict__ sbb,
                                                         const double* __restrict__ scal, int* __restrict__ brun) {
    __shared__ int any;
    if (threadIdx.x == 0) any = 0;
    __syncthreads();
    if (scal[S_STOP] == RUN) {
        int mine = 0;
        for (int s = threadIdx.x; s < S; s += blockDim.x) {
            double rr = 0.0;
            for (int g = sc_goff[s]; g < sc_goff[s + 1]; g++) rr += partB[B + g];
            if (first) sbb[s] = rr;
            if (!(rr <= rtol2 * sbb[s])) mine = 1;
        }
        if (mine) any = 1;
    }
    __syncthreads();
    if (threadIdx.x == 0) *brun = any;
}

// iteration base of the next batch
__global__ void k_pcg_advance(int* kbase, int n) { if (threadIdx.x == 0 && blockIdx.x == 0) *kbase += n; }

__global__ __launch_bounds__(64 * SP_ROWS) void k_pcg_spmv(int B, const int* __restrict__ kbase, int kloc, double rtol2,
                                                           const int* __restrict__ gcount,
                                                           const int* __restrict__ row_ptr,
                                                           const int* __restrict__ col_idx, const double* __restrict__ Hval,
                                                           const double* __restrict__ z, const double* __restrict__ pold,
                                                           double* __restrict__ pnew, double* __restrict__ y,
                                                           const double* __restrict__ partB, double* __restrict__ partA,
                                                           double* __restrict__ scal, const int* __restrict__ brun) {
    __shared__ double sh[SP_ROWS];
    __shared__ double acc_sh[SP_ROWS][5][12];
    if (scal[S_STOP] != RUN) return;
    const int iter = *kbase + kloc;
    // scalars of iteration iter - 1 (identical in every wave of every workgroup)
    const int nbU = *gcount;
    double RZ, RR;
    wave_sum_array2(partB, partB + B, nbU, RZ, RR);
    const double bb = iter == 0 ? RR : scal[S_BB];
    int stop = RUN;
    if (!std::isfinite(RR) || !std::isfinite(RZ)) stop = NONFINITE;
    else if (brun ? *brun == 0 : RR <= rtol2 * bb) stop = CONVERGED;   // batched: every scene's own test
    else if (!(RZ > 0.0)) stop = BREAKDOWN;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        if (iter == 0) scal[S_BB] = RR;
        scal[S_RZ0 + (iter & 1)] = RZ;
        if (stop != RUN) { scal[S_STOP] = stop; scal[S_ITERS] = iter; }
    }
    if (stop != RUN) return;
    const double beta = iter == 0 ? 0.0 : RZ / scal[S_RZ0 + ((iter - 1) & 1)];
    const int team = threadIdx.x / 64, lane = threadIdx.x % 64;
    const int row = blockIdx.x * SP_ROWS + team;
    const int g = lane / 12, r = lane % 12;
    if (row < B && g < 5) {
        double acc = 0.0;
        for (int k = row_ptr[row] + g; k < row_ptr[row + 1]; k += 5) {
            const double* H = Hval + 144 * (size_t)k;
            const size_t cb = 12 * (size_t)col_idx[k];
            double s = 0.0;
#pragma unroll
            for (int c = 0; c < 12; c++) s += H[c * 12 + r] * (z[cb + c] + beta * pold[cb + c]);
            acc += s;
        }
        acc_sh[team][g][r] = acc;
    }
    __syncthreads();
    double prod = 0.0;
    if (row < B && lane < 12) {
        const size_t k = 12 * (size_t)row + lane;
        const double yr = (((acc_sh[team][0][lane] + acc_sh[team][1][lane]) + acc_sh[team][2][lane]) + acc_sh[team][3][lane]) +
                          acc_sh[team][4][lane];
        const double pr = z[k] + beta * pold[k];
        y[k] = yr;
        pnew[k] = pr;
        prod = pr * yr;
    }
    prod = wg_sum(prod, sh);
    if (threadIdx.x == 0) partA[blockIdx.x] = prod;
}

// UPD_T lanes per group: row i = t % GN of the group, column half t / GN of
// z = M^-1 r, the two halves combined in order
constexpr int UPD_T = 2 * GN;
__global__ __launch_bounds__(UPD_T) void k_pcg_update(int B, const int* __restrict__ kbase, int kloc, int nbS, Groups G,
                                                      const double* __restrict__ p, const double* __restrict__ y,
                                                      double* __restrict__ x, double* __restrict__ r, double* __restrict__ z,
                                                      const double* __restrict__ partA, double* __restrict__ partB,
                                                      double* __restrict__ scal) {
    __shared__ double rs[GN];
    __shared__ double half2[GN];
    __shared__ double red[UPD_T / 64][2];
    __shared__ int bod[GN / 12];
    if (scal[S_STOP] != RUN) return;
    const int gi = blockIdx.x, t = threadIdx.x;
    if (gi > 0 && gi >= *G.count) return;     // workgroup 0 also reports a breakdown
    const int iter = *kbase + kloc;
    const double pAp = wave_sum_array(partA, nbS);
    const double RZ = scal[S_RZ0 + (iter & 1)];
    if (!(pAp > 0.0) || !std::isfinite(pAp)) {
        if (gi == 0 && t == 0) { scal[S_STOP] = std::isfinite(pAp) ? BREAKDOWN : NONFINITE; scal[S_ITERS] = iter; }
        return;
    }
    if (gi >= *G.count) return;
    const double alpha = RZ / pAp;
    const int m = G.off[gi + 1] - G.off[gi], n = 12 * m;
    if (t < m) bod[t] = G.body[G.off[gi] + t];
    __syncthreads();
    if (t < n) {
        const size_t k = 12 * (size_t)bod[t / 12] + t % 12;
        x[k] += alpha * p[k];
        const double rk = r[k] - alpha * y[k];
        r[k] = rk;
        rs[t] = rk;
    }
    __syncthreads();
    const double* M = G.Minv + G.moff[gi];
    const int i = t % GN, h = t / GN, j0 = h * ((n + 1) / 2), j1 = h ? n : (n + 1) / 2;
    double zi = 0.0;
    if (i < n) {
        // loads SUM_U at a time, products added in order of j
        int j = j0;
        for (; j + SUM_U <= j1; j += SUM_U) {
            double m[SUM_U];
#pragma unroll
            for (int u = 0; u < SUM_U; u++) m[u] = M[(size_t)(j + u) * n + i];
#pragma unroll
            for (int u = 0; u < SUM_U; u++) zi += m[u] * rs[j + u];
        }
        for (; j < j1; j++) zi += M[(size_t)j * n + i] * rs[j];
    }
    if (h == 1 && i < n) half2[i] = zi;
    __syncthreads();
    double rz = 0.0, rr = 0.0;
    if (h == 0 && i < n) {
        zi += half2[i];
        z[12 * (size_t)bod[i / 12] + i % 12] = zi;
        rz = rs[i] * zi; rr = rs[i] * rs[i];
    }
    rz = wave_sum(rz);
    rr = wave_sum(rr);
    if (t % 64 == 0) { red[t / 64][0] = rz; red[t / 64][1] = rr; }
    __syncthreads();
    if (t == 0) {
        double a0 = 0.0, a1 = 0.0;
        for (int w = 0; w < UPD_T / 64; w++) { a0 += red[w][0]; a1 += red[w][1]; }
        partB[gi] = a0; partB[B + gi] = a1;
    }
}

static Groups groups_of(World& w) {
    Groups G;
    G.off = w.grp_off; G.body = w.grp_body; G.moff = w.grp_moff; G.Minv = w.Ginv.p; G.count = w.grp_n;
    return G;
}

// One batch = pcg_batch iterations; kernel k of the batch computes iteration
// *kbase + k, and the last node advances *kbase.  Because the batch is even, the
// p-buffer parity of local iteration k is k & 1 in every batch, so one
// captured graph serves every batch.
static void enqueue_batch(World& w, int nb, bool prof) {
    const int B = w.B;
    const int nbS = (B + SP_ROWS - 1) / SP_ROWS;
    const double rtol2 = w.P.pcg_rtol * w.P.pcg_rtol;
    double* partA = w.partial.p;
    double* partB = w.partial.p + nbS;
    const Groups G = groups_of(w);
    double* pb[2] = {w.pcg_p.p, w.pcg_p2.p};
    for (int k = 0; k < nb; k++) {
        {
            // H blocks, their column indices, z and p_old read, p_new and y written
            ProfScope ps(w, PK_SPMV, prof ? 1152.0 * w.nnzb + 4.0 * w.nnzb + 4.0 * 96.0 * B : 0.0);
            hipLaunchKernelGGL(k_pcg_spmv, dim3(nbS), dim3(64 * SP_ROWS), 0, w.stream, B, w.pcg_k.p, k, rtol2, w.grp_n, w.row_ptr.p,
                               w.col_idx.p, w.Hval.p, w.pcg_z.p, pb[k & 1], pb[(k & 1) ^ 1], w.pcg_Ap.p, partB, partA, w.scal.p,
                               w.S > 1 ? (const int*)w.pcg_brun.p : (const int*)nullptr);
        }
        {
            // group inverses read; p, y read; x, r read and written; z written
            ProfScope ps(w, PK_PCG_UPDATE, prof ? 8.0 * (double)w.h_grp_moff.back() + 7.0 * 96.0 * B : 0.0);
            hipLaunchKernelGGL(k_pcg_update, dim3(B), dim3(UPD_T), 0, w.stream, B, w.pcg_k.p, k, nbS, G, pb[(k & 1) ^ 1],
                               w.pcg_Ap.p, w.pcg_xo, w.pcg_r.p, w.pcg_z.p, partA, partB, w.scal.p);
        }
        if (w.S > 1)
            hipLaunchKernelGGL(k_pcg_scene_stop, dim3(1), dim3(1024), 0, w.stream, w.S, B, w.sc_goff.p, rtol2, 0, partB, w.pcg_sbb.p,
                               w.scal.p, w.pcg_brun.p);
    }
    hipLaunchKernelGGL(k_pcg_advance, dim3(1), dim3(64), 0, w.stream, w.pcg_k.p, nb);
}

int pcg_solve(World& w, double* dq_out, int* iters, const std::function<void()>& after, bool* after_valid) {
    if (after_valid) *after_valid = false;
    const int B = w.B, n = 12 * B;
    const int nbS = (B + SP_ROWS - 1) / SP_ROWS;
    // the iterate x lives in the caller's output (12 B doubles)
    w.pcg_xo = dq_out;
    w.pcg_r.reserve(n); w.pcg_z.reserve(n); w.pcg_p.reserve(n); w.pcg_p2.reserve(n); w.pcg_Ap.reserve(n);
    w.partial.reserve((size_t)nbS + 2 * (size_t)B); w.scal.reserve(S_COUNT); w.pcg_k.reserve(1);
    double* partB = w.partial.p + nbS;
    if (!w.h_scal) ADB_HIP(hipHostMalloc((void**)&w.h_scal, 2 * S_COUNT * sizeof(double), hipHostMallocDefault));
    for (hipEvent_t& e : w.pcg_ev)
        if (!e) ADB_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    hipLaunchKernelGGL(k_pcg_init, dim3(B), dim3(64), 0, w.stream, B, groups_of(w), w.grad.p, w.pcg_xo, w.pcg_r.p, w.pcg_z.p,
                       w.pcg_p.p, partB, w.scal.p, w.pcg_k.p);
    ADB_HIP(hipGetLastError());
    if (w.S > 1) {
        w.pcg_sbb.reserve(w.S); w.pcg_brun.reserve(1);
        hipLaunchKernelGGL(k_pcg_scene_stop, dim3(1), dim3(1024), 0, w.stream, w.S, B, w.sc_goff.p,
                           w.P.pcg_rtol * w.P.pcg_rtol, 1, partB, w.pcg_sbb.p, w.scal.p, w.pcg_brun.p);
        ADB_HIP(hipGetLastError());
    }
    // the batch graph is captured once per set of buffer addresses (not while
    // per-kernel profiling is on, nor with ADB_PCG_GRAPH=0 for tools that do
    // not support graph capture)
    const bool use_graph = !w.prof.on && w.use_graphs;
    if (use_graph) {
        const void* key[] = {w.row_ptr.p, w.col_idx.p, w.Hval.p, w.Ginv.p, w.pcg_xo, w.pcg_r.p, w.pcg_z.p, w.pcg_p.p,
                             w.pcg_p2.p, w.pcg_Ap.p, w.partial.p, w.scal.p, w.pcg_k.p, w.grp_off, w.grp_body, w.grp_moff,
                             w.grp_n, w.sc_goff.p, w.pcg_sbb.p, w.pcg_brun.p};
        static_assert(sizeof key / sizeof key[0] == World::PCG_GRAPH_KEYS, "graph key size");
        const int BATCH = w.tune.pcg_batch;
        bool same = w.pcg_graph != nullptr && w.pcg_graph_batch == BATCH;
        for (int i = 0; i < World::PCG_GRAPH_KEYS; i++) same = same && w.pcg_graph_key[i] == key[i];
        if (!same) {
            if (w.pcg_graph) { ADB_HIP(hipGraphExecDestroy(w.pcg_graph)); w.pcg_graph = nullptr; }
            hipGraph_t g;
            ADB_HIP(hipStreamBeginCapture(w.stream, hipStreamCaptureModeThreadLocal));
            enqueue_batch(w, BATCH, false);
            ADB_HIP(hipStreamEndCapture(w.stream, &g));
            ADB_HIP(hipGraphInstantiate(&w.pcg_graph, g, nullptr, nullptr, 0));
            ADB_HIP(hipGraphDestroy(g));
            for (int i = 0; i < World::PCG_GRAPH_KEYS; i++) w.pcg_graph_key[i] = key[i];
            w.pcg_graph_batch = BATCH;
        }
    }
    const int BATCH = w.tune.pcg_batch;
    int it = 0, slot = 0;
    bool have_prev = false;
    const int max_it = std::max(w.P.pcg_max_iter, 0);
    // stop test after the last queued update: the spmv of the next iteration
    // tests it first and returns at once when stopped; if not, it does that
    // iteration's spmv, which the next batch's first spmv repeats identically
    // (same iteration, same p-buffer parity: batches are even)
    auto stop_test = [&]() {
        const double rtol2 = w.P.pcg_rtol * w.P.pcg_rtol;
        double* partA = w.partial.p;
        hipLaunchKernelGGL(k_pcg_spmv, dim3(nbS), dim3(64 * SP_ROWS), 0, w.stream, B, w.pcg_k.p, 0, rtol2, w.grp_n, w.row_ptr.p,
                           w.col_idx.p, w.Hval.p, w.pcg_z.p, w.pcg_p.p, w.pcg_p2.p, w.pcg_Ap.p, partB, partA, w.scal.p,
                           w.S > 1 ? (const int*)w.pcg_brun.p : (const int*)nullptr);
        ADB_HIP(hipGetLastError());
    };
    // Near the iteration count of the previous solve (Newton steps change the
    // system little) the batches shrink to 4 iterations, so the stop test,
    // read one batch late, leaves few queued iterations that return at once.
    const int pred = w.pcg_prev_its;
    bool done = false;
    if (pred > 0 && pred <= 4 && max_it > 0) {
        // a short previous solve (the light phase of a scene, where groups are
        // whole components): queue that many iterations (rounded up to even)
        // and test at once, instead of queueing batches ahead of the test
        const int nb0 = std::min(pred + (pred & 1), max_it);
        enqueue_batch(w, nb0, w.prof.on);
        it += nb0;
        stop_test();
        ADB_HIP(hipMemcpyAsync(w.h_scal, w.scal.p, S_COUNT * sizeof(double), hipMemcpyDeviceToHost, w.stream));
        // the caller's next work (on x as it stands) is queued ahead of the
        // wait; it is valid when these iterations finished the solve
        if (after) after();
        wait_stream(w);
        done = w.h_scal[S_STOP] != RUN;
        if (after_valid) *after_valid = done && after;
    }
    while (!done && it < max_it) {
        int nb = std::min(BATCH, max_it - it);
        if (pred > 0 && it + BATCH > pred - 8) nb = std::min(nb, 4);
        if (use_graph && nb == BATCH) ADB_HIP(hipGraphLaunch(w.pcg_graph, w.stream));
        else enqueue_batch(w, nb, w.prof.on);
        it += nb;
        ADB_HIP(hipGetLastError());
        ADB_HIP(hipMemcpyAsync(w.h_scal + S_COUNT * slot, w.scal.p, S_COUNT * sizeof(double), hipMemcpyDeviceToHost, w.stream));
        ADB_HIP(hipEventRecord(w.pcg_ev[slot], w.stream));
        if (have_prev) {
            // the batch before this one: stopped already?
            wait_event(w, w.pcg_ev[slot ^ 1]);
            if (w.h_scal[S_COUNT * (slot ^ 1) + S_STOP] != RUN) break;
        }
        have_prev = true;
        slot ^= 1;
    }
    if (!done) {
        stop_test();
        ADB_HIP(hipMemcpyAsync(w.h_scal, w.scal.p, S_COUNT * sizeof(double), hipMemcpyDeviceToHost, w.stream));
        wait_stream(w);
    }
    const double* h = w.h_scal;
    const int stop = (int)h[S_STOP];
    w.pcg_prev_its = stop == RUN ? it : (int)h[S_ITERS];
    if (iters) *iters = w.pcg_prev_its;
    if (stop == NONFINITE) return ADB_E_NAN;
    if (stop == RUN) return ADB_E_NOT_CONVERGED;
    if (stop == BREAKDOWN) return ADB_E_SOLVER;
    return ADB_OK;
}

} // namespace adb
