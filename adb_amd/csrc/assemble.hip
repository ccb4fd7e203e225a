// assemble.hip - per-body inertia terms and block-sparse Hessian assembly.
//
// Replaces grad_residue_per_body / hess_inertia_per_body (model/mutils.cpp:53-76,
// model/othogonal_energy.cpp:16-47), gen_empty_sm (model/sparse.cpp:13-57), the
// locked scatter of IPC::output_hessian_gradient (model/assemble_matrices.cpp:28-59)
// and the diagonal += c.hess of ABD::solve (model/abd.cpp:328-341).
//
// No atomics: every contribution (4 per PT/EE contact, 1 marker per body) gets
// a 64-bit key (row << 32 | col); a stable radix sort groups equal keys in a
// fixed order, unique keys become the blocks of a BSR matrix (12x12 blocks,
// column-major as the reference's CSC values), and one workgroup per block
// sums its contributions in sorted order (segmented reduction).  Static
// bodies (mass < 0) follow the reference: contributions to them are dropped,
// their diagonal block is I and their gradient 0.
#include <hipcub/hipcub.hpp>
#include "dmath.h"
#include "world.h"

namespace adb {

// inertia + orthogonality per body (H column-major)
__global__ void k_inertia(int B, double dt, double kap, double ga0, double ga1, double ga2, const double* __restrict__ q,
                          const double* __restrict__ q0, const double* __restrict__ qdot, const double* __restrict__ mass,
                          const double* __restrict__ Ic, double* __restrict__ bodyG, double* __restrict__ bodyH) {
    int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= B) return;
    double* G = bodyG + 12 * (size_t)b;
    double* H = bodyH + 144 * (size_t)b;
    if (mass[b] < 0.0) {
        for (int i = 0; i < 12; i++) G[i] = 0.0;
        for (int i = 0; i < 144; i++) H[i] = (i % 13 == 0) ? 1.0 : 0.0;
        return;
    }
    const double* Q = q + 12 * (size_t)b;
    d3 a[4] = {ld3(Q), ld3(Q + 3), ld3(Q + 6), ld3(Q + 9)};
    double go[12];
    go[0] = go[1] = go[2] = 0.0;
    for (int i = 1; i < 4; i++) {
        double s0 = 0, s1 = 0, s2 = 0;
        for (int j = 1; j < 4; j++) {
            double s = dot(a[i], a[j]) - (i == j ? 1.0 : 0.0);
            s0 += s * a[j].x; s1 += s * a[j].y; s2 += s * a[j].z;
        }
        go[3 * i] = 4 * kap * s0; go[3 * i + 1] = 4 * kap * s1; go[3 * i + 2] = 4 * kap * s2;
    }
    const double m = mass[b], ic = Ic[b];
    for (int i = 0; i < 12; i++) {
        double qt = q0[12 * (size_t)b + i] + dt * qdot[12 * (size_t)b + i];
        if (i < 3) qt += dt * dt * (i == 0 ? ga0 : (i == 1 ? ga1 : ga2));
        G[i] = dt * dt * go[i] + (i < 3 ? m : ic) * (Q[i] - qt);
    }
    for (int i = 0; i < 144; i++) H[i] = 0.0;
    for (int i = 0; i < 12; i++) H[13 * i] = i < 3 ? m : ic;
    for (int i = 1; i < 4; i++)
        for (int j = 1; j < 4; j++)
            for (int r = 0; r < 3; r++)
                for (int c = 0; c < 3; c++) {
                    double h;
                    if (i == j) {
                        h = 2 * cmp(a[i], r) * cmp(a[i], c) + (r == c ? dot(a[i], a[i]) - 1 : 0.0);
                        for (int k = 1; k < 4; k++) if (k != i) h += cmp(a[k], r) * cmp(a[k], c);
                    } else {
                        h = (r == c ? dot(a[j], a[i]) : 0.0) + cmp(a[j], r) * cmp(a[i], c);
                    }
                    H[(3 * j + c) * 12 + 3 * i + r] += h * 4 * kap * dt * dt;
                }
}

// contribution keys: B diagonal markers first, then 4 per PT/EE contact
__global__ void k_contrib_keys(int B, int n_pt, int n_ee, const int4* __restrict__ pt, const int4* __restrict__ ee,
                               unsigned long long* __restrict__ keys, int* __restrict__ vals) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int nc = n_pt + n_ee;
    if (i < B) { keys[i] = ((unsigned long long)i << 32) | (unsigned long long)i; vals[i] = -1 - i; return; }
    i -= B;
    if (i >= 4 * nc) return;
    const int c = i / 4, slot = i % 4;
    int4 t = c < n_pt ? pt[c] : ee[c - n_pt];
    const int I = t.x, J = t.z;
    int r, col;
    switch (slot) {
    case 0: r = I; col = I; break;
    case 1: r = J; col = J; break;
    case 2: r = I; col = J; break;
    default: r = J; col = I; break;
    }
    keys[B + i] = ((unsigned long long)r << 32) | (unsigned long long)col;
    vals[B + i] = i;
}
__global__ void k_flags(int n, const unsigned long long* __restrict__ keys, int* __restrict__ flag) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) flag[i] = (i == 0 || keys[i] != keys[i - 1]) ? 1 : 0;
}
// blk_of = inclusive scan of flags - 1
__global__ void k_block_meta(int n, const unsigned long long* __restrict__ keys, const int* __restrict__ blk_of,
                             int* __restrict__ blk_start, int* __restrict__ row_ptr, int* __restrict__ col_idx,
                             int* __restrict__ diag_blk) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    if (i == 0 || keys[i] != keys[i - 1]) {
        const int k = blk_of[i] - 1;
        const int row = (int)(keys[i] >> 32), col = (int)(keys[i] & 0xffffffffull);
        blk_start[k] = i;
        col_idx[k] = col;
        if (row == col) diag_blk[row] = k;
        if (i == 0 || (int)(keys[i - 1] >> 32) != row) row_ptr[row] = k;
    }
}
__global__ void k_vg_start(int B, int n_vg, const int2* __restrict__ vg, int* __restrict__ start) {
    int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b > B) return;
    int lo = 0, hi = n_vg;      // first entry with body >= b
    while (lo < hi) { int mid = (lo + hi) / 2; if (vg[mid].x < b) lo = mid + 1; else hi = mid; }
    start[b] = lo;
}

struct ReduceArgs {
    int nnzb, n_pt, n_ee, B;
    const unsigned long long* keys; const int* vals; const int* blk_start;
    const int4* pt; const int4* ee; const double* mass;
    const double* rec; const double* bodyH; const double* bodyG;
    const int2* vg; const int* vg_start; const double* vgK; const int* voff; const double* xbar;
    double* Hval; double* grad;
};
// Pass 1: one 256-lane workgroup per chunk of at most RCH contributions of
// one block (long segments - the diagonal block of a body with thousands of
// contacts - are split so every workgroup has bounded work).  Wave w sums
// contributions w, w + 4, ... of the chunk (lane t: entries t, t + 64,
// t + 128 < 144), then the four wave sums are combined in a fixed order.
constexpr int RCH = 64;
constexpr int RW = 4;            // waves per reduction workgroup
__global__ void k_chunk_count(int nnzb, const int* __restrict__ blk_start, int* __restrict__ nch) {
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < nnzb) nch[k] = (blk_start[k + 1] - blk_start[k] + RCH - 1) / RCH;
}
// chunk_off = exclusive scan of nch (nnzb + 1 entries)
// (also stores the chunk count in counts[3] for the byte accounting)
__global__ void k_chunk_owner(int nnzb, const int* __restrict__ chunk_off, int* __restrict__ chunk_blk, int* __restrict__ counts) {
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < nnzb) for (int c = chunk_off[k]; c < chunk_off[k + 1]; c++) chunk_blk[c] = k;
    if (k == 0) counts[3] = chunk_off[nnzb];
}
// 144 block entries + 12 gradient entries of one wave, lane t owning t, t + 64, t + 128
struct Acc156 { double e[3]; double g; };
// fixed-order combine of the RW wave accumulators into wave 0
__device__ __forceinline__ void combine_waves(Acc156& A, double (*sh)[156]) {
    const int w = threadIdx.x / 64, t = threadIdx.x % 64;
    if (w > 0) {
        for (int u = 0; u < 3; u++) { int e = t + 64 * u; if (e < 144) sh[w - 1][e] = A.e[u]; }
        if (t < 12) sh[w - 1][144 + t] = A.g;
    }
    __syncthreads();
    if (w == 0) {
        for (int q = 0; q < RW - 1; q++) {
            for (int u = 0; u < 3; u++) { int e = t + 64 * u; if (e < 144) A.e[u] += sh[q][e]; }
            if (t < 12) A.g += sh[q][144 + t];
        }
    }
}
// LDS writes of one wave visible to its own lanes
__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}
constexpr int RG = 8;            // contributions staged in LDS per wave at once
// Each wave sums RCH / RW consecutive contributions of the chunk.  They are
// taken RG at a time: every lane reads the RG contribution ids and record
// headers (broadcast loads), the wave copies the RG records (or inertia
// blocks) into LDS with independent coalesced loads, then every lane expands
// its entries from LDS:
//   block(s1, s2)[(al, r), (be, c)] = sum_ab Z[a][s1][al] Z[b][s2][be] PK[(3a + r) D + 3b + c]
//                                   + G[s1][al] G[s2][be] F[3r + c]
__global__ __launch_bounds__(64 * RW) void k_reduce_chunks(ReduceArgs a, const int* __restrict__ chunk_off,
                                                           const int* __restrict__ chunk_blk, double* __restrict__ part) {
    __shared__ double sh[RW - 1][156];
    __shared__ double stage[RW][RG][160];
    __shared__ int meta[RW][RG][4];     // kind (0 skip, 1 contact, 2 inertia), D, side bits, contact / body id
    const int ch = blockIdx.x;
    if (ch >= chunk_off[a.nnzb]) return;
    const int k = chunk_blk[ch];
    const int beg = a.blk_start[k] + (ch - chunk_off[k]) * RCH;
    const int end = min(beg + RCH, a.blk_start[k + 1]);
    const int w = threadIdx.x / 64, t = threadIdx.x % 64;
    const int wb = min(beg + w * (RCH / RW), end), we = min(wb + RCH / RW, end);
    Acc156 A = {{0.0, 0.0, 0.0}, 0.0};
    double (*S)[160] = stage[w];
    int (*M)[4] = meta[w];
    for (int g0 = wb; g0 < we; g0 += RG) {
        const int ng = min(RG, we - g0);
        if (t < RG) {
            int kind = 0, D = 0, sides = 0, x = 0;
            if (t < ng) {
                const int v = a.vals[g0 + t];
                if (v < 0) { kind = 2; x = -1 - v; }
                else {
                    const int c = v / 4, slot = v % 4;
                    const int code = (int)a.rec[REC * (size_t)c + R_D];
                    const int Dd = code % 100, di = (code / 100) & 1, dj = (code / 100) >> 1;
                    const bool use = slot == 0 ? di : (slot == 1 ? dj : (di && dj));
                    if (use && Dd > 0) {
                        kind = 1; D = Dd; x = c;
                        // bit 0: s1, bit 1: s2, bit 2: gradient (slot 0 = (I, I), 1 = (J, J), 2 = (I, J), 3 = (J, I))
                        sides = ((slot == 1 || slot == 3) ? 1 : 0) | ((slot == 1 || slot == 2) ? 2 : 0) | (slot < 2 ? 4 : 0);
                    }
                }
            }
            M[t][0] = kind; M[t][1] = D; M[t][2] = sides; M[t][3] = x;
        }
        wave_sync();
        {
            // all loads first (independent), then the LDS stores
            double tmp[RG][3];
#pragma unroll
            for (int j = 0; j < RG; j++) {
                const int kind = j < ng ? M[j][0] : 0, x = M[j][3];
                const int len = kind == 1 ? R_PK + M[j][1] * M[j][1] : (kind == 2 ? 156 : 0);
#pragma unroll
                for (int u = 0; u < 3; u++) {
                    const int e = t + 64 * u;
                    double v = 0.0;
                    if (e < len) {
                        if (kind == 1) v = a.rec[REC * (size_t)x + e];
                        else v = e < 144 ? a.bodyH[144 * (size_t)x + e] : a.bodyG[12 * (size_t)x + e - 144];
                    }
                    tmp[j][u] = v;
                }
            }
#pragma unroll
            for (int j = 0; j < RG; j++)
#pragma unroll
                for (int u = 0; u < 3; u++) {
                    const int e = t + 64 * u;
                    if (e < 160) S[j][e] = tmp[j][u];
                }
        }
        wave_sync();
#pragma unroll 1
        for (int j = 0; j < ng; j++) {
            const int kind = M[j][0];
            const double* Rj = S[j];
            if (kind == 2) {
                for (int u = 0; u < 3; u++) { int e = t + 64 * u; if (e < 144) A.e[u] += Rj[e]; }
                if (t < 12) A.g += Rj[144 + t];
            } else if (kind == 1) {
                const int Dj = M[j][1], nA = Dj / 3, sides = M[j][2], a1 = sides & 1, a2 = (sides >> 1) & 1;
                for (int u = 0; u < 3; u++) {
                    const int e = t + 64 * u;
                    if (e < 144) {
                        // column (be, cc), row (al, r)
                        const int col = e / 12, row = e % 12;
                        const int be = col / 3, cc = col % 3, al = row / 3, r = row % 3;
                        double val = Rj[R_G + 4 * a1 + al] * Rj[R_G + 4 * a2 + be] * Rj[R_F + 3 * r + cc];
                        for (int aa = 0; aa < nA; aa++) {
                            double sb = 0.0;
                            for (int bb = 0; bb < nA; bb++)
                                sb += Rj[R_Z + 8 * bb + 4 * a2 + be] * Rj[R_PK + (3 * aa + r) * Dj + 3 * bb + cc];
                            val += Rj[R_Z + 8 * aa + 4 * a1 + al] * sb;
                        }
                        A.e[u] += val;
                    }
                }
                if ((sides & 4) && t < 12) A.g += Rj[R_GRAD + 12 * a1 + t];
            }
        }
        wave_sync();
    }
    combine_waves(A, sh);
    if (w == 0) {
        double* P = part + 156 * (size_t)ch;
        for (int u = 0; u < 3; u++) { int e = t + 64 * u; if (e < 144) P[e] = A.e[u]; }
        if (t < 12) P[144 + t] = A.g;
    }
}

// Pass 2: one workgroup per block sums its chunk partials (wave w: chunks
// w, w + 4, ..., then a fixed-order combine), then the body's vertex-ground
// terms for diagonal blocks.
__global__ __launch_bounds__(64 * RW) void k_reduce_final(ReduceArgs a, const int* __restrict__ chunk_off, const double* __restrict__ part) {
    __shared__ double sh[RW - 1][156];
    const int k = blockIdx.x;
    if (k >= a.nnzb) return;
    const unsigned long long key = a.keys[a.blk_start[k]];
    const int row = (int)(key >> 32), col = (int)(key & 0xffffffffull);
    const int w = threadIdx.x / 64, t = threadIdx.x % 64;
    Acc156 A = {{0.0, 0.0, 0.0}, 0.0};
    for (int c = chunk_off[k] + w; c < chunk_off[k + 1]; c += RW) {
        const double* P = part + 156 * (size_t)c;
        for (int u = 0; u < 3; u++) { int e = t + 64 * u; if (e < 144) A.e[u] += P[e]; }
        if (t < 12) A.g += P[144 + t];
    }
    combine_waves(A, sh);
    if (w != 0) return;
    double acc[3] = {A.e[0], A.e[1], A.e[2]};
    double gacc = A.g;
    if (row == col && a.mass[row] > 0.0) {
        // vertex-ground terms of this body: (w w^T) (x) K, gradient w (x) gx
        for (int s = a.vg_start[row]; s < a.vg_start[row + 1]; s++) {
            const int v = a.voff[row] + a.vg[s].y;
            const double w4[4] = {1.0, a.xbar[3 * (size_t)v], a.xbar[3 * (size_t)v + 1], a.xbar[3 * (size_t)v + 2]};
            const double* K = a.vgK + 12 * (size_t)s;
            for (int u = 0; u < 3; u++) {
                int e = t + 64 * u;
                if (e < 144) {
                    int cc = e / 12, rr = e % 12;
                    acc[u] += w4[rr / 3] * w4[cc / 3] * K[3 * (cc % 3) + (rr % 3)];
                }
            }
            if (t < 12) gacc += w4[t / 3] * K[9 + t % 3];
        }
    }
    for (int u = 0; u < 3; u++) { int e = t + 64 * u; if (e < 144) a.Hval[144 * (size_t)k + e] = acc[u]; }
    if (row == col && t < 12) a.grad[12 * (size_t)row + t] = gacc;
}

// block-Jacobi preconditioner: the explicit inverse of each diagonal block,
// M^-1 = W^T W with W = L^-1 from the Cholesky factor L (symmetric by
// construction).  One thread per body, matrices staged in LDS; falls back to
// the inverse diagonal when the block is not positive definite.
constexpr int CH_T = 32;
__global__ __launch_bounds__(CH_T) void k_block_inv(int B, const int* __restrict__ diag_blk, const double* __restrict__ Hval,
                                                   double* __restrict__ Minv) {
    __shared__ double S[CH_T][144];
    __shared__ double T[CH_T][144];
    const int b = blockIdx.x * CH_T + threadIdx.x;
    if (b >= B) return;
    double* A = S[threadIdx.x];   // A[c * 12 + r]: column-major; L(r, c) for r >= c after factorisation
    double* W = T[threadIdx.x];   // W[c * 12 + r] = (L^-1)(r, c), r >= c
    const double* D = Hval + 144 * (size_t)diag_blk[b];
    for (int i = 0; i < 144; i++) A[i] = D[i];
    bool ok = true;
    for (int j = 0; j < 12 && ok; j++) {
        double s = A[j * 12 + j];
        for (int k = 0; k < j; k++) s -= A[k * 12 + j] * A[k * 12 + j];
        if (!(s > 0.0)) { ok = false; break; }
        double l = sqrt(s);
        A[j * 12 + j] = l;
        for (int i = j + 1; i < 12; i++) {
            double t = A[j * 12 + i];
            for (int k = 0; k < j; k++) t -= A[k * 12 + i] * A[k * 12 + j];
            A[j * 12 + i] = t / l;
        }
    }
    double* out = Minv + 144 * (size_t)b;
    if (ok) {
        for (int c = 0; c < 12; c++) {
            W[c * 12 + c] = 1.0 / A[c * 12 + c];
            for (int r = c + 1; r < 12; r++) {
                double s = 0.0;
                for (int k = c; k < r; k++) s += A[k * 12 + r] * W[c * 12 + k];
                W[c * 12 + r] = -s / A[r * 12 + r];
            }
        }
        for (int c = 0; c < 12; c++)
            for (int r = c; r < 12; r++) {
                double s = 0.0;
                for (int k = r; k < 12; k++) s += W[r * 12 + k] * W[c * 12 + k];
                out[c * 12 + r] = s;
                out[r * 12 + c] = s;
            }
    } else {
        for (int i = 0; i < 144; i++) out[i] = 0.0;
        for (int i = 0; i < 12; i++) { double d = D[13 * i]; out[13 * i] = 1.0 / (d > 0.0 ? d : 1.0); }
    }
}

template <class K, class V>
static void sort_pairs(World& w, K* kin, K* kout, V* vin, V* vout, int n, int end_bit) {
    size_t tmp = 0;
    ADB_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, tmp, kin, kout, vin, vout, n, 0, end_bit, w.stream));
    w.sort_tmp.reserve(tmp + 16);
    ADB_HIP(hipcub::DeviceRadixSort::SortPairs(w.sort_tmp.p, tmp, kin, kout, vin, vout, n, 0, end_bit, w.stream));
}

void assemble_system(World& w) {
    const int B = w.B;
    w.bodyG.reserve(12 * (size_t)B);
    w.bodyH.reserve(144 * (size_t)B);
    hipLaunchKernelGGL(k_inertia, dim3((B + 127) / 128), dim3(128), 0, w.stream, B, w.P.dt, w.P.kappa_ortho, w.P.gravity[0], w.P.gravity[1], w.P.gravity[2],
                       w.q.p, w.q0.p, w.qdot.p, w.mass.p, w.Ic.p, w.bodyG.p, w.bodyH.p);
    ADB_HIP(hipGetLastError());
    const int nc = w.cs.n_pt + w.cs.n_ee;
    const int n = B + 4 * nc;
    w.n_contrib = n;
    w.akeys.reserve(n); w.akeys_s.reserve(n); w.aval.reserve(n); w.aval_s.reserve(n);
    w.uniq_flag.reserve(n); w.blk_of.reserve(n);
    hipLaunchKernelGGL(k_contrib_keys, dim3((n + 255) / 256), dim3(256), 0, w.stream, B, w.cs.n_pt, w.cs.n_ee, w.cs.pt.p, w.cs.ee.p,
                       w.akeys.p, w.aval.p);
    ADB_HIP(hipGetLastError());
    int bb = 1; while ((1ll << bb) <= B) bb++;
    sort_pairs(w, w.akeys.p, w.akeys_s.p, w.aval.p, w.aval_s.p, n, 32 + bb);
    hipLaunchKernelGGL(k_flags, dim3((n + 255) / 256), dim3(256), 0, w.stream, n, w.akeys_s.p, w.uniq_flag.p);
    ADB_HIP(hipGetLastError());
    {
        size_t tmp = 0;
        ADB_HIP(hipcub::DeviceScan::InclusiveSum(nullptr, tmp, w.uniq_flag.p, w.blk_of.p, n, w.stream));
        w.sort_tmp.reserve(tmp + 16);
        ADB_HIP(hipcub::DeviceScan::InclusiveSum(w.sort_tmp.p, tmp, w.uniq_flag.p, w.blk_of.p, n, w.stream));
    }
    int last = 0;
    ADB_HIP(hipMemcpyAsync(&last, w.blk_of.p + n - 1, sizeof(int), hipMemcpyDeviceToHost, w.stream));
    ADB_HIP(hipStreamSynchronize(w.stream));
    const int nnzb = last;   // inclusive scan: count of unique keys
    w.nnzb = nnzb;
    w.blk_start.reserve(nnzb + 1); w.col_idx.reserve(nnzb); w.row_ptr.reserve(B + 1); w.diag_blk.reserve(B);
    w.Hval.reserve(144 * (size_t)nnzb);
    // blk_of currently holds (inclusive count); shift to 0-based block ids in k_block_meta via -1
    hipLaunchKernelGGL(k_block_meta, dim3((n + 255) / 256), dim3(256), 0, w.stream, n, w.akeys_s.p, w.blk_of.p,
                       w.blk_start.p, w.row_ptr.p, w.col_idx.p, w.diag_blk.p);
    ADB_HIP(hipGetLastError());
    ADB_HIP(hipMemcpyAsync(w.blk_start.p + nnzb, &w.n_contrib, sizeof(int), hipMemcpyHostToDevice, w.stream));
    ADB_HIP(hipMemcpyAsync(w.row_ptr.p + B, &w.nnzb, sizeof(int), hipMemcpyHostToDevice, w.stream));
    w.vg_start.reserve(B + 1);
    hipLaunchKernelGGL(k_vg_start, dim3((B + 1 + 255) / 256), dim3(256), 0, w.stream, B, w.cs.n_vg, w.cs.vg.p, w.vg_start.p);
    ADB_HIP(hipGetLastError());
    w.grad.reserve(12 * (size_t)B);
    ReduceArgs a;
    a.nnzb = nnzb; a.n_pt = w.cs.n_pt; a.n_ee = w.cs.n_ee; a.B = B;
    a.keys = w.akeys_s.p; a.vals = w.aval_s.p; a.blk_start = w.blk_start.p;
    a.pt = w.cs.pt.p; a.ee = w.cs.ee.p; a.mass = w.mass.p;
    a.rec = w.crec.p; a.bodyH = w.bodyH.p; a.bodyG = w.bodyG.p;
    a.vg = w.cs.vg.p; a.vg_start = w.vg_start.p; a.vgK = w.vgK.p; a.voff = w.voff.p; a.xbar = w.xbar.p;
    a.Hval = w.Hval.p; a.grad = w.grad.p;
    {
        // chunk layout of the segments
        w.chunk_n.reserve(nnzb + 1); w.chunk_off.reserve(nnzb + 1);
        hipLaunchKernelGGL(k_chunk_count, dim3((nnzb + 255) / 256), dim3(256), 0, w.stream, nnzb, w.blk_start.p, w.chunk_n.p);
        ADB_HIP(hipMemsetAsync(w.chunk_n.p + nnzb, 0, sizeof(int), w.stream));
        size_t tmp = 0;
        ADB_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, tmp, w.chunk_n.p, w.chunk_off.p, nnzb + 1, w.stream));
        w.sort_tmp.reserve(tmp + 16);
        ADB_HIP(hipcub::DeviceScan::ExclusiveSum(w.sort_tmp.p, tmp, w.chunk_n.p, w.chunk_off.p, nnzb + 1, w.stream));
        // each block has at most len / RCH + 1 chunks: launch for the bound, the
        // kernel reads the exact count
        const int max_chunks = nnzb + n / RCH;
        w.chunk_blk.reserve(max_chunks);
        w.chunk_part.reserve(156 * (size_t)max_chunks);
        hipLaunchKernelGGL(k_chunk_owner, dim3((nnzb + 255) / 256), dim3(256), 0, w.stream, nnzb, w.chunk_off.p, w.chunk_blk.p,
                           w.ccount.p);
        ADB_HIP(hipGetLastError());
        {
            // every contact record read once (D = 3: 59 doubles, 6: 94, 9: 147), the
            // contribution list and inertia terms read, one partial block written per chunk
            ProfScope ps(w, PK_ASM_REDUCE, 24.0 * n + 1248.0 * B);
            ps.counts(prof_snapshot(w, w.ccount.p), 472.0, 752.0, 1176.0, 1248.0);
            hipLaunchKernelGGL(k_reduce_chunks, dim3(max_chunks), dim3(64 * RW), 0, w.stream, a, w.chunk_off.p, w.chunk_blk.p,
                               w.chunk_part.p);
        }
        hipLaunchKernelGGL(k_reduce_final, dim3(nnzb), dim3(64 * RW), 0, w.stream, a, w.chunk_off.p, w.chunk_part.p);
        ADB_HIP(hipGetLastError());
    }
    ADB_HIP(hipGetLastError());
    w.Pinv.reserve(144 * (size_t)B);
    hipLaunchKernelGGL(k_block_inv, dim3((B + CH_T - 1) / CH_T), dim3(CH_T), 0, w.stream, B, w.diag_blk.p, w.Hval.p, w.Pinv.p);
    ADB_HIP(hipGetLastError());
}

} // namespace adb
