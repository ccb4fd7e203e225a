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
#include <algorithm>
#include <vector>
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
// value of a contact contribution: slot (bits 0-1), number of Helmert rows
// nA = D / 3 (bits 2-3), used (bit 4: the block's bodies are dynamic and the
// contact active), contact id (bits 5+); body markers are -1 - body
__global__ void k_contrib_keys(int B, int n_pt, int n_ee, const int4* __restrict__ pt, const int4* __restrict__ ee,
                               const int* __restrict__ ccls, const double* __restrict__ mass,
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
    const bool di = mass[I] > 0.0, dj = mass[J] > 0.0;
    const bool use = (slot == 0 ? di : (slot == 1 ? dj : (di && dj))) && ccls[c] > 0;
    vals[B + i] = (c << 5) | ((use ? 1 : 0) << 4) | (ccls[c] << 2) | slot;
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
    const int* row_ptr; const int* col_idx;
    double* Hval; double* grad;
};
// Pass 1: one 256-lane workgroup per chunk of at most RCH contributions of
// one block (long segments - the diagonal block of a body with thousands of
// contacts - are split so every workgroup has bounded work).
constexpr int RW = 4;            // waves per reduction workgroup
constexpr int RS = 5;            // contributions per wave at once (12 lanes each)
constexpr int RCH = RW * 3 * RS; // contributions per chunk
// Blocks below the diagonal get no chunks: block (J, I) = block (I, J)^T is
// written by its upper partner in k_reduce_final, so each off-diagonal pair
// is reduced once.
__global__ void k_chunk_count(int nnzb, const unsigned long long* __restrict__ keys, const int* __restrict__ blk_start,
                              int* __restrict__ nch) {
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nnzb) return;
    const unsigned long long key = keys[blk_start[k]];
    const bool lower = (key >> 32) > (key & 0xffffffffull);
    nch[k] = lower ? 0 : (blk_start[k + 1] - blk_start[k] + RCH - 1) / RCH;
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
// Pass 1 body: each wave takes RCH / RW consecutive contributions of the
// chunk, RS at a time; 12 lanes per contribution, lane (slot, col) owning
// column col = (be, cc) of the 12x12 block:
//   t[(a, r)]       = sum_b PK[(3a + r) D + 3b + cc] Z[b][s2][be]
//   entry (al, r)  += sum_a Z[a][s1][al] t[(a, r)] + G[s1][al] G[s2][be] F[3r + cc]
// (lane col 0 of each slot also adds the 12 gradient entries).  The RW x RS
// slot sums are combined in a fixed order.
__global__ __launch_bounds__(64 * RW) void k_reduce_chunks(ReduceArgs a, const int* __restrict__ chunk_off,
                                                           const int* __restrict__ chunk_blk, double* __restrict__ part) {
    __shared__ double P[RW][RS][156];
    // XCD-aware order: workgroups are dealt to the 8 XCDs round-robin, so XCD
    // x gets the contiguous chunk range [x per, (x + 1) per) - the blocks of
    // the same rows, whose contributions share contact records, meet in one L2
    const int per = gridDim.x / 8;
    const int ch = (blockIdx.x % 8) * per + blockIdx.x / 8;
    if (ch >= chunk_off[a.nnzb]) return;
    const int k = chunk_blk[ch];
    const int beg = a.blk_start[k] + (ch - chunk_off[k]) * RCH;
    const int end = min(beg + RCH, a.blk_start[k + 1]);
    const int w = threadIdx.x / 64, l = threadIdx.x % 64;
    const int slot = l / 12, col = l % 12, be = col / 3, cc = col % 3;
    const int wb = min(beg + w * (RCH / RW), end), we = min(wb + RCH / RW, end);
    double acc[12], gacc = 0.0;
#pragma unroll
    for (int i = 0; i < 12; i++) acc[i] = 0.0;
    for (int r0 = wb; r0 < we; r0 += RS) {
        const int s = r0 + slot;
        if (slot < RS && s < we) {
            const int v = a.vals[s];
            if (v < 0) {     // diagonal marker: inertia block and gradient of body -1 - v
                const int b = -1 - v;
#pragma unroll
                for (int i = 0; i < 12; i++) acc[i] += a.bodyH[144 * (size_t)b + 12 * col + i];
                gacc += a.bodyG[12 * (size_t)b + col];
            } else if ((v >> 4) & 1) {
                const int c = v >> 5, cs = v & 3, nA = (v >> 2) & 3, D = 3 * nA;
                const double* R = a.rec + REC * (size_t)c;
                {
                    // cs 0 = (I, I), 1 = (J, J), 2 = (I, J), 3 = (J, I)
                    const int s1 = (cs == 1 || cs == 3) ? 1 : 0, s2 = (cs == 1 || cs == 2) ? 1 : 0;
                    double t[3][3];
#pragma unroll
                    for (int aa = 0; aa < 3; aa++)
#pragma unroll
                        for (int r = 0; r < 3; r++) {
                            double sum = 0.0;
#pragma unroll
                            for (int bb = 0; bb < 3; bb++)
                                if (aa < nA && bb < nA) sum += R[R_PK + (3 * aa + r) * D + 3 * bb + cc] * R[R_Z + 8 * bb + 4 * s2 + be];
                            t[aa][r] = sum;
                        }
                    const double g2 = R[R_G + 4 * s2 + be];
                    const double F0 = R[R_F + cc], F1 = R[R_F + 3 + cc], F2 = R[R_F + 6 + cc];
#pragma unroll
                    for (int al = 0; al < 4; al++) {
                        const double gg = R[R_G + 4 * s1 + al] * g2;
                        double z[3];
#pragma unroll
                        for (int aa = 0; aa < 3; aa++) z[aa] = aa < nA ? R[R_Z + 8 * aa + 4 * s1 + al] : 0.0;
                        acc[3 * al + 0] += (z[0] * t[0][0] + z[1] * t[1][0] + z[2] * t[2][0]) + gg * F0;
                        acc[3 * al + 1] += (z[0] * t[0][1] + z[1] * t[1][1] + z[2] * t[2][1]) + gg * F1;
                        acc[3 * al + 2] += (z[0] * t[0][2] + z[1] * t[1][2] + z[2] * t[2][2]) + gg * F2;
                    }
                    if (cs < 2) gacc += R[R_GRAD + 12 * s1 + col];
                }
            }
        }
    }
    if (slot < RS) {
#pragma unroll
        for (int i = 0; i < 12; i++) P[w][slot][12 * col + i] = acc[i];
        P[w][slot][144 + col] = gacc;
    }
    __syncthreads();
    for (int e = threadIdx.x; e < 156; e += 64 * RW) {
        double sum = 0.0;
        for (int q = 0; q < RW; q++)
            for (int u = 0; u < RS; u++) sum += P[q][u][e];
        part[156 * (size_t)ch + e] = sum;
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
    if (row > col) return;   // written transposed by block (col, row)
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
    if (row < col) {
        // the mirror block (col, row) in row col of the BSR structure
        int lo = a.row_ptr[col], hi = a.row_ptr[col + 1];
        while (lo < hi) { const int mid = (lo + hi) / 2; if (a.col_idx[mid] < row) lo = mid + 1; else hi = mid; }
        if (lo < a.row_ptr[col + 1] && a.col_idx[lo] == row) {   // always: the structure is symmetric
            double* M = a.Hval + 144 * (size_t)lo;
            for (int u = 0; u < 3; u++) {
                const int e = t + 64 * u;
                if (e < 144) M[12 * (e % 12) + e / 12] = acc[u];
            }
        }
    }
}

// Preconditioner: the explicit inverse of the diagonal sub-matrix of each
// body group (a connected component of the contact graph, or a BFS-ordered
// piece of one with at most GMAX bodies).  On groups that are whole
// components the preconditioned system decouples exactly, so PCG needs only a
// few iterations where the contact graph is a union of small clusters; in
// large clusters it is block-Jacobi with blocks of up to GMAX bodies.
// One workgroup per group: gather the blocks into LDS, Cholesky in place,
// W = L^-1 column by column (thread per column, stored transposed in the
// upper triangle), then M^-1 = W^T W.  A group whose matrix is not positive
// definite falls back to the inverse diagonal.
constexpr int GMAX = 8;
// GM: largest group size of the launch (1 for single bodies, GMAX otherwise);
// list: the groups of this launch
template <int GM>
__global__ __launch_bounds__(256) void k_group_inv(const int* __restrict__ list, const int* __restrict__ grp_off,
                                                   const int* __restrict__ grp_body, const long long* __restrict__ grp_moff,
                                                   const int* __restrict__ row_ptr, const int* __restrict__ col_idx,
                                                   const double* __restrict__ Hval, double* __restrict__ Ginv) {
    constexpr int GN = 12 * GM, GLD = GN + 1;
    __shared__ double A[GN * GLD];
    __shared__ double wd[GN];
    __shared__ int blk[GM][GM];
    __shared__ int bod[GM];
    __shared__ int ok;
    const int g = list[blockIdx.x], t = threadIdx.x;
    const int m = grp_off[g + 1] - grp_off[g], n = 12 * m;
    if (t < m) bod[t] = grp_body[grp_off[g] + t];
    if (t == 0) ok = 1;
    __syncthreads();
    for (int e = t; e < m * m; e += blockDim.x) {
        const int ia = e / m, ib = e % m;
        int k = -1;
        for (int q = row_ptr[bod[ia]]; q < row_ptr[bod[ia] + 1]; q++)
            if (col_idx[q] == bod[ib]) { k = q; break; }
        blk[ia][ib] = k;
    }
    __syncthreads();
    for (int e = t; e < n * n; e += blockDim.x) {
        const int j = e / n, i = e % n;       // A[i][j], column-major in the block
        const int k = blk[i / 12][j / 12];
        A[i * GLD + j] = k >= 0 ? Hval[144 * (size_t)k + 12 * (j % 12) + (i % 12)] : 0.0;
    }
    __syncthreads();
    // Cholesky, lower triangle in place
    for (int j = 0; j < n; j++) {
        if (t == 0) {
            const double d = A[j * GLD + j];
            if (d > 0.0) A[j * GLD + j] = sqrt(d); else ok = 0;
        }
        __syncthreads();
        if (!ok) break;
        const double ljj = A[j * GLD + j];
        for (int i = j + 1 + t; i < n; i += blockDim.x) A[i * GLD + j] /= ljj;
        __syncthreads();
        const int r = n - j - 1;
        for (int e = t; e < r * r; e += blockDim.x) {
            const int i = j + 1 + e / r, k = j + 1 + e % r;
            if (k <= i) A[i * GLD + k] -= A[i * GLD + j] * A[k * GLD + j];
        }
        __syncthreads();
    }
    double* out = Ginv + grp_moff[g];
    if (!ok) {
        // the Cholesky stopped early; the original diagonal is gone, so take it from Hval
        for (int e = t; e < n * n; e += blockDim.x) {
            const int j = e / n, i = e % n;
            double v = 0.0;
            if (i == j) {
                const int k = blk[i / 12][i / 12];
                const double d = k >= 0 ? Hval[144 * (size_t)k + 13 * (i % 12)] : 1.0;
                v = 1.0 / (d > 0.0 ? d : 1.0);
            }
            out[(size_t)j * n + i] = v;
        }
        return;
    }
    // W = L^-1: thread c builds column c (W[i][c], i >= c) into A[c][i] (upper) and wd[c]
    for (int c = t; c < n; c += blockDim.x) {
        const double wcc = 1.0 / A[c * GLD + c];
        wd[c] = wcc;
        for (int i = c + 1; i < n; i++) {
            double s = A[i * GLD + c] * wcc;
            for (int k = c + 1; k < i; k++) s += A[i * GLD + k] * A[c * GLD + k];
            A[c * GLD + i] = -s / A[i * GLD + i];
        }
    }
    __syncthreads();
    // M^-1 = W^T W: (i, j) = sum_{k >= max(i, j)} W[k][i] W[k][j]
    for (int e = t; e < n * n; e += blockDim.x) {
        const int j = e / n, i = e % n;
        if (i > j) continue;
        double s = 0.0;
        for (int k = j; k < n; k++) {
            const double wki = k == i ? wd[i] : A[i * GLD + k];
            const double wkj = k == j ? wd[j] : A[j * GLD + k];
            s += wki * wkj;
        }
        out[(size_t)j * n + i] = s;
        out[(size_t)i * n + j] = s;
    }
}

// Connected components of the block graph (host, from the BSR structure) and
// their split into groups of at most GMAX bodies in BFS order.
static void build_groups(World& w) {
    const int B = w.B, nnzb = w.nnzb;
    w.h_rp.resize(B + 1); w.h_ci.resize(nnzb);
    ADB_HIP(hipMemcpyAsync(w.h_rp.data(), w.row_ptr.p, (B + 1) * sizeof(int), hipMemcpyDeviceToHost, w.stream));
    ADB_HIP(hipMemcpyAsync(w.h_ci.data(), w.col_idx.p, nnzb * sizeof(int), hipMemcpyDeviceToHost, w.stream));
    ADB_HIP(hipStreamSynchronize(w.stream));
    std::vector<char> seen(B, 0);
    std::vector<int> queue;
    queue.reserve(B);
    w.h_grp_off.assign(1, 0);
    w.h_grp_body.clear();
    w.h_grp_moff.assign(1, 0);
    for (int s = 0; s < B; s++) {
        if (seen[s]) continue;
        queue.clear();
        queue.push_back(s);
        seen[s] = 1;
        for (size_t h = 0; h < queue.size(); h++) {
            const int r = queue[h];
            for (int k = w.h_rp[r]; k < w.h_rp[r + 1]; k++) {
                const int c = w.h_ci[k];
                if (!seen[c]) { seen[c] = 1; queue.push_back(c); }
            }
        }
        for (size_t q = 0; q < queue.size(); q += GMAX) {
            const int m = (int)std::min<size_t>(GMAX, queue.size() - q);
            for (int i = 0; i < m; i++) w.h_grp_body.push_back(queue[q + i]);
            w.h_grp_off.push_back((int)w.h_grp_body.size());
            w.h_grp_moff.push_back(w.h_grp_moff.back() + (long long)(12 * m) * (12 * m));
        }
    }
    w.n_groups = (int)w.h_grp_off.size() - 1;
    // groups with several bodies first: their inverses take longest
    w.h_glist.clear();
    for (int g = 0; g < w.n_groups; g++) if (w.h_grp_off[g + 1] - w.h_grp_off[g] > 1) w.h_glist.push_back(g);
    for (int g = 0; g < w.n_groups; g++) if (w.h_grp_off[g + 1] - w.h_grp_off[g] == 1) w.h_glist.push_back(g);
    w.glist.reserve(std::max(w.n_groups, 1));
    ADB_HIP(hipMemcpyAsync(w.glist.p, w.h_glist.data(), w.n_groups * sizeof(int), hipMemcpyHostToDevice, w.stream));
    w.grp_off.reserve(w.n_groups + 1); w.grp_body.reserve(B); w.grp_moff.reserve(w.n_groups + 1); w.grp_n.reserve(1);
    ADB_HIP(hipMemcpyAsync(w.grp_n.p, &w.n_groups, sizeof(int), hipMemcpyHostToDevice, w.stream));
    ADB_HIP(hipMemcpyAsync(w.grp_off.p, w.h_grp_off.data(), (w.n_groups + 1) * sizeof(int), hipMemcpyHostToDevice, w.stream));
    ADB_HIP(hipMemcpyAsync(w.grp_body.p, w.h_grp_body.data(), B * sizeof(int), hipMemcpyHostToDevice, w.stream));
    ADB_HIP(hipMemcpyAsync(w.grp_moff.p, w.h_grp_moff.data(), (w.n_groups + 1) * sizeof(long long), hipMemcpyHostToDevice,
                           w.stream));
    // the host vectors stay alive (members of w) until the copies have run
    w.Ginv.reserve((size_t)w.h_grp_moff.back());
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
                       w.ccls.p, w.mass.p,
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
    a.row_ptr = w.row_ptr.p; a.col_idx = w.col_idx.p;
    a.Hval = w.Hval.p; a.grad = w.grad.p;
    {
        // chunk layout of the segments
        w.chunk_n.reserve(nnzb + 1); w.chunk_off.reserve(nnzb + 1);
        hipLaunchKernelGGL(k_chunk_count, dim3((nnzb + 255) / 256), dim3(256), 0, w.stream, nnzb, w.akeys_s.p, w.blk_start.p,
                           w.chunk_n.p);
        ADB_HIP(hipMemsetAsync(w.chunk_n.p + nnzb, 0, sizeof(int), w.stream));
        size_t tmp = 0;
        ADB_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, tmp, w.chunk_n.p, w.chunk_off.p, nnzb + 1, w.stream));
        w.sort_tmp.reserve(tmp + 16);
        ADB_HIP(hipcub::DeviceScan::ExclusiveSum(w.sort_tmp.p, tmp, w.chunk_n.p, w.chunk_off.p, nnzb + 1, w.stream));
        // each diagonal or upper block has at most len / RCH + 1 chunks (the
        // structure is symmetric with one diagonal block per body, so there are
        // B + (nnzb - B) / 2 of them): launch for the bound rounded up to the 8
        // XCDs, the kernel reads the exact count
        const int max_chunks = (B + (nnzb - B) / 2 + n / RCH + 7) / 8 * 8;
        w.chunk_blk.reserve(max_chunks);
        w.chunk_part.reserve(156 * (size_t)max_chunks);
        hipLaunchKernelGGL(k_chunk_owner, dim3((nnzb + 255) / 256), dim3(256), 0, w.stream, nnzb, w.chunk_off.p, w.chunk_blk.p,
                           w.ccount.p);
        ADB_HIP(hipGetLastError());
        if (w.side_pending) {   // records of the 4-point contacts (side stream)
            ADB_HIP(hipStreamWaitEvent(w.stream, w.ev_side[1], 0));
            w.side_pending = false;
        }
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
    build_groups(w);
    {
        ProfScope ps(w, PK_BLOCK_CHOL, 0.0);
        hipLaunchKernelGGL(k_group_inv<GMAX>, dim3(w.n_groups), dim3(256), 0, w.stream, w.glist.p, w.grp_off.p, w.grp_body.p,
                           w.grp_moff.p, w.row_ptr.p, w.col_idx.p, w.Hval.p, w.Ginv.p);
        ADB_HIP(hipGetLastError());
    }
}

} // namespace adb
