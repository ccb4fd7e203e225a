// assemble.hip - per-body inertia terms and block-sparse Hessian assembly.
//
// Replaces grad_residue_per_body / hess_inertia_per_body (model/mutils.cpp:53-76,
// model/othogonal_energy.cpp:16-47), gen_empty_sm (model/sparse.cpp:13-57), the
// locked scatter of IPC::output_hessian_gradient (model/assemble_matrices.cpp:28-59)
// and the diagonal += c.hess of ABD::solve (model/abd.cpp:248-260).
//
// No atomics, every sum in a fixed order.  Contacts are radix-sorted by body
// pair (lo, hi), lo < hi; each pair's contacts are cut into chunks of at most
// PCH, and one workgroup per chunk reads every contact record once (staged in
// LDS) and accumulates the chunk's three blocks (lo, lo), (hi, hi), (lo, hi)
// and two gradients.  Then one workgroup per pair sums its (lo, hi) chunk
// blocks into the BSR block (lo, hi) and its transpose (hi, lo), and one
// workgroup per body sums inertia, vertex-ground terms and the diagonal chunk
// blocks of every pair it belongs to.  The BSR matrix (12x12 blocks,
// column-major as the reference's CSC values) has, in row b, the pairs
// (lo, b) sorted by lo, the diagonal, then the pairs (b, hi) sorted by hi.
// Static bodies (mass < 0) follow the reference: contributions to them are
// dropped, their diagonal block is I and their gradient 0.
#include <hipcub/hipcub.hpp>
#include <algorithm>
#include <climits>
#include <cstring>
#include <vector>
#include "dmath.h"
#include "world.h"

namespace adb {

// inertia + orthogonality per body (H column-major): one thread per Hessian
// entry (B x 144 threads, consecutive threads on consecutive entries of one
// body, so the stores coalesce); entries 0..11 of a body also write its
// gradient.  Every entry sees the operations of the plain per-body loop in
// the same order: the diagonal mass term, then + h 4 kappa dt^2 for the
// blocks of a1..a3.
__global__ void k_inertia(int B, double dt, double kap, double ga0, double ga1, double ga2, const double* __restrict__ q,
                          const double* __restrict__ q0, const double* __restrict__ qdot, const double* __restrict__ mass,
                          const double* __restrict__ Ic, double* __restrict__ bodyG, double* __restrict__ bodyH) {
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= 144ll * B) return;
    const int b = (int)(t / 144), e = (int)(t % 144);
    double* G = bodyG + 12 * (size_t)b;
    double* H = bodyH + 144 * (size_t)b;
    const int col = e / 12, row = e % 12;
    if (mass[b] < 0.0) {
        if (e < 12) G[e] = 0.0;
        H[e] = (e % 13 == 0) ? 1.0 : 0.0;
        return;
    }
    const double* Q = q + 12 * (size_t)b;
    const d3 a[4] = {ld3(Q), ld3(Q + 3), ld3(Q + 6), ld3(Q + 9)};
    const double m = mass[b], ic = Ic[b];
    if (e < 12) {
        double go = 0.0;
        const int i = e / 3, r = e % 3;
        if (i > 0) {
            double s0 = 0, s1 = 0, s2 = 0;
            for (int j = 1; j < 4; j++) {
                double s = dot(a[i], a[j]) - (i == j ? 1.0 : 0.0);
                s0 += s * a[j].x; s1 += s * a[j].y; s2 += s * a[j].z;
            }
            go = 4 * kap * (r == 0 ? s0 : (r == 1 ? s1 : s2));
        }
        double qt = q0[12 * (size_t)b + e] + dt * qdot[12 * (size_t)b + e];
        if (e < 3) qt += dt * dt * (e == 0 ? ga0 : (e == 1 ? ga1 : ga2));
        G[e] = dt * dt * go + (e < 3 ? m : ic) * (Q[e] - qt);
    }
    double hv = row == col ? (row < 3 ? m : ic) : 0.0;
    // H[(3 j + c) * 12 + 3 i + r]
    const int j = col / 3, c = col % 3, i = row / 3, r = row % 3;
    if (i >= 1 && j >= 1) {
        double h;
        if (i == j) {
            h = 2 * cmp(a[i], r) * cmp(a[i], c) + (r == c ? dot(a[i], a[i]) - 1 : 0.0);
            for (int k = 1; k < 4; k++) if (k != i) h += cmp(a[k], r) * cmp(a[k], c);
        } else {
            h = (r == c ? dot(a[j], a[i]) : 0.0) + cmp(a[j], r) * cmp(a[i], c);
        }
        hv += h * 4 * kap * dt * dt;
    }
    H[e] = hv;
}

// body-pair key of each contact (lo << bb | hi) and the contact id
__global__ void k_pair_keys(int n_pt, int n_ee, const int4* __restrict__ pt, const int4* __restrict__ ee, int bb,
                            unsigned long long* __restrict__ keys, int* __restrict__ vals) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= n_pt + n_ee) return;
    const int4 t = c < n_pt ? pt[c] : ee[c - n_pt];
    const int lo = min(t.x, t.z), hi = max(t.x, t.z);
    keys[c] = ((unsigned long long)lo << bb) | (unsigned long long)hi;
    vals[c] = c;
}
// 1 where a run of equal keys starts (read by the scan through an iterator:
// the flags are never stored)
struct FirstOfRun {
    const unsigned long long* k;
    __host__ __device__ int operator()(int i) const { return (i == 0 || k[i] != k[i - 1]) ? 1 : 0; }
};
// The pair count: *npd when the host sized the launches by an upper bound np
// (the pair count stays on the device; clamped to the bound, which the host
// checks when it reads the count), else np.
__device__ __forceinline__ int pair_count(int np, const int* npd) { return npd ? min(*npd, np) : np; }
// seg = inclusive scan of the flags: pair p = seg - 1 starts at its first key
// (pairs past cap, the lists' length, are not written).  Each pair also gets
// its key by (hi, lo) (the lower-triangle side of each row) in tk, tv; slots
// [np, cap) of lists sized by a bound (npd: the count on the device) get keys
// above every pair's (all ones), so the sort leaves them at the end.
__global__ void k_pair_meta(int n, int bb, int cap, const int* __restrict__ npd, const unsigned long long* __restrict__ keys,
                            const int* __restrict__ seg, int* __restrict__ pair_start, int* __restrict__ pair_lo,
                            int* __restrict__ pair_hi, unsigned long long* __restrict__ tk, int* __restrict__ tv) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    if (i == n - 1) pair_start[min(seg[i], cap)] = n;   // end of the last pair (or of the list)
    if (npd && i < cap && i >= *npd) { tk[i] = ~0ull; tv[i] = i; }
    if (!(i == 0 || keys[i] != keys[i - 1])) return;
    const int p = seg[i] - 1;
    if (p >= cap) return;
    const int lo = (int)(keys[i] >> bb), hi = (int)(keys[i] & ((1ull << bb) - 1));
    pair_start[p] = i;
    pair_lo[p] = lo;
    pair_hi[p] = hi;
    tk[p] = ((unsigned long long)hi << bb) | (unsigned long long)lo;
    tv[p] = p;
}
__device__ __forceinline__ int lower_bound_int(const int* v, int n, int x) {
    int lo = 0, hi = n;
    while (lo < hi) { const int mid = (lo + hi) / 2; if (v[mid] < x) lo = mid + 1; else hi = mid; }
    return lo;
}
__device__ __forceinline__ int lower_bound_hi(const unsigned long long* tk, int n, int bb, int x) {
    int lo = 0, hi = n;
    while (lo < hi) { const int mid = (lo + hi) / 2; if ((int)(tk[mid] >> bb) < x) lo = mid + 1; else hi = mid; }
    return lo;
}
// per body b (b <= B): first pair with lo = b, first transposed pair with hi =
// b, and the row length (lower pairs + diagonal + upper pairs)
__global__ void k_body_ranges(int B, int npb, const int* __restrict__ npd, int bb, const int* __restrict__ pair_lo,
                              const unsigned long long* __restrict__ tk, int* __restrict__ up_start, int* __restrict__ low_start,
                              int* __restrict__ rowcnt) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b > B) return;
    const int np = pair_count(npb, npd);
    const int u0 = lower_bound_int(pair_lo, np, b), l0 = lower_bound_hi(tk, np, bb, b);
    up_start[b] = u0; low_start[b] = l0;
    if (b == B) { rowcnt[b] = 0; return; }
    const int u1 = lower_bound_int(pair_lo, np, b + 1), l1 = lower_bound_hi(tk, np, bb, b + 1);
    rowcnt[b] = (l1 - l0) + 1 + (u1 - u0);
}
// column indices of row b; each pair's block index in its upper and lower row
__global__ void k_fill_cols(int B, int bb, const int* __restrict__ row_ptr, const int* __restrict__ up_start,
                            const int* __restrict__ low_start, const int* __restrict__ pair_hi,
                            const unsigned long long* __restrict__ tk, const int* __restrict__ tv, int* __restrict__ col_idx,
                            int* __restrict__ diag_blk, int* __restrict__ pair_up, int* __restrict__ pair_low) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= B) return;
    int k = row_ptr[b];
    for (int i = low_start[b]; i < low_start[b + 1]; i++, k++) {
        col_idx[k] = (int)(tk[i] & ((1ull << bb) - 1));
        pair_low[tv[i]] = k;
    }
    diag_blk[b] = k;
    col_idx[k++] = b;
    for (int p = up_start[b]; p < up_start[b + 1]; p++, k++) {
        col_idx[k] = pair_hi[p];
        pair_up[p] = k;
    }
}
__global__ void k_vg_start(int B, int n_vg, const int2* __restrict__ vg, int* __restrict__ start) {
    int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b > B) return;
    int lo = 0, hi = n_vg;      // first entry with body >= b
    while (lo < hi) { int mid = (lo + hi) / 2; if (vg[mid].x < b) lo = mid + 1; else hi = mid; }
    start[b] = lo;
}

// Chunks: at most PCH contacts of one pair, RSLOT contacts per round, 36
// lanes per contact: lane (block, column) with block 0 = (lo, lo),
// 1 = (hi, hi), 2 = (lo, hi) and column col = (be, cc) of the 12x12 block:
//   t[(a, r)]       = sum_b PK[(3a + r) D + 3b + cc] Z[b][s2][be]
//   entry (al, r)  += sum_a Z[a][s1][al] t[(a, r)] + G[s1][al] G[s2][be] F[3r + cc]
// (s1, s2 = the contact's sides of the block's row and column bodies); the
// lanes of blocks 0 and 1 also add the gradient entry col of their body.
constexpr int RSLOT = 7;
constexpr int RLANES = 36;
constexpr int PCH = 16 * RSLOT;   // 112: at step 70 of the icosphere pile 4 % faster than 56 (8 rounds); 28 and 14 slower
constexpr int PART = 3 * 144 + 24;   // chunk partial: 3 blocks (column-major) + gradients of lo, hi
constexpr int REC_USED = R_PK + 82;  // doubles of a record that can be in use (even)
// chunks per pair; 0 for the slots [np, npb] of a launch sized by the bound
__global__ void k_pair_chunks(int npb, const int* __restrict__ npd, const int* __restrict__ pair_start, int* __restrict__ nch) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p > npb) return;
    const int np = pair_count(npb, npd);
    nch[p] = p < np ? (pair_start[p + 1] - pair_start[p] + PCH - 1) / PCH : 0;
}
// chunk_off = exclusive scan of nch (npb + 1 entries); the chunk count also
// goes to counts[3] for the byte accounting
__global__ void k_chunk_owner(int npb, const int* __restrict__ npd, const int* __restrict__ chunk_off, int* __restrict__ chunk_pair,
                              int* __restrict__ counts) {
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    const int np = pair_count(npb, npd);
    if (k < np) for (int c = chunk_off[k]; c < chunk_off[k + 1]; c++) chunk_pair[c] = k;
    if (k == 0) counts[3] = chunk_off[np];
}

// ---- the block-row structure, vertex-ground starts, chunk table and the
// host's copy of the pair lists in one workgroup (the launches of the kernels
// above, which stay for large B + pair counts, each cost more than their work
// in the light phase of a scene)
constexpr int STRUCT_T = 1024, STRUCT_MAX = 1 << 16;
// exclusive scan of in[0, n) into out[0, n], out[n] = the total, by one
// workgroup of STRUCT_T threads (contiguous ranges per thread); returns the total
__device__ int block_scan_excl(const int* in, int* out, int n, int* sh) {
    const int t = threadIdx.x, per = (n + STRUCT_T - 1) / STRUCT_T;
    const int i0 = min(n, t * per), i1 = min(n, i0 + per);
    int s = 0;
    for (int i = i0; i < i1; i++) s += in[i];
    int incl = s;
    for (int d = 1; d < 64; d <<= 1) { const int v = __shfl_up(incl, d); if (__lane_id() >= d) incl += v; }
    if (__lane_id() == 63) sh[t / 64] = incl;
    __syncthreads();
    int off = incl - s, total = 0;
    for (int w = 0; w < STRUCT_T / 64; w++) { if (w < t / 64) off += sh[w]; total += sh[w]; }
    for (int i = i0; i < i1; i++) { out[i] = off; off += in[i]; }
    if (t == 0) out[n] = total;
    __syncthreads();
    return total;
}
// in-place suffix minimum of v[0, n) by one workgroup of STRUCT_T threads
// (contiguous ranges per thread): v[i] = min(v[i], ..., v[n - 1])
__device__ void block_suffix_min(int* v, int n, int* sh) {
    const int t = threadIdx.x, per = (n + STRUCT_T - 1) / STRUCT_T;
    const int i0 = min(n, t * per), i1 = min(n, i0 + per);
    int m = INT_MAX;
    for (int i = i0; i < i1; i++) m = min(m, v[i]);
    int suf = m;   // minimum over this lane and the lanes after it
    for (int d = 1; d < 64; d <<= 1) {
        const int o = __shfl_down(suf, d);
        if (__lane_id() + d < 64) suf = min(suf, o);
    }
    if (__lane_id() == 0) sh[t / 64] = suf;
    __syncthreads();
    int right = __shfl_down(suf, 1);   // minimum over the ranges after this thread's
    if (__lane_id() == 63) right = INT_MAX;
    for (int wv = t / 64 + 1; wv < STRUCT_T / 64; wv++) right = min(right, sh[wv]);
    for (int i = i1 - 1; i >= i0; i--) { right = min(right, v[i]); v[i] = right; }
    __syncthreads();
}
struct StructArgs {
    int B, npb, bb, n_vg;
    const int* npd;
    const int* pair_lo; const int* pair_hi; const int* pair_start; const unsigned long long* tk; const int* tv; const int2* vg;
    int* up_start; int* low_start; int* rowcnt; int* row_ptr; int* col_idx; int* diag_blk; int* pair_up; int* pair_low;
    int* vg_start; int* chunk_n; int* chunk_off; int* chunk_pair; int* counts;
    int* host_pack;   // [pair count, npb pair_lo, npb pair_hi, npb + 1 pair_start]
};
// body arrays of at most STRUCT_LB entries (B + 1) and chunk arrays of at most
// 5 STRUCT_LB / 2 entries (np + 1) are staged in LDS (40 KB, reused): the
// phases between barriers then read their inputs from LDS instead of global
// memory, and the outputs are written once
constexpr int STRUCT_LB = 2048;
__global__ __launch_bounds__(STRUCT_T) void k_structure(StructArgs a) {
    __shared__ int sh[STRUCT_T / 64];
    __shared__ int lb[5 * STRUCT_LB];
    const int np = pair_count(a.npb, a.npd), t = threadIdx.x, B = a.B;
    const bool lb_on = B + 1 <= STRUCT_LB;
    int* up = lb_on ? lb : a.up_start;
    int* low = lb_on ? lb + STRUCT_LB : a.low_start;
    int* vgs = lb_on ? lb + 2 * STRUCT_LB : a.vg_start;
    int* rc = lb_on ? lb + 3 * STRUCT_LB : a.rowcnt;
    int* rp = lb_on ? lb + 4 * STRUCT_LB : a.row_ptr;
    // k_body_ranges and k_vg_start without their binary searches (chains of
    // dependent loads): each body's first entry in a list sorted by body is
    // marked, the other bodies get the list length, and a suffix minimum over
    // the bodies gives every body the first entry at or after it
    for (int b = t; b <= B; b += STRUCT_T) { up[b] = np; low[b] = np; vgs[b] = a.n_vg; }
    __syncthreads();
    for (int i = t; i < np; i += STRUCT_T) {
        const int k = a.pair_lo[i];
        if (i == 0 || a.pair_lo[i - 1] != k) up[k] = i;
        const int h = (int)(a.tk[i] >> a.bb);
        if (i == 0 || (int)(a.tk[i - 1] >> a.bb) != h) low[h] = i;
    }
    for (int i = t; i < a.n_vg; i += STRUCT_T) {
        const int k = a.vg[i].x;
        if (i == 0 || a.vg[i - 1].x != k) vgs[k] = i;
    }
    __syncthreads();
    block_suffix_min(up, B + 1, sh);
    block_suffix_min(low, B + 1, sh);
    block_suffix_min(vgs, B + 1, sh);
    for (int b = t; b < B; b += STRUCT_T) rc[b] = (low[b + 1] - low[b]) + 1 + (up[b + 1] - up[b]);
    __syncthreads();
    block_scan_excl(rc, rp, B, sh);
    if (lb_on)
        for (int b = t; b <= B; b += STRUCT_T) {
            a.up_start[b] = up[b]; a.low_start[b] = low[b]; a.vg_start[b] = vgs[b]; a.row_ptr[b] = rp[b];
        }
    // k_fill_cols
    for (int b = t; b < B; b += STRUCT_T) {
        int k = rp[b];
        for (int i = low[b]; i < low[b + 1]; i++, k++) {
            a.col_idx[k] = (int)(a.tk[i] & ((1ull << a.bb) - 1));
            a.pair_low[a.tv[i]] = k;
        }
        a.diag_blk[b] = k;
        a.col_idx[k++] = b;
        for (int p = up[b]; p < up[b + 1]; p++, k++) {
            a.col_idx[k] = a.pair_hi[p];
            a.pair_up[p] = k;
        }
    }
    __syncthreads();   // the LDS is reused below
    // k_pair_chunks, the scan, k_chunk_owner
    const bool lc_on = 2 * (np + 1) <= 5 * STRUCT_LB;
    int* cn = lc_on ? lb : a.chunk_n;
    int* co = lc_on ? lb + np + 1 : a.chunk_off;
    for (int p = t; p < np; p += STRUCT_T) cn[p] = (a.pair_start[p + 1] - a.pair_start[p] + PCH - 1) / PCH;
    __syncthreads();
    const int nch = block_scan_excl(cn, co, np, sh);
    for (int p = t; p < np; p += STRUCT_T)
        for (int c = co[p]; c < co[p + 1]; c++) a.chunk_pair[c] = p;
    if (lc_on)
        for (int p = t; p <= np; p += STRUCT_T) a.chunk_off[p] = co[p];
    if (t == 0) a.counts[3] = nch;
    // the pair lists for the host's group building, one copy
    if (t == 0) a.host_pack[0] = np;
    for (int p = t; p < np; p += STRUCT_T) { a.host_pack[1 + p] = a.pair_lo[p]; a.host_pack[1 + a.npb + p] = a.pair_hi[p]; }
    for (int p = t; p <= np; p += STRUCT_T) a.host_pack[1 + 2 * a.npb + p] = a.pair_start[p];
}

struct PairArgs {
    int np, n_pt, n_ee, B;
    const int* npd;                           // the pair count on the device (np is then a bound), or null
    const int* cid;                           // contact ids sorted by pair
    const int* pair_start; const int* pair_lo; const int* pair_hi; const int* pair_up; const int* pair_low;
    const int* up_start; const int* low_start; const int* tv;
    const int* chunk_off;
    const int4* pt; const int4* ee; const int* ccls; const double* mass;
    const double* rec; const double* bodyH; const double* bodyG;
    const int2* vg; const int* vg_start; const double* vgK; const int* voff; const double* xbar;
    const int* diag_blk;
    double* Hval; double* grad;
};

// record doubles a thread stages per round (RSLOT records of REC_USED / 2 double2 over 256 threads)
constexpr int STAGE_PER_T = (RSLOT * (REC_USED / 2) + 255) / 256;
// One contact's share of a lane's 12 block entries (column (be, cc) of the
// block (s1, s2)): sum_ab Z[a][s1] Z[b][s2]^T (x) PK_ab + G[s1] G[s2]^T (x) F,
// with the Helmert row count NA of the contact's class as a constant, so a
// 2-point contact evaluates 1 of the 9 row pairs (terms in the same order).
template <int NA>
__device__ __forceinline__ void pair_terms(const double* R, int s1, int s2, int be, int cc, double (&acc)[12]) {
    constexpr int D = 3 * NA;
    double tt[NA][3];
#pragma unroll
    for (int aa = 0; aa < NA; aa++)
#pragma unroll
        for (int r = 0; r < 3; r++) {
            double sum = 0.0;
#pragma unroll
            for (int bb = 0; bb < NA; bb++) sum += R[R_PK + (3 * aa + r) * D + 3 * bb + cc] * R[R_Z + 8 * bb + 4 * s2 + be];
            tt[aa][r] = sum;
        }
    const double g2 = R[R_G + 4 * s2 + be];
    const double F[3] = {R[R_F + cc], R[R_F + 3 + cc], R[R_F + 6 + cc]};
#pragma unroll
    for (int al = 0; al < 4; al++) {
        const double gg = R[R_G + 4 * s1 + al] * g2;
        double z[NA];
#pragma unroll
        for (int aa = 0; aa < NA; aa++) z[aa] = R[R_Z + 8 * aa + 4 * s1 + al];
#pragma unroll
        for (int c = 0; c < 3; c++) {
            double t = z[0] * tt[0][c];
#pragma unroll
            for (int aa = 1; aa < NA; aa++) t = t + z[aa] * tt[aa][c];
            acc[3 * al + c] += t + gg * F[c];
        }
    }
}

__global__ __launch_bounds__(256) void k_pair_reduce(PairArgs a, const int* __restrict__ chunk_pair, double* __restrict__ part) {
    // two record buffers (the next round is loaded while this one is used);
    // at the end the same space holds the per-slot column sums
    constexpr int RECBUF = 2 * RSLOT * REC_USED, PSUM = RSLOT * RLANES * 13;
    __shared__ __align__(16) double lds[RECBUF > PSUM ? RECBUF : PSUM];
    __shared__ int sc[PCH], sside[PCH], scls[PCH];   // per contact of the chunk: id, side of lo, Helmert rows
    // XCD-aware order: workgroups are dealt to the 8 XCDs round-robin; XCD x
    // takes the contiguous chunk range [x per, (x + 1) per)
    const int per = gridDim.x / 8;
    const int ch = (blockIdx.x % 8) * per + blockIdx.x / 8;
    if (ch >= a.chunk_off[pair_count(a.np, a.npd)]) return;
    const int p = chunk_pair[ch];
    const int beg = a.pair_start[p] + (ch - a.chunk_off[p]) * PCH;
    const int end = min(beg + PCH, a.pair_start[p + 1]);
    const int lo = a.pair_lo[p], hi = a.pair_hi[p];
    const bool dlo = a.mass[lo] > 0.0, dhi = a.mass[hi] > 0.0;
    const int t = threadIdx.x;
    if (t < end - beg) {
        const int c = a.cid[beg + t];
        const int I = c < a.n_pt ? a.pt[c].x : a.ee[c - a.n_pt].x;
        sc[t] = c; sside[t] = I == lo ? 0 : 1; scls[t] = a.ccls[c];
    }
    __syncthreads();
    const int slot = t / RLANES, rem = t % RLANES, blk = rem / 12, col = rem % 12, be = col / 3, cc = col % 3;
    const bool use = slot < RSLOT && (blk == 0 ? dlo : (blk == 1 ? dhi : (dlo && dhi)));
    // loads of round k into registers: element e = t + 256 u of the round's
    // RSLOT x (REC_USED / 2) double2 (only the doubles its record uses)
    double2 pre[STAGE_PER_T];
    auto fetch = [&](int r0) {
#pragma unroll
        for (int u = 0; u < STAGE_PER_T; u++) {
            const int e = t + 256 * u, sl = e / (REC_USED / 2), o = e % (REC_USED / 2), k = r0 - beg + sl;
            pre[u] = make_double2(0.0, 0.0);
            if (sl < RSLOT && r0 + sl < end) {
                const int nA = scls[k];
                if (nA > 0 && 2 * o < R_PK + 9 * nA * nA)
                    pre[u] = reinterpret_cast<const double2*>(a.rec + REC * (size_t)sc[k])[o];
            }
        }
    };
    auto store = [&](int buf) {
#pragma unroll
        for (int u = 0; u < STAGE_PER_T; u++) {
            const int e = t + 256 * u;
            if (e < RSLOT * (REC_USED / 2)) reinterpret_cast<double2*>(lds + buf * RSLOT * REC_USED)[e] = pre[u];
        }
    };
    double acc[12], gacc = 0.0;
#pragma unroll
    for (int i = 0; i < 12; i++) acc[i] = 0.0;
    fetch(beg);
    store(0);
    __syncthreads();
    int buf = 0;
    for (int r0 = beg; r0 < end; r0 += RSLOT, buf ^= 1) {
        const bool more = r0 + RSLOT < end;
        if (more) fetch(r0 + RSLOT);
        const int k = r0 - beg + slot;
        if (use && r0 + slot < end && scls[k] > 0) {
            const double* R = lds + (buf * RSLOT + slot) * REC_USED;
            const int nA = scls[k];
            const int slo = sside[k], shi = 1 - slo;
            const int s1 = blk == 1 ? shi : slo, s2 = blk == 0 ? slo : shi;
            if (nA == 1) pair_terms<1>(R, s1, s2, be, cc, acc);
            else if (nA == 2) pair_terms<2>(R, s1, s2, be, cc, acc);
            else pair_terms<3>(R, s1, s2, be, cc, acc);
            if (blk < 2) gacc += R[R_GRAD + 12 * s1 + col];
        }
        if (more) store(buf ^ 1);
        __syncthreads();
    }
    double* P = lds;   // [RSLOT][RLANES][13]
    if (slot < RSLOT) {
#pragma unroll
        for (int i = 0; i < 12; i++) P[(slot * RLANES + rem) * 13 + i] = acc[i];
        P[(slot * RLANES + rem) * 13 + 12] = gacc;
    }
    __syncthreads();
    for (int e = t; e < PART; e += 256) {
        int ln, row;
        if (e < 432) { ln = 12 * (e / 144) + (e % 144) / 12; row = e % 12; }
        else { ln = e - 432; row = 12; }      // gradient of lo (lanes 0-11), of hi (lanes 12-23)
        double sum = 0.0;
        for (int sl = 0; sl < RSLOT; sl++) sum += P[(sl * RLANES + ln) * 13 + row];
        part[PART * (size_t)ch + e] = sum;
    }
}

// off-diagonal blocks: one thread per entry of each pair's block (lo, hi):
// the pair's chunk blocks summed in order, written to (lo, hi) and
// transposed to (hi, lo)
// acc += part[PART c + off] for c in [c0, c1), in order; the loads are issued
// CH_U at a time (a chain of dependent adds would wait for one HBM round trip
// per chunk)
constexpr int CH_U = 8;
__device__ __forceinline__ void add_chunks(double& acc, const double* __restrict__ part, int c0, int c1, int off) {
    for (int c = c0; c < c1; c += CH_U) {
        double v[CH_U];
#pragma unroll
        for (int u = 0; u < CH_U; u++) v[u] = c + u < c1 ? part[PART * (size_t)(c + u) + off] : 0.0;
#pragma unroll
        for (int u = 0; u < CH_U; u++)
            if (c + u < c1) acc += v[u];
    }
}

__global__ void k_pair_offdiag(PairArgs a, const double* __restrict__ part) {
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    const int p = idx / 144, e = idx % 144;
    if (p >= pair_count(a.np, a.npd)) return;
    double acc = 0.0;
    add_chunks(acc, part, a.chunk_off[p], a.chunk_off[p + 1], 288 + e);
    a.Hval[144 * (size_t)a.pair_up[p] + e] = acc;
    a.Hval[144 * (size_t)a.pair_low[p] + 12 * (e % 12) + e / 12] = acc;
}

// diagonal blocks and gradients: one thread per entry (144 block entries, 12
// gradient entries) of each body: inertia, the (lo, lo) chunk blocks of its
// pairs (b, hi), the (hi, hi) chunk blocks of its pairs (lo, b), each in
// order, then its vertex-ground terms
__global__ void k_pair_diag(PairArgs a, const double* __restrict__ part) {
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    const int b = idx / 156, e = idx % 156;
    if (b >= a.B) return;
    const bool isg = e >= 144;
    const int ge = e - 144;
    double acc = isg ? a.bodyG[12 * (size_t)b + ge] : a.bodyH[144 * (size_t)b + e];
    const int off_up = isg ? 432 + ge : e, off_low = isg ? 444 + ge : 144 + e;
    // pairs (b, hi): their chunks are contiguous
    add_chunks(acc, part, a.chunk_off[a.up_start[b]], a.chunk_off[a.up_start[b + 1]], off_up);
    // pairs (lo, b), CH_U pairs' chunk ranges fetched at a time
    for (int i0 = a.low_start[b], i1 = a.low_start[b + 1]; i0 < i1; i0 += CH_U) {
        int c0[CH_U], c1[CH_U];
#pragma unroll
        for (int u = 0; u < CH_U; u++) {
            const int p = i0 + u < i1 ? a.tv[i0 + u] : 0;
            c0[u] = i0 + u < i1 ? a.chunk_off[p] : 0;
            c1[u] = i0 + u < i1 ? a.chunk_off[p + 1] : 0;
        }
#pragma unroll
        for (int u = 0; u < CH_U; u++) add_chunks(acc, part, c0[u], c1[u], off_low);
    }
    if (a.mass[b] > 0.0) {
        // vertex-ground terms of this body: (w w^T) (x) K, gradient w (x) gx;
        // the loads of CH_U contacts are issued together, the adds stay in order
        const int cc = e / 12, rr = e % 12;
        const int ka = isg ? ge / 3 : rr / 3, kb = isg ? 0 : cc / 3;
        const int kk = isg ? 9 + ge % 3 : 3 * (cc % 3) + (rr % 3);
        for (int s0 = a.vg_start[b], s1 = a.vg_start[b + 1]; s0 < s1; s0 += CH_U) {
            double wa[CH_U], wb[CH_U], kv[CH_U];
#pragma unroll
            for (int u = 0; u < CH_U; u++) {
                const int s = s0 + u < s1 ? s0 + u : s0;
                const size_t v = (size_t)a.voff[b] + a.vg[s].y;
                wa[u] = ka == 0 ? 1.0 : a.xbar[3 * v + ka - 1];
                wb[u] = kb == 0 ? 1.0 : a.xbar[3 * v + kb - 1];
                kv[u] = a.vgK[12 * (size_t)s + kk];
            }
#pragma unroll
            for (int u = 0; u < CH_U; u++)
                if (s0 + u < s1) acc += isg ? wa[u] * kv[u] : wa[u] * wb[u] * kv[u];
        }
    }
    if (isg) a.grad[12 * (size_t)b + ge] = acc;
    else a.Hval[144 * (size_t)a.diag_blk[b] + e] = acc;
}

// Preconditioner: the explicit inverse of the diagonal sub-matrix of each
// body group (at most GMAX bodies, built by heavy-edge aggregation of the
// contact graph, build_groups).  On groups that are whole components the
// preconditioned system decouples exactly, so PCG needs only a few
// iterations where the contact graph is a union of small clusters; in large
// clusters it is block-Jacobi with blocks of up to GMAX bodies.
//
// The inverse is formed by the block sweep operator (the sweep operator of
// Goodnight, "A tutorial on the SWEEP operator", The American Statistician
// 33(3), 1979, taken one 12 x 12 body block at a time).  Sweeping block K of
// a symmetric A with D = A_KK replaces
//     A_KK <- -D^-1,  A_iK <- A_iK D^-1,  A_Kj <- D^-1 A_Kj,
//     A_ij <- A_ij - A_iK D^-1 A_Kj                    (i, j outside K),
// and after every block A holds -A^-1.  For a symmetric positive definite A
// every D is a Schur complement, itself positive definite; a non-positive
// pivot while inverting D (itself a 12-step scalar sweep) marks the group
// as not positive definite, which then falls back to the inverse of its
// diagonal.  A group of m bodies takes m block steps, each a rank-12 update
// of the whole upper triangle in which every lane of the group works:
// a lane owns one 6 x 6 tile (a 6 x 6 tile lies in one body block) and
// updates it with 12 fused multiply-adds per entry from the column panel
// P = A_:K and W = P D^-1, both in LDS (transposed, so a tile's 6 rows or
// columns are 6 consecutive doubles).  Size classes, one launch, classes by
// block range, largest first: 5..8 bodies one group per 256-lane workgroup
// (136 tiles of 16 x 16), 1..4 bodies one group per wavefront (36 tiles of
// 8 x 8, wavefront-level synchronisation).
constexpr int GMAX = 8;
constexpr int GI_LD = 97;            // padded row length of the transposed panels (class 5..8)
constexpr int GI_LD4 = 49;           // the same for groups of up to 4 bodies
// LDS of one group: PT[12][LD], WT[12][LD], D[12][13]
constexpr int GI_LDS8 = 24 * GI_LD + 156, GI_LDS4 = 24 * GI_LD4 + 156;

__device__ __forceinline__ void gi_sync(bool block) {
    if (block) {
        __syncthreads();
    } else {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    }
}
// upper-triangle tile of a lane: tiles enumerated row by row, (bi, bj), bi <= bj < nt
__device__ __forceinline__ bool upper_tile(int l, int nt, int& bi, int& bj) {
    int r = 0, rem = l;
    while (r < nt && rem >= nt - r) { rem -= nt - r; r++; }
    bi = r; bj = r + rem;
    return r < nt;
}
// In-place scalar sweep of the 12 x 12 symmetric D (LDS, row length 13) by
// the team's NL lanes; afterwards D holds -D^-1.  Returns false (on every
// lane) when a pivot is not positive.
template <int NL>
__device__ __forceinline__ bool sweep12(double* D, int lane, bool block_sync) {
    constexpr int S = (144 + NL - 1) / NL;
    for (int k = 0; k < 12; k++) {
        const double d = D[13 * k + k];
        if (!(d > 0.0) || !(d < INFINITY)) return false;       // the same d on every lane
        const double inv = 1.0 / d;
        double v[S];
#pragma unroll
        for (int s = 0; s < S; s++) {
            const int e = lane + NL * s, i = e / 12, j = e % 12;
            v[s] = 0.0;
            if (e < 144) {
                const double ci = D[13 * i + k], cj = D[13 * k + j];
                double x = D[13 * i + j] - ci * (cj * inv);
                if (i == k || j == k) x = (i == j) ? -inv : (i == k ? cj * inv : ci * inv);
                v[s] = x;
            }
        }
        gi_sync(block_sync);
#pragma unroll
        for (int s = 0; s < S; s++) {
            const int e = lane + NL * s;
            if (e < 144) D[13 * (e / 12) + e % 12] = v[s];
        }
        gi_sync(block_sync);
    }
    return true;
}
// Block sweep of one group of m bodies (n = 12 m <= NMAX) by a team of NL
// lanes.  sm: LDS of the group (PT, WT, D); blk: m x m BSR block positions (-1: zero).
template <int NMAX, int NL, int LD>
__device__ __forceinline__ void group_bsweep(int lane, int m, const int* blk, const double* __restrict__ Hval,
                                             double* __restrict__ out, double* sm, bool block_sync) {
    constexpr int T = 6, NT = NMAX / T;
    double* PT = sm;                 // PT[r][i] = A[i][K0 + r]
    double* WT = sm + 12 * LD;       // WT[c][i] = (P D^-1)[i][c]
    double* D = sm + 24 * LD;        // 12 x 12, row length 13
    const int n = 12 * m, nt = n / T;
    int bi, bj;
    const bool own = upper_tile(lane, NT, bi, bj) && bj < nt;
    const int i0 = T * bi, j0 = T * bj, brow = bi / 2, bcol = bj / 2;
    double a[T][T];
    if (own) {
        // Hval blocks are column-major 12 x 12
        const int k = blk[brow * m + bcol];
#pragma unroll
        for (int y = 0; y < T; y++)
#pragma unroll
            for (int x = 0; x < T; x++)
                a[x][y] = k >= 0 ? Hval[144 * (size_t)k + 12 * ((j0 + y) % 12) + (i0 + x) % 12] : 0.0;
    }
    bool pd = true;
    for (int kb = 0; kb < m; kb++) {
        const int K0 = 12 * kb;
        // the column panel A[:, K] from the upper triangle (and its mirror)
        if (own && (bcol == kb || brow == kb)) {
#pragma unroll
            for (int x = 0; x < T; x++)
#pragma unroll
                for (int y = 0; y < T; y++) {
                    const int i = i0 + x, j = j0 + y;
                    if (i <= j) {
                        if (bcol == kb) PT[(j - K0) * LD + i] = a[x][y];
                        if (brow == kb) PT[(i - K0) * LD + j] = a[x][y];
                    }
                }
        }
        gi_sync(block_sync);
        // D = the panel's rows in K; -D^-1 by the scalar sweep
        for (int e = lane; e < 144; e += NL) D[13 * (e / 12) + e % 12] = PT[(e % 12) * LD + K0 + e / 12];
        gi_sync(block_sync);
        if (!sweep12<NL>(D, lane, block_sync)) { pd = false; break; }
        // W = P D^-1 = -P (D swept), rows outside K (rows in K are not used)
        for (int e = lane; e < 12 * n; e += NL) {
            const int i = e / 12, c = e % 12;
            double s = 0.0;
#pragma unroll
            for (int r = 0; r < 12; r++) s = fma(PT[r * LD + i], D[13 * r + c], s);
            WT[c * LD + i] = -s;
        }
        gi_sync(block_sync);
        if (own) {
            if (brow != kb && bcol != kb) {
                // A_ij -= sum_r W[i][r] P[j][r]
#pragma unroll 2
                for (int r = 0; r < 12; r++) {
                    double wv[T], pv[T];
#pragma unroll
                    for (int x = 0; x < T; x++) wv[x] = WT[r * LD + i0 + x];
#pragma unroll
                    for (int y = 0; y < T; y++) pv[y] = PT[r * LD + j0 + y];
#pragma unroll
                    for (int x = 0; x < T; x++)
#pragma unroll
                        for (int y = 0; y < T; y++) a[x][y] = fma(-wv[x], pv[y], a[x][y]);
                }
            } else if (brow == kb && bcol == kb) {
#pragma unroll
                for (int x = 0; x < T; x++)
#pragma unroll
                    for (int y = 0; y < T; y++) a[x][y] = D[13 * (i0 + x - K0) + j0 + y - K0];     // -D^-1
            } else if (bcol == kb) {
#pragma unroll
                for (int x = 0; x < T; x++)
#pragma unroll
                    for (int y = 0; y < T; y++) a[x][y] = WT[(j0 + y - K0) * LD + i0 + x];       // (A_iK D^-1)
            } else {
#pragma unroll
                for (int x = 0; x < T; x++)
#pragma unroll
                    for (int y = 0; y < T; y++) a[x][y] = WT[(i0 + x - K0) * LD + j0 + y];       // (D^-1 A_Kj)
            }
        }
        gi_sync(block_sync);
    }
    if (!pd) {
        // not positive definite: the inverse diagonal of the original matrix (from Hval)
        if (own)
#pragma unroll
            for (int x = 0; x < T; x++)
#pragma unroll
                for (int y = 0; y < T; y++) {
                    const int i = i0 + x, j = j0 + y;
                    double v = 0.0;
                    if (i == j) {
                        const int k = blk[(i / 12) * m + i / 12];
                        const double dd = k >= 0 ? Hval[144 * (size_t)k + 13 * (i % 12)] : 1.0;
                        v = 1.0 / (dd > 0.0 ? dd : 1.0);
                    }
                    out[(size_t)j * n + i] = v;
                    out[(size_t)i * n + j] = v;
                }
        return;
    }
    // A = -M^-1 (upper triangle); store both triangles, column-major
    if (own)
#pragma unroll
        for (int x = 0; x < T; x++)
#pragma unroll
            for (int y = 0; y < T; y++) {
                const int i = i0 + x, j = j0 + y;
                if (i <= j) {
                    out[(size_t)j * n + i] = -a[x][y];
                    out[(size_t)i * n + j] = -a[x][y];
                }
            }
}
// block index of every (body a, body b) pair of a group, row-major m x m
__device__ __forceinline__ void group_blocks(int lane, int nl, int m, const int* bod, const int* __restrict__ row_ptr,
                                             const int* __restrict__ col_idx, int* blk) {
    for (int e = lane; e < m * m; e += nl) {
        const int ia = e / m, ib = e % m;
        int k = -1;
        for (int q = row_ptr[bod[ia]]; q < row_ptr[bod[ia] + 1]; q++)
            if (col_idx[q] == bod[ib]) { k = q; break; }
        blk[e] = k;
    }
}
__global__ __launch_bounds__(256) void k_group_inv(const int* __restrict__ list, int nb8, int n4,
                                                   const int* __restrict__ grp_off, const int* __restrict__ grp_body,
                                                   const long long* __restrict__ grp_moff, const int* __restrict__ row_ptr,
                                                   const int* __restrict__ col_idx, const double* __restrict__ Hval,
                                                   double* __restrict__ Ginv) {
    static_assert(4 * GI_LDS4 >= GI_LDS8, "LDS of one 5..8-body group within that of four small groups");
    __shared__ double sm[4 * GI_LDS4];
    __shared__ int blk[4][64];
    __shared__ int bod[4][8];
    const int t = threadIdx.x, wv = t / 64, ln = t % 64;
    if ((int)blockIdx.x < nb8) {
        const int g = list[blockIdx.x];
        const int m = grp_off[g + 1] - grp_off[g];
        if (t < m) bod[0][t] = grp_body[grp_off[g] + t];
        __syncthreads();
        group_blocks(t, 256, m, bod[0], row_ptr, col_idx, blk[0]);
        __syncthreads();
        group_bsweep<96, 256, GI_LD>(t, m, blk[0], Hval, Ginv + grp_moff[g], sm, true);
    } else {
        const int q = 4 * ((int)blockIdx.x - nb8) + wv;           // one group per wavefront
        if (q >= n4) return;                                       // whole wavefront: no block barrier below
        const int g = list[nb8 + q];
        const int m = grp_off[g + 1] - grp_off[g];
        if (ln < m) bod[wv][ln] = grp_body[grp_off[g] + ln];
        gi_sync(false);
        group_blocks(ln, 64, m, bod[wv], row_ptr, col_idx, blk[wv]);
        gi_sync(false);
        group_bsweep<48, 64, GI_LD4>(ln, m, blk[wv], Hval, Ginv + grp_moff[g], sm + wv * GI_LDS4, false);
    }
}

// Body groups for the preconditioner (host): heavy-edge aggregation of the
// contact graph.  Body pairs are taken in decreasing order of their contact
// count (ties by pair index) and their groups merged while the union has at
// most GMAX bodies, so the strongest couplings land inside the inverted
// diagonal sub-matrices; every connected component of at most GMAX bodies
// ends up as one group (the preconditioned system then decouples exactly).
// The pair lists go to pinned host memory before the assembly reduction is
// queued (groups_fetch), so the host builds the groups while it runs.
// h_pairs = [pair count, npb pair_lo, npb pair_hi, npb + 1 pair_start]: the
// first npb entries of each list (npb = the pair count, or a bound of it
// with npd the count on the device); build_groups reads the count with them.
static void groups_fetch(World& w, int npb, const int* npd, const int* pack, hipStream_t st) {
    if (3 * (size_t)npb + 2 > w.h_pairs_cap) {
        if (w.h_pairs) ADB_HIP(hipHostFree(w.h_pairs));
        w.h_pairs_cap = 3 * (size_t)npb + 2 + (size_t)npb / 2 + 64;
        ADB_HIP(hipHostMalloc((void**)&w.h_pairs, w.h_pairs_cap * sizeof(int), hipHostMallocDefault));
    }
    if (!w.ev_pairs) ADB_HIP(hipEventCreateWithFlags(&w.ev_pairs, hipEventDisableTiming));
    w.h_pairs_n = npb;
    if (pack) {
        // the same layout, packed on the device (k_structure): one copy
        ADB_HIP(hipMemcpyAsync(w.h_pairs, pack, (3 * (size_t)npb + 2) * sizeof(int), hipMemcpyDeviceToHost, st));
    } else if (npd) {
        ADB_HIP(hipMemcpyAsync(w.h_pairs, npd, sizeof(int), hipMemcpyDeviceToHost, st));
    } else {
        w.h_pairs[0] = npb;
    }
    if (!pack && npb > 0) {
        ADB_HIP(hipMemcpyAsync(w.h_pairs + 1, w.pair_lo.p, npb * sizeof(int), hipMemcpyDeviceToHost, st));
        ADB_HIP(hipMemcpyAsync(w.h_pairs + 1 + npb, w.pair_hi.p, npb * sizeof(int), hipMemcpyDeviceToHost, st));
        ADB_HIP(hipMemcpyAsync(w.h_pairs + 1 + 2 * (size_t)npb, w.pair_start.p, (npb + 1) * sizeof(int), hipMemcpyDeviceToHost,
                               st));
    }
    ADB_HIP(hipEventRecord(w.ev_pairs, st));
}
static void build_groups(World& w) {
    wait_event(w, w.ev_pairs);
    const int B = w.B, npb = w.h_pairs_n, np = w.h_pairs[0];
    if (np > npb) throw AdbError(ADB_E_ARG, "body pairs with contacts exceed the culling pass's pair count");
    w.n_pairs = np;
    w.nnzb = B + 2 * np;
    const int* plo = w.h_pairs + 1;
    const int* phi = w.h_pairs + 1 + npb;
    const int* ps = w.h_pairs + 1 + 2 * (size_t)npb;
    // pairs by decreasing contact count, ties by index (counting sort)
    int wmax = 0;
    for (int p = 0; p < np; p++) wmax = std::max(wmax, ps[p + 1] - ps[p]);
    std::vector<int>& bucket = w.hg_bucket;
    std::vector<int>& order = w.hg_order;
    bucket.assign(wmax + 2, 0);
    order.resize(np);
    for (int p = 0; p < np; p++) bucket[wmax - (ps[p + 1] - ps[p]) + 1]++;
    for (int i = 1; i <= wmax + 1; i++) bucket[i] += bucket[i - 1];
    for (int p = 0; p < np; p++) order[bucket[wmax - (ps[p + 1] - ps[p])]++] = p;
    std::vector<int>& parent = w.hg_parent;
    std::vector<int>& size = w.hg_size;
    parent.resize(B);
    size.assign(B, 1);
    for (int b = 0; b < B; b++) parent[b] = b;
    auto find = [&](int x) { while (parent[x] != x) { parent[x] = parent[parent[x]]; x = parent[x]; } return x; };
    for (int p : order) {
        int ra = find(plo[p]), rb = find(phi[p]);
        if (ra == rb || size[ra] + size[rb] > GMAX) continue;
        if (ra > rb) std::swap(ra, rb);
        parent[rb] = ra; size[ra] += size[rb];
    }
    // groups in order of their smallest body; bodies ascending inside: group
    // numbers by first appearance, then a counting sort of the bodies
    std::vector<int>& gid = w.hg_gid;    // group of each root
    std::vector<int>& bg = w.hg_bg;      // group of each body
    gid.assign(B, -1);
    bg.resize(B);
    int G = 0;
    for (int b = 0; b < B; b++) {
        const int r = find(b);
        if (gid[r] < 0) gid[r] = G++;
        bg[b] = gid[r];
    }
    w.h_grp_off.assign(G + 1, 0);
    for (int b = 0; b < B; b++) w.h_grp_off[bg[b] + 1]++;
    for (int g = 0; g < G; g++) w.h_grp_off[g + 1] += w.h_grp_off[g];
    w.h_grp_body.resize(B);
    gid.assign(w.h_grp_off.begin(), w.h_grp_off.end() - 1);   // fill positions
    for (int b = 0; b < B; b++) w.h_grp_body[gid[bg[b]]++] = b;
    w.h_grp_moff.resize(G + 1);
    w.h_grp_moff[0] = 0;
    for (int g = 0; g < G; g++) {
        const long long n = 12 * (long long)(w.h_grp_off[g + 1] - w.h_grp_off[g]);
        w.h_grp_moff[g + 1] = w.h_grp_moff[g] + n * n;
    }
    w.n_groups = G;
    if (w.S > 1) {
        // groups are numbered in order of their smallest body and never span two
        // scenes (no body pair crosses scenes): scene s owns groups
        // [h_sc_goff[s], h_sc_goff[s + 1])
        w.h_sc_goff.assign(w.S + 1, w.n_groups);
        int g = 0;
        for (int sc = 0; sc < w.S; sc++) {
            while (g < w.n_groups && w.h_grp_body[w.h_grp_off[g]] < w.h_scene_off[sc]) g++;
            w.h_sc_goff[sc] = g;
        }
        w.sc_goff.reserve(w.S + 1);
        ADB_HIP(hipMemcpyAsync(w.sc_goff.p, w.h_sc_goff.data(), (w.S + 1) * sizeof(int), hipMemcpyHostToDevice, w.stream));
    }
    // group lists by size class, largest first (their inverses take longest;
    // the inverse kernel gives each class its own team width): 5..GMAX
    // bodies, 2..4, single bodies
    w.h_glist.clear();
    for (int g = 0; g < w.n_groups; g++) if (w.h_grp_off[g + 1] - w.h_grp_off[g] > 4) w.h_glist.push_back(g);
    w.n_glist[0] = (int)w.h_glist.size();
    for (int g = 0; g < w.n_groups; g++) {
        const int m = w.h_grp_off[g + 1] - w.h_grp_off[g];
        if (m > 1 && m <= 4) w.h_glist.push_back(g);
    }
    w.n_glist[1] = (int)w.h_glist.size() - w.n_glist[0];
    for (int g = 0; g < w.n_groups; g++) if (w.h_grp_off[g + 1] - w.h_grp_off[g] == 1) w.h_glist.push_back(g);
    w.n_glist[2] = (int)w.h_glist.size() - w.n_glist[0] - w.n_glist[1];
    // one upload of all group tables (pinned staging; the previous upload from
    // it has completed before it is rewritten).  Offsets depend on B only (at
    // most B groups), so the tables' device addresses stay the same from one
    // Newton iteration to the next (the captured PCG batch is keyed on them).
    const size_t nll = (size_t)B + 1, nint = (size_t)(B + 1) + B + B + 1, total = nll + (nint + 1) / 2;
    if (!w.ev_grp) ADB_HIP(hipEventCreateWithFlags(&w.ev_grp, hipEventDisableTiming));
    wait_event(w, w.ev_grp);
    if (total > w.h_grp_pack_cap) {
        if (w.h_grp_pack) ADB_HIP(hipHostFree(w.h_grp_pack));
        w.h_grp_pack_cap = total + total / 2 + 64;
        ADB_HIP(hipHostMalloc((void**)&w.h_grp_pack, w.h_grp_pack_cap * sizeof(long long), hipHostMallocDefault));
        std::memset(w.h_grp_pack, 0, w.h_grp_pack_cap * sizeof(long long));
    }
    w.grp_pack.reserve(total);
    {
        long long* hm = w.h_grp_pack;
        int* hi = reinterpret_cast<int*>(w.h_grp_pack + nll);
        std::memcpy(hm, w.h_grp_moff.data(), (size_t)(G + 1) * sizeof(long long));
        std::memcpy(hi, w.h_grp_off.data(), (size_t)(G + 1) * sizeof(int));
        std::memcpy(hi + B + 1, w.h_grp_body.data(), (size_t)B * sizeof(int));
        std::memcpy(hi + 2 * B + 1, w.h_glist.data(), (size_t)G * sizeof(int));
        hi[3 * B + 1] = G;
    }
    w.grp_moff = w.grp_pack.p;
    w.grp_off = reinterpret_cast<int*>(w.grp_pack.p + nll);
    w.grp_body = w.grp_off + B + 1;
    w.glist = w.grp_body + B;
    w.grp_n = w.glist + B;
    ADB_HIP(hipMemcpyAsync(w.grp_pack.p, w.h_grp_pack, total * sizeof(long long), hipMemcpyHostToDevice, w.stream));
    ADB_HIP(hipEventRecord(w.ev_grp, w.stream));
    // the host vectors stay alive (members of w) until the copies have run
    w.Ginv.reserve((size_t)w.h_grp_moff.back());
}

// FP64 operations of k_group_inv for a group of m bodies (n = 12 m), counted
// from the block sweep (a fused multiply-add is 2): per block step the 12 x 12
// scalar sweep (12 steps of 144 entries, 3 operations each, and a
// reciprocal), W = P D^-1 (12 n entries of 12 multiply-adds) and the rank-12
// update of the upper 6 x 6 tiles outside block row and column K (12
// multiply-adds per entry; diagonal tiles whole).
static double group_inv_flops(int m) {
    const int nt = 2 * m;
    double f = 0.0;
    for (int kb = 0; kb < m; kb++) {
        f += 12.0 * (144.0 * 3.0 + 1.0);
        f += 12.0 * (12.0 * m) * 24.0;
        for (int bi = 0; bi < nt; bi++)
            for (int bj = bi; bj < nt; bj++)
                if (bi / 2 != kb && bj / 2 != kb) f += 36.0 * 24.0;
    }
    return f;
}

// Stable radix sort of (key, value) pairs on bits [0, end_bit) with rocPRIM:
// its block sort + merge passes with 4096-item blocks up to 400k items (at
// 1.9e5 body-pair keys 57 us against 65 with its default 1024-item blocks),
// the merges odd-even over that whole range (one launch per pass; its default
// switches to two-launch merge-path passes above ~2e5 items: 2.15e5 keys 84 ->
// 60 us, 4e5 keys 113 -> 83 us), onesweep above (at 6e5 keys 91 us against
// 134-155 for the merge passes).
using PairSortConfig = rocprim::radix_sort_config<rocprim::default_config,
                                                  rocprim::merge_sort_config<512, 512, 8, 128, 128, 4, 400000>,
                                                  rocprim::default_config, 400000>;
template <class K, class V>
static void sort_pairs(World& w, K* kin, K* kout, V* vin, V* vout, int n, int end_bit, hipStream_t st) {
    size_t tmp = 0;
    ADB_HIP(rocprim::radix_sort_pairs<PairSortConfig>(nullptr, tmp, kin, kout, vin, vout, (size_t)n, 0, end_bit, st));
    w.sort_tmp.reserve(tmp + 16);
    ADB_HIP(rocprim::radix_sort_pairs<PairSortConfig>((void*)w.sort_tmp.p, tmp, kin, kout, vin, vout, (size_t)n, 0, end_bit, st));
}

// Queues the pair structure of w.cs on stream4, from the fork point ev_asm_fork
// (recorded by the caller where the contact set was final when forked, else
// here): contacts sorted by body pair, pair lists, BSR block rows, the chunk
// table and the host's copy of the pair lists; its plan (sizes, the device
// pair count, the set's serial) goes to w.ap for assemble_system.
void assemble_structure(World& w, bool forked) {
    const int B = w.B;
    if (!w.stream4) {
        ADB_HIP(hipStreamCreate(&w.stream4));
        ADB_HIP(hipEventCreateWithFlags(&w.ev_asm_join, hipEventDisableTiming));
    }
    if (!w.ev_asm_fork) ADB_HIP(hipEventCreateWithFlags(&w.ev_asm_fork, hipEventDisableTiming));
    if (!forked) ADB_HIP(hipEventRecord(w.ev_asm_fork, w.stream));
    const hipStream_t st = w.stream4;
    ADB_HIP(hipStreamWaitEvent(st, w.ev_asm_fork, 0));
    const int nc = w.cs.n_pt + w.cs.n_ee;
    int bb = 1; while ((1ll << bb) <= B) bb++;
    // 1. contacts sorted by body pair; pair segments
    w.akeys.reserve(std::max(nc, 1)); w.akeys_s.reserve(std::max(nc, 1));
    w.aval.reserve(std::max(nc, 1)); w.aval_s.reserve(std::max(nc, 1));
    w.blk_of.reserve(std::max(nc, 1));
    int np = 0;
    if (nc > 0) {
        hipLaunchKernelGGL(k_pair_keys, dim3((nc + 255) / 256), dim3(256), 0, st, w.cs.n_pt, w.cs.n_ee, w.cs.pt.p, w.cs.ee.p, bb,
                           w.akeys.p, w.aval.p);
        ADB_HIP(hipGetLastError());
        sort_pairs(w, w.akeys.p, w.akeys_s.p, w.aval.p, w.aval_s.p, nc, 2 * bb, st);
        // pair segment of each sorted contact: inclusive scan of the run starts
        const hipcub::TransformInputIterator<int, FirstOfRun, hipcub::CountingInputIterator<int>> starts(
            hipcub::CountingInputIterator<int>(0), FirstOfRun{w.akeys_s.p});
        size_t tmp = 0;
        ADB_HIP(hipcub::DeviceScan::InclusiveSum(nullptr, tmp, starts, w.blk_of.p, nc, st));
        w.sort_tmp.reserve(tmp + 16);
        ADB_HIP(hipcub::DeviceScan::InclusiveSum(w.sort_tmp.p, tmp, starts, w.blk_of.p, nc, st));
    }
    // The pair count np = blk_of[nc - 1] stays on the device when the culling
    // pass that made the set bounded it (its overlapping body pairs): every
    // launch below is sized by that bound and reads np there; the host gets
    // np with the pair lists (build_groups).  Else it is read here.
    const int* npd = nullptr;
    if (nc > 0 && w.cs.np_bound >= 0 && !w.general_paths) {
        npd = w.blk_of.p + nc - 1;
        np = std::min(w.cs.np_bound, nc);
    } else if (nc > 0) {
        if (!w.h_np) ADB_HIP(hipHostMalloc((void**)&w.h_np, sizeof(int), hipHostMallocDefault));
        ADB_HIP(hipMemcpyAsync(w.h_np, w.blk_of.p + nc - 1, sizeof(int), hipMemcpyDeviceToHost, st));
        wait_stream(w, st);
        np = *w.h_np;
    }
    // np: the pair count or its bound (npd set) from here on
    w.pair_start.reserve(np + 1); w.pair_lo.reserve(std::max(np, 1)); w.pair_hi.reserve(std::max(np, 1));
    w.pair_up.reserve(std::max(np, 1)); w.pair_low.reserve(std::max(np, 1));
    w.tkeys.reserve(std::max(np, 1)); w.tkeys_s.reserve(std::max(np, 1)); w.tval.reserve(std::max(np, 1)); w.tval_s.reserve(std::max(np, 1));
    if (nc > 0) {
        hipLaunchKernelGGL(k_pair_meta, dim3((nc + 255) / 256), dim3(256), 0, st, nc, bb, np, npd, w.akeys_s.p, w.blk_of.p,
                           w.pair_start.p, w.pair_lo.p, w.pair_hi.p, w.tkeys.p, w.tval.p);
        ADB_HIP(hipGetLastError());
        sort_pairs(w, w.tkeys.p, w.tkeys_s.p, w.tval.p, w.tval_s.p, np, 2 * bb, st);
    }
    if (nc == 0) ADB_HIP(hipMemsetAsync(w.pair_start.p, 0, sizeof(int), st));
    // 2. BSR structure: row b = pairs (lo, b), diagonal, pairs (b, hi)
    const int nnzb = B + 2 * np;
    w.nnzb = nnzb;
    w.up_start.reserve(B + 1); w.low_start.reserve(B + 1); w.rowcnt.reserve(B + 1);
    w.row_ptr.reserve(B + 1); w.col_idx.reserve(nnzb); w.diag_blk.reserve(B);
    w.Hval.reserve(144 * (size_t)nnzb);
    w.vg_start.reserve(B + 1);
    w.grad.reserve(12 * (size_t)B);
    w.chunk_n.reserve(np + 1); w.chunk_off.reserve(np + 1);
    // each pair has at most len / PCH + 1 chunks: launch for the bound, rounded
    // up to the 8 XCDs; the kernel reads the exact count
    const int max_chunks = (np + nc / PCH + 7) / 8 * 8;
    w.chunk_blk.reserve(std::max(max_chunks, 1));
    w.chunk_part.reserve(PART * (size_t)std::max(max_chunks, 1));
    const bool one_wg = B + np <= STRUCT_MAX && !w.general_paths;
    if (one_wg) {
        // 3. block rows, vertex-ground starts and chunks in one workgroup
        w.pair_pack.reserve(3 * (size_t)np + 2);
        StructArgs sa;
        sa.B = B; sa.npb = np; sa.bb = bb; sa.n_vg = w.cs.n_vg; sa.npd = npd;
        sa.pair_lo = w.pair_lo.p; sa.pair_hi = w.pair_hi.p; sa.pair_start = w.pair_start.p; sa.tk = w.tkeys_s.p; sa.tv = w.tval_s.p;
        sa.vg = w.cs.vg.p;
        sa.up_start = w.up_start.p; sa.low_start = w.low_start.p; sa.rowcnt = w.rowcnt.p; sa.row_ptr = w.row_ptr.p;
        sa.col_idx = w.col_idx.p; sa.diag_blk = w.diag_blk.p; sa.pair_up = w.pair_up.p; sa.pair_low = w.pair_low.p;
        sa.vg_start = w.vg_start.p; sa.chunk_n = w.chunk_n.p; sa.chunk_off = w.chunk_off.p; sa.chunk_pair = w.chunk_blk.p;
        sa.counts = w.ccount.p; sa.host_pack = w.pair_pack.p;
        hipLaunchKernelGGL(k_structure, dim3(1), dim3(STRUCT_T), 0, st, sa);
        ADB_HIP(hipGetLastError());
    } else {
        hipLaunchKernelGGL(k_body_ranges, dim3((B + 1 + 255) / 256), dim3(256), 0, st, B, np, npd, bb, w.pair_lo.p,
                           w.tkeys_s.p, w.up_start.p, w.low_start.p, w.rowcnt.p);
        ADB_HIP(hipGetLastError());
        {
            size_t tmp = 0;
            ADB_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, tmp, w.rowcnt.p, w.row_ptr.p, B + 1, st));
            w.sort_tmp.reserve(tmp + 16);
            ADB_HIP(hipcub::DeviceScan::ExclusiveSum(w.sort_tmp.p, tmp, w.rowcnt.p, w.row_ptr.p, B + 1, st));
        }
        hipLaunchKernelGGL(k_fill_cols, dim3((B + 255) / 256), dim3(256), 0, st, B, bb, w.row_ptr.p, w.up_start.p,
                           w.low_start.p, w.pair_hi.p, w.tkeys_s.p, w.tval_s.p, w.col_idx.p, w.diag_blk.p, w.pair_up.p, w.pair_low.p);
        ADB_HIP(hipGetLastError());
        hipLaunchKernelGGL(k_vg_start, dim3((B + 1 + 255) / 256), dim3(256), 0, st, B, w.cs.n_vg, w.cs.vg.p, w.vg_start.p);
        ADB_HIP(hipGetLastError());
        // 3. chunks of at most PCH contacts per pair
        hipLaunchKernelGGL(k_pair_chunks, dim3((np + 1 + 255) / 256), dim3(256), 0, st, np, npd, w.pair_start.p, w.chunk_n.p);
        ADB_HIP(hipGetLastError());
        {
            size_t tmp = 0;
            ADB_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, tmp, w.chunk_n.p, w.chunk_off.p, np + 1, st));
            w.sort_tmp.reserve(tmp + 16);
            ADB_HIP(hipcub::DeviceScan::ExclusiveSum(w.sort_tmp.p, tmp, w.chunk_n.p, w.chunk_off.p, np + 1, st));
        }
        hipLaunchKernelGGL(k_chunk_owner, dim3((np + 255) / 256 + 1), dim3(256), 0, st, np, npd, w.chunk_off.p, w.chunk_blk.p,
                           w.ccount.p);
        ADB_HIP(hipGetLastError());
    }
    groups_fetch(w, np, npd, one_wg ? w.pair_pack.p : nullptr, st);
    ADB_HIP(hipEventRecord(w.ev_asm_join, st));
    w.ap.pending = true; w.ap.serial = w.cs.serial;
    w.ap.nc = nc; w.ap.np = np; w.ap.npd = npd; w.ap.max_chunks = max_chunks;
}

// The values: inertia terms, then (after the pair structure and the contact
// records of the side streams) the chunk reduction, the off-diagonal and
// diagonal blocks and gradients, the host's groups and their inverses.
void assemble_system(World& w) {
    if (!w.ap.pending || w.ap.serial != w.cs.serial) assemble_structure(w, false);
    w.ap.pending = false;
    const int B = w.B, nc = w.ap.nc, np = w.ap.np, max_chunks = w.ap.max_chunks;
    const int* npd = w.ap.npd;
    w.bodyG.reserve(12 * (size_t)B);
    w.bodyH.reserve(144 * (size_t)B);
    hipLaunchKernelGGL(k_inertia, dim3((144 * B + 255) / 256), dim3(256), 0, w.stream, B, w.P.dt, w.P.kappa_ortho, w.P.gravity[0],
                       w.P.gravity[1], w.P.gravity[2], w.q.p, w.q0.p, w.qdot.p, w.mass.p, w.Ic.p, w.bodyG.p, w.bodyH.p);
    ADB_HIP(hipGetLastError());
    PairArgs a;
    a.np = np; a.npd = npd; a.n_pt = w.cs.n_pt; a.n_ee = w.cs.n_ee; a.B = B;
    a.cid = w.aval_s.p;
    a.pair_start = w.pair_start.p; a.pair_lo = w.pair_lo.p; a.pair_hi = w.pair_hi.p; a.pair_up = w.pair_up.p; a.pair_low = w.pair_low.p;
    a.up_start = w.up_start.p; a.low_start = w.low_start.p; a.tv = w.tval_s.p;
    a.chunk_off = w.chunk_off.p;
    a.pt = w.cs.pt.p; a.ee = w.cs.ee.p; a.ccls = w.ccls.p; a.mass = w.mass.p;
    a.rec = w.crec.p; a.bodyH = w.bodyH.p; a.bodyG = w.bodyG.p;
    a.vg = w.cs.vg.p; a.vg_start = w.vg_start.p; a.vgK = w.vgK.p; a.voff = w.voff.p; a.xbar = w.xbar.p;
    a.diag_blk = w.diag_blk.p;
    a.Hval = w.Hval.p; a.grad = w.grad.p;
    ADB_HIP(hipStreamWaitEvent(w.stream, w.ev_asm_join, 0));   // the pair structure
    if (w.side_pending) {   // records of the 4-point contacts and 3-point projections (side streams)
        ADB_HIP(hipStreamWaitEvent(w.stream, w.ev_side[1], 0));
        ADB_HIP(hipStreamWaitEvent(w.stream, w.ev_pk, 0));
        w.side_pending = false;
    }
    if (max_chunks > 0) {
        // every contact record read once (D = 3: 75 doubles, 6: 102, 9: 147), its id,
        // pair side and class; one partial (3 blocks + 2 gradients) written per chunk
        ProfScope ps(w, PK_ASM_REDUCE, 24.0 * nc);
        ps.counts(prof_snapshot(w, w.ccount.p), 600.0, 816.0, 1176.0, 8.0 * PART);
        hipLaunchKernelGGL(k_pair_reduce, dim3(max_chunks), dim3(256), 0, w.stream, a, w.chunk_blk.p, w.chunk_part.p);
        ADB_HIP(hipGetLastError());
    }
    if (np > 0) hipLaunchKernelGGL(k_pair_offdiag, dim3((144 * np + 255) / 256), dim3(256), 0, w.stream, a, w.chunk_part.p);
    hipLaunchKernelGGL(k_pair_diag, dim3((156 * B + 255) / 256), dim3(256), 0, w.stream, a, w.chunk_part.p);
    ADB_HIP(hipGetLastError());
    build_groups(w);
    {
        ProfScope ps(w, PK_BLOCK_CHOL, 0.0);
        if (w.prof.on) {
            double fl = 0.0;
            for (int g = 0; g < w.n_groups; g++) fl += group_inv_flops(w.h_grp_off[g + 1] - w.h_grp_off[g]);
            ps.flops = fl;
        }
        static_assert(GMAX == 8, "size classes 5..8 and 1..4");
        const int n4 = w.n_glist[1] + w.n_glist[2];
        const int nb = w.n_glist[0] + (n4 + 3) / 4;
        if (nb > 0)
            hipLaunchKernelGGL(k_group_inv, dim3(nb), dim3(256), 0, w.stream, w.glist, w.n_glist[0], n4, w.grp_off,
                               w.grp_body, w.grp_moff, w.row_ptr.p, w.col_idx.p, w.Hval.p, w.Ginv.p);
        ADB_HIP(hipGetLastError());
    }
}

} // namespace adb
