// pcg.hip - block-Jacobi preconditioned conjugate gradients on the BSR
// Hessian, FP64 throughout.  Replaces the direct PardisoLLT / SimplicialLLT
// solve of ABD::solve (model/abd.cpp:343-371): dq = -H^{-1} g.
//
// Per iteration three launches, no host round trip:
//   spmv   : y = H p (one 64-lane workgroup per block row, 5 x 12-lane groups
//            split the row's blocks, fixed-order LDS combine) + row partial p.y
//   update : every workgroup sums the row partials in the same fixed order
//            (deterministic), alpha = rz / pAp, x += alpha p, r -= alpha y,
//            z = L^-T L^-1 r per body, partials r.z and r.r
//   dir    : beta = rz' / rz, p = z + beta p
// Scalars are double-buffered by iteration parity.  The host reads r.r every
// CHECK iterations.
#include <cmath>
#include "dmath.h"
#include "world.h"

namespace adb {

constexpr int CHECK = 8;

__device__ __forceinline__ double sum_fixed(const double* v, int n) {
    // identical order in every workgroup: 4 interleaved accumulators then combine
    double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
    int i = 0;
    for (; i + 3 < n; i += 4) { s0 += v[i]; s1 += v[i + 1]; s2 += v[i + 2]; s3 += v[i + 3]; }
    for (; i < n; i++) s0 += v[i];
    return (s0 + s1) + (s2 + s3);
}
// LDS-cooperative version: each workgroup (256 lanes) computes the same total
__device__ double block_sum_fixed(const double* v, int n, double* sh) {
    const int t = threadIdx.x;
    double s = 0.0;
    for (int i = t; i < n; i += blockDim.x) s += v[i];
    sh[t] = s;
    __syncthreads();
    for (int m = blockDim.x / 2; m >= 1; m >>= 1) {
        if (t < m) sh[t] += sh[t + m];
        __syncthreads();
    }
    double r = sh[0];
    __syncthreads();
    return r;
}

__global__ __launch_bounds__(64) void k_spmv(int B, const int* __restrict__ row_ptr, const int* __restrict__ col_idx,
                                             const double* __restrict__ Hval, const double* __restrict__ p,
                                             double* __restrict__ y, double* __restrict__ part) {
    __shared__ double sh[5][12];
    const int row = blockIdx.x;
    if (row >= B) return;
    const int g = threadIdx.x / 12, r = threadIdx.x % 12;
    double acc = 0.0;
    if (g < 5) {
        for (int k = row_ptr[row] + g; k < row_ptr[row + 1]; k += 5) {
            const double* H = Hval + 144 * (size_t)k;
            const double* x = p + 12 * (size_t)col_idx[k];
            double s = 0.0;
#pragma unroll
            for (int c = 0; c < 12; c++) s += H[c * 12 + r] * x[c];
            acc += s;
        }
        sh[g][r] = acc;
    }
    __syncthreads();
    __shared__ double prod[12];
    if (threadIdx.x < 12) {
        const int rr = threadIdx.x;
        double yr = (((sh[0][rr] + sh[1][rr]) + sh[2][rr]) + sh[3][rr]) + sh[4][rr];
        y[12 * (size_t)row + rr] = yr;
        prod[rr] = p[12 * (size_t)row + rr] * yr;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0.0;
        for (int i = 0; i < 12; i++) t += prod[i];
        part[row] = t;
    }
}

// z_b = L^-T L^-1 r_b
__device__ void precond_body(const double* L, const double* r, double* z) {
    double t[12];
    for (int i = 0; i < 12; i++) {
        double s = r[i];
        for (int m = 0; m < i; m++) s -= L[m * 12 + i] * t[m];
        t[i] = s / L[i * 12 + i];
    }
    for (int i = 11; i >= 0; i--) {
        double s = t[i];
        for (int m = i + 1; m < 12; m++) s -= L[i * 12 + m] * z[m];
        z[i] = s / L[i * 12 + i];
    }
}

// initial: x = 0, r = -g (solve H dq = -g), z = M^-1 r, partials r.z, r.r, b.b
__global__ void k_pcg_init(int B, const double* __restrict__ g, const double* __restrict__ L, double* __restrict__ x,
                           double* __restrict__ r, double* __restrict__ z, double* __restrict__ p, double* __restrict__ part) {
    int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= B) return;
    double rb[12], zb[12];
    for (int i = 0; i < 12; i++) { rb[i] = -g[12 * (size_t)b + i]; x[12 * (size_t)b + i] = 0.0; r[12 * (size_t)b + i] = rb[i]; }
    precond_body(L + 144 * (size_t)b, rb, zb);
    double rz = 0.0, rr = 0.0;
    for (int i = 0; i < 12; i++) { z[12 * (size_t)b + i] = zb[i]; p[12 * (size_t)b + i] = zb[i]; rz += rb[i] * zb[i]; rr += rb[i] * rb[i]; }
    part[B + b] = rz;
    part[2 * B + b] = rr;
}
__global__ __launch_bounds__(256) void k_pcg_init_scal(int B, const double* __restrict__ part, double* __restrict__ scal) {
    __shared__ double sh[256];
    double rz = block_sum_fixed(part + B, B, sh);
    double rr = block_sum_fixed(part + 2 * B, B, sh);
    if (threadIdx.x == 0) { scal[0] = rz; scal[1] = rz; scal[2] = rr; scal[3] = rr; scal[4] = 0.0; }
}
// scal: [0..1] rz by parity, [2] rr, [3] bb, [4] breakdown flag
__global__ __launch_bounds__(256) void k_pcg_update(int B, int par, const double* __restrict__ L, const double* __restrict__ p,
                                                    const double* __restrict__ y, double* __restrict__ x, double* __restrict__ r,
                                                    double* __restrict__ z, double* __restrict__ part, double* __restrict__ scal) {
    __shared__ double sh[256];
    const double pAp = block_sum_fixed(part, B, sh);
    const double rz = scal[par];
    const bool stop = scal[4] != 0.0 || !(pAp > 0.0) || !(rz > 0.0);
    const double alpha = stop ? 0.0 : rz / pAp;
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= B) return;
    double rb[12], zb[12];
    for (int i = 0; i < 12; i++) {
        size_t k = 12 * (size_t)b + i;
        x[k] += alpha * p[k];
        rb[i] = r[k] - alpha * y[k];
        r[k] = rb[i];
    }
    precond_body(L + 144 * (size_t)b, rb, zb);
    double rzn = 0.0, rr = 0.0;
    for (int i = 0; i < 12; i++) { z[12 * (size_t)b + i] = zb[i]; rzn += rb[i] * zb[i]; rr += rb[i] * rb[i]; }
    part[B + b] = rzn;
    part[2 * B + b] = rr;
    if (b == 0 && stop) scal[4] = 1.0;
}
__global__ __launch_bounds__(256) void k_pcg_dir(int B, int par, const double* __restrict__ z, double* __restrict__ p,
                                                 const double* __restrict__ part, double* __restrict__ scal) {
    __shared__ double sh[256];
    const double rzn = block_sum_fixed(part + B, B, sh);
    const double rr = block_sum_fixed(part + 2 * B, B, sh);
    const double rz = scal[par];
    const bool stop = scal[4] != 0.0 || !(rz > 0.0);
    const double beta = stop ? 0.0 : rzn / rz;
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b < B && !stop)
        for (int i = 0; i < 12; i++) { size_t k = 12 * (size_t)b + i; p[k] = z[k] + beta * p[k]; }
    if (blockIdx.x == 0 && threadIdx.x == 0) { scal[par ^ 1] = rzn; scal[2] = rr; }
}

int pcg_solve(World& w, double* dq_out, int* iters) {
    const int B = w.B, n = 12 * B;
    w.pcg_x.reserve(n); w.pcg_r.reserve(n); w.pcg_z.reserve(n); w.pcg_p.reserve(n); w.pcg_Ap.reserve(n);
    w.partial.reserve(3 * (size_t)B); w.scal.reserve(8);
    const int nb = (B + 255) / 256;
    hipLaunchKernelGGL(k_pcg_init, dim3(nb), dim3(256), 0, w.stream, B, w.grad.p, w.Lchol.p, w.pcg_x.p, w.pcg_r.p, w.pcg_z.p, w.pcg_p.p, w.partial.p);
    hipLaunchKernelGGL(k_pcg_init_scal, dim3(1), dim3(256), 0, w.stream, B, w.partial.p, w.scal.p);
    ADB_HIP(hipGetLastError());
    double h[5];
    ADB_HIP(hipMemcpyAsync(h, w.scal.p, 5 * sizeof(double), hipMemcpyDeviceToHost, w.stream));
    ADB_HIP(hipStreamSynchronize(w.stream));
    const double bb = h[3];
    if (!std::isfinite(bb)) {   // NaN / inf in the gradient: the reference aborts (model/abd.cpp:368-371)
        if (iters) *iters = 0;
        return ADB_E_NAN;
    }
    const double tol2 = w.P.pcg_rtol * w.P.pcg_rtol * bb;
    int it = 0, par = 0;
    int status = ADB_OK;
    if (bb > 0.0) {
        while (it < w.P.pcg_max_iter) {
            for (int k = 0; k < CHECK; k++) {
                {
                    ProfScope ps(w, PK_SPMV, 1152.0 * w.nnzb + 4.0 * w.nnzb + 2.0 * 96.0 * B);
                    hipLaunchKernelGGL(k_spmv, dim3(B), dim3(64), 0, w.stream, B, w.row_ptr.p, w.col_idx.p, w.Hval.p, w.pcg_p.p, w.pcg_Ap.p, w.partial.p);
                }
                hipLaunchKernelGGL(k_pcg_update, dim3(nb), dim3(256), 0, w.stream, B, par, w.Lchol.p, w.pcg_p.p, w.pcg_Ap.p,
                                   w.pcg_x.p, w.pcg_r.p, w.pcg_z.p, w.partial.p, w.scal.p);
                hipLaunchKernelGGL(k_pcg_dir, dim3(nb), dim3(256), 0, w.stream, B, par, w.pcg_z.p, w.pcg_p.p, w.partial.p, w.scal.p);
                par ^= 1;
                it++;
            }
            ADB_HIP(hipGetLastError());
            ADB_HIP(hipMemcpyAsync(h, w.scal.p, 5 * sizeof(double), hipMemcpyDeviceToHost, w.stream));
            ADB_HIP(hipStreamSynchronize(w.stream));
            if (h[2] <= tol2 || h[4] != 0.0) break;
        }
        if (!(h[2] <= tol2) && h[4] == 0.0) status = ADB_E_NOT_CONVERGED;
    }
    if (iters) *iters = it;
    ADB_HIP(hipMemcpyAsync(dq_out, w.pcg_x.p, n * sizeof(double), hipMemcpyDeviceToDevice, w.stream));
    return status;
}

} // namespace adb
