// contact.hip - PT / EE / VG contact terms on gfx950.
//
// Replaces IPC::ipc_hess_pt_12x12 / ipc_hess_ee_12x12 (model/core_ipc.cpp:21-83),
// project_to_psd (model/psd.h:9-33), pt_uktk / ee_uktk (model/pt_ipc.cpp:22-107,
// model/ee_ipc.cpp:21-118), friction (model/friction.cpp:15-47), the J^T H J body
// mapping (model/pt_ipc.cpp:153-169, model/ee_ipc.cpp:175-192) and the VG term
// (model/vg_ipc.cpp:12-30, model/abd.cpp:380-401).
//
// World path (contact_terms): k_contact_classify computes each contact's
// distance type, squared distance and class, the dimension D = 3 n of the
// translation-free space of the n + 1 points it involves, and appends it to
// that class's list; then one thread per contact:
//   class 1 (two points, D = 3): closed-form projection (k_contact_small<1>);
//   class 2 (three points, D = 6): record fields (k_contact_scalars2), then
//     cyclic Jacobi on the packed 6 x 6 matrix in registers (k_contact_pk2);
//   class 3 (four points, or any mollified EE, D = 9): the 9 x 9 matrix into
//     the record, then cyclic Jacobi in registers (k_contact_c3r, nine
//     threads per contact building the matrix and a tenth the other record
//     fields, then k_contact_psd9, one thread per contact, on a side stream).
// Each contact writes a compact body-space record (world.h) that the
// assembly expands into its four 12 x 12 blocks; the class 1 and class 2
// record kernels stage it per wavefront in LDS and store it cooperatively
// (Stage, stage_flush).
//
// PSD projection: every barrier and mollifier term depends on the points only
// through differences, so the Hessian annihilates rigid translations.  With
// Q = h (x) I3, h the Helmert basis of the complement of (1, ..., 1) over the
// contact's points, H = Q^T (Q H Q^T) Q and the projection onto the PSD cone
// is Q^T P(Q H Q^T) Q: a D x D symmetric eigenproblem.  The kernel API, as in
// the reference, returns a Hessian whose eigenvalues are all >= 0 unchanged;
// the world path always rebuilds V diag(max(w, 0)) V^T, equal to it within
// the Jacobi tolerance (the reference's own test sees the nine rounding-level
// eigenvalues of the translations).
//
// Kernel-level API (k_hess_api, the abdtk functions): one 16-lane group per
// contact, four contacts per wavefront, lane j < 12 owning column j of the
// 12 x 12 Hessian; its 9 x 9 projection is parallel cyclic Jacobi (5 disjoint
// rotations per round, 9 rounds per sweep over 10 slots, slot 9 idle) on the
// matrix staged in LDS.
#include <cstdio>
#include <algorithm>
#include "dmath.h"
#include "world.h"

namespace adb {

__constant__ signed char kRR9[9][5][2] = {
    {{0, 9}, {1, 8}, {2, 7}, {3, 6}, {4, 5}}, {{0, 1}, {2, 9}, {3, 8}, {4, 7}, {5, 6}},
    {{0, 2}, {1, 3}, {4, 9}, {5, 8}, {6, 7}}, {{0, 3}, {2, 4}, {1, 5}, {6, 9}, {7, 8}},
    {{0, 4}, {3, 5}, {2, 6}, {1, 7}, {8, 9}}, {{0, 5}, {4, 6}, {3, 7}, {2, 8}, {1, 9}},
    {{0, 6}, {5, 7}, {4, 8}, {3, 9}, {1, 2}}, {{0, 7}, {6, 8}, {5, 9}, {1, 4}, {2, 3}},
    {{0, 8}, {7, 9}, {1, 6}, {2, 5}, {3, 4}}};

// Helmert basis of the complement of (1, 1, 1, 1) / 2 in R^4
#define HS2 0.70710678118654752440
#define HS6 0.40824829046386301637
#define HS12 0.28867513459481288225
__constant__ double kHel[3][4] = {{HS2, -HS2, 0.0, 0.0}, {HS6, HS6, -2.0 * HS6, 0.0}, {HS12, HS12, HS12, -3.0 * HS12}};

constexpr int GL = 16;        // lanes per contact
constexpr int CPW = 4;        // contacts per 64-lane workgroup
constexpr int LDH = 13;       // padded leading dimensions in LDS
constexpr int LDK = 11;
constexpr int MAX_SWEEPS = 15;
// Jacobi stops when the squared off-diagonal Frobenius norm is below this
// fraction of the total (off-diagonal entries ~1e-12 relative: rounding-level
// for the projected Hessian, and reachable in a few sweeps, unlike 1e-30,
// which sits at the FP64 noise floor of the rotations themselves)
constexpr double JACOBI_TOL = 1e-24;
// rotations of off-diagonal entries below JACOBI_SKIP * |A|_F are skipped: 36
// such entries add at most 36e-28 |A|_F^2 to the off-diagonal sum, far below
// the stop test, and late sweeps then cost little
constexpr double JACOBI_SKIP = 1e-14;
// Reciprocal and reciprocal square root from the hardware approximations,
// each refined by two Newton steps (each step doubles the correct bits).
// The Jacobi rotation parameters need no correct rounding, and this chain is
// several times shorter than IEEE division and square root.
__device__ __forceinline__ double rcp_r(double x) {
    double y = __builtin_amdgcn_rcp(x);
    double e = fma(-x, y, 1.0); y = fma(y, e, y);
    e = fma(-x, y, 1.0); return fma(y, e, y);
}
__device__ __forceinline__ double rsq_r(double x) {
    double y = __builtin_amdgcn_rsq(x);
    const double h = 0.5 * x;
    double e = fma(-h * y, y, 0.5); y = fma(y, e, y);
    e = fma(-h * y, y, 0.5); return fma(y, e, y);
}
// The same with one Newton step: where only the rotation angle depends on
// the value, an error there leaves a residual of that relative size in the
// annihilated entry (removed by the next sweep), not a non-orthogonal rotation.
__device__ __forceinline__ double rcp_r1(double x) {
    const double y = __builtin_amdgcn_rcp(x);
    return fma(y, fma(-x, y, 1.0), y);
}
__device__ __forceinline__ double rsq_r1(double x) {
    const double y = __builtin_amdgcn_rsq(x);
    return fma(y, fma(-(0.5 * x) * y, y, 0.5), y);
}
// Jacobi rotation annihilating a_pq (Golub and Van Loan, Algorithm 8.5.1):
// tau = (a_qq - a_pp) / (2 a_pq), t = sign(tau) / (|tau| + sqrt(1 + tau^2)),
// c = 1 / sqrt(1 + t^2), s = t c; after it a_pp' = a_pp - t a_pq, a_qq' = a_qq + t a_pq.
// t sets only the angle (one Newton step each); c keeps two, so that
// c^2 + s^2 = 1 to rounding and the accumulated rotations stay orthogonal.
__device__ __forceinline__ void jacobi_cs(double app, double aqq, double apq, double& c, double& s, double& t) {
    const double tau = (aqq - app) * (0.5 * rcp_r1(apq));
    if (fabs(tau) > 1e150) {
        t = 0.5 * rcp_r1(tau);
    } else {
        const double x = fma(tau, tau, 1.0);
        t = (tau >= 0 ? 1.0 : -1.0) * rcp_r1(fma(x, rsq_r1(x), fabs(tau)));      // x rsq(x) = sqrt(x)
    }
    c = rsq_r(fma(t, t, 1.0));
    s = t * c;
}

struct GroupLDS {
    double H[12][LDH];
    double A[9][LDK];
    double V[9][LDK];
    double cs[5][2];
    double g[12];
};

struct KParams {
    double d_hat, kappa, eps_x;
    int psd;
    double mu, eps_v, max_uk;
    int pt_fric, ee_fric;
};

__device__ __forceinline__ double sel12(const double a[12], int j) {
    double r = 0.0;
#pragma unroll
    for (int k = 0; k < 12; k++) r = (k == j) ? a[k] : r;
    return r;
}
__device__ __forceinline__ double sel4(double a0, double a1, double a2, double a3, int j) {
    return j == 0 ? a0 : (j == 1 ? a1 : (j == 2 ? a2 : a3));
}

// ---------------------------------------------------------------- PSD (group)
// Input: the symmetric vertex-space Hessian in S.H (all 64 lanes must call:
// barriers).  HOUT = false: S.H is projected in place, and left unchanged when
// the matrix is already PSD or `on` is false.  HOUT = true: S.A receives the
// Helmert-coordinate matrix K = Q H Q^T, replaced by its projection
// V diag(max(w, 0)) V^T when `on`.
// KIN = true: S.A already holds K (S.H is not read).
template <bool HOUT, bool KIN = false>
__device__ __forceinline__ void group_psd(GroupLDS& S, int lane, bool on) {
    // K = Q H Q^T: lane u = (b, j) < 9 builds column u
    if (!KIN && lane < 9) {
        const int b = lane / 3, j = lane % 3;
        double hb[4];
#pragma unroll
        for (int l = 0; l < 4; l++) hb[l] = kHel[b][l];
        double t[4][3];
#pragma unroll
        for (int k = 0; k < 4; k++)
#pragma unroll
            for (int i = 0; i < 3; i++) {
                double s = 0.0;
#pragma unroll
                for (int l = 0; l < 4; l++) s += hb[l] * S.H[3 * k + i][3 * l + j];
                t[k][i] = s;
            }
#pragma unroll
        for (int a = 0; a < 3; a++)
#pragma unroll
            for (int i = 0; i < 3; i++) {
                double s = 0.0;
#pragma unroll
                for (int k = 0; k < 4; k++) s += kHel[a][k] * t[k][i];
                S.A[3 * a + i][lane] = s;
            }
    }
    if (lane < 9) {
#pragma unroll
        for (int r = 0; r < 9; r++) S.V[r][lane] = (r == lane) ? 1.0 : 0.0;
    }
    __syncthreads();
    for (int sweep = 0; sweep < MAX_SWEEPS; sweep++) {
        double off = 0.0, tot = 0.0;
        if (lane < 9) {
#pragma unroll
            for (int r = 0; r < 9; r++) { double a = S.A[r][lane]; tot += a * a; if (r != lane) off += a * a; }
        }
        for (int m = 8; m >= 1; m >>= 1) { off += __shfl_xor(off, m, GL); tot += __shfl_xor(tot, m, GL); }
        const bool conv = !on || off <= JACOBI_TOL * tot || off == 0.0;
        if (__all(conv)) break;
        for (int rd = 0; rd < 9; rd++) {
            if (lane < 5) {
                const int p = kRR9[rd][lane][0], q = kRR9[rd][lane][1];
                double c = 1.0, s = 0.0;
                if (!conv && q < 9) {
                    const double apq = S.A[p][q];
                    if (apq != 0.0) {
                        const double app = S.A[p][p], aqq = S.A[q][q];
                        double t;
                        jacobi_cs(app, aqq, apq, c, s, t);
                    }
                }
                S.cs[lane][0] = c; S.cs[lane][1] = s;
            }
            __syncthreads();
            if (!conv && lane < 9) {          // A <- A J and V <- V J: lane owns row `lane`
#pragma unroll
                for (int k = 0; k < 5; k++) {
                    const int p = kRR9[rd][k][0], q = kRR9[rd][k][1];
                    if (q < 9) {
                        const double c = S.cs[k][0], s = S.cs[k][1];
                        const double ap = S.A[lane][p], aq = S.A[lane][q];
                        S.A[lane][p] = fma(c, ap, -s * aq); S.A[lane][q] = fma(s, ap, c * aq);
                        const double vp = S.V[lane][p], vq = S.V[lane][q];
                        S.V[lane][p] = fma(c, vp, -s * vq); S.V[lane][q] = fma(s, vp, c * vq);
                    }
                }
            }
            __syncthreads();
            if (!conv && lane < 9) {          // A <- J^T A: lane owns column `lane`
#pragma unroll
                for (int k = 0; k < 5; k++) {
                    const int p = kRR9[rd][k][0], q = kRR9[rd][k][1];
                    if (q < 9) {
                        const double c = S.cs[k][0], s = S.cs[k][1];
                        const double ap = S.A[p][lane], aq = S.A[q][lane];
                        S.A[p][lane] = fma(c, ap, -s * aq); S.A[q][lane] = fma(s, ap, c * aq);
                    }
                }
            }
            __syncthreads();
        }
    }
    double w[9];
    double wmin = 0.0;
#pragma unroll
    for (int k = 0; k < 9; k++) { w[k] = S.A[k][k]; wmin = k == 0 ? w[k] : fmin(wmin, w[k]); }
    const bool project = HOUT ? on : (on && wmin < 0.0);
    __syncthreads();
    // P(K) = V diag(max(w, 0)) V^T, column `lane`, into A
    if (project && lane < 9) {
#pragma unroll
        for (int r = 0; r < 9; r++) {
            double s = 0.0;
#pragma unroll
            for (int k = 0; k < 9; k++) s += S.V[r][k] * fmax(w[k], 0.0) * S.V[lane][k];
            S.A[r][lane] = s;
        }
    }
    __syncthreads();
    if (HOUT) return;
    // H_psd = Q^T P(K) Q: lane v = (l, j) < 12 builds column v
    if (project && lane < 12) {
        const int l = lane / 3, j = lane % 3;
        double s_ai[3][3];
#pragma unroll
        for (int a = 0; a < 3; a++)
#pragma unroll
            for (int i = 0; i < 3; i++) {
                double s = 0.0;
#pragma unroll
                for (int b = 0; b < 3; b++) s += kHel[b][l] * S.A[3 * a + i][3 * b + j];
                s_ai[a][i] = s;
            }
#pragma unroll
        for (int k = 0; k < 4; k++)
#pragma unroll
            for (int i = 0; i < 3; i++) {
                double s = 0.0;
#pragma unroll
                for (int a = 0; a < 3; a++) s += kHel[a][k] * s_ai[a][i];
                S.H[3 * k + i][lane] = s;
            }
    }
    __syncthreads();
}

// ---------------------------------------------------------------- barrier Hessian column
// Contact geometry and barrier scalars, computed redundantly by every lane of
// the group (no private arrays indexed by lane-dependent values).
struct Geo {
    d3 p0, p1, p2, p3;
    int is_ee, type;
    double d, B0, B1, B2;
    double m, m1, m2;       // EE mollifier m(x) and dm/dx, d2m/dx2
    RedMap map;
    RedDist rd;
    CrossNorm mc;           // x = |(a1 - a0) x (b1 - b0)|^2
};

__device__ __forceinline__ void geo_init(Geo& G, const KParams& kp) {
    const double dh = kp.d_hat, ka = kp.kappa;
    G.B1 = barrier_d1(G.d, dh, ka);
    G.B2 = barrier_d2(G.d, dh, ka);
    G.B0 = G.is_ee ? barrier(G.d, dh, ka) : 0.0;
    G.map = G.is_ee ? ee_map(G.type) : pt_map(G.type);
    d3 r0 = mk(0, 0, 0), r1 = mk(0, 0, 0), r2 = mk(0, 0, 0);
    const d3 p0 = G.p0, p1 = G.p1, p2 = G.p2, p3 = G.p3;
    if (!G.is_ee) {
        switch (G.type) {
        case P_T0: r0 = p0 - p1; break;
        case P_T1: r0 = p0 - p2; break;
        case P_T2: r0 = p0 - p3; break;
        case P_E0: r0 = p1 - p0; r1 = p2 - p0; break;
        case P_E1: r0 = p2 - p0; r1 = p3 - p0; break;
        case P_E2: r0 = p3 - p0; r1 = p1 - p0; break;
        default:   r0 = p0 - p1; r1 = p2 - p1; r2 = p3 - p1; break;
        }
    } else {
        switch (G.type) {
        case EA0_EB0: r0 = p0 - p2; break;
        case EA0_EB1: r0 = p0 - p3; break;
        case EA1_EB0: r0 = p1 - p2; break;
        case EA1_EB1: r0 = p1 - p3; break;
        case EA_EB0:  r0 = p0 - p2; r1 = p1 - p2; break;
        case EA_EB1:  r0 = p0 - p3; r1 = p1 - p3; break;
        case EA0_EB:  r0 = p2 - p0; r1 = p3 - p0; break;
        case EA1_EB:  r0 = p2 - p1; r1 = p3 - p1; break;
        default:      r0 = p2 - p0; r1 = p1 - p0; r2 = p3 - p2; break;
        }
    }
    G.rd.init(G.map.kind, r0, r1, r2);
    G.m = 1.0; G.m1 = 0.0; G.m2 = 0.0;
    if (G.is_ee) {
        const double eps = kp.eps_x * sqn(p0 - p1) * sqn(p2 - p3);
        G.mc.init(p1 - p0, p3 - p2);
        const double x = G.mc.value();
        if (x < eps) {
            G.m = (-x / eps + 2) * x / eps;
            G.m1 = 2.0 / eps - 2.0 * x / (eps * eps);
            G.m2 = -2.0 / (eps * eps);
        }
    }
}
// point-space gradient entry (k, i) of d, and of the mollifier
__device__ __forceinline__ double grad_d(const Geo& G, int k, int i) {
    double s = 0.0;
#pragma unroll
    for (int m = 0; m < 3; m++)
        if (m < G.map.nr) s += red_sign(G.map, m, k) * cmp(G.rd.g[m], i);
    return s;
}
__device__ __forceinline__ double grad_m(const Geo& G, int k, int i) {
    if (G.m2 == 0.0) return 0.0;
    // x depends on y = p1 - p0 (points 1, 0) and z = p3 - p2 (points 3, 2)
    const double sy = (k == 1 ? 1.0 : 0.0) - (k == 0 ? 1.0 : 0.0);
    const double sz = (k == 3 ? 1.0 : 0.0) - (k == 2 ? 1.0 : 0.0);
    return G.m1 * (sy * G.mc.grad(0, i) + sz * G.mc.grad(1, i));
}
// column j of the barrier Hessian (model/core_ipc.cpp:36-43,70-80) and g_j
__device__ __forceinline__ void barrier_column(const Geo& G, int j, double (*H)[LDH], double& gj) {
    const int l = j / 3, c = j % 3;
    const double gdj = grad_d(G, l, c), gmj = grad_m(G, l, c);
    const double syl = (l == 1 ? 1.0 : 0.0) - (l == 0 ? 1.0 : 0.0), szl = (l == 3 ? 1.0 : 0.0) - (l == 2 ? 1.0 : 0.0);
    const double gxj = syl * G.mc.grad(0, c) + szl * G.mc.grad(1, c);
#pragma unroll 1
    for (int k = 0; k < 4; k++)
#pragma unroll
        for (int i = 0; i < 3; i++) {
            // d^2 d / dx_(k,i) dx_(l,c) from the reduced Hessian
            double hd = 0.0;
#pragma unroll
            for (int m = 0; m < 3; m++)
#pragma unroll
                for (int n = 0; n < 3; n++)
                    if (m < G.map.nr && n < G.map.nr) {
                        const double sg = red_sign(G.map, m, k) * red_sign(G.map, n, l);
                        if (sg != 0.0) hd += sg * G.rd.hess(m, i, n, c);
                    }
            const double gdi = grad_d(G, k, i);
            if (!G.is_ee) {
                H[3 * k + i][j] = hd * G.B1 + gdi * gdj * G.B2;
            } else {
                double hm = 0.0, gmi = 0.0;
                if (G.m2 != 0.0) {
                    const double syk = (k == 1 ? 1.0 : 0.0) - (k == 0 ? 1.0 : 0.0), szk = (k == 3 ? 1.0 : 0.0) - (k == 2 ? 1.0 : 0.0);
                    const double hx = syk * syl * G.mc.hess(0, i, 0, c) + syk * szl * G.mc.hess(0, i, 1, c)
                                    + szk * syl * G.mc.hess(1, i, 0, c) + szk * szl * G.mc.hess(1, i, 1, c);
                    const double gxi = syk * G.mc.grad(0, i) + szk * G.mc.grad(1, i);
                    hm = G.m2 * gxi * gxj + G.m1 * hx;
                    gmi = G.m1 * gxi;
                }
                H[3 * k + i][j] = hm * G.B0 + G.B1 * (gmi * gdj + gdi * gmj) + G.m * (G.B2 * gdi * gdj + G.B1 * hd);
            }
        }
    gj = G.is_ee ? G.m * gdj * G.B1 + gmj * G.B0 : gdj * G.B1;
}

// vertex-space gradient entry j = (point, component) < 12 (model/core_ipc.cpp:36-43,70-80)
__device__ __forceinline__ double barrier_grad(const Geo& G, int j) {
    const int l = j / 3, c = j % 3;
    const double gdj = grad_d(G, l, c);
    return G.is_ee ? G.m * gdj * G.B1 + grad_m(G, l, c) * G.B0 : gdj * G.B1;
}

// ---------------------------------------------------------------- friction setup
__device__ __forceinline__ d3 unit(d3 a) { double l = sqrt(sqn(a)); return mk(a.x / l, a.y / l, a.z / l); }
__device__ __forceinline__ void basis_pp(d3 p0, d3 p1, d3& c0, d3& c1) {
    d3 d = p1 - p0;
    d3 cx = cross(mk(1, 0, 0), d), cy = cross(mk(0, 1, 0), d);
    if (sqn(cx) > sqn(cy)) { c0 = unit(cx); c1 = unit(cross(d, cx)); }
    else { c0 = unit(cy); c1 = unit(cross(d, cy)); }
}
__device__ __forceinline__ void basis_pe(d3 p, d3 e0, d3 e1, d3& c0, d3& c1) { d3 e = e1 - e0; c0 = unit(e); c1 = unit(cross(e, p - e0)); }
__device__ __forceinline__ double pe_closest(d3 p, d3 e0, d3 e1) { d3 e = e1 - e0; return dot(p - e0, e) / sqn(e); }
__device__ __forceinline__ void solve2(double m00, double m01, double m11, double r0, double r1, double& x, double& y) {
    double det = m00 * m11 - m01 * m01;
    x = (r0 * m11 - m01 * r1) / det;
    y = (m00 * r1 - m01 * r0) / det;
}
// coefficients g0..g3 of the relative tangential displacement sum_k g_k x_k
// and the tangent basis (c0, c1) (model/pt_ipc.cpp:31-103, model/ee_ipc.cpp:38-110)
__device__ __forceinline__ void friction_frame(int is_ee, int type, d3 P0, d3 P1, d3 P2, d3 P3,
                                               double& g0, double& g1, double& g2, double& g3, d3& c0, d3& c1) {
    if (!is_ee) {
        double t0, t1, t2;
        switch (type) {
        case P_T0: basis_pp(P0, P1, c0, c1); t0 = 1; t1 = 0; t2 = 0; break;
        case P_T1: basis_pp(P0, P2, c0, c1); t0 = 0; t1 = 1; t2 = 0; break;
        case P_T2: basis_pp(P0, P3, c0, c1); t0 = 0; t1 = 0; t2 = 1; break;
        case P_E0: { double l = pe_closest(P0, P1, P2); t0 = 1 - l; t1 = l; t2 = 0; basis_pe(P0, P1, P2, c0, c1); } break;
        case P_E1: { double l = pe_closest(P0, P2, P3); t0 = 0; t1 = 1 - l; t2 = l; basis_pe(P0, P2, P3, c0, c1); } break;
        case P_E2: { double l = pe_closest(P0, P3, P1); t0 = l; t1 = 0; t2 = 1 - l; basis_pe(P0, P3, P1, c0, c1); } break;
        default: {
            d3 b0 = P2 - P1, b1 = P3 - P1, r = P0 - P1;
            double x, y; solve2(dot(b0, b0), dot(b0, b1), dot(b1, b1), dot(b0, r), dot(b1, r), x, y);
            t0 = 1 - x - y; t1 = x; t2 = y;
            d3 n = cross(b0, b1);
            c0 = unit(b0); c1 = unit(cross(n, b0));
        }
        }
        g0 = -1.0; g1 = t0; g2 = t1; g3 = t2;
    } else {
        double l0, l1;
        switch (type) {
        case EA0_EB0: basis_pp(P0, P2, c0, c1); l0 = 0; l1 = 0; break;
        case EA0_EB1: basis_pp(P0, P3, c0, c1); l0 = 0; l1 = 1; break;
        case EA1_EB0: basis_pp(P1, P2, c0, c1); l0 = 1; l1 = 0; break;
        case EA1_EB1: basis_pp(P1, P3, c0, c1); l0 = 1; l1 = 1; break;
        case EA_EB0: l0 = pe_closest(P2, P0, P1); l1 = 0; basis_pe(P2, P0, P1, c0, c1); break;
        case EA_EB1: l0 = pe_closest(P3, P0, P1); l1 = 1; basis_pe(P3, P0, P1, c0, c1); break;
        case EA0_EB: l0 = 0; l1 = pe_closest(P0, P2, P3); basis_pe(P0, P2, P3, c0, c1); break;
        case EA1_EB: l0 = 1; l1 = pe_closest(P1, P2, P3); basis_pe(P1, P2, P3, c0, c1); break;
        default: {
            d3 w = P0 - P2, ea = P1 - P0, eb = P3 - P2;
            solve2(sqn(ea), -dot(eb, ea), sqn(eb), -dot(w, ea), dot(w, eb), l0, l1);
            d3 n = cross(ea, eb);
            c0 = unit(ea); c1 = unit(cross(n, ea));
        }
        }
        l0 = fmax(fmin(1.0, l0), 0.0);
        l1 = fmax(fmin(1.0, l1), 0.0);
        g0 = -(1 - l0); g1 = -l0; g2 = 1 - l1; g3 = l1;
    }
}

// 2x2 symmetric PSD projection (closed-form eigenpairs)
__device__ __forceinline__ void psd2(double a, double b, double c, double& o0, double& o1, double& o2) {
    double h = 0.5 * (a + c), dlt = 0.5 * (a - c);
    double r = sqrt(dlt * dlt + b * b);
    double lmin = h - r, lmax = h + r;
    if (lmin >= 0.0) { o0 = a; o1 = b; o2 = c; return; }
    if (lmax <= 0.0) { o0 = 0; o1 = 0; o2 = 0; return; }
    double vx, vy;
    if (dlt >= 0) { vx = lmax - c; vy = b; } else { vx = b; vy = lmax - a; }
    double n = sqrt(vx * vx + vy * vy);
    if (n == 0.0) { vx = 1; vy = 0; } else { vx /= n; vy /= n; }
    o0 = lmax * vx * vx; o1 = lmax * vx * vy; o2 = lmax * vy * vy;
}
// friction 2x2 middle matrix M (model/friction.cpp:53-74) and force scale
__device__ __forceinline__ void friction_M(double u0, double u1, double lam, const KParams& kp, double& fs, double M[3]) {
    double u = sqrt(u0 * u0 + u1 * u1);
    double f1 = f1_over_x(u, kp.eps_v);
    fs = kp.mu * lam * f1;
    if (u >= kp.eps_v) {
        double k = kp.mu * lam * f1 / (u * u);
        double t0 = -u1, t1 = u0;
        M[0] = k * t0 * t0; M[1] = k * t0 * t1; M[2] = k * t1 * t1;
    } else if (u <= kp.max_uk) {
        double k = kp.mu * lam * f1;
        M[0] = k; M[1] = 0; M[2] = k;
    } else {
        double k = (-1.0 / (kp.eps_v * kp.eps_v)) / u;
        double m0 = (k * u0 * u0 + f1) * kp.mu * lam, m1 = (k * u1 * u0) * kp.mu * lam, m2 = (k * u1 * u1 + f1) * kp.mu * lam;
        psd2(m0, m1, m2, M[0], M[1], M[2]);
    }
}

// ---------------------------------------------------------------- kernel-level API
// H and g of IPC::ipc_hess_{pt,ee}_12x12 for n independent inputs.
template <int IS_EE>
__global__ __launch_bounds__(64) void k_hess_api(KParams kp, int n, const double* __restrict__ x,
                                                   const int* __restrict__ type, const double* __restrict__ dist,
                                                   double* __restrict__ H, double* __restrict__ g, double* __restrict__ moll) {
    __shared__ GroupLDS lds[CPW];
    const int grp = threadIdx.x / GL, lane = threadIdx.x % GL;
    const int e = blockIdx.x * CPW + grp;
    const bool valid = e < n;
    GroupLDS& L = lds[grp];
    Geo G;
    double gj = 0.0;
    double mval = 1.0;
    if (valid) {
        const double* xe = x + 12 * (size_t)e;
        G.p0 = ld3(xe); G.p1 = ld3(xe + 3); G.p2 = ld3(xe + 6); G.p3 = ld3(xe + 9);
        G.is_ee = IS_EE; G.type = type[e]; G.d = dist[e];
        geo_init(G, kp);
        mval = G.m;
        if (lane < 12) barrier_column(G, lane, L.H, gj);
    } else if (lane < 12) {
        for (int i = 0; i < 12; i++) L.H[i][lane] = 0.0;
    }
    __syncthreads();
    if (kp.psd) group_psd<false>(L, lane, valid);
    if (valid && lane < 12) {
        for (int i = 0; i < 12; i++) H[144 * (size_t)e + 12 * lane + i] = L.H[i][lane];
        g[12 * (size_t)e + lane] = gj;
        if (IS_EE && lane == 0 && moll) moll[e] = mval;
    }
}

// ---------------------------------------------------------------- world contact terms
// Contact ids: the four world vertex ids of contact `cid` (PT: point, face
// vertices; EE: edge a, edge b).  Points 0..nside0-1 belong to body I.
struct WorldArgs {
    KParams kp;
    int n_pt, n_ee;
    const int4* pt; const int4* ee;
    const int* voff; const int* foff; const int* eoff; const int* face; const int* edge;
    const double* X; const double* Xs; const double* xbar; const double* mass;
    int* ctype; double* cdist; int* ccls; int* list; int* list3; int* counts;
    double* rec; double* tk; double* lam;
};

// R[R_D] = D + 100 (dynamic(I) + 2 dynamic(J)): the assembly drops the
// blocks of static bodies (mass < 0, model/assemble_matrices.cpp:43-58)
__device__ __forceinline__ double rec_code(const WorldArgs& a, int cid, int D) {
    const int4 c = cid < a.n_pt ? a.pt[cid] : a.ee[cid - a.n_pt];
    const int di = a.mass[c.x] > 0.0 ? 1 : 0, dj = a.mass[c.z] > 0.0 ? 1 : 0;
    return (double)(D + 100 * (di + 2 * dj));
}
__device__ __forceinline__ bool contact_ids(const WorldArgs& a, int cid, int& id0, int& id1, int& id2, int& id3) {
    if (cid < a.n_pt) {
        const int4 c = a.pt[cid];
        const int f = a.foff[c.z] + c.w;
        id0 = a.voff[c.x] + c.y; id1 = a.face[3 * f]; id2 = a.face[3 * f + 1]; id3 = a.face[3 * f + 2];
        return false;
    }
    const int4 c = a.ee[cid - a.n_pt];
    const int ea = a.eoff[c.x] + c.y, eb = a.eoff[c.z] + c.w;
    id0 = a.edge[2 * ea]; id1 = a.edge[2 * ea + 1]; id2 = a.edge[2 * eb]; id3 = a.edge[2 * eb + 1];
    return true;
}

// Class of an active contact by the points its barrier depends on:
// 1 = two points (D = 3), 2 = three points (D = 6), 3 = four points (D = 9,
// also every mollified edge-edge contact).  Inactive contacts (d >= d_hat)
// get an empty record here.
__device__ __forceinline__ int classify_one(const WorldArgs& a, int cid) {
    int id0, id1, id2, id3;
    const bool is_ee = contact_ids(a, cid, id0, id1, id2, id3);
    const d3 p0 = ld3(a.X + 3 * (size_t)id0), p1 = ld3(a.X + 3 * (size_t)id1);
    const d3 p2 = ld3(a.X + 3 * (size_t)id2), p3 = ld3(a.X + 3 * (size_t)id3);
    int type; double d;
    if (!is_ee) { type = pt_type(p0, p1, p2, p3); d = pt_dist(p0, p1, p2, p3, type); }
    else { type = ee_type(p0, p1, p2, p3); d = ee_dist(p0, p1, p2, p3, type); }
    a.ctype[cid] = type; a.cdist[cid] = d;
    int cls = 0;
    if (d < a.kp.d_hat) {
        const int kind = is_ee ? ee_map(type).kind : pt_map(type).kind;
        cls = kind == K_PP ? 1 : (kind == K_PL ? 2 : 3);
        if (is_ee && cls < 3) {
            // mollifier active: x = |(p1 - p0) x (p3 - p2)|^2 < eps_x |p0 - p1|^2 |p2 - p3|^2
            const double eps = a.kp.eps_x * sqn(p0 - p1) * sqn(p2 - p3);
            const d3 c = cross(p1 - p0, p3 - p2);
            if (dot(c, c) < eps) cls = 3;
        }
    } else {
        a.rec[REC * (size_t)cid + R_D] = 0.0;
        double* tk = a.tk + TKN * (size_t)cid;
        for (int i = 0; i < TKN; i++) tk[i] = 0.0;
        a.lam[cid] = 0.0;
    }
    a.ccls[cid] = cls;
    return cls;
}

// Classification and the per-class lists (class 1 from the front of `list`,
// class 2 from its back, class 3 in `list3`).  Device-scope atomics on one
// address serialize (~11 ns each on gfx950), so the slots are claimed per
// workgroup: each workgroup classifies a tile of rpt x CL_THREADS contacts,
// counts its classes with LDS atomics (one per wavefront and class), takes
// one global range per class, and then writes the list entries.
constexpr int CL_THREADS = 1024;
constexpr int CL_RPT_MAX = 8;
__global__ __launch_bounds__(CL_THREADS) void k_contact_classify(WorldArgs a, int rpt) {
    __shared__ int lcnt[3], lbase[3];
    const int total = a.n_pt + a.n_ee;
    const int lane = __lane_id();
    if (threadIdx.x < 3) lcnt[threadIdx.x] = 0;
    __syncthreads();
    int cls_r[CL_RPT_MAX], pos_r[CL_RPT_MAX];
    const int base = blockIdx.x * CL_THREADS * rpt;
#pragma unroll
    for (int r = 0; r < CL_RPT_MAX; r++) {
        cls_r[r] = 0; pos_r[r] = 0;
        if (r < rpt) {
            const int cid = base + r * CL_THREADS + threadIdx.x;
            const int cls = cid < total ? classify_one(a, cid) : 0;
            cls_r[r] = cls;
#pragma unroll
            for (int k = 1; k <= 3; k++) {
                const unsigned long long m = __ballot(cls == k);
                if (m == 0ull) continue;
                const int leader = __ffsll((long long)m) - 1;
                int b = 0;
                if (lane == leader) b = atomicAdd(&lcnt[k - 1], __popcll(m));
                b = __shfl(b, leader);
                if (cls == k) pos_r[r] = b + __popcll(m & ((1ull << lane) - 1ull));
            }
        }
    }
    __syncthreads();
    if (threadIdx.x < 3) lbase[threadIdx.x] = lcnt[threadIdx.x] ? atomicAdd(&a.counts[threadIdx.x], lcnt[threadIdx.x]) : 0;
    __syncthreads();
#pragma unroll
    for (int r = 0; r < CL_RPT_MAX; r++) {
        if (r >= rpt) break;
        const int cid = base + r * CL_THREADS + threadIdx.x;
        const int cls = cls_r[r];
        if (cls == 1) a.list[lbase[0] + pos_r[r]] = cid;
        else if (cls == 2) a.list[total - 1 - (lbase[1] + pos_r[r])] = cid;
        else if (cls == 3) a.list3[lbase[2] + pos_r[r]] = cid;
    }
}

// Friction of one contact (model/friction.cpp:15-47) with the tangential
// displacement measured from the step start.
struct Fric {
    double gm0, gm1, gm2, gm3;
    d3 c0, c1;
    double u0, u1, lam, fs, M[3];
    bool on;
};
__device__ __forceinline__ void friction_eval(const WorldArgs& a, int is_ee, int type, double d, double B1,
                                              d3 P0, d3 P1, d3 P2, d3 P3, int id0, int id1, int id2, int id3, Fric& f) {
    friction_frame(is_ee, type, P0, P1, P2, P3, f.gm0, f.gm1, f.gm2, f.gm3, f.c0, f.c1);
    const d3 d0 = P0 - ld3(a.Xs + 3 * (size_t)id0), d1 = P1 - ld3(a.Xs + 3 * (size_t)id1);
    const d3 d2 = P2 - ld3(a.Xs + 3 * (size_t)id2), d3v = P3 - ld3(a.Xs + 3 * (size_t)id3);
    f.u0 = f.gm0 * dot(f.c0, d0) + f.gm1 * dot(f.c0, d1) + f.gm2 * dot(f.c0, d2) + f.gm3 * dot(f.c0, d3v);
    f.u1 = f.gm0 * dot(f.c1, d0) + f.gm1 * dot(f.c1, d1) + f.gm2 * dot(f.c1, d2) + f.gm3 * dot(f.c1, d3v);
    f.lam = -B1 * 2 * sqrt(d);
    f.on = is_ee ? a.kp.ee_fric : a.kp.pt_fric;
    f.fs = 0.0; f.M[0] = 0.0; f.M[1] = 0.0; f.M[2] = 0.0;
    if (f.on) friction_M(f.u0, f.u1, f.lam, a.kp, f.fs, f.M);
}
// F[i][j] = sum_{r, q} M_rq c_r[i] c_q[j]
__device__ __forceinline__ double fric_F(const Fric& f, int i, int j) {
    const double c0i = cmp(f.c0, i), c1i = cmp(f.c1, i), c0j = cmp(f.c0, j), c1j = cmp(f.c1, j);
    return f.M[0] * c0i * c0j + f.M[1] * (c0i * c1j + c1i * c0j) + f.M[2] * c1i * c1j;
}

// Record fields shared by the per-thread and per-group paths, emitted as
// put(field, value).  Friction part: F and G of the record; the lagged basis
// (gm_k, c0, c1: TKN doubles); lambda is stored by the caller.
template <class Put>
__device__ __forceinline__ void fric_fields(const Fric& f, int nside0, const double w[4][4], Put put) {
    const double gmv[4] = {f.gm0, f.gm1, f.gm2, f.gm3};
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++) put(R_F + 3 * i + j, fric_F(f, i, j));
#pragma unroll
    for (int s = 0; s < 2; s++)
#pragma unroll
        for (int al = 0; al < 4; al++) {
            double g = 0.0;
#pragma unroll
            for (int k = 0; k < 4; k++)
                if ((k < nside0) == (s == 0)) g += gmv[k] * w[k][al];
            put(R_G + 4 * s + al, g);
        }
}
template <class Put>
__device__ __forceinline__ void tk_fields(const Fric& f, Put put) {
    const double gmv[4] = {f.gm0, f.gm1, f.gm2, f.gm3};
#pragma unroll
    for (int k = 0; k < 4; k++) put(k, gmv[k]);
#pragma unroll
    for (int r = 0; r < 3; r++) { put(4 + r, cmp(f.c0, r)); put(7 + r, cmp(f.c1, r)); }
}
// Body-space gradient from per-point vertex gradients gv_k (friction included).
template <class Put>
__device__ __forceinline__ void grad_fields(int nside0, const double w[4][4], const d3 gv[4], Put put) {
#pragma unroll
    for (int s = 0; s < 2; s++)
#pragma unroll
        for (int al = 0; al < 4; al++)
#pragma unroll
            for (int r = 0; r < 3; r++) {
                double g = 0.0;
#pragma unroll
                for (int k = 0; k < 4; k++)
                    if ((k < nside0) == (s == 0)) g += w[k][al] * cmp(gv[k], r);
                put(R_GRAD + 12 * s + 3 * al + r, g);
            }
}
__device__ __forceinline__ void write_fric(const WorldArgs& a, int cid, int nside0, const Fric& f, const double w[4][4]) {
    double* R = a.rec + REC * (size_t)cid;
    fric_fields(f, nside0, w, [&](int k, double v) { R[k] = v; });
    double* tk = a.tk + TKN * (size_t)cid;
    tk_fields(f, [&](int k, double v) { tk[k] = v; });
    a.lam[cid] = f.lam;
}
__device__ __forceinline__ void write_grad(const WorldArgs& a, int cid, int nside0, const double w[4][4], const d3 gv[4]) {
    double* R = a.rec + REC * (size_t)cid;
    grad_fields(nside0, w, gv, [&](int k, double v) { R[k] = v; });
}

// Record sections staged per wavefront.  A contact kernel with one thread per
// record would store each field with 64 lanes on 64 different cache lines;
// instead lane l puts field k of a section into S[k][l] (element-major, row
// padded to 65 doubles), and stage_flush stores the section's NF consecutive
// doubles of each of the wavefront's records with consecutive lanes on
// consecutive (record, field) pairs: a store instruction then covers a few
// lines.  Lanes past the end of the contact list compute a duplicate of the
// last contact (nothing of theirs is stored), so every flush runs with the
// whole wavefront active.
constexpr int SNF = 24, SLD = 65;
struct Stage {
    double (*S)[SLD];
    int lane, cid, nvalid;
};
__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}
// base[stride * cid_r + off + k] = S[k][r] for records r < nvalid, fields k < NF,
// as 16-byte stores of field pairs (the texture address unit's work grows with
// the number of store instructions, not with the cache lines they cover; off,
// stride and record bases are even in doubles, so every pair is aligned)
template <int NF>
__device__ __forceinline__ void stage_flush(const Stage& st, double* base, int stride, int off) {
    static_assert(NF <= SNF && NF % 2 == 0, "sections of an even number of fields within the staging rows");
    constexpr int NP = NF / 2;
    wave_lds_sync();
#pragma unroll 1
    for (int it = 0; it < NP; it++) {
        const int t = it * 64 + st.lane;      // < 64 NP, so r < 64
        const int r = t / NP, jp = t - r * NP;
        const int c = __shfl(st.cid, r);
        if (r < st.nvalid) {
            double2 v;
            v.x = st.S[2 * jp][r];
            v.y = st.S[2 * jp + 1][r];
            *reinterpret_cast<double2*>(base + (size_t)stride * c + off + 2 * jp) = v;
        }
    }
    wave_lds_sync();
}
__device__ __forceinline__ void load_w(const double* xbar, int id0, int id1, int id2, int id3, double w[4][4]) {
    const d3 x0 = ld3(xbar + 3 * (size_t)id0), x1 = ld3(xbar + 3 * (size_t)id1);
    const d3 x2 = ld3(xbar + 3 * (size_t)id2), x3 = ld3(xbar + 3 * (size_t)id3);
    const d3 xs[4] = {x0, x1, x2, x3};
#pragma unroll
    for (int k = 0; k < 4; k++) { w[k][0] = 1.0; w[k][1] = xs[k].x; w[k][2] = xs[k].y; w[k][3] = xs[k].z; }
}

// Point slot k of the contact as a selector over the four points.
__device__ __forceinline__ d3 pick(d3 p0, d3 p1, d3 p2, d3 p3, int k) {
    return k == 0 ? p0 : (k == 1 ? p1 : (k == 2 ? p2 : p3));
}

// Class 1 contacts, one thread each: d = |r|^2, r = x_a - x_b.  In the
// Helmert coordinate y = (x_a - x_b) / sqrt(2) the Hessian is
// K = 4 B1 I + 8 B2 r r^T with eigenvalues 4 B1 (twice, normal to r) and
// 4 B1 + 8 B2 d (along r).
__device__ __forceinline__ void contact_small(const WorldArgs& a, const Stage& st, bool valid) {
    const int cid = st.cid, l = st.lane;
    double (*S)[SLD] = st.S;
    int id0, id1, id2, id3;
    const int is_ee = contact_ids(a, cid, id0, id1, id2, id3);
    const d3 P0 = ld3(a.X + 3 * (size_t)id0), P1 = ld3(a.X + 3 * (size_t)id1);
    const d3 P2 = ld3(a.X + 3 * (size_t)id2), P3 = ld3(a.X + 3 * (size_t)id3);
    const int type = a.ctype[cid];
    const double d = a.cdist[cid];
    const RedMap mp = is_ee ? ee_map(type) : pt_map(type);
    const int nside0 = is_ee ? 2 : 1;
    const double B1 = barrier_d1(d, a.kp.d_hat, a.kp.kappa);
    const double B2 = barrier_d2(d, a.kp.d_hat, a.kp.kappa);
    const d3 z = mk(0, 0, 0);
    double w[4][4];
    load_w(a.xbar, id0, id1, id2, id3, w);
    // friction first: D, F, G (fields 0..17), the lagged basis and lambda
    // stored now; kept: gm_k and ft = fs (u0 c0 + u1 c1), the friction
    // gradient of point k being gm_k ft
    double gmv[4];
    d3 ft = z;
    {
        Fric f;
        friction_eval(a, is_ee, type, d, B1, P0, P1, P2, P3, id0, id1, id2, id3, f);
        S[R_D][l] = rec_code(a, cid, 3);
        fric_fields(f, nside0, w, [&](int k, double v) { S[k][l] = v; });
        stage_flush<R_Z>(st, a.rec, REC, 0);
        tk_fields(f, [&](int k, double v) { S[k][l] = v; });
        stage_flush<TKN>(st, a.tk, TKN, 0);
        if (valid) a.lam[cid] = f.lam;
        gmv[0] = f.gm0; gmv[1] = f.gm1; gmv[2] = f.gm2; gmv[3] = f.gm3;
        if (f.on) {
            const double t0 = f.fs * f.u0, t1 = f.fs * f.u1;
            ft = mk(t0 * f.c0.x + t1 * f.c1.x, t0 * f.c0.y + t1 * f.c1.y, t0 * f.c0.z + t1 * f.c1.z);
        }
    }
    d3 gv[4];
    double hz[4];         // Helmert weight of point slot k (zero off the contact's slots)
    {
        const int ka = mp.plus[0], kb = mp.minus[0];
        const d3 r = pick(P0, P1, P2, P3, ka) - pick(P0, P1, P2, P3, kb);
        const double lp = 4.0 * B1, ll = 4.0 * B1 + 8.0 * B2 * d;
        if (a.kp.psd) {
            const double cp = fmax(lp, 0.0), cl = fmax(ll, 0.0);
#pragma unroll
            for (int i = 0; i < 3; i++)
#pragma unroll
                for (int j = 0; j < 3; j++) {
                    const double rr = cmp(r, i) * cmp(r, j) / d;
                    S[3 * i + j][l] = cl * rr + cp * ((i == j ? 1.0 : 0.0) - rr);
                }
        } else {
#pragma unroll
            for (int i = 0; i < 3; i++)
#pragma unroll
                for (int j = 0; j < 3; j++) S[3 * i + j][l] = (i == j ? lp : 0.0) + 8.0 * B2 * cmp(r, i) * cmp(r, j);
        }
        S[9][l] = 0.0;      // pad to a field pair: R_PK + 9 is past a 2-point record
        stage_flush<10>(st, a.rec, REC, R_PK);
#pragma unroll
        for (int k = 0; k < 4; k++) hz[k] = k == ka ? HS2 : (k == kb ? -HS2 : 0.0);
        const d3 g = mk(2 * r.x * B1, 2 * r.y * B1, 2 * r.z * B1);
#pragma unroll
        for (int k = 0; k < 4; k++) gv[k] = k == ka ? g : (k == kb ? mk(-g.x, -g.y, -g.z) : z);
    }
#pragma unroll
    for (int s = 0; s < 2; s++)
#pragma unroll
        for (int al = 0; al < 4; al++) {
            double zz = 0.0;
#pragma unroll
            for (int k = 0; k < 4; k++)
                if ((k < nside0) == (s == 0)) zz += hz[k] * w[k][al];
            S[4 * s + al][l] = zz;
        }
    stage_flush<8>(st, a.rec, REC, R_Z);
#pragma unroll
    for (int k = 0; k < 4; k++)
        gv[k] = mk(gv[k].x + gmv[k] * ft.x, gv[k].y + gmv[k] * ft.y, gv[k].z + gmv[k] * ft.z);
    grad_fields(nside0, w, gv, [&](int k, double v) { S[k - R_GRAD][l] = v; });
    stage_flush<24>(st, a.rec, REC, R_GRAD);
}
// Class 3 contacts (four points: point-plane, edge-edge interior, every
// mollified edge-edge) in two launches.
// k_contact_c3r: ten threads per contact.  Thread role r < 9 builds the entries
// of the 9x9 Helmert-coordinate matrix K = (C (x) I3) Mred (C^T (x) I3) +
// mollifier terms (C = h4 S^T, Cm the same for y = p1 - p0, z = p3 - p2) of
// one component pair (i, j) = (r / 3, r % 3) and stores them with their
// mirror images (K exactly symmetric) in the record's PK field; role 9
// writes the record's other fields (Z, friction, gradient, class code).
// k_contact_psd9: one thread per contact projects K in place, cyclic Jacobi
// on the packed upper triangle in registers with the eigenvectors beside it
// (a scheduling barrier after each rotation keeps the live ranges to one
// rotation).  Split this way the geometry's registers stay out of the Jacobi
// kernel, and the matrix of a contact is built by nine threads at once.
__host__ __device__ constexpr int P9(int p, int q) {
    return p <= q ? p * 9 - p * (p - 1) / 2 + (q - p) : q * 9 - q * (q - 1) / 2 + (p - q);
}
constexpr int C3_ROLES = 10;
struct C3Ctx {
    Geo G;
    int id0, id1, id2, id3;
    double* R;
};
__device__ __forceinline__ void c3_setup(const WorldArgs& a, int cid, C3Ctx& x) {
    x.R = a.rec + REC * (size_t)cid;
    Geo& G = x.G;
    G.is_ee = contact_ids(a, cid, x.id0, x.id1, x.id2, x.id3);
    G.p0 = ld3(a.X + 3 * (size_t)x.id0); G.p1 = ld3(a.X + 3 * (size_t)x.id1);
    G.p2 = ld3(a.X + 3 * (size_t)x.id2); G.p3 = ld3(a.X + 3 * (size_t)x.id3);
    G.type = a.ctype[cid]; G.d = a.cdist[cid];
    geo_init(G, a.kp);
}
// the entries of K for component pair role = (i, j), into the PK field
__device__ __forceinline__ void c3_kbuild(const C3Ctx& x, int role) {
    const Geo& G = x.G;
    double* R = x.R;
    double C[3][3], Cm[3][2];
#pragma unroll
    for (int q = 0; q < 3; q++) {
#pragma unroll
        for (int m = 0; m < 3; m++) {
            double c = 0.0;
#pragma unroll
            for (int k = 0; k < 4; k++) c += kHel[q][k] * red_sign(G.map, m, k);
            C[q][m] = m < G.map.nr ? c : 0.0;
        }
        Cm[q][0] = kHel[q][1] - kHel[q][0];
        Cm[q][1] = kHel[q][3] - kHel[q][2];
    }
    {
        const int i = role / 3, j = role % 3;
        double K[3][3];
#pragma unroll
        for (int q = 0; q < 3; q++)
#pragma unroll
            for (int b = 0; b < 3; b++) K[q][b] = 0.0;
#pragma unroll
        for (int m = 0; m < 3; m++)
#pragma unroll
            for (int n = 0; n < 3; n++) {
                if (m < G.map.nr && n < G.map.nr) {
                    const double Mv = G.B1 * G.rd.hess(m, i, n, j) + G.B2 * cmp(G.rd.g[m], i) * cmp(G.rd.g[n], j);
#pragma unroll
                    for (int q = 0; q < 3; q++)
#pragma unroll
                        for (int b = 0; b < 3; b++) K[q][b] += C[q][m] * C[b][n] * Mv;
                }
            }
        if (G.is_ee) {
            double gdi[3], gdj[3], gxi[3], gxj[3];
#pragma unroll
            for (int q = 0; q < 3; q++) {
                double si = 0.0, sj = 0.0;
#pragma unroll
                for (int m = 0; m < 3; m++)
                    if (m < G.map.nr) { si += C[q][m] * cmp(G.rd.g[m], i); sj += C[q][m] * cmp(G.rd.g[m], j); }
                gdi[q] = si; gdj[q] = sj;
                gxi[q] = G.m2 != 0.0 ? Cm[q][0] * G.mc.grad(0, i) + Cm[q][1] * G.mc.grad(1, i) : 0.0;
                gxj[q] = G.m2 != 0.0 ? Cm[q][0] * G.mc.grad(0, j) + Cm[q][1] * G.mc.grad(1, j) : 0.0;
            }
#pragma unroll
            for (int q = 0; q < 3; q++)
#pragma unroll
                for (int b = 0; b < 3; b++) {
                    double hxk = 0.0;
                    if (G.m2 != 0.0) {
#pragma unroll
                        for (int u = 0; u < 2; u++)
#pragma unroll
                            for (int v = 0; v < 2; v++) hxk += Cm[q][u] * Cm[b][v] * G.mc.hess(u, i, v, j);
                    }
                    const double gmi = G.m1 * gxi[q], gmj = G.m1 * gxj[b];
                    K[q][b] = G.m * K[q][b] + G.B1 * (gmi * gdj[b] + gdi[q] * gmj) + G.B0 * (G.m2 * gxi[q] * gxj[b] + G.m1 * hxk);
                }
        }
#pragma unroll
        for (int q = 0; q < 3; q++)
#pragma unroll
            for (int b = 0; b < 3; b++)
                if (3 * q + i <= 3 * b + j) {
                    // entry (3q + i, 3b + j) of the upper triangle and its mirror; the
                    // lower-triangle entries of this pair are upper ones of pair (j, i)
                    R[R_PK + 9 * (3 * q + i) + 3 * b + j] = K[q][b];
                    R[R_PK + 9 * (3 * b + j) + 3 * q + i] = K[q][b];
                }
    }
}
// everything but PK: Z, friction (F, G, Tk, lambda), the gradient, the class code
__device__ __forceinline__ void c3_fields(const WorldArgs& a, int cid, const C3Ctx& x) {
    const Geo& G = x.G;
    double* R = x.R;
    const int id0 = x.id0, id1 = x.id1, id2 = x.id2, id3 = x.id3;
    const int nside0 = G.is_ee ? 2 : 1;
    double w[4][4];
    load_w(a.xbar, id0, id1, id2, id3, w);
#pragma unroll
    for (int q = 0; q < 3; q++)
#pragma unroll
        for (int s = 0; s < 2; s++)
#pragma unroll
            for (int al = 0; al < 4; al++) {
                double zz = 0.0;
#pragma unroll
                for (int k = 0; k < 4; k++)
                    if ((k < nside0) == (s == 0)) zz += kHel[q][k] * w[k][al];
                R[R_Z + 8 * q + 4 * s + al] = zz;
            }
    Fric f;
    friction_eval(a, G.is_ee, G.type, G.d, G.B1, G.p0, G.p1, G.p2, G.p3, id0, id1, id2, id3, f);
    write_fric(a, cid, nside0, f, w);
    const double gmv[4] = {f.gm0, f.gm1, f.gm2, f.gm3};
    d3 gv[4];
#pragma unroll
    for (int k = 0; k < 4; k++) {
        double g3[3];
#pragma unroll
        for (int c = 0; c < 3; c++) {
            double g = barrier_grad(G, 3 * k + c);
            if (f.on) g += f.fs * (gmv[k] * cmp(f.c0, c) * f.u0 + gmv[k] * cmp(f.c1, c) * f.u1);
            g3[c] = g;
        }
        gv[k] = mk(g3[0], g3[1], g3[2]);
    }
    write_grad(a, cid, nside0, w, gv);
    R[R_D] = rec_code(a, cid, 9);
}
__device__ __forceinline__ void psd9(const WorldArgs& a, int cid) {
    double* R = a.rec + REC * (size_t)cid;
    // PK: packed K and V in registers (static indices); a scheduling barrier
    // after each rotation keeps the live ranges to one rotation
    double A[45];
#pragma unroll
    for (int p = 0; p < 9; p++)
#pragma unroll
        for (int q = p; q < 9; q++) A[P9(p, q)] = R[R_PK + 9 * p + q];
    {
        double Vr[81];
#pragma unroll
        for (int e = 0; e < 81; e++) Vr[e] = (e % 10 == 0) ? 1.0 : 0.0;
        for (int sweep = 0; sweep < MAX_SWEEPS; sweep++) {
            double off = 0.0, tot = 0.0;
#pragma unroll
            for (int p = 0; p < 9; p++)
#pragma unroll
                for (int q = p; q < 9; q++) {
                    const double x = A[P9(p, q)] * A[P9(p, q)];
                    if (p == q) tot += x; else { off += 2.0 * x; tot += 2.0 * x; }
                }
            if (off <= JACOBI_TOL * tot || off == 0.0) break;
            const double thr = JACOBI_SKIP * sqrt(tot);
#pragma unroll
            for (int p = 0; p < 8; p++)
#pragma unroll
                for (int q = p + 1; q < 9; q++) {
                    const double apq = A[P9(p, q)];
                    if (fabs(apq) > thr) {
                        const double app = A[P9(p, p)], aqq = A[P9(q, q)];
                        double c, s, t;
                        jacobi_cs(app, aqq, apq, c, s, t);
#pragma unroll
                        for (int k = 0; k < 9; k++) {
                            if (k == p || k == q) continue;
                            const double akp = A[P9(k, p)], akq = A[P9(k, q)];
                            A[P9(k, p)] = fma(c, akp, -s * akq);
                            A[P9(k, q)] = fma(s, akp, c * akq);
                        }
                        A[P9(p, p)] = fma(-t, apq, app);
                        A[P9(q, q)] = fma(t, apq, aqq);
                        A[P9(p, q)] = 0.0;
#pragma unroll
                        for (int k = 0; k < 9; k++) {
                            const double vkp = Vr[9 * k + p], vkq = Vr[9 * k + q];
                            Vr[9 * k + p] = fma(c, vkp, -s * vkq);
                            Vr[9 * k + q] = fma(s, vkp, c * vkq);
                        }
                    }
                    __builtin_amdgcn_sched_barrier(0);
                }
        }
        double wv[9];
#pragma unroll
        for (int k = 0; k < 9; k++) wv[k] = fmax(A[P9(k, k)], 0.0);
#pragma unroll
        for (int p = 0; p < 9; p++)
#pragma unroll
            for (int q = p; q < 9; q++) {
                double s = 0.0;
#pragma unroll
                for (int k = 0; k < 9; k++) s += Vr[9 * p + k] * wv[k] * Vr[9 * q + k];
                R[R_PK + 9 * p + q] = s;
                R[R_PK + 9 * q + p] = s;
            }
    }
}
// The roles and the Jacobi are separate kernels, so the role threads get the
// registers of their own code (two wavefronts per SIMD) while the Jacobi
// holds ~280 (one).
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2))) void k_contact_c3r(WorldArgs a) {
    const int n3 = a.counts[2];
    const long long n = (long long)C3_ROLES * n3;
    for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (long long)gridDim.x * blockDim.x) {
        const int cid = a.list3[t / C3_ROLES], role = (int)(t % C3_ROLES);
        C3Ctx x;
        c3_setup(a, cid, x);
        if (role < 9) c3_kbuild(x, role);
        else c3_fields(a, cid, x);
    }
}
__global__ __launch_bounds__(64) void k_contact_psd9(WorldArgs a) {
    const int n3 = a.counts[2];
    for (int p = blockIdx.x * 64 + threadIdx.x; p < n3; p += gridDim.x * 64) psd9(a, a.list3[p]);
}

// record fields other than PK of a 3-point contact: D, friction, gradient, Z
__device__ __forceinline__ void scalars_c2(const WorldArgs& a, const Stage& st, bool valid) {
    const int cid = st.cid, l = st.lane;
    double (*S)[SLD] = st.S;
    int id0, id1, id2, id3;
    const int is_ee = contact_ids(a, cid, id0, id1, id2, id3);
    const d3 P0 = ld3(a.X + 3 * (size_t)id0), P1 = ld3(a.X + 3 * (size_t)id1);
    const d3 P2 = ld3(a.X + 3 * (size_t)id2), P3 = ld3(a.X + 3 * (size_t)id3);
    const int type = a.ctype[cid];
    const double d = a.cdist[cid];
    const RedMap mp = is_ee ? ee_map(type) : pt_map(type);
    const int nside0 = is_ee ? 2 : 1;
    const double B1 = barrier_d1(d, a.kp.d_hat, a.kp.kappa);
    const d3 z = mk(0, 0, 0);
    Fric f;
    friction_eval(a, is_ee, type, d, B1, P0, P1, P2, P3, id0, id1, id2, id3, f);
    double w[4][4];
    load_w(a.xbar, id0, id1, id2, id3, w);
    S[R_D][l] = rec_code(a, cid, 6);
    fric_fields(f, nside0, w, [&](int k, double v) { S[k][l] = v; });
    stage_flush<R_Z>(st, a.rec, REC, 0);
    tk_fields(f, [&](int k, double v) { S[k][l] = v; });
    stage_flush<TKN>(st, a.tk, TKN, 0);
    if (valid) a.lam[cid] = f.lam;
    const int ka = mp.plus[0], kb = mp.plus[1], km = mp.minus[0];
    const d3 pm = pick(P0, P1, P2, P3, km);
    RedDist rd;
    rd.init(K_PL, pick(P0, P1, P2, P3, ka) - pm, pick(P0, P1, P2, P3, kb) - pm, z);
    const d3 g0 = mk(B1 * rd.g[0].x, B1 * rd.g[0].y, B1 * rd.g[0].z);
    const d3 g1 = mk(B1 * rd.g[1].x, B1 * rd.g[1].y, B1 * rd.g[1].z);
    const d3 gm = mk(-(g0.x + g1.x), -(g0.y + g1.y), -(g0.z + g1.z));
    const double gmv[4] = {f.gm0, f.gm1, f.gm2, f.gm3};
    d3 ft = z;
    if (f.on) {
        const double t0 = f.fs * f.u0, t1 = f.fs * f.u1;
        ft = mk(t0 * f.c0.x + t1 * f.c1.x, t0 * f.c0.y + t1 * f.c1.y, t0 * f.c0.z + t1 * f.c1.z);
    }
    d3 gv[4];
#pragma unroll
    for (int k = 0; k < 4; k++) {
        const d3 gb = k == ka ? g0 : (k == kb ? g1 : (k == km ? gm : z));
        gv[k] = mk(gb.x + gmv[k] * ft.x, gb.y + gmv[k] * ft.y, gb.z + gmv[k] * ft.z);
    }
    grad_fields(nside0, w, gv, [&](int k, double v) { S[k - R_GRAD][l] = v; });
    stage_flush<24>(st, a.rec, REC, R_GRAD);
    // Helmert rows over the slots (m, a, b): (1, -1, 0) / sqrt 2, (1, 1, -2) / sqrt 6
#pragma unroll
    for (int aa = 0; aa < 2; aa++)
#pragma unroll
        for (int s = 0; s < 2; s++)
#pragma unroll
            for (int al = 0; al < 4; al++) {
                double zz = 0.0;
#pragma unroll
                for (int k = 0; k < 4; k++) {
                    const double h = aa == 0 ? (k == km ? HS2 : (k == ka ? -HS2 : 0.0))
                                             : (k == km ? HS6 : (k == ka ? HS6 : (k == kb ? -2.0 * HS6 : 0.0)));
                    if ((k < nside0) == (s == 0)) zz += h * w[k][al];
                }
                S[8 * aa + 4 * s + al][l] = zz;
            }
    stage_flush<16>(st, a.rec, REC, R_Z);
}

// packed upper-triangle index of an N x N symmetric matrix
template <int N> __host__ __device__ constexpr int PK_(int p, int q) {
    return p <= q ? p * N - p * (p - 1) / 2 + (q - p) : q * N - q * (q - 1) / 2 + (p - q);
}
// Cyclic Jacobi on a packed symmetric matrix A (registers, static indices)
// with the eigenvectors V (row-major N x N, registers); rotations of entries
// below JACOBI_SKIP |A|_F are skipped, and a scheduling barrier after each
// rotation bounds the live ranges.  Returns with the eigenvalues on A's diagonal.
template <int N>
__device__ __forceinline__ void jacobi_packed(double (&A)[N * (N + 1) / 2], double (&V)[N * N]) {
#pragma unroll
    for (int e = 0; e < N * N; e++) V[e] = (e % (N + 1) == 0) ? 1.0 : 0.0;
    for (int sweep = 0; sweep < MAX_SWEEPS; sweep++) {
        double off = 0.0, tot = 0.0;
#pragma unroll
        for (int p = 0; p < N; p++)
#pragma unroll
            for (int q = p; q < N; q++) {
                const double x = A[PK_<N>(p, q)] * A[PK_<N>(p, q)];
                if (p == q) tot += x; else { off += 2.0 * x; tot += 2.0 * x; }
            }
        if (off <= JACOBI_TOL * tot || off == 0.0) break;
        const double thr = JACOBI_SKIP * sqrt(tot);
#pragma unroll
        for (int p = 0; p < N - 1; p++)
#pragma unroll
            for (int q = p + 1; q < N; q++) {
                const double apq = A[PK_<N>(p, q)];
                if (fabs(apq) > thr) {
                    const double app = A[PK_<N>(p, p)], aqq = A[PK_<N>(q, q)];
                    double c, s, t;
                    jacobi_cs(app, aqq, apq, c, s, t);
#pragma unroll
                    for (int k = 0; k < N; k++) {
                        if (k == p || k == q) continue;
                        const double akp = A[PK_<N>(k, p)], akq = A[PK_<N>(k, q)];
                        A[PK_<N>(k, p)] = fma(c, akp, -s * akq);
                        A[PK_<N>(k, q)] = fma(s, akp, c * akq);
                    }
                    A[PK_<N>(p, p)] = fma(-t, apq, app);
                    A[PK_<N>(q, q)] = fma(t, apq, aqq);
                    A[PK_<N>(p, q)] = 0.0;
#pragma unroll
                    for (int k = 0; k < N; k++) {
                        const double vkp = V[N * k + p], vkq = V[N * k + q];
                        V[N * k + p] = fma(c, vkp, -s * vkq);
                        V[N * k + q] = fma(s, vkp, c * vkq);
                    }
                }
                __builtin_amdgcn_sched_barrier(0);
            }
    }
}
// PK = V diag(max(w, 0)) V^T into R[R_PK], row-major N x N
template <int N>
__device__ __forceinline__ void write_projection(const double (&A)[N * (N + 1) / 2], const double (&V)[N * N], double* R) {
    double wv[N];
#pragma unroll
    for (int k = 0; k < N; k++) wv[k] = fmax(A[PK_<N>(k, k)], 0.0);
#pragma unroll
    for (int p = 0; p < N; p++)
#pragma unroll
        for (int q = p; q < N; q++) {
            double s = 0.0;
#pragma unroll
            for (int k = 0; k < N; k++) s += V[N * p + k] * wv[k] * V[N * q + k];
            R[R_PK + N * p + q] = s;
            R[R_PK + N * q + p] = s;
        }
}

// PK of a 3-point contact (point-line distance): K = (C (x) I3) Mred (C^T (x) I3),
// 6x6 Jacobi in registers
__device__ __forceinline__ void pk_c2(const WorldArgs& a, int cid) {
    int id0, id1, id2, id3;
    const int is_ee = contact_ids(a, cid, id0, id1, id2, id3);
    const d3 P0 = ld3(a.X + 3 * (size_t)id0), P1 = ld3(a.X + 3 * (size_t)id1);
    const d3 P2 = ld3(a.X + 3 * (size_t)id2), P3 = ld3(a.X + 3 * (size_t)id3);
    const int type = a.ctype[cid];
    const double d = a.cdist[cid];
    const RedMap mp = is_ee ? ee_map(type) : pt_map(type);
    const double B1 = barrier_d1(d, a.kp.d_hat, a.kp.kappa);
    const double B2 = barrier_d2(d, a.kp.d_hat, a.kp.kappa);
    double* R = a.rec + REC * (size_t)cid;
    const int ka = mp.plus[0], kb = mp.plus[1], km = mp.minus[0];
    const d3 pm = pick(P0, P1, P2, P3, km);
    RedDist rd;
    rd.init(K_PL, pick(P0, P1, P2, P3, ka) - pm, pick(P0, P1, P2, P3, kb) - pm, mk(0, 0, 0));
    double A[21];
    {
        double M[6][6];
#pragma unroll
        for (int u = 0; u < 6; u++)
#pragma unroll
            for (int v = u; v < 6; v++) {
                const int m = u / 3, i = u % 3, n = v / 3, j = v % 3;
                M[u][v] = B1 * rd.hess(m, i, n, j) + B2 * cmp(rd.g[m], i) * cmp(rd.g[n], j);
                M[v][u] = M[u][v];
            }
        const double C[2][2] = {{-2.0 * HS2, -HS2}, {0.0, -3.0 * HS6}};
#pragma unroll
        for (int aa = 0; aa < 2; aa++)
#pragma unroll
            for (int i = 0; i < 3; i++)
#pragma unroll
                for (int bb = 0; bb < 2; bb++)
#pragma unroll
                    for (int j = 0; j < 3; j++) {
                        double s = 0.0;
#pragma unroll
                        for (int m = 0; m < 2; m++)
#pragma unroll
                            for (int n = 0; n < 2; n++) s += C[aa][m] * C[bb][n] * M[3 * m + i][3 * n + j];
                        if (3 * aa + i <= 3 * bb + j) A[PK_<6>(3 * aa + i, 3 * bb + j)] = s;
                    }
    }
    if (a.kp.psd) {
        double V[36];
        jacobi_packed<6>(A, V);
        write_projection<6>(A, V, R);
    } else {
#pragma unroll
        for (int p = 0; p < 6; p++)
#pragma unroll
            for (int q = 0; q < 6; q++) R[R_PK + 6 * p + q] = A[PK_<6>(p, q)];
    }
}

// record fields other than PK of 3-point contacts, one thread each
// one 64-contact slice of the class list per wavefront and pass (block of
// 256 lanes, four staging areas: 49.9 KB, three blocks per CU as the
// registers allow); class 1 at list[0, n1), class 2 at list[total - n2, total)
// (stored from the back)
template <int CLS>
__device__ __forceinline__ void contact_slices(const WorldArgs& a) {
    __shared__ double S[4][SNF][SLD];
    const int total = a.n_pt + a.n_ee;
    const int cnt = a.counts[CLS - 1];
    const int wv = threadIdx.x >> 6;
    Stage st;
    st.S = S[wv];
    st.lane = threadIdx.x & 63;
    for (int p0 = (blockIdx.x * 4 + wv) * 64; p0 < cnt; p0 += gridDim.x * 256) {
        const int p = min(p0 + st.lane, cnt - 1);
        st.nvalid = min(64, cnt - p0);
        st.cid = a.list[CLS == 1 ? p : total - 1 - p];
        if (CLS == 1) contact_small(a, st, p0 + st.lane < cnt);
        else scalars_c2(a, st, p0 + st.lane < cnt);
    }
}
__global__ __launch_bounds__(256) void k_contact_scalars2(WorldArgs a) { contact_slices<2>(a); }
__global__ __launch_bounds__(64) void k_contact_pk2(WorldArgs a) {
    const int total = a.n_pt + a.n_ee;
    const int cnt = a.counts[1];
    for (int p = blockIdx.x * blockDim.x + threadIdx.x; p < cnt; p += gridDim.x * blockDim.x)
        pk_c2(a, a.list[total - 1 - p]);
}

template <int CLS>
__global__ __launch_bounds__(256) void k_contact_small(WorldArgs a) {
    static_assert(CLS == 1, "3-point contacts: k_contact_scalars2 / k_contact_pk2");
    contact_slices<1>(a);
}

// VG contact: compact body-space form.  The VG block of body I is
// (w w^T) (x) K with w = (1, xbar) and K = B''(d^2) (2d)^2 e_y e_y^T + friction,
// gradient w (x) gx (model/vg_ipc.cpp:12-30, model/barrier.cpp:31-69).
struct VgArgs {
    KParams kp; int n; double ground_y; int vg_fric;
    const int2* vg; const int* voff; const double* X; const double* Xs;
    double* K; double* lam;
};
__global__ void k_vg_world(VgArgs a) {
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= a.n) return;
    int2 c = a.vg[k];
    int v = a.voff[c.x] + c.y;
    d3 p = ld3(a.X + 3 * (size_t)v), p0 = ld3(a.Xs + 3 * (size_t)v);
    double dd = p.y - a.ground_y, d = dd * dd;
    double* K = a.K + 12 * (size_t)k;
    if (!(d < a.kp.d_hat)) {
        for (int i = 0; i < 12; i++) K[i] = 0.0;
        a.lam[k] = 0.0;
        return;
    }
    double B1 = barrier_d1(d, a.kp.d_hat, a.kp.kappa);
    double lam = -B1 * 2 * dd;
    a.lam[k] = lam;
    double gy = 1.0 * barrier_d1(dd * dd, a.kp.d_hat, a.kp.kappa) * 2 * dd;
    double B2 = barrier_d2(dd * dd, a.kp.d_hat, a.kp.kappa);
    double s = dd * 2;
    double Kyy = B2 * s * s;
    double gx = 0.0, gz = 0.0, Kxx = 0.0, Kxz = 0.0, Kzz = 0.0;
    if (a.vg_fric) {
        double u0 = p.x - p0.x, u1 = p.z - p0.z;
        double fs, M[3];
        friction_M(u0, u1, lam, a.kp, fs, M);
        gx = fs * u0; gz = fs * u1;
        Kxx = M[0]; Kxz = M[1]; Kzz = M[2];
    }
    // K column-major 3x3 (x, y, z), then gx
    K[0] = Kxx; K[1] = 0.0; K[2] = Kxz;
    K[3] = 0.0; K[4] = Kyy; K[5] = 0.0;
    K[6] = Kxz; K[7] = 0.0; K[8] = Kzz;
    K[9] = gx; K[10] = gy; K[11] = gz;
}

static KParams make_kp(const adb_params& P) {
    KParams kp;
    kp.d_hat = P.d_hat; kp.kappa = P.kappa; kp.eps_x = P.eps_x; kp.psd = P.psd;
    kp.mu = P.mu; kp.eps_v = P.dt * P.evh; kp.max_uk = P.max_uk;
    kp.pt_fric = P.pt_fric; kp.ee_fric = P.ee_fric;
    return kp;
}

void contact_terms(World& w) {
    const int n = w.cs.n_pt + w.cs.n_ee;
    const size_t nn = (size_t)std::max(n, 1);
    w.crec.reserve(REC * nn);
    w.ctype.reserve(nn); w.cdist.reserve(nn); w.ccls.reserve(nn); w.clist.reserve(nn); w.clist3.reserve(nn);
    w.ccount.reserve(4);
    w.ctk.reserve(TKN * nn);
    w.clam.reserve(nn);
    // the contact set is final here: the assembly's pair structure, which
    // needs nothing else, forks from this point onto its own stream (queued
    // below, after the contact kernels)
    if (!w.ev_asm_fork) ADB_HIP(hipEventCreateWithFlags(&w.ev_asm_fork, hipEventDisableTiming));
    ADB_HIP(hipEventRecord(w.ev_asm_fork, w.stream));
    // class counts [0, 3) (entry 3, the chunk count, belongs to the assembly)
    ADB_HIP(hipMemsetAsync(w.ccount.p, 0, 3 * sizeof(int), w.stream));
    if (n > 0) {
        WorldArgs a;
        a.kp = make_kp(w.P);
        a.n_pt = w.cs.n_pt; a.n_ee = w.cs.n_ee;
        a.pt = w.cs.pt.p; a.ee = w.cs.ee.p;
        a.voff = w.voff.p; a.foff = w.foff.p; a.eoff = w.eoff.p; a.face = w.face.p; a.edge = w.edge.p;
        a.X = w.X.p; a.Xs = w.Xs.p; a.xbar = w.xbar.p; a.mass = w.mass.p;
        a.ctype = w.ctype.p; a.cdist = w.cdist.p; a.ccls = w.ccls.p; a.list = w.clist.p; a.list3 = w.clist3.p; a.counts = w.ccount.p;
        a.rec = w.crec.p; a.tk = w.ctk.p; a.lam = w.clam.p;
        {
            // ids and 4 points in; type, distance and list slot out
            ProfScope ps(w, PK_CONTACT_CLASSIFY, (double)n * 152.0);
            const int rpt = std::min(CL_RPT_MAX, std::max(1, (n + CL_THREADS * 512 - 1) / (CL_THREADS * 512)));
            hipLaunchKernelGGL(k_contact_classify, dim3((n + CL_THREADS * rpt - 1) / (CL_THREADS * rpt)), dim3(CL_THREADS), 0,
                               w.stream, a, rpt);
            ADB_HIP(hipGetLastError());
        }
        const int* snap = prof_snapshot(w, w.ccount.p);
        {
            // 4-point contacts on the side stream (four 64-lane workgroups per CU,
            // one wave per SIMD, FP64-VALU-bound): they overlap the 2- and 3-point
            // kernels and the assembly's sort; the assembly waits for them
            // before its reduction
            if (!w.stream2) {
                ADB_HIP(hipStreamCreate(&w.stream2));
                ADB_HIP(hipStreamCreate(&w.stream3));
                for (hipEvent_t& e : w.ev_side) ADB_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
                ADB_HIP(hipEventCreateWithFlags(&w.ev_pk, hipEventDisableTiming));
            }
            ADB_HIP(hipEventRecord(w.ev_side[0], w.stream));
            ADB_HIP(hipStreamWaitEvent(w.stream2, w.ev_side[0], 0));
            {
                ProfScope ps(w, PK_CONTACT_GROUP, 0.0, w.stream2);
                ps.counts(snap, 0.0, 0.0, 1608.0, 0.0);
                // the class count is on the device: grids sized for the worst case, grid-stride loops
                hipLaunchKernelGGL(k_contact_c3r, dim3(std::min((C3_ROLES * n + 255) / 256, 16 * w.num_cu)), dim3(256), 0,
                                   w.stream2, a);
                ADB_HIP(hipGetLastError());
                if (a.kp.psd)
                    hipLaunchKernelGGL(k_contact_psd9, dim3(std::min((n + 63) / 64, 4 * w.num_cu)), dim3(64), 0, w.stream2, a);
                ADB_HIP(hipGetLastError());
            }
            ADB_HIP(hipEventRecord(w.ev_side[1], w.stream2));
            // the 6x6 projections of the 3-point contacts (VALU-bound) on a third
            // stream, beside the 2- and 3-point record kernels (bound by their
            // scattered loads and stores); they write disjoint record fields
            ADB_HIP(hipStreamWaitEvent(w.stream3, w.ev_side[0], 0));
            {
                // per 3-point contact: ids, 4 points, type and distance in; PK (36 doubles) out
                ProfScope ps(w, PK_CONTACT_PK, 0.0, w.stream3);
                ps.counts(snap, 0.0, 440.0, 0.0, 0.0);
                hipLaunchKernelGGL(k_contact_pk2, dim3(std::min((n + 63) / 64, 16 * w.num_cu)), dim3(64), 0, w.stream3, a);
                ADB_HIP(hipGetLastError());
            }
            ADB_HIP(hipEventRecord(w.ev_pk, w.stream3));
            w.side_pending = true;
        }
        {
            // per contact: list, ids, type, distance, X / Xs / xbar of 4 points in;
            // record (D = 3: 59 doubles; D = 6: 58 here, PK on the third stream),
            // lagged basis (10) and lambda out
            ProfScope ps(w, PK_CONTACT, 0.0);
            ps.counts(snap, 904.0, 896.0, 0.0, 0.0);
            const int grid = std::min((n + 255) / 256, 8 * w.num_cu);
            hipLaunchKernelGGL(k_contact_small<1>, dim3(grid), dim3(256), 0, w.stream, a);
            ADB_HIP(hipGetLastError());
            hipLaunchKernelGGL(k_contact_scalars2, dim3(grid), dim3(256), 0, w.stream, a);
            ADB_HIP(hipGetLastError());
        }
    }
    const int nv = w.cs.n_vg;
    // at least V / 8 entries: the count grows step by step while a pile
    // settles, and each regrowth frees a buffer (a device-wide wait)
    const size_t nvc = std::max<size_t>(std::max(nv, 1), (size_t)w.V / 8);
    w.vgK.reserve(12 * nvc);
    w.vglam.reserve(nvc);
    if (nv > 0) {
        VgArgs v;
        v.kp = make_kp(w.P); v.n = nv; v.ground_y = w.P.ground_y; v.vg_fric = w.P.vg_fric;
        v.vg = w.cs.vg.p; v.voff = w.voff.p; v.X = w.X.p; v.Xs = w.Xs.p;
        v.K = w.vgK.p; v.lam = w.vglam.p;
        ProfScope ps(w, PK_VG_CONTACT, (double)nv * (8.0 + 48.0 + 96.0 + 8.0));
        hipLaunchKernelGGL(k_vg_world, dim3((nv + 255) / 256), dim3(256), 0, w.stream, v);
        ADB_HIP(hipGetLastError());
    }
    assemble_structure(w, true);
}

} // namespace adb

// ============================================================================ C ABI
using namespace adb;

extern "C" void adb_kernel_config_default(adb_kernel_config* c) {
    // abdtk build: barrier::d_hat = 1e-4, barrier::kappa = 1e-3 (model/barrier.cpp:7-8),
    // globals.eps_x = 1e-3 and globals.psd = true (model/settings.h:16,39)
    c->d_hat = 1e-4; c->kappa = 1e-3; c->eps_x = 1e-3; c->psd = 1;
}

static KParams api_kp(const adb_kernel_config* c) {
    KParams kp;
    adb_kernel_config d;
    if (!c) { adb_kernel_config_default(&d); c = &d; }
    kp.d_hat = c->d_hat; kp.kappa = c->kappa; kp.eps_x = c->eps_x; kp.psd = c->psd;
    kp.mu = 0; kp.eps_v = 1; kp.max_uk = 0; kp.pt_fric = 0; kp.ee_fric = 0;
    return kp;
}

extern "C" int adb_ipc_hess_pt_12x12_batched(const adb_kernel_config* cfg, int n, const double* pt, const int* type,
                                             const double* dist, double* H, double* g, void* stream) {
    if (n < 0) return ADB_E_ARG;
    if (n == 0) return ADB_OK;
    hipLaunchKernelGGL(k_hess_api<0>, dim3((n + CPW - 1) / CPW), dim3(64), 0, (hipStream_t)stream,
                       api_kp(cfg), n, pt, type, dist, H, g, (double*)nullptr);
    return hipGetLastError() == hipSuccess ? ADB_OK : ADB_E_HIP;
}
extern "C" int adb_ipc_hess_ee_12x12_batched(const adb_kernel_config* cfg, int n, const double* ee, const int* type,
                                             const double* dist, double* H, double* g, double* moll, void* stream) {
    if (n < 0) return ADB_E_ARG;
    if (n == 0) return ADB_OK;
    hipLaunchKernelGGL(k_hess_api<1>, dim3((n + CPW - 1) / CPW), dim3(64), 0, (hipStream_t)stream,
                       api_kp(cfg), n, ee, type, dist, H, g, moll);
    return hipGetLastError() == hipSuccess ? ADB_OK : ADB_E_HIP;
}

template <int IS_EE>
static int single_hess(const adb_kernel_config* cfg, const double x[12], int type, double dist, double H[144], double g[12], double* moll) {
    double *dx = nullptr, *dd = nullptr, *dH = nullptr, *dg = nullptr, *dm = nullptr;
    int* dt = nullptr;
    int rc = ADB_OK;
    if (hipMalloc((void**)&dx, 12 * sizeof(double)) != hipSuccess || hipMalloc((void**)&dd, sizeof(double)) != hipSuccess ||
        hipMalloc((void**)&dH, 144 * sizeof(double)) != hipSuccess || hipMalloc((void**)&dg, 12 * sizeof(double)) != hipSuccess ||
        hipMalloc((void**)&dm, sizeof(double)) != hipSuccess || hipMalloc((void**)&dt, sizeof(int)) != hipSuccess) {
        rc = ADB_E_HIP;
    } else {
        if (hipMemcpy(dx, x, 12 * sizeof(double), hipMemcpyHostToDevice) != hipSuccess ||
            hipMemcpy(dd, &dist, sizeof(double), hipMemcpyHostToDevice) != hipSuccess ||
            hipMemcpy(dt, &type, sizeof(int), hipMemcpyHostToDevice) != hipSuccess) rc = ADB_E_HIP;
        if (rc == ADB_OK) {
            if (IS_EE) rc = adb_ipc_hess_ee_12x12_batched(cfg, 1, dx, dt, dd, dH, dg, dm, nullptr);
            else rc = adb_ipc_hess_pt_12x12_batched(cfg, 1, dx, dt, dd, dH, dg, nullptr);
        }
        if (rc == ADB_OK && hipDeviceSynchronize() != hipSuccess) rc = ADB_E_HIP;
        if (rc == ADB_OK) {
            if (hipMemcpy(H, dH, 144 * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess ||
                hipMemcpy(g, dg, 12 * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess) rc = ADB_E_HIP;
            if (IS_EE && moll && hipMemcpy(moll, dm, sizeof(double), hipMemcpyDeviceToHost) != hipSuccess) rc = ADB_E_HIP;
        }
    }
    (void)hipFree(dx); (void)hipFree(dd); (void)hipFree(dH); (void)hipFree(dg); (void)hipFree(dm); (void)hipFree(dt);
    return rc;
}
extern "C" int adb_ipc_hess_pt_12x12(const adb_kernel_config* cfg, const double pt[12], int type, double dist, double H[144], double g[12]) {
    return single_hess<0>(cfg, pt, type, dist, H, g, nullptr);
}
extern "C" int adb_ipc_hess_ee_12x12(const adb_kernel_config* cfg, const double ee[12], int type, double dist, double H[144], double g[12], double* moll) {
    return single_hess<1>(cfg, ee, type, dist, H, g, moll);
}
