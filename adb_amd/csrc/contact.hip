// contact.hip - PT / EE / VG contact terms on gfx950.
//
// Replaces IPC::ipc_hess_pt_12x12 / ipc_hess_ee_12x12 (model/core_ipc.cpp:21-83),
// project_to_psd (model/psd.h:9-33), pt_uktk / ee_uktk (model/pt_ipc.cpp:22-107,
// model/ee_ipc.cpp:21-118), friction (model/friction.cpp:15-47), the J^T H J body
// mapping (model/pt_ipc.cpp:153-169, model/ee_ipc.cpp:175-192) and the VG term
// (model/vg_ipc.cpp:12-30, model/abd.cpp:460-481).
//
// Layout: one 16-lane group per PT/EE contact, four contacts per 64-lane
// wavefront (one wavefront per workgroup).  Lane j < 12 owns column j of the
// contact's 12x12 vertex-space Hessian.
//
// PSD projection: every barrier and mollifier term depends on the four points
// only through differences, so the Hessian annihilates rigid translations.
// With Q = h (x) I3, h the 3x4 Helmert basis of the complement of (1,1,1,1),
// H = Q^T (Q H Q^T) Q and the projection onto the PSD cone is
// Q^T P(Q H Q^T) Q: a 9x9 symmetric eigenproblem, solved by parallel cyclic
// Jacobi (5 disjoint rotations per round, 9 rounds per sweep over 10 slots,
// slot 9 idle) on the matrix staged in LDS.  As in the reference, a Hessian
// whose eigenvalues are all >= 0 is returned unchanged.
#include <cstdio>
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
// Projects the symmetric vertex-space Hessian in S.H in place (all 64 lanes
// must call: barriers).  S.H is unchanged when the matrix is already PSD or
// `on` is false.
__device__ __forceinline__ void group_psd(GroupLDS& S, int lane, bool on) {
    // K = Q H Q^T: lane u = (b, j) < 9 builds column u
    if (lane < 9) {
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
        const bool conv = !on || off <= 1e-30 * tot || off == 0.0;
        if (__all(conv)) break;
        for (int rd = 0; rd < 9; rd++) {
            if (lane < 5) {
                const int p = kRR9[rd][lane][0], q = kRR9[rd][lane][1];
                double c = 1.0, s = 0.0;
                if (!conv && q < 9) {
                    const double apq = S.A[p][q];
                    if (apq != 0.0) {
                        const double app = S.A[p][p], aqq = S.A[q][q];
                        const double tau = (aqq - app) / (2.0 * apq);
                        double t;
                        if (fabs(tau) > 1e150) t = 0.5 / tau;
                        else t = (tau >= 0 ? 1.0 : -1.0) / (fabs(tau) + sqrt(1.0 + tau * tau));
                        c = 1.0 / sqrt(1.0 + t * t);
                        s = t * c;
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
                        S.A[lane][p] = c * ap - s * aq; S.A[lane][q] = s * ap + c * aq;
                        const double vp = S.V[lane][p], vq = S.V[lane][q];
                        S.V[lane][p] = c * vp - s * vq; S.V[lane][q] = s * vp + c * vq;
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
                        S.A[p][lane] = c * ap - s * aq; S.A[q][lane] = s * ap + c * aq;
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
    const bool project = on && wmin < 0.0;
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
__device__ __forceinline__ void friction_frame(const Geo& G, double& g0, double& g1, double& g2, double& g3, d3& c0, d3& c1) {
    const d3 P0 = G.p0, P1 = G.p1, P2 = G.p2, P3 = G.p3;
    if (!G.is_ee) {
        double t0, t1, t2;
        switch (G.type) {
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
        switch (G.type) {
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
    if (kp.psd) group_psd(L, lane, valid);
    if (valid && lane < 12) {
        for (int i = 0; i < 12; i++) H[144 * (size_t)e + 12 * lane + i] = L.H[i][lane];
        g[12 * (size_t)e + lane] = gj;
        if (IS_EE && lane == 0 && moll) moll[e] = mval;
    }
}

// ---------------------------------------------------------------- world contact terms
struct WorldArgs {
    KParams kp;
    int n_pt, n_ee;
    const int4* pt; const int4* ee;
    const int* voff; const int* foff; const int* eoff; const int* face; const int* edge;
    const double* X; const double* Xs; const double* xbar;
    double* blk; double* grad; double* tk; double* lam;
};

__global__ __launch_bounds__(64) void k_contact_world(WorldArgs a) {
    __shared__ GroupLDS lds[CPW];
    const int grp = threadIdx.x / GL, lane = threadIdx.x % GL;
    const int cid = blockIdx.x * CPW + grp;
    const int total = a.n_pt + a.n_ee;
    const bool valid = cid < total;
    GroupLDS& L = lds[grp];
    Geo G;
    int id0 = 0, id1 = 0, id2 = 0, id3 = 0;
    double gj = 0.0;
    bool active = false;
    if (valid) {
        G.is_ee = cid >= a.n_pt;
        if (!G.is_ee) {
            const int4 c = a.pt[cid];
            const int f = a.foff[c.z] + c.w;
            id0 = a.voff[c.x] + c.y; id1 = a.face[3 * f]; id2 = a.face[3 * f + 1]; id3 = a.face[3 * f + 2];
        } else {
            const int4 c = a.ee[cid - a.n_pt];
            const int ea = a.eoff[c.x] + c.y, eb = a.eoff[c.z] + c.w;
            id0 = a.edge[2 * ea]; id1 = a.edge[2 * ea + 1]; id2 = a.edge[2 * eb]; id3 = a.edge[2 * eb + 1];
        }
        G.p0 = ld3(a.X + 3 * (size_t)id0); G.p1 = ld3(a.X + 3 * (size_t)id1);
        G.p2 = ld3(a.X + 3 * (size_t)id2); G.p3 = ld3(a.X + 3 * (size_t)id3);
        if (!G.is_ee) { G.type = pt_type(G.p0, G.p1, G.p2, G.p3); G.d = pt_dist(G.p0, G.p1, G.p2, G.p3, G.type); }
        else { G.type = ee_type(G.p0, G.p1, G.p2, G.p3); G.d = ee_dist(G.p0, G.p1, G.p2, G.p3, G.type); }
        active = G.d < a.kp.d_hat;
        if (active) {
            geo_init(G, a.kp);
            if (lane < 12) barrier_column(G, lane, L.H, gj);
        }
    }
    if (!active && lane < 12) {
        for (int i = 0; i < 12; i++) L.H[i][lane] = 0.0;
    }
    __syncthreads();
    if (a.kp.psd) group_psd(L, lane, valid && active);
    // friction (model/friction.cpp:15-47) with displacement from the step start
    double lamc = 0.0, gm0 = 0.0, gm1 = 0.0, gm2 = 0.0, gm3 = 0.0;
    d3 c0 = mk(0, 0, 0), c1 = mk(0, 0, 0);
    if (valid && active) {
        friction_frame(G, gm0, gm1, gm2, gm3, c0, c1);
        const d3 d0 = G.p0 - ld3(a.Xs + 3 * (size_t)id0), d1 = G.p1 - ld3(a.Xs + 3 * (size_t)id1);
        const d3 d2 = G.p2 - ld3(a.Xs + 3 * (size_t)id2), d3v = G.p3 - ld3(a.Xs + 3 * (size_t)id3);
        const double u0 = gm0 * dot(c0, d0) + gm1 * dot(c0, d1) + gm2 * dot(c0, d2) + gm3 * dot(c0, d3v);
        const double u1 = gm0 * dot(c1, d0) + gm1 * dot(c1, d1) + gm2 * dot(c1, d2) + gm3 * dot(c1, d3v);
        lamc = -G.B1 * 2 * sqrt(G.d);
        const bool fr = G.is_ee ? a.kp.ee_fric : a.kp.pt_fric;
        if (fr && lane < 12) {
            double fs, M[3];
            friction_M(u0, u1, lamc, a.kp, fs, M);
            // T = Tk^T (12 x 2): T[(k, i)][r] = gam_k * c_r[i]
            const int kj = lane / 3, ij = lane % 3;
            const double gk = sel4(gm0, gm1, gm2, gm3, kj);
            const double Tj0 = gk * cmp(c0, ij), Tj1 = gk * cmp(c1, ij);
            gj += fs * (Tj0 * u0 + Tj1 * u1);
            const double m0 = M[0] * Tj0 + M[1] * Tj1, m1 = M[1] * Tj0 + M[2] * Tj1;
            const double gmv[4] = {gm0, gm1, gm2, gm3};
#pragma unroll
            for (int k = 0; k < 4; k++)
#pragma unroll
                for (int i = 0; i < 3; i++) L.H[3 * k + i][lane] += gmv[k] * cmp(c0, i) * m0 + gmv[k] * cmp(c1, i) * m1;
        }
    }
    if (lane < 12) L.g[lane] = gj;
    __syncthreads();
    if (valid && lane < 12) {
        // J^T H J with J = w^T (x) I3, w = (1, xbar): block (alpha, beta) of the
        // body-space matrix is sum_{a, b} w_a[alpha] w_b[beta] H_ab (Kronecker form)
        const int beta = lane / 3, c = lane % 3;
        const d3 xb0 = ld3(a.xbar + 3 * (size_t)id0), xb1 = ld3(a.xbar + 3 * (size_t)id1);
        const d3 xb2 = ld3(a.xbar + 3 * (size_t)id2), xb3 = ld3(a.xbar + 3 * (size_t)id3);
        const double wv[4][4] = {{1.0, xb0.x, xb0.y, xb0.z}, {1.0, xb1.x, xb1.y, xb1.z},
                                 {1.0, xb2.x, xb2.y, xb2.z}, {1.0, xb3.x, xb3.y, xb3.z}};
        const int nside0 = G.is_ee ? 2 : 1;      // points 0..nside0-1 belong to body I
        double wb[4];
#pragma unroll
        for (int b = 0; b < 4; b++) wb[b] = sel4(wv[b][0], wv[b][1], wv[b][2], wv[b][3], beta);
        double* blk = a.blk + 432 * (size_t)cid;
        // entry (alpha, r) of column (beta, c): hp = (I, I), ht = (J, J), off = (I, J)
#pragma unroll
        for (int alpha = 0; alpha < 4; alpha++)
#pragma unroll
            for (int r = 0; r < 3; r++) {
                double o0 = 0.0, o1 = 0.0, o2 = 0.0;
#pragma unroll
                for (int b = 0; b < 4; b++) {
                    double s0 = 0.0, s1 = 0.0;
#pragma unroll
                    for (int aa = 0; aa < 4; aa++) {
                        const double t = wv[aa][alpha] * L.H[3 * aa + r][3 * b + c];
                        if (aa < nside0) s0 += t; else s1 += t;
                    }
                    if (b < nside0) o0 += wb[b] * s0;
                    else { o1 += wb[b] * s1; o2 += wb[b] * s0; }
                }
                blk[12 * lane + 3 * alpha + r] = active ? o0 : 0.0;
                blk[144 + 12 * lane + 3 * alpha + r] = active ? o1 : 0.0;
                blk[288 + 12 * lane + 3 * alpha + r] = active ? o2 : 0.0;
            }
        // gradients: dg[(alpha, r)] = sum_a w_a[alpha] g_a[r] over each side
        const int alpha = lane / 3, r = lane % 3;
        double gp = 0.0, gt = 0.0;
#pragma unroll
        for (int aa = 0; aa < 4; aa++) {
            const double t = sel4(wv[aa][0], wv[aa][1], wv[aa][2], wv[aa][3], alpha) * L.g[3 * aa + r];
            if (aa < nside0) gp += t; else gt += t;
        }
        a.grad[24 * (size_t)cid + lane] = active ? gp : 0.0;
        a.grad[24 * (size_t)cid + 12 + lane] = active ? gt : 0.0;
        // lagged friction data: Tk (2 x 12, row-major) and lambda
        const double gk = sel4(gm0, gm1, gm2, gm3, alpha);
        a.tk[24 * (size_t)cid + lane] = gk * cmp(c0, r);
        a.tk[24 * (size_t)cid + 12 + lane] = gk * cmp(c1, r);
        if (lane == 0) a.lam[cid] = lamc;
    }
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
    w.cblk.reserve(432 * (size_t)std::max(n, 1));
    w.cgrad.reserve(24 * (size_t)std::max(n, 1));
    w.ctk.reserve(24 * (size_t)std::max(n, 1));
    w.clam.reserve(std::max(n, 1));
    if (n > 0) {
        WorldArgs a;
        a.kp = make_kp(w.P);
        a.n_pt = w.cs.n_pt; a.n_ee = w.cs.n_ee;
        a.pt = w.cs.pt.p; a.ee = w.cs.ee.p;
        a.voff = w.voff.p; a.foff = w.foff.p; a.eoff = w.eoff.p; a.face = w.face.p; a.edge = w.edge.p;
        a.X = w.X.p; a.Xs = w.Xs.p; a.xbar = w.xbar.p;
        a.blk = w.cblk.p; a.grad = w.cgrad.p; a.tk = w.ctk.p; a.lam = w.clam.p;
        const int blocks = (n + CPW - 1) / CPW;
        // per contact: 3 x 4 points of X, Xs, xbar + ids in; 3 blocks, 2 grads, Tk, lambda out
        ProfScope ps(w, PK_CONTACT, (double)n * (288.0 + 16.0 + 3456.0 + 192.0 + 192.0 + 8.0));
        hipLaunchKernelGGL(k_contact_world, dim3(blocks), dim3(64), 0, w.stream, a);
        ADB_HIP(hipGetLastError());
    }
    const int nv = w.cs.n_vg;
    w.vgK.reserve(12 * (size_t)std::max(nv, 1));
    w.vglam.reserve(std::max(nv, 1));
    if (nv > 0) {
        VgArgs v;
        v.kp = make_kp(w.P); v.n = nv; v.ground_y = w.P.ground_y; v.vg_fric = w.P.vg_fric;
        v.vg = w.cs.vg.p; v.voff = w.voff.p; v.X = w.X.p; v.Xs = w.Xs.p;
        v.K = w.vgK.p; v.lam = w.vglam.p;
        ProfScope ps(w, PK_VG_CONTACT, (double)nv * (8.0 + 48.0 + 96.0 + 8.0));
        hipLaunchKernelGGL(k_vg_world, dim3((nv + 255) / 256), dim3(256), 0, w.stream, v);
        ADB_HIP(hipGetLastError());
    }
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
