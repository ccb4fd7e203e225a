// adb_oracle.cpp - CPU FP64 oracle of the ADB per-timestep Newton contact step.
//
// TEST INFRASTRUCTURE ONLY: a from-scratch restatement of the reference's CPU
// path, used as the checker by tests/, __graft_entry__.smoke() and the
// cpu_baseline leg of bench.py.  It is never part of the product path.
//
// Reference files followed (paths relative to the reference repository):
//   step / Newton loop ....... model/abd.cpp:30-140, 142-196, 198-227, 229-303, 305-404
//   line search / energies ... model/line_search.cpp:39-136, model/energy.cpp:19-86,195-270
//   inertia / orthogonality .. model/mutils.cpp:53-76,197-201,245-252, model/othogonal_energy.cpp:8-62
//   contact terms ............ model/core_ipc.cpp:21-83, model/pt_ipc.cpp:22-186,
//                              model/ee_ipc.cpp:21-210, model/vg_ipc.cpp:12-30,
//                              model/friction.cpp:15-54, model/psd.h:9-33, model/barrier.cpp:11-69
//   assembly ................. model/sparse.cpp:13-57, model/assemble_matrices.cpp:28-59
//   culling .................. model/iaabb.cpp:23-568, model/thread_local_culling.cpp:209-386
//   CCD ...................... model/collision.cpp:83-348, model/iaabb.cpp:270-303,507-540
#include "adb_oracle.h"
#include "oracle_math.hpp"
#include <vector>
#include <map>
#include <unordered_map>
#include <set>
#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstring>
#include <string>
#ifdef _OPENMP
#include <omp.h>
#endif

using namespace orc;

namespace {

typedef double M12[144];   // column-major 12x12

// ---------------------------------------------------------------------------
// Distance derivatives by forward AD (independent of the GPU's closed forms).
// ---------------------------------------------------------------------------
typedef HD<12> D12;
typedef T3<D12> V3D;

static void ad_vars(const double x[12], V3D P[4]) {
    for (int k = 0; k < 4; k++)
        P[k] = V3D{D12::var(x[3 * k + 0], 3 * k + 0), D12::var(x[3 * k + 1], 3 * k + 1), D12::var(x[3 * k + 2], 3 * k + 2)};
}
static void ad_out(const D12& d, double* g, double* H) {
    if (g) for (int i = 0; i < 12; i++) g[i] = d.g[i];
    if (H) for (int c = 0; c < 12; c++) for (int r = 0; r < 12; r++) H[c * 12 + r] = d.h[D12::id(r, c)];
}
static D12 pt_dist_ad(const double x[12], int type) {
    V3D P[4]; ad_vars(x, P);
    switch (type) {
    case P_T0: return d_pp(P[0], P[1]);
    case P_T1: return d_pp(P[0], P[2]);
    case P_T2: return d_pp(P[0], P[3]);
    case P_E0: return d_pl(P[0], P[1], P[2]);
    case P_E1: return d_pl(P[0], P[2], P[3]);
    case P_E2: return d_pl(P[0], P[3], P[1]);
    default:   return d_pplane(P[0], P[1], P[2], P[3]);
    }
}
static D12 ee_dist_ad(const double x[12], int type) {
    V3D P[4]; ad_vars(x, P);
    switch (type) {
    case EA0_EB0: return d_pp(P[0], P[2]);
    case EA0_EB1: return d_pp(P[0], P[3]);
    case EA1_EB0: return d_pp(P[1], P[2]);
    case EA1_EB1: return d_pp(P[1], P[3]);
    case EA_EB0:  return d_pl(P[2], P[0], P[1]);
    case EA_EB1:  return d_pl(P[3], P[0], P[1]);
    case EA0_EB:  return d_pl(P[0], P[2], P[3]);
    case EA1_EB:  return d_pl(P[1], P[2], P[3]);
    default:      return d_ll(P[0], P[1], P[2], P[3]);
    }
}
// edge_edge_mollifier and its derivatives (ipc-toolkit v1.2.1):
// x = |(ea1-ea0) x (eb1-eb0)|^2, m = (-x/eps + 2) x/eps for x < eps, else 1.
static void mollifier_ad(const double x[12], double eps, double* m, double* g, double* H) {
    V3D P[4]; ad_vars(x, P);
    V3D c = tcross(sub(P[1], P[0]), sub(P[3], P[2]));
    D12 xx = tdot(c, c);
    if (xx.v < eps) {
        D12 e = D12::cst(eps);
        D12 r = (D12::cst(0.0) - xx / e + D12::cst(2.0)) * xx / e;
        *m = r.v;
        ad_out(r, g, H);
    } else {
        *m = 1.0;
        if (g) std::memset(g, 0, 12 * sizeof(double));
        if (H) std::memset(H, 0, 144 * sizeof(double));
    }
}
static double mollifier_value(V3 a0, V3 a1, V3 b0, V3 b1, double eps) {
    double x = sqn(cross(a1 - a0, b1 - b0));
    if (x < eps) return (-x / eps + 2) * x / eps;
    return 1.0;
}

// ---------------------------------------------------------------------------
// Symmetric eigen-decomposition (cyclic Jacobi) and PSD projection
// (model/psd.h:9-33, which uses Eigen's SelfAdjointEigenSolver).
// ---------------------------------------------------------------------------
static void jacobi_eig(int n, const double* Ain, double* w, double* V) {
    std::vector<double> A(Ain, Ain + n * n);
    for (int i = 0; i < n * n; i++) V[i] = 0.0;
    for (int i = 0; i < n; i++) V[i * n + i] = 1.0;
    auto a = [&](int r, int c) -> double& { return A[c * n + r]; };
    for (int sweep = 0; sweep < 100; sweep++) {
        double off = 0.0, tot = 0.0;
        for (int c = 0; c < n; c++) for (int r = 0; r < n; r++) { tot += a(r, c) * a(r, c); if (r != c) off += a(r, c) * a(r, c); }
        if (off <= 1e-34 * tot || off == 0.0) break;
        for (int p = 0; p < n - 1; p++)
            for (int q = p + 1; q < n; q++) {
                double apq = a(p, q);
                if (apq == 0.0) continue;
                double app = a(p, p), aqq = a(q, q);
                double tau = (aqq - app) / (2.0 * apq);
                double t = (tau >= 0 ? 1.0 : -1.0) / (std::fabs(tau) + std::sqrt(1.0 + tau * tau));
                double cs = 1.0 / std::sqrt(1.0 + t * t), sn = t * cs;
                for (int k = 0; k < n; k++) {   // A <- A J (columns p, q)
                    double akp = a(k, p), akq = a(k, q);
                    a(k, p) = cs * akp - sn * akq;
                    a(k, q) = sn * akp + cs * akq;
                }
                for (int k = 0; k < n; k++) {   // A <- J^T A (rows p, q)
                    double apk = a(p, k), aqk = a(q, k);
                    a(p, k) = cs * apk - sn * aqk;
                    a(q, k) = sn * apk + cs * aqk;
                }
                for (int k = 0; k < n; k++) {
                    double vkp = V[p * n + k], vkq = V[q * n + k];
                    V[p * n + k] = cs * vkp - sn * vkq;
                    V[q * n + k] = sn * vkp + cs * vkq;
                }
            }
    }
    for (int i = 0; i < n; i++) w[i] = a(i, i);
}
static void psd_project(int n, const double* A, double* out) {
    std::vector<double> w(n), V(n * n);
    jacobi_eig(n, A, w.data(), V.data());
    double mn = w[0];
    for (int i = 1; i < n; i++) mn = std::min(mn, w[i]);
    if (mn >= 0.0) { std::memcpy(out, A, sizeof(double) * n * n); return; }
    for (int i = 0; i < n; i++) if (w[i] < 0.0) w[i] = 0.0;
    for (int c = 0; c < n; c++)
        for (int r = 0; r < n; r++) {
            double s = 0.0;
            for (int k = 0; k < n; k++) s += V[k * n + r] * w[k] * V[k * n + c];
            out[c * n + r] = s;
        }
}

// ---------------------------------------------------------------------------
// Friction (model/friction.cpp:15-47).  T is n x 2 (column-major), g/H n-dim.
// ---------------------------------------------------------------------------
static void friction(const double uk[2], double lam, const double* T, int n, double mu, double eps,
                     double max_uk, double* g, double* H) {
    double u = std::sqrt(uk[0] * uk[0] + uk[1] * uk[1]);
    double f1 = f1_over_x(u, eps);
    double s = mu * lam * f1;
    for (int i = 0; i < n; i++) g[i] += s * (T[i] * uk[0] + T[n + i] * uk[1]);
    double M[4];   // 2x2 column-major
    if (u >= eps) {
        double ut0 = -uk[1], ut1 = uk[0];
        double k = mu * lam * f1 / (u * u);
        M[0] = k * ut0 * ut0; M[1] = k * ut1 * ut0; M[2] = k * ut0 * ut1; M[3] = k * ut1 * ut1;
    } else if (u <= max_uk) {
        double k = mu * lam * f1;
        M[0] = k; M[1] = 0; M[2] = 0; M[3] = k;
    } else {
        double f2 = -1.0 / (eps * eps);
        double k = f2 / u;
        double m0 = k * uk[0] * uk[0] + f1, m1 = k * uk[1] * uk[0], m3 = k * uk[1] * uk[1] + f1;
        double A2[4] = {m0 * mu * lam, m1 * mu * lam, m1 * mu * lam, m3 * mu * lam};
        psd_project(2, A2, M);
    }
    for (int c = 0; c < n; c++)
        for (int r = 0; r < n; r++) {
            double tr0 = T[r], tr1 = T[n + r], tc0 = T[c], tc1 = T[n + c];
            H[c * n + r] += tr0 * (M[0] * tc0 + M[2] * tc1) + tr1 * (M[1] * tc0 + M[3] * tc1);
        }
}

// ---------------------------------------------------------------------------
// Closest points and tangent bases (ipc-toolkit v1.2.1 friction/closest_point
// and friction/tangent_basis, called at model/pt_ipc.cpp:31-63 and
// model/ee_ipc.cpp:38-79).  Bases are 3x2 column-major.
// ---------------------------------------------------------------------------
static V3 nrm(V3 a) { double l = std::sqrt(sqn(a)); return V3{a.x / l, a.y / l, a.z / l}; }
static void set_basis(double Pk[6], V3 c0, V3 c1) { store3(Pk, c0); store3(Pk + 3, c1); }
static void basis_pp(V3 p0, V3 p1, double Pk[6]) {
    V3 d = p1 - p0;
    V3 cx = cross(v3(1, 0, 0), d), cy = cross(v3(0, 1, 0), d);
    if (sqn(cx) > sqn(cy)) set_basis(Pk, nrm(cx), nrm(cross(d, cx)));
    else set_basis(Pk, nrm(cy), nrm(cross(d, cy)));
}
static void basis_pe(V3 p, V3 e0, V3 e1, double Pk[6]) {
    V3 e = e1 - e0;
    set_basis(Pk, nrm(e), nrm(cross(e, p - e0)));
}
static void basis_pt(V3 t0, V3 t1, V3 t2, double Pk[6]) {
    V3 e0 = t1 - t0, n = cross(e0, t2 - t0);
    set_basis(Pk, nrm(e0), nrm(cross(n, e0)));
}
static void basis_ee(V3 a0, V3 a1, V3 b0, V3 b1, double Pk[6]) {
    V3 ea = a1 - a0, n = cross(ea, b1 - b0);
    set_basis(Pk, nrm(ea), nrm(cross(n, ea)));
}
static double pe_closest(V3 p, V3 e0, V3 e1) { V3 e = e1 - e0; return dot(p - e0, e) / sqn(e); }
static void solve2(double m00, double m01, double m11, double r0, double r1, double& x, double& y) {
    double det = m00 * m11 - m01 * m01;
    x = (r0 * m11 - m01 * r1) / det;
    y = (m00 * r1 - m01 * r0) / det;
}
static void pt_closest(const V3 P[4], int type, double tl[3], double Pk[6]) {
    switch (type) {
    case P_T0: basis_pp(P[0], P[1], Pk); tl[0] = 1; tl[1] = 0; tl[2] = 0; break;
    case P_T1: basis_pp(P[0], P[2], Pk); tl[0] = 0; tl[1] = 1; tl[2] = 0; break;
    case P_T2: basis_pp(P[0], P[3], Pk); tl[0] = 0; tl[1] = 0; tl[2] = 1; break;
    case P_E0: { double l = pe_closest(P[0], P[1], P[2]); tl[0] = 1 - l; tl[1] = l; tl[2] = 0; basis_pe(P[0], P[1], P[2], Pk); } break;
    case P_E1: { double l = pe_closest(P[0], P[2], P[3]); tl[0] = 0; tl[1] = 1 - l; tl[2] = l; basis_pe(P[0], P[2], P[3], Pk); } break;
    case P_E2: { double l = pe_closest(P[0], P[3], P[1]); tl[0] = l; tl[1] = 0; tl[2] = 1 - l; basis_pe(P[0], P[3], P[1], Pk); } break;
    default: {
        V3 b0 = P[2] - P[1], b1 = P[3] - P[1], r = P[0] - P[1];
        double x, y; solve2(dot(b0, b0), dot(b0, b1), dot(b1, b1), dot(b0, r), dot(b1, r), x, y);
        tl[0] = 1 - x - y; tl[1] = x; tl[2] = y;
        basis_pt(P[1], P[2], P[3], Pk);
    }
    }
}
static void ee_closest(const V3 P[4], int type, double lm[2], double Pk[6]) {
    switch (type) {
    case EA0_EB0: basis_pp(P[0], P[2], Pk); lm[0] = 0; lm[1] = 0; break;
    case EA0_EB1: basis_pp(P[0], P[3], Pk); lm[0] = 0; lm[1] = 1; break;
    case EA1_EB0: basis_pp(P[1], P[2], Pk); lm[0] = 1; lm[1] = 0; break;
    case EA1_EB1: basis_pp(P[1], P[3], Pk); lm[0] = 1; lm[1] = 1; break;
    case EA_EB0: lm[0] = pe_closest(P[2], P[0], P[1]); lm[1] = 0; basis_pe(P[2], P[0], P[1], Pk); break;
    case EA_EB1: lm[0] = pe_closest(P[3], P[0], P[1]); lm[1] = 1; basis_pe(P[3], P[0], P[1], Pk); break;
    case EA0_EB: lm[0] = 0; lm[1] = pe_closest(P[0], P[2], P[3]); basis_pe(P[0], P[2], P[3], Pk); break;
    case EA1_EB: lm[0] = 1; lm[1] = pe_closest(P[1], P[2], P[3]); basis_pe(P[1], P[2], P[3], Pk); break;
    default: {
        V3 w = P[0] - P[2], ea = P[1] - P[0], eb = P[3] - P[2];
        double x, y; solve2(sqn(ea), -dot(eb, ea), sqn(eb), -dot(w, ea), dot(w, eb), x, y);
        lm[0] = x; lm[1] = y;
        basis_ee(P[0], P[1], P[2], P[3], Pk);
    }
    }
    lm[0] = std::max(std::min(1.0, lm[0]), 0.0);
    lm[1] = std::max(std::min(1.0, lm[1]), 0.0);
}

// ---------------------------------------------------------------------------
// IPC::ipc_hess_pt_12x12 / ipc_hess_ee_12x12 (model/core_ipc.cpp:21-83).
// ---------------------------------------------------------------------------
static void hess_pt(const double x[12], int type, double dist, double dh, double kap, int psd, double* H, double* g) {
    D12 d = pt_dist_ad(x, type);
    double gd[12], Hd[144];
    ad_out(d, gd, Hd);
    double B1 = barrier_d1(dist, dh, kap), B2 = barrier_d2(dist, dh, kap);
    double T[144];
    for (int c = 0; c < 12; c++) for (int r = 0; r < 12; r++) T[c * 12 + r] = Hd[c * 12 + r] * B1 + gd[r] * gd[c] * B2;
    for (int i = 0; i < 12; i++) g[i] = gd[i] * B1;
    if (psd) psd_project(12, T, H); else std::memcpy(H, T, sizeof T);
}
static void hess_ee(const double x[12], int type, double dist, double dh, double kap, double epsx, int psd,
                    double* H, double* g, double* moll) {
    V3 a0 = load3(x), a1 = load3(x + 3), b0 = load3(x + 6), b1 = load3(x + 9);
    double eps = epsx * sqn(a0 - a1) * sqn(b0 - b1);
    double m, gm[12], Hm[144];
    mollifier_ad(x, eps, &m, gm, Hm);
    D12 d = ee_dist_ad(x, type);
    double gd[12], Hd[144];
    ad_out(d, gd, Hd);
    double B1 = barrier_d1(dist, dh, kap), B2 = barrier_d2(dist, dh, kap), B0 = barrier(dist, dh, kap);
    double T[144];
    for (int c = 0; c < 12; c++)
        for (int r = 0; r < 12; r++)
            T[c * 12 + r] = Hm[c * 12 + r] * B0 + B1 * (gm[r] * gd[c] + gd[r] * gm[c]) + m * (B2 * gd[r] * gd[c] + B1 * Hd[c * 12 + r]);
    for (int i = 0; i < 12; i++) g[i] = m * gd[i] * B1 + gm[i] * B0;
    *moll = m;
    if (psd) psd_project(12, T, H); else std::memcpy(H, T, sizeof T);
}

// ---------------------------------------------------------------------------
// World
// ---------------------------------------------------------------------------
struct Contact {            // output of one PT / EE contact term
    int I, J; bool active;
    double hp[144], ht[144], off[144], gp[12], gt[12];
    double Tk[24];          // 2x12 row-major: Tk[r*12+c]
    double lam;
};

struct World {
    adb_params P;
    int B = 0, V = 0, F = 0, E = 0;
    std::vector<int> voff, foff, eoff, vbody, fbody, ebody;
    std::vector<double> xbar;          // V*3
    std::vector<int> face, edge;       // global vertex ids
    std::vector<double> mass, Ic;
    std::vector<double> q, q0, dqdt;   // B*12
    int stall_pt = 0, stall_vg = 0;
    // collision set (canonical): PT (I,v,J,f) local ids, EE (I,ei,J,ej), VG (I,v)
    std::vector<int> pt, ee, vg;
    // lagged friction data of the current Newton iteration
    std::vector<double> pt_tk, ee_tk, pt_lam, ee_lam, vg_lam;
    // last assembled system
    std::vector<double> grad;                                  // B*12
    std::map<std::pair<int, int>, std::vector<double>> blocks; // (row, col) -> 144 col-major
    std::string err;
};

static void positions(const World& w, const double* qv, std::vector<double>& X) {
    X.resize(w.V * 3);
    for (int b = 0; b < w.B; b++) {
        const double* Q = qv + 12 * b;
        for (int v = w.voff[b]; v < w.voff[b + 1]; v++) {
            const double* xb = &w.xbar[3 * v];
            for (int i = 0; i < 3; i++)
                X[3 * v + i] = ((Q[3 + i] * xb[0] + Q[6 + i] * xb[1]) + Q[9 + i] * xb[2]) + Q[i];
        }
    }
}
static void q_plus(const World& w, const double* dq, std::vector<double>& out) {
    out.resize(12 * w.B);
    for (int i = 0; i < 12 * w.B; i++) out[i] = w.q[i] + (dq ? dq[i] : 0.0);
}
static inline V3 X3(const std::vector<double>& X, int v) { return V3{X[3 * v], X[3 * v + 1], X[3 * v + 2]}; }

struct Box { double lo[3], hi[3]; };
static inline bool overlap(const Box& a, const Box& b) {
    for (int k = 0; k < 3; k++) if (!(a.lo[k] <= b.hi[k] && a.hi[k] >= b.lo[k])) return false;
    return true;
}
static inline void box_add(Box& b, V3 p) {
    double c[3] = {p.x, p.y, p.z};
    for (int k = 0; k < 3; k++) { b.lo[k] = std::min(b.lo[k], c[k]); b.hi[k] = std::max(b.hi[k], c[k]); }
}
static inline Box box_empty() { Box b; for (int k = 0; k < 3; k++) { b.lo[k] = 1e300; b.hi[k] = -1e300; } return b; }
static inline Box box_inflate(Box b, double r) { for (int k = 0; k < 3; k++) { b.lo[k] -= r; b.hi[k] += r; } return b; }

// conservative culling radius: sqrt(d_hat) with a relative safety margin, so
// that no box test rejects a pair the exact FP64 predicate d < d_hat accepts
static double cull_radius(const World& w) { return std::sqrt(w.P.d_hat) * (1.0 + 1e-6) + 1e-12; }

// ---------------------------------------------------------------------------
// Contact sets: all (I,v,J,f) with I != J and d_PT < d_hat; all (I,ei,J,ej)
// with I < J and d_EE < d_hat; all (I,v) with (y - ground)^2 < d_hat.
// This is the set the reference's iAABB test pins against brute force
// (iAABB/test.cpp:317-411); both methods below produce it, sorted.
// ---------------------------------------------------------------------------
static void pair_tests(const World& w, const std::vector<double>& X, int I, int J,
                       const std::vector<int>& vi, const std::vector<int>& fj,
                       const std::vector<int>& vj, const std::vector<int>& fi,
                       const std::vector<int>& ei, const std::vector<int>& ej,
                       std::vector<int>& pt, std::vector<int>& ee) {
    const double r = cull_radius(w), dh = w.P.d_hat;
    auto face_box = [&](int f) { Box b = box_empty(); for (int k = 0; k < 3; k++) box_add(b, X3(X, w.face[3 * f + k])); return b; };
    auto edge_box = [&](int e) { Box b = box_empty(); for (int k = 0; k < 2; k++) box_add(b, X3(X, w.edge[2 * e + k])); return b; };
    auto pt_dir = [&](int A, const std::vector<int>& vs, int Bb, const std::vector<int>& fs) {
        if (!w.P.enable_pt) return;
        std::vector<Box> fb(fs.size());
        for (size_t k = 0; k < fs.size(); k++) fb[k] = face_box(fs[k]);
        for (int v : vs) {
            V3 p = X3(X, v);
            Box vb = box_empty(); box_add(vb, p); vb = box_inflate(vb, r);
            for (size_t k = 0; k < fs.size(); k++) {
                if (!overlap(vb, fb[k])) continue;
                int f = fs[k];
                V3 t0 = X3(X, w.face[3 * f]), t1 = X3(X, w.face[3 * f + 1]), t2 = X3(X, w.face[3 * f + 2]);
                int ty = pt_type(p, t0, t1, t2);
                double d = pt_dist_typed(p, t0, t1, t2, ty);
                if (d < dh) { pt.push_back(A); pt.push_back(v - w.voff[A]); pt.push_back(Bb); pt.push_back(f - w.foff[Bb]); }
            }
        }
    };
    pt_dir(I, vi, J, fj);
    pt_dir(J, vj, I, fi);
    if (w.P.enable_ee) {
        std::vector<Box> eb(ej.size());
        for (size_t k = 0; k < ej.size(); k++) eb[k] = edge_box(ej[k]);
        for (int a : ei) {
            Box ab = box_inflate(edge_box(a), r);
            V3 a0 = X3(X, w.edge[2 * a]), a1 = X3(X, w.edge[2 * a + 1]);
            for (size_t k = 0; k < ej.size(); k++) {
                if (!overlap(ab, eb[k])) continue;
                int b = ej[k];
                V3 b0 = X3(X, w.edge[2 * b]), b1 = X3(X, w.edge[2 * b + 1]);
                int ty = ee_type(a0, a1, b0, b1);
                double d = ee_dist_typed(a0, a1, b0, b1, ty);
                if (d < dh) { ee.push_back(I); ee.push_back(a - w.eoff[I]); ee.push_back(J); ee.push_back(b - w.eoff[J]); }
            }
        }
    }
}

static void sort4(std::vector<int>& v) {
    size_t n = v.size() / 4;
    std::vector<std::array<int, 4>> t(n);
    for (size_t i = 0; i < n; i++) for (int k = 0; k < 4; k++) t[i][k] = v[4 * i + k];
    std::sort(t.begin(), t.end());
    for (size_t i = 0; i < n; i++) for (int k = 0; k < 4; k++) v[4 * i + k] = t[i][k];
}

static std::vector<Box> body_boxes(const World& w, const std::vector<double>& X, const std::vector<double>* X1, double infl) {
    std::vector<Box> bb(w.B);
    for (int b = 0; b < w.B; b++) {
        Box x = box_empty();
        for (int v = w.voff[b]; v < w.voff[b + 1]; v++) { box_add(x, X3(X, v)); if (X1) box_add(x, X3(*X1, v)); }
        bb[b] = box_inflate(x, infl);
    }
    return bb;
}
static std::vector<std::pair<int, int>> body_pairs(const std::vector<Box>& bb) {
    // sort-and-sweep on x (model/iaabb.cpp:184-288 sweeps all three axes)
    int n = (int)bb.size();
    std::vector<int> ord(n);
    for (int i = 0; i < n; i++) ord[i] = i;
    std::sort(ord.begin(), ord.end(), [&](int a, int b) { return bb[a].lo[0] < bb[b].lo[0] || (bb[a].lo[0] == bb[b].lo[0] && a < b); });
    std::vector<std::pair<int, int>> out;
    for (int s = 0; s < n; s++) {
        int a = ord[s];
        for (int t = s + 1; t < n; t++) {
            int b = ord[t];
            if (bb[b].lo[0] > bb[a].hi[0]) break;
            if (overlap(bb[a], bb[b])) out.push_back({std::min(a, b), std::max(a, b)});
        }
    }
    std::sort(out.begin(), out.end());
    return out;
}

// ---------------------------------------------------------------------------
// Spatial hashing, the reference's alternative culling (model/spatial_hashing.cpp:11-211;
// contact sets model/col_set.cpp:20-339; CCD model/upper_bound.cpp:16-228).
// Cell coordinates are floor(max(x - min_xyz, 0) / dx) per axis
// (spatial_hashing.cpp:23-27); the hash concatenates the low `bits` bits of
// the three coordinates (spatial_hashing.cpp:11-21), so cells 2^bits apart
// alias.  A registered primitive goes into every cell of its box
// (register_interval, :43-66); a query gathers the entries of every cell of
// its box, sorts and de-duplicates them (query_interval, :68-113); its own
// body is dropped and the exact predicate decides.  Differences from the
// reference, none of which changes a result:
//   * cells hold any number of entries (the reference keeps n_buffer per cell
//     plus 8 overflow slots and loses candidates beyond that);
//   * a range longer than 2^bits cells stops after 2^bits (the rest alias);
//   * query boxes of contact sets are inflated by 2 r (r = the culling radius)
//     and candidates pass the same primitive box test as the iAABB path, so
//     both methods test exactly the pairs whose boxes overlap;
//   * an edge's CCD box spans its end points at t0 and t1 (the reference's
//     register_edge_trajectory reads the t0 end points twice,
//     spatial_hashing.cpp:150-151).
// ---------------------------------------------------------------------------
struct HashGrid {
    double mn, dx; int bits;
    std::unordered_map<int, std::vector<int>> cell;
};
static HashGrid sh_grid(const World& w) { HashGrid g; g.mn = w.P.sh_min_xyz; g.dx = w.P.sh_dx; g.bits = w.P.sh_bits; return g; }
static int sh_coord(const HashGrid& g, double x) {
    double u = std::max(x - g.mn, 0.0) / g.dx;
    return u < 1073741824.0 ? (int)u : 1073741824;
}
template <class Fn> static void sh_cells(const HashGrid& g, const Box& b, Fn fn) {
    int lo[3], hi[3];
    for (int k = 0; k < 3; k++) {
        lo[k] = sh_coord(g, b.lo[k]);
        hi[k] = std::min(sh_coord(g, b.hi[k]), lo[k] + (1 << g.bits) - 1);
    }
    const int m = (1 << g.bits) - 1;
    for (int i = lo[0]; i <= hi[0]; i++)
        for (int j = lo[1]; j <= hi[1]; j++)
            for (int k = lo[2]; k <= hi[2]; k++) fn(((i & m) << (2 * g.bits)) | ((j & m) << g.bits) | (k & m));
}
static void sh_register(HashGrid& g, const Box& b, int id) {
    sh_cells(g, b, [&](int h) { g.cell[h].push_back(id); });
}
static void sh_query(const HashGrid& g, const Box& b, std::vector<int>& out) {
    out.clear();
    sh_cells(g, b, [&](int h) {
        auto it = g.cell.find(h);
        if (it != g.cell.end()) out.insert(out.end(), it->second.begin(), it->second.end());
    });
    std::sort(out.begin(), out.end());
    out.erase(std::unique(out.begin(), out.end()), out.end());
}
static int n_threads() {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}
static int thread_id() {
#ifdef _OPENMP
    return omp_get_thread_num();
#else
    return 0;
#endif
}

// contact sets (model/col_set.cpp:65-152 PT, :188-296 EE)
static void sh_collision_set(const World& w, const std::vector<double>& X, std::vector<int>& pt, std::vector<int>& ee) {
    const double r = cull_radius(w), dh = w.P.d_hat;
    const int nt = n_threads();
    std::vector<std::vector<int>> tpt(nt), tee(nt);
    if (w.P.enable_pt && w.F > 0) {
        HashGrid g = sh_grid(w);
        for (int v = 0; v < w.V; v++) { Box b = box_empty(); box_add(b, X3(X, v)); sh_register(g, b, v); }
#pragma omp parallel for schedule(dynamic, 64)
        for (int f = 0; f < w.F; f++) {
            std::vector<int> cand;
            const int J = w.fbody[f];
            V3 t0 = X3(X, w.face[3 * f]), t1 = X3(X, w.face[3 * f + 1]), t2 = X3(X, w.face[3 * f + 2]);
            Box fb = box_empty(); box_add(fb, t0); box_add(fb, t1); box_add(fb, t2);
            sh_query(g, box_inflate(fb, 2 * r), cand);
            for (int v : cand) {
                const int I = w.vbody[v];
                if (I == J) continue;
                V3 p = X3(X, v);
                Box vb = box_empty(); box_add(vb, p);
                if (!overlap(box_inflate(vb, r), fb)) continue;
                int ty = pt_type(p, t0, t1, t2);
                if (pt_dist_typed(p, t0, t1, t2, ty) < dh) {
                    std::vector<int>& o = tpt[thread_id()];
                    o.push_back(I); o.push_back(v - w.voff[I]); o.push_back(J); o.push_back(f - w.foff[J]);
                }
            }
        }
    }
    if (w.P.enable_ee && w.E > 0) {
        HashGrid g = sh_grid(w);
        auto ebox = [&](int e) { Box b = box_empty(); box_add(b, X3(X, w.edge[2 * e])); box_add(b, X3(X, w.edge[2 * e + 1])); return b; };
        for (int e = 0; e < w.E; e++) sh_register(g, ebox(e), e);
#pragma omp parallel for schedule(dynamic, 64)
        for (int b = 0; b < w.E; b++) {
            std::vector<int> cand;
            const int J = w.ebody[b];
            const Box bb = ebox(b);
            sh_query(g, box_inflate(bb, 2 * r), cand);
            V3 b0 = X3(X, w.edge[2 * b]), b1 = X3(X, w.edge[2 * b + 1]);
            for (int a : cand) {
                const int I = w.ebody[a];
                if (I >= J) continue;                       // each pair once, I < J (col_set.cpp:282)
                if (!overlap(box_inflate(ebox(a), r), bb)) continue;
                V3 a0 = X3(X, w.edge[2 * a]), a1 = X3(X, w.edge[2 * a + 1]);
                int ty = ee_type(a0, a1, b0, b1);
                if (ee_dist_typed(a0, a1, b0, b1, ty) < dh) {
                    std::vector<int>& o = tee[thread_id()];
                    o.push_back(I); o.push_back(a - w.eoff[I]); o.push_back(J); o.push_back(b - w.eoff[J]);
                }
            }
        }
    }
    for (int t = 0; t < nt; t++) { pt.insert(pt.end(), tpt[t].begin(), tpt[t].end()); ee.insert(ee.end(), tee[t].begin(), tee[t].end()); }
}

// CCD step bound of the PT / EE pairs (model/upper_bound.cpp:85-132 PT, :175-224 EE)
static double sh_ccd_pairs(const World& w, const std::vector<double>& X0, const std::vector<double>& X1) {
    auto tbox = [&](const int* vs, int n) {
        Box b = box_empty();
        for (int c = 0; c < n; c++) { box_add(b, X3(X0, vs[c])); box_add(b, X3(X1, vs[c])); }
        return b;
    };
    double toi = 1.0;
    if (w.P.enable_pt && w.F > 0) {
        HashGrid g = sh_grid(w);
        for (int v = 0; v < w.V; v++) sh_register(g, tbox(&v, 1), v);
#pragma omp parallel for schedule(dynamic, 64) reduction(min : toi)
        for (int f = 0; f < w.F; f++) {
            std::vector<int> cand;
            const int J = w.fbody[f];
            const int* fv = &w.face[3 * f];
            const Box fb = tbox(fv, 3);
            sh_query(g, fb, cand);
            for (int v : cand) {
                if (w.vbody[v] == J) continue;
                if (!overlap(tbox(&v, 1), fb)) continue;
                V3 P0[4] = {X3(X0, v), X3(X0, fv[0]), X3(X0, fv[1]), X3(X0, fv[2])};
                V3 P1[4] = {X3(X1, v), X3(X1, fv[0]), X3(X1, fv[1]), X3(X1, fv[2])};
                toi = std::min(toi, pt_toi(P0, P1));
            }
        }
    }
    if (w.P.enable_ee && w.E > 0) {
        HashGrid g = sh_grid(w);
        for (int e = 0; e < w.E; e++) sh_register(g, tbox(&w.edge[2 * e], 2), e);
#pragma omp parallel for schedule(dynamic, 64) reduction(min : toi)
        for (int b = 0; b < w.E; b++) {
            std::vector<int> cand;
            const int J = w.ebody[b];
            const int* vb = &w.edge[2 * b];
            const Box bb = tbox(vb, 2);
            sh_query(g, bb, cand);
            for (int a : cand) {
                if (w.ebody[a] >= J) continue;
                const int* va = &w.edge[2 * a];
                if (!overlap(tbox(va, 2), bb)) continue;
                V3 P0[4] = {X3(X0, va[0]), X3(X0, va[1]), X3(X0, vb[0]), X3(X0, vb[1])};
                V3 P1[4] = {X3(X1, va[0]), X3(X1, va[1]), X3(X1, vb[0]), X3(X1, vb[1])};
                toi = std::min(toi, ee_toi(P0, P1));
            }
        }
    }
    return toi;
}

static void collision_set(const World& w, const std::vector<double>& X, int method,
                          std::vector<int>& pt, std::vector<int>& ee, std::vector<int>& vg) {
    pt.clear(); ee.clear(); vg.clear();
    const double r = cull_radius(w);
    if (w.P.enable_ground)
        for (int v = 0; v < w.V; v++) {
            double dd = X[3 * v + 1] - w.P.ground_y;
            double d = dd * dd;
            if (d < w.P.d_hat) { vg.push_back(w.vbody[v]); vg.push_back(v - w.voff[w.vbody[v]]); }
        }
    if (method == 2) {
        sh_collision_set(w, X, pt, ee);
        sort4(pt); sort4(ee);
        return;
    }
    std::vector<Box> bb = body_boxes(w, X, nullptr, r);
    std::vector<std::pair<int, int>> pairs;
    if (method == 0) {
        for (int I = 0; I < w.B; I++) for (int J = I + 1; J < w.B; J++) pairs.push_back({I, J});
    } else {
        pairs = body_pairs(bb);
    }
    int np = (int)pairs.size();
    int nt = 1;
#ifdef _OPENMP
    nt = omp_get_max_threads();
#endif
    std::vector<std::vector<int>> tpt(nt), tee(nt);
#pragma omp parallel for schedule(dynamic, 4)
    for (int k = 0; k < np; k++) {
        int tid = 0;
#ifdef _OPENMP
        tid = omp_get_thread_num();
#endif
        int I = pairs[k].first, J = pairs[k].second;
        std::vector<int> vi, vj, fi, fj, ei, ej;
        if (method == 0) {
            for (int v = w.voff[I]; v < w.voff[I + 1]; v++) vi.push_back(v);
            for (int v = w.voff[J]; v < w.voff[J + 1]; v++) vj.push_back(v);
            for (int f = w.foff[I]; f < w.foff[I + 1]; f++) fi.push_back(f);
            for (int f = w.foff[J]; f < w.foff[J + 1]; f++) fj.push_back(f);
            for (int e = w.eoff[I]; e < w.eoff[I + 1]; e++) ei.push_back(e);
            for (int e = w.eoff[J]; e < w.eoff[J + 1]; e++) ej.push_back(e);
        } else {
            // primitive binning against the overlap box (model/iaabb.cpp:333-404)
            if (!overlap(bb[I], bb[J])) continue;
            Box X0;
            for (int a = 0; a < 3; a++) { X0.lo[a] = std::max(bb[I].lo[a], bb[J].lo[a]); X0.hi[a] = std::min(bb[I].hi[a], bb[J].hi[a]); }
            auto bin = [&](int b, std::vector<int>& vs, std::vector<int>& fs, std::vector<int>& es) {
                for (int v = w.voff[b]; v < w.voff[b + 1]; v++) { Box x = box_empty(); box_add(x, X3(X, v)); if (overlap(x, X0)) vs.push_back(v); }
                for (int f = w.foff[b]; f < w.foff[b + 1]; f++) {
                    Box x = box_empty(); for (int c = 0; c < 3; c++) box_add(x, X3(X, w.face[3 * f + c]));
                    if (overlap(x, X0)) fs.push_back(f);
                }
                for (int e = w.eoff[b]; e < w.eoff[b + 1]; e++) {
                    Box x = box_empty(); for (int c = 0; c < 2; c++) box_add(x, X3(X, w.edge[2 * e + c]));
                    if (overlap(x, X0)) es.push_back(e);
                }
            };
            bin(I, vi, fi, ei);
            bin(J, vj, fj, ej);
        }
        pair_tests(w, X, I, J, vi, fj, vj, fi, ei, ej, tpt[tid], tee[tid]);
    }
    for (int t = 0; t < nt; t++) { pt.insert(pt.end(), tpt[t].begin(), tpt[t].end()); ee.insert(ee.end(), tee[t].begin(), tee[t].end()); }
    sort4(pt); sort4(ee);
}

// ---------------------------------------------------------------------------
// CCD step bound along X0 -> X1 (model/iaabb.cpp:270-303 VG; 507-540 PT/EE;
// pairs are filtered by tight swept boxes as in
// model/thread_local_culling.cpp:28-207).
// ---------------------------------------------------------------------------
static double ccd_toi(World& w, const std::vector<double>& X0, const std::vector<double>& X1, int method, int* status) {
    double toi_g = 1.0;
    if (w.P.enable_ground)
        for (int v = 0; v < w.V; v++) {
            double d2 = X1[3 * v + 1] - w.P.ground_y;
            if (d2 < 0) {
                double d1 = X0[3 * v + 1] - w.P.ground_y;
                double t = d1 / (d1 - d2);
                toi_g = std::min(toi_g, t);
            }
        }
    *status = ADB_OK;
    if (toi_g < 1e-6) { if (++w.stall_vg > 1) *status = ADB_E_TOI_STALL; } else w.stall_vg = 0;

    if (method == 2) {
        const double toi_p = sh_ccd_pairs(w, X0, X1);
        if (toi_p < 1e-6) { if (++w.stall_pt > 1) *status = ADB_E_TOI_STALL; } else w.stall_pt = 0;
        return std::min(toi_g, toi_p);
    }
    std::vector<Box> bb = body_boxes(w, X0, &X1, 0.0);
    std::vector<std::pair<int, int>> pairs;
    if (method == 0) { for (int I = 0; I < w.B; I++) for (int J = I + 1; J < w.B; J++) pairs.push_back({I, J}); }
    else pairs = body_pairs(bb);
    auto vbox = [&](int v) { Box b = box_empty(); box_add(b, X3(X0, v)); box_add(b, X3(X1, v)); return b; };
    auto fbox = [&](int f) { Box b = box_empty(); for (int c = 0; c < 3; c++) { box_add(b, X3(X0, w.face[3 * f + c])); box_add(b, X3(X1, w.face[3 * f + c])); } return b; };
    auto ebox = [&](int e) { Box b = box_empty(); for (int c = 0; c < 2; c++) { box_add(b, X3(X0, w.edge[2 * e + c])); box_add(b, X3(X1, w.edge[2 * e + c])); } return b; };
    double toi_p = 1.0;
    int np = (int)pairs.size();
#pragma omp parallel for schedule(dynamic, 4) reduction(min : toi_p)
    for (int k = 0; k < np; k++) {
        int I = pairs[k].first, J = pairs[k].second;
        if (!overlap(bb[I], bb[J])) continue;
        auto ptdir = [&](int A, int Bb) {
            if (!w.P.enable_pt) return;
            for (int v = w.voff[A]; v < w.voff[A + 1]; v++) {
                Box vb = vbox(v);
                if (!overlap(vb, bb[Bb])) continue;
                for (int f = w.foff[Bb]; f < w.foff[Bb + 1]; f++) {
                    if (!overlap(vb, fbox(f))) continue;
                    V3 P0[4] = {X3(X0, v), X3(X0, w.face[3 * f]), X3(X0, w.face[3 * f + 1]), X3(X0, w.face[3 * f + 2])};
                    V3 P1[4] = {X3(X1, v), X3(X1, w.face[3 * f]), X3(X1, w.face[3 * f + 1]), X3(X1, w.face[3 * f + 2])};
                    toi_p = std::min(toi_p, pt_toi(P0, P1));
                }
            }
        };
        ptdir(I, J); ptdir(J, I);
        if (w.P.enable_ee)
            for (int a = w.eoff[I]; a < w.eoff[I + 1]; a++) {
                Box ab = ebox(a);
                if (!overlap(ab, bb[J])) continue;
                for (int b = w.eoff[J]; b < w.eoff[J + 1]; b++) {
                    if (!overlap(ab, ebox(b))) continue;
                    V3 P0[4] = {X3(X0, w.edge[2 * a]), X3(X0, w.edge[2 * a + 1]), X3(X0, w.edge[2 * b]), X3(X0, w.edge[2 * b + 1])};
                    V3 P1[4] = {X3(X1, w.edge[2 * a]), X3(X1, w.edge[2 * a + 1]), X3(X1, w.edge[2 * b]), X3(X1, w.edge[2 * b + 1])};
                    toi_p = std::min(toi_p, ee_toi(P0, P1));
                }
            }
    }
    if (toi_p < 1e-6) { if (++w.stall_pt > 1) *status = ADB_E_TOI_STALL; } else w.stall_pt = 0;
    return std::min(toi_g, toi_p);
}

// ---------------------------------------------------------------------------
// Inertia + orthogonality (model/mutils.cpp:53-76,197-201,245-252,
// model/othogonal_energy.cpp:16-62)
// ---------------------------------------------------------------------------
static void ortho(const double* q, double kap, double* E, double* g, double* H) {
    V3 a[4] = {load3(q), load3(q + 3), load3(q + 6), load3(q + 9)};
    if (E) {
        double e = 0;
        for (int i = 1; i < 4; i++) for (int j = 1; j < 4; j++) { double t = dot(a[i], a[j]) - 1.0 * (i == j); e += t * t; }
        *E = e * kap;
    }
    if (g) {
        for (int k = 0; k < 3; k++) g[k] = 0;
        for (int i = 1; i < 4; i++) {
            double gg[3] = {0, 0, 0};
            for (int j = 1; j < 4; j++) {
                double s = dot(a[i], a[j]) - (i == j ? 1.0 : 0.0);
                gg[0] += s * a[j].x; gg[1] += s * a[j].y; gg[2] += s * a[j].z;
            }
            for (int k = 0; k < 3; k++) g[3 * i + k] = 4 * kap * gg[k];
        }
    }
    if (H) {
        std::memset(H, 0, 144 * sizeof(double));
        for (int i = 1; i < 4; i++)
            for (int j = 1; j < 4; j++) {
                double h[3][3];
                double ai[3] = {a[i].x, a[i].y, a[i].z}, aj[3] = {a[j].x, a[j].y, a[j].z};
                if (i == j) {
                    double s = dot(a[i], a[i]) - 1;
                    for (int r = 0; r < 3; r++) for (int c = 0; c < 3; c++) h[r][c] = 2 * ai[r] * ai[c] + (r == c ? s : 0.0);
                    for (int k = 1; k < 4; k++) if (k != i) {
                        double ak[3] = {a[k].x, a[k].y, a[k].z};
                        for (int r = 0; r < 3; r++) for (int c = 0; c < 3; c++) h[r][c] += ak[r] * ak[c];
                    }
                } else {
                    double s = dot(a[j], a[i]);
                    for (int r = 0; r < 3; r++) for (int c = 0; c < 3; c++) h[r][c] = (r == c ? s : 0.0) + aj[r] * ai[c];
                }
                for (int r = 0; r < 3; r++) for (int c = 0; c < 3; c++) H[(3 * j + c) * 12 + 3 * i + r] = h[r][c] * 4 * kap;
            }
    }
}
static void q_tilde(const World& w, int b, double* qt) {
    const double dt = w.P.dt;
    for (int i = 0; i < 12; i++) qt[i] = w.q0[12 * b + i] + dt * w.dqdt[12 * b + i];
    for (int i = 0; i < 3; i++) qt[i] += dt * dt * w.P.gravity[i];
}
static double inertia_energy(const World& w, int b, const double* qb) {
    if (w.mass[b] < 0.0) return 0.0;
    double qt[12]; q_tilde(w, b, qt);
    double eo; ortho(qb, w.P.kappa_ortho, &eo, nullptr, nullptr);
    double pp = 0, rr = 0;
    for (int i = 0; i < 3; i++) { double d = qb[i] - qt[i]; pp += d * d; }
    for (int i = 3; i < 12; i++) { double d = qb[i] - qt[i]; rr += d * d; }
    return eo * w.P.dt * w.P.dt + 0.5 * (pp * w.mass[b] + rr * w.Ic[b]);
}

// J(xbar) = [I, xbar0 I, xbar1 I, xbar2 I] (3x12, model/barrier.cpp:42-54)
static void jac(const double* xb, double J[36]) {   // row-major 3x12
    std::memset(J, 0, 36 * sizeof(double));
    double w4[4] = {1.0, xb[0], xb[1], xb[2]};
    for (int k = 0; k < 4; k++) for (int i = 0; i < 3; i++) J[i * 12 + 3 * k + i] = w4[k];
}
// out(12x12) += Ja^T * Hblk(3x3 taken from H at rows ra, cols cb) * Jb
static void add_JtHJ(const double Ja[36], const double* H, int ra, int cb, const double Jb[36], double* out) {
    double tmp[3][12];
    for (int r = 0; r < 3; r++)
        for (int c = 0; c < 12; c++) {
            double s = 0;
            for (int k = 0; k < 3; k++) s += H[(cb + k) * 12 + ra + r] * Jb[k * 12 + c];
            tmp[r][c] = s;
        }
    for (int r = 0; r < 12; r++)
        for (int c = 0; c < 12; c++) {
            double s = 0;
            for (int k = 0; k < 3; k++) s += Ja[k * 12 + r] * tmp[k][c];
            out[c * 12 + r] += s;
        }
}

// ---------------------------------------------------------------------------
// Contact terms at the current q (model/abd.cpp:305-404 and the files it calls)
// ---------------------------------------------------------------------------
static void contact_pt(const World& w, const std::vector<double>& X, const std::vector<double>& Xs, const int* ij, Contact& c) {
    int I = ij[0], v = w.voff[I] + ij[1], J = ij[2], f = w.foff[J] + ij[3];
    int fv[3] = {w.face[3 * f], w.face[3 * f + 1], w.face[3 * f + 2]};
    int ids[4] = {v, fv[0], fv[1], fv[2]};
    V3 P[4]; double x[12], vst[12];
    for (int k = 0; k < 4; k++) { P[k] = X3(X, ids[k]); store3(x + 3 * k, P[k]); store3(vst + 3 * k, X3(X, ids[k]) - X3(Xs, ids[k])); }
    c.I = I; c.J = J;
    int ty = pt_type(P[0], P[1], P[2], P[3]);
    double d = pt_dist_typed(P[0], P[1], P[2], P[3], ty);
    c.active = d < w.P.d_hat;
    std::memset(c.Tk, 0, sizeof c.Tk); c.lam = 0.0;
    if (!c.active) return;
    // pt_uktk (model/pt_ipc.cpp:22-107)
    double tl[3], Pk[6];
    pt_closest(P, ty, tl, Pk);
    double G[36] = {0};     // gamma 3x12 row-major
    for (int i = 0; i < 3; i++) { G[i * 12 + i] = -1.0; for (int j = 0; j < 3; j++) G[i * 12 + i + 3 + 3 * j] = tl[j]; }
    for (int r = 0; r < 2; r++) for (int col = 0; col < 12; col++) {
        double s = 0; for (int k = 0; k < 3; k++) s += Pk[3 * r + k] * G[k * 12 + col];
        c.Tk[r * 12 + col] = s;
    }
    double uk[2];
    for (int r = 0; r < 2; r++) { double s = 0; for (int col = 0; col < 12; col++) s += c.Tk[r * 12 + col] * vst[col]; uk[r] = s; }
    c.lam = -barrier_d1(d, w.P.d_hat, w.P.kappa) * 2 * std::sqrt(d);
    double H[144], g[12];
    hess_pt(x, ty, d, w.P.d_hat, w.P.kappa, w.P.psd, H, g);
    if (w.P.pt_fric) {
        double T[24]; for (int col = 0; col < 12; col++) { T[col] = c.Tk[col]; T[12 + col] = c.Tk[12 + col]; }
        friction(uk, c.lam, T, 12, w.P.mu, w.P.dt * w.P.evh, w.P.max_uk, g, H);
    }
    double Jp[36], Jt[3][36];
    jac(&w.xbar[3 * v], Jp);
    for (int k = 0; k < 3; k++) jac(&w.xbar[3 * fv[k]], Jt[k]);
    std::memset(c.hp, 0, sizeof c.hp); std::memset(c.ht, 0, sizeof c.ht); std::memset(c.off, 0, sizeof c.off);
    add_JtHJ(Jp, H, 0, 0, Jp, c.hp);
    for (int a = 0; a < 3; a++) for (int b = 0; b < 3; b++) add_JtHJ(Jt[a], H, 3 + 3 * a, 3 + 3 * b, Jt[b], c.ht);
    for (int b = 0; b < 3; b++) add_JtHJ(Jp, H, 0, 3 + 3 * b, Jt[b], c.off);
    for (int r = 0; r < 12; r++) {
        double s = 0; for (int k = 0; k < 3; k++) s += Jp[k * 12 + r] * g[k];
        c.gp[r] = s;
        double t = 0; for (int a = 0; a < 3; a++) for (int k = 0; k < 3; k++) t += Jt[a][k * 12 + r] * g[3 + 3 * a + k];
        c.gt[r] = t;
    }
}
static void contact_ee(const World& w, const std::vector<double>& X, const std::vector<double>& Xs, const int* ij, Contact& c) {
    int I = ij[0], a = w.eoff[I] + ij[1], J = ij[2], b = w.eoff[J] + ij[3];
    int ids[4] = {w.edge[2 * a], w.edge[2 * a + 1], w.edge[2 * b], w.edge[2 * b + 1]};
    V3 P[4]; double x[12], vst[12];
    for (int k = 0; k < 4; k++) { P[k] = X3(X, ids[k]); store3(x + 3 * k, P[k]); store3(vst + 3 * k, X3(X, ids[k]) - X3(Xs, ids[k])); }
    c.I = I; c.J = J;
    int ty = ee_type(P[0], P[1], P[2], P[3]);
    double d = ee_dist_typed(P[0], P[1], P[2], P[3], ty);
    c.active = d < w.P.d_hat;
    std::memset(c.Tk, 0, sizeof c.Tk); c.lam = 0.0;
    if (!c.active) return;
    double H[144], g[12], m;
    hess_ee(x, ty, d, w.P.d_hat, w.P.kappa, w.P.eps_x, w.P.psd, H, g, &m);
    // ee_uktk (model/ee_ipc.cpp:21-118)
    double lm[2], Pk[6];
    ee_closest(P, ty, lm, Pk);
    double L[4] = {1 - lm[0], lm[0], 1 - lm[1], lm[1]};
    double G[36] = {0};
    for (int i = 0; i < 3; i++) { G[i * 12 + i] = -L[0]; G[i * 12 + i + 3] = -L[1]; G[i * 12 + i + 6] = L[2]; G[i * 12 + i + 9] = L[3]; }
    for (int r = 0; r < 2; r++) for (int col = 0; col < 12; col++) {
        double s = 0; for (int k = 0; k < 3; k++) s += Pk[3 * r + k] * G[k * 12 + col];
        c.Tk[r * 12 + col] = s;
    }
    double uk[2];
    for (int r = 0; r < 2; r++) { double s = 0; for (int col = 0; col < 12; col++) s += c.Tk[r * 12 + col] * vst[col]; uk[r] = s; }
    c.lam = -barrier_d1(d, w.P.d_hat, w.P.kappa) * 2 * std::sqrt(d);
    if (w.P.ee_fric) {
        double T[24]; for (int col = 0; col < 12; col++) { T[col] = c.Tk[col]; T[12 + col] = c.Tk[12 + col]; }
        friction(uk, c.lam, T, 12, w.P.mu, w.P.dt * w.P.evh, w.P.max_uk, g, H);
    }
    double J0[2][36], J1[2][36];
    for (int k = 0; k < 2; k++) { jac(&w.xbar[3 * ids[k]], J0[k]); jac(&w.xbar[3 * ids[2 + k]], J1[k]); }
    std::memset(c.hp, 0, sizeof c.hp); std::memset(c.ht, 0, sizeof c.ht); std::memset(c.off, 0, sizeof c.off);
    for (int p = 0; p < 2; p++) for (int s = 0; s < 2; s++) {
        add_JtHJ(J0[p], H, 3 * p, 3 * s, J0[s], c.hp);
        add_JtHJ(J1[p], H, 6 + 3 * p, 6 + 3 * s, J1[s], c.ht);
        add_JtHJ(J0[p], H, 3 * p, 6 + 3 * s, J1[s], c.off);
    }
    for (int r = 0; r < 12; r++) {
        double s = 0, t = 0;
        for (int p = 0; p < 2; p++) for (int k = 0; k < 3; k++) { s += J0[p][k * 12 + r] * g[3 * p + k]; t += J1[p][k * 12 + r] * g[6 + 3 * p + k]; }
        c.gp[r] = s; c.gt[r] = t;
    }
}

// ---------------------------------------------------------------------------
// Block-sparse Cholesky with a minimum-degree body ordering; stands in for
// MKL PardisoLLT (model/abd.cpp:268-276).
// ---------------------------------------------------------------------------
static bool chol12(double* A) {   // in place, lower, column-major; returns false if not PD
    for (int j = 0; j < 12; j++) {
        double s = A[j * 12 + j];
        for (int k = 0; k < j; k++) s -= A[k * 12 + j] * A[k * 12 + j];
        if (!(s > 0)) return false;
        double l = std::sqrt(s);
        A[j * 12 + j] = l;
        for (int i = j + 1; i < 12; i++) {
            double t = A[j * 12 + i];
            for (int k = 0; k < j; k++) t -= A[k * 12 + i] * A[k * 12 + j];
            A[j * 12 + i] = t / l;
        }
        for (int i = 0; i < j; i++) A[j * 12 + i] = 0.0;
    }
    return true;
}
static int sparse_solve(int B, const std::map<std::pair<int, int>, std::vector<double>>& blocks, const double* rhs, double* x) {
    // adjacency
    std::vector<std::set<int>> adj(B);
    for (auto& kv : blocks) if (kv.first.first != kv.first.second) { adj[kv.first.first].insert(kv.first.second); adj[kv.first.second].insert(kv.first.first); }
    // minimum degree ordering on the elimination graph
    std::vector<int> order; order.reserve(B);
    std::vector<char> done(B, 0);
    std::set<std::pair<int, int>> pq;
    for (int i = 0; i < B; i++) pq.insert({(int)adj[i].size(), i});
    std::vector<std::set<int>> eg = adj;
    while (!pq.empty()) {
        int v = pq.begin()->second; pq.erase(pq.begin());
        done[v] = 1; order.push_back(v);
        std::vector<int> nb(eg[v].begin(), eg[v].end());
        for (int a : nb) { pq.erase({(int)eg[a].size(), a}); eg[a].erase(v); }
        for (size_t s = 0; s < nb.size(); s++) for (size_t t = s + 1; t < nb.size(); t++) { eg[nb[s]].insert(nb[t]); eg[nb[t]].insert(nb[s]); }
        for (int a : nb) pq.insert({(int)eg[a].size(), a});
        eg[v].clear();
    }
    std::vector<int> pos(B);
    for (int i = 0; i < B; i++) pos[order[i]] = i;
    // lower blocks L[(row, col)] in permuted indices, row >= col
    std::vector<std::map<int, std::vector<double>>> col(B);
    for (auto& kv : blocks) {
        int r = pos[kv.first.first], c = pos[kv.first.second];
        if (r < c) continue;
        col[c][r] = kv.second;
    }
    for (int k = 0; k < B; k++) {
        auto& ck = col[k];
        if (!ck.count(k)) ck[k] = std::vector<double>(144, 0.0);
        double* D = ck[k].data();
        if (!chol12(D)) return ADB_E_NAN;
        // L_ik = A_ik L_kk^{-T}
        for (auto& kv : ck) {
            if (kv.first == k) continue;
            double* A = kv.second.data();
            for (int rr = 0; rr < 12; rr++)
                for (int j = 0; j < 12; j++) {
                    double t = A[j * 12 + rr];
                    for (int m = 0; m < j; m++) t -= A[m * 12 + rr] * D[m * 12 + j];
                    A[j * 12 + rr] = t / D[j * 12 + j];
                }
        }
        // Schur update
        std::vector<int> rows;
        for (auto& kv : ck) if (kv.first != k) rows.push_back(kv.first);
        for (size_t s = 0; s < rows.size(); s++)
            for (size_t t = 0; t <= s; t++) {
                int i = std::max(rows[s], rows[t]), j = std::min(rows[s], rows[t]);
                const double* Li = ck[i].data(); const double* Lj = ck[j].data();
                auto it = col[j].find(i);
                if (it == col[j].end()) it = col[j].emplace(i, std::vector<double>(144, 0.0)).first;
                double* T = it->second.data();
                for (int c = 0; c < 12; c++) for (int r = 0; r < 12; r++) {
                    double sum = 0; for (int m = 0; m < 12; m++) sum += Li[m * 12 + r] * Lj[m * 12 + c];
                    T[c * 12 + r] -= sum;
                }
            }
    }
    // forward / backward substitution in permuted order
    std::vector<double> y(12 * B);
    for (int b = 0; b < B; b++) for (int i = 0; i < 12; i++) y[12 * pos[b] + i] = rhs[12 * b + i];
    for (int k = 0; k < B; k++) {
        const double* D = col[k][k].data();
        double* yk = &y[12 * k];
        for (int i = 0; i < 12; i++) { double t = yk[i]; for (int m = 0; m < i; m++) t -= D[m * 12 + i] * yk[m]; yk[i] = t / D[i * 12 + i]; }
        for (auto& kv : col[k]) {
            if (kv.first == k) continue;
            const double* L = kv.second.data(); double* yi = &y[12 * kv.first];
            for (int r = 0; r < 12; r++) { double t = 0; for (int m = 0; m < 12; m++) t += L[m * 12 + r] * yk[m]; yi[r] -= t; }
        }
    }
    for (int k = B - 1; k >= 0; k--) {
        double* yk = &y[12 * k];
        for (auto& kv : col[k]) {
            if (kv.first == k) continue;
            const double* L = kv.second.data(); const double* yi = &y[12 * kv.first];
            for (int m = 0; m < 12; m++) { double t = 0; for (int r = 0; r < 12; r++) t += L[m * 12 + r] * yi[r]; yk[m] -= t; }
        }
        const double* D = col[k][k].data();
        for (int i = 11; i >= 0; i--) { double t = yk[i]; for (int m = i + 1; m < 12; m++) t -= D[i * 12 + m] * yk[m]; yk[i] = t / D[i * 12 + i]; }
    }
    for (int b = 0; b < B; b++) for (int i = 0; i < 12; i++) x[12 * b + i] = y[12 * pos[b] + i];
    return ADB_OK;
}

// ---------------------------------------------------------------------------
// Assembly at the current q: inertia + contacts + VG (model/abd.cpp:68-74,
// 319-340, 385-485; model/assemble_matrices.cpp:28-59)
// ---------------------------------------------------------------------------
static void assemble(World& w, long long* nactive) {
    std::vector<double> X, Xs;
    positions(w, w.q.data(), X);
    positions(w, w.q0.data(), Xs);
    const int B = w.B;
    std::vector<double> bg(12 * B), bh(144 * B);
    for (int b = 0; b < B; b++) {
        double* g = &bg[12 * b]; double* H = &bh[144 * b];
        if (w.mass[b] < 0.0) {
            for (int i = 0; i < 12; i++) g[i] = 0;
            std::memset(H, 0, 144 * sizeof(double));
            for (int i = 0; i < 12; i++) H[i * 12 + i] = 1.0;
            continue;
        }
        double go[12], Ho[144], qt[12];
        ortho(&w.q[12 * b], w.P.kappa_ortho, nullptr, go, Ho);
        q_tilde(w, b, qt);
        const double dt = w.P.dt;
        for (int i = 0; i < 12; i++) {
            double m = i < 3 ? w.mass[b] : w.Ic[b];
            g[i] = dt * dt * go[i] + m * (w.q[12 * b + i] - qt[i]);
        }
        for (int c = 0; c < 12; c++) for (int r = 0; r < 12; r++)
            H[c * 12 + r] = (r == c ? (r < 3 ? w.mass[b] : w.Ic[b]) : 0.0) + Ho[c * 12 + r] * dt * dt;
    }
    // PT / EE terms (computed independently, accumulated in canonical order)
    int npt = (int)w.pt.size() / 4, nee = (int)w.ee.size() / 4, nvg = (int)w.vg.size() / 2;
    std::vector<Contact> cpt(npt), cee(nee);
#pragma omp parallel for schedule(dynamic, 16)
    for (int k = 0; k < npt; k++) contact_pt(w, X, Xs, &w.pt[4 * k], cpt[k]);
#pragma omp parallel for schedule(dynamic, 16)
    for (int k = 0; k < nee; k++) contact_ee(w, X, Xs, &w.ee[4 * k], cee[k]);
    w.pt_tk.assign(24 * npt, 0.0); w.pt_lam.assign(npt, 0.0);
    w.ee_tk.assign(24 * nee, 0.0); w.ee_lam.assign(nee, 0.0);
    w.vg_lam.assign(nvg, 0.0);
    w.blocks.clear();
    auto blk = [&](int r, int c) -> double* {
        auto it = w.blocks.find({r, c});
        if (it == w.blocks.end()) it = w.blocks.emplace(std::make_pair(r, c), std::vector<double>(144, 0.0)).first;
        return it->second.data();
    };
    // sparsity pattern: PT + EE body pairs + diagonal (model/sparse.cpp:13-57)
    for (int b = 0; b < B; b++) blk(b, b);
    w.grad.assign(12 * B, 0.0);
    for (int b = 0; b < B; b++) for (int i = 0; i < 12; i++) w.grad[12 * b + i] = bg[12 * b + i];
    long long na = 0;
    auto scatter = [&](const Contact& c) {
        blk(c.I, c.J); blk(c.J, c.I);
        if (!c.active) return;
        na++;
        bool di = w.mass[c.I] > 0.0, dj = w.mass[c.J] > 0.0;
        if (dj) {
            if (di) { double* o = blk(c.I, c.J); for (int i = 0; i < 144; i++) o[i] += c.off[i]; }
            double* t = blk(c.J, c.J); for (int i = 0; i < 144; i++) t[i] += c.ht[i];
            for (int i = 0; i < 12; i++) w.grad[12 * c.J + i] += c.gt[i];
        }
        if (di) {
            if (dj) { double* o = blk(c.J, c.I); for (int cc = 0; cc < 12; cc++) for (int r = 0; r < 12; r++) o[cc * 12 + r] += c.off[r * 12 + cc]; }
            double* p = blk(c.I, c.I); for (int i = 0; i < 144; i++) p[i] += c.hp[i];
            for (int i = 0; i < 12; i++) w.grad[12 * c.I + i] += c.gp[i];
        }
    };
    for (int k = 0; k < npt; k++) { scatter(cpt[k]); std::memcpy(&w.pt_tk[24 * k], cpt[k].Tk, sizeof cpt[k].Tk); w.pt_lam[k] = cpt[k].lam; }
    for (int k = 0; k < nee; k++) { scatter(cee[k]); std::memcpy(&w.ee_tk[24 * k], cee[k].Tk, sizeof cee[k].Tk); w.ee_lam[k] = cee[k].lam; }
    // VG terms into the body's own gradient / Hessian (model/vg_ipc.cpp:12-30)
    for (int k = 0; k < nvg; k++) {
        int I = w.vg[2 * k], v = w.voff[I] + w.vg[2 * k + 1];
        V3 p = X3(X, v), p0 = X3(Xs, v);
        double dd = p.y - w.P.ground_y, d = dd * dd;
        if (!(d < w.P.d_hat)) continue;
        na++;
        double uk[2] = {p.x - p0.x, p.z - p0.z};
        w.vg_lam[k] = -barrier_d1(d, w.P.d_hat, w.P.kappa) * 2 * dd;
        if (w.mass[I] < 0.0) continue;
        double J[36]; jac(&w.xbar[3 * v], J);
        double gx = 1.0 * barrier_d1(dd * dd, w.P.d_hat, w.P.kappa) * 2 * dd;
        double gq[12];
        for (int i = 0; i < 12; i++) gq[i] = J[1 * 12 + i] * dd * 2;
        double B2 = barrier_d2(dd * dd, w.P.d_hat, w.P.kappa);
        double* g = &bg[12 * I]; double* H = &bh[144 * I];
        for (int i = 0; i < 12; i++) w.grad[12 * I + i] += gx * J[12 + i];
        for (int c = 0; c < 12; c++) for (int r = 0; r < 12; r++) H[c * 12 + r] += B2 * gq[r] * gq[c];
        if (w.P.vg_fric) {
            double T[24];   // J^T Pk, Pk = [e_x, e_z]
            for (int i = 0; i < 12; i++) { T[i] = J[0 * 12 + i]; T[12 + i] = J[2 * 12 + i]; }
            double gtmp[12] = {0};
            friction(uk, w.vg_lam[k], T, 12, w.P.mu, w.P.dt * w.P.evh, w.P.max_uk, gtmp, H);
            for (int i = 0; i < 12; i++) w.grad[12 * I + i] += gtmp[i];
        }
        (void)g;
    }
    for (int b = 0; b < B; b++) { double* D = blk(b, b); for (int i = 0; i < 144; i++) D[i] += bh[144 * b + i]; }
    if (nactive) *nactive += na;
}

// ---------------------------------------------------------------------------
// Line-search energies (model/energy.cpp:19-86,195-270)
// ---------------------------------------------------------------------------
static double E_barrier_inert(const World& w, const std::vector<double>& q1, const std::vector<double>& X,
                              const std::vector<int>& pt, const std::vector<int>& ee, const std::vector<int>& vg) {
    double e = 0.0;
    for (int b = 0; b < w.B; b++) e += inertia_energy(w, b, &q1[12 * b]);
    if (w.P.enable_pt)
        for (size_t k = 0; k < pt.size() / 4; k++) {
            const int* ij = &pt[4 * k];
            int v = w.voff[ij[0]] + ij[1], f = w.foff[ij[2]] + ij[3];
            V3 p = X3(X, v), t0 = X3(X, w.face[3 * f]), t1 = X3(X, w.face[3 * f + 1]), t2 = X3(X, w.face[3 * f + 2]);
            double d = pt_dist_typed(p, t0, t1, t2, pt_type(p, t0, t1, t2));
            e += barrier(d, w.P.d_hat, w.P.kappa);
        }
    if (w.P.enable_ee)
        for (size_t k = 0; k < ee.size() / 4; k++) {
            const int* ij = &ee[4 * k];
            int a = w.eoff[ij[0]] + ij[1], b = w.eoff[ij[2]] + ij[3];
            V3 a0 = X3(X, w.edge[2 * a]), a1 = X3(X, w.edge[2 * a + 1]), b0 = X3(X, w.edge[2 * b]), b1 = X3(X, w.edge[2 * b + 1]);
            double eps = sqn(a0 - a1) * sqn(b0 - b1) * w.P.eps_x;
            double m = mollifier_value(a0, a1, b0, b1, eps);
            double d = ee_dist_typed(a0, a1, b0, b1, ee_type(a0, a1, b0, b1));
            e += barrier(d, w.P.d_hat, w.P.kappa) * m;
        }
    if (w.P.enable_ground)
        for (size_t k = 0; k < vg.size() / 2; k++) {
            int v = w.voff[vg[2 * k]] + vg[2 * k + 1];
            double dd = X[3 * v + 1] - w.P.ground_y, d = dd * dd;
            if (d < w.P.d_hat) e += barrier(d, w.P.d_hat, w.P.kappa);
        }
    return e;
}
static double E_fric(const World& w, const std::vector<double>& X, const std::vector<double>& Xs) {
    double ef = 0.0;
    const double eps = w.P.dt * w.P.evh;
    auto uk_of = [&](const double* Tk, const int ids[4]) {
        double vs[12];
        for (int k = 0; k < 4; k++) store3(vs + 3 * k, X3(X, ids[k]) - X3(Xs, ids[k]));
        double u0 = 0, u1 = 0;
        for (int c = 0; c < 12; c++) { u0 += Tk[c] * vs[c]; u1 += Tk[12 + c] * vs[c]; }
        return std::sqrt(u0 * u0 + u1 * u1);
    };
    if (w.P.enable_pt && w.P.pt_fric)
        for (size_t k = 0; k < w.pt.size() / 4; k++) {
            const int* ij = &w.pt[4 * k];
            int v = w.voff[ij[0]] + ij[1], f = w.foff[ij[2]] + ij[3];
            int ids[4] = {v, w.face[3 * f], w.face[3 * f + 1], w.face[3 * f + 2]};
            ef += w.P.mu * w.pt_lam[k] * f0_sf(uk_of(&w.pt_tk[24 * k], ids), eps);
        }
    if (w.P.enable_ee && w.P.ee_fric)
        for (size_t k = 0; k < w.ee.size() / 4; k++) {
            const int* ij = &w.ee[4 * k];
            int a = w.eoff[ij[0]] + ij[1], b = w.eoff[ij[2]] + ij[3];
            int ids[4] = {w.edge[2 * a], w.edge[2 * a + 1], w.edge[2 * b], w.edge[2 * b + 1]};
            ef += w.P.mu * w.ee_lam[k] * f0_sf(uk_of(&w.ee_tk[24 * k], ids), eps);
        }
    if (w.P.enable_ground && w.P.vg_fric)
        for (size_t k = 0; k < w.vg.size() / 2; k++) {
            int v = w.voff[w.vg[2 * k]] + w.vg[2 * k + 1];
            V3 u = X3(X, v) - X3(Xs, v);
            ef += w.P.mu * w.vg_lam[k] * f0_sf(std::sqrt(u.x * u.x + u.z * u.z), eps);
        }
    return ef;
}

static double now_s() { return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count(); }

static int step(World& w, adb_step_stats* st) {
    adb_step_stats S; std::memset(&S, 0, sizeof S); S.toi_min = 1.0;
    double t_start = now_s();
    const int B = w.B, n = 12 * B;
    w.q = w.q0;
    std::vector<double> X, Xs, X1, qv;
    positions(w, w.q.data(), X);
    positions(w, w.q0.data(), Xs);
    const int set_method = (w.P.iaabb & 1) ? 1 : 2, ccd_method = (w.P.iaabb & 2) ? 1 : 2;
    collision_set(w, X, set_method, w.pt, w.ee, w.vg);
    int status = ADB_OK;
    std::vector<double> dq(n);
    bool converged = false;
    for (int iter = 0; iter < w.P.max_iter; iter++) {
        double t0 = now_s();
        assemble(w, &S.contacts_processed);
        double t1 = now_s(); S.t_ipc += t1 - t0;
        std::vector<double> sol(n);
        int rc = sparse_solve(B, w.blocks, w.grad.data(), sol.data());
        if (rc != ADB_OK) { status = rc; w.err = "factorization failed"; break; }
        double nrm = 0;
        for (int i = 0; i < n; i++) { dq[i] = -sol[i]; nrm += dq[i] * dq[i]; }
        double sup_dq = std::sqrt(nrm);
        double t2 = now_s(); S.t_solver += t2 - t1;
        if (std::isnan(sup_dq)) { status = ADB_E_NAN; w.err = "solver nan"; break; }
        // CCD step bound (model/abd.cpp:142-196)
        double toi = 1.0, factor = 1.0;
        if (w.P.upper_bound) {
            q_plus(w, dq.data(), qv);
            positions(w, qv.data(), X1);
            int cs = ADB_OK;
            toi = ccd_toi(w, X, X1, ccd_method, &cs);
            if (cs != ADB_OK) { status = cs; w.err = "toi stall"; break; }
        }
        if (toi < 1.0) factor = w.P.backoff;
        S.toi_min = std::min(S.toi_min, toi);
        double sc = factor * toi;
        for (int i = 0; i < n; i++) dq[i] *= sc;
        double t3 = now_s(); S.t_ccd += t3 - t2;
        // line search (model/line_search.cpp:39-136)
        double alpha = 1.0;
        std::vector<int> npt, nee, nvg;
        if (w.P.line_search) {
            std::vector<double> zero(n, 0.0);
            double E0 = E_barrier_inert(w, w.q, X, w.pt, w.ee, w.vg) + E_fric(w, X, Xs);
            double qdg = 0, gn = 0;
            for (int i = 0; i < n; i++) { qdg += dq[i] * w.grad[i]; gn += w.grad[i] * w.grad[i]; }
            gn = std::sqrt(gn);
            std::vector<double> dqk(n);
            int trials = 0;
            while (true) {
                for (int i = 0; i < n; i++) dqk[i] = dq[i] * alpha;
                q_plus(w, dqk.data(), qv);
                positions(w, qv.data(), X1);
                collision_set(w, X1, set_method, npt, nee, nvg);
                double E1 = E_barrier_inert(w, qv, X1, npt, nee, nvg) + E_fric(w, X1, Xs);
                bool wolfe = E1 <= E0 + w.P.c1 * alpha * qdg;
                alpha /= 2;
                trials++;
                if (wolfe || !(gn > 1e-3)) break;
                if (trials >= w.P.max_ls_iter) { status = ADB_E_NOT_CONVERGED; w.err = "line search exhausted"; break; }
            }
            S.ls_trials += trials;
            alpha *= 2;
            if (status != ADB_OK) break;
        } else {
            q_plus(w, dq.data(), qv);
            positions(w, qv.data(), X1);
            collision_set(w, X1, set_method, npt, nee, nvg);
        }
        for (int i = 0; i < n; i++) dq[i] *= alpha;
        for (int i = 0; i < n; i++) w.q[i] += dq[i];
        w.pt.swap(npt); w.ee.swap(nee); w.vg.swap(nvg);
        positions(w, w.q.data(), X);
        S.t_line_search += now_s() - t3;
        S.newton_iters++;
        if (sup_dq < w.P.tol) { converged = true; break; }
    }
    if (status == ADB_OK && !converged) status = ADB_E_NOT_CONVERGED;
    if (status == ADB_OK || status == ADB_E_NOT_CONVERGED) {
        for (int i = 0; i < n; i++) { w.dqdt[i] = (w.q[i] - w.q0[i]) / w.P.dt; w.q0[i] = w.q[i]; }
    }
    S.n_pt = (int)w.pt.size() / 4; S.n_ee = (int)w.ee.size() / 4; S.n_vg = (int)w.vg.size() / 2;
    S.t_total = now_s() - t_start;
    if (st) *st = S;
    return status;
}

static World* W(void* p) { return reinterpret_cast<World*>(p); }

} // namespace

// ===========================================================================
// C API
// ===========================================================================
extern "C" {

int oracle_set_threads(int n) {
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
    return omp_get_max_threads();
#else
    (void)n; return 1;
#endif
}

int oracle_world_create(const adb_scene* s, const adb_params* p, void** out) {
    if (!s || !p || !out || s->n_bodies <= 0) return ADB_E_ARG;
    World* w = new World();
    w->P = *p;
    w->B = s->n_bodies;
    w->voff.assign(s->vert_offset, s->vert_offset + w->B + 1);
    w->foff.assign(s->face_offset, s->face_offset + w->B + 1);
    w->eoff.assign(s->edge_offset, s->edge_offset + w->B + 1);
    w->V = w->voff[w->B]; w->F = w->foff[w->B]; w->E = w->eoff[w->B];
    w->xbar.assign(s->rest_verts, s->rest_verts + 3 * w->V);
    w->face.resize(3 * w->F); w->edge.resize(2 * w->E);
    w->vbody.resize(w->V); w->fbody.resize(w->F); w->ebody.resize(w->E);
    for (int b = 0; b < w->B; b++) {
        for (int v = w->voff[b]; v < w->voff[b + 1]; v++) w->vbody[v] = b;
        for (int f = w->foff[b]; f < w->foff[b + 1]; f++) { w->fbody[f] = b; for (int k = 0; k < 3; k++) w->face[3 * f + k] = w->voff[b] + s->faces[3 * f + k]; }
        for (int e = w->eoff[b]; e < w->eoff[b + 1]; e++) { w->ebody[e] = b; for (int k = 0; k < 2; k++) w->edge[2 * e + k] = w->voff[b] + s->edges[2 * e + k]; }
    }
    w->mass.assign(s->mass, s->mass + w->B);
    w->Ic.assign(s->Ic, s->Ic + w->B);
    w->q0.assign(s->q0, s->q0 + 12 * w->B);
    w->q = w->q0;
    w->dqdt.assign(s->dqdt, s->dqdt + 12 * w->B);
    *out = w;
    return ADB_OK;
}
void oracle_world_destroy(void* w) { delete W(w); }
int oracle_step(void* w, adb_step_stats* st) { return step(*W(w), st); }
int oracle_get_state(void* p, double* q, double* qdot) {
    World& w = *W(p);
    if (q) std::memcpy(q, w.q0.data(), sizeof(double) * 12 * w.B);
    if (qdot) std::memcpy(qdot, w.dqdt.data(), sizeof(double) * 12 * w.B);
    return ADB_OK;
}
int oracle_set_state(void* p, const double* q, const double* qdot) {
    World& w = *W(p);
    if (q) { std::memcpy(w.q0.data(), q, sizeof(double) * 12 * w.B); w.q = w.q0; }
    if (qdot) std::memcpy(w.dqdt.data(), qdot, sizeof(double) * 12 * w.B);
    return ADB_OK;
}
int oracle_collision_set(void* p, int vtn, const double* dq, int method,
                         int* pt, int pt_cap, int* n_pt, int* ee, int ee_cap, int* n_ee,
                         int* vg, int vg_cap, int* n_vg) {
    World& w = *W(p);
    std::vector<double> qv, X;
    q_plus(w, vtn == 2 ? dq : nullptr, qv);
    positions(w, qv.data(), X);
    std::vector<int> a, b, c;
    collision_set(w, X, method, a, b, c);
    *n_pt = (int)a.size() / 4; *n_ee = (int)b.size() / 4; *n_vg = (int)c.size() / 2;
    int rc = ADB_OK;
    if (*n_pt > pt_cap || *n_ee > ee_cap || *n_vg > vg_cap) rc = ADB_E_CAPACITY;
    if (pt) std::memcpy(pt, a.data(), sizeof(int) * std::min<size_t>(a.size(), 4 * (size_t)pt_cap));
    if (ee) std::memcpy(ee, b.data(), sizeof(int) * std::min<size_t>(b.size(), 4 * (size_t)ee_cap));
    if (vg) std::memcpy(vg, c.data(), sizeof(int) * std::min<size_t>(c.size(), 2 * (size_t)vg_cap));
    return rc;
}
int oracle_step_size_upper_bound(void* p, const double* dq, int method, double* toi) {
    World& w = *W(p);
    std::vector<double> qv, X0, X1;
    positions(w, w.q.data(), X0);
    q_plus(w, dq, qv);
    positions(w, qv.data(), X1);
    int st = ADB_OK;
    int sv = w.stall_pt, sg = w.stall_vg;
    *toi = ccd_toi(w, X0, X1, method, &st);
    w.stall_pt = sv; w.stall_vg = sg;
    return ADB_OK;
}
int oracle_assemble(void* p, double* grad, double* hess) {
    World& w = *W(p);
    std::vector<double> X;
    positions(w, w.q.data(), X);
    collision_set(w, X, 1, w.pt, w.ee, w.vg);
    assemble(w, nullptr);
    int n = 12 * w.B;
    if (grad) std::memcpy(grad, w.grad.data(), sizeof(double) * n);
    if (hess) {
        std::memset(hess, 0, sizeof(double) * (size_t)n * n);
        for (auto& kv : w.blocks)
            for (int c = 0; c < 12; c++) for (int r = 0; r < 12; r++)
                hess[(size_t)(12 * kv.first.second + c) * n + 12 * kv.first.first + r] = kv.second[c * 12 + r];
    }
    return ADB_OK;
}
// Assembly at a Newton iterate q with the step start q0 unchanged: friction is
// lagged from the contacts at q and measured against the step-start
// positions, as ABD::ipc inside a step (model/abd.cpp:68-90,305-404).
int oracle_assemble_at(void* p, const double* qv, double* grad, double* hess) {
    World& w = *W(p);
    std::memcpy(w.q.data(), qv, sizeof(double) * 12 * w.B);
    std::vector<double> X;
    positions(w, w.q.data(), X);
    collision_set(w, X, 1, w.pt, w.ee, w.vg);
    assemble(w, nullptr);
    int n = 12 * w.B;
    if (grad) std::memcpy(grad, w.grad.data(), sizeof(double) * n);
    if (hess) {
        std::memset(hess, 0, sizeof(double) * (size_t)n * n);
        for (auto& kv : w.blocks)
            for (int c = 0; c < 12; c++) for (int r = 0; r < 12; r++)
                hess[(size_t)(12 * kv.first.second + c) * n + 12 * kv.first.first + r] = kv.second[c * 12 + r];
    }
    return ADB_OK;
}
// Lagged friction energy (model/energy.cpp:195-270) at q, with the Tk / lambda
// of the last assembly and the step-start positions of q0.
int oracle_friction_energy(void* p, const double* qv, double* E) {
    World& w = *W(p);
    std::vector<double> X, Xs;
    positions(w, qv, X);
    positions(w, w.q0.data(), Xs);
    *E = E_fric(w, X, Xs);
    return ADB_OK;
}
// ABD::line_search (model/abd.cpp:198-227, model/line_search.cpp:39-136) from
// the iterate q along dq (already scaled by the CCD bound): assembles at q
// (gradient and lagged friction), then halves alpha from 1 until the Armijo
// condition holds.  Returns the accepted alpha, E0, E1 and the trial count.
int oracle_line_search(void* p, const double* qv, const double* dq, double* alpha_out, double* E0_out, double* E1_out,
                       int* trials_out) {
    World& w = *W(p);
    const int n = 12 * w.B;
    oracle_assemble_at(p, qv, nullptr, nullptr);
    std::vector<double> X, Xs, X1, q1(n);
    positions(w, w.q.data(), X);
    positions(w, w.q0.data(), Xs);
    const int set_method = (w.P.iaabb & 1) ? 1 : 2;
    double E0 = E_barrier_inert(w, w.q, X, w.pt, w.ee, w.vg) + E_fric(w, X, Xs);
    double qdg = 0, gn = 0;
    for (int i = 0; i < n; i++) { qdg += dq[i] * w.grad[i]; gn += w.grad[i] * w.grad[i]; }
    gn = std::sqrt(gn);
    double alpha = 1.0, E1 = E0;
    int trials = 0, rc = ADB_OK;
    std::vector<int> npt, nee, nvg;
    while (true) {
        for (int i = 0; i < n; i++) q1[i] = w.q[i] + dq[i] * alpha;
        positions(w, q1.data(), X1);
        collision_set(w, X1, set_method, npt, nee, nvg);
        E1 = E_barrier_inert(w, q1, X1, npt, nee, nvg) + E_fric(w, X1, Xs);
        bool wolfe = E1 <= E0 + w.P.c1 * alpha * qdg;
        alpha /= 2;
        trials++;
        if (wolfe || !(gn > 1e-3)) break;
        if (trials >= w.P.max_ls_iter) { rc = ADB_E_NOT_CONVERGED; break; }
    }
    *alpha_out = alpha * 2; *E0_out = E0; *E1_out = E1; *trials_out = trials;
    return rc;
}
// x = H^{-1} rhs with the exact (unprojected) Hessian at the current state:
// the reference gradient of the differentiable step's adjoint.
int oracle_adjoint_solve(void* p, const double* rhs, double* x) {
    World& w = *W(p);
    const int psd = w.P.psd;
    w.P.psd = 0;
    oracle_assemble(p, nullptr, nullptr);
    w.P.psd = psd;
    return sparse_solve(w.B, w.blocks, rhs, x);
}
int oracle_energy(void* p, const double* qv, double* E) {
    World& w = *W(p);
    std::vector<double> q1(qv, qv + 12 * w.B), X;
    positions(w, q1.data(), X);
    std::vector<int> a, b, c;
    collision_set(w, X, 1, a, b, c);
    *E = E_barrier_inert(w, q1, X, a, b, c);
    return ADB_OK;
}
int oracle_newton_direction(void* p, double* dq) {
    World& w = *W(p);
    oracle_assemble(p, nullptr, nullptr);
    std::vector<double> sol(12 * w.B);
    int rc = sparse_solve(w.B, w.blocks, w.grad.data(), sol.data());
    for (int i = 0; i < 12 * w.B; i++) dq[i] = -sol[i];
    return rc;
}

void oracle_pt_distance(const double x[12], int* type, double* d) {
    V3 p = load3(x), t0 = load3(x + 3), t1 = load3(x + 6), t2 = load3(x + 9);
    *type = pt_type(p, t0, t1, t2);
    *d = pt_dist_typed(p, t0, t1, t2, *type);
}
void oracle_ee_distance(const double x[12], int* type, double* d) {
    V3 a0 = load3(x), a1 = load3(x + 3), b0 = load3(x + 6), b1 = load3(x + 9);
    *type = ee_type(a0, a1, b0, b1);
    *d = ee_dist_typed(a0, a1, b0, b1, *type);
}
void oracle_pt_distance_derivs(const double x[12], int type, double g[12], double H[144]) { ad_out(pt_dist_ad(x, type), g, H); }
void oracle_ee_distance_derivs(const double x[12], int type, double g[12], double H[144]) { ad_out(ee_dist_ad(x, type), g, H); }
void oracle_ee_mollifier(const double x[12], double eps, double* m, double g[12], double H[144]) { mollifier_ad(x, eps, m, g, H); }
void oracle_barrier(double d, double dh, double k, double out[3]) {
    out[0] = barrier(d, dh, k); out[1] = barrier_d1(d, dh, k); out[2] = barrier_d2(d, dh, k);
}
void oracle_sym_eig(int n, const double* A, double* w, double* V) { jacobi_eig(n, A, w, V); }
void oracle_psd_project(int n, const double* A, double* out) { psd_project(n, A, out); }
int oracle_ipc_hess_pt_12x12(const adb_kernel_config* c, const double pt[12], int type, double dist, double H[144], double g[12]) {
    hess_pt(pt, type, dist, c->d_hat, c->kappa, c->psd, H, g);
    return ADB_OK;
}
int oracle_ipc_hess_ee_12x12(const adb_kernel_config* c, const double ee[12], int type, double dist, double H[144], double g[12], double* m) {
    hess_ee(ee, type, dist, c->d_hat, c->kappa, c->eps_x, c->psd, H, g, m);
    return ADB_OK;
}
double oracle_pt_collision_time(const double x[24]) {
    V3 P0[4], P1[4];
    for (int k = 0; k < 4; k++) { P0[k] = load3(x + 3 * k); P1[k] = load3(x + 12 + 3 * k); }
    return pt_toi(P0, P1);
}
double oracle_ee_collision_time(const double x[24]) {
    V3 P0[4], P1[4];
    for (int k = 0; k < 4; k++) { P0[k] = load3(x + 3 * k); P1[k] = load3(x + 12 + 3 * k); }
    return ee_toi(P0, P1);
}
void oracle_det_polynomial(const double a[9], const double b[9], double out[4]) {
    double A[3][3], Bm[3][3];
    for (int r = 0; r < 3; r++) for (int c = 0; c < 3; c++) { A[r][c] = a[c * 3 + r]; Bm[r][c] = b[c * 3 + r]; }
    det_poly(A, Bm, out);
}
int oracle_cubic_roots(const double coef[4], double x0, double x1, double roots[3]) { return cubic_roots(roots, coef, x0, x1); }
void oracle_friction(const double uk[2], double lam, const double* T, int n, double mu, double eps, double max_uk, double* g, double* H) {
    friction(uk, lam, T, n, mu, eps, max_uk, g, H);
}
void oracle_ortho(const double q[12], double kap, double* E, double g[12], double H[144]) { ortho(q, kap, E, g, H); }
void oracle_pt_closest(const double x[12], int type, double lams[3], double Pk[6]) {
    V3 P[4] = {load3(x), load3(x + 3), load3(x + 6), load3(x + 9)};
    pt_closest(P, type, lams, Pk);
}
void oracle_ee_closest(const double x[12], int type, double lams[2], double Pk[6]) {
    V3 P[4] = {load3(x), load3(x + 3), load3(x + 6), load3(x + 9)};
    ee_closest(P, type, lams, Pk);
}
double oracle_f0(double x, double eps) { return f0_sf(x, eps); }
double oracle_f1_over_x(double x, double eps) { return f1_over_x(x, eps); }

} // extern "C"
