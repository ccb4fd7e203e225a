// dmath.h - device math for the ADB contact step on gfx950 (FP64).
//
// Everything here is __device__ code compiled with -ffp-contract=off.  The
// predicate paths (distance types, squared distances, CCD polynomial, root
// finding, root verification) use a fixed operation order so that contact
// sets and CCD step bounds are reproducible bit for bit by any FP64
// implementation that follows the same order.
//
// Reference semantics: distance types and distances follow ipc-toolkit
// v1.2.1 as called at model/geometry.cpp:101-112 and model/abd.cpp:353-354;
// the barrier follows model/barrier.cpp:11-28; CCD follows
// model/collision.cpp:83-348 with the cubic solver of
// cyCodeBase/cyPolynomial.h:383-447,587-609,787-880 (host defaults).
// The cubic root finder below restates cyCodeBase's CubicRoots / FindClosed
// (Copyright (c) 2022, Cem Yuksel; MIT license, full notice in
// THIRD_PARTY_NOTICES.md at the repository root).
#pragma once
#include <hip/hip_runtime.h>

namespace adb {

struct d3 { double x, y, z; };
__device__ __forceinline__ d3 mk(double x, double y, double z) { d3 r; r.x = x; r.y = y; r.z = z; return r; }
__device__ __forceinline__ d3 operator+(d3 a, d3 b) { return mk(a.x + b.x, a.y + b.y, a.z + b.z); }
__device__ __forceinline__ d3 operator-(d3 a, d3 b) { return mk(a.x - b.x, a.y - b.y, a.z - b.z); }
__device__ __forceinline__ d3 operator*(double s, d3 a) { return mk(s * a.x, s * a.y, s * a.z); }
__device__ __forceinline__ double dot(d3 a, d3 b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
__device__ __forceinline__ d3 cross(d3 a, d3 b) { return mk(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x); }
__device__ __forceinline__ double sqn(d3 a) { return dot(a, a); }
__device__ __forceinline__ double cmp(d3 a, int k) { return k == 0 ? a.x : (k == 1 ? a.y : a.z); }
__device__ __forceinline__ d3 ld3(const double* p) { return mk(p[0], p[1], p[2]); }

enum { P_T0 = 0, P_T1, P_T2, P_E0, P_E1, P_E2, P_T };
enum { EA0_EB0 = 0, EA0_EB1, EA1_EB0, EA1_EB1, EA_EB0, EA_EB1, EA0_EB, EA1_EB, EA_EB };

// ---------------------------------------------------------------- distances
__device__ __forceinline__ double dist_pp(d3 a, d3 b) { d3 e = a - b; return dot(e, e); }
__device__ __forceinline__ double dist_pl(d3 p, d3 e0, d3 e1) {
    d3 c = cross(e0 - p, e1 - p);
    d3 e = e1 - e0;
    return dot(c, c) / dot(e, e);
}
__device__ __forceinline__ double dist_pplane(d3 p, d3 t0, d3 t1, d3 t2) {
    d3 n = cross(t1 - t0, t2 - t0);
    double s = dot(p - t0, n);
    return s * s / dot(n, n);
}
__device__ __forceinline__ double dist_ll(d3 a0, d3 a1, d3 b0, d3 b1) {
    d3 n = cross(a1 - a0, b1 - b0);
    double s = dot(b0 - a0, n);
    return s * s / dot(n, n);
}

// Closest feature of a triangle: an edge wins when the point projects inside
// the edge's span and on the outer side of it (solve of the 2x2 system in the
// edge's (tangent, in-plane outward normal) frame); then vertex regions.
__device__ __forceinline__ int pt_type(d3 p, d3 t0, d3 t1, d3 t2) {
    d3 n = cross(t1 - t0, t2 - t0);
    double px0 = 0.0, px1 = 0.0, px2 = 0.0;
#pragma unroll
    for (int k = 0; k < 3; k++) {
        d3 e0 = k == 0 ? t0 : (k == 1 ? t1 : t2);
        d3 e1 = k == 0 ? t1 : (k == 1 ? t2 : t0);
        d3 b0 = e1 - e0;
        d3 b1 = cross(b0, n);
        d3 r = p - e0;
        double m00 = dot(b0, b0), m01 = dot(b0, b1), m11 = dot(b1, b1);
        double r0 = dot(b0, r), r1 = dot(b1, r);
        double det = m00 * m11 - m01 * m01;
        double x = (r0 * m11 - m01 * r1) / det;
        double y = (m00 * r1 - m01 * r0) / det;
        if (k == 0) px0 = x; else if (k == 1) px1 = x; else px2 = x;
        if (x > 0.0 && x < 1.0 && y >= 0.0) return P_E0 + k;
    }
    if (px0 <= 0.0 && px2 >= 1.0) return P_T0;
    if (px1 <= 0.0 && px0 >= 1.0) return P_T1;
    if (px2 <= 0.0 && px1 >= 1.0) return P_T2;
    return P_T;
}
__device__ __forceinline__ double pt_dist(d3 p, d3 t0, d3 t1, d3 t2, int type) {
    switch (type) {
    case P_T0: return dist_pp(p, t0);
    case P_T1: return dist_pp(p, t1);
    case P_T2: return dist_pp(p, t2);
    case P_E0: return dist_pl(p, t0, t1);
    case P_E1: return dist_pl(p, t1, t2);
    case P_E2: return dist_pl(p, t2, t0);
    default:   return dist_pplane(p, t0, t1, t2);
    }
}
// parallel edges: the smallest endpoint-to-segment distance (earliest wins ties)
__device__ __forceinline__ int ee_parallel_type(d3 a0, d3 a1, d3 b0, d3 b1) {
    d3 u = a1 - a0, v = b1 - b0;
    double a = dot(u, u), c = dot(v, v);
    int best; double bd, d; int t;
    { double s = dot(a0 - b0, v) / c;
      if (s <= 0.0) { t = EA0_EB0; d = dist_pp(a0, b0); } else if (s >= 1.0) { t = EA0_EB1; d = dist_pp(a0, b1); }
      else { t = EA0_EB; d = dist_pl(a0, b0, b1); }
      best = t; bd = d; }
    { double s = dot(a1 - b0, v) / c;
      if (s <= 0.0) { t = EA1_EB0; d = dist_pp(a1, b0); } else if (s >= 1.0) { t = EA1_EB1; d = dist_pp(a1, b1); }
      else { t = EA1_EB; d = dist_pl(a1, b0, b1); }
      if (d < bd) { best = t; bd = d; } }
    { double s = dot(b0 - a0, u) / a;
      if (s <= 0.0) { t = EA0_EB0; d = dist_pp(b0, a0); } else if (s >= 1.0) { t = EA1_EB0; d = dist_pp(b0, a1); }
      else { t = EA_EB0; d = dist_pl(b0, a0, a1); }
      if (d < bd) { best = t; bd = d; } }
    { double s = dot(b1 - a0, u) / a;
      if (s <= 0.0) { t = EA0_EB1; d = dist_pp(b1, a0); } else if (s >= 1.0) { t = EA1_EB1; d = dist_pp(b1, a1); }
      else { t = EA_EB1; d = dist_pl(b1, a0, a1); }
      if (d < bd) { best = t; bd = d; } }
    return best;
}
__device__ __forceinline__ int ee_type(d3 a0, d3 a1, d3 b0, d3 b1) {
    d3 u = a1 - a0, v = b1 - b0, w = a0 - b0;
    double a = dot(u, u), b = dot(u, v), c = dot(v, v), d = dot(u, w), e = dot(v, w);
    double D = a * c - b * b;
    if (a == 0.0 && c == 0.0) return EA0_EB0;
    if (a == 0.0) return EA0_EB;
    if (c == 0.0) return EA_EB0;
    double ac = a * c;
    double tol = 1.0e-20 * (1.0 >= ac ? 1.0 : ac);
    if (sqn(cross(u, v)) < tol) return ee_parallel_type(a0, a1, b0, b1);
    int def = EA_EB;
    double sN = b * e - c * d, tN, tD;
    if (sN <= 0.0) { tN = e; tD = c; def = EA0_EB; }
    else if (sN >= D) { tN = e + b; tD = c; def = EA1_EB; }
    else { tN = a * e - b * d; tD = D; }
    if (tN <= 0.0) {
        if (-d <= 0.0) return EA0_EB0;
        else if (-d >= a) return EA1_EB0;
        else return EA_EB0;
    } else if (tN >= tD) {
        if (-d + b <= 0.0) return EA0_EB1;
        else if (-d + b >= a) return EA1_EB1;
        else return EA_EB1;
    }
    return def;
}
__device__ __forceinline__ double ee_dist(d3 a0, d3 a1, d3 b0, d3 b1, int type) {
    switch (type) {
    case EA0_EB0: return dist_pp(a0, b0);
    case EA0_EB1: return dist_pp(a0, b1);
    case EA1_EB0: return dist_pp(a1, b0);
    case EA1_EB1: return dist_pp(a1, b1);
    case EA_EB0:  return dist_pl(b0, a0, a1);
    case EA_EB1:  return dist_pl(b1, a0, a1);
    case EA0_EB:  return dist_pl(a0, b0, b1);
    case EA1_EB:  return dist_pl(a1, b0, b1);
    default:      return dist_ll(a0, a1, b0, b1);
    }
}
__device__ __forceinline__ double ee_mollifier_value(d3 a0, d3 a1, d3 b0, d3 b1, double eps) {
    double x = sqn(cross(a1 - a0, b1 - b0));
    if (x < eps) return (-x / eps + 2) * x / eps;
    return 1.0;
}

// ---------------------------------------------------------------- barrier
__device__ __forceinline__ double barrier(double d, double dh, double k) {
    if (d >= dh) return 0.0;
    return k * -(d - dh) * (d - dh) * log(d / dh) / (dh * dh);
}
__device__ __forceinline__ double barrier_d1(double x, double dh, double k) {
    if (x >= dh) return 0.0;
    return -(x - dh) * k * (2 * log(x / dh) + (x - dh) / x) / (dh * dh);
}
__device__ __forceinline__ double barrier_d2(double d, double dh, double k) {
    if (d >= dh) return 0.0;
    return -k * (2 * log(d / dh) + (d - dh) / d + (d - dh) * (2 / d + dh / d / d)) / (dh * dh);
}
__device__ __forceinline__ double f0_sf(double x, double eps) {
    if (fabs(x) >= eps) return x;
    return x * x * (-x / 3 + eps) / (eps * eps) + eps / 3;
}
__device__ __forceinline__ double f1_over_x(double x, double eps) {
    if (fabs(x) >= eps) return 1 / x;
    return (-x / eps + 2) / eps;
}

// ---------------------------------------------------------------- reduced derivatives
// Every squared distance is a function of up to three difference vectors
// r_m = x[plus_m] - x[minus_m] of the four contact points:
//   PP:     d = |r0|^2
//   PL:     d = |r0 x r1|^2 / |r1 - r0|^2            (r0 = e0 - p, r1 = e1 - p)
//   PLANE:  d = (r0 . (r1 x r2))^2 / |r1 x r2|^2      (point-plane and line-line)
// and the EE mollifier argument is |r1 x r2|^2 with r1 = a1 - a0, r2 = b1 - b0.
// grad[m][k] = d d / d r_m[k];  hess[m][n][k][l] = d^2 d / d r_m[k] d r_n[l].
struct RedMap { int nr; int plus[3]; int minus[3]; int kind; };   // kind: 0 PP, 1 PL, 2 PLANE
enum { K_PP = 0, K_PL = 1, K_PLANE = 2 };

__device__ __forceinline__ RedMap pt_map(int type) {
    RedMap m;
    switch (type) {
    case P_T0: m.kind = K_PP; m.nr = 1; m.plus[0] = 0; m.minus[0] = 1; break;
    case P_T1: m.kind = K_PP; m.nr = 1; m.plus[0] = 0; m.minus[0] = 2; break;
    case P_T2: m.kind = K_PP; m.nr = 1; m.plus[0] = 0; m.minus[0] = 3; break;
    case P_E0: m.kind = K_PL; m.nr = 2; m.plus[0] = 1; m.minus[0] = 0; m.plus[1] = 2; m.minus[1] = 0; break;
    case P_E1: m.kind = K_PL; m.nr = 2; m.plus[0] = 2; m.minus[0] = 0; m.plus[1] = 3; m.minus[1] = 0; break;
    case P_E2: m.kind = K_PL; m.nr = 2; m.plus[0] = 3; m.minus[0] = 0; m.plus[1] = 1; m.minus[1] = 0; break;
    default:   m.kind = K_PLANE; m.nr = 3; m.plus[0] = 0; m.minus[0] = 1; m.plus[1] = 2; m.minus[1] = 1; m.plus[2] = 3; m.minus[2] = 1; break;
    }
    return m;
}
__device__ __forceinline__ RedMap ee_map(int type) {
    RedMap m;
    switch (type) {
    case EA0_EB0: m.kind = K_PP; m.nr = 1; m.plus[0] = 0; m.minus[0] = 2; break;
    case EA0_EB1: m.kind = K_PP; m.nr = 1; m.plus[0] = 0; m.minus[0] = 3; break;
    case EA1_EB0: m.kind = K_PP; m.nr = 1; m.plus[0] = 1; m.minus[0] = 2; break;
    case EA1_EB1: m.kind = K_PP; m.nr = 1; m.plus[0] = 1; m.minus[0] = 3; break;
    case EA_EB0:  m.kind = K_PL; m.nr = 2; m.plus[0] = 0; m.minus[0] = 2; m.plus[1] = 1; m.minus[1] = 2; break;
    case EA_EB1:  m.kind = K_PL; m.nr = 2; m.plus[0] = 0; m.minus[0] = 3; m.plus[1] = 1; m.minus[1] = 3; break;
    case EA0_EB:  m.kind = K_PL; m.nr = 2; m.plus[0] = 2; m.minus[0] = 0; m.plus[1] = 3; m.minus[1] = 0; break;
    case EA1_EB:  m.kind = K_PL; m.nr = 2; m.plus[0] = 2; m.minus[0] = 1; m.plus[1] = 3; m.minus[1] = 1; break;
    default:      m.kind = K_PLANE; m.nr = 3; m.plus[0] = 2; m.minus[0] = 0; m.plus[1] = 1; m.minus[1] = 0; m.plus[2] = 3; m.minus[2] = 2; break;
    }
    return m;
}
// sign with which point k enters reduced vector m
__device__ __forceinline__ double red_sign(const RedMap& m, int r, int k) {
    return (m.plus[r] == k ? 1.0 : 0.0) - (m.minus[r] == k ? 1.0 : 0.0);
}

// Cross-norm c(u, w) = |u x w|^2 = (u.u)(w.w) - (u.w)^2: gradient and the
// (a, b) = (u|w, u|w) Hessian entry (row component i of block a, column
// component j of block b).
struct CrossNorm {
    d3 u, w; double uu, ww, uw;
    __device__ __forceinline__ void init(d3 u_, d3 w_) { u = u_; w = w_; uu = dot(u, u); ww = dot(w, w); uw = dot(u, w); }
    __device__ __forceinline__ double value() const { d3 c = cross(u, w); return dot(c, c); }
    __device__ __forceinline__ double grad(int a, int i) const {
        return a == 0 ? 2 * ww * cmp(u, i) - 2 * uw * cmp(w, i) : 2 * uu * cmp(w, i) - 2 * uw * cmp(u, i);
    }
    __device__ __forceinline__ double hess(int a, int i, int b, int j) const {
        double dij = i == j ? 1.0 : 0.0;
        if (a == 0 && b == 0) return 2 * ww * dij - 2 * cmp(w, i) * cmp(w, j);
        if (a == 1 && b == 1) return 2 * uu * dij - 2 * cmp(u, i) * cmp(u, j);
        if (a == 0) return 4 * cmp(u, i) * cmp(w, j) - 2 * cmp(w, i) * cmp(u, j) - 2 * uw * dij;   // (u, w)
        return 4 * cmp(w, i) * cmp(u, j) - 2 * cmp(u, i) * cmp(w, j) - 2 * uw * dij;               // (w, u)
    }
};

// Levi-Civita contraction: (d^2 s / d r_a[i] d r_b[j]) for s = r0 . (r1 x r2)
__device__ __forceinline__ double eps3(int i, int j, int k) {
    return (double)((i - j) * (j - k) * (k - i)) * 0.5;
}

// Reduced gradient / Hessian of a squared distance of kind PP, PL or PLANE.
// The Hessian is produced one column block at a time: hcol(n, l, m, k)
// = d^2 d / d r_m[k] d r_n[l].
struct RedDist {
    int kind;
    d3 r[3];
    double d;           // value
    d3 g[3];            // gradient
    // PL: quotient N / D with N = |r0 x r1|^2, D = |r1 - r0|^2
    // PLANE: quotient s^2 / G with s = r0.(r1 x r2), G = |r1 x r2|^2
    CrossNorm cn;
    double s, G, N, D;
    d3 gs[3];
    // Every field is assigned exactly once, after the branches, so that the
    // compiler cannot merge per-branch stores into an address phi (which
    // would force the whole struct into scratch memory).
    __device__ __forceinline__ void init(int kind_, d3 r0, d3 r1, d3 r2) {
        const d3 z = mk(0, 0, 0);
        double dv = 0.0, Nv = 0.0, Dv = 1.0, Gv = 1.0, sv = 0.0;
        d3 g0 = z, g1 = z, g2 = z, s0 = z, s1 = z, s2 = z;
        CrossNorm c;
        if (kind_ == K_PLANE) c.init(r1, r2); else c.init(r0, r1);
        if (kind_ == K_PP) {
            dv = dot(r0, r0);
            g0 = mk(2 * r0.x, 2 * r0.y, 2 * r0.z);
        } else if (kind_ == K_PL) {
            Nv = c.value();
            d3 e = r1 - r0;
            Dv = dot(e, e);
            dv = Nv / Dv;
            double t0[3], t1[3];
#pragma unroll
            for (int k = 0; k < 3; k++) {
                double gD0 = -2 * cmp(e, k), gD1 = 2 * cmp(e, k);
                t0[k] = (c.grad(0, k) - dv * gD0) / Dv;
                t1[k] = (c.grad(1, k) - dv * gD1) / Dv;
            }
            g0 = mk(t0[0], t0[1], t0[2]); g1 = mk(t1[0], t1[1], t1[2]);
        } else {
            Gv = c.value();
            s0 = cross(r1, r2); s1 = cross(r2, r0); s2 = cross(r0, r1);
            sv = dot(r0, s0);
            Nv = sv * sv;
            dv = Nv / Gv;
            double t0[3], t1[3], t2[3];
#pragma unroll
            for (int k = 0; k < 3; k++) {
                t0[k] = (2 * sv * cmp(s0, k)) / Gv;
                t1[k] = (2 * sv * cmp(s1, k) - dv * c.grad(0, k)) / Gv;
                t2[k] = (2 * sv * cmp(s2, k) - dv * c.grad(1, k)) / Gv;
            }
            g0 = mk(t0[0], t0[1], t0[2]); g1 = mk(t1[0], t1[1], t1[2]); g2 = mk(t2[0], t2[1], t2[2]);
        }
        kind = kind_; r[0] = r0; r[1] = r1; r[2] = r2;
        d = dv; N = Nv; D = Dv; G = Gv; s = sv; cn = c;
        g[0] = g0; g[1] = g1; g[2] = g2;
        gs[0] = s0; gs[1] = s1; gs[2] = s2;
    }
    // d^2 d / d r_m[k] d r_n[l]
    __device__ __forceinline__ double hess(int m, int k, int n, int l) const {
        if (kind == K_PP) return k == l ? 2.0 : 0.0;
        if (kind == K_PL) {
            d3 e = r[1] - r[0];
            double sm = m == 0 ? -1.0 : 1.0, sn = n == 0 ? -1.0 : 1.0;
            double gDm = 2 * sm * cmp(e, k), gDn = 2 * sn * cmp(e, l);
            double hD = (k == l ? 2.0 * sm * sn : 0.0);
            double hN = cn.hess(m, k, n, l);
            return (hN - d * hD - cmp(g[m], k) * gDn - gDm * cmp(g[n], l)) / D;
        }
        // PLANE
        double hs = 0.0;
        if (m != n) {
            // s = r0.(r1 x r2): d2s/dr_m[k] dr_n[l] = eps(k, l, j) r_q[j] with sign by
            // the cyclic order of (m, n, q)
            int q = 3 - m - n;
            double sg = ((m == 0 && n == 1) || (m == 1 && n == 2) || (m == 2 && n == 0)) ? 1.0 : -1.0;
            double acc = 0.0;
            for (int j = 0; j < 3; j++) acc += eps3(k, l, j) * cmp(r[q], j);
            hs = sg * acc;
        }
        double hN = 2 * cmp(gs[m], k) * cmp(gs[n], l) + 2 * s * hs;
        double gGm = m == 0 ? 0.0 : cn.grad(m - 1, k);
        double gGn = n == 0 ? 0.0 : cn.grad(n - 1, l);
        double hG = (m == 0 || n == 0) ? 0.0 : cn.hess(m - 1, k, n - 1, l);
        return (hN - d * hG - cmp(g[m], k) * gGn - gGm * cmp(g[n], l)) / G;
    }
};

// ---------------------------------------------------------------- CCD
__device__ __forceinline__ double poly3(const double c[4], double x) { double r = c[3]; r = r * x + c[2]; r = r * x + c[1]; r = r * x + c[0]; return r; }
__device__ __forceinline__ double poly2(const double c[3], double x) { double r = c[2]; r = r * x + c[1]; r = r * x + c[0]; return r; }
__device__ __forceinline__ bool diff_sign(double a, double b) { return (a < 0) != (b < 0); }
__device__ __forceinline__ double cmin(double a, double b) { return a <= b ? a : b; }
__device__ __forceinline__ double cmax(double a, double b) { return a >= b ? a : b; }
__device__ __forceinline__ double msign(double v, double s) { return v * (s < 0 ? -1.0 : 1.0); }

// single root in [x0, x1] bracketed by a sign change: 16 clamped Newton steps,
// then Newton safeguarded by bisection (tolerance xe on x)
__device__ double find_closed(const double c[4], const double dv[3], double x0, double x1, double y0, double xe) {
    const double ep2 = 2 * xe;
    double xr = (x0 + x1) / 2;
    if (x1 - x0 <= ep2) return xr;
    const double xr0 = xr;
    for (int it = 0; it < 16; ++it) {
        double xn = xr - poly3(c, xr) / poly2(dv, xr);
        xn = cmin(x1, cmax(x0, xn));
        if (fabs(xr - xn) <= xe) return xn;
        xr = xn;
    }
    if (!isfinite(xr)) xr = xr0;
    double yr = poly3(c, xr);
    double xb0 = x0, xb1 = x1;
    for (int guard = 0; guard < 4096; ++guard) {
        const bool side = diff_sign(y0, yr);
        if (side) xb1 = xr; else xb0 = xr;
        const double dy = poly2(dv, xr);
        const double dx = yr / dy;
        const double xn = xr - dx;
        if (xn > xb0 && xn < xb1) {
            const double step = fabs(xr - xn);
            xr = xn;
            if (step > xe) yr = poly3(c, xr);
            else break;
        } else {
            xr = (xb0 + xb1) / 2;
            if (xr == xb0 || xr == xb1 || xb1 - xb0 <= ep2) break;
            yr = poly3(c, xr);
        }
    }
    return xr;
}
__device__ __forceinline__ int quad_roots(double* roots, const double q[3], double x0, double x1) {
    const double c = q[0], b = q[1], a = q[2];
    const double delta = b * b - 4 * a * c;
    if (delta > 0) {
        const double d = sqrt(delta);
        const double qq = -0.5 * (b + msign(d, b));
        const double rv0 = qq / a, rv1 = c / qq;
        const double r0 = cmin(rv0, rv1), r1 = cmax(rv0, rv1);
        const int r0i = (r0 >= x0) & (r0 <= x1);
        const int r1i = (r1 >= x0) & (r1 <= x1);
        roots[0] = r0;
        if (r0i) roots[1] = r1; else roots[0] = r1;
        return r0i + r1i;
    } else if (delta < 0) return 0;
    const double r0 = -0.5 * b / a;
    roots[0] = r0;
    return (r0 >= x0) & (r0 <= x1);
}
__device__ __forceinline__ int cubic_roots(double roots[3], const double c[4], double x0, double x1, double xe) {
    const double y0 = poly3(c, x0), y1 = poly3(c, x1);
    const double a = c[3] * 3, b2 = c[2], cc = c[1];
    const double dv[3] = {cc, 2 * b2, a};
    const double delta4 = b2 * b2 - a * cc;
    if (delta4 > 0) {
        const double d2 = sqrt(delta4);
        const double q = -(b2 + msign(d2, b2));
        const double rv0 = q / a, rv1 = cc / q;
        const double xa = cmin(rv0, rv1), xb = cmax(rv0, rv1);
        if (diff_sign(y0, y1)) {
            if (xa >= x1 || xb <= x0 || (xa <= x0 && xb >= x1)) { roots[0] = find_closed(c, dv, x0, x1, y0, xe); return 1; }
        } else {
            if ((xa >= x1 || xb <= x0) || (xa <= x0 && xb >= x1)) return 0;
        }
        if (xa > x0) {
            const double ya = poly3(c, xa);
            if (diff_sign(y0, ya)) {
                roots[0] = find_closed(c, dv, x0, xa, y0, xe);
                if (diff_sign(ya, y1) || (xb < x1 && diff_sign(ya, poly3(c, xb)))) {
                    double dp[3]; dp[2] = c[3]; dp[1] = c[2] + roots[0] * dp[2]; dp[0] = c[1] + roots[0] * dp[1];
                    return quad_roots(roots + 1, dp, xa, x1) + 1;
                }
                return 1;
            }
            if (xb < x1) {
                const double yb = poly3(c, xb);
                if (diff_sign(ya, yb)) {
                    roots[0] = find_closed(c, dv, xa, xb, ya, xe);
                    if (diff_sign(yb, y1)) {
                        double dp[3]; dp[2] = c[3]; dp[1] = c[2] + roots[0] * dp[2]; dp[0] = c[1] + roots[0] * dp[1];
                        return quad_roots(roots + 1, dp, xb, x1) + 1;
                    }
                    return 1;
                }
                if (diff_sign(yb, y1)) { roots[0] = find_closed(c, dv, xb, x1, yb, xe); return 1; }
            } else {
                if (diff_sign(ya, y1)) { roots[0] = find_closed(c, dv, xa, x1, ya, xe); return 1; }
            }
        } else {
            const double yb = poly3(c, xb);
            if (diff_sign(y0, yb)) {
                roots[0] = find_closed(c, dv, x0, xb, y0, xe);
                if (diff_sign(yb, y1)) {
                    double dp[3]; dp[2] = c[3]; dp[1] = c[2] + roots[0] * dp[2]; dp[0] = c[1] + roots[0] * dp[1];
                    return quad_roots(roots + 1, dp, xb, x1) + 1;
                }
                return 1;
            }
            if (diff_sign(yb, y1)) { roots[0] = find_closed(c, dv, xb, x1, yb, xe); return 1; }
        }
        return 0;
    }
    if (diff_sign(y0, y1)) { roots[0] = find_closed(c, dv, x0, x1, y0, xe); return 1; }
    return 0;
}

// det(A t + B) as a cubic; A, B given as three column points each (column c,
// row r = coordinate r of point c).  Six Sarrus products, positive and
// negative sums kept apart (model/collision.cpp:107-178).
__device__ __forceinline__ void sarrus_term(double A0, double A1, double A2, double B0, double B1, double B2, double p[4]) {
    p[0] += B0 * B1 * B2;
    p[1] += A0 * B1 * B2 + B0 * B1 * A2 + B0 * A1 * B2;
    p[2] += A0 * A1 * B2 + B0 * A1 * A2 + A0 * B1 * A2;
    p[3] += A0 * A1 * A2;
}
__device__ __forceinline__ void det_cubic(d3 a0, d3 a1, d3 a2, d3 b0, d3 b1, d3 b2, double out[4]) {
    // a(r, c): column c is point ac
    double pos[4] = {0, 0, 0, 0}, neg[4] = {0, 0, 0, 0};
    sarrus_term(a0.x, a1.y, a2.z, b0.x, b1.y, b2.z, pos);   // c11 c22 c33
    sarrus_term(a1.x, a2.y, a0.z, b1.x, b2.y, b0.z, pos);   // c12 c23 c31
    sarrus_term(a2.x, a0.y, a1.z, b2.x, b0.y, b1.z, pos);   // c13 c21 c32
    sarrus_term(a0.x, a2.y, a1.z, b0.x, b2.y, b1.z, neg);   // c11 c23 c32
    sarrus_term(a1.x, a0.y, a2.z, b1.x, b0.y, b2.z, neg);   // c12 c21 c33
    sarrus_term(a2.x, a1.y, a0.z, b2.x, b1.y, b0.z, neg);   // c13 c22 c31
    for (int i = 0; i < 4; i++) out[i] = pos[i] - neg[i];
}
// coplanarity cubic of four linearly moving points, roots in [0, 1]
__device__ __forceinline__ int coplanar_roots(const d3 P0[4], const d3 P1[4], double roots[3]) {
    double D[4][4];
#pragma unroll
    for (int m = 0; m < 4; m++) {
        int c0 = m == 0 ? 1 : 0, c1 = m <= 1 ? 2 : 1, c2 = m <= 2 ? 3 : 2;
        d3 b0 = P0[c0], b1 = P0[c1], b2 = P0[c2];
        d3 a0 = P1[c0] - b0, a1 = P1[c1] - b1, a2 = P1[c2] - b2;
        det_cubic(a0, a1, a2, b0, b1, b2, D[m]);
    }
    double t[4];
    for (int i = 0; i < 4; i++) t[i] = D[0][i] - D[1][i] + D[2][i] - D[3][i];
    return cubic_roots(roots, t, 0.0, 1.0, 6e-7);
}
__device__ __forceinline__ d3 lerp(d3 p1, d3 p0, double t) {
    double s = 1 - t;
    return mk(t * p1.x + s * p0.x, t * p1.y + s * p0.y, t * p1.z + s * p0.z);
}
__device__ __forceinline__ double tri_area(d3 a, d3 b, d3 c) { return sqrt(sqn(cross(c - a, b - a))); }
__device__ __forceinline__ bool verify_pt(d3 v_, d3 t0, d3 t1, d3 t2) {
    d3 nn = cross(t0 - t1, t0 - t2);
    double l = sqrt(dot(nn, nn));
    d3 n = mk(nn.x / l, nn.y / l, nn.z / l);
    double d = dot(n, v_ - t0);
    d3 v = v_ - d * n;
    double a1 = tri_area(t1, t0, t2);
    double a2 = tri_area(t0, t1, v) + tri_area(t1, t2, v) + tri_area(t2, t0, v);
    return a2 <= a1 + 1e-8;
}
__device__ __forceinline__ bool straddle(d3 i0, d3 i1, d3 j0, d3 j1) {
    d3 e = i1 - i0;
    return dot(cross(e, j0 - i0), cross(e, j1 - i0)) < 0.0;
}
__device__ __forceinline__ double pt_toi(const d3 P0[4], const d3 P1[4]) {
    double roots[3];
    int found = coplanar_roots(P0, P1, roots);
    bool ok = false; double root = 1.0;
    for (int i = 0; i < found && !ok; i++) {
        root = roots[i];
        ok = verify_pt(lerp(P1[0], P0[0], root), lerp(P1[1], P0[1], root), lerp(P1[2], P0[2], root), lerp(P1[3], P0[3], root));
    }
    return (found && ok) ? root : 1.0;
}
__device__ __forceinline__ double ee_toi(const d3 P0[4], const d3 P1[4]) {
    double roots[3];
    int found = coplanar_roots(P0, P1, roots);
    bool ok = false; double root = 1.0;
    for (int i = 0; i < found && !ok; i++) {
        root = roots[i];
        d3 a0 = lerp(P1[0], P0[0], root), a1 = lerp(P1[1], P0[1], root), b0 = lerp(P1[2], P0[2], root), b1 = lerp(P1[3], P0[3], root);
        ok = straddle(a0, a1, b0, b1) && straddle(b0, b1, a0, a1);
    }
    return (found && ok) ? root : 1.0;
}

} // namespace adb
