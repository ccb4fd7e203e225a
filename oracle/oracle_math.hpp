// oracle_math.hpp - scalar FP64 math of the ADB contact step, restated from the
// reference for the CPU oracle.  TEST INFRASTRUCTURE ONLY (see adb_oracle.h).
//
// Compiled with -ffp-contract=off: every +,-,*,/ is a separately rounded IEEE
// operation, so the GPU kernels (built with the same flag and the same
// operation order for the predicates below) reproduce the culling and CCD
// decisions bit for bit.
#pragma once
#include <cmath>
#include <cstring>
#include <algorithm>

namespace orc {

struct V3 { double x, y, z; };
inline V3 v3(double x, double y, double z) { return V3{x, y, z}; }
inline V3 operator+(V3 a, V3 b) { return V3{a.x + b.x, a.y + b.y, a.z + b.z}; }
inline V3 operator-(V3 a, V3 b) { return V3{a.x - b.x, a.y - b.y, a.z - b.z}; }
inline V3 operator*(double s, V3 a) { return V3{s * a.x, s * a.y, s * a.z}; }
inline double dot(V3 a, V3 b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
inline V3 cross(V3 a, V3 b) { return V3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
inline double sqn(V3 a) { return dot(a, a); }
inline double comp(V3 a, int k) { return k == 0 ? a.x : (k == 1 ? a.y : a.z); }
inline V3 load3(const double* p) { return V3{p[0], p[1], p[2]}; }
inline void store3(double* p, V3 a) { p[0] = a.x; p[1] = a.y; p[2] = a.z; }

// ---------------------------------------------------------------------------
// Second-order forward AD over N inputs (value, gradient, packed upper
// Hessian).  Used by the oracle only, to differentiate the distance and
// mollifier functions independently of the GPU's closed-form derivatives.
// ---------------------------------------------------------------------------
template <int N> struct HD {
    static constexpr int NH = N * (N + 1) / 2;
    double v; double g[N]; double h[NH];
    static int id(int i, int j) { if (i > j) std::swap(i, j); return i * N - i * (i - 1) / 2 + (j - i); }
    static HD cst(double c) { HD r; r.v = c; std::memset(r.g, 0, sizeof r.g); std::memset(r.h, 0, sizeof r.h); return r; }
    static HD var(double c, int k) { HD r = cst(c); r.g[k] = 1.0; return r; }
};
template <int N> HD<N> operator+(const HD<N>& a, const HD<N>& b) {
    HD<N> r; r.v = a.v + b.v;
    for (int i = 0; i < N; i++) r.g[i] = a.g[i] + b.g[i];
    for (int k = 0; k < HD<N>::NH; k++) r.h[k] = a.h[k] + b.h[k];
    return r;
}
template <int N> HD<N> operator-(const HD<N>& a, const HD<N>& b) {
    HD<N> r; r.v = a.v - b.v;
    for (int i = 0; i < N; i++) r.g[i] = a.g[i] - b.g[i];
    for (int k = 0; k < HD<N>::NH; k++) r.h[k] = a.h[k] - b.h[k];
    return r;
}
template <int N> HD<N> operator*(const HD<N>& a, const HD<N>& b) {
    HD<N> r; r.v = a.v * b.v;
    for (int i = 0; i < N; i++) r.g[i] = a.v * b.g[i] + b.v * a.g[i];
    for (int i = 0; i < N; i++)
        for (int j = i; j < N; j++) {
            int k = HD<N>::id(i, j);
            r.h[k] = a.v * b.h[k] + b.v * a.h[k] + a.g[i] * b.g[j] + a.g[j] * b.g[i];
        }
    return r;
}
template <int N> HD<N> operator/(const HD<N>& a, const HD<N>& b) {
    HD<N> r; r.v = a.v / b.v;
    for (int i = 0; i < N; i++) r.g[i] = (a.g[i] - r.v * b.g[i]) / b.v;
    for (int i = 0; i < N; i++)
        for (int j = i; j < N; j++) {
            int k = HD<N>::id(i, j);
            r.h[k] = (a.h[k] - r.v * b.h[k] - r.g[i] * b.g[j] - r.g[j] * b.g[i]) / b.v;
        }
    return r;
}
template <int N> HD<N> operator*(double s, const HD<N>& a) { return HD<N>::cst(s) * a; }

template <class T> struct T3 { T x, y, z; };
template <class T> T3<T> sub(const T3<T>& a, const T3<T>& b) { return T3<T>{a.x - b.x, a.y - b.y, a.z - b.z}; }
template <class T> T tdot(const T3<T>& a, const T3<T>& b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
template <class T> T3<T> tcross(const T3<T>& a, const T3<T>& b) {
    return T3<T>{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}

// ---------------------------------------------------------------------------
// Squared distances (the functions ipc-toolkit v1.2.1 dispatches to from
// point_triangle_distance / edge_edge_distance; called at
// model/geometry.cpp:108-109, model/abd.cpp:353-354,
// model/thread_local_culling.cpp:325-326,375-376).
// ---------------------------------------------------------------------------
template <class T> T d_pp(const T3<T>& a, const T3<T>& b) { T3<T> e = sub(a, b); return tdot(e, e); }
template <class T> T d_pl(const T3<T>& p, const T3<T>& e0, const T3<T>& e1) {
    T3<T> c = tcross(sub(e0, p), sub(e1, p));
    T3<T> e = sub(e1, e0);
    return tdot(c, c) / tdot(e, e);
}
template <class T> T d_pplane(const T3<T>& p, const T3<T>& t0, const T3<T>& t1, const T3<T>& t2) {
    T3<T> n = tcross(sub(t1, t0), sub(t2, t0));
    T s = tdot(sub(p, t0), n);
    return s * s / tdot(n, n);
}
template <class T> T d_ll(const T3<T>& a0, const T3<T>& a1, const T3<T>& b0, const T3<T>& b1) {
    T3<T> n = tcross(sub(a1, a0), sub(b1, b0));
    T s = tdot(sub(b0, a0), n);
    return s * s / tdot(n, n);
}

enum { P_T0 = 0, P_T1, P_T2, P_E0, P_E1, P_E2, P_T };
enum { EA0_EB0 = 0, EA0_EB1, EA1_EB0, EA1_EB1, EA_EB0, EA_EB1, EA0_EB, EA1_EB, EA_EB };

inline T3<double> t3(V3 a) { return T3<double>{a.x, a.y, a.z}; }

// point_triangle_distance_type: project p onto each edge's (tangent, outward
// in-plane normal) frame; an edge wins when the projection is inside its span
// and outside the triangle; then the three vertex regions; else the face.
inline int pt_type(V3 p, V3 t0, V3 t1, V3 t2) {
    V3 n = cross(t1 - t0, t2 - t0);
    V3 E0[3] = {t0, t1, t2}, E1[3] = {t1, t2, t0};
    double px[3];
    for (int k = 0; k < 3; k++) {
        V3 b0 = E1[k] - E0[k];
        V3 b1 = cross(b0, n);
        V3 r = p - E0[k];
        double m00 = dot(b0, b0), m01 = dot(b0, b1), m11 = dot(b1, b1);
        double r0 = dot(b0, r), r1 = dot(b1, r);
        double det = m00 * m11 - m01 * m01;
        double x = (r0 * m11 - m01 * r1) / det;
        double y = (m00 * r1 - m01 * r0) / det;
        px[k] = x;
        if (x > 0.0 && x < 1.0 && y >= 0.0) return P_E0 + k;
    }
    if (px[0] <= 0.0 && px[2] >= 1.0) return P_T0;
    if (px[1] <= 0.0 && px[0] >= 1.0) return P_T1;
    if (px[2] <= 0.0 && px[1] >= 1.0) return P_T2;
    return P_T;
}
inline double pt_dist_typed(V3 p, V3 t0, V3 t1, V3 t2, int type) {
    switch (type) {
    case P_T0: return d_pp(t3(p), t3(t0));
    case P_T1: return d_pp(t3(p), t3(t1));
    case P_T2: return d_pp(t3(p), t3(t2));
    case P_E0: return d_pl(t3(p), t3(t0), t3(t1));
    case P_E1: return d_pl(t3(p), t3(t1), t3(t2));
    case P_E2: return d_pl(t3(p), t3(t2), t3(t0));
    default:   return d_pplane(t3(p), t3(t0), t3(t1), t3(t2));
    }
}

// Exactly or nearly parallel edges: the closest pair is attained at an
// endpoint of one segment, so take the smallest of the four endpoint-to-
// segment distances (ties keep the earliest candidate).
inline int ee_parallel_type(V3 a0, V3 a1, V3 b0, V3 b1) {
    V3 u = a1 - a0, v = b1 - b0;
    double a = dot(u, u), c = dot(v, v);
    int best = -1; double bd = 0.0;
    auto consider = [&](int t, double d) { if (best < 0 || d < bd) { best = t; bd = d; } };
    // a0 against segment b
    { double t = dot(a0 - b0, v) / c;
      if (t <= 0.0) consider(EA0_EB0, d_pp(t3(a0), t3(b0)));
      else if (t >= 1.0) consider(EA0_EB1, d_pp(t3(a0), t3(b1)));
      else consider(EA0_EB, d_pl(t3(a0), t3(b0), t3(b1))); }
    { double t = dot(a1 - b0, v) / c;
      if (t <= 0.0) consider(EA1_EB0, d_pp(t3(a1), t3(b0)));
      else if (t >= 1.0) consider(EA1_EB1, d_pp(t3(a1), t3(b1)));
      else consider(EA1_EB, d_pl(t3(a1), t3(b0), t3(b1))); }
    { double s = dot(b0 - a0, u) / a;
      if (s <= 0.0) consider(EA0_EB0, d_pp(t3(b0), t3(a0)));
      else if (s >= 1.0) consider(EA1_EB0, d_pp(t3(b0), t3(a1)));
      else consider(EA_EB0, d_pl(t3(b0), t3(a0), t3(a1))); }
    { double s = dot(b1 - a0, u) / a;
      if (s <= 0.0) consider(EA0_EB1, d_pp(t3(b1), t3(a0)));
      else if (s >= 1.0) consider(EA1_EB1, d_pp(t3(b1), t3(a1)));
      else consider(EA_EB1, d_pl(t3(b1), t3(a0), t3(a1))); }
    return best;
}

// edge_edge_distance_type: segment-segment closest-point classification
// (sN/tN numerators over D = ac - b^2), with the parallel case above.
inline int ee_type(V3 a0, V3 a1, V3 b0, V3 b1) {
    V3 u = a1 - a0, v = b1 - b0, w = a0 - b0;
    double a = dot(u, u), b = dot(u, v), c = dot(v, v), d = dot(u, w), e = dot(v, w);
    double D = a * c - b * b;
    if (a == 0.0 && c == 0.0) return EA0_EB0;
    if (a == 0.0) return EA0_EB;
    if (c == 0.0) return EA_EB0;
    double tol = 1.0e-20 * std::max(1.0, a * c);
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
inline double ee_dist_typed(V3 a0, V3 a1, V3 b0, V3 b1, int type) {
    switch (type) {
    case EA0_EB0: return d_pp(t3(a0), t3(b0));
    case EA0_EB1: return d_pp(t3(a0), t3(b1));
    case EA1_EB0: return d_pp(t3(a1), t3(b0));
    case EA1_EB1: return d_pp(t3(a1), t3(b1));
    case EA_EB0:  return d_pl(t3(b0), t3(a0), t3(a1));
    case EA_EB1:  return d_pl(t3(b1), t3(a0), t3(a1));
    case EA0_EB:  return d_pl(t3(a0), t3(b0), t3(b1));
    case EA1_EB:  return d_pl(t3(a1), t3(b0), t3(b1));
    default:      return d_ll(t3(a0), t3(a1), t3(b0), t3(b1));
    }
}

// ---------------------------------------------------------------------------
// Barrier (model/barrier.cpp:11-28), d squared.
// ---------------------------------------------------------------------------
inline double barrier(double d, double dh, double k) {
    if (d >= dh) return 0.0;
    return k * -(d - dh) * (d - dh) * std::log(d / dh) / (dh * dh);
}
inline double barrier_d1(double x, double dh, double k) {
    if (x >= dh) return 0.0;
    return -(x - dh) * k * (2 * std::log(x / dh) + (x - dh) / x) / (dh * dh);
}
inline double barrier_d2(double d, double dh, double k) {
    if (d >= dh) return 0.0;
    return -k * (2 * std::log(d / dh) + (d - dh) / d + (d - dh) * (2 / d + dh / d / d)) / (dh * dh);
}

// Smooth friction mollifier (ipc-toolkit f0_SF / f1_SF_over_x, the closed
// forms of model/FrictionUtils.h:5-32 written in x rather than x^2).
inline double f0_sf(double x, double eps) {
    if (std::fabs(x) >= eps) return x;
    return x * x * (-x / 3 + eps) / (eps * eps) + eps / 3;
}
inline double f1_over_x(double x, double eps) {
    if (std::fabs(x) >= eps) return 1 / x;
    return (-x / eps + 2) / eps;
}

// ---------------------------------------------------------------------------
// CCD (model/collision.cpp:83-348, cyCodeBase/cyPolynomial.h:69,117,383-447,
// 587-609,787-880).  Root finding restates Yuksel's cubic solver with the
// host defaults the reference uses (xError = 6e-7, Newton-first FindClosed,
// boundError = false).  cyCodeBase: Copyright (c) 2022, Cem Yuksel; MIT
// license, full notice in THIRD_PARTY_NOTICES.md at the repository root.
// ---------------------------------------------------------------------------
inline double poly3(const double c[4], double x) { double r = c[3]; r = r * x + c[2]; r = r * x + c[1]; r = r * x + c[0]; return r; }
inline double poly2(const double c[3], double x) { double r = c[2]; r = r * x + c[1]; r = r * x + c[0]; return r; }
inline bool diff_sign(double a, double b) { return (a < 0) != (b < 0); }
inline double cy_min(double a, double b) { return a <= b ? a : b; }
inline double cy_max(double a, double b) { return a >= b ? a : b; }
inline double mult_sign(double v, double s) { return v * (s < 0 ? -1.0 : 1.0); }

inline double find_closed(const double c[4], const double dv[3], double x0, double x1, double y0, double xe) {
    const double ep2 = 2 * xe;
    double xr = (x0 + x1) / 2;
    if (x1 - x0 <= ep2) return xr;
    const double xr0 = xr;
    for (int it = 0; it < 16; ++it) {
        double xn = xr - poly3(c, xr) / poly2(dv, xr);
        xn = cy_min(x1, cy_max(x0, xn));
        if (std::fabs(xr - xn) <= xe) return xn;
        xr = xn;
    }
    if (!std::isfinite(xr)) xr = xr0;
    double yr = poly3(c, xr);
    double xb0 = x0, xb1 = x1;
    while (true) {
        const bool side = diff_sign(y0, yr);
        if (side) xb1 = xr; else xb0 = xr;
        const double dy = poly2(dv, xr);
        const double dx = yr / dy;
        const double xn = xr - dx;
        if (xn > xb0 && xn < xb1) {
            const double step = std::fabs(xr - xn);
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
inline int quad_roots(double roots[2], const double q[3], double x0, double x1) {
    const double c = q[0], b = q[1], a = q[2];
    const double delta = b * b - 4 * a * c;
    if (delta > 0) {
        const double d = std::sqrt(delta);
        const double qq = -0.5 * (b + mult_sign(d, b));
        const double rv0 = qq / a, rv1 = c / qq;
        const double r0 = cy_min(rv0, rv1), r1 = cy_max(rv0, rv1);
        const int r0i = (r0 >= x0) & (r0 <= x1);
        const int r1i = (r1 >= x0) & (r1 <= x1);
        roots[0] = r0;
        roots[r0i] = r1;
        return r0i + r1i;
    } else if (delta < 0) return 0;
    const double r0 = -0.5 * b / a;
    roots[0] = r0;
    return (r0 >= x0) & (r0 <= x1);
}
inline void deflate3(double out[3], const double c[4], double root) {
    out[2] = c[3];
    out[1] = c[2] + root * out[2];
    out[0] = c[1] + root * out[1];
}
inline int cubic_roots(double roots[3], const double c[4], double x0, double x1, double xe = 6e-7) {
    const double y0 = poly3(c, x0), y1 = poly3(c, x1);
    const double a = c[3] * 3, b2 = c[2], cc = c[1];
    const double dv[3] = {cc, 2 * b2, a};
    const double delta4 = b2 * b2 - a * cc;
    if (delta4 > 0) {
        const double d2 = std::sqrt(delta4);
        const double q = -(b2 + mult_sign(d2, b2));
        const double rv0 = q / a, rv1 = cc / q;
        const double xa = cy_min(rv0, rv1), xb = cy_max(rv0, rv1);
        if (diff_sign(y0, y1)) {
            if (xa >= x1 || xb <= x0 || (xa <= x0 && xb >= x1)) {
                roots[0] = find_closed(c, dv, x0, x1, y0, xe);
                return 1;
            }
        } else {
            if ((xa >= x1 || xb <= x0) || (xa <= x0 && xb >= x1)) return 0;
        }
        if (xa > x0) {
            const double ya = poly3(c, xa);
            if (diff_sign(y0, ya)) {
                roots[0] = find_closed(c, dv, x0, xa, y0, xe);
                if (diff_sign(ya, y1) || (xb < x1 && diff_sign(ya, poly3(c, xb)))) {
                    double dp[3]; deflate3(dp, c, roots[0]);
                    return quad_roots(roots + 1, dp, xa, x1) + 1;
                }
                return 1;
            }
            if (xb < x1) {
                const double yb = poly3(c, xb);
                if (diff_sign(ya, yb)) {
                    roots[0] = find_closed(c, dv, xa, xb, ya, xe);
                    if (diff_sign(yb, y1)) {
                        double dp[3]; deflate3(dp, c, roots[0]);
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
                    double dp[3]; deflate3(dp, c, roots[0]);
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

// det(a t + b) as a cubic in t (model/collision.cpp:107-178): a, b are 3x3
// matrices accessed as m[r][c]; the six signed diagonal products of the rule
// of Sarrus are expanded term by term, positive and negative sums kept apart.
inline void binom3(const double A[3], const double Bv[3], double p[4]) {
    p[0] += Bv[0] * Bv[1] * Bv[2];
    p[1] += A[0] * Bv[1] * Bv[2] + Bv[0] * Bv[1] * A[2] + Bv[0] * A[1] * Bv[2];
    p[2] += A[0] * A[1] * Bv[2] + Bv[0] * A[1] * A[2] + A[0] * Bv[1] * A[2];
    p[3] += A[0] * A[1] * A[2];
}
inline void det_poly(const double a[3][3], const double b[3][3], double out[4]) {
    double pos[4] = {0, 0, 0, 0}, neg[4] = {0, 0, 0, 0};
    { double A[3] = {a[0][0], a[1][1], a[2][2]}, Bv[3] = {b[0][0], b[1][1], b[2][2]}; binom3(A, Bv, pos); }
    { double A[3] = {a[0][1], a[1][2], a[2][0]}, Bv[3] = {b[0][1], b[1][2], b[2][0]}; binom3(A, Bv, pos); }
    { double A[3] = {a[0][2], a[1][0], a[2][1]}, Bv[3] = {b[0][2], b[1][0], b[2][1]}; binom3(A, Bv, pos); }
    { double A[3] = {a[0][0], a[1][2], a[2][1]}, Bv[3] = {b[0][0], b[1][2], b[2][1]}; binom3(A, Bv, neg); }
    { double A[3] = {a[0][1], a[1][0], a[2][2]}, Bv[3] = {b[0][1], b[1][0], b[2][2]}; binom3(A, Bv, neg); }
    { double A[3] = {a[0][2], a[1][1], a[2][0]}, Bv[3] = {b[0][2], b[1][1], b[2][0]}; binom3(A, Bv, neg); }
    for (int i = 0; i < 4; i++) out[i] = pos[i] - neg[i];
}
// Coplanarity cubic of four linearly moving points (collision.cpp:222-257).
inline int coplanar_roots(const V3 P0[4], const V3 P1[4], double roots[3]) {
    static const int cols[4][3] = {{1, 2, 3}, {0, 2, 3}, {0, 1, 3}, {0, 1, 2}};
    double D[4][4];
    for (int m = 0; m < 4; m++) {
        double a[3][3], b[3][3];
        for (int c = 0; c < 3; c++) {
            V3 t0 = P0[cols[m][c]], t1 = P1[cols[m][c]];
            b[0][c] = t0.x; b[1][c] = t0.y; b[2][c] = t0.z;
            a[0][c] = t1.x - b[0][c]; a[1][c] = t1.y - b[1][c]; a[2][c] = t1.z - b[2][c];
        }
        det_poly(a, b, D[m]);
    }
    double t[4];
    for (int i = 0; i < 4; i++) t[i] = D[0][i] - D[1][i] + D[2][i] - D[3][i];
    return cubic_roots(roots, t, 0.0, 1.0);
}
inline V3 lerp(V3 p1, V3 p0, double t) {
    double s = 1 - t;
    return V3{t * p1.x + s * p0.x, t * p1.y + s * p0.y, t * p1.z + s * p0.z};
}
inline double tri_area(V3 a, V3 b, V3 c) { return std::sqrt(sqn(cross(c - a, b - a))); }
inline bool verify_pt(V3 v_, V3 t0, V3 t1, V3 t2) {
    V3 nn = cross(t0 - t1, t0 - t2);
    double l = std::sqrt(dot(nn, nn));
    V3 n{nn.x / l, nn.y / l, nn.z / l};
    double d = dot(n, v_ - t0);
    V3 v = v_ - d * n;
    double a1 = tri_area(t1, t0, t2);
    double a2 = tri_area(t0, t1, v) + tri_area(t1, t2, v) + tri_area(t2, t0, v);
    return a2 <= a1 + 1e-8;
}
inline bool ee_straddle(V3 i0, V3 i1, V3 j0, V3 j1) {
    V3 e = i1 - i0;
    return dot(cross(e, j0 - i0), cross(e, j1 - i0)) < 0.0;
}
inline bool verify_ee(V3 i0, V3 i1, V3 j0, V3 j1) { return ee_straddle(i0, i1, j0, j1) && ee_straddle(j0, j1, i0, i1); }

// pt_collision_time: P0 = {p, t0, t1, t2} at time 0, P1 at time 1.
inline double pt_toi(const V3 P0[4], const V3 P1[4]) {
    double roots[3];
    int found = coplanar_roots(P0, P1, roots);
    bool ok = false; double root = 1.0;
    for (int i = 0; i < found && !ok; i++) {
        root = roots[i];
        ok = verify_pt(lerp(P1[0], P0[0], root), lerp(P1[1], P0[1], root), lerp(P1[2], P0[2], root), lerp(P1[3], P0[3], root));
    }
    return found && ok ? root : 1.0;
}
// ee_collision_time: P0 = {ei0, ei1, ej0, ej1} at time 0, P1 at time 1.
inline double ee_toi(const V3 P0[4], const V3 P1[4]) {
    double roots[3];
    int found = coplanar_roots(P0, P1, roots);
    bool ok = false; double root = 1.0;
    for (int i = 0; i < found && !ok; i++) {
        root = roots[i];
        ok = verify_ee(lerp(P1[0], P0[0], root), lerp(P1[1], P0[1], root), lerp(P1[2], P0[2], root), lerp(P1[3], P0[3], root));
    }
    return found && ok ? root : 1.0;
}

} // namespace orc
