// energy.hip - line-search energies and deterministic reductions.
//
// Replaces E_barrier_plus_inert, E_fric and E_ground (model/energy.cpp:19-86,
// 195-270) and the inertia energy E (model/mutils.cpp:197-201).  Every item
// (body, contact) writes its energy to a buffer which is then summed by a
// fixed-shape two-level tree, so results do not depend on scheduling.
#include "dmath.h"
#include "world.h"

namespace adb {

struct EnergyArgs {
    int B, n_pt, n_ee, n_vg;
    double dt, kap_o, d_hat, kappa, eps_x, ground_y, g0, g1, g2;
    int en_pt, en_ee, en_g;
    const double* qv; const double* q0; const double* qdot; const double* mass; const double* Ic;
    const double* X;
    const int4* pt; const int4* ee; const int2* vg;
    const int* voff; const int* foff; const int* eoff; const int* face; const int* edge;
    double* out;
};

// energy of item i < B + n_pt + n_ee + n_vg: body i (inertia and
// orthogonality), then the PT, EE and VG barriers
__device__ __forceinline__ double energy_item(const EnergyArgs& a, int i) {
    double e = 0.0;
    if (i < a.B) {
        if (a.mass[i] >= 0.0) {
            const double* Q = a.qv + 12 * (size_t)i;
            d3 c[4] = {ld3(Q), ld3(Q + 3), ld3(Q + 6), ld3(Q + 9)};
            double eo = 0.0;
            for (int s = 1; s < 4; s++)
                for (int t = 1; t < 4; t++) { double v = dot(c[s], c[t]) - 1.0 * (s == t); eo += v * v; }
            eo = eo * a.kap_o;
            double pp = 0.0, rr = 0.0;
            for (int k = 0; k < 12; k++) {
                double qt = a.q0[12 * (size_t)i + k] + a.dt * a.qdot[12 * (size_t)i + k];
                if (k < 3) qt += a.dt * a.dt * (k == 0 ? a.g0 : (k == 1 ? a.g1 : a.g2));
                double d = Q[k] - qt;
                if (k < 3) pp += d * d; else rr += d * d;
            }
            e = eo * a.dt * a.dt + 0.5 * (pp * a.mass[i] + rr * a.Ic[i]);
        }
    } else if (i < a.B + a.n_pt) {
        if (a.en_pt) {
            int4 c = a.pt[i - a.B];
            int v = a.voff[c.x] + c.y, f = a.foff[c.z] + c.w;
            d3 p = ld3(a.X + 3 * (size_t)v), t0 = ld3(a.X + 3 * (size_t)a.face[3 * f]);
            d3 t1 = ld3(a.X + 3 * (size_t)a.face[3 * f + 1]), t2 = ld3(a.X + 3 * (size_t)a.face[3 * f + 2]);
            double d = pt_dist(p, t0, t1, t2, pt_type(p, t0, t1, t2));
            e = barrier(d, a.d_hat, a.kappa);
        }
    } else if (i < a.B + a.n_pt + a.n_ee) {
        if (a.en_ee) {
            int4 c = a.ee[i - a.B - a.n_pt];
            int ea = a.eoff[c.x] + c.y, eb = a.eoff[c.z] + c.w;
            d3 a0 = ld3(a.X + 3 * (size_t)a.edge[2 * ea]), a1 = ld3(a.X + 3 * (size_t)a.edge[2 * ea + 1]);
            d3 b0 = ld3(a.X + 3 * (size_t)a.edge[2 * eb]), b1 = ld3(a.X + 3 * (size_t)a.edge[2 * eb + 1]);
            double eps = sqn(a0 - a1) * sqn(b0 - b1) * a.eps_x;
            double m = ee_mollifier_value(a0, a1, b0, b1, eps);
            double d = ee_dist(a0, a1, b0, b1, ee_type(a0, a1, b0, b1));
            e = barrier(d, a.d_hat, a.kappa) * m;
        }
    } else {
        if (a.en_g) {
            int2 c = a.vg[i - a.B - a.n_pt - a.n_ee];
            int v = a.voff[c.x] + c.y;
            double dd = a.X[3 * (size_t)v + 1] - a.ground_y, d = dd * dd;
            if (d < a.d_hat) e = barrier(d, a.d_hat, a.kappa);
        }
    }
    return e;
}
__global__ void k_energy_items(EnergyArgs a) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int n = a.B + a.n_pt + a.n_ee + a.n_vg;
    if (i >= n) return;
    a.out[i] = energy_item(a, i);
}

struct FricArgs {
    int n_pt, n_ee, n_vg;
    double mu, eps_v;
    int pt_f, ee_f, vg_f, en_pt, en_ee, en_g;
    const double* X; const double* Xs;
    const int4* pt; const int4* ee; const int2* vg;
    const int* voff; const int* foff; const int* eoff; const int* face; const int* edge;
    const double* tk; const double* lam; const double* vglam;
    double* out;
};
// friction energy of item i < n_pt + n_ee + n_vg (lagged terms)
__device__ __forceinline__ double fric_item(const FricArgs& a, int i) {
    double e = 0.0;
    if (i < a.n_pt + a.n_ee) {
        const bool is_ee = i >= a.n_pt;
        if (is_ee ? (a.en_ee && a.ee_f) : (a.en_pt && a.pt_f)) {
            int ids[4];
            if (!is_ee) {
                int4 c = a.pt[i];
                int f = a.foff[c.z] + c.w;
                ids[0] = a.voff[c.x] + c.y; ids[1] = a.face[3 * f]; ids[2] = a.face[3 * f + 1]; ids[3] = a.face[3 * f + 2];
            } else {
                int4 c = a.ee[i - a.n_pt];
                int ea = a.eoff[c.x] + c.y, eb = a.eoff[c.z] + c.w;
                ids[0] = a.edge[2 * ea]; ids[1] = a.edge[2 * ea + 1]; ids[2] = a.edge[2 * eb]; ids[3] = a.edge[2 * eb + 1];
            }
            // lagged basis: the weight of point k times each tangent direction
            const double* T = a.tk + TKN * (size_t)i;
            double u0 = 0.0, u1 = 0.0;
            for (int k = 0; k < 4; k++)
                for (int c = 0; c < 3; c++) {
                    double dv = a.X[3 * (size_t)ids[k] + c] - a.Xs[3 * (size_t)ids[k] + c];
                    u0 += (T[k] * T[4 + c]) * dv;
                    u1 += (T[k] * T[7 + c]) * dv;
                }
            e = a.mu * a.lam[i] * f0_sf(sqrt(u0 * u0 + u1 * u1), a.eps_v);
        }
    } else if (a.en_g && a.vg_f) {
        int k = i - a.n_pt - a.n_ee;
        int2 c = a.vg[k];
        int v = a.voff[c.x] + c.y;
        double ux = a.X[3 * (size_t)v] - a.Xs[3 * (size_t)v];
        double uz = a.X[3 * (size_t)v + 2] - a.Xs[3 * (size_t)v + 2];
        e = a.mu * a.vglam[k] * f0_sf(sqrt(ux * ux + uz * uz), a.eps_v);
    }
    return e;
}
__global__ void k_fric_items(FricArgs a) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int n = a.n_pt + a.n_ee + a.n_vg;
    if (i >= n) return;
    a.out[i] = fric_item(a, i);
}

// two-level fixed-shape reduction
constexpr int RT = 256;
__global__ __launch_bounds__(RT) void k_sum1(int n, int G, const double* __restrict__ a, const double* __restrict__ b,
                                            double* __restrict__ part) {
    __shared__ double sh[RT];
    double s = 0.0;
    for (long long i = (long long)blockIdx.x * RT + threadIdx.x; i < n; i += (long long)G * RT) s += b ? a[i] * b[i] : a[i];
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int m = RT / 2; m >= 1; m >>= 1) {
        if ((int)threadIdx.x < m) sh[threadIdx.x] += sh[threadIdx.x + m];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = sh[0];
}
__global__ __launch_bounds__(RT) void k_sum2(int G, const double* __restrict__ part, double* __restrict__ out) {
    __shared__ double sh[RT];
    double s = 0.0;
    for (int i = threadIdx.x; i < G; i += RT) s += part[i];
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int m = RT / 2; m >= 1; m >>= 1) {
        if ((int)threadIdx.x < m) sh[threadIdx.x] += sh[threadIdx.x + m];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[0] = sh[0];
}

// k_sum1 over two inputs at once (blockIdx.y selects a1 b1 -> part1 or
// a2 b2 -> part2)
__global__ __launch_bounds__(RT) void k_sum1x(int n, int G, const double* __restrict__ a1, const double* __restrict__ b1,
                                             double* __restrict__ part1, const double* __restrict__ a2,
                                             const double* __restrict__ b2, double* __restrict__ part2) {
    __shared__ double sh[RT];
    const double* a = blockIdx.y ? a2 : a1;
    const double* b = blockIdx.y ? b2 : b1;
    double s = 0.0;
    for (long long i = (long long)blockIdx.x * RT + threadIdx.x; i < n; i += (long long)G * RT) s += b ? a[i] * b[i] : a[i];
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int m = RT / 2; m >= 1; m >>= 1) {
        if ((int)threadIdx.x < m) sh[threadIdx.x] += sh[threadIdx.x + m];
        __syncthreads();
    }
    if (threadIdx.x == 0) (blockIdx.y ? part2 : part1)[blockIdx.x] = sh[0];
}

// The barrier + inertia items (blocks [0, G1)), the friction items (blocks
// [G1, G1 + G2)) and the products da[i] db[i], i < dn (blocks from G1 + G2)
// of one evaluation in one launch, one item per thread, each block's tree
// sum written to part1 / part2 / part3: the partials that k_sum1 forms over
// the item buffers (and over da, db) when no thread holds two items.
__global__ __launch_bounds__(RT) void k_energy_parts(EnergyArgs ea, int G1, FricArgs fa, int G2, const double* __restrict__ da,
                                                     const double* __restrict__ db, int dn, double* __restrict__ part1,
                                                     double* __restrict__ part2, double* __restrict__ part3) {
    __shared__ double sh[RT];
    const int blk = blockIdx.x;
    const int list = blk < G1 ? 0 : (blk < G1 + G2 ? 1 : 2);
    const int lb = list == 0 ? blk : (list == 1 ? blk - G1 : blk - G1 - G2);
    const int i = lb * RT + threadIdx.x;
    double s = 0.0;
    if (list == 0) { if (i < ea.B + ea.n_pt + ea.n_ee + ea.n_vg) s += energy_item(ea, i); }
    else if (list == 1) { if (i < fa.n_pt + fa.n_ee + fa.n_vg) s += fric_item(fa, i); }
    else if (i < dn) s += da[i] * db[i];
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int m = RT / 2; m >= 1; m >>= 1) {
        if ((int)threadIdx.x < m) sh[threadIdx.x] += sh[threadIdx.x + m];
        __syncthreads();
    }
    if (threadIdx.x == 0) (list == 0 ? part1 : (list == 1 ? part2 : part3))[lb] = sh[0];
}
// k_sum2 over up to three partial lists at once (block l: list l)
struct PartLists { int G[3]; const double* part[3]; double* out[3]; };
__global__ __launch_bounds__(RT) void k_sum2x(PartLists L) {
    __shared__ double sh[RT];
    const int l = blockIdx.x;
    const int G = L.G[l];
    const double* part = L.part[l];
    double s = 0.0;
    for (int i = threadIdx.x; i < G; i += RT) s += part[i];
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int m = RT / 2; m >= 1; m >>= 1) {
        if ((int)threadIdx.x < m) sh[threadIdx.x] += sh[threadIdx.x + m];
        __syncthreads();
    }
    if (threadIdx.x == 0) L.out[l][0] = sh[0];
}

// Sum of a[i] (b == null) or a[i] b[i] into w.epart[slot] (slots 0..7), on
// the world stream without waiting; read_slots fetches the results.
constexpr int NSLOT = EPART_NSLOT;
static void reduce_async(World& w, const double* a, const double* b, int n, int slot) {
    int G = (n + RT - 1) / RT;
    if (G > EPART_MAXG) G = EPART_MAXG;
    if (w.epart.cap < (size_t)EPART_SLOTS) throw AdbError(ADB_E_ARG, "reduction slots not reserved");
    if (n <= 0) {
        ADB_HIP(hipMemsetAsync(w.epart.p + slot, 0, sizeof(double), w.stream));
        return;
    }
    hipLaunchKernelGGL(k_sum1, dim3(G), dim3(RT), 0, w.stream, n, G, a, b, w.epart.p + NSLOT);
    hipLaunchKernelGGL(k_sum2, dim3(1), dim3(RT), 0, w.stream, G, w.epart.p + NSLOT, w.epart.p + slot);
    ADB_HIP(hipGetLastError());
}

void read_slots(World& w, int k, double* out) {
    ADB_HIP(hipMemcpyAsync(out, w.epart.p, k * sizeof(double), hipMemcpyDeviceToHost, w.stream));
    wait_stream(w);
}
void copy_slots_async(World& w, int first, int k) {
    if (!w.h_slots) ADB_HIP(hipHostMalloc((void**)&w.h_slots, NSLOT * sizeof(double), hipHostMallocDefault));
    ADB_HIP(hipMemcpyAsync(w.h_slots + first, w.epart.p + first, k * sizeof(double), hipMemcpyDeviceToHost, w.stream));
}

void dot_async(World& w, const double* a, const double* b, int n, int slot) { reduce_async(w, a, b, n, slot); }
// a1 . b1 into slot1 and a2 . b2 into slot2 in two launches, each sum as
// dot_async forms it
void dot2_async(World& w, const double* a1, const double* b1, int slot1, const double* a2, const double* b2, int slot2, int n) {
    if (n <= 0) { dot_async(w, a1, b1, n, slot1); dot_async(w, a2, b2, n, slot2); return; }
    int G = (n + RT - 1) / RT;
    if (G > EPART_MAXG) G = EPART_MAXG;
    if (w.epart.cap < (size_t)EPART_SLOTS) throw AdbError(ADB_E_ARG, "reduction slots not reserved");
    double* part1 = w.epart.p + NSLOT;
    double* part2 = w.epart.p + NSLOT + EPART_MAXG;
    hipLaunchKernelGGL(k_sum1x, dim3(G, 2), dim3(RT), 0, w.stream, n, G, a1, b1, part1, a2, b2, part2);
    PartLists L;
    L.G[0] = G; L.part[0] = part1; L.out[0] = w.epart.p + slot1;
    L.G[1] = G; L.part[1] = part2; L.out[1] = w.epart.p + slot2;
    L.G[2] = 0; L.part[2] = part2; L.out[2] = nullptr;
    hipLaunchKernelGGL(k_sum2x, dim3(2), dim3(RT), 0, w.stream, L);
    ADB_HIP(hipGetLastError());
}

static DBuf<double>& items(World& w) { return w.pcg_Ap; }   // scratch reused outside the solve

static EnergyArgs energy_args(World& w, const double* qv, const double* X, const ContactSet& cs) {
    EnergyArgs a;
    a.B = w.B; a.n_pt = cs.n_pt; a.n_ee = cs.n_ee; a.n_vg = cs.n_vg;
    a.dt = w.P.dt; a.kap_o = w.P.kappa_ortho; a.d_hat = w.P.d_hat; a.kappa = w.P.kappa; a.eps_x = w.P.eps_x;
    a.ground_y = w.P.ground_y; a.g0 = w.P.gravity[0]; a.g1 = w.P.gravity[1]; a.g2 = w.P.gravity[2];
    a.en_pt = w.P.enable_pt; a.en_ee = w.P.enable_ee; a.en_g = w.P.enable_ground;
    a.qv = qv; a.q0 = w.q0.p; a.qdot = w.qdot.p; a.mass = w.mass.p; a.Ic = w.Ic.p; a.X = X;
    a.pt = cs.pt.p; a.ee = cs.ee.p; a.vg = cs.vg.p;
    a.voff = w.voff.p; a.foff = w.foff.p; a.eoff = w.eoff.p; a.face = w.face.p; a.edge = w.edge.p;
    a.out = nullptr;
    return a;
}
static FricArgs fric_args(World& w, const double* X) {
    const ContactSet& cs = w.cs;
    FricArgs a;
    a.n_pt = cs.n_pt; a.n_ee = cs.n_ee; a.n_vg = cs.n_vg;
    a.mu = w.P.mu; a.eps_v = w.P.dt * w.P.evh;
    a.pt_f = w.P.pt_fric; a.ee_f = w.P.ee_fric; a.vg_f = w.P.vg_fric;
    a.en_pt = w.P.enable_pt; a.en_ee = w.P.enable_ee; a.en_g = w.P.enable_ground;
    a.X = X; a.Xs = w.Xs.p; a.pt = cs.pt.p; a.ee = cs.ee.p; a.vg = cs.vg.p;
    a.voff = w.voff.p; a.foff = w.foff.p; a.eoff = w.eoff.p; a.face = w.face.p; a.edge = w.edge.p;
    a.tk = w.ctk.p; a.lam = w.clam.p; a.vglam = w.vglam.p;
    a.out = nullptr;
    return a;
}

void energy_barrier_inertia_async(World& w, const double* qv, const double* X, const ContactSet& cs, int slot) {
    const int n = w.B + cs.n_pt + cs.n_ee + cs.n_vg;
    DBuf<double>& buf = items(w);
    buf.reserve(std::max(n, 12 * w.B));
    EnergyArgs a = energy_args(w, qv, X, cs);
    a.out = buf.p;
    ProfScope ps(w, PK_ENERGY, 0.0);
    hipLaunchKernelGGL(k_energy_items, dim3((n + 255) / 256), dim3(256), 0, w.stream, a);
    ADB_HIP(hipGetLastError());
    reduce_async(w, buf.p, nullptr, n, slot);
}

void energy_friction_async(World& w, const double* X, int slot) {
    const ContactSet& cs = w.cs;
    const int n = cs.n_pt + cs.n_ee + cs.n_vg;
    if (n == 0) { reduce_async(w, nullptr, nullptr, 0, slot); return; }
    DBuf<double>& buf = items(w);
    buf.reserve(std::max(n, 12 * w.B));
    FricArgs a = fric_args(w, X);
    a.out = buf.p;
    hipLaunchKernelGGL(k_fric_items, dim3((n + 255) / 256), dim3(256), 0, w.stream, a);
    ADB_HIP(hipGetLastError());
    reduce_async(w, buf.p, nullptr, n, slot);
}

// Barrier + inertia energy at (qv, X) over *cs into slot_bi (cs null: not
// evaluated), friction energy at X (lagged terms of w.cs) into slot_f and,
// with da, da . db (dn entries) into slot_d.  When every list fits one item
// per thread in EPART_MAXG blocks, two launches (k_energy_parts, k_sum2x)
// with the operations of the per-item buffers and reduce_async; else those.
void energy_async(World& w, const double* qv, const double* X, const ContactSet* cs, int slot_bi, int slot_f,
                  const double* da, const double* db, int dn, int slot_d) {
    const int n1 = cs ? w.B + cs->n_pt + cs->n_ee + cs->n_vg : 0;
    const int n2 = w.cs.n_pt + w.cs.n_ee + w.cs.n_vg;
    const int n3 = da ? dn : 0;
    const int G1 = (n1 + RT - 1) / RT, G2 = (n2 + RT - 1) / RT, G3 = (n3 + RT - 1) / RT;
    if (G1 > EPART_MAXG || G2 > EPART_MAXG || G3 > EPART_MAXG) {
        if (cs) energy_barrier_inertia_async(w, qv, X, *cs, slot_bi);
        energy_friction_async(w, X, slot_f);
        if (da) dot_async(w, da, db, dn, slot_d);
        return;
    }
    if (w.epart.cap < (size_t)EPART_SLOTS) throw AdbError(ADB_E_ARG, "reduction slots not reserved");
    EnergyArgs ea = cs ? energy_args(w, qv, X, *cs) : EnergyArgs{};
    const FricArgs fa = fric_args(w, X);
    double* part[3] = {w.epart.p + NSLOT, w.epart.p + NSLOT + EPART_MAXG, w.epart.p + NSLOT + 2 * EPART_MAXG};
    {
        ProfScope ps(w, PK_ENERGY, 0.0);
        if (G1 + G2 + G3 > 0)
            hipLaunchKernelGGL(k_energy_parts, dim3(G1 + G2 + G3), dim3(RT), 0, w.stream, ea, G1, fa, G2, da, db, n3, part[0], part[1],
                               part[2]);
        ADB_HIP(hipGetLastError());
    }
    // the lists that are evaluated, in order (an empty list sums to 0)
    PartLists L;
    int nl = 0;
    if (cs) { L.G[nl] = G1; L.part[nl] = part[0]; L.out[nl] = w.epart.p + slot_bi; nl++; }
    L.G[nl] = G2; L.part[nl] = part[1]; L.out[nl] = w.epart.p + slot_f; nl++;
    if (da) { L.G[nl] = G3; L.part[nl] = part[2]; L.out[nl] = w.epart.p + slot_d; nl++; }
    for (int l = nl; l < 3; l++) { L.G[l] = 0; L.part[l] = part[0]; L.out[l] = nullptr; }
    hipLaunchKernelGGL(k_sum2x, dim3(nl), dim3(RT), 0, w.stream, L);
    ADB_HIP(hipGetLastError());
}

// ---------------------------------------------------------------- per scene
// Sums over the ranges r < R of scene s: [base_r + mult_r off_r[s],
// base_r + mult_r off_r[s + 1]) of a[i] (b == null) or a[i] b[i]; one
// workgroup per scene, terms in a fixed order (strided, then a tree).
struct SceneRanges { int R; const int* off[4]; long long base[4]; int mult[4]; };
__global__ __launch_bounds__(RT) void k_scene_sum(int S, SceneRanges rg, const double* __restrict__ a,
                                                  const double* __restrict__ b, double* __restrict__ out, int accumulate) {
    __shared__ double sh[RT];
    const int s = blockIdx.x;
    double acc = 0.0;
    for (int r = 0; r < rg.R; r++) {
        const long long lo = rg.base[r] + (long long)rg.mult[r] * rg.off[r][s];
        const long long hi = rg.base[r] + (long long)rg.mult[r] * rg.off[r][s + 1];
        for (long long i = lo + threadIdx.x; i < hi; i += RT) acc += b ? a[i] * b[i] : a[i];
    }
    sh[threadIdx.x] = acc;
    __syncthreads();
    for (int m = RT / 2; m >= 1; m >>= 1) {
        if ((int)threadIdx.x < m) sh[threadIdx.x] += sh[threadIdx.x + m];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[s] = accumulate ? out[s] + sh[0] : sh[0];
}
// first entry of each scene in the sorted PT, EE and VG lists (entries are
// sorted by their first body, bodies by scene): seg[k (S + 1) + s], k = list
__global__ void k_scene_ranges(int S, const int* __restrict__ scene_off, int n_pt, const int4* __restrict__ pt, int n_ee,
                               const int4* __restrict__ ee, int n_vg, const int2* __restrict__ vg, int* __restrict__ seg) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= 3 * (S + 1)) return;
    const int k = t / (S + 1), s = t % (S + 1);
    const int n = k == 0 ? n_pt : (k == 1 ? n_ee : n_vg);
    if (s == S) { seg[t] = n; return; }
    const int b0 = scene_off[s];
    int lo = 0, hi = n;
    while (lo < hi) {
        const int mid = (lo + hi) / 2;
        const int I = k == 0 ? pt[mid].x : (k == 1 ? ee[mid].x : vg[mid].x);
        if (I < b0) lo = mid + 1; else hi = mid;
    }
    seg[t] = lo;
}

static void scene_ranges(World& w, const ContactSet& cs) {
    const int S = w.S;
    w.seg.reserve(3 * (size_t)(S + 1));
    hipLaunchKernelGGL(k_scene_ranges, dim3((3 * (S + 1) + 255) / 256), dim3(256), 0, w.stream, S, w.scene_off.p, cs.n_pt, cs.pt.p,
                       cs.n_ee, cs.ee.p, cs.n_vg, cs.vg.p, w.seg.p);
    ADB_HIP(hipGetLastError());
}

void dot_scenes_async(World& w, const double* a, const double* b, int slot) {
    SceneRanges rg;
    rg.R = 1; rg.off[0] = w.scene_off.p; rg.base[0] = 0; rg.mult[0] = 12;
    w.sred.reserve((size_t)SRED_N * w.S);
    hipLaunchKernelGGL(k_scene_sum, dim3(w.S), dim3(RT), 0, w.stream, w.S, rg, a, b, w.sred.p + (size_t)slot * w.S, 0);
    ADB_HIP(hipGetLastError());
}

// E_barrier_inert + E_fric per scene (model/energy.cpp:19-86,195-270)
void energy_scenes_async(World& w, const double* qv, const double* X, const ContactSet& cs, int slot) {
    const int S = w.S;
    w.sred.reserve((size_t)SRED_N * S);
    double* out = w.sred.p + (size_t)slot * S;
    scene_ranges(w, cs);
    const int* seg = w.seg.p;
    DBuf<double>& buf = items(w);
    const int n = w.B + cs.n_pt + cs.n_ee + cs.n_vg;
    buf.reserve(std::max(n, 12 * w.B));
    {
        EnergyArgs a;
        a.B = w.B; a.n_pt = cs.n_pt; a.n_ee = cs.n_ee; a.n_vg = cs.n_vg;
        a.dt = w.P.dt; a.kap_o = w.P.kappa_ortho; a.d_hat = w.P.d_hat; a.kappa = w.P.kappa; a.eps_x = w.P.eps_x;
        a.ground_y = w.P.ground_y; a.g0 = w.P.gravity[0]; a.g1 = w.P.gravity[1]; a.g2 = w.P.gravity[2];
        a.en_pt = w.P.enable_pt; a.en_ee = w.P.enable_ee; a.en_g = w.P.enable_ground;
        a.qv = qv; a.q0 = w.q0.p; a.qdot = w.qdot.p; a.mass = w.mass.p; a.Ic = w.Ic.p; a.X = X;
        a.pt = cs.pt.p; a.ee = cs.ee.p; a.vg = cs.vg.p;
        a.voff = w.voff.p; a.foff = w.foff.p; a.eoff = w.eoff.p; a.face = w.face.p; a.edge = w.edge.p;
        a.out = buf.p;
        ProfScope ps(w, PK_ENERGY, 0.0);
        hipLaunchKernelGGL(k_energy_items, dim3((n + 255) / 256), dim3(256), 0, w.stream, a);
        ADB_HIP(hipGetLastError());
        SceneRanges rg;
        rg.R = 4;
        rg.off[0] = w.scene_off.p; rg.base[0] = 0; rg.mult[0] = 1;
        rg.off[1] = seg; rg.base[1] = w.B; rg.mult[1] = 1;
        rg.off[2] = seg + (S + 1); rg.base[2] = (long long)w.B + cs.n_pt; rg.mult[2] = 1;
        rg.off[3] = seg + 2 * (S + 1); rg.base[3] = (long long)w.B + cs.n_pt + cs.n_ee; rg.mult[3] = 1;
        hipLaunchKernelGGL(k_scene_sum, dim3(S), dim3(RT), 0, w.stream, S, rg, buf.p, (const double*)nullptr, out, 0);
        ADB_HIP(hipGetLastError());
    }
    // lagged friction on the step's contact set w.cs (its own ranges)
    const ContactSet& fc = w.cs;
    const int nf = fc.n_pt + fc.n_ee + fc.n_vg;
    if (nf == 0) return;
    scene_ranges(w, fc);
    FricArgs a;
    a.n_pt = fc.n_pt; a.n_ee = fc.n_ee; a.n_vg = fc.n_vg;
    a.mu = w.P.mu; a.eps_v = w.P.dt * w.P.evh;
    a.pt_f = w.P.pt_fric; a.ee_f = w.P.ee_fric; a.vg_f = w.P.vg_fric;
    a.en_pt = w.P.enable_pt; a.en_ee = w.P.enable_ee; a.en_g = w.P.enable_ground;
    a.X = X; a.Xs = w.Xs.p; a.pt = fc.pt.p; a.ee = fc.ee.p; a.vg = fc.vg.p;
    a.voff = w.voff.p; a.foff = w.foff.p; a.eoff = w.eoff.p; a.face = w.face.p; a.edge = w.edge.p;
    a.tk = w.ctk.p; a.lam = w.clam.p; a.vglam = w.vglam.p;
    a.out = buf.p;
    hipLaunchKernelGGL(k_fric_items, dim3((nf + 255) / 256), dim3(256), 0, w.stream, a);
    ADB_HIP(hipGetLastError());
    SceneRanges rg;
    rg.R = 3;
    rg.off[0] = seg; rg.base[0] = 0; rg.mult[0] = 1;
    rg.off[1] = seg + (S + 1); rg.base[1] = fc.n_pt; rg.mult[1] = 1;
    rg.off[2] = seg + 2 * (S + 1); rg.base[2] = (long long)fc.n_pt + fc.n_ee; rg.mult[2] = 1;
    hipLaunchKernelGGL(k_scene_sum, dim3(S), dim3(RT), 0, w.stream, S, rg, buf.p, (const double*)nullptr, out, 1);
    ADB_HIP(hipGetLastError());
}

void read_scene_sums(World& w, int nslots) {
    if (!w.h_sred) ADB_HIP(hipHostMalloc((void**)&w.h_sred, (size_t)SRED_N * w.S * sizeof(double), hipHostMallocDefault));
    ADB_HIP(hipMemcpyAsync(w.h_sred, w.sred.p, (size_t)nslots * w.S * sizeof(double), hipMemcpyDeviceToHost, w.stream));
    wait_stream(w);
}

} // namespace adb
