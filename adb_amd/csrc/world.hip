// world.hip - host driver of the Newton contact step and the C ABI.
//
// adb_step replaces ABD::implicit_euler (model/abd.cpp:30-140) with its
// stages ABD::ipc, ABD::solve, ABD::ccd and ABD::line_search
// (model/abd.cpp:142-404, model/line_search.cpp:39-136).  The control flow is
// the reference's: q <- q0; contact set at q; repeat { inertia + contact
// terms, solve, CCD bound with backoff, Armijo backtracking with a fresh
// contact set per trial, q += alpha dq } until ||dq|| < tol; then
// qdot = (q - q0) / dt.  All state stays on the device; the host reads only
// counters and the scalars that drive the control flow.
#include <chrono>
#include <cmath>
#include <cstring>
#include <cstdlib>
#include <algorithm>
#include "dmath.h"
#include "world.h"

struct adb_world { adb::World w; };

namespace adb {

__global__ void k_scale(int n, double* v, double s) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) v[i] = v[i] * s;
}
__global__ void k_velocity(int n, double dt, const double* __restrict__ q, double* __restrict__ q0, double* __restrict__ qdot) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) { qdot[i] = (q[i] - q0[i]) / dt; q0[i] = q[i]; }
}
// batched worlds: scenes whose step failed (keep[scene] == 0) keep their state
__global__ void k_velocity_scenes(int n, double dt, const int* __restrict__ scene_of, const double* __restrict__ keep,
                                  const double* __restrict__ q, double* __restrict__ q0, double* __restrict__ qdot) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n && keep[scene_of[i / 12]] != 0.0) { qdot[i] = (q[i] - q0[i]) / dt; q0[i] = q[i]; }
}

static double now_s() { return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count(); }

template <class T> static void upload(World& w, DBuf<T>& b, const T* src, size_t n) {
    b.reserve(std::max<size_t>(n, 1));
    if (n) ADB_HIP(hipMemcpyAsync(b.p, src, n * sizeof(T), hipMemcpyHostToDevice, w.stream));
}

static void scale(World& w, double* v, int n, double s) {
    hipLaunchKernelGGL(k_scale, dim3((n + 255) / 256), dim3(256), 0, w.stream, n, v, s);
    ADB_HIP(hipGetLastError());
}

void wait_stream(World& w) {
    if (!w.spin_wait) { ADB_HIP(hipStreamSynchronize(w.stream)); return; }
    hipError_t e;
    while ((e = hipStreamQuery(w.stream)) == hipErrorNotReady) {}
    ADB_HIP(e);
}
void wait_stream(World& w, hipStream_t s) {
    if (!w.spin_wait) { ADB_HIP(hipStreamSynchronize(s)); return; }
    hipError_t e;
    while ((e = hipStreamQuery(s)) == hipErrorNotReady) {}
    ADB_HIP(e);
}
void wait_event(World& w, hipEvent_t ev) {
    if (!w.spin_wait) { ADB_HIP(hipEventSynchronize(ev)); return; }
    hipError_t e;
    while ((e = hipEventQuery(ev)) == hipErrorNotReady) {}
    ADB_HIP(e);
}

// Armijo backtracking (model/line_search.cpp:39-136) from w.q along w.dq (the
// direction already scaled by the CCD bound).  On entry w.X and w.cs hold the
// positions and the contact set at q, w.grad the gradient there and the
// lagged friction terms are those of q; gn = ||grad||.  On return w.qtrial,
// w.X1 and w.cs_new hold the accepted point q + alpha dq.
static int line_search(World& w, double gn, double* alpha_out, int* trials_out, double* E0_out, double* E1_out) {
    const int n = 12 * w.B;
    double sl[7];
    // E0 = barrier + inertia + friction at q, and dq . grad, queued now and read
    // with the first trial's energies (one wait).  The barrier + inertia part is
    // that of the point the previous line search accepted when there was one:
    // the same q, positions and contact set, so the same value bit for bit.
    // The friction part is recomputed: its lagged terms are this iteration's.
    const bool have_bi = w.e0_valid;
    energy_async(w, w.q.p, w.X.p, have_bi ? nullptr : &w.cs, 5, 6, w.dq.p, w.grad.p, n, 2);
    double E0 = 0.0, qdg = 0.0, E1 = 0.0, alpha = 1.0, e1_bi = 0.0;
    bool have_e0 = false;
    int trials = 0, status = ADB_OK;
    // Trial sets from the candidates of the search segment q -> q + alpha dq
    // (vertex positions are affine in q, so every later trial point lies on
    // it): one swept pass, then an exact filter per trial.  The swept pass
    // costs more than one plain pass, so when the previous search took one
    // trial the first trial gets a plain pass and the candidates are built
    // only if it is rejected.  Both give the same sets.
    const bool cand_ok = w.ls_cand && (w.P.iaabb & 1);
    bool cand = cand_ok && w.ls_prev > 1;
    if (cand) {
        transform_step(w, w.q.p, w.dq.p, 1.0, w.qtrial.p, w.X1.p);
        segment_candidates(w, w.X.p, w.X1.p, w.cand);
    }
    while (true) {
        transform_step(w, w.q.p, w.dq.p, alpha, w.qtrial.p, w.X1.p);
        if (!cand && cand_ok && trials > 0) {
            segment_candidates(w, w.X.p, w.X1.p, w.cand);
            cand = true;
        }
        if (cand) trial_set(w, w.X1.p, w.cand, w.cs_new);
        else collision_set(w, w.X1.p, w.cs_new);
        energy_async(w, w.qtrial.p, w.X1.p, &w.cs_new, 0, 1);
        read_slots(w, have_e0 ? 2 : 7, sl);
        if (!have_e0) {
            E0 = (have_bi ? w.e0_bi : sl[5]) + sl[6];
            qdg = sl[2];
            have_e0 = true;
        }
        e1_bi = sl[0];
        E1 = sl[0] + sl[1];
        const bool wolfe = E1 <= E0 + w.P.c1 * alpha * qdg;
        alpha /= 2;
        trials++;
        if (wolfe || !(gn > 1e-3)) break;
        if (trials >= w.P.max_ls_iter) { status = ADB_E_NOT_CONVERGED; w.err = "line search exhausted"; break; }
    }
    w.ls_prev = trials;
    w.e0_valid = status == ADB_OK;
    w.e0_bi = e1_bi;
    *alpha_out = alpha * 2; *trials_out = trials; *E0_out = E0; *E1_out = E1;
    return status;
}

static int step(World& w, adb_step_stats* st) {
    adb_step_stats S;
    std::memset(&S, 0, sizeof S);
    S.toi_min = 1.0;
    const double t_start = now_s();
    const int n = 12 * w.B;
    // q = q0, and X = Xs = the positions at q0, in one launch
    transform_step(w, w.q0.p, nullptr, 0.0, w.q.p, w.X.p, w.Xs.p);
    collision_set(w, w.X.p, w.cs);
    w.e0_valid = false;
    int status = ADB_OK;
    bool converged = false;
    for (int iter = 0; iter < w.P.max_iter; iter++) {
        const double t0 = now_s();
        if (!w.ev_phase[0])
            for (hipEvent_t& e : w.ev_phase) ADB_HIP(hipEventCreate(&e));
        ADB_HIP(hipEventRecord(w.ev_phase[0], w.stream));
        contact_terms(w);
        assemble_system(w);
        ADB_HIP(hipEventRecord(w.ev_phase[1], w.stream));
        S.contacts_processed += (long long)w.cs.n_pt + w.cs.n_ee + w.cs.n_vg;
        // no wait here: the solve is queued behind the assembly, and the ipc
        // bucket gets the device time between the two events (the rest of the
        // host time to the solve's end goes to the solver bucket)
        // |dq|^2 and |g|^2 (slots 3, 4) copied to the host and the CCD step
        // bound (model/abd.cpp:142-196) queued: read in one wait.  Behind a
        // short solve they are queued ahead of its stop test's wait (and
        // queued again if that solve turns out unfinished).
        auto queue_bound = [&]() {
            dot2_async(w, w.dq.p, w.dq.p, 3, w.grad.p, w.grad.p, 4, n);
            copy_slots_async(w, 3, 2);
            if (w.P.upper_bound) {
                transform_step(w, w.q.p, w.dq.p, 1.0, w.qtrial.p, w.X1.p);
                ccd_queue(w, w.X.p, w.X1.p);
            }
        };
        int pit = 0;
        bool queued = false;
        const int prc = pcg_solve(w, w.dq.p, &pit, queue_bound, &queued);
        S.pcg_iters += pit;
        const double t2 = now_s();
        {
            ADB_HIP(hipEventSynchronize(w.ev_phase[1]));
            float ms = 0.f;
            ADB_HIP(hipEventElapsedTime(&ms, w.ev_phase[0], w.ev_phase[1]));
            const double t_asm = std::min(t2 - t0, 1e-3 * (double)ms);
            S.t_ipc += t_asm;
            S.t_solver += (t2 - t0) - t_asm;
        }
        if (!queued) queue_bound();
        double toi = 1.0, factor = 1.0;
        int cs_status = ADB_OK;
        if (w.P.upper_bound) toi = ccd_finish(w, w.X.p, w.X1.p, &cs_status);
        else wait_stream(w);
        const double sl[2] = {w.h_slots[3], w.h_slots[4]};
        // a solve that stopped early is counted, not hidden: the direction is
        // still the last PCG iterate and the line search still guards descent
        if (prc == ADB_E_NOT_CONVERGED) {
            S.pcg_not_converged++;
            w.err = "PCG reached pcg_max_iter at Newton iteration " + std::to_string(iter);
        } else if (prc == ADB_E_SOLVER) {
            S.pcg_breakdown++;
            w.err = "PCG breakdown (system not positive definite) at Newton iteration " + std::to_string(iter);
        }
        const double nrm2 = prc == ADB_E_NAN ? NAN : sl[0];
        const double gn = std::sqrt(sl[1]);
        const double sup_dq = std::sqrt(nrm2);
        if (std::isnan(sup_dq)) { status = ADB_E_NAN; w.err = "solver nan (model/abd.cpp:288-291)"; break; }
        if (cs_status != ADB_OK) { status = cs_status; w.err = "toi < 1e-6 twice in a row (model/iaabb.cpp:305-311,533-537)"; break; }
        if (toi < 1.0) factor = w.P.backoff;
        S.toi_min = std::min(S.toi_min, toi);
        scale(w, w.dq.p, n, factor * toi);
        const double t3 = now_s();
        S.t_ccd += t3 - t2;
        // Armijo backtracking (model/line_search.cpp:39-136)
        double alpha = 1.0;
        if (w.P.line_search) {
            int trials = 0;
            double E0, E1;
            status = line_search(w, gn, &alpha, &trials, &E0, &E1);
            S.ls_trials += trials;
            if (status != ADB_OK) break;
        } else {
            transform_step(w, w.q.p, w.dq.p, 1.0, w.qtrial.p, w.X1.p);
            collision_set(w, w.X1.p, w.cs_new);
            w.e0_valid = false;
        }
        // q += alpha dq is exactly the last trial's q + dq * alpha
        std::swap(w.q, w.qtrial);
        std::swap(w.X, w.X1);
        std::swap(w.cs, w.cs_new);
        S.t_line_search += now_s() - t3;
        S.newton_iters++;
        if (sup_dq < w.P.tol) { converged = true; break; }
    }
    if (status == ADB_OK && !converged) status = ADB_E_NOT_CONVERGED;
    if (status == ADB_OK || status == ADB_E_NOT_CONVERGED) {
        hipLaunchKernelGGL(k_velocity, dim3((n + 255) / 256), dim3(256), 0, w.stream, n, w.P.dt, w.q.p, w.q0.p, w.qdot.p);
        ADB_HIP(hipGetLastError());
    }
    wait_stream(w);
    S.n_pt = w.cs.n_pt; S.n_ee = w.cs.n_ee; S.n_vg = w.cs.n_vg;
    S.t_total = now_s() - t_start;
    if (st) *st = S;
    return status;
}

// One step of every scene of a batched world (adb_world_create_batched).  The
// control flow of step() per scene: each scene has its own Newton iteration
// count, convergence test on its own ||dq||, CCD bound and backoff, and
// Armijo search with its own alpha, energies and gradient; a scene that has
// converged keeps its state (its gradient is zeroed, so its PCG rows stay at
// dq = 0) while the others iterate.  Every stage is one launch sequence over
// all scenes; per-scene scalars are segmented sums over each scene's bodies
// and contacts.
static int step_batched(World& w, adb_step_stats* st) {
    adb_step_stats S_;
    std::memset(&S_, 0, sizeof S_);
    S_.toi_min = 1.0;
    const double t_start = now_s();
    const int S = w.S, n = 12 * w.B;
    std::vector<char> active(S, 1);
    std::vector<int> newton(S, 0), lst(S, 0), status(S, ADB_OK);
    std::vector<double> fac(S), alpha(S), toi(S), E0(S), qdg(S), gn(S), sup(S);
    std::vector<int> cst(S);
    w.sfac.reserve(S);
    auto put_factors = [&](const std::vector<double>& f) {
        ADB_HIP(hipMemcpyAsync(w.sfac.p, f.data(), S * sizeof(double), hipMemcpyHostToDevice, w.stream));
    };
    ADB_HIP(hipMemcpyAsync(w.q.p, w.q0.p, n * sizeof(double), hipMemcpyDeviceToDevice, w.stream));
    transform_positions(w, w.q.p, w.X.p);
    transform_positions(w, w.q0.p, w.Xs.p);
    collision_set(w, w.X.p, w.cs);
    int iters = 0;
    for (int iter = 0; iter < w.P.max_iter; iter++) {
        bool any = false;
        for (int s = 0; s < S; s++) any = any || active[s];
        if (!any) break;
        iters++;
        const double t0 = now_s();
        contact_terms(w);
        assemble_system(w);
        S_.contacts_processed += (long long)w.cs.n_pt + w.cs.n_ee + w.cs.n_vg;
        // scenes that have converged: zero gradient, so dq = 0 on their rows
        for (int s = 0; s < S; s++) fac[s] = active[s] ? 1.0 : 0.0;
        put_factors(fac);
        scale_scenes(w, w.grad.p, w.sfac.p);
        wait_stream(w);
        const double t1 = now_s();
        S_.t_ipc += t1 - t0;
        int pit = 0;
        const int prc = pcg_solve(w, w.dq.p, &pit);
        S_.pcg_iters += pit;
        if (prc == ADB_E_NOT_CONVERGED) { S_.pcg_not_converged++; w.err = "PCG reached pcg_max_iter"; }
        else if (prc == ADB_E_SOLVER) { S_.pcg_breakdown++; w.err = "PCG breakdown (system not positive definite)"; }
        dot_scenes_async(w, w.dq.p, w.dq.p, 0);
        dot_scenes_async(w, w.grad.p, w.grad.p, 1);
        read_scene_sums(w, 2);
        for (int s = 0; s < S; s++) {
            sup[s] = prc == ADB_E_NAN ? NAN : std::sqrt(w.h_sred[s]);
            gn[s] = std::sqrt(w.h_sred[S + s]);
            if (active[s] && std::isnan(sup[s])) { status[s] = ADB_E_NAN; active[s] = 0; }
        }
        const double t2 = now_s();
        S_.t_solver += t2 - t1;
        // CCD bounds (model/abd.cpp:142-196), per scene
        for (int s = 0; s < S; s++) toi[s] = 1.0;
        if (w.P.upper_bound) {
            q_axpy(w, w.q.p, w.dq.p, 1.0, w.qtrial.p);
            transform_positions(w, w.qtrial.p, w.X1.p);
            ccd_toi_scenes(w, w.X.p, w.X1.p, toi.data(), cst.data());
            for (int s = 0; s < S; s++)
                if (active[s] && cst[s] != ADB_OK) { status[s] = cst[s]; active[s] = 0; }
        }
        for (int s = 0; s < S; s++) {
            fac[s] = active[s] ? (toi[s] < 1.0 ? w.P.backoff : 1.0) * toi[s] : 0.0;
            if (active[s]) S_.toi_min = std::min(S_.toi_min, toi[s]);
        }
        put_factors(fac);
        scale_scenes(w, w.dq.p, w.sfac.p);
        const double t3 = now_s();
        S_.t_ccd += t3 - t2;
        // Armijo backtracking (model/line_search.cpp:39-136), per scene
        std::vector<char> done(S), exhausted(S, 0);
        for (int s = 0; s < S; s++) { alpha[s] = active[s] ? 1.0 : 0.0; done[s] = !active[s]; }
        if (w.P.line_search) {
            energy_scenes_async(w, w.q.p, w.X.p, w.cs, 0);
            dot_scenes_async(w, w.dq.p, w.grad.p, 1);
            read_scene_sums(w, 2);
            for (int s = 0; s < S; s++) { E0[s] = w.h_sred[s]; qdg[s] = w.h_sred[S + s]; }
            const bool cand_ok = w.ls_cand;
            bool cand = false;
            if (cand_ok) {
                q_axpy(w, w.q.p, w.dq.p, 1.0, w.qtrial.p);
                transform_positions(w, w.qtrial.p, w.X1.p);
                segment_candidates(w, w.X.p, w.X1.p, w.cand);
                cand = true;
            }
            std::vector<int> trials(S, 0);
            while (true) {
                put_factors(alpha);
                q_axpy_scenes(w, w.q.p, w.dq.p, w.sfac.p, w.qtrial.p);
                transform_positions(w, w.qtrial.p, w.X1.p);
                if (cand) trial_set(w, w.X1.p, w.cand, w.cs_new);
                else collision_set(w, w.X1.p, w.cs_new);
                energy_scenes_async(w, w.qtrial.p, w.X1.p, w.cs_new, 0);
                read_scene_sums(w, 1);
                bool all = true;
                for (int s = 0; s < S; s++) {
                    if (done[s]) continue;
                    const double E1 = w.h_sred[s];
                    trials[s]++;
                    if (E1 <= E0[s] + w.P.c1 * alpha[s] * qdg[s] || !(gn[s] > 1e-3)) { done[s] = 1; continue; }
                    if (trials[s] >= w.P.max_ls_iter) {
                        // as step(): the scene stops at the iterate before this search
                        status[s] = ADB_E_NOT_CONVERGED; done[s] = 1; active[s] = 0; exhausted[s] = 1;
                        continue;
                    }
                    alpha[s] /= 2;
                    all = false;
                }
                if (all) break;
            }
            for (int s = 0; s < S; s++) { lst[s] += trials[s]; S_.ls_trials += trials[s]; }
            bool redo = false;
            for (int s = 0; s < S; s++)
                if (exhausted[s] && alpha[s] != 0.0) { alpha[s] = 0.0; redo = true; }
            if (redo) {
                put_factors(alpha);
                q_axpy_scenes(w, w.q.p, w.dq.p, w.sfac.p, w.qtrial.p);
                transform_positions(w, w.qtrial.p, w.X1.p);
                collision_set(w, w.X1.p, w.cs_new);
            }
        } else {
            put_factors(alpha);
            q_axpy_scenes(w, w.q.p, w.dq.p, w.sfac.p, w.qtrial.p);
            transform_positions(w, w.qtrial.p, w.X1.p);
            collision_set(w, w.X1.p, w.cs_new);
        }
        std::swap(w.q, w.qtrial);
        std::swap(w.X, w.X1);
        std::swap(w.cs, w.cs_new);
        S_.t_line_search += now_s() - t3;
        for (int s = 0; s < S; s++) {
            if (!active[s]) continue;
            newton[s]++;
            if (sup[s] < w.P.tol) active[s] = 0;
        }
    }
    for (int s = 0; s < S; s++)
        if (active[s] && status[s] == ADB_OK) status[s] = ADB_E_NOT_CONVERGED;
    // velocity update (model/abd.cpp:128-134) for the scenes whose step ended
    // with ADB_OK or ADB_E_NOT_CONVERGED, as step()
    for (int s = 0; s < S; s++) fac[s] = (status[s] == ADB_OK || status[s] == ADB_E_NOT_CONVERGED) ? 1.0 : 0.0;
    put_factors(fac);
    hipLaunchKernelGGL(k_velocity_scenes, dim3((n + 255) / 256), dim3(256), 0, w.stream, n, w.P.dt, w.scene_of.p, w.sfac.p,
                       w.q.p, w.q0.p, w.qdot.p);
    ADB_HIP(hipGetLastError());
    // contacts per scene at the end of the step
    {
        std::vector<int> pt, ee, vg;
        download_sets(w, w.cs, pt, ee, vg);
        std::vector<int> cnt(S, 0);
        auto scene_of_body = [&](int b) {
            return (int)(std::upper_bound(w.h_scene_off.begin(), w.h_scene_off.end(), b) - w.h_scene_off.begin()) - 1;
        };
        for (size_t i = 0; i < pt.size(); i += 4) cnt[scene_of_body(pt[i])]++;
        for (size_t i = 0; i < ee.size(); i += 4) cnt[scene_of_body(ee[i])]++;
        for (size_t i = 0; i < vg.size(); i += 2) cnt[scene_of_body(vg[i])]++;
        w.sc_contacts = cnt;
    }
    wait_stream(w);
    w.sc_newton = newton; w.sc_ls = lst; w.sc_status = status;
    S_.newton_iters = *std::max_element(newton.begin(), newton.end());
    S_.n_pt = w.cs.n_pt; S_.n_ee = w.cs.n_ee; S_.n_vg = w.cs.n_vg;
    S_.t_total = now_s() - t_start;
    if (st) *st = S_;
    int rc = ADB_OK;
    for (int s = 0; s < S; s++)
        if (status[s] != ADB_OK && (rc == ADB_OK || rc == ADB_E_NOT_CONVERGED)) rc = status[s];
    (void)iters;
    return rc;
}

void download_sets(World& w, const ContactSet& cs, std::vector<int>& pt, std::vector<int>& ee, std::vector<int>& vg) {
    pt.resize(4 * (size_t)cs.n_pt); ee.resize(4 * (size_t)cs.n_ee); vg.resize(2 * (size_t)cs.n_vg);
    if (cs.n_pt) ADB_HIP(hipMemcpyAsync(pt.data(), cs.pt.p, pt.size() * sizeof(int), hipMemcpyDeviceToHost, w.stream));
    if (cs.n_ee) ADB_HIP(hipMemcpyAsync(ee.data(), cs.ee.p, ee.size() * sizeof(int), hipMemcpyDeviceToHost, w.stream));
    if (cs.n_vg) ADB_HIP(hipMemcpyAsync(vg.data(), cs.vg.p, vg.size() * sizeof(int), hipMemcpyDeviceToHost, w.stream));
    wait_stream(w);
}

// assemble at the current q (= q0 between steps)
static void assemble_now(World& w) {
    const int n = 12 * w.B;
    ADB_HIP(hipMemcpyAsync(w.q.p, w.q0.p, n * sizeof(double), hipMemcpyDeviceToDevice, w.stream));
    transform_positions(w, w.q.p, w.X.p);
    transform_positions(w, w.q0.p, w.Xs.p);
    collision_set(w, w.X.p, w.cs);
    contact_terms(w);
    assemble_system(w);
}

} // namespace adb

using namespace adb;

#define ADB_TRY(wptr, ...)                                                    \
    try { __VA_ARGS__ }                                                       \
    catch (const HipError& e) { if (wptr) (wptr)->w.err = e.what; return ADB_E_HIP; } \
    catch (const AdbError& e) { if (wptr) (wptr)->w.err = e.what; return e.code; }    \
    catch (const std::bad_alloc& e) { if (wptr) (wptr)->w.err = "host allocation failed"; return ADB_E_CAPACITY; } \
    catch (const std::exception& e) { if (wptr) (wptr)->w.err = e.what(); return ADB_E_ARG; } \
    catch (...) { if (wptr) (wptr)->w.err = "unknown exception"; return ADB_E_ARG; }

extern "C" void adb_params_default(adb_params* p) {
    std::memset(p, 0, sizeof *p);
    p->dt = 1e-2; p->d_hat = 1e-4; p->kappa = 1e-4; p->kappa_ortho = 1e9; p->mu = 0.0; p->evh = 1e-3;
    p->eps_x = 1e-3; p->tol = 1e-2; p->c1 = 1e-4; p->backoff = 0.8; p->max_uk = 0.0;
    p->gravity[0] = 0.0; p->gravity[1] = -9.8; p->gravity[2] = 0.0; p->ground_y = -0.5;
    p->max_iter = 1000; p->max_ls_iter = 200;
    p->enable_pt = p->enable_ee = p->enable_ground = 1;
    p->pt_fric = p->ee_fric = p->vg_fric = 1;
    p->psd = 1; p->line_search = 1; p->upper_bound = 1; p->broad_phase_fp32 = 0;
    p->pcg_rtol = 1e-12; p->pcg_max_iter = 2000;
    p->iaabb = 3; p->sh_bits = 6; p->sh_min_xyz = -10.0; p->sh_dx = 0.8;
}

extern "C" int adb_device_count(void) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return ADB_E_NO_DEVICE;
    int good = 0;
    for (int i = 0; i < n; i++) {
        hipDeviceProp_t prop;
        if (hipGetDeviceProperties(&prop, i) == hipSuccess && std::strncmp(prop.gcnArchName, "gfx950", 6) == 0) good++;
    }
    return good;
}

static void free_world(World& w);

extern "C" int adb_world_create(const adb_scene* s, const adb_params* p, adb_world** out) {
    if (!s || !p || !out || s->n_bodies <= 0) return ADB_E_ARG;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return ADB_E_NO_DEVICE;
    adb_world* W = new adb_world();
    World& w = W->w;
    try {
        w.P = *p;
        if ((w.P.iaabb & 3) != 3 && !(w.P.sh_bits >= 1 && w.P.sh_bits <= 8 && w.P.sh_dx > 0.0 && std::isfinite(w.P.sh_dx) &&
                                       std::isfinite(w.P.sh_min_xyz)))
            throw AdbError(ADB_E_ARG, "spatial hashing needs 1 <= sh_bits <= 8 and a finite sh_dx > 0");
        w.B = s->n_bodies;
        if (const char* g = std::getenv("ADB_PCG_GRAPH")) w.use_graphs = std::strcmp(g, "0") != 0;
        if (const char* g = std::getenv("ADB_LS_CANDIDATES")) w.ls_cand = std::strcmp(g, "0") != 0;
        if (const char* g = std::getenv("ADB_SPIN_WAIT")) w.spin_wait = std::strcmp(g, "0") != 0;
        if (const char* g = std::getenv("ADB_GENERAL_PATHS")) w.general_paths = std::strcmp(g, "0") != 0;
        w.h_voff.assign(s->vert_offset, s->vert_offset + w.B + 1);
        w.h_foff.assign(s->face_offset, s->face_offset + w.B + 1);
        w.h_eoff.assign(s->edge_offset, s->edge_offset + w.B + 1);
        w.V = w.h_voff[w.B]; w.F = w.h_foff[w.B]; w.E = w.h_eoff[w.B];
        // body pairs by sort and sweep above a few thousand bodies (the
        // all-pairs test costs B^2 / 2); both give the same pairs
        w.sweep = w.B > 4096;
        if (w.V <= 0) throw AdbError(ADB_E_ARG, "scene has no vertices");
        ADB_HIP(hipStreamCreate(&w.stream));
        hipDeviceProp_t prop;
        int dev = 0;
        ADB_HIP(hipGetDevice(&dev));
        ADB_HIP(hipGetDeviceProperties(&prop, dev));
        w.num_cu = prop.multiProcessorCount;
        std::vector<int> vbody(w.V), fbody(std::max(w.F, 1)), ebody(std::max(w.E, 1)), face(3 * (size_t)w.F), edge(2 * (size_t)w.E);
        for (int b = 0; b < w.B; b++) {
            w.max_vpb = std::max(w.max_vpb, w.h_voff[b + 1] - w.h_voff[b]);
            for (int v = w.h_voff[b]; v < w.h_voff[b + 1]; v++) vbody[v] = b;
            for (int f = w.h_foff[b]; f < w.h_foff[b + 1]; f++) {
                fbody[f] = b;
                for (int k = 0; k < 3; k++) {
                    int lv = s->faces[3 * (size_t)f + k];
                    if (lv < 0 || lv >= w.h_voff[b + 1] - w.h_voff[b]) throw AdbError(ADB_E_ARG, "face index out of range");
                    face[3 * (size_t)f + k] = w.h_voff[b] + lv;
                }
            }
            for (int e = w.h_eoff[b]; e < w.h_eoff[b + 1]; e++) {
                ebody[e] = b;
                for (int k = 0; k < 2; k++) {
                    int lv = s->edges[2 * (size_t)e + k];
                    if (lv < 0 || lv >= w.h_voff[b + 1] - w.h_voff[b]) throw AdbError(ADB_E_ARG, "edge index out of range");
                    edge[2 * (size_t)e + k] = w.h_voff[b] + lv;
                }
            }
        }
        upload(w, w.voff, w.h_voff.data(), w.B + 1);
        upload(w, w.foff, w.h_foff.data(), w.B + 1);
        upload(w, w.eoff, w.h_eoff.data(), w.B + 1);
        upload(w, w.vbody, vbody.data(), w.V);
        upload(w, w.fbody, fbody.data(), fbody.size());
        upload(w, w.ebody, ebody.data(), ebody.size());
        upload(w, w.face, face.data(), face.size());
        upload(w, w.edge, edge.data(), edge.size());
        upload(w, w.xbar, s->rest_verts, 3 * (size_t)w.V);
        upload(w, w.mass, s->mass, w.B);
        upload(w, w.Ic, s->Ic, w.B);
        upload(w, w.q0, s->q0, 12 * (size_t)w.B);
        upload(w, w.qdot, s->dqdt, 12 * (size_t)w.B);
        w.h_mass.assign(s->mass, s->mass + w.B);
        const size_t n = 12 * (size_t)w.B;
        w.q.reserve(n); w.dq.reserve(n); w.qtrial.reserve(n);
        ADB_HIP(hipMemcpyAsync(w.q.p, w.q0.p, n * sizeof(double), hipMemcpyDeviceToDevice, w.stream));
        w.X.reserve(3 * (size_t)w.V); w.Xs.reserve(3 * (size_t)w.V); w.X1.reserve(3 * (size_t)w.V);
        w.counters.reserve(16);
        w.epart.reserve(EPART_SLOTS);
        wait_stream(w);
    } catch (const HipError& e) {
        free_world(w); delete W; return ADB_E_HIP;
    } catch (const AdbError& e) {
        free_world(w); delete W; return e.code;
    } catch (const std::bad_alloc&) {
        free_world(w); delete W; return ADB_E_CAPACITY;
    } catch (...) {
        free_world(w); delete W; return ADB_E_ARG;
    }
    *out = W;
    return ADB_OK;
}

static void free_world(World& w) {
    DBuf<int>* ib[] = {&w.voff, &w.foff, &w.eoff, &w.vbody, &w.fbody, &w.ebody, &w.face, &w.edge, &w.counters,
                       &w.aval, &w.aval_s, &w.blk_of, &w.row_ptr, &w.col_idx, &w.diag_blk, &w.vg_start,
                       &w.pair_start, &w.pair_lo, &w.pair_hi, &w.pair_up, &w.pair_low, &w.tval, &w.tval_s,
                       &w.up_start, &w.low_start, &w.rowcnt,
                       &w.ctype, &w.ccls, &w.clist, &w.clist3, &w.ccount, &w.chunk_n, &w.chunk_off, &w.chunk_blk,
                       &w.pair_pack};
    for (auto* b : ib) b->free_();
    DBuf<double>* db[] = {&w.xbar, &w.mass, &w.Ic, &w.q, &w.q0, &w.qdot, &w.dq, &w.qtrial, &w.X, &w.Xs, &w.X1, &w.bbox, &w.bsph,
                          &w.crec, &w.cdist, &w.chunk_part, &w.ctk, &w.clam, &w.vgK, &w.vglam, &w.bodyH, &w.bodyG, &w.Hval, &w.grad,
                          &w.Ginv, &w.pcg_r, &w.pcg_z, &w.pcg_p, &w.pcg_p2, &w.pcg_Ap, &w.partial, &w.scal, &w.epart};
    for (auto* b : db) b->free_();
    w.pairs.free_(); w.keys_a.free_(); w.keys_b.free_(); w.hcnt.free_(); w.hend.free_(); w.sort_tmp.free_(); w.toi_bits.free_(); w.vg_keys.free_();
    w.bcnt.free_(); w.sh_cnt.free_(); w.sh_start.free_(); w.sh_ent.free_(); w.sh_rng.free_(); w.sh_total.free_();
    w.akeys.free_(); w.akeys_s.free_(); w.tkeys.free_(); w.tkeys_s.free_();
    for (ContactSet* c : {&w.cs, &w.cs_new, &w.cand}) { c->pt.free_(); c->ee.free_(); c->vg.free_(); }
    w.tcount.free_();
    if (w.h_tcount) (void)hipHostFree(w.h_tcount);
    w.h_tcount = nullptr;
    if (w.h_small) (void)hipHostFree(w.h_small);
    w.h_small = nullptr;
    w.vsel.free_(); w.keepbits.free_();
    w.scene_off.free_(); w.scene_of.free_(); w.swk.free_(); w.swk_s.free_(); w.swi.free_(); w.swi_s.free_();
    w.seg.free_(); w.sred.free_(); w.sfac.free_(); w.sc_goff.free_(); w.pcg_sbb.free_(); w.pcg_brun.free_();
    if (w.h_sred) (void)hipHostFree(w.h_sred);
    w.h_sred = nullptr;
    for (auto& r : w.prof.pending) { w.prof.pool.push_back(r.a); w.prof.pool.push_back(r.b); }
    w.prof.pending.clear();
    for (hipEvent_t e : w.prof.pool) (void)hipEventDestroy(e);
    w.prof.pool.clear();
    if (w.prof.snap) (void)hipHostFree(w.prof.snap);
    w.prof.snap = nullptr;
    if (w.pcg_graph) (void)hipGraphExecDestroy(w.pcg_graph);
    w.pcg_graph = nullptr;
    w.pcg_k.free_();
    w.grp_pack.free_();
    w.grp_moff = nullptr; w.grp_off = w.grp_body = w.glist = w.grp_n = nullptr;
    if (w.h_grp_pack) (void)hipHostFree(w.h_grp_pack);
    w.h_grp_pack = nullptr; w.h_grp_pack_cap = 0;
    if (w.ev_grp) (void)hipEventDestroy(w.ev_grp);
    w.ev_grp = nullptr;
    if (w.h_scal) (void)hipHostFree(w.h_scal);
    w.h_scal = nullptr;
    if (w.h_slots) (void)hipHostFree(w.h_slots);
    w.h_slots = nullptr;
    if (w.h_pairs) (void)hipHostFree(w.h_pairs);
    w.h_pairs = nullptr; w.h_pairs_cap = 0;
    if (w.h_np) (void)hipHostFree(w.h_np);
    w.h_np = nullptr;
    if (w.ev_pairs) (void)hipEventDestroy(w.ev_pairs);
    w.ev_pairs = nullptr;
    for (hipEvent_t& e : w.ev_phase) {
        if (e) (void)hipEventDestroy(e);
        e = nullptr;
    }
    for (hipEvent_t& e : w.pcg_ev) { if (e) (void)hipEventDestroy(e); e = nullptr; }
    for (hipEvent_t& e : w.ev_side) { if (e) (void)hipEventDestroy(e); e = nullptr; }
    if (w.stream2) (void)hipStreamDestroy(w.stream2);
    w.stream2 = nullptr;
    if (w.ev_pk) (void)hipEventDestroy(w.ev_pk);
    w.ev_pk = nullptr;
    if (w.stream3) (void)hipStreamDestroy(w.stream3);
    w.stream3 = nullptr;
    for (hipEvent_t* e : {&w.ev_asm_fork, &w.ev_asm_join}) { if (*e) (void)hipEventDestroy(*e); *e = nullptr; }
    if (w.stream4) (void)hipStreamDestroy(w.stream4);
    w.stream4 = nullptr;
    if (w.stream) (void)hipStreamDestroy(w.stream);
    w.stream = nullptr;
}

extern "C" void adb_world_destroy(adb_world* W) {
    if (!W) return;
    free_world(W->w);
    delete W;
}

extern "C" int adb_step(adb_world* W, adb_step_stats* st) {
    if (!W) return ADB_E_ARG;
    ADB_TRY(W, { return W->w.S > 1 ? step_batched(W->w, st) : step(W->w, st); })
}

extern "C" int adb_world_create_batched(const adb_scene* s, const int* scene_off, int n_scenes, const adb_params* p,
                                        adb_world** out) {
    if (!s || !scene_off || !p || !out || n_scenes <= 0 || scene_off[0] != 0 || scene_off[n_scenes] != s->n_bodies)
        return ADB_E_ARG;
    for (int i = 0; i < n_scenes; i++)
        if (scene_off[i + 1] < scene_off[i]) return ADB_E_ARG;
    if ((p->iaabb & 3) != 3) return ADB_E_ARG;      // contact sets and CCD by iAABB body overlaps
    int rc = adb_world_create(s, p, out);
    if (rc != ADB_OK) return rc;
    World& w = (*out)->w;
    try {
        w.S = n_scenes;
        w.sweep = true;
        w.h_scene_off.assign(scene_off, scene_off + n_scenes + 1);
        std::vector<int> of(w.B);
        for (int sc = 0; sc < n_scenes; sc++)
            for (int b = scene_off[sc]; b < scene_off[sc + 1]; b++) of[b] = sc;
        upload(w, w.scene_off, w.h_scene_off.data(), n_scenes + 1);
        upload(w, w.scene_of, of.data(), w.B);
        w.sred.reserve((size_t)SRED_N * n_scenes);
        w.sfac.reserve(n_scenes);
        w.pcg_sbb.reserve(n_scenes); w.pcg_brun.reserve(1);
        w.sc_goff.reserve(n_scenes + 1);
        wait_stream(w);
    } catch (const HipError& e) {
        adb_world_destroy(*out); *out = nullptr; return ADB_E_HIP;
    } catch (...) {
        adb_world_destroy(*out); *out = nullptr; return ADB_E_ARG;
    }
    return ADB_OK;
}

extern "C" int adb_batch_stats(adb_world* W, int* newton_iters, int* ls_trials, int* status, int* contacts) {
    if (!W) return ADB_E_ARG;
    World& w = W->w;
    const size_t S = w.sc_newton.size();
    for (size_t i = 0; i < S; i++) {
        if (newton_iters) newton_iters[i] = w.sc_newton[i];
        if (ls_trials) ls_trials[i] = w.sc_ls[i];
        if (status) status[i] = w.sc_status[i];
        if (contacts) contacts[i] = w.sc_contacts[i];
    }
    return S == (size_t)w.S ? ADB_OK : ADB_E_ARG;
}

extern "C" int adb_get_state(adb_world* W, double* q, double* qdot) {
    if (!W) return ADB_E_ARG;
    ADB_TRY(W, {
        World& w = W->w;
        const size_t n = 12 * (size_t)w.B * sizeof(double);
        if (q) ADB_HIP(hipMemcpyAsync(q, w.q0.p, n, hipMemcpyDeviceToHost, w.stream));
        if (qdot) ADB_HIP(hipMemcpyAsync(qdot, w.qdot.p, n, hipMemcpyDeviceToHost, w.stream));
        wait_stream(w);
        return ADB_OK;
    })
}

extern "C" int adb_set_state(adb_world* W, const double* q, const double* qdot) {
    if (!W) return ADB_E_ARG;
    ADB_TRY(W, {
        World& w = W->w;
        const size_t n = 12 * (size_t)w.B * sizeof(double);
        if (q) {
            ADB_HIP(hipMemcpyAsync(w.q0.p, q, n, hipMemcpyHostToDevice, w.stream));
            ADB_HIP(hipMemcpyAsync(w.q.p, q, n, hipMemcpyHostToDevice, w.stream));
        }
        if (qdot) ADB_HIP(hipMemcpyAsync(w.qdot.p, qdot, n, hipMemcpyHostToDevice, w.stream));
        wait_stream(w);
        return ADB_OK;
    })
}

extern "C" const char* adb_last_error(adb_world* W) { return W ? W->w.err.c_str() : "null world"; }

namespace adb {
void prof_flush(World& w) {
    if (w.prof.pending.empty()) return;
    wait_stream(w);
    for (auto& r : w.prof.pending) {
        float ms = 0.0f;
        ADB_HIP(hipEventElapsedTime(&ms, r.a, r.b));
        w.prof.ms[r.k] += ms;
        double bytes = r.bytes;
        if (r.snap) for (int i = 0; i < 4; i++) bytes += r.per[i] * (double)r.snap[i];
        w.prof.bytes[r.k] += bytes;
        w.prof.flops[r.k] += r.flops;
        w.prof.count[r.k] += 1;
        w.prof.pool.push_back(r.a);
        w.prof.pool.push_back(r.b);
    }
    w.prof.pending.clear();
    w.prof.snap_used = 0;
}
const int* prof_snapshot(World& w, const int* dev_counts) {
    if (!w.prof.on) return nullptr;
    if (!w.prof.snap) {
        ADB_HIP(hipHostMalloc((void**)&w.prof.snap, 4 * 8192 * sizeof(int), hipHostMallocDefault));
        w.prof.snap_cap = 4 * 8192;
    }
    if (w.prof.snap_used + 4 > w.prof.snap_cap) prof_flush(w);
    int* dst = w.prof.snap + w.prof.snap_used;
    w.prof.snap_used += 4;
    ADB_HIP(hipMemcpyAsync(dst, dev_counts, 4 * sizeof(int), hipMemcpyDeviceToHost, w.stream));
    return dst;
}
} // namespace adb

extern "C" int adb_profile_enable(adb_world* W, int on) {
    if (!W) return ADB_E_ARG;
    ADB_TRY(W, {
        World& w = W->w;
        prof_flush(w);
        w.prof.on = on != 0;
        for (int k = 0; k < PK_COUNT; k++) { w.prof.ms[k] = 0.0; w.prof.bytes[k] = 0.0; w.prof.flops[k] = 0.0; w.prof.count[k] = 0; }
        return ADB_OK;
    })
}

extern "C" int adb_get_tuning(adb_world* W, adb_tuning* t) {
    if (!W || !t) return ADB_E_ARG;
    *t = W->w.tune;
    return ADB_OK;
}

extern "C" int adb_set_tuning(adb_world* W, const adb_tuning* t) {
    if (!W || !t) return ADB_E_ARG;
    if (!(t->narrow_per_pair >= 0.125 && t->narrow_per_pair <= 8.0) || t->ee_slices < 1 || t->ee_slices > 16 ||
        t->pcg_batch < 2 || t->pcg_batch > 64 || (t->pcg_batch & 1)) {
        W->w.err = "tuning out of range";
        return ADB_E_ARG;
    }
    W->w.tune = *t;
    return ADB_OK;
}

extern "C" int adb_profile_flops(adb_world* W, int kernel, double* flops) {
    if (!W || kernel < 0 || kernel >= PK_COUNT || !flops) return ADB_E_ARG;
    ADB_TRY(W, {
        World& w = W->w;
        prof_flush(w);
        *flops = w.prof.flops[kernel];
        return ADB_OK;
    })
}

extern "C" int adb_profile_read(adb_world* W, int kernel, double* ms, long long* launches, double* bytes) {
    if (!W || kernel < 0 || kernel >= PK_COUNT) return ADB_E_ARG;
    ADB_TRY(W, {
        World& w = W->w;
        prof_flush(w);
        if (ms) *ms = w.prof.ms[kernel];
        if (launches) *launches = w.prof.count[kernel];
        if (bytes) *bytes = w.prof.bytes[kernel];
        return ADB_OK;
    })
}

extern "C" int adb_collision_set(adb_world* W, int vtn, const double* dq, int* pt, int pt_cap, int* n_pt, int* ee, int ee_cap,
                                 int* n_ee, int* vg, int vg_cap, int* n_vg) {
    if (!W || (vtn != 1 && vtn != 2) || (vtn == 2 && !dq)) return ADB_E_ARG;
    ADB_TRY(W, {
        World& w = W->w;
        const size_t n = 12 * (size_t)w.B;
        ADB_HIP(hipMemcpyAsync(w.q.p, w.q0.p, n * sizeof(double), hipMemcpyDeviceToDevice, w.stream));
        if (vtn == 2) {
            ADB_HIP(hipMemcpyAsync(w.dq.p, dq, n * sizeof(double), hipMemcpyHostToDevice, w.stream));
            q_axpy(w, w.q.p, w.dq.p, 1.0, w.qtrial.p);
        } else {
            q_axpy(w, w.q.p, nullptr, 0.0, w.qtrial.p);
        }
        transform_positions(w, w.qtrial.p, w.X1.p);
        collision_set(w, w.X1.p, w.cs_new);
        std::vector<int> a, b, c;
        download_sets(w, w.cs_new, a, b, c);
        *n_pt = w.cs_new.n_pt; *n_ee = w.cs_new.n_ee; *n_vg = w.cs_new.n_vg;
        if (pt) std::memcpy(pt, a.data(), sizeof(int) * std::min(a.size(), 4 * (size_t)std::max(pt_cap, 0)));
        if (ee) std::memcpy(ee, b.data(), sizeof(int) * std::min(b.size(), 4 * (size_t)std::max(ee_cap, 0)));
        if (vg) std::memcpy(vg, c.data(), sizeof(int) * std::min(c.size(), 2 * (size_t)std::max(vg_cap, 0)));
        if (*n_pt > pt_cap || *n_ee > ee_cap || *n_vg > vg_cap) return ADB_E_CAPACITY;
        return ADB_OK;
    })
}

extern "C" int adb_step_size_upper_bound(adb_world* W, const double* dq, double* toi) {
    if (!W || !dq || !toi) return ADB_E_ARG;
    ADB_TRY(W, {
        World& w = W->w;
        const size_t n = 12 * (size_t)w.B;
        ADB_HIP(hipMemcpyAsync(w.q.p, w.q0.p, n * sizeof(double), hipMemcpyDeviceToDevice, w.stream));
        ADB_HIP(hipMemcpyAsync(w.dq.p, dq, n * sizeof(double), hipMemcpyHostToDevice, w.stream));
        transform_positions(w, w.q.p, w.X.p);
        q_axpy(w, w.q.p, w.dq.p, 1.0, w.qtrial.p);
        transform_positions(w, w.qtrial.p, w.X1.p);
        int sp = w.stall_pt, sv = w.stall_vg, status = ADB_OK;
        *toi = ccd_toi(w, w.X.p, w.X1.p, &status);
        w.stall_pt = sp; w.stall_vg = sv;
        return ADB_OK;
    })
}

// the assembled gradient and (when hess != null) the dense Hessian to the host
static void read_system(World& w, double* grad, double* hess) {
        const int n = 12 * w.B;
        if (grad) ADB_HIP(hipMemcpyAsync(grad, w.grad.p, n * sizeof(double), hipMemcpyDeviceToHost, w.stream));
        if (hess) {
            std::vector<int> rp(w.B + 1), ci(w.nnzb);
            std::vector<double> hv(144 * (size_t)w.nnzb);
            ADB_HIP(hipMemcpyAsync(rp.data(), w.row_ptr.p, rp.size() * sizeof(int), hipMemcpyDeviceToHost, w.stream));
            ADB_HIP(hipMemcpyAsync(ci.data(), w.col_idx.p, ci.size() * sizeof(int), hipMemcpyDeviceToHost, w.stream));
            ADB_HIP(hipMemcpyAsync(hv.data(), w.Hval.p, hv.size() * sizeof(double), hipMemcpyDeviceToHost, w.stream));
            wait_stream(w);
            std::memset(hess, 0, sizeof(double) * (size_t)n * n);
            for (int r = 0; r < w.B; r++)
                for (int k = rp[r]; k < rp[r + 1]; k++)
                    for (int c = 0; c < 12; c++)
                        for (int i = 0; i < 12; i++)
                            hess[(size_t)(12 * ci[k] + c) * n + 12 * r + i] = hv[144 * (size_t)k + 12 * c + i];
        }
        wait_stream(w);
}

// The contact set, contact terms and assembly at the Newton iterate q of a
// step that started at the world's state q0 (uploaded to w.q; w.X / w.Xs
// hold the positions at q and q0).
static void terms_at(World& w, const double* q_host) {
    const size_t n = 12 * (size_t)w.B;
    ADB_HIP(hipMemcpyAsync(w.q.p, q_host, n * sizeof(double), hipMemcpyHostToDevice, w.stream));
    transform_positions(w, w.q.p, w.X.p);
    transform_positions(w, w.q0.p, w.Xs.p);
    collision_set(w, w.X.p, w.cs);
    contact_terms(w);
    assemble_system(w);
}

extern "C" int adb_assemble(adb_world* W, double* grad, double* hess) {
    if (!W) return ADB_E_ARG;
    ADB_TRY(W, {
        World& w = W->w;
        assemble_now(w);
        read_system(w, grad, hess);
        return ADB_OK;
    })
}

extern "C" int adb_contact_terms(adb_world* W, const double* q, double* grad, double* hess) {
    if (!W || !q) return ADB_E_ARG;
    ADB_TRY(W, {
        World& w = W->w;
        terms_at(w, q);
        read_system(w, grad, hess);
        return ADB_OK;
    })
}

extern "C" int adb_line_search(adb_world* W, const double* q, const double* dq, double* alpha, double* E0, double* E1,
                               int* trials) {
    if (!W || !q || !dq) return ADB_E_ARG;
    ADB_TRY(W, {
        World& w = W->w;
        const int n = 12 * w.B;
        terms_at(w, q);
        ADB_HIP(hipMemcpyAsync(w.dq.p, dq, n * sizeof(double), hipMemcpyHostToDevice, w.stream));
        double sl[1];
        dot_async(w, w.grad.p, w.grad.p, n, 0);
        read_slots(w, 1, sl);
        double a = 1.0, e0 = 0.0, e1 = 0.0;
        int t = 0;
        w.e0_valid = false;
        const int rc = line_search(w, std::sqrt(sl[0]), &a, &t, &e0, &e1);
        w.e0_valid = false;
        if (alpha) *alpha = a;
        if (E0) *E0 = e0;
        if (E1) *E1 = e1;
        if (trials) *trials = t;
        return rc;
    })
}

extern "C" int adb_adjoint_solve(adb_world* W, const double* rhs, double* x, int* iters) {
    if (!W || !rhs || !x) return ADB_E_ARG;
    World& w = W->w;
    // the lagged friction terms depend on the step's start through Tk and
    // lambda; the adjoint covers the frictionless step
    if (w.P.mu != 0.0 && (w.P.pt_fric || w.P.ee_fric || w.P.vg_fric)) {
        w.err = "adb_adjoint_solve differentiates frictionless steps only (mu = 0)";
        return ADB_E_ARG;
    }
    const int psd = w.P.psd;
    w.P.psd = 0;                      // exact second derivatives, not their PSD projections
    int rc = ADB_OK;
    try {
        const int n = 12 * w.B;
        assemble_now(w);
        // pcg_solve returns -H^{-1} grad: solve with grad = -rhs
        ADB_HIP(hipMemcpyAsync(w.grad.p, rhs, n * sizeof(double), hipMemcpyHostToDevice, w.stream));
        scale(w, w.grad.p, n, -1.0);
        int it = 0;
        rc = pcg_solve(w, w.dq.p, &it);
        if (iters) *iters = it;
        ADB_HIP(hipMemcpyAsync(x, w.dq.p, n * sizeof(double), hipMemcpyDeviceToHost, w.stream));
        wait_stream(w);
        if (rc == ADB_E_NOT_CONVERGED) w.err = "adjoint PCG reached pcg_max_iter";
        if (rc == ADB_E_SOLVER) w.err = "adjoint PCG breakdown: the Hessian at the state is not positive definite";
    } catch (const HipError& e) { w.err = e.what; rc = ADB_E_HIP; }
    catch (const AdbError& e) { w.err = e.what; rc = e.code; }
    catch (...) { w.err = "unknown exception"; rc = ADB_E_ARG; }
    w.P.psd = psd;
    return rc;
}

extern "C" int adb_newton_direction(adb_world* W, double* dq) {
    if (!W || !dq) return ADB_E_ARG;
    ADB_TRY(W, {
        World& w = W->w;
        assemble_now(w);
        int it = 0;
        int rc = pcg_solve(w, w.dq.p, &it);
        ADB_HIP(hipMemcpyAsync(dq, w.dq.p, 12 * (size_t)w.B * sizeof(double), hipMemcpyDeviceToHost, w.stream));
        wait_stream(w);
        return rc;
    })
}
