/*
 * adb_capi.h - C ABI of the MI355X-native ADB Newton contact step.
 *
 * This header is the drop-in boundary for the reference's per-timestep contact
 * step (`ABD::implicit_euler` and everything under it, model/abd.cpp:30-140) and
 * for the kernel-level functions the reference exposes through its `abdtk`
 * Python module (python/bindings.cpp:15-40,78-82).  Every entry point takes
 * plain pointers and sizes; no C++ or torch types cross the boundary.
 *
 * Conventions
 *   - Body generalised coordinates are 12 doubles per body in the reference's
 *     order q = [p, a1, a2, a3] (model/affine_body.h:24, model/mutils.cpp:44-51);
 *     a world vertex is x = [a1 a2 a3] * xbar + p.
 *   - 12x12 matrices are column-major (Eigen's default storage, which the
 *     reference's CSC values use: model/assemble_matrices.cpp:8-15).
 *   - Distance arguments are SQUARED distances (model/barrier.cpp:7-8).
 *   - Enum values equal ipc-toolkit's PointTriangleDistanceType and
 *     EdgeEdgeDistanceType (python/bindings.cpp:54-76).
 *   - No entry point ever exits the process: errors are returned as status codes
 *     (the reference calls exit(1): model/abd.cpp:288-291, model/iaabb.cpp:305-311,533-537).
 */
#ifndef ADB_CAPI_H
#define ADB_CAPI_H

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes -------------------------------------------------------- */
#define ADB_OK               0
#define ADB_E_ARG           -1  /* bad argument / size                          */
#define ADB_E_HIP           -2  /* HIP runtime error                            */
#define ADB_E_NAN           -3  /* NaN in the Newton direction (abd.cpp:288)    */
#define ADB_E_TOI_STALL     -4  /* toi < 1e-6 twice in a row (iaabb.cpp:305)    */
#define ADB_E_CAPACITY      -5  /* a fixed-capacity buffer overflowed           */
#define ADB_E_NOT_CONVERGED -6  /* Newton hit max_iter without ||dq|| < tol     */
#define ADB_E_NO_DEVICE     -7  /* no usable gfx950 device                      */
#define ADB_E_SOLVER        -8  /* PCG broke down (system not positive definite) */

/* ---- distance-type enums (ipc-toolkit numbering) ------------------------- */
enum adb_pt_type { ADB_P_T0 = 0, ADB_P_T1, ADB_P_T2, ADB_P_E0, ADB_P_E1, ADB_P_E2, ADB_P_T, ADB_PT_AUTO };
enum adb_ee_type { ADB_EA0_EB0 = 0, ADB_EA0_EB1, ADB_EA1_EB0, ADB_EA1_EB1, ADB_EA_EB0, ADB_EA_EB1,
                   ADB_EA0_EB, ADB_EA1_EB, ADB_EA_EB, ADB_EE_AUTO };

/* ---- scene ----------------------------------------------------------------
 * Replaces the reference's std::vector<std::unique_ptr<AffineBody>>
 * (model/affine_body.h:11-108): rest vertices, triangles, edges, mass, Ic,
 * q0 and dqdt per body.  Offsets are prefix sums (n_bodies + 1 entries).
 * Face and edge indices are LOCAL to their body (as AffineBody::indices and
 * AffineBody::edges).  mass < 0 marks a static body (model/mutils.cpp:55,71). */
typedef struct adb_scene {
    int n_bodies;
    const int*    vert_offset;   /* [n_bodies+1]                      */
    const int*    face_offset;   /* [n_bodies+1]                      */
    const int*    edge_offset;   /* [n_bodies+1]                      */
    const double* rest_verts;    /* [V*3]  xbar                        */
    const int*    faces;         /* [F*3]  local vertex ids            */
    const int*    edges;         /* [E*2]  local vertex ids            */
    const double* mass;          /* [n_bodies]                        */
    const double* Ic;            /* [n_bodies]                        */
    const double* q0;            /* [n_bodies*12]                     */
    const double* dqdt;          /* [n_bodies*12]                     */
} adb_scene;

/* ---- parameters -----------------------------------------------------------
 * Every constant the reference reads from config.json, model/settings.h or
 * hard-codes (model/barrier.cpp:7-8, model/othogonal_energy.cpp:8,
 * model/geometry.cpp:60) is explicit here; nothing is global or static. */
typedef struct adb_params {
    double dt;            /* config.json:5                                        */
    double d_hat;         /* squared barrier threshold, barrier::d_hat            */
    double kappa;         /* barrier stiffness, config.json:55                    */
    double kappa_ortho;   /* orthogonality stiffness, othogonal_energy.cpp:8      */
    double mu;            /* friction coefficient, config.json:21                 */
    double evh;           /* friction eps = dt * evh, config.json:51              */
    double eps_x;         /* EE mollifier scale, config.json:29                   */
    double tol;           /* Newton tolerance on ||dq||, config.json:57           */
    double c1;            /* Armijo constant, config.json:60                      */
    double backoff;       /* CCD backoff factor, config.json:23                   */
    double max_uk;        /* friction lower branch threshold, config.json:61      */
    double gravity[3];    /* settings.h:8                                         */
    double ground_y;      /* ground plane height, geometry.cpp:60 (-0.5)          */
    int    max_iter;      /* Newton iterations per step                           */
    int    max_ls_iter;   /* line-search trials per Newton iteration              */
    int    enable_pt, enable_ee, enable_ground;
    int    pt_fric, ee_fric, vg_fric;
    int    psd;           /* PSD projection of contact Hessians, config.json:30   */
    int    line_search;   /* config.json:14                                       */
    int    upper_bound;   /* CCD step bound, config.json:20 ("uppper_bound")      */
    int    broad_phase_fp32; /* 1: FP32 conservative AABBs + exact FP64 filter    */
    double pcg_rtol;      /* PCG relative residual target                         */
    int    pcg_max_iter;
    /* Culling method, config.json:50 "iaabb" (model/abd.cpp:48-51,159-173,
     * model/line_search.cpp:88-109): bit 0 set = contact sets by iAABB body
     * overlaps, clear = spatial hashing (model/col_set.cpp:20-339); bit 1 set =
     * CCD step bound by iAABB, clear = spatial hashing (model/upper_bound.cpp:16-228).
     * Both methods give the same sets and bounds; default 3.                  */
    int    iaabb;
    int    sh_bits;       /* bits per cell coordinate in the hash, config.json:44 */
    double sh_min_xyz;    /* grid origin, config.json:40                          */
    double sh_dx;         /* cell size, config.json:42                            */
} adb_params;

/* Fills p with the reference's defaults (config.json, model/settings.h). */
void adb_params_default(adb_params* p);

typedef struct adb_step_stats {
    int    newton_iters;
    int    ls_trials;          /* total line-search trials in the step            */
    int    pcg_iters;          /* total PCG iterations in the step                */
    int    n_pt, n_ee, n_vg;   /* contact counts at the end of the step           */
    long long contacts_processed; /* sum over Newton iterations of active contacts */
    double toi_min;            /* smallest CCD step bound seen                    */
    double t_ipc, t_solver, t_ccd, t_line_search; /* seconds, the reference's 4
                                   phase buckets (model/abd.h:68, model/abd.cpp:110-121);
                                   t_ipc is the device time of the contact terms and
                                   the assembly (HIP events: the solve is queued
                                   behind them without a host wait), t_solver the
                                   rest of the host time to the solve's end      */
    double t_total;
    /* Newton iterations of the step whose linear solve did not finish: PCG hit
     * pcg_max_iter before ||r|| <= pcg_rtol ||b||, or broke down (r.z <= 0 or
     * p.Hp <= 0: the assembled system is not positive definite).  The step goes
     * on with the last PCG iterate, as the line search still guards descent;
     * adb_last_error names the last such event.  The reference's direct solve
     * has no counterpart (model/abd.cpp:268-276). */
    int    pcg_not_converged;
    int    pcg_breakdown;
} adb_step_stats;

typedef struct adb_world adb_world;

/* ---- world / step API ----------------------------------------------------
 * Replaces ABD::ABD (model/abd.h:13-14) and ABD::implicit_euler
 * (model/abd.cpp:30-140).  The scene is copied; the caller's arrays are not
 * retained.  One world per (thread, device). */
int  adb_world_create(const adb_scene* scene, const adb_params* params, adb_world** out);
void adb_world_destroy(adb_world* w);
int  adb_step(adb_world* w, adb_step_stats* stats);
int  adb_get_state(adb_world* w, double* q, double* qdot);               /* [B*12] each */
int  adb_set_state(adb_world* w, const double* q, const double* qdot);
const char* adb_last_error(adb_world* w);

/* ---- batched independent scenes ------------------------------------------
 * Many independent scenes (each with its own bodies, contacts, Newton loop,
 * CCD bound, line search and convergence test, as if each were stepped by its
 * own ABD::implicit_euler, model/abd.cpp:30-140) in one world: every stage
 * runs once per Newton iteration over the bodies and contacts of all scenes
 * still iterating, with per-scene scalars on the device.  `scene` holds the
 * bodies of all scenes, scene by scene; scene_offsets[s] .. scene_offsets[s+1]
 * are the bodies of scene s (n_scenes + 1 entries, scene_offsets[0] = 0,
 * nondecreasing, scene_offsets[n_scenes] = n_bodies).  Bodies of different
 * scenes never interact, whatever their positions.  Contact sets by iAABB
 * only (params->iaabb must be 3).  adb_step steps every scene once; the step's
 * statistics are summed over the scenes (newton_iters: the largest per-scene
 * count) and adb_batch_stats gives them per scene. */
int adb_world_create_batched(const adb_scene* scene, const int* scene_offsets, int n_scenes,
                             const adb_params* params, adb_world** out);
/* Per-scene results of the last adb_step (arrays of n_scenes entries, any may
 * be NULL): Newton iterations, line-search trials, status (ADB_OK or an
 * ADB_E_* code for that scene), and the scene's contacts at the end of the
 * step (PT + EE + VG). */
int adb_batch_stats(adb_world* w, int* newton_iters, int* ls_trials, int* status, int* contacts);

/* ---- stage entry points (parity tests) -----------------------------------
 * adb_collision_set: replaces IAABB::iaabb_brute_force(vtn = 1 or 2)
 * (model/iaabb.cpp:545-568).  vtn = 1 culls at the current q, vtn = 2 at
 * q + dq.  Outputs are sorted canonically: PT by (I, v, J, f), EE by
 * (I, ei, J, ej) with I < J, VG by (I, v).  Capacities are in entries; the
 * true counts are always written, and ADB_E_CAPACITY is returned when one
 * exceeds its capacity. */
int adb_collision_set(adb_world* w, int vtn, const double* dq,
                      int* pt, int pt_cap, int* n_pt,      /* int32[4] per entry */
                      int* ee, int ee_cap, int* n_ee,      /* int32[4] per entry */
                      int* vg, int vg_cap, int* n_vg);     /* int32[2] per entry */

/* Replaces IAABB::iaabb_brute_force(vtn = 3) / step_size_upper_bound
 * (model/iaabb.cpp:545-568, model/upper_bound.cpp:16): the CCD step bound
 * along q -> q + dq (dq: [B*12]). */
int adb_step_size_upper_bound(adb_world* w, const double* dq, double* toi);

/* Contact gradient/Hessian assembly at the current q, with friction lagged on
 * the step's start positions: replaces ABD::ipc + per-body inertia terms +
 * ABD::solve's diagonal assembly (model/abd.cpp:68-74,248-260,305-404).
 * Writes the assembled global gradient [B*12] and the dense 12B x 12B
 * Hessian (column-major) when hess_dense != NULL (small scenes only). */
int adb_assemble(adb_world* w, double* grad, double* hess_dense);

/* ABD::ipc plus the per-body inertia terms at a Newton iterate q ([B*12])
 * of a step that starts at the world's state (model/abd.cpp:68-90,305-404):
 * the contact set at q, the barrier terms, and friction lagged from the
 * contacts at q against the step-start positions.  Writes the gradient and,
 * when hess_dense != NULL, the dense Hessian as adb_assemble. */
int adb_contact_terms(adb_world* w, const double* q, double* grad, double* hess_dense);

/* ABD::line_search (model/abd.cpp:198-227, model/line_search.cpp:39-136) from
 * the iterate q along dq (both [B*12], dq already scaled by the CCD bound):
 * contact terms at q as adb_contact_terms, then alpha = 1, 1/2, 1/4, ... until
 * E(q + alpha dq) <= E(q) + c1 alpha dq.g, each trial with the contact set at
 * its own point.  Returns the accepted alpha, the energies E(q) and
 * E(q + alpha dq), and the number of trials; ADB_E_NOT_CONVERGED when
 * max_ls_iter trials fail. */
int adb_line_search(adb_world* w, const double* q, const double* dq, double* alpha, double* E0, double* E1,
                    int* trials);

/* Newton direction dq = -H^{-1} g at the current q (ABD::solve,
 * model/abd.cpp:229-303) using PCG preconditioned by contact-group inverses. */
int adb_newton_direction(adb_world* w, double* dq);

/* Adjoint solve of one converged step (implicit differentiation), for the
 * differentiable step: x = H^{-1} rhs, with H the exact (not PSD-projected)
 * Hessian of the step's incremental potential at the world's current state,
 * i.e. the end of the last adb_step: M + dt^2 Hess(E_ortho) + Hess(barrier)
 * (model/mutils.cpp:53-76, model/core_ipc.cpp:21-83).  The step's optimality
 * condition grad E(q*; q_tilde) = 0 with q_tilde = q0 + dt qdot + dt^2 g
 * (AffineBody::q_tile, model/mutils.cpp:245-252) gives
 * d(q*) / d(q_tilde) = H^{-1} M.  Frictionless worlds only (ADB_E_ARG when mu != 0
 * and friction is on); solved by the step's PCG to pcg_rtol; iters may be
 * NULL.  No counterpart in the reference, which has no adjoint. */
int adb_adjoint_solve(adb_world* w, const double* rhs, double* x, int* iters);

/* ---- kernel-level API (abdtk) ---------------------------------------------
 * Batched device versions: all pointers are DEVICE pointers, n entries,
 * launched on `stream` (a hipStream_t, 0 = default).  Constants follow the
 * reference's frozen abdtk configuration (barrier::d_hat, barrier::kappa,
 * globals.psd, globals.eps_x) unless set through adb_kernel_config. */
typedef struct adb_kernel_config {
    double d_hat, kappa, eps_x;
    int    psd;
} adb_kernel_config;
void adb_kernel_config_default(adb_kernel_config* c);

/* IPC::ipc_hess_pt_12x12 (model/core_ipc.cpp:21-46): pt = [p, t0, t1, t2]
 * (12 doubles per entry), type per entry, squared distance per entry ->
 * H (144, column-major) and g (12) per entry. */
int adb_ipc_hess_pt_12x12_batched(const adb_kernel_config* cfg, int n,
                                  const double* pt, const int* type, const double* dist,
                                  double* H, double* g, void* stream);
/* IPC::ipc_hess_ee_12x12 (model/core_ipc.cpp:47-83) -> H, g, mollifier. */
int adb_ipc_hess_ee_12x12_batched(const adb_kernel_config* cfg, int n,
                                  const double* ee, const int* type, const double* dist,
                                  double* H, double* g, double* mollifier, void* stream);
/* pt_collision_time / ee_collision_time (model/collision.cpp:299-348):
 * x = 8 points (24 doubles: 4 at t0 then 4 at t1) per entry -> toi. */
int adb_pt_collision_time_batched(int n, const double* x, double* toi, void* stream);
int adb_ee_collision_time_batched(int n, const double* x, double* toi, void* stream);

/* Single-entry host-pointer conveniences with the abdtk signatures
 * (python/bindings.cpp:15-40).  They run the same device kernels. */
int adb_ipc_hess_pt_12x12(const adb_kernel_config* cfg, const double pt[12], int pt_type,
                          double dist, double H[144], double g[12]);
int adb_ipc_hess_ee_12x12(const adb_kernel_config* cfg, const double ee[12], int ee_type,
                          double dist, double H[144], double g[12], double* mollifier);
double adb_pt_collision_time(const double x[24]);
double adb_ee_collision_time(const double x[24]);

/* No counterpart in the reference: a variant of adb_pt_collision_time /
 * adb_ee_collision_time for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  is_ee selects the edge-edge form; x holds n entries of 8
 * points (24 doubles, 4 at t0 then 4 at t1) in HOST memory and toi receives n
 * times of impact in host memory.  It allocates nothing: the inputs and
 * outputs are staged through a fixed device array of the library (at most
 * 4096 entries per call, about 800 KB per device, held until the process
 * ends).  n > 4096 returns ADB_E_CAPACITY, bad arguments ADB_E_ARG.  Calls in
 * one process are serialized by a lock; the array is shared by every call on a
 * device, so callers that bypass the lock (another copy of the library in the
 * same process) would race on it. */
int adb_collision_time_noalloc(int is_ee, int n, const double* x, double* toi);

/* Distance type + squared distance (ipc-toolkit point_triangle_distance_type /
 * point_triangle_distance and the edge-edge pair, called at
 * model/geometry.cpp:101-112 and model/abd.cpp:353-354), batched on device. */
int adb_pt_distance_batched(int n, const double* pt, int* type, double* dist, void* stream);
int adb_ee_distance_batched(int n, const double* ee, int* type, double* dist, void* stream);

/* Library build/version probe: returns the number of gfx950 devices visible,
 * or a negative status.  Does not allocate device memory. */
int adb_device_count(void);

/* ---- measurement hooks -----------------------------------------------------
 * No counterpart in the reference (its timers are the four std::chrono phase
 * buckets of model/abd.h:68, model/abd.cpp:110-121, reported in adb_step_stats).  Per-kernel
 * HIP-event timing on the world's own stream, for roofline figures:
 * adb_profile_enable(w, 1) clears and starts, adb_profile_read returns the
 * accumulated time, launch count and algorithmic bytes of one kernel. */
enum adb_prof_kernel {
    ADB_PK_TRANSFORM = 0, ADB_PK_BODY_BOX, ADB_PK_BODY_PAIRS, ADB_PK_NARROW_SET, ADB_PK_VG_SCAN, ADB_PK_SORT,
    ADB_PK_NARROW_CCD, ADB_PK_VG_TOI, ADB_PK_CONTACT, ADB_PK_VG_CONTACT, ADB_PK_INERTIA, ADB_PK_ASM_REDUCE,
    ADB_PK_BLOCK_CHOL, ADB_PK_SPMV, ADB_PK_PCG_UPDATE, ADB_PK_PCG_DIR, ADB_PK_ENERGY,
    ADB_PK_CONTACT_CLASSIFY, ADB_PK_CONTACT_GROUP, ADB_PK_HASH_BUILD, ADB_PK_HASH_QUERY, ADB_PK_CONTACT_PK, ADB_PK_COUNT
};
int adb_profile_enable(adb_world* w, int on);
int adb_profile_read(adb_world* w, int kernel, double* total_ms, long long* launches, double* bytes);
/* ---- launch tuning ----------------------------------------------------------
 * No counterpart in the reference (its only knob is the BVH threshold
 * `thres`, config.json:68).  Launch shapes of a world that change how fast a
 * step runs but never its results (contact sets are sorted, sums are taken in
 * a fixed order whatever the grid): workgroups of the narrow phase per
 * overlapping body pair and pass (narrow_per_pair, 0.125 .. 8), workgroups per
 * body pair of its edge-edge pass (ee_slices, 1 .. 16), PCG iterations per
 * queued batch between host checks (pcg_batch, even, 2 .. 64).  The
 * autotuner (adb_amd/autotune.py) times candidates on a workload's own
 * states and caches the choice per workload. */
typedef struct adb_tuning {
    double narrow_per_pair;
    int    ee_slices;
    int    pcg_batch;
} adb_tuning;
int adb_get_tuning(adb_world* w, adb_tuning* t);
int adb_set_tuning(adb_world* w, const adb_tuning* t);   /* ADB_E_ARG outside the ranges */

/* Counted FP64 operations of the kernel's launches since adb_profile_enable,
 * for the FP64-VALU-bound kernels whose operation count the host knows
 * exactly (the preconditioner's group inverses); 0 for the others. */
int adb_profile_flops(adb_world* w, int kernel, double* flops);

#ifdef __cplusplus
}
#endif
#endif /* ADB_CAPI_H */
