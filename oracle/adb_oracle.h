/*
 * adb_oracle.h - CPU FP64 oracle for the ADB Newton contact step.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a from-scratch C++ restatement of the
 * reference's CPU path (Luke-Skycrawler/adb, files under model/), used by tests/,
 * __graft_entry__.smoke() and the cpu_baseline leg of bench.py as the checker.
 * The product (adb_amd/, include/adb_capi.h) never links, loads or calls it.
 *
 * Functions whose results depend on third-party code the reference does not
 * vendor (ipc-toolkit v1.2.1 distance types/derivatives, closest points,
 * tangent bases, mollifier; Eigen's SelfAdjointEigenSolver; MKL PardisoLLT)
 * are restated from the published algorithms and pinned by mathematics
 * (closed forms, autograd, finite differences, brute force, eigh) in tests/.
 */
#ifndef ADB_ORACLE_H
#define ADB_ORACLE_H
#include "../include/adb_capi.h"

#ifdef __cplusplus
extern "C" {
#endif

/* world */
int  oracle_world_create(const adb_scene* scene, const adb_params* params, void** out);
void oracle_world_destroy(void* w);
int  oracle_step(void* w, adb_step_stats* st);
int  oracle_get_state(void* w, double* q, double* qdot);
int  oracle_set_state(void* w, const double* q, const double* qdot);
/* method: 0 = brute force over all primitive pairs, 1 = iAABB body-overlap culling,
 * 2 = spatial hashing (model/spatial_hashing.cpp, model/col_set.cpp, model/upper_bound.cpp) */
int  oracle_collision_set(void* w, int vtn, const double* dq, int method,
                          int* pt, int pt_cap, int* n_pt, int* ee, int ee_cap, int* n_ee,
                          int* vg, int vg_cap, int* n_vg);
int  oracle_step_size_upper_bound(void* w, const double* dq, int method, double* toi);
int  oracle_assemble(void* w, double* grad, double* hess_dense);
int  oracle_newton_direction(void* w, double* dq);
/* incremental potential without friction at q: inertia + orthogonality + barriers */
int  oracle_energy(void* w, const double* q, double* E);
int  oracle_assemble_at(void* w, const double* q, double* grad, double* hess);
int  oracle_adjoint_solve(void* w, const double* rhs, double* x);
int  oracle_friction_energy(void* w, const double* q, double* E);
int  oracle_line_search(void* w, const double* q, const double* dq, double* alpha, double* E0, double* E1, int* trials);
int  oracle_set_threads(int n);

/* kernel level */
void   oracle_pt_distance(const double pt[12], int* type, double* d);
void   oracle_ee_distance(const double ee[12], int* type, double* d);
void   oracle_pt_distance_derivs(const double pt[12], int type, double g[12], double H[144]);
void   oracle_ee_distance_derivs(const double ee[12], int type, double g[12], double H[144]);
void   oracle_ee_mollifier(const double ee[12], double eps_x, double* m, double g[12], double H[144]);
void   oracle_barrier(double d, double d_hat, double kappa, double out[3]);
void   oracle_sym_eig(int n, const double* A, double* evals, double* evecs);
void   oracle_psd_project(int n, const double* A, double* out);
int    oracle_ipc_hess_pt_12x12(const adb_kernel_config* c, const double pt[12], int type,
                                double dist, double H[144], double g[12]);
int    oracle_ipc_hess_ee_12x12(const adb_kernel_config* c, const double ee[12], int type,
                                double dist, double H[144], double g[12], double* moll);
double oracle_pt_collision_time(const double x[24]);
double oracle_ee_collision_time(const double x[24]);
void   oracle_det_polynomial(const double a[9], const double b[9], double out[4]);
int    oracle_cubic_roots(const double coef[4], double x0, double x1, double roots[3]);
void   oracle_friction(const double uk[2], double lam, const double* T, int n, double mu,
                       double eps, double max_uk, double* g, double* H);
void   oracle_ortho(const double q[12], double kappa, double* E, double g[12], double H[144]);
void   oracle_pt_closest(const double pt[12], int type, double lams[3], double Pk[6]);
void   oracle_ee_closest(const double ee[12], int type, double lams[2], double Pk[6]);
double oracle_f0(double x, double eps);
double oracle_f1_over_x(double x, double eps);

#ifdef __cplusplus
}
#endif
#endif
