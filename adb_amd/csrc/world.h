// world.h - host-side state of one ADB world on one gfx950 device, shared by
// the translation units of libadb.so.  Not part of the public ABI.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>
#include <functional>
#include <string>
#include <vector>
#include "../../include/adb_capi.h"

namespace adb {

#define ADB_HIP(call)                                                                    \
    do {                                                                                 \
        hipError_t e_ = (call);                                                          \
        if (e_ != hipSuccess) throw HipError(e_, #call, __FILE__, __LINE__);             \
    } while (0)

struct HipError {
    hipError_t e; std::string what;
    HipError(hipError_t e_, const char* call, const char* file, int line) : e(e_) {
        what = std::string(hipGetErrorString(e_)) + " in " + call + " at " + file + ":" + std::to_string(line);
    }
};
struct AdbError {
    int code; std::string what;
    AdbError(int c, std::string w) : code(c), what(std::move(w)) {}
};

// Grow-only device buffer.  A regrowth frees the old buffer, which waits for
// the whole device, so each takes half again what is asked for: counts that
// grow step by step (a settling pile) regrow rarely.
template <class T> struct DBuf {
    T* p = nullptr; size_t cap = 0;
    void reserve(size_t n) {
        if (n <= cap) return;
        if (p) ADB_HIP(hipFree(p));
        size_t c = n + n / 2 + 64;
        ADB_HIP(hipMalloc((void**)&p, c * sizeof(T)));
        cap = c;
    }
    void free_() { if (p) (void)hipFree(p); p = nullptr; cap = 0; }
};

// Compact body-space record of one PT/EE contact (doubles).  With D the
// dimension of the contact's translation-free point space (3, 6 or 9; 0 when
// the contact is inactive; R[R_D] = D + 100 (dynamic(I) + 2 dynamic(J))),
// n = D / 3 Helmert rows, sides s = 0 (body I) and
// 1 (body J), and block coordinates (alpha, r) = (affine column, component):
//   block(s1, s2)[(alpha, r), (beta, c)] =
//       sum_{a, b < n} Z[a][s1][alpha] Z[b][s2][beta] PK[(3a + r) * D + 3b + c]
//     + G[s1][alpha] G[s2][beta] F[3r + c]
//   gradient(s)[(alpha, r)] = GRAD[12 s + 3 alpha + r]
// PK is the PSD-projected barrier Hessian in Helmert coordinates, F the
// friction matrix and Z, G the per-side weights of the affine map.
constexpr int REC = 160;
// Lagged friction basis of one PT/EE contact: the weights gm_k of its four
// points and the two tangent directions c0, c1, so that
// Tk[r][3k + i] = gm_k c_r[i] (products formed where they are used).
constexpr int TKN = 10;
constexpr int R_D = 0, R_F = 1, R_G = 10, R_Z = 18, R_GRAD = 42, R_PK = 66;
// Reduction result slots (8) + per-workgroup partials (at most 1024 workgroups):
// reserved once at world creation, so no reduction queued behind another can
// see its slots reallocated.
constexpr int EPART_NSLOT = 8, EPART_MAXG = 1024, EPART_SLOTS = EPART_NSLOT + 3 * EPART_MAXG;

// Per-kernel HIP-event timing on the world's stream (off unless enabled).
enum ProfKernel {
    PK_TRANSFORM = 0, PK_BODY_BOX, PK_BODY_PAIRS, PK_NARROW_SET, PK_VG_SCAN, PK_SORT, PK_NARROW_CCD, PK_VG_TOI,
    PK_CONTACT, PK_VG_CONTACT, PK_INERTIA, PK_ASM_REDUCE, PK_BLOCK_CHOL, PK_SPMV, PK_PCG_UPDATE, PK_PCG_DIR,
    PK_ENERGY, PK_CONTACT_CLASSIFY, PK_CONTACT_GROUP, PK_HASH_BUILD, PK_HASH_QUERY, PK_CONTACT_PK, PK_COUNT
};
struct Prof {
    bool on = false;
    // bytes of a launch = bytes + sum_i per[i] * snap[i], where snap points to
    // a pinned host copy of device counters taken on the stream before the launch
    struct Rec { int k; hipEvent_t a, b; double bytes; const int* snap; double per[4]; double flops; };
    std::vector<Rec> pending;
    std::vector<hipEvent_t> pool;
    int* snap = nullptr;             // pinned host ring of counter snapshots
    int snap_cap = 0, snap_used = 0;
    double ms[PK_COUNT] = {0};
    double bytes[PK_COUNT] = {0};
    double flops[PK_COUNT] = {0};     // counted FP64 operations (FP64-VALU-bound kernels)
    long long count[PK_COUNT] = {0};
};

// Contact-set buffers (sorted canonical lists).
struct ContactSet {
    DBuf<int4> pt, ee;      // (I, v, J, f), (I, ei, J, ej)
    DBuf<int2> vg;          // (I, v)
    int n_pt = 0, n_ee = 0, n_vg = 0;
    int np_bound = -1;      // bound of the body pairs with PT / EE contacts (the culling pass's overlapping pairs), or -1
    long long serial = 0;   // distinct per built set (World::cs_serial)
};

struct World {
    adb_params P;
    int B = 0, V = 0, F = 0, E = 0;
    int max_vpb = 0, max_fpb = 0, max_epb = 0;   // max primitives per body
    hipStream_t stream = nullptr;
    hipStream_t stream2 = nullptr;   // side stream: 4-point contacts overlap the rest of the stage
    hipEvent_t ev_side[2] = {nullptr, nullptr};   // [0] fork, [1] join
    hipStream_t stream3 = nullptr;   // third stream: 3-point projections
    hipEvent_t ev_pk = nullptr;      // its join
    bool side_pending = false;
    // the assembly's pair structure (sort by body pair, block rows, chunks) on
    // a fourth stream, forked where the contact set is final (ev_asm_fork) and
    // joined before the assembly reduction (ev_asm_join)
    hipStream_t stream4 = nullptr;
    hipEvent_t ev_asm_fork = nullptr, ev_asm_join = nullptr;
    // the queued pair structure's plan: the contact set's serial, contacts,
    // pair count (or its bound with npd the count on the device), chunk launch size
    struct AsmPlan {
        bool pending = false; long long serial = -1; int nc = 0, np = 0, max_chunks = 0; const int* npd = nullptr;
    } ap;
    long long cs_serial = 0;         // serial of the last contact set built
    std::string err;

    // ---- scene (device)
    DBuf<int> voff, foff, eoff, vbody, fbody, ebody, face, edge;   // face/edge: global vertex ids
    DBuf<double> xbar, mass, Ic;
    // ---- state (device)
    DBuf<double> q, q0, qdot, dq, qtrial;
    DBuf<double> X, Xs, X1;          // world vertex positions: current, step start, trial
    // ---- culling scratch
    DBuf<double> bbox;               // 6 per body
    DBuf<double> bsph;               // 4 per body: bounding sphere centre, radius
    DBuf<int2> pairs;                // body pairs (I < J)
    int pair_hint[3] = {0, 0, 0};    // body-pair count of the last culling pass of each kind (narrow-phase grid
                                     // size): contact set, CCD, segment candidates
    DBuf<int> counters;              // device counters
    DBuf<int> bcnt;                  // contact-set key bucket counters (2 x 64) + VG counter
    DBuf<unsigned long long> keys_a, keys_b;  // bucketed / sorted keys
    DBuf<int> hcnt, hend;            // key sort: count and slot range per vertex or edge id
    DBuf<unsigned char> sort_tmp;
    DBuf<unsigned long long> toi_bits;

    DBuf<unsigned long long> vg_keys;
    // spatial-hash table (hash.hip): slot counts / starts (2^(3 bits) + 1),
    // per registered primitive its cell range, the slot entries, entry total
    DBuf<int> sh_cnt, sh_start, sh_ent;
    DBuf<int4> sh_rng;
    DBuf<unsigned long long> sh_total;
    ContactSet cs, cs_new;
    DBuf<int> tcount;                // trial-set counts (PT, PT + EE, all), gathered for one read
    int* h_tcount = nullptr;         // pinned host copy
    int* h_small = nullptr;          // pinned host scratch (256 ints) for the small synchronous reads of the culling passes
    ContactSet cand;                 // line search: candidates of the search segment
    bool spin_wait = true;           // host waits spin on the query (ADB_SPIN_WAIT=0: blocking calls)
    // the paths taken above a size at every size (tests): the contact-set keys
    // radix-sorted, the assembly's pair count read by the host, its block
    // rows and chunk table from the multi-launch kernels, the body boxes one
    // workgroup per body
    bool general_paths = false;
    bool ls_cand = true;             // trial sets filtered from them (ADB_LS_CANDIDATES=0: full passes)
    int ls_prev = 2;                 // trials of the previous line search
    // barrier + inertia energy at the point the last line search accepted (the
    // next Newton iteration's q, positions and contact set), valid within a step
    bool e0_valid = false;
    double e0_bi = 0.0;
    double* h_slots = nullptr;       // pinned host copies of reduction slots copied ahead of a wait
    DBuf<int> vsel;                  // trial sets: kept items per block and their scan
    DBuf<unsigned long long> keepbits;   // trial sets: test result bits
    // ---- contact terms (per PT/EE contact, then per VG contact)
    DBuf<double> crec;               // REC per PT/EE contact: compact body-space record
    DBuf<int> ctype;                 // distance type per contact
    DBuf<int> ccls;                  // class per contact: 0 inactive, 1..3 = D / 3
    DBuf<double> cdist;              // squared distance per contact
    DBuf<int> clist;                 // class 1 from the front, class 2 from the back
    DBuf<int> clist3;                // class 3 contacts
    DBuf<int> ccount;                // 4 counters: class 1, 2, 3 sizes
    DBuf<double> ctk;                // TKN per PT/EE contact: lagged friction basis (see TKN)
    DBuf<double> clam;               // 1 per PT/EE contact
    DBuf<double> vgK;                // 12 per VG contact: K (3x3 col-major) + gx (3)
    DBuf<double> vglam;
    // ---- assembly / BSR
    DBuf<double> bodyH, bodyG;       // inertia terms per body (144 + 12)
    DBuf<unsigned long long> akeys, akeys_s;   // body-pair key per contact
    DBuf<int> aval, aval_s;          // contact ids (sorted by body pair)
    DBuf<int> blk_of;                // per sorted contact: pair segment
    DBuf<int> pair_pack;             // pair count and lists packed for one copy to h_pairs
    // body pairs (lo < hi) with contacts, sorted: first contact, bodies, block
    // index in row lo (upper) and row hi (lower); the pairs keyed (hi, lo)
    DBuf<int> pair_start, pair_lo, pair_hi, pair_up, pair_low;
    DBuf<unsigned long long> tkeys, tkeys_s;
    DBuf<int> tval, tval_s;
    DBuf<int> up_start, low_start, rowcnt;   // per body: first pair (b, *), first transposed pair (*, b), row length
    int n_pairs = 0;
    DBuf<int> row_ptr, col_idx, diag_blk, vg_start;
    DBuf<int> chunk_n, chunk_off, chunk_blk;   // per pair: chunk count, first chunk; per chunk: its pair
    DBuf<double> chunk_part;                   // 456 per chunk (3 blocks + 2 gradients)
    DBuf<double> Hval;               // 144 per block
    DBuf<double> grad;               // 12 B
    // preconditioner groups: heavy-edge aggregates of the contact graph of at
    // most GMAX bodies; per group the dense inverse
    // of its diagonal sub-matrix (n x n, n = 12 bodies, column-major)
    // the group tables live in one device buffer, uploaded with one copy from a
    // pinned staging buffer: grp_moff (offset of group g's inverse in Ginv),
    // grp_off / grp_body (group g owns grp_body[grp_off[g], grp_off[g + 1])),
    // glist (group ids in launch order: multi-body groups first), grp_n (count)
    DBuf<long long> grp_pack;
    long long* h_grp_pack = nullptr;
    size_t h_grp_pack_cap = 0;
    hipEvent_t ev_grp = nullptr;     // the last upload from the staging buffer
    long long* grp_moff = nullptr;
    int *grp_off = nullptr, *grp_body = nullptr, *glist = nullptr, *grp_n = nullptr;
    DBuf<double> Ginv;
    int n_groups = 0;
    std::vector<int> h_glist;
    int n_glist[3] = {0, 0, 0};      // groups of 5..8, 2..4 and 1 bodies, in that order in glist
    std::vector<int> h_grp_off, h_grp_body;
    std::vector<long long> h_grp_moff;
    // host scratch of the group building (kept to avoid allocations per call)
    std::vector<int> hg_bucket, hg_order, hg_parent, hg_size, hg_gid, hg_bg;
    int nnzb = 0;
    // ---- PCG
    DBuf<double> pcg_r, pcg_z, pcg_p, pcg_p2, pcg_Ap, partial, scal;
    double* pcg_xo = nullptr;        // the solve's iterate: the caller's output buffer
    DBuf<int> pcg_k;                 // iteration base of the queued PCG batch
    int pcg_prev_its = 0;            // iterations of the previous solve
    int* h_pairs = nullptr;          // pinned host copy of the pair count and lists (lo, hi, start)
    size_t h_pairs_cap = 0;
    int h_pairs_n = 0;               // entries per list in h_pairs (the pair count or its bound)
    int* h_np = nullptr;             // pinned host copy of the pair count (read when unbounded)
    hipEvent_t ev_pairs = nullptr;   // their copy is done
    hipEvent_t ev_phase[2] = {nullptr, nullptr};   // contact terms + assembly on the stream: start, end
    bool use_graphs = true;          // ADB_PCG_GRAPH=0 turns graph capture off
    hipGraphExec_t pcg_graph = nullptr;   // one captured PCG batch
    static constexpr int PCG_GRAPH_KEYS = 20;
    const void* pcg_graph_key[PCG_GRAPH_KEYS] = {};   // buffer addresses the graph was captured with
    double* h_scal = nullptr;        // pinned host copies of the PCG scalars (2 x 8)
    hipEvent_t pcg_ev[2] = {nullptr, nullptr};
    // ---- energies
    DBuf<double> epart;
    // ---- host mirrors
    std::vector<int> h_voff, h_foff, h_eoff;
    std::vector<double> h_mass;
    int stall_pt = 0, stall_vg = 0;
    int num_cu = 256;
    Prof prof;
    // ---- independent scenes (adb_world_create_batched; S == 1 otherwise).
    // Bodies of scene s are [h_scene_off[s], h_scene_off[s + 1]); bodies of
    // different scenes never form a pair.
    int S = 1;
    std::vector<int> h_scene_off;
    DBuf<int> scene_off, scene_of;        // device copies; body -> scene
    bool sweep = false;                   // body pairs by sort and sweep (large or batched worlds)
    DBuf<unsigned long long> swk, swk_s;  // sweep keys (scene, box min x), sorted
    DBuf<int> swi, swi_s;                 // body ids, sorted with the keys
    DBuf<int> seg;                        // per contact list, per scene: first entry (3 x (S + 1))
    DBuf<double> sred;                    // per-scene sums (SRED_N x S)
    DBuf<double> sfac;                    // per-scene factor (CCD scale, line-search alpha)
    DBuf<int> sc_goff;                    // per scene: first preconditioner group (S + 1)
    std::vector<int> h_sc_goff;
    DBuf<double> pcg_sbb;                 // per scene: |b|^2 of the current solve
    DBuf<int> pcg_brun;                   // 1 while some scene's residual is above its tolerance
    double* h_sred = nullptr;             // pinned host copy of sred
    std::vector<int> h_stall_pt, h_stall_vg;
    // per-scene results of the last step
    std::vector<int> sc_newton, sc_ls, sc_status, sc_contacts;
    // launch shapes that change how fast a step runs, never its results
    // (adb_set_tuning; adb_amd/autotune.py picks them per workload)
    adb_tuning tune = {1.125, 4, 16};
    int pcg_graph_batch = 0;              // PCG batch size the graph was captured with
};
constexpr int SRED_N = 4;   // per-scene sum slots

// Brackets a kernel launch with events when profiling is on; `bytes` is the
// launch's algorithmic HBM traffic.
struct ProfScope {
    World& w; int k; double bytes; hipEvent_t a = nullptr; hipStream_t st;
    double flops = 0.0;               // the launch's counted FP64 operations, when known
    const int* snap = nullptr; double per[4] = {0, 0, 0, 0};
    void counts(const int* s, double p0, double p1, double p2, double p3) { snap = s; per[0] = p0; per[1] = p1; per[2] = p2; per[3] = p3; }
    ProfScope(World& w_, int k_, double bytes_, hipStream_t s = nullptr) : w(w_), k(k_), bytes(bytes_), st(s ? s : w_.stream) {
        if (!w.prof.on) return;
        a = get();
        ADB_HIP(hipEventRecord(a, st));
    }
    ~ProfScope() {
        if (!w.prof.on || !a) return;
        hipEvent_t b = get();
        (void)hipEventRecord(b, st);
        w.prof.pending.push_back({k, a, b, bytes, snap, {per[0], per[1], per[2], per[3]}, flops});
    }
    hipEvent_t get() {
        if (!w.prof.pool.empty()) { hipEvent_t e = w.prof.pool.back(); w.prof.pool.pop_back(); return e; }
        hipEvent_t e; ADB_HIP(hipEventCreate(&e)); return e;
    }
};
void prof_flush(World& w);
// Host waits for the world's stream or an event on it: a spin on the query
// (ADB_SPIN_WAIT=1, default) answers as soon as the work is done, where the
// blocking call may first put the thread to sleep; both report errors.
void wait_stream(World& w);
void wait_stream(World& w, hipStream_t s);
void wait_event(World& w, hipEvent_t e);
// Pinned host copy of 4 device counters (enqueued on the world stream) when
// profiling is on, else nullptr.
const int* prof_snapshot(World& w, const int* dev_counts);

// ---- stage launchers (implemented in the .hip files)
void transform_positions(World& w, const double* q, double* X);                         // cull.hip
void q_axpy(World& w, const double* q, const double* dq, double alpha, double* out);     // cull.hip
void transform_step(World& w, const double* q, const double* dq, double alpha, double* qout, double* X,
                    double* X2 = nullptr);                                                  // cull.hip
void collision_set(World& w, const double* X, ContactSet& out);                          // cull.hip
// PT / EE candidates of every set on the segment X0 -> X1 (cull.hip), and a
// set at a point of that segment filtered from them (VG by a full scan)
void segment_candidates(World& w, const double* X0, const double* X1, ContactSet& cand);
void trial_set(World& w, const double* X, const ContactSet& cand, ContactSet& out);
double ccd_toi(World& w, const double* X0, const double* X1, int* status);              // cull.hip
void ccd_queue(World& w, const double* X0, const double* X1);                            // cull.hip
double ccd_finish(World& w, const double* X0, const double* X1, int* status);           // cull.hip
// per-scene CCD bounds of a batched world into toi[S] (host), stalls per scene into status[S]
void ccd_toi_scenes(World& w, const double* X0, const double* X1, double* toi, int* status);   // cull.hip
// out = q + f[scene(body)] dq (f: per-scene factors on the device)
void q_axpy_scenes(World& w, const double* q, const double* dq, const double* f, double* out); // cull.hip
void scale_scenes(World& w, double* v, const double* f);                                     // cull.hip
// per-scene sums into w.sred[slot * S + s]: sum of a[i] (b[i]) over scene s's body
// rows (a, b: 12 per body), or its energy items (barrier + inertia and friction) at q, X
void dot_scenes_async(World& w, const double* a, const double* b, int slot);                 // energy.hip
void energy_scenes_async(World& w, const double* qv, const double* X, const ContactSet& cs, int slot); // energy.hip
void read_scene_sums(World& w, int nslots);                                                 // energy.hip: into w.h_sred
void hash_sets(World& w, const double* X, double r, unsigned long long* pt_keys, unsigned long long* ee_keys,
               int sub, int* bcnt);                                                      // hash.hip
void hash_ccd(World& w, const double* X0, const double* X1, unsigned long long* toi_bits); // hash.hip
void contact_terms(World& w);                                                            // contact.hip
void assemble_structure(World& w, bool forked);                                         // assemble.hip
void assemble_system(World& w);                                                          // assemble.hip
// `after` (if given) queues work behind a short solve's stop test, before
// its wait; *after_valid tells whether that solve was then complete (else
// the work ran on an unfinished x and must be queued again)
int pcg_solve(World& w, double* dq_out, int* iters, const std::function<void()>& after = {},
              bool* after_valid = nullptr);                                                 // pcg.hip
// reductions into result slots 0..7 on the world stream; read_slots waits and copies slots [0, k)
void energy_barrier_inertia_async(World& w, const double* qv, const double* X, const ContactSet& cs, int slot); // energy.hip
void energy_friction_async(World& w, const double* X, int slot);                        // energy.hip
void energy_async(World& w, const double* qv, const double* X, const ContactSet* cs, int slot_bi, int slot_f,
                  const double* da = nullptr, const double* db = nullptr, int dn = 0, int slot_d = 0);   // energy.hip
void dot_async(World& w, const double* a, const double* b, int n, int slot);            // energy.hip
void dot2_async(World& w, const double* a1, const double* b1, int slot1, const double* a2, const double* b2, int slot2,
                int n);                                                                    // energy.hip
void read_slots(World& w, int k, double* out);                                          // energy.hip
void copy_slots_async(World& w, int first, int k);    // energy.hip: slots [first, first + k) into w.h_slots, no wait
void download_sets(World& w, const ContactSet& cs, std::vector<int>& pt, std::vector<int>& ee, std::vector<int>& vg);

} // namespace adb
