// abd_internal.hpp — context and launch wrappers shared by the HIP translation units.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>
#include <vector>

#include "../../include/abd_hip.h"
#include "abd_cnt.hpp"

namespace abd {

// Candidate-pair key: scene(16) | kind(2) | i_local(23) | j_local(23).  Local indices are
// relative to the scene's first vertex / edge / triangle, so sorting the keys groups pairs by
// scene, then kind, then primitive ids.
constexpr int KEY_IBITS = 23;
constexpr uint64_t KEY_IMASK = (1ull << KEY_IBITS) - 1;
__host__ __device__ inline uint64_t make_key(int scene, int kind, int il, int jl) {
    return ((uint64_t)scene << 48) | ((uint64_t)kind << 46) | ((uint64_t)il << 23) | (uint64_t)jl;
}
__host__ __device__ inline int key_scene(uint64_t k) { return (int)(k >> 48); }
__host__ __device__ inline int key_kind(uint64_t k) { return (int)((k >> 46) & 3); }
__host__ __device__ inline int key_il(uint64_t k) { return (int)((k >> 23) & KEY_IMASK); }
__host__ __device__ inline int key_jl(uint64_t k) { return (int)(k & KEY_IMASK); }

// BVH node of one mesh, in preorder: an internal node's first child is the next node and `skip`
// is the mesh-local index one past its subtree, so a traversal needs no stack (continue at i + 1
// into a node, at skip past it).  prim >= 0 marks a leaf holding that mesh-local primitive.
// Rest-frame nodes are built once on the host; world-space copies of every body's nodes (the box
// of the rest box mapped through the body's affine state) are rebuilt by each detection.
struct BNode {
    float lo[3], hi[3];
    int prim, skip;
};
static_assert(sizeof(BNode) == 32, "BNode is 32 bytes");

// Per-body primitive ranges and BVHs.  Vertices, edges and triangles are grouped by body; bodies
// with identical local meshes share one mesh entry (one set of rest-frame nodes).
struct BodyMesh {
    const int* body_v0;   // [nb+1]
    const int* body_e0;   // [nb+1]
    const int* body_t0;   // [nb+1]
    const int* body_mesh;
    const int* mesh_tn0;  // [n_meshes+1] rest triangle-BVH node ranges
    const int* mesh_en0;  // [n_meshes+1] rest edge-BVH node ranges
    const int* body_tw0;  // [nb+1] world triangle-BVH node ranges
    const int* body_ew0;  // [nb+1] world edge-BVH node ranges
    const BNode* tri_nodes;
    const BNode* edge_nodes;
    BNode* tri_world;
    BNode* edge_world;
};

// Static scene tables on the device (read-only during a step).
struct Geo {
    int nb, nv, ne, nt, ns;
    const double4* rest;  // (x, y, z, body)
    const int2* edges;
    const int4* tris;     // (v0, v1, v2, body)
    const double* mass4;  // nb*16
    const int* pinned;
    const int* bscene;
    const int* scene_b0;  // [ns+1] prefix ranges
    const int* scene_v0;
    const int* scene_e0;
    const int* scene_t0;
    int* err;             // device error word: bits set by failed index checks
    const double* kb;     // [nb] contact stiffness of each body; a contact uses the larger of its two
                          // bodies' (adaptive; k_contact everywhere when fixed)
};

// error bits of Geo::err
enum : int { ERR_DECODE = 1, ERR_BVH = 2, ERR_INCIDENCE = 4, ERR_PAIR = 8, ERR_CSR = 16, ERR_BODY = 32,
             ERR_PENETRATION = 64 };

// Device-resident per-scene solver controls
struct SceneCtl {
    int* newton_active;  // scene still iterating
    double* residual;    // ||grad||_2 of the last assembly
    double* alpha;       // line-search step
    double* alpha0;      // CCD bound (initial step)
    unsigned long long* amax_bits;
    int* ls_done;
    int* ls_floor;       // 1 if the last line search ended at the floor
    double* E0;
    double* E1;
    double* kb;          // [nb] body contact stiffness (Geo::kb points here)
    double* dmin_prev;   // [nb] smallest contact distance of the body after the previous iteration
    unsigned long long* dmin_bits;  // [nb] smallest d^2 at the accepted iterate (bit pattern, atomicMin)
    int* k_raises;       // [ns] body stiffness doublings of the scene in the current step
    double* kc;          // [ns] largest body stiffness of the scene (reported)
    double* pred_alpha;  // [ns] step taken toward the predicted first iterate
};

// Buffers of the block-Jacobi PCG (abd_kernels.hip, k_pcg_cg); r, w, sv are ping-pong pairs.  T is
// the working precision of the matrix, the preconditioner and the vectors (double, or float in the
// mixed-precision mode); partial sums and per-scene scalars are double in both.
template <typename T>
struct PcgBufT {
    const T* b;                // right-hand side (gradient)
    const T* diag;
    const T* pinv;
    const int* row_ptr;
    const uint64_t* ent_key;
    const int* ent_val;
    const T* off;
    const T* offT;
    T* x;
    T* p;
    T* u;
    T* r[2];
    T* w[2];
    T* sv[2];
    T* mq;                     // persistent solve: M_e = B_e P_col(e) per CSR entry e (144 T each)
    double* part;
    double* sc;                // [4 * ns]: g by parity, then alpha by parity
    const double* bnorm;
    int* cg_iters;
    int* cg_done;
    int* err;
    // persistent solve: grid barrier counter, 3-slot done counters, failure flag, outputs
    unsigned* bar;
    int* done_cnt;
    int* fail;
    double* bnorm_out;
    double* residual_out;
    const double* bref;        // optional per-scene reference norm for the stopping test (warm start)
    unsigned long long* ts;    // optional launch timestamps (device wall clock): [0] earliest workgroup
                               // start, [1] latest workgroup end
};
using PcgBuf = PcgBufT<double>;
using PcgBuf32 = PcgBufT<float>;

struct GrowBuf {
    void* p = nullptr;
    size_t bytes = 0;
};

// ---------------------------------------------------------------------------------------------
// kernel launchers (abd_kernels.hip)
// ---------------------------------------------------------------------------------------------
// abd_sort.hip
hipError_t excl_scan_int(hipStream_t s, const int* in, int* out, Cnt n, int* bsum);
int scan_scratch_ints(long n);
size_t radix_scratch_bytes(long n);
// sig: the bits that may be set in any key (stable sort over those bits)
hipError_t radix_sort64(hipStream_t s, const uint64_t* kin, uint64_t* kout, const int* vin, int* vout, Cnt n,
                        uint64_t sig, void* scratch);
// two lists of at most SORT_SMALL_MAX_KEYS keys, sorted by one launch of the counting-rank sort
hipError_t small_sort2(hipStream_t s, const uint64_t* ka, uint64_t* oa, Cnt na, const uint64_t* kb, uint64_t* ob,
                       Cnt nb);

void launch_qtilde(hipStream_t s, int nb, const double* qnow, const double* qdot, double dt, const double* g, double* qtilde);
// trials > 1: X + k nv for k < trials holds the vertices at q - 2^-k alpha du (line-search trials)
void launch_xform(hipStream_t s, const Geo& G, const double* q, const double* du, const double* alpha_scene, double4* X,
                  int trials = 1);
// abd_broad.hip
void launch_body_boxes(hipStream_t s, const Geo& G, const BodyMesh& M, const double4* X0, const double4* X1,
                       double half_infl, float* box);
void launch_world_nodes(hipStream_t s, const Geo& G, const BodyMesh& M, const double* q, const double* du);
void launch_body_pair_count(hipStream_t s, const Geo& G, const float* box, int* cnt);
void launch_body_pair_write(hipStream_t s, const Geo& G, const float* box, const int* off, int2* out, int cap);
void launch_pair_queries(hipStream_t s, const Geo& G, const BodyMesh& M, int npairs, int max_queries,
                         const int2* pairs, const double4* X0, const double4* X1, const float* bbox, double infl,
                         uint64_t* cand, int cap, int* count, int* overflow, int box_only);
void launch_cache_contains(hipStream_t s, const Geo& G, const double4* Xc0, const double4* Xc1, const double4* X0,
                           const double4* X1, double m, const int* active, int* outside);
void launch_cache_filter(hipStream_t s, const Geo& G, Cnt n, const uint64_t* keys, const double4* X0,
                         const double4* X1, const AbdParams& P, double infl, int* flag);
void launch_ground_queries(hipStream_t s, const Geo& G, const double4* X0, const double4* X1, const AbdParams& P,
                           double infl, uint64_t* cand, int cap, int* count, int* overflow, int box_only);
// abd_kernels.hip
constexpr int ABD_N_FEATURES = 49;
constexpr int ABD_FEAT_Z = 37;          // features 37..48: the block-Jacobi step D_b^-1 grad_b
constexpr float ABD_FEAT_CLIP = 8.0f;   // standardised features are clipped to this magnitude
void launch_body_features(hipStream_t s, const Geo& G, const double* q, const double* qtilde, const double* qdot,
                          Cnt n_inc, const uint64_t* inc, const double* g24, Cnt n_act, const double* pinv,
                          const double* grad, float* feat);
void launch_features_pack(hipStream_t s, int nb, int rows, int dp, const float* feat, const float* mean,
                          const float* istd, uint16_t* out);
void launch_warm_x0(hipStream_t s, const Geo& G, int dpo, const float* Y, const float* omean, const float* ostd,
                    const float* feat, double* x0);
constexpr int ABD_N_PRECOND = 78;  // learned preconditioner outputs: 12 log-diagonal + 66 lower entries
void launch_learned_pinv(hipStream_t s, const Geo& G, int dpo, const float* Y, const float* omean, const float* ostd,
                         const double* diag, double* pinv);
void launch_residual(hipStream_t s, int nb, const double* b, const double* diag, const int* row_ptr,
                     const uint64_t* ent_key, const int* ent_val, const double* off, const double* offT,
                     const double* x, double* r, int n_pairs, int* err);
void launch_scene_norm(hipStream_t s, const Geo& G, const double* v, double* out);
void launch_warm_select(hipStream_t s, const Geo& G, const double* nb_b, const double* nb_r, const double* b,
                        double* rhs, double* x0, int* used);
void launch_add_inplace(hipStream_t s, int n, double* y, const double* x);
void launch_cand_eval(hipStream_t s, const Geo& G, Cnt n, const uint64_t* keys, const double4* X, const AbdParams& P,
                      int* flag, double* energy, const int* active = nullptr, int trials = 1, size_t estride = 0);
void launch_pair_d2(hipStream_t s, const Geo& G, int n, const uint64_t* keys, const double4* X, const AbdParams& P,
                    double* d2);
void launch_compact(hipStream_t s, Cnt n, const int* flag, const int* pos, const uint64_t* keys, uint64_t* out,
                    int* n_out = nullptr);
void launch_psd_test(hipStream_t s, int m, int precision, int absolute, int n, const double* A, double* out);
void launch_contact_hess(hipStream_t s, const Geo& G, Cnt n, const uint64_t* act, const double4* X, const AbdParams& P,
                         int2* body, double* g24, double* H432);
void launch_incidence(hipStream_t s, const Geo& G, Cnt n, const int2* body, uint64_t* inc, int* n_inc,
                      uint64_t* pairk, int* n_pair);
void launch_run_heads(hipStream_t s, Cnt n, const uint64_t* sorted, int shift, int* head);
void launch_pair_unique(hipStream_t s, Cnt n, const uint64_t* sorted, const int* head, const int* pos, int* pair_b,
                        int* run_start, int* n_unique, int* n_ent);
void launch_offblock_sum(hipStream_t s, Cnt n_pairs, const int* pair_b, const int* run_start,
                         const uint64_t* sorted, const int2* body, const double* H432, double* off, double* offT,
                         Cnt n_act, int* err);
void launch_csr_keys(hipStream_t s, Cnt n_pairs, const int* pair_b, uint64_t* keys, int* vals);
void launch_row_ptr(hipStream_t s, int nb, Cnt n_keys, const uint64_t* sorted, int* row_ptr, int shift = 32);
void launch_body_assemble(hipStream_t s, const Geo& G, const AbdParams& P, const double* q, const double* qtilde,
                          const int* inc_ptr, const uint64_t* inc_sorted, const int2* body, const double* g24,
                          const double* H432, const int* newton_active, double* grad, double* diag, double* pinv,
                          double* e_body, Cnt n_act);
void launch_scene_sum(hipStream_t s, const Geo& G, const double* e_body, const double* e_cand, const int* cand_lo,
                      const int* cand_hi, double* out);
void launch_scene_sum_ls(hipStream_t s, const Geo& G, const double* e_body, const double* e_cand, const int* lo,
                         const int* hi, double* out, int ls, const AbdParams& P, const int* active, SceneCtl c,
                         int floor_accept, int* pending, int trials = 1, size_t estride = 0);
void launch_pcg_tile_sum(hipStream_t s, int ns, const int* scene_tile0, const double* part, int stride, int which,
                         double* out, int do_sqrt);
template <typename T>
void launch_pcg_cg(hipStream_t s, int k, int ntiles, int ns, const int* tile_b0, const int* tile_b1,
                   const int* tile_scene, const int* scene_tile0, const PcgBufT<T>& B, double tol, int nb, int n_pairs);
template <typename T>
hipError_t launch_pcg_persist(hipStream_t s, int ntiles, int ns, const int* tile_b0, const int* tile_b1,
                              const int* tile_scene, const int* scene_tile0, const PcgBufT<T>& B, double tol,
                              int max_iter, int nb, int n_pairs);
template <typename T>
hipError_t pcg_persist_max_blocks(int* out);
template <typename T>
hipError_t launch_pcg_multi(hipStream_t s, int blocks, int ntiles, int ns, const int* tile_b0, const int* tile_b1,
                            const int* tile_scene, const int* scene_tile0, const PcgBufT<T>& B, double tol,
                            int max_iter, int nb, int n_pairs);
template <typename T>
hipError_t pcg_multi_max_blocks(int* out);
template <typename T>
void launch_pcg_mblocks(hipStream_t s, Cnt n_ent, const uint64_t* ent_key, const int* ent_val, const double* off,
                        const double* offT, const double* pinv, int nb, int n_pairs, T* mq, int* err);
// adds one persistent solve to acc: [0] wall-clock ticks (ts[1] - ts[0]), [1] algorithmic bytes
// K (fixed + per_pair P), [2] launches, with K = max(cg_iters[0..ns)) and P = *np; resets ts
void launch_pcg_account(hipStream_t s, unsigned long long* ts, double* acc, const int* cg_iters, int ns,
                        const int* np, double fixed, double per_pair);
constexpr int PCG_TILES_PER_BLOCK = 4;  // waves per workgroup of k_pcg_persist (one tile each) and k_pcg_multi
// mixed precision: FP64 -> FP32 copies of the PCG operands, FP32 -> FP64 of the solution
void launch_to_f32(hipStream_t s, long n, const double* a, float* b);
void launch_to_f64(hipStream_t s, long n, const float* a, double* b);
void launch_pcg_any(hipStream_t s, int ns, int ntiles, const int* scene_tile0, const double* part, int it,
                    const double* bnorm, double tol, const int* cg_done, int* any);
void launch_ccd(hipStream_t s, const Geo& G, Cnt n, const uint64_t* keys, const double4* X0, const double4* X1,
                const AbdParams& P, unsigned long long* amax_bits);
void launch_body_energy(hipStream_t s, const Geo& G, const AbdParams& P, const double* q, const double* du,
                        const double* alpha_scene, const double* qtilde, double* e_body, const int* active = nullptr,
                        int trials = 1);
void launch_predict_du(hipStream_t s, const Geo& G, int mode, double dt, const double* q, const double* qnow,
                       const double* qdot, const double* qtilde, double* du);
void launch_prev_x0(hipStream_t s, const Geo& G, const double* alpha, const double* du, double* x0);
void launch_step_pred(hipStream_t s, const Geo& G, int dpo, const float* Y, const float* omean, const float* ostd,
                      double dt, const double* q, const double* qnow, const double* qdot, double* du);
void launch_dmin(hipStream_t s, const Geo& G, Cnt n, const uint64_t* keys, const double4* X, const AbdParams& P,
                 unsigned long long* dmin_bits);
void launch_stiffness_update(hipStream_t s, const Geo& G, const AbdParams& P, const int* active, SceneCtl c,
                             const int* skip);
void launch_scene_kmax(hipStream_t s, const Geo& G, SceneCtl c);
void launch_apply(hipStream_t s, const Geo& G, double* q, const double* du, const double* alpha, const int* active,
                  const int* skip);
void launch_forward(hipStream_t s, int nb, double dt, const double* q, double* qnow, double* qdot);
// direct solver (abd_direct.hip): per-scene dense Cholesky of the assembled system
constexpr int DIRECT_TILE = 32;        // tile edge; a scene's matrix order is padded to a multiple
constexpr int DIRECT_MAX_DOFS = 6144;  // largest padded order (the solve keeps the vector in LDS)
struct DirectBuf {
    double* A = nullptr;            // dense lower-triangular factors of every directly solved scene
    double* Lkk = nullptr;          // [ns][DIRECT_TILE^2] factored diagonal tile of the current column
    const long long* off = nullptr; // [ns] start of scene s in A
    const int* N = nullptr;         // [ns] padded order, 0 = the scene is not solved directly
    int* mask = nullptr;            // [ns] scenes of the current solve
    int* fail = nullptr;            // [ns] a pivot was not positive
    int n_max = 0;                  // largest N
};
void launch_direct_mask(hipStream_t st, int ns, int fallback, int cg_max_iter, const int* N, const int* active,
                        const int* cg_iters, int* mask, int* any);
void launch_direct_factor(hipStream_t st, const DirectBuf& D, int ns, int nb, int n_pairs, const int* bscene,
                          const int* scene_b0, const double* diag, const int* pair_b, const double* offT);
void launch_direct_solve(hipStream_t st, const DirectBuf& D, int ns, const int* scene_b0, const double* b,
                         double* du, double* bnorm, double* residual, int* cg_iters, int zero_iters);
void launch_contact_points(hipStream_t s, const Geo& G, int n, const uint64_t* keys, const double4* X,
                           const AbdParams& P, double* mid, double* dist);

}  // namespace abd
