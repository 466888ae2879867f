// abd_internal.hpp — context and launch wrappers shared by the HIP translation units.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>
#include <vector>

#include "../../include/abd_hip.h"

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

// LBVH node: children (>= 0 internal, < 0 leaf ~index in sorted order) and their boxes
struct BvhNode {
    int left, right;
    float lo_l[3], hi_l[3], lo_r[3], hi_r[3];
    int pad[2];
};
static_assert(sizeof(BvhNode) == 64, "BvhNode is one 64-byte line");

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
};

// error bits of Geo::err
enum : int { ERR_DECODE = 1, ERR_BVH = 2, ERR_INCIDENCE = 4, ERR_PAIR = 8, ERR_CSR = 16, ERR_BODY = 32 };

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
};

struct GrowBuf {
    void* p = nullptr;
    size_t bytes = 0;
};

// ---------------------------------------------------------------------------------------------
// kernel launchers (abd_kernels.hip)
// ---------------------------------------------------------------------------------------------
// abd_sort.hip
hipError_t excl_scan_int(hipStream_t s, const int* in, int* out, int n, int* bsum);
int scan_scratch_ints(long n);
size_t radix_scratch_bytes(long n);
hipError_t radix_sort64(hipStream_t s, const uint64_t* kin, uint64_t* kout, const int* vin, int* vout, int n,
                        int end_bit, void* scratch);

void launch_qtilde(hipStream_t s, int nb, const double* qnow, const double* qdot, double dt, const double* g, double* qtilde);
void launch_xform(hipStream_t s, const Geo& G, const double* q, const double* du, const double* alpha_scene, double4* X);
void launch_scene_bounds(hipStream_t s, const Geo& G, const double4* X0, const double4* X1, float* bounds);
void launch_prim_boxes(hipStream_t s, const Geo& G, int tri_mode, const double4* X0, const double4* X1, double infl,
                       const float* bounds, float4* boxes, uint64_t* keys, int* idx);
void launch_bvh_build(hipStream_t s, int n, int ns, const uint64_t* skeys, BvhNode* nodes, int* parent, int* leaf_parent,
                      int* scene_root);
void launch_bvh_refit(hipStream_t s, int n, const int* sidx, const float4* boxes, BvhNode* nodes, const int* parent,
                      const int* leaf_parent, int* flags, int* err);
void launch_query_pt(hipStream_t s, const Geo& G, const double4* X0, const double4* X1, double infl, const AbdParams& P,
                     const BvhNode* nodes, const int* sidx, const int* scene_root, const float4* boxes,
                     uint64_t* cand, int cap, int* count, int* overflow);
void launch_query_ee(hipStream_t s, const Geo& G, const BvhNode* nodes, const int* sidx, const int* scene_root,
                     const float4* boxes, uint64_t* cand, int cap, int* count, int* overflow);
// tree sizes for traversal checks are passed through Geo::nt / Geo::ne
void launch_cand_ranges(hipStream_t s, int n, const uint64_t* keys, int* lo, int* hi);
void launch_cand_eval(hipStream_t s, const Geo& G, int n, const uint64_t* keys, const double4* X, const AbdParams& P,
                      int* flag, double* energy);
void launch_pair_d2(hipStream_t s, const Geo& G, int n, const uint64_t* keys, const double4* X, const AbdParams& P,
                    double* d2);
void launch_compact(hipStream_t s, int n, const int* flag, const int* pos, const uint64_t* keys, uint64_t* out);
void launch_contact_hess(hipStream_t s, const Geo& G, int n, const uint64_t* act, const double4* X, const AbdParams& P,
                         int2* body, double* g24, double* H432);
void launch_incidence(hipStream_t s, const Geo& G, int n, const int2* body, uint64_t* inc, int* n_inc,
                      uint64_t* pairk, int* n_pair);
void launch_run_heads(hipStream_t s, int n, const uint64_t* sorted, int shift, int* head);
void launch_pair_unique(hipStream_t s, int n, const uint64_t* sorted, const int* head, const int* pos, int* pair_b,
                        int* run_start, int* n_unique);
void launch_offblock_sum(hipStream_t s, int n_pairs, const int* pair_b, const int* run_start, int n_keys,
                         const uint64_t* sorted, const int2* body, const double* H432, double* off, double* offT,
                         int n_act, int* err);
void launch_csr_keys(hipStream_t s, int n_pairs, const int* pair_b, uint64_t* keys, int* vals);
void launch_row_ptr(hipStream_t s, int nb, int n_ent, const uint64_t* sorted, int* row_ptr);
void launch_body_assemble(hipStream_t s, const Geo& G, const AbdParams& P, const double* q, const double* qtilde,
                          int n_inc, const uint64_t* inc_sorted, const int2* body, const double* g24,
                          const double* H432, const int* newton_active, double* grad, double* diag, double* pinv,
                          double* e_body, int n_act);
void launch_scene_sum(hipStream_t s, const Geo& G, const double* e_body, const double* e_cand, const int* cand_lo,
                      const int* cand_hi, double* out);
void launch_pcg_init(hipStream_t s, int ntiles, const int* tile_b0, const int* tile_b1, const double* g,
                     const double* pinv, double* x, double* r, double* z, double* p, double* part);
void launch_pcg_tile_sum(hipStream_t s, int ns, const int* scene_tile0, const double* part, int stride, int which,
                         double* out, int do_sqrt);
void launch_pcg_A(hipStream_t s, int it, int ntiles, const int* tile_b0, const int* tile_b1, const int* tile_scene,
                  const int* scene_tile0, const double* diag, const int* row_ptr, const uint64_t* ent_key,
                  const int* ent_val, const double* off, const double* offT, const double* z, const double* p_old,
                  double* p_new, double* Ap, double* part, const double* bnorm, double tol, int* cg_iters,
                  int* cg_done, int nb, int n_pairs, int* err);
void launch_pcg_B(hipStream_t s, int it, int ntiles, const int* tile_b0, const int* tile_b1, const int* tile_scene,
                  const int* scene_tile0, const double* pinv, const double* p, const double* Ap, double* x,
                  double* r, double* z, double* part, const double* bnorm, double tol, const int* cg_done);
void launch_pcg_any(hipStream_t s, int ns, int ntiles, const int* scene_tile0, const double* part, int it,
                    const double* bnorm, double tol, const int* cg_done, int* any);
void launch_ccd(hipStream_t s, const Geo& G, int n, const uint64_t* keys, const double4* X0, const double4* X1,
                const AbdParams& P, unsigned long long* amax_bits);
void launch_body_energy(hipStream_t s, const Geo& G, const AbdParams& P, const double* q, const double* du,
                        const double* alpha_scene, const double* qtilde, double* e_body);
void launch_ls_init(hipStream_t s, int ns, const AbdParams& P, const int* newton_active, SceneCtl c);
void launch_ls_decide(hipStream_t s, int ns, const AbdParams& P, SceneCtl c, int* pending);
void launch_apply(hipStream_t s, const Geo& G, double* q, const double* du, const double* alpha, const int* active);
void launch_forward(hipStream_t s, int nb, double dt, const double* q, double* qnow, double* qdot);
void launch_newton_check(hipStream_t s, int ns, const AbdParams& P, SceneCtl c);
void launch_contact_points(hipStream_t s, const Geo& G, int n, const uint64_t* keys, const double4* X,
                           const AbdParams& P, double* mid, double* dist);

}  // namespace abd
