// abd_api.hip — host side of the MI355X ABD Newton step: the context, the device-resident
// Newton loop and the C ABI declared in include/abd_hip.h.
//
// One Newton iteration (NewtonAbd.step, abd.py:680-699):
//   assemble  narrow phase over the candidate list -> active contacts -> per-contact 12x12 blocks
//             -> body-pair blocks (sorted keys, fixed summation order) -> per-body diagonal blocks,
//             gradient, block-Jacobi inverse and energy
//   solve     tiled block-Jacobi PCG, per-scene scalars
//   CCD       swept LBVH broad phase q -> q - du (also the candidate list of the next iteration)
//             + additive CCD -> per-scene step bound
//   line search  halving from the CCD bound until the energy decreases (floor 1/64)
// Host synchronisations per iteration: counts that size the sorts, PCG convergence polls and the
// line-search/Newton flags.
#include <string.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>

#include <hip/hip_runtime.h>
#include <map>
#include <string>
#include <vector>

#include "abd_internal.hpp"
#include "abd_mlp.hpp"

using namespace abd;

namespace {

struct Timer {
    bool on = false;
    bool pcg_only = false;     // time the PCG solve launches only (light enough for a timed run)
    struct Rec {
        std::string name;
        hipEvent_t a, b;
        double bytes;
        // kslot >= 0: bytes = K x (fixed + per_pair x P), with K (iterations of a persistent solve,
        // else 1) and P (body pairs) left in device slot kslot by the launch's own stream
        int kslot = -1;
        double fixed = 0.0, per_pair = 0.0;
    };
    int* kslots = nullptr;     // device: [K, P] per timed launch whose bytes depend on device counts
    int* h_kslots = nullptr;   // pinned host copy
    int nk = 0;
    int last_idx = -1;         // record opened by the last tic (-1: that tic was not timed)
    unsigned long long* pcg_ts = nullptr;  // PCG-only mode: launch timestamps of the current solve
    double* pcg_acc = nullptr;             // PCG-only mode: summed ticks, bytes, launches
    double clock_khz = 0.0;                // device wall-clock rate
    static constexpr int KSLOTS = 1 << 14;
    std::vector<Rec> open;
    std::vector<hipEvent_t> pool;
    std::map<std::string, double> ms, bytes;
    std::map<std::string, int> calls;
    std::vector<std::string> order;
    hipEvent_t get() {
        if (!pool.empty()) {
            hipEvent_t e = pool.back();
            pool.pop_back();
            return e;
        }
        hipEvent_t e;
        (void)hipEventCreate(&e);
        return e;
    }
};

}  // namespace

struct AbdCtx {
    AbdParams P;
    int nb = 0, nv = 0, ne = 0, nt = 0, ns = 0;
    hipStream_t st = nullptr;
    std::string err;
    std::vector<void*> allocs;
    Geo G;
    // host copies of the scene tables
    std::vector<int> h_scene_b0, h_scene_v0, h_scene_e0, h_scene_t0, h_tile_b0, h_scene_tile0;
    int ntiles = 0;
    // static tables
    double4* rest = nullptr;
    int2* edges = nullptr;
    int4* tris = nullptr;
    double* mass4 = nullptr;
    int *pinned = nullptr, *bscene = nullptr, *scene_b0 = nullptr, *scene_v0 = nullptr, *scene_e0 = nullptr,
        *scene_t0 = nullptr;
    int *tile_b0 = nullptr, *tile_b1 = nullptr, *tile_scene = nullptr, *scene_tile0 = nullptr;
    // state
    double *q = nullptr, *qnow = nullptr, *qdot = nullptr, *qtilde = nullptr, *qtmp = nullptr;
    double *du = nullptr, *grad = nullptr, *diag = nullptr, *pinv = nullptr, *e_body = nullptr;
    double4 *X0 = nullptr, *X1 = nullptr, *Xt = nullptr;
    // broad phase: per-body ranges, rest-frame BVHs of the distinct meshes, body pairs
    BodyMesh M{};
    int *body_v0 = nullptr, *body_e0 = nullptr, *body_t0 = nullptr, *body_mesh = nullptr;
    int *mesh_tn0 = nullptr, *mesh_en0 = nullptr, *body_tw0 = nullptr, *body_ew0 = nullptr;
    BNode *tri_nodes = nullptr, *edge_nodes = nullptr, *tri_world = nullptr, *edge_world = nullptr;
    int n_meshes = 0;
    int *bp_cnt = nullptr, *bp_off = nullptr;  // [nb+1] per-body pair counts and offsets
    float* body_box = nullptr;
    int2* bpairs = nullptr;
    int bpair_cap = 0, n_bpairs = 0;
    int max_queries = 0;     // largest per-body-pair query count (2 max verts + max edges)

    // candidate cache (AbdParams.cand_margin > 0): the box-level pairs of the last traversal and the
    // path it was taken over
    int cache_cap = 0, n_cache = 0;
    uint64_t *cache = nullptr, *cache_s = nullptr;
    int *cache_flag = nullptr, *cache_pos = nullptr;
    double4 *Xc0 = nullptr, *Xc1 = nullptr;
    bool cache_valid = false;
    int step_traversals = 0, step_detections = 0;
    int step_slow = 0;  // Newton iterations of the step that took a second host round trip
    // HIP graph of the fast Newton iteration: captured the second time an iteration of the same key
    // (graph_key) runs eagerly, then replayed by one hipGraphLaunch while the key stays
    hipGraphExec_t gexec = nullptr;
    uint64_t gkey = 0, gkey_seen = 0;
    bool capturing = false;
    bool graph_off = false;      // a capture failed on this context: eager launches from then on
    uint64_t alloc_gen = 0;      // bumped by every device (re)allocation (captured arguments hold pointers)
    int* h_now = nullptr;        // pinned copy of the scenes' Newton flags
    int step_graph_launches = 0, step_captures = 0;
    int step_ls_trials = 0, step_ls_syncs = 0;  // line-search trials launched, host round trips waiting for them
    // candidates: nc_bound bounds the current list's length (its kernels are sized by it), the
    // length itself is the device counter C_NCAND
    int cand_cap = 0, nc_bound = 0;
    uint64_t *cand = nullptr, *cand_s = nullptr, *act = nullptr;
    int *flag = nullptr, *pos = nullptr, *cand_lo = nullptr, *cand_hi = nullptr;  // cand_hi = cand_lo + 1
    // set while fast_launches runs: k_iter_reset has cleared the iteration's counters and flags, so
    // the launch functions skip their own fills of them
    bool iter_reset = false;
    double* e_cand = nullptr;
    bool cand_valid = false;
    // active contacts
    int act_cap = 0;
    int2* act_body = nullptr;
    double *act_g = nullptr, *act_H = nullptr;
    uint64_t *inc = nullptr, *inc_s = nullptr, *pk = nullptr, *pk_s = nullptr;
    int *head = nullptr, *hpos = nullptr, *pair_b = nullptr, *run_start = nullptr;
    // exact lengths of the last assembly, read back only for the calls that hand them out (read_counts)
    int n_cand = 0, n_act = 0, n_inc = 0, n_pk = 0, n_pairs = 0;
    // BSR
    int pair_cap = 0;
    double *off = nullptr, *offT = nullptr, *mq = nullptr;  // mq: 144 per CSR entry (persistent PCG)
    uint64_t *ent_k = nullptr, *ent_ks = nullptr;
    int *ent_v = nullptr, *ent_vs = nullptr, *row_ptr = nullptr;
    int* inc_ptr = nullptr;  // each body's range of the sorted incidence keys (nb + 1)
    // PCG
    double *r = nullptr, *z = nullptr, *p0 = nullptr, *p1 = nullptr, *Ap = nullptr, *part = nullptr, *bnorm = nullptr;
    double *r1 = nullptr, *w1 = nullptr, *s0 = nullptr, *pcg_sc = nullptr;
    unsigned* pcg_ctl = nullptr;  // persistent solve: [0] barrier count, [1..3] done counters, [4] failure flag
    unsigned* h_pcg_ctl = nullptr;
    bool coop_ok = true;          // cleared when a cooperative launch fails for a reason other than its size
    bool defer_pcg_check = false; // the persistent PCG's failure flag is read with the iteration's report
    bool pcg_unchecked = false;   // the last solve was persistent and its failure flag is not read yet
    bool force_iter_pcg = false;  // solve with one launch per iteration (redo after a failed persistent solve)
    int pcg_fallbacks = 0;        // solves of the current step rerun with one launch per iteration
    int ls_trials = 2;            // line-search trials launched together before a host round trip
    int coop_blocks[2] = {-1, -1};  // co-resident k_pcg_persist workgroups (FP64, FP32); -1 = not queried
    int multi_blocks[2] = {-1, -1}; // co-resident k_pcg_multi workgroups (FP64, FP32); -1 = not queried
    bool pcg_multi = false;       // solve with one launch per iteration (autotuner choice)
    // mixed precision (AbdParams.precision == 1): FP32 copies of the PCG operands and vectors
    float *f_vec = nullptr, *f_diag = nullptr, *f_pinv = nullptr, *f_off = nullptr, *f_offT = nullptr, *f_mq = nullptr;
    int f_pair_cap = 0;
    // learned block preconditioner (abd_set_preconditioner)
    AbdMlp* pre_mlp = nullptr;
    float *pfmean = nullptr, *pfistd = nullptr, *pomean = nullptr, *postd = nullptr;
    // learned step predictor (abd_set_predictor, AbdParams.predictor 3) and its training log
    AbdMlp* pred_mlp = nullptr;
    float *prfmean = nullptr, *prfistd = nullptr, *promean = nullptr, *prostd = nullptr;
    float* feat_step = nullptr;   // features of every body at the start of the step
    bool log_steps = false;
    std::vector<float> step_feat, step_dq;
    // learned warm start (abd_set_warm_start) and training-sample log (abd_log_samples)
    AbdMlp* mlp = nullptr;
    float *feat_raw = nullptr, *fmean = nullptr, *fistd = nullptr, *omean = nullptr, *ostd = nullptr;
    double *x0 = nullptr, *rhs = nullptr, *nrm_b = nullptr, *nrm_r = nullptr;
    int* warm_used = nullptr;
    bool log_on = false;
    std::vector<float> log_feat, log_du;
    // direct solver (AbdParams.solver 1 or 2): dense per-scene factors, sized on first use
    DirectBuf dn;
    long long* dn_off = nullptr;
    int *dn_N = nullptr, *dn_mask = nullptr, *dn_fail = nullptr;
    double *dn_A = nullptr, *dn_L = nullptr;
    size_t dn_cap = 0;  // doubles allocated at dn_A
    int dn_maxd = -1;   // direct_max_dofs the tables were built for
    std::vector<int> h_dn_N;
    int bodies_per_tile = 5;
    int* cg_iters = nullptr;
    int* cg_done = nullptr;
    // per-scene control
    SceneCtl C;
    int* nact_s = nullptr;
    int* act_before = nullptr;
    double* report = nullptr;  // head of the commit staging block: report (R_N x ns doubles), then
                               // the Newton flags (ns ints), the counters (C_N) and the PCG words (8)
    char* h_stage = nullptr;   // pinned copy of the staging block
    int* cnt = nullptr;       // device counters
    int* h_cnt = nullptr;     // pinned host mirror
    double* h_report = nullptr;
    // sort scratch
    void* sort_tmp = nullptr;
    size_t sort_bytes = 0;
    // logs
    std::vector<double> log_res, log_alpha;
    std::vector<int> log_cg;
    std::vector<AbdStepStats> stats;
    bool assembled = false, solved = false;
    int newton_it = -1;  // Newton iteration of the running step (-1: outside a step)
    Timer T;
};

namespace {

// device counters: list lengths the kernels of one Newton iteration hand to each other (C_NCAND
// candidates, C_NACT active contacts, C_NINC incidences, C_NPK contact body pairs, C_NUNIQ
// distinct body pairs, C_NENT CSR entries), flags and the error word
enum { C_COUNT = 0, C_OVERFLOW = 1, C_NINC = 2, C_NPK = 3, C_NUNIQ = 4, C_ANY = 5, C_PENDING = 6, C_NACT = 7, C_ERR = 8,
       C_OUTSIDE = 9, C_NCAND = 10, C_NENT = 11, C_SKIP = 12, C_NCACHE = 13, C_N = 14 };
enum { R_RES = 0, R_ALPHA, R_ALPHA0, R_FLOOR, R_ACTIVE, R_CG, R_E1, R_NACT, R_NCAND, R_KC, R_RAISES, R_PALPHA, R_N };

#define HIPCHK(expr)                                                                               \
    do {                                                                                           \
        hipError_t _e = (expr);                                                                    \
        if (_e != hipSuccess) {                                                                    \
            ctx->err = std::string(#expr) + ": " + hipGetErrorName(_e) + " " + hipGetErrorString(_e); \
            return ABD_ERR_HIP;                                                                    \
        }                                                                                          \
    } while (0)

#define CHK(expr)                                                                                  \
    do {                                                                                           \
        int _rc = (expr);                                                                          \
        if (_rc != ABD_OK) return _rc;                                                             \
    } while (0)

template <class T>
int dalloc(AbdCtx* ctx, T** p, size_t n) {
    if (ctx->capturing) {
        ctx->err = "device allocation during graph capture";
        return ABD_ERR_STATE;
    }
    ctx->alloc_gen += 1;
    n = std::max<size_t>(n, 1);
    HIPCHK(hipMalloc((void**)p, n * sizeof(T)));
    ctx->allocs.push_back((void*)*p);
    return ABD_OK;
}

template <class T>
int dfree(AbdCtx* ctx, T*& p) {
    if (p) {
        if (ctx->capturing) {
            ctx->err = "device free during graph capture";
            return ABD_ERR_STATE;
        }
        ctx->alloc_gen += 1;
        auto it = std::find(ctx->allocs.begin(), ctx->allocs.end(), (void*)p);
        if (it != ctx->allocs.end()) ctx->allocs.erase(it);
        HIPCHK(hipFree(p));
        p = nullptr;
    }
    return ABD_OK;
}

// events time every stage (mode 1), or only the per-iteration PCG launches in the PCG-only mode,
// whose persistent solves stamp the device wall clock themselves (graph replays included)
bool timed(const AbdCtx* ctx, const char* name) {
    return ctx->T.on && (!ctx->T.pcg_only || std::strcmp(name, "pcg_iter") == 0);
}

void tic(AbdCtx* ctx, const char* name, double bytes = 0.0) {
    ctx->T.last_idx = -1;
    if (!timed(ctx, name)) return;
    Timer::Rec r{name, ctx->T.get(), ctx->T.get(), bytes};
    (void)hipEventRecord(r.a, ctx->st);
    ctx->T.open.push_back(r);
    ctx->T.last_idx = (int)ctx->T.open.size() - 1;
}

// tic/toc pairs do not nest: toc closes the record its tic opened, if that tic was timed
void toc(AbdCtx* ctx) {
    if (ctx->T.last_idx < 0) return;
    (void)hipEventRecord(ctx->T.open[ctx->T.last_idx].b, ctx->st);
}

// the bytes of the record the last tic opened: K x (fixed + per_pair x body pairs), K the most
// iterations of any scene (kmax, a persistent solve) or 1, both read on the device
int time_bytes(AbdCtx* ctx, const int* kmax, double fixed, double per_pair);

void timer_collect(AbdCtx* ctx) {
    if (!ctx->T.on || (ctx->T.open.empty() && ctx->T.nk == 0)) return;
    (void)hipStreamSynchronize(ctx->st);
    if (ctx->T.nk > 0) {
        (void)hipMemcpy(ctx->T.h_kslots, ctx->T.kslots, sizeof(int) * 2 * ctx->T.nk, hipMemcpyDeviceToHost);
        for (auto& r : ctx->T.open)
            if (r.kslot >= 0)
                r.bytes = (double)ctx->T.h_kslots[2 * r.kslot] *
                          (r.fixed + r.per_pair * (double)ctx->T.h_kslots[2 * r.kslot + 1]);
        ctx->T.nk = 0;
    }
    ctx->T.last_idx = -1;
    for (auto& r : ctx->T.open) {
        float ms = 0.f;
        (void)hipEventElapsedTime(&ms, r.a, r.b);
        if (!ctx->T.calls.count(r.name)) ctx->T.order.push_back(r.name);
        ctx->T.ms[r.name] += ms;
        ctx->T.calls[r.name] += 1;
        ctx->T.bytes[r.name] += r.bytes;
        ctx->T.pool.push_back(r.a);
        ctx->T.pool.push_back(r.b);
    }
    ctx->T.open.clear();
}

// --------------------------------------------------------------------------------------------
// small helper kernels
// --------------------------------------------------------------------------------------------
__global__ void k_fill_i32(int n, int* p, int v) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}
// the fast Newton iteration's clears in one launch: the incidence and body-pair counts, the cache
// containment flag, the line search's pending flag, the PCG control words and iteration counts, and
// each scene's CCD bound (1.0)
__global__ void k_iter_reset(int ns, int* cnt, unsigned long long* amax_bits, int* cg_iters, unsigned* pcg_ctl) {
    const int t = threadIdx.x;
    if (t == 0) {
        cnt[C_NINC] = 0;
        cnt[C_NPK] = 0;
        cnt[C_OUTSIDE] = 0;
        cnt[C_PENDING] = 0;
    }
    if (t < 8) pcg_ctl[t] = 0u;
    for (int s = t; s < ns; s += blockDim.x) {
        amax_bits[s] = 0x3FF0000000000000ull;
        cg_iters[s] = 0;
    }
}
__global__ void k_fill_u64(int n, unsigned long long* p, unsigned long long v) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}
// active contacts per scene from the flags and their exclusive scan; thread 0 also stores the
// total (the scan's last entry plus the last flag)
__global__ void k_nact_scene(int ns, const int* lo, const int* hi, const int* flag, const int* pos, int* out, Cnt cn,
                             int* n_act) {
    int s = blockIdx.x * blockDim.x + threadIdx.x;
    const int nc = cnt_of(cn);
    if (s == 0) *n_act = nc > 0 ? pos[nc - 1] + flag[nc - 1] : 0;
    if (s >= ns) return;
    int a = lo[s], b = hi[s];
    out[s] = (b > a) ? (pos[b - 1] + flag[b - 1] - pos[a]) : 0;
}
// timing slot: out[0] = max(a[0..n)) (1 without a), out[1] = *np; one wave
__global__ void k_timing_slot(int n, const int* a, const int* np, int* out) {
    int m = a ? 0 : 1;
    if (a)
        for (int i = threadIdx.x; i < n; i += 64) m = max(m, a[i]);
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) m = max(m, __shfl_xor(m, o, 64));
    if (threadIdx.x == 0) {
        out[0] = m;
        out[1] = *np;
    }
}
// bytes of the commit staging block (report, Newton flags, counters, PCG words)
inline size_t stage_bytes(int ns) { return sizeof(double) * R_N * ns + sizeof(int) * (ns + C_N + 8); }

// C_SKIP: the speculative part of a Newton iteration is discarded when a scene's line search is still
// halving (C_PENDING), the swept path left the cached candidate region (C_OUTSIDE), the persistent
// PCG gave up at a grid barrier (*pcg_fail, when that solve's flag is unread) or the assembly's lists
// outgrew their buffers
__device__ __forceinline__ bool skip_test(const int* cnt, const unsigned* pcg_fail, int act_cap, int pair_cap) {
    return cnt[C_PENDING] != 0 || cnt[C_OUTSIDE] != 0 || (pcg_fail && *pcg_fail != 0u) || cnt[C_NACT] > act_cap ||
           cnt[C_NUNIQ] > pair_cap;
}

// add_du_kernel (abd.py:780-785): q -= alpha du, with the fast iteration's skip test (every thread
// evaluates it; thread 0 stores it in C_SKIP for the kernels after and the host): nothing moves
// when it holds
__global__ void k_apply_tested(Geo G, double* q, const double* __restrict__ du, const double* __restrict__ alpha,
                               const int* __restrict__ active, int* cnt, const unsigned* pcg_fail, int act_cap,
                               int pair_cap) {
    const bool skip = skip_test(cnt, pcg_fail, act_cap, pair_cap);
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i == 0) cnt[C_SKIP] = skip ? 1 : 0;
    if (i >= 12 * G.nb || skip) return;
    const int s = G.bscene[i / 12];
    if (active[s]) q[i] -= alpha[s] * du[i];
}

// The end of a Newton iteration, one workgroup per scene: the scene's largest contact stiffness
// (k_scene_kmax), the convergence test (abd.py:699, or, with stall_alpha > 0, a step pinned below
// it; nothing when the skip flag is set), the report,
// and the staging block that one copy brings to the host (workgroup 0 adds the counters and the
// PCG words, final by now)
__global__ __launch_bounds__(256) void k_commit_tail(Geo G, AbdParams P, SceneCtl c, int* active_before,
                                                     const int* skip, const int* cg, const int* nact, const int* lo,
                                                     const int* hi, const int* cnt, const unsigned* pcg_ctl,
                                                     double* stage) {
    __shared__ double sh[4];
    const int s = blockIdx.x, ns = G.ns;
    double k = 0.0;
    for (int b = G.scene_b0[s] + threadIdx.x; b < G.scene_b0[s + 1]; b += 256) k = fmax(k, c.kb[b]);
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) k = fmax(k, __shfl_xor(k, o, 64));
    if (threadIdx.x % 64 == 0) sh[threadIdx.x / 64] = k;
    __syncthreads();
    int* stage_i = (int*)(stage + (size_t)R_N * ns);
    if (threadIdx.x == 0) {
        const double kc = fmax(fmax(sh[0], sh[1]), fmax(sh[2], sh[3]));
        c.kc[s] = kc;
        if (!(skip && *skip)) {
            const int a = c.newton_active[s];
            active_before[s] = a;
            if (a && (c.residual[s] < P.newton_tol || (P.stall_alpha > 0.0 && c.alpha[s] < P.stall_alpha)))
                c.newton_active[s] = 0;
        }
        double* R = stage + (size_t)s * R_N;
        R[R_RES] = c.residual[s];
        R[R_ALPHA] = c.alpha[s];
        R[R_ALPHA0] = c.alpha0[s];
        R[R_FLOOR] = c.ls_floor[s];
        R[R_ACTIVE] = active_before[s];
        R[R_CG] = cg[s];
        R[R_E1] = c.E1[s];
        R[R_NACT] = nact[s];
        R[R_NCAND] = hi[s] - lo[s];
        R[R_KC] = kc;
        R[R_RAISES] = c.k_raises[s];
        R[R_PALPHA] = c.pred_alpha[s];
        stage_i[s] = c.newton_active[s];
    }
    if (s == 0) {
        for (int i = threadIdx.x; i < C_N; i += 256) stage_i[ns + i] = cnt[i];
        if (threadIdx.x < 8) stage_i[ns + C_N + threadIdx.x] = (int)pcg_ctl[threadIdx.x];
    }
}

// *n_out = pos[n - 1] + flag[n - 1] (the length of a compacted list), 0 for an empty input
__global__ void k_scan_total(Cnt cn, const int* flag, const int* pos, int* n_out) {
    const int n = cnt_of(cn);
    if (blockIdx.x == 0 && threadIdx.x == 0) *n_out = n > 0 ? pos[n - 1] + flag[n - 1] : 0;
}
__global__ void k_copy(int n, const double* a, double* b) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) b[i] = a[i];
}

inline int cdiv(long a, int b) { return (int)((a + b - 1) / b); }

// Debug mode (AbdParams.reserved0 bit 0): synchronise after every launch group and report the
// first failing one by name, with the device error word.
int dbg(AbdCtx* ctx, const char* what) {
    if (!(ctx->P.reserved0 & 1)) return ABD_OK;
    hipError_t e = hipStreamSynchronize(ctx->st);
    if (e == hipSuccess) e = hipGetLastError();
    if (e != hipSuccess) {
        ctx->err = std::string("after ") + what + ": " + hipGetErrorName(e) + " " + hipGetErrorString(e);
        return ABD_ERR_HIP;
    }
    int w = 0;
    if (hipMemcpy(&w, ctx->cnt + C_ERR, sizeof(int), hipMemcpyDeviceToHost) != hipSuccess) return ABD_ERR_HIP;
    if (w) {
        char buf[160];
        std::snprintf(buf, sizeof buf, "after %s: device index check failed (error bits 0x%x)", what, w);
        ctx->err = buf;
        return ABD_ERR_STATE;
    }
    if (ctx->P.reserved0 & 2) { std::fprintf(stderr, "[abd] ok %s\n", what); std::fflush(stderr); }
    return ABD_OK;
}

int fill_i32(AbdCtx* ctx, int* p, int n, int v) {
    if (n <= 0) return ABD_OK;
    hipLaunchKernelGGL(k_fill_i32, dim3(cdiv(n, 256)), dim3(256), 0, ctx->st, n, p, v);
    CHK(dbg(ctx, "k_fill_i32"));
    HIPCHK(hipGetLastError());
    return ABD_OK;
}

int time_bytes(AbdCtx* ctx, const int* kmax, double fixed, double per_pair) {
    if (ctx->T.last_idx < 0 || ctx->T.nk >= Timer::KSLOTS) return ABD_OK;
    if (!ctx->T.kslots) {
        HIPCHK(hipMalloc((void**)&ctx->T.kslots, sizeof(int) * 2 * Timer::KSLOTS));
        HIPCHK(hipHostMalloc((void**)&ctx->T.h_kslots, sizeof(int) * 2 * Timer::KSLOTS, hipHostMallocDefault));
    }
    const int slot = ctx->T.nk++;
    hipLaunchKernelGGL(k_timing_slot, dim3(1), dim3(64), 0, ctx->st, ctx->ns, kmax, ctx->cnt + C_NUNIQ,
                       ctx->T.kslots + 2 * slot);
    Timer::Rec& r = ctx->T.open[ctx->T.last_idx];
    r.kslot = slot;
    r.fixed = fixed;
    r.per_pair = per_pair;
    return ABD_OK;
}

// --------------------------------------------------------------------------------------------
// sort / scan (abd_sort.hip) with scratch grown on demand
// --------------------------------------------------------------------------------------------
int ensure_tmp(AbdCtx* ctx, size_t bytes) {
    if (bytes <= ctx->sort_bytes) return ABD_OK;
    if (ctx->capturing) {
        ctx->err = "scratch growth during graph capture";
        return ABD_ERR_STATE;
    }
    ctx->alloc_gen += 1;
    HIPCHK(hipStreamSynchronize(ctx->st));
    CHK(dfree(ctx, ctx->sort_tmp));
    size_t nb = bytes + bytes / 4 + 4096;
    HIPCHK(hipMalloc(&ctx->sort_tmp, nb));
    ctx->allocs.push_back(ctx->sort_tmp);
    ctx->sort_bytes = nb;
    return ABD_OK;
}

// sig: the bits that may be set in any key
int sort_keys(AbdCtx* ctx, const uint64_t* in, uint64_t* out, Cnt n, uint64_t sig) {
    if (n.cap <= 0) return ABD_OK;
    CHK(ensure_tmp(ctx, radix_scratch_bytes(n.cap)));
    HIPCHK(radix_sort64(ctx->st, in, out, nullptr, nullptr, n, sig, ctx->sort_tmp));
    return dbg(ctx, "radix_sort64(keys)");
}

int sort_pairs(AbdCtx* ctx, const uint64_t* kin, uint64_t* kout, const int* vin, int* vout, Cnt n, uint64_t sig) {
    if (n.cap <= 0) return ABD_OK;
    CHK(ensure_tmp(ctx, radix_scratch_bytes(n.cap)));
    HIPCHK(radix_sort64(ctx->st, kin, kout, vin, vout, n, sig, ctx->sort_tmp));
    return dbg(ctx, "radix_sort64(pairs)");
}

int excl_scan(AbdCtx* ctx, const int* in, int* out, Cnt n) {
    if (n.cap <= 0) return ABD_OK;
    CHK(ensure_tmp(ctx, sizeof(int) * (size_t)scan_scratch_ints(n.cap)));
    HIPCHK(excl_scan_int(ctx->st, in, out, n, (int*)ctx->sort_tmp));
    return dbg(ctx, "excl_scan_int");
}

// the device error word, after h_cnt came back
int check_err_word(AbdCtx* ctx) {
    if (ctx->h_cnt[C_ERR]) {
        char buf[160];
        if (ctx->h_cnt[C_ERR] == ERR_PENETRATION)
            std::snprintf(buf, sizeof buf, "a vertex is on or below the ground plane (infeasible state)");
        else
            std::snprintf(buf, sizeof buf, "device index check failed (error bits 0x%x)", ctx->h_cnt[C_ERR]);
        ctx->err = buf;
        return ABD_ERR_STATE;
    }
    return ABD_OK;
}

int sync_counters(AbdCtx* ctx) {
    if (ctx->capturing) {
        ctx->err = "host round trip during graph capture";
        return ABD_ERR_STATE;
    }
    HIPCHK(hipMemcpyAsync(ctx->h_cnt, ctx->cnt, C_N * sizeof(int), hipMemcpyDeviceToHost, ctx->st));
    HIPCHK(hipStreamSynchronize(ctx->st));
    return check_err_word(ctx);
}

int bits_for(int n) {
    int b = 1;
    while ((1 << b) < n) ++b;
    return b;
}

// mask of the low `bits` bits
uint64_t low_bits(int bits) { return bits >= 64 ? ~0ull : ((1ull << bits) - 1); }

// the bits a candidate key (make_key) can set: the scene index (none for one scene), the kind, and
// the two scene-relative primitive ids up to the largest scene's vertex, edge and triangle counts;
// the radix sort passes over these bits only
uint64_t cand_key_sig(const AbdCtx* ctx) {
    int mi = 1, mj = 1;
    for (int s = 0; s < ctx->ns; ++s) {
        const int nv = ctx->h_scene_v0[s + 1] - ctx->h_scene_v0[s], ne = ctx->h_scene_e0[s + 1] - ctx->h_scene_e0[s];
        const int nt = ctx->h_scene_t0[s + 1] - ctx->h_scene_t0[s];
        mi = std::max(mi, std::max(nv, ne));
        mj = std::max(mj, std::max(nt, ne));
    }
    const uint64_t scene = ctx->ns > 1 ? low_bits(bits_for(ctx->ns)) << 48 : 0ull;
    return scene | (3ull << 46) | (low_bits(bits_for(mi)) << 23) | low_bits(bits_for(mj));
}

// --------------------------------------------------------------------------------------------
// capacities
// --------------------------------------------------------------------------------------------
int grow_candidates(AbdCtx* ctx, int cap) {
    if (cap <= ctx->cand_cap) return ABD_OK;
    cap = std::max(cap, ctx->cand_cap * 2);
    CHK(dfree(ctx, ctx->cand)); CHK(dfree(ctx, ctx->cand_s)); CHK(dfree(ctx, ctx->act));
    CHK(dfree(ctx, ctx->flag)); CHK(dfree(ctx, ctx->pos)); CHK(dfree(ctx, ctx->e_cand));
    CHK(dalloc(ctx, &ctx->cand, cap)); CHK(dalloc(ctx, &ctx->cand_s, cap)); CHK(dalloc(ctx, &ctx->act, cap));
    CHK(dalloc(ctx, &ctx->flag, cap)); CHK(dalloc(ctx, &ctx->pos, cap));
    CHK(dalloc(ctx, &ctx->e_cand, (size_t)LS_TRIALS_MAX * cap));  // one slice per line-search trial
    ctx->cand_cap = cap;
    return ABD_OK;
}

int grow_active(AbdCtx* ctx, int cap) {
    if (cap <= ctx->act_cap) return ABD_OK;
    cap = std::max(cap, ctx->act_cap * 2);
    if (cap > (1 << 24)) {
        ctx->err = "more than 2^24 active contacts";
        return ABD_ERR_CAPACITY;
    }
    CHK(dfree(ctx, ctx->act_body)); CHK(dfree(ctx, ctx->act_g)); CHK(dfree(ctx, ctx->act_H));
    CHK(dfree(ctx, ctx->inc)); CHK(dfree(ctx, ctx->inc_s)); CHK(dfree(ctx, ctx->pk)); CHK(dfree(ctx, ctx->pk_s));
    CHK(dfree(ctx, ctx->head)); CHK(dfree(ctx, ctx->hpos)); CHK(dfree(ctx, ctx->pair_b)); CHK(dfree(ctx, ctx->run_start));
    CHK(dalloc(ctx, &ctx->act_body, cap)); CHK(dalloc(ctx, &ctx->act_g, (size_t)cap * 24));
    CHK(dalloc(ctx, &ctx->act_H, (size_t)cap * 432));
    CHK(dalloc(ctx, &ctx->inc, 2 * (size_t)cap)); CHK(dalloc(ctx, &ctx->inc_s, 2 * (size_t)cap));
    CHK(dalloc(ctx, &ctx->pk, cap)); CHK(dalloc(ctx, &ctx->pk_s, cap));
    CHK(dalloc(ctx, &ctx->head, cap)); CHK(dalloc(ctx, &ctx->hpos, cap));
    CHK(dalloc(ctx, &ctx->pair_b, 2 * (size_t)cap)); CHK(dalloc(ctx, &ctx->run_start, (size_t)cap + 1));
    ctx->act_cap = cap;
    return ABD_OK;
}

int grow_pairs(AbdCtx* ctx, int cap) {
    if (cap <= ctx->pair_cap) return ABD_OK;
    cap = std::max(cap, ctx->pair_cap * 2);
    CHK(dfree(ctx, ctx->off)); CHK(dfree(ctx, ctx->offT)); CHK(dfree(ctx, ctx->ent_k)); CHK(dfree(ctx, ctx->ent_ks));
    CHK(dfree(ctx, ctx->ent_v)); CHK(dfree(ctx, ctx->ent_vs)); CHK(dfree(ctx, ctx->mq));
    CHK(dalloc(ctx, &ctx->off, (size_t)cap * 144)); CHK(dalloc(ctx, &ctx->offT, (size_t)cap * 144));
    CHK(dalloc(ctx, &ctx->mq, 2 * (size_t)cap * 144));
    CHK(dalloc(ctx, &ctx->ent_k, 2 * (size_t)cap)); CHK(dalloc(ctx, &ctx->ent_ks, 2 * (size_t)cap));
    CHK(dalloc(ctx, &ctx->ent_v, 2 * (size_t)cap)); CHK(dalloc(ctx, &ctx->ent_vs, 2 * (size_t)cap));
    ctx->pair_cap = cap;
    return ABD_OK;
}

// --------------------------------------------------------------------------------------------
// broad phase over the linear path q -> q - du (du == nullptr: the static set at q), with
// X0 / X1 the vertex positions at both ends (ContactSolverBase.update_bvh + the collision
// queries, contact.py:143-169, 244-279; the rule itself is stated in abd_broad.hip)
// --------------------------------------------------------------------------------------------
int grow_body_pairs(AbdCtx* ctx, int cap) {
    if (cap <= ctx->bpair_cap) return ABD_OK;
    cap = std::max(cap, ctx->bpair_cap * 2);
    CHK(dfree(ctx, ctx->bpairs));
    CHK(dalloc(ctx, &ctx->bpairs, cap));
    ctx->bpair_cap = cap;
    return ABD_OK;
}

int grow_cache(AbdCtx* ctx, int cap) {
    if (cap <= ctx->cache_cap) return ABD_OK;
    cap = std::max(cap, ctx->cache_cap * 2);
    CHK(dfree(ctx, ctx->cache)); CHK(dfree(ctx, ctx->cache_s)); CHK(dfree(ctx, ctx->cache_flag));
    CHK(dfree(ctx, ctx->cache_pos));
    CHK(dalloc(ctx, &ctx->cache, cap)); CHK(dalloc(ctx, &ctx->cache_s, cap)); CHK(dalloc(ctx, &ctx->cache_flag, cap));
    CHK(dalloc(ctx, &ctx->cache_pos, cap));
    ctx->cache_cap = cap;
    return ABD_OK;
}

// BVH traversal over the path X0 -> X1 (q -> q - du): the pairs meeting the candidate rule with
// inflation infl (box_only: those whose boxes overlap) into the unsorted list of the target
// (the candidate buffers, or the cache's), growing it as needed; *n receives the count
int traverse(AbdCtx* ctx, const double4* X0, const double4* X1, const double* du, double infl, int box_only,
             bool to_cache, int* n_out) {
    ctx->step_traversals += 1;
    tic(ctx, "broad_phase_pairs");
    launch_body_boxes(ctx->st, ctx->G, ctx->M, X0, X1, 0.5 * infl, ctx->body_box);
    CHK(dbg(ctx, "launch_body_boxes"));
    launch_body_pair_count(ctx->st, ctx->G, ctx->body_box, ctx->bp_cnt);
    CHK(dbg(ctx, "launch_body_pair_count"));
    CHK(excl_scan(ctx, ctx->bp_cnt, ctx->bp_off, host_cnt(ctx->nb + 1)));
    HIPCHK(hipMemcpyAsync(ctx->cnt + C_COUNT, ctx->bp_off + ctx->nb, sizeof(int), hipMemcpyDeviceToDevice, ctx->st));
    CHK(sync_counters(ctx));
    ctx->n_bpairs = ctx->h_cnt[C_COUNT];
    if (ctx->n_bpairs < 0) {
        ctx->err = "body-pair count overflowed";
        return ABD_ERR_CAPACITY;
    }
    CHK(grow_body_pairs(ctx, ctx->n_bpairs));
    launch_body_pair_write(ctx->st, ctx->G, ctx->body_box, ctx->bp_off, ctx->bpairs, ctx->bpair_cap);
    CHK(dbg(ctx, "launch_body_pair_write"));
    HIPCHK(hipGetLastError());
    toc(ctx);
    tic(ctx, "broad_phase_nodes");
    launch_world_nodes(ctx->st, ctx->G, ctx->M, ctx->q, du);
    CHK(dbg(ctx, "launch_world_nodes"));
    toc(ctx);
    for (int attempt = 0; attempt < 8; ++attempt) {
        uint64_t* out = to_cache ? ctx->cache : ctx->cand;
        const int cap = to_cache ? ctx->cache_cap : ctx->cand_cap;
        // algorithmic bytes: positions once, primitive indices once, candidate keys out
        tic(ctx, "broad_phase_query", (double)ctx->nv * 24 + (2.0 * ctx->ne + 3.0 * ctx->nt) * 4);
        HIPCHK(hipMemsetAsync(ctx->cnt + C_COUNT, 0, 2 * sizeof(int), ctx->st));
        launch_pair_queries(ctx->st, ctx->G, ctx->M, ctx->n_bpairs, ctx->max_queries, ctx->bpairs, X0, X1,
                            ctx->body_box, infl, out, cap, ctx->cnt + C_COUNT, ctx->cnt + C_OVERFLOW, box_only);
        CHK(dbg(ctx, "launch_pair_queries"));
        launch_ground_queries(ctx->st, ctx->G, X0, X1, ctx->P, infl, out, cap, ctx->cnt + C_COUNT,
                              ctx->cnt + C_OVERFLOW, box_only);
        CHK(dbg(ctx, "launch_ground_queries"));
        HIPCHK(hipGetLastError());
        toc(ctx);
        CHK(sync_counters(ctx));
        if (ctx->h_cnt[C_OVERFLOW] == 2) {
            ctx->err = "BVH traversal failed (malformed tree)";
            return ABD_ERR_CAPACITY;
        }
        const int n = ctx->h_cnt[C_COUNT];
        if (ctx->P.reserved0 & 4)
            std::fprintf(stderr, "[abd] broad phase (%s%s): %d body pairs, %d pairs\n", du ? "swept" : "static",
                         to_cache ? ", cache" : "", ctx->n_bpairs, n);
        if (n <= cap) {
            *n_out = n;
            return ABD_OK;
        }
        if (attempt == 7) {
            ctx->err = "candidate list did not fit";
            return ABD_ERR_CAPACITY;
        }
        CHK(to_cache ? grow_cache(ctx, n + n / 2) : grow_candidates(ctx, n + n / 2));
    }
    return ABD_OK;
}

// the candidate list's length: bounded by nc_bound on the host, exact in C_NCAND on the device
Cnt cand_cnt(const AbdCtx* ctx) { return Cnt{ctx->nc_bound, ctx->cnt + C_NCAND}; }

// per-scene ranges of the sorted candidates: scene s holds [cand_lo[s], cand_lo[s + 1]) (cand_hi is
// cand_lo + 1)
int cand_ranges(AbdCtx* ctx) {
    launch_row_ptr(ctx->st, ctx->ns, cand_cnt(ctx), ctx->cand_s, ctx->cand_lo, 48);
    CHK(dbg(ctx, "launch_row_ptr(scenes)"));
    HIPCHK(hipGetLastError());
    return ABD_OK;
}

// the candidates over X0 -> X1 from the cache: rebuilt by a widened traversal when a vertex of the
// path leaves the cached region, then filtered by the exact rule (sorted order kept); the filtered
// length stays on the device (C_NCAND), bounded by the cache's length
//
// speculative: the containment test's answer (C_OUTSIDE) is not waited for; the cached list is used
// either way and the caller discards the iteration when the flag turns out to be set.
int cached_broad_phase(AbdCtx* ctx, const double4* X0, const double4* X1, const double* du, bool speculative) {
    const double m = ctx->P.cand_margin;
    if (ctx->cache_valid) {
        tic(ctx, "cache_check", (double)ctx->nv * 128);
        if (!ctx->iter_reset) CHK(fill_i32(ctx, ctx->cnt + C_OUTSIDE, 1, 0));
        // batches skip the vertices of scenes whose Newton loop has stopped (one scene: nothing to skip)
        launch_cache_contains(ctx->st, ctx->G, ctx->Xc0, ctx->Xc1, X0, X1, 0.5 * m,
                              ctx->ns > 1 ? ctx->C.newton_active : nullptr, ctx->cnt + C_OUTSIDE);
        CHK(dbg(ctx, "launch_cache_contains"));
        toc(ctx);
        if (!speculative) {
            CHK(sync_counters(ctx));
            if (ctx->h_cnt[C_OUTSIDE]) ctx->cache_valid = false;
        }
    }
    if (!ctx->cache_valid) {
        int n = 0;
        CHK(traverse(ctx, X0, X1, du, ctx->P.dhat + 2.0 * m, 1, true, &n));
        ctx->n_cache = n;
        CHK(fill_i32(ctx, ctx->cnt + C_NCACHE, 1, n));
        tic(ctx, "candidate_sort");
        CHK(sort_keys(ctx, ctx->cache, ctx->cache_s, host_cnt(n), cand_key_sig(ctx)));
        toc(ctx);
        HIPCHK(hipMemcpyAsync(ctx->Xc0, X0, ctx->nv * sizeof(double4), hipMemcpyDeviceToDevice, ctx->st));
        HIPCHK(hipMemcpyAsync(ctx->Xc1, X1, ctx->nv * sizeof(double4), hipMemcpyDeviceToDevice, ctx->st));
        ctx->cache_valid = true;
    }
    // the filter and its compaction are sized by the cache's capacity and read its length on the
    // device, so their launches stay the same from one rebuild of the cache to the next
    const Cnt cn{ctx->cache_cap, ctx->cnt + C_NCACHE};
    CHK(grow_candidates(ctx, std::max(ctx->cache_cap, 1)));
    ctx->nc_bound = ctx->cache_cap;
    if (ctx->cache_cap > 0) {
        // algorithmic bytes: the cached key, four vertices at both ends, the flag
        tic(ctx, "broad_phase_filter", (double)ctx->n_cache * (8 + 8 * 32 + 4));
        launch_cache_filter(ctx->st, ctx->G, cn, ctx->cache_s, X0, X1, ctx->P, ctx->P.dhat, ctx->cache_flag);
        CHK(dbg(ctx, "launch_cache_filter"));
        CHK(excl_scan(ctx, ctx->cache_flag, ctx->cache_pos, cn));
        launch_compact(ctx->st, cn, ctx->cache_flag, ctx->cache_pos, ctx->cache_s, ctx->cand_s, ctx->cnt + C_NCAND);
        CHK(dbg(ctx, "launch_compact"));
        toc(ctx);
    } else {
        CHK(fill_i32(ctx, ctx->cnt + C_NCAND, 1, 0));
    }
    if (ctx->P.reserved0 & 4) {
        CHK(sync_counters(ctx));
        std::fprintf(stderr, "[abd] cached broad phase (%s): %d cached, %d candidates\n", du ? "swept" : "static",
                     ctx->n_cache, ctx->h_cnt[C_NCAND]);
    }
    return cand_ranges(ctx);
}

int broad_phase(AbdCtx* ctx, const double4* X0, const double4* X1, const double* du, bool speculative = false) {
    ctx->step_detections += 1;
    if (ctx->P.cand_margin > 0.0) return cached_broad_phase(ctx, X0, X1, du, speculative);
    int n = 0;
    CHK(traverse(ctx, X0, X1, du, ctx->P.dhat, 0, false, &n));
    ctx->nc_bound = n;
    CHK(fill_i32(ctx, ctx->cnt + C_NCAND, 1, n));
    tic(ctx, "candidate_sort");
    CHK(sort_keys(ctx, ctx->cand, ctx->cand_s, host_cnt(n), cand_key_sig(ctx)));
    toc(ctx);
    return cand_ranges(ctx);
}

int static_broad_phase(AbdCtx* ctx) {
    launch_xform(ctx->st, ctx->G, ctx->q, nullptr, nullptr, ctx->X0);
    CHK(dbg(ctx, "launch_xform"));
    CHK(broad_phase(ctx, ctx->X0, ctx->X0, nullptr));
    ctx->cand_valid = true;
    return ABD_OK;
}

// --------------------------------------------------------------------------------------------
// assembly at the current iterate over the current candidate list
// --------------------------------------------------------------------------------------------
// The assembly's lists live on the device, bounded by the buffers: at most act_cap active contacts
// (2 act_cap incidences, act_cap contact body pairs) and pair_cap distinct body pairs (2 pair_cap CSR
// entries).  Kernels read the lengths on the device and clamp them to these capacities, so an
// assembly whose active set outgrew them stays in bounds and is flagged instead (assembly_overflow);
// the caller grows the buffers to the counts and assembles again.
Cnt act_cnt(const AbdCtx* ctx) { return Cnt{ctx->act_cap, ctx->cnt + C_NACT}; }
Cnt inc_cnt(const AbdCtx* ctx) { return Cnt{2 * ctx->act_cap, ctx->cnt + C_NINC}; }
Cnt pk_cnt(const AbdCtx* ctx) { return Cnt{ctx->act_cap, ctx->cnt + C_NPK}; }
Cnt pair_cnt(const AbdCtx* ctx) { return Cnt{ctx->pair_cap, ctx->cnt + C_NUNIQ}; }
Cnt ent_cnt(const AbdCtx* ctx) { return Cnt{2 * ctx->pair_cap, ctx->cnt + C_NENT}; }

// the counters of the last assembly (h_cnt, after a round trip) exceed the buffers
bool assembly_overflow(const AbdCtx* ctx) {
    return ctx->h_cnt[C_NACT] > ctx->act_cap || ctx->h_cnt[C_NUNIQ] > ctx->pair_cap;
}

// buffers for the counts of the last assembly, with room to grow
int grow_to_counts(AbdCtx* ctx) {
    CHK(grow_active(ctx, ctx->h_cnt[C_NACT] + ctx->h_cnt[C_NACT] / 2 + 64));
    CHK(grow_pairs(ctx, ctx->h_cnt[C_NUNIQ] + ctx->h_cnt[C_NUNIQ] / 2 + 64));
    return ABD_OK;
}

int assemble(AbdCtx* ctx) {
    const int nc = ctx->nc_bound;
    const Cnt cc = cand_cnt(ctx), ca = act_cnt(ctx);
    if (ctx->act_cap == 0) CHK(grow_active(ctx, 1024));
    if (ctx->pair_cap == 0) CHK(grow_pairs(ctx, 1024));
    launch_xform(ctx->st, ctx->G, ctx->q, nullptr, nullptr, ctx->X0);
    CHK(dbg(ctx, "launch_xform"));
    tic(ctx, "narrow_phase", (double)nc * (8 + 4 * 32 + 4 + 8));
    launch_cand_eval(ctx->st, ctx->G, cc, ctx->cand_s, ctx->X0, ctx->P, ctx->flag, ctx->e_cand);
    CHK(dbg(ctx, "launch_cand_eval"));
    toc(ctx);
    CHK(excl_scan(ctx, ctx->flag, ctx->pos, cc));
    hipLaunchKernelGGL(k_nact_scene, dim3(cdiv(ctx->ns, 128)), dim3(128), 0, ctx->st, ctx->ns, ctx->cand_lo,
                       ctx->cand_hi, ctx->flag, ctx->pos, ctx->nact_s, cc, ctx->cnt + C_NACT);
    CHK(dbg(ctx, "k_nact_scene"));
    launch_compact(ctx->st, cc, ctx->flag, ctx->pos, ctx->cand_s, ctx->act);
    CHK(dbg(ctx, "launch_compact"));
    tic(ctx, "contact_hess", (double)ctx->act_cap * (8 + 4 * 32 + 432 * 8 + 24 * 8 + 8));
    launch_contact_hess(ctx->st, ctx->G, ca, ctx->act, ctx->X0, ctx->P, ctx->act_body, ctx->act_g, ctx->act_H);
    CHK(dbg(ctx, "launch_contact_hess"));
    toc(ctx);
    if (!ctx->iter_reset) CHK(fill_i32(ctx, ctx->cnt + C_NINC, 2, 0));  // C_NINC, C_NPK
    launch_incidence(ctx->st, ctx->G, ca, ctx->act_body, ctx->inc, ctx->cnt + C_NINC, ctx->pk, ctx->cnt + C_NPK);
    CHK(dbg(ctx, "launch_incidence"));
    HIPCHK(hipGetLastError());
    const Cnt cpk = pk_cnt(ctx);
    tic(ctx, "block_sort");
    // incidence keys body << 32 | contact << 1 | role; body-pair keys lo << 44 | hi << 24 | contact
    const uint64_t bbits = low_bits(bits_for(ctx->nb)), cbits = low_bits(bits_for(std::max(ctx->act_cap, 1)));
    if (inc_cnt(ctx).cap <= SORT_SMALL_MAX_KEYS && cpk.cap <= SORT_SMALL_MAX_KEYS) {
        // both lists fit the counting-rank sort: one launch for the two
        HIPCHK(small_sort2(ctx->st, ctx->inc, ctx->inc_s, inc_cnt(ctx), ctx->pk, ctx->pk_s, cpk));
        CHK(dbg(ctx, "small_sort2"));
    } else {
        CHK(sort_keys(ctx, ctx->inc, ctx->inc_s, inc_cnt(ctx), (bbits << 32) | (cbits << 1) | 1ull));
        CHK(sort_keys(ctx, ctx->pk, ctx->pk_s, cpk, (bbits << 44) | (bbits << 24) | cbits));
    }
    toc(ctx);
    launch_row_ptr(ctx->st, ctx->nb, inc_cnt(ctx), ctx->inc_s, ctx->inc_ptr);
    CHK(dbg(ctx, "launch_row_ptr"));
    launch_run_heads(ctx->st, cpk, ctx->pk_s, 24, ctx->head);
    CHK(dbg(ctx, "launch_run_heads"));
    CHK(excl_scan(ctx, ctx->head, ctx->hpos, cpk));
    launch_pair_unique(ctx->st, cpk, ctx->pk_s, ctx->head, ctx->hpos, ctx->pair_b, ctx->run_start,
                       ctx->cnt + C_NUNIQ, ctx->cnt + C_NENT);
    CHK(dbg(ctx, "launch_pair_unique"));
    tic(ctx, "offblock_sum", (double)ctx->pair_cap * 144 * 8 * 3);
    launch_offblock_sum(ctx->st, pair_cnt(ctx), ctx->pair_b, ctx->run_start, ctx->pk_s, ctx->act_body, ctx->act_H,
                        ctx->off, ctx->offT, ca, ctx->cnt + C_ERR);
    CHK(dbg(ctx, "launch_offblock_sum"));
    toc(ctx);
    launch_csr_keys(ctx->st, pair_cnt(ctx), ctx->pair_b, ctx->ent_k, ctx->ent_v);
    CHK(dbg(ctx, "launch_csr_keys"));
    CHK(sort_pairs(ctx, ctx->ent_k, ctx->ent_ks, ctx->ent_v, ctx->ent_vs, ent_cnt(ctx),
                   (bbits << 32) | bbits));  // body << 32 | body
    launch_row_ptr(ctx->st, ctx->nb, ent_cnt(ctx), ctx->ent_ks, ctx->row_ptr);
    CHK(dbg(ctx, "launch_row_ptr"));
    tic(ctx, "body_assemble", (double)ctx->nb * (3 * 96 + 128 + 3 * 1152 + 96 + 8));
    launch_body_assemble(ctx->st, ctx->G, ctx->P, ctx->q, ctx->qtilde, ctx->inc_ptr, ctx->inc_s, ctx->act_body,
                         ctx->act_g, ctx->act_H, ctx->C.newton_active, ctx->grad, ctx->diag, ctx->pinv, ctx->e_body,
                         ca);
    CHK(dbg(ctx, "launch_body_assemble"));
    toc(ctx);
    HIPCHK(hipGetLastError());
    ctx->assembled = true;
    return ABD_OK;
}

// an assembly whose lists fit its buffers (one host round trip; a second assembly after growing
// them if they did not)
int assemble_checked(AbdCtx* ctx) {
    for (int attempt = 0; attempt < 4; ++attempt) {
        CHK(assemble(ctx));
        CHK(sync_counters(ctx));
        if (!assembly_overflow(ctx)) return ABD_OK;
        CHK(grow_to_counts(ctx));
    }
    ctx->err = "assembly lists did not fit after growing their buffers";
    return ABD_ERR_CAPACITY;
}

// exact list lengths of the last broad phase and assembly, read back for the calls that hand them
// out (one host round trip)
int read_counts(AbdCtx* ctx) {
    CHK(sync_counters(ctx));
    ctx->n_cand = ctx->h_cnt[C_NCAND];
    ctx->n_act = ctx->h_cnt[C_NACT];
    ctx->n_inc = ctx->h_cnt[C_NINC];
    ctx->n_pk = ctx->h_cnt[C_NPK];
    ctx->n_pairs = ctx->h_cnt[C_NUNIQ];
    return ABD_OK;
}

// --------------------------------------------------------------------------------------------
// block-Jacobi PCG (compute_du, abd.py:631-658); du <- solution, residual <- ||grad||
// --------------------------------------------------------------------------------------------
// Learned warm start: body features of the current assembly -> MLP -> x0; r0 = b - H x0 per scene,
// kept only where ||r0|| < ||b|| (plain PCG from zero otherwise)
int warm_start(AbdCtx* ctx) {
    AbdMlp* m = ctx->mlp;
    if (m) {
        launch_body_features(ctx->st, ctx->G, ctx->q, ctx->qtilde, ctx->qdot, inc_cnt(ctx), ctx->inc_s, ctx->act_g,
                             act_cnt(ctx), ctx->pinv, ctx->grad, ctx->feat_raw);
        CHK(dbg(ctx, "launch_body_features"));
        int rc = mlp_reserve(m, ctx->nb);
        if (rc != ABD_OK) {
            ctx->err = "warm start: " + m->err;
            return rc;
        }
        const int rows = (ctx->nb + 15) / 16 * 16;
        launch_features_pack(ctx->st, ctx->nb, rows, m->dp[0], ctx->feat_raw, ctx->fmean, ctx->fistd, m->A[0]);
        mlp_forward_dev(m, ctx->nb, ctx->st);
        launch_warm_x0(ctx->st, ctx->G, m->dp[m->L], m->Z[m->L - 1], ctx->omean, ctx->ostd, ctx->feat_raw, ctx->x0);
        CHK(dbg(ctx, "warm start forward"));
    } else {
        // AbdParams.pcg_x0 = 1: the part of the previous direction its line search did not take
        launch_prev_x0(ctx->st, ctx->G, ctx->C.alpha, ctx->du, ctx->x0);
        CHK(dbg(ctx, "launch_prev_x0"));
    }
    launch_residual(ctx->st, ctx->nb, ctx->grad, ctx->diag, ctx->row_ptr, ctx->ent_ks, ctx->ent_vs, ctx->off,
                    ctx->offT, ctx->x0, ctx->rhs, ctx->pair_cap, ctx->cnt + C_ERR);
    launch_scene_norm(ctx->st, ctx->G, ctx->grad, ctx->nrm_b);
    launch_scene_norm(ctx->st, ctx->G, ctx->rhs, ctx->nrm_r);
    launch_warm_select(ctx->st, ctx->G, ctx->nrm_b, ctx->nrm_r, ctx->grad, ctx->rhs, ctx->x0, ctx->warm_used);
    CHK(dbg(ctx, "warm start residual"));
    HIPCHK(hipGetLastError());
    return ABD_OK;
}

// Learned block preconditioner: body features of the current assembly (with the block-Jacobi
// inverse still in pinv) -> MLP -> pinv <- L_D^-T T T^T L_D^-1 per unpinned body
int learned_precond(AbdCtx* ctx) {
    AbdMlp* m = ctx->pre_mlp;
    launch_body_features(ctx->st, ctx->G, ctx->q, ctx->qtilde, ctx->qdot, inc_cnt(ctx), ctx->inc_s, ctx->act_g,
                         act_cnt(ctx), ctx->pinv, ctx->grad, ctx->feat_raw);
    CHK(dbg(ctx, "launch_body_features"));
    int rc = mlp_reserve(m, ctx->nb);
    if (rc != ABD_OK) {
        ctx->err = "learned preconditioner: " + m->err;
        return rc;
    }
    const int rows = (ctx->nb + 15) / 16 * 16;
    launch_features_pack(ctx->st, ctx->nb, rows, m->dp[0], ctx->feat_raw, ctx->pfmean, ctx->pfistd, m->A[0]);
    mlp_forward_dev(m, ctx->nb, ctx->st);
    launch_learned_pinv(ctx->st, ctx->G, m->dp[m->L], m->Z[m->L - 1], ctx->pomean, ctx->postd, ctx->diag, ctx->pinv);
    CHK(dbg(ctx, "learned preconditioner"));
    HIPCHK(hipGetLastError());
    return ABD_OK;
}

// after the PCG: add the warm start, restore ||b|| as the residual, log training samples
int check_params(AbdCtx* ctx, const AbdParams* p) {
    auto bad = [&](const char* what) {
        ctx->err = std::string("invalid parameter: ") + what;
        return ABD_ERR_ARG;
    };
    auto finite = [](double v) { return std::isfinite(v); };
    if (!(p->dt > 0.0) || !finite(p->dt)) return bad("dt must be positive");
    if (!(p->dhat > 0.0) || !finite(p->dhat)) return bad("dhat must be positive");
    if (!(p->kappa >= 0.0) || !finite(p->kappa)) return bad("kappa must be >= 0");
    if (!(p->k_contact >= 0.0) || !finite(p->k_contact)) return bad("k_contact must be >= 0");
    if (!(p->newton_tol >= 0.0) || !finite(p->newton_tol)) return bad("newton_tol must be >= 0");
    if (p->max_newton < 1) return bad("max_newton must be >= 1");
    if (p->cg_max_iter < 1) return bad("cg_max_iter must be >= 1");
    if (!(p->cg_tol > 0.0) || !(p->cg_tol < 1.0)) return bad("cg_tol must be in (0, 1)");
    if (!(p->ccd_gap_ratio > 0.0) || !(p->ccd_gap_ratio < 1.0)) return bad("ccd_gap_ratio must be in (0, 1)");
    if (p->ccd_max_iter < 1) return bad("ccd_max_iter must be >= 1");
    if (!(p->ls_floor > 0.0) || !(p->ls_floor < 1.0)) return bad("ls_floor must be in (0, 1)");
    if (!(p->pinned_diag > 0.0) || !finite(p->pinned_diag)) return bad("pinned_diag must be positive");
    for (int k = 0; k < 3; ++k)
        if (!finite(p->gravity[k]) || !finite(p->ground_normal[k])) return bad("gravity and ground_normal must be finite");
    if (!finite(p->ground_offset)) return bad("ground_offset must be finite");
    if (p->ground_enabled &&
        !(p->ground_normal[0] * p->ground_normal[0] + p->ground_normal[1] * p->ground_normal[1] +
              p->ground_normal[2] * p->ground_normal[2] > 0.0))
        return bad("ground_normal must be non-zero");
    if (p->predictor < 0 || p->predictor > 3) return bad("predictor must be 0, 1, 2 or 3");
    if (p->hess_proj != 0 && p->hess_proj != 1) return bad("hess_proj must be 0 or 1");
    if (p->pcg_x0 != 0 && p->pcg_x0 != 1) return bad("pcg_x0 must be 0 or 1");
    if (!finite(p->cand_margin)) return bad("cand_margin must be finite");
    if (!(p->stall_alpha >= 0.0) || !(p->stall_alpha < 1.0)) return bad("stall_alpha must be in [0, 1)");
    if (!finite(p->k_max)) return bad("k_max must be finite");
    if (p->k_max > p->k_contact && (!(p->k_raise_gap > 0.0) || !(p->k_raise_gap < 1.0)))
        return bad("k_raise_gap must be in (0, 1)");
    if (p->precision != 0 && p->precision != 1) {
        ctx->err = "precision must be 0 (FP64) or 1 (mixed)";
        return ABD_ERR_ARG;
    }
    if (p->solver < 0 || p->solver > 2) {
        ctx->err = "solver must be 0 (PCG), 1 (direct) or 2 (PCG with direct fallback)";
        return ABD_ERR_ARG;
    }
    if (p->direct_max_dofs > DIRECT_MAX_DOFS) {
        ctx->err = "direct_max_dofs above " + std::to_string(DIRECT_MAX_DOFS);
        return ABD_ERR_ARG;
    }
    return ABD_OK;
}

// per-scene dense matrix tables of the direct solver: padded order N_s (0 for a scene above
// direct_max_dofs) and offsets; the matrices are allocated once for every scene that fits
int ensure_dense(AbdCtx* ctx) {
    const int maxd = ctx->P.direct_max_dofs <= 0 ? DIRECT_MAX_DOFS : ctx->P.direct_max_dofs;
    if (ctx->dn_maxd == maxd) return ABD_OK;
    const int ns = ctx->ns;
    std::vector<long long> off(ns, 0);
    std::vector<int> N(ns, 0);
    long long total = 0;
    int nmax = 0;
    for (int s = 0; s < ns; ++s) {
        const int nr = 12 * (ctx->h_scene_b0[s + 1] - ctx->h_scene_b0[s]);
        if (nr > maxd) continue;
        const int n = (nr + DIRECT_TILE - 1) / DIRECT_TILE * DIRECT_TILE;
        off[s] = total;
        N[s] = n;
        total += (long long)n * n;
        nmax = std::max(nmax, n);
    }
    if (total * 8 > (64LL << 30)) {
        ctx->err = "direct solver: the dense matrices would take " + std::to_string(total * 8 >> 20) +
                   " MiB (limit 64 GiB); lower direct_max_dofs";
        return ABD_ERR_ARG;
    }
    if (!ctx->dn_N) {
        CHK(dalloc(ctx, &ctx->dn_off, ns));
        CHK(dalloc(ctx, &ctx->dn_N, ns));
        CHK(dalloc(ctx, &ctx->dn_mask, ns));
        CHK(dalloc(ctx, &ctx->dn_fail, ns));
        CHK(dalloc(ctx, &ctx->dn_L, (size_t)ns * DIRECT_TILE * DIRECT_TILE));
    }
    if ((size_t)total > ctx->dn_cap) {
        CHK(dfree(ctx, ctx->dn_A));
        CHK(dalloc(ctx, &ctx->dn_A, (size_t)total));
        ctx->dn_cap = (size_t)total;
    }
    HIPCHK(hipMemcpyAsync(ctx->dn_off, off.data(), ns * sizeof(long long), hipMemcpyHostToDevice, ctx->st));
    HIPCHK(hipMemcpyAsync(ctx->dn_N, N.data(), ns * sizeof(int), hipMemcpyHostToDevice, ctx->st));
    HIPCHK(hipStreamSynchronize(ctx->st));
    DirectBuf& D = ctx->dn;
    D.A = ctx->dn_A; D.Lkk = ctx->dn_L; D.off = ctx->dn_off; D.N = ctx->dn_N; D.mask = ctx->dn_mask;
    D.fail = ctx->dn_fail; D.n_max = nmax;
    ctx->h_dn_N = N;
    ctx->dn_maxd = maxd;
    return ABD_OK;
}

// direct solve of H du = grad (compute_du's direct branch, abd.py:637-654): every active scene
// (fallback false), or the active scenes whose PCG ended at cg_max_iter (fallback true)
int solve_direct(AbdCtx* ctx, bool fallback) {
    CHK(ensure_dense(ctx));
    const int ns = ctx->ns;
    if (!fallback)
        for (int s = 0; s < ns; ++s)
            if (ctx->h_dn_N[s] == 0) {
                ctx->err = "direct solver: scene " + std::to_string(s) + " has " +
                           std::to_string(12 * (ctx->h_scene_b0[s + 1] - ctx->h_scene_b0[s])) +
                           " unknowns, above direct_max_dofs " + std::to_string(ctx->dn_maxd);
                return ABD_ERR_ARG;
            }
    HIPCHK(hipMemsetAsync(ctx->cnt + C_ANY, 0, sizeof(int), ctx->st));
    launch_direct_mask(ctx->st, ns, fallback ? 1 : 0, ctx->P.cg_max_iter, ctx->dn_N, ctx->C.newton_active,
                       ctx->cg_iters, ctx->dn_mask, ctx->cnt + C_ANY);
    CHK(dbg(ctx, "launch_direct_mask"));
    if (fallback) {
        CHK(sync_counters(ctx));
        if (!ctx->h_cnt[C_ANY]) return ABD_OK;
    }
    HIPCHK(hipMemsetAsync(ctx->dn_fail, 0, ns * sizeof(int), ctx->st));
    CHK(read_counts(ctx));  // the scatter launch covers the exact body-pair count
    tic(ctx, "direct_factor");
    launch_direct_factor(ctx->st, ctx->dn, ns, ctx->nb, ctx->n_pairs, ctx->bscene, ctx->scene_b0, ctx->diag,
                         ctx->pair_b, ctx->offT);
    toc(ctx);
    CHK(dbg(ctx, "launch_direct_factor"));
    tic(ctx, "direct_solve");
    launch_direct_solve(ctx->st, ctx->dn, ns, ctx->scene_b0, ctx->grad, ctx->du, ctx->bnorm, ctx->C.residual,
                        ctx->cg_iters, fallback ? 0 : 1);
    toc(ctx);
    CHK(dbg(ctx, "launch_direct_solve"));
    std::vector<int> bad(ns);
    HIPCHK(hipMemcpyAsync(bad.data(), ctx->dn_fail, ns * sizeof(int), hipMemcpyDeviceToHost, ctx->st));
    HIPCHK(hipStreamSynchronize(ctx->st));
    for (int s = 0; s < ns; ++s)
        if (bad[s]) {
            ctx->err = "direct solver: the system of scene " + std::to_string(s) + " is not positive definite";
            return ABD_ERR_STATE;
        }
    return ABD_OK;
}

int solve_finish(AbdCtx* ctx, bool warm) {
    if (warm) {
        launch_add_inplace(ctx->st, 12 * ctx->nb, ctx->du, ctx->x0);
        hipLaunchKernelGGL(k_copy, dim3(cdiv(ctx->ns, 64)), dim3(64), 0, ctx->st, ctx->ns, ctx->nrm_b, ctx->bnorm);
        hipLaunchKernelGGL(k_copy, dim3(cdiv(ctx->ns, 64)), dim3(64), 0, ctx->st, ctx->ns, ctx->nrm_b, ctx->C.residual);
        CHK(dbg(ctx, "warm start finish"));
    }
    if (ctx->P.solver == 2) CHK(solve_direct(ctx, true));
    if (ctx->log_on) {
        // features were computed from the block-Jacobi inverse before the solve (log_features)
        const int nb = ctx->nb;
        std::vector<float> f((size_t)ABD_N_FEATURES * nb);
        std::vector<double> du(12 * (size_t)nb);
        std::vector<int> act(ctx->ns);
        HIPCHK(hipMemcpyAsync(f.data(), ctx->feat_raw, f.size() * sizeof(float), hipMemcpyDeviceToHost, ctx->st));
        HIPCHK(hipMemcpyAsync(du.data(), ctx->du, du.size() * sizeof(double), hipMemcpyDeviceToHost, ctx->st));
        HIPCHK(hipMemcpyAsync(act.data(), ctx->C.newton_active, ctx->ns * sizeof(int), hipMemcpyDeviceToHost, ctx->st));
        HIPCHK(hipStreamSynchronize(ctx->st));
        std::vector<int> pinned(nb), bscene(nb);
        for (int s = 0; s < ctx->ns; ++s)
            for (int b = ctx->h_scene_b0[s]; b < ctx->h_scene_b0[s + 1]; ++b) bscene[b] = s;
        HIPCHK(hipMemcpy(pinned.data(), ctx->pinned, nb * sizeof(int), hipMemcpyDeviceToHost));
        for (int b = 0; b < nb; ++b) {
            if (pinned[b] || !act[bscene[b]]) continue;
            ctx->log_feat.insert(ctx->log_feat.end(), f.begin() + (size_t)ABD_N_FEATURES * b,
                                 f.begin() + (size_t)ABD_N_FEATURES * (b + 1));
            for (int j = 0; j < 12; ++j) ctx->log_du.push_back((float)du[12 * (size_t)b + j]);
        }
    }
    HIPCHK(hipGetLastError());
    ctx->solved = true;
    return ABD_OK;
}

// one cooperative launch for the whole solve (k_pcg_persist); false when the launch was refused or
// a grid barrier timed out, in which case the caller solves again with one launch per iteration
// does one cooperative launch of k_pcg_persist<T> over the current tiles fit on the device?
template <typename T>
bool persist_fits(AbdCtx* ctx) {
    int& cap = ctx->coop_blocks[sizeof(T) == 4 ? 1 : 0];
    if (cap < 0 && pcg_persist_max_blocks<T>(&cap) != hipSuccess) {
        (void)hipGetLastError();
        cap = 0;
    }
    return (ctx->ntiles + PCG_TILES_PER_BLOCK - 1) / PCG_TILES_PER_BLOCK <= cap;
}

// co-resident workgroups of k_pcg_multi<T> (the persistent form for more tiles than fit at once)
template <typename T>
int multi_cap(AbdCtx* ctx) {
    int& cap = ctx->multi_blocks[sizeof(T) == 4 ? 1 : 0];
    if (cap < 0 && pcg_multi_max_blocks<T>(&cap) != hipSuccess) {
        (void)hipGetLastError();
        cap = 0;
    }
    return cap;
}

// one persistent launch per solve is possible (k_pcg_persist or k_pcg_multi)
bool persist_ok(AbdCtx* ctx) {
    if (!ctx->coop_ok) return false;
    return ctx->P.precision == 1 ? (persist_fits<float>(ctx) || multi_cap<float>(ctx) > 0)
                                 : (persist_fits<double>(ctx) || multi_cap<double>(ctx) > 0);
}

template <typename T>
int solve_persistent(AbdCtx* ctx, const PcgBufT<T>& B, bool* ok) {
    *ok = false;
    if (!ctx->coop_ok || ctx->pcg_multi || ctx->force_iter_pcg || (ctx->P.reserved0 & 8)) return ABD_OK;
    // one tile per wave when the tiles fit on the device at once (k_pcg_persist), else waves that
    // take several tiles per stage (k_pcg_multi)
    // debug bit 32: the multi-tile form even where the tiles fit (a test of that form at small sizes)
    const bool fits = persist_fits<T>(ctx) && !(ctx->P.reserved0 & 32);
    const int mcap = fits ? 0 : multi_cap<T>(ctx);
    if (!fits && mcap <= 0) return ABD_OK;
    if (!ctx->iter_reset) {
        CHK(fill_i32(ctx, ctx->cg_iters, ctx->ns, 0));
        CHK(fill_i32(ctx, (int*)ctx->pcg_ctl, 8, 0));
    }
    if (!fits) CHK(fill_i32(ctx, ctx->cg_done, ctx->ns, 0));
    // debug bit 4: raise the failure flag before the launch, so every workgroup gives up at its first
    // barrier that does not fill at once and the solve is redone with one launch per iteration
    const bool forced = (ctx->P.reserved0 & 16) != 0;
    if (forced) hipLaunchKernelGGL(k_fill_i32, dim3(1), dim3(64), 0, ctx->st, 1, (int*)(ctx->pcg_ctl + 4), 1);
    if (fits && sizeof(T) == 8) {  // the FP64 solve multiplies by M_e = B_e P_col (the FP32 one by P, then B)
        launch_pcg_mblocks(ctx->st, ent_cnt(ctx), ctx->ent_ks, ctx->ent_vs, ctx->off, ctx->offT, ctx->pinv,
                           ctx->nb, ctx->pair_cap, B.mq, ctx->cnt + C_ERR);
        CHK(dbg(ctx, "launch_pcg_mblocks"));
    }
    tic(ctx, "pcg_solve");
    hipError_t e;
    if (fits) {
        e = launch_pcg_persist(ctx->st, ctx->ntiles, ctx->ns, ctx->tile_b0, ctx->tile_b1, ctx->tile_scene,
                               ctx->scene_tile0, B, ctx->P.cg_tol, ctx->P.cg_max_iter, ctx->nb, ctx->pair_cap);
    } else {
        const int blocks = std::min(mcap, (ctx->ntiles + PCG_TILES_PER_BLOCK - 1) / PCG_TILES_PER_BLOCK);
        e = launch_pcg_multi(ctx->st, blocks, ctx->ntiles, ctx->ns, ctx->tile_b0, ctx->tile_b1, ctx->tile_scene,
                             ctx->scene_tile0, B, ctx->P.cg_tol, ctx->P.cg_max_iter, ctx->nb, ctx->pair_cap);
    }
    toc(ctx);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        ctx->coop_ok = false;
        ctx->pcg_fallbacks += 1;
        return ABD_OK;
    }
    CHK(dbg(ctx, "launch_pcg_persist"));
    {
        // algorithmic bytes of the launch: K iterations x [nnzb * (144 T + 4) + nb * (8 * 12 T + 144 T)],
        // nnzb = nb + 2 pairs
        const double t = sizeof(T);
        const double fixed = (double)ctx->nb * (144 * t + 4 + 8 * 12 * t + 144 * t), per_pair = 2.0 * (144 * t + 4);
        if (B.ts)
            launch_pcg_account(ctx->st, B.ts, ctx->T.pcg_acc, ctx->cg_iters, ctx->ns, ctx->cnt + C_NUNIQ, fixed, per_pair);
        else
            CHK(time_bytes(ctx, ctx->cg_iters, fixed, per_pair));
    }
    if (ctx->defer_pcg_check && !forced) {
        // the failure flag arrives with the iteration's report; the iteration is discarded and redone
        // with one launch per PCG iteration if it is set
        ctx->pcg_unchecked = true;
        *ok = true;
        return ABD_OK;
    }
    HIPCHK(hipMemcpyAsync(ctx->h_pcg_ctl, ctx->pcg_ctl, 8 * sizeof(unsigned), hipMemcpyDeviceToHost, ctx->st));
    HIPCHK(hipStreamSynchronize(ctx->st));
    if (ctx->h_pcg_ctl[4]) {
        // a grid barrier did not fill: do not try again on this context (unless forced for a test)
        if (!forced) ctx->coop_ok = false;
        ctx->pcg_fallbacks += 1;
        return ABD_OK;
    }
    *ok = true;
    return ABD_OK;
}

// FP32 buffers of the mixed-precision PCG, sized for the current pair count
int ensure_f32(AbdCtx* ctx) {
    const int nb = ctx->nb;
    if (!ctx->f_vec) {
        CHK(dalloc(ctx, &ctx->f_vec, 11 * 12 * (size_t)nb));
        CHK(dalloc(ctx, &ctx->f_diag, 144 * (size_t)nb));
        CHK(dalloc(ctx, &ctx->f_pinv, 144 * (size_t)nb));
    }
    if (ctx->f_pair_cap < ctx->pair_cap || !ctx->f_off) {
        CHK(dfree(ctx, ctx->f_off));
        CHK(dfree(ctx, ctx->f_offT));
        CHK(dfree(ctx, ctx->f_mq));
        CHK(dalloc(ctx, &ctx->f_off, 144 * (size_t)std::max(ctx->pair_cap, 1)));
        CHK(dalloc(ctx, &ctx->f_offT, 144 * (size_t)std::max(ctx->pair_cap, 1)));
        CHK(dalloc(ctx, &ctx->f_mq, 2 * 144 * (size_t)std::max(ctx->pair_cap, 1)));
        ctx->f_pair_cap = ctx->pair_cap;
    }
    return ABD_OK;
}

template <typename T>
void pcg_ctl(AbdCtx* ctx, PcgBufT<T>& B) {
    B.row_ptr = ctx->row_ptr; B.ent_key = ctx->ent_ks; B.ent_val = ctx->ent_vs;
    B.part = ctx->part; B.sc = ctx->pcg_sc; B.bnorm = ctx->bnorm; B.cg_iters = ctx->cg_iters; B.cg_done = ctx->cg_done;
    B.err = ctx->cnt + C_ERR;
    B.bar = ctx->pcg_ctl; B.done_cnt = (int*)(ctx->pcg_ctl + 1); B.fail = (int*)(ctx->pcg_ctl + 4);
    B.bnorm_out = ctx->bnorm; B.residual_out = ctx->C.residual;
    B.bref = nullptr;
    B.ts = (ctx->T.on && ctx->T.pcg_only) ? ctx->T.pcg_ts : nullptr;
}

// the PCG proper: one cooperative launch where possible, else one launch per iteration
template <typename T>
int run_pcg(AbdCtx* ctx, const PcgBufT<T>& B, bool warm) {
    bool ok = false;
    CHK(solve_persistent(ctx, B, &ok));
    if (ok) return ABD_OK;
    const int nt = ctx->ntiles;
    const double tol = ctx->P.cg_tol;
    HIPCHK(hipMemsetAsync(ctx->cg_iters, 0, sizeof(int) * ctx->ns, ctx->st));
    HIPCHK(hipMemsetAsync(ctx->cg_done, 0, sizeof(int) * ctx->ns, ctx->st));
    launch_pcg_cg(ctx->st, 0, nt, ctx->ns, ctx->tile_b0, ctx->tile_b1, ctx->tile_scene, ctx->scene_tile0, B, tol,
                  ctx->nb, ctx->pair_cap);
    CHK(dbg(ctx, "launch_pcg_cg(0)"));
    launch_pcg_tile_sum(ctx->st, ctx->ns, ctx->scene_tile0, ctx->part, nt, 2 * 2 + 0, ctx->bnorm, 1);
    CHK(dbg(ctx, "launch_pcg_tile_sum"));
    if (warm) hipLaunchKernelGGL(k_copy, dim3(cdiv(ctx->ns, 64)), dim3(64), 0, ctx->st, ctx->ns, ctx->nrm_b, ctx->bnorm);
    hipLaunchKernelGGL(k_copy, dim3(cdiv(ctx->ns, 64)), dim3(64), 0, ctx->st, ctx->ns, ctx->bnorm, ctx->C.residual);
    CHK(dbg(ctx, "k_copy"));
    // algorithmic bytes per launch: (nb + 2 pairs) 144 T + pairs 2 (144 T + 36 T) + nb (144 T + 168 T)
    const double t = sizeof(T);
    int poll = 4;
    for (int it = 0; it < ctx->P.cg_max_iter; ++it) {
        tic(ctx, "pcg_iter");
        launch_pcg_cg(ctx->st, it + 1, nt, ctx->ns, ctx->tile_b0, ctx->tile_b1, ctx->tile_scene, ctx->scene_tile0, B,
                      tol, ctx->nb, ctx->pair_cap);
        toc(ctx);
        CHK(time_bytes(ctx, nullptr, (double)ctx->nb * (144 * t + 144 * t + 168 * t), 2.0 * 144 * t + 2.0 * 180 * t));
        CHK(dbg(ctx, "launch_pcg_cg"));
        if (it + 1 == poll || it + 1 == ctx->P.cg_max_iter) {
            HIPCHK(hipMemsetAsync(ctx->cnt + C_ANY, 0, sizeof(int), ctx->st));
            launch_pcg_any(ctx->st, ctx->ns, nt, ctx->scene_tile0, ctx->part, it + 1, ctx->bnorm, tol, ctx->cg_done,
                           ctx->cnt + C_ANY);
            CHK(dbg(ctx, "launch_pcg_any"));
            CHK(sync_counters(ctx));
            if (!ctx->h_cnt[C_ANY]) break;
            poll += std::min(32, std::max(4, poll / 2));
        }
    }
    return ABD_OK;
}

int solve(AbdCtx* ctx) {
    if (ctx->P.solver == 1) {
        CHK(solve_direct(ctx, false));
        return solve_finish(ctx, false);
    }
    const bool warm = ctx->mlp != nullptr || (ctx->P.pcg_x0 == 1 && ctx->newton_it > 0);
    const double* b = ctx->grad;
    const double* bref = nullptr;
    if (warm) {
        // x0 from the model; PCG then solves H e = b - H x0 (stopping relative to ||b||), du = x0 + e
        CHK(warm_start(ctx));
        b = ctx->rhs;
        bref = ctx->nrm_b;
    }
    if (ctx->log_on && !ctx->mlp) {
        // logged features use the block-Jacobi inverse, as at inference (the warm start computed them)
        launch_body_features(ctx->st, ctx->G, ctx->q, ctx->qtilde, ctx->qdot, inc_cnt(ctx), ctx->inc_s, ctx->act_g,
                             act_cnt(ctx), ctx->pinv, ctx->grad, ctx->feat_raw);
        CHK(dbg(ctx, "launch_body_features"));
    }
    if (ctx->pre_mlp) CHK(learned_precond(ctx));  // after the features, which use D^-1
    if (ctx->P.precision == 1) {
        // mixed precision: the assembled FP64 system is rounded once to FP32 and solved in FP32
        // (partial sums and scalars stay FP64); the FP32 solution is widened into du
        CHK(ensure_f32(ctx));
        const size_t n12 = 12 * (size_t)ctx->nb;
        float* v = ctx->f_vec;
        PcgBuf32 B;
        pcg_ctl(ctx, B);
        B.bref = bref;
        B.b = v; B.x = v + n12; B.p = v + 2 * n12; B.u = v + 3 * n12;
        B.r[0] = v + 4 * n12; B.r[1] = v + 5 * n12; B.w[0] = v + 6 * n12; B.w[1] = v + 7 * n12;
        B.sv[0] = v + 8 * n12; B.sv[1] = v + 9 * n12;
        B.diag = ctx->f_diag; B.pinv = ctx->f_pinv; B.off = ctx->f_off; B.offT = ctx->f_offT; B.mq = ctx->f_mq;
        // the off-diagonal blocks are converted up to the pair bound (entries past the pair count are
        // never read)
        tic(ctx, "pcg_downcast", (double)(n12 + 288 * (size_t)ctx->nb + 288 * (size_t)ctx->pair_cap) * 12);
        launch_to_f32(ctx->st, (long)n12, b, v);
        launch_to_f32(ctx->st, 144 * (long)ctx->nb, ctx->diag, ctx->f_diag);
        launch_to_f32(ctx->st, 144 * (long)ctx->nb, ctx->pinv, ctx->f_pinv);
        launch_to_f32(ctx->st, 144 * (long)ctx->pair_cap, ctx->off, ctx->f_off);
        launch_to_f32(ctx->st, 144 * (long)ctx->pair_cap, ctx->offT, ctx->f_offT);
        toc(ctx);
        CHK(dbg(ctx, "launch_to_f32"));
        CHK(run_pcg(ctx, B, warm));
        launch_to_f64(ctx->st, (long)n12, B.x, ctx->du);
        CHK(dbg(ctx, "launch_to_f64"));
    } else {
        PcgBuf B;
        pcg_ctl(ctx, B);
        B.bref = bref;
        B.b = b; B.diag = ctx->diag; B.pinv = ctx->pinv; B.off = ctx->off; B.offT = ctx->offT; B.mq = ctx->mq;
        B.x = ctx->du; B.p = ctx->p0; B.u = ctx->z;
        B.r[0] = ctx->r; B.r[1] = ctx->r1; B.w[0] = ctx->Ap; B.w[1] = ctx->w1; B.sv[0] = ctx->s0; B.sv[1] = ctx->p1;
        CHK(run_pcg(ctx, B, warm));
    }
    CHK(solve_finish(ctx, warm));
    return ABD_OK;
}

// --------------------------------------------------------------------------------------------
// CCD bound + backtracking line search (line_search_iterative, gauss_newton.py:98-115)
// --------------------------------------------------------------------------------------------
// speculative: see cached_broad_phase.  x0_current: X0 already holds the vertices at q (the fast
// iteration's assembly transformed them and q has not moved since)
int ccd_bound(AbdCtx* ctx, bool speculative = false, bool x0_current = false) {
    if (!x0_current) {
        launch_xform(ctx->st, ctx->G, ctx->q, nullptr, nullptr, ctx->X0);
        CHK(dbg(ctx, "launch_xform"));
    }
    launch_xform(ctx->st, ctx->G, ctx->q, ctx->du, nullptr, ctx->X1);
    CHK(dbg(ctx, "launch_xform"));
    if (!ctx->iter_reset) CHK(fill_i32(ctx, ctx->cnt + C_OUTSIDE, 1, 0));
    CHK(broad_phase(ctx, ctx->X0, ctx->X1, ctx->du, speculative));
    if (!ctx->iter_reset) {
        hipLaunchKernelGGL(k_fill_u64, dim3(cdiv(ctx->ns, 128)), dim3(128), 0, ctx->st, ctx->ns, ctx->C.amax_bits,
                           (unsigned long long)0x3FF0000000000000ull);
        CHK(dbg(ctx, "k_fill_u64"));
    }
    if (ctx->P.ccd_enabled) {
        tic(ctx, "ccd", (double)ctx->nc_bound * (8 + 8 * 32));
        launch_ccd(ctx->st, ctx->G, cand_cnt(ctx), ctx->cand_s, ctx->X0, ctx->X1, ctx->P, ctx->C.amax_bits);
        CHK(dbg(ctx, "launch_ccd"));
        toc(ctx);
    }
    HIPCHK(hipGetLastError());
    return ABD_OK;
}

// the line search's start: E0 at q over the swept candidate list, alpha0 from the CCD bound
int ls_begin(AbdCtx* ctx) {
    // batches leave out the energies of scenes whose Newton loop has stopped (their search is decided)
    const int* act = ctx->ns > 1 ? ctx->C.newton_active : nullptr;
    launch_body_energy(ctx->st, ctx->G, ctx->P, ctx->q, nullptr, nullptr, ctx->qtilde, ctx->e_body, act);
    CHK(dbg(ctx, "launch_body_energy"));
    tic(ctx, "energy", (double)ctx->nc_bound * (8 + 4 * 32 + 8));
    launch_cand_eval(ctx->st, ctx->G, cand_cnt(ctx), ctx->cand_s, ctx->X0, ctx->P, nullptr, ctx->e_cand, act);
    CHK(dbg(ctx, "launch_cand_eval"));
    toc(ctx);
    launch_scene_sum_ls(ctx->st, ctx->G, ctx->e_body, ctx->e_cand, ctx->cand_lo, ctx->cand_hi, ctx->C.E0, 1, ctx->P,
                        ctx->C.newton_active, ctx->C, 0, nullptr);
    CHK(dbg(ctx, "launch_scene_sum(ls_init)"));
    return ABD_OK;
}

// `trials` consecutive trials of every undecided scene in one set of launches, the k-th at
// q - 2^-k alpha du (the steps the halvings of one trial per launch would reach), decided in order:
// accept / halve / floor (C_PENDING: some scene is still undecided after the last).  The decisions,
// energies and steps are those of one trial per launch; trials past a scene's decisive one are
// wasted work.  Xt holds the vertices of the k-th trial at Xt + k nv, so only with trials = 1 does
// it hold every scene's accepted iterate afterwards (the adaptive stiffness reads it).
int ls_trial(AbdCtx* ctx, int floor_accept, int trials = 1) {
    if (trials < 1 || trials > LS_TRIALS_MAX) {
        ctx->err = "line-search trials per launch out of range";
        return ABD_ERR_ARG;
    }
    launch_xform(ctx->st, ctx->G, ctx->q, ctx->du, ctx->C.alpha, ctx->Xt, trials);
    CHK(dbg(ctx, "launch_xform"));
    ctx->step_ls_trials += trials;
    tic(ctx, "energy", (double)trials * ctx->nc_bound * (8 + 4 * 32 + 8));
    const int* act = ctx->ns > 1 ? ctx->C.newton_active : nullptr;
    launch_cand_eval(ctx->st, ctx->G, cand_cnt(ctx), ctx->cand_s, ctx->Xt, ctx->P, nullptr, ctx->e_cand, act, trials,
                     (size_t)ctx->cand_cap);
    CHK(dbg(ctx, "launch_cand_eval"));
    toc(ctx);
    // qtmp (12 nb doubles) holds the body energies of up to 12 trials
    launch_body_energy(ctx->st, ctx->G, ctx->P, ctx->q, ctx->du, ctx->C.alpha, ctx->qtilde, ctx->qtmp, act, trials);
    CHK(dbg(ctx, "launch_body_energy"));
    if (!ctx->iter_reset) CHK(fill_i32(ctx, ctx->cnt + C_PENDING, 1, 0));
    launch_scene_sum_ls(ctx->st, ctx->G, ctx->qtmp, ctx->e_cand, ctx->cand_lo, ctx->cand_hi, ctx->C.E1, 2, ctx->P,
                        nullptr, ctx->C, floor_accept, ctx->cnt + C_PENDING, trials, (size_t)ctx->cand_cap);
    CHK(dbg(ctx, "launch_scene_sum(ls_decide)"));
    HIPCHK(hipGetLastError());
    return ABD_OK;
}

// trials until every scene has decided, one host round trip per launch of trials.  The launches
// after the first take ctx->ls_trials trials each (doubled after a launch that leaves a scene
// undecided, lowered by one after a search that one launch finished), so a scene whose searches
// need many halvings pays one round trip for them, and one that needs one pays for one trial.
// With the adaptive stiffness on (it reads the accepted vertices in Xt) or debug bit 128, one trial
// per launch.
int ls_until_done(AbdCtx* ctx, int floor_accept, int first_trial) {
    const bool batch = !(ctx->P.k_max > ctx->P.k_contact) && !(ctx->P.reserved0 & 128);
    int k = batch ? std::min(std::max(ctx->ls_trials, 1), LS_TRIALS_MAX) : 1;
    int launches = 0;
    for (int trial = first_trial; trial < 64;) {
        const int n = std::min(k, 64 - trial);
        CHK(ls_trial(ctx, floor_accept, n));
        trial += n;
        launches += 1;
        ctx->step_ls_syncs += 1;
        CHK(sync_counters(ctx));
        if (!ctx->h_cnt[C_PENDING]) break;
        if (batch) k = std::min(2 * k, LS_TRIALS_MAX);
    }
    if (batch) ctx->ls_trials = launches > 1 ? k : std::max(1, ctx->ls_trials - 1);
    return ABD_OK;
}

// floor_accept: the Newton line search keeps the floor step without a decrease (the reference); the
// predictor's search takes no step then
int line_search(AbdCtx* ctx, int floor_accept = 1) {
    CHK(ls_begin(ctx));
    CHK(ls_until_done(ctx, floor_accept, 0));
    launch_apply(ctx->st, ctx->G, ctx->q, ctx->du, ctx->C.alpha, ctx->C.newton_active, nullptr);
    CHK(dbg(ctx, "launch_apply"));
    HIPCHK(hipGetLastError());
    // the accepted iterate lies inside the swept boxes: the candidates stay valid
    ctx->cand_valid = true;
    return ABD_OK;
}

// adaptive stiffness after a Newton iteration: closest contact of each scene at the accepted iterate
// (Xt holds it: the last trial of the line search used every scene's accepted step); skip: see
// k_stiffness_update
int adapt_stiffness(AbdCtx* ctx, const int* skip = nullptr) {
    if (!(ctx->P.k_max > ctx->P.k_contact)) return ABD_OK;
    hipLaunchKernelGGL(k_fill_u64, dim3(cdiv(ctx->nb, 128)), dim3(128), 0, ctx->st, ctx->nb, ctx->C.dmin_bits,
                       (unsigned long long)0x7FF0000000000000ull);  // +inf
    launch_dmin(ctx->st, ctx->G, cand_cnt(ctx), ctx->cand_s, ctx->Xt, ctx->P, ctx->C.dmin_bits);
    launch_stiffness_update(ctx->st, ctx->G, ctx->P, ctx->C.newton_active, ctx->C, skip);
    CHK(dbg(ctx, "adapt_stiffness"));
    HIPCHK(hipGetLastError());
    return ABD_OK;
}

// features of every body at the current state (the step's start): an assembly, then the 49 body
// features of the warm start into feat_step
int step_features(AbdCtx* ctx) {
    if (!ctx->cand_valid) CHK(static_broad_phase(ctx));
    CHK(assemble_checked(ctx));
    launch_body_features(ctx->st, ctx->G, ctx->q, ctx->qtilde, ctx->qdot, inc_cnt(ctx), ctx->inc_s, ctx->act_g,
                         act_cnt(ctx), ctx->pinv, ctx->grad, ctx->feat_step);
    CHK(dbg(ctx, "launch_body_features"));
    return ABD_OK;
}

// first iterate of the step from the predictor (AbdParams.predictor 1-3): q moves from q_now toward
// the prediction by the largest CCD-safe halving step that lowers the energy
int predict(AbdCtx* ctx) {
    if (ctx->P.predictor == 3) {
        AbdMlp* m = ctx->pred_mlp;
        if (!m) {
            ctx->err = "predictor 3 needs a model (abd_set_predictor)";
            return ABD_ERR_STATE;
        }
        int rc = mlp_reserve(m, ctx->nb);
        if (rc != ABD_OK) {
            ctx->err = "learned predictor: " + m->err;
            return rc;
        }
        const int rows = (ctx->nb + 15) / 16 * 16;
        launch_features_pack(ctx->st, ctx->nb, rows, m->dp[0], ctx->feat_step, ctx->prfmean, ctx->prfistd, m->A[0]);
        mlp_forward_dev(m, ctx->nb, ctx->st);
        launch_step_pred(ctx->st, ctx->G, m->dp[m->L], m->Z[m->L - 1], ctx->promean, ctx->prostd, ctx->P.dt, ctx->q,
                         ctx->qnow, ctx->qdot, ctx->du);
        CHK(dbg(ctx, "learned predictor"));
    } else {
        launch_predict_du(ctx->st, ctx->G, ctx->P.predictor, ctx->P.dt, ctx->q, ctx->qnow, ctx->qdot, ctx->qtilde,
                          ctx->du);
        CHK(dbg(ctx, "launch_predict_du"));
    }
    CHK(ccd_bound(ctx));
    CHK(line_search(ctx, 0));
    hipLaunchKernelGGL(k_copy, dim3(cdiv(ctx->ns, 64)), dim3(64), 0, ctx->st, ctx->ns, ctx->C.alpha, ctx->C.pred_alpha);
    CHK(dbg(ctx, "predictor"));
    return ABD_OK;
}

// training rows of the step predictor: every unpinned body's features at the step's start and its
// step q - q_now
int log_step(AbdCtx* ctx) {
    const int nb = ctx->nb;
    std::vector<float> f((size_t)ABD_N_FEATURES * nb);
    std::vector<double> q(12 * (size_t)nb), qn(12 * (size_t)nb);
    std::vector<int> pinned(nb);
    HIPCHK(hipMemcpyAsync(f.data(), ctx->feat_step, f.size() * sizeof(float), hipMemcpyDeviceToHost, ctx->st));
    HIPCHK(hipMemcpyAsync(q.data(), ctx->q, q.size() * sizeof(double), hipMemcpyDeviceToHost, ctx->st));
    HIPCHK(hipMemcpyAsync(qn.data(), ctx->qnow, qn.size() * sizeof(double), hipMemcpyDeviceToHost, ctx->st));
    HIPCHK(hipMemcpyAsync(pinned.data(), ctx->pinned, nb * sizeof(int), hipMemcpyDeviceToHost, ctx->st));
    HIPCHK(hipStreamSynchronize(ctx->st));
    for (int b = 0; b < nb; ++b) {
        if (pinned[b]) continue;
        ctx->step_feat.insert(ctx->step_feat.end(), f.begin() + (size_t)ABD_N_FEATURES * b,
                              f.begin() + (size_t)ABD_N_FEATURES * (b + 1));
        for (int j = 0; j < 12; ++j) ctx->step_dq.push_back((float)(q[12 * (size_t)b + j] - qn[12 * (size_t)b + j]));
    }
    return ABD_OK;
}

// The kernels after the line search (q -= alpha du, adaptive stiffness, the convergence test, the
// report) and one copy of the staging block (report, Newton flags, counters, PCG control words) into
// pinned host memory.  tested: the fast iteration's skip test decides in k_apply_tested whether
// anything is committed (C_SKIP).
int commit_launches(AbdCtx* ctx, bool tested) {
    const int ns = ctx->ns;
    const int* skip = tested ? ctx->cnt + C_SKIP : nullptr;
    if (tested) {
        hipLaunchKernelGGL(k_apply_tested, dim3(cdiv(12 * ctx->nb, 256)), dim3(256), 0, ctx->st, ctx->G, ctx->q,
                           ctx->du, ctx->C.alpha, ctx->C.newton_active, ctx->cnt,
                           ctx->pcg_unchecked ? (const unsigned*)(ctx->pcg_ctl + 4) : (const unsigned*)nullptr,
                           ctx->act_cap, ctx->pair_cap);
        CHK(dbg(ctx, "k_apply_tested"));
    } else {
        launch_apply(ctx->st, ctx->G, ctx->q, ctx->du, ctx->C.alpha, ctx->C.newton_active, nullptr);
        CHK(dbg(ctx, "launch_apply"));
    }
    CHK(adapt_stiffness(ctx, skip));
    // keep the pre-check flags for the report, then stop converged scenes (abd.py:699)
    hipLaunchKernelGGL(k_commit_tail, dim3(ns), dim3(256), 0, ctx->st, ctx->G, ctx->P, ctx->C, ctx->act_before, skip,
                       ctx->cg_iters, ctx->nact_s, ctx->cand_lo, ctx->cand_hi, ctx->cnt, ctx->pcg_ctl, ctx->report);
    CHK(dbg(ctx, "k_commit_tail"));
    HIPCHK(hipMemcpyAsync(ctx->h_stage, ctx->report, stage_bytes(ns), hipMemcpyDeviceToHost, ctx->st));
    return ABD_OK;
}

// the iteration's one host round trip: wait for the copy of commit_launches and unpack it
int commit_wait(AbdCtx* ctx, std::vector<int>& now) {
    HIPCHK(hipStreamSynchronize(ctx->st));
    const int ns = ctx->ns;
    const int* si = (const int*)(ctx->h_stage + sizeof(double) * R_N * ns);
    std::memcpy(ctx->h_report, ctx->h_stage, sizeof(double) * R_N * ns);
    std::memcpy(ctx->h_now, si, sizeof(int) * ns);
    std::memcpy(ctx->h_cnt, si + ns, sizeof(int) * C_N);
    std::memcpy(ctx->h_pcg_ctl, si + ns + C_N, sizeof(unsigned) * 8);
    std::copy(ctx->h_now, ctx->h_now + ns, now.begin());
    return check_err_word(ctx);
}

int commit_iteration(AbdCtx* ctx, bool tested, std::vector<int>& now) {
    CHK(commit_launches(ctx, tested));
    return commit_wait(ctx, now);
}

// Every launch of a fast Newton iteration up to its report copies, with no host round trip: the
// assembly (lengths on the device), the solve (the persistent PCG's failure flag unread), the CCD
// bound over the candidate cache (its containment test unread), the line search's first trial, the
// skip flag, and the commit under it.
int fast_launches(AbdCtx* ctx) {
    // one kernel clears what the iteration's launches would clear one fill at a time
    hipLaunchKernelGGL(k_iter_reset, dim3(1), dim3(256), 0, ctx->st, ctx->ns, ctx->cnt, ctx->C.amax_bits,
                       ctx->cg_iters, ctx->pcg_ctl);
    CHK(dbg(ctx, "k_iter_reset"));
    struct Reset {
        AbdCtx* c;
        ~Reset() { c->iter_reset = false; }
    } reset{ctx};
    ctx->iter_reset = true;
    CHK(assemble(ctx));
    ctx->pcg_unchecked = false;
    ctx->defer_pcg_check = true;
    const int rc = solve(ctx);
    ctx->defer_pcg_check = false;
    CHK(rc);
    CHK(ccd_bound(ctx, true, true));
    CHK(ls_begin(ctx));
    CHK(ls_trial(ctx, 1));
    return commit_launches(ctx, true);
}

// Key of the fast iteration's launch sequence: every host value its launches bake in (buffer
// pointers through the allocation generation, capacities and bounds that size grids, the
// parameters passed by value, the PCG layout and form, the warm-start branch).
uint64_t graph_key(const AbdCtx* ctx) {
    uint64_t h = 1469598103934665603ull;
    auto mix = [&](const void* p, size_t n) {
        const unsigned char* b = (const unsigned char*)p;
        for (size_t i = 0; i < n; ++i) h = (h ^ b[i]) * 1099511628211ull;
    };
    auto mixi = [&](uint64_t v) { mix(&v, sizeof v); };
    mix(&ctx->P, sizeof ctx->P);
    mixi(ctx->alloc_gen);
    mixi((uint64_t)ctx->nc_bound); mixi((uint64_t)ctx->act_cap); mixi((uint64_t)ctx->pair_cap);
    mixi((uint64_t)ctx->cand_cap); mixi((uint64_t)ctx->cache_cap);
    mixi((uint64_t)ctx->sort_bytes); mixi((uint64_t)ctx->ntiles); mixi((uint64_t)ctx->bodies_per_tile);
    mixi((uint64_t)ctx->pcg_multi); mixi((uint64_t)ctx->coop_ok);
    mixi((uint64_t)(ctx->T.on && ctx->T.pcg_only)); mixi((uint64_t)(uintptr_t)ctx->T.pcg_ts);
    mixi((uint64_t)(ctx->P.pcg_x0 == 1 && ctx->newton_it > 0));
    return h | 1ull;  // never 0 (0 = no key)
}

// graph capture applies: the fast iteration, with the candidate cache valid (no traversal), the
// persistent PCG (one launch), no per-stage timing, debug checks or learned warm start and
// preconditioner, and not switched off (debug bit 6)
bool graph_usable(AbdCtx* ctx) {
    if (ctx->graph_off || (ctx->T.on && !ctx->T.pcg_only) || (ctx->P.reserved0 & (1 | 2 | 8 | 16 | 64))) return false;
    if (!(ctx->P.cand_margin > 0.0) || !ctx->cache_valid || !ctx->cand_valid) return false;
    if (ctx->mlp || ctx->pre_mlp || ctx->pcg_multi || !persist_ok(ctx)) return false;
    return true;
}

// the fast iteration's launches as one graph launch: eager the first time a key is seen, captured
// and instantiated the second time, replayed after that.  Host effects of the launch functions are
// redone on a replay (the detection count, the unread PCG flag).
int fast_graph(AbdCtx* ctx) {
    const uint64_t key = graph_key(ctx);
    if (ctx->gexec && key == ctx->gkey) {
        HIPCHK(hipGraphLaunch(ctx->gexec, ctx->st));
        ctx->step_graph_launches += 1;
        ctx->step_detections += 1;
        ctx->pcg_unchecked = true;
        ctx->assembled = ctx->solved = true;
        return ABD_OK;
    }
    if (key != ctx->gkey_seen) {
        ctx->gkey_seen = key;
        return fast_launches(ctx);
    }
    if (ctx->gexec) {
        (void)hipGraphExecDestroy(ctx->gexec);
        ctx->gexec = nullptr;
        ctx->gkey = 0;
    }
    const int det = ctx->step_detections;
    if (hipStreamBeginCapture(ctx->st, hipStreamCaptureModeThreadLocal) != hipSuccess) {
        (void)hipGetLastError();
        ctx->graph_off = true;
        return fast_launches(ctx);
    }
    ctx->capturing = true;
    const int rc = fast_launches(ctx);
    ctx->capturing = false;
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(ctx->st, &g);
    hipGraphExec_t x = nullptr;
    hipError_t ei = hipErrorNotSupported;  // not instantiated
    if (rc == ABD_OK && e == hipSuccess && g) ei = hipGraphInstantiate(&x, g, nullptr, nullptr, 0);
    if (g) (void)hipGraphDestroy(g);
    ctx->step_detections = det;
    if (rc != ABD_OK || e != hipSuccess || ei != hipSuccess || !x) {
        // the sequence could not be captured here: eager launches on this context from now on
        (void)hipGetLastError();
        if (x) (void)hipGraphExecDestroy(x);
        ctx->graph_off = true;
        ctx->err.clear();
        return fast_launches(ctx);
    }
    ctx->gexec = x;
    ctx->gkey = key;
    ctx->step_captures += 1;
    HIPCHK(hipGraphLaunch(ctx->gexec, ctx->st));
    ctx->step_detections += 1;
    return ABD_OK;
}

// One Newton iteration (abd.py:679-699).  fast: the persistent PCG's failure flag, the candidate
// cache's containment test and the first line-search trial are not waited for; the iteration is
// committed under the device flag C_SKIP and its report comes back in one round trip (the launches
// replayed from a HIP graph when graph_usable).  Only when the flag is set does the host redo the
// part that needs it (the solve with one launch per PCG iteration, a BVH traversal, the remaining
// halvings, or the whole iteration after growing the assembly's buffers) and commit again.
int newton_iteration(AbdCtx* ctx, bool fast, std::vector<int>& now) {
    if (!ctx->cand_valid) CHK(static_broad_phase(ctx));
    if (!fast) {
        CHK(assemble_checked(ctx));
        ctx->pcg_unchecked = false;
        CHK(solve(ctx));
        CHK(ccd_bound(ctx));
        CHK(ls_begin(ctx));
        CHK(ls_until_done(ctx, 1, 0));
        CHK(commit_iteration(ctx, false, now));  // applies the accepted step
        ctx->cand_valid = true;
        return ABD_OK;
    }
    CHK(graph_usable(ctx) ? fast_graph(ctx) : fast_launches(ctx));
    CHK(commit_wait(ctx, now));
    if (ctx->h_cnt[C_SKIP]) {
        ctx->step_slow += 1;
        if (assembly_overflow(ctx)) {
            // the active set outgrew the buffers: grow them and redo the iteration with round trips
            // between its stages
            CHK(grow_to_counts(ctx));
            ctx->pcg_unchecked = false;
            return newton_iteration(ctx, false, now);
        }
        if (ctx->pcg_unchecked && ctx->h_pcg_ctl[4]) {
            // a grid barrier did not fill: solve again with one launch per PCG iteration, and do not
            // try the persistent form again on this context
            ctx->coop_ok = false;
            ctx->pcg_fallbacks += 1;
            ctx->pcg_unchecked = false;
            ctx->force_iter_pcg = true;
            const int rc2 = solve(ctx);
            ctx->force_iter_pcg = false;
            CHK(rc2);
            CHK(ccd_bound(ctx));
            CHK(ls_begin(ctx));
            CHK(ls_until_done(ctx, 1, 0));
        } else if (ctx->h_cnt[C_OUTSIDE]) {
            ctx->cache_valid = false;  // a traversal over the path, then the line search again
            CHK(ccd_bound(ctx));
            CHK(ls_begin(ctx));
            CHK(ls_until_done(ctx, 1, 0));
        } else {
            CHK(ls_until_done(ctx, 1, 1));  // the halvings after the first trial
        }
        ctx->pcg_unchecked = false;
        CHK(commit_iteration(ctx, false, now));
    }
    ctx->pcg_unchecked = false;
    // the accepted iterate lies inside the swept boxes: the candidates stay valid
    ctx->cand_valid = true;
    return ABD_OK;
}

int step_once(AbdCtx* ctx) {
    const int ns = ctx->ns, mi = ctx->P.max_newton;
    launch_qtilde(ctx->st, ctx->nb, ctx->qnow, ctx->qdot, ctx->P.dt, ctx->P.gravity, ctx->qtilde);
    CHK(dbg(ctx, "launch_qtilde"));
    CHK(fill_i32(ctx, ctx->C.newton_active, ns, 1));
    ctx->log_res.assign((size_t)ns * mi, 0.0);
    ctx->log_alpha.assign((size_t)ns * mi, 0.0);
    ctx->log_cg.assign((size_t)ns * mi, 0);
    ctx->stats.assign(ns, AbdStepStats{});
    for (auto& s : ctx->stats) { s.alpha_min = 1.0; s.ccd_alpha_min = 1.0; }
    ctx->pcg_fallbacks = 0;
    ctx->step_traversals = ctx->step_detections = ctx->step_slow = 0;
    ctx->step_graph_launches = ctx->step_captures = 0;
    ctx->step_ls_trials = ctx->step_ls_syncs = 0;
    hipLaunchKernelGGL(k_fill_u64, dim3(cdiv(ctx->nb, 128)), dim3(128), 0, ctx->st, ctx->nb,
                       (unsigned long long*)ctx->C.dmin_prev, (unsigned long long)0x7FF0000000000000ull);  // +inf
    CHK(fill_i32(ctx, ctx->C.k_raises, ns, 0));
    hipLaunchKernelGGL(k_fill_u64, dim3(cdiv(ns, 128)), dim3(128), 0, ctx->st, ns,
                       (unsigned long long*)ctx->C.pred_alpha, 0ull);
    if (ctx->log_steps || ctx->P.predictor == 3) CHK(step_features(ctx));
    if (ctx->P.predictor > 0) CHK(predict(ctx));
    // one host round trip per Newton iteration unless a mode needs the state between stages (the
    // direct solver, sample logging)
    const bool fast = ctx->P.solver == 0 && !ctx->log_on;
    std::vector<int> now(ns);
    for (int it = 0; it < mi; ++it) {
        ctx->newton_it = it;
        CHK(newton_iteration(ctx, fast, now));
        bool any = false;
        for (int s = 0; s < ns; ++s) {
            const double* R = ctx->h_report + (size_t)s * R_N;
            if (R[R_ACTIVE] != 0.0) {
                AbdStepStats& S = ctx->stats[s];
                ctx->log_res[(size_t)s * mi + it] = R[R_RES];
                ctx->log_alpha[(size_t)s * mi + it] = R[R_ALPHA];
                ctx->log_cg[(size_t)s * mi + it] = (int)R[R_CG];
                if (it == 0) S.residual_first = R[R_RES];
                S.residual_last = R[R_RES];
                S.newton_iters = it + 1;
                S.cg_iters += (int)R[R_CG];
                S.ls_floor_hits += (int)R[R_FLOOR];
                S.n_active = (int)R[R_NACT];
                S.n_candidates = (int)R[R_NCAND];
                S.alpha_min = std::min(S.alpha_min, R[R_ALPHA]);
                S.ccd_alpha_min = std::min(S.ccd_alpha_min, R[R_ALPHA0]);
                S.energy = R[R_E1];
                if (now[s] == 0) {
                    if (R[R_RES] < ctx->P.newton_tol) S.converged = 1;
                    else S.stalled = 1;
                }
            }
            ctx->stats[s].k_contact = R[R_KC];
            ctx->stats[s].k_raises = (int)R[R_RAISES];
            ctx->stats[s].predictor_alpha = R[R_PALPHA];
            ctx->stats[s].pcg_fallbacks = ctx->pcg_fallbacks;
            ctx->stats[s].traversals = ctx->step_traversals;
            ctx->stats[s].detections = ctx->step_detections;
            ctx->stats[s].slow_iterations = ctx->step_slow;
            ctx->stats[s].graph_launches = ctx->step_graph_launches;
            ctx->stats[s].graph_captures = ctx->step_captures;
            ctx->stats[s].ls_trials = ctx->step_ls_trials;
            ctx->stats[s].ls_round_trips = ctx->step_ls_syncs;
            any |= now[s] != 0;
        }
        if (!any) break;
    }
    ctx->newton_it = -1;
    if (ctx->log_steps) CHK(log_step(ctx));
    launch_forward(ctx->st, ctx->nb, ctx->P.dt, ctx->q, ctx->qnow, ctx->qdot);
    CHK(dbg(ctx, "launch_forward"));
    // the inertial target of the next step, so that the hooks called between steps (compute_g,
    // assemble) see the new state's potential, as the reference's compute_g does
    launch_qtilde(ctx->st, ctx->nb, ctx->qnow, ctx->qdot, ctx->P.dt, ctx->P.gravity, ctx->qtilde);
    CHK(dbg(ctx, "launch_qtilde"));
    HIPCHK(hipGetLastError());
    return ABD_OK;
}

int check_ctx(AbdCtx* ctx) { return ctx ? ABD_OK : ABD_ERR_ARG; }

// every body's contact stiffness back to k_contact
int reset_stiffness(AbdCtx* ctx) {
    std::vector<double> k(ctx->nb, ctx->P.k_contact);
    HIPCHK(hipMemcpyAsync(ctx->C.kb, k.data(), ctx->nb * sizeof(double), hipMemcpyHostToDevice, ctx->st));
    launch_scene_kmax(ctx->st, ctx->G, ctx->C);
    HIPCHK(hipStreamSynchronize(ctx->st));
    return ABD_OK;
}

// --------------------------------------------------------------------------------------------
// PCG tiles: runs of at most `bpt` consecutive bodies of one scene (one wave each, 12 lanes per
// body); the per-tile partial sums are sized with them
// --------------------------------------------------------------------------------------------
int set_tiles(AbdCtx* ctx, int bpt) {
    if (bpt < 1 || bpt > 5) {
        ctx->err = "bodies per PCG tile must be in [1, 5]";
        return ABD_ERR_ARG;
    }
    const int ns = ctx->ns;
    const std::vector<int>& sb = ctx->h_scene_b0;
    std::vector<int> tb0, tb1, tsc, st0(ns + 1, 0);
    for (int s = 0; s < ns; ++s) {
        st0[s] = (int)tb0.size();
        for (int b = sb[s]; b < sb[s + 1]; b += bpt) {
            tb0.push_back(b);
            tb1.push_back(std::min(b + bpt, sb[s + 1]));
            tsc.push_back(s);
        }
    }
    st0[ns] = (int)tb0.size();
    if (ctx->st) HIPCHK(hipStreamSynchronize(ctx->st));
    CHK(dfree(ctx, ctx->tile_b0)); CHK(dfree(ctx, ctx->tile_b1)); CHK(dfree(ctx, ctx->tile_scene));
    CHK(dfree(ctx, ctx->scene_tile0)); CHK(dfree(ctx, ctx->part));
    const int nt = (int)tb0.size();
    CHK(dalloc(ctx, &ctx->tile_b0, nt)); CHK(dalloc(ctx, &ctx->tile_b1, nt)); CHK(dalloc(ctx, &ctx->tile_scene, nt));
    CHK(dalloc(ctx, &ctx->scene_tile0, ns + 1)); CHK(dalloc(ctx, &ctx->part, 6 * (size_t)nt));
    HIPCHK(hipMemcpy(ctx->tile_b0, tb0.data(), sizeof(int) * nt, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(ctx->tile_b1, tb1.data(), sizeof(int) * nt, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(ctx->tile_scene, tsc.data(), sizeof(int) * nt, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(ctx->scene_tile0, st0.data(), sizeof(int) * (ns + 1), hipMemcpyHostToDevice));
    ctx->ntiles = nt;
    ctx->h_tile_b0 = tb0;
    ctx->h_scene_tile0 = st0;
    ctx->bodies_per_tile = bpt;
    return ABD_OK;
}

// --------------------------------------------------------------------------------------------
// rest-frame BVHs (host, once per distinct mesh)
// --------------------------------------------------------------------------------------------
float f_rd(double x) {
    float f = (float)x;
    if ((double)f > x) f = std::nextafter(f, -HUGE_VALF);
    return f;
}
float f_ru(double x) {
    float f = (float)x;
    if ((double)f < x) f = std::nextafter(f, HUGE_VALF);
    return f;
}

struct PrimBox {
    double lo[3], hi[3], c[3];
};

// Top-down build over [lo, hi) of `order`: median split of the centroids along the longest axis
// of their box (depth ceil(log2 n) + 1).  Nodes are appended to `out` in preorder, so an internal
// node's first child is the next node; `skip` is the mesh-local index one past the node's subtree
// (where a traversal continues after rejecting or finishing the node).  `base` is the index of the
// mesh's root in `out`.
void bvh_build_range(const std::vector<PrimBox>& pb, std::vector<int>& order, int lo, int hi, int base,
                     std::vector<BNode>& out) {
    const int me = (int)out.size();
    out.push_back(BNode{});
    double blo[3] = {HUGE_VAL, HUGE_VAL, HUGE_VAL}, bhi[3] = {-HUGE_VAL, -HUGE_VAL, -HUGE_VAL};
    double clo[3] = {HUGE_VAL, HUGE_VAL, HUGE_VAL}, chi[3] = {-HUGE_VAL, -HUGE_VAL, -HUGE_VAL};
    for (int i = lo; i < hi; ++i) {
        const PrimBox& b = pb[order[i]];
        for (int k = 0; k < 3; ++k) {
            blo[k] = std::min(blo[k], b.lo[k]); bhi[k] = std::max(bhi[k], b.hi[k]);
            clo[k] = std::min(clo[k], b.c[k]); chi[k] = std::max(chi[k], b.c[k]);
        }
    }
    BNode nd;
    for (int k = 0; k < 3; ++k) { nd.lo[k] = f_rd(blo[k]); nd.hi[k] = f_ru(bhi[k]); }
    if (hi - lo == 1) {
        nd.prim = order[lo];
    } else {
        int ax = 0;
        for (int k = 1; k < 3; ++k)
            if (chi[k] - clo[k] > chi[ax] - clo[ax]) ax = k;
        const int mid = lo + (hi - lo) / 2;
        std::nth_element(order.begin() + lo, order.begin() + mid, order.begin() + hi, [&](int a, int b) {
            if (pb[a].c[ax] != pb[b].c[ax]) return pb[a].c[ax] < pb[b].c[ax];
            return a < b;
        });
        nd.prim = -1;
        bvh_build_range(pb, order, lo, mid, base, out);
        bvh_build_range(pb, order, mid, hi, base, out);
    }
    nd.skip = (int)out.size() - base;
    out[me] = nd;
}

// BVH over primitives given as vertex-index tuples (arity 2 or 3) into `xr` (rest coords, xyz),
// appended to `out` (2n - 1 nodes; none if n = 0).
void bvh_build(const double* xr, const int* prim, int arity, int n, std::vector<BNode>& out) {
    if (n <= 0) return;
    std::vector<PrimBox> pb(n);
    for (int i = 0; i < n; ++i) {
        PrimBox& b = pb[i];
        for (int k = 0; k < 3; ++k) { b.lo[k] = HUGE_VAL; b.hi[k] = -HUGE_VAL; b.c[k] = 0.0; }
        for (int a = 0; a < arity; ++a) {
            const double* x = xr + 3 * (size_t)prim[arity * i + a];
            for (int k = 0; k < 3; ++k) {
                b.lo[k] = std::min(b.lo[k], x[k]);
                b.hi[k] = std::max(b.hi[k], x[k]);
                b.c[k] += x[k] / arity;
            }
        }
    }
    std::vector<int> order(n);
    for (int i = 0; i < n; ++i) order[i] = i;
    bvh_build_range(pb, order, 0, n, (int)out.size(), out);
}

uint64_t fnv1a(uint64_t h, const void* p, size_t n) {
    const unsigned char* c = (const unsigned char*)p;
    for (size_t i = 0; i < n; ++i) { h ^= c[i]; h *= 1099511628211ull; }
    return h;
}

}  // namespace

// =============================================================================================
// C ABI
// =============================================================================================
extern "C" {

int abd_create(const AbdSceneDesc* d, const AbdParams* p, AbdCtx** out) {
    if (!d || !p || !out) return ABD_ERR_ARG;
    *out = nullptr;
    AbdCtx* ctx = new AbdCtx();
    auto fail = [&](int rc) {
        // keep the context so the caller can read the message, but mark it unusable
        *out = ctx;
        return rc;
    };
    ctx->P = *p;
    {
        double* n = ctx->P.ground_normal;
        double l = std::sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
        if (l > 0) { n[0] /= l; n[1] /= l; n[2] /= l; }
    }
    if (check_params(ctx, p) != ABD_OK) return fail(ABD_ERR_ARG);
    ctx->nb = d->n_bodies; ctx->nv = d->n_verts; ctx->ne = d->n_edges; ctx->nt = d->n_tris; ctx->ns = d->n_scenes;
    const int nb = ctx->nb, nv = ctx->nv, ne = ctx->ne, nt = ctx->nt, ns = ctx->ns;
    if (nb <= 0 || nv <= 0 || ns <= 0 || ns > 65535 || nb >= (1 << 20)) {
        ctx->err = "invalid scene sizes";
        return fail(ABD_ERR_ARG);
    }
    // host validation and scene tables
    std::vector<int> sb(ns + 1, 0), sv(ns + 1, 0), se(ns + 1, 0), stt(ns + 1, 0);
    for (int b = 0; b < nb; ++b) {
        int s = d->body_scene[b];
        if (s < 0 || s >= ns || (b > 0 && s < d->body_scene[b - 1])) {
            ctx->err = "body_scene must be non-decreasing in [0, n_scenes)";
            return fail(ABD_ERR_ARG);
        }
        sb[s + 1]++;
    }
    std::vector<double4> rest(nv);
    for (int v = 0; v < nv; ++v) {
        int b = d->vert_body[v];
        if (b < 0 || b >= nb || (v > 0 && b < d->vert_body[v - 1])) {
            ctx->err = "vert_body must be non-decreasing in [0, n_bodies)";
            return fail(ABD_ERR_ARG);
        }
        rest[v] = make_double4(d->rest_x[3 * v], d->rest_x[3 * v + 1], d->rest_x[3 * v + 2], (double)b);
        sv[d->body_scene[b] + 1]++;
    }
    std::vector<int2> edges(std::max(ne, 1));
    for (int e = 0; e < ne; ++e) {
        int a = d->edges[2 * e], b = d->edges[2 * e + 1];
        if (a < 0 || a >= nv || b < 0 || b >= nv || d->vert_body[a] != d->vert_body[b]) {
            ctx->err = "edge indices out of range or spanning two bodies";
            return fail(ABD_ERR_ARG);
        }
        int s = d->body_scene[d->vert_body[a]];
        if (e > 0 && s < d->body_scene[d->vert_body[d->edges[2 * e - 2]]]) {
            ctx->err = "edges must be grouped by scene";
            return fail(ABD_ERR_ARG);
        }
        edges[e] = make_int2(a, b);
        se[s + 1]++;
    }
    std::vector<int4> tris(std::max(nt, 1));
    for (int t = 0; t < nt; ++t) {
        int a = d->tris[3 * t], b = d->tris[3 * t + 1], c = d->tris[3 * t + 2];
        if (a < 0 || a >= nv || b < 0 || b >= nv || c < 0 || c >= nv || d->vert_body[a] != d->vert_body[b] ||
            d->vert_body[a] != d->vert_body[c]) {
            ctx->err = "triangle indices out of range or spanning two bodies";
            return fail(ABD_ERR_ARG);
        }
        int s = d->body_scene[d->vert_body[a]];
        if (t > 0 && s < d->body_scene[d->vert_body[d->tris[3 * t - 3]]]) {
            ctx->err = "triangles must be grouped by scene";
            return fail(ABD_ERR_ARG);
        }
        tris[t] = make_int4(a, b, c, d->vert_body[a]);
        stt[s + 1]++;
    }
    for (int s = 0; s < ns; ++s) {
        sb[s + 1] += sb[s]; sv[s + 1] += sv[s]; se[s + 1] += se[s]; stt[s + 1] += stt[s];
        if (sv[s + 1] - sv[s] > (int)KEY_IMASK || se[s + 1] - se[s] > (int)KEY_IMASK || stt[s + 1] - stt[s] > (int)KEY_IMASK) {
            ctx->err = "a scene exceeds 2^23 primitives";
            return fail(ABD_ERR_ARG);
        }
    }
    ctx->h_scene_b0 = sb; ctx->h_scene_v0 = sv; ctx->h_scene_e0 = se; ctx->h_scene_t0 = stt;
    // per-body primitive ranges: edges and triangles must be grouped by body like the vertices
    std::vector<int> bv0(nb + 1, 0), be0(nb + 1, 0), bt0(nb + 1, 0);
    for (int v = 0; v < nv; ++v) bv0[d->vert_body[v] + 1]++;
    for (int e = 0; e < ne; ++e) {
        int b = d->vert_body[d->edges[2 * e]];
        if (e > 0 && b < d->vert_body[d->edges[2 * e - 2]]) {
            ctx->err = "edges must be grouped by body";
            return fail(ABD_ERR_ARG);
        }
        be0[b + 1]++;
    }
    for (int t = 0; t < nt; ++t) {
        int b = d->vert_body[d->tris[3 * t]];
        if (t > 0 && b < d->vert_body[d->tris[3 * t - 3]]) {
            ctx->err = "triangles must be grouped by body";
            return fail(ABD_ERR_ARG);
        }
        bt0[b + 1]++;
    }
    for (int b = 0; b < nb; ++b) { bv0[b + 1] += bv0[b]; be0[b + 1] += be0[b]; bt0[b + 1] += bt0[b]; }
    {
        long long mv = 0, me = 0;
        for (int b = 0; b < nb; ++b) {
            mv = std::max<long long>(mv, bv0[b + 1] - bv0[b]);
            me = std::max<long long>(me, be0[b + 1] - be0[b]);
        }
        const long long mq = 2 * mv + me;
        if (mq > 64LL * 65535) {
            ctx->err = "a body has too many vertices and edges for the broad-phase launch shape";
            return fail(ABD_ERR_ARG);
        }
        ctx->max_queries = (int)mq;
    }
    // distinct local meshes (same rest coordinates and local connectivity) share one BVH pair
    std::vector<int> bmesh(nb), mesh_rep, mtn0(1, 0), men0(1, 0);
    std::vector<BNode> tnodes, enodes;
    {
        auto local = [&](int b, int arity) {
            const int* src = arity == 3 ? d->tris : d->edges;
            const int p0 = arity == 3 ? bt0[b] : be0[b], p1 = arity == 3 ? bt0[b + 1] : be0[b + 1];
            std::vector<int> out((size_t)(p1 - p0) * arity);
            for (size_t i = 0; i < out.size(); ++i) out[i] = src[(size_t)p0 * arity + i] - bv0[b];
            return out;
        };
        std::map<uint64_t, std::vector<int>> by_hash;
        for (int b = 0; b < nb; ++b) {
            const int cnts[3] = {bv0[b + 1] - bv0[b], be0[b + 1] - be0[b], bt0[b + 1] - bt0[b]};
            std::vector<int> lt = local(b, 3), le = local(b, 2);
            const double* xr = d->rest_x + 3 * (size_t)bv0[b];
            uint64_t h = 1469598103934665603ull;
            h = fnv1a(h, cnts, sizeof cnts);
            h = fnv1a(h, xr, sizeof(double) * 3 * (size_t)cnts[0]);
            h = fnv1a(h, lt.data(), sizeof(int) * lt.size());
            h = fnv1a(h, le.data(), sizeof(int) * le.size());
            int found = -1;
            for (int m : by_hash[h]) {
                const int r = mesh_rep[m];
                if (bv0[r + 1] - bv0[r] != cnts[0] || be0[r + 1] - be0[r] != cnts[1] || bt0[r + 1] - bt0[r] != cnts[2])
                    continue;
                if (std::memcmp(d->rest_x + 3 * (size_t)bv0[r], xr, sizeof(double) * 3 * (size_t)cnts[0]) != 0) continue;
                if (local(r, 3) != lt || local(r, 2) != le) continue;
                found = m;
                break;
            }
            if (found < 0) {
                found = (int)mesh_rep.size();
                mesh_rep.push_back(b);
                by_hash[h].push_back(found);
                bvh_build(xr, lt.data(), 3, cnts[2], tnodes);
                bvh_build(xr, le.data(), 2, cnts[1], enodes);
                mtn0.push_back((int)tnodes.size());
                men0.push_back((int)enodes.size());
            }
            bmesh[b] = found;
        }
    }
    ctx->n_meshes = (int)mesh_rep.size();
    {
        hipError_t e = hipStreamCreateWithFlags(&ctx->st, hipStreamNonBlocking);
        if (e != hipSuccess) {
            ctx->st = nullptr;
            ctx->err = std::string("hipStreamCreateWithFlags: ") + hipGetErrorName(e) + ": " + hipGetErrorString(e);
            return fail(ABD_ERR_HIP);
        }
    }
    int rc = ABD_OK;
#define A(ptr, n) if ((rc = dalloc(ctx, &ctx->ptr, (n))) != ABD_OK) return fail(rc)
#define U(ptr, src, n) if (hipMemcpy(ctx->ptr, (src), sizeof(*ctx->ptr) * (size_t)(n), hipMemcpyHostToDevice) != hipSuccess) { ctx->err = "upload failed"; return fail(ABD_ERR_HIP); }
    A(rest, nv); U(rest, rest.data(), nv);
    A(edges, std::max(ne, 1)); U(edges, edges.data(), std::max(ne, 1));
    A(tris, std::max(nt, 1)); U(tris, tris.data(), std::max(nt, 1));
    A(mass4, 16 * nb); U(mass4, d->mass4, 16 * nb);
    A(pinned, nb); U(pinned, d->body_pinned, nb);
    A(bscene, nb); U(bscene, d->body_scene, nb);
    A(scene_b0, ns + 1); U(scene_b0, sb.data(), ns + 1);
    A(scene_v0, ns + 1); U(scene_v0, sv.data(), ns + 1);
    A(scene_e0, ns + 1); U(scene_e0, se.data(), ns + 1);
    A(scene_t0, ns + 1); U(scene_t0, stt.data(), ns + 1);
    A(q, 12 * nb); A(qnow, 12 * nb); A(qdot, 12 * nb); A(qtilde, 12 * nb); A(qtmp, 12 * nb);
    A(du, 12 * nb); A(grad, 12 * nb); A(diag, 144 * nb); A(pinv, 144 * nb); A(e_body, nb);
    A(X0, nv); A(X1, nv); A(Xt, (size_t)LS_TRIALS_MAX * nv); A(Xc0, nv); A(Xc1, nv);
    A(body_v0, nb + 1); U(body_v0, bv0.data(), nb + 1);
    A(body_e0, nb + 1); U(body_e0, be0.data(), nb + 1);
    A(body_t0, nb + 1); U(body_t0, bt0.data(), nb + 1);
    A(body_mesh, nb); U(body_mesh, bmesh.data(), nb);
    A(mesh_tn0, ctx->n_meshes + 1); U(mesh_tn0, mtn0.data(), ctx->n_meshes + 1);
    A(mesh_en0, ctx->n_meshes + 1); U(mesh_en0, men0.data(), ctx->n_meshes + 1);
    A(tri_nodes, std::max<size_t>(tnodes.size(), 1)); if (!tnodes.empty()) { U(tri_nodes, tnodes.data(), tnodes.size()); }
    A(edge_nodes, std::max<size_t>(enodes.size(), 1)); if (!enodes.empty()) { U(edge_nodes, enodes.data(), enodes.size()); }
    {
        // world-space copies of every body's node boxes, refreshed by each detection
        std::vector<long long> tw(nb + 1, 0), ew(nb + 1, 0);
        for (int b = 0; b < nb; ++b) {
            const int m = bmesh[b];
            tw[b + 1] = tw[b] + (mtn0[m + 1] - mtn0[m]);
            ew[b + 1] = ew[b] + (men0[m + 1] - men0[m]);
        }
        if (tw[nb] > INT_MAX || ew[nb] > INT_MAX) {
            ctx->err = "more than 2^31 BVH nodes over all bodies";
            return fail(ABD_ERR_ARG);
        }
        std::vector<int> tw32(tw.begin(), tw.end()), ew32(ew.begin(), ew.end());
        A(body_tw0, nb + 1); U(body_tw0, tw32.data(), nb + 1);
        A(body_ew0, nb + 1); U(body_ew0, ew32.data(), nb + 1);
        A(tri_world, std::max<long long>(tw[nb], 1));
        A(edge_world, std::max<long long>(ew[nb], 1));
    }
    A(bp_cnt, nb + 1); A(bp_off, nb + 1);
    A(body_box, 6 * (size_t)nb);
    ctx->M = BodyMesh{ctx->body_v0, ctx->body_e0, ctx->body_t0, ctx->body_mesh, ctx->mesh_tn0, ctx->mesh_en0,
                      ctx->body_tw0, ctx->body_ew0, ctx->tri_nodes, ctx->edge_nodes, ctx->tri_world, ctx->edge_world};
    A(cand_lo, ns + 1);
    ctx->cand_hi = ctx->cand_lo + 1;
    A(r, 12 * nb); A(z, 12 * nb); A(p0, 12 * nb); A(p1, 12 * nb); A(Ap, 12 * nb);
    A(r1, 12 * nb); A(w1, 12 * nb); A(s0, 12 * nb); A(pcg_sc, 4 * ns); A(pcg_ctl, 8);
    A(feat_raw, (size_t)ABD_N_FEATURES * nb); A(fmean, ABD_N_FEATURES); A(fistd, ABD_N_FEATURES); A(omean, 12);
    A(ostd, 12); A(x0, 12 * nb); A(rhs, 12 * nb); A(nrm_b, ns); A(nrm_r, ns); A(warm_used, ns);
    A(pfmean, ABD_N_FEATURES); A(pfistd, ABD_N_FEATURES); A(pomean, ABD_N_PRECOND); A(postd, ABD_N_PRECOND);
    A(prfmean, ABD_N_FEATURES); A(prfistd, ABD_N_FEATURES); A(promean, 12); A(prostd, 12);
    A(feat_step, (size_t)ABD_N_FEATURES * nb);
    if ((rc = set_tiles(ctx, 5)) != ABD_OK) return fail(rc);
    A(bnorm, ns); A(cg_iters, ns); A(cg_done, ns);
    A(C.newton_active, ns); A(C.residual, ns); A(C.alpha, ns); A(C.alpha0, ns); A(C.amax_bits, ns);
    A(C.ls_done, ns); A(C.ls_floor, ns); A(C.E0, ns); A(C.E1, ns);
    A(C.kb, nb); A(C.dmin_prev, nb); A(C.dmin_bits, nb); A(C.k_raises, ns); A(C.kc, ns); A(C.pred_alpha, ns);
    A(nact_s, ns); A(act_before, ns); A(report, R_N * ns + (ns + C_N + 8 + 1) / 2); A(cnt, C_N); A(row_ptr, nb + 1); A(inc_ptr, nb + 1);
#undef A
#undef U
    if (hipHostMalloc((void**)&ctx->h_pcg_ctl, 8 * sizeof(unsigned), hipHostMallocDefault) != hipSuccess ||
        hipHostMalloc((void**)&ctx->h_cnt, C_N * sizeof(int), hipHostMallocDefault) != hipSuccess ||
        hipHostMalloc((void**)&ctx->h_report, R_N * ns * sizeof(double), hipHostMallocDefault) != hipSuccess ||
        hipHostMalloc((void**)&ctx->h_now, ns * sizeof(int), hipHostMallocDefault) != hipSuccess ||
        hipHostMalloc((void**)&ctx->h_stage, stage_bytes(ns), hipHostMallocDefault) != hipSuccess) {
        ctx->err = "hipHostMalloc failed";
        return fail(ABD_ERR_HIP);
    }
    if ((rc = grow_candidates(ctx, std::max(1024, std::min(p->max_candidates > 0 ? p->max_candidates : (1 << 20), 1 << 26)))) != ABD_OK) return fail(rc);
    if ((rc = grow_active(ctx, 1024)) != ABD_OK) return fail(rc);
    if ((rc = grow_cache(ctx, 1024)) != ABD_OK) return fail(rc);
    if ((rc = grow_pairs(ctx, 1024)) != ABD_OK) return fail(rc);
    if ((rc = grow_body_pairs(ctx, std::max(1024, std::min(16 * nb, 1 << 24)))) != ABD_OK) return fail(rc);
    if (hipMemset(ctx->bp_cnt, 0, (nb + 1) * sizeof(int)) != hipSuccess) {
        ctx->err = "hipMemset failed";
        return fail(ABD_ERR_HIP);
    }
    ctx->G = Geo{nb, nv, ne, nt, ns, ctx->rest, ctx->edges, ctx->tris, ctx->mass4, ctx->pinned, ctx->bscene,
                 ctx->scene_b0, ctx->scene_v0, ctx->scene_e0, ctx->scene_t0, ctx->cnt + C_ERR, ctx->C.kb};
    if ((rc = reset_stiffness(ctx)) != ABD_OK) return fail(rc);
    if (hipMemset(ctx->cnt, 0, C_N * sizeof(int)) != hipSuccess) {
        ctx->err = "hipMemset failed";
        return fail(ABD_ERR_HIP);
    }
    // identity initial state
    std::vector<double> q0(12 * (size_t)nb, 0.0);
    for (int b = 0; b < nb; ++b) q0[12 * b + 3] = q0[12 * b + 7] = q0[12 * b + 11] = 1.0;
    std::vector<double> zero(12 * (size_t)nb, 0.0);
    if ((rc = abd_set_state(ctx, q0.data(), q0.data(), zero.data())) != ABD_OK) return fail(rc);
    *out = ctx;
    return ABD_OK;
}

void abd_destroy(AbdCtx* ctx) {
    if (!ctx) return;
    if (ctx->st) (void)hipStreamSynchronize(ctx->st);
    for (void* p : ctx->allocs) (void)hipFree(p);
    for (auto e : ctx->T.pool) (void)hipEventDestroy(e);
    for (auto& r : ctx->T.open) { (void)hipEventDestroy(r.a); (void)hipEventDestroy(r.b); }
    if (ctx->T.kslots) (void)hipFree(ctx->T.kslots);
    if (ctx->T.h_kslots) (void)hipHostFree(ctx->T.h_kslots);
    if (ctx->h_cnt) (void)hipHostFree(ctx->h_cnt);
    if (ctx->h_pcg_ctl) (void)hipHostFree(ctx->h_pcg_ctl);
    if (ctx->h_report) (void)hipHostFree(ctx->h_report);
    if (ctx->h_now) (void)hipHostFree(ctx->h_now);
    if (ctx->h_stage) (void)hipHostFree(ctx->h_stage);
    if (ctx->gexec) (void)hipGraphExecDestroy(ctx->gexec);
    if (ctx->st) (void)hipStreamDestroy(ctx->st);
    delete ctx;
}

const char* abd_last_error(const AbdCtx* ctx) { return ctx ? ctx->err.c_str() : "null context"; }

int abd_get_params(const AbdCtx* ctx, AbdParams* out) {
    if (!ctx || !out) return ABD_ERR_ARG;
    *out = ctx->P;
    return ABD_OK;
}

int abd_set_params(AbdCtx* ctx, const AbdParams* p) {
    if (!ctx || !p) return ABD_ERR_ARG;
    CHK(check_params(ctx, p));
    ctx->P = *p;
    double* n = ctx->P.ground_normal;
    double l = std::sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
    if (l > 0) { n[0] /= l; n[1] /= l; n[2] /= l; }
    ctx->cand_valid = false;
    ctx->cache_valid = false;  // margin, dhat or the ground plane may have changed
    ctx->assembled = ctx->solved = false;
    return reset_stiffness(ctx);
}

int abd_set_state(AbdCtx* ctx, const double* q, const double* q_now, const double* qdot) {
    CHK(check_ctx(ctx));
    size_t n = 12 * (size_t)ctx->nb * sizeof(double);
    for (const double* a : {q, q_now, qdot})
        if (a)
            for (size_t i = 0; i < 12 * (size_t)ctx->nb; ++i)
                if (!std::isfinite(a[i])) {
                    ctx->err = "abd_set_state: non-finite entry in the state";
                    return ABD_ERR_NAN;
                }
    if (q) HIPCHK(hipMemcpyAsync(ctx->q, q, n, hipMemcpyHostToDevice, ctx->st));
    if (q_now) HIPCHK(hipMemcpyAsync(ctx->qnow, q_now, n, hipMemcpyHostToDevice, ctx->st));
    if (qdot) HIPCHK(hipMemcpyAsync(ctx->qdot, qdot, n, hipMemcpyHostToDevice, ctx->st));
    launch_qtilde(ctx->st, ctx->nb, ctx->qnow, ctx->qdot, ctx->P.dt, ctx->P.gravity, ctx->qtilde);
    CHK(dbg(ctx, "launch_qtilde"));
    CHK(fill_i32(ctx, ctx->C.newton_active, ctx->ns, 1));
    CHK(fill_i32(ctx, ctx->cnt + C_ERR, 1, 0));  // a new state: forget an error reported for the old one
    HIPCHK(hipStreamSynchronize(ctx->st));
    ctx->cand_valid = false;
    ctx->assembled = ctx->solved = false;
    return ABD_OK;
}

int abd_get_state(const AbdCtx* cctx, double* q, double* q_now, double* qdot) {
    AbdCtx* ctx = const_cast<AbdCtx*>(cctx);
    CHK(check_ctx(ctx));
    size_t n = 12 * (size_t)ctx->nb * sizeof(double);
    if (q) HIPCHK(hipMemcpyAsync(q, ctx->q, n, hipMemcpyDeviceToHost, ctx->st));
    if (q_now) HIPCHK(hipMemcpyAsync(q_now, ctx->qnow, n, hipMemcpyDeviceToHost, ctx->st));
    if (qdot) HIPCHK(hipMemcpyAsync(qdot, ctx->qdot, n, hipMemcpyDeviceToHost, ctx->st));
    HIPCHK(hipStreamSynchronize(ctx->st));
    return ABD_OK;
}

int abd_step(AbdCtx* ctx, int nsteps, AbdStepStats* stats) {
    CHK(check_ctx(ctx));
    if (nsteps < 0) return ABD_ERR_ARG;
    for (int k = 0; k < nsteps; ++k) {
        CHK(step_once(ctx));
        timer_collect(ctx);
    }
    CHK(sync_counters(ctx));
    if (stats)
        for (int s = 0; s < ctx->ns && s < (int)ctx->stats.size(); ++s) stats[s] = ctx->stats[s];
    for (auto& s : ctx->stats)
        if (!std::isfinite(s.residual_last) || !std::isfinite(s.energy)) {
            ctx->err = "non-finite residual or energy";
            return ABD_ERR_NAN;
        }
    return ABD_OK;
}

int abd_get_stiffness(const AbdCtx* cctx, double* k) {
    AbdCtx* ctx = const_cast<AbdCtx*>(cctx);
    CHK(check_ctx(ctx));
    if (!k) return ABD_ERR_ARG;
    HIPCHK(hipMemcpyAsync(k, ctx->C.kb, ctx->nb * sizeof(double), hipMemcpyDeviceToHost, ctx->st));
    HIPCHK(hipStreamSynchronize(ctx->st));
    return ABD_OK;
}

int abd_set_stiffness(AbdCtx* ctx, const double* k) {
    CHK(check_ctx(ctx));
    if (!k) return ABD_ERR_ARG;
    for (int b = 0; b < ctx->nb; ++b)
        if (!(k[b] >= 0.0) || !std::isfinite(k[b])) {
            ctx->err = "contact stiffness must be finite and >= 0";
            return ABD_ERR_ARG;
        }
    HIPCHK(hipMemcpyAsync(ctx->C.kb, k, ctx->nb * sizeof(double), hipMemcpyHostToDevice, ctx->st));
    launch_scene_kmax(ctx->st, ctx->G, ctx->C);
    HIPCHK(hipStreamSynchronize(ctx->st));
    ctx->assembled = ctx->solved = false;
    return ABD_OK;
}

int abd_get_iter_log(const AbdCtx* ctx, double* residual, double* alpha, int* cg) {
    if (!ctx) return ABD_ERR_ARG;
    if (residual) std::copy(ctx->log_res.begin(), ctx->log_res.end(), residual);
    if (alpha) std::copy(ctx->log_alpha.begin(), ctx->log_alpha.end(), alpha);
    if (cg) std::copy(ctx->log_cg.begin(), ctx->log_cg.end(), cg);
    return ABD_OK;
}

int abd_detect(AbdCtx* ctx, int* pairs, double* d2, int cap, int* n) {
    CHK(check_ctx(ctx));
    CHK(static_broad_phase(ctx));
    CHK(read_counts(ctx));
    const int nc = ctx->n_cand;
    std::vector<uint64_t> keys(nc);
    std::vector<int> flag(nc);
    std::vector<double> dd(nc);
    if (nc > 0) {
        launch_cand_eval(ctx->st, ctx->G, host_cnt(nc), ctx->cand_s, ctx->X0, ctx->P, ctx->flag, nullptr);
        CHK(dbg(ctx, "launch_cand_eval"));
        launch_pair_d2(ctx->st, ctx->G, nc, ctx->cand_s, ctx->X0, ctx->P, ctx->e_cand);
        CHK(dbg(ctx, "launch_pair_d2"));
        HIPCHK(hipGetLastError());
        HIPCHK(hipMemcpyAsync(keys.data(), ctx->cand_s, nc * sizeof(uint64_t), hipMemcpyDeviceToHost, ctx->st));
        HIPCHK(hipMemcpyAsync(flag.data(), ctx->flag, nc * sizeof(int), hipMemcpyDeviceToHost, ctx->st));
        HIPCHK(hipMemcpyAsync(dd.data(), ctx->e_cand, nc * sizeof(double), hipMemcpyDeviceToHost, ctx->st));
    }
    HIPCHK(hipStreamSynchronize(ctx->st));
    int na = 0;
    for (int i = 0; i < nc; ++i) na += flag[i];
    *n = na;
    if (na > cap) {
        ctx->err = "abd_detect: capacity too small";
        return ABD_ERR_CAPACITY;
    }
    int k = 0;
    for (int c = 0; c < nc; ++c) {
        if (!flag[c]) continue;
        uint64_t key = keys[c];
        int s = key_scene(key), kind = key_kind(key), il = key_il(key), jl = key_jl(key);
        int i = (kind == ABD_EE ? ctx->h_scene_e0[s] : ctx->h_scene_v0[s]) + il;
        int j = (kind == ABD_PT ? ctx->h_scene_t0[s] + jl : (kind == ABD_EE ? ctx->h_scene_e0[s] + jl : 0));
        pairs[3 * k] = kind;
        pairs[3 * k + 1] = i;
        pairs[3 * k + 2] = j;
        if (d2) d2[k] = dd[c];
        ++k;
    }
    return ABD_OK;
}

int abd_assemble(AbdCtx* ctx, double* grad, double* diag, int* pair_ij, double* offdiag, int cap, int* n_pairs,
                 double* energy) {
    CHK(check_ctx(ctx));
    CHK(fill_i32(ctx, ctx->C.newton_active, ctx->ns, 1));
    CHK(static_broad_phase(ctx));
    CHK(assemble_checked(ctx));
    CHK(read_counts(ctx));
    const int nb = ctx->nb;
    if (grad) HIPCHK(hipMemcpyAsync(grad, ctx->grad, 12 * nb * sizeof(double), hipMemcpyDeviceToHost, ctx->st));
    if (diag) HIPCHK(hipMemcpyAsync(diag, ctx->diag, 144 * (size_t)nb * sizeof(double), hipMemcpyDeviceToHost, ctx->st));
    if (n_pairs) *n_pairs = ctx->n_pairs;
    if (ctx->n_pairs > cap && (pair_ij || offdiag)) {
        ctx->err = "abd_assemble: capacity too small";
        HIPCHK(hipStreamSynchronize(ctx->st));
        return ABD_ERR_CAPACITY;
    }
    if (pair_ij && ctx->n_pairs)
        HIPCHK(hipMemcpyAsync(pair_ij, ctx->pair_b, 2 * ctx->n_pairs * sizeof(int), hipMemcpyDeviceToHost, ctx->st));
    if (offdiag && ctx->n_pairs)
        HIPCHK(hipMemcpyAsync(offdiag, ctx->off, 144 * (size_t)ctx->n_pairs * sizeof(double), hipMemcpyDeviceToHost, ctx->st));
    if (energy) {
        launch_scene_sum(ctx->st, ctx->G, ctx->e_body, ctx->e_cand, ctx->cand_lo, ctx->cand_hi, ctx->C.E0);
        CHK(dbg(ctx, "launch_scene_sum"));
        HIPCHK(hipMemcpyAsync(energy, ctx->C.E0, ctx->ns * sizeof(double), hipMemcpyDeviceToHost, ctx->st));
    }
    HIPCHK(hipStreamSynchronize(ctx->st));
    timer_collect(ctx);
    return ABD_OK;
}

int abd_solve(AbdCtx* ctx, double* du, double* residual, int* cg) {
    CHK(check_ctx(ctx));
    if (!ctx->assembled) {
        ctx->err = "abd_solve before abd_assemble";
        return ABD_ERR_STATE;
    }
    CHK(solve(ctx));
    if (du) HIPCHK(hipMemcpyAsync(du, ctx->du, 12 * ctx->nb * sizeof(double), hipMemcpyDeviceToHost, ctx->st));
    if (residual) HIPCHK(hipMemcpyAsync(residual, ctx->bnorm, ctx->ns * sizeof(double), hipMemcpyDeviceToHost, ctx->st));
    if (cg) HIPCHK(hipMemcpyAsync(cg, ctx->cg_iters, ctx->ns * sizeof(int), hipMemcpyDeviceToHost, ctx->st));
    HIPCHK(hipStreamSynchronize(ctx->st));
    timer_collect(ctx);
    return ABD_OK;
}

int abd_energy(AbdCtx* ctx, const double* q, double* energy) {
    CHK(check_ctx(ctx));
    if (q) {
        // evaluate at q without disturbing the iterate: swap buffers
        HIPCHK(hipMemcpyAsync(ctx->qtmp, ctx->q, 12 * ctx->nb * sizeof(double), hipMemcpyDeviceToDevice, ctx->st));
        HIPCHK(hipMemcpyAsync(ctx->q, q, 12 * ctx->nb * sizeof(double), hipMemcpyHostToDevice, ctx->st));
    }
    int rc = static_broad_phase(ctx);
    if (rc == ABD_OK) {
        // a failure below is recorded, and q is still restored before returning
        launch_cand_eval(ctx->st, ctx->G, cand_cnt(ctx), ctx->cand_s, ctx->X0, ctx->P, nullptr, ctx->e_cand);
        rc = dbg(ctx, "launch_cand_eval");
    }
    if (rc == ABD_OK) {
        launch_body_energy(ctx->st, ctx->G, ctx->P, ctx->q, nullptr, nullptr, ctx->qtilde, ctx->e_body);
        rc = dbg(ctx, "launch_body_energy");
    }
    if (rc == ABD_OK) {
        launch_scene_sum(ctx->st, ctx->G, ctx->e_body, ctx->e_cand, ctx->cand_lo, ctx->cand_hi, ctx->C.E0);
        rc = dbg(ctx, "launch_scene_sum");
    }
    if (rc == ABD_OK) HIPCHK(hipMemcpyAsync(energy, ctx->C.E0, ctx->ns * sizeof(double), hipMemcpyDeviceToHost, ctx->st));
    if (q) {
        HIPCHK(hipMemcpyAsync(ctx->q, ctx->qtmp, 12 * ctx->nb * sizeof(double), hipMemcpyDeviceToDevice, ctx->st));
        ctx->cand_valid = false;
    }
    HIPCHK(hipStreamSynchronize(ctx->st));
    ctx->assembled = ctx->solved = false;
    return rc;
}

int abd_ccd(AbdCtx* ctx, const double* du, double* alpha) {
    CHK(check_ctx(ctx));
    if (du) HIPCHK(hipMemcpyAsync(ctx->du, du, 12 * ctx->nb * sizeof(double), hipMemcpyHostToDevice, ctx->st));
    CHK(ccd_bound(ctx));
    ctx->cand_valid = false;  // swept list, not re-validated by a line search
    std::vector<unsigned long long> bits(ctx->ns);
    HIPCHK(hipMemcpyAsync(bits.data(), ctx->C.amax_bits, ctx->ns * sizeof(unsigned long long), hipMemcpyDeviceToHost, ctx->st));
    HIPCHK(hipStreamSynchronize(ctx->st));
    for (int s = 0; s < ctx->ns; ++s) {
        double a;
        std::memcpy(&a, &bits[s], sizeof(double));
        alpha[s] = ctx->P.ccd_enabled ? std::min(1.0, a) : 1.0;
    }
    return ABD_OK;
}

int abd_line_search(AbdCtx* ctx, double* alpha) {
    CHK(check_ctx(ctx));
    if (!ctx->solved) {
        ctx->err = "abd_line_search before abd_solve";
        return ABD_ERR_STATE;
    }
    CHK(ccd_bound(ctx));
    CHK(line_search(ctx));
    if (alpha) HIPCHK(hipMemcpyAsync(alpha, ctx->C.alpha, ctx->ns * sizeof(double), hipMemcpyDeviceToHost, ctx->st));
    HIPCHK(hipStreamSynchronize(ctx->st));
    ctx->assembled = ctx->solved = false;
    return ABD_OK;
}

int abd_line_search_along(AbdCtx* ctx, const double* du, double* alpha) {
    CHK(check_ctx(ctx));
    if (!du) return ABD_ERR_ARG;
    HIPCHK(hipMemcpyAsync(ctx->du, du, 12 * ctx->nb * sizeof(double), hipMemcpyHostToDevice, ctx->st));
    CHK(fill_i32(ctx, ctx->C.newton_active, ctx->ns, 1));
    CHK(ccd_bound(ctx));
    CHK(line_search(ctx));
    if (alpha) HIPCHK(hipMemcpyAsync(alpha, ctx->C.alpha, ctx->ns * sizeof(double), hipMemcpyDeviceToHost, ctx->st));
    HIPCHK(hipStreamSynchronize(ctx->st));
    ctx->assembled = ctx->solved = false;
    return ABD_OK;
}

int abd_forward(AbdCtx* ctx) {
    CHK(check_ctx(ctx));
    launch_forward(ctx->st, ctx->nb, ctx->P.dt, ctx->q, ctx->qnow, ctx->qdot);
    CHK(dbg(ctx, "launch_forward"));
    launch_qtilde(ctx->st, ctx->nb, ctx->qnow, ctx->qdot, ctx->P.dt, ctx->P.gravity, ctx->qtilde);
    CHK(dbg(ctx, "launch_qtilde"));
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(ctx->st));
    return ABD_OK;
}

int abd_get_vertices(AbdCtx* ctx, double* xyz) {
    CHK(check_ctx(ctx));
    launch_xform(ctx->st, ctx->G, ctx->q, nullptr, nullptr, ctx->Xt);
    CHK(dbg(ctx, "launch_xform"));
    std::vector<double4> X(ctx->nv);
    HIPCHK(hipMemcpyAsync(X.data(), ctx->Xt, ctx->nv * sizeof(double4), hipMemcpyDeviceToHost, ctx->st));
    HIPCHK(hipStreamSynchronize(ctx->st));
    for (int v = 0; v < ctx->nv; ++v) {
        xyz[3 * v] = X[v].x; xyz[3 * v + 1] = X[v].y; xyz[3 * v + 2] = X[v].z;
    }
    return ABD_OK;
}

int abd_get_contact_points(AbdCtx* ctx, double* midpoints, double* dist, int cap, int* n) {
    CHK(check_ctx(ctx));
    CHK(static_broad_phase(ctx));
    CHK(read_counts(ctx));
    const int nc = ctx->n_cand;
    launch_cand_eval(ctx->st, ctx->G, host_cnt(nc), ctx->cand_s, ctx->X0, ctx->P, ctx->flag, nullptr);
    CHK(dbg(ctx, "launch_cand_eval"));
    CHK(excl_scan(ctx, ctx->flag, ctx->pos, host_cnt(nc)));
    hipLaunchKernelGGL(k_scan_total, dim3(1), dim3(64), 0, ctx->st, host_cnt(nc), ctx->flag, ctx->pos,
                       ctx->cnt + C_NACT);
    CHK(sync_counters(ctx));
    int na = ctx->h_cnt[C_NACT];
    *n = na;
    if (na > cap) {
        ctx->err = "abd_get_contact_points: capacity too small";
        return ABD_ERR_CAPACITY;
    }
    if (na == 0) return ABD_OK;
    CHK(grow_active(ctx, na));
    launch_compact(ctx->st, host_cnt(nc), ctx->flag, ctx->pos, ctx->cand_s, ctx->act);
    CHK(dbg(ctx, "launch_compact"));
    double *dm = nullptr, *dd = nullptr;
    CHK(dalloc(ctx, &dm, 3 * na));
    CHK(dalloc(ctx, &dd, na));
    launch_contact_points(ctx->st, ctx->G, na, ctx->act, ctx->X0, ctx->P, dm, dd);
    CHK(dbg(ctx, "launch_contact_points"));
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(midpoints, dm, 3 * na * sizeof(double), hipMemcpyDeviceToHost, ctx->st));
    HIPCHK(hipMemcpyAsync(dist, dd, na * sizeof(double), hipMemcpyDeviceToHost, ctx->st));
    HIPCHK(hipStreamSynchronize(ctx->st));
    CHK(dfree(ctx, dm));
    CHK(dfree(ctx, dd));
    return ABD_OK;
}

int abd_autotune(AbdCtx* ctx, int repeats, int* bodies_per_tile, int* per_iteration_launch, double* ms) {
    CHK(check_ctx(ctx));
    if (repeats < 1) repeats = 1;
    CHK(fill_i32(ctx, ctx->C.newton_active, ctx->ns, 1));
    CHK(static_broad_phase(ctx));
    CHK(assemble_checked(ctx));
    double best = -1.0, t_cur = -1.0;
    const int cur_bpt = ctx->bodies_per_tile, cur_multi = ctx->pcg_multi ? 1 : 0;
    int best_bpt = cur_bpt, best_multi = cur_multi;
    for (int bpt = 1; bpt <= 5; ++bpt) {
        CHK(set_tiles(ctx, bpt));
        for (int multi = 0; multi < 2; ++multi) {
            const int slot = 2 * (bpt - 1) + multi;
            if (ms) ms[slot] = -1.0;
            if (!multi && !persist_ok(ctx)) continue;
            ctx->pcg_multi = multi != 0;
            CHK(solve(ctx));  // warm-up (and the cooperative-launch probe)
            HIPCHK(hipStreamSynchronize(ctx->st));
            if (!multi && !persist_ok(ctx)) continue;
            auto t0 = std::chrono::steady_clock::now();
            for (int r = 0; r < repeats; ++r) CHK(solve(ctx));
            HIPCHK(hipStreamSynchronize(ctx->st));
            const double t = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count() / repeats;
            if (ms) ms[slot] = t;
            if (bpt == cur_bpt && multi == cur_multi) t_cur = t;
            if (best < 0.0 || t < best) {
                best = t;
                best_bpt = bpt;
                best_multi = multi;
            }
        }
    }
    // keep the layout in effect unless another is clearly faster (timing noise must not flip it:
    // a different layout sums the PCG partials in a different order)
    if (t_cur > 0.0 && t_cur <= 1.03 * best) {
        best_bpt = cur_bpt;
        best_multi = cur_multi;
    }
    CHK(set_tiles(ctx, best_bpt));
    ctx->pcg_multi = best_multi != 0;
    if (bodies_per_tile) *bodies_per_tile = best_bpt;
    if (per_iteration_launch) *per_iteration_launch = best_multi;
    ctx->assembled = ctx->solved = false;
    return ABD_OK;
}

int abd_set_pcg_layout(AbdCtx* ctx, int bodies_per_tile, int per_iteration_launch) {
    CHK(check_ctx(ctx));
    CHK(set_tiles(ctx, bodies_per_tile));
    ctx->pcg_multi = per_iteration_launch != 0;
    ctx->solved = false;
    return ABD_OK;
}

int abd_set_warm_start(AbdCtx* ctx, AbdMlp* model, const float* feat_mean, const float* feat_std,
                       const float* out_mean, const float* out_std) {
    CHK(check_ctx(ctx));
    if (!model) {
        ctx->mlp = nullptr;
        return ABD_OK;
    }
    if (!feat_mean || !feat_std || !out_mean || !out_std || model->L < 1 || model->d[0] != ABD_N_FEATURES ||
        model->d[model->L] != 12) {
        ctx->err = "warm-start model must map " + std::to_string(ABD_N_FEATURES) +
                   " body features to 12 DOFs, with normalisation arrays";
        return ABD_ERR_ARG;
    }
    std::vector<float> istd(ABD_N_FEATURES);
    for (int j = 0; j < ABD_N_FEATURES; ++j) istd[j] = feat_std[j] > 0.f ? 1.f / feat_std[j] : 0.f;
    HIPCHK(hipMemcpy(ctx->fmean, feat_mean, ABD_N_FEATURES * sizeof(float), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(ctx->fistd, istd.data(), ABD_N_FEATURES * sizeof(float), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(ctx->omean, out_mean, 12 * sizeof(float), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(ctx->ostd, out_std, 12 * sizeof(float), hipMemcpyHostToDevice));
    HIPCHK(hipStreamSynchronize(model->st));
    ctx->mlp = model;
    return ABD_OK;
}

int abd_set_preconditioner(AbdCtx* ctx, AbdMlp* model, const float* feat_mean, const float* feat_std,
                           const float* out_mean, const float* out_std) {
    CHK(check_ctx(ctx));
    if (!model) {
        ctx->pre_mlp = nullptr;
        return ABD_OK;
    }
    if (!feat_mean || !feat_std || !out_mean || !out_std || model->L < 1 || model->d[0] != ABD_N_FEATURES ||
        model->d[model->L] != ABD_N_PRECOND) {
        ctx->err = "preconditioner model must map " + std::to_string(ABD_N_FEATURES) + " body features to " +
                   std::to_string(ABD_N_PRECOND) + " factor entries, with normalisation arrays";
        return ABD_ERR_ARG;
    }
    std::vector<float> istd(ABD_N_FEATURES);
    for (int j = 0; j < ABD_N_FEATURES; ++j) istd[j] = feat_std[j] > 0.f ? 1.f / feat_std[j] : 0.f;
    HIPCHK(hipMemcpy(ctx->pfmean, feat_mean, ABD_N_FEATURES * sizeof(float), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(ctx->pfistd, istd.data(), ABD_N_FEATURES * sizeof(float), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(ctx->pomean, out_mean, ABD_N_PRECOND * sizeof(float), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(ctx->postd, out_std, ABD_N_PRECOND * sizeof(float), hipMemcpyHostToDevice));
    HIPCHK(hipStreamSynchronize(model->st));
    ctx->pre_mlp = model;
    ctx->solved = false;
    return ABD_OK;
}

int abd_set_predictor(AbdCtx* ctx, AbdMlp* model, const float* feat_mean, const float* feat_std,
                      const float* out_mean, const float* out_std) {
    CHK(check_ctx(ctx));
    if (!model) {
        ctx->pred_mlp = nullptr;
        return ABD_OK;
    }
    if (!feat_mean || !feat_std || !out_mean || !out_std || model->L < 1 || model->d[0] != ABD_N_FEATURES ||
        model->d[model->L] != 12) {
        ctx->err = "step-predictor model must map " + std::to_string(ABD_N_FEATURES) +
                   " body features to 12 DOFs, with normalisation arrays";
        return ABD_ERR_ARG;
    }
    std::vector<float> istd(ABD_N_FEATURES);
    for (int j = 0; j < ABD_N_FEATURES; ++j) istd[j] = feat_std[j] > 0.f ? 1.f / feat_std[j] : 0.f;
    HIPCHK(hipMemcpy(ctx->prfmean, feat_mean, ABD_N_FEATURES * sizeof(float), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(ctx->prfistd, istd.data(), ABD_N_FEATURES * sizeof(float), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(ctx->promean, out_mean, 12 * sizeof(float), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(ctx->prostd, out_std, 12 * sizeof(float), hipMemcpyHostToDevice));
    HIPCHK(hipStreamSynchronize(model->st));
    ctx->pred_mlp = model;
    return ABD_OK;
}

int abd_log_steps(AbdCtx* ctx, int on) {
    CHK(check_ctx(ctx));
    ctx->log_steps = on != 0;
    ctx->step_feat.clear();
    ctx->step_dq.clear();
    return ABD_OK;
}

int abd_get_step_samples(const AbdCtx* ctx, float* feat, float* dq, int cap, int* n) {
    if (!ctx || !n) return ABD_ERR_ARG;
    const int total = (int)(ctx->step_dq.size() / 12);
    *n = total;
    const int k = std::min(total, cap);
    if (feat && k > 0) std::memcpy(feat, ctx->step_feat.data(), sizeof(float) * ABD_N_FEATURES * (size_t)k);
    if (dq && k > 0) std::memcpy(dq, ctx->step_dq.data(), sizeof(float) * 12 * (size_t)k);
    return ABD_OK;
}

int abd_log_samples(AbdCtx* ctx, int on) {
    CHK(check_ctx(ctx));
    ctx->log_on = on != 0;
    ctx->log_feat.clear();
    ctx->log_du.clear();
    return ABD_OK;
}

int abd_get_samples(const AbdCtx* ctx, float* feat, float* du, int cap, int* n) {
    if (!ctx || !n) return ABD_ERR_ARG;
    const int total = (int)(ctx->log_du.size() / 12);
    *n = total;
    const int k = std::min(total, cap);
    if (feat && k > 0) std::memcpy(feat, ctx->log_feat.data(), sizeof(float) * ABD_N_FEATURES * (size_t)k);
    if (du && k > 0) std::memcpy(du, ctx->log_du.data(), sizeof(float) * 12 * (size_t)k);
    return ABD_OK;
}

int abd_enable_timing(AbdCtx* ctx, int on) {
    CHK(check_ctx(ctx));
    if (on < 0 || on > 2) {
        ctx->err = "timing mode must be 0, 1 or 2";
        return ABD_ERR_ARG;
    }
    if (ctx->st) HIPCHK(hipStreamSynchronize(ctx->st));
    timer_collect(ctx);  // records still open belong to the previous mode
    ctx->T.on = on != 0;
    ctx->T.pcg_only = on == 2;
    if (on == 2) {
        if (!ctx->T.pcg_ts) {
            CHK(dalloc(ctx, &ctx->T.pcg_ts, 2));
            CHK(dalloc(ctx, &ctx->T.pcg_acc, 4));
            int dev = 0, khz = 0;
            HIPCHK(hipGetDevice(&dev));
            HIPCHK(hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, dev));
            ctx->T.clock_khz = khz > 0 ? (double)khz : 1.0;
        }
        hipLaunchKernelGGL(k_fill_u64, dim3(1), dim3(64), 0, ctx->st, 1, ctx->T.pcg_ts, ~0ull);
        hipLaunchKernelGGL(k_fill_u64, dim3(1), dim3(64), 0, ctx->st, 1, ctx->T.pcg_ts + 1, 0ull);
        hipLaunchKernelGGL(k_fill_u64, dim3(1), dim3(64), 0, ctx->st, 4, (unsigned long long*)ctx->T.pcg_acc, 0ull);
        HIPCHK(hipStreamSynchronize(ctx->st));
    }
    ctx->T.ms.clear(); ctx->T.calls.clear(); ctx->T.bytes.clear(); ctx->T.order.clear();
    return ABD_OK;
}

int abd_get_kernel_times(const AbdCtx* cctx, char* names, int name_cap, double* ms, int* calls, double* bytes, int cap,
                         int* n) {
    if (!cctx) return ABD_ERR_ARG;
    AbdCtx* ctx = const_cast<AbdCtx*>(cctx);
    if (ctx->T.on && ctx->T.pcg_only && ctx->T.pcg_acc) {
        // the persistent solves' device timestamps since timing was enabled
        double acc[4] = {0, 0, 0, 0};
        HIPCHK(hipStreamSynchronize(ctx->st));
        HIPCHK(hipMemcpy(acc, ctx->T.pcg_acc, sizeof acc, hipMemcpyDeviceToHost));
        if (acc[2] > 0.0) {
            if (!ctx->T.calls.count("pcg_solve")) ctx->T.order.push_back("pcg_solve");
            ctx->T.ms["pcg_solve"] = acc[0] / ctx->T.clock_khz;
            ctx->T.calls["pcg_solve"] = (int)acc[2];
            ctx->T.bytes["pcg_solve"] = acc[1];
        }
    }
    std::string all;
    int k = 0;
    for (auto& name : ctx->T.order) {
        if (k < cap) {
            ms[k] = ctx->T.ms.at(name);
            calls[k] = ctx->T.calls.at(name);
            bytes[k] = ctx->T.bytes.at(name);
        }
        all += name;
        all += "\n";
        ++k;
    }
    *n = k;
    if (names && name_cap > 0) {
        std::strncpy(names, all.c_str(), name_cap - 1);
        names[name_cap - 1] = 0;
    }
    return k > cap ? ABD_ERR_CAPACITY : ABD_OK;
}

int abd_psd_project_test(int m, int precision, int absolute, int n, const double* A, double* out) {
    if (!A || !out || n < 1 || (m != 3 && m != 6 && m != 9 && m != 12) || (precision != 0 && precision != 1))
        return ABD_ERR_ARG;
    const size_t bytes = (size_t)n * m * m * sizeof(double);
    double *dA = nullptr, *dO = nullptr;
    hipError_t e = hipMalloc(&dA, bytes);
    if (e == hipSuccess) e = hipMalloc(&dO, bytes);
    if (e == hipSuccess) e = hipMemcpy(dA, A, bytes, hipMemcpyHostToDevice);
    if (e == hipSuccess) {
        launch_psd_test(nullptr, m, precision, absolute, n, dA, dO);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipStreamSynchronize(nullptr);
    if (e == hipSuccess) e = hipMemcpy(out, dO, bytes, hipMemcpyDeviceToHost);
    (void)hipFree(dA);
    (void)hipFree(dO);
    return e == hipSuccess ? ABD_OK : ABD_ERR_HIP;
}

}  // extern "C"
