// abd_broad.hip — broad phase of the ABD Newton step, built for affine bodies.
//
// The reference refits one world-space BVH over all edges at every detection
// (contact.py:244-250) and queries it per edge (contact.py:143-169).  Affine bodies never change
// their rest-frame meshes, so here every distinct mesh carries a rest-frame BVH built once on the
// host (abd_api.hip), and a detection over the linear path X0 -> X1 is
//   1. world boxes of the bodies over X0 and X1, inflated by dhat/2 (k_body_boxes);
//   2. the body pairs of each scene whose boxes overlap and which are not both pinned, in (A, B)
//      order (k_body_pair_count, k_body_pair_write);
//   3. world-space boxes of every body's BVH nodes: the rest-frame node box mapped through the
//      affine state x = c + A xr at both ends of the path (positions are linear in the path
//      parameter, so the swept image of the box lies in the union of its two world AABBs)
//      (k_world_nodes);
//   4. per body pair, every vertex of either body against the other body's triangle BVH, every
//      edge of the lower body against the upper body's edge BVH, by a stackless preorder
//      traversal; leaf pairs are queued in LDS and checked 64 at a time (k_pair_queries).
// The leaf check applies the exact candidate rule in FP64, the same as the oracle's broad phase:
//   * the swept primitive boxes overlap (point box inflated by dhat, edge boxes by dhat/2 each,
//     triangle boxes not inflated), and
//   * the distance at X0 minus the additive-CCD motion bound is below dhat (may_reach).
// Every node and body-level test is conservative (float boxes rounded outward by at least one
// float ulp beyond FP64 rounding), so the candidate list is exactly the set of pairs meeting the
// leaf rule.  With X1 = X0 it is the active set.
#include <hip/hip_runtime.h>
#include <climits>
#include <float.h>
#include <stdint.h>

#include "abd_device.hpp"
#include "abd_internal.hpp"

namespace abd {

static inline int cdivb(long a, int b) { return (int)((a + b - 1) / b); }

// float bounds strictly outside the FP64 value
__device__ __forceinline__ float flo(double v) { return nextafterf(__double2float_rd(v), -FLT_MAX); }
__device__ __forceinline__ float fhi(double v) { return nextafterf(__double2float_ru(v), FLT_MAX); }

__device__ __forceinline__ bool fbox_ov(const float* a, const float* b) {
    return !(a[0] > b[3] || b[0] > a[3] || a[1] > b[4] || b[1] > a[4] || a[2] > b[5] || b[2] > a[5]);
}

// ---------------------------------------------------------------------------------------------
// 1. body boxes (one wave per body)
// ---------------------------------------------------------------------------------------------
__global__ __launch_bounds__(64) void k_body_boxes(Geo G, BodyMesh M, const double4* __restrict__ X0,
                                                    const double4* __restrict__ X1, double half_infl,
                                                    float* __restrict__ box) {
    const int b = blockIdx.x, lane = threadIdx.x;
    double lo[3] = {DBL_MAX, DBL_MAX, DBL_MAX}, hi[3] = {-DBL_MAX, -DBL_MAX, -DBL_MAX};
    for (int v = M.body_v0[b] + lane; v < M.body_v0[b + 1]; v += 64) {
        double4 a = X0[v], c = X1[v];
        double pa[3] = {a.x, a.y, a.z}, pc[3] = {c.x, c.y, c.z};
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            lo[k] = fmin(lo[k], fmin(pa[k], pc[k]));
            hi[k] = fmax(hi[k], fmax(pa[k], pc[k]));
        }
    }
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        for (int o = 32; o > 0; o >>= 1) {
            lo[k] = fmin(lo[k], __shfl_xor(lo[k], o, 64));
            hi[k] = fmax(hi[k], __shfl_xor(hi[k], o, 64));
        }
    }
    if (lane == 0) {
        const bool empty = lo[0] > hi[0];
        for (int k = 0; k < 3; ++k) {
            box[6 * b + k] = empty ? FLT_MAX : flo(lo[k] - half_infl);
            box[6 * b + 3 + k] = empty ? -FLT_MAX : fhi(hi[k] + half_infl);
        }
    }
}

void launch_body_boxes(hipStream_t s, const Geo& G, const BodyMesh& M, const double4* X0, const double4* X1,
                       double half_infl, float* box) {
    hipLaunchKernelGGL(k_body_boxes, dim3(G.nb), dim3(64), 0, s, G, M, X0, X1, half_infl, box);
}

// ---------------------------------------------------------------------------------------------
// 2. body pairs (A, B > A) of a scene whose inflated boxes overlap and which are not both pinned:
//    one wave per body A, lanes over the later bodies of its scene.  The count pass sizes each
//    body's run, an exclusive scan places the runs, the write pass fills them in order, so the
//    list is sorted by (A, B) whatever the scheduling.
// ---------------------------------------------------------------------------------------------
__device__ __forceinline__ bool body_pair_ok(const Geo& G, const float* box, int A, int B) {
    return !(G.pinned[A] && G.pinned[B]) && fbox_ov(box + 6 * A, box + 6 * B);
}

__global__ __launch_bounds__(64) void k_body_pair_count(Geo G, const float* __restrict__ box, int* __restrict__ cnt) {
    const int A = blockIdx.x, lane = threadIdx.x;
    const int end = G.scene_b0[G.bscene[A] + 1];
    int c = 0;
    for (int B = A + 1 + lane; B < end; B += 64) c += body_pair_ok(G, box, A, B) ? 1 : 0;
    for (int o = 32; o > 0; o >>= 1) c += __shfl_xor(c, o, 64);
    if (lane == 0) cnt[A] = c;
}

__global__ __launch_bounds__(64) void k_body_pair_write(Geo G, const float* __restrict__ box, const int* __restrict__ off,
                                                         int2* __restrict__ out, int cap) {
    const int A = blockIdx.x, lane = threadIdx.x;
    const int end = G.scene_b0[G.bscene[A] + 1];
    int pos = off[A];
    for (int B0 = A + 1; B0 < end; B0 += 64) {
        const int B = B0 + lane;
        const bool ok = B < end && body_pair_ok(G, box, A, B);
#if defined(ABD_EMULATION)
        // fibre emulation: lanes do not run in lockstep, so rank by a serial pass over the chunk
        int rank = 0;
        for (int k = B0; k < B && k < end; ++k) rank += body_pair_ok(G, box, A, k) ? 1 : 0;
        int tot = 0;
        for (int k = B0; k < end && k < B0 + 64; ++k) tot += body_pair_ok(G, box, A, k) ? 1 : 0;
#else
        const unsigned long long m = __ballot(ok);
        const int rank = __popcll(m & ((1ull << lane) - 1ull));
        const int tot = __popcll(m);
#endif
        if (ok && pos + rank < cap) out[pos + rank] = make_int2(A, B);
        pos += tot;
    }
}

void launch_body_pair_count(hipStream_t s, const Geo& G, const float* box, int* cnt) {
    hipLaunchKernelGGL(k_body_pair_count, dim3(G.nb), dim3(64), 0, s, G, box, cnt);
}

void launch_body_pair_write(hipStream_t s, const Geo& G, const float* box, const int* off, int2* out, int cap) {
    hipLaunchKernelGGL(k_body_pair_write, dim3(G.nb), dim3(64), 0, s, G, box, off, out, cap);
}

// ---------------------------------------------------------------------------------------------
// 3. world-space node boxes and 4. per body pair traversal
// ---------------------------------------------------------------------------------------------
// affine state of a body: x = c + A xr with A_ik = a_k[i] (q12 = [c, a0, a1, a2])
struct Aff {
    double c[3], A[3][3];
};

__device__ __forceinline__ Aff load_aff(const double* q, const double* du, int b) {
    Aff f;
    const double* qb = q + 12 * b;
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        f.c[i] = du ? qb[i] - du[12 * b + i] : qb[i];
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            const int m = 3 + 3 * k + i;
            f.A[i][k] = du ? qb[m] - du[12 * b + m] : qb[m];
        }
    }
    return f;
}

// World AABB of a rest-frame box under an affine state x = c + A xr, in FP32, widened and rounded
// outward so that it contains the exact image.  Its rounding error is below
// 11 u * sc (u = 2^-24, sc = |c| + sum_k |A_ik| max(|lo_k|, |hi_k|)): the FP32 copies of c and A
// (u each), the box centre and half-width (u), and the two 3-term FMA chains (4u and 3u); the
// slack 4e-6 * sc is six times that, and the final subtraction is covered by the outward ulp step.
struct Aff32 {
    float c[3], A[3][3];
};
__device__ __forceinline__ Aff32 aff32(const Aff& f) {
    Aff32 g;
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        g.c[i] = (float)f.c[i];
#pragma unroll
        for (int k = 0; k < 3; ++k) g.A[i][k] = (float)f.A[i][k];
    }
    return g;
}
__device__ __forceinline__ void map_box32(const Aff32& f, const float* rlo, const float* rhi, float* out) {
    float ctr[3], half[3], mag[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        ctr[k] = 0.5f * (rlo[k] + rhi[k]);
        half[k] = 0.5f * (rhi[k] - rlo[k]);
        mag[k] = fmaxf(fabsf(rlo[k]), fabsf(rhi[k]));
    }
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        float cw = f.c[i], hw = 0.0f, sc = fabsf(f.c[i]);
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            cw = fmaf(f.A[i][k], ctr[k], cw);
            hw = fmaf(fabsf(f.A[i][k]), half[k], hw);
            sc = fmaf(fabsf(f.A[i][k]), mag[k], sc);
        }
        const float slack = 4e-6f * sc + 1e-30f;
        out[i] = nextafterf(cw - hw - slack, -FLT_MAX);
        out[3 + i] = nextafterf(cw + hw + slack, FLT_MAX);
    }
}

__device__ __forceinline__ bool node_ov(const float* q, const BNode& n) {
    return !(q[0] > n.hi[0] || n.lo[0] > q[3] || q[1] > n.hi[1] || n.lo[1] > q[4] || q[2] > n.hi[2] ||
             n.lo[2] > q[5]);
}

// World-space node boxes of every body over the path q -> q - du (du == nullptr: at q).  One
// 256-thread workgroup per (body, triangle / edge tree).
__global__ __launch_bounds__(256) void k_world_nodes(Geo G, BodyMesh M, const double* __restrict__ q,
                                                      const double* __restrict__ du) {
    const int b = blockIdx.x;
    const bool tri = blockIdx.y == 0;
    const int m = M.body_mesh[b];
    const int r0 = tri ? M.mesh_tn0[m] : M.mesh_en0[m];
    const int n = (tri ? M.mesh_tn0[m + 1] : M.mesh_en0[m + 1]) - r0;
    const int w0 = tri ? M.body_tw0[b] : M.body_ew0[b];
    if ((tri ? M.body_tw0[b + 1] : M.body_ew0[b + 1]) - w0 != n) {
        if (threadIdx.x == 0) atomicOr(G.err, ERR_BVH);
        return;
    }
    const BNode* __restrict__ R = (tri ? M.tri_nodes : M.edge_nodes) + r0;
    BNode* __restrict__ W = (tri ? M.tri_world : M.edge_world) + w0;
    const Aff32 s0 = aff32(load_aff(q, nullptr, b));
    const Aff32 s1 = du ? aff32(load_aff(q, du, b)) : s0;
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
        const BNode r = R[i];
        float w[6];
        map_box32(s0, r.lo, r.hi, w);
        if (du) {
            float w1[6];
            map_box32(s1, r.lo, r.hi, w1);
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                w[k] = fminf(w[k], w1[k]);
                w[3 + k] = fmaxf(w[3 + k], w1[3 + k]);
            }
        }
        BNode o;
#pragma unroll
        for (int k = 0; k < 3; ++k) { o.lo[k] = w[k]; o.hi[k] = w[3 + k]; }
        o.prim = r.prim;
        o.skip = r.skip;
        W[i] = o;
    }
}

void launch_world_nodes(hipStream_t s, const Geo& G, const BodyMesh& M, const double* q, const double* du) {
    hipLaunchKernelGGL(k_world_nodes, dim3(G.nb, 2), dim3(256), 0, s, G, M, q, du);
}

// Stackless traversal of one body's world-space BVH (n nodes, nprim primitives) against the box
// qb; f(local primitive) at overlapping leaves.  The index strictly increases at every step, so the
// loop ends within n steps; returns false on a malformed node.
template <class F>
__device__ __forceinline__ bool world_query(const BNode* __restrict__ W, int n, int nprim, const float* qb, F&& f) {
    int i = 0;
    while (i < n) {
        const BNode nd = W[i];
        int next = nd.skip;
        if (node_ov(qb, nd)) {
            if (nd.prim >= 0) {
                if (nd.prim >= nprim) return false;
                f(nd.prim);
            } else {
                next = i + 1;
            }
        }
        if (next <= i || next > n) return false;
        i = next;
    }
    return true;
}

struct DBox {
    double lo[3], hi[3];
};

__device__ __forceinline__ void dbox_init(DBox& b) {
#pragma unroll
    for (int k = 0; k < 3; ++k) { b.lo[k] = DBL_MAX; b.hi[k] = -DBL_MAX; }
}
__device__ __forceinline__ void dbox_add(DBox& b, const double4& p) {
    b.lo[0] = fmin(b.lo[0], p.x); b.hi[0] = fmax(b.hi[0], p.x);
    b.lo[1] = fmin(b.lo[1], p.y); b.hi[1] = fmax(b.hi[1], p.y);
    b.lo[2] = fmin(b.lo[2], p.z); b.hi[2] = fmax(b.hi[2], p.z);
}
__device__ __forceinline__ void dbox_inflate(DBox& b, double r) {
#pragma unroll
    for (int k = 0; k < 3; ++k) { b.lo[k] -= r; b.hi[k] += r; }
}
__device__ __forceinline__ bool dbox_ov(const DBox& a, const DBox& b) {
#pragma unroll
    for (int k = 0; k < 3; ++k)
        if (a.lo[k] > b.hi[k] || b.lo[k] > a.hi[k]) return false;
    return true;
}
__device__ __forceinline__ void dbox_to_f(const DBox& b, double r, float* out) {
#pragma unroll
    for (int k = 0; k < 3; ++k) { out[k] = flo(b.lo[k] - r); out[3 + k] = fhi(b.hi[k] + r); }
}

__device__ __forceinline__ D3 d3of(const double4& a) { return mk(a.x, a.y, a.z); }

struct QueryCtx {
    Geo G;
    BodyMesh M;
    const double4* X0;
    const double4* X1;
    double infl;
    uint64_t* cand;
    int cap;
    int* count;
    int* overflow;
    int box_only;  // keep every leaf pair whose boxes overlap (the candidate cache), not the full rule
};


// Exact candidate rule on one leaf pair (global ids; PT: vertex i, triangle j; EE: edges i < j):
// the FP64 swept boxes overlap (point box inflated by infl, edge boxes by infl/2 each, triangle
// boxes not inflated; the oracle's broad_phase builds the same boxes with the same operations) and
// may_reach holds.  box_only: the box test alone.
__device__ __forceinline__ bool leaf_pass(const Geo& G, const double4* __restrict__ X0, const double4* __restrict__ X1,
                                          double infl, int box_only, int kind, int i, int j) {
    int vid[4];
    if (kind == K_PT) {
        const int4 tr = G.tris[j];
        vid[0] = i; vid[1] = tr.x; vid[2] = tr.y; vid[3] = tr.z;
    } else {
        const int2 e = G.edges[i], f = G.edges[j];
        vid[0] = e.x; vid[1] = e.y; vid[2] = f.x; vid[3] = f.y;
    }
    D3 a[4], b[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) { a[k] = d3of(X0[vid[k]]); b[k] = d3of(X1[vid[k]]); }
    const int split = (kind == K_PT) ? 1 : 2;
    const double r0 = (kind == K_PT) ? infl : 0.5 * infl, r1 = (kind == K_PT) ? 0.0 : 0.5 * infl;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        double lo0 = DBL_MAX, hi0 = -DBL_MAX, lo1 = DBL_MAX, hi1 = -DBL_MAX;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const double u = cmp(a[k], c), w = cmp(b[k], c);
            if (k < split) {
                lo0 = fmin(lo0, fmin(u, w)); hi0 = fmax(hi0, fmax(u, w));
            } else {
                lo1 = fmin(lo1, fmin(u, w)); hi1 = fmax(hi1, fmax(u, w));
            }
        }
        lo0 -= r0; hi0 += r0; lo1 -= r1; hi1 += r1;
        if (lo0 > hi1 || lo1 > hi0) return false;
    }
    return box_only || may_reach(kind, a, b, infl);
}

// appends the key of a leaf pair that passes the rule to the candidate list
__device__ __forceinline__ void leaf_rule(const QueryCtx& Q, int s, int kind, int i, int j) {
    const Geo& G = Q.G;
    if (!leaf_pass(G, Q.X0, Q.X1, Q.infl, Q.box_only, kind, i, j)) return;
    const uint64_t key = (kind == K_PT) ? make_key(s, K_PT, i - G.scene_v0[s], j - G.scene_t0[s])
                                        : make_key(s, K_EE, i - G.scene_e0[s], j - G.scene_e0[s]);
    const int slot = atomicAdd(Q.count, 1);
    if (slot < Q.cap) Q.cand[slot] = key;
    else atomicMax(Q.overflow, 1);
}

// leaf-pair queue entry: kind (bit 63) | i (bits 32-62) | j (bits 0-31)
__device__ __forceinline__ uint64_t leaf_pack(int kind, int i, int j) {
    return ((uint64_t)kind << 63) | ((uint64_t)(unsigned)i << 32) | (uint64_t)(unsigned)j;
}
__device__ __forceinline__ void leaf_run(const QueryCtx& Q, int s, uint64_t ent) {
    const int kind = opaque((int)(ent >> 63));
    const int i = opaque((int)((ent >> 32) & 0x7fffffffu));
    const int j = opaque((int)(ent & 0xffffffffu));
    leaf_rule(Q, s, kind, i, j);
}

#if defined(ABD_EMULATION)
#define WAVE_ANY(p) emu_any(p)
#else
#define WAVE_ANY(p) __any(p)
#endif

// Broad-phase queries: one wave per (body pair, chunk of 64 queries).  Query qi of the pair
// (A < B, same scene) is a vertex of A (qi < nvA), a vertex of B, or an edge of A; it walks the
// other body's world-space BVH (stackless, one node per lane per round) and pushes the leaves whose
// node box overlaps its query box into an LDS queue.  Whenever 64 leaf pairs are queued the wave
// evaluates them together, one per lane (leaf_rule), so the FP64 leaf work runs at full width and
// only surviving candidates touch the global counter.
__global__ __launch_bounds__(64) void k_pair_queries(QueryCtx Q, const int2* __restrict__ pairs, int npairs,
                                                      int chunks, const float* __restrict__ bbox) {
    __shared__ uint64_t queue[128];
    __shared__ int qn;
    // Workgroups are dealt to the 8 XCDs round-robin by linear index; XCD x takes the x-th
    // contiguous range of (pair, chunk) work items, so consecutive items on one XCD share bodies
    // (pairs are ordered by their lower body) and their BVH nodes stay in that XCD's L2.
    const long long nitems = (long long)npairs * chunks;
    const long long per = (nitems + 7) / 8;
    const long long item = (long long)(blockIdx.x & 7) * per + (blockIdx.x >> 3);
    if (item >= nitems) return;
    const int2 pr = pairs[item / chunks];
    const int chunk = (int)(item % chunks);
    const Geo& G = Q.G;
    const BodyMesh& M = Q.M;
    const int A = pr.x, B = pr.y;
    if (A < 0 || A >= G.nb || B <= A || B >= G.nb) {
        if (threadIdx.x == 0) atomicOr(G.err, ERR_PAIR);
        return;
    }
    const int lane = threadIdx.x;
    const int s = G.bscene[A];
    const int vA = M.body_v0[A], vB = M.body_v0[B];
    const int nvA = M.body_v0[A + 1] - vA, nvB = M.body_v0[B + 1] - vB;
    const int neA = M.body_e0[A + 1] - M.body_e0[A];
    const int total = nvA + nvB + neA;
    const int qi = chunk * 64 + lane;
    const double infl = Q.infl, h = 0.5 * Q.infl;
    // per-lane query: kind, primitive id, query box, the BVH to walk and its primitive base
    int kind = -1, pid = 0, prim0 = 0, nprim = 0, n = 0;
    const BNode* W = nullptr;
    float qb[6];
    if (qi < nvA + nvB) {
        const bool fromA = qi < nvA;
        const int O = fromA ? B : A;
        const int v = fromA ? vA + qi : vB + (qi - nvA);
        DBox vb;
        dbox_init(vb);
        dbox_add(vb, Q.X0[v]);
        dbox_add(vb, Q.X1[v]);
        dbox_to_f(vb, infl, qb);
        if (fbox_ov(qb, bbox + 6 * O)) {
            kind = K_PT; pid = v;
            prim0 = M.body_t0[O]; nprim = M.body_t0[O + 1] - prim0;
            W = M.tri_world + M.body_tw0[O]; n = M.body_tw0[O + 1] - M.body_tw0[O];
        }
    } else if (qi < total) {
        const int e = M.body_e0[A] + (qi - nvA - nvB);
        const int2 ev = G.edges[e];
        DBox eb;
        dbox_init(eb);
        dbox_add(eb, Q.X0[ev.x]); dbox_add(eb, Q.X0[ev.y]); dbox_add(eb, Q.X1[ev.x]); dbox_add(eb, Q.X1[ev.y]);
        dbox_to_f(eb, h, qb);
        if (fbox_ov(qb, bbox + 6 * B)) {
            // the other edge's box is inflated by h as well; node boxes are not
            dbox_to_f(eb, infl, qb);
            kind = K_EE; pid = e;
            prim0 = M.body_e0[B]; nprim = M.body_e0[B + 1] - prim0;
            W = M.edge_world + M.body_ew0[B]; n = M.body_ew0[B + 1] - M.body_ew0[B];
        }
    }
    if (lane == 0) qn = 0;
    __syncthreads();
    int i = 0;
    bool bad = false;
    while (WAVE_ANY(i < n)) {
        if (i < n) {
            const BNode nd = W[i];
            int next = nd.skip;
            if (node_ov(qb, nd)) {
                if (nd.prim >= 0) {
                    if (nd.prim < nprim) {
                        const int slot = atomicAdd(&qn, 1);
                        queue[slot] = leaf_pack(kind, pid, prim0 + nd.prim);
                    } else {
                        bad = true;
                    }
                } else {
                    next = i + 1;
                }
            }
            if (next <= i || next > n) {
                bad = true;
                next = n;
            }
            i = bad ? n : next;
        }
        __syncthreads();
        const int cnt = qn;
        if (cnt >= 64) {
            const uint64_t ent = queue[lane];
            __syncthreads();
            if (lane < cnt - 64) queue[lane] = queue[64 + lane];
            if (lane == 0) qn = cnt - 64;
            __syncthreads();
            leaf_run(Q, s, ent);
        }
    }
    const int cnt = qn;
    if (lane < cnt) leaf_run(Q, s, queue[lane]);
    if (bad) {
        atomicMax(Q.overflow, 2);
        atomicOr(G.err, ERR_BVH);
    }
}

void launch_pair_queries(hipStream_t s, const Geo& G, const BodyMesh& M, int npairs, int max_queries,
                         const int2* pairs, const double4* X0, const double4* X1, const float* bbox, double infl,
                         uint64_t* cand, int cap, int* count, int* overflow, int box_only) {
    if (npairs <= 0 || max_queries <= 0) return;
    QueryCtx Q{G, M, X0, X1, infl, cand, cap, count, overflow, box_only};
    const int chunks = (max_queries + 63) / 64;
    const long long nitems = (long long)npairs * chunks;
    const long long blocks = 8 * ((nitems + 7) / 8);
    hipLaunchKernelGGL(k_pair_queries, dim3((unsigned)blocks), dim3(64), 0, s, Q, pairs, npairs, chunks, bbox);
}

// point-ground candidates (one thread per vertex): min(d(X0), d(X1)) < dhat; box_only (the
// candidate cache): the lowest point of the vertex's box over the path comes within infl
__global__ void k_ground_queries(Geo G, const double4* __restrict__ X0, const double4* __restrict__ X1, AbdParams P,
                                 double infl, uint64_t* cand, int cap, int* count, int* overflow, int box_only) {
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= G.nv) return;
    const int b = (int)G.rest[v].w;
    if (G.pinned[b]) return;
    const double4 a = X0[v], c = X1[v];
    const double* n = P.ground_normal;
    double dmin;
    if (box_only) {
        const double pa[3] = {a.x, a.y, a.z}, pc[3] = {c.x, c.y, c.z};
        dmin = -P.ground_offset;
#pragma unroll
        for (int k = 0; k < 3; ++k) dmin += fmin(n[k] * fmin(pa[k], pc[k]), n[k] * fmax(pa[k], pc[k]));
    } else {
        const double d0 = n[0] * a.x + n[1] * a.y + n[2] * a.z - P.ground_offset;
        const double d1 = n[0] * c.x + n[1] * c.y + n[2] * c.z - P.ground_offset;
        dmin = fmin(d0, d1);
    }
    if (!(dmin < infl)) return;
    const int s = G.bscene[b];
    const int slot = atomicAdd(count, 1);
    if (slot < cap) cand[slot] = make_key(s, K_PG, v - G.scene_v0[s], 0);
    else atomicMax(overflow, 1);
}

void launch_ground_queries(hipStream_t s, const Geo& G, const double4* X0, const double4* X1, const AbdParams& P,
                           double infl, uint64_t* cand, int cap, int* count, int* overflow, int box_only) {
    if (!P.ground_enabled || G.nv <= 0) return;
    hipLaunchKernelGGL(k_ground_queries, dim3(cdivb(G.nv, 256)), dim3(256), 0, s, G, X0, X1, P, infl, cand, cap,
                       count, overflow, box_only);
}

// ---------------------------------------------------------------------------------------------
// Candidate cache.  A traversal over a path Xc0 -> Xc1 with every box widened by the margin m
// (query inflation dhat + 2m, box-only leaves) lists every pair whose boxes can overlap on any path
// each of whose vertex positions stays within the margin of that vertex's box over Xc0 -> Xc1.
// A later path X0 -> X1 inside that region (every vertex's end positions within m/2 of its cached
// box, the other half covering rounding) needs no traversal: the cached list, filtered by the exact
// rule above, is its candidate list, in the same sorted order.
// ---------------------------------------------------------------------------------------------
// active (optional, per scene): vertices of scenes whose Newton loop has stopped are skipped: their
// path has not moved since it was last found inside (the accepted point of a checked path lies in
// that path's box, and a stopped scene's direction is zero)
__global__ void k_cache_contains(Geo G, const double4* __restrict__ Xc0, const double4* __restrict__ Xc1,
                                 const double4* __restrict__ X0, const double4* __restrict__ X1, double m,
                                 const int* __restrict__ active, int* __restrict__ outside) {
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= G.nv) return;
    if (active && !active[G.bscene[(int)G.rest[v].w]]) return;
    const double4 c0 = Xc0[v], c1 = Xc1[v], a = X0[v], b = X1[v];
    const double cl[3] = {fmin(c0.x, c1.x) - m, fmin(c0.y, c1.y) - m, fmin(c0.z, c1.z) - m};
    const double ch[3] = {fmax(c0.x, c1.x) + m, fmax(c0.y, c1.y) + m, fmax(c0.z, c1.z) + m};
    const double pa[3] = {a.x, a.y, a.z}, pb[3] = {b.x, b.y, b.z};
    bool in = true;
#pragma unroll
    for (int k = 0; k < 3; ++k) in = in && pa[k] >= cl[k] && pa[k] <= ch[k] && pb[k] >= cl[k] && pb[k] <= ch[k];
    if (!in) atomicOr(outside, 1);
}

void launch_cache_contains(hipStream_t s, const Geo& G, const double4* Xc0, const double4* Xc1, const double4* X0,
                           const double4* X1, double m, const int* active, int* outside) {
    if (G.nv <= 0) return;
    hipLaunchKernelGGL(k_cache_contains, dim3(cdivb(G.nv, 256)), dim3(256), 0, s, G, Xc0, Xc1, X0, X1, m, active,
                       outside);
}

// flag[c] = 1 when cached pair c meets the exact candidate rule on X0 -> X1 (ground pairs: the
// vertex comes within infl of the plane)
__global__ void k_cache_filter(Geo G, Cnt cn, const uint64_t* __restrict__ keys, const double4* __restrict__ X0,
                               const double4* __restrict__ X1, AbdParams P, double infl, int* __restrict__ flag) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= cnt_of(cn)) return;
    const uint64_t key = keys[c];
    const int s = key_scene(key), kind = key_kind(key);
    const int il = opaque(key_il(key)), jl = opaque(key_jl(key));
    bool ok = false;
    if (s < G.ns) {
        if (kind == K_PT) {
            const int i = G.scene_v0[s] + il, j = G.scene_t0[s] + jl;
            ok = i < G.scene_v0[s + 1] && j < G.scene_t0[s + 1] && leaf_pass(G, X0, X1, infl, 0, K_PT, i, j);
        } else if (kind == K_EE) {
            const int i = G.scene_e0[s] + il, j = G.scene_e0[s] + jl;
            ok = i < G.scene_e0[s + 1] && j < G.scene_e0[s + 1] && leaf_pass(G, X0, X1, infl, 0, K_EE, i, j);
        } else if (kind == K_PG) {
            const int v = G.scene_v0[s] + il;
            if (v < G.scene_v0[s + 1]) {
                const double4 a = X0[v], b = X1[v];
                const double* nr = P.ground_normal;
                const double d0 = nr[0] * a.x + nr[1] * a.y + nr[2] * a.z - P.ground_offset;
                const double d1 = nr[0] * b.x + nr[1] * b.y + nr[2] * b.z - P.ground_offset;
                ok = fmin(d0, d1) < infl;
            }
        }
    }
    flag[c] = ok ? 1 : 0;
}

void launch_cache_filter(hipStream_t s, const Geo& G, Cnt n, const uint64_t* keys, const double4* X0,
                         const double4* X1, const AbdParams& P, double infl, int* flag) {
    if (n.cap <= 0) return;
    hipLaunchKernelGGL(k_cache_filter, dim3(cdivb(n.cap, 256)), dim3(256), 0, s, G, n, keys, X0, X1, P, infl, flag);
}

}  // namespace abd
