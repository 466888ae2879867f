// abd_broad.hip — broad phase of the ABD Newton step, built for affine bodies.
//
// The reference refits one world-space BVH over all edges at every detection
// (contact.py:244-250) and queries it per edge (contact.py:143-169).  Affine bodies never change
// their rest-frame meshes, so here every distinct mesh carries a rest-frame BVH built once on the
// host (abd_api.hip), and a detection over the linear path X0 -> X1 is
//   1. world boxes of the bodies over X0 and X1, inflated by dhat/2 (k_body_boxes);
//   2. the body pairs of each scene whose boxes overlap and which are not both pinned
//      (k_body_pairs);
//   3. world-space boxes of every body's BVH nodes: the rest-frame node box mapped through the
//      affine state x = c + A xr at both ends of the path (positions are linear in the path
//      parameter, so the swept image of the box lies in the union of its two world AABBs)
//      (k_world_nodes);
//   4. one workgroup per body pair: every vertex of either body against the other body's
//      triangle BVH, every edge of the lower body against the upper body's edge BVH, by a
//      stackless preorder traversal (k_pair_queries).
// Leaves apply the exact candidate rule in FP64, the same as the oracle's broad phase:
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
// 2. body pairs: thread t (grid-stride) takes the t-th pair of the scene-major numbering of the
//    strict upper triangles; pair_off[s] = pairs of the scenes before s.
// ---------------------------------------------------------------------------------------------
__global__ void k_body_pairs(Geo G, const long long* __restrict__ pair_off, long long total,
                             const float* __restrict__ box, int2* __restrict__ out, int cap, int* count,
                             int* overflow) {
    for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < total;
         t += (long long)gridDim.x * blockDim.x) {
        int lo = 0, hi = G.ns;
        while (hi - lo > 1) {
            int mid = (lo + hi) >> 1;
            if (pair_off[mid] <= t) lo = mid; else hi = mid;
        }
        const int s = lo;
        const long long r = t - pair_off[s];
        const long long n = G.scene_b0[s + 1] - G.scene_b0[s];
        // row i holds the pairs (i, j > i); it starts at i (2n - i - 1) / 2
        const double nn = (double)n - 0.5;
        long long i = (long long)floor(nn - sqrt(fmax(nn * nn - 2.0 * (double)r, 0.0)));
        if (i < 0) i = 0;
        if (i > n - 2) i = n - 2;
        while (i > 0 && i * (2 * n - i - 1) / 2 > r) --i;
        while (i + 1 <= n - 2 && (i + 1) * (2 * n - i - 2) / 2 <= r) ++i;
        const long long j = i + 1 + (r - i * (2 * n - i - 1) / 2);
        if (i < 0 || j <= i || j >= n) {
            atomicOr(G.err, ERR_PAIR);
            continue;
        }
        const int A = G.scene_b0[s] + (int)i, B = G.scene_b0[s] + (int)j;
        if (G.pinned[A] && G.pinned[B]) continue;
        if (!fbox_ov(box + 6 * A, box + 6 * B)) continue;
        const int slot = atomicAdd(count, 1);
        if (slot < cap) out[slot] = make_int2(A, B);
        else atomicMax(overflow, 1);
    }
}

void launch_body_pairs(hipStream_t s, const Geo& G, const long long* pair_off, long long total, const float* box,
                       int2* out, int cap, int* count, int* overflow) {
    if (total <= 0) return;
    long long blocks = (total + 255) / 256;
    if (blocks > 8192) blocks = 8192;
    hipLaunchKernelGGL(k_body_pairs, dim3((unsigned)blocks), dim3(256), 0, s, G, pair_off, total, box, out, cap,
                       count, overflow);
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

// World AABB of a rest-frame box under an affine state, widened by a relative 1e-12 of the
// magnitudes involved (far above the rounding of x = c + A xr in FP64) and rounded outward.
__device__ __forceinline__ void map_box(const Aff& f, const float* rlo, const float* rhi, float* out) {
    double ctr[3], half[3], mag[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        ctr[k] = 0.5 * ((double)rlo[k] + (double)rhi[k]);
        half[k] = 0.5 * ((double)rhi[k] - (double)rlo[k]);
        mag[k] = fmax(fabs((double)rlo[k]), fabs((double)rhi[k]));
    }
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        double cw = f.c[i], hw = 0.0, sc = fabs(f.c[i]);
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            cw += f.A[i][k] * ctr[k];
            hw += fabs(f.A[i][k]) * half[k];
            sc += fabs(f.A[i][k]) * mag[k];
        }
        const double slack = 1e-12 * sc + 1e-280;
        out[i] = flo(cw - hw - slack);
        out[3 + i] = fhi(cw + hw + slack);
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
    const Aff s0 = load_aff(q, nullptr, b);
    const Aff s1 = du ? load_aff(q, du, b) : s0;
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
        const BNode r = R[i];
        float w[6];
        map_box(s0, r.lo, r.hi, w);
        if (du) {
            float w1[6];
            map_box(s1, r.lo, r.hi, w1);
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
};

__device__ __forceinline__ void emit_key(const QueryCtx& Q, uint64_t key) {
    const int slot = atomicAdd(Q.count, 1);
    if (slot < Q.cap) Q.cand[slot] = key;
    else atomicMax(Q.overflow, 1);
}

// One 256-thread workgroup per body pair (A < B, same scene).
__global__ __launch_bounds__(256) void k_pair_queries(QueryCtx Q, const int2* __restrict__ pairs,
                                                       const float* __restrict__ bbox) {
    const int2 pr = pairs[blockIdx.x];
    const Geo& G = Q.G;
    const BodyMesh& M = Q.M;
    const int A = pr.x, B = pr.y;
    if (A < 0 || A >= G.nb || B <= A || B >= G.nb) {
        if (threadIdx.x == 0) atomicOr(G.err, ERR_PAIR);
        return;
    }
    const int s = G.bscene[A];
    const int vA = M.body_v0[A], vB = M.body_v0[B];
    const int nvA = M.body_v0[A + 1] - vA, nvB = M.body_v0[B + 1] - vB;
    const int neA = M.body_e0[A + 1] - M.body_e0[A];
    const int total = nvA + nvB + neA;
    const double infl = Q.infl, h = 0.5 * Q.infl;
    const int sv0 = G.scene_v0[s], se0 = G.scene_e0[s], st0 = G.scene_t0[s];
    for (int qi = threadIdx.x; qi < total; qi += blockDim.x) {
        if (qi < nvA + nvB) {
            // vertex v of body P against the triangles of body O
            const bool fromA = qi < nvA;
            const int O = fromA ? B : A;
            const int v = fromA ? vA + qi : vB + (qi - nvA);
            const double4 x0 = Q.X0[v], x1 = Q.X1[v];
            DBox vb;
            dbox_init(vb);
            dbox_add(vb, x0);
            dbox_add(vb, x1);
            dbox_inflate(vb, infl);
            float qb[6];
            dbox_to_f(vb, 0.0, qb);
            if (!fbox_ov(qb, bbox + 6 * O)) continue;
            const int t0 = M.body_t0[O], t1 = M.body_t0[O + 1];
            const int w0 = M.body_tw0[O];
            bool ok = world_query(M.tri_world + w0, M.body_tw0[O + 1] - w0, t1 - t0, qb, [&](int lt) {
                const int t = t0 + lt;
                const int4 tr = G.tris[t];
                const double4 p0 = Q.X0[tr.x], p1 = Q.X0[tr.y], p2 = Q.X0[tr.z];
                const double4 r0 = Q.X1[tr.x], r1 = Q.X1[tr.y], r2 = Q.X1[tr.z];
                DBox tb;
                dbox_init(tb);
                dbox_add(tb, p0); dbox_add(tb, p1); dbox_add(tb, p2);
                dbox_add(tb, r0); dbox_add(tb, r1); dbox_add(tb, r2);
                if (!dbox_ov(vb, tb)) return;
                D3 y0[4] = {d3of(x0), d3of(p0), d3of(p1), d3of(p2)};
                D3 y1[4] = {d3of(x1), d3of(r0), d3of(r1), d3of(r2)};
                if (may_reach(K_PT, y0, y1, infl)) emit_key(Q, make_key(s, K_PT, v - sv0, t - st0));
            });
            if (!ok) { atomicMax(Q.overflow, 2); atomicOr(G.err, ERR_BVH); }
        } else {
            // edge e of body A against the edges of body B
            const int e = M.body_e0[A] + (qi - nvA - nvB);
            const int2 ev = G.edges[e];
            const double4 p0 = Q.X0[ev.x], p1 = Q.X0[ev.y], r0 = Q.X1[ev.x], r1 = Q.X1[ev.y];
            DBox eb;
            dbox_init(eb);
            dbox_add(eb, p0); dbox_add(eb, p1); dbox_add(eb, r0); dbox_add(eb, r1);
            dbox_inflate(eb, h);
            float qb[6];
            dbox_to_f(eb, 0.0, qb);
            if (!fbox_ov(qb, bbox + 6 * B)) continue;
            // the other edge's box is inflated by h as well; node boxes are not
            dbox_to_f(eb, h, qb);
            const int f0 = M.body_e0[B], f1 = M.body_e0[B + 1];
            const int w0 = M.body_ew0[B];
            bool ok = world_query(M.edge_world + w0, M.body_ew0[B + 1] - w0, f1 - f0, qb, [&](int lf) {
                const int f = f0 + lf;
                const int2 fv = G.edges[f];
                const double4 g0 = Q.X0[fv.x], g1 = Q.X0[fv.y], k0 = Q.X1[fv.x], k1 = Q.X1[fv.y];
                DBox fb;
                dbox_init(fb);
                dbox_add(fb, g0); dbox_add(fb, g1); dbox_add(fb, k0); dbox_add(fb, k1);
                dbox_inflate(fb, h);
                if (!dbox_ov(eb, fb)) return;
                D3 y0[4] = {d3of(p0), d3of(p1), d3of(g0), d3of(g1)};
                D3 y1[4] = {d3of(r0), d3of(r1), d3of(k0), d3of(k1)};
                if (may_reach(K_EE, y0, y1, infl)) emit_key(Q, make_key(s, K_EE, e - se0, f - se0));
            });
            if (!ok) { atomicMax(Q.overflow, 2); atomicOr(G.err, ERR_BVH); }
        }
    }
}

void launch_pair_queries(hipStream_t s, const Geo& G, const BodyMesh& M, int npairs, const int2* pairs,
                         const double4* X0, const double4* X1, const float* bbox, double infl, uint64_t* cand,
                         int cap, int* count, int* overflow) {
    if (npairs <= 0) return;
    QueryCtx Q{G, M, X0, X1, infl, cand, cap, count, overflow};
    hipLaunchKernelGGL(k_pair_queries, dim3(npairs), dim3(256), 0, s, Q, pairs, bbox);
}

// point-ground candidates (one thread per vertex): min(d(X0), d(X1)) < dhat
__global__ void k_ground_queries(Geo G, const double4* __restrict__ X0, const double4* __restrict__ X1, AbdParams P,
                                 double infl, uint64_t* cand, int cap, int* count, int* overflow) {
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= G.nv) return;
    const int b = (int)G.rest[v].w;
    if (G.pinned[b]) return;
    const double4 a = X0[v], c = X1[v];
    const double* n = P.ground_normal;
    const double d0 = n[0] * a.x + n[1] * a.y + n[2] * a.z - P.ground_offset;
    const double d1 = n[0] * c.x + n[1] * c.y + n[2] * c.z - P.ground_offset;
    if (!(fmin(d0, d1) < infl)) return;
    const int s = G.bscene[b];
    const int slot = atomicAdd(count, 1);
    if (slot < cap) cand[slot] = make_key(s, K_PG, v - G.scene_v0[s], 0);
    else atomicMax(overflow, 1);
}

void launch_ground_queries(hipStream_t s, const Geo& G, const double4* X0, const double4* X1, const AbdParams& P,
                           double infl, uint64_t* cand, int cap, int* count, int* overflow) {
    if (!P.ground_enabled || G.nv <= 0) return;
    hipLaunchKernelGGL(k_ground_queries, dim3(cdivb(G.nv, 256)), dim3(256), 0, s, G, X0, X1, P, infl, cand, cap,
                       count, overflow);
}

}  // namespace abd
