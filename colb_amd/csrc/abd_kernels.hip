// abd_kernels.hip — CDNA4 (gfx950) kernels of the ABD Newton step with IPC barriers.
//
// Data layout in HBM (structure of arrays across bodies and primitives):
//   rest   double4[nv]  body-frame rest position + owning body (w)
//   X      double4[nv]  world positions x = c + sum_k xr_k a_k (abd_simple.py:26-35)
//   q      double[nb*12] DOF vectors [c, a0, a1, a2]; grad/du/x/r/z/p the same layout
//   diag   double[nb*144] 12x12 diagonal blocks; off/offT double[npairs*144] off-diagonal blocks
//   cand   uint64[ncand] sorted candidate-pair keys (scene | kind | i | j)
// Every reduction is per scene, over fixed ranges in a fixed order, so results do not depend
// on scheduling (the reference's wp.atomic_add sums, abd.py:64, 75, 301-303, do).
#include <hip/hip_runtime.h>
#include <algorithm>
#include <climits>
#include <float.h>
#include <stdint.h>

#include "abd_device.hpp"
#include "abd_internal.hpp"
#include "abd_mlp.hpp"

namespace abd {

static inline int cdiv(long a, int b) { return (int)((a + b - 1) / b); }

// device-side key field extraction through the optimisation barrier (see opaque())
__device__ __forceinline__ int dkey_scene(uint64_t k) { return opaque((int)(k >> 48)); }
__device__ __forceinline__ int dkey_kind(uint64_t k) { return opaque((int)((k >> 46) & 3)); }
__device__ __forceinline__ int dkey_il(uint64_t k) { return opaque((int)((k >> 23) & KEY_IMASK)); }
__device__ __forceinline__ int dkey_jl(uint64_t k) { return opaque((int)(k & KEY_IMASK)); }

// ---------------------------------------------------------------------------------------------
// state kernels
// ---------------------------------------------------------------------------------------------
// tildeA / tildep (abd.py:160-166): q~ = q_now + dt*qdot + dt^2*(g, 0, 0, 0)
__global__ void k_qtilde(int n, const double* __restrict__ qnow, const double* __restrict__ qdot, double dt,
                         double g0, double g1, double g2, double* __restrict__ qt) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int k = i % 12;
    double v = qnow[i] + dt * qdot[i];
    if (k < 3) v += dt * dt * (k == 0 ? g0 : (k == 1 ? g1 : g2));
    qt[i] = v;
}

void launch_qtilde(hipStream_t s, int nb, const double* qnow, const double* qdot, double dt, const double* g, double* qt) {
    int n = nb * 12;
    hipLaunchKernelGGL(k_qtilde, dim3(cdiv(n, 256)), dim3(256), 0, s, n, qnow, qdot, dt, g[0], g[1], g[2], qt);
}

// compute_V (abd_simple.py:26-35) at q - alpha_s * du
// blockIdx.y = k > 0 (line-search trials in one launch): the vertices at q - 2^-k alpha du into
// X + k nv
__global__ void k_xform(Geo G, const double* __restrict__ q, const double* __restrict__ du,
                        const double* __restrict__ alpha, double4* __restrict__ X) {
    int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= G.nv) return;
    X += (size_t)blockIdx.y * G.nv;
    double4 r = G.rest[v];
    int b = (int)r.w;
    const double* qb = q + 12 * b;
    double a = 0.0;
    if (du) a = alpha ? ldexp(alpha[G.bscene[b]], -(int)blockIdx.y) : 1.0;
    double c[12];
#pragma unroll
    for (int k = 0; k < 12; ++k) c[k] = du ? qb[k] - a * du[12 * b + k] : qb[k];
    X[v] = make_double4(c[0] + r.x * c[3] + r.y * c[6] + r.z * c[9], c[1] + r.x * c[4] + r.y * c[7] + r.z * c[10],
                        c[2] + r.x * c[5] + r.y * c[8] + r.z * c[11], 0.0);
}

void launch_xform(hipStream_t s, const Geo& G, const double* q, const double* du, const double* alpha, double4* X,
                  int trials) {
    hipLaunchKernelGGL(k_xform, dim3(cdiv(G.nv, 256), trials), dim3(256), 0, s, G, q, du, alpha, X);
}

// ---------------------------------------------------------------------------------------------
// narrow phase
// ---------------------------------------------------------------------------------------------
struct PairV {
    int kind, scene, nvp;
    int v[4];
};

// Decodes a candidate key into its vertices.  Every index is range-checked against the scene's
// primitive ranges; a key that fails sets ERR_DECODE and decodes to kind -1 (callers skip it).
__device__ __forceinline__ PairV decode(const Geo& G, uint64_t key) {
    PairV p;
    p.kind = dkey_kind(key);
    p.scene = dkey_scene(key);
    p.nvp = 0;
    p.v[0] = p.v[1] = p.v[2] = p.v[3] = 0;
    int il = dkey_il(key), jl = dkey_jl(key);
    if (p.scene >= G.ns) { atomicOr(G.err, ERR_DECODE); p.kind = -1; p.scene = 0; return p; }
    const int s = p.scene;
    if (p.kind == K_PT) {
        int v = G.scene_v0[s] + il, t = G.scene_t0[s] + jl;
        if (v >= G.scene_v0[s + 1] || t >= G.scene_t0[s + 1]) { atomicOr(G.err, ERR_DECODE); p.kind = -1; return p; }
        p.v[0] = v;
        int4 tr = G.tris[t];
        p.v[1] = tr.x; p.v[2] = tr.y; p.v[3] = tr.z;
        p.nvp = 4;
    } else if (p.kind == K_EE) {
        int e0 = G.scene_e0[s];
        if (e0 + il >= G.scene_e0[s + 1] || e0 + jl >= G.scene_e0[s + 1]) { atomicOr(G.err, ERR_DECODE); p.kind = -1; return p; }
        int2 a = G.edges[e0 + il], b = G.edges[e0 + jl];
        p.v[0] = a.x; p.v[1] = a.y; p.v[2] = b.x; p.v[3] = b.y;
        p.nvp = 4;
    } else if (p.kind == K_PG) {
        int v = G.scene_v0[s] + il;
        if (v >= G.scene_v0[s + 1]) { atomicOr(G.err, ERR_DECODE); p.kind = -1; return p; }
        p.v[0] = v;
        p.v[1] = p.v[2] = p.v[3] = v;
        p.nvp = 1;
    } else {
        atomicOr(G.err, ERR_DECODE);
        p.kind = -1;
    }
    return p;
}

__device__ __forceinline__ D3 ld3(const double4* X, int v) {
    double4 a = X[v];
    return mk(a.x, a.y, a.z);
}

// contact stiffness of a pair: the larger of its bodies' (the first and last stencil slot belong to
// the two bodies; ground pairs have one)
__device__ __forceinline__ int pair_body(const Geo& G, const PairV& p, int slot) {
    return (int)G.rest[p.v[slot]].w;
}
__device__ __forceinline__ double pair_stiffness(const Geo& G, const PairV& p) {
    const double ka = G.kb[pair_body(G, p, 0)];
    return p.kind == K_PG ? ka : fmax(ka, G.kb[pair_body(G, p, 3)]);
}

// signed distance of vertex v above the ground plane n.x = h
__device__ __forceinline__ double ground_gap(const double4* X, int v, const AbdParams& P) {
    D3 x = ld3(X, v);
    return P.ground_normal[0] * x.x + P.ground_normal[1] * x.y + P.ground_normal[2] * x.z - P.ground_offset;
}

__device__ __forceinline__ double cand_d2(const Geo& G, const PairV& p, const double4* X, const AbdParams& P) {
    if (p.kind == K_PG) {
        const double d = ground_gap(X, p.v[0], P);
        return d * d;
    }
    D3 x[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) x[k] = ld3(X, p.v[k]);
    return pair_d2(p.kind, x);
}

// energy_contact (abd.py:66-75): dt^2 k B(d^2) for active pairs (d < dhat).  active (optional, per
// scene, energies only): candidates of scenes whose Newton loop has stopped get energy 0 (their
// line search is decided already)
// blockIdx.y = k > 0 (line-search trials in one launch): vertices X + k nv, energies energy + k estride
__global__ void k_cand_eval(Geo G, Cnt cn, const uint64_t* __restrict__ keys, const double4* __restrict__ X, AbdParams P,
                            int* __restrict__ flag, double* __restrict__ energy, const int* __restrict__ active,
                            size_t estride) {
    const int n = cnt_of(cn);
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    X += (size_t)blockIdx.y * G.nv;
    if (energy) energy += (size_t)blockIdx.y * estride;
    if (active) {
        const int sc = dkey_scene(keys[i]);
        if (sc >= 0 && sc < G.ns && !active[sc]) {
            energy[i] = 0.0;
            return;
        }
    }
    PairV p = decode(G, keys[i]);
    if (p.kind < 0) {
        if (flag) flag[i] = 0;
        if (energy) energy[i] = 0.0;
        return;
    }
    double d2 = cand_d2(G, p, X, P);
    double shat = P.dhat * P.dhat;
    double B, B1, B2;
    barrier3(d2, shat, B, B1, B2);
    if (flag) flag[i] = d2 < shat ? 1 : 0;
    // a vertex on or below the ground plane is infeasible: infinite energy, so no line-search
    // trial that reaches it is accepted
    const bool through = p.kind == K_PG && ground_gap(X, p.v[0], P) <= 0.0;
    if (energy) energy[i] = through ? (double)INFINITY : P.dt * P.dt * pair_stiffness(G, p) * B;
}

void launch_cand_eval(hipStream_t s, const Geo& G, Cnt n, const uint64_t* keys, const double4* X, const AbdParams& P,
                      int* flag, double* energy, const int* active, int trials, size_t estride) {
    if (n.cap <= 0) return;
    hipLaunchKernelGGL(k_cand_eval, dim3(cdiv(n.cap, 256), trials), dim3(256), 0, s, G, n, keys, X, P, flag, energy,
                       flag ? (const int*)nullptr : active, estride);
}

__global__ void k_pair_d2(Geo G, int n, const uint64_t* __restrict__ keys, const double4* __restrict__ X, AbdParams P,
                          double* __restrict__ d2) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    PairV p = decode(G, keys[i]);
    d2[i] = (p.kind < 0) ? 0.0 : cand_d2(G, p, X, P);
}

void launch_pair_d2(hipStream_t s, const Geo& G, int n, const uint64_t* keys, const double4* X, const AbdParams& P,
                    double* d2) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_pair_d2, dim3(cdiv(n, 256)), dim3(256), 0, s, G, n, keys, X, P, d2);
}

// out[pos[i]] = keys[i] where flag[i]; n_out (optional): the compacted length pos[n-1] + flag[n-1]
// (0 for an empty input)
__global__ void k_compact(Cnt cn, const int* __restrict__ flag, const int* __restrict__ pos,
                          const uint64_t* __restrict__ keys, uint64_t* __restrict__ out, int* n_out) {
    const int n = cnt_of(cn);
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n && flag[i]) out[pos[i]] = keys[i];
    if (n_out && i == max(n - 1, 0)) *n_out = n > 0 ? pos[n - 1] + flag[n - 1] : 0;
}

void launch_compact(hipStream_t s, Cnt n, const int* flag, const int* pos, const uint64_t* keys, uint64_t* out,
                    int* n_out) {
    if (n.cap <= 0) return;
    hipLaunchKernelGGL(k_compact, dim3(cdiv(n.cap, 256)), dim3(256), 0, s, n, flag, pos, keys, out, n_out);
}

// contact_hessian_ee (abd.py:198-571), one wave per active contact: d^2 gradient/Hessian,
// barrier chain rule, PSD projection of the whole barrier Hessian, then the affine Jacobian
// J = [1, xr^T] (x) I3 (abd.py:256-263) maps it to the three 12x12 body blocks
// (A,A), (B,B), (A,B) and two 12-vectors.  HT: the type of the barrier Hessian's projection
// (double; float in the mixed-precision assembly, whose Hessian feeds only the FP32 PCG); the
// distances, the gradient and the block sums stay FP64.
template <typename HT>
__global__ __launch_bounds__(64) void k_contact_hess(Geo G, Cnt cn, const uint64_t* __restrict__ act,
                                                      const double4* __restrict__ X, AbdParams P,
                                                      int2* __restrict__ body, double* __restrict__ g24,
                                                      double* __restrict__ H432) {
    __shared__ HT A[144];
    __shared__ double gx[12], jv[4][4];
    __shared__ int side[4];
    const int c = blockIdx.x;
    const int lane = threadIdx.x;
    if (c >= cnt_of(cn)) return;
    PairV p = decode(G, act[c]);
    if (p.kind < 0) {
        // invalid key (already flagged): contribute nothing
        for (int e = lane; e < 432; e += 64) H432[(size_t)c * 432 + e] = 0.0;
        if (lane < 24) g24[(size_t)c * 24 + lane] = 0.0;
        if (lane == 0) body[c] = make_int2(-1, -1);
        return;
    }
    const double kd = pair_stiffness(G, p) * P.dt * P.dt;
    const double shat = P.dhat * P.dhat;
    int m, ns;
    int st[4] = {0, 1, 2, 3};
    if (p.kind == K_PG) {
        D3 x = ld3(X, p.v[0]);
        D3 nrm = mk(P.ground_normal[0], P.ground_normal[1], P.ground_normal[2]);
        double d = dot(nrm, x) - P.ground_offset;
        if (!(d > 0.0)) {
            // on or below the plane: an infeasible state (the host reports it), no contribution
            if (lane == 0) atomicOr(G.err, ERR_PENETRATION);
            d = 0.0;
        }
        double B, B1, B2;
        barrier3(d * d, shat, B, B1, B2);
        const double lam0 = kd * (2.0 * B1 + 4.0 * B2 * d * d);
        const double lam = P.hess_proj == 1 ? fabs(lam0) : fmax(0.0, lam0);
        m = 3; ns = 1;
        if (lane < 9) A[lane] = (HT)(lam * cmp(nrm, lane / 3) * cmp(nrm, lane % 3));
        if (lane < 3) gx[lane] = kd * B1 * 2.0 * d * cmp(nrm, lane);
        __syncthreads();
    } else {
        D3 x[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) x[k] = ld3(X, p.v[k]);
        DistEval E;
        dist_eval(p.kind, x, E);
        ns = E.S.ns;
        m = 3 * ns;
        for (int k = 0; k < 4; ++k) st[k] = E.S.st[k];
        double B, B1, B2;
        barrier3(E.d2, shat, B, B1, B2);
        for (int e = lane; e < m * m; e += 64) {
            int R = e / m, C = e % m;
            A[e] = (HT)(kd * (B1 * d2_hess(E, R, C) + B2 * d2_grad(E, R) * d2_grad(E, C)));
        }
        if (lane < m) gx[lane] = kd * B1 * d2_grad(E, lane);
        __syncthreads();
        wave_psd_project(A, m, lane, P.hess_proj == 1);
    }
    // stencil point k -> vertex -> body side and Jacobian row [1, xr]
    const int bA = (int)G.rest[p.v[0]].w;
    const int bB = (p.kind == K_PG) ? -1 : (int)G.rest[p.v[3]].w;
    if (lane < ns) {
        const int sl = (lane == 0) ? st[0] : (lane == 1) ? st[1] : (lane == 2) ? st[2] : st[3];
        const int vid = (sl == 0) * p.v[0] + (sl == 1) * p.v[1] + (sl == 2) * p.v[2] + (sl == 3) * p.v[3];
        double4 r = G.rest[vid];
        side[lane] = ((int)r.w == bA) ? 0 : 1;
        jv[lane][0] = 1.0; jv[lane][1] = r.x; jv[lane][2] = r.y; jv[lane][3] = r.z;
    }
    __syncthreads();
    double* Hout = H432 + (size_t)c * 432;
    for (int e = lane; e < 432; e += 64) {
        int blk = e / 144, rc = e % 144;
        int R = rc / 12, C = rc % 12;
        int i = R / 3, r = R % 3, j = C / 3, cc = C % 3;
        int sa = (blk == 1) ? 1 : 0, sb = (blk == 0) ? 0 : 1;
        double h = 0.0;
        for (int k = 0; k < ns; ++k) {
            if (side[k] != sa) continue;
            for (int l = 0; l < ns; ++l) {
                if (side[l] != sb) continue;
                h += jv[k][i] * jv[l][j] * (double)A[(3 * k + r) * m + 3 * l + cc];
            }
        }
        Hout[e] = h;
    }
    if (lane < 24) {
        int sd = lane / 12, i = (lane % 12) / 3, r = lane % 3;
        double g = 0.0;
        for (int k = 0; k < ns; ++k)
            if (side[k] == sd) g += jv[k][i] * gx[3 * k + r];
        g24[(size_t)c * 24 + lane] = g;
    }
    if (lane == 0) body[c] = make_int2(bA, bB);
}

void launch_contact_hess(hipStream_t s, const Geo& G, Cnt n, const uint64_t* act, const double4* X, const AbdParams& P,
                         int2* body, double* g24, double* H432) {
    if (n.cap <= 0) return;
    if (P.precision == 1)
        hipLaunchKernelGGL(k_contact_hess<float>, dim3(n.cap), dim3(64), 0, s, G, n, act, X, P, body, g24, H432);
    else
        hipLaunchKernelGGL(k_contact_hess<double>, dim3(n.cap), dim3(64), 0, s, G, n, act, X, P, body, g24, H432);
}

// PSD projection of n symmetric m x m matrices (row-major FP64, rounded to HT), one wave each: the
// routine of k_contact_hess and k_body_assemble on given matrices
template <typename HT>
__global__ __launch_bounds__(64) void k_psd_test(int m, int absolute, int n, const double* __restrict__ A,
                                                 double* __restrict__ out) {
    __shared__ HT S[144];
    const int c = blockIdx.x, lane = threadIdx.x;
    if (c >= n) return;
    for (int e = lane; e < m * m; e += 64) S[e] = (HT)A[(size_t)c * m * m + e];
    __syncthreads();
    wave_psd_project(S, m, lane, absolute != 0);
    for (int e = lane; e < m * m; e += 64) out[(size_t)c * m * m + e] = (double)S[e];
}

void launch_psd_test(hipStream_t s, int m, int precision, int absolute, int n, const double* A, double* out) {
    if (precision == 1)
        hipLaunchKernelGGL(k_psd_test<float>, dim3(n), dim3(64), 0, s, m, absolute, n, A, out);
    else
        hipLaunchKernelGGL(k_psd_test<double>, dim3(n), dim3(64), 0, s, m, absolute, n, A, out);
}

// incidence keys (body << 32 | contact << 1 | role) and body-pair keys (min(20) | max(20) | contact(24))
__global__ void k_incidence(Geo G, Cnt cn, const int2* __restrict__ body, uint64_t* inc, int* n_inc, uint64_t* pairk,
                            int* n_pair) {
    const int n = cnt_of(cn);
    int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= n) return;
    int2 b = body[c];
    if (b.x < 0 || b.x >= G.nb || b.y >= G.nb) {
        if (b.x != -1) atomicOr(G.err, ERR_BODY);
        return;
    }
    bool dA = !G.pinned[b.x];
    bool dB = b.y >= 0 && !G.pinned[b.y];
    if (dA) inc[atomicAdd(n_inc, 1)] = ((uint64_t)b.x << 32) | ((uint64_t)c << 1);
    if (dB) inc[atomicAdd(n_inc, 1)] = ((uint64_t)b.y << 32) | ((uint64_t)c << 1) | 1ull;
    if (dA && dB) {
        uint64_t lo = min(b.x, b.y), hi = max(b.x, b.y);
        pairk[atomicAdd(n_pair, 1)] = (lo << 44) | (hi << 24) | (uint64_t)c;
    }
}

void launch_incidence(hipStream_t s, const Geo& G, Cnt n, const int2* body, uint64_t* inc, int* n_inc, uint64_t* pairk,
                      int* n_pair) {
    if (n.cap <= 0) return;
    hipLaunchKernelGGL(k_incidence, dim3(cdiv(n.cap, 256)), dim3(256), 0, s, G, n, body, inc, n_inc, pairk, n_pair);
}

__global__ void k_run_heads(Cnt cn, const uint64_t* __restrict__ k, int shift, int* head) {
    const int n = cnt_of(cn);
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    head[i] = (i == 0 || (k[i] >> shift) != (k[i - 1] >> shift)) ? 1 : 0;
}

void launch_run_heads(hipStream_t s, Cnt n, const uint64_t* sorted, int shift, int* head) {
    if (n.cap <= 0) return;
    hipLaunchKernelGGL(k_run_heads, dim3(cdiv(n.cap, 256)), dim3(256), 0, s, n, sorted, shift, head);
}

// n_unique receives the number of body pairs and n_ent twice that (the CSR entries); an empty
// list writes zeros (the launch always has a workgroup)
__global__ void k_pair_unique(Cnt cn, const uint64_t* __restrict__ k, const int* __restrict__ head,
                              const int* __restrict__ pos, int* pair_b, int* run_start, int* n_unique, int* n_ent) {
    const int n = cnt_of(cn);
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (n == 0 && i == 0) {
        run_start[0] = 0;
        *n_unique = 0;
        *n_ent = 0;
    }
    if (i >= n) return;
    if (head[i]) {
        int u = pos[i];
        pair_b[2 * u] = (int)(k[i] >> 44);
        pair_b[2 * u + 1] = opaque((int)((k[i] >> 24) & 0xFFFFF));
        run_start[u] = i;
    }
    if (i == n - 1) {
        int total = pos[i] + head[i];
        run_start[total] = n;
        *n_unique = total;
        *n_ent = 2 * total;
    }
}

void launch_pair_unique(hipStream_t s, Cnt n, const uint64_t* sorted, const int* head, const int* pos, int* pair_b,
                        int* run_start, int* n_unique, int* n_ent) {
    hipLaunchKernelGGL(k_pair_unique, dim3(std::max(1, cdiv(n.cap, 256))), dim3(256), 0, s, n, sorted, head, pos,
                       pair_b, run_start, n_unique, n_ent);
}

// off-diagonal block (bmin, bmax) = sum over its contacts of H_AB (or H_AB^T when the contact's
// first body is bmax), in sorted contact order
__global__ void k_offblock_sum(Cnt cpairs, const int* __restrict__ pair_b, const int* __restrict__ run_start,
                               const uint64_t* __restrict__ k, const int2* __restrict__ body,
                               const double* __restrict__ H432, double* __restrict__ off, double* __restrict__ offT,
                               Cnt cact, int* err) {
    const int n_pairs = cnt_of(cpairs), n_act = cnt_of(cact);
    long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long)n_pairs * 144) return;
    int u = (int)(t / 144), e = (int)(t % 144);
    int R = e / 12, C = e % 12;
    int bmin = pair_b[2 * u];
    double h = 0.0;
    for (int m = run_start[u]; m < run_start[u + 1]; ++m) {
        int c = opaque((int)(k[m] & 0xFFFFFF));
        if (c >= n_act) { atomicOr(err, ERR_PAIR); break; }
        const double* Hc = H432 + (size_t)c * 432 + 288;
        h += (body[c].x == bmin) ? Hc[R * 12 + C] : Hc[C * 12 + R];
    }
    off[(size_t)u * 144 + e] = h;
    offT[(size_t)u * 144 + C * 12 + R] = h;
}

void launch_offblock_sum(hipStream_t s, Cnt n_pairs, const int* pair_b, const int* run_start,
                         const uint64_t* sorted, const int2* body, const double* H432, double* off, double* offT,
                         Cnt n_act, int* err) {
    if (n_pairs.cap <= 0) return;
    long n = (long)n_pairs.cap * 144;
    hipLaunchKernelGGL(k_offblock_sum, dim3(cdiv(n, 256)), dim3(256), 0, s, n_pairs, pair_b, run_start, sorted, body,
                       H432, off, offT, n_act, err);
}

__global__ void k_csr_keys(Cnt cpairs, const int* __restrict__ pair_b, uint64_t* keys, int* vals) {
    int u = blockIdx.x * blockDim.x + threadIdx.x;
    if (u >= cnt_of(cpairs)) return;
    uint64_t a = (uint64_t)pair_b[2 * u], b = (uint64_t)pair_b[2 * u + 1];
    keys[2 * u] = (a << 32) | b;
    vals[2 * u] = u << 1;
    keys[2 * u + 1] = (b << 32) | a;
    vals[2 * u + 1] = (u << 1) | 1;
}

void launch_csr_keys(hipStream_t s, Cnt n_pairs, const int* pair_b, uint64_t* keys, int* vals) {
    if (n_pairs.cap <= 0) return;
    hipLaunchKernelGGL(k_csr_keys, dim3(cdiv(n_pairs.cap, 256)), dim3(256), 0, s, n_pairs, pair_b, keys, vals);
}

__device__ __forceinline__ int lower_bound64(const uint64_t* a, int n, uint64_t x) {
    int lo = 0, hi = n;
    while (lo < hi) {
        int mid = (lo + hi) >> 1;
        if (a[mid] < x) lo = mid + 1; else hi = mid;
    }
    return lo;
}

// row_ptr[b] (b = 0..nb) = the first key with row (bits `shift`+) >= b, i.e. lower_bound(b << shift)
// of the sorted keys, one thread per row (log2 n dependent loads; filling the rows from the run
// heads instead leaves one thread a serial loop over every empty row of a gap).  Rows: bodies of
// the CSR entry and incidence keys (shift 32), scenes of the candidate keys (shift 48).
__global__ void k_row_ptr(int nb, Cnt cent, const uint64_t* __restrict__ k, int shift, int* row_ptr) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b > nb) return;
    row_ptr[b] = lower_bound64(k, cnt_of(cent), (uint64_t)b << shift);
}

void launch_row_ptr(hipStream_t s, int nb, Cnt n_keys, const uint64_t* sorted, int* row_ptr, int shift) {
    hipLaunchKernelGGL(k_row_ptr, dim3(cdiv(nb + 1, 256)), dim3(256), 0, s, nb, n_keys, sorted, shift, row_ptr);
}

// ---------------------------------------------------------------------------------------------
// per-body assembly: inertia_grad_hess (abd.py:108-149) with a 4x4 mass matrix, PSD-projected
// orthogonality Hessian, the body's contact incidences, the block-Jacobi inverse and the body
// energy (energy_inertia, abd.py:53-64).  One wave per body.
// ---------------------------------------------------------------------------------------------
__global__ __launch_bounds__(64) void k_body_assemble(Geo G, AbdParams P, const double* __restrict__ q,
                                                       const double* __restrict__ qtilde,
                                                       const int* __restrict__ inc_ptr,
                                                       const uint64_t* __restrict__ inc,
                                                       const double* __restrict__ g24,
                                                       const double* __restrict__ H432,
                                                       const int* __restrict__ newton_active,
                                                       double* __restrict__ grad, double* __restrict__ diag,
                                                       double* __restrict__ pinv, double* __restrict__ e_body,
                                                       Cnt cact) {
    __shared__ double A[81], V[81], H[144], W[12 * 24], gq[12], fac[12], qb[12], dq[12];
    const int n_act = cnt_of(cact);
    const int b = blockIdx.x;
    const int lane = threadIdx.x;
    const double pd = P.pinned_diag;
    if (G.pinned[b]) {
        for (int e = lane; e < 144; e += 64) {
            int r = e / 12, c = e % 12;
            diag[(size_t)b * 144 + e] = (r == c) ? pd : 0.0;
            pinv[(size_t)b * 144 + e] = (r == c) ? 1.0 / pd : 0.0;
        }
        if (lane < 12) grad[12 * b + lane] = 0.0;
        if (lane == 0) e_body[b] = 0.0;
        return;
    }
    if (!newton_active[G.bscene[b]]) {
        // a scene whose Newton loop has stopped needs only a zero gradient from this assembly: its
        // PCG ends at once on b = 0 and its line search takes no step, so its blocks, inverse and
        // energy keep the values of its last iteration (in a batch most scenes stop early)
        if (lane < 12) grad[12 * b + lane] = 0.0;
        return;
    }
    const double dt2 = P.dt * P.dt;
    const double kap = P.kappa;
    const double* M = G.mass4 + 16 * b;
    if (lane < 12) {
        qb[lane] = q[12 * b + lane];
        dq[lane] = qb[lane] - qtilde[12 * b + lane];
    }
    __syncthreads();
    const double* a = qb + 3;
    // ortho gradient / Hessian (ortho.py:19-41)
    if (lane < 12) {
        int i = lane / 3, r = lane % 3;
        double g = 0.0;
        for (int j = 0; j < 4; ++j) g += M[4 * i + j] * dq[3 * j + r];
        if (i > 0) {
            int ii = i - 1;
            double s = -a[3 * ii + r];
            for (int j = 0; j < 3; ++j)
                s += (a[3 * ii] * a[3 * j] + a[3 * ii + 1] * a[3 * j + 1] + a[3 * ii + 2] * a[3 * j + 2]) * a[3 * j + r];
            g += dt2 * 4.0 * kap * s;
        }
        gq[lane] = g;
    }
    for (int e = lane; e < 81; e += 64) {
        int R = e / 9, C = e % 9;
        int i = R / 3, r = R % 3, j = C / 3, c = C % 3;
        double Gij = a[3 * i] * a[3 * j] + a[3 * i + 1] * a[3 * j + 1] + a[3 * i + 2] * a[3 * j + 2];
        double h;
        if (i == j) {
            h = a[3 * i + r] * a[3 * i + c] + (r == c ? Gij - 1.0 : 0.0);
            for (int k = 0; k < 3; ++k) h += a[3 * k + r] * a[3 * k + c];
        } else {
            h = a[3 * j + r] * a[3 * i + c] + (r == c ? Gij : 0.0);
        }
        A[e] = 4.0 * kap * h;
    }
    // PSD projection: the Hessian minus its negative eigenpairs, which come in closed form from the
    // singular values of the body's 3 x 3 block (ortho_negative_modes; mode m in V[9 m .. 9 m + 8]);
    // near singular blocks take the matrix-sign projection (after a Cholesky test)
    __shared__ double lam_neg[9];
    __shared__ int n_neg;
    if (lane == 0) n_neg = ortho_negative_modes(a, 4.0 * kap, V, lam_neg);
    __syncthreads();
    if (n_neg < 0) {
        if (!wave_is_psd(A, V, 9, lane)) wave_psd_project(A, 9, lane);
    } else if (n_neg > 0) {
        for (int e = lane; e < 81; e += 64) {
            const int R = e / 9, C = e % 9;
            double h = A[e];
            for (int m = 0; m < n_neg; ++m) h -= lam_neg[m] * (V[9 * m + R] * V[9 * m + C]);  // symmetric bits
            A[e] = h;
        }
    }
    __syncthreads();
    for (int e = lane; e < 144; e += 64) {
        int R = e / 12, C = e % 12;
        int i = R / 3, r = R % 3, j = C / 3, c = C % 3;
        double h = (r == c) ? M[4 * i + j] : 0.0;
        if (i > 0 && j > 0) h += dt2 * A[(3 * (i - 1) + r) * 9 + 3 * (j - 1) + c];
        H[e] = h;
    }
    __syncthreads();
    // contact incidences of this body (inc_ptr: the body's range of the sorted incidence keys), in
    // sorted (contact) order: each lane sums its entries e = lane, lane + 64, lane + 128 of H (and of
    // the gradient) in registers, in incidence order, with the loads of four incidences in flight
    const int lo = inc_ptr[b], hi = inc_ptr[b + 1];
    double h0 = H[lane], h1 = H[lane + 64], h2 = lane < 16 ? H[lane + 128] : 0.0;
    double gl = lane < 12 ? gq[lane] : 0.0;
    bool bad = false;
    for (int m0 = lo; m0 < hi && !bad; m0 += 4) {
        int cc[4], rl[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const uint32_t w = (uint32_t)inc[min(m0 + u, hi - 1)];
            cc[u] = opaque((int)(w >> 1));
            rl[u] = opaque((int)(w & 1));
        }
        double v0[4], v1[4], v2[4], gv[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const bool ok = cc[u] >= 0 && cc[u] < n_act;  // (out-of-range contacts are not read)
            const double* Hc = H432 + (size_t)(ok ? cc[u] : 0) * 432 + rl[u] * 144;
            v0[u] = ok ? Hc[lane] : 0.0;
            v1[u] = ok ? Hc[lane + 64] : 0.0;
            v2[u] = (ok && lane < 16) ? Hc[lane + 128] : 0.0;
            gv[u] = (ok && lane < 12) ? g24[(size_t)(ok ? cc[u] : 0) * 24 + rl[u] * 12 + lane] : 0.0;
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            if (m0 + u >= hi || bad) break;
            if (cc[u] < 0 || cc[u] >= n_act) {
                if (lane == 0) atomicOr(G.err, ERR_INCIDENCE);
                bad = true;
                break;
            }
            h0 += v0[u];
            h1 += v1[u];
            h2 += v2[u];
            gl += gv[u];
        }
    }
    __syncthreads();  // every lane has read H
    H[lane] = h0;
    H[lane + 64] = h1;
    if (lane < 16) H[lane + 128] = h2;
    if (lane < 12) gq[lane] = gl;
    __syncthreads();
    const bool active = newton_active[G.bscene[b]] != 0;
    if (lane < 12) grad[12 * b + lane] = active ? gq[lane] : 0.0;
    for (int e = lane; e < 144; e += 64) diag[(size_t)b * 144 + e] = H[e];
    if (lane == 0) {
        double ein = 0.0;
        for (int i = 0; i < 4; ++i)
            for (int j = 0; j < 4; ++j)
                ein += M[4 * i + j] * (dq[3 * i] * dq[3 * j] + dq[3 * i + 1] * dq[3 * j + 1] + dq[3 * i + 2] * dq[3 * j + 2]);
        e_body[b] = 0.5 * ein + dt2 * ortho_energy(qb, kap);
    }
    // block-Jacobi inverse: Gauss-Jordan on [H | I] (SPD, no pivoting)
    for (int e = lane; e < 288; e += 64) {
        int r = e / 24, c = e % 24;
        W[e] = (c < 12) ? H[r * 12 + c] : ((c - 12 == r) ? 1.0 : 0.0);
    }
    __syncthreads();
    for (int k = 0; k < 12; ++k) {
        double piv = W[k * 24 + k];
        __syncthreads();
        if (lane < 24) W[k * 24 + lane] /= piv;
        __syncthreads();
        if (lane < 12) fac[lane] = W[lane * 24 + k];
        __syncthreads();
        for (int e = lane; e < 288; e += 64) {
            int r = e / 24, c = e % 24;
            if (r != k) W[e] -= fac[r] * W[k * 24 + c];
        }
        __syncthreads();
    }
    for (int e = lane; e < 144; e += 64) pinv[(size_t)b * 144 + e] = W[(e / 12) * 24 + 12 + e % 12];
}

void launch_body_assemble(hipStream_t s, const Geo& G, const AbdParams& P, const double* q, const double* qtilde,
                          const int* inc_ptr, const uint64_t* inc_sorted, const int2* body, const double* g24,
                          const double* H432, const int* newton_active, double* grad, double* diag, double* pinv,
                          double* e_body, Cnt n_act) {
    hipLaunchKernelGGL(k_body_assemble, dim3(G.nb), dim3(64), 0, s, G, P, q, qtilde, inc_ptr, inc_sorted, g24, H432,
                       newton_active, grad, diag, pinv, e_body, n_act);
}

// per-scene sum of body energies and candidate energies (fixed tree order)
__device__ void ls_init_scene(int s, const AbdParams& P, const int* active, const SceneCtl& c);
__device__ void ls_decide_scene(int s, const AbdParams& P, const SceneCtl& c, int floor_accept, int* pending);

// per-scene energy sum; ls (optional, after the sum, by the scene's thread 0): 1 = the line search's
// start, 2 = the decisions of `trials` consecutive trials (the k-th at 2^-k alpha, its body energies
// at e_body + k nb and candidate energies at e_cand + k estride), taken in order as one trial per
// launch would take them: a scene decided before trial k (here or in an earlier launch) is left
// alone, so its E1 stays the energy of its decisive trial.  `pending` reports the scenes still
// undecided after the last trial.
__global__ __launch_bounds__(256) void k_scene_sum(Geo G, const double* __restrict__ e_body,
                                                    const double* __restrict__ e_cand, const int* __restrict__ lo,
                                                    const int* __restrict__ hi, double* __restrict__ out, int ls,
                                                    AbdParams P, const int* active, SceneCtl c, int floor_accept,
                                                    int* pending, int trials, size_t estride) {
    const int s = blockIdx.x;
    __shared__ double sh[256];
    __shared__ int more;  // halvings decided before the last trial (reported by the last one only)
    const int nk = (ls == 2) ? trials : 1;
    for (int k = 0; k < nk; ++k) {
        // c.ls_done[s] is written by thread 0 before the barrier that ends the previous trial: the
        // exit is uniform
        if (ls == 2 && c.ls_done[s]) return;
        const double* eb = e_body + (size_t)k * G.nb;
        const double* ec = e_cand ? e_cand + (size_t)k * estride : nullptr;
        double acc = 0.0;
        for (int b = G.scene_b0[s] + threadIdx.x; b < G.scene_b0[s + 1]; b += 256) acc += eb[b];
        if (ec)
            for (int i = lo[s] + threadIdx.x; i < hi[s]; i += 256) acc += ec[i];
        sh[threadIdx.x] = acc;
        __syncthreads();
        for (int o = 128; o > 0; o >>= 1) {
            if (threadIdx.x < o) sh[threadIdx.x] += sh[threadIdx.x + o];
            __syncthreads();
        }
        if (threadIdx.x == 0) {
            out[s] = sh[0];
            if (ls == 1) ls_init_scene(s, P, active, c);
            if (ls == 2) ls_decide_scene(s, P, c, floor_accept, k + 1 == nk ? pending : &more);
        }
        __syncthreads();
    }
}

void launch_scene_sum(hipStream_t s, const Geo& G, const double* e_body, const double* e_cand, const int* lo,
                      const int* hi, double* out) {
    hipLaunchKernelGGL(k_scene_sum, dim3(G.ns), dim3(256), 0, s, G, e_body, e_cand, lo, hi, out, 0, AbdParams{},
                       (const int*)nullptr, SceneCtl{}, 0, (int*)nullptr, 1, (size_t)0);
}

void launch_scene_sum_ls(hipStream_t s, const Geo& G, const double* e_body, const double* e_cand, const int* lo,
                         const int* hi, double* out, int ls, const AbdParams& P, const int* active, SceneCtl c,
                         int floor_accept, int* pending, int trials, size_t estride) {
    hipLaunchKernelGGL(k_scene_sum, dim3(G.ns), dim3(256), 0, s, G, e_body, e_cand, lo, hi, out, ls, P, active, c,
                       floor_accept, pending, trials, estride);
}

// ---------------------------------------------------------------------------------------------
// Block-Jacobi PCG (replaces warp.optim.linear.bicgstab, abd.py:636), tiled: one wave per tile
// of at most 5 bodies (60 rows) of one scene.  Partial dot products per tile; each scene's
// scalars are the sum of its tiles' partials, reduced by every wave in the same fixed order.
// part layout: [quantity(3: pAp, rz, rr)][parity(2)][tile]
// ---------------------------------------------------------------------------------------------
#define TILE_BODIES 5

__device__ __forceinline__ double scene_partial(const double* part, int ntiles, int qty, int par, int t0, int t1,
                                                int lane) {
    const double* p = part + (size_t)(qty * 2 + par) * ntiles;
    double acc = 0.0;
    for (int t = t0 + lane; t < t1; t += 64) acc += p[t];
    return wave_sum(acc);
}

// per-scene sum of one partial (which = qty*2+par); optional sqrt (residual norms)
__global__ __launch_bounds__(64) void k_pcg_tile_sum(int ns, const int* __restrict__ st0, const double* __restrict__ part,
                                                      int stride, int which, double* out, int do_sqrt) {
    int s = blockIdx.x;
    double v = 0.0;
    for (int t = st0[s] + threadIdx.x; t < st0[s + 1]; t += 64) v += part[(size_t)which * stride + t];
    v = wave_sum(v);
    if (threadIdx.x == 0) out[s] = do_sqrt ? sqrt(v) : v;
}

void launch_pcg_tile_sum(hipStream_t s, int ns, const int* scene_tile0, const double* part, int stride, int which,
                         double* out, int do_sqrt) {
    hipLaunchKernelGGL(k_pcg_tile_sum, dim3(ns), dim3(64), 0, s, ns, scene_tile0, part, stride, which, out, do_sqrt);
}

// A scene is done once its residual after `it` iterations (parity it&1) is small enough; the
// flag makes that sticky, because later launches see the other parity's (older) partials.
__device__ __forceinline__ bool pcg_done(const double* part, int ntiles, int par, int t0, int t1, int lane, double bn,
                                         double tol, const int* done_flag) {
    if (bn == 0.0 || *done_flag) return true;
    double rr = scene_partial(part, ntiles, 2, par, t0, t1, lane);
    return sqrt(rr) <= tol * bn;
}

// Block-Jacobi PCG in the Chronopoulos-Gear arrangement (one global reduction point, so one
// launch, per iteration; algebraically the standard PCG of the oracle):
//   stage 0:   r0 = b, u0 = P r0, w0 = H u0;                     partials g0=(r,u), d0=(w,u), rho0=(r,r)
//   stage i+1: beta = g_i/g_{i-1}, alpha = g_i/(d_i - beta g_i/alpha_{i-1})  (i = 0: beta 0, alpha g/d)
//              p = u + beta p, s = w + beta s, x += alpha p, r -= alpha s,
//              u = P r, w = H u;                                  partials g, d, rho of the new r, u, w
// H u needs u of neighbouring bodies: each recomputes it from the neighbour's previous r, w, s
// (ping-pong buffers, slot k&1 written at stage k, slot (k-1)&1 read), 12 lanes of the body
// computing one row each and sharing it through LDS.  Partials: qty 0 = d, 1 = g, 2 = rho, at
// parity k&1.  Per-scene scalars sc[s][par] = g_i, sc[ns + ...] = alpha_i, written by the scene's
// first tile.  A scene stops (sticky flag) once sqrt(rho_i) <= tol * ||b||.

// row of the neighbour's new preconditioned residual: (P_j (r_j - alpha (w_j + beta s_j)))_row
template <typename T>
__device__ __forceinline__ T pcg_r_next(T r, T w, T s, T alpha, T beta) {
    return r - alpha * (w + beta * s);
}

template <typename T>
__device__ __forceinline__ T aload(const T* p) {
    return __hip_atomic_load(const_cast<T*>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <typename T>
__device__ __forceinline__ void astore(T* p, T v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// LDS exchange among the lanes of one wave: order the wave's LDS accesses (no block barrier, the
// waves of a workgroup run different tiles)
__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// plain or agent-scope (coherent across XCDs within one launch) access to the PCG's shared state
template <bool A, typename V>
__device__ __forceinline__ V ldp(const V* p) {
    if constexpr (A) return aload(p);
    else return *p;
}
template <bool A, typename V>
__device__ __forceinline__ void stp(V* p, V v) {
    if constexpr (A) astore(p, v);
    else *p = v;
}
template <bool A>
__device__ __forceinline__ double scene_partial_p(const double* part, int ntiles, int qty, int par, int t0, int t1,
                                                  int lane) {
    const double* p = part + (size_t)(qty * 2 + par) * ntiles;
    double acc = 0.0;
    // four rounds of 64 tiles per batch, every load of a batch in flight at once, summed in tile order
    for (int tb = t0 + lane; tb < t1; tb += 256) {
        double v[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) v[u] = ldp<A>(p + min(tb + 64 * u, t1 - 1));
#pragma unroll
        for (int u = 0; u < 4; ++u)
            if (tb + 64 * u < t1) acc += v[u];
    }
    return wave_sum(acc);
}

// Stage k of the PCG for tile t by one wave (lane 0-63; vec: the wave's LDS rows).  A = false:
// one launch per stage (k_pcg_cg; the scene's ||b|| in B.bnorm, computed after stage 0); A = true:
// the stages of one persistent launch (k_pcg_multi; shared state through agent-scope atomics,
// ||b|| taken at stage 1 and kept in B.bnorm_out, every scene found done at stage k counted once in
// done_cnt[(k - 1) % 3]).
template <typename T, bool A>
__device__ void pcg_stage(int k, int t, int lane, T (*vec)[12], const PcgBufT<T>& B, int ntiles, int ns,
                          const int* __restrict__ tb0, const int* __restrict__ tb1, const int* __restrict__ tscene,
                          const int* __restrict__ st0, double tol, int nb, int n_pairs) {
    const int s = tscene[t];
    const int t0 = st0[s], t1 = st0[s + 1];
    const bool init = (k == 0);
    const int rd = (k + 1) & 1, wr = k & 1;
    double alpha = 0.0, beta = 0.0;
    if (!init) {
        const int par = (k - 1) & 1;
        const int i = k - 1;
        double bn;
        if constexpr (A) {
            if (k == 1) {
                bn = B.bref ? B.bref[s] : sqrt(scene_partial_p<A>(B.part, ntiles, 2, 0, t0, t1, lane));
                if (t == t0 && lane == 0) {
                    astore(B.bnorm_out + s, bn);
                    astore(B.residual_out + s, bn);
                }
            } else {
                bn = aload(B.bnorm_out + s);
            }
        } else {
            bn = B.bnorm[s];
        }
        // sticky: later stages see the other parity's (older) partials of a done scene
        const bool done = bn == 0.0 || ldp<A>(B.cg_done + s) != 0 ||
                          sqrt(scene_partial_p<A>(B.part, ntiles, 2, par, t0, t1, lane)) <= tol * bn;
        if (done) {
            if (t == t0 && lane == 0) {
                stp<A>(B.cg_done + s, 1);
                if constexpr (A) atomicAdd(B.done_cnt + (i % 3), 1);
            }
            return;
        }
        const double g = scene_partial_p<A>(B.part, ntiles, 1, par, t0, t1, lane);
        const double d = scene_partial_p<A>(B.part, ntiles, 0, par, t0, t1, lane);
        if (i == 0) {
            alpha = g / d;
        } else {
            const double gp = ldp<A>(B.sc + 2 * s + ((i - 1) & 1)), ap = ldp<A>(B.sc + 2 * ns + 2 * s + ((i - 1) & 1));
            beta = g / gp;
            alpha = g / (d - beta * g / ap);
        }
        if (t == t0 && lane == 0) {
            stp<A>(B.sc + 2 * s + (i & 1), g);
            stp<A>(B.sc + 2 * ns + 2 * s + (i & 1), alpha);
            B.cg_iters[s] = i + 1;
        }
    }
    const int b0 = tb0[t], nbt = tb1[t] - b0;
    const int lb = lane / 12, row = lane % 12;
    const bool on = lb < nbt;
    const int bb = b0 + (on ? lb : 0);
    const int idx = 12 * bb + row;
    const T al = (T)alpha, be = (T)beta;
    T rn = 0;
    if (on) {
        // x, p, u rows belong to this tile alone: plain accesses
        if (init) {
            rn = B.b[idx];
            B.x[idx] = 0;
            B.p[idx] = 0;
            stp<A>(B.sv[wr] + idx, T(0));
        } else {
            const T wo = ldp<A>(B.w[rd] + idx), so = ldp<A>(B.sv[rd] + idx), ro = ldp<A>(B.r[rd] + idx);
            const T pv = B.u[idx] + be * B.p[idx];
            B.p[idx] = pv;
            stp<A>(B.sv[wr] + idx, wo + be * so);
            B.x[idx] += al * pv;
            rn = pcg_r_next(ro, wo, so, al, be);
        }
        stp<A>(B.r[wr] + idx, rn);
        vec[lb][row] = rn;
    }
    wave_lds_sync();
    T un = 0;
    if (on) {
        const T* Pb = B.pinv + (size_t)bb * 144 + row * 12;
#pragma unroll
        for (int c = 0; c < 12; ++c) un += Pb[c] * vec[lb][c];
        B.u[idx] = un;
    }
    wave_lds_sync();
    if (on) vec[lb][row] = un;
    wave_lds_sync();
    T wv = 0;
    if (on) {
        const T* D = B.diag + (size_t)bb * 144 + row * 12;
#pragma unroll
        for (int c = 0; c < 12; ++c) wv += D[c] * vec[lb][c];
    }
    const int e0 = on ? B.row_ptr[bb] : 0, e1 = on ? B.row_ptr[bb + 1] : 0;
    int mx = e1 - e0;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) mx = max(mx, __shfl_xor(mx, o, 64));
    wave_lds_sync();
    for (int q = 0; q < mx; ++q) {
        const bool has = on && q < e1 - e0;
        const T* Bk = nullptr;
        T uc = 0;
        if (has) {
            const int e = e0 + q;
            const int col = opaque((int)(B.ent_key[e] & 0xFFFFFFFFu));
            const int v = B.ent_val[e];
            if (col < 0 || col >= nb || v < 0 || (v >> 1) >= n_pairs) {
                atomicOr(B.err, ERR_CSR);
            } else {
                Bk = (opaque(v & 1) ? B.offT : B.off) + (size_t)opaque(v >> 1) * 144 + row * 12;
                const T* Pc = B.pinv + (size_t)col * 144 + row * 12;
#pragma unroll
                for (int c = 0; c < 12; ++c) {
                    const int j = 12 * col + c;
                    const T rj = init ? B.b[j]
                                      : pcg_r_next(ldp<A>(B.r[rd] + j), ldp<A>(B.w[rd] + j), ldp<A>(B.sv[rd] + j), al, be);
                    uc += Pc[c] * rj;
                }
            }
        }
        vec[lb][row] = uc;
        wave_lds_sync();
        if (Bk) {
#pragma unroll
            for (int c = 0; c < 12; ++c) wv += Bk[c] * vec[lb][c];
        }
        wave_lds_sync();
    }
    if (on) stp<A>(B.w[wr] + idx, wv);
    const double g = wave_sum(on ? (double)rn * un : 0.0), d = wave_sum(on ? (double)wv * un : 0.0),
                 rr = wave_sum(on ? (double)rn * rn : 0.0);
    if (lane == 0) {
        stp<A>(B.part + (size_t)(0 * 2 + wr) * ntiles + t, d);
        stp<A>(B.part + (size_t)(1 * 2 + wr) * ntiles + t, g);
        stp<A>(B.part + (size_t)(2 * 2 + wr) * ntiles + t, rr);
    }
}

template <typename T>
__global__ __launch_bounds__(64) void k_pcg_cg(int k, int ntiles, int ns, const int* __restrict__ tb0,
                                                const int* __restrict__ tb1, const int* __restrict__ tscene,
                                                const int* __restrict__ st0, PcgBufT<T> B, double tol, int nb,
                                                int n_pairs) {
    __shared__ T vec[TILE_BODIES + 1][12];  // lanes 60-63 (lb = 5) only touch the last row
    pcg_stage<T, false>(k, blockIdx.x, threadIdx.x, vec, B, ntiles, ns, tb0, tb1, tscene, st0, tol, nb, n_pairs);
}

template <typename T>
void launch_pcg_cg(hipStream_t s, int k, int ntiles, int ns, const int* tile_b0, const int* tile_b1,
                   const int* tile_scene, const int* scene_tile0, const PcgBufT<T>& B, double tol, int nb, int n_pairs) {
    hipLaunchKernelGGL(k_pcg_cg<T>, dim3(ntiles), dim3(64), 0, s, k, ntiles, ns, tile_b0, tile_b1, tile_scene,
                       scene_tile0, B, tol, nb, n_pairs);
}
template void launch_pcg_cg<double>(hipStream_t, int, int, int, const int*, const int*, const int*, const int*,
                                    const PcgBufT<double>&, double, int, int);
template void launch_pcg_cg<float>(hipStream_t, int, int, int, const int*, const int*, const int*, const int*,
                                   const PcgBufT<float>&, double, int, int);

// ---------------------------------------------------------------------------------------------
// The same PCG as one cooperative launch per solve.  Each workgroup keeps its tile's rows (x, p,
// u, r, w, s, its rows of P and of the diagonal blocks) in registers across iterations and
// publishes r, w, s and its partials with agent-scope atomic stores; neighbours and partial sums
// are read with agent-scope atomic loads, so the exchange is coherent across XCDs.  Iterations are
// separated by a grid barrier on a monotonic counter (the k-th barrier waits for k * grid
// arrivals).  Every wait is bounded: a barrier that does not fill within the spin limit sets the
// failure flag and every workgroup leaves, the host then reruns the solve with one launch per
// iteration.  A scene is done (sticky, decided identically by all its tiles) once sqrt(rho) <=
// tol * ||b||; the first tile of every done scene adds one to done_cnt[i % 3] at iteration i, and
// the loop ends when that count reaches ns.  Slot (i + 1) % 3 is cleared by workgroup 0 during
// iteration i, between its last read (before the previous barrier) and its next increment.
// ---------------------------------------------------------------------------------------------
#define PCG_SPIN_LIMIT (1 << 22)


// grid barrier #k (k = 1, 2, ...); false if it timed out or another workgroup failed
__device__ bool grid_barrier(unsigned* bar, int* fail, unsigned k, unsigned nblocks) {
    __shared__ int ok_sh;
    __syncthreads();
    if (threadIdx.x == 0) {
        int ok = 1;
        // every value another workgroup reads after this barrier went out through agent-scope
        // atomic stores; wait for them, then arrive.  Readers use agent-scope atomic loads, so no
        // cache maintenance is needed on either side.
        __builtin_amdgcn_s_waitcnt(0);
        __hip_atomic_fetch_add(bar, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        const unsigned target = k * nblocks;
        int spins = 0;
        // one agent-scope load per spin, ~1500 cycles apart (s_sleep 24): polling the count more
        // often delays the arrivals on it (8.66 us per chain-net PCG iteration at s_sleep 1, 8.12
        // at 24, 8.27 at 32; issuing four polls at once was slower still).  The failure flag (set
        // by a workgroup whose wait ran out) is read every 32 spins, so a failed solve still ends
        // every workgroup within bounded time, while a barrier that fills costs no extra round
        // trip.  A workgroup that passes a barrier another one gave up on stops at the next
        // barrier, which cannot fill.
        while (__hip_atomic_load(bar, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < target) {
            if ((++spins & 31) == 0) {
                if (__hip_atomic_load(fail, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) { ok = 0; break; }
                if (spins > PCG_SPIN_LIMIT) {
                    __hip_atomic_store(fail, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    ok = 0;
                    break;
                }
            }
            __builtin_amdgcn_s_sleep(24);
        }
        ok_sh = ok;
    }
    __syncthreads();
    return ok_sh != 0;
}

#define PCG_WAVES PCG_TILES_PER_BLOCK
#ifndef PCG_NBR_CACHE
#define PCG_NBR_CACHE 8  // neighbours per chunk of the persistent PCG
#endif

// column of CSR entry e, or -1 (and the error bit) if the entry is out of range
template <typename T>
__device__ __forceinline__ int pcg_entry(const PcgBufT<T>& B, int e, int nb, int n_pairs) {
    const int col = opaque((int)(B.ent_key[e] & 0xFFFFFFFFu));
    const int v = B.ent_val[e];
    if (col < 0 || col >= nb || v < 0 || (v >> 1) >= n_pairs) {
        atomicOr(B.err, ERR_CSR);
        return -1;
    }
    return col;
}

// M_e = B_e P_col(e) for every CSR entry e (B_e = off or off^T of the pair), FP64 products rounded
// to T; one thread per element.  Invalid entries get M = 0 and the error bit.
template <typename T>
__global__ void k_pcg_mblocks(Cnt cent, const uint64_t* __restrict__ ent_key, const int* __restrict__ ent_val,
                              const double* __restrict__ off, const double* __restrict__ offT,
                              const double* __restrict__ pinv, int nb, int n_pairs, T* __restrict__ mq, int* err) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (long)cnt_of(cent) * 144) return;
    const int e = (int)(i / 144), rc = (int)(i % 144), r = rc / 12, c = rc % 12;
    const int col = opaque((int)(ent_key[e] & 0xFFFFFFFFu));
    const int v = ent_val[e];
    double m = 0.0;
    if (col < 0 || col >= nb || v < 0 || (v >> 1) >= n_pairs) {
        if (rc == 0) atomicOr(err, ERR_CSR);
    } else {
        const double* Bk = (opaque(v & 1) ? offT : off) + (size_t)opaque(v >> 1) * 144 + r * 12;
        const double* Pc = pinv + (size_t)col * 144 + c;
#pragma unroll
        for (int k = 0; k < 12; ++k) m += Bk[k] * Pc[12 * k];
    }
    mq[i] = (T)m;
}

template <typename T>
void launch_pcg_mblocks(hipStream_t s, Cnt n_ent, const uint64_t* ent_key, const int* ent_val, const double* off,
                        const double* offT, const double* pinv, int nb, int n_pairs, T* mq, int* err) {
    if (n_ent.cap <= 0) return;
    hipLaunchKernelGGL(k_pcg_mblocks<T>, dim3(cdiv((long)n_ent.cap * 144, 256)), dim3(256), 0, s, n_ent, ent_key,
                       ent_val, off, offT, pinv, nb, n_pairs, mq, err);
}
template void launch_pcg_mblocks<double>(hipStream_t, Cnt, const uint64_t*, const int*, const double*, const double*,
                                         const double*, int, int, double*, int*);
template void launch_pcg_mblocks<float>(hipStream_t, Cnt, const uint64_t*, const int*, const double*, const double*,
                                        const double*, int, int, float*, int*);


template <typename T>
__global__ __launch_bounds__(64 * PCG_WAVES) void k_pcg_persist(int ntiles, int ns, const int* __restrict__ tb0,
                                                                 const int* __restrict__ tb1,
                                                                 const int* __restrict__ tscene,
                                                                 const int* __restrict__ st0, PcgBufT<T> B,
                                                                 double tol, int max_iter, int nb, int n_pairs) {
    __shared__ T vec_all[PCG_WAVES][TILE_BODIES + 1][12];
    // CSR columns of each body of the tile in chunks of PCG_NBR_CACHE (the first chunk stays), and
    // the neighbours' new residual rows of the current chunk
    __shared__ int nbr_all[PCG_WAVES][TILE_BODIES + 1][PCG_NBR_CACHE];
    __shared__ T rb_all[PCG_WAVES][TILE_BODIES + 1][PCG_NBR_CACHE][12];
    const int wave = threadIdx.x / 64, lane = threadIdx.x % 64;
    if (B.ts && threadIdx.x == 0) atomicMin(B.ts, (unsigned long long)wall_clock64());
    const int t = blockIdx.x * PCG_WAVES + wave;
    const bool tile = t < ntiles;  // waves past the last tile only take part in the barriers
    T(*vec)[12] = vec_all[wave];
    int(*nbr)[PCG_NBR_CACHE] = nbr_all[wave];
    T(*rb)[PCG_NBR_CACHE][12] = rb_all[wave];
    const unsigned nblocks = gridDim.x;
    const int s = tile ? tscene[t] : 0;
    const int t0 = tile ? st0[s] : 0, t1 = tile ? st0[s + 1] : 0;
    const int b0 = tile ? tb0[t] : 0, nbt = tile ? tb1[t] - b0 : 0;
    const int lb = lane / 12, row = lane % 12;
    const bool on = lb < nbt;
    const int bb = b0 + (on ? lb : 0);
    const int idx = 12 * bb + row;
    const bool lead = tile && (t == t0 && lane == 0);
    // static rows of this lane: P and the diagonal block
    T Prow[12], Drow[12];
#pragma unroll
    for (int c = 0; c < 12; ++c) {
        Prow[c] = on ? B.pinv[(size_t)bb * 144 + row * 12 + c] : T(0);
        Drow[c] = on ? B.diag[(size_t)bb * 144 + row * 12 + c] : T(0);
    }
    const int e0 = on ? B.row_ptr[bb] : 0, e1 = on ? B.row_ptr[bb + 1] : 0;
    int mx = e1 - e0;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) mx = max(mx, __shfl_xor(mx, o, 64));
    // Off-diagonal part of w = H u = H P r: sum over entries e of M_e r_col, with M_e = B_e P_col
    // (k_pcg_mblocks, FP64 products stored in T) fixed during the solve.
    if (on && row == 0) {
        for (int q = 0; q < min(e1 - e0, PCG_NBR_CACHE); ++q) {
            const int col = opaque((int)(B.ent_key[e0 + q] & 0xFFFFFFFFu));
            nbr[lb][q] = (col >= 0 && col < nb) ? col : -1;
        }
    }
    wave_lds_sync();

    T x = 0, p = 0, u = 0, r = 0, w = 0, sv = 0;
    double alpha = 0.0, beta = 0.0, g_prev = 1.0, a_prev = 1.0, bnorm = 0.0;
    bool done = !tile;
    unsigned nbar = 0;
    // columns of the chunk starting at q0 into nbr (row-0 lanes; -1 for an out-of-range column)
    auto load_cols = [&](int q0) {
        if (on && row == 0)
            for (int q = 0; q < PCG_NBR_CACHE && q0 + q < e1 - e0; ++q) {
                const int col = opaque((int)(B.ent_key[e0 + q0 + q] & 0xFFFFFFFFu));
                nbr[lb][q] = (col >= 0 && col < nb) ? col : -1;
            }
    };
    // neighbour rows of the chunk starting at q0: lane `row` loads row `row` of b (stage 0) or of
    // r, w, s of the previous stage, for the slots any body of the wave uses (wave-uniform count),
    // all loads in flight together
    T vr[PCG_NBR_CACHE], vw[PCG_NBR_CACHE], vs[PCG_NBR_CACHE];
    auto fetch = [&](int q0, int k, int rd) {
        const int nq = on ? min(e1 - e0 - q0, PCG_NBR_CACHE) : 0;
        const int nqw = min(mx - q0, PCG_NBR_CACHE);
#pragma unroll
        for (int q = 0; q < PCG_NBR_CACHE; ++q) {
            if (q >= nqw) break;
            int col = (q < nq) ? nbr[lb][q] : -1;
            col = col < 0 ? bb : col;  // in-range address for unused slots
            const int j = 12 * col + row;
            if (k == 0) {
                vr[q] = B.b[j];
            } else {
                vr[q] = aload(B.r[rd] + j);
                vw[q] = aload(B.w[rd] + j);
                vs[q] = aload(B.sv[rd] + j);
            }
        }
    };
    for (int k = 0; k <= max_iter; ++k) {
        const int wr = k & 1, rd = (k + 1) & 1;
        const int i = k - 1;  // iteration index of this stage (-1: initial stage)
        if (k > 0) {
            // slot (i + 1) % 3 was last read before the previous barrier and is next counted at i + 1
            if (blockIdx.x == 0 && threadIdx.x == 0)
                __hip_atomic_store(B.done_cnt + ((i + 1) % 3), 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
        // the first chunk's neighbour rows do not depend on this stage's scalars: their loads go out
        // together with the partial-sum loads below (one round trip for both)
        if (!done && mx > 0) fetch(0, k, rd);
        if (k > 0 && tile) {
            const int par = (k - 1) & 1;
            // the three partial sums of the scene in one pass over its tiles, four rounds of 64 tiles
            // per batch with every load of the batch in flight at once (one round trip per 256 tiles
            // instead of one per 64), summed in tile order as a one-round-at-a-time loop would
            double a_rho = 0.0, a_g = 0.0, a_d = 0.0;
            for (int tb = t0 + lane; tb < t1; tb += 256) {
                double pd[4], pg[4], pr[4];
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    const int tt = min(tb + 64 * u, t1 - 1);  // past the scene: a repeat of its last tile, dropped
                    pd[u] = aload(B.part + (size_t)(0 * 2 + par) * ntiles + tt);
                    pg[u] = aload(B.part + (size_t)(1 * 2 + par) * ntiles + tt);
                    pr[u] = aload(B.part + (size_t)(2 * 2 + par) * ntiles + tt);
                }
#pragma unroll
                for (int u = 0; u < 4; ++u)
                    if (tb + 64 * u < t1) {
                        a_d += pd[u];
                        a_g += pg[u];
                        a_rho += pr[u];
                    }
            }
            const double rho = wave_sum(a_rho), g = wave_sum(a_g), d = wave_sum(a_d);
            if (k == 1) {
                bnorm = B.bref ? B.bref[s] : sqrt(rho);
                if (lead) { B.bnorm_out[s] = bnorm; B.residual_out[s] = bnorm; }
            }
            if (!done && (bnorm == 0.0 || sqrt(rho) <= tol * bnorm)) done = true;
            if (done && lead) atomicAdd(B.done_cnt + (i % 3), 1);
            if (!done) {
                if (i == 0) {
                    beta = 0.0;
                    alpha = g / d;
                } else {
                    beta = g / g_prev;
                    alpha = g / (d - beta * g / a_prev);
                }
                g_prev = g;
                a_prev = alpha;
                if (lead) B.cg_iters[s] = i + 1;
            }
        }
        if (!done) {
            const T al = (T)alpha, be = (T)beta;
            T rn = 0, un = 0, wv = 0;
            if (k == 0) {
                rn = on ? B.b[idx] : T(0);
            } else {
                p = u + be * p;
                rn = pcg_r_next(r, w, sv, al, be);  // r_i - alpha (w_i + beta s_{i-1})
                sv = w + be * sv;                   // s_i
                x += al * p;
            }
            if (on) vec[lb][row] = rn;
            wave_lds_sync();
            if (on) {
#pragma unroll
                for (int c = 0; c < 12; ++c) un += Prow[c] * vec[lb][c];
            }
            wave_lds_sync();
            if (on) vec[lb][row] = un;
            wave_lds_sync();
            if (on) {
#pragma unroll
                for (int c = 0; c < 12; ++c) wv += Drow[c] * vec[lb][c];
            }
            wave_lds_sync();
            // neighbours in chunks of PCG_NBR_CACHE: all rows of a chunk are fetched in one round of
            // loads (lane `row` fetches row `row` of each neighbour's new residual), then every lane
            // multiplies its rows of M with them from LDS
            for (int q0 = 0; q0 < mx; q0 += PCG_NBR_CACHE) {
                if (q0 > 0) {  // later chunks: their columns, then their rows
                    wave_lds_sync();
                    load_cols(q0);
                    wave_lds_sync();
                    fetch(q0, k, rd);
                }
                const int nq = on ? min(e1 - e0 - q0, PCG_NBR_CACHE) : 0;
                const int nqw = min(mx - q0, PCG_NBR_CACHE);  // slots used by some body of the wave
#pragma unroll
                for (int q = 0; q < PCG_NBR_CACHE; ++q) {
                    if (q >= nqw) break;
                    const bool use = q < nq && nbr[lb][q] >= 0;
                    const T rj = (k == 0) ? vr[q] : pcg_r_next(vr[q], vw[q], vs[q], al, be);
                    if (on) rb[lb][q][row] = use ? rj : T(0);
                }
                wave_lds_sync();
                if constexpr (sizeof(T) == 8) {
                    // unused slots hold rb = 0 and read M of entry 0 (present whenever mx > 0):
                    // no branch, so the unrolled M-row loads go out together
#pragma unroll 4
                    for (int q = 0; q < nqw; ++q) {
                        const T* M = B.mq + (size_t)(q < nq ? e0 + q0 + q : 0) * 144 + row * 12;
#pragma unroll
                        for (int c = 0; c < 12; ++c) wv += M[c] * rb[lb][q][c];
                    }
                } else {
                    // FP32: u_col = P_col r_col first, then B_e u_col (rounding M = B_e P_col to FP32
                    // loses the stiff contact directions)
                    for (int q = 0; q < nqw; ++q) {
                        const bool use = q < nq && nbr[lb][q] >= 0;
                        T uc = 0, bk[12];
                        if (use) {
                            const int col = nbr[lb][q], v = B.ent_val[e0 + q0 + q];
                            const T* Bk = (opaque(v & 1) ? B.offT : B.off) + (size_t)opaque(v >> 1) * 144 + row * 12;
                            const T* Pc = B.pinv + (size_t)col * 144 + row * 12;
#pragma unroll
                            for (int c = 0; c < 12; ++c) bk[c] = Bk[c];
#pragma unroll
                            for (int c = 0; c < 12; ++c) uc += Pc[c] * rb[lb][q][c];
                        }
                        vec[lb][row] = uc;
                        wave_lds_sync();
                        if (use) {
#pragma unroll
                            for (int c = 0; c < 12; ++c) wv += bk[c] * vec[lb][c];
                        }
                        wave_lds_sync();
                    }
                }
                wave_lds_sync();
            }
            if (mx > PCG_NBR_CACHE) {  // the first chunk's columns again, for the next prefetch
                load_cols(0);
                wave_lds_sync();
            }
            // publish this stage's r, w and s for the neighbours of the next stage
            if (on) {
                astore(B.r[wr] + idx, rn);
                astore(B.w[wr] + idx, wv);
                astore(B.sv[wr] + idx, sv);
            }
            r = rn;
            u = un;
            w = wv;
            const double gp = wave_sum(on ? (double)rn * un : 0.0), dp = wave_sum(on ? (double)wv * un : 0.0),
                         rr = wave_sum(on ? (double)rn * rn : 0.0);
            if (lane == 0) {
                astore(B.part + (size_t)(0 * 2 + wr) * ntiles + t, dp);
                astore(B.part + (size_t)(1 * 2 + wr) * ntiles + t, gp);
                astore(B.part + (size_t)(2 * 2 + wr) * ntiles + t, rr);
            }
        }
        // each wave's stores are complete before the workgroup arrives
        __builtin_amdgcn_s_waitcnt(0);
        if (!grid_barrier(B.bar, B.fail, ++nbar, nblocks)) break;
        if (k > 0 && __hip_atomic_load(B.done_cnt + ((k - 1) % 3), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) >= ns)
            break;
    }
    if (on) B.x[idx] = x;
    if (B.ts) {
        __syncthreads();  // every wave of the workgroup has left the loop (its exits are uniform)
        if (threadIdx.x == 0) atomicMax(B.ts + 1, (unsigned long long)wall_clock64());
    }
}

// one persistent solve into the timing accumulator (one thread)
__global__ void k_pcg_account(unsigned long long* ts, double* acc, const int* cg_iters, int ns, const int* np,
                              double fixed, double per_pair) {
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    int K = 0;
    for (int s = 0; s < ns; ++s) K = max(K, cg_iters[s]);
    if (ts[1] > ts[0]) acc[0] += (double)(ts[1] - ts[0]);
    acc[1] += (double)K * (fixed + per_pair * (double)(*np));
    acc[2] += 1.0;
    ts[0] = ~0ull;
    ts[1] = 0ull;
}

void launch_pcg_account(hipStream_t s, unsigned long long* ts, double* acc, const int* cg_iters, int ns,
                        const int* np, double fixed, double per_pair) {
    hipLaunchKernelGGL(k_pcg_account, dim3(1), dim3(64), 0, s, ts, acc, cg_iters, ns, np, fixed, per_pair);
}


// A cooperative launch (all workgroups co-resident, which the grid barriers need).  Under stream
// capture it is recorded as a kernel node with the cooperative launch attribute.
#if !defined(ABD_EMULATION)
static hipError_t launch_coop(const void* f, int blocks, int threads, void** args, hipStream_t s) {
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    hipError_t e = hipStreamIsCapturing(s, &cs);
    if (e != hipSuccess) return e;
    if (cs == hipStreamCaptureStatusNone) return hipLaunchCooperativeKernel(f, dim3(blocks), dim3(threads), args, 0, s);
    hipLaunchAttribute attr = {};
    attr.id = hipLaunchAttributeCooperative;
    attr.val.cooperative = 1;
    hipLaunchConfig_t cfg = {};
    cfg.gridDim = dim3(blocks);
    cfg.blockDim = dim3(threads);
    cfg.dynamicSmemBytes = 0;
    cfg.stream = s;
    cfg.attrs = &attr;
    cfg.numAttrs = 1;
    return hipLaunchKernelExC(&cfg, f, args);
}
#endif

template <typename T>
hipError_t launch_pcg_persist(hipStream_t s, int ntiles, int ns, const int* tile_b0, const int* tile_b1,
                              const int* tile_scene, const int* scene_tile0, const PcgBufT<T>& B, double tol,
                              int max_iter, int nb, int n_pairs) {
    void* args[] = {(void*)&ntiles, (void*)&ns, (void*)&tile_b0, (void*)&tile_b1, (void*)&tile_scene,
                    (void*)&scene_tile0, (void*)&B, (void*)&tol, (void*)&max_iter, (void*)&nb, (void*)&n_pairs};
    const int blocks = (ntiles + PCG_WAVES - 1) / PCG_WAVES;
#if defined(ABD_EMULATION)
    // the CPU emulation runs one workgroup at a time: a single-workgroup grid is co-resident
    (void)args;
    if (blocks != 1) return hipErrorNotSupported;
    hipLaunchKernelGGL(k_pcg_persist<T>, dim3(1), dim3(64 * PCG_WAVES), 0, s, ntiles, ns, tile_b0, tile_b1, tile_scene,
                       scene_tile0, B, tol, max_iter, nb, n_pairs);
    return hipSuccess;
#else
    return launch_coop((const void*)k_pcg_persist<T>, blocks, 64 * PCG_WAVES, args, s);
#endif
}
// how many workgroups of k_pcg_persist<T> the device holds at once (a cooperative launch must fit)
template <typename T>
hipError_t pcg_persist_max_blocks(int* out) {
    int dev = 0, cus = 0, per_cu = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e == hipSuccess) e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    if (e == hipSuccess)
        e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, (const void*)k_pcg_persist<T>, 64 * PCG_WAVES, 0);
    *out = (e == hipSuccess) ? per_cu * cus : 0;
    return e;
}
template hipError_t pcg_persist_max_blocks<double>(int*);
template hipError_t pcg_persist_max_blocks<float>(int*);

template hipError_t launch_pcg_persist<double>(hipStream_t, int, int, const int*, const int*, const int*, const int*,
                                               const PcgBufT<double>&, double, int, int, int);

// ---------------------------------------------------------------------------------------------
// The PCG as one cooperative launch when the tiles outnumber the co-resident waves (the batched
// scenes): a grid of co-resident workgroups whose waves take tiles gw, gw + W, ... (W waves in the
// grid) at every stage, with every vector in memory (pcg_stage<T, true>: the arithmetic of the
// per-stage launches), the stages separated by the grid barrier and ended by the done counters of
// k_pcg_persist.
// ---------------------------------------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(64 * PCG_WAVES) void k_pcg_multi(int ntiles, int ns, const int* __restrict__ tb0,
                                                               const int* __restrict__ tb1,
                                                               const int* __restrict__ tscene,
                                                               const int* __restrict__ st0, PcgBufT<T> B, double tol,
                                                               int max_iter, int nb, int n_pairs) {
    __shared__ T vec_all[PCG_WAVES][TILE_BODIES + 1][12];
    const int wave = threadIdx.x / 64, lane = threadIdx.x % 64;
    const int gw = blockIdx.x * PCG_WAVES + wave, nw = gridDim.x * PCG_WAVES;
    if (B.ts && threadIdx.x == 0) atomicMin(B.ts, (unsigned long long)wall_clock64());
    unsigned nbar = 0;
    for (int k = 0; k <= max_iter; ++k) {
        // slot k % 3 is next counted at stage k + 1; it was last read before the previous barrier
        if (k > 0 && blockIdx.x == 0 && threadIdx.x == 0) astore(B.done_cnt + (k % 3), 0);
        for (int t = gw; t < ntiles; t += nw)
            pcg_stage<T, true>(k, t, lane, vec_all[wave], B, ntiles, ns, tb0, tb1, tscene, st0, tol, nb, n_pairs);
        __builtin_amdgcn_s_waitcnt(0);
        if (!grid_barrier(B.bar, B.fail, ++nbar, gridDim.x)) break;
        if (k > 0 && aload(B.done_cnt + ((k - 1) % 3)) >= ns) break;
    }
    if (B.ts) {
        __syncthreads();
        if (threadIdx.x == 0) atomicMax(B.ts + 1, (unsigned long long)wall_clock64());
    }
}

template <typename T>
hipError_t launch_pcg_multi(hipStream_t s, int blocks, int ntiles, int ns, const int* tile_b0, const int* tile_b1,
                            const int* tile_scene, const int* scene_tile0, const PcgBufT<T>& B, double tol,
                            int max_iter, int nb, int n_pairs) {
    void* args[] = {(void*)&ntiles, (void*)&ns, (void*)&tile_b0, (void*)&tile_b1, (void*)&tile_scene,
                    (void*)&scene_tile0, (void*)&B, (void*)&tol, (void*)&max_iter, (void*)&nb, (void*)&n_pairs};
#if defined(ABD_EMULATION)
    (void)args;
    if (blocks != 1) return hipErrorNotSupported;
    hipLaunchKernelGGL(k_pcg_multi<T>, dim3(1), dim3(64 * PCG_WAVES), 0, s, ntiles, ns, tile_b0, tile_b1, tile_scene,
                       scene_tile0, B, tol, max_iter, nb, n_pairs);
    return hipSuccess;
#else
    return launch_coop((const void*)k_pcg_multi<T>, blocks, 64 * PCG_WAVES, args, s);
#endif
}

// co-resident workgroups of k_pcg_multi<T>
template <typename T>
hipError_t pcg_multi_max_blocks(int* out) {
    int dev = 0, cus = 0, per_cu = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e == hipSuccess) e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    if (e == hipSuccess)
        e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, (const void*)k_pcg_multi<T>, 64 * PCG_WAVES, 0);
    *out = (e == hipSuccess) ? per_cu * cus : 0;
    return e;
}
template hipError_t pcg_multi_max_blocks<double>(int*);
template hipError_t pcg_multi_max_blocks<float>(int*);
template hipError_t launch_pcg_multi<double>(hipStream_t, int, int, int, const int*, const int*, const int*, const int*,
                                             const PcgBufT<double>&, double, int, int, int);
template hipError_t launch_pcg_multi<float>(hipStream_t, int, int, int, const int*, const int*, const int*, const int*,
                                            const PcgBufT<float>&, double, int, int, int);
template hipError_t launch_pcg_persist<float>(hipStream_t, int, int, const int*, const int*, const int*, const int*,
                                              const PcgBufT<float>&, double, int, int, int);

__global__ void k_to_f32(long n, const double* __restrict__ a, float* __restrict__ b) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) b[i] = (float)a[i];
}
__global__ void k_to_f64(long n, const float* __restrict__ a, double* __restrict__ b) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) b[i] = (double)a[i];
}
void launch_to_f32(hipStream_t s, long n, const double* a, float* b) {
    if (n > 0) hipLaunchKernelGGL(k_to_f32, dim3(cdiv(n, 256)), dim3(256), 0, s, n, a, b);
}
void launch_to_f64(hipStream_t s, long n, const float* a, double* b) {
    if (n > 0) hipLaunchKernelGGL(k_to_f64, dim3(cdiv(n, 256)), dim3(256), 0, s, n, a, b);
}

// any scene still iterating after `it` completed iterations (residual of parity it&1)
__global__ __launch_bounds__(64) void k_pcg_any(int ns, int ntiles, const int* __restrict__ st0,
                                                 const double* __restrict__ part, int it,
                                                 const double* __restrict__ bnorm, double tol,
                                                 const int* __restrict__ cg_done, int* any) {
    int s = blockIdx.x;
    if (!pcg_done(part, ntiles, it & 1, st0[s], st0[s + 1], threadIdx.x, bnorm[s], tol, cg_done + s))
        if (threadIdx.x == 0) atomicOr(any, 1);
}

void launch_pcg_any(hipStream_t s, int ns, int ntiles, const int* scene_tile0, const double* part, int it,
                    const double* bnorm, double tol, const int* cg_done, int* any) {
    hipLaunchKernelGGL(k_pcg_any, dim3(ns), dim3(64), 0, s, ns, ntiles, scene_tile0, part, it, bnorm, tol, cg_done,
                       any);
}

// ---------------------------------------------------------------------------------------------
// learned warm start: body features, the step the MLP predicts, and the residual it leaves
// ---------------------------------------------------------------------------------------------
// raw features per body: q - q~ (12), qdot (12), incident active contacts (1), summed contact
// gradient (12), from the incidences of the last assembly
__global__ void k_body_features(Geo G, const double* __restrict__ q, const double* __restrict__ qtilde,
                                const double* __restrict__ qdot, Cnt cinc, const uint64_t* __restrict__ inc,
                                const double* __restrict__ g24, Cnt cact, const double* __restrict__ pinv,
                                const double* __restrict__ grad, float* __restrict__ feat) {
    const int n_inc = cnt_of(cinc), n_act = cnt_of(cact);
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= G.nb) return;
    float* f = feat + (size_t)ABD_N_FEATURES * b;
    for (int j = 0; j < 12; ++j) {
        f[j] = (float)(q[12 * b + j] - qtilde[12 * b + j]);
        f[12 + j] = (float)qdot[12 * b + j];
    }
    const int lo = lower_bound64(inc, n_inc, (uint64_t)b << 32), hi = lower_bound64(inc, n_inc, (uint64_t)(b + 1) << 32);
    double gs[12];
    for (int j = 0; j < 12; ++j) gs[j] = 0.0;
    for (int m = lo; m < hi; ++m) {
        const uint32_t w = (uint32_t)inc[m];
        const int c = opaque((int)(w >> 1)), role = opaque((int)(w & 1));
        if (c >= n_act) {
            atomicOr(G.err, ERR_INCIDENCE);
            break;
        }
        for (int j = 0; j < 12; ++j) gs[j] += g24[(size_t)c * 24 + role * 12 + j];
    }
    f[24] = (float)(hi - lo);
    for (int j = 0; j < 12; ++j) f[25 + j] = (float)gs[j];
    // block-Jacobi step z = D_b^-1 grad_b (the preconditioner's own guess)
    const double* P = pinv + 144 * (size_t)b;
    const double* g = grad + 12 * (size_t)b;
    for (int i = 0; i < 12; ++i) {
        double z = 0.0;
        for (int j = 0; j < 12; ++j) z += P[12 * i + j] * g[j];
        f[ABD_FEAT_Z + i] = (float)z;
    }
}

void launch_body_features(hipStream_t s, const Geo& G, const double* q, const double* qtilde, const double* qdot,
                          Cnt n_inc, const uint64_t* inc, const double* g24, Cnt n_act, const double* pinv,
                          const double* grad, float* feat) {
    hipLaunchKernelGGL(k_body_features, dim3(cdiv(G.nb, 128)), dim3(128), 0, s, G, q, qtilde, qdot, n_inc, inc, g24,
                       n_act, pinv, grad, feat);
}

// standardised features, clipped to +-ABD_FEAT_CLIP -> BF16 model input (rows x dp, zero padding)
__global__ void k_features_pack(int nb, int rows, int dp, const float* __restrict__ feat, const float* __restrict__ mean,
                                const float* __restrict__ istd, uint16_t* __restrict__ out) {
    const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (long)rows * dp) return;
    const int i = (int)(e / dp), j = (int)(e % dp);
    float v = 0.f;
    if (i < nb && j < ABD_N_FEATURES) v = (feat[(size_t)ABD_N_FEATURES * i + j] - mean[j]) * istd[j];
    out[e] = f2bf(fminf(fmaxf(v, -ABD_FEAT_CLIP), ABD_FEAT_CLIP));
}

void launch_features_pack(hipStream_t s, int nb, int rows, int dp, const float* feat, const float* mean,
                          const float* istd, uint16_t* out) {
    const long n = (long)rows * dp;
    hipLaunchKernelGGL(k_features_pack, dim3(cdiv(n, 256)), dim3(256), 0, s, nb, rows, dp, feat, mean, istd, out);
}

// x0 = z + de-standardised model output (the model predicts du - z; pinned bodies: 0)
__global__ void k_warm_x0(Geo G, int dpo, const float* __restrict__ Y, const float* __restrict__ omean,
                          const float* __restrict__ ostd, const float* __restrict__ feat, double* __restrict__ x0) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= 12 * G.nb) return;
    const int b = e / 12, j = e % 12;
    const float z = feat[(size_t)ABD_N_FEATURES * b + ABD_FEAT_Z + j];
    x0[e] = G.pinned[b] ? 0.0 : (double)z + (double)(Y[(size_t)b * dpo + j] * ostd[j] + omean[j]);
}

void launch_warm_x0(hipStream_t s, const Geo& G, int dpo, const float* Y, const float* omean, const float* ostd,
                    const float* feat, double* x0) {
    hipLaunchKernelGGL(k_warm_x0, dim3(cdiv(12 * G.nb, 256)), dim3(256), 0, s, G, dpo, Y, omean, ostd, feat, x0);
}

// x0 = (1 - alpha_s) du: the part of the previous Newton direction that its line search did not
// take (pinned bodies and non-finite values: 0)
__global__ void k_prev_x0(Geo G, const double* __restrict__ alpha, const double* __restrict__ du,
                          double* __restrict__ x0) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= 12 * G.nb) return;
    const int b = e / 12;
    const double v = (1.0 - alpha[G.bscene[b]]) * du[e];
    x0[e] = (G.pinned[b] || !isfinite(v)) ? 0.0 : v;
}

void launch_prev_x0(hipStream_t s, const Geo& G, const double* alpha, const double* du, double* x0) {
    hipLaunchKernelGGL(k_prev_x0, dim3(cdiv(12 * G.nb, 256)), dim3(256), 0, s, G, alpha, du, x0);
}

// Learned block preconditioner: P_b = L_D^-T (T T^T) L_D^-1, where D_b = L_D L_D^T is the body's
// diagonal block and T the lower-triangular factor the model predicts (diagonal exp of its first 12
// outputs, the 66 entries below the diagonal row by row from the rest), both de-standardised.
// T = I gives the block-Jacobi D_b^-1; any T with a positive diagonal gives an SPD P_b.  One wave
// per body; pinned bodies and blocks whose Cholesky fails keep the block-Jacobi inverse.
__global__ __launch_bounds__(64) void k_learned_pinv(Geo G, int dpo, const float* __restrict__ Y,
                                                      const float* __restrict__ omean, const float* __restrict__ ostd,
                                                      const double* __restrict__ diag, double* __restrict__ pinv) {
    __shared__ double Lb[144], Li[144], Tm[144], X[144];
    const int b = blockIdx.x, lane = threadIdx.x;
    if (G.pinned[b]) return;
    for (int e = lane; e < 144; e += 64) {
        Lb[e] = diag[(size_t)b * 144 + e];
        Tm[e] = 0.0;
    }
    __syncthreads();
    // Cholesky D = L L^T in place (lower triangle)
    for (int k = 0; k < 12; ++k) {
        const double d = Lb[k * 12 + k];
        __syncthreads();  // every lane has read the pivot before it is overwritten
        if (!(d > 0.0)) return;  // same value in every lane: uniform exit, block-Jacobi kept
        const double sd = sqrt(d);
        if (lane == 0) Lb[k * 12 + k] = sd;
        if (lane > k && lane < 12) Lb[lane * 12 + k] /= sd;
        __syncthreads();
        for (int e = lane; e < 144; e += 64) {
            const int i = e / 12, j = e % 12;
            if (j > k && i >= j) Lb[e] -= Lb[i * 12 + k] * Lb[j * 12 + k];
        }
        __syncthreads();
    }
    // Li = L^-1 (lower): lane j solves L x = e_j by forward substitution
    if (lane < 12) {
        const int j = lane;
        for (int i = 0; i < 12; ++i) {
            double v = (i == j) ? 1.0 : 0.0;
            for (int k = j; k < i; ++k) v -= Lb[i * 12 + k] * Li[k * 12 + j];
            Li[i * 12 + j] = (i < j) ? 0.0 : v / Lb[i * 12 + i];
        }
    }
    // T from the model output row of this body
    const float* y = Y + (size_t)b * dpo;
    if (lane < 12) Tm[lane * 12 + lane] = exp((double)(y[lane] * ostd[lane] + omean[lane]));
    for (int m = lane; m < 66; m += 64) {
        int i = 1, rem = m;
        while (rem >= i) { rem -= i; ++i; }  // m-th entry below the diagonal, row by row
        Tm[i * 12 + rem] = (double)(y[12 + m] * ostd[12 + m] + omean[12 + m]);
    }
    __syncthreads();
    // X = Li^T T, then P = X X^T
    for (int e = lane; e < 144; e += 64) {
        const int i = e / 12, j = e % 12;
        double v = 0.0;
        for (int k = 0; k < 12; ++k) v += Li[k * 12 + i] * Tm[k * 12 + j];
        X[e] = v;
    }
    __syncthreads();
    for (int e = lane; e < 144; e += 64) {
        const int i = e / 12, j = e % 12;
        double v = 0.0;
        for (int k = 0; k < 12; ++k) v += X[i * 12 + k] * X[j * 12 + k];
        pinv[(size_t)b * 144 + e] = v;
    }
}

void launch_learned_pinv(hipStream_t s, const Geo& G, int dpo, const float* Y, const float* omean, const float* ostd,
                         const double* diag, double* pinv) {
    hipLaunchKernelGGL(k_learned_pinv, dim3(G.nb), dim3(64), 0, s, G, dpo, Y, omean, ostd, diag, pinv);
}

// r = b - H x (H: diagonal blocks + CSR off-diagonal blocks), one thread per row
__global__ void k_residual(int nb, const double* __restrict__ bvec, const double* __restrict__ diag,
                           const int* __restrict__ row_ptr, const uint64_t* __restrict__ ent_key,
                           const int* __restrict__ ent_val, const double* __restrict__ off,
                           const double* __restrict__ offT, const double* __restrict__ x, double* __restrict__ r,
                           int n_pairs, int* err) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= 12 * nb) return;
    const int b = e / 12, row = e % 12;
    double acc = bvec[e];
    const double* D = diag + (size_t)b * 144 + row * 12;
    for (int c = 0; c < 12; ++c) acc -= D[c] * x[12 * b + c];
    for (int k = row_ptr[b]; k < row_ptr[b + 1]; ++k) {
        const int col = opaque((int)(ent_key[k] & 0xFFFFFFFFu));
        const int v = ent_val[k];
        if (col < 0 || col >= nb || v < 0 || (v >> 1) >= n_pairs) {
            atomicOr(err, ERR_CSR);
            break;
        }
        const double* Bk = (opaque(v & 1) ? offT : off) + (size_t)opaque(v >> 1) * 144 + row * 12;
        for (int c = 0; c < 12; ++c) acc -= Bk[c] * x[12 * col + c];
    }
    r[e] = acc;
}

void launch_residual(hipStream_t s, int nb, const double* b, const double* diag, const int* row_ptr,
                     const uint64_t* ent_key, const int* ent_val, const double* off, const double* offT,
                     const double* x, double* r, int n_pairs, int* err) {
    hipLaunchKernelGGL(k_residual, dim3(cdiv(12 * nb, 256)), dim3(256), 0, s, nb, b, diag, row_ptr, ent_key, ent_val,
                       off, offT, x, r, n_pairs, err);
}

// ||v|| per scene over its bodies' rows (one workgroup per scene, fixed order)
__global__ __launch_bounds__(256) void k_scene_norm(Geo G, const double* __restrict__ v, double* __restrict__ out) {
    __shared__ double sh[256];
    const int s = blockIdx.x;
    double acc = 0.0;
    for (int e = 12 * G.scene_b0[s] + threadIdx.x; e < 12 * G.scene_b0[s + 1]; e += 256) acc += v[e] * v[e];
    sh[threadIdx.x] = acc;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (threadIdx.x < o) sh[threadIdx.x] += sh[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[s] = sqrt(sh[0]);
}

void launch_scene_norm(hipStream_t s, const Geo& G, const double* v, double* out) {
    hipLaunchKernelGGL(k_scene_norm, dim3(G.ns), dim3(256), 0, s, G, v, out);
}

// Keep the warm start of a scene only if its residual is below ||b|| (else start from zero:
// rhs = b, x0 = 0); used[s] records the choice
__global__ void k_warm_select(Geo G, const double* __restrict__ nb_b, const double* __restrict__ nb_r,
                              const double* __restrict__ bvec, double* __restrict__ rhs, double* __restrict__ x0,
                              int* __restrict__ used) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= 12 * G.nb) return;
    const int b = e / 12, s = G.bscene[b];
    const bool keep = nb_r[s] < nb_b[s];
    if (!keep) {
        rhs[e] = bvec[e];
        x0[e] = 0.0;
    }
    if (e == 12 * G.scene_b0[s]) used[s] = keep ? 1 : 0;
}

void launch_warm_select(hipStream_t s, const Geo& G, const double* nb_b, const double* nb_r, const double* b,
                        double* rhs, double* x0, int* used) {
    hipLaunchKernelGGL(k_warm_select, dim3(cdiv(12 * G.nb, 256)), dim3(256), 0, s, G, nb_b, nb_r, b, rhs, x0, used);
}

__global__ void k_add_inplace(int n, double* __restrict__ y, const double* __restrict__ x) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) y[i] += x[i];
}

void launch_add_inplace(hipStream_t s, int n, double* y, const double* x) {
    hipLaunchKernelGGL(k_add_inplace, dim3(cdiv(n, 256)), dim3(256), 0, s, n, y, x);
}

// ---------------------------------------------------------------------------------------------
// CCD bound (additive CCD, Li, Kaufman, Jiang 2021) per candidate -> per-scene minimum
// ---------------------------------------------------------------------------------------------
__device__ double accd_dev(int kind, const D3* x0, const D3* x1, double s, double tmax, int max_iter) {
    D3 p[4], x[4];
    double lp = motion_bound(kind, x0, x1, p);
    for (int i = 0; i < 4; ++i) x[i] = x0[i];
    if (lp == 0.0) return tmax;
    double d = sqrt(pair_d2(kind, x));
    double g = s * d;
    double t = 0.0;
    double tl = (1.0 - s) * d / lp;
    for (int it = 0; it < max_iter; ++it) {
        for (int i = 0; i < 4; ++i) x[i] = x[i] + tl * p[i];
        d = sqrt(pair_d2(kind, x));
        if (t > 0.0 && d < g) return t;
        t += tl;
        if (t > tmax) return tmax;
        tl = 0.9 * d / lp;
    }
    return t;
}

__global__ void k_ccd(Geo G, Cnt cn, const uint64_t* __restrict__ keys, const double4* __restrict__ X0,
                      const double4* __restrict__ X1, AbdParams P, unsigned long long* amax_bits) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= cnt_of(cn)) return;
    PairV p = decode(G, keys[i]);
    if (p.kind < 0) return;
    double toi;
    if (p.kind == K_PG) {
        D3 a = ld3(X0, p.v[0]), b = ld3(X1, p.v[0]);
        D3 nrm = mk(P.ground_normal[0], P.ground_normal[1], P.ground_normal[2]);
        double d0 = dot(nrm, a) - P.ground_offset;
        double v = -(dot(nrm, b) - dot(nrm, a));
        // moving away: no bound; moving down from on or below the plane: no step at all
        toi = (v <= 0.0) ? 1.0 : (d0 <= 0.0 ? 0.0 : fmin(1.0, (1.0 - P.ccd_gap_ratio) * d0 / v));
    } else {
        D3 a[4], b[4];
        for (int k = 0; k < 4; ++k) { a[k] = ld3(X0, p.v[k]); b[k] = ld3(X1, p.v[k]); }
        toi = accd_dev(p.kind, a, b, P.ccd_gap_ratio, 1.0, P.ccd_max_iter);
    }
    if (toi < 1.0) atomicMin(&amax_bits[p.scene], (unsigned long long)__double_as_longlong(toi));
}

void launch_ccd(hipStream_t s, const Geo& G, Cnt n, const uint64_t* keys, const double4* X0, const double4* X1,
                const AbdParams& P, unsigned long long* amax_bits) {
    if (n.cap <= 0) return;
    hipLaunchKernelGGL(k_ccd, dim3(cdiv(n.cap, 128)), dim3(128), 0, s, G, n, keys, X0, X1, P, amax_bits);
}

// energy_inertia (abd.py:53-64) at q - alpha_s du
__global__ void k_body_energy(Geo G, AbdParams P, const double* __restrict__ q, const double* __restrict__ du,
                              const double* __restrict__ alpha, const double* __restrict__ qtilde,
                              double* __restrict__ e_body, const int* __restrict__ active) {
    int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= G.nb) return;
    e_body += (size_t)blockIdx.y * G.nb;  // k = blockIdx.y: the trial at 2^-k alpha
    // pinned bodies, and (with `active`) bodies of scenes whose Newton loop has stopped: 0
    if (G.pinned[b] || (active && !active[G.bscene[b]])) { e_body[b] = 0.0; return; }
    double a = alpha ? ldexp(alpha[G.bscene[b]], -(int)blockIdx.y) : 0.0;
    double qb[12], dq[12];
    for (int k = 0; k < 12; ++k) {
        qb[k] = du ? q[12 * b + k] - a * du[12 * b + k] : q[12 * b + k];
        dq[k] = qb[k] - qtilde[12 * b + k];
    }
    const double* M = G.mass4 + 16 * b;
    double ein = 0.0;
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j)
            ein += M[4 * i + j] * (dq[3 * i] * dq[3 * j] + dq[3 * i + 1] * dq[3 * j + 1] + dq[3 * i + 2] * dq[3 * j + 2]);
    e_body[b] = 0.5 * ein + P.dt * P.dt * ortho_energy(qb, P.kappa);
}

void launch_body_energy(hipStream_t s, const Geo& G, const AbdParams& P, const double* q, const double* du,
                        const double* alpha, const double* qtilde, double* e_body, const int* active, int trials) {
    hipLaunchKernelGGL(k_body_energy, dim3(cdiv(G.nb, 128), trials), dim3(128), 0, s, G, P, q, du, alpha, qtilde,
                       e_body, active);
}

// line_search_iterative (gauss_newton.py:98-115) per scene
__device__ void ls_init_scene(int s, const AbdParams& P, const int* active, const SceneCtl& c) {
    double amax = P.ccd_enabled ? fmin(1.0, __longlong_as_double((long long)c.amax_bits[s])) : 1.0;
    c.alpha0[s] = amax;
    c.alpha[s] = active[s] ? amax : 0.0;
    c.ls_done[s] = active[s] ? 0 : 1;
    c.ls_floor[s] = 0;
}


// Accept the first halving step that lowers the energy.  At the floor without a decrease the
// Newton line search keeps the floor step (gauss_newton.py:110-111) unless its energy is not finite
// (a trial through the ground plane), the predictor's search takes no step (floor_accept 0).
__device__ void ls_decide_scene(int s, const AbdParams& P, const SceneCtl& c, int floor_accept, int* pending) {
    if (c.ls_done[s]) return;
    if (c.E1[s] < c.E0[s]) {
        c.ls_done[s] = 1;
    } else if (c.alpha[s] * 0.5 < P.ls_floor * c.alpha0[s]) {
        c.ls_done[s] = 1;
        c.ls_floor[s] = 1;
        if (!floor_accept || !isfinite(c.E1[s])) {
            c.alpha[s] = 0.0;  // no step: the accepted iterate is q itself
            c.E1[s] = c.E0[s];
        }
    } else {
        c.alpha[s] *= 0.5;
        atomicOr(pending, 1);
    }
}


// Direction toward the learned predicted first iterate q_now + dq, dq = dt qdot + de-standardised
// model output (the model predicts the step's correction over the extrapolation); pinned bodies and
// non-finite predictions stay
__global__ void k_step_pred(Geo G, int dpo, const float* __restrict__ Y, const float* __restrict__ omean,
                            const float* __restrict__ ostd, double dt, const double* __restrict__ q,
                            const double* __restrict__ qnow, const double* __restrict__ qdot, double* __restrict__ du) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= 12 * G.nb) return;
    const int b = e / 12, j = e % 12;
    const double dq = dt * qdot[e] + (double)(Y[(size_t)b * dpo + j] * ostd[j] + omean[j]);
    const double d = q[e] - qnow[e] - dq;
    du[e] = (G.pinned[b] || !isfinite(d)) ? 0.0 : d;
}

void launch_step_pred(hipStream_t s, const Geo& G, int dpo, const float* Y, const float* omean, const float* ostd,
                      double dt, const double* q, const double* qnow, const double* qdot, double* du) {
    hipLaunchKernelGGL(k_step_pred, dim3(cdiv(12 * G.nb, 256)), dim3(256), 0, s, G, dpo, Y, omean, ostd, dt, q, qnow,
                       qdot, du);
}

// Direction toward the predicted first iterate, as a Newton step du (q - alpha du moves toward the
// target): mode 1 = q_now + dt qdot, mode 2 = q~; pinned bodies stay
__global__ void k_predict_du(Geo G, int mode, double dt, const double* __restrict__ q, const double* __restrict__ qnow,
                             const double* __restrict__ qdot, const double* __restrict__ qtilde,
                             double* __restrict__ du) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= 12 * G.nb) return;
    const double target = (mode == 2) ? qtilde[i] : qnow[i] + dt * qdot[i];
    const double d = q[i] - target;
    du[i] = (G.pinned[i / 12] || !isfinite(d)) ? 0.0 : d;
}

void launch_predict_du(hipStream_t s, const Geo& G, int mode, double dt, const double* q, const double* qnow,
                       const double* qdot, const double* qtilde, double* du) {
    hipLaunchKernelGGL(k_predict_du, dim3(cdiv(12 * G.nb, 256)), dim3(256), 0, s, G, mode, dt, q, qnow, qdot, qtilde,
                       du);
}

// smallest d^2 per body over the candidate pairs at X (ground pairs: the squared gap)
__global__ void k_dmin(Geo G, Cnt cn, const uint64_t* __restrict__ keys, const double4* __restrict__ X, AbdParams P,
                       unsigned long long* dmin_bits) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= cnt_of(cn)) return;
    PairV p = decode(G, keys[i]);
    if (p.kind < 0) return;
    const unsigned long long d2 = (unsigned long long)__double_as_longlong(cand_d2(G, p, X, P));
    atomicMin(&dmin_bits[pair_body(G, p, 0)], d2);
    if (p.kind != K_PG) atomicMin(&dmin_bits[pair_body(G, p, 3)], d2);
}

void launch_dmin(hipStream_t s, const Geo& G, Cnt n, const uint64_t* keys, const double4* X, const AbdParams& P,
                 unsigned long long* dmin_bits) {
    if (n.cap <= 0) return;
    hipLaunchKernelGGL(k_dmin, dim3(cdiv(n.cap, 256)), dim3(256), 0, s, G, n, keys, X, P, dmin_bits);
}

// Adaptive contact stiffness, per body (after Li et al. 2020, Incremental Potential Contact,
// section 4.3, where one stiffness is raised for the whole scene): a body of an iterating scene
// doubles its stiffness, up to k_max, when its closest contact is nearer than k_raise_gap * dhat and
// nearer than after its previous iteration of the step.
// skip (optional): a device flag; when set the kernel changes nothing (an iteration whose result the
// host discards)
__global__ void k_stiffness_update(Geo G, AbdParams P, const int* __restrict__ active, SceneCtl c,
                                   const int* __restrict__ skip) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= G.nb || (skip && *skip)) return;
    const int s = G.bscene[b];
    if (!active[s]) return;
    const double d = sqrt(__longlong_as_double((long long)c.dmin_bits[b]));
    if (P.k_max > c.kb[b] && d < P.k_raise_gap * P.dhat && d < c.dmin_prev[b]) {
        c.kb[b] = fmin(2.0 * c.kb[b], P.k_max);
        atomicAdd(&c.k_raises[s], 1);
    }
    c.dmin_prev[b] = d;
}

void launch_stiffness_update(hipStream_t s, const Geo& G, const AbdParams& P, const int* active, SceneCtl c,
                             const int* skip) {
    hipLaunchKernelGGL(k_stiffness_update, dim3(cdiv(G.nb, 128)), dim3(128), 0, s, G, P, active, c, skip);
}

// largest body stiffness of each scene (one workgroup per scene)
__global__ __launch_bounds__(256) void k_scene_kmax(Geo G, SceneCtl c) {
    __shared__ double sh[4];
    const int s = blockIdx.x;
    double k = 0.0;
    for (int b = G.scene_b0[s] + threadIdx.x; b < G.scene_b0[s + 1]; b += 256) k = fmax(k, c.kb[b]);
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) k = fmax(k, __shfl_xor(k, o, 64));
    if (threadIdx.x % 64 == 0) sh[threadIdx.x / 64] = k;
    __syncthreads();
    if (threadIdx.x == 0) c.kc[s] = fmax(fmax(sh[0], sh[1]), fmax(sh[2], sh[3]));
}

void launch_scene_kmax(hipStream_t s, const Geo& G, SceneCtl c) {
    hipLaunchKernelGGL(k_scene_kmax, dim3(G.ns), dim3(256), 0, s, G, c);
}

// add_du_kernel (abd.py:780-785): q -= alpha du (nothing when the optional skip flag is set)
__global__ void k_apply(Geo G, double* q, const double* __restrict__ du, const double* __restrict__ alpha,
                        const int* __restrict__ active, const int* __restrict__ skip) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= 12 * G.nb || (skip && *skip)) return;
    int s = G.bscene[i / 12];
    if (active[s]) q[i] -= alpha[s] * du[i];
}

void launch_apply(hipStream_t s, const Geo& G, double* q, const double* du, const double* alpha, const int* active,
                  const int* skip) {
    hipLaunchKernelGGL(k_apply, dim3(cdiv(12 * G.nb, 256)), dim3(256), 0, s, G, q, du, alpha, active, skip);
}

// forward_states (abd.py:574-581)
__global__ void k_forward(int n, double dt, const double* __restrict__ q, double* qnow, double* qdot) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    qdot[i] = (q[i] - qnow[i]) / dt;
    qnow[i] = q[i];
}

void launch_forward(hipStream_t s, int nb, double dt, const double* q, double* qnow, double* qdot) {
    hipLaunchKernelGGL(k_forward, dim3(cdiv(12 * nb, 256)), dim3(256), 0, s, 12 * nb, dt, q, qnow, qdot);
}


// get_contact_points (abd.py:759-766): midpoint of the closest points and the distance
__global__ void k_contact_points(Geo G, int n, const uint64_t* __restrict__ keys, const double4* __restrict__ X,
                                 AbdParams P, double* mid, double* dist) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    PairV p = decode(G, keys[i]);
    if (p.kind < 0) { mid[3 * i] = mid[3 * i + 1] = mid[3 * i + 2] = 0.0; dist[i] = 0.0; return; }
    D3 c0, c1;
    if (p.kind == K_PG) {
        D3 x = ld3(X, p.v[0]);
        D3 nrm = mk(P.ground_normal[0], P.ground_normal[1], P.ground_normal[2]);
        double d = dot(nrm, x) - P.ground_offset;
        c0 = x;
        c1 = x - d * nrm;
    } else {
        D3 x[4];
        for (int k = 0; k < 4; ++k) x[k] = ld3(X, p.v[k]);
        DistEval E;
        dist_eval(p.kind, x, E);
        // closest point on the far feature of the stencil: y0 vs. the rest
        if (E.S.type == T_PP) { c0 = E.y[0]; c1 = E.y[1]; }
        else if (E.S.type == T_PE) {
            D3 e = E.y[2] - E.y[1];
            double t = dot(E.y[0] - E.y[1], e) / dot(e, e);
            c0 = E.y[0]; c1 = E.y[1] + t * e;
        } else if (E.S.type == T_PT) {
            D3 nn = (1.0 / sqrt(E.N)) * E.n;
            c0 = E.y[0];
            c1 = E.y[0] - dot(E.y[0] - E.y[1], nn) * nn;
        } else {
            D3 d1 = E.y[1] - E.y[0], d2 = E.y[3] - E.y[2], r = E.y[0] - E.y[2];
            double a = dot(d1, d1), e = dot(d2, d2), f = dot(d2, r), c = dot(d1, r), b = dot(d1, d2);
            double den = a * e - b * b;
            double s = (b * f - c * e) / den, t = (b * s + f) / e;
            c0 = E.y[0] + s * d1; c1 = E.y[2] + t * d2;
        }
    }
    D3 m = 0.5 * (c0 + c1);
    mid[3 * i] = m.x; mid[3 * i + 1] = m.y; mid[3 * i + 2] = m.z;
    D3 dd = c0 - c1;
    dist[i] = sqrt(dot(dd, dd));
}

void launch_contact_points(hipStream_t s, const Geo& G, int n, const uint64_t* keys, const double4* X,
                           const AbdParams& P, double* mid, double* dist) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_contact_points, dim3(cdiv(n, 256)), dim3(256), 0, s, G, n, keys, X, P, mid, dist);
}

}  // namespace abd
