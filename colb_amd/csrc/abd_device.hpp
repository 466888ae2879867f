// abd_device.hpp — device-side FP64 math of the ABD Newton step on gfx950.
//
// Contact distances follow the IPC distance types (Li et al. 2020): after closest-feature
// classification a pair's squared distance is one of
//   PP  |y0 - y1|^2
//   PE  |(e0 - p) x (e1 - p)|^2 / |e1 - e0|^2
//   PT / EE  (a.n)^2 / |n|^2,  n = u x v, with a, u, v linear in the stencil points.
// The reference obtains these derivatives from the (un-vendored) ipctk_wp module
// (abd.py:15-17, 247, 376, 485) and the closest points from Warp's closest_point_edge_edge
// (abd.py:190).  Here each Hessian entry is an explicit function of a few shared
// intermediates, so the 64 lanes of a wave evaluate the 12x12 entries in parallel.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace abd {

// Optimisation barrier for bit-field extractions that feed address arithmetic.  hipcc (ROCm 7.2)
// miscompiles `(size_t)(x & 0xFFFFFF) * C`: the mask is kept for comparisons but dropped from the
// operand of the v_mad_u64_u32 that forms the address, so the unmasked register is used.  An
// empty asm with a read-write register operand makes the masked value opaque and forces its use.
#if defined(ABD_EMULATION)
#define ABD_OPAQUE(v) asm volatile("" : "+r"(v))
#else
#define ABD_OPAQUE(v) asm volatile("" : "+v"(v))
#endif
__device__ __forceinline__ int opaque(int v) {
    ABD_OPAQUE(v);
    return v;
}

struct D3 {
    double x, y, z;
};
__device__ __forceinline__ D3 mk(double x, double y, double z) { return {x, y, z}; }
__device__ __forceinline__ D3 operator+(D3 a, D3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ __forceinline__ D3 operator-(D3 a, D3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ D3 operator*(double s, D3 a) { return {s * a.x, s * a.y, s * a.z}; }
__device__ __forceinline__ double dot(D3 a, D3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ D3 cross(D3 a, D3 b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
__device__ __forceinline__ double cmp(D3 a, int k) { return k == 0 ? a.x : (k == 1 ? a.y : a.z); }
// ([w]x)_{rc}
__device__ __forceinline__ double skew_rc(D3 w, int r, int c) {
    if (r == c) return 0.0;
    if (r == 0) return c == 1 ? -w.z : w.y;
    if (r == 1) return c == 0 ? w.z : -w.x;
    return c == 0 ? -w.y : w.x;
}

enum { K_PT = 0, K_EE = 1, K_PG = 2 };
enum { T_PP = 0, T_PE = 1, T_PT = 2, T_EE = 3, T_PG = 4 };

// barrier.py:47-66: cubic barrier in s = d^2
__device__ __forceinline__ void barrier3(double s, double shat, double& B, double& B1, double& B2) {
    if (s < shat) {
        double d = s - shat;
        B = -d * d * d;
        B1 = -3.0 * d * d;
        B2 = -6.0 * d;
    } else {
        B = B1 = B2 = 0.0;
    }
}

// ortho.py:9-16: kappa * sum_ij (a_i.a_j - delta_ij)^2, a = rows 1..3 of the DOF vector
__device__ __forceinline__ double ortho_energy(const double* q, double kappa) {
    const double* a = q + 3;
    double e = 0;
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            double t = a[3 * i] * a[3 * j] + a[3 * i + 1] * a[3 * j + 1] + a[3 * i + 2] * a[3 * j + 2] - (i == j ? 1.0 : 0.0);
            e += t * t;
        }
    return kappa * e;
}

// ---------------------------------------------------------------------------------------------
// closest-feature classification (Ericson, Real-Time Collision Detection 5.1.5 and 5.1.9)
// ---------------------------------------------------------------------------------------------
struct Stencil {
    int type, ns;
    int st[4];
};

__device__ __forceinline__ Stencil classify_pt(const D3* x) {
    Stencil S;
    D3 p = x[0], A = x[1], B = x[2], C = x[3];
    D3 ab = B - A, ac = C - A, ap = p - A;
    double d1 = dot(ab, ap), d2 = dot(ac, ap);
    if (d1 <= 0 && d2 <= 0) { S.type = T_PP; S.ns = 2; S.st[0] = 0; S.st[1] = 1; return S; }
    D3 bp = p - B;
    double d3 = dot(ab, bp), d4 = dot(ac, bp);
    if (d3 >= 0 && d4 <= d3) { S.type = T_PP; S.ns = 2; S.st[0] = 0; S.st[1] = 2; return S; }
    double vc = d1 * d4 - d3 * d2;
    if (vc <= 0 && d1 >= 0 && d3 <= 0) { S.type = T_PE; S.ns = 3; S.st[0] = 0; S.st[1] = 1; S.st[2] = 2; return S; }
    D3 cp = p - C;
    double d5 = dot(ab, cp), d6 = dot(ac, cp);
    if (d6 >= 0 && d5 <= d6) { S.type = T_PP; S.ns = 2; S.st[0] = 0; S.st[1] = 3; return S; }
    double vb = d5 * d2 - d1 * d6;
    if (vb <= 0 && d2 >= 0 && d6 <= 0) { S.type = T_PE; S.ns = 3; S.st[0] = 0; S.st[1] = 1; S.st[2] = 3; return S; }
    double va = d3 * d6 - d5 * d4;
    if (va <= 0 && (d4 - d3) >= 0 && (d5 - d6) >= 0) { S.type = T_PE; S.ns = 3; S.st[0] = 0; S.st[1] = 2; S.st[2] = 3; return S; }
    S.type = T_PT; S.ns = 4; S.st[0] = 0; S.st[1] = 1; S.st[2] = 2; S.st[3] = 3;
    return S;
}

#define ABD_PARALLEL_EPS 1e-10

__device__ __forceinline__ Stencil classify_ee(const D3* x) {
    Stencil S;
    D3 d1 = x[1] - x[0], d2 = x[3] - x[2], r = x[0] - x[2];
    double a = dot(d1, d1), e = dot(d2, d2), f = dot(d2, r);
    double c = dot(d1, r), b = dot(d1, d2);
    double denom = a * e - b * b;
    double s, t;
    if (denom > ABD_PARALLEL_EPS * a * e) s = fmin(1.0, fmax(0.0, (b * f - c * e) / denom));
    else s = 0.0;
    t = (b * s + f) / e;
    if (t < 0.0) { t = 0.0; s = fmin(1.0, fmax(0.0, -c / a)); }
    else if (t > 1.0) { t = 1.0; s = fmin(1.0, fmax(0.0, (b - c) / a)); }
    bool s_end = (s == 0.0 || s == 1.0), t_end = (t == 0.0 || t == 1.0);
    int sa = (s == 1.0) ? 1 : 0, tb = (t == 1.0) ? 3 : 2;
    if (s_end && t_end) { S.type = T_PP; S.ns = 2; S.st[0] = sa; S.st[1] = tb; return S; }
    if (s_end) { S.type = T_PE; S.ns = 3; S.st[0] = sa; S.st[1] = 2; S.st[2] = 3; return S; }
    if (t_end) { S.type = T_PE; S.ns = 3; S.st[0] = tb; S.st[1] = 0; S.st[2] = 1; return S; }
    S.type = T_EE; S.ns = 4; S.st[0] = 0; S.st[1] = 1; S.st[2] = 2; S.st[3] = 3;
    return S;
}

// ---------------------------------------------------------------------------------------------
// Distance evaluation: shared intermediates + per-entry gradient/Hessian
// ---------------------------------------------------------------------------------------------
struct DistEval {
    Stencil S;
    D3 y[4];
    // plane form (PT/EE): d2 = s^2/N ; PE: d2 = N/L
    double al[4], mu[4], nu[4], om[4];
    D3 a, u, v, n, w;
    double s, N, L, d2;
    D3 vxa, axu;
};

// Stencil coefficients per type (PT: a = y0 - y1, u = y2 - y1, v = y3 - y1; EE: a = y0 - y2,
// u = y1 - y0, v = y3 - y2; PE: u = y1 - y0, v = y2 - y0, w = y2 - y1; PP: a = y0 - y1), written
// as blends of the type flags: branches that store into the arrays leave them in scratch memory.
__device__ __forceinline__ void set_coeffs(DistEval& E) {
    const double pt = (double)(E.S.type == T_PT), ee = (double)(E.S.type == T_EE), pe = (double)(E.S.type == T_PE),
                 pp = (double)(E.S.type == T_PP);
    E.al[0] = pt + ee + pp; E.al[1] = -pt - pp; E.al[2] = -ee; E.al[3] = 0.0;
    E.mu[0] = -ee - pe; E.mu[1] = -pt + ee + pe; E.mu[2] = pt; E.mu[3] = 0.0;
    E.nu[0] = -pe; E.nu[1] = -pt; E.nu[2] = -ee + pe; E.nu[3] = pt + ee;
    E.om[0] = 0.0; E.om[1] = -pe; E.om[2] = pe; E.om[3] = 0.0;
}

// Runtime-indexed reads of small register arrays as exact 0/1 blends (finite operands): a
// dynamically indexed private array, or a select chain the compiler folds back into one, lives
// in scratch memory.
__device__ __forceinline__ double sel4(const double* a, int i) {
    return (double)(i == 0) * a[0] + (double)(i == 1) * a[1] + (double)(i == 2) * a[2] + (double)(i == 3) * a[3];
}
__device__ __forceinline__ D3 pick4(const D3* x, int i) {
    const double w0 = (double)(i == 0), w1 = (double)(i == 1), w2 = (double)(i == 2), w3 = (double)(i == 3);
    return mk(w0 * x[0].x + w1 * x[1].x + w2 * x[2].x + w3 * x[3].x, w0 * x[0].y + w1 * x[1].y + w2 * x[2].y + w3 * x[3].y,
              w0 * x[0].z + w1 * x[1].z + w2 * x[2].z + w3 * x[3].z);
}

// x: the pair's 4 slot positions (PT: p,t0,t1,t2; EE: a0,a1,b0,b1)
__device__ __forceinline__ void dist_eval(int kind, const D3* x, DistEval& E) {
    E.S = (kind == K_PT) ? classify_pt(x) : classify_ee(x);
#pragma unroll
    for (int k = 0; k < 4; ++k) E.y[k] = pick4(x, k < E.S.ns ? E.S.st[k] : 0);
    set_coeffs(E);
    // Every field is written on every path, unused ones from zero coefficients (slots k >= ns have
    // zero coefficients and finite y): stores under type branches are merged by the compiler into
    // stores through a selected address, which leaves the struct in scratch memory.
    E.a = mk(0, 0, 0); E.u = mk(0, 0, 0); E.v = mk(0, 0, 0); E.w = mk(0, 0, 0);
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        E.a = E.a + E.al[k] * E.y[k];
        E.u = E.u + E.mu[k] * E.y[k];
        E.v = E.v + E.nu[k] * E.y[k];
        E.w = E.w + E.om[k] * E.y[k];
    }
    E.n = cross(E.u, E.v);
    E.N = dot(E.n, E.n);
    E.L = dot(E.w, E.w);
    E.s = dot(E.a, E.n);
    E.vxa = cross(E.v, E.a);
    E.axu = cross(E.a, E.u);
    const int ty = E.S.type;
    E.d2 = (ty == T_PP) ? dot(E.a, E.a) : (ty == T_PE) ? E.N / E.L : E.s * E.s / E.N;
}

// distance only (for CCD and energies)
__device__ __forceinline__ double pair_d2(int kind, const D3* x) {
    DistEval E;
    dist_eval(kind, x, E);
    return E.d2;
}

// Bound on how far the pair's distance can change along x0 -> x1 (additive CCD, Li, Kaufman,
// Jiang 2021): displacements minus their mean (a common translation leaves the distance
// unchanged), then the largest point motion on each primitive, summed.  p receives the
// centred displacements.
__device__ __forceinline__ double motion_bound(int kind, const D3* x0, const D3* x1, D3* p) {
    D3 mean = mk(0, 0, 0);
    for (int i = 0; i < 4; ++i) { p[i] = x1[i] - x0[i]; mean = mean + 0.25 * p[i]; }
    for (int i = 0; i < 4; ++i) p[i] = p[i] - mean;
    double n0 = sqrt(dot(p[0], p[0])), n1 = sqrt(dot(p[1], p[1])), n2 = sqrt(dot(p[2], p[2])),
           n3 = sqrt(dot(p[3], p[3]));
    return (kind == K_PT) ? n0 + fmax(n1, fmax(n2, n3)) : fmax(n0, n1) + fmax(n2, n3);
}

// Candidate rule of the broad phase: the pair can come closer than `infl` somewhere on the
// linear path x0 -> x1 (distance at x0 minus the motion bound; relative slack 1e-9 so a pair
// that is active at any point of the path is never dropped by rounding).
__device__ __forceinline__ bool may_reach(int kind, const D3* x0, const D3* x1, double infl) {
    D3 p[4];
    double lp = motion_bound(kind, x0, x1, p);
    double d2 = pair_d2(kind, x0);
    double lim = infl * (1.0 + 1e-9);
    return lp == 0.0 ? d2 < lim * lim : sqrt(d2) - lp < lim;
}

// (Dn_i)_{kr} = -mu_i [v]x_{kr} + nu_i [u]x_{kr}
__device__ __forceinline__ double Dn(const DistEval& E, int i, int k, int r) {
    return -sel4(E.mu, i) * skew_rc(E.v, k, r) + sel4(E.nu, i) * skew_rc(E.u, k, r);
}
// d s / d y_{i,r}
__device__ __forceinline__ double gs_at(const DistEval& E, int R) {
    int i = R / 3, r = R % 3;
    return sel4(E.al, i) * cmp(E.n, r) + sel4(E.mu, i) * cmp(E.vxa, r) + sel4(E.nu, i) * cmp(E.axu, r);
}
// d N / d y_{i,r} = 2 (Dn_i^T n)_r
__device__ __forceinline__ double gN_at(const DistEval& E, int R) {
    int i = R / 3, r = R % 3;
    return 2.0 * (Dn(E, i, 0, r) * E.n.x + Dn(E, i, 1, r) * E.n.y + Dn(E, i, 2, r) * E.n.z);
}
__device__ __forceinline__ double gL_at(const DistEval& E, int R) { return 2.0 * sel4(E.om, R / 3) * cmp(E.w, R % 3); }

// gradient of d^2 w.r.t. stencil coordinate R (0 <= R < 3 ns)
__device__ __forceinline__ double d2_grad(const DistEval& E, int R) {
    if (E.S.type == T_PP) {
        double g = 2.0 * cmp(E.a, R % 3);
        return R < 3 ? g : -g;
    }
    if (E.S.type == T_PE) return gN_at(E, R) / E.L - E.N / (E.L * E.L) * gL_at(E, R);
    return 2.0 * E.s / E.N * gs_at(E, R) - E.s * E.s / (E.N * E.N) * gN_at(E, R);
}

__device__ __forceinline__ double hN_at(const DistEval& E, int R, int C) {
    int i = R / 3, r = R % 3, j = C / 3, c = C % 3;
    double t = 0;
#pragma unroll
    for (int k = 0; k < 3; ++k) t += Dn(E, i, k, r) * Dn(E, j, k, c);
    return 2.0 * t + 2.0 * (-sel4(E.mu, i) * sel4(E.nu, j) + sel4(E.nu, i) * sel4(E.mu, j)) * skew_rc(E.n, r, c);
}

// Hessian of d^2 w.r.t. stencil coordinates (R, C)
__device__ __forceinline__ double d2_hess(const DistEval& E, int R, int C) {
    int i = R / 3, r = R % 3, j = C / 3, c = C % 3;
    if (E.S.type == T_PP) {
        if (r != c) return 0.0;
        return (i == j) ? 2.0 : -2.0;
    }
    if (E.S.type == T_PE) {
        double hL = (r == c) ? 2.0 * sel4(E.om, i) * sel4(E.om, j) : 0.0;
        double gNR = gN_at(E, R), gNC = gN_at(E, C), gLR = gL_at(E, R), gLC = gL_at(E, C);
        double L = E.L, N = E.N;
        return hN_at(E, R, C) / L - (gNR * gLC + gLR * gNC) / (L * L) + 2.0 * N / (L * L * L) * gLR * gLC -
               N / (L * L) * hL;
    }
    const double ali = sel4(E.al, i), mui = sel4(E.mu, i), nui = sel4(E.nu, i);
    const double alj = sel4(E.al, j), muj = sel4(E.mu, j), nuj = sel4(E.nu, j);
    double hs = ali * muj * (-skew_rc(E.v, r, c)) + ali * nuj * skew_rc(E.u, r, c) + mui * alj * skew_rc(E.v, r, c) +
                mui * nuj * (-skew_rc(E.a, r, c)) + nui * alj * (-skew_rc(E.u, r, c)) + nui * muj * skew_rc(E.a, r, c);
    double gsR = gs_at(E, R), gsC = gs_at(E, C), gNR = gN_at(E, R), gNC = gN_at(E, C);
    double s = E.s, N = E.N;
    return 2.0 / N * gsR * gsC + 2.0 * s / N * hs - 2.0 * s / (N * N) * (gsR * gNC + gNR * gsC) +
           2.0 * s * s / (N * N * N) * gNR * gNC - s * s / (N * N) * hN_at(E, R, C);
}

// ---------------------------------------------------------------------------------------------
// wave-level helpers (workgroups of exactly one 64-lane wave)
// ---------------------------------------------------------------------------------------------
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// Reciprocal and reciprocal square root: the hardware approximations refined by two Newton steps
// (each squares the relative error), full double precision for the Jacobi rotations.
__device__ __forceinline__ double rcp_nr(double a) {
    double x = __builtin_amdgcn_rcp(a);
    x = fma(x, fma(-a, x, 1.0), x);
    x = fma(x, fma(-a, x, 1.0), x);
    return x;
}
__device__ __forceinline__ double rsq_nr(double a) {
    double y = __builtin_amdgcn_rsq(a);
    y = fma(0.5 * y, fma(-a * y, y, 1.0), y);
    y = fma(0.5 * y, fma(-a * y, y, 1.0), y);
    return y;
}

// Cholesky test of A + delta I, delta = 1e-10 trace(A), for the symmetric m x m matrix A (LDS,
// m <= 12) by one wave, in the scratch L: true when it succeeds, i.e. every eigenvalue of A
// exceeds -delta, so that A differs from its PSD projection by at most delta (positive
// semidefinite matrices with a null space, such as the orthogonality Hessian at a rotation, pass).
__device__ inline bool wave_is_psd(const double* A, double* L, int m, int lane) {
    const int mm = m * m;
    const double tr = wave_sum(lane < m ? A[lane * m + lane] : 0.0);
    const double delta = 1e-10 * fabs(tr);
    for (int e = lane; e < mm; e += 64) L[e] = A[e] + ((e / m == e % m) ? delta : 0.0);
    __syncthreads();
    if (!(tr > 0.0)) return false;
    for (int k = 0; k < m; ++k) {
        const double d = L[k * m + k];
        __syncthreads();  // every lane has read d before any lane writes L again (here or in the caller)
        if (!(d > 0.0)) return false;  // same value in every lane: uniform exit
        const double sd = sqrt(d);
        if (lane > k && lane < m) L[lane * m + k] /= sd;
        __syncthreads();
        for (int e = lane; e < mm; e += 64) {
            const int i = e / m, jj = e % m;
            if (jj > k && i >= jj) L[e] -= L[i * m + k] * L[jj * m + k];
        }
        __syncthreads();
    }
    return true;
}

// Eigen-decomposition S = V diag(w) V^T of the symmetric 3 x 3 S by cyclic Jacobi rotations
// (Golub and Van Loan, Matrix Computations, 8.5.2), one thread; S is overwritten.
__device__ inline void sym3_eig(double S[3][3], double V[3][3], double w[3]) {
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) V[i][j] = (i == j) ? 1.0 : 0.0;
    for (int sweep = 0; sweep < 12; ++sweep) {
        const double off = S[0][1] * S[0][1] + S[0][2] * S[0][2] + S[1][2] * S[1][2];
        const double dg = S[0][0] * S[0][0] + S[1][1] * S[1][1] + S[2][2] * S[2][2];
        if (!(off > 1e-32 * dg)) break;
#pragma unroll
        for (int r = 0; r < 3; ++r) {
            const int p = (r == 2) ? 1 : 0, q = (r == 0) ? 1 : 2;
            const double apq = S[p][q];
            if (apq == 0.0) continue;
            // the rotation that zeroes S[p][q]: t = tan, the root of t^2 + 2 theta t - 1 of smaller size
            const double th = (S[q][q] - S[p][p]) / (2.0 * apq);
            const double t = (th >= 0.0 ? 1.0 : -1.0) / (fabs(th) + sqrt(th * th + 1.0));
            const double c = 1.0 / sqrt(t * t + 1.0), sn = t * c;
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                const double skp = S[k][p], skq = S[k][q];
                S[k][p] = c * skp - sn * skq;
                S[k][q] = sn * skp + c * skq;
            }
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                const double spk = S[p][k], sqk = S[q][k];
                S[p][k] = c * spk - sn * sqk;
                S[q][k] = sn * spk + c * sqk;
            }
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                const double vkp = V[k][p], vkq = V[k][q];
                V[k][p] = c * vkp - sn * vkq;
                V[k][q] = sn * vkp + c * vkq;
            }
        }
    }
    w[0] = S[0][0];
    w[1] = S[1][1];
    w[2] = S[2][2];
}

// The negative eigenpairs of the orthogonality Hessian k4 * h (ortho.py:19-41; k4 = 4 kappa) of the
// rows a_0..a_2 of the 3 x 3 matrix A (a[3 i + r] = A(i, r)), from its analytic eigensystem.
// Psi = kappa sum_ij (a_i.a_j - delta_ij)^2 = kappa sum_k (sigma_k^2 - 1)^2 depends on A only
// through its singular values (A = U diag(sigma) V^T), so its Hessian's eigenvectors are those of
// isotropic energies (Smith, de Goes and Kim, Analytic Eigensystems for Isotropic Distortion
// Energies, 2019): u_k v_k^T with eigenvalue k4 (3 sigma_k^2 - 1), and for i < j
// (u_i v_j^T + u_j v_i^T) / sqrt 2 with k4 (sigma_i^2 + sigma_i sigma_j + sigma_j^2 - 1) and
// (u_i v_j^T - u_j v_i^T) / sqrt 2 with k4 (sigma_i^2 - sigma_i sigma_j + sigma_j^2 - 1).  V and
// sigma^2 from the eigen-decomposition of A^T A, u_k = A v_k / sigma_k.  Writes the eigenvalues
// below zero into lam and their unit eigenvectors (9 entries each, in the order of a) into Q, and
// returns how many; -1 when A is near singular (sigma_min^2 <= 1e-12 sigma_max^2), where u_k is not
// determined accurately.  One thread.
__device__ inline int ortho_negative_modes(const double* a, double k4, double* Q, double* lam) {
    double S[3][3], V[3][3], w[3];
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int c = 0; c < 3; ++c) S[r][c] = a[r] * a[c] + a[3 + r] * a[3 + c] + a[6 + r] * a[6 + c];
    sym3_eig(S, V, w);
    const double wm = fmax(w[0], fmax(w[1], w[2]));
    if (!(w[0] > 1e-12 * wm && w[1] > 1e-12 * wm && w[2] > 1e-12 * wm)) return -1;
    double sg[3], U[3][3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        sg[k] = sqrt(w[k]);
#pragma unroll
        for (int i = 0; i < 3; ++i)
            U[i][k] = (a[3 * i] * V[0][k] + a[3 * i + 1] * V[1][k] + a[3 * i + 2] * V[2][k]) / sg[k];
    }
    int n = 0;
    for (int k = 0; k < 3; ++k) {
        const double l = k4 * (3.0 * w[k] - 1.0);
        if (l < 0.0) {
            lam[n] = l;
            for (int e = 0; e < 9; ++e) Q[9 * n + e] = U[e / 3][k] * V[e % 3][k];
            ++n;
        }
    }
    const double h = 0.70710678118654752440;
    for (int i = 0; i < 3; ++i)
        for (int j = i + 1; j < 3; ++j)
            for (int sgn = 1; sgn >= -1; sgn -= 2) {
                const double l = k4 * (w[i] + sgn * sg[i] * sg[j] + w[j] - 1.0);
                if (l < 0.0) {
                    lam[n] = l;
                    for (int e = 0; e < 9; ++e)
                        Q[9 * n + e] = h * (U[e / 3][i] * V[e % 3][j] + sgn * U[e / 3][j] * V[e % 3][i]);
                    ++n;
                }
            }
    return n;
}

// ---- 16 x 16 symmetric matrices on the matrix cores ---------------------------------------
// One wave holds a symmetric matrix of order <= 16 (zero-padded) in the output layout of
// v_mfma_{f64,f32}_16x16x4: lane l, slot u holds row R(l / 16, u), column l % 16, with
// R(g, u) = g + 4 u for the FP64 instruction and 4 g + u for the FP32 one (both checked on the
// hardware by the projection's test).  The instruction reads A(i, k) from lane i + 16 k and B(k, j)
// from lane j + 16 k.  The product of two such matrices takes four instructions, the kb-th summing
// over k = R(l / 16, kb): slot kb of lane l is X(R(l / 16, kb), l % 16) = X(l % 16, R(l / 16, kb)) by
// symmetry, the instruction's A(l % 16, l / 16), and Y(R(l / 16, kb), l % 16) is its B(l / 16,
// l % 16); the operands come straight from the registers, with no exchange through LDS.  (X Y is
// formed as X^T Y, the same thing for a symmetric X.)
template <typename HT>
struct Sym16;
template <>
struct Sym16<double> {
    typedef double V4 __attribute__((vector_size(4 * sizeof(double))));
    static __device__ __forceinline__ V4 step(double a, double b, V4 c) {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
    static __device__ __forceinline__ int row(int g, int u) { return g + 4 * u; }
};
template <>
struct Sym16<float> {
    typedef float V4 __attribute__((vector_size(4 * sizeof(float))));
    static __device__ __forceinline__ V4 step(float a, float b, V4 c) {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
    static __device__ __forceinline__ int row(int g, int u) { return 4 * g + u; }
};

template <typename HT>
__device__ __forceinline__ typename Sym16<HT>::V4 sym16_mul(typename Sym16<HT>::V4 X, typename Sym16<HT>::V4 Y) {
    typename Sym16<HT>::V4 acc = {0, 0, 0, 0};
#pragma unroll
    for (int kb = 0; kb < 4; ++kb) acc = Sym16<HT>::step(X[kb], Y[kb], acc);
    return acc;
}

// (X + X^T) / 2 of a Sym16 matrix, through the 16 x 17 LDS scratch tr: exactly symmetric (each pair
// of mirrored entries is the same sum).  A product of two commuting symmetric matrices is
// symmetric only up to rounding, and the sign iteration amplifies an asymmetric error along
// eigenvectors of small eigenvalues as it amplifies the eigenvalues themselves.
template <typename HT>
__device__ __forceinline__ typename Sym16<HT>::V4 sym16_symmetrize(typename Sym16<HT>::V4 X, HT* tr, int lane) {
    const int col = lane % 16, g = lane / 16;
#pragma unroll
    for (int u = 0; u < 4; ++u) tr[Sym16<HT>::row(g, u) * 17 + col] = X[u];
    __syncthreads();
    typename Sym16<HT>::V4 Y;
#pragma unroll
    for (int u = 0; u < 4; ++u) Y[u] = HT(0.5) * (X[u] + tr[col * 17 + Sym16<HT>::row(g, u)]);
    __syncthreads();
    return Y;
}

// PSD projection of the symmetric m x m matrix A (LDS, row-major, m <= 12) by one wave (the whole
// wave, uniformly), through the matrix sign function: A+ = (A + A S) / 2 with S = sign(A)
// (A S = |A|), the projection onto the PSD cone that clamping the eigenvalues at zero gives (a
// unique matrix function, so it agrees with the oracle's Jacobi-based psd_project to rounding).
// S comes from Newton-Schulz iterations on X0 = A / ||A||_F (eigenvalues in [-1, 1]; Higham,
// Functions of Matrices, 5.3): quintic steps X <- X (a + b X^2 + c X^4), (a, b, c) = (3.4445,
// -4.7750, 2.0315) (the quintic of Jordan et al.'s Muon optimiser), which multiply small
// eigenvalues by 3.44 per step and keep the others in [0.68, 1.21], then cubic steps
// X <- X (3 - X^2) / 2, which converge quadratically to +-1.  FP64: 24 quintic and 6 cubic steps;
// an eigenvalue below ~1e-14 ||A||_F may end with a partly converged sign, and its contribution to
// A+ is bounded by its own size, so the result is accurate to ~1e-14 of ||A||.  FP32 (the
// mixed-precision assembly): 11 and 5 steps, since eigenvalues below ~1e-6 ||A|| are at the level
// of the FP32 rounding (more quintic steps would amplify the rounding noise of the null space past
// the iteration's basin); accurate to ~1e-6 of ||A||.  The iteration runs in registers on the
// matrix cores (Sym16), each iterate made exactly symmetric; the result is written back from the
// upper triangle, mirrored.  absolute: A <- |A| (eigenvalues mirrored rather than clamped).
template <typename HT>
__device__ inline void wave_psd_project(HT* A, int m, int lane, bool absolute = false) {
    typedef typename Sym16<HT>::V4 V4;
    __shared__ HT tr[16 * 17];
    constexpr int n_quintic = sizeof(HT) == 8 ? 24 : 11, n_cubic = sizeof(HT) == 8 ? 6 : 5;
    const int col = lane % 16, g = lane / 16;
    V4 a;
    double f = 0;
#pragma unroll
    for (int u = 0; u < 4; ++u) {
        const int row = Sym16<HT>::row(g, u);
        a[u] = (row < m && col < m) ? A[row * m + col] : HT(0);
        f += (double)a[u] * (double)a[u];
    }
    const double fro = wave_sum(f);
    if (fro == 0.0 || !(fro < INFINITY)) return;
    const HT inv = (HT)rsq_nr(fro);
    V4 X;
#pragma unroll
    for (int u = 0; u < 4; ++u) X[u] = a[u] * inv;
    for (int it = 0; it < n_quintic + n_cubic; ++it) {
        const V4 X2 = sym16_mul<HT>(X, X);
        V4 T;
        if (it < n_quintic) {
            const V4 X4 = sym16_mul<HT>(X2, X2);
#pragma unroll
            for (int u = 0; u < 4; ++u)
                T[u] = HT(2.0315) * X4[u] - HT(4.7750) * X2[u] + ((Sym16<HT>::row(g, u) == col) ? HT(3.4445) : HT(0));
        } else {
#pragma unroll
            for (int u = 0; u < 4; ++u) T[u] = ((Sym16<HT>::row(g, u) == col) ? HT(1.5) : HT(0)) - HT(0.5) * X2[u];
        }
        X = sym16_symmetrize<HT>(sym16_mul<HT>(X, T), tr, lane);
    }
    const V4 AS = sym16_mul<HT>(a, X);  // |A| = A S
    __syncthreads();  // every lane has read A
#pragma unroll
    for (int u = 0; u < 4; ++u) {
        const int row = Sym16<HT>::row(g, u);
        if (row < m && col < m && row <= col) {
            const HT v = absolute ? AS[u] : HT(0.5) * (a[u] + AS[u]);
            A[row * m + col] = v;
            A[col * m + row] = v;
        }
    }
    __syncthreads();
}

}  // namespace abd
