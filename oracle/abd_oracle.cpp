// abd_oracle.cpp — CPU FP64 oracle of the ABD Newton step with IPC barriers.
//
// TEST INFRASTRUCTURE ONLY.  This file is the checker for the HIP path in colb_amd/csrc: only
// tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load it.  The product never
// links or calls it.
//
// It restates the reference solver `NewtonAbd` (abd.py) and its mixins, plus the parts the
// reference lacks (point-triangle and ground contact, PSD projection, block-Jacobi PCG,
// additive CCD), in plain serial-or-OpenMP C++ with every quantity in double precision.
// Citations (file:line) refer to the reference sources.  Functions whose results are pinned by
// an independent check (closed form, finite differences, brute force, numpy/scipy) are listed
// in tests/test_oracle_*.py; see the header comment of each function for what it follows.
#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <map>
#include <unordered_map>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

#include "../include/abd_hip.h"

namespace {

// ---------------------------------------------------------------------------------------------
// small vector helpers
// ---------------------------------------------------------------------------------------------
struct V3 {
    double x, y, z;
};
inline V3 v3(const double* p) { return {p[0], p[1], p[2]}; }
inline V3 operator+(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline V3 operator-(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline V3 operator*(double s, V3 a) { return {s * a.x, s * a.y, s * a.z}; }
inline double dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline V3 cross(V3 a, V3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
inline double comp(V3 a, int k) { return k == 0 ? a.x : (k == 1 ? a.y : a.z); }
// [w]x as a row-major 3x3
inline void skew(V3 w, double m[9]) {
    m[0] = 0;    m[1] = -w.z; m[2] = w.y;
    m[3] = w.z;  m[4] = 0;    m[5] = -w.x;
    m[6] = -w.y; m[7] = w.x;  m[8] = 0;
}

// ---------------------------------------------------------------------------------------------
// barrier.py:47-66 — cubic barrier in s = d^2: B = -(s - s_hat)^3 for s < s_hat, else 0.
// ---------------------------------------------------------------------------------------------
inline void barrier(double s, double shat, double& B, double& B1, double& B2) {
    if (s < shat) {
        double d = s - shat;
        B = -d * d * d;
        B1 = -3.0 * d * d;
        B2 = -6.0 * d;
    } else {
        B = B1 = B2 = 0.0;
    }
}

// ---------------------------------------------------------------------------------------------
// ortho.py:9-41 — Psi = kappa * sum_ij (a_i.a_j - delta_ij)^2 over the rows a_i of q.
// g9[3i+r] = dPsi/da_i[r];  H81 row-major 9x9 in (i,r),(j,c) order.
// ---------------------------------------------------------------------------------------------
double ortho(const double* q12, double kappa, double* g9, double* H81) {
    const double* a = q12 + 3;
    double G[9];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
            G[3 * i + j] = a[3 * i] * a[3 * j] + a[3 * i + 1] * a[3 * j + 1] + a[3 * i + 2] * a[3 * j + 2];
    double e = 0;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            double t = G[3 * i + j] - (i == j ? 1.0 : 0.0);
            e += t * t;
        }
    if (g9) {
        for (int i = 0; i < 3; ++i)
            for (int r = 0; r < 3; ++r) {
                double s = -a[3 * i + r];
                for (int j = 0; j < 3; ++j) s += G[3 * i + j] * a[3 * j + r];
                g9[3 * i + r] = 4.0 * kappa * s;
            }
    }
    if (H81) {
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j)
                for (int r = 0; r < 3; ++r)
                    for (int c = 0; c < 3; ++c) {
                        double h;
                        if (i == j) {
                            h = a[3 * i + r] * a[3 * i + c] + (r == c ? G[3 * i + i] - 1.0 : 0.0);
                            for (int k = 0; k < 3; ++k) h += a[3 * k + r] * a[3 * k + c];
                        } else {
                            // outer(A[j], A[i]) + diag(A[j].A[i])   (ortho.py:38)
                            h = a[3 * j + r] * a[3 * i + c] + (r == c ? G[3 * i + j] : 0.0);
                        }
                        H81[(3 * i + r) * 9 + 3 * j + c] = 4.0 * kappa * h;
                    }
    }
    return kappa * e;
}

// ---------------------------------------------------------------------------------------------
// Squared distances and their derivatives.  All functions work on a stencil of up to four
// points y[k] and return d^2, the gradient g[3*ns] and Hessian H[(3*ns)^2] w.r.t. the stencil.
// Formulas: IPC (Li et al. 2020) distance types, written as rational functions of triple
// products and differentiated by hand (checked against central differences in the tests).
// ---------------------------------------------------------------------------------------------

// d^2 = s^2/N with s = a.(u x v), N = |u x v|^2 and a, u, v linear combinations of the stencil
// points with coefficients al, mu, nu.  Point-triangle (face) and edge-edge (interior) distances.
double dist2_plane(const V3* y, int ns, const double* al, const double* mu, const double* nu,
                   double* g, double* H) {
    V3 a{0, 0, 0}, u{0, 0, 0}, v{0, 0, 0};
    for (int i = 0; i < ns; ++i) {
        a = a + al[i] * y[i];
        u = u + mu[i] * y[i];
        v = v + nu[i] * y[i];
    }
    V3 n = cross(u, v);
    double s = dot(a, n), N = dot(n, n);
    double d2 = s * s / N;
    if (!g && !H) return d2;
    const int m = 3 * ns;
    double gs[12], gN[12];
    double Dn[4][9];  // dn/dy_i (3x3)
    double Sv[9], Su[9], Sa[9], Sn[9];
    skew(v, Sv); skew(u, Su); skew(a, Sa); skew(n, Sn);
    V3 vxa = cross(v, a), axu = cross(a, u);
    for (int i = 0; i < ns; ++i) {
        V3 gsi = al[i] * n + mu[i] * vxa + nu[i] * axu;
        gs[3 * i] = gsi.x; gs[3 * i + 1] = gsi.y; gs[3 * i + 2] = gsi.z;
        for (int k = 0; k < 9; ++k) Dn[i][k] = -mu[i] * Sv[k] + nu[i] * Su[k];
        for (int c = 0; c < 3; ++c) {  // (Dn_i^T n)_c
            double t = Dn[i][0 * 3 + c] * n.x + Dn[i][1 * 3 + c] * n.y + Dn[i][2 * 3 + c] * n.z;
            gN[3 * i + c] = 2.0 * t;
        }
    }
    if (g)
        for (int k = 0; k < m; ++k) g[k] = 2.0 * s / N * gs[k] - s * s / (N * N) * gN[k];
    if (H) {
        // Hessian of s (trilinear): blocks H_PQ over P,Q in {a,u,v}
        // H_au = -[v]x, H_av = [u]x, H_ua = [v]x, H_uv = -[a]x, H_va = -[u]x, H_vu = [a]x
        for (int i = 0; i < ns; ++i)
            for (int j = 0; j < ns; ++j) {
                double hs[9], hN[9];
                for (int k = 0; k < 9; ++k) {
                    hs[k] = al[i] * mu[j] * (-Sv[k]) + al[i] * nu[j] * Su[k] + mu[i] * al[j] * Sv[k] +
                            mu[i] * nu[j] * (-Sa[k]) + nu[i] * al[j] * (-Su[k]) + nu[i] * mu[j] * Sa[k];
                }
                for (int r = 0; r < 3; ++r)
                    for (int c = 0; c < 3; ++c) {
                        double t = 0;
                        for (int k = 0; k < 3; ++k) t += Dn[i][k * 3 + r] * Dn[j][k * 3 + c];
                        hN[r * 3 + c] = 2.0 * t + 2.0 * (-mu[i] * nu[j] + nu[i] * mu[j]) * Sn[r * 3 + c];
                    }
                for (int r = 0; r < 3; ++r)
                    for (int c = 0; c < 3; ++c) {
                        int R = 3 * i + r, C = 3 * j + c;
                        double h = 2.0 / N * gs[R] * gs[C] + 2.0 * s / N * hs[r * 3 + c] -
                                   2.0 * s / (N * N) * (gs[R] * gN[C] + gN[R] * gs[C]) +
                                   2.0 * s * s / (N * N * N) * gN[R] * gN[C] - s * s / (N * N) * hN[r * 3 + c];
                        H[R * m + C] = h;
                    }
            }
    }
    return d2;
}

// point-edge: d^2 = |(e0-p) x (e1-p)|^2 / |e1-e0|^2, stencil (p, e0, e1)
double dist2_pe(const V3* y, double* g, double* H) {
    const double mu[3] = {-1, 1, 0}, nu[3] = {-1, 0, 1}, om[3] = {0, -1, 1};
    V3 u = y[1] - y[0], v = y[2] - y[0], w = y[2] - y[1];
    V3 n = cross(u, v);
    double N = dot(n, n), L = dot(w, w);
    double d2 = N / L;
    if (!g && !H) return d2;
    double Sv[9], Su[9], Sn[9];
    skew(v, Sv); skew(u, Su); skew(n, Sn);
    double Dn[3][9], gN[9], gL[9];
    for (int i = 0; i < 3; ++i) {
        for (int k = 0; k < 9; ++k) Dn[i][k] = -mu[i] * Sv[k] + nu[i] * Su[k];
        for (int c = 0; c < 3; ++c) {
            gN[3 * i + c] = 2.0 * (Dn[i][c] * n.x + Dn[i][3 + c] * n.y + Dn[i][6 + c] * n.z);
            gL[3 * i + c] = 2.0 * om[i] * comp(w, c);
        }
    }
    if (g)
        for (int k = 0; k < 9; ++k) g[k] = gN[k] / L - N / (L * L) * gL[k];
    if (H) {
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j)
                for (int r = 0; r < 3; ++r)
                    for (int c = 0; c < 3; ++c) {
                        double t = 0;
                        for (int k = 0; k < 3; ++k) t += Dn[i][k * 3 + r] * Dn[j][k * 3 + c];
                        double hN = 2.0 * t + 2.0 * (-mu[i] * nu[j] + nu[i] * mu[j]) * Sn[r * 3 + c];
                        double hL = (r == c) ? 2.0 * om[i] * om[j] : 0.0;
                        int R = 3 * i + r, C = 3 * j + c;
                        H[R * 9 + C] = hN / L - (gN[R] * gL[C] + gL[R] * gN[C]) / (L * L) +
                                       2.0 * N / (L * L * L) * gL[R] * gL[C] - N / (L * L) * hL;
                    }
    }
    return d2;
}

// point-point: stencil (p0, p1)
double dist2_pp(const V3* y, double* g, double* H) {
    V3 d = y[0] - y[1];
    if (g) {
        g[0] = 2 * d.x; g[1] = 2 * d.y; g[2] = 2 * d.z;
        g[3] = -2 * d.x; g[4] = -2 * d.y; g[5] = -2 * d.z;
    }
    if (H) {
        for (int k = 0; k < 36; ++k) H[k] = 0;
        for (int i = 0; i < 2; ++i)
            for (int j = 0; j < 2; ++j)
                for (int r = 0; r < 3; ++r) H[(3 * i + r) * 6 + 3 * j + r] = (i == j) ? 2.0 : -2.0;
    }
    return dot(d, d);
}

// Closest-feature classification.  A pair's four slots are (p, t0, t1, t2) for point-triangle
// and (a0, a1, b0, b1) for edge-edge.  The result lists the slots of the stencil (ns of them)
// and the distance type.

// Point-triangle region (Ericson, Real-Time Collision Detection, section 5.1.5: barycentric
// region tests of the closest point on a triangle).
int classify_pt(const V3* x, int* st, int* ns) {
    V3 p = x[0], A = x[1], B = x[2], C = x[3];
    V3 ab = B - A, ac = C - A, ap = p - A;
    double d1 = dot(ab, ap), d2 = dot(ac, ap);
    if (d1 <= 0 && d2 <= 0) { st[0] = 0; st[1] = 1; *ns = 2; return ABD_D_PP; }
    V3 bp = p - B;
    double d3 = dot(ab, bp), d4 = dot(ac, bp);
    if (d3 >= 0 && d4 <= d3) { st[0] = 0; st[1] = 2; *ns = 2; return ABD_D_PP; }
    double vc = d1 * d4 - d3 * d2;
    if (vc <= 0 && d1 >= 0 && d3 <= 0) { st[0] = 0; st[1] = 1; st[2] = 2; *ns = 3; return ABD_D_PE; }
    V3 cp = p - C;
    double d5 = dot(ab, cp), d6 = dot(ac, cp);
    if (d6 >= 0 && d5 <= d6) { st[0] = 0; st[1] = 3; *ns = 2; return ABD_D_PP; }
    double vb = d5 * d2 - d1 * d6;
    if (vb <= 0 && d2 >= 0 && d6 <= 0) { st[0] = 0; st[1] = 1; st[2] = 3; *ns = 3; return ABD_D_PE; }
    double va = d3 * d6 - d5 * d4;
    if (va <= 0 && (d4 - d3) >= 0 && (d5 - d6) >= 0) { st[0] = 0; st[1] = 2; st[2] = 3; *ns = 3; return ABD_D_PE; }
    st[0] = 0; st[1] = 1; st[2] = 2; st[3] = 3; *ns = 4;
    return ABD_D_PT;
}

// Segment-segment closest parameters (Ericson, section 5.1.9), with a relative parallel test.
// Returns the region: both parameters interior -> EE; one clamped -> PE; both clamped -> PP.
const double kParallelEps = 1e-10;
int classify_ee(const V3* x, int* st, int* ns, double* sp, double* tp) {
    V3 d1 = x[1] - x[0], d2 = x[3] - x[2], r = x[0] - x[2];
    double a = dot(d1, d1), e = dot(d2, d2), f = dot(d2, r);
    double c = dot(d1, r), b = dot(d1, d2);
    double denom = a * e - b * b;
    double s, t;
    if (denom > kParallelEps * a * e) s = std::min(1.0, std::max(0.0, (b * f - c * e) / denom));
    else s = 0.0;
    t = (b * s + f) / e;
    if (t < 0.0) { t = 0.0; s = std::min(1.0, std::max(0.0, -c / a)); }
    else if (t > 1.0) { t = 1.0; s = std::min(1.0, std::max(0.0, (b - c) / a)); }
    if (sp) *sp = s;
    if (tp) *tp = t;
    bool s_end = (s == 0.0 || s == 1.0), t_end = (t == 0.0 || t == 1.0);
    int sa = (s == 1.0) ? 1 : 0, tb = (t == 1.0) ? 3 : 2;
    if (s_end && t_end) { st[0] = sa; st[1] = tb; *ns = 2; return ABD_D_PP; }
    if (s_end) { st[0] = sa; st[1] = 2; st[2] = 3; *ns = 3; return ABD_D_PE; }
    if (t_end) { st[0] = tb; st[1] = 0; st[2] = 1; *ns = 3; return ABD_D_PE; }
    st[0] = 0; st[1] = 1; st[2] = 2; st[3] = 3; *ns = 4;
    return ABD_D_EE;
}

// d^2 of a pair with closest-feature classification; g12 / H144 (may be NULL) are w.r.t. the
// pair's four slots.  kind: ABD_PT or ABD_EE.
double pair_dist2(int kind, const V3* x, double* g12, double* H144, int* dtype) {
    int st[4], ns;
    int type = (kind == ABD_PT) ? classify_pt(x, st, &ns) : classify_ee(x, st, &ns, nullptr, nullptr);
    if (dtype) *dtype = type;
    V3 y[4];
    for (int i = 0; i < ns; ++i) y[i] = x[st[i]];
    double gl[12], Hl[144];
    double* gp = g12 ? gl : nullptr;
    double* Hp = H144 ? Hl : nullptr;
    double d2;
    if (type == ABD_D_PP) d2 = dist2_pp(y, gp, Hp);
    else if (type == ABD_D_PE) d2 = dist2_pe(y, gp, Hp);
    else if (type == ABD_D_PT) {
        const double al[4] = {1, -1, 0, 0}, mu[4] = {0, -1, 1, 0}, nu[4] = {0, -1, 0, 1};
        d2 = dist2_plane(y, 4, al, mu, nu, gp, Hp);
    } else {
        const double al[4] = {1, 0, -1, 0}, mu[4] = {-1, 1, 0, 0}, nu[4] = {0, 0, -1, 1};
        d2 = dist2_plane(y, 4, al, mu, nu, gp, Hp);
    }
    const int m = 3 * ns;
    if (g12) {
        for (int k = 0; k < 12; ++k) g12[k] = 0;
        for (int i = 0; i < ns; ++i)
            for (int r = 0; r < 3; ++r) g12[3 * st[i] + r] = gl[3 * i + r];
    }
    if (H144) {
        for (int k = 0; k < 144; ++k) H144[k] = 0;
        for (int i = 0; i < ns; ++i)
            for (int j = 0; j < ns; ++j)
                for (int r = 0; r < 3; ++r)
                    for (int c = 0; c < 3; ++c)
                        H144[(3 * st[i] + r) * 12 + 3 * st[j] + c] = Hl[(3 * i + r) * m + 3 * j + c];
    }
    return d2;
}

// ---------------------------------------------------------------------------------------------
// PSD projection by the cyclic Jacobi eigenvalue method (Golub & Van Loan, section 8.5):
// H <- V max(Lambda, 0) V^T (absolute: V |Lambda| V^T).  Replaces ipctk_wp's
// x_to_grad_psd_hess_ee projection (abd.py:247), applied to the whole barrier Hessian so that PCG
// sees an SPD system.
// ---------------------------------------------------------------------------------------------
void psd_project(int n, double* H, bool absolute = false) {
    double A[144], V[144];
    std::memcpy(A, H, sizeof(double) * n * n);
    for (int i = 0; i < n * n; ++i) V[i] = 0;
    for (int i = 0; i < n; ++i) V[i * n + i] = 1;
    double fro = 0;
    for (int i = 0; i < n * n; ++i) fro += A[i] * A[i];
    if (fro == 0) return;
    for (int sweep = 0; sweep < 32; ++sweep) {
        double off = 0;
        for (int p = 0; p < n; ++p)
            for (int q = p + 1; q < n; ++q) off += A[p * n + q] * A[p * n + q];
        if (off <= 1e-32 * fro) break;
        for (int p = 0; p < n - 1; ++p)
            for (int q = p + 1; q < n; ++q) {
                double apq = A[p * n + q];
                if (apq == 0.0) continue;
                double app = A[p * n + p], aqq = A[q * n + q];
                double theta = (aqq - app) / (2.0 * apq);
                double t = (theta >= 0 ? 1.0 : -1.0) / (std::fabs(theta) + std::sqrt(theta * theta + 1.0));
                double c = 1.0 / std::sqrt(t * t + 1.0), s = t * c;
                for (int k = 0; k < n; ++k) {
                    double akp = A[k * n + p], akq = A[k * n + q];
                    A[k * n + p] = c * akp - s * akq;
                    A[k * n + q] = s * akp + c * akq;
                }
                for (int k = 0; k < n; ++k) {
                    double apk = A[p * n + k], aqk = A[q * n + k];
                    A[p * n + k] = c * apk - s * aqk;
                    A[q * n + k] = s * apk + c * aqk;
                }
                for (int k = 0; k < n; ++k) {
                    double vkp = V[k * n + p], vkq = V[k * n + q];
                    V[k * n + p] = c * vkp - s * vkq;
                    V[k * n + q] = s * vkp + c * vkq;
                }
            }
    }
    double lam[12];
    for (int i = 0; i < n; ++i) lam[i] = absolute ? std::fabs(A[i * n + i]) : std::max(0.0, A[i * n + i]);
    for (int r = 0; r < n; ++r)
        for (int c = 0; c < n; ++c) {
            double h = 0;
            for (int k = 0; k < n; ++k) h += V[r * n + k] * lam[k] * V[c * n + k];
            H[r * n + c] = h;
        }
}

// ---------------------------------------------------------------------------------------------
// Additive CCD (Li, Kaufman, Jiang, "Codimensional Incremental Potential Contact", 2021,
// Algorithm 1) for the linear trajectory x(t) = x0 + t*(x1 - x0), t in [0, tmax].  Returns a
// conservative time of impact: at the returned t the gap is >= s * (initial gap).  Vertex
// trajectories are exactly linear in the line-search step because x = J q is linear in q.
// ---------------------------------------------------------------------------------------------
// Bound on the change of the pair distance along x0 -> x1: the displacements minus their mean
// (a common translation leaves the distance unchanged), largest point motion per primitive, summed.
double motion_bound(int kind, const V3* x0, const V3* x1, V3* p) {
    V3 mean{0, 0, 0};
    for (int i = 0; i < 4; ++i) { p[i] = x1[i] - x0[i]; mean = mean + 0.25 * p[i]; }
    for (int i = 0; i < 4; ++i) p[i] = p[i] - mean;
    auto nrm = [](V3 a) { return std::sqrt(dot(a, a)); };
    if (kind == ABD_PT) return nrm(p[0]) + std::max(nrm(p[1]), std::max(nrm(p[2]), nrm(p[3])));
    return std::max(nrm(p[0]), nrm(p[1])) + std::max(nrm(p[2]), nrm(p[3]));
}

// Candidate rule: the pair can come closer than `infl` on the path x0 -> x1 (distance at x0
// minus the motion bound, relative slack 1e-9 so rounding never drops a pair active on the path).
bool may_reach(int kind, const V3* x0, const V3* x1, double infl) {
    V3 p[4];
    double lp = motion_bound(kind, x0, x1, p);
    double d2 = pair_dist2(kind, x0, nullptr, nullptr, nullptr);
    double lim = infl * (1.0 + 1e-9);
    return lp == 0.0 ? d2 < lim * lim : std::sqrt(d2) - lp < lim;
}

double accd(int kind, const V3* x0, const V3* x1, double s, double tmax, int max_iter) {
    V3 p[4], x[4];
    double lp = motion_bound(kind, x0, x1, p);
    for (int i = 0; i < 4; ++i) x[i] = x0[i];
    if (lp == 0.0) return tmax;
    double d = std::sqrt(pair_dist2(kind, x, nullptr, nullptr, nullptr));
    double g = s * d;
    double t = 0.0;
    double tl = (1.0 - s) * d / lp;
    for (int it = 0; it < max_iter; ++it) {
        for (int i = 0; i < 4; ++i) x[i] = x[i] + tl * p[i];
        d = std::sqrt(pair_dist2(kind, x, nullptr, nullptr, nullptr));
        if (t > 0.0 && d < g) return t;
        t += tl;
        if (t > tmax) return tmax;
        tl = 0.9 * d / lp;
    }
    return t;  // iteration cap: the t reached so far is still conservative
}

// point vs. plane n.x = h, linear trajectory: gap(t) = d0 - t*v; keep >= s*d0.  Moving away: no
// bound; moving down from on or below the plane: no step at all.
double ccd_ground(double d0, double v, double s, double tmax) {
    if (v <= 0.0) return tmax;
    if (d0 <= 0.0) return 0.0;
    return std::min(tmax, (1.0 - s) * d0 / v);
}

// ---------------------------------------------------------------------------------------------
// Scene context
// ---------------------------------------------------------------------------------------------
struct Pair {
    int kind, i, j;
    bool operator<(const Pair& o) const {
        if (kind != o.kind) return kind < o.kind;
        if (i != o.i) return i < o.i;
        return j < o.j;
    }
};

struct Box {
    double lo[3], hi[3];
};

struct Ctx {
    AbdParams P;
    int nb, nv, ne, nt, ns;
    std::vector<double> rest;  // nv*3
    std::vector<int> vbody, edges, tris, bscene, pinned;
    std::vector<double> mass4;
    std::vector<int> scene_b0, scene_b1;  // body range per scene
    std::vector<std::vector<int>> sc_v, sc_e, sc_t;  // primitives per scene
    std::vector<double> q, qnow, qdot, qtilde;
    // last step log
    std::vector<double> log_res, log_alpha;
    std::vector<int> log_cg;
    std::vector<AbdStepStats> stats;
    std::vector<double> kb;  // contact stiffness per body (adaptive rule of AbdParams.k_max)
    std::vector<double> pred_dq;  // the step predictor's dq (AbdParams.predictor 3), set per step
    mutable std::atomic<int> err{0};  // 64: a vertex on or below the ground plane was assembled
    int threads = 1;
};

inline V3 vpos(const Ctx& C, const double* q, int v) {
    const double* qb = q + 12 * C.vbody[v];
    const double* xr = &C.rest[3 * v];
    return {qb[0] + xr[0] * qb[3] + xr[1] * qb[6] + xr[2] * qb[9],
            qb[1] + xr[0] * qb[4] + xr[1] * qb[7] + xr[2] * qb[10],
            qb[2] + xr[0] * qb[5] + xr[1] * qb[8] + xr[2] * qb[11]};
}

void all_positions(const Ctx& C, const double* q, std::vector<V3>& X) {
    X.resize(C.nv);
    for (int v = 0; v < C.nv; ++v) X[v] = vpos(C, q, v);
}

int scene_of_vert(const Ctx& C, int v) { return C.bscene[C.vbody[v]]; }

// Uniform hash grid over boxes of triangles or edges; queries return candidate ids whose
// boxes overlap the query box.  Used for the exact active set and for swept candidates.
struct Grid {
    double h;
    std::unordered_map<int64_t, std::vector<int>> cells;
    static int64_t key(int64_t ix, int64_t iy, int64_t iz) {
        return ((ix + (1 << 20)) << 42) | ((iy + (1 << 20)) << 21) | (iz + (1 << 20));
    }
    void cell_range(const Box& b, int64_t lo[3], int64_t hi[3]) const {
        for (int k = 0; k < 3; ++k) {
            lo[k] = (int64_t)std::floor(b.lo[k] / h);
            hi[k] = (int64_t)std::floor(b.hi[k] / h);
        }
    }
    void build(const std::vector<Box>& boxes, const std::vector<int>& ids) {
        cells.clear();
        for (int id : ids) {
            int64_t lo[3], hi[3];
            cell_range(boxes[id], lo, hi);
            for (int64_t x = lo[0]; x <= hi[0]; ++x)
                for (int64_t y = lo[1]; y <= hi[1]; ++y)
                    for (int64_t z = lo[2]; z <= hi[2]; ++z) cells[key(x, y, z)].push_back(id);
        }
    }
    void query(const Box& b, std::vector<int>& out) const {
        out.clear();
        int64_t lo[3], hi[3];
        cell_range(b, lo, hi);
        for (int64_t x = lo[0]; x <= hi[0]; ++x)
            for (int64_t y = lo[1]; y <= hi[1]; ++y)
                for (int64_t z = lo[2]; z <= hi[2]; ++z) {
                    auto it = cells.find(key(x, y, z));
                    if (it != cells.end()) out.insert(out.end(), it->second.begin(), it->second.end());
                }
        std::sort(out.begin(), out.end());
        out.erase(std::unique(out.begin(), out.end()), out.end());
    }
};

inline bool overlap(const Box& a, const Box& b) {
    for (int k = 0; k < 3; ++k)
        if (a.lo[k] > b.hi[k] || b.lo[k] > a.hi[k]) return false;
    return true;
}
inline void box_add(Box& b, V3 p) {
    double c[3] = {p.x, p.y, p.z};
    for (int k = 0; k < 3; ++k) {
        b.lo[k] = std::min(b.lo[k], c[k]);
        b.hi[k] = std::max(b.hi[k], c[k]);
    }
}
inline Box box_empty() { return {{1e300, 1e300, 1e300}, {-1e300, -1e300, -1e300}}; }
inline void box_inflate(Box& b, double r) {
    for (int k = 0; k < 3; ++k) { b.lo[k] -= r; b.hi[k] += r; }
}

int pair_scene(const Ctx& C, int kind, int i) {
    return scene_of_vert(C, kind == ABD_EE ? C.edges[2 * i] : i);
}

bool pair_dynamic(const Ctx& C, int b0, int b1) { return !(C.pinned[b0] && C.pinned[b1]); }

// Candidate pairs over the linear path X0 -> X1 (X1 may equal X0): the primitives' swept boxes
// overlap (point box inflated by `infl`, edge boxes by infl/2 each) and may_reach holds.
// Same-body pairs, pairs sharing a vertex (contact.py:158-160) and pairs between two pinned
// bodies are skipped.  EE keeps i < j (contact.py:160).
void broad_phase(const Ctx& C, const std::vector<V3>& X0, const std::vector<V3>& X1, double infl,
                 std::vector<Pair>& out) {
    out.clear();
    std::vector<Box> vbox(C.nv), ebox(C.ne), tbox(C.nt);
    double ext = 0;
    int cnt = 0;
    for (int v = 0; v < C.nv; ++v) {
        Box b = box_empty(); box_add(b, X0[v]); box_add(b, X1[v]);
        vbox[v] = b;
    }
    for (int e = 0; e < C.ne; ++e) {
        Box b = box_empty();
        for (int k = 0; k < 2; ++k) { box_add(b, X0[C.edges[2 * e + k]]); box_add(b, X1[C.edges[2 * e + k]]); }
        box_inflate(b, 0.5 * infl);
        ebox[e] = b;
        for (int k = 0; k < 3; ++k) ext += b.hi[k] - b.lo[k];
        ++cnt;
    }
    for (int t = 0; t < C.nt; ++t) {
        Box b = box_empty();
        for (int k = 0; k < 3; ++k) { box_add(b, X0[C.tris[3 * t + k]]); box_add(b, X1[C.tris[3 * t + k]]); }
        tbox[t] = b;
    }
    for (int v = 0; v < C.nv; ++v) box_inflate(vbox[v], infl);
    Grid grid;
    grid.h = cnt ? std::max(1e-6, 2.0 * ext / (3.0 * cnt)) : 1.0;
    for (int s = 0; s < C.ns; ++s) {
        const std::vector<int>& sv = C.sc_v[s];
        const std::vector<int>& se = C.sc_e[s];
        const std::vector<int>& stt = C.sc_t[s];
        std::vector<std::vector<Pair>> local(sv.size() + se.size());
        // point-triangle
        grid.build(tbox, stt);
#pragma omp parallel for schedule(dynamic, 64) num_threads(C.threads)
        for (long k = 0; k < (long)sv.size(); ++k) {
            std::vector<int> cand;
            int v = sv[k];
            grid.query(vbox[v], cand);
            for (int t : cand) {
                int bt = C.vbody[C.tris[3 * t]], bv = C.vbody[v];
                if (bt == bv || !pair_dynamic(C, bv, bt)) continue;
                if (C.tris[3 * t] == v || C.tris[3 * t + 1] == v || C.tris[3 * t + 2] == v) continue;
                if (!overlap(vbox[v], tbox[t])) continue;
                const int* tv = &C.tris[3 * t];
                V3 a[4] = {X0[v], X0[tv[0]], X0[tv[1]], X0[tv[2]]};
                V3 b[4] = {X1[v], X1[tv[0]], X1[tv[1]], X1[tv[2]]};
                if (!may_reach(ABD_PT, a, b, infl)) continue;
                local[k].push_back({ABD_PT, v, t});
            }
        }
        // edge-edge
        grid.build(ebox, se);
#pragma omp parallel for schedule(dynamic, 64) num_threads(C.threads)
        for (long k = 0; k < (long)se.size(); ++k) {
            std::vector<int> cand;
            int e = se[k];
            grid.query(ebox[e], cand);
            int a0 = C.edges[2 * e], a1 = C.edges[2 * e + 1];
            for (int f : cand) {
                if (f <= e) continue;
                int b0 = C.edges[2 * f], b1 = C.edges[2 * f + 1];
                if (C.vbody[a0] == C.vbody[b0] || !pair_dynamic(C, C.vbody[a0], C.vbody[b0])) continue;
                if (a0 == b0 || a0 == b1 || a1 == b0 || a1 == b1) continue;
                if (!overlap(ebox[e], ebox[f])) continue;
                V3 a[4] = {X0[a0], X0[a1], X0[b0], X0[b1]};
                V3 b[4] = {X1[a0], X1[a1], X1[b0], X1[b1]};
                if (!may_reach(ABD_EE, a, b, infl)) continue;
                local[sv.size() + k].push_back({ABD_EE, e, f});
            }
        }
        for (auto& l : local) out.insert(out.end(), l.begin(), l.end());
        // ground
        if (C.P.ground_enabled) {
            V3 n = v3(C.P.ground_normal);
            for (int v : sv) {
                if (C.pinned[C.vbody[v]]) continue;
                double d0 = dot(n, X0[v]) - C.P.ground_offset, d1 = dot(n, X1[v]) - C.P.ground_offset;
                if (std::min(d0, d1) < infl) out.push_back({ABD_PG, v, 0});
            }
        }
    }
    std::sort(out.begin(), out.end());
}

void pair_points(const Ctx& C, const Pair& pr, const std::vector<V3>& X, V3* x, int* vid) {
    if (pr.kind == ABD_PT) {
        vid[0] = pr.i; vid[1] = C.tris[3 * pr.j]; vid[2] = C.tris[3 * pr.j + 1]; vid[3] = C.tris[3 * pr.j + 2];
    } else if (pr.kind == ABD_EE) {
        vid[0] = C.edges[2 * pr.i]; vid[1] = C.edges[2 * pr.i + 1];
        vid[2] = C.edges[2 * pr.j]; vid[3] = C.edges[2 * pr.j + 1];
    } else {
        vid[0] = pr.i; vid[1] = vid[2] = vid[3] = -1;
    }
    for (int k = 0; k < 4; ++k) x[k] = vid[k] >= 0 ? X[vid[k]] : V3{0, 0, 0};
}

// d^2 of a candidate at positions X; ground: signed distance squared
double cand_dist2(const Ctx& C, const Pair& pr, const std::vector<V3>& X, int* dtype) {
    V3 x[4]; int vid[4];
    pair_points(C, pr, X, x, vid);
    if (pr.kind == ABD_PG) {
        double d = dot(v3(C.P.ground_normal), x[0]) - C.P.ground_offset;
        if (dtype) *dtype = ABD_D_PG;
        return d * d;
    }
    return pair_dist2(pr.kind, x, nullptr, nullptr, dtype);
}

// bodies of a pair (b1 = -1 for ground)
void pair_bodies(const Ctx& C, const Pair& pr, int* b0, int* b1) {
    if (pr.kind == ABD_PT) { *b0 = C.vbody[pr.i]; *b1 = C.vbody[C.tris[3 * pr.j]]; }
    else if (pr.kind == ABD_EE) { *b0 = C.vbody[C.edges[2 * pr.i]]; *b1 = C.vbody[C.edges[2 * pr.j]]; }
    else { *b0 = C.vbody[pr.i]; *b1 = -1; }
}

// contact stiffness of a pair: the larger of its bodies' stiffness (abd.py:30 when fixed)
double pair_stiffness(const Ctx& C, const Pair& pr) {
    int b0, b1;
    pair_bodies(C, pr, &b0, &b1);
    return b1 < 0 ? C.kb[b0] : std::max(C.kb[b0], C.kb[b1]);
}

// energy_inertia (abd.py:53-64): 1/2 |q - q~|_M^2 + dt^2 Psi(q), pinned bodies excluded.
double body_energy(const Ctx& C, const double* q, int b) {
    if (C.pinned[b]) return 0.0;
    const double* qb = q + 12 * b;
    const double* qt = &C.qtilde[12 * b];
    const double* M = &C.mass4[16 * b];
    double e = 0;
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j) {
            double d = 0;
            for (int r = 0; r < 3; ++r) d += (qb[3 * i + r] - qt[3 * i + r]) * (qb[3 * j + r] - qt[3 * j + r]);
            e += M[4 * i + j] * d;
        }
    double dt = C.P.dt;
    return 0.5 * e + dt * dt * ortho(qb, C.P.kappa, nullptr, nullptr);
}

// energy_contact (abd.py:66-75): dt^2 k B(d^2) over active pairs.
double contact_energy(const Ctx& C, const Pair& pr, const std::vector<V3>& X) {
    double d2 = cand_dist2(C, pr, X, nullptr);
    if (pr.kind == ABD_PG && !(dot(v3(C.P.ground_normal), X[pr.i]) - C.P.ground_offset > 0.0))
        return HUGE_VAL;  // on or below the plane: infeasible
    double B, B1, B2;
    barrier(d2, C.P.dhat * C.P.dhat, B, B1, B2);
    return C.P.dt * C.P.dt * pair_stiffness(C, pr) * B;
}

// Incremental potential per scene at q, over the active pairs at q (compute_g, abd.py:725-735:
// contacts re-detected at the evaluated state).
void scene_energy(const Ctx& C, const double* q, double* E) {
    std::vector<V3> X;
    all_positions(C, q, X);
    std::vector<Pair> pairs;
    broad_phase(C, X, X, C.P.dhat, pairs);
    for (int s = 0; s < C.ns; ++s) E[s] = 0;
    for (int b = 0; b < C.nb; ++b) E[C.bscene[b]] += body_energy(C, q, b);
    for (auto& pr : pairs) E[pair_scene(C, pr.kind, pr.i)] += contact_energy(C, pr, X);
}

void active_pairs(const Ctx& C, const double* q, std::vector<Pair>& act, std::vector<double>* d2out) {
    std::vector<V3> X;
    all_positions(C, q, X);
    std::vector<Pair> cand;
    broad_phase(C, X, X, C.P.dhat, cand);
    act.clear();
    if (d2out) d2out->clear();
    double dh2 = C.P.dhat * C.P.dhat;
    for (auto& pr : cand) {
        double d2 = cand_dist2(C, pr, X, nullptr);
        if (d2 < dh2) {
            act.push_back(pr);
            if (d2out) d2out->push_back(d2);
        }
    }
}

// Block-sparse system: diagonal 12x12 per body + upper off-diagonal blocks keyed (b0 < b1).
struct System {
    std::vector<double> grad, diag;
    std::map<std::pair<int, int>, std::vector<double>> off;
};

// inertia_grad_hess (abd.py:108-149) generalised to a 4x4 mass matrix; ortho Hessian is
// PSD-projected (9x9) so that the block is SPD.
void assemble_body(const Ctx& C, const double* q, int b, double* g, double* H) {
    for (int k = 0; k < 12; ++k) g[k] = 0;
    for (int k = 0; k < 144; ++k) H[k] = 0;
    if (C.pinned[b]) {
        for (int k = 0; k < 12; ++k) H[k * 13] = C.P.pinned_diag;
        return;
    }
    const double* qb = q + 12 * b;
    const double* qt = &C.qtilde[12 * b];
    const double* M = &C.mass4[16 * b];
    double dt2 = C.P.dt * C.P.dt;
    double g9[9], H81[81];
    ortho(qb, C.P.kappa, g9, H81);
    psd_project(9, H81);
    for (int i = 0; i < 4; ++i)
        for (int r = 0; r < 3; ++r) {
            double s = 0;
            for (int j = 0; j < 4; ++j) s += M[4 * i + j] * (qb[3 * j + r] - qt[3 * j + r]);
            g[3 * i + r] = s + (i > 0 ? dt2 * g9[3 * (i - 1) + r] : 0.0);
        }
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j)
            for (int r = 0; r < 3; ++r) H[(3 * i + r) * 12 + 3 * j + r] += M[4 * i + j];
    for (int R = 0; R < 9; ++R)
        for (int Cc = 0; Cc < 9; ++Cc) H[(R + 3) * 12 + Cc + 3] += dt2 * H81[R * 9 + Cc];
}

// contact_hessian_ee (abd.py:198-571) generalised to PT/EE/PG with the whole barrier Hessian
// PSD-projected in x-space; J = [1, xr^T] (x) I3 per vertex (abd.py:256-263).
struct ContactTerm {
    int b[2];          // bodies (b[1] = -1 for ground)
    double g[2][12];   // body gradients
    double H[3][144];  // (b0,b0), (b1,b1), (b0,b1)
};

void contact_term(const Ctx& C, const Pair& pr, const std::vector<V3>& X, ContactTerm& T) {
    V3 x[4]; int vid[4];
    pair_points(C, pr, X, x, vid);
    double gx[12], Hx[144];
    double d2;
    int nvp = 4;
    if (pr.kind == ABD_PG) {
        V3 n = v3(C.P.ground_normal);
        double d = dot(n, x[0]) - C.P.ground_offset;
        if (!(d > 0.0)) {
            // on or below the plane: an infeasible state (reported), no contribution
            C.err |= 64;
            d = 0.0;
        }
        d2 = d * d;
        for (int k = 0; k < 12; ++k) gx[k] = 0;
        for (int k = 0; k < 144; ++k) Hx[k] = 0;
        for (int r = 0; r < 3; ++r) {
            gx[r] = 2 * d * comp(n, r);
            for (int c = 0; c < 3; ++c) Hx[r * 12 + c] = 2 * comp(n, r) * comp(n, c);
        }
        nvp = 1;
    } else {
        d2 = pair_dist2(pr.kind, x, gx, Hx, nullptr);
    }
    double B, B1, B2;
    barrier(d2, C.P.dhat * C.P.dhat, B, B1, B2);
    double kd = pair_stiffness(C, pr) * C.P.dt * C.P.dt;
    const bool absolute = C.P.hess_proj == 1;
    if (nvp == 1) {
        // point-plane: H = kd (B1 * 2 n n^T + B2 (2 d n)(2 d n)^T) = kd (2 B1 + 4 B2 d^2) n n^T
        // (unit n); its PSD projection clamps the single eigenvalue at zero (or mirrors it).
        V3 n = v3(C.P.ground_normal);
        const double lam0 = kd * (2.0 * B1 + 4.0 * B2 * d2);
        double lam = absolute ? std::fabs(lam0) : std::max(0.0, lam0);
        for (int r = 0; r < 3; ++r)
            for (int c = 0; c < 3; ++c) Hx[r * 12 + c] = lam * comp(n, r) * comp(n, c);
    } else {
        for (int R = 0; R < 12; ++R)
            for (int Cc = 0; Cc < 12; ++Cc) Hx[R * 12 + Cc] = kd * (B1 * Hx[R * 12 + Cc] + B2 * gx[R] * gx[Cc]);
        psd_project(12, Hx, absolute);
    }
    for (int k = 0; k < 12; ++k) gx[k] *= kd * B1;
    int bod[4];
    double jv[4][4];
    for (int k = 0; k < nvp; ++k) {
        bod[k] = C.vbody[vid[k]];
        jv[k][0] = 1.0;
        for (int r = 0; r < 3; ++r) jv[k][1 + r] = C.rest[3 * vid[k] + r];
    }
    T.b[0] = bod[0];
    T.b[1] = (nvp == 1) ? -1 : bod[nvp - 1];
    std::memset(T.g, 0, sizeof(T.g));
    std::memset(T.H, 0, sizeof(T.H));
    auto side = [&](int k) { return (nvp > 1 && bod[k] != T.b[0]) ? 1 : 0; };
    for (int k = 0; k < nvp; ++k) {
        int sk = side(k);
        for (int i = 0; i < 4; ++i)
            for (int r = 0; r < 3; ++r) T.g[sk][3 * i + r] += jv[k][i] * gx[3 * k + r];
    }
    for (int k = 0; k < nvp; ++k)
        for (int l = 0; l < nvp; ++l) {
            int sk = side(k), sl = side(l);
            int blk;
            bool tr = false;
            if (sk == 0 && sl == 0) blk = 0;
            else if (sk == 1 && sl == 1) blk = 1;
            else if (sk == 0 && sl == 1) blk = 2;
            else { blk = 2; tr = true; }
            if (tr) continue;  // (b1,b0) is the transpose of (b0,b1)
            for (int i = 0; i < 4; ++i)
                for (int j = 0; j < 4; ++j) {
                    double w = jv[k][i] * jv[l][j];
                    for (int r = 0; r < 3; ++r)
                        for (int c = 0; c < 3; ++c)
                            T.H[blk][(3 * i + r) * 12 + 3 * j + c] += w * Hx[(3 * k + r) * 12 + 3 * l + c];
                }
        }
}

void assemble(const Ctx& C, const double* q, const std::vector<Pair>& act, System& S) {
    std::vector<V3> X;
    all_positions(C, q, X);
    S.grad.assign(12 * C.nb, 0.0);
    S.diag.assign(144 * C.nb, 0.0);
    S.off.clear();
#pragma omp parallel for num_threads(C.threads)
    for (int b = 0; b < C.nb; ++b) assemble_body(C, q, b, &S.grad[12 * b], &S.diag[144 * b]);
    std::vector<ContactTerm> terms(act.size());
#pragma omp parallel for schedule(dynamic, 16) num_threads(C.threads)
    for (long c = 0; c < (long)act.size(); ++c) contact_term(C, act[c], X, terms[c]);
    for (auto& T : terms) {
        int b0 = T.b[0], b1 = T.b[1];
        bool d0 = !C.pinned[b0], d1 = (b1 >= 0) && !C.pinned[b1];
        if (d0) {
            for (int k = 0; k < 12; ++k) S.grad[12 * b0 + k] += T.g[0][k];
            for (int k = 0; k < 144; ++k) S.diag[144 * b0 + k] += T.H[0][k];
        }
        if (d1) {
            for (int k = 0; k < 12; ++k) S.grad[12 * b1 + k] += T.g[1][k];
            for (int k = 0; k < 144; ++k) S.diag[144 * b1 + k] += T.H[1][k];
        }
        if (d0 && d1) {
            if (b0 < b1) {
                auto& blk = S.off[{b0, b1}];
                if (blk.empty()) blk.assign(144, 0.0);
                for (int k = 0; k < 144; ++k) blk[k] += T.H[2][k];
            } else {
                auto& blk = S.off[{b1, b0}];
                if (blk.empty()) blk.assign(144, 0.0);
                for (int r = 0; r < 12; ++r)
                    for (int c = 0; c < 12; ++c) blk[r * 12 + c] += T.H[2][c * 12 + r];
            }
        }
    }
}

// 12x12 Cholesky inverse for the block-Jacobi preconditioner
bool chol_inverse12(const double* A, double* Ainv) {
    double L[144] = {0};
    for (int j = 0; j < 12; ++j) {
        double s = A[j * 12 + j];
        for (int k = 0; k < j; ++k) s -= L[j * 12 + k] * L[j * 12 + k];
        if (s <= 0) return false;
        L[j * 12 + j] = std::sqrt(s);
        for (int i = j + 1; i < 12; ++i) {
            double t = A[i * 12 + j];
            for (int k = 0; k < j; ++k) t -= L[i * 12 + k] * L[j * 12 + k];
            L[i * 12 + j] = t / L[j * 12 + j];
        }
    }
    for (int col = 0; col < 12; ++col) {
        double y[12], x[12];
        for (int i = 0; i < 12; ++i) {
            double t = (i == col) ? 1.0 : 0.0;
            for (int k = 0; k < i; ++k) t -= L[i * 12 + k] * y[k];
            y[i] = t / L[i * 12 + i];
        }
        for (int i = 11; i >= 0; --i) {
            double t = y[i];
            for (int k = i + 1; k < 12; ++k) t -= L[k * 12 + i] * x[k];
            x[i] = t / L[i * 12 + i];
        }
        for (int i = 0; i < 12; ++i) Ainv[i * 12 + col] = x[i];
    }
    return true;
}

void spmv(const Ctx& C, const System& S, const std::vector<double>& x, std::vector<double>& y) {
    y.assign(12 * C.nb, 0.0);
    for (int b = 0; b < C.nb; ++b)
        for (int r = 0; r < 12; ++r) {
            double s = 0;
            for (int c = 0; c < 12; ++c) s += S.diag[144 * b + r * 12 + c] * x[12 * b + c];
            y[12 * b + r] += s;
        }
    for (auto& kv : S.off) {
        int i = kv.first.first, j = kv.first.second;
        const double* B = kv.second.data();
        for (int r = 0; r < 12; ++r) {
            double s = 0, t = 0;
            for (int c = 0; c < 12; ++c) {
                s += B[r * 12 + c] * x[12 * j + c];
                t += B[c * 12 + r] * x[12 * i + c];
            }
            y[12 * i + r] += s;
            y[12 * j + r] += t;
        }
    }
}

// Block-Jacobi PCG per scene (replaces warp.optim.linear.bicgstab, abd.py:636).  Starts from
// du = 0 (abd.py:632), or with x0 (AbdParams.pcg_x0) from x0 in every scene where the residual
// grad - H x0 is smaller than grad (the stopping test stays relative to ||grad||).  Returns
// per-scene iteration counts.
void pcg(const Ctx& C, const System& S, std::vector<double>& du, std::vector<int>& iters,
         const std::vector<double>* x0in = nullptr) {
    int n = 12 * C.nb;
    std::vector<double> Pinv(144 * C.nb);
    for (int b = 0; b < C.nb; ++b) chol_inverse12(&S.diag[144 * b], &Pinv[144 * b]);
    du.assign(n, 0.0);
    std::vector<double> r = S.grad, z(n), p(n), Ap;
    std::vector<double> x0, bref(C.ns, -1.0);
    if (x0in) {
        x0 = *x0in;
        spmv(C, S, x0, Ap);
        for (int s = 0; s < C.ns; ++s) {
            double nb2 = 0, nr2 = 0;
            for (int i = 12 * C.scene_b0[s]; i < 12 * C.scene_b1[s]; ++i) {
                nb2 += S.grad[i] * S.grad[i];
                const double ri = S.grad[i] - Ap[i];
                nr2 += ri * ri;
            }
            bref[s] = std::sqrt(nb2);
            const bool keep = std::sqrt(nr2) < bref[s];
            for (int i = 12 * C.scene_b0[s]; i < 12 * C.scene_b1[s]; ++i) {
                if (keep) r[i] = S.grad[i] - Ap[i];
                else x0[i] = 0.0;
            }
        }
    }
    auto prec = [&](const std::vector<double>& in, std::vector<double>& out) {
        for (int b = 0; b < C.nb; ++b)
            for (int i = 0; i < 12; ++i) {
                double s = 0;
                for (int k = 0; k < 12; ++k) s += Pinv[144 * b + i * 12 + k] * in[12 * b + k];
                out[12 * b + i] = s;
            }
    };
    auto sdot = [&](const std::vector<double>& a, const std::vector<double>& b, int s) {
        double t = 0;
        for (int i = 12 * C.scene_b0[s]; i < 12 * C.scene_b1[s]; ++i) t += a[i] * b[i];
        return t;
    };
    prec(r, z);
    p = z;
    std::vector<double> rz(C.ns), bn(C.ns);
    std::vector<char> done(C.ns, 0);
    iters.assign(C.ns, 0);
    for (int s = 0; s < C.ns; ++s) {
        rz[s] = sdot(r, z, s);
        bn[s] = bref[s] >= 0.0 ? bref[s] : std::sqrt(sdot(r, r, s));
        if (bn[s] == 0) done[s] = 1;
    }
    for (int it = 0; it < C.P.cg_max_iter; ++it) {
        bool any = false;
        for (int s = 0; s < C.ns; ++s) {
            if (done[s]) continue;
            if (std::sqrt(sdot(r, r, s)) <= C.P.cg_tol * bn[s]) done[s] = 1;
            else any = true;
        }
        if (!any) break;
        spmv(C, S, p, Ap);
        for (int s = 0; s < C.ns; ++s) {
            if (done[s]) continue;
            double pAp = sdot(p, Ap, s);
            double alpha = rz[s] / pAp;
            for (int i = 12 * C.scene_b0[s]; i < 12 * C.scene_b1[s]; ++i) {
                du[i] += alpha * p[i];
                r[i] -= alpha * Ap[i];
            }
        }
        prec(r, z);
        for (int s = 0; s < C.ns; ++s) {
            if (done[s]) continue;
            double rzn = sdot(r, z, s);
            double beta = rzn / rz[s];
            rz[s] = rzn;
            for (int i = 12 * C.scene_b0[s]; i < 12 * C.scene_b1[s]; ++i) p[i] = z[i] + beta * p[i];
            iters[s] = it + 1;
        }
    }
    if (x0in)
        for (int i = 0; i < n; ++i) du[i] += x0[i];
}

// Per-scene CCD bound of q -> q - du over swept candidates (inflated by dhat).
void ccd_bound(const Ctx& C, const double* q, const std::vector<double>& du, std::vector<double>& amax) {
    std::vector<double> q1(12 * C.nb);
    for (int i = 0; i < 12 * C.nb; ++i) q1[i] = q[i] - du[i];
    std::vector<V3> X0, X1;
    all_positions(C, q, X0);
    all_positions(C, q1.data(), X1);
    std::vector<Pair> cand;
    broad_phase(C, X0, X1, C.P.dhat, cand);
    amax.assign(C.ns, 1.0);
    std::vector<double> toi(cand.size(), 1.0);
#pragma omp parallel for schedule(dynamic, 64) num_threads(C.threads)
    for (long c = 0; c < (long)cand.size(); ++c) {
        const Pair& pr = cand[c];
        V3 a[4], b[4]; int vid[4];
        pair_points(C, pr, X0, a, vid);
        pair_points(C, pr, X1, b, vid);
        if (pr.kind == ABD_PG) {
            V3 n = v3(C.P.ground_normal);
            double d0 = dot(n, a[0]) - C.P.ground_offset;
            double v = -(dot(n, b[0]) - dot(n, a[0]));
            toi[c] = ccd_ground(d0, v, C.P.ccd_gap_ratio, 1.0);
        } else {
            toi[c] = accd(pr.kind, a, b, C.P.ccd_gap_ratio, 1.0, C.P.ccd_max_iter);
        }
    }
    for (size_t c = 0; c < cand.size(); ++c) {
        int s = pair_scene(C, cand[c].kind, cand[c].i);
        amax[s] = std::min(amax[s], toi[c]);
    }
}

void compute_qtilde(Ctx& C) {
    double dt = C.P.dt;
    C.qtilde.resize(12 * C.nb);
    for (int b = 0; b < C.nb; ++b)
        for (int k = 0; k < 12; ++k) {
            double v = C.qnow[12 * b + k] + dt * C.qdot[12 * b + k];
            if (k < 3) v += dt * dt * C.P.gravity[k];  // tildep (abd.py:164-166)
            C.qtilde[12 * b + k] = v;
        }
}

// line_search_iterative (gauss_newton.py:98-115) per scene along q -> q - alpha du from the CCD
// bound: halve until the energy decreases.  At the floor without a decrease the Newton search keeps
// the floor step (gauss_newton.py:110-111) unless its energy is not finite; the predictor's search
// (floor_accept false) takes no step.  Applies the step to the iterating scenes.
void line_search(Ctx& C, const std::vector<double>& du, const std::vector<char>& active, bool floor_accept,
                 std::vector<double>& alpha, std::vector<double>& amax, std::vector<int>& floor_hit) {
    amax.assign(C.ns, 1.0);
    if (C.P.ccd_enabled) ccd_bound(C, C.q.data(), du, amax);
    std::vector<double> E0(C.ns), E1(C.ns);
    scene_energy(C, C.q.data(), E0.data());
    std::vector<double> a0(C.ns);
    alpha.assign(C.ns, 0.0);
    floor_hit.assign(C.ns, 0);
    std::vector<char> ls_done(C.ns, 0);
    for (int s = 0; s < C.ns; ++s) {
        a0[s] = alpha[s] = amax[s];
        if (!active[s]) { ls_done[s] = 1; alpha[s] = 0.0; }
    }
    std::vector<double> qt(C.q.size());
    while (true) {
        bool pending = false;
        for (int s = 0; s < C.ns; ++s) pending |= !ls_done[s];
        if (!pending) break;
        for (int b = 0; b < C.nb; ++b) {
            int s = C.bscene[b];
            double al = ls_done[s] ? 0.0 : alpha[s];
            for (int k = 0; k < 12; ++k) qt[12 * b + k] = C.q[12 * b + k] - al * du[12 * b + k];
        }
        scene_energy(C, qt.data(), E1.data());
        for (int s = 0; s < C.ns; ++s) {
            if (ls_done[s]) continue;
            C.stats[s].energy = E1[s];
            if (E1[s] < E0[s]) ls_done[s] = 1;
            else if (alpha[s] * 0.5 < C.P.ls_floor * a0[s]) {
                ls_done[s] = 1;
                floor_hit[s] = 1;
                if (!floor_accept || !std::isfinite(E1[s])) {
                    alpha[s] = 0.0;  // no step: the accepted iterate is q itself
                    C.stats[s].energy = E0[s];
                }
            } else alpha[s] *= 0.5;
        }
    }
    for (int b = 0; b < C.nb; ++b) {
        int s = C.bscene[b];
        if (!active[s]) continue;
        for (int k = 0; k < 12; ++k) C.q[12 * b + k] -= alpha[s] * du[12 * b + k];
    }
}

// Adaptive contact stiffness per body (the rule of AbdParams.k_max): a body of an iterating scene
// doubles its stiffness, up to k_max, when its closest contact at the accepted iterate is nearer
// than k_raise_gap * dhat and nearer than after its previous iteration of the step.  Contacts at or
// beyond dhat never trigger it, so the distances over the active set decide as over any superset.
void adapt_stiffness(Ctx& C, const std::vector<char>& active, std::vector<double>& dmin_prev) {
    if (!(C.P.k_max > C.P.k_contact)) return;
    std::vector<Pair> act;
    std::vector<double> d2;
    active_pairs(C, C.q.data(), act, &d2);
    std::vector<double> dmin(C.nb, HUGE_VAL);
    for (size_t c = 0; c < act.size(); ++c) {
        int b0, b1;
        pair_bodies(C, act[c], &b0, &b1);
        const double d = std::sqrt(d2[c]);
        dmin[b0] = std::min(dmin[b0], d);
        if (b1 >= 0) dmin[b1] = std::min(dmin[b1], d);
    }
    for (int b = 0; b < C.nb; ++b) {
        const int s = C.bscene[b];
        if (!active[s]) continue;
        if (C.P.k_max > C.kb[b] && dmin[b] < C.P.k_raise_gap * C.P.dhat && dmin[b] < dmin_prev[b]) {
            C.kb[b] = std::min(2.0 * C.kb[b], C.P.k_max);
            C.stats[s].k_raises++;
        }
        dmin_prev[b] = dmin[b];
    }
}

// NewtonAbd.step (abd.py:674-702): detect, assemble, solve, line search, until the residual of
// the pre-update iterate is below tol or max_newton solves; then forward_states.  With a predictor
// (AbdParams.predictor 1, 2) the first iterate moves from q_now toward q_now + dt qdot or q~ first.
void step_once(Ctx& C) {
    compute_qtilde(C);
    const int mi = C.P.max_newton;
    C.log_res.assign(C.ns * mi, 0.0);
    C.log_alpha.assign(C.ns * mi, 0.0);
    C.log_cg.assign(C.ns * mi, 0);
    C.stats.assign(C.ns, AbdStepStats{});
    for (auto& st : C.stats) { st.alpha_min = 1.0; st.ccd_alpha_min = 1.0; }
    std::vector<char> active(C.ns, 1);
    std::vector<double> dmin_prev(C.nb, HUGE_VAL);
    std::vector<double> alpha, amax, du_prev;
    std::vector<int> floor_hit;
    if (C.P.predictor >= 1 && C.P.predictor <= 3) {
        std::vector<double> dup(12 * C.nb, 0.0);
        const bool learned = C.P.predictor == 3 && C.pred_dq.size() == 12 * (size_t)C.nb;
        if (C.P.predictor == 3 && !learned) C.err |= 128;  // no prediction set
        for (int b = 0; b < C.nb; ++b)
            if (!C.pinned[b])
                for (int k = 0; k < 12; ++k) {
                    const int i = 12 * b + k;
                    const double target = C.P.predictor == 2 ? C.qtilde[i]
                                          : learned           ? C.qnow[i] + C.pred_dq[i]
                                                              : C.qnow[i] + C.P.dt * C.qdot[i];
                    const double d = C.q[i] - target;
                    dup[i] = std::isfinite(d) ? d : 0.0;
                }
        line_search(C, dup, active, false, alpha, amax, floor_hit);
        for (int s = 0; s < C.ns; ++s) C.stats[s].predictor_alpha = alpha[s];
    }
    for (int it = 0; it < mi; ++it) {
        bool any = false;
        for (int s = 0; s < C.ns; ++s) any |= active[s] != 0;
        if (!any) break;
        std::vector<Pair> act;
        active_pairs(C, C.q.data(), act, nullptr);
        System S;
        assemble(C, C.q.data(), act, S);
        std::vector<double> res(C.ns, 0.0);
        for (int s = 0; s < C.ns; ++s) {
            double t = 0;
            for (int i = 12 * C.scene_b0[s]; i < 12 * C.scene_b1[s]; ++i) t += S.grad[i] * S.grad[i];
            res[s] = std::sqrt(t);
        }
        for (int s = 0; s < C.ns; ++s) C.stats[s].n_active = 0;
        for (auto& pr : act) C.stats[pair_scene(C, pr.kind, pr.i)].n_active++;
        // frozen scenes keep their state: zero their rhs so du = 0 there
        for (int s = 0; s < C.ns; ++s)
            if (!active[s])
                for (int i = 12 * C.scene_b0[s]; i < 12 * C.scene_b1[s]; ++i) S.grad[i] = 0;
        std::vector<double> du;
        std::vector<int> cgit;
        if (C.P.pcg_x0 == 1 && it > 0) {
            // the part of the previous direction its line search did not take
            std::vector<double> x0(12 * C.nb, 0.0);
            for (int b = 0; b < C.nb; ++b) {
                if (C.pinned[b]) continue;
                const double a = alpha[C.bscene[b]];
                for (int k = 0; k < 12; ++k) {
                    const double v = (1.0 - a) * du_prev[12 * b + k];
                    x0[12 * b + k] = std::isfinite(v) ? v : 0.0;
                }
            }
            pcg(C, S, du, cgit, &x0);
        } else {
            pcg(C, S, du, cgit);
        }
        line_search(C, du, active, true, alpha, amax, floor_hit);
        du_prev = du;
        adapt_stiffness(C, active, dmin_prev);
        for (int s = 0; s < C.ns; ++s) {
            if (!active[s]) continue;
            AbdStepStats& st = C.stats[s];
            st.ls_floor_hits += floor_hit[s];
            C.log_res[s * mi + it] = res[s];
            C.log_alpha[s * mi + it] = alpha[s];
            C.log_cg[s * mi + it] = cgit[s];
            if (it == 0) st.residual_first = res[s];
            st.residual_last = res[s];
            st.newton_iters = it + 1;
            st.cg_iters += cgit[s];
            st.alpha_min = std::min(st.alpha_min, alpha[s]);
            st.ccd_alpha_min = std::min(st.ccd_alpha_min, amax[s]);
            if (res[s] < C.P.newton_tol) { st.converged = 1; active[s] = 0; }
            else if (C.P.stall_alpha > 0.0 && alpha[s] < C.P.stall_alpha) { st.stalled = 1; active[s] = 0; }
        }
    }
    for (int s = 0; s < C.ns; ++s) {
        double k = 0.0;
        for (int b = C.scene_b0[s]; b < C.scene_b1[s]; ++b) k = std::max(k, C.kb[b]);
        C.stats[s].k_contact = k;
    }
    // forward_states (abd.py:574-581)
    for (int i = 0; i < 12 * C.nb; ++i) {
        C.qdot[i] = (C.q[i] - C.qnow[i]) / C.P.dt;
        C.qnow[i] = C.q[i];
    }
    compute_qtilde(C);
}

}  // namespace

// =============================================================================================
// C ABI (ctypes) — test infrastructure
// =============================================================================================
extern "C" {

void orc_barrier(double s, double shat, double* out3) { barrier(s, shat, out3[0], out3[1], out3[2]); }

double orc_ortho(const double* q12, double kappa, double* g9, double* H81) { return ortho(q12, kappa, g9, H81); }

// kind ABD_PT / ABD_EE: x12 = 4 points; returns d^2, fills g12/H144 (nullable) and dtype
double orc_pair_dist2(int kind, const double* x12, double* g12, double* H144, int* dtype) {
    V3 x[4];
    for (int k = 0; k < 4; ++k) x[k] = v3(x12 + 3 * k);
    return pair_dist2(kind, x, g12, H144, dtype);
}

void orc_seg_params(const double* x12, double* s, double* t) {
    V3 x[4];
    for (int k = 0; k < 4; ++k) x[k] = v3(x12 + 3 * k);
    int st[4], ns;
    classify_ee(x, st, &ns, s, t);
}

void orc_psd_project(int n, double* H) { psd_project(n, H); }

double orc_accd(int kind, const double* x0, const double* x1, double s, double tmax, int max_iter) {
    V3 a[4], b[4];
    for (int k = 0; k < 4; ++k) { a[k] = v3(x0 + 3 * k); b[k] = v3(x1 + 3 * k); }
    return accd(kind, a, b, s, tmax, max_iter);
}

int orc_chol_inverse12(const double* A, double* Ainv) { return chol_inverse12(A, Ainv) ? 0 : -1; }

void* orc_create(const AbdSceneDesc* d, const AbdParams* p) {
    Ctx* C = new Ctx();
    C->P = *p;
    C->nb = d->n_bodies; C->nv = d->n_verts; C->ne = d->n_edges; C->nt = d->n_tris; C->ns = d->n_scenes;
    C->rest.assign(d->rest_x, d->rest_x + 3 * C->nv);
    C->vbody.assign(d->vert_body, d->vert_body + C->nv);
    C->edges.assign(d->edges, d->edges + 2 * C->ne);
    C->tris.assign(d->tris, d->tris + 3 * C->nt);
    C->bscene.assign(d->body_scene, d->body_scene + C->nb);
    C->pinned.assign(d->body_pinned, d->body_pinned + C->nb);
    C->mass4.assign(d->mass4, d->mass4 + 16 * C->nb);
    C->scene_b0.assign(C->ns, C->nb);
    C->scene_b1.assign(C->ns, 0);
    for (int b = 0; b < C->nb; ++b) {
        int s = C->bscene[b];
        C->scene_b0[s] = std::min(C->scene_b0[s], b);
        C->scene_b1[s] = std::max(C->scene_b1[s], b + 1);
    }
    C->sc_v.resize(C->ns); C->sc_e.resize(C->ns); C->sc_t.resize(C->ns);
    for (int v = 0; v < C->nv; ++v) C->sc_v[scene_of_vert(*C, v)].push_back(v);
    for (int e = 0; e < C->ne; ++e) C->sc_e[scene_of_vert(*C, C->edges[2 * e])].push_back(e);
    for (int t = 0; t < C->nt; ++t) C->sc_t[scene_of_vert(*C, C->tris[3 * t])].push_back(t);
    {
        double* n = C->P.ground_normal;
        double l = std::sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
        if (l > 0) { n[0] /= l; n[1] /= l; n[2] /= l; }
    }
    C->q.assign(12 * C->nb, 0.0);
    for (int b = 0; b < C->nb; ++b) { C->q[12 * b + 3] = C->q[12 * b + 7] = C->q[12 * b + 11] = 1.0; }
    C->qnow = C->q;
    C->qdot.assign(12 * C->nb, 0.0);
    C->kb.assign(C->nb, C->P.k_contact);
    compute_qtilde(*C);
    return C;
}

void orc_destroy(void* c) { delete (Ctx*)c; }

void orc_set_threads(void* c, int n) { ((Ctx*)c)->threads = std::max(1, n); }

void orc_set_params(void* c, const AbdParams* p) {
    Ctx* C = (Ctx*)c;
    C->P = *p;
    double* n = C->P.ground_normal;
    double l = std::sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
    if (l > 0) { n[0] /= l; n[1] /= l; n[2] /= l; }
    C->kb.assign(C->nb, C->P.k_contact);
    compute_qtilde(*C);
}

void orc_get_stiffness(void* c, double* k) {
    Ctx* C = (Ctx*)c;
    std::memcpy(k, C->kb.data(), C->nb * sizeof(double));
}

void orc_set_stiffness(void* c, const double* k) {
    Ctx* C = (Ctx*)c;
    std::memcpy(C->kb.data(), k, C->nb * sizeof(double));
}

// error bits raised since the last call (64: a vertex on or below the ground plane was assembled)
int orc_take_error(void* c) { return ((Ctx*)c)->err.exchange(0); }

void orc_set_state(void* c, const double* q, const double* qnow, const double* qdot) {
    Ctx* C = (Ctx*)c;
    size_t n = 12 * (size_t)C->nb;
    if (q) std::memcpy(C->q.data(), q, n * sizeof(double));
    if (qnow) std::memcpy(C->qnow.data(), qnow, n * sizeof(double));
    if (qdot) std::memcpy(C->qdot.data(), qdot, n * sizeof(double));
    compute_qtilde(*C);
}

void orc_get_state(void* c, double* q, double* qnow, double* qdot) {
    Ctx* C = (Ctx*)c;
    size_t n = 12 * (size_t)C->nb;
    if (q) std::memcpy(q, C->q.data(), n * sizeof(double));
    if (qnow) std::memcpy(qnow, C->qnow.data(), n * sizeof(double));
    if (qdot) std::memcpy(qdot, C->qdot.data(), n * sizeof(double));
}

// active pairs at q (NULL = current iterate) as (kind,i,j) triples + d^2; returns count
int orc_detect(void* c, const double* q, int* pairs, double* d2, int cap) {
    Ctx* C = (Ctx*)c;
    std::vector<Pair> act;
    std::vector<double> dd;
    active_pairs(*C, q ? q : C->q.data(), act, &dd);
    for (size_t k = 0; k < act.size() && (int)k < cap; ++k) {
        pairs[3 * k] = act[k].kind; pairs[3 * k + 1] = act[k].i; pairs[3 * k + 2] = act[k].j;
        if (d2) d2[k] = dd[k];
    }
    return (int)act.size();
}

// candidate pairs of the broad phase over q0 -> q1 inflated by infl; returns count
int orc_candidates(void* c, const double* q0, const double* q1, double infl, int* pairs, int cap) {
    Ctx* C = (Ctx*)c;
    std::vector<V3> X0, X1;
    all_positions(*C, q0, X0);
    all_positions(*C, q1 ? q1 : q0, X1);
    std::vector<Pair> cand;
    broad_phase(*C, X0, X1, infl, cand);
    for (size_t k = 0; k < cand.size() && (int)k < cap; ++k) {
        pairs[3 * k] = cand[k].kind; pairs[3 * k + 1] = cand[k].i; pairs[3 * k + 2] = cand[k].j;
    }
    return (int)cand.size();
}

void orc_energy(void* c, const double* q, double* E) {
    Ctx* C = (Ctx*)c;
    scene_energy(*C, q ? q : C->q.data(), E);
}

// assemble at q (NULL = current): grad[12nb], diag[144nb], off-diagonal upper blocks; returns
// the number of off-diagonal blocks (pair_ij/offdiag filled up to cap)
int orc_assemble(void* c, const double* q, double* grad, double* diag, int* pair_ij, double* offdiag, int cap) {
    Ctx* C = (Ctx*)c;
    const double* qq = q ? q : C->q.data();
    std::vector<Pair> act;
    active_pairs(*C, qq, act, nullptr);
    System S;
    assemble(*C, qq, act, S);
    if (grad) std::memcpy(grad, S.grad.data(), S.grad.size() * sizeof(double));
    if (diag) std::memcpy(diag, S.diag.data(), S.diag.size() * sizeof(double));
    int k = 0;
    for (auto& kv : S.off) {
        if (k < cap) {
            if (pair_ij) { pair_ij[2 * k] = kv.first.first; pair_ij[2 * k + 1] = kv.first.second; }
            if (offdiag) std::memcpy(offdiag + 144 * k, kv.second.data(), 144 * sizeof(double));
        }
        ++k;
    }
    return k;
}

// assemble at the current iterate and solve; du[12nb], cg iterations per scene
void orc_solve(void* c, double* du, int* iters, double* residual) {
    Ctx* C = (Ctx*)c;
    std::vector<Pair> act;
    active_pairs(*C, C->q.data(), act, nullptr);
    System S;
    assemble(*C, C->q.data(), act, S);
    std::vector<double> x;
    std::vector<int> it;
    pcg(*C, S, x, it);
    if (du) std::memcpy(du, x.data(), x.size() * sizeof(double));
    for (int s = 0; s < C->ns; ++s) {
        if (iters) iters[s] = it[s];
        if (residual) {
            double t = 0;
            for (int i = 12 * C->scene_b0[s]; i < 12 * C->scene_b1[s]; ++i) t += S.grad[i] * S.grad[i];
            residual[s] = std::sqrt(t);
        }
    }
}

void orc_ccd(void* c, const double* du, double* alpha) {
    Ctx* C = (Ctx*)c;
    std::vector<double> d(du, du + 12 * C->nb), a;
    ccd_bound(*C, C->q.data(), d, a);
    for (int s = 0; s < C->ns; ++s) alpha[s] = a[s];
}

// the 49 body features of the learned models at the current iterate (float feat[49 nb]): q - q~,
// qdot, the number of active contacts incident to the body and their summed gradient (unpinned
// bodies only), and the block-Jacobi step z = D_b^-1 grad_b
void orc_features(void* c, float* feat) {
    Ctx* C = (Ctx*)c;
    const double* q = C->q.data();
    std::vector<Pair> act;
    active_pairs(*C, q, act, nullptr);
    System S;
    assemble(*C, q, act, S);
    std::vector<V3> X;
    all_positions(*C, q, X);
    std::vector<double> gs(12 * (size_t)C->nb, 0.0);
    std::vector<int> cnt(C->nb, 0);
    for (auto& pr : act) {
        ContactTerm T;
        contact_term(*C, pr, X, T);
        for (int side = 0; side < 2; ++side) {
            const int b = T.b[side];
            if (b < 0 || C->pinned[b]) continue;
            cnt[b]++;
            for (int k = 0; k < 12; ++k) gs[12 * b + k] += T.g[side][k];
        }
    }
    for (int b = 0; b < C->nb; ++b) {
        float* f = feat + 49 * (size_t)b;
        for (int k = 0; k < 12; ++k) {
            f[k] = (float)(q[12 * b + k] - C->qtilde[12 * b + k]);
            f[12 + k] = (float)C->qdot[12 * b + k];
            f[25 + k] = (float)gs[12 * b + k];
        }
        f[24] = (float)cnt[b];
        double P[144];
        if (C->pinned[b]) {
            for (int k = 0; k < 144; ++k) P[k] = (k % 13 == 0) ? 1.0 / C->P.pinned_diag : 0.0;
        } else if (!chol_inverse12(&S.diag[144 * b], P)) {
            C->err |= 256;
            for (int k = 0; k < 144; ++k) P[k] = 0.0;
        }
        for (int i = 0; i < 12; ++i) {
            double z = 0.0;
            for (int j = 0; j < 12; ++j) z += P[12 * i + j] * S.grad[12 * b + j];
            f[37 + i] = (float)z;
        }
    }
}

// the step predictor's dq[12 nb] for the following steps with AbdParams.predictor 3 (NULL clears)
void orc_set_prediction(void* c, const double* dq) {
    Ctx* C = (Ctx*)c;
    if (dq) C->pred_dq.assign(dq, dq + 12 * (size_t)C->nb);
    else C->pred_dq.clear();
}

// the line search along a given direction (the Newton search's rules); alpha[ns]
void orc_line_search(void* c, const double* du, double* alpha_out) {
    Ctx* C = (Ctx*)c;
    std::vector<double> d(du, du + 12 * C->nb), alpha, amax;
    std::vector<int> fh;
    std::vector<char> active(C->ns, 1);
    C->stats.assign(C->ns, AbdStepStats{});
    line_search(*C, d, active, true, alpha, amax, fh);
    for (int s = 0; s < C->ns; ++s) alpha_out[s] = alpha[s];
}

void orc_step(void* c, int nsteps, AbdStepStats* stats) {
    Ctx* C = (Ctx*)c;
    for (int k = 0; k < nsteps; ++k) step_once(*C);
    if (stats)
        for (int s = 0; s < C->ns; ++s) stats[s] = C->stats[s];
}

void orc_get_iter_log(void* c, double* res, double* alpha, int* cg) {
    Ctx* C = (Ctx*)c;
    size_t n = C->log_res.size();
    if (res) std::memcpy(res, C->log_res.data(), n * sizeof(double));
    if (alpha) std::memcpy(alpha, C->log_alpha.data(), n * sizeof(double));
    if (cg) std::memcpy(cg, C->log_cg.data(), n * sizeof(int));
}

void orc_get_vertices(void* c, double* xyz) {
    Ctx* C = (Ctx*)c;
    for (int v = 0; v < C->nv; ++v) {
        V3 p = vpos(*C, C->q.data(), v);
        xyz[3 * v] = p.x; xyz[3 * v + 1] = p.y; xyz[3 * v + 2] = p.z;
    }
}

}  // extern "C"
