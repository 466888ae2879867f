// Direct solve of the Newton system per scene: dense Cholesky H = L L^T and two triangular solves.
//
// The reference's `else` branch of compute_du (abd.py:637-654) hands the assembled CSR matrix to
// a sparse direct solver (dxslv.CUSolver: analyze_pattern, factorize, solve_cuda).  Here the
// block-sparse system of each scene is scattered into a dense matrix of padded order N (a multiple
// of DN_T, the padding rows/columns hold the identity) and factored by a right-looking tiled
// Cholesky: per column tile k one launch factors the diagonal tile and solves the panel below it,
// one launch applies the symmetric rank-DN_T update to the trailing lower tiles.  Every scene of a
// batch runs in the same launches (grid.y = scene); a scene whose order is below the tile column
// leaves its workgroups idle.  Only the lower triangle is read.
//
// Layout: scene s owns A + off[s], an N_s x N_s row-major FP64 matrix (N_s = 0: not solved here).
#include <hip/hip_runtime.h>
#include <algorithm>

#include "abd_internal.hpp"

namespace abd {

static inline int cdiv(long a, int b) { return (int)((a + b - 1) / b); }

namespace {

constexpr int DN_T = DIRECT_TILE;

// zero the scene's matrix, identity on the padding rows past its 12 * (b1 - b0) unknowns
__global__ void k_dense_clear(const long long* __restrict__ off, const int* __restrict__ N,
                              const int* __restrict__ scene_b0, const int* __restrict__ mask, double* A) {
    const int s = blockIdx.y;
    const int n = N[s];
    if (n == 0 || !mask[s]) return;
    const int nr = 12 * (scene_b0[s + 1] - scene_b0[s]);
    double* As = A + off[s];
    const long total = (long)n * n;
    for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (long)gridDim.x * blockDim.x) {
        const int i = (int)(e / n), j = (int)(e % n);
        As[e] = (i == j && i >= nr) ? 1.0 : 0.0;
    }
}

// body diagonal blocks (full 12 x 12) and the lower block (B, A), B > A, of every body pair
__global__ void k_dense_scatter(int nb, int n_pairs, const int* __restrict__ bscene, const int* __restrict__ scene_b0,
                                const long long* __restrict__ off, const int* __restrict__ N,
                                const int* __restrict__ mask, const double* __restrict__ diag,
                                const int* __restrict__ pair_b, const double* __restrict__ offT, double* A) {
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long nd = (long)nb * 144;
    if (t >= nd + (long)n_pairs * 144) return;
    const int e = (int)(t % 144), R = e / 12, Cc = e % 12;
    int row_b, col_b;
    double v;
    if (t < nd) {
        row_b = col_b = (int)(t / 144);
        v = diag[t];
    } else {
        const int u = (int)((t - nd) / 144);
        col_b = pair_b[2 * u];      // A (the smaller body index)
        row_b = pair_b[2 * u + 1];  // B
        v = offT[(size_t)u * 144 + e];  // offT holds the (B, A) block row-major
    }
    const int s = bscene[row_b];
    const int n = N[s];
    if (n == 0 || !mask[s]) return;
    const int b0 = scene_b0[s];
    const long i = 12L * (row_b - b0) + R, j = 12L * (col_b - b0) + Cc;
    A[off[s] + i * n + j] = v;
}

// column tile k: every workgroup (one wave) factors the diagonal tile A_kk (bit-identical copies):
// the tile sits in LDS, lane i updates row i after reading column j into registers.  A pivot that is
// not positive sets fail[s] and is replaced by 1 (the result is then discarded).  Workgroup 0 keeps
// L_kk in Lkk[s] (the following update launch writes it into A: A_kk is still being read here), and
// each thread solves one row i of the panel below: L_ik = A_ik L_kk^-T
__global__ __launch_bounds__(64) void k_chol_panel(int k, const long long* __restrict__ off,
                                                   const int* __restrict__ N, const int* __restrict__ mask,
                                                   double* A, double* Lkk, int* fail) {
    const int s = blockIdx.y;
    const int n = N[s];
    if (n == 0 || !mask[s] || k * DN_T >= n) return;
    __shared__ double L[DN_T][DN_T + 1];
    double* As = A + off[s];
    const int c0 = k * DN_T;
    const int lane = threadIdx.x, li = lane & (DN_T - 1);  // lanes 32..63 mirror lanes 0..31
    for (int idx = lane; idx < DN_T * DN_T; idx += 64) {
        const int r = idx / DN_T, m = idx % DN_T;
        L[r][m] = (m <= r) ? As[(long)(c0 + r) * n + c0 + m] : 0.0;
    }
    bool bad = false;
    const bool own = lane < DN_T;
    // two pivots per round: columns j and j + 1 are factored together (every lane computes the
    // 2 x 2 leading block from three LDS reads), then each row takes a rank-2 update
#pragma unroll
    for (int j = 0; j < DN_T; j += 2) {
        __syncthreads();
        double d0 = L[j][j];
        if (!(d0 > 0.0)) {
            bad = true;
            d0 = 1.0;
        }
        const double s0 = sqrt(d0), i0 = 1.0 / s0;
        const double l10 = L[j + 1][j] * i0;
        double d1 = L[j + 1][j + 1] - l10 * l10;
        if (!(d1 > 0.0)) {
            bad = true;
            d1 = 1.0;
        }
        const double s1 = sqrt(d1), i1 = 1.0 / s1;
        // rows below the block: their entries in columns j, j + 1; x0 = x1 = 0 on the other rows
        double x0 = 0.0, x1 = 0.0;
        if (li > j + 1) {
            x0 = L[li][j] * i0;
            x1 = (L[li][j + 1] - x0 * l10) * i1;
        }
        __syncthreads();
        if (own) {
            if (li > j + 1) {
                L[li][j] = x0;
                L[li][j + 1] = x1;
            } else if (li == j + 1) {
                L[j + 1][j] = l10;
                L[j + 1][j + 1] = s1;
            } else if (li == j) {
                L[j][j] = s0;
            }
        }
        __syncthreads();
        if (own && j + 2 < DN_T) {
            // row li -= x0 * column j + x1 * column j + 1: the columns first (other lanes' rows),
            // then the own row.  The update runs over the whole row without a per-entry test:
            // entries right of the diagonal take values that are never read (every read below is
            // of the lower triangle), and rows li <= j + 1 have x0 = x1 = 0.
            double c0v[DN_T], c1v[DN_T];
#pragma unroll
            for (int m = j + 2; m < DN_T; ++m) {
                c0v[m] = L[m][j];
                c1v[m] = L[m][j + 1];
            }
#pragma unroll
            for (int m = j + 2; m < DN_T; ++m) L[li][m] -= x0 * c0v[m] + x1 * c1v[m];
        }
    }
    __syncthreads();
    if (bad && lane == 0) atomicOr(fail + s, 1);
    if (blockIdx.x == 0)
        for (int idx = lane; idx < DN_T * DN_T; idx += 64)
            Lkk[(size_t)s * DN_T * DN_T + idx] = L[idx / DN_T][idx % DN_T];
    const long i = (long)c0 + DN_T + (long)blockIdx.x * 64 + lane;
    if (i >= n) return;
    double* row = As + i * n + c0;
    double x[DN_T];
#pragma unroll
    for (int c = 0; c < DN_T; ++c) x[c] = row[c];
    // x L_kk^T = a by forward substitution along the row
#pragma unroll
    for (int c = 0; c < DN_T; ++c) {
        double t = x[c];
#pragma unroll
        for (int m = 0; m < c; ++m) t -= x[m] * L[c][m];
        x[c] = t / L[c][c];
    }
#pragma unroll
    for (int c = 0; c < DN_T; ++c) row[c] = x[c];
}

// trailing update after column tile k: A_IJ -= L_Ik L_Jk^T for k < J <= I < N/DN_T; the extra
// workgroup x = 0 stores L_kk into A
__global__ __launch_bounds__(256) void k_chol_update(int k, const long long* __restrict__ off,
                                                     const int* __restrict__ N, const int* __restrict__ mask,
                                                     double* A, const double* __restrict__ Lkk) {
    const int s = blockIdx.y;
    const int n = N[s];
    if (n == 0 || !mask[s] || k * DN_T >= n) return;
    double* As = A + off[s];
    const int tid = threadIdx.x;
    if (blockIdx.x == 0) {
        const long c0 = (long)k * DN_T;
        for (int idx = tid; idx < DN_T * DN_T; idx += 256) {
            const int i = idx / DN_T, m = idx % DN_T;
            if (m <= i) As[(c0 + i) * n + c0 + m] = Lkk[(size_t)s * DN_T * DN_T + idx];
        }
        return;
    }
    const long li = blockIdx.x - 1;
    int Ip = (int)((sqrt(8.0 * (double)li + 1.0) - 1.0) * 0.5);
    while ((long)(Ip + 1) * (Ip + 2) / 2 <= li) ++Ip;
    while ((long)Ip * (Ip + 1) / 2 > li) --Ip;
    const int Jp = (int)(li - (long)Ip * (Ip + 1) / 2);
    const int I = k + 1 + Ip, J = k + 1 + Jp;
    if (I * DN_T >= n) return;
    __shared__ double LI[DN_T][DN_T + 1], LJ[DN_T][DN_T + 1];
    const long c0 = (long)k * DN_T;
    for (int idx = tid; idx < DN_T * DN_T; idx += 256) {
        const int r = idx / DN_T, m = idx % DN_T;
        LI[r][m] = As[((long)I * DN_T + r) * n + c0 + m];
        LJ[r][m] = As[((long)J * DN_T + r) * n + c0 + m];
    }
    __syncthreads();
    const int r = tid / 8, cq = (tid % 8) * 4;
    double acc[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll 8
    for (int m = 0; m < DN_T; ++m) {
        const double a = LI[r][m];
#pragma unroll
        for (int q = 0; q < 4; ++q) acc[q] += a * LJ[cq + q][m];
    }
    double* out = As + ((long)I * DN_T + r) * n + (long)J * DN_T + cq;
#pragma unroll
    for (int q = 0; q < 4; ++q) out[q] -= acc[q];
}

__device__ __forceinline__ double shfl_d(double v, int src) {
    return __shfl(v, src, 64);
}

// L y = b, L^T x = y for one scene per workgroup (the vector in LDS), then du = x on the scene's
// bodies, ||b|| into bnorm and residual (the PCG reports the same norm)
__global__ __launch_bounds__(256) void k_chol_solve(const long long* __restrict__ off, const int* __restrict__ N,
                                                    const int* __restrict__ mask, const int* __restrict__ scene_b0,
                                                    const double* __restrict__ A, const double* __restrict__ b,
                                                    double* du, double* bnorm, double* residual, int* cg_iters,
                                                    int zero_iters) {
    const int s = blockIdx.x;
    const int n = N[s];
    if (n == 0 || !mask[s]) return;
    __shared__ double y[DIRECT_MAX_DOFS];
    __shared__ double red[8][DN_T];
    __shared__ double nrm[4];
    const double* As = A + off[s];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const long v0 = 12L * scene_b0[s];
    const int nr = 12 * (scene_b0[s + 1] - scene_b0[s]);
    double bb = 0.0;
    for (int i = tid; i < n; i += 256) {
        const double v = i < nr ? b[v0 + i] : 0.0;
        y[i] = v;
        bb += v * v;
    }
    for (int o = 32; o > 0; o >>= 1) bb += __shfl_xor(bb, o, 64);
    if (lane == 0) nrm[wave] = bb;
    __syncthreads();
    const int T = n / DN_T;
    // forward substitution, one tile row at a time
    for (int kt = 0; kt < T; ++kt) {
        const int r = tid / 8, sl = tid % 8;
        const double* row = As + ((long)kt * DN_T + r) * n;
        double a4[4] = {0.0, 0.0, 0.0, 0.0};  // four chains: kt * DN_T is a multiple of 32
#pragma unroll 4
        for (int c = sl; c < kt * DN_T; c += 32) {
#pragma unroll
            for (int q = 0; q < 4; ++q) a4[q] += row[c + 8 * q] * y[c + 8 * q];
        }
        double acc = (a4[0] + a4[1]) + (a4[2] + a4[3]);
        acc += __shfl_xor(acc, 1, 64);
        acc += __shfl_xor(acc, 2, 64);
        acc += __shfl_xor(acc, 4, 64);
        if (sl == 0) red[0][r] = acc;
        __syncthreads();
        if (wave == 0) {
            const int rr = lane & (DN_T - 1);
            const double* Lr = As + ((long)kt * DN_T + rr) * n + (long)kt * DN_T;
            double lrow[DN_T];
#pragma unroll
            for (int m = 0; m < DN_T; ++m) lrow[m] = Lr[m];  // row rr of the diagonal tile
            const double rdg = 1.0 / Lr[rr];
            double t = y[kt * DN_T + rr] - red[0][rr];
#pragma unroll
            for (int m = 0; m < DN_T; ++m) {
                const double ym = shfl_d(t * rdg, m);  // lane m holds row m: its quotient is y_m
                if (rr > m) t -= lrow[m] * ym;
                if (rr == m) t = ym;
            }
            if (lane < DN_T) y[kt * DN_T + rr] = t;
        }
        __syncthreads();
    }
    // back substitution with L^T, last tile first
    for (int kt = T - 1; kt >= 0; --kt) {
        const int r = tid % DN_T, sl = tid / DN_T;
        double a4[4] = {0.0, 0.0, 0.0, 0.0};  // four chains: n - (kt + 1) * DN_T is a multiple of 32
#pragma unroll 4
        for (int c = (kt + 1) * DN_T + sl; c < n; c += 32) {
#pragma unroll
            for (int q = 0; q < 4; ++q) a4[q] += As[(long)(c + 8 * q) * n + (long)kt * DN_T + r] * y[c + 8 * q];
        }
        red[sl][r] = (a4[0] + a4[1]) + (a4[2] + a4[3]);
        __syncthreads();
        if (wave == 0) {
            const int rr = lane & (DN_T - 1);
            double t = y[kt * DN_T + rr];
            for (int q = 0; q < 8; ++q) t -= red[q][rr];
            // lane rr keeps column rr of the diagonal tile: lcol[m] = L[m][rr]
            const double* Lc = As + (long)kt * DN_T * n + (long)kt * DN_T + rr;
            double lcol[DN_T];
#pragma unroll
            for (int m = 0; m < DN_T; ++m) lcol[m] = Lc[(long)m * n];
            const double rdg = 1.0 / Lc[(long)rr * n];
#pragma unroll
            for (int m = DN_T - 1; m >= 0; --m) {
                const double xm = shfl_d(t * rdg, m);  // valid on lane m
                if (rr < m) t -= lcol[m] * xm;
                if (rr == m) t = xm;
            }
            if (lane < DN_T) y[kt * DN_T + rr] = t;
        }
        __syncthreads();
    }
    for (int i = tid; i < nr; i += 256) du[v0 + i] = y[i];
    if (tid == 0) {
        const double bn = sqrt(nrm[0] + nrm[1] + nrm[2] + nrm[3]);
        bnorm[s] = bn;
        residual[s] = bn;
        if (zero_iters) cg_iters[s] = 0;
    }
}

// scenes the direct solver takes: every Newton-active scene it can hold (fallback == 0), or those
// whose PCG ended at its iteration cap (fallback == 1)
__global__ void k_direct_mask(int ns, int fallback, int cg_max_iter, const int* __restrict__ N,
                              const int* __restrict__ active, const int* __restrict__ cg_iters, int* mask, int* any) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= ns) return;
    const int m = N[s] > 0 && active[s] && (!fallback || cg_iters[s] >= cg_max_iter);
    mask[s] = m;
    if (m) atomicAdd(any, 1);
}

}  // namespace

void launch_direct_mask(hipStream_t st, int ns, int fallback, int cg_max_iter, const int* N, const int* active,
                        const int* cg_iters, int* mask, int* any) {
    hipLaunchKernelGGL(k_direct_mask, dim3(cdiv(ns, 128)), dim3(128), 0, st, ns, fallback, cg_max_iter, N, active,
                       cg_iters, mask, any);
}

void launch_direct_factor(hipStream_t st, const DirectBuf& D, int ns, int nb, int n_pairs, const int* bscene,
                          const int* scene_b0, const double* diag, const int* pair_b, const double* offT) {
    if (D.n_max <= 0) return;
    const long nn = (long)D.n_max * D.n_max;
    hipLaunchKernelGGL(k_dense_clear, dim3((unsigned)std::min<long>(cdiv(nn, 256), 65535), ns), dim3(256), 0, st,
                       D.off, D.N, scene_b0, D.mask, D.A);
    const long nent = (long)(nb + n_pairs) * 144;
    hipLaunchKernelGGL(k_dense_scatter, dim3(cdiv(nent, 256)), dim3(256), 0, st, nb, n_pairs, bscene, scene_b0, D.off,
                       D.N, D.mask, diag, pair_b, offT, D.A);
    const int T = D.n_max / DN_T;
    for (int k = 0; k < T; ++k) {
        const int below = D.n_max - (k + 1) * DN_T;
        hipLaunchKernelGGL(k_chol_panel, dim3(std::max(1, cdiv(below, 64)), ns), dim3(64), 0, st, k, D.off, D.N,
                           D.mask, D.A, D.Lkk, D.fail);
        const long rem = T - k - 1;
        hipLaunchKernelGGL(k_chol_update, dim3((unsigned)(1 + rem * (rem + 1) / 2), ns), dim3(256), 0, st, k, D.off,
                           D.N, D.mask, D.A, D.Lkk);
    }
}

void launch_direct_solve(hipStream_t st, const DirectBuf& D, int ns, const int* scene_b0, const double* b,
                         double* du, double* bnorm, double* residual, int* cg_iters, int zero_iters) {
    if (D.n_max <= 0) return;
    hipLaunchKernelGGL(k_chol_solve, dim3(ns), dim3(256), 0, st, D.off, D.N, D.mask, scene_b0, D.A, b, du, bnorm,
                       residual, cg_iters, zero_iters);
}

}  // namespace abd
