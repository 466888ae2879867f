// abd_mlp.hip — the per-body MLP of the learned warm start: forward, backward and Adam training
// on the GPU with hand-written kernels (no counterpart in the reference).
//
// Model: fully connected, ReLU between layers, dims d_0 -> d_1 -> ... -> d_L.  Every dimension is
// padded to a multiple of 16 (padded weights, biases and features are zero and stay zero).
// Weights are kept as an FP32 master copy (Adam) and a BF16 copy that the matrix products read;
// products accumulate in FP32 on the matrix cores (v_mfma_f32_16x16x16_bf16, one wave per 16x16
// output tile).  Activations between layers are stored in BF16, pre-activations in FP32.
//
// Loss of a batch of n rows: sum_ij w_j (Y_ij - T_ij)^2 / n with per-output weights w (0 on the
// padded columns).
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "abd_device.hpp"
#include "abd_mlp.hpp"

namespace abd {

typedef short short4v __attribute__((vector_size(4 * sizeof(short))));
typedef float float4v __attribute__((vector_size(4 * sizeof(float))));

static inline int cdivm(long a, int b) { return (int)((a + b - 1) / b); }
static inline int pad16(int n) { return (n + 15) / 16 * 16; }

__device__ __forceinline__ float bf2f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

// C(i, j) = sum_k A(i, k) B(k, j) [+ bias(j)], A(i, k) = A[i*sa0 + k*sa1], B(k, j) = B[k*sb0 + j*sb1],
// C row-major with leading dimension ldc.  M, N, K are multiples of 16; one wave per 16x16 tile of C.
// Lane l holds A(row l%16, k0 + 4(l/16) + 0..3) and B(k0 + 4(l/16) + 0..3, col l%16) and receives
// C(4(l/16) + 0..3, l%16).
__global__ __launch_bounds__(64) void k_gemm_bf16(int M, int N, int K, const uint16_t* __restrict__ A, int sa0,
                                                   int sa1, const uint16_t* __restrict__ B, int sb0, int sb1,
                                                   float* __restrict__ C, int ldc, const float* __restrict__ bias) {
    const int tiles_n = N / 16;
    const int ti = blockIdx.x / tiles_n, tj = blockIdx.x % tiles_n;
    const int lane = threadIdx.x;
    const int r = lane % 16, q = 4 * (lane / 16);
    const int i0 = 16 * ti, j0 = 16 * tj;
    float4v acc = {0.f, 0.f, 0.f, 0.f};
    for (int k0 = 0; k0 < K; k0 += 16) {
        short4v a, b;
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            a[u] = (short)A[(size_t)(i0 + r) * sa0 + (size_t)(k0 + q + u) * sa1];
            b[u] = (short)B[(size_t)(k0 + q + u) * sb0 + (size_t)(j0 + r) * sb1];
        }
        acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc, 0, 0, 0);
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) {
        const int i = i0 + q + u, j = j0 + r;
        C[(size_t)i * ldc + j] = acc[u] + (bias ? bias[j] : 0.f);
    }
}

static void gemm(hipStream_t s, int M, int N, int K, const uint16_t* A, int sa0, int sa1, const uint16_t* B, int sb0,
                 int sb1, float* C, int ldc, const float* bias) {
    hipLaunchKernelGGL(k_gemm_bf16, dim3((M / 16) * (N / 16)), dim3(64), 0, s, M, N, K, A, sa0, sa1, B, sb0, sb1, C,
                       ldc, bias);
}

// rows [0, n) of X (n x d, FP32, host layout) -> BF16 (rows x dp), zero padding
__global__ void k_pack_rows(int n, int rows, int d, int dp, const float* __restrict__ X, const int* __restrict__ idx,
                            uint16_t* __restrict__ out) {
    const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (long)rows * dp) return;
    const int i = (int)(e / dp), j = (int)(e % dp);
    float v = 0.f;
    if (i < n && j < d) v = X[(size_t)(idx ? idx[i] : i) * d + j];
    out[e] = f2bf(v);
}

// same for FP32 targets
__global__ void k_pack_rows_f32(int n, int rows, int d, int dp, const float* __restrict__ X, const int* __restrict__ idx,
                                float* __restrict__ out) {
    const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (long)rows * dp) return;
    const int i = (int)(e / dp), j = (int)(e % dp);
    float v = 0.f;
    if (i < n && j < d) v = X[(size_t)(idx ? idx[i] : i) * d + j];
    out[e] = v;
}

// A = bf16(relu(Z)) on rows < n (padded rows zero)
__global__ void k_relu_bf16(long total, int n, int dp, const float* __restrict__ Z, uint16_t* __restrict__ A) {
    const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= total) return;
    const int i = (int)(e / dp);
    A[e] = (i < n) ? f2bf(fmaxf(Z[e], 0.f)) : (uint16_t)0;
}

// dY = w_j 2 (Y - T) / n on rows < n; per-row squared error partials for the loss
__global__ void k_loss_grad(int n, int rows, int dp, const float* __restrict__ Y, const float* __restrict__ T,
                            const float* __restrict__ w, uint16_t* __restrict__ G, float* __restrict__ row_loss) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= rows) return;
    float acc = 0.f;
    for (int j = 0; j < dp; ++j) {
        const size_t e = (size_t)i * dp + j;
        float g = 0.f;
        if (i < n) {
            const float d = Y[e] - T[e];
            acc += w[j] * d * d;
            g = w[j] * 2.f * d / (float)n;
        }
        G[e] = f2bf(g);
    }
    row_loss[i] = (i < n) ? acc / (float)n : 0.f;
}

// G_prev = bf16(dA * (Z > 0))
__global__ void k_relu_back(long total, const float* __restrict__ dA, const float* __restrict__ Z,
                            uint16_t* __restrict__ G) {
    const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= total) return;
    G[e] = f2bf(Z[e] > 0.f ? dA[e] : 0.f);
}

// column sums of a BF16 matrix (bias gradients), fixed order
// column sums: 16 columns per workgroup, 16 row strides per column summed in a fixed order
__global__ __launch_bounds__(256) void k_colsum(int rows, int dp, const uint16_t* __restrict__ G,
                                                float* __restrict__ out) {
    __shared__ float sh[16][17];
    const int c = threadIdx.x % 16, r = threadIdx.x / 16;
    const int j = blockIdx.x * 16 + c;
    float s = 0.f;
    if (j < dp)
        for (int i = r; i < rows; i += 16) s += bf2f(G[(size_t)i * dp + j]);
    sh[r][c] = s;
    __syncthreads();
    if (r == 0 && j < dp) {
        float t = 0.f;
        for (int k = 0; k < 16; ++k) t += sh[k][c];
        out[j] = t;
    }
}

// fixed-order sum of n floats by one workgroup
__global__ __launch_bounds__(256) void k_sum(int n, const float* __restrict__ x, float* out) {
    __shared__ float sh[256];
    float s = 0.f;
    for (int i = threadIdx.x; i < n; i += 256) s += x[i];
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (threadIdx.x < o) sh[threadIdx.x] += sh[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) *out = sh[0];
}

// Adam on the FP32 masters (mask: 1 on real entries, 0 on padding), then the BF16 copy
__global__ void k_adam(long n, float* __restrict__ w, const float* __restrict__ g, const float* __restrict__ mask,
                       float* __restrict__ m, float* __restrict__ v, float lr, float b1, float b2, float eps, float c1,
                       float c2, uint16_t* __restrict__ wb) {
    const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n) return;
    const float gg = g[e] * mask[e];
    const float mm = b1 * m[e] + (1.f - b1) * gg, vv = b2 * v[e] + (1.f - b2) * gg * gg;
    m[e] = mm;
    v[e] = vv;
    const float wn = w[e] - lr * (mm / c1) / (sqrtf(vv / c2) + eps) * mask[e];
    w[e] = wn;
    if (wb) wb[e] = f2bf(wn);
}

__global__ void k_to_bf16(long n, const float* __restrict__ x, uint16_t* __restrict__ y) {
    const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e < n) y[e] = f2bf(x[e]);
}

// ---------------------------------------------------------------------------------------------
// host object
// ---------------------------------------------------------------------------------------------
#define MCHK(expr)                                                                                 \
    do {                                                                                           \
        hipError_t _e = (expr);                                                                    \
        if (_e != hipSuccess) {                                                                    \
            m->err = std::string(#expr) + ": " + hipGetErrorString(_e);                             \
            return ABD_ERR_HIP;                                                                    \
        }                                                                                          \
    } while (0)

template <class T>
static int malloc_dev(Mlp* m, T** p, size_t n) {
    n = std::max<size_t>(n, 1);
    hipError_t e = hipMalloc((void**)p, n * sizeof(T));
    if (e != hipSuccess) {
        m->err = std::string("hipMalloc: ") + hipGetErrorString(e);
        return ABD_ERR_HIP;
    }
    m->allocs.push_back((void*)*p);
    hipMemset(*p, 0, n * sizeof(T));
    return ABD_OK;
}

int mlp_reserve(Mlp* m, int rows) {
    rows = pad16(std::max(rows, 16));
    if (rows <= m->rows_cap) return ABD_OK;
    (void)hipStreamSynchronize(m->st);
    for (void* p : m->act_allocs) {
        auto it = std::find(m->allocs.begin(), m->allocs.end(), p);
        if (it != m->allocs.end()) m->allocs.erase(it);
        (void)hipFree(p);
    }
    m->act_allocs.clear();
    const int L = m->L;
    m->A.assign(L + 1, nullptr);
    m->Z.assign(L, nullptr);
    m->G.assign(L, nullptr);
    m->dA.assign(L, nullptr);
    int rc;
    for (int l = 0; l <= L; ++l) {
        if ((rc = malloc_dev(m, &m->A[l], (size_t)rows * m->dp[l])) != ABD_OK) return rc;
        m->act_allocs.push_back(m->A[l]);
    }
    for (int l = 0; l < L; ++l) {
        if ((rc = malloc_dev(m, &m->Z[l], (size_t)rows * m->dp[l + 1])) != ABD_OK) return rc;
        if ((rc = malloc_dev(m, &m->G[l], (size_t)rows * m->dp[l + 1])) != ABD_OK) return rc;
        if ((rc = malloc_dev(m, &m->dA[l], (size_t)rows * m->dp[l + 1])) != ABD_OK) return rc;
        m->act_allocs.push_back(m->Z[l]);
        m->act_allocs.push_back(m->G[l]);
        m->act_allocs.push_back(m->dA[l]);
    }
    if ((rc = malloc_dev(m, &m->T, (size_t)rows * m->dp[L])) != ABD_OK) return rc;
    if ((rc = malloc_dev(m, &m->row_loss, (size_t)rows)) != ABD_OK) return rc;
    if ((rc = malloc_dev(m, &m->idx, (size_t)rows)) != ABD_OK) return rc;
    m->act_allocs.push_back(m->T);
    m->act_allocs.push_back(m->row_loss);
    m->act_allocs.push_back(m->idx);
    m->rows_cap = rows;
    return ABD_OK;
}

// forward over the packed rows already in A[0] (rows padded to 16): Z[L-1] is the output
void mlp_forward_bufs(Mlp* m, int n, hipStream_t st, uint16_t* const* A, float* const* Z) {
    const int rows = pad16(std::max(n, 1));
    for (int l = 0; l < m->L; ++l) {
        gemm(st, rows, m->dp[l + 1], m->dp[l], A[l], m->dp[l], 1, m->Wb[l], m->dp[l + 1], 1, Z[l], m->dp[l + 1],
             m->b[l]);
        if (l + 1 < m->L) {
            const long tot = (long)rows * m->dp[l + 1];
            hipLaunchKernelGGL(k_relu_bf16, dim3(cdivm(tot, 256)), dim3(256), 0, st, tot, n, m->dp[l + 1], Z[l],
                               A[l + 1]);
        }
    }
}

void mlp_forward_dev(Mlp* m, int n, hipStream_t st) { mlp_forward_bufs(m, n, st, m->A.data(), m->Z.data()); }

// ---------------------------------------------------------------------------------------------
// Scratch of abd_mlp_forward_static: the staged input rows and every layer's activations, in one
// static device array that each GPU holds (at its own address) until the process ends.  Host code
// reaches it only through the address hipGetSymbolAddress gives on that device.
// ---------------------------------------------------------------------------------------------
constexpr size_t MLP_STATIC_BYTES = 64ull << 20;
static __device__ __align__(256) unsigned char scratch[MLP_STATIC_BYTES];

constexpr int MLP_MAX_DEVICES = 64;
static void* g_scratch_ptr[MLP_MAX_DEVICES];
static std::mutex g_scratch_mu;

// the current device's address of `scratch`, looked up once per device
static int scratch_address(Mlp* m, void** out) {
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess || dev < 0 || dev >= MLP_MAX_DEVICES) {
        m->err = "static scratch: no usable current device";
        return ABD_ERR_HIP;
    }
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* ptr = nullptr;
        e = hipGetSymbolAddress((void**)&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !ptr) {
            m->err = std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e);
            return ABD_ERR_HIP;
        }
        g_scratch_ptr[dev] = ptr;
    }
    m->dev = dev;
    *out = g_scratch_ptr[dev];
    return ABD_OK;
}

static size_t align256(size_t b) { return (b + 255) / 256 * 256; }

// bytes of the static scratch a forward of n rows uses: staged input, then A[0..L], then Z[0..L-1]
static size_t static_bytes(const Mlp* m, int n) {
    const size_t rows = (size_t)pad16(std::max(n, 1));
    size_t b = align256((size_t)n * m->d[0] * sizeof(float));
    for (int l = 0; l <= m->L; ++l) b += align256(rows * m->dp[l] * sizeof(uint16_t));
    for (int l = 0; l < m->L; ++l) b += align256(rows * m->dp[l + 1] * sizeof(float));
    return b;
}

// gradients of the batch loss into gW/gb (targets already packed in T); loss into *loss_dev
void mlp_backward_dev(Mlp* m, int n, float* loss_dev, hipStream_t st) {
    const int rows = pad16(std::max(n, 1));
    const int L = m->L;
    hipLaunchKernelGGL(k_loss_grad, dim3(cdivm(rows, 128)), dim3(128), 0, st, n, rows, m->dp[L], m->Z[L - 1], m->T,
                       m->colw, m->G[L - 1], m->row_loss);
    hipLaunchKernelGGL(k_sum, dim3(1), dim3(256), 0, st, rows, m->row_loss, loss_dev);
    for (int l = L - 1; l >= 0; --l) {
        // dW_l = A_l^T G_l  (dp[l] x dp[l+1]), K = rows
        gemm(st, m->dp[l], m->dp[l + 1], rows, m->A[l], 1, m->dp[l], m->G[l], m->dp[l + 1], 1, m->gW[l],
             m->dp[l + 1], nullptr);
        hipLaunchKernelGGL(k_colsum, dim3(cdivm(m->dp[l + 1], 16)), dim3(256), 0, st, rows, m->dp[l + 1], m->G[l],
                           m->gb[l]);
        if (l > 0) {
            // dA_l = G_l W_l^T (rows x dp[l]); G_{l-1} = dA_l * (Z_{l-1} > 0)
            gemm(st, rows, m->dp[l], m->dp[l + 1], m->G[l], m->dp[l + 1], 1, m->Wb[l], 1, m->dp[l + 1], m->dA[l - 1],
                 m->dp[l], nullptr);
            const long tot = (long)rows * m->dp[l];
            hipLaunchKernelGGL(k_relu_back, dim3(cdivm(tot, 256)), dim3(256), 0, st, tot, m->dA[l - 1], m->Z[l - 1],
                               m->G[l - 1]);
        }
    }
}

}  // namespace abd

using namespace abd;

extern "C" {

int abd_mlp_create(const int* dims, int n_layers, AbdMlp** out) {
    if (!dims || !out || n_layers < 1 || n_layers > 8) return ABD_ERR_ARG;
    Mlp* m = new Mlp();
    *out = m;
    m->L = n_layers;
    for (int l = 0; l <= n_layers; ++l) {
        if (dims[l] < 1 || dims[l] > 4096) {
            m->err = "layer widths must be in [1, 4096]";
            return ABD_ERR_ARG;
        }
        m->d.push_back(dims[l]);
        m->dp.push_back(pad16(dims[l]));
    }
    if (hipStreamCreateWithFlags(&m->st, hipStreamNonBlocking) != hipSuccess) {
        m->st = nullptr;
        m->err = "hipStreamCreateWithFlags failed";
        return ABD_ERR_HIP;
    }
    const int L = n_layers;
    m->W.assign(L, nullptr); m->Wb.assign(L, nullptr); m->b.assign(L, nullptr); m->gW.assign(L, nullptr);
    m->gb.assign(L, nullptr); m->mW.assign(L, nullptr); m->vW.assign(L, nullptr); m->mb.assign(L, nullptr);
    m->vb.assign(L, nullptr); m->maskW.assign(L, nullptr); m->maskb.assign(L, nullptr);
    int rc;
    for (int l = 0; l < L; ++l) {
        const size_t nw = (size_t)m->dp[l] * m->dp[l + 1], nbias = (size_t)m->dp[l + 1];
        if ((rc = malloc_dev(m, &m->W[l], nw)) || (rc = malloc_dev(m, &m->Wb[l], nw)) ||
            (rc = malloc_dev(m, &m->gW[l], nw)) || (rc = malloc_dev(m, &m->mW[l], nw)) ||
            (rc = malloc_dev(m, &m->vW[l], nw)) || (rc = malloc_dev(m, &m->maskW[l], nw)) ||
            (rc = malloc_dev(m, &m->b[l], nbias)) || (rc = malloc_dev(m, &m->gb[l], nbias)) ||
            (rc = malloc_dev(m, &m->mb[l], nbias)) || (rc = malloc_dev(m, &m->vb[l], nbias)) ||
            (rc = malloc_dev(m, &m->maskb[l], nbias)))
            return rc;
        std::vector<float> mw(nw, 0.f), mb(nbias, 0.f);
        for (int i = 0; i < m->d[l]; ++i)
            for (int j = 0; j < m->d[l + 1]; ++j) mw[(size_t)i * m->dp[l + 1] + j] = 1.f;
        for (int j = 0; j < m->d[l + 1]; ++j) mb[j] = 1.f;
        if (hipMemcpy(m->maskW[l], mw.data(), nw * 4, hipMemcpyHostToDevice) != hipSuccess ||
            hipMemcpy(m->maskb[l], mb.data(), nbias * 4, hipMemcpyHostToDevice) != hipSuccess) {
            m->err = "upload failed";
            return ABD_ERR_HIP;
        }
    }
    if ((rc = malloc_dev(m, &m->colw, (size_t)m->dp[L])) != ABD_OK) return rc;
    if ((rc = malloc_dev(m, &m->loss_dev, 4096)) != ABD_OK) return rc;
    if ((rc = scratch_address(m, &m->scratch)) != ABD_OK) return rc;
    return mlp_reserve(m, 1024);
}

void abd_mlp_destroy(AbdMlp* m) {
    if (!m) return;
    if (m->st) (void)hipStreamSynchronize(m->st);
    for (void* p : m->allocs) (void)hipFree(p);
    if (m->st) (void)hipStreamDestroy(m->st);
    delete m;
}

const char* abd_mlp_last_error(const AbdMlp* m) { return m ? m->err.c_str() : "null model"; }

int abd_mlp_num_params(const AbdMlp* m) {
    if (!m) return ABD_ERR_ARG;
    int n = 0;
    for (int l = 0; l < m->L; ++l) n += m->d[l] * m->d[l + 1] + m->d[l + 1];
    return n;
}

// flat layout per layer: W (d_l x d_{l+1}, row-major) then b (d_{l+1})
int abd_mlp_set_params(AbdMlp* m, const float* flat) {
    if (!m || !flat) return ABD_ERR_ARG;
    size_t off = 0;
    for (int l = 0; l < m->L; ++l) {
        std::vector<float> w((size_t)m->dp[l] * m->dp[l + 1], 0.f), b(m->dp[l + 1], 0.f);
        for (int i = 0; i < m->d[l]; ++i)
            for (int j = 0; j < m->d[l + 1]; ++j) w[(size_t)i * m->dp[l + 1] + j] = flat[off + (size_t)i * m->d[l + 1] + j];
        off += (size_t)m->d[l] * m->d[l + 1];
        for (int j = 0; j < m->d[l + 1]; ++j) b[j] = flat[off + j];
        off += m->d[l + 1];
        MCHK(hipMemcpy(m->W[l], w.data(), w.size() * 4, hipMemcpyHostToDevice));
        MCHK(hipMemcpy(m->b[l], b.data(), b.size() * 4, hipMemcpyHostToDevice));
        MCHK(hipMemset(m->mW[l], 0, w.size() * 4)); MCHK(hipMemset(m->vW[l], 0, w.size() * 4));
        MCHK(hipMemset(m->mb[l], 0, b.size() * 4)); MCHK(hipMemset(m->vb[l], 0, b.size() * 4));
        const long nw = (long)w.size();
        hipLaunchKernelGGL(k_to_bf16, dim3(cdivm(nw, 256)), dim3(256), 0, m->st, nw, m->W[l], m->Wb[l]);
    }
    m->adam_t = 0;
    MCHK(hipStreamSynchronize(m->st));
    return ABD_OK;
}

int abd_mlp_get_params(const AbdMlp* cm, float* flat) {
    AbdMlp* m = const_cast<AbdMlp*>(cm);
    if (!m || !flat) return ABD_ERR_ARG;
    MCHK(hipStreamSynchronize(m->st));
    size_t off = 0;
    for (int l = 0; l < m->L; ++l) {
        std::vector<float> w((size_t)m->dp[l] * m->dp[l + 1]), b(m->dp[l + 1]);
        MCHK(hipMemcpy(w.data(), m->W[l], w.size() * 4, hipMemcpyDeviceToHost));
        MCHK(hipMemcpy(b.data(), m->b[l], b.size() * 4, hipMemcpyDeviceToHost));
        for (int i = 0; i < m->d[l]; ++i)
            for (int j = 0; j < m->d[l + 1]; ++j) flat[off + (size_t)i * m->d[l + 1] + j] = w[(size_t)i * m->dp[l + 1] + j];
        off += (size_t)m->d[l] * m->d[l + 1];
        for (int j = 0; j < m->d[l + 1]; ++j) flat[off + j] = b[j];
        off += m->d[l + 1];
    }
    return ABD_OK;
}

int abd_mlp_set_loss_weights(AbdMlp* m, const float* w) {
    if (!m || !w) return ABD_ERR_ARG;
    std::vector<float> cw(m->dp[m->L], 0.f);
    for (int j = 0; j < m->d[m->L]; ++j) cw[j] = w[j];
    MCHK(hipMemcpy(m->colw, cw.data(), cw.size() * 4, hipMemcpyHostToDevice));
    return ABD_OK;
}

// host X (n x d_0) -> host Y (n x d_L)
int abd_mlp_forward(AbdMlp* m, const float* X, int n, float* Y) {
    if (!m || !X || !Y || n < 1) return ABD_ERR_ARG;
    int rc = mlp_reserve(m, n);
    if (rc != ABD_OK) return rc;
    const int rows = pad16(n);
    float* dX = nullptr;
    MCHK(hipMalloc(&dX, (size_t)n * m->d[0] * 4));
    MCHK(hipMemcpyAsync(dX, X, (size_t)n * m->d[0] * 4, hipMemcpyHostToDevice, m->st));
    const long tot = (long)rows * m->dp[0];
    hipLaunchKernelGGL(k_pack_rows, dim3(cdivm(tot, 256)), dim3(256), 0, m->st, n, rows, m->d[0], m->dp[0], dX,
                       (const int*)nullptr, m->A[0]);
    mlp_forward_dev(m, n, m->st);
    std::vector<float> z((size_t)rows * m->dp[m->L]);
    MCHK(hipMemcpyAsync(z.data(), m->Z[m->L - 1], z.size() * 4, hipMemcpyDeviceToHost, m->st));
    MCHK(hipStreamSynchronize(m->st));
    (void)hipFree(dX);
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < m->d[m->L]; ++j) Y[(size_t)i * m->d[m->L] + j] = z[(size_t)i * m->dp[m->L] + j];
    return ABD_OK;
}

// abd_mlp_forward without any allocation: the staged input and the activations live in the
// device's static scratch, the output rows go from the last pre-activation straight to Y
int abd_mlp_forward_static(AbdMlp* m, const float* X, int n, float* Y) {
    if (!m || !X || !Y || n < 1) return ABD_ERR_ARG;
    if (!m->scratch) {
        m->err = "static scratch address unknown (model not created)";
        return ABD_ERR_STATE;
    }
    int dev = -1;
    if (hipGetDevice(&dev) != hipSuccess || dev != m->dev) {
        m->err = "abd_mlp_forward_static: the current device is not the model's";
        return ABD_ERR_STATE;
    }
    const size_t need = static_bytes(m, n);
    if (need > MLP_STATIC_BYTES) {
        int hi = n;
        while (hi > 1 && static_bytes(m, hi) > MLP_STATIC_BYTES) hi /= 2;
        int lo = hi;
        while (static_bytes(m, lo + 1) <= MLP_STATIC_BYTES) ++lo;
        m->err = "abd_mlp_forward_static: " + std::to_string(n) + " rows need " + std::to_string(need >> 10) +
                 " KiB of static scratch (" + std::to_string(MLP_STATIC_BYTES >> 20) + " MiB): at most " +
                 std::to_string(lo) + " rows for this model";
        return ABD_ERR_CAPACITY;
    }
    const int L = m->L, rows = pad16(n);
    unsigned char* base = (unsigned char*)m->scratch;
    size_t off = 0;
    float* dX = (float*)(base + off);
    off += align256((size_t)n * m->d[0] * sizeof(float));
    uint16_t* A[9];
    float* Z[8];
    for (int l = 0; l <= L; ++l) {
        A[l] = (uint16_t*)(base + off);
        off += align256((size_t)rows * m->dp[l] * sizeof(uint16_t));
    }
    for (int l = 0; l < L; ++l) {
        Z[l] = (float*)(base + off);
        off += align256((size_t)rows * m->dp[l + 1] * sizeof(float));
    }
    MCHK(hipMemcpyAsync(dX, X, (size_t)n * m->d[0] * sizeof(float), hipMemcpyHostToDevice, m->st));
    const long tot = (long)rows * m->dp[0];
    hipLaunchKernelGGL(k_pack_rows, dim3(cdivm(tot, 256)), dim3(256), 0, m->st, n, rows, m->d[0], m->dp[0], dX,
                       (const int*)nullptr, A[0]);
    mlp_forward_bufs(m, n, m->st, A, Z);
    MCHK(hipGetLastError());
    MCHK(hipMemcpy2DAsync(Y, (size_t)m->d[L] * sizeof(float), Z[L - 1], (size_t)m->dp[L] * sizeof(float),
                          (size_t)m->d[L] * sizeof(float), (size_t)n, hipMemcpyDeviceToHost, m->st));
    MCHK(hipStreamSynchronize(m->st));
    return ABD_OK;
}

// loss and gradients (flat layout as abd_mlp_get_params) of one batch, without updating
int abd_mlp_gradient(AbdMlp* m, const float* X, const float* T, int n, float* loss, float* grad_flat) {
    if (!m || !X || !T || n < 1) return ABD_ERR_ARG;
    int rc = mlp_reserve(m, n);
    if (rc != ABD_OK) return rc;
    const int rows = pad16(n), L = m->L;
    float *dX = nullptr, *dT = nullptr;
    MCHK(hipMalloc(&dX, (size_t)n * m->d[0] * 4));
    MCHK(hipMalloc(&dT, (size_t)n * m->d[L] * 4));
    MCHK(hipMemcpyAsync(dX, X, (size_t)n * m->d[0] * 4, hipMemcpyHostToDevice, m->st));
    MCHK(hipMemcpyAsync(dT, T, (size_t)n * m->d[L] * 4, hipMemcpyHostToDevice, m->st));
    long tot = (long)rows * m->dp[0];
    hipLaunchKernelGGL(k_pack_rows, dim3(cdivm(tot, 256)), dim3(256), 0, m->st, n, rows, m->d[0], m->dp[0], dX,
                       (const int*)nullptr, m->A[0]);
    tot = (long)rows * m->dp[L];
    hipLaunchKernelGGL(k_pack_rows_f32, dim3(cdivm(tot, 256)), dim3(256), 0, m->st, n, rows, m->d[L], m->dp[L], dT,
                       (const int*)nullptr, m->T);
    mlp_forward_dev(m, n, m->st);
    mlp_backward_dev(m, n, m->loss_dev, m->st);
    float lh = 0.f;
    MCHK(hipMemcpyAsync(&lh, m->loss_dev, 4, hipMemcpyDeviceToHost, m->st));
    MCHK(hipStreamSynchronize(m->st));
    (void)hipFree(dX);
    (void)hipFree(dT);
    if (loss) *loss = lh;
    if (grad_flat) {
        size_t off = 0;
        for (int l = 0; l < L; ++l) {
            std::vector<float> w((size_t)m->dp[l] * m->dp[l + 1]), b(m->dp[l + 1]);
            MCHK(hipMemcpy(w.data(), m->gW[l], w.size() * 4, hipMemcpyDeviceToHost));
            MCHK(hipMemcpy(b.data(), m->gb[l], b.size() * 4, hipMemcpyDeviceToHost));
            for (int i = 0; i < m->d[l]; ++i)
                for (int j = 0; j < m->d[l + 1]; ++j)
                    grad_flat[off + (size_t)i * m->d[l + 1] + j] = w[(size_t)i * m->dp[l + 1] + j];
            off += (size_t)m->d[l] * m->d[l + 1];
            for (int j = 0; j < m->d[l + 1]; ++j) grad_flat[off + j] = b[j];
            off += m->d[l + 1];
        }
    }
    return ABD_OK;
}

// Adam training on the device: `steps` minibatches of `batch` rows; perm (steps*batch row ids of
// the n training rows, chosen by the caller) fixes the order.  losses[steps] receives the batch
// losses before each update.
int abd_mlp_train(AbdMlp* m, const float* X, const float* T, int n, const int* perm, int steps, int batch, float lr,
                  float beta1, float beta2, float eps, float* losses) {
    if (!m || !X || !T || !perm || n < 1 || steps < 1 || batch < 1 || steps > 4096) return ABD_ERR_ARG;
    int rc = mlp_reserve(m, batch);
    if (rc != ABD_OK) return rc;
    const int L = m->L, rows = pad16(batch);
    for (long i = 0; i < (long)steps * batch; ++i)
        if (perm[i] < 0 || perm[i] >= n) {
            m->err = "batch row index out of range";
            return ABD_ERR_ARG;
        }
    float *dX = nullptr, *dT = nullptr;
    int* dperm = nullptr;
    MCHK(hipMalloc(&dX, (size_t)n * m->d[0] * 4));
    MCHK(hipMalloc(&dT, (size_t)n * m->d[L] * 4));
    MCHK(hipMalloc(&dperm, (size_t)steps * batch * 4));
    MCHK(hipMemcpyAsync(dX, X, (size_t)n * m->d[0] * 4, hipMemcpyHostToDevice, m->st));
    MCHK(hipMemcpyAsync(dT, T, (size_t)n * m->d[L] * 4, hipMemcpyHostToDevice, m->st));
    MCHK(hipMemcpyAsync(dperm, perm, (size_t)steps * batch * 4, hipMemcpyHostToDevice, m->st));
    for (int s = 0; s < steps; ++s) {
        const int* idx = dperm + (size_t)s * batch;
        long tot = (long)rows * m->dp[0];
        hipLaunchKernelGGL(k_pack_rows, dim3(cdivm(tot, 256)), dim3(256), 0, m->st, batch, rows, m->d[0], m->dp[0], dX,
                           idx, m->A[0]);
        tot = (long)rows * m->dp[L];
        hipLaunchKernelGGL(k_pack_rows_f32, dim3(cdivm(tot, 256)), dim3(256), 0, m->st, batch, rows, m->d[L], m->dp[L],
                           dT, idx, m->T);
        mlp_forward_dev(m, batch, m->st);
        mlp_backward_dev(m, batch, m->loss_dev + s, m->st);
        m->adam_t += 1;
        const float c1 = 1.f - std::pow(beta1, (float)m->adam_t), c2 = 1.f - std::pow(beta2, (float)m->adam_t);
        for (int l = 0; l < L; ++l) {
            const long nw = (long)m->dp[l] * m->dp[l + 1], nbias = m->dp[l + 1];
            hipLaunchKernelGGL(k_adam, dim3(cdivm(nw, 256)), dim3(256), 0, m->st, nw, m->W[l], m->gW[l], m->maskW[l],
                               m->mW[l], m->vW[l], lr, beta1, beta2, eps, c1, c2, m->Wb[l]);
            hipLaunchKernelGGL(k_adam, dim3(cdivm(nbias, 256)), dim3(256), 0, m->st, nbias, m->b[l], m->gb[l],
                               m->maskb[l], m->mb[l], m->vb[l], lr, beta1, beta2, eps, c1, c2, (uint16_t*)nullptr);
        }
    }
    MCHK(hipGetLastError());
    if (losses) MCHK(hipMemcpyAsync(losses, m->loss_dev, (size_t)steps * 4, hipMemcpyDeviceToHost, m->st));
    MCHK(hipStreamSynchronize(m->st));
    (void)hipFree(dX);
    (void)hipFree(dT);
    (void)hipFree(dperm);
    return ABD_OK;
}

// raw matrix product for the layout test: C (M x N) = A (M x K) B (K x N), BF16 inputs given as
// FP32 (rounded to nearest even), M, N, K multiples of 16
int abd_mlp_gemm_test(AbdMlp* m, int M, int N, int K, const float* A, const float* B, float* C) {
    if (!m || !A || !B || !C || M % 16 || N % 16 || K % 16 || M < 16 || N < 16 || K < 16) return ABD_ERR_ARG;
    float *dA = nullptr, *dB = nullptr, *dC = nullptr;
    uint16_t *hA = nullptr, *hB = nullptr;
    MCHK(hipMalloc(&dA, (size_t)M * K * 4)); MCHK(hipMalloc(&dB, (size_t)K * N * 4));
    MCHK(hipMalloc(&dC, (size_t)M * N * 4));
    MCHK(hipMalloc(&hA, (size_t)M * K * 2)); MCHK(hipMalloc(&hB, (size_t)K * N * 2));
    MCHK(hipMemcpy(dA, A, (size_t)M * K * 4, hipMemcpyHostToDevice));
    MCHK(hipMemcpy(dB, B, (size_t)K * N * 4, hipMemcpyHostToDevice));
    hipLaunchKernelGGL(k_to_bf16, dim3(cdivm((long)M * K, 256)), dim3(256), 0, m->st, (long)M * K, dA, hA);
    hipLaunchKernelGGL(k_to_bf16, dim3(cdivm((long)K * N, 256)), dim3(256), 0, m->st, (long)K * N, dB, hB);
    gemm(m->st, M, N, K, hA, K, 1, hB, N, 1, dC, N, nullptr);
    MCHK(hipStreamSynchronize(m->st));
    MCHK(hipMemcpy(C, dC, (size_t)M * N * 4, hipMemcpyDeviceToHost));
    (void)hipFree(dA); (void)hipFree(dB); (void)hipFree(dC); (void)hipFree(hA); (void)hipFree(hB);
    return ABD_OK;
}

}  // extern "C"
