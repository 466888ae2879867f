// abd_sort.hip — stable LSD radix sort of 64-bit keys (optionally carrying int values) and an
// exclusive prefix sum of ints, for the candidate, incidence and body-pair lists.
//
// Every length is a Cnt: grids are sized by its host bound and the kernels read the actual length
// on the device, so a sort or scan can follow the kernel that counted its input with no host
// round trip in between.
//
// Radix sort: 4-bit digits over the bits the caller says may be set (a mask: one pass per 4-bit
// window that holds a significant bit, windows placed greedily from the lowest such bit, so a key
// made of a few packed fields costs the passes of its fields' widths only); up to SORT_SMALL_MAX
// keys are ranked directly (k_sort_small).  Per pass, each 256-thread block owns a tile of 2048 consecutive
// elements (8 per thread, in order); it counts digits (k_radix_hist), the digit-major histogram
// [digit][block] is prefix-summed, and the block re-reads its tile, ranks every element among
// the equal digits of earlier threads and earlier items (a block-wide scan over a [digit][thread]
// count matrix in LDS) and scatters it.  Each pass is stable, so the sort is stable.
#include <hip/hip_runtime.h>
#include <algorithm>
#include <stdint.h>

#include "abd_cnt.hpp"
#include "abd_device.hpp"

namespace abd {

#define SORT_T 256
#define SORT_I 8
#define SORT_TILE (SORT_T * SORT_I)
#define RADIX 16

static inline int cdivs(long a, int b) { return (int)((a + b - 1) / b); }

// --------------------------------------------------------------------------------------------
// exclusive scan of int (n <= SORT_TILE * SORT_TILE)
// --------------------------------------------------------------------------------------------
__device__ __forceinline__ int block_excl_scan256(int v, int* sh) {
    const int t = threadIdx.x;
    sh[t] = v;
    __syncthreads();
    for (int o = 1; o < SORT_T; o <<= 1) {
        int a = (t >= o) ? sh[t - o] : 0;
        __syncthreads();
        sh[t] += a;
        __syncthreads();
    }
    int incl = sh[t];
    __syncthreads();
    return incl - v;
}

__global__ __launch_bounds__(SORT_T) void k_scan_tile_sums(Cnt cn, const int* __restrict__ in, int* __restrict__ bsum) {
    __shared__ int sh[SORT_T];
    const int n = cnt_of(cn);
    const int base = blockIdx.x * SORT_TILE + threadIdx.x * SORT_I;
    int s = 0;
    for (int i = 0; i < SORT_I; ++i)
        if (base + i < n) s += in[base + i];
    int ex = block_excl_scan256(s, sh);
    if (threadIdx.x == SORT_T - 1) bsum[blockIdx.x] = ex + s;
}

// single block: exclusive scan of nb <= SORT_TILE block sums, in place
__global__ __launch_bounds__(SORT_T) void k_scan_block_sums(int nb, int* bsum) {
    __shared__ int sh[SORT_T];
    const int base = threadIdx.x * SORT_I;
    int v[SORT_I];
    int s = 0;
    for (int i = 0; i < SORT_I; ++i) {
        v[i] = (base + i < nb) ? bsum[base + i] : 0;
        s += v[i];
    }
    int ex = block_excl_scan256(s, sh);
    for (int i = 0; i < SORT_I; ++i) {
        if (base + i < nb) bsum[base + i] = ex;
        ex += v[i];
    }
}

// bsum == nullptr: a single tile (n <= SORT_TILE), offset 0
// `in` and `out` may be the same array (each element is read before the thread writes it)
__global__ __launch_bounds__(SORT_T) void k_scan_apply(Cnt cn, const int* in, const int* __restrict__ bsum, int* out) {
    __shared__ int sh[SORT_T];
    const int n = cnt_of(cn);
    if (blockIdx.x * SORT_TILE >= n) return;  // whole workgroup: no barrier is skipped by part of it
    const int base = blockIdx.x * SORT_TILE + threadIdx.x * SORT_I;
    int v[SORT_I];
    int s = 0;
    for (int i = 0; i < SORT_I; ++i) {
        v[i] = (base + i < n) ? in[base + i] : 0;
        s += v[i];
    }
    int ex = block_excl_scan256(s, sh) + (bsum ? bsum[blockIdx.x] : 0);
    for (int i = 0; i < SORT_I; ++i) {
        if (base + i < n) out[base + i] = ex;
        ex += v[i];
    }
}

// scratch: the tile sums of every level (cdiv(n, SORT_TILE), then recursively of those)
int scan_scratch_ints(long n) {
    const long nb = cdivs(n, SORT_TILE);
    return (int)(nb + 1 + (nb > SORT_TILE ? scan_scratch_ints(nb) : 0));
}

// exclusive scan of n ints; levels above the first are scanned recursively (in place).  Tile sums
// of tiles past the actual length are zero, so the upper levels need only the host bound.
hipError_t excl_scan_int(hipStream_t s, const int* in, int* out, Cnt n, int* bsum) {
    if (n.cap <= 0) return hipSuccess;
    const int nb = cdivs(n.cap, SORT_TILE);
    if (nb == 1) {
        hipLaunchKernelGGL(k_scan_apply, dim3(1), dim3(SORT_T), 0, s, n, in, (const int*)nullptr, out);
        return hipGetLastError();
    }
    hipLaunchKernelGGL(k_scan_tile_sums, dim3(nb), dim3(SORT_T), 0, s, n, in, bsum);
    if (nb <= SORT_TILE) {
        hipLaunchKernelGGL(k_scan_block_sums, dim3(1), dim3(SORT_T), 0, s, nb, bsum);
    } else {
        hipError_t e = excl_scan_int(s, bsum, bsum, host_cnt(nb), bsum + nb + 1);
        if (e != hipSuccess) return e;
    }
    hipLaunchKernelGGL(k_scan_apply, dim3(nb), dim3(SORT_T), 0, s, n, in, bsum, out);
    return hipGetLastError();
}

// --------------------------------------------------------------------------------------------
// radix sort
// --------------------------------------------------------------------------------------------
// lengths of at most SORT_SMALL_MAX are sorted by k_sort_small; the radix kernels then do nothing
#define SORT_SMALL_MAX 8192
static_assert(SORT_SMALL_MAX == SORT_SMALL_MAX_KEYS, "the small-sort bound the assembly checks");
__global__ __launch_bounds__(SORT_T) void k_radix_hist(Cnt cn, const uint64_t* __restrict__ keys, int shift, int nblk,
                                                        int* __restrict__ hist) {
    __shared__ int cnt[RADIX];
    const int n = cnt_of(cn);
    if (n <= SORT_SMALL_MAX) return;
    if (threadIdx.x < RADIX) cnt[threadIdx.x] = 0;
    __syncthreads();
    const int base = blockIdx.x * SORT_TILE + threadIdx.x * SORT_I;
    for (int i = 0; i < SORT_I; ++i)
        if (base + i < n) atomicAdd(&cnt[opaque((int)((keys[base + i] >> shift) & (RADIX - 1)))], 1);
    __syncthreads();
    if (threadIdx.x < RADIX) hist[threadIdx.x * nblk + blockIdx.x] = cnt[threadIdx.x];
}

__global__ __launch_bounds__(SORT_T) void k_radix_scatter(Cnt cn, const uint64_t* __restrict__ kin,
                                                           const int* __restrict__ vin, uint64_t* __restrict__ kout,
                                                           int* __restrict__ vout, int shift, int nblk,
                                                           const int* __restrict__ hoff) {
    __shared__ int m[RADIX * SORT_T];
    __shared__ int sh[SORT_T];
    const int n = cnt_of(cn);
    if (n <= SORT_SMALL_MAX) return;
    const int t = threadIdx.x;
    const int base = blockIdx.x * SORT_TILE + t * SORT_I;
    uint64_t k[SORT_I];
    int dg[SORT_I];
    int c[RADIX];
    for (int d = 0; d < RADIX; ++d) c[d] = 0;
    for (int i = 0; i < SORT_I; ++i) {
        if (base + i < n) {
            k[i] = kin[base + i];
            dg[i] = opaque((int)((k[i] >> shift) & (RADIX - 1)));
            c[dg[i]]++;
        } else {
            dg[i] = -1;
        }
    }
    for (int d = 0; d < RADIX; ++d) m[d * SORT_T + t] = c[d];
    __syncthreads();
    // block exclusive scan over the digit-major matrix: thread t owns entries [16t, 16t+16)
    int run[RADIX];
    int s = 0;
    for (int j = 0; j < RADIX; ++j) {
        run[j] = m[t * RADIX + j];
        s += run[j];
    }
    int ex = block_excl_scan256(s, sh);
    for (int j = 0; j < RADIX; ++j) {
        m[t * RADIX + j] = ex;
        ex += run[j];
    }
    __syncthreads();
    // m[d*T + t] = items of the block with digit < d + items with digit d in threads < t
    int seen[RADIX];
    for (int d = 0; d < RADIX; ++d) seen[d] = 0;
    for (int i = 0; i < SORT_I; ++i) {
        int d = dg[i];
        if (d < 0) continue;
        int rank = m[d * SORT_T + t] - m[d * SORT_T] + seen[d];
        seen[d]++;
        int pos = hoff[d * nblk + blockIdx.x] + rank;
        kout[pos] = k[i];
        if (vin) vout[pos] = vin[base + i];
    }
}

// Whole sort of n <= SORT_SMALL_MAX keys by counting ranks: the final position of key i is
// #{j : k_j < k_i} + #{j < i : k_j == k_i} (stable).  Every workgroup holds all n keys in LDS and
// ranks 16 of them, sixteen lanes per key, lane l counting over keys l, l + 16, ... (n / 16
// comparisons per lane; at one step a wave reads 16 consecutive keys, the same for its 4 keys).
// Lengths above SORT_SMALL_MAX leave it to the radix passes.
#define SORT_SMALL_KEYS 16
__device__ __forceinline__ void sort_small_block(Cnt cn, const uint64_t* __restrict__ kin, const int* __restrict__ vin,
                                                 uint64_t* __restrict__ kout, int* __restrict__ vout, uint64_t* kb) {
    const int n = cnt_of(cn);
    if (n > SORT_SMALL_MAX || blockIdx.x * SORT_SMALL_KEYS >= n) return;  // whole workgroup
    const int t = threadIdx.x;
    for (int i = t; i < n; i += SORT_T) kb[i] = kin[i];
    __syncthreads();
    const int i = blockIdx.x * SORT_SMALL_KEYS + t / 16, part = t % 16;
    const uint64_t k = i < n ? kb[i] : 0;
    int rank = 0;
    for (int j = part; j < n; j += 16) {
        const uint64_t o = kb[j];
        rank += (o < k || (o == k && j < i)) ? 1 : 0;
    }
#pragma unroll
    for (int o = 1; o < 16; o <<= 1) rank += __shfl_xor(rank, o, 64);
    if (i < n && part == 0) {
        rank = opaque(rank);
        kout[rank] = k;
        if (vout) vout[rank] = vin[i];
    }
}

__global__ __launch_bounds__(SORT_T) void k_sort_small(Cnt cn, const uint64_t* __restrict__ kin,
                                                        const int* __restrict__ vin, uint64_t* __restrict__ kout,
                                                        int* __restrict__ vout) {
    __shared__ uint64_t kb[SORT_SMALL_MAX];
    sort_small_block(cn, kin, vin, kout, vout, kb);
}

// two independent key lists in one launch (blockIdx.y picks the list)
__global__ __launch_bounds__(SORT_T) void k_sort_small2(Cnt na, const uint64_t* __restrict__ ka,
                                                         uint64_t* __restrict__ oa, Cnt nb,
                                                         const uint64_t* __restrict__ kb_in, uint64_t* __restrict__ ob) {
    __shared__ uint64_t kb[SORT_SMALL_MAX];
    if (blockIdx.y == 0)
        sort_small_block(na, ka, nullptr, oa, nullptr, kb);
    else
        sort_small_block(nb, kb_in, nullptr, ob, nullptr, kb);
}

// two key lists of at most SORT_SMALL_MAX keys each (host bounds), sorted by one launch
hipError_t small_sort2(hipStream_t s, const uint64_t* ka, uint64_t* oa, Cnt na, const uint64_t* kb, uint64_t* ob,
                       Cnt nb) {
    const int cap = std::max(na.cap, nb.cap);
    if (cap <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_sort_small2, dim3(cdivs(cap, SORT_SMALL_KEYS), 2), dim3(SORT_T), 0, s, na, ka, oa, nb, kb,
                       ob);
    return hipGetLastError();
}

// scratch for radix_sort64: tmp keys (n), tmp vals (n, if values), hist 2*RADIX*nblk ints + scan
size_t radix_scratch_bytes(long n) {
    long nblk = cdivs(n, SORT_TILE);
    return (size_t)n * 8 + (size_t)n * 4 + (size_t)(2 * RADIX * nblk + scan_scratch_ints(RADIX * nblk) + 64) * 4 + 1024;
}

// the 4-bit windows (their lowest bit) that cover every set bit of `sig`, lowest first
static int radix_windows(uint64_t sig, int* shift) {
    int n = 0;
    for (int b = 0; b < 64;) {
        if ((sig >> b) == 0) break;
        if (((sig >> b) & 1ull) == 0) { ++b; continue; }
        shift[n++] = b;
        b += 4;
    }
    return n;
}

// copies the first n keys (and values) when no key bit can differ
__global__ void k_copy_keys(Cnt cn, const uint64_t* __restrict__ kin, const int* __restrict__ vin,
                            uint64_t* __restrict__ kout, int* __restrict__ vout) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= cnt_of(cn)) return;
    kout[i] = kin[i];
    if (vin) vout[i] = vin[i];
}

// n.cap <= SORT_SMALL_MAX: the counting-rank sort alone.  Larger bounds launch both forms; each
// kernel reads the actual length and the form that does not apply returns at once.
hipError_t radix_sort64(hipStream_t s, const uint64_t* kin, uint64_t* kout, const int* vin, int* vout, Cnt n,
                        uint64_t sig, void* scratch) {
    if (n.cap <= 0) return hipSuccess;
    const int nblk = cdivs(n.cap, SORT_TILE);
    char* p = (char*)scratch;
    uint64_t* ktmp = (uint64_t*)p;
    p += (size_t)n.cap * 8;
    int* vtmp = (int*)p;
    p += (size_t)n.cap * 4;
    int* hist = (int*)p;
    int* hoff = hist + RADIX * nblk;
    int* bsum = hoff + RADIX * nblk;
    int shifts[64];
    const int passes = radix_windows(sig, shifts);
    if (passes == 0) {
        hipLaunchKernelGGL(k_copy_keys, dim3(cdivs(n.cap, 256)), dim3(256), 0, s, n, kin, vin, kout,
                           vin ? vout : nullptr);
        return hipGetLastError();
    }
    hipLaunchKernelGGL(k_sort_small, dim3(cdivs(std::min(n.cap, SORT_SMALL_MAX), SORT_SMALL_KEYS)), dim3(SORT_T), 0,
                       s, n, kin, vin, kout, vin ? vout : nullptr);
    if (n.cap <= SORT_SMALL_MAX) return hipGetLastError();
    const uint64_t* ks = kin;
    const int* vs = vin;
    for (int pass = 0; pass < passes; ++pass) {
        // the last pass writes the output buffers
        bool to_out = ((passes - 1 - pass) % 2) == 0;
        uint64_t* kd = to_out ? kout : ktmp;
        int* vd = vin ? (to_out ? vout : vtmp) : nullptr;
        const int shift = shifts[pass];
        hipLaunchKernelGGL(k_radix_hist, dim3(nblk), dim3(SORT_T), 0, s, n, ks, shift, nblk, hist);
        // the histogram scan sees stale entries when the radix form does not apply; its result is
        // then unused
        hipError_t e = excl_scan_int(s, hist, hoff, host_cnt(RADIX * nblk), bsum);
        if (e != hipSuccess) return e;
        hipLaunchKernelGGL(k_radix_scatter, dim3(nblk), dim3(SORT_T), 0, s, n, ks, vs, kd, vd, shift, nblk, hoff);
        ks = kd;
        vs = vd;
    }
    return hipGetLastError();
}

}  // namespace abd
