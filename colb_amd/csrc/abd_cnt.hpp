// abd_cnt.hpp — list lengths that live on the device.
#pragma once
#include <hip/hip_runtime.h>

namespace abd {

// Length of a list that an earlier launch on the same stream computed on the device: kernels are
// sized by `cap` (a bound the host knows) and read the actual length from `dev` (when non-null)
// at run time, so no host round trip sits between the producer and the consumer.
struct Cnt {
    int cap;
    const int* dev;
};
inline Cnt host_cnt(int n) { return Cnt{n, nullptr}; }
__device__ __forceinline__ int cnt_of(const Cnt& c) {
    if (!c.dev) return c.cap;
    const int n = *c.dev;
    return n < 0 ? 0 : (n > c.cap ? c.cap : n);
}

// the longest list the single-launch counting-rank sort takes (abd_sort.hip)
constexpr int SORT_SMALL_MAX_KEYS = 8192;
// line-search trials evaluated by one set of launches (the vertex and candidate-energy buffers of
// the trials hold this many slices)
constexpr int LS_TRIALS_MAX = 8;

}  // namespace abd
