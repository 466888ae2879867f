// Host stand-ins for the rocPRIM device algorithms used by colb_amd/csrc (TEST ONLY).
#pragma once
#include <algorithm>
#include <numeric>
#include <vector>
#include "../../emu_hip.hpp"
namespace rocprim {
template <class T> struct plus { T operator()(const T& a, const T& b) const { return a + b; } };
template <class K>
inline hipError_t radix_sort_keys(void* tmp, size_t& bytes, const K* in, K* out, unsigned n, unsigned b0, unsigned b1, hipStream_t = 0) {
    if (!tmp) { bytes = 16; return hipSuccess; }
    std::vector<K> v(in, in + n);
    K mask = (b1 >= 8 * sizeof(K)) ? ~K(0) : ((K(1) << b1) - 1);
    mask &= ~((K(1) << b0) - 1);
    std::stable_sort(v.begin(), v.end(), [&](K a, K b) { return (a & mask) < (b & mask); });
    std::copy(v.begin(), v.end(), out);
    return hipSuccess;
}
template <class K, class V>
inline hipError_t radix_sort_pairs(void* tmp, size_t& bytes, const K* kin, K* kout, const V* vin, V* vout, unsigned n, unsigned b0, unsigned b1, hipStream_t = 0) {
    if (!tmp) { bytes = 16; return hipSuccess; }
    std::vector<unsigned> idx(n);
    std::iota(idx.begin(), idx.end(), 0u);
    K mask = (b1 >= 8 * sizeof(K)) ? ~K(0) : ((K(1) << b1) - 1);
    mask &= ~((K(1) << b0) - 1);
    std::stable_sort(idx.begin(), idx.end(), [&](unsigned a, unsigned b) { return (kin[a] & mask) < (kin[b] & mask); });
    std::vector<K> k2(n); std::vector<V> v2(n);
    for (unsigned i = 0; i < n; ++i) { k2[i] = kin[idx[i]]; v2[i] = vin[idx[i]]; }
    std::copy(k2.begin(), k2.end(), kout); std::copy(v2.begin(), v2.end(), vout);
    return hipSuccess;
}
template <class T, class I, class Op>
inline hipError_t exclusive_scan(void* tmp, size_t& bytes, const T* in, T* out, I init, size_t n, Op op, hipStream_t = 0) {
    if (!tmp) { bytes = 16; return hipSuccess; }
    T acc = (T)init;
    for (size_t i = 0; i < n; ++i) { T x = in[i]; out[i] = acc; acc = op(acc, x); }
    return hipSuccess;
}
}  // namespace rocprim
