// Emulator driver for the radix sort and scan of colb_amd/csrc/abd_sort.hip (TEST ONLY).
#include <algorithm>
#include <cstdio>
#include <numeric>
#include <random>
#include <vector>
#include "emu_hip.hpp"
namespace abd {
hipError_t excl_scan_int(hipStream_t s, const int* in, int* out, int n, int* bsum);
int scan_scratch_ints(long n);
size_t radix_scratch_bytes(long n);
hipError_t radix_sort64(hipStream_t s, const uint64_t* kin, uint64_t* kout, const int* vin, int* vout, int n,
                        int end_bit, void* scratch);
}  // namespace abd
int main() {
    std::mt19937_64 rng(1);
    int fails = 0;
    for (int n : {1, 2, 7, 100, 2047, 2048, 2049, 5000, 40000}) {
        for (int bits : {13, 33, 64}) {
            std::vector<uint64_t> k(n);
            std::vector<int> v(n);
            for (int i = 0; i < n; ++i) {
                k[i] = bits == 64 ? rng() : (rng() & ((1ull << bits) - 1));
                if (i % 3 == 0 && i) k[i] = k[i - 1];
                v[i] = i;
            }
            uint64_t *dk, *dko;
            int *dv, *dvo;
            void* tmp;
            hipMalloc(&dk, n * 8); hipMalloc(&dko, n * 8); hipMalloc(&dv, n * 4); hipMalloc(&dvo, n * 4);
            hipMalloc(&tmp, abd::radix_scratch_bytes(n));
            hipMemcpy(dk, k.data(), n * 8, hipMemcpyHostToDevice);
            hipMemcpy(dv, v.data(), n * 4, hipMemcpyHostToDevice);
            abd::radix_sort64(nullptr, dk, dko, dv, dvo, n, bits, tmp);
            std::vector<uint64_t> ko(n);
            std::vector<int> vo(n);
            hipMemcpy(ko.data(), dko, n * 8, hipMemcpyDeviceToHost);
            hipMemcpy(vo.data(), dvo, n * 4, hipMemcpyDeviceToHost);
            std::vector<int> idx(n);
            std::iota(idx.begin(), idx.end(), 0);
            std::stable_sort(idx.begin(), idx.end(), [&](int a, int b) { return k[a] < k[b]; });
            bool ok = true;
            for (int i = 0; i < n; ++i) ok &= (ko[i] == k[idx[i]] && vo[i] == idx[i]);
            // keys-only
            abd::radix_sort64(nullptr, dk, dko, nullptr, nullptr, n, bits, tmp);
            hipMemcpy(ko.data(), dko, n * 8, hipMemcpyDeviceToHost);
            for (int i = 0; i < n; ++i) ok &= (ko[i] == k[idx[i]]);
            // scan of small ints
            std::vector<int> a(n), out(n);
            for (int i = 0; i < n; ++i) a[i] = (int)(rng() % 3);
            hipMemcpy(dv, a.data(), n * 4, hipMemcpyHostToDevice);
            abd::excl_scan_int(nullptr, dv, dvo, n, (int*)tmp);
            hipMemcpy(out.data(), dvo, n * 4, hipMemcpyDeviceToHost);
            int acc = 0;
            for (int i = 0; i < n; ++i) { ok &= out[i] == acc; acc += a[i]; }
            std::printf("n=%6d bits=%2d %s\n", n, bits, ok ? "ok" : "FAIL");
            fails += !ok;
            hipFree(dk); hipFree(dko); hipFree(dv); hipFree(dvo); hipFree(tmp);
        }
    }
    // multi-level scan: more than SORT_TILE^2 elements
    {
        const int n = 2048 * 2048 + 4099;
        std::vector<int> a(n), out(n);
        for (int i = 0; i < n; ++i) a[i] = (int)(rng() % 3);
        int *din, *dout, *tmp;
        hipMalloc(&din, (size_t)n * 4); hipMalloc(&dout, (size_t)n * 4);
        hipMalloc(&tmp, sizeof(int) * (size_t)abd::scan_scratch_ints(n));
        hipMemcpy(din, a.data(), (size_t)n * 4, hipMemcpyHostToDevice);
        bool ok = abd::excl_scan_int(nullptr, din, dout, n, tmp) == hipSuccess;
        hipMemcpy(out.data(), dout, (size_t)n * 4, hipMemcpyDeviceToHost);
        long acc = 0;
        for (int i = 0; i < n && ok; ++i) { ok &= out[i] == acc; acc += a[i]; }
        std::printf("scan n=%d (three levels) %s\n", n, ok ? "ok" : "FAIL");
        fails += !ok;
        hipFree(din); hipFree(dout); hipFree(tmp);
    }
    return fails;
}
