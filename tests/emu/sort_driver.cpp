// Emulator driver for the radix sort and scan of colb_amd/csrc/abd_sort.hip (TEST ONLY).
#include <algorithm>
#include <array>
#include <cstdio>
#include <numeric>
#include <random>
#include <vector>
#include "emu_hip.hpp"
#include "../../colb_amd/csrc/abd_cnt.hpp"
namespace abd {
hipError_t excl_scan_int(hipStream_t s, const int* in, int* out, Cnt n, int* bsum);
int scan_scratch_ints(long n);
size_t radix_scratch_bytes(long n);
hipError_t small_sort2(hipStream_t s, const uint64_t* ka, uint64_t* oa, Cnt na, const uint64_t* kb, uint64_t* ob,
                       Cnt nb);
hipError_t radix_sort64(hipStream_t s, const uint64_t* kin, uint64_t* kout, const int* vin, int* vout, Cnt n,
                        uint64_t sig, void* scratch);
}  // namespace abd
using abd::host_cnt;
int main() {
    std::mt19937_64 rng(1);
    int fails = 0;
    // masks: low bits, and packed fields with gaps (windows that overlap the next field)
    const uint64_t masks[] = {(1ull << 13) - 1, (1ull << 33) - 1, ~0ull, (0x3FFull << 32) | 0x1FFFull,
                              (0x3FFull << 44) | (0x3FFull << 24) | 0xFFFull, (0x7ull << 61) | 0x5ull};
    for (int n : {1, 2, 7, 100, 2047, 2048, 2049, 4096, 4097, 8192, 8193, 40000}) {
        for (uint64_t mask : masks) {
            std::vector<uint64_t> k(n);
            std::vector<int> v(n);
            for (int i = 0; i < n; ++i) {
                k[i] = rng() & mask;
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
            abd::radix_sort64(nullptr, dk, dko, dv, dvo, host_cnt(n), mask, tmp);
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
            abd::radix_sort64(nullptr, dk, dko, nullptr, nullptr, host_cnt(n), mask, tmp);
            hipMemcpy(ko.data(), dko, n * 8, hipMemcpyDeviceToHost);
            for (int i = 0; i < n; ++i) ok &= (ko[i] == k[idx[i]]);
            // scan of small ints
            std::vector<int> a(n), out(n);
            for (int i = 0; i < n; ++i) a[i] = (int)(rng() % 3);
            hipMemcpy(dv, a.data(), n * 4, hipMemcpyHostToDevice);
            abd::excl_scan_int(nullptr, dv, dvo, host_cnt(n), (int*)tmp);
            hipMemcpy(out.data(), dvo, n * 4, hipMemcpyDeviceToHost);
            int acc = 0;
            for (int i = 0; i < n; ++i) { ok &= out[i] == acc; acc += a[i]; }
            std::printf("n=%6d mask=%016llx %s\n", n, (unsigned long long)mask, ok ? "ok" : "FAIL");
            fails += !ok;
            hipFree(dk); hipFree(dko); hipFree(dv); hipFree(dvo); hipFree(tmp);
        }
    }
    // lengths read on the device: the grid sized by a larger bound (both sort forms launched when the
    // bound exceeds the counting-rank limit), the actual length in device memory
    for (int cap : {100, 4096, 8192, 9000, 20000}) {
        for (int n : {0, 1, 63, 100, 2049, 4097, 8192, 8193, 20000}) {
            if (n > cap) continue;
            const uint64_t mask = (0x3FFull << 44) | (0x3FFull << 24) | 0xFFFull;
            std::vector<uint64_t> k(cap, ~0ull);
            std::vector<int> v(cap, -7);
            for (int i = 0; i < n; ++i) {
                k[i] = rng() & mask;
                if (i % 3 == 0 && i) k[i] = k[i - 1];
                v[i] = i;
            }
            uint64_t *dk, *dko;
            int *dv, *dvo, *dn;
            void* tmp;
            hipMalloc(&dk, cap * 8); hipMalloc(&dko, cap * 8); hipMalloc(&dv, cap * 4); hipMalloc(&dvo, cap * 4);
            hipMalloc(&dn, 4); hipMalloc(&tmp, abd::radix_scratch_bytes(cap));
            hipMemcpy(dk, k.data(), cap * 8, hipMemcpyHostToDevice);
            hipMemcpy(dv, v.data(), cap * 4, hipMemcpyHostToDevice);
            hipMemcpy(dn, &n, 4, hipMemcpyHostToDevice);
            abd::radix_sort64(nullptr, dk, dko, dv, dvo, abd::Cnt{cap, dn}, mask, tmp);
            std::vector<uint64_t> ko(cap);
            std::vector<int> vo(cap);
            hipMemcpy(ko.data(), dko, cap * 8, hipMemcpyDeviceToHost);
            hipMemcpy(vo.data(), dvo, cap * 4, hipMemcpyDeviceToHost);
            std::vector<int> idx(n);
            std::iota(idx.begin(), idx.end(), 0);
            std::stable_sort(idx.begin(), idx.end(), [&](int a, int b) { return k[a] < k[b]; });
            bool ok = true;
            for (int i = 0; i < n; ++i) ok &= (ko[i] == k[idx[i]] && vo[i] == idx[i]);
            std::vector<int> a(cap, 5), out(cap, -3);
            for (int i = 0; i < n; ++i) a[i] = (int)(rng() % 3);
            hipMemcpy(dv, a.data(), cap * 4, hipMemcpyHostToDevice);
            hipMemcpy(dvo, out.data(), cap * 4, hipMemcpyHostToDevice);
            abd::excl_scan_int(nullptr, dv, dvo, abd::Cnt{cap, dn}, (int*)tmp);
            hipMemcpy(out.data(), dvo, cap * 4, hipMemcpyDeviceToHost);
            int acc = 0;
            for (int i = 0; i < n; ++i) { ok &= out[i] == acc; acc += a[i]; }
            std::printf("device length n=%5d bound=%5d %s\n", n, cap, ok ? "ok" : "FAIL");
            fails += !ok;
            hipFree(dk); hipFree(dko); hipFree(dv); hipFree(dvo); hipFree(dn); hipFree(tmp);
        }
    }
    // two lists in one launch (device lengths below the host bounds, one list empty, one at the limit)
    for (auto [ca, na, cb, nbk] : std::vector<std::array<int, 4>>{{4096, 3134, 2048, 1567}, {100, 0, 8192, 8192},
                                                                  {8192, 8000, 64, 1}}) {
        std::vector<uint64_t> ka(ca, ~0ull), kb(cb, ~0ull);
        for (int i = 0; i < na; ++i) ka[i] = rng() & ((1ull << 40) - 1);
        for (int i = 0; i < nbk; ++i) kb[i] = rng() & ((1ull << 52) - 1);
        uint64_t *da, *dao, *db, *dbo;
        int* dn;
        hipMalloc(&da, ca * 8); hipMalloc(&dao, ca * 8); hipMalloc(&db, cb * 8); hipMalloc(&dbo, cb * 8);
        hipMalloc(&dn, 8);
        int nn[2] = {na, nbk};
        hipMemcpy(da, ka.data(), ca * 8, hipMemcpyHostToDevice);
        hipMemcpy(db, kb.data(), cb * 8, hipMemcpyHostToDevice);
        hipMemcpy(dn, nn, 8, hipMemcpyHostToDevice);
        abd::small_sort2(nullptr, da, dao, abd::Cnt{ca, dn}, db, dbo, abd::Cnt{cb, dn + 1});
        std::vector<uint64_t> oa(ca), ob(cb);
        hipMemcpy(oa.data(), dao, ca * 8, hipMemcpyDeviceToHost);
        hipMemcpy(ob.data(), dbo, cb * 8, hipMemcpyDeviceToHost);
        std::vector<uint64_t> ea(ka.begin(), ka.begin() + na), eb(kb.begin(), kb.begin() + nbk);
        std::sort(ea.begin(), ea.end());
        std::sort(eb.begin(), eb.end());
        bool ok = true;
        for (int i = 0; i < na; ++i) ok &= oa[i] == ea[i];
        for (int i = 0; i < nbk; ++i) ok &= ob[i] == eb[i];
        std::printf("two lists %d/%d, %d/%d %s\n", na, ca, nbk, cb, ok ? "ok" : "FAIL");
        fails += !ok;
        hipFree(da); hipFree(dao); hipFree(db); hipFree(dbo); hipFree(dn);
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
        bool ok = abd::excl_scan_int(nullptr, din, dout, host_cnt(n), tmp) == hipSuccess;
        hipMemcpy(out.data(), dout, (size_t)n * 4, hipMemcpyDeviceToHost);
        long acc = 0;
        for (int i = 0; i < n && ok; ++i) { ok &= out[i] == acc; acc += a[i]; }
        std::printf("scan n=%d (three levels) %s\n", n, ok ? "ok" : "FAIL");
        fails += !ok;
        hipFree(din); hipFree(dout); hipFree(tmp);
    }
    return fails;
}
