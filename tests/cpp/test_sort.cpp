// C++ parity driver for the GPU radix sort in the shape of the reference's
// test/radix_sort/test_sort.cu (:185-261): random keys (and values), sort on
// the device through the C++ facade (include/b40c/radix_sort.hpp), compare
// with std::sort of the keys / std::stable_sort of (key, value) pairs by key.
//
// usage: test_sort [n] [seed]    exit status 0 iff every case matches.
#include <algorithm>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <numeric>
#include <vector>

#include "b40c/radix_sort.hpp"

namespace {

uint64_t splitmix(uint64_t& s) {
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

template <typename T>
T* device_copy(const std::vector<T>& h) {
    void* d = nullptr;
    if (b40c_device_alloc(&d, (int64_t)std::max<size_t>(h.size(), 1) * sizeof(T)) != B40C_OK) return nullptr;
    if (!h.empty() && b40c_memcpy_h2d(d, h.data(), (int64_t)(h.size() * sizeof(T))) != B40C_OK) return nullptr;
    return static_cast<T*>(d);
}

template <typename T>
std::vector<T> host_copy(const T* d, size_t n) {
    std::vector<T> h(n);
    if (n) b40c_memcpy_d2h(h.data(), d, (int64_t)(n * sizeof(T)));
    return h;
}

// Keys only, 64-bit, all bits or bits [0, 40).
bool keys_case(size_t n, uint64_t seed, bool low40) {
    std::vector<uint64_t> h(n);
    for (auto& k : h) k = splitmix(seed) >> (low40 ? 24 : 0);
    uint64_t* d0 = device_copy(h);
    uint64_t* d1 = device_copy(std::vector<uint64_t>(n));
    b40c::util::DoubleBuffer<uint64_t> buf(d0, d1);
    b40c::radix_sort::Enactor sorter;
    const b40c_status s = low40 ? sorter.Sort<0, 40>(buf, n) : sorter.Sort(buf, n);
    std::vector<uint64_t> got = host_copy(buf.d_keys[buf.selector], n);
    b40c_device_free(d0);
    b40c_device_free(d1);
    std::sort(h.begin(), h.end());
    const bool ok = s == B40C_OK && got == h;
    printf("keys n=%zu bits=%s: %s\n", n, low40 ? "[0,40)" : "all", ok ? "CORRECT" : "INCORRECT");
    return ok;
}

// 32-bit keys with 32-bit values: stable, so equal keys keep their order.
bool pairs_case(size_t n, uint64_t seed) {
    std::vector<uint32_t> k(n), v(n);
    for (size_t i = 0; i < n; ++i) {
        k[i] = (uint32_t)(splitmix(seed) % 1000);  // many duplicates
        v[i] = (uint32_t)i;
    }
    uint32_t* dk0 = device_copy(k);
    uint32_t* dk1 = device_copy(std::vector<uint32_t>(n));
    uint32_t* dv0 = device_copy(v);
    uint32_t* dv1 = device_copy(std::vector<uint32_t>(n));
    b40c::util::DoubleBuffer<uint32_t, uint32_t> buf(dk0, dk1, dv0, dv1);
    b40c::radix_sort::Enactor sorter;
    const b40c_status s = sorter.Sort(buf, n);
    std::vector<uint32_t> gk = host_copy(buf.d_keys[buf.selector], n), gv = host_copy(buf.d_values[buf.selector], n);
    for (void* p : {(void*)dk0, (void*)dk1, (void*)dv0, (void*)dv1}) b40c_device_free(p);
    std::vector<size_t> order(n);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](size_t a, size_t b) { return k[a] < k[b]; });
    bool ok = s == B40C_OK;
    for (size_t i = 0; ok && i < n; ++i) ok = gk[i] == k[order[i]] && gv[i] == v[order[i]];
    printf("pairs n=%zu: %s\n", n, ok ? "CORRECT" : "INCORRECT");
    return ok;
}

}  // namespace

int main(int argc, char** argv) {
    const size_t n = argc > 1 ? (size_t)strtoull(argv[1], nullptr, 10) : 1000000;
    const uint64_t seed = argc > 2 ? strtoull(argv[2], nullptr, 10) : 7;
    bool ok = true;
    ok &= keys_case(n, seed, false);
    ok &= keys_case(n, seed + 1, true);
    ok &= pairs_case(n, seed + 2);
    ok &= keys_case(1, seed + 3, false);
    ok &= pairs_case(0, seed + 4);
    printf("summary: %s\n", ok ? "all CORRECT" : "FAILURES");
    return ok ? 0 : 1;
}
