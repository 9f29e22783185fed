// C++ parity driver for the GPU radix sort in the shape of the reference's
// test/radix_sort/test_sort.cu (:185-261): random keys (and values), sort on
// the device through the C++ facade (include/b40c/radix_sort.hpp), compare
// with std::sort of the keys / std::stable_sort of (key, value) pairs by key.
// Also the consecutive removal (test/consecutive_removal/
// test_consecutive_removal.cu:135-159: first of each run) and the exclusive
// scan through include/b40c/primitives.hpp.
//
// usage: test_sort [n] [seed]    exit status 0 iff every case matches.
#include <algorithm>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <numeric>
#include <vector>

#include "b40c/primitives.hpp"
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

// Sorted keys with runs, trimmed with values; reference = first of each run.
bool trim_case(size_t n, uint64_t seed) {
    std::vector<uint64_t> k(n);
    std::vector<uint32_t> v(n);
    for (size_t i = 0; i < n; ++i) {
        k[i] = splitmix(seed) % (n / 3 + 1);
        v[i] = (uint32_t)i;
    }
    std::sort(k.begin(), k.end());
    uint64_t* dk0 = device_copy(k);
    uint64_t* dk1 = device_copy(std::vector<uint64_t>(n));
    uint32_t* dv0 = device_copy(v);
    uint32_t* dv1 = device_copy(std::vector<uint32_t>(n));
    b40c::util::DoubleBuffer<uint64_t, uint32_t> buf(dk0, dk1, dv0, dv1);
    b40c::consecutive_removal::Enactor trimmer;
    int64_t m = -1;
    const b40c_status s = trimmer.Trim(buf, (int64_t)n, &m);
    std::vector<uint64_t> gk = host_copy(buf.d_keys[buf.selector], m > 0 ? (size_t)m : 0);
    std::vector<uint32_t> gv = host_copy(buf.d_values[buf.selector], m > 0 ? (size_t)m : 0);
    for (void* p : {(void*)dk0, (void*)dk1, (void*)dv0, (void*)dv1}) b40c_device_free(p);
    std::vector<uint64_t> ek;
    std::vector<uint32_t> ev;
    for (size_t i = 0; i < n; ++i)
        if (i == 0 || k[i] != k[i - 1]) {
            ek.push_back(k[i]);
            ev.push_back(v[i]);
        }
    const bool ok = s == B40C_OK && m == (int64_t)ek.size() && gk == ek && gv == ev;
    printf("trim n=%zu -> %lld: %s\n", n, (long long)m, ok ? "CORRECT" : "INCORRECT");
    return ok;
}

bool scan_case(size_t n, uint64_t seed) {
    std::vector<uint32_t> x(n);
    for (auto& e : x) e = (uint32_t)splitmix(seed);
    uint32_t* d = device_copy(x);
    b40c::scan::Enactor scanner;
    uint64_t total = 0;
    const b40c_status s = scanner.Scan<true>(d, d, (int64_t)n, &total);
    std::vector<uint32_t> got = host_copy(d, n);
    b40c_device_free(d);
    std::vector<uint32_t> ref(n);
    uint32_t acc = 0;
    uint64_t exact = 0;
    for (size_t i = 0; i < n; ++i) {
        ref[i] = acc;
        acc += x[i];
        exact += x[i];
    }
    const bool ok = s == B40C_OK && got == ref && total == exact;
    printf("exclusive scan n=%zu: %s\n", n, ok ? "CORRECT" : "INCORRECT");
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
    ok &= trim_case(n, seed + 5);
    ok &= trim_case(1, seed + 6);
    ok &= scan_case(n, seed + 7);
    ok &= scan_case(0, seed + 8);
    printf("summary: %s\n", ok ? "all CORRECT" : "FAILURES");
    return ok ? 0 : 1;
}
