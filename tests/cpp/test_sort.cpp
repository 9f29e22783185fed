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
#include <cstring>
#include <numeric>
#include <type_traits>
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

// Keys of the reference's other KeyTraits types (sort_utils.cuh:163-183),
// optionally with values, against std::stable_sort by the type's own `<`
// (values are the input positions, so stability is checked too).  Floating
// keys are finite and nonzero here, where `<` is a total order.
template <typename K>
K random_key(uint64_t& seed) {
    const uint64_t r = splitmix(seed);
    if constexpr (std::is_floating_point<K>::value)
        return (K)((int64_t)(r >> 12) - (int64_t)(1ll << 51)) * (K)1e-9 + ((r & 1) ? (K)0.5 : (K)-0.5);
    else
        return (K)(r % 5000) - (K)(std::is_signed<K>::value ? 2500 : 0);  // duplicates, both signs
}

template <typename K>
bool typed_case(const char* name, size_t n, uint64_t seed, bool with_values, int items = 0, int wgs = 0) {
    std::vector<K> k(n);
    std::vector<uint32_t> v(n);
    for (size_t i = 0; i < n; ++i) {
        k[i] = random_key<K>(seed);
        v[i] = (uint32_t)i;
    }
    K* dk0 = device_copy(k);
    K* dk1 = device_copy(std::vector<K>(n));
    uint32_t* dv0 = device_copy(v);
    uint32_t* dv1 = device_copy(std::vector<uint32_t>(n));
    b40c::radix_sort::Enactor sorter;
    b40c_status s = sorter.SetPolicy(items, wgs);
    std::vector<K> gk;
    std::vector<uint32_t> gv;
    if (s != B40C_OK) {
    } else if (with_values) {
        b40c::util::DoubleBuffer<K, uint32_t> buf(dk0, dk1, dv0, dv1);
        s = sorter.Sort(buf, n);
        gk = host_copy(buf.d_keys[buf.selector], n);
        gv = host_copy(buf.d_values[buf.selector], n);
    } else {
        b40c::util::DoubleBuffer<K> buf(dk0, dk1);
        s = sorter.Sort(buf, n);
        gk = host_copy(buf.d_keys[buf.selector], n);
    }
    for (void* p : {(void*)dk0, (void*)dk1, (void*)dv0, (void*)dv1}) b40c_device_free(p);
    std::vector<size_t> order(n);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](size_t a, size_t b) { return k[a] < k[b]; });
    bool ok = s == B40C_OK;
    for (size_t i = 0; ok && i < n; ++i) ok = gk[i] == k[order[i]] && (!with_values || gv[i] == v[order[i]]);
    printf("%s%s n=%zu policy=%d/%d: %s\n", name, with_values ? " pairs" : " keys", n, items, wgs,
           ok ? "CORRECT" : "INCORRECT");
    return ok;
}

// IEEE special values, in the order the reference's float conversion gives
// them (sort_utils.cuh:142-160: negative keys have every bit flipped,
// positive keys the sign bit), worked out by hand from the bit patterns:
// -NaN, -inf, -1, -0, +0, 1, inf, NaN.
bool float_specials_case() {
    const uint32_t bits[8] = {0x3F800000u, 0x80000000u, 0x7FC00000u, 0x00000000u,
                              0xFF800000u, 0x7F800000u, 0xBF800000u, 0xFFC00000u};
    const uint32_t want[8] = {0xFFC00000u, 0xFF800000u, 0xBF800000u, 0x80000000u,
                              0x00000000u, 0x3F800000u, 0x7F800000u, 0x7FC00000u};
    std::vector<float> k(8);
    std::memcpy(k.data(), bits, sizeof(bits));
    float* d0 = device_copy(k);
    float* d1 = device_copy(std::vector<float>(8));
    b40c::util::DoubleBuffer<float> buf(d0, d1);
    b40c::radix_sort::Enactor sorter;
    const b40c_status s = sorter.Sort(buf, 8);
    std::vector<float> got = host_copy(buf.d_keys[buf.selector], 8);
    b40c_device_free(d0);
    b40c_device_free(d1);
    bool ok = s == B40C_OK;
    for (int i = 0; ok && i < 8; ++i) ok = std::memcmp(&got[i], &want[i], 4) == 0;
    printf("float specials: %s\n", ok ? "CORRECT" : "INCORRECT");
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
    ok &= typed_case<int32_t>("int32", n, seed + 9, true);
    ok &= typed_case<int64_t>("int64", n, seed + 10, false);
    ok &= typed_case<float>("float", n, seed + 11, false);
    ok &= typed_case<double>("double", n, seed + 12, true);
    ok &= typed_case<uint32_t>("uint32", n, seed + 13, false);
    ok &= typed_case<uint64_t>("uint64", n, seed + 14, true, 16, 2);  // a non-default launch policy
    ok &= typed_case<int32_t>("int32", n, seed + 15, false, 16, 8);
    ok &= float_specials_case();
    ok &= scan_case(n, seed + 7);
    ok &= scan_case(0, seed + 8);
    printf("summary: %s\n", ok ? "all CORRECT" : "FAILURES");
    return ok ? 0 : 1;
}
