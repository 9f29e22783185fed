/*
 * C++ facade of the GPU radix sort with the reference's names:
 *   util::DoubleBuffer<K, V>       b40c/util/multiple_buffering.cuh:65-150
 *   radix_sort::Enactor::Sort      b40c/radix_sort/enactor.cuh:153-266
 * Keys: any KeyTraits type (uint32, int32, float, uint64, int64, double),
 * alone or with 32-bit values.  Both buffers of a DoubleBuffer must be
 * allocated by the caller (the reference allocates the second lazily); after
 * Sort, selector names the buffer holding the result.
 * Digits are 8 bits wide (the reference uses 4); the sort is a stable LSD
 * sort either way, so results are identical.
 */
#pragma once

#include <cstdint>
#include <type_traits>

#include "../b40c_mi355x.h"

namespace b40c {
namespace util {

struct NullType {};

template <typename KeyType, typename ValueType = NullType>
struct DoubleBuffer {
    KeyType* d_keys[2] = {nullptr, nullptr};
    ValueType* d_values[2] = {nullptr, nullptr};
    int selector = 0;

    DoubleBuffer() = default;
    DoubleBuffer(KeyType* keys0, KeyType* keys1) {
        d_keys[0] = keys0;
        d_keys[1] = keys1;
    }
    DoubleBuffer(KeyType* keys0, KeyType* keys1, ValueType* values0, ValueType* values1) {
        d_keys[0] = keys0;
        d_keys[1] = keys1;
        d_values[0] = values0;
        d_values[1] = values1;
    }
};

}  // namespace util

namespace radix_sort {

// Key type codes of the C ABI for the reference's KeyTraits key types
// (radix_sort/sort_utils.cuh:163-183).
template <typename K>
struct KeyTypeCode;
template <> struct KeyTypeCode<uint32_t> { static constexpr int32_t value = B40C_KEY_U32; };
template <> struct KeyTypeCode<int32_t> { static constexpr int32_t value = B40C_KEY_I32; };
template <> struct KeyTypeCode<float> { static constexpr int32_t value = B40C_KEY_F32; };
template <> struct KeyTypeCode<uint64_t> { static constexpr int32_t value = B40C_KEY_U64; };
template <> struct KeyTypeCode<int64_t> { static constexpr int32_t value = B40C_KEY_I64; };
template <> struct KeyTypeCode<double> { static constexpr int32_t value = B40C_KEY_F64; };

// Owns its scratch (the spine) across calls, like the reference Enactor
// (radix_sort/enactor.cuh:96-140).  Not copyable.
class Enactor {
public:
    Enactor() { (void)b40c_sort_enactor_create(&e_); }
    ~Enactor() { b40c_sort_enactor_destroy(e_); }
    Enactor(const Enactor&) = delete;
    Enactor& operator=(const Enactor&) = delete;

    // Sorts bits [START_BIT, START_BIT + NUM_BITS) of the keys, stable.
    template <int START_BIT, int NUM_BITS, typename K, typename V, typename SizeT>
    b40c_status Sort(util::DoubleBuffer<K, V>& storage, SizeT num_elements, int max_grid_size = 0) {
        (void)max_grid_size;
        return sort(storage, (int64_t)num_elements, START_BIT, START_BIT + NUM_BITS);
    }

    // All key bits.
    template <typename K, typename V, typename SizeT>
    b40c_status Sort(util::DoubleBuffer<K, V>& storage, SizeT num_elements, int max_grid_size = 0) {
        (void)max_grid_size;
        return sort(storage, (int64_t)num_elements, 0, (int)(8 * sizeof(K)));
    }

    // Launch policy chosen at run time (the reference picks a KernelPolicy
    // per architecture at compile time, radix_sort/autotuned_policy.cuh:
    // 190-230): keys per thread of a tile (8 or 16) and workgroups per CU
    // (1-8); 0 keeps the tuned default.
    b40c_status SetPolicy(int items_per_thread, int wgs_per_cu) {
        if (!e_) return B40C_ERR_HIP;
        return b40c_sort_enactor_set_policy(e_, items_per_thread, wgs_per_cu);
    }

private:
    template <typename K, typename V>
    b40c_status sort(util::DoubleBuffer<K, V>& s, int64_t n, int begin_bit, int end_bit) {
        if (!e_) return B40C_ERR_HIP;
        int32_t sel = 0;
        uint32_t* v0 = nullptr;
        uint32_t* v1 = nullptr;
        if constexpr (!std::is_same<V, util::NullType>::value) {
            static_assert(sizeof(V) == 4, "values are 32-bit");
            v0 = reinterpret_cast<uint32_t*>(s.d_values[s.selector]);
            v1 = reinterpret_cast<uint32_t*>(s.d_values[s.selector ^ 1]);
        }
        const b40c_status st = b40c_sort_enactor_sort(e_, KeyTypeCode<K>::value, s.d_keys[s.selector],
                                                      s.d_keys[s.selector ^ 1], v0, v1, n, begin_bit, end_bit, &sel);
        if (st == B40C_OK) s.selector ^= sel;
        return st;
    }

    b40c_sort_enactor* e_ = nullptr;
};

}  // namespace radix_sort
}  // namespace b40c
