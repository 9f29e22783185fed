/*
 * C++ facade of the GPU radix sort with the reference's names:
 *   util::DoubleBuffer<K, V>       b40c/util/multiple_buffering.cuh:65-150
 *   radix_sort::Enactor::Sort      b40c/radix_sort/enactor.cuh:153-266
 * Keys: uint64 (keys only) or uint32 with uint32 values.  Both buffers of a
 * DoubleBuffer must be allocated by the caller (the reference allocates the
 * second lazily); after Sort, selector names the buffer holding the result.
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

class Enactor {
public:
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

private:
    template <typename K, typename V>
    b40c_status sort(util::DoubleBuffer<K, V>& s, int64_t n, int begin_bit, int end_bit) {
        int32_t sel = 0;
        b40c_status st;
        if constexpr (std::is_same<V, util::NullType>::value) {
            static_assert(sizeof(K) == 8, "keys-only sorts take 64-bit keys");
            st = b40c_radix_sort_u64(reinterpret_cast<uint64_t*>(s.d_keys[s.selector]),
                                     reinterpret_cast<uint64_t*>(s.d_keys[s.selector ^ 1]), n, begin_bit, end_bit,
                                     &sel);
        } else {
            static_assert(sizeof(K) == 4, "key-value sorts take 32-bit keys and values");
            st = b40c_radix_sort_pairs_u32(reinterpret_cast<uint32_t*>(s.d_keys[s.selector]),
                                           reinterpret_cast<uint32_t*>(s.d_keys[s.selector ^ 1]),
                                           reinterpret_cast<uint32_t*>(s.d_values[s.selector]),
                                           reinterpret_cast<uint32_t*>(s.d_values[s.selector ^ 1]), n, begin_bit,
                                           end_bit, &sel);
        }
        if (st == B40C_OK) s.selector ^= sel;
        return st;
    }
};

}  // namespace radix_sort
}  // namespace b40c
