/*
 * C++ facade for the device primitives next to the CSR build, with the
 * reference's names:
 *   consecutive_removal::Enactor::Trim   b40c/consecutive_removal/enactor.cuh:119-126
 *   scan::Enactor::Scan<EXCLUSIVE>       b40c/scan/enactor.cuh:122-136
 * Trim reads d_keys[selector] (and d_values[selector]) and writes the run
 * heads to the other buffer, then flips the selector, as the reference's
 * output lands in its second buffer.  Keys are uint64, optional values
 * uint32; the equality operator is key ==.  Scan sums uint32 values.
 */
#pragma once

#include <cstdint>

#include "../b40c_mi355x.h"
#include "radix_sort.hpp"

namespace b40c {
namespace consecutive_removal {

class Enactor {
public:
    b40c_status Trim(util::DoubleBuffer<uint64_t>& storage, int64_t num_elements, int64_t* h_num_compacted) {
        const int s = storage.selector;
        b40c_status st = b40c_consecutive_removal_u64(storage.d_keys[s], nullptr, storage.d_keys[s ^ 1], nullptr,
                                                      num_elements, h_num_compacted);
        if (st == B40C_OK) storage.selector = s ^ 1;
        return st;
    }
    b40c_status Trim(util::DoubleBuffer<uint64_t, uint32_t>& storage, int64_t num_elements,
                     int64_t* h_num_compacted) {
        const int s = storage.selector;
        b40c_status st = b40c_consecutive_removal_u64(storage.d_keys[s], storage.d_values[s], storage.d_keys[s ^ 1],
                                                      storage.d_values[s ^ 1], num_elements, h_num_compacted);
        if (st == B40C_OK) storage.selector = s ^ 1;
        return st;
    }
};

}  // namespace consecutive_removal

namespace scan {

class Enactor {
public:
    template <bool EXCLUSIVE>
    b40c_status Scan(uint32_t* d_dest, const uint32_t* d_src, int64_t num_elements, uint64_t* total = nullptr) {
        return b40c_scan_u32(d_src, d_dest, num_elements, EXCLUSIVE ? 1 : 0, total);
    }
};

}  // namespace scan
}  // namespace b40c
