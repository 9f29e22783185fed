// Internal host-side declarations shared by the library's translation units.
#pragma once

#include <atomic>

#include "common.hpp"

namespace b40c {

// Owning device allocation (hipMalloc / hipFree).
struct DeviceBuffer {
    void* ptr = nullptr;
    size_t bytes = 0;
    DeviceBuffer() = default;
    DeviceBuffer(const DeviceBuffer&) = delete;
    DeviceBuffer& operator=(const DeviceBuffer&) = delete;
    ~DeviceBuffer() { release(); }
    b40c_status alloc(size_t n) {
        release();
        if (n == 0) n = 1;
        hipError_t e = hipMalloc(&ptr, n);
        if (e != hipSuccess) {
            ptr = nullptr;
            return fail(e == hipErrorOutOfMemory ? B40C_ERR_OOM : B40C_ERR_HIP, "hipMalloc(%zu) failed: %s", n,
                        hipGetErrorString(e));
        }
        bytes = n;
        return B40C_OK;
    }
    // Grow-only reallocation (contents are not preserved).
    b40c_status reserve(size_t n) { return (n <= bytes && ptr) ? B40C_OK : alloc(n); }
    void release() {
        if (ptr) (void)hipFree(ptr);
        ptr = nullptr;
        bytes = 0;
    }
    template <typename T>
    T* as() const { return static_cast<T*>(ptr); }
};

// radix_sort.hip.  A sort's device scratch: the spine (256 counters per
// workgroup), two digit-total buffers, the device-side selector and the
// trivial-pass flag, carved out of one block of `words` uint32 words.
struct SortScratch {
    uint32_t* spine = nullptr;
    uint32_t* digit_total = nullptr;
    int32_t* sel = nullptr;
    uint32_t* trivial = nullptr;
    size_t words = 0;
};
// A sort's launch policy (the reference's per-architecture KernelPolicy /
// AutotunedPolicy, radix_sort/autotuned_policy.cuh:190-230, chosen at run
// time): keys per thread of a downsweep tile (8 or 16; a tile is 256 x that)
// and workgroups per CU (1-8).  Zero fields take the size-based default.
struct SortPolicy {
    int items = 0;
    int wgs_per_cu = 0;
};
// Policy with its zero fields filled in for n keys of key_bytes bytes.
SortPolicy resolve_sort_policy(SortPolicy p, int64_t n, int key_bytes);
// Words of scratch a sort of n keys needs on a device with num_cus CUs.
size_t sort_scratch_words(int64_t n, int num_cus, const SortPolicy& p);
SortScratch sort_scratch_view(uint32_t* words, size_t count);
// Launches the passes (no allocation, no synchronization); the selector is
// left in *sc.sel.  `launched` is false when there was nothing to sort.
b40c_status radix_sort_async(int32_t key_type, void* k0, void* k1, uint32_t* v0, uint32_t* v1, int64_t n,
                             int begin_bit, int end_bit, int num_cus, const SortPolicy& policy,
                             const SortScratch& sc, hipStream_t stream, bool& launched);
b40c_status radix_sort_u64(uint64_t* k0, uint64_t* k1, int64_t n, int begin_bit, int end_bit, int32_t* selector,
                           hipStream_t stream);
b40c_status radix_sort_pairs_u32(uint32_t* k0, uint32_t* k1, uint32_t* v0, uint32_t* v1, int64_t n, int begin_bit,
                                 int end_bit, int32_t* selector, hipStream_t stream);

// primitives.hip: prefix scan of uint32 on `stream` with caller-provided
// scratch (`partials`: scan_u32_max_partials() entries; `total`: the exact sum).
int scan_u32_max_partials();
b40c_status scan_u32_async(const uint32_t* in, uint32_t* out, int64_t n, bool exclusive, uint64_t* partials,
                           uint64_t* total, hipStream_t stream);

}  // namespace b40c

// Device-resident CSR graph (reference counted).
struct b40c_graph {
    int32_t nodes = 0;
    int64_t edges = 0;
    b40c::DeviceBuffer offsets;  // int32[nodes + 1]
    b40c::DeviceBuffer cols;     // int32[edges]
    std::atomic<int> refs{1};
    // 1: every edge (u,v) has its reverse (v,u); 0: not (or not provable);
    // -1: not yet checked.  Bottom-up BFS steps are only valid when 1.
    std::atomic<int> symmetric{-1};
    // Largest out-degree, or -1 while not yet computed (graph_max_degree).
    std::atomic<int> max_degree{-1};
    const int32_t* d_offsets() const { return offsets.as<int32_t>(); }
    const int32_t* d_cols() const { return cols.as<int32_t>(); }
};

namespace b40c {
void graph_retain(b40c_graph* g);
void graph_release(b40c_graph* g);
// Resolves g->symmetric (runs a GPU check the first time it is unknown).
b40c_status graph_symmetric(b40c_graph* g, bool& symmetric);
b40c_status graph_max_degree(b40c_graph* g, int32_t& max_degree);
}  // namespace b40c
