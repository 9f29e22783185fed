// Stable LSD radix sort for gfx950: one upsweep -> spine -> downsweep triple
// per 8-bit digit place, replacing b40c's radix_sort enactor
// (b40c/radix_sort/enactor.cuh:520-622) and its upsweep/spine/downsweep
// kernels (radix_sort/upsweep/kernel.cuh:90, scan/spine/kernel.cuh:92,
// radix_sort/downsweep/kernel.cuh:236).
//
// Design (MI355X-first rather than a translation):
//  * 8-bit digits (256 bins) instead of the reference's 4-bit digits: half the
//    passes, and 160 KB of LDS per CU easily holds the per-wave histograms.
//  * Each workgroup owns a contiguous run of 2048-key tiles, so the spine is
//    256 x G counters (digit-major, like partition/upsweep/cta.cuh:302-305).
//  * Stable local ranking uses wave64 "match" masks built from 8 ballots (one
//    per digit bit) and per-wave running digit counters in LDS; keys are
//    processed in (wave, round, lane) order which equals index order inside a
//    tile, so ranks are stable.  The tile is exchanged through LDS into digit
//    order and written out as contiguous per-digit runs (two-phase scatter,
//    partition/downsweep/tile.cuh:779-1030).
//  * The trivial-pass early exit (downsweep/kernel.cuh DownsweepPass) is
//    decided on the device: the spine kernel flags a pass whose keys all share
//    one digit, the downsweep then does nothing and a device-side selector
//    (DoubleBuffer::selector) is left unflipped.  The host reads the selector
//    once at the end.
#include "common.hpp"
#include "internal.hpp"

namespace b40c {
namespace {

constexpr int RS_THREADS = 256;
constexpr int RS_WAVES = RS_THREADS / kWave;
constexpr int RS_ITEMS = 8;
constexpr int RS_TILE = RS_THREADS * RS_ITEMS;
constexpr int RS_BINS = 256;

template <typename K>
__device__ __forceinline__ uint32_t digit_of(K key, int shift, uint32_t mask) {
    return (uint32_t)(key >> shift) & mask;
}

// Per-workgroup digit histogram of its tile range; counts go to the spine
// (digit-major) and to the global digit totals.
template <typename K>
__global__ __launch_bounds__(RS_THREADS) void k_upsweep(K* const* bufs, const int32_t* sel, int64_t n, int shift,
                                                        uint32_t mask, int64_t tiles_per_wg, uint32_t* spine,
                                                        uint32_t* digit_total, int G) {
    __shared__ uint32_t hist[RS_WAVES][RS_BINS];
    const int tid = threadIdx.x, wid = tid / kWave;
    for (int i = tid; i < RS_WAVES * RS_BINS; i += RS_THREADS) (&hist[0][0])[i] = 0;
    __syncthreads();
    const K* keys = bufs[*sel];
    const int64_t begin = (int64_t)blockIdx.x * tiles_per_wg * RS_TILE;
    const int64_t end = min(n, begin + tiles_per_wg * RS_TILE);
    for (int64_t i = begin + tid; i < end; i += RS_THREADS) {
        uint32_t d = digit_of(keys[i], shift, mask);
        atomicAdd(&hist[wid][d], 1u);
    }
    __syncthreads();
    uint32_t c = 0;
#pragma unroll
    for (int w = 0; w < RS_WAVES; ++w) c += hist[w][tid];
    spine[(int64_t)tid * G + blockIdx.x] = c;
    if (c) atomic_add_u32(&digit_total[tid], c);
}

// Exclusive scan of the spine, one workgroup per digit.  Digit d's base is the
// sum of the totals of the smaller digits.  Flags a trivial pass.
__global__ __launch_bounds__(RS_THREADS) void k_spine(uint32_t* spine, const uint32_t* digit_total, int G, int64_t n,
                                                      uint32_t* trivial) {
    __shared__ int s_wave[RS_WAVES];
    const int d = blockIdx.x, tid = threadIdx.x;
    const uint32_t t = digit_total[tid];
    int total_below;
    int contrib = (tid < d) ? (int)t : 0;
    (void)wg_exclusive_scan<RS_THREADS>(contrib, total_below, s_wave);
    if (tid == 0 && (int64_t)digit_total[d] == n) st_relaxed(trivial, 1u);
    uint32_t carry = (uint32_t)total_below;
    uint32_t* row = spine + (int64_t)d * G;
    for (int base = 0; base < G; base += RS_THREADS) {
        const int i = base + tid;
        const int v = (i < G) ? (int)row[i] : 0;
        int chunk_total;
        const int excl = wg_exclusive_scan<RS_THREADS>(v, chunk_total, s_wave);
        if (i < G) row[i] = carry + (uint32_t)excl;
        carry += (uint32_t)chunk_total;
    }
}

template <typename K, bool VALS>
__global__ __launch_bounds__(RS_THREADS) void k_downsweep(K* const* kbufs, uint32_t* const* vbufs, const int32_t* sel,
                                                          int64_t n, int shift, uint32_t mask, int64_t tiles_per_wg,
                                                          const uint32_t* spine, int G, const uint32_t* trivial) {
    if (*trivial) return;
    __shared__ K stage_k[RS_TILE];
    __shared__ uint32_t stage_v[VALS ? RS_TILE : 1];
    __shared__ uint32_t wave_hist[RS_WAVES][RS_BINS];
    __shared__ uint32_t carry[RS_BINS];
    __shared__ uint32_t tile_excl[RS_BINS];
    __shared__ int s_wave[RS_WAVES];

    const int tid = threadIdx.x, lane = lane_id(), wid = tid / kWave;
    const int s = *sel;
    const K* kin = kbufs[s];
    K* kout = kbufs[s ^ 1];
    const uint32_t* vin = VALS ? vbufs[s] : nullptr;
    uint32_t* vout = VALS ? vbufs[s ^ 1] : nullptr;

    carry[tid] = spine[(int64_t)tid * G + blockIdx.x];
    const int64_t begin = (int64_t)blockIdx.x * tiles_per_wg * RS_TILE;
    const int64_t end = min(n, begin + tiles_per_wg * RS_TILE);
    const uint64_t lt_mask = (lane == 0) ? 0ull : (~0ull >> (64 - lane));

    for (int64_t tile_base = begin; tile_base < end; tile_base += RS_TILE) {
        const int tile_n = (int)min((int64_t)RS_TILE, end - tile_base);
        for (int i = lane; i < RS_BINS; i += kWave) wave_hist[wid][i] = 0;
        __syncthreads();

        K key[RS_ITEMS];
        uint32_t val[RS_ITEMS];
        uint32_t rank[RS_ITEMS];
        uint32_t dig[RS_ITEMS];
#pragma unroll
        for (int j = 0; j < RS_ITEMS; ++j) {
            const int idx = wid * (kWave * RS_ITEMS) + j * kWave + lane;
            const bool valid = idx < tile_n;
            key[j] = valid ? kin[tile_base + idx] : (K)0;
            if (VALS) val[j] = valid ? vin[tile_base + idx] : 0u;
            const uint32_t d = digit_of(key[j], shift, mask);
            dig[j] = d;
            uint64_t peers = __ballot(valid);
#pragma unroll
            for (int b = 0; b < 8; ++b) {
                const bool bit = (d >> b) & 1u;
                const uint64_t bal = __ballot(bit);
                peers &= bit ? bal : ~bal;
            }
            const uint32_t before = wave_hist[wid][d];
            rank[j] = before + (uint32_t)__popcll(peers & lt_mask);
            const bool leader = valid && ((peers & lt_mask) == 0);
            if (leader) wave_hist[wid][d] = before + (uint32_t)__popcll(peers);
            if (!valid) rank[j] = 0xFFFFFFFFu;
        }
        __syncthreads();
        // Per digit: offsets of each wave inside the digit, and tile totals.
        uint32_t cnt = 0;
#pragma unroll
        for (int w = 0; w < RS_WAVES; ++w) {
            const uint32_t c = wave_hist[w][tid];
            wave_hist[w][tid] = cnt;
            cnt += c;
        }
        int tot;
        const int excl = wg_exclusive_scan<RS_THREADS>((int)cnt, tot, s_wave);
        tile_excl[tid] = (uint32_t)excl;
        __syncthreads();
#pragma unroll
        for (int j = 0; j < RS_ITEMS; ++j) {
            if (rank[j] != 0xFFFFFFFFu) {
                const uint32_t pos = tile_excl[dig[j]] + wave_hist[wid][dig[j]] + rank[j];
                stage_k[pos] = key[j];
                if (VALS) stage_v[pos] = val[j];
            }
        }
        __syncthreads();
        for (int k = tid; k < tile_n; k += RS_THREADS) {
            const K kk = stage_k[k];
            const uint32_t d = digit_of(kk, shift, mask);
            const uint32_t out = carry[d] + (uint32_t)k - tile_excl[d];
            kout[out] = kk;
            if (VALS) vout[out] = stage_v[k];
        }
        __syncthreads();
        carry[tid] += cnt;
        __syncthreads();
    }
}

__global__ void k_flip_selector(int32_t* sel, uint32_t* trivial) {
    if (*trivial == 0) *sel ^= 1;
    *trivial = 0;
}

struct SortScratch {
    DeviceBuffer spine, digit_total, flags, bufptrs;
};

template <typename K, bool VALS>
b40c_status radix_sort_impl(K* k0, K* k1, uint32_t* v0, uint32_t* v1, int64_t n, int begin_bit, int end_bit,
                            int32_t* selector, hipStream_t stream) {
    if (n < 0 || begin_bit < 0 || end_bit > (int)(8 * sizeof(K)) || begin_bit > end_bit)
        return fail(B40C_ERR_INVALID_ARG, "radix sort: bad arguments (n=%lld bits=[%d,%d))", (long long)n,
                    begin_bit, end_bit);
    if (n > 0xFFFFFFFFll) return fail(B40C_ERR_UNSUPPORTED, "radix sort: n=%lld exceeds 2^32-1", (long long)n);
    *selector = 0;
    if (n <= 1 || begin_bit == end_bit) return B40C_OK;

    DeviceInfo info;
    B40C_TRY(device_info(info));
    const int64_t num_tiles = (n + RS_TILE - 1) / RS_TILE;
    const int64_t max_grid = (int64_t)info.num_cus * 4;
    const int64_t tiles_per_wg = (num_tiles + max_grid - 1) / max_grid;
    const int G = (int)((num_tiles + tiles_per_wg - 1) / tiles_per_wg);

    SortScratch sc;
    B40C_TRY(sc.spine.alloc((size_t)RS_BINS * G * sizeof(uint32_t)));
    B40C_TRY(sc.digit_total.alloc(RS_BINS * sizeof(uint32_t)));
    B40C_TRY(sc.flags.alloc(2 * sizeof(uint32_t)));
    B40C_TRY(sc.bufptrs.alloc(4 * sizeof(void*)));
    void* host_ptrs[4] = {k0, k1, v0, v1};
    B40C_HIP(hipMemcpyAsync(sc.bufptrs.ptr, host_ptrs, sizeof(host_ptrs), hipMemcpyHostToDevice, stream));
    B40C_HIP(hipMemsetAsync(sc.flags.ptr, 0, 2 * sizeof(uint32_t), stream));
    K* const* kbufs = (K* const*)sc.bufptrs.ptr;
    uint32_t* const* vbufs = (uint32_t* const*)((void**)sc.bufptrs.ptr + 2);
    int32_t* d_sel = (int32_t*)sc.flags.ptr;
    uint32_t* d_trivial = (uint32_t*)sc.flags.ptr + 1;
    uint32_t* spine = (uint32_t*)sc.spine.ptr;
    uint32_t* dtot = (uint32_t*)sc.digit_total.ptr;

    for (int shift = begin_bit; shift < end_bit; shift += 8) {
        const int bits = min(8, end_bit - shift);
        const uint32_t mask = (1u << bits) - 1u;
        B40C_HIP(hipMemsetAsync(dtot, 0, RS_BINS * sizeof(uint32_t), stream));
        hipLaunchKernelGGL(k_upsweep<K>, dim3(G), dim3(RS_THREADS), 0, stream, kbufs, d_sel, n, shift, mask,
                           tiles_per_wg, spine, dtot, G);
        hipLaunchKernelGGL(k_spine, dim3(RS_BINS), dim3(RS_THREADS), 0, stream, spine, dtot, G, n, d_trivial);
        hipLaunchKernelGGL((k_downsweep<K, VALS>), dim3(G), dim3(RS_THREADS), 0, stream, kbufs, vbufs, d_sel, n,
                           shift, mask, tiles_per_wg, spine, G, d_trivial);
        hipLaunchKernelGGL(k_flip_selector, dim3(1), dim3(1), 0, stream, d_sel, d_trivial);
    }
    B40C_HIP(hipGetLastError());
    int32_t h_sel = 0;
    B40C_HIP(hipMemcpyAsync(&h_sel, d_sel, sizeof(int32_t), hipMemcpyDeviceToHost, stream));
    B40C_HIP(hipStreamSynchronize(stream));
    *selector = h_sel;
    return B40C_OK;
}

}  // namespace

b40c_status radix_sort_u64(uint64_t* k0, uint64_t* k1, int64_t n, int begin_bit, int end_bit, int32_t* selector,
                           hipStream_t stream) {
    return radix_sort_impl<uint64_t, false>(k0, k1, nullptr, nullptr, n, begin_bit, end_bit, selector, stream);
}

b40c_status radix_sort_pairs_u32(uint32_t* k0, uint32_t* k1, uint32_t* v0, uint32_t* v1, int64_t n, int begin_bit,
                                 int end_bit, int32_t* selector, hipStream_t stream) {
    return radix_sort_impl<uint32_t, true>(k0, k1, v0, v1, n, begin_bit, end_bit, selector, stream);
}

}  // namespace b40c

extern "C" b40c_status b40c_radix_sort_u64(uint64_t* d_keys0, uint64_t* d_keys1, int64_t n, int32_t begin_bit,
                                           int32_t end_bit, int32_t* selector) {
    if (!selector || (n > 0 && (!d_keys0 || !d_keys1)))
        return b40c::fail(B40C_ERR_INVALID_ARG, "b40c_radix_sort_u64: null pointer");
    return b40c::radix_sort_u64(d_keys0, d_keys1, n, begin_bit, end_bit, selector, nullptr);
}

extern "C" b40c_status b40c_radix_sort_pairs_u32(uint32_t* d_keys0, uint32_t* d_keys1, uint32_t* d_vals0,
                                                 uint32_t* d_vals1, int64_t n, int32_t begin_bit, int32_t end_bit,
                                                 int32_t* selector) {
    if (!selector || (n > 0 && (!d_keys0 || !d_keys1 || !d_vals0 || !d_vals1)))
        return b40c::fail(B40C_ERR_INVALID_ARG, "b40c_radix_sort_pairs_u32: null pointer");
    return b40c::radix_sort_pairs_u32(d_keys0, d_keys1, d_vals0, d_vals1, n, begin_bit, end_bit, selector, nullptr);
}
