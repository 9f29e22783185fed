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
//  * Signed and floating-point keys (KeyTraits, radix_sort/sort_utils.cuh:
//    87-183) are ordered by the reference's bit twiddles, applied in
//    registers whenever a digit is extracted: memory always holds the
//    caller's keys, so no pass rewrites them and the trivial-pass skip stays
//    valid for every key type.
//  * The trivial-pass early exit (downsweep/kernel.cuh DownsweepPass) is
//    decided on the device: the spine kernel flags a pass whose keys all share
//    one digit, the downsweep then does nothing and a device-side selector
//    (DoubleBuffer::selector) is left unflipped.
//  * Scratch (spine, digit totals, selector, trivial flag) comes from the
//    caller as a SortScratch view: a per-call allocation for the stateless
//    entry points, the sort enactor's own buffer, or a fixed device array for
//    the allocation-free entry point.  Launching a sort allocates nothing and
//    does not synchronize; only reading the selector back does.
#include <mutex>
#include <vector>

#include "common.hpp"
#include "internal.hpp"

namespace b40c {
namespace {

constexpr int RS_THREADS = 256;
constexpr int RS_WAVES = RS_THREADS / kWave;
constexpr int RS_BINS = 256;
constexpr int RS_MAX_WGS_PER_CU = 8;

// Order-preserving views of the key bits (sort_utils.cuh:87-183):
// CONV 0 unsigned (as is), 1 signed (flip the sign bit), 2 IEEE float
// (negative: flip all bits; positive: flip the sign bit).
enum KeyConv { KC_UNSIGNED = 0, KC_SIGNED = 1, KC_FLOAT = 2 };

template <typename K, int CONV>
__device__ __forceinline__ K ordered_bits(K k) {
    constexpr K HIGH = (K)1 << (8 * sizeof(K) - 1);
    if (CONV == KC_SIGNED) return k ^ HIGH;
    if (CONV == KC_FLOAT) return k ^ ((k & HIGH) ? (K)~(K)0 : HIGH);
    return k;
}

template <typename K, int CONV>
__device__ __forceinline__ uint32_t digit_of(K key, int shift, uint32_t mask) {
    return (uint32_t)(ordered_bits<K, CONV>(key) >> shift) & mask;
}

// Start of a sort: digit totals of the first pass, selector and trivial flag
// cleared (kernels, not memsets, so a caller may capture the launches).
__global__ void k_sort_init(uint32_t* digit_total0, int32_t* sel, uint32_t* trivial) {
    digit_total0[threadIdx.x] = 0;
    if (threadIdx.x == 0) {
        *sel = 0;
        *trivial = 0;
    }
}

// Per-workgroup digit histogram of its tile range; counts go to the spine
// (digit-major) and to the global digit totals.
template <typename K, int CONV>
__global__ __launch_bounds__(RS_THREADS) void k_upsweep(const K* k0, const K* k1, const int32_t* sel, int64_t n,
                                                        int shift, uint32_t mask, int64_t keys_per_wg,
                                                        uint32_t* spine, uint32_t* digit_total, int G) {
    __shared__ uint32_t hist[RS_WAVES][RS_BINS];
    const int tid = threadIdx.x, wid = tid / kWave;
    for (int i = tid; i < RS_WAVES * RS_BINS; i += RS_THREADS) (&hist[0][0])[i] = 0;
    __syncthreads();
    const K* keys = *sel ? k1 : k0;
    const int64_t begin = (int64_t)blockIdx.x * keys_per_wg;
    const int64_t end = min(n, begin + keys_per_wg);
    // RS_UP keys per thread per round, their loads issued together.
    constexpr int RS_UP = 8;
    int64_t i = begin + tid;
    for (; i + (int64_t)(RS_UP - 1) * RS_THREADS < end; i += (int64_t)RS_UP * RS_THREADS) {
        K kk[RS_UP];
#pragma unroll
        for (int j = 0; j < RS_UP; ++j) kk[j] = keys[i + (int64_t)j * RS_THREADS];
#pragma unroll
        for (int j = 0; j < RS_UP; ++j) atomicAdd(&hist[wid][digit_of<K, CONV>(kk[j], shift, mask)], 1u);
    }
    for (; i < end; i += RS_THREADS) atomicAdd(&hist[wid][digit_of<K, CONV>(keys[i], shift, mask)], 1u);
    __syncthreads();
    uint32_t c = 0;
#pragma unroll
    for (int w = 0; w < RS_WAVES; ++w) c += hist[w][tid];
    spine[(int64_t)tid * G + blockIdx.x] = c;
    if (c) atomic_add_u32(&digit_total[tid], c);
}

// Exclusive scan of the spine, one workgroup per digit.  Digit d's base is the
// sum of the totals of the smaller digits.  Flags a trivial pass, and
// workgroup 0 clears the other digit-total buffer for the next pass.
__global__ __launch_bounds__(RS_THREADS) void k_spine(uint32_t* spine, const uint32_t* digit_total,
                                                      uint32_t* next_digit_total, int G, int64_t n,
                                                      uint32_t* trivial) {
    __shared__ int s_wave[RS_WAVES];
    const int d = blockIdx.x, tid = threadIdx.x;
    const uint32_t t = digit_total[tid];
    if (d == 0) next_digit_total[tid] = 0;
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

// ITEMS keys per thread: a tile of RS_THREADS * ITEMS keys is ranked and
// staged in LDS at once.  Larger tiles give longer per-digit runs in the
// scatter (about ITEMS keys per digit on uniform keys), so fewer partial
// cache lines are written, for more LDS per workgroup.
template <typename K, int CONV, bool VALS, int ITEMS>
__global__ __launch_bounds__(RS_THREADS) void k_downsweep(K* k0, K* k1, uint32_t* v0, uint32_t* v1,
                                                          const int32_t* sel, int64_t n, int shift, uint32_t mask,
                                                          int64_t keys_per_wg, const uint32_t* spine, int G,
                                                          const uint32_t* trivial) {
    constexpr int RS_TILE = RS_THREADS * ITEMS;
    if (*trivial) return;
    __shared__ K stage_k[RS_TILE];
    __shared__ uint32_t stage_v[VALS ? RS_TILE : 1];
    __shared__ uint32_t wave_hist[RS_WAVES][RS_BINS];
    __shared__ uint32_t carry[RS_BINS];
    __shared__ uint32_t tile_excl[RS_BINS];
    __shared__ int s_wave[RS_WAVES];

    const int tid = threadIdx.x, lane = lane_id(), wid = tid / kWave;
    const int s = *sel;
    const K* kin = s ? k1 : k0;
    K* kout = s ? k0 : k1;
    const uint32_t* vin = VALS ? (s ? v1 : v0) : nullptr;
    uint32_t* vout = VALS ? (s ? v0 : v1) : nullptr;

    carry[tid] = spine[(int64_t)tid * G + blockIdx.x];
    const int64_t begin = (int64_t)blockIdx.x * keys_per_wg;
    const int64_t end = min(n, begin + keys_per_wg);
    const uint64_t lt_mask = (lane == 0) ? 0ull : (~0ull >> (64 - lane));

    // The next tile's keys (and values) are loaded into registers while the
    // current tile is ranked and scattered.
    K nkey[ITEMS];
    uint32_t nval[ITEMS];
    auto load_tile = [&](int64_t tb) {
        const int tn = (int)min((int64_t)RS_TILE, end - tb);
#pragma unroll
        for (int j = 0; j < ITEMS; ++j) {
            const int idx = wid * (kWave * ITEMS) + j * kWave + lane;
            const bool valid = idx < tn;
            nkey[j] = valid ? kin[tb + idx] : (K)0;
            if (VALS) nval[j] = valid ? vin[tb + idx] : 0u;
        }
    };
    if (begin < end) load_tile(begin);
    for (int64_t tile_base = begin; tile_base < end; tile_base += RS_TILE) {
        const int tile_n = (int)min((int64_t)RS_TILE, end - tile_base);
        for (int i = lane; i < RS_BINS; i += kWave) wave_hist[wid][i] = 0;
        __syncthreads();

        K key[ITEMS];
        uint32_t val[ITEMS];
        uint32_t rank[ITEMS];
        uint32_t dig[ITEMS];
#pragma unroll
        for (int j = 0; j < ITEMS; ++j) {
            key[j] = nkey[j];
            if (VALS) val[j] = nval[j];
        }
        if (tile_base + RS_TILE < end) load_tile(tile_base + RS_TILE);
#pragma unroll
        for (int j = 0; j < ITEMS; ++j) {
            const int idx = wid * (kWave * ITEMS) + j * kWave + lane;
            const bool valid = idx < tile_n;
            const uint32_t d = digit_of<K, CONV>(key[j], shift, mask);
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
        for (int j = 0; j < ITEMS; ++j) {
            if (rank[j] != 0xFFFFFFFFu) {
                const uint32_t pos = tile_excl[dig[j]] + wave_hist[wid][dig[j]] + rank[j];
                stage_k[pos] = key[j];
                if (VALS) stage_v[pos] = val[j];
            }
        }
        __syncthreads();
        for (int k = tid; k < tile_n; k += RS_THREADS) {
            const K kk = stage_k[k];
            const uint32_t d = digit_of<K, CONV>(kk, shift, mask);
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

// Launch shape for n keys on a device with `num_cus` CUs: at most
// policy.wgs_per_cu workgroups per CU, each owning a whole number of tiles.
struct SortPlan {
    int64_t keys_per_wg = 1;
    int G = 1;
    int items = 8;
};

SortPlan sort_plan(int64_t n, int num_cus, const SortPolicy& pol) {
    SortPlan p;
    p.items = pol.items;
    const int64_t tile = (int64_t)RS_THREADS * pol.items;
    const int64_t num_tiles = std::max<int64_t>(1, (n + tile - 1) / tile);
    const int64_t max_grid = (int64_t)num_cus * pol.wgs_per_cu;
    const int64_t tiles_per_wg = (num_tiles + max_grid - 1) / max_grid;
    p.keys_per_wg = tiles_per_wg * tile;
    p.G = (int)((num_tiles + tiles_per_wg - 1) / tiles_per_wg);
    return p;
}

template <typename K, int CONV, bool VALS, int ITEMS>
void launch_passes(K* k0, K* k1, uint32_t* v0, uint32_t* v1, int64_t n, int begin_bit, int end_bit,
                   const SortPlan& plan, const SortScratch& sc, hipStream_t stream) {
    uint32_t* dtot[2] = {sc.digit_total, sc.digit_total + RS_BINS};
    hipLaunchKernelGGL(k_sort_init, dim3(1), dim3(RS_BINS), 0, stream, dtot[0], sc.sel, sc.trivial);
    int pass = 0;
    for (int shift = begin_bit; shift < end_bit; shift += 8, ++pass) {
        const int bits = std::min(8, end_bit - shift);
        const uint32_t mask = (1u << bits) - 1u;
        uint32_t* cur = dtot[pass & 1];
        uint32_t* next = dtot[(pass + 1) & 1];
        hipLaunchKernelGGL((k_upsweep<K, CONV>), dim3(plan.G), dim3(RS_THREADS), 0, stream, k0, k1, sc.sel, n, shift,
                           mask, plan.keys_per_wg, sc.spine, cur, plan.G);
        hipLaunchKernelGGL(k_spine, dim3(RS_BINS), dim3(RS_THREADS), 0, stream, sc.spine, cur, next, plan.G, n,
                           sc.trivial);
        hipLaunchKernelGGL((k_downsweep<K, CONV, VALS, ITEMS>), dim3(plan.G), dim3(RS_THREADS), 0, stream, k0, k1, v0,
                           v1, sc.sel, n, shift, mask, plan.keys_per_wg, sc.spine, plan.G, sc.trivial);
        hipLaunchKernelGGL(k_flip_selector, dim3(1), dim3(1), 0, stream, sc.sel, sc.trivial);
    }
}

template <typename K, int CONV>
void launch_typed(void* k0, void* k1, uint32_t* v0, uint32_t* v1, int64_t n, int begin_bit, int end_bit,
                  const SortPlan& plan, const SortScratch& sc, hipStream_t stream) {
    if (v0) {
        if (plan.items == 16)
            launch_passes<K, CONV, true, 16>((K*)k0, (K*)k1, v0, v1, n, begin_bit, end_bit, plan, sc, stream);
        else
            launch_passes<K, CONV, true, 8>((K*)k0, (K*)k1, v0, v1, n, begin_bit, end_bit, plan, sc, stream);
    } else {
        if (plan.items == 16)
            launch_passes<K, CONV, false, 16>((K*)k0, (K*)k1, nullptr, nullptr, n, begin_bit, end_bit, plan, sc,
                                              stream);
        else
            launch_passes<K, CONV, false, 8>((K*)k0, (K*)k1, nullptr, nullptr, n, begin_bit, end_bit, plan, sc,
                                             stream);
    }
}

int key_bits(int32_t key_type) {
    switch (key_type) {
        case B40C_KEY_U32:
        case B40C_KEY_I32:
        case B40C_KEY_F32:
            return 32;
        case B40C_KEY_U64:
        case B40C_KEY_I64:
        case B40C_KEY_F64:
            return 64;
    }
    return 0;
}

b40c_status validate(const char* who, int32_t key_type, const void* k0, const void* k1, const uint32_t* v0,
                     const uint32_t* v1, int64_t n, int begin_bit, int end_bit, const int32_t* selector) {
    const int kb = key_bits(key_type);
    if (!kb) return fail(B40C_ERR_INVALID_ARG, "%s: unknown key type %d", who, key_type);
    if (!selector || (n > 0 && (!k0 || !k1)) || ((v0 == nullptr) != (v1 == nullptr)))
        return fail(B40C_ERR_INVALID_ARG, "%s: null pointer", who);
    if (n < 0 || begin_bit < 0 || end_bit > kb || begin_bit > end_bit)
        return fail(B40C_ERR_INVALID_ARG, "%s: bad arguments (n=%lld bits=[%d,%d))", who, (long long)n, begin_bit,
                    end_bit);
    if (n > 0xFFFFFFFFll) return fail(B40C_ERR_UNSUPPORTED, "%s: n=%lld exceeds 2^32-1", who, (long long)n);
    return B40C_OK;
}

}  // namespace

// Defaults: 8 keys per thread and 4 workgroups per CU.
SortPolicy resolve_sort_policy(SortPolicy p, int64_t n, int key_bytes) {
    (void)n;
    (void)key_bytes;
    if (p.items != 8 && p.items != 16) p.items = 8;
    if (p.wgs_per_cu < 1 || p.wgs_per_cu > RS_MAX_WGS_PER_CU) p.wgs_per_cu = 4;
    return p;
}

size_t sort_scratch_words(int64_t n, int num_cus, const SortPolicy& pol) {
    const SortPlan p = sort_plan(n, num_cus, pol);
    return (size_t)RS_BINS * (size_t)p.G + 2 * RS_BINS + 2;
}

SortScratch sort_scratch_view(uint32_t* words, size_t count) {
    SortScratch sc;
    sc.words = count;
    sc.digit_total = words;
    sc.sel = (int32_t*)(words + 2 * RS_BINS);
    sc.trivial = words + 2 * RS_BINS + 1;
    sc.spine = words + 2 * RS_BINS + 2;
    return sc;
}

b40c_status radix_sort_async(int32_t key_type, void* k0, void* k1, uint32_t* v0, uint32_t* v1, int64_t n,
                             int begin_bit, int end_bit, int num_cus, const SortPolicy& policy,
                             const SortScratch& sc, hipStream_t stream, bool& launched) {
    launched = false;
    if (n <= 1 || begin_bit == end_bit) return B40C_OK;
    const SortPolicy pol = resolve_sort_policy(policy, n, key_bits(key_type) / 8);
    const SortPlan plan = sort_plan(n, num_cus, pol);
    if (sort_scratch_words(n, num_cus, pol) > sc.words)
        return fail(B40C_ERR_INVALID_ARG, "radix sort: scratch of %zu words is smaller than the %zu needed", sc.words,
                    sort_scratch_words(n, num_cus, pol));
    switch (key_type) {
        case B40C_KEY_U32: launch_typed<uint32_t, KC_UNSIGNED>(k0, k1, v0, v1, n, begin_bit, end_bit, plan, sc, stream); break;
        case B40C_KEY_I32: launch_typed<uint32_t, KC_SIGNED>(k0, k1, v0, v1, n, begin_bit, end_bit, plan, sc, stream); break;
        case B40C_KEY_F32: launch_typed<uint32_t, KC_FLOAT>(k0, k1, v0, v1, n, begin_bit, end_bit, plan, sc, stream); break;
        case B40C_KEY_U64: launch_typed<uint64_t, KC_UNSIGNED>(k0, k1, v0, v1, n, begin_bit, end_bit, plan, sc, stream); break;
        case B40C_KEY_I64: launch_typed<uint64_t, KC_SIGNED>(k0, k1, v0, v1, n, begin_bit, end_bit, plan, sc, stream); break;
        case B40C_KEY_F64: launch_typed<uint64_t, KC_FLOAT>(k0, k1, v0, v1, n, begin_bit, end_bit, plan, sc, stream); break;
        default: return fail(B40C_ERR_INVALID_ARG, "radix sort: unknown key type %d", key_type);
    }
    B40C_HIP(hipGetLastError());
    launched = true;
    return B40C_OK;
}

// Runs a sort with its scratch and reads the selector back (one stream
// synchronization).
b40c_status radix_sort_sync(int32_t key_type, void* k0, void* k1, uint32_t* v0, uint32_t* v1, int64_t n, int begin_bit,
                            int end_bit, int num_cus, const SortPolicy& policy, const SortScratch& sc,
                            int32_t* selector, hipStream_t stream) {
    bool launched = false;
    B40C_TRY(radix_sort_async(key_type, k0, k1, v0, v1, n, begin_bit, end_bit, num_cus, policy, sc, stream,
                              launched));
    int32_t h_sel = 0;
    if (launched) {
        B40C_HIP(hipMemcpyAsync(&h_sel, sc.sel, sizeof(int32_t), hipMemcpyDeviceToHost, stream));
        B40C_HIP(hipStreamSynchronize(stream));
    }
    *selector = h_sel;
    return B40C_OK;
}

// Stateless form: scratch allocated for this call.
b40c_status radix_sort_once(int32_t key_type, void* k0, void* k1, uint32_t* v0, uint32_t* v1, int64_t n, int begin_bit,
                            int end_bit, int32_t* selector, hipStream_t stream) {
    *selector = 0;
    if (n <= 1 || begin_bit == end_bit) return B40C_OK;
    DeviceInfo info;
    B40C_TRY(device_info(info));
    DeviceBuffer mem;
    const SortPolicy pol = resolve_sort_policy(SortPolicy{}, n, key_bits(key_type) / 8);
    const size_t words = sort_scratch_words(n, info.num_cus, pol);
    B40C_TRY(mem.alloc(words * sizeof(uint32_t)));
    return radix_sort_sync(key_type, k0, k1, v0, v1, n, begin_bit, end_bit, info.num_cus, pol,
                           sort_scratch_view(mem.as<uint32_t>(), words), selector, stream);
}

b40c_status radix_sort_u64(uint64_t* k0, uint64_t* k1, int64_t n, int begin_bit, int end_bit, int32_t* selector,
                           hipStream_t stream) {
    B40C_TRY(validate("radix sort", B40C_KEY_U64, k0, k1, nullptr, nullptr, n, begin_bit, end_bit, selector));
    return radix_sort_once(B40C_KEY_U64, k0, k1, nullptr, nullptr, n, begin_bit, end_bit, selector, stream);
}

b40c_status radix_sort_pairs_u32(uint32_t* k0, uint32_t* k1, uint32_t* v0, uint32_t* v1, int64_t n, int begin_bit,
                                 int end_bit, int32_t* selector, hipStream_t stream) {
    B40C_TRY(validate("radix sort", B40C_KEY_U32, k0, k1, v0, v1, n, begin_bit, end_bit, selector));
    return radix_sort_once(B40C_KEY_U32, k0, k1, v0, v1, n, begin_bit, end_bit, selector, stream);
}

}  // namespace b40c

// ---------------------------------------------------------------------------
// Allocation-free sort (no counterpart in the reference): its scratch is this
// fixed device array, one copy per GPU, which the GPU keeps until the
// process ends.  The grid is capped at RS_MAX_WGS_PER_CU workgroups per CU,
// so kNoAllocWords covers a sort of any n <= 2^32 - 1 under any policy on a
// device with up to kNoAllocMaxCus CUs (256 on MI355X): 2 MiB + 2 KiB.  In host code `scratch`
// is only a stand-in for the device array; its device address comes from
// hipGetSymbolAddress, once per device, and only that address is used.
// ---------------------------------------------------------------------------
namespace b40c {
namespace {
constexpr int kNoAllocMaxCus = 256;
constexpr size_t kNoAllocWords = (size_t)RS_BINS * RS_MAX_WGS_PER_CU * kNoAllocMaxCus + 2 * RS_BINS + 64;
}  // namespace
}  // namespace b40c

static __device__ __align__(256) uint32_t scratch[b40c::kNoAllocWords];

namespace b40c {
namespace {
constexpr int kNoAllocMaxDevices = 64;
struct NoAllocSlot {
    std::once_flag once;
    hipError_t err = hipSuccess;
    uint32_t* ptr = nullptr;  // device address of `scratch` on this device
    int num_cus = 0;
};
NoAllocSlot g_noalloc[kNoAllocMaxDevices];

// Resolves the calling thread's current device's copy of `scratch` (first
// call on a device only).
b40c_status noalloc_slot(NoAllocSlot*& out) {
    int dev = 0;
    B40C_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= kNoAllocMaxDevices)
        return fail(B40C_ERR_UNSUPPORTED, "allocation-free sort: device %d beyond the %d supported", dev,
                    kNoAllocMaxDevices);
    NoAllocSlot& slot = g_noalloc[dev];
    std::call_once(slot.once, [&] {
        slot.err = hipGetSymbolAddress((void**)&slot.ptr, HIP_SYMBOL(scratch));
        if (slot.err == hipSuccess) slot.err = hipDeviceGetAttribute(&slot.num_cus, hipDeviceAttributeMultiprocessorCount, dev);
    });
    if (slot.err != hipSuccess)
        return fail(B40C_ERR_HIP, "allocation-free sort: resolving the scratch array failed: %s",
                    hipGetErrorString(slot.err));
    out = &slot;
    return B40C_OK;
}
}  // namespace
}  // namespace b40c

// Sort enactor: owns its scratch across calls, like the reference Enactor's
// spine storage (radix_sort/enactor.cuh:96-140).  One stream per enactor.
struct b40c_sort_enactor {
    b40c::DeviceBuffer scratch;
    hipStream_t stream = nullptr;
    b40c::SortPolicy policy;  // zero fields: the size-based default
};

using namespace b40c;

extern "C" {

b40c_status b40c_radix_sort_u64(uint64_t* d_keys0, uint64_t* d_keys1, int64_t n, int32_t begin_bit, int32_t end_bit,
                                int32_t* selector) {
    B40C_TRY(validate("b40c_radix_sort_u64", B40C_KEY_U64, d_keys0, d_keys1, nullptr, nullptr, n, begin_bit, end_bit,
                      selector));
    return radix_sort_once(B40C_KEY_U64, d_keys0, d_keys1, nullptr, nullptr, n, begin_bit, end_bit, selector, nullptr);
}

b40c_status b40c_radix_sort_pairs_u32(uint32_t* d_keys0, uint32_t* d_keys1, uint32_t* d_vals0, uint32_t* d_vals1,
                                      int64_t n, int32_t begin_bit, int32_t end_bit, int32_t* selector) {
    if (n > 0 && (!d_vals0 || !d_vals1))
        return fail(B40C_ERR_INVALID_ARG, "b40c_radix_sort_pairs_u32: null pointer");
    B40C_TRY(validate("b40c_radix_sort_pairs_u32", B40C_KEY_U32, d_keys0, d_keys1, d_vals0, d_vals1, n, begin_bit,
                      end_bit, selector));
    return radix_sort_once(B40C_KEY_U32, d_keys0, d_keys1, d_vals0, d_vals1, n, begin_bit, end_bit, selector, nullptr);
}

b40c_status b40c_radix_sort(int32_t key_type, void* d_keys0, void* d_keys1, uint32_t* d_vals0, uint32_t* d_vals1,
                            int64_t n, int32_t begin_bit, int32_t end_bit, int32_t* selector) {
    B40C_TRY(validate("b40c_radix_sort", key_type, d_keys0, d_keys1, d_vals0, d_vals1, n, begin_bit, end_bit, selector));
    return radix_sort_once(key_type, d_keys0, d_keys1, d_vals0, d_vals1, n, begin_bit, end_bit, selector, nullptr);
}

b40c_status b40c_radix_sort_u64_noalloc(uint64_t* d_keys0, uint64_t* d_keys1, int64_t n, int32_t begin_bit,
                                        int32_t end_bit, int32_t* selector) {
    B40C_TRY(validate("b40c_radix_sort_u64_noalloc", B40C_KEY_U64, d_keys0, d_keys1, nullptr, nullptr, n, begin_bit,
                      end_bit, selector));
    *selector = 0;
    if (n <= 1 || begin_bit == end_bit) return B40C_OK;
    NoAllocSlot* slot = nullptr;
    B40C_TRY(noalloc_slot(slot));
    const SortPolicy pol = resolve_sort_policy(SortPolicy{}, n, 8);
    const size_t need = sort_scratch_words(n, slot->num_cus, pol);
    if (need > kNoAllocWords)
        return fail(B40C_ERR_UNSUPPORTED,
                    "b40c_radix_sort_u64_noalloc: n=%lld on %d CUs needs %zu scratch words; the fixed array holds %zu",
                    (long long)n, slot->num_cus, need, kNoAllocWords);
    return radix_sort_sync(B40C_KEY_U64, d_keys0, d_keys1, nullptr, nullptr, n, begin_bit, end_bit, slot->num_cus,
                           pol, sort_scratch_view(slot->ptr, kNoAllocWords), selector, nullptr);
}

b40c_status b40c_sort_enactor_create(b40c_sort_enactor** out) {
    if (!out) return fail(B40C_ERR_INVALID_ARG, "b40c_sort_enactor_create: null out");
    b40c_sort_enactor* e = new b40c_sort_enactor();
    const hipError_t err = hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking);
    if (err != hipSuccess) {
        delete e;
        return fail(B40C_ERR_HIP, "b40c_sort_enactor_create: %s", hipGetErrorString(err));
    }
    *out = e;
    return B40C_OK;
}

b40c_status b40c_sort_enactor_sort(b40c_sort_enactor* e, int32_t key_type, void* d_keys0, void* d_keys1,
                                   uint32_t* d_vals0, uint32_t* d_vals1, int64_t n, int32_t begin_bit, int32_t end_bit,
                                   int32_t* selector) {
    if (!e) return fail(B40C_ERR_INVALID_ARG, "b40c_sort_enactor_sort: null enactor");
    B40C_TRY(validate("b40c_sort_enactor_sort", key_type, d_keys0, d_keys1, d_vals0, d_vals1, n, begin_bit, end_bit,
                      selector));
    *selector = 0;
    if (n <= 1 || begin_bit == end_bit) return B40C_OK;
    DeviceInfo info;
    B40C_TRY(device_info(info));
    const SortPolicy pol = resolve_sort_policy(e->policy, n, key_bits(key_type) / 8);
    const size_t words = sort_scratch_words(n, info.num_cus, pol);
    // Grow-only: later sorts of at most this size allocate nothing.
    B40C_TRY(e->scratch.reserve(words * sizeof(uint32_t)));
    return radix_sort_sync(key_type, d_keys0, d_keys1, d_vals0, d_vals1, n, begin_bit, end_bit, info.num_cus, pol,
                           sort_scratch_view(e->scratch.as<uint32_t>(), e->scratch.bytes / sizeof(uint32_t)), selector,
                           e->stream);
}

b40c_status b40c_sort_enactor_set_policy(b40c_sort_enactor* e, int32_t items_per_thread, int32_t wgs_per_cu) {
    if (!e) return fail(B40C_ERR_INVALID_ARG, "b40c_sort_enactor_set_policy: null enactor");
    if ((items_per_thread != 0 && items_per_thread != 8 && items_per_thread != 16) || wgs_per_cu < 0 ||
        wgs_per_cu > RS_MAX_WGS_PER_CU)
        return fail(B40C_ERR_INVALID_ARG,
                    "b40c_sort_enactor_set_policy: items_per_thread must be 0, 8 or 16 and wgs_per_cu 0-%d (got %d, %d)",
                    RS_MAX_WGS_PER_CU, items_per_thread, wgs_per_cu);
    e->policy.items = items_per_thread;
    e->policy.wgs_per_cu = wgs_per_cu;
    return B40C_OK;
}

void b40c_sort_enactor_destroy(b40c_sort_enactor* e) {
    if (!e) return;
    if (e->stream) {
        (void)hipStreamSynchronize(e->stream);
        (void)hipStreamDestroy(e->stream);
    }
    delete e;
}

}  // extern "C"
