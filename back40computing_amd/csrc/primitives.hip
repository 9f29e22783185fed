// Device primitives next to the CSR build, each a chunked three-pass
// (count or reduce per chunk, one-workgroup spine, ordered scatter) over
// coalesced 256-element tiles:
//  * consecutive removal — consecutive_removal::Enactor::Trim
//    (b40c/consecutive_removal/enactor.cuh:119-126): the first element of
//    every run of equal keys (and its value), in input order;
//  * prefix sum — scan::Enactor::Scan (b40c/scan/enactor.cuh:122-136),
//    exclusive or inclusive, unsigned 32-bit arithmetic.
// Chunks are contiguous, so each workgroup streams its own slice; the chunk
// count is a few workgroups per CU, enough to fill the chip while keeping the
// spine to one workgroup.
#include "internal.hpp"

namespace b40c {
namespace {

constexpr int PT = 256;  // threads per workgroup = elements per tile
constexpr int PW = PT / kWave;

template <typename T>
__device__ __forceinline__ T wave_scan_incl(T x) {
    const int lane = lane_id();
#pragma unroll
    for (int d = 1; d < kWave; d <<= 1) {
        const T y = __shfl_up(x, d, kWave);
        if (lane >= d) x += y;
    }
    return x;
}

// Workgroup exclusive scan; `total` is the workgroup's sum.
template <typename T>
__device__ __forceinline__ T wg_scan_excl(T x, T& total, T* s_wave) {
    const int wid = threadIdx.x / kWave;
    const T incl = wave_scan_incl(x);
    if (lane_id() == kWave - 1) s_wave[wid] = incl;
    __syncthreads();
    T base = 0, tot = 0;
#pragma unroll
    for (int w = 0; w < PW; ++w) {
        const T t = s_wave[w];
        base += (w < wid) ? t : T(0);
        tot += t;
    }
    __syncthreads();
    total = tot;
    return base + incl - x;
}

__device__ __forceinline__ bool run_head(const uint64_t* keys, int64_t i) { return i == 0 || keys[i - 1] != keys[i]; }

__global__ __launch_bounds__(PT) void k_trim_count(const uint64_t* keys, int64_t n, int64_t chunk, uint64_t* counts) {
    __shared__ uint64_t s_wave[PW];
    const int64_t begin = (int64_t)blockIdx.x * chunk, end = min(n, begin + chunk);
    uint64_t c = 0;
    for (int64_t i = begin + threadIdx.x; i < end; i += PT) c += run_head(keys, i) ? 1u : 0u;
    uint64_t total;
    (void)wg_scan_excl<uint64_t>(c, total, s_wave);
    if (threadIdx.x == 0) counts[blockIdx.x] = total;
}

// Exclusive scan, in place, of the G per-chunk values (one workgroup);
// *total receives their sum.
__global__ __launch_bounds__(PT) void k_spine_u64(uint64_t* v, int G, uint64_t* total) {
    __shared__ uint64_t s_wave[PW];
    uint64_t carry = 0;
    for (int base = 0; base < G; base += PT) {
        const int i = base + threadIdx.x;
        const uint64_t x = (i < G) ? v[i] : 0;
        uint64_t tot;
        const uint64_t ex = wg_scan_excl<uint64_t>(x, tot, s_wave);
        if (i < G) v[i] = carry + ex;
        carry += tot;
    }
    if (threadIdx.x == 0) *total = carry;
}

__global__ __launch_bounds__(PT) void k_trim_scatter(const uint64_t* keys, const uint32_t* vals, int64_t n,
                                                     int64_t chunk, const uint64_t* offsets, uint64_t* keys_out,
                                                     uint32_t* vals_out) {
    __shared__ uint32_t s_wave[PW];
    const int64_t begin = (int64_t)blockIdx.x * chunk, end = min(n, begin + chunk);
    uint64_t carry = offsets[blockIdx.x];
    for (int64_t base = begin; base < end; base += PT) {
        const int64_t i = base + threadIdx.x;
        const bool keep = i < end && run_head(keys, i);
        uint32_t tot;
        const uint32_t ex = wg_scan_excl<uint32_t>(keep ? 1u : 0u, tot, s_wave);
        if (keep) {
            keys_out[carry + ex] = keys[i];
            if (vals) vals_out[carry + ex] = vals[i];
        }
        carry += tot;
    }
}

__global__ __launch_bounds__(PT) void k_scan_reduce(const uint32_t* in, int64_t n, int64_t chunk, uint64_t* sums) {
    __shared__ uint64_t s_wave[PW];
    const int64_t begin = (int64_t)blockIdx.x * chunk, end = min(n, begin + chunk);
    uint64_t s = 0;
    for (int64_t i = begin + threadIdx.x; i < end; i += PT) s += in[i];
    uint64_t total;
    (void)wg_scan_excl<uint64_t>(s, total, s_wave);
    if (threadIdx.x == 0) sums[blockIdx.x] = total;
}

// Each tile's values are in registers before the tile's first barrier, so
// `out` may alias `in`.
__global__ __launch_bounds__(PT) void k_scan_down(const uint32_t* in, uint32_t* out, int64_t n, int64_t chunk,
                                                  const uint64_t* offsets, int inclusive) {
    __shared__ uint32_t s_wave[PW];
    const int64_t begin = (int64_t)blockIdx.x * chunk, end = min(n, begin + chunk);
    uint32_t carry = (uint32_t)offsets[blockIdx.x];
    for (int64_t base = begin; base < end; base += PT) {
        const int64_t i = base + threadIdx.x;
        const uint32_t x = (i < end) ? in[i] : 0u;
        uint32_t tot;
        const uint32_t ex = wg_scan_excl<uint32_t>(x, tot, s_wave);
        if (i < end) out[i] = carry + ex + (inclusive ? x : 0u);
        carry += tot;
    }
}

// ---- vector paths: 16-byte loads, STEP elements per workgroup iteration ----
// A thread owns VEC consecutive elements of each of SUB sub-tiles, so every
// load instruction of a wave covers a contiguous 1 KiB and each thread has
// 64 bytes in flight per iteration; one workgroup scan per sub-tile.
constexpr int VEC32 = 4;                   // uint32 per uint4
constexpr int SUB32 = 4;
constexpr int64_t STEP32 = (int64_t)PT * VEC32 * SUB32;  // 4096 elements
constexpr int VEC64 = 4;                   // uint64 per two uint4
constexpr int SUB64 = 2;
constexpr int64_t STEP64 = (int64_t)PT * VEC64 * SUB64;  // 2048 keys

__device__ __forceinline__ void load4_u32(const uint32_t* p, int64_t i, int64_t end, uint32_t (&x)[4]) {
    if (i + 4 <= end) {
        const uint4 q = *reinterpret_cast<const uint4*>(p + i);
        x[0] = q.x, x[1] = q.y, x[2] = q.z, x[3] = q.w;
    } else {
#pragma unroll
        for (int j = 0; j < 4; ++j) x[j] = (i + j < end) ? p[i + j] : 0u;
    }
}

__device__ __forceinline__ void load4_u64(const uint64_t* p, int64_t i, int64_t end, uint64_t (&x)[4]) {
    if (i + 4 <= end) {
        const uint4 a = *reinterpret_cast<const uint4*>(p + i);
        const uint4 b = *reinterpret_cast<const uint4*>(p + i + 2);
        x[0] = ((uint64_t)a.y << 32) | a.x, x[1] = ((uint64_t)a.w << 32) | a.z;
        x[2] = ((uint64_t)b.y << 32) | b.x, x[3] = ((uint64_t)b.w << 32) | b.z;
    } else {
#pragma unroll
        for (int j = 0; j < 4; ++j) x[j] = (i + j < end) ? p[i + j] : 0ull;
    }
}

__global__ __launch_bounds__(PT) void k_scan_reduce_v(const uint32_t* in, int64_t n, int64_t chunk, uint64_t* sums) {
    __shared__ uint64_t s_wave[PW];
    const int64_t begin = (int64_t)blockIdx.x * chunk, end = min(n, begin + chunk);
    uint64_t s = 0;
    for (int64_t base = begin; base < end; base += STEP32) {
        uint32_t x[SUB32][4];
#pragma unroll
        for (int k = 0; k < SUB32; ++k) load4_u32(in, base + (int64_t)k * PT * VEC32 + threadIdx.x * VEC32, end, x[k]);
#pragma unroll
        for (int k = 0; k < SUB32; ++k) s += (uint64_t)x[k][0] + x[k][1] + x[k][2] + x[k][3];
    }
    uint64_t total;
    (void)wg_scan_excl<uint64_t>(s, total, s_wave);
    if (threadIdx.x == 0) sums[blockIdx.x] = total;
}

__global__ __launch_bounds__(PT) void k_scan_down_v(const uint32_t* in, uint32_t* out, int64_t n, int64_t chunk,
                                                    const uint64_t* offsets, int inclusive) {
    __shared__ uint32_t s_wave[PW];
    const int64_t begin = (int64_t)blockIdx.x * chunk, end = min(n, begin + chunk);
    uint32_t carry = (uint32_t)offsets[blockIdx.x];
    for (int64_t base = begin; base < end; base += STEP32) {
        uint32_t x[SUB32][4];
#pragma unroll
        for (int k = 0; k < SUB32; ++k) load4_u32(in, base + (int64_t)k * PT * VEC32 + threadIdx.x * VEC32, end, x[k]);
#pragma unroll
        for (int k = 0; k < SUB32; ++k) {
            const int64_t i = base + (int64_t)k * PT * VEC32 + threadIdx.x * VEC32;
            uint32_t pre[4], run = 0;
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                pre[j] = run + (inclusive ? x[k][j] : 0u);
                run += x[k][j];
            }
            uint32_t tot;
            const uint32_t ex = carry + wg_scan_excl<uint32_t>(run, tot, s_wave);
            if (i + 4 <= end) {
                *reinterpret_cast<uint4*>(out + i) = make_uint4(ex + pre[0], ex + pre[1], ex + pre[2], ex + pre[3]);
            } else {
#pragma unroll
                for (int j = 0; j < 4; ++j)
                    if (i + j < end) out[i + j] = ex + pre[j];
            }
            carry += tot;
        }
    }
}

// Run heads of the VEC64 keys at i (the key before i is read separately).
__device__ __forceinline__ uint32_t heads4(const uint64_t* keys, int64_t i, int64_t end, const uint64_t (&x)[4]) {
    uint64_t prev = (i > 0 && i < end) ? keys[i - 1] : ~x[0];
    uint32_t h = 0;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        if (i + j < end && (i + j == 0 || x[j] != prev)) h |= 1u << j;
        prev = x[j];
    }
    return h;
}

__global__ __launch_bounds__(PT) void k_trim_count_v(const uint64_t* keys, int64_t n, int64_t chunk,
                                                     uint64_t* counts) {
    __shared__ uint64_t s_wave[PW];
    const int64_t begin = (int64_t)blockIdx.x * chunk, end = min(n, begin + chunk);
    uint64_t c = 0;
    for (int64_t base = begin; base < end; base += STEP64) {
        uint64_t x[SUB64][4];
#pragma unroll
        for (int k = 0; k < SUB64; ++k) load4_u64(keys, base + (int64_t)k * PT * VEC64 + threadIdx.x * VEC64, end, x[k]);
#pragma unroll
        for (int k = 0; k < SUB64; ++k)
            c += __popc(heads4(keys, base + (int64_t)k * PT * VEC64 + threadIdx.x * VEC64, end, x[k]));
    }
    uint64_t total;
    (void)wg_scan_excl<uint64_t>(c, total, s_wave);
    if (threadIdx.x == 0) counts[blockIdx.x] = total;
}

__global__ __launch_bounds__(PT) void k_trim_scatter_v(const uint64_t* keys, const uint32_t* vals, int64_t n,
                                                       int64_t chunk, const uint64_t* offsets, uint64_t* keys_out,
                                                       uint32_t* vals_out) {
    __shared__ uint32_t s_wave[PW];
    const int64_t begin = (int64_t)blockIdx.x * chunk, end = min(n, begin + chunk);
    uint64_t carry = offsets[blockIdx.x];
    for (int64_t base = begin; base < end; base += STEP64) {
        uint64_t x[SUB64][4];
#pragma unroll
        for (int k = 0; k < SUB64; ++k) load4_u64(keys, base + (int64_t)k * PT * VEC64 + threadIdx.x * VEC64, end, x[k]);
#pragma unroll
        for (int k = 0; k < SUB64; ++k) {
            const int64_t i = base + (int64_t)k * PT * VEC64 + threadIdx.x * VEC64;
            const uint32_t h = heads4(keys, i, end, x[k]);
            uint32_t tot;
            uint64_t pos = carry + wg_scan_excl<uint32_t>((uint32_t)__popc(h), tot, s_wave);
            uint32_t v[4] = {0u, 0u, 0u, 0u};
            if (vals && h) load4_u32(vals, i, end, v);
#pragma unroll
            for (int j = 0; j < 4; ++j)
                if ((h >> j) & 1u) {
                    keys_out[pos] = x[k][j];
                    if (vals) vals_out[pos] = v[j];
                    ++pos;
                }
            carry += tot;
        }
    }
}

inline bool aligned16(const void* p) { return ((uintptr_t)p & 15u) == 0; }

// Contiguous chunks of whole `step`-element iterations, at most 8 per CU.
void chunking(int64_t n, int64_t step, int64_t& chunk, int& G) {
    DeviceInfo info;
    const int cus = (device_info(info) == B40C_OK) ? info.num_cus : 256;
    const int64_t steps = (n + step - 1) / step;
    const int64_t g = std::max<int64_t>(1, std::min<int64_t>(steps, (int64_t)cus * 8));
    chunk = (steps + g - 1) / g * step;
    G = (int)((n + chunk - 1) / chunk);
    if (G < 1) G = 1;
}

}  // namespace

int scan_u32_max_partials() {
    DeviceInfo info;
    const int cus = (device_info(info) == B40C_OK) ? info.num_cus : 256;
    return cus * 8;
}

b40c_status scan_u32_async(const uint32_t* in, uint32_t* out, int64_t n, bool exclusive, uint64_t* partials,
                           uint64_t* total, hipStream_t stream) {
    if (n <= 0) return B40C_OK;
    const bool vec = aligned16(in) && aligned16(out);
    int64_t chunk;
    int G;
    chunking(n, vec ? STEP32 : PT, chunk, G);
    if (vec)
        hipLaunchKernelGGL(k_scan_reduce_v, dim3(G), dim3(PT), 0, stream, in, n, chunk, partials);
    else
        hipLaunchKernelGGL(k_scan_reduce, dim3(G), dim3(PT), 0, stream, in, n, chunk, partials);
    hipLaunchKernelGGL(k_spine_u64, dim3(1), dim3(PT), 0, stream, partials, G, total);
    if (vec)
        hipLaunchKernelGGL(k_scan_down_v, dim3(G), dim3(PT), 0, stream, in, out, n, chunk, partials,
                           exclusive ? 0 : 1);
    else
        hipLaunchKernelGGL(k_scan_down, dim3(G), dim3(PT), 0, stream, in, out, n, chunk, partials,
                           exclusive ? 0 : 1);
    B40C_HIP(hipGetLastError());
    return B40C_OK;
}

}  // namespace b40c

using b40c::fail;

extern "C" {

b40c_status b40c_consecutive_removal_u64(const uint64_t* d_keys_in, const uint32_t* d_values_in,
                                         uint64_t* d_keys_out, uint32_t* d_values_out, int64_t n,
                                         int64_t* num_compacted) {
    using namespace b40c;
    if (n < 0 || !num_compacted || (n > 0 && (!d_keys_in || !d_keys_out)) || (!d_values_in != !d_values_out))
        return fail(B40C_ERR_INVALID_ARG, "b40c_consecutive_removal_u64: bad arguments");
    if (n > 0 && (d_keys_in == d_keys_out || (d_values_in && d_values_in == d_values_out)))
        return fail(B40C_ERR_INVALID_ARG, "b40c_consecutive_removal_u64: output must not alias input");
    *num_compacted = 0;
    if (n == 0) return B40C_OK;
    const bool vec = aligned16(d_keys_in) && (!d_values_in || aligned16(d_values_in));
    int64_t chunk;
    int G;
    chunking(n, vec ? STEP64 : PT, chunk, G);
    DeviceBuffer counts, total;
    B40C_TRY(counts.alloc((size_t)G * 8));
    B40C_TRY(total.alloc(8));
    if (vec)
        hipLaunchKernelGGL(k_trim_count_v, dim3(G), dim3(PT), 0, nullptr, d_keys_in, n, chunk, counts.as<uint64_t>());
    else
        hipLaunchKernelGGL(k_trim_count, dim3(G), dim3(PT), 0, nullptr, d_keys_in, n, chunk, counts.as<uint64_t>());
    hipLaunchKernelGGL(k_spine_u64, dim3(1), dim3(PT), 0, nullptr, counts.as<uint64_t>(), G, total.as<uint64_t>());
    if (vec)
        hipLaunchKernelGGL(k_trim_scatter_v, dim3(G), dim3(PT), 0, nullptr, d_keys_in, d_values_in, n, chunk,
                           counts.as<uint64_t>(), d_keys_out, d_values_out);
    else
        hipLaunchKernelGGL(k_trim_scatter, dim3(G), dim3(PT), 0, nullptr, d_keys_in, d_values_in, n, chunk,
                           counts.as<uint64_t>(), d_keys_out, d_values_out);
    B40C_HIP(hipGetLastError());
    uint64_t h = 0;
    B40C_HIP(hipMemcpy(&h, total.ptr, 8, hipMemcpyDeviceToHost));
    *num_compacted = (int64_t)h;
    return B40C_OK;
}

b40c_status b40c_scan_u32(const uint32_t* d_in, uint32_t* d_out, int64_t n, int32_t exclusive, uint64_t* total) {
    using namespace b40c;
    if (n < 0 || (n > 0 && (!d_in || !d_out)))
        return fail(B40C_ERR_INVALID_ARG, "b40c_scan_u32: bad arguments");
    if (total) *total = 0;
    if (n == 0) return B40C_OK;
    DeviceBuffer sums, tot;
    B40C_TRY(sums.alloc((size_t)scan_u32_max_partials() * 8));
    B40C_TRY(tot.alloc(8));
    B40C_TRY(scan_u32_async(d_in, d_out, n, exclusive != 0, sums.as<uint64_t>(), tot.as<uint64_t>(), nullptr));
    uint64_t h = 0;
    B40C_HIP(hipMemcpy(&h, tot.ptr, 8, hipMemcpyDeviceToHost));
    if (total) *total = h;
    return B40C_OK;
}

}  // extern "C"
