// CPU replay of the library's device code (test infrastructure only).
//
// Stands in for <hip/hip_runtime.h> when the BFS kernel bodies
// (back40computing_amd/csrc/bfs_kernels.hpp) are compiled with g++ for the
// host.  Every lane of a launch is one host thread; the wave (64 lanes) and
// the workgroup each have a barrier, and the wave collectives (ballot,
// shuffles) exchange values through it.  That reproduces the kernels'
// observable behaviour as long as every lane of a wave reaches each
// collective the same number of times (true of every body the replay runs:
// the level steps, the persistent loop and the grid barrier), while
// AddressSanitizer checks every global access against the exact size of the
// buffer it belongs to.  A barrier that waits longer than two minutes
// reports which one and aborts the replay instead of hanging.
#pragma once

#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <mutex>
#include <thread>
#include <type_traits>
#include <vector>

#include <pthread.h>

#define __device__
#define __host__
#define __global__
#define __forceinline__ inline
#define __launch_bounds__(...)

typedef int hipError_t;
typedef void* hipStream_t;
#define hipSuccess 0
#define hipErrorOutOfMemory 2
inline const char* hipGetErrorString(hipError_t) { return "replay"; }

struct int2 {
    int x, y;
};
inline int2 make_int2(int x, int y) { return int2{x, y}; }

template <class A, class B>
inline typename std::common_type<A, B>::type min(A a, B b) {
    return a < b ? a : b;
}
template <class A, class B>
inline typename std::common_type<A, B>::type max(A a, B b) {
    return a > b ? a : b;
}

namespace emu {

// Reusable barrier with a timeout: a hang in the replay is a broken
// assumption about the kernels (a collective reached by only part of a
// wave, a grid barrier that never completes), reported rather than waited on.
// Waiters spin briefly on the generation word, then sleep on a condition
// variable (the last arrival takes the mutex before notifying, so no wake-up
// is lost).
class Barrier {
  public:
    explicit Barrier(int n) : n_(n) {}
    void wait(const char* what) {
        const uint64_t g = gen_.load(std::memory_order_acquire);
        if (count_.fetch_add(1, std::memory_order_acq_rel) + 1 == n_) {
            count_.store(0, std::memory_order_relaxed);
            {
                std::lock_guard<std::mutex> lk(m_);
                gen_.store(g + 1, std::memory_order_release);
            }
            cv_.notify_all();
            return;
        }
        for (int i = 0; i < 2000; ++i) {
            if (gen_.load(std::memory_order_acquire) != g) return;
            if ((i & 63) == 63) std::this_thread::yield();
        }
        std::unique_lock<std::mutex> lk(m_);
        if (!cv_.wait_for(lk, std::chrono::seconds(120), [&] { return gen_.load(std::memory_order_acquire) != g; })) {
            std::fprintf(stderr, "replay: %s barrier timed out (%d of %d arrived)\n", what, count_.load(), n_);
            std::abort();
        }
    }

  private:
    std::mutex m_;
    std::condition_variable cv_;
    const int n_;
    std::atomic<int> count_{0};
    std::atomic<uint64_t> gen_{0};
};

struct Wave {
    Barrier bar{64};
    uint64_t x[2][64];
};

struct Block {
    explicit Block(int threads) : bar(threads) {}
    Barrier bar;
    std::atomic<int> acc[3];
};

struct Dim {
    uint32_t x = 0, y = 0, z = 0;
};

struct Lane {
    Dim tid, bid, bdim, gdim;
    int lane = 0;
    Wave* wave = nullptr;
    Block* block = nullptr;
    uint64_t wave_ops = 0;
    uint64_t block_ops = 0;
};

inline thread_local Lane self;

// Wave collective: every lane publishes a 64-bit value, then reads all.
inline const uint64_t* wave_exchange(uint64_t v) {
    Wave& w = *self.wave;
    uint64_t* buf = w.x[self.wave_ops++ & 1];
    buf[self.lane] = v;
    w.bar.wait("wave");
    return buf;
}

// Workgroup barrier that also ORs one int over the workgroup.  Three
// accumulators rotate: op k uses acc[k % 3], and thread 0 clears
// acc[(k + 2) % 3] after op k's barrier (its previous readers, op k - 1,
// finished before that barrier; its next writers, op k + 2, start after the
// next one).
inline int block_op(int v) {
    Block& b = *self.block;
    const uint64_t k = self.block_ops++;
    if (v) b.acc[k % 3].fetch_or(1);
    b.bar.wait("workgroup");
    const int r = b.acc[k % 3].load();
    if (self.tid.x == 0) b.acc[(k + 2) % 3].store(0);
    return r;
}

template <class T>
inline uint64_t bits_of(T v) {
    static_assert(sizeof(T) <= 8, "shuffle of a wide type");
    uint64_t u = 0;
    std::memcpy(&u, &v, sizeof(T));
    return u;
}
template <class T>
inline T from_bits(uint64_t u) {
    T v;
    std::memcpy(&v, &u, sizeof(T));
    return v;
}

// Runs body() on grid x threads host threads (threads must be a multiple of
// 64), one per lane.  Returns after every lane has finished.
inline void launch(unsigned grid, unsigned threads, const std::function<void()>& body) {
    if (threads % 64 != 0) {
        std::fprintf(stderr, "replay: %u threads is not a whole number of waves\n", threads);
        std::abort();
    }
    std::vector<Block*> blocks;
    std::vector<Wave*> waves;
    for (unsigned b = 0; b < grid; ++b) {
        blocks.push_back(new Block((int)threads));
        for (int i = 0; i < 3; ++i) blocks.back()->acc[i].store(0);
        for (unsigned w = 0; w < threads / 64; ++w) waves.push_back(new Wave());
    }
    struct Arg {
        Lane init;
        const std::function<void()>* body;
    };
    std::vector<Arg> args(grid * threads);
    std::vector<pthread_t> ths(grid * threads);
    pthread_attr_t attr;
    pthread_attr_init(&attr);
    pthread_attr_setstacksize(&attr, 1 << 20);
    for (unsigned b = 0; b < grid; ++b)
        for (unsigned t = 0; t < threads; ++t) {
            Arg& a = args[b * threads + t];
            a.init.tid.x = t;
            a.init.bid.x = b;
            a.init.bdim.x = threads;
            a.init.gdim.x = grid;
            a.init.lane = (int)(t % 64);
            a.init.wave = waves[b * (threads / 64) + t / 64];
            a.init.block = blocks[b];
            a.body = &body;
            auto fn = [](void* p) -> void* {
                Arg* a = static_cast<Arg*>(p);
                self = a->init;
                (*a->body)();
                return nullptr;
            };
            if (pthread_create(&ths[b * threads + t], &attr, fn, &a) != 0) {
                std::fprintf(stderr, "replay: pthread_create failed\n");
                std::abort();
            }
        }
    pthread_attr_destroy(&attr);
    for (auto& th : ths) pthread_join(th, nullptr);
    for (auto* b : blocks) delete b;
    for (auto* w : waves) delete w;
}

template <class T>
inline T fetch_max(T* p, T v) {
    T cur = __atomic_load_n(p, __ATOMIC_SEQ_CST);
    while (cur < v && !__atomic_compare_exchange_n(p, &cur, v, false, __ATOMIC_SEQ_CST, __ATOMIC_SEQ_CST)) {
    }
    return cur;
}
template <class T>
inline T fetch_min(T* p, T v) {
    T cur = __atomic_load_n(p, __ATOMIC_SEQ_CST);
    while (cur > v && !__atomic_compare_exchange_n(p, &cur, v, false, __ATOMIC_SEQ_CST, __ATOMIC_SEQ_CST)) {
    }
    return cur;
}

}  // namespace emu

#define threadIdx (emu::self.tid)
#define blockIdx (emu::self.bid)
#define blockDim (emu::self.bdim)
#define gridDim (emu::self.gdim)

inline int __lane_id() { return emu::self.lane; }
inline void __syncthreads() { emu::block_op(0); }
inline int __syncthreads_or(int v) { return emu::block_op(v); }

inline unsigned long long __ballot(int pred) {
    const uint64_t* x = emu::wave_exchange(pred ? 1u : 0u);
    unsigned long long m = 0;
    for (int l = 0; l < 64; ++l)
        if (x[l]) m |= 1ull << l;
    return m;
}
template <class T>
inline T __shfl(T v, int src, int width = 64) {
    (void)width;
    const uint64_t* x = emu::wave_exchange(emu::bits_of(v));
    return emu::from_bits<T>(x[src & 63]);
}
template <class T>
inline T __shfl_up(T v, unsigned delta, int width = 64) {
    (void)width;
    const uint64_t* x = emu::wave_exchange(emu::bits_of(v));
    const int l = emu::self.lane;
    return l >= (int)delta ? emu::from_bits<T>(x[l - (int)delta]) : v;
}
template <class T>
inline T __shfl_xor(T v, int mask, int width = 64) {
    (void)width;
    const uint64_t* x = emu::wave_exchange(emu::bits_of(v));
    return emu::from_bits<T>(x[(emu::self.lane ^ mask) & 63]);
}
inline void __builtin_amdgcn_wave_barrier() { emu::wave_exchange(0); }

inline int __popcll(unsigned long long x) { return __builtin_popcountll(x); }
inline int __ffsll(unsigned long long x) { return __builtin_ffsll((long long)x); }

inline uint32_t __builtin_amdgcn_mbcnt_lo(uint32_t mask, uint32_t add) {
    const int l = emu::self.lane;
    const uint32_t below = l >= 32 ? 0xFFFFFFFFu : ((1u << l) - 1u);
    return add + (uint32_t)__builtin_popcount(mask & below);
}
inline uint32_t __builtin_amdgcn_mbcnt_hi(uint32_t mask, uint32_t add) {
    const int l = emu::self.lane - 32;
    const uint32_t below = l <= 0 ? 0u : ((1u << l) - 1u);
    return add + (uint32_t)__builtin_popcount(mask & below);
}

#define __HIP_MEMORY_SCOPE_SINGLETHREAD 1
#define __HIP_MEMORY_SCOPE_WAVEFRONT 2
#define __HIP_MEMORY_SCOPE_WORKGROUP 3
#define __HIP_MEMORY_SCOPE_AGENT 4
#define __HIP_MEMORY_SCOPE_SYSTEM 5

// Device atomics of any scope become sequentially consistent host atomics.
#define __hip_atomic_load(p, order, scope) __atomic_load_n((p), __ATOMIC_SEQ_CST)
#define __hip_atomic_store(p, v, order, scope) __atomic_store_n((p), (v), __ATOMIC_SEQ_CST)
#define __hip_atomic_fetch_add(p, v, order, scope) __atomic_fetch_add((p), (v), __ATOMIC_SEQ_CST)
#define __hip_atomic_fetch_or(p, v, order, scope) __atomic_fetch_or((p), (v), __ATOMIC_SEQ_CST)
#define __hip_atomic_fetch_max(p, v, order, scope) emu::fetch_max((p), (v))
#define __hip_atomic_fetch_min(p, v, order, scope) emu::fetch_min((p), (v))
#define __builtin_amdgcn_fence(order, scope) __atomic_thread_fence(__ATOMIC_SEQ_CST)
// A wave's wait for its outstanding memory operations: host memory accesses
// have completed when they return, so a full fence stands in for it.
inline void __builtin_amdgcn_s_waitcnt(int) { __atomic_thread_fence(__ATOMIC_SEQ_CST); }

template <class T>
inline T atomicAdd(T* p, T v) {
    return __atomic_fetch_add(p, v, __ATOMIC_SEQ_CST);
}

// A wave's s_sleep: give the host thread away (the grid barrier polls).
inline void __builtin_amdgcn_s_sleep(int) { std::this_thread::sleep_for(std::chrono::microseconds(20)); }

inline unsigned long long wall_clock64() {
    return (unsigned long long)std::chrono::steady_clock::now().time_since_epoch().count();
}
