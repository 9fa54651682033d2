// CPU emulation of the small HIP surface the decode kernels use, so their
// source can be compiled with g++ and run in CPU tests.  One workgroup runs at
// a time; each of its lanes is a host thread; __syncthreads is a real
// barrier; __shfl_xor exchanges through a shared array between two barriers
// (valid because the kernels call it from uniform control flow).  Test
// infrastructure only.
#pragma once
#include <math.h>
#include <stdint.h>

#include <atomic>
#include <barrier>
#include <cstring>
#include <functional>
#include <memory>
#include <thread>
#include <vector>

#define __global__
#define __device__
#define __host__
#define __forceinline__ inline
#define __launch_bounds__(...)
#define __align__(n) alignas(n)
#define __shared__ static

struct EmuDim3 {
    unsigned x = 0, y = 0, z = 0;
};
struct uint4 {
    uint32_t x, y, z, w;
};
struct uint2 {
    uint32_t x, y;
};
inline uint2 make_uint2(uint32_t x, uint32_t y) { return uint2{x, y}; }
struct float4 {
    float x, y, z, w;
};

inline thread_local EmuDim3 threadIdx;
inline EmuDim3 blockIdx, blockDim, gridDim;

struct EmuBlock {
    std::unique_ptr<std::barrier<>> bar;
    std::vector<unsigned char> lds;
    std::vector<uint32_t> shfl;
};
inline EmuBlock g_emu_block;

// HIP's device min/max overloads for the integer types the kernels use.
inline int min(int a, int b) { return a < b ? a : b; }
inline int max(int a, int b) { return a > b ? a : b; }
inline uint32_t min(uint32_t a, uint32_t b) { return a < b ? a : b; }
inline uint32_t max(uint32_t a, uint32_t b) { return a > b ? a : b; }

inline unsigned char *emu_dynamic_lds() { return g_emu_block.lds.data(); }
inline void __syncthreads() { g_emu_block.bar->arrive_and_wait(); }
inline int __popc(uint32_t v) { return __builtin_popcount(v); }
inline uint32_t __umulhi(uint32_t a, uint32_t b) { return (uint32_t)(((uint64_t)a * b) >> 32); }

inline int __syncthreads_or(int pred) {
    static std::atomic<int> acc[2];
    static thread_local int phase = 0;
    const int ph = phase;
    phase ^= 1;
    if (pred) acc[ph].fetch_or(1);
    __syncthreads();
    const int r = acc[ph].load();
    __syncthreads();
    if (threadIdx.x == 0) acc[ph].store(0);   // cleared before the next use of this phase
    return r;
}

inline uint32_t atomicXor(uint32_t *addr, uint32_t v) {
    return __atomic_fetch_xor(addr, v, __ATOMIC_SEQ_CST);
}
inline uint32_t atomicAdd(uint32_t *addr, uint32_t v) {
    return __atomic_fetch_add(addr, v, __ATOMIC_SEQ_CST);
}
inline uint32_t atomicExch(uint32_t *addr, uint32_t v) {
    return __atomic_exchange_n(addr, v, __ATOMIC_SEQ_CST);
}
inline void __threadfence() { __atomic_thread_fence(__ATOMIC_SEQ_CST); }

inline uint32_t __shfl_xor(uint32_t v, int mask, int width) {
    (void)width;
    const unsigned t = threadIdx.x;
    g_emu_block.shfl[t] = v;
    __syncthreads();
    const uint32_t r = g_emu_block.shfl[t ^ (unsigned)mask];
    __syncthreads();
    return r;
}

inline float __shfl_xor(float v, int mask, int width) {
    uint32_t u;
    std::memcpy(&u, &v, sizeof u);
    u = __shfl_xor(u, mask, width);
    float r;
    std::memcpy(&r, &u, sizeof r);
    return r;
}

// Runs fn once per lane of each workgroup in [0, grid).
inline void emu_launch(unsigned grid, unsigned block, size_t lds_bytes,
                       const std::function<void()> &fn) {
    blockDim.x = block;
    blockDim.y = blockDim.z = 1;
    gridDim.x = grid;
    gridDim.y = gridDim.z = 1;
    for (unsigned b = 0; b < grid; ++b) {
        blockIdx.x = b;
        g_emu_block.bar = std::make_unique<std::barrier<>>((std::ptrdiff_t)block);
        g_emu_block.lds.assign(lds_bytes + 64, 0xCD);   // poison: reads of unset LDS show up
        g_emu_block.shfl.assign(block, 0);
        std::vector<std::thread> lanes;
        lanes.reserve(block);
        for (unsigned t = 0; t < block; ++t)
            lanes.emplace_back([t, &fn] {
                threadIdx.x = t;
                fn();
            });
        for (auto &l : lanes) l.join();
    }
}
