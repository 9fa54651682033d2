// Device CRC helpers shared by the decode and synthesis kernels (TS 38.212
// 5.1; c_lib/nr_crc.c:129-179): MSB-first, init 0, no reflection.  A
// workgroup computes the CRC of a byte string in parallel: each lane runs the
// byte-table CRC over its slice, multiplies its remainder by
// x^(8 * bytes after the slice) mod g (GF(2) linearity), and the remainders
// are XOR-reduced.  Include after the HIP runtime header (or the CPU
// emulation header).
#pragma once
#include "ldpc_device.h"

namespace nrldpc {

// ---------------------------------------------------------------------------
// CRC (TS 38.212 5.1; c_lib/nr_crc.c:129-179): MSB-first, init 0.
// ---------------------------------------------------------------------------
struct Crc {
    uint32_t poly, mask;
    int w;
};

__device__ inline Crc crc_spec(int type) {
    switch (type) {
    case 0: return {0x1021u, 0xFFFFu, 16};
    case 1: return {0x864CFBu, 0xFFFFFFu, 24};
    case 2: return {0x800063u, 0xFFFFFFu, 24};
    default: return {0xB2B117u, 0xFFFFFFu, 24};
    }
}

__device__ inline uint32_t crc_xtimes(uint32_t r, const Crc &c) {
    const uint32_t top = (r >> (c.w - 1)) & 1u;
    r = (r << 1) & c.mask;
    return top ? r ^ c.poly : r;
}

// a(x) * b(x) mod g(x) for remainders a, b of degree < w.
__device__ inline uint32_t crc_mulmod(uint32_t a, uint32_t b, const Crc &c) {
    uint32_t r = 0;
    for (int i = c.w - 1; i >= 0; --i) {
        r = crc_xtimes(r, c);
        if ((b >> i) & 1u) r ^= a;
    }
    return r;
}

// x^(8m) mod g(x)
__device__ inline uint32_t crc_x8pow(int m, const Crc &c) {
    uint32_t res = 1u, base = 0x100u;
    while (m) {
        if (m & 1) res = crc_mulmod(res, base, c);
        base = crc_mulmod(base, base, c);
        m >>= 1;
    }
    return res;
}

__device__ inline uint32_t wave_xor(uint32_t v) {
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) v ^= __shfl_xor(v, o, 64);
    return v;
}

// Hard-decision byte from 8 consecutive beliefs (bit 7 = first belief).
__device__ inline uint32_t hard_byte(const int16_t *lq, int b) {
    const uint4 v = *reinterpret_cast<const uint4 *>(lq + 8 * b);
    const uint32_t w[4] = {v.x, v.y, v.z, v.w};
    uint32_t out = 0;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        out = (out << 1) | ((w[i] >> 15) & 1u);
        out = (out << 1) | ((w[i] >> 31) & 1u);
    }
    return out;
}

__device__ inline uint32_t hard_byte(const float *lq, int b) {
    const float4 a = *reinterpret_cast<const float4 *>(lq + 8 * b);
    const float4 c = *reinterpret_cast<const float4 *>(lq + 8 * b + 4);
    const float f[8] = {a.x, a.y, a.z, a.w, c.x, c.y, c.z, c.w};
    uint32_t out = 0;
#pragma unroll
    for (int i = 0; i < 8; ++i) out = (out << 1) | (f[i] < 0.0f ? 1u : 0u);
    return out;
}

// Shared per-workgroup CRC context.
struct CrcCtx {
    Crc spec;
    uint32_t mult;     // x^(8 * bytes after this lane's slice) mod g
    int b_lo, b_hi;    // this lane's byte slice [b_lo, b_hi)
};

__device__ inline CrcCtx crc_setup(int crctype, int nbytes, uint32_t *tab) {
    CrcCtx ctx;
    ctx.spec = crc_spec(crctype);
    const int t = threadIdx.x, nt = blockDim.x;
    for (int i = t; i < 256; i += nt) {
        uint32_t r = (uint32_t)i << (ctx.spec.w - 8);
        for (int k = 0; k < 8; ++k) r = crc_xtimes(r, ctx.spec);
        tab[i] = r;
    }
    const int per = (nbytes + nt - 1) / nt;
    ctx.b_lo = t * per < nbytes ? t * per : nbytes;
    ctx.b_hi = (t + 1) * per < nbytes ? (t + 1) * per : nbytes;
    ctx.mult = (ctx.b_hi > ctx.b_lo) ? crc_x8pow(nbytes - ctx.b_hi, ctx.spec) : 0u;
    return ctx;
}

// Whole-codeblock CRC of the hard decisions; every lane gets the result.
// Contains one workgroup barrier.  Caller guarantees lq is settled.
template <typename T>
__device__ inline uint32_t block_crc(const T *lq, const CrcCtx &ctx,
                                     const uint32_t *tab, uint32_t *slots) {
    const Crc &c = ctx.spec;
    uint32_t r = 0;
    for (int b = ctx.b_lo; b < ctx.b_hi; ++b) {
        const uint32_t byte = hard_byte(lq, b);
        r = ((r << 8) ^ tab[((r >> (c.w - 8)) ^ byte) & 0xFFu]) & c.mask;
    }
    uint32_t part = r ? crc_mulmod(r, ctx.mult, c) : 0u;
    part = wave_xor(part);
    const int wave = threadIdx.x >> 6, nwaves = (blockDim.x + 63) >> 6;
    if ((threadIdx.x & 63) == 0) slots[wave] = part;
    __syncthreads();
    uint32_t total = 0;
    for (int w = 0; w < nwaves; ++w) total ^= slots[w];
    return total;
}

}  // namespace nrldpc

namespace nrldpc {
// crc_setup with the tables precomputed on the host (DecOut::crc): the byte
// table is copied to tab (LDS) and the lane's multiplier is one load.
__device__ inline CrcCtx crc_setup_tab(int crctype, int nbytes, uint32_t *tab,
                                       const uint32_t *crc_tables) {
    CrcCtx ctx;
    ctx.spec = crc_spec(crctype);
    const uint32_t *g = crc_tables + (size_t)crctype * kCrcTabWords;
    const int t = threadIdx.x, nt = blockDim.x;
    for (int i = t; i < 256; i += nt) tab[i] = g[i];
    const int per = (nbytes + nt - 1) / nt;
    ctx.b_lo = t * per < nbytes ? t * per : nbytes;
    ctx.b_hi = (t + 1) * per < nbytes ? (t + 1) * per : nbytes;
    ctx.mult = (ctx.b_hi > ctx.b_lo) ? g[256 + nbytes - ctx.b_hi] : 0u;
    return ctx;
}
}  // namespace nrldpc
