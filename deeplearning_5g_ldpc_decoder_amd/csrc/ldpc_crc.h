// Device CRC helpers shared by the decode and synthesis kernels (TS 38.212
// 5.1; c_lib/nr_crc.c:129-179): MSB-first, init 0, no reflection.  A
// workgroup computes the CRC of a byte string in parallel: each lane runs the
// byte-table CRC over its slice, multiplies its remainder by
// x^(8 * bytes after the slice) mod g (GF(2) linearity), and the remainders
// are XOR-reduced.  Include after the HIP runtime header (or the CPU
// emulation header).
#pragma once
#include "ldpc_device.h"
#include "ldpc_graph.h"

namespace nrldpc {

// ---------------------------------------------------------------------------
// CRC (TS 38.212 5.1; c_lib/nr_crc.c:129-179): MSB-first, init 0.
// ---------------------------------------------------------------------------
struct Crc {
    uint32_t poly, mask;
    int w;
};

__device__ __forceinline__ Crc crc_spec(int type) {
    switch (type) {
    case 0: return {0x1021u, 0xFFFFu, 16};
    case 1: return {0x864CFBu, 0xFFFFFFu, 24};
    case 2: return {0x800063u, 0xFFFFFFu, 24};
    default: return {0xB2B117u, 0xFFFFFFu, 24};
    }
}

__device__ __forceinline__ uint32_t crc_xtimes(uint32_t r, const Crc &c) {
    const uint32_t top = (r >> (c.w - 1)) & 1u;
    r = (r << 1) & c.mask;
    return top ? r ^ c.poly : r;
}

// a(x) * b(x) mod g(x) for remainders a, b of degree < w.
__device__ __forceinline__ uint32_t crc_mulmod(uint32_t a, uint32_t b, const Crc &c) {
    uint32_t r = 0;
    for (int i = c.w - 1; i >= 0; --i) {
        r = crc_xtimes(r, c);
        if ((b >> i) & 1u) r ^= a;
    }
    return r;
}

// crc_mulmod unrolled for a width W known at compile time, 5 ops per bit of
// b (Horner from its top bit): r * x is r << 1 with g folded in when bit
// W - 1 of r was set (g | x^W also clears the bit shifted out), and a is
// added under bit i of b, both selected by sign-extended bit masks.
template <int W>
__device__ __forceinline__ uint32_t crc_mulmod_w(uint32_t a, uint32_t b, uint32_t poly) {
    const uint32_t g = poly | (1u << W);
    uint32_t r = 0;
#pragma unroll
    for (int i = W - 1; i >= 0; --i) {
        const uint32_t top = (uint32_t)((int32_t)(r << (31 - (W - 1))) >> 31);
        r = (r << 1) ^ (g & top);
        r ^= a & (uint32_t)((int32_t)(b << (31 - i)) >> 31);
    }
    return r;
}

__device__ __forceinline__ uint32_t crc_mulmod_fast(uint32_t a, uint32_t b, const Crc &c) {
    return c.w == 24 ? crc_mulmod_w<24>(a, b, c.poly) : crc_mulmod_w<16>(a, b, c.poly);
}

// x^(8m) mod g(x)
__device__ __forceinline__ uint32_t crc_x8pow(int m, const Crc &c) {
    uint32_t res = 1u, base = 0x100u;
    while (m) {
        if (m & 1) res = crc_mulmod(res, base, c);
        base = crc_mulmod(base, base, c);
        m >>= 1;
    }
    return res;
}

__device__ __forceinline__ uint32_t wave_xor(uint32_t v) {
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) v ^= __shfl_xor(v, o, 64);
    return v;
}

// XOR of v over the wave, in every lane.  On the GPU: DPP row shifts give
// each row of 16 lanes its XOR in its last lane, two row broadcasts carry
// rows 0-2 into lane 63, and one readlane returns it (the wave must be
// whole: DPP reads inactive lanes' registers too).
__device__ __forceinline__ uint32_t wave_xor_full(uint32_t v) {
#ifdef NRLDPC_CPU_EMU
    return wave_xor(v);
#else
    int x = (int)v;
    x ^= __builtin_amdgcn_update_dpp(0, x, 0x111, 0xF, 0xF, true);   // row_shr:1
    x ^= __builtin_amdgcn_update_dpp(0, x, 0x112, 0xF, 0xF, true);   // row_shr:2
    x ^= __builtin_amdgcn_update_dpp(0, x, 0x114, 0xF, 0xF, true);   // row_shr:4
    x ^= __builtin_amdgcn_update_dpp(0, x, 0x118, 0xF, 0xF, true);   // row_shr:8
    x ^= __builtin_amdgcn_update_dpp(0, x, 0x142, 0xA, 0xF, true);   // row_bcast:15
    x ^= __builtin_amdgcn_update_dpp(0, x, 0x143, 0xC, 0xF, true);   // row_bcast:31
    return (uint32_t)__builtin_amdgcn_readlane(x, 63);
#endif
}

// Hard-decision byte from 8 consecutive beliefs (bit 7 = first belief).
__device__ __forceinline__ uint32_t hard_byte(const int16_t *lq, int b) {
    const uint4 v = *reinterpret_cast<const uint4 *>(lq + 8 * b);
#ifdef NRLDPC_CPU_EMU
    const uint32_t w[4] = {v.x, v.y, v.z, v.w};
    uint32_t out = 0;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        out = (out << 1) | ((w[i] >> 15) & 1u);
        out = (out << 1) | ((w[i] >> 31) & 1u);
    }
    return out;
#else
    // v_perm_b32 selectors 8-11 give 0x00 / 0xFF bytes from the sign bits of
    // the low and high halves of its two words (beliefs 0-3, then 4-7); one
    // bit per byte survives the mask, and v_sad_u8 against 0 adds the bytes,
    // which ORs the distinct bits into the hard byte.
    const uint32_t s01 = __builtin_amdgcn_perm(v.y, v.x, 0x0B0A0908u) & 0x10204080u;
    const uint32_t s23 = __builtin_amdgcn_perm(v.w, v.z, 0x0B0A0908u) & 0x01020408u;
    return __builtin_amdgcn_sad_u8(s01, 0u, __builtin_amdgcn_sad_u8(s23, 0u, 0u));
#endif
}

__device__ __forceinline__ uint32_t hard_byte(const float *lq, int b) {
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

__device__ __forceinline__ CrcCtx crc_setup(int crctype, int nbytes, uint32_t *tab) {
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
__device__ __forceinline__ uint32_t block_crc(const T *lq, const CrcCtx &ctx,
                                     const uint32_t *tab, uint32_t *slots) {
    const Crc &c = ctx.spec;
    uint32_t r = 0;
    // up to 4 bytes' beliefs loaded together, then the dependent table steps
    for (int b0 = ctx.b_lo; b0 < ctx.b_hi; b0 += 4) {
        uint32_t bytes[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) bytes[u] = b0 + u < ctx.b_hi ? hard_byte(lq, b0 + u) : 0u;
#pragma unroll
        for (int u = 0; u < 4; ++u)
            if (b0 + u < ctx.b_hi)
                r = ((r << 8) ^ tab[((r >> (c.w - 8)) ^ bytes[u]) & 0xFFu]) & c.mask;
    }
    uint32_t part = crc_mulmod_fast(r, ctx.mult, c);
    part = wave_xor_full(part);   // block sizes are whole waves
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
__device__ __forceinline__ CrcCtx crc_setup_tab(int crctype, int nbytes, uint32_t *tab,
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

namespace nrldpc {
// CRC zero test of the hard decisions, linear form.  Bit i of the B-bit block
// contributes x^(B - 1 - i) mod g to the block's remainder (up to the factor
// x^w, which is invertible mod g), so the remainder is the XOR of the weights
// of the set bits.  A lane holds bit i = c * Zc + zx of each systematic column
// c and the weights of those bits (0 past the block and for lanes without a
// codeblock) in LDS, column-major with a stride of one workgroup
// (crc_lin_weights); per column it extracts the hard bit as a 0 / -1 mask and
// adds w & mask with one v_bitop3 (truth-table byte 0x78: src0 ^ (src1 &
// src2) with sources 0xF0 / 0xCC / 0xAA).  Against a byte-serial table CRC and
// a 24-step remainder multiply per lane, the test is two or three VALU ops
// per column.
// KB columns: the loads are unconditional (an index clamped into the table)
// so that all of them are in flight together (crc_lin_load); the stores
// follow once they have arrived (crc_lin_store).
template <int KB>
struct CrcLinRegs {
    uint32_t w[KB];
};
template <int KB>
__device__ __forceinline__ void crc_lin_load(CrcLinRegs<KB> &r, const DecOut &o, const CbDesc &d,
                                             int zx, int Zc, bool member) {
    const uint32_t *xb = o.crc + (size_t)d.crctype * kCrcTabWords + kCrcBitOff;
    const int nb = (int)d.B;
#pragma unroll
    for (int c = 0; c < KB; ++c) {
        const int m = nb - 1 - (c * Zc + zx);   // bits after this one
        r.w[c] = xb[m < 0 ? 0 : m];
        if (!member || m < 0) r.w[c] = 0u;
    }
}
template <int KB>
__device__ __forceinline__ void crc_lin_store(const CrcLinRegs<KB> &r, uint32_t *wts, int stride) {
#pragma unroll
    for (int c = 0; c < KB; ++c) wts[c * stride] = r.w[c];
}
template <int KB>
__device__ __forceinline__ void crc_lin_weights(uint32_t *wts, int stride, const DecOut &o,
                                                const CbDesc &d, int zx, int Zc, bool member) {
    CrcLinRegs<KB> r;
    crc_lin_load<KB>(r, o, d, zx, Zc, member);
    crc_lin_store<KB>(r, wts, stride);
}

// Hard bit of a belief as 0 / -1: the sign bit of an int16, v < 0 for a float
// (-0.0 is a 0 bit, as in hard_byte).
__device__ __forceinline__ uint32_t crc_hard_mask(int16_t v) {
#ifdef NRLDPC_CPU_EMU
    return (uint32_t)(-(int)(((uint32_t)(uint16_t)v >> 15) & 1u));
#else
    return (uint32_t)__builtin_amdgcn_sbfe((int)(uint16_t)v, 15u, 1u);
#endif
}
__device__ __forceinline__ uint32_t crc_hard_mask(float v) { return v < 0.0f ? ~0u : 0u; }

__device__ __forceinline__ uint32_t crc_lin_add(uint32_t acc, uint32_t w, uint32_t m) {
#ifdef NRLDPC_CPU_EMU
    return acc ^ (w & m);
#else
    return __builtin_amdgcn_bitop3_b32(acc, w, m, 0x78u);   // acc ^ (w & m)
#endif
}

// The one-codeblock-per-workgroup kernels' test: beliefs lq (generic
// pointer, lifting size Zc at run time), weights wts[c * blockDim.x + zx];
// lanes zx >= Zc read another column's belief under a zero weight (still
// inside the N + 2Zc beliefs for every Zc >= 2).
template <int BG, typename T>
__device__ __forceinline__ uint32_t block_crc_linear(const T *lq, int Zc, const uint32_t *wts,
                                                     uint32_t *slots) {
    const int zx = threadIdx.x, nt = blockDim.x;
    uint32_t acc = 0u;
#pragma unroll
    for (int c = 0; c < Graph<BG>::kInfoCols; ++c)
        acc = crc_lin_add(acc, wts[c * nt + zx], crc_hard_mask(lq[c * Zc + zx]));
    acc = wave_xor_full(acc);   // block sizes are whole waves
    if ((zx & 63) == 0) slots[zx >> 6] = acc;
    __syncthreads();
    uint32_t total = 0u;
    for (int w = 0; w < (nt >> 6); ++w) total ^= slots[w];
    return total;
}

}  // namespace nrldpc
