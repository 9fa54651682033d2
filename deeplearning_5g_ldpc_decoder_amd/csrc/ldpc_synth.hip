// Device synthesis of decoder inputs (gfx950): one workgroup per codeblock
// generates a payload, attaches the CRC, LDPC-encodes it, rate-matches it and
// passes it through the AWGN test channel, so simulations and training get
// their inputs without a host round trip.
//
//   payload   A bits from a counter-based generator (or the caller's bits)
//   CRC       TS 38.212 5.1, the parallel byte-table CRC of ldpc_crc.h
//             (c_lib/nr_crc.c:129-179)
//   encode    TS 38.212 5.3.2 with the base graphs' double-diagonal core
//             (c_lib/ldpc_encoder.c:221-242): lambda_r = info part of core rows
//             0..3; p0 from their sum, p1..p3 by back substitution; each
//             extension row's parity is the XOR of its info and core columns
//   rate match TS 38.212 5.4.2 (c_lib/ldpc_rate_match.c:27-106): bit k of
//             the selection is the (a0 + k) mod Navail-th non-filler position
//             of the circular buffer (a0 = rank of k0), interleaved in Qm rows
//   channel   Gray QAM (TS 38.211 5.1) + complex AWGN, max-log LLR per bit
//             (c_lib/tx_rx_process.c:68-324) in the reference's piecewise
//             linear form and float operation order (kDemapSegs), so given
//             the same noise the LLRs are the reference's bit for bit; int16
//             output quantized x128, rounded half away from zero, clipped to
//             +-10000 (:27-39).  The noise is the counter-based generator's,
//             or the caller's samples (SynthArgs::noise)
//   beliefs   (optional) the float rate recovery of the trainable decoder's
//             inputs (py5gphy/ldpc/nr_ldpc_raterecover.py:15-68): repeats
//             summed, filler positions 10 * max|LLR|, 2Zc leading zeros
//
// Random numbers: splitmix64 of (seed, codeblock, counter) -- uniform bits for
// the payload, Box-Muller normals for the noise; each value depends only on
// its counter, so results do not depend on the launch shape.
#ifdef NRLDPC_CPU_EMU
#include "hip_cpu_emu.h"
#else
#include <hip/hip_runtime.h>
#endif

#include "ldpc_crc.h"
#include "ldpc_device.h"
#include "ldpc_graph.h"

namespace nrldpc {

constexpr int kSynthThreads = 256;

__device__ __forceinline__ uint64_t mix64(uint64_t z) {
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

// Counter-based 64 random bits for (seed, codeblock, stream, counter).
__device__ __forceinline__ uint64_t rnd64(uint64_t seed, int cb, int stream, uint32_t ctr) {
    return mix64(seed ^ mix64(((uint64_t)(uint32_t)cb << 32) ^ ((uint64_t)stream << 28) ^ ctr));
}

__device__ __forceinline__ float unit_open(uint64_t r) {
    return ((float)(uint32_t)(r >> 40) + 0.5f) * (1.0f / 16777216.0f);   // (0, 1)
}

// Gray PAM level of one axis from nb bits, MSB first (TS 38.211 5.1).
__device__ __forceinline__ float pam_level(const int *b, int nb) {
    if (nb == 1) return (float)(1 - 2 * b[0]);
    if (nb == 2) return (float)((1 - 2 * b[0]) * (2 - (1 - 2 * b[1])));
    if (nb == 3) return (float)((1 - 2 * b[0]) * (4 - (1 - 2 * b[1]) * (2 - (1 - 2 * b[2]))));
    return (float)((1 - 2 * b[0]) *
                   (8 - (1 - 2 * b[1]) * (4 - (1 - 2 * b[2]) * (2 - (1 - 2 * b[3])))));
}

// Max-log LLR of bit k of one Gray-PAM axis with 2^nb levels as a function of
// the received coordinate r, as the reference evaluates it
// (c_lib/tx_rx_process.c:101-232): the first segment with r < T*A gives
//     sgn * ((C*A) * (r + O*A)) / noise_var      (r alone when O = 0),
// each product and sum rounded as written there.  Rows: nb = 2, 3, 4 and
// k = 0..nb-1; a row ends at the segment with T = kDemapOpen.
struct DemapSeg {
    int8_t T, C, sgn, O;
};
constexpr int8_t kDemapOpen = 127;
constexpr DemapSeg kDemapSegs[] = {
    // nb = 2 (16QAM), k = 0, 1
    {-2, 8, 1, 1}, {2, 4, 1, 0}, {kDemapOpen, 8, 1, -1},
    {0, 4, 1, 2}, {kDemapOpen, 4, -1, -2},
    // nb = 3 (64QAM), k = 0, 1, 2
    {-6, 16, 1, 3}, {-4, 12, 1, 2}, {-2, 8, 1, 1}, {2, 4, 1, 0}, {4, 8, 1, -1}, {6, 12, 1, -2},
    {kDemapOpen, 16, 1, -3},
    {-6, 8, 1, 5}, {-2, 4, 1, 4}, {0, 8, 1, 3}, {2, 8, -1, -3}, {6, 4, -1, -4},
    {kDemapOpen, 8, -1, -5},
    {-4, 4, 1, 6}, {0, 4, -1, 2}, {4, 4, 1, -2}, {kDemapOpen, 4, -1, -6},
    // nb = 4 (256QAM), k = 0 .. 3
    {-14, 32, 1, 7}, {-12, 28, 1, 6}, {-10, 24, 1, 5}, {-8, 20, 1, 4}, {-6, 16, 1, 3},
    {-4, 12, 1, 2}, {-2, 8, 1, 1}, {2, 4, 1, 0}, {4, 8, 1, -1}, {6, 12, 1, -2}, {8, 16, 1, -3},
    {10, 20, 1, -4}, {12, 24, 1, -5}, {14, 28, 1, -6}, {kDemapOpen, 32, 1, -7},
    {-14, 16, 1, 11}, {-12, 12, 1, 10}, {-10, 8, 1, 9}, {-6, 4, 1, 8}, {-4, 8, 1, 7},
    {-2, 12, 1, 6}, {0, 16, 1, 5}, {2, 16, -1, -5}, {4, 12, -1, -6}, {6, 8, -1, -7},
    {10, 4, -1, -8}, {12, 8, -1, -9}, {14, 12, -1, -10}, {kDemapOpen, 16, -1, -11},
    {-14, 8, 1, 13}, {-10, 4, 1, 12}, {-8, 8, 1, 11}, {-6, 8, -1, 5}, {-2, 4, -1, 4},
    {0, 8, -1, 3}, {2, 8, 1, -3}, {6, 4, 1, -4}, {8, 8, 1, -5}, {10, 8, -1, -11},
    {14, 4, -1, -12}, {kDemapOpen, 8, -1, -13},
    {-12, 4, 1, 14}, {-8, 4, -1, 10}, {-4, 4, 1, 6}, {0, 4, -1, 2}, {4, 4, 1, -2},
    {8, 4, -1, -6}, {12, 4, 1, -10}, {kDemapOpen, 4, -1, -14},
};

// First segment of each row: rows (2, 0), (2, 1), (3, 0) .. (3, 2), (4, 0) .. (4, 3).
struct DemapRows {
    int first[9];
    constexpr DemapRows() : first() {
        int row = 0;
        first[0] = 0;
        for (int i = 0; row < 8; ++i)
            if (kDemapSegs[i].T == kDemapOpen) first[++row] = i + 1;
    }
};
inline constexpr DemapRows kDemapRows{};

__device__ __forceinline__ float demap_pam(float r, int nb, int k, float A, float nv) {
    int i = kDemapRows.first[(nb == 2 ? 0 : (nb == 3 ? 2 : 5)) + k];
    while (kDemapSegs[i].T != kDemapOpen && !(r < (float)kDemapSegs[i].T * A)) ++i;
    const DemapSeg g = kDemapSegs[i];
    const float x = g.O == 0 ? r : r + (float)g.O * A;
    const float v = ((float)g.C * A) * x / nv;
    return g.sgn < 0 ? -v : v;
}

__device__ __forceinline__ int16_t quantize_llr(float llr) {
    const float scaled = llr * 128.0f;
    int q = (int)(scaled >= 0.0f ? scaled + 0.5f : scaled - 0.5f);
    q = q > 10000 ? 10000 : (q < -10000 ? -10000 : q);
    return (int16_t)q;
}

// Extension rows: parity of row R = XOR of its info and core-parity columns.
template <int BG, int R>
__device__ __forceinline__ void ext_rows(uint8_t *cw, const uint32_t *etab, int zx, int Zc) {
    if constexpr (R < Graph<BG>::kRows) {
        constexpr int kb = Graph<BG>::kInfoCols;
        uint32_t acc = 0;
#pragma unroll
        for (int e = Graph<BG>::start(R); e < Graph<BG>::start(R + 1); ++e) {
            if (Graph<BG>::col(e) >= kb + 4) continue;   // this row's own parity column
            const uint32_t t = etab[e];
            int c = zx + (int)(t >> 16);
            c = c >= Zc ? c - Zc : c;
            acc ^= cw[(t & 0xFFFFu) + c];
        }
        cw[(kb + R) * Zc + zx] = (uint8_t)acc;
        ext_rows<BG, R + 1>(cw, etab, zx, Zc);
    }
}

template <int BG, int R>
__device__ __forceinline__ uint32_t core_lambda(const uint8_t *cw, const uint32_t *etab, int zx,
                                                int Zc) {
    constexpr int kb = Graph<BG>::kInfoCols;
    uint32_t acc = 0;
#pragma unroll
    for (int e = Graph<BG>::start(R); e < Graph<BG>::start(R + 1); ++e) {
        if (Graph<BG>::col(e) >= kb) continue;
        const uint32_t t = etab[e];
        int c = zx + (int)(t >> 16);
        c = c >= Zc ? c - Zc : c;
        acc ^= cw[(t & 0xFFFFu) + c];
    }
    return acc;
}

template <int BG>
__global__ void __launch_bounds__(kSynthThreads) ldpc_synth_kernel(SynthArgs a) {
    constexpr int kb = Graph<BG>::kInfoCols, ncol = Graph<BG>::kCols;
    __shared__ uint8_t cw[68 * kMaxZc];
    __shared__ uint8_t lam[4][kMaxZc];
    __shared__ uint32_t crc_tab[256];
    __shared__ uint32_t slots[kSynthThreads / 64];
    __shared__ float fmax_slots[kSynthThreads / 64];
    const int cb = blockIdx.x, t = threadIdx.x, nt = blockDim.x;
    const int Zc = a.Zc, A = a.A, B = a.B, K = a.K, N = a.N, E = a.E, Qm = a.Qm;
    const int ntot = ncol * Zc;

    // 1. payload
    for (int k = t; k < ntot; k += nt) {
        uint8_t v = 0;
        if (k < A)
            v = a.payload ? (uint8_t)(a.payload[(size_t)cb * A + k] & 1u)
                          : (uint8_t)((rnd64(a.seed, cb, 0, (uint32_t)k >> 6) >> (k & 63)) & 1u);
        cw[k] = v;   // CRC, filler and parity positions start at zero
    }
    // 2. CRC of the A/8 payload bytes
    const int nbytes = A / 8, clen = B - A;
    const CrcCtx crc = crc_setup(a.crctype, nbytes, crc_tab);
    __syncthreads();
    {
        const Crc &c = crc.spec;
        uint32_t r = 0;
        for (int b = crc.b_lo; b < crc.b_hi; ++b) {
            uint32_t byte = 0;
            for (int i = 0; i < 8; ++i) byte = (byte << 1) | cw[8 * b + i];
            r = ((r << 8) ^ crc_tab[((r >> (c.w - 8)) ^ byte) & 0xFFu]) & c.mask;
        }
        uint32_t part = r ? crc_mulmod(r, crc.mult, c) : 0u;
        part = wave_xor(part);
        if ((t & 63) == 0) slots[t >> 6] = part;
        __syncthreads();
        uint32_t total = 0;
        for (int w = 0; w < (nt + 63) / 64; ++w) total ^= slots[w];
        for (int i = t; i < clen; i += nt) cw[A + i] = (uint8_t)((total >> (clen - 1 - i)) & 1u);
    }
    __syncthreads();
    if (a.bits_out)
        for (int k = t; k < B; k += nt) a.bits_out[(size_t)cb * B + k] = cw[k];

    // 3. encode
    const uint32_t *etab = a.etab;
    for (int zx = t; zx < Zc; zx += nt) {
        lam[0][zx] = (uint8_t)core_lambda<BG, 0>(cw, etab, zx, Zc);
        lam[1][zx] = (uint8_t)core_lambda<BG, 1>(cw, etab, zx, Zc);
        lam[2][zx] = (uint8_t)core_lambda<BG, 2>(cw, etab, zx, Zc);
        lam[3][zx] = (uint8_t)core_lambda<BG, 3>(cw, etab, zx, Zc);
    }
    __syncthreads();
    // the four core rows sum to one circulant on p0: row 0's p0 entry pairs
    // with the unpaired entry of row 1 (BG1) or row 2 (BG2)
    constexpr int ea = find_edge_ct<BG>(0, kb);
    constexpr int eb = find_edge_ct<BG>(BG == 1 ? 1 : 2, kb);
    static_assert(ea >= 0 && eb >= 0, "core parity structure");
    const int sa = (int)(etab[ea] >> 16), sb = (int)(etab[eb] >> 16);
    uint8_t *p0 = cw + kb * Zc, *p1 = p0 + Zc, *p2 = p1 + Zc, *p3 = p2 + Zc;
    for (int zx = t; zx < Zc; zx += nt) {
        int c = zx + sb;
        c = c >= Zc ? c - Zc : c;
        p0[c] = lam[0][zx] ^ lam[1][zx] ^ lam[2][zx] ^ lam[3][zx];
    }
    __syncthreads();
    for (int zx = t; zx < Zc; zx += nt) {
        int c = zx + sa;
        c = c >= Zc ? c - Zc : c;
        const uint8_t rot = p0[c];
        p1[zx] = lam[0][zx] ^ rot;
        p3[zx] = lam[3][zx] ^ rot;
        p2[zx] = BG == 1 ? (uint8_t)(lam[2][zx] ^ p3[zx]) : (uint8_t)(lam[1][zx] ^ p1[zx]);
    }
    __syncthreads();
    for (int zx = t; zx < Zc; zx += nt) ext_rows<BG, 4>(cw, etab, zx, Zc);
    __syncthreads();
    if (a.cw_out)
        for (int k = t; k < ntot; k += nt) a.cw_out[(size_t)cb * ntot + k] = cw[k];

    // 4. rate matching + channel, one symbol per lane step
    const int S = E / Qm;
    const int lo = B - 2 * Zc, hi = K - 2 * Zc, F = hi - lo, navail = N - F;
    const int a0 = a.k0 < lo ? a.k0 : (a.k0 < hi ? lo : a.k0 - F);
    const float nv = a.noise_var ? a.noise_var[cb] : powf(10.0f, a.noise_db[cb] / 10.0f);
    const float sigma = sqrtf(nv / 2.0f);
    const int nb = Qm == 1 ? 1 : Qm / 2;
    // the reference's 1 / sqrt(2), 1 / sqrt(10), ... rounded as it computes them
    const float amp = Qm <= 2 ? 1.0f / sqrtf(2.0f)
                              : (Qm == 4 ? 1.0f / sqrtf(10.0f)
                                         : (Qm == 6 ? 1.0f / sqrtf(42.0f) : 1.0f / sqrtf(170.0f)));
    float local_max = 0.0f;
    for (int j = t; j < S; j += nt) {
        int bits[8];
        for (int i = 0; i < Qm; ++i) {
            int q = a0 + i * S + j;
            q %= navail;
            const int pos = q < lo ? q : q + F;
            bits[i] = cw[pos + 2 * Zc];
            if (a.rm_out) a.rm_out[(size_t)cb * E + (size_t)j * Qm + i] = (uint8_t)bits[i];
        }
        float na, nb_;
        if (a.noise) {
            na = a.noise[((size_t)cb * S + j) * 2];
            nb_ = a.noise[((size_t)cb * S + j) * 2 + 1];
        } else {
            const uint64_t r = rnd64(a.seed, cb, 1, (uint32_t)j);
            const float u1 = unit_open(r), u2 = unit_open(r << 24);
            const float rad = sqrtf(-2.0f * logf(u1));
            na = sigma * (rad * cosf(6.283185307179586f * u2));
            nb_ = sigma * (rad * sinf(6.283185307179586f * u2));
        }
        float llr[8];
        if (Qm == 1) {
            const float s = (float)(1 - 2 * bits[0]) * amp;
            llr[0] = 4.0f * ((s + na) + (s + nb_)) * amp / nv;
        } else if (Qm == 2) {
            llr[0] = 4.0f * ((float)(1 - 2 * bits[0]) * amp + na) * amp / nv;
            llr[1] = 4.0f * ((float)(1 - 2 * bits[1]) * amp + nb_) * amp / nv;
        } else {
            int ab[2][4];
            for (int k = 0; k < nb; ++k) {
                ab[0][k] = bits[2 * k];
                ab[1][k] = bits[2 * k + 1];
            }
            const float rx[2] = {pam_level(ab[0], nb) * amp + na, pam_level(ab[1], nb) * amp + nb_};
            for (int ax = 0; ax < 2; ++ax)
                for (int k = 0; k < nb; ++k) llr[2 * k + ax] = demap_pam(rx[ax], nb, k, amp, nv);
        }
        for (int i = 0; i < Qm; ++i) {
            const size_t o = (size_t)cb * E + (size_t)j * Qm + i;
            if (a.llr_out) a.llr_out[o] = quantize_llr(llr[i]);
            if (a.llr_f32) a.llr_f32[o] = llr[i];
            const float m = fabsf(llr[i]);
            local_max = m > local_max ? m : local_max;
        }
    }
    if (!a.beliefs) return;

    // 5. float rate recovery (gather form; repeats summed in double)
    for (int o = 32; o >= 1; o >>= 1) {
        const float v = __shfl_xor(local_max, o, 64);
        local_max = v > local_max ? v : local_max;
    }
    if ((t & 63) == 0) fmax_slots[t >> 6] = local_max;
    __syncthreads();   // also orders this workgroup's llr_f32 writes before the reads below
    float amax = 0.0f;
    for (int w = 0; w < (nt + 63) / 64; ++w) amax = fmax_slots[w] > amax ? fmax_slots[w] : amax;
    const float fill = (float)((double)amax * 10.0);
    const float *f = a.llr_f32 + (size_t)cb * E;
    float *out = a.beliefs + (size_t)cb * ntot;
    for (int p = t; p < ntot; p += nt) {
        float v;
        if (p < 2 * Zc) {
            v = 0.0f;
        } else if (p >= B && p < K) {
            v = fill;
        } else {
            const int pos = p - 2 * Zc;
            const int q = pos < lo ? pos : pos - F;
            int k = q - a0;
            if (k < 0) k += navail;
            double acc = 0.0;
            for (; k < E; k += navail) {
                const int j = k % S, i = k / S;
                acc += (double)f[(size_t)j * Qm + i];
            }
            v = (float)acc;
        }
        out[p] = v;
    }
}

}  // namespace nrldpc

#ifndef NRLDPC_CPU_EMU
extern "C" hipError_t nrldpc_launch_synth(int bgn, const nrldpc::SynthArgs *a, hipStream_t s) {
    using namespace nrldpc;
    const dim3 grid(a->n), block(kSynthThreads);
    if (bgn == 1)
        hipLaunchKernelGGL(ldpc_synth_kernel<1>, grid, block, 0, s, *a);
    else
        hipLaunchKernelGGL(ldpc_synth_kernel<2>, grid, block, 0, s, *a);
    return hipGetLastError();
}
#endif
