// Rate recovery + HARQ combine in gather form, shared by the decode kernels
// (ldpc_kernels.hip, ldpc_bp.hip).  Include after the HIP runtime header (or
// the CPU emulation header) and ldpc_device.h.
#pragma once
#include "ldpc_device.h"

namespace nrldpc {

// ---------------------------------------------------------------------------
// Rate recovery + HARQ combine fused into the LDS fill
// (c_lib/ldpc_rate_recovery.c:24-65).  Gather form: buffer position p takes
// the sum of every LLR k with compact index q(p) = (ck0 + k) mod (N - F),
// k < E; the sum wraps mod 2^16 exactly like the reference's int16 +=.
//
// recover_fill writes positions first, first + stride, ... < n of one
// codeblock into lq (as T) and, when the codeblock has a soft buffer, into
// it.  A lane works on kRecoverBatch positions at a time and issues their
// first LLR loads back to back, so the gather runs at memory throughput
// instead of one dependent global load per position; further contributions
// (E > N - F, repetition) are added in a second, usually empty, loop.
// ---------------------------------------------------------------------------
constexpr int kRecoverBatch = 8;

// Belief as the decoder stores it: int16 as is, float / 128 (algos 0 and 4,
// c_lib/floatpoint_ldpc_decoder.c:325-327).
template <typename T> __device__ inline T from_soft(int16_t v);
template <> __device__ inline int16_t from_soft<int16_t>(int16_t v) { return v; }
template <> __device__ inline float from_soft<float>(int16_t v) { return (float)v / 128.0f; }

struct RecoverCtx {
    int Zc, B, K, F, lo, nc, ck0, S, E, Qm;
};

__device__ __forceinline__ RecoverCtx recover_ctx(const CbDesc &d) {
    RecoverCtx c;
    c.Zc = d.Zc;
    c.B = d.B;
    c.K = d.K;
    c.F = d.K - d.B;
    c.lo = d.B - 2 * d.Zc;
    const int hi = d.K - 2 * d.Zc;
    c.nc = d.N - c.F;
    c.ck0 = d.k0 < c.lo ? d.k0 : (d.k0 < hi ? c.lo : d.k0 - c.F);
    c.E = d.E;
    c.Qm = d.Qm;
    c.S = d.E / d.Qm;
    return c;
}

// LLR index of rate-matched bit k (bit-interleaved: symbol k mod S, bit k / S).
__device__ __forceinline__ int llr_index(int k, const RecoverCtx &c) {
    int sym = k, bit = 0;
    while (sym >= c.S) {
        sym -= c.S;
        ++bit;
    }
    return sym * c.Qm + bit;
}

template <typename T>
__device__ __forceinline__ void recover_fill(T *lq, const CbDesc &d, const int16_t *__restrict__ llr,
                             int16_t *soft, int filler, int first, int stride, int n) {
    const RecoverCtx c = recover_ctx(d);
    const bool harq = d.harq_on && soft;
    for (int i0 = first; i0 < n; i0 += stride * kRecoverBatch) {
        int acc[kRecoverBatch], k[kRecoverBatch];
#pragma unroll
        for (int u = 0; u < kRecoverBatch; ++u) {
            const int p = i0 + u * stride;
            acc[u] = 0;
            k[u] = c.E;   // no contribution
            if (p < n) {
                if (harq) acc[u] = soft[p];
                if (p >= c.B && p < c.K) {
                    acc[u] = filler;
                } else if (p >= 2 * c.Zc) {
                    const int pos = p - 2 * c.Zc;
                    int kk = (pos < c.lo ? pos : pos - c.F) - c.ck0;
                    if (kk < 0) kk += c.nc;
                    k[u] = kk;
                }
            }
        }
        int add[kRecoverBatch];
#pragma unroll
        for (int u = 0; u < kRecoverBatch; ++u)
            add[u] = k[u] < c.E ? (int)llr[llr_index(k[u], c)] : 0;
#pragma unroll
        for (int u = 0; u < kRecoverBatch; ++u) {
            const int p = i0 + u * stride;
            int v = acc[u] + add[u];
            for (int kk = k[u] + c.nc; kk < c.E; kk += c.nc) v += llr[llr_index(kk, c)];
            if (p < n) {
                const int16_t w = (int16_t)v;   // the reference's int16 += wraps mod 2^16
                if (soft) soft[p] = w;
                lq[p] = from_soft<T>(w);
            }
        }
    }
}

}  // namespace nrldpc
