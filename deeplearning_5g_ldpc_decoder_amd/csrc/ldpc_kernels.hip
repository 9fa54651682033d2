// HIP kernels (gfx950) for the 5G NR LDPC decode path.
//
// Mapping: one workgroup decodes one codeblock; lane zx owns lifted check zx
// of every base-graph row.  The layered schedule of the reference
// (c_lib/fixedpoint_ldpc_decoder.c:182-212) walks the rows in order and the
// Zc lifted checks of one row touch disjoint variable nodes, so running them
// in parallel and putting a workgroup barrier between rows reproduces the
// reference's serial order bit for bit.
//
// State:
//   * variable-node beliefs LQ: LDS (int16 or float, N+2Zc entries), filled by
//     the fused rate recovery / HARQ combine;
//   * min-sum check-to-variable messages: VGPRs, compressed per lifted check
//     to (min1', min2', argmin, one sign bit per edge) -- the reference's
//     message of edge j is exactly sign_j * (j == argmin ? min2' : min1'),
//     so the compressed form re-expands bit-exactly.
// (The FP32 BP decoder, algo 0, is in ldpc_bp.hip.)
// Rows are unrolled at compile time per base graph so the per-row state sits
// at fixed register indices.
//
// After every iteration the hard decisions of the first B beliefs are CRC'd
// in parallel: each lane runs the byte-table CRC over its slice of bytes,
// shifts its remainder by x^(8*bytes_after) mod g (GF(2) linearity), and the
// remainders are XOR-reduced across the workgroup.  A zero remainder stops the
// codeblock (c_lib/fixedpoint_ldpc_decoder.c:214-219).
#ifdef NRLDPC_CPU_EMU
// Host build of the kernel source for CPU tests (tests/kernel_emu/).
#include "hip_cpu_emu.h"
#define NRLDPC_DYNAMIC_LDS(name) unsigned char *name = emu_dynamic_lds()
#else
#include <hip/hip_runtime.h>
#define NRLDPC_DYNAMIC_LDS(name) extern __shared__ __align__(16) unsigned char name[]
#endif

#include <algorithm>
#include <utility>

#include "ldpc_device.h"
#include "ldpc_graph.h"
#include "ldpc_crc.h"
#include "ldpc_recover.h"

// Opaque per-row redefinition of a loop-invariant value: keeps the compiler
// from hoisting every row's (constant-offset) address arithmetic out of the
// iteration loop, which would hold hundreds of registers across it.
#ifdef NRLDPC_CPU_EMU
#define NRLDPC_OPAQUE_LANE(x) (void)(x)
#define NRLDPC_OPAQUE_UNIFORM(x) (void)(x)
#define NRLDPC_WAVE_UNIFORM(x) (x)
#else
#define NRLDPC_OPAQUE_LANE(x) asm volatile("" : "+v"(x))
#define NRLDPC_OPAQUE_UNIFORM(x) asm volatile("" : "+s"(x))
// a value equal on every lane of the wave, moved to a scalar register
#define NRLDPC_WAVE_UNIFORM(x) __builtin_amdgcn_readfirstlane(x)
#endif

namespace nrldpc {

// Scatter form of the same recovery, for codeblocks without repetition
// (E <= N - F: every buffer position receives at most one LLR).  The LLRs are
// read once, in order and coalesced, kScatterBatch loads in flight per lane;
// LLR j is rate-matched bit k = (j mod Qm) * S + j / Qm, which lands on
// compact index q = (ck0 + k) mod (N - F).
constexpr int kScatterBatch = 16;

// Initial beliefs: filler positions take the filler value, the others the
// soft buffer (HARQ combining) or 0.  Lanes write 8 positions (16 bytes) at a
// time: chunk c = first, first + stride, ... covers positions [8c, 8c + 8);
// the beliefs and soft buffers are 16-byte aligned.
__device__ __forceinline__ uint32_t put_half(uint32_t w, int hi, int v) {
    return hi ? ((w & 0xFFFFu) | ((uint32_t)v << 16)) : ((w & 0xFFFF0000u) | ((uint32_t)v & 0xFFFFu));
}

__device__ __forceinline__ void recover_init(int16_t *lq, const CbDesc &d, const int16_t *soft, int filler,
                             int first, int stride, int n) {
    const bool harq = d.harq_on && soft;
    const int chunks = n >> 3;
    for (int c = first; c < chunks; c += stride) {
        const int p0 = 8 * c;
        uint4 v;
        if (harq) {
            v = *reinterpret_cast<const uint4 *>(soft + p0);
        } else {
            v.x = v.y = v.z = v.w = 0u;
        }
        if (p0 + 8 > d.B && p0 < d.K) {   // the chunk meets the filler range
            uint32_t w[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
            for (int u = 0; u < 8; ++u)
                if (p0 + u >= d.B && p0 + u < d.K) w[u >> 1] = put_half(w[u >> 1], u & 1, filler);
            v.x = w[0];
            v.y = w[1];
            v.z = w[2];
            v.w = w[3];
        }
        *reinterpret_cast<uint4 *>(lq + p0) = v;
    }
    for (int p = 8 * chunks + first; p < n; p += stride)
        lq[p] = (int16_t)((p >= d.B && p < d.K) ? filler : (harq ? (int)soft[p] : 0));
}

// harq: the initial beliefs hold soft-buffer values to add to; otherwise they
// are 0 and each LLR is stored directly.
template <int kBatch = kScatterBatch>
__device__ __forceinline__ void recover_scatter(int16_t *lq, const CbDesc &d, const int16_t *__restrict__ llr,
                                int first, int stride, bool harq) {
    const RecoverCtx c = recover_ctx(d);
    // floor(j / Qm) = umulhi(j, ceil(2^32 / Qm)), exact for j < 2^16 and Qm <= 8
    const uint32_t magic = c.Qm > 1 ? 0xFFFFFFFFu / (uint32_t)c.Qm + 1u : 0u;
    const int p0 = 2 * c.Zc;
    for (int j0 = first; j0 < c.E; j0 += stride * kBatch) {
        int v[kBatch];
#pragma unroll
        for (int u = 0; u < kBatch; ++u) {
            const int j = j0 + u * stride;
            v[u] = j < c.E ? (int)llr[j] : 0;
        }
#pragma unroll
        for (int u = 0; u < kBatch; ++u) {
            const int j = j0 + u * stride;
            if (j < c.E) {
                const int sym = c.Qm > 1 ? (int)__umulhi((uint32_t)j, magic) : j;
                const int bit = j - sym * c.Qm;
                int q = bit * c.S + sym + c.ck0;
                if (q >= c.nc) q -= c.nc;
                const int p = (q < c.lo ? q : q + c.F) + p0;
                lq[p] = (int16_t)(harq ? lq[p] + v[u] : v[u]);
            }
        }
    }
}

// recover_scatter for LLRs in 16-byte chunks (the codeblock's LLRs 16-byte
// aligned, E a multiple of 8): kChunks loads of 8 LLRs in flight per lane, so
// one round trip to memory covers E <= 8 * kChunks * stride.  Within a chunk
// the rate-matched bit k = bit * S + sym advances by S per LLR, and by
// 1 - (Qm - 1) * S when bit wraps to 0 (E = Qm * S).
constexpr int kScatterChunks = 12;

template <int kChunks = kScatterChunks>
__device__ __forceinline__ void recover_scatter_x8(int16_t *lq, const CbDesc &d, const int16_t *__restrict__ llr,
                                   int first, int stride, bool harq, bool qpsk_runs) {
    const RecoverCtx c = recover_ctx(d);
    const uint4 *src = reinterpret_cast<const uint4 *>(llr);
    const int chunks = c.E >> 3;
    const int p0 = 2 * c.Zc;
    const int back = 1 - (c.Qm - 1) * c.S;
    const uint32_t magic = c.Qm > 1 ? 0xFFFFFFFFu / (uint32_t)c.Qm + 1u : 0u;   // as in recover_scatter
    for (int c0 = first; c0 < chunks; c0 += stride * kChunks) {
        uint4 v[kChunks];
#pragma unroll
        for (int u = 0; u < kChunks; ++u) {
            const int ch = c0 + u * stride;
            if (ch < chunks) v[u] = src[ch];
        }
#pragma unroll
        for (int u = 0; u < kChunks; ++u) {
            const int ch = c0 + u * stride;
            if (ch < chunks) {
                const int j0 = 8 * ch;
                const uint32_t w[4] = {v[u].x, v[u].y, v[u].z, v[u].w};
                if (qpsk_runs && c.Qm == 2 && !harq) {
                    // QPSK: the chunk's even LLRs are bits 4ch .. 4ch + 3, its
                    // odd ones bits S + 4ch .. S + 4ch + 3.  Each run lands on
                    // 4 consecutive beliefs unless it crosses the circular
                    // buffer's end or the filler gap; then, 8-byte aligned,
                    // it is one 8-byte store of two packed words.
                    int pp[2];
                    bool ok = true;
#pragma unroll
                    for (int h = 0; h < 2; ++h) {
                        int q = 4 * ch + h * c.S + c.ck0;
                        if (q >= c.nc) q -= c.nc;
                        ok = ok && q + 3 < c.nc && (q + 3 < c.lo || q >= c.lo);
                        pp[h] = (q < c.lo ? q : q + c.F) + p0;
                        ok = ok && (pp[h] & 3) == 0;
                    }
                    if (ok) {
#ifdef NRLDPC_CPU_EMU
                        const uint32_t ev0 = (w[0] & 0xFFFFu) | (w[1] << 16);
                        const uint32_t ev1 = (w[2] & 0xFFFFu) | (w[3] << 16);
                        const uint32_t od0 = (w[0] >> 16) | (w[1] & 0xFFFF0000u);
                        const uint32_t od1 = (w[2] >> 16) | (w[3] & 0xFFFF0000u);
#else
                        const uint32_t ev0 = __builtin_amdgcn_perm(w[1], w[0], 0x05040100u);
                        const uint32_t ev1 = __builtin_amdgcn_perm(w[3], w[2], 0x05040100u);
                        const uint32_t od0 = __builtin_amdgcn_perm(w[1], w[0], 0x07060302u);
                        const uint32_t od1 = __builtin_amdgcn_perm(w[3], w[2], 0x07060302u);
#endif
                        *reinterpret_cast<uint2 *>(lq + pp[0]) = make_uint2(ev0, ev1);
                        *reinterpret_cast<uint2 *>(lq + pp[1]) = make_uint2(od0, od1);
                        continue;
                    }
                }
                const int sym = c.Qm > 1 ? (int)__umulhi((uint32_t)j0, magic) : j0;
                int bit = j0 - sym * c.Qm;
                int k = bit * c.S + sym;   // rate-matched bit of LLR j0
#pragma unroll
                for (int e = 0; e < 8; ++e) {
                    const int val = (int)(int16_t)((e & 1) ? (w[e >> 1] >> 16) : (w[e >> 1] & 0xFFFFu));
                    int q = k + c.ck0;   // k < E <= N - F: one wrap at most
                    if (q >= c.nc) q -= c.nc;
                    const int pp = (q < c.lo ? q : q + c.F) + p0;
                    lq[pp] = (int16_t)(harq ? lq[pp] + val : val);
                    if (++bit == c.Qm) {
                        bit = 0;
                        k += back;
                    } else {
                        k += c.S;
                    }
                }
            }
        }
    }
}

// Rate recovery of one codeblock into int16 beliefs and its soft buffer by
// lanes first, first + stride, ... .  Every lane of the workgroup calls it
// (it holds two workgroup barriers); lanes with active = false only join
// them.  The caller's next barrier orders the soft-buffer write-back before
// decoding changes the beliefs.  mid() runs after the first barrier, before
// the scatter issues its loads: loads it issues share the scatter's memory
// round trip.  qpsk_runs: the 8-byte stores of QPSK runs in the aligned
// scatter (the packed tiles leave them out: the extra registers pushed that
// kernel into scratch).
struct NoHook {
    __device__ void operator()() const {}
};
template <int kBatch = kScatterBatch, typename Mid = NoHook>
__device__ __forceinline__ void recover_i16(int16_t *lq, const CbDesc &d, const int16_t *llr, int16_t *soft,
                            int filler, int first, int stride, int n, bool active,
                            Mid &&mid = NoHook(), bool qpsk_runs = false) {
    const bool scatter = active && d.E <= d.N - (d.K - d.B);
    if (active) {
        if (scatter)
            recover_init(lq, d, soft, filler, first, stride, n);
        else
            recover_fill<int16_t>(lq, d, llr, soft, filler, first, stride, n);
    }
    __syncthreads();
    mid();
    if (scatter) {
        const bool harq = d.harq_on && soft;
        if (((uintptr_t)llr & 15u) == 0 && (d.E & 7) == 0)
            recover_scatter_x8(lq, d, llr, first, stride, harq, qpsk_runs);
        else
            recover_scatter<kBatch>(lq, d, llr, first, stride, harq);
    }
    __syncthreads();
    if (scatter && soft)
        for (int p = first; p < n; p += stride) soft[p] = lq[p];
}

// ---------------------------------------------------------------------------
// Algo 1: int16 normalized/offset min-sum
// (c_lib/fixedpoint_ldpc_decoder.c:28-66, 103-144).
// ---------------------------------------------------------------------------
template <int BG>
struct StateI16 {
    uint32_t mag[Graph<BG>::kRows];   // min1' | min2' << 16
    uint32_t sgn[Graph<BG>::kRows];   // sign bit per edge | argmin << 24
};

// Bit j of v as 0 / -1.
__device__ __forceinline__ int sext_bit(uint32_t v, int j) {
#ifdef NRLDPC_CPU_EMU
    return -(int)((v >> j) & 1u);
#else
    return __builtin_amdgcn_sbfe((int)v, (unsigned)j, 1u);
#endif
}

__device__ __forceinline__ int wrap16(int v) { return (int)(int16_t)v; }

// int16 wrap of a - b, sign-extended: one SDWA subtract on gfx950.
__device__ __forceinline__ int sub_wrap16(int a, int b) {
#ifdef NRLDPC_CPU_EMU
    return wrap16(a - b);
#else
    int r;
    asm("v_sub_u32_sdwa %0, %1, %2 dst_sel:WORD_0 dst_unused:UNUSED_SEXT "
        "src0_sel:DWORD src1_sel:DWORD"
        : "=v"(r) : "v"(a), "v"(b));
    return r;
#endif
}

// Median of three (a clamp of b into [a, c] when a <= c).
__device__ __forceinline__ int med3_i32(int a, int b, int c) {
#ifdef NRLDPC_CPU_EMU
    return max(min(a, b), min(max(a, b), c));
#else
    int r;
    asm("v_med3_i32 %0, %1, %2, %3" : "=v"(r) : "v"(a), "v"(b), "v"(c));
    return r;
#endif
}

__device__ __forceinline__ uint32_t med3_u32(uint32_t a, uint32_t b, uint32_t c) {
#ifdef NRLDPC_CPU_EMU
    return max(min(a, b), min(max(a, b), c));
#else
    uint32_t r;
    asm("v_med3_u32 %0, %1, %2, %3" : "=v"(r) : "v"(a), "v"(b), "v"(c));
    return r;
#endif
}

// (acc << 1) | sign bit of v, in one funnel shift.
__device__ __forceinline__ uint32_t push_sign(uint32_t acc, int v) {
#ifdef NRLDPC_CPU_EMU
    return (acc << 1) | ((uint32_t)v >> 31);
#else
    return __builtin_amdgcn_alignbit(acc, (uint32_t)v, 31u);
#endif
}

// (alpha * max(m - beta, 0) + 2^14) >> 15 as int16 (c_lib/fixedpoint_ldpc_decoder.c:38-46).
// alpha is in [-32768, 32767] and the offset in [0, 65535], so a 24-bit
// multiply is exact and the product fits in 32 bits.
__device__ __forceinline__ int scale_i16(int m, int alpha, int beta) {
    const int off = max(m - beta, 0);
#ifdef NRLDPC_CPU_EMU
    const int scaled = alpha * off;
#else
    // full-rate 24-bit multiply (__mul24 lowers to a quarter-rate v_mul_lo_u32 here)
    int scaled;
    asm("v_mul_i32_i24 %0, %1, %2" : "=v"(scaled) : "v"(alpha), "v"(off));
#endif
    return wrap16((scaled + (1 << 14)) >> 15);
}

// scale_i16 for beta = 0 and m in [0, 32767] (min(m, 32767) = m): one
// 24-bit multiply-add and a shift.
__device__ __forceinline__ int scale_simple(int m, int alpha) {
#ifdef NRLDPC_CPU_EMU
    return (alpha * m + (1 << 14)) >> 15;
#else
    int r;
    asm("v_mad_i32_i24 %0, %1, %2, %3" : "=v"(r) : "v"(alpha), "v"(m), "s"(1 << 14));
    return r >> 15;
#endif
}

// Larger than every two-minimum key (magnitudes reach 32768, see below).
constexpr uint32_t kKeyMax = (32768u << 5) | 31u;

// One base-graph row, one lifted check per lane (c_lib/fixedpoint_ldpc_decoder.c:103-144).
//
// Magnitudes and edge indices travel together as key = |rel| << 5 | j, so the
// two-minimum search is a min and a med3 per edge; unique keys make the
// first-occurrence tie-break of the reference's strict '<' come out the same.
// |rel| is taken without the reference's 32767 cap for rel = -32768: both
// minima are clipped to max_llr <= 32767 before use, and whenever 32768 is
// one of the two smallest magnitudes both clipped minima are equal, so the
// messages are the same; likewise when every magnitude is 32767 (reference
// argmin -1).  Sign bits are kept in reverse edge order (edge j at bit
// deg-1-j).  kGeneral = false is the fast path for decoder rows whose
// scaled minima stay in [0, 32767] (alpha >= 0 and
// (alpha * max(max_llr - beta, 0) + 2^14) >> 15 <= 32767) and whose
// min_llr <= max_llr: messages are then plain negations and the LQ clip is a
// clamp; true is exact for every row (message wrapped to int16 before the
// add, clip as the reference orders its two tests).  The host decides
// (DecParams::general).  (alpha, beta) is the row's pair: the decoder row's
// for algo 1, the table's for algos 2 and 3.
template <int BG, int R, bool kGeneral>
__device__ __forceinline__ void row_i16(int16_t *__restrict__ lq, StateI16<BG> &st,
                                        const uint32_t *__restrict__ etab, int zx,
                                        int Zc, const DecParams &p, int alpha, int beta) {
    constexpr int e0 = Graph<BG>::start(R);
    constexpr int deg = Graph<BG>::start(R + 1) - e0;
    const uint32_t mag = st.mag[R], sb = st.sgn[R];
    const int o1 = wrap16((int)mag), o2 = (int)mag >> 16;
    const int oidx = (int)(sb >> 24);
    int addr[deg], rel[deg];
    uint32_t k1 = kKeyMax, k2 = kKeyMax, negs = 0;
#pragma unroll
    for (int j = 0; j < deg; ++j) {
        const uint32_t t = etab[e0 + j];
        const uint32_t c = (uint32_t)zx + (t >> 16);
        addr[j] = (int)(t & 0xFFFFu) + (int)min(c, c - (uint32_t)Zc);
        // previous message of this edge (the wrap of its int16 form is
        // absorbed by the wrap of rel)
        const int ov = (j == oidx) ? o2 : o1;
        const int s = sext_bit(sb, deg - 1 - j);
        const int v = sub_wrap16((int)lq[addr[j]], (ov ^ s) - s);
        rel[j] = v;
        negs = push_sign(negs, v);
        const uint32_t key = ((uint32_t)max(v, -v) << 5) | (uint32_t)j;
        k2 = med3_u32(k1, key, k2);
        k1 = min(k1, key);
    }
    const int idx = (int)(k1 & 31u);
    const int m1 = min((int)(k1 >> 5), p.max_llr), m2 = min((int)(k2 >> 5), p.max_llr);
    const int n1 = scale_i16(m1, alpha, beta), n2 = scale_i16(m2, alpha, beta);
    constexpr uint32_t mask = (1u << deg) - 1u;
    const uint32_t bits = ((__popc(negs) & 1) ? ~negs : negs) & mask;
#pragma unroll
    for (int j = 0; j < deg; ++j) {
        const int nv = (j == idx) ? n2 : n1;
        const int s = sext_bit(bits, deg - 1 - j);
        int v;
        if constexpr (kGeneral) {
            v = rel[j] + sub_wrap16(nv ^ s, s);
            v = v > p.max_llr ? p.max_llr : max(v, p.min_llr);
        } else {
            v = med3_i32(p.min_llr, rel[j] + ((nv ^ s) - s), p.max_llr);
        }
        lq[addr[j]] = (int16_t)v;
    }
    st.mag[R] = (uint32_t)(uint16_t)n1 | ((uint32_t)(uint16_t)n2 << 16);
    st.sgn[R] = bits | ((uint32_t)idx << 24);
}

// All rows of one iteration; lanes with run = false only take part in the
// barriers.  kTable: row R's (alpha, beta) is the scale word ab[R] (algos 2
// and 3, c_lib/neural_fixedpoint_ldpc_decoder.c:248-250,
// c_lib/degree_fixedpoint_ldpc_decoder.c:230-231); ab is uniform, so the
// word is a scalar load.
template <int BG, int R, bool kGeneral, bool kTable>
__device__ __forceinline__ void rows_i16(int16_t *lq, StateI16<BG> &st, const uint32_t *etab,
                                         int zx, int Zc, bool run, const DecParams &p,
                                         const uint32_t *ab) {
    if constexpr (R < Graph<BG>::kRows) {
        int alpha = p.alpha, beta = p.beta;
        if constexpr (kTable) {
            const uint32_t w = ab[R];
            alpha = (int)(int16_t)(w & 0xFFFFu);
            beta = (int)(int16_t)(w >> 16);
        }
        if (run) row_i16<BG, R, kGeneral>(lq, st, etab, zx, Zc, p, alpha, beta);
        if constexpr (kRowGroups<BG>.barrier_after[R]) __syncthreads();
        rows_i16<BG, R + 1, kGeneral, kTable>(lq, st, etab, zx, Zc, run, p, ab);
    }
}

// Scale words of iteration it (null unless kTable).
template <int BG, bool kTable>
__device__ __forceinline__ const uint32_t *scale_row(const DecParams &p, int it) {
    if constexpr (kTable)
        return p.scale[BG - 1] + min(it, p.scale_iters - 1) * Graph<BG>::kRows;
    else
        return nullptr;
}

template <int BG>
__device__ __forceinline__ void init_state(StateI16<BG> &st) {
#pragma unroll
    for (int r = 0; r < Graph<BG>::kRows; ++r) {
        st.mag[r] = 0u;
        st.sgn[r] = 255u << 24;
    }
}

// ---------------------------------------------------------------------------
// Lifting-size specialized int16 decode: every codeblock of the launch has
// base graph BG and lifting size ZC, so each edge's circulant shift and
// column offset are compile-time constants (LDS addresses become immediates,
// no per-edge scalar work), and each lane keeps its checks' messages
// explicitly, one int16 per edge in half a VGPR, instead of re-expanding
// them from (min1', min2', argmin, signs).  Arithmetic, schedule and results
// are those of row_i16 (c_lib/fixedpoint_ldpc_decoder.c:103-144).
// ---------------------------------------------------------------------------

// Lifting-size set index of TS 38.212 Table 5.3.2-1: zc = a * 2^j with
// a = 2, 3, 5, 7, 9, 11, 13, 15 for sets 0..7; -1 if zc is not a lifting size.
__host__ __device__ constexpr int lifting_set_ct(int zc) {
    constexpr int a[8] = {2, 3, 5, 7, 9, 11, 13, 15};
    for (int i = 0; i < 8; ++i) {
        if (zc <= 0 || zc > kMaxZc || zc % a[i]) continue;
        const int q = zc / a[i];
        if ((q & (q - 1)) == 0) return i;
    }
    return -1;
}
static_assert(lifting_set_ct(2) == 0 && lifting_set_ct(384) == 1 && lifting_set_ct(320) == 2 &&
                  lifting_set_ct(224) == 3 && lifting_set_ct(288) == 4 &&
                  lifting_set_ct(352) == 5 && lifting_set_ct(208) == 6 &&
                  lifting_set_ct(240) == 7 && lifting_set_ct(12) == 1 && lifting_set_ct(17) == -1,
              "lifting-size sets");

template <int BG, int ZC>
struct FixedCode {
    static constexpr int kILS = lifting_set_ct(ZC);
    static_assert(kILS >= 0, "ZC is not a lifting size");
    static constexpr int kMsgWords = (Graph<BG>::kEdges + 1) / 2;
    // byte offset of edge e's column block and its shift in bytes
    __host__ __device__ static constexpr uint32_t col_bytes(int e) {
        return 2u * (uint32_t)Graph<BG>::col(e) * (uint32_t)ZC;
    }
    __host__ __device__ static constexpr uint32_t shift_bytes(int e) {
        return 2u * (uint32_t)((BG == 1 ? kBG1Shift[kILS][e] : kBG2Shift[kILS][e]) % ZC);
    }
};

// The fixed kernel has no static LDS, so its dynamic LDS -- the beliefs --
// starts at LDS address 0 (checked on the host when the kernels are
// configured).  Addressing the beliefs from a constant LDS address 0 lets the
// compiler fold each edge's column offset into the DS instruction's
// immediate; from the array's symbol it adds the (link-time) base per access.
#ifdef NRLDPC_CPU_EMU
typedef char lds_char;
typedef int16_t lds_i16;
__device__ __forceinline__ lds_char *fixed_lds_base(int16_t *lq) {
    return reinterpret_cast<char *>(lq);
}
#else
typedef __attribute__((address_space(3))) char lds_char;
typedef __attribute__((address_space(3))) int16_t lds_i16;
__device__ __forceinline__ lds_char *fixed_lds_base(int16_t *) {
    return (lds_char *)(uintptr_t)0;
}
#endif

// wrap16(a - m) for the int16 message m held in the low / high half of pair:
// one SDWA subtract that reads the half directly.
__device__ __forceinline__ int sub_wrap16_lo(int a, uint32_t pair) {
#ifdef NRLDPC_CPU_EMU
    return wrap16(a - (int)(int16_t)(pair & 0xFFFFu));
#else
    int r;
    asm("v_sub_u32_sdwa %0, %1, %2 dst_sel:WORD_0 dst_unused:UNUSED_SEXT "
        "src0_sel:DWORD src1_sel:WORD_0"
        : "=v"(r) : "v"(a), "v"(pair));
    return r;
#endif
}

__device__ __forceinline__ int sub_wrap16_hi(int a, uint32_t pair) {
#ifdef NRLDPC_CPU_EMU
    return wrap16(a - (int)(pair >> 16));
#else
    int r;
    asm("v_sub_u32_sdwa %0, %1, %2 dst_sel:WORD_0 dst_unused:UNUSED_SEXT "
        "src0_sel:DWORD src1_sel:WORD_1"
        : "=v"(r) : "v"(a), "v"(pair));
    return r;
#endif
}

// Stores the low 16 bits of m as edge E's message (the reference's int16
// store, wrap included).
template <int E, int N>
__device__ __forceinline__ void msg_store(uint32_t (&msg)[N], int m) {
    if constexpr (E & 1)
        msg[E >> 1] = (msg[E >> 1] & 0xFFFFu) | ((uint32_t)m << 16);
    else
        msg[E >> 1] = (msg[E >> 1] & 0xFFFF0000u) | ((uint32_t)m & 0xFFFFu);
}

// Per-row working set between the two passes of a fixed row.
template <int DEG>
struct FixedRowState {
    uint32_t off[DEG];   // byte offset of the edge's variable inside its column block
    int rel[DEG];        // extrinsic input wrap16(LQ - old message)
    uint32_t k1, k2, negs;
};

template <int BG, int ZC, int E, int J, int DEG>
__device__ __forceinline__ void fixed_edge_in(const lds_char *__restrict__ lqb, const uint32_t *msg,
                                              uint32_t zx2, FixedRowState<DEG> &rs) {
    using F = FixedCode<BG, ZC>;
    constexpr uint32_t sh = F::shift_bytes(E), cb = F::col_bytes(E);
    const uint32_t c = zx2 + sh;
    const uint32_t o = min(c, c - 2u * (uint32_t)ZC);
    rs.off[J] = o;
    const int q = *(const lds_i16 *)(lqb + cb + o);
    const int v = (E & 1) ? sub_wrap16_hi(q, msg[E >> 1]) : sub_wrap16_lo(q, msg[E >> 1]);
    rs.rel[J] = v;
    rs.negs = push_sign(rs.negs, v);
    const uint32_t key = ((uint32_t)max(v, -v) << 5) | (uint32_t)J;
    rs.k2 = med3_u32(rs.k1, key, rs.k2);
    rs.k1 = min(rs.k1, key);
}

template <int BG, int ZC, int E, int J, int DEG, bool kGeneral, int N>
__device__ __forceinline__ void fixed_edge_out(lds_char *__restrict__ lqb, uint32_t (&msg)[N],
                                               const FixedRowState<DEG> &rs, int idx, int n1,
                                               int n2, uint32_t bits, const DecParams &p) {
    using F = FixedCode<BG, ZC>;
    constexpr uint32_t cb = F::col_bytes(E);
    const int nv = (J == idx) ? n2 : n1;
    const int s = sext_bit(bits, DEG - 1 - J);
    const int m = (nv ^ s) - s;
    msg_store<E>(msg, m);
    int v;
    if constexpr (kGeneral) {
        v = rs.rel[J] + wrap16(m);
        v = v > p.max_llr ? p.max_llr : max(v, p.min_llr);
    } else {
        v = med3_i32(p.min_llr, rs.rel[J] + m, p.max_llr);
    }
    *(lds_i16 *)(lqb + cb + rs.off[J]) = (int16_t)v;
}

template <int BG, int ZC, int R, bool kGeneral, int N, int... J>
__device__ __forceinline__ void fixed_row(lds_char *__restrict__ lqb, uint32_t (&msg)[N],
                                          uint32_t zx2, const DecParams &p, int alpha, int beta,
                                          std::integer_sequence<int, J...>) {
    constexpr int e0 = Graph<BG>::start(R);
    constexpr int deg = (int)sizeof...(J);
    NRLDPC_OPAQUE_LANE(zx2);
    FixedRowState<deg> rs;
    rs.k1 = kKeyMax;
    rs.k2 = kKeyMax;
    rs.negs = 0;
    (fixed_edge_in<BG, ZC, e0 + J, J, deg>(lqb, msg, zx2, rs), ...);
    const int idx = (int)(rs.k1 & 31u);
    const int m1 = min((int)(rs.k1 >> 5), p.max_llr), m2 = min((int)(rs.k2 >> 5), p.max_llr);
    const int n1 = scale_i16(m1, alpha, beta), n2 = scale_i16(m2, alpha, beta);
    constexpr uint32_t mask = (1u << deg) - 1u;
    const uint32_t bits = ((__popc(rs.negs) & 1) ? ~rs.negs : rs.negs) & mask;
    (fixed_edge_out<BG, ZC, e0 + J, J, deg, kGeneral>(lqb, msg, rs, idx, n1, n2, bits, p), ...);
}

// ---------------------------------------------------------------------------
// Packed form of the fixed row (fast rows: every scaled minimum in
// [0, 32767] and min_llr <= max_llr).  Two edges of a row share one VGPR as
// int16 halves, so the extrinsic subtract, |.| with the reference's 32767 cap,
// the sign and the saturated LQ update run as one packed op for both; each
// row's messages are stored pairwise (edges e0 + 2i, e0 + 2i + 1; an odd last
// edge alone in the low half, high half kept 0).
// ---------------------------------------------------------------------------
#ifdef NRLDPC_CPU_EMU
__device__ __forceinline__ int pk_lo(uint32_t v) { return (int)(int16_t)(v & 0xFFFFu); }
__device__ __forceinline__ int pk_hi(uint32_t v) { return (int)(int16_t)(v >> 16); }
__device__ __forceinline__ uint32_t pk_make(int lo, int hi) {
    return ((uint32_t)lo & 0xFFFFu) | ((uint32_t)hi << 16);
}
__device__ __forceinline__ int sat16(int v) { return v > 32767 ? 32767 : (v < -32768 ? -32768 : v); }
// per half: (a - b) mod 2^16
__device__ __forceinline__ uint32_t pk_subw(uint32_t a, uint32_t b) {
    return pk_make(pk_lo(a) - pk_lo(b), pk_hi(a) - pk_hi(b));
}
// per half: max(r, sat(-r)) = |r| capped at 32767
__device__ __forceinline__ uint32_t pk_abs_sat(uint32_t r) {
    return pk_make(max(pk_lo(r), sat16(-pk_lo(r))), max(pk_hi(r), sat16(-pk_hi(r))));
}
// per half: +1 for a clear sign bit, -1 for a set one
__device__ __forceinline__ uint32_t pk_sgn(uint32_t r) {
    return pk_make(pk_lo(r) < 0 ? -1 : 1, pk_hi(r) < 0 ? -1 : 1);
}
// per half: (a * b) mod 2^16
__device__ __forceinline__ uint32_t pk_mullo(uint32_t a, uint32_t b) {
    return pk_make(pk_lo(a) * pk_lo(b), pk_hi(a) * pk_hi(b));
}
// per half: min(max(sat(a + b), lo), hi)
__device__ __forceinline__ uint32_t pk_update(uint32_t a, uint32_t b, uint32_t lo, uint32_t hi) {
    const int x = min(max(sat16(pk_lo(a) + pk_lo(b)), pk_lo(lo)), pk_lo(hi));
    const int y = min(max(sat16(pk_hi(a) + pk_hi(b)), pk_hi(lo)), pk_hi(hi));
    return pk_make(x, y);
}
// per half: (a + b) mod 2^16
__device__ __forceinline__ uint32_t pk_addw(uint32_t a, uint32_t b) {
    return pk_make(pk_lo(a) + pk_lo(b), pk_hi(a) + pk_hi(b));
}
// low halves of lo and hi as one word
__device__ __forceinline__ uint32_t pk_join(uint32_t lo, uint32_t hi) {
    return (lo & 0xFFFFu) | (hi << 16);
}
// per half: min / max (int16)
__device__ __forceinline__ uint32_t pk_min(uint32_t a, uint32_t b) {
    return pk_make(min(pk_lo(a), pk_lo(b)), min(pk_hi(a), pk_hi(b)));
}
__device__ __forceinline__ uint32_t pk_max(uint32_t a, uint32_t b) {
    return pk_make(max(pk_lo(a), pk_lo(b)), max(pk_hi(a), pk_hi(b)));
}
// per half: max(a, lo) (a saturated sum a + b, then clipped below)
__device__ __forceinline__ uint32_t pk_update_lo(uint32_t a, uint32_t b, uint32_t lo) {
    return pk_make(max(sat16(pk_lo(a) + pk_lo(b)), pk_lo(lo)), max(sat16(pk_hi(a) + pk_hi(b)), pk_hi(lo)));
}
#else
typedef short nr_s2 __attribute__((ext_vector_type(2)));
__device__ __forceinline__ nr_s2 as_s2(uint32_t v) { return __builtin_bit_cast(nr_s2, v); }
__device__ __forceinline__ uint32_t as_u32(nr_s2 v) { return __builtin_bit_cast(uint32_t, v); }
__device__ __forceinline__ uint32_t pk_make(int lo, int hi) {
    return ((uint32_t)lo & 0xFFFFu) | ((uint32_t)hi << 16);
}
__device__ __forceinline__ uint32_t pk_subw(uint32_t a, uint32_t b) {
    return as_u32(as_s2(a) - as_s2(b));
}
__device__ __forceinline__ uint32_t pk_addw(uint32_t a, uint32_t b) {
    return as_u32(as_s2(a) + as_s2(b));
}
// one v_perm_b32: bytes 0-1 of lo, then bytes 0-1 of hi
__device__ __forceinline__ uint32_t pk_join(uint32_t lo, uint32_t hi) {
    return __builtin_amdgcn_perm(hi, lo, 0x05040100u);
}
__device__ __forceinline__ uint32_t pk_abs_sat(uint32_t r) {
    const nr_s2 v = as_s2(r);
    return as_u32(__builtin_elementwise_max(v, __builtin_elementwise_sub_sat((nr_s2)0, v)));
}
// bits 15 and 14 of each half after forcing bit 14: 0b01 -> +1, 0b11 -> -1
__device__ __forceinline__ uint32_t pk_sgn(uint32_t r) {
    return as_u32(as_s2(r | 0x40004000u) >> 14);
}
__device__ __forceinline__ uint32_t pk_mullo(uint32_t a, uint32_t b) {
    return as_u32(as_s2(a) * as_s2(b));
}
__device__ __forceinline__ uint32_t pk_update(uint32_t a, uint32_t b, uint32_t lo, uint32_t hi) {
    const nr_s2 v = __builtin_elementwise_add_sat(as_s2(a), as_s2(b));
    return as_u32(__builtin_elementwise_min(__builtin_elementwise_max(v, as_s2(lo)), as_s2(hi)));
}
__device__ __forceinline__ uint32_t pk_min(uint32_t a, uint32_t b) {
    return as_u32(__builtin_elementwise_min(as_s2(a), as_s2(b)));
}
__device__ __forceinline__ uint32_t pk_max(uint32_t a, uint32_t b) {
    return as_u32(__builtin_elementwise_max(as_s2(a), as_s2(b)));
}
__device__ __forceinline__ uint32_t pk_update_lo(uint32_t a, uint32_t b, uint32_t lo) {
    return as_u32(__builtin_elementwise_max(__builtin_elementwise_add_sat(as_s2(a), as_s2(b)), as_s2(lo)));
}
#endif

// Packed message words of base graph BG: ceil(deg / 2) per row.
template <int BG>
__host__ __device__ constexpr int packed_word0(int r) {
    int w = 0;
    for (int x = 0; x < r; ++x) w += (Graph<BG>::start(x + 1) - Graph<BG>::start(x) + 1) / 2;
    return w;
}

// Barrier groups as index ranges: group g holds rows first[g] .. first[g+1]-1.
template <int BG>
struct GroupPlan {
    int group_of[Graph<BG>::kRows];
    int first[Graph<BG>::kRows + 1];
    int groups, max_rows, max_pairs;
    constexpr GroupPlan() : group_of(), first(), groups(0), max_rows(0), max_pairs(0) {
        int g = 0;
        first[0] = 0;
        for (int r = 0; r < Graph<BG>::kRows; ++r) {
            group_of[r] = g;
            const int pr = (Graph<BG>::start(r + 1) - Graph<BG>::start(r) + 1) / 2;
            max_pairs = pr > max_pairs ? pr : max_pairs;
            if (kRowGroups<BG>.barrier_after[r]) {
                ++g;
                first[g] = r + 1;
                max_rows = first[g] - first[g - 1] > max_rows ? first[g] - first[g - 1] : max_rows;
            }
        }
        groups = g;
    }
};
template <int BG> inline constexpr GroupPlan<BG> kGroupPlan{};

// Edge j of row r reads a column that no row of the previous barrier group
// writes (groups after the first of an iteration): its belief is final once
// that group's barrier is reached, so it may be loaded before the barrier.
template <int BG>
__host__ __device__ constexpr bool edge_early(int r, int j) {
    const int g = kGroupPlan<BG>.group_of[r];
    if (g == 0) return false;
    const int c = Graph<BG>::col(Graph<BG>::start(r) + j);
    for (int x = kGroupPlan<BG>.first[g - 1]; x < kGroupPlan<BG>.first[g]; ++x)
        for (int f = Graph<BG>::start(x); f < Graph<BG>::start(x + 1); ++f)
            if (Graph<BG>::col(f) == c) return false;
    return true;
}
// Late (not early) edges of pair i of row r: 0, 1 or 2.
template <int BG>
__host__ __device__ constexpr int pair_late(int r, int i) {
    const int deg = Graph<BG>::start(r + 1) - Graph<BG>::start(r);
    return (edge_early<BG>(r, 2 * i) ? 0 : 1) + (2 * i + 1 < deg && !edge_early<BG>(r, 2 * i + 1) ? 1 : 0);
}
// k-th pair of row r in first-pass order: pairs with fewer late edges first.
template <int BG>
__host__ __device__ constexpr int pair_order(int r, int k) {
    const int pairs = (Graph<BG>::start(r + 1) - Graph<BG>::start(r) + 1) / 2;
    for (int late = 0; late <= 2; ++late)
        for (int i = 0; i < pairs; ++i)
            if (pair_late<BG>(r, i) == late && k-- == 0) return i;
    return -1;
}

template <int PAIRS>
struct PackedRowState {
    uint32_t oa[PAIRS], ob[PAIRS];   // byte offsets of the pair's two variables
    uint32_t qa[PAIRS], qb[PAIRS];   // the pair's two beliefs as loaded (uint16)
    uint32_t rel[PAIRS];             // extrinsic inputs, int16 halves
    uint32_t mag[PAIRS];             // |rel| capped at 32767
    uint32_t m1, m2;                 // per half: smallest and second smallest |rel|
    uint32_t px;
};

// Offsets of pair I (edges E, E + 1) of a row.
template <int BG, int ZC, int E, int I, int DEG, int PAIRS>
__device__ __forceinline__ void packed_pair_addr(uint32_t zx2, uint32_t zz, PackedRowState<PAIRS> &rs) {
    using F = FixedCode<BG, ZC>;
    constexpr bool has_b = 2 * I + 1 < DEG;
    constexpr uint32_t sha = F::shift_bytes(E);
    if constexpr (has_b && sha != 0 && F::shift_bytes(E + 1) != 0) {
        // both offsets at once: zz holds 2*zx in each half, so t holds
        // 2*zx + shift bytes per half (< 2 * 2Zc: no carry), u the same less
        // 2Zc bytes per half, and each offset is the smaller of its halves
        constexpr uint32_t shb = F::shift_bytes(E + 1);
        constexpr uint32_t wrap = (uint32_t)(uint16_t)(-2 * ZC) * 0x10001u;
        const uint32_t t = zz + (sha | (shb << 16));
        const uint32_t u = pk_addw(t, wrap);
        rs.oa[I] = min(t & 0xFFFFu, u & 0xFFFFu);
        rs.ob[I] = min(t >> 16, u >> 16);
    } else {
        // per-edge offsets; a zero shift needs none (running lanes have zx < Zc)
        rs.oa[I] = sha == 0 ? zx2 : min(zx2 + sha, zx2 + sha - 2u * (uint32_t)ZC);
        if constexpr (has_b) {
            constexpr uint32_t shb = F::shift_bytes(E + 1);
            rs.ob[I] = shb == 0 ? zx2 : min(zx2 + shb, zx2 + shb - 2u * (uint32_t)ZC);
        }
    }
}

// Loads of pair I's edges whose early flag equals kEarly.
template <int BG, int ZC, int R, int E, int I, int DEG, int PAIRS, bool kEarly>
__device__ __forceinline__ void packed_pair_load(const lds_char *__restrict__ lqb,
                                                 PackedRowState<PAIRS> &rs) {
    using F = FixedCode<BG, ZC>;
    if constexpr (edge_early<BG>(R, 2 * I) == kEarly)
        rs.qa[I] = (uint32_t)(uint16_t)*(const lds_i16 *)(lqb + F::col_bytes(E) + rs.oa[I]);
    if constexpr (2 * I + 1 < DEG) {
        if constexpr (edge_early<BG>(R, 2 * I + 1) == kEarly)
            rs.qb[I] = (uint32_t)(uint16_t)*(const lds_i16 *)(lqb + F::col_bytes(E + 1) + rs.ob[I]);
    }
}

// The first pass keeps the two smallest magnitudes per half; no edge index is
// needed.  An edge's outgoing magnitude is the scaled minimum over the other
// edges: f(m2) for an edge with |rel| = m1 (a tie makes m2 = m1), f(m1) for
// every other edge, whose |rel| is then at least m2.  The fast-row scale f
// (Q15 alpha in [0, 32767] after the offset and the max_llr cap) is
// non-decreasing with f(m2) - f(m1) <= m2 - m1, so both cases are
//     max((f(m2) + m1) - |rel|, f(m1)),
// whose exact value lies in [-32767, 32767]: int16 wrap arithmetic is exact.
// K is the pair's place in the first pass (min and max commute).
template <int K, int I, int DEG, int PAIRS>
__device__ __forceinline__ void packed_pair_in(const uint32_t *msg, int w, PackedRowState<PAIRS> &rs) {
    constexpr bool has_b = 2 * I + 1 < DEG;
    const uint32_t q = has_b ? pk_join(rs.qa[I], rs.qb[I]) : rs.qa[I];
    const uint32_t rel = pk_subw(q, msg[w]);
    rs.rel[I] = rel;
    rs.px = K == 0 ? rel : (rs.px ^ rel);
    uint32_t a = pk_abs_sat(rel);
    // an odd last edge: the unused high half is at least every magnitude
    if constexpr (!has_b) a |= 0x7FFF0000u;
    rs.mag[I] = a;
    if constexpr (K == 0) {
        rs.m1 = a;
    } else if constexpr (K == 1) {
        rs.m2 = pk_max(rs.m1, a);
        rs.m1 = pk_min(rs.m1, a);
    } else {
        rs.m2 = pk_min(rs.m2, pk_max(rs.m1, a));
        rs.m1 = pk_min(rs.m1, a);
    }
}

// A row's scaled minima and parity, packed for the second pass.
struct PackedScale {
    uint32_t cp;    // f(m2) + m1 in both halves
    uint32_t n1p;   // f(m1) in both halves
    uint32_t pm;    // row parity in the sign bit of both halves
    uint32_t lo, hi;
};

// kSimple: DecParams::simple (beta 0, max_llr 32767; the capped |rel| keeps
// every minimum within [0, 32767]); otherwise the general fast-row scale.
template <bool kSimple, int PAIRS>
__device__ __forceinline__ PackedScale packed_finalize(const PackedRowState<PAIRS> &rs,
                                                       const DecParams &p, int alpha, int beta) {
    // merge the halves: m1 = min of the four, m2 = second smallest
    const uint32_t m1l = rs.m1 & 0xFFFFu, m1h = rs.m1 >> 16;
    const int m1 = (int)min(m1l, m1h);
    // the halves' second minima as one 16-bit min (an SDWA op; a min3 over
    // the extracted halves takes three)
#ifdef NRLDPC_CPU_EMU
    const uint32_t m2s = min(rs.m2 & 0xFFFFu, rs.m2 >> 16);
#else
    const uint32_t m2s = __builtin_elementwise_min((uint16_t)rs.m2, (uint16_t)(rs.m2 >> 16));
#endif
    const int m2 = (int)min(max(m1l, m1h), m2s);
    int n1, n2;
    if constexpr (kSimple) {
        n1 = scale_simple(m1, alpha);
        n2 = scale_simple(m2, alpha);
    } else {
        n1 = scale_i16(min(m1, p.max_llr), alpha, beta);
        n2 = scale_i16(min(m2, p.max_llr), alpha, beta);
    }
    PackedScale f;
    // odd number of negative inputs: every message flips sign.  px ^ rot16(px)
    // holds that parity in the sign bit of both halves (the other bits are
    // never read: only the sign of rel ^ pm is)
    f.pm = rs.px ^ ((rs.px >> 16) | (rs.px << 16));
    f.n1p = (uint32_t)n1 * 0x10001u;   // n1 in [0, 32767]
    const uint32_t c = (uint32_t)(n2 + m1) & 0xFFFFu;
    f.cp = c | (c << 16);
    f.lo = pk_make(p.min_llr, p.min_llr);
    f.hi = pk_make(p.max_llr, p.max_llr);
    return f;
}

// Second pass of pair I: its new messages (stored) and its two new beliefs
// (returned as int16 halves).
template <int I, int DEG, int PAIRS, bool kSimple, int N>
__device__ __forceinline__ uint32_t packed_pair_update(uint32_t (&msg)[N], int w,
                                                       const PackedRowState<PAIRS> &rs,
                                                       const PackedScale &f) {
    constexpr bool has_b = 2 * I + 1 < DEG;
    uint32_t nv = pk_max(pk_subw(f.cp, rs.mag[I]), f.n1p);
    if constexpr (!has_b) nv &= 0xFFFFu;   // keeps the unused high half at 0
    const uint32_t rel = rs.rel[I];
    // message = magnitude * (+-1): the sign of rel, flipped for odd row parity
    const uint32_t m = pk_mullo(nv, pk_sgn(rel ^ f.pm));
    msg[w] = m;
    // kSimple has max_llr = 32767: the saturated add already clips above
    return kSimple ? pk_update_lo(rel, m, f.lo) : pk_update(rel, m, f.lo, f.hi);
}

template <int BG, int ZC, int E, int I, int DEG, int PAIRS, bool kSimple, int N>
__device__ __forceinline__ void packed_pair_out(lds_char *__restrict__ lqb, uint32_t (&msg)[N],
                                                int w, const PackedRowState<PAIRS> &rs,
                                                const PackedScale &f) {
    using F = FixedCode<BG, ZC>;
    const uint32_t v = packed_pair_update<I, DEG, PAIRS, kSimple>(msg, w, rs, f);
    *(lds_i16 *)(lqb + F::col_bytes(E) + rs.oa[I]) = (int16_t)(v & 0xFFFFu);
    if constexpr (2 * I + 1 < DEG)
        *(lds_i16 *)(lqb + F::col_bytes(E + 1) + rs.ob[I]) = (int16_t)(v >> 16);
}

// Before row R's barrier: every offset of the row and its early loads.
template <int BG, int ZC, int R, int PAIRS, int... I>
__device__ __forceinline__ void packed_pre(const lds_char *__restrict__ lqb, uint32_t zx2, uint32_t zz,
                                           PackedRowState<PAIRS> &rs, std::integer_sequence<int, I...>) {
    constexpr int e0 = Graph<BG>::start(R);
    constexpr int deg = Graph<BG>::start(R + 1) - e0;
    (packed_pair_addr<BG, ZC, e0 + 2 * I, I, deg, PAIRS>(zx2, zz, rs), ...);
    (packed_pair_load<BG, ZC, R, e0 + 2 * I, I, deg, PAIRS, true>(lqb, rs), ...);
}

// After the barrier: the late loads of row R (for every row of the group
// before any row's update: the rows of a group touch disjoint columns).
template <int BG, int ZC, int R, int PAIRS, int... I>
__device__ __forceinline__ void packed_late(const lds_char *__restrict__ lqb, PackedRowState<PAIRS> &rs,
                                            std::integer_sequence<int, I...>) {
    constexpr int e0 = Graph<BG>::start(R);
    constexpr int deg = Graph<BG>::start(R + 1) - e0;
    (packed_pair_load<BG, ZC, R, e0 + 2 * I, I, deg, PAIRS, false>(lqb, rs), ...);
}

// Then the row itself (kSimple as in packed_finalize).
template <int BG, int ZC, int R, bool kSimple, int PAIRS, int N, int... I>
__device__ __forceinline__ void packed_main(lds_char *__restrict__ lqb, uint32_t (&msg)[N],
                                            PackedRowState<PAIRS> &rs, const DecParams &p,
                                            int alpha, int beta, std::integer_sequence<int, I...>) {
    constexpr int e0 = Graph<BG>::start(R);
    constexpr int deg = Graph<BG>::start(R + 1) - e0;
    constexpr int w0 = packed_word0<BG>(R);
    static_assert(sizeof...(I) >= 2, "packed rows need degree >= 3");
    (packed_pair_in<I, pair_order<BG>(R, I), deg, PAIRS>(msg, w0 + pair_order<BG>(R, I), rs), ...);
    const PackedScale f = packed_finalize<kSimple>(rs, p, alpha, beta);
    (packed_pair_out<BG, ZC, e0 + 2 * I, I, deg, PAIRS, kSimple>(lqb, msg, w0 + I, rs, f), ...);
}

template <int BG>
using GroupRowState = PackedRowState<kGroupPlan<BG>.max_pairs>;

template <int BG, int ZC, int R, int REnd>
__device__ __forceinline__ void group_pre(const lds_char *lqb, uint32_t zx2, uint32_t zz,
                                          GroupRowState<BG> *rows) {
    if constexpr (R < REnd) {
        constexpr int deg = Graph<BG>::start(R + 1) - Graph<BG>::start(R);
        packed_pre<BG, ZC, R>(lqb, zx2, zz, *rows, std::make_integer_sequence<int, (deg + 1) / 2>{});
        group_pre<BG, ZC, R + 1, REnd>(lqb, zx2, zz, rows + 1);
    }
}

template <int BG, int ZC, int R, int REnd>
__device__ __forceinline__ void group_late(const lds_char *lqb, GroupRowState<BG> *rows) {
    if constexpr (R < REnd) {
        constexpr int deg = Graph<BG>::start(R + 1) - Graph<BG>::start(R);
        packed_late<BG, ZC, R>(lqb, *rows, std::make_integer_sequence<int, (deg + 1) / 2>{});
        group_late<BG, ZC, R + 1, REnd>(lqb, rows + 1);
    }
}

template <int BG, int ZC, int R, int REnd, bool kTable, bool kSimple, int N>
__device__ __forceinline__ void group_main(lds_char *lqb, uint32_t (&msg)[N], const DecParams &p,
                                           const uint32_t *ab, int alpha0, GroupRowState<BG> *rows) {
    if constexpr (R < REnd) {
        int alpha = alpha0, beta = p.beta;
        if constexpr (kTable) {
            const uint32_t w = ab[R];
            alpha = (int)(int16_t)(w & 0xFFFFu);
            beta = (int)(int16_t)(w >> 16);
        }
        constexpr int deg = Graph<BG>::start(R + 1) - Graph<BG>::start(R);
        packed_main<BG, ZC, R, kSimple>(lqb, msg, *rows, p, alpha, beta,
                                        std::make_integer_sequence<int, (deg + 1) / 2>{});
        group_main<BG, ZC, R + 1, REnd, kTable, kSimple>(lqb, msg, p, ab, alpha0, rows + 1);
    }
}

// Packed rows of one iteration, group by group.  rows holds group G's row
// states (offsets and early loads made before the barrier that opens G);
// before the barrier that closes G, the next group's are made the same way.
template <int BG, int ZC, int G, bool kTable, bool kSimple, int N>
__device__ __forceinline__ void packed_groups(lds_char *lqb, uint32_t (&msg)[N], uint32_t zx2,
                                              uint32_t zz, bool run, const DecParams &p,
                                              const uint32_t *ab, int alpha0,
                                              GroupRowState<BG> *rows) {
    constexpr auto &gp = kGroupPlan<BG>;
    if (run) {
        group_late<BG, ZC, gp.first[G], gp.first[G + 1]>(lqb, rows);
        group_main<BG, ZC, gp.first[G], gp.first[G + 1], kTable, kSimple>(lqb, msg, p, ab, alpha0,
                                                                          rows);
    }
    if constexpr (G + 1 < gp.groups) {
        GroupRowState<BG> next[gp.max_rows];
        if (run) group_pre<BG, ZC, gp.first[G + 1], gp.first[G + 2]>(lqb, zx2, zz, next);
        __syncthreads();
        packed_groups<BG, ZC, G + 1, kTable, kSimple>(lqb, msg, zx2, zz, run, p, ab, alpha0, next);
    } else {
        __syncthreads();
    }
}

template <int BG, int ZC, int R, bool kGeneral, bool kTable, bool kSimple, int N>
__device__ __forceinline__ void fixed_rows(lds_char *lqb, uint32_t (&msg)[N], uint32_t zx2, bool run,
                                           const DecParams &p, const uint32_t *ab, int alpha0) {
    if constexpr (R < Graph<BG>::kRows) {
        int alpha = alpha0, beta = p.beta;
        if constexpr (kTable) {
            const uint32_t w = ab[R];
            alpha = (int)(int16_t)(w & 0xFFFFu);
            beta = (int)(int16_t)(w >> 16);
        }
        constexpr int deg = Graph<BG>::start(R + 1) - Graph<BG>::start(R);
        if (run)
            fixed_row<BG, ZC, R, kGeneral>(lqb, msg, zx2, p, alpha, beta,
                                           std::make_integer_sequence<int, deg>{});
        if constexpr (kRowGroups<BG>.barrier_after[R]) __syncthreads();
        fixed_rows<BG, ZC, R + 1, kGeneral, kTable, kSimple>(lqb, msg, zx2, run, p, ab, alpha0);
    }
}

// Lanes of the specialized kernels: ZC rounded up to whole waves.
template <int ZC>
inline constexpr int kFixedLanes = (ZC + 63) / 64 * 64;

// The specialized kernels' test: beliefs of type T at LDS address 0 (lqb),
// zxb = zx * sizeof(T); one workgroup barrier; every lane gets the result.
template <int BG, int ZC, typename T>
__device__ __forceinline__ uint32_t fixed_crc_linear(const lds_char *lqb, uint32_t zxb,
                                                     const uint32_t *wts, uint32_t *slots) {
#ifdef NRLDPC_CPU_EMU
    typedef T lds_t;
#else
    typedef __attribute__((address_space(3))) T lds_t;
#endif
    constexpr int kLanes = kFixedLanes<ZC>;
    const int zx = threadIdx.x;
    uint32_t acc = 0u;
#pragma unroll
    for (int c = 0; c < Graph<BG>::kInfoCols; ++c) {
        const T q = *(const lds_t *)(lqb + (uint32_t)(sizeof(T) * c * ZC) + zxb);
        acc = crc_lin_add(acc, wts[c * kLanes + zx], crc_hard_mask(q));
    }
    acc = wave_xor_full(acc);
    if ((zx & 63) == 0) slots[zx >> 6] = acc;
    __syncthreads();
    uint32_t total = 0u;
#pragma unroll
    for (int w = 0; w < kLanes / 64; ++w) total ^= slots[w];
    return total;
}

template <int BG, int ZC, bool kGeneral, bool kTable, bool kSimple>
__device__ __forceinline__ int run_fixed_i16(int16_t *lq, const DecParams &p, const uint32_t *wts,
                                             uint32_t *slots, int *status) {
    // exact rows: one message per edge in global edge order; fast rows:
    // row-paired packed messages
    constexpr int kWords =
        kGeneral ? FixedCode<BG, ZC>::kMsgWords : packed_word0<BG>(Graph<BG>::kRows);
    uint32_t msg[kWords];
#pragma unroll
    for (int i = 0; i < kWords; ++i) msg[i] = 0u;
    const int zx = threadIdx.x;
    uint32_t zx2 = 2u * (uint32_t)zx;
    lds_char *lqb = fixed_lds_base(lq);
    int it = 0;
    *status = 1;
    while (it < p.max_it) {
        // packed rows: lane values made opaque once per iteration (keeps the
        // offsets from being hoisted into registers; alpha stays in one VGPR)
        int alpha0 = p.alpha;
        if constexpr (!kGeneral) {
            NRLDPC_OPAQUE_LANE(zx2);
            if constexpr (!kTable) NRLDPC_OPAQUE_LANE(alpha0);
        }
        const bool run = zx < ZC;
        if constexpr (kGeneral) {
            fixed_rows<BG, ZC, 0, kGeneral, kTable, kSimple>(lqb, msg, zx2, run, p,
                                                             scale_row<BG, kTable>(p, it), alpha0);
        } else {
            const uint32_t zz = zx2 | (zx2 << 16);
            GroupRowState<BG> rows[kGroupPlan<BG>.max_rows];
            if (run)
                group_pre<BG, ZC, 0, kGroupPlan<BG>.first[1]>(lqb, zx2, zz, rows);
            packed_groups<BG, ZC, 0, kTable, kSimple>(lqb, msg, zx2, zz, run, p,
                                                      scale_row<BG, kTable>(p, it), alpha0, rows);
        }
        ++it;
        if (fixed_crc_linear<BG, ZC, int16_t>(lqb, 2u * (uint32_t)zx, wts, slots) == 0u) {
            *status = 0;
            break;
        }
    }
    return it;
}

// ---------------------------------------------------------------------------
// Packed rows with run-time shifts: the generic int16 kernel (one codeblock
// of any lifting size per workgroup) and the packed tiles, fast rows only.
// The arithmetic is the specialized kernel's (packed_pair_in,
// packed_finalize with the general scale, packed_pair_update); offsets come
// from the pair table that follows the edge table of the (BG, Zc) block:
// per row pair, (2 * shift_a) | (2 * shift_b) << 16, then the two columns'
// byte offsets (three scalar loads, no scalar decoding).  lqb is the lane's
// codeblock beliefs in LDS, zx2 = 2 * zx, zz = zx2 in both halves,
// z2 = 2 * Zc and wrap = -2 * Zc in both int16 halves.
// ---------------------------------------------------------------------------
struct GenLane {
    char *lqb;
    uint32_t zx2, zz, z2, wrap;
};

// The edge table through the constant address space: uniform indices become
// scalar loads and the column / shift decoding scalar ops.
#ifdef NRLDPC_CPU_EMU
typedef const uint32_t *etab_cptr;
#else
typedef const __attribute__((address_space(4))) uint32_t *etab_cptr;
#endif

// Pair I of a row from its three pair-table words pw[0..2].
template <int I, int DEG, int PAIRS>
__device__ __forceinline__ void gpair_load(const GenLane &g, const uint32_t *pw,
                                           PackedRowState<PAIRS> &rs) {
    const uint32_t sh = pw[0], cola = pw[1];
    if constexpr (2 * I + 1 < DEG) {
        // both offsets at once (see packed_pair_addr): halves below 4 * Zc
        const uint32_t colb = pw[2];
        const uint32_t t = g.zz + sh;
        const uint32_t u = pk_addw(t, g.wrap);
        rs.oa[I] = cola + min(t & 0xFFFFu, u & 0xFFFFu);
        rs.ob[I] = colb + min(t >> 16, u >> 16);
        rs.qa[I] = (uint32_t)(uint16_t)*(const int16_t *)(g.lqb + rs.oa[I]);
        rs.qb[I] = (uint32_t)(uint16_t)*(const int16_t *)(g.lqb + rs.ob[I]);
    } else {
        const uint32_t c = g.zx2 + sh;   // a lone edge: the high half is 0
        rs.oa[I] = cola + min(c, c - g.z2);
        rs.qa[I] = (uint32_t)(uint16_t)*(const int16_t *)(g.lqb + rs.oa[I]);
    }
}

template <int I, int DEG, int PAIRS>
__device__ __forceinline__ void gpair_store(const GenLane &g, const PackedRowState<PAIRS> &rs,
                                            uint32_t v) {
    *(int16_t *)(g.lqb + rs.oa[I]) = (int16_t)(v & 0xFFFFu);
    if constexpr (2 * I + 1 < DEG) *(int16_t *)(g.lqb + rs.ob[I]) = (int16_t)(v >> 16);
}

// Pair-table words of barrier group G (rows first[G] .. first[G+1] - 1),
// held in scalar registers.  A group's words are loaded while the previous
// group runs, so the rows after a barrier find their offsets' scalar
// operands at hand instead of waiting on a scalar load.
template <int BG, int G>
struct GroupWords {
    static constexpr int kP0 = packed_word0<BG>(kGroupPlan<BG>.first[G]);
    static constexpr int kN = 3 * (packed_word0<BG>(kGroupPlan<BG>.first[G + 1]) - kP0);
    uint32_t w[kN];
};

template <int BG, int G>
__device__ __forceinline__ GroupWords<BG, G> load_group_words(etab_cptr ptab) {
    GroupWords<BG, G> gw;
#pragma unroll
    for (int i = 0; i < GroupWords<BG, G>::kN; ++i) gw.w[i] = ptab[3 * GroupWords<BG, G>::kP0 + i];
    return gw;
}

// kSimple as in packed_finalize (DecParams::simple).  pw: the row's pair words.
template <int BG, int R, bool kSimple, int N, int... I>
__device__ __forceinline__ void gpacked_row(const GenLane &g, uint32_t (&msg)[N],
                                            const uint32_t *pw, const DecParams &p, int alpha,
                                            int beta, std::integer_sequence<int, I...>) {
    constexpr int e0 = Graph<BG>::start(R);
    constexpr int deg = Graph<BG>::start(R + 1) - e0;
    constexpr int w0 = packed_word0<BG>(R);
    constexpr int pairs = (int)sizeof...(I);
    static_assert(pairs >= 2, "packed rows need degree >= 3");
    PackedRowState<pairs> rs;
    (gpair_load<I, deg, pairs>(g, pw + 3 * I, rs), ...);
    (packed_pair_in<I, I, deg, pairs>(msg, w0 + I, rs), ...);
    const PackedScale f = packed_finalize<kSimple>(rs, p, alpha, beta);
    (gpair_store<I, deg, pairs>(g, rs,
                                packed_pair_update<I, deg, pairs, kSimple>(msg, w0 + I, rs, f)),
     ...);
}

// Rows R .. REnd - 1 of group G (column-disjoint: no barrier between them).
template <int BG, int G, int R, int REnd, bool kTable, bool kSimple, int N>
__device__ __forceinline__ void gpgroup_rows(const GenLane &g, uint32_t (&msg)[N],
                                             const GroupWords<BG, G> &gw, const DecParams &p,
                                             const uint32_t *ab) {
    if constexpr (R < REnd) {
        int alpha = p.alpha, beta = p.beta;
        if constexpr (kTable) {
            const uint32_t w = ab[R];
            alpha = (int)(int16_t)(w & 0xFFFFu);
            beta = (int)(int16_t)(w >> 16);
        }
        constexpr int deg = Graph<BG>::start(R + 1) - Graph<BG>::start(R);
        constexpr int off = 3 * (packed_word0<BG>(R) - GroupWords<BG, G>::kP0);
        gpacked_row<BG, R, kSimple>(g, msg, gw.w + off, p, alpha, beta,
                                    std::make_integer_sequence<int, (deg + 1) / 2>{});
        gpgroup_rows<BG, G, R + 1, REnd, kTable, kSimple>(g, msg, gw, p, ab);
    }
}

// Group G of one iteration, then (recursively) the groups after it; gw holds
// group G's words.  Lanes with run = false only take part in the barriers.
template <int BG, int G, bool kTable, bool kSimple, int N>
__device__ __forceinline__ void gpgroups(const GenLane &g, uint32_t (&msg)[N], etab_cptr ptab,
                                         bool run, const DecParams &p, const uint32_t *ab,
                                         GroupWords<BG, G> &gw) {
    constexpr auto &gp = kGroupPlan<BG>;
    if constexpr (G + 1 < gp.groups) {
        GroupWords<BG, G + 1> next = load_group_words<BG, G + 1>(ptab);
        if (run) gpgroup_rows<BG, G, gp.first[G], gp.first[G + 1], kTable, kSimple>(g, msg, gw, p, ab);
#pragma unroll
        for (int i = 0; i < GroupWords<BG, G + 1>::kN; ++i) NRLDPC_OPAQUE_UNIFORM(next.w[i]);
        __syncthreads();
        gpgroups<BG, G + 1, kTable, kSimple>(g, msg, ptab, run, p, ab, next);
    } else {
        if (run) gpgroup_rows<BG, G, gp.first[G], gp.first[G + 1], kTable, kSimple>(g, msg, gw, p, ab);
        __syncthreads();
    }
}

// Lane setup of the run-time packed rows.
__device__ __forceinline__ GenLane gen_lane(int16_t *lq, int zx, int Zc) {
    GenLane g;
    g.lqb = reinterpret_cast<char *>(lq);
    g.zx2 = 2u * (uint32_t)zx;
    g.zz = g.zx2 | (g.zx2 << 16);
    g.z2 = 2u * (uint32_t)Zc;
    g.wrap = (uint32_t)(uint16_t)(-2 * Zc) * 0x10001u;
    return g;
}

// One iteration's rows, with the lane's offsets and the edge table made
// opaque first: otherwise their loop-invariant offsets would be hoisted out
// of the iteration loop into hundreds of registers.
template <int BG, bool kTable, bool kSimple, int N>
__device__ __forceinline__ void gprows_iteration(GenLane g, uint32_t (&msg)[N], const uint32_t *etab,
                                                 bool run, const DecParams &p, const uint32_t *ab) {
    NRLDPC_OPAQUE_LANE(g.zx2);
    NRLDPC_OPAQUE_LANE(g.zz);
    etab_cptr ce = (etab_cptr)etab;
    NRLDPC_OPAQUE_UNIFORM(ce);
    const etab_cptr ptab = ce + Graph<BG>::kEdges;
    GroupWords<BG, 0> first = load_group_words<BG, 0>(ptab);
    gpgroups<BG, 0, kTable, kSimple>(g, msg, ptab, run, p, ab, first);
}

template <int BG>
constexpr int kPackedWords = packed_word0<BG>(Graph<BG>::kRows);

template <int BG, bool kGeneral, bool kTable>
__device__ __forceinline__ int run_i16(int16_t *lq, const uint32_t *etab, const CbDesc &d, const DecParams &p,
                       const uint32_t *wts, uint32_t *slots, int *status) {
    const int zx = threadIdx.x;
    int it = 0;
    *status = 1;
    // exact rows: compressed messages; fast rows: row-paired packed messages
    StateI16<BG> st;
    uint32_t msg[kGeneral ? 1 : kPackedWords<BG>];
    if constexpr (kGeneral) {
        init_state<BG>(st);
    } else {
#pragma unroll
        for (int i = 0; i < kPackedWords<BG>; ++i) msg[i] = 0u;
    }
    const GenLane g = gen_lane(lq, zx, d.Zc);
    while (it < p.max_it) {
        if constexpr (kGeneral)
            rows_i16<BG, 0, kGeneral, kTable>(lq, st, etab, zx, d.Zc, zx < d.Zc, p,
                                              scale_row<BG, kTable>(p, it));
        else
            gprows_iteration<BG, kTable, false>(g, msg, etab, zx < d.Zc, p,
                                                scale_row<BG, kTable>(p, it));
        ++it;
        if (block_crc_linear<BG>(lq, d.Zc, wts, slots) == 0u) {
            *status = 0;
            break;
        }
    }
    return it;
}

// ---------------------------------------------------------------------------
// Algo 4: FP32 min-sum (c_lib/floatpoint_ldpc_decoder.c:153-220).  Single
// IEEE operations only (built with -ffp-contract=off): bit-exact with the
// reference's FP32 arithmetic.
// ---------------------------------------------------------------------------
template <int BG>
struct StateF32 {
    float n1[Graph<BG>::kRows];
    float n2[Graph<BG>::kRows];
    uint32_t sgn[Graph<BG>::kRows];   // sign bit per edge | argmin << 24
};

// v negated when bit j of bits is set: IEEE negation is a flip of the sign
// bit, so one shift and one bitwise op give exactly -v or v.
__device__ __forceinline__ float flip_sign(float v, uint32_t bits, int j) {
    return __builtin_bit_cast(float, __builtin_bit_cast(uint32_t, v) ^ ((bits << (31 - j)) & 0x80000000u));
}

// FP32 belief at byte offset a from a generic or an LDS base pointer.
__device__ __forceinline__ float f32_at(const char *b, uint32_t a) {
    return *reinterpret_cast<const float *>(b + a);
}
__device__ __forceinline__ void f32_put(char *b, uint32_t a, float v) {
    *reinterpret_cast<float *>(b + a) = v;
}
#ifndef NRLDPC_CPU_EMU
typedef __attribute__((address_space(3))) float lds_f32;
__device__ __forceinline__ float f32_at(const lds_char *b, uint32_t a) {
    return *(const lds_f32 *)(b + a);
}
__device__ __forceinline__ void f32_put(lds_char *b, uint32_t a, float v) { *(lds_f32 *)(b + a) = v; }
#endif

// The row update on beliefs at byte offsets addr[j] from lqb (a generic or
// an LDS pointer to char).
template <int BG, int R, typename CharPtr>
__device__ __forceinline__ void row_f32ms_at(CharPtr lqb, const uint32_t *addr, StateF32<BG> &st,
                                             const DecParams &p) {
    constexpr int e0 = Graph<BG>::start(R);
    constexpr int deg = Graph<BG>::start(R + 1) - e0;
    const float o1 = st.n1[R], o2 = st.n2[R];
    const uint32_t sb = st.sgn[R];
    const int oidx = (int)(sb >> 24);
    float rel[deg];
#pragma unroll
    for (int j = 0; j < deg; ++j) {
        const float ov = (j == oidx) ? o2 : o1;
        rel[j] = f32_at(lqb, addr[j]) - flip_sign(ov, sb, j);
    }
    float m1 = 3.402823466e+38f, m2 = 3.402823466e+38f;
    int idx = 0;
    uint32_t negs = 0;
#pragma unroll
    for (int j = 0; j < deg; ++j) {
        const bool neg = rel[j] < 0.0f;
        negs |= (uint32_t)neg << j;
        const float a = fabsf(rel[j]);
        const bool lt = a < m1;
        m2 = lt ? m1 : (a < m2 ? a : m2);
        idx = lt ? j : idx;
        m1 = lt ? a : m1;
    }
    if (m1 > p.fmax) m1 = p.fmax;
    if (m1 < 0.0f) m1 = 0.0f;
    if (m2 > p.fmax) m2 = p.fmax;
    if (m2 < 0.0f) m2 = 0.0f;
    float off1 = m1 - p.fbeta, off2 = m2 - p.fbeta;
    if (off1 < 0.0f) off1 = 0.0f;
    if (off2 < 0.0f) off2 = 0.0f;
    const float n1 = p.falpha * off1, n2 = p.falpha * off2;
    const uint32_t mask = (1u << deg) - 1u;
    const uint32_t bits = ((__popc(negs) & 1) ? ~negs : negs) & mask;
#pragma unroll
    for (int j = 0; j < deg; ++j) {
        const float nv = (j == idx) ? n2 : n1;
        float v = rel[j] + flip_sign(nv, bits, j);
        if (v > p.fmax) v = p.fmax;
        if (v < p.fmin) v = p.fmin;
        f32_put(lqb, addr[j], v);
    }
    st.n1[R] = n1;
    st.n2[R] = n2;
    st.sgn[R] = bits | ((uint32_t)idx << 24);
}

// Row R with run-time shifts from the edge table.
template <int BG, int R>
__device__ __forceinline__ void row_f32ms(float *__restrict__ lq, StateF32<BG> &st, etab_cptr etab,
                                          int zx, int Zc, const DecParams &p) {
    constexpr int e0 = Graph<BG>::start(R);
    constexpr int deg = Graph<BG>::start(R + 1) - e0;
    uint32_t addr[deg];
#pragma unroll
    for (int j = 0; j < deg; ++j) {
        const uint32_t t = etab[e0 + j];
        int c = zx + (int)(t >> 16);
        c = (c >= Zc) ? c - Zc : c;
        addr[j] = 4u * ((t & 0xFFFFu) + (uint32_t)c);
    }
    row_f32ms_at<BG, R>(reinterpret_cast<char *>(lq), addr, st, p);
}

// Rows with the barrier after each group of column-disjoint rows only
// (RowGroups): such rows touch disjoint beliefs, so the result is the
// row-serial one.
template <int BG, int R>
__device__ __forceinline__ void rows_f32ms(float *lq, StateF32<BG> &st, etab_cptr etab,
                                           int zx, int Zc, const DecParams &p) {
    if constexpr (R < Graph<BG>::kRows) {
        if (zx < Zc) row_f32ms<BG, R>(lq, st, etab, zx, Zc, p);
        if constexpr (kRowGroups<BG>.barrier_after[R]) __syncthreads();
        rows_f32ms<BG, R + 1>(lq, st, etab, zx, Zc, p);
    }
}

template <int BG>
__device__ __forceinline__ int run_f32ms(float *lq, const uint32_t *etab, const CbDesc &d, const DecParams &p,
                         const uint32_t *wts, uint32_t *slots, int *status) {
    StateF32<BG> st;
#pragma unroll
    for (int r = 0; r < Graph<BG>::kRows; ++r) {
        st.n1[r] = 0.0f;
        st.n2[r] = 0.0f;
        st.sgn[r] = 255u << 24;
    }
    const int zx = threadIdx.x;
    int it = 0;
    *status = 1;
    while (it < p.max_it) {
        // the edge table read by scalar loads, reloaded each iteration
        etab_cptr ce = (etab_cptr)etab;
        NRLDPC_OPAQUE_UNIFORM(ce);
        rows_f32ms<BG, 0>(lq, st, ce, zx, d.Zc, p);
        ++it;
        if (block_crc_linear<BG>(lq, d.Zc, wts, slots) == 0u) {
            *status = 0;
            break;
        }
    }
    return it;
}


// FP32 min-sum with compile-time shifts (every codeblock of the launch has
// base graph BG and lifting size ZC): row_f32ms_at on LDS beliefs at LDS
// address 0, each edge's byte offset a constant column part (the DS
// instruction's immediate) plus the lane's wrapped shift.
template <int BG, int ZC, int E>
__device__ __forceinline__ uint32_t fixed_f32_off(uint32_t zx4) {
    constexpr uint32_t col = 2u * FixedCode<BG, ZC>::col_bytes(E);
    constexpr uint32_t sh = 2u * FixedCode<BG, ZC>::shift_bytes(E);
    if constexpr (sh == 0)
        return col + zx4;
    else
        return col + min(zx4 + sh, zx4 + sh - 4u * (uint32_t)ZC);
}

template <int BG, int ZC, int R, int... J>
__device__ __forceinline__ void fixed_row_f32ms(lds_char *lqb, StateF32<BG> &st, uint32_t zx4,
                                                const DecParams &p,
                                                std::integer_sequence<int, J...>) {
    constexpr int e0 = Graph<BG>::start(R);
    const uint32_t addr[] = {fixed_f32_off<BG, ZC, e0 + J>(zx4)...};
    row_f32ms_at<BG, R>(lqb, addr, st, p);
}

template <int BG, int ZC, int R>
__device__ __forceinline__ void fixed_rows_f32ms(lds_char *lqb, StateF32<BG> &st, uint32_t zx4,
                                                 bool run, const DecParams &p) {
    if constexpr (R < Graph<BG>::kRows) {
        constexpr int deg = Graph<BG>::start(R + 1) - Graph<BG>::start(R);
        if (run) fixed_row_f32ms<BG, ZC, R>(lqb, st, zx4, p, std::make_integer_sequence<int, deg>{});
        if constexpr (kRowGroups<BG>.barrier_after[R]) __syncthreads();
        fixed_rows_f32ms<BG, ZC, R + 1>(lqb, st, zx4, run, p);
    }
}

template <int BG, int ZC>
__device__ __forceinline__ int run_fixed_f32ms(float *lq, const DecParams &p, const uint32_t *wts,
                                               uint32_t *slots, int *status) {
    StateF32<BG> st;
#pragma unroll
    for (int r = 0; r < Graph<BG>::kRows; ++r) {
        st.n1[r] = 0.0f;
        st.n2[r] = 0.0f;
        st.sgn[r] = 255u << 24;
    }
    const int zx = threadIdx.x;
    uint32_t zx4 = 4u * (uint32_t)zx;
    lds_char *lqb = fixed_lds_base(reinterpret_cast<int16_t *>(lq));
    int it = 0;
    *status = 1;
    while (it < p.max_it) {
        NRLDPC_OPAQUE_LANE(zx4);
        fixed_rows_f32ms<BG, ZC, 0>(lqb, st, zx4, zx < ZC, p);
        ++it;
        if (fixed_crc_linear<BG, ZC, float>(lqb, 4u * (uint32_t)zx, wts, slots) == 0u) {
            *status = 0;
            break;
        }
    }
    return it;
}

// ---------------------------------------------------------------------------
// Lane-per-edge schedule (algo 1, fast rows, small lifting sizes): one wave
// per codeblock, and each layer (base-graph row) spread over the wave one
// edge per lane -- lane l holds edge j = l / Zc of lifted check zx = l % Zc
// (deg * Zc <= 64: Zc <= 3 for BG1, <= 6 for BG2).  A row is one belief
// load per lane and a tree reduction of (min1, min2, sign parity) over the
// check's lanes, which sit Zc apart, through LDS; the arithmetic is the
// packed fast rows' (packed_pair_in / packed_finalize / packed_pair_update),
// so results are the reference's (c_lib/fixedpoint_ldpc_decoder.c:103-144)
// whenever DecParams::general is 0.
// ---------------------------------------------------------------------------
template <int BG>
inline constexpr int kEdgeLaneMaxZc =
    64 / (BG == 1 ? NR_BG1_MAX_DEGREE : NR_BG2_MAX_DEGREE);

// Two smallest magnitudes (15-bit halves) and the sign parity (bit 31) of
// two partial sets.
__device__ __forceinline__ uint32_t lpe_combine(uint32_t a, uint32_t b) {
    const uint32_t a1 = a & 0x7FFFu, a2 = (a >> 16) & 0x7FFFu;
    const uint32_t b1 = b & 0x7FFFu, b2 = (b >> 16) & 0x7FFFu;
    const uint32_t m1 = min(a1, b1);
    const uint32_t m2 = min(max(a1, b1), min(a2, b2));
    return m1 | (m2 << 16) | ((a ^ b) & 0x80000000u);
}

template <int BG, int R>
__device__ __forceinline__ void lpe_row(int16_t *lq, int &msg, int addr, int j, int zx, int Zc,
                                        uint32_t *keys, const DecParams &p) {
    constexpr int deg = Graph<BG>::start(R + 1) - Graph<BG>::start(R);
    const int lane = threadIdx.x;
    const bool act = j < deg;
    int rel = 0, a = 0x7FFF;
    uint32_t sb = 0;
    if (act) {
        rel = wrap16((int)lq[addr] - msg);
        a = rel == -32768 ? 32767 : (rel < 0 ? -rel : rel);
        sb = rel < 0 ? 1u : 0u;
    }
    uint32_t key = (uint32_t)a | (0x7FFFu << 16) | (sb << 31);
#pragma unroll
    for (int st = 1; st < deg; st <<= 1) {
        keys[lane] = key;
        __syncthreads();
        if (act && (j & (2 * st - 1)) == 0 && j + st < deg)
            key = lpe_combine(key, keys[lane + st * Zc]);
        __syncthreads();
    }
    keys[lane] = key;   // lane zx (j = 0) holds its check's result
    __syncthreads();
    const uint32_t tot = keys[zx];
    if (act) {
        const int m1 = (int)(tot & 0x7FFFu), m2 = (int)((tot >> 16) & 0x7FFFu);
        int n1, n2;
        if (p.simple) {
            n1 = scale_simple(m1, p.alpha);
            n2 = scale_simple(m2, p.alpha);
        } else {
            n1 = scale_i16(min(m1, p.max_llr), p.alpha, p.beta);
            n2 = scale_i16(min(m2, p.max_llr), p.alpha, p.beta);
        }
        const int mag = a == m1 ? n2 : n1;
        const int m = ((sb ^ (tot >> 31)) & 1u) ? -mag : mag;
        int v = rel + m;   // the saturating int16 add
        v = v > 32767 ? 32767 : (v < -32768 ? -32768 : v);
        v = max(v, p.min_llr);
        if (!p.simple) v = min(v, p.max_llr);
        lq[addr] = (int16_t)v;
        msg = m;
    }
    __syncthreads();   // the row's beliefs and keys are read before the next row writes
}

template <int BG, int R>
__device__ __forceinline__ void lpe_rows(int16_t *lq, int (&msg)[Graph<BG>::kRows],
                                         const int (&addr)[Graph<BG>::kRows], int j, int zx,
                                         int Zc, uint32_t *keys, const DecParams &p) {
    if constexpr (R < Graph<BG>::kRows) {
        lpe_row<BG, R>(lq, msg[R], addr[R], j, zx, Zc, keys, p);
        lpe_rows<BG, R + 1>(lq, msg, addr, j, zx, Zc, keys, p);
    }
}

template <int BG>
__device__ __forceinline__ int run_lpe_i16(int16_t *lq, const uint32_t *etab, const CbDesc &d,
                           const DecParams &p, const CrcCtx &crc, const uint32_t *tab,
                           uint32_t *slots, int *status) {
    __shared__ uint32_t keys[64];
    constexpr int nR = Graph<BG>::kRows;
    const int Zc = d.Zc, lane = threadIdx.x;
    const int j = lane / Zc, zx = lane - j * Zc;
    // each row's belief index of this lane's edge (lanes past the row's
    // edges keep 0; they take part in the barriers only)
    int addr[nR], msg[nR];
#pragma unroll
    for (int r = 0; r < nR; ++r) {
        const int deg = Graph<BG>::start(r + 1) - Graph<BG>::start(r);
        addr[r] = 0;
        msg[r] = 0;
        if (j < deg) {
            const uint32_t t = etab[Graph<BG>::start(r) + j];
            int c = zx + (int)(t >> 16);
            c = c >= Zc ? c - Zc : c;
            addr[r] = (int)(t & 0xFFFFu) + c;
        }
    }
    int it = 0;
    *status = 1;
    while (it < p.max_it) {
        lpe_rows<BG, 0>(lq, msg, addr, j, zx, Zc, keys, p);
        ++it;
        if (block_crc(lq, crc, tab, slots) == 0u) {
            *status = 0;
            break;
        }
    }
    return it;
}

// ---------------------------------------------------------------------------
// Kernels
// ---------------------------------------------------------------------------
// kMinSumI16Edge: algo 1 in the lane-per-edge schedule (run_lpe_i16).
enum class Algo { kMinSumI16 = 1, kMinSumF32 = 4, kMinSumI16Edge = 5 };


// kGeneral selects the exact-for-every-row int16 path (see row_i16) and
// kTable the per-row scale words of algos 2/3; the other algorithms ignore
// both.
template <Algo A, typename T, bool kGeneral = false, bool kTable = false>
__global__ void __launch_bounds__(kMaxThreads)
ldpc_decode_kernel(DecOut o, DecParams p) {
    NRLDPC_DYNAMIC_LDS(smem);
    __shared__ uint32_t crc_tab[256];
    __shared__ uint32_t slots[kMaxThreads / 64];
    T *lq = reinterpret_cast<T *>(smem);

    const CbDesc d = o.desc[blockIdx.x];
    const int n = d.N + 2 * d.Zc;
    const int t = threadIdx.x, nt = blockDim.x;

    // fused rate recovery + HARQ combine into LDS; soft buffer write-back.
    // The min-sum kernels' CRC bit weights follow this codeblock's beliefs in
    // the dynamic LDS (the launch adds kGenericWtsCols * blockDim.x words).
    const int16_t *llr = o.llr + d.llr_off;
    int16_t *soft = (d.soft_off != kNone) ? o.soft + d.soft_off : nullptr;
    uint32_t *wts = reinterpret_cast<uint32_t *>(smem + (((size_t)n * sizeof(T) + 15) & ~(size_t)15));
    if constexpr (A == Algo::kMinSumI16) {
        if (d.bgn == 1) {
            CrcLinRegs<Graph<1>::kInfoCols> wr;
            recover_i16<kScatterBatch>(lq, d, llr, soft, p.filler, t, nt, n, true,
                                       [&]() { crc_lin_load(wr, o, d, t, d.Zc, t < d.Zc); }, true);
            crc_lin_store(wr, wts + t, nt);
        } else {
            CrcLinRegs<Graph<2>::kInfoCols> wr;
            recover_i16<kScatterBatch>(lq, d, llr, soft, p.filler, t, nt, n, true,
                                       [&]() { crc_lin_load(wr, o, d, t, d.Zc, t < d.Zc); }, true);
            crc_lin_store(wr, wts + t, nt);
        }
    } else if constexpr (A == Algo::kMinSumI16Edge) {
        recover_i16(lq, d, llr, soft, p.filler, t, nt, n, true);
    } else {
        if (d.bgn == 1)
            crc_lin_weights<Graph<1>::kInfoCols>(wts + t, nt, o, d, t, d.Zc, t < d.Zc);
        else
            crc_lin_weights<Graph<2>::kInfoCols>(wts + t, nt, o, d, t, d.Zc, t < d.Zc);
        recover_fill<T>(lq, d, llr, soft, p.filler, t, nt, n);
    }
    // the lane-per-edge schedule keeps the byte-form test (its lanes are
    // edges, not lifted checks)
    CrcCtx crc{};
    if constexpr (A == Algo::kMinSumI16Edge) crc = crc_setup_tab(d.crctype, d.B / 8, crc_tab, o.crc);
    __syncthreads();

    const uint32_t *etab = o.etab + d.etab_off;
    int status = 1, it = 0;
    if (p.max_it > 0) {
        if constexpr (A == Algo::kMinSumI16) {
            it = (d.bgn == 1)
                     ? run_i16<1, kGeneral, kTable>(lq, etab, d, p, wts, slots, &status)
                     : run_i16<2, kGeneral, kTable>(lq, etab, d, p, wts, slots, &status);
        } else if constexpr (A == Algo::kMinSumF32) {
            it = (d.bgn == 1) ? run_f32ms<1>(lq, etab, d, p, wts, slots, &status)
                              : run_f32ms<2>(lq, etab, d, p, wts, slots, &status);
        } else if constexpr (A == Algo::kMinSumI16Edge) {
            it = (d.bgn == 1) ? run_lpe_i16<1>(lq, etab, d, p, crc, crc_tab, slots, &status)
                              : run_lpe_i16<2>(lq, etab, d, p, crc, crc_tab, slots, &status);
        }
        // hard decisions of the final iteration (left untouched when no
        // iteration ran, like the reference)
        for (int b = t; b < d.B / 8; b += nt) o.hard[d.hard_off + b] = (uint8_t)hard_byte(lq, b);
    }
    if (t == 0) {
        o.status[blockIdx.x] = status;
        o.iters[blockIdx.x] = it;
    }
    if (d.lq_off != kNone) {
        T *out = reinterpret_cast<T *>(o.lq) + d.lq_off;
        for (int i = t; i < n; i += nt) out[i] = lq[i];
    }
}

// Lifting-size specialized int16 decode kernel: one codeblock of base graph
// BG and lifting size ZC per workgroup (run_fixed_i16).  All of its LDS is
// dynamic with the beliefs at offset 0, so every belief access is a DS
// instruction with an immediate column offset: LQ (N + 2Zc int16), then the
// per-wave CRC slots and the lanes' CRC bit weights (fixed_crc_linear).
template <int BG, int ZC>
struct FixedLds {
    static constexpr int kBeliefs = Graph<BG>::kCols * ZC;           // N + 2Zc
    static constexpr int kSlotsOff = ((kBeliefs * 2 + 15) / 16) * 16;
    static constexpr int kWtsOff = kSlotsOff + 64;
    static constexpr int kBytes = kWtsOff + Graph<BG>::kInfoCols * kFixedLanes<ZC> * 4;
};

template <int BG, int ZC, bool kGeneral, bool kTable, bool kSimple = false>
__global__ void __launch_bounds__(kMaxThreads)
ldpc_decode_fixed_kernel(DecOut o, DecParams p) {
    using LdsL = FixedLds<BG, ZC>;
    constexpr int kLanes = kFixedLanes<ZC>;
    NRLDPC_DYNAMIC_LDS(smem);
    int16_t *lq = reinterpret_cast<int16_t *>(smem);
    uint32_t *slots = reinterpret_cast<uint32_t *>(smem + LdsL::kSlotsOff);
    uint32_t *wts = reinterpret_cast<uint32_t *>(smem + LdsL::kWtsOff);

    const CbDesc d = o.desc[blockIdx.x];
    constexpr int n = LdsL::kBeliefs;
    const int t = threadIdx.x, nt = blockDim.x;
    const int16_t *llr = o.llr + d.llr_off;
    int16_t *soft = (d.soft_off != kNone) ? o.soft + d.soft_off : nullptr;
    // CRC bit weights first (their own LDS region): the loads overlap the
    // recovery, whose barriers publish them.  Bit i of the B-bit block has
    // weight x^(B - 1 - i) mod g.
    // CRC bit weights: loaded with the recovery's LLR loads, stored after it
    CrcLinRegs<Graph<BG>::kInfoCols> wr;
    recover_i16<kScatterBatch>(lq, d, llr, soft, p.filler, t, nt, n, true,
                               [&]() { crc_lin_load(wr, o, d, t, ZC, t < ZC); }, true);
    crc_lin_store(wr, wts + t, kLanes);
    __syncthreads();

    int status = 1, it = 0;
    if (p.max_it > 0) {
        it = run_fixed_i16<BG, ZC, kGeneral, kTable, kSimple>(lq, p, wts, slots, &status);
        for (int b = t; b < d.B / 8; b += nt) o.hard[d.hard_off + b] = (uint8_t)hard_byte(lq, b);
    }
    if (t == 0) {
        o.status[blockIdx.x] = status;
        o.iters[blockIdx.x] = it;
    }
    if (d.lq_off != kNone) {
        int16_t *out = reinterpret_cast<int16_t *>(o.lq) + d.lq_off;
        for (int i = t; i < n; i += nt) out[i] = lq[i];
    }
}

// Lifting-size specialized FP32 min-sum kernel (algo 4): one codeblock of
// base graph BG and lifting size ZC per workgroup, beliefs as floats at LDS
// address 0 (all of its LDS is dynamic), then the CRC slots and bit weights.
template <int BG, int ZC>
struct FixedLdsF32 {
    static constexpr int kBeliefs = Graph<BG>::kCols * ZC;           // N + 2Zc
    static constexpr int kSlotsOff = ((kBeliefs * 4 + 15) / 16) * 16;
    static constexpr int kWtsOff = kSlotsOff + 64;
    static constexpr int kBytes = kWtsOff + Graph<BG>::kInfoCols * kFixedLanes<ZC> * 4;
};

template <int BG, int ZC>
__global__ void __launch_bounds__(kMaxThreads)
ldpc_decode_fixed_f32_kernel(DecOut o, DecParams p) {
    using LdsL = FixedLdsF32<BG, ZC>;
    NRLDPC_DYNAMIC_LDS(smem);
    float *lq = reinterpret_cast<float *>(smem);
    uint32_t *slots = reinterpret_cast<uint32_t *>(smem + LdsL::kSlotsOff);
    uint32_t *wts = reinterpret_cast<uint32_t *>(smem + LdsL::kWtsOff);

    const CbDesc d = o.desc[blockIdx.x];
    constexpr int n = LdsL::kBeliefs;
    const int t = threadIdx.x, nt = blockDim.x;
    const int16_t *llr = o.llr + d.llr_off;
    int16_t *soft = (d.soft_off != kNone) ? o.soft + d.soft_off : nullptr;
    crc_lin_weights<Graph<BG>::kInfoCols>(wts + t, kFixedLanes<ZC>, o, d, t, ZC, t < ZC);
    recover_fill<float>(lq, d, llr, soft, p.filler, t, nt, n);
    __syncthreads();

    int status = 1, it = 0;
    if (p.max_it > 0) {
        it = run_fixed_f32ms<BG, ZC>(lq, p, wts, slots, &status);
        for (int b = t; b < d.B / 8; b += nt) o.hard[d.hard_off + b] = (uint8_t)hard_byte(lq, b);
    }
    if (t == 0) {
        o.status[blockIdx.x] = status;
        o.iters[blockIdx.x] = it;
    }
    if (d.lq_off != kNone) {
        float *out = reinterpret_cast<float *>(o.lq) + d.lq_off;
        for (int i = t; i < n; i += nt) out[i] = lq[i];
    }
}

// ---------------------------------------------------------------------------
// Packed tiles (algo 1): several codeblocks of one (BG, Zc) per workgroup, for
// batches of mixed sizes.  Each codeblock keeps its own LDS beliefs, CRC and
// early stop; the tile iterates until all of its codeblocks have stopped.
// Per-codeblock CRC remainders are XOR-accumulated with LDS atomics into one
// of two banks (bank = iteration parity); the bank not in use is cleared
// right after each CRC barrier, when every lane has read it.
// ---------------------------------------------------------------------------
template <int BG, bool kGeneral, bool kTable, bool kSimple>
__device__ __forceinline__ void run_tile_i16(int16_t *lq, const uint32_t *etab, int zx, int Zc, int slot,
                             bool member, const uint32_t *wts,
                             uint32_t (*bank)[kMaxTileCodeblocks], const DecParams &p,
                             int *status, int *iters) {
    StateI16<BG> st;
    uint32_t msg[kGeneral ? 1 : kPackedWords<BG>];
    if constexpr (kGeneral) {
        init_state<BG>(st);
    } else {
#pragma unroll
        for (int i = 0; i < kPackedWords<BG>; ++i) msg[i] = 0u;
    }
    const GenLane g = gen_lane(lq, zx, Zc);
    bool done = !member;
    bool tile_active = __syncthreads_or(member);
    for (int it = 0; it < p.max_it && tile_active; ++it) {
        if constexpr (kGeneral)
            rows_i16<BG, 0, kGeneral, kTable>(lq, st, etab, zx, Zc, !done, p,
                                              scale_row<BG, kTable>(p, it));
        else
            gprows_iteration<BG, kTable, kSimple>(g, msg, etab, !done, p,
                                                  scale_row<BG, kTable>(p, it));
        const int b = it & 1;
        if (!done) {
            // linear CRC test (crc_lin_weights): this lane's bits zx + c * Zc
            uint32_t r = 0u;
#pragma unroll
            for (int c = 0; c < Graph<BG>::kInfoCols; ++c)
                r = crc_lin_add(r, wts[c * kTileThreads], crc_hard_mask(lq[c * Zc + zx]));
            if (r) atomicXor(&bank[b][slot], r);
        }
        __syncthreads();
        if (!done) {
            *iters = it + 1;
            if (bank[b][slot] == 0u) {
                done = true;
                *status = 0;
            }
        }
        if (member && zx == 0) bank[b ^ 1][slot] = 0u;
        tile_active = __syncthreads_or(!done);
    }
}

// One tile (workgroup descriptor `tile`) on the kTileThreads lanes of this
// workgroup; each wave reads the descriptor of its own group (TileDesc).
// Every group of a tile has the same base graph, so all waves run the same
// rows and barriers.  crc_tab holds the four CRC tables already.
template <bool kGeneral, bool kTable, bool kSimple>
__device__ __forceinline__ void tile_body(const DecOut &o, const TileDesc *tiles,
                                          const uint32_t *order, const DecParams &p,
                                          uint32_t tile, unsigned char *smem, uint32_t *wts,
                                          uint32_t (*bank)[kMaxTileCodeblocks]) {
    const int lane = threadIdx.x, nt = blockDim.x;
    const int wave = NRLDPC_WAVE_UNIFORM(lane >> 6);
    const TileDesc td = tiles[(size_t)tile * kTileWaves + wave];
    const int Zc = td.Zc, local = lane - td.lane_base;
    const int cbl = local / Zc;
    const bool member = cbl < td.count;
    const int zx = local - cbl * Zc;
    const int slot = member ? td.slot_base + cbl : 0;
    const int cbi = member ? (int)order[td.first + cbl] : 0;
    const CbDesc d = o.desc[cbi];
    const int n = d.N + 2 * Zc;
    int16_t *lq = reinterpret_cast<int16_t *>(smem) + td.lds_base + (member ? cbl : 0) * td.lds_stride;

    for (int i = lane; i < 2 * kMaxTileCodeblocks; i += nt)
        bank[i / kMaxTileCodeblocks][i % kMaxTileCodeblocks] = 0u;
    // the lane's CRC bit weights: loaded with the recovery's LLR loads, stored
    // after it (the barrier below publishes them).  With the fully unrolled
    // CRC loop below this kernel spills 16-36 bytes per lane; loading the
    // weights after the recovery and unrolling the CRC loop by 4 removes every
    // spill but measured 2% slower on the mixed batch.
    const int16_t *llr = o.llr + d.llr_off;
    int16_t *soft = (d.soft_off != kNone) ? o.soft + d.soft_off : nullptr;
    if (td.bgn == 1) {
        CrcLinRegs<Graph<1>::kInfoCols> wr;
        recover_i16<kScatterBatch>(lq, d, llr, soft, p.filler, zx, Zc, n, member,
                                   [&]() { crc_lin_load(wr, o, d, zx, Zc, member); });
        crc_lin_store(wr, wts + lane, kTileThreads);
    } else {
        CrcLinRegs<Graph<2>::kInfoCols> wr;
        recover_i16<kScatterBatch>(lq, d, llr, soft, p.filler, zx, Zc, n, member,
                                   [&]() { crc_lin_load(wr, o, d, zx, Zc, member); });
        crc_lin_store(wr, wts + lane, kTileThreads);
    }
    __syncthreads();

    int status = 1, iters = 0;
    const uint32_t *etab = o.etab + td.etab_off;
    if (p.max_it > 0) {
        if (td.bgn == 1)
            run_tile_i16<1, kGeneral, kTable, kSimple>(lq, etab, zx, Zc, slot, member, wts + lane,
                                                       bank, p, &status, &iters);
        else
            run_tile_i16<2, kGeneral, kTable, kSimple>(lq, etab, zx, Zc, slot, member, wts + lane,
                                                       bank, p, &status, &iters);
    }
    if (!member) return;
    if (p.max_it > 0)
        for (int b = zx; b < d.B / 8; b += Zc) o.hard[d.hard_off + b] = (uint8_t)hard_byte(lq, b);
    if (zx == 0) {
        o.status[cbi] = status;
        o.iters[cbi] = iters;
    }
    if (d.lq_off != kNone) {
        int16_t *out = reinterpret_cast<int16_t *>(o.lq) + d.lq_off;
        for (int i = zx; i < n; i += Zc) out[i] = lq[i];
    }
}

// Packed-tile kernel.  With q.ctr null, workgroup b decodes tile b (one
// workgroup per tile).  Otherwise the launch is persistent: a grid of at most
// one workgroup per resident slot pulls tile indices from the device queue
// counter q.ctr[0] until they run past q.ntiles, and the last workgroup to
// leave (counted in q.ctr[1]) sets both counters back to 0 for the next
// launch that uses them.
template <bool kGeneral, bool kTable, bool kSimple = false>
__global__ void __launch_bounds__(kTileThreads)
ldpc_tile_kernel(DecOut o, const TileDesc *tiles, const uint32_t *order, DecParams p,
                 TileQueue q) {
    NRLDPC_DYNAMIC_LDS(smem);
    __shared__ uint32_t bank[2][kMaxTileCodeblocks];
    __shared__ uint32_t next_tile;

    const int lane = threadIdx.x;
    // CRC bit weights after the beliefs: kMaxInfoCols columns of kTileThreads
    uint32_t *wts = reinterpret_cast<uint32_t *>(smem + q.wts_lds);
    if (!q.ctr) {
        tile_body<kGeneral, kTable, kSimple>(o, tiles, order, p, blockIdx.x, smem, wts, bank);
        return;
    }
    for (;;) {
        if (lane == 0) next_tile = atomicAdd(q.ctr, 1u);
        __syncthreads();
        const uint32_t tile = next_tile;
        if (tile >= (uint32_t)q.ntiles) break;   // uniform over the workgroup
        tile_body<kGeneral, kTable, kSimple>(o, tiles, order, p, tile, smem, wts, bank);
        // every lane is done with this tile's LDS (and with next_tile)
        __syncthreads();
    }
    if (lane == 0) {
        __threadfence();   // this workgroup's last draw precedes its exit count
        if (atomicAdd(q.ctr + 1, 1u) == gridDim.x - 1) {
            atomicExch(q.ctr, 0u);
            atomicExch(q.ctr + 1, 0u);
        }
    }
}

}  // namespace nrldpc

#if !defined(NRLDPC_CPU_EMU) && defined(NRLDPC_FIXED_TU)
// ---------------------------------------------------------------------------
// The lifting-size specialized kernels, compiled in their own translation
// unit (ldpc_fixed.hip includes this file with NRLDPC_FIXED_TU) so the
// library's two largest groups of kernels build in parallel.
// ---------------------------------------------------------------------------
namespace {
using namespace nrldpc;

// Every (general, table, simple) variant of ldpc_decode_fixed_kernel<BG, ZC>.
template <int BG, int ZC>
const void *fixed_kernel_of(bool general, bool table, bool simple) {
    if (general)
        return table ? reinterpret_cast<const void *>(&ldpc_decode_fixed_kernel<BG, ZC, true, true>)
                     : reinterpret_cast<const void *>(&ldpc_decode_fixed_kernel<BG, ZC, true, false>);
    if (simple)
        return table
                   ? reinterpret_cast<const void *>(&ldpc_decode_fixed_kernel<BG, ZC, false, true, true>)
                   : reinterpret_cast<const void *>(
                         &ldpc_decode_fixed_kernel<BG, ZC, false, false, true>);
    return table ? reinterpret_cast<const void *>(&ldpc_decode_fixed_kernel<BG, ZC, false, true>)
                 : reinterpret_cast<const void *>(&ldpc_decode_fixed_kernel<BG, ZC, false, false>);
}

// Only the fast rows without tables (algo 1 with a fast-row decoder row);
// other variants run the generic kernel.
template <int BG, int ZC>
const void *fixed_kernel_fast(bool general, bool table, bool simple) {
    if (general || table) return nullptr;
    return simple ? reinterpret_cast<const void *>(&ldpc_decode_fixed_kernel<BG, ZC, false, false, true>)
                  : reinterpret_cast<const void *>(&ldpc_decode_fixed_kernel<BG, ZC, false, false>);
}

constexpr uint32_t fixed_pair(int bg, int zc) { return ((uint32_t)bg << 16) | (uint32_t)zc; }
// BG1 at the largest lifting size (every variant), and the BG2 lifting
// sizes of the FP32 BLER workload (104 and 208; fast algo-1 rows)
constexpr uint32_t kFixedPairs[] = {fixed_pair(1, 384), fixed_pair(2, 104), fixed_pair(2, 208)};

// Set by nrldpc_configure_fixed once every specialized kernel is known to
// have no static LDS (their belief addressing assumes LDS address 0).
bool g_fixed_ok = false;

std::pair<const void *, int> fixed_kernel_any(bool general, bool table, bool simple,
                                              uint32_t uniform) {
    switch (uniform) {
    case fixed_pair(1, 384):
        return {fixed_kernel_of<1, 384>(general, table, simple), FixedLds<1, 384>::kBytes};
    case fixed_pair(2, 104):
        return {fixed_kernel_fast<2, 104>(general, table, simple), FixedLds<2, 104>::kBytes};
    case fixed_pair(2, 208):
        return {fixed_kernel_fast<2, 208>(general, table, simple), FixedLds<2, 208>::kBytes};
    default:
        return {nullptr, 0};
    }
}

// The FP32 min-sum (algo 4) kernels: the BG2 lifting sizes of the FP32 BLER
// workload.
std::pair<const void *, int> fixed_f32_kernel_any(uint32_t uniform) {
    switch (uniform) {
    case fixed_pair(2, 104):
        return {reinterpret_cast<const void *>(&ldpc_decode_fixed_f32_kernel<2, 104>),
                FixedLdsF32<2, 104>::kBytes};
    case fixed_pair(2, 208):
        return {reinterpret_cast<const void *>(&ldpc_decode_fixed_f32_kernel<2, 208>),
                FixedLdsF32<2, 208>::kBytes};
    default:
        return {nullptr, 0};
    }
}
}  // namespace

extern "C" hipError_t nrldpc_configure_fixed() {
    const int max_lds = 160 * 1024 - 2048;
    bool fixed_ok = true;
    for (uint32_t u : kFixedPairs) {
        const void *fn = fixed_f32_kernel_any(u).first;
        if (!fn) continue;
        hipError_t e = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, max_lds);
        if (e != hipSuccess) return e;
        hipFuncAttributes fa;
        e = hipFuncGetAttributes(&fa, fn);
        if (e != hipSuccess) return e;
        if (fa.sharedSizeBytes != 0) fixed_ok = false;
    }
    for (uint32_t u : kFixedPairs)
        for (int g = 0; g < 2; ++g)
            for (int t = 0; t < 2; ++t)
                for (int m = 0; m < 2; ++m) {
                    const void *fn = fixed_kernel_any(g, t, m, u).first;
                    if (!fn) continue;
                    hipError_t e =
                        hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, max_lds);
                    if (e != hipSuccess) return e;
                    hipFuncAttributes fa;
                    e = hipFuncGetAttributes(&fa, fn);
                    if (e != hipSuccess) return e;
                    if (fa.sharedSizeBytes != 0) fixed_ok = false;
                }
    g_fixed_ok = fixed_ok;
    return hipSuccess;
}

// The specialized kernel for `uniform` ((bgn << 16) | Zc) and the variant
// (general, table, simple), with its dynamic LDS bytes in *lds_bytes; NULL
// when there is none (or the kernels were not configured).
extern "C" const void *nrldpc_fixed_kernel(int general, int table, int simple, uint32_t uniform,
                                           int *lds_bytes) {
    if (!g_fixed_ok) return nullptr;
    const auto fk = fixed_kernel_any(general != 0, table != 0, simple != 0, uniform);
    if (fk.first && lds_bytes) *lds_bytes = fk.second;
    return fk.first;
}

// The specialized FP32 min-sum kernel for `uniform`, as nrldpc_fixed_kernel.
extern "C" const void *nrldpc_fixed_f32_kernel(uint32_t uniform, int *lds_bytes) {
    if (!g_fixed_ok) return nullptr;
    const auto fk = fixed_f32_kernel_any(uniform);
    if (fk.first && lds_bytes) *lds_bytes = fk.second;
    return fk.first;
}
#endif  // !NRLDPC_CPU_EMU && NRLDPC_FIXED_TU

#if !defined(NRLDPC_CPU_EMU) && !defined(NRLDPC_FIXED_TU)
extern "C" hipError_t nrldpc_configure_fixed();
extern "C" const void *nrldpc_fixed_kernel(int general, int table, int simple, uint32_t uniform,
                                           int *lds_bytes);
extern "C" const void *nrldpc_fixed_f32_kernel(uint32_t uniform, int *lds_bytes);

namespace {
using namespace nrldpc;

// int16 decode kernel of (general, table).
const void *i16_kernel(bool general, bool table) {
    if (table)
        return general ? reinterpret_cast<const void *>(
                             &ldpc_decode_kernel<Algo::kMinSumI16, int16_t, true, true>)
                       : reinterpret_cast<const void *>(
                             &ldpc_decode_kernel<Algo::kMinSumI16, int16_t, false, true>);
    return general ? reinterpret_cast<const void *>(
                         &ldpc_decode_kernel<Algo::kMinSumI16, int16_t, true, false>)
                   : reinterpret_cast<const void *>(
                         &ldpc_decode_kernel<Algo::kMinSumI16, int16_t, false, false>);
}

// Packed-tile kernel of (general, table, simple); simple applies to fast rows.
const void *tile_kernel(bool general, bool table, bool simple) {
    if (general)
        return table ? reinterpret_cast<const void *>(&ldpc_tile_kernel<true, true>)
                     : reinterpret_cast<const void *>(&ldpc_tile_kernel<true, false>);
    if (simple)
        return table ? reinterpret_cast<const void *>(&ldpc_tile_kernel<false, true, true>)
                     : reinterpret_cast<const void *>(&ldpc_tile_kernel<false, false, true>);
    return table ? reinterpret_cast<const void *>(&ldpc_tile_kernel<false, true>)
                 : reinterpret_cast<const void *>(&ldpc_tile_kernel<false, false>);
}
}  // namespace

// Host-callable launchers (C linkage inside the library only).  Algos 1-3
// share the int16 kernels (2 and 3 with per-row scale words).
extern "C" hipError_t nrldpc_launch_bp(const nrldpc::DecOut *o, const nrldpc::DecParams *p,
                                       int ncb, int threads, int max_soft, hipStream_t stream);
extern "C" hipError_t nrldpc_configure_bp();

// uniform: (bgn << 16) | Zc when every codeblock of the launch shares one base
// graph and lifting size, else 0; a specialized int16 kernel runs when one
// exists for that pair.
extern "C" hipError_t nrldpc_launch_decode(int algo, const nrldpc::DecOut *o,
                                           const nrldpc::DecParams *p, int ncb, int threads,
                                           int max_soft, uint32_t uniform, hipStream_t stream) {
    using namespace nrldpc;
    const bool i16 = algo >= 1 && algo <= 3;
    const size_t esz = i16 ? sizeof(int16_t) : sizeof(float);
    // beliefs, then the CRC bit weights of the min-sum kernels: one column
    // per systematic column of the launch's base graph (both when mixed)
    const int wcols = (uniform >> 16) == 2u ? Graph<2>::kInfoCols : kGenericWtsCols;
    size_t lds = (((size_t)max_soft * esz + 15) & ~(size_t)15) +
                 (size_t)wcols * (size_t)threads * 4;
    const dim3 grid(ncb), block(threads);
    DecOut ov = *o;
    DecParams pv = *p;
    void *args[] = {&ov, &pv};
    const void *fn = nullptr;
    switch (algo) {
    case 1:
    case 2:
    case 3:
        if (algo != 1 && (pv.scale_iters < 1 || !pv.scale[0] || !pv.scale[1]))
            return hipErrorInvalidValue;
        fn = i16_kernel(pv.general != 0, algo != 1);
        // the specialized kernels run one lane per lifted check, Zc rounded
        // up to whole waves
        if (uniform && threads == (int)(((uniform & 0xFFFFu) + 63u) / 64u * 64u)) {
            int fl = 0;
            if (const void *fk = nrldpc_fixed_kernel(pv.general != 0, algo != 1, pv.simple != 0,
                                                     uniform, &fl)) {
                fn = fk;
                lds = (size_t)fl;
            }
        }
        break;
    case 4:
        fn = reinterpret_cast<const void *>(&ldpc_decode_kernel<Algo::kMinSumF32, float>);
        if (uniform) {
            int fl = 0;
            if (const void *fk = nrldpc_fixed_f32_kernel(uniform, &fl)) {
                fn = fk;
                lds = (size_t)fl;
            }
        }
        break;
    case 0:
        return nrldpc_launch_bp(o, p, ncb, threads, max_soft, stream);
    default:
        return hipErrorInvalidValue;
    }
    const hipError_t e = hipLaunchKernel(fn, grid, block, args, lds, stream);
    if (e != hipSuccess) return e;
    return hipGetLastError();
}

extern "C" hipError_t nrldpc_configure_kernels() {
    using namespace nrldpc;
    const int max_lds = 160 * 1024 - 2048;
    hipError_t e = hipSuccess;
    for (int g = 0; g < 2 && e == hipSuccess; ++g)
        for (int t = 0; t < 2 && e == hipSuccess; ++t)
            e = hipFuncSetAttribute(i16_kernel(g, t), hipFuncAttributeMaxDynamicSharedMemorySize,
                                    max_lds);
    if (e != hipSuccess) return e;
    e = nrldpc_configure_fixed();
    if (e != hipSuccess) return e;
    e = hipFuncSetAttribute(
        reinterpret_cast<const void *>(&ldpc_decode_kernel<Algo::kMinSumF32, float>),
        hipFuncAttributeMaxDynamicSharedMemorySize, max_lds);
    if (e != hipSuccess) return e;
    e = nrldpc_configure_bp();
    if (e != hipSuccess) return e;
    // a tile workgroup holds at most 512 lanes' worth of beliefs (~70 KB for
    // base graph 1; at most 90 KB accepted) and the lanes' CRC bit weights
    const int tile_lds = 90 * 1024 + kTileWtsBytes;
    for (int g = 0; g < 2 && e == hipSuccess; ++g)
        for (int t = 0; t < 2 && e == hipSuccess; ++t)
            for (int m = 0; m < 2 && e == hipSuccess; ++m)
                e = hipFuncSetAttribute(tile_kernel(g, t, m),
                                        hipFuncAttributeMaxDynamicSharedMemorySize, tile_lds);
    return e;
}

// Lane-per-edge launch (algo 1, fast rows): one 64-lane workgroup per
// codeblock; every codeblock must have Zc <= kEdgeLaneMaxZc of its base graph.
extern "C" hipError_t nrldpc_launch_edge_lanes(const nrldpc::DecOut *o,
                                               const nrldpc::DecParams *p, int ncb,
                                               int max_soft, hipStream_t stream) {
    using namespace nrldpc;
    if (p->general) return hipErrorInvalidValue;
    const size_t lds = ((size_t)max_soft * sizeof(int16_t) + 15) & ~(size_t)15;
    DecOut ov = *o;
    DecParams pv = *p;
    void *args[] = {&ov, &pv};
    const hipError_t e = hipLaunchKernel(
        reinterpret_cast<const void *>(&ldpc_decode_kernel<Algo::kMinSumI16Edge, int16_t>),
        dim3(ncb), dim3(64), args, lds, stream);
    if (e != hipSuccess) return e;
    return hipGetLastError();
}

// Largest lifting size of base graph bgn that the lane-per-edge schedule takes.
extern "C" int nrldpc_edge_lanes_max_zc(int bgn) {
    return bgn == 1 ? nrldpc::kEdgeLaneMaxZc<1> : nrldpc::kEdgeLaneMaxZc<2>;
}

// Packed-tile launch (algos 1-3): ntiles workgroups of kTileThreads lanes,
// or, with queue non-null (two zeroed words), a persistent grid of at most
// one workgroup per resident slot that pulls the ntiles tiles from it.
extern "C" hipError_t nrldpc_launch_tiles(int algo, const nrldpc::DecOut *o,
                                          const nrldpc::TileDesc *tiles, const uint32_t *order,
                                          const nrldpc::DecParams *p, int ntiles,
                                          int lds_entries, uint32_t *queue, hipStream_t stream) {
    using namespace nrldpc;
    if (algo < 1 || algo > 3) return hipErrorInvalidValue;
    if (algo != 1 && (p->scale_iters < 1 || !p->scale[0] || !p->scale[1]))
        return hipErrorInvalidValue;
    const size_t beliefs = ((size_t)lds_entries * sizeof(int16_t) + 15) & ~(size_t)15;
    if (beliefs > 90 * 1024) return hipErrorInvalidValue;
    const size_t lds = beliefs + kTileWtsBytes;
    const void *fn = tile_kernel(p->general != 0, algo != 1, p->simple != 0);
    int grid = ntiles;
    if (queue) {
        int dev = 0, cus = 0, per_cu = 0;
        hipError_t e = hipGetDevice(&dev);
        if (e == hipSuccess) e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
        if (e == hipSuccess)
            e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, kTileThreads, lds);
        if (e != hipSuccess) return e;
        grid = std::min(ntiles, std::max(1, cus * per_cu));
    }
    DecOut ov = *o;
    DecParams pv = *p;
    TileQueue q;
    q.ctr = queue;
    q.ntiles = ntiles;
    q.wts_lds = (uint32_t)beliefs;
    void *args[] = {&ov, &tiles, &order, &pv, &q};
    const hipError_t e = hipLaunchKernel(fn, dim3(grid), dim3(kTileThreads), args, lds, stream);
    if (e != hipSuccess) return e;
    return hipGetLastError();
}
#endif  // !NRLDPC_CPU_EMU && !NRLDPC_FIXED_TU
