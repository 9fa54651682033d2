// Unrolled FP32 layered min-sum with a hand-written reverse pass (gfx950):
// the trainable decoder of deep_learning/deeplearning_ldpc_model.py:315-410
// (CTableNeuralLayeredMinSumDecoder.forward) and the gradient torch autograd
// takes through it, for training alpha/beta on the GPU.
//
// Per iteration t and base-graph row r, every lifted check of the row does
//   rel_j = LQ[col_j] - Lr_j;  s_j = rel_j < 0;  a_j = |rel_j|
//   m1,i1 = min_j a_j (first index);  m2,i2 = min_{j != i1} a_j (first index)
//   Lr_j  = (-1)^(s_j ^ S) * alpha[t][r] * max((j == i1 ? m2 : m1) - beta[t][r], 0)
//   LQ[col_j] = clamp(rel_j + Lr_j, -clip, clip)     (clamp only with clip_llr)
// with S the parity of the s_j, rows in order with a workgroup barrier in
// between (lifted checks of one row touch disjoint columns).  Each operation is
// a single IEEE float operation (built with -ffp-contract=off), so the forward
// matches the reference's FP32 tensors bit for bit.
//
// Mapping: a workgroup holds `cbs` codeblocks of `lpc` lanes each (lpc = Zc
// rounded up to a power of two below 64, to whole waves above), lane zx owns
// lifted check zx of every row, beliefs LQ (forward) or their gradient (reverse)
// live in LDS.  The forward keeps each row's messages compressed in VGPRs
// (alpha*max(m1-beta,0), alpha*max(m2-beta,0), argmin, one sign bit per edge)
// and writes a per-(codeblock, t, r, check) record for the reverse pass:
//   w0 = m1, w1 = m2 (float bits), w2 = sign bits | i1 << 19 | i2 << 24,
//   w3 = clamp-pass bits (edge j's clamp passes gradient), only with clip_llr,
// stored [codeblock][t][r][word][zx] so a row's words are coalesced.
//
// Reverse pass, rows and iterations backwards: for the edges of one check
//   gq_j   = dL/dLQ[col_j] if the clamp passed, else 0
//   gn_j   = gq_j + dL/dLr_j (from the next iteration's rel, kept in HBM)
//   gmag_j = (-1)^(s_j ^ S) * gn_j
//   dL/dalpha += gmag_j * max(u_j - beta, 0);  gd_j = [u_j - beta >= 0] alpha gmag_j
//   dL/dbeta  -= gd_j;  m2 collects gd_{i1}, m1 the other gd_j
//   grel_j = gq_j + (j == i1 ? dL/dm1 sgn(rel_i1) : j == i2 ? dL/dm2 sgn(rel_i2) : 0)
//   dL/dLQ_old[col_j] = grel_j,  dL/dLr_old_j = -grel_j
// (torch's conventions: abs' (0) = 0, clamp passes at its bounds, min routes
// to the first minimal index).  With detach the check update sees rel as a
// constant (deeplearning_ldpc_model.py:365), so grel_j = 0.  Per-row alpha and
// beta sums are reduced over lane segments of <= 64 lanes by xor shuffles and
// written per (t, r, codeblock, segment); a last kernel sums them in a fixed
// order, so gradients are deterministic.
#ifdef NRLDPC_CPU_EMU
#include "hip_cpu_emu.h"
#define NRLDPC_DYNAMIC_LDS(name) unsigned char *name = emu_dynamic_lds()
#else
#include <hip/hip_runtime.h>
#define NRLDPC_DYNAMIC_LDS(name) extern __shared__ __align__(16) unsigned char name[]
#endif

#include "ldpc_device.h"
#include "ldpc_graph.h"

// Marks a value as redefined at this point so that the compiler recomputes
// what depends on it here instead of hoisting it out of the iteration loop
// (rows are fully unrolled: hoisted per-edge address arithmetic of all rows
// does not fit in registers).  No code is emitted.
#ifdef NRLDPC_CPU_EMU
#define NRLDPC_OPAQUE_V(x) (void)(x)
#define NRLDPC_OPAQUE_S(x) (void)(x)
#else
#define NRLDPC_OPAQUE_V(x) asm volatile("" : "+v"(x))
#define NRLDPC_OPAQUE_S(x) asm volatile("" : "+s"(x))
#endif

namespace nrldpc {

struct UnrolledArgs {
    const float *llr;       // [batch][n] input beliefs
    float *lq;              // [batch][n] final beliefs
    float *iter_lq;         // [T][batch][n] beliefs after each iteration, or null
    uint32_t *rec;          // forward records, or null (no reverse pass)
    const float *alpha;     // [T][rows]
    const float *beta;      // [T][rows]
    const uint32_t *etab;   // [edges] (column*Zc) | (V_ij mod Zc) << 16
    const float *g_lq;      // [batch][n]
    const float *g_iter;    // [T][batch][n] or null
    float *g_llr;           // [batch][n] or null
    float *g_lr;            // [batch][edges][Zc] message gradients (null with detach)
    float *partial;         // [T*rows][batch*segs][2] alpha/beta partial sums
    int batch, n, Zc, T;
    int lpc, cbs, seg;      // lanes per codeblock, codeblocks per workgroup, segment
    int stride;             // LDS floats per codeblock
    int words;              // record words per check: 3, or 4 with clip
    int has_clip, detach;
    float clip;
};

__device__ __forceinline__ uint32_t f2u(float f) { return __builtin_bit_cast(uint32_t, f); }

// Median of three (exact for the non-negative, non-NaN magnitudes used here).
__device__ __forceinline__ float fmed3(float a, float b, float c) {
#ifdef NRLDPC_CPU_EMU
    return fmaxf(fminf(a, b), fminf(fmaxf(a, b), c));
#else
    return __builtin_amdgcn_fmed3f(a, b, c);
#endif
}
__device__ __forceinline__ float u2f(uint32_t u) { return __builtin_bit_cast(float, u); }

// The edge table through the constant address space: uniform indices become
// scalar loads and the column / shift decoding scalar ops.
#ifdef NRLDPC_CPU_EMU
typedef const uint32_t *uetab_ptr;
#else
typedef const __attribute__((address_space(4))) uint32_t *uetab_ptr;
#endif

// v negated when bit j of bits is set: IEEE negation flips the sign bit, so
// a shift and one bitwise op give exactly -v or v.
__device__ __forceinline__ float uflip_sign(float v, uint32_t bits, int j) {
    return __builtin_bit_cast(float, __builtin_bit_cast(uint32_t, v) ^ ((bits << (31 - j)) & 0x80000000u));
}

template <int BG>
struct UnrState {
    float g1[Graph<BG>::kRows];   // alpha * max(m1 - beta, 0)
    float g2[Graph<BG>::kRows];   // alpha * max(m2 - beta, 0)
    uint32_t sgn[Graph<BG>::kRows];   // message sign bit per edge | argmin << 24
};

// One lifted check of row R, forward.
template <int BG, int R>
__device__ __forceinline__ void urow_fwd(float *__restrict__ lq, UnrState<BG> &st,
                                         const UnrolledArgs &a, uetab_ptr etab,
                                         const float *ab, int zx, uint32_t *__restrict__ rec) {
    constexpr int e0 = Graph<BG>::start(R);
    constexpr int deg = Graph<BG>::start(R + 1) - e0;
    constexpr int nR = Graph<BG>::kRows;
    int Zc = a.Zc;
    NRLDPC_OPAQUE_S(Zc);
    NRLDPC_OPAQUE_V(zx);
    NRLDPC_OPAQUE_S(etab);
    const float al = ab[R], be = ab[nR + R];
    const float o1 = st.g1[R], o2 = st.g2[R];
    const uint32_t sb = st.sgn[R];
    const int oidx = (int)(sb >> 24);
    int addr[deg];
    float rel[deg];
#pragma unroll
    for (int j = 0; j < deg; ++j) {
        const uint32_t e = etab[e0 + j];
        int c = zx + (int)(e >> 16);
        c = (c >= Zc) ? c - Zc : c;
        addr[j] = (int)(e & 0xFFFFu) + c;
        const float ov = (j == oidx) ? o2 : o1;
        const float lr = uflip_sign(ov, sb, j);
        rel[j] = lq[addr[j]] - lr;
    }
    // two smallest magnitudes (with multiplicity), then the first index of
    // each: i1 = first j with |rel_j| == m1, i2 = first j != i1 with
    // |rel_j| == m2 (torch.min's first-index rule)
    float m1 = INFINITY, m2 = INFINITY;
    uint32_t negs = 0;
#pragma unroll
    for (int j = 0; j < deg; ++j) {
        negs |= (uint32_t)(rel[j] < 0.0f) << j;
        m2 = fmed3(m1, fabsf(rel[j]), m2);
        m1 = fminf(m1, fabsf(rel[j]));
    }
    int i1 = 0, i2 = 0;
#pragma unroll
    for (int j = deg - 1; j >= 0; --j) i1 = (fabsf(rel[j]) == m1) ? j : i1;
#pragma unroll
    for (int j = deg - 1; j >= 0; --j) i2 = (fabsf(rel[j]) == m2 && j != i1) ? j : i2;
    float d1 = m1 - be, d2 = m2 - be;
    d1 = (d1 < 0.0f) ? 0.0f : d1;
    d2 = (d2 < 0.0f) ? 0.0f : d2;
    const float g1 = al * d1, g2 = al * d2;
    constexpr uint32_t mask = (1u << deg) - 1u;
    const uint32_t bits = ((__popc(negs) & 1) ? ~negs : negs) & mask;
    uint32_t pass = 0;
#pragma unroll
    for (int j = 0; j < deg; ++j) {
        const float mag = (j == i1) ? g2 : g1;
        const float nl = uflip_sign(mag, bits, j);
        float q = rel[j] + nl;
        if (a.has_clip) {
            pass |= (uint32_t)(q >= -a.clip && q <= a.clip) << j;
            q = (q < -a.clip) ? -a.clip : q;
            q = (q > a.clip) ? a.clip : q;
        }
        lq[addr[j]] = q;
    }
    st.g1[R] = g1;
    st.g2[R] = g2;
    st.sgn[R] = bits | ((uint32_t)i1 << 24);
    if (a.rec) {
        uint32_t *w = rec + (size_t)R * a.words * Zc + zx;
        w[0] = f2u(m1);
        w[Zc] = f2u(m2);
        w[2 * Zc] = negs | ((uint32_t)i1 << 19) | ((uint32_t)i2 << 24);
        if (a.has_clip) w[3 * Zc] = pass;
    }
}

template <int BG, int R>
__device__ __forceinline__ void urows_fwd(float *lq, UnrState<BG> &st, const UnrolledArgs &a,
                                          uetab_ptr etab, const float *ab, int zx,
                                          bool active, uint32_t *rec) {
    if constexpr (R < Graph<BG>::kRows) {
        if (active) urow_fwd<BG, R>(lq, st, a, etab, ab, zx, rec);
        // column-disjoint rows commute (RowGroups): a barrier between groups only
        if constexpr (kRowGroups<BG>.barrier_after[R]) __syncthreads();
        urows_fwd<BG, R + 1>(lq, st, a, etab, ab, zx, active, rec);
    }
}

// Sum over an aligned segment of `seg` lanes (a power of two <= 64).
__device__ __forceinline__ float seg_sum(float v, int seg) {
    for (int m = 1; m < seg; m <<= 1) v += __shfl_xor(v, m, 64);
    return v;
}

// One lifted check of row R, reverse, in two phases so that every row of a
// barrier group issues its loads before any row of the group computes: the
// row's message-gradient stores go to global memory, which the next row's
// gradient loads would otherwise wait behind (the rows of a group touch
// disjoint beliefs and message slots, so their order within the group does
// not matter).
template <int DEG>
struct UbwdLoad {
    float m1, m2;
    uint32_t w2, pass;
    float gin[DEG];     // message gradients of the later iteration
    int addr[DEG];      // belief indices of the row's edges
    float g[DEG];       // belief gradients (clamp gating applied in the compute phase,
                        // so no load result is waited for before the group's last load)
};

template <int BG, int R>
__device__ __forceinline__ void urow_bwd_load(const float *__restrict__ glq, const UnrolledArgs &a,
                                              uetab_ptr etab, int zx, bool active, bool lr_in,
                                              const uint32_t *__restrict__ rec,
                                              const float *__restrict__ glr,
                                              UbwdLoad<Graph<BG>::start(R + 1) -
                                                       Graph<BG>::start(R)> &ld) {
    constexpr int e0 = Graph<BG>::start(R);
    constexpr int deg = Graph<BG>::start(R + 1) - e0;
    int Zc = a.Zc;
    NRLDPC_OPAQUE_S(Zc);
    NRLDPC_OPAQUE_V(zx);
    NRLDPC_OPAQUE_S(etab);
    // every lane loads (inactive ones at lane 0 of their codeblock, results
    // unused): the load phase stays one straight-line block, so no wait for
    // its loads is forced at the end of a divergent branch
    zx = active ? zx : 0;
    const uint32_t *w = rec + (size_t)R * a.words * Zc + zx;
    ld.m1 = u2f(w[0]);
    ld.m2 = u2f(w[Zc]);
    ld.w2 = w[2 * Zc];
    ld.pass = a.has_clip ? w[3 * Zc] : 0xFFFFFFFFu;
    if (lr_in) {
        const float *lr = glr + (size_t)e0 * Zc + zx;
#pragma unroll
        for (int j = 0; j < deg; ++j) ld.gin[j] = lr[j * Zc];
    }
#pragma unroll
    for (int j = 0; j < deg; ++j) {
        const uint32_t e = etab[e0 + j];
        int c = zx + (int)(e >> 16);
        c = (c >= Zc) ? c - Zc : c;
        ld.addr[j] = (int)(e & 0xFFFFu) + c;
        ld.g[j] = glq[ld.addr[j]];
    }
}

// Called by every lane of the workgroup (the segment sums shuffle across
// lanes).  lr_in / lr_out are uniform: whether message gradients are read (not
// in the last iteration) and written (not in the first).
template <int BG, int R>
__device__ __forceinline__ void urow_bwd_compute(float *__restrict__ glq, const UnrolledArgs &a,
                                                 const float *ab, int zx, int b, int t,
                                                 bool active, bool member, bool lr_in,
                                                 bool lr_out,
                                                 const UbwdLoad<Graph<BG>::start(R + 1) -
                                                                Graph<BG>::start(R)> &ld,
                                                 float *__restrict__ glr) {
    constexpr int e0 = Graph<BG>::start(R);
    constexpr int deg = Graph<BG>::start(R + 1) - e0;
    constexpr int nR = Graph<BG>::kRows;
    int Zc = a.Zc;
    NRLDPC_OPAQUE_S(Zc);
    float ga = 0.0f, gb = 0.0f;
    if (active) {
        const float al = ab[R], be = ab[nR + R];
        const float m1 = ld.m1, m2 = ld.m2;
        const uint32_t w2 = ld.w2;
        float *lr = glr + (size_t)e0 * Zc + zx;
        const uint32_t negs = w2 & 0x7FFFFu;
        const int i1 = (int)((w2 >> 19) & 31u), i2 = (int)((w2 >> 24) & 31u);
        constexpr uint32_t mask = (1u << deg) - 1u;
        const uint32_t bits = ((__popc(negs) & 1) ? ~negs : negs) & mask;
        const float d1 = m1 - be, d2 = m2 - be;
        const float c1 = (d1 < 0.0f) ? 0.0f : d1, c2 = (d2 < 0.0f) ? 0.0f : d2;
        const bool p1 = d1 >= 0.0f, p2 = d2 >= 0.0f;
        float gq[deg];
#pragma unroll
        for (int j = 0; j < deg; ++j) gq[j] = ((ld.pass >> j) & 1u) ? ld.g[j] : 0.0f;
        float gm1 = 0.0f, gm2 = 0.0f;
#pragma unroll
        for (int j = 0; j < deg; ++j) {
            const float gn = lr_in ? gq[j] + ld.gin[j] : gq[j];
            const float gmag = uflip_sign(gn, bits, j);
            const bool first = (j == i1);
            ga += gmag * (first ? c2 : c1);
            const float gd = (first ? p2 : p1) ? gmag * al : 0.0f;
            gb -= gd;
            if (first)
                gm2 = gd;
            else
                gm1 += gd;
        }
        const float s1 = (m1 == 0.0f) ? 0.0f : (((negs >> i1) & 1u) ? -1.0f : 1.0f);
        const float s2 = (m2 == 0.0f) ? 0.0f : (((negs >> i2) & 1u) ? -1.0f : 1.0f);
#pragma unroll
        for (int j = 0; j < deg; ++j) {
            const float gabs = (j == i1) ? gm1 * s1 : ((j == i2) ? gm2 * s2 : 0.0f);
            const float gr = a.detach ? 0.0f : gq[j] + gabs;
            glq[ld.addr[j]] = gr;
            if (lr_out) lr[j * Zc] = -gr;
        }
    }
    ga = seg_sum(ga, a.seg);
    gb = seg_sum(gb, a.seg);
    if (member && (zx & (a.seg - 1)) == 0) {
        const int segs = a.lpc / a.seg;
        float *p = a.partial + 2 * (((size_t)(t * nR + R) * a.batch + b) * segs + zx / a.seg);
        p[0] = ga;
        p[1] = gb;
    }
}

// First row of the barrier group that holds row R.
template <int BG>
__host__ __device__ constexpr int group_first(int r) {
    while (r > 0 && !kRowGroups<BG>.barrier_after[r - 1]) --r;
    return r;
}

// Rows R down to RFirst of one barrier group: this row's loads, then (by
// recursion) the loads and computes of the rows below it, then this row's
// compute -- every load of the group is issued before the first compute.
template <int BG, int R, int RFirst>
__device__ __forceinline__ void ugroup_bwd(float *glq, const UnrolledArgs &a, uetab_ptr etab,
                                           const float *ab, int zx, int b, int t, bool active,
                                           bool member, bool lr_in, bool lr_out,
                                           const uint32_t *rec, float *glr) {
    UbwdLoad<Graph<BG>::start(R + 1) - Graph<BG>::start(R)> ld;
    urow_bwd_load<BG, R>(glq, a, etab, zx, active, lr_in, rec, glr, ld);
    if constexpr (R > RFirst)
        ugroup_bwd<BG, R - 1, RFirst>(glq, a, etab, ab, zx, b, t, active, member, lr_in, lr_out,
                                      rec, glr);
    urow_bwd_compute<BG, R>(glq, a, ab, zx, b, t, active, member, lr_in, lr_out, ld, glr);
}

// Rows R down to 0 of one iteration, group by group, a barrier after each.
template <int BG, int R>
__device__ __forceinline__ void urows_bwd(float *glq, const UnrolledArgs &a, uetab_ptr etab,
                                          const float *ab, int zx, int b, int t, bool active,
                                          bool member, bool lr_in, bool lr_out,
                                          const uint32_t *rec, float *glr) {
    if constexpr (R >= 0) {
        constexpr int first = group_first<BG>(R);
        ugroup_bwd<BG, R, first>(glq, a, etab, ab, zx, b, t, active, member, lr_in, lr_out, rec,
                                 glr);
        __syncthreads();
        urows_bwd<BG, first - 1>(glq, a, etab, ab, zx, b, t, active, member, lr_in, lr_out, rec,
                                 glr);
    }
}

// alpha[t][*] and beta[t][*] into LDS for the rows of iteration t.  The
// caller has a barrier between the previous iteration's rows and this.
template <int BG>
__device__ __forceinline__ void load_ab(float *ab, const UnrolledArgs &a, int t) {
    constexpr int nR = Graph<BG>::kRows;
    for (int i = threadIdx.x; i < 2 * nR; i += blockDim.x)
        ab[i] = (i < nR) ? a.alpha[t * nR + i] : a.beta[t * nR + i - nR];
    __syncthreads();
}

template <int BG>
__global__ void __launch_bounds__(kMaxThreads) unrolled_fwd_kernel(UnrolledArgs a) {
    NRLDPC_DYNAMIC_LDS(smem);
    __shared__ float ab[2 * Graph<BG>::kRows];
    constexpr int nR = Graph<BG>::kRows;
    const uetab_ptr etab = (uetab_ptr)a.etab;
    const int lane = threadIdx.x, cbl = lane / a.lpc, zx = lane - cbl * a.lpc;
    const int b = blockIdx.x * a.cbs + cbl;
    const bool member = b < a.batch;
    const bool active = member && zx < a.Zc;
    const int n = a.n;
    float *lq = reinterpret_cast<float *>(smem) + (size_t)cbl * a.stride;
    if (member) {
        const float *src = a.llr + (size_t)b * n;
        for (int i = zx; i < n; i += a.lpc) {
            float v = src[i];
            if (a.has_clip) {
                v = (v < -a.clip) ? -a.clip : v;
                v = (v > a.clip) ? a.clip : v;
            }
            lq[i] = v;
        }
    }
    __syncthreads();
    UnrState<BG> st;
#pragma unroll
    for (int r = 0; r < nR; ++r) {
        st.g1[r] = 0.0f;
        st.g2[r] = 0.0f;
        st.sgn[r] = 255u << 24;
    }
    for (int t = 0; t < a.T; ++t) {
        uint32_t *rec = (a.rec && member)
                            ? a.rec + ((size_t)b * a.T + t) * nR * a.words * a.Zc
                            : nullptr;
        load_ab<BG>(ab, a, t);
        urows_fwd<BG, 0>(lq, st, a, etab, ab, zx, active, rec);
        if (a.iter_lq) {
            if (member) {
                float *dst = a.iter_lq + ((size_t)t * a.batch + b) * n;
                for (int i = zx; i < n; i += a.lpc) dst[i] = lq[i];
            }
            __syncthreads();
        }
    }
    if (member) {
        float *dst = a.lq + (size_t)b * n;
        for (int i = zx; i < n; i += a.lpc) dst[i] = lq[i];
    }
}

template <int BG>
__global__ void __launch_bounds__(kMaxThreads) unrolled_bwd_kernel(UnrolledArgs a) {
    NRLDPC_DYNAMIC_LDS(smem);
    __shared__ float ab[2 * Graph<BG>::kRows];
    constexpr int nR = Graph<BG>::kRows;
    const uetab_ptr etab = (uetab_ptr)a.etab;
    constexpr int nE = Graph<BG>::kEdges;
    const int lane = threadIdx.x, cbl = lane / a.lpc, zx = lane - cbl * a.lpc;
    const int b = blockIdx.x * a.cbs + cbl;
    const bool member = b < a.batch;
    const bool active = member && zx < a.Zc;
    const int n = a.n;
    float *glq = reinterpret_cast<float *>(smem) + (size_t)cbl * a.stride;
    if (member) {
        const float *g = a.g_lq + (size_t)b * n;
        const float *gi =
            (a.g_iter && a.T > 0) ? a.g_iter + ((size_t)(a.T - 1) * a.batch + b) * n : nullptr;
        for (int i = zx; i < n; i += a.lpc) glq[i] = gi ? g[i] + gi[i] : g[i];
    }
    __syncthreads();
    float *glr = a.g_lr ? a.g_lr + (size_t)(member ? b : a.batch - 1) * nE * a.Zc : nullptr;
    for (int t = a.T - 1; t >= 0; --t) {
        if (a.g_iter && t < a.T - 1) {
            if (member) {
                const float *gi = a.g_iter + ((size_t)t * a.batch + b) * n;
                for (int i = zx; i < n; i += a.lpc) glq[i] += gi[i];
            }
            __syncthreads();
        }
        // the last codeblock's record for lanes past the batch (loaded, unused)
        const int bl = member ? b : a.batch - 1;
        const uint32_t *rec = a.rec + ((size_t)bl * a.T + t) * nR * a.words * a.Zc;
        const bool lr_on = a.g_lr != nullptr;
        load_ab<BG>(ab, a, t);
        urows_bwd<BG, nR - 1>(glq, a, etab, ab, zx, b, t, active, member, lr_on && t < a.T - 1,
                              lr_on && t > 0, rec, glr);
    }
    if (member && a.g_llr) {
        const float *src = a.llr + (size_t)b * n;
        float *dst = a.g_llr + (size_t)b * n;
        for (int i = zx; i < n; i += a.lpc) {
            const float v = src[i];
            const bool in = !a.has_clip || (v >= -a.clip && v <= a.clip);
            dst[i] = in ? glq[i] : 0.0f;
        }
    }
}

// Per-(t, r) sums of the partial alpha/beta gradients: one workgroup per
// (t, r), fixed summation order (strided per lane, then a tree), in double.
constexpr int kReduceThreads = 256;
__global__ void __launch_bounds__(kReduceThreads)
unrolled_reduce_kernel(const float *partial, int m, float *g_alpha, float *g_beta) {
    __shared__ double ra[kReduceThreads], rb[kReduceThreads];
    const int tid = threadIdx.x;
    const float *p = partial + (size_t)blockIdx.x * m * 2;
    double sa = 0.0, sb = 0.0;
    for (int i = tid; i < m; i += kReduceThreads) {
        sa += (double)p[2 * i];
        sb += (double)p[2 * i + 1];
    }
    ra[tid] = sa;
    rb[tid] = sb;
    __syncthreads();
    for (int s = kReduceThreads / 2; s > 0; s >>= 1) {
        if (tid < s) {
            ra[tid] += ra[tid + s];
            rb[tid] += rb[tid + s];
        }
        __syncthreads();
    }
    if (tid == 0) {
        g_alpha[blockIdx.x] = (float)ra[0];
        g_beta[blockIdx.x] = (float)rb[0];
    }
}

// ---------------------------------------------------------------------------
// Launch shape and buffer sizes (host side; shared with the CPU build).
// ---------------------------------------------------------------------------
constexpr int kUnrDefaultLdsBytes = 64 * 1024;   // per workgroup, so two fit on a CU
constexpr int kUnrMaxLdsBytes = 160 * 1024 - 4096;

struct UnrolledShape {
    int bgn, Zc, n, rows, edges, T;
    int has_clip, detach;
    float clip;
    int lpc, seg, cbs;
};

// Lanes per codeblock: Zc rounded up to a power of two below 64 lanes (so a
// codeblock is an aligned lane segment of a wave), to whole waves above.
inline int unr_lanes_per_codeblock(int zc) {
    if (zc >= 64) return ((zc + 63) / 64) * 64;
    int l = 1;
    while (l < zc) l <<= 1;
    return l;
}

inline int unr_lds_stride(int n) { return (n + 3) & ~3; }

inline int unr_default_cbs(const UnrolledShape &h) {
    int k = kMaxThreads / h.lpc;
    const int per = unr_lds_stride(h.n) * (int)sizeof(float);
    while (k > 1 && k * per > kUnrDefaultLdsBytes) --k;
    return k < 1 ? 1 : k;
}

inline bool unr_cbs_fits(const UnrolledShape &h, int k) {
    return k >= 1 && k * h.lpc <= kMaxThreads &&
           (size_t)k * unr_lds_stride(h.n) * sizeof(float) <= (size_t)kUnrMaxLdsBytes;
}

inline UnrolledShape unr_shape(int bgn, int zc, int N, int T, float clip, int detach) {
    UnrolledShape h{};
    h.bgn = bgn;
    h.Zc = zc;
    h.n = N + 2 * zc;
    h.rows = bgn == 1 ? Graph<1>::kRows : Graph<2>::kRows;
    h.edges = bgn == 1 ? Graph<1>::kEdges : Graph<2>::kEdges;
    h.T = T;
    h.has_clip = clip > 0.0f;
    h.clip = clip;
    h.detach = detach ? 1 : 0;
    h.lpc = unr_lanes_per_codeblock(zc);
    h.seg = h.lpc < 64 ? h.lpc : 64;
    h.cbs = unr_default_cbs(h);
    return h;
}

inline size_t unr_align256(size_t v) { return (v + 255) & ~(size_t)255; }

inline size_t unr_record_bytes(const UnrolledShape &h, int batch) {
    return (size_t)batch * h.T * h.rows * (h.has_clip ? 4 : 3) * h.Zc * sizeof(uint32_t);
}

inline size_t unr_glr_bytes(const UnrolledShape &h, int batch) {
    return h.detach ? 0 : unr_align256((size_t)batch * h.edges * h.Zc * sizeof(float));
}

inline size_t unr_partial_bytes(const UnrolledShape &h, int batch) {
    return (size_t)h.T * h.rows * batch * (h.lpc / h.seg) * 2 * sizeof(float);
}

inline size_t unr_scratch_bytes(const UnrolledShape &h, int batch) {
    return unr_glr_bytes(h, batch) + unr_partial_bytes(h, batch);
}

inline UnrolledArgs unr_args(const UnrolledShape &h, int batch, const uint32_t *etab) {
    UnrolledArgs a{};
    a.etab = etab;
    a.batch = batch;
    a.n = h.n;
    a.Zc = h.Zc;
    a.T = h.T;
    a.lpc = h.lpc;
    a.cbs = h.cbs;
    a.seg = h.seg;
    a.stride = unr_lds_stride(h.n);
    a.words = h.has_clip ? 4 : 3;
    a.has_clip = h.has_clip;
    a.detach = h.detach;
    a.clip = h.clip;
    return a;
}

// Backward buffers carved from one scratch allocation.
inline void unr_bind_scratch(const UnrolledShape &h, int batch, void *scratch, UnrolledArgs &a) {
    unsigned char *sc = static_cast<unsigned char *>(scratch);
    a.g_lr = h.detach ? nullptr : reinterpret_cast<float *>(sc);
    a.partial = reinterpret_cast<float *>(sc + unr_glr_bytes(h, batch));
}

}  // namespace nrldpc

#ifndef NRLDPC_CPU_EMU
// ---------------------------------------------------------------------------
// C ABI (include/nr_ldpc_mi355.h, "unrolled trainable min-sum")
// ---------------------------------------------------------------------------
#include <stdio.h>

#include <vector>

#include "../../include/nr_ldpc_mi355.h"
#include "ldpc_host.h"

struct ldpc_mi355_unrolled {
    nrldpc::UnrolledShape shape;
    int B;
    uint32_t *d_etab;
};

namespace {

using namespace nrldpc;

#define UNR_HIP(expr)                                                                 \
    do {                                                                              \
        hipError_t _e = (expr);                                                       \
        if (_e != hipSuccess) {                                                       \
            fprintf(stderr, "libnrldpc_mi355: %s failed: %s\n", #expr,                \
                    hipGetErrorString(_e));                                           \
            return -1;                                                                \
        }                                                                             \
    } while (0)

hipError_t configure_unrolled() {
    static hipError_t status = [] {
        hipError_t e = hipSuccess;
        const void *fns[] = {reinterpret_cast<const void *>(&unrolled_fwd_kernel<1>),
                             reinterpret_cast<const void *>(&unrolled_fwd_kernel<2>),
                             reinterpret_cast<const void *>(&unrolled_bwd_kernel<1>),
                             reinterpret_cast<const void *>(&unrolled_bwd_kernel<2>)};
        for (const void *f : fns) {
            e = hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize,
                                    kUnrMaxLdsBytes);
            if (e != hipSuccess) break;
        }
        return e;
    }();
    return status;
}

}  // namespace

extern "C" int32_t ldpc_mi355_unrolled_create(ldpc_mi355_unrolled_t **handle,
                                              const ldpc_in_config_t *cfg, int32_t iterations,
                                              float clip_llr, int32_t flags) {
    if (!handle || !cfg) return -4;
    *handle = nullptr;
    Params p;
    if (const char *why = derive(*cfg, p)) {
        print_invalid(*cfg, why);
        return -2;
    }
    if (iterations < 0 || iterations > 1024 || (flags & ~LDPC_MI355_UNROLLED_DETACH) != 0)
        return -4;
    UNR_HIP(configure_unrolled());
    auto *h = new ldpc_mi355_unrolled{};
    h->shape = unr_shape(p.bgn, p.Zc, p.N, iterations, clip_llr,
                         (flags & LDPC_MI355_UNROLLED_DETACH) != 0);
    h->B = p.B;
    std::vector<uint32_t> tab(h->shape.edges);
    edge_table(p.bgn, p.Zc, lifting_set(p.Zc), tab.data());
    if (hipMalloc(&h->d_etab, tab.size() * sizeof(uint32_t)) != hipSuccess ||
        hipMemcpy(h->d_etab, tab.data(), tab.size() * sizeof(uint32_t),
                  hipMemcpyHostToDevice) != hipSuccess) {
        fprintf(stderr, "libnrldpc_mi355: unrolled edge table upload failed\n");
        if (h->d_etab) (void)hipFree(h->d_etab);
        delete h;
        return -1;
    }
    *handle = h;
    return 0;
}

extern "C" void ldpc_mi355_unrolled_destroy(ldpc_mi355_unrolled_t *h) {
    if (!h) return;
    if (h->d_etab) (void)hipFree(h->d_etab);
    delete h;
}

extern "C" int32_t ldpc_mi355_unrolled_dims(const ldpc_mi355_unrolled_t *h, int32_t *dims) {
    if (!h || !dims) return -4;
    const UnrolledShape &s = h->shape;
    dims[0] = s.n;
    dims[1] = s.rows;
    dims[2] = s.edges;
    dims[3] = s.Zc;
    dims[4] = h->B;
    dims[5] = s.T;
    dims[6] = s.cbs;
    dims[7] = s.lpc;
    return 0;
}

extern "C" int32_t ldpc_mi355_unrolled_set_codeblocks_per_workgroup(ldpc_mi355_unrolled_t *h,
                                                                    int32_t k) {
    if (!h || k < 0) return -4;
    if (k == 0) {
        h->shape.cbs = unr_default_cbs(h->shape);
        return 0;
    }
    if (!unr_cbs_fits(h->shape, k)) return -4;
    h->shape.cbs = k;
    return 0;
}

extern "C" uint64_t ldpc_mi355_unrolled_record_bytes(const ldpc_mi355_unrolled_t *h,
                                                     int32_t batch) {
    return (!h || batch < 0) ? 0 : (uint64_t)unr_record_bytes(h->shape, batch);
}

extern "C" uint64_t ldpc_mi355_unrolled_scratch_bytes(const ldpc_mi355_unrolled_t *h,
                                                      int32_t batch) {
    return (!h || batch < 0) ? 0 : (uint64_t)unr_scratch_bytes(h->shape, batch);
}

extern "C" int32_t ldpc_mi355_unrolled_forward(const ldpc_mi355_unrolled_t *h, int32_t batch,
                                               const float *llr, const float *alpha,
                                               const float *beta, float *lq_out,
                                               float *iter_lq_out, void *record, void *stream) {
    if (!h || batch < 0) return -4;
    if (batch == 0) return 0;
    if (!llr || !lq_out || (h->shape.T > 0 && (!alpha || !beta))) return -4;
    const UnrolledShape &sh = h->shape;
    UnrolledArgs a = unr_args(sh, batch, h->d_etab);
    a.llr = llr;
    a.lq = lq_out;
    a.iter_lq = iter_lq_out;
    a.rec = static_cast<uint32_t *>(record);
    a.alpha = alpha;
    a.beta = beta;
    const dim3 grid((batch + sh.cbs - 1) / sh.cbs), block(sh.cbs * sh.lpc);
    const size_t lds = (size_t)sh.cbs * a.stride * sizeof(float);
    auto s = static_cast<hipStream_t>(stream);
    if (sh.bgn == 1)
        hipLaunchKernelGGL(unrolled_fwd_kernel<1>, grid, block, lds, s, a);
    else
        hipLaunchKernelGGL(unrolled_fwd_kernel<2>, grid, block, lds, s, a);
    UNR_HIP(hipGetLastError());
    return 0;
}

extern "C" int32_t ldpc_mi355_unrolled_backward(const ldpc_mi355_unrolled_t *h, int32_t batch,
                                                const float *llr, const float *alpha,
                                                const float *beta, const void *record,
                                                const float *g_lq, const float *g_iter_lq,
                                                float *g_llr, float *g_alpha, float *g_beta,
                                                void *scratch, void *stream) {
    if (!h || batch < 0) return -4;
    const UnrolledShape &sh = h->shape;
    const bool params = sh.T > 0;
    if (params && (!alpha || !beta || !g_alpha || !g_beta)) return -4;
    if (batch > 0 && (!llr || !g_lq || (params && (!record || !scratch)))) return -4;
    auto s = static_cast<hipStream_t>(stream);
    if (batch == 0) {
        // no codeblock: the parameter gradients are zero (a kernel, not a
        // memset node, so the call stays capturable)
        if (params) {
            hipLaunchKernelGGL(unrolled_reduce_kernel, dim3(sh.T * sh.rows),
                               dim3(kReduceThreads), 0, s, nullptr, 0, g_alpha, g_beta);
            UNR_HIP(hipGetLastError());
        }
        return 0;
    }
    UnrolledArgs a = unr_args(sh, batch, h->d_etab);
    a.llr = llr;
    a.rec = const_cast<uint32_t *>(static_cast<const uint32_t *>(record));
    a.alpha = alpha;
    a.beta = beta;
    a.g_lq = g_lq;
    a.g_iter = g_iter_lq;
    a.g_llr = g_llr;
    if (params) unr_bind_scratch(sh, batch, scratch, a);
    const dim3 grid((batch + sh.cbs - 1) / sh.cbs), block(sh.cbs * sh.lpc);
    const size_t lds = (size_t)sh.cbs * a.stride * sizeof(float);
    if (sh.bgn == 1)
        hipLaunchKernelGGL(unrolled_bwd_kernel<1>, grid, block, lds, s, a);
    else
        hipLaunchKernelGGL(unrolled_bwd_kernel<2>, grid, block, lds, s, a);
    UNR_HIP(hipGetLastError());
    if (params) {
        const int m = batch * (sh.lpc / sh.seg);
        hipLaunchKernelGGL(unrolled_reduce_kernel, dim3(sh.T * sh.rows), dim3(kReduceThreads), 0,
                           s, a.partial, m, g_alpha, g_beta);
        UNR_HIP(hipGetLastError());
    }
    return 0;
}
#endif  // NRLDPC_CPU_EMU
