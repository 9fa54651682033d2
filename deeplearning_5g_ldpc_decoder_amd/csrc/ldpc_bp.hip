// FP32 layered belief-propagation decoder (algo 0) for gfx950.
//
// Reference: c_lib/floatpoint_ldpc_decoder.c:54-151 (the sum-product check
// update with its exact-zero cases) and :237-298 (the layered loop with CRC
// early stop).  Same mapping as the other decode kernels: one workgroup per
// codeblock, lane zx owns lifted check zx of every base-graph row, beliefs LQ
// in LDS (float, N + 2Zc entries), a workgroup barrier after each group of
// column-disjoint rows.
//
// Layout of the code, chosen for what the compiler must produce:
//   * the rows run in a loop over a small row table (first edge, degree,
//     barrier flag), and the check update is instantiated once per row
//     degree (3..10 and 19, shared by both base graphs).  Each instance keeps
//     its per-edge values (extrinsic beliefs, tanh terms, prefix and suffix
//     products) in registers, indexed at compile time only;
//   * everything is inlined into the kernel: the kernel makes no calls, uses
//     no scratch and is small enough that every branch is a short branch.
//     (An earlier form unrolled all 46 rows at compile time, which the
//     compiler outlined into 240-370 KB device functions; in the BG2 one its
//     long-branch expansion reused the return-address registers s[30:31], so
//     the function's return jumped back into its own exit block and never
//     left it.  tests/test_bp_codegen.py checks the compiled kernel for calls,
//     long branches and scratch.)
//
// BP messages have no compressed form, so each codeblock keeps its nnz * Zc
// check-to-variable messages in HBM (DecOut::work), edge-major: edge e of
// lifted check zx at e * Zc + zx, so one edge of a row is one coalesced
// 4 * Zc-byte access over the lanes.
#ifdef NRLDPC_CPU_EMU
#include "hip_cpu_emu.h"
#ifndef NRLDPC_DYNAMIC_LDS
#define NRLDPC_DYNAMIC_LDS(name) unsigned char *name = emu_dynamic_lds()
#endif
#else
#include <hip/hip_runtime.h>
#ifndef NRLDPC_DYNAMIC_LDS
#define NRLDPC_DYNAMIC_LDS(name) extern __shared__ __align__(16) unsigned char name[]
#endif
#endif

#include "ldpc_device.h"
#include "ldpc_graph.h"
#include "ldpc_crc.h"
#include "ldpc_recover.h"

namespace nrldpc {

// Edge-table words read through the constant address space, so uniform reads
// become scalar loads.
#ifdef NRLDPC_CPU_EMU
typedef const uint32_t *bp_etab_ptr;
#else
typedef const __attribute__((address_space(4))) uint32_t *bp_etab_ptr;
#endif

// Row words of the layered schedule, BG1 rows then BG2 rows:
// first edge | degree << 16 | (a barrier follows the row) << 24.
constexpr int kBpRows = Graph<1>::kRows + Graph<2>::kRows;

template <int BG>
constexpr uint32_t bp_row_word_of(int r) {
    return (uint32_t)Graph<BG>::start(r) |
           ((uint32_t)(Graph<BG>::start(r + 1) - Graph<BG>::start(r)) << 16) |
           ((kRowGroups<BG>.barrier_after[r] ? 1u : 0u) << 24);
}

struct BpRowTable {
    uint32_t w[kBpRows];
    constexpr BpRowTable() : w() {
        for (int r = 0; r < Graph<1>::kRows; ++r) w[r] = bp_row_word_of<1>(r);
        for (int r = 0; r < Graph<2>::kRows; ++r) w[Graph<1>::kRows + r] = bp_row_word_of<2>(r);
    }
};

__device__ __forceinline__ uint32_t bp_row_word(int i) {
    static constexpr BpRowTable kTable{};
    return kTable.w[i];
}

// Outgoing message from an extrinsic tanh product
// (c_lib/floatpoint_ldpc_decoder.c:54-62): 2 atanh(x), saturated at
// +-2 * 19.07 where |x| reaches 1.  Device atanhf; a zero product gives +0.
__device__ __forceinline__ float bp_message(float x) {
    const float sat = 2.0f * 19.07f;
    if (x >= 1.0f) return sat;
    if (x <= -1.0f) return -sat;
    return 2.0f * atanhf(x);
}

// One lifted check of degree DEG (c_lib/floatpoint_ldpc_decoder.c:65-151,
// 263-283).  et: the row's edge-table words (column * Zc | shift << 16);
// lr: the row's messages (edge j of this check at j * Zc + zx); first: the
// first iteration, whose old messages are 0.
//   rel_j = LQ[c_j] - old_j;  t_j = tanh(rel_j / 2), exact zeros counted;
//   no zero:    msg_j = m(prefix_j * suffix_(j+1)), products in edge order;
//   one zero:   the zero edge gets m(product of the other t), others +0;
//   two or more zeros: every message is +0;
//   LQ[c_j] = rel_j + msg_j (not clipped).
// The three cases share one form: the extrinsic product is
// prefix_j * suffix_(j+1) when no rel is zero, else the non-zero product on
// the single zero edge and +0 elsewhere, and m(+0) = +0.
template <int DEG>
__device__ __forceinline__ void bp_check(float *__restrict__ lq, float *__restrict__ lr,
                                         bp_etab_ptr et, int zx, int Zc, bool first) {
    int addr[DEG];
    float rel[DEG], t[DEG];
#pragma unroll
    for (int j = 0; j < DEG; ++j) {
        const uint32_t w = et[j];
        int c = zx + (int)(w >> 16);
        c = (c >= Zc) ? c - Zc : c;
        addr[j] = (int)(w & 0xFFFFu) + c;
    }
    float old[DEG];
#pragma unroll
    for (int j = 0; j < DEG; ++j) old[j] = first ? 0.0f : lr[j * Zc + zx];
#pragma unroll
    for (int j = 0; j < DEG; ++j) rel[j] = lq[addr[j]] - old[j];

    int nz = 0, zidx = -1;
    float prod = 1.0f;
#pragma unroll
    for (int j = 0; j < DEG; ++j) {
        const bool z = rel[j] == 0.0f;
        const float th = tanhf(rel[j] / 2.0f);
        t[j] = z ? 0.0f : th;
        nz += z ? 1 : 0;
        zidx = z ? j : zidx;
        prod = z ? prod : prod * th;
    }
    float pre[DEG + 1], suf[DEG + 1];
    pre[0] = 1.0f;
#pragma unroll
    for (int j = 0; j < DEG; ++j) pre[j + 1] = pre[j] * t[j];
    suf[DEG] = 1.0f;
#pragma unroll
    for (int j = DEG - 1; j >= 0; --j) suf[j] = suf[j + 1] * t[j];
#pragma unroll
    for (int j = 0; j < DEG; ++j) {
        const float x = (nz == 0) ? pre[j] * suf[j + 1] : ((nz == 1 && j == zidx) ? prod : 0.0f);
        const float m = bp_message(x);
        lr[j * Zc + zx] = m;
        lq[addr[j]] = rel[j] + m;
    }
}

// The row's check, dispatched on its (uniform) degree: the degrees of BG1
// and BG2 are 3..10 and 19.
__device__ __forceinline__ void bp_row(int deg, float *lq, float *lr, bp_etab_ptr et, int zx,
                                       int Zc, bool first) {
    switch (deg) {
    case 3: bp_check<3>(lq, lr, et, zx, Zc, first); break;
    case 4: bp_check<4>(lq, lr, et, zx, Zc, first); break;
    case 5: bp_check<5>(lq, lr, et, zx, Zc, first); break;
    case 6: bp_check<6>(lq, lr, et, zx, Zc, first); break;
    case 7: bp_check<7>(lq, lr, et, zx, Zc, first); break;
    case 8: bp_check<8>(lq, lr, et, zx, Zc, first); break;
    case 9: bp_check<9>(lq, lr, et, zx, Zc, first); break;
    case 10: bp_check<10>(lq, lr, et, zx, Zc, first); break;
    default: bp_check<19>(lq, lr, et, zx, Zc, first); break;
    }
}

// One codeblock per workgroup (blockDim: Zc rounded up to whole waves).
// Fused rate recovery + HARQ combine into float beliefs (/128,
// c_lib/floatpoint_ldpc_decoder.c:251-254), then iterations of all rows with
// the parallel CRC after each (:286-294).
__global__ void __launch_bounds__(kMaxThreads) ldpc_bp_kernel(DecOut o, DecParams p) {
    NRLDPC_DYNAMIC_LDS(smem);
    __shared__ uint32_t crc_tab[256];
    __shared__ uint32_t slots[kMaxThreads / 64];
    float *lq = reinterpret_cast<float *>(smem);

    const CbDesc d = o.desc[blockIdx.x];
    const int n = d.N + 2 * d.Zc;
    const int t = threadIdx.x, nt = blockDim.x;
    const int16_t *llr = o.llr + d.llr_off;
    int16_t *soft = (d.soft_off != kNone) ? o.soft + d.soft_off : nullptr;
    recover_fill<float>(lq, d, llr, soft, p.filler, t, nt, n);
    const CrcCtx crc = crc_setup_tab(d.crctype, d.B / 8, crc_tab, o.crc);
    __syncthreads();

    int status = 1, it = 0;
    if (p.max_it > 0) {
        float *lr = o.work + d.work_off;
        const bp_etab_ptr et = (bp_etab_ptr)(o.etab + d.etab_off);
        const int Zc = d.Zc;
        const int row0 = d.bgn == 1 ? 0 : Graph<1>::kRows;
        const int rows = d.bgn == 1 ? Graph<1>::kRows : Graph<2>::kRows;
        while (it < p.max_it) {
            for (int r = 0; r < rows; ++r) {
                const uint32_t w = bp_row_word(row0 + r);
                const int e0 = (int)(w & 0xFFFFu);
                if (t < Zc) bp_row((int)((w >> 16) & 0xFFu), lq, lr + e0 * Zc, et + e0, t, Zc, it == 0);
                if (w >> 24) __syncthreads();
            }
            ++it;
            if (block_crc(lq, crc, crc_tab, slots) == 0u) {
                status = 0;
                break;
            }
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
        float *out = reinterpret_cast<float *>(o.lq) + d.lq_off;
        for (int i = t; i < n; i += nt) out[i] = lq[i];
    }
}

}  // namespace nrldpc

#ifndef NRLDPC_CPU_EMU
// Host launchers (C linkage inside the library only).
extern "C" hipError_t nrldpc_configure_bp() {
    return hipFuncSetAttribute(reinterpret_cast<const void *>(&nrldpc::ldpc_bp_kernel),
                               hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024 - 2048);
}

// ncb codeblocks, one per workgroup of `threads` lanes; max_soft: the largest
// N + 2Zc of the launch.  o->work must hold every codeblock's nnz * Zc
// messages at its work_off.
extern "C" hipError_t nrldpc_launch_bp(const nrldpc::DecOut *o, const nrldpc::DecParams *p,
                                       int ncb, int threads, int max_soft, hipStream_t stream) {
    if (!o->work) return hipErrorInvalidValue;
    const size_t lds = ((size_t)max_soft * sizeof(float) + 15) & ~(size_t)15;
    nrldpc::DecOut ov = *o;
    nrldpc::DecParams pv = *p;
    void *args[] = {&ov, &pv};
    const hipError_t e = hipLaunchKernel(reinterpret_cast<const void *>(&nrldpc::ldpc_bp_kernel),
                                         dim3(ncb), dim3(threads), args, lds, stream);
    if (e != hipSuccess) return e;
    return hipGetLastError();
}
#endif  // NRLDPC_CPU_EMU
