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
//     left it.  tests/test_codegen.py checks the compiled kernels for calls,
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
// the parallel CRC after each (:286-294).  Inlined into both kernels below.
__device__ __forceinline__ void bp_decode(const CbDesc &d, const DecOut &o, const DecParams &p,
                                          float *lq, uint32_t *crc_tab, uint32_t *slots) {
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

// Codeblock blockIdx.x of a batch, its descriptor in o.desc.
__global__ void __launch_bounds__(kMaxThreads) ldpc_bp_kernel(DecOut o, DecParams p) {
    NRLDPC_DYNAMIC_LDS(smem);
    __shared__ uint32_t crc_tab[256];
    __shared__ uint32_t slots[kMaxThreads / 64];
    const CbDesc d = o.desc[blockIdx.x];
    bp_decode(d, o, p, reinterpret_cast<float *>(smem), crc_tab, slots);
}

// Codeblock i = blockIdx.x of a batch of one config: the descriptor is
// `base` with its buffers at codeblock i's place -- LLRs at i * E, hard bytes
// at i * B / 8, soft slot i (when base has one), posteriors at i * (N + 2Zc),
// messages at i * nnz * Zc.
__global__ void __launch_bounds__(kMaxThreads)
ldpc_bp_uniform_kernel(DecOut o, DecParams p, CbDesc base, uint32_t edges) {
    NRLDPC_DYNAMIC_LDS(smem);
    __shared__ uint32_t crc_tab[256];
    __shared__ uint32_t slots[kMaxThreads / 64];
    const uint32_t i = blockIdx.x;
    CbDesc d = base;
    d.llr_off = i * d.E;
    d.hard_off = i * (d.B / 8u);
    if (base.soft_off != kNone) d.soft_off = i * (uint32_t)kMaxSoft;
    if (base.lq_off != kNone) d.lq_off = i * (uint32_t)(d.N + 2 * d.Zc);
    d.work_off = i * edges * d.Zc;
    bp_decode(d, o, p, reinterpret_cast<float *>(smem), crc_tab, slots);
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

// ---------------------------------------------------------------------------
// Allocation-free BP decode (ldpc_mi355_bp_decode_noalloc).  The messages
// live in `scratch`, a file-scope device array (one per GPU, loaded with the
// code object and held until the process ends), and the edge and CRC tables
// in `noalloc_tables`.  In host code these names are host-side stand-ins,
// not device addresses: every host use goes through the per-device pointers
// from hipGetSymbolAddress.
// ---------------------------------------------------------------------------
#include <stdio.h>

#include <mutex>
#include <vector>

#include "ldpc_host.h"

namespace nrldpc {
constexpr size_t kScratchFloats = (size_t)8 << 20;   // 32 MiB of float messages
constexpr int kNumLiftingSizes = 51;
constexpr size_t kNoAllocTableWords =
    (size_t)kNumLiftingSizes * (NR_BG1_EDGES + NR_BG2_EDGES) + 4 * (size_t)kCrcTabWords;
}  // namespace nrldpc

static __device__ __align__(256) float scratch[nrldpc::kScratchFloats];
static __device__ __align__(256) uint32_t noalloc_tables[nrldpc::kNoAllocTableWords];

namespace {
using namespace nrldpc;

// Host image of noalloc_tables (the same on every GPU).
struct NoAllocTables {
    std::vector<uint32_t> words;
    uint32_t etab_off[2][kMaxZc + 1] = {};
    uint32_t crc_off = 0;
    NoAllocTables() {
        int nz = 0;
        const uint16_t *zs = all_lifting_sizes(&nz);
        for (int bg = 1; bg <= 2; ++bg) {
            const int ne = bg == 1 ? NR_BG1_EDGES : NR_BG2_EDGES;
            for (int i = 0; i < nz; ++i) {
                etab_off[bg - 1][zs[i]] = (uint32_t)words.size();
                words.resize(words.size() + ne);
                edge_table(bg, zs[i], lifting_set(zs[i]), words.data() + etab_off[bg - 1][zs[i]]);
            }
        }
        crc_off = (uint32_t)words.size();
        words.resize(words.size() + 4 * (size_t)kCrcTabWords);
        crc_device_tables(words.data() + crc_off);
    }
};

const NoAllocTables &noalloc_host_tables() {
    static const NoAllocTables t;
    return t;
}

// Device addresses of scratch and noalloc_tables on one GPU.
struct NoAllocDevice {
    bool ready = false;
    float *scratch = nullptr;
    uint32_t *tables = nullptr;
};
constexpr int kNoAllocMaxDevices = 64;
std::mutex g_noalloc_mu;
NoAllocDevice g_noalloc[kNoAllocMaxDevices];

#define NOALLOC_HIP(expr)                                                              \
    do {                                                                               \
        const hipError_t _e = (expr);                                                  \
        if (_e != hipSuccess) {                                                        \
            fprintf(stderr, "libnrldpc_mi355: %s failed: %s\n", #expr,                \
                    hipGetErrorString(_e));                                            \
            return nullptr;                                                            \
        }                                                                              \
    } while (0)

// Once per GPU, with that GPU current: the symbol addresses, the tables and
// the kernel's LDS limit.
const NoAllocDevice *noalloc_device(int dev) {
    if (dev < 0 || dev >= kNoAllocMaxDevices) return nullptr;
    std::lock_guard<std::mutex> guard(g_noalloc_mu);
    NoAllocDevice &d = g_noalloc[dev];
    if (d.ready) return &d;
    const NoAllocTables &t = noalloc_host_tables();
    if (t.words.size() != kNoAllocTableWords) {
        fprintf(stderr, "libnrldpc_mi355: table layout mismatch\n");
        return nullptr;
    }
    void *ps = nullptr, *pt = nullptr;
    NOALLOC_HIP(hipGetSymbolAddress(&ps, HIP_SYMBOL(scratch)));
    NOALLOC_HIP(hipGetSymbolAddress(&pt, HIP_SYMBOL(noalloc_tables)));
    NOALLOC_HIP(hipMemcpy(pt, t.words.data(), t.words.size() * sizeof(uint32_t),
                          hipMemcpyHostToDevice));
    NOALLOC_HIP(hipFuncSetAttribute(reinterpret_cast<const void *>(&ldpc_bp_uniform_kernel),
                                    hipFuncAttributeMaxDynamicSharedMemorySize,
                                    160 * 1024 - 2048));
    d.scratch = static_cast<float *>(ps);
    d.tables = static_cast<uint32_t *>(pt);
    d.ready = true;
    return &d;
}
}  // namespace

extern "C" int32_t ldpc_mi355_bp_decode_noalloc(const ldpc_in_config_t *cfg,
                                                const ldpc_decoder_cfg_t *decoder_config,
                                                int32_t n, const int16_t *llr,
                                                int16_t *soft_pool, uint8_t *hard,
                                                int32_t *status, int32_t *iterations, float *lq,
                                                void *stream) {
    using namespace nrldpc;
    if (!cfg || !decoder_config || n <= 0 || !llr || !hard || !status || !iterations) return -4;
    if (decoder_config->algo != 0) {
        fprintf(stderr, "libnrldpc_mi355: the allocation-free decode runs algo 0 (FP32 BP) only, "
                        "not algo %u\n", decoder_config->algo);
        return -3;
    }
    Params p;
    if (const char *why = derive(*cfg, p)) {
        print_invalid(*cfg, why);
        return -2;
    }
    if (p.harq_on && !soft_pool) {
        fprintf(stderr, "libnrldpc_mi355: HARQ_on needs soft-buffer slots (soft_pool is NULL)\n");
        return -4;
    }
    const size_t per_cb = (size_t)p.edges * (size_t)p.Zc;
    if ((size_t)n * per_cb > kScratchFloats) {
        fprintf(stderr,
                "libnrldpc_mi355: %d codeblocks of BG%d Zc=%d need %zu message floats; the "
                "allocation-free decode holds %zu (at most %zu codeblocks of this code per call)\n",
                n, p.bgn, p.Zc, (size_t)n * per_cb, kScratchFloats, kScratchFloats / per_cb);
        return -4;
    }
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return -1;
    const NoAllocDevice *nd = noalloc_device(dev);
    if (!nd) return -1;
    const NoAllocTables &t = noalloc_host_tables();

    DecOut o;
    o.desc = nullptr;
    o.llr = llr;
    o.soft = soft_pool;
    o.hard = hard;
    o.status = status;
    o.iters = iterations;
    o.lq = lq;
    o.work = nd->scratch;
    o.etab = nd->tables;
    o.crc = nd->tables + t.crc_off;
    DecParams dp = decoder_params(*decoder_config);
    CbDesc base = codeblock_desc(p, t.etab_off[p.bgn - 1][p.Zc]);
    base.llr_off = 0;
    base.hard_off = 0;
    base.work_off = 0;
    base.soft_off = soft_pool ? 0u : kNone;
    base.lq_off = lq ? 0u : kNone;
    uint32_t edges = (uint32_t)p.edges;
    const size_t lds = ((size_t)(p.N + 2 * p.Zc) * sizeof(float) + 15) & ~(size_t)15;
    void *args[] = {&o, &dp, &base, &edges};
    hipError_t e = hipLaunchKernel(reinterpret_cast<const void *>(&ldpc_bp_uniform_kernel),
                                   dim3(n), dim3(workgroup_threads(p.Zc)), args, lds,
                                   reinterpret_cast<hipStream_t>(stream));
    if (e == hipSuccess) e = hipGetLastError();
    if (e != hipSuccess) {
        fprintf(stderr, "libnrldpc_mi355: allocation-free BP launch failed: %s\n",
                hipGetErrorString(e));
        return -1;
    }
    return 0;
}
#endif  // NRLDPC_CPU_EMU
