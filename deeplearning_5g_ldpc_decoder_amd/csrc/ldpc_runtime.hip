// Device runtime of libnrldpc_mi355.so: device context, the per-codeblock
// reference ABI (c_ldpc_5g_decode) and batched decode plans.
#include <hip/hip_runtime.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <atomic>
#include <mutex>
#include <vector>

#include "../../include/nr_ldpc_mi355.h"
#include "ldpc_device.h"
#include "ldpc_host.h"

using namespace nrldpc;

extern "C" hipError_t nrldpc_launch_decode(int algo, const DecOut *o, const DecParams *p,
                                           int ncb, int threads, int max_soft, uint32_t uniform,
                                           hipStream_t stream);
extern "C" hipError_t nrldpc_configure_kernels();
extern "C" hipError_t nrldpc_launch_edge_lanes(const DecOut *o, const DecParams *p, int ncb,
                                               int max_soft, hipStream_t stream);
extern "C" int nrldpc_edge_lanes_max_zc(int bgn);
extern "C" hipError_t nrldpc_launch_synth(int bgn, const SynthArgs *a, hipStream_t s);
extern "C" hipError_t nrldpc_launch_tiles(int algo, const DecOut *o, const TileDesc *tiles,
                                          const uint32_t *order, const DecParams *p, int ntiles,
                                          int lds_entries, uint32_t *queue, hipStream_t stream);

namespace {

#define HIP_OK(expr)                                                                  \
    do {                                                                              \
        hipError_t _e = (expr);                                                       \
        if (_e != hipSuccess) {                                                       \
            fprintf(stderr, "libnrldpc_mi355: %s failed: %s (%s:%d)\n", #expr,        \
                    hipGetErrorString(_e), __FILE__, __LINE__);                       \
            return -1;                                                                \
        }                                                                             \
    } while (0)

constexpr int kStatusBytes = 16;                  // status, iterations
constexpr int kHardBytes = 8448 / 8 + 16;         // B/8 max, padded
constexpr size_t kSoftBytes = (size_t)kMaxSoft * 2;

// Per-codeblock staging layouts (one H2D and one D2H copy per call).
constexpr size_t kScaleBytes =
    (size_t)kMaxTrainedIterations * (BG1_NUM_ROWS + BG2_NUM_ROWS) * sizeof(uint32_t);
struct InLayout {
    static constexpr size_t desc = 0;
    static constexpr size_t llr = 64;
    static constexpr size_t soft = llr + 65536 * 2;
    static constexpr size_t scale = soft + kSoftBytes;
    static constexpr size_t total = scale + kScaleBytes;
};
struct OutLayout {
    static constexpr size_t status = 0;
    static constexpr size_t hard = kStatusBytes;
    static constexpr size_t soft = hard + kHardBytes;
    static constexpr size_t lq = soft + kSoftBytes;
    static constexpr size_t total = lq + (size_t)kMaxSoft * 4;
};

struct Runtime {
    std::mutex mu;             // serializes the per-codeblock path
    std::once_flag once;
    int init_status = -1;
    int device = 0;
    hipStream_t stream = nullptr;
    uint32_t *d_etab = nullptr;        // edge tables, then the CRC tables
    uint32_t etab_off[2][kMaxZc + 1] = {};
    uint32_t crc_off = 0;
    unsigned char *h_in = nullptr, *h_out = nullptr;   // pinned
    unsigned char *d_in = nullptr, *d_out = nullptr;
    float *d_work = nullptr;                           // BP messages, one codeblock

    int init() {
        HIP_OK(hipGetDevice(&device));
        HIP_OK(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
        HIP_OK(nrldpc_configure_kernels());
        int nz = 0;
        const uint16_t *zs = all_lifting_sizes(&nz);
        std::vector<uint32_t> tab;
        for (int bg = 1; bg <= 2; ++bg) {
            const int ne = edge_block_words(bg);
            for (int i = 0; i < nz; ++i) {
                const int zc = zs[i];
                etab_off[bg - 1][zc] = (uint32_t)tab.size();
                tab.resize(tab.size() + ne);
                edge_block(bg, zc, lifting_set(zc), tab.data() + etab_off[bg - 1][zc]);
            }
        }
        crc_off = (uint32_t)tab.size();
        tab.resize(tab.size() + 4 * (size_t)kCrcTabWords);
        crc_device_tables(tab.data() + crc_off);
        HIP_OK(hipMalloc(&d_etab, tab.size() * sizeof(uint32_t)));
        HIP_OK(hipMemcpy(d_etab, tab.data(), tab.size() * sizeof(uint32_t),
                         hipMemcpyHostToDevice));
        HIP_OK(hipHostMalloc(&h_in, InLayout::total, hipHostMallocDefault));
        HIP_OK(hipHostMalloc(&h_out, OutLayout::total, hipHostMallocDefault));
        HIP_OK(hipMalloc(&d_in, InLayout::total));
        HIP_OK(hipMalloc(&d_out, OutLayout::total));
        HIP_OK(hipMalloc(&d_work, (size_t)BG1_MAX_NON_NULL_ELEMENTS * kMaxZc * sizeof(float)));
        return 0;
    }

    int ensure() {
        std::call_once(once, [this] { init_status = init(); });
        return init_status;
    }
};

Runtime &rt() {
    static Runtime r;
    return r;
}

bool algo_supported(int algo) { return algo >= 0 && algo <= 4; }

// Algos 1-3 decode in int16 (posteriors int16); 0 and 4 in FP32.
bool i16_algo(int algo) { return algo >= 1 && algo <= 3; }
bool table_algo(int algo) { return algo == 2 || algo == 3; }

void report_table_iterations(int bgn, int max_it, int table_iters) {
    fprintf(stderr,
            "libnrldpc_mi355: algo 2: max_iterations %d exceeds the %d iterations of the BG%d "
            "trained table\n",
            max_it, table_iters, bgn);
}

// Points the kernel parameters at scale words uploaded to dev (algos 2/3).
void bind_scale(DecParams &dp, const ScaleSet &s, const uint32_t *dev) {
    dp.scale[0] = dev + s.off[0];
    dp.scale[1] = dev + s.off[1];
    dp.scale_iters = s.iters;
    dp.general = s.fast ? 0 : 1;
    dp.simple = (s.fast && s.simple) ? 1 : 0;
}

DecParams make_params(const ldpc_decoder_cfg_t &c) { return decoder_params(c); }

void report_missing_soft_pool(int max_slot) {
    fprintf(stderr,
            "libnrldpc_mi355: codeblocks use soft-buffer slots up to %d but soft_pool is NULL\n",
            max_slot);
}

CbDesc make_desc(const Params &p, const Runtime &r) {
    return codeblock_desc(p, r.etab_off[p.bgn - 1][p.Zc]);
}

int threads_for(int zc) { return workgroup_threads(zc); }

uint32_t uniform_pair(int bgn, int zc) { return ((uint32_t)bgn << 16) | (uint32_t)zc; }

}  // namespace

// ---------------------------------------------------------------------------
// Per-codeblock path (the reference ABI): one H2D, one launch, one D2H.
// ---------------------------------------------------------------------------
namespace {
// The reference's c_lib and CUDA decoder wrappers differ only in which algo
// IDs they accept and what they run for them; everything else is this path.
enum class AlgoMap { kCLib, kCuda };

int32_t decode_one(int16_t *LLr, uint8_t *hardbits, ldpc_in_config_t *inconfig,
                   ldpc_decoder_cfg_t *decoder_config, int32_t *iterations, void *posteriors,
                   AlgoMap map);
}  // namespace

extern "C" int32_t ldpc_mi355_decode_ex(int16_t *LLr, uint8_t *hardbits,
                                        ldpc_in_config_t *inconfig,
                                        ldpc_decoder_cfg_t *decoder_config, int32_t *iterations,
                                        void *posteriors) {
    return decode_one(LLr, hardbits, inconfig, decoder_config, iterations, posteriors,
                      AlgoMap::kCLib);
}

extern "C" int32_t c_ldpc_5g_decode(int16_t *LLr, uint8_t *hardbits, ldpc_in_config_t *inconfig,
                                    ldpc_decoder_cfg_t *decoder_config) {
    return decode_one(LLr, hardbits, inconfig, decoder_config, nullptr, nullptr, AlgoMap::kCLib);
}

// CUDA/cuda_ldpc_5g_decoder.cu:31-58: algo 1 is the int16 min-sum, algo 4 the
// FP32 *BP* decoder (CUDA/floatpoint_ldpc_decoder.cu), anything else -3 after
// rate recovery.
extern "C" int32_t cuda_ldpc_5g_decode(int16_t *LLr, uint8_t *hardbits,
                                       ldpc_in_config_t *inconfig,
                                       ldpc_decoder_cfg_t *decoder_config) {
    return decode_one(LLr, hardbits, inconfig, decoder_config, nullptr, nullptr, AlgoMap::kCuda);
}

namespace {
int32_t decode_one(int16_t *LLr, uint8_t *hardbits, ldpc_in_config_t *inconfig,
                   ldpc_decoder_cfg_t *decoder_config, int32_t *iterations, void *posteriors,
                   AlgoMap map) {
    if (iterations) *iterations = 0;
    Params p;
    if (const char *why = derive(*inconfig, p)) {
        print_invalid(*inconfig, why);
        return -2;
    }
    Runtime &r = rt();
    if (r.ensure() != 0) return -1;
    std::lock_guard<std::mutex> guard(r.mu);

    const int asked = decoder_config->algo;
    const bool valid_algo =
        map == AlgoMap::kCLib ? algo_supported(asked) : (asked == 1 || asked == 4);
    const int algo = !valid_algo ? 1 : (map == AlgoMap::kCuda && asked == 4) ? 0 : asked;
    ScaleSet scale;
    if (table_algo(algo)) {
        make_scale_set(*decoder_config, scale);
        if (!scale_iterations_ok(scale, p.bgn, decoder_config->max_iterations)) {
            report_table_iterations(p.bgn, decoder_config->max_iterations,
                                    scale.table_iters[p.bgn - 1]);
            return -2;
        }
    }
    DecParams dp = make_params(*decoder_config);
    if (!valid_algo) dp.max_it = 0;   // rate recovery only, like the reference
    const int n = p.N + 2 * p.Zc;

    CbDesc d = make_desc(p, r);
    d.llr_off = 0;
    d.hard_off = 0;
    d.soft_off = 0;
    d.lq_off = (posteriors && valid_algo) ? 0u : kNone;
    d.work_off = 0;
    memcpy(r.h_in + InLayout::desc, &d, sizeof d);
    memcpy(r.h_in + InLayout::llr, LLr, (size_t)p.E * 2);
    if (p.harq_on) memcpy(r.h_in + InLayout::soft, HARQ_memory, (size_t)n * 2);
    const size_t in_bytes = p.harq_on ? InLayout::soft + (size_t)n * 2
                                      : InLayout::llr + (size_t)p.E * 2;
    HIP_OK(hipMemcpyAsync(r.d_in, r.h_in, in_bytes, hipMemcpyHostToDevice, r.stream));
    if (table_algo(algo)) {
        const size_t sb = scale.words.size() * sizeof(uint32_t);
        memcpy(r.h_in + InLayout::scale, scale.words.data(), sb);
        HIP_OK(hipMemcpyAsync(r.d_in + InLayout::scale, r.h_in + InLayout::scale, sb,
                              hipMemcpyHostToDevice, r.stream));
        bind_scale(dp, scale, reinterpret_cast<const uint32_t *>(r.d_in + InLayout::scale));
    }
    // the kernel reads the soft buffer only when HARQ_on; it always writes it
    DecOut o;
    o.desc = reinterpret_cast<const CbDesc *>(r.d_in + InLayout::desc);
    o.llr = reinterpret_cast<const int16_t *>(r.d_in + InLayout::llr);
    o.soft = reinterpret_cast<int16_t *>(r.d_in + InLayout::soft);
    o.hard = r.d_out + OutLayout::hard;
    o.status = reinterpret_cast<int32_t *>(r.d_out + OutLayout::status);
    o.iters = o.status + 1;
    o.lq = r.d_out + OutLayout::lq;
    o.work = r.d_work;
    o.etab = r.d_etab;
    o.crc = r.d_etab + r.crc_off;
    HIP_OK(nrldpc_launch_decode(algo, &o, &dp, 1, threads_for(p.Zc), n,
                                uniform_pair(p.bgn, p.Zc), r.stream));
    // results: status/iterations/hard contiguous, then the soft buffer
    HIP_OK(hipMemcpyAsync(r.h_out, r.d_out, OutLayout::soft, hipMemcpyDeviceToHost, r.stream));
    HIP_OK(hipMemcpyAsync(r.h_out + OutLayout::soft, r.d_in + InLayout::soft, (size_t)n * 2,
                          hipMemcpyDeviceToHost, r.stream));
    const size_t lq_bytes = (size_t)n * (i16_algo(algo) ? 2 : 4);
    if (d.lq_off != kNone)
        HIP_OK(hipMemcpyAsync(r.h_out + OutLayout::lq, r.d_out + OutLayout::lq, lq_bytes,
                              hipMemcpyDeviceToHost, r.stream));
    HIP_OK(hipStreamSynchronize(r.stream));

    int32_t st, its;
    memcpy(&st, r.h_out + OutLayout::status, 4);
    memcpy(&its, r.h_out + OutLayout::status + 4, 4);
    memcpy(HARQ_memory, r.h_out + OutLayout::soft, (size_t)n * 2);
    if (!valid_algo) {
        if (map == AlgoMap::kCLib)
            fprintf(stderr, "Invalid decoding algorithm selection: %u\n", decoder_config->algo);
        else
            fprintf(stderr, "Unsupported CUDA LDPC decoding algorithm selection: %u\n",
                    decoder_config->algo);
        return -3;
    }
    if (its > 0) memcpy(hardbits, r.h_out + OutLayout::hard, (size_t)p.B / 8);
    if (iterations) *iterations = its;
    if (posteriors) memcpy(posteriors, r.h_out + OutLayout::lq, lq_bytes);
    return st;
}

}  // namespace

// ---------------------------------------------------------------------------
// Batched plans
// ---------------------------------------------------------------------------
struct ldpc_mi355_plan {
    int n = 0, algo = 1, threads = 64, max_soft = 0;
    int max_slot = -1;             // largest soft-buffer slot of its codeblocks, -1 for none
    uint32_t uniform = 0;          // (bgn << 16) | Zc of a single-size batch, else 0
    DecParams dp;
    CbDesc *d_desc = nullptr;
    float *d_work = nullptr;
    uint32_t *d_scale = nullptr;   // algos 2/3: scale words (ScaleSet)
    uint64_t hard_bytes = 0, lq_elems = 0;
    std::vector<uint32_t> hard_off, lq_off;
    bool want_lq = false;
    // packed tiles (algo 1): several codeblocks of one (BG, Zc) per workgroup
    int ntiles = 0, tile_lds = 0;
    TileDesc *d_tiles = nullptr;
    uint32_t *d_order = nullptr;
    // persistent tile launches (flags bit 4): kQueueRing queues of two words,
    // zeroed at creation and rearmed by each launch's last workgroup; launch
    // k uses queue k mod kQueueRing
    static constexpr int kQueueRing = 64;
    uint32_t *d_queue = nullptr;
    mutable std::atomic<uint32_t> launches{0};
    // lane-per-edge schedule (flags bit 5, when every codeblock qualifies)
    bool edge_lanes = false;
};

namespace {
// Codeblocks per packed tile, by (BG, Zc); 0 = default (as many as fit).
int g_tile_k[2][kMaxZc + 1] = {};
std::mutex g_tile_mu;

int tile_codeblocks(int bgn, int zc) {
    std::lock_guard<std::mutex> g(g_tile_mu);
    const int cap = std::min(kMaxThreads / zc, kMaxTileCodeblocks);
    const int k = g_tile_k[bgn - 1][zc];
    return k > 0 ? std::min(k, cap) : cap;
}

}  // namespace

extern "C" int32_t ldpc_mi355_set_tile_codeblocks(int32_t bgn, int32_t zc, int32_t k) {
    if ((bgn != 1 && bgn != 2) || zc < 2 || zc > kMaxZc || lifting_set(zc) < 0 || k < 0)
        return -1;
    std::lock_guard<std::mutex> g(g_tile_mu);
    g_tile_k[bgn - 1][zc] = k;
    return 0;
}

extern "C" int32_t ldpc_mi355_plan_tiles(const ldpc_mi355_plan_t *plan) {
    return plan ? plan->ntiles : -1;
}

namespace {

// Descriptors of a batch as the kernels read them, laid out back to back in
// the hard / posterior outputs.  Returns 0 or -2 (config of codeblock
// *bad_index is invalid).
struct BatchLayout {
    std::vector<CbDesc> descs;
    std::vector<uint32_t> hard_off, lq_off;
    uint64_t hard_bytes = 0, lq_elems = 0, work = 0;
    int max_zc = 2, max_soft = 0;
    int max_slot = -1;      // largest soft-buffer slot a codeblock uses, -1 for none
    uint32_t uniform = 0;   // (bgn << 16) | Zc when every codeblock shares them, else 0
};

int build_layout(const ldpc_mi355_cb_t *cbs, int n, int algo, bool want_lq, const Runtime &r,
                 const ScaleSet *scale, int max_it, BatchLayout &L, int32_t *bad_index) {
    L.descs.resize((size_t)n);
    L.hard_off.resize((size_t)n);
    L.lq_off.resize((size_t)n);
    L.hard_bytes = L.lq_elems = L.work = 0;
    L.max_zc = 2;
    L.max_soft = 0;
    L.max_slot = -1;
    for (int i = 0; i < n; ++i) {
        Params p;
        if (const char *why = derive(cbs[i].cfg, p)) {
            print_invalid(cbs[i].cfg, why);
            if (bad_index) *bad_index = i;
            return -2;
        }
        if (scale && !scale_iterations_ok(*scale, p.bgn, max_it)) {
            report_table_iterations(p.bgn, max_it, scale->table_iters[p.bgn - 1]);
            if (bad_index) *bad_index = i;
            return -2;
        }
        CbDesc d = make_desc(p, r);
        d.llr_off = cbs[i].llr_offset;
        d.hard_off = (uint32_t)L.hard_bytes;
        d.soft_off = cbs[i].soft_slot >= 0 ? (uint32_t)cbs[i].soft_slot * (uint32_t)kMaxSoft
                                           : kNone;
        d.lq_off = want_lq ? (uint32_t)L.lq_elems : kNone;
        d.work_off = (uint32_t)L.work;
        L.hard_off[i] = d.hard_off;
        L.lq_off[i] = want_lq ? (uint32_t)L.lq_elems : 0u;
        L.hard_bytes += (uint64_t)p.B / 8;
        L.lq_elems += (uint64_t)(p.N + 2 * p.Zc);
        if (algo == 0) L.work += (uint64_t)p.edges * p.Zc;
        if (cbs[i].soft_slot > L.max_slot) L.max_slot = cbs[i].soft_slot;
        if (p.Zc > L.max_zc) L.max_zc = p.Zc;
        if (p.N + 2 * p.Zc > L.max_soft) L.max_soft = p.N + 2 * p.Zc;
        const uint32_t u = uniform_pair(p.bgn, p.Zc);
        L.uniform = (i == 0 || L.uniform == u) ? u : 0u;
        L.descs[i] = d;
    }
    return 0;
}

// Packed tiles for the int16 algos when asked (flags bit 1) or when some lifting size
// leaves room for two or more codeblocks per workgroup (bit 2 forbids it).
bool want_packing(int algo, uint32_t flags, const std::vector<CbDesc> &descs) {
    if (!i16_algo(algo) || (flags & 4u)) return false;
    if (flags & 2u) return true;
    for (const CbDesc &d : descs)
        if (tile_codeblocks(d.bgn, d.Zc) > 1) return true;
    return false;
}

}  // namespace

extern "C" int32_t ldpc_mi355_plan_create(ldpc_mi355_plan_t **plan, const ldpc_mi355_cb_t *cbs,
                                          int32_t n, const ldpc_decoder_cfg_t *dec,
                                          uint32_t flags, int32_t *bad_index) {
    *plan = nullptr;
    if (bad_index) *bad_index = -1;
    if (n <= 0) return -2;
    if (!algo_supported(dec->algo)) return -3;
    Runtime &r = rt();
    if (r.ensure() != 0) return -1;
    auto *pl = new ldpc_mi355_plan;
    pl->n = n;
    pl->algo = dec->algo;
    pl->dp = make_params(*dec);
    pl->want_lq = (flags & 1u) != 0;
    ScaleSet scale;
    if (table_algo(pl->algo)) make_scale_set(*dec, scale);
    BatchLayout L;
    if (int st = build_layout(cbs, n, pl->algo, pl->want_lq, r,
                              table_algo(pl->algo) ? &scale : nullptr, dec->max_iterations, L,
                              bad_index)) {
        delete pl;
        return st;
    }
    pl->hard_off = L.hard_off;
    pl->lq_off = L.lq_off;
    pl->hard_bytes = L.hard_bytes;
    pl->lq_elems = L.lq_elems;
    pl->max_soft = L.max_soft;
    pl->max_slot = L.max_slot;
    pl->threads = threads_for(L.max_zc);
    pl->uniform = (flags & 8u) ? 0u : L.uniform;
    if (hipMalloc(&pl->d_desc, sizeof(CbDesc) * (size_t)n) != hipSuccess ||
        hipMemcpy(pl->d_desc, L.descs.data(), sizeof(CbDesc) * (size_t)n,
                  hipMemcpyHostToDevice) != hipSuccess ||
        (L.work && hipMalloc(&pl->d_work, L.work * sizeof(float)) != hipSuccess)) {
        fprintf(stderr, "libnrldpc_mi355: plan allocation failed\n");
        ldpc_mi355_plan_destroy(pl);
        return -1;
    }
    if (table_algo(pl->algo)) {
        const size_t sb = scale.words.size() * sizeof(uint32_t);
        if (hipMalloc(&pl->d_scale, sb) != hipSuccess ||
            hipMemcpy(pl->d_scale, scale.words.data(), sb, hipMemcpyHostToDevice) !=
                hipSuccess) {
            fprintf(stderr, "libnrldpc_mi355: plan allocation failed\n");
            ldpc_mi355_plan_destroy(pl);
            return -1;
        }
        bind_scale(pl->dp, scale, pl->d_scale);
    }
    // lane-per-edge: algo 1 fast rows, every codeblock small enough for one
    // wave per layer
    bool edge = (flags & 32u) && pl->algo == 1 && !pl->dp.general;
    for (int i = 0; edge && i < n; ++i)
        edge = L.descs[i].Zc <= nrldpc_edge_lanes_max_zc(L.descs[i].bgn);
    pl->edge_lanes = edge;
    if (!edge && want_packing(pl->algo, flags, L.descs)) {
        std::vector<TileDesc> tiles;
        std::vector<uint32_t> order;
        build_tiles(L.descs, r.etab_off, tile_codeblocks, tiles, order, pl->tile_lds);
        pl->ntiles = (int)(tiles.size() / kTileWaves);
        if (hipMalloc(&pl->d_tiles, sizeof(TileDesc) * tiles.size()) != hipSuccess ||
            hipMemcpy(pl->d_tiles, tiles.data(), sizeof(TileDesc) * tiles.size(),
                      hipMemcpyHostToDevice) != hipSuccess ||
            hipMalloc(&pl->d_order, sizeof(uint32_t) * order.size()) != hipSuccess ||
            hipMemcpy(pl->d_order, order.data(), sizeof(uint32_t) * order.size(),
                      hipMemcpyHostToDevice) != hipSuccess ||
            ((flags & 16u) &&
             (hipMalloc(&pl->d_queue, 2 * sizeof(uint32_t) * ldpc_mi355_plan::kQueueRing) !=
                  hipSuccess ||
              hipMemset(pl->d_queue, 0, 2 * sizeof(uint32_t) * ldpc_mi355_plan::kQueueRing) !=
                  hipSuccess ||
              hipDeviceSynchronize() != hipSuccess))) {
            fprintf(stderr, "libnrldpc_mi355: plan allocation failed\n");
            ldpc_mi355_plan_destroy(pl);
            return -1;
        }
    }
    *plan = pl;
    return 0;
}

extern "C" int32_t ldpc_mi355_plan_layout(const ldpc_mi355_plan_t *plan, uint64_t *hard_bytes,
                                          uint64_t *lq_elems, int32_t *lq_elem_size,
                                          uint32_t *hard_offsets, uint32_t *lq_offsets) {
    if (!plan) return -1;
    if (hard_bytes) *hard_bytes = plan->hard_bytes;
    if (lq_elems) *lq_elems = plan->want_lq ? plan->lq_elems : 0;
    if (lq_elem_size) *lq_elem_size = i16_algo(plan->algo) ? 2 : 4;
    if (hard_offsets) memcpy(hard_offsets, plan->hard_off.data(), 4 * (size_t)plan->n);
    if (lq_offsets) memcpy(lq_offsets, plan->lq_off.data(), 4 * (size_t)plan->n);
    return 0;
}

extern "C" int32_t ldpc_mi355_plan_decode(const ldpc_mi355_plan_t *plan, const int16_t *llr,
                                          int16_t *soft_pool, uint8_t *hard, int32_t *status,
                                          int32_t *iterations, void *lq, void *stream) {
    if (!plan) return -1;
    if (plan->want_lq && !lq) return -1;
    if (plan->max_slot >= 0 && !soft_pool) {
        report_missing_soft_pool(plan->max_slot);
        return -4;
    }
    Runtime &r = rt();
    if (r.ensure() != 0) return -1;
    DecOut o;
    o.desc = plan->d_desc;
    o.llr = llr;
    o.soft = soft_pool;
    o.hard = hard;
    o.status = status;
    o.iters = iterations;
    o.lq = lq;
    o.work = plan->d_work;
    o.etab = r.d_etab;
    o.crc = r.d_etab + r.crc_off;
    // NULL is HIP's default (null) stream, as everywhere in HIP -- it is also
    // the handle PyTorch reports for its default stream.
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    uint32_t *queue = nullptr;
    if (plan->d_queue)
        queue = plan->d_queue +
                2 * (plan->launches.fetch_add(1u) % (uint32_t)ldpc_mi355_plan::kQueueRing);
    if (plan->edge_lanes)
        HIP_OK(nrldpc_launch_edge_lanes(&o, &plan->dp, plan->n, plan->max_soft, s));
    else if (plan->ntiles > 0)
        HIP_OK(nrldpc_launch_tiles(plan->algo, &o, plan->d_tiles, plan->d_order, &plan->dp,
                                   plan->ntiles, plan->tile_lds, queue, s));
    else
        HIP_OK(nrldpc_launch_decode(plan->algo, &o, &plan->dp, plan->n, plan->threads,
                                    plan->max_soft, plan->uniform, s));
    return 0;
}

extern "C" void ldpc_mi355_plan_destroy(ldpc_mi355_plan_t *plan) {
    if (!plan) return;
    if (plan->d_desc) (void)hipFree(plan->d_desc);
    if (plan->d_work) (void)hipFree(plan->d_work);
    if (plan->d_tiles) (void)hipFree(plan->d_tiles);
    if (plan->d_order) (void)hipFree(plan->d_order);
    if (plan->d_queue) (void)hipFree(plan->d_queue);
    if (plan->d_scale) (void)hipFree(plan->d_scale);
    delete plan;
}

// ---------------------------------------------------------------------------
// Reusable batch contexts (continuous batching): descriptors and tiles are
// built on the host into the context's pinned staging buffer, copied to its
// device buffer with one asynchronous copy on the caller's stream, and the
// decode is launched behind it -- no allocation and no synchronization per
// call.  An event recorded after the launch tells when the context (and its
// staging buffer) may be reused; a decode on a context whose previous decode
// is still running first waits for it.
// ---------------------------------------------------------------------------
struct ldpc_mi355_batch {
    int capacity = 0, algo = 1, max_it = 0;
    uint32_t flags = 0;
    bool want_lq = false;
    DecParams dp;
    ScaleSet scale;                 // algos 2/3: tables current at creation
    uint32_t *d_scale = nullptr;
    size_t stage_bytes = 0;
    unsigned char *h_stage = nullptr, *d_stage = nullptr;
    hipEvent_t done = nullptr;
    bool pending = false;
    BatchLayout layout;
    int n = 0, ntiles = 0;
};

namespace {
size_t batch_stage_bytes(int capacity) {
    const size_t d = ((sizeof(CbDesc) * (size_t)capacity) + 255) & ~(size_t)255;
    // at most one tile workgroup per codeblock
    const size_t t = ((sizeof(TileDesc) * kTileWaves * (size_t)capacity) + 255) & ~(size_t)255;
    return d + t + sizeof(uint32_t) * (size_t)capacity;
}
}  // namespace

extern "C" int32_t ldpc_mi355_batch_create(ldpc_mi355_batch_t **batch, int32_t capacity,
                                           const ldpc_decoder_cfg_t *dec, uint32_t flags) {
    if (!batch || !dec) return -4;
    *batch = nullptr;
    if (capacity <= 0) return -4;
    if (!algo_supported(dec->algo)) return -3;
    if (dec->algo == 0) return -3;   // BP keeps per-codeblock message buffers: plans only
    Runtime &r = rt();
    if (r.ensure() != 0) return -1;
    auto *b = new ldpc_mi355_batch;
    b->capacity = capacity;
    b->algo = dec->algo;
    b->flags = flags;
    b->want_lq = (flags & 1u) != 0;
    b->max_it = dec->max_iterations;
    b->dp = make_params(*dec);
    b->stage_bytes = batch_stage_bytes(capacity);
    if (hipHostMalloc(&b->h_stage, b->stage_bytes, hipHostMallocDefault) != hipSuccess ||
        hipMalloc(&b->d_stage, b->stage_bytes) != hipSuccess ||
        hipEventCreateWithFlags(&b->done, hipEventDisableTiming) != hipSuccess) {
        fprintf(stderr, "libnrldpc_mi355: batch context allocation failed\n");
        ldpc_mi355_batch_destroy(b);
        return -1;
    }
    if (table_algo(b->algo)) {
        make_scale_set(*dec, b->scale);
        const size_t sb = b->scale.words.size() * sizeof(uint32_t);
        if (hipMalloc(&b->d_scale, sb) != hipSuccess ||
            hipMemcpy(b->d_scale, b->scale.words.data(), sb, hipMemcpyHostToDevice) !=
                hipSuccess) {
            fprintf(stderr, "libnrldpc_mi355: batch context allocation failed\n");
            ldpc_mi355_batch_destroy(b);
            return -1;
        }
        bind_scale(b->dp, b->scale, b->d_scale);
    }
    *batch = b;
    return 0;
}

extern "C" int32_t ldpc_mi355_batch_decode(ldpc_mi355_batch_t *b, const ldpc_mi355_cb_t *cbs,
                                           int32_t n, const int16_t *llr, int16_t *soft_pool,
                                           uint8_t *hard, int32_t *status, int32_t *iterations,
                                           void *lq, void *stream, int32_t *bad_index) {
    if (bad_index) *bad_index = -1;
    if (!b || !cbs || n <= 0) return -4;
    if (n > b->capacity) return -4;
    if (b->want_lq && !lq) return -4;
    Runtime &r = rt();
    if (r.ensure() != 0) return -1;
    if (b->pending) {
        HIP_OK(hipEventSynchronize(b->done));
        b->pending = false;
    }
    BatchLayout &L = b->layout;
    b->n = 0;
    b->ntiles = 0;
    if (int st = build_layout(cbs, n, b->algo, b->want_lq, r,
                              table_algo(b->algo) ? &b->scale : nullptr, b->max_it, L, bad_index))
        return st;
    if (L.max_slot >= 0 && !soft_pool) {
        report_missing_soft_pool(L.max_slot);
        return -4;
    }
    // Staging packed to this batch's size (descriptors, tiles, order back to
    // back): one small copy that the runtime runs as a copy kernel on the
    // stream.  Copies sized for the whole capacity go to the DMA engine,
    // which the contexts' streams share and which orders them one after
    // another behind each other's pending device-to-host copies.
    const size_t doff = 0;
    size_t toff = 0, ooff = 0;
    memcpy(b->h_stage + doff, L.descs.data(), sizeof(CbDesc) * (size_t)n);
    size_t bytes = sizeof(CbDesc) * (size_t)n;
    int tile_lds = 0;
    if (want_packing(b->algo, b->flags, L.descs)) {
        std::vector<TileDesc> tiles;
        std::vector<uint32_t> order;
        build_tiles(L.descs, r.etab_off, tile_codeblocks, tiles, order, tile_lds);
        b->ntiles = (int)(tiles.size() / kTileWaves);
        toff = (bytes + 255) & ~(size_t)255;
        ooff = toff + ((sizeof(TileDesc) * tiles.size() + 255) & ~(size_t)255);
        bytes = ooff + sizeof(uint32_t) * order.size();
        if (bytes > b->stage_bytes) {   // at most one workgroup per codeblock
            fprintf(stderr, "libnrldpc_mi355: batch staging overflow (%zu > %zu bytes)\n", bytes,
                    b->stage_bytes);
            return -1;
        }
        memcpy(b->h_stage + toff, tiles.data(), sizeof(TileDesc) * tiles.size());
        memcpy(b->h_stage + ooff, order.data(), sizeof(uint32_t) * order.size());
    }
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    HIP_OK(hipMemcpyAsync(b->d_stage, b->h_stage, bytes, hipMemcpyHostToDevice, s));
    DecOut o;
    o.desc = reinterpret_cast<const CbDesc *>(b->d_stage + doff);
    o.llr = llr;
    o.soft = soft_pool;
    o.hard = hard;
    o.status = status;
    o.iters = iterations;
    o.lq = lq;
    o.work = nullptr;
    o.etab = r.d_etab;
    o.crc = r.d_etab + r.crc_off;
    if (b->ntiles > 0)
        HIP_OK(nrldpc_launch_tiles(b->algo, &o,
                                   reinterpret_cast<const TileDesc *>(b->d_stage + toff),
                                   reinterpret_cast<const uint32_t *>(b->d_stage + ooff),
                                   &b->dp, b->ntiles, tile_lds, nullptr, s));
    else
        HIP_OK(nrldpc_launch_decode(b->algo, &o, &b->dp, n, threads_for(L.max_zc), L.max_soft,
                                    (b->flags & 8u) ? 0u : L.uniform, s));
    HIP_OK(hipEventRecord(b->done, s));
    b->pending = true;
    b->n = n;
    return 0;
}

extern "C" int32_t ldpc_mi355_batch_layout(const ldpc_mi355_batch_t *b, uint64_t *hard_bytes,
                                           uint64_t *lq_elems, uint32_t *hard_offsets,
                                           uint32_t *lq_offsets) {
    if (!b) return -4;
    if (hard_bytes) *hard_bytes = b->layout.hard_bytes;
    if (lq_elems) *lq_elems = b->want_lq ? b->layout.lq_elems : 0;
    if (hard_offsets) memcpy(hard_offsets, b->layout.hard_off.data(), 4 * (size_t)b->n);
    if (lq_offsets) memcpy(lq_offsets, b->layout.lq_off.data(), 4 * (size_t)b->n);
    return b->n;
}

extern "C" int32_t ldpc_mi355_batch_query(ldpc_mi355_batch_t *b) {
    if (!b) return -4;
    if (!b->pending) return 1;
    const hipError_t e = hipEventQuery(b->done);
    if (e == hipSuccess) {
        b->pending = false;
        return 1;
    }
    if (e == hipErrorNotReady) return 0;
    fprintf(stderr, "libnrldpc_mi355: batch query failed: %s\n", hipGetErrorString(e));
    return -1;
}

extern "C" int32_t ldpc_mi355_batch_wait(ldpc_mi355_batch_t *b) {
    if (!b) return -4;
    if (b->pending) {
        HIP_OK(hipEventSynchronize(b->done));
        b->pending = false;
    }
    return 0;
}

extern "C" void ldpc_mi355_batch_destroy(ldpc_mi355_batch_t *b) {
    if (!b) return;
    if (b->pending) (void)hipEventSynchronize(b->done);
    if (b->done) (void)hipEventDestroy(b->done);
    if (b->h_stage) (void)hipHostFree(b->h_stage);
    if (b->d_stage) (void)hipFree(b->d_stage);
    if (b->d_scale) (void)hipFree(b->d_scale);
    delete b;
}

// Many small asynchronous copies in one call (the served loop moves one
// pinned span per request in and one result span out per batch).
extern "C" int32_t ldpc_mi355_copy_spans(void *const *dst, const void *const *src,
                                         const uint64_t *bytes, int32_t n, int32_t kind,
                                         void *stream) {
    if (n < 0 || (n > 0 && (!dst || !src || !bytes)) || (kind != 1 && kind != 2)) return -1;
    const hipMemcpyKind k = kind == 1 ? hipMemcpyHostToDevice : hipMemcpyDeviceToHost;
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    for (int32_t i = 0; i < n; ++i)
        if (bytes[i] && hipMemcpyAsync(dst[i], src[i], bytes[i], k, s) != hipSuccess) return -1;
    return 0;
}

// ---------------------------------------------------------------------------
// Device input synthesis (one config, n codeblocks, one launch)
// ---------------------------------------------------------------------------
namespace {
int32_t synth_device(const ldpc_in_config_t *cfg, int32_t n, const float *noise_power_db,
                     const float *noise_var, const float *noise, uint64_t seed,
                     const uint8_t *payload_bits, uint8_t *bits_out, uint8_t *codeword_out,
                     uint8_t *rm_bits_out, int16_t *llr_out, float *llr_f32_out,
                     float *beliefs_out, void *stream) {
    if (!cfg || n < 0) return -4;
    Params p;
    if (const char *why = derive(*cfg, p)) {
        print_invalid(*cfg, why);
        return -2;
    }
    if (n == 0) return 0;
    if ((!noise_power_db && !noise_var) || (beliefs_out && !llr_f32_out)) return -4;
    Runtime &r = rt();
    if (r.ensure() != 0) return -1;
    SynthArgs a{};
    a.noise_db = noise_power_db;
    a.noise_var = noise_var;
    a.noise = noise;
    a.payload = payload_bits;
    a.bits_out = bits_out;
    a.cw_out = codeword_out;
    a.rm_out = rm_bits_out;
    a.llr_out = llr_out;
    a.llr_f32 = llr_f32_out;
    a.beliefs = beliefs_out;
    a.etab = r.d_etab + r.etab_off[p.bgn - 1][p.Zc];
    a.seed = seed;
    a.n = n;
    a.A = p.A;
    a.B = p.B;
    a.K = p.K;
    a.N = p.N;
    a.E = p.E;
    a.Zc = p.Zc;
    a.Qm = p.Qm;
    a.k0 = p.k0;
    a.crctype = p.crctype;
    HIP_OK(nrldpc_launch_synth(p.bgn, &a, reinterpret_cast<hipStream_t>(stream)));
    return 0;
}
}  // namespace

extern "C" int32_t ldpc_mi355_synth_device(const ldpc_in_config_t *cfg, int32_t n,
                                           const float *noise_power_db, uint64_t seed,
                                           const uint8_t *payload_bits, uint8_t *bits_out,
                                           uint8_t *codeword_out, uint8_t *rm_bits_out,
                                           int16_t *llr_out, float *llr_f32_out,
                                           float *beliefs_out, void *stream) {
    if (!noise_power_db) return -4;
    return synth_device(cfg, n, noise_power_db, nullptr, nullptr, seed, payload_bits, bits_out,
                        codeword_out, rm_bits_out, llr_out, llr_f32_out, beliefs_out, stream);
}

extern "C" int32_t ldpc_mi355_synth_device_noise(const ldpc_in_config_t *cfg, int32_t n,
                                                 const float *noise_var, const float *noise,
                                                 uint64_t seed, const uint8_t *payload_bits,
                                                 uint8_t *bits_out, uint8_t *codeword_out,
                                                 uint8_t *rm_bits_out, int16_t *llr_out,
                                                 float *llr_f32_out, float *beliefs_out,
                                                 void *stream) {
    if (!noise_var || !noise) return -4;
    return synth_device(cfg, n, nullptr, noise_var, noise, seed, payload_bits, bits_out,
                        codeword_out, rm_bits_out, llr_out, llr_f32_out, beliefs_out, stream);
}

// ---------------------------------------------------------------------------
// Algo-2 trained tables and algo-3 degree alpha (process-wide)
// ---------------------------------------------------------------------------
extern "C" int32_t ldpc_mi355_set_trained_tables(int32_t bgn, const int16_t *alpha,
                                                 const int16_t *beta, int32_t iterations) {
    return set_trained_tables(bgn, alpha, beta, iterations);
}

extern "C" int32_t ldpc_mi355_get_trained_tables(int32_t bgn, int16_t *alpha, int16_t *beta,
                                                 int32_t capacity_iterations) {
    return get_trained_tables(bgn, alpha, beta, capacity_iterations);
}

extern "C" int32_t ldpc_mi355_set_degree_alpha(const int16_t *degree_alpha_values) {
    set_degree_alpha(degree_alpha_values);
    return 0;
}

extern "C" int32_t ldpc_mi355_get_degree_alpha(int16_t *degree_alpha_values) {
    if (!degree_alpha_values) return -4;
    get_degree_alpha(degree_alpha_values);
    return 0;
}

extern "C" int32_t ldpc_mi355_sync(void) {
    Runtime &r = rt();
    if (r.ensure() != 0) return -1;
    HIP_OK(hipStreamSynchronize(r.stream));
    return 0;
}

extern "C" int32_t ldpc_mi355_soft_slot_elems(void) { return kMaxSoft; }

extern "C" int32_t ldpc_mi355_algo_available(int32_t algo) {
    return algo_supported(algo) ? 1 : 0;
}

extern "C" const char *ldpc_mi355_version(void) { return "nrldpc-mi355 0.1 (gfx950)"; }
