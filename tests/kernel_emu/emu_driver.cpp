// Host build of the HIP decode kernels (ldpc_kernels.hip compiled with g++
// under hip_cpu_emu.h), driven like ldpc_mi355_plan_decode.  Test
// infrastructure: lets CPU tests check the kernel source logic against the
// oracle, and shows whether a kernel's control flow terminates.
#include "ldpc_kernels.hip"
#include "ldpc_bp.hip"

#include "ldpc_host.h"

using namespace nrldpc;

namespace {
struct EdgeTables {
    std::vector<uint32_t> tab;
    std::vector<uint32_t> crc = std::vector<uint32_t>(4 * (size_t)kCrcTabWords);
    uint32_t off[2][kMaxZc + 1] = {};
    EdgeTables() {
        crc_device_tables(crc.data());
        int nz = 0;
        const uint16_t *zs = all_lifting_sizes(&nz);
        for (int bg = 1; bg <= 2; ++bg) {
            const int ne = edge_block_words(bg);
            for (int i = 0; i < nz; ++i) {
                off[bg - 1][zs[i]] = (uint32_t)tab.size();
                tab.resize(tab.size() + ne);
                edge_block(bg, zs[i], lifting_set(zs[i]), tab.data() + off[bg - 1][zs[i]]);
            }
        }
    }
};

EdgeTables &tables() {
    static EdgeTables et;
    return et;
}

struct Batch {
    std::vector<CbDesc> descs;
    DecParams p;
    ScaleSet scale;   // algos 2/3
    int max_zc = 2, max_soft = 0;
    uint64_t work = 0;
    uint32_t uniform = 0;   // (bgn << 16) | Zc when all codeblocks share them
};

// Descriptors exactly as ldpc_mi355_plan_create lays them out.
int make_batch(int n, const int32_t *cfgs, const uint32_t *llr_off, const int32_t *row,
               const int32_t *soft_slot, bool want_lq, uint32_t *hard_off, uint32_t *lq_off,
               Batch &b) {
    ldpc_decoder_cfg_t dc{(uint16_t)row[0], (uint16_t)row[1], (uint16_t)row[2], (uint16_t)row[3],
                          (int16_t)row[4], (int16_t)row[5], (int16_t)row[6]};
    b.p = decoder_params(dc);
    if (dc.algo == 2 || dc.algo == 3) {
        make_scale_set(dc, b.scale);
        b.p.scale[0] = b.scale.words.data() + b.scale.off[0];
        b.p.scale[1] = b.scale.words.data() + b.scale.off[1];
        b.p.scale_iters = b.scale.iters;
        b.p.general = b.scale.fast ? 0 : 1;
        b.p.simple = (b.scale.fast && b.scale.simple) ? 1 : 0;
    }
    b.descs.resize(n);
    uint32_t h = 0, q = 0;
    for (int i = 0; i < n; ++i) {
        const int32_t *c = cfgs + 7 * i;
        ldpc_in_config_t in{(uint16_t)c[0], (uint16_t)c[1], (uint16_t)c[2], (uint16_t)c[3],
                            (uint16_t)c[4], (uint16_t)c[5], (uint16_t)c[6]};
        Params pr;
        if (derive(in, pr)) return -2;
        if ((dc.algo == 2 || dc.algo == 3) && !scale_iterations_ok(b.scale, pr.bgn, dc.max_iterations))
            return -2;
        CbDesc d = codeblock_desc(pr, tables().off[pr.bgn - 1][pr.Zc]);
        d.llr_off = llr_off[i];
        d.hard_off = h;
        d.soft_off = soft_slot && soft_slot[i] >= 0 ? (uint32_t)soft_slot[i] * kMaxSoft : kNone;
        d.lq_off = want_lq ? q : kNone;
        d.work_off = (uint32_t)b.work;
        hard_off[i] = h;
        lq_off[i] = q;
        h += pr.B / 8;
        q += pr.N + 2 * pr.Zc;
        b.work += (uint64_t)pr.edges * pr.Zc;
        b.max_zc = std::max(b.max_zc, pr.Zc);
        b.max_soft = std::max(b.max_soft, pr.N + 2 * pr.Zc);
        const uint32_t u = ((uint32_t)pr.bgn << 16) | (uint32_t)pr.Zc;
        b.uniform = (i == 0 || b.uniform == u) ? u : 0u;
        b.descs[i] = d;
    }
    return 0;
}

DecOut outputs(const Batch &b, const int16_t *llr, int16_t *soft_pool, uint8_t *hard,
               int32_t *status, int32_t *iters, void *lq, float *work) {
    DecOut o;
    o.desc = b.descs.data();
    o.llr = llr;
    o.soft = soft_pool;
    o.hard = hard;
    o.status = status;
    o.iters = iters;
    o.lq = lq;
    o.work = work;
    o.etab = tables().tab.data();
    o.crc = tables().crc.data();
    return o;
}

int g_force_general = -1;   // -1: the library's choice; 0 / 1 forces a path
int g_edge_lanes = 0;       // 1: algo-1 fast rows in the lane-per-edge schedule
int g_use_fixed = 1;        // run the lifting-size specialized kernel when one exists

// The specialized kernel of (BG, ZC) for a uniform batch.  The library
// builds BG1 Zc=384; the small pairs exist here only, to test the template
// cheaply on the CPU.
template <int BG, int ZC>
void launch_fixed(unsigned n, unsigned threads, const DecOut &o, const DecParams &p, bool g,
                  bool t) {
    const size_t lds = (size_t)FixedLds<BG, ZC>::kBytes;
    if (!g && p.simple && !t)
        emu_launch(n, threads, lds,
                   [&] { ldpc_decode_fixed_kernel<BG, ZC, false, false, true>(o, p); });
    else if (!g && p.simple && t)
        emu_launch(n, threads, lds,
                   [&] { ldpc_decode_fixed_kernel<BG, ZC, false, true, true>(o, p); });
    else if (!g && !t)
        emu_launch(n, threads, lds, [&] { ldpc_decode_fixed_kernel<BG, ZC, false, false>(o, p); });
    else if (g && !t)
        emu_launch(n, threads, lds, [&] { ldpc_decode_fixed_kernel<BG, ZC, true, false>(o, p); });
    else if (!g && t)
        emu_launch(n, threads, lds, [&] { ldpc_decode_fixed_kernel<BG, ZC, false, true>(o, p); });
    else
        emu_launch(n, threads, lds, [&] { ldpc_decode_fixed_kernel<BG, ZC, true, true>(o, p); });
}

bool try_fixed(uint32_t uniform, unsigned n, unsigned threads, const DecOut &o,
               const DecParams &p, bool g, bool t) {
    if (!g_use_fixed) return false;
    switch (uniform) {
    case (1u << 16) | 384u: launch_fixed<1, 384>(n, threads, o, p, g, t); return true;
    case (2u << 16) | 26u: launch_fixed<2, 26>(n, threads, o, p, g, t); return true;
    case (1u << 16) | 15u: launch_fixed<1, 15>(n, threads, o, p, g, t); return true;
    case (1u << 16) | 16u: launch_fixed<1, 16>(n, threads, o, p, g, t); return true;
    default: return false;
    }
}
int g_fixed_k = 0;
int fixed_k(int bgn, int zc) {
    (void)bgn;
    const int cap = std::min(kMaxThreads / zc, kMaxTileCodeblocks);
    return g_fixed_k > 0 ? std::min(g_fixed_k, cap) : cap;
}
}  // namespace

// n codeblocks, cfgs[7*i..] = ldpc_in_config_t fields, one decoder row, one
// codeblock per workgroup.  Returns 0, -2 for an invalid config, -3 for an
// unknown algo, -1 when the BP message store is too small.
extern "C" int emu_decode_batch(int n, const int32_t *cfgs, const uint32_t *llr_off,
                                const int32_t *row, const int16_t *llr, int16_t *soft_pool,
                                const int32_t *soft_slot, uint8_t *hard, int32_t *status,
                                int32_t *iters, void *lq, uint32_t *hard_off,
                                uint32_t *lq_off, float *work, uint64_t work_elems) {
    Batch b;
    if (int rc = make_batch(n, cfgs, llr_off, row, soft_slot, lq != nullptr, hard_off, lq_off, b))
        return rc;
    if (row[0] == 0 && b.work > work_elems) return -1;
    const DecOut o = outputs(b, llr, soft_pool, hard, status, iters, lq, work);
    const DecParams p = b.p;
    const unsigned threads = (unsigned)workgroup_threads(b.max_zc);
    // beliefs, then the min-sum kernels' CRC bit weights (nrldpc_launch_decode)
    const size_t wts_bytes = (size_t)kGenericWtsCols * threads * 4;
    const size_t lds16 = (((size_t)b.max_soft * 2 + 15) & ~(size_t)15) + wts_bytes;
    const size_t lds32 = (((size_t)b.max_soft * 4 + 15) & ~(size_t)15) + wts_bytes;
    switch (row[0]) {
    case 1:
    case 2:
    case 3: {
        // the launcher's choice (nrldpc_launch_decode): DecParams::general
        // and the table flag of algos 2/3
        const bool g = g_force_general >= 0 ? g_force_general != 0 : p.general != 0;
        const bool t = row[0] != 1;
        if (g_edge_lanes && !g && !t) {
            // lane-per-edge schedule: one 64-lane workgroup per codeblock
            for (int i = 0; i < n; ++i)
                if (b.descs[i].Zc > (b.descs[i].bgn == 1 ? kEdgeLaneMaxZc<1> : kEdgeLaneMaxZc<2>))
                    return -3;
            emu_launch(n, 64, lds16,
                       [&] { ldpc_decode_kernel<Algo::kMinSumI16Edge, int16_t>(o, p); });
            return 0;
        }
        if (try_fixed(b.uniform, n, threads, o, p, g, t)) return 0;
        if (!g && !t)
            emu_launch(n, threads, lds16,
                       [&] { ldpc_decode_kernel<Algo::kMinSumI16, int16_t, false, false>(o, p); });
        else if (g && !t)
            emu_launch(n, threads, lds16,
                       [&] { ldpc_decode_kernel<Algo::kMinSumI16, int16_t, true, false>(o, p); });
        else if (!g && t)
            emu_launch(n, threads, lds16,
                       [&] { ldpc_decode_kernel<Algo::kMinSumI16, int16_t, false, true>(o, p); });
        else
            emu_launch(n, threads, lds16,
                       [&] { ldpc_decode_kernel<Algo::kMinSumI16, int16_t, true, true>(o, p); });
        return 0;
    }
    case 4:
        if (g_use_fixed && b.uniform == ((2u << 16) | 26u)) {
            emu_launch(n, threads, (size_t)FixedLdsF32<2, 26>::kBytes,
                       [&] { ldpc_decode_fixed_f32_kernel<2, 26>(o, p); });
            return 0;
        }
        if (g_use_fixed && b.uniform == ((1u << 16) | 15u)) {
            emu_launch(n, threads, (size_t)FixedLdsF32<1, 15>::kBytes,
                       [&] { ldpc_decode_fixed_f32_kernel<1, 15>(o, p); });
            return 0;
        }
        emu_launch(n, threads, lds32,
                   [&] { ldpc_decode_kernel<Algo::kMinSumF32, float>(o, p); });
        return 0;
    case 0:
        emu_launch(n, threads, (size_t)b.max_soft * 4,
                   [&] { ldpc_bp_kernel(o, p); });
        return 0;
    default:
        return -3;
    }
}

// The tile plan of a batch (build_tiles, as plans and batch contexts build
// it) with k codeblocks per group at most (0 = as many as fit): up to
// max_entries TileDesc (kTileWaves per workgroup) into tiles, n indices into
// order, the largest LQ footprint into *lds_entries.  Returns the number of
// workgroups, -1 when max_entries is too small, or another negative code.
extern "C" int emu_plan_tiles(int n, const int32_t *cfgs, int k, TileDesc *tiles_out,
                              int max_entries, uint32_t *order_out, int *lds_entries) {
    std::vector<uint32_t> llr_off((size_t)n, 0u), hard_off((size_t)n), lq_off((size_t)n);
    const int32_t row[7] = {1, 8, 22936, 0, 1280, 32767, -32767};
    Batch b;
    if (int rc = make_batch(n, cfgs, llr_off.data(), row, nullptr, false, hard_off.data(),
                            lq_off.data(), b))
        return rc < 0 ? rc : -rc;
    std::vector<TileDesc> tiles;
    std::vector<uint32_t> order;
    g_fixed_k = k;
    build_tiles(b.descs, tables().off, fixed_k, tiles, order, *lds_entries);
    if ((int)tiles.size() > max_entries) return -1;
    memcpy(tiles_out, tiles.data(), sizeof(TileDesc) * tiles.size());
    memcpy(order_out, order.data(), sizeof(uint32_t) * order.size());
    return (int)(tiles.size() / kTileWaves);
}

// > 0: tile launches are persistent with this many workgroups at most.
int g_persistent_grid = 0;
extern "C" void emu_set_persistent_tiles(int grid) { g_persistent_grid = grid; }

// Same batch through the packed-tile kernel (algos 1-3) with k codeblocks
// per group at most (0 = as many as fit).  Returns the number of tile
// workgroups, or < 0.
extern "C" int emu_decode_tiles(int n, const int32_t *cfgs, const uint32_t *llr_off,
                                const int32_t *row, const int16_t *llr, int16_t *soft_pool,
                                const int32_t *soft_slot, uint8_t *hard, int32_t *status,
                                int32_t *iters, void *lq, uint32_t *hard_off, uint32_t *lq_off,
                                int k) {
    if (row[0] < 1 || row[0] > 3) return -3;
    Batch b;
    if (int rc = make_batch(n, cfgs, llr_off, row, soft_slot, lq != nullptr, hard_off, lq_off, b))
        return rc;
    std::vector<TileDesc> tiles;
    std::vector<uint32_t> order;
    int lds_entries = 0;
    g_fixed_k = k;
    build_tiles(b.descs, tables().off, fixed_k, tiles, order, lds_entries);
    const DecOut o = outputs(b, llr, soft_pool, hard, status, iters, lq, nullptr);
    const DecParams p = b.p;
    const TileDesc *td = tiles.data();
    const uint32_t *ord = order.data();
    const bool g = g_force_general >= 0 ? g_force_general != 0 : p.general != 0;
    const bool t = row[0] != 1;
    const unsigned nt = (unsigned)(tiles.size() / kTileWaves);
    const size_t beliefs = ((size_t)lds_entries * 2 + 15) & ~(size_t)15;
    const size_t lds = beliefs + kTileWtsBytes;
    const bool simple = p.simple != 0;
    // persistent launch: g_persistent_grid workgroups pull the tiles from a queue
    uint32_t ctr[2] = {0u, 0u};
    TileQueue q;
    q.ctr = g_persistent_grid > 0 ? ctr : nullptr;
    q.ntiles = (int)nt;
    q.wts_lds = (uint32_t)beliefs;
    const unsigned grid = g_persistent_grid > 0 ? std::min(nt, (unsigned)g_persistent_grid) : nt;
    if (!g && !t && simple)
        emu_launch(grid, kTileThreads, lds, [&] { ldpc_tile_kernel<false, false, true>(o, td, ord, p, q); });
    else if (!g && t && simple)
        emu_launch(grid, kTileThreads, lds, [&] { ldpc_tile_kernel<false, true, true>(o, td, ord, p, q); });
    else if (!g && !t)
        emu_launch(grid, kTileThreads, lds, [&] { ldpc_tile_kernel<false, false>(o, td, ord, p, q); });
    else if (g && !t)
        emu_launch(grid, kTileThreads, lds, [&] { ldpc_tile_kernel<true, false>(o, td, ord, p, q); });
    else if (!g && t)
        emu_launch(grid, kTileThreads, lds, [&] { ldpc_tile_kernel<false, true>(o, td, ord, p, q); });
    else
        emu_launch(grid, kTileThreads, lds, [&] { ldpc_tile_kernel<true, true>(o, td, ord, p, q); });
    // the last workgroup out rearms the queue
    if (q.ctr && (ctr[0] != 0u || ctr[1] != 0u)) return -5;
    return (int)nt;
}

// The library's process-wide algo-2 / algo-3 tables (this build's copy).
extern "C" int emu_set_trained_tables(int bgn, const int16_t *alpha, const int16_t *beta,
                                      int iterations) {
    return set_trained_tables(bgn, alpha, beta, iterations);
}

extern "C" void emu_set_degree_alpha(const int16_t *values) { set_degree_alpha(values); }

// 0: never run a lifting-size specialized kernel (generic kernels only).
extern "C" void emu_use_fixed(int on) { g_use_fixed = on; }

// The device CRC tables as the library builds them (crc_device_tables):
// 4 * kCrcTabWords words; returns the number copied, or -1 if n is short.
extern "C" int emu_crc_tables(uint32_t *out, int n) {
    const std::vector<uint32_t> &t = tables().crc;
    if (n < (int)t.size()) return -1;
    std::copy(t.begin(), t.end(), out);
    return (int)t.size();
}
extern "C" int emu_crc_tab_words() { return kCrcTabWords; }
extern "C" int emu_crc_bit_off() { return kCrcBitOff; }

// Forces the int16 row form (0 fast, 1 exact) for negative controls; -1
// restores the library's choice.
extern "C" void emu_force_general(int g) { g_force_general = g; }

// DecParams::general of a decoder row (algo 1) as the library computes it.
extern "C" int emu_row_general(const int32_t *row) {
    ldpc_decoder_cfg_t dc{(uint16_t)row[0], (uint16_t)row[1], (uint16_t)row[2], (uint16_t)row[3],
                          (int16_t)row[4], (int16_t)row[5], (int16_t)row[6]};
    return decoder_params(dc).general;
}

extern "C" void emu_set_edge_lanes(int on) { g_edge_lanes = on; }
