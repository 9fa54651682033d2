// Host build of the HIP decode kernels (ldpc_kernels.hip compiled with g++
// under hip_cpu_emu.h), driven like ldpc_mi355_plan_decode.  Test
// infrastructure: lets CPU tests check the kernel source logic against the
// oracle, and shows whether a kernel's control flow terminates.
#include "ldpc_kernels.hip"

#include "ldpc_host.h"

using namespace nrldpc;

namespace {
struct EdgeTables {
    std::vector<uint32_t> tab;
    uint32_t off[2][kMaxZc + 1] = {};
    EdgeTables() {
        int nz = 0;
        const uint16_t *zs = all_lifting_sizes(&nz);
        for (int bg = 1; bg <= 2; ++bg) {
            const int ne = bg == 1 ? NR_BG1_EDGES : NR_BG2_EDGES;
            for (int i = 0; i < nz; ++i) {
                off[bg - 1][zs[i]] = (uint32_t)tab.size();
                tab.resize(tab.size() + ne);
                edge_table(bg, zs[i], lifting_set(zs[i]), tab.data() + off[bg - 1][zs[i]]);
            }
        }
    }
};
}  // namespace

// n codeblocks, cfgs[7*i..] = ldpc_in_config_t fields, one decoder row.
// Offsets of each codeblock's hard bytes / posteriors are returned.
// Returns 0, -2 for an invalid config, -3 for an unknown algo.
extern "C" int emu_decode_batch(int n, const int32_t *cfgs, const uint32_t *llr_off,
                                const int32_t *row, const int16_t *llr, int16_t *soft_pool,
                                const int32_t *soft_slot, uint8_t *hard, int32_t *status,
                                int32_t *iters, void *lq, uint32_t *hard_off,
                                uint32_t *lq_off, float *work, uint64_t work_elems) {
    static EdgeTables et;
    ldpc_decoder_cfg_t dc{(uint16_t)row[0], (uint16_t)row[1], (uint16_t)row[2], (uint16_t)row[3],
                          (int16_t)row[4], (int16_t)row[5], (int16_t)row[6]};
    const DecParams p = decoder_params(dc);
    std::vector<CbDesc> descs(n);
    uint32_t h = 0, q = 0;
    uint64_t w = 0;
    int max_zc = 2, max_soft = 0;
    for (int i = 0; i < n; ++i) {
        const int32_t *c = cfgs + 7 * i;
        ldpc_in_config_t in{(uint16_t)c[0], (uint16_t)c[1], (uint16_t)c[2], (uint16_t)c[3],
                            (uint16_t)c[4], (uint16_t)c[5], (uint16_t)c[6]};
        Params pr;
        if (derive(in, pr)) return -2;
        CbDesc d = codeblock_desc(pr, et.off[pr.bgn - 1][pr.Zc]);
        d.llr_off = llr_off[i];
        d.hard_off = h;
        d.soft_off = soft_slot && soft_slot[i] >= 0 ? (uint32_t)soft_slot[i] * kMaxSoft : kNone;
        d.lq_off = lq ? q : kNone;
        d.work_off = (uint32_t)w;
        hard_off[i] = h;
        lq_off[i] = q;
        h += pr.B / 8;
        q += pr.N + 2 * pr.Zc;
        w += (uint64_t)pr.edges * pr.Zc;
        max_zc = pr.Zc > max_zc ? pr.Zc : max_zc;
        max_soft = pr.N + 2 * pr.Zc > max_soft ? pr.N + 2 * pr.Zc : max_soft;
        descs[i] = d;
    }
    if (row[0] == 0 && w > work_elems) return -1;
    DecOut o;
    o.desc = descs.data();
    o.llr = llr;
    o.soft = soft_pool;
    o.hard = hard;
    o.status = status;
    o.iters = iters;
    o.lq = lq;
    o.work = work;
    o.etab = et.tab.data();
    const unsigned threads = (unsigned)workgroup_threads(max_zc);
    switch (row[0]) {
    case 1:
        if (p.alpha >= 0 && p.min_llr <= p.max_llr)
            emu_launch(n, threads, (size_t)max_soft * 2,
                       [&] { ldpc_decode_kernel<Algo::kMinSumI16, int16_t, false>(o, p); });
        else
            emu_launch(n, threads, (size_t)max_soft * 2,
                       [&] { ldpc_decode_kernel<Algo::kMinSumI16, int16_t, true>(o, p); });
        return 0;
    case 4:
        emu_launch(n, threads, (size_t)max_soft * 4,
                   [&] { ldpc_decode_kernel<Algo::kMinSumF32, float>(o, p); });
        return 0;
    case 0:
        emu_launch(n, threads, (size_t)max_soft * 4,
                   [&] { ldpc_decode_kernel<Algo::kBpF32, float>(o, p); });
        return 0;
    default:
        return -3;
    }
}
