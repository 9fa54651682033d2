// Sanitizer run of the CPU build of the decode kernels and of the oracle
// (test infrastructure): built with -fsanitize=address,undefined and
// -fno-sanitize-recover, so any out-of-bounds access, use after free or
// undefined behaviour in the kernel source, the host helpers or the oracle
// ends the process with a report.  Decodes synthesized codeblocks of several
// codes with algos 0, 1 and 4 (one codeblock per workgroup) and algo 1 in
// packed tiles, and checks algo 1 against the oracle's whole decode path.
#include <stdio.h>
#include <stdint.h>
#include <string.h>

#include <vector>

#include "../../include/nr_ldpc_mi355.h"
extern "C" {
#include "../../oracle/ldpc_oracle.h"
}

extern "C" int emu_decode_batch(int n, const int32_t *cfgs, const uint32_t *llr_off,
                                const int32_t *row, const int16_t *llr, int16_t *soft_pool,
                                const int32_t *soft_slot, uint8_t *hard, int32_t *status,
                                int32_t *iters, void *lq, uint32_t *hard_off, uint32_t *lq_off,
                                float *work, uint64_t work_elems);
extern "C" int emu_decode_tiles(int n, const int32_t *cfgs, const uint32_t *llr_off,
                                const int32_t *row, const int16_t *llr, int16_t *soft_pool,
                                const int32_t *soft_slot, uint8_t *hard, int32_t *status,
                                int32_t *iters, void *lq, uint32_t *hard_off, uint32_t *lq_off,
                                int k);

namespace {
struct Case {
    int cfg[7];
    float noise_db;
};

int run_case(const Case &c, int n) {
    ldpc_in_config_t ic{(uint16_t)c.cfg[0], (uint16_t)c.cfg[1], (uint16_t)c.cfg[2],
                        (uint16_t)c.cfg[3], (uint16_t)c.cfg[4], (uint16_t)c.cfg[5],
                        (uint16_t)c.cfg[6]};
    const int E = c.cfg[3], B = c.cfg[1] + (c.cfg[2] == 0 ? 16 : 24);
    std::vector<int16_t> llr((size_t)n * E);
    if (ldpc_mi355_synth_batch(&ic, n, c.noise_db, 7, 2, llr.data(), nullptr) != 0) return 1;
    std::vector<int32_t> cfgs((size_t)7 * n);
    std::vector<uint32_t> off(n), hoff(n), loff(n);
    for (int i = 0; i < n; ++i) {
        memcpy(&cfgs[(size_t)7 * i], c.cfg, sizeof c.cfg);
        off[i] = (uint32_t)(i * E);
    }
    std::vector<int16_t> pool((size_t)n * 68 * 384, 5);
    std::vector<int32_t> slots(n);
    for (int i = 0; i < n; ++i) slots[i] = i;
    std::vector<uint8_t> hard((size_t)n * 1056);
    std::vector<int32_t> st(n), it(n);
    std::vector<float> lq((size_t)n * 68 * 384);
    std::vector<float> work((size_t)n * 316 * 384);
    const int32_t rows[3][7] = {{1, 8, 22936, 0, 1280, 32767, -32767},
                                {4, 12, 22936, 0, 1280, 32767, -32767},
                                {0, 8, 1, 0, 1280, 20, -32768}};
    for (const auto &row : rows) {
        if (emu_decode_batch(n, cfgs.data(), off.data(), row, llr.data(), pool.data(),
                             slots.data(), hard.data(), st.data(), it.data(), lq.data(),
                             hoff.data(), loff.data(), work.data(), work.size()) != 0)
            return 2;
        if (row[0] == 1) {
            orc_dec_t d{(uint16_t)row[0], (uint16_t)row[1], (uint16_t)row[2], (uint16_t)row[3],
                        (int16_t)row[4], (int16_t)row[5], (int16_t)row[6]};
            for (int i = 0; i < n; ++i) {
                std::vector<uint8_t> h(B / 8);
                std::vector<int16_t> soft(68 * 384), lq16(68 * 384);
                int its = 0;
                const int s = orc_decode_path(&llr[(size_t)i * E], h.data(), c.cfg[0], c.cfg[1],
                                              c.cfg[2], E, c.cfg[4], c.cfg[5], 0, &d,
                                              soft.data(), lq16.data(), nullptr, &its);
                if (s != st[i] || its != it[i] || memcmp(h.data(), &hard[hoff[i]], B / 8) != 0) {
                    fprintf(stderr, "algo 1 mismatch with the oracle: codeblock %d\n", i);
                    return 3;
                }
            }
            if (emu_decode_tiles(n, cfgs.data(), off.data(), row, llr.data(), pool.data(),
                                 slots.data(), hard.data(), st.data(), it.data(), nullptr,
                                 hoff.data(), loff.data(), 0) < 0)
                return 4;
        }
    }
    return 0;
}
}  // namespace

int main() {
    const Case cases[] = {
        {{1, 488, 1, 1536, 2, 2, 0}, -1.0f},     // BG1 Zc=24, rv 2
        {{2, 176, 1, 512, 0, 2, 0}, -2.0f},      // BG2 Zc=26
        {{2, 1016, 2, 3120, 3, 4, 0}, 0.0f},     // BG2 Zc=104, 16QAM, rv 3
        {{2, 56, 0, 1200, 0, 8, 0}, 3.0f},       // repetition (E > N - F), 256QAM
        {{1, 16, 0, 300, 1, 2, 0}, 1.0f},        // tiny BG1, filler bits
    };
    for (const Case &c : cases) {
        const int rc = run_case(c, 3);
        if (rc) {
            fprintf(stderr, "case bgn=%d A=%d E=%d failed (%d)\n", c.cfg[0], c.cfg[1], c.cfg[3], rc);
            return rc;
        }
    }
    printf("sanitize ok\n");
    return 0;
}
