// Host half of libnrldpc_mi355.so: the reference-compatible parameter,
// CRC, encoder, rate-matching and test-channel entry points.  These produce
// or describe the decoder's inputs; decoding itself runs on the GPU
// (ldpc_runtime.hip).
#include <math.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <atomic>
#include <mutex>
#include <thread>
#include <vector>

#include "../../include/nr_ldpc_mi355.h"
#include "ldpc_host.h"
#include "nr_bg_tables.h"

// ---------------------------------------------------------------------------
// Base graphs in the reference's exported layout (c_lib/nr_ldpc.h:93-99),
// filled from the 3GPP tables at load time.
// ---------------------------------------------------------------------------
extern "C" {
uint16_t BG1_HBG_offset_array[BG1_NUM_ROWS + 1];
uint16_t BG1_HBG_array[BG1_MAX_NON_NULL_ELEMENTS];
uint16_t BG1_HBG_Vij_array[8][BG1_MAX_NON_NULL_ELEMENTS];
uint16_t BG2_HBG_offset_array[BG2_NUM_ROWS + 1];
uint16_t BG2_HBG_array[BG2_MAX_NON_NULL_ELEMENTS];
uint16_t BG2_HBG_Vij_array[8][BG2_MAX_NON_NULL_ELEMENTS];
int16_t HARQ_memory[BG1_NUM_COLS * MAX_ZC];
}

namespace {
struct TableInit {
    TableInit() {
        for (int r = 0; r <= NR_BG1_ROWS; ++r) BG1_HBG_offset_array[r] = kBG1RowStart[r];
        for (int r = 0; r <= NR_BG2_ROWS; ++r) BG2_HBG_offset_array[r] = kBG2RowStart[r];
        for (int e = 0; e < NR_BG1_EDGES; ++e) BG1_HBG_array[e] = kBG1EdgeCol[e];
        for (int e = 0; e < NR_BG2_EDGES; ++e) BG2_HBG_array[e] = kBG2EdgeCol[e];
        for (int s = 0; s < 8; ++s) {
            for (int e = 0; e < NR_BG1_EDGES; ++e) BG1_HBG_Vij_array[s][e] = kBG1Shift[s][e];
            for (int e = 0; e < NR_BG2_EDGES; ++e) BG2_HBG_Vij_array[s][e] = kBG2Shift[s][e];
        }
    }
} g_table_init;
}  // namespace

namespace nrldpc {

// TS 38.212 Table 5.3.2-1 (lifting sizes by set index iLS).
static const uint16_t kSets[8][8] = {
    {2, 4, 8, 16, 32, 64, 128, 256}, {3, 6, 12, 24, 48, 96, 192, 384},
    {5, 10, 20, 40, 80, 160, 320, 0}, {7, 14, 28, 56, 112, 224, 0, 0},
    {9, 18, 36, 72, 144, 288, 0, 0},  {11, 22, 44, 88, 176, 352, 0, 0},
    {13, 26, 52, 104, 208, 0, 0, 0},  {15, 30, 60, 120, 240, 0, 0, 0}};

int lifting_set(int zc) {
    for (int s = 0; s < 8; ++s)
        for (int i = 0; i < 8; ++i)
            if (kSets[s][i] == zc) return s;
    return -1;
}

const uint16_t *all_lifting_sizes(int *count) {
    static uint16_t sizes[51];
    static std::once_flag once;
    std::call_once(once, [] {
        int n = 0;
        for (int z = 2; z <= 384; ++z)
            if (lifting_set(z) >= 0) sizes[n++] = (uint16_t)z;
    });
    *count = 51;
    return sizes;
}

// Derivation of c_lib/ldpc_internal.c:66-169 (TS 38.212 5.2.2, 5.3.2,
// Table 5.4.2.1-2 with Ncb = N).  Returns a message for an invalid config.
const char *derive(const ldpc_in_config_t &in, Params &p) {
    p = Params{};
    p.bgn = in.bgn; p.A = in.A; p.crctype = in.crctype; p.E = in.E;
    p.rv = in.rv; p.Qm = in.Qm; p.harq_on = in.HARQ_on;
    static char msg[160];
    if (in.bgn != 1 && in.bgn != 2) {
        snprintf(msg, sizeof msg, "Invalid base graph number: %d. Must be 1 or 2.", in.bgn);
        return msg;
    }
    if (in.crctype > 3) {
        snprintf(msg, sizeof msg, "Invalid CRC type: %d. Must be in [0,1,2,3].", in.crctype);
        return msg;
    }
    if (!(in.Qm == 1 || in.Qm == 2 || in.Qm == 4 || in.Qm == 6 || in.Qm == 8)) {
        snprintf(msg, sizeof msg, "Invalid modulation order: %d. Must be in [1,2,4,6,8].", in.Qm);
        return msg;
    }
    if (in.HARQ_on > 1) {
        snprintf(msg, sizeof msg, "Invalid HARQ_on value: %d. Must be 0 or 1.", in.HARQ_on);
        return msg;
    }
    if (in.rv > 3) {
        snprintf(msg, sizeof msg, "Invalid redundancy version: %d. Must be in [0,1,2,3].", in.rv);
        return msg;
    }
    if (in.E % in.Qm) {
        snprintf(msg, sizeof msg, "Invalid E value: %d. Must be a multiple of Qm.", in.E);
        return msg;
    }
    if (in.A % 8) {
        snprintf(msg, sizeof msg, "Invalid A value: %d. Must be a multiple of 8.", in.A);
        return msg;
    }
    p.B = in.A + (in.crctype == 0 ? 16 : 24);
    const int bmax = in.bgn == 1 ? 8448 : 3840;
    if (p.B > bmax) {
        snprintf(msg, sizeof msg,
                 "Invalid A value: %d for BG%d. B Must be less than or equal to %d.", in.A,
                 in.bgn, bmax);
        return msg;
    }
    int kb = 22;
    if (in.bgn == 2) kb = p.B > 640 ? 10 : p.B > 560 ? 9 : p.B > 192 ? 8 : 6;
    int best = 0;
    for (int s = 0; s < 8; ++s)
        for (int i = 0; i < 8 && kSets[s][i]; ++i)
            if (kb * kSets[s][i] >= p.B) {
                if (!best || kSets[s][i] < best) { best = kSets[s][i]; p.iLS = s; }
                break;
            }
    p.Zc = best;
    p.K = (in.bgn == 1 ? 22 : 10) * p.Zc;
    p.N = (in.bgn == 1 ? 66 : 50) * p.Zc;
    p.F = p.K - p.B;
    static const int k0t[2][4] = {{0, 17, 33, 56}, {0, 13, 25, 43}};
    p.k0 = k0t[in.bgn - 1][in.rv] * p.Zc;
    p.rows = in.bgn == 1 ? NR_BG1_ROWS : NR_BG2_ROWS;
    p.edges = in.bgn == 1 ? NR_BG1_EDGES : NR_BG2_EDGES;
    return nullptr;
}

void print_invalid(const ldpc_in_config_t &in, const char *why) {
    if (why) fprintf(stderr, "%s\n", why);
    fprintf(stderr, "Invalid LDPC configuration:\n");
    fprintf(stderr, "bgn=%d, A=%d, crctype=%d, E=%d, rv=%d, Qm=%d, HARQ_on=%d\n", in.bgn, in.A,
            in.crctype, in.E, in.rv, in.Qm, in.HARQ_on);
}

bool i16_fast_ok(int alpha, int beta, int max_llr, int min_llr) {
    if (alpha < 0 || min_llr > max_llr) return false;
    const int64_t off = std::max(max_llr - beta, 0);
    return (((int64_t)alpha * off + (1 << 14)) >> 15) <= 32767;
}

DecParams decoder_params(const ldpc_decoder_cfg_t &c) {
    DecParams p;
    memset(&p, 0, sizeof p);
    p.max_it = c.max_iterations;
    p.alpha = (int16_t)c.alpha;   // the int16 view the reference's helpers take
    p.beta = (int16_t)c.beta;
    p.max_llr = c.max_llr;
    p.min_llr = c.min_llr;
    p.filler = c.filling_bit_value;
    p.falpha = (float)c.alpha / 32768.0f;
    p.fbeta = (float)c.beta / 128.0f;
    p.fmax = (float)c.max_llr / 128.0f;
    p.fmin = (float)c.min_llr / 128.0f;
    p.general = i16_fast_ok(p.alpha, p.beta, p.max_llr, p.min_llr) ? 0 : 1;
    p.simple = (!p.general && p.beta == 0 && p.max_llr == 32767) ? 1 : 0;
    return p;
}

// ---------------------------------------------------------------------------
// Algo-2 / algo-3 scale tables.  The reference compiles its trained tables
// (c_lib/BG1_trained_ldpc_weights.c, BG2_...) and its degree alpha defaults
// (c_lib/degree_fixedpoint_ldpc_decoder.c:19-43) into the library; here they
// are process-wide state with the same defaults, replaceable at run time.
// The defaults are generated from the rule both reference files state: Q15
// alpha 26623 for check degree <= 4, 24575 for 5..8, 22527 above 8; trained
// tables repeat it over 32 iterations with beta 0.
// ---------------------------------------------------------------------------
namespace {
int16_t rule_alpha(int degree) {
    return degree <= 4 ? 26623 : (degree <= 8 ? 24575 : 22527);
}

int row_degree(int bgn, int r) {
    const uint16_t *rs = bgn == 1 ? kBG1RowStart : kBG2RowStart;
    return rs[r + 1] - rs[r];
}

int bg_rows(int bgn) { return bgn == 1 ? NR_BG1_ROWS : NR_BG2_ROWS; }

struct TableState {
    std::mutex mu;
    std::vector<int16_t> alpha[2], beta[2];   // [iterations][rows]
    int iters[2] = {0, 0};
    int16_t degree_alpha[kMaxScaleDegree];

    TableState() {
        reset_trained(1);
        reset_trained(2);
        reset_degree();
    }
    void reset_trained(int bgn) {
        const int rows = bg_rows(bgn);
        iters[bgn - 1] = kDefaultTrainedIterations;
        alpha[bgn - 1].assign((size_t)iters[bgn - 1] * rows, 0);
        beta[bgn - 1].assign((size_t)iters[bgn - 1] * rows, 0);
        for (int t = 0; t < iters[bgn - 1]; ++t)
            for (int r = 0; r < rows; ++r)
                alpha[bgn - 1][(size_t)t * rows + r] = rule_alpha(row_degree(bgn, r));
    }
    void reset_degree() {
        for (int d = 1; d <= kMaxScaleDegree; ++d) degree_alpha[d - 1] = rule_alpha(d);
    }
};

TableState &tables() {
    static TableState t;
    return t;
}

uint32_t scale_word(int alpha, int beta) {
    return (uint32_t)(uint16_t)alpha | ((uint32_t)(uint16_t)beta << 16);
}
}  // namespace

int set_trained_tables(int bgn, const int16_t *alpha, const int16_t *beta, int iterations) {
    if (bgn != 1 && bgn != 2) return -4;
    TableState &ts = tables();
    std::lock_guard<std::mutex> g(ts.mu);
    if (!alpha) {
        ts.reset_trained(bgn);
        return 0;
    }
    if (!beta || iterations < 1 || iterations > kMaxTrainedIterations) return -4;
    const size_t n = (size_t)iterations * bg_rows(bgn);
    ts.alpha[bgn - 1].assign(alpha, alpha + n);
    ts.beta[bgn - 1].assign(beta, beta + n);
    ts.iters[bgn - 1] = iterations;
    return 0;
}

int get_trained_tables(int bgn, int16_t *alpha, int16_t *beta, int capacity) {
    if (bgn != 1 && bgn != 2) return -4;
    TableState &ts = tables();
    std::lock_guard<std::mutex> g(ts.mu);
    const int it = ts.iters[bgn - 1];
    const size_t n = (size_t)std::min(std::max(capacity, 0), it) * bg_rows(bgn);
    if (alpha) std::copy(ts.alpha[bgn - 1].begin(), ts.alpha[bgn - 1].begin() + n, alpha);
    if (beta) std::copy(ts.beta[bgn - 1].begin(), ts.beta[bgn - 1].begin() + n, beta);
    return it;
}

void set_degree_alpha(const int16_t *values) {
    TableState &ts = tables();
    std::lock_guard<std::mutex> g(ts.mu);
    if (values)
        std::copy(values, values + kMaxScaleDegree, ts.degree_alpha);
    else
        ts.reset_degree();
}

void get_degree_alpha(int16_t *values) {
    TableState &ts = tables();
    std::lock_guard<std::mutex> g(ts.mu);
    std::copy(ts.degree_alpha, ts.degree_alpha + kMaxScaleDegree, values);
}

void make_scale_set(const ldpc_decoder_cfg_t &c, ScaleSet &s) {
    TableState &ts = tables();
    std::lock_guard<std::mutex> g(ts.mu);
    const int beta_row = (int16_t)c.beta;
    s.algo = c.algo;
    s.words.clear();
    s.fast = true;
    s.simple = c.max_llr == 32767;
    // algo 2 reads iteration t's words for t < max_iterations; algo 3 one set
    s.iters = c.algo == 2 ? std::max((int)c.max_iterations, 1) : 1;
    for (int bgn = 1; bgn <= 2; ++bgn) {
        const int rows = bg_rows(bgn);
        s.off[bgn - 1] = (uint32_t)s.words.size();
        s.table_iters[bgn - 1] = ts.iters[bgn - 1];
        for (int t = 0; t < s.iters; ++t)
            for (int r = 0; r < rows; ++r) {
                int a, b;
                if (c.algo == 2) {
                    // rows past a table that is too short are never read: a
                    // decode whose max_iterations exceeds its base graph's
                    // table is refused
                    const int tt = std::min(t, ts.iters[bgn - 1] - 1);
                    a = ts.alpha[bgn - 1][(size_t)tt * rows + r];
                    b = ts.beta[bgn - 1][(size_t)tt * rows + r];
                } else {
                    // exact degree d at entry d-1; the BG1/BG2 degrees are
                    // 3..19, inside the table
                    // (c_lib/degree_fixedpoint_ldpc_decoder.c:65-80)
                    a = ts.degree_alpha[row_degree(bgn, r) - 1];
                    b = beta_row;
                }
                s.words.push_back(scale_word(a, b));
                if (!i16_fast_ok(a, b, c.max_llr, c.min_llr)) s.fast = false;
                if ((int16_t)b != 0) s.simple = false;
            }
    }
}

bool scale_iterations_ok(const ScaleSet &s, int bgn, int max_iterations) {
    return s.algo != 2 || max_iterations <= s.table_iters[bgn - 1];
}

CbDesc codeblock_desc(const Params &p, uint32_t etab_off) {
    CbDesc d;
    memset(&d, 0, sizeof d);
    d.bgn = (uint16_t)p.bgn; d.Zc = (uint16_t)p.Zc; d.B = (uint16_t)p.B;
    d.K = (uint16_t)p.K; d.N = (uint16_t)p.N; d.E = (uint16_t)p.E;
    d.k0 = (uint16_t)p.k0; d.Qm = (uint16_t)p.Qm;
    d.crctype = (uint8_t)p.crctype; d.harq_on = (uint8_t)p.harq_on;
    d.etab_off = etab_off;
    d.soft_off = kNone;
    d.lq_off = kNone;
    return d;
}

// Waves of one group of lifting size zc: the fewest waves (at most 6) whose
// codeblocks fill at least 7/8 of their lanes, else the best-filled count.
static int group_waves(int zc) {
    int best = 0, best_used = 0;
    for (int w = 1; w <= 6; ++w) {
        const int k = 64 * w / zc;
        if (k < 1) continue;
        const int used = k * zc;
        if (8 * used >= 7 * 64 * w) return w;
        if (best == 0 || used * 64 * best > best_used * 64 * w) {
            best = w;
            best_used = used;
        }
    }
    return best;
}

void build_tiles(const std::vector<CbDesc> &descs, const uint32_t (*etab_off)[kMaxZc + 1],
                 int (*k_of)(int bgn, int zc), std::vector<TileDesc> &tiles,
                 std::vector<uint32_t> &order, int &lds_entries) {
    std::vector<uint32_t> idx(descs.size());
    for (size_t i = 0; i < idx.size(); ++i) idx[i] = (uint32_t)i;
    std::stable_sort(idx.begin(), idx.end(), [&](uint32_t a, uint32_t b) {
        if (descs[a].bgn != descs[b].bgn) return descs[a].bgn < descs[b].bgn;
        return descs[a].Zc > descs[b].Zc;
    });
    // groups: up to k codeblocks of one (BG, Zc) on whole waves
    struct Group {
        uint32_t first;
        uint16_t count, bgn, Zc, lds_stride, waves;
    };
    std::vector<Group> groups;
    size_t i = 0;
    while (i < idx.size()) {
        const CbDesc &d0 = descs[idx[i]];
        const int w0 = group_waves(d0.Zc);
        const int k = std::max(1, std::min(64 * w0 / (int)d0.Zc, k_of(d0.bgn, d0.Zc)));
        size_t j = i;
        while (j < idx.size() && j - i < (size_t)k && descs[idx[j]].bgn == d0.bgn &&
               descs[idx[j]].Zc == d0.Zc)
            ++j;
        Group g;
        g.first = (uint32_t)i;
        g.count = (uint16_t)(j - i);
        g.bgn = d0.bgn;
        g.Zc = d0.Zc;
        g.lds_stride = (uint16_t)(((d0.N + 2 * d0.Zc) + 7) & ~7);
        g.waves = (uint16_t)((g.count * d0.Zc + 63) / 64);
        groups.push_back(g);
        i = j;
    }
    // workgroups: per base graph, groups by decreasing waves into the open
    // workgroup with the least room that holds them (kTileWaves waves and
    // kMaxTileCodeblocks CRC slots each)
    std::stable_sort(groups.begin(), groups.end(), [](const Group &a, const Group &b) {
        if (a.bgn != b.bgn) return a.bgn < b.bgn;
        return a.waves > b.waves;
    });
    struct Wg {
        int bgn, waves = 0, slots = 0;
        std::vector<int> members;
    };
    std::vector<Wg> wgs;
    std::vector<int> open[kTileWaves];   // open workgroups by free waves - 1
    int cur_bgn = 0;
    for (int gi = 0; gi < (int)groups.size(); ++gi) {
        const Group &g = groups[(size_t)gi];
        if (g.bgn != cur_bgn) {
            for (auto &o : open) o.clear();
            cur_bgn = g.bgn;
        }
        int chosen = -1;
        for (int f = g.waves; f <= kTileWaves && chosen < 0; ++f) {
            auto &lst = open[f - 1];
            for (size_t q = 0; q < lst.size(); ++q)
                if (wgs[(size_t)lst[q]].slots + g.count <= kMaxTileCodeblocks) {
                    chosen = lst[q];
                    lst.erase(lst.begin() + (long)q);
                    break;
                }
        }
        if (chosen < 0) {
            chosen = (int)wgs.size();
            wgs.push_back(Wg());
            wgs.back().bgn = g.bgn;
        }
        Wg &w = wgs[(size_t)chosen];
        w.members.push_back(gi);
        w.waves += g.waves;
        w.slots += g.count;
        const int free_waves = kTileWaves - w.waves;
        if (free_waves > 0 && w.slots < kMaxTileCodeblocks) open[free_waves - 1].push_back(chosen);
    }
    tiles.assign(wgs.size() * kTileWaves, TileDesc());
    lds_entries = 0;
    for (size_t wi = 0; wi < wgs.size(); ++wi) {
        const Wg &w = wgs[wi];
        TileDesc *t = &tiles[wi * kTileWaves];
        int wave = 0, slot = 0, lds = 0;
        for (int gi : w.members) {
            const Group &g = groups[(size_t)gi];
            TileDesc d;
            d.first = g.first;
            d.etab_off = etab_off[g.bgn - 1][g.Zc];
            d.lds_base = (uint32_t)lds;
            d.count = g.count;
            d.bgn = g.bgn;
            d.Zc = g.Zc;
            d.lds_stride = g.lds_stride;
            d.lane_base = (uint16_t)(wave * 64);
            d.slot_base = (uint16_t)slot;
            for (int v = wave; v < wave + g.waves; ++v) t[v] = d;
            wave += g.waves;
            slot += g.count;
            lds += g.count * g.lds_stride;
        }
        // waves outside any group run the first group's rows with no
        // codeblock: their lanes address that group's first LQ region, as a
        // group's lanes past its last codeblock do
        for (int v = wave; v < kTileWaves; ++v) {
            t[v] = t[0];
            t[v].count = 0;
            t[v].lane_base = (uint16_t)(v * 64);
        }
        lds_entries = std::max(lds_entries, lds);
    }
    order = idx;
}

void edge_table(int bgn, int zc, int ils, uint32_t *out) {
    const int ne = bgn == 1 ? NR_BG1_EDGES : NR_BG2_EDGES;
    const uint8_t *col = bgn == 1 ? kBG1EdgeCol : kBG2EdgeCol;
    const uint16_t *sh = bgn == 1 ? kBG1Shift[ils] : kBG2Shift[ils];
    for (int e = 0; e < ne; ++e)
        out[e] = (uint32_t)(col[e] * zc) | ((uint32_t)(sh[e] % zc) << 16);
}

int edge_block_words(int bgn) {
    const int rows = bgn == 1 ? NR_BG1_ROWS : NR_BG2_ROWS;
    const uint16_t *start = bgn == 1 ? kBG1RowStart : kBG2RowStart;
    int pairs = 0;
    for (int r = 0; r < rows; ++r) pairs += (start[r + 1] - start[r] + 1) / 2;
    return (bgn == 1 ? NR_BG1_EDGES : NR_BG2_EDGES) + 3 * pairs;
}

void edge_block(int bgn, int zc, int ils, uint32_t *out) {
    edge_table(bgn, zc, ils, out);
    const int ne = bgn == 1 ? NR_BG1_EDGES : NR_BG2_EDGES;
    const int rows = bgn == 1 ? NR_BG1_ROWS : NR_BG2_ROWS;
    const uint16_t *start = bgn == 1 ? kBG1RowStart : kBG2RowStart;
    uint32_t *pt = out + ne;
    for (int r = 0; r < rows; ++r)
        for (int e = start[r]; e < start[r + 1]; e += 2) {
            const uint32_t a = out[e];
            const uint32_t b = e + 1 < start[r + 1] ? out[e + 1] : 0u;
            *pt++ = 2u * (a >> 16) | (2u * (b >> 16)) << 16;
            *pt++ = 2u * (a & 0xFFFFu);
            *pt++ = 2u * (b & 0xFFFFu);
        }
}

}  // namespace nrldpc

using namespace nrldpc;

// ---------------------------------------------------------------------------
// CRC (c_lib/nr_crc.c:93-179): byte table per polynomial, built once.
// ---------------------------------------------------------------------------
namespace {
struct CrcTables {
    uint32_t tab[4][256];
    int width[4] = {16, 24, 24, 24};
    uint32_t poly[4] = {0x1021u, 0x864CFBu, 0x800063u, 0xB2B117u};
    CrcTables() {
        for (int t = 0; t < 4; ++t) {
            const uint32_t top = 1u << (width[t] - 1), mask = (1u << width[t]) - 1u;
            for (int i = 0; i < 256; ++i) {
                uint32_t r = (uint32_t)i << (width[t] - 8);
                for (int b = 0; b < 8; ++b) r = (r & top) ? ((r << 1) ^ poly[t]) : (r << 1);
                tab[t][i] = r & mask;
            }
        }
    }
};
const CrcTables &crc_tables() {
    static const CrcTables t;
    return t;
}
}  // namespace

namespace nrldpc {
// Per CRC type: the byte table, then x^(8m) mod g for m = 0..kMaxCrcBytes
// (t * x^8 mod g = ((t << 8) mod x^w) ^ tab[t >> (w - 8)]).  A lane whose
// byte slice has m bytes after it multiplies its partial remainder by entry m.
// Then x^m mod g for m < kMaxCrcBits: bit i of a B-bit block contributes
// x^(B - 1 - i) to the block's remainder (times x^w, which is invertible
// mod g, so a zero test needs no x^w).
void crc_device_tables(uint32_t *out) {
    const CrcTables &T = crc_tables();
    for (int t = 0; t < 4; ++t) {
        uint32_t *o = out + (size_t)t * kCrcTabWords;
        const int w = T.width[t];
        const uint32_t mask = (1u << w) - 1u;
        for (int i = 0; i < 256; ++i) o[i] = T.tab[t][i];
        uint32_t r = 1u;
        for (int m = 0; m <= kMaxCrcBytes; ++m) {
            o[256 + m] = r;
            r = ((r << 8) & mask) ^ T.tab[t][(r >> (w - 8)) & 0xFFu];
        }
        // x^m mod g, one bit at a time
        r = 1u;
        for (int m = 0; m < kMaxCrcBits; ++m) {
            o[kCrcBitOff + m] = r;
            r = ((r >> (w - 1)) & 1u) ? ((r << 1) & mask) ^ T.poly[t] : (r << 1) & mask;
        }
    }
}
}  // namespace nrldpc

extern "C" uint32_t calculate_crc(const uint8_t *data, int len, int crctype) {
    if (crctype < 0 || crctype > 3) return 0;
    const CrcTables &T = crc_tables();
    const int sh = T.width[crctype] - 8;
    const uint32_t mask = (1u << T.width[crctype]) - 1u;
    uint32_t r = 0;
    for (int i = 0; i < len; ++i) r = ((r << 8) ^ T.tab[crctype][((r >> sh) ^ data[i]) & 0xFFu]) & mask;
    return r;
}

extern "C" int gen_ldpc_parameters(ldpc_config_t *params, ldpc_in_config_t *inconfig) {
    Params p;
    const char *why = derive(*inconfig, p);
    params->bgn = inconfig->bgn; params->A = inconfig->A; params->crctype = inconfig->crctype;
    params->E = inconfig->E; params->rv = inconfig->rv; params->Qm = inconfig->Qm;
    params->HARQ_on = inconfig->HARQ_on;
    if (why) {
        fprintf(stderr, "%s\n", why);
        return -1;
    }
    params->B = (uint16_t)p.B; params->Zc = (uint16_t)p.Zc; params->iLS = (uint16_t)p.iLS;
    params->K = (uint16_t)p.K; params->F = (uint16_t)p.F; params->N = (uint16_t)p.N;
    params->k0 = (uint16_t)p.k0;
    if (p.bgn == 1) {
        params->Vij_array = BG1_HBG_Vij_array[p.iLS];
        params->HBG_array = BG1_HBG_array;
        params->HBG_offset_array = BG1_HBG_offset_array;
    } else {
        params->Vij_array = BG2_HBG_Vij_array[p.iLS];
        params->HBG_array = BG2_HBG_array;
        params->HBG_offset_array = BG2_HBG_offset_array;
    }
    return 0;
}

extern "C" float dequantize_int16_to_llr(int16_t q) { return (float)q / 128.0f; }

// ---------------------------------------------------------------------------
// Encoder (TS 38.212 5.3.2; c_lib/ldpc_encoder.c:221-242) and rate matching
// (TS 38.212 5.4.2; c_lib/ldpc_rate_match.c:27-106).  Input generation only.
// Codeword layout: (N + 2Zc) bits, base-graph column c at [c*Zc, (c+1)*Zc).
// ---------------------------------------------------------------------------
namespace nrldpc {

static inline int find_edge(int bgn, int row, int col) {
    const uint16_t *rs = bgn == 1 ? kBG1RowStart : kBG2RowStart;
    const uint8_t *ec = bgn == 1 ? kBG1EdgeCol : kBG2EdgeCol;
    for (int e = rs[row]; e < rs[row + 1]; ++e)
        if (ec[e] == col) return e;
    return -1;
}

void encode_codeword(const Params &p, const uint8_t *bits_b, uint8_t *cw) {
    const int Zc = p.Zc, bgn = p.bgn;
    const int kb = bgn == 1 ? 22 : 10, ncol = bgn == 1 ? 68 : 52;
    const uint16_t *rs = bgn == 1 ? kBG1RowStart : kBG2RowStart;
    const uint8_t *ec = bgn == 1 ? kBG1EdgeCol : kBG2EdgeCol;
    const uint16_t *sh = bgn == 1 ? kBG1Shift[p.iLS] : kBG2Shift[p.iLS];
    memset(cw, 0, (size_t)ncol * Zc);
    memcpy(cw, bits_b, (size_t)p.B);
    auto at = [&](int e, int zx) { return ec[e] * Zc + (sh[e] % Zc + zx) % Zc; };
    // lambda_r = systematic part of core rows 0..3
    std::vector<uint8_t> lam(4 * (size_t)Zc, 0);
    for (int r = 0; r < 4; ++r)
        for (int e = rs[r]; e < rs[r + 1]; ++e)
            if (ec[e] < kb)
                for (int zx = 0; zx < Zc; ++zx) lam[r * Zc + zx] ^= cw[at(e, zx)];
    // Core parity (double-diagonal structure of both base graphs): summing
    // the four core rows leaves a single circulant on the first parity group.
    const int ea = find_edge(bgn, 0, kb);                  // row 0 on p0
    const int eb = find_edge(bgn, bgn == 1 ? 1 : 2, kb);   // unpaired p0 circulant
    const int sa = sh[ea] % Zc, sb = sh[eb] % Zc;
    uint8_t *p0 = cw + kb * Zc, *p1 = p0 + Zc, *p2 = p1 + Zc, *p3 = p2 + Zc;
    for (int zx = 0; zx < Zc; ++zx)
        p0[(zx + sb) % Zc] = lam[zx] ^ lam[Zc + zx] ^ lam[2 * Zc + zx] ^ lam[3 * Zc + zx];
    for (int zx = 0; zx < Zc; ++zx) {
        const uint8_t rot = p0[(zx + sa) % Zc];
        p1[zx] = lam[zx] ^ rot;
        p3[zx] = lam[3 * Zc + zx] ^ rot;
    }
    for (int zx = 0; zx < Zc; ++zx)
        p2[zx] = bgn == 1 ? (uint8_t)(lam[2 * Zc + zx] ^ p3[zx]) : (uint8_t)(lam[Zc + zx] ^ p1[zx]);
    // extension rows: one identity parity column each
    for (int r = 4; r < p.rows; ++r) {
        uint8_t *pr = cw + (kb + r) * Zc;
        for (int e = rs[r]; e < rs[r + 1]; ++e) {
            if (ec[e] >= kb + 4) continue;
            for (int zx = 0; zx < Zc; ++zx) pr[zx] ^= cw[at(e, zx)];
        }
    }
}

void rate_match_bits(const Params &p, const uint8_t *cw, uint8_t *out_bits) {
    const int Zc = p.Zc, N = p.N, E = p.E, Qm = p.Qm, S = E / Qm;
    const int lo = p.B - 2 * Zc, hi = p.K - 2 * Zc;
    std::vector<uint8_t> e((size_t)E + 1);
    int pos = p.k0;
    for (int k = 0; k < E;) {
        if (pos < lo || pos >= hi) e[k++] = cw[pos + 2 * Zc];
        if (++pos == N) pos = 0;
    }
    for (int j = 0; j < S; ++j)
        for (int i = 0; i < Qm; ++i) out_bits[j * Qm + i] = e[(size_t)i * S + j];
}

}  // namespace nrldpc

extern "C" int32_t c_ldpc_5g_encode(uint8_t *infobytes, uint8_t *rm_bytes,
                                    ldpc_in_config_t *inconfig) {
    Params p;
    if (const char *why = derive(*inconfig, p)) {
        print_invalid(*inconfig, why);
        return -2;
    }
    std::vector<uint8_t> b(p.B), cw((size_t)(p.N + 2 * p.Zc)), rm((size_t)p.E);
    for (int i = 0; i < p.A; ++i) b[i] = (infobytes[i >> 3] >> (7 - (i & 7))) & 1u;
    const uint32_t crc = calculate_crc(infobytes, p.A / 8, p.crctype);
    const int clen = p.B - p.A;
    for (int i = 0; i < clen; ++i) b[p.A + i] = (crc >> (clen - 1 - i)) & 1u;
    encode_codeword(p, b.data(), cw.data());
    rate_match_bits(p, cw.data(), rm.data());
    memset(rm_bytes, 0, (size_t)(p.E + 7) / 8);
    for (int i = 0; i < p.E; ++i) rm_bytes[i >> 3] |= (uint8_t)(rm[i] << (7 - (i & 7)));
    return 0;
}

// CUDA/cuda_ldpc_5g_encoder.cu:47-84: the CUDA library's top-level encoder
// produces the same bytes as c_ldpc_5g_encode (CRC attach, encode, rate
// match); bound by scripts/sim_cuda_ldpc_encoder_decoder.py:65.
extern "C" int32_t cuda_ldpc_5g_encode(uint8_t *infobytes, uint8_t *rm_bytes,
                                       ldpc_in_config_t *inconfig) {
    return c_ldpc_5g_encode(infobytes, rm_bytes, inconfig);
}

// ---------------------------------------------------------------------------
// Test channel (c_lib/tx_rx_process.c:286-324).  Gray QAM of TS 38.211
// 5.1.2-5.1.6: bits 0,2,4,.. drive the I axis and 1,3,5,.. the Q axis, each
// axis a Gray PAM; max-log LLR per bit = (d1^2 - d0^2) / noise_var over the
// axis levels (equal to the reference's piecewise forms).  Noise:
// xoshiro256** + Box-Muller, seeded by ldpc_mi355_channel_seed.
// ---------------------------------------------------------------------------
namespace {
struct Rng {
    uint64_t s[4];
    explicit Rng(uint64_t seed) { reseed(seed); }
    void reseed(uint64_t seed) {
        uint64_t z = seed + 0x9E3779B97F4A7C15ull;
        for (auto &v : s) {
            z += 0x9E3779B97F4A7C15ull;
            uint64_t x = z;
            x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
            x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
            v = x ^ (x >> 31);
        }
    }
    static uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
    uint64_t next() {
        const uint64_t r = rotl(s[1] * 5, 7) * 9, t = s[1] << 17;
        s[2] ^= s[0]; s[3] ^= s[1]; s[1] ^= s[2]; s[0] ^= s[3];
        s[2] ^= t; s[3] = rotl(s[3], 45);
        return r;
    }
    double uniform() { return ((double)(next() >> 11) + 0.5) * (1.0 / 9007199254740992.0); }
    void normal2(double &a, double &b) {
        const double u1 = uniform(), u2 = uniform();
        const double rad = sqrt(-2.0 * log(u1));
        a = rad * cos(6.283185307179586 * u2);
        b = rad * sin(6.283185307179586 * u2);
    }
};
std::mutex g_rng_mu;
Rng g_rng(0);

// Gray PAM level of an axis from its bits (MSB first): TS 38.211 5.1.
inline double pam_level(const uint8_t *b, int nb) {
    // nb = 1: 1-2b0; nb = 2: (1-2b0)(2-(1-2b1)); nb = 3/4 nested likewise
    double v = 0;
    if (nb == 1) v = 1 - 2 * b[0];
    else if (nb == 2) v = (1 - 2 * b[0]) * (2 - (1 - 2 * b[1]));
    else if (nb == 3) v = (1 - 2 * b[0]) * (4 - (1 - 2 * b[1]) * (2 - (1 - 2 * b[2])));
    else v = (1 - 2 * b[0]) * (8 - (1 - 2 * b[1]) * (4 - (1 - 2 * b[2]) * (2 - (1 - 2 * b[3]))));
    return v;
}

inline int16_t quantize(double llr) {
    const float scaled = (float)llr * 128.0f;
    int32_t q = (int32_t)(scaled >= 0.0f ? scaled + 0.5f : scaled - 0.5f);
    if (q > 10000) q = 10000;
    if (q < -10000) q = -10000;
    return (int16_t)q;
}

// Channel LLRs of nsym symbols; emit(i, llr) receives LLR i (double).
template <typename Emit>
void channel_llrs(const uint8_t *bits, size_t nsym, int Qm, double nv, Rng &rng, Emit emit) {
    const double sigma = sqrt(nv / 2.0);
    if (Qm == 1) {
        const double A = 1.0 / sqrt(2.0);
        for (size_t i = 0; i < nsym; ++i) {
            const double s = (1 - 2 * bits[i]) * A;
            double na, nb;
            rng.normal2(na, nb);
            emit(i, 4.0 * ((s + sigma * na) + (s + sigma * nb)) * A / nv);
        }
        return;
    }
    const int nb = Qm / 2;
    const double A = Qm == 2 ? 1 / sqrt(2.0) : Qm == 4 ? 1 / sqrt(10.0) : Qm == 6 ? 1 / sqrt(42.0)
                                                                                  : 1 / sqrt(170.0);
    const int L = 1 << nb;
    double lev[16];
    uint8_t lbits[16][4];
    for (int v = 0; v < L; ++v) {
        for (int k = 0; k < nb; ++k) lbits[v][k] = (v >> (nb - 1 - k)) & 1;
        lev[v] = pam_level(lbits[v], nb) * A;
    }
    uint8_t axb[2][4];
    for (size_t i = 0; i < nsym; ++i) {
        const uint8_t *b = bits + i * Qm;
        for (int k = 0; k < nb; ++k) { axb[0][k] = b[2 * k]; axb[1][k] = b[2 * k + 1]; }
        double na, nbv;
        rng.normal2(na, nbv);
        const double r[2] = {pam_level(axb[0], nb) * A + sigma * na,
                             pam_level(axb[1], nb) * A + sigma * nbv};
        for (int ax = 0; ax < 2; ++ax)
            for (int k = 0; k < nb; ++k) {
                double d0 = 1e300, d1 = 1e300;
                for (int v = 0; v < L; ++v) {
                    const double d = (r[ax] - lev[v]) * (r[ax] - lev[v]);
                    if (lbits[v][k]) d1 = d < d1 ? d : d1;
                    else d0 = d < d0 ? d : d0;
                }
                emit(i * Qm + 2 * k + ax, (d1 - d0) / nv);
            }
    }
}

void channel_symbols(const uint8_t *bits, size_t nsym, int Qm, double nv, Rng &rng, int16_t *out) {
    channel_llrs(bits, nsym, Qm, nv, rng, [out](size_t i, double v) { out[i] = quantize(v); });
}

// Random payload + CRC (B bits, one per byte) from rng.
void random_block(const Params &p, Rng &rng, uint8_t *b, std::vector<uint8_t> &bytes) {
    for (int k = 0; k < p.A; k += 64) {
        uint64_t w = rng.next();
        for (int j = 0; j < 64 && k + j < p.A; ++j) b[k + j] = (w >> j) & 1u;
    }
    memset(bytes.data(), 0, bytes.size());
    for (int k = 0; k < p.A; ++k) bytes[k >> 3] |= (uint8_t)(b[k] << (7 - (k & 7)));
    const uint32_t crc = calculate_crc(bytes.data(), p.A / 8, p.crctype);
    const int clen = p.B - p.A;
    for (int k = 0; k < clen; ++k) b[p.A + k] = (crc >> (clen - 1 - k)) & 1u;
}

}  // namespace

extern "C" void ldpc_mi355_channel_seed(uint64_t seed) {
    std::lock_guard<std::mutex> g(g_rng_mu);
    g_rng.reseed(seed);
}

extern "C" int tx_rx_process(const uint8_t *inbytes, size_t num_bits, uint16_t Qm,
                             float noise_power_db, int16_t *llr_out) {
    if (!inbytes || !llr_out || num_bits == 0) return -1;
    if (!(Qm == 1 || Qm == 2 || Qm == 4 || Qm == 6 || Qm == 8)) return -1;
    if (num_bits % Qm) return -1;
    std::vector<uint8_t> bits(num_bits);
    for (size_t i = 0; i < num_bits; ++i) bits[i] = (inbytes[i >> 3] >> (7 - (i & 7))) & 1u;
    const double nv = pow(10.0, (double)noise_power_db / 10.0);
    std::lock_guard<std::mutex> g(g_rng_mu);
    channel_symbols(bits.data(), num_bits / Qm, Qm, nv, g_rng, llr_out);
    return 0;
}

// Batch synthesis for simulations and benchmarks: n codeblocks of one
// config, random payloads, encoded, rate-matched and passed through the
// channel at noise_power_db, on worker threads (one RNG stream per
// codeblock, derived from seed and the codeblock index, so the result does
// not depend on the thread count).  llr_out: n*E int16; payload_out
// (optional): n*A payload bits, one per byte.  0 or -2.
extern "C" int32_t ldpc_mi355_synth_batch(const ldpc_in_config_t *cfg, int32_t n,
                                          float noise_power_db, uint64_t seed, int32_t threads,
                                          int16_t *llr_out, uint8_t *payload_out) {
    Params p;
    if (const char *why = derive(*cfg, p)) {
        print_invalid(*cfg, why);
        return -2;
    }
    const double nv = pow(10.0, (double)noise_power_db / 10.0);
    std::atomic<int> next{0};
    auto worker = [&] {
        std::vector<uint8_t> b(p.B), cw((size_t)(p.N + 2 * p.Zc)), rm((size_t)p.E);
        std::vector<uint8_t> bytes((size_t)p.A / 8 + 1);
        for (int i; (i = next.fetch_add(1)) < n;) {
            Rng rng(seed * 0x100000001B3ull + (uint64_t)i);
            random_block(p, rng, b.data(), bytes);
            encode_codeword(p, b.data(), cw.data());
            rate_match_bits(p, cw.data(), rm.data());
            channel_symbols(rm.data(), (size_t)p.E / p.Qm, p.Qm, nv, rng,
                            llr_out + (size_t)i * p.E);
            if (payload_out) memcpy(payload_out + (size_t)i * p.A, b.data(), (size_t)p.A);
        }
    };
    if (threads < 1) threads = 1;
    std::vector<std::thread> pool;
    for (int t = 1; t < threads; ++t) pool.emplace_back(worker);
    worker();
    for (auto &t : pool) t.join();
    return 0;
}

// Float inputs of the trainable decoder in its belief layout
// (deep_learning/ldpc_training_data.py:24-82): random payload + CRC,
// encoded, rate-matched, unquantized channel LLRs (py5gphy's demodulators
// give the same max-log values as the test channel above, in float), then the
// float rate recovery of py5gphy/ldpc/nr_ldpc_raterecover.py:15-68: Qm
// deinterleave, circular buffer from k0 over Ncb = N skipping filler
// positions with repeats summed, filler positions set to 10 * max|LLR| of the
// codeblock, and 2Zc leading zeros for the punctured columns.  One RNG stream
// per codeblock from (seed, index).  noise_power_db: n values; llr_out:
// n*(N+2Zc) floats; bits_out (optional): n*B bits (payload + CRC), one per
// byte.  0 or -2.
extern "C" int32_t ldpc_mi355_synth_float_batch(const ldpc_in_config_t *cfg, int32_t n,
                                                const float *noise_power_db, uint64_t seed,
                                                int32_t threads, float *llr_out,
                                                uint8_t *bits_out) {
    Params p;
    if (const char *why = derive(*cfg, p)) {
        print_invalid(*cfg, why);
        return -2;
    }
    if (n < 0 || (n > 0 && (!noise_power_db || !llr_out))) return -2;
    const int nn = p.N + 2 * p.Zc, S = p.E / p.Qm;
    const int f0 = p.B - 2 * p.Zc, f1 = p.K - 2 * p.Zc;   // filler span in the buffer
    std::atomic<int> next{0};
    auto worker = [&] {
        std::vector<uint8_t> b(p.B), cw((size_t)nn), rm((size_t)p.E);
        std::vector<uint8_t> bytes((size_t)p.A / 8 + 1);
        std::vector<double> fe((size_t)p.E), buf((size_t)p.N);
        for (int i; (i = next.fetch_add(1)) < n;) {
            Rng rng(seed * 0x100000001B3ull + (uint64_t)i);
            random_block(p, rng, b.data(), bytes);
            encode_codeword(p, b.data(), cw.data());
            rate_match_bits(p, cw.data(), rm.data());
            const double nv = pow(10.0, (double)noise_power_db[i] / 10.0);
            double amax = 0.0;
            channel_llrs(rm.data(), (size_t)S, p.Qm, nv, rng, [&](size_t k, double v) {
                const float f = (float)v;   // the demodulators return float32
                fe[k] = f;
                amax = fabs((double)f) > amax ? fabs((double)f) : amax;
            });
            std::fill(buf.begin(), buf.end(), 0.0);
            int pos = p.k0;
            for (int k = 0; k < p.E; ++k) {
                while (pos >= f0 && pos < f1) pos = (pos + 1) % p.N;
                buf[pos] += fe[(size_t)(k % S) * p.Qm + k / S];
                pos = (pos + 1) % p.N;
            }
            const double fill = (double)(float)amax * 10.0;
            for (int k = f0; k < f1; ++k) buf[k] = fill;
            float *out = llr_out + (size_t)i * nn;
            for (int k = 0; k < 2 * p.Zc; ++k) out[k] = 0.0f;
            for (int k = 0; k < p.N; ++k) out[2 * p.Zc + k] = (float)buf[k];
            if (bits_out) memcpy(bits_out + (size_t)i * p.B, b.data(), (size_t)p.B);
        }
    };
    if (threads < 1) threads = 1;
    std::vector<std::thread> pool;
    for (int t = 1; t < threads; ++t) pool.emplace_back(worker);
    worker();
    for (auto &t : pool) t.join();
    return 0;
}
