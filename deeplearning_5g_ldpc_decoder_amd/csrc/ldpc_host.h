// Host-side helpers shared by the library's translation units.
#pragma once
#include <stdint.h>

#include <vector>

#include "../../include/nr_ldpc_mi355.h"
#include "ldpc_device.h"

namespace nrldpc {

// Derived code parameters (c_lib/nr_ldpc.h:54-79 without table pointers).
struct Params {
    int bgn = 0, A = 0, crctype = 0, E = 0, rv = 0, Qm = 0, harq_on = 0;
    int B = 0, Zc = 0, iLS = 0, K = 0, F = 0, N = 0, k0 = 0;
    int rows = 0, edges = 0;
};

// nullptr on success, else the reason the config is invalid.
const char *derive(const ldpc_in_config_t &in, Params &p);
void print_invalid(const ldpc_in_config_t &in, const char *why);

// Lifting-size set index of zc (Table 5.3.2-1), or -1.
int lifting_set(int zc);
const uint16_t *all_lifting_sizes(int *count);

// Device edge table of (bgn, zc): per edge (column * zc) | (V_ij mod zc) << 16.
void edge_table(int bgn, int zc, int ils, uint32_t *out);

// Words of one (bgn, zc) table block: the edge table (NR_BG*_EDGES words),
// then the pair table of the run-time packed rows, 3 words per row pair in
// row order (edges e0 + 2i, e0 + 2i + 1 of a row; an odd last edge alone):
// (2 * shift_a) | (2 * shift_b) << 16, 2 * column_a * zc, 2 * column_b * zc
// (byte units of int16 beliefs; the b words 0 for a lone edge).
int edge_block_words(int bgn);
void edge_block(int bgn, int zc, int ils, uint32_t *out);

// Device CRC tables (DecOut::crc): 4 * kCrcTabWords words.
void crc_device_tables(uint32_t *out);

void encode_codeword(const Params &p, const uint8_t *bits_b, uint8_t *cw);
void rate_match_bits(const Params &p, const uint8_t *cw, uint8_t *out_bits);

// Kernel-side views of a decoder row and of one codeblock (ldpc_device.h).
// The descriptor's buffer offsets are set to "none"; callers fill them.
// decoder_params leaves the scale tables of algos 2/3 unset (see ScaleSet).
DecParams decoder_params(const ldpc_decoder_cfg_t &c);

// True when the int16 fast row form is exact for the pair (alpha, beta)
// under the clip bounds: alpha >= 0, min_llr <= max_llr and every scaled
// minimum (alpha * max(m - beta, 0) + 2^14) >> 15 with m <= max_llr stays
// within [0, 32767], so no message wraps to int16 -32768.
bool i16_fast_ok(int alpha, int beta, int max_llr, int min_llr);

// Algo-2 trained tables ([iterations][rows] int16 alpha, beta per base
// graph) and algo-3 degree alpha (kMaxScaleDegree Q15 values, entry d-1 for
// check degree d): process-wide, defaults as the reference ships.
constexpr int kMaxScaleDegree = 19;
constexpr int kDefaultTrainedIterations = 32;
constexpr int kMaxTrainedIterations = 1024;
int set_trained_tables(int bgn, const int16_t *alpha, const int16_t *beta, int iterations);
int get_trained_tables(int bgn, int16_t *alpha, int16_t *beta, int capacity);
void set_degree_alpha(const int16_t *values);
void get_degree_alpha(int16_t *values);

// The scale words a decode of algo 2 or 3 reads (DecParams::scale): base
// graph 1's iters * 46 words at off[0], base graph 2's iters * 42 at off[1],
// built from the tables current when it is made.  fast: every word passes
// i16_fast_ok under the row's clip bounds.
struct ScaleSet {
    int algo = 0;
    std::vector<uint32_t> words;
    uint32_t off[2] = {0, 0};
    int iters = 0;
    int table_iters[2] = {0, 0};
    bool fast = true;
    bool simple = true;   // fast, every beta 0 and max_llr 32767 (DecParams::simple)
};
void make_scale_set(const ldpc_decoder_cfg_t &c, ScaleSet &s);
// Algo 2 decodes at most as many iterations as its base graph's table holds
// (the reference reads past the table); true for other algos.
bool scale_iterations_ok(const ScaleSet &s, int bgn, int max_iterations);
CbDesc codeblock_desc(const Params &p, uint32_t etab_off);
// Packs codeblock descriptors into tile workgroups: codeblocks of one
// (BG, Zc) form groups of whole waves (the fewest waves, up to 6, that the
// codeblocks fill to 7/8, and at most k_of(bgn, zc) codeblocks), and groups
// of one base graph are packed by decreasing size into workgroups of
// kTileWaves waves.  tiles receives kTileWaves descriptors per workgroup,
// order the descriptor indices group by group, lds_entries the largest
// workgroup's LQ footprint.  Base graph 1's workgroups come first.
void build_tiles(const std::vector<CbDesc> &descs, const uint32_t (*etab_off)[kMaxZc + 1],
                 int (*k_of)(int bgn, int zc), std::vector<TileDesc> &tiles,
                 std::vector<uint32_t> &order, int &lds_entries);

// Workgroup size for lifting size zc: one lane per lifted check, whole waves.
inline int workgroup_threads(int zc) { return ((zc + 63) / 64) * 64; }

}  // namespace nrldpc
