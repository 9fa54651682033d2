// Device-side data layout shared by the HIP kernels and the host runtime.
//
// One decode launch processes a batch of codeblocks.  Everything a
// workgroup needs about its codeblock is in one CbDesc (read once into
// SGPRs); the arrays it reads and writes are flat device buffers addressed
// by the offsets in the descriptor:
//
//   llr   int16  E values per codeblock, rate-matched + Qm-interleaved order
//   soft  int16  N+2Zc values per codeblock (HARQ soft buffer, in/out)
//   hard  uint8  B/8 bytes per codeblock, MSB-first
//   lq    int16 | float  N+2Zc posteriors per codeblock (optional)
//   work  float  nnz*Zc check-to-variable messages (BP only)
//   etab  uint32 per (BG, Zc): one word per base-graph edge,
//                low 16 bits = column*Zc, high 16 bits = V_ij mod Zc
#pragma once
#include <stdint.h>

namespace nrldpc {

constexpr uint32_t kNone = 0xFFFFFFFFu;
constexpr int kMaxZc = 384;
// CRC tables on the device, per CRC type (CRC16, 24A, 24B, 24C): the 256-entry
// byte table, then x^(8m) mod g for m = 0 .. kMaxCrcBytes, then x^m mod g for
// m = 0 .. kMaxCrcBits - 1 (the weight of a bit with m bits after it).
constexpr int kMaxCrcBytes = 8448 / 8;
constexpr int kMaxCrcBits = 8448;
constexpr int kCrcBitOff = 256 + kMaxCrcBytes + 1;
constexpr int kCrcTabWords = kCrcBitOff + kMaxCrcBits;
constexpr int kMaxThreads = 384;         // one lane per lifted check
constexpr int kMaxSoft = 68 * kMaxZc;    // N + 2Zc for BG1 Zc=384

struct CbDesc {
    uint32_t llr_off;    // int16 element offset of the E channel LLRs
    uint32_t hard_off;   // byte offset of the B/8 hard-decision bytes
    uint32_t soft_off;   // int16 element offset of the soft buffer or kNone
    uint32_t lq_off;     // element offset of the posterior output or kNone
    uint32_t etab_off;   // uint32 element offset of the (BG, Zc) edge table
    uint32_t work_off;   // float element offset of the BP message store
    uint16_t bgn, Zc, B, K, N, E, k0, Qm;
    uint8_t crctype, harq_on, pad0, pad1;
};
static_assert(sizeof(CbDesc) == 44, "CbDesc layout");

// Decoder row (c_lib/nr_ldpc.h:82-90) pre-converted for the kernels.
struct DecParams {
    int32_t max_it;
    int32_t alpha, beta;        // algo 1: uint16 fields wrapped to int16
    int32_t max_llr, min_llr;   // int16 clip bounds
    int32_t filler;             // filler-bit LLR (int16)
    float falpha, fbeta;        // algo 4: alpha/32768, beta/128
    float fmax, fmin;           // algo 4: max_llr/128, min_llr/128
    // int16 algos: 1 when some row's scaled minimum can wrap to int16 -32768
    // (or alpha < 0, or min_llr > max_llr), which needs the exact row form
    int32_t general;
    // int16 fast rows: 1 when every (alpha, beta) pair has beta = 0 and
    // max_llr = 32767, so a scaled minimum is (alpha * m + 2^14) >> 15
    int32_t simple;
    // algos 2/3: per (iteration, row) scale words of base graph 1 / 2,
    // alpha (int16) | beta (int16) << 16, iteration t at t * rows; iterations
    // at or past scale_iters use the last one
    int32_t scale_iters;
    const uint32_t *scale[2];
};

// Packed tiles: a workgroup of kTileThreads lanes (kTileWaves waves, two per
// SIMD) holds groups of codeblocks of one base graph.  A group is `count`
// codeblocks of one lifting size on whole waves starting at lane lane_base,
// codeblock i on lanes lane_base + [i*Zc, (i+1)*Zc), with its beliefs in its
// own LDS region at lds_base + i*lds_stride; order[first .. first+count) are
// their indices in the codeblock descriptor array, and slot_base + i their
// CRC slots.  A workgroup's descriptor is kTileWaves TileDesc, one per wave:
// each wave reads the descriptor of the group it belongs to (waves outside
// any group: count 0).  All kTileWaves entries carry the workgroup's bgn.
constexpr int kTileWaves = 8;
constexpr int kTileThreads = kTileWaves * 64;
constexpr int kMaxTileCodeblocks = kTileThreads / 2;
// CRC bit weights of a tile workgroup: one word per lane and systematic
// column (22 for base graph 1).
constexpr int kTileWtsBytes = 22 * kTileThreads * 4;
// CRC bit-weight columns of the one-codeblock-per-workgroup min-sum kernels
// (the larger base graph's systematic columns), per lane.
constexpr int kGenericWtsCols = 22;
struct TileDesc {
    uint32_t first;
    uint32_t etab_off;
    uint32_t lds_base;                     // LQ entries before the group's first codeblock
    uint16_t count, bgn, Zc, lds_stride;   // lds_stride: LQ entries per codeblock
    uint16_t lane_base, slot_base;
};
static_assert(sizeof(TileDesc) == 24, "TileDesc layout");

// Device queue of a persistent tile launch: ctr[0] is the next tile index,
// ctr[1] counts workgroups that have left; both 0 between launches.  A null
// ctr launches one workgroup per tile instead.
struct TileQueue {
    uint32_t *ctr;
    int32_t ntiles;
    uint32_t wts_lds;   // byte offset of the CRC bit weights in dynamic LDS
};

// Per-codeblock result flags (status) follow c_ldpc_5g_decode: 0 CRC pass,
// 1 not converged.
struct DecOut {
    const CbDesc *desc;
    const int16_t *llr;
    int16_t *soft;
    uint8_t *hard;
    int32_t *status;
    int32_t *iters;
    void *lq;
    float *work;
    const uint32_t *etab;
    const uint32_t *crc;   // [4][kCrcTabWords] CRC tables
};

// Device input synthesis (ldpc_synth.hip): n codeblocks of one config.
struct SynthArgs {
    const float *noise_db;      // n values: noise power (dB) per codeblock
    const uint8_t *payload;     // n*A bits (one per byte) or null: generated
    uint8_t *bits_out;          // n*B payload + CRC bits, or null
    uint8_t *cw_out;            // n*(N+2Zc) codeword bits, or null
    uint8_t *rm_out;            // n*E rate-matched bits, or null
    int16_t *llr_out;           // n*E quantized LLRs, or null
    float *llr_f32;             // n*E float LLRs, or null (needed for beliefs)
    float *beliefs;             // n*(N+2Zc) float beliefs, or null
    const uint32_t *etab;       // edge table of (BG, Zc)
    // Given channel noise (both or neither): noise_var n values, the channel's
    // noise variance per codeblock; noise n*(E/Qm)*2 values, the noise added
    // to each symbol's real and imaginary part.  Null: noise_db and the
    // counter-based generator.
    const float *noise_var;
    const float *noise;
    uint64_t seed;
    int n, A, B, K, N, E, Zc, Qm, k0, crctype;
};

}  // namespace nrldpc
