/*
 * nr_ldpc_mi355.h — C ABI of libnrldpc_mi355.so, the MI355X (gfx950) 5G NR
 * LDPC decode path.
 *
 * Part 1 is a drop-in for the reference library's ctypes surface
 * (hahaliu2001/deeplearning_5g_ldpc_decoder, c_lib/nr_ldpc.h): the same
 * struct layouts, symbol names, argument meanings and return codes, so
 * scripts/ldpc_c_sim_common.py:load_c_library and
 * deep_learning/ldpc_table_config.py bind it unchanged.  Decoding runs on the
 * GPU; parameter derivation, CRC, encoding, rate matching and the test
 * channel are host code (they generate inputs; they are not the hot path).
 *
 * Part 2 is new: a batched, allocation-free-per-call decode over device
 * buffers (plans), device-resident HARQ soft buffers, and queries.
 *
 * Plain C types only; the library owns its HIP stream and device tables.
 */
#ifndef NR_LDPC_MI355_H
#define NR_LDPC_MI355_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------ */
/* Part 1: reference-compatible surface                                */
/* ------------------------------------------------------------------ */

/* c_lib/nr_ldpc.h:8-21 */
#define BG1_NUM_ROWS 46
#define BG1_NUM_COLS 68
#define BG1_MAX_NON_NULL_ELEMENTS 316
#define BG2_NUM_ROWS 42
#define BG2_NUM_COLS 52
#define BG2_MAX_NON_NULL_ELEMENTS 197
#define MAX_DEGREE 19
#define MAX_ZC 384

/* c_lib/nr_ldpc.h:42-50 — 7 x uint16, 14 bytes, no padding. */
typedef struct {
    uint16_t bgn;     /* base graph 1 or 2 */
    uint16_t A;       /* payload bits before CRC (multiple of 8) */
    uint16_t crctype; /* 0 CRC16, 1 CRC24A, 2 CRC24B, 3 CRC24C */
    uint16_t E;       /* rate-matched bits */
    uint16_t rv;      /* redundancy version 0..3 */
    uint16_t Qm;      /* 1, 2, 4, 6, 8 */
    uint16_t HARQ_on; /* 1 = combine into the soft buffer, 0 = fresh */
} ldpc_in_config_t;

/* c_lib/nr_ldpc.h:54-79 — derived parameters plus host table pointers. */
typedef struct {
    uint16_t bgn, A, crctype, E, rv, Qm, HARQ_on;
    uint16_t B, Zc, iLS, K, F, N, k0;
    uint16_t *Vij_array;
    uint16_t *HBG_array;
    uint16_t *HBG_offset_array;
} ldpc_config_t;

/* c_lib/nr_ldpc.h:82-90 — the 7-field decoder row
 * [algo, max_iterations, alpha, beta, filling_bit_value, max_llr, min_llr].
 * algo: 0 FP32 layered BP, 1 int16 min-sum, 2 int16 min-sum with trained
 * per-(iteration, row) alpha/beta tables (alpha, beta ignored), 3 int16
 * min-sum with per-degree alpha (alpha ignored, beta used), 4 FP32
 * min-sum. */
typedef struct {
    uint16_t algo;
    uint16_t max_iterations;
    uint16_t alpha;   /* Q15 scale (algo 1 reads it as int16, algo 4 /32768) */
    uint16_t beta;    /* offset in int16 LLR units (algo 4: /128) */
    int16_t filling_bit_value;
    int16_t max_llr;
    int16_t min_llr;
} ldpc_decoder_cfg_t;

/* Compressed base graphs, reference layout (c_lib/nr_ldpc.h:93-99); the
 * pointers gen_ldpc_parameters returns point into these host arrays. */
extern uint16_t BG1_HBG_offset_array[BG1_NUM_ROWS + 1];
extern uint16_t BG1_HBG_array[BG1_MAX_NON_NULL_ELEMENTS];
extern uint16_t BG1_HBG_Vij_array[8][BG1_MAX_NON_NULL_ELEMENTS];
extern uint16_t BG2_HBG_offset_array[BG2_NUM_ROWS + 1];
extern uint16_t BG2_HBG_array[BG2_MAX_NON_NULL_ELEMENTS];
extern uint16_t BG2_HBG_Vij_array[8][BG2_MAX_NON_NULL_ELEMENTS];

/* Host mirror of the process soft buffer (c_lib/nr_ldpc.h:105,
 * c_lib/ldpc_internal.c:11).  After each c_ldpc_5g_decode it holds the
 * rate-recovered buffer of that call; a call with HARQ_on = 1 combines into
 * whatever it holds (the caller may edit it between calls). */
extern int16_t HARQ_memory[BG1_NUM_COLS * MAX_ZC];

/* c_lib/nr_ldpc.h:102, c_lib/nr_crc.c:170 — CRC remainder over len bytes;
 * 0 for an unknown crctype. */
uint32_t calculate_crc(const uint8_t *data, int len, int crctype);

/* c_lib/nr_ldpc.h:108, c_lib/ldpc_internal.c:66 — 0 or -1 (message on
 * stderr). */
int gen_ldpc_parameters(ldpc_config_t *params, ldpc_in_config_t *inconfig);

/* c_lib/nr_ldpc.h:124, c_lib/tx_rx_process.c:42 — x / 128. */
float dequantize_int16_to_llr(int16_t quantized_llr);

/* c_lib/nr_ldpc.h:127, c_lib/c_ldpc_5g_decoder.c:34 — rate recovery + HARQ
 * combine + decode of one codeblock on the GPU.  LLr: E int16 (x128);
 * hardbits: B/8 bytes MSB-first (untouched when max_iterations = 0).
 * Returns 0 CRC pass, 1 not converged, -2 invalid config, -3 invalid algo
 * (rate recovery still updates HARQ_memory, as in the reference), -1 GPU or
 * runtime error.  Algo 2 with max_iterations beyond the trained table of
 * the codeblock's base graph (32 iterations by default) returns -2 before
 * rate recovery: the reference reads past its table there
 * (c_lib/neural_fixedpoint_ldpc_decoder.c:248).  Serialized internally
 * (safe from several threads). */
int32_t c_ldpc_5g_decode(int16_t *LLr, uint8_t *hardbits, ldpc_in_config_t *inconfig,
                         ldpc_decoder_cfg_t *decoder_config);

/* c_lib/nr_ldpc.h:129, c_lib/c_ldpc_5g_encoder.c:66 — CRC attach, encode,
 * rate match; rm_bytes receives ceil(E/8) bytes MSB-first.  0 or -2.
 * (Host code; the encoder buffer is sized 68*Zc, so Zc >= 373 is safe.) */
int32_t c_ldpc_5g_encode(uint8_t *infobytes, uint8_t *rm_bytes, ldpc_in_config_t *inconfig);

/* CUDA/cuda_ldpc_5g_decoder.cu:31 — the CUDA library's twin of
 * c_ldpc_5g_decode, bound by scripts/sim_cuda_ldpc_encoder_decoder.py:66
 * (decode_symbol="cuda_ldpc_5g_decode").  Same arguments, HARQ semantics and
 * return codes, with the CUDA library's algo IDs: 1 is the int16 min-sum and
 * 4 is FP32 layered *BP* (CUDA/cuda_ldpc_5g_decoder.cu:49-51 routes 4 to
 * its float BP kernel, where c_lib's 4 is FP32 min-sum); any other algo,
 * 0 included, returns -3 after rate recovery. */
int32_t cuda_ldpc_5g_decode(int16_t *LLr, uint8_t *hardbits, ldpc_in_config_t *inconfig,
                            ldpc_decoder_cfg_t *decoder_config);

/* CUDA/cuda_ldpc_5g_encoder.cu:47 — the CUDA library's encoder symbol
 * (encode_symbol="cuda_ldpc_5g_encode"); the same bytes as c_ldpc_5g_encode. */
int32_t cuda_ldpc_5g_encode(uint8_t *infobytes, uint8_t *rm_bytes, ldpc_in_config_t *inconfig);

/* c_lib/nr_ldpc.h:123, c_lib/tx_rx_process.c:286 — test channel: Gray QAM,
 * complex AWGN of power 10^(noise_power_db/10), max-log LLRs x128 rounded
 * and clipped to +-10000.  The noise comes from this library's own
 * generator (seed with ldpc_mi355_channel_seed), not glibc rand().
 * 0 or -1 for invalid arguments. */
int tx_rx_process(const uint8_t *inbytes, size_t num_bits, uint16_t Qm, float noise_power_db,
                  int16_t *llr_out);

/* ------------------------------------------------------------------ */
/* Part 2: MI355X extensions (no counterpart in the reference)         */
/* ------------------------------------------------------------------ */

/* Algo-2 trained tables of base graph bgn: alpha and beta are int16
 * [iterations][rows] (rows 46 for BG1, 42 for BG2), flattened iteration
 * first, Q15 alpha and beta in int16 LLR units -- the layout of the
 * reference's compiled-in BG<n>_trained_alpha_list / _beta_list
 * (c_lib/BG1_trained_ldpc_weights.h:7-17, indexed at
 * c_lib/neural_fixedpoint_ldpc_decoder.c:248-250), which it can only change
 * by recompiling.  iterations 1..1024; alpha NULL restores the default
 * tables (32 iterations, alpha by the row-degree rule, beta 0).  The tables
 * are process-wide: c_ldpc_5g_decode reads the tables current at the call,
 * plans and batch contexts keep the tables current when they were created.
 * 0, or -4 for invalid arguments. */
int32_t ldpc_mi355_set_trained_tables(int32_t bgn, const int16_t *alpha, const int16_t *beta,
                                      int32_t iterations);

/* Copies up to capacity_iterations iterations of base graph bgn's current
 * algo-2 tables into alpha / beta (either may be NULL) and returns the
 * tables' iteration count, or -4 for an invalid bgn. */
int32_t ldpc_mi355_get_trained_tables(int32_t bgn, int16_t *alpha, int16_t *beta,
                                      int32_t capacity_iterations);

/* Algo-3 degree alpha: MAX_DEGREE Q15 values, entry d-1 for check degree d
 * -- the table layered_degree_minsum_ldpc_decode_with_alpha takes
 * (c_lib/nr_ldpc.h:113, c_lib/degree_fixedpoint_ldpc_decoder.c:15-17,
 * 65-80); NULL restores the defaults (26623 for degree <= 4, 24575 for 5..8,
 * 22527 above).  Process-wide, read like the trained tables.  Returns 0. */
int32_t ldpc_mi355_set_degree_alpha(const int16_t *degree_alpha_values);

/* Copies the current MAX_DEGREE degree alpha values; 0, or -4 for NULL. */
int32_t ldpc_mi355_get_degree_alpha(int16_t *degree_alpha_values);

/* c_ldpc_5g_decode plus what the reference keeps internal: the number of
 * iterations executed and the N+2Zc posteriors at exit (int16 for algos
 * 1-3, float for algos 0/4); either output may be NULL. */
int32_t ldpc_mi355_decode_ex(int16_t *LLr, uint8_t *hardbits, ldpc_in_config_t *inconfig,
                             ldpc_decoder_cfg_t *decoder_config, int32_t *iterations,
                             void *posteriors);

/* Seeds the tx_rx_process noise generator (default seed 0). */
void ldpc_mi355_channel_seed(uint64_t seed);

/* Input synthesis for simulations and benchmarks (host, worker threads):
 * n codeblocks of one config with random payloads, CRC, encoding, rate
 * matching and the test channel at noise_power_db.  Each codeblock has its
 * own noise stream derived from (seed, index), so results do not depend on
 * the thread count.  llr_out: n*E int16; payload_out (may be NULL): n*A
 * payload bits, one per byte.  0, or -2 for an invalid config. */
int32_t ldpc_mi355_synth_batch(const ldpc_in_config_t *cfg, int32_t n, float noise_power_db,
                               uint64_t seed, int32_t threads, int16_t *llr_out,
                               uint8_t *payload_out);

/* Float inputs of the trainable decoder, in its N+2Zc belief layout
 * (deep_learning/ldpc_training_data.py:24-82): random payload and CRC,
 * encoding, rate matching, unquantized channel LLRs at noise_power_db[i]
 * for codeblock i, then the float rate recovery of
 * py5gphy/ldpc/nr_ldpc_raterecover.py:15-68 (repeats summed, filler
 * positions 10 * max|LLR|, 2Zc leading zeros).  llr_out: n*(N+2Zc) floats;
 * bits_out (may be NULL): n*B target bits (payload + CRC), one per byte.
 * Noise streams are derived from (seed, index).  0, or -2. */
int32_t ldpc_mi355_synth_float_batch(const ldpc_in_config_t *cfg, int32_t n,
                                     const float *noise_power_db, uint64_t seed, int32_t threads,
                                     float *llr_out, uint8_t *bits_out);

/* The same inputs generated on the device (no host round trip): n codeblocks
 * of one config, one workgroup each -- payload (the caller's payload_bits,
 * n*A bits one per byte, or a counter-based generator keyed by seed), CRC,
 * LDPC encoding, rate matching, Gray-QAM AWGN channel at noise_power_db[i]
 * (device array of n floats), max-log LLRs.  Every output is optional
 * (device pointers, may be NULL): bits_out n*B payload+CRC bits,
 * codeword_out n*(N+2Zc) bits, rm_bits_out n*E rate-matched bits, llr_out
 * n*E int16 LLRs (x128, rounded, clipped to +-10000, the c_ldpc_5g_decode
 * input), llr_f32_out n*E float LLRs, beliefs_out n*(N+2Zc) float beliefs in
 * the trainable decoder's layout (needs llr_f32_out).  Asynchronous on
 * `stream`.  0, -2 invalid config, -4 invalid argument, -1 GPU error. */
int32_t ldpc_mi355_synth_device(const ldpc_in_config_t *cfg, int32_t n,
                                const float *noise_power_db, uint64_t seed,
                                const uint8_t *payload_bits, uint8_t *bits_out,
                                uint8_t *codeword_out, uint8_t *rm_bits_out, int16_t *llr_out,
                                float *llr_f32_out, float *beliefs_out, void *stream);

/* ldpc_mi355_synth_device with the channel noise given instead of drawn
 * (replaces the noise of the reference's tx_rx_process,
 * c_lib/tx_rx_process.c:286-324, whose samples come from the C library's
 * rand()): noise_var (device, n floats) is each codeblock's noise variance
 * (the reference's powf(10, noise_power_db / 10)), noise (device,
 * n*(E/Qm)*2 floats) the values added to each symbol's real and imaginary
 * part (the reference's noise_sigma * randn() in call order).  Modulation,
 * max-log LLRs and quantization are the reference's operations, so the same
 * noise gives the reference's LLRs bit for bit.  0, -2, -4 (a noise pointer
 * NULL), -1. */
int32_t ldpc_mi355_synth_device_noise(const ldpc_in_config_t *cfg, int32_t n,
                                      const float *noise_var, const float *noise, uint64_t seed,
                                      const uint8_t *payload_bits, uint8_t *bits_out,
                                      uint8_t *codeword_out, uint8_t *rm_bits_out,
                                      int16_t *llr_out, float *llr_f32_out, float *beliefs_out,
                                      void *stream);

/* One codeblock of a batch. */
typedef struct {
    ldpc_in_config_t cfg;  /* HARQ_on: combine into the soft buffer slot */
    int16_t reserved;
    uint32_t llr_offset;   /* element offset of its E LLRs in the LLR buffer */
    int32_t soft_slot;     /* soft-buffer slot index, or -1 for none */
} ldpc_mi355_cb_t;

typedef struct ldpc_mi355_plan ldpc_mi355_plan_t;

/* Validates n codeblock descriptors against one decoder row and uploads
 * them to the device (for algos 2/3 also the scale tables current at this
 * call).  flags: bit 0 = produce posteriors; bit 1 = pack codeblocks of one
 * (BG, Zc) into shared workgroups (int16 algos 1-3); bit 2 = never pack.
 * Without bits 1/2, int16 plans pack whenever some lifting size leaves room
 * for two or more codeblocks per workgroup.  bit 3 = never use a kernel
 * specialized for the batch's single (base graph, lifting size) (int16
 * algos; used when all codeblocks share one, e.g. BG1 Zc=384).  bit 4 =
 * persistent packed tiles: a grid of at most one workgroup per resident slot
 * pulls the tiles from a device queue (the plan holds 64 queues, used in
 * turn by successive launches, so at most 64 launches of one plan may be in
 * flight at once).  bit 5 = lane-per-edge schedule (algo 1 with fast rows,
 * used when every codeblock has Zc <= 3 for base graph 1, <= 6 for base
 * graph 2; ignored otherwise): one wave per codeblock, each layer's edges
 * one per lane.  Returns 0, -2 if any codeblock
 * config is invalid or, for algo 2, needs more iterations than its base
 * graph's trained table holds (index in *bad_index when non-NULL), -3 for
 * an unsupported algo, -1 on a GPU error. */
int32_t ldpc_mi355_plan_create(ldpc_mi355_plan_t **plan, const ldpc_mi355_cb_t *cbs, int32_t n,
                               const ldpc_decoder_cfg_t *decoder_config, uint32_t flags,
                               int32_t *bad_index);

/* Output layout of a plan: total hard bytes, total posterior elements,
 * element size of a posterior (2 or 4), and per-codeblock offsets (arrays
 * of n, may be NULL). */
int32_t ldpc_mi355_plan_layout(const ldpc_mi355_plan_t *plan, uint64_t *hard_bytes,
                               uint64_t *lq_elems, int32_t *lq_elem_size,
                               uint32_t *hard_offsets, uint32_t *lq_offsets);

/* Launches the batch on `stream` (a hipStream_t; NULL is HIP's default null
 * stream).  All pointers are device pointers: llr (int16), soft_pool (int16,
 * slots of 68*384 elements; may be NULL if no codeblock uses a slot),
 * hard (hard_bytes), status and iterations (n int32 each), lq (optional).
 * Asynchronous and allocation-free (capturable in a hipGraph).  Returns 0;
 * -4 (nothing launched) when a codeblock has a soft slot and soft_pool is
 * NULL; -1 when posteriors were asked for and lq is NULL, or on a GPU error.
 * soft_pool must hold (largest slot + 1) * 68*384 elements. */
int32_t ldpc_mi355_plan_decode(const ldpc_mi355_plan_t *plan, const int16_t *llr,
                               int16_t *soft_pool, uint8_t *hard, int32_t *status,
                               int32_t *iterations, void *lq, void *stream);

void ldpc_mi355_plan_destroy(ldpc_mi355_plan_t *plan);

/* Number of packed workgroups of a plan (0 = one codeblock per workgroup),
 * -1 for NULL. */
int32_t ldpc_mi355_plan_tiles(const ldpc_mi355_plan_t *plan);

/* Codeblocks per packed workgroup for lifting size zc of base graph bgn,
 * used by plans created afterwards (k = 0 restores the default, as many as
 * fit in 384 lanes; larger k is capped there).  Set by the autotuner.
 * 0, or -1 for an invalid (bgn, zc, k). */
int32_t ldpc_mi355_set_tile_codeblocks(int32_t bgn, int32_t zc, int32_t k);

/* Reusable batch contexts, for continuous batching: every decode takes a new
 * set of up to `capacity` codeblocks (configs may differ from call to call),
 * builds descriptors on the host into the context's pinned staging buffer,
 * and issues one asynchronous copy plus the decode launch on `stream` -- no
 * allocation, no synchronization.  A context runs one decode at a time: a
 * decode on a context whose previous decode is still in flight first waits
 * for it; ldpc_mi355_batch_query polls (1 done, 0 running), batch_wait
 * blocks.  flags as for plans (bit 0 posteriors, bit 1 pack, bit 2 never
 * pack).  Algos 1-4; algos 2/3 keep the scale tables current at creation.
 * Returns 0, -2 invalid codeblock config or algo-2 iterations beyond the
 * table (*bad_index), -3 unsupported algo, -4 invalid argument (including a
 * codeblock with a soft slot while soft_pool is NULL), -1 GPU error. */
typedef struct ldpc_mi355_batch ldpc_mi355_batch_t;
int32_t ldpc_mi355_batch_create(ldpc_mi355_batch_t **batch, int32_t capacity,
                                const ldpc_decoder_cfg_t *decoder_config, uint32_t flags);
int32_t ldpc_mi355_batch_decode(ldpc_mi355_batch_t *batch, const ldpc_mi355_cb_t *cbs, int32_t n,
                                const int16_t *llr, int16_t *soft_pool, uint8_t *hard,
                                int32_t *status, int32_t *iterations, void *lq, void *stream,
                                int32_t *bad_index);
/* Output layout of the context's last decode (like plan_layout); returns
 * its number of codeblocks. */
int32_t ldpc_mi355_batch_layout(const ldpc_mi355_batch_t *batch, uint64_t *hard_bytes,
                                uint64_t *lq_elems, uint32_t *hard_offsets, uint32_t *lq_offsets);
int32_t ldpc_mi355_batch_query(ldpc_mi355_batch_t *batch);
int32_t ldpc_mi355_batch_wait(ldpc_mi355_batch_t *batch);
void ldpc_mi355_batch_destroy(ldpc_mi355_batch_t *batch);

/* n asynchronous copies on stream, span i: bytes[i] from src[i] to dst[i];
 * kind 1 = host to device, 2 = device to host (pinned host memory for the
 * copies to be asynchronous).  No counterpart in the reference: it serves
 * callers that move many small spans per launch, such as a served loop
 * copying each request's LLRs into its batch, in one call.  Returns 0, or
 * -1 on invalid arguments or a HIP error. */
int32_t ldpc_mi355_copy_spans(void *const *dst, const void *const *src, const uint64_t *bytes,
                              int32_t n, int32_t kind, void *stream);

/* ------------------------------------------------------------------ */
/* Native served loop.  No counterpart in the reference: the serving   */
/* pattern (continuous batching, admission by free device memory,      */
/* device-resident HARQ soft buffers) run in C++ on the calling thread.*/
/* ------------------------------------------------------------------ */
typedef struct {
    int32_t n_contexts;            /* batches in flight at most (one HIP stream each) */
    int32_t max_batch_codeblocks;  /* codeblocks per batch at most */
    int32_t n_ues, n_procs;        /* HARQ processes per UE */
    int32_t max_cbs;               /* codeblocks per request at most: soft slot (ue, proc, cb)
                                    * is ((ue * n_procs) + proc) * max_cbs + cb */
    int32_t llr_source;            /* where llr_host is and how batches read it: one of
                                    * LDPC_MI355_LLR_* below */
    double deadline_s;             /* recorded with each request (not enforced) */
    double harq_rtt_s;             /* a failed transmission's next rv arrives this much later */
    double max_wall_s;             /* the loop stops after this long */
    uint64_t llr_pool_elems;       /* admission budget: int16 elements of llr_pool (host LLRs) */
} ldpc_mi355_serve_cfg_t;

/* llr_source values.  HOST_COPY: llr_host is pinned host memory; each
 * batch copies its requests' spans into a region of llr_pool, which
 * admission budgets.  DEVICE: llr_host is device memory (as a demapper on
 * the GPU leaves it); batches decode it in place.  HOST_MAPPED: llr_host is
 * pinned host memory (hipHostMalloc / torch pin_memory); batches decode it
 * in place, the kernel reading it over PCIe.  In place, llr_pool is unused
 * and every LLR offset must be below 2^32 elements. */
#define LDPC_MI355_LLR_HOST_COPY 0
#define LDPC_MI355_LLR_DEVICE 1
#define LDPC_MI355_LLR_HOST_MAPPED 2

/* A request template: n_cbs codeblocks with configs cb_cfg[cb_first ..]
 * (their rv and HARQ_on are set per transmission); the LLRs of redundancy
 * version index r (rv 0, 2, 3, 1) are llr_elems int16 values at element
 * offset llr_host[r] of the host LLR array, codeblock after codeblock. */
typedef struct {
    int32_t n_cbs, cb_first;
    uint64_t llr_elems;
    uint64_t llr_host[4];
} ldpc_mi355_serve_template_t;

/* One transmission, as served.  Codeblock results: status and iterations at
 * [cb_first, cb_first + n_cbs) of the result arrays; hard bytes at byte
 * hard_first of hard_out (-1 when not recorded).  Times in seconds from the
 * start of the run: arrival_s the arrival (new data) or the due time of a
 * retransmission (its failed transmission's completion + harq_rtt_s);
 * queued_s when it joined the decode queue: a retransmission at arrival_s,
 * new data once its UE had a free HARQ process (the later of arrival_s and
 * the completion that freed the process); admitted_s when its batch was
 * launched; done_s when the loop saw the batch complete. */
typedef struct {
    int32_t template_id, ue, proc, rv_index;
    int32_t n_cbs, n_pass;          /* codeblocks, codeblocks that passed CRC */
    int64_t cb_first, hard_first;
    int32_t batch, final_tx;        /* batch index; 1 = last transmission of its chain */
    double arrival_s, admitted_s, done_s;
    double queued_s;
} ldpc_mi355_serve_record_t;

typedef struct {
    int32_t n_records, batches, admission_waits, max_blocked, max_inflight, reserved;
    int64_t n_cb_results, hard_bytes;
    uint64_t peak_pool_elems;
    double wall_s, host_launch_s, host_collect_s;
} ldpc_mi355_serve_stats_t;

/* Runs the served loop until every arrival's HARQ chain is done (or
 * max_wall_s): new-data arrivals (arrival_s relative to the start, template,
 * UE) wait for a free HARQ process of their UE, then queue; whenever one of
 * n_contexts batch contexts is free, the queue head is admitted as one batch
 * of whole requests within max_batch_codeblocks and a first-fit region of
 * llr_pool (device, llr_pool_elems int16) into which its request LLRs
 * (llr_host) are copied -- with llr_source DEVICE or HOST_MAPPED each
 * batch reads its requests' LLRs where they are, with no region and no
 * copy -- decoded with decoder_config
 * and its results written by the decode into the context's pinned host
 * buffer, asynchronously on its context's stream; a
 * transmission with a failed codeblock is retransmitted with the next rv of
 * 0, 2, 3, 1 after harq_rtt_s and combined into the same soft slots of
 * soft_pool (device, 68*384 int16 per slot).  Fills records (one per
 * transmission, max_records), per-codeblock cb_status / cb_iterations
 * (max_cb_results) and, when hard_out is not NULL, the hard bytes
 * (hard_capacity bytes; later ones are not recorded), and stats.
 * Returns 0; -4 invalid argument or a result array too small; -5 a request
 * larger than the LLR pool; -1 GPU error.  Allocates its contexts, streams
 * and result buffers once at the start, nothing inside the loop. */
int32_t ldpc_mi355_serve_run(const ldpc_mi355_serve_cfg_t *cfg,
                             const ldpc_decoder_cfg_t *decoder_config,
                             const ldpc_mi355_serve_template_t *templates, int32_t n_templates,
                             const ldpc_in_config_t *cb_cfg, const int16_t *llr_host,
                             int16_t *llr_pool, int16_t *soft_pool, const double *arrival_s,
                             const int32_t *arrival_template, const int32_t *arrival_ue,
                             int32_t n_arrivals, ldpc_mi355_serve_record_t *records,
                             int32_t max_records, int32_t *cb_status, int32_t *cb_iterations,
                             int64_t max_cb_results, uint8_t *hard_out, int64_t hard_capacity,
                             ldpc_mi355_serve_stats_t *stats);

/* 1 when the library runs decoder algo `algo` (0-4) on the GPU, else 0. */
int32_t ldpc_mi355_algo_available(int32_t algo);

/* Allocation-free FP32 BP decode.  No counterpart in the reference: it is
 * for callers that cannot allocate device memory -- code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc.  Decodes n
 * codeblocks of one config *cfg with decoder row *decoder_config (algo 0
 * only) like ldpc_mi355_plan_create + ldpc_mi355_plan_decode on a plan whose
 * codeblock i has LLR offset i*E and soft slot i, and gives the same
 * results, but calls no allocator: the check-to-variable messages live in a
 * 32 MiB device array of the library (one per GPU, held until the process
 * ends), so n * nnz(BG) * Zc * 4 bytes must fit in it (nnz 316 for BG1, 197
 * for BG2: n <= 69 at BG1 Zc=384, n <= 409 at BG2 Zc=104); larger n returns
 * -4 with a message.  Its first call on a GPU sets up that GPU's tables
 * (one synchronous copy), later calls only launch.  Device pointers: llr
 * (n*E int16), soft_pool (slots of 68*384 int16, slot i for codeblock i; may
 * be NULL unless HARQ_on), hard (n*B/8 bytes), status and iterations (n
 * int32 each), lq (n*(N+2Zc) float posteriors, or NULL).  Asynchronous on
 * `stream` (a hipStream_t; NULL = HIP's null stream).  Every call on one GPU
 * shares its message array: calls that overlap on a GPU (other streams,
 * other host threads) race on it and must be ordered by the caller.
 * Returns 0, -2 invalid config, -3 algo other than 0, -4 invalid argument or
 * n too large, -1 GPU error. */
int32_t ldpc_mi355_bp_decode_noalloc(const ldpc_in_config_t *cfg,
                                     const ldpc_decoder_cfg_t *decoder_config, int32_t n,
                                     const int16_t *llr, int16_t *soft_pool, uint8_t *hard,
                                     int32_t *status, int32_t *iterations, float *lq,
                                     void *stream);

/* Blocks until the library stream is idle.  0 or -1. */
int32_t ldpc_mi355_sync(void);

/* Elements per soft-buffer slot (68*384). */
int32_t ldpc_mi355_soft_slot_elems(void);

/* Version string of this library. */
const char *ldpc_mi355_version(void);

/* ------------------------------------------------------------------ */
/* Unrolled trainable FP32 min-sum.  Replaces the forward of           */
/* CTableNeuralLayeredMinSumDecoder (deep_learning/                    */
/* deeplearning_ldpc_model.py:315-410) and the gradient torch autograd */
/* takes through it, for one LDPC config and a batch of codeblocks.    */
/* alpha/beta are the per-(iteration, base-graph row) values the row   */
/* update uses (the reference's alpha()/beta() mapped through          */
/* _alpha_for_row, :285-313); callers map their parameter sharing mode */
/* onto this [iterations][rows] table.  All data pointers are device   */
/* pointers; calls are asynchronous on `stream` (NULL = default        */
/* stream) and allocate nothing.                                       */
/* ------------------------------------------------------------------ */
typedef struct ldpc_mi355_unrolled ldpc_mi355_unrolled_t;

/* flags bit: the check update treats its inputs as constants
 * (detach_check_update_inputs=True, deeplearning_ldpc_model.py:365). */
#define LDPC_MI355_UNROLLED_DETACH 1

/* Model for config `cfg` (gen_ldpc_parameters layout: n = N + 2Zc beliefs)
 * unrolled over `iterations`; clip_llr > 0 clamps beliefs to
 * [-clip_llr, clip_llr] like the reference's clip_llr (<= 0: no clamp).
 * 0, -2 invalid config, -4 invalid argument, -1 GPU error. */
int32_t ldpc_mi355_unrolled_create(ldpc_mi355_unrolled_t **handle, const ldpc_in_config_t *cfg,
                                   int32_t iterations, float clip_llr, int32_t flags);
void ldpc_mi355_unrolled_destroy(ldpc_mi355_unrolled_t *handle);

/* dims[8] = n, rows, edges, Zc, B, iterations, codeblocks per workgroup,
 * lanes per codeblock. */
int32_t ldpc_mi355_unrolled_dims(const ldpc_mi355_unrolled_t *handle, int32_t *dims);

/* Codeblocks per workgroup (0 = default, as many as fit 384 lanes and
 * 64 KB of LDS).  -4 if k does not fit. */
int32_t ldpc_mi355_unrolled_set_codeblocks_per_workgroup(ldpc_mi355_unrolled_t *handle,
                                                         int32_t k);

/* Device bytes of the forward record a backward needs, and of the backward
 * scratch, for `batch` codeblocks. */
uint64_t ldpc_mi355_unrolled_record_bytes(const ldpc_mi355_unrolled_t *handle, int32_t batch);
uint64_t ldpc_mi355_unrolled_scratch_bytes(const ldpc_mi355_unrolled_t *handle, int32_t batch);

/* llr: [batch][n] float beliefs in; lq_out: [batch][n] final beliefs;
 * iter_lq_out (may be NULL): [iterations][batch][n] beliefs after each
 * iteration; record (may be NULL for inference): record_bytes(batch),
 * written for ldpc_mi355_unrolled_backward.  alpha/beta: [iterations][rows]. */
int32_t ldpc_mi355_unrolled_forward(const ldpc_mi355_unrolled_t *handle, int32_t batch,
                                    const float *llr, const float *alpha, const float *beta,
                                    float *lq_out, float *iter_lq_out, void *record,
                                    void *stream);

/* Gradients of a loss L given dL/d lq_out (g_lq, [batch][n]) and, when the
 * forward produced per-iteration beliefs, dL/d iter_lq_out (g_iter_lq, may be
 * NULL).  Writes dL/d llr (g_llr, may be NULL) and dL/d alpha, dL/d beta
 * ([iterations][rows] each, summed over the batch in a fixed order).  record
 * is the forward's, for the same llr/alpha/beta; scratch: scratch_bytes(batch). */
int32_t ldpc_mi355_unrolled_backward(const ldpc_mi355_unrolled_t *handle, int32_t batch,
                                     const float *llr, const float *alpha, const float *beta,
                                     const void *record, const float *g_lq,
                                     const float *g_iter_lq, float *g_llr, float *g_alpha,
                                     float *g_beta, void *scratch, void *stream);

#ifdef __cplusplus
}
#endif
#endif /* NR_LDPC_MI355_H */
