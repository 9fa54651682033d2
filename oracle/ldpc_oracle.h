/*
 * ORACLE — test infrastructure only.
 *
 * A CPU restatement of the reference 5G NR LDPC decode path
 * (hahaliu2001/deeplearning_5g_ldpc_decoder, c_lib/) used as the checker for
 * the MI355X library.  Only tests/, the smoke check and the CPU-baseline leg
 * of the benchmark may load it; the product never links or calls it.
 *
 * Every function is reentrant (no file-static state), unlike the reference,
 * and exposes what the reference keeps internal: posteriors and the number of
 * iterations executed.  Each function cites the reference code it restates.
 *
 * Pinning (what checks each function, beyond restating it):
 *   orc_code_init       exhaustive equality with the from-source reference
 *                       gen_ldpc_parameters over every B (golden fixture)
 *   orc_crc             bitwise long division (a different method from the
 *                       reference byte table); reference CRC vectors;
 *                       linearity; appended-CRC codewords divide to zero
 *   orc_rate_recover    reference HARQ_memory fixtures incl. HARQ chains;
 *                       closed-form gather (tests) and mod-2^16 additivity
 *   orc_decode_i16      reference hardbits/status/iterations/LQ fixtures
 *   orc_decode_i16_table / _degree
 *                       reference algo-2 / algo-3 fixtures (default tables,
 *                       and algo 2 with exported non-uniform tables linked
 *                       into the from-source reference); reduction to
 *                       orc_decode_i16 for uniform tables (tests)
 *   orc_decode_f32_ms   reference FP32 fixtures (bit-exact) + FP64 twin
 *   orc_decode_f32_bp   reference FP32 fixtures (same glibc libm) + FP64 twin
 *   orc_encode / _rm    reference encoder fixtures; H*c = 0 syndrome check
 *   orc_decode_f64_*    hand-built check rows with closed-form answers
 *                       (zero-edge BP cases, min-sum ties) in tests
 */
#ifndef LDPC_ORACLE_H
#define LDPC_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
    /* caller fields (ldpc_in_config_t, c_lib/nr_ldpc.h:42-50) */
    int bgn, A, crctype, E, rv, Qm, harq_on;
    /* derived fields (ldpc_config_t, c_lib/nr_ldpc.h:64-72) */
    int B, Zc, iLS, K, F, N, k0;
    /* base graph slice */
    int rows, nnz;
    const uint16_t *row_off; /* rows + 1 */
    const uint16_t *col;     /* nnz */
    const uint16_t *vij;     /* nnz, for this iLS */
} orc_code_t;

#define ORC_MAX_ROWS 46     /* BG1 rows (c_lib/nr_ldpc.h) */
#define ORC_MAX_DEGREE 19   /* MAX_DEGREE (c_lib/nr_ldpc.h:18) */

typedef struct {
    /* decoder row, ldpc_decoder_cfg_t (c_lib/nr_ldpc.h:82-90) */
    uint16_t algo, max_iterations, alpha, beta;
    int16_t filling_bit_value, max_llr, min_llr;
} orc_dec_t;

/* c_lib/ldpc_internal.c:66-169.  0 on success, -1 on an invalid config. */
int orc_code_init(orc_code_t *c, int bgn, int A, int crctype, int E, int rv,
                  int Qm, int harq_on);
/* Plain export of the derived fields for ctypes callers:
 * out[0..6] = B, Zc, iLS, K, F, N, k0.  Returns orc_code_init's status. */
int orc_params(int bgn, int A, int crctype, int E, int rv, int Qm,
               int harq_on, int *out7);

/* c_lib/nr_crc.c:170-179 semantics: MSB-first, init 0, no reflection, no
 * xor-out; crctype 0..3 = CRC16, CRC24A, CRC24B, CRC24C; other types -> 0. */
uint32_t orc_crc(const uint8_t *bytes, int nbytes, int crctype);

/* c_lib/ldpc_rate_recovery.c:24-65: soft[0..N+2Zc) receives the recovered
 * buffer; when harq_on, soft holds the previous buffer on entry. */
void orc_rate_recover(const orc_code_t *c, int16_t filler, const int16_t *llr,
                      int16_t *soft);

/* Decoders.  soft = N+2Zc rate-recovered int16 values.  hard receives B/8
 * bytes MSB-first (left untouched when max_iterations == 0).  lq_out (may be
 * NULL) receives the N+2Zc posteriors at exit, iters the number of
 * iterations executed.  Return 0 = CRC pass, 1 = not converged. */
int orc_decode_i16(const orc_code_t *c, const orc_dec_t *d,
                   const int16_t *soft, uint8_t *hard, int16_t *lq_out,
                   int *iters);
/* Algo 2: alpha/beta [table_iterations][rows]; -2 when max_iterations
 * exceeds table_iterations (the reference reads past its table then). */
int orc_decode_i16_table(const orc_code_t *c, const orc_dec_t *d,
                         const int16_t *soft, uint8_t *hard, int16_t *lq_out,
                         int *iters, const int16_t *alpha,
                         const int16_t *beta, int table_iterations);
/* Algo 3: degree_alpha = ORC_MAX_DEGREE Q15 values (entry d-1 for degree d)
 * or NULL for the reference defaults; beta from the decoder row. */
int orc_decode_i16_degree(const orc_code_t *c, const orc_dec_t *d,
                          const int16_t *soft, uint8_t *hard, int16_t *lq_out,
                          int *iters, const int16_t *degree_alpha);
/* The reference's default alpha of a check degree and its default trained
 * tables ([iterations][rows], beta 0). */
int16_t orc_default_degree_alpha(int degree);
void orc_default_trained_tables(int bgn, int iterations, int16_t *alpha,
                                int16_t *beta);
int orc_decode_f32_ms(const orc_code_t *c, const orc_dec_t *d,
                      const int16_t *soft, uint8_t *hard, float *lq_out,
                      int *iters);
int orc_decode_f32_bp(const orc_code_t *c, const orc_dec_t *d,
                      const int16_t *soft, uint8_t *hard, float *lq_out,
                      int *iters);
int orc_decode_f64_ms(const orc_code_t *c, const orc_dec_t *d,
                      const int16_t *soft, uint8_t *hard, double *lq_out,
                      int *iters);
int orc_decode_f64_bp(const orc_code_t *c, const orc_dec_t *d,
                      const int16_t *soft, uint8_t *hard, double *lq_out,
                      int *iters);

/* Whole-path oracle of c_ldpc_5g_decode (c_lib/c_ldpc_5g_decoder.c:34-79):
 * params -> rate recovery into soft (N+2Zc, caller-owned, the HARQ buffer)
 * -> decoder by algo.  Returns 0/1, -2 invalid config, -3 invalid algo. */
int orc_decode_path(const int16_t *llr, uint8_t *hard, int bgn, int A,
                    int crctype, int E, int rv, int Qm, int harq_on,
                    const orc_dec_t *d, int16_t *soft, int16_t *lq_i16,
                    float *lq_f32, int *iters);
/* The same with explicit algo-2 tables ([trained_iterations][rows]) and
 * algo-3 degree alpha (NULL = defaults); orc_decode_path uses the
 * reference's default tables (32 iterations). */
int orc_decode_path_tables(const int16_t *llr, uint8_t *hard, int bgn, int A,
                           int crctype, int E, int rv, int Qm, int harq_on,
                           const orc_dec_t *d, int16_t *soft, int16_t *lq_i16,
                           float *lq_f32, int *iters, const int16_t *trained_alpha,
                           const int16_t *trained_beta, int trained_iterations,
                           const int16_t *degree_alpha);

/* Check-node kernels on one lifted check, exposed for hand-built tests.
 * rel[deg] in, msg[deg] out (new check-to-variable messages). */
void orc_check_minsum_i16(int deg, const int16_t *rel, int16_t alpha,
                          int16_t beta, int16_t max_llr, int16_t *msg);
void orc_check_bp_f64(int deg, const double *rel, double *msg);

/* Input generation for tests (c_lib/c_ldpc_5g_encoder.c:66-88,
 * ldpc_encoder.c:221-242, ldpc_rate_match.c:27-106).  cw receives the
 * N+2Zc = (66|50+2)*Zc codeword bits (one per byte).  rm_bits receives E
 * bits (one per byte) after selection + interleaving. */
int orc_encode(const orc_code_t *c, const uint8_t *payload_bits, uint8_t *cw);
void orc_rate_match(const orc_code_t *c, const uint8_t *cw, uint8_t *rm_bits);
/* 1 when H * cw = 0 over GF(2) for every lifted check. */
int orc_syndrome_ok(const orc_code_t *c, const uint8_t *cw);

/* The reference's test channel with its own noise stream
 * (c_lib/tx_rx_process.c:286-324): srand(seed), then per symbol Gray QAM
 * (:68-99), noise_sigma * randn() added to the real, then the imaginary
 * part (randn = Box-Muller of two rand() draws, :59-65), piecewise max-log
 * LLRs (:101-284) and x128 quantization (:27-39), every float operation as
 * the reference orders it.  bits: nbits (one per byte; nbits % Qm == 0).
 * noise_out (NULL or 2 per symbol): the values added; noise_var_out (NULL or
 * one value); llr_out (NULL or nbits floats); q_out: nbits int16.  Uses the
 * C library's global rand() state (not reentrant).  0, or -1 for invalid
 * arguments.  Pinned by tests/golden/channel_vectors.npz: the
 * from-source reference's tx_rx_process after the same srand(seed). */
int orc_tx_rx(const uint8_t *bits, int nbits, int Qm, float noise_db, unsigned seed,
              float *noise_out, float *noise_var_out, float *llr_out, int16_t *q_out);

/* CPU baseline: decode n codeblocks of one config with algo 1 on P POSIX
 * threads (one codeblock per task).  llr: n*E int16, hard: n*B/8 bytes,
 * status/iters: n each.  Returns elapsed seconds (monotonic clock). */
double orc_decode_batch_i16(const int16_t *llr, uint8_t *hard, int32_t *status,
                            int32_t *iters, int n, int bgn, int A, int crctype,
                            int E, int rv, int Qm, const orc_dec_t *d,
                            int threads);

/* Unrolled float layered min-sum of the trainable decoder
 * (deep_learning/deeplearning_ldpc_model.py:315-410) on one codeblock
 * (unrolled_oracle.c).  llr: N+2Zc beliefs; alpha/beta: [iters][rows]
 * values as used by the row update; clip > 0 clamps beliefs to
 * [-clip, clip] (the reference's clip_llr), clip <= 0 disables it.
 * lq_out: N+2Zc final beliefs; iter_out (may be NULL): [iters][N+2Zc]
 * beliefs after each iteration; over (may be NULL, int[2]) counts input
 * beliefs beyond +-clip and check updates with a minimum above clip (the
 * cases where c_lib's FP32 min-sum, which clips minima and not inputs,
 * computes differently).  The backward adds d loss / d alpha and
 * d loss / d beta into g_alpha/g_beta ([iters][rows]) for the loss whose
 * gradient is g_lq (final beliefs) plus g_iter (per-iteration beliefs, may
 * be NULL), and writes d loss / d llr into g_llr (may be NULL).  detach = 1
 * follows detach_check_update_inputs=True. */
void orc_unrolled_forward_f32(const orc_code_t *c, int iters, const float *llr,
                              const float *alpha, const float *beta,
                              float clip, float *lq_out, float *iter_out,
                              int *over);
void orc_unrolled_forward_f64(const orc_code_t *c, int iters,
                              const double *llr, const double *alpha,
                              const double *beta, double clip, double *lq_out,
                              double *iter_out, int *over);
void orc_unrolled_backward_f32(const orc_code_t *c, int iters,
                               const float *llr, const float *alpha,
                               const float *beta, float clip, int detach,
                               const float *g_lq, const float *g_iter,
                               float *g_llr, double *g_alpha, double *g_beta);
void orc_unrolled_backward_f64(const orc_code_t *c, int iters,
                               const double *llr, const double *alpha,
                               const double *beta, double clip, int detach,
                               const double *g_lq, const double *g_iter,
                               double *g_llr, double *g_alpha,
                               double *g_beta);

#ifdef __cplusplus
}
#endif
#endif
