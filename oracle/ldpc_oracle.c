/*
 * ORACLE — test infrastructure only (see ldpc_oracle.h for the contract and
 * for what pins each function).  Restates the reference C library
 * (hahaliu2001/deeplearning_5g_ldpc_decoder, c_lib/) on the CPU, reentrant,
 * with posteriors and iteration counts exposed.  Built with
 * -ffp-contract=off so every float expression rounds exactly as written.
 */
#include "ldpc_oracle.h"
#include "oracle_bg_tables.h"

#include <math.h>
#include <float.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

/* ------------------------------------------------------------------ */
/* Parameters: c_lib/ldpc_internal.c:13-169 (TS 38.212 5.2.2, 5.3.2,  */
/* Table 5.4.2.1-2 with Ncb = N).                                     */
/* ------------------------------------------------------------------ */

/* TS 38.212 Table 5.3.2-1, all 51 lifting sizes with their set index. */
static const struct { int zc, ils; } kLifting[] = {
    {2, 0}, {4, 0}, {8, 0}, {16, 0}, {32, 0}, {64, 0}, {128, 0}, {256, 0},
    {3, 1}, {6, 1}, {12, 1}, {24, 1}, {48, 1}, {96, 1}, {192, 1}, {384, 1},
    {5, 2}, {10, 2}, {20, 2}, {40, 2}, {80, 2}, {160, 2}, {320, 2},
    {7, 3}, {14, 3}, {28, 3}, {56, 3}, {112, 3}, {224, 3},
    {9, 4}, {18, 4}, {36, 4}, {72, 4}, {144, 4}, {288, 4},
    {11, 5}, {22, 5}, {44, 5}, {88, 5}, {176, 5}, {352, 5},
    {13, 6}, {26, 6}, {52, 6}, {104, 6}, {208, 6},
    {15, 7}, {30, 7}, {60, 7}, {120, 7}, {240, 7},
};

int orc_code_init(orc_code_t *c, int bgn, int A, int crctype, int E, int rv,
                  int Qm, int harq_on)
{
    memset(c, 0, sizeof(*c));
    c->bgn = bgn; c->A = A; c->crctype = crctype; c->E = E; c->rv = rv;
    c->Qm = Qm; c->harq_on = harq_on;
    /* validation order and rules: c_lib/ldpc_internal.c:86-115, 124-131 */
    if (bgn != 1 && bgn != 2) return -1;
    if (crctype < 0 || crctype > 3) return -1;
    if (!(Qm == 1 || Qm == 2 || Qm == 4 || Qm == 6 || Qm == 8)) return -1;
    if (harq_on != 0 && harq_on != 1) return -1;
    if (rv < 0 || rv > 3) return -1;
    if (E % Qm != 0) return -1;
    if (A % 8 != 0) return -1;
    c->B = A + (crctype == 0 ? 16 : 24);
    if (bgn == 1 && c->B > 8448) return -1;
    if (bgn == 2 && c->B > 3840) return -1;
    /* Kb of TS 38.212 5.2.2 (c_lib/ldpc_internal.c:29-36) */
    int kb;
    if (bgn == 1) kb = 22;
    else if (c->B > 640) kb = 10;
    else if (c->B > 560) kb = 9;
    else if (c->B > 192) kb = 8;
    else kb = 6;
    /* minimum Zc of Table 5.3.2-1 with kb*Zc >= B */
    int best = -1;
    for (size_t i = 0; i < sizeof(kLifting) / sizeof(kLifting[0]); i++) {
        if (kb * kLifting[i].zc >= c->B &&
            (best < 0 || kLifting[i].zc < kLifting[best].zc))
            best = (int)i;
    }
    c->Zc = kLifting[best].zc;
    c->iLS = kLifting[best].ils;
    c->K = (bgn == 1 ? 22 : 10) * c->Zc;
    c->N = (bgn == 1 ? 66 : 50) * c->Zc;
    c->F = c->K - c->B;
    /* k0 with Ncb = N (c_lib/ldpc_internal.c:146-155) */
    static const int k0_bg1[4] = {0, 17, 33, 56};
    static const int k0_bg2[4] = {0, 13, 25, 43};
    c->k0 = (bgn == 1 ? k0_bg1[rv] : k0_bg2[rv]) * c->Zc;
    if (bgn == 1) {
        c->rows = 46; c->nnz = 316;
        c->row_off = orc_bg1_row_off; c->col = orc_bg1_col;
        c->vij = orc_bg1_vij + (size_t)c->iLS * 316;
    } else {
        c->rows = 42; c->nnz = 197;
        c->row_off = orc_bg2_row_off; c->col = orc_bg2_col;
        c->vij = orc_bg2_vij + (size_t)c->iLS * 197;
    }
    return 0;
}

int orc_params(int bgn, int A, int crctype, int E, int rv, int Qm,
               int harq_on, int *out7)
{
    orc_code_t c;
    int st = orc_code_init(&c, bgn, A, crctype, E, rv, Qm, harq_on);
    out7[0] = c.B; out7[1] = c.Zc; out7[2] = c.iLS; out7[3] = c.K;
    out7[4] = c.F; out7[5] = c.N; out7[6] = c.k0;
    return st;
}

/* ------------------------------------------------------------------ */
/* CRC: c_lib/nr_crc.c:170-179 semantics, computed by bitwise long      */
/* division of M(x)*x^w by g(x) (TS 38.212 5.1 generator polynomials). */
/* ------------------------------------------------------------------ */
uint32_t orc_crc(const uint8_t *bytes, int nbytes, int crctype)
{
    uint32_t poly; int w;
    switch (crctype) {
    case 0: poly = 0x1021; w = 16; break;      /* gCRC16  */
    case 1: poly = 0x864CFB; w = 24; break;    /* gCRC24A */
    case 2: poly = 0x800063; w = 24; break;    /* gCRC24B */
    case 3: poly = 0xB2B117; w = 24; break;    /* gCRC24C */
    default: return 0;
    }
    uint32_t mask = (w == 32) ? 0xFFFFFFFFu : ((1u << w) - 1u);
    uint32_t reg = 0;
    for (int i = 0; i < nbytes; i++) {
        for (int b = 7; b >= 0; b--) {
            uint32_t in = (bytes[i] >> b) & 1u;
            uint32_t top = (reg >> (w - 1)) & 1u;
            reg = (reg << 1) & mask;
            if (top ^ in) reg ^= poly;
        }
    }
    return reg & mask;
}

/* ------------------------------------------------------------------ */
/* Rate recovery + HARQ combine: c_lib/ldpc_rate_recovery.c:24-65.      */
/* ------------------------------------------------------------------ */
static inline int16_t wrap16(int32_t v) { return (int16_t)(uint16_t)(uint32_t)v; }

void orc_rate_recover(const orc_code_t *c, int16_t filler, const int16_t *llr,
                      int16_t *soft)
{
    const int Zc = c->Zc, N = c->N, S = c->Qm ? c->E / c->Qm : 0;
    if (!c->harq_on) memset(soft, 0, sizeof(int16_t) * (size_t)(N + 2 * Zc));
    for (int i = 0; i < c->F; i++) soft[c->B + i] = filler;
    const int fill_lo = c->B - 2 * Zc, fill_hi = c->K - 2 * Zc;
    int pos = c->k0;
    for (int k = 0; k < c->E;) {
        if (pos < fill_lo || pos >= fill_hi) {
            int idx = (k % S) * c->Qm + k / S;   /* inverse bit interleave */
            soft[pos + 2 * Zc] = wrap16((int32_t)soft[pos + 2 * Zc] + llr[idx]);
            k++;
        }
        pos = (pos + 1) % N;
    }
}

/* ------------------------------------------------------------------ */
/* Hard decision + CRC early stop shared by all decoders                */
/* (c_lib/fixedpoint_ldpc_decoder.c:79-84, 214-219).                    */
/* ------------------------------------------------------------------ */
#define HARD_PACK(T)                                                       \
    static void hard_pack_##T(const T *lq, int B, uint8_t *hard) {        \
        for (int byte = 0; byte < B / 8; byte++) {                         \
            uint8_t v = 0;                                                 \
            for (int b = 0; b < 8; b++)                                    \
                v = (uint8_t)((v << 1) | (lq[byte * 8 + b] < 0 ? 1 : 0));  \
            hard[byte] = v;                                                \
        }                                                                  \
    }
HARD_PACK(int16_t)
HARD_PACK(float)
HARD_PACK(double)

static inline int lifted_col(const orc_code_t *c, int e, int zx)
{
    return c->col[e] * c->Zc + (c->vij[e] + zx) % c->Zc;
}

/* ------------------------------------------------------------------ */
/* Algo 1: int16 normalized/offset min-sum,                            */
/* c_lib/fixedpoint_ldpc_decoder.c:28-66 (helpers), 103-144 (check),   */
/* 162-222 (layered schedule).                                         */
/* ------------------------------------------------------------------ */
void orc_check_minsum_i16(int deg, const int16_t *rel, int16_t alpha,
                          int16_t beta, int16_t max_llr, int16_t *msg)
{
    int16_t m1 = 32767, m2 = 32767;
    int idx = -1, sgn = 1;
    for (int j = 0; j < deg; j++) {
        int16_t v = rel[j];
        int16_t a;
        if (v >= 0) a = v;
        else { sgn = -sgn; a = (v == -32768) ? 32767 : (int16_t)(-v); }
        if (a < m1) { m2 = m1; m1 = a; idx = j; }
        else if (a < m2) m2 = a;
    }
    if (m1 > max_llr) m1 = max_llr;
    if (m2 > max_llr) m2 = max_llr;
    int16_t s[2];
    int16_t mm[2] = {m1, m2};
    for (int i = 0; i < 2; i++) {
        int32_t off = (int32_t)mm[i] - (int32_t)beta;
        if (off < 0) off = 0;
        int32_t scaled = (int32_t)alpha * off;
        s[i] = (int16_t)((scaled + (1 << 14)) >> 15);
    }
    for (int j = 0; j < deg; j++) {
        int16_t nv = (j == idx) ? s[1] : s[0];
        int32_t signed_v = (rel[j] >= 0) ? sgn * (int32_t)nv : -sgn * (int32_t)nv;
        msg[j] = wrap16(signed_v);
    }
}

static inline int16_t clip_i16(int32_t v, int16_t hi, int16_t lo)
{
    if (v > hi) return hi;
    if (v < lo) return lo;
    return (int16_t)v;
}

/* The layered schedule of c_lib/fixedpoint_ldpc_decoder.c:162-222 with the
 * (alpha, beta) pair of iteration it, row r taken from
 * alpha[it * stride + r], beta[it * stride + r] (stride 0: one pair per row
 * for every iteration).  Algos 1, 2 and 3 differ only in that lookup. */
static int decode_i16_rows(const orc_code_t *c, const orc_dec_t *d,
                           const int16_t *soft, uint8_t *hard, int16_t *lq_out,
                           int *iters, const int16_t *alpha,
                           const int16_t *beta, int stride)
{
    const int Zc = c->Zc, len = c->N + 2 * Zc;
    int16_t *LQ = (int16_t *)malloc(sizeof(int16_t) * (size_t)len);
    int16_t *Lr = (int16_t *)calloc((size_t)c->nnz * Zc, sizeof(int16_t));
    memcpy(LQ, soft, sizeof(int16_t) * (size_t)len);
    int16_t rel[32], msg[32];
    int colv[32];
    int status = 1, it;
    for (it = 0; it < d->max_iterations; it++) {
        for (int r = 0; r < c->rows; r++) {
            const int e0 = c->row_off[r], deg = c->row_off[r + 1] - e0;
            const int16_t a = alpha[(size_t)it * stride + r];
            const int16_t b = beta[(size_t)it * stride + r];
            for (int zx = 0; zx < Zc; zx++) {
                int16_t *lr = Lr + (size_t)e0 * Zc + (size_t)zx * deg;
                for (int j = 0; j < deg; j++) {
                    colv[j] = lifted_col(c, e0 + j, zx);
                    rel[j] = wrap16((int32_t)LQ[colv[j]] - lr[j]);
                }
                orc_check_minsum_i16(deg, rel, a, b, d->max_llr, msg);
                for (int j = 0; j < deg; j++) {
                    lr[j] = msg[j];
                    LQ[colv[j]] = clip_i16((int32_t)rel[j] + msg[j],
                                           d->max_llr, d->min_llr);
                }
            }
        }
        hard_pack_int16_t(LQ, c->B, hard);
        if (orc_crc(hard, c->B / 8, c->crctype) == 0) { status = 0; it++; break; }
    }
    if (iters) *iters = it;
    if (lq_out) memcpy(lq_out, LQ, sizeof(int16_t) * (size_t)len);
    free(LQ); free(Lr);
    return status;
}

int orc_decode_i16(const orc_code_t *c, const orc_dec_t *d,
                   const int16_t *soft, uint8_t *hard, int16_t *lq_out,
                   int *iters)
{
    int16_t a[ORC_MAX_ROWS], b[ORC_MAX_ROWS];
    for (int r = 0; r < c->rows; r++) {
        a[r] = (int16_t)d->alpha;   /* the int16 view the reference's helpers take */
        b[r] = (int16_t)d->beta;
    }
    return decode_i16_rows(c, d, soft, hard, lq_out, iters, a, b, 0);
}

/* ------------------------------------------------------------------ */
/* Algo 2: trained-table min-sum,                                      */
/* c_lib/neural_fixedpoint_ldpc_decoder.c:223-281 (alpha/beta of       */
/* iteration it, row r at it * rows + r, :248-250).                    */
/* ------------------------------------------------------------------ */
int orc_decode_i16_table(const orc_code_t *c, const orc_dec_t *d,
                         const int16_t *soft, uint8_t *hard, int16_t *lq_out,
                         int *iters, const int16_t *alpha,
                         const int16_t *beta, int table_iterations)
{
    if (iters) *iters = 0;
    if (d->max_iterations > table_iterations) return -2;
    return decode_i16_rows(c, d, soft, hard, lq_out, iters, alpha, beta,
                           c->rows);
}

/* Degree rule of the reference's default tables: Q15 alpha 26623 for
 * check degree <= 4, 24575 for 5..8, 22527 above 8
 * (c_lib/degree_fixedpoint_ldpc_decoder.c:9-43; the default trained
 * tables of c_lib/BG1_trained_ldpc_weights.c:27-33 apply the same rule to
 * every iteration with beta 0). */
int16_t orc_default_degree_alpha(int degree)
{
    if (degree <= 4) return 26623;
    if (degree <= 8) return 24575;
    return 22527;
}

void orc_default_trained_tables(int bgn, int iterations, int16_t *alpha,
                                int16_t *beta)
{
    const uint16_t *off = bgn == 1 ? orc_bg1_row_off : orc_bg2_row_off;
    const int rows = bgn == 1 ? 46 : 42;
    for (int it = 0; it < iterations; it++)
        for (int r = 0; r < rows; r++) {
            alpha[it * rows + r] = orc_default_degree_alpha(off[r + 1] - off[r]);
            beta[it * rows + r] = 0;
        }
}

/* ------------------------------------------------------------------ */
/* Algo 3: degree min-sum,                                             */
/* c_lib/degree_fixedpoint_ldpc_decoder.c:65-80 (alpha by exact degree */
/* d at entry d-1; degree 0 -> entry 0, degree > 19 -> entry 18),      */
/* 200-268 (schedule; beta from the decoder row).                      */
/* ------------------------------------------------------------------ */
int orc_decode_i16_degree(const orc_code_t *c, const orc_dec_t *d,
                          const int16_t *soft, uint8_t *hard, int16_t *lq_out,
                          int *iters, const int16_t *degree_alpha)
{
    int16_t a[ORC_MAX_ROWS], b[ORC_MAX_ROWS];
    for (int r = 0; r < c->rows; r++) {
        const int deg = c->row_off[r + 1] - c->row_off[r];
        if (degree_alpha) {
            const int k = deg == 0 ? 0 : (deg > ORC_MAX_DEGREE ? ORC_MAX_DEGREE - 1 : deg - 1);
            a[r] = degree_alpha[k];
        } else {
            a[r] = orc_default_degree_alpha(deg);
        }
        b[r] = (int16_t)d->beta;
    }
    return decode_i16_rows(c, d, soft, hard, lq_out, iters, a, b, 0);
}

/* ------------------------------------------------------------------ */
/* Algo 4: float min-sum, c_lib/floatpoint_ldpc_decoder.c:153-220,     */
/* 307-366.  Instantiated for float (bit-exact) and double (FP64 twin). */
/* ------------------------------------------------------------------ */
#define DEFINE_MS(T, NAME, BIGV, ABS)                                          \
int NAME(const orc_code_t *c, const orc_dec_t *d, const int16_t *soft,        \
         uint8_t *hard, T *lq_out, int *iters)                                \
{                                                                             \
    const int Zc = c->Zc, len = c->N + 2 * Zc;                                \
    T *LQ = (T *)malloc(sizeof(T) * (size_t)len);                             \
    T *Lr = (T *)calloc((size_t)c->nnz * Zc, sizeof(T));                      \
    const T alpha = (T)d->alpha / (T)32768.0;                                 \
    const T beta = (T)d->beta / (T)128.0;                                     \
    const T hi = (T)d->max_llr / (T)128.0, lo = (T)d->min_llr / (T)128.0;     \
    for (int i = 0; i < len; i++) LQ[i] = (T)soft[i] / (T)128.0;              \
    T rel[32]; int colv[32];                                                  \
    int status = 1, it;                                                       \
    for (it = 0; it < d->max_iterations; it++) {                              \
        for (int r = 0; r < c->rows; r++) {                                   \
            const int e0 = c->row_off[r], deg = c->row_off[r + 1] - e0;       \
            for (int zx = 0; zx < Zc; zx++) {                                 \
                T *lr = Lr + (size_t)e0 * Zc + (size_t)zx * deg;              \
                T m1 = BIGV, m2 = BIGV; int idx = 0, sgn = 1;                 \
                for (int j = 0; j < deg; j++) {                               \
                    colv[j] = lifted_col(c, e0 + j, zx);                      \
                    rel[j] = LQ[colv[j]] - lr[j];                             \
                }                                                             \
                for (int j = 0; j < deg; j++) {                               \
                    if (rel[j] < (T)0) sgn = -sgn;                            \
                    T a = ABS(rel[j]);                                        \
                    if (a < m1) { m2 = m1; m1 = a; idx = j; }                 \
                    else if (a < m2) m2 = a;                                  \
                }                                                             \
                if (m1 > hi) m1 = hi;                                         \
                if (m1 < (T)0) m1 = (T)0;                                     \
                if (m2 > hi) m2 = hi;                                         \
                if (m2 < (T)0) m2 = (T)0;                                     \
                T o1 = m1 - beta; if (o1 < (T)0) o1 = (T)0;                   \
                T o2 = m2 - beta; if (o2 < (T)0) o2 = (T)0;                   \
                const T n1 = alpha * o1, n2 = alpha * o2;                     \
                for (int j = 0; j < deg; j++) {                               \
                    T nv = (j == idx) ? n2 : n1;                              \
                    T es = (rel[j] < (T)0) ? -(T)sgn : (T)sgn;                \
                    lr[j] = es * nv;                                          \
                    T v = rel[j] + lr[j];                                     \
                    if (v > hi) v = hi;                                       \
                    if (v < lo) v = lo;                                       \
                    LQ[colv[j]] = v;                                          \
                }                                                             \
            }                                                                 \
        }                                                                     \
        hard_pack_##T(LQ, c->B, hard);                                        \
        if (orc_crc(hard, c->B / 8, c->crctype) == 0) {                       \
            status = 0; it++; break;                                          \
        }                                                                     \
    }                                                                         \
    if (iters) *iters = it;                                                   \
    if (lq_out) memcpy(lq_out, LQ, sizeof(T) * (size_t)len);                  \
    free(LQ); free(Lr);                                                       \
    return status;                                                            \
}
DEFINE_MS(float, orc_decode_f32_ms, FLT_MAX, fabsf)
DEFINE_MS(double, orc_decode_f64_ms, DBL_MAX, fabs)

/* ------------------------------------------------------------------ */
/* Algo 0: float BP, c_lib/floatpoint_ldpc_decoder.c:30-151, 237-298.   */
/* Exact zeros are special-cased (one zero: only that edge gets a       */
/* message; two or more: all messages zero); messages saturate at       */
/* +-2*19.07 when the extrinsic product reaches +-1; LQ is not clipped. */
/* ------------------------------------------------------------------ */
#define DEFINE_BP_CHECK(T, NAME, TANH, ATANH, CLIPV)                          \
static void NAME(int deg, const T *rel, T *msg)                               \
{                                                                             \
    T t[32], pre[33], suf[33];                                                \
    int nz = 0, zidx = -1;                                                    \
    T prod = (T)1;                                                            \
    for (int j = 0; j < deg; j++) {                                           \
        if (rel[j] == (T)0) { nz++; zidx = j; t[j] = (T)0; }                  \
        else { t[j] = TANH(rel[j] / (T)2); prod *= t[j]; }                    \
    }                                                                         \
    const T sat = (T)2 * CLIPV;                                               \
    if (nz >= 1) {                                                            \
        for (int j = 0; j < deg; j++) msg[j] = (T)0;                          \
        if (nz == 1)                                                          \
            msg[zidx] = prod >= (T)1 ? sat : prod <= (T)-1 ? -sat             \
                        : (T)2 * ATANH(prod);                                 \
        return;                                                               \
    }                                                                         \
    pre[0] = (T)1;                                                            \
    for (int j = 0; j < deg; j++) pre[j + 1] = pre[j] * t[j];                 \
    suf[deg] = (T)1;                                                          \
    for (int j = deg - 1; j >= 0; j--) suf[j] = suf[j + 1] * t[j];            \
    for (int j = 0; j < deg; j++) {                                           \
        T p = pre[j] * suf[j + 1];                                            \
        msg[j] = p >= (T)1 ? sat : p <= (T)-1 ? -sat : (T)2 * ATANH(p);       \
    }                                                                         \
}
DEFINE_BP_CHECK(float, bp_check_f32, tanhf, atanhf, 19.07f)
DEFINE_BP_CHECK(double, bp_check_f64, tanh, atanh, 19.07)

void orc_check_bp_f64(int deg, const double *rel, double *msg)
{
    bp_check_f64(deg, rel, msg);
}

#define DEFINE_BP(T, NAME, CHECK)                                             \
int NAME(const orc_code_t *c, const orc_dec_t *d, const int16_t *soft,        \
         uint8_t *hard, T *lq_out, int *iters)                                \
{                                                                             \
    const int Zc = c->Zc, len = c->N + 2 * Zc;                                \
    T *LQ = (T *)malloc(sizeof(T) * (size_t)len);                             \
    T *Lr = (T *)calloc((size_t)c->nnz * Zc, sizeof(T));                      \
    for (int i = 0; i < len; i++) LQ[i] = (T)soft[i] / (T)128.0;              \
    T rel[32], msg[32]; int colv[32];                                         \
    int status = 1, it;                                                       \
    for (it = 0; it < d->max_iterations; it++) {                              \
        for (int r = 0; r < c->rows; r++) {                                   \
            const int e0 = c->row_off[r], deg = c->row_off[r + 1] - e0;       \
            for (int zx = 0; zx < Zc; zx++) {                                 \
                T *lr = Lr + (size_t)e0 * Zc + (size_t)zx * deg;              \
                for (int j = 0; j < deg; j++) {                               \
                    colv[j] = lifted_col(c, e0 + j, zx);                      \
                    rel[j] = LQ[colv[j]] - lr[j];                             \
                }                                                             \
                CHECK(deg, rel, msg);                                         \
                for (int j = 0; j < deg; j++) {                               \
                    lr[j] = msg[j];                                           \
                    LQ[colv[j]] = rel[j] + msg[j];                            \
                }                                                             \
            }                                                                 \
        }                                                                     \
        hard_pack_##T(LQ, c->B, hard);                                        \
        if (orc_crc(hard, c->B / 8, c->crctype) == 0) {                       \
            status = 0; it++; break;                                          \
        }                                                                     \
    }                                                                         \
    if (iters) *iters = it;                                                   \
    if (lq_out) memcpy(lq_out, LQ, sizeof(T) * (size_t)len);                  \
    free(LQ); free(Lr);                                                       \
    return status;                                                            \
}
DEFINE_BP(float, orc_decode_f32_bp, bp_check_f32)
DEFINE_BP(double, orc_decode_f64_bp, bp_check_f64)

/* ------------------------------------------------------------------ */
/* Whole path: c_lib/c_ldpc_5g_decoder.c:34-79.                          */
/* ------------------------------------------------------------------ */
int orc_decode_path_tables(const int16_t *llr, uint8_t *hard, int bgn, int A,
                           int crctype, int E, int rv, int Qm, int harq_on,
                           const orc_dec_t *d, int16_t *soft, int16_t *lq_i16,
                           float *lq_f32, int *iters, const int16_t *trained_alpha,
                           const int16_t *trained_beta, int trained_iterations,
                           const int16_t *degree_alpha)
{
    orc_code_t c;
    if (iters) *iters = 0;
    if (orc_code_init(&c, bgn, A, crctype, E, rv, Qm, harq_on) != 0) return -2;
    if (d->algo == 2 && d->max_iterations > trained_iterations) return -2;
    orc_rate_recover(&c, d->filling_bit_value, llr, soft);
    switch (d->algo) {
    case 0: return orc_decode_f32_bp(&c, d, soft, hard, lq_f32, iters);
    case 1: return orc_decode_i16(&c, d, soft, hard, lq_i16, iters);
    case 2: return orc_decode_i16_table(&c, d, soft, hard, lq_i16, iters, trained_alpha,
                                        trained_beta, trained_iterations);
    case 3: return orc_decode_i16_degree(&c, d, soft, hard, lq_i16, iters, degree_alpha);
    case 4: return orc_decode_f32_ms(&c, d, soft, hard, lq_f32, iters);
    default: return -3;
    }
}

int orc_decode_path(const int16_t *llr, uint8_t *hard, int bgn, int A,
                    int crctype, int E, int rv, int Qm, int harq_on,
                    const orc_dec_t *d, int16_t *soft, int16_t *lq_i16,
                    float *lq_f32, int *iters)
{
    int16_t ta[32 * ORC_MAX_ROWS], tb[32 * ORC_MAX_ROWS];
    if (d->algo == 2 && (bgn == 1 || bgn == 2))
        orc_default_trained_tables(bgn, 32, ta, tb);
    return orc_decode_path_tables(llr, hard, bgn, A, crctype, E, rv, Qm, harq_on, d, soft,
                                  lq_i16, lq_f32, iters, ta, tb, 32, NULL);
}

/* ------------------------------------------------------------------ */
/* Encoder (test input generation): TS 38.212 5.3.2 parity solve,       */
/* as in c_lib/ldpc_encoder.c:221-242, with the 4 core parity groups    */
/* solved generically from the table and filler bits encoded as 0.      */
/* ------------------------------------------------------------------ */
int orc_encode(const orc_code_t *c, const uint8_t *payload_bits, uint8_t *cw)
{
    const int Zc = c->Zc, kbf = (c->bgn == 1) ? 22 : 10;
    const int ncol = (c->bgn == 1) ? 68 : 52;
    memset(cw, 0, (size_t)ncol * Zc);
    /* info + CRC bits (payload_bits holds the B bits), filler zeros */
    memcpy(cw, payload_bits, (size_t)c->B);
    uint8_t *syn = (uint8_t *)calloc((size_t)4 * Zc, 1);
    for (int r = 0; r < 4; r++)
        for (int zx = 0; zx < Zc; zx++) {
            int s = 0;
            for (int e = c->row_off[r]; e < c->row_off[r + 1]; e++)
                if (c->col[e] < kbf) s ^= cw[lifted_col(c, e, zx)];
            syn[r * Zc + zx] = (uint8_t)s;
        }
    /* p0: sum of the 4 core rows leaves one circulant on column kbf */
    int cnt[512];
    memset(cnt, 0, sizeof(cnt));
    for (int r = 0; r < 4; r++)
        for (int e = c->row_off[r]; e < c->row_off[r + 1]; e++)
            if (c->col[e] == kbf) cnt[c->vij[e] % Zc] ^= 1;
    int s0 = -1, odd = 0;
    for (int s = 0; s < Zc; s++) if (cnt[s]) { s0 = s; odd++; }
    if (odd != 1) { free(syn); return -1; }
    for (int zx = 0; zx < Zc; zx++) {
        int v = syn[zx] ^ syn[Zc + zx] ^ syn[2 * Zc + zx] ^ syn[3 * Zc + zx];
        cw[kbf * Zc + (zx + s0) % Zc] = (uint8_t)v;
    }
    /* remaining core groups: repeatedly solve a row with one unknown */
    int known[4] = {1, 0, 0, 0};
    for (int pass = 0; pass < 4; pass++)
        for (int r = 0; r < 4; r++) {
            int unk = -1, nunk = 0, ue = -1;
            for (int e = c->row_off[r]; e < c->row_off[r + 1]; e++) {
                int g = c->col[e] - kbf;
                if (g >= 0 && g < 4 && !known[g]) { unk = g; ue = e; nunk++; }
            }
            if (nunk != 1) continue;
            for (int zx = 0; zx < Zc; zx++) {
                int v = syn[r * Zc + zx];
                for (int e = c->row_off[r]; e < c->row_off[r + 1]; e++) {
                    int g = c->col[e] - kbf;
                    if (g >= 0 && g < 4 && known[g]) v ^= cw[lifted_col(c, e, zx)];
                }
                cw[lifted_col(c, ue, zx)] = (uint8_t)v;
            }
            known[unk] = 1;
        }
    free(syn);
    if (!(known[1] && known[2] && known[3])) return -1;
    /* extension rows: each has one degree-1 parity column past kbf+3 */
    for (int r = 4; r < c->rows; r++) {
        int pe = -1;
        for (int e = c->row_off[r]; e < c->row_off[r + 1]; e++)
            if (c->col[e] >= kbf + 4) pe = e;
        for (int zx = 0; zx < Zc; zx++) {
            int v = 0;
            for (int e = c->row_off[r]; e < c->row_off[r + 1]; e++)
                if (e != pe) v ^= cw[lifted_col(c, e, zx)];
            cw[lifted_col(c, pe, zx)] = (uint8_t)v;
        }
    }
    return 0;
}

int orc_syndrome_ok(const orc_code_t *c, const uint8_t *cw)
{
    for (int r = 0; r < c->rows; r++)
        for (int zx = 0; zx < c->Zc; zx++) {
            int s = 0;
            for (int e = c->row_off[r]; e < c->row_off[r + 1]; e++)
                s ^= cw[lifted_col(c, e, zx)];
            if (s) return 0;
        }
    return 1;
}

/* Rate matching (c_lib/ldpc_rate_match.c:27-106): circular-buffer bit
 * selection from k0 over cw[2Zc..N+2Zc) skipping filler, then the Qm-row
 * bit interleaver of TS 38.212 5.4.2.2. */
void orc_rate_match(const orc_code_t *c, const uint8_t *cw, uint8_t *rm_bits)
{
    const int Zc = c->Zc, N = c->N, E = c->E, Qm = c->Qm, S = E / Qm;
    const int fill_lo = c->B - 2 * Zc, fill_hi = c->K - 2 * Zc;
    uint8_t *e = (uint8_t *)malloc((size_t)(E > 0 ? E : 1));
    int pos = c->k0;
    for (int k = 0; k < E;) {
        if (pos < fill_lo || pos >= fill_hi) e[k++] = cw[pos + 2 * Zc];
        pos = (pos + 1) % N;
    }
    for (int j = 0; j < S; j++)
        for (int i = 0; i < Qm; i++) rm_bits[j * Qm + i] = e[i * S + j];
    free(e);
}

/* ------------------------------------------------------------------ */
/* CPU baseline: many codeblocks of one config on POSIX threads.        */
/* ------------------------------------------------------------------ */
typedef struct {
    const int16_t *llr; uint8_t *hard; int32_t *status; int32_t *iters;
    int n, next; pthread_mutex_t mu;
    orc_code_t code; orc_dec_t dec;
} batch_job_t;

static void *batch_worker(void *arg)
{
    batch_job_t *job = (batch_job_t *)arg;
    const orc_code_t *c = &job->code;
    int16_t *soft = (int16_t *)malloc(sizeof(int16_t) * (size_t)(c->N + 2 * c->Zc));
    for (;;) {
        pthread_mutex_lock(&job->mu);
        int i = job->next++;
        pthread_mutex_unlock(&job->mu);
        if (i >= job->n) break;
        orc_rate_recover(c, job->dec.filling_bit_value,
                         job->llr + (size_t)i * c->E, soft);
        int it = 0;
        job->status[i] = orc_decode_i16(c, &job->dec, soft,
                                        job->hard + (size_t)i * (c->B / 8),
                                        NULL, &it);
        job->iters[i] = it;
    }
    free(soft);
    return NULL;
}

double orc_decode_batch_i16(const int16_t *llr, uint8_t *hard, int32_t *status,
                            int32_t *iters, int n, int bgn, int A, int crctype,
                            int E, int rv, int Qm, const orc_dec_t *d,
                            int threads)
{
    batch_job_t job;
    memset(&job, 0, sizeof(job));
    if (orc_code_init(&job.code, bgn, A, crctype, E, rv, Qm, 0) != 0) return -1.0;
    job.llr = llr; job.hard = hard; job.status = status; job.iters = iters;
    job.n = n; job.dec = *d;
    pthread_mutex_init(&job.mu, NULL);
    if (threads < 1) threads = 1;
    pthread_t *tid = (pthread_t *)malloc(sizeof(pthread_t) * (size_t)threads);
    struct timespec t0, t1;
    clock_gettime(CLOCK_MONOTONIC, &t0);
    for (int t = 0; t < threads; t++) pthread_create(&tid[t], NULL, batch_worker, &job);
    for (int t = 0; t < threads; t++) pthread_join(tid[t], NULL);
    clock_gettime(CLOCK_MONOTONIC, &t1);
    free(tid);
    pthread_mutex_destroy(&job.mu);
    return (double)(t1.tv_sec - t0.tv_sec) + 1e-9 * (double)(t1.tv_nsec - t0.tv_nsec);
}

/* ------------------------------------------------------------------ */
/* Test channel (c_lib/tx_rx_process.c:20-324).                       */
/* ------------------------------------------------------------------ */

/* One N(0, 1) draw from two rand() calls (Box-Muller, :59-65). */
static float orc_gauss(void)
{
    const float u1 = ((float)rand() + 1.0f) / ((float)RAND_MAX + 2.0f);
    const float u2 = ((float)rand() + 1.0f) / ((float)RAND_MAX + 2.0f);
    return sqrtf(-2.0f * logf(u1)) * cosf(6.28318530717958647692f * u2);
}

/* Max-log LLR of one axis bit as the reference writes it (:101-232):
 * breakpoints bp[0..n-2] in units of A; segment i (r below bp[i] * A, the
 * last one above all) is sgn[i] * ((mul[i] * A) * (r + off[i] * A)) / nv,
 * with r alone for off 0. */
typedef struct {
    int n;
    int bp[15], mul[15], sgn[15], off[15];
} orc_axis_bit_t;

static const orc_axis_bit_t kQam16[2] = {
    {3, {-2, 2}, {8, 4, 8}, {1, 1, 1}, {1, 0, -1}},
    {2, {0}, {4, 4}, {1, -1}, {2, -2}},
};
static const orc_axis_bit_t kQam64[3] = {
    {7, {-6, -4, -2, 2, 4, 6}, {16, 12, 8, 4, 8, 12, 16}, {1, 1, 1, 1, 1, 1, 1},
     {3, 2, 1, 0, -1, -2, -3}},
    {6, {-6, -2, 0, 2, 6}, {8, 4, 8, 8, 4, 8}, {1, 1, 1, -1, -1, -1}, {5, 4, 3, -3, -4, -5}},
    {4, {-4, 0, 4}, {4, 4, 4, 4}, {1, -1, 1, -1}, {6, 2, -2, -6}},
};
static const orc_axis_bit_t kQam256[4] = {
    {15, {-14, -12, -10, -8, -6, -4, -2, 2, 4, 6, 8, 10, 12, 14},
     {32, 28, 24, 20, 16, 12, 8, 4, 8, 12, 16, 20, 24, 28, 32},
     {1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1},
     {7, 6, 5, 4, 3, 2, 1, 0, -1, -2, -3, -4, -5, -6, -7}},
    {14, {-14, -12, -10, -6, -4, -2, 0, 2, 4, 6, 10, 12, 14},
     {16, 12, 8, 4, 8, 12, 16, 16, 12, 8, 4, 8, 12, 16},
     {1, 1, 1, 1, 1, 1, 1, -1, -1, -1, -1, -1, -1, -1},
     {11, 10, 9, 8, 7, 6, 5, -5, -6, -7, -8, -9, -10, -11}},
    {12, {-14, -10, -8, -6, -2, 0, 2, 6, 8, 10, 14},
     {8, 4, 8, 8, 4, 8, 8, 4, 8, 8, 4, 8},
     {1, 1, 1, -1, -1, -1, 1, 1, 1, -1, -1, -1},
     {13, 12, 11, 5, 4, 3, -3, -4, -5, -11, -12, -13}},
    {8, {-12, -8, -4, 0, 4, 8, 12}, {4, 4, 4, 4, 4, 4, 4, 4},
     {1, -1, 1, -1, 1, -1, 1, -1}, {14, 10, 6, 2, -2, -6, -10, -14}},
};

static float orc_axis_llr(const orc_axis_bit_t *b, float r, float A, float nv)
{
    int i = 0;
    while (i < b->n - 1 && !(r < (float)b->bp[i] * A)) i++;
    const float x = b->off[i] == 0 ? r : r + (float)b->off[i] * A;
    const float v = ((float)b->mul[i] * A) * x / nv;
    return b->sgn[i] < 0 ? -v : v;
}

/* Gray PAM level of one axis from its nb bits (MSB first), an exact small
 * integer (TS 38.211 5.1; :68-99). */
static int orc_pam_level(const uint8_t *axis_bits, int nb)
{
    int lvl = 0;
    for (int m = nb - 1; m >= 0; m--) {
        const int s = axis_bits[m] ? -1 : 1;
        lvl = (m == nb - 1) ? s : s * ((1 << (nb - 1 - m)) - lvl);
    }
    return lvl;
}

int orc_tx_rx(const uint8_t *bits, int nbits, int Qm, float noise_db, unsigned seed,
              float *noise_out, float *noise_var_out, float *llr_out, int16_t *q_out)
{
    if (!bits || !q_out || nbits <= 0) return -1;
    if (!(Qm == 1 || Qm == 2 || Qm == 4 || Qm == 6 || Qm == 8) || nbits % Qm) return -1;
    srand(seed);
    const float nv = powf(10.0f, noise_db / 10.0f);
    const float sigma = sqrtf(nv / 2.0f);
    if (noise_var_out) *noise_var_out = nv;
    const float A = Qm <= 2 ? 1.0f / sqrtf(2.0f)
                  : Qm == 4 ? 1.0f / sqrtf(10.0f)
                  : Qm == 6 ? 1.0f / sqrtf(42.0f) : 1.0f / sqrtf(170.0f);
    const orc_axis_bit_t *tab = Qm == 4 ? kQam16 : Qm == 6 ? kQam64 : kQam256;
    const int nb = Qm / 2;
    for (int i = 0; i < nbits / Qm; i++) {
        const uint8_t *sb = bits + (size_t)i * Qm;
        float re, im;
        if (Qm == 1) {
            re = im = (sb[0] ? -1.0f : 1.0f) * A;
        } else {
            uint8_t ax[2][4];
            for (int k = 0; k < nb; k++) {
                ax[0][k] = sb[2 * k];
                ax[1][k] = sb[2 * k + 1];
            }
            re = (float)orc_pam_level(ax[0], nb) * A;
            im = (float)orc_pam_level(ax[1], nb) * A;
        }
        const float n_re = sigma * orc_gauss();
        const float n_im = sigma * orc_gauss();
        re += n_re;
        im += n_im;
        if (noise_out) {
            noise_out[2 * i] = n_re;
            noise_out[2 * i + 1] = n_im;
        }
        float llr[8];
        if (Qm == 1) {
            llr[0] = 4.0f * (re + im) * A / nv;
        } else if (Qm == 2) {
            llr[0] = 4.0f * re * A / nv;
            llr[1] = 4.0f * im * A / nv;
        } else {
            for (int k = 0; k < nb; k++) {
                llr[2 * k] = orc_axis_llr(&tab[k], re, A, nv);
                llr[2 * k + 1] = orc_axis_llr(&tab[k], im, A, nv);
            }
        }
        for (int j = 0; j < Qm; j++) {
            const float scaled = llr[j] * 128.0f;
            int32_t q = (int32_t)(scaled >= 0.0f ? scaled + 0.5f : scaled - 0.5f);
            q = q > 10000 ? 10000 : (q < -10000 ? -10000 : q);
            q_out[(size_t)i * Qm + j] = (int16_t)q;
            if (llr_out) llr_out[(size_t)i * Qm + j] = llr[j];
        }
    }
    return 0;
}
