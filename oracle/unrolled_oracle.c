/*
 * ORACLE — test infrastructure only (see ldpc_oracle.h for the contract).
 *
 * The unrolled float layered min-sum of the reference's trainable decoder
 * (deep_learning/deeplearning_ldpc_model.py:315-410, CTableNeuralLayeredMin-
 * SumDecoder.forward) on one codeblock, and its reverse-mode gradient with
 * respect to the input LLRs and the per-(iteration, base-graph row) alpha
 * and beta values, derived by hand from the forward's autograd graph:
 *
 *   rel    = LQ[cols] - Lr_old                        (:364)
 *   s      = rel < 0 ? -1 : +1                        (:366-370; no gradient)
 *   a      = |rel|          d/d rel = sgn(rel), sgn(0) = 0
 *   m1,i1  = min_j a        gradient to the first minimal index (:372)
 *   m2,i2  = min_{j!=i1} a  gradient to its first minimal index (:376-378)
 *   u_j    = j == i1 ? m2 : m1                        (:381)
 *   c_j    = max(u_j - beta, 0)   passes where u_j - beta >= 0 (:382)
 *   Lr_new = (S * s_j) * (alpha * c_j),  S = prod_j s_j (:380, :383-384)
 *   q_j    = rel_j + Lr_new_j, clamped to [-clip, clip] when clipping
 *            (:386-388); the clamp passes gradient where -clip <= q <= clip
 *   LQ[cols] = q (the old LQ at cols receives gradient only through rel)
 *
 * With detach (deeplearning_ldpc_model.py:365) rel is a constant inside the
 * check update, so no gradient reaches LQ_old or Lr_old through the row.
 *
 * T = float restates the forward bit for bit (single IEEE operations,
 * -ffp-contract=off); the float backward chains per-edge gradients in float
 * in the same operation order as the HIP kernel and accumulates the alpha and
 * beta sums in double.  T = double is the FP64 twin used for finite
 * differences.  Pinning: the forward against the reference FP32 min-sum
 * fixtures (same row update as c_lib/floatpoint_ldpc_decoder.c:189-220 when
 * the two minima stay below max_llr), the gradient against FP64 central
 * finite differences of the FP64 twin and against torch autograd of a torch
 * restatement of the forward (tests/test_unrolled_oracle.py).
 */
#include "ldpc_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

static inline int ucol(const orc_code_t *c, int e, int zx)
{
    return c->col[e] * c->Zc + (c->vij[e] + zx) % c->Zc;
}

#define DEFINE_UNROLLED(T, ACC, SFX)                                           \
/* Forward over one codeblock.  rel_store (iters*nnz*Zc, may be NULL)      */ \
/* receives every rel value, [t][edge][zx]; over (may be NULL) counts      */ \
/* [0] input beliefs beyond +-clip, [1] check updates whose m1 or m2       */ \
/* exceeds clip (where c_lib's FP32 min-sum, which clips the minima,       */ \
/* differs from this update).                                              */ \
static void unrolled_fwd_##SFX(const orc_code_t *c, int iters, const T *llr,  \
                               const T *alpha, const T *beta, T clip,        \
                               T *lq, T *iter_out, T *rel_store, int *over)  \
{                                                                             \
    const int Zc = c->Zc, n = c->N + 2 * Zc, R = c->rows;                     \
    T *Lr = (T *)calloc((size_t)c->nnz * Zc, sizeof(T));                      \
    T rel[32], a[32];                                                         \
    int colv[32], neg[32];                                                    \
    if (over) over[0] = over[1] = 0;                                          \
    for (int i = 0; i < n; i++) {                                             \
        T v = llr[i];                                                         \
        if (clip > (T)0) {                                                    \
            if (over && (v < -clip || v > clip)) over[0]++;                   \
            if (v < -clip) v = -clip;                                         \
            if (v > clip) v = clip;                                           \
        }                                                                     \
        lq[i] = v;                                                            \
    }                                                                         \
    for (int t = 0; t < iters; t++) {                                         \
        for (int r = 0; r < R; r++) {                                         \
            const int e0 = c->row_off[r], deg = c->row_off[r + 1] - e0;       \
            const T al = alpha[t * R + r], be = beta[t * R + r];              \
            for (int zx = 0; zx < Zc; zx++) {                                 \
                int S = 0, i1 = 0, i2 = -1;                                   \
                T m1 = (T)INFINITY, m2 = (T)INFINITY;                         \
                for (int j = 0; j < deg; j++) {                               \
                    colv[j] = ucol(c, e0 + j, zx);                            \
                    rel[j] = lq[colv[j]] - Lr[(size_t)(e0 + j) * Zc + zx];    \
                    if (rel_store)                                            \
                        rel_store[((size_t)t * c->nnz + e0 + j) * Zc + zx] =  \
                            rel[j];                                           \
                    neg[j] = rel[j] < (T)0;                                   \
                    S ^= neg[j];                                              \
                    a[j] = rel[j] < (T)0 ? -rel[j] : rel[j];                  \
                    if (a[j] < m1) { m1 = a[j]; i1 = j; }                     \
                }                                                             \
                for (int j = 0; j < deg; j++)                                 \
                    if (j != i1 && (i2 < 0 || a[j] < m2)) { m2 = a[j]; i2 = j; }\
                if (over && clip > (T)0 && (m1 > clip || m2 > clip)) over[1]++; \
                T d1 = m1 - be, d2 = m2 - be;                                 \
                if (d1 < (T)0) d1 = (T)0;                                     \
                if (d2 < (T)0) d2 = (T)0;                                     \
                const T g1 = al * d1, g2 = al * d2;                           \
                for (int j = 0; j < deg; j++) {                               \
                    const T mag = (j == i1) ? g2 : g1;                        \
                    const T nl = (neg[j] ^ S) ? -mag : mag;                   \
                    T q = rel[j] + nl;                                        \
                    if (clip > (T)0) {                                        \
                        if (q < -clip) q = -clip;                             \
                        if (q > clip) q = clip;                               \
                    }                                                         \
                    lq[colv[j]] = q;                                          \
                    Lr[(size_t)(e0 + j) * Zc + zx] = nl;                      \
                }                                                             \
            }                                                                 \
        }                                                                     \
        if (iter_out) memcpy(iter_out + (size_t)t * n, lq, sizeof(T) * n);    \
    }                                                                         \
    free(Lr);                                                                 \
}                                                                             \
                                                                              \
void orc_unrolled_forward_##SFX(const orc_code_t *c, int iters, const T *llr, \
                                const T *alpha, const T *beta, T clip,       \
                                T *lq_out, T *iter_out, int *over)           \
{                                                                             \
    unrolled_fwd_##SFX(c, iters, llr, alpha, beta, clip, lq_out, iter_out,   \
                       NULL, over);                                           \
}                                                                             \
                                                                              \
void orc_unrolled_backward_##SFX(const orc_code_t *c, int iters,              \
                                 const T *llr, const T *alpha, const T *beta, \
                                 T clip, int detach, const T *g_lq,          \
                                 const T *g_iter, T *g_llr, ACC *g_alpha,    \
                                 ACC *g_beta)                                \
{                                                                             \
    const int Zc = c->Zc, n = c->N + 2 * Zc, R = c->rows;                     \
    T *rs = (T *)malloc(sizeof(T) * (size_t)iters * c->nnz * Zc);             \
    T *lq = (T *)malloc(sizeof(T) * (size_t)n);                               \
    T *glq = (T *)malloc(sizeof(T) * (size_t)n);                              \
    T *glr = (T *)calloc((size_t)c->nnz * Zc, sizeof(T));                     \
    unrolled_fwd_##SFX(c, iters, llr, alpha, beta, clip, lq, NULL, rs, NULL); \
    memcpy(glq, g_lq, sizeof(T) * n);                                         \
    T rel[32], a[32], gq[32];                                                 \
    int colv[32], neg[32];                                                    \
    for (int t = iters - 1; t >= 0; t--) {                                    \
        if (g_iter)                                                           \
            for (int i = 0; i < n; i++) glq[i] += g_iter[(size_t)t * n + i];  \
        for (int r = R - 1; r >= 0; r--) {                                    \
            const int e0 = c->row_off[r], deg = c->row_off[r + 1] - e0;       \
            const T al = alpha[t * R + r], be = beta[t * R + r];              \
            ACC ga = 0, gb = 0;                                               \
            for (int zx = 0; zx < Zc; zx++) {                                 \
                int S = 0, i1 = 0, i2 = -1;                                   \
                T m1 = (T)INFINITY, m2 = (T)INFINITY;                         \
                for (int j = 0; j < deg; j++) {                               \
                    colv[j] = ucol(c, e0 + j, zx);                            \
                    rel[j] = rs[((size_t)t * c->nnz + e0 + j) * Zc + zx];     \
                    neg[j] = rel[j] < (T)0;                                   \
                    S ^= neg[j];                                              \
                    a[j] = rel[j] < (T)0 ? -rel[j] : rel[j];                  \
                    if (a[j] < m1) { m1 = a[j]; i1 = j; }                     \
                }                                                             \
                for (int j = 0; j < deg; j++)                                 \
                    if (j != i1 && (i2 < 0 || a[j] < m2)) { m2 = a[j]; i2 = j; }\
                const T d1 = m1 - be, d2 = m2 - be;                           \
                const T c1 = d1 < (T)0 ? (T)0 : d1, c2 = d2 < (T)0 ? (T)0 : d2;\
                const T g1 = al * c1, g2 = al * c2;                           \
                T gm1 = 0, gm2 = 0;                                           \
                for (int j = 0; j < deg; j++) {                               \
                    const int first = (j == i1);                              \
                    const T mag = first ? g2 : g1;                            \
                    const int flip = neg[j] ^ S;                              \
                    const T nl = flip ? -mag : mag;                           \
                    const T q = rel[j] + nl;                                  \
                    T g = glq[colv[j]];                                       \
                    if (clip > (T)0 && !(q >= -clip && q <= clip)) g = (T)0;  \
                    gq[j] = g;                                                \
                    const T gn = g + glr[(size_t)(e0 + j) * Zc + zx];         \
                    const T gmag = flip ? -gn : gn;                           \
                    ga += (ACC)gmag * (ACC)(first ? c2 : c1);                 \
                    const int pass = first ? (d2 >= (T)0) : (d1 >= (T)0);     \
                    const T gd = pass ? gmag * al : (T)0;                     \
                    gb -= (ACC)gd;                                            \
                    if (first) gm2 = gd; else gm1 += gd;                      \
                }                                                             \
                for (int j = 0; j < deg; j++) {                               \
                    T gabs = (j == i1) ? gm1 : (j == i2) ? gm2 : (T)0;        \
                    const T sg = rel[j] > (T)0 ? (T)1                         \
                                 : rel[j] < (T)0 ? (T)-1 : (T)0;              \
                    T gr = detach ? (T)0 : gq[j] + gabs * sg;                 \
                    glq[colv[j]] = gr;                                        \
                    glr[(size_t)(e0 + j) * Zc + zx] = -gr;                    \
                }                                                             \
            }                                                                 \
            g_alpha[t * R + r] += ga;                                         \
            g_beta[t * R + r] += gb;                                          \
        }                                                                     \
    }                                                                         \
    if (g_llr)                                                                \
        for (int i = 0; i < n; i++) {                                         \
            const T v = llr[i];                                               \
            g_llr[i] = (clip > (T)0 && !(v >= -clip && v <= clip)) ? (T)0     \
                                                                    : glq[i]; \
        }                                                                     \
    free(rs); free(lq); free(glq); free(glr);                                 \
}

DEFINE_UNROLLED(float, double, f32)
DEFINE_UNROLLED(double, double, f64)
