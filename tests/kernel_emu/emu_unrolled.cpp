// Host build of the unrolled trainable min-sum kernels (ldpc_unrolled.hip
// compiled with g++ under hip_cpu_emu.h), driven like
// ldpc_mi355_unrolled_forward / _backward.  Test infrastructure: lets CPU
// tests check the kernel source against the oracle.
#include "ldpc_unrolled.hip"

#include <vector>

#include "ldpc_host.h"

using namespace nrldpc;

namespace {

int make_shape(int bgn, int A, int crctype, int E, int T, float clip, int detach, int cbs,
               UnrolledShape &sh, std::vector<uint32_t> &etab) {
    ldpc_in_config_t in{(uint16_t)bgn, (uint16_t)A, (uint16_t)crctype, (uint16_t)E, 0, 2, 0};
    Params p;
    if (derive(in, p)) return -2;
    sh = unr_shape(p.bgn, p.Zc, p.N, T, clip, detach);
    if (cbs > 0) {
        if (!unr_cbs_fits(sh, cbs)) return -4;
        sh.cbs = cbs;
    }
    etab.resize(sh.edges);
    edge_table(p.bgn, p.Zc, lifting_set(p.Zc), etab.data());
    return 0;
}

}  // namespace

// Forward of `batch` codeblocks; rec receives the record (may be NULL).
extern "C" int emu_unrolled_forward(int bgn, int A, int crctype, int E, int T, float clip,
                                    int detach, int cbs, int batch, const float *llr,
                                    const float *alpha, const float *beta, float *lq,
                                    float *iter_lq, uint32_t *rec) {
    UnrolledShape sh;
    std::vector<uint32_t> etab;
    if (int st = make_shape(bgn, A, crctype, E, T, clip, detach, cbs, sh, etab)) return st;
    UnrolledArgs a = unr_args(sh, batch, etab.data());
    a.llr = llr;
    a.lq = lq;
    a.iter_lq = iter_lq;
    a.rec = rec;
    a.alpha = alpha;
    a.beta = beta;
    const unsigned grid = (batch + sh.cbs - 1) / sh.cbs, block = sh.cbs * sh.lpc;
    const size_t lds = (size_t)sh.cbs * a.stride * sizeof(float);
    emu_launch(grid, block, lds, [&] {
        if (sh.bgn == 1)
            unrolled_fwd_kernel<1>(a);
        else
            unrolled_fwd_kernel<2>(a);
    });
    return 0;
}

extern "C" long emu_unrolled_record_bytes(int bgn, int A, int crctype, int E, int T, float clip,
                                          int detach, int batch) {
    UnrolledShape sh;
    std::vector<uint32_t> etab;
    if (make_shape(bgn, A, crctype, E, T, clip, detach, 0, sh, etab)) return -1;
    return (long)unr_record_bytes(sh, batch);
}

// Backward (forward record `rec` of the same inputs).
extern "C" int emu_unrolled_backward(int bgn, int A, int crctype, int E, int T, float clip,
                                     int detach, int cbs, int batch, const float *llr,
                                     const float *alpha, const float *beta, const uint32_t *rec,
                                     const float *g_lq, const float *g_iter, float *g_llr,
                                     float *g_alpha, float *g_beta) {
    UnrolledShape sh;
    std::vector<uint32_t> etab;
    if (int st = make_shape(bgn, A, crctype, E, T, clip, detach, cbs, sh, etab)) return st;
    UnrolledArgs a = unr_args(sh, batch, etab.data());
    std::vector<unsigned char> scratch(unr_scratch_bytes(sh, batch) + 256, 0xCD);
    unr_bind_scratch(sh, batch, scratch.data(), a);
    a.llr = llr;
    a.rec = const_cast<uint32_t *>(rec);
    a.alpha = alpha;
    a.beta = beta;
    a.g_lq = g_lq;
    a.g_iter = g_iter;
    a.g_llr = g_llr;
    const unsigned grid = (batch + sh.cbs - 1) / sh.cbs, block = sh.cbs * sh.lpc;
    const size_t lds = (size_t)sh.cbs * a.stride * sizeof(float);
    emu_launch(grid, block, lds, [&] {
        if (sh.bgn == 1)
            unrolled_bwd_kernel<1>(a);
        else
            unrolled_bwd_kernel<2>(a);
    });
    const int m = batch * (sh.lpc / sh.seg);
    emu_launch(sh.T * sh.rows, kReduceThreads, 0,
               [&] { unrolled_reduce_kernel(a.partial, m, g_alpha, g_beta); });
    return 0;
}
