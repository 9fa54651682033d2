// Host build of the input-synthesis kernel (ldpc_synth.hip compiled with g++
// under hip_cpu_emu.h), driven like ldpc_mi355_synth_device.  Test
// infrastructure.
#include "ldpc_synth.hip"

#include <vector>

#include "ldpc_host.h"

using namespace nrldpc;

// noise_var / noise: the channel's given noise (both or neither, as
// ldpc_mi355_synth_device_noise); otherwise noise_db and the generator.
extern "C" int emu_synth_noise(int bgn, int A, int crctype, int E, int rv, int Qm, int n,
                               const float *noise_db, const float *noise_var,
                               const float *noise, uint64_t seed, const uint8_t *payload,
                               uint8_t *bits_out, uint8_t *cw_out, uint8_t *rm_out,
                               int16_t *llr_out, float *llr_f32, float *beliefs) {
    ldpc_in_config_t in{(uint16_t)bgn, (uint16_t)A, (uint16_t)crctype, (uint16_t)E,
                        (uint16_t)rv, (uint16_t)Qm, 0};
    Params p;
    if (derive(in, p)) return -2;
    std::vector<uint32_t> etab(p.bgn == 1 ? NR_BG1_EDGES : NR_BG2_EDGES);
    edge_table(p.bgn, p.Zc, lifting_set(p.Zc), etab.data());
    SynthArgs a{};
    a.noise_db = noise_db;
    a.noise_var = noise_var;
    a.noise = noise;
    a.payload = payload;
    a.bits_out = bits_out;
    a.cw_out = cw_out;
    a.rm_out = rm_out;
    a.llr_out = llr_out;
    a.llr_f32 = llr_f32;
    a.beliefs = beliefs;
    a.etab = etab.data();
    a.seed = seed;
    a.n = n;
    a.A = p.A;
    a.B = p.B;
    a.K = p.K;
    a.N = p.N;
    a.E = p.E;
    a.Zc = p.Zc;
    a.Qm = p.Qm;
    a.k0 = p.k0;
    a.crctype = p.crctype;
    emu_launch(n, kSynthThreads, 0, [&] {
        if (p.bgn == 1)
            ldpc_synth_kernel<1>(a);
        else
            ldpc_synth_kernel<2>(a);
    });
    return 0;
}

extern "C" int emu_synth(int bgn, int A, int crctype, int E, int rv, int Qm, int n,
                         const float *noise_db, uint64_t seed, const uint8_t *payload,
                         uint8_t *bits_out, uint8_t *cw_out, uint8_t *rm_out, int16_t *llr_out,
                         float *llr_f32, float *beliefs) {
    return emu_synth_noise(bgn, A, crctype, E, rv, Qm, n, noise_db, nullptr, nullptr, seed,
                           payload, bits_out, cw_out, rm_out, llr_out, llr_f32, beliefs);
}
