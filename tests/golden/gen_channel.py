"""Writes tests/golden/channel_vectors.npz: the reference's test channel
(c_lib/tx_rx_process.c:286-324) on rate-matched codeword bits, with the noise
stream of srand(seed).

Runs ONLY in the development container: it loads the from-source reference
library oracle/_ref/libnrLDPC_ref.so (make -f oracle/Makefile.ref).  For each
case it
  * draws a payload, attaches the CRC, encodes and rate-matches it with the
    oracle (pinned by encoder_vectors.npz);
  * calls the reference's tx_rx_process after srand(seed): the expected int16
    LLRs;
  * runs the oracle's restatement orc_tx_rx with the same seed and asserts
    its int16 LLRs equal the reference's, bit for bit;
  * stores the payload, the noise values the channel added (as the oracle
    drew them: the reference's noise_sigma * randn() in call order), the noise
    variance, the oracle's float LLRs and the reference's int16 LLRs.
The GPU test feeds the payload and noise to ldpc_mi355_synth_device_noise and
compares its LLRs with these.

    python tests/golden/gen_channel.py
"""
import ctypes
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, HERE)

from gen_golden import Ref  # noqa: E402
from oracle import oracle as O  # noqa: E402

# (bgn, A, crctype, E, rv, Qm, harq_on), noise_power_db, seed
CASES = [
    ((2, 40, 0, 200, 0, 1, 0), -2.0, 11),
    ((2, 1016, 2, 3120, 0, 2, 0), -1.0, 12),
    ((1, 2088, 2, 6336, 1, 2, 0), 3.0, 13),
    ((2, 1016, 2, 2080, 0, 4, 0), -9.0, 14),
    ((2, 2056, 2, 4160, 2, 4, 0), -6.0, 15),
    ((1, 1496, 2, 4560, 0, 6, 0), -14.0, 16),
    ((2, 488, 1, 1200, 3, 6, 0), -10.0, 17),
    ((1, 3816, 1, 7632, 0, 8, 0), -20.0, 18),
    ((2, 256, 1, 808, 0, 8, 0), -16.0, 19),
    ((2, 256, 1, 800, 0, 8, 0), 10.0, 20),
]


def main():
    ref = Ref()
    ref.L.tx_rx_process.restype = ctypes.c_int
    rng = np.random.default_rng(2024)
    out = {k: [] for k in ("cfg", "noise_db", "seed", "noise_var", "payload", "noise", "llr_f32",
                           "llr_ref")}
    for cfg, ndb, seed in CASES:
        bgn, A, crctype, E, rv, Qm, _ = cfg
        c = O.code(*cfg)
        pay = rng.integers(0, 2, A).astype(np.uint8)
        rm = O.rate_match(c, O.encode(c, O.attach_crc(pay, crctype)))
        packed = np.packbits(rm)
        q_ref = np.zeros(E, np.int16)
        ref.libc.srand(seed)
        assert ref.L.tx_rx_process(packed.ctypes.data_as(ctypes.c_void_p), E, Qm,
                                   ctypes.c_float(ndb),
                                   q_ref.ctypes.data_as(ctypes.c_void_p)) == 0
        noise, nv, llr, q = O.tx_rx(rm, Qm, ndb, seed)
        assert np.array_equal(q, q_ref), (cfg, int(np.sum(q != q_ref)))
        print(f"{cfg} noise {ndb} dB seed {seed}: {E} LLRs, oracle == reference; "
              f"|LLR| max {np.abs(q_ref).max()}, clipped {int(np.sum(np.abs(q_ref) == 10000))}")
        out["cfg"].append(cfg)
        out["noise_db"].append(ndb)
        out["seed"].append(seed)
        out["noise_var"].append(nv)
        out["payload"].append(pay)
        out["noise"].append(noise.reshape(-1))
        out["llr_f32"].append(llr)
        out["llr_ref"].append(q_ref)
    np.savez_compressed(
        os.path.join(HERE, "channel_vectors.npz"),
        cfg=np.array(out["cfg"], np.int32), noise_db=np.array(out["noise_db"], np.float32),
        seed=np.array(out["seed"], np.uint32), noise_var=np.array(out["noise_var"], np.float32),
        payload=np.concatenate(out["payload"]), noise=np.concatenate(out["noise"]),
        llr_f32=np.concatenate(out["llr_f32"]), llr_ref=np.concatenate(out["llr_ref"]))
    print("wrote", os.path.join(HERE, "channel_vectors.npz"))


if __name__ == "__main__":
    main()
