"""Synthetic workloads for batched decoding.

``mixed_batch`` builds the mixed-size batch of one persistent launch:
base graph 1/2 half and half, lifting size uniform over the 51 sizes of
TS 38.212 Table 5.3.2-1 (resampled where a base graph cannot reach it), code
rate uniform in [0.2, 0.9], Qm in {2, 4, 6, 8}, and a per-codeblock Es/N0
drawn within +-1 dB of the rate's waterfall.  The waterfall is placed at the
BICM Shannon limit of the rate, 10*log10(2^(R*Qm) - 1) dB, plus a finite-length
gap of 1.5 dB.
"""
import ctypes

import numpy as np

from . import _lib
from ._lib import LdpcConfig, LdpcInConfig

LIFTING_SIZES = (2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13, 14, 15, 16, 18, 20, 22, 24, 26, 28,
                 30, 32, 36, 40, 44, 48, 52, 56, 60, 64, 72, 80, 88, 96, 104, 112, 120, 128,
                 144, 160, 176, 192, 208, 224, 240, 256, 288, 320, 352, 384)


def _params(bgn, A, crctype):
    pc = LdpcConfig()
    st = _lib.load().gen_ldpc_parameters(ctypes.byref(pc), ctypes.byref(
        LdpcInConfig(bgn, A, crctype, 64 * 8, 0, 2, 0)))
    return (pc.Zc, pc.B) if st == 0 else (None, None)


_PAYLOAD = {}


def payload_for(bgn, zc, crctype=2):
    """Largest payload A (multiple of 8) whose code uses lifting size zc, or
    None when the base graph cannot reach zc with this CRC."""
    key = (bgn, zc, crctype)
    if key not in _PAYLOAD:
        clen = 16 if crctype == 0 else 24
        kb = 22 if bgn == 1 else 10
        found = None
        for B in range((kb * zc) // 8 * 8, clen + 7, -8):
            z, _ = _params(bgn, B - clen, crctype)
            if z == zc:
                found = B - clen
                break
        _PAYLOAD[key] = found
    return _PAYLOAD[key]


def waterfall_esn0_db(rate, qm, B=8448):
    """Es/N0 (dB) near the 10 % BLER point of the int16 min-sum decoder with
    8 iterations: BICM Shannon limit plus a gap fitted to oracle decodes
    (1.8-2.5 dB for long QPSK codes, more for short blocks, high rates and
    high-order QAM)."""
    gap = 1.5 + 0.3 * (qm - 2) + 8.0 / np.sqrt(B) + 1.7 * max(rate - 0.5, 0.0)
    return 10.0 * np.log10(2.0 ** (rate * qm) - 1.0) + gap


def draw_config(rng):
    """One codeblock config (bgn, A, crctype, E, rv, Qm, 0) and its Es/N0."""
    while True:
        bgn = int(rng.integers(1, 3))
        zc = int(rng.choice(LIFTING_SIZES))
        crc = 2
        A = payload_for(bgn, zc, crc)
        if A is None:
            crc = 0
            A = payload_for(bgn, zc, crc)
        if A is None:
            continue            # unreachable for this base graph: resample
        B = A + (16 if crc == 0 else 24)
        qm = int(rng.choice([2, 4, 6, 8]))
        rate = float(rng.uniform(0.2, 0.9))
        E = int(B / rate)
        E -= E % qm
        if E < B or E > 65535:
            continue
        esn0 = waterfall_esn0_db(B / E, qm, B) + float(rng.uniform(-1.0, 1.0))
        return (bgn, A, crc, E, 0, qm, 0), esn0


def mixed_batch(n, seed=0, threads=8):
    """n codeblocks of mixed size.  Returns (cfgs, llr list, payload list)."""
    from concurrent.futures import ThreadPoolExecutor

    from .batch import synth_batch
    rng = np.random.default_rng(seed)
    draws = [draw_config(rng) for _ in range(n)]

    def one(i):   # the library call releases the GIL
        cfg, esn0 = draws[i]
        llr, pay = synth_batch(cfg, 1, -esn0, seed=seed * 1000003 + i, threads=1)
        return llr[0], pay[0]

    with ThreadPoolExecutor(max_workers=max(1, threads)) as ex:
        out = list(ex.map(one, range(n)))
    return [d[0] for d in draws], [o[0] for o in out], [o[1] for o in out]
