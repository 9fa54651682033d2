"""The test channel (c_lib/tx_rx_process.c) against the reference itself.

tests/golden/channel_vectors.npz holds, per case, the int16 LLRs of the
from-source reference's tx_rx_process after srand(seed) on rate-matched
codeword bits, and the noise values that call added (tests/golden/
gen_channel.py).  On the CPU:
  * the oracle's restatement (orc_tx_rx: glibc rand() Box-Muller noise, Gray
    QAM, the piecewise max-log LLRs in the reference's float order, x128
    quantization) reproduces every case: noise, variance, float and int16
    LLRs;
  * the CPU build of the device synthesis kernel, given that noise
    (ldpc_mi355_synth_device_noise's path), gives the reference's int16 LLRs
    and the oracle's float LLRs bit for bit;
  * negative control: the same max-log LLRs evaluated as (d1^2 - d0^2) /
    noise_var over the levels -- mathematically equal, rounded differently --
    miss the reference's int16 LLRs on some cases, so the check is sensitive
    to the operation order."""
import os
import sys

import numpy as np
import pytest

from conftest import GOLDEN, ROOT

sys.path.insert(0, os.path.join(ROOT, "tests", "kernel_emu"))
sys.path.insert(0, os.path.join(ROOT, "tests"))


def channel_cases():
    z = np.load(os.path.join(GOLDEN, "channel_vectors.npz"))
    out, op, on, ol = [], 0, 0, 0
    for i, cfg in enumerate(z["cfg"]):
        cfg = tuple(int(v) for v in cfg)
        A, E, Qm = cfg[1], cfg[3], cfg[5]
        S = E // Qm
        out.append(dict(cfg=cfg, noise_db=float(z["noise_db"][i]), seed=int(z["seed"][i]),
                        noise_var=np.float32(z["noise_var"][i]),
                        payload=z["payload"][op:op + A], noise=z["noise"][on:on + 2 * S],
                        llr_f32=z["llr_f32"][ol:ol + E], llr_ref=z["llr_ref"][ol:ol + E]))
        op, on, ol = op + A, on + 2 * S, ol + E
    return out


CASES = channel_cases()


def _rm_bits(O, case):
    c = O.code(*case["cfg"])
    return O.rate_match(c, O.encode(c, O.attach_crc(case["payload"], case["cfg"][2])))


@pytest.mark.parametrize("case", CASES, ids=lambda c: f"Qm{c['cfg'][5]}-E{c['cfg'][3]}")
def test_oracle_channel_matches_reference(oracle_mod, case):
    rm = _rm_bits(oracle_mod, case)
    noise, nv, llr, q = oracle_mod.tx_rx(rm, case["cfg"][5], case["noise_db"], case["seed"])
    assert np.array_equal(q, case["llr_ref"])
    assert np.float32(nv) == case["noise_var"]
    assert np.array_equal(noise.reshape(-1), case["noise"])
    assert np.array_equal(llr, case["llr_f32"])


@pytest.mark.parametrize("case", CASES, ids=lambda c: f"Qm{c['cfg'][5]}-E{c['cfg'][3]}")
def test_emu_synth_given_noise_matches_reference(oracle_mod, case):
    import emu
    cfg = case["cfg"]
    out = emu.synth(cfg[:6], 1, 0.0, payload=case["payload"][None, :],
                    noise=(np.array([case["noise_var"]]), case["noise"]))
    assert np.array_equal(out["rm"][0], _rm_bits(oracle_mod, case))
    assert np.array_equal(out["llr"][0], case["llr_ref"])
    assert np.array_equal(out["llr_f32"][0], case["llr_f32"])


def _pam(bits):
    s = [1 - 2 * int(b) for b in bits]
    lvl = s[-1]
    for m in range(len(s) - 2, -1, -1):
        lvl = s[m] * ((1 << (len(s) - 1 - m)) - lvl)
    return lvl


def test_distance_form_demapper_misses_reference(oracle_mod):
    f32 = np.float32
    misses = 0
    for case in CASES:
        Qm = case["cfg"][5]
        if Qm < 4:
            continue
        nb, S = Qm // 2, case["cfg"][3] // Qm
        rm = _rm_bits(oracle_mod, case)
        noise = case["noise"].reshape(S, 2)
        amp = f32(1) / np.sqrt(f32({4: 10, 6: 42, 8: 170}[Qm]))
        nv = case["noise_var"]
        levels = [(v, [(v >> (nb - 1 - m)) & 1 for m in range(nb)]) for v in range(1 << nb)]
        for j in range(S):
            for ax in range(2):
                r = f32(f32(_pam(rm[j * Qm + ax:(j + 1) * Qm:2])) * amp + noise[j, ax])
                for k in range(nb):
                    d0 = d1 = f32(3e38)
                    for _, lb in levels:
                        d = f32(r - f32(_pam(lb)) * amp)
                        if lb[k]:
                            d1 = min(d1, f32(d * d))
                        else:
                            d0 = min(d0, f32(d * d))
                    sc = f32(f32(f32(d1 - d0) / nv) * f32(128))
                    q = int(f32(sc + f32(0.5)) if sc >= 0 else f32(sc - f32(0.5)))
                    q = max(-10000, min(10000, q))
                    misses += q != case["llr_ref"][j * Qm + 2 * k + ax]
    assert misses > 0
