"""Device input synthesis (ldpc_mi355_synth_device) on the GPU: payload+CRC,
codeword and rate-matched bits equal the oracle's encoder and rate matcher
(pinned by the reference encoder fixtures); channel LLRs equal the numpy
restatement of the kernel's random streams and max-log demapper to float
rounding; int16 quantization and the float rate recovery of the trainable
decoder's inputs are exact; and device-synthesized codeblocks decode with the
BLER of host-synthesized ones at the same SNR."""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

CASES = [(1, 8424, 2, 25344, 0, 2, 0), (2, 1016, 2, 2080, 2, 4, 0), (1, 488, 1, 1536, 1, 1, 0),
         (2, 16, 0, 160, 3, 2, 0), (2, 40, 0, 600, 0, 6, 0), (1, 2088, 2, 6336, 0, 2, 0),
         (2, 3816, 1, 7632, 0, 8, 0)]


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("cfg", CASES)
def test_synth_device_matches_oracle(oracle_mod, cfg):
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    sys.path.insert(0, __file__.rsplit("/", 1)[0])
    import synth_ref
    from deeplearning_5g_ldpc_decoder_amd.batch import synth_device
    O = oracle_mod
    _say(f"synth_device cfg={cfg}")
    c = O.code(*cfg)
    st, p = O.params(*cfg)
    p.update(E=cfg[3], Qm=cfg[5])
    n = 3
    pay = np.random.default_rng(cfg[1]).integers(0, 2, (n, cfg[1])).astype(np.uint8)
    out = synth_device(cfg, n, -1.0, seed=9, payload=pay,
                       outputs=("bits", "cw", "rm", "llr", "llr_f32", "beliefs"))
    torch.cuda.synchronize()
    h = {k: v.cpu().numpy() for k, v in out.items() if k != "_keep"}
    for i in range(n):
        b = O.attach_crc(pay[i], cfg[2])
        assert np.array_equal(h["bits"][i], b)
        cw = O.encode(c, b)
        assert np.array_equal(h["cw"][i], cw)
        assert np.array_equal(h["rm"][i], O.rate_match(c, cw))
        ref = synth_ref.channel_llr(h["rm"][i], cfg[5], -1.0, 9, i)
        assert np.allclose(h["llr_f32"][i], ref, rtol=1e-4, atol=1e-4)
        assert np.array_equal(h["llr"][i], synth_ref.quantize(h["llr_f32"][i]))
        assert np.array_equal(h["beliefs"][i], synth_ref.float_rate_recover(h["llr_f32"][i], p))
    gen = synth_device(cfg, 2, -1.0, seed=5, outputs=("bits", "cw"))
    torch.cuda.synchronize()
    for i in range(2):
        bits = gen["bits"][i].cpu().numpy()
        assert np.array_equal(bits[:cfg[1]], synth_ref.payload_bits(5, i, cfg[1]))
        assert np.array_equal(bits, O.attach_crc(bits[:cfg[1]], cfg[2]))
        assert O.syndrome_ok(c, gen["cw"][i].cpu().numpy())


def test_synth_device_bler_matches_host_synthesis():
    """Same config and SNR on the waterfall: BLER of device-synthesized
    inputs inside the 99.9 % interval of host-synthesized ones."""
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from deeplearning_5g_ldpc_decoder_amd.batch import BatchPlan, synth_batch, synth_device
    from deeplearning_5g_ldpc_decoder_amd.ldpc_sim import bler_confidence_interval
    cfg = (2, 1016, 2, 2080, 0, 4, 0)
    row = [1, 8, 22936, 0, 1280, 32767, -32767]
    noise = -6.8
    n = 4096
    plan = BatchPlan([cfg] * n, row)
    d = synth_device(cfg, n, noise, seed=3)
    st_dev = plan.decode(d["llr"].reshape(-1))["status"].cpu().numpy()
    llr, _ = synth_batch(cfg, n, noise, seed=4)
    st_host = plan.decode(torch.from_numpy(llr.reshape(-1)).cuda())["status"].cpu().numpy()
    plan.close()
    e1, e2 = int((st_dev != 0).sum()), int((st_host != 0).sum())
    lo1, hi1 = bler_confidence_interval(e1, n, z=3.29)
    lo2, hi2 = bler_confidence_interval(e2, n, z=3.29)
    _say(f"BLER device-synthesized {e1 / n:.4f}, host-synthesized {e2 / n:.4f}")
    assert 0.02 < e2 / n < 0.98
    assert lo1 <= hi2 and lo2 <= hi1


def _channel_cases():
    sys.path.insert(0, __file__.rsplit("/", 1)[0])
    from test_channel import CASES
    return CASES


@pytest.mark.parametrize("case", _channel_cases(), ids=lambda c: f"Qm{c['cfg'][5]}-E{c['cfg'][3]}")
def test_synth_device_given_noise_matches_reference(oracle_mod, case):
    """ldpc_mi355_synth_device_noise on the GPU, given the noise the
    reference's tx_rx_process drew after srand(seed): its int16 LLRs are the
    reference's (tests/golden/channel_vectors.npz, written from the
    from-source reference) and its float LLRs the oracle's, bit for bit."""
    from deeplearning_5g_ldpc_decoder_amd.batch import synth_device
    cfg = case["cfg"]
    _say(f"synth_device_noise cfg={cfg}")
    out = synth_device(cfg, 1, 0.0, payload=case["payload"][None, :], outputs=("rm", "llr", "llr_f32"),
                       noise=(np.array([case["noise_var"]]), case["noise"]))
    torch.cuda.synchronize()
    c = oracle_mod.code(*cfg)
    rm = oracle_mod.rate_match(c, oracle_mod.encode(c, oracle_mod.attach_crc(case["payload"], cfg[2])))
    assert np.array_equal(out["rm"][0].cpu().numpy(), rm)
    assert np.array_equal(out["llr"][0].cpu().numpy(), case["llr_ref"])
    assert np.array_equal(out["llr_f32"][0].cpu().numpy(), case["llr_f32"])


def test_device_noise_statistics():
    """The device's own noise (counter-based generator, Box-Muller) against
    the closed-form statistics of complex AWGN with variance noise_var / 2
    per dimension, recovered exactly from QPSK float LLRs
    (llr = 4 * A * y / noise_var): mean, variance, kurtosis, a
    Kolmogorov-Smirnov distance to N(0, noise_var / 2), and no correlation
    between real and imaginary parts, neighbouring symbols or codeblocks.
    Bounds are 6 standard errors (KS: the 1e-6 critical value)."""
    from scipy import stats

    from deeplearning_5g_ldpc_decoder_amd.batch import synth_device
    cfg = (1, 8424, 2, 25344, 0, 2, 0)
    n, ndb = 16, 1.5
    _say(f"device noise statistics cfg={cfg} n={n}")
    out = synth_device(cfg, n, ndb, seed=77, outputs=("rm", "llr_f32"))
    torch.cuda.synchronize()
    rm = out["rm"].cpu().numpy().astype(np.float64)
    llr = out["llr_f32"].cpu().numpy().astype(np.float64)
    nv = float(np.float32(10.0) ** np.float32(ndb / 10.0))
    A = 1.0 / np.sqrt(2.0)
    noise = llr * nv / (4.0 * A) - (1.0 - 2.0 * rm) * A      # [n, E]: re, im alternate
    sd = np.sqrt(nv / 2.0)
    x = noise.reshape(-1) / sd
    N = x.size
    assert abs(x.mean()) < 6.0 / np.sqrt(N)
    assert abs(x.var() - 1.0) < 6.0 * np.sqrt(2.0 / N)
    assert abs(stats.kurtosis(x, fisher=False) - 3.0) < 6.0 * np.sqrt(24.0 / N)
    assert stats.kstest(x, "norm").statistic < 2.63 / np.sqrt(N)
    re, im = noise[:, 0::2] / sd, noise[:, 1::2] / sd
    lim = 6.0 / np.sqrt(re.size)
    assert abs(np.mean(re * im)) < lim
    assert abs(np.mean(re[:, 1:] * re[:, :-1])) < 6.0 / np.sqrt(re[:, 1:].size)
    assert abs(np.mean(re[1:] * re[:-1])) < 6.0 / np.sqrt(re[1:].size)
