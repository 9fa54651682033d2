"""Pins the oracle of the trainable decoder (oracle/unrolled_oracle.c), CPU only.

* forward: equal, bit for bit, to the reference's FP32 min-sum fixtures
  (c_lib/floatpoint_ldpc_decoder.c through tests/golden/decode_algo4.npz) in
  every case where the two updates coincide (no minimum above the clip, no
  input beyond it), with alpha = alpha/32768, beta = beta/128 and
  clip = max_llr/128;
* gradient: FP64 central finite differences of the FP64 twin (every alpha
  and beta entry and a sample of input beliefs);
* conventions (detach, clipping, punctured zeros and ties, per-iteration
  outputs), which finite differences cannot see: against the reference's
  own module in tests/test_unrolled_reference.py.
"""
import numpy as np
import pytest

from conftest import load_decode_cases



def test_forward_matches_reference_fp32_fixtures(oracle_mod):
    O = oracle_mod
    exact = 0
    for cs in load_decode_cases("decode_algo4.npz"):
        c = O.code(*cs["cfg"])
        row = cs["row"]
        T = cs["iters"]
        llr = cs["soft"].astype(np.float32) / np.float32(128.0)
        a = np.full((T, c.rows), np.float32(row[2] & 0xFFFF) / np.float32(32768.0), np.float32)
        b = np.full((T, c.rows), np.float32(row[3] & 0xFFFF) / np.float32(128.0), np.float32)
        assert row[6] == -row[5]
        clip = float(np.float32(row[5]) / np.float32(128.0))
        lq, _, over = O.unrolled_forward(c, T, llr, a, b, clip=clip, precision="f32",
                                         want_over=True)
        if over[0] == 0 and over[1] == 0:
            assert np.array_equal(lq, cs["lq"]), cs["idx"]
            exact += 1
    assert exact >= 8


CASES = [
    # (bgn, A, crctype, E), iterations
    ((2, 40, 0, 200), 3),     # BG2 Zc=10
    ((1, 104, 0, 390), 2),    # BG1 Zc=6
]


def _inputs(c, T, seed, clip):
    rng = np.random.default_rng(seed)
    n = c.N + 2 * c.Zc
    llr = rng.normal(0.0, 3.0, n)
    llr[np.abs(llr) < 1e-3] = 0.5          # no exact zeros (|x| kink)
    alpha = rng.uniform(0.55, 0.95, (T, c.rows))
    beta = rng.uniform(0.05, 0.4, (T, c.rows))
    w = rng.normal(size=n)
    wi = rng.normal(size=(T, n)) * 0.3
    return llr, alpha, beta, w, wi


@pytest.mark.parametrize("cfg,T", CASES)
@pytest.mark.parametrize("clip", [None, 6.0])
def test_gradient_matches_fd64(oracle_mod, cfg, T, clip):
    O = oracle_mod
    c = O.code(*cfg)
    llr, alpha, beta, w, wi = _inputs(c, T, 7, clip)

    def loss(a, b, x):
        q, it = O.unrolled_forward(c, T, x, a, b, clip=clip, precision="f64", want_iters=True)
        return float(w @ q + np.sum(wi * it))

    g_llr, ga, gb = O.unrolled_backward(c, T, llr, alpha, beta, w, wi, clip=clip,
                                        precision="f64")
    h = 1e-6
    fa = np.zeros_like(ga)
    fb = np.zeros_like(gb)
    for t in range(T):
        for r in range(c.rows):
            p = alpha.copy(); p[t, r] += h
            m = alpha.copy(); m[t, r] -= h
            fa[t, r] = (loss(p, beta, llr) - loss(m, beta, llr)) / (2 * h)
            p = beta.copy(); p[t, r] += h
            m = beta.copy(); m[t, r] -= h
            fb[t, r] = (loss(alpha, p, llr) - loss(alpha, m, llr)) / (2 * h)
    rng = np.random.default_rng(3)
    idx = rng.choice(len(llr), 40, replace=False)
    fl = np.zeros(len(idx))
    for k, i in enumerate(idx):
        p = llr.copy(); p[i] += h
        m = llr.copy(); m[i] -= h
        fl[k] = (loss(alpha, beta, p) - loss(alpha, beta, m)) / (2 * h)
    for an, fd in ((ga, fa), (gb, fb), (g_llr[idx], fl)):
        rel = np.linalg.norm(an - fd) / max(np.linalg.norm(fd), 1e-12)
        assert rel < 1e-6, rel
    assert np.abs(ga).max() > 0 and np.abs(gb).max() > 0


@pytest.mark.parametrize("cfg,T", CASES)
def test_fp32_gradient_close_to_fp64(oracle_mod, cfg, T):
    O = oracle_mod
    c = O.code(*cfg)
    llr, alpha, beta, w, wi = _inputs(c, T, 5, None)
    args32 = [v.astype(np.float32) for v in (llr, alpha, beta, w)]
    args64 = [v.astype(np.float64) for v in args32]
    g32 = O.unrolled_backward(c, T, *args32[:3], args32[3], precision="f32")
    g64 = O.unrolled_backward(c, T, *args64[:3], args64[3], precision="f64")
    for x32, x64 in zip(g32, g64):
        assert np.linalg.norm(x32 - x64) / np.linalg.norm(x64) < 1e-4
