"""GPU parity of the trainable decoder (ldpc_mi355_unrolled_*, HIP kernels in
csrc/ldpc_unrolled.hip) through the C ABI, against the oracle
(oracle/unrolled_oracle.c) and, for the gradient, FP64 finite differences:

* forward beliefs, per-iteration beliefs and dL/dllr equal to the oracle's
  FP32 restatement bit for bit; dL/dalpha and dL/dbeta (float partial sums,
  double reduction) within 1e-5 of the oracle's;
* the gradient within 1e-4 (relative, norm-wise) of FP64 central finite
  differences of the FP64 oracle;
* the torch module (CTableNeuralLayeredMinSumDecoder) in every sharing mode
  against the reference's own module (fixtures from
  tests/golden/gen_unrolled_reference.py): LLRs, loss and raw_alpha /
  raw_beta / input gradients; the new iteration_degree modes against the
  reference's checkpoint mode with tied tables;
* a short training run of the BG1 Zc=96 workload lowers the loss.
"""
import os
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


def _need_gpu():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _inputs(c, T, batch, seed, scale=3.0):
    rng = np.random.default_rng(seed)
    n = c.N + 2 * c.Zc
    llr = rng.normal(0.0, scale, (batch, n)).astype(np.float32)
    llr[:, :2 * c.Zc] = 0.0
    alpha = rng.uniform(0.55, 0.95, (T, c.rows)).astype(np.float32)
    beta = rng.uniform(0.0, 0.3, (T, c.rows)).astype(np.float32)
    g = rng.normal(size=(batch, n)).astype(np.float32)
    gi = (rng.normal(size=(T, batch, n)) * 0.3).astype(np.float32)
    return llr, alpha, beta, g, gi


def _kernel(cfg, T, clip, detach, cbs=0):
    from deeplearning_5g_ldpc_decoder_amd.unrolled import UnrolledKernel, load_ldpc_table_config
    tc = load_ldpc_table_config(*cfg, 0, 2)
    k = UnrolledKernel(tc, T, clip, detach)
    if cbs:
        k.set_codeblocks_per_workgroup(cbs)
    return tc, k


def _gpu_run(k, llr, alpha, beta, g, gi):
    d = lambda x: torch.from_numpy(np.ascontiguousarray(x)).cuda()  # noqa: E731
    x, a, b = d(llr), d(alpha), d(beta)
    lq, it, rec = k.forward(x, a, b, want_iters=True)
    g_llr, ga, gb = k.backward(x, a, b, rec, d(g), d(gi))
    torch.cuda.synchronize()
    return {n: v.cpu().numpy() for n, v in
            dict(lq=lq, it=it, g_llr=g_llr, ga=ga, gb=gb).items()}


PARITY = [
    # cfg (bgn, A, crctype, E), T, clip, detach, batch, codeblocks per workgroup
    ((1, 2088, 2, 6336), 10, None, False, 6, 0),          # BG1 Zc=96, the training shape
    ((1, 2088, 2, 6336), 10, 32767 / 128, True, 5, 0),    # reference training options
    ((1, 2088, 2, 6336), 4, 6.0, False, 5, 1),
    ((2, 1016, 2, 3120), 5, 20.0, False, 3, 0),           # BG2 Zc=104
    ((1, 8424, 2, 25344), 2, None, False, 2, 0),          # BG1 Zc=384
    ((2, 40, 0, 200), 3, 4.0, False, 50, 0),              # BG2 Zc=10, 24 per workgroup
    ((1, 104, 0, 390), 3, None, True, 70, 5),             # BG1 Zc=6
]


@pytest.mark.parametrize("cfg,T,clip,detach,batch,cbs", PARITY)
def test_unrolled_matches_oracle(oracle_mod, cfg, T, clip, detach, batch, cbs):
    _need_gpu()
    O = oracle_mod
    _say(f"unrolled parity cfg={cfg} T={T} clip={clip} detach={detach} batch={batch} cbs={cbs}")
    c = O.code(*cfg)
    tc, k = _kernel(cfg, T, clip, detach, cbs)
    llr, alpha, beta, g, gi = _inputs(c, T, batch, seed=batch + T)
    r = _gpu_run(k, llr, alpha, beta, g, gi)
    ga = np.zeros((T, c.rows))
    gb = np.zeros((T, c.rows))
    for i in range(batch):
        q, it = O.unrolled_forward(c, T, llr[i], alpha, beta, clip=clip, want_iters=True)
        assert np.array_equal(q, r["lq"][i]), i
        assert np.array_equal(it, r["it"][:, i]), i
        gl, a_i, b_i = O.unrolled_backward(c, T, llr[i], alpha, beta, g[i], gi[:, i], clip=clip,
                                           detach=detach)
        assert np.array_equal(gl, r["g_llr"][i]), i
        ga += a_i
        gb += b_i
    for mine, ref in ((r["ga"], ga), (r["gb"], gb)):
        assert np.abs(mine - ref).max() <= 1e-5 * max(np.abs(ref).max(), 1e-3)


def test_unrolled_codeblocks_per_workgroup_invariant(oracle_mod):
    _need_gpu()
    cfg, T = (2, 1016, 2, 3120), 3
    c = oracle_mod.code(*cfg)
    llr, alpha, beta, g, gi = _inputs(c, T, 7, seed=1)
    ref = None
    for cbs in (1, 2, 3):
        _say(f"unrolled cbs={cbs}")
        _, k = _kernel(cfg, T, 10.0, False, cbs)
        r = _gpu_run(k, llr, alpha, beta, g, gi)
        if ref is None:
            ref = r
            continue
        for name in ("lq", "it", "g_llr"):
            assert np.array_equal(r[name], ref[name]), (cbs, name)
        for name in ("ga", "gb"):
            assert np.allclose(r[name], ref[name], rtol=1e-5, atol=1e-5), (cbs, name)


@pytest.mark.parametrize("cfg,T,clip", [((2, 40, 0, 200), 3, None), ((2, 40, 0, 200), 3, 6.0),
                                        ((1, 104, 0, 390), 2, None)])
def test_unrolled_gradient_matches_fd64(oracle_mod, cfg, T, clip):
    """GPU gradient (FP32) against central differences of the FP64 oracle."""
    _need_gpu()
    O = oracle_mod
    _say(f"unrolled fd64 cfg={cfg} T={T} clip={clip}")
    c = O.code(*cfg)
    rng = np.random.default_rng(5)
    n = c.N + 2 * c.Zc
    llr = rng.normal(0.0, 3.0, (1, n)).astype(np.float32)
    llr[np.abs(llr) < 1e-3] = 0.5
    alpha = rng.uniform(0.55, 0.95, (T, c.rows)).astype(np.float32)
    beta = rng.uniform(0.05, 0.3, (T, c.rows)).astype(np.float32)
    w = rng.normal(size=(1, n)).astype(np.float32)
    wi = np.zeros((T, 1, n), np.float32)
    _, k = _kernel(cfg, T, clip, False)
    r = _gpu_run(k, llr, alpha, beta, w, wi)
    x64, a64, b64, w64 = (v.astype(np.float64) for v in (llr[0], alpha, beta, w[0]))

    def loss(a, b, x):
        q, _ = O.unrolled_forward(c, T, x, a, b, clip=clip, precision="f64")
        return float(w64 @ q)

    h = 1e-6
    fa = np.zeros((T, c.rows))
    fb = np.zeros((T, c.rows))
    for t in range(T):
        for row in range(c.rows):
            p = a64.copy(); p[t, row] += h
            m = a64.copy(); m[t, row] -= h
            fa[t, row] = (loss(p, b64, x64) - loss(m, b64, x64)) / (2 * h)
            p = b64.copy(); p[t, row] += h
            m = b64.copy(); m[t, row] -= h
            fb[t, row] = (loss(a64, p, x64) - loss(a64, m, x64)) / (2 * h)
    idx = rng.choice(n, 40, replace=False)
    fl = np.zeros(len(idx))
    for j, i in enumerate(idx):
        p = x64.copy(); p[i] += h
        m = x64.copy(); m[i] -= h
        fl[j] = (loss(a64, b64, p) - loss(a64, b64, m)) / (2 * h)
    for an, fd in ((r["ga"], fa), (r["gb"], fb), (r["g_llr"][0][idx], fl)):
        rel = np.linalg.norm(an - fd) / max(np.linalg.norm(fd), 1e-12)
        assert rel < 1e-4, rel


def _reference_case(k):
    from conftest import GOLDEN
    fx = np.load(os.path.join(GOLDEN, "unrolled_reference.npz"))
    p = f"c{k}_"
    return {key[len(p):]: fx[key] for key in fx.files if key.startswith(p)}


def _module_vs_reference(k, alpha_mode, beta_mode, raw_alpha, raw_beta):
    """Runs the GPU module on fixture case k with the given raw parameters;
    returns (fixture, module, final LLR, per-iteration LLRs, loss, input)."""
    from deeplearning_5g_ldpc_decoder_amd.unrolled import (CTableNeuralLayeredMinSumDecoder,
                                                           load_ldpc_table_config)
    f = _reference_case(k)
    meta = [int(v) for v in f["meta"]]
    cfg, T = tuple(meta[:7]), meta[7]
    clip = float(f["clip"]) if meta[9] else None
    tc = load_ldpc_table_config(*cfg)
    m = CTableNeuralLayeredMinSumDecoder(tc, num_iterations=T, alpha_mode=alpha_mode,
                                         beta_mode=beta_mode, clip_llr=clip,
                                         detach_check_update_inputs=bool(meta[10]))
    with torch.no_grad():
        m.raw_alpha.copy_(torch.from_numpy(np.asarray(raw_alpha, np.float32)))
        m.raw_beta.copy_(torch.from_numpy(np.asarray(raw_beta, np.float32)))
    m = m.cuda()
    x = torch.from_numpy(f["llr"]).cuda().requires_grad_(True)
    _, q, its = m(x, return_llr=True, return_iteration_llrs=True)
    loss = (q * torch.from_numpy(f["w"]).cuda()).sum() + sum(
        (its[t] * torch.from_numpy(f["wi"][t]).cuda()).sum() for t in range(T))
    loss.backward()
    return f, m, q, its, loss, x


def _close(a, b, rel):
    return np.abs(a - b).max() <= rel * max(np.abs(b).max(), 1e-6)


REF_CASES = [0, 1, 2, 3, 4, 5, 6, 7]
MODE_NAMES = ("single", "checkpoint", "degree")


@pytest.mark.parametrize("k", REF_CASES)
def test_module_matches_reference_model(k):
    """The GPU module against the reference's own trainable decoder
    (deep_learning/deeplearning_ldpc_model.py, fixtures written by
    tests/golden/gen_unrolled_reference.py) in its sharing modes, with and
    without clip_llr and detach: the same raw parameters and inputs give the
    reference's final and per-iteration LLRs, loss, and raw_alpha / raw_beta
    / input-LLR gradients.  The device sigmoid may round an alpha or beta in
    the last bit differently from the CPU's, hence the tolerances."""
    _need_gpu()
    f0 = _reference_case(k)
    mode = MODE_NAMES[int(f0["meta"][8])]
    _say(f"reference model case {k}: mode {mode} meta {list(f0['meta'])}")
    f, m, q, its, loss, x = _module_vs_reference(k, mode, "checkpoint", f0["raw_alpha"],
                                                 f0["raw_beta"])
    assert np.allclose(q.detach().cpu().numpy(), f["lq"], rtol=1e-5, atol=1e-4)
    for t in range(len(its)):
        assert np.allclose(its[t].detach().cpu().numpy(), f["its"][t], rtol=1e-5, atol=1e-4)
    assert abs(float(loss.detach()) - float(f["loss"])) <= 1e-5 * max(abs(float(f["loss"])), 1.0)
    assert _close(m.raw_alpha.grad.cpu().numpy(), f["g_raw_alpha"], 1e-4)
    assert _close(m.raw_beta.grad.cpu().numpy(), f["g_raw_beta"], 1e-4)
    assert _close(x.grad.cpu().numpy(), f["g_llr"], 1e-4)


@pytest.mark.parametrize("k", [8, 9])
def test_iteration_degree_mode_matches_reference_sums(k):
    """The iteration_degree modes (one alpha and one beta per iteration and
    row degree; not in the reference) against the reference in checkpoint
    mode with its tables tied per (iteration, degree): same LLRs, and each
    per-(iteration, degree) gradient is the sum of the tied rows'
    gradients."""
    _need_gpu()
    f0 = _reference_case(k)
    assert int(f0["meta"][11]) == 1
    from deeplearning_5g_ldpc_decoder_amd.unrolled import (degree_alpha_indices,
                                                           load_ldpc_table_config, row_degrees)
    tc = load_ldpc_table_config(*[int(v) for v in f0["meta"][:7]])
    _, idx = degree_alpha_indices(row_degrees(tc))
    idx = np.asarray(idx)
    nd = int(idx.max()) + 1
    first = [int(np.flatnonzero(idx == d)[0]) for d in range(nd)]
    ra = f0["raw_alpha"][:, first]
    rb = f0["raw_beta"][:, first]
    assert np.array_equal(ra[:, idx], f0["raw_alpha"]) and np.array_equal(rb[:, idx], f0["raw_beta"])
    f, m, q, its, loss, x = _module_vs_reference(k, "iteration_degree", "iteration_degree", ra, rb)
    assert np.allclose(q.detach().cpu().numpy(), f["lq"], rtol=1e-5, atol=1e-4)
    want_a = np.stack([f["g_raw_alpha"][:, idx == d].sum(axis=1) for d in range(nd)], axis=1)
    want_b = np.stack([f["g_raw_beta"][:, idx == d].sum(axis=1) for d in range(nd)], axis=1)
    assert _close(m.raw_alpha.grad.cpu().numpy(), want_a, 1e-4)
    assert _close(m.raw_beta.grad.cpu().numpy(), want_b, 1e-4)
    assert _close(x.grad.cpu().numpy(), f["g_llr"], 1e-4)


def test_edge_cases(oracle_mod):
    """Batch of zero (parameter gradients zero) and zero iterations."""
    _need_gpu()
    _say("unrolled edge cases")
    cfg = (2, 40, 0, 200)
    c = oracle_mod.code(*cfg)
    _, k = _kernel(cfg, 2, 5.0, False)
    a = torch.full((2, c.rows), 0.7, device="cuda")
    b = torch.full((2, c.rows), 0.1, device="cuda")
    x = torch.zeros((0, c.N + 2 * c.Zc), device="cuda")
    lq, _, rec = k.forward(x, a, b)
    g_llr, ga, gb = k.backward(x, a, b, rec, torch.zeros_like(x))
    torch.cuda.synchronize()
    assert lq.shape == x.shape and float(ga.abs().sum()) == 0 and float(gb.abs().sum()) == 0
    _, k0 = _kernel(cfg, 0, 5.0, False)
    llr = torch.linspace(-9, 9, 3 * (c.N + 2 * c.Zc), device="cuda").reshape(3, -1)
    e = torch.zeros((0, c.rows), device="cuda")
    lq, _, rec = k0.forward(llr, e, e)
    g_llr, _, _ = k0.backward(llr, e, e, rec, torch.ones_like(llr))
    torch.cuda.synchronize()
    assert torch.equal(lq, llr.clamp(-5, 5))
    assert torch.equal(g_llr, (llr.abs() <= 5).float())


def test_training_lowers_loss_and_exports(tmp_path):
    """A short run of the BG1 Zc=96 workload (batch 256) on the GPU."""
    _need_gpu()
    from deeplearning_5g_ldpc_decoder_amd.trainer import (BG1_ZC96_WORKLOAD,
                                                          train_c_table_ldpc_model)
    from deeplearning_5g_ldpc_decoder_amd.unrolled import algo2_tables, export_algo2_tables
    _say("training 30 steps, BG1 Zc=96, batch 256")
    cfg = dict(BG1_ZC96_WORKLOAD)
    cfg.update(epochs=30, batch_size=256, lr=2e-2, save_path=None, verbose=False,
               noise_power_db=-1.0)
    hist = []
    model, _, best = train_c_table_ldpc_model(history=hist, **cfg)
    first = np.mean([h["loss"] for h in hist[:5]])
    last = np.mean([h["loss"] for h in hist[-5:]])
    _say(f"loss first5={first:.5f} last5={last:.5f} best={best:.5f}")
    assert last < first
    fa, fb = algo2_tables(model)
    assert fa.shape == (32, 46) and fa.dtype == np.int16 and fb.shape == (32, 46)
    h, c = export_algo2_tables(model, str(tmp_path))
    text = open(c).read()
    assert "BG1_trained_alpha_list" in text and "BG1_trained_beta_list" in text
