"""Writes tests/golden/unrolled_reference.npz from the reference's own
trainable decoder, CTableNeuralLayeredMinSumDecoder
(deep_learning/deeplearning_ldpc_model.py:72-410), imported from
/root/reference with its table loader pointed at the from-source reference
library oracle/_ref/libnrLDPC_ref.so (make -f oracle/Makefile.ref).  Run in
the build container only; the fixtures are data.

Per case (small BG1 / BG2 codes, 3 iterations, batch 3, alpha modes single /
checkpoint / degree, with and without clip_llr and
detach_check_update_inputs; two checkpoint cases whose raw_alpha and
raw_beta are tied per (iteration, row degree)):
  inputs   llr [batch, n] (rate-recovered beliefs / 128 of oracle
           codeblocks: punctured zeros, filler 10), raw_alpha, raw_beta
           (spread so every entry differs), loss weights w [batch, n] and
           wi [T, batch, n];
  forward  final LLR, per-iteration LLRs, loss
           L = sum(w * LLR) + sum_t sum(wi[t] * LLR_t);
  backward raw_alpha, raw_beta and llr gradients of L through the module
           (its alpha()/beta() mappings and sharing mode), and the gradients
           of L with respect to the per-(iteration, row) alpha/beta tables the
           forward used (a checkpoint-mode run on those values).
"""
import os
import sys

import numpy as np
import torch

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, "/root/reference")
sys.path.insert(0, ROOT)
from deep_learning.deeplearning_ldpc_model import (  # noqa: E402
    CTableNeuralLayeredMinSumDecoder, load_ldpc_table_config)
from oracle import oracle as O  # noqa: E402

LIB = os.path.join(ROOT, "oracle", "_ref", "libnrLDPC_ref.so")
MODES = ("single", "checkpoint", "degree")
CASES = [  # cfg, mode, clip, detach, tied
    ((2, 56, 1, 240, 0, 2, 0), "single", None, False, False),
    ((2, 56, 1, 240, 0, 2, 0), "checkpoint", None, False, False),
    ((2, 56, 1, 240, 0, 2, 0), "degree", 8.0, False, False),
    ((2, 56, 1, 240, 0, 2, 0), "checkpoint", 8.0, True, False),
    ((1, 128, 1, 456, 1, 2, 0), "single", 20.0, False, False),
    ((1, 128, 1, 456, 1, 2, 0), "checkpoint", None, True, False),
    ((1, 128, 1, 456, 1, 2, 0), "degree", None, False, False),
    ((1, 128, 1, 456, 1, 2, 0), "degree", 20.0, True, False),
    # checkpoint tables tied per (iteration, row degree): the gradient of a
    # per-(iteration, degree) parameter is the sum over its tied rows
    ((2, 56, 1, 240, 0, 2, 0), "checkpoint", None, False, True),
    ((1, 128, 1, 456, 1, 2, 0), "checkpoint", 20.0, False, True),
]
T, BATCH = 3, 3


def inputs(cfg, seed):
    rng = np.random.default_rng(seed)
    c = O.code(*cfg)
    llr = []
    for k in range(BATCH):
        cb = O.make_codeblock(*cfg[:6], float(rng.uniform(-2.0, 1.0)), rng)
        llr.append(O.rate_recover(c, 1280, cb["llr"]).astype(np.float32) / np.float32(128.0))
    n = c.N + 2 * c.Zc
    w = rng.normal(size=(BATCH, n)).astype(np.float32)
    wi = (rng.normal(size=(T, BATCH, n)) * 0.3).astype(np.float32)
    return np.stack(llr), w, wi


def run(model, llr, w, wi):
    x = torch.tensor(llr, requires_grad=True)
    _, lq, its = model(x, return_llr=True, return_iteration_llrs=True)
    loss = (lq * torch.from_numpy(w)).sum() + sum(
        (its[t] * torch.from_numpy(wi[t])).sum() for t in range(T))
    loss.backward()
    return x, lq, its, loss


def main():
    out = {}
    for k, (cfg, mode, clip, detach, tied) in enumerate(CASES):
        tc = load_ldpc_table_config(*cfg, lib_path=LIB)
        m = CTableNeuralLayeredMinSumDecoder(tc, num_iterations=T, init_alpha=0.7, init_beta=0.0,
                                             alpha_mode=mode, clip_llr=clip,
                                             detach_check_update_inputs=detach)
        with torch.no_grad():
            if tied:
                idx = m.row_alpha_degree_indices
                nd = len(m.unique_row_degrees)
                ra = torch.linspace(-0.6, 0.6, T * nd).reshape(T, nd)
                rb = torch.linspace(0.0, 3.0, T * nd).reshape(T, nd)
                m.raw_alpha.add_(ra[:, idx])
                m.raw_beta.add_(rb[:, idx])
            else:
                m.raw_alpha.add_(torch.linspace(-0.6, 0.6, m.raw_alpha.numel()).reshape(
                    m.raw_alpha.shape))
                m.raw_beta.add_(torch.linspace(0.0, 3.0, m.raw_beta.numel()).reshape(
                    m.raw_beta.shape))
        llr, w, wi = inputs(cfg, 100 + k)
        x, lq, its, loss = run(m, llr, w, wi)
        # the per-(iteration, row) tables this forward used, and the loss
        # gradient with respect to them (checkpoint mode on the same values)
        with torch.no_grad():
            a = m.alpha()
            at = torch.stack([torch.stack([m._alpha_for_row(a, t, r).reshape(())
                                           for r in range(tc.num_rows)]) for t in range(T)])
            bt = m.beta().clone()
        mt = CTableNeuralLayeredMinSumDecoder(tc, num_iterations=T, alpha_mode="checkpoint",
                                              clip_llr=clip, detach_check_update_inputs=detach)
        mt.alpha = lambda: at_leaf        # noqa: E731  the forward reads these tables
        mt.beta = lambda: bt_leaf         # noqa: E731
        at_leaf = at.clone().requires_grad_(True)
        bt_leaf = bt.clone().requires_grad_(True)
        x2, lq2, its2, loss2 = run(mt, llr, w, wi)
        assert torch.equal(lq2, lq) and float(loss2.detach()) == float(loss.detach())
        p = f"c{k}_"
        out[p + "meta"] = np.array(list(cfg) + [T, MODES.index(mode), 0 if clip is None else 1,
                                              int(detach), int(tied)], np.int64)
        out[p + "clip"] = np.float32(0.0 if clip is None else clip)
        out[p + "raw_alpha"] = m.raw_alpha.detach().numpy()
        out[p + "raw_beta"] = m.raw_beta.detach().numpy()
        out[p + "alpha_table"] = at.numpy().astype(np.float32)
        out[p + "beta_table"] = bt.numpy().astype(np.float32)
        out[p + "llr"] = llr
        out[p + "w"] = w
        out[p + "wi"] = wi
        out[p + "lq"] = lq.detach().numpy()
        out[p + "its"] = np.stack([t.detach().numpy() for t in its])
        out[p + "loss"] = np.float64(float(loss.detach()))
        out[p + "g_raw_alpha"] = m.raw_alpha.grad.numpy()
        out[p + "g_raw_beta"] = m.raw_beta.grad.numpy()
        out[p + "g_llr"] = x.grad.numpy()
        out[p + "g_alpha_table"] = at_leaf.grad.numpy()
        out[p + "g_beta_table"] = bt_leaf.grad.numpy()
        print(k, cfg, mode, clip, detach, tied, "loss", float(loss.detach()), "Zc", tc.Zc,
              flush=True)
    out["cases"] = np.int64(len(CASES))
    np.savez_compressed(os.path.join(HERE, "unrolled_reference.npz"), **out)


if __name__ == "__main__":
    main()
