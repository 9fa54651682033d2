"""The oracle's unrolled FP32 min-sum (oracle/unrolled_oracle.c) against the
reference's own trainable decoder (deep_learning/deeplearning_ldpc_model.py:
315-410, run by tests/golden/gen_unrolled_reference.py into
unrolled_reference.npz): on the per-(iteration, row) alpha/beta tables the
reference forward used, the final and per-iteration LLRs are bit-exact, and
the gradients of the fixture loss with respect to the tables and the input
LLRs agree to 1e-5 (the reference sums in FP32 in autograd's order, the
oracle per codeblock in double)."""
import os

import numpy as np
import pytest

from conftest import GOLDEN

FX = np.load(os.path.join(GOLDEN, "unrolled_reference.npz"))
CASES = list(range(int(FX["cases"])))


def case(k):
    p = f"c{k}_"
    return {key[len(p):]: FX[key] for key in FX.files if key.startswith(p)}


@pytest.mark.parametrize("k", CASES)
def test_oracle_matches_reference_model(oracle_mod, k):
    O = oracle_mod
    f = case(k)
    meta = [int(v) for v in f["meta"]]
    cfg, T = tuple(meta[:7]), meta[7]
    clip = float(f["clip"]) if meta[9] else None
    detach = bool(meta[10])
    c = O.code(*cfg)
    ga = np.zeros(f["alpha_table"].shape)
    gb = np.zeros(f["beta_table"].shape)
    for b in range(f["llr"].shape[0]):
        q, it = O.unrolled_forward(c, T, f["llr"][b], f["alpha_table"], f["beta_table"], clip=clip,
                                   precision="f32", want_iters=True)
        assert np.array_equal(q, f["lq"][b]), (k, b)
        assert np.array_equal(it, f["its"][:, b]), (k, b)
        gl, a_b, b_b = O.unrolled_backward(c, T, f["llr"][b], f["alpha_table"], f["beta_table"],
                                           f["w"][b], f["wi"][:, b], clip=clip, detach=detach,
                                           precision="f32")
        ref = f["g_llr"][b]
        assert np.allclose(gl, ref, rtol=1e-5, atol=1e-5 * max(np.abs(ref).max(), 1.0)), (k, b)
        ga += a_b
        gb += b_b
    for mine, ref in ((ga, f["g_alpha_table"]), (gb, f["g_beta_table"])):
        assert np.abs(mine - ref).max() <= 1e-5 * max(np.abs(ref).max(), 1e-3), k
