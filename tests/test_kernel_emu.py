"""CPU tests of the decode-kernel source: ldpc_kernels.hip built for the host
(tests/kernel_emu/) and run lane by lane, compared with the reference's own
fixtures and the oracle.  This checks the kernels' indexing, schedule, CRC
reduction and fused rate recovery without a GPU; the GPU tests check the
compiled gfx950 code.  With glibc tanhf/atanhf on the host, the BP kernel's
source is bit-exact with the reference's FP32 BP here too.
"""
import os
import sys

import numpy as np
import pytest

from conftest import ROOT, case_id, load_decode_cases

sys.path.insert(0, os.path.join(ROOT, "tests", "kernel_emu"))
import emu  # noqa: E402

ALGO1 = load_decode_cases("decode_algo1.npz")
ALGO4 = load_decode_cases("decode_algo4.npz")
ALGO0 = load_decode_cases("decode_algo0.npz")


def _small(c):
    return c["cfg"][1] <= 4096 and c["harq_in"] is None


def _run_case(case):
    cfg = case["cfg"]
    pool = np.zeros(68 * 384, np.int16)
    if case["harq_in"] is not None:
        pool[:len(case["harq_in"])] = case["harq_in"]
    out = emu.decode_batch([cfg], case["row"], [case["llr"]], soft_pool=pool, soft_slots=[0])[0]
    n = len(case["soft"])
    assert np.array_equal(pool[:n], case["soft"])
    assert out["status"] == case["status"]
    assert out["iters"] == case["iters"]
    if case["row"][1]:
        assert np.array_equal(out["hard"], case["hard"])
    if case["row"][0] == 1:
        assert np.array_equal(out["lq"], case["lq"])
    else:
        assert np.array_equal(out["lq"].view(np.uint32), case["lq"].view(np.uint32))


@pytest.mark.parametrize("case", [c for c in ALGO1 if _small(c) or c["harq_in"] is not None],
                         ids=case_id)
def test_emu_algo1_matches_reference(case):
    _run_case(case)


def test_emu_algo1_full_size_case():
    case = next(c for c in ALGO1 if c["cfg"][1] == 8424 and c["cfg"][3] == 25344)
    _run_case(case)


@pytest.mark.parametrize("case", [c for c in ALGO4 if _small(c)], ids=case_id)
def test_emu_algo4_matches_reference(case):
    _run_case(case)


@pytest.mark.parametrize("case", [c for c in ALGO0 if _small(c)], ids=case_id)
def test_emu_algo0_matches_reference(case):
    _run_case(case)


MIXED = [(1, 8424, 2, 25344, 0, 2), (2, 1016, 2, 3120, 1, 2), (1, 488, 1, 1536, 2, 4),
         (2, 16, 0, 200, 0, 2), (2, 2056, 3, 4160, 0, 4), (1, 64, 0, 720, 0, 8)]


@pytest.mark.parametrize("row", [[1, 8, 22936, 0, 1280, 32767, -32767],
                                 [4, 10, 22936, 0, 1280, 32767, -32767],
                                 [0, 6, 1, 0, 1280, 20, -32768]])
def test_emu_mixed_batch_matches_oracle(oracle_mod, row):
    """Small-Zc codeblocks inside a 384-lane workgroup (sized by the largest
    Zc of the batch): idle lanes and extra waves must not change results."""
    rng = np.random.default_rng(7 + row[0])
    cfgs, llrs = [], []
    for i, c in enumerate(MIXED):
        cb = oracle_mod.make_codeblock(*c, float(rng.uniform(-3, 0)), rng)
        cfgs.append(c + (0,))
        llrs.append(cb["llr"])
    outs = emu.decode_batch(cfgs, row, llrs)
    prec = "i16" if row[0] == 1 else "f32"
    for cfg, llr, out in zip(cfgs, llrs, outs):
        c = oracle_mod.code(*cfg)
        s, h, q, k = oracle_mod.decode_soft(c, row, oracle_mod.rate_recover(c, row[4], llr),
                                            precision=prec)
        assert (out["status"], out["iters"]) == (s, k), cfg
        assert np.array_equal(out["hard"], h), cfg
        assert np.array_equal(out["lq"].view(np.uint8), q.view(np.uint8)), cfg
