"""GPU parity of the table-driven int16 decoders (algo 2: trained
per-(iteration, row) alpha/beta tables; algo 3: alpha by check degree)
through the C ABI, bit-exact against the reference fixtures and the oracle:
hard bits, status, iterations, posteriors and the HARQ buffer.

The algo-2 fixtures come from three builds of the reference decoder: its own
default tables, and two table sets written by this library's exporter and
linked into it (a smooth trained-like set and an adversarial one with
negative alpha, negative beta and int16 extremes).  The algo-3 fixtures use
the default degree alpha through c_ldpc_5g_decode and custom tables through
layered_degree_minsum_ldpc_decode_with_alpha.
"""
import ctypes
import sys

import numpy as np
import pytest

from conftest import case_id, load_decode_cases

from deeplearning_5g_ldpc_decoder_amd import _lib, tables
from deeplearning_5g_ldpc_decoder_amd._lib import (CbDescriptor, DecoderConfig, LdpcInConfig,
                                                   ptr)

pytestmark = pytest.mark.gpu

ALGO2 = load_decode_cases("decode_algo2.npz")
ALGO3 = load_decode_cases("decode_algo3.npz")
MIXED = [(1, 8424, 2, 25344, 0, 2), (2, 1016, 2, 3120, 1, 2), (1, 488, 1, 1536, 2, 4),
         (2, 16, 0, 200, 0, 2), (1, 16, 0, 300, 3, 2), (2, 2056, 3, 4160, 0, 4),
         (1, 3816, 1, 7998, 0, 6), (2, 56, 1, 240, 0, 2), (1, 64, 0, 720, 0, 8)]


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def L():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a visible MI355X"
    return _lib.load()


@pytest.fixture
def restore_tables():
    yield
    tables.set_trained_tables(1)
    tables.set_trained_tables(2)
    tables.set_degree_alpha()


def _fixture_tables(idx):
    z = np.load(__file__.rsplit("/", 1)[0] + "/golden/decode_algo2.npz")
    return {b: (z[f"alpha_bg{b}"][idx], z[f"beta_bg{b}"][idx]) for b in (1, 2)}


def _fixture_degree_alpha(idx):
    z = np.load(__file__.rsplit("/", 1)[0] + "/golden/decode_algo3.npz")
    return z["degree_alpha"][idx]


@pytest.mark.parametrize("case", ALGO2 + ALGO3, ids=case_id)
def test_table_algos_abi_match_reference(L, case, restore_tables):
    from test_gpu_parity import abi_decode
    if "tables" in case:
        tables.set_trained_tables(case["cfg"][0], *case["tables"])
    if case.get("degree_alpha") is not None:
        tables.set_degree_alpha(case["degree_alpha"])
    _say(f"case {case_id(case)}")
    out = abi_decode(L, case["llr"], case["cfg"], case["row"], case["harq_in"])
    n = len(case["soft"])
    assert out["status"] == case["status"]
    assert out["iters"] == case["iters"]
    assert np.array_equal(out["soft"][:n], case["soft"])
    if case["row"][1] == 0:
        assert np.all(out["hard"] == 0xA5)
    else:
        assert np.array_equal(out["hard"], case["hard"])
    assert np.array_equal(out["lq"][:n], case["lq"])


def test_algo2_refuses_iterations_past_its_table(L, restore_tables):
    """max_iterations beyond the codeblock's base-graph table: -2 before rate
    recovery (HARQ_memory and hard bits untouched); a shorter table set at run
    time moves the limit, the other base graph keeps its own."""
    from test_gpu_parity import harq_view
    case = next(c for c in ALGO2 if c["cfg"][0] == 2 and c["harq_in"] is None)
    hv = harq_view(L)
    hv[:] = 1234
    hard = np.full(len(case["hard"]), 0x5A, np.uint8)
    llr = np.ascontiguousarray(case["llr"], np.int16)
    ic = LdpcInConfig(*case["cfg"])
    st = L.c_ldpc_5g_decode(ptr(llr), ptr(hard), ctypes.byref(ic),
                            ctypes.byref(DecoderConfig(2, 33, 0, 0, 1280, 32767, -32767)))
    assert st == -2
    assert np.all(hv == 1234) and np.all(hard == 0x5A)
    a, b = tables.get_trained_tables(2)
    tables.set_trained_tables(2, a[:4], b[:4])
    st = L.c_ldpc_5g_decode(ptr(llr), ptr(hard), ctypes.byref(ic),
                            ctypes.byref(DecoderConfig(2, 5, 0, 0, 1280, 32767, -32767)))
    assert st == -2
    st = L.c_ldpc_5g_decode(ptr(llr), ptr(hard), ctypes.byref(ic),
                            ctypes.byref(DecoderConfig(2, 4, 0, 0, 1280, 32767, -32767)))
    assert st in (0, 1)
    cfg1 = next(c for c in ALGO2 if c["cfg"][0] == 1 and c["harq_in"] is None)
    llr1 = np.ascontiguousarray(cfg1["llr"], np.int16)
    hard1 = np.zeros(len(cfg1["hard"]), np.uint8)
    st = L.c_ldpc_5g_decode(ptr(llr1), ptr(hard1), ctypes.byref(LdpcInConfig(*cfg1["cfg"])),
                            ctypes.byref(DecoderConfig(2, 20, 0, 0, 1280, 32767, -32767)))
    assert st in (0, 1)
    # plans refuse the same way and name the codeblock
    from deeplearning_5g_ldpc_decoder_amd.batch import BatchPlan
    with pytest.raises(ValueError, match="codeblock 1"):
        BatchPlan([cfg1["cfg"], case["cfg"]], [2, 5, 0, 0, 1280, 32767, -32767])


def _random_batch(oracle_mod, rng, n):
    cfgs, llrs = [], []
    for _ in range(n):
        bgn, A, crc, E, rv, Qm = MIXED[rng.integers(0, len(MIXED))]
        cb = oracle_mod.make_codeblock(bgn, A, crc, E, rv, Qm, float(rng.uniform(-3, 1)), rng)
        cfgs.append((bgn, A, crc, E, rv, Qm, 0))
        llrs.append(cb["llr"])
    return cfgs, llrs


def _oracle_batch(oracle_mod, cfgs, llrs, row, tabs=None, dalpha=None):
    out = []
    for cfg, llr in zip(cfgs, llrs):
        c = oracle_mod.code(*cfg)
        soft = oracle_mod.rate_recover(c, row[4], llr)
        out.append(oracle_mod.decode_soft(c, row, soft, tables=None if tabs is None
                                          else tabs[cfg[0]], degree_alpha=dalpha))
    return out


def _check_plan(plan, out, want):
    hard = plan.split_hard(out["hard"].cpu().numpy())
    lq = plan.split_lq(out["lq"].cpu().numpy())
    st = out["status"].cpu().numpy()
    its = out["iters"].cpu().numpy()
    for i, (s, h, q, k) in enumerate(want):
        assert (st[i], its[i]) == (s, k), i
        assert np.array_equal(hard[i], h), i
        assert np.array_equal(lq[i], q), i


@pytest.mark.parametrize("algo,tab,flags", [(2, 1, 0), (2, 2, 0), (2, 2, 4), (3, 1, 0),
                                            (3, 2, 4), (2, 0, 2)])
def test_table_plans_match_oracle(L, oracle_mod, restore_tables, algo, tab, flags):
    """Mixed BG1/BG2 batches (Zc 2..384) in one launch, packed tiles (flags
    0/2) and one codeblock per workgroup (4)."""
    import torch
    from deeplearning_5g_ldpc_decoder_amd.batch import BatchPlan
    rng = np.random.default_rng(70 + 3 * algo + tab)
    cfgs, llrs = _random_batch(oracle_mod, rng, 24)
    row = [algo, 10, 0, 24, 1280, 32767, -32767]
    tabs = dalpha = None
    if algo == 2:
        tabs = _fixture_tables(tab)
        for b in (1, 2):
            tables.set_trained_tables(b, *tabs[b])
    else:
        dalpha = _fixture_degree_alpha(tab)
        tables.set_degree_alpha(dalpha)
    _say(f"plan algo={algo} table={tab} flags={flags}")
    plan = BatchPlan(cfgs, row, want_lq=True, flags=flags)
    out = plan.decode(torch.from_numpy(np.concatenate(llrs)).cuda())
    torch.cuda.synchronize()
    _check_plan(plan, out, _oracle_batch(oracle_mod, cfgs, llrs, row, tabs, dalpha))
    plan.close()


def test_plan_keeps_tables_of_its_creation(L, oracle_mod, restore_tables):
    """Plans snapshot the tables; changing them afterwards affects plans
    created later only."""
    import torch
    from deeplearning_5g_ldpc_decoder_amd.batch import BatchPlan
    rng = np.random.default_rng(77)
    cfgs, llrs = _random_batch(oracle_mod, rng, 12)
    row = [2, 8, 0, 0, 1280, 32767, -32767]
    t1, t2 = _fixture_tables(1), _fixture_tables(2)
    for b in (1, 2):
        tables.set_trained_tables(b, *t1[b])
    p1 = BatchPlan(cfgs, row, want_lq=True)
    for b in (1, 2):
        tables.set_trained_tables(b, *t2[b])
    p2 = BatchPlan(cfgs, row, want_lq=True)
    d = torch.from_numpy(np.concatenate(llrs)).cuda()
    o1, o2 = p1.decode(d), p2.decode(d)
    torch.cuda.synchronize()
    _check_plan(p1, o1, _oracle_batch(oracle_mod, cfgs, llrs, row, t1))
    _check_plan(p2, o2, _oracle_batch(oracle_mod, cfgs, llrs, row, t2))
    p1.close()
    p2.close()


@pytest.mark.parametrize("algo", [2, 3])
def test_batch_context_table_algos(L, oracle_mod, restore_tables, algo):
    """Reusable batch contexts (continuous batching) with algos 2/3: two
    decodes of different codeblock sets on one context, plus the refusal of
    algo-2 iterations beyond the table at decode time (bad index reported)."""
    import torch
    rng = np.random.default_rng(90 + algo)
    row = [algo, 12, 0, 16, 1280, 32767, -32767]
    tabs = _fixture_tables(2) if algo == 2 else None
    dalpha = _fixture_degree_alpha(2) if algo == 3 else None
    if algo == 2:
        for b in (1, 2):
            tables.set_trained_tables(b, *tabs[b])
    else:
        tables.set_degree_alpha(dalpha)
    h = ctypes.c_void_p()
    assert L.ldpc_mi355_batch_create(ctypes.byref(h), 32, ctypes.byref(DecoderConfig(*row)),
                                     1) == 0
    # the context keeps the tables of its creation
    tables.set_trained_tables(1)
    tables.set_trained_tables(2)
    tables.set_degree_alpha()
    try:
        for rep, n in enumerate((20, 9)):
            cfgs, llrs = _random_batch(oracle_mod, rng, n)
            descs = (CbDescriptor * n)()
            off = 0
            for i, c in enumerate(cfgs):
                descs[i].cfg = LdpcInConfig(*c)
                descs[i].llr_offset = off
                descs[i].soft_slot = -1
                off += c[3]
            d_llr = torch.from_numpy(np.concatenate(llrs)).cuda()
            hard = torch.zeros(n * 1056, dtype=torch.uint8, device="cuda")
            st = torch.zeros(n, dtype=torch.int32, device="cuda")
            it = torch.zeros(n, dtype=torch.int32, device="cuda")
            lq = torch.zeros(n * 68 * 384, dtype=torch.int16, device="cuda")
            bad = ctypes.c_int32(-1)
            _say(f"batch context algo={algo} decode {rep} n={n}")
            rc = L.ldpc_mi355_batch_decode(h, descs, n, ptr(d_llr), None, ptr(hard), ptr(st),
                                           ptr(it), ptr(lq), None, ctypes.byref(bad))
            assert rc == 0
            assert L.ldpc_mi355_batch_wait(h) == 0
            hoff = np.zeros(n, np.uint32)
            loff = np.zeros(n, np.uint32)
            assert L.ldpc_mi355_batch_layout(h, None, None, ptr(hoff), ptr(loff)) == n
            want = _oracle_batch(oracle_mod, cfgs, llrs, row, tabs, dalpha)
            hh, ss, ii, qq = (hard.cpu().numpy(), st.cpu().numpy(), it.cpu().numpy(),
                              lq.cpu().numpy())
            for i, (s, hb, q, k) in enumerate(want):
                assert (ss[i], ii[i]) == (s, k), (rep, i)
                assert np.array_equal(hh[hoff[i]:hoff[i] + len(hb)], hb), (rep, i)
                assert np.array_equal(qq[loff[i]:loff[i] + len(q)], q), (rep, i)
    finally:
        L.ldpc_mi355_batch_destroy(h)
    if algo == 2:
        # a context whose decoder row needs more iterations than the table
        tables.set_trained_tables(2, *[t[:6] for t in tabs[2]])
        h2 = ctypes.c_void_p()
        assert L.ldpc_mi355_batch_create(ctypes.byref(h2), 4, ctypes.byref(
            DecoderConfig(2, 8, 0, 0, 1280, 32767, -32767)), 0) == 0
        try:
            descs = (CbDescriptor * 2)()
            descs[0].cfg = LdpcInConfig(1, 488, 1, 1536, 0, 2, 0)
            descs[1].cfg = LdpcInConfig(2, 176, 1, 512, 0, 2, 0)
            descs[0].soft_slot = descs[1].soft_slot = -1
            descs[1].llr_offset = 1536
            d = torch.zeros(2048, dtype=torch.int16, device="cuda")
            out = torch.zeros(64, dtype=torch.int32, device="cuda")
            bad = ctypes.c_int32(-1)
            rc = L.ldpc_mi355_batch_decode(h2, descs, 2, ptr(d), None, ptr(out), ptr(out),
                                           ptr(out), None, None, ctypes.byref(bad))
            assert rc == -2 and bad.value == 1
        finally:
            L.ldpc_mi355_batch_destroy(h2)


def test_message_wrap_row_takes_exact_path(L, oracle_mod):
    """alpha 22936 with beta 40000 (-25536 as int16): minima of 21279 scale
    to int16 -32768, whose negation the reference keeps at -32768.  The
    library must run such rows in the exact form (regression: rows with
    alpha >= 0 used to take the fast form whatever beta was)."""
    from test_gpu_parity import abi_decode
    row = [1, 3, 22936, 40000, 21279, 32767, -32767]
    cfg = (2, 16, 0, 200, 0, 2, 1)
    c = oracle_mod.code(*cfg)
    n = c.N + 2 * c.Zc
    soft = np.full(n, 21279, np.int16)
    soft[1::3] = -21279
    llr = np.zeros(200, np.int16)
    want = oracle_mod.decode_soft(c, row, oracle_mod.rate_recover(c, row[4], llr, soft))
    out = abi_decode(L, llr, cfg, row, harq_in=soft)
    assert (out["status"], out["iters"]) == (want[0], want[3])
    assert np.array_equal(out["lq"][:n], want[2])
    assert np.array_equal(out["hard"], want[1])


def test_trained_model_tables_decode_as_exported(L, oracle_mod, restore_tables):
    """The trainable decoder's tables (algo2_tables: Q15 alpha, beta x128,
    the last trained iteration repeated) installed with load_algo2_tables
    decode with algo 2 exactly as the oracle decodes with the same tables --
    the reference's export / rebuild / decode loop, without the rebuild (the
    exported files themselves compile into the reference decoder: see the
    algo-2 fixtures)."""
    import torch
    from deeplearning_5g_ldpc_decoder_amd.batch import BatchPlan
    from deeplearning_5g_ldpc_decoder_amd.unrolled import (CTableNeuralLayeredMinSumDecoder,
                                                           load_algo2_tables,
                                                           load_ldpc_table_config)
    tc = load_ldpc_table_config(1, 2088, 2, 6336, 0, 2)
    model = CTableNeuralLayeredMinSumDecoder(tc, num_iterations=10,
                                             alpha_mode="iteration_degree",
                                             beta_mode="iteration_degree", init_alpha=0.7,
                                             init_beta=0.2)
    g = torch.Generator().manual_seed(5)
    with torch.no_grad():
        for p in model.parameters():
            p.add_(0.6 * torch.rand(p.shape, generator=g) - 0.3)
    fa, fb = load_algo2_tables(model)
    assert fa.shape == (32, 46)
    assert np.array_equal(fa[10:], np.repeat(fa[9:10], 22, axis=0))
    assert fb.any()
    a, b = tables.get_trained_tables(1)
    assert np.array_equal(a, fa) and np.array_equal(b, fb)
    rng = np.random.default_rng(12)
    cfg = (1, 2088, 2, 6336, 0, 2, 0)
    llrs = [oracle_mod.make_codeblock(*cfg[:6], float(rng.uniform(-2, 0)), rng)["llr"]
            for _ in range(16)]
    row = [2, 16, 0, 0, 1280, 32767, -32767]
    plan = BatchPlan([cfg] * 16, row, want_lq=True)
    out = plan.decode(torch.from_numpy(np.concatenate(llrs)).cuda())
    torch.cuda.synchronize()
    want = _oracle_batch(oracle_mod, [cfg] * 16, llrs, row, {1: (fa, fb)})
    _check_plan(plan, out, want)
    assert sum(w[0] == 0 for w in want) >= 4
    plan.close()
