"""CPU tests of the decode-kernel source: ldpc_kernels.hip built for the host
(tests/kernel_emu/) and run lane by lane, compared with the reference's own
fixtures and the oracle.  This checks the kernels' indexing, schedule, CRC
reduction and fused rate recovery without a GPU; the GPU tests check the
compiled gfx950 code.  With glibc tanhf/atanhf on the host, the BP kernel's
source (ldpc_bp.hip) is bit-exact with the reference's FP32 BP here too.
"""
import os
import sys

import numpy as np
import pytest

from conftest import ROOT, case_id, load_decode_cases, rate_recovery_cases

sys.path.insert(0, os.path.join(ROOT, "tests", "kernel_emu"))
import emu  # noqa: E402

ALGO1 = load_decode_cases("decode_algo1.npz")
ALGO4 = load_decode_cases("decode_algo4.npz")
ALGO0 = load_decode_cases("decode_algo0.npz")
ALGO2 = load_decode_cases("decode_algo2.npz")
ALGO3 = load_decode_cases("decode_algo3.npz")


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
    if case["row"][0] in (1, 2, 3):
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


def test_emu_algo0_full_size_case():
    case = next(c for c in ALGO0 if c["cfg"][1] == 8424)
    _run_case(case)


def test_emu_algo0_random_llrs_bg2_zc26(oracle_mod):
    """BG2 Zc=26 (A=176, CRC24A, E=512) with random LLRs, 8 BP iterations
    without a CRC pass: the kernel source equals the oracle's FP32 BP."""
    O = oracle_mod
    rng = np.random.default_rng(2)
    llr = rng.integers(-500, 500, 512).astype(np.int16)
    cfg = (2, 176, 1, 512, 0, 2, 0)
    row = [0, 8, 1, 0, 1280, 20, -32768]
    out = emu.decode_batch([cfg], row, [llr])[0]
    ref = O.decode_path(llr, cfg, row)
    assert out["status"] == ref["status"] == 1 and out["iters"] == ref["iters"] == 8
    assert np.array_equal(out["hard"], ref["hard"])
    assert np.array_equal(out["lq"].view(np.uint32), ref["lq"].view(np.uint32))


@pytest.fixture
def case_tables():
    """Loads a fixture case's algo-2 / algo-3 tables into the CPU build and
    restores the defaults afterwards."""
    def load(case):
        if "tables" in case:
            emu.set_trained_tables(case["cfg"][0], *case["tables"])
        if "degree_alpha" in case:
            emu.set_degree_alpha(case["degree_alpha"])
    yield load
    emu.set_trained_tables(1)
    emu.set_trained_tables(2)
    emu.set_degree_alpha(None)


@pytest.mark.parametrize("case", [c for c in ALGO2 + ALGO3
                                  if c["cfg"][1] <= 1100 or c["harq_in"] is not None],
                         ids=case_id)
def test_emu_table_algos_match_reference(case, case_tables):
    """Algo 2 (trained tables: reference defaults, exported smooth and
    adversarial sets) and algo 3 (default and custom degree alpha); the
    larger fixture codes run on the GPU only."""
    case_tables(case)
    _run_case(case)


def test_emu_algo2_refuses_iterations_past_table():
    case = ALGO2[0]
    with pytest.raises(RuntimeError):
        emu.decode_batch([case["cfg"]], [2, 33, 0, 0, 1280, 32767, -32767], [case["llr"]])


def test_fast_path_choice_covers_message_wrap(oracle_mod):
    """alpha 22936, beta 40000 (-25536 as int16): a check whose smallest
    magnitude is 21279 scales to 32768, which the reference stores as int16
    -32768 and negates to -32768 again.  The library must pick the exact row
    form for this decoder row; the fast form (forced here as a negative
    control) gets the negated message wrong."""
    row = [1, 1, 22936, 40000, 21279, 32767, -32767]   # filler bits at 21279 too
    assert emu.row_general(row)
    assert not emu.row_general([1, 1, 22936, 0, 0, 32767, -32767])
    assert not emu.row_general([1, 1, 32767, 65535, 0, 32767, -32767])   # beta -1: 32767 max
    assert emu.row_general([1, 1, 32767, 65534, 0, 32767, -32767])       # beta -2: 32768
    cfg = (2, 16, 0, 200, 0, 2, 1)
    c = oracle_mod.code(*cfg)
    n = c.N + 2 * c.Zc
    soft = np.full(n, 21279, np.int16)
    soft[1::3] = -21279
    llr = np.zeros(200, np.int16)
    s, h, q, k = oracle_mod.decode_soft(c, row, oracle_mod.rate_recover(c, row[4], llr, soft))
    results = {}
    for g in (None, False):
        emu.force_general(g)
        try:
            pool = np.zeros(68 * 384, np.int16)
            pool[:n] = soft
            results[g] = emu.decode_batch([cfg], row, [llr], soft_pool=pool, soft_slots=[0])[0]
        finally:
            emu.force_general(None)
    assert np.array_equal(results[None]["lq"], q)
    assert not np.array_equal(results[False]["lq"], q)


@pytest.mark.parametrize("algo", [2, 3])
def test_emu_packed_tiles_table_algos(oracle_mod, algo):
    """Packed tiles over the lifting sizes with the adversarial algo-2 tables
    (exact row form) and custom degree alpha, bit-exact with the oracle."""
    rng = np.random.default_rng(40 + algo)
    cfgs, llrs = _all_zc_batch(oracle_mod, rng, per_zc=2, zcs=[2, 7, 26, 88])
    z = np.load(os.path.join(ROOT, "tests", "golden", "decode_algo2.npz"))
    tabs = {b: (z[f"alpha_bg{b}"][2], z[f"beta_bg{b}"][2]) for b in (1, 2)}
    dz = np.load(os.path.join(ROOT, "tests", "golden", "decode_algo3.npz"))["degree_alpha"][1]
    row = [algo, 4, 0, 24, 1280, 32767, -32767]
    try:
        for b in (1, 2):
            emu.set_trained_tables(b, *tabs[b])
        emu.set_degree_alpha(dz)
        outs = emu.decode_batch(cfgs, row, llrs, tile_k=0)
    finally:
        emu.set_trained_tables(1)
        emu.set_trained_tables(2)
        emu.set_degree_alpha(None)
    for cfg, llr, out in zip(cfgs, llrs, outs):
        c = oracle_mod.code(*cfg)
        s, h, q, k = oracle_mod.decode_soft(c, row, oracle_mod.rate_recover(c, row[4], llr),
                                            tables=tabs[cfg[0]], degree_alpha=dz)
        assert (out["status"], out["iters"]) == (s, k), cfg
        assert np.array_equal(out["lq"], q), cfg


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


ALL_ZC = [2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13, 14, 15, 16, 18, 20, 22, 24, 26, 28, 30, 32,
          36, 40, 44, 48, 52, 56, 60, 64, 72, 80, 88, 96, 104, 112, 120, 128, 144, 160, 176,
          192, 208, 224, 240, 256, 288, 320, 352, 384]
# spread of small / odd / large lifting sizes for the CPU-emulated runs
SOME_ZC = [2, 3, 7, 15, 26, 44, 64, 120, 208, 384]


def _all_zc_batch(oracle_mod, rng, per_zc=2, zcs=ALL_ZC):
    """Codeblocks covering the given lifting sizes (TS 38.212 Table 5.3.2-1)
    for each base graph where reachable, rates 0.25..0.85, Qm 2..8."""
    cfgs, llrs = [], []
    for bgn in (1, 2):
        for zc in zcs:
            found = None
            kb = 22 if bgn == 1 else 10
            for crc, clen in ((2, 24), (0, 16)):
                for B in range((kb * zc) // 8 * 8, clen + 7, -8):
                    A = B - clen
                    st, p = oracle_mod.params(bgn, A, crc, 1000, 0, 2)
                    if st == 0 and p["Zc"] == zc:
                        found = (A, crc, B)
                        break
                if found:
                    break
            if found is None:
                continue
            A, crc, B = found
            for _ in range(per_zc):
                Qm = int(rng.choice([2, 4, 6, 8]))
                R = float(rng.uniform(0.25, 0.85))
                E = int(B / R)
                E -= E % Qm
                rv = int(rng.integers(0, 4))
                cb = oracle_mod.make_codeblock(bgn, A, crc, E, rv, Qm,
                                               float(rng.uniform(-4, -1)) + 10 * np.log10(Qm) / 2,
                                               rng)
                cfgs.append((bgn, A, crc, E, rv, Qm, 0))
                llrs.append(cb["llr"])
    return cfgs, llrs


@pytest.mark.parametrize("tile_k", [0, 3])
def test_emu_packed_tiles_lifting_sizes(oracle_mod, tile_k):
    """Packed tiles (groups of codeblocks of one BG/Zc on whole waves, groups
    of several lifting sizes per 512-lane workgroup, per-codeblock CRC and
    early stop) over every reachable lifting size, bit-exact with the oracle
    per codeblock."""
    rng = np.random.default_rng(31 + tile_k)
    cfgs, llrs = _all_zc_batch(oracle_mod, rng, per_zc=3, zcs=SOME_ZC)
    assert len(cfgs) >= 50
    row = [1, 8, 22936, 0, 1280, 32767, -32767]
    outs = emu.decode_batch(cfgs, row, llrs, tile_k=tile_k)
    for cfg, llr, out in zip(cfgs, llrs, outs):
        c = oracle_mod.code(*cfg)
        s, h, q, k = oracle_mod.decode_soft(c, row, oracle_mod.rate_recover(c, row[4], llr))
        assert (out["status"], out["iters"]) == (s, k), cfg
        assert np.array_equal(out["hard"], h), cfg
        assert np.array_equal(out["lq"], q), cfg


@pytest.mark.parametrize("grid", [1, 3])
def test_emu_persistent_tiles_match_oracle(oracle_mod, grid):
    """Persistent tile launch: `grid` workgroups pull every tile from the
    device queue (each workgroup decodes several tiles in turn, reusing its
    LDS and CRC banks), bit-exact with the oracle per codeblock; the last
    workgroup out leaves the queue counters at zero (checked by the driver,
    which fails the launch otherwise)."""
    rng = np.random.default_rng(77 + grid)
    cfgs, llrs = _all_zc_batch(oracle_mod, rng, per_zc=2, zcs=SOME_ZC)
    row = [1, 8, 22936, 0, 1280, 32767, -32767]
    emu.persistent_tiles(grid)
    try:
        outs = emu.decode_batch(cfgs, row, llrs, tile_k=0)
    finally:
        emu.persistent_tiles(0)
    for cfg, llr, out in zip(cfgs, llrs, outs):
        c = oracle_mod.code(*cfg)
        s, h, q, k = oracle_mod.decode_soft(c, row, oracle_mod.rate_recover(c, row[4], llr))
        assert (out["status"], out["iters"]) == (s, k), cfg
        assert np.array_equal(out["hard"], h), cfg
        assert np.array_equal(out["lq"], q), cfg


def test_emu_packed_tiles_general_row(oracle_mod):
    """Packed tiles with a decoder row that needs the exact general path
    (alpha negative as int16, min_llr > max_llr)."""
    rng = np.random.default_rng(5)
    cfgs, llrs = _all_zc_batch(oracle_mod, rng, per_zc=1, zcs=SOME_ZC)
    for row in ([1, 6, 40000, 16, 1280, 32767, -32767], [1, 6, 22936, 0, 1280, 300, 2000]):
        outs = emu.decode_batch(cfgs, row, llrs, tile_k=0)
        for cfg, llr, out in zip(cfgs, llrs, outs):
            c = oracle_mod.code(*cfg)
            s, h, q, k = oracle_mod.decode_soft(c, row, oracle_mod.rate_recover(c, row[4], llr))
            assert (out["status"], out["iters"]) == (s, k), (row, cfg)
            assert np.array_equal(out["lq"], q), (row, cfg)


# --------------------------------------------------------------------------
# Unrolled trainable min-sum kernels (ldpc_unrolled.hip) against the oracle:
# forward beliefs, per-iteration beliefs and dL/dllr bit for bit, the alpha
# and beta sums (float partials, double reduction) to 1e-6 relative.
# --------------------------------------------------------------------------
UNROLLED_CASES = [
    # cfg (bgn, A, crctype, E), iterations, codeblocks per workgroup (0 = default), batch,
    # clip, detach
    ((2, 40, 0, 200), 2, 0, 5, None, False),   # BG2 Zc=10: 16 lanes per codeblock, 24 packed
    ((2, 40, 0, 200), 3, 1, 2, 5.0, False),    # one codeblock per workgroup
    ((2, 40, 0, 200), 2, 2, 3, 5.0, True),     # detach
    ((1, 104, 0, 390), 2, 3, 4, 5.0, False),   # BG1 Zc=6, last workgroup partly empty
    ((1, 104, 0, 390), 2, 3, 4, None, True),
    ((1, 1496, 2, 4560), 1, 0, 1, 5.0, False),  # BG1 Zc=72: two 64-lane segments
]


@pytest.mark.parametrize("cfg,T,cbs,batch,clip,detach", UNROLLED_CASES)
def test_emu_unrolled_matches_oracle(oracle_mod, cfg, T, cbs, batch, clip, detach):
    O = oracle_mod
    c = O.code(*cfg)
    n = c.N + 2 * c.Zc
    rng = np.random.default_rng(T * 7 + batch)
    llr = rng.normal(0.0, 3.0, (batch, n)).astype(np.float32)
    llr[:, :2 * c.Zc] = 0.0
    alpha = rng.uniform(0.55, 0.95, (T, c.rows)).astype(np.float32)
    beta = rng.uniform(0.0, 0.3, (T, c.rows)).astype(np.float32)
    g = rng.normal(size=(batch, n)).astype(np.float32)
    gi = (rng.normal(size=(T, batch, n)) * 0.3).astype(np.float32)
    r = emu.unrolled(cfg, T, llr, alpha, beta, clip=clip, detach=detach, cbs=cbs, g_lq=g,
                     g_iter=gi)
    ga = np.zeros((T, c.rows))
    gb = np.zeros((T, c.rows))
    for k in range(batch):
        q, it = O.unrolled_forward(c, T, llr[k], alpha, beta, clip=clip, want_iters=True)
        assert np.array_equal(q, r["lq"][k])
        assert np.array_equal(it, r["iter_lq"][:, k])
        gl, a_k, b_k = O.unrolled_backward(c, T, llr[k], alpha, beta, g[k], gi[:, k], clip=clip,
                                           detach=detach)
        assert np.array_equal(gl, r["g_llr"][k])
        ga += a_k
        gb += b_k
    for mine, ref in ((r["g_alpha"], ga), (r["g_beta"], gb)):
        assert np.abs(mine - ref).max() <= 1e-6 * max(np.abs(ref).max(), 1e-3)


# --------------------------------------------------------------------------
# Device input synthesis (ldpc_synth.hip) against the oracle's encoder / rate
# matcher (pinned by the reference encoder fixtures) and a numpy restatement
# of its random streams and channel.
# --------------------------------------------------------------------------
SYNTH_CASES = [
    (1, 8424, 2, 25344, 0, 2),    # BG1 Zc=384
    (2, 1016, 2, 2080, 2, 4),     # BG2 Zc=104, rv2
    (1, 488, 1, 1536, 1, 1),      # BG1 Zc=24 with filler bits, BPSK
    (2, 16, 0, 160, 3, 2),        # BG2 small, CRC16, rv3
    (2, 40, 0, 600, 0, 6),        # repetition (E > N)
]


@pytest.mark.parametrize("cfg", SYNTH_CASES)
def test_emu_synth_matches_oracle(oracle_mod, cfg):
    import synth_ref
    O = oracle_mod
    c = O.code(*cfg, 0)
    st, p = O.params(*cfg, 0)
    p.update(E=cfg[3], Qm=cfg[5])
    rng = np.random.default_rng(cfg[1])
    n = 2
    pay = rng.integers(0, 2, (n, cfg[1])).astype(np.uint8)
    out = emu.synth(cfg, n, -1.0, seed=9, payload=pay, beliefs=True)
    for i in range(n):
        b = O.attach_crc(pay[i], cfg[2])
        assert np.array_equal(out["bits"][i], b)
        cw = O.encode(c, b)
        assert np.array_equal(out["cw"][i], cw)
        assert np.array_equal(out["rm"][i], O.rate_match(c, cw))
        ref = synth_ref.channel_llr(out["rm"][i], cfg[5], -1.0, 9, i)
        assert np.allclose(out["llr_f32"][i], ref, rtol=1e-4, atol=1e-4)
        q = synth_ref.quantize(out["llr_f32"][i])
        assert np.array_equal(out["llr"][i], q)
        assert np.array_equal(out["beliefs"][i], synth_ref.float_rate_recover(out["llr_f32"][i], p))
    # generated payloads: the kernel's own stream, a valid CRC and codeword
    gen = emu.synth(cfg, 1, -1.0, seed=5)
    assert np.array_equal(gen["bits"][0][:cfg[1]], synth_ref.payload_bits(5, 0, cfg[1]))
    assert np.array_equal(gen["bits"][0], O.attach_crc(gen["bits"][0][:cfg[1]], cfg[2]))
    assert O.syndrome_ok(c, gen["cw"][0])


@pytest.mark.parametrize("pair", [(2, 26), (1, 15), (1, 16)])
@pytest.mark.parametrize("row,tab", [
    ([1, 8, 22936, 0, 1280, 32767, -32767], 2),      # packed, simple scale
    ([1, 8, 26000, 32, -900, 32767, -32767], 2),     # packed, offset min-sum
    ([1, 8, 22936, 0, 1280, 2000, -1500], 2),        # packed, tight clip
    ([1, 6, 22936, 40000, 21279, 32767, -32767], 2),  # exact: message wrap
    ([1, 6, 40000, 16, 1280, 300, 2000], 2),          # exact: alpha < 0, min > max
    ([2, 6, 0, 24, 1280, 32767, -32767], 0),          # default tables: packed, simple
    ([2, 6, 0, 24, 1280, 32767, -32767], 1),          # smooth tables: packed, beta != 0
    ([2, 6, 0, 24, 1280, 32767, -32767], 2),          # adversarial tables: exact
    ([3, 6, 0, 0, 1280, 32767, -32767], 1),           # degree alpha: packed, simple
    ([3, 6, 0, 40000, 1280, 32767, -32767], 2)])      # exact
def test_emu_fixed_kernel_matches_generic_and_oracle(oracle_mod, pair, row, tab, case_tables):
    """The lifting-size specialized kernel (compile-time shifts; packed
    row-paired messages for fast rows, simple scale when beta = 0 and
    max_llr = 32767; per-edge messages for exact rows) on uniform batches:
    bit-exact with the generic kernel and with the oracle."""
    bgn, zc = pair
    rng = np.random.default_rng(1000 * bgn + zc + row[0])
    cfgs, llrs = _all_zc_batch(oracle_mod, rng, per_zc=4, zcs=[zc])
    cfgs, llrs = zip(*[(c, l) for c, l in zip(cfgs, llrs) if c[0] == bgn])
    z = np.load(os.path.join(ROOT, "tests", "golden", "decode_algo2.npz"))
    tabs = (z[f"alpha_bg{bgn}"][tab], z[f"beta_bg{bgn}"][tab])
    dalpha = np.load(os.path.join(ROOT, "tests", "golden", "decode_algo3.npz"))["degree_alpha"][tab]
    case_tables(dict(cfg=(bgn,), tables=tabs, degree_alpha=dalpha))
    outs = {}
    try:
        for fixed in (True, False):
            emu.use_fixed(fixed)
            outs[fixed] = emu.decode_batch(list(cfgs), row, list(llrs))
    finally:
        emu.use_fixed(True)
    for cfg, llr, a, b in zip(cfgs, llrs, outs[True], outs[False]):
        assert (a["status"], a["iters"]) == (b["status"], b["iters"]), cfg
        assert np.array_equal(a["hard"], b["hard"]) and np.array_equal(a["lq"], b["lq"]), cfg
        c = oracle_mod.code(*cfg)
        s, h, q, k = oracle_mod.decode_soft(c, row, oracle_mod.rate_recover(c, row[4], llr),
                                            tables=tabs, degree_alpha=dalpha)
        assert (a["status"], a["iters"]) == (s, k), cfg
        assert np.array_equal(a["lq"], q), cfg


@pytest.mark.parametrize("pair", [(2, 26), (1, 15)])
@pytest.mark.parametrize("row", [[4, 20, 22936, 0, 1280, 32767, -32767],
                                 [4, 8, 26000, 300, -900, 2000, -1500]])
def test_emu_fixed_f32_kernel_matches_generic_and_oracle(oracle_mod, pair, row):
    """The lifting-size specialized FP32 min-sum kernel (algo 4; compile-time
    shifts, beliefs at LDS address 0) on uniform batches: bit-exact with the
    generic FP32 kernel and with the oracle's FP32 decoder (posteriors,
    status, iterations)."""
    bgn, zc = pair
    rng = np.random.default_rng(2000 * bgn + zc + row[1])
    cfgs, llrs = _all_zc_batch(oracle_mod, rng, per_zc=4, zcs=[zc])
    cfgs, llrs = zip(*[(c, l) for c, l in zip(cfgs, llrs) if c[0] == bgn])
    outs = {}
    try:
        for fixed in (True, False):
            emu.use_fixed(fixed)
            outs[fixed] = emu.decode_batch(list(cfgs), row, list(llrs))
    finally:
        emu.use_fixed(True)
    for cfg, llr, a, b in zip(cfgs, llrs, outs[True], outs[False]):
        assert (a["status"], a["iters"]) == (b["status"], b["iters"]), cfg
        assert np.array_equal(a["hard"], b["hard"]) and np.array_equal(a["lq"], b["lq"]), cfg
        c = oracle_mod.code(*cfg)
        s, h, q, k = oracle_mod.decode_soft(c, row, oracle_mod.rate_recover(c, row[4], llr))
        assert (a["status"], a["iters"]) == (s, k), cfg
        assert np.array_equal(a["lq"], q), cfg


EDGE_ZC = {1: [2, 3], 2: [2, 3, 4, 5, 6]}


@pytest.mark.parametrize("row", [[1, 8, 22936, 0, 1280, 32767, -32767],
                                 [1, 8, 26000, 32, -900, 32767, -32767],
                                 [1, 8, 22936, 0, 1280, 2000, -1500]])
def test_emu_edge_lanes_match_oracle(oracle_mod, row):
    """The lane-per-edge schedule (one wave per codeblock, a layer's edges one
    per lane, tree reduction of min1/min2/parity over each check's lanes) on
    every lifting size it takes (BG1 Zc 2-3, BG2 Zc 4 and 6), Qm 2-8, in
    one batch: bit-exact with the oracle (status, iterations, hard bits,
    posteriors) for the simple, offset and clipped fast-row decoder rows."""
    rng = np.random.default_rng(41 + row[3])
    cfgs, llrs = [], []
    for bgn, zcs in EDGE_ZC.items():
        c_, l_ = _all_zc_batch(oracle_mod, rng, per_zc=3, zcs=zcs)
        for c, l in zip(c_, l_):
            if c[0] == bgn:
                cfgs.append(c)
                llrs.append(l)
    # every reachable class (BG2 Zc 2, 3 and 5 take no codeblock size)
    assert {(c[0], oracle_mod.code(*c).Zc) for c in cfgs} == {(1, 2), (1, 3), (2, 4), (2, 6)}
    emu.edge_lanes(True)
    try:
        outs = emu.decode_batch(cfgs, row, llrs)
    finally:
        emu.edge_lanes(False)
    for cfg, llr, out in zip(cfgs, llrs, outs):
        c = oracle_mod.code(*cfg)
        s, h, q, k = oracle_mod.decode_soft(c, row, oracle_mod.rate_recover(c, row[4], llr))
        assert (out["status"], out["iters"]) == (s, k), cfg
        assert np.array_equal(out["hard"], h), cfg
        assert np.array_equal(out["lq"], q), cfg


def test_emu_scatter_chunked_and_unaligned(oracle_mod):
    """Rate recovery without repetition reads 16-byte chunks of 8 LLRs when
    the codeblock's LLRs are 16-byte aligned and E is a multiple of 8, and
    single LLRs otherwise.  Both forms, in the per-codeblock and the packed
    tile kernels, must give the oracle's buffer: Qm 1-8, every rv, filler
    bits, HARQ combining into a soft buffer."""
    rng = np.random.default_rng(11)
    cfgs = [(1, 8424, 2, 25344, 0, 2, 0)]
    # E a multiple of 8 and not, just inside N - F (no repetition)
    for bgn, A, crc, rv, qm, harq, frac in [(1, 3816, 2, 2, 4, 1, 0.9), (2, 1000, 1, 3, 6, 0, 0.8),
                                            (2, 296, 1, 1, 8, 1, 0.95), (1, 504, 1, 0, 1, 0, 0.7),
                                            (2, 200, 1, 0, 2, 1, 0.9), (2, 200, 1, 2, 2, 0, 0.85)]:
        c = oracle_mod.code(bgn, A, crc, 200 * qm, rv, qm)
        e = int((c.N - (c.K - c.B)) * frac) // qm * qm
        cfgs.append((bgn, A, crc, e, rv, qm, harq))
    assert {c[3] % 8 == 0 for c in cfgs} == {True, False}
    pad_cfg = (2, 40, 0, 150, 0, 2, 0)   # 150 LLRs: misaligns what follows
    pad_llr = rng.integers(-100, 100, 150).astype(np.int16)
    checked = 0
    for cfg in cfgs:
        c = oracle_mod.code(*cfg)
        assert c.E <= c.N - (c.K - c.B), cfg   # no repetition: the scatter form
        n = c.N + 2 * c.Zc
        llr = rng.integers(-3000, 3000, c.E).astype(np.int16)
        soft = rng.integers(-2000, 2000, n).astype(np.int16) if cfg[6] else None
        want = oracle_mod.rate_recover(c, 77, llr, soft)
        row = [1, 0, 22936, 0, 77, 32767, -32767]
        for lead in (False, True):
            for tile in (None, 0):
                pool = np.zeros(2 * 68 * 384, np.int16)
                if soft is not None:
                    pool[68 * 384:68 * 384 + n] = soft
                cs, ls, slots = [cfg], [llr], [1]
                if lead:
                    cs, ls, slots = [pad_cfg, cfg], [pad_llr, llr], [-1, 1]
                out = emu.decode_batch(cs, row, ls, soft_pool=pool, soft_slots=slots,
                                       tile_k=tile)[-1]
                assert np.array_equal(out["lq"], want), (cfg, lead, tile)
                if soft is not None:
                    assert np.array_equal(pool[68 * 384:68 * 384 + n], want), (cfg, lead, tile)
                checked += 1
    assert checked == 4 * len(cfgs)


def test_emu_rate_recovery_matches_reference():
    """The kernels' batched gather (recover_fill) on the reference's
    HARQ_memory fixtures: heavy repetition with int16 wrap, every rv, Qm 1-8,
    and an rv chain combining into the soft buffer."""
    for cfg, filler, llr, before, want in rate_recovery_cases():
        pool = np.zeros(68 * 384, np.int16)
        if before is not None:
            pool[:len(before)] = before
        row = [1, 0, 22936, 0, filler, 32767, -32767]
        emu.decode_batch([cfg], row, [llr], soft_pool=pool, soft_slots=[0])
        assert np.array_equal(pool[:len(want)], want), cfg


def test_tile_plan_invariants(oracle_mod):
    """The tile planner on 2048 mixed configs (the mixed-size workload's
    generator): every codeblock in exactly one group; a group's codeblocks
    share (BG, Zc) and sit on whole waves of one workgroup without overlap;
    all eight entries of a workgroup carry its base graph (so every wave runs
    the same rows and barriers); CRC slots and LDS regions of a workgroup are
    disjoint and within kMaxTileCodeblocks / the reported footprint; and
    workgroups do hold groups of different lifting sizes."""
    from deeplearning_5g_ldpc_decoder_amd.workloads import draw_config
    rng = np.random.default_rng(12)
    cfgs = [draw_config(rng)[0] for _ in range(2048)]
    tiles, order, lds = emu.plan_tiles(cfgs)
    codes = [oracle_mod.code(*c) for c in cfgs]
    seen = np.zeros(len(cfgs), int)
    mixed_wgs = 0
    for wg in tiles:
        assert len(set(int(t["bgn"]) for t in wg)) == 1
        groups = {}
        for v, t in enumerate(wg):
            if t["count"] == 0:   # a wave outside any group: the first group's rows
                assert t["lane_base"] == 64 * v and v > 0
                assert (t["Zc"], t["etab_off"], t["lds_base"]) == (wg[0]["Zc"], wg[0]["etab_off"],
                                                                   wg[0]["lds_base"])
                continue
            key = (int(t["first"]), int(t["lane_base"]))
            groups.setdefault(key, []).append(v)
        slots, regions, zcs = [], [], set()
        for (first, lane_base), waves in groups.items():
            t = wg[waves[0]]
            assert all(wg[v] == t for v in waves)
            assert lane_base == 64 * waves[0] and waves == list(range(waves[0], waves[-1] + 1))
            cnt, zc = int(t["count"]), int(t["Zc"])
            assert cnt * zc <= 64 * len(waves) < cnt * zc + 64
            zcs.add(zc)
            for i in range(cnt):
                c = codes[order[first + i]]
                assert (c.bgn, c.Zc) == (int(t["bgn"]), zc)
                assert int(t["lds_stride"]) >= c.N + 2 * c.Zc
                seen[order[first + i]] += 1
            slots.append((int(t["slot_base"]), cnt))
            regions.append((int(t["lds_base"]), cnt * int(t["lds_stride"])))
        for lst, cap in ((slots, 256), (regions, lds)):
            lst.sort()
            for (a, n), (b, _) in zip(lst, lst[1:]):
                assert a + n <= b
            assert not lst or lst[-1][0] + lst[-1][1] <= cap
        mixed_wgs += len(zcs) > 1
    assert np.all(seen == 1)
    assert mixed_wgs > 0
    assert lds * 2 <= 90 * 1024


def test_crc_bit_weights_give_the_reference_crc(oracle_mod):
    """The per-bit weight table behind the kernels' linear CRC test: entry m
    of a CRC type is x^m mod g, and for random byte strings the XOR of the
    set bits' weights (bit i of B weighs x^(B-1-i)), times x^w mod g, equals
    the reference's byte-serial CRC of the string (the oracle's, pinned to
    the reference's vectors).  So a block passes the kernels' zero test
    exactly when its CRC remainder is zero."""
    tabs, off = emu.crc_tables()
    polys = {0: (0x1021, 16), 1: (0x864CFB, 24), 2: (0x800063, 24), 3: (0xB2B117, 24)}
    rng = np.random.default_rng(7)
    for t, (poly, w) in polys.items():
        g = poly | (1 << w)
        weights = tabs[t][off:].astype(np.int64)
        r = 1
        for m in range(weights.size):
            assert weights[m] == r, (t, m)
            r <<= 1
            if (r >> w) & 1:
                r ^= g
        for nbytes in (1, 3, 66, 1056):
            data = rng.integers(0, 256, nbytes, dtype=np.uint8)
            bits = np.unpackbits(data)
            acc = 0
            for i in np.flatnonzero(bits):
                acc ^= int(weights[bits.size - 1 - i])
            for _ in range(w):
                acc <<= 1
                if (acc >> w) & 1:
                    acc ^= g
            assert acc == oracle_mod.crc(data, t), (t, nbytes)
        # a string with its CRC appended has every weight sum zero
        data = rng.integers(0, 256, 40, dtype=np.uint8)
        c = oracle_mod.crc(data, t)
        full = np.concatenate([data, np.array([(c >> s) & 0xFF for s in range(w - 8, -1, -8)],
                                              np.uint8)])
        bits = np.unpackbits(full)
        acc = 0
        for i in np.flatnonzero(bits):
            acc ^= int(weights[bits.size - 1 - i])
        assert acc == 0, t
