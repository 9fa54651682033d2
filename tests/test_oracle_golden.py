"""Pin the oracle (CPU restatement) against the reference's own outputs.

Fixtures come from the reference C library built from source
(tests/golden/gen_golden.py); these tests run on the CPU only.
"""
import os

import numpy as np
import pytest

from conftest import GOLDEN, case_id, load_decode_cases

ALGO1 = load_decode_cases("decode_algo1.npz")
ALGO4 = load_decode_cases("decode_algo4.npz")
ALGO0 = load_decode_cases("decode_algo0.npz")
ALGO2 = load_decode_cases("decode_algo2.npz")
ALGO3 = load_decode_cases("decode_algo3.npz")


def test_params_table_matches_reference(oracle_mod):
    rows = np.load(os.path.join(GOLDEN, "params_table.npz"))["rows"]
    mismatches = 0
    for r in rows:
        cfg = [int(v) for v in r[:7]]
        st, p = oracle_mod.params(*cfg)
        if st != int(r[7]):
            mismatches += 1
            continue
        if st == 0 and [p[k] for k in ("B", "Zc", "iLS", "K", "F", "N", "k0")] != \
                [int(v) for v in r[8:15]]:
            mismatches += 1
    assert mismatches == 0
    assert len(rows) > 30000


def test_crc_vectors(oracle_mod):
    z = np.load(os.path.join(GOLDEN, "crc_vectors.npz"))
    o = 0
    for n, t, want in zip(z["lens"], z["crctype"], z["crc"]):
        data = z["data"][o:o + n]
        o += n
        assert oracle_mod.crc(data, int(t)) == int(want)


def test_crc_linearity_and_codeword(oracle_mod):
    rng = np.random.default_rng(5)
    for t in range(4):
        a = rng.integers(0, 256, 100).astype(np.uint8)
        b = rng.integers(0, 256, 100).astype(np.uint8)
        assert oracle_mod.crc(a ^ b, t) == oracle_mod.crc(a, t) ^ oracle_mod.crc(b, t)
        # payload with its CRC appended divides to zero
        bits = rng.integers(0, 2, 800).astype(np.uint8)
        full = oracle_mod.attach_crc(bits, t)
        assert oracle_mod.crc(np.packbits(full), t) == 0
        # single bit error is always detected
        full[rng.integers(0, len(full))] ^= 1
        assert oracle_mod.crc(np.packbits(full), t) != 0


def test_encoder_matches_reference(oracle_mod):
    z = np.load(os.path.join(GOLDEN, "encoder_vectors.npz"))
    po = ro = 0
    for m in z["meta"]:
        bgn, A, crctype, E, rv, Qm = (int(v) for v in m)
        payload = z["payload"][po:po + A]
        po += A
        nb = (E + 7) // 8
        want = np.unpackbits(z["rm"][ro:ro + nb])[:E]
        ro += nb
        c = oracle_mod.code(bgn, A, crctype, E, rv, Qm)
        cw = oracle_mod.encode(c, oracle_mod.attach_crc(payload, crctype))
        assert oracle_mod.syndrome_ok(c, cw), m
        assert np.array_equal(oracle_mod.rate_match(c, cw), want), m


def test_rate_recovery_matches_reference(oracle_mod):
    z = np.load(os.path.join(GOLDEN, "rate_recovery.npz"))
    lo = so = 0
    prev = None
    for m in z["meta"]:
        bgn, A, crctype, E, rv, Qm, harq_on, filler = (int(v) for v in m)
        c = oracle_mod.code(bgn, A, crctype, E, rv, Qm, harq_on)
        n = c.N + 2 * c.Zc
        llr = z["llr"][lo:lo + E]
        lo += E
        want = z["soft"][so:so + n]
        so += n
        got = oracle_mod.rate_recover(c, filler, llr, prev if harq_on else None)
        assert np.array_equal(got, want), m
        prev = got


def test_rate_recovery_closed_form(oracle_mod):
    """Gather form: compact position q receives k = q - k0' (mod N-F)."""
    rng = np.random.default_rng(3)
    for cfg in [(1, 8424, 2, 25344, 0, 2), (2, 16, 3, 9000, 1, 6), (1, 488, 1, 1536, 3, 4)]:
        c = oracle_mod.code(*cfg)
        llr = rng.integers(-3000, 3000, c.E).astype(np.int16)
        soft = oracle_mod.rate_recover(c, 77, llr)
        Zc, N, F, B, K, E, Qm = c.Zc, c.N, c.F, c.B, c.K, c.E, c.Qm
        S = E // Qm
        nc = N - F
        lo, hi = B - 2 * Zc, K - 2 * Zc
        ck0 = c.k0 if c.k0 < lo else (lo if c.k0 < hi else c.k0 - F)
        want = np.zeros(N + 2 * Zc, np.int64)
        k = np.arange(E)
        q = (ck0 + k) % nc
        pos = np.where(q < lo, q, q + F)
        np.add.at(want, pos + 2 * Zc, llr[(k % S) * Qm + k // S].astype(np.int64))
        want[B:K] = 77
        wrapped = ((want + 32768) % 65536 - 32768).astype(np.int16)
        assert np.array_equal(soft, wrapped)


@pytest.mark.parametrize("case", ALGO1, ids=case_id)
def test_algo1_matches_reference(oracle_mod, case):
    out = oracle_mod.decode_path(case["llr"], case["cfg"], case["row"], harq=case["harq_in"])
    assert out["status"] == case["status"]
    assert np.array_equal(out["soft"], case["soft"])
    assert np.array_equal(out["hard"], case["hard"]) or case["row"][1] == 0
    assert out["iters"] == case["iters"]
    assert np.array_equal(out["lq"], case["lq"])


@pytest.mark.parametrize("case", ALGO2 + ALGO3, ids=case_id)
def test_table_algos_match_reference(oracle_mod, case):
    """Algo 2 with the reference's default tables and with two exported table
    sets linked into the reference build; algo 3 with the default and custom
    degree alpha tables (layered_degree_minsum_ldpc_decode_with_alpha)."""
    out = oracle_mod.decode_path(case["llr"], case["cfg"], case["row"], harq=case["harq_in"],
                                 tables=case.get("tables"),
                                 degree_alpha=case.get("degree_alpha"))
    assert out["status"] == case["status"]
    assert np.array_equal(out["soft"], case["soft"])
    assert np.array_equal(out["hard"], case["hard"]) or case["row"][1] == 0
    assert out["iters"] == case["iters"]
    assert np.array_equal(out["lq"], case["lq"])


def test_table_fixtures_cover_the_table_kinds():
    """Default, smooth and adversarial trained tables; converged and
    non-converged decodes; default and custom degree tables."""
    assert {c["table_id"] for c in ALGO2} == {0, 1, 2}
    assert {c["table_id"] for c in ALGO3} == {-1, 1, 2}
    for cs in (ALGO2, ALGO3):
        st = [c["status"] for c in cs]
        assert st.count(0) >= 10 and st.count(1) >= 10
    # the reference's default trained tables are the degree rule, beta 0
    t = ALGO2[0]["tables"] if ALGO2[0]["table_id"] == 0 else None
    assert t is not None and not t[1].any()


def test_uniform_tables_reduce_to_algo1(oracle_mod):
    """An algo-2 table holding one (alpha, beta) pair everywhere, and an
    algo-3 degree table holding one alpha, decode exactly like algo 1 with
    that pair (algo 1 is pinned by its own fixtures)."""
    for case in ALGO1[8:30]:
        row = case["row"]
        if row[1] == 0:
            continue
        bgn = case["cfg"][0]
        rows = 46 if bgn == 1 else 42
        a = np.full((max(row[1], 1), rows), np.int16(np.uint16(row[2]).view(np.int16)))
        b = np.full_like(a, np.int16(np.uint16(row[3]).view(np.int16)))
        o2 = oracle_mod.decode_path(case["llr"], case["cfg"], [2] + row[1:],
                                    harq=case["harq_in"], tables=(a, b))
        o3 = oracle_mod.decode_path(case["llr"], case["cfg"], [3] + row[1:],
                                    harq=case["harq_in"], degree_alpha=np.full(19, a[0, 0]))
        for o in (o2, o3):
            assert o["status"] == case["status"] and o["iters"] == case["iters"]
            assert np.array_equal(o["lq"], case["lq"])


def test_algo2_refuses_iterations_past_its_table(oracle_mod):
    case = ALGO2[0]
    a, b = oracle_mod.default_trained_tables(case["cfg"][0], 8)
    out = oracle_mod.decode_path(case["llr"], case["cfg"], [2, 9, 0, 0, 1280, 32767, -32767],
                                 tables=(a, b))
    assert out["status"] == -2


@pytest.mark.parametrize("case", ALGO4 + ALGO0, ids=case_id)
def test_float_algos_match_reference(oracle_mod, case):
    out = oracle_mod.decode_path(case["llr"], case["cfg"], case["row"])
    assert out["status"] == case["status"]
    assert out["iters"] == case["iters"]
    if case["row"][1] > 0:
        assert np.array_equal(out["hard"], case["hard"])
    # algo 4 is IEEE single ops only; algo 0 uses the same glibc tanhf/atanhf
    assert np.array_equal(out["lq"].view(np.uint32), case["lq"].view(np.uint32))


def test_cuda_reference_cases_decode_cleanly():
    """The reference CUDA test's 8 configs all decode with status 0."""
    cuda = [c for c in ALGO1 if c["row"][1] == 32 and c["cfg"][5] == 1]
    assert len(cuda) == 8
    assert all(c["status"] == 0 for c in cuda)
