"""CPU tests of libnrldpc_mi355.so's host half and its C ABI surface.

No GPU call is made here: only symbol export checks and the host entry
points (parameters, CRC, encoder, test channel).
"""
import ctypes
import os
import re
import subprocess

import numpy as np
import pytest

from conftest import GOLDEN, ROOT

from deeplearning_5g_ldpc_decoder_amd import _lib
from deeplearning_5g_ldpc_decoder_amd._lib import LdpcConfig, LdpcInConfig, ptr

HEADER = os.path.join(ROOT, "include", "nr_ldpc_mi355.h")


@pytest.fixture(scope="module")
def L():
    return _lib.load()


def _declared_symbols():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    funcs = set(re.findall(r"\b(\w+)\s*\([^;{]*\)\s*;", text))
    funcs -= {"sizeof"}
    data = set(re.findall(r"extern\s+\w+\s+(\w+)\s*\[", text))
    return funcs, data


def test_every_declared_symbol_is_exported(L):
    funcs, data = _declared_symbols()
    assert "c_ldpc_5g_decode" in funcs and "HARQ_memory" in data
    out = subprocess.run(["nm", "-D", "--defined-only", _lib.LIB_PATH], capture_output=True,
                         text=True, check=True).stdout
    exported = {ln.split()[-1] for ln in out.splitlines() if ln.strip()}
    missing = (funcs | data) - exported
    assert not missing, missing
    assert set(_lib.EXPORTED_FUNCTIONS) == funcs
    assert set(_lib.EXPORTED_DATA) == data
    for name in funcs:
        getattr(L, name)


def test_struct_layouts():
    assert ctypes.sizeof(LdpcInConfig) == 14
    assert ctypes.sizeof(_lib.DecoderConfig) == 14
    assert ctypes.sizeof(LdpcConfig) == 56
    assert ctypes.sizeof(_lib.CbDescriptor) == 24


def test_gen_ldpc_parameters_matches_reference_table(L):
    rows = np.load(os.path.join(GOLDEN, "params_table.npz"))["rows"]
    bad = 0
    for r in rows[::3]:
        cfg = [int(v) for v in r[:7]]
        pc = LdpcConfig()
        st = L.gen_ldpc_parameters(ctypes.byref(pc), ctypes.byref(LdpcInConfig(*cfg)))
        if st != int(r[7]):
            bad += 1
        elif st == 0 and [pc.B, pc.Zc, pc.iLS, pc.K, pc.F, pc.N, pc.k0] != \
                [int(v) for v in r[8:15]]:
            bad += 1
    assert bad == 0


def test_gen_ldpc_parameters_table_pointers(L):
    """Pointers returned into the exported tables (used by
    deep_learning/ldpc_table_config.py:129-131)."""
    pc = LdpcConfig()
    assert L.gen_ldpc_parameters(ctypes.byref(pc), ctypes.byref(
        LdpcInConfig(1, 8424, 2, 25344, 0, 2, 0))) == 0
    off = [pc.HBG_offset_array[i] for i in range(47)]
    assert off[0] == 0 and off[-1] == 316
    vij = (ctypes.c_uint16 * (8 * 316)).in_dll(L, "BG1_HBG_Vij_array")
    assert ctypes.addressof(pc.Vij_array.contents) == ctypes.addressof(vij) + 2 * 316 * pc.iLS


def test_tables_equal_oracle_tables(L, oracle_mod):
    """Product tables (from the 3GPP .mat data) == oracle tables (from the
    reference's C table text)."""
    for bgn, rows, nnz in ((1, 46, 316), (2, 42, 197)):
        off = np.ctypeslib.as_array((ctypes.c_uint16 * (rows + 1)).in_dll(
            L, f"BG{bgn}_HBG_offset_array"))
        col = np.ctypeslib.as_array((ctypes.c_uint16 * nnz).in_dll(L, f"BG{bgn}_HBG_array"))
        vij = np.ctypeslib.as_array((ctypes.c_uint16 * (8 * nnz)).in_dll(
            L, f"BG{bgn}_HBG_Vij_array"))
        for ils in range(8):
            zc = {1: 384, 2: 208}[bgn] if ils in (1, 6) else None
            A = 24 if bgn == 2 else 8424
            c = oracle_mod.code(bgn, A, 0, 3000, 0, 2)
            ro = np.ctypeslib.as_array(ctypes.cast(c.row_off, ctypes.POINTER(ctypes.c_uint16)),
                                       (rows + 1,))
            co = np.ctypeslib.as_array(ctypes.cast(c.col, ctypes.POINTER(ctypes.c_uint16)),
                                       (nnz,))
            assert np.array_equal(off, ro) and np.array_equal(col, co)
            del zc
        # full shift tables against the oracle's per-iLS slices
        for ils in range(8):
            base = ctypes.cast(c.vij, ctypes.c_void_p).value - 2 * nnz * c.iLS
            ov = np.ctypeslib.as_array(
                ctypes.cast(base + 2 * nnz * ils, ctypes.POINTER(ctypes.c_uint16)), (nnz,))
            assert np.array_equal(vij[ils * nnz:(ils + 1) * nnz], ov)


def test_calculate_crc_vectors(L):
    z = np.load(os.path.join(GOLDEN, "crc_vectors.npz"))
    o = 0
    for n, t, want in zip(z["lens"], z["crctype"], z["crc"]):
        data = np.ascontiguousarray(z["data"][o:o + n])
        o += n
        assert L.calculate_crc(ptr(data), int(n), int(t)) == int(want)


def test_dequantize(L):
    for v in (-32768, -1, 0, 1, 128, 32767):
        assert L.dequantize_int16_to_llr(v) == np.float32(v) / np.float32(128)


@pytest.mark.parametrize("symbol", ["c_ldpc_5g_encode", "cuda_ldpc_5g_encode"])
def test_encoder_matches_reference(L, symbol):
    """Both encoder symbols the reference's simulation drivers bind
    (scripts/sim_c_ldpc_encoder_decoder.py, sim_cuda_ldpc_encoder_decoder.py:65)
    against the reference encoder's outputs."""
    encode = getattr(L, symbol)
    z = np.load(os.path.join(GOLDEN, "encoder_vectors.npz"))
    po = ro = 0
    for m in z["meta"]:
        bgn, A, crctype, E, rv, Qm = (int(v) for v in m)
        payload = z["payload"][po:po + A]
        po += A
        nb = (E + 7) // 8
        want = z["rm"][ro:ro + nb]
        ro += nb
        out = np.zeros(nb, np.uint8)
        info = np.packbits(payload)
        st = encode(ptr(info), ptr(out), ctypes.byref(LdpcInConfig(bgn, A, crctype, E, rv, Qm, 0)))
        assert st == 0
        # compare the E valid bits (the last byte's padding is unspecified)
        assert np.array_equal(np.unpackbits(out)[:E], np.unpackbits(want)[:E]), m


def test_encoder_invalid_config(L):
    out = np.zeros(8, np.uint8)
    info = np.zeros(8, np.uint8)
    assert L.c_ldpc_5g_encode(ptr(info), ptr(out), ctypes.byref(
        LdpcInConfig(3, 8, 0, 64, 0, 2, 0))) == -2


@pytest.mark.parametrize("Qm", [1, 2, 4, 6, 8])
def test_channel_noiseless_signs(L, Qm):
    """At very high SNR the LLR sign is the transmitted bit (0 -> positive)."""
    rng = np.random.default_rng(Qm)
    E = 240 * Qm
    bits = rng.integers(0, 2, E).astype(np.uint8)
    llr = np.zeros(E, np.int16)
    L.ldpc_mi355_channel_seed(7)
    assert L.tx_rx_process(ptr(np.packbits(bits)), E, Qm, ctypes.c_float(-60.0), ptr(llr)) == 0
    assert np.all((llr < 0) == (bits == 1))
    assert np.abs(llr).max() == 10000


def test_channel_matches_oracle_maxlog_at_zero_noise_limit(L, oracle_mod):
    """Max-log magnitudes: QPSK LLR = 4*A*r/N0 (c_lib/tx_rx_process.c:240-242)."""
    bits = np.array([0, 1, 1, 0] * 10, np.uint8)
    llr = np.zeros(len(bits), np.int16)
    nv_db = 20.0  # large noise power: values stay inside the +-10000 clip
    L.tx_rx_process(ptr(np.packbits(bits)), len(bits), 2, ctypes.c_float(nv_db), ptr(llr))
    assert np.abs(llr).max() < 10000
    assert L.tx_rx_process(ptr(np.packbits(bits)), 3, 2, ctypes.c_float(0.0), ptr(llr)) == -1
    assert L.tx_rx_process(ptr(np.packbits(bits)), 4, 3, ctypes.c_float(0.0), ptr(llr)) == -1


def test_synth_batch_codewords(oracle_mod):
    """Synthesized LLRs decode (oracle) to the returned payloads at high SNR
    and are independent of the thread count."""
    from deeplearning_5g_ldpc_decoder_amd.batch import synth_batch
    cfg = (2, 1016, 2, 3120, 0, 2, 0)
    llr1, pay1 = synth_batch(cfg, 6, -8.0, seed=3, threads=1)
    llr4, pay4 = synth_batch(cfg, 6, -8.0, seed=3, threads=4)
    assert np.array_equal(llr1, llr4) and np.array_equal(pay1, pay4)
    for i in range(6):
        out = oracle_mod.decode_path(llr1[i], cfg, [1, 8, 22936, 0, 1280, 32767, -32767])
        assert out["status"] == 0
        assert np.array_equal(np.unpackbits(out["hard"])[:1016], pay1[i])


def test_default_scale_tables_equal_reference_defaults(oracle_mod):
    """The library's algo-2 / algo-3 defaults equal the tables the reference
    ships (its trained-table fixtures carry BG1/BG2 default tables as decoded
    by the from-source reference build)."""
    from deeplearning_5g_ldpc_decoder_amd import tables
    z = np.load(os.path.join(GOLDEN, "decode_algo2.npz"))
    for bgn in (1, 2):
        a, b = tables.get_trained_tables(bgn)
        assert a.shape == (32, 46 if bgn == 1 else 42)
        assert np.array_equal(a, z[f"alpha_bg{bgn}"][0])
        assert np.array_equal(b, z[f"beta_bg{bgn}"][0])
    d = np.load(os.path.join(GOLDEN, "decode_algo3.npz"))["degree_alpha"][0]
    assert np.array_equal(tables.get_degree_alpha(), d)
    assert np.array_equal(d, oracle_mod.default_degree_alpha())


def test_scale_table_setters_round_trip(L):
    from deeplearning_5g_ldpc_decoder_amd import tables
    rng = np.random.default_rng(3)
    try:
        a = rng.integers(-32768, 32768, (7, 42)).astype(np.int16)
        b = rng.integers(-32768, 32768, (7, 42)).astype(np.int16)
        tables.set_trained_tables(2, a, b)
        ga, gb = tables.get_trained_tables(2)
        assert np.array_equal(ga, a) and np.array_equal(gb, b)
        assert tables.get_trained_tables(1)[0].shape == (32, 46)   # BG1 untouched
        with pytest.raises(ValueError):
            tables.set_trained_tables(2, a[:, :41], b[:, :41])
        assert L.ldpc_mi355_set_trained_tables(3, ptr(a), ptr(b), 7) == -4
        assert L.ldpc_mi355_set_trained_tables(2, ptr(a), ptr(b), 0) == -4
        assert L.ldpc_mi355_set_trained_tables(2, ptr(a), ptr(b), 1025) == -4
        assert L.ldpc_mi355_set_trained_tables(2, ptr(a), None, 7) == -4
        assert L.ldpc_mi355_get_trained_tables(0, None, None, 0) == -4
        v = rng.integers(-32768, 32768, 19).astype(np.int16)
        tables.set_degree_alpha(v)
        assert np.array_equal(tables.get_degree_alpha(), v)
        assert L.ldpc_mi355_get_degree_alpha(None) == -4
    finally:
        tables.set_trained_tables(2)
        tables.set_degree_alpha()
    assert tables.get_trained_tables(2)[0].shape == (32, 42)


def test_copy_spans_argument_checks(L):
    """ldpc_mi355_copy_spans refuses bad arguments before any HIP call, and
    a call with no spans does nothing."""
    assert L.ldpc_mi355_copy_spans(None, None, None, 0, 1, None) == 0
    assert L.ldpc_mi355_copy_spans(None, None, None, 0, 2, None) == 0
    assert L.ldpc_mi355_copy_spans(None, None, None, -1, 1, None) == -1
    assert L.ldpc_mi355_copy_spans(None, None, None, 0, 3, None) == -1   # unknown kind
    assert L.ldpc_mi355_copy_spans(None, None, None, 2, 1, None) == -1   # no span arrays
