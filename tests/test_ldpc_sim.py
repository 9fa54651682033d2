"""The simulation harness mirror (deeplearning_5g_ldpc_decoder_amd.ldpc_sim)
against the behaviour the reference pins for scripts/ldpc_c_sim_common.py:
make_config_grid's filtering, warning and strict modes
(test/test_ldpc_config_grid.py:14-60), the rejection reasons (:158-224), and
run_simulation's trial loop and result pickle
(test_result, decoder_config_list, snr_power_db_list, test_num) (:309-343).
CPU only: run_simulation runs here on a stub library with the same three
bound calls; tests/test_gpu_sim.py runs it on the real library."""
import ctypes
import pickle

import numpy as np
import pytest

from deeplearning_5g_ldpc_decoder_amd import ldpc_sim as S


def test_make_config_grid_filters_unsupported_product_pairs():
    with pytest.warns(UserWarning, match="BG2 supports B <= 3840"):
        grid = S.make_config_grid(Qm_list=[1], rv_list=[0], B_list=[3840, 4096],
                                  coderate_list=[1 / 2], bgn_list=[1, 2], crctype_list=[1])
    assert grid == [(1, 0, 3840, 1 / 2, 1, 1), (1, 0, 3840, 1 / 2, 2, 1),
                    (1, 0, 4096, 1 / 2, 1, 1)]


def test_make_config_grid_raises_when_every_config_is_unsupported():
    with pytest.warns(UserWarning, match="Qm must be one of"):
        with pytest.raises(ValueError,
                           match="make_config_grid did not produce any supported configs"):
            S.make_config_grid(Qm_list=[10], rv_list=[0], B_list=[256], coderate_list=[1 / 2],
                               bgn_list=[1], crctype_list=[1])


def test_make_config_grid_strict_mode_raises_on_first_unsupported_config():
    with pytest.raises(ValueError, match="unsupported LDPC config"):
        S.make_config_grid(Qm_list=[1], rv_list=[0], B_list=[4096], coderate_list=[1 / 2],
                           bgn_list=[2], crctype_list=[1], strict=True)


@pytest.mark.parametrize("t,why", [
    ((3, 0, 256, 0.5, 1, 1), "Qm must be one of [1, 2, 4, 6, 8]"),
    ((2, 4, 256, 0.5, 1, 1), "rv must be one of [0, 1, 2, 3]"),
    ((2, 0, 256, 0.5, 3, 1), "bgn must be one of [1, 2]"),
    ((2, 0, 256, 0.5, 1, 4), "crctype must be one of [0, 1, 2, 3]"),
    ((2, 0, 256, 1.5, 1, 1), "coderate must be in the range (0, 1]"),
    ((2, 0, 24, 0.5, 1, 1), "B must be larger than the selected CRC length"),
    ((2, 0, 260, 0.5, 1, 1), "A, derived as B minus CRC length, must be a multiple of 8"),
    ((2, 0, 8472, 0.5, 1, 1), "BG1 supports B <= 8448"),
    ((8, 0, 256, 0.3, 1, 1), "E, derived as int(B / coderate), must be a multiple of Qm"),
    ((2, 0, 256, 0.5, 1, 1), None),
    ((1, 0, 40, 1.0, 2, 0), None),
])
def test_rejection_reasons(t, why):
    assert S.get_config_grid_rejection_reason(*t) == why


def test_make_ldpc_config_and_coderate_format():
    c = S.make_ldpc_config(4, 2, 3840, 1 / 3, 1, 1)
    assert (c.bgn, c.A, c.crctype, c.E, c.rv, c.Qm, c.HARQ_on) == (1, 3816, 1, 11520, 2, 4, 0)
    assert S.format_coderate(1 / 3) == "1/3" and S.format_coderate(0.123456789) == "0.1235"


class _StubLibrary:
    """The three calls load_c_library binds, on the CPU: encode copies the
    payload bits (MSB-first) into the first A rate-matched bits, the channel
    maps bit b to LLR +-1000 and flips every bit when noise_power_db > 0, and
    a decode passes (0) when no LLR is negative and returns the payload read
    from the LLR signs."""

    def __init__(self):
        self.calls = []

    def ldpc_5g_encode(self, info, out, cfg):
        cfg = cfg._obj
        bits = np.unpackbits(info)[:cfg.A]
        e = np.zeros(cfg.E, np.uint8)
        e[:cfg.A] = bits
        out[:] = np.packbits(e)
        return 0

    def tx_rx_process_fn(self, encoded, E, Qm, noise_db, llr):
        bits = np.unpackbits(encoded)[:E].astype(np.int16)
        if noise_db.value > 0:
            bits ^= 1
        llr[:] = 1000 - 2000 * bits
        return 0

    def ldpc_5g_decode(self, llr, hard, cfg, dec):
        cfg, dec = cfg._obj, dec._obj
        self.calls.append(int(dec.algo))
        bits = (llr[:cfg.A] < 0).astype(np.uint8)
        nb = (cfg.A + 7) // 8
        hard[:nb] = np.packbits(bits)
        return 0 if not bits.any() else 1


def test_run_simulation_pickle_matches_reference_format(tmp_path, monkeypatch):
    stub = _StubLibrary()
    monkeypatch.setattr(S, "load_c_library", lambda *a, **k: stub)
    np.random.seed(5)
    grid = [(2, 0, 256, 1 / 2, 1, 1), (2, 0, 4096, 1 / 2, 2, 1), (4, 1, 512, 1 / 3, 2, 0)]
    rows = [[1, 8, 22936, 0, 1280, 32767, -32767], [4, 20, 22936, 0, 1280, 32767, -32767]]
    snr = [-1.0, -10.0]   # noise power -1 - 10log10(Qm) < 0 (clean), 10 - ... > 0 (flipped)
    out = tmp_path / "res.pkl"
    res = S.run_simulation(grid, rows, snr, 7, str(out))
    with open(out, "rb") as f:
        test_result, dec_list, snr_list, test_num = pickle.load(f)
    assert dec_list == rows and snr_list == snr and test_num == 7
    assert len(test_result) == len(res)
    for a, b in zip(test_result, res):
        assert a[0] == b[0]
        assert all(np.array_equal(x, y) for x, y in zip(a[1] + a[2], b[1] + b[2]))
    # the BG2 B > 3840 tuple is skipped, as in the reference
    assert [r[0] for r in test_result] == [grid[0], grid[2]]
    for cfg_t, bler, ber in test_result:
        assert len(bler) == len(snr) and len(ber) == len(snr)
        # clean point: a block fails exactly when its payload has a 1 bit, and
        # its bit errors are 0; flipped point: every bit of every block wrong
        assert np.all(ber[0] == 0) and np.all(ber[1] == 1.0)
        assert np.all(bler[1] == 1.0)
        assert bler[0].shape == (2,)
    assert stub.calls[:2] == [1, 4]


def test_load_c_library_leaves_shared_bindings_alone():
    """load_c_library binds the harness's ndpointer argument types on function
    objects of its own: the package's shared library handle keeps its
    pointer-typed bindings for every other caller."""
    from deeplearning_5g_ldpc_decoder_amd import _lib
    L = _lib.load()
    before = {n: list(getattr(L, n).argtypes) for n in
              ("c_ldpc_5g_decode", "c_ldpc_5g_encode", "tx_rx_process")}
    lib = S.load_c_library()
    assert lib.ldpc_5g_decode.argtypes[0] is not ctypes.c_void_p
    for n, types in before.items():
        assert list(getattr(L, n).argtypes) == types, n
