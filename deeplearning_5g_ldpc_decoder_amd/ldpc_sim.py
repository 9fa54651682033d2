"""BLER simulation harness with the reference's loader/plugin interface.

Same names, arguments, return values and result-file format as
scripts/ldpc_c_sim_common.py of the reference (load_c_library :60,
make_decoder_config :112, make_ldpc_config :135, make_config_grid :192,
run_ldpc_trial :227, run_bler_sweep :275, run_simulation :309), so a driver
written for the reference runs against libnrldpc_mi355.so by changing the
import (the library path defaults to this package's build).  The per-trial
path calls the reference-compatible ABI one codeblock at a time;
``run_bler_sweep_batched`` is the MI355X path: whole SNR points synthesized
on the host and decoded in one batched launch.
"""
import ctypes
import itertools
import math
import pickle
import warnings
from fractions import Fraction

import numpy as np

from . import _lib
from ._lib import DecoderConfig, LdpcInConfig, ptr

LIB_PATH = _lib.LIB_PATH

VALID_QM = {1, 2, 4, 6, 8}
VALID_RV = {0, 1, 2, 3}
VALID_BGN = {1, 2}
VALID_CRCTYPE = {0, 1, 2, 3}
MAX_B_BY_BGN = {1: 8448, 2: 3840}


class _BoundLibrary:
    """What load_c_library returns: the CDLL plus the three bound calls."""

    def __init__(self, lib, encode_fn, decode_fn, channel_fn, channel_lib):
        self._lib = lib
        self.ldpc_5g_encode = encode_fn
        self.ldpc_5g_decode = decode_fn
        self.tx_rx_process_fn = channel_fn
        self._channel_lib = channel_lib

    def __getattr__(self, name):
        return getattr(self._lib, name)


def load_c_library(lib_path=LIB_PATH, encode_symbol="c_ldpc_5g_encode",
                   decode_symbol="c_ldpc_5g_decode", channel_lib_path=None):
    """Load an LDPC shared library and bind encode/channel/decode by symbol
    name; the channel may come from a second library."""
    from numpy.ctypeslib import ndpointer
    lib = _lib.load() if lib_path == LIB_PATH else ctypes.CDLL(lib_path)
    chan = lib if channel_lib_path is None else ctypes.CDLL(channel_lib_path)
    u8 = ndpointer(dtype=np.uint8, flags="C_CONTIGUOUS")
    i16 = ndpointer(dtype=np.int16, flags="C_CONTIGUOUS")
    # lib[name] is a function object of its own (getattr returns the one the
    # CDLL caches): the ndpointer argument types stay on these bindings and do
    # not leak to other users of the package's shared library handle
    enc = lib[encode_symbol]
    enc.argtypes = [u8, u8, ctypes.POINTER(LdpcInConfig)]
    enc.restype = ctypes.c_int32
    dec = lib[decode_symbol]
    dec.argtypes = [i16, u8, ctypes.POINTER(LdpcInConfig), ctypes.POINTER(DecoderConfig)]
    dec.restype = ctypes.c_int32
    ch = chan["tx_rx_process"]
    ch.argtypes = [u8, ctypes.c_size_t, ctypes.c_uint16, ctypes.c_float, i16]
    ch.restype = ctypes.c_int32
    return _BoundLibrary(lib, enc, dec, ch, chan)


def make_decoder_config(values):
    """Decoder row [algo, max_it, alpha, beta, filler, max_llr, min_llr]."""
    return DecoderConfig(*[int(v) for v in values[:7]])


def format_coderate(coderate, max_denominator=1000):
    f = Fraction(coderate).limit_denominator(max_denominator)
    if abs(float(f) - coderate) < 1e-12:
        return f"{f.numerator}/{f.denominator}"
    return f"{coderate:.4g}"


def get_crc_length(crctype):
    return 16 if crctype == 0 else 24


def make_ldpc_config(Qm, rv, B, coderate, bgn, crctype):
    """Sweep point -> LdpcInConfig (A = B - CRC bits, E = int(B / R))."""
    return LdpcInConfig(bgn=bgn, A=B - get_crc_length(crctype), crctype=crctype,
                        E=int(B / coderate), rv=rv, Qm=Qm, HARQ_on=0)


def get_config_grid_rejection_reason(Qm, rv, B, coderate, bgn, crctype):
    """None when the tuple is a runnable sweep point, else the reason."""
    checks = [
        (Qm not in VALID_QM, f"Qm must be one of {sorted(VALID_QM)}"),
        (rv not in VALID_RV, f"rv must be one of {sorted(VALID_RV)}"),
        (bgn not in VALID_BGN, f"bgn must be one of {sorted(VALID_BGN)}"),
        (crctype not in VALID_CRCTYPE, f"crctype must be one of {sorted(VALID_CRCTYPE)}"),
        (coderate <= 0 or coderate > 1, "coderate must be in the range (0, 1]"),
    ]
    for bad, why in checks:
        if bad:
            return why
    if B <= get_crc_length(crctype):
        return "B must be larger than the selected CRC length"
    if (B - get_crc_length(crctype)) % 8:
        return "A, derived as B minus CRC length, must be a multiple of 8"
    if B > MAX_B_BY_BGN[bgn]:
        return f"BG{bgn} supports B <= {MAX_B_BY_BGN[bgn]}"
    E = int(B / coderate)
    if E < B:
        return "E, derived as int(B / coderate), must be at least B"
    if E % Qm:
        return "E, derived as int(B / coderate), must be a multiple of Qm"
    return None


def make_config_grid(Qm_list, rv_list, B_list, coderate_list, bgn_list, crctype_list,
                     strict=False):
    grid, skipped = [], {}
    for t in itertools.product(Qm_list, rv_list, B_list, coderate_list, bgn_list, crctype_list):
        why = get_config_grid_rejection_reason(*t)
        if why is None:
            grid.append(t)
            continue
        if strict:
            raise ValueError(f"unsupported LDPC config {t}: {why}")
        skipped[why] = skipped.get(why, 0) + 1
    if skipped:
        text = ", ".join(f"{n} skipped because {w}" for w, n in sorted(skipped.items()))
        warnings.warn(f"make_config_grid skipped unsupported LDPC configs: {text}",
                      stacklevel=2)
    if not grid:
        raise ValueError("make_config_grid did not produce any supported configs")
    return grid


def run_ldpc_trial(libnrLDPC, ldpc_config, decoder_config_list, noise_power_db):
    """One random codeblock: encode, channel, decode with every decoder row.
    Returns (block error flags, payload bit-error counts) per row."""
    A, E = ldpc_config.A, ldpc_config.E
    info_bits = np.random.randint(0, 2, A).astype(np.uint8)
    encoded = np.zeros((E + 7) // 8, np.uint8)
    st = libnrLDPC.ldpc_5g_encode(np.packbits(info_bits, bitorder="big"), encoded,
                                  ctypes.byref(ldpc_config))
    if st != 0:
        raise RuntimeError(f"LDPC encode failed with status {st}")
    llr = np.zeros(E, np.int16)
    st = libnrLDPC.tx_rx_process_fn(encoded, E, ldpc_config.Qm,
                                    ctypes.c_float(noise_power_db), llr)
    if st != 0:
        raise RuntimeError(f"tx_rx_process failed with status {st}")
    hard = np.zeros((A + get_crc_length(ldpc_config.crctype) + 7) // 8, np.uint8)
    flags = np.zeros(len(decoder_config_list))
    bit_errors = np.zeros(len(decoder_config_list))
    for i, row in enumerate(decoder_config_list):
        hard[:] = 0
        st = libnrLDPC.ldpc_5g_decode(llr, hard, ctypes.byref(ldpc_config),
                                      ctypes.byref(make_decoder_config(row)))
        flags[i] = 0 if st == 0 else 1
        bit_errors[i] = np.count_nonzero(np.unpackbits(hard, bitorder="big")[:A] != info_bits)
    return flags, bit_errors


def _report(cfg, coderate, snr, bler, ber):
    print(f"Qm {cfg.Qm}, rv {cfg.rv}, A {cfg.A}, E {cfg.E}, "
          f"coderate {format_coderate(coderate)}, bgn {cfg.bgn}, crc {cfg.crctype}, "
          f"snr_power_db {snr:.2f}, BLER={bler}, BER={ber}")


def run_bler_sweep(libnrLDPC, ldpc_config, coderate, decoder_config_list, snr_power_db_list,
                   test_num):
    """BLER/BER per SNR point; noise power = -snr - 10 log10(Qm)."""
    bler_list, ber_list = [], []
    for snr in snr_power_db_list:
        noise_db = -snr - 10 * np.log10(ldpc_config.Qm)
        errs = np.zeros(len(decoder_config_list))
        bits = np.zeros(len(decoder_config_list))
        for _ in range(test_num):
            e, b = run_ldpc_trial(libnrLDPC, ldpc_config, decoder_config_list, noise_db)
            errs += e
            bits += b
        bler, ber = errs / test_num, bits / (test_num * ldpc_config.A)
        _report(ldpc_config, coderate, snr, bler, ber)
        bler_list.append(bler)
        ber_list.append(ber)
    return bler_list, ber_list


def run_bler_sweep_batched(ldpc_config, coderate, decoder_config_list, snr_power_db_list,
                           test_num, seed=0, device="cuda", threads=8):
    """Same results format as run_bler_sweep, decoded as one GPU batch per SNR
    point and decoder row (synthesized inputs from the library's generator)."""
    import torch
    from .batch import BatchPlan, synth_batch
    cfg = tuple(int(getattr(ldpc_config, f)) for f in
                ("bgn", "A", "crctype", "E", "rv", "Qm", "HARQ_on"))
    B = cfg[1] + get_crc_length(cfg[2])
    plans = [BatchPlan([cfg] * test_num, row) for row in decoder_config_list]
    bler_list, ber_list = [], []
    for si, snr in enumerate(snr_power_db_list):
        noise_db = float(-snr - 10 * np.log10(cfg[5]))
        llr, payload = synth_batch(cfg, test_num, noise_db, seed=seed * 1000 + si,
                                   threads=threads)
        d_llr = torch.from_numpy(llr.reshape(-1)).to(device)
        errs = np.zeros(len(plans))
        bits = np.zeros(len(plans))
        for i, plan in enumerate(plans):
            out = plan.decode(d_llr)
            hard = out["hard"].cpu().numpy().reshape(test_num, B // 8)
            st = out["status"].cpu().numpy()
            dec_bits = np.unpackbits(hard, axis=1, bitorder="big")[:, :cfg[1]]
            errs[i] = np.count_nonzero(st != 0)
            bits[i] = np.count_nonzero(dec_bits != payload)
        bler, ber = errs / test_num, bits / (test_num * cfg[1])
        _report(ldpc_config, coderate, snr, bler, ber)
        bler_list.append(bler)
        ber_list.append(ber)
    for p in plans:
        p.close()
    return bler_list, ber_list


def run_simulation(config_grid, decoder_config_list, snr_power_db_list, test_num, result_file,
                   lib_path=LIB_PATH, encode_symbol="c_ldpc_5g_encode",
                   decode_symbol="c_ldpc_5g_decode", channel_lib_path=None, batched=False):
    """Full sweep; pickles (test_result, decoder_config_list, snr_power_db_list,
    test_num) like the reference so its plotting notebooks read the file."""
    lib = None if batched else load_c_library(lib_path, encode_symbol, decode_symbol,
                                              channel_lib_path)
    print(f"Generated {len(config_grid)} configurations.")
    results = []
    for t in config_grid:
        Qm, rv, B, coderate, bgn, crctype = t
        if bgn == 2 and B > 3840:
            continue
        cfg = make_ldpc_config(Qm, rv, B, coderate, bgn, crctype)
        if batched:
            bl, be = run_bler_sweep_batched(cfg, coderate, decoder_config_list,
                                            snr_power_db_list, test_num)
        else:
            bl, be = run_bler_sweep(lib, cfg, coderate, decoder_config_list, snr_power_db_list,
                                    test_num)
        results.append([t, bl, be])
    with open(result_file, "wb") as f:
        pickle.dump((results, decoder_config_list, snr_power_db_list, test_num), f)
    return results


def bler_confidence_interval(errors, trials, z=2.576):
    """Wilson score interval for a BLER estimate (99 % by default)."""
    if trials == 0:
        return 0.0, 1.0
    p = errors / trials
    den = 1 + z * z / trials
    centre = (p + z * z / (2 * trials)) / den
    half = z * math.sqrt(p * (1 - p) / trials + z * z / (4 * trials * trials)) / den
    return max(0.0, centre - half), min(1.0, centre + half)
