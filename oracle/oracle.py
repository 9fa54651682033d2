"""ORACLE — test infrastructure only.

ctypes front end to ``oracle/liboracle.so`` (the CPU restatement of the
reference decode path, see ``ldpc_oracle.h``) plus numpy test-input
generators.  Only ``tests/``, ``__graft_entry__.smoke()`` and the CPU-baseline
leg of ``bench.py`` may import this module, and only as the checker.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")


class OrcDec(ctypes.Structure):
    _fields_ = [
        ("algo", ctypes.c_uint16),
        ("max_iterations", ctypes.c_uint16),
        ("alpha", ctypes.c_uint16),
        ("beta", ctypes.c_uint16),
        ("filling_bit_value", ctypes.c_int16),
        ("max_llr", ctypes.c_int16),
        ("min_llr", ctypes.c_int16),
    ]


class OrcCode(ctypes.Structure):
    _fields_ = [(n, ctypes.c_int) for n in (
        "bgn", "A", "crctype", "E", "rv", "Qm", "harq_on",
        "B", "Zc", "iLS", "K", "F", "N", "k0", "rows", "nnz")] + [
        ("row_off", ctypes.c_void_p), ("col", ctypes.c_void_p),
        ("vij", ctypes.c_void_p)]


def build():
    """Compile liboracle.so (gcc); the oracle is built, never shipped."""
    subprocess.run(["make", "-s", "-C", HERE], check=True)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        P = ctypes.c_void_p
        L.orc_code_init.argtypes = [ctypes.POINTER(OrcCode)] + [ctypes.c_int] * 7
        L.orc_tx_rx.argtypes = [P, ctypes.c_int, ctypes.c_int, ctypes.c_float, ctypes.c_uint,
                                P, P, P, P]
        L.orc_params.argtypes = [ctypes.c_int] * 7 + [P]
        L.orc_crc.argtypes = [P, ctypes.c_int, ctypes.c_int]
        L.orc_crc.restype = ctypes.c_uint32
        L.orc_rate_recover.argtypes = [ctypes.POINTER(OrcCode), ctypes.c_int16, P, P]
        for name in ("orc_decode_i16", "orc_decode_f32_ms", "orc_decode_f32_bp",
                     "orc_decode_f64_ms", "orc_decode_f64_bp"):
            getattr(L, name).argtypes = [ctypes.POINTER(OrcCode),
                                         ctypes.POINTER(OrcDec), P, P, P, P]
        L.orc_decode_path.argtypes = [P, P] + [ctypes.c_int] * 7 + [
            ctypes.POINTER(OrcDec), P, P, P, P]
        L.orc_decode_path_tables.argtypes = [P, P] + [ctypes.c_int] * 7 + [
            ctypes.POINTER(OrcDec), P, P, P, P, P, P, ctypes.c_int, P]
        L.orc_decode_i16_table.argtypes = [ctypes.POINTER(OrcCode), ctypes.POINTER(OrcDec),
                                           P, P, P, P, P, P, ctypes.c_int]
        L.orc_decode_i16_degree.argtypes = [ctypes.POINTER(OrcCode), ctypes.POINTER(OrcDec),
                                            P, P, P, P, P]
        L.orc_default_degree_alpha.argtypes = [ctypes.c_int]
        L.orc_default_degree_alpha.restype = ctypes.c_int16
        L.orc_default_trained_tables.argtypes = [ctypes.c_int, ctypes.c_int, P, P]
        L.orc_check_minsum_i16.argtypes = [ctypes.c_int, P, ctypes.c_int16,
                                           ctypes.c_int16, ctypes.c_int16, P]
        L.orc_check_bp_f64.argtypes = [ctypes.c_int, P, P]
        L.orc_encode.argtypes = [ctypes.POINTER(OrcCode), P, P]
        L.orc_rate_match.argtypes = [ctypes.POINTER(OrcCode), P, P]
        L.orc_syndrome_ok.argtypes = [ctypes.POINTER(OrcCode), P]
        L.orc_decode_batch_i16.argtypes = [P, P, P, P] + [ctypes.c_int] * 7 + [
            ctypes.POINTER(OrcDec), ctypes.c_int]
        L.orc_decode_batch_i16.restype = ctypes.c_double
        for sfx, ft in (("f32", ctypes.c_float), ("f64", ctypes.c_double)):
            getattr(L, "orc_unrolled_forward_" + sfx).argtypes = [
                ctypes.POINTER(OrcCode), ctypes.c_int, P, P, P, ft, P, P, P]
            getattr(L, "orc_unrolled_backward_" + sfx).argtypes = [
                ctypes.POINTER(OrcCode), ctypes.c_int, P, P, P, ft, ctypes.c_int,
                P, P, P, P, P]
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def dec(row):
    """Decoder row [algo, max_it, alpha, beta, filler, max_llr, min_llr]."""
    return OrcDec(*[int(v) for v in row])


def code(bgn, A, crctype, E, rv=0, Qm=2, harq_on=0):
    c = OrcCode()
    st = lib().orc_code_init(ctypes.byref(c), bgn, A, crctype, E, rv, Qm, harq_on)
    if st != 0:
        raise ValueError(f"invalid LDPC config bgn={bgn} A={A} crc={crctype} "
                         f"E={E} rv={rv} Qm={Qm} harq={harq_on}")
    return c


def params(bgn, A, crctype, E, rv=0, Qm=2, harq_on=0):
    out = np.zeros(7, np.int32)
    st = lib().orc_params(bgn, A, crctype, E, rv, Qm, harq_on, _p(out))
    return st, dict(zip(("B", "Zc", "iLS", "K", "F", "N", "k0"), out.tolist()))


def crc(data_bytes, crctype):
    b = np.ascontiguousarray(data_bytes, dtype=np.uint8)
    return int(lib().orc_crc(_p(b), len(b), crctype))


def rate_recover(c, filler, llr, soft_in=None):
    soft = np.zeros(c.N + 2 * c.Zc, np.int16) if soft_in is None else \
        np.array(soft_in, dtype=np.int16, copy=True)
    llr = np.ascontiguousarray(llr, dtype=np.int16)
    lib().orc_rate_recover(ctypes.byref(c), filler, _p(llr), _p(soft))
    return soft


_DEC_FN = {
    ("i16", 1): ("orc_decode_i16", np.int16),
    ("f32", 4): ("orc_decode_f32_ms", np.float32),
    ("f32", 0): ("orc_decode_f32_bp", np.float32),
    ("f64", 4): ("orc_decode_f64_ms", np.float64),
    ("f64", 0): ("orc_decode_f64_bp", np.float64),
}


MAX_DEGREE = 19
NUM_ROWS = {1: 46, 2: 42}


def default_degree_alpha():
    """The reference's algo-3 defaults, entry d-1 for check degree d."""
    return np.array([lib().orc_default_degree_alpha(d) for d in range(1, MAX_DEGREE + 1)],
                    np.int16)


def default_trained_tables(bgn, iterations=32):
    """The reference's default algo-2 tables: (alpha, beta) int16
    [iterations, rows]."""
    a = np.zeros((iterations, NUM_ROWS[bgn]), np.int16)
    b = np.zeros_like(a)
    lib().orc_default_trained_tables(bgn, iterations, _p(a), _p(b))
    return a, b


def _tables_for(bgn, tables):
    if tables is None:
        tables = default_trained_tables(bgn)
    a = np.ascontiguousarray(tables[0], np.int16)
    b = np.ascontiguousarray(tables[1], np.int16)
    if a.shape != b.shape or a.ndim != 2 or a.shape[1] != NUM_ROWS[bgn]:
        raise ValueError("algo-2 tables must be [iterations, rows] int16 pairs")
    return a, b


def decode_soft(c, row, soft, precision=None, tables=None, degree_alpha=None):
    """Decode a rate-recovered buffer.  Returns (status, hard, lq, iters).
    Algo 2 takes tables = (alpha, beta) [iterations, rows] (default: the
    reference's), algo 3 degree_alpha (19 values; default: the reference's)."""
    d = dec(row)
    soft = np.ascontiguousarray(soft, dtype=np.int16)
    hard = np.zeros(c.B // 8, np.uint8)
    it = ctypes.c_int(0)
    if d.algo in (2, 3):
        lq = np.zeros(c.N + 2 * c.Zc, np.int16)
        if d.algo == 2:
            a, b = _tables_for(c.bgn, tables)
            st = lib().orc_decode_i16_table(ctypes.byref(c), ctypes.byref(d), _p(soft), _p(hard),
                                            _p(lq), ctypes.byref(it), _p(a), _p(b), a.shape[0])
        else:
            da = None if degree_alpha is None else np.ascontiguousarray(degree_alpha, np.int16)
            st = lib().orc_decode_i16_degree(ctypes.byref(c), ctypes.byref(d), _p(soft),
                                             _p(hard), _p(lq), ctypes.byref(it),
                                             None if da is None else _p(da))
        return st, hard, lq, it.value
    if precision is None:
        precision = "i16" if d.algo == 1 else "f32"
    name, dt = _DEC_FN[(precision, d.algo)]
    lq = np.zeros(c.N + 2 * c.Zc, dt)
    st = getattr(lib(), name)(ctypes.byref(c), ctypes.byref(d), _p(soft),
                              _p(hard), _p(lq), ctypes.byref(it))
    return st, hard, lq, it.value


def decode_path(llr, cfg, row, harq=None, tables=None, degree_alpha=None):
    """Whole c_ldpc_5g_decode path.  cfg = (bgn, A, crctype, E, rv, Qm, harq_on).
    tables / degree_alpha as in decode_soft.  Returns dict(status, hard, soft,
    lq, iters)."""
    bgn, A, crctype, E, rv, Qm, harq_on = cfg
    st, p = params(bgn, A, crctype, E, rv, Qm, harq_on)
    n = (p["N"] + 2 * p["Zc"]) if st == 0 else 68 * 384
    soft = np.zeros(68 * 384, np.int16)
    if harq is not None:
        soft[:len(harq)] = harq
    B = p["B"] if st == 0 else A + 24
    hard = np.zeros(max(B // 8, 1), np.uint8)
    lq16 = np.zeros(68 * 384, np.int16)
    lq32 = np.zeros(68 * 384, np.float32)
    it = ctypes.c_int(0)
    d = dec(row)
    llr = np.ascontiguousarray(llr, dtype=np.int16)
    ta = tb = da = None
    titers = 0
    if d.algo == 2 and bgn in (1, 2):
        ta, tb = _tables_for(bgn, tables)
        titers = ta.shape[0]
    if d.algo == 3 and degree_alpha is not None:
        da = np.ascontiguousarray(degree_alpha, np.int16)
    status = lib().orc_decode_path_tables(
        _p(llr), _p(hard), bgn, A, crctype, E, rv, Qm, harq_on, ctypes.byref(d), _p(soft),
        _p(lq16), _p(lq32), ctypes.byref(it), None if ta is None else _p(ta),
        None if tb is None else _p(tb), titers, None if da is None else _p(da))
    lq = lq16[:n] if d.algo in (1, 2, 3) else lq32[:n]
    return dict(status=status, hard=hard, soft=soft[:n], lq=lq, iters=it.value)


def check_minsum_i16(rel, alpha, beta, max_llr):
    rel = np.ascontiguousarray(rel, dtype=np.int16)
    msg = np.zeros_like(rel)
    lib().orc_check_minsum_i16(len(rel), _p(rel), alpha, beta, max_llr, _p(msg))
    return msg


def check_bp_f64(rel):
    rel = np.ascontiguousarray(rel, dtype=np.float64)
    msg = np.zeros_like(rel)
    lib().orc_check_bp_f64(len(rel), _p(rel), _p(msg))
    return msg


def encode(c, payload_bits):
    """payload_bits: B bits (payload + CRC).  Returns (N+2Zc) codeword bits."""
    ncol = 68 if c.bgn == 1 else 52
    cw = np.zeros(ncol * c.Zc, np.uint8)
    b = np.ascontiguousarray(payload_bits, dtype=np.uint8)
    if lib().orc_encode(ctypes.byref(c), _p(b), _p(cw)) != 0:
        raise RuntimeError("oracle encoder failed")
    return cw


def rate_match(c, cw):
    out = np.zeros(c.E, np.uint8)
    cw = np.ascontiguousarray(cw, dtype=np.uint8)
    lib().orc_rate_match(ctypes.byref(c), _p(cw), _p(out))
    return out


def syndrome_ok(c, cw):
    cw = np.ascontiguousarray(cw, dtype=np.uint8)
    return bool(lib().orc_syndrome_ok(ctypes.byref(c), _p(cw)))


def attach_crc(payload_bits, crctype):
    """payload bits (A, multiple of 8) -> B bits with the CRC appended."""
    width = 16 if crctype == 0 else 24
    r = crc(np.packbits(payload_bits), crctype)
    tail = np.array([(r >> (width - 1 - i)) & 1 for i in range(width)], np.uint8)
    return np.concatenate([payload_bits.astype(np.uint8), tail])


# --------------------------------------------------------------------------
# Test channel: Gray QAM (TS 38.211 5.1) + complex AWGN + max-log LLRs,
# quantized x128, rounded half away from zero, clipped to +-10000
# (c_lib/tx_rx_process.c:27-39, 68-99, 234-268, 286-324), with numpy's RNG.
# --------------------------------------------------------------------------
def _constellation(Qm):
    """All 2^Qm points of the TS 38.211 Gray mapping, unit average energy."""
    pts = np.zeros(1 << Qm, np.complex128)
    for v in range(1 << Qm):
        b = [(v >> (Qm - 1 - i)) & 1 for i in range(Qm)]
        s = [1 - 2 * x for x in b]
        if Qm == 1:
            pts[v] = (s[0] + 1j * s[0]) / np.sqrt(2)
        elif Qm == 2:
            pts[v] = (s[0] + 1j * s[1]) / np.sqrt(2)
        elif Qm == 4:
            pts[v] = (s[0] * (2 - s[2]) + 1j * s[1] * (2 - s[3])) / np.sqrt(10)
        elif Qm == 6:
            pts[v] = (s[0] * (4 - s[2] * (2 - s[4])) +
                      1j * s[1] * (4 - s[3] * (2 - s[5]))) / np.sqrt(42)
        elif Qm == 8:
            pts[v] = (s[0] * (8 - s[2] * (4 - s[4] * (2 - s[6]))) +
                      1j * s[1] * (8 - s[3] * (4 - s[5] * (2 - s[7])))) / np.sqrt(170)
    return pts


def tx_rx(bits, Qm, noise_power_db, seed):
    """The reference's test channel with its own noise stream after
    srand(seed) (orc_tx_rx, c_lib/tx_rx_process.c:286-324).  Returns
    (noise [E/Qm, 2] as added, noise variance, float LLRs, int16 LLRs)."""
    bits = np.ascontiguousarray(bits, np.uint8)
    E = len(bits)
    noise = np.zeros((E // Qm, 2), np.float32)
    nv = ctypes.c_float()
    llr = np.zeros(E, np.float32)
    q = np.zeros(E, np.int16)
    if lib().orc_tx_rx(_p(bits), E, Qm, ctypes.c_float(noise_power_db), ctypes.c_uint(seed),
                       _p(noise), ctypes.byref(nv), _p(llr), _p(q)) != 0:
        raise ValueError("orc_tx_rx: invalid arguments")
    return noise, float(nv.value), llr, q


def channel_llr(bits, Qm, noise_power_db, rng):
    """bits (E,) -> int16 LLRs (E,).  Max-log LLR by exhaustive distance
    (equal to the reference's piecewise formulas for Gray QAM)."""
    bits = np.asarray(bits, np.uint8)
    E = len(bits)
    assert E % Qm == 0
    pts = _constellation(Qm)
    sym_idx = np.zeros(E // Qm, np.int64)
    for i in range(Qm):
        sym_idx = (sym_idx << 1) | bits[i::Qm]
    tx = pts[sym_idx]
    nv = 10.0 ** (noise_power_db / 10.0)
    sigma = np.sqrt(nv / 2.0)
    rx = tx + sigma * (rng.standard_normal(len(tx)) + 1j * rng.standard_normal(len(tx)))
    d2 = np.abs(rx[:, None] - pts[None, :]) ** 2
    llr = np.zeros((len(tx), Qm))
    for i in range(Qm):
        bit_i = (np.arange(1 << Qm) >> (Qm - 1 - i)) & 1
        d0 = d2[:, bit_i == 0].min(axis=1)
        d1 = d2[:, bit_i == 1].min(axis=1)
        llr[:, i] = (d1 - d0) / nv
    if Qm == 1:
        llr = llr * 1.0  # BPSK on both axes: same max-log form
    scaled = llr.reshape(-1) * 128.0
    q = np.where(scaled >= 0, np.floor(scaled + 0.5), np.ceil(scaled - 0.5))
    return np.clip(q, -10000, 10000).astype(np.int16)


def make_codeblock(bgn, A, crctype, E, rv, Qm, noise_power_db, rng):
    """Random payload -> CRC -> encode -> rate match -> channel LLRs."""
    c = code(bgn, A, crctype, E, rv, Qm)
    payload = rng.integers(0, 2, A).astype(np.uint8)
    b = attach_crc(payload, crctype)
    cw = encode(c, b)
    rm = rate_match(c, cw)
    llr = channel_llr(rm, Qm, noise_power_db, rng)
    return dict(code=c, payload=payload, bits=b, cw=cw, rm=rm, llr=llr)


# --------------------------------------------------------------------------
# Unrolled float min-sum of the trainable decoder
# (deep_learning/deeplearning_ldpc_model.py:315-410) and its gradient.
# --------------------------------------------------------------------------
_UNROLLED_DT = {"f32": np.float32, "f64": np.float64}


def unrolled_forward(c, iters, llr, alpha, beta, clip=None, precision="f32",
                     want_iters=False, want_over=False):
    """One codeblock.  llr (N+2Zc,), alpha/beta (iters, rows).
    Returns (lq, iter_lq or None) [+ over counts when want_over]."""
    dt = _UNROLLED_DT[precision]
    n = c.N + 2 * c.Zc
    llr = np.ascontiguousarray(llr, dtype=dt)
    a = np.ascontiguousarray(alpha, dtype=dt).reshape(iters, c.rows)
    b = np.ascontiguousarray(beta, dtype=dt).reshape(iters, c.rows)
    lq = np.zeros(n, dt)
    it = np.zeros((iters, n), dt) if want_iters else None
    over = np.zeros(2, np.int32)
    getattr(lib(), "orc_unrolled_forward_" + precision)(
        ctypes.byref(c), iters, _p(llr), _p(a), _p(b), -1.0 if clip is None else clip,
        _p(lq), _p(it) if it is not None else None, _p(over))
    return (lq, it, over) if want_over else (lq, it)


def unrolled_backward(c, iters, llr, alpha, beta, g_lq, g_iter=None, clip=None,
                      detach=False, precision="f32"):
    """One codeblock.  Returns (g_llr, g_alpha, g_beta); the alpha/beta
    gradients are float64 (iters, rows)."""
    dt = _UNROLLED_DT[precision]
    n = c.N + 2 * c.Zc
    llr = np.ascontiguousarray(llr, dtype=dt)
    a = np.ascontiguousarray(alpha, dtype=dt).reshape(iters, c.rows)
    b = np.ascontiguousarray(beta, dtype=dt).reshape(iters, c.rows)
    g_lq = np.ascontiguousarray(g_lq, dtype=dt)
    gi = None if g_iter is None else np.ascontiguousarray(g_iter, dtype=dt).reshape(iters, n)
    g_llr = np.zeros(n, dt)
    ga = np.zeros((iters, c.rows), np.float64)
    gb = np.zeros((iters, c.rows), np.float64)
    getattr(lib(), "orc_unrolled_backward_" + precision)(
        ctypes.byref(c), iters, _p(llr), _p(a), _p(b), -1.0 if clip is None else clip,
        1 if detach else 0, _p(g_lq), _p(gi) if gi is not None else None, _p(g_llr),
        _p(ga), _p(gb))
    return g_llr, ga, gb


def row_degrees(c):
    ro = np.ctypeslib.as_array(ctypes.cast(c.row_off, ctypes.POINTER(ctypes.c_uint16)),
                               shape=(c.rows + 1,))
    return np.diff(ro.astype(np.int64))
