"""ctypes front end of the CPU build of the decode kernels (test-only)."""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "libldpc_emu.so")
_lib = None


def lib():
    global _lib
    if _lib is None:
        subprocess.run(["make", "-s", "-C", HERE], check=True)
        L = ctypes.CDLL(LIB)
        P = ctypes.c_void_p
        L.emu_decode_batch.argtypes = [ctypes.c_int] + [P] * 13 + [ctypes.c_uint64]
        L.emu_decode_batch.restype = ctypes.c_int
        L.emu_decode_tiles.argtypes = [ctypes.c_int] + [P] * 12 + [ctypes.c_int]
        L.emu_decode_tiles.restype = ctypes.c_int
        L.emu_plan_tiles.argtypes = [ctypes.c_int, P, ctypes.c_int, P, ctypes.c_int, P, P]
        L.emu_plan_tiles.restype = ctypes.c_int
        I, F = ctypes.c_int, ctypes.c_float
        L.emu_unrolled_forward.argtypes = [I, I, I, I, I, F, I, I, I] + [P] * 6
        L.emu_unrolled_forward.restype = I
        L.emu_unrolled_record_bytes.argtypes = [I, I, I, I, I, F, I, I]
        L.emu_unrolled_record_bytes.restype = ctypes.c_long
        L.emu_unrolled_backward.argtypes = [I, I, I, I, I, F, I, I, I] + [P] * 9
        L.emu_unrolled_backward.restype = I
        L.emu_synth.argtypes = [I] * 7 + [P, ctypes.c_uint64] + [P] * 7
        L.emu_synth.restype = I
        L.emu_synth_noise.argtypes = [I] * 7 + [P, P, P, ctypes.c_uint64] + [P] * 7
        L.emu_synth_noise.restype = I
        L.emu_set_trained_tables.argtypes = [I, P, P, I]
        L.emu_set_trained_tables.restype = I
        L.emu_set_degree_alpha.argtypes = [P]
        L.emu_set_degree_alpha.restype = None
        L.emu_row_general.argtypes = [P]
        L.emu_row_general.restype = I
        L.emu_force_general.argtypes = [I]
        L.emu_force_general.restype = None
        L.emu_use_fixed.argtypes = [I]
        L.emu_crc_tables.argtypes = [P, I]
        L.emu_crc_tables.restype = I
        L.emu_use_fixed.restype = None
        L.emu_set_persistent_tiles.argtypes = [I]
        L.emu_set_persistent_tiles.restype = None
        L.emu_set_edge_lanes.argtypes = [I]
        L.emu_set_edge_lanes.restype = None
        _lib = L
    return _lib


def _p(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def dims(cfg):
    from oracle import oracle as O
    st, p = O.params(*cfg)
    return p


def decode_batch(cfgs, row, llrs, soft_pool=None, soft_slots=None, want_lq=True,
                 tile_k=None):
    """Run the kernel source on the CPU (packed tiles of tile_k codeblocks
    when tile_k is not None, 0 = as many as fit).  Returns per-codeblock
    dicts."""
    n = len(cfgs)
    cf = np.array(cfgs, np.int32).reshape(n, 7)
    ps = [dims(tuple(c)) for c in cfgs]
    off = np.cumsum([0] + [len(x) for x in llrs[:-1]]).astype(np.uint32)
    llr = np.ascontiguousarray(np.concatenate(llrs), np.int16)
    hard = np.zeros(sum(p["B"] // 8 for p in ps), np.uint8)
    st = np.zeros(n, np.int32)
    it = np.zeros(n, np.int32)
    nlq = sum(p["N"] + 2 * p["Zc"] for p in ps)
    lq = np.zeros(nlq, np.int16 if row[0] in (1, 2, 3) else np.float32) if want_lq else None
    hoff = np.zeros(n, np.uint32)
    loff = np.zeros(n, np.uint32)
    nwork = sum((316 if c[0] == 1 else 197) * p["Zc"] for c, p in zip(cfgs, ps))
    work = np.zeros(max(nwork, 1), np.float32)
    slots = None if soft_slots is None else np.ascontiguousarray(soft_slots, np.int32)
    if tile_k is None:
        rc = lib().emu_decode_batch(n, _p(cf), _p(off), _p(np.array(row, np.int32)), _p(llr),
                                    _p(soft_pool), _p(slots), _p(hard), _p(st), _p(it), _p(lq),
                                    _p(hoff), _p(loff), _p(work), ctypes.c_uint64(len(work)))
        if rc != 0:
            raise RuntimeError(f"emu_decode_batch returned {rc}")
    else:
        rc = lib().emu_decode_tiles(n, _p(cf), _p(off), _p(np.array(row, np.int32)), _p(llr),
                                    _p(soft_pool), _p(slots), _p(hard), _p(st), _p(it), _p(lq),
                                    _p(hoff), _p(loff), int(tile_k))
        if rc <= 0:
            raise RuntimeError(f"emu_decode_tiles returned {rc}")
    out = []
    for i, p in enumerate(ps):
        h = hard[hoff[i]:hoff[i] + p["B"] // 8]
        q = lq[loff[i]:loff[i] + p["N"] + 2 * p["Zc"]] if want_lq else None
        out.append(dict(status=int(st[i]), iters=int(it[i]), hard=h, lq=q))
    return out


def unrolled(cfg, T, llr, alpha, beta, clip=None, detach=False, cbs=0, g_lq=None, g_iter=None,
             want_iters=False):
    """Unrolled trainable min-sum kernels on the CPU.  cfg = (bgn, A, crctype, E);
    llr (batch, n) float32; alpha/beta (T, rows).  Returns dict(lq, iter_lq,
    and with g_lq: g_llr, g_alpha, g_beta)."""
    bgn, A, crctype, E = cfg[:4]
    llr = np.ascontiguousarray(llr, np.float32)
    batch, n = llr.shape
    a = np.ascontiguousarray(alpha, np.float32)
    b = np.ascontiguousarray(beta, np.float32)
    c = -1.0 if clip is None else float(clip)
    L = lib()
    nb = L.emu_unrolled_record_bytes(bgn, A, crctype, E, T, c, int(detach), batch)
    assert nb >= 0
    rec = np.zeros(max(nb // 4, 1), np.uint32)
    lq = np.zeros((batch, n), np.float32)
    it = np.zeros((T, batch, n), np.float32) if (want_iters or g_iter is not None) else None
    rc = L.emu_unrolled_forward(bgn, A, crctype, E, T, c, int(detach), cbs, batch, _p(llr),
                                _p(a), _p(b), _p(lq), _p(it), _p(rec))
    assert rc == 0, rc
    out = dict(lq=lq, iter_lq=it)
    if g_lq is not None:
        g = np.ascontiguousarray(g_lq, np.float32)
        gi = None if g_iter is None else np.ascontiguousarray(g_iter, np.float32)
        g_llr = np.zeros((batch, n), np.float32)
        ga = np.zeros(a.shape, np.float32)
        gb = np.zeros(b.shape, np.float32)
        rc = L.emu_unrolled_backward(bgn, A, crctype, E, T, c, int(detach), cbs, batch, _p(llr),
                                     _p(a), _p(b), _p(rec), _p(g), _p(gi), _p(g_llr), _p(ga),
                                     _p(gb))
        assert rc == 0, rc
        out.update(g_llr=g_llr, g_alpha=ga, g_beta=gb)
    return out


def synth(cfg, n, noise_db, seed=0, payload=None, beliefs=False, noise=None):
    """Input-synthesis kernel on the CPU.  cfg = (bgn, A, crctype, E, rv, Qm).
    noise: optional (noise_var[n], samples[n, E/Qm, 2]) given channel noise.
    Returns dict(bits, cw, rm, llr, llr_f32[, beliefs])."""
    from oracle import oracle as O
    bgn, A, crctype, E, rv, Qm = cfg[:6]
    st, p = O.params(bgn, A, crctype, E, rv, Qm, 0)
    assert st == 0
    B, nt = p["B"], p["N"] + 2 * p["Zc"]
    nz = np.ascontiguousarray(np.broadcast_to(np.float32(noise_db), (n,)), np.float32)
    pay = None if payload is None else np.ascontiguousarray(payload, np.uint8)
    out = dict(bits=np.zeros((n, B), np.uint8), cw=np.zeros((n, nt), np.uint8),
               rm=np.zeros((n, E), np.uint8), llr=np.zeros((n, E), np.int16),
               llr_f32=np.zeros((n, E), np.float32))
    bel = np.zeros((n, nt), np.float32) if beliefs else None
    nv = smp = None
    if noise is not None:
        nv = np.ascontiguousarray(noise[0], np.float32).reshape(n)
        smp = np.ascontiguousarray(noise[1], np.float32).reshape(-1)
        assert smp.size == n * (E // Qm) * 2
    rc = lib().emu_synth_noise(bgn, A, crctype, E, rv, Qm, n, _p(nz), _p(nv), _p(smp),
                               ctypes.c_uint64(seed), _p(pay), _p(out["bits"]), _p(out["cw"]),
                               _p(out["rm"]), _p(out["llr"]), _p(out["llr_f32"]), _p(bel))
    assert rc == 0, rc
    if beliefs:
        out["beliefs"] = bel
    return out


def set_trained_tables(bgn, alpha=None, beta=None):
    """Algo-2 tables [iterations, rows] of the CPU build (None: defaults)."""
    if alpha is None:
        lib().emu_set_trained_tables(bgn, None, None, 0)
        return
    a = np.ascontiguousarray(alpha, np.int16)
    b = np.ascontiguousarray(beta, np.int16)
    if lib().emu_set_trained_tables(bgn, _p(a), _p(b), a.shape[0]) != 0:
        raise ValueError("invalid tables")


def set_degree_alpha(values=None):
    lib().emu_set_degree_alpha(None if values is None else
                               _p(np.ascontiguousarray(values, np.int16)))


def row_general(row):
    return bool(lib().emu_row_general(_p(np.array(row, np.int32))))


def force_general(g):
    """Force the int16 row form: True exact, False fast, None the library's."""
    lib().emu_force_general(-1 if g is None else int(bool(g)))


def persistent_tiles(grid):
    """Tile launches pull their tiles from a device queue with at most
    `grid` workgroups (0: one workgroup per tile)."""
    lib().emu_set_persistent_tiles(int(grid))


def edge_lanes(on):
    """Algo-1 fast-row batches (one codeblock per workgroup) in the
    lane-per-edge schedule."""
    lib().emu_set_edge_lanes(int(bool(on)))


def use_fixed(on):
    """Whether uniform batches run the lifting-size specialized kernel."""
    lib().emu_use_fixed(int(bool(on)))


TILE_DESC = np.dtype([("first", "<u4"), ("etab_off", "<u4"), ("lds_base", "<u4"),
                      ("count", "<u2"), ("bgn", "<u2"), ("Zc", "<u2"), ("lds_stride", "<u2"),
                      ("lane_base", "<u2"), ("slot_base", "<u2")])
TILE_WAVES = 8


def plan_tiles(cfgs, k=0):
    """The tile plan of a batch of configs: (descriptors [workgroups, 8] as
    TILE_DESC, order, largest LQ footprint in entries)."""
    cf = np.ascontiguousarray(np.array(cfgs, np.int32).reshape(-1, 7))
    n = len(cf)
    cap = TILE_WAVES * max(n, 1)
    tiles = np.zeros(cap, TILE_DESC)
    order = np.zeros(max(n, 1), np.uint32)
    lds = ctypes.c_int(0)
    rc = lib().emu_plan_tiles(n, _p(cf), k, _p(tiles), cap, _p(order), ctypes.byref(lds))
    if rc < 0:
        raise RuntimeError(f"emu_plan_tiles returned {rc}")
    return tiles[:rc * TILE_WAVES].reshape(rc, TILE_WAVES), order[:n], lds.value


def crc_tables():
    """The device CRC tables (4 CRC types x words per type) and the offset of
    the per-bit weights x^m mod g within a type's words."""
    L = lib()
    words = L.emu_crc_tab_words()
    out = np.zeros(4 * words, np.uint32)
    assert L.emu_crc_tables(_p(out), out.size) == out.size
    return out.reshape(4, words), L.emu_crc_bit_off()
