"""Batched decode on the MI355X through the C ABI plan entry points.

PyTorch is only the device-memory and stream plumbing here: buffers are torch
tensors on ``cuda`` (HIP), their addresses go to
``ldpc_mi355_plan_decode``, and the work is the library's HIP kernel.
"""
import ctypes

import numpy as np

from . import _lib
from ._lib import CbDescriptor, DecoderConfig, LdpcInConfig, ptr

ALGO_LQ_DTYPE = {1: "int16", 2: "int16", 3: "int16", 0: "float32", 4: "float32"}


def soft_slot_elems():
    return int(_lib.load().ldpc_mi355_soft_slot_elems())


def _decoder(row):
    return DecoderConfig(*[int(v) for v in row])


_tuning_applied = False
# tile schedule of plans that do not choose one: persistent (a resident grid
# pulling tiles from a device queue) when the autotuner measured it faster
_persistent_default = False
# (bgn, Zc) classes whose autotuned schedule is lane-per-edge: a plan whose
# codeblocks all belong to such classes uses it (algo 1)
_edge_lane_classes = set()
_zc_cache = {}


def _bgn_zc(cfg):
    key = tuple(int(v) for v in cfg[:6])
    if key not in _zc_cache:
        pc = _lib.LdpcConfig()
        st = _lib.load().gen_ldpc_parameters(ctypes.byref(pc), ctypes.byref(LdpcInConfig(
            *(key + (0,)))))
        _zc_cache[key] = (key[0], int(pc.Zc)) if st == 0 else None
    return _zc_cache[key]


def apply_tuning(path=None):
    """Load the autotuned codeblocks-per-workgroup table (tile_table.json)
    into the library and its tile schedule as the plans' default; plans
    built afterwards use them.  Missing table: the library's default (as many
    codeblocks as fit in 384 lanes, one workgroup per tile)."""
    global _tuning_applied, _persistent_default
    import json
    import os
    if path is None:
        path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tile_table.json")
    _tuning_applied = True
    if not os.path.exists(path):
        return 0
    with open(path) as f:
        table = json.load(f)
    _persistent_default = bool(table.get("schedule", {}).get("persistent", False))
    _edge_lane_classes.clear()
    for key, v in table.get("classes", {}).items():
        if v.get("schedule") == "edge_lanes":
            _edge_lane_classes.add(tuple(int(x) for x in key.split(":")))
    L = _lib.load()
    n = 0
    for key, v in table.get("classes", {}).items():
        bgn, zc = (int(x) for x in key.split(":"))
        if L.ldpc_mi355_set_tile_codeblocks(bgn, zc, int(v["k"])) == 0:
            n += 1
    return n


class BatchPlan:
    """A fixed batch of codeblocks and one decoder row.

    cfgs: sequence of (bgn, A, crctype, E, rv, Qm, HARQ_on).
    llr_offsets: element offset of each codeblock's E LLRs in the LLR buffer
    (default: packed back to back).  soft_slots: HARQ soft-buffer slot per
    codeblock or -1 (default: none).  want_lq: also return posteriors.
    schedule: "oneshot" (one workgroup per packed tile), "persistent" (a
    resident grid pulling the tiles from a device queue) or None (the
    autotuned default from tile_table.json).
    """

    PACK = 2        # plan flag: pack codeblocks of one (BG, Zc) per workgroup
    NO_PACK = 4     # plan flag: one codeblock per workgroup
    GENERIC = 8     # plan flag: no kernel specialized for a single (BG, Zc)
    PERSISTENT = 16  # plan flag: packed tiles pulled from a device queue by a resident grid
    EDGE_LANES = 32  # plan flag: lane-per-edge schedule (small lifting sizes, algo 1 fast rows)

    def __init__(self, cfgs, decoder_row, llr_offsets=None, soft_slots=None, want_lq=False,
                 flags=0, schedule=None):
        L = _lib.load()
        if not _tuning_applied:
            apply_tuning()
        if schedule is None:
            schedule = "persistent" if _persistent_default else "oneshot"
        if schedule not in ("oneshot", "persistent"):
            raise ValueError(f"unknown tile schedule {schedule!r}")
        if schedule == "persistent":
            flags |= self.PERSISTENT
        if (_edge_lane_classes and int(decoder_row[0]) == 1 and not flags & self.PACK and
                all(_bgn_zc(c) in _edge_lane_classes for c in cfgs)):
            flags |= self.EDGE_LANES
        n = len(cfgs)
        if n == 0:
            raise ValueError("empty batch")
        self.n = n
        self.cfgs = [tuple(int(v) for v in c) for c in cfgs]
        self.row = [int(v) for v in decoder_row]
        arr = (CbDescriptor * n)()
        off = 0
        self.llr_offsets = []
        for i, c in enumerate(self.cfgs):
            arr[i].cfg = LdpcInConfig(*c)
            o = off if llr_offsets is None else int(llr_offsets[i])
            arr[i].llr_offset = o
            self.llr_offsets.append(o)
            off += c[3]
            arr[i].soft_slot = -1 if soft_slots is None else int(soft_slots[i])
        self.llr_elems = max(o + c[3] for o, c in zip(self.llr_offsets, self.cfgs))
        self.max_slot = -1 if soft_slots is None else max(int(v) for v in soft_slots)
        self._handle = ctypes.c_void_p()
        bad = ctypes.c_int32(-1)
        dec = _decoder(self.row)
        st = L.ldpc_mi355_plan_create(ctypes.byref(self._handle), arr, n, ctypes.byref(dec),
                                      (1 if want_lq else 0) | int(flags), ctypes.byref(bad))
        if st != 0:
            raise ValueError(f"plan_create failed ({st}), codeblock {bad.value}")
        self.tiles = int(L.ldpc_mi355_plan_tiles(self._handle))
        hb, le, es = ctypes.c_uint64(), ctypes.c_uint64(), ctypes.c_int32()
        hoff = np.zeros(n, np.uint32)
        loff = np.zeros(n, np.uint32)
        L.ldpc_mi355_plan_layout(self._handle, ctypes.byref(hb), ctypes.byref(le),
                                 ctypes.byref(es), ptr(hoff), ptr(loff))
        self.hard_bytes = int(hb.value)
        self.lq_elems = int(le.value)
        self.lq_elem_size = int(es.value)
        self.hard_offsets = hoff.astype(np.int64)
        self.lq_offsets = loff.astype(np.int64)
        self.want_lq = want_lq

    def close(self):
        if getattr(self, "_handle", None) and self._handle.value:
            _lib.load().ldpc_mi355_plan_destroy(self._handle)
            self._handle = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def alloc_outputs(self, device="cuda"):
        import torch
        out = dict(
            hard=torch.zeros(max(self.hard_bytes, 1), dtype=torch.uint8, device=device),
            status=torch.zeros(self.n, dtype=torch.int32, device=device),
            iters=torch.zeros(self.n, dtype=torch.int32, device=device))
        if self.want_lq:
            dt = torch.int16 if self.lq_elem_size == 2 else torch.float32
            out["lq"] = torch.zeros(self.lq_elems, dtype=dt, device=device)
        return out

    def launch(self, llr, out, soft_pool=None, stream=None):
        """Asynchronous decode of device tensor ``llr`` (int16) into ``out`` on
        ``stream`` (a raw hipStream_t handle; default: torch's current stream,
        whose handle is 0 for the default stream)."""
        import torch
        if llr.dtype.itemsize != 2 or llr.numel() < self.llr_elems:
            raise ValueError("llr must be int16 with every codeblock's E values")
        if not llr.is_cuda:
            raise ValueError("llr must be a device tensor")
        if self.max_slot >= 0:
            need = (self.max_slot + 1) * soft_slot_elems()
            if soft_pool is None or not soft_pool.is_cuda or soft_pool.dtype.itemsize != 2:
                raise ValueError("codeblocks use soft-buffer slots: soft_pool must be an int16 "
                                 "device tensor")
            if soft_pool.numel() < need:
                raise ValueError(f"soft_pool holds {soft_pool.numel()} elements; slot "
                                 f"{self.max_slot} needs {need}")
        if stream is None:
            stream = torch.cuda.current_stream(llr.device).cuda_stream
        st = _lib.load().ldpc_mi355_plan_decode(
            self._handle, ptr(llr), ptr(soft_pool) if soft_pool is not None else None,
            ptr(out["hard"]), ptr(out["status"]), ptr(out["iters"]),
            ptr(out["lq"]) if self.want_lq else None, ctypes.c_void_p(stream))
        if st != 0:
            raise RuntimeError(f"plan_decode failed ({st})")
        return out

    def decode(self, llr, soft_pool=None):
        """Synchronous convenience: torch device tensors in and out."""
        import torch
        out = self.alloc_outputs(llr.device)
        stream = torch.cuda.current_stream(llr.device).cuda_stream
        self.launch(llr, out, soft_pool, stream)
        torch.cuda.current_stream(llr.device).synchronize()
        return out

    def split_hard(self, hard):
        """Per-codeblock B/8-byte views of a host copy of ``hard``."""
        res = []
        for i, c in enumerate(self.cfgs):
            B = c[1] + (16 if c[2] == 0 else 24)
            o = int(self.hard_offsets[i])
            res.append(hard[o:o + B // 8])
        return res

    def split_lq(self, lq):
        res = []
        for i, c in enumerate(self.cfgs):
            zc, N = code_dims(c)
            o = int(self.lq_offsets[i])
            res.append(lq[o:o + N + 2 * zc])
        return res


def code_dims(cfg):
    """(Zc, N) of a config via the library's gen_ldpc_parameters."""
    L = _lib.load()
    pc = _lib.LdpcConfig()
    st = L.gen_ldpc_parameters(ctypes.byref(pc), ctypes.byref(LdpcInConfig(*cfg)))
    if st != 0:
        raise ValueError(f"invalid config {cfg}")
    return int(pc.Zc), int(pc.N)


def synth_batch(cfg, n, noise_power_db, seed=0, threads=8, want_payload=True):
    """Host synthesis of n codeblocks (numpy int16 LLRs [n, E], payload bits)."""
    L = _lib.load()
    ic = LdpcInConfig(*[int(v) for v in cfg])
    llr = np.zeros((n, ic.E), np.int16)
    pay = np.zeros((n, ic.A), np.uint8) if want_payload else None
    st = L.ldpc_mi355_synth_batch(ctypes.byref(ic), n, ctypes.c_float(noise_power_db),
                                  ctypes.c_uint64(seed), threads, ptr(llr),
                                  ptr(pay) if pay is not None else None)
    if st != 0:
        raise ValueError(f"invalid config {cfg}")
    return llr, pay


def synth_device(cfg, n, noise_power_db, seed=0, payload=None, outputs=("llr",), stream=None,
                 noise=None):
    """Device synthesis of n codeblocks of config cfg (ldpc_mi355_synth_device).
    noise_power_db: scalar or n values.  outputs: any of "bits", "cw", "rm",
    "llr" (int16 E per codeblock), "llr_f32", "beliefs".  noise: optional
    (noise_var[n], samples[n, E/Qm, 2]) given channel noise
    (ldpc_mi355_synth_device_noise; noise_power_db is then ignored).
    Returns a dict of torch device tensors shaped [n, ...]."""
    import torch
    L = _lib.load()
    ic = LdpcInConfig(*[int(v) for v in cfg])
    pc = _lib.LdpcConfig()
    if L.gen_ldpc_parameters(ctypes.byref(pc), ctypes.byref(ic)) != 0:
        raise ValueError(f"invalid config {cfg}")
    dev = torch.device("cuda")
    nz = torch.as_tensor(np.broadcast_to(np.asarray(noise_power_db, np.float32), (n,)).copy(),
                         device=dev)
    nt = pc.N + 2 * pc.Zc
    shapes = dict(bits=(pc.B, torch.uint8), cw=(nt, torch.uint8), rm=(ic.E, torch.uint8),
                  llr=(ic.E, torch.int16), llr_f32=(ic.E, torch.float32),
                  beliefs=(nt, torch.float32))
    want = set(outputs)
    if "beliefs" in want:
        want.add("llr_f32")
    out = {k: torch.empty((n, shapes[k][0]), dtype=shapes[k][1], device=dev) for k in want}
    pay = None
    if payload is not None:
        pay = torch.as_tensor(np.ascontiguousarray(payload, np.uint8), device=dev)
    if stream is None:
        stream = torch.cuda.current_stream().cuda_stream
    g = lambda k: ptr(out[k]) if k in out else None  # noqa: E731
    if noise is None:
        st = L.ldpc_mi355_synth_device(ctypes.byref(ic), n, ptr(nz), ctypes.c_uint64(seed),
                                       ptr(pay) if pay is not None else None, g("bits"), g("cw"),
                                       g("rm"), g("llr"), g("llr_f32"), g("beliefs"),
                                       ctypes.c_void_p(stream))
        keep = (nz, pay)
    else:
        nv = torch.as_tensor(np.ascontiguousarray(noise[0], np.float32).reshape(n), device=dev)
        smp = np.ascontiguousarray(noise[1], np.float32)
        if smp.size != n * (ic.E // ic.Qm) * 2:
            raise ValueError("noise samples must be n * (E / Qm) * 2 floats")
        smp = torch.as_tensor(smp.reshape(-1), device=dev)
        st = L.ldpc_mi355_synth_device_noise(ctypes.byref(ic), n, ptr(nv), ptr(smp),
                                             ctypes.c_uint64(seed),
                                             ptr(pay) if pay is not None else None, g("bits"),
                                             g("cw"), g("rm"), g("llr"), g("llr_f32"),
                                             g("beliefs"), ctypes.c_void_p(stream))
        keep = (nv, smp, pay)
    if st != 0:
        raise RuntimeError(f"ldpc_mi355_synth_device failed ({st})")
    out["_keep"] = keep    # inputs stay alive until the caller drops the dict
    return out
