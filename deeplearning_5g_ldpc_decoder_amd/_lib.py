"""ctypes binding of libnrldpc_mi355.so (declared in include/nr_ldpc_mi355.h).

The library is built in-tree (``make -C deeplearning_5g_ldpc_decoder_amd/csrc``
or ``__graft_entry__.build()``).  Loading fails loudly when it is missing:
there is no CPU fallback for the decode path.
"""
import ctypes
import os

import numpy as np

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG_DIR, "libnrldpc_mi355.so")
CSRC_DIR = os.path.join(PKG_DIR, "csrc")

# Every exported symbol of include/nr_ldpc_mi355.h (functions and data).
EXPORTED_FUNCTIONS = (
    "calculate_crc", "gen_ldpc_parameters", "dequantize_int16_to_llr",
    "c_ldpc_5g_decode", "c_ldpc_5g_encode", "tx_rx_process",
    "cuda_ldpc_5g_decode", "cuda_ldpc_5g_encode",
    "ldpc_mi355_decode_ex", "ldpc_mi355_channel_seed", "ldpc_mi355_plan_create",
    "ldpc_mi355_plan_layout", "ldpc_mi355_plan_decode", "ldpc_mi355_plan_destroy",
    "ldpc_mi355_sync", "ldpc_mi355_soft_slot_elems", "ldpc_mi355_version",
    "ldpc_mi355_algo_available", "ldpc_mi355_plan_tiles", "ldpc_mi355_set_tile_codeblocks",
    "ldpc_mi355_synth_batch", "ldpc_mi355_synth_float_batch", "ldpc_mi355_synth_device",
    "ldpc_mi355_synth_device_noise", "ldpc_mi355_unrolled_create", "ldpc_mi355_unrolled_destroy", "ldpc_mi355_unrolled_dims",
    "ldpc_mi355_unrolled_set_codeblocks_per_workgroup", "ldpc_mi355_unrolled_record_bytes",
    "ldpc_mi355_unrolled_scratch_bytes", "ldpc_mi355_unrolled_forward",
    "ldpc_mi355_unrolled_backward",
    "ldpc_mi355_batch_create", "ldpc_mi355_batch_decode", "ldpc_mi355_batch_layout",
    "ldpc_mi355_batch_query", "ldpc_mi355_batch_wait", "ldpc_mi355_batch_destroy",
    "ldpc_mi355_copy_spans",
    "ldpc_mi355_set_trained_tables", "ldpc_mi355_get_trained_tables",
    "ldpc_mi355_set_degree_alpha", "ldpc_mi355_get_degree_alpha",
    "ldpc_mi355_bp_decode_noalloc", "ldpc_mi355_serve_run",
)
EXPORTED_DATA = (
    "BG1_HBG_offset_array", "BG1_HBG_array", "BG1_HBG_Vij_array",
    "BG2_HBG_offset_array", "BG2_HBG_array", "BG2_HBG_Vij_array", "HARQ_memory",
)


class LdpcInConfig(ctypes.Structure):
    """ldpc_in_config_t (c_lib/nr_ldpc.h:42-50)."""
    _fields_ = [(n, ctypes.c_uint16) for n in
                ("bgn", "A", "crctype", "E", "rv", "Qm", "HARQ_on")]


class DecoderConfig(ctypes.Structure):
    """ldpc_decoder_cfg_t (c_lib/nr_ldpc.h:82-90)."""
    _fields_ = [("algo", ctypes.c_uint16), ("max_iterations", ctypes.c_uint16),
                ("alpha", ctypes.c_uint16), ("beta", ctypes.c_uint16),
                ("filling_bit_value", ctypes.c_int16), ("max_llr", ctypes.c_int16),
                ("min_llr", ctypes.c_int16)]


class LdpcConfig(ctypes.Structure):
    """ldpc_config_t (c_lib/nr_ldpc.h:54-79)."""
    _fields_ = [(n, ctypes.c_uint16) for n in
                ("bgn", "A", "crctype", "E", "rv", "Qm", "HARQ_on",
                 "B", "Zc", "iLS", "K", "F", "N", "k0")] + [
        ("Vij_array", ctypes.POINTER(ctypes.c_uint16)),
        ("HBG_array", ctypes.POINTER(ctypes.c_uint16)),
        ("HBG_offset_array", ctypes.POINTER(ctypes.c_uint16))]


class CbDescriptor(ctypes.Structure):
    """ldpc_mi355_cb_t: one codeblock of a batch."""
    _fields_ = [("cfg", LdpcInConfig), ("reserved", ctypes.c_int16),
                ("llr_offset", ctypes.c_uint32), ("soft_slot", ctypes.c_int32)]


# ldpc_mi355_serve_cfg_t.llr_source
LLR_SOURCES = {"host_copy": 0, "device": 1, "host_mapped": 2}


class ServeCfg(ctypes.Structure):
    """ldpc_mi355_serve_cfg_t."""
    _fields_ = [("n_contexts", ctypes.c_int32), ("max_batch_codeblocks", ctypes.c_int32),
                ("n_ues", ctypes.c_int32), ("n_procs", ctypes.c_int32),
                ("max_cbs", ctypes.c_int32), ("llr_source", ctypes.c_int32),
                ("deadline_s", ctypes.c_double), ("harq_rtt_s", ctypes.c_double),
                ("max_wall_s", ctypes.c_double), ("llr_pool_elems", ctypes.c_uint64)]


class ServeTemplate(ctypes.Structure):
    """ldpc_mi355_serve_template_t."""
    _fields_ = [("n_cbs", ctypes.c_int32), ("cb_first", ctypes.c_int32),
                ("llr_elems", ctypes.c_uint64), ("llr_host", ctypes.c_uint64 * 4)]


class ServeStatsC(ctypes.Structure):
    """ldpc_mi355_serve_stats_t."""
    _fields_ = [("n_records", ctypes.c_int32), ("batches", ctypes.c_int32),
                ("admission_waits", ctypes.c_int32), ("max_blocked", ctypes.c_int32),
                ("max_inflight", ctypes.c_int32), ("reserved", ctypes.c_int32),
                ("n_cb_results", ctypes.c_int64), ("hard_bytes", ctypes.c_int64),
                ("peak_pool_elems", ctypes.c_uint64), ("wall_s", ctypes.c_double),
                ("host_launch_s", ctypes.c_double), ("host_collect_s", ctypes.c_double)]


# ldpc_mi355_serve_record_t as a numpy record
SERVE_RECORD_DTYPE = np.dtype([
    ("template_id", "<i4"), ("ue", "<i4"), ("proc", "<i4"), ("rv_index", "<i4"),
    ("n_cbs", "<i4"), ("n_pass", "<i4"), ("cb_first", "<i8"), ("hard_first", "<i8"),
    ("batch", "<i4"), ("final_tx", "<i4"), ("arrival_s", "<f8"), ("admitted_s", "<f8"),
    ("done_s", "<f8"), ("queued_s", "<f8")])
assert SERVE_RECORD_DTYPE.itemsize == 80


_lib = None


def build(quiet=True):
    """Compile the HIP library for gfx950 (hipcc cross-compiles without a GPU)."""
    import subprocess
    out = subprocess.run(["make", "-C", CSRC_DIR, "-j8"], capture_output=quiet, text=True)
    if out.returncode != 0:
        raise RuntimeError("libnrldpc_mi355 build failed:\n" + (out.stdout or "") +
                           (out.stderr or ""))


def load():
    """Load the library (once) and declare argument types."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(
            f"{LIB_PATH} is missing: build it with __graft_entry__.build() or "
            f"make -C {CSRC_DIR}; the decode path has no CPU fallback")
    # PyTorch-ROCm ships its own libamdhip64 (soname libamdhip64.so.7) and
    # reaches it under another file name; loading torch first makes this
    # library bind to that same HIP runtime, so device pointers and streams
    # are shared instead of living in two runtimes.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    L = ctypes.CDLL(LIB_PATH)
    P = ctypes.c_void_p
    L.calculate_crc.argtypes = [P, ctypes.c_int, ctypes.c_int]
    L.calculate_crc.restype = ctypes.c_uint32
    L.gen_ldpc_parameters.argtypes = [ctypes.POINTER(LdpcConfig), ctypes.POINTER(LdpcInConfig)]
    L.gen_ldpc_parameters.restype = ctypes.c_int
    L.dequantize_int16_to_llr.argtypes = [ctypes.c_int16]
    L.dequantize_int16_to_llr.restype = ctypes.c_float
    L.c_ldpc_5g_decode.argtypes = [P, P, ctypes.POINTER(LdpcInConfig),
                                   ctypes.POINTER(DecoderConfig)]
    L.c_ldpc_5g_decode.restype = ctypes.c_int32
    L.ldpc_mi355_decode_ex.argtypes = [P, P, ctypes.POINTER(LdpcInConfig),
                                       ctypes.POINTER(DecoderConfig), P, P]
    L.ldpc_mi355_decode_ex.restype = ctypes.c_int32
    L.c_ldpc_5g_encode.argtypes = [P, P, ctypes.POINTER(LdpcInConfig)]
    L.c_ldpc_5g_encode.restype = ctypes.c_int32
    L.cuda_ldpc_5g_decode.argtypes = [P, P, ctypes.POINTER(LdpcInConfig),
                                      ctypes.POINTER(DecoderConfig)]
    L.cuda_ldpc_5g_decode.restype = ctypes.c_int32
    L.cuda_ldpc_5g_encode.argtypes = [P, P, ctypes.POINTER(LdpcInConfig)]
    L.cuda_ldpc_5g_encode.restype = ctypes.c_int32
    L.tx_rx_process.argtypes = [P, ctypes.c_size_t, ctypes.c_uint16, ctypes.c_float, P]
    L.tx_rx_process.restype = ctypes.c_int
    L.ldpc_mi355_channel_seed.argtypes = [ctypes.c_uint64]
    L.ldpc_mi355_channel_seed.restype = None
    L.ldpc_mi355_plan_create.argtypes = [ctypes.POINTER(P), ctypes.POINTER(CbDescriptor),
                                         ctypes.c_int32, ctypes.POINTER(DecoderConfig),
                                         ctypes.c_uint32, ctypes.POINTER(ctypes.c_int32)]
    L.ldpc_mi355_plan_create.restype = ctypes.c_int32
    L.ldpc_mi355_plan_layout.argtypes = [P, P, P, P, P, P]
    L.ldpc_mi355_plan_layout.restype = ctypes.c_int32
    L.ldpc_mi355_plan_decode.argtypes = [P, P, P, P, P, P, P, P]
    L.ldpc_mi355_plan_decode.restype = ctypes.c_int32
    L.ldpc_mi355_plan_destroy.argtypes = [P]
    L.ldpc_mi355_plan_destroy.restype = None
    L.ldpc_mi355_sync.argtypes = []
    L.ldpc_mi355_sync.restype = ctypes.c_int32
    L.ldpc_mi355_soft_slot_elems.argtypes = []
    L.ldpc_mi355_soft_slot_elems.restype = ctypes.c_int32
    L.ldpc_mi355_plan_tiles.argtypes = [P]
    L.ldpc_mi355_plan_tiles.restype = ctypes.c_int32
    L.ldpc_mi355_set_tile_codeblocks.argtypes = [ctypes.c_int32] * 3
    L.ldpc_mi355_set_tile_codeblocks.restype = ctypes.c_int32
    L.ldpc_mi355_algo_available.argtypes = [ctypes.c_int32]
    L.ldpc_mi355_algo_available.restype = ctypes.c_int32
    L.ldpc_mi355_bp_decode_noalloc.argtypes = [ctypes.POINTER(LdpcInConfig),
                                               ctypes.POINTER(DecoderConfig), ctypes.c_int32,
                                               P, P, P, P, P, P, P]
    L.ldpc_mi355_bp_decode_noalloc.restype = ctypes.c_int32
    L.ldpc_mi355_serve_run.argtypes = [ctypes.POINTER(ServeCfg), ctypes.POINTER(DecoderConfig),
                                       ctypes.POINTER(ServeTemplate), ctypes.c_int32, P, P, P,
                                       P, P, P, P, ctypes.c_int32, P, ctypes.c_int32, P, P,
                                       ctypes.c_int64, P, ctypes.c_int64,
                                       ctypes.POINTER(ServeStatsC)]
    L.ldpc_mi355_serve_run.restype = ctypes.c_int32
    L.ldpc_mi355_version.argtypes = []
    L.ldpc_mi355_version.restype = ctypes.c_char_p
    L.ldpc_mi355_synth_batch.argtypes = [ctypes.POINTER(LdpcInConfig), ctypes.c_int32,
                                         ctypes.c_float, ctypes.c_uint64, ctypes.c_int32, P, P]
    L.ldpc_mi355_synth_batch.restype = ctypes.c_int32
    L.ldpc_mi355_synth_float_batch.argtypes = [ctypes.POINTER(LdpcInConfig), ctypes.c_int32, P,
                                               ctypes.c_uint64, ctypes.c_int32, P, P]
    L.ldpc_mi355_synth_float_batch.restype = ctypes.c_int32
    L.ldpc_mi355_synth_device.argtypes = [ctypes.POINTER(LdpcInConfig), ctypes.c_int32, P,
                                          ctypes.c_uint64] + [P] * 8
    L.ldpc_mi355_synth_device.restype = ctypes.c_int32
    L.ldpc_mi355_synth_device_noise.argtypes = [ctypes.POINTER(LdpcInConfig), ctypes.c_int32, P,
                                                P, ctypes.c_uint64] + [P] * 8
    L.ldpc_mi355_synth_device_noise.restype = ctypes.c_int32
    L.ldpc_mi355_unrolled_create.argtypes = [ctypes.POINTER(P), ctypes.POINTER(LdpcInConfig),
                                             ctypes.c_int32, ctypes.c_float, ctypes.c_int32]
    L.ldpc_mi355_unrolled_create.restype = ctypes.c_int32
    L.ldpc_mi355_unrolled_destroy.argtypes = [P]
    L.ldpc_mi355_unrolled_destroy.restype = None
    L.ldpc_mi355_unrolled_dims.argtypes = [P, P]
    L.ldpc_mi355_unrolled_dims.restype = ctypes.c_int32
    L.ldpc_mi355_unrolled_set_codeblocks_per_workgroup.argtypes = [P, ctypes.c_int32]
    L.ldpc_mi355_unrolled_set_codeblocks_per_workgroup.restype = ctypes.c_int32
    L.ldpc_mi355_unrolled_record_bytes.argtypes = [P, ctypes.c_int32]
    L.ldpc_mi355_unrolled_record_bytes.restype = ctypes.c_uint64
    L.ldpc_mi355_unrolled_scratch_bytes.argtypes = [P, ctypes.c_int32]
    L.ldpc_mi355_unrolled_scratch_bytes.restype = ctypes.c_uint64
    L.ldpc_mi355_unrolled_forward.argtypes = [P, ctypes.c_int32] + [P] * 7
    L.ldpc_mi355_unrolled_forward.restype = ctypes.c_int32
    L.ldpc_mi355_unrolled_backward.argtypes = [P, ctypes.c_int32] + [P] * 11
    L.ldpc_mi355_unrolled_backward.restype = ctypes.c_int32
    L.ldpc_mi355_batch_create.argtypes = [ctypes.POINTER(P), ctypes.c_int32,
                                          ctypes.POINTER(DecoderConfig), ctypes.c_uint32]
    L.ldpc_mi355_batch_create.restype = ctypes.c_int32
    L.ldpc_mi355_batch_decode.argtypes = [P, ctypes.POINTER(CbDescriptor), ctypes.c_int32] + \
        [P] * 8
    L.ldpc_mi355_batch_decode.restype = ctypes.c_int32
    L.ldpc_mi355_batch_layout.argtypes = [P] * 5
    L.ldpc_mi355_batch_layout.restype = ctypes.c_int32
    L.ldpc_mi355_batch_query.argtypes = [P]
    L.ldpc_mi355_batch_query.restype = ctypes.c_int32
    L.ldpc_mi355_batch_wait.argtypes = [P]
    L.ldpc_mi355_batch_wait.restype = ctypes.c_int32
    L.ldpc_mi355_batch_destroy.argtypes = [P]
    L.ldpc_mi355_batch_destroy.restype = None
    L.ldpc_mi355_copy_spans.argtypes = [P, P, P, ctypes.c_int32, ctypes.c_int32, P]
    L.ldpc_mi355_copy_spans.restype = ctypes.c_int32
    L.ldpc_mi355_set_trained_tables.argtypes = [ctypes.c_int32, P, P, ctypes.c_int32]
    L.ldpc_mi355_set_trained_tables.restype = ctypes.c_int32
    L.ldpc_mi355_get_trained_tables.argtypes = [ctypes.c_int32, P, P, ctypes.c_int32]
    L.ldpc_mi355_get_trained_tables.restype = ctypes.c_int32
    L.ldpc_mi355_set_degree_alpha.argtypes = [P]
    L.ldpc_mi355_set_degree_alpha.restype = ctypes.c_int32
    L.ldpc_mi355_get_degree_alpha.argtypes = [P]
    L.ldpc_mi355_get_degree_alpha.restype = ctypes.c_int32
    _lib = L
    return L


def ptr(a):
    """Address of a numpy array or torch tensor as c_void_p."""
    if hasattr(a, "data_ptr"):
        return ctypes.c_void_p(a.data_ptr())
    return a.ctypes.data_as(ctypes.c_void_p)
