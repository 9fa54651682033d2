"""Trainable unrolled layered min-sum on the MI355X.

The module mirrors the reference's neural decoder API
(deep_learning/deeplearning_ldpc_model.py, ``CTableNeuralLayeredMinSumDecoder``
:72-431, and deep_learning/ldpc_table_config.py, ``load_ldpc_table_config``
:110-140): same constructor arguments, parameter names (``raw_alpha``,
``raw_beta``), ``alpha()``/``beta()`` mappings and ``forward`` outputs.  The
difference is where the work runs: the unrolled iterations and their gradient
are the library's HIP kernels (``ldpc_mi355_unrolled_forward`` /
``_backward``) behind one ``torch.autograd.Function``, instead of ~20 torch
ops per base-graph row and a full ``LQ.clone()`` per row.  PyTorch holds the
parameters, the optimizer and the device buffers; the decode runs only on
the GPU (a CPU tensor is an error, there is no fallback).

Parameter sharing.  The kernels take alpha and beta as one value per
(iteration, base-graph row).  The modes map onto that table with plain torch
indexing, so autograd sums the table's gradient back into the parameters:

  ``single``            alpha [1]            every (t, r)
  ``checkpoint``        alpha [T, rows]      (t, r)  (reference :290-309)
  ``degree``            alpha [degrees]      row r's degree, every t
  ``iteration_degree``  alpha [T, degrees]   (t, degree of r)  (new)

beta is ``[T, rows]`` as in the reference (``beta_mode="checkpoint"``) or
``[T, degrees]`` with ``beta_mode="iteration_degree"``.
"""
import ctypes
from dataclasses import dataclass
from typing import Optional, Sequence, Tuple

import numpy as np

from . import _lib
from ._lib import LdpcConfig, LdpcInConfig, ptr

BG1_NUM_ROWS, BG1_NUM_COLS, BG1_MAX_NON_NULL_ELEMENTS = 46, 68, 316
BG2_NUM_ROWS, BG2_NUM_COLS, BG2_MAX_NON_NULL_ELEMENTS = 42, 52, 197

ALPHA_MODES = ("single", "checkpoint", "degree", "iteration_degree")
BETA_MODES = ("checkpoint", "iteration_degree")


@dataclass(frozen=True)
class LdpcTableConfig:
    """Sparse LDPC table metadata (deep_learning/ldpc_table_config.py:57-79)."""
    bgn: int
    A: int
    crctype: int
    E: int
    rv: int
    Qm: int
    HARQ_on: int
    B: int
    Zc: int
    iLS: int
    K: int
    F: int
    N: int
    k0: int
    num_rows: int
    num_cols: int
    num_edges: int
    hbg_columns: Tuple[int, ...]
    vij: Tuple[int, ...]
    row_offsets: Tuple[int, ...]


def load_ldpc_table_config(bgn, A, crctype, E, rv, Qm, HARQ_on=0, lib_path=None):
    """Derived parameters and base-graph tables from this library's
    ``gen_ldpc_parameters`` (deep_learning/ldpc_table_config.py:110-140).
    ``lib_path`` is accepted for signature compatibility; the tables always
    come from libnrldpc_mi355."""
    L = _lib.load()
    cfg = LdpcConfig()
    st = L.gen_ldpc_parameters(ctypes.byref(cfg),
                               ctypes.byref(LdpcInConfig(bgn, A, crctype, E, rv, Qm, HARQ_on)))
    if st != 0:
        raise ValueError("gen_ldpc_parameters failed for the supplied LDPC config")
    if cfg.bgn == 1:
        rows, cols, edges = BG1_NUM_ROWS, BG1_NUM_COLS, BG1_MAX_NON_NULL_ELEMENTS
    else:
        rows, cols, edges = BG2_NUM_ROWS, BG2_NUM_COLS, BG2_MAX_NON_NULL_ELEMENTS
    return LdpcTableConfig(
        bgn=int(cfg.bgn), A=int(cfg.A), crctype=int(cfg.crctype), E=int(cfg.E),
        rv=int(cfg.rv), Qm=int(cfg.Qm), HARQ_on=int(cfg.HARQ_on), B=int(cfg.B),
        Zc=int(cfg.Zc), iLS=int(cfg.iLS), K=int(cfg.K), F=int(cfg.F), N=int(cfg.N),
        k0=int(cfg.k0), num_rows=rows, num_cols=cols, num_edges=edges,
        hbg_columns=tuple(int(cfg.HBG_array[i]) for i in range(edges)),
        vij=tuple(int(cfg.Vij_array[i]) for i in range(edges)),
        row_offsets=tuple(int(cfg.HBG_offset_array[i]) for i in range(rows + 1)))


def row_degrees(table_config):
    """Non-null base-graph entries per row (deeplearning_ldpc_model.py:51-56)."""
    ro = table_config.row_offsets
    return [ro[r + 1] - ro[r] for r in range(table_config.num_rows)]


def degree_alpha_indices(degrees):
    """Sorted unique degrees and each row's index into them (:58-70)."""
    unique = tuple(sorted(set(degrees)))
    pos = {d: i for i, d in enumerate(unique)}
    return unique, [pos[d] for d in degrees]


# --------------------------------------------------------------------------
# Library handle
# --------------------------------------------------------------------------
class UnrolledKernel:
    """One ``ldpc_mi355_unrolled_t``: the kernels for one LDPC config and
    iteration count."""

    def __init__(self, table_config, num_iterations, clip_llr=None, detach=False):
        L = _lib.load()
        self._h = ctypes.c_void_p()
        ic = LdpcInConfig(table_config.bgn, table_config.A, table_config.crctype, table_config.E,
                          table_config.rv, table_config.Qm, table_config.HARQ_on)
        clip = -1.0 if clip_llr is None else float(clip_llr)
        st = L.ldpc_mi355_unrolled_create(ctypes.byref(self._h), ctypes.byref(ic),
                                          int(num_iterations), ctypes.c_float(clip),
                                          1 if detach else 0)
        if st != 0:
            raise RuntimeError(f"ldpc_mi355_unrolled_create failed ({st})")
        d = np.zeros(8, np.int32)
        L.ldpc_mi355_unrolled_dims(self._h, ptr(d))
        self.n, self.rows, self.edges, self.Zc, self.B, self.T = (int(v) for v in d[:6])
        self.codeblocks_per_workgroup = int(d[6])
        self.lanes_per_codeblock = int(d[7])

    def set_codeblocks_per_workgroup(self, k):
        L = _lib.load()
        if L.ldpc_mi355_unrolled_set_codeblocks_per_workgroup(self._h, int(k)) != 0:
            raise ValueError(f"{k} codeblocks per workgroup do not fit")
        d = np.zeros(8, np.int32)
        L.ldpc_mi355_unrolled_dims(self._h, ptr(d))
        self.codeblocks_per_workgroup = int(d[6])

    def record_bytes(self, batch):
        return int(_lib.load().ldpc_mi355_unrolled_record_bytes(self._h, int(batch)))

    def scratch_bytes(self, batch):
        return int(_lib.load().ldpc_mi355_unrolled_scratch_bytes(self._h, int(batch)))

    def forward(self, llr, alpha, beta, want_iters=False, want_record=True):
        """llr [batch, n] float32 device tensor, alpha/beta [T, rows].
        Returns (lq, iter_lq or None, record or None)."""
        import torch
        batch = llr.shape[0]
        lq = torch.empty_like(llr)
        it = (torch.empty((self.T, batch, self.n), dtype=llr.dtype, device=llr.device)
              if want_iters else None)
        rec = (torch.empty(max(self.record_bytes(batch), 1), dtype=torch.uint8, device=llr.device)
               if want_record else None)
        stream = torch.cuda.current_stream(llr.device).cuda_stream
        st = _lib.load().ldpc_mi355_unrolled_forward(
            self._h, batch, ptr(llr), ptr(alpha), ptr(beta), ptr(lq),
            ptr(it) if it is not None else None, ptr(rec) if rec is not None else None,
            ctypes.c_void_p(stream))
        if st != 0:
            raise RuntimeError(f"ldpc_mi355_unrolled_forward failed ({st})")
        return lq, it, rec

    def backward(self, llr, alpha, beta, record, g_lq, g_iter=None, want_g_llr=True):
        """Returns (g_llr or None, g_alpha [T, rows], g_beta [T, rows])."""
        import torch
        batch = llr.shape[0]
        g_llr = torch.empty_like(llr) if want_g_llr else None
        ga = torch.empty((self.T, self.rows), dtype=torch.float32, device=llr.device)
        gb = torch.empty_like(ga)
        scratch = torch.empty(max(self.scratch_bytes(batch), 1), dtype=torch.uint8,
                              device=llr.device)
        stream = torch.cuda.current_stream(llr.device).cuda_stream
        st = _lib.load().ldpc_mi355_unrolled_backward(
            self._h, batch, ptr(llr), ptr(alpha), ptr(beta), ptr(record), ptr(g_lq),
            ptr(g_iter) if g_iter is not None else None,
            ptr(g_llr) if g_llr is not None else None, ptr(ga), ptr(gb), ptr(scratch),
            ctypes.c_void_p(stream))
        if st != 0:
            raise RuntimeError(f"ldpc_mi355_unrolled_backward failed ({st})")
        return g_llr, ga, gb

    def close(self):
        if getattr(self, "_h", None) and self._h.value:
            _lib.load().ldpc_mi355_unrolled_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _check_device_tensor(x, name):
    if not x.is_cuda:
        raise RuntimeError(f"{name} must be a device tensor: the unrolled decoder runs only on "
                           "the GPU (libnrldpc_mi355), there is no CPU path")


def _autograd_function():
    import torch

    class UnrolledMinSumFunction(torch.autograd.Function):
        """(llr [batch, n], alpha [T, rows], beta [T, rows]) ->
        (final beliefs [batch, n], per-iteration beliefs [T, batch, n] or an
        empty tensor)."""

        @staticmethod
        def forward(ctx, llr, alpha, beta, kernel, want_iters):
            lq, it, rec = kernel.forward(llr, alpha, beta, want_iters=want_iters,
                                         want_record=True)
            ctx.kernel = kernel
            ctx.want_iters = want_iters
            ctx.save_for_backward(llr, alpha, beta, rec)
            if it is None:
                it = llr.new_empty((0,))
                ctx.mark_non_differentiable(it)
            return lq, it

        @staticmethod
        def backward(ctx, g_lq, g_it):
            llr, alpha, beta, rec = ctx.saved_tensors
            if g_lq is None:
                g_lq = torch.zeros_like(llr)
            g_lq = g_lq.contiguous().float()
            if not ctx.want_iters or g_it is None or g_it.numel() == 0:
                g_it = None
            else:
                g_it = g_it.contiguous().float()
            g_llr, ga, gb = ctx.kernel.backward(llr, alpha, beta, rec, g_lq, g_it,
                                                want_g_llr=ctx.needs_input_grad[0])
            return (g_llr if ctx.needs_input_grad[0] else None,
                    ga if ctx.needs_input_grad[1] else None,
                    gb if ctx.needs_input_grad[2] else None, None, None)

    return UnrolledMinSumFunction


_FN = None


def unrolled_min_sum(llr, alpha, beta, kernel, want_iters=False):
    """Differentiable unrolled min-sum: returns (lq, iter_lq or None)."""
    global _FN
    if _FN is None:
        _FN = _autograd_function()
    _check_device_tensor(llr, "llr")
    T, R = kernel.T, kernel.rows
    if tuple(alpha.shape) != (T, R) or tuple(beta.shape) != (T, R):
        raise ValueError(f"alpha and beta must have shape ({T}, {R})")
    if llr.dim() != 2 or llr.shape[1] != kernel.n:
        raise ValueError(f"llr must have shape [batch, {kernel.n}]")
    lq, it = _FN.apply(llr.contiguous().float(), alpha.contiguous().float(),
                       beta.contiguous().float(), kernel, bool(want_iters))
    return lq, (it if want_iters else None)


# --------------------------------------------------------------------------
# The module
# --------------------------------------------------------------------------
def _inverse_sigmoid(value):
    import torch
    v = torch.tensor(float(value)).clamp(1e-6, 1.0 - 1e-6)
    return torch.log(v / (1.0 - v))


def _inverse_bounded_sigmoid_tensor(values, lo, hi):
    import torch
    values = torch.as_tensor(values, dtype=torch.float32)
    if bool(torch.any(values <= lo)) or bool(torch.any(values >= hi)):
        raise ValueError(f"bounded sigmoid values must satisfy {lo} < value < {hi}")
    x = ((values - lo) / (hi - lo)).clamp(1e-6, 1.0 - 1e-6)
    return torch.log(x / (1.0 - x))


def _module_base():
    import torch.nn as nn
    return nn.Module


class CTableNeuralLayeredMinSumDecoder(_module_base()):
    """Neural layered min-sum decoder on the c_lib table layout, trained on
    the GPU (deep_learning/deeplearning_ldpc_model.py:72-431).

    Args follow the reference: num_iterations, learnable_scaling,
    learnable_beta, init_alpha, init_beta, alpha_mode, alpha_min, alpha_max,
    init_alpha_degree_values, clip_llr, detach_check_update_inputs; plus
    ``beta_mode`` ("checkpoint" as in the reference, or "iteration_degree")
    and the extra alpha mode "iteration_degree" (one alpha per iteration and
    row degree).  Input beliefs: [batch, N + 2Zc] (or [N + 2Zc]) on the GPU.
    """

    def __init__(self, table_config, num_iterations=8, learnable_scaling=True,
                 learnable_beta=True, init_alpha=0.75, init_beta=0.0, alpha_mode="single",
                 alpha_min=0.5, alpha_max=1.0 - 1e-6, init_alpha_degree_values=None,
                 clip_llr=None, detach_check_update_inputs=False, beta_mode="checkpoint"):
        import torch
        import torch.nn as nn
        super().__init__()
        if alpha_mode not in ALPHA_MODES:
            raise ValueError("alpha_mode must be 'single', 'checkpoint', 'degree' or "
                             "'iteration_degree'")
        if beta_mode not in BETA_MODES:
            raise ValueError("beta_mode must be 'checkpoint' or 'iteration_degree'")
        if not 0.0 < alpha_min < alpha_max < 1.0:
            raise ValueError("alpha_min and alpha_max must satisfy 0 < alpha_min < alpha_max < 1")
        self.config = table_config
        self.num_iterations = int(num_iterations)
        self.clip_llr = clip_llr
        self.detach_check_update_inputs = bool(detach_check_update_inputs)
        self.alpha_mode = alpha_mode
        self.beta_mode = beta_mode
        self.alpha_min = float(alpha_min)
        self.alpha_max = float(alpha_max)
        self.n = table_config.N + 2 * table_config.Zc
        self.B = table_config.B
        self.num_rows = table_config.num_rows
        self.num_edges = table_config.num_edges
        self.Zc = table_config.Zc
        degrees = row_degrees(table_config)
        self._row_degrees = degrees
        self.unique_row_degrees, idx = degree_alpha_indices(degrees)
        self.register_buffer("row_alpha_degree_indices", torch.tensor(idx, dtype=torch.long))
        raw_alpha = self._make_raw_alpha(init_alpha, init_alpha_degree_values)
        nb = len(self.unique_row_degrees) if beta_mode == "iteration_degree" else self.num_rows
        raw_beta = _inverse_sigmoid(init_beta).repeat(self.num_iterations, nb)
        if learnable_scaling:
            self.raw_alpha = nn.Parameter(raw_alpha)
        else:
            self.register_buffer("raw_alpha", raw_alpha)
        if learnable_beta:
            self.raw_beta = nn.Parameter(raw_beta)
        else:
            self.register_buffer("raw_beta", raw_beta)
        self._kernel = None

    # -- parameters (reference :186-313) ---------------------------------
    def _make_raw_alpha(self, init_alpha, init_alpha_degree_values):
        import torch
        lo, hi = self.alpha_min, self.alpha_max
        T, R = self.num_iterations, self.num_rows
        U = len(self.unique_row_degrees)
        scalar = isinstance(init_alpha, (int, float))
        if self.alpha_mode in ("single", "checkpoint") and init_alpha_degree_values is not None:
            raise ValueError("init_alpha_degree_values is only valid for the degree modes")
        if self.alpha_mode == "single":
            if scalar:
                return _inverse_bounded_sigmoid_tensor([float(init_alpha)], lo, hi)
            return self._table(init_alpha, (1,))
        if self.alpha_mode == "checkpoint":
            if scalar:
                return _inverse_bounded_sigmoid_tensor(
                    torch.full((T, R), float(init_alpha)), lo, hi)
            return self._table(init_alpha, (T, R))
        per_degree = None
        if init_alpha_degree_values is not None:
            v = torch.as_tensor(init_alpha_degree_values, dtype=torch.float32)
            if v.dim() != 1:
                raise ValueError("init_alpha_degree_values must be a one-dimensional sequence")
            if v.numel() == U:
                per_degree = v
            elif v.numel() >= max(self.unique_row_degrees):
                per_degree = torch.stack([v[d - 1] for d in self.unique_row_degrees])
            else:
                raise ValueError(
                    "init_alpha_degree_values must contain either one value per unique row "
                    f"degree {self.unique_row_degrees} or a 1-based exact-degree table with at "
                    f"least {max(self.unique_row_degrees)} entries; got {v.numel()}")
        if self.alpha_mode == "degree":
            if per_degree is not None:
                return _inverse_bounded_sigmoid_tensor(per_degree, lo, hi)
            if scalar:
                return _inverse_bounded_sigmoid_tensor(torch.full((U,), float(init_alpha)), lo, hi)
            return self._table(init_alpha, (U,))
        # iteration_degree
        if per_degree is not None:
            return _inverse_bounded_sigmoid_tensor(per_degree.repeat(T, 1), lo, hi)
        if scalar:
            return _inverse_bounded_sigmoid_tensor(torch.full((T, U), float(init_alpha)), lo, hi)
        return self._table(init_alpha, (T, U))

    def _table(self, values, shape):
        import torch
        t = torch.as_tensor(values, dtype=torch.float32)
        if tuple(t.shape) != tuple(shape):
            raise ValueError(f"init_alpha must be a scalar or have shape {tuple(shape)} when "
                             f"alpha_mode='{self.alpha_mode}'; got {tuple(t.shape)}")
        return _inverse_bounded_sigmoid_tensor(t, self.alpha_min, self.alpha_max)

    def alpha_parameter_shape(self):
        return tuple(self.raw_alpha.shape)

    def alpha(self):
        """Bounded alpha values in the shape of the sharing mode (:285-288)."""
        import torch
        a = self.alpha_min + (self.alpha_max - self.alpha_min) * torch.sigmoid(self.raw_alpha)
        return a.clamp(min=self.alpha_min, max=self.alpha_max)

    def beta(self):
        """beta values, [T, rows] or [T, degrees] (:311-313)."""
        import torch
        return torch.sigmoid(self.raw_beta).clamp(min=0.0, max=1.0 - 1e-6)

    def alpha_table(self):
        """alpha per (iteration, row), [T, rows] (the kernels' layout)."""
        a = self.alpha()
        T, R = self.num_iterations, self.num_rows
        if self.alpha_mode == "single":
            return a.reshape(1, 1).expand(T, R)
        if self.alpha_mode == "checkpoint":
            return a
        idx = self.row_alpha_degree_indices.to(a.device)
        if self.alpha_mode == "degree":
            return a[idx].unsqueeze(0).expand(T, R)
        return a[:, idx]

    def beta_table(self):
        b = self.beta()
        if self.beta_mode == "checkpoint":
            return b
        return b[:, self.row_alpha_degree_indices.to(b.device)]

    # -- decode ----------------------------------------------------------
    def kernel(self):
        if self._kernel is None:
            self._kernel = UnrolledKernel(self.config, self.num_iterations, self.clip_llr,
                                          self.detach_check_update_inputs)
        return self._kernel

    def forward(self, llr, return_llr=False, return_iteration_llrs=False):
        """Decode a batch (:315-410).  Returns hard bits (int64), and the final
        beliefs and/or the list of per-iteration beliefs when asked."""
        import torch
        single = False
        if llr.dim() == 1:
            llr = llr.unsqueeze(0)
            single = True
        if llr.dim() != 2:
            raise ValueError("llr must have shape [n] or [batch, n]")
        if llr.shape[1] != self.n:
            raise ValueError(f"expected llr length {self.n}, got {llr.shape[1]}")
        _check_device_tensor(llr, "llr")
        lq, it = unrolled_min_sum(llr, self.alpha_table(), self.beta_table(), self.kernel(),
                                  want_iters=return_iteration_llrs)
        final_llr = lq
        bits = (final_llr < 0).to(torch.int64)
        iteration_llrs = list(it.unbind(0)) if return_iteration_llrs else None
        if single:
            bits = bits.squeeze(0)
            final_llr = final_llr.squeeze(0)
            if iteration_llrs is not None:
                iteration_llrs = [x.squeeze(0) for x in iteration_llrs]
        if return_llr and return_iteration_llrs:
            return bits, final_llr, iteration_llrs
        if return_llr:
            return bits, final_llr
        if return_iteration_llrs:
            return bits, iteration_llrs
        return bits

    def trainable_parameters(self):
        return [(name, tuple(p.shape)) for name, p in self.named_parameters() if p.requires_grad]

    def memory_summary(self, batch_size, dtype_bytes=4):
        """Device memory of one training step in MB: beliefs, the forward
        record kept for the backward, and the backward scratch (the reference
        reports LQ and Lr message memory, :418-426)."""
        k = self.kernel()
        lq = batch_size * self.n * dtype_bytes / 2 ** 20
        return {"LQ_MB": lq,
                "record_MB": k.record_bytes(batch_size) / 2 ** 20,
                "scratch_MB": k.scratch_bytes(batch_size) / 2 ** 20,
                "total_message_MB": lq + (k.record_bytes(batch_size) +
                                          k.scratch_bytes(batch_size)) / 2 ** 20}


# --------------------------------------------------------------------------
# Export to the trained-table decoder's int16 format (algo 2)
# --------------------------------------------------------------------------
ALGO2_NUM_ITERATIONS = 32


def algo2_tables(model, num_iterations=ALGO2_NUM_ITERATIONS, beta_scale=128.0):
    """int16 [num_iterations, rows] alpha and beta tables of the trained
    decoder (c_lib/neural_fixedpoint_ldpc_decoder.c:223-281, index
    iteration * rows + row, :248): alpha_fixed = round(alpha * 32767) (Q15,
    as c_lib/BG1_trained_ldpc_weights.c documents), beta_fixed =
    round(beta * beta_scale): the decoder subtracts beta from int16
    magnitudes in units of LLR * 128 (:115-124), so 128 maps the float model's
    beta onto the same LLR offset.  Iterations past the trained ones repeat
    the last trained iteration."""
    import torch
    with torch.no_grad():
        a = model.alpha_table().detach().double().cpu().numpy()
        b = model.beta_table().detach().double().cpu().numpy()
    T = a.shape[0]
    if T == 0:
        raise ValueError("the model has no iterations")
    pick = np.minimum(np.arange(num_iterations), T - 1)
    a = a[pick]
    b = b[pick]
    fa = np.clip(np.round(a * 32767.0), -32768, 32767).astype(np.int16)
    fb = np.clip(np.round(b * beta_scale), -32768, 32767).astype(np.int16)
    return fa, fb


def _c_rows(vals, per_line=14):
    flat = [str(int(v)) for v in vals.reshape(-1)]
    lines = []
    for i in range(0, len(flat), per_line):
        lines.append("    " + ", ".join(flat[i:i + per_line]) + ",")
    return "\n".join(lines)


def write_algo2_weight_files(bgn, alpha, beta, out_dir, meta=None, note=""):
    """Write BG<bgn>_trained_ldpc_weights.h/.c in the layout of
    c_lib/BG1_trained_ldpc_weights.{h,c}: int16 alpha / beta tables
    [iterations, rows] flattened iteration-major, with the NUM_ITERATIONS,
    NUM_ROWS and ALPHA_BETA_LIST_SIZE macros and the trained-config metadata
    macros (meta: dict with A, B, E, Zc, crctype, rv, Qm; zeros if absent).
    Returns the two paths."""
    import os
    alpha = np.asarray(alpha)
    beta = np.asarray(beta)
    rows = BG1_NUM_ROWS if bgn == 1 else BG2_NUM_ROWS
    if alpha.shape != beta.shape or alpha.ndim != 2 or alpha.shape[1] != rows:
        raise ValueError(f"tables must be [iterations, {rows}]")
    if alpha.dtype != np.int16 or beta.dtype != np.int16:
        raise ValueError("tables must be int16")
    iters = alpha.shape[0]
    m = dict(A=0, B=0, E=0, Zc=0, crctype=0, rv=0, Qm=0)
    m.update(meta or {})
    tag = f"BG{bgn}_TRAINED_LDPC"
    name = f"BG{bgn}_trained_ldpc_weights"
    os.makedirs(out_dir, exist_ok=True)
    h = f"""#ifndef {name.upper()}_H
#define {name.upper()}_H

#include <stdint.h>

#define {tag}_BGN {bgn}
#define {tag}_NUM_ITERATIONS {iters}
#define {tag}_NUM_ROWS {rows}

/* table_index = iteration * {tag}_NUM_ROWS + HBG_row */
#define {tag}_ALPHA_BETA_LIST_SIZE {iters * rows}

/* configuration the weights were trained for (metadata) */
#define {tag}_A {m['A']}
#define {tag}_B {m['B']}
#define {tag}_E {m['E']}
#define {tag}_ZC {m['Zc']}
#define {tag}_CRC_TYPE {m['crctype']}
#define {tag}_RV {m['rv']}
#define {tag}_QM {m['Qm']}

extern const int16_t BG{bgn}_trained_alpha_list[{tag}_ALPHA_BETA_LIST_SIZE];
extern const int16_t BG{bgn}_trained_beta_list[{tag}_ALPHA_BETA_LIST_SIZE];

#endif /* {name.upper()}_H */
"""
    c = f"""#include "{name}.h"

/*
 * Min-sum weights, [iteration][HBG row]:
 *   alpha_fixed = round(alpha * 32767)   (Q15)
 *   beta_fixed  = beta in int16 LLR units
{note} */
const int16_t BG{bgn}_trained_alpha_list[{tag}_ALPHA_BETA_LIST_SIZE] = {{
{_c_rows(alpha)}
}};

const int16_t BG{bgn}_trained_beta_list[{tag}_ALPHA_BETA_LIST_SIZE] = {{
{_c_rows(beta)}
}};
"""
    hp = os.path.join(out_dir, name + ".h")
    cp = os.path.join(out_dir, name + ".c")
    with open(hp, "w") as f:
        f.write(h)
    with open(cp, "w") as f:
        f.write(c)
    return hp, cp


def export_algo2_tables(model, out_dir, num_iterations=ALGO2_NUM_ITERATIONS, beta_scale=128.0):
    """Write BG<n>_trained_ldpc_weights.h/.c (the layout of
    c_lib/BG1_trained_ldpc_weights.{h,c}) for the model's base graph.
    Returns the two paths."""
    cfg = model.config
    fa, fb = algo2_tables(model, num_iterations, beta_scale)
    note = (f" * beta_fixed = round(beta * {beta_scale:g}); trained iterations: "
            f"{model.num_iterations}, later iterations repeat the last one.\n")
    meta = dict(A=cfg.A, B=cfg.B, E=cfg.E, Zc=cfg.Zc, crctype=cfg.crctype, rv=cfg.rv, Qm=cfg.Qm)
    return write_algo2_weight_files(cfg.bgn, fa, fb, out_dir, meta, note)


def load_algo2_tables(model, num_iterations=ALGO2_NUM_ITERATIONS, beta_scale=128.0):
    """Installs the model's exported tables (algo2_tables) as the library's
    algo-2 tables for its base graph, so c_ldpc_5g_decode with algo 2 (and
    plans / batch contexts created afterwards) decodes with them -- the
    reference's export, rebuild and decode loop without the rebuild.
    Returns (alpha, beta)."""
    from .tables import set_trained_tables
    fa, fb = algo2_tables(model, num_iterations, beta_scale)
    set_trained_tables(model.config.bgn, fa, fb)
    return fa, fb


def synth_float_batch(table_config, noise_power_db, seed=0, threads=8):
    """Host synthesis of training inputs: float beliefs [n, N+2Zc] and target
    bits [n, B] (payload + CRC) for per-codeblock noise powers."""
    L = _lib.load()
    nz = np.ascontiguousarray(noise_power_db, np.float32).reshape(-1)
    n = nz.size
    c = table_config
    llr = np.zeros((n, c.N + 2 * c.Zc), np.float32)
    bits = np.zeros((n, c.B), np.uint8)
    ic = LdpcInConfig(c.bgn, c.A, c.crctype, c.E, c.rv, c.Qm, c.HARQ_on)
    st = L.ldpc_mi355_synth_float_batch(ctypes.byref(ic), n, ptr(nz), ctypes.c_uint64(seed),
                                        int(threads), ptr(llr), ptr(bits))
    if st != 0:
        raise ValueError("invalid config")
    return llr, bits


def synth_float_batch_device(table_config, noise_power_db, seed=0):
    """Device synthesis of training inputs (ldpc_mi355_synth_device): float
    beliefs [n, N+2Zc] and target bits [n, B] as float32 device tensors, for
    per-codeblock noise powers (a sequence of n values)."""
    import torch
    from .batch import synth_device
    c = table_config
    nz = np.ascontiguousarray(noise_power_db, np.float32).reshape(-1)
    out = synth_device((c.bgn, c.A, c.crctype, c.E, c.rv, c.Qm, c.HARQ_on), nz.size, nz,
                       seed=seed, outputs=("bits", "beliefs"))
    return out["beliefs"], out["bits"].to(torch.float32)
