"""Scale tables of the table-driven int16 decoders (algos 2 and 3).

The reference compiles them into its library:

  algo 2  ``BG<n>_trained_alpha_list`` / ``_beta_list``, int16
          [iterations][rows] (c_lib/BG1_trained_ldpc_weights.h:7-17), read
          at ``iteration * rows + row``
          (c_lib/neural_fixedpoint_ldpc_decoder.c:248-250);
  algo 3  ``default_degree_minsum_alpha_values``, MAX_DEGREE Q15 values by
          exact check degree (c_lib/degree_fixedpoint_ldpc_decoder.c:19-43),
          or a table passed to ``layered_degree_minsum_ldpc_decode_with_alpha``
          (c_lib/nr_ldpc.h:113).

Here they are process-wide state of libnrldpc_mi355.so with the same
defaults, set at run time.  ``c_ldpc_5g_decode`` reads the tables current
at the call; plans and batch contexts keep the tables current at their
creation.
"""
import numpy as np

from . import _lib

NUM_ROWS = {1: 46, 2: 42}
MAX_DEGREE = 19


def set_trained_tables(bgn, alpha=None, beta=None):
    """Algo-2 tables of base graph bgn: int16 alpha, beta [iterations, rows]
    (Q15 alpha, beta in int16 LLR units); alpha None restores the defaults."""
    L = _lib.load()
    if alpha is None:
        st = L.ldpc_mi355_set_trained_tables(int(bgn), None, None, 0)
    else:
        a = np.ascontiguousarray(alpha, np.int16)
        b = np.ascontiguousarray(np.zeros_like(a) if beta is None else beta, np.int16)
        if a.ndim != 2 or a.shape != b.shape or a.shape[1] != NUM_ROWS.get(bgn, -1):
            raise ValueError(f"tables must be [iterations, {NUM_ROWS.get(bgn)}] int16")
        st = L.ldpc_mi355_set_trained_tables(int(bgn), _lib.ptr(a), _lib.ptr(b), a.shape[0])
    if st != 0:
        raise ValueError(f"ldpc_mi355_set_trained_tables returned {st}")


def get_trained_tables(bgn):
    """Current algo-2 tables of base graph bgn: (alpha, beta) int16
    [iterations, rows]."""
    L = _lib.load()
    it = L.ldpc_mi355_get_trained_tables(int(bgn), None, None, 0)
    if it < 0:
        raise ValueError(f"invalid base graph {bgn}")
    a = np.zeros((it, NUM_ROWS[bgn]), np.int16)
    b = np.zeros_like(a)
    L.ldpc_mi355_get_trained_tables(int(bgn), _lib.ptr(a), _lib.ptr(b), it)
    return a, b


def set_degree_alpha(values=None):
    """Algo-3 Q15 alpha by check degree (entry d-1 for degree d, MAX_DEGREE
    values); None restores the defaults."""
    L = _lib.load()
    if values is None:
        L.ldpc_mi355_set_degree_alpha(None)
        return
    v = np.ascontiguousarray(values, np.int16)
    if v.shape != (MAX_DEGREE,):
        raise ValueError(f"degree alpha must hold {MAX_DEGREE} values")
    L.ldpc_mi355_set_degree_alpha(_lib.ptr(v))


def get_degree_alpha():
    v = np.zeros(MAX_DEGREE, np.int16)
    _lib.load().ldpc_mi355_get_degree_alpha(_lib.ptr(v))
    return v
