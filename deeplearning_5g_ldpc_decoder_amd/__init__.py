"""MI355X-native 5G NR LDPC decode path.

A drop-in for the reference library's ctypes surface (c_ldpc_5g_decode and
friends, see include/nr_ldpc_mi355.h) backed by hand-written HIP kernels for
gfx950, plus batched plans, a differentiable min-sum and a serving loop.
"""
from ._lib import (CbDescriptor, DecoderConfig, LdpcConfig, LdpcInConfig, LIB_PATH,  # noqa
                   build, load)

__all__ = ["CbDescriptor", "DecoderConfig", "LdpcConfig", "LdpcInConfig", "LIB_PATH",
           "build", "load"]
