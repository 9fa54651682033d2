// The lifting-size specialized int16 decode kernels (ldpc_decode_fixed_kernel
// in ldpc_kernels.hip), compiled here on their own so the library's kernels
// build in parallel: this unit instantiates only them, with their
// configuration (nrldpc_configure_fixed) and lookup (nrldpc_fixed_kernel).
#define NRLDPC_FIXED_TU
#include "ldpc_kernels.hip"
