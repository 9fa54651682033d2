/*
 * Test-only probe around the reference fixed-point decoder.
 *
 * The reference keeps its variable-node beliefs in a file-static array
 * (c_lib/fixedpoint_ldpc_decoder.c:19), so the posterior LLRs after a decode
 * are not observable through its public API.  This translation unit compiles
 * that reference source file where it lies (the include path is set by
 * oracle/Makefile.ref) and adds one accessor for the array.  The resulting
 * library lives only under oracle/_ref/ and is used solely to write golden
 * fixtures and to cross-check the oracle.
 */
#include "fixedpoint_ldpc_decoder.c"

const int16_t *ref_probe_fixed_lq(void) { return LQ; }
