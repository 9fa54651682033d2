/*
 * Test-only probe around the reference degree min-sum decoder (algo 3).
 *
 * c_lib/degree_fixedpoint_ldpc_decoder.c:46 keeps the int16 beliefs in a
 * file-static array.  This unit compiles that reference source where it lies
 * and exposes the array read-only, so golden fixtures can record algo-3
 * posteriors.  Built only into oracle/_ref/.
 */
#include "degree_fixedpoint_ldpc_decoder.c"

const int16_t *ref_probe_degree_lq(void) { return degree_LQ; }
