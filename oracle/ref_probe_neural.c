/*
 * Test-only probe around the reference trained-table decoder (algo 2).
 *
 * c_lib/neural_fixedpoint_ldpc_decoder.c:30 keeps the int16 beliefs in a
 * file-static array.  This unit compiles that reference source where it lies
 * and exposes the array read-only, so golden fixtures can record algo-2
 * posteriors.  Built only into oracle/_ref/.
 */
#include "neural_fixedpoint_ldpc_decoder.c"

const int16_t *ref_probe_neural_lq(void) { return neural_LQ; }
