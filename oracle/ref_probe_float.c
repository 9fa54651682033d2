/*
 * Test-only probe around the reference floating-point decoders.
 *
 * c_lib/floatpoint_ldpc_decoder.c:21 keeps the FP32 beliefs of algo 0 (BP) and
 * algo 4 (float min-sum) in a file-static array.  This unit compiles that
 * reference source where it lies and exposes the array read-only, so golden
 * fixtures can record FP32 posteriors.  Built only into oracle/_ref/.
 */
#include "floatpoint_ldpc_decoder.c"

const float *ref_probe_float_lq(void) { return LQ; }
