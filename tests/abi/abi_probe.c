/* Layout probe for the C-ABI boundary: compiled against a header given as
 * ABI_HEADER (the reference's c_lib/nr_ldpc.h when the fixture is written,
 * include/nr_ldpc_mi355.h in the test) and run; prints sizeof / offsetof /
 * field sizes of the boundary structs, the sizes of the exported tables and
 * the header's constants as JSON.  The function-pointer assignments fail to
 * compile (-Werror) when a boundary prototype's types differ from these. */
#include <stddef.h>
#include <stdint.h>
#include <stdio.h>
#include ABI_HEADER

#define FIELD(T, f) printf("  \"%s.%s\": [%zu, %zu],\n", #T, #f, offsetof(T, f), sizeof(((T *)0)->f))

int main(void) {
    int32_t (*dec)(int16_t *, uint8_t *, ldpc_in_config_t *, ldpc_decoder_cfg_t *) =
        c_ldpc_5g_decode;
    int32_t (*enc)(uint8_t *, uint8_t *, ldpc_in_config_t *) = c_ldpc_5g_encode;
    int (*chan)(const uint8_t *, size_t, uint16_t, float, int16_t *) = tx_rx_process;
    uint32_t (*crc)(const uint8_t *, int, int) = calculate_crc;
    int (*gen)(ldpc_config_t *, ldpc_in_config_t *) = gen_ldpc_parameters;
    float (*deq)(int16_t) = dequantize_int16_to_llr;
    (void)dec; (void)enc; (void)chan; (void)crc; (void)gen; (void)deq;

    printf("{\n");
    printf("  \"sizeof.ldpc_in_config_t\": %zu,\n", sizeof(ldpc_in_config_t));
    printf("  \"sizeof.ldpc_config_t\": %zu,\n", sizeof(ldpc_config_t));
    printf("  \"sizeof.ldpc_decoder_cfg_t\": %zu,\n", sizeof(ldpc_decoder_cfg_t));
    FIELD(ldpc_in_config_t, bgn); FIELD(ldpc_in_config_t, A); FIELD(ldpc_in_config_t, crctype);
    FIELD(ldpc_in_config_t, E); FIELD(ldpc_in_config_t, rv); FIELD(ldpc_in_config_t, Qm);
    FIELD(ldpc_in_config_t, HARQ_on);
    FIELD(ldpc_config_t, bgn); FIELD(ldpc_config_t, A); FIELD(ldpc_config_t, crctype);
    FIELD(ldpc_config_t, E); FIELD(ldpc_config_t, rv); FIELD(ldpc_config_t, Qm);
    FIELD(ldpc_config_t, HARQ_on); FIELD(ldpc_config_t, B); FIELD(ldpc_config_t, Zc);
    FIELD(ldpc_config_t, iLS); FIELD(ldpc_config_t, K); FIELD(ldpc_config_t, F);
    FIELD(ldpc_config_t, N); FIELD(ldpc_config_t, k0); FIELD(ldpc_config_t, Vij_array);
    FIELD(ldpc_config_t, HBG_array); FIELD(ldpc_config_t, HBG_offset_array);
    FIELD(ldpc_decoder_cfg_t, algo); FIELD(ldpc_decoder_cfg_t, max_iterations);
    FIELD(ldpc_decoder_cfg_t, alpha); FIELD(ldpc_decoder_cfg_t, beta);
    FIELD(ldpc_decoder_cfg_t, filling_bit_value); FIELD(ldpc_decoder_cfg_t, max_llr);
    FIELD(ldpc_decoder_cfg_t, min_llr);
    printf("  \"sizeof.BG1_HBG_offset_array\": %zu,\n", sizeof(BG1_HBG_offset_array));
    printf("  \"sizeof.BG1_HBG_array\": %zu,\n", sizeof(BG1_HBG_array));
    printf("  \"sizeof.BG1_HBG_Vij_array\": %zu,\n", sizeof(BG1_HBG_Vij_array));
    printf("  \"sizeof.BG2_HBG_offset_array\": %zu,\n", sizeof(BG2_HBG_offset_array));
    printf("  \"sizeof.BG2_HBG_array\": %zu,\n", sizeof(BG2_HBG_array));
    printf("  \"sizeof.BG2_HBG_Vij_array\": %zu,\n", sizeof(BG2_HBG_Vij_array));
    printf("  \"sizeof.HARQ_memory\": %zu,\n", sizeof(HARQ_memory));
    printf("  \"BG1_NUM_ROWS\": %d, \"BG1_NUM_COLS\": %d, \"BG1_MAX_NON_NULL_ELEMENTS\": %d,\n",
           BG1_NUM_ROWS, BG1_NUM_COLS, BG1_MAX_NON_NULL_ELEMENTS);
    printf("  \"BG2_NUM_ROWS\": %d, \"BG2_NUM_COLS\": %d, \"BG2_MAX_NON_NULL_ELEMENTS\": %d,\n",
           BG2_NUM_ROWS, BG2_NUM_COLS, BG2_MAX_NON_NULL_ELEMENTS);
    printf("  \"MAX_DEGREE\": %d, \"MAX_ZC\": %d\n", MAX_DEGREE, MAX_ZC);
    printf("}\n");
    return 0;
}
