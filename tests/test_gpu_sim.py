"""run_simulation of the harness mirror on the GPU library, bound the way the
reference's drivers bind their libraries: the c_lib symbols
(scripts/sim_c_ldpc_encoder_decoder.py) and the CUDA symbols with the channel
from a second library path (scripts/sim_cuda_ldpc_encoder_decoder.py:57-68).
The result pickle has the reference's layout, and far above / below the
waterfall every block decodes / fails."""
import pickle

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("symbols", [("c_ldpc_5g_encode", "c_ldpc_5g_decode", False),
                                     ("cuda_ldpc_5g_encode", "cuda_ldpc_5g_decode", True)])
def test_run_simulation_on_gpu_library(tmp_path, symbols):
    from deeplearning_5g_ldpc_decoder_amd import ldpc_sim as S
    enc, dec, second_lib = symbols
    np.random.seed(11)
    grid = S.make_config_grid([2], [0], [256], [1 / 3], [1, 2], [1])
    rows = [[1, 8, 22936, 0, 1280, 32767, -32767], [4, 20, 22936, 0, 1280, 32767, -32767]]
    snr = [-12.0, 8.0]
    out = tmp_path / "res.pkl"
    S.run_simulation(grid, rows, snr, 16, str(out), encode_symbol=enc, decode_symbol=dec,
                     channel_lib_path=S.LIB_PATH if second_lib else None)
    with open(out, "rb") as f:
        test_result, dec_list, snr_list, test_num = pickle.load(f)
    assert dec_list == rows and snr_list == snr and test_num == 16
    assert [r[0] for r in test_result] == grid
    for _, bler, ber in test_result:
        assert np.all(bler[0] == 1.0)           # -12 dB: nothing decodes
        assert np.all(bler[1] == 0.0) and np.all(ber[1] == 0.0)   # +8 dB: everything does
