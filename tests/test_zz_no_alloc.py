"""The allocation-free FP32 BP entry point (ldpc_mi355_bp_decode_noalloc)
against the plan path it mirrors (ldpc_mi355_plan_create/_decode, algo 0)
through the C ABI: bit-identical hard bytes, status, iterations, posteriors
and soft-buffer slots, on two calls in a row, the second with a smaller batch
of another code, so the second reuses the library's message array after the
first.  Oversized batches and wrong arguments are refused on the host."""
import ctypes

import numpy as np
import pytest

from deeplearning_5g_ldpc_decoder_amd import _lib
from deeplearning_5g_ldpc_decoder_amd._lib import DecoderConfig, LdpcInConfig

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

ROW0 = [0, 12, 1, 0, 1280, 20, -32768]


def _dptr(t):
    return ctypes.c_void_p(t.data_ptr())


def _noalloc(L, cfg, llr, soft):
    bgn, A, crc, E, rv, Qm, h = cfg
    B = A + (16 if crc == 0 else 24)
    n = llr.shape[0]
    d_llr = torch.from_numpy(np.ascontiguousarray(llr.reshape(-1))).cuda()
    hard = torch.zeros(n * (B // 8), dtype=torch.uint8, device="cuda")
    st = torch.full((n,), -7, dtype=torch.int32, device="cuda")
    it = torch.full((n,), -7, dtype=torch.int32, device="cuda")
    slots = soft.shape[0] // (68 * 384)
    assert slots >= n
    from oracle import oracle as O
    c = O.code(*cfg)
    lq = torch.zeros(n * (c.N + 2 * c.Zc), dtype=torch.float32, device="cuda")
    rc = L.ldpc_mi355_bp_decode_noalloc(ctypes.byref(LdpcInConfig(*cfg)),
                                        ctypes.byref(DecoderConfig(*ROW0)), n, _dptr(d_llr),
                                        _dptr(soft), _dptr(hard), _dptr(st), _dptr(it), _dptr(lq),
                                        None)
    assert rc == 0
    torch.cuda.synchronize()
    return dict(hard=hard.cpu().numpy(), status=st.cpu().numpy(), iters=it.cpu().numpy(),
                lq=lq.cpu().numpy(), soft=soft.cpu().numpy())


def _plan(cfg, llr, soft):
    from deeplearning_5g_ldpc_decoder_amd.batch import BatchPlan
    n = llr.shape[0]
    plan = BatchPlan([cfg] * n, ROW0, soft_slots=list(range(n)), want_lq=True)
    out = plan.decode(torch.from_numpy(np.ascontiguousarray(llr.reshape(-1))).cuda(),
                      soft_pool=soft)
    torch.cuda.synchronize()
    res = {k: v.cpu().numpy() for k, v in out.items()}
    res["soft"] = soft.cpu().numpy()
    plan.close()
    return res


@pytest.mark.parametrize("calls", [[((2, 1016, 2, 5200, 0, 2, 0), 48, -2.5, 11),
                                    ((1, 488, 1, 1536, 0, 2, 0), 12, -1.0, 12)]])
def test_noalloc_bp_matches_plan_twice(calls):
    from deeplearning_5g_ldpc_decoder_amd.batch import synth_batch
    L = _lib.load()
    for cfg, n, noise, seed in calls:
        print(f"noalloc BP cfg={cfg} n={n}", flush=True)
        llr, _ = synth_batch(cfg, n, noise, seed=seed, want_payload=False)
        pool_a = torch.full((n * 68 * 384,), 3, dtype=torch.int16, device="cuda")
        pool_b = pool_a.clone()
        got = _noalloc(L, cfg, llr, pool_a)
        want = _plan(cfg, llr, pool_b)
        assert np.array_equal(got["status"], want["status"]), cfg
        assert np.array_equal(got["iters"], want["iters"]), cfg
        assert np.array_equal(got["hard"], want["hard"]), cfg
        assert np.array_equal(got["lq"].view(np.uint32), want["lq"].view(np.uint32)), cfg
        assert np.array_equal(got["soft"], want["soft"]), cfg
        assert np.any(got["status"] == 0)


def test_noalloc_bp_refuses_on_host():
    L = _lib.load()
    cfg = LdpcInConfig(1, 8424, 2, 25344, 0, 2, 0)
    row = DecoderConfig(*ROW0)
    dummy = ctypes.c_void_p(16)   # never dereferenced: every call below is refused first
    # 70 BG1 Zc=384 codeblocks need more than the 32 MiB message array
    assert L.ldpc_mi355_bp_decode_noalloc(ctypes.byref(cfg), ctypes.byref(row), 70, dummy, None,
                                          dummy, dummy, dummy, None, None) == -4
    assert L.ldpc_mi355_bp_decode_noalloc(ctypes.byref(cfg), ctypes.byref(row), 0, dummy, None,
                                          dummy, dummy, dummy, None, None) == -4
    row4 = DecoderConfig(4, 8, 22936, 0, 1280, 32767, -32767)
    assert L.ldpc_mi355_bp_decode_noalloc(ctypes.byref(cfg), ctypes.byref(row4), 4, dummy, None,
                                          dummy, dummy, dummy, None, None) == -3
    bad = LdpcInConfig(3, 8424, 2, 25344, 0, 2, 0)
    assert L.ldpc_mi355_bp_decode_noalloc(ctypes.byref(bad), ctypes.byref(row), 4, dummy, None,
                                          dummy, dummy, dummy, None, None) == -2
    harq = LdpcInConfig(1, 8424, 2, 25344, 0, 2, 1)
    assert L.ldpc_mi355_bp_decode_noalloc(ctypes.byref(harq), ctypes.byref(row), 4, dummy, None,
                                          dummy, dummy, dummy, None, None) == -4
