"""Served decode loop on the GPU (deeplearning_5g_ldpc_decoder_amd/serving.py):
Poisson arrivals of mixed-size requests with HARQ retransmissions are batched
continuously over several launch contexts; afterwards every request is decoded
again alone, in the same per-process order, with fresh device HARQ buffers, and
must give the same status, iteration counts and hard bits bit for bit.  First
transmissions are also checked against the oracle."""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

ROW = [1, 8, 22936, 0, 1280, 32767, -32767]


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def test_serving_results_same_alone_and_in_the_mix(oracle_mod):
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from deeplearning_5g_ldpc_decoder_amd.serving import (RV_SEQUENCE, DevicePool, GpuExecutor,
                                                          HarqSlots, RequestPool, Server,
                                                          poisson_arrivals)
    _say("serving: synthesizing 24 request templates")
    pool = RequestPool(24, seed=2, max_cbs=16, snr_offset_db=-1.0, threads=16)
    harq = HarqSlots(16, 4, 16)
    llr_pool = DevicePool(64 << 20)
    ex = GpuExecutor(pool, harq, llr_pool, n_contexts=3, max_batch_codeblocks=96)
    srv = Server(ex, deadline_s=0.5e-3, harq_rtt_s=0.3e-3)
    rng = np.random.default_rng(4)
    arr = poisson_arrivals(20000.0, 0.02, seed=4)
    _say(f"serving: {len(arr)} arrivals")
    stats = srv.run(arr, rng.integers(0, len(pool), len(arr)), ues=rng.integers(0, 16, len(arr)))
    s = stats.summary(0.5e-3)
    _say(f"serving: {s}")
    reqs = stats.requests
    assert s["requests"] == len(reqs) and all(r.done >= 0 for r in reqs)
    assert s["max_batches_in_flight"] >= 2
    assert any(r.rv_index > 0 for r in reqs)
    ex.close()

    _say("serving: replaying every request alone")
    harq2 = HarqSlots(16, 4, 16)
    ex2 = GpuExecutor(pool, harq2, llr_pool, n_contexts=1, max_batch_codeblocks=96)
    for r in sorted(reqs, key=lambda q: q.admitted):
        q = type(r)(rid=r.rid, template=r.template, ue=r.ue, proc=r.proc, rv_index=r.rv_index,
                    arrival=0.0, deadline=0.0, n_cbs=r.n_cbs)
        region = llr_pool.alloc(pool.elems[q.template])
        ctx = ex2.free_context()
        ex2.launch(ctx, [q], region)
        torch.cuda.synchronize()
        assert ex2.poll() == [ctx]
        ex2.collect(ctx)
        llr_pool.release(*region)
        assert np.array_equal(q.status, r.status), r.rid
        assert np.array_equal(q.iters, r.iters), r.rid
        assert all(np.array_equal(a, b) for a, b in zip(q.hard, r.hard)), r.rid
    ex2.close()

    _say("serving: first transmissions against the oracle")
    O = oracle_mod
    for r in [q for q in reqs if q.rv_index == 0][:6]:
        llr, lens = pool.llr(r.template, 0)
        pos = 0
        for c, (cfg, ln) in enumerate(zip(pool.cfgs[r.template], lens)):
            out = O.decode_path(llr[pos:pos + ln], tuple(cfg[:4]) + (RV_SEQUENCE[0], cfg[5], 0),
                                ROW)
            assert out["status"] == r.status[c] and out["iters"] == r.iters[c]
            if out["iters"]:
                assert np.array_equal(out["hard"], r.hard[c])
            pos += ln


@pytest.mark.parametrize("llr_source", ["host_copy", "device", "host_mapped"])
def test_native_serving_results_same_alone_and_in_the_mix(oracle_mod, llr_source):
    """The C++ served loop (ldpc_mi355_serve_run) on Poisson arrivals with
    HARQ retransmissions over 4 contexts, request LLRs copied from pinned host
    memory, read in place in HBM or read in place in pinned host memory: every
    transmission, replayed alone in the same per-process order with fresh HARQ
    buffers, gives the same status, iterations and hard bits; first
    transmissions match the oracle."""
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from deeplearning_5g_ldpc_decoder_amd.serving import (RV_SEQUENCE, DevicePool, GpuExecutor,
                                                          HarqSlots, NativeServer, RequestPool,
                                                          poisson_arrivals)
    _say("native serving: synthesizing 24 request templates")
    pool = RequestPool(24, seed=3, max_cbs=16, snr_offset_db=-1.0, threads=16)
    harq = HarqSlots(16, 4, 16)
    llr_pool = DevicePool(64 << 20)
    srv = NativeServer(pool, harq, llr_pool, n_contexts=4, max_batch_codeblocks=96,
                       harq_rtt_s=0.3e-3, record_hard=True, llr_source=llr_source)
    rng = np.random.default_rng(6)
    arr = poisson_arrivals(20000.0, 0.02, seed=6)
    _say(f"native serving: {len(arr)} arrivals")
    stats = srv.run(arr, rng.integers(0, len(pool), len(arr)), ues=rng.integers(0, 16, len(arr)))
    s = stats.summary(0.5e-3)
    _say(f"native serving: {s}")
    reqs = stats.requests
    assert sum(1 for r in reqs if r.rv_index == 0) == len(arr)
    assert all(r.done >= r.admitted >= r.arrival for r in reqs)
    assert s["max_batches_in_flight"] >= 2
    assert any(r.rv_index > 0 for r in reqs)
    # every chain ends with a final transmission
    assert sum(1 for r in reqs if r.final) == len(arr)

    _say("native serving: replaying every transmission alone")
    harq2 = HarqSlots(16, 4, 16)
    ex2 = GpuExecutor(pool, harq2, llr_pool, n_contexts=1, max_batch_codeblocks=96)
    for r in sorted(reqs, key=lambda q: (q.admitted, q.rid)):
        q = type(r)(rid=r.rid, template=r.template, ue=r.ue, proc=r.proc, rv_index=r.rv_index,
                    arrival=0.0, deadline=0.0, n_cbs=r.n_cbs)
        region = llr_pool.alloc(pool.elems[q.template])
        ctx = ex2.free_context()
        ex2.launch(ctx, [q], region)
        torch.cuda.synchronize()
        assert ex2.poll() == [ctx]
        ex2.collect(ctx)
        llr_pool.release(*region)
        assert np.array_equal(q.status, r.status), r.rid
        assert np.array_equal(q.iters, r.iters), r.rid
        assert all(np.array_equal(a, b) for a, b in zip(q.hard, r.hard)), r.rid
    ex2.close()

    O = oracle_mod
    for r in [q for q in reqs if q.rv_index == 0][:6]:
        llr, lens = pool.llr(r.template, 0)
        pos = 0
        for c, (cfg, ln) in enumerate(zip(pool.cfgs[r.template], lens)):
            out = O.decode_path(llr[pos:pos + ln], tuple(cfg[:4]) + (RV_SEQUENCE[0], cfg[5], 0),
                                ROW)
            assert out["status"] == r.status[c] and out["iters"] == r.iters[c]
            pos += ln
