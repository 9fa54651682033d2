"""GPU tests of batched decoding: packed tiles over every lifting size, the
plain vs packed kernels, hipGraph capture/replay.  Bit-exact vs the oracle."""
import numpy as np
import pytest

from test_kernel_emu import ALL_ZC, _all_zc_batch

pytestmark = pytest.mark.gpu

ROW1 = [1, 8, 22936, 0, 1280, 32767, -32767]


@pytest.fixture(scope="module")
def torch_dev():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a visible MI355X"
    from deeplearning_5g_ldpc_decoder_amd import _lib
    _lib.load()
    return torch


def _check(oracle_mod, plan, out, cfgs, llrs, row):
    hard = plan.split_hard(out["hard"].cpu().numpy())
    lq = plan.split_lq(out["lq"].cpu().numpy()) if "lq" in out else None
    st = out["status"].cpu().numpy()
    its = out["iters"].cpu().numpy()
    for i, (cfg, llr) in enumerate(zip(cfgs, llrs)):
        c = oracle_mod.code(*cfg)
        s, h, q, k = oracle_mod.decode_soft(c, row, oracle_mod.rate_recover(c, row[4], llr))
        assert (st[i], its[i]) == (s, k), (i, cfg)
        assert np.array_equal(hard[i], h), (i, cfg)
        if lq is not None:
            assert np.array_equal(lq[i], q), (i, cfg)


@pytest.mark.parametrize("flags_k", [(2, 0), (2, 4), (4, 0), (18, 0)])
def test_all_lifting_sizes_packed_and_plain(torch_dev, oracle_mod, flags_k):
    """Every reachable (BG, Zc), rates 0.25..0.85, Qm 2..8, rv 0..3 in one
    plan: packed tiles (default and k = 4 per tile, one workgroup per tile or
    a persistent grid pulling tiles from a device queue) and one codeblock per
    workgroup, all bit-exact with the oracle."""
    from deeplearning_5g_ldpc_decoder_amd import _lib
    from deeplearning_5g_ldpc_decoder_amd.batch import BatchPlan
    flags, k = flags_k
    L = _lib.load()
    rng = np.random.default_rng(77 + k)
    cfgs, llrs = _all_zc_batch(oracle_mod, rng, per_zc=3, zcs=ALL_ZC)
    assert len(cfgs) > 250
    for zc in ALL_ZC:
        for bgn in (1, 2):
            L.ldpc_mi355_set_tile_codeblocks(bgn, zc, k)
    try:
        plan = BatchPlan(cfgs, ROW1, want_lq=True, flags=flags,
                         schedule="persistent" if flags & 16 else "oneshot")
        assert (plan.tiles > 0) == bool(flags & 2)
        out = plan.decode(torch_dev.from_numpy(np.concatenate(llrs)).cuda())
        _check(oracle_mod, plan, out, cfgs, llrs, ROW1)
        plan.close()
    finally:
        for zc in ALL_ZC:
            for bgn in (1, 2):
                L.ldpc_mi355_set_tile_codeblocks(bgn, zc, 0)


def test_packed_general_rows(torch_dev, oracle_mod):
    from deeplearning_5g_ldpc_decoder_amd.batch import BatchPlan
    rng = np.random.default_rng(8)
    cfgs, llrs = _all_zc_batch(oracle_mod, rng, per_zc=1, zcs=ALL_ZC[::3])
    d = torch_dev.from_numpy(np.concatenate(llrs)).cuda()
    for row in ([1, 6, 40000, 16, 1280, 32767, -32767], [1, 6, 22936, 0, 1280, 300, 2000],
                [1, 0, 22936, 0, 1280, 32767, -32767]):
        plan = BatchPlan(cfgs, row, want_lq=True, flags=2)
        out = plan.decode(d)
        if row[1] == 0:
            st = out["status"].cpu().numpy()
            its = out["iters"].cpu().numpy()
            assert np.all(st == 1) and np.all(its == 0)
        else:
            _check(oracle_mod, plan, out, cfgs, llrs, row)
        plan.close()


@pytest.mark.parametrize("flags", [0, 18])
def test_hip_graph_capture_replay(torch_dev, oracle_mod, flags):
    """A packed plan captured in a HIP graph: every replay, with eager work
    and synchronizations in between, equals the eager launch.  flags 18: the
    persistent tile launch, whose device queue the last workgroup of each
    replay rearms."""
    torch = torch_dev
    from deeplearning_5g_ldpc_decoder_amd.batch import BatchPlan
    rng = np.random.default_rng(12)
    cfgs, llrs = _all_zc_batch(oracle_mod, rng, per_zc=1, zcs=ALL_ZC[::2])
    d_llr = torch.from_numpy(np.concatenate(llrs)).cuda()
    plan = BatchPlan(cfgs, ROW1, want_lq=True, flags=flags,
                     schedule="persistent" if flags & 16 else "oneshot")
    eager = plan.decode(d_llr)
    ref = {k: v.cpu().numpy() for k, v in eager.items()}
    out = plan.alloc_outputs("cuda")
    s = torch.cuda.Stream()
    with torch.cuda.stream(s):
        plan.launch(d_llr, out, stream=s.cuda_stream)
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g, stream=s):
        plan.launch(d_llr, out, stream=s.cuda_stream)
    for rep in range(4):
        for t in out.values():
            t.fill_(7 + rep)            # eager work between replays
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        for k, v in out.items():
            assert np.array_equal(v.cpu().numpy(), ref[k]), (rep, k)
    plan.close()


def test_copy_spans_round_trip(torch_dev):
    """ldpc_mi355_copy_spans: ragged pinned spans (one empty) into one device
    buffer at their own offsets, then back out in one span, on one stream."""
    import ctypes
    torch = torch_dev
    from deeplearning_5g_ldpc_decoder_amd import _lib
    L = _lib.load()
    rng = np.random.default_rng(4)
    sizes = [1, 0, 4097, 77, 65536]
    src = [torch.from_numpy(rng.integers(0, 256, max(s, 1)).astype(np.uint8)).pin_memory()
           for s in sizes]
    dev = torch.zeros(sum(sizes) + 64, dtype=torch.uint8, device="cuda")
    offs = np.concatenate([[0], np.cumsum(sizes)[:-1]])
    spans = np.array([[dev.data_ptr() + int(o) for o in offs], [t.data_ptr() for t in src], sizes],
                     np.uint64)
    torch.cuda.synchronize()   # the zero fill ran on the default stream
    stream = torch.cuda.Stream()
    sh = ctypes.c_void_p(stream.cuda_stream)
    assert L.ldpc_mi355_copy_spans(spans[0].ctypes.data, spans[1].ctypes.data,
                                   spans[2].ctypes.data, len(sizes), 1, sh) == 0
    back = torch.zeros(dev.numel(), dtype=torch.uint8).pin_memory()
    out = np.array([[back.data_ptr()], [dev.data_ptr()], [dev.numel()]], np.uint64)
    assert L.ldpc_mi355_copy_spans(out[0].ctypes.data, out[1].ctypes.data, out[2].ctypes.data,
                                   1, 2, sh) == 0
    stream.synchronize()
    b = back.numpy()
    for o, s, t in zip(offs, sizes, src):
        assert np.array_equal(b[o:o + s], t.numpy()[:s])
    assert not b[sum(sizes):].any()


@pytest.mark.parametrize("uniform,flags", [(True, 0), (False, 0), (False, 18)])
def test_overlapping_launches_of_one_plan(torch_dev, oracle_mod, uniform, flags):
    """Launches of one plan in flight on two streams, each with its own
    outputs (as the throughput bench runs them): every output set equals the
    single-stream decode.  uniform: BG1 Zc=384 codeblocks only (the
    specialized kernel); otherwise a mixed batch (packed tiles; flags 18:
    persistent tile launches, each on its own device queue)."""
    torch = torch_dev
    from deeplearning_5g_ldpc_decoder_amd.batch import BatchPlan
    rng = np.random.default_rng(21)
    if uniform:
        cfgs = [(1, 8424, 2, 25344, 0, 2, 0)] * 48
        llrs = [rng.integers(-900, 1100, 25344).astype(np.int16) for _ in cfgs]
    else:
        cfgs, llrs = _all_zc_batch(oracle_mod, rng, per_zc=1, zcs=ALL_ZC[1::3])
    d_llr = torch.from_numpy(np.concatenate(llrs)).cuda()
    plan = BatchPlan(cfgs, ROW1, want_lq=True, flags=flags,
                     schedule="persistent" if flags & 16 else "oneshot")
    ref = {k: v.cpu().numpy() for k, v in plan.decode(d_llr).items()}
    streams = [torch.cuda.Stream() for _ in range(2)]
    outs = [plan.alloc_outputs("cuda") for _ in range(2)]
    torch.cuda.synchronize()
    for k in range(8):
        plan.launch(d_llr, outs[k % 2], stream=streams[k % 2].cuda_stream)
    torch.cuda.synchronize()
    for o in outs:
        for k, v in o.items():
            assert np.array_equal(v.cpu().numpy(), ref[k]), k
    plan.close()


# CRC24B as in config 2, plus CRC16 and CRC24A codeblocks (and filler bits)
# for the CRC stop test's per-type bit weights
_SPECIALIZED_CFGS = [(2, 1016, 2, 2080, 0, 4, 0), (2, 2056, 2, 10400, 1, 2, 0),
                     (1, 8424, 2, 25344, 0, 2, 0), (1, 8432, 0, 25344, 0, 2, 0),
                     (1, 8424, 1, 25344, 0, 2, 0), (2, 1024, 0, 2080, 0, 4, 0),
                     (2, 2040, 1, 10400, 1, 2, 0)]
_SPECIALIZED_ROWS = [[1, 8, 22936, 0, 1280, 32767, -32767],
                     [1, 8, 26000, 40, -900, 2000, -1500],
                     [4, 20, 22936, 0, 1280, 32767, -32767]]


# the FP32 specialized kernels cover BG2 Zc 104/208 only
@pytest.mark.parametrize("cfg,row", [(c, r) for r in _SPECIALIZED_ROWS for c in _SPECIALIZED_CFGS
                                     if not (r[0] == 4 and c[0] == 1)])
def test_specialized_kernels_match_generic(torch_dev, oracle_mod, cfg, row):
    """Uniform batches of the lifting sizes with compile-time-shift kernels
    (BG2 Zc 104 and 208 for int16 fast rows and FP32 min-sum, BG1 Zc 384 for
    int16): the specialized kernel's outputs equal the generic kernel's
    (plan flag 8) bit for bit, and the oracle's on a sample."""
    torch = torch_dev
    from deeplearning_5g_ldpc_decoder_amd.batch import BatchPlan, synth_batch
    n = 96
    A, E = cfg[1], cfg[3]
    B = A + 24
    nz = -(10 * np.log10(2.0 ** (B / E * cfg[5]) - 1.0) + 1.6 + 0.3 * (cfg[5] - 2))
    llr, _ = synth_batch(cfg, n, nz, seed=31 + cfg[1])
    d = torch.from_numpy(llr.reshape(-1)).cuda()
    outs, lqs = {}, None
    for name, flags in (("specialized", 0), ("generic", BatchPlan.GENERIC)):
        plan = BatchPlan([cfg] * n, row, want_lq=True, flags=BatchPlan.NO_PACK | flags)
        outs[name] = {k: v.cpu().numpy() for k, v in plan.decode(d).items()}
        if name == "specialized":
            lqs = plan.split_lq(outs[name]["lq"])
        plan.close()
    for k in outs["generic"]:
        assert np.array_equal(outs["specialized"][k], outs["generic"][k]), k
    st, its = outs["specialized"]["status"], outs["specialized"]["iters"]
    assert 0 < int(np.sum(its > 1))
    c = oracle_mod.code(*cfg)
    for i in range(0, n, 16):
        s, h, q, k = oracle_mod.decode_soft(c, row, oracle_mod.rate_recover(c, row[4], llr[i]))
        assert (st[i], its[i]) == (s, k), i
        assert np.array_equal(lqs[i], q), i


def test_edge_lanes_match_oracle_and_tiles(torch_dev, oracle_mod):
    """The lane-per-edge schedule (plan flag 32: one wave per codeblock, a
    layer's edges one per lane) on every lifting size it takes, in one
    plan: bit-exact with the oracle and with the packed-tile decode of the
    same batch."""
    torch = torch_dev
    from test_kernel_emu import EDGE_ZC
    from deeplearning_5g_ldpc_decoder_amd.batch import BatchPlan
    rng = np.random.default_rng(57)
    cfgs, llrs = [], []
    for bgn, zcs in EDGE_ZC.items():
        c_, l_ = _all_zc_batch(oracle_mod, rng, per_zc=6, zcs=zcs)
        for c, l in zip(c_, l_):
            if c[0] == bgn:
                cfgs.append(c)
                llrs.append(l)
    d = torch.from_numpy(np.concatenate(llrs)).cuda()
    for row in (ROW1, [1, 8, 26000, 32, -900, 2000, -1500]):
        edge = BatchPlan(cfgs, row, want_lq=True, flags=BatchPlan.NO_PACK | BatchPlan.EDGE_LANES,
                         schedule="oneshot")
        tiles = BatchPlan(cfgs, row, want_lq=True, flags=BatchPlan.PACK, schedule="oneshot")
        assert tiles.tiles > 0
        out = edge.decode(d)
        ref = tiles.decode(d)
        for k in ref:
            assert np.array_equal(out[k].cpu().numpy(), ref[k].cpu().numpy()), (row, k)
        _check(oracle_mod, edge, out, cfgs, llrs, row)
        edge.close()
        tiles.close()
