"""Served decode loop on the MI355X: continuous batching, admission by free
device memory, HARQ soft buffers resident on the device.

Requests are slots of 1-64 codeblocks of mixed base graph / lifting size /
rate / modulation, each from one UE's HARQ process and carrying a deadline.
The loop (``Server.run``) runs in real time:

* new-data arrivals (Poisson process) wait for a free HARQ process of their
  UE, then join a FIFO decode queue; retransmissions join it when due.
  Latency is counted from joining that queue (``Request.latency``) and, as
  well, from arrival (``Request.latency_from_arrival``);
* whenever a launch context is free, the queue head is admitted as one batch
  for as long as the batch stays within ``max_batch_codeblocks`` and its
  LLRs fit in the device LLR pool (``DevicePool``, first-fit over one HBM
  allocation): admission is by free device memory, and requests that do not
  fit wait until a running batch frees its region;
* each batch gets its own launch context (``ldpc_mi355_batch_*``) and HIP
  stream: its LLRs are copied host->device, decoded (int16 layered
  min-sum, CRC early stop) and its results copied back, all asynchronously,
  so batches admitted while others run overlap on the device;
* completions are polled; a codeblock that fails CRC is retransmitted with
  the next redundancy version of the sequence 0, 2, 3, 1 after
  ``harq_rtt_s``, soft-combined on the device into the same HARQ slot
  (HARQ_on = 1); slot = (UE, process, codeblock index), ``ldpc_mi355``
  soft-buffer slots of 68*384 int16 values.

Results of a request do not depend on what it was batched with: every
codeblock's decode reads only its own LLRs and HARQ slot (tests check this
against each request decoded alone).
"""
import bisect
import ctypes
import time
from dataclasses import dataclass, field
from typing import List, Optional

import numpy as np

from . import _lib
from ._lib import CbDescriptor, DecoderConfig, LdpcInConfig, ptr

RV_SEQUENCE = (0, 2, 3, 1)
DEFAULT_ROW = (1, 8, 22936, 0, 1280, 32767, -32767)

# ldpc_mi355_cb_t as a numpy record (24 bytes): descriptors of a batch are
# built by concatenating per-template records instead of field by field.
CB_DTYPE = np.dtype([("cfg", "<u2", (7,)), ("reserved", "<i2"), ("llr_offset", "<u4"),
                     ("soft_slot", "<i4")])
assert CB_DTYPE.itemsize == 24


# --------------------------------------------------------------------------
# Request material
# --------------------------------------------------------------------------
class RequestPool:
    """Pre-synthesized request templates (host side).  Template k holds
    n_k codeblocks (1..max_cbs) drawn like the mixed-size workload (base graph,
    lifting size, rate, Qm, per-codeblock Es/N0 around the rate's waterfall
    + ``snr_offset_db``), their payloads, and the channel LLRs of each
    redundancy version in RV_SEQUENCE (same payload and noise seed per rv)."""

    def __init__(self, n_templates, seed=0, max_cbs=64, snr_offset_db=0.0, threads=16,
                 pinned=True):
        from concurrent.futures import ThreadPoolExecutor

        from .batch import synth_batch
        from .workloads import draw_config
        rng = np.random.default_rng(seed)
        self.cfgs, self.esn0, self.sizes = [], [], []
        for _ in range(n_templates):
            k = int(rng.integers(1, max_cbs + 1))
            draws = [draw_config(rng) for _ in range(k)]
            self.cfgs.append([d[0] for d in draws])
            self.esn0.append([d[1] + snr_offset_db for d in draws])
            self.sizes.append(k)
        # ordered so that each (template, rv index) is one contiguous run of
        # its codeblocks' LLRs
        jobs = [(t, c, rvi) for t in range(n_templates) for rvi in range(len(RV_SEQUENCE))
                for c in range(self.sizes[t])]

        def one(job):
            t, c, rvi = job
            cfg = list(self.cfgs[t][c])
            cfg[4] = RV_SEQUENCE[rvi]
            llr, pay = synth_batch(tuple(cfg), 1, -self.esn0[t][c],
                                   seed=(seed * 7919 + t) * 131 + c, threads=1)
            return llr[0], pay[0]

        with ThreadPoolExecutor(max_workers=max(1, threads)) as ex:
            out = list(ex.map(one, jobs))
        # one flat int16 array: template t, rv index r -> its codeblocks' LLRs
        self.offsets = {}
        total = sum(len(o[0]) for o in out)
        flat = np.empty(total, np.int16)
        self.payloads = [[None] * self.sizes[t] for t in range(n_templates)]
        pos = 0
        it = iter(zip(jobs, out))
        for (t, c, rvi), (llr, pay) in it:
            self.offsets.setdefault((t, rvi), []).append((pos, len(llr)))
            flat[pos:pos + len(llr)] = llr
            pos += len(llr)
            if rvi == 0:
                self.payloads[t][c] = pay
        self.flat = flat
        self.pinned = None
        if pinned:
            import torch
            self.pinned = torch.from_numpy(flat).pin_memory()
            self.flat = self.pinned.numpy()
        self.info_bits = [sum(c[1] for c in cfgs) for cfgs in self.cfgs]
        self.cb_info = [np.array([c[1] for c in cfgs], np.int64) for cfgs in self.cfgs]
        self.cb_hard = [np.array([(c[1] + (16 if c[2] == 0 else 24)) // 8 for c in cfgs])
                        for cfgs in self.cfgs]
        self.elems = [sum(c[3] for c in cfgs) for cfgs in self.cfgs]
        # descriptor records per (template, rv index): llr_offset relative to
        # the request's first LLR, soft_slot = codeblock index (the caller adds
        # the base of the UE/process slots)
        self.desc = {}
        for t, cfgs in enumerate(self.cfgs):
            lens = np.array([c[3] for c in cfgs], np.int64)
            rel = np.concatenate([[0], np.cumsum(lens)[:-1]])
            for rvi, rv in enumerate(RV_SEQUENCE):
                d = np.zeros(len(cfgs), CB_DTYPE)
                d["cfg"] = [(c[0], c[1], c[2], c[3], rv, c[5], 1 if rvi > 0 else 0)
                            for c in cfgs]
                d["llr_offset"] = rel
                d["soft_slot"] = np.arange(len(cfgs))
                self.desc[(t, rvi)] = d

    def __len__(self):
        return len(self.cfgs)

    def span(self, t, rvi):
        """[a, b) of template t's LLRs for rv index rvi in the flat array."""
        parts = self.offsets[(t, rvi)]
        return parts[0][0], parts[-1][0] + parts[-1][1]

    def llr(self, t, rvi):
        """Concatenated LLRs of template t's codeblocks for rv index rvi."""
        a, b = self.span(t, rvi)
        return self.flat[a:b], [p[1] for p in self.offsets[(t, rvi)]]


@dataclass
class Request:
    rid: int
    template: int
    ue: int
    proc: int
    rv_index: int
    arrival: float
    deadline: float
    n_cbs: int = 0
    info_bits: int = 0
    queued: float = -1.0      # joined the decode queue (-1: at arrival)
    admitted: float = -1.0
    done: float = -1.0
    batch_id: int = -1
    status: Optional[np.ndarray] = None
    iters: Optional[np.ndarray] = None
    hard_bytes: Optional[np.ndarray] = None   # the request's hard decisions, back to back
    hard_len: Optional[np.ndarray] = None     # B/8 per codeblock
    final: bool = False                       # last transmission of its HARQ chain
    _hard: Optional[List[np.ndarray]] = None

    @property
    def latency(self):
        """Decode latency: from joining the decode queue to the results on
        the host.  New data joins it once its UE has a free HARQ process (a
        UE sends a transport block only on one); a retransmission when it is
        due."""
        return self.done - (self.queued if self.queued >= 0 else self.arrival)

    @property
    def latency_from_arrival(self):
        """From arrival, including the wait of new data for a free HARQ
        process of its UE."""
        return self.done - self.arrival

    @property
    def hard(self):
        """Per-codeblock B/8-byte hard decisions."""
        if self._hard is None and self.hard_bytes is not None:
            ends = np.cumsum(self.hard_len)
            self._hard = np.split(self.hard_bytes, ends[:-1])
        return self._hard

    @hard.setter
    def hard(self, v):
        self._hard = v

    @property
    def ok(self):
        return self.status is not None and bool(np.all(self.status == 0))


# --------------------------------------------------------------------------
# Device memory: LLR pool (admission) and HARQ slots
# --------------------------------------------------------------------------
class DevicePool:
    """First-fit allocator over one device int16 tensor of ``nbytes``: the
    admission budget.  Offsets are in int16 elements, 128-element aligned."""

    ALIGN = 128

    def __init__(self, nbytes, device="cuda", tensor=None):
        self.elems = int(nbytes) // 2 // self.ALIGN * self.ALIGN
        if tensor is None:
            import torch
            tensor = torch.empty(self.elems, dtype=torch.int16, device=device)
        self.tensor = tensor
        self.free = [(0, self.elems)]   # sorted (offset, length)
        self.used = 0
        self.peak = 0

    def alloc(self, elems):
        need = (int(elems) + self.ALIGN - 1) // self.ALIGN * self.ALIGN
        for i, (off, ln) in enumerate(self.free):
            if ln >= need:
                if ln == need:
                    self.free.pop(i)
                else:
                    self.free[i] = (off + need, ln - need)
                self.used += need
                self.peak = max(self.peak, self.used)
                return off, need
        return None

    def release(self, off, need):
        i = bisect.bisect_left(self.free, (off, 0))
        self.free.insert(i, (off, need))
        # coalesce with neighbours
        if i + 1 < len(self.free) and self.free[i][0] + self.free[i][1] == self.free[i + 1][0]:
            self.free[i] = (self.free[i][0], self.free[i][1] + self.free[i + 1][1])
            self.free.pop(i + 1)
        if i > 0 and self.free[i - 1][0] + self.free[i - 1][1] == self.free[i][0]:
            self.free[i - 1] = (self.free[i - 1][0], self.free[i - 1][1] + self.free[i][1])
            self.free.pop(i)
        self.used -= need

    @property
    def free_bytes(self):
        return 2 * (self.elems - self.used)


class HarqSlots:
    """Device int16 soft buffers, one slot per (UE, HARQ process, codeblock)."""

    def __init__(self, n_ues=16, n_procs=4, max_cbs=64, device="cuda"):
        import torch
        self.n_ues, self.n_procs, self.max_cbs = n_ues, n_procs, max_cbs
        self.slot_elems = int(_lib.load().ldpc_mi355_soft_slot_elems())
        self.tensor = torch.zeros(n_ues * n_procs * max_cbs * self.slot_elems, dtype=torch.int16,
                                  device=device)

    def slot(self, ue, proc, cb):
        return (ue * self.n_procs + proc) * self.max_cbs + cb

    @property
    def nbytes(self):
        return self.tensor.numel() * 2


# --------------------------------------------------------------------------
# GPU executor: launch contexts with their own streams
# --------------------------------------------------------------------------
class _Context:
    def __init__(self, idx, cap_cbs, row, device):
        import torch
        L = _lib.load()
        self.idx = idx
        self.handle = ctypes.c_void_p()
        st = L.ldpc_mi355_batch_create(ctypes.byref(self.handle), cap_cbs,
                                       ctypes.byref(DecoderConfig(*row)), 0)
        if st != 0:
            raise RuntimeError(f"ldpc_mi355_batch_create failed ({st})")
        self.stream = torch.cuda.Stream(device=device)    # default priority
        self.cap = cap_cbs
        # results in one span (status, iterations, hard bytes), so that one
        # copy brings a batch's results back
        nbytes = 8 * cap_cbs + cap_cbs * 1056
        self.out_dev = torch.empty(nbytes, dtype=torch.uint8, device=device)
        self.status = self.out_dev[:4 * cap_cbs].view(torch.int32)
        self.iters = self.out_dev[4 * cap_cbs:8 * cap_cbs].view(torch.int32)
        self.hard = self.out_dev[8 * cap_cbs:]
        self.out_host = torch.empty(nbytes, dtype=torch.uint8).pin_memory()
        h = self.out_host.numpy()
        self.h_status = h[:4 * cap_cbs].view(np.int32)
        self.h_iters = h[4 * cap_cbs:8 * cap_cbs].view(np.int32)
        self.h_hard = h[8 * cap_cbs:]
        self.descs = (CbDescriptor * cap_cbs)()
        self.done_event = torch.cuda.Event()
        self.busy = False
        self.requests = []
        self.region = None
        self.hard_off = np.zeros(cap_cbs, np.uint32)
        self.n = 0

    def close(self):
        if self.handle.value:
            _lib.load().ldpc_mi355_batch_destroy(self.handle)
            self.handle = ctypes.c_void_p()


class GpuExecutor:
    """Runs batches on ``n_contexts`` launch contexts (one HIP stream each)."""

    def __init__(self, pool: RequestPool, harq: HarqSlots, llr_pool: DevicePool,
                 row=DEFAULT_ROW, n_contexts=3, max_batch_codeblocks=256, device="cuda"):
        import torch
        self.torch = torch
        self.pool = pool
        self.harq = harq
        self.llr_pool = llr_pool
        self.row = tuple(int(v) for v in row)
        self.ctxs = [_Context(i, max_batch_codeblocks, self.row, device)
                     for i in range(n_contexts)]
        self.max_cbs = max_batch_codeblocks
        if pool.pinned is None:
            raise ValueError("the GPU executor needs the request LLRs in pinned memory")
        self.batches = 0

    def free_context(self):
        for c in self.ctxs:
            if not c.busy:
                return c
        return None

    def launch(self, ctx, reqs, region):
        """Copy the batch's LLRs into its pool region, decode, copy results
        back -- all queued on the context's stream."""
        L = _lib.load()
        off, need = region
        nreq = len(reqs)
        stream = ctx.stream
        sh = ctypes.c_void_p(stream.cuda_stream)
        # one host->device span per request, straight from its pinned LLRs
        # into the batch's pool region, all in one call
        spans = np.empty((3, nreq), np.uint64)
        parts, llr_base, slot_base = [], np.empty(nreq, np.int64), np.empty(nreq, np.int64)
        dev0, host0 = self.llr_pool.tensor.data_ptr(), self.pool.pinned.data_ptr()
        pos = 0
        for i, r in enumerate(reqs):
            a, b = self.pool.span(r.template, r.rv_index)
            spans[0, i] = dev0 + 2 * (off + pos)
            spans[1, i] = host0 + 2 * a
            spans[2, i] = 2 * (b - a)
            parts.append(self.pool.desc[(r.template, r.rv_index)])
            llr_base[i] = off + pos
            slot_base[i] = self.harq.slot(r.ue, r.proc, 0)
            pos += b - a
        assert pos <= need
        if L.ldpc_mi355_copy_spans(spans[0].ctypes.data, spans[1].ctypes.data,
                                   spans[2].ctypes.data, nreq, 1, sh) != 0:
            raise RuntimeError("ldpc_mi355_copy_spans (LLRs) failed")
        descs = np.concatenate(parts)
        counts = [len(d) for d in parts]
        descs["llr_offset"] += np.repeat(llr_base, counts).astype(descs["llr_offset"].dtype)
        descs["soft_slot"] += np.repeat(slot_base, counts).astype(descs["soft_slot"].dtype)
        n = len(descs)
        assert n <= ctx.cap
        ctx.desc_buf = descs
        bad = ctypes.c_int32(-1)
        st = L.ldpc_mi355_batch_decode(
            ctx.handle, descs.ctypes.data_as(ctypes.POINTER(CbDescriptor)), n,
            ptr(self.llr_pool.tensor), ptr(self.harq.tensor), ptr(ctx.hard), ptr(ctx.status),
            ptr(ctx.iters), None, sh, ctypes.byref(bad))
        if st != 0:
            raise RuntimeError(f"ldpc_mi355_batch_decode failed ({st}), codeblock {bad.value}")
        # codeblocks' hard bytes are laid out back to back in batch order:
        # status, iterations and those bytes come back as one span
        hb = int(sum(int(self.pool.cb_hard[r.template].sum()) for r in reqs))
        out = np.array([[ctx.out_host.data_ptr()], [ctx.out_dev.data_ptr()],
                        [8 * ctx.cap + hb]], np.uint64)
        if L.ldpc_mi355_copy_spans(out[0].ctypes.data, out[1].ctypes.data, out[2].ctypes.data,
                                   1, 2, sh) != 0:
            raise RuntimeError("ldpc_mi355_copy_spans (results) failed")
        ctx.done_event.record(stream)
        ctx.busy = True
        ctx.requests = reqs
        ctx.region = region
        ctx.n = n
        self.batches += 1

    def poll(self):
        """Contexts whose batch (and result copies) finished."""
        out = []
        for c in self.ctxs:
            if c.busy and c.done_event.query():
                out.append(c)
        return out

    def collect(self, ctx):
        """Per-request results of a finished context; frees it."""
        st, its, hard = ctx.h_status, ctx.h_iters, ctx.h_hard
        k = 0
        h = 0
        for r in ctx.requests:
            m = r.n_cbs
            hl = self.pool.cb_hard[r.template]
            nb = int(hl.sum())
            r.status = st[k:k + m].copy()
            r.iters = its[k:k + m].copy()
            r.hard_bytes = hard[h:h + nb].copy()
            r.hard_len = hl
            r._hard = None
            k += m
            h += nb
        reqs, region = ctx.requests, ctx.region
        ctx.busy = False
        ctx.requests = []
        ctx.region = None
        return reqs, region

    def close(self):
        for c in self.ctxs:
            c.close()


# --------------------------------------------------------------------------
# The loop
# --------------------------------------------------------------------------
@dataclass
class ServeStats:
    requests: List[Request] = field(default_factory=list)
    cb_info: list = field(default_factory=list)
    wall_s: float = 0.0
    batches: int = 0
    admission_waits: int = 0
    max_blocked_arrivals: int = 0
    max_inflight: int = 0
    host_launch_s: float = 0.0      # host time inside executor.launch
    host_collect_s: float = 0.0     # host time inside executor.collect
    peak_pool_bytes: int = 0

    def summary(self, deadline_s):
        done = [r for r in self.requests if r.done >= 0]
        lat = np.array([r.latency for r in done]) if done else np.array([np.nan])
        lat_a = np.array([r.latency_from_arrival for r in done]) if done else np.array([np.nan])
        # goodput: payload bits of the codeblocks that passed CRC on the last
        # transmission of their HARQ chain
        info = sum(int(self.cb_info[r.template][r.status == 0].sum())
                   for r in done if r.final)
        first = [r for r in done if r.rv_index == 0]
        return {
            "requests": len(done), "batches": self.batches,
            "latency_ms": {"p50": float(np.percentile(lat, 50) * 1e3),
                           "p99": float(np.percentile(lat, 99) * 1e3),
                           "mean": float(np.mean(lat) * 1e3)},
            "deadline_miss_rate": float(np.mean(lat > deadline_s)),
            "latency_from_arrival_ms": {"p50": float(np.percentile(lat_a, 50) * 1e3),
                                        "p99": float(np.percentile(lat_a, 99) * 1e3),
                                        "mean": float(np.mean(lat_a) * 1e3)},
            "deadline_miss_rate_from_arrival": float(np.mean(lat_a > deadline_s)),
            "info_gbps": info / self.wall_s / 1e9 if self.wall_s > 0 else 0.0,
            "codeblocks": int(sum(r.n_cbs for r in done)),
            "transmission_fail_rate": float(np.mean([not r.ok for r in done])) if done else 0.0,
            "codeblock_bler_first_tx": float(np.mean(np.concatenate(
                [r.status != 0 for r in first]))) if first else 0.0,
            "chains_failed": int(sum(1 for r in done if r.final and not r.ok)),
            "retransmissions": int(sum(1 for r in done if r.rv_index > 0)),
            "admission_waits": self.admission_waits,
            "max_blocked_arrivals": self.max_blocked_arrivals,
            "max_batches_in_flight": self.max_inflight,
            "host_us_per_batch": {
                "launch": 1e6 * self.host_launch_s / max(self.batches, 1),
                "collect": 1e6 * self.host_collect_s / max(self.batches, 1)},
            "peak_llr_pool_MB": self.peak_pool_bytes / 2 ** 20,
            "wall_s": self.wall_s,
        }


class Server:
    """Continuous-batching request loop over an executor (GpuExecutor)."""

    def __init__(self, executor, deadline_s=0.5e-3, harq_rtt_s=1.0e-3, n_ues=16, n_procs=4,
                 clock=time.perf_counter):
        self.ex = executor
        self.deadline_s = deadline_s
        self.harq_rtt_s = harq_rtt_s
        self.n_ues, self.n_procs = n_ues, n_procs
        self.clock = clock

    def run(self, arrivals, templates, ues=None, until_drained=True, max_wall_s=60.0):
        """arrivals: sorted new-data arrival instants (s, relative to start);
        templates: request template per arrival; ues: UE per arrival (default
        round robin).  Returns ServeStats."""
        pool = self.ex.pool
        stats = ServeStats(cb_info=pool.cb_info)
        proc_busy = np.zeros((self.n_ues, self.n_procs), bool)
        proc_freed = np.zeros((self.n_ues, self.n_procs))   # when each process was last freed
        pending = []        # (time, seq, Request) not yet arrived, heap-ordered by time
        import heapq
        seq = 0
        for i, (t, tpl) in enumerate(zip(arrivals, templates)):
            ue = int(ues[i]) if ues is not None else i % self.n_ues
            heapq.heappush(pending, (float(t), seq, ("new", ue, int(tpl))))
            seq += 1
        queue = []          # admitted-in-order FIFO of arrived Requests
        blocked = []        # new-data arrivals waiting for a free HARQ process of their UE
        rid = 0
        batches0 = self.ex.batches
        t0 = self.clock()
        inflight = 0

        def make(kind_tpl, ue, proc, rvi, t_arr, t_queued):
            nonlocal rid
            r = Request(rid=rid, template=kind_tpl, ue=ue, proc=proc, rv_index=rvi,
                        arrival=t_arr, deadline=t_queued + self.deadline_s, queued=t_queued,
                        n_cbs=len(pool.cfgs[kind_tpl]), info_bits=pool.info_bits[kind_tpl])
            rid += 1
            stats.requests.append(r)
            return r

        while True:
            now = self.clock() - t0
            # completions
            for ctx in self.ex.poll():
                tc = self.clock()
                reqs, region = self.ex.collect(ctx)
                stats.host_collect_s += self.clock() - tc
                self.ex.llr_pool.release(*region)
                inflight -= 1
                for r in reqs:
                    r.done = now
                    r.final = r.ok or r.rv_index + 1 == len(RV_SEQUENCE)
                    if not r.final:
                        heapq.heappush(pending, (now + self.harq_rtt_s, seq,
                                                 ("retx", r.ue, r.template, r.proc,
                                                  r.rv_index + 1)))
                        seq += 1
                    else:
                        proc_busy[r.ue, r.proc] = False
                        proc_freed[r.ue, r.proc] = now
            # arrivals
            while pending and pending[0][0] <= now:
                t_arr, _, ev = heapq.heappop(pending)
                if ev[0] == "new":
                    blocked.append((t_arr, ev[1], ev[2]))
                else:
                    _, ue, tpl, proc, rvi = ev
                    queue.append(make(tpl, ue, proc, rvi, t_arr, t_arr))
            still = []
            for t_arr, ue, tpl in blocked:
                free = np.flatnonzero(~proc_busy[ue])
                if len(free):
                    proc_busy[ue, free[0]] = True
                    queue.append(make(tpl, ue, int(free[0]), 0, t_arr,
                                      max(t_arr, float(proc_freed[ue, free[0]]))))
                else:
                    still.append((t_arr, ue, tpl))
            stats.max_blocked_arrivals = max(stats.max_blocked_arrivals, len(still))
            blocked = still
            # admission: continuous batching into free contexts
            while queue:
                ctx = self.ex.free_context()
                if ctx is None:
                    break
                batch, cbs, elems = [], 0, 0
                for r in queue:
                    e = pool.elems[r.template]
                    if cbs + r.n_cbs > self.ex.max_cbs:
                        break
                    batch.append(r)
                    cbs += r.n_cbs
                    elems += e
                region = self.ex.llr_pool.alloc(elems) if batch else None
                while batch and region is None:
                    # admission by free device memory: shrink the batch to fit
                    r = batch.pop()
                    elems -= pool.elems[r.template]
                    region = self.ex.llr_pool.alloc(elems) if batch else None
                if not batch:
                    stats.admission_waits += 1
                    if inflight == 0:
                        raise RuntimeError("request does not fit in the device LLR pool")
                    break
                del queue[:len(batch)]
                now = self.clock() - t0
                for r in batch:
                    r.admitted = now
                tl = self.clock()
                self.ex.launch(ctx, batch, region)
                stats.host_launch_s += self.clock() - tl
                inflight += 1
                stats.max_inflight = max(stats.max_inflight, inflight)
                stats.peak_pool_bytes = max(stats.peak_pool_bytes, 2 * self.ex.llr_pool.used)
            if not pending and not queue and not blocked and inflight == 0:
                break
            if not until_drained and not pending:
                break
            if now > max_wall_s:
                break
        stats.wall_s = self.clock() - t0
        stats.batches = self.ex.batches - batches0
        return stats


class NativeServer:
    """The same served loop run natively (ldpc_mi355_serve_run): request
    templates, pinned LLRs, the device LLR pool and HARQ slots as for
    GpuExecutor/Server, host work per batch in C++.  run() returns ServeStats
    with one Request per transmission, like Server.run."""

    def __init__(self, pool: RequestPool, harq: HarqSlots, llr_pool: DevicePool,
                 row=DEFAULT_ROW, n_contexts=4, max_batch_codeblocks=256, deadline_s=0.5e-3,
                 harq_rtt_s=1.0e-3, record_hard=False, llr_source="host_copy"):
        if pool.pinned is None:
            raise ValueError("the served loop needs the request LLRs in pinned memory")
        # llr_source "device": the request LLRs sit in HBM, as a demapper on
        # the GPU would leave them, and batches decode them in place;
        # "host_mapped": batches decode the pinned host LLRs in place, the
        # kernel reading them over PCIe; "host_copy": each batch copies them
        # from pinned host memory into its region of the device LLR pool (the
        # pool goes unused in place)
        if llr_source not in _lib.LLR_SOURCES:
            raise ValueError(f"llr_source: one of {sorted(_lib.LLR_SOURCES)}")
        self.llr_src = pool.pinned.cuda() if llr_source == "device" else pool.pinned
        self.llr_source = llr_source
        self.pool, self.harq, self.llr_pool = pool, harq, llr_pool
        self.row = tuple(int(v) for v in row)
        self.n_contexts = int(n_contexts)
        self.max_cbs = int(max_batch_codeblocks)
        self.deadline_s, self.harq_rtt_s = deadline_s, harq_rtt_s
        self.record_hard = record_hard
        n = len(pool)
        self.templates = (_lib.ServeTemplate * n)()
        cfgs = []
        for t in range(n):
            tp = self.templates[t]
            tp.n_cbs = pool.sizes[t]
            tp.cb_first = len(cfgs)
            a0, b0 = pool.span(t, 0)
            tp.llr_elems = b0 - a0
            for rvi in range(len(RV_SEQUENCE)):
                a, b = pool.span(t, rvi)
                assert b - a == b0 - a0
                tp.llr_host[rvi] = a
            cfgs.extend(pool.cfgs[t])
        self.cb_cfg = (LdpcInConfig * len(cfgs))(*[LdpcInConfig(*[int(v) for v in c])
                                                     for c in cfgs])
        self.batches = 0

    def run(self, arrivals, templates, ues=None, max_wall_s=60.0):
        L = _lib.load()
        pool = self.pool
        n_arr = len(arrivals)
        t_arr = np.ascontiguousarray(arrivals, np.float64)
        tpl = np.ascontiguousarray(templates, np.int32)
        ue = (np.arange(n_arr, dtype=np.int32) % self.harq.n_ues if ues is None
              else np.ascontiguousarray(ues, np.int32))
        max_rec = 4 * n_arr + 1
        recs = np.zeros(max_rec, _lib.SERVE_RECORD_DTYPE)
        max_cb = int(sum(pool.sizes[int(k)] for k in tpl)) * 4 + 1
        st = np.zeros(max_cb, np.int32)
        its = np.zeros(max_cb, np.int32)
        hard_cap = int(sum(pool.cb_hard[int(k)].sum() for k in tpl)) * 4 if self.record_hard else 0
        hard = np.zeros(max(hard_cap, 1), np.uint8)
        cfg = _lib.ServeCfg(n_contexts=self.n_contexts, max_batch_codeblocks=self.max_cbs,
                            n_ues=self.harq.n_ues, n_procs=self.harq.n_procs,
                            max_cbs=self.harq.max_cbs, llr_source=_lib.LLR_SOURCES[self.llr_source],
                            deadline_s=self.deadline_s,
                            harq_rtt_s=self.harq_rtt_s, max_wall_s=max_wall_s,
                            llr_pool_elems=self.llr_pool.elems)
        out = _lib.ServeStatsC()
        rc = L.ldpc_mi355_serve_run(
            ctypes.byref(cfg), ctypes.byref(DecoderConfig(*self.row)), self.templates, len(pool),
            self.cb_cfg, ptr(self.llr_src), ptr(self.llr_pool.tensor), ptr(self.harq.tensor),
            ptr(t_arr), ptr(tpl), ptr(ue), n_arr, ptr(recs), max_rec, ptr(st), ptr(its), max_cb,
            ptr(hard) if self.record_hard else None, hard_cap, ctypes.byref(out))
        if rc != 0:
            raise RuntimeError(f"ldpc_mi355_serve_run failed ({rc})")
        stats = ServeStats(cb_info=pool.cb_info)
        for i, r in enumerate(recs[:out.n_records]):
            t = int(r["template_id"])
            a, m = int(r["cb_first"]), int(r["n_cbs"])
            q = Request(rid=i, template=t, ue=int(r["ue"]), proc=int(r["proc"]),
                        rv_index=int(r["rv_index"]), arrival=float(r["arrival_s"]),
                        queued=float(r["queued_s"]),
                        deadline=float(r["queued_s"]) + self.deadline_s, n_cbs=m,
                        info_bits=pool.info_bits[t], admitted=float(r["admitted_s"]),
                        done=float(r["done_s"]), batch_id=int(r["batch"]))
            q.status = st[a:a + m].copy()
            q.iters = its[a:a + m].copy()
            q.final = bool(r["final_tx"])
            if self.record_hard and r["hard_first"] >= 0:
                h0 = int(r["hard_first"])
                q.hard_len = pool.cb_hard[t]
                q.hard_bytes = hard[h0:h0 + int(q.hard_len.sum())].copy()
            stats.requests.append(q)
        stats.wall_s = out.wall_s
        stats.batches = out.batches
        stats.admission_waits = out.admission_waits
        stats.max_blocked_arrivals = out.max_blocked
        stats.max_inflight = out.max_inflight
        stats.host_launch_s = out.host_launch_s
        stats.host_collect_s = out.host_collect_s
        stats.peak_pool_bytes = 2 * int(out.peak_pool_elems)
        self.batches += out.batches
        return stats


def poisson_arrivals(rate_per_s, duration_s, seed=0):
    rng = np.random.default_rng(seed)
    t, out = 0.0, []
    while True:
        t += rng.exponential(1.0 / rate_per_s)
        if t > duration_s:
            return np.array(out)
        out.append(t)
