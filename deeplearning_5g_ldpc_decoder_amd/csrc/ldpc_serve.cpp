// Native served decode loop (ldpc_mi355_serve_run): continuous batching of
// decode requests into the library's batch contexts, admission by free
// device memory, HARQ soft buffers resident on the device.
//
// Requests are transmissions of one HARQ process of one UE: 1..max_cbs
// codeblocks whose channel LLRs sit in pinned host memory or in device
// memory, one contiguous span per redundancy version.  The loop runs in real
// time on the calling thread:
//   * new-data arrivals wait for a free HARQ process of their UE, then join
//     a FIFO queue; retransmissions join it harq_rtt_s after the failed
//     transmission completed;
//   * whenever a context (one batch context and one HIP stream each) is
//     free, the queue head is admitted as one batch, as many whole requests
//     as fit max_batch_codeblocks and a first-fit region of the device LLR
//     pool: requests that do not fit wait until a running batch frees its
//     region;
//   * a batch's LLR spans are copied into its region (host LLRs) or read in
//     place (device LLRs, as a demapper on the GPU leaves them: no region,
//     no copy) and decoded (the decoder row of the contexts), the decode
//     writing status, iterations and hard bytes straight into the context's
//     pinned host buffer (no device-to-host copy), asynchronously on the
//     context's stream; an event per batch is polled for completion;
//   * a transmission whose codeblocks do not all pass CRC is retransmitted
//     with the next redundancy version of 0, 2, 3, 1 and soft-combined on
//     the device into the same slots (HARQ_on = 1); slot (ue, proc, cb) is
//     ((ue * n_procs) + proc) * max_cbs + cb of the soft-buffer array.
// Host work per batch is descriptor building, one copy per request (host
// LLRs only), the batch decode and an event record; nothing is allocated
// inside the loop.
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <chrono>
#include <deque>
#include <queue>
#include <vector>

#include "../../include/nr_ldpc_mi355.h"

namespace {

constexpr int kRv[4] = {0, 2, 3, 1};
constexpr int kHardMax = 8448 / 8;

// First-fit allocator over [0, elems) int16 elements, 128-element granules.
class FirstFit {
public:
    explicit FirstFit(uint64_t elems) : total_(elems / kAlign * kAlign) {
        free_.push_back({0, total_});
    }
    // offset or -1
    int64_t alloc(uint64_t elems, uint64_t *got) {
        const uint64_t need = (elems + kAlign - 1) / kAlign * kAlign;
        for (size_t i = 0; i < free_.size(); ++i) {
            if (free_[i].len >= need) {
                const uint64_t off = free_[i].off;
                if (free_[i].len == need)
                    free_.erase(free_.begin() + (long)i);
                else
                    free_[i] = {off + need, free_[i].len - need};
                used_ += need;
                peak_ = std::max(peak_, used_);
                *got = need;
                return (int64_t)off;
            }
        }
        return -1;
    }
    void release(uint64_t off, uint64_t len) {
        auto it = std::lower_bound(free_.begin(), free_.end(), off,
                                   [](const Span &s, uint64_t o) { return s.off < o; });
        it = free_.insert(it, {off, len});
        size_t i = (size_t)(it - free_.begin());
        if (i + 1 < free_.size() && free_[i].off + free_[i].len == free_[i + 1].off) {
            free_[i].len += free_[i + 1].len;
            free_.erase(free_.begin() + (long)i + 1);
        }
        if (i > 0 && free_[i - 1].off + free_[i - 1].len == free_[i].off) {
            free_[i - 1].len += free_[i].len;
            free_.erase(free_.begin() + (long)i);
        }
        used_ -= len;
    }
    uint64_t peak() const { return peak_; }

private:
    static constexpr uint64_t kAlign = 128;
    struct Span {
        uint64_t off, len;
    };
    uint64_t total_, used_ = 0, peak_ = 0;
    std::vector<Span> free_;
};

struct Event {
    double t;
    uint64_t seq;
    bool retx;
    int ue, tpl, proc, rvi;
    bool operator>(const Event &o) const { return t != o.t ? t > o.t : seq > o.seq; }
};

struct Context {
    ldpc_mi355_batch_t *batch = nullptr;
    hipStream_t stream = nullptr;
    hipEvent_t done = nullptr;
    // status (n int32), iterations (n int32), hard bytes of the running
    // batch of n codeblocks, sized for cap codeblocks: pinned host memory
    // that the decode writes through its device address (d_out)
    unsigned char *h_out = nullptr;
    unsigned char *d_out = nullptr;
    int n = 0;
    std::vector<ldpc_mi355_cb_t> descs;
    std::vector<int> records;         // records of the running batch
    uint64_t region_off = 0, region_len = 0;
    bool busy = false;
};

double seconds_since(std::chrono::steady_clock::time_point t0) {
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

}  // namespace

extern "C" int32_t ldpc_mi355_serve_run(
    const ldpc_mi355_serve_cfg_t *cfg, const ldpc_decoder_cfg_t *decoder_config,
    const ldpc_mi355_serve_template_t *templates, int32_t n_templates,
    const ldpc_in_config_t *cb_cfg, const int16_t *llr_host, int16_t *llr_pool,
    int16_t *soft_pool, const double *arrival_s, const int32_t *arrival_template,
    const int32_t *arrival_ue, int32_t n_arrivals, ldpc_mi355_serve_record_t *records,
    int32_t max_records, int32_t *cb_status, int32_t *cb_iterations, int64_t max_cb_results,
    uint8_t *hard_out, int64_t hard_capacity, ldpc_mi355_serve_stats_t *stats) {
    const bool in_place = cfg && cfg->llr_source != LDPC_MI355_LLR_HOST_COPY;
    if (!cfg || !decoder_config || !templates || n_templates <= 0 || !cb_cfg || !llr_host ||
        (!in_place && (!llr_pool || cfg->llr_pool_elems == 0)) || !soft_pool ||
        n_arrivals < 0 || (n_arrivals && (!arrival_s || !arrival_template || !arrival_ue)) ||
        !records || !cb_status || !cb_iterations || !stats || cfg->n_contexts <= 0 ||
        cfg->max_batch_codeblocks <= 0 || cfg->n_ues <= 0 || cfg->n_procs <= 0 ||
        cfg->max_cbs <= 0 || cfg->llr_source < LDPC_MI355_LLR_HOST_COPY ||
        cfg->llr_source > LDPC_MI355_LLR_HOST_MAPPED)
        return -4;
    for (int32_t i = 0; i < n_templates; ++i) {
        if (templates[i].n_cbs <= 0 || templates[i].n_cbs > cfg->max_cbs ||
            templates[i].n_cbs > cfg->max_batch_codeblocks)
            return -4;
        // read in place, a codeblock's LLRs are addressed by a 32-bit element
        // offset from llr_host
        if (in_place)
            for (int r = 0; r < 4; ++r)
                if (templates[i].llr_host[r] + templates[i].llr_elems > UINT32_MAX) return -4;
    }
    for (int32_t i = 0; i < n_arrivals; ++i)
        if (arrival_template[i] < 0 || arrival_template[i] >= n_templates || arrival_ue[i] < 0 ||
            arrival_ue[i] >= cfg->n_ues)
            return -4;
    memset(stats, 0, sizeof *stats);

    // set-up: contexts, streams, events, result spans
    const int cap = cfg->max_batch_codeblocks;
    const size_t out_bytes = (size_t)8 * cap + (size_t)cap * kHardMax;
    std::vector<Context> ctx((size_t)cfg->n_contexts);
    int32_t rc = 0;
    for (auto &c : ctx) {
        if (ldpc_mi355_batch_create(&c.batch, cap, decoder_config, 0) != 0 ||
            hipStreamCreateWithFlags(&c.stream, hipStreamNonBlocking) != hipSuccess ||
            hipEventCreateWithFlags(&c.done, hipEventDisableTiming) != hipSuccess ||
            hipHostMalloc(&c.h_out, out_bytes, hipHostMallocDefault) != hipSuccess ||
            hipHostGetDevicePointer(reinterpret_cast<void **>(&c.d_out), c.h_out, 0) !=
                hipSuccess) {
            rc = -1;
            break;
        }
        c.descs.resize((size_t)cap);
    }
    // per template: codeblock LLR offsets within one transmission, hard bytes
    std::vector<std::vector<uint32_t>> cb_rel((size_t)n_templates);
    std::vector<int> tpl_hard((size_t)n_templates, 0);
    for (int32_t t = 0; t < n_templates && rc == 0; ++t) {
        uint32_t pos = 0;
        for (int c = 0; c < templates[t].n_cbs; ++c) {
            const ldpc_in_config_t &k = cb_cfg[templates[t].cb_first + c];
            cb_rel[(size_t)t].push_back(pos);
            pos += k.E;
            tpl_hard[(size_t)t] += (k.A + (k.crctype == 0 ? 16 : 24)) / 8;
        }
        if (pos != templates[t].llr_elems) rc = -4;
    }

    // LLRs from pinned host memory, copied over PCIe into a region of
    // llr_pool, or read where they are: in device memory, or in pinned host
    // memory through its device address
    const int16_t *llr_base = in_place ? llr_host : llr_pool;
    if (rc == 0 && cfg->llr_source == LDPC_MI355_LLR_HOST_MAPPED) {
        void *mapped = nullptr;
        if (hipHostGetDevicePointer(&mapped, const_cast<int16_t *>(llr_host), 0) != hipSuccess ||
            !mapped) {
            fprintf(stderr, "libnrldpc_mi355: llr_host is not pinned host memory mapped for the "
                            "device\n");
            rc = -4;
        }
        llr_base = static_cast<const int16_t *>(mapped);
    }
    FirstFit pool(in_place ? 0 : cfg->llr_pool_elems);
    std::priority_queue<Event, std::vector<Event>, std::greater<Event>> pending;
    uint64_t seq = 0;
    for (int32_t i = 0; i < n_arrivals; ++i)
        pending.push({arrival_s[i], seq++, false, arrival_ue[i], arrival_template[i], 0, 0});
    std::deque<int> queue;                 // admitted-in-order records
    std::deque<Event> blocked;             // new data waiting for a free HARQ process
    std::vector<char> proc_busy((size_t)cfg->n_ues * cfg->n_procs, 0);
    // when each HARQ process was last freed (a blocked arrival joins the
    // decode queue then)
    std::vector<double> proc_freed((size_t)cfg->n_ues * cfg->n_procs, 0.0);
    int32_t n_rec = 0;
    int64_t n_cb = 0, n_hard = 0;
    int inflight = 0;
    const auto t0 = std::chrono::steady_clock::now();

    auto new_record = [&](const Event &e, int proc, double t_arr, double t_queued) -> int {
        if (n_rec >= max_records) return -1;
        const int32_t nc = templates[e.tpl].n_cbs;
        if (n_cb + nc > max_cb_results) return -1;
        ldpc_mi355_serve_record_t &r = records[n_rec];
        memset(&r, 0, sizeof r);
        r.template_id = e.tpl;
        r.ue = e.ue;
        r.proc = proc;
        r.rv_index = e.rvi;
        r.n_cbs = nc;
        r.cb_first = n_cb;
        r.hard_first = -1;
        r.batch = -1;
        r.arrival_s = t_arr;
        r.queued_s = t_queued;
        r.admitted_s = -1.0;
        r.done_s = -1.0;
        n_cb += nc;
        return n_rec++;
    };

    while (rc == 0) {
        double now = seconds_since(t0);
        // completions
        for (auto &c : ctx) {
            if (!c.busy) continue;
            const hipError_t q = hipEventQuery(c.done);
            if (q == hipErrorNotReady) continue;
            if (q != hipSuccess) {
                rc = -1;
                break;
            }
            const auto tc = std::chrono::steady_clock::now();
            const int32_t *st = reinterpret_cast<const int32_t *>(c.h_out);
            const int32_t *it = st + c.n;
            const uint8_t *hb = c.h_out + (size_t)8 * c.n;
            int k = 0;
            size_t h = 0;
            for (int ri : c.records) {
                ldpc_mi355_serve_record_t &r = records[ri];
                memcpy(cb_status + r.cb_first, st + k, sizeof(int32_t) * (size_t)r.n_cbs);
                memcpy(cb_iterations + r.cb_first, it + k, sizeof(int32_t) * (size_t)r.n_cbs);
                const int nb = tpl_hard[(size_t)r.template_id];
                if (hard_out && n_hard + nb <= hard_capacity) {
                    memcpy(hard_out + n_hard, hb + h, (size_t)nb);
                    r.hard_first = n_hard;
                    n_hard += nb;
                }
                int pass = 0;
                for (int j = 0; j < r.n_cbs; ++j) pass += st[k + j] == 0;
                r.n_pass = pass;
                r.done_s = now;
                r.final_tx = (pass == r.n_cbs || r.rv_index + 1 == 4) ? 1 : 0;
                if (!r.final_tx)
                    pending.push({now + cfg->harq_rtt_s, seq++, true, r.ue, r.template_id, r.proc,
                                  r.rv_index + 1});
                else {
                    proc_busy[(size_t)r.ue * cfg->n_procs + r.proc] = 0;
                    proc_freed[(size_t)r.ue * cfg->n_procs + r.proc] = now;
                }
                k += r.n_cbs;
                h += (size_t)nb;
            }
            if (c.region_len) pool.release(c.region_off, c.region_len);
            c.busy = false;
            c.records.clear();
            --inflight;
            stats->host_collect_s += seconds_since(tc);
        }
        if (rc) break;
        // arrivals and retransmissions that are due
        while (!pending.empty() && pending.top().t <= now) {
            const Event e = pending.top();
            pending.pop();
            if (!e.retx) {
                blocked.push_back(e);
            } else {
                const int ri = new_record(e, e.proc, e.t, e.t);
                if (ri < 0) {
                    rc = -4;
                    break;
                }
                queue.push_back(ri);
            }
        }
        if (rc) break;
        for (size_t i = 0; i < blocked.size();) {
            const Event &e = blocked[i];
            int proc = -1;
            for (int p = 0; p < cfg->n_procs; ++p)
                if (!proc_busy[(size_t)e.ue * cfg->n_procs + p]) {
                    proc = p;
                    break;
                }
            if (proc < 0) {
                ++i;
                continue;
            }
            proc_busy[(size_t)e.ue * cfg->n_procs + proc] = 1;
            const double t_queued =
                std::max(e.t, proc_freed[(size_t)e.ue * cfg->n_procs + proc]);
            const int ri = new_record(e, proc, e.t, t_queued);
            if (ri < 0) {
                rc = -4;
                break;
            }
            queue.push_back(ri);
            blocked.erase(blocked.begin() + (long)i);
        }
        if (rc) break;
        stats->max_blocked = std::max(stats->max_blocked, (int32_t)blocked.size());
        // admission: continuous batching into free contexts
        while (!queue.empty()) {
            Context *c = nullptr;
            for (auto &x : ctx)
                if (!x.busy) {
                    c = &x;
                    break;
                }
            if (!c) break;
            size_t take = 0;
            int cbs = 0;
            uint64_t elems = 0;
            while (take < queue.size()) {
                const ldpc_mi355_serve_template_t &tp = templates[records[queue[take]].template_id];
                if (cbs + tp.n_cbs > cap) break;
                cbs += tp.n_cbs;
                elems += tp.llr_elems;
                ++take;
            }
            uint64_t got = 0;
            int64_t off = in_place ? 0 : take ? pool.alloc(elems, &got) : -1;
            while (take && off < 0) {   // admission by free device memory
                --take;
                elems -= templates[records[queue[take]].template_id].llr_elems;
                off = take ? pool.alloc(elems, &got) : -1;
            }
            if (!take) {
                ++stats->admission_waits;
                if (inflight == 0) rc = -5;   // a request larger than the whole pool
                break;
            }
            const auto tl = std::chrono::steady_clock::now();
            now = seconds_since(t0);
            // LLR spans into the region, then descriptors
            uint64_t pos = (uint64_t)off;
            int n = 0;
            for (size_t q = 0; q < take && rc == 0; ++q) {
                ldpc_mi355_serve_record_t &r = records[queue[q]];
                const ldpc_mi355_serve_template_t &tp = templates[r.template_id];
                r.admitted_s = now;
                r.batch = stats->batches;
                uint64_t span = pos;
                if (in_place) {
                    span = tp.llr_host[r.rv_index];
                } else if (hipMemcpyAsync(llr_pool + pos, llr_host + tp.llr_host[r.rv_index],
                                          tp.llr_elems * sizeof(int16_t), hipMemcpyHostToDevice,
                                          c->stream) != hipSuccess) {
                    rc = -1;
                    break;
                }
                const int slot0 = (r.ue * cfg->n_procs + r.proc) * cfg->max_cbs;
                for (int j = 0; j < tp.n_cbs; ++j) {
                    ldpc_mi355_cb_t &d = c->descs[(size_t)n++];
                    d.cfg = cb_cfg[tp.cb_first + j];
                    d.cfg.rv = (uint16_t)kRv[r.rv_index];
                    d.cfg.HARQ_on = r.rv_index > 0 ? 1 : 0;
                    d.reserved = 0;
                    d.llr_offset = (uint32_t)(span + cb_rel[(size_t)r.template_id][(size_t)j]);
                    d.soft_slot = slot0 + j;
                }
                pos += tp.llr_elems;
                c->records.push_back(queue[q]);
            }
            if (rc) break;
            int32_t *d_status = reinterpret_cast<int32_t *>(c->d_out);
            int32_t bad = -1;
            if (ldpc_mi355_batch_decode(c->batch, c->descs.data(), n, llr_base, soft_pool,
                                        c->d_out + (size_t)8 * n, d_status, d_status + n,
                                        nullptr, c->stream, &bad) != 0) {
                rc = -1;
                break;
            }
            if (hipEventRecord(c->done, c->stream) != hipSuccess) {
                rc = -1;
                break;
            }
            queue.erase(queue.begin(), queue.begin() + (long)take);
            c->busy = true;
            c->n = n;
            c->region_off = (uint64_t)off;
            c->region_len = got;
            ++inflight;
            ++stats->batches;
            stats->max_inflight = std::max(stats->max_inflight, (int32_t)inflight);
            stats->host_launch_s += seconds_since(tl);
        }
        if (rc) break;
        if (pending.empty() && queue.empty() && blocked.empty() && inflight == 0) break;
        if (seconds_since(t0) > cfg->max_wall_s) break;
    }
    // drain whatever is still in flight before tearing down
    for (auto &c : ctx)
        if (c.stream) (void)hipStreamSynchronize(c.stream);
    stats->wall_s = seconds_since(t0);
    stats->n_records = n_rec;
    stats->n_cb_results = n_cb;
    stats->hard_bytes = n_hard;
    stats->peak_pool_elems = pool.peak();
    for (auto &c : ctx) {
        if (c.batch) ldpc_mi355_batch_destroy(c.batch);
        if (c.done) (void)hipEventDestroy(c.done);
        if (c.stream) (void)hipStreamDestroy(c.stream);
        if (c.h_out) (void)hipHostFree(c.h_out);
    }
    if (rc == -5)
        fprintf(stderr, "libnrldpc_mi355: a request does not fit in the device LLR pool\n");
    return rc;
}
