// serve_sim.h -- discrete-event model of the group-by request server's host
// logic, for CPU tests: requests arrive at given times, `workers` batches run
// at once and share the device equally (each batch's work is rows x a
// per-table cost), every admission goes through plan_batch over the server's
// range allocator, and a caller consumes results strictly in submission
// order, releasing each result block as soon as it has it (the benchmark's
// replay collector).  The model reports whether every request finished or
// the run stopped with work left and nothing able to move (a stall).
//
// holds_by_head = true sizes the holds of passed-over requests with the
// head request's signature instead of their own (how the server sized them
// before each request carried its own scratch), so tests can show that this
// sizing stalls where per-request sizing does not.
#pragma once

#include "../visondf_amd/csrc/serve_core.h"

#include <cstdio>
#include <deque>
#include <string>
#include <vector>

namespace vdf_sim {

struct SimTable {
    uint64_t sig = 0;        // batch signature
    uint64_t rows = 0;
    uint64_t out_bytes = 0;  // result block bytes
    double cost_per_row = 0; // device seconds per row when the batch runs alone
};

struct SimRequest {
    double arrive = 0;  // seconds
    int table = 0;
};

struct SimConfig {
    uint64_t arena_bytes = 0;
    int workers = 1;
    vdf::PlanLimits lim;
    double batch_overhead_s = 0;  // fixed device time per batch
    bool holds_by_head = false;
};

struct SimResult {
    bool stalled = false;
    size_t consumed = 0;      // requests the in-order caller received
    size_t failed = 0;        // requests rejected as never fitting
    size_t batches = 0;
    double makespan_s = 0;
    uint64_t arena_peak = 0;
    size_t max_queue = 0;
    std::string state;        // at a stall: what was queued, running and held
};

// scratch(sig, rows of the batch's tables) -> bytes of the batch's scratch
template <typename ScratchFn>
SimResult simulate(const std::vector<SimTable>& tables, const std::vector<SimRequest>& reqs, const SimConfig& cfg,
                   ScratchFn scratch)
{
    using vdf::RangeAllocator;
    struct Running {
        double remaining;  // device seconds left at full share
        std::vector<size_t> reqs;
        uint64_t scr_off;
    };
    SimResult res;
    RangeAllocator arena(cfg.arena_bytes);
    const size_t R = reqs.size();
    std::vector<uint64_t> out_off(R, RangeAllocator::kNone);
    std::vector<char> done(R, 0);
    std::deque<size_t> queue;
    std::vector<Running> running;
    size_t next_arrival = 0, consumer = 0;
    double t = 0;
    for (;;) {
        while (next_arrival < R && reqs[next_arrival].arrive <= t) queue.push_back(next_arrival++);
        res.max_queue = std::max(res.max_queue, queue.size());
        // the in-order caller takes every finished result it is waiting for
        while (consumer < R && done[consumer]) {
            if (out_off[consumer] != RangeAllocator::kNone) arena.release(out_off[consumer]);
            out_off[consumer] = RangeAllocator::kNone;
            ++consumer;
        }
        res.consumed = consumer;
        if (consumer == R) break;
        // every free worker plans a batch
        while (static_cast<int>(running.size()) < cfg.workers && !queue.empty()) {
            const SimTable& ht = tables[reqs[queue.front()].table];
            std::vector<vdf::PlanItem> items;
            for (size_t q : queue) {
                const SimTable& tb = tables[reqs[q].table];
                const uint64_t own = scratch(tb.sig, std::vector<uint64_t>{tb.rows});
                const uint64_t by_head = scratch(ht.sig, std::vector<uint64_t>{tb.rows});
                items.push_back(vdf::PlanItem{tb.rows, tb.sig, tb.out_bytes, cfg.holds_by_head ? by_head : own});
            }
            vdf::Plan p = vdf::plan_batch(items, cfg.lim, arena, [&](const std::vector<uint64_t>& ns) {
                return scratch(ht.sig, ns);
            });
            if (p.head_never_fits) {
                done[queue.front()] = 1;
                ++res.failed;
                queue.pop_front();
                continue;
            }
            if (p.picked.empty()) break;
            Running b;
            b.remaining = cfg.batch_overhead_s;
            b.scr_off = p.scr_off;
            for (size_t k = p.picked.size(); k-- > 0;) {
                const size_t r = queue[p.picked[k]];
                out_off[r] = p.out_off[k];
                b.reqs.push_back(r);
                b.remaining += static_cast<double>(tables[reqs[r].table].rows) * tables[reqs[r].table].cost_per_row;
                queue.erase(queue.begin() + static_cast<std::ptrdiff_t>(p.picked[k]));
            }
            running.push_back(b);
            ++res.batches;
        }
        res.arena_peak = std::max(res.arena_peak, arena.peak());
        // next event: an arrival or the first batch to finish under equal sharing
        const double share = running.empty() ? 1.0 : 1.0 / static_cast<double>(running.size());
        double t_next = next_arrival < R ? reqs[next_arrival].arrive : 1e300;
        for (const Running& b : running) t_next = std::min(t_next, t + b.remaining / share);
        if (t_next >= 1e299) {
            res.stalled = true;
            char buf[256];
            std::snprintf(buf, sizeof buf,
                          "stall at t=%.4f s: waiting for request %zu, %zu queued, arena used %llu of %llu bytes, "
                          "largest free %llu",
                          t, consumer, queue.size(), static_cast<unsigned long long>(arena.used()),
                          static_cast<unsigned long long>(arena.size()),
                          static_cast<unsigned long long>(arena.largest_free()));
            res.state = buf;
            break;
        }
        const double dt = t_next - t;
        t = t_next;
        for (size_t i = running.size(); i-- > 0;) {
            running[i].remaining -= dt * share;
            if (running[i].remaining <= 1e-12) {
                arena.release(running[i].scr_off);
                for (size_t r : running[i].reqs) done[r] = 1;
                running.erase(running.begin() + static_cast<std::ptrdiff_t>(i));
            }
        }
    }
    res.makespan_s = t;
    return res;
}

}  // namespace vdf_sim
