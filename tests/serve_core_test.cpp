// Unit test of the request server's host logic (visondf_amd/csrc/serve_core.h):
// the arena range allocator against a byte-occupancy model, and the batch
// planner's admission rules.  Built with g++ by tests/test_serve_core.py.
#include "serve_sim.h"

#include <cstdio>
#include <cstdlib>
#include <random>
#include <vector>

using namespace vdf;

static int g_fail = 0;
#define CHECK(c)                                                            \
    do {                                                                    \
        if (!(c)) {                                                         \
            std::printf("FAIL %s:%d: %s\n", __FILE__, __LINE__, #c);        \
            ++g_fail;                                                       \
        }                                                                   \
    } while (0)

static void allocator_random(uint64_t seed)
{
    const uint64_t A = 256, SIZE = 1 << 20;
    RangeAllocator ra(SIZE, A);
    std::vector<int> owner(SIZE / A, -1);  // model: which live range holds each 256-byte unit
    struct Live {
        uint64_t off, len;
    };
    std::vector<Live> live;
    std::mt19937_64 rng(seed);
    for (int step = 0; step < 20000; ++step) {
        const bool do_alloc = live.empty() || (rng() % 100) < 55;
        if (do_alloc) {
            const uint64_t bytes = 1 + rng() % (SIZE / 16);
            const uint64_t off = ra.alloc(bytes);
            const uint64_t len = (bytes + A - 1) / A * A;
            // model: first fit over free units
            uint64_t want = RangeAllocator::kNone;
            for (uint64_t u = 0, run = 0; u < SIZE / A; ++u) {
                run = owner[u] < 0 ? run + 1 : 0;
                if (run * A >= len) {
                    want = (u + 1) * A - len;
                    break;
                }
            }
            CHECK(off == want);
            if (off == RangeAllocator::kNone) continue;
            CHECK(off % A == 0 && off + len <= SIZE);
            for (uint64_t u = off / A; u < (off + len) / A; ++u) {
                CHECK(owner[u] < 0);
                owner[u] = step;
            }
            live.push_back({off, len});
        } else {
            const size_t k = rng() % live.size();
            CHECK(ra.release(live[k].off));
            for (uint64_t u = live[k].off / A; u < (live[k].off + live[k].len) / A; ++u) owner[u] = -1;
            live.erase(live.begin() + static_cast<long>(k));
        }
        uint64_t used = 0;
        for (const Live& l : live) used += l.len;
        CHECK(ra.used() == used);
        CHECK(ra.live_ranges() == live.size());
    }
    CHECK(!ra.release(12345 * A + 7));  // not a live range start
    for (const Live& l : live) CHECK(ra.release(l.off));
    CHECK(ra.used() == 0);
    CHECK(ra.free_ranges() == 1);  // fully merged
    CHECK(ra.largest_free() == SIZE);
}

// Queue items whose alone_scratch is what `scr` gives for the request alone.
template <typename F>
static std::vector<PlanItem> with_scratch(std::vector<PlanItem> q, F scr)
{
    for (PlanItem& it : q) it.alone_scratch = scr(std::vector<uint64_t>{it.rows});
    return q;
}

static uint64_t scratch_of(const std::vector<uint64_t>& ns)
{
    uint64_t s = 1000;
    for (uint64_t n : ns) s += 10 * n;
    return s;
}

static void planner_rules()
{
    const uint64_t A = 256;
    // FIFO head, signature filter, table limit
    {
        RangeAllocator ra(1 << 24, A);
        std::vector<PlanItem> q = with_scratch({{100, 1, 1000}, {200, 2, 1000}, {300, 1, 1000}, {400, 1, 1000}, {500, 1, 1000}},
                                                 scratch_of);
        PlanLimits lim;
        lim.max_tables = 3;
        Plan p = plan_batch(q, lim, ra, scratch_of);
        CHECK(p.picked.size() == 3);
        CHECK(p.picked[0] == 0 && p.picked[1] == 2 && p.picked[2] == 3);
        CHECK(p.scr_bytes == scratch_of({100, 300, 400}));
        CHECK(p.scr_off != RangeAllocator::kNone);
        CHECK(ra.live_ranges() == 4);
    }
    // row limit skips large later requests but always takes the head
    {
        RangeAllocator ra(1 << 24, A);
        std::vector<PlanItem> q = with_scratch({{5000, 7, 10}, {100, 7, 10}, {9000, 7, 10}, {200, 7, 10}}, scratch_of);
        PlanLimits lim;
        lim.max_rows = 5400;
        Plan p = plan_batch(q, lim, ra, scratch_of);
        CHECK(p.picked.size() == 3);
        CHECK(p.picked[0] == 0 && p.picked[1] == 1 && p.picked[2] == 3);
    }
    // memory: the scan stops at the first request whose result block does not fit
    {
        RangeAllocator ra(64 * A, A);
        auto scr = [](const std::vector<uint64_t>&) { return uint64_t(4 * 256); };
        std::vector<PlanItem> q = with_scratch({{1, 0, 8 * A}, {1, 0, 60 * A}, {1, 0, 2 * A}}, scr);
        PlanLimits lim;
        Plan p = plan_batch(q, lim, ra, scr);
        CHECK(p.picked.size() == 1 && p.picked[0] == 0);
        CHECK(ra.used() == 12 * A);
    }
    // scratch does not fit with every pick: newest picks are dropped until it does
    {
        RangeAllocator ra(40 * A, A);
        auto scr = [](const std::vector<uint64_t>& ns) { return uint64_t(ns.size() * 5 * 256); };
        std::vector<PlanItem> q = with_scratch({{10, 0, 10 * A}, {10, 0, 10 * A}, {10, 0, 10 * A}}, scr);
        PlanLimits lim;
        Plan p = plan_batch(q, lim, ra, scr);
        CHECK(p.picked.size() == 2);  // 3 blocks = 30 units + 15 scratch > 40; 2 blocks + 10 = 30
        CHECK(ra.used() == 30 * A);
        CHECK(p.scr_bytes == 10 * A);
    }
    // nothing fits now (memory held elsewhere): empty plan, not head_never_fits
    {
        RangeAllocator ra(40 * A, A);
        const uint64_t held = ra.alloc(35 * A);
        std::vector<PlanItem> q = with_scratch({{10, 0, 10 * A}}, scratch_of);
        Plan p = plan_batch(q, PlanLimits{}, ra, scratch_of);
        CHECK(p.picked.empty() && !p.head_never_fits);
        CHECK(ra.used() == 35 * A);
        ra.release(held);
        Plan p2 = plan_batch(q, PlanLimits{}, ra, scratch_of);
        CHECK(p2.picked.size() == 1);
    }
    // the head alone exceeds the arena
    {
        RangeAllocator ra(40 * A, A);
        auto scr = [](const std::vector<uint64_t>&) { return uint64_t(2 * 256); };
        std::vector<PlanItem> q = with_scratch({{10, 0, 39 * A}, {1, 0, A}}, scr);
        Plan p = plan_batch(q, PlanLimits{}, ra, scr);
        CHECK(p.head_never_fits && p.picked.empty());
        CHECK(ra.used() == 0);
    }
}

// A caller that consumes results in submission order: newer same-signature
// requests must not take the memory an older request of another signature
// needs, or the older one could never be admitted.
static void planner_keeps_room_for_skipped()
{
    const uint64_t A = 256;
    auto scr = [](const std::vector<uint64_t>& ns) { return uint64_t(ns.size() * 2 * 256); };
    // arena 40 units; queue: head (sig 0, 6 units), older skipped (sig 1, 20 units), newer sig 0 requests
    RangeAllocator ra(40 * A, A);
    std::vector<PlanItem> q =
        with_scratch({{1, 0, 6 * A}, {1, 1, 20 * A}, {1, 0, 6 * A}, {1, 0, 6 * A}, {1, 0, 6 * A}}, scr);
    Plan p = plan_batch(q, PlanLimits{}, ra, scr);
    // head (6) + scratch; the skipped request needs 20 + 2 = 22 contiguous units
    CHECK(!p.picked.empty() && p.picked[0] == 0);
    CHECK(ra.largest_free() >= 22 * A);
    for (size_t k : p.picked) CHECK(k != 1);
    // the in-order consumer: results of the batch stay held; the skipped request is now the head
    std::vector<PlanItem> q2(q.begin() + 1, q.end());
    for (size_t k = p.picked.size(); k-- > 1;) q2.erase(q2.begin() + static_cast<long>(p.picked[k] - 1));
    ra.release(p.scr_off);  // the batch finished; its results (out blocks) are still held
    Plan p2 = plan_batch(q2, PlanLimits{}, ra, scr);
    CHECK(!p2.picked.empty() && p2.picked[0] == 0);  // the older request runs
}

// Requests passed for the row limit, and several passed requests at once,
// keep their room too: after the batch, with its results still held, each
// passed request still fits when it becomes the oldest.
static void planner_keeps_room_for_row_limit_and_several()
{
    const uint64_t A = 256;
    auto scr = [](const std::vector<uint64_t>& ns) { return uint64_t(ns.size() * 2 * 256); };
    // row limit: the head (100 rows) and a 900-row request over max_rows 500,
    // then small newer requests of the same signature
    {
        RangeAllocator ra(40 * A, A);
        std::vector<PlanItem> q = with_scratch(
            {{100, 0, 6 * A}, {900, 0, 20 * A}, {50, 0, 6 * A}, {50, 0, 6 * A}, {50, 0, 6 * A}}, scr);
        PlanLimits lim;
        lim.max_rows = 500;
        Plan p = plan_batch(q, lim, ra, scr);
        CHECK(!p.picked.empty() && p.picked[0] == 0);
        for (size_t k : p.picked) CHECK(k != 1);
        CHECK(ra.largest_free() >= 22 * A);
        ra.release(p.scr_off);  // finished; results held
        Plan p2 = plan_batch({q[1]}, lim, ra, scr);
        CHECK(p2.picked.size() == 1);
    }
    // two passed requests of another signature need room side by side
    {
        RangeAllocator ra(50 * A, A);
        std::vector<PlanItem> q = with_scratch({{1, 0, 4 * A}, {1, 1, 15 * A}, {1, 1, 15 * A}, {1, 0, 4 * A},
                                                {1, 0, 4 * A}, {1, 0, 4 * A}, {1, 0, 4 * A}},
                                               scr);
        Plan p = plan_batch(q, PlanLimits{}, ra, scr);
        CHECK(!p.picked.empty() && p.picked[0] == 0);
        ra.release(p.scr_off);  // finished; results held
        // both older sig-1 requests run (one batch of two, or one after the other
        // with the first one's results held)
        Plan p2 = plan_batch({q[1], q[2]}, PlanLimits{}, ra, scr);
        CHECK(!p2.picked.empty());
        if (p2.picked.size() == 1) {
            ra.release(p2.scr_off);
            Plan p3 = plan_batch({q[2]}, PlanLimits{}, ra, scr);
            CHECK(p3.picked.size() == 1);
        }
    }
    // the peak counts admitted memory only, not the holds of planning
    {
        RangeAllocator ra(40 * A, A);
        std::vector<PlanItem> q = with_scratch({{1, 0, 4 * A}, {1, 1, 20 * A}}, scr);
        Plan p = plan_batch(q, PlanLimits{}, ra, scr);
        CHECK(p.picked.size() == 1);
        CHECK(ra.peak() == 6 * A);
    }
}

// Two signatures whose scratch per row differs (8-byte keys: 2 units per
// table of 10 rows; 16-byte keys: 40 units), a nearly full arena and a caller
// that consumes results in submission order.  A 16-byte-key request passed
// over behind an 8-byte-key head must be held at its own size: held at the
// head's size (3 units instead of 41), the batch admits eight newer
// 8-byte-key requests whose results (70 units, held by the caller, who waits
// for the 16-byte-key request first) leave 30 units, and the 16-byte-key
// request never fits.
static void holds_sized_by_own_signature()
{
    using namespace vdf_sim;
    const uint64_t A = 256;
    std::vector<SimTable> tables(2);
    tables[0] = SimTable{0, 10, 10 * A, 1e-3};  // 8-byte keys: large results, small scratch
    tables[1] = SimTable{1, 10, 1 * A, 1e-3};   // 16-byte keys: small results, large scratch
    auto scr = [&](uint64_t sig, const std::vector<uint64_t>& ns) {
        return uint64_t(ns.size() * (sig == 0 ? 2 : 40) * A);
    };
    std::vector<SimRequest> reqs;
    reqs.push_back({0.0, 0});
    reqs.push_back({0.0, 1});
    for (int i = 0; i < 11; ++i) reqs.push_back({0.0, 0});
    SimConfig cfg;
    cfg.arena_bytes = 100 * A;
    cfg.workers = 1;
    cfg.holds_by_head = true;
    SimResult head = simulate(tables, reqs, cfg, scr);
    CHECK(head.stalled);  // the sizing the server used before
    CHECK(head.consumed == 1);
    cfg.holds_by_head = false;
    SimResult own = simulate(tables, reqs, cfg, scr);
    CHECK(!own.stalled);
    CHECK(own.consumed == reqs.size());
    // the same with three workers and staggered arrivals
    cfg.workers = 3;
    for (size_t i = 0; i < reqs.size(); ++i) reqs[i].arrive = 1e-4 * static_cast<double>(i);
    SimResult own3 = simulate(tables, reqs, cfg, scr);
    CHECK(!own3.stalled && own3.consumed == reqs.size());
}

// Requests skipped after the batch's last pick are passed by nothing: they get
// no hold, so holds never take the head's own room.  Arena 40 units; the head
// needs 2 + 20; ten younger requests of another signature would each hold
// 5 + 5 units.
static void head_not_blocked_by_younger_holds()
{
    const uint64_t A = 256;
    auto scr0 = [](const std::vector<uint64_t>& ns) { return uint64_t(ns.size() * 20 * 256); };
    RangeAllocator ra(40 * A, A);
    std::vector<PlanItem> q;
    q.push_back(PlanItem{1, 0, 2 * A, 20 * A});
    for (int i = 0; i < 10; ++i) q.push_back(PlanItem{1, 1, 5 * A, 5 * A});
    Plan p = plan_batch(q, PlanLimits{}, ra, scr0);
    CHECK(p.picked.size() == 1 && p.picked[0] == 0);
    CHECK(ra.used() == 22 * A);  // the head's block and scratch; no hold outlives planning
    // a younger pick of the head's signature behind the others: they are passed
    // now, and the pick is taken only beside their holds (here: no room)
    RangeAllocator rb(40 * A, A);
    q.push_back(PlanItem{1, 0, 2 * A, 20 * A});
    Plan p2 = plan_batch(q, PlanLimits{}, rb, scr0);
    CHECK(p2.picked.size() == 1 && p2.picked[0] == 0);
    CHECK(rb.used() == 22 * A);
}

// Random traces of three signatures with different per-row scratch, several
// workers and an in-order caller: per-request holds never stall.
static void random_traces_make_progress()
{
    using namespace vdf_sim;
    const uint64_t A = 256;
    for (uint64_t seed = 1; seed <= 40; ++seed) {
        std::mt19937_64 rng(seed);
        std::vector<SimTable> tables;
        for (int i = 0; i < 24; ++i) {
            const uint64_t rows = 1 + rng() % 400;
            const uint64_t sig = rng() % 3;
            tables.push_back(SimTable{sig, rows, (1 + rows / 8) * A, 1e-6 * static_cast<double>(1 + sig)});
        }
        auto scr = [&](uint64_t sig, const std::vector<uint64_t>& ns) {
            uint64_t b = 4 * A;
            for (uint64_t n : ns) b += n * (1 + 2 * sig) * A / 4;
            return b;
        };
        std::vector<SimRequest> reqs;
        double t = 0;
        for (int i = 0; i < 300; ++i) {
            t += 1e-5 * static_cast<double>(rng() % 100);
            reqs.push_back({t, static_cast<int>(rng() % tables.size())});
        }
        SimConfig cfg;
        cfg.arena_bytes = (600 + rng() % 600) * A;
        cfg.workers = 1 + static_cast<int>(rng() % 4);
        cfg.lim.max_tables = 1 + rng() % 16;
        cfg.lim.max_rows = 200 + rng() % 1000;
        SimResult r = simulate(tables, reqs, cfg, scr);
        CHECK(!r.stalled);
        if (r.stalled) std::printf("seed %llu: %s\n", static_cast<unsigned long long>(seed), r.state.c_str());
        CHECK(r.consumed + 0 == reqs.size());
    }
}

int main()
{
    for (uint64_t s = 1; s <= 4; ++s) allocator_random(s);
    planner_rules();
    planner_keeps_room_for_skipped();
    planner_keeps_room_for_row_limit_and_several();
    holds_sized_by_own_signature();
    head_not_blocked_by_younger_holds();
    random_traces_make_progress();
    if (g_fail) {
        std::printf("%d failures\n", g_fail);
        return 1;
    }
    std::printf("serve_core ok\n");
    return 0;
}
