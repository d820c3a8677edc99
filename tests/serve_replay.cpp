// Replays a served-benchmark trace through the server's planner on the CPU
// (tests/serve_sim.h), with the library's own result-block layout and batch
// scratch sizes (vdf_groupby_batch_scratch_bytes of libvdf_hip.so, host
// arithmetic only), and prints one line: whether every request finished or
// the run stalled, with counts, makespan and arena peak.
//
// Input (text, from tests/test_serve_core.py):
//   arena_bytes workers max_tables max_rows holds_by_head
//   T, then T lines: key_type rows max_groups cost_per_row
//   R, then R lines: arrival_s table
#include "serve_sim.h"

#include <cstdio>
#include <cstdlib>

int main(int argc, char** argv)
{
    if (argc != 2) {
        std::fprintf(stderr, "usage: serve_replay TRACE\n");
        return 2;
    }
    FILE* f = std::fopen(argv[1], "r");
    if (!f) return 2;
    unsigned long long arena = 0, max_tables = 0, max_rows = 0;
    int workers = 0, by_head = 0;
    if (std::fscanf(f, "%llu %d %llu %llu %d", &arena, &workers, &max_tables, &max_rows, &by_head) != 5) return 2;
    int T = 0;
    if (std::fscanf(f, "%d", &T) != 1 || T <= 0) return 2;
    std::vector<vdf_sim::SimTable> tables(static_cast<size_t>(T));
    std::vector<int> key_type(static_cast<size_t>(T));
    const int vt[1] = {VDF_VAL_F64};
    for (int i = 0; i < T; ++i) {
        unsigned long long rows = 0, max_groups = 0;
        double cost = 0;
        if (std::fscanf(f, "%d %llu %llu %lf", &key_type[i], &rows, &max_groups, &cost) != 4) return 2;
        vdf::TableShape t;
        t.key_type = key_type[i];
        t.key_bytes = key_type[i] == VDF_KEY_CHAR16 ? 16 : 8;
        t.n = rows;
        t.nvals = 1;
        t.val_types[0] = VDF_VAL_F64;
        t.val_bytes[0] = 8;
        const uint64_t cap = (max_groups == 0 || max_groups > rows) ? rows : max_groups;
        tables[i].sig = vdf::signature(t, VDF_AGG_ALL);
        tables[i].rows = rows;
        tables[i].out_bytes = vdf::result_layout(t, VDF_AGG_ALL, false, cap, nullptr, nullptr);
        tables[i].cost_per_row = cost;
    }
    int R = 0;
    if (std::fscanf(f, "%d", &R) != 1 || R <= 0) return 2;
    std::vector<vdf_sim::SimRequest> reqs(static_cast<size_t>(R));
    for (int i = 0; i < R; ++i)
        if (std::fscanf(f, "%lf %d", &reqs[i].arrive, &reqs[i].table) != 2 || reqs[i].table >= T) return 2;
    std::fclose(f);

    // signature -> key type (one key type per signature in this trace)
    auto scratch = [&](uint64_t sig, const std::vector<uint64_t>& ns) -> uint64_t {
        int kt = -1;
        for (int i = 0; i < T; ++i)
            if (tables[i].sig == sig) kt = key_type[i];
        const uint64_t b = vdf_groupby_batch_scratch_bytes(kt, ns.data(), static_cast<int>(ns.size()), vt, 1);
        if (b == 0) {
            std::fprintf(stderr, "no scratch size for key type %d\n", kt);
            std::exit(2);
        }
        return b;
    };
    vdf_sim::SimConfig cfg;
    cfg.arena_bytes = arena;
    cfg.workers = workers;
    cfg.lim.max_tables = max_tables;
    cfg.lim.max_rows = max_rows;
    cfg.batch_overhead_s = 2e-4;
    cfg.holds_by_head = by_head != 0;
    const vdf_sim::SimResult r = vdf_sim::simulate(tables, reqs, cfg, scratch);
    std::printf("%s consumed=%zu/%d failed=%zu batches=%zu makespan_s=%.4f arena_peak=%llu max_queue=%zu %s\n",
                r.stalled ? "STALL" : "OK", r.consumed, R, r.failed, r.batches, r.makespan_s,
                static_cast<unsigned long long>(r.arena_peak), r.max_queue, r.state.c_str());
    return r.stalled ? 1 : 0;
}
