// tune.hip -- autotuner of the radix / group-by launch shape.
//
// Times the caller's own workload (a sort or a batched group-by, on the
// caller's stream and buffers) under candidate launch shapes and keeps the
// fastest in the handle's tuning table.  Knobs: rows per thread of a radix
// pass tile (tile size), persistent blocks per CU (block count), XCD tile
// chunk and tile order, and rows per histogram work item.  The digit width is
// fixed at 8 bits: 256 buckets map one-to-one onto the 256 threads of a block
// and a wave's 64-bit match masks, and an 8-bit pass's tile histogram and
// look-back row fit LDS and one cache line per thread.  Search: coordinate
// descent from the current table (each knob in turn over its candidates, the
// others held), two sweeps or until nothing improves by more than 1%; every
// candidate is timed as the minimum over `reps` runs after a warm-up run.
// Results do not depend on any knob (the sort is stable and the group-by
// tile is fixed), so every run leaves correct outputs.
//
// No reference counterpart: the reference fixes digit width and thread count
// at compile time (radix_sort_uint32_mt.inl:3-5).
#include "vdf_internal.h"

#include <functional>

namespace vdf {
namespace {

struct Knob {
    int Tuning::*field;
    std::vector<int> values;
};

void run_groupby(Handle* h, int kt, const std::vector<GbTable>& ts, const int* vt, int nv, uint32_t mask,
                 const vdf_groupby_out* outs, hipStream_t s)
{
    std::vector<uint64_t> ns(ts.size());
    for (size_t i = 0; i < ts.size(); ++i) ns[i] = ts[i].n;
    const int nt = static_cast<int>(ts.size());
    const uint64_t hr = static_cast<uint64_t>(h->tune.hist_rows);
    void* scr = h->ensure_scratch(groupby_batch_scratch(kt, ns.data(), nt, hr, groupby_carry_bytes(h->tune, kt, vt, nv)));
    h->staging_begin(groupby_batch_staging(kt, ns.data(), nt, hr));
    try {
        groupby_batch_enqueue(h, kt, ts.data(), nt, vt, nv, mask, outs, scr, s);
    } catch (...) {
        h->staging_end(s);
        throw;
    }
    h->staging_end(s);
}

// one sort; with pay_bytes > 0 a value column rides along (pay_in -> pay_out)
void run_argsort(Handle* h, int kt, const void* keys, uint64_t n, uint32_t flags, void* perm, hipStream_t s,
                 int pay_bytes = 0, const void* pay_in = nullptr, void* pay_out = nullptr)
{
    const uint64_t hr = static_cast<uint64_t>(h->tune.hist_rows);
    void* scr = h->ensure_scratch(argsort_batch_scratch(kt, &n, 1, hr, pay_bytes));
    h->staging_begin(argsort_batch_staging(kt, &n, 1, hr));
    try {
        SortSeg seg{keys, n, perm, nullptr};
        seg.pay_in = pay_in;
        seg.pay_out = pay_out;
        argsort_batch_enqueue(h, kt, &seg, 1, flags, scr, s, pay_bytes);
    } catch (...) {
        h->staging_end(s);
        throw;
    }
    h->staging_end(s);
}

// Coordinate descent over the knobs; returns trials run.  kb: key bytes;
// pay_bytes: payload bytes moved with the keys (selects the tile sizes that
// have pass kernels with that payload).
int descend(Handle* h, int kb, int reps, hipStream_t s, const std::function<void()>& run, vdf_tune_report* rep,
            int pay_bytes = 0)
{
    hipEvent_t e0, e1;
    VDF_HIP_CHECK(hipEventCreate(&e0));
    VDF_HIP_CHECK(hipEventCreate(&e1));
    int trials = 0;
    auto time_it = [&](const Tuning& t) -> double {
        h->tune = t;
        run();  // warm-up (scratch growth, caches)
        double best = 1e30;
        for (int r = 0; r < reps; ++r) {
            VDF_HIP_CHECK(hipEventRecord(e0, s));
            run();
            VDF_HIP_CHECK(hipEventRecord(e1, s));
            VDF_HIP_CHECK(hipEventSynchronize(e1));
            float ms = 0.f;
            VDF_HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
            best = std::min(best, static_cast<double>(ms));
        }
        ++trials;
        return best;
    };
    std::vector<Knob> knobs;
    if (kb == 4)
        knobs.push_back({&Tuning::sort_items_32,
                         pay_bytes == 0   ? std::vector<int>{16, 24, 32}
                         : pay_bytes == 4 ? std::vector<int>{8, 12, 24}
                         : pay_bytes == 8 ? std::vector<int>{12, 16, 24}
                                          : std::vector<int>{8, 12, 16}});
    if (kb == 8)
        knobs.push_back({&Tuning::sort_items_64, pay_bytes == 0   ? std::vector<int>{8, 12, 16}
                                                 : pay_bytes == 4 ? std::vector<int>{8, 12}
                                                                  : std::vector<int>{8, 12, 16}});
    if (kb == 16) knobs.push_back({&Tuning::sort_items_128, {8, 16}});
    if (kb == 4) knobs.push_back({&Tuning::sort_local, {1, 0}});
    knobs.push_back({&Tuning::sort_bpc, {0, 1, 2, 3}});
    knobs.push_back({&Tuning::sort_xcd, {1, 0}});
    knobs.push_back({&Tuning::sort_chunk, {4, 16, 64, 256}});
    knobs.push_back({&Tuning::hist_rows, {16384, 65536, 262144}});
    try {
        Tuning best = h->tune;
        double best_ms = time_it(best);
        rep->default_ms = best_ms;
        for (int sweep = 0; sweep < 2; ++sweep) {
            bool moved = false;
            for (const Knob& k : knobs) {
                for (int v : k.values) {
                    if (best.*(k.field) == v) continue;
                    Tuning t = best;
                    t.*(k.field) = v;
                    const double ms = time_it(t);
                    if (ms < best_ms * 0.99) {
                        best = t;
                        best_ms = ms;
                        moved = true;
                    }
                }
            }
            if (!moved) break;
        }
        h->tune = best;
        rep->best_ms = best_ms;
        rep->sort_items = kb == 4 ? best.sort_items_32 : (kb == 8 ? best.sort_items_64 : best.sort_items_128);
        rep->sort_blocks_per_cu = best.sort_bpc;
        rep->sort_xcd = best.sort_xcd;
        rep->sort_chunk = best.sort_chunk;
        rep->hist_rows = best.hist_rows;
        rep->digit_bits = kRadixBits;
        rep->sort_local = best.sort_local;
    } catch (...) {
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        throw;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    rep->trials = trials;
    return trials;
}

// RAII: lock the handle and make its device current (as the entry points do)
struct TuneScope {
    std::lock_guard<std::mutex> lk;
    int prev = -1;
    explicit TuneScope(Handle* h) : lk(h->mu)
    {
        VDF_HIP_CHECK(hipGetDevice(&prev));
        if (prev != h->device) VDF_HIP_CHECK(hipSetDevice(h->device));
    }
    ~TuneScope()
    {
        int cur = -1;
        if (hipGetDevice(&cur) == hipSuccess && prev >= 0 && cur != prev) (void)hipSetDevice(prev);
    }
};

}  // namespace
}  // namespace vdf

using namespace vdf;

extern "C" {

vdf_status vdf_autotune_groupby_batch(vdf_handle* h, int key_type, const vdf_table* tables, int ntables,
                                      const int* val_types, int nvals, uint32_t agg_mask, vdf_groupby_out* outs,
                                      int reps, void* stream, vdf_tune_report* report)
{
    // one ordinary call validates the arguments (and warms up)
    const vdf_status st = vdf_groupby_batch(h, key_type, tables, ntables, val_types, nvals, agg_mask, outs, stream);
    if (st != VDF_OK) return st;
    return guarded([&] {
        require(report != nullptr, "vdf_autotune_groupby_batch: null report");
        require(reps >= 1 && reps <= 100, "vdf_autotune_groupby_batch: reps must be 1 .. 100");
        *report = vdf_tune_report{};
        if (ntables == 0) return;
        std::vector<GbTable> ts(static_cast<size_t>(ntables));
        for (int i = 0; i < ntables; ++i) {
            ts[i].keys = tables[i].keys;
            ts[i].n = tables[i].n;
            for (int c = 0; c < VDF_MAX_VALUE_COLS; ++c) ts[i].vals[c] = c < nvals ? tables[i].vals[c] : nullptr;
        }
        Handle* hh = h;
        TuneScope scope(hh);
        const hipStream_t s = static_cast<hipStream_t>(stream);
        descend(hh, static_cast<int>(key_bytes(key_type)), reps, s,
                [&] { run_groupby(hh, key_type, ts, val_types, nvals, agg_mask, outs, s); }, report);
    });
}

vdf_status vdf_autotune_argsort(vdf_handle* h, int key_type, const void* d_keys, uint64_t n, uint32_t flags,
                                void* d_perm_out, int reps, void* stream, vdf_tune_report* report)
{
    const vdf_status st = vdf_argsort(h, key_type, d_keys, n, flags, d_perm_out, nullptr, stream);
    if (st != VDF_OK) return st;
    return guarded([&] {
        require(report != nullptr, "vdf_autotune_argsort: null report");
        require(reps >= 1 && reps <= 100, "vdf_autotune_argsort: reps must be 1 .. 100");
        *report = vdf_tune_report{};
        if (n == 0) return;
        Handle* hh = h;
        TuneScope scope(hh);
        const hipStream_t s = static_cast<hipStream_t>(stream);
        descend(hh, static_cast<int>(key_bytes(key_type)), reps, s,
                [&] { run_argsort(hh, key_type, d_keys, n, flags, d_perm_out, s); }, report);
    });
}

vdf_status vdf_autotune_sort_by(vdf_handle* h, int key_type, const void* d_keys, uint64_t n, uint32_t flags,
                                void* d_perm_out, uint32_t val_bytes, const void* d_vals, void* d_vals_out, int reps,
                                void* stream, vdf_tune_report* report)
{
    const vdf_status st =
        vdf_sort_by(h, key_type, d_keys, n, flags, d_perm_out, nullptr, val_bytes, d_vals, d_vals_out, stream);
    if (st != VDF_OK) return st;
    return guarded([&] {
        require(report != nullptr, "vdf_autotune_sort_by: null report");
        require(reps >= 1 && reps <= 100, "vdf_autotune_sort_by: reps must be 1 .. 100");
        *report = vdf_tune_report{};
        if (n == 0) return;
        Handle* hh = h;
        TuneScope scope(hh);
        const hipStream_t s = static_cast<hipStream_t>(stream);
        const int pb = static_cast<int>(val_bytes);
        descend(
            hh, static_cast<int>(key_bytes(key_type)), reps, s,
            [&] { run_argsort(hh, key_type, d_keys, n, flags, d_perm_out, s, pb, d_vals, d_vals_out); }, report,
            pb);
    });
}

}  // extern "C"
