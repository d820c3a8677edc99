// serve_core.h -- host-only parts of the group-by request server: the range
// allocator that carves the server's HBM arena, and the planner that picks
// the next batch of queued requests.  No HIP here; the same code is compiled
// into libvdf_hip.so and into a plain g++ unit test.
//
// No reference counterpart: VisonDF runs one group-by per call on the calling
// thread (transform_group_by.hpp:11-14) and has no request loop.
#pragma once

#include <algorithm>
#include <cstdint>
#include <iterator>
#include <map>
#include <set>
#include <vector>

#include "../../include/vdf.h"

namespace vdf {

// First-fit allocator of byte ranges inside one arena of `size` bytes.  Every
// range starts on an `align` boundary and its length is rounded up to one.
// Adjacent free ranges are merged on free, so a drained arena is one range.
class RangeAllocator {
public:
    static constexpr uint64_t kNone = ~0ull;

    explicit RangeAllocator(uint64_t size = 0, uint64_t align = 256) : align_(align) { reset(size); }

    void reset(uint64_t size)
    {
        size_ = size / align_ * align_;
        free_.clear();
        live_.clear();
        used_ = peak_ = held_ = 0;
        holds_.clear();
        if (size_) free_[0] = size_;
    }

    uint64_t round(uint64_t bytes) const { return (std::max<uint64_t>(bytes, 1) + align_ - 1) / align_ * align_; }

    // Offset of a new range of at least `bytes`, or kNone when no free range is
    // large enough (hold: a transient reservation that the peak leaves out).
    uint64_t alloc(uint64_t bytes, bool hold = false)
    {
        const uint64_t need = round(bytes);
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            if (it->second < need) continue;
            const uint64_t off = it->first;
            const uint64_t rest = it->second - need;
            free_.erase(it);
            if (rest) free_[off + need] = rest;
            live_[off] = need;
            used_ += need;
            if (hold) {
                held_ += need;
                holds_.insert(off);
            }
            peak_ = std::max(peak_, used_ - held_);
            return off;
        }
        return kNone;
    }

    // Returns false for an offset that is not the start of a live range.
    bool release(uint64_t off)
    {
        auto it = live_.find(off);
        if (it == live_.end()) return false;
        uint64_t beg = off, len = it->second;
        live_.erase(it);
        used_ -= len;
        if (holds_.erase(off)) held_ -= len;
        auto nx = free_.lower_bound(beg);
        if (nx != free_.end() && nx->first == beg + len) {
            len += nx->second;
            nx = free_.erase(nx);
        }
        if (nx != free_.begin()) {
            auto pv = std::prev(nx);
            if (pv->first + pv->second == beg) {
                beg = pv->first;
                len += pv->second;
                free_.erase(pv);
            }
        }
        free_[beg] = len;
        return true;
    }

    uint64_t size() const { return size_; }
    uint64_t used() const { return used_; }
    uint64_t peak() const { return peak_; }
    size_t live_ranges() const { return live_.size(); }
    size_t free_ranges() const { return free_.size(); }
    uint64_t largest_free() const
    {
        uint64_t m = 0;
        for (const auto& f : free_) m = std::max(m, f.second);
        return m;
    }

private:
    uint64_t align_;
    uint64_t size_ = 0;
    uint64_t used_ = 0, peak_ = 0, held_ = 0;
    std::set<uint64_t> holds_;  // live ranges that are planning holds
    std::map<uint64_t, uint64_t> free_;  // offset -> length
    std::map<uint64_t, uint64_t> live_;  // offset -> length
};

// One queued request as the planner sees it.
struct PlanItem {
    uint64_t rows;          // rows of the request's table
    uint64_t sig;           // batch signature: requests of one batch share it
    uint64_t out_bytes;     // bytes of the request's result block
    uint64_t alone_scratch; // scratch of a batch that holds this request alone, from its
                            // own key type, value types and carried columns
};

struct PlanLimits {
    size_t max_tables = 64;          // requests per batch
    uint64_t max_rows = 1ull << 30;  // rows per batch (the oldest request is always taken)
};

struct Plan {
    std::vector<size_t> picked;    // queue positions, increasing
    std::vector<uint64_t> out_off; // arena offset of each picked request's result block
    uint64_t scr_off = RangeAllocator::kNone;
    uint64_t scr_bytes = 0;
    bool head_never_fits = false;  // the oldest request alone exceeds the whole arena
};

// Picks the next batch: the oldest queued request and, in queue order, later
// requests with its signature while the batch stays within the limits; each
// picked request gets its result block, then the batch gets one scratch block
// of scratch_bytes(rows of the picked requests), a function of the head's
// signature (every pick shares it).  Memory is taken in queue order and the
// scan stops at the first request whose block does not fit, so a large old
// request is never passed by newer ones for memory; when the scratch block
// does not fit, the newest picks are dropped until it does.
//
// Requests of another signature, and requests over the row limit, are
// skipped (they wait for a later batch), but memory is kept for each of them:
// its result block plus its own single-request scratch (alone_scratch, sized
// by its own signature: a 16-byte-key request needs more per row than an
// 8-byte-key head) is held while the batch is planned.  Without this, newer
// results held by a caller that consumes results in submission order could
// fill the arena while the older request waits for memory that only the
// caller can free.
//
// An empty plan means nothing can start until memory is released (or, with
// head_never_fits, that the oldest request can never run in this arena).
template <typename ScratchFn>
Plan plan_batch(const std::vector<PlanItem>& queue, const PlanLimits& lim, RangeAllocator& arena,
                ScratchFn scratch_bytes)
{
    Plan p;
    if (queue.empty()) return p;
    const PlanItem& head = queue[0];
    if (arena.round(head.out_bytes) + arena.round(head.alone_scratch) > arena.size()) {
        p.head_never_fits = true;
        return p;
    }
    // Every older request the batch passes (another signature, or over the
    // row limit, and older than a pick) gets its own need (result block +
    // scratch alone) held in the arena while the batch is planned, placed
    // just before the first younger pick's result block, so the batch only
    // takes memory beside those needs.  Younger requests' results then never
    // take the room an older request needs, whatever the caller holds: a
    // caller that consumes results in submission order cannot wait on a
    // request that no longer fits.  Requests skipped after the last pick are
    // not passed by anything and get no hold (holds for them would only take
    // room from the batch, down to the head's own scratch).
    uint64_t rows = 0;
    std::vector<uint64_t> ns;
    std::vector<uint64_t> held;        // arena offsets of the holds, in placement order
    std::vector<size_t> holds_before;  // per pick: holds placed before it
    std::vector<size_t> pending;       // skipped since the last pick
    auto release_holds_from = [&](size_t k) {
        for (size_t j = k; j < held.size(); ++j) arena.release(held[j]);
        held.resize(k);
    };
    for (size_t i = 0; i < queue.size() && p.picked.size() < lim.max_tables; ++i) {
        const PlanItem& q = queue[i];
        if (q.sig != head.sig || (i > 0 && rows + q.rows > lim.max_rows)) {
            pending.push_back(i);
            continue;
        }
        const size_t mark = held.size();
        bool room = true;
        for (size_t j : pending) {
            const PlanItem& o = queue[j];
            const uint64_t off = arena.alloc(arena.round(o.out_bytes) + arena.round(o.alone_scratch), true);
            if (off == RangeAllocator::kNone) {
                room = false;  // no room for an older request: take nothing younger
                break;
            }
            held.push_back(off);
        }
        const uint64_t off = room ? arena.alloc(q.out_bytes) : RangeAllocator::kNone;
        if (off == RangeAllocator::kNone) {
            release_holds_from(mark);
            break;
        }
        pending.clear();
        p.picked.push_back(i);
        p.out_off.push_back(off);
        holds_before.push_back(mark);
        ns.push_back(q.rows);
        rows += q.rows;
    }
    // the batch's scratch, beside the holds of the requests it passes; a pick
    // that is dropped takes the holds placed for it along
    while (!p.picked.empty()) {
        const uint64_t sb = scratch_bytes(ns);
        const uint64_t off = arena.alloc(sb);
        if (off != RangeAllocator::kNone) {
            p.scr_off = off;
            p.scr_bytes = sb;
            break;
        }
        arena.release(p.out_off.back());
        release_holds_from(holds_before.back());
        p.out_off.pop_back();
        p.picked.pop_back();
        holds_before.pop_back();
        ns.pop_back();
    }
    release_holds_from(0);
    return p;
}

// Shape of a registered table as the server's accounting sees it.
struct TableShape {
    int key_type = 0;
    uint32_t key_bytes = 0;  // bytes of one key cell
    uint64_t n = 0;
    int nvals = 0;
    int val_types[VDF_MAX_VALUE_COLS] = {};
    uint32_t val_bytes[VDF_MAX_VALUE_COLS] = {};
};

// Lays out a request's result block at `base` (base null: only measures it):
// group keys, offsets, counts, the permutation when asked, and the per-column
// aggregates of agg_mask, each array 256-byte aligned, for `cap` groups.
// Returns the block's bytes.
inline uint64_t result_layout(const TableShape& t, uint32_t agg_mask, bool want_perm, uint64_t cap, char* base,
                              vdf_groupby_out* o)
{
    const uint64_t c = std::max<uint64_t>(cap, 1);
    uint64_t off = 0;
    auto take = [&](uint64_t bytes) -> void* {
        off = (off + 255) / 256 * 256;
        void* p = base ? base + off : nullptr;
        off += bytes;
        return p;
    };
    vdf_groupby_out r{};
    r.group_keys = take(c * t.key_bytes);
    r.offsets = static_cast<uint32_t*>(take((cap + 1) * 4));
    if (agg_mask & VDF_AGG_COUNT) r.counts = static_cast<uint32_t*>(take(c * 4));
    if (want_perm) r.perm = static_cast<uint32_t*>(take(std::max<uint64_t>(t.n, 1) * 4));
    for (int i = 0; i < t.nvals; ++i) {
        const uint64_t vb = t.val_bytes[i];
        if (agg_mask & VDF_AGG_SUM) r.aggs[i].sum = take(c * 8);
        if (agg_mask & VDF_AGG_MEAN) r.aggs[i].mean = static_cast<double*>(take(c * 8));
        if (agg_mask & VDF_AGG_MIN) r.aggs[i].min = take(c * vb);
        if (agg_mask & VDF_AGG_MAX) r.aggs[i].max = take(c * vb);
    }
    r.capacity = cap;
    if (o) *o = r;
    return (off + 255) / 256 * 256;
}

// Batch signature: requests of one batch share key type, value types and
// aggregate mask.
inline uint64_t signature(const TableShape& t, uint32_t agg_mask)
{
    uint64_t s = static_cast<uint64_t>(t.key_type) | (static_cast<uint64_t>(t.nvals) << 4) |
                 (static_cast<uint64_t>(agg_mask) << 8);
    for (int i = 0; i < t.nvals; ++i) s |= static_cast<uint64_t>(t.val_types[i] & 0xF) << (16 + 4 * i);
    return s;
}

}  // namespace vdf
