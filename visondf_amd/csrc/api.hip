// api.hip -- the extern "C" boundary of libvdf_hip.so (declared in include/vdf.h).
//
// Every entry point: validate arguments, take the handle's lock, make the
// handle's device current (restored on return), enqueue on the caller's
// stream, and turn exceptions into a status code plus a thread-local message.
#include "vdf_internal.h"

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <exception>
#include <memory>

namespace vdf {

static thread_local std::string g_last_error;

// Scratch of the allocation-free sort (vdf_sort_by_noalloc): one fixed array
// per device, held from the library's load until the process ends.  Its name
// on the host is only a stand-in for the symbol; each device's address comes
// from hipGetSymbolAddress (taken in vdf_create with that device current) and
// only that address is handed to kernels.
static __device__ __align__(256) char scratch[VDF_NOALLOC_SCRATCH_BYTES];

constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static void* g_scratch_ptr[kMaxDevices] = {};
static std::string g_scratch_err[kMaxDevices];

// Looks up the device address of `scratch` on `dev`, which must be current,
// once per device.  A failed lookup is kept as that device's error, reported
// by the allocation-free entry point (other entry points do not need it).
static void noalloc_scratch_setup(int dev)
{
    if (dev < 0 || dev >= kMaxDevices) return;
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (g_scratch_ptr[dev]) return;
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || p == nullptr) {
        (void)hipGetLastError();
        g_scratch_err[dev] = std::string("hipGetSymbolAddress of the fixed scratch failed: ") + hipGetErrorString(e);
        return;
    }
    g_scratch_ptr[dev] = p;
    g_scratch_err[dev].clear();
}

// The fixed scratch's address on `dev` (set up by vdf_create on that device).
static void* noalloc_scratch(int dev)
{
    require(dev >= 0 && dev < kMaxDevices, "device index beyond the library's device table");
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[dev])
        throw Error{VDF_ERR_HIP, g_scratch_err[dev].empty() ? "the fixed scratch of this device was not set up"
                                                            : g_scratch_err[dev]};
    return g_scratch_ptr[dev];
}

void set_error(const std::string& msg) { g_last_error = msg; }

void* Handle::ensure_scratch(size_t bytes)
{
    if (bytes <= scratch_cap) return scratch;
    if (budget && bytes > budget)
        throw Error{VDF_ERR_BUDGET, "scratch of " + std::to_string(bytes) + " bytes exceeds the handle budget of " +
                                        std::to_string(budget) + " bytes"};
    // Grow: earlier work of this handle may still use the old arena.
    VDF_HIP_CHECK(hipDeviceSynchronize());
    if (scratch) VDF_HIP_CHECK(hipFree(scratch));
    scratch = nullptr;
    scratch_cap = 0;
    size_t want = bytes + bytes / 8;
    if (budget) want = std::min<size_t>(want, budget);
    void* p = nullptr;
    if (hipMalloc(&p, want) != hipSuccess) {
        (void)hipGetLastError();
        want = bytes;
        if (hipMalloc(&p, want) != hipSuccess) {
            (void)hipGetLastError();
            throw Error{VDF_ERR_OOM, "hipMalloc of " + std::to_string(want) + " scratch bytes failed"};
        }
    }
    scratch = p;
    scratch_cap = want;
    return scratch;
}

void Handle::staging_begin(size_t bytes)
{
    if (staging_fixed) {
        if (bytes > staging_fixed_cap) throw Error{VDF_ERR_HIP, "plan staging buffer too small"};
        staging_slot = kStaging;  // marks the fixed buffer
        staging_off = 0;
        return;
    }
    const int slot = static_cast<int>(staging_next++ % kStaging);
    if (staging_pending[slot]) {
        VDF_HIP_CHECK(hipEventSynchronize(staging_done[slot]));
        staging_pending[slot] = false;
    }
    if (bytes > staging_cap[slot]) {
        if (staging[slot]) VDF_HIP_CHECK(hipHostFree(staging[slot]));
        staging[slot] = nullptr;
        staging_cap[slot] = 0;
        const size_t want = std::max<size_t>(bytes + bytes / 4, 1 << 16);
        VDF_HIP_CHECK(hipHostMalloc(&staging[slot], want, hipHostMallocDefault));
        staging_cap[slot] = want;
    }
    staging_slot = slot;
    staging_off = 0;
}

void* Handle::staging_take(size_t bytes)
{
    if (staging_slot < 0) throw Error{VDF_ERR_HIP, "staging_take outside a call"};
    staging_off = align_up(staging_off, 256);
    const bool fixed = staging_slot == kStaging;
    if (staging_off + bytes > (fixed ? staging_fixed_cap : staging_cap[staging_slot]))
        throw Error{VDF_ERR_HIP, "staging region too small"};
    void* p = static_cast<char*>(fixed ? staging_fixed : staging[staging_slot]) + staging_off;
    staging_off += bytes;
    return p;
}

void Handle::staging_end(hipStream_t s)
{
    if (staging_slot < 0) return;
    if (staging_slot == kStaging) {  // a plan's buffer: the plan tracks its reuse
        staging_slot = -1;
        return;
    }
    if (!staging_done[staging_slot])
        VDF_HIP_CHECK(hipEventCreateWithFlags(&staging_done[staging_slot], hipEventDisableTiming));
    VDF_HIP_CHECK(hipEventRecord(staging_done[staging_slot], s));
    staging_pending[staging_slot] = true;
    staging_slot = -1;
}

void Handle::release_resources()
{
    for (ProfRec& r : recs) {
        (void)hipEventDestroy(r.beg);
        (void)hipEventDestroy(r.end);
        if (r.groups_done) (void)hipEventDestroy(r.groups_done);
    }
    recs.clear();
    for (hipEvent_t e : free_events) (void)hipEventDestroy(e);
    free_events.clear();
    if (scratch) (void)hipFree(scratch);
    scratch = nullptr;
    scratch_cap = 0;
    if (passes_ring) (void)hipHostFree(passes_ring);
    passes_ring = nullptr;
    for (int i = 0; i < kStaging; ++i) {
        if (staging[i]) (void)hipHostFree(staging[i]);
        if (staging_done[i]) (void)hipEventDestroy(staging_done[i]);
        staging[i] = nullptr;
        staging_done[i] = nullptr;
        staging_cap[i] = 0;
        staging_pending[i] = false;
    }
}

hipEvent_t Handle::get_event()
{
    if (!free_events.empty()) {
        hipEvent_t e = free_events.back();
        free_events.pop_back();
        return e;
    }
    hipEvent_t e;
    VDF_HIP_CHECK(hipEventCreate(&e));
    return e;
}

void Handle::prof_collect()
{
    for (ProfRec& r : recs) {
        VDF_HIP_CHECK(hipEventSynchronize(r.end));
        float ms = 0.f;
        VDF_HIP_CHECK(hipEventElapsedTime(&ms, r.beg, r.end));
        uint64_t bytes = r.bytes;
        bool active = true;
        if (r.plan_slot >= 0 && r.acct) {
            // a pass slot did work for the tables whose plan has more passes
            const PassAcct& pa = *r.acct;
            const uint32_t s = static_cast<uint32_t>(r.plan_slot);
            bytes = 0;
            active = false;
            for (size_t i = 0; i < pa.rows.size(); ++i) {
                // slots from kLocalSlot0 on, and the local finish itself, run by
                // the plan as the local finish's size check left it
                const bool late = pa.passes2 && (s == kProfLocalSlot || s >= static_cast<uint32_t>(kLocalSlot0));
                const uint32_t* w = (late ? pa.passes2 : pa.passes) + 2 * i;
                const uint32_t end = w[0], mode = w[1], s0 = mode & 0xFFu;
                const bool local = (mode & kModeLocal) != 0;
                if (s == kProfLocalSlot) {
                    if (!local) continue;
                    active = true;
                    bytes += pa.rows[i] * pa.per_row_local;
                    continue;
                }
                if (s >= end || s < s0) continue;
                active = true;
                const bool first = s == s0, last = s == end - 1 && !local;
                const int pos = first && last ? 0 : first ? 1 : last ? 3 : 2;
                bytes += pa.rows[i] * pa.per_row[pos];
            }
        }
        if (r.groups) {
            VDF_HIP_CHECK(hipEventSynchronize(r.groups_done));
            bytes += *r.groups * r.group_bytes;
            free_events.push_back(r.groups_done);
        }
        if (active) {
            auto it = std::find_if(totals.begin(), totals.end(), [&](const ProfTotal& t) { return t.name == r.name; });
            if (it == totals.end()) {
                totals.push_back(ProfTotal{r.name, 0.0, 0, 0});
                it = totals.end() - 1;
            }
            it->ms += ms;
            it->launches += 1;
            it->bytes += bytes;
        }
        free_events.push_back(r.beg);
        free_events.push_back(r.end);
    }
    recs.clear();
}

// RAII: lock the handle and make its device current for the call
struct CallScope {
    std::lock_guard<std::mutex> lk;
    int prev = -1;
    explicit CallScope(Handle* h) : lk(h->mu)
    {
        VDF_HIP_CHECK(hipGetDevice(&prev));
        if (prev != h->device) VDF_HIP_CHECK(hipSetDevice(h->device));
    }
    ~CallScope()
    {
        int cur = -1;
        if (hipGetDevice(&cur) == hipSuccess && prev >= 0 && cur != prev) (void)hipSetDevice(prev);
    }
};

// ends the call's staging region on every exit path
struct StagingScope {
    Handle* h;
    hipStream_t s;
    StagingScope(Handle* h_, hipStream_t s_) : h(h_), s(s_) {}
    ~StagingScope()
    {
        try {
            h->staging_end(s);
        } catch (...) {
        }
    }
};

}  // namespace vdf

using namespace vdf;

extern "C" {

const char* vdf_last_error(void) { return g_last_error.c_str(); }

const char* vdf_version(void) { return "vdf_hip 0.1 (gfx950)"; }

vdf_status vdf_create(vdf_handle** out, int device, uint64_t hbm_budget_bytes)
{
    return guarded([&] {
        require(out != nullptr, "vdf_create: out is null");
        *out = nullptr;
        int ndev = 0;
        VDF_HIP_CHECK(hipGetDeviceCount(&ndev));
        require(device >= 0 && device < ndev, "vdf_create: no such device");
        {
            // the fixed scratch's address on this device, with the device current
            int prev = 0;
            VDF_HIP_CHECK(hipGetDevice(&prev));
            if (prev != device) VDF_HIP_CHECK(hipSetDevice(device));
            noalloc_scratch_setup(device);
            (void)lds_rank_check(device);
            if (prev != device) VDF_HIP_CHECK(hipSetDevice(prev));
        }
        auto* h = new vdf_handle();
        h->device = device;
        h->budget = hbm_budget_bytes;
        *out = h;
    });
}

vdf_status vdf_destroy(vdf_handle* h)
{
    return guarded([&] {
        if (!h) return;
        {
            CallScope cs(h);
            VDF_HIP_CHECK(hipDeviceSynchronize());
            h->release_resources();
        }
        delete h;
    });
}

uint64_t vdf_argsort_scratch_bytes(int key_type, uint64_t n, uint32_t flags, int want_keys)
{
    (void)want_keys;
    if (key_bytes(key_type) == 0) return 0;
    return argsort_scratch(key_type, n, flags);
}

uint64_t vdf_groupby_scratch_bytes(int key_type, uint64_t n)
{
    if (key_bytes(key_type) == 0) return 0;
    return groupby_scratch(key_type, n);
}

uint64_t vdf_groupby_batch_scratch_bytes(int key_type, const uint64_t* n, int ntables, const int* val_types,
                                         int nvals)
{
    if (key_bytes(key_type) == 0 || ntables < 0 || (ntables > 0 && n == nullptr)) return 0;
    if (nvals < 0 || nvals > VDF_MAX_VALUE_COLS || (nvals > 0 && val_types == nullptr)) return 0;
    for (int c = 0; c < nvals; ++c)
        if (val_bytes(val_types[c]) == 0) return 0;
    const Tuning t;
    return groupby_batch_scratch(key_type, n, ntables, static_cast<uint64_t>(t.hist_rows),
                                 groupby_carry_bytes(t, key_type, val_types, nvals));
}

vdf_status vdf_argsort(vdf_handle* h, int key_type, const void* d_keys, uint64_t n, uint32_t flags,
                       void* d_perm_out, void* d_sorted_keys_out, void* stream)
{
    return guarded([&] {
        require(h != nullptr, "vdf_argsort: null handle");
        require(key_bytes(key_type) != 0, "vdf_argsort: unknown key type");
        require((flags & ~(VDF_SORT_DESC | VDF_SORT_DESC_REVERSE | VDF_SORT_PERM64)) == 0,
                "vdf_argsort: unknown flag bits");
        require(!((flags & VDF_SORT_DESC) && (flags & VDF_SORT_DESC_REVERSE)),
                "vdf_argsort: VDF_SORT_DESC and VDF_SORT_DESC_REVERSE are exclusive");
        require(n <= VDF_MAX_ROWS, "vdf_argsort: n exceeds VDF_MAX_ROWS (2^30 - 1)");
        if (n == 0) return;
        require(d_keys != nullptr && d_perm_out != nullptr, "vdf_argsort: null device pointer");
        CallScope cs(h);
        const uint64_t hr = static_cast<uint64_t>(h->tune.hist_rows);
        void* scr = h->ensure_scratch(argsort_batch_scratch(key_type, &n, 1, hr));
        h->staging_begin(argsort_batch_staging(key_type, &n, 1, hr));
        StagingScope ss(h, static_cast<hipStream_t>(stream));
        argsort_enqueue(h, key_type, d_keys, n, flags, d_perm_out, d_sorted_keys_out, scr,
                        static_cast<hipStream_t>(stream));
    });
}

vdf_status vdf_sort_by(vdf_handle* h, int key_type, const void* d_keys, uint64_t n, uint32_t flags,
                       void* d_perm_out, void* d_sorted_keys_out, uint32_t val_bytes, const void* d_vals,
                       void* d_vals_out, void* stream)
{
    return guarded([&] {
        require(h != nullptr, "vdf_sort_by: null handle");
        require(key_bytes(key_type) != 0, "vdf_sort_by: unknown key type");
        require((flags & ~(VDF_SORT_DESC | VDF_SORT_DESC_REVERSE | VDF_SORT_PERM64)) == 0,
                "vdf_sort_by: unknown flag bits");
        require(!((flags & VDF_SORT_DESC) && (flags & VDF_SORT_DESC_REVERSE)),
                "vdf_sort_by: VDF_SORT_DESC and VDF_SORT_DESC_REVERSE are exclusive");
        require(val_bytes == 4 || val_bytes == 8, "vdf_sort_by: val_bytes must be 4 or 8");
        require(n <= VDF_MAX_ROWS, "vdf_sort_by: n exceeds VDF_MAX_ROWS (2^30 - 1)");
        if (n == 0) return;
        require(d_keys != nullptr && d_vals != nullptr && d_vals_out != nullptr, "vdf_sort_by: null device pointer");
        CallScope cs(h);
        const hipStream_t s = static_cast<hipStream_t>(stream);
        const uint64_t hr = static_cast<uint64_t>(h->tune.hist_rows);
        const int pb = static_cast<int>(val_bytes);
        void* scr = h->ensure_scratch(argsort_batch_scratch(key_type, &n, 1, hr, pb));
        h->staging_begin(argsort_batch_staging(key_type, &n, 1, hr));
        StagingScope ss(h, s);
        SortSeg seg{d_keys, n, d_perm_out, d_sorted_keys_out};
        seg.pay_in = d_vals;
        seg.pay_out = d_vals_out;
        argsort_batch_enqueue(h, key_type, &seg, 1, flags, scr, s, pb);
    });
}

vdf_status vdf_sort_by_noalloc(vdf_handle* h, int key_type, const void* d_keys, uint64_t n, uint32_t flags,
                               void* d_perm_out, void* d_sorted_keys_out, uint32_t val_bytes, const void* d_vals,
                               void* d_vals_out, void* stream)
{
    return guarded([&] {
        require(h != nullptr, "vdf_sort_by_noalloc: null handle");
        require(key_bytes(key_type) != 0, "vdf_sort_by_noalloc: unknown key type");
        require((flags & ~(VDF_SORT_DESC | VDF_SORT_DESC_REVERSE | VDF_SORT_PERM64)) == 0,
                "vdf_sort_by_noalloc: unknown flag bits");
        require(!((flags & VDF_SORT_DESC) && (flags & VDF_SORT_DESC_REVERSE)),
                "vdf_sort_by_noalloc: VDF_SORT_DESC and VDF_SORT_DESC_REVERSE are exclusive");
        require(val_bytes == 4 || val_bytes == 8, "vdf_sort_by_noalloc: val_bytes must be 4 or 8");
        require(n <= VDF_MAX_ROWS, "vdf_sort_by_noalloc: n exceeds VDF_MAX_ROWS (2^30 - 1)");
        if (n == 0) return;
        require(d_keys != nullptr && d_vals != nullptr && d_vals_out != nullptr,
                "vdf_sort_by_noalloc: null device pointer");
        const int pb = static_cast<int>(val_bytes);
        const uint64_t hr = static_cast<uint64_t>(h->tune.hist_rows);
        const size_t need = argsort_batch_scratch(key_type, &n, 1, hr, pb);
        if (need > VDF_NOALLOC_SCRATCH_BYTES) {
            char msg[192];
            std::snprintf(msg, sizeof msg,
                          "vdf_sort_by_noalloc: %llu rows need %llu scratch bytes, above the fixed %llu "
                          "(VDF_NOALLOC_SCRATCH_BYTES); use vdf_sort_by",
                          static_cast<unsigned long long>(n), static_cast<unsigned long long>(need),
                          static_cast<unsigned long long>(VDF_NOALLOC_SCRATCH_BYTES));
            throw Error{VDF_ERR_INVALID, msg};
        }
        CallScope cs(h);
        void* scr = noalloc_scratch(h->device);
        const hipStream_t s = static_cast<hipStream_t>(stream);
        // no pinned staging (descriptors written on the device), no profiling events
        struct Restore {
            Handle* h;
            bool prof;
            ~Restore()
            {
                h->device_desc = false;
                h->prof = prof;
            }
        } restore{h, h->prof};
        h->device_desc = true;
        h->prof = false;
        SortSeg seg{d_keys, n, d_perm_out, d_sorted_keys_out};
        seg.pay_in = d_vals;
        seg.pay_out = d_vals_out;
        argsort_batch_enqueue(h, key_type, &seg, 1, flags, scr, s, pb);
    });
}

vdf_status vdf_gather(vdf_handle* h, uint32_t elem_bytes, const void* d_src, const void* d_perm, int perm64,
                      uint64_t n, void* d_dst, void* stream)
{
    return guarded([&] {
        require(h != nullptr, "vdf_gather: null handle");
        require(elem_bytes >= 1 && elem_bytes <= VDF_MAX_STRING_BYTES,
                "vdf_gather: elem_bytes must be 1 .. VDF_MAX_STRING_BYTES");
        if (n == 0) return;
        require(d_src && d_perm && d_dst, "vdf_gather: null device pointer");
        CallScope cs(h);
        gather_enqueue(h, elem_bytes, d_src, d_perm, perm64, n, d_dst, static_cast<hipStream_t>(stream));
    });
}

vdf_status vdf_groupby(vdf_handle* h, int key_type, const void* d_keys, uint64_t n, const vdf_value_col* vals,
                       int nvals, uint32_t agg_mask, vdf_groupby_out* out, void* stream)
{
    return guarded([&] {
        require(h != nullptr, "vdf_groupby: null handle");
        require(key_bytes(key_type) != 0, "vdf_groupby: unknown key type");
        require(out != nullptr, "vdf_groupby: null output descriptor");
        require(nvals >= 0 && nvals <= VDF_MAX_VALUE_COLS, "vdf_groupby: nvals out of range");
        require(nvals == 0 || vals != nullptr, "vdf_groupby: null value columns");
        require((agg_mask & ~VDF_AGG_ALL) == 0, "vdf_groupby: unknown aggregate bits");
        require(n <= VDF_MAX_ROWS, "vdf_groupby: n exceeds VDF_MAX_ROWS (2^30 - 1)");
        require(out->n_groups != nullptr && out->offsets != nullptr, "vdf_groupby: n_groups and offsets are required");
        for (int c = 0; c < nvals; ++c) {
            require(val_bytes(vals[c].type) != 0, "vdf_groupby: unknown value type");
            require(n == 0 || vals[c].data != nullptr, "vdf_groupby: null value column data");
        }
        require(n == 0 || d_keys != nullptr, "vdf_groupby: null keys");
        CallScope cs(h);
        const uint64_t hr = static_cast<uint64_t>(h->tune.hist_rows);
        int vt[VDF_MAX_VALUE_COLS] = {0, 0, 0, 0};
        for (int c = 0; c < nvals; ++c) vt[c] = vals[c].type;
        const int pay = groupby_carry_bytes(h->tune, key_type, vt, nvals);
        void* scr = h->ensure_scratch(groupby_batch_scratch(key_type, &n, 1, hr, pay));
        h->staging_begin(groupby_batch_staging(key_type, &n, 1, hr));
        StagingScope ss(h, static_cast<hipStream_t>(stream));
        groupby_enqueue(h, key_type, d_keys, n, vals, nvals, agg_mask, out, scr, static_cast<hipStream_t>(stream));
    });
}

vdf_status vdf_argsort_batch(vdf_handle* h, int key_type, const vdf_sort_seg* segs, int nsegs, uint32_t flags,
                             void* stream)
{
    return guarded([&] {
        require(h != nullptr, "vdf_argsort_batch: null handle");
        require(key_bytes(key_type) != 0, "vdf_argsort_batch: unknown key type");
        require((flags & ~(VDF_SORT_DESC | VDF_SORT_DESC_REVERSE | VDF_SORT_PERM64)) == 0,
                "vdf_argsort_batch: unknown flag bits");
        require(!((flags & VDF_SORT_DESC) && (flags & VDF_SORT_DESC_REVERSE)),
                "vdf_argsort_batch: VDF_SORT_DESC and VDF_SORT_DESC_REVERSE are exclusive");
        require(nsegs >= 0 && nsegs <= (1 << 22), "vdf_argsort_batch: nsegs out of range");
        if (nsegs == 0) return;
        require(segs != nullptr, "vdf_argsort_batch: null segment list");
        std::vector<SortSeg> ss(nsegs);
        std::vector<uint64_t> ns(nsegs);
        uint64_t total = 0;
        for (int i = 0; i < nsegs; ++i) {
            require(segs[i].n <= VDF_MAX_ROWS, "vdf_argsort_batch: a table exceeds VDF_MAX_ROWS (2^30 - 1)");
            require(segs[i].n == 0 || (segs[i].keys && segs[i].perm_out), "vdf_argsort_batch: null device pointer");
            ss[i] = SortSeg{segs[i].keys, segs[i].n, segs[i].perm_out, segs[i].sorted_keys_out};
            ns[i] = segs[i].n;
            total += segs[i].n;
        }
        if (total == 0) return;
        CallScope cs(h);
        const uint64_t hr = static_cast<uint64_t>(h->tune.hist_rows);
        void* scr = h->ensure_scratch(argsort_batch_scratch(key_type, ns.data(), nsegs, hr));
        h->staging_begin(argsort_batch_staging(key_type, ns.data(), nsegs, hr));
        StagingScope st(h, static_cast<hipStream_t>(stream));
        argsort_batch_enqueue(h, key_type, ss.data(), nsegs, flags, scr, static_cast<hipStream_t>(stream));
    });
}

vdf_status vdf_groupby_batch(vdf_handle* h, int key_type, const vdf_table* tables, int ntables, const int* val_types,
                             int nvals, uint32_t agg_mask, vdf_groupby_out* outs, void* stream)
{
    return guarded([&] {
        require(h != nullptr, "vdf_groupby_batch: null handle");
        require(key_bytes(key_type) != 0, "vdf_groupby_batch: unknown key type");
        require(ntables >= 0 && ntables <= (1 << 22), "vdf_groupby_batch: ntables out of range");
        require(nvals >= 0 && nvals <= VDF_MAX_VALUE_COLS, "vdf_groupby_batch: nvals out of range");
        require((agg_mask & ~VDF_AGG_ALL) == 0, "vdf_groupby_batch: unknown aggregate bits");
        if (ntables == 0) return;
        require(tables != nullptr && outs != nullptr, "vdf_groupby_batch: null table or output list");
        require(nvals == 0 || val_types != nullptr, "vdf_groupby_batch: null value types");
        for (int c = 0; c < nvals; ++c) require(val_bytes(val_types[c]) != 0, "vdf_groupby_batch: unknown value type");
        std::vector<GbTable> ts(ntables);
        std::vector<uint64_t> ns(ntables);
        for (int i = 0; i < ntables; ++i) {
            require(tables[i].n <= VDF_MAX_ROWS, "vdf_groupby_batch: a table exceeds VDF_MAX_ROWS (2^30 - 1)");
            require(outs[i].n_groups != nullptr && outs[i].offsets != nullptr,
                    "vdf_groupby_batch: n_groups and offsets are required");
            require(tables[i].n == 0 || tables[i].keys != nullptr, "vdf_groupby_batch: null keys");
            for (int c = 0; c < nvals; ++c)
                require(tables[i].n == 0 || tables[i].vals[c] != nullptr, "vdf_groupby_batch: null value column");
            ts[i].keys = tables[i].keys;
            ts[i].n = tables[i].n;
            for (int c = 0; c < VDF_MAX_VALUE_COLS; ++c) ts[i].vals[c] = c < nvals ? tables[i].vals[c] : nullptr;
            ns[i] = tables[i].n;
        }
        CallScope cs(h);
        const uint64_t hr = static_cast<uint64_t>(h->tune.hist_rows);
        void* scr = h->ensure_scratch(groupby_batch_scratch(key_type, ns.data(), ntables, hr,
                                                            groupby_carry_bytes(h->tune, key_type, val_types, nvals)));
        h->staging_begin(groupby_batch_staging(key_type, ns.data(), ntables, hr));
        StagingScope st(h, static_cast<hipStream_t>(stream));
        groupby_batch_enqueue(h, key_type, ts.data(), ntables, val_types, nvals, agg_mask, outs, scr,
                              static_cast<hipStream_t>(stream));
    });
}

// Key column types and cell widths of a multi-column key.
struct KeyCols {
    std::vector<int> types;
    std::vector<uint32_t> widths;
    std::vector<const void*> data;
    uint32_t width = 0;  // composite bytes
};

static KeyCols key_cols_of(const vdf_key_col* keys, int nkeys, uint64_t n, const char* who)
{
    require(keys != nullptr, "null key column list");
    require(nkeys >= 1 && nkeys <= VDF_MAX_KEY_COLS, "nkeys out of range (1 .. VDF_MAX_KEY_COLS)");
    KeyCols k;
    for (int c = 0; c < nkeys; ++c) {
        const int t = keys[c].type;
        if (t == VDF_KEY_BYTES)
            require(keys[c].width >= 1 && keys[c].width <= VDF_MAX_STRING_BYTES,
                    "VDF_KEY_BYTES column width must be 1 .. VDF_MAX_STRING_BYTES");
        else
            require(key_bytes(t) != 0, "unknown key type");
        require(n == 0 || keys[c].data != nullptr, "null key column data");
        k.types.push_back(t);
        k.widths.push_back(key_col_width(t, keys[c].width));
        k.data.push_back(keys[c].data);
    }
    k.width = composite_width(k.types.data(), k.widths.data(), nkeys);
    if (k.width > VDF_MAX_STRING_BYTES)
        throw Error{VDF_ERR_UNSUPPORTED, std::string(who) + ": key columns wider than VDF_MAX_STRING_BYTES together"};
    return k;
}

vdf_status vdf_groupby_multi(vdf_handle* h, const vdf_key_col* keys, int nkeys, uint64_t n,
                             const vdf_value_col* vals, int nvals, uint32_t agg_mask, vdf_groupby_out* out,
                             void* const* group_keys, void* stream)
{
    return guarded([&] {
        require(h != nullptr && out != nullptr, "vdf_groupby_multi: null handle or output");
        const KeyCols kc = key_cols_of(keys, nkeys, n, "vdf_groupby_multi");
        require(nvals >= 0 && nvals <= VDF_MAX_VALUE_COLS, "vdf_groupby_multi: nvals out of range");
        require(nvals == 0 || vals != nullptr, "vdf_groupby_multi: null value columns");
        require((agg_mask & ~VDF_AGG_ALL) == 0, "vdf_groupby_multi: unknown aggregate bits");
        require(n <= VDF_MAX_ROWS, "vdf_groupby_multi: n exceeds VDF_MAX_ROWS (2^30 - 1)");
        require(out->n_groups != nullptr && out->offsets != nullptr,
                "vdf_groupby_multi: n_groups and offsets are required");
        std::vector<int> vts(static_cast<size_t>(std::max(nvals, 1)), 0);
        GbTable t{};
        for (int c = 0; c < nvals; ++c) {
            require(val_bytes(vals[c].type) != 0, "vdf_groupby_multi: unknown value type");
            require(n == 0 || vals[c].data != nullptr, "vdf_groupby_multi: null value column data");
            vts[c] = vals[c].type;
            t.vals[c] = vals[c].data;
        }
        CallScope cs(h);
        const hipStream_t s = static_cast<hipStream_t>(stream);
        const int kt = composite_key_type(kc.width);
        const size_t cb = kt >= 0 ? key_bytes(kt) : kc.width;  // composite bytes per row
        const uint64_t cap = (out->capacity == 0 || out->capacity > n) ? n : out->capacity;
        const uint64_t hr = static_cast<uint64_t>(h->tune.hist_rows);
        if (kt < 0 && n == 0) {  // zero groups through the aggregate stage
            void* scr = h->ensure_scratch(groupby_agg_scratch(&n, 1));
            h->staging_begin(groupby_agg_staging(&n, 1));
            StagingScope st(h, s);
            const uint32_t* perms[1] = {nullptr};
            vdf_groupby_out o = *out;
            o.group_keys = nullptr;
            groupby_aggregate_enqueue(h, VDF_KEY_U32, nullptr, perms, &t, 1, vts.data(), nvals, agg_mask, &o, scr, s);
            return;
        }
        const size_t gb = kt >= 0 ? groupby_batch_scratch(kt, &n, 1, hr,
                                                          groupby_carry_bytes(h->tune, kt, vts.data(), nvals))
                                  : bytes_groupby_scratch(kc.width, n, hr);
        void* scr = h->ensure_scratch(align_up(n * cb, 256) + align_up(std::max<uint64_t>(cap, 1) * cb, 256) + gb +
                                      512);
        Carve cv(scr);
        void* comp = cv.take<char>(n * cb);
        void* gkeys = cv.take<char>(std::max<uint64_t>(cap, 1) * cb);
        void* gscr = cv.take<char>(gb);
        pack_keys_enqueue(h, kc.types.data(), kc.widths.data(), kc.data.data(), nkeys, n, comp, s);
        t.keys = comp;
        t.n = n;
        vdf_groupby_out o = *out;
        o.group_keys = gkeys;
        if (kt >= 0) {
            h->staging_begin(groupby_batch_staging(kt, &n, 1, hr));
            StagingScope st(h, s);
            groupby_batch_enqueue(h, kt, &t, 1, vts.data(), nvals, agg_mask, &o, gscr, s);
        } else {  // a wide composite: the byte-string group-by over its rows
            h->staging_begin(bytes_groupby_staging(kc.width, n, hr));
            StagingScope st(h, s);
            bytes_groupby_enqueue(h, comp, kc.width, n, t, vts.data(), nvals, agg_mask, &o, gscr, s);
        }
        if (group_keys)
            unpack_keys_enqueue(h, kc.types.data(), kc.widths.data(), group_keys, nkeys, gkeys, out->n_groups, cap, s);
    });
}

vdf_status vdf_argsort_multi(vdf_handle* h, const vdf_key_col* keys, int nkeys, uint64_t n, uint32_t flags,
                             void* d_perm_out, void* stream)
{
    return guarded([&] {
        require(h != nullptr, "vdf_argsort_multi: null handle");
        const KeyCols kc = key_cols_of(keys, nkeys, n, "vdf_argsort_multi");
        require((flags & ~(VDF_SORT_DESC | VDF_SORT_DESC_REVERSE | VDF_SORT_PERM64)) == 0,
                "vdf_argsort_multi: unknown flag bits");
        require(!((flags & VDF_SORT_DESC) && (flags & VDF_SORT_DESC_REVERSE)),
                "vdf_argsort_multi: VDF_SORT_DESC and VDF_SORT_DESC_REVERSE are exclusive");
        require(n <= VDF_MAX_ROWS, "vdf_argsort_multi: n exceeds VDF_MAX_ROWS (2^30 - 1)");
        if (n == 0) return;
        require(d_perm_out != nullptr, "vdf_argsort_multi: null permutation output");
        CallScope cs(h);
        const hipStream_t s = static_cast<hipStream_t>(stream);
        const int kt = composite_key_type(kc.width);
        const size_t cb = kt >= 0 ? key_bytes(kt) : kc.width;
        const uint64_t hr = static_cast<uint64_t>(h->tune.hist_rows);
        const size_t as = kt >= 0 ? argsort_batch_scratch(kt, &n, 1, hr) : bytes_argsort_scratch(n, hr);
        void* scr = h->ensure_scratch(align_up(n * cb, 256) + as + 512);
        Carve cv(scr);
        void* comp = cv.take<char>(n * cb);
        void* ascr = cv.take<char>(as);
        pack_keys_enqueue(h, kc.types.data(), kc.widths.data(), kc.data.data(), nkeys, n, comp, s);
        if (kt >= 0) {
            h->staging_begin(argsort_batch_staging(kt, &n, 1, hr));
            StagingScope st(h, s);
            argsort_enqueue(h, kt, comp, n, flags, d_perm_out, nullptr, ascr, s);
        } else {
            h->staging_begin(bytes_argsort_staging(kc.width, n, hr));
            StagingScope st(h, s);
            bytes_argsort_enqueue(h, comp, kc.width, n, flags, d_perm_out, nullptr, ascr, s);
        }
    });
}

vdf_status vdf_argsort_bytes(vdf_handle* h, const void* d_keys, uint32_t width, uint64_t n, uint32_t flags,
                             void* d_perm_out, void* d_sorted_out, void* stream)
{
    return guarded([&] {
        require(h != nullptr, "vdf_argsort_bytes: null handle");
        require(width >= 1 && width <= VDF_MAX_STRING_BYTES, "vdf_argsort_bytes: width must be 1 .. VDF_MAX_STRING_BYTES");
        require((flags & ~(VDF_SORT_DESC | VDF_SORT_DESC_REVERSE | VDF_SORT_PERM64)) == 0,
                "vdf_argsort_bytes: unknown flag bits");
        require(!((flags & VDF_SORT_DESC) && (flags & VDF_SORT_DESC_REVERSE)),
                "vdf_argsort_bytes: VDF_SORT_DESC and VDF_SORT_DESC_REVERSE are exclusive");
        require(n <= VDF_MAX_ROWS, "vdf_argsort_bytes: n exceeds VDF_MAX_ROWS (2^30 - 1)");
        if (n == 0) return;
        require(d_keys != nullptr && d_perm_out != nullptr, "vdf_argsort_bytes: null device pointer");
        CallScope cs(h);
        const hipStream_t s = static_cast<hipStream_t>(stream);
        const uint64_t hr = static_cast<uint64_t>(h->tune.hist_rows);
        void* scr = h->ensure_scratch(bytes_argsort_scratch(n, hr));
        h->staging_begin(bytes_argsort_staging(width, n, hr));
        StagingScope st(h, s);
        bytes_argsort_enqueue(h, d_keys, width, n, flags, d_perm_out, d_sorted_out, scr, s);
    });
}

vdf_status vdf_groupby_bytes(vdf_handle* h, const void* d_keys, uint32_t width, uint64_t n,
                             const vdf_value_col* vals, int nvals, uint32_t agg_mask, vdf_groupby_out* out,
                             void* stream)
{
    return guarded([&] {
        require(h != nullptr && out != nullptr, "vdf_groupby_bytes: null handle or output");
        require(width >= 1 && width <= VDF_MAX_STRING_BYTES, "vdf_groupby_bytes: width must be 1 .. VDF_MAX_STRING_BYTES");
        require(nvals >= 0 && nvals <= VDF_MAX_VALUE_COLS, "vdf_groupby_bytes: nvals out of range");
        require(nvals == 0 || vals != nullptr, "vdf_groupby_bytes: null value columns");
        require((agg_mask & ~VDF_AGG_ALL) == 0, "vdf_groupby_bytes: unknown aggregate bits");
        require(out->n_groups != nullptr && out->offsets != nullptr, "vdf_groupby_bytes: n_groups and offsets are required");
        require(n <= VDF_MAX_ROWS, "vdf_groupby_bytes: n exceeds VDF_MAX_ROWS (2^30 - 1)");
        require(n == 0 || d_keys != nullptr, "vdf_groupby_bytes: null keys");
        std::vector<int> vts(static_cast<size_t>(std::max(nvals, 1)), 0);
        GbTable t{};
        t.n = n;
        for (int c = 0; c < nvals; ++c) {
            require(val_bytes(vals[c].type) != 0, "vdf_groupby_bytes: unknown value type");
            require(n == 0 || vals[c].data != nullptr, "vdf_groupby_bytes: null value column data");
            vts[c] = vals[c].type;
            t.vals[c] = vals[c].data;
        }
        CallScope cs(h);
        const hipStream_t s = static_cast<hipStream_t>(stream);
        if (n == 0) {  // zero groups (the aggregate stage writes them for empty tables)
            void* scr = h->ensure_scratch(groupby_agg_scratch(&n, 1));
            h->staging_begin(groupby_agg_staging(&n, 1));
            StagingScope st(h, s);
            const uint32_t* perms[1] = {nullptr};
            vdf_groupby_out o = *out;
            o.group_keys = nullptr;
            groupby_aggregate_enqueue(h, VDF_KEY_U32, nullptr, perms, &t, 1, vts.data(), nvals, agg_mask, &o, scr, s);
            return;
        }
        const uint64_t hr = static_cast<uint64_t>(h->tune.hist_rows);
        void* scr = h->ensure_scratch(bytes_groupby_scratch(width, n, hr));
        h->staging_begin(bytes_groupby_staging(width, n, hr));
        StagingScope st(h, s);
        bytes_groupby_enqueue(h, d_keys, width, n, t, vts.data(), nvals, agg_mask, out, scr, s);
    });
}

vdf_status vdf_broadcast(vdf_handle* h, int src_type, int dst_type, const void* d_group_vals,
                         const uint32_t* d_offsets, const uint64_t* d_n_groups, uint64_t capacity,
                         const uint32_t* d_perm, uint64_t n, void* d_out, void* stream)
{
    return guarded([&] {
        require(h != nullptr, "vdf_broadcast: null handle");
        require(val_bytes(src_type) != 0, "vdf_broadcast: unknown source type");
        require(dst_type == src_type || dst_type == VDF_VAL_F64, "vdf_broadcast: dst_type must be src_type or F64");
        require(n <= VDF_MAX_ROWS, "vdf_broadcast: n exceeds VDF_MAX_ROWS (2^30 - 1)");
        if (n == 0) return;
        require(d_group_vals && d_offsets && d_n_groups && d_perm && d_out, "vdf_broadcast: null device pointer");
        CallScope cs(h);
        broadcast_enqueue(h, src_type, dst_type, d_group_vals, d_offsets, d_n_groups, capacity, d_perm, n, d_out,
                          static_cast<hipStream_t>(stream));
    });
}

}  // extern "C"

// A cached grouping: the sorted keys and the group-contiguous permutation of
// one key column, kept on the device for further aggregations.
struct vdf_grouping {
    vdf_handle* h = nullptr;
    int key_type = 0;
    uint64_t n = 0;
    void* mem = nullptr;
    void* skeys = nullptr;
    uint32_t* perm = nullptr;
};

extern "C" {

vdf_status vdf_grouping_create(vdf_handle* h, int key_type, const void* d_keys, uint64_t n, void* stream,
                               vdf_grouping** out)
{
    return guarded([&] {
        require(h != nullptr && out != nullptr, "vdf_grouping_create: null argument");
        *out = nullptr;
        require(key_bytes(key_type) != 0, "vdf_grouping_create: unknown key type");
        require(n <= VDF_MAX_ROWS, "vdf_grouping_create: n exceeds VDF_MAX_ROWS (2^30 - 1)");
        require(n == 0 || d_keys != nullptr, "vdf_grouping_create: null keys");
        CallScope cs(h);
        const hipStream_t s = static_cast<hipStream_t>(stream);
        const size_t kb = key_bytes(key_type);
        const size_t kbytes = align_up(std::max<uint64_t>(n, 1) * kb, 256);
        auto g = std::make_unique<vdf_grouping>();
        g->h = h;
        g->key_type = key_type;
        g->n = n;
        if (hipMalloc(&g->mem, kbytes + align_up(std::max<uint64_t>(n, 1) * 4, 256)) != hipSuccess) {
            (void)hipGetLastError();
            throw Error{VDF_ERR_OOM, "hipMalloc of a grouping failed"};
        }
        g->skeys = g->mem;
        g->perm = reinterpret_cast<uint32_t*>(static_cast<char*>(g->mem) + kbytes);
        try {
            if (n) {
                const uint64_t hr = static_cast<uint64_t>(h->tune.hist_rows);
                void* scr = h->ensure_scratch(argsort_batch_scratch(key_type, &n, 1, hr));
                h->staging_begin(argsort_batch_staging(key_type, &n, 1, hr));
                StagingScope st(h, s);
                argsort_enqueue(h, key_type, d_keys, n, 0, g->perm, g->skeys, scr, s);
            }
        } catch (...) {
            (void)hipFree(g->mem);
            throw;
        }
        *out = g.release();
    });
}

vdf_status vdf_grouping_aggregate(vdf_grouping* g, const vdf_value_col* vals, int nvals, uint32_t agg_mask,
                                  vdf_groupby_out* out, void* stream)
{
    return guarded([&] {
        require(g != nullptr && out != nullptr, "vdf_grouping_aggregate: null grouping or output");
        require(nvals >= 0 && nvals <= VDF_MAX_VALUE_COLS, "vdf_grouping_aggregate: nvals out of range");
        require(nvals == 0 || vals != nullptr, "vdf_grouping_aggregate: null value columns");
        require((agg_mask & ~VDF_AGG_ALL) == 0, "vdf_grouping_aggregate: unknown aggregate bits");
        require(out->n_groups != nullptr && out->offsets != nullptr,
                "vdf_grouping_aggregate: n_groups and offsets are required");
        const uint64_t n = g->n;
        std::vector<int> vts(static_cast<size_t>(std::max(nvals, 1)), 0);
        GbTable t{};
        t.n = n;
        for (int c = 0; c < nvals; ++c) {
            require(val_bytes(vals[c].type) != 0, "vdf_grouping_aggregate: unknown value type");
            require(n == 0 || vals[c].data != nullptr, "vdf_grouping_aggregate: null value column data");
            vts[c] = vals[c].type;
            t.vals[c] = vals[c].data;
        }
        vdf_handle* h = g->h;
        CallScope cs(h);
        const hipStream_t s = static_cast<hipStream_t>(stream);
        void* scr = h->ensure_scratch(groupby_agg_scratch(&n, 1));
        if (out->perm && n)
            VDF_HIP_CHECK(hipMemcpyAsync(out->perm, g->perm, n * 4, hipMemcpyDeviceToDevice, s));
        h->staging_begin(groupby_agg_staging(&n, 1));
        StagingScope st(h, s);
        const uint32_t* perms[1] = {g->perm};
        groupby_aggregate_enqueue(h, g->key_type, g->skeys, perms, &t, 1, vts.data(), nvals, agg_mask, out, scr, s);
    });
}

vdf_status vdf_grouping_destroy(vdf_grouping* g)
{
    return guarded([&] {
        if (!g) return;
        if (g->mem) (void)hipFree(g->mem);
        delete g;
    });
}

static int* tune_slot(Tuning& t, const std::string& n)
{
    if (n == "sort_items_32") return &t.sort_items_32;
    if (n == "sort_items_64") return &t.sort_items_64;
    if (n == "sort_items_128") return &t.sort_items_128;
    if (n == "sort_chunk") return &t.sort_chunk;
    if (n == "sort_xcd") return &t.sort_xcd;
    if (n == "sort_blocks_per_cu") return &t.sort_bpc;
    if (n == "hist_rows") return &t.hist_rows;
    if (n == "gather_rows") return &t.gather_rows;
    if (n == "gather_nt") return &t.gather_nt;
    if (n == "csv_stage") return &t.csv_stage;
    if (n == "sort_local") return &t.sort_local;
    if (n == "sort_rank") return &t.sort_rank;
    if (n == "groupby_carry") return &t.gb_carry;
    return nullptr;
}

vdf_status vdf_tune_set(vdf_handle* h, const char* name, int64_t value)
{
    return guarded([&] {
        require(h != nullptr && name != nullptr, "vdf_tune_set: null argument");
        std::lock_guard<std::mutex> lk(h->mu);
        int* slot = tune_slot(h->tune, name);
        require(slot != nullptr, "vdf_tune_set: unknown tuning knob");
        require(tune_supported(name, value), "vdf_tune_set: value not supported for this knob");
        *slot = static_cast<int>(value);
    });
}

vdf_status vdf_tune_get(vdf_handle* h, const char* name, int64_t* value)
{
    return guarded([&] {
        require(h != nullptr && name != nullptr && value != nullptr, "vdf_tune_get: null argument");
        std::lock_guard<std::mutex> lk(h->mu);
        if (std::string(name) == "sort_rank_atomic") {  // read-only: the ranking the sorts use
            *value = (h->tune.sort_rank && lds_rank_verified(h->device)) ? 1 : 0;
            return;
        }
        int* slot = tune_slot(h->tune, name);
        require(slot != nullptr, "vdf_tune_get: unknown tuning knob");
        *value = *slot;
    });
}

vdf_status vdf_profile_enable(vdf_handle* h, int on)
{
    return guarded([&] {
        require(h != nullptr, "vdf_profile_enable: null handle");
        CallScope cs(h);
        if (on && !h->passes_ring) {
            void* p = nullptr;
            VDF_HIP_CHECK(hipHostMalloc(&p, Handle::kRing * sizeof(uint32_t), hipHostMallocDefault));
            h->passes_ring = static_cast<uint32_t*>(p);
        }
        h->prof = on != 0;
    });
}

vdf_status vdf_profile_reset(vdf_handle* h)
{
    return guarded([&] {
        require(h != nullptr, "vdf_profile_reset: null handle");
        CallScope cs(h);
        h->prof_collect();
        h->totals.clear();
    });
}

vdf_status vdf_profile_read(vdf_handle* h, const char* kernel, double* total_ms, uint64_t* launches, uint64_t* bytes)
{
    return guarded([&] {
        require(h != nullptr && kernel != nullptr, "vdf_profile_read: null argument");
        CallScope cs(h);
        h->prof_collect();
        double ms = 0;
        uint64_t l = 0, b = 0;
        for (const ProfTotal& t : h->totals)
            if (t.name == kernel) {
                ms = t.ms;
                l = t.launches;
                b = t.bytes;
            }
        if (total_ms) *total_ms = ms;
        if (launches) *launches = l;
        if (bytes) *bytes = b;
    });
}

}  // extern "C"
