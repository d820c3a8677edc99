// plan.hip -- HIP-graph-captured launch sequences of a batched group-by.
//
// A plan is captured once for a batch shape (the tables' row counts, key and
// value types, aggregates) and replayed with any tables and outputs of that
// shape.  Every kernel of the sequence reads the batch description (table
// pointers, outputs, tile maps) from device memory that one host-to-device
// copy fills from a pinned buffer, and the copy is the first node of the
// graph: a replay only rewrites the pinned buffer on the host (a "dry"
// enqueue that builds the description and launches nothing) and launches the
// graph.  The plan owns its scratch, so a plan replays independently of the
// handle's other work; its radix passes are plain launches whose tile claiming
// needs no co-residency beside other work.  Replays of
// one plan share its scratch: launch them on one stream (or order them).  Two
// pinned buffers, each captured into its own graph,
// let the host prepare the next replay while the previous one runs.
//
// No reference counterpart: VisonDF issues no device work; this is the
// launch-shape side of batching (many tables per launch sequence).
#include "vdf_internal.h"

#include <memory>

namespace vdf {

struct Plan {
    Handle ph;  // private handle: tuning snapshot, staging override, dry flag
    int key_type = 0;
    int nvals = 0;
    int val_types[VDF_MAX_VALUE_COLS] = {0, 0, 0, 0};
    uint32_t agg_mask = 0;
    std::vector<uint64_t> ns;
    size_t staging_bytes = 0;
    void* scratch = nullptr;
    size_t scratch_bytes = 0;
    // graph[v]: v = 1 moves row ids through the sort (a replay that asks for a
    // permutation), v = 0 does not (the row-id-free kernels; null when the
    // value columns need row ids anyway)
    struct Buf {
        void* stg = nullptr;
        hipGraph_t graph[2] = {nullptr, nullptr};
        hipGraphExec_t exec[2] = {nullptr, nullptr};
        hipEvent_t done = nullptr;
        bool pending = false;
    } buf[2];
    int next = 0;

    void release()
    {
        for (Buf& b : buf) {
            if (b.pending && b.done) (void)hipEventSynchronize(b.done);
            for (int v = 0; v < 2; ++v) {
                if (b.exec[v]) (void)hipGraphExecDestroy(b.exec[v]);
                if (b.graph[v]) (void)hipGraphDestroy(b.graph[v]);
            }
            if (b.done) (void)hipEventDestroy(b.done);
            if (b.stg) (void)hipHostFree(b.stg);
            b = Buf{};
        }
        if (scratch) (void)hipFree(scratch);
        scratch = nullptr;
        ph.release_resources();
    }
};

namespace {

std::vector<GbTable> to_tables(const vdf_table* t, int n, int nvals)
{
    std::vector<GbTable> ts(static_cast<size_t>(n));
    for (int i = 0; i < n; ++i) {
        ts[i].keys = t[i].keys;
        ts[i].n = t[i].n;
        for (int c = 0; c < VDF_MAX_VALUE_COLS; ++c) ts[i].vals[c] = c < nvals ? t[i].vals[c] : nullptr;
    }
    return ts;
}

// Captures buffer k's graph of variant v (row ids moved or not): the enqueue
// runs on a capturing stream with the plan's scratch and buffer k as staging.
void capture(Plan& p, int k, int v, const std::vector<GbTable>& ts, const std::vector<vdf_groupby_out>& outs)
{
    Plan::Buf& b = p.buf[k];
    p.ph.keep_rowids = v == 1;
    p.ph.staging_fixed = b.stg;
    p.ph.staging_fixed_cap = p.staging_bytes;
    hipStream_t cs = nullptr;
    VDF_HIP_CHECK(hipStreamCreateWithFlags(&cs, hipStreamNonBlocking));
    hipError_t e = hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal);
    if (e != hipSuccess) {
        (void)hipStreamDestroy(cs);
        throw Error{VDF_ERR_HIP, std::string("hipStreamBeginCapture: ") + hipGetErrorString(e)};
    }
    std::string err;
    try {
        p.ph.staging_begin(p.staging_bytes);
        groupby_batch_enqueue(&p.ph, p.key_type, ts.data(), static_cast<int>(ts.size()), p.val_types, p.nvals,
                              p.agg_mask, outs.data(), p.scratch, cs);
        p.ph.staging_end(cs);
    } catch (const Error& x) {
        err = x.msg;
    }
    hipGraph_t g = nullptr;
    e = hipStreamEndCapture(cs, &g);
    (void)hipStreamDestroy(cs);
    if (!err.empty() || e != hipSuccess) {
        if (g) (void)hipGraphDestroy(g);
        (void)hipGetLastError();
        throw Error{VDF_ERR_HIP, err.empty() ? std::string("hipStreamEndCapture: ") + hipGetErrorString(e) : err};
    }
    b.graph[v] = g;
    e = hipGraphInstantiate(&b.exec[v], g, nullptr, nullptr, 0);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        throw Error{VDF_ERR_HIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(e)};
    }
}

}  // namespace
}  // namespace vdf

using namespace vdf;

struct vdf_plan : Plan {};

extern "C" {

vdf_status vdf_plan_groupby_batch(vdf_handle* h, int key_type, const vdf_table* tables, int ntables,
                                  const int* val_types, int nvals, uint32_t agg_mask, vdf_plan** out)
{
    return guarded([&] {
        require(h != nullptr && out != nullptr, "vdf_plan_groupby_batch: null argument");
        *out = nullptr;
        require(key_bytes(key_type) != 0, "vdf_plan_groupby_batch: unknown key type");
        require(ntables >= 1 && ntables <= (1 << 22), "vdf_plan_groupby_batch: ntables out of range");
        require(tables != nullptr, "vdf_plan_groupby_batch: null table list");
        require(nvals >= 0 && nvals <= VDF_MAX_VALUE_COLS, "vdf_plan_groupby_batch: nvals out of range");
        require(nvals == 0 || val_types != nullptr, "vdf_plan_groupby_batch: null value types");
        require((agg_mask & ~VDF_AGG_ALL) == 0, "vdf_plan_groupby_batch: unknown aggregate bits");
        for (int c = 0; c < nvals; ++c)
            require(val_bytes(val_types[c]) != 0, "vdf_plan_groupby_batch: unknown value type");
        auto p = std::make_unique<vdf_plan>();
        int prev = 0;
        VDF_HIP_CHECK(hipGetDevice(&prev));
        {
            std::lock_guard<std::mutex> lk(h->mu);
            p->ph.device = h->device;
            p->ph.tune = h->tune;
        }
        VDF_HIP_CHECK(hipSetDevice(p->ph.device));
        p->key_type = key_type;
        p->nvals = nvals;
        for (int c = 0; c < nvals; ++c) p->val_types[c] = val_types[c];
        p->agg_mask = agg_mask;
        p->ns.resize(static_cast<size_t>(ntables));
        for (int i = 0; i < ntables; ++i) {
            require(tables[i].n <= VDF_MAX_ROWS, "vdf_plan_groupby_batch: a table exceeds VDF_MAX_ROWS (2^30 - 1)");
            p->ns[i] = tables[i].n;
        }
        const uint64_t hr = static_cast<uint64_t>(p->ph.tune.hist_rows);
        p->scratch_bytes = groupby_batch_scratch(key_type, p->ns.data(), ntables, hr,
                                                 groupby_carry_bytes(p->ph.tune, key_type, p->val_types, p->nvals));
        p->staging_bytes = groupby_batch_staging(key_type, p->ns.data(), ntables, hr);
        try {
            if (hipMalloc(&p->scratch, p->scratch_bytes) != hipSuccess) {
                (void)hipGetLastError();
                p->scratch = nullptr;
                throw Error{VDF_ERR_OOM, "hipMalloc of " + std::to_string(p->scratch_bytes) + " plan scratch bytes failed"};
            }
            // description of the capture run: the tables given, outputs unset (every
            // replay rewrites the description before launching)
            const std::vector<GbTable> ts = to_tables(tables, ntables, nvals);
            std::vector<vdf_groupby_out> outs(static_cast<size_t>(ntables));
            for (int k = 0; k < 2; ++k) {
                VDF_HIP_CHECK(hipHostMalloc(&p->buf[k].stg, p->staging_bytes, hipHostMallocDefault));
                VDF_HIP_CHECK(hipEventCreateWithFlags(&p->buf[k].done, hipEventDisableTiming));
            }
            // value columns beyond the carried ones are gathered through row ids,
            // so such a plan moves them in either case and needs one variant
            const int pay = groupby_carry_bytes(p->ph.tune, key_type, p->val_types, nvals);
            const bool rows_always = nvals > (pay >= 12 ? 2 : (pay != 0 ? 1 : 0));
            for (int k = 0; k < 2; ++k) {
                capture(*p, k, 1, ts, outs);
                if (!rows_always) capture(*p, k, 0, ts, outs);
            }
        } catch (...) {
            p->release();
            (void)hipSetDevice(prev);
            throw;
        }
        (void)hipSetDevice(prev);
        *out = p.release();
    });
}

vdf_status vdf_plan_launch(vdf_plan* p, const vdf_table* tables, int ntables, vdf_groupby_out* outs, void* stream)
{
    return guarded([&] {
        require(p != nullptr, "vdf_plan_launch: null plan");
        require(ntables == static_cast<int>(p->ns.size()), "vdf_plan_launch: table count differs from the plan");
        require(tables != nullptr && outs != nullptr, "vdf_plan_launch: null table or output list");
        for (int i = 0; i < ntables; ++i) {
            require(tables[i].n == p->ns[i], "vdf_plan_launch: a table's row count differs from the plan");
            require(outs[i].n_groups != nullptr && outs[i].offsets != nullptr,
                    "vdf_plan_launch: n_groups and offsets are required");
            require(tables[i].n == 0 || tables[i].keys != nullptr, "vdf_plan_launch: null keys");
            for (int c = 0; c < p->nvals; ++c)
                require(tables[i].n == 0 || tables[i].vals[c] != nullptr, "vdf_plan_launch: null value column");
        }
        std::lock_guard<std::mutex> lk(p->ph.mu);
        int prev = 0;
        VDF_HIP_CHECK(hipGetDevice(&prev));
        VDF_HIP_CHECK(hipSetDevice(p->ph.device));
        Plan::Buf& b = p->buf[p->next];
        p->next ^= 1;
        bool want_perm = false;
        for (int i = 0; i < ntables; ++i) want_perm = want_perm || outs[i].perm != nullptr;
        const int v = (want_perm || !b.exec[0]) ? 1 : 0;
        try {
            if (b.pending) {
                VDF_HIP_CHECK(hipEventSynchronize(b.done));  // this buffer's previous upload is done
                b.pending = false;
            }
            // rebuild the batch description in the buffer the graph uploads from
            const std::vector<GbTable> ts = to_tables(tables, ntables, p->nvals);
            p->ph.keep_rowids = v == 1;
            p->ph.staging_fixed = b.stg;
            p->ph.staging_fixed_cap = p->staging_bytes;
            p->ph.dry = true;
            try {
                p->ph.staging_begin(p->staging_bytes);
                groupby_batch_enqueue(&p->ph, p->key_type, ts.data(), ntables, p->val_types, p->nvals, p->agg_mask,
                                      outs, p->scratch, nullptr);
                p->ph.staging_end(nullptr);
            } catch (...) {
                p->ph.dry = false;
                throw;
            }
            p->ph.dry = false;
            const hipStream_t s = static_cast<hipStream_t>(stream);
            VDF_HIP_CHECK(hipGraphLaunch(b.exec[v], s));
            VDF_HIP_CHECK(hipEventRecord(b.done, s));
            b.pending = true;
        } catch (...) {
            (void)hipSetDevice(prev);
            throw;
        }
        (void)hipSetDevice(prev);
    });
}

vdf_status vdf_plan_scratch_bytes(vdf_plan* p, uint64_t* bytes)
{
    return guarded([&] {
        require(p != nullptr && bytes != nullptr, "vdf_plan_scratch_bytes: null argument");
        *bytes = p->scratch_bytes;
    });
}

vdf_status vdf_plan_destroy(vdf_plan* p)
{
    return guarded([&] {
        if (!p) return;
        int prev = 0;
        VDF_HIP_CHECK(hipGetDevice(&prev));
        VDF_HIP_CHECK(hipSetDevice(p->ph.device));
        p->release();
        (void)hipSetDevice(prev);
        delete p;
    });
}

}  // extern "C"
