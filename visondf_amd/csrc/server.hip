// server.hip -- group-by request server: device-resident tables, one HBM
// arena as the admission budget, and continuous batching of queued requests
// into segmented group-by launch sequences run by several workers (one host
// thread and one stream each), so batches overlap on the device.
//
// No reference counterpart: VisonDF runs each group-by on the calling thread
// (transform_group_by.hpp:11-14).  Request semantics are those of
// vdf_groupby on the request's table; results are bit-identical to it because
// a table's result does not depend on the batch it runs in.
#include "serve_core.h"
#include "vdf_internal.h"

#include <chrono>
#include <condition_variable>
#include <deque>
#include <memory>
#include <thread>
#include <unordered_map>

namespace vdf {
namespace {

uint64_t now_ns()
{
    return static_cast<uint64_t>(std::chrono::duration_cast<std::chrono::nanoseconds>(
                                     std::chrono::steady_clock::now().time_since_epoch())
                                     .count());
}

struct Table : TableShape {
    const void* keys;
    const void* vals[VDF_MAX_VALUE_COLS];
    uint64_t max_groups;
};

struct Request {
    uint64_t id = 0;
    uint32_t table = 0;
    uint32_t agg_mask = 0;
    bool want_perm = false;
    uint64_t cap = 0;        // groups the result block holds
    uint64_t sig = 0;        // batch signature (key type, value types, aggregates)
    uint64_t out_bytes = 0;  // result block bytes
    uint64_t alone_scratch = 0;  // scratch of a batch of this request alone
    uint64_t out_off = RangeAllocator::kNone;
    uint64_t submit_ns = 0, start_ns = 0, done_ns = 0;
    uint64_t n_groups = 0;
    uint32_t batch_tables = 0;
    bool done = false;
    bool retried = false;
    vdf_status status = VDF_OK;
    std::string msg;
    vdf_groupby_out out{};
};

// bytes per row the default batch row limit assumes for one batch's scratch
constexpr uint64_t kDefaultRowBytes = 96;

struct Worker {
    std::unique_ptr<Handle> h;
    hipStream_t stream = nullptr;
    uint64_t* d_ng = nullptr;  // device: group counts of the running batch
    uint64_t* h_ng = nullptr;  // pinned host copy
};

}  // namespace

struct Server {
    int device = 0;
    PlanLimits lim;
    size_t scan_window = 4096;  // queued requests the planner looks at
    char* arena_base = nullptr;
    RangeAllocator arena;
    std::vector<Table> tables;
    std::vector<Worker> workers;
    std::vector<std::thread> threads;

    std::mutex mu;
    std::condition_variable cv_work;  // queue or arena changed
    std::condition_variable cv_done;  // a request finished
    std::deque<Request*> queue;
    std::unordered_map<uint64_t, std::unique_ptr<Request>> reqs;
    uint64_t next_id = 1;
    int inflight = 0;
    bool stop = false;
    vdf_server_stats st{};
    Tuning tune;  // the workers' launch shape (all workers share it)

    // scratch of a batch that holds `t` alone, as the workers size it
    uint64_t alone_scratch(const Table& t) const
    {
        return static_cast<uint64_t>(groupby_batch_scratch(t.key_type, &t.n, 1, static_cast<uint64_t>(tune.hist_rows),
                                                           groupby_carry_bytes(tune, t.key_type, t.val_types,
                                                                               t.nvals)));
    }

    void finish(Request* r, vdf_status s, const std::string& m, uint64_t t)
    {
        r->status = s;
        r->msg = m;
        r->done = true;
        r->done_ns = t;
        if (s == VDF_OK) {
            ++st.completed;
            st.rows += tables[r->table].n;
        } else {
            ++st.failed;
            if (r->out_off != RangeAllocator::kNone) arena.release(r->out_off);
            r->out_off = RangeAllocator::kNone;
            r->out = vdf_groupby_out{};
        }
    }

    void run(int w);
    void teardown();
};

// One worker: take the next batch the arena admits, run it on this worker's
// stream, publish the results, repeat.
void Server::run(int w)
{
    (void)hipSetDevice(device);
    Worker& W = workers[w];
    Handle& h = *W.h;
    const uint64_t hist_rows = static_cast<uint64_t>(h.tune.hist_rows);
    for (;;) {
        std::vector<Request*> batch;
        Plan plan;
        // copies of the batch's table descriptions (the registry may grow meanwhile)
        std::vector<GbTable> ts;
        std::vector<vdf_groupby_out> outs;
        std::vector<uint64_t> ns;
        Table t0{};
        {
            std::unique_lock<std::mutex> lk(mu);
            for (;;) {
                if (!queue.empty()) {
                    const Table& ht = tables[queue.front()->table];
                    std::vector<PlanItem> items;
                    const size_t m = std::min(queue.size(), scan_window);
                    items.reserve(m);
                    for (size_t i = 0; i < m; ++i) {
                        const Request* r = queue[i];
                        items.push_back(PlanItem{tables[r->table].n, r->sig, r->out_bytes, r->alone_scratch});
                    }
                    plan = plan_batch(items, lim, arena, [&](const std::vector<uint64_t>& ns) {
                        return static_cast<uint64_t>(
                            groupby_batch_scratch(ht.key_type, ns.data(), static_cast<int>(ns.size()), hist_rows,
                                                  groupby_carry_bytes(h.tune, ht.key_type, ht.val_types, ht.nvals)));
                    });
                    if (plan.head_never_fits) {
                        Request* r = queue.front();
                        queue.pop_front();
                        finish(r, VDF_ERR_BUDGET,
                               "request needs more than the server arena of " + std::to_string(arena.size()) +
                                   " bytes",
                               now_ns());
                        cv_done.notify_all();
                        continue;
                    }
                    if (!plan.picked.empty()) break;
                    if (stop && inflight == 0) {
                        // nothing runs and nothing fits: results held by the caller block the queue
                        const uint64_t t = now_ns();
                        while (!queue.empty()) {
                            finish(queue.front(), VDF_ERR_BUDGET, "server stopped before the request fit", t);
                            queue.pop_front();
                        }
                        cv_done.notify_all();
                        continue;
                    }
                }
                if (stop && queue.empty()) return;
                cv_work.wait(lk);
            }
            batch.resize(plan.picked.size());
            for (size_t k = plan.picked.size(); k-- > 0;) {
                batch[k] = queue[plan.picked[k]];
                queue.erase(queue.begin() + static_cast<std::ptrdiff_t>(plan.picked[k]));
            }
            const uint64_t t = now_ns();
            const size_t B = batch.size();
            ts.resize(B);
            outs.resize(B);
            ns.resize(B);
            t0 = tables[batch[0]->table];
            for (size_t k = 0; k < B; ++k) {
                Request* r = batch[k];
                const Table& tb = tables[r->table];
                r->out_off = plan.out_off[k];
                r->start_ns = t;
                r->batch_tables = static_cast<uint32_t>(B);
                result_layout(tb, r->agg_mask, r->want_perm, r->cap, arena_base + r->out_off, &r->out);
                ts[k].keys = tb.keys;
                ts[k].n = tb.n;
                for (int c = 0; c < VDF_MAX_VALUE_COLS; ++c) ts[k].vals[c] = c < tb.nvals ? tb.vals[c] : nullptr;
                ns[k] = tb.n;
                outs[k] = r->out;
                outs[k].n_groups = W.d_ng + k;
            }
            ++inflight;
            ++st.launches;
        }

        // ---- run the batch (no lock held) ----
        const int B = static_cast<int>(batch.size());
        vdf_status status = VDF_OK;
        std::string msg;
        try {
            h.staging_begin(groupby_batch_staging(t0.key_type, ns.data(), B, hist_rows));
            try {
                groupby_batch_enqueue(&h, t0.key_type, ts.data(), B, t0.val_types, t0.nvals, batch[0]->agg_mask,
                                      outs.data(), arena_base + plan.scr_off, W.stream);
            } catch (...) {
                h.staging_end(W.stream);
                throw;
            }
            h.staging_end(W.stream);
            VDF_HIP_CHECK(hipMemcpyAsync(W.h_ng, W.d_ng, static_cast<size_t>(B) * 8, hipMemcpyDeviceToHost, W.stream));
            VDF_HIP_CHECK(hipStreamSynchronize(W.stream));
        } catch (const Error& e) {
            status = e.code;
            msg = e.msg;
        } catch (const std::exception& e) {
            status = VDF_ERR_INVALID;
            msg = e.what();
        }
        const uint64_t t_done = now_ns();

        // ---- publish ----
        {
            std::lock_guard<std::mutex> lk(mu);
            arena.release(plan.scr_off);
            for (int i = B; i-- > 0;) {
                Request* r = batch[i];
                if (status != VDF_OK) {
                    finish(r, status, msg, t_done);
                    continue;
                }
                r->n_groups = W.h_ng[i];
                if (r->n_groups > r->cap && !r->retried) {
                    // more groups than the table's max_groups: rerun once with the exact count
                    arena.release(r->out_off);
                    r->out_off = RangeAllocator::kNone;
                    r->retried = true;
                    r->cap = r->n_groups;
                    r->out_bytes = result_layout(tables[r->table], r->agg_mask, r->want_perm, r->cap, nullptr, nullptr);
                    ++st.retries;
                    queue.push_front(r);  // batch walked newest first: the oldest ends up in front
                } else if (r->n_groups > r->cap) {
                    finish(r, VDF_ERR_INVALID, "group count exceeds the result capacity after a rerun", t_done);
                } else {
                    finish(r, VDF_OK, std::string(), t_done);
                }
            }
            --inflight;
        }
        cv_work.notify_all();
        cv_done.notify_all();
    }
}

void Server::teardown()
{
    for (Worker& W : workers) {
        if (W.stream) (void)hipStreamSynchronize(W.stream);
        if (W.h) W.h->release_resources();
        if (W.stream) (void)hipStreamDestroy(W.stream);
        if (W.d_ng) (void)hipFree(W.d_ng);
        if (W.h_ng) (void)hipHostFree(W.h_ng);
        W = Worker{};
    }
    workers.clear();
    if (arena_base) (void)hipFree(arena_base);
    arena_base = nullptr;
}

}  // namespace vdf

using namespace vdf;

struct vdf_server : Server {};

extern "C" {

vdf_status vdf_server_create(vdf_server** out, const vdf_server_config* cfg)
{
    return guarded([&] {
        require(out != nullptr && cfg != nullptr, "vdf_server_create: null argument");
        *out = nullptr;
        require(cfg->workers >= 1 && cfg->workers <= 16, "vdf_server_create: workers must be 1 .. 16");
        require(cfg->max_batch_tables >= 1 && cfg->max_batch_tables <= 65536,
                "vdf_server_create: max_batch_tables must be 1 .. 65536");
        require(cfg->arena_bytes >= (1ull << 20), "vdf_server_create: arena_bytes must be at least 1 MiB");
        int ndev = 0;
        VDF_HIP_CHECK(hipGetDeviceCount(&ndev));
        require(cfg->device >= 0 && cfg->device < ndev, "vdf_server_create: no such device");
        int prev = 0;
        VDF_HIP_CHECK(hipGetDevice(&prev));
        VDF_HIP_CHECK(hipSetDevice(cfg->device));
        (void)lds_rank_check(cfg->device);
        auto S = std::make_unique<vdf_server>();
        S->device = cfg->device;
        S->lim.max_tables = static_cast<size_t>(cfg->max_batch_tables);
        // default row limit: a batch's scratch (up to ~90 bytes per row for
        // 16-byte keys with a carried f64 column) stays within half of the
        // arena's share per worker, so several batches and held results fit
        S->lim.max_rows = cfg->max_batch_rows
                              ? cfg->max_batch_rows
                              : std::max<uint64_t>(cfg->arena_bytes / (kDefaultRowBytes * 2 * cfg->workers), 1ull << 20);
        // Global tile order: batches of several workers run at once, so a pass
        // kernel may hold only a few blocks on an XCD, and a chunk of
        // consecutive tiles claimed there would hold up every later tile's
        // look-back while those few blocks work through it.
        S->tune.sort_xcd = 0;
        try {
            void* p = nullptr;
            if (hipMalloc(&p, cfg->arena_bytes) != hipSuccess) {
                (void)hipGetLastError();
                throw Error{VDF_ERR_OOM, "hipMalloc of the " + std::to_string(cfg->arena_bytes) +
                                             "-byte server arena failed"};
            }
            S->arena_base = static_cast<char*>(p);
            S->arena.reset(cfg->arena_bytes);
            // staging for a full batch of large tables, so workers never regrow it mid-run
            const int nt = cfg->max_batch_tables;
            const uint64_t rows = std::min<uint64_t>(S->lim.max_rows, 1ull << 31);
            std::vector<uint64_t> ns(static_cast<size_t>(nt), std::max<uint64_t>(rows / nt, 1));
            S->workers.resize(static_cast<size_t>(cfg->workers));
            for (Worker& W : S->workers) {
                W.h = std::make_unique<Handle>();
                W.h->device = cfg->device;
                W.h->tune = S->tune;
                W.h->rowfree_kernels = false;  // the kernels the served workload was validated with
                VDF_HIP_CHECK(hipStreamCreateWithFlags(&W.stream, hipStreamNonBlocking));
                VDF_HIP_CHECK(hipMalloc(&W.d_ng, static_cast<size_t>(nt) * 8));
                VDF_HIP_CHECK(hipHostMalloc(&W.h_ng, static_cast<size_t>(nt) * 8, hipHostMallocDefault));
                const size_t sb = groupby_batch_staging(VDF_KEY_CHAR16, ns.data(), nt,
                                                        static_cast<uint64_t>(W.h->tune.hist_rows));
                for (int k = 0; k < Handle::kStaging; ++k) {
                    W.h->staging_begin(sb);
                    W.h->staging_end(W.stream);
                }
            }
            VDF_HIP_CHECK(hipDeviceSynchronize());
            for (int w = 0; w < cfg->workers; ++w) S->threads.emplace_back([s = S.get(), w] { s->run(w); });
        } catch (...) {
            S->teardown();
            (void)hipSetDevice(prev);
            throw;
        }
        (void)hipSetDevice(prev);
        *out = S.release();
    });
}

vdf_status vdf_server_destroy(vdf_server* s)
{
    return guarded([&] {
        if (!s) return;
        {
            std::lock_guard<std::mutex> lk(s->mu);
            s->stop = true;
        }
        s->cv_work.notify_all();
        for (std::thread& t : s->threads) t.join();
        s->threads.clear();
        int prev = 0;
        VDF_HIP_CHECK(hipGetDevice(&prev));
        VDF_HIP_CHECK(hipSetDevice(s->device));
        s->teardown();
        (void)hipSetDevice(prev);
        delete s;
    });
}

vdf_status vdf_server_add_table(vdf_server* s, int key_type, const void* d_keys, uint64_t n, const int* val_types,
                                const void* const* d_vals, int nvals, uint64_t max_groups, uint32_t* table_id)
{
    return guarded([&] {
        require(s != nullptr && table_id != nullptr, "vdf_server_add_table: null argument");
        require(key_bytes(key_type) != 0, "vdf_server_add_table: unknown key type");
        require(n <= VDF_MAX_ROWS, "vdf_server_add_table: n exceeds VDF_MAX_ROWS (2^30 - 1)");
        require(n == 0 || d_keys != nullptr, "vdf_server_add_table: null keys");
        require(nvals >= 0 && nvals <= VDF_MAX_VALUE_COLS, "vdf_server_add_table: nvals out of range");
        require(nvals == 0 || (val_types != nullptr && d_vals != nullptr), "vdf_server_add_table: null value list");
        Table t{};
        t.key_type = key_type;
        t.key_bytes = static_cast<uint32_t>(key_bytes(key_type));
        t.keys = d_keys;
        t.n = n;
        t.nvals = nvals;
        for (int c = 0; c < nvals; ++c) {
            require(val_bytes(val_types[c]) != 0, "vdf_server_add_table: unknown value type");
            require(n == 0 || d_vals[c] != nullptr, "vdf_server_add_table: null value column");
            t.val_types[c] = val_types[c];
            t.val_bytes[c] = static_cast<uint32_t>(val_bytes(val_types[c]));
            t.vals[c] = d_vals[c];
        }
        t.max_groups = max_groups;
        std::lock_guard<std::mutex> lk(s->mu);
        require(s->tables.size() < (1ull << 31), "vdf_server_add_table: too many tables");
        *table_id = static_cast<uint32_t>(s->tables.size());
        s->tables.push_back(t);
    });
}

vdf_status vdf_server_submit(vdf_server* s, uint32_t table_id, uint32_t agg_mask, int want_perm,
                             uint64_t* request_id)
{
    return guarded([&] {
        require(s != nullptr && request_id != nullptr, "vdf_server_submit: null argument");
        require((agg_mask & ~VDF_AGG_ALL) == 0, "vdf_server_submit: unknown aggregate bits");
        auto r = std::make_unique<Request>();
        r->submit_ns = now_ns();
        {
            std::lock_guard<std::mutex> lk(s->mu);
            require(!s->stop, "vdf_server_submit: server is stopping");
            require(table_id < s->tables.size(), "vdf_server_submit: unknown table id");
            const Table& t = s->tables[table_id];
            r->id = s->next_id++;
            r->table = table_id;
            r->agg_mask = agg_mask;
            r->want_perm = want_perm != 0;
            r->cap = (t.max_groups == 0 || t.max_groups > t.n) ? t.n : t.max_groups;
            r->sig = signature(t, agg_mask);
            r->out_bytes = result_layout(t, agg_mask, r->want_perm, r->cap, nullptr, nullptr);
            r->alone_scratch = s->alone_scratch(t);
            *request_id = r->id;
            s->queue.push_back(r.get());
            s->reqs.emplace(r->id, std::move(r));
            ++s->st.submitted;
        }
        s->cv_work.notify_one();
    });
}

vdf_status vdf_server_wait(vdf_server* s, uint64_t request_id, int64_t timeout_us, vdf_served* out)
{
    return guarded([&] {
        require(s != nullptr && out != nullptr, "vdf_server_wait: null argument");
        std::unique_lock<std::mutex> lk(s->mu);
        auto it = s->reqs.find(request_id);
        require(it != s->reqs.end(), "vdf_server_wait: unknown request id");
        Request* r = it->second.get();
        auto ready = [r] { return r->done; };
        if (timeout_us < 0) {
            s->cv_done.wait(lk, ready);
        } else if (!s->cv_done.wait_for(lk, std::chrono::microseconds(timeout_us), ready)) {
            throw Error{VDF_ERR_TIMEOUT, "vdf_server_wait: request not done"};
        }
        vdf_served v{};
        v.status = r->status;
        v.batch_tables = r->batch_tables;
        v.n_groups = r->n_groups;
        v.submit_ns = r->submit_ns;
        v.start_ns = r->start_ns;
        v.done_ns = r->done_ns;
        v.out = r->out;
        v.out.n_groups = nullptr;
        *out = v;
        if (r->status != VDF_OK) set_error(r->msg);
    });
}

vdf_status vdf_server_release(vdf_server* s, uint64_t request_id)
{
    return guarded([&] {
        require(s != nullptr, "vdf_server_release: null server");
        {
            std::lock_guard<std::mutex> lk(s->mu);
            auto it = s->reqs.find(request_id);
            require(it != s->reqs.end(), "vdf_server_release: unknown request id");
            require(it->second->done, "vdf_server_release: request still queued or running");
            if (it->second->out_off != RangeAllocator::kNone) s->arena.release(it->second->out_off);
            s->reqs.erase(it);
        }
        s->cv_work.notify_all();
    });
}

vdf_status vdf_server_stats_get(vdf_server* s, vdf_server_stats* out)
{
    return guarded([&] {
        require(s != nullptr && out != nullptr, "vdf_server_stats_get: null argument");
        std::lock_guard<std::mutex> lk(s->mu);
        vdf_server_stats v = s->st;
        v.queued = s->queue.size();
        v.arena_bytes = s->arena.size();
        v.arena_used = s->arena.used();
        v.arena_peak = s->arena.peak();
        *out = v;
    });
}

}  // extern "C"
