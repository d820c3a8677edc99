// vdf_internal.h -- shared host/device definitions of libvdf_hip.so.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <exception>
#include <memory>
#include <mutex>
#include <new>
#include <string>
#include <vector>

#include "../../include/vdf.h"

namespace vdf {

// ---------------------------------------------------------------------------
// errors
// ---------------------------------------------------------------------------
void set_error(const std::string& msg);

struct Error {
    vdf_status code;
    std::string msg;
};

// Runs f, turning exceptions into a status code plus the thread's message.
template <typename F>
vdf_status guarded(F&& f)
{
    try {
        f();
        return VDF_OK;
    } catch (const Error& e) {
        set_error(e.msg);
        return e.code;
    } catch (const std::bad_alloc&) {
        set_error("host allocation failed");
        return VDF_ERR_OOM;
    } catch (const std::exception& e) {
        set_error(e.what());
        return VDF_ERR_INVALID;
    }
}

inline void require(bool ok, const char* msg)
{
    if (!ok) throw Error{VDF_ERR_INVALID, msg};
}

// bytes of one cell of a value column type (0: unknown type)
inline int val_bytes(int t)
{
    switch (t) {
    case VDF_VAL_F64:
    case VDF_VAL_I64:
    case VDF_VAL_U64: return 8;
    case VDF_VAL_I32:
    case VDF_VAL_U32: return 4;
    default: return 0;
    }
}

#define VDF_HIP_CHECK(expr)                                                              \
    do {                                                                                 \
        hipError_t e_ = (expr);                                                          \
        if (e_ != hipSuccess)                                                            \
            throw ::vdf::Error{VDF_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)}; \
    } while (0)

// ---------------------------------------------------------------------------
// radix constants: 8-bit digits, 256 buckets
// ---------------------------------------------------------------------------
constexpr int kRadixBits = 8;
constexpr int kRadix = 1 << kRadixBits;
constexpr int kBlock = 256;            // threads per sort / group-by block (4 waves)
constexpr int kWaves = kBlock / 64;
constexpr int kMaxDigits = 16;         // 128-bit keys
constexpr uint32_t kFlagA = 1u << 30;  // look-back word: tile aggregate available
constexpr uint32_t kFlagP = 2u << 30;  // look-back word: inclusive prefix available
constexpr uint32_t kValMask = (1u << 30) - 1u;
constexpr uint32_t kTicketStride = 32;   // 32-bit words of tile-claim counters per pass slot

constexpr uint64_t kDefaultHistRows = 65536;  // rows per histogram work item

// One segment (table) of a sort batch: caller's keys and outputs.
struct SortSeg {
    const void* keys;
    uint64_t n;
    void* perm_out;   // local row ids, u32 or (VDF_SORT_PERM64) u64; nullable
    void* keys_out;   // nullable
    const void* pay_in = nullptr;  // payload column carried through the passes (batch pay_bytes per row)
    void* pay_out = nullptr;       // the payload in sorted order
    const void* pay_in2 = nullptr; // second 8-byte payload column (pay_bytes 16)
    void* pay_out2 = nullptr;
};

// ---------------------------------------------------------------------------
// handle: scratch arena + timing
// ---------------------------------------------------------------------------
// Byte accounting of the radix pass launches of one sort: each table's pass
// count (pinned copy of the device plan) and rows, and the bytes per row of a
// pass that is the table's only / first / middle / last one.
// A table's plan is two words: the end of its pass slots and its mode (first
// slot in bits 0..7, kModeLocal when a local finish writes the outputs).
constexpr uint32_t kModeLocal = 1u << 8;
constexpr int kLocalSlot0 = 2;       // first pass slot of the plain passes after a local finish fell back
constexpr int kProfLocalSlot = 64;   // ProfRec.plan_slot of a local-finish launch
struct PassAcct {
    const uint32_t* passes = nullptr;   // 2 * nsegs words after the plan, valid once the stream reached the copy
    const uint32_t* passes2 = nullptr;  // the same after a local finish's size check (null: none launched)
    std::vector<uint64_t> rows;
    uint64_t per_row[4] = {0, 0, 0, 0};
    uint64_t per_row_local = 0;  // bytes per row of a local finish
};

struct ProfRec {
    std::string name;
    hipEvent_t beg, end;
    uint64_t bytes;                        // algorithmic bytes of the launch (no acct)
    int plan_slot;                         // >= 0: radix pass slot
    std::shared_ptr<const PassAcct> acct;  // pass slots: per-table activity and bytes
    const uint64_t* groups = nullptr;      // group-by tile launch: pinned copy of the groups written,
    hipEvent_t groups_done = nullptr;      // valid once this event completed,
    uint64_t group_bytes = 0;              // and the output bytes per group
};

struct ProfTotal {
    std::string name;
    double ms = 0;
    uint64_t launches = 0;
    uint64_t bytes = 0;
};

// Tuning table of one handle (set by vdf_tune_set or the autotuner).
struct Tuning {
    int sort_items_32 = 0;   // rows per thread of a radix pass tile, 4-byte keys (0: by payload width)
    int sort_items_64 = 0;   // same, 8-byte keys
    int sort_items_128 = 0;  // same, 16-byte keys
    int sort_chunk = 64;     // consecutive tiles one XCD claims at a time
    int sort_xcd = 1;        // 1: XCD-grouped tile order; 0: global tile order (both plain
                             //    launches, safe beside other persistent work)
    int sort_bpc = 0;        // radix pass blocks per CU (persistent grid); 0 = every resident block
    int hist_rows = 65536;   // rows per histogram work item
    int gather_rows = 8;     // rows per thread of the gather kernel
    int gather_nt = 0;       // non-temporal loads of the gathered column
    int csv_stage = 1;       // CSV row kernels parse each wave's rows from a copy in LDS
    int gb_carry = 1;        // group-by: value columns carried through the sort (1) or gathered (0)
    int sort_local = 1;      // 4-byte keys: finish a sort in LDS after two global passes when its buckets fit
    int sort_rank = 1;       // 1: one LDS atomic per row ranks (where the device check passed); 0: wave ballots
};

bool tune_supported(const char* name, int64_t value);

struct Handle {
    int device = 0;
    Tuning tune;
    uint64_t budget = 0;
    std::mutex mu;
    // scratch arena (grown on demand, never shrunk)
    void* scratch = nullptr;
    size_t scratch_cap = 0;
    // timing
    bool prof = false;
    std::vector<ProfRec> recs;
    std::vector<hipEvent_t> free_events;
    std::vector<ProfTotal> totals;
    uint32_t* passes_ring = nullptr;   // pinned host ring for per-table pass counts
    size_t ring_pos = 0;
    static constexpr size_t kRing = size_t(1) << 20;

    // pinned host staging for batch descriptions: a ring of buffers, one per
    // call; a buffer is reused once the upload of its previous call finished
    static constexpr int kStaging = 4;
    void* staging[kStaging] = {};
    size_t staging_cap[kStaging] = {};
    hipEvent_t staging_done[kStaging] = {};
    bool staging_pending[kStaging] = {};
    int staging_slot = -1;
    size_t staging_off = 0;
    unsigned staging_next = 0;
    // a captured plan's own pinned buffer replaces the ring (staging_fixed != null)
    void* staging_fixed = nullptr;
    size_t staging_fixed_cap = 0;
    // dry run: enqueue functions write the host staging of a batch and return
    // before uploading or launching anything (refreshes a captured plan)
    bool dry = false;
    // Group-by sorts always move row ids (graph plans: the outputs of a replay,
    // with or without a permutation, are only known after capture).
    bool keep_rowids = false;
    // row-id-free pass kernels for sorts without row ids (off for server
    // workers, whose sorts run concurrently on several streams)
    bool rowfree_kernels = true;
    // one-table sorts write their descriptors on the device instead of
    // uploading them from pinned staging (set only for allocation-free calls)
    bool device_desc = false;

    void* ensure_scratch(size_t bytes);
    void staging_begin(size_t bytes);      // start a call's staging region
    void* staging_take(size_t bytes);      // carve from it (256-byte aligned)
    void staging_end(hipStream_t s);       // the region's uploads are enqueued on s
    hipEvent_t get_event();
    void prof_collect();
    void release_resources();  // events, scratch, pinned buffers (work must be finished)
};

// scoped launch timer: records events around a launch when profiling is on
struct LaunchTimer {
    Handle* h;
    hipStream_t s;
    const char* name;
    uint64_t bytes;
    int slot;
    std::shared_ptr<const PassAcct> acct;
    hipEvent_t b = nullptr;
    LaunchTimer(Handle* h_, hipStream_t s_, const char* n_, uint64_t bytes_, int slot_ = -1,
                std::shared_ptr<const PassAcct> acct_ = nullptr)
        : h(h_), s(s_), name(n_), bytes(bytes_), slot(slot_), acct(std::move(acct_))
    {
        if (h->prof) {
            b = h->get_event();
            VDF_HIP_CHECK(hipEventRecord(b, s));
        }
    }
    ~LaunchTimer() noexcept(false)
    {
        if (h->prof) {
            hipEvent_t e = h->get_event();
            VDF_HIP_CHECK(hipEventRecord(e, s));
            h->recs.push_back(ProfRec{name, b, e, bytes, slot, acct});
        }
    }
};

inline void check_launch(const char* what)
{
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) throw Error{VDF_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e)};
}

inline size_t align_up(size_t x, size_t a) { return (x + a - 1) / a * a; }

// Device pointers read from descriptors in memory are generic: accesses
// through them compile to flat instructions, which count against lgkmcnt as
// well as vmcnt, so every LDS wait (barriers, shuffles) would also wait for
// outstanding global traffic.  These casts put them in the global address
// space (all of them point into device memory).
template <typename T>
__device__ inline __attribute__((address_space(1))) T* gmem(T* p)
{
    return (__attribute__((address_space(1))) T*)p;
}
template <typename T>
__device__ inline const __attribute__((address_space(1))) T* gmem(const T* p)
{
    return (const __attribute__((address_space(1))) T*)p;
}

// Bump allocator over the handle's scratch arena.
struct Carve {
    char* base;
    size_t off = 0;
    explicit Carve(void* b) : base(static_cast<char*>(b)) {}
    template <typename T>
    T* take(size_t count)
    {
        off = align_up(off, 256);
        T* p = reinterpret_cast<T*>(base + off);
        off += count * sizeof(T);
        return p;
    }
};

// ---------------------------------------------------------------------------
// entry points implemented by the kernel files
// ---------------------------------------------------------------------------
size_t key_bytes(int key_type);
size_t argsort_scratch(int key_type, uint64_t n, uint32_t flags);

size_t argsort_batch_scratch(int key_type, const uint64_t* n, int nsegs, uint64_t hist_rows, int pay_bytes = 0);
size_t argsort_batch_staging(int key_type, const uint64_t* n, int nsegs, uint64_t hist_rows);

// Sorts into d_perm_out / d_keys_out using scratch at `scr` (argsort_scratch bytes).
void argsort_enqueue(Handle* h, int key_type, const void* d_keys, uint64_t n, uint32_t flags,
                     void* d_perm_out, void* d_keys_out, void* scr, hipStream_t s);
// Sorts every segment of a batch (argsort_batch_scratch bytes of scratch).
// pay_bytes (0, 4, 8 or 16): every segment's payload column (pay_in ->
// pay_out; with 16, two 8-byte columns, the second pay_in2 -> pay_out2) moves
// with its keys through the passes.
// rowids false: no segment takes a permutation (perm_out null), so the
// passes do not move row ids.
void argsort_batch_enqueue(Handle* h, int key_type, const SortSeg* segs, int nsegs, uint32_t flags, void* scr,
                           hipStream_t s, int pay_bytes = 0, bool rowids = true);

void gather_enqueue(Handle* h, uint32_t elem_bytes, const void* d_src, const void* d_perm,
                    int perm64, uint64_t n, void* d_dst, hipStream_t s);
// out[perm[p]] = group value of sorted position p (per-row broadcast)
void broadcast_enqueue(Handle* h, int src_type, int dst_type, const void* gval, const uint32_t* offsets,
                       const uint64_t* n_groups, uint64_t cap, const uint32_t* perm, uint64_t n, void* out,
                       hipStream_t s);

// One table of a group-by batch.
struct GbTable {
    const void* keys;
    uint64_t n;
    const void* vals[VDF_MAX_VALUE_COLS];
};

size_t groupby_scratch(int key_type, uint64_t n);
// Payload bytes per row of the value columns a group-by carries through its
// sort (see groupby.hip); 0 when every column is gathered.
int groupby_carry_bytes(const Tuning& t, int key_type, const int* val_types, int nvals);
size_t groupby_batch_scratch(int key_type, const uint64_t* n, int ntables, uint64_t hist_rows, int pay_bytes);
size_t groupby_batch_staging(int key_type, const uint64_t* n, int ntables, uint64_t hist_rows);
void groupby_enqueue(Handle* h, int key_type, const void* d_keys, uint64_t n,
                     const vdf_value_col* vals, int nvals, uint32_t agg_mask,
                     const vdf_groupby_out* out, void* scr, hipStream_t s);
void groupby_batch_enqueue(Handle* h, int key_type, const GbTable* tables, int ntables, const int* val_types,
                           int nvals, uint32_t agg_mask, const vdf_groupby_out* outs, void* scr, hipStream_t s);
// the aggregate stage alone, over already sorted tables (keys at skeys with
// 16-row aligned table starts, permutations perms[i])
size_t groupby_agg_scratch(const uint64_t* n, int ntables);
size_t groupby_agg_staging(const uint64_t* n, int ntables);
// spay/pay_bytes: value columns 0 (and 1 with 16 bytes) already in key order
// (as the sort that carried them wrote them), else null/0.
void groupby_aggregate_enqueue(Handle* h, int key_type, const void* skeys, const uint32_t* const* perms,
                               const GbTable* tables, int ntables, const int* val_types, int nvals, uint32_t agg_mask,
                               const vdf_groupby_out* outs, void* scr, hipStream_t s, const void* spay = nullptr,
                               int pay_bytes = 0);

// multi-column keys: a column's bytes per cell (VDF_KEY_BYTES: its width),
// the composite's width, the composite key type of a width (-1 when wider
// than 16 bytes: a byte-string row of that width), pack the columns into
// composites, unpack composite group keys into the columns
uint32_t key_col_width(int type, uint32_t width);
uint32_t composite_width(const int* types, const uint32_t* widths, int ncols);
int composite_key_type(uint32_t width);
// LDS-only block barrier: orders LDS traffic between the block's waves
// without waiting for the block's outstanding global loads and stores (which
// __syncthreads waits for).
__device__ inline void lds_barrier() { asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory"); }

// Exclusive scan of one value per thread over a 256-thread block.
// Inclusive sum over lanes 0 .. lane of a wave (every lane active), by DPP:
// VALU operand routing instead of six dependent LDS-crossbar permutes.
// Shifts by 1, 2, 4 and 8 inside each row of 16 lanes (lanes without a
// source add 0), then lane 15's sum into the next row for rows 1 and 3
// (row_bcast:15), then lane 31's into rows 2 and 3 (row_bcast:31).
__device__ inline uint32_t wave_incl_scan(uint32_t x)
{
    x += static_cast<uint32_t>(__builtin_amdgcn_update_dpp(0, static_cast<int>(x), 0x111, 0xF, 0xF, false));
    x += static_cast<uint32_t>(__builtin_amdgcn_update_dpp(0, static_cast<int>(x), 0x112, 0xF, 0xF, false));
    x += static_cast<uint32_t>(__builtin_amdgcn_update_dpp(0, static_cast<int>(x), 0x114, 0xF, 0xF, false));
    x += static_cast<uint32_t>(__builtin_amdgcn_update_dpp(0, static_cast<int>(x), 0x118, 0xF, 0xF, false));
    x += static_cast<uint32_t>(__builtin_amdgcn_update_dpp(0, static_cast<int>(x), 0x142, 0xA, 0xF, false));
    x += static_cast<uint32_t>(__builtin_amdgcn_update_dpp(0, static_cast<int>(x), 0x143, 0xC, 0xF, false));
    return x;
}

__device__ inline uint32_t block_excl_scan(uint32_t x, uint32_t* s_wsum)
{
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const uint32_t inc = wave_incl_scan(x);
    if (lane == 63) s_wsum[w] = inc;
    __syncthreads();
    uint32_t pre = 0;
#pragma unroll
    for (int i = 0; i < kWaves; ++i) pre += (i < w) ? s_wsum[i] : 0u;
    __syncthreads();
    return pre + inc - x;
}

// fixed-width byte strings (strings.hip)
size_t bytes_argsort_scratch(uint64_t n, uint64_t hist_rows);
size_t bytes_argsort_staging(uint32_t width, uint64_t n, uint64_t hist_rows);
void bytes_argsort_enqueue(Handle* h, const void* d_keys, uint32_t width, uint64_t n, uint32_t flags,
                           void* d_perm_out, void* d_sorted_out, void* scr, hipStream_t s);
struct GbTable;
size_t bytes_groupby_scratch(uint32_t width, uint64_t n, uint64_t hist_rows);
size_t bytes_groupby_staging(uint32_t width, uint64_t n, uint64_t hist_rows);
void bytes_groupby_enqueue(Handle* h, const void* d_keys, uint32_t width, uint64_t n, const GbTable& t,
                           const int* val_types, int nvals, uint32_t agg_mask, const vdf_groupby_out* out, void* scr,
                           hipStream_t s);

void pack_keys_enqueue(Handle* h, const int* types, const uint32_t* widths, const void* const* data, int ncols,
                       uint64_t n, void* comp, hipStream_t s);
void unpack_keys_enqueue(Handle* h, const int* types, const uint32_t* widths, void* const* outs, int ncols,
                         const void* comp, const uint64_t* n_groups, uint64_t cap, hipStream_t s);

// Verifies once per device (with that device current; allocates a word and
// synchronises) that LDS atomic adds serve the lanes of one instruction in
// lane order, which makes one atomic per row a stable rank.  Called where
// handles are made (vdf_create, server workers); sorts on a device where it
// has not run use the ballot ranking.
bool lds_rank_check(int device);
bool lds_rank_verified(int device);  // the check's result where it ran, else false

// fill kernel (graph-safe replacement for hipMemsetAsync)
void fill_u32(uint32_t* p, size_t count, uint32_t value, hipStream_t s);

}  // namespace vdf

// the C ABI's opaque handle is the library's Handle
struct vdf_handle : vdf::Handle {};
