// groupby.hip -- sort-based group-by for gfx950 (MI355X).
//
// Semantics restate the reference's hash group-by aggregates
// (inlines/transform/transform_group_by/onecol_functions.inl:4-130 Occurence /
// Sum / Gather, create_value_col.inl:36-61) with these stated contracts:
//   groups in ascending key order; count; sum (f64: double-double accumulation
//   rounded once; integers: two's complement mod 2^64); mean = sum / count
//   (integers: exact 128-bit sum / count); min / max = first of equal values
//   in row order (std::min_element over the row-ordered Gather vector).
//
// Every group-by is a batch of tables (segments); one table is a batch of one.
// Pipeline on one stream:
//   1. radix argsort of every table (radix.hip, one launch sequence for the
//      batch) -> perm (row ids in key order, rows in row order within a key)
//      and the sorted keys, concatenated in scratch
//   2. k_gb_tile: one 256-thread block per 4096-row tile of a table (dynamic
//      tile ids over the tiles of all tables); each lane owns 16 consecutive
//      sorted rows (16-byte vector loads; table starts are 16-row aligned in
//      scratch):
//        head flags (key != previous key) per lane; head count of the tile
//        (wave shuffles + 4 LDS words) published to a decoupled look-back
//        array -> global group ids; group keys and offsets written at heads;
//        per value column the lane gathers its 16 values through perm,
//        reduces them sequentially, and a wave segmented scan by shuffles
//        (cross-wave carry through LDS) joins the lanes.  Groups that start
//        and end inside the tile are final here; each tile leaves two
//        partials (before its first head, after its last head).
//   3. k_gb_fixup: one thread per tile owning the group that starts at the
//        tile's last head; a group ending in the next tile is joined there,
//        longer ones are queued for k_gb_fixup_long, which joins the
//        following tiles' leading partials in tile order with a fixed
//        reduction tree (a block per queued tile).
// The reduction order is a function of the row count and tile size only, so a
// table's results do not depend on what else runs with it.
#include "vdf_internal.h"

#include <algorithm>

namespace vdf {

namespace {

constexpr int kGbRows = 16;                 // consecutive rows per thread
constexpr int kGbTile = kBlock * kGbRows;   // 4096 rows per tile

enum : int { KIND_F64 = 0, KIND_SIGNED = 1, KIND_UNSIGNED = 2, KIND_COUNT = 3 };

struct K16 {
    uint64_t lo, hi;
};
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
typedef uint64_t u64x2 __attribute__((ext_vector_type(2)));

// key loads / stores through global-address-space pointers
__device__ inline uint32_t get_key(const uint32_t* p, uint64_t i) { return gmem(p)[i]; }
__device__ inline uint64_t get_key(const uint64_t* p, uint64_t i) { return gmem(p)[i]; }
__device__ inline K16 get_key(const K16* p, uint64_t i)
{
    const auto* w = gmem(reinterpret_cast<const uint64_t*>(p)) + 2 * i;
    return K16{w[0], w[1]};
}
__device__ inline void put_key(uint32_t* p, uint64_t i, uint32_t k) { gmem(p)[i] = k; }
__device__ inline void put_key(uint64_t* p, uint64_t i, uint64_t k) { gmem(p)[i] = k; }
__device__ inline void put_key(K16* p, uint64_t i, K16 k)
{
    auto* w = gmem(reinterpret_cast<uint64_t*>(p)) + 2 * i;
    w[0] = k.lo;
    w[1] = k.hi;
}

__device__ inline bool keq(uint32_t a, uint32_t b) { return a == b; }
__device__ inline bool keq(uint64_t a, uint64_t b) { return a == b; }
__device__ inline bool keq(const K16& a, const K16& b) { return a.lo == b.lo && a.hi == b.hi; }

// Running aggregate of one segment.  F64: a = hi, b = lo of a double-double
// sum; integers: a = low, b = high word of a 128-bit sum.  mn / mx hold the raw
// 64-bit value (f64 bits or the sign / zero extended integer).
struct Agg {
    uint64_t cnt;
    uint64_t a, b;
    uint64_t mn, mx;
};

__device__ inline Agg agg_empty() { return Agg{0, 0, 0, 0, 0}; }
__device__ inline double as_d(uint64_t x) { return __longlong_as_double(static_cast<long long>(x)); }
__device__ inline uint64_t as_u(double d) { return static_cast<uint64_t>(__double_as_longlong(d)); }

__device__ inline void two_sum(double x, double y, double& s, double& e)
{
    s = x + y;
    const double bp = s - x;
    e = (x - (s - bp)) + (y - bp);
}

__device__ inline void agg_add(Agg& s, uint64_t x, int kind)
{
    if (kind == KIND_F64) {
        const double v = as_d(x);
        if (s.cnt == 0) {
            s.mn = x;
            s.mx = x;
        } else {
            if (v < as_d(s.mn)) s.mn = x;
            if (v > as_d(s.mx)) s.mx = x;
        }
        double t, e;
        two_sum(as_d(s.a), v, t, e);
        s.a = as_u(t);
        s.b = as_u(as_d(s.b) + e);
    } else if (kind == KIND_SIGNED) {
        const int64_t v = static_cast<int64_t>(x);
        if (s.cnt == 0) {
            s.mn = x;
            s.mx = x;
        } else {
            if (v < static_cast<int64_t>(s.mn)) s.mn = x;
            if (v > static_cast<int64_t>(s.mx)) s.mx = x;
        }
        const uint64_t lo = s.a + x;
        s.b += static_cast<uint64_t>(v >> 63) + (lo < s.a ? 1u : 0u);
        s.a = lo;
    } else if (kind == KIND_UNSIGNED) {
        if (s.cnt == 0) {
            s.mn = x;
            s.mx = x;
        } else {
            if (x < s.mn) s.mn = x;
            if (x > s.mx) s.mx = x;
        }
        const uint64_t lo = s.a + x;
        s.b += (lo < s.a ? 1u : 0u);
        s.a = lo;
    }
    s.cnt += 1;
}

// A precedes B in row order.
__device__ inline Agg agg_join(const Agg& A, const Agg& B, int kind)
{
    if (B.cnt == 0) return A;
    if (A.cnt == 0) return B;
    Agg r;
    r.cnt = A.cnt + B.cnt;
    if (kind == KIND_F64) {
        double t, e;
        two_sum(as_d(A.a), as_d(B.a), t, e);
        r.a = as_u(t);
        r.b = as_u((as_d(A.b) + as_d(B.b)) + e);
        r.mn = (as_d(B.mn) < as_d(A.mn)) ? B.mn : A.mn;
        r.mx = (as_d(B.mx) > as_d(A.mx)) ? B.mx : A.mx;
    } else {
        const uint64_t lo = A.a + B.a;
        r.b = A.b + B.b + (lo < A.a ? 1u : 0u);
        r.a = lo;
        if (kind == KIND_SIGNED) {
            r.mn = (static_cast<int64_t>(B.mn) < static_cast<int64_t>(A.mn)) ? B.mn : A.mn;
            r.mx = (static_cast<int64_t>(B.mx) > static_cast<int64_t>(A.mx)) ? B.mx : A.mx;
        } else {
            r.mn = (B.mn < A.mn) ? B.mn : A.mn;
            r.mx = (B.mx > A.mx) ? B.mx : A.mx;
        }
    }
    return r;
}

// Correctly rounded conversion of an unsigned 128-bit integer to double.
__device__ inline double u128_to_double(uint64_t lo, uint64_t hi)
{
    if (hi == 0) return static_cast<double>(lo);
    const int lz = __clzll(static_cast<long long>(hi));
    uint64_t m = lz ? ((hi << lz) | (lo >> (64 - lz))) : hi;
    const uint64_t rest = lz ? (lo << lz) : lo;
    m |= (rest != 0) ? 1u : 0u;  // sticky bit below the rounding position
    return ldexp(static_cast<double>(m), 64 - lz);
}

__device__ inline double i128_to_double(uint64_t lo, uint64_t hi)
{
    if (static_cast<int64_t>(hi) >= 0) return u128_to_double(lo, hi);
    const uint64_t nlo = ~lo + 1u;
    const uint64_t nhi = ~hi + (nlo == 0 ? 1u : 0u);
    return -u128_to_double(nlo, nhi);
}

// One value column of the batch (types are shared by all tables).
struct GbColType {
    int type;  // VDF_VAL_*
    int kind;  // KIND_*
};

// One table: its rows in the concatenated scratch and the caller's outputs.
struct GbSeg {
    uint64_t n;
    uint64_t row_off;        // first row of the table's sorted keys in scratch
    const uint32_t* perm;    // the table's permutation (local row ids)
    uint32_t gtile_off;      // first group-by tile
    uint32_t ntiles;         // group-by tiles of the table
    uint64_t cap;            // groups the outputs hold: ids >= cap are not written
    const void* vals[VDF_MAX_VALUE_COLS];   // caller's columns (row order), gathered through perm
    const void* svals[VDF_MAX_VALUE_COLS];  // columns already in key order (null: gather)
    vdf_groupby_out out;
};

struct GbArgs {
    const GbSeg* segs;
    const uint32_t* tile_seg;  // tile -> table (null: one table)
    const void* skeys;         // sorted keys of all tables (caller's key type)
    uint32_t ntiles;
    int ncols;                 // >= 1 (a count-only column when no values are aggregated)
    GbColType cols[VDF_MAX_VALUE_COLS];
    uint32_t agg_mask;
    uint32_t* status;          // [ntiles] look-back words for head counts
    uint32_t* ticket;
    uint32_t* tile_gbase;      // [ntiles] table-local id of the tile's first head's group
    uint32_t* tile_heads;      // [ntiles] heads in the tile
    Agg* prefix;               // [ncols][ntiles] rows before the tile's first head (or the whole tile)
    Agg* suffix;               // [ncols][ntiles] rows from the tile's last head to its end
    uint32_t* long_count;      // tiles queued in long_list (zeroed before the tile kernel)
    uint32_t* long_list;       // [ntiles] tiles whose last group spans following tiles
    unsigned long long* group_total;  // groups written over all tables (profiling only, else null)
};

// Output pointers of one value column, read from the table descriptor once
// (in registers, so the per-group stores cannot force reloads of the
// descriptor, which the compiler could not otherwise tell apart from the
// outputs).
struct ColOut {
    uint32_t* counts;  // column 0 only, when counts are asked for
    void* sum;
    double* mean;
    void* min;
    void* max;
    uint64_t cap;
};

// a wave-uniform 64-bit value, held in scalar registers
template <typename T>
__device__ inline T uniform64(T x)
{
    uint64_t u;
    __builtin_memcpy(&u, &x, 8);
    const uint32_t lo = __builtin_amdgcn_readfirstlane(static_cast<uint32_t>(u));
    const uint32_t hi = __builtin_amdgcn_readfirstlane(static_cast<uint32_t>(u >> 32));
    u = (static_cast<uint64_t>(hi) << 32) | lo;
    __builtin_memcpy(&x, &u, 8);
    return x;
}

// (sg is the same table for the whole wave)
// Output pointers of column c of table sg.  UNIFORM: every lane of the wave
// works on the same table, so the pointers are moved to scalar registers.
template <bool UNIFORM = true>
__device__ inline ColOut col_out(const GbArgs& a, const GbSeg& sg, int c)
{
    ColOut o;
    const vdf_agg_out& ag = sg.out.aggs[c];
    o.counts = (c == 0 && (a.agg_mask & VDF_AGG_COUNT)) ? sg.out.counts : nullptr;
    o.sum = (a.agg_mask & VDF_AGG_SUM) ? ag.sum : nullptr;
    o.mean = (a.agg_mask & VDF_AGG_MEAN) ? ag.mean : nullptr;
    o.min = (a.agg_mask & VDF_AGG_MIN) ? ag.min : nullptr;
    o.max = (a.agg_mask & VDF_AGG_MAX) ? ag.max : nullptr;
    o.cap = sg.cap;
    if (UNIFORM) {
        o.counts = uniform64(o.counts);
        o.sum = uniform64(o.sum);
        o.mean = uniform64(o.mean);
        o.min = uniform64(o.min);
        o.max = uniform64(o.max);
        o.cap = uniform64(o.cap);
    }
    return o;
}

// The output words of one finished group: sum (f64 bits or the low 64 bits
// of the integer sum), mean (f64 bits), min / max (raw), count.
struct Final {
    uint64_t sum, mean, mn, mx;
    uint32_t cnt;
};

__device__ inline Final final_of(const GbArgs& a, int c, const Agg& s)
{
    Final f{0, 0, s.mn, s.mx, static_cast<uint32_t>(s.cnt)};
    const GbColType col = a.cols[c];
    if (col.kind == KIND_COUNT) return f;
    if (col.kind == KIND_F64) {
        const double sum = as_d(s.a) + as_d(s.b);
        f.sum = as_u(sum);
        if (a.agg_mask & VDF_AGG_MEAN) f.mean = as_u(sum / static_cast<double>(s.cnt));
    } else {
        f.sum = s.a;
        if (a.agg_mask & VDF_AGG_MEAN) {
            const double tot = (col.kind == KIND_SIGNED) ? i128_to_double(s.a, s.b) : u128_to_double(s.a, s.b);
            f.mean = as_u(tot / static_cast<double>(s.cnt));
        }
    }
    return f;
}

__device__ inline void emit_final(const GbArgs& a, const ColOut& o, int c, uint32_t gid, const Final& f)
{
    if (gid >= o.cap) return;
    const GbColType col = a.cols[c];
    if (o.counts) gmem(o.counts)[gid] = f.cnt;
    if (col.kind == KIND_COUNT) return;
    const bool w4 = (col.type == VDF_VAL_I32 || col.type == VDF_VAL_U32);
    if (o.sum) gmem(static_cast<uint64_t*>(o.sum))[gid] = f.sum;
    if (o.mean) gmem(reinterpret_cast<uint64_t*>(o.mean))[gid] = f.mean;
    if (o.min) {
        if (w4) gmem(static_cast<uint32_t*>(o.min))[gid] = static_cast<uint32_t>(f.mn);
        else gmem(static_cast<uint64_t*>(o.min))[gid] = f.mn;
    }
    if (o.max) {
        if (w4) gmem(static_cast<uint32_t*>(o.max))[gid] = static_cast<uint32_t>(f.mx);
        else gmem(static_cast<uint64_t*>(o.max))[gid] = f.mx;
    }
}

__device__ inline void write_final(const GbArgs& a, const ColOut& o, int c, uint32_t gid, const Agg& s)
{
    emit_final(a, o, c, gid, final_of(a, c, s));
}

// Finished groups of a tile are staged in LDS by tile-local group index and
// written out afterwards by consecutive lanes, so each output array takes one
// store instruction per 256 groups instead of one per row position that holds
// a head somewhere in the wave.  Tiles with more than kGbStage groups stage
// them in windows of kGbStage ids, one after the other.
constexpr int kGbStage = 1024;

struct GbStage {
    uint64_t a[kGbStage], b[kGbStage], mn[kGbStage], mx[kGbStage];  // the group's aggregate
    uint32_t cnt[kGbStage];
};

// a group key in the stage (16-byte keys use two words)
__device__ inline void stage_key(GbStage* st, uint32_t h, uint32_t k) { st->a[h] = k; }
__device__ inline void stage_key(GbStage* st, uint32_t h, uint64_t k) { st->a[h] = k; }
__device__ inline void stage_key(GbStage* st, uint32_t h, const K16& k)
{
    st->a[h] = k.lo;
    st->b[h] = k.hi;
}
template <typename K>
__device__ inline K staged_key(const GbStage* st, uint32_t h)
{
    if constexpr (sizeof(K) == 16) return K{st->a[h], st->b[h]};
    else return static_cast<K>(st->a[h]);
}

// Stages group gid when it falls in the window of kGbStage ids from wbase
// (groups outside the window are staged when their window's turn comes).
__device__ inline void stage_final(GbStage* st, uint32_t wbase, uint32_t gid, const Agg& s)
{
    const uint32_t h = gid - wbase;  // groups before the window wrap to large values
    if (h < static_cast<uint32_t>(kGbStage)) {
        st->cnt[h] = static_cast<uint32_t>(s.cnt);  // a group inside one tile: at most 4096 rows
        st->a[h] = s.a;
        st->b[h] = s.b;
        st->mn[h] = s.mn;
        st->mx[h] = s.mx;
    }
}

__device__ inline void store_relaxed(uint32_t* p, uint32_t v)
{
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ inline uint32_t load_relaxed(uint32_t* p)
{
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// 64-bit wave shuffles as two 32-bit ones
__device__ inline uint64_t shfl_down64(uint64_t x, int off)
{
    const uint32_t lo = __shfl_down(static_cast<uint32_t>(x), off, 64);
    const uint32_t hi = __shfl_down(static_cast<uint32_t>(x >> 32), off, 64);
    return (static_cast<uint64_t>(hi) << 32) | lo;
}

// Lane moves by DPP (data-parallel primitives: VALU operand routing, no LDS
// round trip).  CTRL: row_shr:n = 0x110 + n (within 16-lane rows),
// wave_shr:1 = 0x138 (across the wave), row_bcast:15 = 0x142,
// row_bcast:31 = 0x143; RM: rows written.  Lanes a move does not reach get 0
// and are not used by the caller.
template <int CTRL, int RM>
__device__ inline uint32_t dpp32(uint32_t x)
{
    return static_cast<uint32_t>(__builtin_amdgcn_update_dpp(0, static_cast<int>(x), CTRL, RM, 0xF, false));
}
template <int CTRL, int RM>
__device__ inline uint64_t dpp64(uint64_t x)
{
    return (static_cast<uint64_t>(dpp32<CTRL, RM>(static_cast<uint32_t>(x >> 32))) << 32) |
           dpp32<CTRL, RM>(static_cast<uint32_t>(x));
}
template <int KIND, int CTRL, int RM>
__device__ inline Agg dpp_agg(const Agg& x)
{
    Agg y = agg_empty();
    y.cnt = dpp64<CTRL, RM>(x.cnt);
    if (KIND != KIND_COUNT) {
        y.a = dpp64<CTRL, RM>(x.a);
        y.b = dpp64<CTRL, RM>(x.b);
        y.mn = dpp64<CTRL, RM>(x.mn);
        y.mx = dpp64<CTRL, RM>(x.mx);
    }
    return y;
}

// The segmented-scan operator: (f1, x1) then (f2, x2) -> (f1 | f2, f2 ? x2 : x1 + x2).
template <int KIND>
__device__ inline void seg_combine(Agg& x, uint32_t& f, const Agg& y, uint32_t fy)
{
    // (y, fy) precedes (x, f)
    if (!f) x = agg_join(y, x, KIND);
    f |= fy;
}

// Loads this lane's rows p[0 .. cnt) (cnt <= kGbRows) of 4-byte words; 16-byte
// vector loads when the lane has all 16 rows and p is 16-byte aligned.
__device__ inline void load_rows_u32(const uint32_t* p, uint32_t cnt, uint32_t (&o)[kGbRows])
{
    if (cnt == kGbRows && (reinterpret_cast<uintptr_t>(p) & 15u) == 0) {
        const auto* q = gmem(reinterpret_cast<const u32x4*>(p));
#pragma unroll
        for (int i = 0; i < kGbRows / 4; ++i) {
            const u32x4 t = q[i];
            o[4 * i + 0] = t.x;
            o[4 * i + 1] = t.y;
            o[4 * i + 2] = t.z;
            o[4 * i + 3] = t.w;
        }
    } else {
#pragma unroll
        for (int j = 0; j < kGbRows; ++j) o[j] = (static_cast<uint32_t>(j) < cnt) ? gmem(p)[j] : 0u;
    }
}

// 16 consecutive 8-byte values of a lane (rows past cnt read as 0)
__device__ inline void load_rows_u64(const uint64_t* p, uint32_t cnt, uint64_t (&o)[kGbRows])
{
    if (cnt == kGbRows && (reinterpret_cast<uintptr_t>(p) & 15u) == 0) {
        const auto* q = gmem(reinterpret_cast<const u64x2*>(p));
#pragma unroll
        for (int i = 0; i < kGbRows / 2; ++i) {
            const u64x2 t = q[i];
            o[2 * i + 0] = t.x;
            o[2 * i + 1] = t.y;
        }
    } else {
#pragma unroll
        for (int j = 0; j < kGbRows; ++j) o[j] = (static_cast<uint32_t>(j) < cnt) ? gmem(p)[j] : 0ull;
    }
}

// Head bits of this lane's rows (bit j: row r0 + j starts a group).  Rows are
// read in 4-row chunks; a chunk is one or more 16-byte vector loads when the
// lane has all 16 rows (table starts are 16-row aligned in scratch).
template <typename K>
__device__ inline uint32_t head_bits(const K* keys, uint64_t r0, uint32_t cnt)
{
    if (cnt == 0) return 0u;
    K prev = get_key(keys, r0 == 0 ? 0 : r0 - 1);
    uint32_t hb = 0;
    if (cnt == kGbRows) {
        constexpr int V = static_cast<int>(4 * sizeof(K) / 16);  // uint4 per 4 rows
        const auto* q = gmem(reinterpret_cast<const u32x4*>(keys + r0));
#pragma unroll
        for (int c = 0; c < kGbRows / 4; ++c) {
            u32x4 t[V];
#pragma unroll
            for (int i = 0; i < V; ++i) t[i] = q[c * V + i];
            K k[4];
            __builtin_memcpy(k, t, sizeof(k));
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const int j = 4 * c + i;
                if (!keq(k[i], prev) || (r0 + j == 0)) hb |= 1u << j;
                prev = k[i];
            }
        }
    } else {
        for (uint32_t j = 0; j < cnt; ++j) {
            const K k = get_key(keys, r0 + j);
            if (!keq(k, prev) || (r0 + j == 0)) hb |= 1u << j;
            prev = k;
        }
    }
    return hb;
}

// One value column of a tile: gather through perm, sequential reduce over the
// lane's 16 rows, wave segmented scan by shuffles, cross-wave carry through LDS.
template <int KIND>
__device__ __attribute__((always_inline)) inline void gb_column(const GbArgs& a, const GbSeg& sg, int c, uint32_t tile, uint64_t tbase, uint32_t cnt,
                          uint32_t hb, uint32_t gfirst, uint32_t gbase, uint32_t H, Agg* s_wagg, uint32_t* s_wflag,
                          GbStage* st)
{
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const uint64_t r0 = tbase + static_cast<uint64_t>(tid) * kGbRows;
    const ColOut co = col_out(a, sg, c);
    uint64_t v[kGbRows];
    auto load_vals = [&]() {
    if (KIND != KIND_COUNT) {
        // A gathered column: all 16 gathers are issued before any conversion
        // waits on one.  Rows past the tile's end hold row id 0 (in bounds:
        // the table is not empty) and are never aggregated.
        const void* vd = sg.vals[c];
        const int type = a.cols[c].type;
        const bool w4 = type == VDF_VAL_I32 || type == VDF_VAL_U32;
        if (sg.svals[c]) {
            // the column is in key order: the lane's 16 rows are consecutive
            // (16-byte vector loads; table starts are 16-row aligned)
            if (w4) {
                uint32_t w[kGbRows];
                load_rows_u32(static_cast<const uint32_t*>(sg.svals[c]) + r0, cnt, w);
                if (type == VDF_VAL_I32) {
#pragma unroll
                    for (int j = 0; j < kGbRows; ++j)
                        v[j] = static_cast<uint64_t>(static_cast<int64_t>(static_cast<int32_t>(w[j])));
                } else {
#pragma unroll
                    for (int j = 0; j < kGbRows; ++j) v[j] = w[j];
                }
            } else {
                load_rows_u64(static_cast<const uint64_t*>(sg.svals[c]) + r0, cnt, v);
            }
        } else if (w4) {
            uint32_t p[kGbRows];
            load_rows_u32(sg.perm + r0, cnt, p);
            const auto* q = gmem(static_cast<const uint32_t*>(vd));
            uint32_t w[kGbRows];
#pragma unroll
            for (int j = 0; j < kGbRows; ++j) w[j] = q[p[j]];
            if (type == VDF_VAL_I32) {
#pragma unroll
                for (int j = 0; j < kGbRows; ++j)
                    v[j] = static_cast<uint64_t>(static_cast<int64_t>(static_cast<int32_t>(w[j])));
            } else {
#pragma unroll
                for (int j = 0; j < kGbRows; ++j) v[j] = w[j];
            }
        } else {
            uint32_t p[kGbRows];
            load_rows_u32(sg.perm + r0, cnt, p);
            const auto* q = gmem(static_cast<const uint64_t*>(vd));
#pragma unroll
            for (int j = 0; j < kGbRows; ++j) v[j] = q[p[j]];
        }
    }
    };
    load_vals();
    // Groups finishing in the tile (0 .. H-2) are staged in windows of
    // kGbStage tile-local ids: window 0 in the first run of the row loop, each
    // further window (tiles with more groups, e.g. keys nearly all distinct)
    // by running the row loop again over the same values, staging only that
    // window's groups, so every group is written by consecutive lanes.
    Agg cur = agg_empty(), first = agg_empty();
    bool seen = false;
    auto row_loop = [&](uint32_t lo) {
        cur = agg_empty();
        first = agg_empty();
        seen = false;
        uint32_t g = gfirst;  // id of the next head's group
#pragma unroll
        for (int j = 0; j < kGbRows; ++j) {
            if (static_cast<uint32_t>(j) < cnt) {
                if ((hb >> j) & 1u) {
                    if (!seen) {
                        first = cur;
                        seen = true;
                    } else {
                        stage_final(st, gbase + lo, g - 1, cur);
                    }
                    cur = agg_empty();
                    ++g;
                }
                agg_add(cur, KIND == KIND_COUNT ? 0ull : v[j], KIND);
            }
        }
    };
    row_loop(0);

    // inclusive segmented scan over the wave's lanes, by DPP: a Hillis-Steele
    // scan inside each 16-lane row (shifts 1, 2, 4, 8), then row 0's total into
    // row 1 and row 2's into row 3 (row_bcast:15), then lane 31's prefix into
    // rows 2 and 3 (row_bcast:31).  The operator is associative, so every lane
    // ends with the combination of lanes 0 .. itself, in a fixed order.
    Agg x = cur;
    uint32_t f = seen ? 1u : 0u;
    const int rl = lane & 15, row = lane >> 4;
    {
        Agg y = dpp_agg<KIND, 0x111, 0xF>(x);
        uint32_t fy = dpp32<0x111, 0xF>(f);
        if (rl >= 1) seg_combine<KIND>(x, f, y, fy);
        y = dpp_agg<KIND, 0x112, 0xF>(x);
        fy = dpp32<0x112, 0xF>(f);
        if (rl >= 2) seg_combine<KIND>(x, f, y, fy);
        y = dpp_agg<KIND, 0x114, 0xF>(x);
        fy = dpp32<0x114, 0xF>(f);
        if (rl >= 4) seg_combine<KIND>(x, f, y, fy);
        y = dpp_agg<KIND, 0x118, 0xF>(x);
        fy = dpp32<0x118, 0xF>(f);
        if (rl >= 8) seg_combine<KIND>(x, f, y, fy);
        y = dpp_agg<KIND, 0x142, 0xA>(x);
        fy = dpp32<0x142, 0xA>(f);
        if (row & 1) seg_combine<KIND>(x, f, y, fy);
        y = dpp_agg<KIND, 0x143, 0xC>(x);
        fy = dpp32<0x143, 0xC>(f);
        if (row >= 2) seg_combine<KIND>(x, f, y, fy);
    }
    if (lane == 63) {
        s_wagg[wave] = x;
        s_wflag[wave] = f;
    }
    // each lane's predecessor by DPP wave_shr:1 (lane 0 gets 0, replaced below)
    Agg ex = dpp_agg<KIND, 0x138, 0xF>(x);
    uint32_t exf = dpp32<0x138, 0xF>(f);
    if (lane == 0) {
        ex = agg_empty();
        exf = 0;
    }
    lds_barrier();
    // carry from the waves before this one, in wave order
    Agg wc = agg_empty();
    uint32_t wcf = 0;
    for (int i = 0; i < wave; ++i) {
        if (s_wflag[i]) {
            wc = s_wagg[i];
            wcf = 1;
        } else {
            wc = agg_join(wc, s_wagg[i], KIND);
        }
    }
    // exclusive carry into this lane
    Agg carry = ex;
    uint32_t cf = exf;
    if (!exf) {
        carry = agg_join(wc, ex, KIND);
        cf = wcf;
    }
    Agg* pre = a.prefix + static_cast<uint64_t>(c) * a.ntiles;
    Agg* suf = a.suffix + static_cast<uint64_t>(c) * a.ntiles;
    // the group ending at this lane's first head (carry from earlier lanes)
    Agg seg = agg_empty();
    const bool seg_fin = seen && cf;
    if (seen) {
        seg = agg_join(carry, first, KIND);
        if (cf) stage_final(st, gbase, gfirst - 1, seg);
        else pre[tile] = seg;  // the tile's first head: rows before it
    }
    if (tid == kBlock - 1) {
        Agg bx = x;
        uint32_t bf = f;
        if (!f) {
            bx = agg_join(wc, x, KIND);
            bf = wcf;
        }
        if (bf) suf[tile] = bx;
        else pre[tile] = bx;
    }
    lds_barrier();  // every staged group of window 0 is in LDS
    // groups 0 .. H-2 finish in this tile (group H-1 runs into the next tiles)
    const uint32_t nfin = H ? H - 1 : 0u;
    // the row loop stages this lane's groups of tile-local ids h0 .. h0 + nheads - 2
    const uint32_t nheads = static_cast<uint32_t>(__popc(hb));
    const uint32_t h0 = gfirst - gbase;
    for (uint32_t lo = 0;;) {
        const uint32_t nw = std::min<uint32_t>(nfin - lo, kGbStage);
        for (uint32_t h = tid; h < nw; h += kBlock) {
            const Agg g{st->cnt[h], st->a[h], st->b[h], st->mn[h], st->mx[h]};
            write_final(a, co, c, gbase + lo + h, g);
        }
        lds_barrier();  // s_wagg and the stage are reused (next window, next column)
        lo += kGbStage;
        if (lo >= nfin) break;  // block-uniform: H is the tile's
        // only lanes with a group in this window read their values again
        if (nheads > 1 && h0 + nheads - 1 > lo && h0 < lo + kGbStage) {
            load_vals();
            row_loop(lo);
        }
        if (seg_fin) stage_final(st, gbase + lo, gfirst - 1, seg);
        lds_barrier();
    }
}

// What a block keeps of one claimed tile.
struct GbTileState {
    uint32_t tile;   // global tile id
    uint32_t sgi;    // table
    uint32_t lt;     // tile index inside the table
    uint32_t H;      // heads in the tile
    uint32_t gbase;  // table-local id of the tile's first head's group (after resolve)
    uint64_t tbase;  // first table row of the tile
    uint64_t r0;     // this lane's first row
    uint32_t cnt;    // this lane's rows inside the table
    uint32_t hb;     // this lane's head bits
    uint32_t hx;     // heads before this lane in the tile
};

// Persistent over tiles and pipelined: a block claims tile k+1 and publishes
// its head count before it aggregates tile k, and resolves tile k+1's group
// ids only afterwards, so that the look-back finds its predecessors' prefixes
// published instead of waiting for them.  Tiles are claimed in increasing
// order by running blocks and a block waits only on smaller tiles, so the
// smallest unresolved tile always completes, at any residency.
template <typename K>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(4))) void k_gb_tile(GbArgs a)
{
    __shared__ Agg s_wagg[kWaves];
    __shared__ uint32_t s_wflag[kWaves];
    __shared__ uint32_t s_wsum[kWaves];
    __shared__ uint32_t s_tile, s_gbase;
    __shared__ GbStage s_stage;

    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;

    // claim a tile, find its heads, publish its head count; false at the end
    auto begin = [&](GbTileState& t) -> bool {
        if (tid == 0) s_tile = atomicAdd(a.ticket, 1u);
        lds_barrier();
        t.tile = __builtin_amdgcn_readfirstlane(s_tile);
        if (t.tile >= a.ntiles) return false;
        t.sgi = a.tile_seg ? __builtin_amdgcn_readfirstlane(a.tile_seg[t.tile]) : 0u;
        const GbSeg& sg = a.segs[t.sgi];
        const uint64_t n = sg.n;
        t.lt = t.tile - sg.gtile_off;
        t.tbase = static_cast<uint64_t>(t.lt) * kGbTile;
        const K* keys = static_cast<const K*>(a.skeys) + sg.row_off;
        // this lane's rows: tbase + tid*16 .. +15 (cnt of them inside the table)
        const uint32_t tn = static_cast<uint32_t>(std::min<uint64_t>(kGbTile, n - t.tbase));
        const uint32_t lr0 = static_cast<uint32_t>(tid) * kGbRows;
        t.cnt = lr0 < tn ? std::min<uint32_t>(kGbRows, tn - lr0) : 0u;
        t.r0 = t.tbase + lr0;
        t.hb = head_bits<K>(keys, t.r0, t.cnt);
        const uint32_t hcount = __popc(t.hb);
        const uint32_t inc = wave_incl_scan(hcount);  // inclusive scan of head counts over the wave
        if (lane == 63) s_wsum[wave] = inc;
        lds_barrier();
        uint32_t pre_w = 0;
#pragma unroll
        for (int i = 0; i < kWaves; ++i) pre_w += (i < wave) ? s_wsum[i] : 0u;
        t.hx = pre_w + inc - hcount;
        t.H = __builtin_amdgcn_readfirstlane(s_wsum[0] + s_wsum[1] + s_wsum[2] + s_wsum[3]);
        if (tid == 0) store_relaxed(a.status + t.tile, (t.lt == 0 ? kFlagP : kFlagA) | t.H);
        return true;
    };

    // decoupled look-back over the table's tiles' head counts -> first group id.
    // Wave 0 reads the 64 previous tiles' status words at once (one round
    // trip; they are contiguous), takes the nearest one with an inclusive
    // prefix (lane L), and adds the aggregates of lanes 0..L.  A window with
    // an unpublished word before L is read again; a window without an
    // inclusive prefix is added whole and the next 64 tiles are read.  The
    // table's first tile publishes its inclusive prefix at once, so the walk
    // ends there at the latest.
    auto resolve = [&](GbTileState& t) {
        const GbSeg& sg = a.segs[t.sgi];
        if (wave == 0) {
            uint32_t excl = 0;
            if (t.lt > 0) {
                const int64_t first = static_cast<int64_t>(sg.gtile_off);
                int64_t k = static_cast<int64_t>(t.tile) - 1;
                for (;;) {
                    const int64_t idx = k - lane;
                    // lanes before the table's first tile: a zero prefix (never
                    // reached past the first tile's own inclusive prefix)
                    const uint32_t w = idx >= first ? load_relaxed(a.status + idx) : kFlagP;
                    const uint32_t f = w & ~kValMask;
                    const uint64_t pm = __ballot(f == kFlagP);
                    const uint64_t zm = __ballot(f == 0u);
                    const int L = pm ? __ffsll(static_cast<long long>(pm)) - 1 : 64;
                    const uint64_t upto = L >= 63 ? ~0ull : ((2ull << L) - 1);  // lanes 0..L
                    if (zm & upto) {
                        __builtin_amdgcn_s_sleep(1);
                        continue;
                    }
                    uint32_t part = lane <= L ? (w & kValMask) : 0u;
#pragma unroll
                    for (int off = 32; off > 0; off >>= 1) part += __shfl_xor(part, off, 64);
                    excl += part;
                    if (L < 64) break;
                    k -= 64;
                }
                if (lane == 0) store_relaxed(a.status + t.tile, kFlagP | (excl + t.H));
            }
            if (lane == 0) {
                s_gbase = excl;
                a.tile_gbase[t.tile] = excl;
                a.tile_heads[t.tile] = t.H;
                if (t.lt == sg.ntiles - 1) {
                    *gmem(sg.out.n_groups) = static_cast<uint64_t>(excl) + t.H;
                    if (excl + t.H <= sg.cap) gmem(sg.out.offsets)[excl + t.H] = static_cast<uint32_t>(sg.n);
                    if (a.group_total)
                        atomicAdd(a.group_total, static_cast<unsigned long long>(
                                                     std::min<uint64_t>(static_cast<uint64_t>(excl) + t.H, sg.cap)));
                }
            }
        }
        lds_barrier();
        t.gbase = __builtin_amdgcn_readfirstlane(s_gbase);
    };

    // group keys, offsets and every column's aggregates of a resolved tile
    auto finish = [&](const GbTileState& t) {
        const GbSeg& sg = a.segs[t.sgi];
        const K* keys = static_cast<const K*>(a.skeys) + sg.row_off;
        const uint32_t gfirst = t.gbase + t.hx;
        {
            // 4-byte keys: the lane's 16 keys come back with four 16-byte loads
            // (the lane's rows are 16-row aligned and inside the table's padded
            // rows), and each head picks its key from registers; wider keys
            // are read per head
            uint32_t kw[kGbRows];
            if constexpr (sizeof(K) == 4) {
                if (t.hb) {
                    const auto* q = gmem(reinterpret_cast<const u32x4*>(keys + t.r0));
#pragma unroll
                    for (int i = 0; i < kGbRows / 4; ++i) {
                        const u32x4 v = q[i];
                        kw[4 * i + 0] = v.x;
                        kw[4 * i + 1] = v.y;
                        kw[4 * i + 2] = v.z;
                        kw[4 * i + 3] = v.w;
                    }
                }
            }
            // offsets and keys of the tile's groups go through the LDS stage
            // (tile-local index h = g - gbase, in windows of kGbStage ids when
            // the tile has more groups) and out by consecutive lanes
            K* gkeys = static_cast<K*>(sg.out.group_keys);
            const uint32_t nheads = static_cast<uint32_t>(__popc(t.hb));
            for (uint32_t lo = 0; lo < t.H; lo += kGbStage) {  // block-uniform
                const uint32_t h0 = gfirst - t.gbase;  // this lane's first head's tile-local id
                if (nheads && h0 + nheads > lo && h0 < lo + kGbStage) {
                    uint32_t bits = t.hb;
                    uint32_t g = gfirst;
                    while (bits) {
                        const int k = __ffs(bits) - 1;
                        bits &= bits - 1;
                        const uint32_t h = g - t.gbase - lo;
                        ++g;
                        if (h >= static_cast<uint32_t>(kGbStage)) continue;  // another window's
                        const uint64_t r = t.r0 + k;
                        s_stage.cnt[h] = static_cast<uint32_t>(r);
                        if (gkeys) {
                            K kk;
                            if constexpr (sizeof(K) == 4) {
                                kk = kw[0];
#pragma unroll
                                for (int j = 1; j < kGbRows; ++j) kk = (k == j) ? kw[j] : kk;
                            } else {
                                kk = get_key(keys, r);
                            }
                            stage_key(&s_stage, h, kk);
                        }
                    }
                }
                lds_barrier();
                const uint32_t nst = std::min<uint32_t>(t.H - lo, kGbStage);
                for (uint32_t h = tid; h < nst; h += kBlock) {
                    const uint32_t gg = t.gbase + lo + h;
                    if (gg < sg.cap) {
                        gmem(sg.out.offsets)[gg] = s_stage.cnt[h];
                        if (gkeys) put_key(gkeys, gg, staged_key<K>(&s_stage, h));
                    }
                }
                lds_barrier();  // the stage is reused (next window, the columns)
            }
        }
        for (int c = 0; c < a.ncols; ++c) {
            switch (a.cols[c].kind) {
            case KIND_F64:
                gb_column<KIND_F64>(a, sg, c, t.tile, t.tbase, t.cnt, t.hb, gfirst, t.gbase, t.H, s_wagg, s_wflag, &s_stage);
                break;
            case KIND_SIGNED:
                gb_column<KIND_SIGNED>(a, sg, c, t.tile, t.tbase, t.cnt, t.hb, gfirst, t.gbase, t.H, s_wagg, s_wflag, &s_stage);
                break;
            case KIND_UNSIGNED:
                gb_column<KIND_UNSIGNED>(a, sg, c, t.tile, t.tbase, t.cnt, t.hb, gfirst, t.gbase, t.H, s_wagg, s_wflag, &s_stage);
                break;
            default:
                gb_column<KIND_COUNT>(a, sg, c, t.tile, t.tbase, t.cnt, t.hb, gfirst, t.gbase, t.H, s_wagg, s_wflag, &s_stage);
                break;
            }
        }
    };

    GbTileState c0, c1;
    if (!begin(c0)) return;
    resolve(c0);
    for (;;) {
        const bool has1 = begin(c1);
        finish(c0);
        if (!has1) break;
        resolve(c1);
        c0 = c1;
    }
}

// Joins the group that starts at tile b's last head with the leading
// partials of the table's following tiles up to (and including) the next tile
// that has a head.  One thread per tile: a group that ends in the next tile
// (or at the table's end) is finished here; a tile whose group spans tiles
// without heads is queued for k_gb_fixup_long.
__global__ __launch_bounds__(kBlock) void k_gb_fixup(GbArgs a)
{
    const uint32_t b = blockIdx.x * kBlock + threadIdx.x;
    if (b >= a.ntiles) return;
    const uint32_t heads = a.tile_heads[b];
    if (heads == 0) return;
    const uint32_t sgi = a.tile_seg ? a.tile_seg[b] : 0u;
    const GbSeg& sg = a.segs[sgi];
    const uint32_t tend = sg.gtile_off + sg.ntiles;  // one past the table's last tile
    const bool has_next = b + 1 < tend;
    if (has_next && a.tile_heads[b + 1] == 0) {
        gmem(a.long_list)[atomicAdd(a.long_count, 1u)] = b;
        return;
    }
    const uint32_t gid = a.tile_gbase[b] + heads - 1;
    for (int c = 0; c < a.ncols; ++c) {
        const int kind = a.cols[c].kind;
        Agg acc = a.suffix[static_cast<uint64_t>(c) * a.ntiles + b];
        if (has_next) acc = agg_join(acc, a.prefix[static_cast<uint64_t>(c) * a.ntiles + b + 1], kind);
        write_final(a, col_out<false>(a, sg, c), c, gid, acc);  // lanes differ in table
    }
}

// Groups that span whole tiles: a block per queued tile b walks the following
// tiles in windows of kBlock * kFixRows; thread i joins its kFixRows
// consecutive tiles in order, then the block joins the threads' partials by a
// fixed tree (lane pairs, then the four waves in order).
constexpr int kFixRows = 8;

template <int KIND>
__device__ inline Agg wave_join_tree(Agg x)
{
    // lane l joins lane l + off when l is a multiple of 2 * off: after the
    // steps lane 0 holds lanes 0..63 joined in lane order
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        Agg y;
        y.cnt = shfl_down64(x.cnt, off);
        y.a = shfl_down64(x.a, off);
        y.b = shfl_down64(x.b, off);
        y.mn = shfl_down64(x.mn, off);
        y.mx = shfl_down64(x.mx, off);
        if ((lane & (2 * off - 1)) == 0) x = agg_join(x, y, KIND);
    }
    return x;
}

template <int KIND>
__device__ inline void fix_window(const GbArgs& a, int c, uint32_t w0, uint32_t last, Agg* s_part, Agg& acc)
{
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const Agg* pre = a.prefix + static_cast<uint64_t>(c) * a.ntiles;
    Agg x = agg_empty();
    const uint32_t t0 = w0 + static_cast<uint32_t>(tid) * kFixRows;
    Agg p[kFixRows];  // every load issued before the joins (tiles past `last` read tile `last`)
#pragma unroll
    for (int k = 0; k < kFixRows; ++k) p[k] = pre[std::min(t0 + k, last)];
#pragma unroll
    for (int k = 0; k < kFixRows; ++k)
        if (t0 + k <= last) x = agg_join(x, p[k], KIND);
    x = wave_join_tree<KIND>(x);
    if (lane == 0) s_part[wave] = x;
    __syncthreads();
    if (tid == 0)
        for (int w = 0; w < kWaves; ++w) acc = agg_join(acc, s_part[w], KIND);
    __syncthreads();
}

__global__ __launch_bounds__(kBlock) void k_gb_fixup_long(GbArgs a)
{
    __shared__ Agg s_part[kWaves];
    __shared__ Agg s_acc[VDF_MAX_VALUE_COLS];  // the group's running joins (thread 0)
    __shared__ uint32_t s_stop;
    const int tid = threadIdx.x;
    const uint32_t nlong = *a.long_count;
    for (uint32_t e = blockIdx.x; e < nlong; e += gridDim.x) {
        const uint32_t b = a.long_list[e];
        const uint32_t sgi = a.tile_seg ? a.tile_seg[b] : 0u;
        const GbSeg& sg = a.segs[sgi];
        const uint32_t tend = sg.gtile_off + sg.ntiles;
        const uint32_t gid = a.tile_gbase[b] + a.tile_heads[b] - 1;
        if (tid == 0)
            for (int c = 0; c < a.ncols; ++c) s_acc[c] = a.suffix[static_cast<uint64_t>(c) * a.ntiles + b];
        for (uint32_t w0 = b + 1;; w0 += kBlock * kFixRows) {
            // the chain ends at the first tile with a head, or at the table's end
            if (tid == 0) s_stop = 0xFFFFFFFFu;
            __syncthreads();
            const uint32_t t0 = w0 + static_cast<uint32_t>(tid) * kFixRows;
            uint32_t hd[kFixRows];
#pragma unroll
            for (int k = 0; k < kFixRows; ++k) hd[k] = t0 + k < tend ? a.tile_heads[t0 + k] : 1u;
            uint32_t mine = 0xFFFFFFFFu;  // this thread's first stop
#pragma unroll
            for (int k = kFixRows - 1; k >= 0; --k)
                if (hd[k] > 0) mine = t0 + k;
            if (mine != 0xFFFFFFFFu) atomicMin(&s_stop, mine);
            __syncthreads();
            const uint32_t stop = s_stop;
            // tiles w0 .. last join the group (the stop tile with its leading rows)
            const uint32_t last = std::min(stop, tend - 1);
            for (int c = 0; c < a.ncols; ++c) {
                switch (a.cols[c].kind) {
                case KIND_F64: fix_window<KIND_F64>(a, c, w0, last, s_part, s_acc[c]); break;
                case KIND_SIGNED: fix_window<KIND_SIGNED>(a, c, w0, last, s_part, s_acc[c]); break;
                case KIND_UNSIGNED: fix_window<KIND_UNSIGNED>(a, c, w0, last, s_part, s_acc[c]); break;
                default: fix_window<KIND_COUNT>(a, c, w0, last, s_part, s_acc[c]); break;
                }
            }
            if (stop != 0xFFFFFFFFu) break;
        }
        if (tid == 0)
            for (int c = 0; c < a.ncols; ++c) write_final(a, col_out(a, sg, c), c, gid, s_acc[c]);
        __syncthreads();  // s_acc is set again for the next queued tile
    }
}

// Tables without rows: zero groups.
__global__ void k_gb_empty(const GbSeg* segs, uint32_t nsegs)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < nsegs && segs[i].n == 0) {
        *gmem(segs[i].out.n_groups) = 0;
        gmem(segs[i].out.offsets)[0] = 0;
    }
}

// Resident blocks of a tile kernel on a device (cached per kernel and device).
unsigned gb_resident_blocks(const void* fn, int device)
{
    struct Entry {
        const void* fn;
        int device;
        unsigned blocks;
    };
    static std::mutex mu;
    static std::vector<Entry> cache;
    std::lock_guard<std::mutex> lk(mu);
    for (const Entry& e : cache)
        if (e.fn == fn && e.device == device) return e.blocks;
    int per_cu = 0, cus = 0;
    VDF_HIP_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, kBlock, 0));
    VDF_HIP_CHECK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device));
    const unsigned blocks = static_cast<unsigned>(std::max(1, per_cu) * std::max(1, cus));
    cache.push_back(Entry{fn, device, blocks});
    return blocks;
}

int kind_of(int type)
{
    switch (type) {
    case VDF_VAL_F64: return KIND_F64;
    case VDF_VAL_I64:
    case VDF_VAL_I32: return KIND_SIGNED;
    default: return KIND_UNSIGNED;
    }
}

struct GbLayout {
    uint64_t rows = 0, tiles = 0;
    size_t sort_bytes = 0, agg_bytes = 0, bytes = 0;
};

// scratch of the aggregate stage alone (descriptors, tile map, look-back,
// per-tile partials)
size_t agg_bytes_for(uint64_t tiles, int ntables)
{
    size_t b = 0;
    b += align_up(static_cast<size_t>(ntables) * sizeof(GbSeg), 256);
    b += align_up(std::max<uint64_t>(tiles, 1) * 4, 256);      // tile -> table
    b += align_up((tiles + 2) * 4, 256);                       // status + ticket + long_count
    b += 3 * align_up(std::max<uint64_t>(tiles, 1) * 4, 256);  // tile_gbase, tile_heads, long_list
    b += 2 * align_up(VDF_MAX_VALUE_COLS * std::max<uint64_t>(tiles, 1) * sizeof(Agg), 256);
    b += 256;  // group total (profiling)
    return b + 256;
}

GbLayout gb_layout(int key_type, const uint64_t* n, int ntables, uint64_t hist_rows, int pay_bytes)
{
    GbLayout L;
    for (int i = 0; i < ntables; ++i) {
        L.rows += align_up(n[i], kGbRows);  // table starts 16-row aligned (vector loads)
        L.tiles += (n[i] + kGbTile - 1) / kGbTile;
    }
    L.sort_bytes = align_up(argsort_batch_scratch(key_type, n, ntables, hist_rows, pay_bytes), 256);
    L.agg_bytes = agg_bytes_for(L.tiles, ntables);
    size_t b = L.sort_bytes;
    b += align_up(L.rows * key_bytes(key_type), 256);                   // sorted keys
    b += align_up(L.rows * 4, 256);                                     // perms the caller does not take
    b += align_up(L.rows * static_cast<uint64_t>(pay_bytes), 256);      // carried columns in key order
    b += L.agg_bytes;
    L.bytes = b + 256;
    return L;
}

}  // namespace

// Value columns a group-by carries through its sort into key order (the
// aggregate kernel then reads them sequentially instead of gathering them
// through the permutation, a random 8-byte read per row and column that costs
// a whole memory line): column 0 when it is 4 or 8 bytes, columns 0 and 1 when
// both are 8 bytes, or (4-byte keys) when column 0 is 8 bytes and column 1 is
// 4 bytes.  Returns the payload bytes per row (0: gather everything).
int groupby_carry_bytes(const Tuning& t, int key_type, const int* vt, int nv)
{
    if (!t.gb_carry || nv <= 0) return 0;
    const int w0 = val_bytes(vt[0]);
    if (nv >= 2 && w0 == 8 && val_bytes(vt[1]) == 8) return 16;
    if (nv >= 2 && w0 == 8 && val_bytes(vt[1]) == 4 && key_bytes(key_type) == 4) return 12;
    return w0;
}

// (the budget query assumes the largest payload)
size_t groupby_scratch(int key_type, uint64_t n) { return gb_layout(key_type, &n, 1, kDefaultHistRows, 16).bytes; }

size_t groupby_batch_scratch(int key_type, const uint64_t* n, int ntables, uint64_t hist_rows, int pay_bytes)
{
    return gb_layout(key_type, n, ntables, hist_rows, pay_bytes).bytes;
}

size_t groupby_agg_staging(const uint64_t* n, int ntables)
{
    uint64_t tiles = 0;
    for (int i = 0; i < ntables; ++i) tiles += (n[i] + kGbTile - 1) / kGbTile;
    return 512 + align_up(static_cast<size_t>(ntables) * sizeof(GbSeg), 256) + std::max<uint64_t>(tiles, 1) * 4;
}

size_t groupby_agg_scratch(const uint64_t* n, int ntables)
{
    uint64_t tiles = 0;
    for (int i = 0; i < ntables; ++i) tiles += (n[i] + kGbTile - 1) / kGbTile;
    return agg_bytes_for(tiles, ntables);
}

size_t groupby_batch_staging(int key_type, const uint64_t* n, int ntables, uint64_t hist_rows)
{
    return argsort_batch_staging(key_type, n, ntables, hist_rows) + groupby_agg_staging(n, ntables);
}

// The aggregate stage over tables already sorted: sorted keys of every table
// at `skeys` (table starts 16-row aligned), each table's permutation in
// perms[i]; scratch at `scr` (groupby_agg_scratch bytes).
void groupby_aggregate_enqueue(Handle* h, int key_type, const void* skeys, const uint32_t* const* perms,
                               const GbTable* tables, int ntables, const int* val_types, int nvals, uint32_t agg_mask,
                               const vdf_groupby_out* outs, void* scr, hipStream_t s, const void* spay,
                               int pay_bytes)
{
    const size_t kb = key_bytes(key_type);
    uint64_t tiles_total = 0, rows_total = 0;
    for (int i = 0; i < ntables; ++i) {
        tiles_total += (tables[i].n + kGbTile - 1) / kGbTile;
        rows_total += align_up(tables[i].n, kGbRows);
    }
    Carve cv(scr);
    GbSeg* dsegs = cv.take<GbSeg>(ntables);
    uint32_t* dtile = cv.take<uint32_t>(std::max<uint64_t>(tiles_total, 1));
    uint32_t* status = cv.take<uint32_t>(tiles_total + 2);
    GbArgs a{};
    a.tile_gbase = cv.take<uint32_t>(std::max<uint64_t>(tiles_total, 1));
    a.tile_heads = cv.take<uint32_t>(std::max<uint64_t>(tiles_total, 1));
    a.long_list = cv.take<uint32_t>(std::max<uint64_t>(tiles_total, 1));
    a.prefix = cv.take<Agg>(VDF_MAX_VALUE_COLS * std::max<uint64_t>(tiles_total, 1));
    a.suffix = cv.take<Agg>(VDF_MAX_VALUE_COLS * std::max<uint64_t>(tiles_total, 1));
    unsigned long long* gtotal = cv.take<unsigned long long>(1);

    // table descriptions and the tile map (pinned staging -> device)
    const size_t seg_bytes = align_up(static_cast<size_t>(ntables) * sizeof(GbSeg), 256);
    char* stg = static_cast<char*>(h->staging_take(seg_bytes + std::max<uint64_t>(tiles_total, 1) * 4));
    GbSeg* hs = reinterpret_cast<GbSeg*>(stg);
    uint32_t* ht = reinterpret_cast<uint32_t*>(stg + seg_bytes);
    uint64_t row = 0, tiles = 0;
    for (int i = 0; i < ntables; ++i) {
        const uint64_t n = tables[i].n;
        const uint64_t nt = (n + kGbTile - 1) / kGbTile;
        GbSeg g{};
        g.n = n;
        g.row_off = row;
        g.perm = perms[i];
        g.gtile_off = static_cast<uint32_t>(tiles);
        g.ntiles = static_cast<uint32_t>(nt);
        g.cap = (outs[i].capacity == 0 || outs[i].capacity > n) ? n : outs[i].capacity;
        for (int c = 0; c < nvals; ++c) g.vals[c] = tables[i].vals[c];
        if (spay != nullptr && pay_bytes != 0) {
            // carried columns: 0 (4 or 8 bytes per row), and with a 12- or
            // 16-byte payload 1 (4 or 8 bytes per row) after column 0's rows
            const char* base = static_cast<const char*>(spay);
            const uint64_t w0 = pay_bytes >= 12 ? 8 : static_cast<uint64_t>(pay_bytes);
            g.svals[0] = base + row * w0;
            if (pay_bytes >= 12) g.svals[1] = base + rows_total * 8 + row * static_cast<uint64_t>(pay_bytes - 8);
        }
        g.out = outs[i];
        hs[i] = g;
        for (uint64_t t = 0; t < nt; ++t) ht[tiles + t] = static_cast<uint32_t>(i);
        row += align_up(n, kGbRows);
        tiles += nt;
    }
    if (h->dry) return;  // a plan refresh only wants the staging
    // dsegs and dtile are consecutive carves with the same alignment as the staging layout
    VDF_HIP_CHECK(hipMemcpyAsync(dsegs, stg, seg_bytes + tiles * 4, hipMemcpyHostToDevice, s));
    if (static_cast<char*>(static_cast<void*>(dtile)) != reinterpret_cast<char*>(dsegs) + seg_bytes)
        throw Error{VDF_ERR_HIP, "group-by scratch layout mismatch"};

    {
        LaunchTimer t(h, s, "groupby_empty", 0);
        hipLaunchKernelGGL(k_gb_empty, dim3((ntables + 255) / 256), dim3(256), 0, s, dsegs,
                           static_cast<uint32_t>(ntables));
        check_launch("k_gb_empty");
    }
    if (tiles == 0) return;

    a.segs = dsegs;
    a.tile_seg = ntables > 1 ? dtile : nullptr;
    a.skeys = skeys;
    a.ntiles = static_cast<uint32_t>(tiles);
    a.agg_mask = agg_mask;
    a.status = status;
    a.ticket = status + tiles;
    a.long_count = status + tiles + 1;
    uint64_t val_bytes = 0;
    if (nvals == 0) {
        a.ncols = 1;
        a.cols[0] = GbColType{VDF_VAL_U32, KIND_COUNT};
    } else {
        a.ncols = nvals;
        const int ncarried = pay_bytes >= 12 ? 2 : (pay_bytes != 0 ? 1 : 0);
        for (int c = 0; c < nvals; ++c) {
            a.cols[c] = GbColType{val_types[c], kind_of(val_types[c])};
            // a carried column is read in order; a gathered one costs its row id too
            val_bytes += (c < ncarried ? 0 : 4) + ((val_types[c] == VDF_VAL_I32 || val_types[c] == VDF_VAL_U32) ? 4 : 8);
        }
    }
    // group outputs per group (of table 0: every table of a batch asks for the same ones)
    uint64_t group_bytes = 0;
    {
        const vdf_groupby_out& o = outs[0];
        group_bytes += (o.group_keys ? kb : 0) + (o.offsets ? 4 : 0) + (o.counts ? 4 : 0);
        for (int c = 0; c < nvals; ++c) {
            const uint64_t vb = (val_types[c] == VDF_VAL_I32 || val_types[c] == VDF_VAL_U32) ? 4 : 8;
            group_bytes += (o.aggs[c].sum ? 8 : 0) + (o.aggs[c].mean ? 8 : 0) + (o.aggs[c].min ? vb : 0) +
                           (o.aggs[c].max ? vb : 0);
        }
    }
    const bool count_groups = h->prof && h->passes_ring;
    a.group_total = count_groups ? gtotal : nullptr;
    {
        LaunchTimer t(h, s, "fill", 0);
        fill_u32(status, tiles + 2, 0u, s);
        if (count_groups) fill_u32(reinterpret_cast<uint32_t*>(gtotal), 2, 0u, s);
    }
    {
        // algorithmic bytes: sorted keys + per column its value (and row id when
        // gathered), plus (profiling: counted on the device) the group outputs
        LaunchTimer t(h, s, "groupby_tile", rows_total * (kb + val_bytes));
        // persistent grid: the resident blocks (tiles are claimed dynamically)
        const void* fn = kb == 4 ? reinterpret_cast<const void*>(k_gb_tile<uint32_t>)
                                 : (kb == 8 ? reinterpret_cast<const void*>(k_gb_tile<uint64_t>)
                                            : reinterpret_cast<const void*>(k_gb_tile<K16>));
        const unsigned grid = static_cast<unsigned>(std::min<uint64_t>(a.ntiles, gb_resident_blocks(fn, h->device)));
        switch (kb) {
        case 4: hipLaunchKernelGGL(k_gb_tile<uint32_t>, dim3(grid), dim3(kBlock), 0, s, a); break;
        case 8: hipLaunchKernelGGL(k_gb_tile<uint64_t>, dim3(grid), dim3(kBlock), 0, s, a); break;
        default: hipLaunchKernelGGL(k_gb_tile<K16>, dim3(grid), dim3(kBlock), 0, s, a); break;
        }
        check_launch("k_gb_tile");
    }
    if (count_groups) {
        // the group count comes back into the pinned ring; the tile launch is
        // charged the group outputs of every table (the fix-up writes some of them)
        if (h->ring_pos % 2) ++h->ring_pos;
        if (h->ring_pos + 2 > Handle::kRing) h->ring_pos = 0;
        uint64_t* dst = reinterpret_cast<uint64_t*>(h->passes_ring + h->ring_pos);
        h->ring_pos += 2;
        VDF_HIP_CHECK(hipMemcpyAsync(dst, gtotal, 8, hipMemcpyDeviceToHost, s));
        hipEvent_t done = h->get_event();
        VDF_HIP_CHECK(hipEventRecord(done, s));
        h->recs.back().groups = dst;
        h->recs.back().groups_done = done;
        h->recs.back().group_bytes = group_bytes;
    }
    {
        LaunchTimer t(h, s, "groupby_fixup", 0);
        hipLaunchKernelGGL(k_gb_fixup, dim3((a.ntiles + kBlock - 1) / kBlock), dim3(kBlock), 0, s, a);
        check_launch("k_gb_fixup");
        // queued long chains (their count is on the device): blocks take
        // queue entries in turn
        hipLaunchKernelGGL(k_gb_fixup_long, dim3(std::min<uint32_t>(a.ntiles, 1024)), dim3(kBlock), 0, s, a);
        check_launch("k_gb_fixup_long");
    }
}

void groupby_batch_enqueue(Handle* h, int key_type, const GbTable* tables, int ntables, const int* val_types,
                           int nvals, uint32_t agg_mask, const vdf_groupby_out* outs, void* scr, hipStream_t s)
{
    std::vector<uint64_t> ns(ntables);
    for (int i = 0; i < ntables; ++i) ns[i] = tables[i].n;
    const uint64_t hist_rows = static_cast<uint64_t>(h->tune.hist_rows);
    const int pay = groupby_carry_bytes(h->tune, key_type, val_types, nvals);
    const GbLayout L = gb_layout(key_type, ns.data(), ntables, hist_rows, pay);
    const size_t kb = key_bytes(key_type);
    Carve cv(scr);
    void* sort_scr = cv.take<char>(L.sort_bytes);
    char* skeys = cv.take<char>(L.rows * kb);
    uint32_t* perm_scr = cv.take<uint32_t>(L.rows);
    char* spay = cv.take<char>(L.rows * static_cast<uint64_t>(pay));
    void* agg_scr = cv.take<char>(L.agg_bytes);

    // 1. sort every table: sorted keys into scratch, perms to the caller or
    //    scratch, carried value columns into key order (16-byte payload: the
    //    second column in the upper half of spay)
    std::vector<SortSeg> ss(ntables);
    std::vector<const uint32_t*> perms(ntables);
    const uint64_t w0 = pay >= 12 ? 8 : static_cast<uint64_t>(pay);
    // The permutation is needed for gathered columns and for callers that
    // take it; otherwise the passes do not move row ids at all.
    const int ncarried = pay >= 12 ? 2 : (pay != 0 ? 1 : 0);
    bool rowids = h->keep_rowids || nvals > ncarried;
    for (int i = 0; i < ntables; ++i) rowids = rowids || outs[i].perm != nullptr;
    uint64_t row = 0;
    for (int i = 0; i < ntables; ++i) {
        uint32_t* p = outs[i].perm ? outs[i].perm : (rowids ? perm_scr + row : nullptr);
        perms[i] = p;
        ss[i] = SortSeg{tables[i].keys, tables[i].n, p, skeys + row * kb};
        if (pay != 0) {
            ss[i].pay_in = tables[i].vals[0];
            ss[i].pay_out = spay + row * w0;
        }
        if (pay >= 12) {
            ss[i].pay_in2 = tables[i].vals[1];
            ss[i].pay_out2 = spay + L.rows * 8 + row * static_cast<uint64_t>(pay - 8);
        }
        row += align_up(tables[i].n, kGbRows);
    }
    argsort_batch_enqueue(h, key_type, ss.data(), ntables, 0, sort_scr, s, pay, rowids);

    // 2. boundaries and aggregates
    groupby_aggregate_enqueue(h, key_type, skeys, perms.data(), tables, ntables, val_types, nvals, agg_mask, outs,
                              agg_scr, s, pay != 0 ? spay : nullptr, pay);
}

void groupby_enqueue(Handle* h, int key_type, const void* d_keys, uint64_t n, const vdf_value_col* vals,
                     int nvals, uint32_t agg_mask, const vdf_groupby_out* out, void* scr, hipStream_t s)
{
    GbTable t{};
    t.keys = d_keys;
    t.n = n;
    int types[VDF_MAX_VALUE_COLS] = {0, 0, 0, 0};
    for (int c = 0; c < nvals; ++c) {
        t.vals[c] = vals[c].data;
        types[c] = vals[c].type;
    }
    groupby_batch_enqueue(h, key_type, &t, 1, types, nvals, agg_mask, out, scr, s);
}

}  // namespace vdf
