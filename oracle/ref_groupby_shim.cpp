// ref_groupby_shim.cpp -- the reference's hash group-by, timed as the CPU
// baseline of the group-by workloads.
//
// TEST INFRASTRUCTURE ONLY (bench.py's cpu_baseline legs and tests/).  Built
// by oracle/Makefile (target `ref`) into oracle/_ref/ with the reference's
// benchmark flags, only where the reference tree exists; never part of the
// product.  No reference source is copied.
//
// The reference's group-by sources do not compile (syntax errors in
// onecol_functions.inl:5-7,24 and transform_group_by.hpp:12, undefined
// identifiers), but the two components its hot loop spends its time in do,
// and are included here where they lie:
//   - the hash map, ankerl::unordered_dense::map (external/ankerl/
//     unordered_dense.h, v4.8.0), keyed by std::string_view over the key
//     bytes (transform_group_by_onecol_mt.hpp:42,57-105);
//   - the key hash, simd_hash (custom_simd_hash/simd_hash.hpp:134-144; with
//     AVX2 it calls hash_avx2, whose <= 16-byte path equals hash_scalar),
//     for keys of up to 8 and of more than 16 bytes.  For 9 to 16 bytes its
//     rd64_tail (:18-22) copies the whole key into one 8-byte word, which
//     overruns the stack (the process aborts under the stack protector), so
//     there the second word is folded in with one more mix64, as the
//     oracle's orc_hash_bytes16 does: the same work per key.
// The loops around them are restated from the reference:
//   - per row: try_emplace(string_view{&key[i], sizeof key}, zero) and the
//     aggregate update, in row order, recording key_vec[i] = &it->first
//     (onecol_functions.inl:34-40 Occurence, :77-83 Add, :120-126 Fill);
//   - CORES > 1: one map per thread over contiguous row chunks of
//     nrow / CORES + 1 rows (transform_group_by_onecol_mt.hpp:226-233),
//     merged in thread order by try_emplace + += or vector append
//     (merge_functions.inl:3-36);
//   - Gather reserves NPerGroup = 4 values per group (ReservingVec,
//     types/ReservingVec.inl:4-13) and min / max is the user functor over the
//     row-ordered vector (create_value_col.inl:48-58).
// Two forms are exported:
//   refapi = 1: the reference API's cost for count, sum, mean, min and max of
//     every value column: one full hash pass (with merge) per aggregate per
//     column, Gather + functor for min and for max;
//   refapi = 0: a fused restatement, one hash pass that keeps every aggregate
//     of every column per group (the stronger CPU baseline).
// Integer sums wrap modulo 2^64 (the reference's would overflow its type).
#include <algorithm>
#include <cstddef>
#include <cstdint>
#include <cstring>
#include <immintrin.h>
#include <omp.h>
#include <string_view>
#include <vector>

#include "external/ankerl/unordered_dense.h"

namespace vref {
#include "custom_simd_hash/simd_hash.hpp"
}  // namespace vref

namespace {

using Key = std::string_view;

// simd_hash, except for 9- to 16-byte keys (see above)
struct KeyHash {
    using is_avalanching = void;
    uint64_t operator()(Key s) const noexcept
    {
        if (s.size() <= 8 || s.size() > 16) return vref::simd_hash{}(s);
        uint64_t v = 0, w = 0;
        std::memcpy(&v, s.data(), 8);
        std::memcpy(&w, s.data() + 8, s.size() - 8);
        return vref::mix64(v ^ vref::mix64(w) ^ 0x9e3779b97f4a7c15ULL);
    }
};

template <typename V>
using Map = ankerl::unordered_dense::map<Key, V, KeyHash>;

// value column kinds
enum { V_NONE = -1, V_F64 = 0, V_I64 = 1, V_I32 = 2 };

struct Cols {
    const char* keys;
    size_t kb;
    size_t n;
    const double* v0;   // float64 column 0
    int k1;             // kind of column 1 (V_NONE, V_I64, V_I32)
    const void* v1;
    int threads;
};

inline Key key_at(const Cols& c, size_t i) { return Key{c.keys + i * c.kb, c.kb}; }

inline int64_t v1_at(const Cols& c, size_t i)
{
    return c.k1 == V_I64 ? static_cast<const int64_t*>(c.v1)[i] : static_cast<const int32_t*>(c.v1)[i];
}

// Runs body(map, begin, end) on per-thread maps over the reference's row
// chunks, then merge(lookup, key, value) of each thread's map in thread order.
template <typename V, typename Body, typename Merge>
Map<V> threaded_pass(const Cols& c, Body body, Merge merge)
{
    const int T = std::max(c.threads, 1);
    if (T == 1) {
        Map<V> m;
        body(m, size_t(0), c.n);
        return m;
    }
    std::vector<Map<V>> parts(static_cast<size_t>(T));
    const size_t chunks = c.n / static_cast<size_t>(T) + 1;
#pragma omp parallel num_threads(T)
    {
        const size_t t = static_cast<size_t>(omp_get_thread_num());
        const size_t start = std::min(c.n, t * chunks);
        const size_t end = std::min(c.n, start + chunks);
        body(parts[t], start, end);
    }
    Map<V> lookup;
    for (auto& p : parts)
        for (const auto& kv : p) merge(lookup, kv.first, kv.second);
    return lookup;
}

// ---- fused restatement: every aggregate of every column in one pass ----
struct Agg {
    uint64_t cnt = 0;
    double s0 = 0, mn0 = 0, mx0 = 0;
    uint64_t s1 = 0;
    int64_t mn1 = 0, mx1 = 0;
};

Map<Agg> fused(const Cols& c)
{
    auto body = [&](Map<Agg>& m, size_t b, size_t e) {
        for (size_t i = b; i < e; ++i) {
            const double x = c.v0[i];
            auto [it, ins] = m.try_emplace(key_at(c, i));
            Agg& a = it->second;
            if (ins) a.mn0 = a.mx0 = x;
            ++a.cnt;
            a.s0 += x;
            if (x < a.mn0) a.mn0 = x;
            if (x > a.mx0) a.mx0 = x;
            if (c.k1 != V_NONE) {
                const int64_t y = v1_at(c, i);
                if (ins) a.mn1 = a.mx1 = y;
                a.s1 += static_cast<uint64_t>(y);
                if (y < a.mn1) a.mn1 = y;
                if (y > a.mx1) a.mx1 = y;
            }
        }
    };
    auto merge = [&](Map<Agg>& l, Key k, const Agg& v) {
        auto [it, ins] = l.try_emplace(k, v);
        if (ins) return;
        Agg& a = it->second;
        a.cnt += v.cnt;
        a.s0 += v.s0;
        if (v.mn0 < a.mn0) a.mn0 = v.mn0;
        if (v.mx0 > a.mx0) a.mx0 = v.mx0;
        a.s1 += v.s1;
        if (v.mn1 < a.mn1) a.mn1 = v.mn1;
        if (v.mx1 > a.mx1) a.mx1 = v.mx1;
    };
    return threaded_pass<Agg>(c, body, merge);
}

// ---- the reference API: one pass per aggregate ----
template <typename T>
struct Gathered {
    std::vector<T> v;
    Gathered() { v.reserve(4); }  // NPerGroup
};

template <typename T, typename Get>
Map<T> sum_pass(const Cols& c, Get get, std::vector<const Key*>& key_vec)
{
    auto body = [&](Map<T>& m, size_t b, size_t e) {
        for (size_t i = b; i < e; ++i) {
            auto [it, ins] = m.try_emplace(key_at(c, i), T{});
            it->second += get(i);
            key_vec[i] = &it->first;
        }
    };
    auto merge = [](Map<T>& l, Key k, const T& v) { l.try_emplace(k, T{}).first->second += v; };
    return threaded_pass<T>(c, body, merge);
}

template <typename T, typename Get>
Map<Gathered<T>> gather_pass(const Cols& c, Get get, std::vector<const Key*>& key_vec)
{
    auto body = [&](Map<Gathered<T>>& m, size_t b, size_t e) {
        for (size_t i = b; i < e; ++i) {
            auto [it, ins] = m.try_emplace(key_at(c, i));
            it->second.v.push_back(get(i));
            key_vec[i] = &it->first;
        }
    };
    auto merge = [](Map<Gathered<T>>& l, Key k, const Gathered<T>& v) {
        auto& dst = l.try_emplace(k).first->second.v;
        dst.insert(dst.end(), v.v.begin(), v.v.end());
    };
    return threaded_pass<Gathered<T>>(c, body, merge);
}

}  // namespace

extern "C" {

// The reference's hash group-by over one key column of kb bytes per row
// (raw bytes, as its string_view keys see them), a float64 column v0 and an
// optional second column (k1: -1 none, 1 int64, 2 int32).  Returns the group
// count; when the out arrays are non-null (capacity >= groups) writes, in
// the reference's iteration order (first occurrence after the thread-order
// merge): keys, counts, sum / min / max of v0, and (with a second column)
// the wrapped sum / min / max of v1 as int64.
uint64_t vref_hash_groupby(const void* keys, size_t kb, size_t n, const double* v0, int k1, const void* v1,
                           int threads, int refapi, void* out_keys, uint64_t* out_cnt, double* out_s0,
                           double* out_mn0, double* out_mx0, int64_t* out_s1, int64_t* out_mn1, int64_t* out_mx1)
{
    const Cols c{static_cast<const char*>(keys), kb, n, v0, k1, v1, threads};
    if (!refapi) {
        const Map<Agg> m = fused(c);
        size_t g = 0;
        for (const auto& kv : m) {
            if (out_keys) std::memcpy(static_cast<char*>(out_keys) + g * kb, kv.first.data(), kb);
            if (out_cnt) out_cnt[g] = kv.second.cnt;
            if (out_s0) out_s0[g] = kv.second.s0;
            if (out_mn0) out_mn0[g] = kv.second.mn0;
            if (out_mx0) out_mx0[g] = kv.second.mx0;
            if (out_s1) out_s1[g] = static_cast<int64_t>(kv.second.s1);
            if (out_mn1) out_mn1[g] = kv.second.mn1;
            if (out_mx1) out_mx1[g] = kv.second.mx1;
            ++g;
        }
        return m.size();
    }
    std::vector<const Key*> key_vec(n);
    // Occurence
    const Map<uint32_t> occ = sum_pass<uint32_t>(c, [](size_t) { return 1u; }, key_vec);
    // Sum and Mean of v0 (Mean runs the same pass; the reference divides at the broadcast)
    const Map<double> s0 = sum_pass<double>(c, [&](size_t i) { return v0[i]; }, key_vec);
    const Map<double> m0 = sum_pass<double>(c, [&](size_t i) { return v0[i]; }, key_vec);
    // Gather + functor for min, then for max
    const auto gmin0 = gather_pass<double>(c, [&](size_t i) { return v0[i]; }, key_vec);
    const auto gmax0 = gather_pass<double>(c, [&](size_t i) { return v0[i]; }, key_vec);
    Map<uint64_t> s1, m1;
    Map<Gathered<int64_t>> gmin1, gmax1;
    if (k1 != V_NONE) {
        auto y = [&](size_t i) { return static_cast<uint64_t>(v1_at(c, i)); };
        auto z = [&](size_t i) { return v1_at(c, i); };
        s1 = sum_pass<uint64_t>(c, y, key_vec);
        m1 = sum_pass<uint64_t>(c, y, key_vec);
        gmin1 = gather_pass<int64_t>(c, z, key_vec);
        gmax1 = gather_pass<int64_t>(c, z, key_vec);
    }
    (void)m0;
    (void)m1;
    size_t g = 0;
    for (const auto& kv : occ) {
        const Key k = kv.first;
        if (out_keys) std::memcpy(static_cast<char*>(out_keys) + g * kb, k.data(), kb);
        if (out_cnt) out_cnt[g] = kv.second;
        if (out_s0) out_s0[g] = s0.at(k);
        const auto& a = gmin0.at(k).v;
        const auto& b = gmax0.at(k).v;
        if (out_mn0) out_mn0[g] = *std::min_element(a.begin(), a.end());
        if (out_mx0) out_mx0[g] = *std::max_element(b.begin(), b.end());
        if (k1 != V_NONE) {
            const auto& p = gmin1.at(k).v;
            const auto& q = gmax1.at(k).v;
            if (out_s1) out_s1[g] = static_cast<int64_t>(s1.at(k));
            if (out_mn1) out_mn1[g] = *std::min_element(p.begin(), p.end());
            if (out_mx1) out_mx1[g] = *std::max_element(q.begin(), q.end());
        }
        ++g;
    }
    return occ.size();
}

// the reference's simd_hash of a key of len bytes (len <= 8 or > 16: the
// lengths it handles without overrunning), for pinning the oracle's
// restatement of it
uint64_t vref_simd_hash(const void* p, size_t len)
{
    if (len > 8 && len <= 16) return 0;
    return vref::simd_hash{}(std::string_view{static_cast<const char*>(p), len});
}

}  // extern "C"
