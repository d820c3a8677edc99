/*
 * oracle.c -- CPU restatement of VisonDF's radix-argsort / gather / group-by path.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP library and the timed CPU
 * baseline.  Nothing in the product links or loads it.
 *
 * Reference files are cited relative to include/vison/ of the reference tree.
 *
 * Pinning (each function is checked by tests/test_oracle.py against something
 * other than itself):
 *   orc_radix_sort_*     numpy's stable argsort; golden vectors produced by
 *                        the reference's own radix .inl files (oracle/_ref);
 *                        the INTERNALS.doc:1554-1631 sort_by example
 *   orc_argsort modes    numpy stable argsort of the order-preserving map;
 *                        INTERNALS.doc descending example; explicit reversal
 *   orc_radix_sort_charbuf  Python sorted() over bytes (stable) + golden vectors
 *   orc_gather           numpy.take
 *   orc_exact_sum        math.fsum (Shewchuk), closed-form dyadic sums
 *   orc_groupby          pandas groupby(sort=True), brute-force dict on tiny
 *                        inputs, exact integer sums through Python ints
 *   orc_hash_groupby_f64 orc_groupby after reordering (same groups/counts/
 *                        min/max; sums within row-order rounding)
 */
#include "oracle.h"

#include <math.h>
#include <omp.h>
#include <stdlib.h>
#include <string.h>

#define RADIX16 65536u
#define MAX_THREADS 32 /* details/constants.inl:9 */

/* ------------------------------------------------------------------------ */
/* Serial 16-bit-digit LSD radix, radix_sort_uint32.inl:5-99 (Simd=false).  */
/* Per pass: histogram (:55-57), exclusive scan (:62-67), stable scatter of  */
/* (key, idx) (:85-92), keys swapped and idx copied back (:95-96).           */
/* ------------------------------------------------------------------------ */
void orc_radix_sort_u32(uint32_t* tkeys, uint64_t* idx, size_t n)
{
    if (n == 0) return; /* radix_sort_uint32.inl:15-18 warns and returns */
    uint32_t* tk = (uint32_t*)malloc(n * sizeof(uint32_t));
    uint64_t* ti = (uint64_t*)malloc(n * sizeof(uint64_t));
    uint64_t* cnt = (uint64_t*)malloc(RADIX16 * sizeof(uint64_t));
    uint32_t* src = tkeys;
    uint32_t* dst = tk;
    for (unsigned pass = 0; pass < 2; ++pass) {
        const unsigned sh = 16u * pass;
        memset(cnt, 0, RADIX16 * sizeof(uint64_t));
        for (size_t i = 0; i < n; ++i) cnt[(src[i] >> sh) & 0xFFFFu]++;
        uint64_t run = 0;
        for (size_t b = 0; b < RADIX16; ++b) { uint64_t c = cnt[b]; cnt[b] = run; run += c; }
        for (size_t i = 0; i < n; ++i) {
            uint32_t k = src[i];
            uint64_t p = cnt[(k >> sh) & 0xFFFFu]++;
            ti[p] = idx[i];
            dst[p] = k;
        }
        memcpy(idx, ti, n * sizeof(uint64_t));
        uint32_t* t = src; src = dst; dst = t;
    }
    /* two passes: the sorted keys are back in the caller's buffer */
    free(tk); free(ti); free(cnt);
}

/* Serial u64 variant, radix_sort_uint64.inl:5-85: four 16-bit passes (:26). */
void orc_radix_sort_u64(uint64_t* tkeys, uint64_t* idx, size_t n)
{
    if (n == 0) return;
    uint64_t* tk = (uint64_t*)malloc(n * sizeof(uint64_t));
    uint64_t* ti = (uint64_t*)malloc(n * sizeof(uint64_t));
    uint64_t* cnt = (uint64_t*)malloc(RADIX16 * sizeof(uint64_t));
    uint64_t* src = tkeys;
    uint64_t* dst = tk;
    for (unsigned pass = 0; pass < 4; ++pass) {
        const unsigned sh = 16u * pass;
        memset(cnt, 0, RADIX16 * sizeof(uint64_t));
        for (size_t i = 0; i < n; ++i) cnt[(src[i] >> sh) & 0xFFFFu]++;
        uint64_t run = 0;
        for (size_t b = 0; b < RADIX16; ++b) { uint64_t c = cnt[b]; cnt[b] = run; run += c; }
        for (size_t i = 0; i < n; ++i) {
            uint64_t k = src[i];
            uint64_t p = cnt[(k >> sh) & 0xFFFFu]++;
            ti[p] = idx[i];
            dst[p] = k;
        }
        memcpy(idx, ti, n * sizeof(uint64_t));
        uint64_t* t = src; src = dst; dst = t;
    }
    free(tk); free(ti); free(cnt);
}

/* ------------------------------------------------------------------------ */
/* Multithreaded variant, radix_sort_uint32_mt.inl:6-164 (Simd=false):      */
/* rows split into THREADS contiguous ranges (:42-48); per pass a per-thread */
/* histogram (:57-88), bucket totals (:93-101), serial exclusive scan        */
/* (:106-110), per-thread bucket offsets in thread order (:115-122) and a    */
/* per-thread stable scatter (:147-157).  THREADS <= MAX_THREADS.            */
/* ------------------------------------------------------------------------ */
#define DEFINE_RADIX_MT(NAME, KT, PASSES)                                          \
void NAME(KT* tkeys, uint64_t* idx, size_t n, int threads)                         \
{                                                                                  \
    if (n == 0) return;                                                            \
    int T = threads < 1 ? 1 : (threads > MAX_THREADS ? MAX_THREADS : threads);      \
    KT* tk = (KT*)malloc(n * sizeof(KT));                                          \
    uint64_t* ti = (uint64_t*)malloc(n * sizeof(uint64_t));                        \
    uint64_t* hist = (uint64_t*)malloc((size_t)T * RADIX16 * sizeof(uint64_t));    \
    uint64_t* base = (uint64_t*)malloc(RADIX16 * sizeof(uint64_t));                \
    KT* src = tkeys;                                                               \
    KT* dst = tk;                                                                  \
    for (unsigned pass = 0; pass < (PASSES); ++pass) {                             \
        const unsigned sh = 16u * pass;                                            \
        _Pragma("omp parallel num_threads(T)")                                     \
        {                                                                          \
            int t = omp_get_thread_num();                                          \
            size_t chunk = n / (size_t)T, rem = n % (size_t)T;                     \
            size_t beg = (size_t)t * chunk + ((size_t)t < rem ? (size_t)t : rem);  \
            size_t end = beg + chunk + ((size_t)t < rem ? 1 : 0);                  \
            uint64_t* h = hist + (size_t)t * RADIX16;                              \
            memset(h, 0, RADIX16 * sizeof(uint64_t));                              \
            for (size_t i = beg; i < end; ++i) h[(src[i] >> sh) & 0xFFFFu]++;      \
            _Pragma("omp barrier")                                                 \
            _Pragma("omp single")                                                  \
            {                                                                      \
                uint64_t run = 0;                                                  \
                for (size_t b = 0; b < RADIX16; ++b) {                             \
                    base[b] = run;                                                 \
                    for (int u = 0; u < T; ++u) run += hist[(size_t)u * RADIX16 + b]; \
                }                                                                  \
            }                                                                      \
            /* thread offsets: base[b] + sum of earlier threads' counts */         \
            _Pragma("omp for schedule(static)")                                    \
            for (size_t b = 0; b < RADIX16; ++b) {                                 \
                uint64_t o = base[b];                                              \
                for (int u = 0; u < T; ++u) {                                      \
                    uint64_t c = hist[(size_t)u * RADIX16 + b];                    \
                    hist[(size_t)u * RADIX16 + b] = o;                             \
                    o += c;                                                        \
                }                                                                  \
            }                                                                      \
            for (size_t i = beg; i < end; ++i) {                                   \
                KT k = src[i];                                                     \
                uint64_t p = h[(k >> sh) & 0xFFFFu]++;                             \
                ti[p] = idx[i];                                                    \
                dst[p] = k;                                                        \
            }                                                                      \
            _Pragma("omp barrier")                                                 \
            _Pragma("omp for schedule(static)")                                    \
            for (size_t i = 0; i < n; ++i) idx[i] = ti[i];                         \
        }                                                                          \
        KT* t = src; src = dst; dst = t;                                           \
    }                                                                              \
    free(tk); free(ti); free(hist); free(base);                                    \
}

DEFINE_RADIX_MT(orc_radix_sort_u32_mt, uint32_t, 2)
DEFINE_RADIX_MT(orc_radix_sort_u64_mt, uint64_t, 4)

/* ------------------------------------------------------------------------ */
/* Byte-wise LSD over fixed-width transformed keys, radix_sort_charbuf_flat  */
/* .inl:4-107 (Simd=false): 256 buckets, byte position from the last byte to */
/* the first (:94), each pass reads the key byte through the current perm    */
/* (:59-61, :86-89).                                                         */
/* ------------------------------------------------------------------------ */
void orc_radix_sort_charbuf(const uint8_t* bufv, size_t n, size_t width, uint64_t* idx)
{
    if (n == 0) return;
    uint64_t* tmp = (uint64_t*)malloc(n * sizeof(uint64_t));
    uint64_t* in = idx;
    uint64_t* out = tmp;
    uint64_t cnt[256];
    for (long pos = (long)width - 1; pos >= 0; --pos) {
        memset(cnt, 0, sizeof(cnt));
        for (size_t i = 0; i < n; ++i) cnt[bufv[in[i] * width + (size_t)pos]]++;
        uint64_t run = 0;
        for (int b = 0; b < 256; ++b) { uint64_t c = cnt[b]; cnt[b] = run; run += c; }
        for (size_t i = 0; i < n; ++i) {
            uint8_t k = bufv[in[i] * width + (size_t)pos];
            out[cnt[k]++] = in[i];
        }
        uint64_t* t = in; in = out; out = t;
    }
    if (in != idx) memcpy(idx, in, n * sizeof(uint64_t));
    free(tmp);
}

/* ------------------------------------------------------------------------ */
/* sort_by type dispatch.                                                    */
/*  u32/u64: keys copied into tkeys (sort_uintegers.inl:46-84)               */
/*  i32/i64: tkeys = UIntT(x) ^ signbit (sort_integers.inl:46,53,109,123)    */
/*  char[16]: bytes ^ 0x80 (sort_char.inl:68-74,104-110), then charbuf LSD   */
/*  descending, reference code: std::reverse of the ascending perm           */
/*  (sort_uintegers.inl:216-218, sort_integers.inl:225-227) = mode 2;        */
/*  stable descending (INTERNALS.doc:1610-1630) = mode 1: keys complemented. */
/* ------------------------------------------------------------------------ */
static void reverse_u64(uint64_t* a, size_t n)
{
    for (size_t i = 0, j = n ? n - 1 : 0; i < j; ++i, --j) { uint64_t t = a[i]; a[i] = a[j]; a[j] = t; }
}

/* Float keys (sort_flt.inl:44-50): mask = -(bits >> 31);
 * tkeys = bits ^ (mask | signbit), then the unsigned radix. */
static uint32_t flt_key32(uint32_t b) { return b ^ ((0u - (b >> 31)) | 0x80000000u); }
static uint64_t flt_key64(uint64_t b) { return b ^ ((0ull - (b >> 63)) | 0x8000000000000000ull); }

int orc_argsort(int key_type, const void* keys, size_t n, int mode, int threads,
                uint64_t* perm_out)
{
    if (mode < 0 || mode > 2) return -1;
    for (size_t i = 0; i < n; ++i) perm_out[i] = i; /* sort_by_mt.hpp:21-24 iota */
    if (n == 0) return 0;
    const int flip = (mode == 1);
    if (key_type == 5 || key_type == 6) {
        if (key_type == 5) {
            const uint32_t* k = (const uint32_t*)keys;
            uint32_t* tk = (uint32_t*)malloc(n * sizeof(uint32_t));
            for (size_t i = 0; i < n; ++i) tk[i] = flt_key32(k[i]) ^ (flip ? 0xFFFFFFFFu : 0u);
            if (threads > 1) orc_radix_sort_u32_mt(tk, perm_out, n, threads);
            else orc_radix_sort_u32(tk, perm_out, n);
            free(tk);
        } else {
            const uint64_t* k = (const uint64_t*)keys;
            uint64_t* tk = (uint64_t*)malloc(n * sizeof(uint64_t));
            for (size_t i = 0; i < n; ++i) tk[i] = flt_key64(k[i]) ^ (flip ? ~0ull : 0ull);
            if (threads > 1) orc_radix_sort_u64_mt(tk, perm_out, n, threads);
            else orc_radix_sort_u64(tk, perm_out, n);
            free(tk);
        }
    } else if (key_type == 0 || key_type == 1) {
        const uint32_t* k = (const uint32_t*)keys;
        uint32_t x = (key_type == 1 ? 0x80000000u : 0u) ^ (flip ? 0xFFFFFFFFu : 0u);
        uint32_t* tk = (uint32_t*)malloc(n * sizeof(uint32_t));
        for (size_t i = 0; i < n; ++i) tk[i] = k[i] ^ x;
        if (threads > 1) orc_radix_sort_u32_mt(tk, perm_out, n, threads);
        else orc_radix_sort_u32(tk, perm_out, n);
        free(tk);
    } else if (key_type == 2 || key_type == 3) {
        const uint64_t* k = (const uint64_t*)keys;
        uint64_t x = (key_type == 3 ? 0x8000000000000000ull : 0ull) ^ (flip ? ~0ull : 0ull);
        uint64_t* tk = (uint64_t*)malloc(n * sizeof(uint64_t));
        for (size_t i = 0; i < n; ++i) tk[i] = k[i] ^ x;
        if (threads > 1) orc_radix_sort_u64_mt(tk, perm_out, n, threads);
        else orc_radix_sort_u64(tk, perm_out, n);
        free(tk);
    } else if (key_type == 4) {
        const uint8_t* k = (const uint8_t*)keys;
        uint8_t x = (uint8_t)(0x80u ^ (flip ? 0xFFu : 0u));
        uint8_t* flat = (uint8_t*)malloc(n * 16);
        for (size_t i = 0; i < n * 16; ++i) flat[i] = k[i] ^ x;
        orc_radix_sort_charbuf(flat, n, 16, perm_out);
        free(flat);
    } else {
        return -1;
    }
    if (mode == 2) reverse_u64(perm_out, n);
    return 0;
}

/* ------------------------------------------------------------------------ */
/* Permutation gather, permute_block_mt.inl:39-41: dst[r] = src[idx[r]].     */
/* (The reference parallelises over columns; rows are split here.)          */
/* ------------------------------------------------------------------------ */
void orc_gather(size_t elem_bytes, const void* src, const uint64_t* idx, size_t n,
                void* dst, int threads)
{
    int T = threads < 1 ? 1 : threads;
    if (elem_bytes == 8) {
        const uint64_t* s = (const uint64_t*)src; uint64_t* d = (uint64_t*)dst;
        #pragma omp parallel for num_threads(T) schedule(static) if(T > 1)
        for (size_t r = 0; r < n; ++r) d[r] = s[idx[r]];
    } else if (elem_bytes == 4) {
        const uint32_t* s = (const uint32_t*)src; uint32_t* d = (uint32_t*)dst;
        #pragma omp parallel for num_threads(T) schedule(static) if(T > 1)
        for (size_t r = 0; r < n; ++r) d[r] = s[idx[r]];
    } else {
        const uint8_t* s = (const uint8_t*)src; uint8_t* d = (uint8_t*)dst;
        #pragma omp parallel for num_threads(T) schedule(static) if(T > 1)
        for (size_t r = 0; r < n; ++r) memcpy(d + r * elem_bytes, s + idx[r] * elem_bytes, elem_bytes);
    }
}

/* ------------------------------------------------------------------------ */
/* Exact, correctly rounded double sum: Shewchuk (1997) non-overlapping      */
/* expansion grown by two-sum, then rounded half-even from the top partial.  */
/* Finite inputs only.                                                       */
/* ------------------------------------------------------------------------ */
typedef struct { double* p; size_t n, cap; } expansion;

static void exp_add(expansion* e, double x)
{
    size_t i = 0;
    for (size_t j = 0; j < e->n; ++j) {
        double y = e->p[j];
        if (fabs(x) < fabs(y)) { double t = x; x = y; y = t; }
        double hi = x + y;
        double lo = y - (hi - x);
        if (lo != 0.0) e->p[i++] = lo;
        x = hi;
    }
    if (i >= e->cap) {
        e->cap = e->cap ? e->cap * 2 : 32;
        e->p = (double*)realloc(e->p, e->cap * sizeof(double));
    }
    e->p[i++] = x;
    e->n = i;
}

static double exp_round(const expansion* e)
{
    size_t n = e->n;
    if (n == 0) return 0.0;
    double hi = e->p[--n], lo = 0.0;
    while (n > 0) {
        double x = hi, y = e->p[--n];
        hi = x + y;
        double yr = hi - x;
        lo = y - yr;
        if (lo != 0.0) break;
    }
    /* round half to even across the remaining partials */
    if (n > 0 && ((lo < 0.0 && e->p[n - 1] < 0.0) || (lo > 0.0 && e->p[n - 1] > 0.0))) {
        double y = lo * 2.0;
        double x = hi + y;
        double yr = x - hi;
        if (y == yr) hi = x;
    }
    return hi;
}

double orc_exact_sum(const double* x, size_t n)
{
    expansion e = {0, 0, 0};
    for (size_t i = 0; i < n; ++i) exp_add(&e, x[i]);
    double r = exp_round(&e);
    free(e.p);
    return r;
}

/* ------------------------------------------------------------------------ */
/* Sort-based group-by: the build's contract for the aggregates whose       */
/* semantics come from onecol_functions.inl:4-130 (Occurence / Sum / Gather)  */
/* and create_value_col.inl:36-61.  Decisions (stated, not the reference's  */
/* quirks): groups in ascending key order; mean = sum / count (the          */
/* reference divides by nrow, create_value_col.inl:46); min/max as           */
/* std::min_element / max_element over the row-ordered Gather vector         */
/* (first of equal values wins); integer sums wrap mod 2^64; integer means   */
/* use the exact 128-bit sum; f64 sums are correctly rounded.                */
/* ------------------------------------------------------------------------ */
static size_t key_bytes(int kt) { return (kt == 0 || kt == 1 || kt == 5) ? 4 : (kt == 4 ? 16 : 8); }
static size_t val_bytes(int vt) { return (vt == 2 || vt == 3) ? 4 : 8; }

static int64_t ival(int vt, const void* col, uint64_t r, int* is_unsigned)
{
    switch (vt) {
    case 1: *is_unsigned = 0; return ((const int64_t*)col)[r];
    case 2: *is_unsigned = 0; return (int64_t)((const int32_t*)col)[r];
    case 3: *is_unsigned = 1; return (int64_t)(uint64_t)((const uint32_t*)col)[r];
    default: *is_unsigned = 1; return (int64_t)((const uint64_t*)col)[r];
    }
}

uint64_t orc_groupby(int key_type, const void* keys, size_t n,
                     int nvals, const int* val_types, const void* const* vals,
                     void* group_keys, uint64_t* offsets, uint64_t* counts,
                     uint64_t* perm,
                     void* const* sums, double* const* means,
                     void* const* mins, void* const* maxs, int threads)
{
    const size_t kb = key_bytes(key_type);
    uint64_t* p = perm ? perm : (uint64_t*)malloc((n ? n : 1) * sizeof(uint64_t));
    orc_argsort(key_type, keys, n, 0, threads, p);
    const uint8_t* kbytes = (const uint8_t*)keys;
    expansion e = {0, 0, 0};
    uint64_t g = 0;
    size_t s = 0;
    while (s < n) {
        size_t t = s + 1;
        while (t < n && memcmp(kbytes + p[t] * kb, kbytes + p[s] * kb, kb) == 0) ++t;
        if (group_keys) memcpy((uint8_t*)group_keys + g * kb, kbytes + p[s] * kb, kb);
        if (offsets) offsets[g] = s;
        if (counts) counts[g] = t - s;
        for (int c = 0; c < nvals; ++c) {
            const int vt = val_types[c];
            if (vt == 0) {
                const double* v = (const double*)vals[c];
                e.n = 0;
                double mn = v[p[s]], mx = v[p[s]];
                for (size_t i = s; i < t; ++i) {
                    double x = v[p[i]];
                    exp_add(&e, x);
                    if (x < mn) mn = x;
                    if (x > mx) mx = x;
                }
                double sum = exp_round(&e);
                if (sums && sums[c]) ((double*)sums[c])[g] = sum;
                if (means && means[c]) means[c][g] = sum / (double)(t - s);
                if (mins && mins[c]) ((double*)mins[c])[g] = mn;
                if (maxs && maxs[c]) ((double*)maxs[c])[g] = mx;
            } else {
                int uns = 0;
                __int128 acc = 0;
                int64_t first = ival(vt, vals[c], p[s], &uns);
                int64_t mn = first, mx = first;
                for (size_t i = s; i < t; ++i) {
                    int64_t x = ival(vt, vals[c], p[i], &uns);
                    acc += uns ? (__int128)(uint64_t)x : (__int128)x;
                    if (uns ? ((uint64_t)x < (uint64_t)mn) : (x < mn)) mn = x;
                    if (uns ? ((uint64_t)x > (uint64_t)mx) : (x > mx)) mx = x;
                }
                if (sums && sums[c]) ((uint64_t*)sums[c])[g] = (uint64_t)acc;
                if (means && means[c]) means[c][g] = (double)acc / (double)(t - s);
                const size_t vb = val_bytes(vt);
                if (mins && mins[c]) {
                    if (vb == 4) ((uint32_t*)mins[c])[g] = (uint32_t)mn;
                    else ((int64_t*)mins[c])[g] = mn;
                }
                if (maxs && maxs[c]) {
                    if (vb == 4) ((uint32_t*)maxs[c])[g] = (uint32_t)mx;
                    else ((int64_t*)maxs[c])[g] = mx;
                }
            }
        }
        ++g;
        s = t;
    }
    if (offsets) offsets[g] = n;
    free(e.p);
    if (!perm) free(p);
    return g;
}

/* ------------------------------------------------------------------------ */
/* Hash group-by restated from the reference (the CPU baseline).            */
/*   key hash: hash_scalar for len <= 16, simd_hash.hpp:31-38 --             */
/*     mix64(v ^ seed ^ 0x9e3779b97f4a7c15) with mix64 from :8-15.  The      */
/*     reference memcpy's all len bytes into one 8-byte v (rd64_tail, :18-22),*/
/*     which overruns v for 9..16-byte keys; here bytes 8..15 are folded in  */
/*     with an extra mix64 instead.                                          */
/*   map: first-occurrence (insertion) order as ankerl::unordered_dense      */
/*     iterates; lookups try_emplace then accumulate in row order            */
/*     (onecol_functions.inl:34-40, :77-83, :120-126).                       */
/*   threads > 1: per-thread maps on contiguous row ranges, merged in thread */
/*     order by try_emplace + += (merge_functions.inl:3-36).                 */
/* ------------------------------------------------------------------------ */
static uint64_t mix64(uint64_t x)
{
    x ^= x >> 33;
    x *= 0xff51afd7ed558ccdull;
    x ^= x >> 33;
    x *= 0xc4ceb9fe1a85ec53ull;
    x ^= x >> 33;
    return x;
}

uint64_t orc_hash_bytes16(const void* p, size_t len)
{
    uint64_t v = 0, w = 0;
    memcpy(&v, p, len < 8 ? len : 8);
    if (len > 8) { memcpy(&w, (const uint8_t*)p + 8, len - 8); v ^= mix64(w); }
    return mix64(v ^ 0x9e3779b97f4a7c15ull);
}

typedef struct {
    size_t kb;
    uint64_t mask;
    int64_t* slot;   /* group index or -1 */
    uint8_t* gkeys;  /* group keys in insertion order */
    uint64_t ng, cap;
    uint64_t* cnt;
    double* sum;
    double* mn;
    double* mx;
    /* Gather vectors (refapi min/max): per group dynamic arrays */
    double** gv;
    uint64_t* gvn;
    uint64_t* gvc;
} hmap;

static void hmap_init(hmap* m, size_t kb, uint64_t expect, int gather)
{
    uint64_t c = 16;
    while (c < expect * 2) c <<= 1;
    m->kb = kb; m->mask = c - 1;
    m->slot = (int64_t*)malloc(c * sizeof(int64_t));
    memset(m->slot, 0xff, c * sizeof(int64_t));
    m->cap = expect < 16 ? 16 : expect;
    m->ng = 0;
    m->gkeys = (uint8_t*)malloc(m->cap * kb);
    m->cnt = (uint64_t*)calloc(m->cap, sizeof(uint64_t));
    m->sum = (double*)calloc(m->cap, sizeof(double));
    m->mn = (double*)malloc(m->cap * sizeof(double));
    m->mx = (double*)malloc(m->cap * sizeof(double));
    m->gv = gather ? (double**)calloc(m->cap, sizeof(double*)) : NULL;
    m->gvn = gather ? (uint64_t*)calloc(m->cap, sizeof(uint64_t)) : NULL;
    m->gvc = gather ? (uint64_t*)calloc(m->cap, sizeof(uint64_t)) : NULL;
}

static void hmap_free(hmap* m)
{
    if (m->gv) { for (uint64_t g = 0; g < m->ng; ++g) free(m->gv[g]); }
    free(m->slot); free(m->gkeys); free(m->cnt); free(m->sum); free(m->mn); free(m->mx);
    free(m->gv); free(m->gvn); free(m->gvc);
}

static void hmap_grow_groups(hmap* m)
{
    uint64_t nc = m->cap * 2;
    m->gkeys = (uint8_t*)realloc(m->gkeys, nc * m->kb);
    m->cnt = (uint64_t*)realloc(m->cnt, nc * sizeof(uint64_t));
    m->sum = (double*)realloc(m->sum, nc * sizeof(double));
    m->mn = (double*)realloc(m->mn, nc * sizeof(double));
    m->mx = (double*)realloc(m->mx, nc * sizeof(double));
    if (m->gv) {
        m->gv = (double**)realloc(m->gv, nc * sizeof(double*));
        m->gvn = (uint64_t*)realloc(m->gvn, nc * sizeof(uint64_t));
        m->gvc = (uint64_t*)realloc(m->gvc, nc * sizeof(uint64_t));
        for (uint64_t g = m->cap; g < nc; ++g) { m->gv[g] = NULL; m->gvn[g] = 0; m->gvc[g] = 0; }
    }
    for (uint64_t g = m->cap; g < nc; ++g) { m->cnt[g] = 0; m->sum[g] = 0.0; }
    m->cap = nc;
}

static void hmap_rehash(hmap* m)
{
    uint64_t c = (m->mask + 1) * 2;
    free(m->slot);
    m->slot = (int64_t*)malloc(c * sizeof(int64_t));
    memset(m->slot, 0xff, c * sizeof(int64_t));
    m->mask = c - 1;
    for (uint64_t g = 0; g < m->ng; ++g) {
        uint64_t h = orc_hash_bytes16(m->gkeys + g * m->kb, m->kb) & m->mask;
        while (m->slot[h] >= 0) h = (h + 1) & m->mask;
        m->slot[h] = (int64_t)g;
    }
}

/* try_emplace: returns the group index, inserting in first-occurrence order */
static uint64_t hmap_find(hmap* m, const uint8_t* key, double v0)
{
    uint64_t h = orc_hash_bytes16(key, m->kb) & m->mask;
    for (;;) {
        int64_t s = m->slot[h];
        if (s < 0) break;
        if (memcmp(m->gkeys + (uint64_t)s * m->kb, key, m->kb) == 0) return (uint64_t)s;
        h = (h + 1) & m->mask;
    }
    if (m->ng == m->cap) hmap_grow_groups(m);
    uint64_t g = m->ng++;
    memcpy(m->gkeys + g * m->kb, key, m->kb);
    m->cnt[g] = 0; m->sum[g] = 0.0; m->mn[g] = v0; m->mx[g] = v0;
    m->slot[h] = (int64_t)g;
    if (m->ng * 2 > m->mask + 1) hmap_rehash(m);
    return g;
}

static void gather_push(hmap* m, uint64_t g, double v)
{
    if (m->gvn[g] == m->gvc[g]) {
        m->gvc[g] = m->gvc[g] ? m->gvc[g] * 2 : 4;
        m->gv[g] = (double*)realloc(m->gv[g], m->gvc[g] * sizeof(double));
    }
    m->gv[g][m->gvn[g]++] = v;
}

/* aggregate kinds for one pass over rows */
enum { AGG_OCC = 1, AGG_SUM = 2, AGG_MINMAX = 4, AGG_GATHER = 8 };

static void hash_pass(hmap* m, const uint8_t* keys, const double* val, size_t beg, size_t end, int what)
{
    const size_t kb = m->kb;
    for (size_t i = beg; i < end; ++i) {
        double v = val ? val[i] : 0.0;
        uint64_t g = hmap_find(m, keys + i * kb, v);
        if (what & AGG_OCC) m->cnt[g]++;
        if (what & AGG_SUM) m->sum[g] += v;
        if (what & AGG_MINMAX) {
            if (v < m->mn[g]) m->mn[g] = v;
            if (v > m->mx[g]) m->mx[g] = v;
        }
        if (what & AGG_GATHER) gather_push(m, g, v);
    }
}

/* merge cur into lookup in cur's insertion order (merge_functions.inl:15-19) */
static void hash_merge(hmap* lookup, hmap* cur, int what)
{
    for (uint64_t g = 0; g < cur->ng; ++g) {
        uint64_t d = hmap_find(lookup, cur->gkeys + g * cur->kb, cur->mn[g]);
        if (what & AGG_OCC) lookup->cnt[d] += cur->cnt[g];
        if (what & AGG_SUM) lookup->sum[d] += cur->sum[g];
        if (what & AGG_MINMAX) {
            if (cur->mn[g] < lookup->mn[d]) lookup->mn[d] = cur->mn[g];
            if (cur->mx[g] > lookup->mx[d]) lookup->mx[d] = cur->mx[g];
        }
        if (what & AGG_GATHER)
            for (uint64_t j = 0; j < cur->gvn[g]; ++j) gather_push(lookup, d, cur->gv[g][j]);
    }
}

/* One full (possibly multithreaded) pass; result in *out (caller frees). */
static void hash_run(hmap* out, size_t kb, const uint8_t* keys, const double* val, size_t n,
                     int T, int what)
{
    if (T <= 1) {
        hmap_init(out, kb, 1024, (what & AGG_GATHER) != 0);
        hash_pass(out, keys, val, 0, n, what);
        return;
    }
    hmap* parts = (hmap*)malloc((size_t)T * sizeof(hmap));
    #pragma omp parallel num_threads(T)
    {
        int t = omp_get_thread_num();
        size_t chunk = n / (size_t)T, rem = n % (size_t)T;
        size_t beg = (size_t)t * chunk + ((size_t)t < rem ? (size_t)t : rem);
        size_t end = beg + chunk + ((size_t)t < rem ? 1 : 0);
        hmap_init(&parts[t], kb, 1024, (what & AGG_GATHER) != 0);
        hash_pass(&parts[t], keys, val, beg, end, what);
    }
    *out = parts[0];
    for (int t = 1; t < T; ++t) { hash_merge(out, &parts[t], what); hmap_free(&parts[t]); }
    free(parts);
}

uint64_t orc_hash_groupby_f64(int key_type, const void* keys, size_t n,
                              const double* val, int threads, int refapi,
                              void* group_keys, uint64_t* counts, double* sums,
                              double* means, double* mins, double* maxs)
{
    const size_t kb = key_bytes(key_type);
    const uint8_t* kp = (const uint8_t*)keys;
    int T = threads < 1 ? 1 : threads;
    hmap m;
    uint64_t ng;
    if (!refapi) {
        hash_run(&m, kb, kp, val, n, T, AGG_OCC | AGG_SUM | AGG_MINMAX);
        ng = m.ng;
        for (uint64_t g = 0; g < ng; ++g) {
            if (group_keys) memcpy((uint8_t*)group_keys + g * kb, m.gkeys + g * kb, kb);
            if (counts) counts[g] = m.cnt[g];
            if (sums) sums[g] = m.sum[g];
            if (means) means[g] = m.sum[g] / (double)m.cnt[g];
            if (mins) mins[g] = m.mn[g];
            if (maxs) maxs[g] = m.mx[g];
        }
        hmap_free(&m);
        return ng;
    }
    /* reference API: Occurence pass, Sum pass, Mean pass, Gather pass + f for
       min, Gather pass + f for max (all in first-occurrence order). */
    hash_run(&m, kb, kp, val, n, T, AGG_OCC);
    ng = m.ng;
    for (uint64_t g = 0; g < ng; ++g) {
        if (group_keys) memcpy((uint8_t*)group_keys + g * kb, m.gkeys + g * kb, kb);
        if (counts) counts[g] = m.cnt[g];
    }
    uint64_t* cnt = (uint64_t*)malloc((ng ? ng : 1) * sizeof(uint64_t));
    memcpy(cnt, m.cnt, ng * sizeof(uint64_t));
    hmap_free(&m);
    hash_run(&m, kb, kp, val, n, T, AGG_SUM);
    for (uint64_t g = 0; g < ng; ++g) if (sums) sums[g] = m.sum[g];
    hmap_free(&m);
    hash_run(&m, kb, kp, val, n, T, AGG_SUM);
    for (uint64_t g = 0; g < ng; ++g) if (means) means[g] = m.sum[g] / (double)cnt[g];
    hmap_free(&m);
    for (int which = 0; which < 2; ++which) {
        hash_run(&m, kb, kp, val, n, T, AGG_GATHER);
        for (uint64_t g = 0; g < ng; ++g) {
            const double* v = m.gv[g];
            double best = v[0];
            for (uint64_t j = 1; j < m.gvn[g]; ++j) {
                if (which == 0 ? (v[j] < best) : (v[j] > best)) best = v[j];
            }
            if (which == 0 && mins) mins[g] = best;
            if (which == 1 && maxs) maxs[g] = best;
        }
        hmap_free(&m);
    }
    free(cnt);
    return ng;
}

/* ---- CSV ingest (read/read_csv/read_csv.hpp:9-216) ---- */

/* calls f(ctx, column, begin, length) for each field of text[b, e) */
typedef void (*csv_field_fn)(void* ctx, int c, const char* s, size_t len);

static int csv_fields(const char* text, uint64_t b, uint64_t e, char delim, char quote, csv_field_fn f, void* ctx)
{
    if (e > b && text[e - 1] == '\r') --e;
    int c = 0, inq = 0;
    uint64_t fs = b;
    for (uint64_t i = b; i < e; ++i) {
        const char ch = text[i];
        if (ch == quote) inq = !inq;
        else if (ch == delim && !inq) {
            f(ctx, c++, text + fs, i - fs);
            fs = i + 1;
        }
    }
    f(ctx, c, text + fs, e - fs);
    return c + 1;
}

static int csv_digit(char c) { return c >= '0' && c <= '9'; }
static char csv_lower(char c) { return (c >= 'A' && c <= 'Z') ? (char)(c - 'A' + 'a') : c; }

/* [-]digits[.digits][(e|E)[+-]digits], digits on at least one side of the
 * dot, the whole field (classify_column.inl's test, with '-' and exponents) */
static int csv_is_number(const char* s, size_t n, int* dot_or_exp, int* neg)
{
    size_t i = 0, digits = 0;
    *dot_or_exp = 0;
    *neg = 0;
    if (i < n && s[i] == '-') {
        *neg = 1;
        ++i;
    }
    while (i < n && csv_digit(s[i])) ++i, ++digits;
    if (i < n && s[i] == '.') {
        *dot_or_exp = 1;
        ++i;
        while (i < n && csv_digit(s[i])) ++i, ++digits;
    }
    if (digits == 0) return 0;
    if (i < n && (s[i] == 'e' || s[i] == 'E')) {
        size_t ed = 0;
        *dot_or_exp = 1;
        ++i;
        if (i < n && (s[i] == '+' || s[i] == '-')) ++i;
        while (i < n && csv_digit(s[i])) ++i, ++ed;
        if (ed == 0) return 0;
    }
    return i == n;
}

typedef struct {
    int want;
    const char* s;
    size_t len;
    int found;
} csv_pick;

static void csv_pick_fn(void* ctx, int c, const char* s, size_t len)
{
    csv_pick* p = (csv_pick*)ctx;
    if (c == p->want) {
        p->s = s;
        p->len = len;
        p->found = 1;
    }
}

static void csv_count_fn(void* ctx, int c, const char* s, size_t len)
{
    (void)ctx;
    (void)c;
    (void)s;
    (void)len;
}

int orc_csv_infer(const char* text, size_t bytes, char delim, char quote, int header, int char_width, int* kinds,
                  int max_cols)
{
    /* the first 12 lines (header + 10 data rows + one) */
    uint64_t ls[13];
    int nl = 0;
    uint64_t fs = 0;
    for (uint64_t i = 0; i < bytes && nl < 12; ++i)
        if (text[i] == '\n') {
            ls[nl++] = fs;
            fs = i + 1;
        }
    uint64_t le[13];
    for (int k = 0; k < nl; ++k) le[k] = (k + 1 < nl) ? ls[k + 1] - 1 : fs - 1;
    if (nl < 12 && fs < bytes) {
        ls[nl] = fs;
        le[nl] = bytes;
        ++nl;
    }
    if (nl == 0) return 0;
    const int ncol = csv_fields(text, ls[0], le[0], delim, quote, csv_count_fn, NULL);
    if (ncol > max_cols) return -1;
    const int first = header ? 1 : 0;
    for (int c = 0; c < ncol; ++c) {
        int numeric = 1, any_dot = 0, any_neg = 0, boolean = 1, fits = 1, seen = 0;
        for (int l = first; l < nl && seen < 10; ++l, ++seen) {
            csv_pick p = {c, NULL, 0, 0};
            csv_fields(text, ls[l], le[l], delim, quote, csv_pick_fn, &p);
            const char* v = p.found ? p.s : "";
            const size_t n = p.found ? p.len : 0;
            int dot = 0, neg = 0;
            if (!csv_is_number(v, n, &dot, &neg)) numeric = 0;
            any_dot |= dot;
            any_neg |= neg;
            if (!(n == 1 && (v[0] == '0' || v[0] == '1'))) boolean = 0;
            if (n > (size_t)char_width) fits = 0;
        }
        if (!numeric) kinds[c] = (fits && char_width > 0) ? 1 : 0;
        else if (any_dot) kinds[c] = 5;
        else if (any_neg) kinds[c] = 3;
        else if (boolean && seen > 0) kinds[c] = 2;
        else kinds[c] = 4;
    }
    return ncol;
}

uint64_t orc_csv_lines(const char* text, size_t bytes, uint64_t* line, int threads)
{
    if (threads < 1) threads = 1;
    const size_t chunk = 1 << 20;
    const size_t nb = (bytes + chunk - 1) / chunk;
    uint64_t* cnt = (uint64_t*)calloc(nb + 1, sizeof(uint64_t));
    #pragma omp parallel for num_threads(threads) schedule(static)
    for (size_t b = 0; b < nb; ++b) {
        uint64_t c = 0;
        const size_t e = (b + 1) * chunk < bytes ? (b + 1) * chunk : bytes;
        for (size_t i = b * chunk; i < e; ++i) c += text[i] == '\n';
        cnt[b + 1] = c;
    }
    for (size_t b = 0; b < nb; ++b) cnt[b + 1] += cnt[b];
    line[0] = 0;
    #pragma omp parallel for num_threads(threads) schedule(static)
    for (size_t b = 0; b < nb; ++b) {
        uint64_t k = cnt[b];
        const size_t e = (b + 1) * chunk < bytes ? (b + 1) * chunk : bytes;
        for (size_t i = b * chunk; i < e; ++i)
            if (text[i] == '\n') line[1 + k++] = i + 1;
    }
    uint64_t nl = cnt[nb];
    free(cnt);
    if ((nl ? line[nl] : 0) < bytes) line[++nl] = bytes + 1;  /* an unterminated last line */
    return nl;
}

typedef struct {
    int ncol;
    const int* kinds;
    uint32_t* maxlen;
} csv_measure_ctx;

static void csv_measure_fn(void* ctx, int c, const char* s, size_t len)
{
    (void)s;
    csv_measure_ctx* m = (csv_measure_ctx*)ctx;
    if (c < m->ncol && m->kinds[c] == 0 && len > m->maxlen[c]) m->maxlen[c] = (uint32_t)len;
}

int64_t orc_csv_measure(const char* text, size_t bytes, const uint64_t* line, uint64_t nrow, char delim,
                        char quote, int ncol, const int* kinds, uint32_t* maxlen, int threads)
{
    if (threads < 1) threads = 1;
    int64_t bad = -1;
    for (int c = 0; c < ncol; ++c) maxlen[c] = 0;
    #pragma omp parallel num_threads(threads)
    {
        uint32_t* mine = (uint32_t*)calloc((size_t)(ncol > 0 ? ncol : 1), sizeof(uint32_t));
        csv_measure_ctx m = {ncol, kinds, mine};
        int64_t bad_t = -1;
        #pragma omp for schedule(static)
        for (uint64_t r = 0; r < nrow; ++r) {
            uint64_t e = line[r + 1] - 1;
            if (e > bytes) e = bytes;
            const int nf = csv_fields(text, line[r], e, delim, quote, csv_measure_fn, &m);
            if (nf != ncol && (bad_t < 0 || (int64_t)r < bad_t)) bad_t = (int64_t)r;
        }
        #pragma omp critical
        {
            for (int c = 0; c < ncol; ++c)
                if (mine[c] > maxlen[c]) maxlen[c] = mine[c];
            if (bad_t >= 0 && (bad < 0 || bad_t < bad)) bad = bad_t;
        }
        free(mine);
    }
    return bad;
}

/* std::from_chars, base 10 */
static int64_t csv_int64(const char* p, size_t n)
{
    size_t i = 0;
    int neg = 0;
    if (i < n && p[i] == '-') {
        neg = 1;
        ++i;
    }
    if (i >= n || !csv_digit(p[i])) return 0;
    const uint64_t lim = neg ? 9223372036854775808ull : 9223372036854775807ull;
    uint64_t v = 0;
    for (; i < n && csv_digit(p[i]); ++i) {
        const uint64_t d = (uint64_t)(p[i] - '0');
        if (v > (lim - d) / 10) return 0;
        v = v * 10 + d;
    }
    return neg ? (int64_t)(0 - v) : (int64_t)v;
}

static uint64_t csv_uint64(const char* p, size_t n)
{
    if (n == 0 || !csv_digit(p[0])) return 0;
    uint64_t v = 0;
    for (size_t i = 0; i < n && csv_digit(p[i]); ++i) {
        const uint64_t d = (uint64_t)(p[i] - '0');
        if (v > (18446744073709551615ull - d) / 10) return 0;
        v = v * 10 + d;
    }
    return v;
}

/* fast_float::from_chars: the longest prefix [-](digits[.digits] | .digits)
 * [(e|E)[+-]digits], or inf / infinity / nan in any case; strtod of that
 * prefix (correctly rounded by glibc) */
static double csv_double(const char* p, size_t n)
{
    size_t i = 0;
    int neg = 0;
    if (i < n && p[i] == '-') {
        neg = 1;
        ++i;
    }
    if (i < n && (csv_lower(p[i]) == 'i' || csv_lower(p[i]) == 'n')) {
        if (i + 3 <= n && csv_lower(p[i]) == 'n' && csv_lower(p[i + 1]) == 'a' && csv_lower(p[i + 2]) == 'n')
            return neg ? -NAN : NAN;
        if (i + 3 <= n && csv_lower(p[i]) == 'i' && csv_lower(p[i + 1]) == 'n' && csv_lower(p[i + 2]) == 'f')
            return neg ? -INFINITY : INFINITY;
        return 0.0;
    }
    size_t digits = 0;
    int dot = 0;
    for (; i < n; ++i) {
        if (p[i] == '.') {
            if (dot) break;
            dot = 1;
            continue;
        }
        if (!csv_digit(p[i])) break;
        ++digits;
    }
    if (digits == 0) return 0.0;
    size_t end = i;
    if (i < n && (p[i] == 'e' || p[i] == 'E')) {
        size_t j = i + 1;
        if (j < n && (p[j] == '-' || p[j] == '+')) ++j;
        if (j < n && csv_digit(p[j])) {
            while (j < n && csv_digit(p[j])) ++j;
            end = j;
        }
    }
    char small[128];
    char* buf = end < sizeof small ? small : (char*)malloc(end + 1);
    memcpy(buf, p, end);
    buf[end] = '\0';
    const double x = strtod(buf, NULL);
    if (buf != small) free(buf);
    return x;
}

typedef struct {
    uint64_t r;
    int ncol;
    const int* kinds;
    const uint32_t* widths;
    void* const* outs;
    char trailing;
} csv_convert_ctx;

static void csv_convert_fn(void* ctx, int c, const char* s, size_t len)
{
    const csv_convert_ctx* x = (const csv_convert_ctx*)ctx;
    if (c >= x->ncol) return;
    const uint32_t w = x->widths[c];
    void* out = x->outs[c];
    switch (x->kinds[c]) {
    case 5: ((double*)out)[x->r] = csv_double(s, len); break;
    case 3:
        if (w == 8) ((int64_t*)out)[x->r] = csv_int64(s, len);
        else ((uint32_t*)out)[x->r] = (uint32_t)csv_int64(s, len);
        break;
    case 4:
        if (w == 8) ((uint64_t*)out)[x->r] = csv_uint64(s, len);
        else ((uint32_t*)out)[x->r] = (uint32_t)csv_uint64(s, len);
        break;
    case 2: ((uint32_t*)out)[x->r] = csv_int64(s, len) != 0 ? 1u : 0u; break;
    default: {
        char* d = (char*)out + x->r * w;
        const char pad = x->kinds[c] == 1 ? x->trailing : '\0';
        const size_t m = len < w ? len : w;
        memcpy(d, s, m);
        memset(d + m, pad, w - m);
        break;
    }
    }
}

void orc_csv_convert(const char* text, size_t bytes, const uint64_t* line, uint64_t nrow, char delim,
                     char quote, char trailing, int ncol, const int* kinds, const uint32_t* widths,
                     void* const* outs, int threads)
{
    if (threads < 1) threads = 1;
    #pragma omp parallel for num_threads(threads) schedule(static)
    for (uint64_t r = 0; r < nrow; ++r) {
        csv_convert_ctx x = {r, ncol, kinds, widths, outs, trailing};
        uint64_t e = line[r + 1] - 1;
        if (e > bytes) e = bytes;
        csv_fields(text, line[r], e, delim, quote, csv_convert_fn, &x);
    }
}
