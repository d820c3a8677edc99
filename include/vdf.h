/*
 * vdf.h -- C ABI of the MI355X (gfx950) radix-argsort / group-by library
 *          (libvdf_hip.so).
 *
 * Conventions
 *   - Every entry point returns vdf_status (0 = VDF_OK).  On failure,
 *     vdf_last_error() returns a message for the calling host thread.
 *   - Data pointers named d_* are device pointers (hipMalloc'ed or equivalent)
 *     owned by the caller.  Scratch memory is owned by the handle.
 *   - `stream` is a hipStream_t passed as void* (NULL = the null stream).
 *     All work is stream-ordered; an entry point returns once its work is
 *     enqueued.  Results are ready when the stream has reached that point.
 *   - A handle serialises its own calls (internal mutex) and owns one scratch
 *     arena; use one handle per concurrently used stream.
 *   - Rows per single sort: n <= VDF_MAX_ROWS (2^30 - 1); row ids are 32 bit
 *     unless VDF_SORT_PERM64 asks for 64-bit (size_t) ids.
 *
 * Reference interfaces replaced (paths relative to include/vison/ of VisonDF):
 *   vdf_argsort     sort_{uintegers,integers,char}(std::vector<size_t>& idx,
 *                   const T* col, size_t nrow)
 *                     inlines/transform/sort_by/sort_uintegers.inl:8-13
 *                     inlines/transform/sort_by/sort_integers.inl:8-13
 *                     inlines/transform/sort_by/sort_char.inl:9-15
 *                   and sort_flt(idx, col, nrow) for float keys
 *                     inlines/transform/sort_by/sort_flt.inl:8-13 (transform :44-50)
 *                   and the radix kernels they call,
 *                     inlines/radix_sorts/radix_sort_uint32{,_mt}.inl:5-7 / 6-8
 *                     inlines/radix_sorts/radix_sort_uint64{,_mt}.inl:5-7 / 6-8
 *                     inlines/radix_sorts/radix_sort_charbuf_flat.inl:4-8
 *   vdf_gather      permute_block_mt(storage, ..., idx, nrow)
 *                     inlines/transform/sort_by/permute_block_mt.inl:5-10,39-41
 *   vdf_groupby     transform_group_by<..., Function, ...>(key_cols, n_col, ...)
 *                     transform/transform_group_by/transform_group_by.hpp:11-14
 *                     per-row loops inlines/transform/transform_group_by/
 *                     onecol_functions.inl:4-130, create_value_col.inl:36-61,
 *                     group-contiguous rows hard/create_value_col_hard.inl:15-106
 */
#ifndef VDF_H
#define VDF_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct vdf_handle vdf_handle;
typedef int vdf_status;

/* status codes */
#define VDF_OK               0
#define VDF_ERR_INVALID      1 /* bad argument: null pointer, unknown type, size limit */
#define VDF_ERR_UNSUPPORTED  2 /* combination not implemented */
#define VDF_ERR_OOM          3 /* device allocation failed */
#define VDF_ERR_HIP          4 /* a HIP runtime call failed */
#define VDF_ERR_BUDGET       5 /* request exceeds the handle's HBM budget */
#define VDF_ERR_TIMEOUT      6 /* vdf_server_wait: the request is not done yet */

/* key column types (sort_by_mt.hpp:72-115 type letters u / i / c / d) */
#define VDF_KEY_U32     0
#define VDF_KEY_I32     1
#define VDF_KEY_U64     2
#define VDF_KEY_I64     3
#define VDF_KEY_CHAR16  4 /* fixed-width 16-byte char[16] cells (CharT = char[16]) */
#define VDF_KEY_F32     5 /* IEEE float, total order of sort_flt.inl:44-50:              */
#define VDF_KEY_F64     6 /* -NaN < -inf < ... < -0 < +0 < ... < +inf < +NaN; groups by bits */
#define VDF_KEY_BYTES   7 /* multi-column keys only: fixed-width byte strings of           */
                          /* vdf_key_col.width bytes (padded std::string cells), ordered */
                          /* as vdf_argsort_bytes orders them                             */

/* value column types */
#define VDF_VAL_F64 0
#define VDF_VAL_I64 1
#define VDF_VAL_I32 2
#define VDF_VAL_U32 3
#define VDF_VAL_U64 4

/* vdf_argsort flags */
#define VDF_SORT_DESC          0x1u /* stable descending (INTERNALS.doc:1610-1630)          */
#define VDF_SORT_DESC_REVERSE  0x2u /* ascending, then reversed: the reference code's        */
                                    /* descending path (sort_uintegers.inl:216-218)         */
#define VDF_SORT_PERM64        0x4u /* d_perm_out is uint64_t[n] (size_t ids, as            */
                                    /* sort_by_mt.hpp:21-24), else uint32_t[n]              */

/* vdf_groupby aggregate mask (GroupFunction, types/supported_group_methods.inl:1-7;
 * MIN / MAX replace Gather + a user min/max functor, create_value_col.inl:48-58) */
#define VDF_AGG_COUNT 0x1u
#define VDF_AGG_SUM   0x2u
#define VDF_AGG_MEAN  0x4u
#define VDF_AGG_MIN   0x8u
#define VDF_AGG_MAX   0x10u
#define VDF_AGG_ALL   0x1Fu

#define VDF_MAX_ROWS        ((1ull << 30) - 1ull)
#define VDF_MAX_VALUE_COLS  4
#define VDF_MAX_KEY_COLS    4
#define VDF_MAX_STRING_BYTES 4096

/* ---- CSV ingest ----
 * Replaces read_csv (read/read_csv/read_csv.hpp:9-216, with the column type
 * inference of inlines/classify_column.inl:13-130 and the conversions of
 * std::from_chars / fast_float::from_chars).  The text (d_text, bytes, on the
 * device, 16-byte aligned) is split into rows at '\n' (a '\r' before it is
 * dropped; an unterminated last line is a row), rows into fields at `delim`
 * outside pairs of `quote`; with `header` the first line names the columns.
 * Column types come from the first 10 data rows: numbers
 * ([-]digits[.digits][e[+-]digits]) give DOUBLE when one has a '.' or an
 * exponent, INT when one starts with '-', BOOL when all are 0 or 1, else
 * UINT; other columns are CHAR (fixed char_width bytes, padded with
 * `trailing`) when their first 10 fields fit char_width bytes, else STRING
 * (padded with 0x00 to the column's longest field).  Every field converts
 * from its longest valid numeric prefix (0 when there is none or an integer
 * is out of range); doubles are correctly rounded.  A row with another field
 * count fails with VDF_ERR_INVALID.  The columns live in the table's device
 * memory until vdf_csv_free.  Synchronous (the null stream). */
#define VDF_CSV_STRING 0 /* type letters of type_classification.inl: s */
#define VDF_CSV_CHAR   1 /* c */
#define VDF_CSV_BOOL   2 /* b: uint32 cells, 0 or 1 */
#define VDF_CSV_INT    3 /* i: int32 or int64 cells (int_bytes) */
#define VDF_CSV_UINT   4 /* u: uint32 or uint64 cells (int_bytes) */
#define VDF_CSV_DOUBLE 5 /* d: float64 cells */

typedef struct vdf_csv_options {
    char delim;      /* ',' */
    char quote;      /* str_context ('\'' in the reference) */
    char trailing;   /* TrailingChar: pads CHAR cells ('0' in the reference) */
    char reserved0;
    int header;      /* header_name: the first line holds column names */
    int int_bytes;   /* 4 or 8: cells of INT / UINT columns (IntT / UIntT) */
    int char_width;  /* df_charbuf_size: CHAR cell bytes (0: no CHAR columns) */
} vdf_csv_options;

typedef struct vdf_csv_column {
    int kind;          /* VDF_CSV_* */
    uint32_t width;    /* bytes per cell */
    void* data;        /* device, nrow cells (owned by the table) */
    const char* name;  /* header name ("NA" for an empty one; "" without a header) */
} vdf_csv_column;

typedef struct vdf_csv_table vdf_csv_table;
vdf_status vdf_read_csv(vdf_handle* h, const void* d_text, uint64_t bytes, const vdf_csv_options* opt,
                        vdf_csv_table** out);
vdf_status vdf_csv_info(vdf_csv_table* t, uint64_t* nrow, int* ncol);
vdf_status vdf_csv_column_get(vdf_csv_table* t, int c, vdf_csv_column* out);
vdf_status vdf_csv_free(vdf_csv_table* t);

/* ---- library / handle ---- */
const char* vdf_last_error(void);
const char* vdf_version(void);

/* Create a handle on `device`.  hbm_budget_bytes caps the handle's scratch
 * arena (0 = no cap beyond the device's free memory). */
vdf_status vdf_create(vdf_handle** out, int device, uint64_t hbm_budget_bytes);
vdf_status vdf_destroy(vdf_handle* h);

/* Scratch bytes vdf_argsort / vdf_groupby need for n rows (for budgeting). */
uint64_t vdf_argsort_scratch_bytes(int key_type, uint64_t n, uint32_t flags, int want_keys);
uint64_t vdf_groupby_scratch_bytes(int key_type, uint64_t n);
/* Scratch bytes vdf_groupby_batch needs for these tables and value types on a
 * handle with the default launch shape (what a request server reserves for a
 * batch); 0 for an unknown key or value type.  Host arithmetic only. */
uint64_t vdf_groupby_batch_scratch_bytes(int key_type, const uint64_t* n, int ntables, const int* val_types,
                                         int nvals);

/* ---- stable LSD radix argsort ----
 * d_perm_out[r] = row id of the r-th smallest key (stable: equal keys keep
 * row order).  d_sorted_keys_out (nullable) receives the keys in that order.
 * Replaces sort_{uintegers,integers,char} + radix_sort_* (see header). */
vdf_status vdf_argsort(vdf_handle* h, int key_type, const void* d_keys, uint64_t n,
                       uint32_t flags, void* d_perm_out, void* d_sorted_keys_out,
                       void* stream);

/* ---- permutation gather: d_dst[r] = d_src[d_perm[r]] ----
 * elem_bytes 1 .. VDF_MAX_STRING_BYTES (4, 8 and 16 have vectorised kernels;
 * other widths, e.g. padded strings, are copied row by row); perm is
 * uint32_t[n] or (perm64) uint64_t[n].
 * Replaces permute_block_mt.inl:39-41 for one column. */
vdf_status vdf_gather(vdf_handle* h, uint32_t elem_bytes, const void* d_src,
                      const void* d_perm, int perm64, uint64_t n, void* d_dst,
                      void* stream);

/* ---- sort with one value column carried along ----
 * The hard sort_by of the reference sorts the row ids and then permutes every
 * column through them (sort_by_mt.hpp:117-158, permute_block_mt.inl:39-41).
 * Here one value column of val_bytes (4 or 8) bytes per row moves with the
 * keys through the radix passes, so d_vals_out[r] = d_vals[perm[r]] is
 * produced by the sort itself, with no random-access gather.  Results are
 * identical to vdf_argsort followed by vdf_gather.  d_perm_out and
 * d_sorted_keys_out are nullable; flags as vdf_argsort. */
vdf_status vdf_sort_by(vdf_handle* h, int key_type, const void* d_keys, uint64_t n, uint32_t flags,
                       void* d_perm_out, void* d_sorted_keys_out, uint32_t val_bytes, const void* d_vals,
                       void* d_vals_out, void* stream);

/* ---- vdf_sort_by without allocating ----
 * No counterpart in the reference.  For callers that cannot allocate: code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc.  Same arguments and results as vdf_sort_by, but the call makes
 * no device or pinned-host allocation: its scratch is one fixed device array
 * of VDF_NOALLOC_SCRATCH_BYTES per device (held from the library's load until
 * the process ends; its address is taken in vdf_create), and the sort's
 * descriptors are written by a kernel instead of an upload.  A sort whose
 * scratch exceeds the array fails with VDF_ERR_INVALID before any work: with
 * an 8-byte column that is above about 2.0M rows for 4-byte keys, 1.6M for
 * 8-byte keys and 1.1M for 16-byte keys.  Every call on one device shares
 * the array: calls that overlap on one device (other streams, other handles
 * or threads) race and must be serialised by the caller.  Profiling is off
 * for these calls. */
#define VDF_NOALLOC_SCRATCH_BYTES (64ull << 20)
vdf_status vdf_sort_by_noalloc(vdf_handle* h, int key_type, const void* d_keys, uint64_t n, uint32_t flags,
                               void* d_perm_out, void* d_sorted_keys_out, uint32_t val_bytes, const void* d_vals,
                               void* d_vals_out, void* stream);

/* ---- sort-based group-by ---- */
typedef struct vdf_value_col {
    int type;          /* VDF_VAL_* */
    const void* data;  /* device, n elements */
} vdf_value_col;

typedef struct vdf_agg_out {
    void* sum;    /* f64 column: double[G]; integer column: int64_t[G], mod 2^64 */
    double* mean; /* double[G]: sum / count (integer columns: exact sum / count) */
    void* min;    /* value type [G]; first of equal values in row order */
    void* max;    /* value type [G] */
} vdf_agg_out;

typedef struct vdf_groupby_out {
    uint64_t* n_groups;  /* device, 1 element: G (always the true group count)      */
    void* group_keys;    /* device, key type, C groups: keys in ascending order      */
    uint32_t* offsets;   /* device, C + 1 entries: group g is perm[offsets[g] ..     */
                         /* offsets[g+1]) ; offsets[G] = n                          */
    uint32_t* counts;    /* device, C groups (nullable)                             */
    uint32_t* perm;      /* device, n row ids (nullable): group-contiguous row ids  */
                         /* in row order within each group (row_view_idx analogue) */
    vdf_agg_out aggs[VDF_MAX_VALUE_COLS]; /* per value column, C groups; null      */
                                          /* members skipped                       */
    uint64_t capacity;   /* C: groups the per-group outputs hold; 0 means n.        */
                         /* When G > C only groups 0 .. C-1 are written (offsets    */
                         /* up to C, no sentinel) and *n_groups still reports G, so */
                         /* the caller sees the truncation and can retry larger.    */
} vdf_groupby_out;

/* Group rows by one key column and aggregate up to VDF_MAX_VALUE_COLS value
 * columns.  Groups come out in ascending key order (the reference emits
 * first-occurrence order; compare after ordering by key).  agg_mask selects
 * which members of out->aggs are written (null members are skipped too).
 * Replaces transform_group_by (see header). */
vdf_status vdf_groupby(vdf_handle* h, int key_type, const void* d_keys, uint64_t n,
                       const vdf_value_col* vals, int nvals, uint32_t agg_mask,
                       vdf_groupby_out* out, void* stream);

/* ---- multi-column keys ----
 * Rows are keyed by several columns of any key type, byte strings included
 * (their byte widths summing to at most VDF_MAX_STRING_BYTES); equal keys are
 * equal in every column, as the reference's concatenated key bytes
 * (transform_group_by_sametype_mt.hpp:13, transform_group_by_difftype_mt.hpp:13,
 * key_build.inl:3-110).  Order: lexicographic over the columns, first column
 * most significant, each column in its own order (as vdf_argsort; byte
 * strings as vdf_argsort_bytes).  The columns are packed into one composite
 * key on the device in order-preserving big-endian form: up to 16 bytes the
 * sort moves 4, 8 or 16 key bytes per row; wider composites are fixed-width
 * byte rows sorted and grouped as vdf_argsort_bytes / vdf_groupby_bytes do. */
typedef struct vdf_key_col {
    int type;           /* VDF_KEY_* */
    const void* data;   /* device, n cells */
    uint32_t width;     /* VDF_KEY_BYTES: bytes per cell (1 .. VDF_MAX_STRING_BYTES); else ignored */
} vdf_key_col;

/* Group by several key columns.  out->group_keys is ignored; group_keys[c]
 * (device, capacity C cells of column c's type; nullable) receives column c's
 * group keys. */
vdf_status vdf_groupby_multi(vdf_handle* h, const vdf_key_col* keys, int nkeys, uint64_t n,
                             const vdf_value_col* vals, int nvals, uint32_t agg_mask,
                             vdf_groupby_out* out, void* const* group_keys, void* stream);
/* Stable argsort by several key columns (flags as vdf_argsort; descending
 * applies to the whole key). */
vdf_status vdf_argsort_multi(vdf_handle* h, const vdf_key_col* keys, int nkeys, uint64_t n, uint32_t flags,
                             void* d_perm_out, void* stream);

/* Stable argsort of n fixed-width byte strings: `width` bytes per row
 * (1 .. VDF_MAX_STRING_BYTES), rows contiguous, shorter strings padded by the
 * caller (the reference pads with PaddingChar = 0x00).  Order: bytewise as
 * signed chars (each byte ^ 0x80), first byte most significant -- the order of
 * sort_string (sort_string.inl:10-85, with radix_sort_charbuf_flat.inl).
 * Flags as vdf_argsort.  d_sorted_out (device, n * width bytes; nullable)
 * receives the rows in sorted order. */
vdf_status vdf_argsort_bytes(vdf_handle* h, const void* d_keys, uint32_t width, uint64_t n, uint32_t flags,
                             void* d_perm_out, void* d_sorted_out, void* stream);
/* Group by fixed-width byte strings (rows and order as vdf_argsort_bytes):
 * the outputs of vdf_groupby, with out->group_keys (capacity * width bytes;
 * nullable) receiving each group's string.  Replaces transform_group_by over a
 * std::string key column (onecol_functions.inl:4-130, keys hashed as string_view
 * of the string's bytes); groups come in ascending string order. */
vdf_status vdf_groupby_bytes(vdf_handle* h, const void* d_keys, uint32_t width, uint64_t n,
                             const vdf_value_col* vals, int nvals, uint32_t agg_mask,
                             vdf_groupby_out* out, void* stream);

/* ---- cached grouping ----
 * Sort one key column once and aggregate value columns against it any number
 * of times (boundaries, counts and aggregates are recomputed from the kept
 * sorted keys and permutation; no further sort).  Replaces the
 * already-grouped path of the reference (transform_group_by_alrd_mt.hpp:10,
 * functions_alrd.inl:13-40), which keeps its grouping to aggregate more
 * columns.  The grouping holds n * (key bytes + 4) bytes of device memory and
 * uses its handle (which must outlive it) for scratch. */
typedef struct vdf_grouping vdf_grouping;
vdf_status vdf_grouping_create(vdf_handle* h, int key_type, const void* d_keys, uint64_t n, void* stream,
                               vdf_grouping** out);
/* As vdf_groupby on the grouping's key column (out->perm, when set, receives
 * a copy of the kept permutation). */
vdf_status vdf_grouping_aggregate(vdf_grouping* g, const vdf_value_col* vals, int nvals, uint32_t agg_mask,
                                  vdf_groupby_out* out, void* stream);
vdf_status vdf_grouping_destroy(vdf_grouping* g);

/* ---- per-row broadcast of group results ----
 * d_out[perm[p]] = value of the group that sorted position p belongs to, for
 * the outputs of one group-by (values d_group_vals of src_type, offsets,
 * n_groups, perm = the group-contiguous row ids); dst_type is src_type or
 * VDF_VAL_F64 (converted).  Replaces CreateValueCol (create_value_col.inl:36-61),
 * the per-row column of transform_group_by.  Writes nothing when the group
 * count exceeds `capacity`. */
vdf_status vdf_broadcast(vdf_handle* h, int src_type, int dst_type, const void* d_group_vals,
                         const uint32_t* d_offsets, const uint64_t* d_n_groups, uint64_t capacity,
                         const uint32_t* d_perm, uint64_t n, void* d_out, void* stream);

/* ---- batches of independent tables (segments) ----
 * One launch sequence sorts / groups every table of the batch: one histogram
 * launch, one plan launch and one persistent launch per radix pass cover all
 * tables; each table gets its own digit plan (digits constant within the table
 * are skipped for it), its own look-back chain and its own outputs.  A table's
 * results are identical to a call with that table alone.
 * No reference counterpart: the reference sorts one frame per call
 * (sort_by_mt.hpp:10); this is the batched form of vdf_argsort / vdf_groupby. */
typedef struct vdf_sort_seg {
    const void* keys;       /* device, n keys */
    uint64_t n;             /* <= VDF_MAX_ROWS */
    void* perm_out;         /* device, n row ids (table-local), u32 or u64 (VDF_SORT_PERM64) */
    void* sorted_keys_out;  /* device, n keys (nullable) */
} vdf_sort_seg;

vdf_status vdf_argsort_batch(vdf_handle* h, int key_type, const vdf_sort_seg* segs, int nsegs,
                             uint32_t flags, void* stream);

typedef struct vdf_table {
    const void* keys;                       /* device, n keys */
    uint64_t n;                             /* <= VDF_MAX_ROWS */
    const void* vals[VDF_MAX_VALUE_COLS];   /* device value columns, types given per call */
} vdf_table;

/* outs[i] receives table i's groups (same layout as vdf_groupby). */
vdf_status vdf_groupby_batch(vdf_handle* h, int key_type, const vdf_table* tables, int ntables,
                             const int* val_types, int nvals, uint32_t agg_mask,
                             vdf_groupby_out* outs, void* stream);

/* ---- tuning table (per handle) ----
 * Knobs (unknown names or unsupported values -> VDF_ERR_INVALID):
 *   "sort_items_32"  rows per thread of a radix pass tile, 4-byte keys: 0 (default: by
 *                    payload width -- 32 without a payload, 24 with an 8-byte one, 12
 *                    with a 4- or 16-byte one), 8, 12, 16, 24, 32; a size without a
 *                    kernel for a sort's payload width uses that width's default
 *   "sort_items_64"  same for 8-byte keys: 0 (default: 12; 16 with an 8-byte payload, 8 with
 *                    a 16-byte one),
 *                    8, 12, 16
 *   "sort_items_128" same for 16-byte keys: 0 (default: 8), 8, 16
 *   "sort_chunk"     consecutive tiles one XCD claims at a time: 1 .. 1024
 *   "sort_xcd"       1 (default): XCD-grouped tile order (each XCD works through chunks of
 *                    consecutive tiles); 0: global tile order.  Both are plain launches
 *                    that need no co-residency, so sorts may run beside other work
 *   "sort_blocks_per_cu"  radix pass blocks per CU: 0 (default: every resident block) .. 32
 *   "hist_rows"      rows per digit-histogram work item: 4096 .. 2^24
 *   "gather_rows"    rows per thread of the gather kernel: 4, 8, 16
 *   "gather_nt"      1 = non-temporal loads of the gathered column
 *   "csv_stage"      1 (default): vdf_read_csv's row kernels parse each wave's rows from a
 *                    copy in LDS; 0: from the text in memory
 *   "groupby_carry"  1 (default): the group-by's sort carries value column 0 (4 or 8
 *                    bytes; columns 0 and 1 when both are 8 bytes) into key order, so
 *                    the aggregate kernel reads them sequentially; 0: every column is
 *                    gathered through the permutation
 *   "sort_local"     1 (default): a one-table sort of 4-byte keys whose four 8-bit
 *                    digits all vary scatters only its top two globally and sorts the
 *                    buckets of equal top digits (up to 3840 rows each; 4096 from 2^24
 *                    rows on) in LDS, one
 *                    streaming pass that writes the outputs; decided on the device from
 *                    the digit histograms, with the plain passes as the fallback;
 *                    0: plain passes only
 *   "sort_rank"      1 (default): the radix passes and the local finish rank rows with one
 *                    LDS atomic per row where the device serves same-address lanes in lane
 *                    order (checked once per device by vdf_create); 0: wave ballots.
 *                    vdf_tune_get("sort_rank_atomic") reads back (only) whether the
 *                    atomic ranking is in use on the handle's device
 * No reference counterpart: the reference fixes its digit width and thread
 * count at compile time (radix_sort_uint32_mt.inl:3-5). */
vdf_status vdf_tune_set(vdf_handle* h, const char* name, int64_t value);
vdf_status vdf_tune_get(vdf_handle* h, const char* name, int64_t* value);

/* ---- autotuner ----
 * Times the caller's workload (one ordinary call first validates the
 * arguments) under candidate launch shapes and stores the fastest in the
 * handle's tuning table: rows per thread of a radix pass tile (tile size),
 * radix pass blocks per CU (block count), XCD tile order and chunk, rows per
 * histogram work item.  The digit width is fixed at 8 bits (256 buckets, one
 * per thread of a 256-thread block).  Coordinate descent, each candidate timed
 * as the minimum of `reps` runs; outputs are written by every run and are the
 * same for every candidate.
 * No reference counterpart (radix_sort_uint32_mt.inl:3-5 fixes the shape). */
typedef struct vdf_tune_report {
    int trials;
    int sort_items;          /* rows per thread of a pass tile for this key width (0: default) */
    int sort_blocks_per_cu;  /* 0 = every resident block */
    int sort_xcd;
    int sort_chunk;
    int hist_rows;
    int digit_bits;
    int sort_local;          /* 4-byte keys: 1 = LDS local finish after two global passes */
    double default_ms;       /* the workload under the tuning table before the call */
    double best_ms;
} vdf_tune_report;

vdf_status vdf_autotune_argsort(vdf_handle* h, int key_type, const void* d_keys, uint64_t n, uint32_t flags,
                                void* d_perm_out, int reps, void* stream, vdf_tune_report* report);
/* As vdf_autotune_argsort for vdf_sort_by (a value column carried along). */
vdf_status vdf_autotune_sort_by(vdf_handle* h, int key_type, const void* d_keys, uint64_t n, uint32_t flags,
                                void* d_perm_out, uint32_t val_bytes, const void* d_vals, void* d_vals_out, int reps,
                                void* stream, vdf_tune_report* report);
vdf_status vdf_autotune_groupby_batch(vdf_handle* h, int key_type, const vdf_table* tables, int ntables,
                                      const int* val_types, int nvals, uint32_t agg_mask,
                                      vdf_groupby_out* outs, int reps, void* stream,
                                      vdf_tune_report* report);

/* ---- HIP-graph-captured batch plans ----
 * vdf_plan_groupby_batch captures the whole launch sequence of a
 * vdf_groupby_batch call for one batch shape (the tables' row counts, key and
 * value types, aggregates) into HIP graphs with their own scratch (it runs
 * nothing).  vdf_plan_launch replays it for any tables and outputs of that
 * shape: the host rewrites the batch description the graph uploads first,
 * then launches the graph on `stream`.  Replays of one plan share its scratch:
 * launch them on one stream or order them.  The tuning table is the handle's
 * at capture.
 * No reference counterpart. */
typedef struct vdf_plan vdf_plan;
vdf_status vdf_plan_groupby_batch(vdf_handle* h, int key_type, const vdf_table* tables, int ntables,
                                  const int* val_types, int nvals, uint32_t agg_mask, vdf_plan** out);
vdf_status vdf_plan_launch(vdf_plan* p, const vdf_table* tables, int ntables, vdf_groupby_out* outs,
                           void* stream);
vdf_status vdf_plan_scratch_bytes(vdf_plan* p, uint64_t* bytes);
vdf_status vdf_plan_destroy(vdf_plan* p);

/* ---- per-kernel timing (HIP events on the launch stream) ----
 * While enabled, every kernel launch of this handle is bracketed by events.
 * vdf_profile_read synchronises the recorded events and returns, for one
 * kernel name, total milliseconds, launch count and algorithmic HBM bytes
 * summed over launches that did work. */
vdf_status vdf_profile_enable(vdf_handle* h, int on);
vdf_status vdf_profile_reset(vdf_handle* h);
vdf_status vdf_profile_read(vdf_handle* h, const char* kernel, double* total_ms,
                            uint64_t* launches, uint64_t* bytes);

/* ---- group-by request server: device-resident tables, continuous batching ----
 * A server owns one HBM arena of arena_bytes (allocated at create) and
 * `workers` host threads, each with its own stream.  Tables are registered
 * once (their columns stay in the caller's device memory and must outlive the
 * server).  Submitted requests wait in one FIFO queue; whenever a worker is
 * free it takes the oldest request plus later requests of the same key type,
 * value types and aggregate mask (up to max_batch_tables requests and
 * max_batch_rows rows, the oldest always taken), reserves each request's
 * result block and the batch's scratch in the arena, and runs them as one
 * vdf_groupby_batch launch sequence.  Requests that arrive while batches run
 * are admitted by the next free worker as soon as arena memory allows, so
 * batches of different workers overlap on the device.  A request that does
 * not fit waits (the oldest request is never passed for memory); one that can
 * never fit fails with VDF_ERR_BUDGET.  Results stay in the arena until
 * vdf_server_release.  A request's results are identical to vdf_groupby on
 * its table alone.
 * No reference counterpart: VisonDF has no request loop; each group-by runs
 * on the calling thread (transform_group_by.hpp:11-14). */
typedef struct vdf_server vdf_server;

typedef struct vdf_server_config {
    int device;
    int workers;              /* launch sequences in flight at once: 1 .. 16            */
    int max_batch_tables;     /* requests per launch sequence: 1 .. 65536               */
    int reserved;
    uint64_t arena_bytes;     /* HBM for scratch and results (the admission budget)     */
    uint64_t max_batch_rows;  /* rows per launch sequence (0 = arena_bytes / (192 *     */
                              /* workers), at least 2^20; the oldest request is always  */
                              /* taken whatever its rows)                               */
} vdf_server_config;

typedef struct vdf_served {
    vdf_status status;        /* of the request (VDF_OK or why it failed)               */
    uint32_t batch_tables;    /* requests in the launch sequence that ran it            */
    uint64_t n_groups;
    uint64_t submit_ns;       /* steady-clock stamps: submitted, launch started, done   */
    uint64_t start_ns;
    uint64_t done_ns;
    vdf_groupby_out out;      /* device pointers into the arena (valid until release);  */
                              /* out.n_groups is NULL (the count is n_groups above)     */
} vdf_served;

typedef struct vdf_server_stats {
    uint64_t submitted, completed, failed;
    uint64_t launches;        /* launch sequences (batches) run                         */
    uint64_t retries;         /* requests rerun with a larger group capacity            */
    uint64_t rows;            /* rows of completed requests                             */
    uint64_t queued;          /* waiting now                                            */
    uint64_t arena_bytes, arena_used, arena_peak;
} vdf_server_stats;

vdf_status vdf_server_create(vdf_server** out, const vdf_server_config* cfg);
/* Finishes every queued request, then frees the arena and the workers. */
vdf_status vdf_server_destroy(vdf_server* s);
/* Registers a device-resident table.  max_groups sizes its per-group result
 * arrays (0 = n); a request whose table has more groups is rerun once with
 * the exact count. */
vdf_status vdf_server_add_table(vdf_server* s, int key_type, const void* d_keys, uint64_t n,
                                const int* val_types, const void* const* d_vals, int nvals,
                                uint64_t max_groups, uint32_t* table_id);
/* Queues a group-by of a registered table (want_perm: also the group-contiguous
 * row ids).  Returns at once. */
vdf_status vdf_server_submit(vdf_server* s, uint32_t table_id, uint32_t agg_mask, int want_perm,
                             uint64_t* request_id);
/* Waits up to timeout_us (< 0: no limit) for a request; VDF_ERR_TIMEOUT if it
 * is not done.  On VDF_OK, *out describes the request (its own status too). */
vdf_status vdf_server_wait(vdf_server* s, uint64_t request_id, int64_t timeout_us, vdf_served* out);
/* Frees a finished request's results (their arena memory admits new work). */
vdf_status vdf_server_release(vdf_server* s, uint64_t request_id);
vdf_status vdf_server_stats_get(vdf_server* s, vdf_server_stats* out);

#ifdef __cplusplus
}
#endif
#endif /* VDF_H */
