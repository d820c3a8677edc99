/*
 * oracle.h -- CPU restatement of VisonDF's radix-argsort / gather / group-by path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: tests/, the smoke
 * check in __graft_entry__.py and the cpu_baseline leg of bench.py may load it.
 * The product (visondf_amd/libvdf_hip.so) never links, loads or calls it.
 *
 * Every function names the reference file:line it restates (paths relative to
 * the reference's include/vison/ directory).
 *
 * Key type codes match include/vdf.h:
 *   0 = u32, 1 = i32, 2 = u64, 3 = i64, 4 = char[16] (fixed-width bytes)
 * Value type codes match include/vdf.h:
 *   0 = f64, 1 = i64, 2 = i32, 3 = u32, 4 = u64
 */
#ifndef VDF_ORACLE_H
#define VDF_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- stable LSD radix sorts with 16-bit digits (the reference's scalar path) ---- */
void orc_radix_sort_u32(uint32_t* tkeys, uint64_t* idx, size_t n);
void orc_radix_sort_u32_mt(uint32_t* tkeys, uint64_t* idx, size_t n, int threads);
void orc_radix_sort_u64(uint64_t* tkeys, uint64_t* idx, size_t n);
void orc_radix_sort_u64_mt(uint64_t* tkeys, uint64_t* idx, size_t n, int threads);
void orc_radix_sort_charbuf(const uint8_t* bufv, size_t n, size_t width, uint64_t* idx);

/* ---- sort_by dispatch: key transform + radix + order mode ----
 * mode 0 = ascending, 1 = stable descending, 2 = ascending then reversed
 * (the reference code's descending path).  idx is written (iota in, perm out).
 * threads <= 1 selects the single-threaded reference routine. */
int orc_argsort(int key_type, const void* keys, size_t n, int mode, int threads,
                uint64_t* perm_out);

/* ---- permutation gather: dst[r] = src[idx[r]] ---- */
void orc_gather(size_t elem_bytes, const void* src, const uint64_t* idx, size_t n,
                void* dst, int threads);

/* ---- exact double summation (correctly rounded, Shewchuk expansion) ---- */
double orc_exact_sum(const double* x, size_t n);

/* ---- sort-based group-by (the build's contract) ----
 * Groups in ascending key order.  Per group: key, first offset in perm,
 * count, and per value column sum / mean / min / max.
 *   sums[c]  : double* for f64 columns, int64_t* (mod 2^64) for integer columns
 *   means[c] : double*
 *   mins[c], maxs[c] : same element type as the value column
 * Any output pointer may be NULL.  Returns the number of groups. */
uint64_t orc_groupby(int key_type, const void* keys, size_t n,
                     int nvals, const int* val_types, const void* const* vals,
                     void* group_keys, uint64_t* offsets, uint64_t* counts,
                     uint64_t* perm,
                     void* const* sums, double* const* means,
                     void* const* mins, void* const* maxs, int threads);

/* ---- hash group-by restated from the reference (CPU baseline) ----
 * First-occurrence group order, row-order accumulation in the value type,
 * per-thread maps merged in thread order.  One f64 value column.
 * refapi != 0 runs one hash pass per aggregate as the reference API would
 * (Occurence, Sum, Mean, Gather+min, Gather+max); otherwise a fused pass.
 * Outputs are in first-occurrence order; returns the number of groups. */
uint64_t orc_hash_groupby_f64(int key_type, const void* keys, size_t n,
                              const double* val, int threads, int refapi,
                              void* group_keys, uint64_t* counts, double* sums,
                              double* means, double* mins, double* maxs);

/* the reference's scalar key hash for keys of <= 16 bytes */
uint64_t orc_hash_bytes16(const void* p, size_t len);

/* ---- CSV ingest restated from read_csv (read/read_csv/read_csv.hpp:9-216) ----
 * Rows are the text between newlines (a '\r' before a newline is dropped; a
 * last line without a newline is a row); fields split at delim outside quote
 * pairs and keep their quote characters.  Kinds (0 string, 1 char, 2 bool,
 * 3 int, 4 uint, 5 double) come from the first 10 data rows
 * (inlines/classify_column.inl:13-130, with signed and exponent numbers
 * counted as numbers).  Conversions: the longest valid prefix in the grammar
 * of std::from_chars / fast_float::from_chars, 0 when there is none or an
 * integer is out of range; doubles through strtod of that prefix (correctly
 * rounded).  OpenMP over rows.
 *   orc_csv_infer   column kinds from the header and first rows; returns ncol
 *                   (at most max_cols; -1 when more)
 *   orc_csv_lines   line starts: line[0] = 0, line[k + 1] one past the k-th
 *                   newline, and a sentinel bytes + 1 after an unterminated
 *                   last line; returns the line count (line needs count + 2)
 *   orc_csv_measure longest field of each string column (maxlen[c]) over the
 *                   rows line[first ..]; returns the first row whose field
 *                   count differs from ncol, or -1
 *   orc_csv_convert every field into its column (outs[c]: nrow cells of
 *                   widths[c] bytes; char cells padded with `trailing`,
 *                   string cells with 0x00) */
int orc_csv_infer(const char* text, size_t bytes, char delim, char quote, int header, int char_width,
                  int* kinds, int max_cols);
uint64_t orc_csv_lines(const char* text, size_t bytes, uint64_t* line, int threads);
int64_t orc_csv_measure(const char* text, size_t bytes, const uint64_t* line, uint64_t nrow, char delim,
                        char quote, int ncol, const int* kinds, uint32_t* maxlen, int threads);
void orc_csv_convert(const char* text, size_t bytes, const uint64_t* line, uint64_t nrow, char delim,
                     char quote, char trailing, int ncol, const int* kinds, const uint32_t* widths,
                     void* const* outs, int threads);

#ifdef __cplusplus
}
#endif
#endif
