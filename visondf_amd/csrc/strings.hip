// strings.hip -- stable argsort of fixed-width byte strings of any width.
//
// The reference sorts std::string keys by padding every string to the longest
// one with PaddingChar (0x00 by default), mapping each byte b to b ^ 0x80
// (signed char order) and running a byte-wise LSD radix sort from the last
// byte to the first (sort_string.inl:10-85, radix_sort_charbuf_flat.inl:4-107).
// Here the caller hands over the padded rows (`width` bytes each, contiguous)
// and the sort runs as an LSD over 16-byte chunks, last chunk first: each
// chunk, read through the permutation so far, is sorted stably by the
// char[16] radix path (which applies the same ^ 0x80 byte order), and the
// chunk's permutation is composed onto the running one.  Stable sorts compose
// into the lexicographic order of the chunks, i.e. of the bytes.  Bytes past
// the width in the last chunk are 0x00 in every row, so they never change an
// order (and the char[16] plan skips their constant digits).
#include "vdf_internal.h"

namespace vdf {

namespace {

// chunk c of row perm[i] (identity when perm is null): 16 bytes, zero-filled
// past the width
__global__ void k_chunk(const uint8_t* keys, uint32_t width, uint32_t c, const uint32_t* perm, uint64_t n,
                        uint64_t* out)
{
    const uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint64_t row = perm ? gmem(perm)[i] : i;
    const auto* src = gmem(keys) + row * width + 16ull * c;
    const uint32_t have = std::min<uint32_t>(16u, width - 16u * c);
    uint64_t w[2] = {0, 0};
    for (uint32_t j = 0; j < have; ++j) w[j >> 3] |= static_cast<uint64_t>(src[j]) << (8 * (j & 7));
    auto* o = gmem(out) + 2 * i;
    o[0] = w[0];
    o[1] = w[1];
}

// next[i] = cur[pc[i]]: the running permutation after one more stable sort
__global__ void k_compose(const uint32_t* cur, const uint32_t* pc, uint64_t n, uint32_t* next)
{
    const uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) gmem(next)[i] = gmem(cur)[gmem(pc)[i]];
}

// the caller's permutation: reversed for VDF_SORT_DESC_REVERSE, 32- or 64-bit
__global__ void k_perm_out(const uint32_t* cur, uint64_t n, uint32_t reverse, uint32_t perm64, void* out)
{
    const uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t v = gmem(cur)[reverse ? n - 1 - i : i];
    if (perm64) gmem(static_cast<uint64_t*>(out))[i] = v;
    else gmem(static_cast<uint32_t*>(out))[i] = v;
}

unsigned blocks_for(uint64_t n) { return static_cast<unsigned>((n + 255) / 256); }

// ---- group-by over byte strings ----
// Sorted rows become dense group ids (u32: 0 for the first group, +1 at every
// row whose string differs from the previous row's), which the aggregate stage
// of the numeric group-by then segments exactly as the strings would be.
constexpr int kIdRows = 16;                     // rows per thread
constexpr int kIdTile = kBlock * kIdRows;       // rows per block

__device__ inline uint32_t is_head(const uint8_t* rows, uint32_t width, uint64_t i)
{
    if (i == 0) return 1u;
    const auto* a = gmem(rows) + i * width;
    const auto* b = a - width;
    for (uint32_t j = 0; j < width; ++j)
        if (a[j] != b[j]) return 1u;
    return 0u;
}

// heads of the thread's rows (bit j: row t0 + j), rows past n count nothing
__device__ inline uint32_t head_bits(const uint8_t* rows, uint32_t width, uint64_t t0, uint64_t n)
{
    uint32_t hb = 0;
    for (int j = 0; j < kIdRows; ++j)
        if (t0 + j < n && is_head(rows, width, t0 + j)) hb |= 1u << j;
    return hb;
}

// heads per block of kIdTile rows
__global__ __launch_bounds__(kBlock) void k_head_counts(const uint8_t* rows, uint32_t width, uint64_t n, uint32_t* cnt)
{
    __shared__ uint32_t s_sum[kWaves];
    const uint64_t t0 = static_cast<uint64_t>(blockIdx.x) * kIdTile + static_cast<uint64_t>(threadIdx.x) * kIdRows;
    uint32_t c = __popc(head_bits(rows, width, t0, n));
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) c += __shfl_xor(c, off, 64);
    if ((threadIdx.x & 63) == 0) s_sum[threadIdx.x >> 6] = c;
    __syncthreads();
    if (threadIdx.x == 0) gmem(cnt)[blockIdx.x] = s_sum[0] + s_sum[1] + s_sum[2] + s_sum[3];
}

// exclusive scan of the per-block counts in place (one block; each thread
// scans a contiguous run)
__global__ __launch_bounds__(kBlock) void k_scan_counts(uint32_t* cnt, uint32_t nb)
{
    __shared__ uint32_t s_wsum[kWaves];
    const uint32_t per = (nb + kBlock - 1) / kBlock;
    const uint32_t b0 = threadIdx.x * per, b1 = std::min(nb, b0 + per);
    uint32_t sum = 0;
    for (uint32_t b = b0; b < b1; ++b) sum += gmem(cnt)[b];
    uint32_t run = block_excl_scan(sum, s_wsum);
    for (uint32_t b = b0; b < b1; ++b) {
        const uint32_t c = gmem(cnt)[b];
        gmem(cnt)[b] = run;
        run += c;
    }
}

// dense group id of every row (rows n .. padded end repeat the last id)
__global__ __launch_bounds__(kBlock) void k_group_ids(const uint8_t* rows, uint32_t width, uint64_t n,
                                                      uint64_t padded, const uint32_t* cnt, uint32_t* ids)
{
    __shared__ uint32_t s_wsum[kWaves];
    const uint64_t t0 = static_cast<uint64_t>(blockIdx.x) * kIdTile + static_cast<uint64_t>(threadIdx.x) * kIdRows;
    const uint32_t hb = head_bits(rows, width, t0, n);
    uint32_t id = gmem(cnt)[blockIdx.x] + block_excl_scan(__popc(hb), s_wsum) - 1u;  // id before the thread's rows
    for (int j = 0; j < kIdRows; ++j) {
        id += (hb >> j) & 1u;
        if (t0 + j < padded) gmem(ids)[t0 + j] = id;
    }
}

// group g's string: the sorted row at offsets[g]
__global__ void k_group_rows(const uint8_t* rows, uint32_t width, const uint32_t* offsets, const uint64_t* n_groups,
                             uint64_t cap, uint8_t* out)
{
    const uint64_t g = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    const uint64_t ng = *gmem(n_groups);
    if (g >= std::min(ng, cap)) return;
    const auto* src = gmem(rows) + static_cast<uint64_t>(gmem(offsets)[g]) * width;
    auto* dst = gmem(out) + g * width;
    for (uint32_t j = 0; j < width; ++j) dst[j] = src[j];
}

}  // namespace

size_t bytes_argsort_scratch(uint64_t n, uint64_t hist_rows)
{
    return align_up(n * 16, 256) + 3 * align_up(n * 4, 256) +
           argsort_batch_scratch(VDF_KEY_CHAR16, &n, 1, hist_rows) + 1024;
}

size_t bytes_argsort_staging(uint32_t width, uint64_t n, uint64_t hist_rows)
{
    return static_cast<size_t>((width + 15) / 16) * (argsort_batch_staging(VDF_KEY_CHAR16, &n, 1, hist_rows) + 256);
}

void bytes_argsort_enqueue(Handle* h, const void* d_keys, uint32_t width, uint64_t n, uint32_t flags,
                           void* d_perm_out, void* d_sorted_out, void* scr, hipStream_t s)
{
    const uint64_t hr = static_cast<uint64_t>(h->tune.hist_rows);
    Carve cv(scr);
    uint64_t* chunk = cv.take<uint64_t>(2 * n);
    uint32_t* pc = cv.take<uint32_t>(n);
    uint32_t* cur = cv.take<uint32_t>(n);
    uint32_t* next = cv.take<uint32_t>(n);
    void* ascr = cv.take<char>(argsort_batch_scratch(VDF_KEY_CHAR16, &n, 1, hr));
    const uint8_t* keys = static_cast<const uint8_t*>(d_keys);
    // every chunk sort ascending or stable descending; the reference's
    // reversed order is the ascending one reversed at the end
    const uint32_t cflags = flags & VDF_SORT_DESC;
    const uint32_t nch = (width + 15) / 16;
    for (uint32_t k = 0; k < nch; ++k) {
        const uint32_t c = nch - 1 - k;
        hipLaunchKernelGGL(k_chunk, dim3(blocks_for(n)), dim3(256), 0, s, keys, width, c, k ? cur : nullptr, n,
                           chunk);
        check_launch("k_chunk");
        argsort_enqueue(h, VDF_KEY_CHAR16, chunk, n, cflags, k ? pc : cur, nullptr, ascr, s);
        if (k) {
            hipLaunchKernelGGL(k_compose, dim3(blocks_for(n)), dim3(256), 0, s, cur, pc, n, next);
            check_launch("k_compose");
            std::swap(cur, next);
        }
    }
    const uint32_t perm64 = (flags & VDF_SORT_PERM64) ? 1u : 0u;
    hipLaunchKernelGGL(k_perm_out, dim3(blocks_for(n)), dim3(256), 0, s, cur, n,
                       (flags & VDF_SORT_DESC_REVERSE) ? 1u : 0u, perm64, d_perm_out);
    check_launch("k_perm_out");
    if (d_sorted_out) gather_enqueue(h, width, keys, d_perm_out, static_cast<int>(perm64), n, d_sorted_out, s);
}

}  // namespace vdf

namespace vdf {

size_t bytes_groupby_scratch(uint32_t width, uint64_t n, uint64_t hist_rows)
{
    const uint64_t nb = (n + kIdTile - 1) / kIdTile;
    return align_up(n * 4, 256) + align_up(n * width, 256) + align_up(align_up(n, 16) * 4, 256) +
           align_up(std::max<uint64_t>(nb, 1) * 4, 256) + bytes_argsort_scratch(n, hist_rows) +
           groupby_agg_scratch(&n, 1) + 1024;
}

size_t bytes_groupby_staging(uint32_t width, uint64_t n, uint64_t hist_rows)
{
    return bytes_argsort_staging(width, n, hist_rows) + groupby_agg_staging(&n, 1) + 256;
}

void bytes_groupby_enqueue(Handle* h, const void* d_keys, uint32_t width, uint64_t n, const GbTable& t,
                           const int* val_types, int nvals, uint32_t agg_mask, const vdf_groupby_out* out, void* scr,
                           hipStream_t s)
{
    const uint64_t hr = static_cast<uint64_t>(h->tune.hist_rows);
    const uint64_t nb = (n + kIdTile - 1) / kIdTile;
    const uint64_t padded = align_up(n, 16);
    Carve cv(scr);
    uint32_t* perm = cv.take<uint32_t>(n);
    uint8_t* rows = cv.take<uint8_t>(n * width);
    uint32_t* ids = cv.take<uint32_t>(padded);
    uint32_t* cnt = cv.take<uint32_t>(std::max<uint64_t>(nb, 1));
    void* sscr = cv.take<char>(bytes_argsort_scratch(n, hr));
    void* ascr = cv.take<char>(groupby_agg_scratch(&n, 1));
    bytes_argsort_enqueue(h, d_keys, width, n, 0, perm, rows, sscr, s);
    hipLaunchKernelGGL(k_head_counts, dim3(static_cast<unsigned>(nb)), dim3(kBlock), 0, s, rows, width, n, cnt);
    check_launch("k_head_counts");
    hipLaunchKernelGGL(k_scan_counts, dim3(1), dim3(kBlock), 0, s, cnt, static_cast<uint32_t>(nb));
    check_launch("k_scan_counts");
    hipLaunchKernelGGL(k_group_ids, dim3(static_cast<unsigned>(nb)), dim3(kBlock), 0, s, rows, width, n, padded,
                       cnt, ids);
    check_launch("k_group_ids");
    if (out->perm) VDF_HIP_CHECK(hipMemcpyAsync(out->perm, perm, n * 4, hipMemcpyDeviceToDevice, s));
    vdf_groupby_out o = *out;
    o.group_keys = nullptr;  // the ids; the strings are written below
    const uint32_t* perms[1] = {perm};
    groupby_aggregate_enqueue(h, VDF_KEY_U32, ids, perms, &t, 1, val_types, nvals, agg_mask, &o, ascr, s);
    if (out->group_keys) {
        const uint64_t cap = (out->capacity == 0 || out->capacity > n) ? n : out->capacity;
        hipLaunchKernelGGL(k_group_rows, dim3(blocks_for(cap)), dim3(256), 0, s, rows, width, out->offsets,
                           out->n_groups, cap, static_cast<uint8_t*>(out->group_keys));
        check_launch("k_group_rows");
    }
}

}  // namespace vdf
