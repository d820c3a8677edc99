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
