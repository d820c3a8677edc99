// gather.hip -- permutation gather dst[r] = src[perm[r]] for gfx950.
//
// Replaces the per-column loop of permute_block_mt (inlines/transform/sort_by/
// permute_block_mt.inl:39-41).  Row ids are read coalesced (striped over the
// block), each thread keeps ROWS independent random reads in flight, and the
// outputs are written coalesced.
#include "vdf_internal.h"

#include <algorithm>

namespace vdf {

typedef uint64_t V16 __attribute__((ext_vector_type(2)));  // 16-byte cells (char[16] keys)

template <typename V, typename P, int ROWS, bool NT>
__global__ __launch_bounds__(256) void k_gather(const V* __restrict__ src, const P* __restrict__ perm, uint64_t n,
                                                V* __restrict__ dst)
{
    const uint64_t base = static_cast<uint64_t>(blockIdx.x) * (256 * ROWS) + threadIdx.x;
    P idx[ROWS];
#pragma unroll
    for (int u = 0; u < ROWS; ++u) {
        const uint64_t i = base + static_cast<uint64_t>(u) * 256;
        idx[u] = (i < n) ? __builtin_nontemporal_load(perm + i) : P(0);
    }
    V v[ROWS];
#pragma unroll
    for (int u = 0; u < ROWS; ++u) {
        const uint64_t i = base + static_cast<uint64_t>(u) * 256;
        if (i < n) v[u] = NT ? __builtin_nontemporal_load(src + idx[u]) : src[idx[u]];
    }
#pragma unroll
    for (int u = 0; u < ROWS; ++u) {
        const uint64_t i = base + static_cast<uint64_t>(u) * 256;
        if (i < n) dst[i] = v[u];
    }
}

template <typename V, typename P, int ROWS, bool NT>
static void launch_rows(const void* src, const void* perm, uint64_t n, void* dst, hipStream_t s)
{
    const uint64_t per_block = 256 * ROWS;
    const unsigned grid = static_cast<unsigned>((n + per_block - 1) / per_block);
    hipLaunchKernelGGL((k_gather<V, P, ROWS, NT>), dim3(grid), dim3(256), 0, s, static_cast<const V*>(src),
                       static_cast<const P*>(perm), n, static_cast<V*>(dst));
    check_launch("k_gather");
}

template <typename V, typename P>
static void launch_gather(const Tuning& t, const void* src, const void* perm, uint64_t n, void* dst, hipStream_t s)
{
    if (t.gather_nt) {
        if (t.gather_rows == 4) return launch_rows<V, P, 4, true>(src, perm, n, dst, s);
        if (t.gather_rows == 16) return launch_rows<V, P, 16, true>(src, perm, n, dst, s);
        return launch_rows<V, P, 8, true>(src, perm, n, dst, s);
    }
    if (t.gather_rows == 4) return launch_rows<V, P, 4, false>(src, perm, n, dst, s);
    if (t.gather_rows == 16) return launch_rows<V, P, 16, false>(src, perm, n, dst, s);
    return launch_rows<V, P, 8, false>(src, perm, n, dst, s);
}

void gather_enqueue(Handle* h, uint32_t elem_bytes, const void* d_src, const void* d_perm, int perm64,
                    uint64_t n, void* d_dst, hipStream_t s)
{
    if (n == 0) return;
    const uint64_t pb = perm64 ? 8 : 4;
    LaunchTimer t(h, s, "gather", n * (pb + 2ull * elem_bytes));
    switch (elem_bytes) {
    case 4:
        perm64 ? launch_gather<uint32_t, uint64_t>(h->tune, d_src, d_perm, n, d_dst, s)
               : launch_gather<uint32_t, uint32_t>(h->tune, d_src, d_perm, n, d_dst, s);
        break;
    case 8:
        perm64 ? launch_gather<uint64_t, uint64_t>(h->tune, d_src, d_perm, n, d_dst, s)
               : launch_gather<uint64_t, uint32_t>(h->tune, d_src, d_perm, n, d_dst, s);
        break;
    case 16:
        perm64 ? launch_gather<V16, uint64_t>(h->tune, d_src, d_perm, n, d_dst, s)
               : launch_gather<V16, uint32_t>(h->tune, d_src, d_perm, n, d_dst, s);
        break;
    default: throw Error{VDF_ERR_INVALID, "vdf_gather: elem_bytes must be 4, 8 or 16"};
    }
}

}  // namespace vdf
