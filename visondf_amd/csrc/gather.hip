// gather.hip -- permutation gather dst[r] = src[perm[r]] for gfx950.
//
// Replaces the per-column loop of permute_block_mt (inlines/transform/sort_by/
// permute_block_mt.inl:39-41).  Row ids are read coalesced (striped over the
// block), each thread keeps ROWS independent random reads in flight, and the
// outputs are written coalesced.
#include "vdf_internal.h"

#include <algorithm>
#include <type_traits>

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

// ---------------------------------------------------------------------------
// Per-row broadcast of per-group values: out[perm[p]] = val[g] for every
// sorted position p of group g (offsets[g] <= p < offsets[g+1]), converted to
// the output type.  This is CreateValueCol of the reference
// (create_value_col.inl:36-61), driven by the group-by's outputs instead of a
// per-row hash lookup.  Each thread takes 16 consecutive sorted positions:
// one binary search over the offsets for the first, then a forward walk.
// Nothing is written when the group count exceeds the arrays' capacity.
// ---------------------------------------------------------------------------
template <typename S, typename D>
__global__ __launch_bounds__(256) void k_broadcast(const S* __restrict__ gval, const uint32_t* __restrict__ offsets,
                                                   const uint64_t* __restrict__ n_groups, uint64_t cap,
                                                   const uint32_t* __restrict__ perm, uint64_t n, D* __restrict__ out)
{
    constexpr int R = 16;
    const uint64_t p0 = (static_cast<uint64_t>(blockIdx.x) * 256 + threadIdx.x) * R;
    if (p0 >= n) return;
    const uint64_t G = *n_groups;
    if (G == 0 || G > cap) return;
    // last group whose first row is <= p0 (offsets[0] = 0, offsets[G] = n > p0)
    uint64_t lo = 0, hi = G;
    while (hi - lo > 1) {
        const uint64_t mid = (lo + hi) >> 1;
        if (offsets[mid] <= p0) lo = mid;
        else hi = mid;
    }
    uint64_t g = lo;
    uint64_t next = offsets[g + 1];
    S v = gval[g];
    for (int j = 0; j < R; ++j) {
        const uint64_t p = p0 + j;
        if (p >= n) break;
        while (p >= next) {  // groups are non-empty: one step per group boundary
            ++g;
            next = offsets[g + 1];
            v = gval[g];
        }
        out[perm[p]] = static_cast<D>(v);
    }
}

template <typename S>
static void launch_broadcast_from(int dst_type, const void* gval, const uint32_t* offsets, const uint64_t* ng,
                                  uint64_t cap, const uint32_t* perm, uint64_t n, void* out, hipStream_t s)
{
    const unsigned grid = static_cast<unsigned>((n + 256 * 16 - 1) / (256 * 16));
    if (dst_type == VDF_VAL_F64 && !std::is_same<S, double>::value)
        hipLaunchKernelGGL((k_broadcast<S, double>), dim3(grid), dim3(256), 0, s, static_cast<const S*>(gval),
                           offsets, ng, cap, perm, n, static_cast<double*>(out));
    else
        hipLaunchKernelGGL((k_broadcast<S, S>), dim3(grid), dim3(256), 0, s, static_cast<const S*>(gval), offsets,
                           ng, cap, perm, n, static_cast<S*>(out));
    check_launch("k_broadcast");
}

void broadcast_enqueue(Handle* h, int src_type, int dst_type, const void* gval, const uint32_t* offsets,
                       const uint64_t* n_groups, uint64_t cap, const uint32_t* perm, uint64_t n, void* out,
                       hipStream_t s)
{
    if (n == 0) return;
    const uint64_t ob = dst_type == VDF_VAL_F64 ? 8 : static_cast<uint64_t>(val_bytes(src_type));
    LaunchTimer t(h, s, "broadcast", n * (4 + ob));
    switch (src_type) {
    case VDF_VAL_F64: launch_broadcast_from<double>(dst_type, gval, offsets, n_groups, cap, perm, n, out, s); break;
    case VDF_VAL_I64: launch_broadcast_from<int64_t>(dst_type, gval, offsets, n_groups, cap, perm, n, out, s); break;
    case VDF_VAL_U64: launch_broadcast_from<uint64_t>(dst_type, gval, offsets, n_groups, cap, perm, n, out, s); break;
    case VDF_VAL_I32: launch_broadcast_from<int32_t>(dst_type, gval, offsets, n_groups, cap, perm, n, out, s); break;
    case VDF_VAL_U32: launch_broadcast_from<uint32_t>(dst_type, gval, offsets, n_groups, cap, perm, n, out, s); break;
    default: throw Error{VDF_ERR_INVALID, "vdf_broadcast: unknown source type"};
    }
}

// Cells of any other width (padded strings): one row per thread, copied in
// 4-byte words when the width and both buffers allow (words), else bytes.
template <typename IDX>
__global__ void k_gather_rows(const uint8_t* src, uint32_t elem, const IDX* perm, uint64_t n, uint8_t* dst,
                              uint32_t words)
{
    const uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint64_t r = gmem(perm)[i];
    if (words) {
        const auto* a = gmem(reinterpret_cast<const uint32_t*>(src + r * elem));
        auto* b = gmem(reinterpret_cast<uint32_t*>(dst + i * elem));
        for (uint32_t j = 0; j < elem / 4; ++j) b[j] = a[j];
    } else {
        const auto* a = gmem(src) + r * elem;
        auto* b = gmem(dst) + i * elem;
        for (uint32_t j = 0; j < elem; ++j) b[j] = a[j];
    }
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
    default: {
        const dim3 grid(static_cast<unsigned>((n + 255) / 256));
        const auto* src = static_cast<const uint8_t*>(d_src);
        auto* dst = static_cast<uint8_t*>(d_dst);
        const uint32_t words =
            ((elem_bytes | reinterpret_cast<uintptr_t>(src) | reinterpret_cast<uintptr_t>(dst)) & 3u) == 0 ? 1u : 0u;
        if (perm64)
            hipLaunchKernelGGL(k_gather_rows<uint64_t>, grid, dim3(256), 0, s, src, elem_bytes,
                               static_cast<const uint64_t*>(d_perm), n, dst, words);
        else
            hipLaunchKernelGGL(k_gather_rows<uint32_t>, grid, dim3(256), 0, s, src, elem_bytes,
                               static_cast<const uint32_t*>(d_perm), n, dst, words);
        check_launch("k_gather_rows");
        break;
    }
    }
}

}  // namespace vdf
