// radix.hip -- stable LSD radix argsort for gfx950 (MI355X).
//
// Replaces the reference's 16-bit-digit CPU radix (inlines/radix_sorts/
// radix_sort_uint32{,_mt}.inl, radix_sort_uint64{,_mt}.inl,
// radix_sort_charbuf_flat.inl) and the key transforms of
// inlines/transform/sort_by/sort_{uintegers,integers,char}.inl.
//
// Launch sequence for one sort of n rows (all on one stream, no host sync):
//   1. k_fill        zero the digit histograms and pass tickets
//   2. k_histogram   one read of the keys -> 256-bucket histogram of every
//                    8-bit digit (LDS-privatised), clears look-back buffer 0
//   3. k_plan        exclusive bucket offsets per digit; digits whose
//                    histogram holds all n rows in one bucket are skipped
//   4. k_pass x slots  one scatter pass per digit slot; slots beyond the
//                    plan's pass count exit at once.  Each 256-thread block
//                    takes a tile id from an atomic ticket, ranks its tile
//                    stably with wave64 ballots, publishes per-bucket counts
//                    to a decoupled look-back array (Merrill & Garland 2016;
//                    Adinets & Merrill 2022 "Onesweep"), stages the tile in
//                    LDS in bucket order and writes coalesced runs.
//
// Keys are mapped to an unsigned order-preserving form on first load
// (u32/u64: identity, i32/i64: sign bit flipped, char[16]: bytes ^ 0x80 read
// big-endian; descending: all bits complemented) and mapped back when the
// last pass writes the sorted keys.
#include "vdf_internal.h"

#include <algorithm>

namespace vdf {

struct K128 {
    uint64_t lo, hi;
};

template <typename K>
struct KeyOps;

template <>
struct KeyOps<uint32_t> {
    static constexpr int kDigits = 4;
    __device__ static uint32_t load(const void* p, uint64_t i, uint64_t mlo, uint64_t)
    {
        return static_cast<const uint32_t*>(p)[i] ^ static_cast<uint32_t>(mlo);
    }
    __device__ static void store(void* p, uint64_t i, uint32_t k, uint64_t mlo, uint64_t)
    {
        static_cast<uint32_t*>(p)[i] = k ^ static_cast<uint32_t>(mlo);
    }
    __device__ static uint32_t digit(uint32_t k, int d) { return (k >> (8 * d)) & 0xFFu; }
    __device__ static uint32_t zero() { return 0u; }
};

template <>
struct KeyOps<uint64_t> {
    static constexpr int kDigits = 8;
    __device__ static uint64_t load(const void* p, uint64_t i, uint64_t mlo, uint64_t)
    {
        return static_cast<const uint64_t*>(p)[i] ^ mlo;
    }
    __device__ static void store(void* p, uint64_t i, uint64_t k, uint64_t mlo, uint64_t)
    {
        static_cast<uint64_t*>(p)[i] = k ^ mlo;
    }
    __device__ static uint32_t digit(uint64_t k, int d)
    {
        return static_cast<uint32_t>(k >> (8 * d)) & 0xFFu;
    }
    __device__ static uint64_t zero() { return 0ull; }
};

// char[16]: byte 0 is the most significant digit (radix_sort_charbuf_flat.inl:94
// sorts from the last byte to the first).  The transformed key holds bytes
// 0..7 big-endian in `hi` and bytes 8..15 in `lo`.
template <>
struct KeyOps<K128> {
    static constexpr int kDigits = 16;
    __device__ static K128 load(const void* p, uint64_t i, uint64_t mlo, uint64_t mhi)
    {
        const uint64_t* w = static_cast<const uint64_t*>(p) + 2 * i;
        K128 k;
        k.hi = __builtin_bswap64(w[0]) ^ mhi;
        k.lo = __builtin_bswap64(w[1]) ^ mlo;
        return k;
    }
    __device__ static void store(void* p, uint64_t i, K128 k, uint64_t mlo, uint64_t mhi)
    {
        uint64_t* w = static_cast<uint64_t*>(p) + 2 * i;
        w[0] = __builtin_bswap64(k.hi ^ mhi);
        w[1] = __builtin_bswap64(k.lo ^ mlo);
    }
    __device__ static uint32_t digit(K128 k, int d)
    {
        return d < 8 ? static_cast<uint32_t>(k.lo >> (8 * d)) & 0xFFu
                     : static_cast<uint32_t>(k.hi >> (8 * (d - 8))) & 0xFFu;
    }
    __device__ static K128 zero() { return K128{0ull, 0ull}; }
};

// rows per thread in a scatter tile, by key width (LDS and VGPR budget)
template <typename K>
struct PassItems {
    static constexpr int value = 16;
};
template <>
struct PassItems<uint64_t> {
    static constexpr int value = 12;
};
template <>
struct PassItems<K128> {
    static constexpr int value = 8;
};

__device__ inline void store_relaxed(uint32_t* p, uint32_t v)
{
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ inline uint32_t load_relaxed(uint32_t* p)
{
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Exclusive scan of one value per thread over a 256-thread block.
__device__ inline uint32_t block_excl_scan(uint32_t x, uint32_t* s_wsum)
{
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    uint32_t inc = x;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const uint32_t y = __shfl_up(inc, off, 64);
        if (lane >= off) inc += y;
    }
    if (lane == 63) s_wsum[w] = inc;
    __syncthreads();
    uint32_t pre = 0;
#pragma unroll
    for (int i = 0; i < kWaves; ++i) pre += (i < w) ? s_wsum[i] : 0u;
    __syncthreads();
    return pre + inc - x;
}

__global__ void k_fill(uint32_t* p, uint64_t count, uint32_t value)
{
    const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
    for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < count; i += stride)
        p[i] = value;
}

void fill_u32(uint32_t* p, size_t count, uint32_t value, hipStream_t s)
{
    if (count == 0) return;
    const unsigned grid = static_cast<unsigned>(std::min<size_t>((count + 255) / 256, 2048));
    hipLaunchKernelGGL(k_fill, dim3(grid), dim3(256), 0, s, p, static_cast<uint64_t>(count), value);
    check_launch("k_fill");
}

// One read of the keys; per-block LDS histograms of every digit, folded into
// the global histogram with one atomic per non-empty bucket.  Also clears the
// first look-back buffer.
template <typename K>
__global__ __launch_bounds__(kBlock) void k_histogram(const void* keys, uint64_t n, uint64_t mlo,
                                                      uint64_t mhi, uint32_t* hist, uint32_t* lb0,
                                                      uint64_t lb_words)
{
    using KO = KeyOps<K>;
    constexpr int D = KO::kDigits;
    __shared__ uint32_t sh[D * kRadix];
    for (int j = threadIdx.x; j < D * kRadix; j += kBlock) sh[j] = 0;
    __syncthreads();
    const uint64_t stride = static_cast<uint64_t>(gridDim.x) * kBlock;
    uint64_t i = static_cast<uint64_t>(blockIdx.x) * kBlock + threadIdx.x;
    for (; i + 3 * stride < n; i += 4 * stride) {
        K k[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) k[u] = KO::load(keys, i + u * stride, mlo, mhi);
#pragma unroll
        for (int u = 0; u < 4; ++u)
#pragma unroll
            for (int d = 0; d < D; ++d) atomicAdd(&sh[d * kRadix + KO::digit(k[u], d)], 1u);
    }
    for (; i < n; i += stride) {
        const K k = KO::load(keys, i, mlo, mhi);
#pragma unroll
        for (int d = 0; d < D; ++d) atomicAdd(&sh[d * kRadix + KO::digit(k, d)], 1u);
    }
    __syncthreads();
    for (int j = threadIdx.x; j < D * kRadix; j += kBlock) {
        const uint32_t c = sh[j];
        if (c) atomicAdd(&hist[j], c);
    }
    for (uint64_t w = static_cast<uint64_t>(blockIdx.x) * kBlock + threadIdx.x; w < lb_words; w += stride)
        lb0[w] = 0u;
}

// Single block: exclusive bucket offsets for every digit and the list of
// digits that need a pass.
__global__ __launch_bounds__(kBlock) void k_plan(const uint32_t* hist, int ndig, uint32_t n, Plan* plan)
{
    __shared__ uint32_t s_wsum[kWaves];
    __shared__ uint32_t s_triv[kMaxDigits];
    const int t = threadIdx.x;
    if (t < kMaxDigits) s_triv[t] = 0;
    __syncthreads();
    for (int d = 0; d < ndig; ++d) {
        const uint32_t c = hist[d * kRadix + t];
        if (c == n) s_triv[d] = 1;
        plan->base[d][t] = block_excl_scan(c, s_wsum);
    }
    __syncthreads();
    if (t == 0) {
        uint32_t D = 0;
        for (int d = 0; d < ndig; ++d)
            if (!s_triv[d]) plan->digit_of[D++] = static_cast<uint32_t>(d);
        if (D == 0) {  // every digit constant: one pass still writes perm and keys
            plan->digit_of[0] = 0;
            D = 1;
        }
        plan->passes = D;
    }
}

struct PassArgs {
    const void* keys_in;   // caller's keys (read by the first pass, transformed on load)
    void* kbuf[2];         // transformed keys between passes
    void* keys_out;        // caller's sorted-keys output (nullable), written by the last pass
    uint32_t* vbuf[2];     // row ids between passes
    void* perm_out;        // caller's permutation, written by the last pass
    uint32_t* lb[2];       // look-back words, [tiles][256], alternating by slot parity
    uint32_t* tickets;     // [kMaxDigits] dynamic tile ids
    const Plan* plan;
    uint64_t n;
    uint64_t mlo, mhi;     // key transform masks
    uint32_t perm64;
    uint32_t reverse;
};

template <typename K, int ITEMS>
__global__ __launch_bounds__(kBlock) void k_pass(PassArgs a, uint32_t slot)
{
    using KO = KeyOps<K>;
    constexpr int TILE = kBlock * ITEMS;
    const uint32_t D = a.plan->passes;
    if (slot >= D) return;
    const int dig = static_cast<int>(a.plan->digit_of[slot]);
    const bool first = (slot == 0);
    const bool last = (slot == D - 1);

    __shared__ K s_keys[TILE];
    __shared__ uint32_t s_vals[TILE];
    __shared__ uint32_t s_wc[kWaves][kRadix];
    __shared__ uint32_t s_gbase[kRadix];
    __shared__ uint32_t s_wsum[kWaves];
    __shared__ uint32_t s_tile;

    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    if (tid == 0) s_tile = atomicAdd(&a.tickets[slot], 1u);
#pragma unroll
    for (int w = 0; w < kWaves; ++w) s_wc[w][tid] = 0;
    __syncthreads();
    const uint32_t tile = s_tile;
    const uint64_t tbase = static_cast<uint64_t>(tile) * TILE;

    uint32_t* lb = a.lb[slot & 1];
    a.lb[(slot + 1) & 1][static_cast<uint64_t>(tile) * kRadix + tid] = 0u;  // next pass's row

    // ---- load: wave w owns rows [w*64*ITEMS, (w+1)*64*ITEMS) of the tile,
    //      item j of lane l is row j*64 + l (coalesced, row order = (j, l)) ----
    K key[ITEMS];
    uint32_t val[ITEMS];
    uint32_t rnk[ITEMS];
    const uint64_t wbase = tbase + static_cast<uint64_t>(wave) * 64 * ITEMS + lane;
    if (first) {
#pragma unroll
        for (int j = 0; j < ITEMS; ++j) {
            const uint64_t i = wbase + static_cast<uint64_t>(j) * 64;
            key[j] = (i < a.n) ? KO::load(a.keys_in, i, a.mlo, a.mhi) : KO::zero();
            val[j] = static_cast<uint32_t>(i);
        }
    } else {
        const K* kin = static_cast<const K*>(a.kbuf[(slot - 1) & 1]);
        const uint32_t* vin = a.vbuf[(slot - 1) & 1];
#pragma unroll
        for (int j = 0; j < ITEMS; ++j) {
            const uint64_t i = wbase + static_cast<uint64_t>(j) * 64;
            const bool v = i < a.n;
            key[j] = v ? kin[i] : KO::zero();
            val[j] = v ? vin[i] : 0u;
        }
    }

    // ---- stable rank inside the wave.  For item j, `dif` collects the lanes
    //      whose digit differs from this lane's in any bit (8 ballots of the
    //      sign-extended bit); peers = valid lanes outside `dif`.  The lowest
    //      peer adds the group size to the wave's bucket counter and every peer
    //      reads the counter back right after; LDS operations of one wave
    //      complete in program order, so these reads see exactly the items
    //      before j and the 2*ITEMS LDS operations pipeline without waits. ----
    uint32_t* wc = s_wc[wave];
#pragma unroll
    for (int j = 0; j < ITEMS; ++j) {
        const bool v = (wbase + static_cast<uint64_t>(j) * 64) < a.n;
        const uint32_t d = KO::digit(key[j], dig);
        uint32_t dif_lo = 0u, dif_hi = 0u;
#pragma unroll
        for (int b = 0; b < kRadixBits; ++b) {
            const uint32_t m = static_cast<uint32_t>(static_cast<int32_t>(d << (31 - b)) >> 31);
            const uint64_t bb = __ballot(m != 0u);
            dif_lo |= static_cast<uint32_t>(bb) ^ m;
            dif_hi |= static_cast<uint32_t>(bb >> 32) ^ m;
        }
        const uint64_t vm = __ballot(v);
        const uint32_t p_lo = ~dif_lo & static_cast<uint32_t>(vm);
        const uint32_t p_hi = ~dif_hi & static_cast<uint32_t>(vm >> 32);
        const uint32_t lt = __builtin_amdgcn_mbcnt_hi(p_hi, __builtin_amdgcn_mbcnt_lo(p_lo, 0u));
        const uint32_t cnt = static_cast<uint32_t>(__popc(p_lo) + __popc(p_hi));
        if (v && lt == 0u) atomicAdd(&wc[d], cnt);
        uint32_t after = 0u;
        if (v) after = wc[d];
        rnk[j] = after - cnt + lt;
    }
    __syncthreads();

    // ---- thread tid owns bucket tid: wave-exclusive prefixes, tile total ----
    uint32_t tot = 0;
#pragma unroll
    for (int w = 0; w < kWaves; ++w) {
        const uint32_t c = s_wc[w][tid];
        s_wc[w][tid] = tot;
        tot += c;
    }
    uint32_t* my = lb + static_cast<uint64_t>(tile) * kRadix + tid;
    store_relaxed(my, (tile == 0 ? kFlagP : kFlagA) | tot);
    const uint32_t lstart = block_excl_scan(tot, s_wsum);
#pragma unroll
    for (int w = 0; w < kWaves; ++w) s_wc[w][tid] += lstart;

    // ---- decoupled look-back over earlier tiles of this bucket ----
    uint32_t excl = 0;
    if (tile > 0) {
        uint32_t k = tile - 1;
        for (;;) {
            const uint32_t w = load_relaxed(lb + static_cast<uint64_t>(k) * kRadix + tid);
            const uint32_t f = w & ~kValMask;
            if (f == 0u) {
                __builtin_amdgcn_s_sleep(1);
                continue;
            }
            excl += w & kValMask;
            if (f == kFlagP) break;
            --k;
        }
        store_relaxed(my, kFlagP | (excl + tot));
    }
    s_gbase[tid] = a.plan->base[dig][tid] + excl - lstart;
    __syncthreads();

    // ---- stage the tile in LDS in bucket order ----
#pragma unroll
    for (int j = 0; j < ITEMS; ++j) {
        if (wbase + static_cast<uint64_t>(j) * 64 < a.n) {
            const uint32_t d = KO::digit(key[j], dig);
            const uint32_t p = s_wc[wave][d] + rnk[j];
            s_keys[p] = key[j];
            s_vals[p] = val[j];
        }
    }
    __syncthreads();

    // ---- write: consecutive LDS slots of one bucket go to consecutive rows ----
    const uint32_t tn = static_cast<uint32_t>(std::min<uint64_t>(TILE, a.n - tbase));
    if (!last) {
        K* kout = static_cast<K*>(a.kbuf[slot & 1]);
        uint32_t* vout = a.vbuf[slot & 1];
        for (uint32_t i = tid; i < tn; i += kBlock) {
            const K k = s_keys[i];
            const uint32_t p = s_gbase[KO::digit(k, dig)] + i;
            kout[p] = k;
            vout[p] = s_vals[i];
        }
    } else {
        const uint32_t nm1 = static_cast<uint32_t>(a.n - 1);
        for (uint32_t i = tid; i < tn; i += kBlock) {
            const K k = s_keys[i];
            uint32_t p = s_gbase[KO::digit(k, dig)] + i;
            if (a.reverse) p = nm1 - p;
            if (a.keys_out) KO::store(a.keys_out, p, k, a.mlo, a.mhi);
            if (a.perm64)
                static_cast<uint64_t*>(a.perm_out)[p] = s_vals[i];
            else
                static_cast<uint32_t*>(a.perm_out)[p] = s_vals[i];
        }
    }
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
size_t key_bytes(int key_type)
{
    switch (key_type) {
    case VDF_KEY_U32:
    case VDF_KEY_I32: return 4;
    case VDF_KEY_U64:
    case VDF_KEY_I64: return 8;
    case VDF_KEY_CHAR16: return 16;
    default: return 0;
    }
}

static int pass_items(int key_type)
{
    switch (key_type) {
    case VDF_KEY_U32:
    case VDF_KEY_I32: return PassItems<uint32_t>::value;
    case VDF_KEY_U64:
    case VDF_KEY_I64: return PassItems<uint64_t>::value;
    default: return PassItems<K128>::value;
    }
}


struct SortLayout {
    uint64_t ntiles;
    size_t zero_words;  // hist + tickets
    size_t bytes;
};

static SortLayout sort_layout(int key_type, uint64_t n)
{
    SortLayout L;
    const uint64_t tile = static_cast<uint64_t>(kBlock) * pass_items(key_type);
    L.ntiles = (n + tile - 1) / tile;
    L.zero_words = kMaxDigits * kRadix + kMaxDigits;
    const size_t kb = key_bytes(key_type);
    size_t b = 0;
    b += align_up(L.zero_words * 4, 256);
    b += align_up(sizeof(Plan), 256);
    b += 2 * align_up(L.ntiles * kRadix * 4, 256);
    b += 2 * align_up(n * kb, 256);
    b += 2 * align_up(n * 4, 256);
    L.bytes = b + 256;
    return L;
}

size_t argsort_scratch(int key_type, uint64_t n, uint32_t) { return sort_layout(key_type, n).bytes; }

template <typename K>
static void launch_sort(Handle* h, const void* d_keys, uint64_t n, uint64_t mlo, uint64_t mhi,
                        uint32_t flags, void* d_perm_out, void* d_keys_out, void* scr, hipStream_t s,
                        int key_type)
{
    constexpr int ITEMS = PassItems<K>::value;
    constexpr int D = KeyOps<K>::kDigits;
    const SortLayout L = sort_layout(key_type, n);
    Carve c(scr);
    uint32_t* zero = c.take<uint32_t>(L.zero_words);
    uint32_t* hist = zero;
    uint32_t* tickets = zero + kMaxDigits * kRadix;
    Plan* plan = c.take<Plan>(1);
    PassArgs a{};
    a.lb[0] = c.take<uint32_t>(L.ntiles * kRadix);
    a.lb[1] = c.take<uint32_t>(L.ntiles * kRadix);
    a.kbuf[0] = c.take<char>(n * sizeof(K));
    a.kbuf[1] = c.take<char>(n * sizeof(K));
    a.vbuf[0] = c.take<uint32_t>(n);
    a.vbuf[1] = c.take<uint32_t>(n);
    a.keys_in = d_keys;
    a.keys_out = d_keys_out;
    a.perm_out = d_perm_out;
    a.tickets = tickets;
    a.plan = plan;
    a.n = n;
    a.mlo = mlo;
    a.mhi = mhi;
    a.perm64 = (flags & VDF_SORT_PERM64) ? 1u : 0u;
    a.reverse = (flags & VDF_SORT_DESC_REVERSE) ? 1u : 0u;

    const size_t kb = sizeof(K);
    {
        LaunchTimer t(h, s, "fill", 0);
        fill_u32(zero, L.zero_words, 0u, s);
    }
    {
        const unsigned grid = static_cast<unsigned>(std::min<uint64_t>(std::max<uint64_t>(L.ntiles, 1), 2048));
        LaunchTimer t(h, s, "radix_histogram", n * kb);
        hipLaunchKernelGGL(k_histogram<K>, dim3(grid), dim3(kBlock), 0, s, d_keys, n, mlo, mhi, hist, a.lb[0],
                           L.ntiles * kRadix);
        check_launch("k_histogram");
    }
    uint32_t* passes_host = nullptr;
    {
        LaunchTimer t(h, s, "radix_plan", 0);
        hipLaunchKernelGGL(k_plan, dim3(1), dim3(kBlock), 0, s, hist, D, static_cast<uint32_t>(n), plan);
        check_launch("k_plan");
    }
    if (h->prof) {
        passes_host = h->passes_ring + (h->ring_pos++ % Handle::kRing);
        VDF_HIP_CHECK(hipMemcpyAsync(passes_host, &plan->passes, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
    }
    // algorithmic bytes of a slot when it is the only / first / middle / last pass:
    // keys read + keys written + row ids read + row ids written
    const uint64_t pb_out = a.perm64 ? 8 : 4;
    const uint64_t kout = d_keys_out ? kb : 0;
    const uint64_t pass_bytes[4] = {n * (kb + kout + pb_out), n * (kb + kb + 4), n * (2 * kb + 8),
                                    n * (kb + 4 + kout + pb_out)};
    for (int slot = 0; slot < D; ++slot) {
        LaunchTimer t(h, s, "radix_pass", 0, slot, passes_host, pass_bytes);
        hipLaunchKernelGGL((k_pass<K, ITEMS>), dim3(static_cast<unsigned>(L.ntiles)), dim3(kBlock), 0, s, a,
                           static_cast<uint32_t>(slot));
        check_launch("k_pass");
    }
}

void argsort_enqueue(Handle* h, int key_type, const void* d_keys, uint64_t n, uint32_t flags,
                     void* d_perm_out, void* d_keys_out, void* scr, hipStream_t s)
{
    const bool desc = (flags & VDF_SORT_DESC) != 0;
    switch (key_type) {
    case VDF_KEY_U32:
    case VDF_KEY_I32: {
        uint64_t m = (key_type == VDF_KEY_I32 ? 0x80000000ull : 0ull) ^ (desc ? 0xFFFFFFFFull : 0ull);
        launch_sort<uint32_t>(h, d_keys, n, m, 0, flags, d_perm_out, d_keys_out, scr, s, key_type);
        break;
    }
    case VDF_KEY_U64:
    case VDF_KEY_I64: {
        uint64_t m = (key_type == VDF_KEY_I64 ? 0x8000000000000000ull : 0ull) ^ (desc ? ~0ull : 0ull);
        launch_sort<uint64_t>(h, d_keys, n, m, 0, flags, d_perm_out, d_keys_out, scr, s, key_type);
        break;
    }
    case VDF_KEY_CHAR16: {
        uint64_t m = 0x8080808080808080ull ^ (desc ? ~0ull : 0ull);
        launch_sort<K128>(h, d_keys, n, m, m, flags, d_perm_out, d_keys_out, scr, s, key_type);
        break;
    }
    default: throw Error{VDF_ERR_INVALID, "unknown key type"};
    }
}

}  // namespace vdf
