// radix.hip -- stable LSD radix argsort for gfx950 (MI355X).
//
// Replaces the reference's 16-bit-digit CPU radix (inlines/radix_sorts/
// radix_sort_uint32{,_mt}.inl, radix_sort_uint64{,_mt}.inl,
// radix_sort_charbuf_flat.inl) and the key transforms of
// inlines/transform/sort_by/sort_{uintegers,integers,char}.inl.
//
// Launch sequence for one sort of n rows (all on one stream, no host sync):
//   1. k_fill        zero the digit histograms and tile counters
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
// Adds one key per lane to the block histogram.  Digits flagged in `shared`
// (wave-uniform) were constant across a sample of the wave's keys; for those
// the wave first checks whether every lane has the same digit and then adds
// the count once, instead of 64 atomics on one LDS word.
template <typename K>
__device__ inline void hist_add(uint32_t* sh, const K& k, bool valid, uint32_t shared)
{
    using KO = KeyOps<K>;
    const uint64_t vm = __ballot(valid);
#pragma unroll
    for (int d = 0; d < KO::kDigits; ++d) {
        const uint32_t b = KO::digit(k, d);
        if ((shared >> d) & 1u) {
            const uint32_t b0 = __builtin_amdgcn_readfirstlane(b);
            if (__ballot(valid && b == b0) == vm) {
                if (__lane_id() == static_cast<uint32_t>(__ffsll(static_cast<long long>(vm)) - 1))
                    atomicAdd(&sh[d * kRadix + b0], static_cast<uint32_t>(__popcll(vm)));
                continue;
            }
        }
        if (valid) atomicAdd(&sh[d * kRadix + b], 1u);
    }
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
    // every lane of a wave runs the same number of iterations (ballots below)
    const uint64_t start = static_cast<uint64_t>(blockIdx.x) * kBlock + threadIdx.x;
    const uint64_t wave_start = start - (threadIdx.x & 63);
    uint32_t shared = 0;  // digits constant over the wave's first 64 keys
    {
        const bool v = start < n;
        const K k = v ? KO::load(keys, start, mlo, mhi) : KO::zero();
        const uint64_t vm = __ballot(v);
#pragma unroll
        for (int d = 0; d < D; ++d) {
            const uint32_t b = KO::digit(k, d);
            const uint32_t b0 = __builtin_amdgcn_readfirstlane(b);
            if (vm != 0 && __ballot(v && b == b0) == vm) shared |= 1u << d;
        }
    }
    const uint32_t lane = threadIdx.x & 63;
    uint64_t i0 = wave_start;
    // full rounds: every lane of the wave has 4 valid keys
    for (; i0 + 63 + 3 * stride < n; i0 += 4 * stride) {
        K k[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) k[u] = KO::load(keys, i0 + lane + u * stride, mlo, mhi);
#pragma unroll
        for (int u = 0; u < 4; ++u) hist_add<K>(sh, k[u], true, shared);
    }
    for (; i0 < n; i0 += stride) {
        const uint64_t i = i0 + lane;
        const bool v = i < n;
        hist_add<K>(sh, v ? KO::load(keys, i, mlo, mhi) : KO::zero(), v, shared);
    }
    __syncthreads();
    for (int j = threadIdx.x; j < D * kRadix; j += kBlock) {
        const uint32_t c = sh[j];
        if (c) atomicAdd(&hist[j], c);
    }
    for (uint64_t w = start; w < lb_words; w += stride) lb0[w] = 0u;
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
    uint32_t* tickets;     // per pass slot: one claim counter per tile group
    const Plan* plan;
    uint64_t n;
    uint64_t mlo, mhi;     // key transform masks
    uint32_t ntiles;
    uint32_t groups;       // tile groups (one per XCD when the grid allows)
    uint32_t chunk;        // consecutive tiles a group claims at a time
    uint32_t perm64;
    uint32_t reverse;
};

// LDS-only block barrier: orders LDS traffic between the block's waves
// without waiting for the block's outstanding global stores.
__device__ inline void lds_barrier() { asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory"); }

__device__ inline uint32_t block_excl_scan_lds(uint32_t x, uint32_t* s_wsum)
{
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    uint32_t inc = x;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const uint32_t y = __shfl_up(inc, off, 64);
        if (lane >= off) inc += y;
    }
    if (lane == 63) s_wsum[w] = inc;
    lds_barrier();
    uint32_t pre = 0;
#pragma unroll
    for (int i = 0; i < kWaves; ++i) pre += (i < w) ? s_wsum[i] : 0u;
    lds_barrier();
    return pre + inc - x;
}

// One radix scatter pass, persistent and software-pipelined.
//
// Per tile: load (coalesced, striped rows), stable rank inside each wave,
// publish the tile's per-bucket counts (flag A), stage the tile in LDS in
// bucket order.  The block then claims, loads and ranks its next tile and
// publishes that aggregate before it resolves the staged tile's decoupled
// look-back (flag P) and writes it out, so look-back waits overlap work.
//
// Ranking of item j in a wave: every lane ORs its lane bit into the wave's
// 64-bit mask of its digit; reading the mask back gives the lanes sharing
// the digit (peers).  The lowest peer adds the group size to the wave's
// bucket counter and clears the mask; every peer reads the counter back.
// LDS operations of one wave complete in program order, so each read sees
// exactly the items before j, and the result is the stable rank.
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
    __shared__ uint64_t s_mask[kWaves][kRadix];
    __shared__ uint32_t s_wc[kWaves][kRadix];
    __shared__ uint32_t s_gbase[kRadix];
    __shared__ uint32_t s_wsum[kWaves];
    __shared__ uint32_t s_tile;

    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    uint32_t* lb = a.lb[slot & 1];
    uint32_t* lbn = a.lb[(slot + 1) & 1];
    const K* kin = static_cast<const K*>(a.kbuf[(slot - 1) & 1]);
    const uint32_t* vin = a.vbuf[(slot - 1) & 1];
    K key[ITEMS];
    uint32_t val[ITEMS];
    uint32_t rnk[ITEMS];

    // Tile claiming.  Block b belongs to group g = b % groups; dispatch puts
    // consecutive blocks on consecutive XCDs, so a group is one XCD.  The
    // k-th claim of group g gets tile (k / chunk * groups + g) * chunk +
    // k % chunk: runs of `chunk` consecutive tiles are written by CUs sharing
    // one L2, which merges the partial lines of adjacent bucket runs.  The
    // kernel is launched cooperatively, so every block is resident; each group
    // claims its tiles in increasing order and every claimed tile publishes
    // its aggregate before its block waits, so the smallest unfinished tile
    // always completes: no deadlock.
    auto claim = [&]() -> uint32_t {
        if (tid == 0) {
            const uint32_t g = blockIdx.x % a.groups;
            const uint32_t k = atomicAdd(a.tickets + slot * kTicketStride + g, 1u);
            const uint64_t t = (static_cast<uint64_t>(k / a.chunk) * a.groups + g) * a.chunk + (k % a.chunk);
            s_tile = static_cast<uint32_t>(std::min<uint64_t>(t, 0xFFFFFFFFull));
        }
        lds_barrier();
        const uint32_t t = s_tile;
        lds_barrier();
        return t;
    };

    // load tile t, rank it, publish its aggregate; returns the bucket total and
    // tile-local bucket start of bucket `tid`; s_wc then holds each wave's
    // first LDS slot per bucket
    auto load_rank = [&](uint32_t t, uint32_t& tot, uint32_t& lstart) {
#pragma unroll
        for (int w = 0; w < kWaves; ++w) {
            s_wc[w][tid] = 0;
            s_mask[w][tid] = 0;
        }
        lbn[static_cast<uint64_t>(t) * kRadix + tid] = 0u;  // the next pass's row of this tile
        const uint64_t wbase = static_cast<uint64_t>(t) * TILE + static_cast<uint64_t>(wave) * 64 * ITEMS + lane;
        if (first) {
#pragma unroll
            for (int j = 0; j < ITEMS; ++j) {
                const uint64_t i = wbase + static_cast<uint64_t>(j) * 64;
                key[j] = (i < a.n) ? KO::load(a.keys_in, i, a.mlo, a.mhi) : KO::zero();
                val[j] = static_cast<uint32_t>(i);
            }
        } else {
#pragma unroll
            for (int j = 0; j < ITEMS; ++j) {
                const uint64_t i = wbase + static_cast<uint64_t>(j) * 64;
                const bool v = i < a.n;
                key[j] = v ? kin[i] : KO::zero();
                val[j] = v ? vin[i] : 0u;
            }
        }
        lds_barrier();
        uint32_t* wc = s_wc[wave];
        unsigned long long* mk = reinterpret_cast<unsigned long long*>(s_mask[wave]);
        const unsigned long long lanebit = 1ull << lane;
#pragma unroll
        for (int j = 0; j < ITEMS; ++j) {
            const bool v = (wbase + static_cast<uint64_t>(j) * 64) < a.n;
            const uint32_t d = KO::digit(key[j], dig);
            if (v) atomicOr(&mk[d], lanebit);
            unsigned long long peers = 0;
            if (v) peers = mk[d];
            const uint32_t pl = static_cast<uint32_t>(peers), ph = static_cast<uint32_t>(peers >> 32);
            const uint32_t lt = __builtin_amdgcn_mbcnt_hi(ph, __builtin_amdgcn_mbcnt_lo(pl, 0u));
            const uint32_t cnt = static_cast<uint32_t>(__popc(pl) + __popc(ph));
            if (v && lt == 0u) {
                atomicAdd(&wc[d], cnt);
                mk[d] = 0ull;
            }
            uint32_t after = 0;
            if (v) after = wc[d];
            rnk[j] = after - cnt + lt;
        }
        lds_barrier();
        tot = 0;
#pragma unroll
        for (int w = 0; w < kWaves; ++w) {
            const uint32_t c = s_wc[w][tid];
            s_wc[w][tid] = tot;
            tot += c;
        }
        store_relaxed(lb + static_cast<uint64_t>(t) * kRadix + tid, (t == 0 ? kFlagP : kFlagA) | tot);
        lstart = block_excl_scan_lds(tot, s_wsum);
#pragma unroll
        for (int w = 0; w < kWaves; ++w) s_wc[w][tid] += lstart;
        lds_barrier();
    };

    auto stage = [&](uint32_t t) {
        const uint64_t wbase = static_cast<uint64_t>(t) * TILE + static_cast<uint64_t>(wave) * 64 * ITEMS + lane;
#pragma unroll
        for (int j = 0; j < ITEMS; ++j) {
            if (wbase + static_cast<uint64_t>(j) * 64 < a.n) {
                const uint32_t p = s_wc[wave][KO::digit(key[j], dig)] + rnk[j];
                s_keys[p] = key[j];
                s_vals[p] = val[j];
            }
        }
        lds_barrier();
    };

    // decoupled look-back for the staged tile t (8 predecessor words per
    // step, consumed in tile order until an inclusive prefix), then write it
    auto resolve_write = [&](uint32_t t, uint32_t tot, uint32_t lstart) {
        uint32_t excl = 0;
        if (t > 0) {
            uint32_t k = t - 1;
            for (;;) {
                uint32_t w[8];
#pragma unroll
                for (int q = 0; q < 8; ++q)
                    w[q] = (k >= static_cast<uint32_t>(q)) ? load_relaxed(lb + static_cast<uint64_t>(k - q) * kRadix + tid)
                                                           : kFlagP;
                uint32_t consumed = 0;
                bool done = false, stall = false;
#pragma unroll
                for (int q = 0; q < 8; ++q) {
                    if (!done && !stall) {
                        const uint32_t f = w[q] & ~kValMask;
                        if (f == 0u) {
                            stall = true;
                        } else {
                            excl += w[q] & kValMask;
                            ++consumed;
                            done = (f == kFlagP);
                        }
                    }
                }
                if (done) break;
                k -= consumed;  // tile 0 always holds P, so k never passes below 0
                if (stall) __builtin_amdgcn_s_sleep(1);
            }
            store_relaxed(lb + static_cast<uint64_t>(t) * kRadix + tid, kFlagP | (excl + tot));
        }
        s_gbase[tid] = a.plan->base[dig][tid] + excl - lstart;
        lds_barrier();
        const uint64_t tbase = static_cast<uint64_t>(t) * TILE;
        const uint32_t tn = static_cast<uint32_t>(std::min<uint64_t>(TILE, a.n - tbase));
        if (!last) {
            K* kout = static_cast<K*>(a.kbuf[slot & 1]);
            uint32_t* vout = a.vbuf[slot & 1];
            for (uint32_t i = tid; i < tn; i += kBlock) {
                const K kk = s_keys[i];
                const uint32_t p = s_gbase[KO::digit(kk, dig)] + i;
                kout[p] = kk;
                vout[p] = s_vals[i];
            }
        } else {
            const uint32_t nm1 = static_cast<uint32_t>(a.n - 1);
            for (uint32_t i = tid; i < tn; i += kBlock) {
                const K kk = s_keys[i];
                uint32_t p = s_gbase[KO::digit(kk, dig)] + i;
                if (a.reverse) p = nm1 - p;
                if (a.keys_out) KO::store(a.keys_out, p, kk, a.mlo, a.mhi);
                if (a.perm64)
                    static_cast<uint64_t*>(a.perm_out)[p] = s_vals[i];
                else
                    static_cast<uint32_t*>(a.perm_out)[p] = s_vals[i];
            }
        }
        lds_barrier();
    };

    uint32_t t0 = claim();
    if (t0 >= a.ntiles) return;
    uint32_t tot0, ls0;
    load_rank(t0, tot0, ls0);
    stage(t0);
    for (;;) {
        const uint32_t t1 = claim();
        const bool has1 = t1 < a.ntiles;
        uint32_t tot1 = 0, ls1 = 0;
        if (has1) load_rank(t1, tot1, ls1);
        resolve_write(t0, tot0, ls0);
        if (!has1) break;
        stage(t1);
        t0 = t1;
        tot0 = tot1;
        ls0 = ls1;
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

struct SortLayout {
    uint64_t ntiles;
    size_t zero_words;  // hist + tickets
    size_t bytes;
};

// Scratch is sized for the smallest supported tile (8 rows per thread), so it
// holds every tuning choice.
static SortLayout sort_layout(int key_type, uint64_t n, int items = 8)
{
    SortLayout L;
    const uint64_t tile = static_cast<uint64_t>(kBlock) * items;
    L.ntiles = (n + tile - 1) / tile;
    L.zero_words = kMaxDigits * kRadix + kMaxDigits * kTicketStride;
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

using PassFn = void (*)(PassArgs, uint32_t);

// Pass kernel instantiations the tuning table can select.
template <typename K>
static PassFn pick_pass(int items);

template <>
PassFn pick_pass<uint32_t>(int items)
{
    if (items == 12) return k_pass<uint32_t, 12>;
    if (items == 16) return k_pass<uint32_t, 16>;
    if (items == 24) return k_pass<uint32_t, 24>;
    return nullptr;
}

template <>
PassFn pick_pass<uint64_t>(int items)
{
    if (items == 8) return k_pass<uint64_t, 8>;
    if (items == 12) return k_pass<uint64_t, 12>;
    if (items == 16) return k_pass<uint64_t, 16>;
    return nullptr;
}

template <>
PassFn pick_pass<K128>(int items)
{
    return items == 8 ? k_pass<K128, 8> : nullptr;
}

bool tune_supported(const char* name, int64_t v)
{
    const std::string n(name);
    if (n == "sort_items_32") return pick_pass<uint32_t>(static_cast<int>(v)) != nullptr;
    if (n == "sort_items_64") return pick_pass<uint64_t>(static_cast<int>(v)) != nullptr;
    if (n == "sort_chunk") return v >= 1 && v <= 1024;
    if (n == "hist_grid") return v >= 64 && v <= 65535;
    if (n == "gather_rows") return v == 4 || v == 8 || v == 16;
    if (n == "gather_nt") return v == 0 || v == 1;
    return false;
}

// Resident blocks of a pass kernel on the whole device (persistent grid).
static unsigned resident_blocks(PassFn fn, int device)
{
    struct Entry {
        PassFn fn;
        int device;
        unsigned blocks;
    };
    static std::mutex mu;
    static std::vector<Entry> cache;
    std::lock_guard<std::mutex> lk(mu);
    for (const Entry& e : cache)
        if (e.fn == fn && e.device == device) return e.blocks;
    int per_cu = 0, cus = 0;
    VDF_HIP_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, reinterpret_cast<const void*>(fn), kBlock, 0));
    VDF_HIP_CHECK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device));
    const unsigned blocks = static_cast<unsigned>(std::max(1, per_cu) * std::max(1, cus));
    cache.push_back(Entry{fn, device, blocks});
    return blocks;
}

template <typename K>
static void launch_sort(Handle* h, const void* d_keys, uint64_t n, uint64_t mlo, uint64_t mhi,
                        uint32_t flags, void* d_perm_out, void* d_keys_out, void* scr, hipStream_t s,
                        int key_type)
{
    constexpr int D = KeyOps<K>::kDigits;
    int items = sizeof(K) == 4 ? h->tune.sort_items_32 : (sizeof(K) == 8 ? h->tune.sort_items_64 : 8);
    PassFn pass = pick_pass<K>(items);
    if (!pass) {  // tuning value without an instantiation for this key width
        items = sizeof(K) == 4 ? 16 : 8;
        pass = pick_pass<K>(items);
    }
    const SortLayout L = sort_layout(key_type, n);  // scratch carve (smallest tile)
    const uint64_t ntiles = (n + static_cast<uint64_t>(kBlock) * items - 1) / (static_cast<uint64_t>(kBlock) * items);
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
    a.ntiles = static_cast<uint32_t>(ntiles);
    a.perm64 = (flags & VDF_SORT_PERM64) ? 1u : 0u;
    a.reverse = (flags & VDF_SORT_DESC_REVERSE) ? 1u : 0u;
    // persistent grid: every resident block, a multiple of the group count
    unsigned grid = static_cast<unsigned>(std::min<uint64_t>(resident_blocks(pass, h->device), ntiles));
    a.groups = std::min(8u, grid);
    grid = grid / a.groups * a.groups;
    // enough chunks that every group gets several
    a.chunk = static_cast<uint32_t>(std::max<uint64_t>(
        1, std::min<uint64_t>(static_cast<uint64_t>(h->tune.sort_chunk), ntiles / (static_cast<uint64_t>(a.groups) * 8))));

    const size_t kb = sizeof(K);
    {
        LaunchTimer t(h, s, "fill", 0);
        fill_u32(zero, L.zero_words, 0u, s);
    }
    {
        const unsigned hgrid = static_cast<unsigned>(
            std::min<uint64_t>(std::max<uint64_t>((n + 1023) / 1024, 1), static_cast<uint64_t>(h->tune.hist_grid)));
        LaunchTimer t(h, s, "radix_histogram", n * kb);
        hipLaunchKernelGGL(k_histogram<K>, dim3(hgrid), dim3(kBlock), 0, s, d_keys, n, mlo, mhi, hist, a.lb[0],
                           ntiles * kRadix);
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
        uint32_t sl = static_cast<uint32_t>(slot);
        void* args[] = {&a, &sl};
        VDF_HIP_CHECK(
            hipLaunchCooperativeKernel(reinterpret_cast<const void*>(pass), dim3(grid), dim3(kBlock), args, 0, s));
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
