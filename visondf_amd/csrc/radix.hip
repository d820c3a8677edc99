// radix.hip -- stable LSD radix argsort for gfx950 (MI355X).
//
// Replaces the reference's 16-bit-digit CPU radix (inlines/radix_sorts/
// radix_sort_uint32{,_mt}.inl, radix_sort_uint64{,_mt}.inl,
// radix_sort_charbuf_flat.inl) and the key transforms of
// inlines/transform/sort_by/sort_{uintegers,integers,char}.inl.
//
// Every sort is a batch of segments (independent tables; a plain argsort is
// one segment).  Launch sequence (one stream, no host synchronisation):
//   1. k_fill        zero the per-segment digit histograms and tile counters
//   2. k_histogram   work items of up to `hist_rows` rows of one segment: one
//                    read of the keys -> 256-bucket LDS histogram of every
//                    8-bit digit, flushed into the segment's histogram; also
//                    clears look-back buffer 0
//   3. k_plan        one block per segment: exclusive bucket offsets per
//                    digit; digits whose histogram holds all the segment's
//                    rows in one bucket are skipped for that segment
//   4. k_pass x slots  one scatter pass per digit slot over the tiles of all
//                    segments (segments with fewer passes skip their tiles);
//                    persistent, see k_pass.
//
// Keys are mapped to an unsigned order-preserving form on first load
// (u32/u64: identity, i32/i64: sign bit flipped, f32/f64: sign bit set for
// positive values and all bits complemented for negative ones, char[16]:
// bytes ^ 0x80 read big-endian; descending: all bits complemented) and mapped
// back when the last pass writes the sorted keys.
#include "vdf_internal.h"

#include <algorithm>
#include <type_traits>

namespace vdf {

struct K128 {
    uint64_t lo, hi;
};

template <typename K>
struct KeyOps;

// 4- and 8-byte keys: `mlo` is the xor mask (sign flip, descending
// complement); `mhi` != 0 marks IEEE floats, whose order-preserving form is
// the bits with the sign bit set for positive values and all bits
// complemented for negative ones (sort_flt.inl:44-50).
template <>
struct KeyOps<uint32_t> {
    static constexpr int kDigits = 4;
    // fetch() reads the caller's key; xform() maps it (kept apart so that a
    // tile's loads are all in flight before any transform waits on one)
    __device__ static uint32_t fetch(const void* p, uint64_t i) { return gmem(static_cast<const uint32_t*>(p))[i]; }
    __device__ static uint32_t xform(uint32_t x, uint64_t mlo, uint64_t mhi)
    {
        const uint32_t fm = (x >> 31) ? 0xFFFFFFFFu : 0x80000000u;
        return x ^ (mhi ? fm : 0u) ^ static_cast<uint32_t>(mlo);
    }
    __device__ static uint32_t load(const void* p, uint64_t i, uint64_t mlo, uint64_t mhi)
    {
        return xform(fetch(p, i), mlo, mhi);
    }
    // the caller's key of a transformed one
    __device__ static uint32_t unxform(uint32_t k, uint64_t mlo, uint64_t mhi)
    {
        const uint32_t x = k ^ static_cast<uint32_t>(mlo);
        return mhi ? ((x >> 31) ? (x ^ 0x80000000u) : ~x) : x;
    }
    __device__ static void store(void* p, uint64_t i, uint32_t k, uint64_t mlo, uint64_t mhi)
    {
        gmem(static_cast<uint32_t*>(p))[i] = unxform(k, mlo, mhi);
    }
    __device__ static uint32_t digit(uint32_t k, int d) { return (k >> (8 * d)) & 0xFFu; }
    __device__ static uint32_t zero() { return 0u; }
    __device__ static uint32_t raw(const uint32_t* p, uint64_t i) { return gmem(p)[i]; }
    __device__ static void put(uint32_t* p, uint64_t i, uint32_t k) { gmem(p)[i] = k; }
};

template <>
struct KeyOps<uint64_t> {
    static constexpr int kDigits = 8;
    __device__ static uint64_t fetch(const void* p, uint64_t i) { return gmem(static_cast<const uint64_t*>(p))[i]; }
    __device__ static uint64_t xform(uint64_t x, uint64_t mlo, uint64_t mhi)
    {
        const uint64_t fm = (x >> 63) ? ~0ull : 0x8000000000000000ull;
        return x ^ (mhi ? fm : 0ull) ^ mlo;
    }
    __device__ static uint64_t load(const void* p, uint64_t i, uint64_t mlo, uint64_t mhi)
    {
        return xform(fetch(p, i), mlo, mhi);
    }
    __device__ static void store(void* p, uint64_t i, uint64_t k, uint64_t mlo, uint64_t mhi)
    {
        uint64_t x = k ^ mlo;
        if (mhi) x = (x >> 63) ? (x ^ 0x8000000000000000ull) : ~x;
        gmem(static_cast<uint64_t*>(p))[i] = x;
    }
    __device__ static uint32_t digit(uint64_t k, int d)
    {
        return static_cast<uint32_t>(k >> (8 * d)) & 0xFFu;
    }
    __device__ static uint64_t zero() { return 0ull; }
    __device__ static uint64_t raw(const uint64_t* p, uint64_t i) { return gmem(p)[i]; }
    __device__ static void put(uint64_t* p, uint64_t i, uint64_t k) { gmem(p)[i] = k; }
};

// char[16]: byte 0 is the most significant digit (radix_sort_charbuf_flat.inl:94
// sorts from the last byte to the first).  The transformed key holds bytes
// 0..7 big-endian in `hi` and bytes 8..15 in `lo`.
template <>
struct KeyOps<K128> {
    static constexpr int kDigits = 16;
    __device__ static K128 fetch(const void* p, uint64_t i)
    {
        const auto* w = gmem(static_cast<const uint64_t*>(p)) + 2 * i;
        return K128{w[1], w[0]};  // bytes 8..15, bytes 0..7 (unswapped)
    }
    __device__ static K128 xform(K128 r, uint64_t mlo, uint64_t mhi)
    {
        K128 k;
        k.hi = __builtin_bswap64(r.hi) ^ mhi;
        k.lo = __builtin_bswap64(r.lo) ^ mlo;
        return k;
    }
    __device__ static K128 load(const void* p, uint64_t i, uint64_t mlo, uint64_t mhi)
    {
        return xform(fetch(p, i), mlo, mhi);
    }
    __device__ static void store(void* p, uint64_t i, K128 k, uint64_t mlo, uint64_t mhi)
    {
        auto* w = gmem(static_cast<uint64_t*>(p)) + 2 * i;
        w[0] = __builtin_bswap64(k.hi ^ mhi);
        w[1] = __builtin_bswap64(k.lo ^ mlo);
    }
    __device__ static uint32_t digit(K128 k, int d)
    {
        return d < 8 ? static_cast<uint32_t>(k.lo >> (8 * d)) & 0xFFu
                     : static_cast<uint32_t>(k.hi >> (8 * (d - 8))) & 0xFFu;
    }
    __device__ static K128 zero() { return K128{0ull, 0ull}; }
    __device__ static K128 raw(const K128* p, uint64_t i)
    {
        const auto* w = gmem(reinterpret_cast<const uint64_t*>(p)) + 2 * i;
        return K128{w[0], w[1]};
    }
    __device__ static void put(K128* p, uint64_t i, K128 k)
    {
        auto* w = gmem(reinterpret_cast<uint64_t*>(p)) + 2 * i;
        w[0] = k.lo;
        w[1] = k.hi;
    }
};

__device__ inline void store_relaxed(uint32_t* p, uint32_t v)
{
    __hip_atomic_store(gmem(p), v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ inline uint32_t load_relaxed(uint32_t* p)
{
    return __hip_atomic_load(gmem(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
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

// ---------------------------------------------------------------------------
// device-side batch description
// ---------------------------------------------------------------------------
struct SegDesc {
    const void* keys;   // the segment's keys in the caller's form
    void* perm_out;     // the segment's permutation (local row ids), u32 or u64 (nullable)
    void* keys_out;     // the segment's sorted keys (nullable)
    const void* pay_in; // payload column (batches with a payload)
    void* pay_out;      // payload in sorted order
    const void* pay_in2;  // second 8-byte payload column (16-byte payloads)
    void* pay_out2;
    uint64_t n;         // rows
    uint64_t row_off;   // first row in the concatenated scratch buffers
    uint32_t tile_off;  // first pass tile of the segment
    uint32_t pad;
};

struct HistItem {
    uint32_t seg;
    uint32_t pad;
    uint64_t begin, end;  // local rows
};

// Per-segment plan header (words):
//   kHdrP     pass slots in use: slots [first, P) scatter
//   kHdrMode  first slot (bits 0..7); bit 8: local finish (the last global
//             pass writes to scratch and k_local writes the outputs)
//   kHdrDig   digit of each slot
//   kHdrNS, kHdrS  the plain plan (non-constant digits, ascending), kept for
//             the fallback when a local finish does not fit
//   kHdrGa, kHdrGb  the two digits a local finish scatters globally
constexpr int kHdrP = 0;
constexpr int kHdrMode = 1;
constexpr int kHdrDig = 2;
constexpr int kHdrNS = kHdrDig + kMaxDigits;
constexpr int kHdrS = kHdrNS + 1;
constexpr int kHdrGa = kHdrS + 4;
constexpr int kHdrGb = kHdrGa + 1;
constexpr int kPlanHdr = 32;
static_assert(kHdrGb < kPlanHdr, "plan header fields");

// Local finish (4-byte keys, one segment, four non-constant digits): two
// global passes over the two top digits leave every bucket of equal top digits
// contiguous; k_local then sorts chunks of whole buckets of at most
// kLocalRows rows in LDS.  Below kLocalGreedyRows rows chunks are cut at a
// fixed stride (a bucket larger than kLocalRows - kLocalMinH rows makes
// k_local_chunks restart the plain passes at slot 2); from kLocalGreedyRows
// rows on they are packed greedily inside each run of 256 buckets of equal
// top digit (a bucket larger than kLocalRows restarts the plain passes).
constexpr uint32_t kLocalRows = 4096;
constexpr int kLocalBlock = 512;  // threads of a k_local block (8 waves)
constexpr uint32_t kLocalMinH = 256;  // smallest chunk stride
constexpr uint64_t kLocalMinSortRows = 1u << 16;  // smallest sort that may finish locally
constexpr uint32_t kLocalGreedyRows = 1u << 24;   // greedy packing from here on
// chunk words (start, end pairs): the stride rule makes at most n / kLocalMinH
// + 1 chunks, the greedy one at most 2 n / kLocalRows + 256
constexpr uint64_t local_chunk_words(uint64_t rows) { return 2 * (rows / kLocalMinH + 2) + 2 * kRadix; }
constexpr int kLocalMaxSlots = kLocalSlot0 + 4;
// Ticket words of a pass slot: one 64-bit claim word per XCD (kXcds), then
// the slot's chunk counter (also the tile counter of global order).  The last
// word of slot 0 is free: k_plan keeps the largest pass count of the batch
// there.
constexpr uint32_t kXcds = 8;
constexpr uint32_t kChunkWord = 2 * kXcds;
constexpr uint32_t kMaxPassWord = kTicketStride - 1;
static_assert(kTicketStride > kChunkWord + 1, "ticket words of slot 0 hold the XCD words, the chunk counter and the pass count");

struct SortArgs {
    const SegDesc* segs;
    const uint32_t* tile_seg;  // pass tile -> segment (null: one segment)
    const HistItem* items;
    uint32_t* hist;            // [S][D][256]
    uint32_t* plan_hdr;        // [S][kPlanHdr]
    uint32_t* plan_base;       // [S][D][256] segment-local bucket starts
    void* kbuf[2];             // transformed keys between passes (concatenated); with
                               // 4-byte keys and row ids, (key, row id) records
    uint32_t* vbuf[2];         // local row ids between passes (8-byte and 16-byte keys)
    void* pbuf[2];             // payload between passes (concatenated)
    void* pbuf2[2];            // second payload column between passes (16-byte payloads)
    uint32_t* lb[2];           // look-back words [tiles][256], alternating by slot parity
    uint32_t* tickets;         // [kMaxDigits][kTicketStride] tile-claim counters
    uint64_t mlo, mhi;         // key transform masks
    uint64_t lb_words;         // words of one look-back buffer
    uint32_t nsegs;
    uint32_t ntiles;
    uint32_t groups;           // kXcds: XCD-grouped claiming; 1: global tile order
    uint32_t chunk;            // consecutive tiles an XCD claims at a time
    uint32_t perm64;
    uint32_t reverse;
    uint32_t norows;           // 1: no segment needs its permutation; row ids are not moved
    uint32_t local_ok;         // k_plan may choose a local finish (one segment, 4-byte keys)
    uint32_t lds_rank;         // LDS atomics serve lanes in lane order on this device (lds_rank_check)
    uint32_t* lstart;          // local finish: [256][257] start of bucket (y, x) (x = 256: end of y)
    uint32_t* lchunk;          // local finish: chunk start rows (chunks + 1 words)
    uint32_t* lctl;            // local finish: [0] chunk count
};

// Adds one key per lane to the block histogram.  Digits flagged in `shared`
// (wave-uniform) were constant across a sample of the wave's keys; for those
// the wave first checks whether every lane has the same digit and then adds
// the count once, instead of 64 atomics on one LDS word.
// DS: words between two digits' counters in `sh` (the lane's histogram copy).
template <typename K, int DS>
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
                    atomicAdd(&sh[d * DS + b0], static_cast<uint32_t>(__popcll(vm)));
                continue;
            }
        }
        if (valid) atomicAdd(&sh[d * DS + b], 1u);
    }
}

// One work item per block: rows [begin, end) of one segment, one read of the
// keys, LDS histogram of every digit, flushed with one atomic per non-empty
// bucket into the segment's histogram.  Also clears look-back buffer 0.
// Descriptors of a one-segment sort written on the device (allocation-free
// calls, which have no pinned host staging): the segment's SegDesc and its
// histogram work items of `hist_rows` rows each.
__global__ void k_setup_one(SegDesc seg, SegDesc* segs, HistItem* items, uint64_t nitems, uint64_t hist_rows)
{
    const uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i == 0) segs[0] = seg;
    if (i < nitems) {
        const uint64_t b = i * hist_rows;
        items[i] = HistItem{0u, 0u, b, b + hist_rows < seg.n ? b + hist_rows : seg.n};
    }
}

template <typename K>
__global__ __launch_bounds__(kBlock) void k_histogram(SortArgs a)
{
    using KO = KeyOps<K>;
    constexpr int D = KO::kDigits;
    // Each quarter of a wave counts into its own copy of the histogram, offset
    // by 16 banks, so lanes with equal digits (narrow or skewed digits) spread
    // over four addresses and banks.  16-byte keys keep one copy (LDS).
    constexpr int C = D <= 8 ? 4 : 1;
    constexpr int kRow = kRadix + 16;
    constexpr int DS = C * kRow;
    __shared__ uint32_t sh[D * DS];
    for (int j = threadIdx.x; j < D * DS; j += kBlock) sh[j] = 0;
    __syncthreads();
    const HistItem it = a.items[blockIdx.x];
    const void* keys = a.segs[it.seg].keys;
    const uint32_t lane = threadIdx.x & 63;
    uint32_t* mine = sh + ((lane >> 4) % C) * kRow;
    const uint64_t wave0 = it.begin + (threadIdx.x & ~63u);  // first row of this wave
    uint32_t shared = 0;  // digits constant over the wave's first 64 keys
    {
        const uint64_t i = wave0 + lane;
        const bool v = i < it.end;
        const K k = v ? KO::load(keys, i, a.mlo, a.mhi) : KO::zero();
        const uint64_t vm = __ballot(v);
#pragma unroll
        for (int d = 0; d < D; ++d) {
            const uint32_t b = KO::digit(k, d);
            const uint32_t b0 = __builtin_amdgcn_readfirstlane(b);
            if (vm != 0 && __ballot(v && b == b0) == vm) shared |= 1u << d;
        }
    }
    // waves step through the item in strides of the block; full rounds first
    uint64_t i0 = wave0;
    constexpr int U = 8;  // loads in flight per lane
    for (; i0 + 63 + (U - 1) * kBlock < it.end; i0 += U * kBlock) {
        K k[U];
#pragma unroll
        for (int u = 0; u < U; ++u) k[u] = KO::fetch(keys, i0 + lane + u * kBlock);
#pragma unroll
        for (int u = 0; u < U; ++u) hist_add<K, DS>(mine, KO::xform(k[u], a.mlo, a.mhi), true, shared);
    }
    for (; i0 < it.end; i0 += kBlock) {
        const uint64_t i = i0 + lane;
        const bool v = i < it.end;
        hist_add<K, DS>(mine, v ? KO::load(keys, i, a.mlo, a.mhi) : KO::zero(), v, shared);
    }
    __syncthreads();
    uint32_t* h = a.hist + static_cast<uint64_t>(it.seg) * D * kRadix;
    for (int j = threadIdx.x; j < D * kRadix; j += kBlock) {
        const int d = j / kRadix, b = j % kRadix;
        uint32_t c = 0;
#pragma unroll
        for (int q = 0; q < C; ++q) c += sh[d * DS + q * kRow + b];
        if (c) atomicAdd(&h[j], c);
    }
    const uint64_t stride = static_cast<uint64_t>(gridDim.x) * kBlock;
    for (uint64_t w = static_cast<uint64_t>(blockIdx.x) * kBlock + threadIdx.x; w < a.lb_words; w += stride)
        a.lb[0][w] = 0u;
}

// One block per segment: exclusive bucket offsets of every digit and the list
// of digits that need a pass (all-in-one-bucket digits are skipped).
template <int D>
__global__ __launch_bounds__(kBlock) void k_plan(SortArgs a)
{
    __shared__ uint32_t s_wsum[kWaves];
    __shared__ uint32_t s_triv[kMaxDigits];
    __shared__ uint32_t s_max[kMaxDigits];  // largest bucket of each digit
    const uint32_t sgi = blockIdx.x;
    const int t = threadIdx.x;
    const uint64_t n = a.segs[sgi].n;
    const uint32_t* h = a.hist + static_cast<uint64_t>(sgi) * D * kRadix;
    uint32_t* base = a.plan_base + static_cast<uint64_t>(sgi) * D * kRadix;
    if (t < kMaxDigits) s_triv[t] = s_max[t] = 0;
    __syncthreads();
    for (int d = 0; d < D; ++d) {
        const uint32_t c = h[d * kRadix + t];
        if (c == n) s_triv[d] = 1;
        if (D == 4) atomicMax(&s_max[d], c);
        base[d * kRadix + t] = block_excl_scan(c, s_wsum);
    }
    __syncthreads();
    if (t == 0) {
        uint32_t* hdr = a.plan_hdr + static_cast<uint64_t>(sgi) * kPlanHdr;
        uint32_t P = 0;
        for (int d = 0; d < D; ++d)
            if (!s_triv[d]) hdr[kHdrDig + P++] = static_cast<uint32_t>(d);
        if (P == 0 && n > 0) {  // every digit constant: one pass still writes perm and keys
            hdr[kHdrDig] = 0;
            P = 1;
        }
        hdr[kHdrMode] = 0;
        // Local finish: with all four digits non-constant, scatter the top two
        // globally and sort the low two inside buckets of equal top digits in
        // LDS (one streaming pass that costs about 1.3 global passes, so it
        // pays when it replaces two).  Chosen when the buckets look small: the
        // largest bucket of two independent digits is about max_ga * max_gb / n.
        if (D == 4 && a.local_ok && P == 4) {
            const uint32_t ga = hdr[kHdrDig + P - 2], gb = hdr[kHdrDig + P - 1];
            const uint64_t est = static_cast<uint64_t>(s_max[ga]) * s_max[gb] / (n ? n : 1);
            if (est <= kLocalRows * 3 / 4) {
                hdr[kHdrNS] = P;
                for (uint32_t i = 0; i < P; ++i) hdr[kHdrS + i] = hdr[kHdrDig + i];
                hdr[kHdrGa] = ga;
                hdr[kHdrGb] = gb;
                hdr[kHdrDig] = ga;
                hdr[kHdrDig + 1] = gb;
                hdr[kHdrMode] = kModeLocal;
                P = 2;
            }
        }
        hdr[kHdrP] = P;
        atomicMax(a.tickets + kMaxPassWord, P);
    }
}


__device__ inline uint32_t block_excl_scan_lds(uint32_t x, uint32_t* s_wsum)
{
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const uint32_t inc = wave_incl_scan(x);
    if (lane == 63) s_wsum[w] = inc;
    lds_barrier();
    uint32_t pre = 0;
#pragma unroll
    for (int i = 0; i < kWaves; ++i) pre += (i < w) ? s_wsum[i] : 0u;
    lds_barrier();
    return pre + inc - x;
}

// ---------------------------------------------------------------------------
// stable ranks inside a wave
// ---------------------------------------------------------------------------
// Ballot ranking: one ballot per digit bit (BITS of them) gives each lane the lanes
// sharing its digit; every lane reads the wave's counter of its digit and the
// lowest peer writes it back advanced by the peer count, so old + (peers below
// the lane) is the lane's stable rank.  (LDS operations of one wave complete
// in program order.)
// (c: the lane's counter, the wave's counter of digit d)
template <int BITS = kRadixBits>
__device__ inline uint32_t rank_ballot_at(uint32_t d, bool v, uint32_t* c)
{
    const uint64_t vm = __ballot(v);
    uint32_t pl = static_cast<uint32_t>(vm), ph = static_cast<uint32_t>(vm >> 32);
#pragma unroll
    for (int b = 0; b < BITS; ++b) {
        // v_bitop3 truth table 0x90 = src0 & ~(src1 ^ src2): keep the lanes that agree on bit b
        const uint32_t bm = static_cast<uint32_t>(static_cast<int32_t>(d << (31 - b)) >> 31);
        const uint64_t mb = __ballot(bm != 0u);
        pl = __builtin_amdgcn_bitop3_b32(pl, static_cast<uint32_t>(mb), bm, 0x90);
        ph = __builtin_amdgcn_bitop3_b32(ph, static_cast<uint32_t>(mb >> 32), bm, 0x90);
    }
    const uint32_t lt = __builtin_amdgcn_mbcnt_hi(ph, __builtin_amdgcn_mbcnt_lo(pl, 0u));
    const uint32_t cnt = static_cast<uint32_t>(__popc(pl) + __popc(ph));
    const uint32_t old = *c;
    if (v && lt == 0u) *c = old + cnt;
    return old + lt;
}
__device__ inline uint32_t rank_ballot(uint32_t d, bool v, uint32_t* wc) { return rank_ballot_at(d, v, wc + d); }

// Atomic ranking: one LDS add-with-return per lane.  Stable when the LDS
// serves the lanes of one instruction that share an address in lane order,
// which lds_rank_check verifies on each device before it is used.
__device__ inline uint32_t rank_atomic(uint32_t d, bool v, uint32_t* wc)
{
    return v ? atomicAdd(&wc[d], 1u) : 0u;
}

// What a block knows about one claimed tile.
struct TileCtx {
    uint32_t t;         // global pass tile
    uint32_t t0;        // the segment's first tile
    uint32_t dig;       // digit of this pass for the segment
    uint32_t first, last;
    uint64_t n;         // segment rows
    uint64_t row_off;   // segment's first row in the scratch buffers
    uint64_t lrow0;     // first local row of the tile
    const void* keys;   // segment's caller keys (first pass)
    void* perm_out;
    void* keys_out;
    const void* pay_in;
    void* pay_out;
    const void* pay_in2;
    void* pay_out2;
    const uint32_t* base;  // segment's bucket starts for this digit
};

// One radix scatter pass over the tiles of every segment, persistent and
// software-pipelined (SPLIT kernels, below, take one tile at a time).
//
// Tile claiming (XCD-grouped).  The tiles are cut into chunks of `chunk`
// consecutive tiles, and each XCD (read from the hardware's XCC_ID register,
// not assumed from the block index) works through chunks of its own, so a
// chunk's bucket runs are written by CUs sharing one L2, which merges the
// partial lines of adjacent runs.  XCD x's claim word holds its current
// chunk id (low half) and R = r - chunk mod 2^32 (high half), r the claims
// taken from that chunk; a claim is one 64-bit atomicAdd of 2^32:
//  - r < chunk: the claim is tile id * chunk + r;
//  - r = chunk (exactly one claim per chunk, and the very first claim of the
//    pass, since the word starts at 0): this block takes the next chunk id
//    from the slot's chunk counter and installs it with r = 1, taking tile 0;
//  - r > chunk: the install is in progress (by a running block); claim again.
// Each XCD thus takes chunks in increasing order and claims increasing tiles,
// every tile of an XCD's earlier chunks is held by a running block, and
// every claimed tile publishes its aggregate before its block waits on a
// predecessor.  So the smallest unfinished tile is held by a running block
// whose predecessors are done, at any residency: the kernel needs no
// cooperative launch and runs beside other work.  With one group the tiles
// are claimed in global order from the chunk counter.
//
// Per tile: load (coalesced, striped rows), stable rank inside each wave,
// publish the tile's per-bucket counts (flag A; a segment's first tile
// publishes its inclusive prefix, flag P), stage the tile in LDS in bucket
// order.  The block then claims, loads and ranks its next tile and publishes
// that aggregate before it resolves the staged tile's decoupled look-back
// (Merrill & Garland 2016; onesweep, Adinets & Merrill 2022) and writes it
// out, so look-back waits overlap work.  Look-back never crosses the start of
// the tile's segment.
//
// Ranking of item j in a wave: rank_atomic (one LDS add-with-return per row)
// where the device check passed, else rank_ballot; items are ranked in
// order, so item j's rank counts exactly the items before it.
// Blocks of a pass kernel that fit one CU's 160 KiB of LDS; the kernel is
// compiled to keep that many waves per SIMD (one wave of each block per SIMD),
// so registers never cut occupancy below what LDS allows.
// SPLIT kernels stage a tile one 8-byte column at a time (LDS: 9 bytes per
// row, the ninth the row's digit), so that larger tiles fit two blocks per CU.
template <typename K, int ITEMS, int PB, bool SPLIT, bool ROWS>
struct PassShape {
    static constexpr int kRid = ROWS ? 4 : 0;  // staged row-id bytes
    static constexpr int kLds =
        (SPLIT ? kBlock * ITEMS * 9 : kBlock * ITEMS * static_cast<int>(sizeof(K) + kRid + PB)) +
        kWaves * kRadix * 4 + kRadix * 4 + 64;
    // register room (bytes of rows a thread holds): the larger tiles and
    // payloads get fewer blocks, so that no instantiation spills
    static constexpr int kRowBytes = ITEMS * static_cast<int>(sizeof(K) + kRid + PB);
    static constexpr int kCap = (SPLIT || ITEMS >= 24 || PB == 16 || kRowBytes > 160 || (PB != 0 && ITEMS >= 12)) ? 2 : 3;
    static constexpr int kBlocksPerCu = (160 * 1024 / kLds) < kCap ? (160 * 1024 / kLds) : kCap;
};

// PB: payload bytes per row moved with the keys: 0, 4, 8, 12 (an 8-byte and a
// 4-byte column) or 16 (two 8-byte columns).  SPLIT: the staged tile is written out in phases, one 8-byte
// column at a time staged at the rows' slots -- (key, row id) records for
// 4-byte keys; the key (8-byte keys) or its two words (16-byte keys) and the
// row ids otherwise; then each payload column -- and tiles are processed one
// at a time (the registers of a column stay live until its phase).
// ROWS = false: a kernel for sorts that move no row ids, whose tiles stage no
// row-id column (4-byte keys: no (key, row id) records either), so larger
// tiles fit the LDS.
template <typename K, int ITEMS, int PB, bool SPLIT = false, bool ROWS = true>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(PassShape<K, ITEMS, PB, SPLIT, ROWS>::kBlocksPerCu)))
void k_pass(SortArgs a, uint32_t slot)
{
    using KO = KeyOps<K>;
    using P = typename std::conditional<PB >= 8, uint64_t, uint32_t>::type;
    using P2 = typename std::conditional<PB == 12, uint32_t, uint64_t>::type;
    constexpr bool TWO = PB == 12 || PB == 16;  // a second payload column (4 or 8 bytes)
    constexpr int PI = PB ? ITEMS : 1;    // payload registers
    constexpr int PI2 = TWO ? ITEMS : 1;
    constexpr int TILE = kBlock * ITEMS;
    constexpr int D = KO::kDigits;
    // 4-byte keys travel between passes packed with their row ids: one 8-byte
    // record (key low, row id high) per row, so a bucket run is written as one
    // stream of full 8-byte words instead of two half-size streams
    constexpr bool KV = std::is_same<K, uint32_t>::value && ROWS;
    __shared__ K s_keys[(KV || SPLIT) ? 1 : TILE];
    __shared__ uint32_t s_vals[(KV || SPLIT || !ROWS) ? 1 : TILE];
    __shared__ uint64_t s_kv[(KV || SPLIT) ? TILE : 1];  // SPLIT: every phase's column
    __shared__ P s_pay[(PB && !SPLIT) ? TILE : 1];
    __shared__ P2 s_pay2[(TWO && !SPLIT) ? TILE : 1];
    __shared__ uint8_t s_dig[SPLIT ? TILE : 1];  // SPLIT: digit of each staged slot
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
    P pay[PI];
    P2 pay2[PI2];
    uint32_t rnk[ITEMS];

    auto claim = [&]() -> uint32_t {
        if (tid == 0) {
            uint32_t* tk = a.tickets + slot * kTicketStride;
            uint64_t t;
            if (a.groups == 1) {
                t = atomicAdd(tk + kChunkWord, 1u);
            } else {
                // XCC_ID (hardware register 20), bits 3:0
                const uint32_t x = static_cast<uint32_t>(__builtin_amdgcn_s_getreg((3 << 11) | 20)) % kXcds;
                auto* w = reinterpret_cast<unsigned long long*>(tk) + x;
                for (;;) {
                    const unsigned long long v = atomicAdd(w, 1ull << 32);
                    const uint32_t r = static_cast<uint32_t>(v >> 32) + a.chunk;
                    if (r < a.chunk) {
                        t = static_cast<uint64_t>(static_cast<uint32_t>(v)) * a.chunk + r;
                        break;
                    }
                    if (r == a.chunk) {
                        const uint32_t c = atomicAdd(tk + kChunkWord, 1u);
                        atomicExch(w, (static_cast<unsigned long long>(1u - a.chunk) << 32) | c);
                        t = static_cast<uint64_t>(c) * a.chunk;
                        break;
                    }
                    __builtin_amdgcn_s_sleep(1);
                }
            }
            s_tile = static_cast<uint32_t>(std::min<uint64_t>(t, 0xFFFFFFFFull));
        }
        lds_barrier();
        const uint32_t t = __builtin_amdgcn_readfirstlane(s_tile);
        lds_barrier();
        return t;
    };

    // claims until a tile whose segment needs this pass; false at the end
    auto next_tile = [&](TileCtx& c) -> bool {
        for (;;) {
            const uint32_t t = claim();
            if (t >= a.ntiles) return false;
            const uint32_t sg = a.tile_seg ? __builtin_amdgcn_readfirstlane(a.tile_seg[t]) : 0u;
            const uint32_t* hdr = a.plan_hdr + static_cast<uint64_t>(sg) * kPlanHdr;
            const uint32_t P = hdr[kHdrP];
            const uint32_t mode = hdr[kHdrMode];
            const uint32_t s0 = mode & 0xFFu;
            if (slot >= P || slot < s0) continue;
            const SegDesc& sd = a.segs[sg];
            c.t = t;
            c.t0 = sd.tile_off;
            c.dig = hdr[kHdrDig + slot];
            c.first = (slot == s0);
            c.last = (slot == P - 1) && !(mode & kModeLocal);
            c.n = sd.n;
            c.row_off = sd.row_off;
            c.lrow0 = static_cast<uint64_t>(t - sd.tile_off) * TILE;
            c.keys = sd.keys;
            c.perm_out = sd.perm_out;
            c.keys_out = sd.keys_out;
            c.pay_in = sd.pay_in;
            c.pay_out = sd.pay_out;
            c.pay_in2 = sd.pay_in2;
            c.pay_out2 = sd.pay_out2;
            c.base = a.plan_base + (static_cast<uint64_t>(sg) * D + c.dig) * kRadix;
            return true;
        }
    };

    // load tile c, rank it, publish its aggregate; returns the bucket total and
    // tile-local bucket start of bucket `tid`; s_wc then holds each wave's
    // first LDS slot per bucket
    auto load_rank = [&](const TileCtx& c, uint32_t& tot, uint32_t& lstart) {
#pragma unroll
        for (int w = 0; w < kWaves; ++w) {
            s_wc[w][tid] = 0;
        }
        lbn[static_cast<uint64_t>(c.t) * kRadix + tid] = 0u;  // the next pass's row of this tile
        const uint64_t wl = c.lrow0 + static_cast<uint64_t>(wave) * 64 * ITEMS + lane;  // local row of item 0
        // Every load is unconditional, so all are in flight at once: a full
        // tile reads at constant offsets; in a segment's last, partial tile
        // rows past the end read the segment's last row and are masked below.
        const bool full = c.lrow0 + TILE <= c.n;
        const uint64_t last = c.n - 1;
        // payload loads, issued after the key loads: vmcnt counts in issue
        // order, so the rank phase waits for the keys alone and the payload
        // stays in flight through the rank and the look-back
        auto load_pay = [&]() __attribute__((always_inline)) {
            if constexpr (PB != 0) {
                const auto* pp = c.first ? gmem(static_cast<const P*>(c.pay_in))
                                         : gmem(static_cast<const P*>(a.pbuf[(slot - 1) & 1]) + c.row_off);
                if (full) {
#pragma unroll
                    for (int j = 0; j < PI; ++j) pay[j] = pp[wl + static_cast<uint64_t>(j) * 64];
                } else {
#pragma unroll
                    for (int j = 0; j < PI; ++j) pay[j] = pp[std::min(wl + static_cast<uint64_t>(j) * 64, last)];
                }
            }
            if constexpr (TWO) {
                const auto* pp = c.first ? gmem(static_cast<const P2*>(c.pay_in2))
                                         : gmem(static_cast<const P2*>(a.pbuf2[(slot - 1) & 1]) + c.row_off);
                if (full) {
#pragma unroll
                    for (int j = 0; j < PI2; ++j) pay2[j] = pp[wl + static_cast<uint64_t>(j) * 64];
                } else {
#pragma unroll
                    for (int j = 0; j < PI2; ++j) pay2[j] = pp[std::min(wl + static_cast<uint64_t>(j) * 64, last)];
                }
            }
        };
        if (c.first) {
            if (full) {
#pragma unroll
                for (int j = 0; j < ITEMS; ++j) key[j] = KO::fetch(c.keys, wl + static_cast<uint64_t>(j) * 64);
            } else {
#pragma unroll
                for (int j = 0; j < ITEMS; ++j)
                    key[j] = KO::fetch(c.keys, std::min(wl + static_cast<uint64_t>(j) * 64, last));
            }
            load_pay();
#pragma unroll
            for (int j = 0; j < ITEMS; ++j) {
                key[j] = KO::xform(key[j], a.mlo, a.mhi);
                val[j] = static_cast<uint32_t>(wl + static_cast<uint64_t>(j) * 64);
            }
        } else if (KV && !a.norows) {
            if constexpr (KV) {
                const auto* rp = gmem(static_cast<const uint64_t*>(a.kbuf[(slot - 1) & 1]) + c.row_off);
                uint64_t r[ITEMS];
                if (full) {
#pragma unroll
                    for (int j = 0; j < ITEMS; ++j) r[j] = rp[wl + static_cast<uint64_t>(j) * 64];
                } else {
#pragma unroll
                    for (int j = 0; j < ITEMS; ++j) r[j] = rp[std::min(wl + static_cast<uint64_t>(j) * 64, last)];
                }
                load_pay();
#pragma unroll
                for (int j = 0; j < ITEMS; ++j) {
                    key[j] = static_cast<K>(r[j]);
                    val[j] = static_cast<uint32_t>(r[j] >> 32);
                }
            }
        } else {
            const K* kp = kin + c.row_off;
            const auto* vp = gmem(vin + c.row_off);
            if (full) {
#pragma unroll
                for (int j = 0; j < ITEMS; ++j) key[j] = KO::raw(kp, wl + static_cast<uint64_t>(j) * 64);
                if (ROWS && !a.norows) {
#pragma unroll
                    for (int j = 0; j < ITEMS; ++j) val[j] = vp[wl + static_cast<uint64_t>(j) * 64];
                }
            } else {
#pragma unroll
                for (int j = 0; j < ITEMS; ++j) key[j] = KO::raw(kp, std::min(wl + static_cast<uint64_t>(j) * 64, last));
                if (ROWS && !a.norows) {
#pragma unroll
                    for (int j = 0; j < ITEMS; ++j) val[j] = vp[std::min(wl + static_cast<uint64_t>(j) * 64, last)];
                }
            }
            load_pay();
            if (!ROWS || a.norows) {
#pragma unroll
                for (int j = 0; j < ITEMS; ++j) val[j] = 0u;
            }
        }
        lds_barrier();
        uint32_t* wc = s_wc[wave];
        const int dig = static_cast<int>(c.dig);
        if (a.lds_rank) {
#pragma unroll
            for (int j = 0; j < ITEMS; ++j)
                rnk[j] = rank_atomic(KO::digit(key[j], dig), (wl + static_cast<uint64_t>(j) * 64) < c.n, wc);
        } else {
#pragma unroll
            for (int j = 0; j < ITEMS; ++j)
                rnk[j] = rank_ballot(KO::digit(key[j], dig), (wl + static_cast<uint64_t>(j) * 64) < c.n, wc);
        }
        lds_barrier();
        tot = 0;
#pragma unroll
        for (int w = 0; w < kWaves; ++w) {
            const uint32_t cw = s_wc[w][tid];
            s_wc[w][tid] = tot;
            tot += cw;
        }
        store_relaxed(lb + static_cast<uint64_t>(c.t) * kRadix + tid, (c.t == c.t0 ? kFlagP : kFlagA) | tot);
        lstart = block_excl_scan_lds(tot, s_wsum);
#pragma unroll
        for (int w = 0; w < kWaves; ++w) s_wc[w][tid] += lstart;
        lds_barrier();
    };

    auto stage = [&](const TileCtx& c) {
        const uint64_t wl = c.lrow0 + static_cast<uint64_t>(wave) * 64 * ITEMS + lane;
        const int dig = static_cast<int>(c.dig);
        // every item's slot is read first (one LDS wait for all), then the
        // valid items are written
#pragma unroll
        for (int j = 0; j < ITEMS; ++j) rnk[j] += s_wc[wave][KO::digit(key[j], dig)];
        if constexpr (SPLIT) {
            // the columns are staged by phases at write-out; here each slot's digit
#pragma unroll
            for (int j = 0; j < ITEMS; ++j)
                if (wl + static_cast<uint64_t>(j) * 64 < c.n) s_dig[rnk[j]] = static_cast<uint8_t>(KO::digit(key[j], dig));
            lds_barrier();
            return;
        }
#pragma unroll
        for (int j = 0; j < ITEMS; ++j) {
            if (wl + static_cast<uint64_t>(j) * 64 < c.n) {
                const uint32_t p = rnk[j];
                if constexpr (KV) {
                    s_kv[p] = key[j] | (static_cast<uint64_t>(val[j]) << 32);
                } else {
                    s_keys[p] = key[j];
                    if constexpr (ROWS) s_vals[p] = val[j];
                }
                if constexpr (PB != 0 && !SPLIT) s_pay[p] = pay[j];
                if constexpr (TWO && !SPLIT) s_pay2[p] = pay2[j];
            }
        }
        lds_barrier();
    };

    // decoupled look-back for the staged tile (8 predecessor words per step,
    // consumed in tile order until an inclusive prefix), then write it out
    auto resolve_write = [&](const TileCtx& c, uint32_t tot, uint32_t lstart) {
        uint32_t excl = 0;
        if (c.t > c.t0) {
            uint32_t k = c.t - 1;
            for (;;) {
                uint32_t w[8];
#pragma unroll
                for (int q = 0; q < 8; ++q)
                    w[q] = (k >= c.t0 + static_cast<uint32_t>(q))
                               ? load_relaxed(lb + static_cast<uint64_t>(k - q) * kRadix + tid)
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
                k -= consumed;  // the segment's first tile holds P, so k stays >= t0
                if (stall) __builtin_amdgcn_s_sleep(1);
            }
            store_relaxed(lb + static_cast<uint64_t>(c.t) * kRadix + tid, kFlagP | (excl + tot));
        }
        s_gbase[tid] = c.base[tid] + excl - lstart;
        lds_barrier();
        const uint32_t tn = static_cast<uint32_t>(std::min<uint64_t>(TILE, c.n - c.lrow0));
        const int dig = static_cast<int>(c.dig);
        // Rows go out U per thread at a time: the LDS reads of U rows are
        // issued before their stores, so one LDS wait covers U rows (8 where
        // the tile allows it, except in split phases, where 8 measured slower).
        // Tiles are a multiple of U * kBlock rows, so every LDS index is inside
        // the staging arrays; rows past a partial tile's end are not stored.
        constexpr int U = (!SPLIT && ITEMS % 8 == 0) ? 8 : 4;
        static_assert(ITEMS % U == 0, "tile rows per thread must be a multiple of U");
        // staged row i: key and row id
        auto staged = [&](uint32_t i, K& kk, uint32_t& vv) {
            if constexpr (KV) {
                const uint64_t r = s_kv[i];
                kk = static_cast<K>(r);
                vv = static_cast<uint32_t>(r >> 32);
            } else {
                kk = s_keys[i];
                vv = ROWS ? s_vals[i] : 0u;
            }
        };
        if constexpr (SPLIT) {
            // Phases: one 8-byte column of the tile at a time is staged at its
            // rows' slots and each slot goes to the position its digit gives.
            const uint32_t nm1 = static_cast<uint32_t>(c.n - 1);
            const uint64_t wl = c.lrow0 + static_cast<uint64_t>(wave) * 64 * ITEMS + lane;
            auto phase = [&](auto get, auto put) {
                lds_barrier();  // the previous phase's LDS reads are done
#pragma unroll
                for (int j = 0; j < ITEMS; ++j)
                    if (wl + static_cast<uint64_t>(j) * 64 < c.n) s_kv[rnk[j]] = get(j);
                lds_barrier();
                for (uint32_t i0 = tid; i0 < tn; i0 += U * kBlock) {
                    uint64_t v[U];
                    uint32_t pos[U];
#pragma unroll
                    for (int u = 0; u < U; ++u) {
                        v[u] = s_kv[i0 + u * kBlock];
                        const uint32_t p = s_gbase[s_dig[i0 + u * kBlock]] + i0 + u * kBlock;
                        pos[u] = (c.last && a.reverse) ? nm1 - p : p;
                    }
#pragma unroll
                    for (int u = 0; u < U; ++u)
                        if (i0 + u * kBlock < tn) put(pos[u], v[u]);
                }
            };
            auto row_ids = [&]() {  // the row ids (permutation in the last pass)
                if constexpr (!ROWS) return;
                if (a.norows) return;
                auto* vout = gmem(a.vbuf[slot & 1] + c.row_off);
                phase([&](int j) { return static_cast<uint64_t>(val[j]); },
                      [&](uint32_t p, uint64_t v) {
                          if (!c.last) vout[p] = static_cast<uint32_t>(v);
                          else if (c.perm_out) {
                              if (a.perm64) gmem(static_cast<uint64_t*>(c.perm_out))[p] = v;
                              else gmem(static_cast<uint32_t*>(c.perm_out))[p] = static_cast<uint32_t>(v);
                          }
                      });
            };
            if constexpr (KV) {
                // (key, row id) records
                K* kout = static_cast<K*>(a.kbuf[slot & 1]) + c.row_off;
                auto* rout = gmem(static_cast<uint64_t*>(a.kbuf[slot & 1]) + c.row_off);
                phase([&](int j) { return key[j] | (static_cast<uint64_t>(val[j]) << 32); },
                      [&](uint32_t p, uint64_t r) {
                          const uint32_t k = static_cast<uint32_t>(r), v = static_cast<uint32_t>(r >> 32);
                          if (!c.last) {
                              if (!a.norows) rout[p] = r;
                              else KO::put(kout, p, k);
                          } else {
                              if (c.keys_out) KO::store(c.keys_out, p, k, a.mlo, a.mhi);
                              if (c.perm_out) {
                                  if (a.perm64) gmem(static_cast<uint64_t*>(c.perm_out))[p] = v;
                                  else gmem(static_cast<uint32_t*>(c.perm_out))[p] = v;
                              }
                          }
                      });
            } else if constexpr (sizeof(K) <= 8) {
                // 8-byte keys, or 4-byte keys without row ids: the keys, then the row ids
                K* kout = static_cast<K*>(a.kbuf[slot & 1]) + c.row_off;
                phase([&](int j) { return static_cast<uint64_t>(key[j]); },
                      [&](uint32_t p, uint64_t k) {
                          if (!c.last) KO::put(kout, p, static_cast<K>(k));
                          else if (c.keys_out) KO::store(c.keys_out, p, static_cast<K>(k), a.mlo, a.mhi);
                      });
                row_ids();
            } else {
                // 16-byte keys, a word per phase: scratch holds (lo, hi); the
                // caller's form is bytes 0..7 = bswap(hi), 8..15 = bswap(lo)
                auto* kw = gmem(static_cast<uint64_t*>(a.kbuf[slot & 1])) + 2 * c.row_off;
                auto* ow = gmem(static_cast<uint64_t*>(c.keys_out));
                phase([&](int j) { return key[j].lo; },
                      [&](uint32_t p, uint64_t x) {
                          if (!c.last) kw[2 * static_cast<uint64_t>(p)] = x;
                          else if (ow) ow[2 * static_cast<uint64_t>(p) + 1] = __builtin_bswap64(x ^ a.mlo);
                      });
                phase([&](int j) { return key[j].hi; },
                      [&](uint32_t p, uint64_t x) {
                          if (!c.last) kw[2 * static_cast<uint64_t>(p) + 1] = x;
                          else if (ow) ow[2 * static_cast<uint64_t>(p)] = __builtin_bswap64(x ^ a.mhi);
                      });
                row_ids();
            }
            if constexpr (PB != 0) {
                auto* mid = gmem(static_cast<P*>(a.pbuf[slot & 1]) + c.row_off);
                auto* fin = gmem(static_cast<P*>(c.pay_out));
                phase([&](int j) { return static_cast<uint64_t>(pay[j]); },
                      [&](uint32_t p, uint64_t v) { (c.last ? fin : mid)[p] = static_cast<P>(v); });
            }
            if constexpr (TWO) {
                auto* mid = gmem(static_cast<P2*>(a.pbuf2[slot & 1]) + c.row_off);
                auto* fin = gmem(static_cast<P2*>(c.pay_out2));
                phase([&](int j) { return static_cast<uint64_t>(pay2[j]); },
                      [&](uint32_t p, uint64_t v) { (c.last ? fin : mid)[p] = static_cast<P2>(v); });
            }
            lds_barrier();
            return;
        }
        if (!c.last) {
            K* kout = static_cast<K*>(a.kbuf[slot & 1]) + c.row_off;
            auto* rout = gmem(static_cast<uint64_t*>(a.kbuf[slot & 1]) + c.row_off);
            auto* vout = gmem(a.vbuf[slot & 1] + c.row_off);
            auto* pout = gmem(static_cast<P*>(a.pbuf[slot & 1]) + c.row_off);
            auto* pout2 = gmem(static_cast<P2*>(a.pbuf2[slot & 1]) + c.row_off);
            for (uint32_t i0 = tid; i0 < tn; i0 += U * kBlock) {
                K kk[U];
                uint32_t pos[U], vv[U];
                P pv[U];
                P2 pv2[U];
#pragma unroll
                for (int u = 0; u < U; ++u) staged(i0 + u * kBlock, kk[u], vv[u]);
#pragma unroll
                for (int u = 0; u < U; ++u) {
                    pos[u] = s_gbase[KO::digit(kk[u], dig)] + i0 + u * kBlock;
                    if constexpr (PB != 0) pv[u] = s_pay[i0 + u * kBlock];
                    if constexpr (TWO) pv2[u] = s_pay2[i0 + u * kBlock];
                }
#pragma unroll
                for (int u = 0; u < U; ++u) {
                    if (i0 + u * kBlock < tn) {
                        if (KV && !a.norows) {
                            if constexpr (KV) rout[pos[u]] = kk[u] | (static_cast<uint64_t>(vv[u]) << 32);
                        } else {
                            KO::put(kout, pos[u], kk[u]);
                            if (!a.norows) vout[pos[u]] = vv[u];
                        }
                        if constexpr (PB != 0) pout[pos[u]] = pv[u];
                        if constexpr (TWO) pout2[pos[u]] = pv2[u];
                    }
                }
            }
        } else {
            const uint32_t nm1 = static_cast<uint32_t>(c.n - 1);
            for (uint32_t i0 = tid; i0 < tn; i0 += U * kBlock) {
                K kk[U];
                uint32_t pos[U], vv[U];
                P pv[U];
                P2 pv2[U];
#pragma unroll
                for (int u = 0; u < U; ++u) staged(i0 + u * kBlock, kk[u], vv[u]);
#pragma unroll
                for (int u = 0; u < U; ++u) {
                    const uint32_t p = s_gbase[KO::digit(kk[u], dig)] + i0 + u * kBlock;
                    pos[u] = a.reverse ? nm1 - p : p;
                    if constexpr (PB != 0) pv[u] = s_pay[i0 + u * kBlock];
                    if constexpr (TWO) pv2[u] = s_pay2[i0 + u * kBlock];
                }
#pragma unroll
                for (int u = 0; u < U; ++u) {
                    if (i0 + u * kBlock < tn) {
                        const uint32_t p = pos[u];
                        if (c.keys_out) KO::store(c.keys_out, p, kk[u], a.mlo, a.mhi);
                        if (c.perm_out) {
                            if (a.perm64)
                                gmem(static_cast<uint64_t*>(c.perm_out))[p] = vv[u];
                            else
                                gmem(static_cast<uint32_t*>(c.perm_out))[p] = vv[u];
                        }
                        if constexpr (PB != 0) gmem(static_cast<P*>(c.pay_out))[p] = pv[u];
                        if constexpr (TWO) gmem(static_cast<P2*>(c.pay_out2))[p] = pv2[u];
                    }
                }
            }
        }
        lds_barrier();
    };

    // a slot beyond every segment's pass count has no tile to move
    if (slot >= load_relaxed(a.tickets + kMaxPassWord)) return;
    if constexpr (SPLIT) {
        TileCtx c;
        while (next_tile(c)) {
            uint32_t tot, ls;
            load_rank(c, tot, ls);
            stage(c);
            resolve_write(c, tot, ls);
        }
        return;
    }
    TileCtx c0, c1;
    if (!next_tile(c0)) return;
    uint32_t tot0, ls0;
    load_rank(c0, tot0, ls0);
    stage(c0);
    for (;;) {
        const bool has1 = next_tile(c1);
        uint32_t tot1 = 0, ls1 = 0;
        if (has1) load_rank(c1, tot1, ls1);
        resolve_write(c0, tot0, ls0);
        if (!has1) break;
        stage(c1);
        c0 = c1;
        tot0 = tot1;
        ls0 = ls1;
    }
}

// Compares the two rankings on digit patterns with 1 .. 256 distinct values
// per wave (block b: 2^(b mod 9)); counts the lanes whose ranks differ.
__global__ __launch_bounds__(kBlock) void k_lds_rank_check(uint32_t* bad)
{
    __shared__ uint32_t s_a[kWaves][kRadix], s_b[kWaves][kRadix];
    const int tid = threadIdx.x, wave = tid >> 6;
#pragma unroll
    for (int w = 0; w < kWaves; ++w) s_a[w][tid] = s_b[w][tid] = 0;
    __syncthreads();
    const uint32_t range = 1u << (blockIdx.x % 9);
    uint32_t miss = 0;
    for (uint32_t j = 0; j < 32; ++j) {
        uint32_t x = (blockIdx.x * 32u + j) * kBlock + static_cast<uint32_t>(tid);
        x = (x ^ 61u) ^ (x >> 16);  // integer hash (Wang)
        x *= 9u;
        x ^= x >> 4;
        x *= 0x27d4eb2du;
        x ^= x >> 15;
        const uint32_t d = x & (range - 1u);
        const bool v = (x >> 28) != 0;  // some lanes inactive
        const uint32_t rb = rank_ballot(d, v, s_b[wave]);
        const uint32_t ra = rank_atomic(d, v, s_a[wave]);
        miss += (v && ra != rb) ? 1u : 0u;
    }
    if (miss) atomicAdd(bad, miss);
}

static std::mutex g_lds_rank_mu;
static std::vector<std::pair<int, bool>> g_lds_rank;  // device -> atomic ranking verified

bool lds_rank_check(int device)
{
    std::lock_guard<std::mutex> lk(g_lds_rank_mu);
    for (const auto& e : g_lds_rank)
        if (e.first == device) return e.second;
    uint32_t* d = nullptr;
    VDF_HIP_CHECK(hipMalloc(&d, sizeof(uint32_t)));
    uint32_t bad = 1;
    const hipError_t e1 = hipMemset(d, 0, sizeof(uint32_t));
    if (e1 == hipSuccess) hipLaunchKernelGGL(k_lds_rank_check, dim3(64), dim3(kBlock), 0, nullptr, d);
    const hipError_t e2 = e1 == hipSuccess ? hipGetLastError() : e1;
    const hipError_t e3 = e2 == hipSuccess ? hipMemcpy(&bad, d, sizeof(uint32_t), hipMemcpyDeviceToHost) : e2;
    (void)hipFree(d);
    VDF_HIP_CHECK(e3);
    g_lds_rank.emplace_back(device, bad == 0);
    return bad == 0;
}

// the check's result when it ran on `device`, else false (the ballot ranking)
bool lds_rank_verified(int device)
{
    std::lock_guard<std::mutex> lk(g_lds_rank_mu);
    for (const auto& e : g_lds_rank)
        if (e.first == device) return e.second;
    return false;
}

// ---------------------------------------------------------------------------
// local finish (4-byte keys, one segment)
// ---------------------------------------------------------------------------
// Bucket starts.  After the global passes over digits ga (slot 0) and gb
// (slot 1) the rows are ordered by (gb, ga); bucket (y, x), the rows with
// digit gb = y and digit ga = x, starts at base_gb[y] + P_y(B_x), where B_x is
// the first row of ga-bucket x in slot 1's input (slot 0's output, ordered by
// ga) and P_y(r) counts that input's rows before r whose digit gb is y.  At a
// tile boundary P_y is the inclusive look-back prefix slot 1 left for the
// tile before; the rows of B_x's own tile before it are counted here.  One
// block per boundary x = 0 .. 256 (B_256 = n); lstart[y][x] is the start.
__global__ __launch_bounds__(kBlock) void k_local_bounds(SortArgs a, uint32_t tile_rows)
{
    __shared__ uint32_t s_cnt[kRadix];
    const uint32_t* hdr = a.plan_hdr;
    if (!(hdr[kHdrMode] & kModeLocal)) return;
    const int tid = threadIdx.x;
    const uint32_t x = blockIdx.x;
    if (x == 0 && tid == 0) a.lctl[0] = a.lctl[1] = 0;  // chunk count, k_local_max's maximum
    const uint32_t ga = hdr[kHdrGa], gb = hdr[kHdrGb];
    const uint64_t n = a.segs[0].n;
    const uint32_t* bgb = a.plan_base + static_cast<uint64_t>(gb) * kRadix;
    const uint64_t B = x < kRadix ? a.plan_base[static_cast<uint64_t>(ga) * kRadix + x] : n;
    uint32_t start;
    if (B >= n) {
        start = bgb[tid] + a.hist[static_cast<uint64_t>(gb) * kRadix + tid];
    } else {
        s_cnt[tid] = 0;
        __syncthreads();
        const uint64_t t = B / tile_rows;
        for (uint64_t i = t * tile_rows + tid; i < B; i += kBlock) {
            const uint32_t k = a.norows ? gmem(static_cast<const uint32_t*>(a.kbuf[0]))[i]
                                        : static_cast<uint32_t>(gmem(static_cast<const uint64_t*>(a.kbuf[0]))[i]);
            atomicAdd(&s_cnt[(k >> (8 * gb)) & 0xFFu], 1u);
        }
        __syncthreads();
        const uint32_t prev = t > 0 ? (gmem(a.lb[1])[(t - 1) * kRadix + tid] & kValMask) : 0u;
        start = bgb[tid] + prev + s_cnt[tid];
    }
    a.lstart[static_cast<uint64_t>(tid) * (kRadix + 1) + x] = start;
}

// The largest bucket of a local finish: block y takes buckets (y, 0..255),
// their sizes lstart[y][x + 1] - lstart[y][x], into lctl[1].
__global__ __launch_bounds__(kBlock) void k_local_max(SortArgs a)
{
    __shared__ uint32_t s_m[kWaves];
    if (!(a.plan_hdr[kHdrMode] & kModeLocal)) return;
    const int tid = threadIdx.x;
    const uint32_t* row = a.lstart + static_cast<uint64_t>(blockIdx.x) * (kRadix + 1);
    uint32_t m = row[tid + 1] - row[tid];
    for (int off = 32; off > 0; off >>= 1) m = std::max(m, static_cast<uint32_t>(__shfl_xor(m, off, 64)));
    if ((tid & 63) == 0) s_m[tid >> 6] = m;
    __syncthreads();
    if (tid == 0) atomicMax(a.lctl + 1, std::max(std::max(s_m[0], s_m[1]), std::max(s_m[2], s_m[3])));
}

// Cuts the rows into chunks of whole buckets of at most kLocalRows rows, as
// (start, end) pairs in lchunk with their count in lctl[0], or falls back.
// Thread x of block y takes bucket b = y * 256 + x, which starts at
// lstart[y][x]; lstart[y][256] is the end of the run of buckets y.
//  - Stride rule (fewer than kLocalGreedyRows rows): with M the largest
//    bucket and H = kLocalRows - M, chunk k holds the buckets that start in
//    [kH, (k+1)H): it starts at the first bucket start >= kH and ends before
//    the first one >= (k+1)H, so it has fewer than H + M = kLocalRows rows.
//    M > kLocalRows - kLocalMinH falls back.
//  - Greedy rule: wave 0 of block y packs run y, each chunk taking as many
//    whole buckets as fit in kLocalRows rows (a chunk and the next hold more
//    than kLocalRows rows together, so a run of R rows gives at most
//    2R / kLocalRows + 1 chunks).  Each lane finds, for its buckets, where a
//    chunk beginning there would end (binary search over the run's starts);
//    the wave then follows that chain from bucket 0 in registers, reserves
//    the run's slots with one atomic add and writes the chunks (slot order
//    does not matter: every chunk writes its own rows).  M > kLocalRows
//    falls back.
// The fallback restarts the plain passes at slot kLocalSlot0 (they read the
// caller's columns again).
__global__ __launch_bounds__(kBlock) void k_local_chunks(SortArgs a)
{
    uint32_t* hdr = a.plan_hdr;
    if (!(hdr[kHdrMode] & kModeLocal)) return;
    const uint32_t m = a.lctl[1];
    const uint32_t n = static_cast<uint32_t>(a.segs[0].n);
    const uint32_t tid = threadIdx.x, y = blockIdx.x;
    const bool greedy = n >= kLocalGreedyRows;
    if (m > (greedy ? kLocalRows : kLocalRows - kLocalMinH)) {
        // (blocks that read the mode after this rewrite return as well)
        if (y == 0 && tid == 0) {
            const uint32_t np = hdr[kHdrNS];
            for (uint32_t i = 0; i < np; ++i) hdr[kHdrDig + kLocalSlot0 + i] = hdr[kHdrS + i];
            hdr[kHdrMode] = kLocalSlot0;
            hdr[kHdrP] = kLocalSlot0 + np;
            atomicMax(a.tickets + kMaxPassWord, kLocalSlot0 + np);
            a.lctl[0] = 0;
        }
        return;
    }
    const uint32_t* row = a.lstart + static_cast<uint64_t>(y) * (kRadix + 1);
    uint32_t* ch = a.lchunk;
    if (greedy) {
        __shared__ uint32_t s_st[kRadix + 1];
        s_st[tid] = row[tid];  // (blockDim.x == kRadix)
        if (tid == 0) s_st[kRadix] = row[kRadix];
        __syncthreads();
        if (tid >= 64) return;
        const uint32_t lane = tid;
        // bucket x = 64 i + lane: its start, and nx = the last bucket index
        // (256 = the run's end) whose start lies within start + kLocalRows,
        // where a chunk that begins at x ends (a bucket has at most kLocalRows
        // rows, so nx > x; the max() only keeps the walk finite regardless)
        uint32_t st[4], nx[4];
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const uint32_t x = static_cast<uint32_t>(i) * 64 + lane;
            st[i] = s_st[x];
            const uint32_t lim = st[i] + kLocalRows;
            uint32_t lo = x, hi = kRadix;  // s_st[lo] <= lim; find the largest such index
            while (lo < hi) {
                const uint32_t mid = (lo + hi + 1) >> 1;
                if (s_st[mid] <= lim) lo = mid;
                else hi = mid - 1;
            }
            nx[i] = std::max(lo, x + 1);
        }
        // the chain of chunk starts from bucket 0, marked in the lanes that hold them
        uint32_t mark = 0;
        if (s_st[0] < s_st[kRadix]) {
            uint32_t x = 0;
            while (x < kRadix) {
                const uint32_t q = x >> 6, l = x & 63;
                if (lane == l) mark |= 1u << q;
                const uint32_t v = q == 0 ? nx[0] : q == 1 ? nx[1] : q == 2 ? nx[2] : nx[3];
                x = __builtin_amdgcn_readlane(v, l);
            }
        }
        // slots: one atomic add per run, chunks in bucket order inside it
        uint32_t below[4], cnt = 0;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const uint64_t b = __ballot((mark >> i) & 1u);
            below[i] = cnt + __builtin_amdgcn_mbcnt_hi(static_cast<uint32_t>(b >> 32),
                                                       __builtin_amdgcn_mbcnt_lo(static_cast<uint32_t>(b), 0u));
            cnt += static_cast<uint32_t>(__popcll(b));
        }
        if (cnt == 0) return;
        uint32_t slot = 0;
        if (lane == 0) slot = atomicAdd(a.lctl, cnt);
        slot = __builtin_amdgcn_readfirstlane(slot);
#pragma unroll
        for (int i = 0; i < 4; ++i)
            if ((mark >> i) & 1u) {
                const uint64_t k = slot + below[i];
                ch[2 * k] = st[i];
                ch[2 * k + 1] = s_st[nx[i]];
            }
        return;
    }
    const uint32_t H = kLocalRows - m;
    const uint32_t K = (n + H - 1) / H;
    // chunks k with start(b - 1) < kH <= start(b) start at start(b) (each k
    // once: an empty bucket's start equals the next one's); chunk k - 1 ends there
    auto cut = [&](uint32_t s, uint32_t p) {
        for (uint32_t k = p / H + 1; k <= s / H && k < K; ++k) {
            ch[2 * static_cast<uint64_t>(k)] = s;
            ch[2 * static_cast<uint64_t>(k) - 1] = s;
        }
    };
    const uint32_t s = row[tid];
    // the bucket before (y, 0) is (y - 1, 255): row[-2] (row[-1] is its end)
    if (y > 0 || tid > 0) cut(s, tid > 0 ? row[tid - 1] : row[-2]);
    if (y == kRadix - 1 && tid == kRadix - 1) cut(n, row[tid]);
    if (y == 0 && tid == 0) {
        ch[0] = 0;
        ch[2 * static_cast<uint64_t>(K) - 1] = n;
        a.lctl[0] = K;
    }
}

// Untyped buffer stores (MUBUF): word 3 of the resource for gfx9-family
// GPUs (data format 32 bits); a lane whose byte offset is at or past the
// resource's size stores nothing.
constexpr int kBufferWord3 = 0x00020000;
__device__ inline void buffer_store(uint32_t v, __amdgpu_buffer_rsrc_t r, int off)
{
    __builtin_amdgcn_raw_buffer_store_b32(v, r, off, 0, 0);
}
__device__ inline void buffer_store(uint64_t v, __amdgpu_buffer_rsrc_t r, int off)
{
    typedef uint32_t u32x2 __attribute__((ext_vector_type(2)));
    const u32x2 w = {static_cast<uint32_t>(v), static_cast<uint32_t>(v >> 32)};
    __builtin_amdgcn_raw_buffer_store_b64(w, r, off, 0, 0);
}

// Sorts each chunk in LDS and writes the sort's outputs.  A chunk's rows
// arrive ordered by their top digits (the global passes); stable 8-bit LDS
// passes, least significant first, over every digit that varies inside the
// chunk order it by the whole key (slots wave-major, each wave ranking its
// rows in order, per-wave bucket counters combined in wave order).  Row ids
// stay in the registers of the thread that loaded them; each column reaches
// its rows' final slots through LDS and leaves in slot order, so every output
// is written by consecutive lanes.  A chunk of two adjacent buckets (digit ga
// x and x + 1, one digit gb) needs the order of digit ga restored after the
// low digits; its last low pass instead ranks a 9-bit digit, the low digit
// plus bit 8 = (digit ga) - x, which saves that pass.  8 waves per block of
// at most 8 rows per thread, 2 blocks per CU.
template <int PB, bool ROWS>
__global__ __launch_bounds__(kLocalBlock) __attribute__((amdgpu_waves_per_eu(4)))
void k_local(SortArgs a)
{
    using KO = KeyOps<uint32_t>;
    using P = typename std::conditional<PB >= 8, uint64_t, uint32_t>::type;
    using P2 = typename std::conditional<PB == 12, uint32_t, uint64_t>::type;
    constexpr bool TWO = PB == 12 || PB == 16;
    constexpr int W = kLocalBlock / 64;         // waves
    constexpr int IM = kLocalRows / kLocalBlock;  // rows per thread at most
    constexpr int PI = PB ? IM : 1;
    constexpr int PI2 = TWO ? IM : 1;
    // 32 KiB: keys (16 KiB), their loaded slots (8 KiB) and the final slot of
    // each loaded slot (8 KiB) while sorting; one column of 8-byte cells at write-out
    __shared__ uint64_t s_raw[kLocalRows];
    // Per-wave bucket counters, two sets (a pass zeroes the next one's); the W
    // counters of bucket b sit at b * CS .. b * CS + W - 1, a row that one
    // thread reads and writes with paired-dword operations, and with CS odd
    // the counters of one wave fall in distinct banks for 64 consecutive buckets.
    constexpr int CS = W + 1;
    constexpr int NB = 2 * kRadix;  // buckets of a 9-bit pass (one thread each in the scan)
    static_assert(NB == kLocalBlock, "one bucket per thread in the 9-bit scan");
    __shared__ uint32_t s_wc[2][NB * CS];
    __shared__ uint32_t s_wsum[W];
    uint32_t* s_key = reinterpret_cast<uint32_t*>(s_raw);
    uint16_t* s_org = reinterpret_cast<uint16_t*>(s_key + kLocalRows);
    uint16_t* s_pos = s_org + kLocalRows;

    const uint32_t* hdr = a.plan_hdr;
    if (!(hdr[kHdrMode] & kModeLocal)) return;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const SegDesc& sd = a.segs[0];
    const uint32_t n = static_cast<uint32_t>(sd.n);
    const uint32_t K = a.lctl[0];
    const auto* recs = gmem(static_cast<const uint64_t*>(a.kbuf[1]));
    const auto* kin = gmem(static_cast<const uint32_t*>(a.kbuf[1]));
    const auto* pin = gmem(static_cast<const P*>(a.pbuf[1]));
    const auto* pin2 = gmem(static_cast<const P2*>(a.pbuf2[1]));
    // output pointers in scalar registers (buffer resources are scalar; a
    // pointer the compiler kept in vector registers costs a loop per store)
    auto sgpr_ptr = [](void* p) __attribute__((always_inline)) -> void* {
        const uint64_t v = reinterpret_cast<uint64_t>(p);
        const uint32_t lo = __builtin_amdgcn_readfirstlane(static_cast<uint32_t>(v));
        const uint32_t hi = __builtin_amdgcn_readfirstlane(static_cast<uint32_t>(v >> 32));
        return reinterpret_cast<void*>((static_cast<uint64_t>(hi) << 32) | lo);
    };
    void* keys_out = sgpr_ptr(sd.keys_out);
    void* perm_out = sgpr_ptr(sd.perm_out);
    void* pay_out = sgpr_ptr(sd.pay_out);
    void* pay_out2 = sgpr_ptr(sd.pay_out2);
    uint32_t key[IM], rid[IM], rnk[IM], org[IM];
    P pay[PI];
    P2 pay2[PI2];
    // the digits every chunk sorts: the non-constant ones below the two global ones
    uint32_t lmask = 0;
    for (uint32_t i = 0; i < hdr[kHdrNS]; ++i) lmask |= 1u << hdr[kHdrS + i];
    const uint32_t ga = hdr[kHdrGa], gb = hdr[kHdrGb];
    lmask &= ~((1u << ga) | (1u << gb));
    for (int i = tid; i < NB * CS; i += kLocalBlock) s_wc[0][i] = s_wc[1][i] = 0;
    lds_barrier();
    uint32_t q = 0;  // passes so far (counter set q & 1)

    // Software pipeline over the block's chunks k = blockIdx.x + i * gridDim.x.
    // Iteration i issues, in this order, vector loads of: chunk i's payload,
    // chunk i+1's (key, row id) records and first / last keys, chunk i+2's
    // bounds -- all unconditional (slots past a chunk read its last row), so
    // the waits before each use leave the later loads in flight.  Rows per
    // thread: wave w holds slots [w*64*I, (w+1)*64*I); slot sl0 + 64 j is row
    // r0 + sl0 + 64 j.
    const uint32_t G = gridDim.x;
    // chunk kk's start and end (past the last chunk: empty, at n); the load
    // itself is unconditional (a clamped index) so every path issues it
    auto bound = [&](uint32_t kk, uint32_t off) __attribute__((always_inline)) -> uint32_t {
        const uint32_t v = gmem(a.lchunk)[2 * static_cast<uint64_t>(std::min(kk, std::max(K, 1u) - 1)) + off];
        return kk < K ? v : n;
    };
    auto clampm = [](uint32_t r0x, uint32_t r1x) -> uint32_t {
        return r1x > r0x && r1x - r0x <= kLocalRows ? r1x - r0x : 0u;  // chunks are < kLocalRows rows
    };
    // the row of slot s of a chunk (an empty chunk reads the table's last row)
    auto row_at = [&](uint32_t r0x, uint32_t mx, uint32_t sx) __attribute__((always_inline)) -> uint32_t {
        return mx ? r0x + std::min(sx, mx - 1) : n - 1;
    };
    auto load_recs = [&](uint32_t r0x, uint32_t mx, uint64_t* dst) __attribute__((always_inline)) {
        const uint32_t Ix = (mx + kLocalBlock - 1) / kLocalBlock;
        const uint32_t s0 = static_cast<uint32_t>(wave) * 64 * Ix + lane;
#pragma unroll
        for (int j = 0; j < IM; ++j) {
            const uint32_t r = row_at(r0x, mx, s0 + j * 64);
            if constexpr (ROWS)
                dst[j] = recs[r];
            else
                dst[j] = kin[r];
        }
    };
    uint64_t nx[IM];
    // chunk 0 of this block: bounds, records and first / last keys, chunk 1's bounds
    uint32_t cr0 = bound(blockIdx.x, 0), cm = clampm(cr0, bound(blockIdx.x, 1));
    load_recs(cr0, cm, nx);
    // a row's key alone (32-bit loads: a 64-bit record whose upper half went
    // unused made the compiler wait for it, and so for every earlier load)
    auto key_at = [&](uint32_t r) __attribute__((always_inline)) -> uint32_t {
        return ROWS ? kin[2 * static_cast<uint64_t>(r)] : kin[r];
    };
    uint32_t kfl = key_at(row_at(cr0, cm, 0));
    uint32_t kll = key_at(row_at(cr0, cm, kLocalRows));
    uint32_t b0 = bound(blockIdx.x + G, 0), b1 = bound(blockIdx.x + G, 1);
    // waited for here, so that the loop's first wait is placed for the loads
    // of the loop alone (and leaves the previous chunk's stores in flight)
#pragma unroll
    for (int j = 0; j < IM; ++j) asm volatile("" ::"v"(nx[j]));
    asm volatile("" ::"v"(kfl), "v"(kll), "v"(b0), "v"(b1));
    for (uint32_t k = blockIdx.x; k < K; k += G) {
        const uint32_t r0 = cr0, m = cm;  // this chunk
        const uint32_t I = (m + kLocalBlock - 1) / kLocalBlock;
        const uint32_t sl0 = static_cast<uint32_t>(wave) * 64 * I + lane;
        auto valid = [&](int j) __attribute__((always_inline)) { return j < static_cast<int>(I) && sl0 + j * 64 < m; };
#pragma unroll
        for (int j = 0; j < IM; ++j) {
            key[j] = static_cast<uint32_t>(nx[j]);
            rid[j] = static_cast<uint32_t>(nx[j] >> 32);
        }
        const uint32_t kf = __builtin_amdgcn_readfirstlane(kfl);
        const uint32_t kl = __builtin_amdgcn_readfirstlane(kll);
        // this chunk's payload
#pragma unroll
        for (int j = 0; j < IM; ++j) {
            const uint32_t r = row_at(r0, m, sl0 + j * 64);
            if constexpr (PB != 0) pay[j] = pin[r];
            if constexpr (TWO) pay2[j] = pin2[r];
        }
        // the next chunk's records and first / last keys, the one after's bounds
        const uint32_t r0n = __builtin_amdgcn_readfirstlane(b0);
        const uint32_t mn = clampm(r0n, __builtin_amdgcn_readfirstlane(b1));
        load_recs(r0n, mn, nx);
        kfl = key_at(row_at(r0n, mn, 0));
        kll = key_at(row_at(r0n, mn, kLocalRows));
        b0 = bound(k + 2 * G, 0);
        b1 = bound(k + 2 * G, 1);
        cr0 = r0n;
        cm = mn;
        // (an empty chunk runs through as well: its loads read one row and its
        // stores are all dropped, so every path issues the same stores)
        // digits that vary inside the chunk: the low ones, and the global ones
        // where its first and last keys differ (its rows are ordered by them)
        // (two adjacent digits ga: the last low pass also ranks bit 8 = digit ga - fa)
        uint32_t vary = lmask;
        const uint32_t fa = KO::digit(kf, ga);
        bool comb = false;
        if (KO::digit(kf, gb) != KO::digit(kl, gb)) vary |= (1u << ga) | (1u << gb);
        else if (KO::digit(kl, ga) == fa + 1 && lmask) comb = true;
        else if (KO::digit(kl, ga) != fa) vary |= 1u << ga;
        if (vary == 0) vary = 1;  // one key value: one pass leaves the order and gives the final slots
#pragma unroll
        for (int j = 0; j < IM; ++j) org[j] = sl0 + j * 64;
        bool from_regs = true;
#pragma unroll 1
        for (int d = 0; d < 4; ++d) {
            if (!((vary >> d) & 1u)) continue;
            const bool last = (vary >> (d + 1)) == 0;
            if (!from_regs) {
                // the previous pass's order
#pragma unroll
                for (int j = 0; j < IM; ++j) {
                    if (j < static_cast<int>(I)) {  // slots < 512 I <= kLocalRows
                        key[j] = s_key[sl0 + j * 64];
                        org[j] = s_org[sl0 + j * 64];
                    }
                }
            }
            uint32_t* cs = s_wc[q & 1];
            uint32_t* wc = cs + wave;  // this wave's counter of bucket b: wc[b * CS]
            // the pass's digit: 8 bits, or 9 in the last pass of a two-bucket chunk
            const bool wide = comb && last;
            const uint32_t nb = wide ? NB : kRadix;
            auto digit_of = [&](uint32_t kv) __attribute__((always_inline)) -> uint32_t {
                const uint32_t g = KO::digit(kv, d);
                return wide ? g | (((KO::digit(kv, ga) - fa) & 1u) << kRadixBits) : g;
            };
            if (a.lds_rank) {
#pragma unroll
                for (int j = 0; j < IM; ++j)
                    if (j < static_cast<int>(I)) {
                        const bool v = sl0 + j * 64 < m;
                        rnk[j] = v ? atomicAdd(&wc[digit_of(key[j]) * CS], 1u) : 0u;
                    }
            } else {
                // (nine ballots: bit 8 is zero in every lane of an 8-bit pass)
#pragma unroll
                for (int j = 0; j < IM; ++j)
                    if (j < static_cast<int>(I)) {
                        const uint32_t dg = digit_of(key[j]);
                        rnk[j] = rank_ballot_at<kRadixBits + 1>(dg, sl0 + j * 64 < m, &wc[dg * CS]);
                    }
            }
            lds_barrier();
            // bucket tid: per-wave offsets (waves in order), then the exclusive
            // scan over the nb buckets (one per thread)
            uint32_t tot = 0, ex[W];
            if (static_cast<uint32_t>(tid) < nb) {
                uint32_t c[W];
#pragma unroll
                for (int w = 0; w < W; ++w) c[w] = cs[tid * CS + w];
#pragma unroll
                for (int w = 0; w < W; ++w) {
                    ex[w] = tot;
                    tot += c[w];
                }
            }
            const uint32_t inc = wave_incl_scan(tot);
            if (lane == 63) s_wsum[wave] = inc;
            lds_barrier();
            {
                // the offsets, and the other set (last read before the previous
                // barrier) cleared for the next pass, all NB buckets
                uint32_t* nxt = s_wc[(q + 1) & 1];
                if (static_cast<uint32_t>(tid) < nb) {
                    uint32_t pre = 0;
#pragma unroll
                    for (int i = 0; i < W; ++i) pre += (i < wave) ? s_wsum[i] : 0u;
                    const uint32_t lstart = pre + inc - tot;
#pragma unroll
                    for (int w = 0; w < W; ++w) cs[tid * CS + w] = ex[w] + lstart;
                }
#pragma unroll
                for (int w = 0; w < W; ++w) nxt[tid * CS + w] = 0;
            }
            lds_barrier();
            // the slots of the next pass; the last pass gives each loaded slot its final slot
#pragma unroll
            for (int j = 0; j < IM; ++j) {
                if (valid(j)) {
                    const uint32_t p = wc[digit_of(key[j]) * CS] + rnk[j];
                    s_key[p] = key[j];
                    if (last)
                        s_pos[org[j]] = static_cast<uint16_t>(p);
                    else
                        s_org[p] = static_cast<uint16_t>(org[j]);
                }
            }
            from_regs = false;
            ++q;
            if (last) break;
            lds_barrier();
        }
        lds_barrier();

        // Write-out: final slot s goes to output row r0 + s (n - 1 - r0 - s when
        // reversed).  Every column leaves through buffer stores over the
        // chunk's m output rows, issued by every lane whatever the chunk and
        // the outputs asked for: slots past the chunk fall outside the buffer
        // and are dropped, and so is a column that is not asked for (an empty
        // buffer).  Vector memory operations complete in issue order; with a
        // store count fixed on every path, the next iteration's wait for the
        // loads issued before these stores leaves the stores in flight (with
        // conditional stores the compiler waited for all of them).
        const uint64_t orow = a.reverse ? static_cast<uint64_t>(n) - r0 - m : r0;
        auto rsrc = [&](void* base, uint32_t es) __attribute__((always_inline)) {
            return __builtin_amdgcn_make_buffer_rsrc(static_cast<char*>(base) + (base ? orow * es : 0), 0,
                                                     base ? static_cast<int>(m * es) : 0, kBufferWord3);
        };
        auto boff = [&](int u, uint32_t es) __attribute__((always_inline)) -> int {
            const uint32_t sl = static_cast<uint32_t>(tid + u * kLocalBlock);
            return static_cast<int>(sl < m ? (a.reverse ? m - 1 - sl : sl) * es : 0x80000000u);
        };
        {
            const auto r = rsrc(keys_out, 4);
            uint32_t kv[IM];
#pragma unroll
            for (int u = 0; u < IM; ++u) kv[u] = 0;
            if (keys_out) {
#pragma unroll
                for (int u = 0; u < IM; ++u) kv[u] = KO::unxform(s_key[tid + u * kLocalBlock], a.mlo, a.mhi);
            }
#pragma unroll
            for (int u = 0; u < IM; ++u) buffer_store(kv[u], r, boff(u, 4));
        }
#pragma unroll
        for (int j = 0; j < IM; ++j)
            if (j < static_cast<int>(I)) rnk[j] = s_pos[sl0 + j * 64];
        // one column at a time: each row's cell to its final slot in LDS, then
        // out in slot order
        auto column = [&](auto* stage, auto get, void* base) __attribute__((always_inline)) {
            using T = typename std::remove_reference<decltype(stage[0])>::type;
            const auto r = rsrc(base, sizeof(T));
            T v[IM];
#pragma unroll
            for (int u = 0; u < IM; ++u) v[u] = 0;
            if (base) {
                lds_barrier();  // the previous phase's LDS reads are done
#pragma unroll
                for (int j = 0; j < IM; ++j)
                    if (valid(j)) stage[rnk[j]] = get(j);
                lds_barrier();
#pragma unroll
                for (int u = 0; u < IM; ++u) v[u] = stage[tid + u * kLocalBlock];  // (slots < kLocalRows)
            }
#pragma unroll
            for (int u = 0; u < IM; ++u) buffer_store(v[u], r, boff(u, sizeof(T)));
        };
        if constexpr (ROWS) {
            if (perm_out && a.perm64)
                column(s_raw, [&](int j) __attribute__((always_inline)) { return static_cast<uint64_t>(rid[j]); },
                       perm_out);
            else
                column(s_key, [&](int j) __attribute__((always_inline)) { return rid[j]; }, perm_out);
        }
        if constexpr (PB != 0)
            column(reinterpret_cast<P*>(s_raw), [&](int j) __attribute__((always_inline)) { return pay[j]; },
                   pay_out);
        if constexpr (TWO)
            column(reinterpret_cast<P2*>(s_raw), [&](int j) __attribute__((always_inline)) { return pay2[j]; },
                   pay_out2);
        // (the next chunk's first LDS writes to s_raw come three barriers later)
    }
}

using LocalFn = void (*)(SortArgs);
static LocalFn pick_local(int pay_bytes, bool rows)
{
    switch (pay_bytes) {
    case 0: return rows ? k_local<0, true> : k_local<0, false>;
    case 4: return rows ? k_local<4, true> : k_local<4, false>;
    case 8: return rows ? k_local<8, true> : k_local<8, false>;
    case 12: return rows ? k_local<12, true> : k_local<12, false>;
    case 16: return rows ? k_local<16, true> : k_local<16, false>;
    default: return nullptr;
    }
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
size_t key_bytes(int key_type)
{
    switch (key_type) {
    case VDF_KEY_U32:
    case VDF_KEY_I32:
    case VDF_KEY_F32: return 4;
    case VDF_KEY_U64:
    case VDF_KEY_I64:
    case VDF_KEY_F64: return 8;
    case VDF_KEY_CHAR16: return 16;
    default: return 0;
    }
}

static int key_digits(int key_type) { return static_cast<int>(key_bytes(key_type)); }

constexpr uint64_t kMinTileRows = static_cast<uint64_t>(kBlock) * 8;  // smallest pass tile

// Scratch of a sort batch.  Tile-indexed arrays are sized for the smallest
// pass tile, so they hold every tuning choice.
struct SortLayout {
    uint64_t rows = 0, tiles = 0, items = 0;
    size_t zero_words = 0;  // tickets and histograms (zeroed by every sort)
    uint64_t lchunks = 0;   // local finish: chunk words (0: no local finish possible)
    size_t bytes = 0;
};

static uint64_t hist_items(uint64_t n, uint64_t rows_per_item)
{
    return n == 0 ? 0 : (n + rows_per_item - 1) / rows_per_item;
}

// Rows per histogram work item for a sort of `rows` rows in all: the tuned
// grain, or a smaller one (down to 4096 rows) when the grain would give
// fewer than kMinHistItems items, so small sorts still spread over the GPU.
constexpr uint64_t kMinHistItems = 512;
static uint64_t hist_grain(uint64_t hist_rows, uint64_t rows)
{
    if (rows / hist_rows >= kMinHistItems) return hist_rows;
    const uint64_t g = (rows / kMinHistItems + 255) / 256 * 256;
    return std::min(hist_rows, std::max<uint64_t>(g, 4096));
}

static SortLayout sort_layout(int key_type, const uint64_t* n, int nsegs, uint64_t hist_rows, int pay_bytes = 0)
{
    SortLayout L;
    for (int s = 0; s < nsegs; ++s) L.rows += n[s];
    hist_rows = hist_grain(hist_rows, L.rows);
    for (int s = 0; s < nsegs; ++s) {
        L.tiles += (n[s] + kMinTileRows - 1) / kMinTileRows;
        L.items += hist_items(n[s], hist_rows);
    }
    const int D = key_digits(key_type);
    const size_t kb = key_bytes(key_type);
    L.zero_words = kMaxDigits * kTicketStride + static_cast<size_t>(nsegs) * D * kRadix;
    size_t b = 0;
    b += align_up(L.zero_words * 4, 256);
    b += align_up(static_cast<size_t>(nsegs) * kPlanHdr * 4, 256);
    b += align_up(static_cast<size_t>(nsegs) * D * kRadix * 4, 256);
    b += align_up(static_cast<size_t>(nsegs) * sizeof(SegDesc), 256);
    b += align_up(std::max<uint64_t>(L.items, 1) * sizeof(HistItem), 256);
    b += align_up(std::max<uint64_t>(L.tiles, 1) * 4, 256);
    b += 2 * align_up(std::max<uint64_t>(L.tiles, 1) * kRadix * 4, 256);
    b += 2 * align_up(L.rows * kb, 256);
    b += 2 * align_up(L.rows * 4, 256);
    b += 2 * align_up(L.rows * static_cast<uint64_t>(pay_bytes), 256);
    if (kb == 4 && nsegs == 1 && L.rows >= kLocalMinSortRows) {
        // local finish: bucket starts, chunk (start, end) pairs, control words
        L.lchunks = local_chunk_words(L.rows);
        b += align_up(static_cast<size_t>(kRadix) * (kRadix + 1) * 4, 256);
        b += align_up(L.lchunks * 4, 256);
        b += 256;
    }
    L.bytes = b + 256;
    return L;
}

// host staging bytes of a sort batch (descriptors, work items, tile map)
static size_t staging_bytes(const SortLayout& L, int nsegs)
{
    return align_up(static_cast<size_t>(nsegs) * sizeof(SegDesc), 256) +
           align_up(std::max<uint64_t>(L.items, 1) * sizeof(HistItem), 256) +
           align_up(std::max<uint64_t>(L.tiles, 1) * 4, 256);
}

size_t argsort_scratch(int key_type, uint64_t n, uint32_t)
{
    return sort_layout(key_type, &n, 1, kDefaultHistRows).bytes;
}

size_t argsort_batch_scratch(int key_type, const uint64_t* n, int nsegs, uint64_t hist_rows, int pay_bytes)
{
    return sort_layout(key_type, n, nsegs, hist_rows, pay_bytes).bytes;
}

size_t argsort_batch_staging(int key_type, const uint64_t* n, int nsegs, uint64_t hist_rows)
{
    return staging_bytes(sort_layout(key_type, n, nsegs, hist_rows), nsegs);
}

using PassFn = void (*)(SortArgs, uint32_t);

// Pass kernel instantiations the tuning table can select (tile rows per
// thread), without and with a 4- or 8-byte payload.
template <typename K>
static PassFn pick_pass(int items, int pay_bytes = 0);

// kernels that move no row ids (sorts without a permutation), where they exist
template <typename K>
static PassFn pick_pass_norows(int, int) { return nullptr; }
template <>
PassFn pick_pass_norows<uint32_t>(int items, int pay_bytes)
{
    if (pay_bytes == 8 && items == 24) return k_pass<uint32_t, 24, 8, false, false>;
    if (pay_bytes == 16 && items == 12) return k_pass<uint32_t, 12, 16, false, false>;
    if (pay_bytes == 16 && items == 24) return k_pass<uint32_t, 24, 16, true, false>;
    if (pay_bytes == 12 && items == 16) return k_pass<uint32_t, 16, 12, false, false>;
    return nullptr;
}

template <typename K>
static PassFn pick_pass_for(int items, int pay_bytes, bool rows)
{
    if (!rows)
        if (PassFn f = pick_pass_norows<K>(items, pay_bytes)) return f;
    return pick_pass<K>(items, pay_bytes);
}

template <>
PassFn pick_pass<uint32_t>(int items, int pay_bytes)
{
    switch (pay_bytes) {
    case 0:
        if (items == 12) return k_pass<uint32_t, 12, 0>;
        if (items == 16) return k_pass<uint32_t, 16, 0>;
        if (items == 24) return k_pass<uint32_t, 24, 0>;
        if (items == 32) return k_pass<uint32_t, 32, 0>;
        return nullptr;
    case 4:
        if (items == 8) return k_pass<uint32_t, 8, 4>;
        if (items == 12) return k_pass<uint32_t, 12, 4>;
        if (items == 24) return k_pass<uint32_t, 24, 4, true>;
        return nullptr;
    case 8:
        if (items == 12) return k_pass<uint32_t, 12, 8>;
        if (items == 16) return k_pass<uint32_t, 16, 8>;
        if (items == 24) return k_pass<uint32_t, 24, 8, true>;
        return nullptr;
    case 12:  // an 8-byte and a 4-byte column (4-byte keys only)
        if (items == 8) return k_pass<uint32_t, 8, 12>;
        if (items == 12) return k_pass<uint32_t, 12, 12>;
        return nullptr;
    case 16:
        if (items == 8) return k_pass<uint32_t, 8, 16>;
        if (items == 12) return k_pass<uint32_t, 12, 16>;
        if (items == 16) return k_pass<uint32_t, 16, 16, true>;
        return nullptr;
    default: return nullptr;
    }
}

template <>
PassFn pick_pass<uint64_t>(int items, int pay_bytes)
{
    switch (pay_bytes) {
    case 0:
        if (items == 8) return k_pass<uint64_t, 8, 0>;
        if (items == 12) return k_pass<uint64_t, 12, 0>;
        if (items == 16) return k_pass<uint64_t, 16, 0>;
        return nullptr;
    case 4:
        if (items == 8) return k_pass<uint64_t, 8, 4>;
        if (items == 12) return k_pass<uint64_t, 12, 4>;
        return nullptr;
    case 8:
        if (items == 8) return k_pass<uint64_t, 8, 8>;
        if (items == 12) return k_pass<uint64_t, 12, 8>;
        if (items == 16) return k_pass<uint64_t, 16, 8, true>;
        return nullptr;
    case 16:
        if (items == 8) return k_pass<uint64_t, 8, 16>;
        if (items == 16) return k_pass<uint64_t, 16, 16, true>;
        return nullptr;
    default: return nullptr;
    }
}

template <>
PassFn pick_pass<K128>(int items, int pay_bytes)
{
    if (items == 8) {
        switch (pay_bytes) {
        case 0: return k_pass<K128, 8, 0>;
        case 4: return k_pass<K128, 8, 4>;
        case 8: return k_pass<K128, 8, 8>;
        case 16: return k_pass<K128, 8, 16>;
        default: return nullptr;
        }
    }
    if (items == 16) {
        switch (pay_bytes) {
        case 0: return k_pass<K128, 16, 0, true>;
        case 4: return k_pass<K128, 16, 4, true>;
        case 8: return k_pass<K128, 16, 8, true>;
        default: return nullptr;
        }
    }
    return nullptr;
}

bool tune_supported(const char* name, int64_t v)
{
    const std::string n(name);
    // a tile size is accepted when some payload width has a kernel for it (a
    // sort whose payload has none uses that payload's default size)
    auto any_pass = [&](auto pick) {
        for (int pb : {0, 4, 8, 16})
            if (pick(static_cast<int>(v), pb) != nullptr) return true;
        return false;
    };
    if (n == "sort_items_32") return v == 0 || any_pass(pick_pass<uint32_t>);
    if (n == "sort_items_64") return v == 0 || any_pass(pick_pass<uint64_t>);
    if (n == "sort_items_128") return v == 0 || any_pass(pick_pass<K128>);
    if (n == "sort_chunk") return v >= 1 && v <= 1024;
    if (n == "sort_xcd") return v == 0 || v == 1;
    if (n == "sort_blocks_per_cu") return v >= 0 && v <= 32;
    if (n == "hist_rows") return v >= 4096 && v <= (1 << 24);
    if (n == "gather_rows") return v == 4 || v == 8 || v == 16;
    if (n == "gather_nt") return v == 0 || v == 1;
    if (n == "csv_stage") return v == 0 || v == 1;
    if (n == "groupby_carry") return v == 0 || v == 1;
    if (n == "sort_local") return v == 0 || v == 1;
    if (n == "sort_rank") return v == 0 || v == 1;
    return false;
}

// Compute units of a device (cached).
static unsigned device_cus(int device)
{
    static std::mutex mu;
    static std::vector<std::pair<int, unsigned>> cache;
    std::lock_guard<std::mutex> lk(mu);
    for (const auto& e : cache)
        if (e.first == device) return e.second;
    int cus = 0;
    VDF_HIP_CHECK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device));
    cache.emplace_back(device, static_cast<unsigned>(std::max(1, cus)));
    return cache.back().second;
}

// Resident blocks of a pass kernel on the whole device (persistent grid).
static unsigned resident_blocks(const void* fn, int device, int threads = kBlock)
{
    struct Entry {
        const void* fn;
        int device;
        unsigned blocks;
    };
    static std::mutex mu;
    static std::vector<Entry> cache;
    std::lock_guard<std::mutex> lk(mu);
    for (const Entry& e : cache)
        if (e.fn == fn && e.device == device) return e.blocks;
    int per_cu = 0, cus = 0;
    VDF_HIP_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, threads, 0));
    VDF_HIP_CHECK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device));
    const unsigned blocks = static_cast<unsigned>(std::max(1, per_cu) * std::max(1, cus));
    cache.push_back(Entry{fn, device, blocks});
    return blocks;
}

// Default rows per thread of a pass tile by key and payload width (measured
// on MI355X, 100M-row sorts): the largest tile whose kernel keeps two blocks
// per CU without spilling.
template <typename K>
static int default_items(int pay_bytes, bool rows)
{
    if (sizeof(K) == 4) {
        switch (pay_bytes) {
        case 0: return 32;   // 8-byte (key, row id) records: 8192-row tiles in 64 KiB
        case 8: return 24;   // split staging; without row ids, the row-id-free kernel
        case 12: return rows ? 12 : 16;  // 8 + 4 bytes; without row ids 4096-row tiles
        case 16: return rows ? 12 : 24;  // two 8-byte columns; without row ids split staging
                                         // (6144-row tiles; 2 % faster than 3072, and with
                                         // a 12-byte payload split tiles lost to 4096 unsplit)
        default: return 12;  // 4-byte payloads
        }
    }
    (void)rows;
    if (sizeof(K) == 8) return pay_bytes == 16 ? 8 : (pay_bytes == 8 ? 16 : 12);  // 16: split staging
    return 8;
}

template <typename K>
static void launch_sort(Handle* h, const SortSeg* segs, int nsegs, uint64_t mlo, uint64_t mhi, uint32_t flags,
                        void* scr, hipStream_t s, int pay_bytes, bool rowids)
{
    constexpr int D = KeyOps<K>::kDigits;
    int items = sizeof(K) == 4 ? h->tune.sort_items_32 : (sizeof(K) == 8 ? h->tune.sort_items_64 : h->tune.sort_items_128);
    const bool rows_kernel = rowids || !h->rowfree_kernels;
    PassFn pass = items ? pick_pass_for<K>(items, pay_bytes, rows_kernel) : nullptr;
    if (!pass) {  // the default, or a tuning value without a kernel for this key width / payload
        items = default_items<K>(pay_bytes, rows_kernel);
        pass = pick_pass_for<K>(items, pay_bytes, rows_kernel);
    }
    require(pass != nullptr, "sort payload must be 0, 4, 8 or 16 bytes per row");
    const uint64_t tile_rows = static_cast<uint64_t>(kBlock) * items;
    // one segment: its row count in place (no host allocation on that path)
    uint64_t n_one = nsegs == 1 ? segs[0].n : 0;
    std::vector<uint64_t> ns_many;
    uint64_t all_rows = n_one;
    if (nsegs != 1) {
        ns_many.resize(nsegs);
        for (int i = 0; i < nsegs; ++i) {
            ns_many[i] = segs[i].n;
            all_rows += segs[i].n;
        }
    }
    const uint64_t* ns = nsegs == 1 ? &n_one : ns_many.data();
    const SortLayout L = sort_layout(sizeof(K) == 4 ? VDF_KEY_U32 : (sizeof(K) == 8 ? VDF_KEY_U64 : VDF_KEY_CHAR16),
                                     ns, nsegs, static_cast<uint64_t>(h->tune.hist_rows), pay_bytes);
    // the grain sort_layout counted the work items with
    const uint64_t hist_rows = hist_grain(static_cast<uint64_t>(h->tune.hist_rows), all_rows);

    // ---- batch description: host-built and uploaded through pinned staging,
    // or (device_desc, one segment) written on the device by k_setup_one ----
    const bool devdesc = h->device_desc;
    require(!devdesc || nsegs == 1, "device-written descriptors take one segment");
    char* stg = nullptr;
    uint32_t* ht = nullptr;
    uint64_t row = 0, tiles = 0, nitems = 0;
    SegDesc one{};
    if (devdesc) {
        const uint64_t n = segs[0].n;
        tiles = (n + tile_rows - 1) / tile_rows;
        nitems = (n + hist_rows - 1) / hist_rows;
        one = SegDesc{segs[0].keys, segs[0].perm_out, segs[0].keys_out, segs[0].pay_in, segs[0].pay_out,
                      segs[0].pay_in2, segs[0].pay_out2, n, 0, 0, 0};
    } else {
        stg = static_cast<char*>(h->staging_take(staging_bytes(L, nsegs)));
        Carve hc(stg);
        SegDesc* hs = hc.take<SegDesc>(nsegs);
        HistItem* hi = hc.take<HistItem>(std::max<uint64_t>(L.items, 1));
        ht = hc.take<uint32_t>(std::max<uint64_t>(L.tiles, 1));
        for (int i = 0; i < nsegs; ++i) {
            const uint64_t n = segs[i].n;
            const uint64_t nt = (n + tile_rows - 1) / tile_rows;
            hs[i] = SegDesc{segs[i].keys,    segs[i].perm_out, segs[i].keys_out, segs[i].pay_in, segs[i].pay_out,
                            segs[i].pay_in2, segs[i].pay_out2, n,             row,            static_cast<uint32_t>(tiles),
                            0};
            for (uint64_t t = 0; t < nt; ++t) ht[tiles + t] = static_cast<uint32_t>(i);
            for (uint64_t b = 0; b < n; b += hist_rows) hi[nitems++] = HistItem{static_cast<uint32_t>(i), 0, b, std::min(n, b + hist_rows)};
            row += n;
            tiles += nt;
        }
    }
    const uint64_t ntiles = tiles;
    if (ntiles == 0 || h->dry) return;  // every segment empty, or only the staging was wanted

    Carve c(scr);
    uint32_t* zero = c.take<uint32_t>(L.zero_words);
    SortArgs a{};
    a.tickets = zero;
    a.hist = zero + kMaxDigits * kTicketStride;
    a.plan_hdr = c.take<uint32_t>(static_cast<size_t>(nsegs) * kPlanHdr);
    a.plan_base = c.take<uint32_t>(static_cast<size_t>(nsegs) * D * kRadix);
    SegDesc* dsegs = c.take<SegDesc>(nsegs);
    HistItem* ditems = c.take<HistItem>(std::max<uint64_t>(L.items, 1));
    uint32_t* dtiles = c.take<uint32_t>(std::max<uint64_t>(L.tiles, 1));
    a.lb[0] = c.take<uint32_t>(std::max<uint64_t>(L.tiles, 1) * kRadix);
    a.lb[1] = c.take<uint32_t>(std::max<uint64_t>(L.tiles, 1) * kRadix);
    // kbuf[i] directly followed by vbuf[i]: with 4-byte keys the pair is one
    // buffer of 8-byte (key, row id) records
    a.kbuf[0] = c.take<char>(L.rows * sizeof(K));
    a.vbuf[0] = c.take<uint32_t>(L.rows);
    a.kbuf[1] = c.take<char>(L.rows * sizeof(K));
    a.vbuf[1] = c.take<uint32_t>(L.rows);
    if (sizeof(K) == 4)
        for (int i = 0; i < 2; ++i)
            require(static_cast<char*>(a.kbuf[i]) + L.rows * 8 <= reinterpret_cast<char*>(a.vbuf[i] + L.rows),
                    "record buffer layout");
    a.pbuf[0] = c.take<char>(L.rows * static_cast<uint64_t>(pay_bytes));
    a.pbuf[1] = c.take<char>(L.rows * static_cast<uint64_t>(pay_bytes));
    // a 16-byte payload is two 8-byte columns: the second in the upper half
    a.pbuf2[0] = static_cast<char*>(a.pbuf[0]) + (pay_bytes >= 12 ? L.rows * 8 : 0);
    a.pbuf2[1] = static_cast<char*>(a.pbuf[1]) + (pay_bytes >= 12 ? L.rows * 8 : 0);
    if (L.lchunks) {
        a.lstart = c.take<uint32_t>(static_cast<size_t>(kRadix) * (kRadix + 1));
        a.lchunk = c.take<uint32_t>(L.lchunks);
        a.lctl = c.take<uint32_t>(64);
    }
    // local finish: one segment of 4-byte keys with enough rows that two
    // global passes and a streaming LDS pass beat the plain passes
    LocalFn local = nullptr;
    if (L.lchunks && h->tune.sort_local && all_rows >= kLocalMinSortRows) local = pick_local(pay_bytes, rowids);
    a.local_ok = local ? 1u : 0u;
    a.lds_rank = (h->tune.sort_rank && lds_rank_verified(h->device)) ? 1u : 0u;
    if (devdesc) {
        const unsigned blocks = static_cast<unsigned>((nitems + kBlock - 1) / kBlock);
        hipLaunchKernelGGL(k_setup_one, dim3(blocks), dim3(kBlock), 0, s, one, dsegs, ditems, nitems, hist_rows);
        check_launch("k_setup_one");
    } else {
        const size_t up = reinterpret_cast<char*>(ht + ntiles) - stg;
        VDF_HIP_CHECK(hipMemcpyAsync(dsegs, stg, up, hipMemcpyHostToDevice, s));  // descs, items, tile map
    }
    // the host carve mirrors the device carve (same order, same alignment)
    a.segs = dsegs;
    a.items = ditems;
    a.tile_seg = nsegs > 1 ? dtiles : nullptr;
    a.mlo = mlo;
    a.mhi = mhi;
    a.lb_words = ntiles * kRadix;
    a.nsegs = static_cast<uint32_t>(nsegs);
    a.ntiles = static_cast<uint32_t>(ntiles);
    a.perm64 = (flags & VDF_SORT_PERM64) ? 1u : 0u;
    a.reverse = (flags & VDF_SORT_DESC_REVERSE) ? 1u : 0u;
    a.norows = rowids ? 0u : 1u;
    // persistent grid: every resident block (or the tuned blocks per CU)
    uint64_t cap = resident_blocks(reinterpret_cast<const void*>(pass), h->device);
    if (h->tune.sort_bpc > 0) cap = std::min<uint64_t>(cap, static_cast<uint64_t>(h->tune.sort_bpc) * device_cus(h->device));
    const unsigned grid = static_cast<unsigned>(std::min<uint64_t>(cap, ntiles));
    if (h->tune.sort_xcd != 0) {
        a.groups = kXcds;
        // enough chunks that every XCD gets several
        a.chunk = static_cast<uint32_t>(std::max<uint64_t>(
            1, std::min<uint64_t>(static_cast<uint64_t>(h->tune.sort_chunk), ntiles / (kXcds * 8ull))));
    } else {
        // tiles claimed in global order from one counter
        a.groups = 1;
        a.chunk = 1;
    }

    const size_t kb = sizeof(K);
    {
        LaunchTimer t(h, s, "fill", 0);
        fill_u32(zero, L.zero_words, 0u, s);
    }
    {
        LaunchTimer t(h, s, "radix_histogram", L.rows * kb);
        hipLaunchKernelGGL(k_histogram<K>, dim3(static_cast<unsigned>(nitems)), dim3(kBlock), 0, s, a);
        check_launch("k_histogram");
    }
    {
        LaunchTimer t(h, s, "radix_plan", 0);
        hipLaunchKernelGGL(k_plan<D>, dim3(static_cast<unsigned>(nsegs)), dim3(kBlock), 0, s, a);
        check_launch("k_plan");
    }
    // Active passes per table are decided on the device; for timing, the
    // tables' pass counts are copied back (into the handle's pinned ring) so
    // each pass launch is charged the bytes of the tables it worked on.
    std::shared_ptr<PassAcct> acct;
    const uint64_t pb_out = segs[0].perm_out ? (a.perm64 ? 8 : 4) : 0;
    const uint64_t kout = segs[0].keys_out ? kb : 0;
    const uint64_t rid = rowids ? 4 : 0;  // row ids read / written by a pass
    // pinned copy of every table's (pass slots end, mode) header words
    auto copy_plan = [&]() -> const uint32_t* {
        if (h->ring_pos + 2 * static_cast<size_t>(nsegs) > Handle::kRing) h->ring_pos = 0;
        uint32_t* dst = h->passes_ring + h->ring_pos;
        h->ring_pos += 2 * static_cast<size_t>(nsegs);
        VDF_HIP_CHECK(hipMemcpy2DAsync(dst, 2 * sizeof(uint32_t), a.plan_hdr, kPlanHdr * sizeof(uint32_t),
                                       2 * sizeof(uint32_t), static_cast<size_t>(nsegs), hipMemcpyDeviceToHost, s));
        return dst;
    };
    if (h->prof && h->passes_ring && 2 * static_cast<size_t>(nsegs) <= Handle::kRing / 2) {
        acct = std::make_shared<PassAcct>();
        acct->passes = copy_plan();
        acct->rows.assign(ns, ns + nsegs);
        // keys read + keys written + row ids read + row ids written, per row, when
        // the pass is the table's only / first / middle / last one
        // (plus the payload read and written in every pass)
        const uint64_t pay2 = 2 * static_cast<uint64_t>(pay_bytes);
        acct->per_row[0] = kb + kout + pb_out + pay2;
        acct->per_row[1] = kb + kb + rid + pay2;
        acct->per_row[2] = 2 * kb + 2 * rid + pay2;
        acct->per_row[3] = kb + rid + kout + pb_out + pay2;
        // a local finish reads the last global pass's keys, row ids and payload
        // and writes the outputs
        acct->per_row_local = kb + rid + pay_bytes + kout + pb_out + pay_bytes;
    }
    const uint64_t middle_bytes = L.rows * (2 * kb + 2 * rid + 2 * static_cast<uint64_t>(pay_bytes));
    auto launch_pass = [&](int slot) {
        LaunchTimer t(h, s, "radix_pass", middle_bytes, slot, acct);
        uint32_t sl = static_cast<uint32_t>(slot);
        void* args[] = {&a, &sl};
        VDF_HIP_CHECK(hipLaunchKernel(reinterpret_cast<const void*>(pass), dim3(grid), dim3(kBlock), args, 0, s));
    };
    if (!local) {
        for (int slot = 0; slot < D; ++slot) launch_pass(slot);
        return;
    }
    // local finish: slots 0 and 1 scatter the top two digits, k_local sorts the
    // buckets; slots kLocalSlot0.. run only when k_local_chunks fell back
    launch_pass(0);
    launch_pass(1);
    {
        LaunchTimer t(h, s, "radix_local_plan", 0);
        hipLaunchKernelGGL(k_local_bounds, dim3(kRadix + 1), dim3(kBlock), 0, s, a, static_cast<uint32_t>(tile_rows));
        check_launch("k_local_bounds");
        hipLaunchKernelGGL(k_local_max, dim3(kRadix), dim3(kBlock), 0, s, a);
        check_launch("k_local_max");
        hipLaunchKernelGGL(k_local_chunks, dim3(kRadix), dim3(kBlock), 0, s, a);
        check_launch("k_local_chunks");
    }
    std::shared_ptr<PassAcct> acct2;
    if (acct) {
        acct2 = std::make_shared<PassAcct>(*acct);
        acct2->passes2 = copy_plan();
        acct = acct2;
    }
    {
        LaunchTimer t(h, s, "radix_local", L.rows * (acct ? acct->per_row_local : 0), kProfLocalSlot, acct);
        const unsigned lgrid = resident_blocks(reinterpret_cast<const void*>(local), h->device, kLocalBlock);
        void* args[] = {&a};
        VDF_HIP_CHECK(hipLaunchKernel(reinterpret_cast<const void*>(local), dim3(lgrid), dim3(kLocalBlock), args, 0, s));
    }
    for (int slot = kLocalSlot0; slot < kLocalMaxSlots; ++slot) launch_pass(slot);
}

static void masks_for(int key_type, uint32_t flags, uint64_t& mlo, uint64_t& mhi)
{
    const bool desc = (flags & VDF_SORT_DESC) != 0;
    switch (key_type) {
    case VDF_KEY_U32:
    case VDF_KEY_I32:
        mlo = (key_type == VDF_KEY_I32 ? 0x80000000ull : 0ull) ^ (desc ? 0xFFFFFFFFull : 0ull);
        mhi = 0;
        break;
    case VDF_KEY_U64:
    case VDF_KEY_I64:
        mlo = (key_type == VDF_KEY_I64 ? 0x8000000000000000ull : 0ull) ^ (desc ? ~0ull : 0ull);
        mhi = 0;
        break;
    case VDF_KEY_F32:
        mlo = desc ? 0xFFFFFFFFull : 0ull;
        mhi = 1;  // float order transform
        break;
    case VDF_KEY_F64:
        mlo = desc ? ~0ull : 0ull;
        mhi = 1;
        break;
    case VDF_KEY_CHAR16:
        mlo = mhi = 0x8080808080808080ull ^ (desc ? ~0ull : 0ull);
        break;
    default: throw Error{VDF_ERR_INVALID, "unknown key type"};
    }
}

void argsort_batch_enqueue(Handle* h, int key_type, const SortSeg* segs, int nsegs, uint32_t flags, void* scr,
                           hipStream_t s, int pay_bytes, bool rowids)
{
    uint64_t mlo = 0, mhi = 0;
    masks_for(key_type, flags, mlo, mhi);
    if (!rowids)
        for (int i = 0; i < nsegs; ++i)
            require(segs[i].perm_out == nullptr, "a sort without row ids cannot write a permutation");
    switch (key_bytes(key_type)) {
    case 4: launch_sort<uint32_t>(h, segs, nsegs, mlo, mhi, flags, scr, s, pay_bytes, rowids); break;
    case 8: launch_sort<uint64_t>(h, segs, nsegs, mlo, mhi, flags, scr, s, pay_bytes, rowids); break;
    default: launch_sort<K128>(h, segs, nsegs, mlo, mhi, flags, scr, s, pay_bytes, rowids); break;
    }
}

void argsort_enqueue(Handle* h, int key_type, const void* d_keys, uint64_t n, uint32_t flags, void* d_perm_out,
                     void* d_keys_out, void* scr, hipStream_t s)
{
    const SortSeg seg{d_keys, n, d_perm_out, d_keys_out};
    argsort_batch_enqueue(h, key_type, &seg, 1, flags, scr, s);
}

}  // namespace vdf
