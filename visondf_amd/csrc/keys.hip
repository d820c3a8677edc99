// keys.hip -- multi-column keys for the sort and the group-by.
//
// The reference groups by several key columns by concatenating each row's key
// bytes with '\x1F' separators into one hash key
// (transform_group_by_sametype_mt.hpp:13, transform_group_by_difftype_mt.hpp:13,
// inlines/transform/transform_group_by/key_build.inl:3-110): equal rows are
// equal in every column.  Here the columns are packed into one composite key
// whose unsigned order is the lexicographic order of the columns (first
// column most significant), so the single-key radix sort and group-by do the
// work:
//   each column is mapped to its order-preserving unsigned form (u: identity,
//   i: sign bit flipped, f: the float transform of sort_flt.inl:44-50,
//   char[16]: bytes ^ 0x80) and its bytes are laid out big-endian, columns in
//   order;
//   total width W <= 4: a u32 composite, W <= 8: a u64 composite (numeric
//   order = byte order), W <= 16: a 16-byte composite on the char[16] path,
//   whose load maps each byte b to b ^ 0x80, so the composite stores the
//   encoded bytes ^ 0x80 (padding 0x80, which the sort's plan skips).
// Group keys come back as composites and are unpacked per column.
#include "vdf_internal.h"

#include <algorithm>

namespace vdf {

namespace {

struct ColDesc {
    int type;   // VDF_KEY_*
    int width;  // bytes
    const void* data;
    void* out;  // group keys of this column (unpack)
};

struct PackArgs {
    ColDesc cols[VDF_MAX_KEY_COLS];
    int ncols;
    int comp;  // composite bytes: 4, 8 or 16
};

// order-preserving unsigned form of one cell, right-aligned in a u64
// (char[16] columns are handled byte-wise)
__device__ inline uint64_t enc_scalar(int type, const void* d, uint64_t r)
{
    switch (type) {
    case VDF_KEY_U32: return static_cast<const uint32_t*>(d)[r];
    case VDF_KEY_I32: return static_cast<const uint32_t*>(d)[r] ^ 0x80000000u;
    case VDF_KEY_F32: {
        const uint32_t x = static_cast<const uint32_t*>(d)[r];
        return x ^ ((x >> 31) ? 0xFFFFFFFFu : 0x80000000u);
    }
    case VDF_KEY_U64: return static_cast<const uint64_t*>(d)[r];
    case VDF_KEY_I64: return static_cast<const uint64_t*>(d)[r] ^ 0x8000000000000000ull;
    default: {  // VDF_KEY_F64
        const uint64_t x = static_cast<const uint64_t*>(d)[r];
        return x ^ ((x >> 63) ? ~0ull : 0x8000000000000000ull);
    }
    }
}

__device__ inline void dec_scalar(int type, uint64_t e, void* d, uint64_t g)
{
    switch (type) {
    case VDF_KEY_U32: static_cast<uint32_t*>(d)[g] = static_cast<uint32_t>(e); break;
    case VDF_KEY_I32: static_cast<uint32_t*>(d)[g] = static_cast<uint32_t>(e) ^ 0x80000000u; break;
    case VDF_KEY_F32: {
        const uint32_t x = static_cast<uint32_t>(e);
        static_cast<uint32_t*>(d)[g] = (x >> 31) ? (x ^ 0x80000000u) : ~x;
        break;
    }
    case VDF_KEY_U64: static_cast<uint64_t*>(d)[g] = e; break;
    case VDF_KEY_I64: static_cast<uint64_t*>(d)[g] = e ^ 0x8000000000000000ull; break;
    default: static_cast<uint64_t*>(d)[g] = (e >> 63) ? (e ^ 0x8000000000000000ull) : ~e; break;
    }
}

// Encoded big-endian bytes of one row into b[0 .. comp).
__device__ inline void encode_row(const PackArgs& a, uint64_t r, uint8_t* b)
{
    int pos = 0;
    for (int c = 0; c < a.ncols; ++c) {
        const ColDesc& cd = a.cols[c];
        if (cd.type == VDF_KEY_CHAR16) {
            const uint8_t* s = static_cast<const uint8_t*>(cd.data) + 16 * r;
            for (int i = 0; i < 16; ++i) b[pos + i] = s[i] ^ 0x80u;
        } else {
            const uint64_t e = enc_scalar(cd.type, cd.data, r);
            for (int i = 0; i < cd.width; ++i) b[pos + i] = static_cast<uint8_t>(e >> (8 * (cd.width - 1 - i)));
        }
        pos += cd.width;
    }
    for (; pos < a.comp; ++pos) b[pos] = 0;
}

__global__ __launch_bounds__(256) void k_pack(PackArgs a, uint64_t n, void* comp)
{
    const uint64_t r = static_cast<uint64_t>(blockIdx.x) * 256 + threadIdx.x;
    if (r >= n) return;
    uint8_t b[16];
    encode_row(a, r, b);
    if (a.comp == 16) {  // char[16] path: the sort maps each byte back with ^ 0x80
        uint64_t w0 = 0, w1 = 0;
        for (int i = 0; i < 8; ++i) {
            w0 |= static_cast<uint64_t>(b[i] ^ 0x80u) << (8 * i);
            w1 |= static_cast<uint64_t>(b[8 + i] ^ 0x80u) << (8 * i);
        }
        static_cast<uint64_t*>(comp)[2 * r] = w0;
        static_cast<uint64_t*>(comp)[2 * r + 1] = w1;
    } else {
        uint64_t v = 0;
        for (int i = 0; i < a.comp; ++i) v = (v << 8) | b[i];
        if (a.comp == 8) static_cast<uint64_t*>(comp)[r] = v;
        else static_cast<uint32_t*>(comp)[r] = static_cast<uint32_t>(v);
    }
}

// Group g's composite key -> each column's group key (g < min(G, cap)).
__global__ __launch_bounds__(256) void k_unpack(PackArgs a, const void* comp, const uint64_t* n_groups, uint64_t cap)
{
    const uint64_t g = static_cast<uint64_t>(blockIdx.x) * 256 + threadIdx.x;
    const uint64_t G = std::min<uint64_t>(*n_groups, cap);
    if (g >= G) return;
    uint8_t b[16];
    if (a.comp == 16) {
        const uint64_t w0 = static_cast<const uint64_t*>(comp)[2 * g];
        const uint64_t w1 = static_cast<const uint64_t*>(comp)[2 * g + 1];
        for (int i = 0; i < 8; ++i) {
            b[i] = static_cast<uint8_t>(w0 >> (8 * i)) ^ 0x80u;
            b[8 + i] = static_cast<uint8_t>(w1 >> (8 * i)) ^ 0x80u;
        }
    } else {
        const uint64_t v = a.comp == 8 ? static_cast<const uint64_t*>(comp)[g]
                                       : static_cast<const uint32_t*>(comp)[g];
        for (int i = 0; i < a.comp; ++i) b[i] = static_cast<uint8_t>(v >> (8 * (a.comp - 1 - i)));
    }
    int pos = 0;
    for (int c = 0; c < a.ncols; ++c) {
        const ColDesc& cd = a.cols[c];
        if (cd.out) {
            if (cd.type == VDF_KEY_CHAR16) {
                uint8_t* d = static_cast<uint8_t*>(cd.out) + 16 * g;
                for (int i = 0; i < 16; ++i) d[i] = b[pos + i] ^ 0x80u;
            } else {
                uint64_t e = 0;
                for (int i = 0; i < cd.width; ++i) e = (e << 8) | b[pos + i];
                dec_scalar(cd.type, e, cd.out, g);
            }
        }
        pos += cd.width;
    }
}

}  // namespace

int composite_key_type(const int* types, int ncols)
{
    size_t w = 0;
    for (int c = 0; c < ncols; ++c) w += key_bytes(types[c]);
    return w <= 4 ? VDF_KEY_U32 : (w <= 8 ? VDF_KEY_U64 : (w <= 16 ? VDF_KEY_CHAR16 : -1));
}

void pack_keys_enqueue(Handle* h, const int* types, const void* const* data, int ncols, uint64_t n, void* comp,
                       hipStream_t s)
{
    PackArgs a{};
    a.ncols = ncols;
    a.comp = static_cast<int>(key_bytes(composite_key_type(types, ncols)));
    uint64_t in_bytes = 0;
    for (int c = 0; c < ncols; ++c) {
        a.cols[c] = ColDesc{types[c], static_cast<int>(key_bytes(types[c])), data[c], nullptr};
        in_bytes += key_bytes(types[c]);
    }
    if (n == 0) return;
    LaunchTimer t(h, s, "pack_keys", n * (in_bytes + a.comp));
    hipLaunchKernelGGL(k_pack, dim3(static_cast<unsigned>((n + 255) / 256)), dim3(256), 0, s, a, n, comp);
    check_launch("k_pack");
}

void unpack_keys_enqueue(Handle* h, const int* types, void* const* outs, int ncols, const void* comp,
                         const uint64_t* n_groups, uint64_t cap, hipStream_t s)
{
    PackArgs a{};
    a.ncols = ncols;
    a.comp = static_cast<int>(key_bytes(composite_key_type(types, ncols)));
    for (int c = 0; c < ncols; ++c)
        a.cols[c] = ColDesc{types[c], static_cast<int>(key_bytes(types[c])), nullptr, outs ? outs[c] : nullptr};
    if (cap == 0) return;
    LaunchTimer t(h, s, "unpack_keys", 0);
    hipLaunchKernelGGL(k_unpack, dim3(static_cast<unsigned>((cap + 255) / 256)), dim3(256), 0, s, a, comp, n_groups,
                       cap);
    check_launch("k_unpack");
}

}  // namespace vdf
