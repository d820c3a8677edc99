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
//   encoded bytes ^ 0x80 (padding 0x80, which the sort's plan skips);
//   W > 16: a W-byte row for the byte-string sort and group-by, which order
//   rows by each byte ^ 0x80 as well, so the row stores the encoded bytes
//   ^ 0x80 too.
// Byte-string columns (VDF_KEY_BYTES, padded std::string cells) encode as
// char[16] does, byte by byte with ^ 0x80 (sort_string.inl:10-85 order).
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
    int comp;  // composite bytes: 4, 8 or 16, or the row width W (> 16) of a wide composite
};

__device__ inline bool bytewise(int type) { return type == VDF_KEY_CHAR16 || type == VDF_KEY_BYTES; }

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
        if (bytewise(cd.type)) {
            const uint8_t* s = static_cast<const uint8_t*>(cd.data) + static_cast<uint64_t>(cd.width) * r;
            for (int i = 0; i < cd.width; ++i) b[pos + i] = s[i] ^ 0x80u;
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
            if (bytewise(cd.type)) {
                uint8_t* d = static_cast<uint8_t*>(cd.out) + static_cast<uint64_t>(cd.width) * g;
                for (int i = 0; i < cd.width; ++i) d[i] = b[pos + i] ^ 0x80u;
            } else {
                uint64_t e = 0;
                for (int i = 0; i < cd.width; ++i) e = (e << 8) | b[pos + i];
                dec_scalar(cd.type, e, cd.out, g);
            }
        }
        pos += cd.width;
    }
}

// Wide composite (W > 16 bytes): row r's encoded bytes ^ 0x80, one thread per
// row, written byte by byte (the byte-string sort maps each back with ^ 0x80).
__global__ __launch_bounds__(256) void k_pack_wide(PackArgs a, uint64_t n, uint8_t* comp)
{
    const uint64_t r = static_cast<uint64_t>(blockIdx.x) * 256 + threadIdx.x;
    if (r >= n) return;
    uint8_t* o = comp + static_cast<uint64_t>(a.comp) * r;
    int pos = 0;
    for (int c = 0; c < a.ncols; ++c) {
        const ColDesc& cd = a.cols[c];
        if (bytewise(cd.type)) {
            const uint8_t* s = static_cast<const uint8_t*>(cd.data) + static_cast<uint64_t>(cd.width) * r;
            for (int i = 0; i < cd.width; ++i) o[pos + i] = s[i];  // (s ^ 0x80) ^ 0x80
        } else {
            const uint64_t e = enc_scalar(cd.type, cd.data, r);
            for (int i = 0; i < cd.width; ++i)
                o[pos + i] = static_cast<uint8_t>(e >> (8 * (cd.width - 1 - i))) ^ 0x80u;
        }
        pos += cd.width;
    }
}

// Group g's wide composite row -> each column's group key.
__global__ __launch_bounds__(256) void k_unpack_wide(PackArgs a, const uint8_t* comp, const uint64_t* n_groups,
                                                     uint64_t cap)
{
    const uint64_t g = static_cast<uint64_t>(blockIdx.x) * 256 + threadIdx.x;
    const uint64_t G = std::min<uint64_t>(*n_groups, cap);
    if (g >= G) return;
    const uint8_t* row = comp + static_cast<uint64_t>(a.comp) * g;
    int pos = 0;
    for (int c = 0; c < a.ncols; ++c) {
        const ColDesc& cd = a.cols[c];
        if (cd.out) {
            if (bytewise(cd.type)) {
                uint8_t* d = static_cast<uint8_t*>(cd.out) + static_cast<uint64_t>(cd.width) * g;
                for (int i = 0; i < cd.width; ++i) d[i] = row[pos + i];
            } else {
                uint64_t e = 0;
                for (int i = 0; i < cd.width; ++i) e = (e << 8) | static_cast<uint8_t>(row[pos + i] ^ 0x80u);
                dec_scalar(cd.type, e, cd.out, g);
            }
        }
        pos += cd.width;
    }
}

PackArgs pack_args(const int* types, const uint32_t* widths, int ncols)
{
    PackArgs a{};
    a.ncols = ncols;
    const uint32_t w = composite_width(types, widths, ncols);
    const int kt = composite_key_type(w);
    a.comp = kt >= 0 ? static_cast<int>(key_bytes(kt)) : static_cast<int>(w);
    for (int c = 0; c < ncols; ++c)
        a.cols[c] = ColDesc{types[c], static_cast<int>(key_col_width(types[c], widths ? widths[c] : 0)), nullptr,
                            nullptr};
    return a;
}

}  // namespace

uint32_t key_col_width(int type, uint32_t width)
{
    return type == VDF_KEY_BYTES ? width : static_cast<uint32_t>(key_bytes(type));
}

uint32_t composite_width(const int* types, const uint32_t* widths, int ncols)
{
    uint32_t w = 0;
    for (int c = 0; c < ncols; ++c) w += key_col_width(types[c], widths ? widths[c] : 0);
    return w;
}

int composite_key_type(uint32_t w)
{
    return w <= 4 ? VDF_KEY_U32 : (w <= 8 ? VDF_KEY_U64 : (w <= 16 ? VDF_KEY_CHAR16 : -1));
}

void pack_keys_enqueue(Handle* h, const int* types, const uint32_t* widths, const void* const* data, int ncols,
                       uint64_t n, void* comp, hipStream_t s)
{
    PackArgs a = pack_args(types, widths, ncols);
    uint64_t in_bytes = 0;
    for (int c = 0; c < ncols; ++c) {
        a.cols[c].data = data[c];
        in_bytes += static_cast<uint64_t>(a.cols[c].width);
    }
    if (n == 0) return;
    LaunchTimer t(h, s, "pack_keys", n * (in_bytes + static_cast<uint64_t>(a.comp)));
    const dim3 grid(static_cast<unsigned>((n + 255) / 256));
    if (a.comp > 16) hipLaunchKernelGGL(k_pack_wide, grid, dim3(256), 0, s, a, n, static_cast<uint8_t*>(comp));
    else hipLaunchKernelGGL(k_pack, grid, dim3(256), 0, s, a, n, comp);
    check_launch("k_pack");
}

void unpack_keys_enqueue(Handle* h, const int* types, const uint32_t* widths, void* const* outs, int ncols,
                         const void* comp, const uint64_t* n_groups, uint64_t cap, hipStream_t s)
{
    PackArgs a = pack_args(types, widths, ncols);
    for (int c = 0; c < ncols; ++c) a.cols[c].out = outs ? outs[c] : nullptr;
    if (cap == 0) return;
    LaunchTimer t(h, s, "unpack_keys", 0);
    const dim3 grid(static_cast<unsigned>((cap + 255) / 256));
    if (a.comp > 16)
        hipLaunchKernelGGL(k_unpack_wide, grid, dim3(256), 0, s, a, static_cast<const uint8_t*>(comp), n_groups, cap);
    else hipLaunchKernelGGL(k_unpack, grid, dim3(256), 0, s, a, comp, n_groups, cap);
    check_launch("k_unpack");
}

}  // namespace vdf
