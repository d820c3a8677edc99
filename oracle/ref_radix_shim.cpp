// ref_radix_shim.cpp -- exposes the reference's own scalar radix sorts through
// a C ABI so tests/golden/make_golden.py can record golden permutations.
//
// TEST INFRASTRUCTURE ONLY.  It is compiled by oracle/Makefile (target `ref`)
// with -I pointing at the reference's include/vison directory, only where that
// tree exists; the result goes to oracle/_ref/ and is never part of the product.
// No reference source is copied: the .inl files are included where they lie.
//
// The reference templates are instantiated with Simd=false (the scalar loops);
// the survey found the AVX-512 u32 scatter and the multithreaded SIMD paths
// wrong or racy, so they are not built.
#include <algorithm>
#include <array>
#include <cstddef>
#include <cstdint>
#include <cstring>
#include <iostream>
#include <numeric>
#include <immintrin.h>
#include <omp.h>
#include <string>
#include <utility>
#include <vector>

namespace vref {
#include "inlines/warning.inl"
#include "inlines/radix_sorts/details/constants.inl"
#include "inlines/radix_sorts/details/u16/get_local_u16.inl"
// the SIMD histogram helpers the radix templates name (declared before them,
// as vison.hpp does; the Simd = false paths used here do not call them)
#include "inlines/radix_sorts/details/u8/get_local_u8.inl"
#include "inlines/radix_sorts/details/u8/histogram_pass_u8_avx2.inl"
#include "inlines/radix_sorts/details/u8/histogram_pass_u8_avx2_8x.inl"
#include "inlines/radix_sorts/details/u32/histogram_pass_u32_avx2_8buckets.inl"
#include "inlines/radix_sorts/details/u32/histogram_pass_u32_avx2.inl"
#include "inlines/radix_sorts/details/u64/histogram_pass_u64_avx2.inl"
#include "inlines/radix_sorts/details/u64/histogram_pass_u64_avx2_8x.inl"
#include "inlines/radix_sorts/radix_sort_uint32.inl"
#include "inlines/radix_sorts/radix_sort_uint32_mt.inl"
#include "inlines/radix_sorts/radix_sort_uint64.inl"
#include "inlines/radix_sorts/radix_sort_uint64_mt.inl"
#include "inlines/radix_sorts/radix_sort_charbuf_flat.inl"
}  // namespace vref

extern "C" {

// radix_sort_uint32<false>(tkeys, idx, n); idx must hold the incoming order.
void vref_radix_sort_u32(uint32_t* keys, size_t* idx, size_t n)
{
    std::vector<uint32_t> t(keys, keys + n);
    vref::radix_sort_uint32<false>(t, idx, n);
    std::memcpy(keys, t.data(), n * sizeof(uint32_t));
}

void vref_radix_sort_u32_mt8(uint32_t* keys, size_t* idx, size_t n)
{
    std::vector<uint32_t> t(keys, keys + n);
    vref::radix_sort_uint32_mt<8, false>(t, idx, n);
    std::memcpy(keys, t.data(), n * sizeof(uint32_t));
}

// the reference's multithreaded path at 16 threads (CORES <= MAX_THREADS = 32):
// the CPU baseline of the benchmark
void vref_radix_sort_u32_mt16(uint32_t* keys, size_t* idx, size_t n)
{
    std::vector<uint32_t> t(keys, keys + n);
    vref::radix_sort_uint32_mt<16, false>(t, idx, n);
    std::memcpy(keys, t.data(), n * sizeof(uint32_t));
}

// at the reference's thread cap (MAX_THREADS = 32, details/constants.inl:9)
void vref_radix_sort_u32_mt32(uint32_t* keys, size_t* idx, size_t n)
{
    std::vector<uint32_t> t(keys, keys + n);
    vref::radix_sort_uint32_mt<32, false>(t, idx, n);
    std::memcpy(keys, t.data(), n * sizeof(uint32_t));
}

void vref_radix_sort_u64(uint64_t* keys, size_t* idx, size_t n)
{
    std::vector<uint64_t> t(keys, keys + n);
    vref::radix_sort_uint64<false>(t, idx, n);
    std::memcpy(keys, t.data(), n * sizeof(uint64_t));
}

void vref_radix_sort_u64_mt8(uint64_t* keys, size_t* idx, size_t n)
{
    std::vector<uint64_t> t(keys, keys + n);
    vref::radix_sort_uint64_mt<8, false>(t, idx, n);
    std::memcpy(keys, t.data(), n * sizeof(uint64_t));
}

// radix_sort_charbuf_flat<false>(bufv, n, width, idx); bufv already transformed
void vref_radix_sort_charbuf(const uint8_t* bufv, size_t n, size_t width, size_t* idx)
{
    vref::radix_sort_charbuf_flat<false>(bufv, n, width, idx);
}

}  // extern "C"
