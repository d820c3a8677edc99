// csv_parse.h -- text-to-number conversion of one CSV field, for the CSV
// ingest kernels (csv.hip) and the CPU unit test (tests/csv_parse_test.cpp).
//
// Semantics follow the reference's conversions (inlines/classify_column.inl:
// std::from_chars for integers, fast_float::from_chars for doubles): a field
// converts from its longest valid prefix; a field with no valid prefix, or an
// integer out of range, converts to 0.
//
// Doubles are correctly rounded (round to nearest, ties to even), as
// fast_float guarantees.  A decimal with at most 15 significant digits and
// |exponent| <= 22 takes one exact double operation (Clinger's fast path);
// any other takes a double estimate refined by exact comparisons of the
// decimal value with the halfway points between the estimate and its
// neighbours, in big integers of W 32-bit words.  The digits are read from
// the field's text, every one of them (no truncation), so the comparison is
// exact; when W words are too few (long digit strings, extreme exponents)
// the conversion reports it and the caller retries with a larger W.
//
// VDF_HD marks the functions for host and device when the includer defines it
// (csv.hip: __host__ __device__); the test includes the header as plain C++.
#pragma once

#include <cstdint>
#include <cstring>

#ifndef VDF_HD
#define VDF_HD
#endif

namespace vdf {
namespace csv {

template <int W>
struct Big {
    uint32_t w[W];
    int n;  // words in use (w[n-1] != 0 unless n == 0)
};

template <int W>
VDF_HD inline void big_set(Big<W>& b, uint64_t v)
{
    b.n = 0;
    while (v) {
        b.w[b.n++] = static_cast<uint32_t>(v);
        v >>= 32;
    }
}

// b = b * m + a (m, a < 2^32); false when the result needs more than W words
template <int W>
VDF_HD inline bool big_muladd(Big<W>& b, uint32_t m, uint32_t a)
{
    uint64_t carry = a;
    for (int i = 0; i < b.n; ++i) {
        const uint64_t t = static_cast<uint64_t>(b.w[i]) * m + carry;
        b.w[i] = static_cast<uint32_t>(t);
        carry = t >> 32;
    }
    if (carry) {
        if (b.n == W) return false;
        b.w[b.n++] = static_cast<uint32_t>(carry);
    }
    return true;
}

template <int W>
VDF_HD inline bool big_pow5(Big<W>& b, int e)
{
    while (e >= 13) {  // 5^13 is the largest power of five below 2^32
        if (!big_muladd(b, 1220703125u, 0)) return false;
        e -= 13;
    }
    uint32_t p = 1;
    while (e-- > 0) p *= 5;
    return big_muladd(b, p, 0);
}

template <int W>
VDF_HD inline bool big_shl(Big<W>& b, int s)
{
    if (b.n == 0 || s == 0) return true;
    const int ws = s / 32, bs = s % 32;
    const int nn = b.n + ws + (bs ? 1 : 0);
    if (nn > W) return false;
    for (int i = nn - 1; i >= 0; --i) {
        const int src = i - ws;
        const uint32_t hi = (src >= 0 && src < b.n) ? b.w[src] : 0u;
        const uint32_t lo = (src >= 1 && src - 1 < b.n) ? b.w[src - 1] : 0u;
        b.w[i] = bs ? ((hi << bs) | (lo >> (32 - bs))) : hi;
    }
    b.n = nn;
    while (b.n && b.w[b.n - 1] == 0) --b.n;
    return true;
}

template <int W>
VDF_HD inline int big_cmp(const Big<W>& a, const Big<W>& b)
{
    if (a.n != b.n) return a.n < b.n ? -1 : 1;
    for (int i = a.n - 1; i >= 0; --i)
        if (a.w[i] != b.w[i]) return a.w[i] < b.w[i] ? -1 : 1;
    return 0;
}

VDF_HD inline bool is_digit(char c) { return c >= '0' && c <= '9'; }
VDF_HD inline char lower(char c) { return (c >= 'A' && c <= 'Z') ? static_cast<char>(c + 32) : c; }

// A decimal as it lies in the text: value = D * 10^exp10, where D is the
// integer of the nd significant digits read from text[first ..] skipping the
// one decimal point.
struct Decimal {
    const char* text;
    int first;  // index of the first significant digit
    int nd;     // significant digits (trailing zeros dropped)
    int exp10;
};

VDF_HD inline double make_double(uint64_t bits)
{
    double x;
    __builtin_memcpy(&x, &bits, 8);
    return x;
}
VDF_HD inline uint64_t bits_of(double x)
{
    uint64_t b;
    __builtin_memcpy(&b, &x, 8);
    return b;
}

// D as a big integer
template <int W>
VDF_HD inline bool decimal_digits(const Decimal& dec, Big<W>& L)
{
    L.n = 0;
    int k = 0;
    for (int i = dec.first; k < dec.nd; ++i) {
        const char c = dec.text[i];
        if (c == '.') continue;
        if (!big_muladd(L, 10, static_cast<uint32_t>(c - '0'))) return false;
        ++k;
    }
    return true;
}

// Sign of D * 10^exp10 - (2m + 1) * 2^(e - 1), the value against the point
// halfway between m * 2^e and (m + 1) * 2^e.  False when W words are too few.
template <int W>
VDF_HD inline bool cmp_halfway(const Decimal& dec, uint64_t m, int e, int& out)
{
    Big<W> L, R;
    if (!decimal_digits(dec, L)) return false;
    big_set(R, 2 * m + 1);
    const int q = dec.exp10;
    int s = e - 1 - q;
    // q >= 0:  D * 5^q * 2^q  vs (2m+1) * 2^(e-1)  ->  D * 5^q vs (2m+1) * 2^s
    // q <  0:  D * 2^q / 5^-q vs (2m+1) * 2^(e-1)  ->  D       vs (2m+1) * 5^-q * 2^s
    if (q >= 0) {
        if (!big_pow5(L, q)) return false;
    } else {
        if (!big_pow5(R, -q)) return false;
    }
    if (s >= 0) {
        if (!big_shl(R, s)) return false;
    } else {
        if (!big_shl(L, -s)) return false;
    }
    out = big_cmp(L, R);
    return true;
}

// Correctly rounded binary64 magnitude of the decimal (nd >= 1); ok false
// when the big integers of W words are too small for the exact comparisons.
template <int W>
VDF_HD inline double decimal_to_double(const Decimal& dec, bool& ok)
{
    ok = true;
    const int top = dec.exp10 + dec.nd;  // value in [10^(top-1), 10^top)
    if (top > 310) return make_double(0x7FF0000000000000ull);
    if (top < -324) return 0.0;
    // leading digits as a double (exact up to 15 digits)
    const int nl = dec.nd < 17 ? dec.nd : 17;
    double lead = 0.0;
    {
        int k = 0;
        for (int i = dec.first; k < nl; ++i) {
            if (dec.text[i] == '.') continue;
            lead = lead * 10.0 + (dec.text[i] - '0');
            ++k;
        }
    }
    if (dec.nd <= 15 && dec.exp10 >= -22 && dec.exp10 <= 22) {  // Clinger: one exact operation
        double p = 1.0;
        const int aq = dec.exp10 < 0 ? -dec.exp10 : dec.exp10;
        for (int i = 0; i < aq; ++i) p *= 10.0;  // 10^k is exact up to 1e22
        return dec.exp10 < 0 ? lead / p : lead * p;
    }
    // estimate (a few ulps off), then exact refinement
    int pe = dec.exp10 + (dec.nd - nl);
    double est = lead;
    while (pe > 0) {
        const int k = pe > 22 ? 22 : pe;
        double p = 1.0;
        for (int i = 0; i < k; ++i) p *= 10.0;
        est *= p;
        pe -= k;
    }
    while (pe < 0) {
        const int k = -pe > 22 ? 22 : -pe;
        double p = 1.0;
        for (int i = 0; i < k; ++i) p *= 10.0;
        est /= p;
        pe += k;
    }
    uint64_t b = bits_of(est) & 0x7FFFFFFFFFFFFFFFull;
    if (b >= 0x7FF0000000000000ull) b = 0x7FEFFFFFFFFFFFFFull;  // start from the largest finite
    // b is the answer when the value lies between the halfway points below and
    // above it (ties to the even significand)
    for (int iter = 0; iter < 256; ++iter) {
        const uint64_t ef = b >> 52, fr = b & 0xFFFFFFFFFFFFFull;
        const uint64_t m = ef ? (fr | (1ull << 52)) : fr;
        const int e = ef ? static_cast<int>(ef) - 1075 : -1074;
        int cu = 0;
        if (!cmp_halfway<W>(dec, m, e, cu)) {
            ok = false;
            return make_double(b);
        }
        if (cu > 0 || (cu == 0 && (m & 1))) {  // above the upper halfway, or a tie to the even neighbour
            if (b == 0x7FEFFFFFFFFFFFFFull) return make_double(0x7FF0000000000000ull);  // rounds to infinity
            ++b;
            continue;
        }
        if (b == 0) return 0.0;
        // halfway down to the neighbour below; below a power of two (not the
        // smallest normal) the spacing halves
        uint64_t ml = m - 1;
        int el = e;
        if (ef > 1 && fr == 0) {
            ml = 2 * m - 1;
            el = e - 1;
        }
        int cl = 0;
        if (!cmp_halfway<W>(dec, ml, el, cl)) {
            ok = false;
            return make_double(b);
        }
        if (cl < 0 || (cl == 0 && (m & 1))) {
            --b;
            continue;
        }
        return make_double(b);
    }
    ok = false;
    return make_double(b);
}

// Result of a double conversion
enum { kParsed = 0, kNoNumber = 1, kNeedsWider = 2 };

// fast_float::from_chars of p[0 .. n): the longest valid prefix (optional
// '-', digits with an optional '.', optional exponent; "inf", "infinity",
// "nan" in any case).  status kNoNumber (value 0) when no prefix is valid,
// kNeedsWider when W words were too few (retry with a larger W).
template <int W>
VDF_HD inline double parse_double(const char* p, int n, int& status)
{
    status = kParsed;
    int i = 0;
    bool neg = false;
    if (i < n && p[i] == '-') {
        neg = true;
        ++i;
    }
    if (i < n && (lower(p[i]) == 'i' || lower(p[i]) == 'n')) {
        if (i + 3 <= n && lower(p[i]) == 'n' && lower(p[i + 1]) == 'a' && lower(p[i + 2]) == 'n')
            return make_double(neg ? 0xFFF8000000000000ull : 0x7FF8000000000000ull);
        if (i + 3 <= n && lower(p[i]) == 'i' && lower(p[i + 1]) == 'n' && lower(p[i + 2]) == 'f')
            return make_double(neg ? 0xFFF0000000000000ull : 0x7FF0000000000000ull);
        status = kNoNumber;
        return 0.0;
    }
    Decimal dec{p, -1, 0, 0};
    int sig = 0;         // significant digits so far (trailing zeros included)
    int nz = 0;          // trailing zeros among them
    int exp_adj = 0;     // power of ten of the digit after the last counted one
    bool any = false, dot = false;
    for (; i < n; ++i) {
        const char c = p[i];
        if (c == '.') {
            if (dot) break;
            dot = true;
            continue;
        }
        if (!is_digit(c)) break;
        any = true;
        if (dec.first < 0 && c == '0') {  // leading zeros
            if (dot) --exp_adj;
            continue;
        }
        if (dec.first < 0) dec.first = i;
        ++sig;
        nz = c == '0' ? nz + 1 : 0;
        if (dot) --exp_adj;
    }
    if (!any) {
        status = kNoNumber;
        return 0.0;
    }
    if (i < n && (p[i] == 'e' || p[i] == 'E')) {
        int j = i + 1;
        bool eneg = false;
        if (j < n && (p[j] == '-' || p[j] == '+')) {
            eneg = p[j] == '-';
            ++j;
        }
        if (j < n && is_digit(p[j])) {
            int64_t ev = 0;
            while (j < n && is_digit(p[j])) {
                if (ev < 100000000) ev = ev * 10 + (p[j] - '0');
                ++j;
            }
            exp_adj += static_cast<int>(eneg ? -ev : ev);
        }
    }
    if (sig == 0) return neg ? -0.0 : 0.0;
    dec.nd = sig - nz;  // trailing zeros carry no digits
    dec.exp10 = exp_adj + nz;
    bool ok = true;
    const double x = decimal_to_double<W>(dec, ok);
    if (!ok) status = kNeedsWider;
    return neg ? -x : x;
}

// The same conversion when it needs no big integers: fields without a
// number, zero, inf and nan, and decimals of at most 15 significant digits
// with |exponent| <= 22 (one exact operation).  Any other decimal reports
// kNeedsWider (the caller converts it with parse_double).
VDF_HD inline double parse_double_fast(const char* p, int n, int& status)
{
    status = kParsed;
    int i = 0;
    bool neg = false;
    if (i < n && p[i] == '-') {
        neg = true;
        ++i;
    }
    if (i < n && (lower(p[i]) == 'i' || lower(p[i]) == 'n')) {
        if (i + 3 <= n && lower(p[i]) == 'n' && lower(p[i + 1]) == 'a' && lower(p[i + 2]) == 'n')
            return make_double(neg ? 0xFFF8000000000000ull : 0x7FF8000000000000ull);
        if (i + 3 <= n && lower(p[i]) == 'i' && lower(p[i + 1]) == 'n' && lower(p[i + 2]) == 'f')
            return make_double(neg ? 0xFFF0000000000000ull : 0x7FF0000000000000ull);
        status = kNoNumber;
        return 0.0;
    }
    uint64_t d = 0;      // the significant digits read, trailing zeros included (16 at most)
    int sig = 0, nz = 0, exp_adj = 0;
    bool any = false, dot = false, started = false;
    for (; i < n; ++i) {
        const char c = p[i];
        if (c == '.') {
            if (dot) break;
            dot = true;
            continue;
        }
        if (!is_digit(c)) break;
        any = true;
        if (!started && c == '0') {
            if (dot) --exp_adj;
            continue;
        }
        started = true;
        ++sig;
        nz = c == '0' ? nz + 1 : 0;
        if (sig <= 16) d = d * 10 + static_cast<uint64_t>(c - '0');
        if (dot) --exp_adj;
    }
    if (!any) {
        status = kNoNumber;
        return 0.0;
    }
    if (i < n && (p[i] == 'e' || p[i] == 'E')) {
        int j = i + 1;
        bool eneg = false;
        if (j < n && (p[j] == '-' || p[j] == '+')) {
            eneg = p[j] == '-';
            ++j;
        }
        if (j < n && is_digit(p[j])) {
            int64_t ev = 0;
            while (j < n && is_digit(p[j])) {
                if (ev < 100000000) ev = ev * 10 + (p[j] - '0');
                ++j;
            }
            exp_adj += static_cast<int>(eneg ? -ev : ev);
        }
    }
    if (sig == 0) return neg ? -0.0 : 0.0;
    const int nd = sig - nz;
    const int e10 = exp_adj + nz;
    if (sig > 16 || nd > 15 || e10 < -22 || e10 > 22) {
        status = kNeedsWider;
        return 0.0;
    }
    for (int k = 0; k < nz; ++k) d /= 10;  // drop the trailing zeros (exact)
    double p10 = 1.0;
    const int aq = e10 < 0 ? -e10 : e10;
    for (int k = 0; k < aq; ++k) p10 *= 10.0;  // 10^k is exact up to 1e22
    const double lead = static_cast<double>(d);  // exact: at most 15 digits
    const double x = e10 < 0 ? lead / p10 : lead * p10;
    return neg ? -x : x;
}

// std::from_chars of a signed / unsigned 64-bit integer (base 10): optional
// '-' (signed only), digits; 0 when there are no digits or the value is out
// of range.
VDF_HD inline int64_t parse_int64(const char* p, int n)
{
    int i = 0;
    bool neg = false;
    if (i < n && p[i] == '-') {
        neg = true;
        ++i;
    }
    if (i >= n || !is_digit(p[i])) return 0;
    uint64_t v = 0;
    const uint64_t lim = neg ? 9223372036854775808ull : 9223372036854775807ull;
    while (i < n && is_digit(p[i])) {
        const uint64_t d = static_cast<uint64_t>(p[i] - '0');
        if (v > (lim - d) / 10) return 0;  // out of range
        v = v * 10 + d;
        ++i;
    }
    return neg ? static_cast<int64_t>(0 - v) : static_cast<int64_t>(v);
}

VDF_HD inline uint64_t parse_uint64(const char* p, int n)
{
    if (n <= 0 || !is_digit(p[0])) return 0;
    uint64_t v = 0;
    for (int i = 0; i < n && is_digit(p[i]); ++i) {
        const uint64_t d = static_cast<uint64_t>(p[i] - '0');
        if (v > (18446744073709551615ull - d) / 10) return 0;
        v = v * 10 + d;
    }
    return v;
}

}  // namespace csv
}  // namespace vdf
