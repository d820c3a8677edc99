// csv.hip -- CSV ingest on the device: delimited text already in HBM becomes
// typed device columns.
//
// Replaces the reference's read_csv (read/read_csv/read_csv.hpp:9-216): a
// header line of names (:58-96), rows split at newlines (simd_count_newlines,
// :138-146), fields split at the delimiter outside quote pairs of
// str_context (standard_parser/parse_rows_chunk), column types inferred from
// the first 10 rows (inlines/classify_column.inl:13-130, type letters
// s c b i u d of type_classification.inl) and every field converted with
// std::from_chars / fast_float::from_chars semantics (csv_parse.h).
//
// Launches (one stream; the host reads back only counts):
//   k_nl_count   per 64 KiB chunk: newlines (16-byte vector loads)
//   k_nl_write   line starts at each chunk's offset (host scan of the counts)
//   k_csv_measure  one thread per row: field count, longest field of each
//                  text column
//   k_csv_convert  one thread per row: every field into its column; doubles
//                  that need wide big integers are queued
//   k_csv_slow   one thread per queued field: the wide conversion
// Type inference reads the header and first rows on the host (they are a
// few hundred bytes): a column is numeric when each of its first 10 fields
// is a number ([-]digits[.digits][e[+-]digits]); double if one of them has a
// '.' or an exponent, int if one starts with '-', bool if all are 0 / 1,
// else uint; other columns are fixed-width char columns when their first 10
// fields fit char_width bytes, else strings (padded with 0x00 to the longest
// field of the column).  The reference's own number test (simd_can_be_nb,
// inlines/simd_can_be_nb.inl:1-41) admits only digits and one '.', so a '-'
// or an exponent makes its column text and its int branch
// (classify_column.inl:49-60) is never taken; here signed and exponent
// numbers are numbers.
#define VDF_HD __host__ __device__
#include "csv_parse.h"
#include "vdf_internal.h"

#include <algorithm>
#include <cstring>
#include <string>
#include <vector>

namespace vdf {
namespace {

constexpr uint64_t kChunk = 65536;  // bytes of text per block of the newline kernels
constexpr int kSmallWords = 24;     // big integers of the row kernel (768 bits)
constexpr int kWideWords = 360;     // the queued conversions (11520 bits)

struct CsvCol {
    int kind;
    uint32_t width;  // bytes per cell
    void* out;
};

struct CsvArgs {
    const char* text;
    uint64_t bytes;
    const uint64_t* line;  // line starts; row r spans [line[r], line[r + 1] - 1), minus a '\r'
    uint64_t nrow;
    int ncol;
    char delim, quote, trailing;
    const CsvCol* cols;
    uint32_t* maxlen;              // [ncol] longest field of each column
    unsigned long long* bad_row;   // first row whose field count differs (else ~0)
    uint32_t* slow_count;          // queued exact conversions
    unsigned long long* slow_need; // fields the measure pass found needing the exact conversion
    uint64_t* slow;                // [slow_cap] (row << 16 | column)
    uint64_t slow_cap;
    uint32_t* failed;              // conversions beyond the wide big integers
};

__global__ __launch_bounds__(256) void k_nl_count(const char* text, uint64_t bytes, uint32_t* counts)
{
    __shared__ uint32_t s_sum[4];
    const uint64_t base = static_cast<uint64_t>(blockIdx.x) * kChunk + static_cast<uint64_t>(threadIdx.x) * 256;
    uint32_t c = 0;
    if (base + 256 <= bytes) {
        const auto* q = reinterpret_cast<const uint4*>(text + base);
        for (int i = 0; i < 16; ++i) {
            const uint4 v = q[i];
            const uint32_t w[4] = {v.x, v.y, v.z, v.w};
            for (int k = 0; k < 4; ++k)
                for (int b = 0; b < 4; ++b) c += ((w[k] >> (8 * b)) & 0xFFu) == '\n';
        }
    } else {
        for (uint64_t i = base; i < base + 256 && i < bytes; ++i) c += text[i] == '\n';
    }
    for (int off = 32; off > 0; off >>= 1) c += __shfl_xor(c, off, 64);
    if ((threadIdx.x & 63) == 0) s_sum[threadIdx.x >> 6] = c;
    __syncthreads();
    if (threadIdx.x == 0) counts[blockIdx.x] = s_sum[0] + s_sum[1] + s_sum[2] + s_sum[3];
}

// line[1 + k] = (position of the k-th newline) + 1.  A thread's 256 bytes
// are read as 16 16-byte vectors (one read; every newline found in
// registers), as k_nl_count reads them; the text's last partial 256 bytes
// byte by byte.  The block scan of the counts is outside both branches: a
// wave can hold threads of each, and the scan's shuffles and barriers need
// every thread of the block.
__global__ __launch_bounds__(256) void k_nl_write(const char* text, uint64_t bytes, const uint64_t* block_off,
                                                  uint64_t* line)
{
    __shared__ uint32_t s_wsum[4];
    const uint64_t base = static_cast<uint64_t>(blockIdx.x) * kChunk + static_cast<uint64_t>(threadIdx.x) * 256;
    const bool full = base + 256 <= bytes;
    // bit j of m[i]: byte 16 i + j of the thread's 256 bytes is a newline
    uint32_t m[16];
    uint32_t c = 0;
    if (full) {
        const auto* q = reinterpret_cast<const uint4*>(text + base);
        uint4 v[16];
#pragma unroll
        for (int i = 0; i < 16; ++i) v[i] = q[i];
#pragma unroll
        for (int i = 0; i < 16; ++i) {
            const uint32_t w[4] = {v[i].x, v[i].y, v[i].z, v[i].w};
            uint32_t bits = 0;
#pragma unroll
            for (int k = 0; k < 4; ++k)
#pragma unroll
                for (int b = 0; b < 4; ++b) bits |= (((w[k] >> (8 * b)) & 0xFFu) == '\n' ? 1u : 0u) << (4 * k + b);
            m[i] = bits;
            c += static_cast<uint32_t>(__popc(bits));
        }
    } else {
        for (uint64_t i = base; i < base + 256 && i < bytes; ++i) c += text[i] == '\n';
    }
    const uint32_t pre = block_excl_scan(c, s_wsum);
    uint64_t k = block_off[blockIdx.x] + pre;
    if (full) {
#pragma unroll
        for (int i = 0; i < 16; ++i) {
            uint32_t bits = m[i];
            while (bits) {
                const int j = __ffs(bits) - 1;
                bits &= bits - 1;
                line[1 + k++] = base + 16 * i + j + 1;
            }
        }
    } else {
        for (uint64_t i = base; i < base + 256 && i < bytes; ++i)
            if (text[i] == '\n') line[1 + k++] = i + 1;
    }
}

// Walks row r's fields column by column: f(column, begin, length) for
// columns 0 .. ncol-1 that the row has; returns the row's field count (ncol
// + 1 when it has more).  The loop over columns is the same for every lane
// of a wave (one row per lane), so each column's conversion runs once for
// all lanes together; a lane only scans ahead to its next delimiter.  Text
// byte i is src[i - off] (the text itself, or a copy in LDS).  A delimiter
// inside a pair of quote characters does not split; a field never ends
// inside quotes, so the quote state restarts at each field.
template <typename F>
__device__ inline int for_fields(const CsvArgs& a, const char* src, uint64_t off, uint64_t r, F f)
{
    uint64_t b = a.line[r], e = a.line[r + 1] - 1;  // e: the newline (or one past the text)
    if (e > a.bytes) e = a.bytes;
    if (e > b && src[e - 1 - off] == '\r') --e;
    uint64_t pos = b;
    int nf = 0;
    bool more = true;  // the row has a field starting at pos
    for (int c = 0; c < a.ncol; ++c) {
        if (more) {
            uint64_t i = pos;
            bool inq = false;
            for (; i < e; ++i) {
                const char ch = src[i - off];
                if (ch == a.quote) inq = !inq;
                else if (ch == a.delim && !inq) break;
            }
            f(c, pos, i - pos);
            ++nf;
            more = i < e;  // ended at a delimiter: another field follows
            pos = i + 1;
        }
    }
    return more ? nf + 1 : nf;
}

// Rows of one wave (lane = row) are parsed from a copy of their text in
// LDS: the wave copies the span of its 64 rows with coalesced 4-byte loads,
// so the per-row walks read LDS instead of issuing one scattered byte load
// per character.  A span longer than kWaveText bytes is read from memory.
constexpr uint32_t kWaveText = 16384;

// Copies the text of rows [r0, r1) into buf; returns the text index of
// buf[0], or ~0 when the span does not fit (every lane gets the same answer).
__device__ inline uint64_t stage_rows(const CsvArgs& a, uint64_t r0, uint64_t r1, char* buf)
{
    if (r0 >= r1) return ~0ull;
    const int lane = threadIdx.x & 63;
    const uint64_t b = a.line[r0];
    uint64_t e = a.line[r1];  // one past the last row's newline (a sentinel past an unterminated one)
    if (e > a.bytes) e = a.bytes;
    const uint64_t a0 = b & ~3ull;
    const uint64_t nw = (e - a0 + 3) / 4;
    if (nw * 4 > kWaveText) return ~0ull;
    auto* bw = reinterpret_cast<uint32_t*>(buf);
    const auto* tw = reinterpret_cast<const uint32_t*>(a.text + a0);  // the text is 16-byte aligned
    for (uint64_t w = static_cast<uint64_t>(lane); w < nw; w += 64) {
        const uint64_t gi = a0 + 4 * w;
        if (gi + 4 <= a.bytes) {
            bw[w] = tw[w];
        } else {  // the text's last bytes: no read past its end
            uint32_t x = 0;
            for (uint64_t k = 0; k < 4 && gi + k < a.bytes; ++k)
                x |= static_cast<uint32_t>(static_cast<uint8_t>(a.text[gi + k])) << (8 * k);
            bw[w] = x;
        }
    }
    return a0;
}

// Longest field of each text column: maxima kept in LDS per block (tables
// of up to kMeasureCols columns), then one global atomic per column and
// block; wider tables take a global atomic per field.
constexpr int kMeasureCols = 1024;

// Also counts the double fields that need the exact conversion (the queue
// of k_csv_slow is sized from it).
template <bool STAGE>
__global__ __launch_bounds__(256) void k_csv_measure(CsvArgs a)
{
    __shared__ uint32_t s_max[kMeasureCols];
    __shared__ uint32_t s_slow;
    __shared__ __align__(16) char s_txt[STAGE ? 4 : 1][STAGE ? kWaveText : 16];
    const bool local = a.ncol <= kMeasureCols;
    if (local)
        for (int c = threadIdx.x; c < a.ncol; c += 256) s_max[c] = 0;
    if (threadIdx.x == 0) s_slow = 0;
    const uint64_t r = static_cast<uint64_t>(blockIdx.x) * 256 + threadIdx.x;
    const int wave = threadIdx.x >> 6;
    const uint64_t w0 = static_cast<uint64_t>(blockIdx.x) * 256 + static_cast<uint64_t>(wave) * 64;
    uint64_t off = ~0ull;
    if constexpr (STAGE) off = stage_rows(a, w0, std::min<uint64_t>(w0 + 64, a.nrow), s_txt[wave]);
    __syncthreads();
    const char* src = off == ~0ull ? a.text : s_txt[STAGE ? wave : 0];
    const uint64_t base = off == ~0ull ? 0 : off;
    uint32_t slow = 0;
    if (r < a.nrow) {
        const int nf = for_fields(a, src, base, r, [&](int c, uint64_t fb, uint64_t len) {
            if (c >= a.ncol) return;
            const int kind = a.cols[c].kind;
            if (kind == VDF_CSV_STRING) {
                const uint32_t l = static_cast<uint32_t>(std::min<uint64_t>(len, 0xFFFFFFFFull));
                if (local) atomicMax(s_max + c, l);
                else atomicMax(a.maxlen + c, l);
            } else if (kind == VDF_CSV_DOUBLE) {
                int st = 0;
                (void)csv::parse_double_fast(src + (fb - base), static_cast<int>(std::min<uint64_t>(len, 0x7FFFFFFFull)), st);
                slow += st == csv::kNeedsWider ? 1u : 0u;
            }
        });
        if (nf != a.ncol) atomicMin(a.bad_row, static_cast<unsigned long long>(r));
    }
    if (slow) atomicAdd(&s_slow, slow);
    __syncthreads();
    if (local)
        for (int c = threadIdx.x; c < a.ncol; c += 256)
            if (s_max[c]) atomicMax(a.maxlen + c, s_max[c]);
    if (threadIdx.x == 0 && s_slow) atomicAdd(a.slow_need, static_cast<unsigned long long>(s_slow));
}

__device__ inline void put_int(void* out, uint64_t r, uint32_t width, uint64_t v)
{
    if (width == 8) static_cast<uint64_t*>(out)[r] = v;
    else static_cast<uint32_t*>(out)[r] = static_cast<uint32_t>(v);
}

// Every field into its column; doubles outside the fast conversion are
// queued for k_csv_slow (no big integers here, so no scratch memory).
template <bool STAGE>
__global__ __launch_bounds__(256) void k_csv_convert(CsvArgs a)
{
    __shared__ __align__(16) char s_txt[STAGE ? 4 : 1][STAGE ? kWaveText : 16];
    const uint64_t r = static_cast<uint64_t>(blockIdx.x) * 256 + threadIdx.x;
    const int wave = threadIdx.x >> 6;
    const uint64_t w0 = static_cast<uint64_t>(blockIdx.x) * 256 + static_cast<uint64_t>(wave) * 64;
    uint64_t off = ~0ull;
    if constexpr (STAGE) {
        off = stage_rows(a, w0, std::min<uint64_t>(w0 + 64, a.nrow), s_txt[wave]);
        __syncthreads();
    }
    if (r >= a.nrow) return;
    const char* src = off == ~0ull ? a.text : s_txt[STAGE ? wave : 0];
    const uint64_t base = off == ~0ull ? 0 : off;
    for_fields(a, src, base, r, [&](int c, uint64_t fb, uint64_t len) {
        if (c >= a.ncol) return;
        const CsvCol col = a.cols[c];
        const char* s = src + (fb - base);
        const int n = static_cast<int>(std::min<uint64_t>(len, 0x7FFFFFFFull));
        switch (col.kind) {
        case VDF_CSV_DOUBLE: {
            int st = 0;
            double x = csv::parse_double_fast(s, n, st);
            if (st == csv::kNeedsWider) {
                const uint64_t k = atomicAdd(reinterpret_cast<unsigned long long*>(a.slow_count), 1ull);
                if (k < a.slow_cap) a.slow[k] = (r << 16) | static_cast<uint64_t>(c);
                x = 0.0;
            } else if (st == csv::kNoNumber) {
                x = 0.0;
            }
            static_cast<double*>(col.out)[r] = x;
            break;
        }
        case VDF_CSV_INT: put_int(col.out, r, col.width, static_cast<uint64_t>(csv::parse_int64(s, n))); break;
        case VDF_CSV_UINT: put_int(col.out, r, col.width, csv::parse_uint64(s, n)); break;
        case VDF_CSV_BOOL: static_cast<uint32_t*>(col.out)[r] = csv::parse_int64(s, n) != 0 ? 1u : 0u; break;
        default: {  // CHAR (padded with the trailing char) and STRING (padded with 0x00)
            char* d = static_cast<char*>(col.out) + r * col.width;
            const char pad = col.kind == VDF_CSV_CHAR ? a.trailing : '\0';
            const uint32_t m = static_cast<uint32_t>(std::min<uint64_t>(len, col.width));
            for (uint32_t i = 0; i < m; ++i) d[i] = s[i];
            for (uint32_t i = m; i < col.width; ++i) d[i] = pad;
            break;
        }
        }
    });
}

// The queued doubles: the exact conversion in big integers of W words.  The
// first pass (kSmallWords) marks the entries it cannot finish (bit 63) and
// counts them in *wide; the second (kWideWords, launched only when some are
// marked) converts the marked ones and counts failures in *failed.
template <int W, bool SECOND>
__global__ __launch_bounds__(64) void k_csv_slow(CsvArgs a, uint64_t count, uint32_t* wide)
{
    const uint64_t k = static_cast<uint64_t>(blockIdx.x) * 64 + threadIdx.x;
    if (k >= count) return;
    const uint64_t q = a.slow[k];
    if (SECOND && !(q >> 63)) return;
    const uint64_t r = (q & ~(1ull << 63)) >> 16;
    const int want = static_cast<int>(q & 0xFFFFu);
    for_fields(a, a.text, 0, r, [&](int c, uint64_t fb, uint64_t len) {
        if (c != want) return;
        int st = 0;
        const double x = csv::parse_double<W>(a.text + fb, static_cast<int>(std::min<uint64_t>(len, 0x7FFFFFFFull)), st);
        if (st == csv::kNeedsWider) {
            if (SECOND) {
                atomicAdd(a.failed, 1u);
            } else {
                a.slow[k] = q | (1ull << 63);
                atomicAdd(wide, 1u);
            }
        }
        static_cast<double*>(a.cols[c].out)[r] = st == csv::kParsed ? x : 0.0;
    });
}

// ---- host-side inference over the first rows ----
bool is_number(const std::string& s, bool& has_dot_or_exp, bool& neg)
{
    size_t i = 0;
    neg = false;
    has_dot_or_exp = false;
    if (i < s.size() && s[i] == '-') {
        neg = true;
        ++i;
    }
    size_t digits = 0;
    while (i < s.size() && csv::is_digit(s[i])) ++i, ++digits;
    if (i < s.size() && s[i] == '.') {
        has_dot_or_exp = true;
        ++i;
        while (i < s.size() && csv::is_digit(s[i])) ++i, ++digits;
    }
    if (digits == 0) return false;
    if (i < s.size() && (s[i] == 'e' || s[i] == 'E')) {
        has_dot_or_exp = true;
        ++i;
        if (i < s.size() && (s[i] == '+' || s[i] == '-')) ++i;
        size_t ed = 0;
        while (i < s.size() && csv::is_digit(s[i])) ++i, ++ed;
        if (ed == 0) return false;
    }
    return i == s.size();
}

std::vector<std::string> split_line(const std::string& line, char delim, char quote)
{
    std::vector<std::string> out;
    bool inq = false;
    size_t fs = 0;
    for (size_t i = 0; i < line.size(); ++i) {
        if (line[i] == quote) inq = !inq;
        else if (line[i] == delim && !inq) {
            out.push_back(line.substr(fs, i - fs));
            fs = i + 1;
        }
    }
    out.push_back(line.substr(fs));
    return out;
}

}  // namespace
}  // namespace vdf

using namespace vdf;

struct vdf_csv_table {
    uint64_t nrow = 0;
    std::vector<vdf_csv_column> cols;
    std::vector<std::string> names;
    std::vector<void*> bufs;
    int device = 0;
    ~vdf_csv_table()
    {
        for (void* p : bufs)
            if (p) (void)hipFree(p);
    }
};

extern "C" {

vdf_status vdf_read_csv(vdf_handle* h, const void* d_text, uint64_t bytes, const vdf_csv_options* opt,
                        vdf_csv_table** out)
{
    return guarded([&] {
        require(h != nullptr && opt != nullptr && out != nullptr, "vdf_read_csv: null argument");
        *out = nullptr;
        require(bytes == 0 || d_text != nullptr, "vdf_read_csv: null text");
        require(reinterpret_cast<uintptr_t>(d_text) % 16 == 0, "vdf_read_csv: the text must be 16-byte aligned");
        require(opt->int_bytes == 4 || opt->int_bytes == 8, "vdf_read_csv: int_bytes must be 4 or 8");
        require(opt->char_width >= 0 && opt->char_width <= VDF_MAX_STRING_BYTES,
                "vdf_read_csv: char_width out of range");
        require(opt->delim != '\n' && opt->delim != '\r', "vdf_read_csv: the delimiter cannot be a line break");
        std::lock_guard<std::mutex> lk(h->mu);
        int prev = 0;
        VDF_HIP_CHECK(hipGetDevice(&prev));
        VDF_HIP_CHECK(hipSetDevice(h->device));
        struct Restore {
            int d;
            ~Restore() { (void)hipSetDevice(d); }
        } restore{prev};
        auto T = std::make_unique<vdf_csv_table>();
        T->device = h->device;
        const char* text = static_cast<const char*>(d_text);

        // ---- header and the first data rows, on the host ----
        std::vector<std::string> lines;
        {
            std::vector<char> pre;
            uint64_t take = std::min<uint64_t>(bytes, 1 << 16);
            for (;;) {
                pre.resize(take);
                if (take) VDF_HIP_CHECK(hipMemcpy(pre.data(), text, take, hipMemcpyDeviceToHost));
                size_t nl = 0;
                for (char ch : pre) nl += ch == '\n';
                if (nl >= 11 + (opt->header ? 1u : 0u) || take == bytes) break;
                take = std::min<uint64_t>(bytes, take * 4);
            }
            size_t fs = 0;
            for (size_t i = 0; i < pre.size(); ++i)
                if (pre[i] == '\n') {
                    std::string l(pre.data() + fs, i - fs);
                    if (!l.empty() && l.back() == '\r') l.pop_back();
                    lines.push_back(l);
                    fs = i + 1;
                    if (lines.size() >= 12) break;
                }
            if (lines.size() < 12 && fs < pre.size() && pre.size() == bytes) {
                std::string l(pre.data() + fs, pre.size() - fs);
                if (!l.empty() && l.back() == '\r') l.pop_back();
                lines.push_back(l);
            }
        }
        if (lines.empty()) {
            *out = T.release();
            return;
        }
        std::vector<std::string> head = split_line(lines[0], opt->delim, opt->quote);
        const int ncol = static_cast<int>(head.size());
        require(ncol <= 65535, "vdf_read_csv: more than 65535 columns");
        for (int c = 0; c < ncol; ++c)
            T->names.push_back(opt->header ? (head[c].empty() ? std::string("NA") : head[c]) : std::string());
        const size_t first = opt->header ? 1 : 0;
        std::vector<int> kind(static_cast<size_t>(ncol), VDF_CSV_UINT);
        for (int c = 0; c < ncol; ++c) {
            bool numeric = true, any_dot = false, any_neg = false, boolean = true, fits = true;
            size_t seen = 0;
            for (size_t l = first; l < lines.size() && seen < 10; ++l, ++seen) {
                const std::vector<std::string> f = split_line(lines[l], opt->delim, opt->quote);
                const std::string v = c < static_cast<int>(f.size()) ? f[c] : std::string();
                bool dot = false, neg = false;
                if (!is_number(v, dot, neg)) numeric = false;
                any_dot |= dot;
                any_neg |= neg;
                if (v != "0" && v != "1") boolean = false;
                if (v.size() > static_cast<size_t>(opt->char_width)) fits = false;
            }
            if (!numeric) kind[c] = (fits && opt->char_width > 0) ? VDF_CSV_CHAR : VDF_CSV_STRING;
            else if (any_dot) kind[c] = VDF_CSV_DOUBLE;
            else if (any_neg) kind[c] = VDF_CSV_INT;
            else if (boolean && seen > 0) kind[c] = VDF_CSV_BOOL;
            else kind[c] = VDF_CSV_UINT;
        }

        // ---- line starts on the device ----
        const uint64_t nblk = std::max<uint64_t>((bytes + kChunk - 1) / kChunk, 1);
        std::vector<uint32_t> counts(nblk, 0);
        std::vector<uint64_t> boff(nblk, 0);
        void* d_counts = nullptr;
        void* d_boff = nullptr;
        VDF_HIP_CHECK(hipMalloc(&d_counts, nblk * 4));
        T->bufs.push_back(d_counts);
        VDF_HIP_CHECK(hipMalloc(&d_boff, nblk * 8));
        T->bufs.push_back(d_boff);
        if (bytes) {
            {
                LaunchTimer lt(h, nullptr, "csv_newlines", bytes);
                hipLaunchKernelGGL(k_nl_count, dim3(static_cast<unsigned>(nblk)), dim3(256), 0, nullptr, text,
                                   bytes, static_cast<uint32_t*>(d_counts));
                check_launch("k_nl_count");
            }
            VDF_HIP_CHECK(hipMemcpy(counts.data(), d_counts, nblk * 4, hipMemcpyDeviceToHost));
        }
        uint64_t nl = 0;
        for (uint64_t b = 0; b < nblk; ++b) {
            boff[b] = nl;
            nl += counts[b];
        }
        // line starts: line[0] = 0, line[1 + k] after the k-th newline, a
        // sentinel one past the text for an unterminated last line
        uint64_t last_nl_end = 0;
        void* d_line = nullptr;
        VDF_HIP_CHECK(hipMalloc(&d_line, (nl + 2) * 8));
        T->bufs.push_back(d_line);
        const uint64_t zero = 0;
        VDF_HIP_CHECK(hipMemcpy(d_line, &zero, 8, hipMemcpyHostToDevice));
        if (nl) {
            VDF_HIP_CHECK(hipMemcpy(d_boff, boff.data(), nblk * 8, hipMemcpyHostToDevice));
            {
                LaunchTimer lt(h, nullptr, "csv_lines", bytes + nl * 8);
                hipLaunchKernelGGL(k_nl_write, dim3(static_cast<unsigned>(nblk)), dim3(256), 0, nullptr, text,
                                   bytes, static_cast<const uint64_t*>(d_boff), static_cast<uint64_t*>(d_line));
                check_launch("k_nl_write");
            }
            VDF_HIP_CHECK(hipMemcpy(&last_nl_end, static_cast<uint64_t*>(d_line) + nl, 8, hipMemcpyDeviceToHost));
        }
        uint64_t nlines = nl;
        if (last_nl_end < bytes) {  // an unterminated last line
            const uint64_t s = bytes + 1;
            VDF_HIP_CHECK(hipMemcpy(static_cast<uint64_t*>(d_line) + nl + 1, &s, 8, hipMemcpyHostToDevice));
            ++nlines;
        }
        const uint64_t nrow = nlines > first ? nlines - first : 0;
        require(nrow <= VDF_MAX_ROWS, "vdf_read_csv: more rows than VDF_MAX_ROWS");
        T->nrow = nrow;

        // ---- columns ----
        void* d_cols = nullptr;
        void* d_maxlen = nullptr;
        void* d_misc = nullptr;  // bad_row, slow_count, failed
        VDF_HIP_CHECK(hipMalloc(&d_cols, std::max(ncol, 1) * sizeof(CsvCol)));
        T->bufs.push_back(d_cols);
        VDF_HIP_CHECK(hipMalloc(&d_maxlen, std::max(ncol, 1) * 4));
        T->bufs.push_back(d_maxlen);
        VDF_HIP_CHECK(hipMalloc(&d_misc, 64));
        T->bufs.push_back(d_misc);
        std::vector<CsvCol> cols(static_cast<size_t>(ncol));
        for (int c = 0; c < ncol; ++c) {
            cols[c].kind = kind[c];
            cols[c].width = kind[c] == VDF_CSV_DOUBLE ? 8u
                            : (kind[c] == VDF_CSV_INT || kind[c] == VDF_CSV_UINT) ? static_cast<uint32_t>(opt->int_bytes)
                            : kind[c] == VDF_CSV_BOOL ? 4u
                            : kind[c] == VDF_CSV_CHAR ? static_cast<uint32_t>(opt->char_width)
                                                      : 0u;
            cols[c].out = nullptr;
        }
        CsvArgs a{};
        a.text = text;
        a.bytes = bytes;
        a.line = static_cast<const uint64_t*>(d_line) + first;
        a.nrow = nrow;
        a.ncol = ncol;
        a.delim = opt->delim;
        a.quote = opt->quote;
        a.trailing = opt->trailing;
        a.cols = static_cast<const CsvCol*>(d_cols);
        a.maxlen = static_cast<uint32_t*>(d_maxlen);
        a.bad_row = static_cast<unsigned long long*>(d_misc);
        a.slow_count = reinterpret_cast<uint32_t*>(static_cast<char*>(d_misc) + 8);  // 64-bit counter
        a.failed = reinterpret_cast<uint32_t*>(static_cast<char*>(d_misc) + 16);
        a.slow_need = reinterpret_cast<unsigned long long*>(static_cast<char*>(d_misc) + 24);
        const unsigned long long none = ~0ull;
        VDF_HIP_CHECK(hipMemcpy(a.bad_row, &none, 8, hipMemcpyHostToDevice));
        VDF_HIP_CHECK(hipMemset(static_cast<char*>(d_misc) + 8, 0, 24));
        VDF_HIP_CHECK(hipMemset(d_maxlen, 0, std::max(ncol, 1) * 4));
        VDF_HIP_CHECK(hipMemcpy(d_cols, cols.data(), ncol * sizeof(CsvCol), hipMemcpyHostToDevice));
        const unsigned rgrid = static_cast<unsigned>((nrow + 255) / 256);
        const bool stage = h->tune.csv_stage != 0;
        if (nrow) {
            LaunchTimer lt(h, nullptr, "csv_measure", bytes + (nrow + 1) * 8);
            if (stage) hipLaunchKernelGGL(k_csv_measure<true>, dim3(rgrid), dim3(256), 0, nullptr, a);
            else hipLaunchKernelGGL(k_csv_measure<false>, dim3(rgrid), dim3(256), 0, nullptr, a);
            check_launch("k_csv_measure");
        }
        unsigned long long bad = none;
        std::vector<uint32_t> maxlen(static_cast<size_t>(ncol), 0);
        VDF_HIP_CHECK(hipMemcpy(&bad, a.bad_row, 8, hipMemcpyDeviceToHost));
        if (ncol) VDF_HIP_CHECK(hipMemcpy(maxlen.data(), d_maxlen, ncol * 4, hipMemcpyDeviceToHost));
        if (bad != none)
            throw Error{VDF_ERR_INVALID, "vdf_read_csv: data row " + std::to_string(bad) + " does not have " +
                                             std::to_string(ncol) + " fields"};
        for (int c = 0; c < ncol; ++c) {
            if (kind[c] == VDF_CSV_STRING) {
                require(maxlen[c] <= VDF_MAX_STRING_BYTES, "vdf_read_csv: a text field longer than VDF_MAX_STRING_BYTES");
                cols[c].width = std::max<uint32_t>(maxlen[c], 1);
            }
            void* p = nullptr;
            VDF_HIP_CHECK(hipMalloc(&p, std::max<uint64_t>(nrow, 1) * cols[c].width));
            T->bufs.push_back(p);
            cols[c].out = p;
        }
        VDF_HIP_CHECK(hipMemcpy(d_cols, cols.data(), ncol * sizeof(CsvCol), hipMemcpyHostToDevice));
        unsigned long long need = 0;
        VDF_HIP_CHECK(hipMemcpy(&need, a.slow_need, 8, hipMemcpyDeviceToHost));
        a.slow_cap = std::max<uint64_t>(need, 1);
        void* d_slow = nullptr;
        VDF_HIP_CHECK(hipMalloc(&d_slow, a.slow_cap * 8ull));
        T->bufs.push_back(d_slow);
        a.slow = static_cast<uint64_t*>(d_slow);
        if (nrow) {
            uint64_t out_bytes = 0;
            for (int c = 0; c < ncol; ++c) out_bytes += nrow * cols[c].width;
            LaunchTimer lt(h, nullptr, "csv_convert", bytes + (nrow + 1) * 8 + out_bytes);
            if (stage) hipLaunchKernelGGL(k_csv_convert<true>, dim3(rgrid), dim3(256), 0, nullptr, a);
            else hipLaunchKernelGGL(k_csv_convert<false>, dim3(rgrid), dim3(256), 0, nullptr, a);
            check_launch("k_csv_convert");
        }
        uint64_t nslow = 0;
        VDF_HIP_CHECK(hipMemcpy(&nslow, a.slow_count, 8, hipMemcpyDeviceToHost));
        require(nslow == need, "vdf_read_csv: the queued exact conversions differ from the count");
        if (nslow) {
            uint32_t* wide = reinterpret_cast<uint32_t*>(static_cast<char*>(d_misc) + 20);
            const dim3 sg(static_cast<unsigned>((nslow + 63) / 64));
            {
                LaunchTimer lt(h, nullptr, "csv_slow", nslow * 16);
                hipLaunchKernelGGL((k_csv_slow<kSmallWords, false>), sg, dim3(64), 0, nullptr, a, nslow, wide);
                check_launch("k_csv_slow");
            }
            uint32_t nwide = 0;
            VDF_HIP_CHECK(hipMemcpy(&nwide, wide, 4, hipMemcpyDeviceToHost));
            if (nwide) hipLaunchKernelGGL((k_csv_slow<kWideWords, true>), sg, dim3(64), 0, nullptr, a, nslow, wide);
            check_launch("k_csv_slow");
            uint32_t failed = 0;
            VDF_HIP_CHECK(hipMemcpy(&failed, a.failed, 4, hipMemcpyDeviceToHost));
            require(failed == 0, "vdf_read_csv: a number with too many digits to round exactly");
        }
        VDF_HIP_CHECK(hipDeviceSynchronize());
        for (int c = 0; c < ncol; ++c) {
            vdf_csv_column vc{};
            vc.kind = cols[c].kind;
            vc.width = cols[c].width;
            vc.data = cols[c].out;
            T->cols.push_back(vc);
        }
        for (int c = 0; c < ncol; ++c) T->cols[c].name = T->names[c].c_str();
        *out = T.release();
    });
}

vdf_status vdf_csv_info(vdf_csv_table* t, uint64_t* nrow, int* ncol)
{
    return guarded([&] {
        require(t != nullptr, "vdf_csv_info: null table");
        if (nrow) *nrow = t->nrow;
        if (ncol) *ncol = static_cast<int>(t->cols.size());
    });
}

vdf_status vdf_csv_column_get(vdf_csv_table* t, int c, vdf_csv_column* out)
{
    return guarded([&] {
        require(t != nullptr && out != nullptr, "vdf_csv_column_get: null argument");
        require(c >= 0 && c < static_cast<int>(t->cols.size()), "vdf_csv_column_get: no such column");
        *out = t->cols[static_cast<size_t>(c)];
    });
}

vdf_status vdf_csv_free(vdf_csv_table* t)
{
    return guarded([&] {
        if (!t) return;
        int prev = 0;
        VDF_HIP_CHECK(hipGetDevice(&prev));
        VDF_HIP_CHECK(hipSetDevice(t->device));
        delete t;
        (void)hipSetDevice(prev);
    });
}

}  // extern "C"
