// CPU unit test of the CSV field conversions (visondf_amd/csrc/csv_parse.h):
// doubles against glibc strtod (correctly rounded) on random and boundary
// decimal strings, integers against strtoll / strtoull with the range rules
// of std::from_chars.  Built with g++ by tests/test_csv_cpu.py.
#include "../visondf_amd/csrc/csv_parse.h"

#include <cerrno>
#include <cinttypes>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>
#include <string>

using namespace vdf::csv;

static int g_fail = 0;
static long g_checked = 0;

static double conv(const std::string& s, int& status)
{
    double x = parse_double<40>(s.data(), static_cast<int>(s.size()), status);
    if (status == kNeedsWider) x = parse_double<200>(s.data(), static_cast<int>(s.size()), status);
    return x;
}

static long g_fast = 0;

static void check_double(const std::string& s)
{
    int st = 0;
    const double got = conv(s, st);
    const double want = std::strtod(s.c_str(), nullptr);
    ++g_checked;
    // the big-integer-free conversion: where it answers, the same bits
    int fst = 0;
    const double fast = parse_double_fast(s.data(), static_cast<int>(s.size()), fst);
    if (fst == kParsed) {
        ++g_fast;
        if (bits_of(fast) != bits_of(got) && !(std::isnan(fast) && std::isnan(got))) {
            if (g_fail < 20) std::printf("FAIL fast double '%s': got %.17g want %.17g\n", s.c_str(), fast, got);
            ++g_fail;
        }
    } else if (fst != kNeedsWider || st == kNoNumber) {
        if (fst != st) {
            if (g_fail < 20) std::printf("FAIL fast status '%s': %d vs %d\n", s.c_str(), fst, st);
            ++g_fail;
        }
    }
    if (st != kParsed || bits_of(got) != bits_of(want)) {
        if (!(std::isnan(got) && std::isnan(want))) {
            if (g_fail < 20)
                std::printf("FAIL double '%s': got %.17g (%016" PRIx64 ") status %d, want %.17g (%016" PRIx64 ")\n",
                            s.c_str(), got, bits_of(got), st, want, bits_of(want));
            ++g_fail;
        }
    }
}

static std::string digits(std::mt19937_64& rng, int n)
{
    std::string s;
    for (int i = 0; i < n; ++i) s.push_back(static_cast<char>('0' + rng() % 10));
    return s;
}

int main()
{
    std::mt19937_64 rng(12345);
    // round trips of random doubles at 17 significant digits (what repr-style writers emit)
    char buf[64];
    for (int i = 0; i < 300000; ++i) {
        const uint64_t b = rng() & 0x7FEFFFFFFFFFFFFFull;
        double x;
        memcpy(&x, &b, 8);
        std::snprintf(buf, sizeof buf, "%.17g", x);
        check_double(buf);
        std::snprintf(buf, sizeof buf, "%.15g", x);
        check_double(buf);
        std::snprintf(buf, sizeof buf, "%.6e", x);
        check_double(buf);
    }
    // uniform [0, 1) values as the benchmark writes them
    std::uniform_real_distribution<double> u01(0.0, 1.0);
    for (int i = 0; i < 200000; ++i) {
        std::snprintf(buf, sizeof buf, "%.17g", u01(rng));
        check_double(buf);
    }
    // random digit strings: 1 .. 60 digits, a decimal point anywhere, exponents -360 .. 330
    for (int i = 0; i < 200000; ++i) {
        const int nd = 1 + static_cast<int>(rng() % 60);
        std::string s = digits(rng, nd);
        if (rng() % 2) s.insert(rng() % (s.size() + 1), ".");
        if (rng() % 3) s += "e" + std::to_string(static_cast<int>(rng() % 691) - 360);
        if (rng() % 2) s = "-" + s;
        check_double(s);
    }
    // halfway cases: exact decimal expansions of (2m + 1) * 2^(e - 1), and one digit either side
    for (int i = 0; i < 3000; ++i) {
        const uint64_t m = (rng() & ((1ull << 52) - 1)) | (1ull << 52);
        const int e = static_cast<int>(rng() % 120) - 60;
        // halfway = (2m + 1) * 2^(e-1): print exactly with long double is not enough; build in decimal
        std::string num = std::to_string(2 * m + 1);
        int p2 = e - 1;
        // multiply / divide the decimal string by powers of two exactly
        if (p2 > 0) {
            for (int k = 0; k < p2; ++k) {  // num *= 2
                int carry = 0;
                for (int j = static_cast<int>(num.size()) - 1; j >= 0; --j) {
                    const int d = (num[j] - '0') * 2 + carry;
                    num[j] = static_cast<char>('0' + d % 10);
                    carry = d / 10;
                }
                if (carry) num.insert(num.begin(), static_cast<char>('0' + carry));
            }
            check_double(num);
        } else {
            // num / 2^-p2 = num * 5^-p2 / 10^-p2
            for (int k = 0; k < -p2; ++k) {  // num *= 5
                int carry = 0;
                for (int j = static_cast<int>(num.size()) - 1; j >= 0; --j) {
                    const int d = (num[j] - '0') * 5 + carry;
                    num[j] = static_cast<char>('0' + d % 10);
                    carry = d / 10;
                }
                if (carry) num.insert(num.begin(), static_cast<char>('0' + carry));
            }
            const std::string s = num + "e-" + std::to_string(-p2);
            check_double(s);
            // nudged above and below the tie
            std::string up = num + "1e-" + std::to_string(-p2 + 1);
            check_double(up);
            std::string dn = num;
            int j = static_cast<int>(dn.size()) - 1;
            while (j >= 0 && dn[j] == '0') dn[j--] = '9';
            if (j >= 0) {
                dn[j] = static_cast<char>(dn[j] - 1);
                check_double(dn + "9e-" + std::to_string(-p2 + 1));
            }
        }
    }
    // boundaries: subnormals, the smallest normal, the largest finite, overflow
    const char* fixed[] = {"4.9406564584124654e-324", "2.4703282292062327e-324", "2.4703282292062328e-324",
                           "2.2250738585072011e-308", "2.2250738585072014e-308", "1.7976931348623157e308",
                           "1.7976931348623158e308", "1.7976931348623159e308", "1e309", "1e-400", "0.0",
                           "-0.0", "000123.4500e2", "1e23", "8.98846567431158e307", "9007199254740993",
                           "9007199254740992.5", "0.1", "3.14159", "123456789012345678901234567890",
                           "1.00000000000000011102230246251565404236316680908203125",
                           "1.00000000000000011102230246251565404236316680908203124",
                           "1.00000000000000011102230246251565404236316680908203126", ".5", "5.", "-.25e1"};
    for (const char* f : fixed) check_double(f);
    // prefixes and non-numbers
    struct {
        const char* s;
        double v;
        int st;
    } pre[] = {{"12abc", 12.0, kParsed}, {"1.5e", 1.5, kParsed}, {"1.5e+", 1.5, kParsed}, {"abc", 0.0, kNoNumber},
               {"-", 0.0, kNoNumber}, {".", 0.0, kNoNumber}, {"", 0.0, kNoNumber}, {"1.2.3", 1.2, kParsed},
               {"inf", INFINITY, kParsed}, {"-Infinity", -INFINITY, kParsed}};
    for (const auto& c : pre) {
        int fst = 0;
        const double fx = parse_double_fast(c.s, static_cast<int>(std::strlen(c.s)), fst);
        if (fst != c.st || bits_of(fx) != bits_of(c.v)) {
            std::printf("FAIL fast prefix '%s': got %g status %d\n", c.s, fx, fst);
            ++g_fail;
        }
        int st = 0;
        const double x = conv(c.s, st);
        if (st != c.st || bits_of(x) != bits_of(c.v)) {
            std::printf("FAIL prefix '%s': got %g status %d\n", c.s, x, st);
            ++g_fail;
        }
    }
    int st = 0;
    if (!std::isnan(conv("nan", st)) || st != kParsed) {
        std::printf("FAIL nan\n");
        ++g_fail;
    }
    // integers
    for (int i = 0; i < 200000; ++i) {
        std::string s = digits(rng, 1 + static_cast<int>(rng() % 21));
        if (rng() % 2) s = "-" + s;
        if (rng() % 5 == 0) s += "x7";
        errno = 0;
        char* end = nullptr;
        long long want = std::strtoll(s.c_str(), &end, 10);
        if (errno == ERANGE) want = 0;
        const int64_t got = parse_int64(s.data(), static_cast<int>(s.size()));
        if (got != want) {
            if (g_fail < 40) std::printf("FAIL int64 '%s': got %" PRId64 " want %lld\n", s.c_str(), got, want);
            ++g_fail;
        }
        uint64_t uwant = 0;
        if (s[0] != '-') {
            errno = 0;
            uwant = std::strtoull(s.c_str(), &end, 10);
            if (errno == ERANGE) uwant = 0;
        }
        const uint64_t ugot = parse_uint64(s.data(), static_cast<int>(s.size()));
        if (ugot != uwant) {
            if (g_fail < 40) std::printf("FAIL uint64 '%s': got %" PRIu64 " want %" PRIu64 "\n", s.c_str(), ugot, uwant);
            ++g_fail;
        }
    }
    if (parse_int64("-9223372036854775808", 20) != INT64_MIN || parse_int64("9223372036854775808", 19) != 0 ||
        parse_uint64("18446744073709551615", 20) != UINT64_MAX || parse_uint64("18446744073709551616", 20) != 0) {
        std::printf("FAIL integer limits\n");
        ++g_fail;
    }
    if (g_fail) {
        std::printf("%d failures of %ld doubles checked\n", g_fail, g_checked);
        return 1;
    }
    if (g_fast < 50000) {
        std::printf("FAIL only %ld doubles took the fast conversion\n", g_fast);
        return 1;
    }
    std::printf("csv_parse ok (%ld doubles, %ld by the fast conversion)\n", g_checked, g_fast);
    return 0;
}
