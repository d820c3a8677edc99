// GPU test program of the header-only C++ façade (include/vison_gpu.hpp):
// the reference-shaped Dataframe calls (add_col, sort_by<ASC, Soft>,
// materialize, group_by, transform_group_by) checked against host
// computations on the same seeded frame.  Built by visondf_amd/csrc/Makefile
// (target facade) and run by tests/test_facade_gpu.py.
#include "vison_gpu.hpp"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <map>
#include <numeric>
#include <random>
#include <unistd.h>

using vison_gpu::Char16;
using vison_gpu::Dataframe;
using vison_gpu::GroupFunction;

static int g_fail = 0;
#define CHECK(c)                                                             \
    do {                                                                     \
        if (!(c)) {                                                          \
            std::printf("FAIL %s:%d: %s\n", __FILE__, __LINE__, #c);         \
            ++g_fail;                                                        \
        }                                                                    \
    } while (0)

struct Frame {
    std::vector<uint32_t> k32;
    std::vector<int64_t> k64;
    std::vector<double> vf;
    std::vector<int32_t> vi;
    std::vector<Char16> kc;
};

static Frame make(size_t n, uint64_t seed)
{
    std::mt19937_64 rng(seed);
    Frame f;
    f.k32.resize(n);
    f.k64.resize(n);
    f.vf.resize(n);
    f.vi.resize(n);
    f.kc.resize(n);
    std::normal_distribution<double> nd(0.0, 1.0);
    for (size_t i = 0; i < n; ++i) {
        f.k32[i] = static_cast<uint32_t>(rng() % 5000) * 2654435761u;
        f.k64[i] = static_cast<int64_t>(rng() % 201) - 100;
        f.vf[i] = nd(rng);
        f.vi[i] = static_cast<int32_t>(rng() % 2000001) - 1000000;
        const uint64_t id = rng() % 700;
        for (int j = 0; j < 16; ++j) f.kc[i].c[j] = static_cast<char>(32 + (id * 2654435761ull >> (j * 3)) % 95);
    }
    return f;
}

static Dataframe load(const Frame& f)
{
    Dataframe df(0);
    df.add_col(f.k32, "k32");
    df.add_col(f.k64, "k64");
    df.add_col(f.vf, "vf");
    df.add_col(f.vi, "vi");
    df.add_col(f.kc, "kc");
    return df;
}

template <typename K>
static std::vector<size_t> stable_order(const std::vector<K>& k)
{
    std::vector<size_t> idx(k.size());
    std::iota(idx.begin(), idx.end(), 0);
    std::stable_sort(idx.begin(), idx.end(), [&](size_t a, size_t b) { return k[a] < k[b]; });
    return idx;
}

static bool same_bits(double a, double b) { return std::memcmp(&a, &b, 8) == 0; }

int main()
{
    std::setvbuf(stdout, nullptr, _IONBF, 0);
    const size_t n = 200003;
    const Frame f = make(n, 7);

    // hard ascending sort by the uint32 column: every column follows
    std::printf("sort_by<true, false> uint32\n");
    {
        Dataframe df = load(f);
        unsigned int c = 0;
        df.sort_by<true, false>(c);
        const auto idx = stable_order(f.k32);
        const auto k = df.get_col<uint32_t>(0);
        const auto vf = df.get_col<double>(2);
        const auto vi = df.get_col<int32_t>(3);
        bool ok = true;
        for (size_t i = 0; i < n && ok; ++i)
            ok = k[i] == f.k32[idx[i]] && same_bits(vf[i], f.vf[idx[i]]) && vi[i] == f.vi[idx[i]];
        CHECK(ok);
    }
    // hard descending sort by the int64 column: the reference's ascending-then-reverse order
    std::printf("sort_by<false, false> int64\n");
    {
        Dataframe df = load(f);
        unsigned int c = 1;
        df.sort_by<false, false>(c);
        auto idx = stable_order(f.k64);
        std::reverse(idx.begin(), idx.end());
        const auto vi = df.get_col<int32_t>(3);
        bool ok = true;
        for (size_t i = 0; i < n && ok; ++i) ok = vi[i] == f.vi[idx[i]];
        CHECK(ok);
    }
    // soft sort (row view) then materialize == hard sort
    std::printf("soft sort + materialize\n");
    {
        Dataframe a = load(f), b = load(f);
        unsigned int c = 4;
        a.sort_by<true, true>(c);
        CHECK(a.in_view());
        a.materialize();
        CHECK(!a.in_view());
        b.sort_by<true, false>(c);
        CHECK(a.get_col<double>(2) == b.get_col<double>(2));
    }
    // two soft sorts in a row: the second orders the viewed rows by its key,
    // stably, so the result is ordered by (int64 key, then the uint32 key's
    // order, then row); checked against the host's stable sorts composed the
    // same way and against a lexicographic stable sort
    std::printf("soft sort twice + materialize\n");
    {
        Dataframe df = load(f);
        unsigned int c0 = 0, c1 = 1;
        df.sort_by<true, true>(c0);
        df.sort_by<true, true>(c1);
        CHECK(df.in_view());
        const auto v1 = stable_order(f.k32);
        std::vector<int64_t> seen(n);
        for (size_t j = 0; j < n; ++j) seen[j] = f.k64[v1[j]];
        const auto q = stable_order(seen);
        std::vector<size_t> want(n);
        for (size_t j = 0; j < n; ++j) want[j] = v1[q[j]];
        const auto& view = df.get_row_view_idx();
        CHECK(view.size() == n && std::equal(view.begin(), view.end(), want.begin()));
        std::vector<size_t> lex(n);
        std::iota(lex.begin(), lex.end(), 0);
        std::stable_sort(lex.begin(), lex.end(), [&](size_t a, size_t b) {
            return f.k64[a] != f.k64[b] ? f.k64[a] < f.k64[b] : f.k32[a] < f.k32[b];
        });
        CHECK(want == lex);
        df.materialize();
        const auto k64 = df.get_col<int64_t>(1);
        const auto vi = df.get_col<int32_t>(3);
        bool ok = true;
        for (size_t i = 0; i < n && ok; ++i) ok = k64[i] == f.k64[want[i]] && vi[i] == f.vi[want[i]];
        CHECK(ok);
        // and a descending soft sort over a view: the reversed ascending order of the viewed keys
        Dataframe d2 = load(f);
        d2.sort_by<true, true>(c0);
        d2.sort_by<false, true>(c1);
        std::vector<size_t> rq = q;
        std::reverse(rq.begin(), rq.end());
        const auto& view2 = d2.get_row_view_idx();
        bool ok2 = view2.size() == n;
        for (size_t j = 0; j < n && ok2; ++j) ok2 = view2[j] == v1[rq[j]];
        CHECK(ok2);
    }
    // group_by on the int64 key: counts, integer sums, min / max exact; f64 within 1e-12
    std::printf("group_by int64\n");
    {
        Dataframe df = load(f);
        const auto R = df.group_by<int64_t>(1, {2, 3});
        struct Acc {
            uint64_t cnt = 0;
            long double sf = 0, sabs = 0;
            double mnf = 0, mxf = 0;
            int64_t si = 0, mni = 0, mxi = 0;
        };
        std::map<int64_t, Acc> m;
        for (size_t i = 0; i < n; ++i) {
            Acc& a = m[f.k64[i]];
            if (a.cnt == 0) {
                a.mnf = a.mxf = f.vf[i];
                a.mni = a.mxi = f.vi[i];
            }
            a.mnf = std::min(a.mnf, f.vf[i]);
            a.mxf = std::max(a.mxf, f.vf[i]);
            a.mni = std::min<int64_t>(a.mni, f.vi[i]);
            a.mxi = std::max<int64_t>(a.mxi, f.vi[i]);
            a.sf += f.vf[i];
            a.sabs += std::fabs(f.vf[i]);
            a.si += f.vi[i];
            ++a.cnt;
        }
        CHECK(R.n_groups == m.size());
        size_t g = 0;
        bool ok = true;
        for (const auto& [key, a] : m) {
            if (g >= R.n_groups) break;
            ok = ok && R.keys[g] == key && R.counts[g] == a.cnt;
            ok = ok && R.sum_i[1][g] == a.si && R.min_i[1][g] == a.mni && R.max_i[1][g] == a.mxi;
            ok = ok && R.min_f[0][g] == a.mnf && R.max_f[0][g] == a.mxf;
            const double exact = static_cast<double>(a.sf);
            // 1e-12 relative, plus the long-double reference's own rounding bound
            ok = ok && std::fabs(R.sum_f[0][g] - exact) <= 1e-12 * std::fabs(exact) + 1e-15 * static_cast<double>(a.sabs);
            ok = ok && std::fabs(R.mean[1][g] - static_cast<double>(a.si) / a.cnt) <=
                           1e-15 * std::fabs(static_cast<double>(a.si) / a.cnt);
            // the row view lists the group's rows in row order
            for (uint32_t p = R.offsets[g]; p < R.offsets[g + 1] && ok; ++p) {
                ok = f.k64[R.row_view[p]] == key;
                if (p > R.offsets[g]) ok = ok && R.row_view[p] > R.row_view[p - 1];
            }
            ++g;
        }
        CHECK(ok);
    }
    // transform_group_by: per-row group sum and count columns
    std::printf("transform_group_by\n");
    {
        Dataframe df = load(f);
        df.transform_group_by<GroupFunction::Sum>({0}, 3, "sum_vi");
        df.transform_group_by<GroupFunction::Occurence>({0}, 0, "count");
        CHECK(df.get_ncol() == 7);
        std::map<uint32_t, std::pair<int64_t, uint64_t>> m;
        for (size_t i = 0; i < n; ++i) {
            m[f.k32[i]].first += f.vi[i];
            m[f.k32[i]].second += 1;
        }
        const auto s = df.get_col<double>(5);
        const auto c = df.get_col<double>(6);
        bool ok = true;
        for (size_t i = 0; i < n && ok; ++i)
            ok = s[i] == static_cast<double>(m[f.k32[i]].first) && c[i] == static_cast<double>(m[f.k32[i]].second);
        CHECK(ok);
    }
    // two key columns: group_by_multi and a per-row transform column over (k64, k32)
    std::printf("group_by_multi + transform_group_by on two keys\n");
    {
        Dataframe df = load(f);
        const auto M = df.group_by_multi({1, 0}, {3});
        std::map<std::pair<int64_t, uint32_t>, std::pair<int64_t, uint64_t>> m;
        for (size_t i = 0; i < n; ++i) {
            auto& e = m[{f.k64[i], f.k32[i]}];
            e.first += f.vi[i];
            e.second += 1;
        }
        CHECK(M.r.n_groups == m.size());
        size_t g = 0;
        bool ok = true;
        for (const auto& [key, e] : m) {
            if (g >= M.r.n_groups) break;
            int64_t k64;
            uint32_t k32;
            std::memcpy(&k64, M.keys_of[0].data() + 8 * g, 8);
            std::memcpy(&k32, M.keys_of[1].data() + 4 * g, 4);
            ok = ok && k64 == key.first && k32 == key.second && M.r.counts[g] == e.second && M.r.sum_i[0][g] == e.first;
            ++g;
        }
        CHECK(ok);
        df.transform_group_by<GroupFunction::Mean>({1, 0}, 3, "mean_vi");
        const auto mc = df.get_col<double>(5);
        ok = true;
        for (size_t i = 0; i < n && ok; ++i) {
            const auto& e = m[{f.k64[i], f.k32[i]}];
            const double exp = static_cast<double>(e.first) / static_cast<double>(e.second);
            ok = std::fabs(mc[i] - exp) <= 1e-15 * std::fabs(exp);
        }
        CHECK(ok);
    }
    // 16-byte string keys: group count and ascending byte order
    std::printf("group_by char16\n");
    {
        Dataframe df = load(f);
        const auto R = df.group_by<Char16>(4, {});
        std::map<std::string, uint64_t> m;
        for (size_t i = 0; i < n; ++i) m[std::string(f.kc[i].c, 16)] += 1;
        CHECK(R.n_groups == m.size());
        size_t g = 0;
        bool ok = true;
        for (const auto& [key, cnt] : m) {
            if (g >= R.n_groups) break;
            ok = ok && std::memcmp(R.keys[g].c, key.data(), 16) == 0 && R.counts[g] == cnt;
            ++g;
        }
        CHECK(ok);
    }
    // float64 key column: the order of sort_flt.inl:44-50 (no NaN or signed zero here: numeric order)
    std::printf("sort_by<true, false> float64\n");
    {
        Dataframe df = load(f);
        unsigned int c = 2;
        df.sort_by<true, false>(c);
        const auto idx = stable_order(f.vf);
        const auto vi = df.get_col<int32_t>(3);
        bool ok = true;
        for (size_t i = 0; i < n && ok; ++i) ok = vi[i] == f.vi[idx[i]];
        CHECK(ok);
    }
    // std::string key column: padded with 0x00 to the longest, ordered bytewise
    // as signed chars (sort_string); ascending is stable, descending reverses it
    std::printf("sort_by std::string\n");
    {
        std::vector<std::string> s = {"pear", "apple", "", "apple pie", "Zebra", "apple", "\xc3\xa9" "clair", "pea",
                                      "pear", "a"};
        std::vector<int32_t> tag(s.size());
        for (size_t i = 0; i < s.size(); ++i) tag[i] = static_cast<int32_t>(i);
        auto key_of = [](const std::string& x) {
            std::vector<int> k;
            for (unsigned char ch : x) k.push_back(static_cast<signed char>(ch));
            return k;  // shorter prefix first: padding 0x00 is above the negative bytes only
        };
        std::vector<size_t> idx(s.size());
        for (size_t i = 0; i < idx.size(); ++i) idx[i] = i;
        const size_t w = 9;  // the longest string
        auto padded = [&](size_t i) {
            std::vector<int> k = key_of(s[i]);
            k.resize(w, 0);
            return k;
        };
        std::stable_sort(idx.begin(), idx.end(), [&](size_t a, size_t b) { return padded(a) < padded(b); });
        for (int mode = 0; mode < 2; ++mode) {
            Dataframe df;
            df.add_col(s, "s");
            df.add_col(tag, "tag");
            unsigned int c = 0;
            if (mode == 0) df.sort_by<true, false>(c);
            else df.sort_by<false, false>(c);
            const auto got = df.get_str_col(0);
            const auto t = df.get_col<int32_t>(1);
            bool ok = true;
            for (size_t i = 0; i < s.size() && ok; ++i) {
                const size_t e = mode == 0 ? idx[i] : idx[s.size() - 1 - i];
                ok = t[i] == static_cast<int32_t>(e) && got[i] == s[e];
            }
            CHECK(ok);
        }
    }
    // group-by over the std::string column: groups ascending, counts and
    // sums of the tag column, and a transform_group_by count broadcast per row
    std::printf("group_by_str\n");
    {
        std::vector<std::string> s = {"kiwi", "fig", "kiwi", "", "fig", "date palm", "kiwi", ""};
        std::vector<int64_t> v = {1, 2, 3, 4, 5, 6, 7, 8};
        Dataframe df;
        df.add_col(s, "s");
        df.add_col(v, "v");
        const auto R = df.group_by_str(0, {1});
        const std::vector<std::string> keys = {"", "date palm", "fig", "kiwi"};
        const std::vector<uint32_t> counts = {2, 1, 2, 3};
        const std::vector<int64_t> sums = {12, 6, 7, 11};
        bool ok = R.n_groups == 4 && R.keys == keys && R.counts == counts && R.sum_i[0] == sums;
        CHECK(ok);
        df.transform_group_by<vison_gpu::GroupFunction::Occurence>({0}, 0, "n");
        const auto nper = df.get_col<double>(2);
        const std::vector<double> want = {3, 2, 3, 2, 2, 1, 3, 2};
        CHECK(nper == want);
    }
    // errors surface as exceptions with the library's message
    {
        Dataframe df = load(f);
        bool thrown = false;
        try {
            df.group_by<int64_t>(1, {4});  // a char[16] column is not an aggregatable value
        } catch (const vison_gpu::Error& e) {
            thrown = e.status == VDF_ERR_UNSUPPORTED;
        }
        CHECK(thrown);
    }
    // read_csv: a file of uint64 keys, int64 values with negatives, float64
    // values, short and long text; every column as written, then a group-by
    // over the file's key column
    std::printf("read_csv\n");
    {
        char path[] = "/tmp/vdf_facade_XXXXXX";
        const int fd = mkstemp(path);
        CHECK(fd >= 0);
        if (fd >= 0) {
            const size_t m = 50007;
            std::mt19937_64 rng(5);
            std::vector<uint64_t> key(m);
            std::vector<int64_t> iv(m);
            std::vector<double> dv(m);
            std::vector<std::string> name(m), note(m);
            std::string text = "key,iv,dv,name,note\n";
            char buf[64];
            for (size_t i = 0; i < m; ++i) {
                key[i] = (rng() % 300) * 0x9E3779B97F4A7C15ull;
                iv[i] = i == 0 ? -7 : static_cast<int64_t>(rng() % 2000001) - 1000000;  // a '-' in the first rows
                dv[i] = std::ldexp(static_cast<double>(rng() >> 11), -53) * ((rng() & 1) ? 1.0 : -1.0);
                name[i] = std::string(1 + rng() % 10, static_cast<char>('a' + rng() % 26));
                note[i] = std::string(17 + rng() % 30, static_cast<char>('A' + rng() % 26));
                std::snprintf(buf, sizeof buf, "%.17g", dv[i]);
                text += std::to_string(key[i]) + "," + std::to_string(iv[i]) + "," + buf + "," + name[i] + "," +
                        note[i] + "\n";
            }
            CHECK(write(fd, text.data(), text.size()) == static_cast<ssize_t>(text.size()));
            close(fd);
            Dataframe df;
            df.read_csv(path);
            // the reference's template parameters: a row range, and the CPU
            // parser options (no effect here)
            Dataframe part;
            part.read_csv<100, 350, 4, false, false, '0'>(path, ',', true, '\'');
            CHECK(part.get_nrow() == 250 && part.get_ncol() == 5);
            CHECK(part.get_col<uint64_t>(0) == std::vector<uint64_t>(key.begin() + 100, key.begin() + 350));
            Dataframe tail;
            tail.read_csv<50000>(path);
            CHECK(tail.get_nrow() == m - 50000);
            CHECK(tail.get_col<int64_t>(1) == std::vector<int64_t>(iv.begin() + 50000, iv.end()));
            std::remove(path);
            CHECK(df.get_nrow() == m && df.get_ncol() == 5);
            CHECK(df.get_colname(0) == "key" && df.get_colname(4) == "note");
            CHECK(df.get_col<uint64_t>(0) == key);
            CHECK(df.get_col<int64_t>(1) == iv);
            const auto d2 = df.get_col<double>(2);
            bool ok = d2.size() == m;
            for (size_t i = 0; i < m && ok; ++i) ok = same_bits(d2[i], dv[i]);
            CHECK(ok);
            const auto nm = df.get_col<Char16>(3);
            ok = nm.size() == m;
            for (size_t i = 0; i < m && ok; ++i) {
                std::string want = name[i] + std::string(16 - name[i].size(), '0');
                ok = std::memcmp(nm[i].c, want.data(), 16) == 0;
            }
            CHECK(ok);
            CHECK(df.get_str_col(4) == note);
            const auto R = df.group_by<uint64_t>(0, {1});
            std::map<uint64_t, std::pair<uint64_t, int64_t>> acc;
            for (size_t i = 0; i < m; ++i) {
                acc[key[i]].first++;
                acc[key[i]].second += iv[i];
            }
            ok = R.n_groups == acc.size();
            size_t g = 0;
            for (auto it = acc.begin(); ok && it != acc.end(); ++it, ++g)
                ok = R.keys[g] == it->first && R.counts[g] == it->second.first && R.sum_i[0][g] == it->second.second;
            CHECK(ok);
        }
    }
    if (g_fail) {
        std::printf("%d failures\n", g_fail);
        return 1;
    }
    std::printf("facade ok\n");
    return 0;
}
