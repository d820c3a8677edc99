/*
 * vdf_datagen.c -- seeded synthetic columns for benchmarks and tests (host,
 * OpenMP).  Same streams as visondf_amd/datagen.py: SplitMix64 keyed by
 * (seed, column id); the Python module cross-checks this library.
 */
#include <math.h>
#include <omp.h>
#include <stdint.h>
#include <stdlib.h>

#define GOLDEN 0x9E3779B97F4A7C15ull

static inline uint64_t mix(uint64_t z)
{
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

static inline uint64_t stream_base(uint64_t seed, uint64_t col)
{
    return mix(seed * GOLDEN + (col & 0xFFFFull));
}

/* x_i = mix(s0 + (start + i + 1) * golden) */
void vdg_splitmix64(uint64_t seed, uint64_t col, uint64_t start, uint64_t n, uint64_t* out, int threads)
{
    const uint64_t s0 = stream_base(seed, col);
    #pragma omp parallel for num_threads(threads) schedule(static)
    for (uint64_t i = 0; i < n; ++i) out[i] = mix(s0 + (start + i + 1) * GOLDEN);
}

void vdg_uniform01(uint64_t seed, uint64_t col, uint64_t n, double* out, int threads)
{
    const uint64_t s0 = stream_base(seed, col);
    #pragma omp parallel for num_threads(threads) schedule(static)
    for (uint64_t i = 0; i < n; ++i) out[i] = (double)(mix(s0 + (i + 1) * GOLDEN) >> 11) * 0x1.0p-53;
}

/* Box-Muller: sqrt(-2 ln u1) cos(2 pi u2), u1 from column col, u2 from col + 1000 */
void vdg_normal01(uint64_t seed, uint64_t col, uint64_t n, double* out, int threads)
{
    const uint64_t s1 = stream_base(seed, col), s2 = stream_base(seed, col + 1000);
    const double two_pi = 2.0 * 3.141592653589793;
    #pragma omp parallel for num_threads(threads) schedule(static)
    for (uint64_t i = 0; i < n; ++i) {
        double u1 = (double)(mix(s1 + (i + 1) * GOLDEN) >> 11) * 0x1.0p-53;
        double u2 = (double)(mix(s2 + (i + 1) * GOLDEN) >> 11) * 0x1.0p-53;
        if (u1 <= 0.0) u1 = 0x1.0p-53;
        out[i] = sqrt(-2.0 * log(u1)) * cos(two_pi * u2);
    }
}

/* Zipf(s) ranks in [0, domain): first k with cdf[k] > u (u uniform01 of the
 * stream), using a guide table over 2^22 equal-probability buckets to bound
 * the binary search. */
void vdg_zipf_ranks(uint64_t seed, uint64_t col, uint64_t n, double s, uint64_t domain, int64_t* out,
                    int threads)
{
    double* cdf = (double*)malloc(domain * sizeof(double));
    double acc = 0.0;
    for (uint64_t k = 0; k < domain; ++k) {
        acc += pow((double)(k + 1), -s);
        cdf[k] = acc;
    }
    for (uint64_t k = 0; k < domain; ++k) cdf[k] /= acc;
    const uint64_t G = 1ull << 22;
    uint64_t* guide = (uint64_t*)malloc((G + 1) * sizeof(uint64_t));
    {
        uint64_t k = 0;
        for (uint64_t b = 0; b <= G; ++b) {
            const double t = (double)b / (double)G;
            while (k < domain && !(cdf[k] > t)) ++k;
            guide[b] = k;
        }
    }
    const uint64_t s0 = stream_base(seed, col);
    #pragma omp parallel for num_threads(threads) schedule(static)
    for (uint64_t i = 0; i < n; ++i) {
        const double u = (double)(mix(s0 + (i + 1) * GOLDEN) >> 11) * 0x1.0p-53;
        uint64_t b = (uint64_t)(u * (double)G);
        if (b >= G) b = G - 1;
        uint64_t lo = guide[b], hi = guide[b + 1];  /* answer in [lo, hi] */
        if (hi > domain) hi = domain;
        while (lo < hi) {
            const uint64_t mid = lo + (hi - lo) / 2;
            if (cdf[mid] > u) hi = mid; else lo = mid + 1;
        }
        out[i] = (int64_t)(lo < domain ? lo : domain - 1);
    }
    free(guide);
    free(cdf);
}

/* ---- CSV text shaped like the US airline delay-cause table ----
 * 21 columns per row: year, month, carrier (2 chars), carrier_name, airport
 * (3 chars), airport_name (quoted, with a comma inside the quotes), then 15
 * counts and delays written with two decimals ("123.45").  Field j of row r
 * draws from the SplitMix64 stream (seed, j) at index r, so the text does not
 * depend on the thread count.  Returns the bytes written (0 when cap is too
 * small: at most 64 + 400 bytes per row are needed). */
static const char* const kCarrierNames[16] = {
    "Endeavor Air Inc.", "American Airlines Inc.", "Alaska Airlines Inc.", "JetBlue Airways",
    "Delta Air Lines Inc.", "ExpressJet Airlines LLC", "Frontier Airlines Inc.", "Allegiant Air",
    "Hawaiian Airlines Inc.", "Envoy Air", "Spirit Air Lines", "PSA Airlines Inc.",
    "SkyWest Airlines Inc.", "United Air Lines Inc.", "Southwest Airlines Co.", "Mesa Airlines Inc."};
static const char* const kCities[16] = {
    "Albuquerque, NM", "Atlanta, GA", "Austin, TX", "Boston, MA", "Charlotte, NC", "Chicago, IL",
    "Denver, CO", "Detroit, MI", "Honolulu, HI", "Houston, TX", "Las Vegas, NV", "Los Angeles, CA",
    "Miami, FL", "New York, NY", "Seattle, WA", "St. Louis, MO"};
static const char* const kAirportKinds[4] = {"International", "Regional", "Municipal", "County"};

static inline uint64_t field_draw(uint64_t seed, int col, uint64_t r)
{
    return mix(stream_base(seed, 100 + (uint64_t)col) + (r + 1) * GOLDEN);
}

static char* put_u64(char* p, uint64_t v)
{
    char tmp[24];
    int n = 0;
    do {
        tmp[n++] = (char)('0' + v % 10);
        v /= 10;
    } while (v);
    while (n) *p++ = tmp[--n];
    return p;
}

static char* put_str(char* p, const char* s)
{
    while (*s) *p++ = *s++;
    return p;
}

static char* csv_row(char* p, uint64_t seed, uint64_t r)
{
    uint64_t d;
    d = field_draw(seed, 0, r);
    p = put_u64(p, 2003 + d % 21);
    *p++ = ',';
    d = field_draw(seed, 1, r);
    p = put_u64(p, 1 + d % 12);
    *p++ = ',';
    d = field_draw(seed, 2, r);
    const int carrier = (int)(d % 16);
    *p++ = (char)('A' + (d >> 8) % 26);
    *p++ = (char)((d >> 16) % 3 == 0 ? '0' + (d >> 20) % 10 : 'A' + (d >> 24) % 26);
    *p++ = ',';
    p = put_str(p, kCarrierNames[carrier]);
    *p++ = ',';
    d = field_draw(seed, 4, r);
    for (int i = 0; i < 3; ++i) *p++ = (char)('A' + (d >> (8 * i)) % 26);
    *p++ = ',';
    d = field_draw(seed, 5, r);
    *p++ = '"';
    p = put_str(p, kCities[d % 16]);
    *p++ = ':';
    *p++ = ' ';
    p = put_str(p, kCities[(d >> 8) % 16]);  /* the airport's name: a city, a kind, "Airport" */
    *p++ = ' ';
    p = put_str(p, kAirportKinds[(d >> 16) % 4]);
    p = put_str(p, " Airport\"");
    /* 15 two-decimal numbers: flights up to 20000, counts up to 2000, delays up to 200000 */
    for (int j = 0; j < 15; ++j) {
        d = field_draw(seed, 6 + j, r);
        const uint64_t scale = j == 0 ? 2000000ull : (j < 9 ? 200000ull : 20000000ull);
        const uint64_t v = d % scale;
        *p++ = ',';
        p = put_u64(p, v / 100);
        *p++ = '.';
        *p++ = (char)('0' + (v / 10) % 10);
        *p++ = (char)('0' + v % 10);
    }
    *p++ = '\n';
    return p;
}

uint64_t vdg_csv_airline(uint64_t seed, uint64_t nrows, char* out, uint64_t cap, int threads)
{
    static const char header[] =
        "year,month,carrier,carrier_name,airport,airport_name,arr_flights,arr_del15,carrier_ct,weather_ct,"
        "nas_ct,security_ct,late_aircraft_ct,arr_cancelled,arr_diverted,arr_delay,carrier_delay,weather_delay,"
        "nas_delay,security_delay,late_aircraft_delay\n";
    const uint64_t hb = sizeof(header) - 1;
    if (cap < hb + nrows * 400ull) return 0;
    char* p = out;
    for (uint64_t i = 0; i < hb; ++i) *p++ = header[i];
    if (threads < 1) threads = 1;
    /* rows in blocks: each block formatted into its slot of at most 400 bytes
     * per row, then the blocks moved together in order */
    const uint64_t blk = 1 << 16;
    const uint64_t nb = (nrows + blk - 1) / blk;
    uint64_t* len = (uint64_t*)calloc(nb ? nb : 1, sizeof(uint64_t));
    if (!len) return 0;
    char* base = out + hb;
    #pragma omp parallel for num_threads(threads) schedule(dynamic, 1)
    for (uint64_t b = 0; b < nb; ++b) {
        char* q = base + b * blk * 400ull;
        char* s = q;
        const uint64_t r1 = (b + 1) * blk < nrows ? (b + 1) * blk : nrows;
        for (uint64_t r = b * blk; r < r1; ++r) q = csv_row(q, seed, r);
        len[b] = (uint64_t)(q - s);
    }
    uint64_t w = 0;
    for (uint64_t b = 0; b < nb; ++b) {
        char* s = base + b * blk * 400ull;
        if (base + w != s)
            for (uint64_t i = 0; i < len[b]; ++i) base[w + i] = s[i];  /* left shift, never overlapping forward */
        w += len[b];
    }
    free(len);
    return hb + w;
}
