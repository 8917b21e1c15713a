// Host-side numerics of the search path that do not need a GPU: the
// crates/dedisperse planning functions, the SIGPROC header parser of
// crates/fbank, search::sift, and the FFT geometry/twiddle tables used by the
// device kernels.  Built with -ffp-contract=off so every f64 expression rounds
// exactly as the reference's Rust (which never fuses multiply-add).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <string>
#include <vector>

#include "../../include/psr_gpu.h"
#include "common.hpp"
#include "host_internal.hpp"

namespace psr {

static thread_local std::string g_last_error;

int fail(int code, const std::string& msg) {
    g_last_error = msg;
    return code;
}
void clear_error() { g_last_error.clear(); }
const char* last_error() { return g_last_error.c_str(); }

namespace dedisperse {
static const char* kErr = "invalid dedispersion input: ";

// crates/dedisperse/src/lib.rs:9-34
int delays(const double* freqs, size_t nch, double tsamp, double dm, uint64_t* out) {
    bool bad = nch == 0 || !std::isfinite(tsamp) || tsamp <= 0.0 || !std::isfinite(dm) || dm < 0.0;
    for (size_t c = 0; !bad && c < nch; c++) bad = !std::isfinite(freqs[c]) || freqs[c] <= 0.0;
    if (bad)
        return fail(PSR_E_INVALID,
                    std::string(kErr) + "positive frequencies/tsamp and nonnegative DM required");
    double reference = 0.0;
    for (size_t c = 0; c < nch; c++) reference = std::fmax(reference, freqs[c]);
    const double ref_inv = 1.0 / (reference * reference);  // powi(-2) == 1/(f*f)
    for (size_t c = 0; c < nch; c++) {
        const double inv = 1.0 / (freqs[c] * freqs[c]);
        const double samples = std::round(DISPERSION_SECONDS_MHZ * dm * (inv - ref_inv) / tsamp);
        if (samples > 18446744073709551615.0)
            return fail(PSR_E_INVALID, std::string(kErr) + "delay overflow");
        out[c] = (uint64_t)std::fmax(samples, 0.0);
    }
    return PSR_OK;
}

// crates/dedisperse/src/lib.rs:35-44
int dm_step(const double* freqs, size_t nch, double tsamp, double* out) {
    std::vector<uint64_t> tmp(nch ? nch : 1);
    int rc = delays(freqs, nch, tsamp, 0.0, tmp.data());
    if (rc) return rc;
    double lo = INFINITY, hi = 0.0;
    for (size_t c = 0; c < nch; c++) {
        lo = std::fmin(lo, freqs[c]);
        hi = std::fmax(hi, freqs[c]);
    }
    const double span = DISPERSION_SECONDS_MHZ * (1.0 / (lo * lo) - 1.0 / (hi * hi));
    if (span <= 0.0) return fail(PSR_E_INVALID, std::string(kErr) + "DM planning requires finite bandwidth");
    *out = tsamp / span;
    return PSR_OK;
}

// crates/dedisperse/src/lib.rs:45-60
int plan(double min, double max, double step, double* out, size_t cap, size_t* count) {
    if (!std::isfinite(min) || !std::isfinite(max) || !std::isfinite(step) || min < 0.0 ||
        max < min || step <= 0.0)
        return fail(PSR_E_INVALID, std::string(kErr) + "invalid DM bounds/step");
    const double q = std::floor((max - min) / step);
    const size_t n = (size_t)q + 1;
    if (n > 100000) return fail(PSR_E_INVALID, std::string(kErr) + "more than 100000 DM trials");
    *count = n;
    if (out) {
        if (n > cap) return fail(PSR_E_SHAPE, "plan output buffer too small");
        for (size_t i = 0; i < n; i++) out[i] = min + (double)i * step;
    }
    return PSR_OK;
}
}  // namespace dedisperse

namespace fbank {
static const char* kFmt = "invalid or unsupported filterbank: ";

struct Cursor {
    const uint8_t* p;
    size_t len, pos;
    bool read(void* dst, size_t n) {
        if (pos + n > len) return false;
        std::memcpy(dst, p + pos, n);
        pos += n;
        return true;
    }
};
static int read_string(Cursor& c, std::string& s) {
    int32_t n;
    if (!c.read(&n, 4)) return fail(PSR_E_FORMAT, "failed to fill whole buffer");
    if (n < 1 || n > 4096) return fail(PSR_E_FORMAT, std::string(kFmt) + "bad string length " + std::to_string(n));
    s.resize((size_t)n);
    if (!c.read(&s[0], (size_t)n)) return fail(PSR_E_FORMAT, "failed to fill whole buffer");
    return PSR_OK;
}

// crates/fbank/src/lib.rs:72-147: strict keyword table, duplicates rejected.
int parse_header(const uint8_t* bytes, size_t len, psr_header* out) {
    static const char* ints[] = {"machine_id", "telescope_id", "data_type", "barycentric",
                                 "pulsarcentric", "nbits", "nsamples", "nchans", "nifs",
                                 "nbeams", "ibeam", "signed"};
    static const char* floats[] = {"az_start", "za_start", "src_raj", "src_dej", "tstart",
                                   "tsamp", "fch1", "foff", "refdm", "period"};
    Cursor c{bytes, len, 0};
    std::string key;
    int rc = read_string(c, key);
    if (rc) return rc;
    if (key != "HEADER_START") return fail(PSR_E_FORMAT, std::string(kFmt) + "missing HEADER_START");
    std::map<std::string, int32_t> iv;
    std::map<std::string, double> fv;
    std::map<std::string, bool> seen;
    for (;;) {
        if ((rc = read_string(c, key))) return rc;
        if (key == "HEADER_END") break;
        bool is_int = false, is_float = false;
        for (const char* k : ints) is_int |= key == k;
        for (const char* k : floats) is_float |= key == k;
        if (key == "rawdatafile" || key == "source_name") {
            std::string v;
            if ((rc = read_string(c, v))) return rc;
        } else if (is_int) {
            int32_t v;
            if (!c.read(&v, 4)) return fail(PSR_E_FORMAT, "failed to fill whole buffer");
            iv[key] = v;
        } else if (is_float) {
            double v;
            if (!c.read(&v, 8)) return fail(PSR_E_FORMAT, "failed to fill whole buffer");
            fv[key] = v;
        } else {
            return fail(PSR_E_FORMAT, std::string(kFmt) + "unknown keyword " + key +
                                          "; cannot infer its type safely");
        }
        if (seen[key]) return fail(PSR_E_FORMAT, std::string(kFmt) + "duplicate " + key);
        seen[key] = true;
    }
    auto need_int = [&](const char* k, int32_t* v) {
        auto it = iv.find(k);
        if (it == iv.end()) return fail(PSR_E_FORMAT, std::string(kFmt) + "missing integer " + k);
        *v = it->second;
        return PSR_OK;
    };
    auto need_float = [&](const char* k, double* v) {
        auto it = fv.find(k);
        if (it == fv.end() || !std::isfinite(it->second))
            return fail(PSR_E_FORMAT, std::string(kFmt) + "missing/invalid float " + k);
        *v = it->second;
        return PSR_OK;
    };
    int32_t nchans, nbits;
    double tsamp, fch1, foff;
    if ((rc = need_int("nchans", &nchans)) || (rc = need_int("nbits", &nbits)) ||
        (rc = need_float("tsamp", &tsamp)) || (rc = need_float("fch1", &fch1)) ||
        (rc = need_float("foff", &foff)))
        return rc;
    if (nchans <= 0 || tsamp <= 0.0 ||
        !(nbits == 1 || nbits == 2 || nbits == 4 || nbits == 8 || nbits == 16 || nbits == 32))
        return fail(PSR_E_FORMAT, std::string(kFmt) + "invalid nchans/tsamp/nbits");
    if (iv.count("nifs") && iv["nifs"] != 1) return fail(PSR_E_FORMAT, std::string(kFmt) + "only nifs=1 supported");
    if (iv.count("signed") && iv["signed"] != 0)
        return fail(PSR_E_FORMAT, std::string(kFmt) + "signed integer samples unsupported");
    if (iv.count("data_type") && iv["data_type"] != 1)
        return fail(PSR_E_FORMAT, std::string(kFmt) + "only filterbank data_type=1 supported");
    for (int32_t ch = 0; ch < nchans; ch++) {
        const double f = fch1 + (double)ch * foff;
        if (!std::isfinite(f) || f <= 0.0) return fail(PSR_E_FORMAT, std::string(kFmt) + "nonpositive frequency");
    }
    out->nchans = (uint32_t)nchans;
    out->nbits = (uint32_t)nbits;
    out->tsamp = tsamp;
    out->fch1 = fch1;
    out->foff = foff;
    out->data_offset = c.pos;
    const uint64_t frame_bits = (uint64_t)nchans * (uint64_t)nbits;
    out->nsamples = (uint64_t)(len - c.pos) * 8 / frame_bits;
    return PSR_OK;
}
}  // namespace fbank

namespace search {
static inline int64_t total_key64(double f) {
    int64_t b;
    std::memcpy(&b, &f, 8);
    b ^= (int64_t)(((uint64_t)(b >> 63)) >> 1);
    return b;
}
// crates/search/src/lib.rs:147-176: stable descending sort, harmonic relation 1..16.
int sift(const psr_candidate* in, size_t n, double resolution_hz, size_t limit, int reject_dm_zero,
         psr_candidate* out, size_t* count) {
    std::vector<size_t> idx(n);
    for (size_t i = 0; i < n; i++) idx[i] = i;
    std::stable_sort(idx.begin(), idx.end(), [&](size_t a, size_t b) {
        return total_key64(in[a].score) > total_key64(in[b].score);
    });
    std::vector<double> rejected;
    size_t kept = 0;
    for (size_t ii = 0; ii < n; ii++) {
        const psr_candidate& c = in[idx[ii]];
        auto related = [&](double f) {
            for (int h = 1; h <= 16; h++)
                if (std::fabs(c.frequency_hz - f * (double)h) <= resolution_hz ||
                    std::fabs(f - c.frequency_hz * (double)h) <= resolution_hz)
                    return true;
            return false;
        };
        bool rel = false;
        for (size_t k = 0; k < kept && !rel; k++) rel = related(out[k].frequency_hz);
        for (size_t k = 0; k < rejected.size() && !rel; k++) rel = related(rejected[k]);
        if (rel) continue;
        if (reject_dm_zero && c.dm <= 0.0) {
            rejected.push_back(c.frequency_hz);
            continue;
        }
        out[kept++] = c;
        if (kept >= limit) break;
    }
    *count = kept;
    return PSR_OK;
}
}  // namespace search

// ---------------------------------------------------------------- geometry
static uint32_t next_pow2(uint32_t n) {
    uint32_t p = 1;
    while (p < n) p <<= 1;
    return p;
}

SpecGeom make_geometry_split(uint32_t n, uint32_t m2) {
    SpecGeom g = make_geometry(n);
    if (g.m1 > 1 && m2 && g.m % m2 == 0) {
        g.m2 = m2;
        g.m1 = g.m / m2;
    }
    return g;
}

SpecGeom make_geometry(uint32_t n) {
    SpecGeom g{};
    g.n = n;
    g.nfft = next_pow2(n) * 2;  // crates/search/src/lib.rs:35
    g.m = g.nfft / 2;
    g.nbins = g.m + 1;
    if (g.m <= FFT_SINGLE_MAX) {
        g.m1 = 1;
        g.m2 = g.m;
    } else {
        g.m2 = FFT_PASS1_LEN;
        while ((uint64_t)g.m / g.m2 > FFT_PASS2_MAX) g.m2 <<= 1;
        g.m1 = g.m / g.m2;
    }
    g.nblocks = (g.nbins + 63) / 64;
    g.series_pitch = (n + 255u) & ~255u;
    g.spec_pitch = ((uint64_t)g.nbins + 255u) & ~(uint64_t)255u;
    return g;
}

// Geometry-specific twiddles in FP64, stored FP32 (used only by the four-step
// path):  [W_m^j, j < 2048][W_m^(2048 j), j < m/2048][W_nfft^k2, k2 < m2]
size_t twiddle_table_size(const SpecGeom& g) {
    const size_t hi = g.m >= 2048 ? g.m / 2048 : 1;
    return 2048 + hi + (size_t)g.m2;
}
static void put(float2* t, size_t i, double turns) {
    const double a = -2.0 * M_PI * turns;
    t[i].x = (float)std::cos(a);
    t[i].y = (float)std::sin(a);
}
void fill_twiddles(const SpecGeom& g, float2* t) {
    size_t o = 0;
    for (uint32_t j = 0; j < 2048; j++) put(t, o + j, (double)j / g.m);
    o += 2048;
    const size_t hi = g.m >= 2048 ? g.m / 2048 : 1;
    for (size_t j = 0; j < hi; j++) put(t, o + j, (double)(2048.0 * j) / g.m);
    o += hi;
    for (uint32_t j = 0; j < g.m2; j++) put(t, o + j, (double)j / g.nfft);
}
void fill_costabs(float* t) {
    for (int L = 2; L <= COSTAB_MAX; L *= 2) {
        const int off = costab_offset(L);
        for (int j = 0; j <= L / 4; j++) t[off + j] = (float)std::cos(2.0 * M_PI * (double)j / (double)L);
        t[off + L / 4] = 0.0f;  // cos(pi/2)
    }
    for (int i = COSTAB_SIZE; i < TW4096_OFFSET; i++) t[i] = 0.0f;
    for (int k = 1; k < 16; k++)
        for (int j = 0; j < 256; j++) {
            const double a = -2.0 * M_PI * (double)((k * j) % 4096) / 4096.0;
            t[TW4096_OFFSET + 2 * (256 * (k - 1) + j)] = (float)std::cos(a);
            t[TW4096_OFFSET + 2 * (256 * (k - 1) + j) + 1] = (float)std::sin(a);
        }
}

}  // namespace psr
