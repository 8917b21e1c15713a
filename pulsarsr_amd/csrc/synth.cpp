// Seeded synthetic SIGPROC filterbanks for benchmarks and tests.
//
// The reference's generator (crates/synth/src/lib.rs:38-85) writes 32-bit
// floats with one pulsar; the workloads here need 1/2/4/8-bit packed data with
// several pulsars, RFI and red noise, so this is the project's own generator.
// Noise is N(0,1) from a counter-based hash (SplitMix64 of seed and sample
// index) through an FP64-built inverse-CDF table with linear interpolation, so
// the output depends only on the seed and parameters, not on thread count.
// Pulses follow the reference's model: a Gaussian in phase with sigma =
// duty/2.355 at phase 0.25, delayed per channel by the fractional physical
// dispersion delay (lib.rs:58-62).  Samples are round(mean + sigma*value),
// clipped to [0, 2^nbits-1], packed LSB-first in time-major frames.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <thread>
#include <vector>

extern "C" {

typedef struct {
    double period_s, dm, duty, amplitude, phase;
} psr_synth_pulsar;

typedef struct {
    uint64_t seed;
    uint64_t nsamples;
    uint32_t nchans, nbits;
    double tsamp, fch1, foff;
    double mean, sigma;            // quantisation level of N(0,1) noise
    uint32_t npulsars, nbursts;
    const psr_synth_pulsar* pulsars;
    const double* bursts;          // nbursts * (start_s, width_s, amplitude)
    double sine_hz, sine_amp;      // zero-DM sinusoid in every channel
    const uint32_t* rfi_chans;     // narrowband RFI channels
    uint32_t nrfi;
    double rfi_amp;                // extra noise amplitude in RFI channels
    const float* common;           // optional common-mode series (red noise), nsamples
    uint32_t threads;
} psr_synth_config;

int psr_synth_generate(const psr_synth_config* cfg, uint8_t* out, uint64_t out_bytes);
}

namespace {

inline uint64_t splitmix(uint64_t x) {
    x += 0x9e3779b97f4a7c15ull;
    x = (x ^ (x >> 30)) * 0xbf58476d1ce4e5b9ull;
    x = (x ^ (x >> 27)) * 0x94d049bb133111ebull;
    return x ^ (x >> 31);
}

struct NormTable {
    std::vector<double> q;  // inverse normal CDF at (i + 0.5) / N, i < N
    static constexpr int BITS = 16;
    NormTable() {
        const int N = 1 << BITS;
        q.resize(N + 1);
        for (int i = 0; i <= N; i++) {
            double p = ((double)i + 0.5) / (double)(N + 1);
            // inverse CDF by Newton iterations on erfc
            double x = 0.0;
            for (int it = 0; it < 60; it++) {
                const double cdf = 0.5 * std::erfc(-x / std::sqrt(2.0));
                const double pdf = std::exp(-0.5 * x * x) / std::sqrt(2.0 * M_PI);
                const double step = (cdf - p) / pdf;
                x -= step;
                if (std::fabs(step) < 1e-14) break;
            }
            q[i] = x;
        }
    }
    inline float sample(uint64_t h) const {
        const uint32_t i = (uint32_t)(h >> (64 - BITS));
        const double f = (double)((h >> 16) & 0xffffffffull) * (1.0 / 4294967296.0);
        return (float)(q[i] + f * (q[i + 1] - q[i]));
    }
};

const NormTable& table() {
    static NormTable t;
    return t;
}

}  // namespace

extern "C" int psr_synth_generate(const psr_synth_config* c, uint8_t* out, uint64_t out_bytes) {
    if (!c || !out || c->nchans == 0 || c->nsamples < 2) return -1;
    const uint32_t nb = c->nbits;
    if (!(nb == 1 || nb == 2 || nb == 4 || nb == 8 || nb == 16 || nb == 32)) return -1;
    const uint64_t need = (c->nsamples * c->nchans * nb + 7) / 8;
    if (out_bytes < need) return -2;
    std::memset(out, 0, need);
    const NormTable& nt = table();
    std::vector<double> freqs(c->nchans), delay(c->nchans);
    double fmax = 0.0;
    for (uint32_t ch = 0; ch < c->nchans; ch++) {
        freqs[ch] = c->fch1 + (double)ch * c->foff;
        fmax = std::max(fmax, freqs[ch]);
    }
    std::vector<uint8_t> is_rfi(c->nchans, 0);
    for (uint32_t i = 0; i < c->nrfi; i++)
        if (c->rfi_chans[i] < c->nchans) is_rfi[c->rfi_chans[i]] = 1;
    const double maxv = nb == 32 ? 0.0 : (double)((1ull << nb) - 1);
    const uint64_t TB = 4096;
    const uint64_t ntb = (c->nsamples + TB - 1) / TB;
    unsigned nth = c->threads ? c->threads : std::max(1u, std::thread::hardware_concurrency());
    nth = std::min<unsigned>(nth, 64);
    auto work = [&](unsigned w) {
        std::vector<float> row(TB);
        for (uint64_t tb = w; tb < ntb; tb += nth) {
            const uint64_t t0 = tb * TB, t1 = std::min(c->nsamples, t0 + TB);
            const uint64_t len = t1 - t0;
            for (uint32_t ch = 0; ch < c->nchans; ch++) {
                const uint64_t key = c->seed * 0x632be59bd9b4e019ull + (uint64_t)ch * c->nsamples;
                const double amp_rfi = is_rfi[ch] ? c->rfi_amp : 0.0;
                for (uint64_t i = 0; i < len; i++) {
                    const uint64_t t = t0 + i;
                    float v = nt.sample(splitmix(key + t));
                    if (amp_rfi != 0.0) v += (float)(amp_rfi * nt.sample(splitmix(~(key + t))));
                    if (c->common) v += c->common[t];
                    if (c->sine_amp != 0.0)
                        v += (float)(c->sine_amp * std::sin(2.0 * M_PI * c->sine_hz * (double)t * c->tsamp));
                    row[i] = v;
                }
                for (uint32_t b = 0; b < c->nbursts; b++) {
                    const double s = c->bursts[3 * b], wdt = c->bursts[3 * b + 1], a = c->bursts[3 * b + 2];
                    const int64_t bs = (int64_t)std::ceil(s / c->tsamp), be = (int64_t)std::floor((s + wdt) / c->tsamp);
                    for (int64_t t = std::max<int64_t>(bs, (int64_t)t0); t < std::min<int64_t>(be, (int64_t)t1); t++)
                        row[t - t0] += (float)a;
                }
                const double f = freqs[ch];
                for (uint32_t p = 0; p < c->npulsars; p++) {
                    const psr_synth_pulsar& q = c->pulsars[p];
                    const double d = 4148.808 * q.dm * (1.0 / (f * f) - 1.0 / (fmax * fmax));
                    const double sig = q.duty / 2.355;
                    const double half = 5.0 * sig * q.period_s;  // +-5 sigma support
                    // pulse centres: t*tsamp - d = (k + phase) * P
                    const double tstart = (double)t0 * c->tsamp, tend = (double)t1 * c->tsamp;
                    const int64_t k0 = (int64_t)std::floor((tstart - d - half) / q.period_s - q.phase);
                    const int64_t k1 = (int64_t)std::ceil((tend - d + half) / q.period_s - q.phase);
                    for (int64_t k = k0; k <= k1; k++) {
                        const double tc = ((double)k + q.phase) * q.period_s + d;
                        const int64_t a0 = std::max<int64_t>((int64_t)t0, (int64_t)std::ceil((tc - half) / c->tsamp));
                        const int64_t a1 = std::min<int64_t>((int64_t)t1 - 1, (int64_t)std::floor((tc + half) / c->tsamp));
                        for (int64_t t = a0; t <= a1; t++) {
                            double ph = ((double)t * c->tsamp - d) / q.period_s - q.phase;
                            ph -= std::floor(ph);
                            const double dist = std::min(ph, 1.0 - ph) / sig;
                            row[t - t0] += (float)(q.amplitude * std::exp(-0.5 * dist * dist));
                        }
                    }
                }
                for (uint64_t i = 0; i < len; i++) {
                    const uint64_t t = t0 + i;
                    const uint64_t idx = t * c->nchans + ch;
                    if (nb == 32) {
                        const float x = (float)(c->mean + c->sigma * row[i]);
                        std::memcpy(out + 4 * idx, &x, 4);
                        continue;
                    }
                    double x = std::round(c->mean + c->sigma * (double)row[i]);
                    x = std::min(std::max(x, 0.0), maxv);
                    const uint64_t u = (uint64_t)x;
                    if (nb == 16) {
                        out[2 * idx] = (uint8_t)(u & 0xff);
                        out[2 * idx + 1] = (uint8_t)(u >> 8);
                    } else if (nb == 8) {
                        out[idx] = (uint8_t)u;
                    } else {
                        const uint64_t bit = idx * nb;
                        // frames of whole bytes: a sample never straddles bytes
                        out[bit >> 3] |= (uint8_t)(u << (bit & 7));
                    }
                }
            }
        }
    };
    if ((uint64_t)c->nchans * nb % 8 != 0 && nb < 8) {
        // frames not byte aligned: samples of neighbouring frames share bytes,
        // so generate serially to keep the OR-packing race free
        nth = 1;
    }
    std::vector<std::thread> pool;
    for (unsigned w = 1; w < nth; w++) pool.emplace_back(work, w);
    work(0);
    for (auto& th : pool) th.join();
    return 0;
}
