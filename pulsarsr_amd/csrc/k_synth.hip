// Seeded synthetic SIGPROC payloads generated on the GPU (workload tool, not
// part of the search path): the model of synth.cpp -- N(0,1) noise from a
// counter-based hash (SplitMix64 of seed, channel and sample) through an
// FP64 inverse-CDF table, extra noise in RFI channels, a zero-DM sinusoid,
// broadband bursts and Gaussian pulses delayed by the fractional physical
// dispersion delay -- quantised to round(mean + sigma * value) clipped to the
// bit depth and packed LSB-first in time-major frames.  One thread per
// payload byte at a time (it computes the byte's 8 / nbits samples; a
// grid-stride loop covers payloads of 2^32 bytes and more), so sub-byte frames
// pack without races.  A sample's pulse term counts every pulse whose centre
// is within 5 sigma of phase, which the host generator also uses as its
// support; rounding near a quantisation edge may still differ from the host
// generator in rare samples, so the two are separate generators with the same
// model, each deterministic in its seed.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
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
    double mean, sigma;
    uint32_t npulsars, nbursts;
    const psr_synth_pulsar* pulsars;
    const double* bursts;
    double sine_hz, sine_amp;
    const uint32_t* rfi_chans;
    uint32_t nrfi;
    double rfi_amp;
    const float* common;
    uint32_t threads;
} psr_synth_config;
int psr_synth_generate_gpu(const psr_synth_config* cfg, uint8_t* host_out, uint64_t out_bytes);
}

namespace {

constexpr int NT_BITS = 16;
constexpr int MAXP = 8;   // pulsars
constexpr int MAXB = 8;   // bursts

struct DevCfg {
    uint64_t seed, nsamples;
    uint32_t nchans, nbits, npulsars, nbursts;
    double tsamp, mean, sigma, maxv, sine_hz, sine_amp, rfi_amp;
    double p_period[MAXP], p_phase[MAXP], p_sig[MAXP], p_amp[MAXP];
    double b_start[MAXB], b_end[MAXB], b_amp[MAXB];
};

__device__ __forceinline__ uint64_t splitmix(uint64_t x) {
    x += 0x9e3779b97f4a7c15ull;
    x = (x ^ (x >> 30)) * 0xbf58476d1ce4e5b9ull;
    x = (x ^ (x >> 27)) * 0x94d049bb133111ebull;
    return x ^ (x >> 31);
}

__device__ __forceinline__ float norm_sample(const double* __restrict__ q, uint64_t h) {
    const uint32_t i = (uint32_t)(h >> (64 - NT_BITS));
    const double f = (double)((h >> 16) & 0xffffffffull) * (1.0 / 4294967296.0);
    return (float)(q[i] + f * (q[i + 1] - q[i]));
}

__device__ void synth_byte(const DevCfg& c, const double* __restrict__ q, const double* __restrict__ delay,
                           const uint8_t* __restrict__ is_rfi, const float* __restrict__ common,
                           uint8_t* __restrict__ out, uint64_t byte) {
    const uint32_t per = c.nbits >= 8 ? 1u : 8u / c.nbits;
    uint32_t packed = 0;
    for (uint32_t s = 0; s < per; s++) {
        const uint64_t idx = c.nbits >= 8 ? byte : byte * per + s;  // sample index t * nchans + ch
        const uint64_t t = idx / c.nchans;
        const uint32_t ch = (uint32_t)(idx % c.nchans);
        const uint64_t key = c.seed * 0x632be59bd9b4e019ull + (uint64_t)ch * c.nsamples;
        float v = norm_sample(q, splitmix(key + t));
        if (is_rfi[ch]) v += (float)(c.rfi_amp * norm_sample(q, splitmix(~(key + t))));
        if (common) v += common[t];
        if (c.sine_amp != 0.0) v += (float)(c.sine_amp * sin(2.0 * M_PI * c.sine_hz * (double)t * c.tsamp));
        for (uint32_t b = 0; b < c.nbursts; b++)
            if ((double)t >= c.b_start[b] && (double)t < c.b_end[b]) v += (float)c.b_amp[b];
        for (uint32_t p = 0; p < c.npulsars; p++) {
            double ph = ((double)t * c.tsamp - delay[(size_t)p * c.nchans + ch]) / c.p_period[p] - c.p_phase[p];
            ph -= floor(ph);
            const double dist = fmin(ph, 1.0 - ph) / c.p_sig[p];
            if (dist <= 5.0) v += (float)(c.p_amp[p] * exp(-0.5 * dist * dist));
        }
        double x = round(c.mean + c.sigma * (double)v);  // halves away from zero, as std::round
        x = fmin(fmax(x, 0.0), c.maxv);
        const uint32_t u = (uint32_t)x;
        if (c.nbits >= 8) {
            packed = u;
        } else {
            packed |= u << (s * c.nbits);
        }
    }
    out[byte] = (uint8_t)packed;
}

__global__ __launch_bounds__(256) void k_synth(DevCfg c, const double* __restrict__ q,
                                               const double* __restrict__ delay,  // [nchans] per pulsar
                                               const uint8_t* __restrict__ is_rfi,
                                               const float* __restrict__ common, uint8_t* __restrict__ out,
                                               uint64_t nbytes) {
    // grid-stride: payloads of 4 GiB and more exceed one launch's 2^32 threads
    for (uint64_t byte = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; byte < nbytes;
         byte += (uint64_t)gridDim.x * blockDim.x)
        synth_byte(c, q, delay, is_rfi, common, out, byte);
}

std::vector<double> make_table() {
    const int N = 1 << NT_BITS;
    std::vector<double> q(N + 1);
    for (int i = 0; i <= N; i++) {
        const double p = ((double)i + 0.5) / (double)(N + 1);
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
    return q;
}

}  // namespace

// Generates into device memory, then copies the packed payload to host_out.
// Returns 0, -1 on invalid parameters, -2 on a short output, -3 on a HIP error.
extern "C" int psr_synth_generate_gpu(const psr_synth_config* c, uint8_t* host_out, uint64_t out_bytes) {
    if (!c || !host_out || c->nchans == 0 || c->nsamples < 2) return -1;
    const uint32_t nb = c->nbits;
    if (!(nb == 1 || nb == 2 || nb == 4 || nb == 8)) return -1;  // packed sub-byte and 8-bit samples
    if (c->npulsars > MAXP || c->nbursts > MAXB) return -1;
    const uint64_t need = (c->nsamples * c->nchans * nb + 7) / 8;
    if (out_bytes < need) return -2;
    if ((uint64_t)c->nchans * nb % 8 != 0) return -1;  // whole-byte frames
    static const std::vector<double> table = make_table();
    DevCfg d;
    std::memset(&d, 0, sizeof d);
    d.seed = c->seed;
    d.nsamples = c->nsamples;
    d.nchans = c->nchans;
    d.nbits = nb;
    d.npulsars = c->npulsars;
    d.nbursts = c->nbursts;
    d.tsamp = c->tsamp;
    d.mean = c->mean;
    d.sigma = c->sigma;
    d.maxv = (double)((1u << nb) - 1);
    d.sine_hz = c->sine_hz;
    d.sine_amp = c->sine_amp;
    d.rfi_amp = c->rfi_amp;
    std::vector<double> fr(c->nchans);
    double fmax = 0.0;
    for (uint32_t ch = 0; ch < c->nchans; ch++) {
        fr[ch] = c->fch1 + (double)ch * c->foff;
        fmax = std::fmax(fmax, fr[ch]);
    }
    std::vector<double> delay((size_t)std::max<uint32_t>(1, c->npulsars) * c->nchans, 0.0);
    for (uint32_t p = 0; p < c->npulsars; p++) {
        const psr_synth_pulsar& s = c->pulsars[p];
        d.p_period[p] = s.period_s;
        d.p_phase[p] = s.phase;
        d.p_sig[p] = s.duty / 2.355;
        d.p_amp[p] = s.amplitude;
        for (uint32_t ch = 0; ch < c->nchans; ch++)
            delay[(size_t)p * c->nchans + ch] = 4148.808 * s.dm * (1.0 / (fr[ch] * fr[ch]) - 1.0 / (fmax * fmax));
    }
    for (uint32_t b = 0; b < c->nbursts; b++) {  // samples [ceil(s / tsamp), floor((s + w) / tsamp))
        d.b_start[b] = std::ceil(c->bursts[3 * b] / c->tsamp);
        d.b_end[b] = std::floor((c->bursts[3 * b] + c->bursts[3 * b + 1]) / c->tsamp);
        d.b_amp[b] = c->bursts[3 * b + 2];
    }
    std::vector<uint8_t> rfi(c->nchans, 0);
    for (uint32_t i = 0; i < c->nrfi; i++)
        if (c->rfi_chans[i] < c->nchans) rfi[c->rfi_chans[i]] = 1;
    double *dq = nullptr, *dd = nullptr;
    uint8_t *dr = nullptr, *dout = nullptr;
    float* dc = nullptr;
    int rc = 0;
    hipError_t e = hipMalloc(&dq, table.size() * sizeof(double));
    if (e == hipSuccess) e = hipMalloc(&dd, delay.size() * sizeof(double));
    if (e == hipSuccess) e = hipMalloc(&dr, rfi.size());
    if (e == hipSuccess) e = hipMalloc(&dout, need);
    if (e == hipSuccess && c->common) e = hipMalloc(&dc, c->nsamples * sizeof(float));
    if (e == hipSuccess) e = hipMemcpy(dq, table.data(), table.size() * sizeof(double), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(dd, delay.data(), delay.size() * sizeof(double), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(dr, rfi.data(), rfi.size(), hipMemcpyHostToDevice);
    if (e == hipSuccess && dc) e = hipMemcpy(dc, c->common, c->nsamples * sizeof(float), hipMemcpyHostToDevice);
    if (e == hipSuccess) {
        const uint64_t blocks = std::min<uint64_t>((need + 255) / 256, 1u << 20);
        hipLaunchKernelGGL(k_synth, dim3((unsigned)blocks), dim3(256), 0, 0, d, dq, dd, dr, dc, dout, need);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpy(host_out, dout, need, hipMemcpyDeviceToHost);
    if (e != hipSuccess) rc = -3;
    if (dq) (void)hipFree(dq);
    if (dd) (void)hipFree(dd);
    if (dr) (void)hipFree(dr);
    if (dout) (void)hipFree(dout);
    if (dc) (void)hipFree(dc);
    return rc;
}
