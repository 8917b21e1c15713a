// C ABI of the MI355X search path (include/psr_gpu.h): context and device
// memory management, observation ingest, and the orchestration of the
// delays -> dedisperse -> FFT -> whiten -> harmonic sum -> rank chain that
// replaces the reference's per-trial loop (crates/cli/src/main.rs:462-469).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>
#include <string>
#include <vector>

#include "ctx_internal.hpp"

using namespace psr;

namespace psr {

int ensure(psr_ctx* c, DevBuf& b, size_t bytes) {
    if (b.bytes >= bytes) return PSR_OK;
    if (b.p) {
        (void)hipFree(b.p);
        c->mem_used -= b.bytes;
        b.p = nullptr;
        b.bytes = 0;
    }
    if (c->mem_cap && c->mem_used + bytes > c->mem_cap)
        return fail(PSR_E_OOM, "device memory cap exceeded (" + std::to_string(c->mem_used + bytes) +
                                   " > " + std::to_string(c->mem_cap) + " bytes)");
    hipError_t e = hipMalloc(&b.p, bytes ? bytes : 1);
    if (e != hipSuccess) {
        b.p = nullptr;
        return fail(PSR_E_OOM, std::string("hipMalloc failed: ") + hipGetErrorString(e));
    }
    b.bytes = bytes;
    c->mem_used += bytes;
    return PSR_OK;
}

void release(psr_ctx* c, DevBuf& b) {
    if (b.p) {
        (void)hipFree(b.p);
        c->mem_used -= b.bytes;
    }
    b = DevBuf{};
}

int check_device(psr_ctx* c) {
    HIP_TRY(hipSetDevice(c->device));
    return PSR_OK;
}

}  // namespace psr

// apply()'s length rules (crates/dedisperse/src/lib.rs:72-78) for the largest DM
// of the batch; delays are non-decreasing in DM so it bounds every trial.
int psr::resolve_length(const psr_obs* o, const double* dms, size_t ntrials, size_t length,
                        size_t* n_out, uint64_t* maxdelay_out, bool fft_limit) {
    double dmax = 0.0;
    for (size_t i = 0; i < ntrials; i++) {
        if (!std::isfinite(dms[i]) || dms[i] < 0.0)
            return fail(PSR_E_INVALID,
                        "invalid dedispersion input: positive frequencies/tsamp and nonnegative DM required");
        dmax = std::max(dmax, dms[i]);
    }
    std::vector<uint64_t> d(o->nchans);
    int rc = dedisperse::delays(o->freqs.data(), o->nchans, o->tsamp, dmax, d.data());
    if (rc) return rc;
    const uint64_t maxd = *std::max_element(d.begin(), d.end());
    if (maxd > o->nsamples) return fail(PSR_E_INVALID, "invalid dedispersion input: delay exceeds observation");
    const uint64_t available = o->nsamples - maxd;
    const uint64_t n = length ? length : available;
    if (n < 2 || n > available) return fail(PSR_E_SHAPE, "invalid dedispersion input: invalid output length");
    if (fft_limit && n > (1u << 23))
        return fail(PSR_E_SHAPE, "series longer than 2^23 samples is not supported");
    if (n > (uint64_t)INT32_MAX || maxd > (uint64_t)INT32_MAX)
        return fail(PSR_E_SHAPE, "series or delays past 2^31 samples are not supported");
    *n_out = (size_t)n;
    *maxdelay_out = maxd;
    return PSR_OK;
}

// Per-channel bound on the delay spread (samples) over any block of `block`
// consecutive sorted trials: the widest DM span of a block times the channel's
// dispersion slope, plus 3 samples for rounding.
std::vector<uint32_t> psr::spread_bounds(const psr_obs* o, const double* dms_sorted, size_t ntrials,
                                    size_t block) {
    double worst = 0.0;
    for (size_t b = 0; b < ntrials; b += block) {
        const size_t e = std::min(ntrials, b + block) - 1;
        worst = std::max(worst, dms_sorted[e] - dms_sorted[b]);
    }
    std::vector<uint32_t> out(o->freqs.size());
    for (size_t c = 0; c < out.size(); c++) {
        const double f = o->freqs[c];
        const double a = std::max(0.0, 1.0 / (f * f) - o->ref_inv);
        out[c] = (uint32_t)std::ceil(DISPERSION_SECONDS_MHZ * a * worst / o->tsamp) + 3u;
    }
    return out;
}

uint64_t psr::max_delay_at(const psr_obs* o, double dm) {
    // the device table's expression (k_delays: relative to the observation's
    // reference frequency, also for channel-range views), plus one
    double worst = 0.0;
    for (const double f : o->freqs) {
        const double samples = std::round(DISPERSION_SECONDS_MHZ * dm * (1.0 / (f * f) - o->ref_inv) / o->tsamp);
        if (!(samples <= 4e18)) return UINT64_MAX;
        worst = std::max(worst, samples);
    }
    return (uint64_t)worst + 1u;
}

namespace {

int upload_twiddles(psr_ctx* c, const SpecGeom& g) {
    if (!c->costabs.p) {
        std::vector<float> ct(COSTAB_ALL);
        fill_costabs(ct.data());
        int rc = ensure(c, c->costabs, ct.size() * sizeof(float));
        if (rc) return rc;
        HIP_TRY(hipMemcpy(c->costabs.p, ct.data(), ct.size() * sizeof(float), hipMemcpyHostToDevice));
    }
    if (c->tw_valid && c->tw_geom.m == g.m && c->tw_geom.m1 == g.m1 && c->tw_geom.nfft == g.nfft)
        return PSR_OK;
    const size_t nt = twiddle_table_size(g);
    std::vector<float2> t(nt);
    fill_twiddles(g, t.data());
    int rc = ensure(c, c->twiddles, nt * sizeof(float2));
    if (rc) return rc;
    HIP_TRY(hipMemcpy(c->twiddles.p, t.data(), nt * sizeof(float2), hipMemcpyHostToDevice));
    c->tw_geom = g;
    c->tw_valid = true;
    return PSR_OK;
}

}  // namespace

// search::periodicity's bin limits (crates/search/src/lib.rs:73-80) and birdie
// bins (lib.rs:62-72) for one geometry.
int psr::make_params(const SpecGeom& g, double tsamp, const psr_search_cfg* cfg, SearchParams* p,
                std::vector<uint32_t>* birdie_bins) {
    std::memset(p, 0, sizeof(*p));
    p->tsamp = tsamp;
    p->bin_hz = 1.0 / ((double)g.nfft * tsamp);
    p->threshold = cfg->threshold;
    p->resolution = 1.0 / ((double)g.n * tsamp);
    const double fq = std::ceil(cfg->min_hz / p->bin_hz);
    const double firstd = std::max(fq, 1.0);
    const double top = std::min(cfg->max_hz, 0.5 / tsamp);
    const double lastd = std::floor(top / p->bin_hz);
    const uint64_t first = firstd >= 4e9 ? 4000000000ull : (uint64_t)firstd;
    uint64_t last = lastd >= 4e9 ? 4000000000ull : (lastd > 0 ? (uint64_t)lastd : 0);
    last = std::min<uint64_t>(last, g.nbins - 2);
    p->first = (uint32_t)std::min<uint64_t>(first, 4000000000ull);
    p->last = (uint32_t)last;
    const uint32_t maxh = cfg->max_harmonics ? cfg->max_harmonics : 16;
    p->nlevels = 0;
    for (int l = 0; l < 5; l++) {
        const uint64_t h = 1ull << l;
        if (h > maxh) break;
        const uint64_t end = std::min<uint64_t>(last, (uint64_t)(g.nbins - 2) / h);
        if (end <= first + 1) break;  // lib.rs:79-80 (later levels end earlier)
        p->end_h[l] = (uint32_t)end;
        p->nlevels = l + 1;
    }
    p->max_candidates = (uint32_t)std::min<uint64_t>(cfg->max_candidates, 1u << 30);
    p->peak_cap = PEAK_CAP;
    p->fp16 = cfg->precision == PSR_PRECISION_FP16;
    birdie_bins->clear();
    for (uint64_t b = 0; b < cfg->nbirdies; b++) {
        const double cen = cfg->birdies[2 * b], w = cfg->birdies[2 * b + 1];
        auto hit = [&](int64_t i) { return std::fabs((double)i * p->bin_hz - cen) <= w; };
        const int64_t nb = g.nbins;
        int64_t ic = (int64_t)std::floor(cen / p->bin_hz);
        ic = std::min<int64_t>(std::max<int64_t>(ic, 0), nb - 1);
        int64_t seed = -1;
        for (int64_t i = std::max<int64_t>(ic - 1, 0); i <= std::min<int64_t>(ic + 2, nb - 1); i++)
            if (hit(i)) {
                seed = i;
                break;
            }
        if (seed < 0) continue;
        int64_t lo = 0, hi = seed;  // smallest hit in [0, seed]
        while (lo < hi) {
            const int64_t mid = (lo + hi) / 2;
            if (hit(mid)) hi = mid;
            else lo = mid + 1;
        }
        int64_t a = seed, z = nb - 1;  // largest hit in [seed, nb)
        while (a < z) {
            const int64_t mid = (a + z + 1) / 2;
            if (hit(mid)) a = mid;
            else z = mid - 1;
        }
        if (birdie_bins->size() / 2 >= (size_t)MAX_BIRDIE_INTERVALS)
            return fail(PSR_E_INVALID, "more than 64 birdie intervals");
        birdie_bins->push_back((uint32_t)lo);
        birdie_bins->push_back((uint32_t)a);
    }
    p->nbirdie = (uint32_t)(birdie_bins->size() / 2);
    return PSR_OK;
}

namespace {

// A trial-series slot: dedispersed series, their means/corrections and DMs.
struct Slot {
    DevBuf* series;
    DevBuf* mean32;
    DevBuf* corr16;
    DevBuf* dms;
};
Slot slot_of(psr_ctx* c, int i) {
    return i == 0 ? Slot{&c->series, &c->mean32, &c->corr16, &c->dms}
                  : Slot{&c->series_b, &c->mean32_b, &c->corr16_b, &c->dms_b};
}

}  // namespace

// Enqueue the spectral chain (FFT, whitening, harmonic sums, ranking) over `nt`
// series in b.series on stream `s`; the ranked candidates land in b.cands and
// the counts (nt + 1 words, the last one the peak-overflow count) in b.counts.
// Allocates nothing, copies nothing to the host and waits for nothing.  With
// `ev`, events ev[0..3] bracket FFT, whitening, harmonic sums and ranking.
int psr::enqueue_spectral(const SpecGeom& g, const SearchParams& p,
                          const std::vector<uint32_t>& birdies, const SpecBufs& b, uint32_t nt,
                          size_t cap_per_trial, hipStream_t s, hipEvent_t* ev) {
    const uint32_t npart = g.m1 == 1 ? 1u : g.m2 / 2 + 1;
    std::vector<uint32_t> bh(2 * MAX_BIRDIE_INTERVALS, 0);
    std::copy(birdies.begin(), birdies.end(), bh.begin());
    // counters cleared by a kernel (no memset nodes if a caller captures the stream)
    launch_fill_u32(s, b.peak_counts, nt, 0u);
    launch_fill_u32(s, b.counts, nt + 1, 0u);
    if (ev) HIP_TRY(hipEventRecord(ev[0], s));
    launch_fft(s, g, b.series, (int)p.fp16, b.mean32, b.corr16, nt, b.work, b.spec_t, b.fft_partial,
               b.twiddles, b.costabs);
    HIP_TRY(hipGetLastError());
    if (ev) HIP_TRY(hipEventRecord(ev[1], s));
    launch_whiten(s, g, b.spec_t, (int)p.fp16, nt, b.fft_partial, npart, b.raw_nat, b.blk_sorted,
                  b.floors, b.scales);
    HIP_TRY(hipGetLastError());
    if (ev) HIP_TRY(hipEventRecord(ev[2], s));
    launch_harmsum(s, g, p, b.raw_nat, b.scales, bh.data(), nt, b.peaks, b.peak_counts);
    HIP_TRY(hipGetLastError());
    if (ev) HIP_TRY(hipEventRecord(ev[3], s));
    launch_finalize(s, g, p, b.dms, nt, b.peaks, b.peak_counts, b.cands, (uint32_t)cap_per_trial,
                    b.counts);
    HIP_TRY(hipGetLastError());
    return PSR_OK;
}

size_t psr::spectral_work_bytes(const SpecGeom& g, uint32_t nt) {
    return g.m1 == 1 ? 8 : (size_t)fft_work_trials(g, nt) * g.m * sizeof(float2);
}

namespace {

// Enqueue the spectral chain over `nt` series of slot `sl` on stream `s` into
// the ctx's buffers, ending with the copy of the ranked candidates to host `out`
// and of the counts (nt + 1 words) to host `counts`.  Nothing waits: the caller
// synchronises.  With `ev`, events ev[0..4] bracket FFT, whitening, harmonic
// sums and ranking, ev[4] after the copies.
int run_spectral(psr_ctx* c, const SpecGeom& g, const SearchParams& p,
                 const std::vector<uint32_t>& birdies, const Slot& sl, uint32_t nt,
                 psr_candidate* out, size_t cap_per_trial, uint32_t* counts, hipStream_t s,
                 hipEvent_t* ev) {
    const size_t es = p.fp16 ? 2 : 4;
    const uint32_t npart = g.m1 == 1 ? 1u : g.m2 / 2 + 1;
    if (!fft_shape_supported(g))
        return fail(PSR_E_SHAPE, "no FFT kernels for a " + std::to_string(g.nfft) + "-point transform");
    int rc;
    if ((rc = upload_twiddles(c, g))) return rc;
    if ((rc = ensure(c, c->work, spectral_work_bytes(g, nt))) ||
        (rc = ensure(c, c->spec_t, (size_t)nt * g.spec_pitch * es)) ||
        (rc = ensure(c, c->raw_nat, (size_t)nt * g.spec_pitch * es)) ||
        (rc = ensure(c, c->scales, (size_t)nt * g.nblocks * sizeof(float))) ||
        (rc = ensure(c, c->blk_sorted, whiten_fused(g) ? 8 : (size_t)nt * g.spec_pitch * es)) ||
        (rc = ensure(c, c->floors, (size_t)nt * sizeof(float))) ||
        (rc = ensure(c, c->fft_partial, (size_t)nt * npart * sizeof(double))) ||
        (rc = ensure(c, c->peaks, (size_t)nt * PEAK_CAP * sizeof(RawPeak))) ||
        (rc = ensure(c, c->peak_counts, (size_t)nt * sizeof(uint32_t))) ||
        (rc = ensure(c, c->cands, (size_t)nt * cap_per_trial * sizeof(CandidateOut))) ||
        (rc = ensure(c, c->counts, (size_t)(nt + 1) * sizeof(uint32_t))))
        return rc;
    SpecBufs b;
    b.series = sl.series->p;
    b.mean32 = (const float*)sl.mean32->p;
    b.corr16 = (const float*)sl.corr16->p;
    b.dms = (const double*)sl.dms->p;
    b.work = (float2*)c->work.p;
    b.spec_t = c->spec_t.p;
    b.raw_nat = c->raw_nat.p;
    b.scales = (float*)c->scales.p;
    b.blk_sorted = c->blk_sorted.p;
    b.floors = (float*)c->floors.p;
    b.fft_partial = (double*)c->fft_partial.p;
    b.peaks = (RawPeak*)c->peaks.p;
    b.peak_counts = (uint32_t*)c->peak_counts.p;
    b.cands = (CandidateOut*)c->cands.p;
    b.counts = (uint32_t*)c->counts.p;
    b.twiddles = (const float2*)c->twiddles.p;
    b.costabs = (const float*)c->costabs.p;
    if ((rc = enqueue_spectral(g, p, birdies, b, nt, cap_per_trial, s, ev))) return rc;
    HIP_TRY(hipMemcpyAsync(counts, c->counts.p, (size_t)(nt + 1) * sizeof(uint32_t),
                           hipMemcpyDeviceToHost, s));
    HIP_TRY(hipMemcpyAsync(out, c->cands.p, (size_t)nt * cap_per_trial * sizeof(CandidateOut),
                           hipMemcpyDeviceToHost, s));
    if (ev) HIP_TRY(hipEventRecord(ev[4], s));
    return PSR_OK;
}

int ensure_pinned(psr_ctx* c, size_t bytes) {
    if (c->pinned_bytes >= bytes) return PSR_OK;
    if (c->pinned) (void)hipHostFree(c->pinned);
    c->pinned = nullptr;
    c->pinned_bytes = 0;
    hipError_t e = hipHostMalloc(&c->pinned, bytes, hipHostMallocDefault);
    if (e != hipSuccess) {
        c->pinned = nullptr;
        return fail(PSR_E_OOM, std::string("hipHostMalloc failed: ") + hipGetErrorString(e));
    }
    c->pinned_bytes = bytes;
    return PSR_OK;
}

// hipEvent_t array owned for the duration of one call
struct EventSet {
    std::vector<hipEvent_t> ev;
    ~EventSet() {
        for (hipEvent_t e : ev)
            if (e) (void)hipEventDestroy(e);
    }
    int create(size_t n) {
        ev.assign(n, nullptr);
        for (size_t i = 0; i < n; i++) HIP_TRY(hipEventCreate(&ev[i]));
        return PSR_OK;
    }
};

float elapsed(hipEvent_t a, hipEvent_t b) {
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, a, b) != hipSuccess) return 0.f;
    return ms;
}

}  // namespace

// ===================================================================== ABI
extern "C" {

const char* psr_version(void) { return "pulsarsr_amd 0.1.0 (gfx950)"; }
const char* psr_last_error(void) { return last_error(); }

int psr_ctx_create(int device, uint32_t flags, psr_ctx** out) {
    clear_error();
    if (!out) return fail(PSR_E_INVALID, "null output pointer");
    if (flags & ~(uint32_t)PSR_CTX_OVERLAP) return fail(PSR_E_INVALID, "unknown context flags");
    int ndev = 0;
    hipError_t e = hipGetDeviceCount(&ndev);
    if (e != hipSuccess || ndev <= 0) return fail(PSR_E_DEVICE, "no HIP device available");
    if (device < 0 || device >= ndev) return fail(PSR_E_DEVICE, "device index out of range");
    hipDeviceProp_t prop;
    HIP_TRY(hipGetDeviceProperties(&prop, device));
    if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0)
        return fail(PSR_E_DEVICE, std::string("device is ") + prop.gcnArchName + ", library is built for gfx950");
    psr_ctx* c = new psr_ctx();
    c->device = device;
    c->flags = flags;
    e = hipSetDevice(device);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&c->stream2, hipStreamNonBlocking);
    for (int i = 0; e == hipSuccess && i < 8; i++) e = hipEventCreate(&c->ev[i]);
    if (e != hipSuccess) {
        delete c;
        return fail(PSR_E_DEVICE, std::string("stream/event creation failed: ") + hipGetErrorString(e));
    }
    *out = c;
    return PSR_OK;
}

void psr_ctx_destroy(psr_ctx* c) {
    if (!c) return;
    (void)hipSetDevice(c->device);
    (void)hipStreamSynchronize(c->stream);
    if (c->stream2) (void)hipStreamSynchronize(c->stream2);
    DevBuf* all[] = {&c->delays, &c->dd_table, &c->dms, &c->series, &c->dd_partial, &c->mean32,
                     &c->series_b, &c->mean32_b, &c->corr16_b, &c->dms_b,
                     &c->corr16, &c->work, &c->spec_t, &c->raw_nat, &c->scales, &c->blk_sorted,
                     &c->floors, &c->fft_partial, &c->peaks, &c->peak_counts, &c->cands,
                     &c->counts, &c->twiddles, &c->costabs, &c->birdies,
                     &c->fold_series, &c->fold_per, &c->fold_off, &c->fold_sums, &c->fold_counts,
                     &c->an_values, &c->an_means, &c->an_resp};
    for (DevBuf* b : all) release(c, *b);
    for (int i = 0; i < 8; i++)
        if (c->ev[i]) (void)hipEventDestroy(c->ev[i]);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    if (c->stream2) (void)hipStreamDestroy(c->stream2);
    if (c->pinned) (void)hipHostFree(c->pinned);
    delete c;
}

int psr_ctx_set_memory_cap(psr_ctx* c, uint64_t bytes) {
    if (!c) return fail(PSR_E_INVALID, "null context");
    c->mem_cap = bytes;
    return PSR_OK;
}

int psr_delays(const double* freqs, size_t nch, double tsamp, double dm, uint64_t* out) {
    clear_error();
    if ((!freqs && nch) || !out) return fail(PSR_E_INVALID, "null pointer");
    return dedisperse::delays(freqs, nch, tsamp, dm, out);
}
int psr_dm_step(const double* freqs, size_t nch, double tsamp, double* out) {
    clear_error();
    if ((!freqs && nch) || !out) return fail(PSR_E_INVALID, "null pointer");
    return dedisperse::dm_step(freqs, nch, tsamp, out);
}
int psr_plan(double min, double max, double step, double* out, size_t cap, size_t* count) {
    clear_error();
    if (!count) return fail(PSR_E_INVALID, "null pointer");
    return dedisperse::plan(min, max, step, out, cap, count);
}
int psr_frequencies(double fch1, double foff, size_t nch, double* out) {
    clear_error();
    if (!out && nch) return fail(PSR_E_INVALID, "null pointer");
    for (size_t c = 0; c < nch; c++) out[c] = fch1 + (double)c * foff;  // fbank/src/lib.rs:42
    return PSR_OK;
}
int psr_parse_header(const uint8_t* bytes, size_t len, psr_header* out) {
    clear_error();
    if (!bytes || !out) return fail(PSR_E_INVALID, "null pointer");
    return fbank::parse_header(bytes, len, out);
}
int psr_sift(const psr_candidate* in, size_t n, double resolution_hz, size_t limit,
             int reject_dm_zero, psr_candidate* out, size_t* count) {
    clear_error();
    if ((!in && n) || !out || !count) return fail(PSR_E_INVALID, "null pointer");
    return search::sift(in, n, resolution_hz, limit, reject_dm_zero, out, count);
}

// ------------------------------------------------------------ observations
static int obs_finish(psr_ctx* c, psr_obs* o) {
    // channel frequencies and 1/(f*f) (fbank/src/lib.rs:40-44; dedisperse/src/lib.rs:21-26)
    o->freqs.resize(o->nchans);
    for (uint32_t ch = 0; ch < o->nchans; ch++) o->freqs[ch] = o->fch1 + (double)ch * o->foff;
    double fmax = 0.0;
    for (double f : o->freqs) {
        if (!std::isfinite(f) || f <= 0.0) return fail(PSR_E_FORMAT, "invalid or unsupported filterbank: nonpositive frequency");
        fmax = std::fmax(fmax, f);
    }
    o->ref_inv = 1.0 / (fmax * fmax);
    std::vector<double> inv(o->nchans);
    for (uint32_t ch = 0; ch < o->nchans; ch++) inv[ch] = 1.0 / (o->freqs[ch] * o->freqs[ch]);
    HIP_TRY(hipMalloc((void**)&o->inv_f2, o->nchans * sizeof(double)));
    HIP_TRY(hipMemcpy(o->inv_f2, inv.data(), o->nchans * sizeof(double), hipMemcpyHostToDevice));
    return obs_storage_scale(c, o);
}

}  // extern "C"

// FP16 storage offset/scale of an observation: sum of channel means (rounded to
// an integer for integer samples) and sqrt of the summed channel variances.
int psr::obs_storage_scale(psr_ctx* c, psr_obs* o) {
    double *dsum = nullptr, *dsq = nullptr;
    HIP_TRY(hipMalloc((void**)&dsum, o->nchans * sizeof(double)));
    HIP_TRY(hipMalloc((void**)&dsq, o->nchans * sizeof(double)));
    launch_chan_stats(c->stream, o->data, o->kind, o->nchans, o->nsamples, o->pitch, dsum, dsq);
    std::vector<double> hs(o->nchans), hq(o->nchans);
    HIP_TRY(hipMemcpyAsync(hs.data(), dsum, o->nchans * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipMemcpyAsync(hq.data(), dsq, o->nchans * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    (void)hipFree(dsum);
    (void)hipFree(dsq);
    double m = 0.0, v = 0.0;
    for (uint32_t ch = 0; ch < o->nchans; ch++) {
        const double mu = hs[ch] / (double)o->nsamples;
        m += mu;
        v += std::max(hq[ch] / (double)o->nsamples - mu * mu, 0.0);
    }
    o->m0 = o->kind == KIND_F32 ? m : std::round(m);
    o->s0 = v > 1e-12 ? std::sqrt(v) : 1.0;
    return PSR_OK;
}

extern "C" {

int psr_obs_create(psr_ctx* c, const psr_filterbank* fb, psr_obs** out) {
    clear_error();
    if (!c || !fb || !out || !fb->payload) return fail(PSR_E_INVALID, "null pointer");
    int rc;
    if ((rc = check_device(c))) return rc;
    const uint32_t nb = fb->nbits;
    if (!(nb == 1 || nb == 2 || nb == 4 || nb == 8 || nb == 16 || nb == 32) || fb->nchans == 0 ||
        !(fb->tsamp > 0.0) || !std::isfinite(fb->tsamp))
        return fail(PSR_E_FORMAT, "invalid or unsupported filterbank: invalid nchans/tsamp/nbits");
    const uint64_t frame_bits = (uint64_t)fb->nchans * nb;
    if ((fb->payload_bytes * 8) % frame_bits != 0)
        return fail(PSR_E_FORMAT, "invalid or unsupported filterbank: truncated frame or packed padding unsupported");
    const uint64_t total = fb->payload_bytes * 8 / frame_bits;
    const uint64_t ns = fb->nsamples ? std::min(fb->nsamples, total) : total;
    if (ns < 2) return fail(PSR_E_FORMAT, "invalid or unsupported filterbank: need at least two samples");
    if (nb == 32) {  // fbank/src/lib.rs:185-187: reject non-finite samples
        const float* f = (const float*)fb->payload;
        for (uint64_t i = 0; i < ns * fb->nchans; i++)
            if (!std::isfinite(f[i]))
                return fail(PSR_E_FORMAT, "invalid or unsupported filterbank: nonfinite sample at " +
                                              std::to_string(i / fb->nchans) + "/" +
                                              std::to_string(i % fb->nchans));
    }
    psr_obs* o = new psr_obs();
    o->ctx = c;
    o->nchans = fb->nchans;
    o->nbits = nb;
    o->nsamples = ns;
    o->tsamp = fb->tsamp;
    o->fch1 = fb->fch1;
    o->foff = fb->foff;
    o->kind = nb <= 8 ? KIND_U8 : (nb == 16 ? KIND_U16 : KIND_F32);
    o->maxval = nb <= 8 ? (1u << nb) - 1u : 0u;
    const size_t es = o->kind == KIND_U8 ? 1 : (o->kind == KIND_U16 ? 2 : 4);
    o->pitch = ((ns + 255) & ~(uint64_t)255) + 256;
    const size_t bytes = (size_t)o->pitch * o->nchans * es + (es == 1 ? DD_ALLOC_SLACK : 0);
    const size_t pbytes = (size_t)((ns * frame_bits + 7) / 8);
    void* dpay = nullptr;
    hipError_t e = hipMalloc(&o->data, bytes);
    if (e == hipSuccess) e = hipMalloc(&dpay, pbytes + 16);
    if (e == hipSuccess) e = hipMemsetAsync(o->data, 0, bytes, c->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(dpay, fb->payload, pbytes, hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess) {
        launch_ingest(c->stream, (const uint8_t*)dpay, (int)nb, o->nchans, ns, o->data, o->pitch, o->kind);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (dpay) (void)hipFree(dpay);
    if (e != hipSuccess) {
        psr_obs_destroy(o);
        return fail(e == hipErrorOutOfMemory ? PSR_E_OOM : PSR_E_DEVICE,
                    std::string("observation upload failed: ") + hipGetErrorString(e));
    }
    if ((rc = obs_finish(c, o))) {
        psr_obs_destroy(o);
        return rc;
    }
    *out = o;
    return PSR_OK;
}

int psr_obs_create_f32(psr_ctx* c, const float* data, uint64_t nsamples, uint32_t nchans,
                       double tsamp, double fch1, double foff, psr_obs** out) {
    clear_error();
    if (!c || !data || !out) return fail(PSR_E_INVALID, "null pointer");
    if (nsamples == 0 || nchans == 0) return fail(PSR_E_SHAPE, "invalid dedispersion input: array shape mismatch");
    if (!(tsamp > 0.0) || !std::isfinite(tsamp)) return fail(PSR_E_INVALID, "invalid dedispersion input: positive frequencies/tsamp and nonnegative DM required");
    int rc;
    if ((rc = check_device(c))) return rc;
    // Integer samples in [0, 255] sum exactly in any order: use the u8 layout.
    const uint64_t total = nsamples * nchans;
    bool small_int = true;
    float vmax = 0.f;
    for (uint64_t i = 0; i < total && small_int; i++) {
        const float v = data[i];
        small_int = v >= 0.f && v <= 255.f && v == std::floor(v);
        vmax = std::max(vmax, v);
    }
    psr_obs* o = new psr_obs();
    o->ctx = c;
    o->nchans = nchans;
    o->nbits = 32;
    o->nsamples = nsamples;
    o->tsamp = tsamp;
    o->fch1 = fch1;
    o->foff = foff;
    o->kind = small_int ? KIND_U8 : KIND_F32;
    o->maxval = small_int ? std::max<uint32_t>(1u, (uint32_t)vmax) : 0u;
    const size_t es = small_int ? 1 : 4;
    o->pitch = ((nsamples + 255) & ~(uint64_t)255) + 256;
    const size_t bytes = (size_t)o->pitch * nchans * es + (small_int ? DD_ALLOC_SLACK : 0);
    hipError_t e = hipMalloc(&o->data, bytes);
    if (e == hipSuccess) e = hipMemset(o->data, 0, bytes);
    if (e == hipSuccess) {
        if (small_int) {
            std::vector<uint8_t> row(nsamples);
            for (uint32_t ch = 0; ch < nchans && e == hipSuccess; ch++) {
                for (uint64_t t = 0; t < nsamples; t++) row[t] = (uint8_t)data[ch * nsamples + t];
                e = hipMemcpy((uint8_t*)o->data + ch * o->pitch, row.data(), nsamples, hipMemcpyHostToDevice);
            }
        } else {
            e = hipMemcpy2D(o->data, o->pitch * 4, data, nsamples * 4, nsamples * 4, nchans,
                            hipMemcpyHostToDevice);
        }
    }
    if (e != hipSuccess) {
        psr_obs_destroy(o);
        return fail(e == hipErrorOutOfMemory ? PSR_E_OOM : PSR_E_DEVICE,
                    std::string("observation upload failed: ") + hipGetErrorString(e));
    }
    if ((rc = obs_finish(c, o))) {
        psr_obs_destroy(o);
        return rc;
    }
    *out = o;
    return PSR_OK;
}

void psr_obs_destroy(psr_obs* o) {
    if (!o) return;
    if (o->data) (void)hipFree(o->data);
    if (o->inv_f2) (void)hipFree(o->inv_f2);
    delete o;
}

int psr_obs_download_f32(psr_obs* o, float* out) {
    clear_error();
    if (!o || !out) return fail(PSR_E_INVALID, "null pointer");
    const size_t es = o->kind == KIND_U8 ? 1 : (o->kind == KIND_U16 ? 2 : 4);
    std::vector<uint8_t> row(o->nsamples * es);
    for (uint32_t ch = 0; ch < o->nchans; ch++) {
        HIP_TRY(hipMemcpy(row.data(), (uint8_t*)o->data + (size_t)ch * o->pitch * es, o->nsamples * es,
                          hipMemcpyDeviceToHost));
        for (uint64_t t = 0; t < o->nsamples; t++) {
            float v;
            if (o->kind == KIND_U8) v = row[t];
            else if (o->kind == KIND_U16) v = ((uint16_t*)row.data())[t];
            else v = ((float*)row.data())[t];
            out[(size_t)ch * o->nsamples + t] = v;
        }
    }
    return PSR_OK;
}

// ------------------------------------------------------------ dedispersion
// Device buffers for dedispersing up to `nt` trials into slot `sl`.
static int ensure_dedisp(psr_ctx* c, psr_obs* o, const SpecGeom& g, const Slot& sl, uint32_t nt,
                         int fp16) {
    const size_t es = fp16 ? 2 : 4;
    const uint32_t ntiles_f32 = (g.n + 255) / 256;
    int rc;
    if ((rc = ensure(c, c->delays, (size_t)nt * o->nchans * sizeof(int32_t))) ||
        (rc = ensure(c, *sl.series, (size_t)nt * g.series_pitch * es + 64)) ||
        (rc = ensure(c, c->dd_partial, (size_t)nt * ntiles_f32 * sizeof(double))) ||
        (rc = ensure(c, *sl.mean32, (size_t)nt * sizeof(float))) ||
        (rc = ensure(c, *sl.corr16, (size_t)nt * sizeof(float))) ||
        (rc = ensure(c, *sl.dms, (size_t)nt * sizeof(double))))
        return rc;
    if (o->kind == KIND_U8 &&
        (rc = ensure(c, c->dd_table, dedisp_u8_table_words(o->nchans, nt) * sizeof(int32_t))))
        return rc;
    return PSR_OK;
}

// Enqueue on `s` the dedispersion of `nt` trials (ascending DMs already in
// sl.dms, the same values in host `dms_host`) into sl.series at
// `g.series_pitch`, with per-trial means for the FFT input.  Buffers must have
// been sized by ensure_dedisp.  With `ev`, events ev[0..2] bracket the delay
// table and the dedispersion.
static int dedisperse_into(psr_ctx* c, psr_obs* o, const SpecGeom& g, const Slot& sl,
                           const double* dms_host, uint32_t nt, int fp16, hipStream_t s,
                           hipEvent_t* ev) {
    const uint32_t n = g.n;
    const uint32_t ntiles_u8 = (n + 511) / 512, ntiles_f32 = (n + 255) / 256;
    if (ev) HIP_TRY(hipEventRecord(ev[0], s));
    launch_delays(s, o->inv_f2, o->nchans, o->ref_inv, o->tsamp, (const double*)sl.dms->p, nt,
                  (int32_t*)c->delays.p);
    HIP_TRY(hipGetLastError());
    if (ev) HIP_TRY(hipEventRecord(ev[1], s));
    const double inv_s0 = 1.0 / o->s0;
    bool int_done = false;
    if (o->kind == KIND_U8) {
        const std::vector<uint32_t> spread = spread_bounds(o, dms_host, nt, DEDISP_BLOCK_TRIALS);
        int_done = launch_dedisp_u8(s, (const uint8_t*)o->data, o->pitch, o->nchans, o->maxval,
                                    (const int32_t*)c->delays.p, (int32_t*)c->dd_table.p, nt, n,
                                    sl.series->p, g.series_pitch, fp16, (float)o->m0, (float)inv_s0,
                                    (double*)c->dd_partial.p, spread.data(), max_delay_at(o, dms_host[nt - 1]));
    }
    if (!int_done) {
        launch_dedisp_f32(s, o->data, o->kind, o->pitch, o->nchans, (const int32_t*)c->delays.p, nt,
                          n, sl.series->p, g.series_pitch, fp16, (float)o->m0, (float)inv_s0,
                          (double*)c->dd_partial.p);
    }
    const uint32_t ntiles = int_done ? ntiles_u8 : ntiles_f32;
    HIP_TRY(hipGetLastError());
    launch_trial_mean(s, (const double*)c->dd_partial.p, ntiles, nt, n, o->m0, inv_s0,
                      (float*)sl.mean32->p, (float*)sl.corr16->p);
    HIP_TRY(hipGetLastError());
    if (ev) HIP_TRY(hipEventRecord(ev[2], s));
    return PSR_OK;
}

int psr_dedisperse(psr_ctx* c, psr_obs* o, const double* dms, size_t ntrials, size_t length,
                   float* out, size_t* out_len) {
    clear_error();
    if (!c || !o || (!dms && ntrials) || !out_len) return fail(PSR_E_INVALID, "null pointer");
    int rc;
    if ((rc = check_device(c))) return rc;
    if (ntrials == 0) {
        *out_len = 0;
        return PSR_OK;
    }
    size_t n;
    uint64_t maxd;
    if ((rc = resolve_length(o, dms, ntrials, length, &n, &maxd, false))) return rc;
    *out_len = n;
    if (!out) return PSR_OK;
    if (n <= ((size_t)1 << 23)) return dedisperse_to_host(c, o, dms, ntrials, n, out);
    // longer series than a search takes: one trial at a time in output segments
    std::vector<uint64_t> d(o->nchans);
    for (size_t i = 0; i < ntrials; i++) {
        if ((rc = dedisperse::delays(o->freqs.data(), o->nchans, o->tsamp, dms[i], d.data()))) return rc;
        if ((rc = apply_table_to_host(c, o, 0, o->nchans, d.data(), n, out + i * n))) return rc;
    }
    return PSR_OK;
}

int psr_apply(psr_ctx* c, psr_obs* o, const uint64_t* delays, uint32_t ch_lo, uint32_t ch_hi,
              size_t length, float* out, size_t* out_len) {
    clear_error();
    if (!c || !o || !out_len || (!delays && ch_hi > ch_lo)) return fail(PSR_E_INVALID, "null pointer");
    // apply()'s checks in its order (crates/dedisperse/src/lib.rs:68-78)
    if (ch_hi <= ch_lo || ch_hi > o->nchans || o->nsamples == 0)
        return fail(PSR_E_SHAPE, "invalid dedispersion input: array shape mismatch");
    const uint32_t nch = ch_hi - ch_lo;
    const uint64_t maxd = *std::max_element(delays, delays + nch);
    if (maxd > o->nsamples) return fail(PSR_E_INVALID, "invalid dedispersion input: delay exceeds observation");
    const uint64_t available = o->nsamples - maxd;
    const uint64_t n = length ? length : available;
    if (n < 2 || n > available) return fail(PSR_E_SHAPE, "invalid dedispersion input: invalid output length");
    if (n > (uint64_t)INT32_MAX || maxd > (uint64_t)INT32_MAX)
        return fail(PSR_E_SHAPE, "series or delays past 2^31 samples are not supported");
    *out_len = (size_t)n;
    if (!out) return PSR_OK;
    int rc;
    if ((rc = check_device(c))) return rc;
    return apply_table_to_host(c, o, ch_lo, nch, delays, (size_t)n, out);
}

}  // extern "C"

// out[t] = sum over c < nch, in channel order, of row (ch_lo + c)[t + delays[c]]
// for t < n (f32 adds, as dedisperse::apply), in output segments of 2^24.
// The caller has checked the delays and n against the observation.
int psr::apply_table_to_host(psr_ctx* c, psr_obs* o, uint32_t ch_lo, uint32_t nch,
                             const uint64_t* delays, size_t n, float* out) {
    constexpr size_t SEG = (size_t)1 << 24;
    const size_t seg = std::min(n, SEG);
    const size_t es = sample_bytes(o);
    int rc;
    if ((rc = ensure(c, c->delays, (size_t)nch * sizeof(int32_t))) ||
        (rc = ensure(c, c->series, seg * sizeof(float) + 64)) ||
        (rc = ensure(c, c->dd_partial, ((seg + 255) / 256) * sizeof(double))))
        return rc;
    std::vector<int32_t> d32(nch);
    for (uint32_t i = 0; i < nch; i++) d32[i] = (int32_t)delays[i];
    HIP_TRY(hipMemcpyAsync(c->delays.p, d32.data(), nch * sizeof(int32_t), hipMemcpyHostToDevice,
                           c->stream));
    for (size_t t0 = 0; t0 < n; t0 += seg) {
        const uint32_t len = (uint32_t)std::min(seg, n - t0);
        const char* base = (const char*)o->data + ((size_t)ch_lo * o->pitch + t0) * es;
        launch_dedisp_f32(c->stream, base, o->kind, o->pitch, nch, (const int32_t*)c->delays.p, 1, len,
                          c->series.p, len, 0, 0.f, 1.f, (double*)c->dd_partial.p);
        HIP_TRY(hipGetLastError());
        HIP_TRY(hipMemcpyAsync(out + t0, c->series.p, len * sizeof(float), hipMemcpyDeviceToHost,
                               c->stream));
        HIP_TRY(hipStreamSynchronize(c->stream));
    }
    return PSR_OK;
}

int psr::dedisperse_to_host(psr_ctx* c, psr_obs* o, const double* dms, size_t ntrials, size_t n,
                            float* out) {
    int rc;
    std::vector<size_t> order(ntrials);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](size_t a, size_t b) { return dms[a] < dms[b]; });
    std::vector<double> sorted(ntrials);
    for (size_t i = 0; i < ntrials; i++) sorted[i] = dms[order[i]];
    const SpecGeom g = make_geometry((uint32_t)n);
    // trials per launch: up to 1024, with at most 4 GiB of series
    const size_t chunk = std::max<size_t>(1, std::min<size_t>(1024, ((size_t)4 << 30) / ((size_t)g.series_pitch * 4)));
    std::vector<float> rowbuf;
    const Slot sl = slot_of(c, 0);
    for (size_t b = 0; b < ntrials; b += chunk) {
        const uint32_t nt = (uint32_t)std::min(chunk, ntrials - b);
        if ((rc = ensure_dedisp(c, o, g, sl, nt, 0))) return rc;
        HIP_TRY(hipMemcpy(sl.dms->p, sorted.data() + b, nt * sizeof(double), hipMemcpyHostToDevice));
        if ((rc = dedisperse_into(c, o, g, sl, sorted.data() + b, nt, 0, c->stream, nullptr))) return rc;
        rowbuf.resize((size_t)nt * g.series_pitch);
        HIP_TRY(hipMemcpyAsync(rowbuf.data(), c->series.p, rowbuf.size() * sizeof(float),
                               hipMemcpyDeviceToHost, c->stream));
        HIP_TRY(hipStreamSynchronize(c->stream));
        for (uint32_t i = 0; i < nt; i++)
            std::memcpy(out + order[b + i] * n, rowbuf.data() + (size_t)i * g.series_pitch, n * sizeof(float));
    }
    return PSR_OK;
}

extern "C" {

// ------------------------------------------------------------------ search
int psr_search(psr_ctx* c, psr_obs* o, const double* dms, size_t ntrials, size_t length,
               const psr_search_cfg* cfg, psr_candidate* out, size_t cap_per_trial,
               uint32_t* counts) {
    clear_error();
    if (!c || !o || !cfg || (!dms && ntrials) || (!out && ntrials) || (!counts && ntrials))
        return fail(PSR_E_INVALID, "null pointer");
    if (cap_per_trial == 0) return fail(PSR_E_INVALID, "cap_per_trial must be positive");
    if (cfg->precision != PSR_PRECISION_FP32 && cfg->precision != PSR_PRECISION_FP16)
        return fail(PSR_E_INVALID, "unknown precision");
    if (!(o->tsamp > 0.0) || !std::isfinite(o->tsamp))
        return fail(PSR_E_CONTRACT, "periodicity requires finite positive tsamp");
    int rc;
    if ((rc = check_device(c))) return rc;
    c->last = psr_stage_times{};
    if (ntrials == 0) return PSR_OK;
    size_t n;
    uint64_t maxd;
    if ((rc = resolve_length(o, dms, ntrials, length, &n, &maxd))) return rc;
    const SpecGeom g = make_geometry((uint32_t)n);
    SearchParams p;
    std::vector<uint32_t> birdies;
    if ((rc = make_params(g, o->tsamp, cfg, &p, &birdies))) return rc;
    std::vector<size_t> order(ntrials);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](size_t a, size_t b) { return dms[a] < dms[b]; });
    std::vector<double> sorted(ntrials);
    for (size_t i = 0; i < ntrials; i++) sorted[i] = dms[order[i]];
    // trials per chunk: keep the working set (the series slots, one set of
    // spectral buffers) within 128 GiB of the 288 GiB HBM -- fewer, longer
    // launches measured faster than 16 or 32 GiB chunks, and 4000 trials of
    // 2^22 samples then run as one launch in FP16 and in FP32
    const bool serial = (c->flags & PSR_CTX_OVERLAP) == 0;
    const size_t es = p.fp16 ? 2 : 4;
    const size_t per_trial = (serial ? 1 : 2) * (size_t)g.series_pitch * es +
                             3 * (size_t)g.spec_pitch * es + (size_t)PEAK_CAP * sizeof(RawPeak) +
                             cap_per_trial * sizeof(CandidateOut) + (size_t)o->nchans * 4 + 4096;
    // (half the free device memory, and half the ctx memory cap when one is set);
    // whole 64-trial dedispersion blocks
    size_t budget = (size_t)128 << 30;
    size_t free_b = 0, total_b = 0;
    if (hipMemGetInfo(&free_b, &total_b) == hipSuccess && free_b) {
        // the ctx's own buffers are reused, so they count as free for this call
        budget = std::min<size_t>(budget, (free_b + (size_t)c->mem_used) / 2);
    }
    if (c->mem_cap) budget = std::min<size_t>(budget, c->mem_cap / 2);
    size_t chunk = std::max<size_t>(1, budget / per_trial);
    if (chunk >= DEDISP_BLOCK_TRIALS) chunk -= chunk % DEDISP_BLOCK_TRIALS;
    chunk = std::min(chunk, ntrials);
    const size_t nchunks = (ntrials + chunk - 1) / chunk;
    hipStream_t sa = c->stream, sb = serial ? c->stream : c->stream2;
    // both slots and the spectral buffers are sized for the first (largest)
    // chunk before anything is queued, so no allocation happens while kernels run
    const Slot slots[2] = {slot_of(c, 0), slot_of(c, nchunks > 1 && !serial ? 1 : 0)};
    const uint32_t nt0 = (uint32_t)chunk;
    if ((rc = ensure_dedisp(c, o, g, slots[0], nt0, (int)p.fp16)) ||
        (rc = ensure_dedisp(c, o, g, slots[1], nt0, (int)p.fp16)))
        return rc;
    // pinned host staging: sorted DMs, then per slot the ranked candidates and counts
    const size_t cand_bytes = chunk * cap_per_trial * sizeof(psr_candidate);
    const size_t cnt_bytes = (chunk + 1) * sizeof(uint32_t);
    const size_t dm_bytes = (ntrials * sizeof(double) + 255) & ~(size_t)255;
    const size_t slot_bytes = (cand_bytes + cnt_bytes + 255) & ~(size_t)255;
    if ((rc = ensure_pinned(c, dm_bytes + 2 * slot_bytes))) return rc;
    double* hdms = (double*)c->pinned;
    for (size_t i = 0; i < ntrials; i++) hdms[i] = sorted[i];
    auto hcand = [&](int k) { return (psr_candidate*)((char*)c->pinned + dm_bytes + k * slot_bytes); };
    auto hcnt = [&](int k) { return (uint32_t*)((char*)hcand(k) + cand_bytes); };
    // events per slot: [0..2] dedispersion, [3..7] spectral; plus the start
    EventSet evs;
    if ((rc = evs.create(17))) return rc;
    hipEvent_t* ev[2] = {evs.ev.data(), evs.ev.data() + 8};
    hipEvent_t ev_start = evs.ev[16];
    uint32_t overflow = 0;
    double t_delay = 0, t_dd = 0, t_fft = 0, t_wh = 0, t_hs = 0, t_fin = 0;
    // wait for chunk k (slot k & 1), add its stage times and scatter its results
    auto drain = [&](size_t k) -> int {
        const int si = serial ? 0 : (int)(k & 1);
        hipEvent_t* e = ev[si];
        HIP_TRY(hipEventSynchronize(e[7]));
        t_delay += elapsed(e[0], e[1]);
        t_dd += elapsed(e[1], e[2]);
        t_fft += elapsed(e[3], e[4]);
        t_wh += elapsed(e[4], e[5]);
        t_hs += elapsed(e[5], e[6]);
        t_fin += elapsed(e[6], e[7]);
        const size_t b = k * chunk;
        const uint32_t nt = (uint32_t)std::min(chunk, ntrials - b);
        const uint32_t* cn = hcnt(si);
        const psr_candidate* cd = hcand(si);
        overflow += cn[nt];
        for (uint32_t i = 0; i < nt; i++) {
            const size_t dst = order[b + i];
            counts[dst] = cn[i];
            std::memcpy(out + dst * cap_per_trial, cd + (size_t)i * cap_per_trial,
                        cn[i] * sizeof(psr_candidate));
        }
        return PSR_OK;
    };
    HIP_TRY(hipEventRecord(ev_start, sa));
    for (size_t k = 0; k < nchunks; k++) {
        const int si = serial ? 0 : (int)(k & 1);
        // the slot's previous user (chunk k - 2, or k - 1 when serial) must be done
        if (serial && k >= 1 && (rc = drain(k - 1))) return rc;
        if (!serial && k >= 2 && (rc = drain(k - 2))) return rc;
        const size_t b = k * chunk;
        const uint32_t nt = (uint32_t)std::min(chunk, ntrials - b);
        const Slot& sl = slots[si];
        HIP_TRY(hipMemcpyAsync(sl.dms->p, hdms + b, nt * sizeof(double), hipMemcpyHostToDevice, sa));
        if ((rc = dedisperse_into(c, o, g, sl, sorted.data() + b, nt, (int)p.fp16, sa, ev[si]))) return rc;
        if (sb != sa) HIP_TRY(hipStreamWaitEvent(sb, ev[si][2], 0));
        if ((rc = run_spectral(c, g, p, birdies, sl, nt, hcand(si), cap_per_trial, hcnt(si), sb,
                               ev[si] + 3)))
            return rc;
    }
    for (size_t k = nchunks >= 2 && !serial ? nchunks - 2 : nchunks - 1; k < nchunks; k++)
        if ((rc = drain(k))) return rc;
    c->last.delays_ms = t_delay;
    c->last.dedisperse_ms = t_dd;
    c->last.fft_ms = t_fft;
    c->last.whiten_ms = t_wh;
    c->last.harmsum_ms = t_hs;
    c->last.finalize_ms = t_fin;
    c->last.total_ms = elapsed(ev_start, ev[serial ? 0 : (int)((nchunks - 1) & 1)][7]);
    c->last.trials = ntrials;
    c->last.series_len = n;
    c->last.fft_len = g.nfft;
    c->last.launches = nchunks;
    if (overflow)
        return fail(PSR_E_OVERFLOW, std::to_string(overflow) +
                                        " trial(s) produced more peaks than the per-trial buffer holds");
    return PSR_OK;
}

int psr_last_stage_times(const psr_ctx* c, psr_stage_times* out) {
    if (!c || !out) return fail(PSR_E_INVALID, "null pointer");
    *out = c->last;
    return PSR_OK;
}

// Upload one host series as trial 0 of c->series with its f64 mean computed the
// reference's way (sequential sum, crates/search/src/lib.rs:36).
static int upload_series(psr_ctx* c, const float* series, size_t n, const SpecGeom& g) {
    int rc;
    if ((rc = ensure(c, c->series, (size_t)g.series_pitch * 4 + 64)) ||
        (rc = ensure(c, c->mean32, sizeof(float))) || (rc = ensure(c, c->corr16, sizeof(float))) ||
        (rc = ensure(c, c->dms, sizeof(double))))
        return rc;
    double mean = 0.0;
    for (size_t i = 0; i < n; i++) mean += (double)series[i];
    mean /= (double)n;
    const float m32 = (float)mean, z = 0.f;
    HIP_TRY(hipMemcpy(c->series.p, series, n * sizeof(float), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(c->mean32.p, &m32, sizeof(float), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(c->corr16.p, &z, sizeof(float), hipMemcpyHostToDevice));
    return PSR_OK;
}

int psr_periodicity(psr_ctx* c, const float* series, size_t n, double tsamp, double dm,
                    const psr_search_cfg* cfg, psr_candidate* out, size_t cap, size_t* count) {
    clear_error();
    if (!c || !series || !cfg || !out || !count) return fail(PSR_E_INVALID, "null pointer");
    if (n < 2 || !std::isfinite(tsamp) || tsamp <= 0.0)  // lib.rs:34 assert!
        return fail(PSR_E_CONTRACT, "periodicity requires at least two samples and finite positive tsamp");
    if (n > (1u << 23)) return fail(PSR_E_SHAPE, "series longer than 2^23 samples is not supported");
    if (cap == 0) return fail(PSR_E_INVALID, "cap must be positive");
    int rc;
    if ((rc = check_device(c))) return rc;
    const SpecGeom g = make_geometry((uint32_t)n);
    psr_search_cfg cf = *cfg;
    cf.precision = PSR_PRECISION_FP32;
    SearchParams p;
    std::vector<uint32_t> birdies;
    if ((rc = make_params(g, tsamp, &cf, &p, &birdies))) return rc;
    if ((rc = upload_series(c, series, n, g))) return rc;
    HIP_TRY(hipMemcpy(c->dms.p, &dm, sizeof(double), hipMemcpyHostToDevice));
    std::vector<psr_candidate> hout(cap);
    uint32_t cnt[2] = {0, 0};
    if ((rc = run_spectral(c, g, p, birdies, slot_of(c, 0), 1, hout.data(), cap, cnt, c->stream,
                           nullptr)))
        return rc;
    HIP_TRY(hipStreamSynchronize(c->stream));
    std::memcpy(out, hout.data(), cnt[0] * sizeof(psr_candidate));
    *count = cnt[0];
    if (cnt[1]) return fail(PSR_E_OVERFLOW, "peak buffer overflow");
    return PSR_OK;
}

static int spectrum_split(psr_ctx* c, const float* series, size_t n, double tsamp, const psr_search_cfg* cfg,
                          uint32_t m2, float* raw, float* white, size_t* nbins) {
    clear_error();
    if (!c || !series || !cfg || !nbins) return fail(PSR_E_INVALID, "null pointer");
    if (n < 2 || !std::isfinite(tsamp) || tsamp <= 0.0)
        return fail(PSR_E_CONTRACT, "periodicity requires at least two samples and finite positive tsamp");
    if (n > (1u << 23)) return fail(PSR_E_SHAPE, "series longer than 2^23 samples is not supported");
    int rc;
    if ((rc = check_device(c))) return rc;
    SpecGeom g = make_geometry((uint32_t)n);
    if (m2) {  // the four-step split with m2 rows (what the batch autotuner may pick)
        if (m2 != 512 && m2 != 1024 && m2 != 2048) return fail(PSR_E_SHAPE, "m2 must be 0, 512, 1024 or 2048");
        g = make_geometry_split((uint32_t)n, m2);
        if (g.m1 == 1 || g.m2 != m2 || !fft_shape_supported(g))
            return fail(PSR_E_SHAPE, "no four-step split with these m2 rows for this length");
    }
    *nbins = g.nbins;
    if (!raw && !white) return PSR_OK;
    psr_search_cfg cf = *cfg;
    const bool f16 = cfg->precision == PSR_PRECISION_FP16;
    if (!f16) cf.precision = PSR_PRECISION_FP32;
    cf.threshold = INFINITY;  // no peaks needed
    SearchParams p;
    std::vector<uint32_t> birdies;
    if ((rc = make_params(g, tsamp, &cf, &p, &birdies))) return rc;
    if ((rc = upload_series(c, series, n, g))) return rc;
    if (f16) {  // FP16 storage: the mean-subtracted series rounded to half precision
        double mean = 0.0;
        for (size_t i = 0; i < n; i++) mean += (double)series[i];
        mean /= (double)n;
        std::vector<_Float16> h(n);
        for (size_t i = 0; i < n; i++) h[i] = (_Float16)(float)((double)series[i] - mean);
        HIP_TRY(hipMemcpy(c->series.p, h.data(), n * sizeof(_Float16), hipMemcpyHostToDevice));
    }
    const double zero = 0.0;
    HIP_TRY(hipMemcpy(c->dms.p, &zero, sizeof(double), hipMemcpyHostToDevice));
    psr_candidate dummy[1];
    uint32_t cnt[2] = {0, 0};
    if ((rc = run_spectral(c, g, p, birdies, slot_of(c, 0), 1, dummy, 1, cnt, c->stream, nullptr)))
        return rc;
    HIP_TRY(hipStreamSynchronize(c->stream));
    std::vector<float> r(g.nbins), sc(g.nblocks);
    if (f16) {  // stored amplitudes a; the search's power is a * a in f32
        std::vector<_Float16> a(g.nbins);
        HIP_TRY(hipMemcpy(a.data(), c->raw_nat.p, g.nbins * sizeof(_Float16), hipMemcpyDeviceToHost));
        for (uint32_t k = 0; k < g.nbins; k++) r[k] = (float)a[k] * (float)a[k];
    } else {
        HIP_TRY(hipMemcpy(r.data(), c->raw_nat.p, g.nbins * sizeof(float), hipMemcpyDeviceToHost));
    }
    HIP_TRY(hipMemcpy(sc.data(), c->scales.p, g.nblocks * sizeof(float), hipMemcpyDeviceToHost));
    if (raw) std::memcpy(raw, r.data(), g.nbins * sizeof(float));
    if (white) {
        for (uint32_t k = 0; k < g.nbins; k++) white[k] = r[k] / sc[k >> 6];
        white[0] = 0.0f;
        for (size_t b = 0; b + 1 < birdies.size(); b += 2)
            for (uint32_t k = birdies[b]; k <= birdies[b + 1]; k++) white[k] = 0.0f;
    }
    return PSR_OK;
}

int psr_spectrum(psr_ctx* c, const float* series, size_t n, double tsamp, const psr_search_cfg* cfg,
                 float* raw, float* white, size_t* nbins) {
    return spectrum_split(c, series, n, tsamp, cfg, 0, raw, white, nbins);
}

int psr_spectrum_split(psr_ctx* c, const float* series, size_t n, double tsamp, const psr_search_cfg* cfg,
                       uint32_t m2, float* raw, float* white, size_t* nbins) {
    return spectrum_split(c, series, n, tsamp, cfg, m2, raw, white, nbins);
}

// ------------------------------------------------------------------- fold
static inline int32_t total_key32(float f) {
    int32_t b;
    std::memcpy(&b, &f, 4);
    b ^= (int32_t)(((uint32_t)(b >> 31)) >> 1);
    return b;
}

// Profile statistics of crates/fold/src/lib.rs:32-79 from device bin sums.
static void profile_stats(const double* sums, const uint64_t* cnt, uint32_t bins, double smean,
                          double svar, double period, psr_profile_stats* st, float* values) {
    for (uint32_t b = 0; b < bins; b++) values[b] = cnt[b] > 0 ? (float)(sums[b] / (double)cnt[b]) : 0.0f;
    std::vector<float> v(values, values + bins);
    std::vector<int32_t> keys(bins);
    for (uint32_t b = 0; b < bins; b++) keys[b] = total_key32(values[b]);
    std::nth_element(keys.begin(), keys.begin() + bins / 2, keys.end());
    int32_t ck = keys[bins / 2];
    ck ^= (int32_t)(((uint32_t)(ck >> 31)) >> 1);
    float center;
    std::memcpy(&center, &ck, 4);
    std::vector<float> off;
    for (uint32_t b = 0; b < bins; b++)
        if (cnt[b] > 0) off.push_back(values[b]);
    std::sort(off.begin(), off.end(), [](float a, float b) { return total_key32(a) < total_key32(b); });
    const size_t keep = std::max<size_t>(off.size() * 3 / 4, 1);
    off.resize(std::min(keep, off.size()));
    double mean = 0.0;
    for (float x : off) mean += (double)x;
    mean /= (double)keep;
    double ss = 0.0;
    for (float x : off) {
        const double d = (double)x - mean;
        ss += d * d;
    }
    const double sigma = std::sqrt(ss / (double)keep);
    float peak = -INFINITY;
    for (uint32_t b = 0; b < bins; b++) peak = std::fmax(peak, values[b]);
    st->snr = sigma > 1e-12 ? ((double)peak - (double)center) / sigma : 0.0;
    uint32_t populated = 0;
    for (uint32_t b = 0; b < bins; b++) populated += cnt[b] > 0;
    double chi = 0.0;
    if (svar > 1e-12 && populated > 1) {
        for (uint32_t b = 0; b < bins; b++) {
            const double d = (double)values[b] - smean;
            chi += (double)cnt[b] * (d * d) / svar;
        }
        chi /= (double)(populated - 1);
    }
    st->reduced_chi_square = chi;
    st->period_s = period;
}

int psr_fold(psr_ctx* c, const float* series, size_t n, double tsamp, const double* periods,
             const uint64_t* offsets, size_t nprof, uint32_t bins, psr_profile_stats* stats,
             float* values, uint64_t* counts) {
    clear_error();
    if (!c || !series || !periods || !stats || !values || !counts) return fail(PSR_E_INVALID, "null pointer");
    if (bins < 4 || bins > 1024 || !(tsamp > 0.0) || !std::isfinite(tsamp))  // fold/src/lib.rs:23
        return fail(PSR_E_CONTRACT, "fold requires 4 <= bins <= 1024 and finite positive tsamp");
    for (size_t i = 0; i < nprof; i++)
        if (!(periods[i] > 0.0) || !std::isfinite(periods[i]))
            return fail(PSR_E_CONTRACT, "fold requires a finite positive period");
    if (n == 0 || nprof == 0) return PSR_OK;
    int rc;
    if ((rc = check_device(c))) return rc;
    if ((rc = ensure(c, c->fold_series, n * sizeof(float))) ||
        (rc = ensure(c, c->fold_per, nprof * sizeof(double))) ||
        (rc = ensure(c, c->fold_off, nprof * sizeof(uint64_t))) ||
        (rc = ensure(c, c->fold_sums, nprof * bins * sizeof(double))) ||
        (rc = ensure(c, c->fold_counts, nprof * bins * sizeof(uint64_t))))
        return rc;
    std::vector<uint64_t> offs(nprof, 0);
    if (offsets) std::copy(offsets, offsets + nprof, offs.begin());
    HIP_TRY(hipMemcpy(c->fold_series.p, series, n * sizeof(float), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(c->fold_per.p, periods, nprof * sizeof(double), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(c->fold_off.p, offs.data(), nprof * sizeof(uint64_t), hipMemcpyHostToDevice));
    launch_fold(c->stream, (const float*)c->fold_series.p, n, tsamp, (const double*)c->fold_per.p,
                (const uint64_t*)c->fold_off.p, (uint32_t)nprof, bins, (double*)c->fold_sums.p,
                (uint64_t*)c->fold_counts.p);
    HIP_TRY(hipGetLastError());
    std::vector<double> sums(nprof * bins);
    HIP_TRY(hipMemcpyAsync(sums.data(), c->fold_sums.p, sums.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipMemcpyAsync(counts, c->fold_counts.p, nprof * bins * sizeof(uint64_t), hipMemcpyDeviceToHost, c->stream));
    // series mean and variance, sequential f64 (fold/src/lib.rs:56-62)
    double smean = 0.0;
    for (size_t t = 0; t < n; t++) smean += (double)series[t];
    smean /= (double)n;
    double svar = 0.0;
    for (size_t t = 0; t < n; t++) {
        const double d = (double)series[t] - smean;
        svar += d * d;
    }
    svar /= (double)n;
    HIP_TRY(hipStreamSynchronize(c->stream));
    for (size_t i = 0; i < nprof; i++)
        profile_stats(sums.data() + i * bins, counts + i * bins, bins, smean, svar, periods[i],
                      stats + i, values + i * bins);
    return PSR_OK;
}

int psr_refine(psr_ctx* c, const float* series, size_t n, double tsamp, double frequency,
               double span_hz, size_t steps, uint32_t bins, psr_profile_stats* stats,
               float* values, uint64_t* counts) {
    clear_error();
    if (!(steps > 0 && frequency > span_hz && span_hz >= 0.0))  // fold/src/lib.rs:89
        return fail(PSR_E_CONTRACT, "refine requires steps > 0 and frequency > span >= 0");
    const size_t np = steps + 1;
    std::vector<double> per(np);
    for (size_t i = 0; i < np; i++) {
        const double f = frequency - span_hz + 2.0 * span_hz * (double)i / (double)steps;
        per[i] = 1.0 / f;
    }
    std::vector<psr_profile_stats> st(np);
    std::vector<float> v((size_t)np * bins);
    std::vector<uint64_t> cn((size_t)np * bins);
    int rc = psr_fold(c, series, n, tsamp, per.data(), nullptr, np, bins, st.data(), v.data(), cn.data());
    if (rc) return rc;
    size_t best = 0;  // Iterator::max_by keeps the last maximum
    for (size_t i = 1; i < np; i++) {
        int64_t a, b;
        double x = st[i].reduced_chi_square, y = st[best].reduced_chi_square;
        std::memcpy(&a, &x, 8);
        std::memcpy(&b, &y, 8);
        a ^= (int64_t)(((uint64_t)(a >> 63)) >> 1);
        b ^= (int64_t)(((uint64_t)(b >> 63)) >> 1);
        if (a >= b) best = i;
    }
    *stats = st[best];
    std::memcpy(values, v.data() + best * bins, bins * sizeof(float));
    std::memcpy(counts, cn.data() + best * bins, bins * sizeof(uint64_t));
    return PSR_OK;
}

}  // extern "C"
