// psr_periodicity_noalloc: search::periodicity over one host series with no
// device allocation.  Every device buffer of the spectral chain (series,
// FFT intermediate, spectra, scales, peaks, candidates, twiddle tables) is
// carved out of one file-scope device array, `scratch`, whose device address
// is looked up once per device with hipGetSymbolAddress and used from then on.
//
// Limits: series of at most PSR_NOALLOC_MAX_SERIES samples, at most
// PSR_NOALLOC_MAX_CANDIDATES candidates.  Every call on a device shares that
// device's array and leaves its contents behind (the counters the kernels
// accumulate into are cleared by a fill kernel at the start of each call), so
// calls that overlap on one device race; calls on different devices do not.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <vector>

#include "ctx_internal.hpp"

using namespace psr;

namespace {

// Byte size of the layout below at the largest series (2^17 samples: a 2^18-point
// transform, 131073 bins) is about 2.8 MiB; 4 MiB leaves room for alignment.
constexpr size_t SCRATCH_BYTES = (size_t)4 << 20;
__device__ __align__(256) unsigned char scratch[SCRATCH_BYTES];

constexpr int MAX_DEVICES = 64;
std::mutex g_lock;
unsigned char* g_scratch[MAX_DEVICES] = {};

// Device address of `scratch` on the current device `device` (looked up once).
int scratch_address(int device, unsigned char** out) {
    if (device < 0 || device >= MAX_DEVICES) return fail(PSR_E_DEVICE, "device index out of range");
    std::lock_guard<std::mutex> lk(g_lock);
    if (!g_scratch[device]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return fail(PSR_E_DEVICE, std::string("hipGetSymbolAddress(scratch) failed: ") +
                                          hipGetErrorString(e));
        g_scratch[device] = (unsigned char*)p;
    }
    *out = g_scratch[device];
    return PSR_OK;
}

// Offsets of every buffer of one single-series spectral chain, 256-byte aligned.
struct Layout {
    size_t series, mean32, corr16, dms, work, spec_t, raw_nat, scales, blk_sorted, floors,
        fft_partial, peaks, peak_counts, cands, counts, twiddles, costabs, total;
};

Layout layout_of(const SpecGeom& g, size_t cap) {
    Layout L{};
    size_t off = 0;
    auto take = [&](size_t bytes) {
        const size_t at = off;
        off += (bytes + 255) & ~(size_t)255;
        return at;
    };
    const uint32_t npart = g.m1 == 1 ? 1u : g.m2 / 2 + 1;
    L.series = take((size_t)g.series_pitch * sizeof(float) + 64);
    L.mean32 = take(sizeof(float));
    L.corr16 = take(sizeof(float));
    L.dms = take(sizeof(double));
    L.work = take(spectral_work_bytes(g, 1));
    L.spec_t = take((size_t)g.spec_pitch * sizeof(float));
    L.raw_nat = take((size_t)g.spec_pitch * sizeof(float));
    L.scales = take((size_t)g.nblocks * sizeof(float));
    L.blk_sorted = take(whiten_fused(g) ? 8 : (size_t)g.spec_pitch * sizeof(float));
    L.floors = take(sizeof(float));
    L.fft_partial = take((size_t)npart * sizeof(double));
    L.peaks = take((size_t)PEAK_CAP * sizeof(RawPeak));
    L.peak_counts = take(sizeof(uint32_t));
    L.cands = take(cap * sizeof(CandidateOut));
    L.counts = take(2 * sizeof(uint32_t));
    L.twiddles = take(twiddle_table_size(g) * sizeof(float2));
    L.costabs = take((size_t)COSTAB_ALL * sizeof(float));
    L.total = off;
    return L;
}

}  // namespace

extern "C" int psr_periodicity_noalloc(psr_ctx* c, const float* series, size_t n, double tsamp,
                                       double dm, const psr_search_cfg* cfg, psr_candidate* out,
                                       size_t cap, size_t* count) {
    clear_error();
    if (!c || !series || !cfg || !out || !count) return fail(PSR_E_INVALID, "null pointer");
    if (n < 2 || !std::isfinite(tsamp) || tsamp <= 0.0)  // crates/search/src/lib.rs:34 assert!
        return fail(PSR_E_CONTRACT, "periodicity requires at least two samples and finite positive tsamp");
    if (n > PSR_NOALLOC_MAX_SERIES)
        return fail(PSR_E_SHAPE, "psr_periodicity_noalloc takes at most " +
                                     std::to_string(PSR_NOALLOC_MAX_SERIES) + " samples, got " +
                                     std::to_string(n));
    if (cap == 0 || cap > PSR_NOALLOC_MAX_CANDIDATES)
        return fail(PSR_E_INVALID, "cap must be between 1 and " + std::to_string(PSR_NOALLOC_MAX_CANDIDATES));
    int rc;
    if ((rc = check_device(c))) return rc;
    const SpecGeom g = make_geometry((uint32_t)n);
    if (!fft_shape_supported(g))
        return fail(PSR_E_SHAPE, "no FFT kernels for a " + std::to_string(g.nfft) + "-point transform");
    const Layout L = layout_of(g, cap);
    if (L.total > SCRATCH_BYTES)  // cannot happen within the limits above
        return fail(PSR_E_SHAPE, "scratch layout exceeds the static scratch array");
    psr_search_cfg cf = *cfg;
    cf.precision = PSR_PRECISION_FP32;
    SearchParams p;
    std::vector<uint32_t> birdies;
    if ((rc = make_params(g, tsamp, &cf, &p, &birdies))) return rc;
    unsigned char* base = nullptr;
    if ((rc = scratch_address(c->device, &base))) return rc;
    // host-side inputs: the series' f64 mean the reference's way (lib.rs:36), the
    // DM and the geometry's twiddle tables
    double mean = 0.0;
    for (size_t i = 0; i < n; i++) mean += (double)series[i];
    mean /= (double)n;
    const float head[2] = {(float)mean, 0.0f};  // mean32, corr16
    std::vector<float2> tw(twiddle_table_size(g));
    fill_twiddles(g, tw.data());
    std::vector<float> ct(COSTAB_ALL);
    fill_costabs(ct.data());
    hipStream_t s = c->stream;
    HIP_TRY(hipMemcpyAsync(base + L.series, series, n * sizeof(float), hipMemcpyHostToDevice, s));
    HIP_TRY(hipMemcpyAsync(base + L.mean32, &head[0], sizeof(float), hipMemcpyHostToDevice, s));
    HIP_TRY(hipMemcpyAsync(base + L.corr16, &head[1], sizeof(float), hipMemcpyHostToDevice, s));
    HIP_TRY(hipMemcpyAsync(base + L.dms, &dm, sizeof(double), hipMemcpyHostToDevice, s));
    HIP_TRY(hipMemcpyAsync(base + L.twiddles, tw.data(), tw.size() * sizeof(float2),
                           hipMemcpyHostToDevice, s));
    HIP_TRY(hipMemcpyAsync(base + L.costabs, ct.data(), ct.size() * sizeof(float),
                           hipMemcpyHostToDevice, s));
    SpecBufs b;
    b.series = base + L.series;
    b.mean32 = (const float*)(base + L.mean32);
    b.corr16 = (const float*)(base + L.corr16);
    b.dms = (const double*)(base + L.dms);
    b.work = (float2*)(base + L.work);
    b.spec_t = base + L.spec_t;
    b.raw_nat = base + L.raw_nat;
    b.scales = (float*)(base + L.scales);
    b.blk_sorted = base + L.blk_sorted;
    b.floors = (float*)(base + L.floors);
    b.fft_partial = (double*)(base + L.fft_partial);
    b.peaks = (RawPeak*)(base + L.peaks);
    b.peak_counts = (uint32_t*)(base + L.peak_counts);
    b.cands = (CandidateOut*)(base + L.cands);
    b.counts = (uint32_t*)(base + L.counts);
    b.twiddles = (const float2*)(base + L.twiddles);
    b.costabs = (const float*)(base + L.costabs);
    if ((rc = enqueue_spectral(g, p, birdies, b, 1, cap, s, nullptr))) return rc;
    uint32_t cnt[2] = {0, 0};
    HIP_TRY(hipMemcpyAsync(cnt, b.counts, sizeof(cnt), hipMemcpyDeviceToHost, s));
    HIP_TRY(hipMemcpyAsync(out, b.cands, cap * sizeof(psr_candidate), hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    *count = cnt[0];
    if (cnt[1]) return fail(PSR_E_OVERFLOW, "peak buffer overflow");
    return PSR_OK;
}
