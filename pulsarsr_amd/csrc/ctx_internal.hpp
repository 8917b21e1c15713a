// Internal view of the C ABI's context and observation objects, shared by the
// translation units that implement include/psr_gpu.h.
#pragma once
#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "../../include/psr_gpu.h"
#include "common.hpp"
#include "host_internal.hpp"

#define HIP_TRY(expr)                                                                    \
    do {                                                                                 \
        hipError_t e_ = (expr);                                                          \
        if (e_ != hipSuccess)                                                            \
            return psr::fail(e_ == hipErrorOutOfMemory ? PSR_E_OOM : PSR_E_DEVICE,       \
                             std::string("HIP error in ") + #expr + ": " + hipGetErrorString(e_)); \
    } while (0)

namespace psr {
struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
};
}  // namespace psr

struct psr_ctx {
    int device = 0;
    uint32_t flags = 0;
    hipStream_t stream = nullptr;
    uint64_t mem_cap = 0;
    uint64_t mem_used = 0;
    psr::DevBuf delays, dd_table, dms, series, dd_partial, mean32, corr16, work, spec_t, raw_nat,
        scales, blk_sorted, floors, fft_partial, peaks, peak_counts, cands, counts, twiddles, costabs,
        birdies, fold_series, fold_per, fold_off, fold_sums, fold_counts, an_values, an_means,
        an_resp;
    // second trial-series slot and stream: psr_search dedisperses chunk k+1 on
    // `stream` while chunk k's spectral stages run on `stream2`
    hipStream_t stream2 = nullptr;
    psr::DevBuf series_b, mean32_b, corr16_b, dms_b;
    void* pinned = nullptr;  // page-locked host staging (candidates, counts, DMs)
    size_t pinned_bytes = 0;
    psr::SpecGeom tw_geom{};
    bool tw_valid = false;
    hipEvent_t ev[8] = {};
    psr_stage_times last{};
};

struct psr_obs {
    psr_ctx* ctx = nullptr;
    int kind = psr::KIND_U8;
    uint32_t nchans = 0, nbits = 0, maxval = 0;
    uint64_t nsamples = 0, pitch = 0;
    double tsamp = 0, fch1 = 0, foff = 0;
    void* data = nullptr;
    double* inv_f2 = nullptr;  // device 1/(f*f) per channel
    std::vector<double> freqs;
    double ref_inv = 0;
    double m0 = 0, s0 = 1;  // FP16 storage offset and scale
};

namespace psr {
// device buffer growth within the ctx's memory cap
int ensure(psr_ctx* c, DevBuf& b, size_t bytes);
void release(psr_ctx* c, DevBuf& b);
int check_device(psr_ctx* c);
// FP32 dedispersion of `ntrials` DMs (any order) at output length n into host
// rows out[i * n ..]; the caller has validated n against the delays.
int dedisperse_to_host(psr_ctx* c, psr_obs* o, const double* dms, size_t ntrials, size_t n,
                       float* out);
// dedisperse::apply over channels [ch_lo, ch_lo + nch) with a caller delay
// table (one f32 trial in channel order) into host out[n]
int apply_table_to_host(psr_ctx* c, psr_obs* o, uint32_t ch_lo, uint32_t nch,
                        const uint64_t* delays, size_t n, float* out);
// apply()'s length rules over a set of DMs: output length n (length 0 = None)
// and the largest delay; fft_limit rejects series a search cannot take (> 2^23)
int resolve_length(const psr_obs* o, const double* dms, size_t ntrials, size_t length, size_t* n_out,
                   uint64_t* maxdelay_out, bool fft_limit = true);
// per-channel delay-spread bound over any `block` consecutive sorted trials
std::vector<uint32_t> spread_bounds(const psr_obs* o, const double* dms_sorted, size_t ntrials,
                                    size_t block);
// the largest channel delay at DM dm (delays grow with the DM: the last of a
// sorted list bounds them all)
uint64_t max_delay_at(const psr_obs* o, double dm);
// search::periodicity's bin limits and birdie bins for a geometry
int make_params(const SpecGeom& g, double tsamp, const psr_search_cfg* cfg, SearchParams* p,
                std::vector<uint32_t>* birdie_bins);
// device buffers of one run of the spectral chain
struct SpecBufs {
    const void* series = nullptr;  // [nt][series_pitch] f32 or f16
    const float* mean32 = nullptr;
    const float* corr16 = nullptr;
    const double* dms = nullptr;  // [nt]
    float2* work = nullptr;       // spectral_work_bytes
    void* spec_t = nullptr;       // [nt][spec_pitch]
    void* raw_nat = nullptr;      // [nt][spec_pitch]
    float* scales = nullptr;      // [nt][nblocks]
    void* blk_sorted = nullptr;   // [nt][spec_pitch] unless whiten_fused
    float* floors = nullptr;      // [nt]
    double* fft_partial = nullptr;  // [nt][m1 == 1 ? 1 : m2 / 2 + 1]
    RawPeak* peaks = nullptr;     // [nt][PEAK_CAP]
    uint32_t* peak_counts = nullptr;  // [nt]
    CandidateOut* cands = nullptr;    // [nt][cap_per_trial]
    uint32_t* counts = nullptr;       // [nt + 1]
    const float2* twiddles = nullptr;  // twiddle_table_size(g)
    const float* costabs = nullptr;    // COSTAB_ALL
};
size_t spectral_work_bytes(const SpecGeom& g, uint32_t nt);
// FFT, whitening, harmonic sums and ranking on stream s; allocates nothing
int enqueue_spectral(const SpecGeom& g, const SearchParams& p, const std::vector<uint32_t>& birdies,
                     const SpecBufs& b, uint32_t nt, size_t cap_per_trial, hipStream_t s,
                     hipEvent_t* ev);
// recompute the FP16 storage offset/scale (m0, s0) from the device samples
int obs_storage_scale(psr_ctx* c, psr_obs* o);
// bytes per sample of an observation's channel rows
inline size_t sample_bytes(const psr_obs* o) {
    return o->kind == KIND_U8 ? 1 : (o->kind == KIND_U16 ? 2 : 4);
}
}  // namespace psr
