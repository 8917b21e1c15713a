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
// recompute the FP16 storage offset/scale (m0, s0) from the device samples
int obs_storage_scale(psr_ctx* c, psr_obs* o);
// bytes per sample of an observation's channel rows
inline size_t sample_bytes(const psr_obs* o) {
    return o->kind == KIND_U8 ? 1 : (o->kind == KIND_U16 ? 2 : 4);
}
}  // namespace psr
