// Phase folding for gfx950 (crates/fold/src/lib.rs:25-31) and the per-trial
// mean used by the FFT input (crates/search/src/lib.rs:36).
//
// Fold: one workgroup per (period, sample_offset) profile.  The bin of sample t
// is computed in f64 exactly as the reference does,
//   x = ((t + offset) * tsamp) / period;  phase = x - trunc(x);
//   b = min((phase * bins) as usize, bins - 1),
// so bin assignment is bit-identical.  Bin sums are f64 and counts u64; sums
// of integer-valued series are exact in any order.
#include <hip/hip_runtime.h>

#include "common.hpp"

namespace psr {

constexpr uint32_t FOLD_MAX_BINS = 1024;

__global__ __launch_bounds__(256) void k_fold(const float* __restrict__ series, uint64_t n,
                                              double tsamp, const double* __restrict__ periods,
                                              const uint64_t* __restrict__ offsets, uint32_t bins,
                                              double* __restrict__ sums,
                                              uint64_t* __restrict__ counts) {
    __shared__ double s_sum[FOLD_MAX_BINS];
    __shared__ unsigned long long s_cnt[FOLD_MAX_BINS];
    const uint32_t prof = blockIdx.x;
    const double period = periods[prof];
    const uint64_t off = offsets[prof];
    for (uint32_t b = threadIdx.x; b < bins; b += blockDim.x) {
        s_sum[b] = 0.0;
        s_cnt[b] = 0ull;
    }
    __syncthreads();
    const double fb = (double)bins;
    for (uint64_t t = threadIdx.x; t < n; t += blockDim.x) {
        const double x = (double)(t + off) * tsamp / period;
        const double phase = x - trunc(x);
        uint64_t b = (uint64_t)(phase * fb);
        if (b > bins - 1) b = bins - 1;
        atomicAdd(&s_sum[b], (double)series[t]);
        atomicAdd(&s_cnt[b], 1ull);
    }
    __syncthreads();
    for (uint32_t b = threadIdx.x; b < bins; b += blockDim.x) {
        sums[(size_t)prof * bins + b] = s_sum[b];
        counts[(size_t)prof * bins + b] = s_cnt[b];
    }
}

void launch_fold(hipStream_t s, const float* series, uint64_t n, double tsamp,
                 const double* periods, const uint64_t* offsets, uint32_t nprof, uint32_t bins,
                 double* sums, uint64_t* counts) {
    hipLaunchKernelGGL(k_fold, dim3(nprof), dim3(256), 0, s, series, n, tsamp, periods, offsets,
                       bins, sums, counts);
}

// Channel profiles of the confirmation stage (crates/search/src/confirmation.rs:391-396):
// one workgroup per channel row folds samples [0, n) of the observation's
// channel-major data (u8, u16 or f32) with fold::profile's bin rule.  Counts
// are the same for every row; row 0 writes them.
__global__ __launch_bounds__(256) void k_fold_rows(const void* __restrict__ data, int kind,
                                                   uint64_t pitch, uint64_t n, double tsamp,
                                                   double period, uint32_t bins,
                                                   double* __restrict__ sums,
                                                   uint64_t* __restrict__ counts) {
    __shared__ double s_sum[FOLD_MAX_BINS];
    __shared__ unsigned long long s_cnt[FOLD_MAX_BINS];
    const uint32_t row = blockIdx.x;
    for (uint32_t b = threadIdx.x; b < bins; b += blockDim.x) {
        s_sum[b] = 0.0;
        s_cnt[b] = 0ull;
    }
    __syncthreads();
    const double fb = (double)bins;
    const size_t base = (size_t)row * pitch;
    for (uint64_t t = threadIdx.x; t < n; t += blockDim.x) {
        const double x = (double)t * tsamp / period;
        const double phase = x - trunc(x);
        uint64_t b = (uint64_t)(phase * fb);
        if (b > bins - 1) b = bins - 1;
        const double v = kind == KIND_U8    ? (double)((const uint8_t*)data)[base + t]
                         : kind == KIND_U16 ? (double)((const uint16_t*)data)[base + t]
                                            : (double)((const float*)data)[base + t];
        atomicAdd(&s_sum[b], v);
        if (row == 0) atomicAdd(&s_cnt[b], 1ull);
    }
    __syncthreads();
    for (uint32_t b = threadIdx.x; b < bins; b += blockDim.x) {
        sums[(size_t)row * bins + b] = s_sum[b];
        if (row == 0) counts[b] = s_cnt[b];
    }
}

void launch_fold_rows(hipStream_t s, const void* data, int kind, uint64_t pitch, uint32_t nrows,
                      uint64_t n, double tsamp, double period, uint32_t bins, double* sums,
                      uint64_t* counts) {
    hipLaunchKernelGGL(k_fold_rows, dim3(nrows), dim3(256), 0, s, data, kind, pitch, n, tsamp,
                       period, bins, sums, counts);
}

// DM response of the confirmation stage (confirmation.rs:397-433): for every DM
// trial, the channel profiles shifted by their delay phase (linear
// interpolation) and summed in channel order in f64, then the on-window minus
// off-window mean.  One workgroup per trial, one thread per phase bin; the
// arithmetic is the reference's, operation for operation.
__global__ __launch_bounds__(256) void k_dm_response(
    const float* __restrict__ values, const double* __restrict__ means,
    const int32_t* __restrict__ delays, uint32_t nch, uint32_t bins, double tsamp, double period,
    uint32_t start, uint32_t width, double* __restrict__ resp) {
    __shared__ double comb[FOLD_MAX_BINS];
    const uint32_t trial = blockIdx.x;
    const int32_t* d = delays + (size_t)trial * nch;
    const double fb = (double)bins;
    for (uint32_t b = threadIdx.x; b < bins; b += blockDim.x) {
        double acc = 0.0;
        for (uint32_t c = 0; c < nch; c++) {
            const double shift = (double)d[c] * tsamp / period * fb;
            double pos = fmod((double)b + shift, fb);
            if (pos < 0.0) pos += fb;  // rem_euclid
            const uint32_t index = (uint32_t)floor(pos);
            const double fraction = pos - (double)index;
            const float* p = values + (size_t)c * bins;
            const double mean = means[c];
            acc += ((double)p[index] - mean) * (1.0 - fraction) +
                   ((double)p[(index + 1) % bins] - mean) * fraction;
        }
        comb[b] = acc;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double on = 0.0;
        for (uint32_t i = 0; i < width; i++) on += comb[(start + i) % bins];
        on /= (double)width;
        double tot = 0.0;
        for (uint32_t b = 0; b < bins; b++) tot += comb[b];
        const double off = (tot - on * (double)width) / (double)(bins - width);
        resp[trial] = on - off;
    }
}

void launch_dm_response(hipStream_t s, const float* values, const double* means,
                        const int32_t* delays, uint32_t nch, uint32_t ntrials, uint32_t bins,
                        double tsamp, double period, uint32_t start, uint32_t width,
                        double* resp) {
    hipLaunchKernelGGL(k_dm_response, dim3(ntrials), dim3(256), 0, s, values, means, delays, nch,
                       bins, tsamp, period, start, width, resp);
}

// mean32[t] = (sum/n) as f32 (FP32 path); corr16[t] = (m0 - sum/n) * inv_s0
// (FP16 path: the offset that turns stored (x - m0)/s0 into (x - mean)/s0).
// One wave per trial reduces the dedispersion kernel's per-tile partials in a
// fixed order (lane-strided sums, then a fixed butterfly), so the result is
// deterministic; for integer-valued series every partial is an exact integer.
__global__ __launch_bounds__(256) void k_trial_mean(const double* __restrict__ partial,
                                                    uint32_t ntiles, uint32_t ntrials, uint32_t n,
                                                    double m0, double inv_s0,
                                                    float* __restrict__ mean32,
                                                    float* __restrict__ corr16) {
    const uint32_t t = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (t >= ntrials) return;
    double s = 0.0;
    for (uint32_t i = lane; i < ntiles; i += 64) s += partial[(size_t)t * ntiles + i];
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o);
    if (lane == 0) {
        const double mean = s / (double)n;
        mean32[t] = (float)mean;
        corr16[t] = (float)((m0 - mean) * inv_s0);
    }
}

// The same for the trials of many observations (one job each range of trials).
__global__ __launch_bounds__(256) void k_trial_mean_multi(const DedispJob* __restrict__ jobs,
                                                          const uint32_t* __restrict__ trial_job,
                                                          uint32_t ntrials, float* __restrict__ mean32,
                                                          float* __restrict__ corr16) {
    const uint32_t t = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (t >= ntrials) return;
    const DedispJob& j = jobs[trial_job[t]];
    const uint32_t local = t - j.trial0;
    double s = 0.0;
    for (uint32_t i = lane; i < j.ntiles; i += 64) s += j.partial[(size_t)local * j.ntiles + i];
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o);
    if (lane == 0) {
        const double mean = s / (double)j.n;
        mean32[t] = (float)mean;
        corr16[t] = (float)((j.m0d - mean) * j.inv_s0d);
    }
}

void launch_trial_mean_multi(hipStream_t s, const DedispJob* jobs, const uint32_t* trial_job,
                             uint32_t ntrials, float* mean32, float* corr16) {
    hipLaunchKernelGGL(k_trial_mean_multi, dim3((ntrials + 3) / 4), dim3(256), 0, s, jobs, trial_job, ntrials,
                       mean32, corr16);
}

void launch_trial_mean(hipStream_t s, const double* partial, uint32_t ntiles, uint32_t ntrials,
                       uint32_t n, double m0, double inv_s0, float* mean32, float* corr16) {
    hipLaunchKernelGGL(k_trial_mean, dim3((ntrials + 3) / 4), dim3(256), 0, s, partial, ntiles,
                       ntrials, n, m0, inv_s0, mean32, corr16);
}

}  // namespace psr
