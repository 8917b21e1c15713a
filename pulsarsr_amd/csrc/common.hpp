// Internal declarations shared by the host orchestration and the gfx950 kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

namespace psr {

// Sample storage of a device-resident observation (channel-major rows).
enum SampleKind : int { KIND_U8 = 0, KIND_U16 = 1, KIND_F32 = 2 };

// Candidate record appended by the harmonic-sum kernel before ranking.
struct RawPeak {
    double score;
    double offset;  // parabolic interpolation offset of the fundamental, [-0.5, 0.5]
    uint32_t k;     // fundamental bin
    uint32_t h;     // harmonics summed
};

// Output record of the ranking kernel; layout-identical to psr_candidate.
struct CandidateOut {
    double frequency_hz;
    double period_s;
    double dm;
    uint64_t harmonics;
    double score;
};

// Geometry of one periodicity search over a batch of equal-length series.
struct SpecGeom {
    uint32_t n;        // series length fed to periodicity
    uint32_t nfft;     // 2 * next_pow2(n)
    uint32_t m;        // nfft / 2 (complex FFT length)
    uint32_t nbins;    // m + 1
    uint32_t m1;       // second-pass FFT length (1 when single pass)
    uint32_t m2;       // first-pass FFT length (m when single pass)
    uint32_t nblocks;  // ceil(nbins / 64)
    uint32_t series_pitch;  // elements between consecutive series rows
    uint64_t spec_pitch;    // elements between consecutive spectrum rows
};

struct SearchParams {
    double tsamp;
    double bin_hz;
    double threshold;
    double resolution;  // 1 / (n * tsamp): per-trial dedup width
    uint32_t first;
    uint32_t last;
    uint32_t end_h[5];  // end bin per harmonic level (0 = level skipped)
    uint32_t nlevels;   // number of harmonic levels searched (h = 1,2,4,8,16)
    uint32_t max_candidates;
    uint32_t nbirdie;   // number of excluded bin intervals
    uint32_t fp16;      // series/spectra stored FP16
    uint32_t peak_cap;  // raw peak capacity per trial
};

// Workgroups are dispatched round-robin over the 8 XCDs, each with its own L2.
// xcd_block() maps a 2-D grid's dispatch order to work coordinates so that each
// XCD walks a contiguous run of work items (x fastest): neighbouring items that
// share cache lines then run on one XCD at about the same time.  Bijective over
// the grid for any size.
constexpr uint32_t NUM_XCDS = 8;
__device__ inline uint2 xcd_block() {
    const uint32_t gx = gridDim.x, total = gx * gridDim.y;
    const uint32_t id = blockIdx.y * gx + blockIdx.x;
    const uint32_t xcd = id % NUM_XCDS, local = id / NUM_XCDS;
    const uint32_t q = total / NUM_XCDS, r = total % NUM_XCDS;
    const uint32_t work = (xcd < r ? xcd * (q + 1) : r * (q + 1) + (xcd - r) * q) + local;
    return make_uint2(work % gx, work / gx);
}

constexpr int MAX_BIRDIE_INTERVALS = 64;
constexpr uint32_t PEAK_CAP = 4096;

// ---- kernel launchers (defined in the .hip files) ----
// ingest: packed time-major payload -> channel-major rows
void launch_ingest(hipStream_t s, const uint8_t* payload, int nbits, uint32_t nchans,
                   uint64_t nsamples, void* out, uint64_t pitch, int kind);
void launch_chan_stats(hipStream_t s, const void* data, int kind, uint32_t nchans,
                       uint64_t nsamples, uint64_t pitch, double* sums, double* sumsq);
void launch_delays(hipStream_t s, const double* inv_f2, uint32_t nchans, double ref_inv,
                   double tsamp, const double* dms, uint32_t ntrials, int32_t* out);
// integer dedispersion (u8 rows), outputs f32 exact sums or FP16 scaled
// DM trials per dedispersion workgroup (the delay spread bound is per block)
constexpr uint32_t DEDISP_BLOCK_TRIALS = 64;
// Bytes allocated after the last channel row of every u8 observation: the
// LDS-staged dedispersion reads a channel row a few hundred samples past the
// last delayed sample it sums.
constexpr uint64_t DD_ALLOC_SLACK = 4096;
// spread[c] (host, per channel) bounds the delay spread of channel c over any
// block of DEDISP_BLOCK_TRIALS trials; max_delay bounds every delay of the
// launch.  Returns false (and launches nothing) when the spreads need more LDS
// than a workgroup has, the totals could pass 2^20 or the reads could pass the
// allocation; the caller then uses launch_dedisp_f32.  `table` holds
// dedisp_u8_table_words(nchans, ntrials) words: the blocks' channel bases and
// the per-block plans (stages of channel-pair rows) a planner kernel writes.
bool launch_dedisp_u8(hipStream_t s, const uint8_t* data, uint64_t pitch, uint32_t nchans,
                      uint32_t maxval, const int32_t* delays, int32_t* table, uint32_t ntrials,
                      uint32_t n, void* out, uint32_t out_pitch, int fp16, float m0, float inv_s0,
                      double* partial, const uint32_t* spread, uint64_t max_delay);
size_t dedisp_u8_table_words(uint32_t nchans, uint32_t ntrials);

// One observation's share (a range of its DM trials) of a many-observation
// dedispersion launch.  Host fills every field; device pointers point into the
// batch's buffers.
struct DedispJob {
    const uint8_t* data;   // u8 channel rows of the observation
    uint64_t pitch;        // bytes between channel rows
    const double* dms;     // [ntrials], ascending
    const double* inv_f2;  // [nchans] 1/(f*f)
    double ref_inv, tsamp;
    int32_t* delays;       // [ntrials][nchans]
    int32_t* bases;        // [nblk][nchans]
    uint32_t* plan;        // [nblk][plan_words]
    uint64_t plan_words;
    void* out;             // series row of the job's first trial
    double* partial;       // [ntrials][ntiles]
    uint32_t nchans, ntrials, nblk, ntiles, n, out_pitch;
    uint32_t phase_slots, kmax;
    float m0, inv_s0;
    uint32_t entry0, blk0, item0, trial0;  // first delay entry, plan block, work item, trial in the batch
    double m0d, inv_s0d;                   // FP16 storage offset and 1/scale in f64 (trial means)
};
// Fills kmax, phase_slots and plan_words for a job with nchans u8 rows of
// `pitch` bytes, samples <= maxval, per-channel delay spreads `spread` over any
// 64-trial block, series length n and delays <= max_delay; false when the
// LDS-staged kernel cannot take it (the caller then dedisperses it with
// launch_dedisp_f32).
bool dedisp_job_setup(uint32_t nchans, uint32_t maxval, uint64_t pitch, const uint32_t* spread, uint32_t n,
                      uint64_t max_delay, DedispJob* j);
uint32_t dedisp_tiles(uint32_t n);  // 512-output time tiles of a series
size_t dedisp_plan_words(size_t nchans);  // plan words per 64-trial block
// delays, plans and the persistent work-queue dedispersion of a batch of jobs:
// entry0/blk0/item0 hold each job's first delay entry / plan block / work item
// (njobs ascending values starting at 0); counters = 8 device words.
void launch_dedisp_batch(hipStream_t s, const DedispJob* jobs, const uint32_t* entry0,
                         const uint32_t* blk0, const uint32_t* item0, uint32_t njobs,
                         uint32_t total_entries, uint32_t total_blocks, uint32_t total_items,
                         uint32_t span, int fp16, uint32_t max_nchans, uint32_t grid,
                         uint32_t* counters);
// sequential-order f32 dedispersion (16-bit and float rows)
void launch_dedisp_f32(hipStream_t s, const void* data, int kind, uint64_t pitch,
                       uint32_t nchans, const int32_t* delays, uint32_t ntrials, uint32_t n,
                       void* out, uint32_t out_pitch, int fp16, float m0, float inv_s0,
                       double* partial);
// fill n words with v (clears counters without memset nodes)
void launch_fill_u32(hipStream_t s, uint32_t* p, size_t n, uint32_t v);
// FFT + power
// trial_n (optional): per-trial series lengths when the trials of one launch
// come from several observations (the geometry's n otherwise); batch_trials:
// trials per pass-1/pass-2 round (0 = fft_work_trials)
void launch_fft(hipStream_t s, const SpecGeom& g, const void* series, int fp16,
                const float* mean32, const float* corr16, uint32_t ntrials, float2* work,
                void* spec_t, double* partial_sums, const float2* tw, const float* cos_tables,
                const uint32_t* trial_n = nullptr, uint32_t batch_trials = 0);

// Quarter-wave cosine tables C_L[j] = cos(2 pi j / L), j = 0..L/4, for every
// power of two L = 2..COSTAB_MAX, concatenated in increasing L.
constexpr int COSTAB_MAX = 8192;
__host__ __device__ constexpr int costab_offset(int L) {
    return L <= 2 ? 0 : costab_offset(L / 2) + (L / 2) / 4 + 1;
}
constexpr int COSTAB_SIZE = costab_offset(2 * COSTAB_MAX);
// After the cosine tables (8-byte aligned): W_4096^(k t) as (re, im) float
// pairs at float index TW4096_OFFSET + 2 (256 (k - 1) + t), k = 1..15, t < 256
// (the twiddles of the first stage of the length-4096 row FFTs).
constexpr int TW4096_OFFSET = (COSTAB_SIZE + 1) & ~1;
constexpr int COSTAB_ALL = TW4096_OFFSET + 2 * 15 * 256;
void fill_costabs(float* host_table);  // COSTAB_ALL floats
// Intermediate bytes per four-step FFT batch (the work buffer holds this many
// trials' intermediates): 224 MiB (14 trials of 2^21 points, 3584 workgroups
// per pass: whole rounds of two per CU) stays within the 256 MiB Infinity
// Cache between the two passes; measured against 128 MiB to 1 GiB batches it
// gives the shortest headline step.
constexpr uint64_t FFT_BATCH_BYTES = 224ull << 20;
uint32_t fft_work_trials(const SpecGeom& g, uint32_t ntrials);
// true when launch_fft has kernels for the geometry's split (m2 of 512, 1024 or
// 2048 in the four-step form; every power of two up to 4096 in one pass)
bool fft_shape_supported(const SpecGeom& g);
// whitening scales + natural-order raw spectrum; `sorted` (spectrum-sized) and
// `floors` (one float per trial) are scratch, unused when whiten_fused(g)
bool whiten_fused(const SpecGeom& g);
void launch_whiten(hipStream_t s, const SpecGeom& g, const void* spec_t, int fp16,
                   uint32_t ntrials, const double* partial_sums, uint32_t npartials,
                   void* raw_nat, void* sorted, float* floors, float* scales);
// harmonic sums + peaks
void launch_harmsum(hipStream_t s, const SpecGeom& g, const SearchParams& p, const void* raw_nat,
                    const float* scales, const uint32_t* birdies, uint32_t ntrials,
                    RawPeak* peaks, uint32_t* peak_counts);
// harmonic sums over trials of several observations: `jobs` holds
// harm_args_bytes() per observation (fill_harm_args), trial_job[t] the
// observation of trial t, max_tiles the largest harmsum_tiles() among them
size_t harm_args_bytes();
void fill_harm_args(const SearchParams& p, const uint32_t* birdies, void* out);
uint32_t harmsum_tiles(const SearchParams& p);
void launch_harmsum_multi(hipStream_t s, const SpecGeom& g, int fp16, uint32_t max_tiles, const void* raw_nat,
                          const float* scales, const void* jobs, const uint32_t* trial_job, uint32_t ntrials,
                          RawPeak* peaks, uint32_t* peak_counts);
void launch_finalize_multi(hipStream_t s, const double* dms, uint32_t ntrials, const RawPeak* peaks,
                           const uint32_t* peak_counts, const void* jobs, const uint32_t* trial_job,
                           CandidateOut* out, uint32_t cap_per_trial, uint32_t* counts);
// rank, dedup, cap
void launch_finalize(hipStream_t s, const SpecGeom& g, const SearchParams& p, const double* dms,
                     uint32_t ntrials, const RawPeak* peaks, const uint32_t* peak_counts,
                     CandidateOut* out, uint32_t cap_per_trial, uint32_t* counts);
// trial means of a batch of jobs: trial_job[t] = job of batch trial t
void launch_trial_mean_multi(hipStream_t s, const DedispJob* jobs, const uint32_t* trial_job,
                             uint32_t ntrials, float* mean32, float* corr16);
void launch_trial_mean(hipStream_t s, const double* partial, uint32_t ntiles, uint32_t ntrials,
                       uint32_t n, double m0, double inv_s0, float* mean32, float* corr16);
// folding
void launch_fold(hipStream_t s, const float* series, uint64_t n, double tsamp,
                 const double* periods, const uint64_t* offsets, uint32_t nprof, uint32_t bins,
                 double* sums, uint64_t* counts);
// RFI mask (crates/rfi/src/lib.rs:31-96) on a device observation
void launch_rfi_row_stats(hipStream_t s, const void* data, int kind, uint64_t pitch, uint32_t nch,
                          uint64_t n, float* means, float* vars);
void launch_rfi_block_sums(hipStream_t s, const void* data, int kind, uint64_t pitch, uint32_t nch,
                           uint64_t n, uint32_t block, uint32_t nblocks, float* sums);
void launch_rfi_block_means(hipStream_t s, const float* sums, const uint8_t* bad, uint32_t nch,
                            uint64_t n, uint32_t block, uint32_t nblocks, uint32_t active,
                            float* means);
void launch_rfi_median(hipStream_t s, const void* data, int kind, uint64_t pitch, uint32_t nch,
                       uint64_t n, uint32_t* median_key);
void launch_rfi_fill(hipStream_t s, void* data, int kind, uint64_t pitch, uint32_t nch, uint64_t n,
                     uint32_t block, const uint8_t* bad, const uint8_t* bad_blocks,
                     const uint32_t* median_key);
// confirmation stage: per-channel profiles of raw rows and the DM response
void launch_fold_rows(hipStream_t s, const void* data, int kind, uint64_t pitch, uint32_t nrows,
                      uint64_t n, double tsamp, double period, uint32_t bins, double* sums,
                      uint64_t* counts);
void launch_dm_response(hipStream_t s, const float* values, const double* means,
                        const int32_t* delays, uint32_t nch, uint32_t ntrials, uint32_t bins,
                        double tsamp, double period, uint32_t start, uint32_t width,
                        double* resp);

// Twiddle tables needed by launch_fft for a geometry (size in float2 elements).
size_t twiddle_table_size(const SpecGeom& g);
void fill_twiddles(const SpecGeom& g, float2* host_table);
SpecGeom make_geometry(uint32_t n);
// the same with the four-step split's pass-1 length m2 chosen by the caller
// (ignored for one-pass sizes); check fft_shape_supported() on the result
SpecGeom make_geometry_split(uint32_t n, uint32_t m2);

}  // namespace psr
