/*
 * psr_gpu.h -- C ABI of the MI355X (gfx950) search hot path of PulsarSr.
 *
 * The entry points below replace the reference's per-DM-trial loop
 *   crates/cli/src/main.rs:462-469  (delays -> apply -> periodicity, rayon par_iter)
 * and the crate functions it calls.  Each declaration names the reference
 * interface it stands in for.  Conventions:
 *   - every function returns PSR_OK (0) or a negative PSR_E_* code, never aborts
 *     or throws across the ABI; psr_last_error() returns the message of the
 *     calling thread's last failure (the text mirrors the reference's Error);
 *   - all host buffers are caller-owned; device memory belongs to the psr_ctx
 *     and is reused across calls;
 *   - a psr_ctx is not thread-safe: use one per thread, or serialise calls;
 *   - there is no CPU fallback: GPU entry points fail with PSR_E_DEVICE when no
 *     gfx950 device is present.
 */
#ifndef PSR_GPU_H
#define PSR_GPU_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PSR_OK 0
#define PSR_E_INVALID (-1)  /* dedisperse::Error: invalid input (crates/dedisperse/src/lib.rs:5-7) */
#define PSR_E_SHAPE (-2)    /* array shape mismatch / invalid output length (lib.rs:68,76-78) */
#define PSR_E_DEVICE (-3)   /* HIP runtime error or no usable device */
#define PSR_E_OOM (-4)      /* device allocation failed or exceeded the ctx memory cap */
#define PSR_E_FORMAT (-5)   /* fbank::Error::Format (crates/fbank/src/lib.rs:12-18) */
#define PSR_E_CONTRACT (-6) /* a reference assert! would panic (search/src/lib.rs:34, fold/src/lib.rs:23,89) */
#define PSR_E_OVERFLOW (-7) /* per-trial peak buffer overflowed; result would not be exact */

#define PSR_PRECISION_FP32 0u /* series and spectra stored FP32 (the reference's arithmetic) */
#define PSR_PRECISION_FP16 1u /* series and spectra stored FP16 after per-trial scaling */

typedef struct psr_ctx psr_ctx;
typedef struct psr_obs psr_obs; /* one device-resident observation */

/* Mirrors search::Candidate (crates/search/src/lib.rs:6-13); usize -> uint64_t. */
typedef struct {
    double frequency_hz;
    double period_s;
    double dm;
    uint64_t harmonics;
    double score;
} psr_candidate;

/* Mirrors search::Config (crates/search/src/lib.rs:14-32) plus two knobs the
 * reference hard-codes: max_harmonics (16 = the reference's 1,2,4,8,16) and the
 * storage precision of series/spectra. */
typedef struct {
    double min_hz;
    double max_hz;
    double threshold;
    uint64_t max_candidates;
    const double *birdies; /* nbirdies pairs (center_hz, half_width_hz) */
    uint64_t nbirdies;
    uint32_t max_harmonics;
    uint32_t precision; /* PSR_PRECISION_* */
} psr_search_cfg;

/* Mirrors fbank::Header's numerical part (crates/fbank/src/lib.rs:29-38). */
typedef struct {
    uint32_t nchans;
    uint32_t nbits;
    double tsamp;
    double fch1;
    double foff;
    uint64_t data_offset; /* bytes before the payload */
    uint64_t nsamples;    /* whole frames in the payload */
} psr_header;

/* A packed SIGPROC payload: time-major frames, LSB-first for 1/2/4-bit
 * (crates/fbank/src/lib.rs:149-196).  nbits in {1,2,4,8,16,32}. */
typedef struct {
    const void *payload;
    uint64_t payload_bytes;
    uint32_t nbits;
    uint32_t nchans;
    uint64_t nsamples; /* frames to use (<= whole frames in payload) */
    double tsamp;
    double fch1;
    double foff;
} psr_filterbank;

/* Mirrors fold::Profile's scalar fields (crates/fold/src/lib.rs:3-11). */
typedef struct {
    double period_s;
    double snr;
    double reduced_chi_square;
} psr_profile_stats;

/* Per-stage device time of the last psr_search call, milliseconds, measured
 * with HIP events on the stream each stage ran on (summed over batches).
 * total_ms is the wall time from the first batch's start to the last batch's
 * end; with overlapped batches it is less than the sum of the stages. */
typedef struct {
    double total_ms;
    double delays_ms;
    double dedisperse_ms;
    double fft_ms;
    double whiten_ms;
    double harmsum_ms;
    double finalize_ms;
    uint64_t trials;
    uint64_t series_len;
    uint64_t fft_len;
    uint64_t launches; /* device batches (one dedispersion launch each) */
} psr_stage_times;

/* ---- library / context ------------------------------------------------- */
const char *psr_version(void);
/* Message of the calling thread's most recent failure ("" if none). */
const char *psr_last_error(void);
/* Context flags.  By default psr_search runs its trial batches one after
 * another on one stream, so each stage time is that stage's own kernel time.
 * PSR_CTX_OVERLAP dedisperses the next batch on a second stream while the
 * current batch's spectral stages run (stage times then overlap; total_ms is
 * wall time).  At 1024 channels x 4000 trials the two give the same
 * throughput: the dedispersion kernel holds every CU's registers. */
#define PSR_CTX_OVERLAP 1u
/* Create a context on HIP device `device` with PSR_CTX_* `flags` (0 = default). */
int psr_ctx_create(int device, uint32_t flags, psr_ctx **out);
void psr_ctx_destroy(psr_ctx *ctx);
/* Cap the device memory a ctx may hold (bytes; 0 = unlimited).  Host-forcible OOM path. */
int psr_ctx_set_memory_cap(psr_ctx *ctx, uint64_t bytes);

/* ---- host math: crates/dedisperse (no GPU needed) ----------------------- */
/* dedisperse::delays (crates/dedisperse/src/lib.rs:9-34) */
int psr_delays(const double *freqs, size_t nch, double tsamp, double dm, uint64_t *out);
/* dedisperse::dm_step (crates/dedisperse/src/lib.rs:35-44) */
int psr_dm_step(const double *freqs, size_t nch, double tsamp, double *out);
/* dedisperse::plan (crates/dedisperse/src/lib.rs:45-60); writes *count trials (<= cap). */
int psr_plan(double min, double max, double step, double *out, size_t cap, size_t *count);
/* fbank::Header::frequencies (crates/fbank/src/lib.rs:40-44) */
int psr_frequencies(double fch1, double foff, size_t nch, double *out);
/* fbank::parse_header (crates/fbank/src/lib.rs:72-147) */
int psr_parse_header(const uint8_t *bytes, size_t len, psr_header *out);
/* search::sift (crates/search/src/lib.rs:147-176); returns kept count in *count. */
int psr_sift(const psr_candidate *in, size_t n, double resolution_hz, size_t limit,
             int reject_dm_zero, psr_candidate *out, size_t *count);

/* ---- observations (device-resident filterbank) -------------------------- */
/* fbank::decode's unpacking (crates/fbank/src/lib.rs:149-196) done on the GPU:
 * uploads the packed payload and lays it out channel-major in HBM. */
int psr_obs_create(psr_ctx *ctx, const psr_filterbank *fb, psr_obs **out);
/* An observation from f32 channel-major data: the `data` argument of
 * dedisperse::apply (crates/dedisperse/src/lib.rs:62-66). */
int psr_obs_create_f32(psr_ctx *ctx, const float *data, uint64_t nsamples, uint32_t nchans,
                       double tsamp, double fch1, double foff, psr_obs **out);
void psr_obs_destroy(psr_obs *obs);
/* Copy the channel-major f32 samples back (nchans*nsamples floats), for checks. */
int psr_obs_download_f32(psr_obs *obs, float *out);

/* rfi::Mask (crates/rfi/src/lib.rs:23-29); the channel and block lists go to
 * caller arrays. */
typedef struct {
    uint64_t nchannels, nblocks, block_samples;
    double masked_fraction;
} psr_mask;
/* rfi::mask (crates/rfi/src/lib.rs:31-96) applied in place to a device-resident
 * observation: masked samples take their channel's median, bit for bit as the
 * reference does on its f32 array.  channels[nchans] and
 * blocks[ceil(nsamples / block_samples)] receive the masked indices. */
int psr_obs_mask(psr_ctx *ctx, psr_obs *obs, uint64_t block_samples, float threshold,
                 psr_mask *out, uint64_t *channels, uint64_t *blocks);

/* ---- the hot path ------------------------------------------------------- */
/* dedisperse::delays + dedisperse::apply for `ntrials` DMs
 * (crates/dedisperse/src/lib.rs:9-34, 62-90).  length==0 means None (full
 * available length).  out[ntrials*length] f32; *out_len gets the length used
 * (out may be NULL to query it).  Lengths up to 2^31 samples. */
int psr_dedisperse(psr_ctx *ctx, psr_obs *obs, const double *dms, size_t ntrials, size_t length,
                   float *out, size_t *out_len);
/* dedisperse::apply(data, nsamples, delays, length) (crates/dedisperse/src/lib.rs:62-90),
 * one to one: `data` is channels [ch_lo, ch_hi) of the device observation (the
 * subband slices of crates/search/src/confirmation.rs:362-372 are such ranges;
 * the whole band is 0..nchans) and delays[ch_hi - ch_lo] is the caller's
 * per-channel delay table in samples (any values, e.g. the all-zero tables of
 * confirmation.rs:253,321).  length==0 means None.  Sums run in channel order in
 * f32, as the reference adds them.  Errors as apply(): an empty or out-of-range
 * channel range is "array shape mismatch" (PSR_E_SHAPE), a delay past the
 * observation "delay exceeds observation" (PSR_E_INVALID), a length outside
 * [2, nsamples - max delay] "invalid output length" (PSR_E_SHAPE).  out[length]
 * f32, or NULL to query *out_len. */
int psr_apply(psr_ctx *ctx, psr_obs *obs, const uint64_t *delays, uint32_t ch_lo, uint32_t ch_hi,
              size_t length, float *out, size_t *out_len);

/* The per-trial loop of crates/cli/src/main.rs:462-469: for every DM, delays ->
 * apply(Some(length)) -> search::periodicity (crates/search/src/lib.rs:33-145).
 * Writes up to cap_per_trial candidates per trial at out[t*cap_per_trial], in the
 * reference's order, and counts[t]. */
int psr_search(psr_ctx *ctx, psr_obs *obs, const double *dms, size_t ntrials, size_t length,
               const psr_search_cfg *cfg, psr_candidate *out, size_t cap_per_trial,
               uint32_t *counts);
/* Stage timings of the last psr_search on this ctx. */
int psr_last_stage_times(const psr_ctx *ctx, psr_stage_times *out);

/* search::periodicity on one host series (crates/search/src/lib.rs:33-145), 1:1. */
int psr_periodicity(psr_ctx *ctx, const float *series, size_t n, double tsamp, double dm,
                    const psr_search_cfg *cfg, psr_candidate *out, size_t cap, size_t *count);
/* No counterpart in the reference.  psr_periodicity without any device
 * allocation, for callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc).  Same
 * arguments, results and errors as psr_periodicity; every device buffer comes
 * from one static per-device scratch array of 4 MiB that the library holds
 * until the process ends.  Limits: n <= PSR_NOALLOC_MAX_SERIES samples and
 * cap <= PSR_NOALLOC_MAX_CANDIDATES (larger requests fail with PSR_E_SHAPE /
 * PSR_E_INVALID).  All calls on one device share that array: calls that overlap
 * on one device race (serialise them); calls on different devices do not. */
#define PSR_NOALLOC_MAX_SERIES 131072u
#define PSR_NOALLOC_MAX_CANDIDATES 256u
int psr_periodicity_noalloc(psr_ctx *ctx, const float *series, size_t n, double tsamp, double dm,
                            const psr_search_cfg *cfg, psr_candidate *out, size_t cap, size_t *count);
/* The spectral stages of periodicity for checks: raw |X|^2 and whitened power
 * (crates/search/src/lib.rs:35-72), each nbins = next_pow2(n)+1 floats.  With
 * cfg->precision = PSR_PRECISION_FP16 the series (less its mean) is stored in
 * half precision and the spectrum as FP16 amplitudes, as psr_search stores
 * them; raw then holds the power a * a the whitening and scoring use. */
int psr_spectrum(psr_ctx *ctx, const float *series, size_t n, double tsamp,
                 const psr_search_cfg *cfg, float *raw, float *white, size_t *nbins);
/* No counterpart in the reference.  psr_spectrum with the four-step FFT split
 * forced to m2 rows of m1 = next_pow2(n) / m2 (m2 = 512, 1024 or 2048; 0 = the
 * default split): the splits psr_batch's autotuner chooses among, for checks
 * of each against the oracle.  PSR_E_SHAPE for an m2 with no split at this
 * length (spectra of at most 4096 points are one pass). */
int psr_spectrum_split(psr_ctx *ctx, const float *series, size_t n, double tsamp,
                       const psr_search_cfg *cfg, uint32_t m2, float *raw, float *white, size_t *nbins);

/* ---- many observations ------------------------------------------------- */
/* The reference searches one observation per `psr search` run (the per-trial
 * loop of crates/cli/src/main.rs:462-469) and its many-observation harness
 * (crates/synth/examples/confirmation.rs) loops over observations in turn.  A
 * psr_batch runs that per-trial loop for many device-resident observations of
 * uneven length, channel count, bit depth and DM plan in one device schedule:
 * per round of trials, one launch per stage (a persistent dedispersion kernel
 * fed by per-XCD work queues of (observation, DM block, time-tile span) items,
 * then FFT, whitening, harmonic sums and ranking over all the round's trials).
 * Every observation gets exactly the candidates psr_search would give it (with
 * the default shapes: bit for bit; tuned FFT splits round the spectrum
 * differently). */
typedef struct psr_batch psr_batch;
#define PSR_BATCH_GRAPH 1u    /* capture the whole schedule in one hipGraph at creation; runs replay it */
#define PSR_BATCH_AUTOTUNE 2u /* time candidate shapes on the batch's own data at creation (FFT split
                               * and round size per FFT length; dedispersion tile span and workgroups) */
typedef struct {
    uint32_t flags;         /* PSR_BATCH_* */
    uint32_t pad_;
    uint64_t memory_budget; /* device bytes per round; 0 = half the free memory, at most 64 GiB */
} psr_batch_opts;
/* Plan a batch over obs[nobs]: observation i searches dms[i][ntrials[i]] with
 * series length lengths[i] (0 = None, as psr_search; lengths may be NULL),
 * all with `cfg` and cap_per_trial candidates per trial.  Device buffers, the
 * tuned shapes and the graph belong to the batch; the observations must
 * outlive it.  Errors as psr_search. */
int psr_batch_create(psr_ctx *ctx, psr_obs *const *obs, size_t nobs, const double *const *dms,
                     const size_t *ntrials, const size_t *lengths, const psr_search_cfg *cfg,
                     size_t cap_per_trial, const psr_batch_opts *opts, psr_batch **out);
/* Run the batch: observation i's trial t (its caller order) lands at row
 * sum_{k<i} ntrials[k] + t of out[rows * cap_per_trial] and counts[rows]. */
int psr_batch_run(psr_batch *batch, psr_candidate *out, uint32_t *counts);
/* JSON description: rounds, FFT splits, dedispersion shape, autotuner timings,
 * the last run's device time (last_ms).  *len = full length; json may be NULL. */
int psr_batch_info(const psr_batch *batch, char *json, size_t cap, size_t *len);
void psr_batch_destroy(psr_batch *batch);

/* ---- folding ------------------------------------------------------------ */
/* fold::profile_at for `nprof` (period, sample_offset) pairs over one host series
 * (crates/fold/src/lib.rs:16-80).  values/counts are [nprof*bins]. */
int psr_fold(psr_ctx *ctx, const float *series, size_t n, double tsamp, const double *periods,
             const uint64_t *offsets, size_t nprof, uint32_t bins, psr_profile_stats *stats,
             float *values, uint64_t *counts);
/* fold::refine (crates/fold/src/lib.rs:81-97): steps+1 trial frequencies, last max wins. */
int psr_refine(psr_ctx *ctx, const float *series, size_t n, double tsamp, double frequency,
               double span_hz, size_t steps, uint32_t bins, psr_profile_stats *stats,
               float *values, uint64_t *counts);

/* ---- confirmation (crates/search/src/confirmation.rs) ------------------- */
/* Window (confirmation.rs:41-51) */
typedef struct {
    uint64_t start_bin, width_bins, bins;
} psr_window;
/* Measurement (confirmation.rs:52-59) */
typedef struct {
    double snr;
    uint64_t on_samples, off_samples;
    double mean_difference, standard_error;
} psr_measurement;
/* Rules (confirmation.rs:17-40); psr_rules_default fills Rules::default(). */
typedef struct {
    double minimum_validation_snr, minimum_part_snr;
    uint64_t minimum_time_parts, minimum_subbands;
    double minimum_dm_contrast, minimum_dispersion_widths, maximum_zero_dm_spectral_ratio;
} psr_rules;
/* DmPoint (confirmation.rs:60-69); has_spectral = 0 stands for None. */
typedef struct {
    double dm, spectral_score, spectral_frequency_hz, discovery_profile_response;
    uint32_t has_spectral, pad_;
} psr_dm_point;
/* One harmonic member: its family and its index in the input candidates. */
typedef struct {
    uint32_t family, pad_;
    uint64_t index;
} psr_member;
/* ConfirmedCandidate (confirmation.rs:70-90) without its variable-length parts,
 * which psr_analyze writes to caller arrays.  Bit i of rejection_mask is the
 * i-th rejection reason of confirmation.rs:434-466, in source order. */
typedef struct {
    psr_candidate candidate;
    double refined_dm;
    psr_profile_stats profile;
    psr_window window;
    psr_measurement validation, zero_dm_validation;
    double zero_dm_period_s;
    psr_window zero_dm_window;
    double dm_contrast, dispersion_widths, zero_dm_related_spectral_ratio;
    psr_measurement time_parts[4], subbands[4];
    uint64_t nmembers;
    uint32_t accepted, rejection_mask;
} psr_confirmed;
/* Outcome (confirmation.rs:91-100) */
typedef struct {
    uint64_t discovery_samples, validation_start_sample, validation_samples, common_samples,
        trial_count, ncandidates;
} psr_outcome;

void psr_rules_default(psr_rules *out);
/* confirmation::measure (confirmation.rs:107-141): the fixed on/off window
 * measurement of x[n], sample i at phase bin of (i + offset) * tsamp / period
 * (confirmation.rs:103-105).  Host code: sequential f64 sums in the
 * reference's order.  window->bins must be positive. */
int psr_measure(const float *x, size_t n, uint64_t offset, double tsamp, double period,
                const psr_window *window, psr_measurement *out);
/* confirmation::group (confirmation.rs:177-199): families of harmonically
 * related candidates, strongest first.  leaders[limit]; members[n] lists
 * (family, input index) in the order members were appended. */
int psr_group(const psr_candidate *in, size_t n, double resolution_hz, size_t limit,
              psr_candidate *leaders, size_t *nfamilies, psr_member *members, size_t *nmembers);
/* confirmation::analyze (confirmation.rs:203-496) over a device-resident
 * observation: every dedispersion, fold, channel profile and DM response runs
 * on the GPU; window selection, measurements and rules on the host.  Caller
 * arrays: confirmed[limit], profile_bins/profile_counts[limit*bins],
 * time_profiles/subband_profiles[limit*4*bins], dm_response[limit*ntrials],
 * members[ngenerated]. */
int psr_analyze(psr_ctx *ctx, psr_obs *obs, const double *trials, size_t ntrials,
                const psr_candidate *generated, size_t ngenerated, size_t limit, uint32_t bins,
                const psr_rules *rules, psr_outcome *outcome, psr_confirmed *confirmed,
                float *profile_bins, uint64_t *profile_counts, float *time_profiles,
                float *subband_profiles, psr_dm_point *dm_response, psr_member *members,
                size_t *nmembers);

#ifdef __cplusplus
}
#endif
#endif
