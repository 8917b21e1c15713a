/*
 * oracle.h -- CPU restatement of the PulsarSr search path.
 *
 * TEST INFRASTRUCTURE ONLY. This library is the checker: only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it.
 * The product (pulsarsr_amd, libpsrgpu.so) never links or calls it.
 *
 * Every function restates one reference function line by line, in the
 * reference's own arithmetic (f32 where the Rust uses f32, f64 where it uses
 * f64, no fused multiply-add: build with -ffp-contract=off).  The one place
 * that cannot be restated is the third-party FFT (realfft 3.5.0 -> rustfft
 * 6.4.1, crates/search/src/lib.rs:41-43): the oracle computes the DFT in f64
 * and rounds each component to f32, i.e. it models an exactly rounded f32
 * transform.  Spectral parity is therefore checked to a tolerance, and the
 * candidate lists are pinned against the saved reference reports.
 *
 * Functions without any independent pin are marked "parity unpinned" below;
 * at present there are none (see the tests named next to each function).
 */
#ifndef PSR_ORACLE_H
#define PSR_ORACLE_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ORACLE_DISPERSION_SECONDS_MHZ 4148.808

typedef struct {
    double frequency_hz;
    double period_s;
    double dm;
    uint64_t harmonics;
    double score;
} oracle_candidate;

typedef struct {
    double min_hz;
    double max_hz;
    double threshold;
    uint64_t max_candidates;
    const double *birdies; /* pairs (center_hz, half_width_hz) */
    uint64_t nbirdies;
    uint32_t max_harmonics; /* 16 = reference behaviour (h in 1,2,4,8,16) */
    uint32_t pad_;
} oracle_config;

/* crates/dedisperse/src/lib.rs:9-34.  Returns 0, or -1 on invalid input. */
int oracle_delays(const double *freqs, size_t nch, double tsamp, double dm, uint64_t *out);
/* crates/dedisperse/src/lib.rs:35-44 */
int oracle_dm_step(const double *freqs, size_t nch, double tsamp, double *out);
/* crates/dedisperse/src/lib.rs:45-60.  Returns trial count (<= cap) or -1. */
long oracle_plan(double min, double max, double step, double *out, size_t cap);
/* crates/dedisperse/src/lib.rs:62-90.  length==0 means None.  *out_len set. */
int oracle_apply(const float *data, size_t nsamples, size_t nch, const uint64_t *delays,
                 size_t length, float *out, size_t *out_len);
/* crates/fbank/src/lib.rs:149-196 (payload part): packed time-major -> f32 channel-major. */
int oracle_decode(const uint8_t *payload, size_t payload_len, int nbits, size_t nchans,
                  size_t limit, float *out, size_t *nsamples_out);
/* crates/rfi/src/lib.rs:4-12 (upper median, total order) */
float oracle_median(const float *v, size_t n);
/* The same upper median over doubles (the FP64 oracle's whitening). */
double oracle_median64(const double *v, size_t n);
/* crates/rfi/src/lib.rs:31-96.  Mutates data; channels/blocks written as 0/1 flags. */
double oracle_mask(float *data, size_t nsamples, size_t nch, size_t block_samples, float threshold,
                   uint8_t *bad_channels, uint8_t *bad_blocks);

/* f64 real FFT of x[0..N) (N even): X[k], k=0..N/2.  Exact-arithmetic model. */
void oracle_rfft64(const double *x, size_t N, double *re, double *im);

/* crates/search/src/lib.rs:35-72: the f32 spectrum stages.
 * raw[N/2+1] = |X|^2 (f32), white[N/2+1] = whitened + DC/birdies.  Returns N_fft. */
size_t oracle_spectrum(const float *series, size_t n, double tsamp, const oracle_config *cfg,
                       float *raw, float *white);
/* crates/search/src/lib.rs:33-145.  Returns the number of candidates written. */
size_t oracle_periodicity(const float *series, size_t n, double tsamp, double dm,
                          const oracle_config *cfg, oracle_candidate *out, size_t cap);
/* All-f64 twin of oracle_spectrum (double raw power and whitened power). Returns N_fft. */
size_t oracle_spectrum_f64(const float *series, size_t n, double tsamp, const oracle_config *cfg,
                           double *raw, double *white);
/* All-f64 twin of oracle_periodicity (FP64 oracle for reduced-precision error reports). */
size_t oracle_periodicity_f64(const float *series, size_t n, double tsamp, double dm,
                              const oracle_config *cfg, oracle_candidate *out, size_t cap);
/* crates/search/src/lib.rs:147-176 */
size_t oracle_sift(const oracle_candidate *in, size_t n, double resolution_hz, size_t limit,
                   int reject_dm_zero, oracle_candidate *out);

/* crates/fold/src/lib.rs:16-80 */
void oracle_profile_at(const float *series, size_t n, double tsamp, double period, size_t bins,
                       size_t sample_offset, double *snr, double *chi2, float *values,
                       uint64_t *counts);
/* crates/fold/src/lib.rs:81-97.  Returns the winning period. */
double oracle_refine(const float *series, size_t n, double tsamp, double frequency, double span_hz,
                     size_t steps, size_t bins, double *snr, double *chi2, float *values,
                     uint64_t *counts);

/* ---- confirmation stage (crates/search/src/confirmation.rs; oracle/analyze.c) ---- */
typedef struct {
    uint64_t start_bin, width_bins, bins;
} oracle_window;
typedef struct {
    double snr;
    uint64_t on_samples, off_samples;
    double mean_difference, standard_error;
} oracle_measurement;
typedef struct {
    double minimum_validation_snr, minimum_part_snr;
    uint64_t minimum_time_parts, minimum_subbands;
    double minimum_dm_contrast, minimum_dispersion_widths, maximum_zero_dm_spectral_ratio;
} oracle_rules;
typedef struct {
    double dm, spectral_score, spectral_frequency_hz, discovery_profile_response;
    uint32_t has_spectral, pad_;
} oracle_dm_point;
typedef struct {
    uint32_t family, pad_;
    uint64_t index; /* into the generated candidates */
} oracle_member;
typedef struct {
    oracle_candidate candidate;
    double refined_dm;
    double profile_period_s, profile_snr, profile_chi2;
    oracle_window window;
    oracle_measurement validation, zero_dm_validation;
    double zero_dm_period_s;
    oracle_window zero_dm_window;
    double dm_contrast, dispersion_widths, zero_dm_related_spectral_ratio;
    oracle_measurement time_parts[4], subbands[4];
    uint64_t nmembers;
    uint32_t accepted, rejection_mask; /* bit i = i-th reason of confirmation.rs:434-466 */
} oracle_confirmed;
typedef struct {
    uint64_t discovery_samples, validation_start_sample, validation_samples, common_samples,
        trial_count, ncandidates;
} oracle_outcome;

/* confirmation.rs:107-141 */
void oracle_measure(const float *x, size_t n, size_t offset, double tsamp, double period,
                    const oracle_window *w, oracle_measurement *out);
/* confirmation.rs:142-176 */
void oracle_select_window(const float *values, size_t bins, oracle_window *out);
/* confirmation.rs:177-199.  Returns the family count. */
size_t oracle_group(const oracle_candidate *in, size_t n, double resolution, size_t limit,
                    oracle_candidate *leaders, oracle_member *members, size_t *nmembers);
/* confirmation.rs:203-496 over f32 channel-major data.  Arrays are sized by
 * limit (conf), limit*bins (profile_*), limit*4*bins (time/subband profiles),
 * limit*ntrials (dm_response) and ngen (members).  Returns 0 or -1 (err set). */
int oracle_analyze(const float *data, size_t nsamples, const double *freqs, size_t nch, double tsamp,
                   const double *trials, size_t ntrials, const oracle_candidate *generated,
                   size_t ngen, size_t limit, size_t bins, const oracle_rules *rules,
                   oracle_outcome *outcome, oracle_confirmed *conf, float *profile_bins,
                   uint64_t *profile_counts, float *time_profiles, float *subband_profiles,
                   oracle_dm_point *dm_response, oracle_member *members, size_t *nmembers, char *err,
                   size_t errlen);

/* Whole per-trial loop of crates/cli/src/main.rs:462-469 over f32 channel-major data,
 * on `threads` POSIX threads (the CPU baseline).  counts[ntrials], out[ntrials*cap]. */
int oracle_search_trials(const float *data, size_t nsamples, const double *freqs, size_t nch,
                         double tsamp, const double *dms, size_t ntrials, size_t length,
                         const oracle_config *cfg, oracle_candidate *out, size_t cap,
                         uint32_t *counts, int threads);
/* The same loop; series_out (optional) receives every trial's dedispersed series
 * [ntrials][length]; f64 != 0 searches with oracle_periodicity_f64. */
int oracle_search_trials_ex(const float *data, size_t nsamples, const double *freqs, size_t nch,
                            double tsamp, const double *dms, size_t ntrials, size_t length,
                            const oracle_config *cfg, oracle_candidate *out, size_t cap,
                            uint32_t *counts, float *series_out, int f64, int threads);

#ifdef __cplusplus
}
#endif
#endif
