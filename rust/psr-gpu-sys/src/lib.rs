//! Raw bindings of include/psr_gpu.h (one `#[repr(C)]` struct per C struct,
//! same field order and types) and a thin safe layer in `gpu` that the
//! reference's call sites use (crates/cli/src/main.rs:462-473,
//! crates/search/src/confirmation.rs:528-536; see ../patches).
#![allow(non_camel_case_types)]
use std::os::raw::{c_char, c_int};

pub const PSR_OK: c_int = 0;
pub const PSR_E_INVALID: c_int = -1;
pub const PSR_E_SHAPE: c_int = -2;
pub const PSR_E_DEVICE: c_int = -3;
pub const PSR_E_OOM: c_int = -4;
pub const PSR_E_FORMAT: c_int = -5;
pub const PSR_E_CONTRACT: c_int = -6;
pub const PSR_E_OVERFLOW: c_int = -7;
pub const PSR_PRECISION_FP32: u32 = 0;
pub const PSR_PRECISION_FP16: u32 = 1;
pub const PSR_CTX_OVERLAP: u32 = 1;
pub const PSR_BATCH_GRAPH: u32 = 1;
pub const PSR_BATCH_AUTOTUNE: u32 = 2;
pub const PSR_NOALLOC_MAX_SERIES: usize = 131072;
pub const PSR_NOALLOC_MAX_CANDIDATES: usize = 256;

#[repr(C)]
pub struct psr_ctx {
    _p: [u8; 0],
}
#[repr(C)]
pub struct psr_obs {
    _p: [u8; 0],
}
#[repr(C)]
pub struct psr_batch {
    _p: [u8; 0],
}

#[repr(C)]
#[derive(Clone, Copy, Debug, Default, PartialEq)]
pub struct psr_candidate {
    pub frequency_hz: f64,
    pub period_s: f64,
    pub dm: f64,
    pub harmonics: u64,
    pub score: f64,
}

#[repr(C)]
#[derive(Clone, Copy, Debug)]
pub struct psr_search_cfg {
    pub min_hz: f64,
    pub max_hz: f64,
    pub threshold: f64,
    pub max_candidates: u64,
    pub birdies: *const f64,
    pub nbirdies: u64,
    pub max_harmonics: u32,
    pub precision: u32,
}

#[repr(C)]
#[derive(Clone, Copy, Debug, Default)]
pub struct psr_header {
    pub nchans: u32,
    pub nbits: u32,
    pub tsamp: f64,
    pub fch1: f64,
    pub foff: f64,
    pub data_offset: u64,
    pub nsamples: u64,
}

#[repr(C)]
#[derive(Clone, Copy, Debug)]
pub struct psr_filterbank {
    pub payload: *const u8,
    pub payload_bytes: u64,
    pub nbits: u32,
    pub nchans: u32,
    pub nsamples: u64,
    pub tsamp: f64,
    pub fch1: f64,
    pub foff: f64,
}

#[repr(C)]
#[derive(Clone, Copy, Debug, Default)]
pub struct psr_profile_stats {
    pub period_s: f64,
    pub snr: f64,
    pub reduced_chi_square: f64,
}

#[repr(C)]
#[derive(Clone, Copy, Debug, Default)]
pub struct psr_stage_times {
    pub total_ms: f64,
    pub delays_ms: f64,
    pub dedisperse_ms: f64,
    pub fft_ms: f64,
    pub whiten_ms: f64,
    pub harmsum_ms: f64,
    pub finalize_ms: f64,
    pub trials: u64,
    pub series_len: u64,
    pub fft_len: u64,
    pub launches: u64,
}

#[repr(C)]
#[derive(Clone, Copy, Debug, Default)]
pub struct psr_mask {
    pub nchannels: u64,
    pub nblocks: u64,
    pub block_samples: u64,
    pub masked_fraction: f64,
}

#[repr(C)]
#[derive(Clone, Copy, Debug, Default)]
pub struct psr_batch_opts {
    pub flags: u32,
    pub pad_: u32,
    pub memory_budget: u64,
}

#[repr(C)]
#[derive(Clone, Copy, Debug, Default)]
pub struct psr_window {
    pub start_bin: u64,
    pub width_bins: u64,
    pub bins: u64,
}

#[repr(C)]
#[derive(Clone, Copy, Debug, Default)]
pub struct psr_measurement {
    pub snr: f64,
    pub on_samples: u64,
    pub off_samples: u64,
    pub mean_difference: f64,
    pub standard_error: f64,
}

#[repr(C)]
#[derive(Clone, Copy, Debug, Default)]
pub struct psr_rules {
    pub minimum_validation_snr: f64,
    pub minimum_part_snr: f64,
    pub minimum_time_parts: u64,
    pub minimum_subbands: u64,
    pub minimum_dm_contrast: f64,
    pub minimum_dispersion_widths: f64,
    pub maximum_zero_dm_spectral_ratio: f64,
}

#[repr(C)]
#[derive(Clone, Copy, Debug, Default)]
pub struct psr_dm_point {
    pub dm: f64,
    pub spectral_score: f64,
    pub spectral_frequency_hz: f64,
    pub discovery_profile_response: f64,
    pub has_spectral: u32,
    pub pad_: u32,
}

#[repr(C)]
#[derive(Clone, Copy, Debug, Default)]
pub struct psr_member {
    pub family: u32,
    pub pad_: u32,
    pub index: u64,
}

#[repr(C)]
#[derive(Clone, Copy, Debug, Default)]
pub struct psr_confirmed {
    pub candidate: psr_candidate,
    pub refined_dm: f64,
    pub profile: psr_profile_stats,
    pub window: psr_window,
    pub validation: psr_measurement,
    pub zero_dm_validation: psr_measurement,
    pub zero_dm_period_s: f64,
    pub zero_dm_window: psr_window,
    pub dm_contrast: f64,
    pub dispersion_widths: f64,
    pub zero_dm_related_spectral_ratio: f64,
    pub time_parts: [psr_measurement; 4],
    pub subbands: [psr_measurement; 4],
    pub nmembers: u64,
    pub accepted: u32,
    pub rejection_mask: u32,
}

#[repr(C)]
#[derive(Clone, Copy, Debug, Default)]
pub struct psr_outcome {
    pub discovery_samples: u64,
    pub validation_start_sample: u64,
    pub validation_samples: u64,
    pub common_samples: u64,
    pub trial_count: u64,
    pub ncandidates: u64,
}

extern "C" {
    pub fn psr_version() -> *const c_char;
    pub fn psr_last_error() -> *const c_char;
    pub fn psr_ctx_create(device: c_int, flags: u32, out: *mut *mut psr_ctx) -> c_int;
    pub fn psr_ctx_destroy(ctx: *mut psr_ctx);
    pub fn psr_ctx_set_memory_cap(ctx: *mut psr_ctx, bytes: u64) -> c_int;

    pub fn psr_delays(freqs: *const f64, nch: usize, tsamp: f64, dm: f64, out: *mut u64) -> c_int;
    pub fn psr_dm_step(freqs: *const f64, nch: usize, tsamp: f64, out: *mut f64) -> c_int;
    pub fn psr_plan(min: f64, max: f64, step: f64, out: *mut f64, cap: usize, count: *mut usize) -> c_int;
    pub fn psr_frequencies(fch1: f64, foff: f64, nch: usize, out: *mut f64) -> c_int;
    pub fn psr_parse_header(bytes: *const u8, len: usize, out: *mut psr_header) -> c_int;
    pub fn psr_sift(input: *const psr_candidate, n: usize, resolution_hz: f64, limit: usize,
                    reject_dm_zero: c_int, out: *mut psr_candidate, count: *mut usize) -> c_int;

    pub fn psr_obs_create(ctx: *mut psr_ctx, fb: *const psr_filterbank, out: *mut *mut psr_obs) -> c_int;
    pub fn psr_obs_create_f32(ctx: *mut psr_ctx, data: *const f32, nsamples: u64, nchans: u32, tsamp: f64,
                              fch1: f64, foff: f64, out: *mut *mut psr_obs) -> c_int;
    pub fn psr_obs_destroy(obs: *mut psr_obs);
    pub fn psr_obs_download_f32(obs: *mut psr_obs, out: *mut f32) -> c_int;
    pub fn psr_obs_mask(ctx: *mut psr_ctx, obs: *mut psr_obs, block_samples: u64, threshold: f32,
                        out: *mut psr_mask, channels: *mut u64, blocks: *mut u64) -> c_int;

    pub fn psr_dedisperse(ctx: *mut psr_ctx, obs: *mut psr_obs, dms: *const f64, ntrials: usize, length: usize,
                          out: *mut f32, out_len: *mut usize) -> c_int;
    pub fn psr_apply(ctx: *mut psr_ctx, obs: *mut psr_obs, delays: *const u64, ch_lo: u32, ch_hi: u32,
                     length: usize, out: *mut f32, out_len: *mut usize) -> c_int;
    pub fn psr_search(ctx: *mut psr_ctx, obs: *mut psr_obs, dms: *const f64, ntrials: usize, length: usize,
                      cfg: *const psr_search_cfg, out: *mut psr_candidate, cap_per_trial: usize,
                      counts: *mut u32) -> c_int;
    pub fn psr_last_stage_times(ctx: *const psr_ctx, out: *mut psr_stage_times) -> c_int;
    pub fn psr_periodicity(ctx: *mut psr_ctx, series: *const f32, n: usize, tsamp: f64, dm: f64,
                           cfg: *const psr_search_cfg, out: *mut psr_candidate, cap: usize,
                           count: *mut usize) -> c_int;
    pub fn psr_periodicity_noalloc(ctx: *mut psr_ctx, series: *const f32, n: usize, tsamp: f64, dm: f64,
                                   cfg: *const psr_search_cfg, out: *mut psr_candidate, cap: usize,
                                   count: *mut usize) -> c_int;
    pub fn psr_spectrum(ctx: *mut psr_ctx, series: *const f32, n: usize, tsamp: f64, cfg: *const psr_search_cfg,
                        raw: *mut f32, white: *mut f32, nbins: *mut usize) -> c_int;
    pub fn psr_spectrum_split(ctx: *mut psr_ctx, series: *const f32, n: usize, tsamp: f64,
                              cfg: *const psr_search_cfg, m2: u32, raw: *mut f32, white: *mut f32,
                              nbins: *mut usize) -> c_int;

    pub fn psr_batch_create(ctx: *mut psr_ctx, obs: *const *mut psr_obs, nobs: usize, dms: *const *const f64,
                            ntrials: *const usize, lengths: *const usize, cfg: *const psr_search_cfg,
                            cap_per_trial: usize, opts: *const psr_batch_opts, out: *mut *mut psr_batch) -> c_int;
    pub fn psr_batch_run(batch: *mut psr_batch, out: *mut psr_candidate, counts: *mut u32) -> c_int;
    pub fn psr_batch_info(batch: *const psr_batch, json: *mut c_char, cap: usize, len: *mut usize) -> c_int;
    pub fn psr_batch_destroy(batch: *mut psr_batch);

    pub fn psr_fold(ctx: *mut psr_ctx, series: *const f32, n: usize, tsamp: f64, periods: *const f64,
                    offsets: *const u64, nprof: usize, bins: u32, stats: *mut psr_profile_stats, values: *mut f32,
                    counts: *mut u64) -> c_int;
    pub fn psr_refine(ctx: *mut psr_ctx, series: *const f32, n: usize, tsamp: f64, frequency: f64, span_hz: f64,
                      steps: usize, bins: u32, stats: *mut psr_profile_stats, values: *mut f32,
                      counts: *mut u64) -> c_int;

    pub fn psr_rules_default(out: *mut psr_rules);
    pub fn psr_measure(x: *const f32, n: usize, offset: u64, tsamp: f64, period: f64, window: *const psr_window,
                       out: *mut psr_measurement) -> c_int;
    pub fn psr_group(input: *const psr_candidate, n: usize, resolution_hz: f64, limit: usize,
                     leaders: *mut psr_candidate, nfamilies: *mut usize, members: *mut psr_member,
                     nmembers: *mut usize) -> c_int;
    pub fn psr_analyze(ctx: *mut psr_ctx, obs: *mut psr_obs, trials: *const f64, ntrials: usize,
                       generated: *const psr_candidate, ngenerated: usize, limit: usize, bins: u32,
                       rules: *const psr_rules, outcome: *mut psr_outcome, confirmed: *mut psr_confirmed,
                       profile_bins: *mut f32, profile_counts: *mut u64, time_profiles: *mut f32,
                       subband_profiles: *mut f32, dm_response: *mut psr_dm_point, members: *mut psr_member,
                       nmembers: *mut usize) -> c_int;
}

/// Thin safe layer: owning handles, `Result`s carrying the library's message,
/// and the crate functions' argument shapes.
pub mod gpu {
    use super::*;
    use std::ffi::CStr;
    use std::fmt;

    /// A non-zero status and psr_last_error()'s text (which already reads like
    /// the reference's errors: "invalid dedispersion input: ...").
    #[derive(Debug, Clone)]
    pub struct Error {
        pub code: c_int,
        pub message: String,
    }
    impl fmt::Display for Error {
        fn fmt(&self, f: &mut fmt::Formatter<'_>) -> fmt::Result {
            f.write_str(&self.message)
        }
    }
    impl std::error::Error for Error {}

    /// Ok on PSR_OK; PSR_E_CONTRACT panics, because the reference asserts there
    /// (crates/search/src/lib.rs:34, crates/fold/src/lib.rs:23,89).
    pub fn check(rc: c_int) -> Result<(), Error> {
        if rc == PSR_OK {
            return Ok(());
        }
        let message = unsafe { CStr::from_ptr(psr_last_error()) }.to_string_lossy().into_owned();
        if rc == PSR_E_CONTRACT {
            panic!("{}", message);
        }
        Err(Error { code: rc, message })
    }

    /// Search settings: search::Config plus the harmonic count and storage precision.
    #[derive(Debug, Clone)]
    pub struct Config {
        pub min_hz: f64,
        pub max_hz: f64,
        pub threshold: f64,
        pub max_candidates: usize,
        pub birdies: Vec<(f64, f64)>,
        pub max_harmonics: u32,
        pub precision: u32,
    }
    impl Config {
        fn raw(&self, flat: &mut Vec<f64>) -> psr_search_cfg {
            flat.clear();
            for &(c, w) in &self.birdies {
                flat.push(c);
                flat.push(w);
            }
            psr_search_cfg {
                min_hz: self.min_hz,
                max_hz: self.max_hz,
                threshold: self.threshold,
                max_candidates: self.max_candidates as u64,
                birdies: if flat.is_empty() { std::ptr::null() } else { flat.as_ptr() },
                nbirdies: self.birdies.len() as u64,
                max_harmonics: self.max_harmonics,
                precision: self.precision,
            }
        }
    }

    /// A psr_ctx (one HIP stream and its device buffers); not thread-safe.
    pub struct Context {
        raw: *mut psr_ctx,
    }
    unsafe impl Send for Context {}
    impl Context {
        pub fn new(device: i32, flags: u32) -> Result<Self, Error> {
            let mut raw = std::ptr::null_mut();
            check(unsafe { psr_ctx_create(device, flags, &mut raw) })?;
            Ok(Context { raw })
        }
        pub fn as_ptr(&self) -> *mut psr_ctx {
            self.raw
        }
        /// Upload a packed SIGPROC payload (fbank::decode's layout, unpacked on the GPU).
        pub fn observe(&self, payload: &[u8], nbits: u32, nchans: u32, tsamp: f64, fch1: f64, foff: f64)
                       -> Result<Observation<'_>, Error> {
            let fb = psr_filterbank {
                payload: payload.as_ptr(),
                payload_bytes: payload.len() as u64,
                nbits,
                nchans,
                nsamples: 0,
                tsamp,
                fch1,
                foff,
            };
            let mut raw = std::ptr::null_mut();
            check(unsafe { psr_obs_create(self.raw, &fb, &mut raw) })?;
            Ok(Observation { ctx: self, raw, nchans })
        }
        /// Upload f32 channel-major data (the `data` of dedisperse::apply).
        pub fn observe_f32(&self, data: &[f32], nsamples: usize, tsamp: f64, fch1: f64, foff: f64)
                           -> Result<Observation<'_>, Error> {
            let nchans = if nsamples == 0 { 0 } else { (data.len() / nsamples) as u32 };
            let mut raw = std::ptr::null_mut();
            check(unsafe {
                psr_obs_create_f32(self.raw, data.as_ptr(), nsamples as u64, nchans, tsamp, fch1, foff, &mut raw)
            })?;
            Ok(Observation { ctx: self, raw, nchans })
        }
        /// search::periodicity on one host series.
        pub fn periodicity(&self, series: &[f32], tsamp: f64, dm: f64, cfg: &Config)
                           -> Result<Vec<psr_candidate>, Error> {
            let mut flat = vec![];
            let c = cfg.raw(&mut flat);
            let mut out = vec![psr_candidate::default(); cfg.max_candidates.max(1)];
            let mut n = 0usize;
            check(unsafe {
                psr_periodicity(self.raw, series.as_ptr(), series.len(), tsamp, dm, &c, out.as_mut_ptr(),
                                out.len(), &mut n)
            })?;
            out.truncate(n);
            Ok(out)
        }
    }
    impl Drop for Context {
        fn drop(&mut self) {
            unsafe { psr_ctx_destroy(self.raw) }
        }
    }

    /// A device-resident observation of a Context.
    pub struct Observation<'a> {
        ctx: &'a Context,
        raw: *mut psr_obs,
        nchans: u32,
    }
    impl<'a> Observation<'a> {
        pub fn as_ptr(&self) -> *mut psr_obs {
            self.raw
        }
        /// dedisperse::apply(data, nsamples, delays, length) over channels lo..hi
        /// (the whole band: 0..nchans).
        pub fn apply(&self, delays: &[usize], channels: std::ops::Range<u32>, length: Option<usize>)
                     -> Result<Vec<f32>, Error> {
            let d: Vec<u64> = delays.iter().map(|&x| x as u64).collect();
            let mut n = 0usize;
            let len = length.unwrap_or(0);
            check(unsafe {
                psr_apply(self.ctx.raw, self.raw, d.as_ptr(), channels.start, channels.end, len,
                          std::ptr::null_mut(), &mut n)
            })?;
            let mut out = vec![0f32; n];
            check(unsafe {
                psr_apply(self.ctx.raw, self.raw, d.as_ptr(), channels.start, channels.end, len,
                          out.as_mut_ptr(), &mut n)
            })?;
            Ok(out)
        }
        /// The per-trial loop of crates/cli/src/main.rs:462-469 in one call: the
        /// candidates of every DM, in the given DM order (periodicity's order
        /// within a trial).
        pub fn search(&self, dms: &[f64], length: usize, cfg: &Config) -> Result<Vec<Vec<psr_candidate>>, Error> {
            let mut flat = vec![];
            let c = cfg.raw(&mut flat);
            let cap = cfg.max_candidates.max(1);
            let mut out = vec![psr_candidate::default(); cap * dms.len()];
            let mut counts = vec![0u32; dms.len()];
            check(unsafe {
                psr_search(self.ctx.raw, self.raw, dms.as_ptr(), dms.len(), length, &c, out.as_mut_ptr(), cap,
                           counts.as_mut_ptr())
            })?;
            Ok((0..dms.len()).map(|t| out[t * cap..t * cap + counts[t] as usize].to_vec()).collect())
        }
        /// rfi::mask in place on the device samples: (channels, blocks, fraction).
        pub fn mask(&self, block_samples: usize, threshold: f32, nsamples: usize)
                    -> Result<(Vec<usize>, Vec<usize>, f64), Error> {
            let mut m = psr_mask::default();
            let mut ch = vec![0u64; self.nchans as usize];
            let mut bl = vec![0u64; (nsamples + block_samples - 1) / block_samples.max(1)];
            check(unsafe {
                psr_obs_mask(self.ctx.raw, self.raw, block_samples as u64, threshold, &mut m, ch.as_mut_ptr(),
                             bl.as_mut_ptr())
            })?;
            ch.truncate(m.nchannels as usize);
            bl.truncate(m.nblocks as usize);
            Ok((ch.into_iter().map(|x| x as usize).collect(), bl.into_iter().map(|x| x as usize).collect(),
                m.masked_fraction))
        }
    }
    impl Drop for Observation<'_> {
        fn drop(&mut self) {
            unsafe { psr_obs_destroy(self.raw) }
        }
    }
}
