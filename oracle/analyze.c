/*
 * analyze.c -- CPU restatement of the confirmation stage
 * (crates/search/src/confirmation.rs:103-496): measure, select_window, group
 * and analyze, composed from the oracle's own apply / periodicity / fold.
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h).  Same arithmetic rules: f64 where
 * the reference uses f64, f32 where it uses f32, sums in the reference's order,
 * no contraction (-ffp-contract=off).
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

static inline int64_t key64(double f) {
    int64_t b;
    memcpy(&b, &f, 8);
    b ^= (int64_t)(((uint64_t)(b >> 63)) >> 1);
    return b;
}

/* confirmation.rs:103-105 */
static size_t phase_bin(size_t t, double tsamp, double period, size_t bins) {
    double x = (double)t * tsamp / period;
    size_t b = (size_t)((x - trunc(x)) * (double)bins);
    return b < bins - 1 ? b : bins - 1;
}

/* confirmation.rs:47-51 */
static int window_contains(const oracle_window *w, size_t bin) {
    return (bin + w->bins - w->start_bin) % w->bins < w->width_bins;
}

/* confirmation.rs:107-141 */
void oracle_measure(const float *x, size_t n, size_t offset, double tsamp, double period,
                    const oracle_window *w, oracle_measurement *out) {
    size_t cnt[2] = {0, 0};
    double sum[2] = {0.0, 0.0}, sum2[2] = {0.0, 0.0};
    for (size_t i = 0; i < n; i++) {
        int k = window_contains(w, phase_bin(i + offset, tsamp, period, w->bins));
        double v = (double)x[i];
        cnt[k] += 1;
        sum[k] += v;
        sum2[k] += v * v;
    }
    out->on_samples = cnt[1];
    out->off_samples = cnt[0];
    if (cnt[0] < 2 || cnt[1] < 2) {
        out->snr = 0.0;
        out->mean_difference = 0.0;
        out->standard_error = 0.0;
        return;
    }
    double difference = sum[1] / (double)cnt[1] - sum[0] / (double)cnt[0];
    double var[2];
    for (int k = 0; k < 2; k++) {
        double v = (sum2[k] - sum[k] * sum[k] / (double)cnt[k]) / (double)(cnt[k] - 1);
        var[k] = v > 0.0 ? v : 0.0; /* f64::max(v, 0.) (NaN -> 0) */
    }
    double error = sqrt(var[1] / (double)cnt[1] + var[0] / (double)cnt[0]);
    out->snr = error > 1e-10 ? difference / error : 0.0;
    out->mean_difference = difference;
    out->standard_error = error;
}

/* confirmation.rs:142-176 */
void oracle_select_window(const float *values, size_t bins, oracle_window *out) {
    double best = -INFINITY;
    out->start_bin = 0;
    out->width_bins = 1;
    out->bins = bins;
    for (size_t width = 1; width <= bins / 4; width <<= 1) {
        for (size_t start = 0; start < bins; start++) {
            double on = 0.0;
            for (size_t j = 0; j < width; j++) on += (double)values[(start + j) % bins];
            double total = 0.0;
            for (size_t b = 0; b < bins; b++) total += (double)values[b];
            double contrast = on / (double)width - (total - on) / (double)(bins - width);
            double score = contrast / sqrt(1.0 / (double)width + 1.0 / (double)(bins - width));
            if (score > best) {
                best = score;
                out->start_bin = start;
                out->width_bins = width;
            }
        }
    }
}

/* stable sort of candidate indices by score descending (slice::sort_by + total_cmp) */
static void sort_by_score_desc(const oracle_candidate *c, size_t *idx, size_t n) {
    size_t *tmp = (size_t *)malloc((n ? n : 1) * sizeof(size_t));
    for (size_t w = 1; w < n; w <<= 1) {
        for (size_t lo = 0; lo < n; lo += 2 * w) {
            size_t mid = lo + w < n ? lo + w : n, hi = lo + 2 * w < n ? lo + 2 * w : n;
            size_t a = lo, b = mid, o = lo;
            while (a < mid && b < hi) {
                /* take from the right run only if strictly greater (keeps stability) */
                if (key64(c[idx[b]].score) > key64(c[idx[a]].score))
                    tmp[o++] = idx[b++];
                else
                    tmp[o++] = idx[a++];
            }
            while (a < mid) tmp[o++] = idx[a++];
            while (b < hi) tmp[o++] = idx[b++];
        }
        memcpy(idx, tmp, n * sizeof(size_t));
    }
    free(tmp);
}

static int related(const oracle_candidate *a, const oracle_candidate *c, double resolution) {
    for (int h = 1; h <= 16; h++)
        if (fabs(a->frequency_hz - c->frequency_hz * (double)h) <= resolution ||
            fabs(c->frequency_hz - a->frequency_hz * (double)h) <= resolution)
            return 1;
    return 0;
}

/* confirmation.rs:177-199.  Returns the family count; members[] lists
 * (family, index into `in`) in the order they were appended. */
size_t oracle_group(const oracle_candidate *in, size_t n, double resolution, size_t limit,
                    oracle_candidate *leaders, oracle_member *members, size_t *nmembers) {
    size_t *idx = (size_t *)malloc((n ? n : 1) * sizeof(size_t));
    for (size_t i = 0; i < n; i++) idx[i] = i;
    sort_by_score_desc(in, idx, n);
    size_t nf = 0, nm = 0;
    for (size_t s = 0; s < n; s++) {
        const oracle_candidate *c = &in[idx[s]];
        size_t f = 0;
        while (f < nf && !related(&leaders[f], c, resolution)) f++;
        if (f < nf) {
            members[nm].family = (uint32_t)f;
            members[nm].index = idx[s];
            nm++;
        } else if (nf < limit) {
            leaders[nf] = *c;
            members[nm].family = (uint32_t)nf;
            members[nm].index = idx[s];
            nm++;
            nf++;
        }
    }
    free(idx);
    *nmembers = nm;
    return nf;
}

static int apply_dm(const float *data, size_t nsamples, const double *freqs, size_t nch, double tsamp,
                    double dm, size_t length, float *out, uint64_t *d) {
    size_t len = 0;
    if (oracle_delays(freqs, nch, tsamp, dm, d)) return -1;
    return oracle_apply(data, nsamples, nch, d, length, out, &len);
}

static void fill_err(char *err, size_t errlen, const char *msg) {
    if (err && errlen) snprintf(err, errlen, "%s", msg);
}

/* confirmation.rs:203-496 */
int oracle_analyze(const float *data, size_t nsamples, const double *freqs, size_t nch, double tsamp,
                   const double *trials, size_t ntrials, const oracle_candidate *generated,
                   size_t ngen, size_t limit, size_t bins, const oracle_rules *rules,
                   oracle_outcome *outcome, oracle_confirmed *conf, float *profile_bins,
                   uint64_t *profile_counts, float *time_profiles, float *subband_profiles,
                   oracle_dm_point *dm_response, oracle_member *members, size_t *nmembers, char *err,
                   size_t errlen) {
    if (ntrials == 0 || limit == 0 || bins < 8 || nch < 4) {
        fill_err(err, errlen, "confirmation requires >=4 channels, >=8 bins and valid array/trials");
        return -1;
    }
    for (size_t i = 1; i < ntrials; i++)
        if (!(trials[i] > trials[i - 1])) {
            fill_err(err, errlen, "DM trials must increase");
            return -1;
        }
    const double max_dm = trials[ntrials - 1];
    uint64_t *d = (uint64_t *)malloc(nch * sizeof(uint64_t));
    if (oracle_delays(freqs, nch, tsamp, max_dm, d)) {
        free(d);
        fill_err(err, errlen, "invalid delay inputs");
        return -1;
    }
    uint64_t maxd = 0;
    for (size_t c = 0; c < nch; c++) maxd = d[c] > maxd ? d[c] : maxd;
    if (maxd > nsamples || nsamples - maxd < 512) {
        free(d);
        fill_err(err, errlen, "observation too short for confirmation");
        return -1;
    }
    const size_t length = nsamples - maxd, guard = maxd;
    if (length < guard || length - guard < 256) {
        free(d);
        fill_err(err, errlen, "DM delay leaves no independent split");
        return -1;
    }
    const size_t discovery = (length - guard) / 2;
    const size_t vstart = discovery + guard, validation = length - vstart;
    const double resolution = 1.0 / ((double)discovery * tsamp);
    const double step = ntrials > 1 ? trials[1] - trials[0] : 1.0;
    outcome->discovery_samples = discovery;
    outcome->validation_start_sample = vstart;
    outcome->validation_samples = validation;
    outcome->common_samples = length;
    outcome->trial_count = ntrials;

    float *zero = (float *)malloc(length * sizeof(float));
    float *series = (float *)malloc(length * sizeof(float));
    float *x = (float *)malloc(length * sizeof(float));
    uint64_t *zd = (uint64_t *)calloc(nch, sizeof(uint64_t));
    size_t len = 0;
    /* zero-DM discovery control (confirmation.rs:252-265) */
    oracle_apply(data, nsamples, nch, zd, discovery, zero, &len);
    oracle_config zc = {1.0, 0.5 / tsamp, 8.0, 64, NULL, 0, 16, 0};
    oracle_candidate zcand[64];
    size_t nz = oracle_periodicity(zero, discovery, tsamp, 0.0, &zc, zcand, 64);
    /* zero-DM series over the common length, shared by every family */
    oracle_apply(data, nsamples, nch, zd, length, zero, &len);

    oracle_candidate *leaders = (oracle_candidate *)malloc(limit * sizeof(oracle_candidate));
    size_t nm = 0;
    size_t nf = oracle_group(generated, ngen, resolution * 1.5, limit, leaders, members, &nm);
    *nmembers = nm;
    outcome->ncandidates = nf;

    float *v = (float *)malloc(bins * sizeof(float));
    uint64_t *cn = (uint64_t *)malloc(bins * sizeof(uint64_t));
    float *chp = (float *)malloc(nch * bins * sizeof(float));
    double *combined = (double *)malloc(bins * sizeof(double));
    uint64_t *dd = (uint64_t *)malloc(nch * sizeof(uint64_t));
    for (size_t f = 0; f < nf; f++) {
        const oracle_candidate cand = leaders[f];
        oracle_confirmed *o = &conf[f];
        memset(o, 0, sizeof(*o));
        o->candidate = cand;
        /* rationally related zero-DM signal (confirmation.rs:268-280) */
        double zr = 0.0;
        for (size_t i = 0; i < nz; i++) {
            int rel = 0;
            for (int m = 1; m <= 16 && !rel; m++)
                for (int q = 1; q <= 16 && !rel; q++)
                    if (fabs(cand.frequency_hz - zcand[i].frequency_hz * (double)m / (double)q) <=
                        resolution * 1.5)
                        rel = 1;
            if (rel) {
                double s = zcand[i].score / fmax(cand.score, 1.0);
                zr = fmax(zr, s);
            }
        }
        /* refine at the candidate DM on the discovery half (:281-294) */
        apply_dm(data, nsamples, freqs, nch, tsamp, cand.dm, length, series, dd);
        double bsnr, bchi;
        float *pb = profile_bins + f * bins;
        uint64_t *pc = profile_counts + f * bins;
        double bper = oracle_refine(series, discovery, tsamp, cand.frequency_hz, resolution, 40, bins,
                                    &bsnr, &bchi, pb, pc);
        double best_dm = cand.dm;
        const double period = bper;
        /* 21 DM steps of half the trial spacing, first strict maximum (:296-310) */
        for (int i = -10; i <= 10; i++) {
            double dm = cand.dm + (double)i * step / 2.0;
            dm = dm < trials[0] ? trials[0] : (dm > max_dm ? max_dm : dm);
            apply_dm(data, nsamples, freqs, nch, tsamp, dm, discovery, x, dd);
            double s, c2;
            oracle_profile_at(x, discovery, tsamp, period, bins, 0, &s, &c2, v, cn);
            if (c2 > bchi) {
                bchi = c2;
                bsnr = s;
                bper = period;
                best_dm = dm;
                memcpy(pb, v, bins * sizeof(float));
                memcpy(pc, cn, bins * sizeof(uint64_t));
            }
        }
        o->refined_dm = best_dm;
        o->profile_period_s = bper;
        o->profile_snr = bsnr;
        o->profile_chi2 = bchi;
        oracle_select_window(pb, bins, &o->window);
        /* validation at the refined DM (:312-320) */
        apply_dm(data, nsamples, freqs, nch, tsamp, best_dm, length, series, dd);
        uint64_t dmax = 0;
        for (size_t c = 0; c < nch; c++) dmax = dd[c] > dmax ? dd[c] : dmax;
        oracle_measure(series + vstart, length - vstart, vstart, tsamp, period, &o->window,
                       &o->validation);
        /* zero-DM control on the validation half (:321-350) */
        double zs, zc2;
        o->zero_dm_period_s =
            oracle_refine(zero + vstart, length - vstart, tsamp, 1.0 / period, resolution * 2.0, 40,
                          bins, &zs, &zc2, v, cn);
        oracle_profile_at(zero + vstart, length - vstart, tsamp, o->zero_dm_period_s, bins, vstart,
                          &zs, &zc2, v, cn);
        oracle_select_window(v, bins, &o->zero_dm_window);
        oracle_measure(zero + vstart, length - vstart, vstart, tsamp, o->zero_dm_period_s,
                       &o->zero_dm_window, &o->zero_dm_validation);
        o->dm_contrast = fmax(o->validation.snr, 0.0) / fmax(o->zero_dm_validation.snr, 1.0);
        o->dispersion_widths = ((double)dmax * tsamp) /
                               (period * (double)o->window.width_bins / (double)bins);
        /* time parts (:352-361) */
        for (size_t part = 0; part < 4; part++) {
            size_t lo = vstart + part * validation / 4, hi = vstart + (part + 1) * validation / 4;
            oracle_measure(series + lo, hi - lo, lo, tsamp, period, &o->window, &o->time_parts[part]);
            double s, c2;
            oracle_profile_at(series + lo, hi - lo, tsamp, period, bins, lo, &s, &c2,
                              time_profiles + (f * 4 + part) * bins, cn);
        }
        /* subbands (:362-390) */
        for (size_t part = 0; part < 4; part++) {
            size_t lo = part * nch / 4, hi = (part + 1) * nch / 4;
            oracle_apply(data + lo * nsamples, nsamples, hi - lo, dd + lo, length, x, &len);
            oracle_measure(x + vstart, length - vstart, vstart, tsamp, period, &o->window,
                           &o->subbands[part]);
            double s, c2;
            oracle_profile_at(x + vstart, length - vstart, tsamp, period, bins, vstart, &s, &c2,
                              subband_profiles + (f * 4 + part) * bins, cn);
        }
        /* first-half channel profiles and the DM response (:391-433) */
        for (size_t c = 0; c < nch; c++) {
            double s, c2;
            oracle_profile_at(data + c * nsamples, discovery, tsamp, period, bins, 0, &s, &c2,
                              chp + c * bins, cn);
        }
        for (size_t t = 0; t < ntrials; t++) {
            oracle_dm_point *pt = &dm_response[f * ntrials + t];
            const double dm = trials[t];
            oracle_delays(freqs, nch, tsamp, dm, dd);
            for (size_t b = 0; b < bins; b++) combined[b] = 0.0;
            for (size_t c = 0; c < nch; c++) {
                const float *p = chp + c * bins;
                double mean = 0.0;
                for (size_t b = 0; b < bins; b++) mean += (double)p[b];
                mean /= (double)bins;
                double shift = (double)dd[c] * tsamp / period * (double)bins;
                for (size_t b = 0; b < bins; b++) {
                    double pos = fmod((double)b + shift, (double)bins);
                    if (pos < 0.0) pos += (double)bins; /* rem_euclid */
                    size_t index = (size_t)floor(pos);
                    double fraction = pos - (double)index;
                    combined[b] += ((double)p[index] - mean) * (1.0 - fraction) +
                                   ((double)p[(index + 1) % bins] - mean) * fraction;
                }
            }
            double on = 0.0;
            for (size_t i = 0; i < o->window.width_bins; i++)
                on += combined[(o->window.start_bin + i) % bins];
            on /= (double)o->window.width_bins;
            double tot = 0.0;
            for (size_t b = 0; b < bins; b++) tot += combined[b];
            double off = (tot - on * (double)o->window.width_bins) / (double)(bins - o->window.width_bins);
            pt->dm = dm;
            pt->discovery_profile_response = on - off;
            /* the family's best member at this DM (max_by keeps the last maximum) */
            pt->has_spectral = 0;
            for (size_t i = 0; i < nm; i++) {
                if (members[i].family != f) continue;
                const oracle_candidate *c = &generated[members[i].index];
                if (fabs(c->dm - dm) < step * 0.01 &&
                    fabs(c->frequency_hz - cand.frequency_hz) < resolution * 1.5) {
                    if (!pt->has_spectral || key64(c->score) >= key64(pt->spectral_score)) {
                        pt->has_spectral = 1;
                        pt->spectral_score = c->score;
                        pt->spectral_frequency_hz = c->frequency_hz;
                    }
                }
            }
            if (!pt->has_spectral) {
                pt->spectral_score = 0.0;
                pt->spectral_frequency_hz = 0.0;
            }
        }
        /* rules (:434-466) */
        uint32_t mask = 0;
        if (o->validation.snr < rules->minimum_validation_snr) mask |= 1u << 0;
        uint64_t np = 0, ns = 0;
        for (int i = 0; i < 4; i++) {
            np += o->time_parts[i].snr >= rules->minimum_part_snr;
            ns += o->subbands[i].snr >= rules->minimum_part_snr;
        }
        if (np < rules->minimum_time_parts) mask |= 1u << 1;
        if (ns < rules->minimum_subbands) mask |= 1u << 2;
        if (best_dm <= 0.0 || o->dm_contrast < rules->minimum_dm_contrast) mask |= 1u << 3;
        if (o->dispersion_widths < rules->minimum_dispersion_widths) mask |= 1u << 4;
        if (zr > rules->maximum_zero_dm_spectral_ratio) mask |= 1u << 5;
        o->zero_dm_related_spectral_ratio = zr;
        o->rejection_mask = mask;
        o->accepted = mask == 0;
        uint64_t cnt = 0;
        for (size_t i = 0; i < nm; i++) cnt += members[i].family == f;
        o->nmembers = cnt;
    }
    free(v);
    free(cn);
    free(chp);
    free(combined);
    free(dd);
    free(leaders);
    free(zero);
    free(series);
    free(x);
    free(zd);
    free(d);
    return 0;
}
