// Confirmation stage of the C ABI (crates/search/src/confirmation.rs:103-496):
// group, window selection, on/off measurements and analyze.  Every
// dedispersion, fold, channel profile and the DM response run on the GPU
// (psr_dedisperse's kernels, k_fold, k_fold_rows, k_dm_response); the window
// search, the measurements (sequential f64 sums, as the reference adds them)
// and the acceptance rules run on the host.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "ctx_internal.hpp"

using namespace psr;

namespace {

int64_t key64(double f) {
    int64_t b;
    std::memcpy(&b, &f, 8);
    b ^= (int64_t)(((uint64_t)(b >> 63)) >> 1);
    return b;
}

// confirmation.rs:103-105
size_t phase_bin(size_t t, double tsamp, double period, size_t bins) {
    const double x = (double)t * tsamp / period;
    const size_t b = (size_t)((x - std::trunc(x)) * (double)bins);
    return std::min(b, bins - 1);
}

// confirmation.rs:47-51
bool contains(const psr_window& w, size_t bin) {
    return (bin + w.bins - w.start_bin) % w.bins < w.width_bins;
}

// confirmation.rs:107-141
void measure(const float* x, size_t n, size_t offset, double tsamp, double period,
             const psr_window& w, psr_measurement* out) {
    size_t cnt[2] = {0, 0};
    double sum[2] = {0.0, 0.0}, sum2[2] = {0.0, 0.0};
    for (size_t i = 0; i < n; i++) {
        const int k = contains(w, phase_bin(i + offset, tsamp, period, w.bins)) ? 1 : 0;
        const double v = (double)x[i];
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
    const double difference = sum[1] / (double)cnt[1] - sum[0] / (double)cnt[0];
    double var[2];
    for (int k = 0; k < 2; k++) {
        const double v = (sum2[k] - sum[k] * sum[k] / (double)cnt[k]) / (double)(cnt[k] - 1);
        var[k] = v > 0.0 ? v : 0.0;  // f64::max(v, 0.): NaN gives 0
    }
    const double error = std::sqrt(var[1] / (double)cnt[1] + var[0] / (double)cnt[0]);
    out->snr = error > 1e-10 ? difference / error : 0.0;
    out->mean_difference = difference;
    out->standard_error = error;
}

// confirmation.rs:142-176: the contiguous (cyclic) power-of-two window with
// the largest on/off contrast score; the first best wins.
psr_window select_window(const float* values, size_t bins) {
    psr_window w{0, 1, bins};
    double best = -INFINITY;
    double total = 0.0;
    for (size_t b = 0; b < bins; b++) total += (double)values[b];
    for (size_t width = 1; width <= bins / 4; width <<= 1) {
        for (size_t start = 0; start < bins; start++) {
            double on = 0.0;
            for (size_t j = 0; j < width; j++) on += (double)values[(start + j) % bins];
            const double contrast = on / (double)width - (total - on) / (double)(bins - width);
            const double score =
                contrast / std::sqrt(1.0 / (double)width + 1.0 / (double)(bins - width));
            if (score > best) {
                best = score;
                w.start_bin = start;
                w.width_bins = width;
            }
        }
    }
    return w;
}

bool related(const psr_candidate& a, const psr_candidate& c, double resolution) {
    for (int h = 1; h <= 16; h++)
        if (std::fabs(a.frequency_hz - c.frequency_hz * (double)h) <= resolution ||
            std::fabs(c.frequency_hz - a.frequency_hz * (double)h) <= resolution)
            return true;
    return false;
}

// confirmation.rs:177-199 (stable sort by score, descending, in total order)
size_t group_impl(const psr_candidate* in, size_t n, double resolution, size_t limit,
                  psr_candidate* leaders, psr_member* members, size_t* nmembers) {
    std::vector<size_t> idx(n);
    for (size_t i = 0; i < n; i++) idx[i] = i;
    std::stable_sort(idx.begin(), idx.end(),
                     [&](size_t a, size_t b) { return key64(in[a].score) > key64(in[b].score); });
    size_t nf = 0, nm = 0;
    for (size_t i : idx) {
        const psr_candidate& c = in[i];
        size_t f = 0;
        while (f < nf && !related(leaders[f], c, resolution)) f++;
        if (f < nf) {
            members[nm++] = psr_member{(uint32_t)f, 0, i};
        } else if (nf < limit) {
            leaders[nf] = c;
            members[nm++] = psr_member{(uint32_t)nf, 0, i};
            nf++;
        }
    }
    *nmembers = nm;
    return nf;
}

int fold_one(psr_ctx* c, const float* x, size_t n, double tsamp, double period, uint64_t offset,
             uint32_t bins, psr_profile_stats* st, float* values, uint64_t* counts) {
    return psr_fold(c, x, n, tsamp, &period, &offset, 1, bins, st, values, counts);
}

}  // namespace

extern "C" {

void psr_rules_default(psr_rules* r) {
    if (!r) return;
    *r = psr_rules{6.0, 1.0, 3, 3, 1.25, 1.5, 1.0};  // confirmation.rs:28-40
}

int psr_measure(const float* x, size_t n, uint64_t offset, double tsamp, double period,
                const psr_window* window, psr_measurement* out) {
    clear_error();
    if ((!x && n) || !window || !out) return fail(PSR_E_INVALID, "null pointer");
    if (window->bins == 0) return fail(PSR_E_CONTRACT, "measure requires a window with bins > 0");
    measure(x, n, (size_t)offset, tsamp, period, *window, out);
    return PSR_OK;
}

int psr_group(const psr_candidate* in, size_t n, double resolution_hz, size_t limit,
              psr_candidate* leaders, size_t* nfamilies, psr_member* members, size_t* nmembers) {
    clear_error();
    if ((!in && n) || (!leaders && limit) || !nfamilies || (!members && n) || !nmembers)
        return fail(PSR_E_INVALID, "null pointer");
    *nfamilies = group_impl(in, n, resolution_hz, limit, leaders, members, nmembers);
    return PSR_OK;
}

int psr_analyze(psr_ctx* c, psr_obs* o, const double* trials, size_t ntrials,
                const psr_candidate* generated, size_t ngenerated, size_t limit, uint32_t bins,
                const psr_rules* rules, psr_outcome* outcome, psr_confirmed* confirmed,
                float* profile_bins, uint64_t* profile_counts, float* time_profiles,
                float* subband_profiles, psr_dm_point* dm_response, psr_member* members,
                size_t* nmembers) {
    clear_error();
    if (!c || !o || !rules || !outcome || !nmembers || (!trials && ntrials) ||
        (!generated && ngenerated) || (!members && ngenerated))
        return fail(PSR_E_INVALID, "null pointer");
    if (ntrials == 0 || limit == 0 || bins < 8 || o->nchans < 4)  // confirmation.rs:217-225
        return fail(PSR_E_INVALID,
                    "confirmation requires >=4 channels, >=8 bins and valid array/trials");
    if (!confirmed || !profile_bins || !profile_counts || !time_profiles || !subband_profiles ||
        !dm_response)
        return fail(PSR_E_INVALID, "null pointer");
    if (bins > 1024) return fail(PSR_E_CONTRACT, "fold requires 4 <= bins <= 1024");
    for (size_t i = 1; i < ntrials; i++)
        if (!(trials[i] > trials[i - 1])) return fail(PSR_E_INVALID, "DM trials must increase");
    int rc;
    if ((rc = check_device(c))) return rc;
    const size_t nch = o->nchans;
    const double tsamp = o->tsamp;
    const double max_dm = trials[ntrials - 1];
    std::vector<uint64_t> d(nch), dd(nch);
    if ((rc = dedisperse::delays(o->freqs.data(), nch, tsamp, max_dm, d.data()))) return rc;
    const uint64_t maxd = *std::max_element(d.begin(), d.end());
    if (maxd > o->nsamples || o->nsamples - maxd < 512)
        return fail(PSR_E_INVALID, "observation too short for confirmation");
    const size_t length = o->nsamples - maxd, guard = maxd;
    if (length < guard || length - guard < 256)
        return fail(PSR_E_INVALID, "DM delay leaves no independent split");
    const size_t discovery = (length - guard) / 2;
    const size_t vstart = discovery + guard, validation = length - vstart;
    const double resolution = 1.0 / ((double)discovery * tsamp);
    const double step = ntrials > 1 ? trials[1] - trials[0] : 1.0;
    *outcome = psr_outcome{discovery, vstart, validation, length, ntrials, 0};

    // zero-DM discovery control (confirmation.rs:252-265) and the zero-DM series
    const double zdm = 0.0;
    psr_search_cfg zc{};
    zc.min_hz = 1.0;
    zc.max_hz = 0.5 / tsamp;
    zc.threshold = 8.0;
    zc.max_candidates = 64;
    zc.max_harmonics = 16;
    zc.precision = PSR_PRECISION_FP32;
    std::vector<psr_candidate> zcand(64);
    uint32_t nz = 0;
    if ((rc = psr_search(c, o, &zdm, 1, discovery, &zc, zcand.data(), 64, &nz))) return rc;
    std::vector<float> zero(length), series(length), x(length);
    if ((rc = dedisperse_to_host(c, o, &zdm, 1, length, zero.data()))) return rc;

    std::vector<psr_candidate> leaders(limit);
    size_t nm = 0;
    const size_t nf =
        group_impl(generated, ngenerated, resolution * 1.5, limit, leaders.data(), members, &nm);
    *nmembers = nm;
    outcome->ncandidates = nf;

    std::vector<float> v(bins), chv(nch * bins), x21(21 * discovery);
    std::vector<uint64_t> cn(bins), ccnt(bins);
    std::vector<double> chsum(nch * bins), means(nch), resp(ntrials);
    for (size_t f = 0; f < nf; f++) {
        const psr_candidate cand = leaders[f];
        psr_confirmed* out = &confirmed[f];
        std::memset(out, 0, sizeof(*out));
        out->candidate = cand;
        // rationally related zero-DM signal (confirmation.rs:268-280)
        double zr = 0.0;
        for (uint32_t i = 0; i < nz; i++) {
            bool rel = false;
            for (int m = 1; m <= 16 && !rel; m++)
                for (int q = 1; q <= 16 && !rel; q++)
                    rel = std::fabs(cand.frequency_hz - zcand[i].frequency_hz * (double)m / (double)q) <=
                          resolution * 1.5;
            if (rel) zr = std::fmax(zr, zcand[i].score / std::fmax(cand.score, 1.0));
        }
        // refine at the candidate DM on the discovery half (:281-294)
        float* pb = profile_bins + f * bins;
        uint64_t* pc = profile_counts + f * bins;
        if ((rc = dedisperse_to_host(c, o, &cand.dm, 1, length, series.data()))) return rc;
        psr_profile_stats best{};
        if ((rc = psr_refine(c, series.data(), discovery, tsamp, cand.frequency_hz, resolution, 40,
                             bins, &best, pb, pc)))
            return rc;
        double best_dm = cand.dm;
        const double period = best.period_s;
        // 21 DM steps of half the trial spacing, first strict maximum (:296-310)
        double dms21[21];
        for (int i = -10; i <= 10; i++) {
            double dm = cand.dm + (double)i * step / 2.0;
            dms21[i + 10] = dm < trials[0] ? trials[0] : (dm > max_dm ? max_dm : dm);
        }
        if ((rc = dedisperse_to_host(c, o, dms21, 21, discovery, x21.data()))) return rc;
        for (int i = 0; i < 21; i++) {
            psr_profile_stats st;
            if ((rc = fold_one(c, x21.data() + (size_t)i * discovery, discovery, tsamp, period, 0,
                               bins, &st, v.data(), cn.data())))
                return rc;
            if (st.reduced_chi_square > best.reduced_chi_square) {
                best = st;
                best_dm = dms21[i];
                std::copy(v.begin(), v.end(), pb);
                std::copy(cn.begin(), cn.end(), pc);
            }
        }
        out->refined_dm = best_dm;
        out->profile = best;
        out->window = select_window(pb, bins);
        // validation at the refined DM (:311-320)
        if ((rc = dedisperse::delays(o->freqs.data(), nch, tsamp, best_dm, dd.data()))) return rc;
        const uint64_t dmax = *std::max_element(dd.begin(), dd.end());
        if ((rc = dedisperse_to_host(c, o, &best_dm, 1, length, series.data()))) return rc;
        measure(series.data() + vstart, length - vstart, vstart, tsamp, period, out->window,
                &out->validation);
        // zero-DM control on the held-out half (:321-350)
        psr_profile_stats zst;
        if ((rc = psr_refine(c, zero.data() + vstart, length - vstart, tsamp, 1.0 / period,
                             resolution * 2.0, 40, bins, &zst, v.data(), cn.data())))
            return rc;
        out->zero_dm_period_s = zst.period_s;
        if ((rc = fold_one(c, zero.data() + vstart, length - vstart, tsamp, zst.period_s, vstart,
                           bins, &zst, v.data(), cn.data())))
            return rc;
        out->zero_dm_window = select_window(v.data(), bins);
        measure(zero.data() + vstart, length - vstart, vstart, tsamp, out->zero_dm_period_s,
                out->zero_dm_window, &out->zero_dm_validation);
        out->dm_contrast =
            std::fmax(out->validation.snr, 0.0) / std::fmax(out->zero_dm_validation.snr, 1.0);
        out->dispersion_widths =
            ((double)dmax * tsamp) / (period * (double)out->window.width_bins / (double)bins);
        // four held-out time parts (:351-361)
        for (size_t part = 0; part < 4; part++) {
            const size_t lo = vstart + part * validation / 4;
            const size_t hi = vstart + (part + 1) * validation / 4;
            measure(series.data() + lo, hi - lo, lo, tsamp, period, out->window, &out->time_parts[part]);
            psr_profile_stats st;
            if ((rc = fold_one(c, series.data() + lo, hi - lo, tsamp, period, lo, bins, &st,
                               time_profiles + (f * 4 + part) * bins, cn.data())))
                return rc;
        }
        // four subbands with the full band's delays (:362-390)
        for (size_t part = 0; part < 4; part++) {
            const size_t lo = part * nch / 4, hi = (part + 1) * nch / 4;
            psr_obs sub = *o;
            sub.nchans = (uint32_t)(hi - lo);
            sub.data = (char*)o->data + lo * o->pitch * sample_bytes(o);
            sub.inv_f2 = o->inv_f2 + lo;
            sub.freqs.assign(o->freqs.begin() + lo, o->freqs.begin() + hi);
            if ((rc = dedisperse_to_host(c, &sub, &best_dm, 1, length, x.data()))) return rc;
            measure(x.data() + vstart, length - vstart, vstart, tsamp, period, out->window,
                    &out->subbands[part]);
            psr_profile_stats st;
            if ((rc = fold_one(c, x.data() + vstart, length - vstart, tsamp, period, vstart, bins,
                               &st, subband_profiles + (f * 4 + part) * bins, cn.data())))
                return rc;
        }
        // first-half channel profiles (:391-396) and the DM response (:397-433)
        if ((rc = ensure(c, c->fold_sums, nch * bins * sizeof(double))) ||
            (rc = ensure(c, c->fold_counts, bins * sizeof(uint64_t))) ||
            (rc = ensure(c, c->an_values, nch * bins * sizeof(float))) ||
            (rc = ensure(c, c->an_means, nch * sizeof(double))) ||
            (rc = ensure(c, c->an_resp, ntrials * sizeof(double))) ||
            (rc = ensure(c, c->dms, ntrials * sizeof(double))) ||
            (rc = ensure(c, c->delays, ntrials * nch * sizeof(int32_t))))
            return rc;
        hipStream_t s = c->stream;
        launch_fold_rows(s, o->data, o->kind, o->pitch, (uint32_t)nch, discovery, tsamp, period,
                         bins, (double*)c->fold_sums.p, (uint64_t*)c->fold_counts.p);
        HIP_TRY(hipGetLastError());
        HIP_TRY(hipMemcpyAsync(chsum.data(), c->fold_sums.p, nch * bins * sizeof(double),
                               hipMemcpyDeviceToHost, s));
        HIP_TRY(hipMemcpyAsync(ccnt.data(), c->fold_counts.p, bins * sizeof(uint64_t),
                               hipMemcpyDeviceToHost, s));
        HIP_TRY(hipStreamSynchronize(s));
        for (size_t ch = 0; ch < nch; ch++) {
            double m = 0.0;
            for (uint32_t b = 0; b < bins; b++) {
                const float val = ccnt[b] > 0 ? (float)(chsum[ch * bins + b] / (double)ccnt[b]) : 0.0f;
                chv[ch * bins + b] = val;
                m += (double)val;
            }
            means[ch] = m / (double)bins;
        }
        HIP_TRY(hipMemcpyAsync(c->an_values.p, chv.data(), nch * bins * sizeof(float),
                               hipMemcpyHostToDevice, s));
        HIP_TRY(hipMemcpyAsync(c->an_means.p, means.data(), nch * sizeof(double),
                               hipMemcpyHostToDevice, s));
        HIP_TRY(hipMemcpyAsync(c->dms.p, trials, ntrials * sizeof(double), hipMemcpyHostToDevice, s));
        launch_delays(s, o->inv_f2, (uint32_t)nch, o->ref_inv, tsamp, (const double*)c->dms.p,
                      (uint32_t)ntrials, (int32_t*)c->delays.p);
        HIP_TRY(hipGetLastError());
        launch_dm_response(s, (const float*)c->an_values.p, (const double*)c->an_means.p,
                           (const int32_t*)c->delays.p, (uint32_t)nch, (uint32_t)ntrials, bins, tsamp,
                           period, (uint32_t)out->window.start_bin, (uint32_t)out->window.width_bins,
                           (double*)c->an_resp.p);
        HIP_TRY(hipGetLastError());
        HIP_TRY(hipMemcpyAsync(resp.data(), c->an_resp.p, ntrials * sizeof(double),
                               hipMemcpyDeviceToHost, s));
        HIP_TRY(hipStreamSynchronize(s));
        for (size_t t = 0; t < ntrials; t++) {
            psr_dm_point& pt = dm_response[f * ntrials + t];
            pt = psr_dm_point{trials[t], 0.0, 0.0, resp[t], 0, 0};
            // the family's best member at this DM (Iterator::max_by keeps the last maximum)
            for (size_t i = 0; i < nm; i++) {
                if (members[i].family != f) continue;
                const psr_candidate& m = generated[members[i].index];
                if (std::fabs(m.dm - trials[t]) < step * 0.01 &&
                    std::fabs(m.frequency_hz - cand.frequency_hz) < resolution * 1.5 &&
                    (!pt.has_spectral || key64(m.score) >= key64(pt.spectral_score))) {
                    pt.has_spectral = 1;
                    pt.spectral_score = m.score;
                    pt.spectral_frequency_hz = m.frequency_hz;
                }
            }
        }
        // acceptance rules (:434-466)
        uint32_t mask = 0;
        if (out->validation.snr < rules->minimum_validation_snr) mask |= 1u << 0;
        uint64_t np = 0, ns = 0;
        for (int i = 0; i < 4; i++) {
            np += out->time_parts[i].snr >= rules->minimum_part_snr;
            ns += out->subbands[i].snr >= rules->minimum_part_snr;
        }
        if (np < rules->minimum_time_parts) mask |= 1u << 1;
        if (ns < rules->minimum_subbands) mask |= 1u << 2;
        if (best_dm <= 0.0 || out->dm_contrast < rules->minimum_dm_contrast) mask |= 1u << 3;
        if (out->dispersion_widths < rules->minimum_dispersion_widths) mask |= 1u << 4;
        if (zr > rules->maximum_zero_dm_spectral_ratio) mask |= 1u << 5;
        out->zero_dm_related_spectral_ratio = zr;
        out->rejection_mask = mask;
        out->accepted = mask == 0;
        uint64_t cnt = 0;
        for (size_t i = 0; i < nm; i++) cnt += members[i].family == f;
        out->nmembers = cnt;
    }
    return PSR_OK;
}

}  // extern "C"
