/*
 * oracle.c -- CPU restatement of the PulsarSr search path (test infrastructure).
 *
 * This file is the checker, not the product: only tests/, __graft_entry__.smoke()
 * and bench.py's cpu_baseline leg load it.  Each function cites the reference
 * file:line it restates.  Build with -O2 -ffp-contract=off (no contraction, no
 * fast-math): the Rust reference never fuses multiply-add.
 */
#include "oracle.h"

#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

/* ---------------------------------------------------------------- helpers */

/* f32::total_cmp key (crates/rfi/src/lib.rs:9 sorts with total_cmp). */
static inline int32_t total_key32(float f) {
    int32_t b;
    memcpy(&b, &f, 4);
    b ^= (int32_t)(((uint32_t)(b >> 31)) >> 1);
    return b;
}
static inline int64_t total_key64(double f) {
    int64_t b;
    memcpy(&b, &f, 8);
    b ^= (int64_t)(((uint64_t)(b >> 63)) >> 1);
    return b;
}

/* k-th smallest key (0-based) by three-way quickselect; keys are permuted. */
#define DEFINE_SELECT(NAME, T)                                                              \
    static T NAME(T *a, size_t n, size_t k) {                                                 \
        size_t lo = 0, hi = n;                                                                \
        while (hi - lo > 1) {                                                                 \
            T p = a[lo + (hi - lo) / 2];                                                      \
            size_t lt = lo, i = lo, gt = hi;                                                  \
            while (i < gt) {                                                                  \
                if (a[i] < p) {                                                               \
                    T t = a[lt];                                                              \
                    a[lt++] = a[i];                                                           \
                    a[i++] = t;                                                               \
                } else if (a[i] > p) {                                                        \
                    T t = a[--gt];                                                            \
                    a[gt] = a[i];                                                             \
                    a[i] = t;                                                                 \
                } else {                                                                      \
                    i++;                                                                      \
                }                                                                             \
            }                                                                                 \
            if (k < lt) hi = lt;                                                              \
            else if (k >= gt) lo = gt;                                                        \
            else return p;                                                                    \
        }                                                                                     \
        return a[lo];                                                                         \
    }
DEFINE_SELECT(select_k32, int32_t)
DEFINE_SELECT(select_k64, int64_t)

static inline float key_to_f32(int32_t k) {
    k ^= (int32_t)(((uint32_t)(k >> 31)) >> 1);
    float f;
    memcpy(&f, &k, 4);
    return f;
}
static inline double key_to_f64(int64_t k) {
    k ^= (int64_t)(((uint64_t)(k >> 63)) >> 1);
    double f;
    memcpy(&f, &k, 8);
    return f;
}

/* crates/rfi/src/lib.rs:4-12: v[len/2] after select_nth_unstable_by(total_cmp). */
float oracle_median(const float *v, size_t n) {
    if (n == 0) return 0.0f;
    int32_t *k = (int32_t *)malloc(n * sizeof(int32_t));
    for (size_t i = 0; i < n; i++) k[i] = total_key32(v[i]);
    float r = key_to_f32(select_k32(k, n, n / 2));
    free(k);
    return r;
}
double oracle_median64(const double *v, size_t n) {
    if (n == 0) return 0.0;
    int64_t *k = (int64_t *)malloc(n * sizeof(int64_t));
    for (size_t i = 0; i < n; i++) k[i] = total_key64(v[i]);
    double r = key_to_f64(select_k64(k, n, n / 2));
    free(k);
    return r;
}

/* ---------------------------------------------------------- dedispersion */

/* crates/dedisperse/src/lib.rs:9-34 */
int oracle_delays(const double *freqs, size_t nch, double tsamp, double dm, uint64_t *out) {
    if (nch == 0 || !isfinite(tsamp) || tsamp <= 0.0 || !isfinite(dm) || dm < 0.0) return -1;
    for (size_t c = 0; c < nch; c++)
        if (!isfinite(freqs[c]) || freqs[c] <= 0.0) return -1;
    double reference = 0.0; /* fold(0.0, f64::max) */
    for (size_t c = 0; c < nch; c++) reference = fmax(reference, freqs[c]);
    double ref_inv = 1.0 / (reference * reference); /* powi(-2) == 1/(f*f) */
    for (size_t c = 0; c < nch; c++) {
        double f = freqs[c];
        double inv = 1.0 / (f * f);
        double samples = round(ORACLE_DISPERSION_SECONDS_MHZ * dm * (inv - ref_inv) / tsamp);
        if (samples > 18446744073709551615.0) return -1;
        double s = fmax(samples, 0.0);
        out[c] = (uint64_t)s;
    }
    return 0;
}

/* crates/dedisperse/src/lib.rs:35-44 */
int oracle_dm_step(const double *freqs, size_t nch, double tsamp, double *out) {
    uint64_t *tmp = (uint64_t *)malloc((nch ? nch : 1) * sizeof(uint64_t));
    int rc = oracle_delays(freqs, nch, tsamp, 0.0, tmp);
    free(tmp);
    if (rc) return rc;
    double lo = INFINITY, hi = 0.0;
    for (size_t c = 0; c < nch; c++) {
        lo = fmin(lo, freqs[c]);
        hi = fmax(hi, freqs[c]);
    }
    double span = ORACLE_DISPERSION_SECONDS_MHZ * (1.0 / (lo * lo) - 1.0 / (hi * hi));
    if (span <= 0.0) return -1;
    *out = tsamp / span;
    return 0;
}

/* crates/dedisperse/src/lib.rs:45-60 */
long oracle_plan(double min, double max, double step, double *out, size_t cap) {
    if (!isfinite(min) || !isfinite(max) || !isfinite(step) || min < 0.0 || max < min ||
        step <= 0.0)
        return -1;
    double q = floor((max - min) / step);
    size_t n = (size_t)q + 1;
    if (n > 100000) return -1;
    if (n > cap) return -1;
    for (size_t i = 0; i < n; i++) out[i] = min + (double)i * step;
    return (long)n;
}

/* crates/dedisperse/src/lib.rs:62-90: channel-order f32 sums from 0.0 */
int oracle_apply(const float *data, size_t nsamples, size_t nch, const uint64_t *delays,
                 size_t length, float *out, size_t *out_len) {
    if (nch == 0 || nsamples == 0) return -1;
    uint64_t maxd = 0;
    for (size_t c = 0; c < nch; c++)
        if (delays[c] > maxd) maxd = delays[c];
    if (maxd > nsamples) return -1;
    size_t available = nsamples - (size_t)maxd;
    size_t n = length ? length : available;
    if (n < 2 || n > available) return -1;
    for (size_t t = 0; t < n; t++) out[t] = 0.0f;
    for (size_t c = 0; c < nch; c++) {
        const float *row = data + c * nsamples + delays[c];
        for (size_t t = 0; t < n; t++) out[t] += row[t];
    }
    *out_len = n;
    return 0;
}

/* crates/fbank/src/lib.rs:149-196 (payload decode, time-major -> channel-major) */
typedef struct {
    const uint8_t *payload;
    size_t nchans, ns;
    float *out;
} decode8_job;
typedef struct {
    decode8_job *j;
    size_t t0, t1;
} decode8_part;
static void *decode8_worker(void *arg) {
    decode8_part *p = (decode8_part *)arg;
    const decode8_job *j = p->j;
    for (size_t t0 = p->t0; t0 < p->t1; t0 += 256) {
        size_t t1 = t0 + 256 < p->t1 ? t0 + 256 : p->t1;
        for (size_t ch = 0; ch < j->nchans; ch++) {
            float *o = j->out + ch * j->ns;
            const uint8_t *q = j->payload + ch;
            for (size_t t = t0; t < t1; t++) o[t] = (float)q[t * j->nchans];
        }
    }
    return NULL;
}

int oracle_decode(const uint8_t *payload, size_t payload_len, int nbits, size_t nchans,
                  size_t limit, float *out, size_t *nsamples_out) {
    if (!(nbits == 1 || nbits == 2 || nbits == 4 || nbits == 8 || nbits == 16 || nbits == 32))
        return -1;
    size_t bits_per_frame = nchans * (size_t)nbits;
    size_t bits = payload_len * 8;
    if (bits_per_frame == 0 || bits % bits_per_frame) return -1;
    size_t total = bits / bits_per_frame;
    size_t ns = limit ? (limit < total ? limit : total) : total;
    if (ns < 2) return -1;
    /* frames in tiles of 256 so the channel-major stores run contiguously */
    if (nbits == 8) { /* every byte is a sample: no checks needed; split over threads */
        decode8_job j = {payload, nchans, ns, out};
        size_t nth = ns * nchans > ((size_t)1 << 26) ? 16 : 1;
        pthread_t th[16];
        decode8_part parts[16];
        for (size_t i = 0; i < nth; i++) {
            parts[i].j = &j;
            parts[i].t0 = (ns * i / nth) & ~(size_t)255;
            parts[i].t1 = i + 1 == nth ? ns : (ns * (i + 1) / nth) & ~(size_t)255;
            pthread_create(&th[i], NULL, decode8_worker, &parts[i]);
        }
        for (size_t i = 0; i < nth; i++) pthread_join(th[i], NULL);
        *nsamples_out = ns;
        return 0;
    }
    for (size_t t0 = 0; t0 < ns; t0 += 256)
        for (size_t ch = 0; ch < nchans; ch++)
            for (size_t t = t0; t < ns && t < t0 + 256; t++) {
            size_t i = t * nchans + ch;
            float v;
            if (nbits <= 4) {
                size_t bit = i * (size_t)nbits;
                v = (float)((payload[bit / 8] >> (bit % 8)) & ((1u << nbits) - 1u));
            } else if (nbits == 8) {
                v = (float)payload[i];
            } else if (nbits == 16) {
                v = (float)(uint16_t)(payload[2 * i] | (payload[2 * i + 1] << 8));
            } else {
                memcpy(&v, payload + 4 * i, 4);
            }
            if (!isfinite(v)) return -1;
            out[ch * ns + t] = v;
        }
    *nsamples_out = ns;
    return 0;
}

/* ------------------------------------------------------------------ RFI */

/* crates/rfi/src/lib.rs:13-22 */
static void outliers(const float *v, size_t n, float threshold, uint8_t *flag) {
    float center = oracle_median(v, n);
    float *dev = (float *)malloc((n ? n : 1) * sizeof(float));
    for (size_t i = 0; i < n; i++) dev[i] = fabsf(v[i] - center);
    float mad = oracle_median(dev, n);
    free(dev);
    float scale = fmaxf(fmaxf(1.4826f * mad, fabsf(center) * 1e-4f), 1e-6f);
    for (size_t i = 0; i < n; i++) flag[i] = fabsf(v[i] - center) > threshold * scale;
}

/* crates/rfi/src/lib.rs:31-96.  Returns masked_fraction. */
double oracle_mask(float *data, size_t nsamples, size_t nch, size_t block_samples, float threshold,
                   uint8_t *bad, uint8_t *bad_blocks) {
    float *means = (float *)malloc(nch * sizeof(float));
    float *vars = (float *)malloc(nch * sizeof(float));
    for (size_t c = 0; c < nch; c++) {
        const float *row = data + c * nsamples;
        float s = 0.0f;
        for (size_t t = 0; t < nsamples; t++) s += row[t];
        means[c] = s / (float)nsamples;
    }
    for (size_t c = 0; c < nch; c++) {
        const float *row = data + c * nsamples;
        float m = means[c], s = 0.0f;
        for (size_t t = 0; t < nsamples; t++) {
            float d = row[t] - m;
            s += d * d;
        }
        vars[c] = s / (float)nsamples;
    }
    uint8_t *mb = (uint8_t *)malloc(nch), *vb = (uint8_t *)malloc(nch);
    outliers(means, nch, threshold, mb);
    outliers(vars, nch, threshold, vb);
    size_t nbad = 0;
    for (size_t c = 0; c < nch; c++) {
        bad[c] = mb[c] || vb[c] || vars[c] < 1e-12f;
        nbad += bad[c];
    }
    size_t blocks_n = (nsamples + block_samples - 1) / block_samples;
    size_t active = nch - nbad;
    float *bm = (float *)malloc(blocks_n * sizeof(float));
    for (size_t b = 0; b < blocks_n; b++) {
        size_t start = b * block_samples;
        size_t end = start + block_samples < nsamples ? start + block_samples : nsamples;
        float sum = 0.0f;
        for (size_t c = 0; c < nch; c++) {
            if (bad[c]) continue;
            float part = 0.0f;
            for (size_t t = start; t < end; t++) part += data[c * nsamples + t];
            sum += part;
        }
        bm[b] = sum / (float)((end - start) * (active ? active : 1));
    }
    outliers(bm, blocks_n, threshold, bad_blocks);
    size_t count = nbad * nsamples;
    for (size_t c = 0; c < nch; c++) {
        float *row = data + c * nsamples;
        float typical = oracle_median(row, nsamples);
        if (bad[c]) {
            for (size_t t = 0; t < nsamples; t++) row[t] = typical;
        } else {
            for (size_t b = 0; b < blocks_n; b++) {
                if (!bad_blocks[b]) continue;
                size_t start = b * block_samples;
                size_t end = start + block_samples < nsamples ? start + block_samples : nsamples;
                for (size_t t = start; t < end; t++) row[t] = typical;
                count += end - start;
            }
        }
    }
    free(means);
    free(vars);
    free(mb);
    free(vb);
    free(bm);
    return (double)count / (double)(nsamples * nch);
}

/* ------------------------------------------------------------------ FFT */

/* Iterative radix-2 complex DFT in f64 (forward, unnormalised). */
static void cfft64(double *re, double *im, size_t M) {
    for (size_t i = 1, j = 0; i < M; i++) {
        size_t bit = M >> 1;
        for (; j & bit; bit >>= 1) j ^= bit;
        j ^= bit;
        if (i < j) {
            double t = re[i];
            re[i] = re[j];
            re[j] = t;
            t = im[i];
            im[i] = im[j];
            im[j] = t;
        }
    }
    double *wr = (double *)malloc((M / 2 + 1) * sizeof(double));
    double *wi = (double *)malloc((M / 2 + 1) * sizeof(double));
    for (size_t k = 0; k < M / 2; k++) {
        double a = -2.0 * M_PI * (double)k / (double)M;
        wr[k] = cos(a);
        wi[k] = sin(a);
    }
    for (size_t len = 2; len <= M; len <<= 1) {
        size_t half = len >> 1, stride = M / len;
        for (size_t s = 0; s < M; s += len)
            for (size_t k = 0; k < half; k++) {
                double c = wr[k * stride], d = wi[k * stride];
                size_t a = s + k, b = a + half;
                double xr = re[b] * c - im[b] * d;
                double xi = re[b] * d + im[b] * c;
                re[b] = re[a] - xr;
                im[b] = im[a] - xi;
                re[a] += xr;
                im[a] += xi;
            }
    }
    free(wr);
    free(wi);
}

void oracle_rfft64(const double *x, size_t N, double *re, double *im) {
    size_t M = N / 2;
    double *zr = (double *)malloc(M * sizeof(double));
    double *zi = (double *)malloc(M * sizeof(double));
    for (size_t m = 0; m < M; m++) {
        zr[m] = x[2 * m];
        zi[m] = x[2 * m + 1];
    }
    cfft64(zr, zi, M);
    for (size_t k = 0; k <= M; k++) {
        size_t a = k % M, b = (M - k) % M;
        double er = 0.5 * (zr[a] + zr[b]), ei = 0.5 * (zi[a] - zi[b]);
        double orr = 0.5 * (zi[a] + zi[b]), oi = -0.5 * (zr[a] - zr[b]);
        double ang = -2.0 * M_PI * (double)k / (double)N;
        double c = cos(ang), s = sin(ang);
        re[k] = er + (orr * c - oi * s);
        im[k] = ei + (orr * s + oi * c);
    }
    free(zr);
    free(zi);
}

static size_t next_pow2(size_t n) {
    size_t p = 1;
    while (p < n) p <<= 1;
    return p;
}

/* crates/search/src/lib.rs:35-72 (f32 stages; FFT modelled exactly rounded). */
size_t oracle_spectrum(const float *series, size_t n, double tsamp, const oracle_config *cfg,
                       float *raw, float *power) {
    size_t size = next_pow2(n) * 2;
    double mean = 0.0;
    for (size_t i = 0; i < n; i++) mean += (double)series[i];
    mean /= (double)n;
    float mean32 = (float)mean;
    double *x = (double *)calloc(size, sizeof(double));
    for (size_t i = 0; i < n; i++) x[i] = (double)(series[i] - mean32);
    size_t nb = size / 2 + 1;
    double *re = (double *)malloc(nb * sizeof(double)), *im = (double *)malloc(nb * sizeof(double));
    oracle_rfft64(x, size, re, im);
    for (size_t k = 0; k < nb; k++) {
        float r = (float)re[k], i = (float)im[k];
        float rr = r * r, ii = i * i;
        raw[k] = rr + ii;
        power[k] = raw[k];
    }
    free(x);
    free(re);
    free(im);
    double total = 0.0;
    for (size_t k = 0; k < nb; k++) total += (double)raw[k];
    float floor32 = (float)(total / (double)nb * 1e-6);
    const float LN2 = 0.693147180559945309417232121458176568f;
    for (size_t block = 0; block * 64 < nb; block++) {
        size_t center = block * 64;
        size_t lo = center >= 128 ? center - 128 : 0;
        if (lo < 1) lo = 1;
        size_t hi = center + 192 < nb ? center + 192 : nb;
        float med = oracle_median(raw + lo, hi > lo ? hi - lo : 0);
        float scale = fmaxf(fmaxf(med / LN2, floor32), 1e-20f);
        size_t end = center + 64 < nb ? center + 64 : nb;
        for (size_t k = center; k < end; k++) power[k] /= scale;
    }
    power[0] = 0.0f;
    double bin_hz = 1.0 / ((double)size * tsamp);
    for (size_t i = 0; i < nb; i++) {
        double f = (double)i * bin_hz;
        for (uint64_t b = 0; b < cfg->nbirdies; b++)
            if (fabs(f - cfg->birdies[2 * b]) <= cfg->birdies[2 * b + 1]) {
                power[i] = 0.0f;
                break;
            }
    }
    return size;
}

/* Candidates are collected then stably sorted by descending score (lib.rs:130). */
typedef struct {
    oracle_candidate c;
    size_t order;
} ranked;
static int cmp_ranked(const void *pa, const void *pb) {
    const ranked *a = (const ranked *)pa, *b = (const ranked *)pb;
    int64_t ka = total_key64(a->c.score), kb = total_key64(b->c.score);
    if (ka != kb) return ka > kb ? -1 : 1; /* descending */
    return a->order < b->order ? -1 : (a->order > b->order);
}

static inline double interp_offset(double a, double b, double c) {
    double curvature = a - 2.0 * b + c;
    if (curvature < 0.0) {
        double o = 0.5 * (a - c) / curvature;
        return o < -0.5 ? -0.5 : (o > 0.5 ? 0.5 : o);
    }
    return 0.0;
}

/* crates/search/src/lib.rs:73-145 over given raw/whitened arrays (f32 or f64 via stride). */
static size_t harmonic_search(const double *raw, const double *power, size_t nb, size_t size,
                              size_t n, double tsamp, double dm, const oracle_config *cfg,
                              oracle_candidate *out, size_t cap) {
    double bin_hz = 1.0 / ((double)size * tsamp);
    double fq = ceil(cfg->min_hz / bin_hz);
    size_t first = (size_t)(fq > 1.0 ? fq : 1.0);
    double top = cfg->max_hz < 0.5 / tsamp ? cfg->max_hz : 0.5 / tsamp;
    size_t last = (size_t)floor(top / bin_hz);
    if (last > nb - 2) last = nb - 2;
    size_t ncap = 1024, nc = 0;
    ranked *cands = (ranked *)malloc(ncap * sizeof(ranked));
    static const size_t HS[5] = {1, 2, 4, 8, 16};
    double *scores = NULL;
    for (int hi = 0; hi < 5; hi++) {
        size_t h = HS[hi];
        if (h > cfg->max_harmonics) break;
        size_t end = (nb - 2) / h;
        if (last < end) end = last;
        if (end <= first + 1) continue;
        size_t len = end - first + 1;
        scores = (double *)realloc(scores, len * sizeof(double));
        for (size_t k = first; k <= end; k++) {
            double s;
            if (h > 1 && power[k] < 4.0) {
                s = -INFINITY;
            } else {
                double offset = interp_offset(raw[k - 1], raw[k], raw[k + 1]);
                double sum = 0.0;
                for (size_t j = 1; j <= h; j++) {
                    double position = ((double)k + offset) * (double)j;
                    size_t index = (size_t)round(position);
                    if (index > nb - 1) index = nb - 1;
                    sum += power[index];
                }
                s = (sum - (double)h) / sqrt((double)h);
            }
            scores[k - first] = s;
        }
        for (size_t i = 1; i + 1 < len; i++) {
            if (scores[i] >= cfg->threshold && scores[i] > scores[i - 1] &&
                scores[i] >= scores[i + 1]) {
                size_t k = first + i;
                double offset = interp_offset(raw[k - 1], raw[k], raw[k + 1]);
                double frequency = ((double)k + offset) * bin_hz;
                if (nc == ncap) {
                    ncap *= 2;
                    cands = (ranked *)realloc(cands, ncap * sizeof(ranked));
                }
                cands[nc].c.frequency_hz = frequency;
                cands[nc].c.period_s = 1.0 / frequency;
                cands[nc].c.dm = dm;
                cands[nc].c.harmonics = h;
                cands[nc].c.score = scores[i];
                cands[nc].order = nc;
                nc++;
            }
        }
    }
    free(scores);
    qsort(cands, nc, sizeof(ranked), cmp_ranked);
    size_t kept = 0;
    double res = 1.0 / ((double)n * tsamp);
    for (size_t i = 0; i < nc && kept < cap; i++) {
        int dup = 0;
        for (size_t j = 0; j < kept; j++)
            if (fabs(out[j].frequency_hz - cands[i].c.frequency_hz) < res) {
                dup = 1;
                break;
            }
        if (dup) continue;
        out[kept++] = cands[i].c;
        if (kept >= cfg->max_candidates) break; /* lib.rs:140-142: push, then test */
    }
    free(cands);
    return kept;
}

/* crates/search/src/lib.rs:33-145 */
size_t oracle_periodicity(const float *series, size_t n, double tsamp, double dm,
                          const oracle_config *cfg, oracle_candidate *out, size_t cap) {
    if (n < 2 || !isfinite(tsamp) || tsamp <= 0.0) return 0;
    size_t size = next_pow2(n) * 2, nb = size / 2 + 1;
    float *raw = (float *)malloc(nb * sizeof(float)), *pw = (float *)malloc(nb * sizeof(float));
    oracle_spectrum(series, n, tsamp, cfg, raw, pw);
    double *rawd = (double *)malloc(nb * sizeof(double)), *pwd = (double *)malloc(nb * sizeof(double));
    for (size_t k = 0; k < nb; k++) {
        rawd[k] = (double)raw[k];
        pwd[k] = (double)pw[k];
    }
    free(raw);
    free(pw);
    size_t r = harmonic_search(rawd, pwd, nb, size, n, tsamp, dm, cfg, out, cap);
    free(rawd);
    free(pwd);
    return r;
}

/* All-f64 twin of oracle_spectrum: crates/search/src/lib.rs:35-72 with every
 * f32 step in double (the FP64 oracle's spectrum stages). */
size_t oracle_spectrum_f64(const float *series, size_t n, double tsamp, const oracle_config *cfg,
                           double *raw, double *pw) {
    size_t size = next_pow2(n) * 2, nb = size / 2 + 1;
    double mean = 0.0;
    for (size_t i = 0; i < n; i++) mean += (double)series[i];
    mean /= (double)n;
    double *x = (double *)calloc(size, sizeof(double));
    for (size_t i = 0; i < n; i++) x[i] = (double)series[i] - mean;
    double *re = (double *)malloc(nb * sizeof(double)), *im = (double *)malloc(nb * sizeof(double));
    oracle_rfft64(x, size, re, im);
    free(x);
    double total = 0.0;
    for (size_t k = 0; k < nb; k++) {
        raw[k] = re[k] * re[k] + im[k] * im[k];
        pw[k] = raw[k];
        total += raw[k];
    }
    free(re);
    free(im);
    double floor64 = total / (double)nb * 1e-6;
    for (size_t block = 0; block * 64 < nb; block++) {
        size_t center = block * 64;
        size_t lo = center >= 128 ? center - 128 : 0;
        if (lo < 1) lo = 1;
        size_t hi = center + 192 < nb ? center + 192 : nb;
        double med = oracle_median64(raw + lo, hi > lo ? hi - lo : 0);
        double scale = fmax(fmax(med / M_LN2, floor64), 1e-20);
        size_t end = center + 64 < nb ? center + 64 : nb;
        for (size_t k = center; k < end; k++) pw[k] /= scale;
    }
    pw[0] = 0.0;
    double bin_hz = 1.0 / ((double)size * tsamp);
    for (size_t i = 0; i < nb; i++) {
        double f = (double)i * bin_hz;
        for (uint64_t b = 0; b < cfg->nbirdies; b++)
            if (fabs(f - cfg->birdies[2 * b]) <= cfg->birdies[2 * b + 1]) {
                pw[i] = 0.0;
                break;
            }
    }
    return size;
}

/* All-f64 twin: same algorithm, double everywhere (the FP64 oracle). */
size_t oracle_periodicity_f64(const float *series, size_t n, double tsamp, double dm,
                              const oracle_config *cfg, oracle_candidate *out, size_t cap) {
    if (n < 2 || !isfinite(tsamp) || tsamp <= 0.0) return 0;
    size_t size = next_pow2(n) * 2, nb = size / 2 + 1;
    double *raw = (double *)malloc(nb * sizeof(double)), *pw = (double *)malloc(nb * sizeof(double));
    oracle_spectrum_f64(series, n, tsamp, cfg, raw, pw);
    size_t r = harmonic_search(raw, pw, nb, size, n, tsamp, dm, cfg, out, cap);
    free(raw);
    free(pw);
    return r;
}

/* crates/search/src/lib.rs:147-176 */
size_t oracle_sift(const oracle_candidate *in, size_t n, double resolution_hz, size_t limit,
                   int reject_dm_zero, oracle_candidate *out) {
    ranked *s = (ranked *)malloc((n ? n : 1) * sizeof(ranked));
    for (size_t i = 0; i < n; i++) {
        s[i].c = in[i];
        s[i].order = i;
    }
    qsort(s, n, sizeof(ranked), cmp_ranked);
    double *rejected = (double *)malloc((n ? n : 1) * sizeof(double));
    size_t kept = 0, nrej = 0;
    for (size_t i = 0; i < n; i++) {
        double f = s[i].c.frequency_hz;
        int related = 0;
        for (size_t j = 0; j < kept + nrej && !related; j++) {
            double g = j < kept ? out[j].frequency_hz : rejected[j - kept];
            for (int h = 1; h <= 16; h++)
                if (fabs(f - g * (double)h) <= resolution_hz ||
                    fabs(g - f * (double)h) <= resolution_hz) {
                    related = 1;
                    break;
                }
        }
        if (related) continue;
        if (reject_dm_zero && s[i].c.dm <= 0.0) {
            rejected[nrej++] = f;
            continue;
        }
        out[kept++] = s[i].c;
        if (kept >= limit) break; /* lib.rs:171-173: push, then test */
    }
    free(s);
    free(rejected);
    return kept;
}

/* ----------------------------------------------------------------- fold */

static int cmp_f32_total(const void *a, const void *b) {
    int32_t x = total_key32(*(const float *)a), y = total_key32(*(const float *)b);
    return (x > y) - (x < y);
}

/* crates/fold/src/lib.rs:16-80 */
void oracle_profile_at(const float *series, size_t n, double tsamp, double period, size_t bins,
                       size_t sample_offset, double *snr_out, double *chi2_out, float *values,
                       uint64_t *counts) {
    double *sums = (double *)calloc(bins, sizeof(double));
    for (size_t b = 0; b < bins; b++) counts[b] = 0;
    for (size_t t = 0; t < n; t++) {
        double x = (double)(t + sample_offset) * tsamp / period;
        double phase = x - trunc(x);
        size_t b = (size_t)(phase * (double)bins);
        if (b > bins - 1) b = bins - 1;
        sums[b] += (double)series[t];
        counts[b] += 1;
    }
    for (size_t b = 0; b < bins; b++)
        values[b] = counts[b] > 0 ? (float)(sums[b] / (double)counts[b]) : 0.0f;
    float center = oracle_median(values, bins);
    float *off = (float *)malloc(bins * sizeof(float));
    size_t noff = 0;
    for (size_t b = 0; b < bins; b++)
        if (counts[b] > 0) off[noff++] = values[b];
    qsort(off, noff, sizeof(float), cmp_f32_total);
    size_t keep = noff * 3 / 4;
    if (keep < 1) keep = 1;
    double mean = 0.0;
    for (size_t i = 0; i < keep; i++) mean += (double)off[i];
    mean /= (double)keep;
    double ss = 0.0;
    for (size_t i = 0; i < keep; i++) {
        double d = (double)off[i] - mean;
        ss += d * d;
    }
    double sigma = sqrt(ss / (double)keep);
    float peak = -INFINITY;
    for (size_t b = 0; b < bins; b++) peak = fmaxf(peak, values[b]);
    *snr_out = sigma > 1e-12 ? ((double)peak - (double)center) / sigma : 0.0;
    double smean = 0.0;
    for (size_t t = 0; t < n; t++) smean += (double)series[t];
    smean /= (double)n;
    double svar = 0.0;
    for (size_t t = 0; t < n; t++) {
        double d = (double)series[t] - smean;
        svar += d * d;
    }
    svar /= (double)n;
    size_t populated = 0;
    for (size_t b = 0; b < bins; b++) populated += counts[b] > 0;
    double chi2 = 0.0;
    if (svar > 1e-12 && populated > 1) {
        for (size_t b = 0; b < bins; b++) {
            double d = (double)values[b] - smean;
            chi2 += (double)counts[b] * (d * d) / svar;
        }
        chi2 /= (double)(populated - 1);
    }
    *chi2_out = chi2;
    free(sums);
    free(off);
}

/* crates/fold/src/lib.rs:81-97: last maximum of reduced chi-square wins. */
double oracle_refine(const float *series, size_t n, double tsamp, double frequency, double span_hz,
                     size_t steps, size_t bins, double *snr, double *chi2, float *values,
                     uint64_t *counts) {
    float *v = (float *)malloc(bins * sizeof(float));
    uint64_t *c = (uint64_t *)malloc(bins * sizeof(uint64_t));
    double best_chi = 0.0, best_period = 0.0;
    int have = 0;
    for (size_t i = 0; i <= steps; i++) {
        double f = frequency - span_hz + 2.0 * span_hz * (double)i / (double)steps;
        double period = 1.0 / f;
        double s, x;
        oracle_profile_at(series, n, tsamp, period, bins, 0, &s, &x, v, c);
        if (!have || total_key64(x) >= total_key64(best_chi)) {
            have = 1;
            best_chi = x;
            best_period = period;
            *snr = s;
            *chi2 = x;
            memcpy(values, v, bins * sizeof(float));
            memcpy(counts, c, bins * sizeof(uint64_t));
        }
    }
    free(v);
    free(c);
    return best_period;
}

/* --------------------------------------------------- per-trial CPU loop */

typedef struct {
    const float *data;
    size_t nsamples, nch, length, cap;
    const double *freqs, *dms;
    double tsamp;
    const oracle_config *cfg;
    oracle_candidate *out;
    uint32_t *counts;
    float *series_out; /* optional [ntrials][length] */
    int f64;
    size_t ntrials;
    size_t next;
    int err;
    pthread_mutex_t lock;
} trial_job;

static void *trial_worker(void *arg) {
    trial_job *j = (trial_job *)arg;
    uint64_t *d = (uint64_t *)malloc(j->nch * sizeof(uint64_t));
    float *x = (float *)malloc(j->length * sizeof(float));
    for (;;) {
        pthread_mutex_lock(&j->lock);
        size_t t = j->next++;
        pthread_mutex_unlock(&j->lock);
        if (t >= j->ntrials) break;
        size_t got = 0;
        if (oracle_delays(j->freqs, j->nch, j->tsamp, j->dms[t], d) ||
            oracle_apply(j->data, j->nsamples, j->nch, d, j->length, x, &got)) {
            j->err = 1;
            continue;
        }
        if (j->series_out) memcpy(j->series_out + t * j->length, x, got * sizeof(float));
        j->counts[t] = (uint32_t)(j->f64 ? oracle_periodicity_f64 : oracle_periodicity)(
            x, got, j->tsamp, j->dms[t], j->cfg, j->out + t * j->cap, j->cap);
    }
    free(d);
    free(x);
    return NULL;
}

/* crates/cli/src/main.rs:462-469 (rayon par_iter restated on POSIX threads) */
int oracle_search_trials(const float *data, size_t nsamples, const double *freqs, size_t nch,
                         double tsamp, const double *dms, size_t ntrials, size_t length,
                         const oracle_config *cfg, oracle_candidate *out, size_t cap,
                         uint32_t *counts, int threads) {
    return oracle_search_trials_ex(data, nsamples, freqs, nch, tsamp, dms, ntrials, length, cfg, out,
                                   cap, counts, NULL, 0, threads);
}

int oracle_search_trials_ex(const float *data, size_t nsamples, const double *freqs, size_t nch,
                            double tsamp, const double *dms, size_t ntrials, size_t length,
                            const oracle_config *cfg, oracle_candidate *out, size_t cap,
                            uint32_t *counts, float *series_out, int f64, int threads) {
    trial_job j;
    j.series_out = series_out;
    j.f64 = f64;
    j.data = data;
    j.nsamples = nsamples;
    j.nch = nch;
    j.length = length;
    j.cap = cap;
    j.freqs = freqs;
    j.dms = dms;
    j.tsamp = tsamp;
    j.cfg = cfg;
    j.out = out;
    j.counts = counts;
    j.ntrials = ntrials;
    j.next = 0;
    j.err = 0;
    pthread_mutex_init(&j.lock, NULL);
    if (threads < 1) threads = 1;
    pthread_t *th = (pthread_t *)malloc((size_t)threads * sizeof(pthread_t));
    for (int i = 0; i < threads; i++) pthread_create(&th[i], NULL, trial_worker, &j);
    for (int i = 0; i < threads; i++) pthread_join(th[i], NULL);
    free(th);
    pthread_mutex_destroy(&j.lock);
    return j.err ? -1 : 0;
}
