// rfi::mask (crates/rfi/src/lib.rs:31-96) on a device-resident observation:
// channel statistics, block means, channel medians and the fill run on the GPU
// (csrc/k_rfi.hip); the robust outlier tests on the few thousand statistics run
// on the host in f32, as the reference computes them.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "ctx_internal.hpp"

using namespace psr;

namespace {

int32_t key32(float f) {
    int32_t b;
    std::memcpy(&b, &f, 4);
    b ^= (int32_t)(((uint32_t)(b >> 31)) >> 1);
    return b;
}

// rfi::median (lib.rs:4-12): element len/2 in f32::total_cmp order
float median(std::vector<float> v) {
    if (v.empty()) return 0.0f;
    const size_t i = v.size() / 2;
    std::nth_element(v.begin(), v.begin() + i, v.end(),
                     [](float a, float b) { return key32(a) < key32(b); });
    return v[i];
}

// lib.rs:13-22
std::vector<uint8_t> outliers(const std::vector<float>& v, float threshold) {
    const float center = median(v);
    std::vector<float> dev(v.size());
    for (size_t i = 0; i < v.size(); i++) dev[i] = std::fabs(v[i] - center);
    const float mad = median(dev);
    const float scale = std::fmax(std::fmax(1.4826f * mad, std::fabs(center) * 1e-4f), 1e-6f);
    std::vector<uint8_t> out(v.size());
    for (size_t i = 0; i < v.size(); i++) out[i] = std::fabs(v[i] - center) > threshold * scale;
    return out;
}

struct Scratch {
    psr_ctx* c;
    std::vector<void*> ptrs;
    ~Scratch() {
        for (void* p : ptrs) (void)hipFree(p);
    }
    template <typename T>
    int alloc(T** p, size_t n) {
        hipError_t e = hipMalloc((void**)p, (n ? n : 1) * sizeof(T));
        if (e != hipSuccess) return fail(PSR_E_OOM, std::string("hipMalloc failed: ") + hipGetErrorString(e));
        ptrs.push_back(*p);
        return PSR_OK;
    }
};

}  // namespace

extern "C" {

int psr_obs_mask(psr_ctx* c, psr_obs* o, uint64_t block_samples, float threshold, psr_mask* out,
                 uint64_t* channels, uint64_t* blocks) {
    clear_error();
    if (!c || !o || !out || !channels || !blocks) return fail(PSR_E_INVALID, "null pointer");
    if (o->nsamples == 0) return fail(PSR_E_CONTRACT, "mask requires nsamples > 0");  // lib.rs:32
    if (block_samples == 0 || !std::isfinite(threshold) || !(threshold > 0.0f))   // lib.rs:33
        return fail(PSR_E_CONTRACT, "mask requires block_samples > 0 and a finite positive threshold");
    int rc;
    if ((rc = check_device(c))) return rc;
    const uint32_t nch = o->nchans;
    const uint64_t n = o->nsamples;
    const uint64_t nblk64 = (n + block_samples - 1) / block_samples;
    if (block_samples > 0xffffffffull || nblk64 > 0xffffffffull)
        return fail(PSR_E_SHAPE, "block size out of range");
    const uint32_t block = (uint32_t)block_samples, nblocks = (uint32_t)nblk64;
    hipStream_t s = c->stream;
    Scratch sc{c, {}};
    float *d_means, *d_vars, *d_sums, *d_bmeans;
    uint8_t *d_bad, *d_badb;
    uint32_t* d_med;
    if ((rc = sc.alloc(&d_means, nch)) || (rc = sc.alloc(&d_vars, nch)) ||
        (rc = sc.alloc(&d_sums, (size_t)nch * nblocks)) || (rc = sc.alloc(&d_bmeans, nblocks)) ||
        (rc = sc.alloc(&d_bad, nch)) || (rc = sc.alloc(&d_badb, nblocks)) || (rc = sc.alloc(&d_med, nch)))
        return rc;
    // channel means and variances (lib.rs:35-44)
    launch_rfi_row_stats(s, o->data, o->kind, o->pitch, nch, n, d_means, d_vars);
    launch_rfi_block_sums(s, o->data, o->kind, o->pitch, nch, n, block, nblocks, d_sums);
    launch_rfi_median(s, o->data, o->kind, o->pitch, nch, n, d_med);
    HIP_TRY(hipGetLastError());
    std::vector<float> means(nch), vars(nch);
    HIP_TRY(hipMemcpyAsync(means.data(), d_means, nch * sizeof(float), hipMemcpyDeviceToHost, s));
    HIP_TRY(hipMemcpyAsync(vars.data(), d_vars, nch * sizeof(float), hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    // bad channels (lib.rs:45-52)
    const std::vector<uint8_t> mb = outliers(means, threshold), vb = outliers(vars, threshold);
    std::vector<uint8_t> bad(nch);
    uint64_t nbad = 0;
    for (uint32_t ch = 0; ch < nch; ch++) {
        bad[ch] = mb[ch] || vb[ch] || vars[ch] < 1e-12f;
        if (bad[ch]) channels[nbad++] = ch;
    }
    // block means over the good channels (lib.rs:53-66) and bad blocks (:67-71)
    HIP_TRY(hipMemcpyAsync(d_bad, bad.data(), nch, hipMemcpyHostToDevice, s));
    launch_rfi_block_means(s, d_sums, d_bad, nch, n, block, nblocks, (uint32_t)(nch - nbad), d_bmeans);
    HIP_TRY(hipGetLastError());
    std::vector<float> bmeans(nblocks);
    HIP_TRY(hipMemcpyAsync(bmeans.data(), d_bmeans, nblocks * sizeof(float), hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    const std::vector<uint8_t> badb = outliers(bmeans, threshold);
    uint64_t nbb = 0, count = nbad * n;
    for (uint32_t b = 0; b < nblocks; b++)
        if (badb[b]) {
            blocks[nbb++] = b;
            const uint64_t start = (uint64_t)b * block, end = std::min<uint64_t>(start + block, n);
            count += (end - start) * (nch - nbad);
        }
    // fill with each channel's median (lib.rs:72-86)
    HIP_TRY(hipMemcpyAsync(d_badb, badb.data(), nblocks, hipMemcpyHostToDevice, s));
    launch_rfi_fill(s, o->data, o->kind, o->pitch, nch, n, block, d_bad, d_badb, d_med);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(s));
    if ((rc = obs_storage_scale(c, o))) return rc;
    out->nchannels = nbad;
    out->nblocks = nbb;
    out->block_samples = block_samples;
    out->masked_fraction = (double)count / ((double)nch * (double)n);  // lib.rs:87-92
    return PSR_OK;
}

}  // extern "C"
