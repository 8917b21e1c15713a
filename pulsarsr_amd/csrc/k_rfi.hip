// RFI mask kernels for gfx950 (crates/rfi/src/lib.rs:31-96), applied in place
// to a device-resident observation (channel-major rows of u8, u16 or f32).
//
// Every sum the reference makes is an f32 sum in sample order, whose rounding
// depends on that order, so each is made by one thread, sequentially:
//   - per-channel mean and variance (lib.rs:35-44): one thread per channel row;
//   - per-block sums of each channel (lib.rs:57-61): one thread per (channel, block),
//     then one thread per block adds the good channels' sums in channel order.
// The channel median (lib.rs:76) is an order statistic, found exactly by radix
// selection over order-preserving integer keys (one workgroup per channel).
#include <hip/hip_runtime.h>

#include "common.hpp"

namespace psr {

__device__ __forceinline__ float sample_f32(const void* data, int kind, size_t idx) {
    if (kind == KIND_U8) return (float)((const uint8_t*)data)[idx];
    if (kind == KIND_U16) return (float)((const uint16_t*)data)[idx];
    return ((const float*)data)[idx];
}

// Order-preserving key of f32::total_cmp (u8/u16 samples are their own keys).
__device__ __forceinline__ uint32_t sample_key(const void* data, int kind, size_t idx) {
    if (kind == KIND_U8) return ((const uint8_t*)data)[idx];
    if (kind == KIND_U16) return ((const uint16_t*)data)[idx];
    const uint32_t b = __float_as_uint(((const float*)data)[idx]);
    return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
}

// mean = (sum of the row in order) / n and variance = (sum of (v - mean)^2 in
// order) / n, all in f32 (lib.rs:35-44).
__global__ __launch_bounds__(64) void k_rfi_row_stats(const void* __restrict__ data, int kind,
                                                      uint64_t pitch, uint32_t nch, uint64_t n,
                                                      float* __restrict__ means,
                                                      float* __restrict__ vars) {
    const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= nch) return;
    const size_t row = (size_t)c * pitch;
    float s = 0.0f;
    for (uint64_t t = 0; t < n; t++) s += sample_f32(data, kind, row + t);
    const float m = s / (float)n;
    float q = 0.0f;
    for (uint64_t t = 0; t < n; t++) {
        const float d = sample_f32(data, kind, row + t) - m;
        q += d * d;
    }
    means[c] = m;
    vars[c] = q / (float)n;
}

// sums[c][b] = sum of channel c over block b, in sample order (lib.rs:57-61)
__global__ __launch_bounds__(256) void k_rfi_block_sums(const void* __restrict__ data, int kind,
                                                        uint64_t pitch, uint32_t nch, uint64_t n,
                                                        uint32_t block, uint32_t nblocks,
                                                        float* __restrict__ sums) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (uint64_t)nch * nblocks) return;
    const uint32_t c = (uint32_t)(i / nblocks), b = (uint32_t)(i % nblocks);
    const uint64_t start = (uint64_t)b * block;
    const uint64_t end = start + block < n ? start + block : n;
    const size_t row = (size_t)c * pitch;
    float s = 0.0f;
    for (uint64_t t = start; t < end; t++) s += sample_f32(data, kind, row + t);
    sums[i] = s;
}

// block mean = (good channels' block sums added in channel order) /
// ((end - start) * max(active, 1)) as f32 (lib.rs:53-64)
__global__ __launch_bounds__(256) void k_rfi_block_means(const float* __restrict__ sums,
                                                         const uint8_t* __restrict__ bad,
                                                         uint32_t nch, uint64_t n, uint32_t block,
                                                         uint32_t nblocks, uint32_t active,
                                                         float* __restrict__ means) {
    const uint32_t b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= nblocks) return;
    const uint64_t start = (uint64_t)b * block;
    const uint64_t end = start + block < n ? start + block : n;
    float s = 0.0f;
    for (uint32_t c = 0; c < nch; c++)
        if (!bad[c]) s += sums[(size_t)c * nblocks + b];
    means[b] = s / (float)((end - start) * (uint64_t)(active > 0 ? active : 1));
}

// Upper median (element n/2 of the sorted row, lib.rs:4-12) by radix selection
// over the row's keys, 8 bits per pass (one pass for u8, two for u16, four for
// f32).  One workgroup per channel; the result is the median's key.
__global__ __launch_bounds__(256) void k_rfi_median(const void* __restrict__ data, int kind,
                                                    uint64_t pitch, uint64_t n,
                                                    uint32_t* __restrict__ median_key) {
    __shared__ uint32_t hist[256];
    __shared__ uint32_t s_prefix;
    __shared__ uint64_t s_rank;
    const uint32_t c = blockIdx.x;
    const size_t row = (size_t)c * pitch;
    const int bits = kind == KIND_U8 ? 8 : (kind == KIND_U16 ? 16 : 32);
    if (threadIdx.x == 0) {
        s_prefix = 0;
        s_rank = n / 2;
    }
    for (int shift = bits - 8; shift >= 0; shift -= 8) {
        for (uint32_t d = threadIdx.x; d < 256; d += blockDim.x) hist[d] = 0;
        __syncthreads();
        const uint32_t prefix = s_prefix;
        const int hi = shift + 8;  // bits above the current digit must match the prefix
        for (uint64_t t = threadIdx.x; t < n; t += blockDim.x) {
            const uint32_t k = sample_key(data, kind, row + t);
            if (hi >= 32 || ((k ^ prefix) >> hi) == 0) atomicAdd(&hist[(k >> shift) & 255u], 1u);
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            uint64_t cum = 0, rank = s_rank;
            for (uint32_t d = 0; d < 256; d++) {
                if (cum + hist[d] > rank) {
                    s_prefix = prefix | (d << shift);
                    s_rank = rank - cum;
                    break;
                }
                cum += hist[d];
            }
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) median_key[c] = s_prefix;
}

// Replace the samples of bad channels, and of bad blocks in the other channels,
// with the channel's median (lib.rs:73-86).
__global__ __launch_bounds__(256) void k_rfi_fill(void* __restrict__ data, int kind, uint64_t pitch,
                                                  uint64_t n, uint32_t block,
                                                  const uint8_t* __restrict__ bad,
                                                  const uint8_t* __restrict__ bad_blocks,
                                                  const uint32_t* __restrict__ median_key) {
    const uint32_t c = blockIdx.y;
    const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    if (!bad[c] && !bad_blocks[t / block]) return;
    const uint32_t k = median_key[c];
    const size_t idx = (size_t)c * pitch + t;
    if (kind == KIND_U8)
        ((uint8_t*)data)[idx] = (uint8_t)k;
    else if (kind == KIND_U16)
        ((uint16_t*)data)[idx] = (uint16_t)k;
    else
        ((uint32_t*)data)[idx] = (k & 0x80000000u) ? (k & 0x7fffffffu) : ~k;
}

void launch_rfi_row_stats(hipStream_t s, const void* data, int kind, uint64_t pitch, uint32_t nch,
                          uint64_t n, float* means, float* vars) {
    hipLaunchKernelGGL(k_rfi_row_stats, dim3((nch + 63) / 64), dim3(64), 0, s, data, kind, pitch, nch,
                       n, means, vars);
}

void launch_rfi_block_sums(hipStream_t s, const void* data, int kind, uint64_t pitch, uint32_t nch,
                           uint64_t n, uint32_t block, uint32_t nblocks, float* sums) {
    const uint64_t total = (uint64_t)nch * nblocks;
    hipLaunchKernelGGL(k_rfi_block_sums, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s, data,
                       kind, pitch, nch, n, block, nblocks, sums);
}

void launch_rfi_block_means(hipStream_t s, const float* sums, const uint8_t* bad, uint32_t nch,
                            uint64_t n, uint32_t block, uint32_t nblocks, uint32_t active,
                            float* means) {
    hipLaunchKernelGGL(k_rfi_block_means, dim3((nblocks + 255) / 256), dim3(256), 0, s, sums, bad, nch,
                       n, block, nblocks, active, means);
}

void launch_rfi_median(hipStream_t s, const void* data, int kind, uint64_t pitch, uint32_t nch,
                       uint64_t n, uint32_t* median_key) {
    hipLaunchKernelGGL(k_rfi_median, dim3(nch), dim3(256), 0, s, data, kind, pitch, n, median_key);
}

void launch_rfi_fill(hipStream_t s, void* data, int kind, uint64_t pitch, uint32_t nch, uint64_t n,
                     uint32_t block, const uint8_t* bad, const uint8_t* bad_blocks,
                     const uint32_t* median_key) {
    hipLaunchKernelGGL(k_rfi_fill, dim3((unsigned)((n + 255) / 256), nch), dim3(256), 0, s, data, kind,
                       pitch, n, block, bad, bad_blocks, median_key);
}

}  // namespace psr
