// Batched real FFT and power spectrum for gfx950 (crates/search/src/lib.rs:35-44):
//   size = 2*next_pow2(n); input = x - (mean as f32), zero-padded; X = rfft(input);
//   power = re^2 + im^2.
// A length-N real transform is computed as a length-M = N/2 complex transform
// of z[m] = x[2m] + i x[2m+1] followed by the even/odd split.  M <= 4096 runs
// in one LDS-resident pass per trial.  Larger M uses a four-step split
// M = m1*m2 (m = m1idx + m1*m2idx):
//   pass 1: length-m2 FFTs over the strided columns, times W_M^(m1idx*k2),
//           stored transposed as work[k2][m1idx];
//   pass 2: length-m1 FFTs over rows k2 and m2-k2 together, then the r2c
//           split (the partner of k = k2 + m2*k1 is M-k, held in the same
//           workgroup) and the power, written as spec_t[k2*m1 + k1].
// The transposed spectrum is put back in natural order by the whitening kernel.
// Butterflies are radix-4 Stockham (one radix-2 stage for odd log2 lengths) in
// FP32 with FP64-generated twiddles; the upper half of the padded input is
// zero and is never read from memory.
#include <hip/hip_fp16.h>
#include <hip/hip_runtime.h>

#include "common.hpp"

namespace psr {

__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
    return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }

// In-LDS Stockham FFT of V contiguous vectors of length L = 2^logL.
// Each thread owns up to BPT radix-4 butterflies per stage; all of a stage's
// inputs are read into registers before any output is written.
template <int BPT>
__device__ void lds_fft(float2* buf, int L, int logL, int V, const float2* __restrict__ tw) {
    const int nt = blockDim.x;
    const int tid = threadIdx.x;
    int Ns = 1;
    if (logL & 1) {
        const int h = L >> 1;
        float2 a0[2 * BPT], a1[2 * BPT];
#pragma unroll
        for (int i = 0; i < 2 * BPT; i++) {
            const int j = tid + i * nt;
            if (j < V * h) {
                const int v = j / h, jj = j - v * h;
                a0[i] = buf[v * L + jj];
                a1[i] = buf[v * L + jj + h];
            }
        }
        __syncthreads();
#pragma unroll
        for (int i = 0; i < 2 * BPT; i++) {
            const int j = tid + i * nt;
            if (j < V * h) {
                const int v = j / h, jj = j - v * h;
                const int base = v * L + 2 * jj;
                buf[base] = cadd(a0[i], a1[i]);
                buf[base + 1] = csub(a0[i], a1[i]);
            }
        }
        __syncthreads();
        Ns = 2;
    }
    const int q = L >> 2;
    for (; Ns < L; Ns <<= 2) {
        float2 a[BPT][4];
#pragma unroll
        for (int i = 0; i < BPT; i++) {
            const int j = tid + i * nt;
            if (j < V * q) {
                const int v = j / q, jj = j - v * q;
#pragma unroll
                for (int r = 0; r < 4; r++) a[i][r] = buf[v * L + jj + r * q];
            }
        }
        __syncthreads();
        const int tstep = L / (4 * Ns);
#pragma unroll
        for (int i = 0; i < BPT; i++) {
            const int j = tid + i * nt;
            if (j < V * q) {
                const int v = j / q, jj = j - v * q;
                const int k = jj & (Ns - 1);
                float2 x0 = a[i][0], x1 = a[i][1], x2 = a[i][2], x3 = a[i][3];
                if (Ns > 1) {
                    x1 = cmul(x1, tw[k * tstep]);
                    x2 = cmul(x2, tw[2 * k * tstep]);
                    x3 = cmul(x3, tw[3 * k * tstep]);
                }
                const float2 s0 = cadd(x0, x2), d0 = csub(x0, x2);
                const float2 s1 = cadd(x1, x3), d1 = csub(x1, x3);
                const int base = v * L + (jj / Ns) * 4 * Ns + k;
                buf[base] = cadd(s0, s1);
                buf[base + 2 * Ns] = csub(s0, s1);
                // forward: X1 = d0 - i d1, X3 = d0 + i d1
                buf[base + Ns] = make_float2(d0.x + d1.y, d0.y - d1.x);
                buf[base + 3 * Ns] = make_float2(d0.x - d1.y, d0.y + d1.x);
            }
        }
        __syncthreads();
    }
}

struct TwOffsets {
    uint32_t p1, p2, lo, hi, rlo, rhi;
};
static TwOffsets tw_offsets(const SpecGeom& g) {
    TwOffsets o;
    o.p1 = 0;
    o.p2 = g.m2;
    o.lo = g.m2 + g.m1;
    o.hi = o.lo + 2048;
    const uint32_t hi = g.m >= 2048 ? g.m / 2048 : 1;
    o.rlo = o.hi + hi;
    o.rhi = o.rlo + g.m2;
    return o;
}

// Series sample i of a trial as FFT input: x[i] - mean32 in f32 for FP32
// storage, h[i] + corr (already scaled) for FP16 storage, zero beyond n.
__device__ __forceinline__ float2 load_pair(const void* series, int fp16, size_t row,
                                            uint32_t i, uint32_t n, float mean, float corr) {
    float2 z = make_float2(0.f, 0.f);
    if (i >= n) return z;
    if (fp16) {
        const __half2 h = *(const __half2*)((const __half*)series + row + i);
        const float2 f = __half22float2(h);
        z.x = f.x + corr;
        z.y = i + 1 < n ? f.y + corr : 0.f;
    } else {
        const float2 f = *(const float2*)((const float*)series + row + i);
        z.x = f.x - mean;
        z.y = i + 1 < n ? f.y - mean : 0.f;
    }
    return z;
}

__device__ __forceinline__ void store_power(void* spec, int fp16, size_t idx, float p, float inv_n) {
    if (fp16)
        ((__half*)spec)[idx] = __float2half_rn(sqrtf(p * inv_n));
    else
        ((float*)spec)[idx] = p;
}

__device__ double block_sum(double v, double* red) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) red[w] = v;
    __syncthreads();
    double s = 0.0;
    if (threadIdx.x == 0)
        for (int i = 0; i < (int)(blockDim.x >> 6); i++) s += red[i];
    return s;
}

// ---------------------------------------------------------- single pass
__global__ __launch_bounds__(256) void k_fft_single(const void* __restrict__ series, int fp16,
                                                    const float* __restrict__ mean32,
                                                    const float* __restrict__ corr16, uint32_t n,
                                                    uint32_t series_pitch, uint32_t M, int logM,
                                                    void* __restrict__ spec, uint64_t spec_pitch,
                                                    double* __restrict__ partial,
                                                    const float2* __restrict__ tw, TwOffsets o) {
    extern __shared__ float2 buf[];
    __shared__ double red[4];
    const uint32_t trial = blockIdx.x;
    const size_t row = (size_t)trial * series_pitch;
    const float mean = fp16 ? 0.f : mean32[trial];
    const float corr = fp16 ? corr16[trial] : 0.f;
    for (uint32_t m = threadIdx.x; m < M; m += blockDim.x)
        buf[m] = load_pair(series, fp16, row, 2 * m, n, mean, corr);
    __syncthreads();
    lds_fft<4>(buf, (int)M, logM, 1, tw + o.p1);
    const float inv_n = 1.0f / (float)n;
    double acc = 0.0;
    const size_t srow = (size_t)trial * spec_pitch;
    for (uint32_t k = threadIdx.x; k <= M; k += blockDim.x) {
        const float2 zk = buf[k & (M - 1)], zp = buf[(M - k) & (M - 1)];
        const float2 e = make_float2(0.5f * (zk.x + zp.x), 0.5f * (zk.y - zp.y));
        const float2 od = make_float2(0.5f * (zk.y + zp.y), -0.5f * (zk.x - zp.x));
        float2 x;
        if (k == M)
            x = csub(e, od);
        else
            x = cadd(e, cmul(tw[o.rlo + k], od));
        const float xx = x.x * x.x, yy = x.y * x.y;
        const float p = xx + yy;
        store_power(spec, fp16, srow + k, p, inv_n);
        acc += fp16 ? (double)(p * inv_n) : (double)p;
    }
    const double s = block_sum(acc, red);
    if (threadIdx.x == 0) partial[trial] = s;
}

// --------------------------------------------------------------- pass 1
__global__ __launch_bounds__(256) void k_fft_pass1(const void* __restrict__ series, int fp16,
                                                   const float* __restrict__ mean32,
                                                   const float* __restrict__ corr16, uint32_t n,
                                                   uint32_t series_pitch, uint32_t M, uint32_t m1,
                                                   uint32_t m2, int log_m2, uint32_t C,
                                                   float2* __restrict__ work,
                                                   const float2* __restrict__ tw, TwOffsets o) {
    extern __shared__ float2 buf[];  // [C][m2]
    const uint32_t trial = blockIdx.y;
    const uint32_t c0 = blockIdx.x * C;
    const size_t row = (size_t)trial * series_pitch;
    const float mean = fp16 ? 0.f : mean32[trial];
    const float corr = fp16 ? corr16[trial] : 0.f;
    for (uint32_t idx = threadIdx.x; idx < C * m2; idx += blockDim.x) {
        const uint32_t r = idx / C, col = idx - r * C;
        const uint32_t i = 2u * (c0 + col + m1 * r);
        buf[col * m2 + r] = load_pair(series, fp16, row, i, n, mean, corr);
    }
    __syncthreads();
    lds_fft<8>(buf, (int)m2, log_m2, (int)C, tw + o.p1);
    float2* w = work + (size_t)trial * M;
    for (uint32_t idx = threadIdx.x; idx < C * m2; idx += blockDim.x) {
        const uint32_t k2 = idx / C, col = idx - k2 * C;
        const uint32_t a = (uint32_t)(((uint64_t)(c0 + col) * k2) & (M - 1));
        const float2 t = cmul(tw[o.hi + (a >> 11)], tw[o.lo + (a & 2047)]);
        w[(size_t)k2 * m1 + c0 + col] = cmul(buf[col * m2 + k2], t);
    }
}

// --------------------------------------------------------------- pass 2
__global__ __launch_bounds__(256) void k_fft_pass2(const float2* __restrict__ work, int fp16,
                                                   uint32_t n, uint32_t M, uint32_t m1,
                                                   int log_m1, uint32_t m2,
                                                   void* __restrict__ spec, uint64_t spec_pitch,
                                                   double* __restrict__ partial,
                                                   const float2* __restrict__ tw, TwOffsets o) {
    extern __shared__ float2 buf[];  // [2][m1]
    __shared__ double red[4];
    const uint32_t trial = blockIdx.y;
    const uint32_t r1 = blockIdx.x, r2 = (m2 - r1) & (m2 - 1);
    const float2* w = work + (size_t)trial * M;
    for (uint32_t k1 = threadIdx.x; k1 < m1; k1 += blockDim.x) {
        buf[k1] = w[(size_t)r1 * m1 + k1];
        buf[m1 + k1] = w[(size_t)r2 * m1 + k1];
    }
    __syncthreads();
    lds_fft<8>(buf, (int)m1, log_m1, 2, tw + o.p2);
    const float inv_n = 1.0f / (float)n;
    const size_t srow = (size_t)trial * spec_pitch;
    double acc = 0.0;
    const uint32_t nsel = r1 == r2 ? 1u : 2u;
    for (uint32_t idx = threadIdx.x; idx < nsel * m1; idx += blockDim.x) {
        const uint32_t sel = idx / m1, k1 = idx - sel * m1;
        const uint32_t rowk = sel ? r2 : r1;
        const float2 zk = buf[sel * m1 + k1];
        float2 zp;
        if (rowk == 0)
            zp = buf[(m1 - k1) & (m1 - 1)];
        else
            zp = buf[(nsel == 1 ? 0u : (1u - sel)) * m1 + (m1 - 1 - k1)];
        const float2 e = make_float2(0.5f * (zk.x + zp.x), 0.5f * (zk.y - zp.y));
        const float2 od = make_float2(0.5f * (zk.y + zp.y), -0.5f * (zk.x - zp.x));
        const float2 wk = cmul(tw[o.rlo + rowk], tw[o.rhi + k1]);
        const float2 x = cadd(e, cmul(wk, od));
        const float xx = x.x * x.x, yy = x.y * x.y;
        const float p = xx + yy;
        store_power(spec, fp16, srow + (size_t)rowk * m1 + k1, p, inv_n);
        acc += fp16 ? (double)(p * inv_n) : (double)p;
        if (rowk == 0 && k1 == 0) {  // Nyquist bin k = M
            const float2 xm = csub(e, od);
            const float pm = xm.x * xm.x + xm.y * xm.y;
            store_power(spec, fp16, srow + M, pm, inv_n);
            acc += fp16 ? (double)(pm * inv_n) : (double)pm;
        }
    }
    const double s = block_sum(acc, red);
    if (threadIdx.x == 0) partial[(size_t)trial * (m2 / 2 + 1) + r1] = s;
}

static int ilog2(uint32_t v) {
    int l = 0;
    while ((1u << l) < v) l++;
    return l;
}

void launch_fft(hipStream_t s, const SpecGeom& g, const void* series, int fp16,
                const float* mean32, const float* corr16, uint32_t ntrials, float2* work,
                void* spec, double* partial, const float2* tw) {
    const TwOffsets o = tw_offsets(g);
    if (g.m1 == 1) {
        const size_t shm = (size_t)g.m * sizeof(float2);
        hipLaunchKernelGGL(k_fft_single, dim3(ntrials), dim3(256), shm, s, series, fp16, mean32,
                           corr16, g.n, g.series_pitch, g.m, ilog2(g.m), spec, g.spec_pitch,
                           partial, tw, o);
        return;
    }
    uint32_t C = 8192 / g.m2;
    if (C > g.m1) C = g.m1;
    if (C < 1) C = 1;
    hipLaunchKernelGGL(k_fft_pass1, dim3(g.m1 / C, ntrials), dim3(256),
                       (size_t)C * g.m2 * sizeof(float2), s, series, fp16, mean32, corr16, g.n,
                       g.series_pitch, g.m, g.m1, g.m2, ilog2(g.m2), C, work, tw, o);
    hipLaunchKernelGGL(k_fft_pass2, dim3(g.m2 / 2 + 1, ntrials), dim3(256),
                       (size_t)2 * g.m1 * sizeof(float2), s, work, fp16, g.n, g.m, g.m1,
                       ilog2(g.m1), g.m2, spec, g.spec_pitch, partial, tw, o);
}

}  // namespace psr
