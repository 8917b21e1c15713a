// Whitening, harmonic summing, peak detection and per-trial ranking for gfx950
// (crates/search/src/lib.rs:45-144).
//
// Whitening (lib.rs:45-60): for 64-bin block b the scale is
//   max(max(median(raw[lo..hi]) / ln2, floor), 1e-20),  lo = max(64b-128, 1),
//   hi = min(64b+192, nbins), floor = (mean_f64(raw) * 1e-6) as f32,
// with the reference's upper median (element len/2 of the sorted window).  Each
// workgroup loads a run of blocks plus a two-block halo on each side from the
// (possibly transposed) power spectrum, writes the raw spectrum back in natural
// order, sorts every 64-bin block in LDS (wave bitonic network), and finds each
// window's median exactly with a cut search over the five sorted blocks: a
// position p_j in every block with sum p_j = len/2 such that everything left of
// the cut is <= everything right of it; the median is then the smallest element
// right of the cut.  The harmonic-sum kernel applies the whitening on the fly:
// P_w[i] = raw[i] / scale[i/64] (f32 division, as lib.rs:57-59), 0 at DC and in
// birdie bins (lib.rs:61-72).
#include <hip/hip_fp16.h>
#include <hip/hip_runtime.h>

#include "common.hpp"

namespace psr {

constexpr uint32_t KEY_INF = 0xffffffffu;

// --------------------------------------------------------------- whitening
// Keys are the bit patterns of non-negative values (f32 power, or FP16
// amplitude), whose unsigned order is the numeric order.
__device__ __forceinline__ uint32_t load_key(const void* spec, int fp16, size_t idx) {
    return fp16 ? (uint32_t)((const uint16_t*)spec)[idx] : ((const uint32_t*)spec)[idx];
}
__device__ __forceinline__ float key_power(uint32_t key, int fp16) {
    if (fp16) {
        const float a = __half2float(__ushort_as_half((unsigned short)key));
        return a * a;
    }
    return __uint_as_float(key);
}

__global__ __launch_bounds__(256) void k_whiten(const void* __restrict__ spec_t, int fp16,
                                                uint32_t m, uint32_t m1, uint32_t m2,
                                                uint32_t nbins, uint32_t nblocks,
                                                uint64_t spec_pitch,
                                                const double* __restrict__ partial,
                                                uint32_t npart, void* __restrict__ raw_nat,
                                                float* __restrict__ scales, uint32_t cw) {
    extern __shared__ uint32_t tile[];
    __shared__ float s_floor;
    // tile position p is stored at p + p / m2: the gather below writes positions
    // m2 apart from neighbouring lanes, which the padding spreads over the banks
    // (64-bin blocks never straddle a pad, as m2 >= 512 is a multiple of 64).
    // The natural single-column layout (m1 == 1) needs no padding.
    const int64_t padu = m1 == 1 ? ((int64_t)1 << 40) : (int64_t)m2;
    auto P = [padu](int64_t p) { return p + p / padu; };
    const uint32_t trial = blockIdx.y;
    const uint32_t k1a = blockIdx.x * cw;
    const uint32_t k1b = min(k1a + cw, m1);
    const uint32_t B0 = (k1a * m2) / 64;
    const uint32_t B1 = k1b == m1 ? nblocks : (k1b * m2) / 64;
    const int64_t tile_lo = ((int64_t)B0 - 2) * 64;  // natural bin of tile[0]
    const uint32_t nt_blocks = (B1 - B0) + 4;
    const uint32_t TL = nt_blocks * 64;
    const size_t srow = (size_t)trial * spec_pitch;

    for (uint32_t i = threadIdx.x; i < (uint32_t)P(TL) + 1; i += blockDim.x) tile[i] = KEY_INF;
    if (threadIdx.x == 0) {
        double S = 0.0;
        for (uint32_t i = 0; i < npart; i++) S += partial[(size_t)trial * npart + i];
        s_floor = (float)(S / (double)nbins * 1e-6);
    }
    __syncthreads();
    // gather the tile: columns k1a-1 .. k1b of the [k2][k1] layout, k1 fastest
    const int64_t c_lo = (int64_t)k1a - 1;
    const uint32_t ncols = (k1b - k1a) + 2;
    for (uint32_t idx = threadIdx.x; idx < ncols * m2; idx += blockDim.x) {
        const uint32_t k2 = idx / ncols;
        const int64_t k1 = c_lo + (idx - k2 * ncols);
        if (k1 < 0 || k1 >= (int64_t)m1) continue;
        const int64_t k = (int64_t)k2 + (int64_t)m2 * k1;
        const int64_t pos = k - tile_lo;
        if (pos < 0 || pos >= (int64_t)TL || k < 1) continue;
        tile[P(pos)] = load_key(spec_t, fp16, srow + (size_t)k2 * m1 + (size_t)k1);
    }
    if (B1 == nblocks && threadIdx.x == 0) tile[P((int64_t)m - tile_lo)] = load_key(spec_t, fp16, srow + m);
    __syncthreads();
    // natural-order raw spectrum for the workgroup's own blocks (bin 0 included)
    {
        const uint32_t kb = B0 * 64, ke = min(B1 * 64, nbins);
        for (uint32_t k = kb + threadIdx.x; k < ke; k += blockDim.x) {
            uint32_t key = tile[P((int64_t)k - tile_lo)];
            if (k == 0) key = load_key(spec_t, fp16, srow);
            if (fp16)
                ((uint16_t*)raw_nat)[srow + k] = (uint16_t)key;
            else
                ((uint32_t*)raw_nat)[srow + k] = key;
        }
    }
    // sort each 64-bin block ascending (bin 0 and absent bins hold KEY_INF)
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    for (uint32_t b = wv; b < nt_blocks; b += blockDim.x >> 6) {
        uint32_t* blk = tile + P((int64_t)b * 64);
        uint32_t key = blk[lane];
#pragma unroll
        for (int k = 2; k <= 64; k <<= 1) {
#pragma unroll
            for (int j = k >> 1; j > 0; j >>= 1) {
                const uint32_t other = __shfl_xor(key, j);
                const bool up = (lane & k) == 0;
                const bool lower = (lane & j) == 0;
                key = (lower == up) ? min(key, other) : max(key, other);
            }
        }
        blk[lane] = key;
    }
    __syncthreads();
    // one thread per block of this workgroup: exact window median by cut search
    const uint32_t b = B0 + threadIdx.x;
    if (b >= B1) return;
    int nn[5];
    uint32_t L[5], U[5];
    int p[5];
    const uint32_t* A[5];
    int len = 0;
#pragma unroll
    for (int j = 0; j < 5; j++) {
        const int64_t gb = (int64_t)b - 2 + j;
        int cnt = 0;
        if (gb >= 0 && gb < (int64_t)nblocks) {
            const int64_t s0 = gb * 64 < 1 ? 1 : gb * 64;
            const int64_t s1 = gb * 64 + 64 < (int64_t)nbins ? gb * 64 + 64 : (int64_t)nbins;
            cnt = s1 > s0 ? (int)(s1 - s0) : 0;
        }
        nn[j] = cnt;
        A[j] = tile + P((int64_t)(b - B0 + j) * 64);
        p[j] = cnt >> 1;
        len += cnt;
    }
    const int r = len >> 1;
    int T = 0;
#pragma unroll
    for (int j = 0; j < 5; j++) {
        T += p[j];
        L[j] = p[j] > 0 ? A[j][p[j] - 1] + 1u : 0u;
        U[j] = p[j] < nn[j] ? A[j][p[j]] + 1u : KEY_INF;
    }
    // raise the cut to rank r, taking the smallest element right of it
    for (int it = 0; it < 512 && T < r; it++) {
        int js = 0;
        uint32_t best = U[0];
#pragma unroll
        for (int j = 1; j < 5; j++)
            if (U[j] < best) {
                best = U[j];
                js = j;
            }
#pragma unroll
        for (int j = 0; j < 5; j++)
            if (j == js) {
                L[j] = U[j];
                p[j]++;
                U[j] = p[j] < nn[j] ? A[j][p[j]] + 1u : KEY_INF;
            }
        T++;
    }
    // exchange across the cut until max(left) <= min(right)
    for (int it = 0; it < 1024; it++) {
        int jl = 0, ju = 0;
        uint32_t lmax = L[0], umin = U[0];
#pragma unroll
        for (int j = 1; j < 5; j++) {
            if (L[j] > lmax) {
                lmax = L[j];
                jl = j;
            }
            if (U[j] < umin) {
                umin = U[j];
                ju = j;
            }
        }
        if (lmax <= umin) break;
#pragma unroll
        for (int j = 0; j < 5; j++) {
            if (j == jl) {
                U[j] = L[j];
                p[j]--;
                L[j] = p[j] > 0 ? A[j][p[j] - 1] + 1u : 0u;
            }
            if (j == ju) {
                L[j] = U[j];
                p[j]++;
                U[j] = p[j] < nn[j] ? A[j][p[j]] + 1u : KEY_INF;
            }
        }
    }
    uint32_t umin = U[0];
#pragma unroll
    for (int j = 1; j < 5; j++) umin = min(umin, U[j]);
    float med = len > 0 ? key_power(umin - 1u, fp16) : 0.0f;
    if (len == 0) med = 0.0f;  // rfi::median of an empty slice is 0.0
    const float LN2 = 0.693147180559945309417232121458176568f;
    const float scale = fmaxf(fmaxf(med / LN2, s_floor), 1e-20f);
    scales[(size_t)trial * nblocks + b] = scale;
}

void launch_whiten(hipStream_t s, const SpecGeom& g, const void* spec_t, int fp16,
                   uint32_t ntrials, const double* partial, uint32_t npartials, void* raw_nat,
                   float* scales) {
    uint32_t cw = g.m1 == 1 ? 1 : (g.m1 < 16 ? g.m1 : 16);
    // one workgroup's blocks must fit the 256 threads that select them
    while (cw > 1 && (cw * g.m2) / 64 + 1 > 256) cw >>= 1;
    const uint32_t blocks = (cw * g.m2) / 64 + 1 + 4;
    const size_t pads = g.m1 == 1 ? 0 : (size_t)blocks * 64 / g.m2;
    const size_t shm = ((size_t)blocks * 64 + pads + 1) * sizeof(uint32_t);
    dim3 grid((g.m1 + cw - 1) / cw, ntrials);
    hipLaunchKernelGGL(k_whiten, grid, dim3(256), shm, s, spec_t, fp16, g.m, g.m1, g.m2, g.nbins,
                       g.nblocks, g.spec_pitch, partial, npartials, raw_nat, scales, cw);
}

// ----------------------------------------------------- harmonic summing
__device__ __forceinline__ float raw_at(const void* raw, int fp16, size_t idx) {
    if (fp16) {
        const float a = __half2float(((const __half*)raw)[idx]);
        return a * a;
    }
    return ((const float*)raw)[idx];
}

struct HarmArgs {
    SearchParams p;
    double sqrt_h[5];
    uint32_t birdie[2 * MAX_BIRDIE_INTERVALS];  // inclusive bin intervals
};

__device__ __forceinline__ float whitened(const void* raw, int fp16, size_t row, const float* sc,
                                          uint32_t i, const HarmArgs& a) {
    if (i == 0) return 0.0f;
    for (uint32_t b = 0; b < a.p.nbirdie; b++)
        if (i >= a.birdie[2 * b] && i <= a.birdie[2 * b + 1]) return 0.0f;
    return raw_at(raw, fp16, row + i) / sc[i >> 6];
}

// offset of the parabolic peak through raw[k-1..k+1] (lib.rs:87-94), in f64
__device__ __forceinline__ double interp_offset(double a, double b, double c) {
    const double two_b = 2.0 * b;
    const double curvature = (a - two_b) + c;
    if (curvature < 0.0) {
        const double o = 0.5 * (a - c) / curvature;
        return o < -0.5 ? -0.5 : (o > 0.5 ? 0.5 : o);
    }
    return 0.0;
}

constexpr int HS_T = 256;       // threads: scores for HS_T bins
constexpr int HS_OUT = HS_T - 2;  // peak tests per workgroup

__global__ __launch_bounds__(256) void k_harmsum(const void* __restrict__ raw, int fp16,
                                                 uint64_t spec_pitch, uint32_t nbins,
                                                 uint32_t nblocks,
                                                 const float* __restrict__ scales, HarmArgs a,
                                                 RawPeak* __restrict__ peaks,
                                                 uint32_t* __restrict__ peak_counts) {
    __shared__ double sc[5][HS_T];
    const uint32_t trial = blockIdx.y;
    const size_t row = (size_t)trial * spec_pitch;
    const float* scl = scales + (size_t)trial * nblocks;
    const int tid = threadIdx.x;
    const int64_t k = (int64_t)a.p.first - 1 + (int64_t)blockIdx.x * HS_OUT + tid;
    const double NEG_INF = -__builtin_inf();
    double offset = 0.0;
#pragma unroll
    for (int l = 0; l < 5; l++) sc[l][tid] = NEG_INF;
    if (k >= (int64_t)a.p.first && k <= (int64_t)a.p.end_h[0]) {
        const uint32_t ku = (uint32_t)k;
        offset = interp_offset((double)raw_at(raw, fp16, row + ku - 1),
                               (double)raw_at(raw, fp16, row + ku),
                               (double)raw_at(raw, fp16, row + ku + 1));
        const float pwk = whitened(raw, fp16, row, scl, ku, a);
        double S = 0.0;
        uint32_t j = 1;
        for (uint32_t l = 0; l < a.p.nlevels; l++) {
            const uint32_t h = 1u << l;
            if (k > (int64_t)a.p.end_h[l]) break;
            if (h > 1 && pwk < 4.0f) {
                sc[l][tid] = NEG_INF;
                continue;
            }
            for (; j <= h; j++) {
                const double position = ((double)ku + offset) * (double)j;
                uint64_t index = (uint64_t)round(position);
                if (index > nbins - 1) index = nbins - 1;
                S += (double)whitened(raw, fp16, row, scl, (uint32_t)index, a);
            }
            sc[l][tid] = (S - (double)h) / a.sqrt_h[l];
        }
    }
    __syncthreads();
    if (tid < 1 || tid > HS_OUT) return;
    if (k < (int64_t)a.p.first + 1) return;
    for (uint32_t l = 0; l < a.p.nlevels; l++) {
        if (k > (int64_t)a.p.end_h[l] - 1) break;
        const double s = sc[l][tid];
        if (s >= a.p.threshold && s > sc[l][tid - 1] && s >= sc[l][tid + 1]) {
            const uint32_t slot = atomicAdd(&peak_counts[trial], 1u);
            if (slot < a.p.peak_cap) {
                RawPeak rp;
                rp.score = s;
                rp.offset = offset;
                rp.k = (uint32_t)k;
                rp.h = 1u << l;
                peaks[(size_t)trial * a.p.peak_cap + slot] = rp;
            }
        }
    }
}

void launch_harmsum(hipStream_t s, const SpecGeom& g, const SearchParams& p, const void* raw_nat,
                    const float* scales, const uint32_t* birdies, uint32_t ntrials,
                    RawPeak* peaks, uint32_t* peak_counts) {
    if (p.nlevels == 0 || p.end_h[0] <= p.first + 1) return;
    HarmArgs a;
    a.p = p;
    for (int l = 0; l < 5; l++) a.sqrt_h[l] = 0.0;
    for (int i = 0; i < 2 * MAX_BIRDIE_INTERVALS; i++) a.birdie[i] = i < (int)(2 * p.nbirdie) ? birdies[i] : 0u;
    // sqrt(h) from the host's correctly rounded sqrt
    static const double SQ[5] = {1.0, 1.4142135623730951, 2.0, 2.8284271247461903, 4.0};
    for (int l = 0; l < 5; l++) a.sqrt_h[l] = SQ[l];
    const uint32_t span = p.end_h[0] - p.first + 1;
    dim3 grid((span + HS_OUT - 1) / HS_OUT, ntrials);
    hipLaunchKernelGGL(k_harmsum, grid, dim3(HS_T), 0, s, raw_nat, (int)p.fp16, g.spec_pitch,
                       g.nbins, g.nblocks, scales, a, peaks, peak_counts);
}

// ------------------------------------------------------------- ranking
// Stable descending sort by score of candidates pushed in (h, k) order
// (lib.rs:130) == sort by (score desc, h asc, k asc); then the greedy
// frequency dedup and the cap (lib.rs:131-143).
__global__ __launch_bounds__(256) void k_finalize(const RawPeak* __restrict__ peaks,
                                                  const uint32_t* __restrict__ peak_counts,
                                                  uint32_t cap, double bin_hz, double resolution,
                                                  uint32_t max_candidates,
                                                  const double* __restrict__ dms,
                                                  CandidateOut* __restrict__ out,
                                                  uint32_t cap_per_trial,
                                                  uint32_t* __restrict__ counts,
                                                  uint32_t* __restrict__ overflow) {
    __shared__ double s_score[PEAK_CAP];
    __shared__ uint32_t s_hk[PEAK_CAP];
    __shared__ uint16_t s_idx[PEAK_CAP];
    __shared__ double s_kept[256];
    __shared__ uint32_t s_nk;
    const uint32_t trial = blockIdx.x;
    uint32_t cnt = peak_counts[trial];
    if (cnt > cap) {
        if (threadIdx.x == 0) atomicAdd(overflow, 1u);
        cnt = cap;
    }
    const RawPeak* pk = peaks + (size_t)trial * cap;
    uint32_t P = 1;
    while (P < cnt) P <<= 1;
    for (uint32_t i = threadIdx.x; i < P; i += blockDim.x) {
        if (i < cnt) {
            s_score[i] = pk[i].score;
            s_hk[i] = (pk[i].h << 27) | pk[i].k;
        } else {
            s_score[i] = -__builtin_inf();
            s_hk[i] = 0xffffffffu;
        }
        s_idx[i] = (uint16_t)i;
    }
    __syncthreads();
    // bitonic sort of (score desc, hk asc) over P entries
    for (uint32_t kk = 2; kk <= P; kk <<= 1) {
        for (uint32_t j = kk >> 1; j > 0; j >>= 1) {
            for (uint32_t i = threadIdx.x; i < P; i += blockDim.x) {
                const uint32_t ixj = i ^ j;
                if (ixj > i) {
                    const bool asc = (i & kk) == 0;
                    const double si = s_score[i], sj = s_score[ixj];
                    const uint32_t hi = s_hk[i], hj = s_hk[ixj];
                    // "i before j" in the final order
                    const bool i_first = si > sj || (si == sj && hi < hj);
                    if (asc != i_first) {
                        s_score[i] = sj;
                        s_score[ixj] = si;
                        s_hk[i] = hj;
                        s_hk[ixj] = hi;
                        const uint16_t t = s_idx[i];
                        s_idx[i] = s_idx[ixj];
                        s_idx[ixj] = t;
                    }
                }
            }
            __syncthreads();
        }
    }
    if (threadIdx.x >= 64) return;
    // greedy dedup by one wave; kept frequencies checked lane-parallel
    const int lane = threadIdx.x;
    uint32_t nk = 0;
    const uint32_t limit = max_candidates < cap_per_trial ? max_candidates : cap_per_trial;
    for (uint32_t i = 0; i < cnt; i++) {
        const RawPeak rp = pk[s_idx[i]];
        const double f = ((double)rp.k + rp.offset) * bin_hz;
        bool dup = false;
        for (uint32_t base = 0; base < nk; base += 64) {
            const uint32_t j = base + lane;
            const bool d = j < nk && fabs(s_kept[j] - f) < resolution;
            dup = dup || __any(d);
        }
        if (dup) continue;
        if (lane == 0) {
            CandidateOut c;
            c.frequency_hz = f;
            c.period_s = 1.0 / f;
            c.dm = dms[trial];
            c.harmonics = rp.h;
            c.score = rp.score;
            out[(size_t)trial * cap_per_trial + nk] = c;
            s_kept[nk] = f;
        }
        __builtin_amdgcn_s_waitcnt(0);
        __builtin_amdgcn_wave_barrier();
        nk++;
        if (nk >= limit || nk >= 256) break;
    }
    if (lane == 0) counts[trial] = nk;
}

void launch_finalize(hipStream_t s, const SpecGeom& g, const SearchParams& p, const double* dms,
                     uint32_t ntrials, const RawPeak* peaks, const uint32_t* peak_counts,
                     CandidateOut* out, uint32_t cap_per_trial, uint32_t* counts) {
    (void)g;
    // the overflow counter lives right after the per-trial counts
    hipLaunchKernelGGL(k_finalize, dim3(ntrials), dim3(256), 0, s, peaks, peak_counts, p.peak_cap,
                       p.bin_hz, p.resolution, p.max_candidates, dms, out, cap_per_trial, counts,
                       (uint32_t*)(counts + ntrials));
}

}  // namespace psr
