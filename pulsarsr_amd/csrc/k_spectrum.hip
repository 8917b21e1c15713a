// Whitening, harmonic summing, peak detection and per-trial ranking for gfx950
// (crates/search/src/lib.rs:45-144).
//
// Whitening (lib.rs:45-60): for 64-bin block b the scale is
//   max(max(median(raw[lo..hi]) / ln2, floor), 1e-20),  lo = max(64b-128, 1),
//   hi = min(64b+192, nbins), floor = (mean_f64(raw) * 1e-6) as f32,
// with the reference's upper median (element len/2 of the sorted window).  A
// first kernel reads the (possibly transposed) power spectrum column tile by
// column tile, writes it back in natural order and sorts every 64-bin block in
// registers; a second finds each window's median exactly from the five sorted
// blocks it spans (cut search).  The harmonic-sum kernel applies the whitening
// on the fly: P_w[i] = raw[i] / scale[i/64] (f32 division, as lib.rs:57-59), 0
// at DC and in birdie bins (lib.rs:61-72).
#include <hip/hip_fp16.h>
#include <hip/hip_runtime.h>

#include <cstring>

#include <type_traits>

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

// Ascending bitonic sorting network over registers (indices are compile-time
// constants once unrolled, so the array stays in VGPRs).
template <int N>
__device__ __forceinline__ void bitonic_sort_regs(uint32_t (&v)[N]) {
#pragma unroll
    for (int k = 2; k <= N; k <<= 1)
#pragma unroll
        for (int j = k >> 1; j > 0; j >>= 1)
#pragma unroll
            for (int i = 0; i < N; i++) {
                const int l = i ^ j;
                if (l > i) {
                    const uint32_t a = v[i], b = v[l];
                    if ((i & k) == 0) {
                        v[i] = min(a, b);
                        v[l] = max(a, b);
                    } else {
                        v[i] = max(a, b);
                        v[l] = min(a, b);
                    }
                }
            }
}

// Ascending sort of 64 16-bit keys held two to a register, w[i] = (key i,
// key i + 32).  The bitonic network is used in its all-ascending form (each
// merge of two sorted runs of k/2 keys starts by comparing key i with its
// mirror i ^ (k - 1), then halves with i ^ j, j = k/4 .. 1), so every
// comparison of the merges up to k = 32 pairs the same half of two registers
// and one packed 16-bit min/max pair performs two compare-exchanges; only the
// mirror step of the final merge pairs key a with key 63 - a, i.e. the low half
// of w[a] with the high half of w[31 - a].
typedef unsigned short u16x2 __attribute__((ext_vector_type(2)));
__device__ __forceinline__ void ce_packed(u16x2& a, u16x2& b) {
    const u16x2 x = a, y = b;
    a = __builtin_elementwise_min(x, y);
    b = __builtin_elementwise_max(x, y);
}
__device__ __forceinline__ void sort64_packed(u16x2 (&w)[32]) {
#pragma unroll
    for (int k = 2; k <= 32; k <<= 1) {
#pragma unroll
        for (int i = 0; i < 32; i++)
            if ((i ^ (k - 1)) > i) ce_packed(w[i], w[i ^ (k - 1)]);
#pragma unroll
        for (int j = k >> 2; j > 0; j >>= 1)
#pragma unroll
            for (int i = 0; i < 32; i++)
                if ((i ^ j) > i) ce_packed(w[i], w[i ^ j]);
    }
#pragma unroll
    for (int a = 0; a < 16; a++) {  // mirror step of the 64-key merge
        const int b = 31 - a;
        const u16x2 sb = w[b].yx;
        const u16x2 x = __builtin_elementwise_min(w[a], sb), y = __builtin_elementwise_max(w[a], sb);
        w[a] = u16x2{x.x, y.y};  // keys a, 32 + a
        w[b] = u16x2{x.y, y.x};  // keys 31 - a, 63 - a
    }
#pragma unroll
    for (int j = 16; j > 0; j >>= 1)
#pragma unroll
        for (int i = 0; i < 32; i++)
            if ((i ^ j) > i) ce_packed(w[i], w[i ^ j]);
}

constexpr uint32_t SB_BLOCKS = 256;  // natural 64-bin blocks per sorting workgroup (max)
constexpr uint32_t MB_BLOCKS = 128;  // blocks per median workgroup
constexpr uint32_t MB_STRIDE = 65;   // LDS words per staged block (odd: spreads banks)

// LDS word of element i of tile block B: 16-byte chunks of a block are XOR
// permuted so that the 8 lanes of a ds_read_b128 group (8 consecutive blocks)
// hit different banks, and so that the gather's stores of different columns
// spread over the banks.
__device__ __forceinline__ uint32_t sb_h(uint32_t B) { return (B ^ (B >> 3) ^ (B >> 6)) & 7u; }
__device__ __forceinline__ uint32_t sb_word(uint32_t B, uint32_t i) {
    return B * 64 + ((((i >> 2) ^ sb_h(B)) << 2) | (i & 3u));
}

template <bool F16>
__device__ __forceinline__ void store_keys8(void* dst, size_t o, const uint32_t* k) {
    // 8 consecutive keys (F16: one 16-byte store; f32: two)
    if (F16) {
        *(uint4*)((uint16_t*)dst + o) =
            make_uint4(k[0] | k[1] << 16, k[2] | k[3] << 16, k[4] | k[5] << 16, k[6] | k[7] << 16);
    } else {
        *(uint4*)((uint32_t*)dst + o) = make_uint4(k[0], k[1], k[2], k[3]);
        *(uint4*)((uint32_t*)dst + o + 4) = make_uint4(k[4], k[5], k[6], k[7]);
    }
}

// Stage 1: one workgroup per `cw` columns of the [k2][k1] (k = k2 + m2 k1)
// spectrum, i.e. the natural bins [m2 k1a, m2 (k1a + cw)).  It writes them back
// in natural order (raw_nat), and each lane sorts one 64-bin block in registers
// and stores it ascending to `sorted` (DC and bins beyond Nyquist as KEY_INF).
// Workgroup 0 of each trial also fixes the whitening floor
// (mean power * 1e-6, partial sums added in order).  The four-step layout
// (cw * m2 = SB_BLOCKS * 64) is read with 16-byte row segments, all of a
// thread's loads issued before any is used; a short single-pass spectrum
// (one column) takes the scalar path.
template <bool F16>
__global__ __launch_bounds__(256) void k_sort_blocks(
    const void* __restrict__ spec_t, uint32_t m, uint32_t m1, uint32_t m2, uint32_t cw,
    uint32_t nbins, uint64_t spec_pitch, const double* __restrict__ partial, uint32_t npart,
    void* __restrict__ raw_nat, void* __restrict__ sorted, float* __restrict__ floors) {
    extern __shared__ __align__(16) uint32_t tile[];
    constexpr int E = F16 ? 8 : 4;                            // keys per 16-byte vector
    constexpr int VPT = (int)(SB_BLOCKS * 64 / E / 256);      // vectors per thread
    const uint2 wb = xcd_block();  // column groups sharing cache lines on one XCD
    const uint32_t trial = wb.y;
    const uint32_t k1a = wb.x * cw;
    const uint32_t k_lo = k1a * m2;
    const uint32_t span = cw * m2;             // natural bins of the columns
    const uint32_t nbk = (span + 63) / 64;     // tile blocks
    const size_t srow = (size_t)trial * spec_pitch;
    if (wb.x == 0 && threadIdx.x == 0) {
        double S = 0.0;
        for (uint32_t i = 0; i < npart; i++) S += partial[(size_t)trial * npart + i];
        floors[trial] = (float)(S / (double)nbins * 1e-6);
    }
    if (span == SB_BLOCKS * 64 && cw % E == 0) {
        const uint32_t vpr = cw / E;  // vectors per row segment
        uint4 q[VPT];
#pragma unroll
        for (int t = 0; t < VPT; t++) {
            const uint32_t v = threadIdx.x + 256u * t;
            const uint32_t k2 = v / vpr, qq = v - k2 * vpr;
            const size_t idx = srow + (size_t)k2 * m1 + k1a + qq * E;
            q[t] = F16 ? *(const uint4*)((const uint16_t*)spec_t + idx)
                       : *(const uint4*)((const uint32_t*)spec_t + idx);
        }
#pragma unroll
        for (int t = 0; t < VPT; t++) {
            const uint32_t v = threadIdx.x + 256u * t;
            const uint32_t k2 = v / vpr, qq = v - k2 * vpr;
            const uint32_t w[4] = {q[t].x, q[t].y, q[t].z, q[t].w};
#pragma unroll
            for (int e = 0; e < E; e++) {
                const uint32_t key = F16 ? (w[e >> 1] >> (16 * (e & 1))) & 0xffffu : w[e];
                const uint32_t j = k2 + m2 * (qq * E + e);
                tile[sb_word(j >> 6, j & 63)] = key;
            }
        }
    } else {
        for (uint32_t idx = threadIdx.x; idx < span; idx += blockDim.x) {
            const uint32_t k2 = idx / cw, c = idx - k2 * cw;
            const uint32_t j = k2 + m2 * c;
            tile[sb_word(j >> 6, j & 63)] = load_key(spec_t, F16, srow + (size_t)k2 * m1 + k1a + c);
        }
        // a short single-pass spectrum (m < 64) shares its only block with Nyquist
        for (uint32_t j = span + threadIdx.x; j < nbk * 64; j += blockDim.x)
            tile[sb_word(j >> 6, j & 63)] =
                k_lo + j == m ? load_key(spec_t, F16, srow + m) : KEY_INF;
    }
    __syncthreads();
    // natural-order raw spectrum, 8 keys per thread and step
    const uint32_t jend = min(nbk * 64, m + 1 - k_lo);
    for (uint32_t j = 8 * threadIdx.x; j < jend; j += 8 * blockDim.x) {
        const uint32_t Bj = j >> 6, i = j & 63;
        const uint4 a = *(const uint4*)(tile + Bj * 64 + ((((i >> 2)) ^ sb_h(Bj)) << 2));
        const uint4 b = *(const uint4*)(tile + Bj * 64 + ((((i >> 2) + 1) ^ sb_h(Bj)) << 2));
        const uint32_t k[8] = {a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w};
        if (j + 8 <= jend) {
            store_keys8<F16>(raw_nat, srow + k_lo + j, k);
        } else {
            for (uint32_t e = 0; j + e < jend; e++) {
                if (F16)
                    ((uint16_t*)raw_nat)[srow + k_lo + j + e] = (uint16_t)k[e];
                else
                    ((uint32_t*)raw_nat)[srow + k_lo + j + e] = k[e];
            }
        }
    }
    if (k_lo + nbk * 64 <= m && k1a + cw == m1 && threadIdx.x == 0) {
        const uint32_t key = load_key(spec_t, F16, srow + m);  // Nyquist, outside the tile
        if (F16)
            ((uint16_t*)raw_nat)[srow + m] = (uint16_t)key;
        else
            ((uint32_t*)raw_nat)[srow + m] = key;
    }
    const uint32_t B = threadIdx.x;
    if (B >= nbk) return;
    uint32_t v[64];
    const uint32_t h = sb_h(B);
#pragma unroll
    for (int c = 0; c < 16; c++) {
        const uint4 q = *(const uint4*)(tile + B * 64 + ((c ^ h) << 2));
        v[4 * c] = q.x;
        v[4 * c + 1] = q.y;
        v[4 * c + 2] = q.z;
        v[4 * c + 3] = q.w;
    }
    if (k_lo == 0 && B == 0) v[0] = KEY_INF;  // DC is not part of any window
    bitonic_sort_regs<64>(v);
    const size_t o = srow + (size_t)(k_lo / 64 + B) * 64;
#pragma unroll
    for (int c = 0; c < 8; c++) store_keys8<F16>(sorted, o + 8 * c, v + 8 * c);
}

// Staged sorted blocks: block lb starts at element lb * STRIDE + one extra word
// per 8 blocks.  The stride alone (an odd number of words) spreads consecutive
// blocks over the banks, as the window medians read them; the extra word also
// spreads blocks 8 apart, which the lanes of a wave store together (lane l holds
// block ch + 8 col for 512-row columns, 16 or 32 apart for longer ones).
template <typename KeyT, int STRIDE>
__device__ __forceinline__ uint32_t key_block(uint32_t lb) {
    return lb * (uint32_t)STRIDE + (lb >> 3) * (uint32_t)(4 / sizeof(KeyT));
}
template <typename KeyT, int STRIDE>
constexpr uint32_t key_blocks_elems(uint32_t nb) {  // elements of nb staged blocks
    return nb * (uint32_t)STRIDE + ((nb + 7) >> 3) * (uint32_t)(4 / sizeof(KeyT));
}
// 16-bit keys: the two words after each block's 64 keys hold 0xFFFF, the value
// every key past a block's count has, which the window median's reads rely on
// (threads t of nt fill them; the block stores never touch them)
template <int STRIDE>
__device__ __forceinline__ void pad_key_blocks(uint16_t* s, uint32_t nb, uint32_t t, uint32_t nt) {
    static_assert(STRIDE >= 66, "two pad words per block");
    for (uint32_t i = t; i < 2 * nb; i += nt) s[key_block<uint16_t, STRIDE>(i >> 1) + 64 + (i & 1)] = 0xFFFFu;
}

// Whitening scale of block b (lib.rs:45-60) from the sorted blocks b-2 .. b+2
// at s + key_block(lb + j): the exact window median is found by a cut search --
// a position p_j in every block with sum p_j = len/2 such that all keys left of
// the cut are <= all keys right of it; the median is then the smallest key
// right of the cut.  Blocks outside [0, nblocks) hold nothing and are not read.
template <bool F16, typename KeyT, int STRIDE>
__device__ float window_scale(const KeyT* s, uint32_t lb, uint32_t b, uint32_t nblocks, uint32_t nbins,
                              float floor) {
    int nn[5];
    uint32_t L[5], U[5];
    int p[5];
    const KeyT* A[5];
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
        A[j] = s + key_block<KeyT, STRIDE>(lb + j);
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
    // exchange across the cut until max(left) <= min(right): the block holding
    // the largest key left of the cut gives up its top keys, the block holding
    // the smallest key right of the cut takes in its bottom keys -- 4 or 2 at a
    // time while every key moved right is larger than every key moved left
    // (each such step only swaps inverted pairs, so the search still ends at
    // the unique median cut), else 1.  Branch-free: both blocks are selected
    // arithmetically, the eight keys next to their cuts are read together
    // (indices clamped into the block, out-of-range keys replaced by the
    // sentinels), and the five cut states are updated with selects.
    if constexpr (F16) {
        // 16-bit keys: each block's state packed as (bound << 10 | p << 3 | j), so a
        // max3/min3 pair picks the block (ties by p, then j: any choice ends at the
        // same median) and carries its cut; U's infinity is 0x10000, which is also
        // key + 1 for every stored key past a block's count (0xFFFF, pads included),
        // so the reads need no count checks: a key past the count never exceeds
        // the other side, and a new U there is infinite
        uint32_t Lp[5], Up[5];
#pragma unroll
        for (int j = 0; j < 5; j++) {
            Lp[j] = L[j] << 10 | (uint32_t)p[j] << 3 | (uint32_t)j;
            Up[j] = (U[j] == KEY_INF ? 0x10000u : U[j]) << 10 | (uint32_t)p[j] << 3 | (uint32_t)j;
        }
        for (int it = 0; it < 1024; it++) {
            const uint32_t lm = max(max(max(Lp[0], Lp[1]), Lp[2]), max(Lp[3], Lp[4]));
            const uint32_t um = min(min(min(Up[0], Up[1]), Up[2]), min(Up[3], Up[4]));
            if ((lm >> 10) <= (um >> 10)) break;
            const uint32_t jl = lm & 7u, ju = um & 7u;
            const int pl = (int)((lm >> 3) & 127u), pu = (int)((um >> 3) & 127u);
            const KeyT* Al = s + key_block<KeyT, STRIDE>(lb + jl);
            const KeyT* Au = s + key_block<KeyT, STRIDE>(lb + ju);
            auto below = [](int i) { return i < 0 ? 0 : i; };
            auto above = [](int i) { return i > 65 ? 65 : i; };
            const uint32_t a2 = Al[below(pl - 2)], a3 = Al[below(pl - 3)], a4 = Al[below(pl - 4)],
                           a5 = Al[below(pl - 5)];
            const uint32_t b1 = Au[above(pu + 1)], b2 = Au[above(pu + 2)], b3 = Au[above(pu + 3)],
                           b4 = Au[above(pu + 4)];
            const bool s4 = pl >= 4 && a4 > b3;
            const bool s2 = pl >= 2 && a2 > b1;
            const int st = s4 ? 4 : (s2 ? 2 : 1);
            const uint32_t kl = s4 ? a5 : (s2 ? a3 : a2);
            const uint32_t newLl = pl - st >= 1 ? kl + 1u : 0u;
            const uint32_t newUl = s4 ? a4 + 1u : (s2 ? a2 + 1u : lm >> 10);
            const uint32_t newLu = s4 ? b3 + 1u : (s2 ? b1 + 1u : um >> 10);
            const uint32_t newUu = (s4 ? b4 : (s2 ? b2 : b1)) + 1u;
            const uint32_t ql = (uint32_t)(pl - st) << 3 | jl, qu = (uint32_t)(pu + st) << 3 | ju;
            const uint32_t lpl = newLl << 10 | ql, upl = newUl << 10 | ql;
            const uint32_t lpu = newLu << 10 | qu, upu = newUu << 10 | qu;
#pragma unroll
            for (int j = 0; j < 5; j++) {
                const bool isl = (uint32_t)j == jl, isu = (uint32_t)j == ju;
                Lp[j] = isl ? lpl : (isu ? lpu : Lp[j]);
                Up[j] = isl ? upl : (isu ? upu : Up[j]);
            }
        }
#pragma unroll
        for (int j = 0; j < 5; j++) U[j] = Up[j] >> 10;
    } else {
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
            int pl = p[0], pu = p[0], nu = nn[0];
#pragma unroll
            for (int j = 1; j < 5; j++) {
                pl = j == jl ? p[j] : pl;
                pu = j == ju ? p[j] : pu;
                nu = j == ju ? nn[j] : nu;
            }
            // jl != ju (within one sorted block L <= U)
            const KeyT* Al = s + key_block<KeyT, STRIDE>(lb + jl);
            const KeyT* Au = s + key_block<KeyT, STRIDE>(lb + ju);
            auto below = [](int i) { return i < 0 ? 0 : i; };
            auto above = [](int i) { return i > 63 ? 63 : i; };
            const uint32_t a2 = Al[below(pl - 2)], a3 = Al[below(pl - 3)], a4 = Al[below(pl - 4)],
                           a5 = Al[below(pl - 5)];
            const uint32_t b1 = Au[above(pu + 1)], b2 = Au[above(pu + 2)], b3 = Au[above(pu + 3)],
                           b4 = Au[above(pu + 4)];
            const bool s4 = pl >= 4 && pu + 4 <= nu && a4 > b3;
            const bool s2 = pl >= 2 && pu + 2 <= nu && a2 > b1;
            const int st = s4 ? 4 : (s2 ? 2 : 1);
            // jl keeps keys [0, pl - st): new L = key pl - st - 1, new U = key pl - st
            const uint32_t kl = s4 ? a5 : (s2 ? a3 : a2);
            const uint32_t newLl = pl - st >= 1 ? kl + 1u : 0u;
            const uint32_t newUl = s4 ? a4 + 1u : (s2 ? a2 + 1u : lmax);
            // ju's left part grows to [0, pu + st): new L = key pu + st - 1, new U = key pu + st
            const uint32_t newLu = s4 ? b3 + 1u : (s2 ? b1 + 1u : umin);
            const uint32_t ku = s4 ? b4 : (s2 ? b2 : b1);
            const uint32_t newUu = pu + st < nu ? ku + 1u : KEY_INF;
#pragma unroll
            for (int j = 0; j < 5; j++) {
                const bool isl = j == jl, isu = j == ju;
                p[j] += isu ? st : (isl ? -st : 0);
                const uint32_t l2 = isl ? newLl : (isu ? newLu : L[j]);
                const uint32_t u2 = isl ? newUl : (isu ? newUu : U[j]);
                L[j] = l2;
                U[j] = u2;
            }
        }
    }
    uint32_t umin = U[0];
#pragma unroll
    for (int j = 1; j < 5; j++) umin = min(umin, U[j]);
    float med = len > 0 ? key_power(umin - 1u, F16) : 0.0f;
    if (len == 0) med = 0.0f;  // rfi::median of an empty slice is 0.0
    const float LN2 = 0.693147180559945309417232121458176568f;
    return fmaxf(fmaxf(med / LN2, floor), 1e-20f);
}

// Stage 2: one thread per 64-bin block b; the five sorted blocks b-2 .. b+2 of
// its window are staged in LDS.  Block `nsorted` (present when m is a multiple
// of 64) holds Nyquist alone.
template <bool F16>
__global__ __launch_bounds__(MB_BLOCKS) void k_block_median(
    const void* __restrict__ sorted, const void* __restrict__ spec_t, uint32_t m, uint32_t nbins,
    uint32_t nblocks, uint64_t spec_pitch, const float* __restrict__ floors,
    float* __restrict__ scales) {
    // FP16 keys are staged as 16-bit values (66 per block: an odd word stride),
    // which halves the footprint and doubles the workgroups per CU
    using KeyT = typename std::conditional<F16, uint16_t, uint32_t>::type;
    constexpr int STRIDE = F16 ? 2 * 33 : (int)MB_STRIDE;
    __shared__ KeyT s[key_blocks_elems<KeyT, STRIDE>(MB_BLOCKS + 4)];
    if constexpr (F16) pad_key_blocks<STRIDE>(s, MB_BLOCKS + 4, threadIdx.x, MB_BLOCKS);
    constexpr int E = F16 ? 8 : 4;                                  // keys per 16-byte load
    constexpr int NV = (int)((MB_BLOCKS + 4) * 64 / E);             // vectors staged
    constexpr int VPT = (NV + (int)MB_BLOCKS - 1) / (int)MB_BLOCKS;  // per thread
    const uint2 wb = xcd_block();
    const uint32_t trial = wb.y;
    const uint32_t b0 = wb.x * MB_BLOCKS;
    const size_t srow = (size_t)trial * spec_pitch;
    const uint32_t nsorted = (m + 63) / 64;
    uint4 q[VPT];
#pragma unroll
    for (int t = 0; t < VPT; t++) {
        const int v = threadIdx.x + MB_BLOCKS * t;
        const int64_t gb = (int64_t)b0 - 2 + (v * E) / 64;
        q[t] = make_uint4(KEY_INF, KEY_INF, KEY_INF, KEY_INF);
        if (v < NV && gb >= 0 && gb < (int64_t)nsorted) {
            const size_t idx = srow + (size_t)gb * 64 + (v * E) % 64;
            q[t] = F16 ? *(const uint4*)((const uint16_t*)sorted + idx)
                       : *(const uint4*)((const uint32_t*)sorted + idx);
        }
    }
#pragma unroll
    for (int t = 0; t < VPT; t++) {
        const int v = threadIdx.x + MB_BLOCKS * t;
        if (v >= NV) break;
        const uint32_t lb = (v * E) / 64, i0 = (v * E) % 64;
        const int64_t gb = (int64_t)b0 - 2 + lb;
        const uint32_t w[4] = {q[t].x, q[t].y, q[t].z, q[t].w};
#pragma unroll
        for (int e = 0; e < E; e++) {
            uint32_t key = F16 ? (w[e >> 1] >> (16 * (e & 1))) & 0xffffu : w[e];
            if (!(gb >= 0 && gb < (int64_t)nsorted)) key = KEY_INF;
            // block `nsorted` (m a multiple of 64) holds Nyquist alone
            if (gb == (int64_t)nsorted && gb < (int64_t)nblocks && i0 + e == 0)
                key = load_key(spec_t, F16, srow + m);
            s[key_block<KeyT, STRIDE>(lb) + i0 + e] = (KeyT)key;  // FP16: KEY_INF keeps its low 16 bits
        }
    }
    __syncthreads();
    const uint32_t b = b0 + threadIdx.x;
    if (b >= nblocks) return;
    scales[(size_t)trial * nblocks + b] =
        window_scale<F16, KeyT, STRIDE>(s, threadIdx.x, b, nblocks, nbins, floors[trial]);
}


// Fused whitening for the four-step layout (cw * m2 = SB_BLOCKS * 64, m2 a
// multiple of 64): the workgroup of columns [k1a, k1a + cw) owns the natural
// blocks of those columns.  A natural block is 64 consecutive rows of one
// column of the [k2][k1] spectrum, so lane l of a wave loads the block of
// column k1a + l % cw straight into registers (each load instruction reads
// whole row segments), writes it back in natural order, and sorts it; the two
// blocks on either side of the range (needed by its windows) are sorted across
// the lanes of a wave; every block's window median is then found from LDS --
// no transpose through LDS and no sorted copy through HBM.  The last workgroup
// of a trial also scores the Nyquist block (m a multiple of 64).  Every
// workgroup recomputes the trial's whitening floor (mean power * 1e-6, partial
// sums added in order).
template <bool F16>
constexpr int wf_stride() { return F16 ? 2 * 33 : (int)MB_STRIDE; }
template <bool F16>
constexpr size_t wf_lds_bytes() {
    using KeyT = typename std::conditional<F16, uint16_t, uint32_t>::type;
    return (size_t)key_blocks_elems<KeyT, wf_stride<F16>()>(SB_BLOCKS + 6) * sizeof(KeyT) + 4 * 64 * 4;  // + side blocks
}
template <bool F16>
__global__ __launch_bounds__(256) void k_whiten_fused(
    const void* __restrict__ spec_t, uint32_t m, uint32_t m1, uint32_t m2, uint32_t cw,
    uint32_t nbins, uint32_t nblocks, uint64_t spec_pitch, const double* __restrict__ partial,
    uint32_t npart, void* __restrict__ raw_nat, float* __restrict__ scales) {
    using KeyT = typename std::conditional<F16, uint16_t, uint32_t>::type;
    constexpr int STRIDE = wf_stride<F16>();
    constexpr size_t MAIN = wf_lds_bytes<F16>() - 4 * 64 * 4;
    extern __shared__ __align__(16) uint32_t lds_w[];
    KeyT* sk = (KeyT*)lds_w;             // sorted blocks, local block lb = global - b0 + 2
    uint32_t* side = lds_w + MAIN / 4;  // [4][64]: blocks b0-2, b0-1, b0+SB, b0+SB+1
    __shared__ float sfloor;
    if constexpr (F16) pad_key_blocks<STRIDE>(sk, SB_BLOCKS + 6, threadIdx.x, 256);
    const uint2 wb = xcd_block();
    const uint32_t trial = wb.y;
    const uint32_t k1a = wb.x * cw;
    const uint32_t k_lo = k1a * m2, b0 = k_lo / 64;
    const uint32_t nsorted = (m + 63) / 64;
    const uint32_t cpb = m2 / 64;  // blocks per column
    const size_t srow = (size_t)trial * spec_pitch;
    // this lane's block: column k1a + col, rows 64 ch .. 64 ch + 63
    const uint32_t col = threadIdx.x % cw, ch = threadIdx.x / cw;
    const uint32_t B = ch + cpb * col;  // local natural block
    uint32_t v[64];
    {
        const size_t base = srow + (size_t)(64 * ch) * m1 + k1a + col;
#pragma unroll
        for (int r = 0; r < 64; r++) v[r] = load_key(spec_t, F16, base + (size_t)r * m1);
    }
    {  // side blocks: thread t takes element t % 64 of side block t / 64
        const uint32_t sb = threadIdx.x >> 6, i = threadIdx.x & 63;
        const int64_t gb = sb < 2 ? (int64_t)b0 - 2 + sb : (int64_t)b0 + SB_BLOCKS + (sb - 2);
        uint32_t key = KEY_INF;
        if (gb >= 0 && gb < (int64_t)nsorted) {
            const uint32_t K = (uint32_t)gb * 64 + i;
            if (K > 0 && K < m) key = load_key(spec_t, F16, srow + (size_t)(K % m2) * m1 + K / m2);
        } else if (gb == (int64_t)nsorted && gb < (int64_t)nblocks && i == 0) {
            key = load_key(spec_t, F16, srow + m);  // block nsorted holds Nyquist alone
        }
        side[threadIdx.x] = key;
    }
    if (threadIdx.x == 0) {
        double S = 0.0;
        for (uint32_t i = 0; i < npart; i++) S += partial[(size_t)trial * npart + i];
        sfloor = (float)(S / (double)nbins * 1e-6);
    }
    // natural-order raw spectrum: the block's 64 keys
    {
        const size_t o = srow + (size_t)(b0 + B) * 64;
#pragma unroll
        for (int c = 0; c < 8; c++) store_keys8<F16>(raw_nat, o + 8 * c, v + 8 * c);
        if (k1a + cw == m1 && threadIdx.x == 0) {
            const uint32_t key = load_key(spec_t, F16, srow + m);  // Nyquist, outside every block
            if (F16)
                ((uint16_t*)raw_nat)[srow + m] = (uint16_t)key;
            else
                ((uint32_t*)raw_nat)[srow + m] = key;
        }
    }
    if constexpr (F16) {
        // 16-bit keys two to a register, (key i, key i + 32)
        u16x2 w[32];
#pragma unroll
        for (int i = 0; i < 32; i++) w[i] = u16x2{(unsigned short)v[i], (unsigned short)v[i + 32]};
        if (k_lo == 0 && B == 0) w[0].x = 0xffffu;  // DC is not part of any window
        sort64_packed(w);
#pragma unroll
        for (int i = 0; i < 32; i++) {
            sk[key_block<KeyT, STRIDE>(B + 2) + i] = w[i].x;
            sk[key_block<KeyT, STRIDE>(B + 2) + i + 32] = w[i].y;
        }
    } else {
        if (k_lo == 0 && B == 0) v[0] = KEY_INF;  // DC is not part of any window
        bitonic_sort_regs<64>(v);
#pragma unroll
        for (int i = 0; i < 64; i++) sk[key_block<KeyT, STRIDE>(B + 2) + i] = (KeyT)v[i];
    }
    __syncthreads();  // side keys and the floor are in LDS
    // side block w (waves 0..3): bitonic network across the wave's lanes
    const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    uint32_t x = side[threadIdx.x];
#pragma unroll
    for (int k = 2; k <= 64; k <<= 1)
#pragma unroll
        for (int j = k >> 1; j > 0; j >>= 1) {
            const uint32_t pv = (uint32_t)__shfl_xor((int)x, j);
            const bool up = (lane & k) == 0, lower = (lane & j) == 0;
            x = lower == up ? min(x, pv) : max(x, pv);
        }
    sk[key_block<KeyT, STRIDE>(wave < 2 ? wave : SB_BLOCKS + wave) + lane] = (KeyT)x;
    if (threadIdx.x < 2 * STRIDE)  // local blocks SB + 4 and SB + 5 (contiguous): never counted, kept defined
        sk[key_block<KeyT, STRIDE>(SB_BLOCKS + 4) + threadIdx.x] = (KeyT)KEY_INF;
    __syncthreads();
    const float fl = sfloor;
    const uint32_t bm = threadIdx.x;  // medians in natural block order
    scales[(size_t)trial * nblocks + b0 + bm] = window_scale<F16, KeyT, STRIDE>(sk, bm, b0 + bm, nblocks, nbins, fl);
    if (bm == 0 && k1a + cw == m1 && nsorted < nblocks)  // the Nyquist block
        scales[(size_t)trial * nblocks + nsorted] =
            window_scale<F16, KeyT, STRIDE>(sk, SB_BLOCKS, nsorted, nblocks, nbins, fl);
}

// FP16 form of k_whiten_fused for WH_GROUPS column groups per workgroup: the
// keys stay packed two to a register, (key i, key i + 32), from the loads on,
// and the next group's 64 loads per lane are issued as soon as the current
// group's keys are in the sorting registers, so their latency overlaps its
// sort, LDS exchange and window medians.
constexpr uint32_t WH_GROUPS = 4;
__global__ __launch_bounds__(256) void k_whiten_fused16(
    const uint16_t* __restrict__ spec_t, uint32_t m, uint32_t m1, uint32_t m2, uint32_t cw,
    uint32_t nbins, uint32_t nblocks, uint64_t spec_pitch, const double* __restrict__ partial,
    uint32_t npart, uint16_t* __restrict__ raw_nat, float* __restrict__ scales) {
    constexpr int STRIDE = wf_stride<true>();
    constexpr size_t MAIN = wf_lds_bytes<true>() - 4 * 64 * 4;
    extern __shared__ __align__(16) uint32_t lds_w[];
    uint16_t* sk = (uint16_t*)lds_w;     // sorted blocks, local block lb = global - b0 + 2
    uint32_t* side = lds_w + MAIN / 4;  // [4][64]: blocks b0-2, b0-1, b0+SB, b0+SB+1
    __shared__ float sfloor;
    pad_key_blocks<STRIDE>(sk, SB_BLOCKS + 6, threadIdx.x, 256);
    const uint2 wb = xcd_block();
    const uint32_t trial = wb.y;
    const uint32_t ngroups = m1 / cw;
    const uint32_t nsorted = (m + 63) / 64;
    const uint32_t cpb = m2 / 64;  // blocks per column
    const size_t srow = (size_t)trial * spec_pitch;
    const uint32_t col = threadIdx.x % cw, ch = threadIdx.x / cw;
    const uint32_t B = ch + cpb * col;  // local natural block
    const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    // this lane's block of group gi: column gi cw + col, rows 64 ch .. 64 ch + 63
    // (one register per key while the loads are in flight: packing them at once
    // would wait for them)
    auto load_group = [&](uint32_t gi, uint32_t (&r)[64]) {
        const uint16_t* src = spec_t + srow + (size_t)(64 * ch) * m1 + gi * cw + col;
#pragma unroll
        for (int i = 0; i < 64; i++) r[i] = src[(size_t)i * m1];
    };
    if (threadIdx.x == 0) {
        double S = 0.0;
        for (uint32_t i = 0; i < npart; i++) S += partial[(size_t)trial * npart + i];
        sfloor = (float)(S / (double)nbins * 1e-6);
    }
    uint32_t pre[64];
    load_group(wb.x * WH_GROUPS, pre);
    for (uint32_t it = 0; it < WH_GROUPS; it++) {
        const uint32_t gi = wb.x * WH_GROUPS + it;
        if (gi >= ngroups) break;
        u16x2 w[32];
#pragma unroll
        for (int i = 0; i < 32; i++) w[i] = u16x2{(unsigned short)pre[i], (unsigned short)pre[i + 32]};
        const uint32_t k1a = gi * cw, k_lo = k1a * m2, b0 = k_lo / 64;
        {  // side blocks: thread t takes element t % 64 of side block t / 64
            const uint32_t sb = threadIdx.x >> 6, i = threadIdx.x & 63;
            const int64_t gb = sb < 2 ? (int64_t)b0 - 2 + sb : (int64_t)b0 + SB_BLOCKS + (sb - 2);
            uint32_t key = KEY_INF;
            if (gb >= 0 && gb < (int64_t)nsorted) {
                const uint32_t K = (uint32_t)gb * 64 + i;
                if (K > 0 && K < m) key = spec_t[srow + (size_t)(K % m2) * m1 + K / m2];
            } else if (gb == (int64_t)nsorted && gb < (int64_t)nblocks && i == 0) {
                key = spec_t[srow + m];  // block nsorted holds Nyquist alone
            }
            side[threadIdx.x] = key;
        }
        {  // natural-order raw spectrum: keys 8c .. 8c + 7 of the block
            uint16_t* o = raw_nat + srow + (size_t)(b0 + B) * 64;
#pragma unroll
            for (int c = 0; c < 8; c++) {
                const int h = c >> 2, i0 = 8 * (c & 3);
                uint32_t q[4];
#pragma unroll
                for (int e = 0; e < 4; e++)
                    q[e] = h ? (uint32_t)w[i0 + 2 * e].y | (uint32_t)w[i0 + 2 * e + 1].y << 16
                             : (uint32_t)w[i0 + 2 * e].x | (uint32_t)w[i0 + 2 * e + 1].x << 16;
                *(uint4*)(o + 8 * c) = make_uint4(q[0], q[1], q[2], q[3]);
            }
            if (k1a + cw == m1 && threadIdx.x == 0) raw_nat[srow + m] = spec_t[srow + m];  // Nyquist
        }
        if (k_lo == 0 && B == 0) w[0].x = 0xffffu;  // DC is not part of any window
        u16x2 v[32];
#pragma unroll
        for (int i = 0; i < 32; i++) v[i] = w[i];
        if (it + 1 < WH_GROUPS && gi + 1 < ngroups) load_group(gi + 1, pre);  // in flight from here
        sort64_packed(v);
#pragma unroll
        for (int i = 0; i < 32; i++) {
            sk[key_block<uint16_t, STRIDE>(B + 2) + i] = v[i].x;
            sk[key_block<uint16_t, STRIDE>(B + 2) + i + 32] = v[i].y;
        }
        __syncthreads();  // side keys and the floor are in LDS
        // side block w (waves 0..3): bitonic network across the wave's lanes
        uint32_t x = side[threadIdx.x];
#pragma unroll
        for (int k = 2; k <= 64; k <<= 1)
#pragma unroll
            for (int j = k >> 1; j > 0; j >>= 1) {
                const uint32_t pv = (uint32_t)__shfl_xor((int)x, j);
                const bool up = (lane & k) == 0, lower = (lane & j) == 0;
                x = lower == up ? min(x, pv) : max(x, pv);
            }
        sk[key_block<uint16_t, STRIDE>(wave < 2 ? wave : SB_BLOCKS + wave) + lane] = (uint16_t)x;
        if (threadIdx.x < 2 * STRIDE)  // local blocks SB + 4 and SB + 5 (contiguous): never counted, kept defined
            sk[key_block<uint16_t, STRIDE>(SB_BLOCKS + 4) + threadIdx.x] = (uint16_t)KEY_INF;
        __syncthreads();
        const float fl = sfloor;
        const uint32_t bm = threadIdx.x;  // medians in natural block order
        scales[(size_t)trial * nblocks + b0 + bm] =
            window_scale<true, uint16_t, STRIDE>(sk, bm, b0 + bm, nblocks, nbins, fl);
        if (bm == 0 && k1a + cw == m1 && nsorted < nblocks)  // the Nyquist block
            scales[(size_t)trial * nblocks + nsorted] =
                window_scale<true, uint16_t, STRIDE>(sk, SB_BLOCKS, nsorted, nblocks, nbins, fl);
        __syncthreads();  // the next group rewrites the sorted blocks and side keys
    }
}

static uint32_t sort_columns(const SpecGeom& g) {  // columns per sorting workgroup
    uint32_t cw = g.m1;
    while (cw > 1 && (uint64_t)cw * g.m2 > (uint64_t)SB_BLOCKS * 64) cw >>= 1;
    return cw;
}
bool whiten_fused(const SpecGeom& g) {
    const uint32_t cw = sort_columns(g);
    return g.m1 > 1 && (uint64_t)cw * g.m2 == (uint64_t)SB_BLOCKS * 64 && g.m2 % 64 == 0;
}

void launch_whiten(hipStream_t s, const SpecGeom& g, const void* spec_t, int fp16,
                   uint32_t ntrials, const double* partial, uint32_t npartials, void* raw_nat,
                   void* sorted, float* floors, float* scales) {
    // columns per sorting workgroup: up to SB_BLOCKS natural blocks
    const uint32_t cw = sort_columns(g);
    const uint32_t nbk = (cw * g.m2 + 63) / 64;
    const size_t shm = (size_t)nbk * 64 * sizeof(uint32_t);
    const dim3 gs(g.m1 / cw, ntrials), gm((g.nblocks + MB_BLOCKS - 1) / MB_BLOCKS, ntrials);
    if (whiten_fused(g)) {
        if (fp16) {
            constexpr size_t shm = wf_lds_bytes<true>();
            if (shm > 64 * 1024)
                (void)hipFuncSetAttribute(reinterpret_cast<const void*>(k_whiten_fused16),
                                          hipFuncAttributeMaxDynamicSharedMemorySize, (int)shm);
            const dim3 gh((g.m1 / cw + WH_GROUPS - 1) / WH_GROUPS, ntrials);
            hipLaunchKernelGGL(k_whiten_fused16, gh, dim3(256), shm, s, (const uint16_t*)spec_t, g.m, g.m1,
                               g.m2, cw, g.nbins, g.nblocks, g.spec_pitch, partial, npartials,
                               (uint16_t*)raw_nat, scales);
        } else {
            constexpr size_t shm = wf_lds_bytes<false>();
            if (shm > 64 * 1024)
                (void)hipFuncSetAttribute(reinterpret_cast<const void*>(k_whiten_fused<false>),
                                          hipFuncAttributeMaxDynamicSharedMemorySize, (int)shm);
            hipLaunchKernelGGL(k_whiten_fused<false>, gs, dim3(256), shm, s, spec_t, g.m, g.m1, g.m2, cw,
                               g.nbins, g.nblocks, g.spec_pitch, partial, npartials, raw_nat, scales);
        }
        return;
    }
    if (fp16) {
        hipLaunchKernelGGL(k_sort_blocks<true>, gs, dim3(256), shm, s, spec_t, g.m, g.m1, g.m2, cw,
                           g.nbins, g.spec_pitch, partial, npartials, raw_nat, sorted, floors);
        hipLaunchKernelGGL(k_block_median<true>, gm, dim3(MB_BLOCKS), 0, s, sorted, spec_t, g.m,
                           g.nbins, g.nblocks, g.spec_pitch, floors, scales);
    } else {
        hipLaunchKernelGGL(k_sort_blocks<false>, gs, dim3(256), shm, s, spec_t, g.m, g.m1, g.m2, cw,
                           g.nbins, g.spec_pitch, partial, npartials, raw_nat, sorted, floors);
        hipLaunchKernelGGL(k_block_median<false>, gm, dim3(MB_BLOCKS), 0, s, sorted, spec_t, g.m,
                           g.nbins, g.nblocks, g.spec_pitch, floors, scales);
    }
}

// ----------------------------------------------------- harmonic summing
// raw spectrum word (FP16 amplitude bits or FP32 power bits) and its power;
// loads and conversions are kept apart so that several loads can be in flight
template <bool F16>
__device__ __forceinline__ uint32_t raw_word(const void* raw, size_t idx) {
    if constexpr (F16) return ((const uint16_t*)raw)[idx];
    else return ((const uint32_t*)raw)[idx];
}
template <bool F16>
__device__ __forceinline__ float raw_power(uint32_t w) {
    if constexpr (F16) {
        const float a = __half2float(__ushort_as_half((unsigned short)w));
        return a * a;
    } else {
        return __uint_as_float(w);
    }
}

struct HarmArgs {
    SearchParams p;
    double sqrt_h[5];
    uint32_t birdie[2 * MAX_BIRDIE_INTERVALS];  // inclusive bin intervals
};

// Whitened power from a raw value and its block scale: 0 at DC and in birdie
// bins, else raw / scale (f32 division).
__device__ __forceinline__ float whiten_val(float r, float scale, uint32_t i, const HarmArgs& a) {
    if (i == 0) return 0.0f;
    for (uint32_t b = 0; b < a.p.nbirdie; b++)
        if (i >= a.birdie[2 * b] && i <= a.birdie[2 * b + 1]) return 0.0f;
    return r / scale;
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
constexpr int HS_NJ = 15;         // harmonics j = 2 .. 16
constexpr int HS_LIST = 64;       // listed bins summed per round
constexpr int HS_TILES = 2;       // tiles of HS_OUT bins per workgroup, the next one's inputs prefetched

// One workgroup scores HS_TILES tiles of HS_T consecutive bins in turn (tiles
// overlap by two bins, the peak test's neighbours), loading the next tile's
// level-0 inputs while the current one is scored.  Level 0 (j = 1) is computed for
// every bin; higher levels only for bins whose whitened power reaches 4
// (lib.rs:97-99), typically a few percent.  Those bins are compacted into a
// list; HS_LIST of them at a time, their (bin, harmonic) loads are spread over
// all threads (one round of memory latency instead of fifteen), and one thread
// per listed bin then adds its harmonics in order j = 2, 3, ... in f64, as the
// reference does.  The small list keeps the LDS footprint low enough for eight
// workgroups per CU.
// MULTI: the trials of one launch come from several observations; trial t
// takes the bin limits, levels and birdies of jobs[trial_job[t]] (a otherwise).
template <bool F16, bool MULTI>
__global__ __launch_bounds__(256) void k_harmsum(const void* __restrict__ raw,
                                                 uint64_t spec_pitch, uint32_t nbins,
                                                 uint32_t nblocks,
                                                 const float* __restrict__ scales, HarmArgs a,
                                                 RawPeak* __restrict__ peaks,
                                                 uint32_t* __restrict__ peak_counts, uint32_t ntiles_all,
                                                 const HarmArgs* __restrict__ jobs,
                                                 const uint32_t* __restrict__ trial_job) {
    __shared__ double sc[5][HS_T];
    __shared__ double lpos[HS_T];   // k + offset of listed bins
    __shared__ double lsum[HS_T];   // their j = 1 term
    __shared__ uint8_t lbin[HS_T];  // their thread index
    __shared__ float hv[HS_LIST][HS_NJ + 2];  // odd row stride: conflict-free column reads
    __shared__ uint32_t nlist;
    const uint2 wb = xcd_block();  // neighbouring bins' harmonics share lines on one XCD
    const uint32_t trial = wb.y;
    const HarmArgs& A = MULTI ? jobs[trial_job[trial]] : a;
    uint32_t ntiles = ntiles_all;
    if constexpr (MULTI) {  // this trial's own tile count (uniform: the workgroup leaves together)
        if (A.p.nlevels == 0 || A.p.end_h[0] <= A.p.first + 1) return;
        ntiles = (A.p.end_h[0] - A.p.first + 1 + HS_OUT - 1) / HS_OUT;
        if (wb.x * HS_TILES >= ntiles) return;
    }
    const size_t row = (size_t)trial * spec_pitch;
    const float* scl = scales + (size_t)trial * nblocks;
    const int tid = threadIdx.x;
    const double NEG_INF = -__builtin_inf();
    // this thread's bin of tile `tile` and its level-0 inputs, loaded one tile
    // ahead: raw k-1 .. k+1 (one 8-byte load of four FP16 words, or three FP32
    // words; k clamped into range so that the loads need no branch and stay in
    // flight) and the scales of k and k + 1
    struct In {
        uint32_t r0, r1, r2;
        float s1, s2;
    };
    auto bin_of = [&](uint32_t tile) { return (int64_t)A.p.first - 1 + (int64_t)tile * HS_OUT + tid; };
    auto load_in = [&](uint32_t tile) {
        In v;
        const int64_t k = bin_of(tile);
        const uint32_t ku = (uint32_t)min(max(k, (int64_t)A.p.first), (int64_t)A.p.end_h[0]);
        if constexpr (F16) {
            // words ku-1 .. ku+2; ku + 2 <= nbins < spec_pitch
            uint64_t q;
            __builtin_memcpy(&q, (const uint16_t*)raw + row + ku - 1, 8);
            v.r0 = (uint32_t)q;
            v.r1 = (uint32_t)(q >> 16);
            v.r2 = (uint32_t)(q >> 32);
        } else {
            const uint32_t* r = (const uint32_t*)raw + row + ku;
            v.r0 = r[-1];
            v.r1 = r[0];
            v.r2 = r[1];
        }
        v.s1 = scl[ku >> 6];
        v.s2 = scl[(ku + 1) >> 6];
        return v;
    };
    In cur = load_in(wb.x * HS_TILES);
    for (uint32_t it = 0; it < (uint32_t)HS_TILES; it++) {
        const uint32_t tile = wb.x * HS_TILES + it;
        if (tile >= ntiles) break;
        const In nxt = load_in(tile + 1);  // in flight while this tile is scored
        const int64_t k = bin_of(tile);
        double offset = 0.0;
#pragma unroll
        for (int l = 0; l < 5; l++) sc[l][tid] = NEG_INF;
        if (tid == 0) nlist = 0;
        __syncthreads();
        bool listed = false;
        double S1 = 0.0;
        if (k >= (int64_t)A.p.first && k <= (int64_t)A.p.end_h[0]) {
            const uint32_t ku = (uint32_t)k;
            // (raw_power<true> reads the low 16 bits of the word)
            const float p0 = raw_power<F16>(cur.r0), p1 = raw_power<F16>(cur.r1), p2 = raw_power<F16>(cur.r2);
            offset = interp_offset((double)p0, (double)p1, (double)p2);
            const float pwk = whiten_val(p1, cur.s1, ku, A);
            // j = 1 sums the bin at round(k + offset): offset is in [-0.5, 0.5] and
            // round() takes halves away from zero, so that is k or k + 1 (<= nbins - 1),
            // and k + 1 exactly when the rounded sum reaches k + 0.5
            const bool up = (double)ku + offset >= (double)ku + 0.5;
            const float w1 = up ? whiten_val(p2, cur.s2, ku + 1, A) : pwk;
            S1 = (double)w1;
            sc[0][tid] = S1 - 1.0;  // / sqrt(1): exact
            listed = A.p.nlevels > 1 && k <= (int64_t)A.p.end_h[1] && pwk >= 4.0f;
        }
        if (listed) {
            const uint32_t slot = atomicAdd(&nlist, 1u);
            lpos[slot] = (double)k + offset;
            lsum[slot] = S1;
            lbin[slot] = (uint8_t)tid;
        }
        __syncthreads();
        const uint32_t G = nlist;
        for (uint32_t g0 = 0; g0 < G; g0 += HS_LIST) {
            const uint32_t ng = min((uint32_t)HS_LIST, G - g0);
            for (uint32_t q = tid; q < ng * HS_NJ; q += HS_T) {
                const uint32_t g = q / HS_NJ, jj = q - g * HS_NJ;
                const double position = lpos[g0 + g] * (double)(jj + 2);
                // round(position) >= 0 clamped to nbins - 1 (any u64 clamp of the
                // rounded value gives the same bin)
                const uint32_t iu = (uint32_t)fmin(round(position), (double)(nbins - 1));
                const uint32_t w = raw_word<F16>(raw, row + iu);
                hv[g][jj] = whiten_val(raw_power<F16>(w), scl[iu >> 6], iu, A);
            }
            __syncthreads();
            if ((uint32_t)tid < ng) {
                const uint32_t t = lbin[g0 + tid];
                const int64_t kk = (int64_t)A.p.first - 1 + (int64_t)tile * HS_OUT + t;
                float v[HS_NJ];
#pragma unroll
                for (int j = 0; j < HS_NJ; j++) v[j] = hv[tid][j];
                // j = 2, 3, ... added in order; level l (h = 2^l) is scored after
                // j = h while k is within its end bin (end bins fall with l)
                double S = lsum[g0 + tid];
#pragma unroll
                for (int l = 1; l < 5; l++) {
                    const int h = 1 << l;
#pragma unroll
                    for (int j = (h >> 1) + 1; j <= h; j++) S += (double)v[j - 2];
                    if ((uint32_t)l < A.p.nlevels && kk <= (int64_t)A.p.end_h[l])
                        // sqrt(4) and sqrt(16) are powers of two: the product is the exact quotient
                        sc[l][t] = (l & 1) ? (S - (double)h) / A.sqrt_h[l] : ldexp(S - (double)h, -(l >> 1));
                }
            }
            __syncthreads();
        }
        if (tid >= 1 && tid <= HS_OUT && k >= (int64_t)A.p.first + 1) {
            double sl[5];
#pragma unroll
            for (int l = 0; l < 5; l++) sl[l] = sc[l][tid];
#pragma unroll
            for (int l = 0; l < 5; l++) {
                const double s = sl[l];
                if ((uint32_t)l < A.p.nlevels && k <= (int64_t)A.p.end_h[l] - 1 && s >= A.p.threshold &&
                    s > sc[l][tid - 1] && s >= sc[l][tid + 1]) {
                    const uint32_t slot = atomicAdd(&peak_counts[trial], 1u);
                    if (slot < A.p.peak_cap) {
                        RawPeak rp;
                        rp.score = s;
                        rp.offset = offset;
                        rp.k = (uint32_t)k;
                        rp.h = 1u << l;
                        peaks[(size_t)trial * A.p.peak_cap + slot] = rp;
                    }
                }
            }
        }
        __syncthreads();  // the next tile rewrites the scores
        cur = nxt;
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
    const uint32_t ntiles = (span + HS_OUT - 1) / HS_OUT;
    dim3 grid((ntiles + HS_TILES - 1) / HS_TILES, ntrials);
    if (p.fp16)
        hipLaunchKernelGGL((k_harmsum<true, false>), grid, dim3(HS_T), 0, s, raw_nat, g.spec_pitch,
                           g.nbins, g.nblocks, scales, a, peaks, peak_counts, ntiles, nullptr, nullptr);
    else
        hipLaunchKernelGGL((k_harmsum<false, false>), grid, dim3(HS_T), 0, s, raw_nat, g.spec_pitch,
                           g.nbins, g.nblocks, scales, a, peaks, peak_counts, ntiles, nullptr, nullptr);
}

size_t harm_args_bytes() { return sizeof(HarmArgs); }

void fill_harm_args(const SearchParams& p, const uint32_t* birdies, void* out) {
    HarmArgs a;
    std::memset(&a, 0, sizeof(a));
    a.p = p;
    for (int i = 0; i < 2 * MAX_BIRDIE_INTERVALS; i++) a.birdie[i] = i < (int)(2 * p.nbirdie) ? birdies[i] : 0u;
    static const double SQ[5] = {1.0, 1.4142135623730951, 2.0, 2.8284271247461903, 4.0};
    for (int l = 0; l < 5; l++) a.sqrt_h[l] = SQ[l];
    std::memcpy(out, &a, sizeof(a));
}

void launch_harmsum_multi(hipStream_t s, const SpecGeom& g, int fp16, uint32_t max_tiles, const void* raw_nat,
                          const float* scales, const void* jobs, const uint32_t* trial_job, uint32_t ntrials,
                          RawPeak* peaks, uint32_t* peak_counts) {
    if (max_tiles == 0) return;
    HarmArgs none;
    std::memset(&none, 0, sizeof(none));
    dim3 grid((max_tiles + HS_TILES - 1) / HS_TILES, ntrials);
    if (fp16)
        hipLaunchKernelGGL((k_harmsum<true, true>), grid, dim3(HS_T), 0, s, raw_nat, g.spec_pitch, g.nbins,
                           g.nblocks, scales, none, peaks, peak_counts, max_tiles, (const HarmArgs*)jobs,
                           trial_job);
    else
        hipLaunchKernelGGL((k_harmsum<false, true>), grid, dim3(HS_T), 0, s, raw_nat, g.spec_pitch, g.nbins,
                           g.nblocks, scales, none, peaks, peak_counts, max_tiles, (const HarmArgs*)jobs,
                           trial_job);
}

uint32_t harmsum_tiles(const SearchParams& p) {
    if (p.nlevels == 0 || p.end_h[0] <= p.first + 1) return 0;
    return (p.end_h[0] - p.first + 1 + HS_OUT - 1) / HS_OUT;
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
                                                  uint32_t* __restrict__ overflow,
                                                  const HarmArgs* __restrict__ jobs,
                                                  const uint32_t* __restrict__ trial_job) {
    __shared__ double s_score[PEAK_CAP];
    __shared__ uint32_t s_hk[PEAK_CAP];
    __shared__ uint16_t s_idx[PEAK_CAP];
    __shared__ double s_kept[256];
    __shared__ uint32_t s_nk;
    const uint32_t trial = blockIdx.x;
    if (jobs) {  // several observations: this trial's own limits
        const SearchParams& q = jobs[trial_job[trial]].p;
        cap = q.peak_cap;
        bin_hz = q.bin_hz;
        resolution = q.resolution;
        max_candidates = q.max_candidates;
    }
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

__global__ void k_fill_u32(uint32_t* __restrict__ p, size_t n, uint32_t v) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

void launch_fill_u32(hipStream_t s, uint32_t* p, size_t n, uint32_t v) {
    if (n == 0) return;
    hipLaunchKernelGGL(k_fill_u32, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, p, n, v);
}

void launch_finalize(hipStream_t s, const SpecGeom& g, const SearchParams& p, const double* dms,
                     uint32_t ntrials, const RawPeak* peaks, const uint32_t* peak_counts,
                     CandidateOut* out, uint32_t cap_per_trial, uint32_t* counts) {
    (void)g;
    // the overflow counter lives right after the per-trial counts
    hipLaunchKernelGGL(k_finalize, dim3(ntrials), dim3(256), 0, s, peaks, peak_counts, p.peak_cap,
                       p.bin_hz, p.resolution, p.max_candidates, dms, out, cap_per_trial, counts,
                       (uint32_t*)(counts + ntrials), (const HarmArgs*)nullptr, (const uint32_t*)nullptr);
}

void launch_finalize_multi(hipStream_t s, const double* dms, uint32_t ntrials, const RawPeak* peaks,
                           const uint32_t* peak_counts, const void* jobs, const uint32_t* trial_job,
                           CandidateOut* out, uint32_t cap_per_trial, uint32_t* counts) {
    hipLaunchKernelGGL(k_finalize, dim3(ntrials), dim3(256), 0, s, peaks, peak_counts, PEAK_CAP, 0.0, 0.0, 0u,
                       dms, out, cap_per_trial, counts, (uint32_t*)(counts + ntrials), (const HarmArgs*)jobs,
                       trial_job);
}

}  // namespace psr
