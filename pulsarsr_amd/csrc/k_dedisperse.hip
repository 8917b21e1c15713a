// Ingest, channel statistics, delay table and incoherent dedispersion kernels
// for gfx950.  Dedispersion restates crates/dedisperse/src/lib.rs:62-90:
//   out[t] = sum_c data[c][t + delay_c]
// For samples of at most 8 bits the channel sums are integers below 2^24, so
// they are exact (bit-identical to the reference's channel-ordered f32 sums) in
// any order; the kernel accumulates them with packed 16-bit adds and widens to
// 32 bits before any half-word can overflow.  16-bit and float samples use a
// kernel that adds in channel order in f32, exactly as the reference does.
#include <algorithm>

#include <hip/hip_fp16.h>
#include <hip/hip_runtime.h>

#include "common.hpp"

namespace psr {

typedef unsigned short u16x2 __attribute__((ext_vector_type(2)));

// ---------------------------------------------------------------- ingest
// Packed time-major SIGPROC frames -> channel-major rows (u8 for <=8 bits,
// u16 for 16 bits, f32 for 32 bits).  Sample i = t*nchans + ch occupies bits
// [i*nbits, (i+1)*nbits), LSB-first within a byte (crates/fbank/src/lib.rs:175-182).
// One workgroup transposes a 64-frame x 64-channel tile through LDS.
__global__ __launch_bounds__(256) void k_ingest(const uint8_t* __restrict__ payload, int nbits,
                                                uint32_t nchans, uint64_t nsamples,
                                                void* __restrict__ out, uint64_t pitch, int kind) {
    __shared__ uint32_t tile[64][65];
    const uint64_t t0 = (uint64_t)blockIdx.x * 64;
    const uint32_t c0 = blockIdx.y * 64;
    const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
    // read: lane tx walks channels of frame t0 + r
    for (int r = ty; r < 64; r += 4) {
        const uint64_t t = t0 + r;
        const uint32_t c = c0 + tx;
        uint32_t v = 0;
        if (t < nsamples && c < nchans) {
            const uint64_t i = t * nchans + c;
            if (nbits <= 8) {
                const uint64_t bit = i * (uint64_t)nbits;
                v = (payload[bit >> 3] >> (bit & 7)) & ((1u << nbits) - 1u);
            } else if (nbits == 16) {
                v = (uint32_t)payload[2 * i] | ((uint32_t)payload[2 * i + 1] << 8);
            } else {
                v = (uint32_t)payload[4 * i] | ((uint32_t)payload[4 * i + 1] << 8) |
                    ((uint32_t)payload[4 * i + 2] << 16) | ((uint32_t)payload[4 * i + 3] << 24);
            }
        }
        tile[r][tx] = v;
    }
    __syncthreads();
    // write: lane tx walks time of channel c0 + r
    for (int r = ty; r < 64; r += 4) {
        const uint32_t c = c0 + r;
        const uint64_t t = t0 + tx;
        if (c >= nchans || t >= nsamples) continue;
        const uint32_t v = tile[tx][r];
        if (kind == KIND_U8)
            ((uint8_t*)out)[c * pitch + t] = (uint8_t)v;
        else if (kind == KIND_U16)
            ((uint16_t*)out)[c * pitch + t] = (uint16_t)v;
        else
            ((uint32_t*)out)[c * pitch + t] = v;
    }
}

void launch_ingest(hipStream_t s, const uint8_t* payload, int nbits, uint32_t nchans,
                   uint64_t nsamples, void* out, uint64_t pitch, int kind) {
    dim3 grid((unsigned)((nsamples + 63) / 64), (nchans + 63) / 64);
    hipLaunchKernelGGL(k_ingest, grid, dim3(256), 0, s, payload, nbits, nchans, nsamples, out,
                       pitch, kind);
}

// ------------------------------------------------------- channel statistics
// Per-channel sum and sum of squares in f64, reduced in a fixed tree order so
// the result is deterministic.  Used only to choose the FP16 storage offset
// and scale; it does not enter any exact result.
__global__ __launch_bounds__(256) void k_chan_stats(const void* __restrict__ data, int kind,
                                                    uint64_t nsamples, uint64_t pitch,
                                                    double* __restrict__ sums,
                                                    double* __restrict__ sumsq) {
    const uint32_t c = blockIdx.x;
    double s = 0.0, q = 0.0;
    for (uint64_t t = threadIdx.x; t < nsamples; t += blockDim.x) {
        double v;
        if (kind == KIND_U8)
            v = ((const uint8_t*)data)[c * pitch + t];
        else if (kind == KIND_U16)
            v = ((const uint16_t*)data)[c * pitch + t];
        else
            v = ((const float*)data)[c * pitch + t];
        s += v;
        q += v * v;
    }
    __shared__ double rs[256], rq[256];
    rs[threadIdx.x] = s;
    rq[threadIdx.x] = q;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
            rs[threadIdx.x] += rs[threadIdx.x + w];
            rq[threadIdx.x] += rq[threadIdx.x + w];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        sums[c] = rs[0];
        sumsq[c] = rq[0];
    }
}

void launch_chan_stats(hipStream_t s, const void* data, int kind, uint32_t nchans,
                       uint64_t nsamples, uint64_t pitch, double* sums, double* sumsq) {
    hipLaunchKernelGGL(k_chan_stats, dim3(nchans), dim3(256), 0, s, data, kind, nsamples, pitch,
                       sums, sumsq);
}

// ------------------------------------------------------------ delay table
// crates/dedisperse/src/lib.rs:21-33 in f64 without contraction:
//   round(((4148.808 * dm) * (1/f^2 - 1/fmax^2)) / tsamp), clamped at 0.
// inv_f2[c] = 1/(f*f) and ref_inv = 1/(fmax*fmax) are computed on the host the
// same way, so the table is bit-identical to the host's delays().
__global__ __launch_bounds__(256) void k_delays(const double* __restrict__ inv_f2,
                                                uint32_t nchans, double ref_inv, double tsamp,
                                                const double* __restrict__ dms, uint32_t ntrials,
                                                int32_t* __restrict__ out) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (uint64_t)nchans * ntrials) return;
    const uint32_t d = (uint32_t)(i / nchans), c = (uint32_t)(i % nchans);
    const double k_dm = 4148.808 * dms[d];
    const double samples = round(k_dm * (inv_f2[c] - ref_inv) / tsamp);
    out[i] = (int32_t)fmax(samples, 0.0);
}

void launch_delays(hipStream_t s, const double* inv_f2, uint32_t nchans, double ref_inv,
                   double tsamp, const double* dms, uint32_t ntrials, int32_t* out) {
    const uint64_t total = (uint64_t)nchans * ntrials;
    hipLaunchKernelGGL(k_delays, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s, inv_f2,
                       nchans, ref_inv, tsamp, dms, ntrials, out);
}

// ------------------------------------------------- integer dedispersion (u8)
// Workgroup = 8 waves sharing one time tile of DD_TILE = 512 outputs and a block
// of DD_TRIALS = 64 consecutive (ascending) DM trials; each wave owns DD_M = 8
// trials.  Channels are staged chunk by chunk into LDS as 32-bit words
//     w[i] = s[base + i] | s[base + i + 256] << 16,
// with base = the block's smallest delay of the channel rounded down to 4.  A
// trial whose delay is base + e needs words e .. e + 255; lane l takes words
// e + 4l .. e + 4l + 3 with one 16-byte ds_read_b128 and adds them with four
// v_pk_add_u16 (8 outputs: 4l + j and 256 + 4l + j).  ds_read_b128 needs a
// 16-byte aligned address, so every channel row is staged four times, copy k
// starting k words later; the trial reads copy e & 3 at word e & ~3.  The 64
// lanes of a wave read 1 KiB of consecutive LDS, free of bank conflicts.
//
// Per-trial LDS byte offsets (16 bits each) are staged with the rows; a wave
// reads its 8 offsets of a channel with one 16-byte broadcast read, so no scalar
// load sits among the LDS reads and the waits stay fine-grained.  16-bit accumulators widen to
// 32 bits every `phase_chans` channels, before a half-word can overflow.
//
// Workgroups are numbered XCD-aware: consecutive workgroup ids go to different
// XCDs, so id is remapped so that each XCD walks a contiguous run of
// (tile, block) work items with the block index fastest -- the channel data of
// a tile is then fetched into that XCD's L2 once and reused by all DM blocks.
constexpr int DD_M = 8;
constexpr int DD_WAVES = 8;
constexpr int DD_THREADS = 64 * DD_WAVES;
constexpr int DD_TILE = 512;
constexpr int DD_TRIALS = DD_WAVES * DD_M;
static_assert(DD_TRIALS == (int)DEDISP_BLOCK_TRIALS, "block size shared with the host");
constexpr int DD_XCDS = 8;

__device__ __forceinline__ uint32_t pk_add(uint32_t a, uint32_t b) {
    u16x2 x = __builtin_bit_cast(u16x2, a), y = __builtin_bit_cast(u16x2, b);
    return __builtin_bit_cast(uint32_t, x + y);
}

// bases[blk][c] = delay of the block's first trial at channel c, rounded down to
// a multiple of 4; rel[blk][c][m] = delay of trial blk*32 + m minus that base
// (trials past the end repeat the last trial).
__global__ __launch_bounds__(256) void k_dd_table(const int32_t* __restrict__ delays,
                                                  uint32_t nchans, uint32_t ntrials,
                                                  int32_t* __restrict__ bases,
                                                  int32_t* __restrict__ rel) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t nblk = (ntrials + DD_TRIALS - 1) / DD_TRIALS;
    if (i >= (uint64_t)nblk * nchans) return;
    const uint32_t blk = (uint32_t)(i / nchans), c = (uint32_t)(i % nchans);
    const uint32_t d0 = blk * DD_TRIALS;
    const int32_t base = delays[(size_t)d0 * nchans + c] & ~3;
    bases[i] = base;
    for (int m = 0; m < DD_TRIALS; m++) {
        const uint32_t d = min(d0 + m, ntrials - 1);
        rel[i * DD_TRIALS + m] = delays[(size_t)d * nchans + c] - base;
    }
}

__device__ __forceinline__ uint32_t load_u32(const uint8_t* row, uint64_t off, uint64_t pitch) {
    return off + 4 <= pitch ? *(const uint32_t*)(row + off) : 0u;
}

// pair word of byte i (0..3) of windows a (low half) and b (high half)
__device__ __forceinline__ uint32_t pair_word(uint32_t b, uint32_t a, int i) {
    return __builtin_amdgcn_perm(b, a, 0x0c040c00u + 0x00010001u * (uint32_t)i);
}

__global__ __launch_bounds__(DD_THREADS) void k_dedisp_u8(
    const uint8_t* __restrict__ data, uint64_t pitch, uint32_t nchans, uint32_t phase_chans,
    const int32_t* __restrict__ bases, const int32_t* __restrict__ rel, uint32_t ntrials,
    uint32_t nblk, uint32_t ntiles, uint32_t n, void* __restrict__ out, uint32_t out_pitch,
    int fp16, float m0, float inv_s0, double* __restrict__ partial, uint32_t W4, uint32_t C) {
    extern __shared__ __align__(16) uint32_t lds[];  // [C][4][W4] words, then [C][32] addresses
    // [C][DD_TRIALS] LDS byte offsets (< 64 KiB) of the runs each trial reads
    uint16_t* laddr = (uint16_t*)(lds + (size_t)C * 4 * W4);

    // XCD-aware work index (bijective over total = ntiles * nblk)
    const uint32_t total = ntiles * nblk;
    const uint32_t id = blockIdx.x;
    const uint32_t xcd = id % DD_XCDS, local = id / DD_XCDS;
    const uint32_t q = total / DD_XCDS, r = total % DD_XCDS;
    const uint32_t work = (xcd < r ? xcd * (q + 1) : r * (q + 1) + (xcd - r) * q) + local;
    const uint32_t tile = work / nblk, blk = work % nblk;

    const uint32_t t0 = tile * DD_TILE;
    const uint32_t d0 = blk * DD_TRIALS;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    const int32_t* bblk = bases + (size_t)blk * nchans;
    const int32_t* rblk = rel + (size_t)blk * nchans * DD_TRIALS;

    uint32_t acc16[DD_M][4];
    uint32_t acc32[DD_M][8];
#pragma unroll
    for (int m = 0; m < DD_M; m++) {
#pragma unroll
        for (int j = 0; j < 4; j++) acc16[m][j] = 0;
#pragma unroll
        for (int j = 0; j < 8; j++) acc32[m][j] = 0;
    }
    uint32_t in_phase = 0;
    const uint32_t quads = W4 / 4;
    const uint32_t c0 = threadIdx.x / quads, q0 = threadIdx.x % quads;
    const uint32_t sc = blockDim.x / quads, sq = blockDim.x % quads;
    const char* lds_lane = (const char*)lds + 16 * lane;
    for (uint32_t cc = 0; cc < nchans; cc += C) {
        const uint32_t nc = min(C, nchans - cc);
        __syncthreads();
        // item i = (channel c, quad qd), walked without division: (c0, q0) is the
        // thread's first item, each step adds (sc, sq)
        uint32_t c = c0, qd = q0;
        for (uint32_t i = threadIdx.x; i < nc * quads; i += blockDim.x) {
            const uint8_t* row = data + (size_t)(cc + c) * pitch;
            const uint64_t off = (uint64_t)t0 + (uint32_t)bblk[cc + c] + 4u * qd;
            const uint32_t a0 = load_u32(row, off, pitch), a1 = load_u32(row, off + 4, pitch);
            const uint32_t b0 = load_u32(row, off + 256, pitch), b1 = load_u32(row, off + 260, pitch);
            uint32_t* dst = lds + (size_t)c * 4 * W4 + 4 * qd;
            // w[j] pairs byte j of the 8-byte window at off with byte j of the window
            // at off + 256; copy k holds w[k .. k + 3]
            uint32_t w[7];
#pragma unroll
            for (int j = 0; j < 4; j++) w[j] = pair_word(b0, a0, j);
#pragma unroll
            for (int j = 0; j < 3; j++) w[4 + j] = pair_word(b1, a1, j);
#pragma unroll
            for (int k = 0; k < 4; k++)
                *(uint4*)(dst + (size_t)k * W4) = make_uint4(w[k], w[k + 1], w[k + 2], w[k + 3]);
            c += sc;
            qd += sq;
            if (qd >= quads) {
                qd -= quads;
                c++;
            }
        }
        for (uint32_t i = threadIdx.x; i < nc * DD_TRIALS; i += blockDim.x) {
            const uint32_t c = i / DD_TRIALS;
            const int32_t e = rblk[(size_t)cc * DD_TRIALS + i];
            laddr[i] = (uint16_t)((((uint32_t)c * 4 + (e & 3)) * W4 + (uint32_t)(e & ~3)) * 4);
        }
        __syncthreads();
#pragma unroll 2
        for (uint32_t c = 0; c < nc; c++) {
            // the wave's 8 offsets: one 16-byte broadcast read, 16 bits each
            const uint4 A = *(const uint4*)(laddr + c * DD_TRIALS + wave * DD_M);
            const uint32_t addr[DD_M] = {A.x & 0xffffu, A.x >> 16, A.y & 0xffffu, A.y >> 16,
                                         A.z & 0xffffu, A.z >> 16, A.w & 0xffffu, A.w >> 16};
#pragma unroll
            for (int m = 0; m < DD_M; m++) {
                const uint4 v = *(const uint4*)(lds_lane + addr[m]);
                acc16[m][0] = pk_add(acc16[m][0], v.x);
                acc16[m][1] = pk_add(acc16[m][1], v.y);
                acc16[m][2] = pk_add(acc16[m][2], v.z);
                acc16[m][3] = pk_add(acc16[m][3], v.w);
            }
            if (++in_phase == phase_chans) {
                in_phase = 0;
#pragma unroll
                for (int m = 0; m < DD_M; m++)
#pragma unroll
                    for (int j = 0; j < 4; j++) {
                        acc32[m][j] += acc16[m][j] & 0xffffu;
                        acc32[m][4 + j] += acc16[m][j] >> 16;
                        acc16[m][j] = 0;
                    }
            }
        }
    }
#pragma unroll
    for (int m = 0; m < DD_M; m++)
#pragma unroll
        for (int j = 0; j < 4; j++) {
            acc32[m][j] += acc16[m][j] & 0xffffu;
            acc32[m][4 + j] += acc16[m][j] >> 16;
        }
    const int m0i = (int)m0;
#pragma unroll
    for (int m = 0; m < DD_M; m++) {
        const uint32_t d = d0 + wave * DD_M + m;
        double lsum = 0.0;
        if (d < ntrials) {
#pragma unroll
            for (int hh = 0; hh < 2; hh++) {
                const uint32_t t = t0 + 256 * hh + 4 * lane;  // outputs t .. t + 3
                const uint32_t* v = &acc32[m][4 * hh];
                const size_t o = (size_t)d * out_pitch + t;
                if (t + 3 < n) {
                    lsum += (double)v[0] + (double)v[1] + (double)v[2] + (double)v[3];
                    if (fp16) {
                        __half2 h01 = __floats2half2_rn((float)((int)v[0] - m0i) * inv_s0,
                                                        (float)((int)v[1] - m0i) * inv_s0);
                        __half2 h23 = __floats2half2_rn((float)((int)v[2] - m0i) * inv_s0,
                                                        (float)((int)v[3] - m0i) * inv_s0);
                        uint2 pk;
                        pk.x = __builtin_bit_cast(uint32_t, h01);
                        pk.y = __builtin_bit_cast(uint32_t, h23);
                        *(uint2*)((__half*)out + o) = pk;
                    } else {
                        *(float4*)((float*)out + o) =
                            make_float4((float)v[0], (float)v[1], (float)v[2], (float)v[3]);
                    }
                } else {
#pragma unroll
                    for (int j = 0; j < 4; j++)
                        if (t + j < n) {
                            lsum += (double)v[j];
                            if (fp16)
                                ((__half*)out)[o + j] = __float2half_rn((float)((int)v[j] - m0i) * inv_s0);
                            else
                                ((float*)out)[o + j] = (float)v[j];
                        }
                }
            }
        }
        // wave reduction of the tile's sum (integers < 2^53: exact in any order)
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) lsum += __shfl_xor(lsum, o);
        if (lane == 0 && d < ntrials) partial[(size_t)d * ntiles + tile] = lsum;
    }
}

size_t dedisp_u8_table_words(uint32_t nchans, uint32_t ntrials) {
    const size_t nblk = (ntrials + DD_TRIALS - 1) / DD_TRIALS;
    return nblk * nchans * (1 + DD_TRIALS);
}

bool launch_dedisp_u8(hipStream_t s, const uint8_t* data, uint64_t pitch, uint32_t nchans,
                      uint32_t maxval, const int32_t* delays, int32_t* table, uint32_t ntrials,
                      uint32_t n, void* out, uint32_t out_pitch, int fp16, float m0, float inv_s0,
                      double* partial, uint32_t spread_max) {
    // a trial reads words (e & ~3) .. (e & ~3) + 255 with e <= spread_max + 3
    const uint64_t W4 = ((uint64_t)spread_max + 3 + 3) / 4 * 4 + DD_TILE / 2;
    // four row copies + 16-bit offsets per channel; the row copies must stay below
    // 64 KiB (16-bit offsets), which also lets two 512-thread workgroups share a CU
    const uint64_t per_chan = 16 * W4 + 2 * DD_TRIALS;
    const uint64_t budget = 64 * 1024;
    if (per_chan > budget) return false;
    uint32_t C = (uint32_t)std::min<uint64_t>(budget / per_chan, 64);
    C = std::min(C, nchans);
    const size_t nblk = (ntrials + DD_TRIALS - 1) / DD_TRIALS;
    int32_t* bases = table;
    int32_t* rel = table + nblk * nchans;
    hipLaunchKernelGGL(k_dd_table, dim3((unsigned)((nblk * nchans + 255) / 256)), dim3(256), 0, s,
                       delays, nchans, ntrials, bases, rel);
    uint32_t phase = 65535u / (maxval ? maxval : 1);
    phase = phase > 4096 ? 4096 : phase;
    const size_t shm = (size_t)C * per_chan;
    if (shm > 64 * 1024)
        (void)hipFuncSetAttribute(reinterpret_cast<const void*>(k_dedisp_u8),
                                  hipFuncAttributeMaxDynamicSharedMemorySize, (int)shm);
    const uint32_t ntiles = (n + DD_TILE - 1) / DD_TILE;
    const uint64_t total = (uint64_t)ntiles * nblk;
    hipLaunchKernelGGL(k_dedisp_u8, dim3((unsigned)total), dim3(DD_THREADS), shm, s, data, pitch, nchans,
                       phase, bases, rel, ntrials, (uint32_t)nblk, ntiles, n, out, out_pitch, fp16,
                       m0, inv_s0, partial, (uint32_t)W4, C);
    return true;
}

// ---------------------------------------- channel-ordered f32 dedispersion
// For 16-bit and float samples (and 8-bit ones whose delay spread is too wide
// for the LDS-staged kernel): one thread per output, f32 adds in channel
// order starting from 0.0 -- the reference's exact operation sequence.
__global__ __launch_bounds__(256) void k_dedisp_f32(const void* __restrict__ data, int kind,
                                                    uint64_t pitch, uint32_t nchans,
                                                    const int32_t* __restrict__ delays,
                                                    uint32_t ntrials, uint32_t n,
                                                    void* __restrict__ out, uint32_t out_pitch,
                                                    int fp16, float m0, float inv_s0,
                                                    double* __restrict__ partial) {
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t d = blockIdx.y;
    const int32_t* dd = delays + (size_t)d * nchans;
    float acc = 0.0f;
    if (t < n) {
        for (uint32_t c = 0; c < nchans; c++) {
            const uint64_t idx = (uint64_t)c * pitch + t + (uint32_t)dd[c];
            const float v = kind == KIND_U8    ? (float)((const uint8_t*)data)[idx]
                            : kind == KIND_U16 ? (float)((const uint16_t*)data)[idx]
                                               : ((const float*)data)[idx];
            acc += v;
        }
        if (fp16)
            ((__half*)out)[(size_t)d * out_pitch + t] = __float2half_rn((acc - m0) * inv_s0);
        else
            ((float*)out)[(size_t)d * out_pitch + t] = acc;
    }
    double lsum = t < n ? (double)acc : 0.0;
    for (int o = 32; o > 0; o >>= 1) lsum += __shfl_xor(lsum, o);
    __shared__ double ws[4];
    if ((threadIdx.x & 63) == 0) ws[threadIdx.x >> 6] = lsum;
    __syncthreads();
    if (threadIdx.x == 0) partial[(size_t)d * gridDim.x + blockIdx.x] = ((ws[0] + ws[1]) + ws[2]) + ws[3];
}

void launch_dedisp_f32(hipStream_t s, const void* data, int kind, uint64_t pitch,
                       uint32_t nchans, const int32_t* delays, uint32_t ntrials, uint32_t n,
                       void* out, uint32_t out_pitch, int fp16, float m0, float inv_s0,
                       double* partial) {
    dim3 grid((n + 255) / 256, ntrials);
    hipLaunchKernelGGL(k_dedisp_f32, grid, dim3(256), 0, s, data, kind, pitch, nchans, delays,
                       ntrials, n, out, out_pitch, fp16, m0, inv_s0, partial);
}

}  // namespace psr
