// Ingest, channel statistics, delay table and incoherent dedispersion kernels
// for gfx950.  Dedispersion restates crates/dedisperse/src/lib.rs:62-90:
//   out[t] = sum_c data[c][t + delay_c]
// For samples of at most 8 bits the channel sums are integers below 2^24, so
// they are exact (bit-identical to the reference's channel-ordered f32 sums) in
// any order; the kernel accumulates them with packed 16-bit adds and widens to
// 32 bits before any half-word can overflow.  16-bit and float samples use a
// kernel that adds in channel order in f32, exactly as the reference does.
#include <algorithm>
#include <vector>

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
// trials.  Channels are staged a few at a time into LDS as 32-bit words
//     w[i] = s[base + i] | s[base + i + 256] << 16,
// with base = the block's smallest delay of the channel rounded down to 4.  A
// trial whose delay is base + e needs words e .. e + 255; lane l takes words
// e + 4l .. e + 4l + 3 with one 16-byte ds_read_b128.  ds_read_b128 needs a
// 16-byte aligned address, so every channel row is staged four times, copy k
// starting k words later; the trial reads copy e & 3 at word e & ~3.  The 64
// lanes of a wave read 1 KiB of consecutive LDS, free of bank conflicts.  Rows
// are W4_c words long, from channel c's own delay spread over the block (low
// frequencies spread most), so a stage holds as many channels as fit.
//
// Channels are added in pairs with v_add3_u32 on the u16 pairs: within a
// "phase" of at most 65535 / maxval channels no half can carry, so the 32-bit
// add is two exact 16-bit adds.  At each phase end the halves are added into
// running sums mod 2^16 whose carries are counted in 4-bit fields, so totals up
// to 2^20 are exact.
//
// Staging is pipelined: while a stage is summed, each wave loads the raw bytes
// and trial offsets of one channel of the next stage straight into LDS (MUBUF
// loads with the LDS bit, no registers held); at the stage boundary the wave
// turns them into the four row copies and the per-trial LDS offsets (16 bits
// each; a wave reads its 8 offsets of a channel with one 16-byte broadcast read).
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
constexpr int DD_AHEAD = 4;  // trials whose LDS reads are issued ahead of their adds
constexpr int DD_ROW_BYTES = 56 * 1024;  // row copies of one stage (at most)
constexpr int DD_RAW_BYTES = 8192;       // raw byte windows of one stage (at most)
constexpr int DD_RAW_CHANS = DD_WAVES;   // channels per stage (at most): one per wave
constexpr int DD_RSRC_WORD3 = 0x00020000;  // buffer descriptor word 3: raw 32-bit data

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

__global__ __launch_bounds__(DD_THREADS, 4) void k_dedisp_u8(
    const uint8_t* __restrict__ data, uint64_t pitch, uint32_t nchans, uint32_t phase_chans,
    const int32_t* __restrict__ bases, const int32_t* __restrict__ rel, uint32_t ntrials,
    uint32_t nblk, uint32_t ntiles, uint32_t n, void* __restrict__ out, uint32_t out_pitch,
    int fp16, float m0, float inv_s0, double* __restrict__ partial,
    const uint32_t* __restrict__ cinfo, const uint32_t* __restrict__ sfirst, uint32_t nstages,
    uint32_t RB) {
    // Static LDS: the stage's row copies, the [stage channel][DD_TRIALS] LDS byte
    // offsets of the runs each trial reads, and the next stage's raw bytes
    // [channel][RB] and trial offsets [channel][DD_TRIALS] written by direct
    // global->LDS loads.  They are separate LDS objects, so the compiler can tell
    // that the sums' reads of the rows need not wait for those loads.  Dynamic
    // LDS: per channel its base, its row length W4 and row offset (cinfo), then
    // the stages' first channels.
    __shared__ __align__(16) uint32_t lds[DD_ROW_BYTES / 4];
    __shared__ __align__(16) uint16_t laddr[DD_RAW_CHANS * DD_TRIALS];
    __shared__ __align__(16) uint8_t raw[DD_RAW_BYTES];
    __shared__ __align__(16) int32_t rrel[DD_RAW_CHANS * DD_TRIALS];
    extern __shared__ __align__(16) int32_t lbase[];
    uint32_t* lcinfo = (uint32_t*)(lbase + nchans);
    uint32_t* lsfirst = lcinfo + nchans;

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

    // acc16: this phase's sums (u16 pairs, no half can overflow within a phase);
    // lo: running sums mod 2^16 per half; cy: carries out of each half, 4 bits per
    // output (bits 4j + 16h for word j, half h; the host keeps every total below 2^20)
    uint32_t acc16[DD_M][4], lo[DD_M][4], cy[DD_M];
#pragma unroll
    for (int m = 0; m < DD_M; m++) {
#pragma unroll
        for (int j = 0; j < 4; j++) acc16[m][j] = lo[m][j] = 0;
        cy[m] = 0;
    }
    auto widen = [&]() {
#pragma unroll
        for (int m = 0; m < DD_M; m++)
#pragma unroll
            for (int j = 0; j < 4; j++) {
                const uint32_t a = acc16[m][j], t = lo[m][j];
                const uint32_t x = (t & 0xffffu) + (a & 0xffffu);
                const uint32_t y = (t >> 16) + (a >> 16);
                lo[m][j] = (x & 0xffffu) | (y << 16);
                cy[m] += ((x >> 16) << (4 * j)) + ((y >> 16) << (4 * j + 16));
                acc16[m][j] = 0;
            }
    };
    uint32_t in_phase = 0;
    const char* lds_lane = (const char*)lds + 16 * lane;
    for (uint32_t i = threadIdx.x; i < nchans; i += DD_THREADS) {
        lbase[i] = bblk[i];
        lcinfo[i] = cinfo[i];
    }
    for (uint32_t i = threadIdx.x; i <= nstages; i += DD_THREADS) lsfirst[i] = sfirst[i];
    // Wave w stages channel w of each stage.  Its raw window starts at t0 + base,
    // read from the 16-byte aligned start s16 in 1 KiB wave loads into raw slot w;
    // lanes past the row end read its last 16 bytes instead (padding, only under
    // outputs at or past n).  MUBUF loads (not FLAT) keep the sums' waits partial.
    auto issue = [&](uint32_t st) {
        const uint32_t cc = __builtin_amdgcn_readfirstlane(lsfirst[st]);
        const uint32_t nc = __builtin_amdgcn_readfirstlane(lsfirst[st + 1]) - cc;
        if ((uint32_t)wave >= nc) return;
        uint32_t ln = threadIdx.x;  // recomputed per call: hoisted per-lane values would spill
        asm volatile("" : "+v"(ln));
        ln &= 63;
        const uint32_t c = cc + wave;
        const uint32_t w4 = __builtin_amdgcn_readfirstlane(lcinfo[c]) & 0xffffu;
        const uint32_t s16 = ((uint32_t)t0 + (uint32_t)__builtin_amdgcn_readfirstlane(lbase[c])) & ~15u;
        const __amdgpu_buffer_rsrc_t rr = __builtin_amdgcn_make_buffer_rsrc(
            (void*)(data + (size_t)c * pitch), 0, (int)pitch, DD_RSRC_WORD3);
        for (uint32_t k = 0; k < w4 + 276; k += 1024) {
            const uint32_t p = (uint32_t)min((uint64_t)s16 + k + 16u * ln, pitch - 16);
            __builtin_amdgcn_raw_ptr_buffer_load_lds(rr, raw + (size_t)wave * RB + k, 16, p, 0, 0, 0);
        }
        const __amdgpu_buffer_rsrc_t rt = __builtin_amdgcn_make_buffer_rsrc(
            (void*)(rblk + (size_t)c * DD_TRIALS), 0, 4 * DD_TRIALS, DD_RSRC_WORD3);
        __builtin_amdgcn_raw_ptr_buffer_load_lds(rt, rrel + (size_t)wave * DD_TRIALS, 4, 4 * ln, 0, 0, 0);
    };
    __syncthreads();
    issue(0);
    for (uint32_t st = 0; st < nstages; st++) {
        const uint32_t cc = __builtin_amdgcn_readfirstlane(lsfirst[st]);
        const uint32_t nc = __builtin_amdgcn_readfirstlane(lsfirst[st + 1]) - cc;
        // this wave's loads have landed; after the barrier every wave's have, and the
        // previous stage's sums are done with the rows
        __builtin_amdgcn_s_waitcnt(0x0F70);  // vmcnt(0)
        __syncthreads();
        if ((uint32_t)wave < nc) {
            uint32_t ln = threadIdx.x;  // as in issue
            asm volatile("" : "+v"(ln));
            ln &= 63;
            const uint32_t c = cc + wave;
            const uint32_t info = __builtin_amdgcn_readfirstlane(lcinfo[c]);
            const uint32_t w4 = info & 0xffffu, rowoff = info >> 16, quads = w4 / 4;
            const uint32_t mis = ((uint32_t)t0 + (uint32_t)__builtin_amdgcn_readfirstlane(lbase[c])) & 15u;
            const uint32_t* rw0 = (const uint32_t*)(raw + (size_t)wave * RB + mis);
            uint32_t* row = lds + rowoff / 4;
            for (uint32_t qd = ln; qd < quads; qd += 64) {
                const uint32_t* rw = rw0 + qd;
                const uint32_t a0 = rw[0], a1 = rw[1], b0 = rw[64], b1 = rw[65];
                uint32_t* dst = row + 4 * qd;
                // w[k] pairs byte k of the 8-byte window at 4 qd with byte k of the
                // window 256 bytes later; copy k holds w[k .. k + 3]
                uint32_t w[7];
#pragma unroll
                for (int k = 0; k < 4; k++) w[k] = pair_word(b0, a0, k);
#pragma unroll
                for (int k = 0; k < 3; k++) w[4 + k] = pair_word(b1, a1, k);
#pragma unroll
                for (int k = 0; k < 4; k++)
                    *(uint4*)(dst + (size_t)k * w4) = make_uint4(w[k], w[k + 1], w[k + 2], w[k + 3]);
            }
            const int32_t e = rrel[wave * DD_TRIALS + ln];
            laddr[wave * DD_TRIALS + ln] = (uint16_t)(rowoff + ((uint32_t)(e & 3) * w4 + (uint32_t)(e & ~3)) * 4);
        }
        __syncthreads();
        if (st + 1 < nstages) issue(st + 1);
        // channels in pairs: one v_add3_u32 adds both channels' u16 pairs (no half
        // carries within a phase, so the 32-bit add is two exact 16-bit adds).  The
        // wave's 8 offsets per channel come from 16-byte broadcast reads (16 bits
        // each), fetched one pair ahead.
        const uint16_t* lw = laddr + wave * DD_M;
        uint4 A = *(const uint4*)(lw), B = *(const uint4*)(lw + DD_TRIALS);
        for (uint32_t c = 0; c + 1 < nc; c += 2) {
            const uint32_t pa[4] = {A.x, A.y, A.z, A.w}, pb[4] = {B.x, B.y, B.z, B.w};
            uint4 x[DD_M], y[DD_M];
#pragma unroll
            for (int m = 0; m < DD_M; m++) {
                const uint32_t sa = m & 1 ? pa[m >> 1] >> 16 : pa[m >> 1] & 0xffffu;
                const uint32_t sb = m & 1 ? pb[m >> 1] >> 16 : pb[m >> 1] & 0xffffu;
                x[m] = *(const uint4*)(lds_lane + sa);
                y[m] = *(const uint4*)(lds_lane + sb);
            }
            const uint32_t cn = c + 3 < nc ? c + 2 : c;  // next pair (a repeat at the end)
            A = *(const uint4*)(lw + cn * DD_TRIALS);
            B = *(const uint4*)(lw + (cn + 1) * DD_TRIALS);
#pragma unroll
            for (int m = 0; m < DD_M; m++) {
                acc16[m][0] = acc16[m][0] + x[m].x + y[m].x;
                acc16[m][1] = acc16[m][1] + x[m].y + y[m].y;
                acc16[m][2] = acc16[m][2] + x[m].z + y[m].z;
                acc16[m][3] = acc16[m][3] + x[m].w + y[m].w;
            }
            // issue order: the reads of DD_AHEAD trials, then each trial's four adds
            // followed by the two reads DD_AHEAD trials later; the next offsets go
            // out after the last data reads
#pragma unroll
            for (int m = 0; m < DD_AHEAD; m++) {
                __builtin_amdgcn_sched_group_barrier(0x0002, 2, 0);  // VALU: addresses
                __builtin_amdgcn_sched_group_barrier(0x0100, 2, 0);  // DS read: x, y
            }
#pragma unroll
            for (int m = 0; m < DD_M; m++) {
                __builtin_amdgcn_sched_group_barrier(0x0002, 4, 0);  // VALU: adds of trial m
                if (m + DD_AHEAD < DD_M) {
                    __builtin_amdgcn_sched_group_barrier(0x0002, 2, 0);
                    __builtin_amdgcn_sched_group_barrier(0x0100, 2, 0);
                }
                if (m + DD_AHEAD == DD_M - 1) __builtin_amdgcn_sched_group_barrier(0x0100, 2, 0);
            }
            in_phase += 2;
            if (in_phase == phase_chans) {
                in_phase = 0;
                widen();
            }
        }
        if (nc & 1) {  // odd channel count: the last stage's final channel alone
            const uint32_t c = nc - 1;
            const uint4 A = *(const uint4*)(laddr + c * DD_TRIALS + wave * DD_M);
            const uint32_t pa[4] = {A.x, A.y, A.z, A.w};
#pragma unroll
            for (int m = 0; m < DD_M; m++) {
                const uint32_t sa = m & 1 ? pa[m >> 1] >> 16 : pa[m >> 1] & 0xffffu;
                const uint4 x = *(const uint4*)(lds_lane + sa);
                acc16[m][0] += x.x;
                acc16[m][1] += x.y;
                acc16[m][2] += x.z;
                acc16[m][3] += x.w;
            }
            in_phase += 1;  // phase_chans is even and in_phase was even: no boundary here
        }
    }
    widen();
    // totals: low 16 bits from lo, the carries above them
    uint32_t acc32[DD_M][8];
#pragma unroll
    for (int m = 0; m < DD_M; m++)
#pragma unroll
        for (int j = 0; j < 4; j++) {
            acc32[m][j] = (lo[m][j] & 0xffffu) + (((cy[m] >> (4 * j)) & 15u) << 16);
            acc32[m][4 + j] = (lo[m][j] >> 16) + (((cy[m] >> (4 * j + 16)) & 15u) << 16);
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
    // bases, relative delays, then the channels' row info and the stages' first channels
    return nblk * nchans * (1 + DD_TRIALS) + 2 * (size_t)nchans + 2;
}

bool launch_dedisp_u8(hipStream_t s, const uint8_t* data, uint64_t pitch, uint32_t nchans,
                      uint32_t maxval, const int32_t* delays, int32_t* table, uint32_t ntrials,
                      uint32_t n, void* out, uint32_t out_pitch, int fp16, float m0, float inv_s0,
                      double* partial, const uint32_t* spread) {
    // 4-bit carry counters per output: every total must stay below 2^20;
    // 16-byte aligned raw loads inside each row; 32-bit buffer offsets
    if (nchans == 0 || (uint64_t)nchans * maxval >= (1u << 20)) return false;
    if (pitch % 16 || pitch < 16 || pitch >= (1ull << 31)) return false;
    // Row length of channel c: a trial reads words (e & ~3) .. (e & ~3) + 255 with
    // e <= spread[c] + 3.  Its raw window (read from a 16-byte aligned start) takes
    // W4 + 276 bytes, in whole 1 KiB wave loads; RB is the raw slot size.
    std::vector<uint32_t> w4(nchans), info(nchans), first;
    uint64_t RB = 1024;
    for (uint32_t c = 0; c < nchans; c++) {
        const uint64_t w = ((uint64_t)spread[c] + 3 + 3) / 4 * 4 + DD_TILE / 2;
        if (w > 0xffff) return false;
        w4[c] = (uint32_t)w;
        RB = std::max<uint64_t>(RB, (w + 276 + 1023) / 1024 * 1024);
    }
    if (RB * 2 > DD_RAW_BYTES) return false;
    const uint32_t slots = std::min<uint32_t>(DD_RAW_CHANS, (uint32_t)(DD_RAW_BYTES / RB));
    // stages: consecutive channels whose four row copies fit DD_ROW_BYTES (< 64 KiB,
    // for the 16-bit offsets), one per wave; an even count (channels are added in
    // pairs) except in the last stage
    for (uint32_t c = 0; c < nchans;) {
        uint32_t k = 0, rows = 0;
        while (c + k < nchans && k < slots && rows + 16 * w4[c + k] <= (uint32_t)DD_ROW_BYTES)
            rows += 16 * w4[c + k++];
        if (c + k < nchans) k &= ~1u;
        if (k == 0) return false;
        first.push_back(c);
        for (uint32_t j = 0, off = 0; j < k; off += 16 * w4[c + j], j++) info[c + j] = w4[c + j] | off << 16;
        c += k;
    }
    const uint32_t nstages = (uint32_t)first.size();
    first.push_back(nchans);
    // dynamic LDS: bases and row info per channel, first channel per stage
    const size_t shm = 8 * (size_t)nchans + 4 * (size_t)(nstages + 1);
    if (shm > 80 * 1024) return false;
    const size_t nblk = (ntrials + DD_TRIALS - 1) / DD_TRIALS;
    int32_t* bases = table;
    int32_t* rel = table + nblk * nchans;
    uint32_t* cinfo = (uint32_t*)(rel + nblk * nchans * DD_TRIALS);
    uint32_t* sfirst = cinfo + nchans;
    if (hipMemcpyAsync(cinfo, info.data(), (size_t)nchans * 4, hipMemcpyHostToDevice, s) != hipSuccess ||
        hipMemcpyAsync(sfirst, first.data(), first.size() * 4, hipMemcpyHostToDevice, s) != hipSuccess)
        return false;
    hipLaunchKernelGGL(k_dd_table, dim3((unsigned)((nblk * nchans + 255) / 256)), dim3(256), 0, s,
                       delays, nchans, ntrials, bases, rel);
    // channels per phase: no 16-bit half can overflow; even, so pairs never straddle
    uint32_t phase = 65535u / (maxval ? maxval : 1);
    phase = (phase > 4096 ? 4096 : phase) & ~1u;
    (void)hipFuncSetAttribute(reinterpret_cast<const void*>(k_dedisp_u8),
                              hipFuncAttributeMaxDynamicSharedMemorySize, (int)shm);
    const uint32_t ntiles = (n + DD_TILE - 1) / DD_TILE;
    const uint64_t total = (uint64_t)ntiles * nblk;
    hipLaunchKernelGGL(k_dedisp_u8, dim3((unsigned)total), dim3(DD_THREADS), shm, s, data, pitch, nchans,
                       phase, bases, rel, ntrials, (uint32_t)nblk, ntiles, n, out, out_pitch, fp16,
                       m0, inv_s0, partial, cinfo, sfirst, nstages, (uint32_t)RB);
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
