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
// Workgroup = 8 waves sharing one time tile of DD_TILE = 512 outputs and a
// block of DD_TRIALS = 64 consecutive (ascending) DM trials; each wave owns
// DD_M = 8 trials, and lane l owns outputs t0 + l + 64 k (k < 8).  A channel's
// samples are staged into LDS as 16-byte words of eight u16 lanes,
//     w[i] = (s[base + i + 64 k], k < 8)      (dword j: k = 2j low, 2j + 1 high),
// with base = the block's smallest delay of the channel rounded down to 4.  A
// trial whose delay is base + e reads word e + l in lane l with one aligned
// ds_read_b128: the 64 lanes read 1 KiB of consecutive LDS, free of bank
// conflicts, and a row is only as long as the delays of the trials reading it
// require (their e range + 64 words).
//
// Channels are summed in pairs (c, c + 1).  Within a block the relative delay
// delta = delay[c + 1] - delay[c] takes few distinct values (two or three when
// the block spans a small DM range), so for each value the workgroup builds one
// "pair row" P[i] = w_c[i] + w_(c+1)[i + delta - (base_(c+1) - base_c)] (u16
// lanes of at most 2 * 255) and every trial reads one pair row per channel
// pair: half the LDS reads and half the adds of reading both channel rows.
// Pairs whose delta takes more than kmax values in a block are summed as two
// single-channel rows.
//
// Rows are added in pairs with v_add3_u32 on the u16 lanes: within a "phase"
// of at most 65535 / (2 maxval) rows no lane can carry, so the 32-bit add is
// two exact 16-bit adds.  At each phase end the halves are added into running
// sums mod 2^16 whose carries are counted in 4-bit fields, so totals up to 2^20
// are exact.
//
// A planner kernel lays out each block as stages: slots (a slot is a channel
// pair or a single channel) whose rows fit the LDS row area, taken from both
// ends of the band in turn so that every stage mixes wide and narrow rows.  A
// stage is one fixed-size record -- header, row records, every trial's 16-bit
// LDS read address per slot and the row of every quad (four row words).  At a
// stage boundary all 512 lanes build the stage's rows, one quad per lane at a
// time, straight from the channel rows in global memory (L2-resident: the
// workgroups of one XCD work on the same time tile): a quad's eight dwords of
// channel c are aligned, those of channel c + 1 are unaligned dword loads at
// the pair's relative delay, so the build costs the LDS only its row stores.
// The record of the next stage is loaded into registers during the sums and
// stored at the next boundary (no scalar loads, whose out-of-order returns
// would force full LDS waits in the sums).  Two workgroups share a CU, so
// one's row build overlaps the other's sums.
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
constexpr int DD_AHEAD = 2;  // trials whose LDS reads are issued ahead of their adds
constexpr int DD_ROW_BYTES = 64 * 1024;  // rows of one stage (at most; 16-bit LDS addresses)
constexpr int DD_RECS = 2;  // stage records held: the one summed, the next one
// Bytes readable past the last channel row of a u8 observation (its
// allocation's slack): the row build reads up to sample t0 + delay + 519 of a
// channel, which the host bounds by the row pitch plus this slack.
constexpr uint64_t DD_READ_REACH = 520;
constexpr int DDP_MAX_ROWS = 24;                  // rows per stage
constexpr int DDP_MAX_SLOTS = 16;                 // slots per stage
constexpr int DDP_KMAX = 3;                       // pair rows per channel pair (at most)
// stage record (u32 words): [0] nslot << 8 | nrow << 16, [1] quads (4-word
// units of the row buffer), [DDP_ROW_REC + 4 r ..] row r: rowoff | quads << 16,
// channel c | paired << 16, the sample of word 0 lane 0 of channel c relative
// to the tile start, the same for channel c + 1 (paired rows);
// [DDP_LAD_REC + 32 s ..] read addresses of slot s (DD_TRIALS x u16);
// [DDP_QMAP_REC ..] the row of every quad (u8).  Rows are packed in record
// order, so quad q of the stage is quad q - rowoff / 64 of its row and lands at
// byte 64 q of the row buffer.
constexpr int DDP_ROW_REC = 4;
constexpr int DDP_LAD_REC = DDP_ROW_REC + 4 * DDP_MAX_ROWS;
constexpr int DDP_MAX_QUADS = DD_ROW_BYTES / 64;  // 4 row words of 16 bytes
constexpr int DDP_QMAP_REC = DDP_LAD_REC + DDP_MAX_SLOTS * DD_TRIALS / 2;
constexpr int DDP_STAGE_WORDS = DDP_QMAP_REC + DDP_MAX_QUADS / 4;
constexpr int DDP_STAGE_CHUNKS = (DDP_STAGE_WORDS + 63) / 64;
constexpr int DDP_STAGE_PITCH = 64 * DDP_STAGE_CHUNKS;
constexpr int DDP_PLAN_HEAD = 64;  // word 0: stage count

static size_t ddp_plan_shm(uint32_t nchans) {  // dynamic LDS of the planner
    const size_t npair = (nchans + 1) / 2;
    return 4 * (2 * (size_t)nchans + 2 * npair + 3 * DDP_KMAX * npair) + DDP_KMAX * npair;
}

static size_t ddp_plan_words(uint32_t nchans) {  // per block; stages <= slots <= nchans
    return DDP_PLAN_HEAD + (size_t)nchans * DDP_STAGE_PITCH;
}

// One workgroup per block of DD_TRIALS trials.  Dynamic LDS: bases[nchans],
// per slot its stage | position << 16, then per channel pair its kind (K pair
// rows, 0 = single-channel rows) and first slot, and per row its delta,
// L | W << 16 and LDS offset.
__device__ __forceinline__ void dd_plan_block(const int32_t* __restrict__ delays, uint32_t nchans,
                                              uint32_t ntrials, uint32_t kmax, uint32_t row_budget, int32_t* __restrict__ bases,
                                              uint32_t* __restrict__ plan, uint64_t plan_words,
                                              uint32_t blk) {
    extern __shared__ int32_t sbase[];
    const uint32_t npair = (nchans + 1) / 2;
    uint32_t* sloc = (uint32_t*)(sbase + nchans);
    uint32_t* ukind = sloc + nchans;
    uint32_t* uslot = ukind + npair;
    int32_t* udel = (int32_t*)(uslot + npair);
    uint32_t* uLW = (uint32_t*)(udel + (size_t)npair * DDP_KMAX);
    uint32_t* uRO = uLW + (size_t)npair * DDP_KMAX;
    uint8_t* uRN = (uint8_t*)(uRO + (size_t)npair * DDP_KMAX);
    const uint32_t d0 = blk * DD_TRIALS;
    uint32_t* pb = plan + (size_t)blk * plan_words;
    auto delay = [&](uint32_t m, uint32_t c) {
        return delays[(size_t)min(d0 + m, ntrials - 1) * nchans + c];
    };
    // row of words [L, L + W) read by trials with e in [lo, hi]: L = lo & ~3 and
    // W = hi - L + 64 rounded up to 4 (a trial reads words e .. e + 63)
    auto row_lw = [](uint32_t lo, uint32_t hi) {
        const uint32_t L = lo & ~3u;
        return L | ((hi - L + DD_TILE / 8 + 3) & ~3u) << 16;
    };
    for (uint32_t c = threadIdx.x; c < nchans; c += blockDim.x) {
        const int32_t b = delay(0, c) & ~3;
        sbase[c] = b;
        bases[(size_t)blk * nchans + c] = b;
    }
    __syncthreads();
    // channel pairs: distinct deltas and, per delta, the e range of channel c
    // (the delays are read 16 trials at a time, all loads issued together)
    constexpr int TB = 16;
    for (uint32_t u = threadIdx.x; u < npair; u += blockDim.x) {
        const uint32_t c = 2 * u;
        uint32_t K = 0;
        bool split = c + 1 >= nchans;
        int32_t del[DDP_KMAX] = {0, 0, 0}, lo[DDP_KMAX] = {0, 0, 0}, hi[DDP_KMAX] = {0, 0, 0};
        int32_t emin[2] = {1 << 30, 1 << 30}, emax[2] = {0, 0};
        const int32_t b0 = sbase[c], b1 = c + 1 < nchans ? sbase[c + 1] : 0;
        for (uint32_t m0 = 0; m0 < (uint32_t)DD_TRIALS; m0 += TB) {
            int32_t dc[TB], d1[TB];
#pragma unroll
            for (int i = 0; i < TB; i++) {
                dc[i] = delay(m0 + i, c);
                d1[i] = c + 1 < nchans ? delay(m0 + i, c + 1) : 0;
            }
#pragma unroll
            for (int i = 0; i < TB; i++) {
                const int32_t e = dc[i] - b0, dl = d1[i] - dc[i];
                emin[0] = min(emin[0], e);
                emax[0] = max(emax[0], e);
                emin[1] = min(emin[1], d1[i] - b1);
                emax[1] = max(emax[1], d1[i] - b1);
                if (split) continue;
                uint32_t k = 0;
                while (k < K && del[k] != dl) k++;
                if (k == K) {
                    if (K == kmax) {
                        split = true;
                        continue;
                    }
                    del[K] = dl;
                    lo[K] = hi[K] = e;
                    K++;
                } else {
                    lo[k] = min(lo[k], e);
                    hi[k] = max(hi[k], e);
                }
            }
        }
        if (split) K = 0;
        ukind[u] = K;
#pragma unroll
        for (int k = 0; k < DDP_KMAX; k++) {
            udel[u * DDP_KMAX + k] = del[k];
            uLW[u * DDP_KMAX + k] = (uint32_t)k < K ? row_lw((uint32_t)lo[k], (uint32_t)hi[k])
                                    : (K == 0 && k < 2) ? row_lw((uint32_t)emin[k], (uint32_t)emax[k]) : 0u;
        }
    }
    __syncthreads();
    // stages (one thread; each pair's data is read up front).  Pairs are taken
    // from both ends of the band in turn (0, npair - 1, 1, npair - 2, ...):
    // the wide rows of the long-delay channels and the many narrow ones of the
    // short-delay channels then share every stage, which keeps the row build
    // of each stage and the sums of the one before it of similar length (the
    // integer sums are exact in any channel order).
    if (threadIdx.x == 0) {
        uint32_t nst = 0, nslot_all = 0;
        uint32_t nslot = 0, nrow = 0, bytes = 0;
        uint32_t* st = pb + DDP_PLAN_HEAD;
        auto close = [&]() {
            if (!nslot) return;
            st[0] = nslot << 8 | nrow << 16;
            st[1] = bytes / 64;
            nst++;
            st = pb + DDP_PLAN_HEAD + (size_t)nst * DDP_STAGE_PITCH;
            nslot = nrow = bytes = 0;
        };
        // row record: word 0 lane 0 of the row is sample off_c of channel ch
        // (and off_c1 of channel ch + 1 for a pair row), relative to the tile
        auto put_row = [&](uint32_t ui, uint32_t W, uint32_t ch, int32_t off_c, int32_t off_c1, uint32_t paired) {
            uRO[ui] = bytes;
            uRN[ui] = (uint8_t)nrow;
            uint32_t* r = st + DDP_ROW_REC + 4 * nrow;
            r[0] = bytes | (W / 4) << 16;
            r[1] = ch | paired << 16;
            r[2] = (uint32_t)off_c;
            r[3] = (uint32_t)off_c1;
            bytes += 16 * W;
            nrow++;
        };
        auto put_slot = [&]() {
            sloc[nslot_all] = nst | nslot << 16;
            nslot++;
            nslot_all++;
        };
        auto fits = [&](uint32_t r, uint32_t b) {
            return nrow + r <= (uint32_t)DDP_MAX_ROWS && bytes + b <= row_budget && nslot < (uint32_t)DDP_MAX_SLOTS;
        };
        for (uint32_t i = 0; i < npair; i++) {
            const uint32_t u = i & 1 ? npair - 1 - i / 2 : i / 2;
            const uint32_t c = 2 * u, K = ukind[u];
            uint32_t LW[DDP_KMAX];
            int32_t dl[DDP_KMAX];
#pragma unroll
            for (int k = 0; k < DDP_KMAX; k++) {
                LW[k] = uLW[u * DDP_KMAX + k];
                dl[k] = udel[u * DDP_KMAX + k];
            }
            const int32_t sb0 = sbase[c], sb1 = c + 1 < nchans ? sbase[c + 1] : 0;
            if (K) {
                uint32_t need = 0;
#pragma unroll
                for (int k = 0; k < DDP_KMAX; k++) need += (uint32_t)k < K ? 16 * (LW[k] >> 16) : 0u;
                if (!fits(K, need)) close();
                uslot[u] = nslot_all;
#pragma unroll
                for (int k = 0; k < DDP_KMAX; k++) {
                    if ((uint32_t)k >= K) break;
                    const uint32_t L = LW[k] & 0xffffu, W = LW[k] >> 16;
                    put_row(u * DDP_KMAX + k, W, c, sb0 + (int32_t)L, sb0 + (int32_t)L + dl[k], 1);
                }
                put_slot();
            } else {
                for (uint32_t j = 0; j < 2 && c + j < nchans; j++) {
                    const uint32_t L = LW[j] & 0xffffu, W = LW[j] >> 16;
                    if (!fits(1, 16 * W)) close();
                    if (j == 0) uslot[u] = nslot_all;
                    const int32_t sb = j ? sb1 : sb0;
                    put_row(u * DDP_KMAX + j, W, c + j, sb + (int32_t)L, sb + (int32_t)L, 0);
                    put_slot();
                }
            }
        }
        close();
        pb[0] = nst;
    }
    __syncthreads();
    // the row of every quad of every stage: one pair per thread
    for (uint32_t u = threadIdx.x; u < npair; u += blockDim.x) {
        const uint32_t c = 2 * u, K = ukind[u];
        const uint32_t nr = K ? K : (c + 1 < nchans ? 2u : 1u);
        for (uint32_t k = 0; k < nr; k++) {
            const uint32_t ui = u * DDP_KMAX + k;
            const uint32_t stg = sloc[uslot[u] + (K ? 0u : k)] & 0xffffu;
            uint8_t* map = (uint8_t*)(pb + DDP_PLAN_HEAD + (size_t)stg * DDP_STAGE_PITCH + DDP_QMAP_REC);
            const uint32_t q0 = uRO[ui] / 64, nq = (uLW[ui] >> 16) / 4;
            for (uint32_t q = 0; q < nq; q++) map[q0 + q] = uRN[ui];
        }
    }
    // every trial's LDS read address per slot: one pair per thread, 16 trials
    // at a time
    auto lad = [&](uint32_t slot) {
        const uint32_t loc = sloc[slot];
        return (uint16_t*)(pb + DDP_PLAN_HEAD + (size_t)(loc & 0xffffu) * DDP_STAGE_PITCH + DDP_LAD_REC) +
               (size_t)(loc >> 16) * DD_TRIALS;
    };
    for (uint32_t u = threadIdx.x; u < npair; u += blockDim.x) {
        const uint32_t c = 2 * u, K = ukind[u];
        const int32_t b0 = sbase[c], b1 = c + 1 < nchans ? sbase[c + 1] : 0;
        uint16_t* l0 = lad(uslot[u]);
        uint16_t* l1 = K == 0 && c + 1 < nchans ? lad(uslot[u] + 1) : l0;
        for (uint32_t m0 = 0; m0 < (uint32_t)DD_TRIALS; m0 += TB) {
            int32_t dc[TB], d1[TB];
#pragma unroll
            for (int i = 0; i < TB; i++) {
                dc[i] = delay(m0 + i, c);
                d1[i] = c + 1 < nchans ? delay(m0 + i, c + 1) : 0;
            }
#pragma unroll
            for (int i = 0; i < TB; i++) {
                const uint32_t m = m0 + i, e = (uint32_t)(dc[i] - b0);
                if (K) {
                    const int32_t dl = d1[i] - dc[i];
                    uint32_t k = 0;
                    while (k + 1 < K && udel[u * DDP_KMAX + k] != dl) k++;
                    const uint32_t L = uLW[u * DDP_KMAX + k] & 0xffffu;
                    l0[m] = (uint16_t)(uRO[u * DDP_KMAX + k] + (e - L) * 16);
                } else {
                    l0[m] = (uint16_t)(uRO[u * DDP_KMAX] + (e - (uLW[u * DDP_KMAX] & 0xffffu)) * 16);
                    if (c + 1 < nchans) {
                        const uint32_t e1 = (uint32_t)(d1[i] - b1);
                        l1[m] = (uint16_t)(uRO[u * DDP_KMAX + 1] + (e1 - (uLW[u * DDP_KMAX + 1] & 0xffffu)) * 16);
                    }
                }
            }
        }
    }
}


__global__ __launch_bounds__(256) void k_dd_plan(const int32_t* __restrict__ delays, uint32_t nchans,
                                                 uint32_t ntrials, uint32_t kmax, uint32_t row_budget,
                                                 int32_t* __restrict__ bases, uint32_t* __restrict__ plan,
                                                 uint64_t plan_words) {
    dd_plan_block(delays, nchans, ntrials, kmax, row_budget, bases, plan, plan_words, blockIdx.x);
}

// Record of stage st, moved through registers: thread t of nt (nt >= half a
// record's dwords) holds dwords 2 t and 2 t + 1 (indices clamped into range;
// the surplus is never stored).
__device__ __forceinline__ uint2 dd_rec_load(const uint32_t* __restrict__ pb, uint32_t st, uint32_t t) {
    const uint32_t* src = pb + DDP_PLAN_HEAD + (size_t)st * DDP_STAGE_PITCH;
    const uint32_t i = min(2 * t, (uint32_t)DDP_STAGE_PITCH - 2);
    return make_uint2(src[i], src[i + 1]);
}
__device__ __forceinline__ void dd_rec_store(uint32_t* slot, uint2 v, uint32_t t) {
    if (2 * t < (uint32_t)DDP_STAGE_PITCH) *(uint2*)(slot + 2 * t) = v;
}
static_assert(2 * DD_THREADS >= DDP_STAGE_PITCH, "a record is two dwords per thread");

// The rows of the stage whose record is sr, built into `rows` by threads
// t = 0 .. nt - 1: the threads take the stage's quads (4 row words) in turn;
// quad q belongs to row qmap[q] and lands at row word 16 q.  A thread builds
// its quad's words from eight dwords of channel c (samples off_c + 4 g + 64 k,
// k < 8, g = the quad's index in its row: aligned) and eight of channel c + 1
// (samples off_c1 + 4 g + 64 k: unaligned dword loads), read from the channel
// rows in global memory (L2-resident: the workgroups of one XCD work on the
// same time tile), and stores them in a lane-rotated order that keeps the
// 16-byte stores of every 8 or 16 lanes on distinct LDS banks.  Single-channel rows read their
// own channel twice and zero the second copy's bytes with an all-zero byte
// select.
__device__ __forceinline__ void dd_build(const uint32_t* sr, uint32_t* rows, const uint8_t* __restrict__ tile_base,
                                         uint64_t pitch, uint32_t t, uint32_t nt) {
    const uint32_t nq = __builtin_amdgcn_readfirstlane(sr[1]);
    const uint8_t* qmap = (const uint8_t*)(sr + DDP_QMAP_REC);
    for (uint32_t q = t; q < nq; q += nt) {
        const uint32_t r = qmap[q];
        const uint4 rec = *(const uint4*)(sr + DDP_ROW_REC + 4 * r);
        const uint32_t g = q - (rec.x & 0xffffu) / 64;
        const bool paired = (rec.y >> 16) & 1u;
        const uint8_t* pc = tile_base + (size_t)(rec.y & 0xffffu) * pitch + (int32_t)(rec.z + 4 * g);
        const uint8_t* p1 = paired ? pc + pitch + ((int32_t)rec.w - (int32_t)rec.z) : pc;
        uint32_t d[8], e[8];
#pragma unroll
        for (int k = 0; k < 8; k++) {
            d[k] = *(const uint32_t*)(pc + 64 * k);
            __builtin_memcpy(&e[k], p1 + 64 * k, 4);
        }
        // store order of the four words: lanes q with equal q % 4 within 16
        // lanes, and the even (odd) lanes within 8, start at different words,
        // so the 16-byte stores of 8 or 16 lanes at a time hit distinct banks
        const uint32_t rot = ((q >> 1) + (q >> 3)) & 3u;
        uint32_t* row = rows + 16 * q;
#pragma unroll
        for (uint32_t bb = 0; bb < 4; bb++) {
            const uint32_t b = (bb + rot) & 3u;
            const uint32_t sd = 0x0c040c00u + 0x00010001u * b, se = paired ? sd : 0x0c0c0c0cu;
            *(uint4*)(row + 4 * b) = make_uint4(
                __builtin_amdgcn_perm(d[1], d[0], sd) + __builtin_amdgcn_perm(e[1], e[0], se),
                __builtin_amdgcn_perm(d[3], d[2], sd) + __builtin_amdgcn_perm(e[3], e[2], se),
                __builtin_amdgcn_perm(d[5], d[4], sd) + __builtin_amdgcn_perm(e[5], e[4], se),
                __builtin_amdgcn_perm(d[7], d[6], sd) + __builtin_amdgcn_perm(e[7], e[6], se));
        }
    }
}

// One summing wave's 8 trials over a 512-output tile.  acc16: this phase's
// sums (u16 lanes, none can overflow within a phase); lo: running sums mod
// 2^16 per lane; cy: carries out of each lane, 4 bits per output (bits
// 4j + 16h for dword j, half h; the host keeps every total below 2^20).
struct DdSums {
    uint32_t acc16[DD_M][4], lo[DD_M][4], cy[DD_M];
    uint32_t in_phase;

    __device__ __forceinline__ void init() {
#pragma unroll
        for (int m = 0; m < DD_M; m++) {
#pragma unroll
            for (int j = 0; j < 4; j++) acc16[m][j] = lo[m][j] = 0;
            cy[m] = 0;
        }
        in_phase = 0;
    }
    __device__ __forceinline__ void widen() {
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
        in_phase = 0;
    }
    // The slots of the stage whose record is sr, read from rows at lds_lane
    // (this lane's 16 bytes of the row buffer), in pairs: one v_add3_u32 adds
    // both slots' u16 lanes (no lane can carry within a phase: every slot adds
    // at most 2 * maxval per lane).  The wave's 8 read addresses per slot come
    // from 16-byte broadcast reads (16 bits each), fetched one pair ahead.
    __device__ __forceinline__ void stage(const uint32_t* sr, const char* lds_lane, int wave,
                                          uint32_t phase_slots) {
        const uint32_t nslot = (__builtin_amdgcn_readfirstlane(sr[0]) >> 8) & 0xffu;
        const uint16_t* lw = (const uint16_t*)(sr + DDP_LAD_REC) + wave * DD_M;
        uint4 A = *(const uint4*)(lw), B = *(const uint4*)(lw + DD_TRIALS);
        for (uint32_t c = 0; c + 1 < nslot; c += 2) {
            if (in_phase + 2 > phase_slots) widen();
            const uint32_t pa[4] = {A.x, A.y, A.z, A.w}, pb2[4] = {B.x, B.y, B.z, B.w};
            uint4 x[DD_M], y[DD_M];
#pragma unroll
            for (int m = 0; m < DD_M; m++) {
                const uint32_t sa = m & 1 ? pa[m >> 1] >> 16 : pa[m >> 1] & 0xffffu;
                const uint32_t sb = m & 1 ? pb2[m >> 1] >> 16 : pb2[m >> 1] & 0xffffu;
                x[m] = *(const uint4*)(lds_lane + sa);
                y[m] = *(const uint4*)(lds_lane + sb);
            }
            const uint32_t cn = c + 3 < nslot ? c + 2 : c;  // next pair (a repeat at the end)
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
            // followed by the two reads DD_AHEAD trials later; the next addresses go
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
        }
        if (nslot & 1) {  // odd slot count: the stage's last slot alone
            if (in_phase + 1 > phase_slots) widen();
            const uint4 A2 = *(const uint4*)(lw + (nslot - 1) * DD_TRIALS);
            const uint32_t pa[4] = {A2.x, A2.y, A2.z, A2.w};
#pragma unroll
            for (int m = 0; m < DD_M; m++) {
                const uint32_t sa = m & 1 ? pa[m >> 1] >> 16 : pa[m >> 1] & 0xffffu;
                const uint4 x = *(const uint4*)(lds_lane + sa);
                acc16[m][0] += x.x;
                acc16[m][1] += x.y;
                acc16[m][2] += x.z;
                acc16[m][3] += x.w;
            }
            in_phase += 1;
        }
    }
    // Totals (low 16 bits from lo, the carries above them) of trials
    // d0 + wave * 8 + m, output t0 + lane + 64 k, written as f32 or as the FP16
    // storage value, and each trial's tile sum (exact in f64) into partial.
    __device__ __forceinline__ void finish(uint32_t d0, int wave, uint32_t t0, uint32_t tile, uint32_t ntrials,
                                           uint32_t ntiles, uint32_t n, void* __restrict__ out, uint32_t out_pitch,
                                           int fp16, float m0, float inv_s0, double* __restrict__ partial) {
        widen();
        uint32_t acc32[DD_M][8];  // acc32[m][k]: dword j holds k = 2j low, 2j + 1 high
#pragma unroll
        for (int m = 0; m < DD_M; m++)
#pragma unroll
            for (int j = 0; j < 4; j++) {
                acc32[m][2 * j] = (lo[m][j] & 0xffffu) + (((cy[m] >> (4 * j)) & 15u) << 16);
                acc32[m][2 * j + 1] = (lo[m][j] >> 16) + (((cy[m] >> (4 * j + 16)) & 15u) << 16);
            }
        const int m0i = (int)m0;
        uint32_t ln = threadIdx.x;
        asm volatile("" : "+v"(ln));
        ln &= 63;
#pragma unroll
        for (int m = 0; m < DD_M; m++) {
            const uint32_t d = d0 + wave * DD_M + m;
            double lsum = 0.0;
            if (d < ntrials) {
#pragma unroll
                for (int k = 0; k < 8; k++) {
                    const uint32_t t = t0 + ln + 64 * k;
                    if (t < n) {
                        const uint32_t v = acc32[m][k];
                        const size_t o = (size_t)d * out_pitch + t;
                        lsum += (double)v;
                        if (fp16)
                            ((__half*)out)[o] = __float2half_rn((float)((int)v - m0i) * inv_s0);
                        else
                            ((float*)out)[o] = (float)v;
                    }
                }
            }
            // wave reduction of the tile's sum (integers < 2^53: exact in any order)
#pragma unroll
            for (int o = 32; o > 0; o >>= 1) lsum += __shfl_xor(lsum, o);
            if (ln == 0 && d < ntrials) partial[(size_t)d * ntiles + tile] = lsum;
        }
    }
};

// One work item of the u8 dedispersion: DM block `blk` (trials blk * 64 ..) over
// time tile `tile` (outputs tile * 512 ..).  plan is the whole launch's table
// (indexed by blk), partial has ntiles words per trial.  Eight waves (two
// workgroups per CU): at each stage boundary all of them build the stage's
// rows, then each sums its 8 trials; the other workgroup of the CU overlaps.
__device__ __forceinline__ void dd_item(const uint8_t* __restrict__ data, uint64_t pitch,
                                        uint32_t phase_slots, const uint32_t* __restrict__ plan,
                                        uint64_t plan_words, uint32_t ntrials, uint32_t ntiles, uint32_t n,
                                        void* __restrict__ out, uint32_t out_pitch, int fp16, float m0,
                                        float inv_s0, double* __restrict__ partial, uint32_t tile,
                                        uint32_t blk) {
    // LDS: the stage's rows and two stage records (the one summed and the
    // next).  Records go through registers (loaded during one stage's sums,
    // stored at the next stage boundary): the sums read their LDS addresses
    // from the records, and a direct global->LDS load would make the sums'
    // first LDS read wait for it.
    __shared__ __align__(16) uint32_t lds[DD_ROW_BYTES / 4];
    __shared__ __align__(16) uint32_t srec[DD_RECS][DDP_STAGE_PITCH];
    const uint32_t t0 = tile * DD_TILE;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const uint32_t* pb = plan + (size_t)blk * plan_words;
    const uint32_t nstages = __builtin_amdgcn_readfirstlane(pb[0]);
    const char* lds_lane = (const char*)lds + 16 * (threadIdx.x & 63);
    DdSums sums;
    sums.init();
    dd_rec_store(srec[0], dd_rec_load(pb, 0, threadIdx.x), threadIdx.x);  // nstages >= 1: nchans >= 1
    uint2 rnext = dd_rec_load(pb, min(1u, nstages - 1), threadIdx.x);  // record st + 1, stored at stage st
    for (uint32_t st = 0; st < nstages; st++) {
        const uint32_t* sr = srec[st % DD_RECS];
        // this wave's record load has landed; after the barrier every wave's
        // has, record st is visible, and the previous stage's sums are done with
        // the rows and with record st - 1's slot, which record st + 1 takes
        __builtin_amdgcn_s_waitcnt(0x0F70);  // vmcnt(0)
        __syncthreads();
        if (st + 1 < nstages) dd_rec_store(srec[(st + 1) % DD_RECS], rnext, threadIdx.x);
        // the build's loads wait on memory while the other workgroup of the CU
        // sums from the LDS: issuing them first (wave priority 1) measured ~1 ms
        // faster per headline launch
        __builtin_amdgcn_s_setprio(1);
        dd_build(sr, lds, data + t0, pitch, threadIdx.x, DD_THREADS);
        __builtin_amdgcn_s_setprio(0);
        __syncthreads();
        rnext = dd_rec_load(pb, min(st + 2, nstages - 1), threadIdx.x);
        sums.stage(sr, lds_lane, wave, phase_slots);
    }
    sums.finish(blk * DD_TRIALS, wave, t0, tile, ntrials, ntiles, n, out, out_pitch, fp16, m0, inv_s0, partial);
}

// One launch over the (tile, block) items of one observation; workgroup ids
// are remapped XCD-aware (each XCD walks a contiguous run of items, block
// fastest, so a tile's channel data is reused from that XCD's L2).
__global__ __launch_bounds__(DD_THREADS, 4) void k_dedisp_u8(
    const uint8_t* __restrict__ data, uint64_t pitch, uint32_t phase_slots, const uint32_t* __restrict__ plan,
    uint64_t plan_words, uint32_t ntrials, uint32_t nblk, uint32_t ntiles, uint32_t n, void* __restrict__ out,
    uint32_t out_pitch, int fp16, float m0, float inv_s0, double* __restrict__ partial) {
    const uint32_t total = ntiles * nblk;
    const uint32_t id = blockIdx.x;
    const uint32_t xcd = id % DD_XCDS, local = id / DD_XCDS;
    const uint32_t q = total / DD_XCDS, r = total % DD_XCDS;
    const uint32_t work = (xcd < r ? xcd * (q + 1) : r * (q + 1) + (xcd - r) * q) + local;
    dd_item(data, pitch, phase_slots, plan, plan_words, ntrials, ntiles, n, out, out_pitch, fp16, m0, inv_s0, partial,
            work / nblk, work % nblk);
}

size_t dedisp_u8_table_words(uint32_t nchans, uint32_t ntrials) {
    const size_t nblk = (ntrials + DD_TRIALS - 1) / DD_TRIALS;
    return nblk * nchans + nblk * ddp_plan_words(nchans);  // bases, then the plans
}

// Shape checks shared by the single- and many-observation launches: false
// when the LDS-staged kernel cannot take the observation.  A row has
// W <= spread + 3 + 64 + 3 words of 16 bytes; the row build reads a channel up
// to sample t0 + delay + DD_READ_REACH (t0 <= n - 1), which must stay inside
// the row pitch plus the slack after the last row.
static bool dd_shape(uint32_t nchans, uint32_t maxval, uint64_t pitch, const uint32_t* spread, uint32_t n,
                     uint64_t max_delay, uint32_t* kmax_out, uint32_t* phase_out) {
    // 4-bit carry counters per output: every total must stay below 2^20;
    // aligned dword loads of channel rows; 16-bit channel numbers
    if (nchans == 0 || n == 0 || (uint64_t)nchans * maxval >= (1u << 20)) return false;
    if (pitch % 16 || pitch < 16 || nchans > 0xffff) return false;
    if (max_delay > pitch || (uint64_t)n - 1 + max_delay + DD_READ_REACH > pitch + DD_ALLOC_SLACK) return false;
    uint32_t smax = 0;
    for (uint32_t c = 0; c < nchans; c++) smax = std::max(smax, spread[c]);
    const uint64_t wmax = (uint64_t)smax + 6 + DD_TILE / 8 + 3;
    if (16 * wmax > (uint64_t)DD_ROW_BYTES) return false;
    // a slot of kmax pair rows must fit the row area alone
    uint32_t kmax = DDP_KMAX;
    while (kmax > 1 && (uint64_t)kmax * 16 * wmax > (uint64_t)DD_ROW_BYTES) kmax--;
    if ((uint64_t)kmax * 16 * wmax > (uint64_t)DD_ROW_BYTES) kmax = 0;
    if (ddp_plan_shm(nchans) > 160 * 1024) return false;
    // slots per phase: no 16-bit lane can overflow (a slot adds at most 2 * maxval)
    uint32_t phase = 65535u / (2 * (maxval ? maxval : 1));
    phase = phase > 4096 ? 4096 : phase;
    if (phase < 2) return false;
    *kmax_out = kmax;
    *phase_out = phase;
    return true;
}

bool launch_dedisp_u8(hipStream_t s, const uint8_t* data, uint64_t pitch, uint32_t nchans,
                      uint32_t maxval, const int32_t* delays, int32_t* table, uint32_t ntrials,
                      uint32_t n, void* out, uint32_t out_pitch, int fp16, float m0, float inv_s0,
                      double* partial, const uint32_t* spread, uint64_t max_delay) {
    uint32_t kmax = 0, phase = 0;
    if (ntrials == 0 || !dd_shape(nchans, maxval, pitch, spread, n, max_delay, &kmax, &phase)) return false;
    const size_t nblk = (ntrials + DD_TRIALS - 1) / DD_TRIALS;
    const size_t plan_shm = ddp_plan_shm(nchans);
    int32_t* bases = table;
    uint32_t* plan = (uint32_t*)(table + nblk * nchans);
    const uint64_t pw = ddp_plan_words(nchans);
    if (plan_shm > 64 * 1024)
        (void)hipFuncSetAttribute(reinterpret_cast<const void*>(k_dd_plan),
                                  hipFuncAttributeMaxDynamicSharedMemorySize, (int)plan_shm);
    hipLaunchKernelGGL(k_dd_plan, dim3((unsigned)nblk), dim3(256), plan_shm, s, delays, nchans, ntrials,
                       kmax, (uint32_t)DD_ROW_BYTES, bases, plan, pw);
    const uint32_t ntiles = (n + DD_TILE - 1) / DD_TILE;
    const uint64_t total = (uint64_t)ntiles * nblk;
    hipLaunchKernelGGL(k_dedisp_u8, dim3((unsigned)total), dim3(DD_THREADS), 0, s, data, pitch, phase, plan, pw,
                       ntrials, (uint32_t)nblk, ntiles, n, out, out_pitch, fp16, m0, inv_s0, partial);
    return true;
}

// ------------------------------------------- many observations per launch
// Index of the job whose range [first[j], first[j + 1]) holds i (first[0] = 0,
// ascending; njobs entries): a binary search over scalar-loaded offsets.
__device__ __forceinline__ uint32_t job_of(const uint32_t* __restrict__ first, uint32_t njobs, uint32_t i) {
    uint32_t lo = 0, hi = njobs;
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (first[mid] <= i) lo = mid;
        else hi = mid;
    }
    return lo;
}

// Delay tables of every job: entry e of job j is trial e / nchans, channel
// e % nchans, computed exactly as k_delays.
__global__ __launch_bounds__(256) void k_delays_multi(const DedispJob* __restrict__ jobs,
                                                      const uint32_t* __restrict__ entry0, uint32_t njobs,
                                                      uint32_t total) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= total) return;
    const DedispJob& j = jobs[job_of(entry0, njobs, i)];
    const uint32_t e = i - j.entry0;
    const uint32_t d = e / j.nchans, c = e % j.nchans;
    const double k_dm = 4148.808 * j.dms[d];
    const double samples = round(k_dm * (j.inv_f2[c] - j.ref_inv) / j.tsamp);
    j.delays[e] = (int32_t)fmax(samples, 0.0);
}

// Plans of every job's DM blocks: one workgroup per block of the batch.
__global__ __launch_bounds__(256) void k_dd_plan_multi(const DedispJob* __restrict__ jobs,
                                                       const uint32_t* __restrict__ blk0, uint32_t njobs) {
    const DedispJob& j = jobs[job_of(blk0, njobs, blockIdx.x)];
    dd_plan_block(j.delays, j.nchans, j.ntrials, j.kmax, (uint32_t)DD_ROW_BYTES, j.bases, j.plan, j.plan_words,
                  blockIdx.x - j.blk0);
}

// Persistent dedispersion over the work items of many observations: item i of
// job j covers DM block (i - item0) % nblk and `span` consecutive time tiles
// starting at span * ((i - item0) / nblk).  The items are split into one
// contiguous queue per XCD (queue x: items [x Q, (x + 1) Q)); a workgroup pops
// items from its own XCD's queue (workgroups are dispatched round-robin over
// the XCDs, so blockIdx.x % 8 is its XCD) and, once that queue is empty,
// steals from the next ones, so uneven observations balance at the end.
// Every workgroup exits once all queues are empty.
__global__ __launch_bounds__(DD_THREADS, 4) void k_dedisp_queue(const DedispJob* __restrict__ jobs,
                                                               const uint32_t* __restrict__ item0,
                                                               uint32_t njobs, uint32_t total,
                                                               uint32_t span, int fp16,
                                                               uint32_t* __restrict__ counters) {
    __shared__ uint32_t s_item;
    const uint32_t home = blockIdx.x % DD_XCDS;
    const uint32_t Q = (total + DD_XCDS - 1) / DD_XCDS;
    uint32_t tried = 0;  // thread 0: queues found empty so far
    for (;;) {
        __syncthreads();  // the previous item is done with every LDS buffer
        if (threadIdx.x == 0) {
            uint32_t it = 0xffffffffu;
            while (tried < (uint32_t)DD_XCDS) {
                const uint32_t qq = (home + tried) % DD_XCDS;
                const uint32_t beg = qq * Q, end = min(beg + Q, total);
                const uint32_t v = beg < end ? atomicAdd(&counters[qq], 1u) : end - beg;
                if (v < end - beg) {
                    it = beg + v;
                    break;
                }
                tried++;
            }
            s_item = it;
        }
        __syncthreads();
        const uint32_t item = __builtin_amdgcn_readfirstlane(s_item);
        if (item == 0xffffffffu) break;
        const DedispJob& j = jobs[job_of(item0, njobs, item)];
        const uint32_t local = item - j.item0;
        const uint32_t blk = local % j.nblk, tg = local / j.nblk;
        for (uint32_t k = 0; k < span; k++) {
            const uint32_t tile = tg * span + k;
            if (tile >= j.ntiles) break;
            if (k) __syncthreads();
            dd_item(j.data, j.pitch, j.phase_slots, j.plan, j.plan_words, j.ntrials,
                    j.ntiles, j.n, j.out, j.out_pitch, fp16, j.m0, j.inv_s0, j.partial, tile, blk);
        }
    }
}

void launch_dedisp_batch(hipStream_t s, const DedispJob* jobs, const uint32_t* entry0,
                         const uint32_t* blk0, const uint32_t* item0, uint32_t njobs,
                         uint32_t total_entries, uint32_t total_blocks, uint32_t total_items,
                         uint32_t span, int fp16, uint32_t max_nchans, uint32_t grid,
                         uint32_t* counters) {
    if (total_entries)
        hipLaunchKernelGGL(k_delays_multi, dim3((total_entries + 255) / 256), dim3(256), 0, s, jobs, entry0,
                           njobs, total_entries);
    if (!total_blocks || !total_items || !grid) return;
    const size_t plan_shm = ddp_plan_shm(max_nchans);
    if (plan_shm > 64 * 1024)
        (void)hipFuncSetAttribute(reinterpret_cast<const void*>(k_dd_plan_multi),
                                  hipFuncAttributeMaxDynamicSharedMemorySize, (int)plan_shm);
    hipLaunchKernelGGL(k_dd_plan_multi, dim3(total_blocks), dim3(256), plan_shm, s, jobs, blk0, njobs);
    launch_fill_u32(s, counters, DD_XCDS, 0u);
    hipLaunchKernelGGL(k_dedisp_queue, dim3(grid), dim3(DD_THREADS), 0, s, jobs, item0, njobs,
                       total_items, span, fp16, counters);
}

bool dedisp_job_setup(uint32_t nchans, uint32_t maxval, uint64_t pitch, const uint32_t* spread, uint32_t n,
                      uint64_t max_delay, DedispJob* j) {
    uint32_t kmax = 0, phase = 0;
    if (!dd_shape(nchans, maxval, pitch, spread, n, max_delay, &kmax, &phase)) return false;
    j->kmax = kmax;
    j->phase_slots = phase;
    j->plan_words = ddp_plan_words(nchans);
    return true;
}

uint32_t dedisp_tiles(uint32_t n) { return (n + DD_TILE - 1) / DD_TILE; }
size_t dedisp_plan_words(size_t nchans) { return ddp_plan_words((uint32_t)nchans); }

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
