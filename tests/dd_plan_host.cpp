// Host check of the u8 dedispersion's plan and row build: the planner
// (dd_plan_block), the shape check (dd_shape) and the row build (dd_build) of
// pulsarsr_amd/csrc/k_dedisperse.hip, compiled for the CPU (one "thread" runs
// every per-thread loop), checked record by record and, for the tiles asked
// for, by summing every trial's slots from the built rows at the planned LDS
// addresses against the brute-force channel sums.  The kernel text is
// included from KERNEL_PART (written by tests/test_dd_plan_host.py from the
// source file), with the shims below standing in for the HIP builtins.
//
// usage: dd_plan_host INPUT   (binary, see read_input)
#include <sys/mman.h>
#include <unistd.h>

#include <algorithm>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#define __device__
#define __forceinline__ inline
#define __shared__
#define __restrict__
struct Dim3 {
    uint32_t x;
};
static Dim3 threadIdx = {0}, blockDim = {1};
static inline void __syncthreads() {}
template <class T>
static inline T min(T a, T b) {
    return a < b ? a : b;
}
template <class T>
static inline T max(T a, T b) {
    return a > b ? a : b;
}
static inline uint32_t __builtin_amdgcn_readfirstlane(uint32_t x) { return x; }
// v_perm_b32: byte i of the result is byte sel_i of {a (bytes 4..7), b (bytes 0..3)};
// selector 0x0c gives 0x00 (the only constant selector the kernel uses)
static inline uint32_t __builtin_amdgcn_perm(uint32_t a, uint32_t b, uint32_t sel) {
    const uint64_t v = (uint64_t)a << 32 | b;
    uint32_t r = 0;
    for (int i = 0; i < 4; i++) {
        const uint32_t s = sel >> (8 * i) & 0xffu;
        uint32_t byte;
        if (s < 8) byte = (uint32_t)(v >> (8 * s)) & 0xffu;
        else if (s == 0x0c) byte = 0;
        else {
            fprintf(stderr, "unexpected byte selector %#x\n", s);
            exit(3);
        }
        r |= byte << (8 * i);
    }
    return r;
}
struct uint2 {
    uint32_t x, y;
};
struct uint4 {
    uint32_t x, y, z, w;
};
static inline uint2 make_uint2(uint32_t a, uint32_t b) { return {a, b}; }
static inline uint4 make_uint4(uint32_t a, uint32_t b, uint32_t c, uint32_t d) { return {a, b, c, d}; }
constexpr uint32_t DEDISP_BLOCK_TRIALS = 64;
constexpr uint64_t DD_ALLOC_SLACK = 4096;
// the planner's dynamic LDS (extern __shared__ sbase[] in the kernel; the
// shape check caps it at 160 KiB)
int32_t g_plan_lds[160 * 1024 / 4];
#define sbase g_plan_lds

#include KERNEL_PART

#undef sbase

namespace {

struct Input {
    uint32_t nchans, ntrials, nsamples, maxval, seed;
    uint64_t max_delay;
    std::vector<uint32_t> tiles, spread;
    std::vector<int32_t> delays;  // [ntrials][nchans], trials ascending
};

bool read_input(const char* path, Input* in) {
    FILE* f = fopen(path, "rb");
    if (!f) return false;
    uint32_t h[6];
    bool ok = fread(h, 4, 6, f) == 6;
    in->nchans = h[0];
    in->ntrials = h[1];
    in->nsamples = h[2];
    in->maxval = h[3];
    in->seed = h[4];
    const uint32_t ntiles = h[5];
    ok = ok && fread(&in->max_delay, 8, 1, f) == 1;
    in->tiles.resize(ntiles);
    in->spread.resize(in->nchans);
    in->delays.resize((size_t)in->ntrials * in->nchans);
    ok = ok && fread(in->tiles.data(), 4, ntiles, f) == ntiles;
    ok = ok && fread(in->spread.data(), 4, in->nchans, f) == in->nchans;
    ok = ok && fread(in->delays.data(), 4, in->delays.size(), f) == in->delays.size();
    fclose(f);
    return ok;
}

int fail(const char* what, uint32_t blk, uint32_t st) {
    printf("FAIL block %u stage %u: %s\n", blk, st, what);
    return 1;
}

}  // namespace

int main(int argc, char** argv) {
    Input in;
    if (argc != 2 || !read_input(argv[1], &in)) {
        fprintf(stderr, "usage: dd_plan_host INPUT\n");
        return 2;
    }
    const uint32_t nchans = in.nchans, ntrials = in.ntrials;
    const uint64_t pitch = ((uint64_t)in.nsamples + 15) & ~15ull;
    const uint32_t n = (uint32_t)(in.nsamples - (in.max_delay - 1));
    uint32_t kmax = 0, phase = 0;
    if (!dd_shape(nchans, in.maxval, pitch, in.spread.data(), n, in.max_delay, &kmax, &phase)) {
        printf("SHAPE rejected\n");
        return 4;
    }
    // channel rows with the observation's slack; the allocation ends at a
    // PROT_NONE page, so a read past the slack faults
    const size_t bytes = nchans * pitch + DD_ALLOC_SLACK, page = (size_t)sysconf(_SC_PAGESIZE);
    const size_t span = (bytes + page - 1) / page * page;
    uint8_t* base = (uint8_t*)mmap(nullptr, span + page, PROT_READ | PROT_WRITE, MAP_PRIVATE | MAP_ANONYMOUS, -1, 0);
    if (base == MAP_FAILED || mprotect(base + span, page, PROT_NONE)) return 2;
    uint8_t* data = base + span - bytes;
    uint64_t x = in.seed * 0x9e3779b97f4a7c15ull + 1;
    for (size_t i = 0; i < bytes; i++) {
        x ^= x << 13;
        x ^= x >> 7;
        x ^= x << 17;
        data[i] = (uint8_t)(x % (in.maxval + 1));
    }
    const uint32_t nblk = (ntrials + DD_TRIALS - 1) / DD_TRIALS;
    const uint64_t pw = ddp_plan_words(nchans);
    std::vector<uint32_t> plan((size_t)nblk * pw, 0xa5a5a5a5u);
    std::vector<int32_t> bases((size_t)nblk * nchans);
    uint64_t nstages = 0, nrows = 0, nslots = 0, nquads = 0;
    for (uint32_t blk = 0; blk < nblk; blk++) {
        std::fill(g_plan_lds, g_plan_lds + ddp_plan_shm(nchans) / 4, 0x5a5a5a5a);
        dd_plan_block(in.delays.data(), nchans, ntrials, kmax, DD_ROW_BYTES, bases.data(), plan.data(), pw, blk);
        const uint32_t* pb = plan.data() + (size_t)blk * pw;
        const uint32_t nst = pb[0];
        if (nst == 0 || nst > nchans) return fail("stage count", blk, 0);
        // every trial sums every channel exactly once, at its own delay
        std::vector<uint32_t> seen((size_t)DD_TRIALS * nchans, 0);
        for (uint32_t st = 0; st < nst; st++) {
            const uint32_t* sr = pb + DDP_PLAN_HEAD + (size_t)st * DDP_STAGE_PITCH;
            const uint32_t nslot = sr[0] >> 8 & 0xffu, nrow = sr[0] >> 16, nq = sr[1];
            if (!nslot || nslot > (uint32_t)DDP_MAX_SLOTS || !nrow || nrow > (uint32_t)DDP_MAX_ROWS ||
                !nq || nq > (uint32_t)DDP_MAX_QUADS)
                return fail("stage header", blk, st);
            nstages++;
            nrows += nrow;
            nslots += nslot;
            nquads += nq;
            // rows packed in record order, each quad mapped to its row
            const uint8_t* qmap = (const uint8_t*)(sr + DDP_QMAP_REC);
            uint32_t q = 0;
            for (uint32_t r = 0; r < nrow; r++) {
                const uint32_t* rec = sr + DDP_ROW_REC + 4 * r;
                const uint32_t off = rec[0] & 0xffffu, rq = rec[0] >> 16;
                if (off != 64 * q || rq == 0) return fail("row packing", blk, st);
                if ((rec[1] & 0xffffu) >= nchans || ((rec[1] >> 16 & 1u) && (rec[1] & 0xffffu) + 1 >= nchans))
                    return fail("row channel", blk, st);
                for (uint32_t i = 0; i < rq; i++)
                    if (qmap[q + i] != r) return fail("quad map", blk, st);
                q += rq;
            }
            if (q != nq) return fail("quad count", blk, st);
            const uint16_t* lad = (const uint16_t*)(sr + DDP_LAD_REC);
            for (uint32_t s = 0; s < nslot; s++)
                for (uint32_t m = 0; m < (uint32_t)DD_TRIALS; m++) {
                    const uint32_t a = lad[s * DD_TRIALS + m];
                    if (a % 16 || a + 16 * 64 > 64 * nq) return fail("read address past the rows", blk, st);
                    uint32_t r = 0;
                    while (r + 1 < nrow && (sr[DDP_ROW_REC + 4 * (r + 1)] & 0xffffu) <= a) r++;
                    const uint32_t* rec = sr + DDP_ROW_REC + 4 * r;
                    const uint32_t word = (a - (rec[0] & 0xffffu)) / 16;
                    if (a + 16 * 64 > (rec[0] & 0xffffu) + 64 * (rec[0] >> 16))
                        return fail("read runs past its row", blk, st);
                    if (blk * DD_TRIALS + m >= ntrials) continue;  // the partial block repeats its last trial
                    const uint32_t c = rec[1] & 0xffffu;
                    const bool paired = rec[1] >> 16 & 1u;
                    const int32_t* d = &in.delays[(size_t)(blk * DD_TRIALS + m) * nchans];
                    if ((int32_t)rec[2] + (int32_t)word != d[c]) return fail("channel c delay", blk, st);
                    seen[(size_t)m * nchans + c]++;
                    if (paired) {
                        if ((int32_t)rec[3] + (int32_t)word != d[c + 1]) return fail("channel c+1 delay", blk, st);
                        seen[(size_t)m * nchans + c + 1]++;
                    }
                }
        }
        for (uint32_t m = 0; m < (uint32_t)DD_TRIALS && blk * DD_TRIALS + m < ntrials; m++)
            for (uint32_t c = 0; c < nchans; c++)
                if (seen[(size_t)m * nchans + c] != 1) return fail("channel not summed exactly once", blk, 0);
    }
    // the built rows, summed at the planned addresses, against brute force
    uint64_t checked = 0;
    std::vector<uint32_t> rows(DD_ROW_BYTES / 4);
    const uint32_t ntiles = (n + DD_TILE - 1) / DD_TILE;
    for (const uint32_t tile : in.tiles) {
        if (tile >= ntiles) continue;
        const uint32_t t0 = tile * DD_TILE;
        for (uint32_t blk = 0; blk < nblk; blk++) {
            const uint32_t* pb = plan.data() + (size_t)blk * pw;
            std::vector<uint64_t> tot((size_t)DD_TRIALS * DD_TILE, 0);
            for (uint32_t st = 0; st < pb[0]; st++) {
                const uint32_t* sr = pb + DDP_PLAN_HEAD + (size_t)st * DDP_STAGE_PITCH;
                std::fill(rows.begin(), rows.end(), 0xdeadbeefu);
                dd_build(sr, rows.data(), data + t0, pitch, 0, 1);
                const uint32_t nslot = sr[0] >> 8 & 0xffu;
                const uint16_t* lad = (const uint16_t*)(sr + DDP_LAD_REC);
                for (uint32_t s = 0; s < nslot; s++)
                    for (uint32_t m = 0; m < (uint32_t)DD_TRIALS; m++) {
                        const uint8_t* w = (const uint8_t*)rows.data() + lad[s * DD_TRIALS + m];
                        for (uint32_t l = 0; l < 64; l++)
                            for (uint32_t k = 0; k < 8; k++) {
                                uint16_t v;
                                memcpy(&v, w + 16 * l + 2 * k, 2);
                                tot[(size_t)m * DD_TILE + l + 64 * k] += v;
                            }
                    }
            }
            for (uint32_t m = 0; m < (uint32_t)DD_TRIALS && blk * DD_TRIALS + m < ntrials; m++) {
                const int32_t* d = &in.delays[(size_t)(blk * DD_TRIALS + m) * nchans];
                for (uint32_t o = 0; o < (uint32_t)DD_TILE && t0 + o < n; o++) {
                    uint64_t ref = 0;
                    for (uint32_t c = 0; c < nchans; c++) ref += data[c * pitch + t0 + o + (uint32_t)d[c]];
                    checked++;
                    if (ref != tot[(size_t)m * DD_TILE + o]) {
                        printf("FAIL tile %u trial %u output %u: rows sum to %llu, channels to %llu\n", tile,
                               blk * DD_TRIALS + m, t0 + o, (unsigned long long)tot[(size_t)m * DD_TILE + o],
                               (unsigned long long)ref);
                        return 1;
                    }
                }
            }
        }
    }
    printf("OK blocks %u kmax %u phase %u stages/block %.2f rows/stage %.2f slots/block %.1f quads/stage %.1f "
           "checked %llu\n",
           nblk, kmax, phase, (double)nstages / nblk, (double)nrows / nstages, (double)nslots / nblk,
           (double)nquads / nstages, (unsigned long long)checked);
    munmap(base, span + page);
    return 0;
}
