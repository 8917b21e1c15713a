// Many observations in one search (include/psr_gpu.h: psr_batch_*).
//
// The reference runs its per-trial loop (crates/cli/src/main.rs:462-469) once
// per observation; its many-observation harness (crates/synth/examples/
// confirmation.rs) loops over observations one after another.  Here every
// observation's DM trials join one device schedule:
//
// - observations are grouped by FFT length (classes); within a class their
//   trials are cut into rounds of whole 64-trial DM blocks that fit the memory
//   budget, so one round mixes several observations;
// - per round, one launch per stage: the delay tables of every observation, the
//   block plans, ONE persistent dedispersion kernel that pops (observation,
//   DM block, time-tile span) items from per-XCD work queues, the trial means,
//   the FFT, the whitening, the harmonic sums and the ranking, each over all
//   trials of the round (per-trial lengths and search limits from tables);
// - the whole schedule of every round is captured once into a hipGraph and
//   replayed by psr_batch_run;
// - an autotuner times candidate shapes on the batch's own data before the
//   capture: the FFT's four-step split (pass-1 length m2) and trials per
//   FFT round per class, and the dedispersion's time-tile span per work item
//   and persistent workgroups per CU.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <functional>
#include <numeric>
#include <string>
#include <vector>

#include "ctx_internal.hpp"

using namespace psr;

namespace {

struct ObsPlan {
    psr_obs* obs = nullptr;
    size_t ntrials = 0, out0 = 0;     // caller trials and their first output row
    size_t n = 0;                     // series length
    std::vector<size_t> order;        // sorted position -> caller trial index
    std::vector<double> sorted;       // ascending DMs
    std::vector<uint32_t> spread;     // per-channel delay spread bound over 64-trial blocks
    SearchParams p{};
    std::vector<uint32_t> birdies;
    int cls = -1;
    bool u8 = false;                  // LDS-staged kernel (else channel-ordered f32 kernel)
    DedispJob proto{};                // kmax, phase_slots, plan_words when u8
};

struct Segment {  // trials [t0, t0 + nt) (sorted order) of one observation
    int obs;
    size_t t0, nt;
};

struct Round {
    int cls;
    std::vector<Segment> segs;
    uint32_t ntrials = 0, njobs = 0;
    uint32_t total_entries = 0, total_blocks = 0, total_items = 0, max_nchans = 0, max_tiles_hs = 0;
    size_t trial_base = 0;            // first row of this round in the pinned output
    size_t table_off = 0;             // byte offset of this round's tables in the table buffer
    std::vector<int> f32_jobs;        // jobs dedispersed by the channel-ordered kernel
};

struct Class {
    uint32_t nfft = 0;
    SpecGeom g{};                     // tuned split; series_pitch = widest of the class
    uint32_t fft_batch = 0;           // trials per pass-1/pass-2 round (0 = default)
    DevBuf tw;                        // twiddles of the geometry
    double tuned_ms = 0.0;
};

// byte offsets of a round's tables inside the table buffer
struct TableLayout {
    size_t jobs, entry0, blk0, item0, trial_job, trial_n, dms, harm, harm_job, total;
};

TableLayout table_layout(uint32_t njobs, uint32_t ntrials) {
    TableLayout L{};
    size_t off = 0;
    auto take = [&](size_t b) {
        const size_t at = off;
        off += (b + 255) & ~(size_t)255;
        return at;
    };
    L.jobs = take(njobs * sizeof(DedispJob));
    L.entry0 = take(njobs * 4);
    L.blk0 = take(njobs * 4);
    L.item0 = take(njobs * 4);
    L.trial_job = take((size_t)ntrials * 4);
    L.trial_n = take((size_t)ntrials * 4);
    L.dms = take((size_t)ntrials * 8);
    L.harm = take(njobs * harm_args_bytes());
    L.harm_job = take((size_t)ntrials * 4);
    L.total = off;
    return L;
}

}  // namespace

struct psr_batch {
    psr_ctx* ctx = nullptr;
    uint32_t flags = 0, fp16 = 0;
    size_t cap = 0, total_trials = 0;
    std::vector<ObsPlan> obs;
    std::vector<Class> classes;
    std::vector<Round> rounds;
    // dedispersion shape (tuned)
    uint32_t span = 1, wg_per_cu = 2, cus = 256;
    // device buffers sized for the largest round
    DevBuf tables, delays, ddtab, series, partial, mean32, corr16, work, spec_t, raw_nat, scales, sorted,
        floors, fft_partial, peaks, peak_counts, cands, counts, counters, costabs;
    size_t ddtab_words = 0;
    void* pinned = nullptr;  // [total_trials * cap] candidates, then counts + one overflow word per round
    size_t pinned_bytes = 0;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    std::string tuning;  // JSON fragment of the autotuner's choices
    double last_ms = 0.0;
};

namespace {

int batch_fail_cleanup(psr_batch* b, int rc) {
    psr_batch_destroy(b);
    return rc;
}

size_t es_of(const psr_batch* b) { return b->fp16 ? 2 : 4; }

// Device bytes one trial of class c needs for the spectral stages and outputs.
size_t spectral_trial_bytes(const psr_batch* b, const SpecGeom& g) {
    const size_t es = es_of(b);
    const uint32_t npart = g.m1 == 1 ? 1u : g.m2 / 2 + 1;
    return (size_t)g.series_pitch * es + 2 * (size_t)g.spec_pitch * es +
           (whiten_fused(g) ? 0 : (size_t)g.spec_pitch * es) + (size_t)g.nblocks * 4 + npart * 8 +
           (size_t)PEAK_CAP * sizeof(RawPeak) + b->cap * sizeof(CandidateOut) + 64;
}

// Cut every class's observations into rounds within the memory budget.
void plan_rounds(psr_batch* b, size_t budget) {
    b->rounds.clear();
    for (int c = 0; c < (int)b->classes.size(); c++) {
        const SpecGeom& g = b->classes[c].g;
        const size_t per_trial = spectral_trial_bytes(b, g);
        Round r;
        r.cls = c;
        size_t used = 0;
        auto flush = [&]() {
            if (!r.segs.empty()) b->rounds.push_back(r);
            r = Round();
            r.cls = c;
            used = 0;
        };
        for (int i = 0; i < (int)b->obs.size(); i++) {
            const ObsPlan& o = b->obs[i];
            if (o.cls != c) continue;
            const size_t nch = o.obs->nchans;
            const size_t tiles = (o.n + 255) / 256;  // f32-kernel tiles bound the partial rows
            const size_t trial_bytes = per_trial + nch * 4 + tiles * 8 + 64 +
                                       (nch + dedisp_plan_words(nch)) * 4 / DEDISP_BLOCK_TRIALS;
            size_t t = 0;
            while (t < o.ntrials) {
                size_t room = used < budget ? (budget - used) / trial_bytes : 0;
                if (room < DEDISP_BLOCK_TRIALS && !r.segs.empty()) {
                    flush();
                    continue;
                }
                room = std::max<size_t>(room, DEDISP_BLOCK_TRIALS);
                size_t nt = std::min(o.ntrials - t, room);
                if (nt < o.ntrials - t) nt -= nt % DEDISP_BLOCK_TRIALS;
                r.segs.push_back(Segment{i, t, nt});
                used += nt * trial_bytes;
                t += nt;
            }
        }
        flush();
    }
}

}  // namespace

namespace {

// Fill the round's host tables (jobs etc.) given the device buffer bases.
struct HostTables {
    std::vector<DedispJob> jobs;
    std::vector<uint32_t> entry0, blk0, item0, trial_job, trial_n, harm_job;
    std::vector<double> dms;
    std::vector<unsigned char> harm;
};

void build_tables(psr_batch* b, Round& r, HostTables& h) {
    const Class& cl = b->classes[r.cls];
    const SpecGeom& g = cl.g;
    const size_t es = es_of(b);
    h = HostTables();
    uint32_t entry = 0, blk = 0, item = 0, trial = 0;
    size_t ddw = 0, part = 0;
    r.max_nchans = 0;
    r.max_tiles_hs = 0;
    r.f32_jobs.clear();
    std::vector<int> harm_of_obs(b->obs.size(), -1);
    for (size_t si = 0; si < r.segs.size(); si++) {
        const Segment& sg = r.segs[si];
        const ObsPlan& o = b->obs[sg.obs];
        psr_obs* ob = o.obs;
        DedispJob j = o.proto;
        j.data = (const uint8_t*)ob->data;
        j.pitch = ob->pitch;
        j.dms = (const double*)((char*)b->tables.p + r.table_off) + 0;  // patched below
        j.inv_f2 = ob->inv_f2;
        j.ref_inv = ob->ref_inv;
        j.tsamp = ob->tsamp;
        j.nchans = ob->nchans;
        j.ntrials = (uint32_t)sg.nt;
        j.nblk = (uint32_t)((sg.nt + DEDISP_BLOCK_TRIALS - 1) / DEDISP_BLOCK_TRIALS);
        j.n = (uint32_t)o.n;
        j.ntiles = o.u8 ? dedisp_tiles(j.n) : (j.n + 255) / 256;
        j.out_pitch = g.series_pitch;
        j.m0 = (float)ob->m0;
        j.inv_s0 = (float)(1.0 / ob->s0);
        j.m0d = ob->m0;
        j.inv_s0d = 1.0 / ob->s0;
        j.entry0 = entry;
        j.trial0 = trial;
        j.delays = (int32_t*)b->delays.p + entry;
        j.out = (char*)b->series.p + (size_t)trial * g.series_pitch * es;
        j.partial = (double*)b->partial.p + part;
        if (o.u8) {
            j.blk0 = blk;
            j.item0 = item;
            j.bases = (int32_t*)b->ddtab.p + ddw;
            j.plan = (uint32_t*)((int32_t*)b->ddtab.p + ddw + (size_t)j.nblk * j.nchans);
            ddw += (size_t)j.nblk * j.nchans + (size_t)j.nblk * j.plan_words;
            blk += j.nblk;
            item += j.nblk * ((j.ntiles + b->span - 1) / b->span);
            r.max_nchans = std::max(r.max_nchans, j.nchans);
        } else {
            j.nblk = 0;  // no plan blocks, no work items
            j.blk0 = blk;
            j.item0 = item;
            j.bases = nullptr;
            j.plan = nullptr;
            r.f32_jobs.push_back((int)si);
        }
        h.entry0.push_back(j.entry0);
        h.blk0.push_back(j.blk0);
        h.item0.push_back(j.item0);
        if (harm_of_obs[sg.obs] < 0) {
            harm_of_obs[sg.obs] = (int)(h.harm.size() / harm_args_bytes());
            h.harm.resize(h.harm.size() + harm_args_bytes());
            fill_harm_args(o.p, o.birdies.data(), h.harm.data() + h.harm.size() - harm_args_bytes());
            r.max_tiles_hs = std::max(r.max_tiles_hs, harmsum_tiles(o.p));
        }
        for (size_t t = 0; t < sg.nt; t++) {
            h.trial_job.push_back((uint32_t)si);
            h.trial_n.push_back(j.n);
            h.harm_job.push_back((uint32_t)harm_of_obs[sg.obs]);
            h.dms.push_back(o.sorted[sg.t0 + t]);
        }
        entry += j.ntrials * j.nchans;
        trial += j.ntrials;
        part += (size_t)j.ntrials * j.ntiles;
        h.jobs.push_back(j);
    }
    r.njobs = (uint32_t)h.jobs.size();
    r.ntrials = trial;
    r.total_entries = entry;
    r.total_blocks = blk;
    r.total_items = item;
    const TableLayout L = table_layout(r.njobs, r.ntrials);
    const double* ddms = (const double*)((char*)b->tables.p + r.table_off + L.dms);
    for (uint32_t k = 0; k < r.njobs; k++) h.jobs[k].dms = ddms + h.jobs[k].trial0;
}

// Sizes of the round-shared device buffers over every round.
struct Need {
    size_t delays = 0, ddtab = 0, series = 0, partial = 0, trials = 0, tables = 0;
};

Need need_of(psr_batch* b) {
    Need nd;
    for (Round& r : b->rounds) {
        const SpecGeom& g = b->classes[r.cls].g;
        size_t ent = 0, ddw = 0, part = 0, tr = 0;
        for (const Segment& sg : r.segs) {
            const ObsPlan& o = b->obs[sg.obs];
            const size_t nch = o.obs->nchans, nblk = (sg.nt + DEDISP_BLOCK_TRIALS - 1) / DEDISP_BLOCK_TRIALS;
            ent += sg.nt * nch;
            if (o.u8) ddw += nblk * nch + nblk * o.proto.plan_words;
            part += sg.nt * (o.u8 ? dedisp_tiles((uint32_t)o.n) : (o.n + 255) / 256);
            tr += sg.nt;
        }
        nd.delays = std::max(nd.delays, ent);
        nd.ddtab = std::max(nd.ddtab, ddw);
        nd.series = std::max(nd.series, tr * g.series_pitch * es_of(b));
        nd.partial = std::max(nd.partial, part);
        nd.trials = std::max(nd.trials, tr);
        r.table_off = nd.tables;
        nd.tables += table_layout((uint32_t)r.segs.size(), (uint32_t)tr).total;
    }
    return nd;
}

int alloc_buffers(psr_batch* b) {
    psr_ctx* c = b->ctx;
    const Need nd = need_of(b);
    size_t spec = 0, blocks = 0, npart = 0, work = 8;
    bool unfused = (b->flags & PSR_BATCH_AUTOTUNE) != 0;  // conservative: a tried split may not fuse
    for (const Round& r : b->rounds) {
        const Class& cl = b->classes[r.cls];
        const SpecGeom& g = cl.g;
        const size_t nt = r.ntrials ? r.ntrials : 0;
        size_t tr = 0;
        for (const Segment& sg : r.segs) tr += sg.nt;
        spec = std::max(spec, tr * (size_t)g.spec_pitch);
        blocks = std::max(blocks, tr * (size_t)g.nblocks);
        // (the tuner may pick a pass-1 length up to 2048)
        const size_t m2max = (b->flags & PSR_BATCH_AUTOTUNE) ? 2048 : g.m2;
        npart = std::max(npart, tr * (size_t)(g.m1 == 1 ? 1 : m2max / 2 + 1));
        if (!whiten_fused(g)) unfused = true;
        const uint32_t fb = cl.fft_batch ? cl.fft_batch : fft_work_trials(g, (uint32_t)tr);
        work = std::max(work, g.m1 == 1 ? (size_t)8 : (size_t)std::min<size_t>(fb, tr) * g.m * sizeof(float2));
        (void)nt;
    }
    // the tuner tries FFT rounds up to 384 MiB of intermediate
    if (b->flags & PSR_BATCH_AUTOTUNE) work = std::max(work, (size_t)384 << 20);
    const size_t es = es_of(b);
    int rc;
    if ((rc = ensure(c, b->tables, nd.tables + 256)) || (rc = ensure(c, b->delays, nd.delays * 4 + 64)) ||
        (rc = ensure(c, b->ddtab, nd.ddtab * 4 + 64)) || (rc = ensure(c, b->series, nd.series + 64)) ||
        (rc = ensure(c, b->partial, nd.partial * 8 + 64)) || (rc = ensure(c, b->mean32, nd.trials * 4 + 64)) ||
        (rc = ensure(c, b->corr16, nd.trials * 4 + 64)) || (rc = ensure(c, b->work, work)) ||
        (rc = ensure(c, b->spec_t, spec * es + 64)) || (rc = ensure(c, b->raw_nat, spec * es + 64)) ||
        (rc = ensure(c, b->sorted, unfused ? spec * es + 64 : 64)) || (rc = ensure(c, b->scales, blocks * 4 + 64)) ||
        (rc = ensure(c, b->floors, nd.trials * 4 + 64)) || (rc = ensure(c, b->fft_partial, npart * 8 + 64)) ||
        (rc = ensure(c, b->peaks, nd.trials * PEAK_CAP * sizeof(RawPeak))) ||
        (rc = ensure(c, b->peak_counts, nd.trials * 4 + 64)) ||
        (rc = ensure(c, b->cands, nd.trials * b->cap * sizeof(CandidateOut))) ||
        (rc = ensure(c, b->counts, (nd.trials + 1) * 4 + 64)) || (rc = ensure(c, b->counters, 256)))
        return rc;
    return PSR_OK;
}

// Upload every round's tables (device pointers are final once the buffers exist).
int upload_tables(psr_batch* b) {
    std::vector<unsigned char> host(b->tables.bytes, 0);
    for (Round& r : b->rounds) {
        HostTables h;
        build_tables(b, r, h);
        const TableLayout L = table_layout(r.njobs, r.ntrials);
        unsigned char* base = host.data() + r.table_off;
        std::memcpy(base + L.jobs, h.jobs.data(), h.jobs.size() * sizeof(DedispJob));
        std::memcpy(base + L.entry0, h.entry0.data(), h.entry0.size() * 4);
        std::memcpy(base + L.blk0, h.blk0.data(), h.blk0.size() * 4);
        std::memcpy(base + L.item0, h.item0.data(), h.item0.size() * 4);
        std::memcpy(base + L.trial_job, h.trial_job.data(), h.trial_job.size() * 4);
        std::memcpy(base + L.trial_n, h.trial_n.data(), h.trial_n.size() * 4);
        std::memcpy(base + L.dms, h.dms.data(), h.dms.size() * 8);
        std::memcpy(base + L.harm, h.harm.data(), h.harm.size());
        std::memcpy(base + L.harm_job, h.harm_job.data(), h.harm_job.size() * 4);
    }
    HIP_TRY(hipMemcpy(b->tables.p, host.data(), host.size(), hipMemcpyHostToDevice));
    return PSR_OK;
}

// The dedispersion stages of round r (delays, plans, persistent kernel, the
// channel-ordered kernel for jobs it cannot take, trial means).
int enqueue_dedisp(psr_batch* b, const Round& r, hipStream_t s) {
    const TableLayout L = table_layout(r.njobs, r.ntrials);
    char* tb = (char*)b->tables.p + r.table_off;
    const DedispJob* jobs = (const DedispJob*)(tb + L.jobs);
    launch_dedisp_batch(s, jobs, (const uint32_t*)(tb + L.entry0), (const uint32_t*)(tb + L.blk0),
                        (const uint32_t*)(tb + L.item0), r.njobs, r.total_entries, r.total_blocks, r.total_items,
                        b->span, (int)b->fp16, std::max<uint32_t>(r.max_nchans, 1), b->wg_per_cu * b->cus,
                        (uint32_t*)b->counters.p);
    HIP_TRY(hipGetLastError());
    if (!r.f32_jobs.empty()) {  // host copies of those jobs' fields
        HostTables h;
        Round rr = r;
        build_tables(b, rr, h);
        for (int k : r.f32_jobs) {
            const DedispJob& j = h.jobs[k];
            const psr_obs* o = b->obs[r.segs[k].obs].obs;
            launch_dedisp_f32(s, o->data, o->kind, o->pitch, j.nchans, j.delays, j.ntrials, j.n, j.out, j.out_pitch,
                              (int)b->fp16, j.m0, j.inv_s0, j.partial);
        }
        HIP_TRY(hipGetLastError());
    }
    launch_trial_mean_multi(s, jobs, (const uint32_t*)(tb + L.trial_job), r.ntrials, (float*)b->mean32.p,
                            (float*)b->corr16.p);
    HIP_TRY(hipGetLastError());
    return PSR_OK;
}

// The spectral stages of round r with its class's (possibly trial) shapes.
int enqueue_spectral_round(psr_batch* b, const Round& r, const Class& cl, hipStream_t s, bool with_copies) {
    const TableLayout L = table_layout(r.njobs, r.ntrials);
    char* tb = (char*)b->tables.p + r.table_off;
    const SpecGeom& g = cl.g;
    const uint32_t nt = r.ntrials;
    const uint32_t npart = g.m1 == 1 ? 1u : g.m2 / 2 + 1;
    launch_fill_u32(s, (uint32_t*)b->peak_counts.p, nt, 0u);
    launch_fill_u32(s, (uint32_t*)b->counts.p, nt + 1, 0u);
    launch_fft(s, g, b->series.p, (int)b->fp16, (const float*)b->mean32.p, (const float*)b->corr16.p, nt,
               (float2*)b->work.p, b->spec_t.p, (double*)b->fft_partial.p, (const float2*)cl.tw.p,
               (const float*)b->costabs.p, (const uint32_t*)(tb + L.trial_n), cl.fft_batch);
    HIP_TRY(hipGetLastError());
    launch_whiten(s, g, b->spec_t.p, (int)b->fp16, nt, (const double*)b->fft_partial.p, npart, b->raw_nat.p,
                  b->sorted.p, (float*)b->floors.p, (float*)b->scales.p);
    HIP_TRY(hipGetLastError());
    launch_harmsum_multi(s, g, (int)b->fp16, r.max_tiles_hs, b->raw_nat.p, (const float*)b->scales.p, tb + L.harm,
                         (const uint32_t*)(tb + L.harm_job), nt, (RawPeak*)b->peaks.p,
                         (uint32_t*)b->peak_counts.p);
    HIP_TRY(hipGetLastError());
    launch_finalize_multi(s, (const double*)(tb + L.dms), nt, (const RawPeak*)b->peaks.p,
                          (const uint32_t*)b->peak_counts.p, tb + L.harm, (const uint32_t*)(tb + L.harm_job),
                          (CandidateOut*)b->cands.p, (uint32_t)b->cap, (uint32_t*)b->counts.p);
    HIP_TRY(hipGetLastError());
    if (with_copies) {
        size_t ri = (size_t)(&r - b->rounds.data());
        char* hc = (char*)b->pinned;
        uint32_t* hcnt = (uint32_t*)(hc + b->total_trials * b->cap * sizeof(psr_candidate));
        HIP_TRY(hipMemcpyAsync(hc + r.trial_base * b->cap * sizeof(psr_candidate), b->cands.p,
                               (size_t)nt * b->cap * sizeof(CandidateOut), hipMemcpyDeviceToHost, s));
        HIP_TRY(hipMemcpyAsync(hcnt + r.trial_base + ri, b->counts.p, (size_t)(nt + 1) * 4, hipMemcpyDeviceToHost,
                               s));
    }
    return PSR_OK;
}

int enqueue_all(psr_batch* b, hipStream_t s) {
    int rc;
    for (const Round& r : b->rounds) {
        if ((rc = enqueue_dedisp(b, r, s))) return rc;
        if ((rc = enqueue_spectral_round(b, r, b->classes[r.cls], s, true))) return rc;
    }
    return PSR_OK;
}

int upload_class_twiddles(psr_batch* b, Class& cl) {
    const size_t nt = twiddle_table_size(cl.g);
    std::vector<float2> t(nt);
    fill_twiddles(cl.g, t.data());
    release(b->ctx, cl.tw);
    int rc = ensure(b->ctx, cl.tw, nt * sizeof(float2));
    if (rc) return rc;
    HIP_TRY(hipMemcpy(cl.tw.p, t.data(), nt * sizeof(float2), hipMemcpyHostToDevice));
    return PSR_OK;
}

float time_ms(hipStream_t s, const std::function<int()>& f, int reps, int* rc_out) {
    hipEvent_t a, z;
    (void)hipEventCreate(&a);
    (void)hipEventCreate(&z);
    float best = 1e30f;
    for (int i = 0; i < reps; i++) {
        (void)hipEventRecord(a, s);
        const int rc = f();
        (void)hipEventRecord(z, s);
        (void)hipEventSynchronize(z);
        if (rc) {
            *rc_out = rc;
            break;
        }
        float ms = 0.f;
        (void)hipEventElapsedTime(&ms, a, z);
        best = std::min(best, ms);
    }
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(z);
    return best;
}

// Autotuner: per class, the four-step split and the trials per FFT round on
// the class's first round; then the dedispersion's time-tile span and
// persistent workgroups per CU on the round with the most work items.
int autotune(psr_batch* b) {
    hipStream_t s = b->ctx->stream;
    int rc = PSR_OK;
    std::string js = "{\"classes\": [";
    for (int ci = 0; ci < (int)b->classes.size(); ci++) {
        Class& cl = b->classes[ci];
        const Round* r0 = nullptr;
        for (const Round& r : b->rounds)
            if (r.cls == ci) {
                r0 = &r;
                break;
            }
        if (!r0) continue;
        if ((rc = enqueue_dedisp(b, *r0, s))) return rc;
        HIP_TRY(hipStreamSynchronize(s));
        const SpecGeom base = cl.g;
        SpecGeom best_g = base;
        uint32_t best_fb = 0;
        float best = 1e30f;
        std::string tried;
        const uint32_t m2s[3] = {512, 1024, 2048};
        const uint64_t fbytes[3] = {96ull << 20, 192ull << 20, 384ull << 20};
        for (int a = 0; a < (base.m1 == 1 ? 1 : 3); a++) {
            SpecGeom g = base.m1 == 1 ? base : make_geometry_split(base.n, m2s[a]);
            g.series_pitch = base.series_pitch;
            if (!fft_shape_supported(g)) continue;
            cl.g = g;
            if ((rc = upload_class_twiddles(b, cl))) return rc;
            for (int f = 0; f < (g.m1 == 1 ? 1 : 3); f++) {
                const uint32_t fb = g.m1 == 1 ? 0u : (uint32_t)std::max<uint64_t>(1, fbytes[f] / ((uint64_t)g.m * 8));
                cl.fft_batch = fb;
                const float ms = time_ms(s, [&]() { return enqueue_spectral_round(b, *r0, cl, s, false); }, 2, &rc);
                if (rc) return rc;
                char buf[96];
                std::snprintf(buf, sizeof buf, "%s{\"m2\": %u, \"fft_trials\": %u, \"ms\": %.3f}",
                              tried.empty() ? "" : ", ", g.m2, fb, ms);
                tried += buf;
                if (ms < best) {
                    best = ms;
                    best_g = g;
                    best_fb = fb;
                }
            }
        }
        cl.g = best_g;
        cl.fft_batch = best_fb;
        cl.tuned_ms = best;
        if ((rc = upload_class_twiddles(b, cl))) return rc;
        char head[160];
        std::snprintf(head, sizeof head, "%s{\"nfft\": %u, \"m1\": %u, \"m2\": %u, \"fft_trials\": %u, \"tried\": [",
                      ci ? ", " : "", cl.nfft, cl.g.m1, cl.g.m2, cl.fft_batch);
        js += head + tried + "]}";
    }
    js += "]";
    // dedispersion shape on the round with the most items
    const Round* rd = nullptr;
    for (const Round& r : b->rounds)
        if (!rd || r.total_items > rd->total_items) rd = &r;
    if (rd && rd->total_items) {
        uint32_t best_span = b->span, best_wg = b->wg_per_cu;
        float best = 1e30f;
        std::string tried;
        const uint32_t spans[3] = {1, 2, 4}, wgs[2] = {1, 2};
        for (uint32_t sp : spans)
            for (uint32_t wg : wgs) {
                b->span = sp;
                b->wg_per_cu = wg;
                if ((rc = upload_tables(b))) return rc;  // item numbering depends on the span
                const float ms = time_ms(s, [&]() { return enqueue_dedisp(b, *rd, s); }, 1, &rc);
                if (rc) return rc;
                char buf[96];
                std::snprintf(buf, sizeof buf, "%s{\"span\": %u, \"wg_per_cu\": %u, \"ms\": %.3f}",
                              tried.empty() ? "" : ", ", sp, wg, ms);
                tried += buf;
                if (ms < best) {
                    best = ms;
                    best_span = sp;
                    best_wg = wg;
                }
            }
        b->span = best_span;
        b->wg_per_cu = best_wg;
        char head[128];
        std::snprintf(head, sizeof head, ", \"dedisperse\": {\"span\": %u, \"wg_per_cu\": %u, \"tried\": [",
                      b->span, b->wg_per_cu);
        js += head + tried + "]}";
    }
    js += "}";
    b->tuning = js;
    return upload_tables(b);
}

}  // namespace

extern "C" {

int psr_batch_create(psr_ctx* c, psr_obs* const* obs, size_t nobs, const double* const* dms,
                     const size_t* ntrials, const size_t* lengths, const psr_search_cfg* cfg,
                     size_t cap_per_trial, const psr_batch_opts* opts, psr_batch** out) {
    clear_error();
    if (!c || !out || !cfg || (nobs && (!obs || !dms || !ntrials))) return fail(PSR_E_INVALID, "null pointer");
    if (cap_per_trial == 0) return fail(PSR_E_INVALID, "cap_per_trial must be positive");
    if (cfg->precision != PSR_PRECISION_FP32 && cfg->precision != PSR_PRECISION_FP16)
        return fail(PSR_E_INVALID, "unknown precision");
    const uint32_t flags = opts ? opts->flags : 0u;
    if (flags & ~(uint32_t)(PSR_BATCH_GRAPH | PSR_BATCH_AUTOTUNE)) return fail(PSR_E_INVALID, "unknown batch flags");
    int rc;
    if ((rc = check_device(c))) return rc;
    psr_batch* b = new psr_batch();
    b->ctx = c;
    b->flags = flags;
    b->fp16 = cfg->precision == PSR_PRECISION_FP16;
    b->cap = cap_per_trial;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, c->device) == hipSuccess && prop.multiProcessorCount > 0)
        b->cus = (uint32_t)prop.multiProcessorCount;
    // per observation: length, geometry class, search limits, kernel choice
    b->obs.resize(nobs);
    for (size_t i = 0; i < nobs; i++) {
        ObsPlan& o = b->obs[i];
        o.obs = obs[i];
        o.ntrials = ntrials[i];
        o.out0 = b->total_trials;
        b->total_trials += o.ntrials;
        if (!o.obs || (o.ntrials && !dms[i])) return batch_fail_cleanup(b, fail(PSR_E_INVALID, "null observation or DM list"));
        if (o.obs->ctx != c) return batch_fail_cleanup(b, fail(PSR_E_INVALID, "observation belongs to another context"));
        if (!(o.obs->tsamp > 0.0) || !std::isfinite(o.obs->tsamp))
            return batch_fail_cleanup(b, fail(PSR_E_CONTRACT, "periodicity requires finite positive tsamp"));
        if (o.ntrials == 0) continue;
        uint64_t maxd;
        if ((rc = resolve_length(o.obs, dms[i], o.ntrials, lengths ? lengths[i] : 0, &o.n, &maxd)))
            return batch_fail_cleanup(b, rc);
        o.order.resize(o.ntrials);
        std::iota(o.order.begin(), o.order.end(), 0);
        const double* d = dms[i];
        std::stable_sort(o.order.begin(), o.order.end(), [&](size_t a, size_t z) { return d[a] < d[z]; });
        o.sorted.resize(o.ntrials);
        for (size_t t = 0; t < o.ntrials; t++) o.sorted[t] = d[o.order[t]];
        const SpecGeom g = make_geometry((uint32_t)o.n);
        if ((rc = make_params(g, o.obs->tsamp, cfg, &o.p, &o.birdies))) return batch_fail_cleanup(b, rc);
        o.birdies.resize(2 * MAX_BIRDIE_INTERVALS, 0u);
        int cls = -1;
        for (int k = 0; k < (int)b->classes.size(); k++)
            if (b->classes[k].nfft == g.nfft) cls = k;
        if (cls < 0) {
            Class cl;
            cl.nfft = g.nfft;
            cl.g = g;
            b->classes.push_back(cl);
            cls = (int)b->classes.size() - 1;
        }
        Class& cl = b->classes[cls];
        if (g.series_pitch > cl.g.series_pitch) cl.g.series_pitch = g.series_pitch;
        if (g.n > cl.g.n) cl.g.n = g.n;
        o.cls = cls;
        o.u8 = false;
        if (o.obs->kind == KIND_U8) {
            o.spread = spread_bounds(o.obs, o.sorted.data(), o.ntrials, DEDISP_BLOCK_TRIALS);
            o.u8 = dedisp_job_setup(o.obs->nchans, o.obs->maxval, o.obs->pitch, o.spread.data(), (uint32_t)o.n,
                                    max_delay_at(o.obs, o.sorted[o.ntrials - 1]), &o.proto);
        }
    }
    for (Class& cl : b->classes) {
        const uint32_t pitch = cl.g.series_pitch;
        cl.g = make_geometry(cl.g.n);  // every member shares nfft, nbins and the split
        cl.g.series_pitch = pitch;
        if (!fft_shape_supported(cl.g))
            return batch_fail_cleanup(b, fail(PSR_E_SHAPE, "no FFT kernels for a " + std::to_string(cl.g.nfft) + "-point transform"));
    }
    // memory budget per round: half the free memory, at most 64 GiB (or the caller's)
    size_t freeb = 0, totalb = 0;
    size_t budget = (size_t)64 << 30;
    if (hipMemGetInfo(&freeb, &totalb) == hipSuccess && freeb) budget = std::min(budget, freeb / 2);
    if (opts && opts->memory_budget) budget = opts->memory_budget;
    if (c->mem_cap) budget = std::min<size_t>(budget, c->mem_cap / 2);
    plan_rounds(b, budget);
    size_t base = 0;
    for (Round& r : b->rounds) {
        r.trial_base = base;
        for (const Segment& sg : r.segs) base += sg.nt;
    }
    if ((rc = alloc_buffers(b))) return batch_fail_cleanup(b, rc);
    {
        std::vector<float> ct(COSTAB_ALL);
        fill_costabs(ct.data());
        if ((rc = ensure(c, b->costabs, ct.size() * sizeof(float)))) return batch_fail_cleanup(b, rc);
        if (hipMemcpy(b->costabs.p, ct.data(), ct.size() * sizeof(float), hipMemcpyHostToDevice) != hipSuccess)
            return batch_fail_cleanup(b, fail(PSR_E_DEVICE, "twiddle upload failed"));
    }
    for (Class& cl : b->classes)
        if ((rc = upload_class_twiddles(b, cl))) return batch_fail_cleanup(b, rc);
    b->pinned_bytes = b->total_trials * cap_per_trial * sizeof(psr_candidate) +
                      (b->total_trials + b->rounds.size() + 1) * sizeof(uint32_t);
    if (hipHostMalloc(&b->pinned, b->pinned_bytes, hipHostMallocDefault) != hipSuccess) {
        b->pinned = nullptr;
        return batch_fail_cleanup(b, fail(PSR_E_OOM, "hipHostMalloc of the batch outputs failed"));
    }
    if ((rc = upload_tables(b))) return batch_fail_cleanup(b, rc);
    if ((flags & PSR_BATCH_AUTOTUNE) && (rc = autotune(b))) return batch_fail_cleanup(b, rc);
    if (flags & PSR_BATCH_GRAPH) {
        // one uncaptured pass first (it also settles the kernels' LDS attributes)
        if ((rc = enqueue_all(b, c->stream))) return batch_fail_cleanup(b, rc);
        if (hipStreamSynchronize(c->stream) != hipSuccess)
            return batch_fail_cleanup(b, fail(PSR_E_DEVICE, "batch warm-up failed"));
        if (hipStreamBeginCapture(c->stream, hipStreamCaptureModeRelaxed) != hipSuccess)
            return batch_fail_cleanup(b, fail(PSR_E_DEVICE, "hipStreamBeginCapture failed"));
        rc = enqueue_all(b, c->stream);
        hipGraph_t gr = nullptr;
        const hipError_t e = hipStreamEndCapture(c->stream, &gr);
        if (rc || e != hipSuccess || !gr) {
            if (gr) (void)hipGraphDestroy(gr);
            return batch_fail_cleanup(b, rc ? rc : fail(PSR_E_DEVICE, std::string("graph capture failed: ") + hipGetErrorString(e)));
        }
        b->graph = gr;
        if (hipGraphInstantiate(&b->exec, gr, nullptr, nullptr, 0) != hipSuccess) {
            b->exec = nullptr;
            return batch_fail_cleanup(b, fail(PSR_E_DEVICE, "hipGraphInstantiate failed"));
        }
    }
    *out = b;
    return PSR_OK;
}

int psr_batch_run(psr_batch* b, psr_candidate* out, uint32_t* counts) {
    clear_error();
    if (!b || (b->total_trials && (!out || !counts))) return fail(PSR_E_INVALID, "null pointer");
    psr_ctx* c = b->ctx;
    int rc;
    if ((rc = check_device(c))) return rc;
    hipEvent_t e0, e1;
    HIP_TRY(hipEventCreate(&e0));
    HIP_TRY(hipEventCreate(&e1));
    HIP_TRY(hipEventRecord(e0, c->stream));
    if (b->exec) {
        HIP_TRY(hipGraphLaunch(b->exec, c->stream));
    } else if ((rc = enqueue_all(b, c->stream))) {
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        return rc;
    }
    HIP_TRY(hipEventRecord(e1, c->stream));
    HIP_TRY(hipEventSynchronize(e1));
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, e0, e1);
    b->last_ms = ms;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    // scatter: round rows are sorted trials of their segments
    const psr_candidate* hc = (const psr_candidate*)b->pinned;
    const uint32_t* hcnt = (const uint32_t*)((const char*)b->pinned + b->total_trials * b->cap * sizeof(psr_candidate));
    uint32_t overflow = 0;
    for (size_t ri = 0; ri < b->rounds.size(); ri++) {
        const Round& r = b->rounds[ri];
        size_t row = r.trial_base;
        for (const Segment& sg : r.segs) {
            const ObsPlan& o = b->obs[sg.obs];
            for (size_t t = 0; t < sg.nt; t++, row++) {
                const uint32_t k = hcnt[row + ri];
                const size_t dst = o.out0 + o.order[sg.t0 + t];
                counts[dst] = k;
                std::memcpy(out + dst * b->cap, hc + row * b->cap, k * sizeof(psr_candidate));
            }
        }
        overflow += hcnt[r.trial_base + ri + (row - r.trial_base)];
    }
    if (overflow)
        return fail(PSR_E_OVERFLOW, std::to_string(overflow) + " trial(s) produced more peaks than the per-trial buffer holds");
    return PSR_OK;
}

int psr_batch_info(const psr_batch* b, char* json, size_t cap, size_t* len) {
    clear_error();
    if (!b || !len) return fail(PSR_E_INVALID, "null pointer");
    std::string s = "{\"observations\": " + std::to_string(b->obs.size()) +
                    ", \"trials\": " + std::to_string(b->total_trials) + ", \"classes\": " +
                    std::to_string(b->classes.size()) + ", \"rounds\": " + std::to_string(b->rounds.size()) +
                    ", \"graph\": " + (b->exec ? "true" : "false") + ", \"last_ms\": " +
                    std::to_string(b->last_ms) + ", \"dedisperse_span\": " + std::to_string(b->span) +
                    ", \"dedisperse_wg_per_cu\": " + std::to_string(b->wg_per_cu) + ", \"fft\": [";
    for (size_t i = 0; i < b->classes.size(); i++) {
        const Class& cl = b->classes[i];
        s += (i ? ", " : "") + std::string("{\"nfft\": ") + std::to_string(cl.nfft) + ", \"m1\": " +
             std::to_string(cl.g.m1) + ", \"m2\": " + std::to_string(cl.g.m2) + ", \"fft_trials\": " +
             std::to_string(cl.fft_batch) + "}";
    }
    s += "], \"round_shapes\": [";
    for (size_t i = 0; i < b->rounds.size(); i++) {
        const Round& r = b->rounds[i];
        s += (i ? ", " : "") + std::string("{\"nfft\": ") + std::to_string(b->classes[r.cls].nfft) +
             ", \"segments\": " + std::to_string(r.segs.size()) + ", \"trials\": " + std::to_string(r.ntrials) +
             ", \"items\": " + std::to_string(r.total_items) + "}";
    }
    s += "]";
    if (!b->tuning.empty()) s += ", \"autotune\": " + b->tuning;
    s += "}";
    *len = s.size();
    if (json && cap) {
        const size_t k = std::min(cap - 1, s.size());
        std::memcpy(json, s.data(), k);
        json[k] = 0;
    }
    return PSR_OK;
}

void psr_batch_destroy(psr_batch* b) {
    if (!b) return;
    psr_ctx* c = b->ctx;
    (void)hipSetDevice(c->device);
    (void)hipStreamSynchronize(c->stream);
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    DevBuf* all[] = {&b->tables, &b->delays, &b->ddtab, &b->series, &b->partial, &b->mean32, &b->corr16,
                     &b->work, &b->spec_t, &b->raw_nat, &b->scales, &b->sorted, &b->floors, &b->fft_partial,
                     &b->peaks, &b->peak_counts, &b->cands, &b->counts, &b->counters, &b->costabs};
    for (DevBuf* d : all) release(c, *d);
    for (Class& cl : b->classes) release(c, cl.tw);
    if (b->pinned) (void)hipHostFree(b->pinned);
    delete b;
}

}  // extern "C"
