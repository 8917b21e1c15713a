// Host planner of the exact subband-tree dedispersion (k_dedisp_tree).
//
// For a block of D consecutive DM trials and a subband of S = 8 channels, the
// channel delays of each trial split into a base (the smallest delay in a tree
// node) and a relative pattern.  Trials whose relative pattern inside a node is
// identical share that node's partial sum
//     P_node,pattern[u] = sum_{c in node} x_c[u + pattern_c],
// which is built from the two child nodes' partials with one add per sample.
// Every trial's output is then the sum over subbands of its root partial read at
// the trial's base delay.  All sums are of integers, so the result equals the
// reference's channel-ordered f32 sum bit for bit (crates/dedisperse/src/lib.rs:80-88)
// while most trials reuse partials instead of re-adding every channel.
//
// The planner emits, per (DM block, subband), a small program the kernel
// interprets: leaf loads (global rows -> LDS pair words), one list of partial-sum
// records per tree level, and each trial's LDS offset of its root partial.  The
// plan depends only on the channel frequencies, tsamp and the DM list, not on
// the data, so it is built once per search and reused for every time tile.
#include <algorithm>
#include <cmath>
#include <map>
#include <thread>
#include <vector>

#include "dedisp_plan.hpp"
#include "host_internal.hpp"

namespace psr {

namespace {

struct NodePat {
    std::vector<int32_t> rel;  // relative delays of the node's channels
    int child[2] = {-1, -1};   // child pattern index
    int off[2] = {0, 0};       // child offset (its base minus this node's base)
    int64_t w0 = 0, w1 = 0;    // window, in samples relative to the tile start: word i covers w0+i
    bool used = false;
    int lds = 0;               // LDS word offset
};

struct Node {
    std::vector<NodePat> pats;
    std::vector<int> trial_pat;     // pattern of each trial
    std::vector<int64_t> base;      // base delay of each trial
};

constexpr int LANE_WORDS = 448;  // words read per trial: lane (0..63) + 128 * p (p < 4)

}  // namespace

// Build the program of one (block, subband).  dl[m][j]: delays of trial m,
// channel j of the subband (S entries; absent channels repeat the last delay).
static int build_program(const std::vector<std::vector<int64_t>>& dl, const std::vector<int>& chans,
                         std::vector<int32_t>& out, uint32_t* lds_words) {
    const int D = (int)dl.size();
    const int S = (int)chans.size();
    int levels = 0;
    while ((1 << levels) < S) levels++;
    // nodes[L][g]
    std::vector<std::vector<Node>> nodes(levels + 1);
    nodes[0].resize(S);
    for (int j = 0; j < S; j++) {
        Node& nd = nodes[0][j];
        nd.pats.resize(1);
        nd.pats[0].rel = {0};
        nd.trial_pat.assign(D, 0);
        nd.base.resize(D);
        for (int m = 0; m < D; m++) nd.base[m] = dl[m][j];
    }
    for (int L = 1; L <= levels; L++) {
        const int width = 1 << L, half = width / 2;
        nodes[L].resize(S / width);
        for (int g = 0; g < S / width; g++) {
            Node& nd = nodes[L][g];
            nd.trial_pat.resize(D);
            nd.base.resize(D);
            std::map<std::vector<int32_t>, int> index;
            for (int m = 0; m < D; m++) {
                int64_t b = dl[m][g * width];
                for (int j = 1; j < width; j++) b = std::min(b, dl[m][g * width + j]);
                std::vector<int32_t> rel(width);
                for (int j = 0; j < width; j++) rel[j] = (int32_t)(dl[m][g * width + j] - b);
                auto it = index.find(rel);
                int p;
                if (it == index.end()) {
                    p = (int)nd.pats.size();
                    index.emplace(rel, p);
                    NodePat np;
                    np.rel = rel;
                    for (int side = 0; side < 2; side++) {
                        const Node& ch = nodes[L - 1][2 * g + side];
                        // the child's pattern for this trial, and its base relative to ours
                        np.child[side] = ch.trial_pat[m];
                        np.off[side] = (int)(ch.base[m] - b);
                    }
                    nd.pats.push_back(np);
                } else {
                    p = it->second;
                }
                nd.trial_pat[m] = p;
                nd.base[m] = b;
            }
            (void)half;
        }
    }
    // windows, top-down: the root pattern p is read at words [base_m - w0, +LANE_WORDS)
    Node& root = nodes[levels][0];
    for (int m = 0; m < D; m++) {
        NodePat& np = root.pats[root.trial_pat[m]];
        const int64_t lo = root.base[m], hi = root.base[m] + LANE_WORDS;
        if (!np.used) {
            np.used = true;
            np.w0 = lo;
            np.w1 = hi;
        } else {
            np.w0 = std::min(np.w0, lo);
            np.w1 = std::max(np.w1, hi);
        }
    }
    for (int L = levels; L >= 1; L--)
        for (size_t g = 0; g < nodes[L].size(); g++)
            for (NodePat& np : nodes[L][g].pats) {
                if (!np.used) continue;
                for (int side = 0; side < 2; side++) {
                    NodePat& c = nodes[L - 1][2 * g + side].pats[np.child[side]];
                    const int64_t lo = np.w0 + np.off[side], hi = np.w1 + np.off[side];
                    if (!c.used) {
                        c.used = true;
                        c.w0 = lo;
                        c.w1 = hi;
                    } else {
                        c.w0 = std::min(c.w0, lo);
                        c.w1 = std::max(c.w1, hi);
                    }
                }
            }
    // leaves start on a multiple of 4 samples (aligned dword loads) and cover whole quads
    for (int j = 0; j < S; j++) {
        NodePat& np = nodes[0][j].pats[0];
        if (!np.used) continue;
        np.w0 = np.w0 & ~(int64_t)3;
        np.w1 = np.w0 + ((np.w1 - np.w0 + 3) & ~(int64_t)3);
    }
    // LDS: even levels in region A, odd levels in region B (a level's inputs are dead
    // once it has been computed)
    int64_t region[2] = {0, 0};
    std::vector<int64_t> level_words(levels + 1, 0);
    for (int L = 0; L <= levels; L++)
        for (Node& nd : nodes[L])
            for (NodePat& np : nd.pats)
                if (np.used) level_words[L] += (np.w1 - np.w0 + 3) & ~(int64_t)3;
    for (int L = 0; L <= levels; L++) region[L & 1] = std::max(region[L & 1], level_words[L]);
    int64_t cursor[2] = {0, region[0]};
    for (int L = 0; L <= levels; L++) {
        int64_t c = cursor[L & 1];
        for (Node& nd : nodes[L])
            for (NodePat& np : nd.pats)
                if (np.used) {
                    np.lds = (int)c;
                    c += (np.w1 - np.w0 + 3) & ~(int64_t)3;
                }
    }
    const int64_t total = region[0] + region[1];
    if (total > (1 << 30)) return fail(PSR_E_INVALID, "dedispersion plan too large");
    *lds_words = std::max(*lds_words, (uint32_t)total);
    // emit
    out.clear();
    int nleaf = 0;
    for (int j = 0; j < S; j++) nleaf += nodes[0][j].pats[0].used ? 1 : 0;
    out.push_back(nleaf);
    out.push_back(levels);
    for (int j = 0; j < S; j++) {
        const NodePat& np = nodes[0][j].pats[0];
        if (!np.used) continue;
        out.push_back(chans[j]);
        out.push_back((int32_t)np.w0);
        out.push_back(np.lds);
        out.push_back((int32_t)(np.w1 - np.w0));
    }
    for (int L = 1; L <= levels; L++) {
        size_t count_at = out.size();
        out.push_back(0);
        int nrec = 0;
        for (size_t g = 0; g < nodes[L].size(); g++)
            for (const NodePat& np : nodes[L][g].pats) {
                if (!np.used) continue;
                const NodePat& a = nodes[L - 1][2 * g].pats[np.child[0]];
                const NodePat& b = nodes[L - 1][2 * g + 1].pats[np.child[1]];
                out.push_back(np.lds);
                out.push_back((int32_t)(a.lds + (np.w0 + np.off[0] - a.w0)));
                out.push_back((int32_t)(b.lds + (np.w0 + np.off[1] - b.w0)));
                out.push_back((int32_t)(np.w1 - np.w0));
                nrec++;
            }
        out[count_at] = nrec;
    }
    for (int m = 0; m < D; m++) {
        const NodePat& np = root.pats[root.trial_pat[m]];
        out.push_back((int32_t)(np.lds + (root.base[m] - np.w0)));
    }
    return PSR_OK;
}

int build_tree_plan(const std::vector<double>& freqs, double tsamp, const double* dms,
                    uint32_t ntrials, uint32_t maxval, TreePlan* plan) {
    const uint32_t nch = (uint32_t)freqs.size();
    const uint32_t S = TREE_SUBBAND, D = TREE_TRIALS;
    plan->nsub = (nch + S - 1) / S;
    plan->nblocks = (ntrials + D - 1) / D;
    plan->lds_words = 0;
    // delays of every trial (host FP64, identical to dedisperse::delays)
    std::vector<std::vector<int64_t>> delays(ntrials, std::vector<int64_t>(nch));
    {
        std::vector<uint64_t> tmp(nch);
        for (uint32_t d = 0; d < ntrials; d++) {
            int rc = dedisperse::delays(freqs.data(), nch, tsamp, dms[d], tmp.data());
            if (rc) return rc;
            for (uint32_t c = 0; c < nch; c++) delays[d][c] = (int64_t)tmp[c];
        }
    }
    const size_t nprog = (size_t)plan->nblocks * plan->nsub;
    std::vector<std::vector<int32_t>> progs(nprog);
    std::vector<uint32_t> lds(nprog, 0);
    std::vector<int> rcs(nprog, 0);
    unsigned nth = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    auto work = [&](unsigned w) {
        std::vector<std::vector<int64_t>> dl(D, std::vector<int64_t>(S));
        std::vector<int> chans(S);
        for (size_t i = w; i < nprog; i += nth) {
            const uint32_t b = (uint32_t)(i / plan->nsub), s = (uint32_t)(i % plan->nsub);
            for (uint32_t j = 0; j < S; j++) {
                const uint32_t c = s * S + j;
                chans[j] = c < nch ? (int)c : -1;  // absent channel: zero row
            }
            for (uint32_t m = 0; m < D; m++) {
                const uint32_t d = std::min(b * D + m, ntrials - 1);
                for (uint32_t j = 0; j < S; j++) {
                    const uint32_t c = std::min(s * S + j, nch - 1);
                    dl[m][j] = delays[d][c];
                }
            }
            rcs[i] = build_program(dl, chans, progs[i], &lds[i]);
        }
    };
    std::vector<std::thread> pool;
    for (unsigned w = 1; w < nth; w++) pool.emplace_back(work, w);
    work(0);
    for (auto& t : pool) t.join();
    for (size_t i = 0; i < nprog; i++)
        if (rcs[i]) return rcs[i];
    plan->offsets.assign(nprog + 1, 0);
    size_t total = 0;
    for (size_t i = 0; i < nprog; i++) {
        plan->offsets[i] = (uint32_t)total;
        total += progs[i].size();
        plan->lds_words = std::max(plan->lds_words, lds[i]);
    }
    plan->offsets[nprog] = (uint32_t)total;
    plan->prog.resize(total);
    for (size_t i = 0; i < nprog; i++)
        std::copy(progs[i].begin(), progs[i].end(), plan->prog.begin() + plan->offsets[i]);
    // root partials hold at most S * maxval per sample; u16 accumulators widen
    // before the summed subbands could exceed 65535
    const uint32_t root_max = S * std::max(1u, maxval);
    plan->phase_subbands = std::max(1u, 65535u / root_max);
    plan->S = S;
    plan->D = D;
    return PSR_OK;
}

}  // namespace psr
