// Plan of the exact subband-tree dedispersion (see dedisp_plan.cpp).
#pragma once
#include <stdint.h>

#include <vector>

namespace psr {

constexpr uint32_t TREE_SUBBAND = 8;  // channels per subband (3 tree levels)
constexpr uint32_t TREE_TRIALS = 32;  // DM trials per workgroup block

// Program of one (DM block, subband), int32 words:
//   nleaf, nlevels,
//   nleaf x {channel (-1 = zero row), window start (samples from tile start, multiple of 4),
//            LDS word offset, words}
//   nlevels x {nrec, nrec x {dst, srcA, srcB, words}}   dst[i] = srcA[i] + srcB[i] (u16 pairs)
//   TREE_TRIALS x {LDS word offset of the trial's root partial at its base delay}
// An LDS word i of an array starting at sample w holds the u16 values of
// samples w+i (low half) and w+i+64 (high half).
struct TreePlan {
    uint32_t S = 0, D = 0;
    uint32_t nblocks = 0, nsub = 0;
    uint32_t lds_words = 0;
    uint32_t phase_subbands = 0;
    std::vector<int32_t> prog;
    std::vector<uint32_t> offsets;  // nblocks * nsub + 1
};

int build_tree_plan(const std::vector<double>& freqs, double tsamp, const double* dms,
                    uint32_t ntrials, uint32_t maxval, TreePlan* plan);

}  // namespace psr
