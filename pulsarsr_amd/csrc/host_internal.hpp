// Host-only helpers shared by the C ABI layer and the host numerics.
#pragma once
#include <string>

#include "../../include/psr_gpu.h"

namespace psr {

constexpr double DISPERSION_SECONDS_MHZ = 4148.808;  // crates/dedisperse/src/lib.rs:3

// FFT geometry limits: a single LDS-resident pass up to FFT_SINGLE_MAX complex
// points, otherwise a four-step split m = m1 * m2 with m2 = FFT_PASS1_LEN
// (grown if m1 would exceed FFT_PASS2_MAX).
constexpr unsigned FFT_SINGLE_MAX = 4096;
constexpr unsigned FFT_PASS1_LEN = 512;
constexpr unsigned FFT_PASS2_MAX = 4096;
constexpr unsigned FFT_MAX_M = 1u << 23;

int fail(int code, const std::string& msg);
void clear_error();
const char* last_error();

namespace dedisperse {
int delays(const double* freqs, size_t nch, double tsamp, double dm, uint64_t* out);
int dm_step(const double* freqs, size_t nch, double tsamp, double* out);
int plan(double min, double max, double step, double* out, size_t cap, size_t* count);
}  // namespace dedisperse
namespace fbank {
int parse_header(const uint8_t* bytes, size_t len, psr_header* out);
}
namespace search {
int sift(const psr_candidate* in, size_t n, double resolution_hz, size_t limit, int reject_dm_zero,
         psr_candidate* out, size_t* count);
}

}  // namespace psr
