// Batched real FFT and power spectrum for gfx950 (crates/search/src/lib.rs:35-44):
//   size = 2*next_pow2(n); input = x - (mean as f32), zero-padded; X = rfft(input);
//   power = re^2 + im^2.
// A length-N real transform is computed as a length-M = N/2 complex transform
// of z[m] = x[2m] + i x[2m+1] followed by the even/odd split.  M <= 4096 runs
// in one LDS-resident pass per trial.  Larger M uses a four-step split
// M = m1*m2 (m = m1idx + m1*m2idx):
//   pass 1: length-m2 FFTs over the strided columns, times W_M^(m1idx*k2),
//           stored transposed as work[k2][m1idx];
//   pass 2: length-m1 FFTs over rows k2 and m2-k2 together, then the r2c
//           split (the partner of k = k2 + m2*k1 is M-k, held in the same
//           workgroup) and the power, written as spec_t[k2*m1 + k1].
// The transposed spectrum is put back in natural order by the whitening kernel.
//
// The common shapes (m2 = 512 in pass 1, m1 = 4096 in pass 2) run
// register-first: each thread loads the 16 points of one radix-16 butterfly
// straight from memory, so the first stage needs no LDS, and the remaining
// stages are regrouped by one (pass 1) or two (pass 2) LDS exchanges plus a
// lane swizzle, with one barrier per exchange.  Other sizes use Stockham
// passes of radix 8 or 16 (2 or 4 for the remainder): each thread loads R
// points, applies the pass twiddles, runs a register DFT_R and stores the
// outputs in autosorted order.  Butterflies are FP32; every twiddle comes from
// FP64-generated tables or exact constants.  The upper half of the
// zero-padded input is never read from memory.
#include <hip/hip_fp16.h>
#include <hip/hip_runtime.h>

#include "common.hpp"

// The transforms may fuse multiply-adds (the library is otherwise built without
// contraction, for the order-exact f32 sums elsewhere): FFT parity is normwise,
// and a fused multiply-add only drops a rounding.
#pragma clang fp contract(fast)

namespace psr {

__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
    return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
// multiply by -i
__device__ __forceinline__ float2 mul_mi(float2 a) { return make_float2(a.y, -a.x); }

// ------------------------------------------------------------ DFT codelets
__device__ __forceinline__ void dft2(float2& a, float2& b) {
    const float2 t = a;
    a = cadd(t, b);
    b = csub(t, b);
}
__device__ __forceinline__ void dft4(float2& a0, float2& a1, float2& a2, float2& a3) {
    const float2 s0 = cadd(a0, a2), d0 = csub(a0, a2);
    const float2 s1 = cadd(a1, a3), d1 = csub(a1, a3);
    a0 = cadd(s0, s1);
    a2 = csub(s0, s1);
    a1 = make_float2(d0.x + d1.y, d0.y - d1.x);  // d0 - i d1
    a3 = make_float2(d0.x - d1.y, d0.y + d1.x);  // d0 + i d1
}

constexpr float RS2 = 0.70710678118654752440f;  // sqrt(1/2)
constexpr float C16 = 0.92387953251128675613f;  // cos(pi/8)
constexpr float S16 = 0.38268343236508977173f;  // sin(pi/8)

template <int R>
__device__ __forceinline__ void dft(float2* x);

template <>
__device__ __forceinline__ void dft<2>(float2* x) {
    dft2(x[0], x[1]);
}
template <>
__device__ __forceinline__ void dft<4>(float2* x) {
    dft4(x[0], x[1], x[2], x[3]);
}
// n = a + 2m, k = b + 4c:  X[b+4c] = sum_a W2^(ac) W8^(ab) DFT4_m(x[a+2m])[b]
template <>
__device__ __forceinline__ void dft<8>(float2* x) {
    float2 e0 = x[0], e1 = x[2], e2 = x[4], e3 = x[6];
    float2 o0 = x[1], o1 = x[3], o2 = x[5], o3 = x[7];
    dft4(e0, e1, e2, e3);
    dft4(o0, o1, o2, o3);
    o1 = make_float2(RS2 * (o1.x + o1.y), RS2 * (o1.y - o1.x));   // * W8^1
    o2 = mul_mi(o2);                                               // * W8^2
    o3 = make_float2(RS2 * (o3.y - o3.x), -RS2 * (o3.x + o3.y));  // * W8^3
    x[0] = cadd(e0, o0);
    x[4] = csub(e0, o0);
    x[1] = cadd(e1, o1);
    x[5] = csub(e1, o1);
    x[2] = cadd(e2, o2);
    x[6] = csub(e2, o2);
    x[3] = cadd(e3, o3);
    x[7] = csub(e3, o3);
}
// n = a + 4m, k = b + 4c:  X[b+4c] = DFT4_a(W16^(ab) DFT4_m(x[a+4m])[b])[c]
// (dft16_finish: the twiddles W16^(ab) and the DFT4 over a)
__device__ __forceinline__ void dft16_finish(float2 (&y)[4][4], float2* x) {
    // W16^j = (cos 2pi j/16, -sin 2pi j/16)
    y[1][1] = cmul(y[1][1], make_float2(C16, -S16));
    y[1][2] = cmul(y[1][2], make_float2(RS2, -RS2));
    y[1][3] = cmul(y[1][3], make_float2(S16, -C16));
    y[2][1] = cmul(y[2][1], make_float2(RS2, -RS2));
    y[2][2] = mul_mi(y[2][2]);
    y[2][3] = cmul(y[2][3], make_float2(-RS2, -RS2));
    y[3][1] = cmul(y[3][1], make_float2(S16, -C16));
    y[3][2] = cmul(y[3][2], make_float2(-RS2, -RS2));
    y[3][3] = cmul(y[3][3], make_float2(-C16, S16));
#pragma unroll
    for (int b = 0; b < 4; b++) {
        float2 a0 = y[0][b], a1 = y[1][b], a2 = y[2][b], a3 = y[3][b];
        dft4(a0, a1, a2, a3);
        x[b] = a0;
        x[b + 4] = a1;
        x[b + 8] = a2;
        x[b + 12] = a3;
    }
}
template <>
__device__ __forceinline__ void dft<16>(float2* x) {
    float2 y[4][4];
#pragma unroll
    for (int a = 0; a < 4; a++) {
        y[a][0] = x[a];
        y[a][1] = x[a + 4];
        y[a][2] = x[a + 8];
        y[a][3] = x[a + 12];
        dft4(y[a][0], y[a][1], y[a][2], y[a][3]);
    }
    dft16_finish(y, x);
}
// the same DFT when x[8..15] are zero (zero padding): each first-stage DFT4 of
// (p, q, 0, 0) is (p + q, p - i q, p - q, p + i q), the values the full network
// gives, with half the additions
__device__ __forceinline__ void dft16_lower(float2* x) {
    float2 y[4][4];
#pragma unroll
    for (int a = 0; a < 4; a++) {
        const float2 p = x[a], q = x[a + 4];
        y[a][0] = cadd(p, q);
        y[a][1] = make_float2(p.x + q.y, p.y - q.x);
        y[a][2] = csub(p, q);
        y[a][3] = make_float2(p.x - q.y, p.y + q.x);
    }
    dft16_finish(y, x);
}

// ------------------------------------------------------------- twiddles
// W_L^m from a quarter-wave table C[j] = cos(2 pi j / L), j = 0..L/4.
template <int L>
__device__ __forceinline__ float2 twiddle(const float* C, int m) {
    constexpr int Q = L / 4;
    const int q = (m / Q) & 3, r = m & (Q - 1);
    const float c0 = C[r], s0 = C[Q - r];  // cos, sin of 2 pi r / L
    float c, s;
    if (q == 0) {
        c = c0;
        s = s0;
    } else if (q == 1) {
        c = -s0;
        s = c0;
    } else if (q == 2) {
        c = -c0;
        s = -s0;
    } else {
        c = s0;
        s = -c0;
    }
    return make_float2(c, -s);
}

// ------------------------------------------------------------ LDS FFTs
// Element e of vector v lives at LDS index e*V + v (padded by one slot per 16
// when PAD, which breaks the power-of-two strides of the V <= 2 layouts).
template <int V, bool PAD>
__device__ __forceinline__ int lidx(int v, int e) {
    const int a = e * V + v;
    return PAD ? a + (a >> 4) : a;
}

constexpr int ilog2c(int v) { return v <= 1 ? 0 : 1 + ilog2c(v / 2); }
constexpr int npasses(int logl) { return logl <= 4 ? 1 : (logl + 3) / 4; }
constexpr int pass_bits(int logl, int p) {
    return logl / npasses(logl) + (p < logl % npasses(logl) ? 1 : 0);
}

// Pass twiddles: pass p (span NS_p, radix R_p) multiplies input r of butterfly
// k by W_(NS_p R_p)^(r k).  Each pass has two small LDS tables indexed by k,
// W^k and W^(4k), so neighbouring lanes read neighbouring entries; the other
// powers are products of these (at most three complex products deep).
constexpr int ns_before(int logl, int p) {
    return p == 0 ? 1 : ns_before(logl, p - 1) << pass_bits(logl, p - 1);
}
constexpr int tw_offset(int logl, int p) {  // entries of passes 1..p-1 (pass 0 has none)
    return p <= 1 ? 0 : tw_offset(logl, p - 1) + ns_before(logl, p - 1);
}
constexpr int tw_total(int L) { return tw_offset(ilog2c(L), npasses(ilog2c(L))); }

template <int L>
__device__ __forceinline__ void build_pass_tw(float2* T1, float2* T4, const float* cos_tables) {
    constexpr int LOGL = ilog2c(L);
    const float* C = cos_tables + costab_offset(L);
    constexpr int total = tw_total(L);
    for (int i = threadIdx.x; i < total; i += blockDim.x) {
        int p = 1;
        while (p + 1 < npasses(LOGL) && i >= tw_offset(LOGL, p + 1)) p++;
        const int k = i - tw_offset(LOGL, p);
        const int step = L / (ns_before(LOGL, p) << pass_bits(LOGL, p));
        T1[i] = twiddle<L>(C, (k * step) & (L - 1));
        T4[i] = twiddle<L>(C, (4 * k * step) & (L - 1));
    }
}

template <int R>
__device__ __forceinline__ void apply_tw(float2* x, float2 w1, float2 w4) {
    if constexpr (R == 2) {
        x[1] = cmul(x[1], w1);
    } else {
        const float2 w2 = cmul(w1, w1), w3 = cmul(w2, w1);
        x[1] = cmul(x[1], w1);
        x[2] = cmul(x[2], w2);
        x[3] = cmul(x[3], w3);
        if constexpr (R >= 8) {
            x[4] = cmul(x[4], w4);
            x[5] = cmul(x[5], cmul(w4, w1));
            x[6] = cmul(x[6], cmul(w4, w2));
            x[7] = cmul(x[7], cmul(w4, w3));
        }
        if constexpr (R == 16) {
            const float2 w8 = cmul(w4, w4);
            x[8] = cmul(x[8], w8);
            x[9] = cmul(x[9], cmul(w8, w1));
            x[10] = cmul(x[10], cmul(w8, w2));
            x[11] = cmul(x[11], cmul(w8, w3));
            const float2 w12 = cmul(w8, w4);
            x[12] = cmul(x[12], w12);
            x[13] = cmul(x[13], cmul(w12, w1));
            x[14] = cmul(x[14], cmul(w12, w2));
            x[15] = cmul(x[15], cmul(w12, w3));
        }
    }
}

// x[k] *= tw[k * S], k = 1 .. 15, with the powers read from an LDS table whose
// rows are the powers k (the threads of a wave that read different columns hit
// consecutive words: no bank conflicts)
template <int S>
__device__ __forceinline__ void apply_tw_col(float2* x, const float2* tw) {
#pragma unroll
    for (int k = 1; k < 16; k++) x[k] = cmul(x[k], tw[k * S]);
}

constexpr int FFT_THREADS = 512;

template <int L, int V, bool PAD, int R, int NS>
__device__ __forceinline__ void stockham_pass(float2* buf, const float2* T1, const float2* T4) {
    constexpr int NT = FFT_THREADS;
    constexpr int NBF = V * L / R;
    constexpr int PER = (NBF + NT - 1) / NT;
    float2 x[PER][R];
    const int tid = threadIdx.x;
#pragma unroll
    for (int i = 0; i < PER; i++) {
        const int j = tid + i * NT;
        if (NBF % NT == 0 || j < NBF) {
            const int v = j % V, jj = j / V;
#pragma unroll
            for (int r = 0; r < R; r++) x[i][r] = buf[lidx<V, PAD>(v, jj + r * (L / R))];
        }
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < PER; i++) {
        const int j = tid + i * NT;
        if (NBF % NT == 0 || j < NBF) {
            const int v = j % V, jj = j / V;
            const int k = jj & (NS - 1);
            if constexpr (NS > 1) apply_tw<R>(x[i], T1[k], T4[k]);
            dft<R>(x[i]);
            const int base = (jj / NS) * NS * R + k;
#pragma unroll
            for (int r = 0; r < R; r++) buf[lidx<V, PAD>(v, base + r * NS)] = x[i][r];
        }
    }
    __syncthreads();
}

template <int L, int V, bool PAD, int P, int NS>
__device__ __forceinline__ void fft_passes(float2* buf, const float2* T1, const float2* T4) {
    constexpr int LOGL = ilog2c(L);
    if constexpr (P < npasses(LOGL) && L > 1) {
        constexpr int R = 1 << pass_bits(LOGL, P);
        constexpr int off = tw_offset(LOGL, P);
        stockham_pass<L, V, PAD, R, NS>(buf, T1 + off, T4 + off);
        fft_passes<L, V, PAD, P + 1, NS * R>(buf, T1, T4);
    }
}

// Series sample i of a trial as FFT input: x[i] - mean32 in f32 for FP32
// storage, h[i] + corr (already scaled) for FP16 storage, zero beyond n.  The
// second element of a pair may sit one past the row's end (inside the pitch or
// the buffer's slack) and is then discarded.
template <bool F16>
__device__ __forceinline__ float2 load_pair(const void* series, size_t row, uint32_t i, uint32_t n,
                                            float mean, float corr) {
    // unconditional load from a valid address (so unrolled loads can be
    // batched), then select; the series is read once, so the loads are marked
    // streaming (nontemporal: 0.5 ms per headline FP16 step in same-box runs)
    const uint32_t is = i < n ? i : 0u;
    float2 f;
    if (F16) {
        const uint32_t u = __builtin_nontemporal_load((const uint32_t*)((const __half*)series + row + is));
        __half2 h;
        __builtin_memcpy(&h, &u, 4);
        f = __half22float2(h);
        f.x += corr;
        f.y += corr;
    } else {
        const uint64_t u = __builtin_nontemporal_load((const uint64_t*)((const float*)series + row + is));
        __builtin_memcpy(&f, &u, 8);
        f.x -= mean;
        f.y -= mean;
    }
    return make_float2(i < n ? f.x : 0.f, i + 1 < n ? f.y : 0.f);
}

// Stored spectrum value: the power (FP32) or the amplitude sqrt(power / n)
// (FP16, which keeps every value in half range); returns what the whitening
// floor sums.  The FP16 amplitude takes the hardware square root (within 1 ulp
// of f32, far below the half-precision rounding that follows) instead of the
// correctly rounded sequence, a seventh of the instructions.
template <bool F16>
__device__ __forceinline__ double store_power(void* spec, size_t idx, float p, float inv_n) {
    if (F16) {
        const float q = p * inv_n;
        ((__half*)spec)[idx] = __float2half_rn(__builtin_amdgcn_sqrtf(q));
        return (double)q;
    }
    ((float*)spec)[idx] = p;
    return (double)p;
}

__device__ double block_sum(double v, double* red) {  // red: blockDim/64 entries
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) red[w] = v;
    __syncthreads();
    double s = 0.0;
    if (threadIdx.x == 0)
        for (int i = 0; i < (int)(blockDim.x >> 6); i++) s += red[i];
    return s;
}

// Power of X[k] from Z[k] and Z[M-k] (zp) with W = W_N^k:
//   X = (Z + conj Zp)/2 + W (Z - conj Zp)/(2i)
__device__ __forceinline__ float r2c_power(float2 zk, float2 zp, float2 w) {
    // e, od and x are twice E[k], O[k] and X[k]; scaling by powers of two is
    // exact, so 0.25 |2 X|^2 is the same float as |X|^2 from the halved terms
    const float2 e = make_float2(zk.x + zp.x, zk.y - zp.y);
    const float2 od = make_float2(zk.y + zp.y, zp.x - zk.x);
    const float2 x = cadd(e, cmul(w, od));
    const float xx = x.x * x.x, yy = x.y * x.y;
    return 0.25f * (xx + yy);
}

constexpr int cdiv(int a, int b) { return (a + b - 1) / b; }

// Global loads of a kernel's input are issued together (fully unrolled into
// registers) before the twiddle tables are built and the data is written to
// LDS, so their latencies overlap instead of adding up.

// ---------------------------------------------------------- single pass
template <int L, bool F16>
__global__ __launch_bounds__(512, 4) void k_fft_single(const void* __restrict__ series,
                                                    const float* __restrict__ mean32,
                                                    const float* __restrict__ corr16, uint32_t n_all,
                                                    const uint32_t* __restrict__ trial_n,
                                                    uint32_t series_pitch,
                                                    void* __restrict__ spec, uint64_t spec_pitch,
                                                    double* __restrict__ partial,
                                                    const float* __restrict__ cos_tables) {
    constexpr int M = L;
    constexpr int NBUF = M + (M >> 4) + 1;
    constexpr int NT = tw_total(M) + 1;
    constexpr int PER = cdiv(M, FFT_THREADS);
    extern __shared__ float2 smem[];
    float2* buf = smem;
    float2* T1 = smem + NBUF;                // pass twiddles W^k
    float2* T4 = T1 + NT;                    // pass twiddles W^(4k)
    const float* CN = cos_tables + costab_offset(2 * M);  // length-2M table (r2c twiddles)
    __shared__ double red[FFT_THREADS / 64];
    const uint32_t trial = blockIdx.x;
    const uint32_t n = trial_n ? trial_n[trial] : n_all;  // per-trial length (many-observation batches)
    const size_t row = (size_t)trial * series_pitch;
    const float mean = F16 ? 0.f : mean32[trial];
    const float corr = F16 ? corr16[trial] : 0.f;
    float2 v[PER];
#pragma unroll
    for (int i = 0; i < PER; i++) {
        const int mm = threadIdx.x + i * FFT_THREADS;
        v[i] = mm < M ? load_pair<F16>(series, row, 2u * mm, n, mean, corr) : make_float2(0.f, 0.f);
    }
    if constexpr (M >= 4) build_pass_tw<M>(T1, T4, cos_tables);
#pragma unroll
    for (int i = 0; i < PER; i++) {
        const int mm = threadIdx.x + i * FFT_THREADS;
        if (mm < M) buf[lidx<1, true>(0, mm)] = v[i];
    }
    __syncthreads();
    fft_passes<M, 1, true, 0, 1>(buf, T1, T4);
    const float inv_n = 1.0f / (float)n;
    double acc = 0.0;
    const size_t srow = (size_t)trial * spec_pitch;
#pragma unroll
    for (int i = 0; i < cdiv(M + 1, FFT_THREADS); i++) {
        const int k = threadIdx.x + i * FFT_THREADS;
        if (k > M) break;
        const float2 zk = buf[lidx<1, true>(0, k & (M - 1))];
        const float2 zp = buf[lidx<1, true>(0, (M - k) & (M - 1))];
        float p;
        if (k == M) {  // Nyquist: X[M] = E[0] - O[0]
            const float xm = zk.x - zk.y;
            p = xm * xm;
        } else {
            p = r2c_power(zk, zp, twiddle<2 * M>(CN, k));
        }
        acc += store_power<F16>(spec, srow + k, p, inv_n);
    }
    const double s = block_sum(acc, red);
    if (threadIdx.x == 0) partial[trial] = s;
}

// ------------------------------------------------- pass 1, register-first
// Length-512 column FFTs, 16 columns per workgroup.  With row r = j + 32 m
// and output k2 = k1 + 16 q:
//   X[k1 + 16 q] = sum_j W_32^(j q) W_512^(j k1) sum_m x[j + 32 m] W_16^(m k1).
// Thread (j, col) loads rows j + 32 m (m < 16) of column col and runs the
// inner DFT_16 on its registers.  One LDS exchange gives thread (k1, col, j0)
// the values j = j0 + 2 j1 (j1 < 16) for a DFT_16 over j1; the final radix 2
// over j0 pairs lanes l and l ^ 16 (j0 is lane bit 4) through a swizzle.
constexpr float C32A = 0.98078528040323044913f;  // cos(pi/16)
constexpr float S32A = 0.19509032201612826785f;  // sin(pi/16)
constexpr float C32B = 0.83146961230254523708f;  // cos(3pi/16)
constexpr float S32B = 0.55557023301960222474f;  // sin(3pi/16)
__device__ __forceinline__ float2 w32(int q) {  // W_32^q, q < 16
    switch (q) {
        case 0: return make_float2(1.f, 0.f);
        case 1: return make_float2(C32A, -S32A);
        case 2: return make_float2(C16, -S16);
        case 3: return make_float2(C32B, -S32B);
        case 4: return make_float2(RS2, -RS2);
        case 5: return make_float2(S32B, -C32B);
        case 6: return make_float2(S16, -C16);
        case 7: return make_float2(S32A, -C32A);
        case 8: return make_float2(0.f, -1.f);
        case 9: return make_float2(-S32A, -C32A);
        case 10: return make_float2(-S16, -C16);
        case 11: return make_float2(-S32B, -C32B);
        case 12: return make_float2(-RS2, -RS2);
        case 13: return make_float2(-C32B, -S32B);
        case 14: return make_float2(-C16, -S16);
        default: return make_float2(-C32A, -S32A);
    }
}
__device__ __forceinline__ float swz16(float x) {  // value of lane l ^ 16 (within 32-lane halves)
    return __int_as_float(__builtin_amdgcn_ds_swizzle(__float_as_int(x), 0x401f));
}
// Powers of the 16 bins v[q] of this lane, whose r2c partners are bins 15 - q
// of lane l ^ 16 (W_N^k of bin q = base * W_32^q).  With 2E = Z + conj Zp and
// 2O = (Z - conj Zp) / i, X_k = E + W O and the partner's X = conj(E - W O), so
// each lane forms the pairs (its q, partner's 15 - q) for q < 8 and returns
// the partner's 8 powers through the swizzle: half the butterflies and
// twiddles of forming all 16 from both sides.
__device__ __forceinline__ void r2c_powers_paired(const float2 (&v)[16], float2 base, float (&pw)[16]) {
    float pp[8];
#pragma unroll
    for (int q = 0; q < 8; q++) {
        const float2 zk = v[q], s = v[15 - q];
        const float2 zp = make_float2(swz16(s.x), swz16(s.y));
        const float2 w = q == 0 ? base : cmul(base, w32(q));
        const float2 e = make_float2(zk.x + zp.x, zk.y - zp.y);
        const float2 od = make_float2(zk.y + zp.y, zp.x - zk.x);
        const float2 t = cmul(w, od);
        const float2 x1 = cadd(e, t), x2 = csub(e, t);
        pw[q] = 0.25f * (x1.x * x1.x + x1.y * x1.y);
        pp[q] = 0.25f * (x2.x * x2.x + x2.y * x2.y);
    }
#pragma unroll
    for (int q = 0; q < 8; q++) pw[15 - q] = swz16(pp[q]);
}
template <bool F16>
__global__ __launch_bounds__(512, 4) void k_fft_pass1r(const void* __restrict__ series,
                                                    const float* __restrict__ mean32,
                                                    const float* __restrict__ corr16, uint32_t n_all,
                                                    const uint32_t* __restrict__ trial_n,
                                                    uint32_t series_pitch, uint32_t M, uint32_t m1,
                                                    float2* __restrict__ work,
                                                    const float2* __restrict__ tw4,
                                                    const float* __restrict__ cos_tables) {
    constexpr int C = 16, JP = 272;  // exchange rows of 256 + 16 (bank offset between rows)
    __shared__ float2 buf[32 * JP];
    __shared__ float2 P[256];  // W_M^(16 (c0 + col) q1) at q1 * 16 + col
    __shared__ float2 TJ[16 * 32];  // W_512^(j k) at k * 32 + j
    const uint2 wb = xcd_block();
    const uint32_t trial = wb.y, c0 = wb.x * C;
    const uint32_t n = trial_n ? trial_n[trial] : n_all;
    const size_t row = (size_t)trial * series_pitch;
    const float mean = F16 ? 0.f : mean32[trial];
    const float corr = F16 ? corr16[trial] : 0.f;
    const int t = threadIdx.x, col = t & 15, j = t >> 4;
    float2 v[16];
    // rows j + 32 m (j < 32) start at sample 2 (c0 + m1 (j + 32 m)).  When the
    // first sample of row 256 is at or past n, row groups m = 8..15 hold only
    // padding zeros for the whole workgroup (the zero-padded half of the series
    // usually does): they are not loaded, and the DFT takes half the adds.
    const bool upper_empty = 2ull * (c0 + (uint64_t)m1 * 256u) >= (uint64_t)n;  // uniform
#pragma unroll
    for (int m = 0; m < 8; m++)
        v[m] = load_pair<F16>(series, row, 2u * (c0 + col + m1 * (uint32_t)(j + 32 * m)), n, mean, corr);
    if (upper_empty) {
#pragma unroll
        for (int m = 8; m < 16; m++) v[m] = make_float2(0.f, 0.f);
    } else {
#pragma unroll
        for (int m = 8; m < 16; m++)
            v[m] = load_pair<F16>(series, row, 2u * (c0 + col + m1 * (uint32_t)(j + 32 * m)), n, mean, corr);
    }
    const float* C512 = cos_tables + costab_offset(512);
    TJ[t] = twiddle<512>(C512, (t >> 5) * (t & 31));  // one table entry per thread
    const float2* tlo = tw4;
    const float2* thi = tw4 + 2048;
    const int j0 = (t >> 4) & 1, k1 = t >> 5;
    const uint32_t c = c0 + col;
    const uint32_t ab = (uint32_t)(((uint64_t)c * (uint32_t)(k1 + 256 * j0)) & (M - 1));
    const float2 base = cmul(thi[ab >> 11], tlo[ab & 2047]);
    if (t < 256) {
        const uint32_t ap = (uint32_t)(((uint64_t)(c0 + (t & 15)) * (16u * (uint32_t)(t >> 4))) & (M - 1));
        P[t] = cmul(thi[ap >> 11], tlo[ap & 2047]);
    }
    if (upper_empty)
        dft16_lower(v);
    else
        dft<16>(v);
    __syncthreads();  // TJ
    apply_tw_col<32>(v, TJ + j);
#pragma unroll
    for (int k = 0; k < 16; k++) buf[j * JP + k * 16 + col] = v[k];
    __syncthreads();
#pragma unroll
    for (int j1 = 0; j1 < 16; j1++) v[j1] = buf[(j0 + 2 * j1) * JP + k1 * 16 + col];
    dft<16>(v);
    const float sgn = j0 ? -1.f : 1.f;
    float2* w = work + (size_t)trial * M;
#pragma unroll
    for (int q = 0; q < 16; q++) {
        const float2 u = q == 0 ? v[q] : cmul(v[q], j0 ? w32(q) : make_float2(1.f, 0.f));
        const float2 p = make_float2(swz16(u.x), swz16(u.y));
        const float2 o = make_float2(p.x + sgn * u.x, p.y + sgn * u.y);
        const uint32_t k2 = (uint32_t)(k1 + 16 * q + 256 * j0);
        w[(size_t)k2 * m1 + c] = cmul(o, cmul(base, P[q * 16 + col]));
    }
}

// ------------------------------------------- pass 1, general (m2 = 1024)
template <int M2, bool F16>
__global__ __launch_bounds__(512, 4) void k_fft_pass1(const void* __restrict__ series,
                                                   const float* __restrict__ mean32,
                                                   const float* __restrict__ corr16, uint32_t n_all,
                                                    const uint32_t* __restrict__ trial_n,
                                                   uint32_t series_pitch, uint32_t M, uint32_t m1,
                                                   float2* __restrict__ work,
                                                   const float2* __restrict__ tw4,
                                                   const float* __restrict__ cos_tables) {
    constexpr int C = 8192 / M2;  // columns per workgroup
    constexpr int NT = tw_total(M2) + 1;
    constexpr int PER = C * M2 / FFT_THREADS;
    extern __shared__ float2 smem[];
    float2* buf = smem;  // [M2][C], element r of column col at r*C + col
    float2* T1 = smem + C * M2;
    float2* T4 = T1 + NT;
    const uint2 wb = xcd_block();  // column groups sharing cache lines on one XCD
    const uint32_t trial = wb.y;
    const uint32_t n = trial_n ? trial_n[trial] : n_all;
    const uint32_t c0 = wb.x * C;
    const size_t row = (size_t)trial * series_pitch;
    const float mean = F16 ? 0.f : mean32[trial];
    const float corr = F16 ? corr16[trial] : 0.f;
    // rows r with 2*(c0 + m1*r) >= n hold only padding zeros
    float2 v[PER];
#pragma unroll
    for (int i = 0; i < PER; i++) {
        const int idx = threadIdx.x + i * FFT_THREADS;
        const int r = idx / C, col = idx % C;
        v[i] = load_pair<F16>(series, row, 2u * (c0 + col + m1 * r), n, mean, corr);
    }
    build_pass_tw<M2>(T1, T4, cos_tables);
#pragma unroll
    for (int i = 0; i < PER; i++) buf[threadIdx.x + i * FFT_THREADS] = v[i];
    // the output twiddles, loaded while the transform runs: thread's column is
    // fixed (FFT_THREADS is a multiple of C); k2 steps by FFT_THREADS / C
    const float2* tlo = tw4;
    const float2* thi = tw4 + 2048;
    const int col = threadIdx.x % C, k20 = threadIdx.x / C;
    constexpr int KS = FFT_THREADS / C;
    float2 tl[PER], th[PER];
#pragma unroll
    for (int i = 0; i < PER; i++) {
        const uint32_t a = (uint32_t)(((uint64_t)(c0 + col) * (k20 + KS * i)) & (M - 1));
        tl[i] = tlo[a & 2047];
        th[i] = thi[a >> 11];
    }
    __syncthreads();
    fft_passes<M2, C, false, 0, 1>(buf, T1, T4);
    float2* w = work + (size_t)trial * M;
#pragma unroll
    for (int i = 0; i < PER; i++) {
        const int k2 = k20 + KS * i;
        w[(size_t)k2 * m1 + c0 + col] = cmul(buf[k2 * C + col], cmul(th[i], tl[i]));
    }
}

// --------------------------------------------------------------- pass 2
// Row pairs (r1, m2 - r1) per workgroup for row length M1 < 4096: enough that a
// workgroup holds 8192 points (every thread has whole butterflies in each
// Stockham pass), at most 256.
constexpr int pass2_pairs(int m1) { return m1 >= 4096 ? 1 : (4096 / m1 > 256 ? 256 : 4096 / m1); }
constexpr int pass2_elems(int m1) { return 2 * pass2_pairs(m1) * m1; }
constexpr int pass2_buf(int m1) { return pass2_elems(m1) + (pass2_elems(m1) >> 4) + 1; }

// P = pass2_pairs(M1) row pairs, pair p = rows r1 = P blockIdx.x + p and
// r2 = m2 - r1 (r1 <= m2 / 2), as 2P interleaved vectors through the LDS
// Stockham passes; then every bin's r2c power (the partner of bin k1 of row r1
// is bin M1 - 1 - k1 of row r2, of row 0 bin M1 - k1 of itself) and each
// pair's sum of the stored values for the whitening floor, reduced in a fixed
// order (pairwise over the pair's bins in LDS).
template <int M1, bool F16>
__global__ __launch_bounds__(512, 4) void k_fft_pass2(const float2* __restrict__ work, uint32_t n_all,
                                                   const uint32_t* __restrict__ trial_n,
                                                   uint32_t M, uint32_t m2,
                                                   void* __restrict__ spec, uint64_t spec_pitch,
                                                   double* __restrict__ partial,
                                                   const float2* __restrict__ twrow,
                                                   const float* __restrict__ cos_tables) {
    constexpr int P = pass2_pairs(M1), V = 2 * P, NE = pass2_elems(M1), NBUF = pass2_buf(M1);
    constexpr int NT = tw_total(M1) + 1;
    constexpr int PE = NE / FFT_THREADS, NW = FFT_THREADS / 64;
    static_assert(NE % FFT_THREADS == 0, "whole values per thread");
    static_assert(P * NW <= NBUF && PE * NW <= NBUF, "the wave sums reuse the transform buffer");
    extern __shared__ float2 smem[];
    float2* buf = smem;  // the 2P rows interleaved, padded
    float2* T1 = smem + NBUF;
    float2* T4 = T1 + NT;
    const float* C2 = cos_tables + costab_offset(2 * M1);  // W_N^(m2 k1) = W_(2 M1)^k1
    const uint32_t trial = blockIdx.y, npairs = m2 / 2 + 1, p0 = blockIdx.x * P;
    const uint32_t n = trial_n ? trial_n[trial] : n_all;
    const float2* w = work + (size_t)trial * M;
    // vector v holds row r1 (v even) or r2 (v odd) of pair v / 2
    auto row_of = [&](int v) {
        const uint32_t r1 = p0 + (uint32_t)(v >> 1);
        return (v & 1) ? (m2 - r1) & (m2 - 1) : r1;
    };
    float2 va[PE];
#pragma unroll
    for (int i = 0; i < PE; i++) {
        const int j = threadIdx.x + i * FFT_THREADS, v = j / M1, k1 = j % M1;
        va[i] = p0 + (uint32_t)(v >> 1) < npairs ? w[(size_t)row_of(v) * M1 + k1] : make_float2(0.f, 0.f);
    }
    if constexpr (M1 >= 4) build_pass_tw<M1>(T1, T4, cos_tables);
#pragma unroll
    for (int i = 0; i < PE; i++) {
        const int j = threadIdx.x + i * FFT_THREADS;
        buf[lidx<V, true>(j / M1, j % M1)] = va[i];
    }
    __syncthreads();
    fft_passes<M1, V, true, 0, 1>(buf, T1, T4);
    const float inv_n = 1.0f / (float)n;
    const size_t srow = (size_t)trial * spec_pitch;
    // the stored values (f32, exactly what the floor sums) and the Nyquist one
    float val[PE];
    float nyq = 0.f;
#pragma unroll 4
    for (int i = 0; i < PE; i++) {
        const int j = threadIdx.x + i * FFT_THREADS;
        val[i] = 0.f;
        const int v = j / M1, k1 = j % M1, pr = v >> 1, sel = v & 1;
        const uint32_t r1 = p0 + (uint32_t)pr, r2 = (m2 - r1) & (m2 - 1);
        const int nsel = r1 == r2 ? 1 : 2;
        if (r1 >= npairs || sel >= nsel) continue;
        const uint32_t rowk = sel ? r2 : r1;
        const float2 zk = buf[lidx<V, true>(v, k1)];
        float2 zp;
        if (rowk == 0)
            zp = buf[lidx<V, true>(2 * pr, (M1 - k1) & (M1 - 1))];
        else
            zp = buf[lidx<V, true>(2 * pr + (nsel == 1 ? 0 : 1 - sel), M1 - 1 - k1)];
        const float2 wk = cmul(twrow[rowk], twiddle<2 * M1>(C2, k1));
        const float pw = r2c_power(zk, zp, wk);
        val[i] = (float)store_power<F16>(spec, srow + (size_t)rowk * M1 + k1, pw, inv_n);
        if (rowk == 0 && k1 == 0) {  // Nyquist bin k = M: X[M] = E[0] - O[0]
            const float xm = zk.x - zk.y;
            nyq = (float)store_power<F16>(spec, srow + M, xm * xm, inv_n);
        }
    }
    // each pair's 2 M1 stored values (zeros where a bin is not stored) summed in
    // a fixed order: within a thread, across the lanes, then across the waves
    constexpr int G2 = 2 * M1;  // values of one pair
    const bool nq = p0 == 0 && threadIdx.x == 0;  // row 0's bin 0: takes the Nyquist bin
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    double* red = (double*)smem;  // [value group][wave]
    double* out = partial + (size_t)trial * npairs + p0;
    if constexpr (G2 > 64) __syncthreads();  // the transform buffer becomes the wave sums
    if constexpr (G2 >= FFT_THREADS) {
        // a thread's values of pair q are val[q G .. q G + G - 1]
        constexpr int G = G2 / FFT_THREADS;
#pragma unroll
        for (int q = 0; q < P; q++) {
            double t = 0.0;
#pragma unroll
            for (int g = 0; g < G; g++) t += (double)val[q * G + g];
            if (q == 0 && nq) t += (double)nyq;
            for (int o = 32; o > 0; o >>= 1) t += __shfl_xor(t, o);
            if (lane == 0) red[q * NW + wv] = t;
        }
        __syncthreads();
        for (int q = threadIdx.x; q < P; q += FFT_THREADS) {
            double t = 0.0;
            for (int k = 0; k < NW; k++) t += red[q * NW + k];
            if (p0 + (uint32_t)q < npairs) out[q] = t;
        }
    } else {
        // val[i] of thread x belongs to pair i PS + x / G2: G2 consecutive lanes
        constexpr int PS = FFT_THREADS / G2, SEG = G2 < 64 ? G2 : 64;
#pragma unroll
        for (int i = 0; i < PE; i++) {
            double t = (double)val[i];
            if (i == 0 && nq) t += (double)nyq;
            for (int o = SEG / 2; o > 0; o >>= 1) t += __shfl_xor(t, o);
            if constexpr (G2 <= 64) {
                const int q = i * PS + (int)threadIdx.x / G2;
                if ((lane & (G2 - 1)) == 0 && p0 + (uint32_t)q < npairs) out[q] = t;
            } else if (lane == 0) {
                red[i * NW + wv] = t;
            }
        }
        if constexpr (G2 > 64) {
            constexpr int W = G2 / 64;  // waves per pair
            __syncthreads();
            for (int q = threadIdx.x; q < P; q += FFT_THREADS) {
                const int i = q / PS, w0 = (q % PS) * W;
                double t = 0.0;
                for (int k = 0; k < W; k++) t += red[i * NW + w0 + k];
                if (p0 + (uint32_t)q < npairs) out[q] = t;
            }
        }
    }
}

// ------------------------------------------------- pass 2, register-first
// Rows r1 and r2 = m2 - r1 (length 4096) as 16 x 16 x 16: thread (row, tt)
// loads elements tt + 256 m (m < 16) and runs the first radix-16 stage on
// them; two LDS exchanges regroup for the other two stages.  The last stage
// places the r2c partner of every output (row r2, column 4095 - k) in lane
// l ^ 16, which a swizzle reaches.  Workgroup 0 takes the two rows that are
// their own partners, 0 (partner 4096 - k) and m2 / 2 (4095 - k), through one
// more LDS exchange, so a trial is exactly m2 / 2 workgroups.
template <bool F16>
__global__ __launch_bounds__(512, 4) void k_fft_pass2r(const float2* __restrict__ work, uint32_t n_all,
                                                   const uint32_t* __restrict__ trial_n,
                                                    uint32_t M, uint32_t m2,
                                                    void* __restrict__ spec, uint64_t spec_pitch,
                                                    double* __restrict__ partial,
                                                    const float2* __restrict__ twrow,
                                                    const float* __restrict__ cos_tables) {
    constexpr int M1 = 4096, JP = 272, RS = 16 * JP;
    __shared__ float2 buf[2 * RS];
    __shared__ float2 TA[16 * 16];  // W_256^(a k) at k * 16 + a
    __shared__ double red[FFT_THREADS / 64];
    const uint32_t trial = blockIdx.y;
    const uint32_t n = trial_n ? trial_n[trial] : n_all;
    const uint32_t r1 = blockIdx.x, r2 = r1 == 0 ? m2 / 2 : m2 - r1;
    const float2* w = work + (size_t)trial * M;
    const int t = threadIdx.x;
    float2 v[16];
    {  // stage A: radix 16 over m, twiddle W_4096^(tt k1) from the FP64-made table
        const int row = t >> 8, tt = t & 255;
        const float2* src = w + (size_t)(row ? r2 : r1) * M1 + tt;
#pragma unroll
        for (int m = 0; m < 16; m++) v[m] = src[256 * m];
        const float2* T = (const float2*)(cos_tables + TW4096_OFFSET) + tt;
        float2 tw[15];
#pragma unroll
        for (int k = 0; k < 15; k++) tw[k] = T[256 * k];
        if (t < 256) TA[t] = twiddle<256>(cos_tables + costab_offset(256), (t >> 4) * (t & 15));
        dft<16>(v);
#pragma unroll
        for (int k = 1; k < 16; k++) v[k] = cmul(v[k], tw[k - 1]);
#pragma unroll
        for (int k = 0; k < 16; k++) buf[row * RS + k * JP + (tt >> 4) * 16 + (tt & 15)] = v[k];
    }
    __syncthreads();
    {  // stage B1: radix 16 over b (tt = a + 16 b), twiddle W_256^(a q1)
        const int row = t >> 8, a = t & 15, k1 = (t >> 4) & 15;
#pragma unroll
        for (int b = 0; b < 16; b++) v[b] = buf[row * RS + k1 * JP + b * 16 + a];
        dft<16>(v);
        apply_tw_col<16>(v, TA + a);  // written before the previous barrier
        __syncthreads();
#pragma unroll
        for (int q = 0; q < 16; q++) buf[row * RS + q * JP + k1 * 16 + (a ^ k1)] = v[q];
    }
    __syncthreads();
    // stage B2: radix 16 over a; lane l holds row (l >> 4) & 1 and, for row 1,
    // the mirrored (k1, q1) of lane l ^ 16's row-0 pair
    const int l = t & 63, row = (l >> 4) & 1, i16 = l & 15, qq = 2 * (t >> 6) + (l >> 5);
    const int k1 = row ? 15 - i16 : i16, q1 = row ? 15 - qq : qq;
#pragma unroll
    for (int a = 0; a < 16; a++) v[a] = buf[row * RS + q1 * JP + k1 * 16 + (a ^ k1)];
    dft<16>(v);  // v[q2] = X_row[k1 + 16 q1 + 256 q2]
    const float inv_n = 1.0f / (float)n;
    const size_t srow = (size_t)trial * spec_pitch;
    const uint32_t rowk = row ? r2 : r1;
    const float* C8192 = cos_tables + costab_offset(8192);
    const float2 base = cmul(twrow[rowk], twiddle<8192>(C8192, k1 + 16 * q1));
    double acc = 0.0;
    if (r1 == 0) {  // rows 0 (partner 4096 - k) and m2 / 2 (partner 4095 - k)
        __syncthreads();
#pragma unroll
        for (int q = 0; q < 16; q++) buf[row * M1 + k1 + 16 * q1 + 256 * q] = v[q];
        __syncthreads();
#pragma unroll
        for (int q = 0; q < 16; q++) {
            const int k = k1 + 16 * q1 + 256 * q;
            const float2 zk = v[q], zp = buf[row ? M1 + (M1 - 1 - k) : (M1 - k) & (M1 - 1)];
            const float p = r2c_power(zk, zp, cmul(base, w32(q)));
            acc += store_power<F16>(spec, srow + (size_t)rowk * M1 + k, p, inv_n);
            if (row == 0 && k == 0) {  // Nyquist bin k = M: X[M] = E[0] - O[0]
                const float xm = zk.x - zk.y;
                acc += store_power<F16>(spec, srow + M, xm * xm, inv_n);
            }
        }
    } else {
        // all 16 powers first (the swizzles need every lane), then one predicated
        // run of stores from one base address; the floor's partial sums the
        // thread's 16 values in f32 first (16 positive terms: relative error
        // below 1e-6 of a floor that is 1e-6 of the mean power)
        float pw[16];
        r2c_powers_paired(v, base, pw);
        {
            const size_t o = srow + (size_t)rowk * M1 + k1 + 16 * q1;
            float part = 0.f;
            if (F16) {
                __half* dst = (__half*)spec + o;
#pragma unroll
                for (int q = 0; q < 16; q++) {
                    const float x = pw[q] * inv_n;
                    dst[256 * q] = __float2half_rn(__builtin_amdgcn_sqrtf(x));
                    part += x;
                }
            } else {
                float* dst = (float*)spec + o;
#pragma unroll
                for (int q = 0; q < 16; q++) {
                    dst[256 * q] = pw[q];
                    part += pw[q];
                }
            }
            acc = (double)part;
        }
    }
    double* out = partial + (size_t)trial * (m2 / 2 + 1);
    if (r1 == 0) {  // the two rows' sums apart
        const double s0 = block_sum(row ? 0.0 : acc, red);
        __syncthreads();  // red is reused
        const double s1 = block_sum(row ? acc : 0.0, red);
        if (threadIdx.x == 0) {
            out[0] = s0;
            out[m2 / 2] = s1;
        }
    } else {
        const double sum = block_sum(acc, red);
        if (threadIdx.x == 0) out[r1] = sum;
    }
}

// ------------------------------------- pass 2, register-first, m1 = 256 R3
// The three stages of k_fft_pass2r for m1 = 256 R3 (R3 = 1, 2, 4, 8), on
// NR = 32 / R3 row vectors (P = NR / 2 pairs r1 = P blockIdx.x + p, r2 =
// m2 - r1; pair 0 is rows 0 and m2 / 2), 16 points per thread.  Row element n = tt + S m (S = 16 R3):
// stage A is the radix 16 over m (output k1) with W_m1^(tt k1); with
// tt = a + 16 b, stage B1 the radix R3 over b (output q1) with W_S^(a q1);
// stage B2 the radix 16 over a (output q2), bin k = k1 + 16 q1 + S q2.  In
// stage B2 lane l holds vector 2 p + ((l >> 4) & 1) and, for r2, the mirrored
// (15 - k1, R3 - 1 - q1) of lane l ^ 16's r1 bins, so the r2c partner
// m1 - 1 - k of every bin is a swizzle away; pair 0, whose rows are their own
// partners (row 0: m1 - k, row m2 / 2: m1 - 1 - k), takes one more LDS
// exchange, so a trial is exactly m2 / (2 P) workgroups.  Every 16 lanes touch
// 16 distinct LDS bank pairs in each exchange.
template <int R3, bool F16>
__global__ __launch_bounds__(512, 4) void k_fft_pass2x(const float2* __restrict__ work, uint32_t n_all,
                                                    const uint32_t* __restrict__ trial_n,
                                                    uint32_t M, uint32_t m2,
                                                    void* __restrict__ spec, uint64_t spec_pitch,
                                                    double* __restrict__ partial,
                                                    const float2* __restrict__ twrow,
                                                    const float* __restrict__ cos_tables) {
    constexpr int M1 = 256 * R3, S = 16 * R3, NR = 32 / R3, P = NR / 2;
    static_assert(R3 == 1 || R3 == 2 || R3 == 4 || R3 == 8, "m1 = 256, 512, 1024 or 2048");
    __shared__ float2 buf[NR * M1];
    __shared__ float2 TA[S];  // W_S^(a q1) at q1 * 16 + a
    __shared__ double red[2 * (FFT_THREADS / 64)];
    const uint32_t trial = blockIdx.y, npairs = m2 / 2 + 1, p0 = blockIdx.x * P;
    const uint32_t n = trial_n ? trial_n[trial] : n_all;
    const float2* w = work + (size_t)trial * M;
    const int t = threadIdx.x;
    auto partner_row = [&](uint32_t r1) { return r1 == 0 ? m2 / 2 : m2 - r1; };
    float2 v[16];
    {  // stage A: radix 16 over m, twiddle W_m1^(tt k1) = W_4096^(tt (16 / R3) k1)
        const int vec = t / S, tt = t % S;
        const uint32_t r1 = p0 + (uint32_t)(vec >> 1);
        const uint32_t row = (vec & 1) ? partner_row(r1) : r1;
        const float2* src = w + (size_t)row * M1 + tt;
#pragma unroll
        for (int m = 0; m < 16; m++) v[m] = src[S * m];
        const float2* T = (const float2*)(cos_tables + TW4096_OFFSET) + tt * (16 / R3);
        float2 tw[15];
#pragma unroll
        for (int k = 0; k < 15; k++) tw[k] = T[256 * k];
        if constexpr (R3 > 1)
            if (t < S) TA[t] = twiddle<S>(cos_tables + costab_offset(S), (t >> 4) * (t & 15));
        dft<16>(v);
#pragma unroll
        for (int k = 1; k < 16; k++) v[k] = cmul(v[k], tw[k - 1]);
        if constexpr (R3 == 1) {  // straight into the stage-B2 layout (a = tt, q1 = 0)
#pragma unroll
            for (int k = 0; k < 16; k++) buf[vec * M1 + k * 16 + (tt ^ k)] = v[k];
        } else {
#pragma unroll
            for (int k = 0; k < 16; k++) buf[vec * M1 + k * S + tt] = v[k];
        }
    }
    __syncthreads();
    if constexpr (R3 > 1) {  // stage B1: 16 / R3 radix-R3 butterflies (vector, k1, a) per thread
        constexpr int NB = 16 / R3;
        const int a = t & 15, k1 = (t >> 4) & 15;
#pragma unroll
        for (int i = 0; i < NB; i++) {
            const int vec = (t >> 8) + 2 * i;
#pragma unroll
            for (int b = 0; b < R3; b++) v[i * R3 + b] = buf[vec * M1 + k1 * S + 16 * b + a];
        }
#pragma unroll
        for (int i = 0; i < NB; i++) {
            dft<R3>(v + i * R3);
#pragma unroll
            for (int q = 1; q < R3; q++) v[i * R3 + q] = cmul(v[i * R3 + q], TA[q * 16 + a]);
        }
        __syncthreads();
#pragma unroll
        for (int i = 0; i < NB; i++) {
            const int vec = (t >> 8) + 2 * i;
#pragma unroll
            for (int q = 0; q < R3; q++) buf[vec * M1 + q * 256 + k1 * 16 + (a ^ k1)] = v[i * R3 + q];
        }
        __syncthreads();
    }
    // stage B2: radix 16 over a
    const int l = t & 63, sel = (l >> 4) & 1, i16 = l & 15, c = 2 * (t >> 6) + (l >> 5);
    const int pr = c / R3, qq = c % R3, vec = 2 * pr + sel;
    const int k1 = sel ? 15 - i16 : i16, q1 = sel ? R3 - 1 - qq : qq;
#pragma unroll
    for (int a = 0; a < 16; a++) v[a] = buf[vec * M1 + q1 * 256 + k1 * 16 + (a ^ k1)];
    dft<16>(v);  // v[q2] = X_row[k1 + 16 q1 + S q2]
    const float inv_n = 1.0f / (float)n;
    const size_t srow = (size_t)trial * spec_pitch;
    const uint32_t r1 = p0 + (uint32_t)pr, rowk = sel ? partner_row(r1) : r1;
    const int kb = k1 + 16 * q1;
    const float2 base = cmul(twrow[rowk], twiddle<2 * M1>(cos_tables + costab_offset(2 * M1), kb));
    double acc = 0.0;
    // all 16 powers first (the swizzles need every lane); pair 0's are redone below
    float pw[16];
    r2c_powers_paired(v, base, pw);
    const bool self = p0 == 0 && pr == 0;  // pair 0: rows 0 and m2 / 2, their own partners
    if (p0 == 0) {
        __syncthreads();
        if (self)
#pragma unroll
            for (int q = 0; q < 16; q++) buf[sel * M1 + kb + S * q] = v[q];
        __syncthreads();
        if (self) {
#pragma unroll
            for (int q = 0; q < 16; q++) {
                const int k = kb + S * q;
                const float2 zk = v[q], zp = buf[sel ? M1 + (M1 - 1 - k) : (M1 - k) & (M1 - 1)];
                const float p = r2c_power(zk, zp, cmul(base, w32(q)));
                acc += store_power<F16>(spec, srow + (size_t)rowk * M1 + k, p, inv_n);
                if (sel == 0 && k == 0) {  // Nyquist bin k = M: X[M] = E[0] - O[0]
                    const float xm = zk.x - zk.y;
                    acc += store_power<F16>(spec, srow + M, xm * xm, inv_n);
                }
            }
        }
    }
    // one predicated run of stores per thread; the floor's partial sums the
    // thread's 16 values in f32 first (16 positive terms)
    if (!self) {
        const size_t o = srow + (size_t)rowk * M1 + kb;
        float part = 0.f;
        if (F16) {
            __half* dst = (__half*)spec + o;
#pragma unroll
            for (int q = 0; q < 16; q++) {
                const float x = pw[q] * inv_n;
                dst[S * q] = __float2half_rn(__builtin_amdgcn_sqrtf(x));
                part += x;
            }
        } else {
            float* dst = (float*)spec + o;
#pragma unroll
            for (int q = 0; q < 16; q++) {
                dst[S * q] = pw[q];
                part += pw[q];
            }
        }
        acc = (double)part;
    }
    // pair p's values are those of threads 32 R3 p .. 32 R3 (p + 1) - 1; the
    // sums of its two rows (lanes with sel 0 and 1) are kept apart for pair 0
    constexpr int SEG = 32 * R3;
    double* out = partial + (size_t)trial * npairs;
    double a = sel ? 0.0 : acc, b = sel ? acc : 0.0;
    if constexpr (SEG <= 64) {
        for (int o = SEG / 2; o > 0; o >>= 1) {
            a += __shfl_xor(a, o);
            b += __shfl_xor(b, o);
        }
        if ((t & (SEG - 1)) == 0) {
            if (self) {
                out[0] = a;
                out[m2 / 2] = b;
            } else {
                out[r1] = a + b;
            }
        }
    } else {
        for (int o = 32; o > 0; o >>= 1) {
            a += __shfl_xor(a, o);
            b += __shfl_xor(b, o);
        }
        if (l == 0) {
            red[t >> 6] = a;
            red[FFT_THREADS / 64 + (t >> 6)] = b;
        }
        __syncthreads();
        if (t < P) {
            double sa = 0.0, sb = 0.0;
#pragma unroll
            for (int k = 0; k < SEG / 64; k++) {
                sa += red[t * (SEG / 64) + k];
                sb += red[FFT_THREADS / 64 + t * (SEG / 64) + k];
            }
            if (p0 == 0 && t == 0) {
                out[0] = sa;
                out[m2 / 2] = sb;
            } else {
                out[p0 + t] = sa + sb;
            }
        }
    }
}

// ------------------------------------------------------------- launchers
template <typename K>
static void allow_lds(K kernel, size_t bytes) {
    if (bytes > 64 * 1024)
        (void)hipFuncSetAttribute(reinterpret_cast<const void*>(kernel),
                                  hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
}

template <int L, bool F16>
static void launch_single(hipStream_t s, const SpecGeom& g, const void* series,
                          const float* mean32, const float* corr16, const uint32_t* trial_n,
                          uint32_t ntrials, void* spec, double* partial, const float* cos_tables) {
    const size_t shm = (size_t)(L + (L >> 4) + 1 + 2 * (tw_total(L) + 1)) * sizeof(float2);
    allow_lds(k_fft_single<L, F16>, shm);
    hipLaunchKernelGGL((k_fft_single<L, F16>), dim3(ntrials), dim3(FFT_THREADS), shm, s, series,
                       mean32, corr16, g.n, trial_n, g.series_pitch, spec, g.spec_pitch, partial, cos_tables);
}

template <int M1, bool F16>
static void launch_pass2(hipStream_t s, const SpecGeom& g, const uint32_t* trial_n, uint32_t ntrials,
                         const float2* work, void* spec, double* partial, const float2* twrow,
                         const float* cos_tables) {
    if constexpr (M1 == 4096) {
        hipLaunchKernelGGL((k_fft_pass2r<F16>), dim3(g.m2 / 2, ntrials), dim3(FFT_THREADS), 0,
                           s, work, g.n, trial_n, g.m, g.m2, spec, g.spec_pitch, partial, twrow, cos_tables);
    } else if constexpr (M1 >= 256) {
        constexpr uint32_t P = 8192 / M1 / 2;  // divides m2 / 2 (>= 256)
        hipLaunchKernelGGL((k_fft_pass2x<M1 / 256, F16>), dim3(g.m2 / 2 / P, ntrials),
                           dim3(FFT_THREADS), 0, s, work, g.n, trial_n, g.m, g.m2, spec, g.spec_pitch, partial,
                           twrow, cos_tables);
    } else {
        const size_t shm = (size_t)(pass2_buf(M1) + 2 * (tw_total(M1) + 1)) * sizeof(float2);
        allow_lds(k_fft_pass2<M1, F16>, shm);
        const uint32_t npairs = g.m2 / 2 + 1, P = (uint32_t)pass2_pairs(M1);
        hipLaunchKernelGGL((k_fft_pass2<M1, F16>), dim3((npairs + P - 1) / P, ntrials), dim3(FFT_THREADS), shm,
                           s, work, g.n, trial_n, g.m, g.m2, spec, g.spec_pitch, partial, twrow, cos_tables);
    }
}

template <int M2, bool F16>
static void launch_pass1(hipStream_t s, const SpecGeom& g, const void* series, const float* mean32,
                         const float* corr16, const uint32_t* trial_n, uint32_t ntrials, float2* work,
                         const float2* tw4, const float* cos_tables) {
    if constexpr (M2 == 512) {  // m1 = M / 512 >= 16 columns
        hipLaunchKernelGGL((k_fft_pass1r<F16>), dim3(g.m1 / 16, ntrials), dim3(FFT_THREADS), 0, s,
                           series, mean32, corr16, g.n, trial_n, g.series_pitch, g.m, g.m1, work, tw4,
                           cos_tables);
        return;
    }
    constexpr int C = 8192 / M2;
    const size_t shm = (size_t)(C * M2 + 2 * (tw_total(M2) + 1)) * sizeof(float2);
    allow_lds(k_fft_pass1<M2, F16>, shm);
    hipLaunchKernelGGL((k_fft_pass1<M2, F16>), dim3(g.m1 / C, ntrials), dim3(FFT_THREADS), shm, s,
                       series, mean32, corr16, g.n, trial_n, g.series_pitch, g.m, g.m1, work, tw4,
                       cos_tables);
}

template <bool F16>
static void launch_fft_t(hipStream_t s, const SpecGeom& g, const void* series, const float* mean32,
                         const float* corr16, const uint32_t* trial_n, uint32_t ntrials, float2* work,
                         void* spec, double* partial, const float2* tw, const float* cos_tables,
                         uint32_t batch_trials) {
    if (g.m1 == 1) {
        switch (g.m) {
#define SINGLE(L)                                                                             \
    case L:                                                                                   \
        launch_single<L, F16>(s, g, series, mean32, corr16, trial_n, ntrials, spec, partial, cos_tables); \
        break;
            SINGLE(2) SINGLE(4) SINGLE(8) SINGLE(16) SINGLE(32) SINGLE(64) SINGLE(128)
            SINGLE(256) SINGLE(512) SINGLE(1024) SINGLE(2048) SINGLE(4096)
#undef SINGLE
        }
        return;
    }
    // tw layout: [W_M^j, j<2048][W_M^(2048 j), j<M/2048][W_N^k2, k2<m2]
    const float2* tw4 = tw;
    const float2* twrow = tw + 2048 + g.m / 2048;
    // Trials go through both passes in batches of FFT_BATCH_BYTES of
    // intermediate, which bounds the work buffer independently of the chunk.
    const uint32_t tb = batch_trials ? (batch_trials < ntrials ? batch_trials : ntrials) : fft_work_trials(g, ntrials);
    const size_t es = F16 ? 2 : 4;
    const uint32_t npart = g.m2 / 2 + 1;
    for (uint32_t b = 0; b < ntrials; b += tb) {
        const uint32_t nb = ntrials - b < tb ? ntrials - b : tb;
        const void* ser = (const char*)series + (size_t)b * g.series_pitch * es;
        void* sp = (char*)spec + (size_t)b * g.spec_pitch * es;
        double* part = partial + (size_t)b * npart;
        const uint32_t* tn = trial_n ? trial_n + b : nullptr;
        switch (g.m2) {  // fft_shape_supported() admits exactly these
            case 512:
                launch_pass1<512, F16>(s, g, ser, mean32 + b, corr16 + b, tn, nb, work, tw4, cos_tables);
                break;
            case 1024:
                launch_pass1<1024, F16>(s, g, ser, mean32 + b, corr16 + b, tn, nb, work, tw4, cos_tables);
                break;
            case 2048:
                launch_pass1<2048, F16>(s, g, ser, mean32 + b, corr16 + b, tn, nb, work, tw4, cos_tables);
                break;
        }
        switch (g.m1) {
#define PASS2(L) \
    case L: launch_pass2<L, F16>(s, g, tn, nb, work, sp, part, twrow, cos_tables); break;
            PASS2(8) PASS2(16) PASS2(32) PASS2(64) PASS2(128) PASS2(256) PASS2(512) PASS2(1024)
            PASS2(2048) PASS2(4096)
#undef PASS2
        }
    }
}

bool fft_shape_supported(const SpecGeom& g) {
    if (g.m1 == 1) return g.m >= 2 && g.m <= 4096 && (g.m & (g.m - 1)) == 0;
    const bool m2_ok = g.m2 == 512 || g.m2 == 1024 || g.m2 == 2048;
    const bool m1_ok = g.m1 >= 8 && g.m1 <= 4096 && (g.m1 & (g.m1 - 1)) == 0;
    return m2_ok && m1_ok && (uint64_t)g.m1 * g.m2 == g.m;
}

uint32_t fft_work_trials(const SpecGeom& g, uint32_t ntrials) {
    if (g.m1 == 1) return ntrials;
    const uint64_t per = (uint64_t)g.m * sizeof(float2);
    uint64_t tb = FFT_BATCH_BYTES / per;
    if (tb < 1) tb = 1;
    return (uint32_t)(tb < ntrials ? tb : ntrials);
}

void launch_fft(hipStream_t s, const SpecGeom& g, const void* series, int fp16,
                const float* mean32, const float* corr16, uint32_t ntrials, float2* work,
                void* spec, double* partial, const float2* tw, const float* cos_tables,
                const uint32_t* trial_n, uint32_t batch_trials) {
    if (fp16)
        launch_fft_t<true>(s, g, series, mean32, corr16, trial_n, ntrials, work, spec, partial, tw, cos_tables,
                           batch_trials);
    else
        launch_fft_t<false>(s, g, series, mean32, corr16, trial_n, ntrials, work, spec, partial, tw, cos_tables,
                            batch_trials);
}

}  // namespace psr
