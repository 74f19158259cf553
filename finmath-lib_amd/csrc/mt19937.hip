// MT19937 stream of commons-math3 MersenneTwister(long) on the GPU, split into
// sub-streams of L = 2^k words whose start states come from GF(2) jump-ahead,
// and the Brownian increments of BrownianMotionFromMersenneRandomNumbers
// (montecarlo/BrownianMotionFromMersenneRandomNumbers.java:141-191):
//   u   = ((w0 >> 6) << 26 | (w1 >> 6)) * 2^-52   (nextDouble, two words)
//   dW  = AS241(u) * sqrt(dt_t)                    (functions/NormalDistribution.java:67-162)
// written in the reference's generation order (path, then time, then factor).

#include <hip/hip_runtime.h>
#include <cstring>
#include <algorithm>
#include "fm_internal.hpp"
#include "hip_check.hpp"

namespace fm {

namespace {

constexpr int N = 624;
constexpr int M = 397;
constexpr uint32_t MATRIX_A = 0x9908b0dfu;
constexpr int kDeg = 19937;
constexpr int kPolyWords = 312;                 // 64-bit words of a reduced jump polynomial
constexpr int kSeqLen = kDeg + N;               // y_0 .. y_{19936+623}

__device__ __forceinline__ uint32_t temper(uint32_t y) {
    y ^= y >> 11;
    y ^= (y << 7) & 0x9d2c5680u;
    y ^= (y << 15) & 0xefc60000u;
    y ^= y >> 18;
    return y;
}

__device__ __forceinline__ uint32_t twist_word(uint32_t a, uint32_t b, uint32_t c) {
    // new = c ^ ((upper(a) | lower(b)) >> 1) ^ (lsb ? MATRIX_A : 0)
    uint32_t y = (a & 0x80000000u) | (b & 0x7fffffffu);
    return c ^ (y >> 1) ^ ((y & 1u) ? MATRIX_A : 0u);
}

// ---------------------------------------------------------------------------
// Jump tree level: workgroup g computes the window of sub-stream dst = src + gap
// from the window of src = g * 2 * gap (in sub-stream units), using the
// polynomial c = x^(gap * L) mod phi. One workgroup per jump; the generated
// sequence y_0..y_20560 of the source window lives in LDS (82 KB).
// ---------------------------------------------------------------------------
constexpr int kJumpThreads = 256;

__global__ void __launch_bounds__(kJumpThreads)
jump_level_kernel(uint32_t* __restrict__ states, const uint64_t* __restrict__ poly,
                  int64_t num_sub, int64_t gap)
{
    __shared__ uint32_t y[kSeqLen];
    const int64_t src = (int64_t)blockIdx.x * 2 * gap;
    const int64_t dst = src + gap;
    if (dst >= num_sub) return;                     // uniform per workgroup
    const int t = threadIdx.x;
    const uint32_t* sw = states + (size_t)src * N;
    for (int w = t; w < N; w += kJumpThreads) y[w] = sw[w];
    __syncthreads();
    // y[n+624] = twist(y[n], y[n+1], y[n+397]); blocks of 227 are independent
    for (int base = 0; base < kDeg; base += (N - M)) {
        const int n = base + t;
        uint32_t v = 0;
        const bool act = (t < (N - M)) && (n < kDeg);
        if (act) v = twist_word(y[n], y[n + 1], y[n + M]);
        __syncthreads();
        if (act) y[n + N] = v;
        __syncthreads();
    }
    // window_J[w] = XOR_{i : c_i = 1} y[i + w]
    uint32_t acc0 = 0, acc1 = 0, acc2 = 0;
    const int w0 = t, w1 = t + kJumpThreads, w2 = t + 2 * kJumpThreads;
    for (int iw = 0; iw < kPolyWords; iw++) {
        uint64_t cw = poly[iw];                      // uniform
        while (cw) {
            const int b = __builtin_ctzll(cw);
            cw &= cw - 1;
            const int i = iw * 64 + b;
            acc0 ^= y[i + w0];
            acc1 ^= y[i + w1];
            if (w2 < N) acc2 ^= y[i + w2];
        }
    }
    uint32_t* dw = states + (size_t)dst * N;
    dw[w0] = acc0;
    dw[w1] = acc1;
    if (w2 < N) dw[w2] = acc2;
}

// ---------------------------------------------------------------------------
// Brownian generator: one workgroup per sub-stream. The 624-word state is
// ping-ponged between two LDS buffers; a twist runs in three dependency phases
// (words [0,227) need only old words, [227,454) need new [0,227), [454,624) need
// new [227,397) and new word 0). 312 threads then each temper one word pair into a
// uniform; central-branch draws are transformed and stored at once, tail-branch draws
// go to an LDS queue that all threads then work through, so the tail code runs in
// about one wavefront per twist instead of in every wavefront.
// ---------------------------------------------------------------------------
constexpr int kGenThreads = 320;

// Coefficient x as an R. An FP64 VALU operand cannot be a literal on this target, so a
// double coefficient is built in an SGPR pair by two scalar moves at its point of use: the
// compiler would otherwise hoist the 48 coefficients of the three rational functions out of
// the generator's loops into 96 VGPRs, halving the resident wavefronts. The moves are
// volatile so they stay where they are; the value is the same double either way.
template <typename R, uint64_t B> struct Coef {
    __device__ __forceinline__ static R get() { return (R)__builtin_bit_cast(double, B); }
};
template <uint64_t B> struct Coef<double, B> {
    __device__ __forceinline__ static double get() {
        uint32_t lo, hi;
        asm volatile("s_mov_b32 %0, %2\n\ts_mov_b32 %1, %3"
                     : "=s"(lo), "=s"(hi)
                     : "i"((uint32_t)(B & 0xffffffffu)), "i"((uint32_t)(B >> 32)));
        return __builtin_bit_cast(double, ((uint64_t)hi << 32) | lo);
    }
};
#define AS241_C(R, x) (Coef<R, __builtin_bit_cast(uint64_t, (double)(x))>::get())

// Wichura AS241 (PPND16), as functions/NormalDistribution.java:67-162; the
// published coefficient table. Split into the central branch |q| <= 0.425 and the
// tail branch so the generator can run the tail densely (about 15 % of the draws).
// p <= 0 (or >= 1) in the tail branch gives 0.
template <typename R> __device__ __forceinline__ R as241_central(R q) {
    const R r = AS241_C(R, 0.180625) - q * q;
    R num = AS241_C(R, 2.5090809287301226727e+03);
    num = num * r + AS241_C(R, 3.3430575583588128105e+04);
    num = num * r + AS241_C(R, 6.7265770927008700853e+04);
    num = num * r + AS241_C(R, 4.5921953931549871457e+04);
    num = num * r + AS241_C(R, 1.3731693765509461125e+04);
    num = num * r + AS241_C(R, 1.9715909503065514427e+03);
    num = num * r + AS241_C(R, 1.3314166789178437745e+02);
    num = num * r + AS241_C(R, 3.3871328727963666080e+00);
    R den = AS241_C(R, 5.2264952788528545610e+03);
    den = den * r + AS241_C(R, 2.8729085735721942674e+04);
    den = den * r + AS241_C(R, 3.9307895800092710610e+04);
    den = den * r + AS241_C(R, 2.1213794301586595867e+04);
    den = den * r + AS241_C(R, 5.3941960214247511077e+03);
    den = den * r + AS241_C(R, 6.8718700749205790830e+02);
    den = den * r + AS241_C(R, 4.2313330701600911252e+01);
    den = den * r + (R)1.0;
    return q * num / den;
}

// Central branch in float for FP32 increments: the same rational function with fused
// multiply-adds and the hardware reciprocal (about 1e-6 relative; the FP64 path above
// keeps the reference's operation order).
__device__ __forceinline__ float as241_central_f32(float q) {
    const float r = 0.180625f - q * q;
    float num = 2.5090809287301226727e+03f;
    num = fmaf(num, r, 3.3430575583588128105e+04f);
    num = fmaf(num, r, 6.7265770927008700853e+04f);
    num = fmaf(num, r, 4.5921953931549871457e+04f);
    num = fmaf(num, r, 1.3731693765509461125e+04f);
    num = fmaf(num, r, 1.9715909503065514427e+03f);
    num = fmaf(num, r, 1.3314166789178437745e+02f);
    num = fmaf(num, r, 3.3871328727963666080e+00f);
    float den = 5.2264952788528545610e+03f;
    den = fmaf(den, r, 2.8729085735721942674e+04f);
    den = fmaf(den, r, 3.9307895800092710610e+04f);
    den = fmaf(den, r, 2.1213794301586595867e+04f);
    den = fmaf(den, r, 5.3941960214247511077e+03f);
    den = fmaf(den, r, 6.8718700749205790830e+02f);
    den = fmaf(den, r, 4.2313330701600911252e+01f);
    den = fmaf(den, r, 1.0f);
    return q * num * __builtin_amdgcn_rcpf(den);
}

// (bits * 2^-52) - (c - 1) for a 52-bit integer bits and c in {1, 1.5}: the double with
// exponent 0 and mantissa bits is 1 + bits * 2^-52, and subtracting c is exact (the result is a
// multiple of 2^-52 below 1 in magnitude), so c = 1 gives the uniform u and c = 1.5 gives
// u - 0.5, both as the conversion and subtraction in double give them.
__device__ __forceinline__ double unit_offset(uint64_t bits, double c) {
    return __builtin_bit_cast(double, 0x3ff0000000000000ull | bits) - c;
}

template <typename R> __device__ __forceinline__ R dev_sqrt_neg_log(double r);
template <> __device__ __forceinline__ double dev_sqrt_neg_log<double>(double r) { return sqrt(-log(r)); }
template <> __device__ __forceinline__ float dev_sqrt_neg_log<float>(double r) { return sqrtf(-logf((float)r)); }

// tail branch; p is the uniform, q = p - 0.5 with |q| > 0.425
template <typename R> __device__ __forceinline__ R as241_tail(double p, double q) {
    const double rp = (q < 0.0) ? p : 1.0 - p;
    if (rp <= 0.0) return (R)0;
    R r = dev_sqrt_neg_log<R>(rp);
    R num, den;
    if (r <= AS241_C(R, 5.0)) {
        r -= AS241_C(R, 1.6);
        num = AS241_C(R, 7.74545014278341407640e-04);
        num = num * r + AS241_C(R, 2.27238449892691845833e-02);
        num = num * r + AS241_C(R, 2.41780725177450611770e-01);
        num = num * r + AS241_C(R, 1.27045825245236838258e+00);
        num = num * r + AS241_C(R, 3.64784832476320460504e+00);
        num = num * r + AS241_C(R, 5.76949722146069140550e+00);
        num = num * r + AS241_C(R, 4.63033784615654529590e+00);
        num = num * r + AS241_C(R, 1.42343711074968357734e+00);
        den = AS241_C(R, 1.05075007164441684324e-09);
        den = den * r + AS241_C(R, 5.47593808499534494600e-04);
        den = den * r + AS241_C(R, 1.51986665636164571966e-02);
        den = den * r + AS241_C(R, 1.48103976427480074590e-01);
        den = den * r + AS241_C(R, 6.89767334985100004550e-01);
        den = den * r + AS241_C(R, 1.67638483018380384940e+00);
        den = den * r + AS241_C(R, 2.05319162663775882187e+00);
        den = den * r + (R)1.0;
    } else {
        r -= AS241_C(R, 5.0);
        num = AS241_C(R, 2.01033439929228813265e-07);
        num = num * r + AS241_C(R, 2.71155556874348757815e-05);
        num = num * r + AS241_C(R, 1.24266094738807843860e-03);
        num = num * r + AS241_C(R, 2.65321895265761230930e-02);
        num = num * r + AS241_C(R, 2.96560571828504891230e-01);
        num = num * r + AS241_C(R, 1.78482653991729133580e+00);
        num = num * r + AS241_C(R, 5.46378491116411436990e+00);
        num = num * r + AS241_C(R, 6.65790464350110377720e+00);
        den = AS241_C(R, 2.04426310338993978564e-15);
        den = den * r + AS241_C(R, 1.42151175831644588870e-07);
        den = den * r + AS241_C(R, 1.84631831751005468180e-05);
        den = den * r + AS241_C(R, 7.86869131145613259100e-04);
        den = den * r + AS241_C(R, 1.48753612908506148525e-02);
        den = den * r + AS241_C(R, 1.36929880922735805310e-01);
        den = den * r + AS241_C(R, 5.99832206555887937690e-01);
        den = den * r + (R)1.0;
    }
    const R v = num / den;
    return (q < 0.0) ? -v : v;
}

// Arithmetic of the inverse normal per output type: FP64 increments use the FP64
// transform (bit-exact with the reference's); FP32 increments use the same formula in
// float (about 1e-6 relative).
template <typename OutT> struct RealOf { using type = double; };
template <> struct RealOf<float> { using type = float; };
template <typename OutT> using Real = typename RealOf<OutT>::type;
template <typename OutT> __device__ __forceinline__ OutT icdf_out(double x, double sq) { return (OutT)(x * sq); }
template <typename OutT> __device__ __forceinline__ OutT icdf_out(float x, double sq) { return x * (float)sq; }

// Twist buffer `cur` (old window) into `nxt` (new window). All kGenThreads threads call.
__device__ __forceinline__ void twist_block(const uint32_t* cur, uint32_t* nxt, int t) {
    // phase A: k in [0,227)
    if (t < N - M) nxt[t] = twist_word(cur[t], cur[t + 1], cur[t + M]);
    __syncthreads();
    // phase B: k in [227,454)
    if (t < N - M) { const int k = t + (N - M); nxt[k] = twist_word(cur[k], cur[k + 1], nxt[k - (N - M)]); }
    __syncthreads();
    // phase C: k in [454,624); k = 623 uses the new word 0
    if (t < N - 2 * (N - M)) {
        const int k = t + 2 * (N - M);
        const uint32_t b = (k + 1 < N) ? cur[k + 1] : nxt[0];
        nxt[k] = twist_word(cur[k], b, nxt[k - (N - M)]);
    }
    __syncthreads();
}

// ---------------------------------------------------------------------------
// Brownian generator, one wavefront per sub-stream (four per workgroup, no workgroup
// barriers). The wavefront twists its 624-word state in place in rounds of 64 words,
// in increasing word order as the sequential algorithm does: every word of a round
// reads the old k and k+1 and either the old k+397 or the new k-227 (written in an
// earlier round), and all lanes read before any lane writes. It then tempers the new
// block into 312 uniforms in five rounds. Central-branch draws are transformed and
// stored at once; tail-branch draws are appended (ballot prefix) to the wavefront's
// LDS queue and transformed 64 at a time.
// ---------------------------------------------------------------------------
constexpr int kGenWaves = 4;
constexpr int kTailQ = 128;

// order this wavefront's LDS accesses (the lanes of a wavefront execute together)
__device__ __forceinline__ void wave_lds_order() {
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
}

template <typename OutT>
__global__ void __launch_bounds__(64 * kGenWaves)
brownian_kernel(const uint32_t* __restrict__ states, int64_t num_sub, uint64_t total_words, int log2_len,
                int nT, int F, const double* __restrict__ sqrt_dt, int uniform_dt, OutT* __restrict__ out)
{
    constexpr bool kF32 = sizeof(OutT) == 4;
    __shared__ uint32_t st[kGenWaves][N];
    __shared__ uint64_t tq_bits[kGenWaves][kTailQ];    // 52 uniform bits of a tail draw
    __shared__ uint32_t tq_off[kGenWaves][kTailQ];     // double index relative to the sub-stream's first
    __shared__ uint32_t tq_ti[kGenWaves][kTailQ];     // time index of the draw (any nT)
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int64_t s = (int64_t)blockIdx.x * kGenWaves + wv;
    if (s >= num_sub) return;                          // the whole wavefront leaves
    uint32_t* mt = st[wv];
    const uint64_t w_begin = (uint64_t)s << log2_len;
    uint64_t w_end = w_begin + ((uint64_t)1 << log2_len);
    if (w_end > total_words) w_end = total_words;
    const uint64_t d_begin = w_begin >> 1;
    for (int w = lane; w < N; w += 64) mt[w] = states[(size_t)s * N + w];
    wave_lds_order();
    const double sq0 = sqrt_dt[0];                     // the scale of every step on a uniform grid
    const uint32_t per = (uint32_t)nT * (uint32_t)F;   // doubles per path
    // position within its path of this lane's double in the current round (non-uniform grids)
    uint32_t rem = uniform_dt ? 0u : (uint32_t)((d_begin + (uint64_t)lane) % per);
    const uint32_t step64 = 64u % per, step56 = (uint32_t)(N / 2 - 4 * 64) % per;
    int count = 0;                                     // queued tail draws (uniform)
    // transform the first nq <= 64 queued draws, one per lane
    auto drain = [&](int nq) {
        wave_lds_order();
        if (lane < nq) {
            const double u = unit_offset(tq_bits[wv][lane], 1.0);
            const double sq = uniform_dt ? sq0 : sqrt_dt[tq_ti[wv][lane]];
            out[d_begin + tq_off[wv][lane]] = icdf_out<OutT>(as241_tail<Real<OutT>>(u, u - 0.5), sq);
        }
        wave_lds_order();
    };
    for (uint64_t w = w_begin; w < w_end; w += N) {
        // in-place twist, rounds unrolled so the LDS indices fold into instruction offsets
#pragma unroll
        for (int k0 = 0; k0 < N; k0 += 64) {
            const int k = k0 + lane;
            uint32_t a = 0, b = 0, c = 0;
            if (k < N) {
                a = mt[k];
                b = mt[(k + 1 < N) ? k + 1 : 0];
                c = mt[(k + M < N) ? k + M : k + M - N];
            }
            wave_lds_order();
            if (k < N) mt[k] = twist_word(a, b, c);
            wave_lds_order();
        }
        // 312 uniforms: doubles d = 64 r + lane, words 2d and 2d + 1
#pragma unroll 1
        for (int r = 0; r < 5; r++) {
            const int d = r * 64 + lane;
            const uint64_t wd = w + 2 * (uint64_t)d;
            bool tail = false;
            uint64_t bits = 0;
            const int ti = uniform_dt ? 0 : (int)(rem / (uint32_t)F);
            if (d < N / 2 && wd < w_end) {
                const uint32_t a = temper(mt[2 * d]);
                const uint32_t b = temper(mt[2 * d + 1]);
                bits = ((uint64_t)(a >> 6) << 26) | (uint64_t)(b >> 6);
                const double sq = uniform_dt ? sq0 : sqrt_dt[ti];
                if (kF32) {
                    // q = u - 1/2 from the top 32 of the 52 bits (float keeps 24)
                    const float q = (float)(int32_t)((uint32_t)(bits >> 20) ^ 0x80000000u) * 0x1.0p-32f;
                    if (fabsf(q) <= 0.425f) out[wd >> 1] = (OutT)(as241_central_f32(q) * (float)sq);
                    else tail = true;
                } else {
                    const double q = unit_offset(bits, 1.5);
                    if (fabs(q) <= 0.425) out[wd >> 1] = icdf_out<OutT>(as241_central<double>(q), sq);
                    else tail = true;
                }
            }
            const uint64_t m = __ballot(tail);
            if (tail) {
                const int pos = count + __popcll(m & ((1ull << lane) - 1ull));
                tq_bits[wv][pos] = bits;
                tq_off[wv][pos] = (uint32_t)((wd >> 1) - d_begin);
                tq_ti[wv][pos] = (uint32_t)ti;
            }
            count += __popcll(m);
            if (count >= 64) {
                // a full wavefront of tail draws; the rest (count - 64 < 64) moves to the front
                drain(64);
                count -= 64;
                uint64_t b = 0;
                uint32_t o = 0, t = 0;
                if (lane < count) { b = tq_bits[wv][64 + lane]; o = tq_off[wv][64 + lane]; t = tq_ti[wv][64 + lane]; }
                wave_lds_order();
                if (lane < count) { tq_bits[wv][lane] = b; tq_off[wv][lane] = o; tq_ti[wv][lane] = t; }
                wave_lds_order();
            }
            if (!uniform_dt) {
                rem += (r < 4) ? step64 : step56;
                if (rem >= per) rem -= per;
            }
        }
    }
    if (count > 0) drain(count);
}

__global__ void __launch_bounds__(kGenThreads)
words_kernel(const uint32_t* __restrict__ states, uint64_t total_words, int log2_len, uint32_t* __restrict__ out)
{
    __shared__ uint32_t buf[2][N];
    const int t = threadIdx.x;
    const int64_t s = blockIdx.x;
    const uint64_t w_begin = (uint64_t)s << log2_len;
    uint64_t w_end = w_begin + ((uint64_t)1 << log2_len);
    if (w_end > total_words) w_end = total_words;
    for (int w = t; w < N; w += kGenThreads) buf[0][w] = states[(size_t)s * N + w];
    __syncthreads();
    int c = 0;
    for (uint64_t w = w_begin; w < w_end; w += N) {
        twist_block(buf[c], buf[c ^ 1], t);
        c ^= 1;
        for (int k = t; k < N; k += kGenThreads)
            if (w + k < w_end) out[w + k] = temper(buf[c][k]);
    }
}

} // namespace

void mt_seed_window(int64_t seed, uint32_t* mt) {
    // commons-math3 MersenneTwister.setSeed(long) -> setSeed(int[]{hi, lo}) = init_by_array
    const uint64_t u = (uint64_t)seed;
    const uint32_t key[2] = {(uint32_t)(u >> 32), (uint32_t)(u & 0xffffffffu)};
    mt[0] = 19650218u;
    for (int i = 1; i < N; i++) mt[i] = 1812433253u * (mt[i - 1] ^ (mt[i - 1] >> 30)) + (uint32_t)i;
    int i = 1, j = 0;
    for (int k = N; k; k--) {
        mt[i] = (mt[i] ^ ((mt[i - 1] ^ (mt[i - 1] >> 30)) * 1664525u)) + key[j] + (uint32_t)j;
        i++; j++;
        if (i >= N) { mt[0] = mt[N - 1]; i = 1; }
        if (j >= 2) j = 0;
    }
    for (int k = N - 1; k; k--) {
        mt[i] = (mt[i] ^ ((mt[i - 1] ^ (mt[i - 1] >> 30)) * 1566083941u)) - (uint32_t)i;
        i++;
        if (i >= N) { mt[0] = mt[N - 1]; i = 1; }
    }
    mt[0] = 0x80000000u;
}

StreamLayout choose_layout(uint64_t total_words) {
    StreamLayout l;
    l.total_words = total_words;
    // up to 8192 sub-streams (one wavefront each), each at least 2^12 words (7 twists)
    int k = 12;
    while (k < 32 && (total_words >> k) > 8192) k++;
    l.log2_len = k;
    l.num_sub = (int64_t)((total_words + ((uint64_t)1 << k) - 1) >> k);
    if (l.num_sub < 1) l.num_sub = 1;
    return l;
}

void launch_jump_tree(const StreamLayout& lay, const uint32_t* seed_window, uint32_t* d_states,
                      uint64_t* d_poly, void* stream_) {
    hipStream_t stream = (hipStream_t)stream_;
    HIP_CHECK(hipMemcpyAsync(d_states, seed_window, N * sizeof(uint32_t), hipMemcpyHostToDevice, stream));
    if (lay.num_sub <= 1) return;
    int m = 0;
    while (((int64_t)1 << m) < lay.num_sub) m++;
    // host staging of the m polynomials x^(2^(log2_len + e)) mod phi, e < m
    std::vector<uint64_t> polys((size_t)m * kPolyWords);
    for (int e = 0; e < m; e++) {
        const auto& p = mt_jump_poly_pow2(lay.log2_len + e);
        std::memcpy(polys.data() + (size_t)e * kPolyWords, p.data(), kPolyWords * sizeof(uint64_t));
    }
    HIP_CHECK(hipMemcpyAsync(d_poly, polys.data(), polys.size() * sizeof(uint64_t), hipMemcpyHostToDevice, stream));
    // level e = m-1 .. 0: gap = 2^e sub-streams; sources are multiples of 2^(e+1)
    for (int e = m - 1; e >= 0; e--) {
        const int64_t gap = (int64_t)1 << e;
        const int64_t nsrc = (lay.num_sub + 2 * gap - 1) / (2 * gap);
        hipLaunchKernelGGL(jump_level_kernel, dim3((unsigned)nsrc), dim3(kJumpThreads), 0, stream,
                           d_states, d_poly + (size_t)e * kPolyWords, lay.num_sub, gap);
        HIP_CHECK(hipGetLastError());
    }
    // keep the host staging alive until the copies have been consumed
    HIP_CHECK(hipStreamSynchronize(stream));
}

void launch_brownian(const StreamLayout& lay, const uint32_t* d_states, int nT, int F,
                     const double* d_sqrt_dt, int uniform_dt, void* out, int out_is_float, void* stream_) {
    hipStream_t stream = (hipStream_t)stream_;
    // the tail queue keeps 32-bit offsets within a sub-stream and 16-bit time indices
    if (lay.log2_len > 32 || nT > 65535)
        throw FmError(FM_ERR_UNSUPPORTED, "stream layout outside the generator's index range");
    const dim3 grid((unsigned)((lay.num_sub + kGenWaves - 1) / kGenWaves));
    if (out_is_float)
        hipLaunchKernelGGL(brownian_kernel<float>, grid, dim3(64 * kGenWaves), 0, stream,
                           d_states, lay.num_sub, lay.total_words, lay.log2_len, nT, F, d_sqrt_dt, uniform_dt,
                           (float*)out);
    else
        hipLaunchKernelGGL(brownian_kernel<double>, grid, dim3(64 * kGenWaves), 0, stream,
                           d_states, lay.num_sub, lay.total_words, lay.log2_len, nT, F, d_sqrt_dt, uniform_dt,
                           (double*)out);
    HIP_CHECK(hipGetLastError());
}

void launch_words(const StreamLayout& lay, const uint32_t* d_states, uint64_t count,
                  uint32_t* out, void* stream_) {
    hipStream_t stream = (hipStream_t)stream_;
    hipLaunchKernelGGL(words_kernel, dim3((unsigned)lay.num_sub), dim3(kGenThreads), 0, stream,
                       d_states, std::min<uint64_t>(count, lay.total_words), lay.log2_len, out);
    HIP_CHECK(hipGetLastError());
}

} // namespace fm
