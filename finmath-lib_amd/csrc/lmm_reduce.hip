// Reductions of per-wavefront partials into Monte Carlo averages and variances.
//
// The reference averages with sequential Kahan summation and computes the
// variance as a second (Kahan) pass over (x - mean)^2
// (montecarlo/RandomVariableFromDoubleArray.java:287-307,333-355). Here each
// wavefront contributes (count, sum, centred M2); partials are merged in a fixed
// order (per-thread sequential Chan/Neumaier merge, then a fixed LDS tree), so the
// result is deterministic run to run.

#include <hip/hip_runtime.h>
#include <algorithm>
#include "fm_internal.hpp"
#include "hip_check.hpp"
#include "lmm_step.hpp"

namespace fm {

namespace {

constexpr int kReduceThreads = 256;

struct Acc { double n, s, c, m2; };   // count, sum, Neumaier compensation, centred M2

__device__ __forceinline__ void merge(Acc& a, const Acc& b) {
    if (b.n == 0.0) return;
    if (a.n == 0.0) { a = b; return; }
    const double sa = a.s + a.c, sb = b.s + b.c;
    const double d = sb / b.n - sa / a.n;
    const double nn = a.n + b.n;
    a.m2 = a.m2 + b.m2 + d * d * (a.n * b.n / nn);
    // Neumaier sum of the two sums
    const double t = a.s + b.s;
    if (fabs(a.s) >= fabs(b.s)) a.c += (a.s - t) + b.s; else a.c += (b.s - t) + a.s;
    a.c += b.c;
    a.s = t;
    a.n = nn;
}

__global__ void __launch_bounds__(kReduceThreads)
reduce_partials_kernel(const double* __restrict__ wave_sum, const double* __restrict__ wave_m2,
                       int num_waves, int stride, int num_slots, int P,
                       double* __restrict__ out_sum, double* __restrict__ out_m2)
{
    __shared__ Acc sh[kReduceThreads];
    const int col = blockIdx.x;
    const bool has_m2 = col < num_slots;
    Acc a{0.0, 0.0, 0.0, 0.0};
    for (int w = threadIdx.x; w < num_waves; w += kReduceThreads) {
        const int64_t left = (int64_t)P - (int64_t)w * 64;
        Acc b{(double)(left > 64 ? 64 : left), wave_sum[(size_t)w * stride + col], 0.0,
              has_m2 ? wave_m2[(size_t)w * stride + col] : 0.0};
        merge(a, b);
    }
    sh[threadIdx.x] = a;
    __syncthreads();
    for (int h = kReduceThreads / 2; h >= 1; h >>= 1) {
        if ((int)threadIdx.x < h) { Acc x = sh[threadIdx.x]; merge(x, sh[threadIdx.x + h]); sh[threadIdx.x] = x; }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        out_sum[col] = sh[0].s + sh[0].c;
        out_m2[col] = sh[0].m2;
    }
}

// Per path: value of each swap = sum over its periods (in period order) of the
// contribution payoff/N(T_pay) times the numeraire-adjustment scale of that date,
// then the wavefront partials of the swap's slot (SimpleSwap.java:140-164).
__global__ void __launch_bounds__(kStepThreads)
swap_finalize_kernel(const double* __restrict__ contrib, int P, int num_swaps,
                     const int32_t* __restrict__ col0, const int32_t* __restrict__ ncol,
                     const int32_t* __restrict__ slot, const double* __restrict__ scale,
                     double* __restrict__ wave_sum, double* __restrict__ wave_m2, int stride)
{
    const int64_t p = (int64_t)blockIdx.x * kStepThreads + threadIdx.x;
    const int lane = threadIdx.x & 63;
    const int64_t wave = p >> 6;
    const bool active = p < P;
    const int64_t pc = active ? p : (int64_t)P - 1;
    const int cnt = (int)((int64_t)P - wave * 64 > 64 ? 64 : ((int64_t)P - wave * 64));
    for (int q = 0; q < num_swaps; q++) {
        double v = 0.0;
        for (int c = 0; c < ncol[q]; c++) {
            const int col = col0[q] + c;
            v = v + contrib[(size_t)col * P + pc] * scale[col];
        }
        reduce_slot(v, active, cnt, lane, wave, wave_sum, wave_m2, stride, slot[q]);
    }
}

// Column sums of a row-major table [num_rows][stride] (the per-wavefront gradient
// partials): workgroup (x, y) sums 64 adjacent columns over row chunk y, each wavefront
// taking every fourth row (coalesced 512-byte reads), Neumaier-compensated; a second
// kernel adds the chunk partials in chunk order. Deterministic.
constexpr int kColChunks = 16;

__global__ void __launch_bounds__(256)
reduce_columns_kernel(const double* __restrict__ rows, int num_rows, int stride, int cols,
                      double* __restrict__ partial)
{
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int col = blockIdx.x * 64 + lane;
    const int chunk = (num_rows + (int)gridDim.y - 1) / (int)gridDim.y;
    const int r0 = (int)blockIdx.y * chunk;
    const int r1 = min(num_rows, r0 + chunk);
    double sum = 0.0, c = 0.0;
    if (col < cols) {
#pragma unroll 4
        for (int r = r0 + w; r < r1; r += 4) {
            const double x = rows[(size_t)r * stride + col];
            const double t = sum + x;
            if (fabs(sum) >= fabs(x)) c += (sum - t) + x; else c += (x - t) + sum;
            sum = t;
        }
    }
    __shared__ double red[4][64];
    red[w][lane] = sum + c;
    __syncthreads();
    if (w == 0 && col < cols)
        partial[(size_t)blockIdx.y * cols + col] = ((red[0][lane] + red[1][lane]) + red[2][lane]) + red[3][lane];
}

__global__ void __launch_bounds__(256)
sum_chunks_kernel(const double* __restrict__ partial, int chunks, int cols, double* __restrict__ out)
{
    const int col = blockIdx.x * 256 + threadIdx.x;
    if (col >= cols) return;
    double s = 0.0;
    for (int y = 0; y < chunks; y++) s += partial[(size_t)y * cols + col];
    out[col] = s;
}

} // namespace

void launch_reduce_columns(const double* rows, int num_rows, int stride, int cols, double* scratch, double* out,
                           void* stream) {
    if (cols <= 0) return;
    const dim3 grid((unsigned)((cols + 63) / 64), (unsigned)kColChunks);
    hipLaunchKernelGGL(reduce_columns_kernel, grid, dim3(256), 0, (hipStream_t)stream, rows, num_rows, stride, cols,
                       scratch);
    HIP_CHECK(hipGetLastError());
    hipLaunchKernelGGL(sum_chunks_kernel, dim3((unsigned)((cols + 255) / 256)), dim3(256), 0, (hipStream_t)stream,
                       scratch, kColChunks, cols, out);
    HIP_CHECK(hipGetLastError());
}

int reduce_columns_scratch(int cols) { return kColChunks * cols; }

static __global__ void __launch_bounds__(256) fill_zero_kernel(double* __restrict__ p, size_t count) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += (size_t)gridDim.x * blockDim.x)
        p[i] = 0.0;
}

void launch_fill_zero(double* p, size_t count, void* stream) {
    if (count == 0) return;
    const size_t blocks = std::min<size_t>((count + 255) / 256, 4096);
    hipLaunchKernelGGL(fill_zero_kernel, dim3((unsigned)blocks), dim3(256), 0, (hipStream_t)stream, p, count);
    HIP_CHECK(hipGetLastError());
}

void launch_reduce_partials(const double* wave_sum, const double* wave_m2, int num_waves, int stride,
                            int num_slots, int num_sums, int P, double* out_sum, double* out_m2, void* stream) {
    const int cols = num_slots + num_sums;
    if (cols <= 0) return;
    hipLaunchKernelGGL(reduce_partials_kernel, dim3((unsigned)cols), dim3(kReduceThreads), 0, (hipStream_t)stream,
                       wave_sum, wave_m2, num_waves, stride, num_slots, P, out_sum, out_m2);
    HIP_CHECK(hipGetLastError());
}

void launch_swap_finalize(const double* contrib, int P, int num_swaps, const int32_t* col0, const int32_t* ncol,
                          const int32_t* slot, const double* scale, double* wave_sum, double* wave_m2,
                          int stride, void* stream) {
    const int64_t blocks = ((int64_t)P + kStepThreads - 1) / kStepThreads;
    hipLaunchKernelGGL(swap_finalize_kernel, dim3((unsigned)blocks), dim3(kStepThreads), 0, (hipStream_t)stream,
                       contrib, P, num_swaps, col0, ncol, slot, scale, wave_sum, wave_m2, stride);
    HIP_CHECK(hipGetLastError());
}

} // namespace fm
