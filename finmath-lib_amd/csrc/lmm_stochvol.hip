// Stochastic volatility scaling of the covariance model
// (covariance/LIBORCovarianceModelStochasticVolatility.java:174-256): the scalar process
// S = exp(X) simulated with the functional Euler scheme of EulerSchemeFromProcessModel
// (montecarlo/process/EulerSchemeFromProcessModel.java:232-256): X = log S(t_i), then
// X + drift dt, then + l0 dW_a, then + l1 dW_b, S(t_{i+1}) = exp(X), S(0) = exp(0) = 1.
// One thread per path; the driver's increments of a path are contiguous. Compiled with
// -ffp-contract=off like the FP64 path kernel.
#include <hip/hip_runtime.h>
#include "fm_internal.hpp"
#include "fm_math.hpp"
#include "hip_check.hpp"

namespace fm {

template <typename out_t>
__global__ void __launch_bounds__(256)
stochastic_volatility_kernel(const double* __restrict__ dw, int nT, int Fd, int fa, int fb, int P,
                             const double* __restrict__ times, double drift, double l0, double l1,
                             out_t* __restrict__ out) {
    const int64_t p = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (p >= P) return;
    const double* w = dw + (size_t)p * nT * Fd;
    double S = 1.0;
    for (int i = 0; i < nT; i++) {
        out[(size_t)i * P + p] = (out_t)S;
        double X = fm_log(S);
        X = X + drift * (times[i + 1] - times[i]);
        X = X + l0 * w[(size_t)i * Fd + fa];
        X = X + l1 * w[(size_t)i * Fd + fb];
        S = exp(X);
    }
}

void launch_stochastic_volatility(const double* dw, int nT, int Fd, int fa, int fb, int P, const double* times,
                                  double drift, double l0, double l1, void* out, int out_fp32, void* stream) {
    const unsigned blocks = (unsigned)(((int64_t)P + 255) / 256);
    if (out_fp32)
        hipLaunchKernelGGL(stochastic_volatility_kernel<float>, dim3(blocks), dim3(256), 0, (hipStream_t)stream, dw,
                           nT, Fd, fa, fb, P, times, drift, l0, l1, (float*)out);
    else
        hipLaunchKernelGGL(stochastic_volatility_kernel<double>, dim3(blocks), dim3(256), 0, (hipStream_t)stream,
                           dw, nT, Fd, fa, fb, P, times, drift, l0, l1, (double*)out);
    HIP_CHECK(hipGetLastError());
}

// The variance process of LIBORCovarianceModelStochasticHestonVolatility
// (covariance/LIBORCovarianceModelStochasticHestonVolatility.java:172-256): V(0) = 1, functional
// Euler with the identity transform, drift (V - theta)(-kappa) and loading sqrt(max(V, 0)) xi on
// one factor; the scaling of the loadings at t_i is sqrt(max(V(t_i), 0)).
template <typename out_t>
__global__ void __launch_bounds__(256)
heston_volatility_kernel(const double* __restrict__ dw, int nT, int Fd, int fa, int P,
                         const double* __restrict__ times, double kappa, double theta, double xi,
                         out_t* __restrict__ out) {
    const int64_t p = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (p >= P) return;
    const double* w = dw + (size_t)p * nT * Fd;
    double V = 1.0;
    for (int i = 0; i < nT; i++) {
        const double Vf = (V != V) ? V : (V >= 0.0 ? V : 0.0);       // Math.max(V, 0.0)
        const double root = sqrt(Vf);
        out[(size_t)i * P + p] = (out_t)root;
        const double drift = (V - theta) * (kappa * -1.0);
        double X = V;
        X = X + drift * (times[i + 1] - times[i]);
        X = X + (root * xi) * w[(size_t)i * Fd + fa];
        V = X;
    }
}

void launch_heston_volatility(const double* dw, int nT, int Fd, int fa, int P, const double* times, double kappa,
                              double theta, double xi, void* out, int out_fp32, void* stream) {
    const unsigned blocks = (unsigned)(((int64_t)P + 255) / 256);
    if (out_fp32)
        hipLaunchKernelGGL(heston_volatility_kernel<float>, dim3(blocks), dim3(256), 0, (hipStream_t)stream, dw, nT,
                           Fd, fa, P, times, kappa, theta, xi, (float*)out);
    else
        hipLaunchKernelGGL(heston_volatility_kernel<double>, dim3(blocks), dim3(256), 0, (hipStream_t)stream, dw,
                           nT, Fd, fa, P, times, kappa, theta, xi, (double*)out);
    HIP_CHECK(hipGetLastError());
}

} // namespace fm
