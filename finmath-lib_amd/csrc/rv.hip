// Device-resident random variables: immutable vectors of double realisations, element
// operations recorded as register programs and evaluated in one pass (RandomVariable
// FromDoubleArray semantics, montecarlo/RandomVariableFromDoubleArray.java; lazy
// evaluation as in montecarlo/RandomVariableLazyEvaluation.java).
//
// The interpreter kernel holds 16 double registers per element in VGPRs. Every
// instruction is uniform over the launch (it comes from the kernel argument), so the
// register selection (indexed VGPR moves) and the opcode dispatch (scalar branches)
// cost no vector work beyond a move per operand, and each element costs
// only the arithmetic of the program plus one load per input and one store: a program
// of any length is one HBM pass. The inputs of an element are loaded together before
// the program runs, so their latencies overlap, and each thread runs two elements.
// Reductions (averages, variances) run the same program and sum the result with
// compensated (Neumaier) partial sums in a fixed order, so they are reproducible.
// Compiled with -ffp-contract=off: a + b * c rounds twice, as the JVM evaluates it.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "fm_internal.hpp"
#include "finmath_rv.h"
#include "hip_check.hpp"

struct fm_rv {
    double time = 0.0;
    int64_t size = 1;
    bool det = true;
    double value = 0.0;
    double* d = nullptr;      // device realisations (owned), NULL for a constant
    int device = 0;
};

namespace fm {
namespace {

constexpr int kRvThreads = 256;
constexpr int kRvPerThread = 2;                 // elements per thread per dispatch
constexpr int kRvMaxBlocks = 8192;              // reduction grid (fixed order of partials)

struct RvProg {
    int32_t n, nin, result, mode;               // mode: 0 store, 1 sum, 2 sum of squared deviations
    double mean;                                // mode 2
    int64_t size;
    const double* prob;                         // mode 1: weights (NULL: none)
    double prob_const;
    fm_rv_instr ins[FM_RV_MAX_INSTRUCTIONS];
    const double* in[FM_RV_MAX_INPUTS];         // NULL: constant input
    double cst[FM_RV_MAX_INPUTS];
};
static_assert(sizeof(RvProg) < 4000, "program must fit the kernel argument segment");

__host__ __device__ inline bool sign_bit(double x) {
    uint64_t u;
    __builtin_memcpy(&u, &x, sizeof(u));
    return (u >> 63) != 0;
}

// Java Math.min / Math.max: NaN wins, -0.0 < +0.0.
__host__ __device__ inline double java_min(double a, double b) {
    if (a != a) return a;
    if (b != b) return b;
    if (a == 0.0 && b == 0.0) return sign_bit(a) ? a : b;
    return a <= b ? a : b;
}
__host__ __device__ inline double java_max(double a, double b) {
    if (a != a) return a;
    if (b != b) return b;
    if (a == 0.0 && b == 0.0) return sign_bit(a) ? b : a;
    return a >= b ? a : b;
}

__host__ __device__ inline double apply_op(int op, double a, double b, double c, double imm) {
    switch (op) {
    case FM_RV_CONST: return imm;
    case FM_RV_ADD: return a + b;
    case FM_RV_SUB: return a - b;
    case FM_RV_MULT: return a * b;
    case FM_RV_DIV: return a / b;
    case FM_RV_CAP: return java_min(a, b);
    case FM_RV_FLOOR: return java_max(a, b);
    case FM_RV_POW: return pow(a, b);
    case FM_RV_EXP: return exp(a);
    case FM_RV_LOG: return log(a);
    case FM_RV_SQRT: return sqrt(a);
    case FM_RV_ABS: return fabs(a);
    case FM_RV_INVERT: return 1.0 / a;
    case FM_RV_SQUARED: return a * a;
    case FM_RV_ADD_PRODUCT: return a + b * c;
    case FM_RV_ADD_RATIO: return a + b / c;
    case FM_RV_SUB_RATIO: return a - b / c;
    case FM_RV_ACCRUE: return a * (1 + b * imm);
    case FM_RV_DISCOUNT: return a / (1.0 + b * imm);
    case FM_RV_CHOOSE: return a >= 0.0 ? b : c;
    case FM_RV_ISNAN: return (a != a) ? 1.0 : 0.0;
    case FM_RV_EXPM1: return expm1(a);
    case FM_RV_SIN: return sin(a);
    case FM_RV_COS: return cos(a);
    default: return a;
    }
}

// The 16 registers of an element are one vector value: a uniform index into it lowers to
// the indexed VGPR moves of the hardware (no private memory).
typedef double RegFile __attribute__((ext_vector_type(FM_RV_REGISTERS)));

// Opcodes whose library code needs many registers (pow, expm1, sin, cos): programs that
// use them run on the FULL instantiation, all others on the lighter one.
__host__ __device__ inline bool heavy_op(int op) {
    return op == FM_RV_POW || op == FM_RV_EXPM1 || op == FM_RV_SIN || op == FM_RV_COS;
}

// Run the program on E elements (idx[e] < size when ok[e]); returns the result register.
template <int E, bool FULL>
__device__ __forceinline__ void run_program(const RvProg& g, const int64_t (&idx)[E], const bool (&ok)[E],
                                            double (&out)[E]) {
    RegFile R[E], IN[E];
#pragma unroll
    for (int e = 0; e < E; e++) R[e] = (RegFile)(0.0);
    // every input first, with constant indices: all loads are in flight together
#pragma unroll
    for (int i = 0; i < FM_RV_MAX_INPUTS; i++) {
        if (i < g.nin) {                        // uniform; no early exit, so the loop unrolls
            const double* src = g.in[i];
#pragma unroll
            for (int e = 0; e < E; e++) IN[e][i] = (src != nullptr) ? (ok[e] ? src[idx[e]] : 0.0) : g.cst[i];
        }
    }
    for (int k = 0; k < g.n; k++) {
        const fm_rv_instr in = g.ins[k];
        const int op = __builtin_amdgcn_readfirstlane(in.op);
        const int ia = __builtin_amdgcn_readfirstlane(in.a), ib = __builtin_amdgcn_readfirstlane(in.b);
        const int ic = __builtin_amdgcn_readfirstlane(in.c), id = __builtin_amdgcn_readfirstlane(in.dst);
        double r[E];
        if (op == FM_RV_LOAD) {
#pragma unroll
            for (int e = 0; e < E; e++) r[e] = IN[e][ia];
        } else {
            double a[E], b[E], c[E];
#pragma unroll
            for (int e = 0; e < E; e++) { a[e] = R[e][ia]; b[e] = R[e][ib]; c[e] = R[e][ic]; }
            switch (op) {     // uniform: one arm per launch-wide opcode
#define FM_RV_OP(o) case o: for (int e = 0; e < E; e++) r[e] = apply_op(o, a[e], b[e], c[e], in.imm); break;
            FM_RV_OP(FM_RV_CONST) FM_RV_OP(FM_RV_ADD) FM_RV_OP(FM_RV_SUB) FM_RV_OP(FM_RV_MULT) FM_RV_OP(FM_RV_DIV)
            FM_RV_OP(FM_RV_CAP) FM_RV_OP(FM_RV_FLOOR) FM_RV_OP(FM_RV_EXP) FM_RV_OP(FM_RV_LOG)
            FM_RV_OP(FM_RV_SQRT) FM_RV_OP(FM_RV_ABS) FM_RV_OP(FM_RV_INVERT) FM_RV_OP(FM_RV_SQUARED)
            FM_RV_OP(FM_RV_ADD_PRODUCT) FM_RV_OP(FM_RV_ADD_RATIO) FM_RV_OP(FM_RV_SUB_RATIO) FM_RV_OP(FM_RV_ACCRUE)
            FM_RV_OP(FM_RV_DISCOUNT) FM_RV_OP(FM_RV_CHOOSE) FM_RV_OP(FM_RV_ISNAN)
            default:
                if constexpr (FULL) {
                    switch (op) {
                    FM_RV_OP(FM_RV_POW) FM_RV_OP(FM_RV_EXPM1) FM_RV_OP(FM_RV_SIN) FM_RV_OP(FM_RV_COS)
                    default: for (int e = 0; e < E; e++) r[e] = a[e]; break;
                    }
                } else {
                    for (int e = 0; e < E; e++) r[e] = a[e];
                }
                break;
#undef FM_RV_OP
            }
        }
#pragma unroll
        for (int e = 0; e < E; e++) R[e][id] = r[e];
    }
    const int ir = __builtin_amdgcn_readfirstlane(g.result);
#pragma unroll
    for (int e = 0; e < E; e++) out[e] = R[e][ir];
}

__device__ __forceinline__ void neumaier_add(double& s, double& c, double x) {
    const double t = s + x;
    if (fabs(s) >= fabs(x)) c += (s - t) + x;
    else c += (x - t) + s;
    s = t;
}

// Store mode: out[i] = program(i). Reduction modes: per-block (sum, compensation) of the
// program's value (mode 1, times the weight), or of (value - mean)^2 (mode 2, times the
// weight). Each thread visits its elements in increasing order (grid-stride) and the
// block tree is fixed, so the partials depend only on the size and the grid.
// Two elements per thread share each instruction's dispatch; the light instantiation is
// held to 256 VGPRs (2 waves per SIMD; at 168 for three it spills), the full one takes what
// it needs.
template <bool FULL>
__global__ void __launch_bounds__(kRvThreads) __attribute__((amdgpu_waves_per_eu(FULL ? 1 : 2))) rv_eval_kernel(const RvProg g, double* __restrict__ out,
                                                               double* __restrict__ partials) {
    constexpr int E = kRvPerThread;
    const int64_t span = (int64_t)gridDim.x * kRvThreads * E;
    double s = 0.0, comp = 0.0;
    for (int64_t base = (int64_t)blockIdx.x * kRvThreads * E; base < g.size; base += span) {
        int64_t idx[E];
        bool ok[E];
#pragma unroll
        for (int e = 0; e < E; e++) {
            idx[e] = base + (int64_t)e * kRvThreads + threadIdx.x;
            ok[e] = idx[e] < g.size;
        }
        double v[E];
        run_program<E, FULL>(g, idx, ok, v);
        if (g.mode == 0) {
#pragma unroll
            for (int e = 0; e < E; e++) if (ok[e]) out[idx[e]] = v[e];
        } else {
#pragma unroll
            for (int e = 0; e < E; e++) {
                if (!ok[e]) continue;
                double x = v[e];
                if (g.mode == 2) { const double d = x - g.mean; x = d * d; }
                if (g.prob || g.prob_const != 1.0) x = x * (g.prob ? g.prob[idx[e]] : g.prob_const);
                neumaier_add(s, comp, x);
            }
        }
    }
    if (g.mode == 0) return;
    // fixed-order tree over the block: (sum, compensation) pairs
    __shared__ double ss[kRvThreads], sc[kRvThreads];
    ss[threadIdx.x] = s;
    sc[threadIdx.x] = comp;
    __syncthreads();
    for (int w = kRvThreads / 2; w >= 1; w >>= 1) {
        if ((int)threadIdx.x < w) {
            double a = ss[threadIdx.x], ca = sc[threadIdx.x];
            neumaier_add(a, ca, ss[threadIdx.x + w]);
            ca += sc[threadIdx.x + w];
            ss[threadIdx.x] = a;
            sc[threadIdx.x] = ca;
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        partials[2 * blockIdx.x] = ss[0];
        partials[2 * blockIdx.x + 1] = sc[0];
    }
}

// One workgroup: thread t merges the block partials of its contiguous chunk in order,
// then the fixed tree of the block reduction; result = sum + compensation.
__global__ void __launch_bounds__(kRvThreads) rv_finish_kernel(const double* __restrict__ partials, int blocks,
                                                                double* __restrict__ result) {
    const int chunk = (blocks + kRvThreads - 1) / kRvThreads;
    const int b0 = threadIdx.x * chunk, b1 = min(blocks, b0 + chunk);
    double s = 0.0, c = 0.0;
    for (int b = b0; b < b1; b++) {
        neumaier_add(s, c, partials[2 * b]);
        c += partials[2 * b + 1];
    }
    __shared__ double ss[kRvThreads], sc[kRvThreads];
    ss[threadIdx.x] = s;
    sc[threadIdx.x] = c;
    __syncthreads();
    for (int w = kRvThreads / 2; w >= 1; w >>= 1) {
        if ((int)threadIdx.x < w) {
            double a = ss[threadIdx.x], ca = sc[threadIdx.x];
            neumaier_add(a, ca, ss[threadIdx.x + w]);
            ca += sc[threadIdx.x + w];
            ss[threadIdx.x] = a;
            sc[threadIdx.x] = ca;
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) result[0] = ss[0] + sc[0];
}

__global__ void __launch_bounds__(256) rv_widen_kernel(const float* __restrict__ src, int64_t n,
                                                        double* __restrict__ dst) {
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256)
        dst[i] = (double)src[i];
}

// ---- host side --------------------------------------------------------------------------
std::mutex g_mu;
std::map<int, hipStream_t> g_streams;
thread_local std::string g_rv_error;

hipStream_t lib_stream(int device) {
    std::lock_guard<std::mutex> lk(g_mu);
    auto it = g_streams.find(device);
    if (it != g_streams.end()) return it->second;
    hipStream_t s = nullptr;
    HIP_CHECK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    g_streams[device] = s;
    return s;
}

struct DevGuard {
    int prev = -1;
    explicit DevGuard(int dev) { HIP_CHECK(hipGetDevice(&prev)); if (prev != dev) HIP_CHECK(hipSetDevice(dev)); }
    ~DevGuard() { if (prev >= 0) (void)hipSetDevice(prev); }
};

void need(bool cond, int code, const std::string& msg) { if (!cond) throw FmError(code, msg); }

int rv_fail(const std::exception& e) {
    if (auto* fe = dynamic_cast<const FmError*>(&e)) { set_last_error(fe->what()); return fe->code; }
    if (dynamic_cast<const std::bad_alloc*>(&e)) { set_last_error("host out of memory"); return FM_ERR_OUT_OF_MEMORY; }
    set_last_error(e.what());
    return FM_ERR_INTERNAL;
}
#define RV_TRY try {
#define RV_CATCH } catch (const std::exception& e) { return rv_fail(e); } catch (...) { set_last_error("unknown error"); return FM_ERR_INTERNAL; }

fm_rv* new_device_rv(int device, double time, int64_t size) {
    std::unique_ptr<fm_rv> r(new fm_rv());
    r->time = time; r->size = size; r->det = false; r->device = device;
    DevGuard g(device);
    const hipError_t e = hipMallocAsync((void**)&r->d, (size_t)size * sizeof(double), lib_stream(device));
    if (e != hipSuccess) {
        (void)hipGetLastError();
        throw FmError(FM_ERR_OUT_OF_MEMORY, "device allocation of " + std::to_string(size) + " realisations failed");
    }
    return r.release();
}

// Validate a program and its inputs; returns the common size (0: all inputs constant).
int64_t check_program(int32_t nin, const fm_rv* const* inputs, int32_t nins, const fm_rv_instr* prog, int32_t result,
                      int* device) {
    need(nin >= 0 && nin <= FM_RV_MAX_INPUTS, FM_ERR_INVALID_ARGUMENT, "at most 16 inputs");
    need(nins >= 1 && nins <= FM_RV_MAX_INSTRUCTIONS, FM_ERR_INVALID_ARGUMENT, "program must have 1..64 instructions");
    need(prog != nullptr && (nin == 0 || inputs != nullptr), FM_ERR_INVALID_ARGUMENT, "NULL program or inputs");
    need(result >= 0 && result < FM_RV_REGISTERS, FM_ERR_INVALID_ARGUMENT, "result register out of range");
    int64_t size = 0;
    *device = -1;
    for (int i = 0; i < nin; i++) {
        need(inputs[i] != nullptr, FM_ERR_INVALID_ARGUMENT, "NULL input");
        if (inputs[i]->det) continue;
        need(size == 0 || size == inputs[i]->size, FM_ERR_INVALID_ARGUMENT, "inputs of different sizes");
        need(*device < 0 || *device == inputs[i]->device, FM_ERR_INVALID_ARGUMENT, "inputs on different devices");
        size = inputs[i]->size;
        *device = inputs[i]->device;
    }
    for (int k = 0; k < nins; k++) {
        const fm_rv_instr& in = prog[k];
        need(in.op >= FM_RV_LOAD && in.op <= FM_RV_COS, FM_ERR_INVALID_ARGUMENT, "unknown opcode " + std::to_string(in.op));
        need(in.dst >= 0 && in.dst < FM_RV_REGISTERS, FM_ERR_INVALID_ARGUMENT, "destination register out of range");
        if (in.op == FM_RV_LOAD) {
            need(in.a >= 0 && in.a < nin, FM_ERR_INVALID_ARGUMENT, "LOAD of an input out of range");
        } else {
            need(in.a >= 0 && in.a < FM_RV_REGISTERS && in.b >= 0 && in.b < FM_RV_REGISTERS && in.c >= 0 &&
                 in.c < FM_RV_REGISTERS, FM_ERR_INVALID_ARGUMENT, "source register out of range");
        }
    }
    return size;
}

double eval_host(int32_t nin, const fm_rv* const* inputs, int32_t nins, const fm_rv_instr* prog, int32_t result) {
    double R[FM_RV_REGISTERS] = {};
    for (int k = 0; k < nins; k++) {
        const fm_rv_instr& in = prog[k];
        R[in.dst] = in.op == FM_RV_LOAD ? inputs[in.a]->value : apply_op(in.op, R[in.a], R[in.b], R[in.c], in.imm);
    }
    (void)nin;
    return R[result];
}

RvProg make_prog(int32_t nin, const fm_rv* const* inputs, int32_t nins, const fm_rv_instr* prog, int32_t result,
                 int64_t size) {
    RvProg g;
    std::memset(&g, 0, sizeof(g));
    g.n = nins; g.nin = nin; g.result = result; g.size = size;
    for (int k = 0; k < nins; k++) g.ins[k] = prog[k];
    for (int i = 0; i < nin; i++) {
        g.in[i] = inputs[i]->det ? nullptr : inputs[i]->d;
        g.cst[i] = inputs[i]->value;
    }
    return g;
}

void launch_eval(const RvProg& g, int blocks, hipStream_t s, double* out, double* partials) {
    bool full = false;
    for (int k = 0; k < g.n; k++) full = full || heavy_op(g.ins[k].op);
    if (full) hipLaunchKernelGGL(rv_eval_kernel<true>, dim3(blocks), dim3(kRvThreads), 0, s, g, out, partials);
    else hipLaunchKernelGGL(rv_eval_kernel<false>, dim3(blocks), dim3(kRvThreads), 0, s, g, out, partials);
    HIP_CHECK(hipGetLastError());
}

int reduce_blocks(int64_t size) {
    const int64_t per = (int64_t)kRvThreads * kRvPerThread;
    return (int)std::max<int64_t>(1, std::min<int64_t>(kRvMaxBlocks, (size + per - 1) / per));
}

// Sum of the program's (mode 1) or squared-deviation (mode 2) values over the realisations.
double reduce_program(RvProg g, int device) {
    DevGuard dg(device);
    hipStream_t s = lib_stream(device);
    const int blocks = reduce_blocks(g.size);
    double* buf = nullptr;
    HIP_CHECK(hipMallocAsync((void**)&buf, (size_t)(2 * blocks + 1) * sizeof(double), s));
    launch_eval(g, blocks, s, nullptr, buf);
    hipLaunchKernelGGL(rv_finish_kernel, dim3(1), dim3(kRvThreads), 0, s, buf, blocks, buf + 2 * blocks);
    HIP_CHECK(hipGetLastError());
    double r = 0.0;
    HIP_CHECK(hipMemcpyAsync(&r, buf + 2 * blocks, sizeof(double), hipMemcpyDeviceToHost, s));
    HIP_CHECK(hipFreeAsync(buf, s));
    HIP_CHECK(hipStreamSynchronize(s));
    return r;
}

} // namespace

// Random variable from device realisations of a context (double or float), copied.
fm_rv* rv_from_device(int device, const void* src, bool is_float, int64_t size, double time) {
    fm_rv* r = new_device_rv(device, time, size);
    DevGuard g(device);
    hipStream_t s = lib_stream(device);
    if (is_float) {
        const int blocks = (int)std::min<int64_t>(4096, (size + 255) / 256);
        hipLaunchKernelGGL(rv_widen_kernel, dim3(blocks), dim3(256), 0, s, (const float*)src, size, r->d);
        HIP_CHECK(hipGetLastError());
    } else {
        HIP_CHECK(hipMemcpyAsync(r->d, src, (size_t)size * sizeof(double), hipMemcpyDeviceToDevice, s));
    }
    HIP_CHECK(hipStreamSynchronize(s));        // the source may be overwritten by the next simulation
    return r;
}

fm_rv* rv_constant(double time, double value) {
    fm_rv* r = new fm_rv();
    r->time = time; r->det = true; r->value = value; r->size = 1;
    return r;
}

// out = (x * zb) / adj (or (1/adj) (x zb) when scalar_form), the numeraire adjustment.
fm_rv* rv_scale_numeraire(fm_rv* x, double zb, double adj, bool scalar_form) {
    const fm_rv* in[1] = {x};
    fm_rv_instr p[5];
    std::memset(p, 0, sizeof(p));
    p[0].op = FM_RV_LOAD; p[0].dst = 0; p[0].a = 0;
    p[1].op = FM_RV_CONST; p[1].dst = 1; p[1].imm = zb;
    p[2].op = FM_RV_MULT; p[2].dst = 0; p[2].a = 0; p[2].b = 1;
    p[3].op = FM_RV_CONST; p[3].dst = 1; p[3].imm = scalar_form ? 1.0 / adj : adj;
    p[4].op = scalar_form ? FM_RV_MULT : FM_RV_DIV; p[4].dst = 0;
    if (scalar_form) { p[4].a = 1; p[4].b = 0; } else { p[4].a = 0; p[4].b = 1; }
    fm_rv* out = nullptr;
    const int rc = fm_rv_evaluate(1, in, 5, p, 0, x->time, &out);
    if (rc != FM_OK) throw FmError(rc, "numeraire adjustment failed");
    return out;
}

} // namespace fm

using namespace fm;

extern "C" {

int fm_rv_from_host(double time, int64_t size, const double* values, fm_rv** out) {
    RV_TRY
    need(out != nullptr && values != nullptr && size >= 1, FM_ERR_INVALID_ARGUMENT, "invalid arguments");
    *out = nullptr;
    int dev = 0;
    HIP_CHECK(hipGetDevice(&dev));
    std::unique_ptr<fm_rv> r(new_device_rv(dev, time, size));
    hipStream_t s = lib_stream(dev);
    HIP_CHECK(hipMemcpyAsync(r->d, values, (size_t)size * sizeof(double), hipMemcpyHostToDevice, s));
    HIP_CHECK(hipStreamSynchronize(s));
    *out = r.release();
    return FM_OK;
    RV_CATCH
}

int fm_rv_constant(double time, double value, fm_rv** out) {
    RV_TRY
    need(out != nullptr, FM_ERR_INVALID_ARGUMENT, "NULL output");
    *out = rv_constant(time, value);
    return FM_OK;
    RV_CATCH
}

void fm_rv_release(fm_rv* rv) {
    if (!rv) return;
    if (rv->d) {
        try {
            DevGuard g(rv->device);
            (void)hipFreeAsync(rv->d, lib_stream(rv->device));
        } catch (...) {
        }
    }
    delete rv;
}

int fm_rv_info(const fm_rv* rv, double* time, int64_t* size, int32_t* is_deterministic, double* value) {
    RV_TRY
    need(rv != nullptr, FM_ERR_INVALID_ARGUMENT, "NULL random variable");
    if (time) *time = rv->time;
    if (size) *size = rv->size;
    if (is_deterministic) *is_deterministic = rv->det ? 1 : 0;
    if (value) *value = rv->det ? rv->value : std::nan("");
    return FM_OK;
    RV_CATCH
}

int fm_rv_to_host(const fm_rv* rv, int64_t size, double* out) {
    RV_TRY
    need(rv != nullptr && out != nullptr, FM_ERR_INVALID_ARGUMENT, "NULL argument");
    if (rv->det) { for (int64_t i = 0; i < size; i++) out[i] = rv->value; return FM_OK; }
    need(size == rv->size, FM_ERR_INVALID_ARGUMENT, "size differs from the number of realisations");
    DevGuard g(rv->device);
    hipStream_t s = lib_stream(rv->device);
    HIP_CHECK(hipMemcpyAsync(out, rv->d, (size_t)size * sizeof(double), hipMemcpyDeviceToHost, s));
    HIP_CHECK(hipStreamSynchronize(s));
    return FM_OK;
    RV_CATCH
}

int fm_rv_device_ptr(const fm_rv* rv, const double** out) {
    RV_TRY
    need(rv != nullptr && out != nullptr, FM_ERR_INVALID_ARGUMENT, "NULL argument");
    *out = rv->d;
    return FM_OK;
    RV_CATCH
}

int fm_rv_evaluate(int32_t nin, const fm_rv* const* inputs, int32_t nins, const fm_rv_instr* prog, int32_t result,
                   double time, fm_rv** out) {
    RV_TRY
    need(out != nullptr, FM_ERR_INVALID_ARGUMENT, "NULL output");
    *out = nullptr;
    int dev = -1;
    const int64_t size = check_program(nin, inputs, nins, prog, result, &dev);
    if (size == 0) {                    // all inputs deterministic
        *out = rv_constant(time, eval_host(nin, inputs, nins, prog, result));
        return FM_OK;
    }
    std::unique_ptr<fm_rv> r(new_device_rv(dev, time, size));
    RvProg g = make_prog(nin, inputs, nins, prog, result, size);
    g.mode = 0;
    DevGuard dg(dev);
    const int64_t per = (int64_t)kRvThreads * kRvPerThread;
    const int blocks = (int)std::min<int64_t>(8192, (size + per - 1) / per);
    launch_eval(g, blocks, lib_stream(dev), r->d, nullptr);
    *out = r.release();
    return FM_OK;
    RV_CATCH
}

int fm_rv_evaluate_average(int32_t nin, const fm_rv* const* inputs, int32_t nins, const fm_rv_instr* prog,
                           int32_t result, const fm_rv* probabilities, double* average, double* variance) {
    RV_TRY
    need(average != nullptr, FM_ERR_INVALID_ARGUMENT, "NULL average");
    int dev = -1;
    int64_t size = check_program(nin, inputs, nins, prog, result, &dev);
    if (probabilities && !probabilities->det) {
        need(size == 0 || size == probabilities->size, FM_ERR_INVALID_ARGUMENT, "probabilities of a different size");
        need(dev < 0 || dev == probabilities->device, FM_ERR_INVALID_ARGUMENT, "probabilities on another device");
        if (size == 0) { size = probabilities->size; dev = probabilities->device; }
    }
    if (size == 0) {                    // deterministic (RandomVariableFromDoubleArray.getAverage :288-290)
        const double v = eval_host(nin, inputs, nins, prog, result);
        *average = probabilities ? v * probabilities->value : v;
        if (variance) *variance = 0.0;
        return FM_OK;
    }
    RvProg g = make_prog(nin, inputs, nins, prog, result, size);
    g.mode = 1;
    g.prob = probabilities ? probabilities->d : nullptr;
    g.prob_const = probabilities ? probabilities->value : 1.0;
    const double mean = reduce_program(g, dev) / (double)size;
    *average = mean;
    if (variance) {
        // getVariance: sum((x - avg)^2) / P; getVariance(probabilities): sum((x - avg)^2 p)
        if (size == 1 && !probabilities) { *variance = 0.0; return FM_OK; }
        g.mode = 2;
        g.mean = mean;
        const double ss = reduce_program(g, dev);
        *variance = probabilities ? ss : ss / (double)size;
    }
    return FM_OK;
    RV_CATCH
}

int fm_rv_average(const fm_rv* rv, double* average) {
    const fm_rv* in[1] = {rv};
    fm_rv_instr p{};
    p.op = FM_RV_LOAD; p.dst = 0; p.a = 0;
    return fm_rv_evaluate_average(1, in, 1, &p, 0, nullptr, average, nullptr);
}

int fm_rv_variance(const fm_rv* rv, double* variance) {
    const fm_rv* in[1] = {rv};
    fm_rv_instr p{};
    p.op = FM_RV_LOAD; p.dst = 0; p.a = 0;
    double avg = 0.0;
    return fm_rv_evaluate_average(1, in, 1, &p, 0, nullptr, &avg, variance);
}

} // extern "C"
