// Test-only probe of FP64 rounding on the device: square root and division
// as compiled for the solver (same flags), and the solver's own
// constant-divisor quotient pearl_gpu::div_const (two Markstein corrections
// of a*RN(1/d)).  The host compares the results with correctly rounded host
// arithmetic.
#include <hip/hip_runtime.h>

#include "pearl_kernels.hpp"

namespace {
__global__ void k_sqrt(const double* __restrict__ x, double* __restrict__ y, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) y[i] = __builtin_sqrt(x[i]);
}
__global__ void k_div(const double* __restrict__ a, const double* __restrict__ b, double* __restrict__ q, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) q[i] = a[i] / b[i];
}
__global__ void k_div_const(const double* __restrict__ a, double d, double y, double* __restrict__ q, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) q[i] = pearl_gpu::div_const(a[i], d, y);
}

template <typename F>
int run(F launch, const double* h_in0, const double* h_in1, double* h_out, int n) {
    double *a = nullptr, *b = nullptr, *o = nullptr;
    const size_t bytes = static_cast<size_t>(n) * sizeof(double);
    if (hipMalloc(&a, bytes) != hipSuccess) return 1;
    if (hipMalloc(&o, bytes) != hipSuccess) return 1;
    if (h_in1 && hipMalloc(&b, bytes) != hipSuccess) return 1;
    if (hipMemcpy(a, h_in0, bytes, hipMemcpyHostToDevice) != hipSuccess) return 2;
    if (h_in1 && hipMemcpy(b, h_in1, bytes, hipMemcpyHostToDevice) != hipSuccess) return 2;
    launch(a, b, o, n);
    if (hipGetLastError() != hipSuccess) return 3;
    if (hipMemcpy(h_out, o, bytes, hipMemcpyDeviceToHost) != hipSuccess) return 4;
    (void)hipFree(a);
    (void)hipFree(o);
    if (b) (void)hipFree(b);
    return 0;
}
}  // namespace

extern "C" int probe_sqrt(const double* x, double* y, int n) {
    return run([](double* a, double*, double* o, int m) {
        hipLaunchKernelGGL(k_sqrt, dim3((m + 255) / 256), dim3(256), 0, 0, a, o, m);
    }, x, nullptr, y, n);
}

extern "C" int probe_div(const double* a, const double* b, double* q, int n) {
    return run([](double* x, double* y, double* o, int m) {
        hipLaunchKernelGGL(k_div, dim3((m + 255) / 256), dim3(256), 0, 0, x, y, o, m);
    }, a, b, q, n);
}

extern "C" int probe_div_const(const double* a, double d, double inv, double* q, int n) {
    return run([d, inv](double* x, double*, double* o, int m) {
        hipLaunchKernelGGL(k_div_const, dim3((m + 255) / 256), dim3(256), 0, 0, x, d, inv, o, m);
    }, a, nullptr, q, n);
}
