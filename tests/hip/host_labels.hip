// Test harness: the device label function (fill_label, pearl_kernels.hpp)
// compiled for the host, so that the labelling logic of k_fill is checked
// against the reference's filling_phase_at on the CPU.  The host build uses
// the C library's sin, as the reference does.
#include "pearl_gpu.h"
#include "pearl_kernels.hpp"

extern "C" int host_labels(const pearl_fill* ops, int count, int nx, int ny, double* phi) {
    const long n = static_cast<long>(nx) * ny;
    for (int gy = 0; gy < ny; ++gy)
        for (int gx = 0; gx < nx; ++gx) {
            int s = 2;
            for (int k = 0; k < count; ++k)
                s = pearl_gpu::fill_label(*reinterpret_cast<const pearl_gpu::FillOp*>(&ops[k]), gx, gy, s);
            const long i = static_cast<long>(gy) * nx + gx;
            phi[i] = s == 0 ? 1.0 : 0.0;
            phi[n + i] = s == 1 ? 1.0 : 0.0;
            phi[2 * n + i] = s == 2 ? 1.0 : 0.0;
        }
    return 0;
}
