// Test-only probe of the pair kernel's FP64 elementary functions (md_sim_amd/csrc/mdf_math.h):
// evaluates them, and the hardware estimates they refine, on host arrays so that a test can
// compare them with correctly rounded values.  Not part of the engine.
#include <hip/hip_runtime.h>

#include "../../md_sim_amd/csrc/mdf_math.h"

namespace {

__global__ void k_math_probe(const double *x, int n, double *ex, double *rs, double *rc, double *rs0, double *rc0) {
    __shared__ double tab[64];
    mdf::exp2_table_load(tab);
    __syncthreads();
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double v = x[i];
    ex[i] = mdf::exp_neg(-v, tab);
    rs[i] = mdf::rsqrt_nr(v);
    rc[i] = mdf::rcp_nr(v);
    rs0[i] = __builtin_amdgcn_rsq(v);
    rc0[i] = __builtin_amdgcn_rcp(v);
}

}  // namespace

// For each x[i] >= 0: out[0][i] = e^-x, out[1][i] = 1/sqrt(x), out[2][i] = 1/x (the engine's
// functions), out[3][i] and out[4][i] the hardware estimates of 1/sqrt(x) and 1/x.  `out` is
// 5 n doubles on the host.  Returns the first HIP error code, 0 on success.
extern "C" int mdf_math_probe(const double *x, int n, double *out) {
    if (n <= 0) return 0;
    double *d = nullptr;
    const size_t b = (size_t)n * sizeof(double);
    hipError_t e = hipMalloc((void **)&d, 6 * b);
    if (e != hipSuccess) return (int)e;
    e = hipMemcpy(d, x, b, hipMemcpyHostToDevice);
    if (e == hipSuccess) {
        const int blocks = (n + 255) / 256;
        k_math_probe<<<blocks, 256>>>(d, n, d + n, d + 2 * (size_t)n, d + 3 * (size_t)n, d + 4 * (size_t)n,
                                      d + 5 * (size_t)n);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpy(out, d + n, 5 * b, hipMemcpyDeviceToHost);
    const hipError_t ef = hipFree(d);
    return (int)(e != hipSuccess ? e : ef);
}
