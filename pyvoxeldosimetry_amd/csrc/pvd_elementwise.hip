// Elementwise weighted sum of T float maps: out = sum_t w[t] * maps[t].
// Used for activity time integration (trapezoid weights computed in FP64 on
// the host, rounded once to FP32).  float4 path for 16-byte-aligned maps.
#include "pvd_internal.h"

struct pvd_wsum_args {
    const float *m[PVD_MAX_TIMEPOINTS];
    float w[PVD_MAX_TIMEPOINTS];
    int T;
};

__global__ __launch_bounds__(256) void pvd_wsum4(pvd_wsum_args a, long long n4, float4 *__restrict__ out) {
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n4;
         i += (long long)gridDim.x * blockDim.x) {
        float4 v = reinterpret_cast<const float4 *>(a.m[0])[i];
        float4 acc = make_float4(a.w[0] * v.x, a.w[0] * v.y, a.w[0] * v.z, a.w[0] * v.w);
        for (int t = 1; t < a.T; ++t) {
            v = reinterpret_cast<const float4 *>(a.m[t])[i];
            acc.x = fmaf(a.w[t], v.x, acc.x);
            acc.y = fmaf(a.w[t], v.y, acc.y);
            acc.z = fmaf(a.w[t], v.z, acc.z);
            acc.w = fmaf(a.w[t], v.w, acc.w);
        }
        out[i] = acc;
    }
}

__global__ __launch_bounds__(256) void pvd_wsum1(pvd_wsum_args a, long long start, long long n,
                                                 float *__restrict__ out) {
    for (long long i = start + blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
         i += (long long)gridDim.x * blockDim.x) {
        float acc = a.w[0] * a.m[0][i];
        for (int t = 1; t < a.T; ++t) acc = fmaf(a.w[t], a.m[t][i], acc);
        out[i] = acc;
    }
}

hipError_t pvd_launch_wsum(const float *const *maps, const float *w, int T, size_t n, float *out, hipStream_t st) {
    pvd_wsum_args a;
    bool aligned = (reinterpret_cast<uintptr_t>(out) & 15) == 0;
    for (int t = 0; t < PVD_MAX_TIMEPOINTS; ++t) {
        a.m[t] = t < T ? maps[t] : nullptr;
        a.w[t] = t < T ? w[t] : 0.f;
        if (t < T) aligned = aligned && (reinterpret_cast<uintptr_t>(maps[t]) & 15) == 0;
    }
    a.T = T;
    long long done = 0;
    if (aligned && n >= 4) {
        const long long n4 = (long long)(n / 4);
        long long g = (n4 + 255) / 256;
        if (g > 8192) g = 8192;
        hipLaunchKernelGGL(pvd_wsum4, dim3((unsigned)g), dim3(256), 0, st, a, n4, reinterpret_cast<float4 *>(out));
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
        done = n4 * 4;
    }
    if ((size_t)done < n) {
        long long rem = (long long)n - done;
        long long g = (rem + 255) / 256;
        if (g > 8192) g = 8192;
        hipLaunchKernelGGL(pvd_wsum1, dim3((unsigned)g), dim3(256), 0, st, a, done, (long long)n, out);
        return hipGetLastError();
    }
    return hipSuccess;
}
