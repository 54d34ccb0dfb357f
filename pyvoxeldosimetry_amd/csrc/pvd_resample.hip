// Grid-to-grid resampling (nearest / trilinear / cubic B-spline) for the activity
// <-> kernel resolution round trip of the dose path.  A workgroup writes one output
// row along z (coalesced stores); the input neighbours of consecutive outputs share
// cache lines, so the pass is bound by HBM (4 B written per output voxel, input read
// about once).  Coordinates are formed in FP64 so that equal spacings map every
// output voxel exactly onto an input voxel.  The cubic B-spline coefficients come
// from a separable mirror-boundary FIR prefilter (three passes, one per axis).
#include "pvd_internal.h"
#include "../../include/pvd_resample.h"

#include <algorithm>
#include <cmath>

namespace {

struct ResampleArgs {
    const float *in;
    float *out;
    int n[3], m[3];
    double ratio[3];   // out_spacing / in_spacing
    float scale;
    int interp;
};

// Index i (any integer) of a whole-sample symmetric (mirror) extension of [0, n).
__device__ __forceinline__ int mirror_index(int i, int n) {
    if (n == 1) return 0;
    const int p = 2 * (n - 1);
    int m = i % p;
    if (m < 0) m += p;
    return m < n ? m : p - m;
}

// Cubic B-spline weights of the 4 neighbours floor(c)-1 .. floor(c)+2 at fraction t.
__device__ __forceinline__ void bspline3_weights(float t, float w[4]) {
    const float t2 = t * t, t3 = t2 * t, u = 1.0f - t;
    w[0] = u * u * u * (1.0f / 6.0f);
    w[1] = (3.0f * t3 - 6.0f * t2 + 4.0f) * (1.0f / 6.0f);
    w[2] = (-3.0f * t3 + 3.0f * t2 + 3.0f * t + 1.0f) * (1.0f / 6.0f);
    w[3] = t3 * (1.0f / 6.0f);
}

// continuous index -> (lower neighbour, weight of the upper one); false if outside
__device__ __forceinline__ bool axis_pos(double c, int n, int &i0, float &f) {
    if (!(c >= -0.5 && c < n - 0.5)) return false;
    const double fl = floor(c);
    i0 = (int)fl;
    f = (float)(c - fl);
    return true;
}

// One workgroup per output row (i, j): the x/y neighbours and weights are
// uniform across the row, the threads stride along z.  Workgroups are handed
// to the 8 XCDs round-robin, so block b works on row chunk (b % 8) * G/8 + b / 8:
// each XCD walks a contiguous band of rows and neighbouring output rows find
// their shared input rows in that XCD's L2.
__global__ __launch_bounds__(256) void resample_kernel(ResampleArgs a) {
    const int rows = a.m[0] * a.m[1];
    const int G = gridDim.x;
    const int b = (G % 8 == 0) ? (int)(blockIdx.x % 8) * (G / 8) + (int)(blockIdx.x / 8) : (int)blockIdx.x;
    for (int row = b; row < rows; row += G) {
        const int i = row / a.m[1], j = row - i * a.m[1];
        float *orow = a.out + (long long)row * a.m[2];
        int b0, b1;
        float f0, f1;
        const bool in_xy = axis_pos(i * a.ratio[0], a.n[0], b0, f0) && axis_pos(j * a.ratio[1], a.n[1], b1, f1);
        if (!in_xy) {
            for (int k = threadIdx.x; k < a.m[2]; k += blockDim.x) orow[k] = 0.0f;
            continue;
        }
        if (a.interp == PVD_INTERP_CUBIC) {
            // 4 x 4 neighbour rows (mirrored indices) with their x/y weights, uniform over the row
            float wx[4], wy[4];
            bspline3_weights(f0, wx);
            bspline3_weights(f1, wy);
            const float *rowp[16];
            float wxy[16];
#pragma unroll
            for (int u = 0; u < 4; ++u)
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    const int x = mirror_index(b0 - 1 + u, a.n[0]), y = mirror_index(b1 - 1 + v, a.n[1]);
                    rowp[u * 4 + v] = a.in + ((long long)x * a.n[1] + y) * a.n[2];
                    wxy[u * 4 + v] = wx[u] * wy[v];
                }
            for (int k = threadIdx.x; k < a.m[2]; k += blockDim.x) {
                int b2;
                float f2;
                float v = 0.0f;
                if (axis_pos(k * a.ratio[2], a.n[2], b2, f2)) {
                    float wz[4];
                    bspline3_weights(f2, wz);
                    int z[4];
#pragma unroll
                    for (int q = 0; q < 4; ++q) z[q] = mirror_index(b2 - 1 + q, a.n[2]);
#pragma unroll
                    for (int r = 0; r < 16; ++r) {
                        const float *p = rowp[r];
                        const float s = fmaf(wz[0], p[z[0]], fmaf(wz[1], p[z[1]], fmaf(wz[2], p[z[2]], wz[3] * p[z[3]])));
                        v = fmaf(wxy[r], s, v);
                    }
                }
                orow[k] = a.scale * v;
            }
            continue;
        }
        if (a.interp == PVD_INTERP_NEAREST) {
            const int x = min(max(b0 + (f0 >= 0.5f), 0), a.n[0] - 1);
            const int y = min(max(b1 + (f1 >= 0.5f), 0), a.n[1] - 1);
            const float *p = a.in + ((long long)x * a.n[1] + y) * a.n[2];
            for (int k = threadIdx.x; k < a.m[2]; k += blockDim.x) {
                int b2;
                float f2;
                float v = 0.0f;
                if (axis_pos(k * a.ratio[2], a.n[2], b2, f2)) v = p[min(max(b2 + (f2 >= 0.5f), 0), a.n[2] - 1)];
                orow[k] = a.scale * v;
            }
            continue;
        }
        const int x0 = max(b0, 0), x1 = min(b0 + 1, a.n[0] - 1);
        const int y0 = max(b1, 0), y1 = min(b1 + 1, a.n[1] - 1);
        const float *p00 = a.in + ((long long)x0 * a.n[1] + y0) * a.n[2];
        const float *p01 = a.in + ((long long)x0 * a.n[1] + y1) * a.n[2];
        const float *p10 = a.in + ((long long)x1 * a.n[1] + y0) * a.n[2];
        const float *p11 = a.in + ((long long)x1 * a.n[1] + y1) * a.n[2];
        for (int k = threadIdx.x; k < a.m[2]; k += blockDim.x) {
            int b2;
            float f2;
            float v = 0.0f;
            if (axis_pos(k * a.ratio[2], a.n[2], b2, f2)) {
                const int z0 = max(b2, 0), z1 = min(b2 + 1, a.n[2] - 1);
                const float c00 = fmaf(f2, p00[z1] - p00[z0], p00[z0]);
                const float c01 = fmaf(f2, p01[z1] - p01[z0], p01[z0]);
                const float c10 = fmaf(f2, p10[z1] - p10[z0], p10[z0]);
                const float c11 = fmaf(f2, p11[z1] - p11[z0], p11[z0]);
                const float e0 = fmaf(f1, c01 - c00, c00);
                const float e1 = fmaf(f1, c11 - c10, c10);
                v = fmaf(f0, e1 - e0, e0);
            }
            orow[k] = a.scale * v;
        }
    }
}

// Cubic B-spline prefilter taps h[k] = sqrt(3) z^k, z = sqrt(3) - 2, k = 0..14 (FP64-built).
constexpr int kBsK = 14;
struct BsTaps { float h[kBsK + 1]; };
constexpr BsTaps bs_taps() {
    BsTaps t{};
    const double z = 1.7320508075688772935 - 2.0;
    double p = 1.7320508075688772935;
    for (int k = 0; k <= kBsK; ++k) { t.h[k] = (float)p; p *= z; }
    return t;
}

// One axis of the prefilter: dst[i] = sum_{|k| <= 14} h[|k|] src[mirror(i_axis + k)];
// a thread per element, consecutive threads along z (coalesced on every axis).
template <int AXIS>
__global__ __launch_bounds__(256) void bspline_axis_kernel(const float *__restrict__ src, float *__restrict__ dst,
                                                           int n0, int n1, int n2) {
    constexpr BsTaps T = bs_taps();
    const long long total = (long long)n0 * n1 * n2;
    const int n = AXIS == 0 ? n0 : AXIS == 1 ? n1 : n2;
    const long long stride = AXIS == 0 ? (long long)n1 * n2 : AXIS == 1 ? n2 : 1;
    for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < total;
         e += (long long)gridDim.x * blockDim.x) {
        const int z = (int)(e % n2);
        const long long xy = e / n2;
        const int y = (int)(xy % n1), x = (int)(xy / n1);
        const int i = AXIS == 0 ? x : AXIS == 1 ? y : z;
        const float *line = src + (e - (long long)i * stride);
        float acc = T.h[0] * line[(long long)i * stride];
#pragma unroll
        for (int k = 1; k <= kBsK; ++k)
            acc = fmaf(T.h[k], line[(long long)mirror_index(i - k, n) * stride] +
                                   line[(long long)mirror_index(i + k, n) * stride], acc);
        dst[e] = acc;
    }
}

}  // namespace

extern "C" int pvd_bspline_prefilter(const float *in, int n0, int n1, int n2, float *coef, float *tmp,
                                     void *stream) {
    if (!in || !coef || !tmp) return pvd_set_error(PVD_EINVAL, "null argument");
    if (n0 < 1 || n1 < 1 || n2 < 1) return pvd_set_error(PVD_EINVAL, "grid extents must be >= 1");
    if (coef == in || tmp == in || coef == tmp) return pvd_set_error(PVD_EINVAL, "coef, tmp and in must differ");
    const long long total = (long long)n0 * n1 * n2;
    const unsigned grid = (unsigned)std::min<long long>((total + 255) / 256, 65536);
    hipStream_t st = (hipStream_t)stream;
    hipLaunchKernelGGL(bspline_axis_kernel<0>, dim3(grid), dim3(256), 0, st, in, coef, n0, n1, n2);
    hipLaunchKernelGGL(bspline_axis_kernel<1>, dim3(grid), dim3(256), 0, st, (const float *)coef, tmp, n0, n1, n2);
    hipLaunchKernelGGL(bspline_axis_kernel<2>, dim3(grid), dim3(256), 0, st, (const float *)tmp, coef, n0, n1, n2);
    const hipError_t e = hipGetLastError();
    return e == hipSuccess ? PVD_OK : pvd_hip_error(e, "launch B-spline prefilter");
}

extern "C" int pvd_resample(const float *in, int n0, int n1, int n2, const double *in_spacing, float *out, int m0,
                            int m1, int m2, const double *out_spacing, int interp, double scale, void *stream) {
    if (!in || !out || !in_spacing || !out_spacing) return pvd_set_error(PVD_EINVAL, "null argument");
    if (n0 < 1 || n1 < 1 || n2 < 1 || m0 < 1 || m1 < 1 || m2 < 1)
        return pvd_set_error(PVD_EINVAL, "grid extents must be >= 1");
    if (interp != PVD_INTERP_NEAREST && interp != PVD_INTERP_LINEAR && interp != PVD_INTERP_CUBIC)
        return pvd_set_error(PVD_EINVAL, "interp must be PVD_INTERP_NEAREST, PVD_INTERP_LINEAR or PVD_INTERP_CUBIC");
    ResampleArgs a{};
    a.in = in;
    a.out = out;
    a.n[0] = n0; a.n[1] = n1; a.n[2] = n2;
    a.m[0] = m0; a.m[1] = m1; a.m[2] = m2;
    for (int d = 0; d < 3; ++d) {
        if (!(in_spacing[d] > 0.0) || !(out_spacing[d] > 0.0))
            return pvd_set_error(PVD_EINVAL, "spacings must be > 0");
        a.ratio[d] = out_spacing[d] / in_spacing[d];
    }
    a.scale = (float)scale;
    a.interp = interp;
    if ((long long)m0 * m1 > 0x7fffffffLL) return pvd_set_error(PVD_ENOTSUP, "too many output rows");
    long long g = (long long)m0 * m1;
    if (g > 65536) g = 65536;
    hipLaunchKernelGGL(resample_kernel, dim3((unsigned)g), dim3(256), 0, (hipStream_t)stream, a);
    const hipError_t e = hipGetLastError();
    return e == hipSuccess ? PVD_OK : pvd_hip_error(e, "launch resample");
}
