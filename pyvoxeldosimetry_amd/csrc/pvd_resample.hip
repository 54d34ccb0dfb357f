// Grid-to-grid resampling (nearest / trilinear) for the activity <-> kernel
// resolution round trip of the dose path.  A workgroup writes one output row
// along z (coalesced stores); the eight input neighbours of consecutive outputs
// share cache lines, so the pass is bound by HBM (4 B written per output voxel,
// input read about once).  Coordinates are formed in FP64 so that equal spacings map
// every output voxel exactly onto an input voxel.
#include "pvd_internal.h"
#include "../../include/pvd_resample.h"

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

}  // namespace

extern "C" int pvd_resample(const float *in, int n0, int n1, int n2, const double *in_spacing, float *out, int m0,
                            int m1, int m2, const double *out_spacing, int interp, double scale, void *stream) {
    if (!in || !out || !in_spacing || !out_spacing) return pvd_set_error(PVD_EINVAL, "null argument");
    if (n0 < 1 || n1 < 1 || n2 < 1 || m0 < 1 || m1 < 1 || m2 < 1)
        return pvd_set_error(PVD_EINVAL, "grid extents must be >= 1");
    if (interp != PVD_INTERP_NEAREST && interp != PVD_INTERP_LINEAR)
        return pvd_set_error(PVD_EINVAL, "interp must be PVD_INTERP_NEAREST or PVD_INTERP_LINEAR");
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
