// HU -> (metal fill) -> class bits and mass density, one pass over the CT.
//
// The reference (tissue/composition.py:72-94) marks metal (HU > 2000), sets it
// to NaN, smooths nan_to_num(image) with scipy.ndimage.gaussian_filter (sigma
// 1, 'reflect', truncate 4) and writes the smoothed values back into the metal
// voxels only.  The smoothing is needed at metal voxels alone, so this kernel
// evaluates the separable Gaussian there as a direct (2r+1)^3 weighted sum in
// FP64 (per-axis reflected indices: the same linear map as three 1-D passes)
// and streams every other voxel straight through.  The pass is HBM-bound:
// 4 B read and up to 9 B written per voxel.
#include "pvd_internal.h"
#include "../../include/pvd_tissue.h"

#include <cfloat>
#include <cmath>

namespace {

struct TissueArgs {
    const float *ct;
    float *corrected;
    unsigned char *bits;
    float *density;
    int n0, n1, n2;
    int radius;
    int handle;
    float metal;
    int ncls;
    int nmat;
    float clo[PVD_TISSUE_MAX_CLASSES];
    float chi[PVD_TISSUE_MAX_CLASSES];
    float mlo[PVD_TISSUE_MAX_MATERIALS];
    float mden[PVD_TISSUE_MAX_MATERIALS];
    double w[2 * PVD_TISSUE_MAX_RADIUS + 1];
};

// scipy.ndimage 'reflect' (half-sample symmetric: d c b a | a b c d | d c b a),
// folded as often as needed when the radius exceeds the extent.
__device__ __forceinline__ int reflect(int i, int n) {
    const int p = 2 * n;
    i %= p;
    if (i < 0) i += p;
    return i < n ? i : p - 1 - i;
}

// nan_to_num of the image whose metal voxels were set to NaN (float32 data).
__device__ __forceinline__ double fill_source(float c, float metal) {
    if (c > metal || c != c) return 0.0;
    if (c == -INFINITY) return -(double)FLT_MAX;
    return (double)c;
}

// Smoothed values of V consecutive voxels of one z-row (x, y, z0 .. z0+V-1),
// for the voxels flagged in `which`.  For each of the (2r+1)^2 (dx, dy) rows the
// lane issues all V+2r reflected z-loads of that row before using them, so a
// metal voxel costs (2r+1)^2 load round trips, shared by the V voxels.  R is
// the radius when known at compile time (the window then lives in registers),
// or -1 for a run-time radius.
template <int V, int R>
__device__ void metal_fill_row(const TissueArgs &a, int x, int y, int z0, unsigned which, float *h) {
    const int r = R >= 0 ? R : a.radius, side = 2 * r + 1;
    constexpr int W = V + 2 * (R >= 0 ? R : PVD_TISSUE_MAX_RADIUS);
    double acc[V];
#pragma unroll
    for (int k = 0; k < V; ++k) acc[k] = 0.0;
    for (int dx = -r; dx <= r; ++dx) {
        const long long px = (long long)reflect(x + dx, a.n0) * a.n1;
        for (int dy = -r; dy <= r; ++dy) {
            const float *row = a.ct + (px + reflect(y + dy, a.n1)) * a.n2;
            const double wxy = a.w[dx + r] * a.w[dy + r];
            float win[W];
#pragma unroll
            for (int j = 0; j < W; ++j)
                if (R >= 0 || j < V + 2 * r) win[j] = row[reflect(z0 - r + j, a.n2)];
#pragma unroll
            for (int k = 0; k < V; ++k) {
                if (!(which & (1u << k))) continue;
                double s = 0.0;
                if constexpr (R >= 0) {
#pragma unroll
                    for (int dz = 0; dz < 2 * R + 1; ++dz) s += a.w[dz] * fill_source(win[k + dz], a.metal);
                } else {
                    for (int dz = 0; dz < side; ++dz) s += a.w[dz] * fill_source(win[k + dz], a.metal);
                }
                acc[k] += wxy * s;
            }
        }
    }
#pragma unroll
    for (int k = 0; k < V; ++k)
        if (which & (1u << k)) h[k] = (float)acc[k];
}

// Class bits and density of one HU value.  The loops have compile-time trip
// counts (NM materials, all PVD_TISSUE_MAX_CLASSES classes; the host pads unused
// entries with bounds that never match) so the table stays in scalar registers.
template <int NM>
__device__ __forceinline__ void classify(const TissueArgs &a, float v, unsigned char &bits, float &dens) {
    unsigned b = 0;
#pragma unroll
    for (int c = 0; c < PVD_TISSUE_MAX_CLASSES; ++c) b |= (unsigned)(v >= a.clo[c] && v <= a.chi[c]) << c;
    bits = (unsigned char)b;
    float d = a.mden[0];
#pragma unroll
    for (int i = 1; i < NM; ++i) d = (v >= a.mlo[i]) ? a.mden[i] : d;
    dens = (v == v) ? d : v;
}

// Voxels are streamed V per lane (V = 4 with 16-byte aligned maps and n2 % 4 == 0,
// so a lane's voxels share one z-row); a lane holding metal fills them in place.
// Each lane issues the loads of kItems grid-stride items before using any, so a
// wavefront keeps kItems x 64 x 4V bytes in flight.
constexpr int kItems = 4;

template <int V, int R, int NM>
__device__ __forceinline__ void tissue_item(const TissueArgs &a, long long i, float *h) {
    unsigned which = 0;
#pragma unroll
    for (int k = 0; k < V; ++k) which |= (unsigned)(a.handle && h[k] > a.metal) << k;
    if (which) {
        const long long v0 = i * V;
        const long long xy = v0 / a.n2;
        metal_fill_row<V, R>(a, (int)(xy / a.n1), (int)(xy % a.n1), (int)(v0 % a.n2), which, h);
    }
    unsigned char b[V];
    float d[V];
#pragma unroll
    for (int k = 0; k < V; ++k) classify<NM>(a, h[k], b[k], d[k]);
    if constexpr (V == 4) {
        if (a.corrected) reinterpret_cast<float4 *>(a.corrected)[i] = make_float4(h[0], h[1], h[2], h[3]);
        if (a.bits) reinterpret_cast<uchar4 *>(a.bits)[i] = make_uchar4(b[0], b[1], b[2], b[3]);
        if (a.density) reinterpret_cast<float4 *>(a.density)[i] = make_float4(d[0], d[1], d[2], d[3]);
    } else {
        if (a.corrected) a.corrected[i] = h[0];
        if (a.bits) a.bits[i] = b[0];
        if (a.density) a.density[i] = d[0];
    }
}

template <bool VEC, int R, int NM>
__global__ __launch_bounds__(256) void tissue_map(TissueArgs a, long long nitems) {
    constexpr int V = VEC ? 4 : 1;
    const long long stride = (long long)gridDim.x * blockDim.x;
    for (long long i0 = blockIdx.x * (long long)blockDim.x + threadIdx.x; i0 < nitems; i0 += kItems * stride) {
        float h[kItems][V];
#pragma unroll
        for (int u = 0; u < kItems; ++u) {
            const long long i = i0 + u * stride;
            if (i >= nitems) continue;
            if constexpr (VEC) {
                const float4 q = reinterpret_cast<const float4 *>(a.ct)[i];
                h[u][0] = q.x;
                h[u][1] = q.y;
                h[u][2] = q.z;
                h[u][3] = q.w;
            } else {
                h[u][0] = a.ct[i];
            }
        }
#pragma unroll
        for (int u = 0; u < kItems; ++u) {
            const long long i = i0 + u * stride;
            if (i < nitems) tissue_item<V, R, NM>(a, i, h[u]);
        }
    }
}

bool aligned(const void *p, uintptr_t k) { return (reinterpret_cast<uintptr_t>(p) & (k - 1)) == 0; }

}  // namespace

extern "C" int pvd_tissue_map(const pvd_tissue_params *p, const float *ct, int n0, int n1, int n2, float *corrected,
                              unsigned char *class_bits, float *density, void *stream) {
    if (!p || !ct) return pvd_set_error(PVD_EINVAL, "null parameters or CT");
    if (n0 < 1 || n1 < 1 || n2 < 1) return pvd_set_error(PVD_EINVAL, "grid extents must be >= 1");
    if (p->n_classes < 0 || p->n_classes > PVD_TISSUE_MAX_CLASSES)
        return pvd_set_error(PVD_EINVAL, "n_classes must be 0..%d", PVD_TISSUE_MAX_CLASSES);
    if (p->n_materials < 1 || p->n_materials > PVD_TISSUE_MAX_MATERIALS)
        return pvd_set_error(PVD_EINVAL, "n_materials must be 1..%d", PVD_TISSUE_MAX_MATERIALS);
    for (int i = 1; i < p->n_materials; ++i)
        if (!(p->mat_lo[i] > p->mat_lo[i - 1])) return pvd_set_error(PVD_EINVAL, "mat_lo must be ascending");
    TissueArgs a{};
    a.ct = ct;
    a.corrected = corrected;
    a.bits = class_bits;
    a.density = density;
    a.n0 = n0;
    a.n1 = n1;
    a.n2 = n2;
    a.handle = p->handle_artifacts ? 1 : 0;
    a.metal = p->metal_hu;
    a.ncls = p->n_classes;
    a.nmat = p->n_materials;
    for (int c = 0; c < PVD_TISSUE_MAX_CLASSES; ++c) {          // unused classes never match
        a.clo[c] = c < a.ncls ? p->class_lo[c] : INFINITY;
        a.chi[c] = c < a.ncls ? p->class_hi[c] : -INFINITY;
    }
    for (int i = 0; i < PVD_TISSUE_MAX_MATERIALS; ++i) {        // unused materials are never reached
        a.mlo[i] = i < a.nmat ? p->mat_lo[i] : INFINITY;
        a.mden[i] = p->mat_density[i < a.nmat ? i : a.nmat - 1];   // +inf HU keeps the last material
    }
    if (a.handle) {
        if (!(p->sigma > 0.0) || !(p->truncate >= 0.0))
            return pvd_set_error(PVD_EINVAL, "sigma must be > 0 and truncate >= 0");
        const int r = (int)(p->truncate * p->sigma + 0.5);   // scipy: int(truncate * sd + 0.5)
        if (r > PVD_TISSUE_MAX_RADIUS)
            return pvd_set_error(PVD_ENOTSUP, "Gaussian radius %d exceeds %d", r, PVD_TISSUE_MAX_RADIUS);
        a.radius = r;
        double sum = 0.0;
        for (int k = -r; k <= r; ++k) sum += (a.w[k + r] = std::exp(-0.5 / (p->sigma * p->sigma) * k * k));
        for (int k = 0; k <= 2 * r; ++k) a.w[k] /= sum;
    }
    const long long n = (long long)n0 * n1 * n2;
    hipStream_t st = (hipStream_t)stream;
    const bool vec = (n2 % 4 == 0) && aligned(ct, 16) && (!corrected || aligned(corrected, 16)) &&
                     (!density || aligned(density, 16)) && (!class_bits || aligned(class_bits, 4));
    const long long items = vec ? n / 4 : n;
    const long long g = std::min<long long>((items + 255) / 256, 4096);
    const bool r4 = a.radius == 4;          // sigma 1, truncate 4: the reference's filter
    const dim3 grid((unsigned)g), block(256);
    if (a.nmat <= 16) {
        if (vec && r4) hipLaunchKernelGGL((tissue_map<true, 4, 16>), grid, block, 0, st, a, items);
        else if (vec) hipLaunchKernelGGL((tissue_map<true, -1, 16>), grid, block, 0, st, a, items);
        else if (r4) hipLaunchKernelGGL((tissue_map<false, 4, 16>), grid, block, 0, st, a, items);
        else hipLaunchKernelGGL((tissue_map<false, -1, 16>), grid, block, 0, st, a, items);
    } else {
        if (vec && r4) hipLaunchKernelGGL((tissue_map<true, 4, 32>), grid, block, 0, st, a, items);
        else if (vec) hipLaunchKernelGGL((tissue_map<true, -1, 32>), grid, block, 0, st, a, items);
        else if (r4) hipLaunchKernelGGL((tissue_map<false, 4, 32>), grid, block, 0, st, a, items);
        else hipLaunchKernelGGL((tissue_map<false, -1, 32>), grid, block, 0, st, a, items);
    }
    if (hipError_t e = hipGetLastError(); e != hipSuccess) return pvd_hip_error(e, "launch tissue map");
    return PVD_OK;
}
