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

// Smoothed value at voxel idx, computed by the whole wavefront: lane l takes the
// taps l, l+64, ... of the (2r+1)^3 window, then an FP64 butterfly reduction.
// Called with every lane of the wavefront active and the same idx.
__device__ double metal_fill_wave(const TissueArgs &a, long long idx, int lane) {
    const int z = (int)(idx % a.n2);
    const long long xy = idx / a.n2;
    const int y = (int)(xy % a.n1);
    const int x = (int)(xy / a.n1);
    const int r = a.radius, side = 2 * r + 1, taps = side * side * side;
    double part = 0.0;
    for (int tap = lane; tap < taps; tap += 64) {
        const int dz = tap % side - r, dy = (tap / side) % side - r, dx = tap / (side * side) - r;
        const long long j = ((long long)reflect(x + dx, a.n0) * a.n1 + reflect(y + dy, a.n1)) * a.n2 +
                            reflect(z + dz, a.n2);
        part += a.w[dx + r] * a.w[dy + r] * a.w[dz + r] * fill_source(a.ct[j], a.metal);
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) part += __shfl_xor(part, off, 64);
    return part;
}

__device__ __forceinline__ void classify(const TissueArgs &a, float v, unsigned char &bits, float &dens) {
    unsigned b = 0;
    for (int c = 0; c < a.ncls; ++c) b |= (unsigned)(v >= a.clo[c] && v <= a.chi[c]) << c;
    bits = (unsigned char)b;
    int m = 0;
    for (int i = 1; i < a.nmat; ++i) m = (v >= a.mlo[i]) ? i : m;
    dens = (v == v) ? a.mden[m] : v;
}

// Voxels are streamed 4 per lane; metal voxels found in a wavefront's chunk are
// then filled one after another by the whole wavefront (wave-uniform loop).
// The grid-stride loop advances whole wavefronts so every lane reaches the
// shuffles; lanes past the end only skip their loads and stores.
template <bool VEC>
__global__ __launch_bounds__(256) void tissue_map(TissueArgs a, long long nitems) {
    const int lane = threadIdx.x & 63;
    const long long wave0 = (blockIdx.x * (long long)blockDim.x + (threadIdx.x & ~63));
    const long long stride = (long long)gridDim.x * blockDim.x;
    constexpr int V = VEC ? 4 : 1;
    for (long long base = wave0; base < nitems; base += stride) {
        const long long i = base + lane;
        const bool valid = i < nitems;
        float h[V];
        if constexpr (VEC) {
            const float4 q = valid ? reinterpret_cast<const float4 *>(a.ct)[i] : make_float4(0.f, 0.f, 0.f, 0.f);
            h[0] = q.x;
            h[1] = q.y;
            h[2] = q.z;
            h[3] = q.w;
        } else {
            h[0] = valid ? a.ct[i] : 0.f;
        }
        unsigned flags = 0;
#pragma unroll
        for (int k = 0; k < V; ++k) flags |= (unsigned)(valid && a.handle && h[k] > a.metal) << k;
        unsigned long long pending = __ballot(flags != 0);
        while (pending) {
            const int src = __ffsll((long long)pending) - 1;
            const unsigned f = __shfl(flags, src, 64);
            const long long i_src = base + src;
#pragma unroll
            for (int k = 0; k < V; ++k) {
                if (f & (1u << k)) {
                    const double v = metal_fill_wave(a, i_src * V + k, lane);
                    if (lane == src) h[k] = (float)v;
                }
            }
            pending &= pending - 1;
        }
        if (!valid) continue;
        unsigned char b[V];
        float d[V];
#pragma unroll
        for (int k = 0; k < V; ++k) classify(a, h[k], b[k], d[k]);
        if constexpr (VEC) {
            if (a.corrected) reinterpret_cast<float4 *>(a.corrected)[i] = make_float4(h[0], h[1], h[2], h[3]);
            if (a.bits) reinterpret_cast<uchar4 *>(a.bits)[i] = make_uchar4(b[0], b[1], b[2], b[3]);
            if (a.density) reinterpret_cast<float4 *>(a.density)[i] = make_float4(d[0], d[1], d[2], d[3]);
        } else {
            if (a.corrected) a.corrected[i] = h[0];
            if (a.bits) a.bits[i] = b[0];
            if (a.density) a.density[i] = d[0];
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
    for (int c = 0; c < a.ncls; ++c) {
        a.clo[c] = p->class_lo[c];
        a.chi[c] = p->class_hi[c];
    }
    for (int i = 0; i < a.nmat; ++i) {
        a.mlo[i] = p->mat_lo[i];
        a.mden[i] = p->mat_density[i];
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
    const bool vec = (n % 4 == 0) && aligned(ct, 16) && (!corrected || aligned(corrected, 16)) &&
                     (!density || aligned(density, 16)) && (!class_bits || aligned(class_bits, 4));
    const long long items = vec ? n / 4 : n;
    const long long g = std::min<long long>((items + 255) / 256, 16384);
    if (vec) hipLaunchKernelGGL(tissue_map<true>, dim3((unsigned)g), dim3(256), 0, st, a, items);
    else hipLaunchKernelGGL(tissue_map<false>, dim3((unsigned)g), dim3(256), 0, st, a, items);
    if (hipError_t e = hipGetLastError(); e != hipSuccess) return pvd_hip_error(e, "launch tissue map");
    return PVD_OK;
}
