// LDS-tiled direct 3-D convolution for small kernels (extent <= 16 per axis).
//
// out[o] = sum_{m'} kr[m'] * A[b(o) + m'],  kr = kernel reversed on every axis,
// b(o) = o - (K-1) + sh; the activity is read periodically (reference mode,
// circular of size N) or as zero outside the grid (linear mode).  A workgroup
// computes an 8x8x32 output tile; each thread owns one (x, y) column and 8
// consecutive z outputs and slides a register window along z.
#include "pvd_internal.h"

namespace {

constexpr int TX = 8, TY = 8, TZ = 32, VZ = 8;
constexpr int BLOCK = TX * TY * (TZ / VZ);   // 256

struct pvd_direct_args {
    int N[3];
    int K[3];       // kernel extent used (after the reference crop)
    int sh[3];      // 0 (reference) or K//2 (linear)
    int wrap;       // 1 = periodic (reference), 0 = zero outside
    int T;
    const float *act[PVD_MAX_TIMEPOINTS];
    float w[PVD_MAX_TIMEPOINTS];
    const float *kr;  // reversed kernel, K[0]*K[1]*K[2] floats
    const float *dens;
    float *out;
    float scale;
};

__device__ __forceinline__ int src_index(int g, int n, bool wrap, bool &ok) {
    if (wrap) {
        int r = g % n;
        ok = true;
        return r < 0 ? r + n : r;
    }
    ok = (g >= 0 && g < n);
    return g;
}

template <int KZ>
__global__ __launch_bounds__(BLOCK) void pvd_direct_kernel(pvd_direct_args a) {
    extern __shared__ float sm[];
    const int KX = a.K[0], KY = a.K[1];
    const int SX = TX + KX - 1, SY = TY + KY - 1, SZ = TZ + KZ - 1;
    float *st = sm;
    float *kw = sm + SX * SY * SZ;
    const int ox0 = blockIdx.z * TX, oy0 = blockIdx.y * TY, oz0 = blockIdx.x * TZ;
    const int bx = ox0 - (KX - 1) + a.sh[0];
    const int by = oy0 - (KY - 1) + a.sh[1];
    const int bz = oz0 - (KZ - 1) + a.sh[2];
    const bool wrap = a.wrap != 0;
    const long long plane = (long long)a.N[1] * a.N[2];
    // halo tile: kU elements per thread per round, every load issued (at element 0 of
    // the map where the element is outside the grid or past the tile) before any is
    // used and masked afterwards (loads under a branch are waited for one by one)
    constexpr int kU = 8;
    const int total = SX * SY * SZ;
    for (int e0 = threadIdx.x; e0 < total; e0 += kU * BLOCK) {
        long long off[kU];
        bool ok[kU];
#pragma unroll
        for (int u = 0; u < kU; ++u) {
            const int e = e0 + u * BLOCK;
            const int i = e / (SY * SZ);
            const int r = e - i * SY * SZ;
            const int j = r / SZ;
            const int l = r - j * SZ;
            bool okx, oky, okz;
            const int gx = src_index(bx + i, a.N[0], wrap, okx);
            const int gy = src_index(by + j, a.N[1], wrap, oky);
            const int gz = src_index(bz + l, a.N[2], wrap, okz);
            ok[u] = e < total && okx && oky && okz;
            off[u] = ok[u] ? gx * plane + (long long)gy * a.N[2] + gz : 0;
        }
        float val[kU];
#pragma unroll
        for (int u = 0; u < kU; ++u) val[u] = a.act[0][off[u]];
#pragma unroll
        for (int u = 0; u < kU; ++u) val[u] = ok[u] ? a.w[0] * val[u] : 0.f;
        for (int t = 1; t < a.T; ++t) {
            float x[kU];
#pragma unroll
            for (int u = 0; u < kU; ++u) x[u] = a.act[t][off[u]];
#pragma unroll
            for (int u = 0; u < kU; ++u)
                if (ok[u]) val[u] = fmaf(a.w[t], x[u], val[u]);
        }
#pragma unroll
        for (int u = 0; u < kU; ++u)
            if (e0 + u * BLOCK < total) st[e0 + u * BLOCK] = val[u];
    }
    for (int e = threadIdx.x; e < KX * KY * KZ; e += BLOCK) kw[e] = a.kr[e];
    __syncthreads();

    const int iz = threadIdx.x % (TZ / VZ);
    const int iy = (threadIdx.x / (TZ / VZ)) % TY;
    const int ix = threadIdx.x / ((TZ / VZ) * TY);
    float acc[VZ];
#pragma unroll
    for (int i = 0; i < VZ; ++i) acc[i] = 0.f;
    for (int dx = 0; dx < KX; ++dx) {
        for (int dy = 0; dy < KY; ++dy) {
            const float *row = st + ((ix + dx) * SY + (iy + dy)) * SZ + iz * VZ;
            float win[VZ + KZ - 1];
#pragma unroll
            for (int q = 0; q < VZ + KZ - 1; ++q) win[q] = row[q];
            const float *kk = kw + (dx * KY + dy) * KZ;
#pragma unroll
            for (int dz = 0; dz < KZ; ++dz) {
                const float wv = kk[dz];
#pragma unroll
                for (int i = 0; i < VZ; ++i) acc[i] = fmaf(wv, win[i + dz], acc[i]);
            }
        }
    }
    const int ox = ox0 + ix, oy = oy0 + iy;
    if (ox >= a.N[0] || oy >= a.N[1]) return;
    const long long obase = ox * plane + (long long)oy * a.N[2];
    float f[VZ];
#pragma unroll
    for (int i = 0; i < VZ; ++i) f[i] = a.scale;
    if (a.dens) {           // all density loads issued (z clamped into the column) before use
        float rho[VZ];
#pragma unroll
        for (int i = 0; i < VZ; ++i) rho[i] = a.dens[obase + min(oz0 + iz * VZ + i, a.N[2] - 1)];
#pragma unroll
        for (int i = 0; i < VZ; ++i) f[i] *= density_factor(rho[i]);
    }
#pragma unroll
    for (int i = 0; i < VZ; ++i) {
        const int oz = oz0 + iz * VZ + i;
        if (oz < a.N[2]) a.out[obase + oz] = acc[i] * f[i];
    }
}

template <int KZ>
hipError_t launch_kz(const pvd_direct_args &a, hipStream_t st) {
    const int SX = TX + a.K[0] - 1, SY = TY + a.K[1] - 1, SZ = TZ + KZ - 1;
    const size_t lds = sizeof(float) * ((size_t)SX * SY * SZ + (size_t)a.K[0] * a.K[1] * KZ);
    dim3 grid((a.N[2] + TZ - 1) / TZ, (a.N[1] + TY - 1) / TY, (a.N[0] + TX - 1) / TX);
    hipLaunchKernelGGL(pvd_direct_kernel<KZ>, grid, dim3(BLOCK), lds, st, a);
    return hipGetLastError();
}

}  // namespace

// LDS bytes the direct kernel needs for kernel extents k[3] (<= 16 each)
size_t pvd_direct_lds_bytes(const int k[3]) {
    const size_t SX = TX + k[0] - 1, SY = TY + k[1] - 1, SZ = TZ + k[2] - 1;
    return sizeof(float) * (SX * SY * SZ + (size_t)k[0] * k[1] * k[2]);
}

hipError_t pvd_direct_set_lds_limit(size_t bytes) {
    hipError_t e = hipSuccess;
#define PVD_SETATTR(KZ) \
    if (e == hipSuccess) e = hipFuncSetAttribute((const void *)pvd_direct_kernel<KZ>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes)
    PVD_SETATTR(1); PVD_SETATTR(2); PVD_SETATTR(3); PVD_SETATTR(4); PVD_SETATTR(5); PVD_SETATTR(6);
    PVD_SETATTR(7); PVD_SETATTR(8); PVD_SETATTR(9); PVD_SETATTR(10); PVD_SETATTR(11); PVD_SETATTR(12);
    PVD_SETATTR(13); PVD_SETATTR(14); PVD_SETATTR(15); PVD_SETATTR(16);
#undef PVD_SETATTR
    return e;
}

hipError_t pvd_launch_direct(const int N[3], const int K[3], const int sh[3], int wrap, const float *const *acts,
                             const float *w, int T, const float *kr, const float *dens, float *out, float scale,
                             hipStream_t st) {
    pvd_direct_args a;
    for (int i = 0; i < 3; ++i) { a.N[i] = N[i]; a.K[i] = K[i]; a.sh[i] = sh[i]; }
    a.wrap = wrap;
    a.T = T;
    for (int t = 0; t < T; ++t) { a.act[t] = acts[t]; a.w[t] = w[t]; }
    for (int t = T; t < PVD_MAX_TIMEPOINTS; ++t) { a.act[t] = nullptr; a.w[t] = 0.f; }
    a.kr = kr;
    a.dens = dens;
    a.out = out;
    a.scale = scale;
    switch (K[2]) {
        case 1: return launch_kz<1>(a, st);
        case 2: return launch_kz<2>(a, st);
        case 3: return launch_kz<3>(a, st);
        case 4: return launch_kz<4>(a, st);
        case 5: return launch_kz<5>(a, st);
        case 6: return launch_kz<6>(a, st);
        case 7: return launch_kz<7>(a, st);
        case 8: return launch_kz<8>(a, st);
        case 9: return launch_kz<9>(a, st);
        case 10: return launch_kz<10>(a, st);
        case 11: return launch_kz<11>(a, st);
        case 12: return launch_kz<12>(a, st);
        case 13: return launch_kz<13>(a, st);
        case 14: return launch_kz<14>(a, st);
        case 15: return launch_kz<15>(a, st);
        case 16: return launch_kz<16>(a, st);
        default: return hipErrorInvalidValue;
    }
}
