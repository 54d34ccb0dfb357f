// One-shot direct dose convolution that allocates nothing (pvd_conv_dose_direct_noalloc).
//
// The direct path's only buffer beyond the caller's inputs and outputs is the
// reversed, cropped FP32 kernel that a direct plan keeps (pvd_plan_set_kernel).
// Here it lives in `scratch`, a file-scope device array of PVD_NOALLOC_MAX_TAPS
// floats that every GPU holds for the life of the process; a small kernel fills
// it from the caller's FP64 kernel on the call's stream, then the direct kernel
// runs exactly as for a plan.  No hipMalloc / hipFree, no host synchronisation.
// Calls that overlap on one GPU share its array and race.
#include <math.h>

#include <mutex>

#include "pvd_internal.h"

hipError_t pvd_launch_direct(const int N[3], const int K[3], const int sh[3], int wrap, const float *const *acts,
                             const float *w, int T, const float *kr, const float *dens, float *out, float scale,
                             hipStream_t st);
size_t pvd_direct_lds_bytes(const int k[3]);
hipError_t pvd_direct_set_lds_limit(size_t bytes);

static __device__ __align__(256) float scratch[PVD_NOALLOC_MAX_TAPS];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_mu;
float *g_scratch[kMaxDevices] = {};   // device address of `scratch` on each GPU
size_t g_max_lds[kMaxDevices] = {};   // LDS per workgroup of each GPU

// Set-up on the first call on a device (with it current): the device address of
// its `scratch`, its LDS size, and the direct kernels' dynamic-LDS limit.
int setup_device(int dev, float **kr, size_t *max_lds) {
    if (dev < 0 || dev >= kMaxDevices) return pvd_set_error(PVD_ENOTSUP, "device %d beyond %d", dev, kMaxDevices);
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_scratch[dev]) {
        int v = 0;
        hipError_t e = hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev);
        if (e != hipSuccess) return pvd_hip_error(e, "query LDS size");
        e = pvd_direct_set_lds_limit((size_t)v);
        if (e != hipSuccess) return pvd_hip_error(e, "set direct LDS limit");
        void *ptr = nullptr;
        e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return pvd_hip_error(e, "look up the scratch array");
        g_max_lds[dev] = (size_t)v;
        g_scratch[dev] = static_cast<float *>(ptr);
    }
    *kr = g_scratch[dev];
    *max_lds = g_max_lds[dev];
    return PVD_OK;
}

// kr[i][j][l] = (float) k[lo0 + Ku0-1-i][lo1 + Ku1-1-j][lo2 + Ku2-1-l]: the kernel
// reversed on every axis and cropped to its used extent, as a direct plan stores it.
__global__ void reverse_crop_kernel(const double *__restrict__ k, int K1, int K2, int lo0, int lo1, int lo2,
                                    int Ku0, int Ku1, int Ku2, float *__restrict__ kr) {
    const int n = Ku0 * Ku1 * Ku2;
    for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < n; e += gridDim.x * blockDim.x) {
        const int l = e % Ku2, j = (e / Ku2) % Ku1, i = e / (Ku1 * Ku2);
        const long long src = ((long long)(lo0 + Ku0 - 1 - i) * K1 + (lo1 + Ku1 - 1 - j)) * K2 + (lo2 + Ku2 - 1 - l);
        kr[e] = (float)k[src];
    }
}

}  // namespace

extern "C" int pvd_conv_dose_direct_noalloc(const int n[3], const int k[3], int mode, const double *kernel,
                                            const float *act, const float *density, float *dose, double scale,
                                            void *stream) {
    if (!n || !k || !kernel || !act || !dose) return pvd_set_error(PVD_EINVAL, "null argument");
    if (mode != PVD_MODE_REFERENCE && mode != PVD_MODE_LINEAR) return pvd_set_error(PVD_EINVAL, "bad mode %d", mode);
    if (!isfinite(scale)) return pvd_set_error(PVD_EINVAL, "non-finite scale");
    int lo[3], Ku[3], sh[3];
    for (int a = 0; a < 3; ++a) {
        if (n[a] < 1 || k[a] < 1) return pvd_set_error(PVD_EINVAL, "grid %d and kernel %d on axis %d", n[a], k[a], a);
        if (mode == PVD_MODE_REFERENCE) {          // np.fft.fftn(kernel, s=N) crops to [0:N]
            lo[a] = 0;
            Ku[a] = k[a] < n[a] ? k[a] : n[a];
            sh[a] = 0;
        } else {                                   // taps within N-1 of the centre K//2
            const int c = k[a] / 2;
            const int mlo = c - n[a] + 1 > 0 ? c - n[a] + 1 : 0, mhi = c + n[a] - 1 < k[a] - 1 ? c + n[a] - 1 : k[a] - 1;
            lo[a] = mlo;
            Ku[a] = mhi - mlo + 1;
            sh[a] = c - mlo;
        }
        if (Ku[a] > 16)
            return pvd_set_error(PVD_ENOTSUP, "kernel extent %d used on axis %d exceeds 16 (direct path)", Ku[a], a);
    }
    if ((size_t)Ku[0] * Ku[1] * Ku[2] > PVD_NOALLOC_MAX_TAPS)
        return pvd_set_error(PVD_ENOTSUP, "kernel of %d taps exceeds the scratch array (%d)", Ku[0] * Ku[1] * Ku[2],
                             PVD_NOALLOC_MAX_TAPS);
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return pvd_hip_error(e, "current device");
    size_t max_lds = 0;
    float *kr = nullptr;
    if (int rc = setup_device(dev, &kr, &max_lds)) return rc;
    if (pvd_direct_lds_bytes(Ku) > max_lds) return pvd_set_error(PVD_ENOTSUP, "direct tile exceeds the LDS");
    hipStream_t st = (hipStream_t)stream;
    hipLaunchKernelGGL(reverse_crop_kernel, dim3(16), dim3(256), 0, st, kernel, k[1], k[2], lo[0], lo[1], lo[2],
                       Ku[0], Ku[1], Ku[2], kr);
    e = hipGetLastError();
    if (e != hipSuccess) return pvd_hip_error(e, "launch kernel reversal");
    const float *acts[1] = {act};
    const float w[1] = {1.0f};
    e = pvd_launch_direct(n, Ku, sh, mode == PVD_MODE_REFERENCE, acts, w, 1, kr, density, dose, (float)scale, st);
    if (e != hipSuccess) return pvd_hip_error(e, "launch direct convolution");
    return PVD_OK;
}
