// FFT dose-convolution passes: the z-pass (FZ, IZ) kernel instances.  RZ selects
// the register-two-step-only form (the plan has a register z transform), TP the
// threads per line pair of the I/O phases; single-volume and batch instances of
// each combination.
#include "pvd_fft_dev.h"

namespace {
template <int PASS, bool RZ>
void launch_tp(int tp, const pvd_vol_d *v, const pvd_batch_args *a, int grid, int block, size_t lds,
               hipStream_t st) {
#define PVD_LZ(TPV)                                                                                      \
    if (a) hipLaunchKernelGGL((pvd_fft_pass_batch<PASS, RZ, TPV>), dim3(grid), dim3(block), lds, st, *a); \
    else hipLaunchKernelGGL((pvd_fft_pass<PASS, RZ, TPV>), dim3(grid), dim3(block), lds, st, *v);
    if (tp == 64) { PVD_LZ(64) }
    else if (tp == 32) { PVD_LZ(32) }
    else { PVD_LZ(16) }
#undef PVD_LZ
}
}  // namespace

hipError_t pvd_fft_launch_z(int pass, bool rz, int tp, const pvd_vol_d *v, const pvd_batch_args *a, int grid,
                            int block, size_t lds, hipStream_t st) {
    if (pass == PVD_PASS_FZ) {
        if (rz) launch_tp<PVD_PASS_FZ, true>(tp, v, a, grid, block, lds, st);
        else launch_tp<PVD_PASS_FZ, false>(tp, v, a, grid, block, lds, st);
    } else if (pass == PVD_PASS_IZ) {
        if (rz) launch_tp<PVD_PASS_IZ, true>(tp, v, a, grid, block, lds, st);
        else launch_tp<PVD_PASS_IZ, false>(tp, v, a, grid, block, lds, st);
    } else {
        return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t pvd_fft_z_set_lds(size_t bytes) {
    hipError_t e = hipSuccess;
#define PVD_SETATTR(K) \
    if (e == hipSuccess) e = hipFuncSetAttribute((const void *)K, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes)
#define PVD_SETATTR_Z(P, R)                                                                          \
    PVD_SETATTR((pvd_fft_pass<P, R, 16>)); PVD_SETATTR((pvd_fft_pass<P, R, 32>));                      \
    PVD_SETATTR((pvd_fft_pass<P, R, 64>)); PVD_SETATTR((pvd_fft_pass_batch<P, R, 16>));                \
    PVD_SETATTR((pvd_fft_pass_batch<P, R, 32>)); PVD_SETATTR((pvd_fft_pass_batch<P, R, 64>))
    PVD_SETATTR_Z(PVD_PASS_FZ, false);
    PVD_SETATTR_Z(PVD_PASS_FZ, true);
    PVD_SETATTR_Z(PVD_PASS_IZ, false);
    PVD_SETATTR_Z(PVD_PASS_IZ, true);
#undef PVD_SETATTR_Z
#undef PVD_SETATTR
    return e;
}
