// Register two-step y/x kernel instances for P = 16 * B (FY, X, IY; single-volume
// and batch), for the B list PVD_XY_FILE_BS that the including source file names.
// Compiled in three files so the instances build in parallel.
#pragma once
#include "pvd_fft_dev.h"

hipError_t PVD_XY_LAUNCH(int pass, int rb, const pvd_vol_d *v, const pvd_batch_args *a, int grid, int block,
                         size_t lds, hipStream_t st) {
#define PVD_XY_ONE(P, B)                                                                                  \
    if (a) hipLaunchKernelGGL((pvd_fft_pass_xy_batch<P, B>), dim3(grid), dim3(block), lds, st, *a);        \
    else hipLaunchKernelGGL((pvd_fft_pass_xy<P, B>), dim3(grid), dim3(block), lds, st, *v);
#define PVD_XY_CASE(B)                                                                                     \
    case B:                                                                                                \
        if (pass == PVD_PASS_FY) { PVD_XY_ONE(PVD_PASS_FY, B) }                                             \
        else if (pass == PVD_PASS_X) { PVD_XY_ONE(PVD_PASS_X, B) }                                          \
        else if (pass == PVD_PASS_IY) { PVD_XY_ONE(PVD_PASS_IY, B) }                                        \
        else return hipErrorInvalidValue;                                                                  \
        break;
    switch (rb) {
        PVD_XY_FILE_BS(PVD_XY_CASE)
        default: return hipErrorInvalidValue;
    }
#undef PVD_XY_CASE
#undef PVD_XY_ONE
    return hipGetLastError();
}

hipError_t PVD_XY_SET_LDS(size_t bytes) {
    hipError_t e = hipSuccess;
#define PVD_SETATTR(K) \
    if (e == hipSuccess) e = hipFuncSetAttribute((const void *)K, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes)
#define PVD_SETATTR_XY(B)                                                                              \
    PVD_SETATTR((pvd_fft_pass_xy<PVD_PASS_FY, B>)); PVD_SETATTR((pvd_fft_pass_xy<PVD_PASS_X, B>));         \
    PVD_SETATTR((pvd_fft_pass_xy<PVD_PASS_IY, B>)); PVD_SETATTR((pvd_fft_pass_xy_batch<PVD_PASS_FY, B>));  \
    PVD_SETATTR((pvd_fft_pass_xy_batch<PVD_PASS_X, B>)); PVD_SETATTR((pvd_fft_pass_xy_batch<PVD_PASS_IY, B>));
    PVD_XY_FILE_BS(PVD_SETATTR_XY)
#undef PVD_SETATTR_XY
#undef PVD_SETATTR
    return e;
}
