// FFT dose-convolution passes (gfx950): the generic y/x instances and the host
// launchers that route each pass to its kernel instance (the z instances are in
// pvd_fft_z.hip, the y/x register two-step instances in pvd_fft_xy1/2/3.hip).
//
// The dose is computed as five descriptor-driven passes over an in-place
// complex workspace work[Nx][Py][Hp] (H = Pz/2+1 used columns, rows padded to
// Hp = 16*ceil(H/16) complex = whole 128-byte lines; x-planes wplane apart):
//   FZ  z: weighted sum of the T activity maps, zero-pad, two-for-one real FFT
//   FY  y: forward FFT of every (x, kz) column
//   X   x: forward FFT, times the kernel spectrum, inverse FFT, fold/roll map
//   IY  y: inverse FFT, fold/roll map
//   IZ  z: two-for-one inverse real FFT, fold/roll map, scale * rho_w/rho, store
// Transforms: 256-point y and x axes run a register-resident 16x16 four-step
// (one LDS exchange); z lengths P = A * B of the kZregA/kZregB table a register
// two-step; every other length LDS Stockham stages over lines staged k-major
// (element (line l, index k) at s[k*ld + l]) so global accesses stay coalesced
// along the contiguous axis.  The global loads of a tile are issued before any
// is consumed, and the barriers order LDS only (lds_barrier), so loads stay in
// flight across them.
#include "pvd_fft_dev.h"

namespace {
// threads per line pair of the z passes' I/O: the largest of 64, 32, 16 with
// tp * lines <= block (16 when the block is smaller; its threads then loop)
int z_threads_per_pair(int block, int lines) {
    return block >= 64 * lines ? 64 : block >= 32 * lines ? 32 : 16;
}

// source file holding the two-step instance for P = 16 * rb (0: none)
int xy_file(int rb) {
    switch (rb) {
#define PVD_XY_F1(B) case B: return 1;
#define PVD_XY_F2(B) case B: return 2;
#define PVD_XY_F3(B) case B: return 3;
        PVD_XY_BS_1(PVD_XY_F1)
        PVD_XY_BS_2(PVD_XY_F2)
        PVD_XY_BS_3(PVD_XY_F3)
#undef PVD_XY_F1
#undef PVD_XY_F2
#undef PVD_XY_F3
        default: return 0;
    }
}

hipError_t launch_xy(int pass, int rb, const pvd_vol_d *v, const pvd_batch_args *a, int grid, int block, size_t lds,
                     hipStream_t st) {
    switch (xy_file(rb)) {
        case 1: return pvd_fft_launch_xy1(pass, rb, v, a, grid, block, lds, st);
        case 2: return pvd_fft_launch_xy2(pass, rb, v, a, grid, block, lds, st);
        case 3: return pvd_fft_launch_xy3(pass, rb, v, a, grid, block, lds, st);
        default: return hipErrorInvalidValue;
    }
}

template <int PASS>
void launch_generic(const pvd_vol_d *v, const pvd_batch_args *a, int grid, int block, size_t lds, hipStream_t st) {
    if (a) hipLaunchKernelGGL((pvd_fft_pass_batch<PASS>), dim3(grid), dim3(block), lds, st, *a);
    else hipLaunchKernelGGL((pvd_fft_pass<PASS>), dim3(grid), dim3(block), lds, st, *v);
}

// One launch of `pass` for the volume `v` (a == nullptr) or for the members of the
// batch launch `a`, of which `v` is one (it selects the instance).
hipError_t launch_pass(int pass, const pvd_vol_d &v, const pvd_batch_args *a, int grid, int block, size_t lds,
                       hipStream_t st) {
    const pvd_vol_d *vp = a ? nullptr : &v;
    switch (pass) {
        case PVD_PASS_FZ:
        case PVD_PASS_IZ:
            return pvd_fft_launch_z(pass, v.ax[2].reg != 0, z_threads_per_pair(block, 1 << v.llz), vp, a, grid,
                                    block, lds, st);
        case PVD_PASS_FY:
        case PVD_PASS_IY:
            if (v.ax[1].reg == 2) return launch_xy(pass, v.ax[1].rb, vp, a, grid, block, lds, st);
            if (pass == PVD_PASS_FY) launch_generic<PVD_PASS_FY>(vp, a, grid, block, lds, st);
            else launch_generic<PVD_PASS_IY>(vp, a, grid, block, lds, st);
            return hipGetLastError();
        default:   // X: grid rounded up to a multiple of 8 (XCD bands; surplus blocks exit)
            grid = (grid + 7) & ~7;
            if (v.ax[0].reg == 2) return launch_xy(pass, v.ax[0].rb, vp, a, grid, block, lds, st);
            launch_generic<PVD_PASS_X>(vp, a, grid, block, lds, st);
            return hipGetLastError();
    }
}
}  // namespace

hipError_t pvd_launch_fft_pass(int pass, const pvd_vol_d &v, int grid, int block, size_t lds, hipStream_t st) {
    return launch_pass(pass, v, nullptr, grid, block, lds, st);
}

hipError_t pvd_launch_fft_pass_batch(int pass, const pvd_vol_d &v, const pvd_batch_args &a, int grid, int block,
                                     size_t lds, hipStream_t st) {
    return launch_pass(pass, v, &a, grid, block, lds, st);
}

hipError_t pvd_set_fft_lds_limits(size_t bytes) {
    hipError_t e = hipSuccess;
#define PVD_SETATTR(K) \
    if (e == hipSuccess) e = hipFuncSetAttribute((const void *)K, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes)
    PVD_SETATTR(pvd_fft_pass<PVD_PASS_FY>);
    PVD_SETATTR(pvd_fft_pass<PVD_PASS_X>);
    PVD_SETATTR(pvd_fft_pass<PVD_PASS_IY>);
    PVD_SETATTR(pvd_fft_pass_batch<PVD_PASS_FY>);
    PVD_SETATTR(pvd_fft_pass_batch<PVD_PASS_X>);
    PVD_SETATTR(pvd_fft_pass_batch<PVD_PASS_IY>);
#undef PVD_SETATTR
    if (e == hipSuccess) e = pvd_fft_z_set_lds(bytes);
    if (e == hipSuccess) e = pvd_fft_xy1_set_lds(bytes);
    if (e == hipSuccess) e = pvd_fft_xy2_set_lds(bytes);
    if (e == hipSuccess) e = pvd_fft_xy3_set_lds(bytes);
    return e;
}
