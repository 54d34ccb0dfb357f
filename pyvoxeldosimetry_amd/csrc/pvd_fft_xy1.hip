// Register two-step y/x kernel instances, lengths 16 * B for B in PVD_XY_BS_1.
#define PVD_XY_FILE_BS PVD_XY_BS_1
#define PVD_XY_LAUNCH pvd_fft_launch_xy1
#define PVD_XY_SET_LDS pvd_fft_xy1_set_lds
#include "pvd_fft_xy_inst.h"
