/* pvd.h -- C ABI of the MI355X kernel-convolution dose library (libpvd.so).
 *
 * Plain pointers and sizes only.  Every device pointer is HIP device memory
 * (for instance a torch tensor's data_ptr()); every `stream` is a
 * hipStream_t passed as void* (NULL = the null stream).  All calls are
 * stream-ordered and return immediately unless documented otherwise.
 *
 * Every export returns int: PVD_OK (0) or a negative PVD_E* code, with a
 * message available from pvd_last_error() (thread-local).  Nothing aborts
 * and no C++ exception crosses this boundary.
 *
 * Array layout: C order (nx, ny, nz), z contiguous -- the NumPy layout the
 * reference's activity maps use (pyvoxeldosimetry/core/kernel_convolution.py:48-76).
 */
#ifndef PVD_H
#define PVD_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PVD_OK        0
#define PVD_EINVAL   (-1)   /* bad argument (maps to ValueError)   */
#define PVD_ENOMEM   (-2)   /* device allocation failed (MemoryError) */
#define PVD_EHIP     (-3)   /* HIP runtime error (RuntimeError)     */
#define PVD_ENOTSUP  (-4)   /* valid but unsupported request        */

/* Convolution semantics.
 * PVD_MODE_REFERENCE: the reference's own operator,
 *   real(ifftn(fftn(A) * fftn(K, s=A.shape)))  (kernel_convolution.py:71-74):
 *   circular of size N, kernel anchored at its array origin, zero-padded at
 *   the high end or cropped to [0:N] on each axis.
 * PVD_MODE_LINEAR: centred ('same') linear convolution, zero outside the grid,
 *   kernel centre at index K//2 (the generators' centre, y90_kernel.py:34). */
#define PVD_MODE_REFERENCE 0
#define PVD_MODE_LINEAR    1

#define PVD_ALGO_AUTO   0   /* cost-model choice (the Python autotuner overrides) */
#define PVD_ALGO_FFT    1   /* LDS-staged Stockham FFT pipeline                   */
#define PVD_ALGO_DIRECT 2   /* LDS-tiled direct convolution (kernel extent <= 16) */

#define PVD_FLAG_PROFILE 1  /* record HIP events around every pass (pvd_plan_timings) */
#define PVD_FLAG_GENERAL_SPECTRUM 2  /* keep the full complex spectrum even for symmetric kernels */
#define PVD_FLAG_STOCKHAM 4  /* LDS Stockham stages on every axis (no register-resident transforms) */
#define PVD_FLAG_SMOOTH_LENGTHS 8  /* automatic FFT lengths from the 7-smooth rule only: no padding
                                      toward lengths with a register-resident transform */

/* Transform of one FFT axis (pvd_plan_info.fft_path). */
#define PVD_PATH_STOCKHAM 0  /* LDS Stockham radix stages                                  */
#define PVD_PATH_SQ256    1  /* y/x, length 256: register 16 x 16 four-step, one LDS exchange */
#define PVD_PATH_TWO_STEP 2  /* register two-step P = A * B (y/x: 16 * B; z: A in {8, 16})  */

#define PVD_MAX_TIMEPOINTS 16
#define PVD_MAX_FFT_LEN    2048

typedef struct pvd_plan pvd_plan;
typedef struct pvd_batch pvd_batch;

typedef struct pvd_plan_params {
    int n[3];       /* activity grid (x, y, z)                                  */
    int k[3];       /* dose-kernel grid (x, y, z)                               */
    int mode;       /* PVD_MODE_*                                               */
    int algo;       /* PVD_ALGO_*                                               */
    int fft_n[3];   /* FFT lengths per axis, 0 = plan's choice (7-smooth)       */
    int flags;      /* PVD_FLAG_*                                               */
    int tile[3];    /* lines per FFT tile: x pass, y pass (kz columns), z pass  */
                    /* (complex line pairs); power of two <= 64, 0 = 16 (x: 8   */
                    /* for the two-step lengths 288..512)                        */
    int max_radix;  /* largest power-of-two Stockham radix: 16 or 8, 0 = 16     */
} pvd_plan_params;

typedef struct pvd_plan_info {
    int algo;               /* resolved PVD_ALGO_FFT or PVD_ALGO_DIRECT          */
    int symmetric_spectrum; /* 1: real-octant kernel spectrum (after set_kernel) */
    int fft_n[3];           /* FFT lengths actually used                         */
    int kernel_used[3];     /* kernel extent after the reference's crop          */
    int n_passes;           /* kernel launches per dose call                     */
    size_t workspace_bytes; /* device bytes owned by the plan                    */
    int tile[3];            /* lines per tile used by the x, y and z passes      */
    int max_radix;          /* largest power-of-two radix used                   */
    int fft_path[3];        /* PVD_PATH_* of the x, y and z transforms (FFT)     */
    int block[5];           /* threads per workgroup of the FZ, FY, X, IY, IZ passes */
    int grid[5];            /* workgroups (tiles) of each pass                   */
} pvd_plan_info;

/* Dose-point-kernel terms (the analytic generators of
 * pyvoxeldosimetry/data/dose_kernels/y90_kernel.py:93-146 and
 * lu177_kernel.py:129-184).  With r the distance in mm from the grid centre
 * g//2 (y90_kernel.py:33-43):
 *   PVD_DPK_BETA:   amp * (1 - r/param)^2 * exp(-2 r/param)   for r <= param
 *   PVD_DPK_PHOTON: amp * exp(-param * r / 10) / (4 pi r^2)   for r > 0;
 *                   at r == 0 it is NaN when flags & PVD_DPK_NAN_CENTRE (the
 *                   reference's inf*0 in y90_kernel.py:134-138), else 0. */
#define PVD_DPK_BETA   0
#define PVD_DPK_PHOTON 1
#define PVD_DPK_NAN_CENTRE 1

typedef struct pvd_dpk_term {
    int kind;
    int flags;
    double amp;
    double param;   /* BETA: tissue range in mm; PHOTON: mu in 1/cm */
} pvd_dpk_term;

const char *pvd_last_error(void);
int pvd_version(void);

/* Replaces Y90KernelGenerator.generate_kernel (y90_kernel.py:20-55) and
 * Lu177KernelGenerator.generate_kernel (lu177_kernel.py:52-75): evaluates the
 * sum of `terms` times `factor` in FP64 on the device into out (gx*gy*gz doubles). */
int pvd_dpk_generate(double voxel_mm, int gx, int gy, int gz,
                     const pvd_dpk_term *terms, int nterms, double factor,
                     double *out, void *stream);

/* Plan for one grid/kernel shape; replaces the per-call setup inside
 * KernelConvolutionCalculator.calculate_dose_rate (kernel_convolution.py:48-76). */
int pvd_plan_create(pvd_plan **plan, const pvd_plan_params *params);
int pvd_plan_info_get(const pvd_plan *plan, pvd_plan_info *info);
/* Host only, no device call: the plan pvd_plan_create would build on a gfx950
 * device for these parameters (FFT lengths, algorithm, transform paths, tiles,
 * launch shapes, workspace bytes; symmetric_spectrum is 0 as no kernel is given).
 * No counterpart in the reference, which plans nothing (kernel_convolution.py:71-74). */
int pvd_plan_describe(const pvd_plan_params *params, pvd_plan_info *info);
int pvd_plan_destroy(pvd_plan *plan);

/* Upload the FP64 dose kernel (device pointer, k[0]*k[1]*k[2] doubles) and
 * build its spectrum in FP64 (rounded once to FP32, 1/(Px*Py*Pz) folded in).
 * Replaces np.fft.fftn(self.kernel, activity_map.shape) (kernel_convolution.py:73),
 * which the reference recomputes on every call. */
int pvd_plan_set_kernel(pvd_plan *plan, const double *kernel, void *stream);

/* dose = scale * (rho_w / rho(x)) * conv(act, kernel)
 * (rho term only when density != NULL; rho <= 0 gives 0).
 * Replaces KernelConvolutionCalculator.calculate_dose_rate (kernel_convolution.py:48-76). */
int pvd_conv_dose(pvd_plan *plan, const float *act, const float *density,
                  float *dose, double scale, void *stream);

/* dose = scale * (rho_w/rho) * conv(sum_t weight[t] * acts[t]).
 * acts: host array of T device pointers. */
int pvd_conv_dose_weighted(pvd_plan *plan, const float *const *acts,
                           const double *weights, int T, const float *density,
                           float *dose, double scale, void *stream);

/* Trapezoid time-integrated dose over T timepoints (hours); by linearity equal
 * to KernelConvolutionCalculator.calculate_absorbed_dose
 * (kernel_convolution.py:78-106: sum_i (D_i + D_{i+1})/2 * (t_{i+1}-t_i)*3600),
 * computed as one convolution of the trapezoid-weighted activity. */
int pvd_conv_dose_timeseries(pvd_plan *plan, const float *const *acts,
                             const double *t_hours, int T, const float *density,
                             float *dose, double scale, void *stream);

/* out = sum_t weights[t] * maps[t] over n floats (FP64 weights rounded once to
 * FP32).  With trapezoid weights (time in the caller's units) this replaces
 * ActivitySampler._trapezoid_integration (core/activity_sampler.py:69-79).
 * maps: host array of T device pointers. */
int pvd_weighted_sum(const float *const *maps, const double *weights, int T, size_t n,
                     float *out, void *stream);

/* Batches: many volumes of uneven sizes (each with its own FFT plan and kernel)
 * go through ONE launch per FFT pass; every launch's grid covers the tiles of
 * all volumes and each workgroup reads its volume's descriptor from a device
 * table.  pvd_batch_run_graph captures the five launches once into a hipGraph
 * and replays it; the descriptor table is uploaded before each replay, so new
 * pointers/weights take effect without re-capture.  No counterpart in the
 * reference, whose calculate_absorbed_dose (kernel_convolution.py:78-106) is
 * called once per volume. */
int pvd_batch_create(pvd_batch **batch, pvd_plan *const *plans, int nvol);
/* Set volume i's inputs: dose_i = scale * (rho_w/rho) * conv(sum_t w_t acts_t). */
int pvd_batch_set_volume(pvd_batch *batch, int i, const float *const *acts, const double *weights, int T,
                         const float *density, float *dose, double scale);
int pvd_batch_run(pvd_batch *batch, void *stream);
int pvd_batch_run_graph(pvd_batch *batch, void *stream);
int pvd_batch_destroy(pvd_batch *batch);

/* No counterpart in the reference.  For callers that cannot allocate device memory
 * (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc): the direct-path dose of pvd_plan_create(algo = PVD_ALGO_DIRECT) +
 * pvd_plan_set_kernel + pvd_conv_dose in one stream-ordered call, with no plan and
 * no allocation.  n: grid (x, y, z); k: kernel grid; kernel: k[0]*k[1]*k[2] FP64 on
 * the device; mode: PVD_MODE_*; dose = scale * (rho_w/rho) * conv(act, kernel) as
 * pvd_conv_dose.  The kernel extent used (after the reference-mode crop, or the
 * taps a linear-mode output reaches) must be <= 16 per axis: at most
 * PVD_NOALLOC_MAX_TAPS taps, held in a static device array of that many floats per
 * GPU.  Every call on a GPU shares that array: calls that overlap on one GPU race. */
#define PVD_NOALLOC_MAX_TAPS 4096
int pvd_conv_dose_direct_noalloc(const int n[3], const int k[3], int mode, const double *kernel,
                                 const float *act, const float *density, float *dose, double scale, void *stream);

/* Per-pass GPU milliseconds of the last dose call of a PVD_FLAG_PROFILE plan
 * (synchronises on it).  Writes min(n, n_passes) values; returns n_passes. */
int pvd_plan_timings(pvd_plan *plan, float *ms, int n);
/* Names of the passes, for reports. */
const char *pvd_pass_name(const pvd_plan *plan, int i);

#ifdef __cplusplus
}
#endif
#endif /* PVD_H */
