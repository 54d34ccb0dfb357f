/* Resampling between voxel grids on the GPU.
 *
 * Fills the gap the reference leaves at KernelConvolutionCalculator._resample_activity
 * (core/kernel_convolution.py:108-115, a stub returning None) and replaces the
 * SimpleITK resampler of ImageRegistration.resample_to_reference
 * (core/image_registration.py:128-173, linear / nearest interpolation).
 *
 * Geometry: origin 0 and identity direction for both grids; voxel i of array
 * axis a sits at i * spacing[a] (spacing[a] belongs to array axis a).  Output
 * voxel j samples the input at the continuous index c = j * out_spacing /
 * in_spacing per axis.  Points with c outside [-0.5, n - 0.5) on any axis get 0
 * (SimpleITK's default pixel value); inside, LINEAR interpolates trilinearly with
 * neighbour indices clamped to [0, n-1], NEAREST takes the nearest voxel
 * (ties round half up, then clamped), CUBIC evaluates the cubic B-spline whose
 * coefficients `in` holds (from pvd_bspline_prefilter) on the 4 x 4 x 4 neighbours,
 * indices mirrored at the edges -- SimpleITK's sitkBSpline interpolator, which
 * resample_to_reference selects for interpolation='cubic' (image_registration.py:
 * 155-160).  out = scale * interpolated value (FP32; coordinates in FP64).
 * Stream-ordered, allocates nothing.  Returns PVD_OK or a negative PVD_E* code
 * (include/pvd.h). */
#ifndef PVD_RESAMPLE_H
#define PVD_RESAMPLE_H

#include "pvd.h"

#ifdef __cplusplus
extern "C" {
#endif

#define PVD_INTERP_NEAREST 0
#define PVD_INTERP_LINEAR 1
#define PVD_INTERP_CUBIC 2

int pvd_resample(const float *in, int n0, int n1, int n2, const double *in_spacing, float *out, int m0, int m1,
                 int m2, const double *out_spacing, int interp, double scale, void *stream);

/* Cubic B-spline interpolation coefficients of an n0 x n1 x n2 image with mirror
 * (whole-sample symmetric) boundaries: the c with f[i] = (c[i-1] + 4 c[i] + c[i+1])/6
 * along every axis (Unser's exact interpolating spline; the prefilter of ITK's
 * BSplineInterpolateImageFunction, which sitkBSpline uses).  Computed per axis as
 * the symmetric FIR c[i] = sum_k h[k] f[mirror(i + k)], h[k] = sqrt(3) z^|k| with
 * z = sqrt(3) - 2, |k| <= 14 (|z|^15 < 3e-9).  coef and tmp: n0*n1*n2 floats each,
 * neither aliasing `in`.  No reference counterpart beyond the SimpleITK call. */
int pvd_bspline_prefilter(const float *in, int n0, int n1, int n2, float *coef, float *tmp, void *stream);

#ifdef __cplusplus
}
#endif
#endif
