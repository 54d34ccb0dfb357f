/* pvd_learned.h -- C ABI of the learned heterogeneity correction in libpvd.so:
 * the project's Monte Carlo target generator and the BF16 MFMA kernels of the
 * 3-D residual CNN (forward, backward, optimizer).
 *
 * Conventions as in pvd.h: device pointers, a hipStream_t passed as void*,
 * int return codes (PVD_OK or PVD_E*), messages from pvd_last_error().
 *
 * None of these has a working counterpart in the reference: its GPU Monte
 * Carlo (pyvoxeldosimetry/core/monte_carlo.py:10-173) cannot be constructed
 * and deposits energy one step from the emission point, and it has no learned
 * correction.  The Monte Carlo replaces that placeholder as the source of
 * heterogeneous-medium dose.
 */
#ifndef PVD_LEARNED_H
#define PVD_LEARNED_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------------------------------------------------------- Monte Carlo */
#define PVD_MC_MAX_LINES 4
typedef struct pvd_mc_params {
    int n[3];                 /* grid (x, y, z), z contiguous                         */
    double voxel_mm;          /* isotropic voxel size                                 */
    int histories;            /* electrons per voxel with activity > 0               */
    unsigned long long seed;
    const float *e_of_u;      /* device: inverse CDF of the beta spectrum, n_u points */
    int n_u;
    const float *e_of_r;      /* device: E(r) on r in [0, r_max_mm], n_r points       */
    int n_r;
    double r_max_mm;          /* largest water-equivalent range in the table          */
    double katz_a, katz_b, katz_c;  /* R(E) = 10*a*E^(b - c ln E) mm in water (E MeV)  */
    double step_mm;           /* largest water-equivalent step                        */
    double scatter;           /* multiple-scattering strength                         */
    double rho_min;           /* densities below are clamped (g/cc)                   */
    const long long *active;  /* device: flat indices of the emitting voxels, or NULL   */
    long long n_active;       /*   (NULL: every voxel, those with activity 0 emit none) */
    /* Photons, as kerma: each packet leaves the decay point isotropically along a
     * straight ray and deposits E (1 - exp(-mu_en rho L)) of its remaining energy E
     * over each voxel chord L (exact voxel traversal), mu_en the energy-absorption
     * coefficient of water scaled by the local density rho. */
    int n_lines;                             /* photon lines per decay (<= PVD_MC_MAX_LINES)     */
    double line_energy[PVD_MC_MAX_LINES];    /* MeV                                              */
    double line_yield[PVD_MC_MAX_LINES];     /* photons per decay                                */
    double brems_z;           /* > 0: a bremsstrahlung packet per electron carrying Y(E0) E0, Y the
                                 thick-target yield 6e-4 Z E/(1 + 6e-4 Z E) of this Z, its mu_en taken
                                 at k = E0 (1 - sqrt(u)); the electron then deposits (1 - Y) of its
                                 losses.  0: none                                                */
    int photon_every;         /* photon packets for every k-th history of a voxel, weight x k (>= 1) */
    int electrons;            /* 1: transport the beta electrons; 0: photons only                */
    const float *mu_en;       /* device: water mu_en/rho (cm^2/g) at n_mu log-spaced energies    */
    int n_mu;
    double mu_log_e_lo, mu_log_e_hi;  /* ln of the first and last table energy (MeV)          */
} pvd_mc_params;

/* edep[v] += energy (MeV x activity weight) deposited in voxel v (FP64: a voxel
 * may receive ~1e7 deposits of ~1e-8 of its total); totals[0..2] += emitted,
 * deposited, escaped (FP64; photon energy leaving the grid or below 1e-6 of its
 * start counts as escaped).  density may be NULL (water). */
int pvd_mc_dose(const pvd_mc_params *p, const float *activity, const float *density, double *edep,
                double *totals, void *stream);

/* ---------------------------------------------------------------- 3-D CNN
 * Activations are channels-last BF16 [n0][n1][n2][C] with n2 % 32 == 0.
 * Weights of a 3x3x3 layer are BF16 [27][32][cin] (tap-major, cout, cin
 * contiguous), tap = (dx+1)*9 + (dy+1)*3 + (dz+1); convolution is the
 * cross-correlation y[v] = sum_tap,ci x[v + d(tap)][ci] w[tap][co][ci] with zero
 * padding outside the grid.  cout is always 32; cin is 16 or 32, or 4: a 16-channel
 * input whose channels 4..15 are zero (only 0..3 are read, and channels 0..3 of the
 * [27][32][16] weights), with the ReLU epilogue and no act (the input layer). */
#define PVD_CNN_EPI_RELU        0  /* a = relu(z + b): writes a (bf16) to y and to act  */
#define PVD_CNN_EPI_RESID_RELU  1  /* a = relu(z + b): act = a, y = resid + a (bf16)     */
/* (act may be NULL in both: inference keeps no activations for a backward pass;
   y and act must be 16-byte aligned: whole 32-voxel tiles are written in 16-byte chunks) */
#define PVD_CNN_EPI_LINEAR_F32  2  /* y32[v] = z[v][0] + b[0]  (FP32, channel 0 only)    */
#define PVD_CNN_EPI_ACCUM_F32   3  /* y32[v][c] += z[v][c]      (FP32, all 32 channels)   */

/* blocks / nblocks (device, both NULL = every block): compute only the listed blocks of
 * 2 x 4 x 32 output voxels (block index (bx * nby + by) * nbz + bz with nby = ceil(n1/4),
 * nbz = n2/32; *nblocks entries), e.g. from pvd_cnn_region_blocks. */
int pvd_cnn_conv_fwd(const void *x, int cin, const void *w, const float *bias, int epi, const void *resid,
                     void *y, void *act, float *y32, int n0, int n1, int n2, const int *blocks, const int *nblocks,
                     void *stream);

/* The blocks within the network's 6-voxel receptive field of the region dp >= frac *
 * (*dose_max): block_list receives their indices (in no fixed order) and *nblocks their
 * number, both on the device (no host synchronisation).  flags: int scratch of
 * pvd_cnn_region_block_count() entries, as is block_list.  A forward pass over these
 * blocks gives the network's output exactly on the region (elsewhere it is undefined). */
size_t pvd_cnn_region_block_count(int n0, int n1, int n2);
int pvd_cnn_region_blocks(const float *dp, int n0, int n1, int n2, const float *dose_max, float frac, int *flags,
                          int *block_list, int *nblocks, void *stream);

/* Weights of the transposed convolution used for the input gradient:
 * wt[tap][ci][co] = w[26 - tap][co][ci] (cin = 32). */
int pvd_cnn_flip_weights(const void *w, void *wt, void *stream);

/* dz = bf16(dh * (act > 0)) over n voxels x 32 channels; dbias_part[b][c]
 * receives per-block channel sums of dz (nblocks = ceil(n / 1024)). */
int pvd_cnn_relu_grad(const float *dh, const void *act, void *dz, float *dbias_part, long long n_voxels,
                      void *stream);

/* Weight gradient dw[tap][co][ci] = sum_v x[v + d(tap)][ci] dz[v][co] (FP32),
 * accumulated over voxels in a fixed order (deterministic).  work: FP32 scratch
 * of pvd_cnn_wgrad_work_floats() floats. */
size_t pvd_cnn_wgrad_work_floats(int n0, int n1, int n2);
int pvd_cnn_wgrad(const void *x, int cin, const void *dz, float *dw, float *work, int n0, int n1, int n2,
                  void *stream);

/* Loss of the multiplicative correction D = Dp (1 + r) against the target Dt on
 * voxels with weight m: loss = sum m (Dp r - (Dt - Dp))^2 / sum m.
 * Writes dr (bf16, channel 0 of a 32-channel tensor, others 0) and loss_part
 * (PVD_CNN_LOSS_BLOCKS rows of 3 floats: partial sums of m*e^2, of m and of dr). */
#define PVD_CNN_LOSS_BLOCKS 256
int pvd_cnn_loss(const float *r, const float *dp, const float *dt, const float *m, long long n_voxels,
                 float inv_count, void *dr, float *loss_part, void *stream);

/* Adam on FP32 master weights; also refreshes the BF16 copy. */
int pvd_cnn_adam(float *param, const float *grad, float *m, float *v, void *param_bf16, long long n, float lr,
                 float beta1, float beta2, float eps, int step, void *stream);

/* Network input: x[v] = bf16(activity/amax, density, dose_plain/dmax, 0 x 13),
 * amax = *act_max and dmax = *dose_max read on the device (floored at 1e-30),
 * so the caller needs no host synchronisation for the scales. */
int pvd_cnn_pack_inputs(const float *activity, const float *density, const float *dose_plain, const float *act_max,
                        const float *dose_max, void *x, long long n_voxels, void *stream);

/* The network input's scales in one pass: maxima[0] = max(activity), maxima[1] =
 * max(dose_plain), for pvd_cnn_pack_inputs' act_max = maxima and dose_max = maxima + 1
 * (NaN values are skipped, as fmaxf does; -inf if every value is NaN).
 * maxima holds PVD_CNN_MAXIMA_FLOATS floats; those past the first two are scratch. */
#define PVD_CNN_MAXIMA_FLOATS (2 + 2 * 1024)
int pvd_cnn_input_maxima(const float *activity, const float *dose_plain, long long n_voxels, float *maxima,
                         void *stream);

/* Corrected dose with fallback: out = dp*(1+r) where dp >= apply_frac * (*dose_max) and
 * |r| <= bound, else dp (r is read only where applied: 0 = everywhere).
 * counts[0] += voxels with |r| > bound and dp >= region_frac * (*dose_max);
 * counts[1] += voxels with dp >= region_frac * (*dose_max). */
int pvd_cnn_apply_correction(const float *dp, const float *r, float bound, const float *dose_max, float region_frac,
                             float apply_frac, float *out, unsigned long long *counts, long long n_voxels,
                             void *stream);

/* Whole-volume fallback on the device, after pvd_cnn_apply_correction on the same stream:
 * out = dp when counts[1] > 0 and counts[0] > max_exceed_frac * counts[1], else out is left
 * as it is.  The decision needs no host synchronisation. */
int pvd_cnn_apply_fallback(const float *dp, const unsigned long long *counts, double max_exceed_frac, float *out,
                           long long n_voxels, void *stream);

/* Sum rows of a [rows][cols] FP32 array in fixed order into out[cols] (+= if accumulate). */
int pvd_reduce_rows(const float *part, int rows, int cols, float *out, int accumulate, void *stream);

#ifdef __cplusplus
}
#endif
#endif /* PVD_LEARNED_H */
