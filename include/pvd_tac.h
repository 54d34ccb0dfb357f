/* Voxelwise time-activity-curve fitting and integration on the GPU.
 *
 * Replaces the per-voxel scipy.optimize.curve_fit loops of
 * TimeCurveFitting.fit_time_activity_curve (time_integration/curve_fitting.py:19-84)
 * and fit_time_activity_curve_advanced (:112-220): one thread fits one voxel by
 * Levenberg-Marquardt in FP64 with analytic Jacobians, minimising
 * sum_i (w_i (y_i - f(t_i; p)))^2 (curve_fit's sigma = 1/w), then integrates.
 *
 * Models (t and the rates in the caller's time unit; lp = lambda_phys):
 *   PVD_TAC_MONO        A0 exp(-lam t)                               p = (A0, lam)
 *   PVD_TAC_NUKFIT_F2   A1 exp(-(l1+lp) t)                            (A1, l1)
 *   PVD_TAC_NUKFIT_F3   A1 (exp(-(l1+lp) t) - exp(-(l2+lp) t))        (A1, l1, l2)
 *   PVD_TAC_NUKFIT_F3A  A1 (a exp(-(l1+lp) t) + (1-a) exp(-lp t))     (A1, a, l1)
 *   PVD_TAC_NUKFIT_F4A  A1 e(l1) + A2 e(l2)                           (A1, l1, A2, l2)
 *   PVD_TAC_NUKFIT_F4B  A1 e(l1) - A2 e(l2) - (A1-A2) e(lbc)          (A1, l1, A2, l2, lbc)
 *   PVD_TAC_NUKFIT_F4C  A1 e(l1) + A2 e(l2) - (A1+A2) e(lbc)          (A1, l1, A2, l2, lbc)
 *   PVD_TAC_LMFIT_MONO  A1 exp(-ln2 t/T1)                             (A1, T1)
 *   PVD_TAC_LMFIT_BI    A1 (exp(-ln2 t/T1) - exp(-ln2 t/T2))          (A1, T1, T2)
 *   PVD_TAC_LMFIT_TRI   A1 g(T1) - A2 g(T2) - (A1-A2) g(Tp)           (A1, T1, A2, T2, Tp)
 *   with e(l) = exp(-(l+lp) t), g(T) = exp(-ln2 t/T).
 *
 * Start: p0 from the parameters, except PVD_TAC_MONO with p0_from_data, which
 * starts at (y[0], p0[1]) as the reference's basic fit does.  A fit fails when
 * it has more parameters than time points, does not converge within max_iter
 * iterations, or ends non-finite; a failed voxel gets fail_params.
 * Integral per voxel (auc):
 *   PVD_TAC_AUC_MONO_TAIL  A0/lam (1 - exp(-lam t_int))   (curve_fitting.py:74-84)
 *   PVD_TAC_AUC_TRAPEZOID  trapezoid of f(t_i; p) over the sample times (:214-215)
 *
 * Inputs: T device pointers to float32 maps of n voxels (host array of
 * pointers).  Outputs (device, float32): params[k*n + v] for k < n_params,
 * auc[v]; status[v] (uint8, nullable) = 1 on success.  Stream-ordered,
 * allocates nothing.  Returns PVD_OK or a negative PVD_E* code (include/pvd.h). */
#ifndef PVD_TAC_H
#define PVD_TAC_H

#include "pvd.h"

#ifdef __cplusplus
extern "C" {
#endif

#define PVD_TAC_MAX_T 16
#define PVD_TAC_MAX_P 5

enum {
    PVD_TAC_MONO = 0,
    PVD_TAC_NUKFIT_F2 = 1,
    PVD_TAC_NUKFIT_F3 = 2,
    PVD_TAC_NUKFIT_F3A = 3,
    PVD_TAC_NUKFIT_F4A = 4,
    PVD_TAC_NUKFIT_F4B = 5,
    PVD_TAC_NUKFIT_F4C = 6,
    PVD_TAC_LMFIT_MONO = 7,
    PVD_TAC_LMFIT_BI = 8,
    PVD_TAC_LMFIT_TRI = 9
};

#define PVD_TAC_AUC_MONO_TAIL 0
#define PVD_TAC_AUC_TRAPEZOID 1

typedef struct {
    int model;                         /* PVD_TAC_*                                        */
    int n_times;                       /* 1..PVD_TAC_MAX_T                                 */
    double t[PVD_TAC_MAX_T];           /* sample times                                     */
    double w[PVD_TAC_MAX_T];           /* residual weights (curve_fit sigma = 1/w)         */
    double lambda_phys;                /* lp of the NUKFIT models                          */
    double p0[PVD_TAC_MAX_P];          /* start point                                      */
    int p0_from_data;                  /* PVD_TAC_MONO: p0[0] = y[0]                       */
    int force_zero_at_t0;              /* y <- y - y[0] (so y[0] = 0) before the fit      */
    int max_iter;                      /* LM iterations (>= 1)                             */
    double tol;                        /* relative step / cost-decrease tolerance          */
    double fail_params[PVD_TAC_MAX_P]; /* written for failed voxels                        */
    int auc_mode;                      /* PVD_TAC_AUC_*                                    */
    double t_int;                      /* integration limit of PVD_TAC_AUC_MONO_TAIL       */
} pvd_tac_params;

/* Number of parameters of a model, or a negative PVD_E* code. */
int pvd_tac_n_params(int model);

int pvd_tac_fit(const pvd_tac_params *p, const float *const *maps, long long n, float *params, float *auc,
                unsigned char *status, void *stream);

#ifdef __cplusplus
}
#endif
#endif
