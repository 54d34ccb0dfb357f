// Voxelwise time-activity-curve fits: one thread per voxel, FP64
// Levenberg-Marquardt with analytic Jacobians (the reference runs
// scipy.optimize.curve_fit voxel by voxel, time_integration/curve_fitting.py).
// The normal equations are accumulated on the fly (no Jacobian storage), so a
// thread holds the T samples, the 5x5 system and the parameters in registers.
#include "pvd_internal.h"
#include "../../include/pvd_tac.h"

#include <cmath>

namespace {

constexpr double kLn2 = 0.69314718055994530942;

struct TacArgs {
    const float *m[PVD_TAC_MAX_T];
    pvd_tac_params p;
    int np;
    long long n;
    float *params;
    float *auc;
    unsigned char *status;
};

int n_params_of(int model) {
    switch (model) {
        case PVD_TAC_MONO: case PVD_TAC_NUKFIT_F2: case PVD_TAC_LMFIT_MONO: return 2;
        case PVD_TAC_NUKFIT_F3: case PVD_TAC_NUKFIT_F3A: case PVD_TAC_LMFIT_BI: return 3;
        case PVD_TAC_NUKFIT_F4A: return 4;
        case PVD_TAC_NUKFIT_F4B: case PVD_TAC_NUKFIT_F4C: case PVD_TAC_LMFIT_TRI: return 5;
        default: return -1;
    }
}

// f(t; p) and, when g != nullptr, df/dp (the expressions of curve_fitting.py:136-170)
__device__ __forceinline__ double model_eval(int model, const double *p, double t, double lp, double *g) {
    switch (model) {
        case PVD_TAC_MONO: {
            const double e = exp(-p[1] * t);
            if (g) { g[0] = e; g[1] = -t * p[0] * e; }
            return p[0] * e;
        }
        case PVD_TAC_NUKFIT_F2: {
            const double e = exp(-(p[1] + lp) * t);
            if (g) { g[0] = e; g[1] = -t * p[0] * e; }
            return p[0] * e;
        }
        case PVD_TAC_NUKFIT_F3: {
            const double e1 = exp(-(p[1] + lp) * t), e2 = exp(-(p[2] + lp) * t);
            if (g) { g[0] = e1 - e2; g[1] = -t * p[0] * e1; g[2] = t * p[0] * e2; }
            return p[0] * (e1 - e2);
        }
        case PVD_TAC_NUKFIT_F3A: {
            const double e1 = exp(-(p[2] + lp) * t), ep = exp(-lp * t);
            const double s = p[1] * e1 + (1.0 - p[1]) * ep;
            if (g) { g[0] = s; g[1] = p[0] * (e1 - ep); g[2] = -t * p[0] * p[1] * e1; }
            return p[0] * s;
        }
        case PVD_TAC_NUKFIT_F4A: {
            const double e1 = exp(-(p[1] + lp) * t), e2 = exp(-(p[3] + lp) * t);
            if (g) { g[0] = e1; g[1] = -t * p[0] * e1; g[2] = e2; g[3] = -t * p[2] * e2; }
            return p[0] * e1 + p[2] * e2;
        }
        case PVD_TAC_NUKFIT_F4B: {
            const double e1 = exp(-(p[1] + lp) * t), e2 = exp(-(p[3] + lp) * t), eb = exp(-(p[4] + lp) * t);
            if (g) {
                g[0] = e1 - eb; g[1] = -t * p[0] * e1; g[2] = eb - e2; g[3] = t * p[2] * e2;
                g[4] = t * (p[0] - p[2]) * eb;
            }
            return p[0] * e1 - p[2] * e2 - (p[0] - p[2]) * eb;
        }
        case PVD_TAC_NUKFIT_F4C: {
            const double e1 = exp(-(p[1] + lp) * t), e2 = exp(-(p[3] + lp) * t), eb = exp(-(p[4] + lp) * t);
            if (g) {
                g[0] = e1 - eb; g[1] = -t * p[0] * e1; g[2] = e2 - eb; g[3] = -t * p[2] * e2;
                g[4] = t * (p[0] + p[2]) * eb;
            }
            return p[0] * e1 + p[2] * e2 - (p[0] + p[2]) * eb;
        }
        case PVD_TAC_LMFIT_MONO: {
            const double e = exp(-kLn2 * t / p[1]);
            if (g) { g[0] = e; g[1] = p[0] * e * kLn2 * t / (p[1] * p[1]); }
            return p[0] * e;
        }
        case PVD_TAC_LMFIT_BI: {
            const double e1 = exp(-kLn2 * t / p[1]), e2 = exp(-kLn2 * t / p[2]);
            if (g) {
                g[0] = e1 - e2; g[1] = p[0] * e1 * kLn2 * t / (p[1] * p[1]);
                g[2] = -p[0] * e2 * kLn2 * t / (p[2] * p[2]);
            }
            return p[0] * (e1 - e2);
        }
        default: {   // PVD_TAC_LMFIT_TRI
            const double e1 = exp(-kLn2 * t / p[1]), e2 = exp(-kLn2 * t / p[3]), ep = exp(-kLn2 * t / p[4]);
            if (g) {
                g[0] = e1 - ep; g[1] = p[0] * e1 * kLn2 * t / (p[1] * p[1]); g[2] = ep - e2;
                g[3] = -p[2] * e2 * kLn2 * t / (p[3] * p[3]); g[4] = -(p[0] - p[2]) * ep * kLn2 * t / (p[4] * p[4]);
            }
            return p[0] * e1 - p[2] * e2 - (p[0] - p[2]) * ep;
        }
    }
}

// weighted cost sum_i (w_i (y_i - f_i))^2 and, if A != nullptr, A = J^T J, b = J^T e
template <int NP>
__device__ double accumulate(const TacArgs &a, const double *y, const double *p, double (*A)[NP], double *b) {
    const pvd_tac_params &q = a.p;
    double cost = 0.0;
    if (A) {
        for (int r = 0; r < NP; ++r) {
            b[r] = 0.0;
            for (int c = 0; c < NP; ++c) A[r][c] = 0.0;
        }
    }
    for (int i = 0; i < q.n_times; ++i) {
        double g[PVD_TAC_MAX_P];
        const double f = model_eval(q.model, p, q.t[i], q.lambda_phys, A ? g : nullptr);
        const double e = q.w[i] * (y[i] - f);
        cost += e * e;
        if (A) {
#pragma unroll
            for (int r = 0; r < NP; ++r) {
                const double jr = q.w[i] * g[r];
                b[r] += jr * e;
#pragma unroll
                for (int c = 0; c <= r; ++c) A[r][c] += jr * q.w[i] * g[c];
            }
        }
    }
    return cost;
}

// solve (A + mu diag(d2)) x = b by Cholesky; false if not positive definite
template <int NP>
__device__ bool damped_solve(const double (*A)[NP], const double *b, const double *d2, double mu, double *x) {
    double L[NP][NP];
    for (int r = 0; r < NP; ++r) {
        for (int c = 0; c <= r; ++c) {
            double s = A[r][c];
            if (r == c) s += mu * d2[r];
            for (int k = 0; k < c; ++k) s -= L[r][k] * L[c][k];
            if (r == c) {
                if (!(s > 0.0)) return false;
                L[r][r] = sqrt(s);
            } else {
                L[r][c] = s / L[c][c];
            }
        }
    }
    double z[NP];
    for (int r = 0; r < NP; ++r) {
        double s = b[r];
        for (int k = 0; k < r; ++k) s -= L[r][k] * z[k];
        z[r] = s / L[r][r];
    }
    for (int r = NP - 1; r >= 0; --r) {
        double s = z[r];
        for (int k = r + 1; k < NP; ++k) s -= L[k][r] * x[k];
        x[r] = s / L[r][r];
    }
    return true;
}

// Levenberg-Marquardt with MINPACK-style scaling: the damping term is
// mu * diag(D^2), D_j the largest Jacobian column norm seen so far, which keeps
// the step scale-invariant and stops the damping from collapsing in directions
// whose curvature vanishes along the path (degenerate valleys of the
// multi-exponential models).
template <int NP>
__device__ bool lm(const TacArgs &a, const double *y, double *p) {
    const double tol = a.p.tol;
    double A[NP][NP], b[NP], d2[NP];
    double cost = accumulate<NP>(a, y, p, A, b);
    if (!isfinite(cost)) return false;
    for (int k = 0; k < NP; ++k) d2[k] = fmax(A[k][k], 1e-300);
    double mu = 1e-3;
    for (int it = 0; it < a.p.max_iter; ++it) {
        if (cost == 0.0) return true;
        for (int k = 0; k < NP; ++k) d2[k] = fmax(d2[k], A[k][k]);
        bool stepped = false;
        for (int attempt = 0; attempt < 60; ++attempt) {
            double d[NP], pn[NP];
            if (damped_solve<NP>(A, b, d2, mu, d)) {
                for (int k = 0; k < NP; ++k) pn[k] = p[k] + d[k];
                const double cn = accumulate<NP>(a, y, pn, nullptr, nullptr);
                if (isfinite(cn) && cn <= cost) {
                    bool small = true;
                    for (int k = 0; k < NP; ++k) small = small && fabs(d[k]) <= tol * (fabs(pn[k]) + tol);
                    const bool flat = (cost - cn) <= tol * cost;
                    for (int k = 0; k < NP; ++k) p[k] = pn[k];
                    cost = accumulate<NP>(a, y, p, A, b);
                    if (!isfinite(cost)) return false;
                    if (small || flat) return true;
                    mu = fmax(mu * 0.3, 1e-15);
                    stepped = true;
                    break;
                }
            }
            mu *= 10.0;
            if (mu > 1e20) return true;      // no descent direction left: at a minimum
        }
        if (!stepped) return true;
    }
    return false;
}

template <int NP>
__device__ bool fit_np(const TacArgs &a, const double *y, double *p) {
    return lm<NP>(a, y, p);
}

__global__ __launch_bounds__(256) void tac_fit_kernel(TacArgs a) {
    const pvd_tac_params &q = a.p;
    for (long long v = blockIdx.x * 256LL + threadIdx.x; v < a.n; v += (long long)gridDim.x * 256) {
        double y[PVD_TAC_MAX_T];
        for (int i = 0; i < q.n_times; ++i) y[i] = (double)a.m[i][v];
        double p[PVD_TAC_MAX_P];
        for (int k = 0; k < PVD_TAC_MAX_P; ++k) p[k] = q.p0[k];
        if (q.model == PVD_TAC_MONO && q.p0_from_data) p[0] = y[0];
        if (q.force_zero_at_t0) {
            const double y0 = y[0];
            for (int i = 0; i < q.n_times; ++i) y[i] -= y0;
            y[0] = 0.0;
        }
        bool ok = a.np <= q.n_times;
        if (ok) {
            switch (a.np) {
                case 2: ok = fit_np<2>(a, y, p); break;
                case 3: ok = fit_np<3>(a, y, p); break;
                case 4: ok = fit_np<4>(a, y, p); break;
                default: ok = fit_np<5>(a, y, p); break;
            }
        }
        for (int k = 0; k < a.np && ok; ++k) ok = isfinite(p[k]);
        if (!ok)
            for (int k = 0; k < PVD_TAC_MAX_P; ++k) p[k] = q.fail_params[k];
        double s;
        if (q.auc_mode == PVD_TAC_AUC_MONO_TAIL) {
            s = p[0] / p[1] * (1.0 - exp(-p[1] * q.t_int));
        } else {
            s = 0.0;
            double f0 = model_eval(q.model, p, q.t[0], q.lambda_phys, nullptr);
            for (int i = 1; i < q.n_times; ++i) {
                const double f1 = model_eval(q.model, p, q.t[i], q.lambda_phys, nullptr);
                s += (q.t[i] - q.t[i - 1]) * (f1 + f0) * 0.5;
                f0 = f1;
            }
        }
        for (int k = 0; k < a.np; ++k) a.params[k * a.n + v] = (float)p[k];
        if (a.auc) a.auc[v] = (float)s;
        if (a.status) a.status[v] = ok ? 1 : 0;
    }
}

}  // namespace

extern "C" int pvd_tac_n_params(int model) {
    const int np = n_params_of(model);
    return np > 0 ? np : pvd_set_error(PVD_EINVAL, "unknown TAC model %d", model);
}

extern "C" int pvd_tac_fit(const pvd_tac_params *p, const float *const *maps, long long n, float *params, float *auc,
                           unsigned char *status, void *stream) {
    if (!p || !maps || !params || n < 1) return pvd_set_error(PVD_EINVAL, "null argument or n < 1");
    const int np = n_params_of(p->model);
    if (np < 0) return pvd_set_error(PVD_EINVAL, "unknown TAC model %d", p->model);
    if (p->n_times < 1 || p->n_times > PVD_TAC_MAX_T)
        return pvd_set_error(PVD_EINVAL, "n_times must be 1..%d", PVD_TAC_MAX_T);
    if (p->max_iter < 1 || !(p->tol > 0.0)) return pvd_set_error(PVD_EINVAL, "max_iter >= 1 and tol > 0 required");
    if (p->auc_mode != PVD_TAC_AUC_MONO_TAIL && p->auc_mode != PVD_TAC_AUC_TRAPEZOID)
        return pvd_set_error(PVD_EINVAL, "unknown auc_mode %d", p->auc_mode);
    if (p->auc_mode == PVD_TAC_AUC_MONO_TAIL && np != 2)
        return pvd_set_error(PVD_EINVAL, "the mono-exponential tail integral needs a 2-parameter model");
    TacArgs a{};
    for (int i = 0; i < p->n_times; ++i) {
        if (!maps[i]) return pvd_set_error(PVD_EINVAL, "null map %d", i);
        a.m[i] = maps[i];
    }
    a.p = *p;
    a.np = np;
    a.n = n;
    a.params = params;
    a.auc = auc;
    a.status = status;
    long long g = (n + 255) / 256;
    if (g > 16384) g = 16384;
    hipLaunchKernelGGL(tac_fit_kernel, dim3((unsigned)g), dim3(256), 0, (hipStream_t)stream, a);
    const hipError_t e = hipGetLastError();
    return e == hipSuccess ? PVD_OK : pvd_hip_error(e, "launch TAC fit");
}
