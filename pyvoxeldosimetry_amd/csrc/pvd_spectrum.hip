// FP64 device work done once per plan: dose-point-kernel generation and the
// kernel spectrum (separable direct DFT over the kernel's support).
#include "pvd_internal.h"

#define PVD_MAX_DPK_TERMS 16

struct pvd_dpk_args {
    double v;
    int g[3];
    int nterms;
    double factor;
    pvd_dpk_term t[PVD_MAX_DPK_TERMS];
};

// y90_kernel.py:33-55 / lu177_kernel.py:52-75: r about the centre g//2, sum of terms, x factor
__global__ __launch_bounds__(256) void pvd_dpk_kernel(pvd_dpk_args a, double *__restrict__ out) {
    const long long total = (long long)a.g[0] * a.g[1] * a.g[2];
    for (long long idx = blockIdx.x * (long long)blockDim.x + threadIdx.x; idx < total;
         idx += (long long)gridDim.x * blockDim.x) {
        const int k = (int)(idx % a.g[2]);
        const int j = (int)((idx / a.g[2]) % a.g[1]);
        const int i = (int)(idx / ((long long)a.g[2] * a.g[1]));
        const double x = (double)(i - a.g[0] / 2) * a.v;
        const double y = (double)(j - a.g[1] / 2) * a.v;
        const double z = (double)(k - a.g[2] / 2) * a.v;
        const double r = sqrt(x * x + y * y + z * z);
        double val = 0.0;
        for (int q = 0; q < a.nterms; ++q) {
            const pvd_dpk_term &t = a.t[q];
            if (t.kind == PVD_DPK_BETA) {
                if (r <= t.param) {
                    const double u = 1.0 - r / t.param;
                    val += u * u * exp(-2.0 * r / t.param) * t.amp;
                }
            } else {
                if (r > 0.0) val += t.amp * exp(-t.param * r / 10.0) / (4.0 * M_PI * r * r);
                else if (t.flags & PVD_DPK_NAN_CENTRE) val += __longlong_as_double(0x7ff8000000000000LL);
            }
        }
        out[idx] = val * a.factor;
    }
}

hipError_t pvd_launch_dpk(double v, const int g[3], const pvd_dpk_term *terms, int nterms, double factor,
                          double *out, hipStream_t st) {
    pvd_dpk_args a;
    a.v = v;
    a.g[0] = g[0]; a.g[1] = g[1]; a.g[2] = g[2];
    a.nterms = nterms;
    a.factor = factor;
    for (int i = 0; i < nterms; ++i) a.t[i] = terms[i];
    const long long total = (long long)g[0] * g[1] * g[2];
    const int block = 256;
    const long long want = (total + block - 1) / block;
    const int grid = (int)(want < 65536 ? want : 65536);
    hipLaunchKernelGGL(pvd_dpk_kernel, dim3(grid > 0 ? grid : 1), dim3(block), 0, st, a, out);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------------
// Kernel spectrum.  Per axis the kernel samples m in [lo, lo+u) sit at FFT
// position pos(m) = (m - s) mod P; the spectrum is
//   S[kx][ky][kz] = (1/(Px Py Pz)) * sum_m K[m] W_Px^{kx pos_x} W_Py^{ky pos_y} W_Pz^{kz pos_z}
// evaluated axis by axis in FP64 (z, then y, then x) and rounded once to FP32.
struct pvd_spec_axis {
    int P, lo, u, s;
};

struct pvd_spec_args {
    pvd_spec_axis a[3];
    int K[3];
    int H;
    int Hp;   // row pitch of the FP32 spectrum
    int sym;  // 1: write only the real part, octant kx <= Px/2, ky <= Py/2, as FP32 scalars
};

__device__ __forceinline__ int spec_pos(const pvd_spec_axis &ax, int m) {
    int p = (m - ax.s) % ax.P;
    return p < 0 ? p + ax.P : p;
}

// W_P^q = exp(-2 pi i q / P), q in [0, P), FP64 via sincospi (exact argument reduction)
__device__ __forceinline__ void fill_twiddles64(double2 *tw, int P) {
    for (int q = threadIdx.x; q < P; q += blockDim.x) {
        double sn, cs;
        sincospi(2.0 * (double)q / (double)P, &sn, &cs);
        tw[q] = make_double2(cs, -sn);
    }
    __syncthreads();
}

__device__ __forceinline__ double2 zmul(double2 a, double2 b) {
    return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

// (q + d) mod P for q, d in [0, P)
__device__ __forceinline__ int add_mod(int q, int d, int P) {
    q += d;
    return q >= P ? q - P : q;
}

// T1[a][b][kz], a < ux, b < uy, kz < H: the twiddle index kz * pos(m) mod P advances
// by kz (mod P) from one tap to the next (pos(m + 1) = pos(m) + 1 mod P)
__global__ __launch_bounds__(256) void pvd_spec_z(pvd_spec_args p, const double *__restrict__ K,
                                                  double2 *__restrict__ T1) {
    extern __shared__ double2 tw64[];
    const pvd_spec_axis &az = p.a[2];
    fill_twiddles64(tw64, az.P);
    const long long total = (long long)p.a[0].u * p.a[1].u * p.H;
    for (long long idx = blockIdx.x * (long long)blockDim.x + threadIdx.x; idx < total;
         idx += (long long)gridDim.x * blockDim.x) {
        const int kz = (int)(idx % p.H);
        const int b = (int)((idx / p.H) % p.a[1].u);
        const int a = (int)(idx / ((long long)p.H * p.a[1].u));
        const double *row = K + ((long long)(a + p.a[0].lo) * p.K[1] + (b + p.a[1].lo)) * p.K[2];
        double2 acc = make_double2(0.0, 0.0);
        const int dq = kz % az.P;
        int q = (int)(((long long)kz * spec_pos(az, az.lo)) % az.P);
        for (int m = 0; m < az.u; ++m) {
            const double kv = row[m + az.lo];
            const double2 w = tw64[q];
            acc.x += kv * w.x;
            acc.y += kv * w.y;
            q = add_mod(q, dq, az.P);
        }
        T1[idx] = acc;
    }
}

// T2[a][ky][kz], ky < Py
__global__ __launch_bounds__(256) void pvd_spec_y(pvd_spec_args p, const double2 *__restrict__ T1,
                                                  double2 *__restrict__ T2) {
    extern __shared__ double2 tw64[];
    const pvd_spec_axis &ay = p.a[1];
    fill_twiddles64(tw64, ay.P);
    const long long total = (long long)p.a[0].u * ay.P * p.H;
    for (long long idx = blockIdx.x * (long long)blockDim.x + threadIdx.x; idx < total;
         idx += (long long)gridDim.x * blockDim.x) {
        const int kz = (int)(idx % p.H);
        const int ky = (int)((idx / p.H) % ay.P);
        const int a = (int)(idx / ((long long)p.H * ay.P));
        double2 acc = make_double2(0.0, 0.0);
        int q = (int)(((long long)ky * spec_pos(ay, ay.lo)) % ay.P);
        for (int b = 0; b < ay.u; ++b) {
            const double2 t = T1[((long long)a * ay.u + b) * p.H + kz];
            const double2 r = zmul(t, tw64[q]);
            acc.x += r.x;
            acc.y += r.y;
            q = add_mod(q, ky, ay.P);
        }
        T2[idx] = acc;
    }
}

// S[kx][ky][kz] (FP32, rows of pitch Hp) = norm * sum_a T2[a][ky][kz] W_Px^{kx pos_x(a)}
__global__ __launch_bounds__(256) void pvd_spec_x(pvd_spec_args p, const double2 *__restrict__ T2,
                                                  double norm, float2 *__restrict__ S) {
    extern __shared__ double2 tw64[];
    const pvd_spec_axis &ax = p.a[0];
    fill_twiddles64(tw64, ax.P);
    const long long plane = (long long)p.a[1].P * p.H;
    const long long total = (long long)ax.P * plane;
    for (long long idx = blockIdx.x * (long long)blockDim.x + threadIdx.x; idx < total;
         idx += (long long)gridDim.x * blockDim.x) {
        const long long c = idx % plane;
        const int kx = (int)(idx / plane);
        double2 acc = make_double2(0.0, 0.0);
        int q = (int)(((long long)kx * spec_pos(ax, ax.lo)) % ax.P);
        for (int a = 0; a < ax.u; ++a) {
            const double2 r = zmul(T2[(long long)a * plane + c], tw64[q]);
            acc.x += r.x;
            acc.y += r.y;
            q = add_mod(q, kx, ax.P);
        }
        const long long ky = c / p.H, kz = c - ky * p.H;
        S[((long long)kx * p.a[1].P + ky) * p.Hp + kz] = make_float2((float)(acc.x * norm), (float)(acc.y * norm));
    }
}

// ---------------------------------------------------------------------------------
// Mirror-symmetric kernel anchored at its centre c = K//2 (s = c, lo = 0, u = K, K odd,
// K[c + j] == K[c - j] checked exactly on the host): every axis transform is the real
// cosine sum  sum_j w_j v[c + j] cos(2 pi k j / P),  w_0 = 1, w_j = 2 (j = 1..c),
// over the half taps j = 0..c, and only the octant kx <= Px/2, ky <= Py/2 is needed.
//   R1[a][b][kz]  a, b in 0..c   (kz < H)
//   R2[a][ky][kz] ky <= Py/2
//   S[kx][ky][kz] kx <= Px/2     (FP32 real octant, rows of pitch Hp)
__device__ __forceinline__ void fill_cos64(double *cs, int P) {
    for (int q = threadIdx.x; q < P; q += blockDim.x) cs[q] = cospi(2.0 * (double)q / (double)P);
    __syncthreads();
}

__global__ __launch_bounds__(256) void pvd_spec_z_sym(pvd_spec_args p, const double *__restrict__ K,
                                                      double *__restrict__ R1) {
    extern __shared__ double cs64[];
    const int P = p.a[2].P, cz = p.K[2] / 2, cy = p.K[1] / 2, cx = p.K[0] / 2;
    fill_cos64(cs64, P);
    const long long total = (long long)(cx + 1) * (cy + 1) * p.H;
    for (long long idx = blockIdx.x * (long long)blockDim.x + threadIdx.x; idx < total;
         idx += (long long)gridDim.x * blockDim.x) {
        const int kz = (int)(idx % p.H);
        const int b = (int)((idx / p.H) % (cy + 1));
        const int a = (int)(idx / ((long long)p.H * (cy + 1)));
        const double *row = K + ((long long)(cx + a) * p.K[1] + (cy + b)) * p.K[2] + cz;
        const int dq = kz % P;
        double acc = row[0];
        int q = dq;
        for (int j = 1; j <= cz; ++j) {
            acc += 2.0 * row[j] * cs64[q];
            q = add_mod(q, dq, P);
        }
        R1[idx] = acc;
    }
}

__global__ __launch_bounds__(256) void pvd_spec_y_sym(pvd_spec_args p, const double *__restrict__ R1,
                                                      double *__restrict__ R2) {
    extern __shared__ double cs64[];
    const int P = p.a[1].P, cy = p.K[1] / 2, cx = p.K[0] / 2, hy = P / 2 + 1;
    fill_cos64(cs64, P);
    const long long total = (long long)(cx + 1) * hy * p.H;
    for (long long idx = blockIdx.x * (long long)blockDim.x + threadIdx.x; idx < total;
         idx += (long long)gridDim.x * blockDim.x) {
        const int kz = (int)(idx % p.H);
        const int ky = (int)((idx / p.H) % hy);
        const int a = (int)(idx / ((long long)p.H * hy));
        const double *col = R1 + (long long)a * (cy + 1) * p.H + kz;
        double acc = col[0];
        int q = ky;
        for (int j = 1; j <= cy; ++j) {
            acc += 2.0 * col[(long long)j * p.H] * cs64[q];
            q = add_mod(q, ky, P);
        }
        R2[idx] = acc;
    }
}

__global__ __launch_bounds__(256) void pvd_spec_x_sym(pvd_spec_args p, const double *__restrict__ R2, double norm,
                                                      float *__restrict__ S) {
    extern __shared__ double cs64[];
    const int P = p.a[0].P, cx = p.K[0] / 2, hx = P / 2 + 1, hy = p.a[1].P / 2 + 1;
    fill_cos64(cs64, P);
    const long long plane = (long long)hy * p.H;
    const long long total = (long long)hx * plane;
    for (long long idx = blockIdx.x * (long long)blockDim.x + threadIdx.x; idx < total;
         idx += (long long)gridDim.x * blockDim.x) {
        const long long c = idx % plane;
        const int kx = (int)(idx / plane);
        double acc = R2[c];
        int q = kx;
        for (int j = 1; j <= cx; ++j) {
            acc += 2.0 * R2[(long long)j * plane + c] * cs64[q];
            q = add_mod(q, kx, P);
        }
        const long long ky = c / p.H, kz = c - ky * p.H;
        S[((long long)kx * hy + ky) * p.Hp + kz] = (float)(acc * norm);
    }
}

static int grid_for(long long total, int block) {
    long long g = (total + block - 1) / block;
    if (g > 16384) g = 16384;
    return (int)(g > 0 ? g : 1);
}

// Scratch (FP64, caller-provided): general kernels T1 = ux*uy*H and T2 = ux*Py*H
// double2; symmetric kernels (sym, centre-anchored) R1 = (cx+1)(cy+1)H and
// R2 = (cx+1)(Py/2+1)H doubles, in the same two buffers.  With sym, S receives the
// real octant [Px/2+1][Py/2+1][Hp] as FP32 scalars.
hipError_t pvd_launch_spectrum(const int P[3], const int lo[3], const int u[3], const int s[3], const int K[3],
                               int Hp, int sym, const double *kernel, double2 *T1, double2 *T2, float2 *S,
                               hipStream_t st) {
    pvd_spec_args p;
    for (int i = 0; i < 3; ++i) {
        p.a[i].P = P[i]; p.a[i].lo = lo[i]; p.a[i].u = u[i]; p.a[i].s = s[i];
        p.K[i] = K[i];
    }
    p.H = P[2] / 2 + 1;
    p.Hp = Hp;
    p.sym = sym;
    const int block = 256;
    const double norm = 1.0 / ((double)P[0] * (double)P[1] * (double)P[2]);
    hipError_t e;
    if (sym) {
        const long long cx = K[0] / 2, cy = K[1] / 2;
        double *R1 = reinterpret_cast<double *>(T1), *R2 = reinterpret_cast<double *>(T2);
        hipLaunchKernelGGL(pvd_spec_z_sym, dim3(grid_for((cx + 1) * (cy + 1) * p.H, block)), dim3(block),
                           sizeof(double) * P[2], st, p, kernel, R1);
        if ((e = hipGetLastError()) != hipSuccess) return e;
        hipLaunchKernelGGL(pvd_spec_y_sym, dim3(grid_for((cx + 1) * (P[1] / 2 + 1) * p.H, block)), dim3(block),
                           sizeof(double) * P[1], st, p, (const double *)R1, R2);
        if ((e = hipGetLastError()) != hipSuccess) return e;
        hipLaunchKernelGGL(pvd_spec_x_sym, dim3(grid_for((long long)(P[0] / 2 + 1) * (P[1] / 2 + 1) * p.H, block)),
                           dim3(block), sizeof(double) * P[0], st, p, (const double *)R2, norm,
                           reinterpret_cast<float *>(S));
        return hipGetLastError();
    }
    hipLaunchKernelGGL(pvd_spec_z, dim3(grid_for((long long)u[0] * u[1] * p.H, block)), dim3(block),
                       sizeof(double2) * P[2], st, p, kernel, T1);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    hipLaunchKernelGGL(pvd_spec_y, dim3(grid_for((long long)u[0] * P[1] * p.H, block)), dim3(block),
                       sizeof(double2) * P[1], st, p, T1, T2);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    hipLaunchKernelGGL(pvd_spec_x, dim3(grid_for((long long)P[0] * P[1] * p.H, block)), dim3(block),
                       sizeof(double2) * P[0], st, p, T2, norm, S);
    return hipGetLastError();
}
