// Internal descriptors and device helpers shared by the HIP sources of libpvd.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/pvd.h"

#define PVD_MAXSTAGE 16

// Tiles: complex line pairs per tile of the z passes (each pair is two real
// lines), kz columns per tile of the y and x passes.  Powers of two, at most
// PVD_TILE_MAX (plan parameter; 16 by default).
#define PVD_TILE_DEFAULT 16
#define PVD_TILE_MAX 64

// The register-resident 256-point column FFT of the y and x passes exchanges
// through a [16][16*16 + PVD_SQ256_PAD] float2 LDS array, PVD_SQ256_PAD x 16 float2
// more than a plain 16-column tile of 256 rows.
#define PVD_SQ256_PAD 16

// Padding (complex elements) between x-planes of the FFT workspace.
#ifndef PVD_PLANE_PAD
#define PVD_PLANE_PAD 0
#endif
#define PVD_SQ256_EXTRA_LDS (16 * PVD_SQ256_PAD * 8)

// One FFT axis of one volume.
//   n     FFT length P (7-smooth)
//   N     logical length (activity grid extent)
//   off   output o reads FFT index (o + off) mod n
//   foldn outputs o < foldn also add FFT index (o + N + off) mod n
//         (reference-mode circular wrap when the FFT is padded, P != N)
//   reg   register-resident transform: on y/x, 1 = the 256-point 16x16 path may
//         run, 2 = the two-step n = 16 * rb (its own kernel instances); on z, index
//         into kZregA/kZregB (0 = LDS Stockham stages)
//   ra, ra_log, rb  the split n = ra * rb (ra a power of two), when it is a two-step
struct pvd_axis_d {
    int n;
    int N;
    int nst;
    int off;
    int foldn;
    int reg;
    int ra, ra_log, rb, pad_;
    const float2 *tw;                 // n forward twiddles W_n^m, FP64-built, FP32-stored
    unsigned char radix[PVD_MAXSTAGE];
};

enum { PVD_PASS_FZ = 0, PVD_PASS_FY = 1, PVD_PASS_X = 2, PVD_PASS_IY = 3, PVD_PASS_IZ = 4,
       PVD_NPASS_FFT = 5 };

// Everything one FFT dose call needs (passed by value as a kernel argument for a
// single volume, or read from a device table for a batch).
struct pvd_vol_d {
    pvd_axis_d ax[3];                 // x, y, z
    int H;                            // Pz/2 + 1 (half spectrum along z)
    int Hp;                           // row pitch of work/spec: H rounded up to 16 (128-byte rows)
    int T;                            // timepoints folded into the load
    int llz;                          // log2 of complex line pairs per z-pass tile
    int ltky;                         // log2 of columns per y-pass tile
    int ltkx;                         // log2 of columns per x-pass tile
    int vec_io;                       // 1: Nz % 4 == 0 and act/dens/out 16-byte aligned (float4 FZ loads
                                      // along the lines); 2: act 16-byte aligned (float4 runs of a tile)
    int sym;                          // 1: spec holds the real octant [Px/2+1][Py/2+1][Hp] (FP32)
    int wplane;                       // complex elements per x-plane of work (Py*Hp + padding)
    const float *act[PVD_MAX_TIMEPOINTS];
    float w[PVD_MAX_TIMEPOINTS];
    const float *dens;
    float *out;
    float scale;
    int pad2_;
    float2 *work;                     // [Nx][Py][Hp] complex workspace, x-planes wplane apart
    const float2 *spec;               // [Px][Py][Hp] kernel spectrum (1/P folded in); see sym
    int tiles[PVD_NPASS_FFT];         // tiles of this volume per pass
    int pad3_;
};

// One batch launch of one FFT pass: the members are volumes that run the same
// kernel instance with the same block; the launch's grid concatenates their tiles.
struct pvd_batch_args {
    const pvd_vol_d *vols;            // descriptor table of the whole batch (device)
    const int *base;                  // first grid tile of each member, ascending (device)
    const int *vol;                   // descriptor index of each member (device)
    int nmem;                         // members of this launch
    int ntiles;                       // tiles of this launch (before any rounding of the grid)
};

// ---------------------------------------------------------------------------------
// Workgroup barrier that orders LDS only.  __syncthreads() carries a release fence
// for all memory, which waits for every outstanding global load (vmcnt(0)) before
// the barrier; the FFT passes exchange data only through LDS and consume every
// global load before its value is shared, so their barriers wait for LDS
// (lgkmcnt) alone and leave prefetched global loads in flight.
__device__ __forceinline__ void lds_barrier() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// ---------------------------------------------------------------------------------
// complex helpers
__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
    return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ float2 cscale(float2 a, float s) { return make_float2(a.x * s, a.y * s); }
// multiply by -i (forward) or +i (inverse)
template <bool INV> __device__ __forceinline__ float2 mul_mi(float2 a) {
    return INV ? make_float2(-a.y, a.x) : make_float2(a.y, -a.x);
}

// ---------------------------------------------------------------------------------
// small DFTs, forward: X[k] = sum_n x[n] e^{-2 pi i nk/R}; inverse: e^{+...}
template <int R, bool INV> struct Dft;

template <bool INV> struct Dft<2, INV> {
    __device__ __forceinline__ static void run(float2 *v) {
        float2 a = v[0], b = v[1];
        v[0] = cadd(a, b);
        v[1] = csub(a, b);
    }
};

template <bool INV> struct Dft<4, INV> {
    __device__ __forceinline__ static void run(float2 *v) {
        float2 t0 = cadd(v[0], v[2]), t1 = csub(v[0], v[2]);
        float2 t2 = cadd(v[1], v[3]), t3 = mul_mi<INV>(csub(v[1], v[3]));
        v[0] = cadd(t0, t2);
        v[2] = csub(t0, t2);
        v[1] = cadd(t1, t3);
        v[3] = csub(t1, t3);
    }
};

template <bool INV> struct Dft<8, INV> {
    __device__ __forceinline__ static void run(float2 *v) {
        const float r = 0.70710678118654752440f;
        float2 e[4] = {v[0], v[2], v[4], v[6]};
        float2 o[4] = {v[1], v[3], v[5], v[7]};
        Dft<4, INV>::run(e);
        Dft<4, INV>::run(o);
        // twiddles W8^k (forward e^{-i pi k/4}); inverse conjugated
        float2 o1 = INV ? make_float2(r * (o[1].x - o[1].y), r * (o[1].x + o[1].y))
                        : make_float2(r * (o[1].x + o[1].y), r * (o[1].y - o[1].x));
        float2 o2 = mul_mi<INV>(o[2]);
        float2 o3 = INV ? make_float2(-r * (o[3].x + o[3].y), r * (o[3].x - o[3].y))
                        : make_float2(r * (o[3].y - o[3].x), -r * (o[3].x + o[3].y));
        v[0] = cadd(e[0], o[0]); v[4] = csub(e[0], o[0]);
        v[1] = cadd(e[1], o1);   v[5] = csub(e[1], o1);
        v[2] = cadd(e[2], o2);   v[6] = csub(e[2], o2);
        v[3] = cadd(e[3], o3);   v[7] = csub(e[3], o3);
    }
};

template <bool INV> struct Dft<16, INV> {
    __device__ __forceinline__ static void run(float2 *v) {
        float2 e[8], o[8];
#pragma unroll
        for (int i = 0; i < 8; ++i) { e[i] = v[2 * i]; o[i] = v[2 * i + 1]; }
        Dft<8, INV>::run(e);
        Dft<8, INV>::run(o);
        // W16^k = cos(pi k/8) -/+ i sin(pi k/8)
        const float c[8] = {1.0f, 0.92387953251128675613f, 0.70710678118654752440f,
                            0.38268343236508977173f, 0.0f, -0.38268343236508977173f,
                            -0.70710678118654752440f, -0.92387953251128675613f};
        const float s[8] = {0.0f, 0.38268343236508977173f, 0.70710678118654752440f,
                            0.92387953251128675613f, 1.0f, 0.92387953251128675613f,
                            0.70710678118654752440f, 0.38268343236508977173f};
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            float2 w = make_float2(c[k], INV ? s[k] : -s[k]);
            float2 t = (k == 0) ? o[0] : cmul(o[k], w);
            v[k] = cadd(e[k], t);
            v[k + 8] = csub(e[k], t);
        }
    }
};

template <bool INV> struct Dft<3, INV> {
    __device__ __forceinline__ static void run(float2 *v) {
        const float s3 = 0.86602540378443864676f;
        float2 t = cadd(v[1], v[2]);
        float2 m = make_float2(v[0].x - 0.5f * t.x, v[0].y - 0.5f * t.y);
        float2 d = cscale(mul_mi<INV>(csub(v[1], v[2])), s3);
        v[0] = cadd(v[0], t);
        v[1] = cadd(m, d);
        v[2] = csub(m, d);
    }
};

template <bool INV> struct Dft<5, INV> {
    __device__ __forceinline__ static void run(float2 *v) {
        const float c1 = 0.30901699437494742410f, c2 = -0.80901699437494742410f;
        const float s1 = 0.95105651629515357212f, s2 = 0.58778525229247312917f;
        float2 b1 = cadd(v[1], v[4]), b2 = cadd(v[2], v[3]);
        float2 d1 = csub(v[1], v[4]), d2 = csub(v[2], v[3]);
        float2 a0 = v[0];
        float2 m1 = make_float2(a0.x + c1 * b1.x + c2 * b2.x, a0.y + c1 * b1.y + c2 * b2.y);
        float2 m2 = make_float2(a0.x + c2 * b1.x + c1 * b2.x, a0.y + c2 * b1.y + c1 * b2.y);
        float2 n1 = mul_mi<INV>(make_float2(s1 * d1.x + s2 * d2.x, s1 * d1.y + s2 * d2.y));
        float2 n2 = mul_mi<INV>(make_float2(s2 * d1.x - s1 * d2.x, s2 * d1.y - s1 * d2.y));
        v[0] = make_float2(a0.x + b1.x + b2.x, a0.y + b1.y + b2.y);
        v[1] = cadd(m1, n1);
        v[4] = csub(m1, n1);
        v[2] = cadd(m2, n2);
        v[3] = csub(m2, n2);
    }
};

template <bool INV> struct Dft<7, INV> {
    __device__ __forceinline__ static void run(float2 *v) {
        // cos/sin(2 pi j/7), j = 1..3
        const float c[3] = {0.62348980185873353053f, -0.22252093395631440429f, -0.90096886790241912624f};
        const float s[3] = {0.78183148246802980871f, 0.97492791218182360702f, 0.43388373911755812048f};
        float2 b[3], d[3];
#pragma unroll
        for (int j = 0; j < 3; ++j) { b[j] = cadd(v[j + 1], v[6 - j]); d[j] = csub(v[j + 1], v[6 - j]); }
        float2 a0 = v[0];
        float2 X[7];
        X[0] = make_float2(a0.x + b[0].x + b[1].x + b[2].x, a0.y + b[0].y + b[1].y + b[2].y);
#pragma unroll
        for (int k = 1; k <= 3; ++k) {
            float2 m = a0, n = make_float2(0.f, 0.f);
#pragma unroll
            for (int j = 0; j < 3; ++j) {
                const int q = ((j + 1) * k) % 7;           // angle index
                const float cc = (q <= 3) ? c[q - 1] : c[7 - q - 1];
                const float ss = (q <= 3) ? s[q - 1] : -s[7 - q - 1];
                m.x += cc * b[j].x; m.y += cc * b[j].y;
                n.x += ss * d[j].x; n.y += ss * d[j].y;
            }
            float2 t = mul_mi<INV>(n);
            X[k] = cadd(m, t);
            X[7 - k] = csub(m, t);
        }
#pragma unroll
        for (int k = 0; k < 7; ++k) v[k] = X[k];
    }
};

// ---------------------------------------------------------------------------------
// Compile-time twiddles: cos / sin(2 pi j / N) from Taylor series in double on
// [-pi, pi] (terms to x^41, error < 1e-16), rounded once to float.
constexpr double kCtPi = 3.14159265358979323846;
constexpr double ct_sin(double x) {
    double term = x, sum = x;
    for (int i = 1; i < 21; ++i) { term *= -x * x / ((2.0 * i) * (2.0 * i + 1.0)); sum += term; }
    return sum;
}
constexpr double ct_cos(double x) {
    double term = 1.0, sum = 1.0;
    for (int i = 1; i < 21; ++i) { term *= -x * x / ((2.0 * i - 1.0) * (2.0 * i)); sum += term; }
    return sum;
}
struct CtTwiddles { float c[64], s[64]; };
template <int N> constexpr CtTwiddles ct_twiddles() {
    CtTwiddles t{};
    for (int j = 0; j < N; ++j) {
        double a = 2.0 * kCtPi * j / N;
        if (a > kCtPi) a -= 2.0 * kCtPi;
        t.c[j] = (float)ct_cos(a);
        t.s[j] = (float)ct_sin(a);
    }
    return t;
}

template <int N> constexpr bool dft_basic() {
    return N == 2 || N == 3 || N == 4 || N == 5 || N == 7 || N == 8 || N == 16;
}
template <int N> constexpr int dft_first_factor() {
    return (N % 8 == 0) ? 8 : (N % 4 == 0) ? 4 : (N % 5 == 0) ? 5 : (N % 3 == 0) ? 3 : (N % 7 == 0) ? 7 : 2;
}

// In-register DFT of any length N whose factors are 2, 3, 5, 7 (natural order in
// and out): the basic lengths directly, others as one mixed-radix split
// N = R1 * R2, n = R2 n1 + n2, k = k1 + R1 k2:
// X[k] = sum_n2 W_R2^(n2 k2) W_N^(n2 k1) sum_n1 x[R2 n1 + n2] W_R1^(n1 k1),
// with the W_N twiddles folded to constants at compile time.
template <int N, bool INV>
__device__ __forceinline__ void dft_any(float2 *v) {
    if constexpr (dft_basic<N>()) {
        Dft<N, INV>::run(v);
    } else {
        constexpr int R1 = dft_first_factor<N>(), R2 = N / R1;
        static_assert(R1 * R2 == N && R2 > 1 && N <= 64, "dft_any: unsupported length");
        constexpr CtTwiddles w = ct_twiddles<N>();
        float2 u[N];
#pragma unroll
        for (int n2 = 0; n2 < R2; ++n2) {
            float2 t[R1];
#pragma unroll
            for (int n1 = 0; n1 < R1; ++n1) t[n1] = v[R2 * n1 + n2];
            dft_any<R1, INV>(t);
#pragma unroll
            for (int k1 = 0; k1 < R1; ++k1) {
                const int j = (n2 * k1) % N;
                u[k1 * R2 + n2] = (j == 0) ? t[k1] : cmul(t[k1], make_float2(w.c[j], INV ? w.s[j] : -w.s[j]));
            }
        }
#pragma unroll
        for (int k1 = 0; k1 < R1; ++k1) {
            float2 t[R2];
#pragma unroll
            for (int n2 = 0; n2 < R2; ++n2) t[n2] = u[k1 * R2 + n2];
            dft_any<R2, INV>(t);
#pragma unroll
            for (int k2 = 0; k2 < R2; ++k2) v[k1 + R1 * k2] = t[k2];
        }
    }
}

// z lengths with a register two-step transform n = A * B (A a power of two):
// code c of pvd_axis_d.reg selects (kZregA[c], kZregB[c]); 0 = none.  B stays at
// most 30: the z kernels hold the B-point DFT of every code in one switch, so the
// largest B sets their register allocation.
#define PVD_ZREG_CODES 21
constexpr int kZregA[PVD_ZREG_CODES] = {0, 8, 8, 8, 8, 16, 8, 16, 16, 16, 8, 8, 8, 8, 8, 8, 16, 16, 16, 16, 16};
constexpr int kZregB[PVD_ZREG_CODES] = {0, 16, 20, 24, 25, 16, 30, 20, 25, 27, 12, 14, 15, 18, 21, 27, 14, 18, 21, 24, 30};
// largest block of the y/x register two-step kernel instances
#define PVD_XY_MAX_BLOCK 512
// y/x lengths 16 * B other than 256 with a register two-step instance, grouped by
// the source file that compiles them (pvd_fft_xy1/2/3.hip)
// two-step lengths 16 * B with B >= PVD_XY_LDS_B run the LDS-resident form (the tile
// loaded into LDS by all threads); smaller B the form fused with its HBM loads
#define PVD_XY_LDS_B 35
// concurrent branches per pass in the batch graph (the groups of one pass overlap)
#define PVD_BATCH_BRANCHES 4
#define PVD_XY_BS_1(X) X(6) X(8) X(9) X(10) X(12) X(14) X(16)
#define PVD_XY_BS_2(X) X(15) X(18) X(20) X(21) X(24) X(25)
#define PVD_XY_BS_3(X) X(27) X(28) X(30) X(32) X(35) X(36)
#define PVD_XY_BS(X) PVD_XY_BS_1(X) PVD_XY_BS_2(X) PVD_XY_BS_3(X)

// ---------------------------------------------------------------------------------
// a / d for 0 <= a < 2^22 and 1 <= d <= 2^16 from a float reciprocal plus one
// correction step (the float estimate is within 1 of the true quotient).
__device__ __forceinline__ int fast_div(int a, int d, float inv_d) {
    int q = (int)((float)a * inv_d);
    const int r = a - q * d;
    if (r >= d) ++q;
    else if (r < 0) --q;
    return q;
}

// One Stockham stage of radix R over nl = 2^lnl lines held in LDS, element
// (line l, index k) at s[k*ld + l].  Ns = product of the radices of the stages
// already done.  Every thread keeps up to IT butterflies in registers between
// the read and the write phase so the stage is in place; the launch guarantees
// nl*n <= 16*blockDim, hence nl*n/R <= IT*blockDim.
template <int R> struct StageItems { static constexpr int v = (16 + R - 1) / R; };

template <int R, bool INV>
__device__ __forceinline__ void stockham_stage(float2 *__restrict__ s, int lnl, int ld, int n, int Ns,
                                               const float2 *__restrict__ tw) {
    constexpr int IT = StageItems<R>::v;
    const int m = n / R;
    const int work = m << lnl;
    const int tstride = n / (Ns * R);
    const float inv_ns = 1.0f / (float)Ns;
    float2 v[IT][R];
    int dst[IT];
#pragma unroll
    for (int it = 0; it < IT; ++it) {
        const int w = threadIdx.x + it * blockDim.x;
        dst[it] = -1;
        if (w < work) {
            const int j = w >> lnl;
            const int l = w & ((1 << lnl) - 1);
            const int q = fast_div(j, Ns, inv_ns);
            const int k = j - q * Ns;
#pragma unroll
            for (int r = 0; r < R; ++r) v[it][r] = s[(j + r * m) * ld + l];
            if (k != 0) {
                const int t = k * tstride;
#pragma unroll
                for (int r = 1; r < R; ++r) {
                    float2 wv = tw[r * t];
                    if (INV) wv.y = -wv.y;
                    v[it][r] = cmul(v[it][r], wv);
                }
            }
            Dft<R, INV>::run(v[it]);
            dst[it] = (q * Ns * R + k) * ld + l;
        }
    }
    lds_barrier();
#pragma unroll
    for (int it = 0; it < IT; ++it) {
        if (dst[it] >= 0) {
#pragma unroll
            for (int r = 0; r < R; ++r) s[dst[it] + r * Ns * ld] = v[it][r];
        }
    }
    lds_barrier();
}

// Full FFT of length ax.n over 2^lnl lines in LDS (natural order in and out,
// unnormalised).  Must be entered after a barrier that published `s` and `tw`.
template <bool INV>
__device__ __forceinline__ void fft_lines(float2 *__restrict__ s, int lnl, int ld, const pvd_axis_d &ax,
                                          const float2 *__restrict__ tw) {
    const int n = ax.n;
    int Ns = 1;
    for (int st = 0; st < ax.nst; ++st) {
        const int R = ax.radix[st];
        switch (R) {
            case 16: stockham_stage<16, INV>(s, lnl, ld, n, Ns, tw); break;
            case 8: stockham_stage<8, INV>(s, lnl, ld, n, Ns, tw); break;
            case 4: stockham_stage<4, INV>(s, lnl, ld, n, Ns, tw); break;
            case 2: stockham_stage<2, INV>(s, lnl, ld, n, Ns, tw); break;
            case 3: stockham_stage<3, INV>(s, lnl, ld, n, Ns, tw); break;
            case 5: stockham_stage<5, INV>(s, lnl, ld, n, Ns, tw); break;
            case 7: stockham_stage<7, INV>(s, lnl, ld, n, Ns, tw); break;
            default: break;
        }
        Ns *= R;
    }
}

// ---------------------------------------------------------------------------------
// Kernel launchers (host).  A launch runs one pass either for one volume (`v`, the
// descriptor passed by value) or for the members of a batch launch (`a`, with `v`
// any member's host descriptor: it selects the kernel instance).
hipError_t pvd_launch_fft_pass(int pass, const pvd_vol_d &v, int grid, int block, size_t lds, hipStream_t st);
hipError_t pvd_launch_fft_pass_batch(int pass, const pvd_vol_d &v, const pvd_batch_args &a, int grid, int block,
                                     size_t lds, hipStream_t st);
hipError_t pvd_set_fft_lds_limits(size_t bytes);
// per source file of the z instances (pvd_fft_z.hip) and of the y/x two-step
// instances (pvd_fft_xy1/2/3.hip); `batch` selects the table-reading instance
hipError_t pvd_fft_launch_z(int pass, bool rz, int tp, const pvd_vol_d *v, const pvd_batch_args *a, int grid,
                            int block, size_t lds, hipStream_t st);
hipError_t pvd_fft_launch_xy1(int pass, int rb, const pvd_vol_d *v, const pvd_batch_args *a, int grid, int block,
                              size_t lds, hipStream_t st);
hipError_t pvd_fft_launch_xy2(int pass, int rb, const pvd_vol_d *v, const pvd_batch_args *a, int grid, int block,
                              size_t lds, hipStream_t st);
hipError_t pvd_fft_launch_xy3(int pass, int rb, const pvd_vol_d *v, const pvd_batch_args *a, int grid, int block,
                              size_t lds, hipStream_t st);
hipError_t pvd_fft_z_set_lds(size_t bytes);
hipError_t pvd_fft_xy1_set_lds(size_t bytes);
hipError_t pvd_fft_xy2_set_lds(size_t bytes);
hipError_t pvd_fft_xy3_set_lds(size_t bytes);

// Record an error message for pvd_last_error() and return `code` (pvd_api.hip).
int pvd_set_error(int code, const char *fmt, ...);
// Map a HIP error to PVD_ENOMEM / PVD_EHIP with a message.
int pvd_hip_error(hipError_t e, const char *what);

// rho_w / rho with rho_w = 1 g/cc; 0 where rho <= 0 (declared epilogue).
// The hardware reciprocal (within 1 ulp) replaces the IEEE division sequence.
__device__ __forceinline__ float density_factor(float rho) { return rho > 0.0f ? __builtin_amdgcn_rcpf(rho) : 0.0f; }
