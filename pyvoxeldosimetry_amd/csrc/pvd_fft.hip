// FFT dose-convolution passes (gfx950).
//
// The dose is computed as five descriptor-driven passes over an in-place
// complex workspace work[Nx][Py][H] (H = Pz/2+1):
//   FZ  z: weighted sum of the T activity maps, zero-pad, two-for-one real FFT
//   FY  y: forward FFT of every (x, kz) column
//   X   x: forward FFT, times the kernel spectrum, inverse FFT, fold/roll map
//   IY  y: inverse FFT, fold/roll map
//   IZ  z: two-for-one inverse real FFT, fold/roll map, scale * rho_w/rho, store
// Every pass stages its lines in LDS (k-major: element (line l, index k) at
// s[k*ld + l]) so global loads and stores are coalesced along the contiguous
// axis of the workspace and the Stockham stages run out of LDS.
#include "pvd_internal.h"

namespace {

__device__ __forceinline__ int wrap_add(int a, int b, int n) {  // (a + b) mod n, a,b in [0, n)
    int r = a + b;
    return r >= n ? r - n : r;
}

__device__ __forceinline__ void load_twiddles(float2 *tw, const pvd_axis_d &ax) {
    for (int i = threadIdx.x; i < ax.n; i += blockDim.x) tw[i] = ax.tw[i];
}

// ---------------------------------------------------------------------------------
// FZ: lines are the Nx*Ny real z-lines in flattened order L = x*Ny + y; a tile
// holds 2*lz consecutive real lines paired into lz complex lines.
__device__ void pass_fz(const pvd_vol_d &v, int tile, float2 *tw, float2 *s) {
    const pvd_axis_d &az = v.ax[2];
    const int Nz = az.N, Pz = az.n, H = v.H, LZ = v.lz, ld = LZ + 1;
    const int Ny = v.ax[1].N, Py = v.ax[1].n;
    const long long nlines = (long long)v.ax[0].N * Ny;
    const long long L0 = (long long)tile * 2 * LZ;
    const int nreal = (int)min((long long)(2 * LZ), nlines - L0);
    load_twiddles(tw, az);
    for (int e = threadIdx.x; e < Pz * ld; e += blockDim.x) s[e] = make_float2(0.f, 0.f);
    __syncthreads();
    const long long base = L0 * Nz;
    const int total = nreal * Nz;
    for (int e = threadIdx.x; e < total; e += blockDim.x) {
        const int ln = e / Nz;
        const int z = e - ln * Nz;
        float a = v.w[0] * v.act[0][base + e];
        for (int t = 1; t < v.T; ++t) a = fmaf(v.w[t], v.act[t][base + e], a);
        float *p = reinterpret_cast<float *>(&s[z * ld + (ln >> 1)]);
        p[ln & 1] = a;
    }
    __syncthreads();
    fft_lines<false>(s, LZ, ld, az, tw);
    // split the two real spectra: A = (Z_k + conj Z_{-k})/2, B = (Z_k - conj Z_{-k})/(2i)
    for (int e = threadIdx.x; e < nreal * H; e += blockDim.x) {
        const int ln = e / H;
        const int k = e - ln * H;
        const int q = ln >> 1;
        const float2 zk = s[k * ld + q];
        const float2 zm = s[(k == 0 ? 0 : Pz - k) * ld + q];
        const float2 r = (ln & 1) ? make_float2(0.5f * (zk.y + zm.y), 0.5f * (zm.x - zk.x))
                                  : make_float2(0.5f * (zk.x + zm.x), 0.5f * (zk.y - zm.y));
        const long long L = L0 + ln;
        const long long x = L / Ny, y = L - x * Ny;
        v.work[(x * Py + y) * H + k] = r;
    }
}

// ---------------------------------------------------------------------------------
// FY / IY: tile = (x, block of tky consecutive kz columns); lines run along y.
template <bool INV>
__device__ void pass_y(const pvd_vol_d &v, int tile, float2 *tw, float2 *s) {
    const pvd_axis_d &ay = v.ax[1];
    const int Ny = ay.N, Py = ay.n, H = v.H, TK = v.tky;
    const int ntc = (H + TK - 1) / TK;
    const int x = tile / ntc;
    const int c0 = (tile - x * ntc) * TK;
    const int nc = min(TK, H - c0);
    float2 *g = v.work + ((long long)x * Py) * H + c0;
    load_twiddles(tw, ay);
    const int rows_in = INV ? Py : Ny;
    for (int e = threadIdx.x; e < Py * TK; e += blockDim.x) {
        const int k = e / TK;
        const int l = e - k * TK;
        s[e] = (k < rows_in && l < nc) ? g[(long long)k * H + l] : make_float2(0.f, 0.f);
    }
    __syncthreads();
    fft_lines<INV>(s, TK, TK, ay, tw);
    if (!INV) {
        for (int e = threadIdx.x; e < Py * TK; e += blockDim.x) {
            const int k = e / TK;
            const int l = e - k * TK;
            if (l < nc) g[(long long)k * H + l] = s[e];
        }
    } else {
        for (int e = threadIdx.x; e < Ny * TK; e += blockDim.x) {
            const int o = e / TK;
            const int l = e - o * TK;
            if (l < nc) {
                float2 val = s[wrap_add(o, ay.off, Py) * TK + l];
                if (o < ay.foldn) val = cadd(val, s[((o + Ny + ay.off) % Py) * TK + l]);
                g[(long long)o * H + l] = val;
            }
        }
    }
}

// ---------------------------------------------------------------------------------
// X: tile = block of tkx consecutive columns of the flattened (y, kz) plane;
// forward FFT along x, times the kernel spectrum, inverse FFT, fold/roll map.
__device__ void pass_x(const pvd_vol_d &v, int tile, float2 *tw, float2 *s) {
    const pvd_axis_d &ax = v.ax[0];
    const int Nx = ax.N, Px = ax.n, TK = v.tkx;
    const long long plane = (long long)v.ax[1].n * v.H;
    const long long c0 = (long long)tile * TK;
    const int nc = (int)min((long long)TK, plane - c0);
    float2 *g = v.work + c0;
    const float2 *sp = v.spec + c0;
    load_twiddles(tw, ax);
    for (int e = threadIdx.x; e < Px * TK; e += blockDim.x) {
        const int k = e / TK;
        const int l = e - k * TK;
        s[e] = (k < Nx && l < nc) ? g[k * plane + l] : make_float2(0.f, 0.f);
    }
    __syncthreads();
    fft_lines<false>(s, TK, TK, ax, tw);
    for (int e = threadIdx.x; e < Px * TK; e += blockDim.x) {
        const int k = e / TK;
        const int l = e - k * TK;
        if (l < nc) s[e] = cmul(s[e], sp[k * plane + l]);
    }
    __syncthreads();
    fft_lines<true>(s, TK, TK, ax, tw);
    for (int e = threadIdx.x; e < Nx * TK; e += blockDim.x) {
        const int o = e / TK;
        const int l = e - o * TK;
        if (l < nc) {
            float2 val = s[wrap_add(o, ax.off, Px) * TK + l];
            if (o < ax.foldn) val = cadd(val, s[((o + Nx + ax.off) % Px) * TK + l]);
            g[o * plane + l] = val;
        }
    }
}

// ---------------------------------------------------------------------------------
// IZ: tile = 2*lz consecutive output lines (x, y); inverse two-for-one real FFT.
__device__ void pass_iz(const pvd_vol_d &v, int tile, float2 *tw, float2 *s) {
    const pvd_axis_d &az = v.ax[2];
    const int Nz = az.N, Pz = az.n, H = v.H, LZ = v.lz, ld = LZ + 1;
    const int Ny = v.ax[1].N, Py = v.ax[1].n;
    const long long nlines = (long long)v.ax[0].N * Ny;
    const long long L0 = (long long)tile * 2 * LZ;
    const int nreal = (int)min((long long)(2 * LZ), nlines - L0);
    const int npair = (nreal + 1) >> 1;
    load_twiddles(tw, az);
    // Z_k = A_k + i B_k on [0, H); Z_{P-k} = conj(A_k) + i conj(B_k) on the rest
    for (int e = threadIdx.x; e < npair * H; e += blockDim.x) {
        const int q = e / H;
        const int k = e - q * H;
        const long long La = L0 + 2 * q;
        const long long xa = La / Ny, ya = La - xa * Ny;
        const float2 A = v.work[(xa * Py + ya) * H + k];
        float2 B = make_float2(0.f, 0.f);
        if (2 * q + 1 < nreal) {
            const long long Lb = La + 1;
            const long long xb = Lb / Ny, yb = Lb - xb * Ny;
            B = v.work[(xb * Py + yb) * H + k];
        }
        s[k * ld + q] = make_float2(A.x - B.y, A.y + B.x);
        if (k >= 1 && k <= Pz - H) s[(Pz - k) * ld + q] = make_float2(A.x + B.y, B.x - A.y);
    }
    for (int e = threadIdx.x; e < Pz * (LZ - npair); e += blockDim.x) {   // unused pairs
        const int k = e / (LZ - npair);
        const int q = npair + (e - k * (LZ - npair));
        s[k * ld + q] = make_float2(0.f, 0.f);
    }
    __syncthreads();
    fft_lines<true>(s, LZ, ld, az, tw);
    const long long obase = L0 * Nz;
    const int total = nreal * Nz;
    for (int e = threadIdx.x; e < total; e += blockDim.x) {
        const int ln = e / Nz;
        const int o = e - ln * Nz;
        const int q = ln >> 1;
        const float *p0 = reinterpret_cast<const float *>(&s[wrap_add(o, az.off, Pz) * ld + q]);
        float val = p0[ln & 1];
        if (o < az.foldn) {
            const float *p1 = reinterpret_cast<const float *>(&s[((o + Nz + az.off) % Pz) * ld + q]);
            val += p1[ln & 1];
        }
        float f = v.scale;
        if (v.dens) f *= density_factor(v.dens[obase + e]);
        v.out[obase + e] = val * f;
    }
}

template <int PASS>
__device__ __forceinline__ void run_pass(const pvd_vol_d &v, int tile) {
    extern __shared__ float2 smem[];
    const int axis = (PASS == PVD_PASS_X) ? 0 : (PASS == PVD_PASS_FY || PASS == PVD_PASS_IY) ? 1 : 2;
    float2 *tw = smem;
    float2 *s = smem + ((v.ax[axis].n + 7) & ~7);
    if (PASS == PVD_PASS_FZ) pass_fz(v, tile, tw, s);
    else if (PASS == PVD_PASS_FY) pass_y<false>(v, tile, tw, s);
    else if (PASS == PVD_PASS_X) pass_x(v, tile, tw, s);
    else if (PASS == PVD_PASS_IY) pass_y<true>(v, tile, tw, s);
    else pass_iz(v, tile, tw, s);
}

}  // namespace

// single volume: descriptor by value (kernarg segment, scalar loads)
template <int PASS>
__global__ __launch_bounds__(1024) void pvd_fft_pass(pvd_vol_d v) {
    run_pass<PASS>(v, blockIdx.x);
}

// batch: grid covers the tiles of every volume; find this block's volume
template <int PASS>
__global__ __launch_bounds__(1024) void pvd_fft_pass_batch(pvd_batch_args a) {
    const int b = blockIdx.x;
    int lo = 0, hi = a.nvol - 1;
    while (lo < hi) {   // last volume with tile_base <= b
        const int mid = (lo + hi + 1) >> 1;
        if (a.vols[mid].tile_base[PASS] <= b) lo = mid; else hi = mid - 1;
    }
    const pvd_vol_d &v = a.vols[lo];
    const int tile = b - v.tile_base[PASS];
    if (tile < v.tiles[PASS]) run_pass<PASS>(v, tile);
}

// host-side launchers (called from pvd_api.hip)
hipError_t pvd_launch_fft_pass(int pass, const pvd_vol_d &v, int grid, int block, size_t lds,
                               hipStream_t st) {
    switch (pass) {
        case PVD_PASS_FZ: hipLaunchKernelGGL(pvd_fft_pass<PVD_PASS_FZ>, dim3(grid), dim3(block), lds, st, v); break;
        case PVD_PASS_FY: hipLaunchKernelGGL(pvd_fft_pass<PVD_PASS_FY>, dim3(grid), dim3(block), lds, st, v); break;
        case PVD_PASS_X: hipLaunchKernelGGL(pvd_fft_pass<PVD_PASS_X>, dim3(grid), dim3(block), lds, st, v); break;
        case PVD_PASS_IY: hipLaunchKernelGGL(pvd_fft_pass<PVD_PASS_IY>, dim3(grid), dim3(block), lds, st, v); break;
        default: hipLaunchKernelGGL(pvd_fft_pass<PVD_PASS_IZ>, dim3(grid), dim3(block), lds, st, v); break;
    }
    return hipGetLastError();
}

hipError_t pvd_launch_fft_pass_batch(int pass, const pvd_batch_args &a, int grid, int block, size_t lds,
                                     hipStream_t st) {
    switch (pass) {
        case PVD_PASS_FZ: hipLaunchKernelGGL(pvd_fft_pass_batch<PVD_PASS_FZ>, dim3(grid), dim3(block), lds, st, a); break;
        case PVD_PASS_FY: hipLaunchKernelGGL(pvd_fft_pass_batch<PVD_PASS_FY>, dim3(grid), dim3(block), lds, st, a); break;
        case PVD_PASS_X: hipLaunchKernelGGL(pvd_fft_pass_batch<PVD_PASS_X>, dim3(grid), dim3(block), lds, st, a); break;
        case PVD_PASS_IY: hipLaunchKernelGGL(pvd_fft_pass_batch<PVD_PASS_IY>, dim3(grid), dim3(block), lds, st, a); break;
        default: hipLaunchKernelGGL(pvd_fft_pass_batch<PVD_PASS_IZ>, dim3(grid), dim3(block), lds, st, a); break;
    }
    return hipGetLastError();
}

hipError_t pvd_set_fft_lds_limits(size_t bytes) {
    hipError_t e = hipSuccess;
#define PVD_SETATTR(K) \
    if (e == hipSuccess) e = hipFuncSetAttribute((const void *)K, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes)
    PVD_SETATTR(pvd_fft_pass<PVD_PASS_FZ>);
    PVD_SETATTR(pvd_fft_pass<PVD_PASS_FY>);
    PVD_SETATTR(pvd_fft_pass<PVD_PASS_X>);
    PVD_SETATTR(pvd_fft_pass<PVD_PASS_IY>);
    PVD_SETATTR(pvd_fft_pass<PVD_PASS_IZ>);
    PVD_SETATTR(pvd_fft_pass_batch<PVD_PASS_FZ>);
    PVD_SETATTR(pvd_fft_pass_batch<PVD_PASS_FY>);
    PVD_SETATTR(pvd_fft_pass_batch<PVD_PASS_X>);
    PVD_SETATTR(pvd_fft_pass_batch<PVD_PASS_IY>);
    PVD_SETATTR(pvd_fft_pass_batch<PVD_PASS_IZ>);
#undef PVD_SETATTR
    return e;
}
