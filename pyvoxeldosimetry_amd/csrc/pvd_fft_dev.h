// Device code of the FFT dose-convolution passes (gfx950), shared by the kernel
// instances in pvd_fft.hip (generic and z-register kernels) and pvd_fft_xy*.hip
// (register two-step y/x kernels).  Everything here is inline and internal.
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
#pragma once

#include "pvd_internal.h"

namespace {

#ifndef PVD_LOAD_BATCH
#define PVD_LOAD_BATCH 16
#endif
#ifdef PVD_FFT_WAVES
#define PVD_FFT_OCC __attribute__((amdgpu_waves_per_eu(PVD_FFT_WAVES)))
#else
#define PVD_FFT_OCC
#endif
constexpr int kBatch = PVD_LOAD_BATCH;   // global loads issued per thread before the first is consumed

__device__ __forceinline__ int wrap_add(int a, int b, int n) {  // (a + b) mod n, a,b in [0, n)
    int r = a + b;
    return r >= n ? r - n : r;
}

// Element at a 32-bit byte offset from a wave-uniform base: the load or store takes
// the base in scalar registers and one VGPR of offset (no 64-bit address math).
// The 256-point passes address at most Nx * wplane * 8 <= 256 * (256 * 1032 + pad) * 8
// bytes of workspace and 129 * 129 * 1032 * 4 bytes of spectrum; an FZ tile spans at
// most 2 * PVD_TILE_MAX lines of PVD_MAX_FFT_LEN floats: all well below 4 GiB.
// Every pointer it takes is to global memory (HBM): the access is typed as such, so
// that pointers read from a device-side descriptor table (the batch kernels) still
// give global rather than flat loads and stores.
// (The qualifier is for the device pass only: the host pass parses these inline
// functions as well and has no global address space.)
#ifdef __HIP_DEVICE_COMPILE__
#define PVD_GLOBAL __attribute__((address_space(1)))
#else
#define PVD_GLOBAL
#endif
template <typename T>
__device__ __forceinline__ PVD_GLOBAL T *gmem(T *p) {
    return (PVD_GLOBAL T *)p;
}
template <typename T>
__device__ __forceinline__ PVD_GLOBAL T &at(T *base, uint32_t byte_off) {
    return *(PVD_GLOBAL T *)((char *)base + byte_off);
}
template <typename T>
__device__ __forceinline__ const PVD_GLOBAL T &at(const T *base, uint32_t byte_off) {
    return *(const PVD_GLOBAL T *)((const char *)base + byte_off);
}

__device__ __forceinline__ void load_twiddles(float2 *tw, const pvd_axis_d &ax) {
    for (int i = threadIdx.x; i < ax.n; i += blockDim.x) tw[i] = gmem(ax.tw)[i];
}

// Load rows [0, rows_in) x columns [0, nc) of a strided tile into LDS
// s[k*nl + l] (zero elsewhere up to n rows); row k starts at g + k*pitch.
// Moves 16-byte column pairs: g and pitch are even (in complex elements) and
// rows are padded to whole tiles, so a pair straddling nc is read whole and its
// upper half replaced by zero; pairs wholly past nc are not read.
__device__ __forceinline__ void load_col_tile(float2 *__restrict__ s, const float2 *__restrict__ g, long long pitch,
                                              int n, int rows_in, int lnl, int nc) {
    const int lp = lnl - 1;
    const int total = n << lp;
    const int pmask = (1 << lp) - 1;
    float4 *s4 = reinterpret_cast<float4 *>(s);
    for (int e0 = threadIdx.x; e0 < total; e0 += kBatch * blockDim.x) {
        float4 v[kBatch];
#pragma unroll
        for (int u = 0; u < kBatch; ++u) {
            const int e = e0 + u * blockDim.x;
            const int k = e >> lp, c = (e & pmask) << 1;
            v[u] = (e < total && k < rows_in && c < nc) ? float4(*gmem(reinterpret_cast<const float4 *>(g + k * pitch + c)))
                                                        : make_float4(0.f, 0.f, 0.f, 0.f);
        }
#pragma unroll
        for (int u = 0; u < kBatch; ++u) {
            const int e = e0 + u * blockDim.x;
            if (e < total) {
                const int c = (e & pmask) << 1;
                float4 x = v[u];
                if (c >= nc) { x.x = 0.f; x.y = 0.f; }
                if (c + 1 >= nc) { x.z = 0.f; x.w = 0.f; }
                s4[e] = x;
            }
        }
    }
}

// Store rows o in [0, rows_out) of a column tile from LDS, whole 16-byte column
// pairs (the upper half of a pair straddling nc lands in row padding; pairs
// wholly past nc are skipped): row o comes from LDS row (o + off) mod P, plus
// row (o + rows_out + off) mod P when o < foldn.
__device__ __forceinline__ void store_col_tile(float2 *__restrict__ g, long long pitch, const float2 *__restrict__ s,
                                               int rows_out, int lnl, int P, int off, int foldn, int nc) {
    const int lp = lnl - 1;
    const int total = rows_out << lp;
    const int pmask = (1 << lp) - 1;
    const float4 *s4 = reinterpret_cast<const float4 *>(s);
    for (int e = threadIdx.x; e < total; e += blockDim.x) {
        const int o = e >> lp, cp = e & pmask;
        if ((cp << 1) >= nc) continue;
        float4 val = s4[(wrap_add(o, off, P) << lp) + cp];
        if (o < foldn) {
            const float4 f = s4[(((o + rows_out + off) % P) << lp) + cp];
            val.x += f.x; val.y += f.y; val.z += f.z; val.w += f.w;
        }
        *gmem(reinterpret_cast<float4 *>(g + o * pitch + (cp << 1))) = val;
    }
}

// ---------------------------------------------------------------------------------
// Offsets in the workspace of the (up to 2 * PVD_TILE_MAX) real z-lines of a z-pass tile.
__device__ __forceinline__ void tile_line_offsets(long long *lineoff, long long L0, int nreal, int Ny, int wplane,
                                                  int Hp) {
    for (int i = threadIdx.x; i < nreal; i += blockDim.x) {
        const int L = (int)(L0 + i);
        const int x = fast_div(L, Ny, 1.0f / (float)Ny);
        lineoff[i] = (long long)x * wplane + (long long)(L - x * Ny) * Hp;
    }
}

// ---------------------------------------------------------------------------------
// Register two-step transform of the z passes, n = A * B, for 16 lines in LDS
// (element (line q, index k) at s[k*17 + q]).  Natural order in; the result
// X[k1 + A k2] is left in row k1*B + k2 (zrow() maps an index to its row):
//   step 1: thread (q, b) reads rows b + B n1 (n1 < A), DFT-A in registers,
//           multiplies by W_n^(b k1) (b k1 < n, table in LDS) and writes the
//           same rows back, Y[k1][b] in row k1*B + b;
//   step 2: thread (q, k1) reads rows k1*B .. k1*B + B-1, DFT-B in registers
//           (compile-time twiddles) and writes X[k1 + A k2] to row k1*B + k2.
// Every thread writes only the rows it read, so one barrier after each step.
// Needs 16*A <= blockDim and 16*B <= 2*blockDim; entered after a barrier.
template <int A, int B, bool INV>
__device__ __forceinline__ void zreg_fft(float2 *__restrict__ s, const float2 *__restrict__ tw) {
    constexpr int LD = 17;
#pragma unroll
    for (int it = 0; it < 2; ++it) {
        const int w = threadIdx.x + it * blockDim.x;
        const int b = w >> 4, q = w & 15;
        if (b < B) {
            float2 *p = s + b * LD + q;
            float2 v[A];
#pragma unroll
            for (int n1 = 0; n1 < A; ++n1) v[n1] = p[n1 * B * LD];
            dft_any<A, INV>(v);
            p[0] = v[0];
#pragma unroll
            for (int k1 = 1; k1 < A; ++k1) {
                float2 t = tw[b * k1];
                if (INV) t.y = -t.y;
                p[k1 * B * LD] = (b == 0) ? v[k1] : cmul(v[k1], t);
            }
        }
    }
    lds_barrier();
    {
        const int w = threadIdx.x;
        const int k1 = w >> 4, q = w & 15;
        if (k1 < A) {
            float2 *p = s + k1 * B * LD + q;
            float2 v[B];
#pragma unroll
            for (int n2 = 0; n2 < B; ++n2) v[n2] = p[n2 * LD];
            dft_any<B, INV>(v);
#pragma unroll
            for (int k2 = 0; k2 < B; ++k2) p[k2 * LD] = v[k2];
        }
    }
    lds_barrier();
}

// Row of LDS that holds transform index k after z_fft.
__device__ __forceinline__ int zrow(const pvd_axis_d &az, int k) {
    return az.reg ? (k & (az.ra - 1)) * az.rb + (k >> az.ra_log) : k;
}

// z-axis FFT over the lines of a z-pass tile: the register two-step where the
// plan selected one (16 lines per tile), else the LDS Stockham stages.  RZ
// kernels (launched only for az.reg != 0) leave the Stockham code out, which
// takes the z passes from 106 to 75 VGPRs.
template <bool INV, bool RZ>
__device__ __forceinline__ void z_fft(float2 *s, int lnl, int ld, const pvd_axis_d &az, const float2 *tw) {
    switch (az.reg) {
        case 1: zreg_fft<8, 16, INV>(s, tw); break;
        case 2: zreg_fft<8, 20, INV>(s, tw); break;
        case 3: zreg_fft<8, 24, INV>(s, tw); break;
        case 4: zreg_fft<8, 25, INV>(s, tw); break;
        case 5: zreg_fft<16, 16, INV>(s, tw); break;
        case 6: zreg_fft<8, 30, INV>(s, tw); break;
        case 7: zreg_fft<16, 20, INV>(s, tw); break;
        case 8: zreg_fft<16, 25, INV>(s, tw); break;
        case 9: zreg_fft<16, 27, INV>(s, tw); break;
        case 10: zreg_fft<8, 12, INV>(s, tw); break;
        case 11: zreg_fft<8, 14, INV>(s, tw); break;
        case 12: zreg_fft<8, 15, INV>(s, tw); break;
        case 13: zreg_fft<8, 18, INV>(s, tw); break;
        case 14: zreg_fft<8, 21, INV>(s, tw); break;
        case 15: zreg_fft<8, 27, INV>(s, tw); break;
        case 16: zreg_fft<16, 14, INV>(s, tw); break;
        case 17: zreg_fft<16, 18, INV>(s, tw); break;
        case 18: zreg_fft<16, 21, INV>(s, tw); break;
        case 19: zreg_fft<16, 24, INV>(s, tw); break;
        case 20: zreg_fft<16, 30, INV>(s, tw); break;
        default:
            if (!RZ) fft_lines<INV>(s, lnl, ld, az, tw);
            break;
    }
}

// Row step of z_fft's result per +TP in the index, and the drop in rows where an
// index wraps past n: TP / A and B for the register two-step (row of k is
// (k mod A) * B + k / A; A divides TP), TP and n for the Stockham stages (row of
// k is k).
__device__ __forceinline__ int zrow_step(const pvd_axis_d &az, int tp) { return az.reg ? tp >> az.ra_log : tp; }
__device__ __forceinline__ int zrow_wrap(const pvd_axis_d &az) { return az.reg ? az.rb : az.n; }

// The z passes move the elements t + TP j (j = 0, 1, ...) of one line pair per
// thread (q, t), TP threads per pair: 16, 32 or 64, the largest that blockDim
// covers for the tile's lines (the plan sizes z blocks to a power-of-two multiple
// of the lines; smaller blocks loop).  Global accesses are TP consecutive words
// per line, LDS rows advance by a constant stride (zrow_step), both real lines of
// a pair travel in one float2, and no index is divided.  A divides TP.
constexpr int kZBatch = 8;      // elements per line in flight per thread
template <int TP> constexpr int z_tpl() { return TP == 16 ? 4 : TP == 32 ? 5 : 6; }

// FZ: lines are the Nx*Ny real z-lines in flattened order L = x*Ny + y; a tile
// holds 2*nl consecutive real lines paired into nl complex lines (ld = nl + 1),
// real line 2q in the real and 2q+1 in the imaginary part of complex line q.
// FZ phases with TP threads per line pair (see kZBatch): the T-weighted load of
// rows z < Nz, and the split of the two real spectra into the workspace.
// FZ load of 16-byte aligned maps whose Nz is not a multiple of 4: the tile's real
// lines L0 .. L0 + nreal - 1 are one contiguous run of nreal * Nz floats starting at a
// multiple of 2 * nl lines (so 16-byte aligned), read as float4 (every load of a map
// issued, at the run's first float4 where past its end, before any is used), summed
// over the T maps, and each float written to its line pair's LDS row z (real part for
// even lines, imaginary for odd).  The run's last nreal * Nz % 4 floats are read one by
// one; lines past the last real line get zeros.
__device__ __forceinline__ void fz_load_chunk(const pvd_vol_d &v, float2 *s, long long L0, int nreal, int nl, int ld,
                                              int Nz) {
    constexpr int kC4 = 8;                     // float4 per thread in flight per map
    float *sf = reinterpret_cast<float *>(s);
    const int n = nreal * Nz, n4 = n >> 2;
    for (int e = threadIdx.x; e < (2 * nl - nreal) * Nz; e += blockDim.x) {
        const int line = nreal + e / Nz, z = e - (e / Nz) * Nz;
        sf[2 * (z * ld + (line >> 1)) + (line & 1)] = 0.f;
    }
    const float inv = 1.0f / (float)Nz;
    auto put = [&](int f, float x) {           // float f of the run -> its LDS slot
        int line = (int)((float)f * inv);
        if (line * Nz > f) --line;
        else if ((line + 1) * Nz <= f) ++line;
        const int z = f - line * Nz;
        sf[2 * (z * ld + (line >> 1)) + (line & 1)] = x;
    };
    for (int f0 = threadIdx.x; f0 < n4; f0 += kC4 * blockDim.x) {
        float4 a[kC4];
        auto load = [&](const float *map, float4 *x) {
            const float4 *m = reinterpret_cast<const float4 *>(map + L0 * Nz);
#pragma unroll
            for (int u = 0; u < kC4; ++u) {
                const int f = f0 + u * blockDim.x;
                x[u] = at(m, (uint32_t)(f < n4 ? f : 0) * 16u);
            }
        };
        load(v.act[0], a);
        const float w0 = v.w[0];
#pragma unroll
        for (int u = 0; u < kC4; ++u) a[u] = make_float4(w0 * a[u].x, w0 * a[u].y, w0 * a[u].z, w0 * a[u].w);
        for (int tt = 1; tt < v.T; ++tt) {
            float4 x[kC4];
            load(v.act[tt], x);
            const float wt = v.w[tt];
#pragma unroll
            for (int u = 0; u < kC4; ++u) {
                a[u].x = fmaf(wt, x[u].x, a[u].x); a[u].y = fmaf(wt, x[u].y, a[u].y);
                a[u].z = fmaf(wt, x[u].z, a[u].z); a[u].w = fmaf(wt, x[u].w, a[u].w);
            }
        }
#pragma unroll
        for (int u = 0; u < kC4; ++u) {
            const int f = f0 + u * blockDim.x;
            if (f < n4) {
                // the four floats of float4 f: one division, then step along the line
                const int e0 = 4 * f;
                int line = (int)((float)e0 * inv);
                if (line * Nz > e0) --line;
                else if ((line + 1) * Nz <= e0) ++line;
                int z = e0 - line * Nz;
                const float x4[4] = {a[u].x, a[u].y, a[u].z, a[u].w};
#pragma unroll
                for (int c = 0; c < 4; ++c) {
                    sf[2 * (z * ld + (line >> 1)) + (line & 1)] = x4[c];
                    if (++z == Nz) { z = 0; ++line; }
                }
            }
        }
    }
    for (int f = 4 * n4 + (int)threadIdx.x; f < n; f += blockDim.x) {
        const float *m = v.act[0] + L0 * Nz;
        float acc = v.w[0] * at(m, (uint32_t)f * 4u);
        for (int tt = 1; tt < v.T; ++tt) acc = fmaf(v.w[tt], at(v.act[tt] + L0 * Nz, (uint32_t)f * 4u), acc);
        put(f, acc);
    }
}

template <int TP>
__device__ __forceinline__ void fz_load(const pvd_vol_d &v, float2 *s, long long L0, int nreal, int nl, int ld,
                                        int Nz) {
    constexpr int tpl = z_tpl<TP>();
    // weighted sum of the T maps into rows z < Nz; pairs past the last line get zeros.
    // Every load of a map is issued (at element 0 of the map where the line or index is
    // past the end) before any is used, and masked afterwards: loads under a branch are
    // waited for one by one.
    if (v.vec_io == 2) {
        fz_load_chunk(v, s, L0, nreal, nl, ld, Nz);
    } else if (v.vec_io) {     // Nz % 4 == 0 and 16-byte aligned maps: float4 z = 4 (t + TP j)
        const int nz4 = Nz >> 2;
        constexpr int kB4 = kZBatch / 2;
        for (int w = threadIdx.x; w < nl * TP; w += blockDim.x) {
            const int q = w >> tpl, t = w & (TP - 1);
            const bool h0 = 2 * q < nreal, h1 = 2 * q + 1 < nreal;
            // 32-bit byte offsets from the tile's first element (wave-uniform): one VGPR
            // per address; the tile spans at most 2 * PVD_TILE_MAX lines of Nz floats
            const uint32_t l0 = (uint32_t)(2 * q * nz4);     // float4 index of real line 2q in the tile
            for (int f0 = t; f0 < nz4; f0 += TP * kB4) {
                float4 a[kB4], b[kB4];
                auto load = [&](const float *map, float4 *xa, float4 *xb) {
                    const float4 *m = reinterpret_cast<const float4 *>(map + L0 * Nz);
#pragma unroll
                    for (int u = 0; u < kB4; ++u) {
                        const int f = f0 + TP * u;
                        xa[u] = at(m, ((f < nz4 && h0) ? l0 + f : 0u) * 16u);
                        xb[u] = at(m, ((f < nz4 && h1) ? l0 + nz4 + f : 0u) * 16u);
                    }
                };
                load(v.act[0], a, b);
                const float w0 = v.w[0];
#pragma unroll
                for (int u = 0; u < kB4; ++u) {
                    const int f = f0 + TP * u;
                    const float4 x = a[u], y = b[u];
                    a[u] = (f < nz4 && h0) ? make_float4(w0 * x.x, w0 * x.y, w0 * x.z, w0 * x.w)
                                           : make_float4(0.f, 0.f, 0.f, 0.f);
                    b[u] = (f < nz4 && h1) ? make_float4(w0 * y.x, w0 * y.y, w0 * y.z, w0 * y.w)
                                           : make_float4(0.f, 0.f, 0.f, 0.f);
                }
                for (int tt = 1; tt < v.T; ++tt) {
                    float4 xa[kB4], xb[kB4];
                    load(v.act[tt], xa, xb);
                    const float wt = v.w[tt];
#pragma unroll
                    for (int u = 0; u < kB4; ++u) {
                        const int f = f0 + TP * u;
                        if (f < nz4 && h0) {
                            const float4 x = xa[u];
                            a[u].x = fmaf(wt, x.x, a[u].x); a[u].y = fmaf(wt, x.y, a[u].y);
                            a[u].z = fmaf(wt, x.z, a[u].z); a[u].w = fmaf(wt, x.w, a[u].w);
                        }
                        if (f < nz4 && h1) {
                            const float4 y = xb[u];
                            b[u].x = fmaf(wt, y.x, b[u].x); b[u].y = fmaf(wt, y.y, b[u].y);
                            b[u].z = fmaf(wt, y.z, b[u].z); b[u].w = fmaf(wt, y.w, b[u].w);
                        }
                    }
                }
#pragma unroll
                for (int u = 0; u < kB4; ++u) {
                    const int f = f0 + TP * u;
                    if (f < nz4) {
                        float2 *d = s + 4 * f * ld + q;
                        d[0] = make_float2(a[u].x, b[u].x);
                        d[ld] = make_float2(a[u].y, b[u].y);
                        d[2 * ld] = make_float2(a[u].z, b[u].z);
                        d[3 * ld] = make_float2(a[u].w, b[u].w);
                    }
                }
            }
        }
    } else for (int w = threadIdx.x; w < nl * TP; w += blockDim.x) {
        const int q = w >> tpl, t = w & (TP - 1);
        const bool h0 = 2 * q < nreal, h1 = 2 * q + 1 < nreal;
        const uint32_t o0 = (uint32_t)(2 * q * Nz);     // element 0 of real line 2q in the tile
        for (int z0 = t; z0 < Nz; z0 += TP * kZBatch) {
            float xa[kZBatch], xb[kZBatch];
            auto load = [&](const float *map) {        // byte offsets from the tile's first element
                const float *m = map + L0 * Nz;
#pragma unroll
                for (int u = 0; u < kZBatch; ++u) {
                    const int z = z0 + TP * u;
                    xa[u] = at(m, ((z < Nz && h0) ? o0 + z : 0u) * 4u);
                    xb[u] = at(m, ((z < Nz && h1) ? o0 + Nz + z : 0u) * 4u);
                }
            };
            float2 a[kZBatch];
            load(v.act[0]);
#pragma unroll
            for (int u = 0; u < kZBatch; ++u) {
                const int z = z0 + TP * u;
                a[u].x = (z < Nz && h0) ? v.w[0] * xa[u] : 0.f;
                a[u].y = (z < Nz && h1) ? v.w[0] * xb[u] : 0.f;
            }
            for (int tt = 1; tt < v.T; ++tt) {
                load(v.act[tt]);
                const float wt = v.w[tt];
#pragma unroll
                for (int u = 0; u < kZBatch; ++u) {
                    const int z = z0 + TP * u;
                    if (z < Nz && h0) a[u].x = fmaf(wt, xa[u], a[u].x);
                    if (z < Nz && h1) a[u].y = fmaf(wt, xb[u], a[u].y);
                }
            }
            float2 *d = s + z0 * ld + q;
#pragma unroll
            for (int u = 0; u < kZBatch; ++u)
                if (z0 + TP * u < Nz) d[TP * u * ld] = a[u];
        }
    }
}

template <int TP>
__device__ __forceinline__ void fz_split(const pvd_vol_d &v, float2 *s, const long long *lineoff,
                                         const pvd_axis_d &az, int npair, int nreal, int ld, int H, int Pz) {
    constexpr int tpl = z_tpl<TP>();
    // split the two real spectra: A = (Z_k + conj Z_{-k})/2, B = (Z_k - conj Z_{-k})/(2i),
    // k = t + TP j; index P - k has row zrow(P - t) - j * step (zrow(P) is B or P)
    const int rs = zrow_step(az, TP) * ld;
    for (int w = threadIdx.x; w < npair * TP; w += blockDim.x) {
        const int q = w >> tpl, t = w & (TP - 1);
        if (t >= H) continue;
        const bool h1 = 2 * q + 1 < nreal;
        const float2 *pk = s + zrow(az, t) * ld + q;
        const float2 *pm = s + zrow(az, Pz - t) * ld + q;
        PVD_GLOBAL float2 *g0 = gmem(v.work) + lineoff[2 * q] + t;
        PVD_GLOBAL float2 *g1 = gmem(v.work) + (h1 ? lineoff[2 * q + 1] : 0) + t;
        for (int k0 = t, j0 = 0; k0 < H; k0 += TP * kZBatch, j0 += kZBatch) {
            float2 zk[kZBatch], zm[kZBatch];       // all LDS reads of the batch before any store
#pragma unroll
            for (int u = 0; u < kZBatch; ++u) {
                const int k = k0 + TP * u, j = j0 + u;
                zk[u] = make_float2(0.f, 0.f);
                zm[u] = make_float2(0.f, 0.f);
                if (k < H) {
                    zk[u] = pk[j * rs];
                    zm[u] = k != 0 ? pm[-j * rs] : zk[u];
                }
            }
#pragma unroll
            for (int u = 0; u < kZBatch; ++u) {
                const int k = k0 + TP * u, j = j0 + u;
                if (k < H) {
                    g0[TP * j] = make_float2(0.5f * (zk[u].x + zm[u].x), 0.5f * (zk[u].y - zm[u].y));
                    if (h1) g1[TP * j] = make_float2(0.5f * (zk[u].y + zm[u].y), 0.5f * (zm[u].x - zk[u].x));
                }
            }
        }
    }
}

template <bool RZ, int TP>
__device__ void pass_fz(const pvd_vol_d &v, int tile, float2 *tw, float2 *s) {
    __shared__ long long lineoff[2 * PVD_TILE_MAX];
    const pvd_axis_d &az = v.ax[2];
    const int Nz = az.N, Pz = az.n, H = v.H, Hp = v.Hp;
    const int lnl = v.llz, nl = 1 << lnl, ld = nl + 1;
    const int Ny = v.ax[1].N, Py = v.ax[1].n;
    const long long nlines = (long long)v.ax[0].N * Ny;
    const long long L0 = (long long)tile * 2 * nl;
    const int nreal = (int)min((long long)(2 * nl), nlines - L0);
    const int npair = (nreal + 1) >> 1;
    load_twiddles(tw, az);
    tile_line_offsets(lineoff, L0, nreal, Ny, v.wplane, Hp);
    // zero the padding rows z >= Nz of every line
    for (int e = threadIdx.x; e < (Pz - Nz) * ld; e += blockDim.x) s[Nz * ld + e] = make_float2(0.f, 0.f);
    fz_load<TP>(v, s, L0, nreal, nl, ld, Nz);
    lds_barrier();
    z_fft<false, RZ>(s, lnl, ld, az, tw);
    fz_split<TP>(v, s, lineoff, az, npair, nreal, ld, H, Pz);
}

// ---------------------------------------------------------------------------------
// Length-256 column FFTs held in registers (the y and x passes of 256-point
// axes, 16 columns per tile, 256 threads).  Four-step split 256 = 16 x 16:
// thread (l = column, g) owns the 16 elements g, g+16, ... of its column, does a
// 16-point DFT in registers, applies W_256^(g k1), exchanges through LDS once
// ([k1][n2][l], rows padded by kSqPad), and finishes with a second 16-point DFT:
// X[k1 + 16 k2] = sum_n2 W_16^(n2 k2) W_256^(n2 k1) sum_n1 x[16 n1 + n2] W_16^(n1 k1).
// One barrier per transform instead of one per radix stage, and no generic
// index arithmetic.  The inverse of data held as X[k1 + 16 k2] (k2 in
// registers) is the same split in the other order:
// x[j1 + 16 j2] = sum_k1 W_16^(-j2 k1) W_256^(-j1 k1) sum_k2 X[k1 + 16 k2] W_16^(-j1 k2).
constexpr int kSqPad = PVD_SQ256_PAD;      // float2 padding between k1 planes (bank spread)
constexpr int kSqPlane = 256 + kSqPad;

__device__ __forceinline__ bool sq256_ok(const pvd_axis_d &ax, int lnl) {
    return ax.reg && ax.n == 256 && lnl == 4 && blockDim.x == 256 && ax.foldn == 0;
}

// v[k1] = sum_n1 v[n1] W16^(+-n1 k1), then times W256^(+-g k1)
template <bool INV>
__device__ __forceinline__ void sq_first(float2 *v, int g, const float2 *tw) {
    Dft<16, INV>::run(v);
#pragma unroll
    for (int k = 1; k < 16; ++k) {
        float2 w = tw[g * k];
        if (INV) w.y = -w.y;
        v[k] = cmul(v[k], w);
    }
}

// forward 256-point transform of registers x[n1] = column element 16 n1 + g;
// returns X[k1 + 16 k2] in v[k2] for k1 = g (after one LDS exchange)
template <bool INV>
__device__ __forceinline__ void sq_transform(float2 *v, int l, int g, const float2 *tw, float2 *s) {
    sq_first<INV>(v, g, tw);
#pragma unroll
    for (int k1 = 0; k1 < 16; ++k1) s[k1 * kSqPlane + g * 16 + l] = v[k1];
    lds_barrier();
#pragma unroll
    for (int n2 = 0; n2 < 16; ++n2) v[n2] = s[g * kSqPlane + n2 * 16 + l];
    Dft<16, INV>::run(v);
}

template <bool INV>
__device__ void pass_y_sq(const pvd_vol_d &v, int tile, float2 *tw, float2 *s) {
    const pvd_axis_d &ay = v.ax[1];
    const int Ny = ay.N, H = v.H, Hp = v.Hp;
    const int ntc = (H + 15) >> 4;
    const int x = tile / ntc;
    const int c0 = (tile - x * ntc) << 4;
    const int nc = min(16, H - c0);
    float2 *g = v.work + (long long)x * v.wplane + c0;    // uniform base of the tile
    const int l = threadIdx.x & 15, grp = threadIdx.x >> 4;
    const float2 tw1 = gmem(ay.tw)[threadIdx.x];             // 256 twiddles, one per thread, in flight
    const int rows_in = INV ? 256 : Ny;
    // lanes past the tile's last column (nc) read that column: no 64-byte sector of row
    // padding is fetched (their values are masked; they store nothing)
    const uint32_t lo = (uint32_t)(grp * Hp + min(l, nc - 1)) * 8u, rstep = (uint32_t)Hp * 128u;   // + 16 rows
    // every row r < 256 = Py and column c0 + l < Hp lies in the workspace: the loads
    // are unconditional (no branches around them) and masked afterwards
    float2 a[16];
#pragma unroll
    for (int n1 = 0; n1 < 16; ++n1) {
        const int r = 16 * n1 + grp;
        const float2 t = at((const float2 *)g, lo + n1 * rstep);
        a[n1] = (r < rows_in && l < nc) ? t : make_float2(0.f, 0.f);
    }
    tw[threadIdx.x] = tw1;
    lds_barrier();                                     // twiddles visible
    sq_transform<INV>(a, l, grp, tw, s);
    if (l >= nc) return;
#pragma unroll
    for (int k2 = 0; k2 < 16; ++k2) {
        const int k = grp + 16 * k2;
        if (!INV) {
            at(g, lo + k2 * rstep) = a[k2];
        } else {
            const int o = k >= ay.off ? k - ay.off : k + 256 - ay.off;     // output row reading index k
            if (o < Ny) at(g, (uint32_t)(o * Hp + l) * 8u) = a[k2];
        }
    }
}

// X-pass tiles run through the y rows in mirror order 0, 1, Py-1, 2, Py-2, ...:
// rows y and Py - y read the same rows of the real-octant spectrum, and
// neighbouring column tiles share its 128-byte lines, so the tiles that share
// spectrum lines are adjacent in the tile order.
__device__ __forceinline__ int x_tile_row(int u, int Py) {
    return u == 0 ? 0 : (u & 1) ? (u + 1) >> 1 : Py - (u >> 1);
}

__device__ void pass_x_sq(const pvd_vol_d &v, int tile, float2 *tw, float2 *s) {
    const pvd_axis_d &ax = v.ax[0];
    const int Nx = ax.N, H = v.H, Hp = v.Hp;
    const int ntc = (H + 15) >> 4;
    const int Py = v.ax[1].n;
    const int u = tile / ntc;
    const int y = x_tile_row(u, Py);
    const int c0 = (tile - u * ntc) << 4;
    const int nc = min(16, H - c0);
    float2 *g = v.work + (long long)y * Hp + c0;       // uniform base; x-planes wplane apart
    const uint32_t pstride = (uint32_t)v.wplane * 8u;  // bytes per x-plane
    const int l = threadIdx.x & 15, grp = threadIdx.x >> 4;
    const int lc = min(l, nc - 1);                     // lanes past nc read the last column
    const float2 tw1 = gmem(ax.tw)[threadIdx.x];             // 256 twiddles, one per thread, in flight
    // loads are unconditional (rows clamped to Nx - 1, lanes past the tile's last column
    // read that column, so no 64-byte sector of row padding is fetched) and masked
    // afterwards, so no branch skips them
    float2 a[16];
#pragma unroll
    for (int n1 = 0; n1 < 16; ++n1)
        a[n1] = at((const float2 *)g, (uint32_t)min(16 * n1 + grp, Nx - 1) * pstride + (uint32_t)lc * 8u);
    // kernel spectrum of rows k = grp + 16 k2, in flight through the forward transform;
    // columns l >= nc (row padding) multiply whatever the padding holds into columns
    // that are never stored, and no column mixes with another
    float2 sp[16];
    if (v.sym) {
        // real octant [Px/2+1][Py/2+1][Hp] FP32: row min(k, 256-k), column (min(y, Py-y), c0 + l)
        const float *sr = reinterpret_cast<const float *>(v.spec) + (long long)min(y, Py - y) * Hp + c0;
        const uint32_t krow = (uint32_t)(Py / 2 + 1) * (uint32_t)Hp * 4u;    // bytes per kx row
#pragma unroll
        for (int k2 = 0; k2 < 16; ++k2) {
            const int k = grp + 16 * k2;
            sp[k2] = make_float2(at(sr, (uint32_t)min(k, 256 - k) * krow + (uint32_t)lc * 4u), 0.f);
        }
    } else {
        const float2 *sg = v.spec + (long long)y * Hp + c0;
        const uint32_t kplane = (uint32_t)Py * (uint32_t)Hp * 8u;          // bytes per kx plane
#pragma unroll
        for (int k2 = 0; k2 < 16; ++k2) {
            const int k = grp + 16 * k2;
            sp[k2] = at(sg, (uint32_t)k * kplane + (uint32_t)lc * 8u);
        }
    }
    // all 33 loads are issued before any result is consumed
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int n1 = 0; n1 < 16; ++n1)
        if (!(16 * n1 + grp < Nx && l < nc)) a[n1] = make_float2(0.f, 0.f);   // zero rows past Nx
    tw[threadIdx.x] = tw1;
    lds_barrier();                                     // twiddles visible
    sq_transform<false>(a, l, grp, tw, s);             // a[k2] = X[grp + 16 k2]
#pragma unroll
    for (int k2 = 0; k2 < 16; ++k2) a[k2] = v.sym ? cscale(a[k2], sp[k2].x) : cmul(a[k2], sp[k2]);
    // inverse: thread-local 16-point DFT over k2 and twiddle W256^(-j1 k1), k1 = grp
    lds_barrier();                                     // every thread done reading s
    sq_transform<true>(a, l, grp, tw, s);              // a[j2] = x[grp + 16 j2]
    if (l >= nc) return;
#pragma unroll
    for (int j2 = 0; j2 < 16; ++j2) {
        const int j = grp + 16 * j2;
        const int o = j >= ax.off ? j - ax.off : j + 256 - ax.off;
        if (o < Nx) at(g, (uint32_t)o * pstride + (uint32_t)l * 8u) = a[j2];
    }
}

// ---------------------------------------------------------------------------------
// FY / IY: tile = (x, block of 2^ltky consecutive kz columns); lines run along y.
template <bool INV>
__device__ void pass_y(const pvd_vol_d &v, int tile, float2 *tw, float2 *s) {
    const pvd_axis_d &ay = v.ax[1];
    const int Ny = ay.N, Py = ay.n, H = v.H, Hp = v.Hp;
    const int lnl = v.ltky, TK = 1 << lnl;
    const int ntc = (H + TK - 1) >> lnl;
    const int x = tile / ntc;
    const int c0 = (tile - x * ntc) << lnl;
    const int nc = min(TK, H - c0);
    float2 *g = v.work + (long long)x * v.wplane + c0;
    load_twiddles(tw, ay);
    load_col_tile(s, g, Hp, Py, INV ? Py : Ny, lnl, nc);
    lds_barrier();
    fft_lines<INV>(s, lnl, TK, ay, tw);
    if (!INV) store_col_tile(g, Hp, s, Py, lnl, Py, 0, 0, nc);
    else store_col_tile(g, Hp, s, Ny, lnl, Py, ay.off, ay.foldn, nc);
}

// ---------------------------------------------------------------------------------
// X: tile = (y, block of 2^ltkx consecutive kz columns) of the (y, kz) plane;
// forward FFT along x, times the kernel spectrum, inverse FFT, fold/roll map.
__device__ void pass_x(const pvd_vol_d &v, int tile, float2 *tw, float2 *s) {
    const pvd_axis_d &ax = v.ax[0];
    const int Nx = ax.N, Px = ax.n, H = v.H, Hp = v.Hp;
    const int lnl = v.ltkx, TK = 1 << lnl;
    const int ntc = (H + TK - 1) >> lnl;
    const int u = tile / ntc;
    const int y = x_tile_row(u, v.ax[1].n);
    const int c0 = (tile - u * ntc) << lnl;
    const int nc = min(TK, H - c0);
    const long long plane = v.wplane;                 // work
    const long long splane = (long long)v.ax[1].n * Hp;   // spectrum
    const long long off0 = (long long)y * Hp + c0;
    float2 *g = v.work + off0;
    const PVD_GLOBAL float2 *sp = gmem(v.spec) + off0;
    load_twiddles(tw, ax);
    load_col_tile(s, g, plane, Px, Nx, lnl, nc);
    lds_barrier();
    fft_lines<false>(s, lnl, TK, ax, tw);
    const int total = Px << lnl;
    if (v.sym) {
        // real spectrum, even in kx and ky: read the octant entry of (min(kx, Px-kx), min(y, Py-y))
        const PVD_GLOBAL float *sr = (const PVD_GLOBAL float *)v.spec;
        const int Py = v.ax[1].n;
        const long long pyh = Py / 2 + 1;
        const long long row0 = (long long)min(y, Py - y) * Hp + c0;
        // (every load issued, at the tile's first entry where masked, before any is
        // used: loads under a branch are waited for one by one)
        for (int e0 = threadIdx.x; e0 < total; e0 += kBatch * blockDim.x) {
            float sv[kBatch];
#pragma unroll
            for (int u = 0; u < kBatch; ++u) {
                const int e = e0 + u * blockDim.x;
                const int k = e >> lnl, l = e & (TK - 1);
                const bool ok = e < total && l < nc;
                sv[u] = sr[ok ? (long long)min(k, Px - k) * pyh * Hp + row0 + l : row0];
            }
#pragma unroll
            for (int u = 0; u < kBatch; ++u) {
                const int e = e0 + u * blockDim.x;
                const bool ok = e < total && (e & (TK - 1)) < nc;
                if (e < total) s[e] = cscale(s[e], ok ? sv[u] : 0.f);
            }
        }
    } else {
        for (int e0 = threadIdx.x; e0 < total; e0 += kBatch * blockDim.x) {
            float2 sv[kBatch];
#pragma unroll
            for (int u = 0; u < kBatch; ++u) {
                const int e = e0 + u * blockDim.x;
                const int k = e >> lnl, l = e & (TK - 1);
                sv[u] = sp[(e < total && l < nc) ? k * splane + l : 0];
            }
#pragma unroll
            for (int u = 0; u < kBatch; ++u) {
                const int e = e0 + u * blockDim.x;
                const bool ok = e < total && (e & (TK - 1)) < nc;
                if (e < total) s[e] = cmul(s[e], ok ? sv[u] : make_float2(0.f, 0.f));
            }
        }
    }
    lds_barrier();
    fft_lines<true>(s, lnl, TK, ax, tw);
    store_col_tile(g, plane, s, Nx, lnl, Px, ax.off, ax.foldn, nc);
}

// ---------------------------------------------------------------------------------
// Register two-step transforms of the y and x passes for lengths P = 16 * B other
// than 256 (B = 6 ... 36), over column tiles of TK = 2^lnl lines.  With n = B n1 + b
// and k = k1 + 16 k2:
//   X[k1 + 16 k2] = sum_b W_B^(b k2) W_P^(b k1) sum_n1 x[B n1 + b] W_16^(n1 k1).
// Step 1 runs straight from HBM: item (b, l) loads x[b + B n1] (n1 < 16) of column
// l, does the DFT-16, multiplies by W_P^(b k1) and writes LDS row b + B k1
// (element (row r, line l) at s[r * TK + l]).  After one barrier, step 2 in
// registers: item (k1, l) reads rows k1 B .. k1 B + B-1 and does the DFT-B, leaving
// X[k1 + 16 k2] in v[k2].
// FY stores X straight to HBM (row k1 + 16 k2).  IY stores output row o, which reads
// index j = (o + off) mod P, from the item holding j.  X multiplies by the kernel
// spectrum in the same registers, runs the inverse DFT-B and twiddles W_P^(-b k1),
// writes the rows back, and after one barrier the inverse DFT-16 of item (b, l)
// leaves x[b + B n1] in registers, stored the same way.
// Reference-mode fold: rows o < foldn also take index (o + N + off) mod P, which
// another item holds, so a tile with a fold is written back to its own LDS rows (IY:
// X[k] in row xy_row<B>(k); X: x[n] in row n) and stored by the row maps of
// store_col_tile(_perm) after one barrier.  (A separate 2 * foldn * TK fold buffer
// past the tile, for direct stores of the other rows, cost a workgroup per CU on the
// 24-point tiles and 4 % of the 48-phantom batch.)
// block bound of the two-step instances (their item loops are strided, so the host
// caps the block at this; the register budget per instance is xy_waves())
constexpr int kXyThreads = PVD_XY_MAX_BLOCK;

// LDS row of FFT index k after the two-step: (k mod 16) B + k / 16
template <int B>
__device__ __forceinline__ int xy_row(int k) { return (k & 15) * B + (k >> 4); }

// store_col_tile for a tile left in xy_row<B> order: FFT index k sits in LDS row
// xy_row<B>(k) (same output mapping: row o takes index (o + off) mod P, plus index
// (o + rows_out + off) mod P when o < foldn).
template <int B>
__device__ __forceinline__ void store_col_tile_perm(float2 *__restrict__ g, long long pitch,
                                                    const float2 *__restrict__ s, int rows_out, int lnl, int P,
                                                    int off, int foldn, int nc) {
    const int lp = lnl - 1;
    const int total = rows_out << lp;
    const int pmask = (1 << lp) - 1;
    const float4 *s4 = reinterpret_cast<const float4 *>(s);
    for (int e = threadIdx.x; e < total; e += blockDim.x) {
        const int o = e >> lp, cp = e & pmask;
        if ((cp << 1) >= nc) continue;
        float4 val = s4[(xy_row<B>(wrap_add(o, off, P)) << lp) + cp];
        if (o < foldn) {
            const float4 f = s4[(xy_row<B>((o + rows_out + off) % P) << lp) + cp];
            val.x += f.x; val.y += f.y; val.z += f.z; val.w += f.w;
        }
        *gmem(reinterpret_cast<float4 *>(g + o * pitch + (cp << 1))) = val;
    }
}


// Addressing: every global access is a 32-bit byte offset from the tile's wave-uniform
// base (one VGPR, no 64-bit address math: at()).  The host takes the fused form only
// where the X pass's workspace walk, (Nx - 1) * wplane * 8 bytes, and the general
// spectrum, (Px - 1) * Py * Hp * 8 bytes, stay below 4 GiB (xy_offsets_fit); the
// y passes stay inside one x-plane.

// step 1 from HBM (all loads of an item issued before any is used; rows >= rows_in and
// columns >= nc are zero); entered after a barrier that published tw
template <int B, bool INV>
__device__ __forceinline__ void xy_step1_global(float2 *__restrict__ s, int lnl, const float2 *__restrict__ tw,
                                                const float2 *__restrict__ g, uint32_t pitch_b, int rows_in, int nc) {
    const int TK = 1 << lnl;
    for (int w = threadIdx.x; w < (B << lnl); w += blockDim.x) {
        const int b = w >> lnl, l = w & (TK - 1);
        const uint32_t lb = (uint32_t)min(l, nc - 1) * 8u;     // past nc: the last column's sector
        float2 v[16];
#pragma unroll
        for (int n1 = 0; n1 < 16; ++n1) v[n1] = at(g, (uint32_t)min(b + B * n1, rows_in - 1) * pitch_b + lb);
#pragma unroll
        for (int n1 = 0; n1 < 16; ++n1)
            if (b + B * n1 >= rows_in || l >= nc) v[n1] = make_float2(0.f, 0.f);
        Dft<16, INV>::run(v);
#pragma unroll
        for (int k1 = 1; k1 < 16; ++k1) {
            float2 t = tw[b * k1];
            if (INV) t.y = -t.y;
            v[k1] = cmul(v[k1], t);
        }
#pragma unroll
        for (int k1 = 0; k1 < 16; ++k1) s[(b + B * k1) * TK + l] = v[k1];
    }
}

// Output of FFT index j (column l of a tile, rows pitch_b bytes apart) without a
// fold: row (j - off) mod P when below rows_out.
__device__ __forceinline__ void xy_put(float2 *__restrict__ g, uint32_t pitch_b, int j, int l, int nc,
                                       const pvd_axis_d &ax, float2 val) {
    const int o = j >= ax.off ? j - ax.off : j + ax.n - ax.off;
    if (o < ax.N && l < nc) at(g, (uint32_t)o * pitch_b + (uint32_t)l * 8u) = val;
}

template <int B, bool INV>
__device__ void pass_y_2s(const pvd_vol_d &v, int tile, float2 *tw, float2 *s) {
    const pvd_axis_d &ay = v.ax[1];
    const int Ny = ay.N, Py = ay.n, H = v.H, Hp = v.Hp;
    const int lnl = v.ltky, TK = 1 << lnl;
    const int ntc = (H + TK - 1) >> lnl;
    const int x = tile / ntc;
    const int c0 = (tile - x * ntc) << lnl;
    const int nc = min(TK, H - c0);
    float2 *g = v.work + (long long)x * v.wplane + c0;    // uniform base of the tile
    const uint32_t pitch_b = (uint32_t)Hp * 8u;
    const bool direct = !INV || ay.foldn == 0;           // block-uniform
    load_twiddles(tw, ay);
    lds_barrier();
    xy_step1_global<B, INV>(s, lnl, tw, g, pitch_b, INV ? Py : Ny, nc);
    lds_barrier();
    for (int w = threadIdx.x; w < (16 << lnl); w += blockDim.x) {
        const int k1 = w >> lnl, l = w & (TK - 1);
        float2 u[B];
#pragma unroll
        for (int n2 = 0; n2 < B; ++n2) u[n2] = s[(k1 * B + n2) * TK + l];
        dft_any<B, INV>(u);
        if (!INV) {
            if (l < nc) {
                const uint32_t o0 = (uint32_t)k1 * pitch_b + (uint32_t)l * 8u;
#pragma unroll
                for (int k2 = 0; k2 < B; ++k2) at(g, o0 + (uint32_t)k2 * 16u * pitch_b) = u[k2];
            }
        } else if (direct) {
#pragma unroll
            for (int k2 = 0; k2 < B; ++k2) xy_put(g, pitch_b, k1 + 16 * k2, l, nc, ay, u[k2]);
        } else {
            // rows k1 B .. k1 B + B-1 are read only by this item: X[k1 + 16 k2] back to
            // row xy_row<B>(k1 + 16 k2) = k1 B + k2
#pragma unroll
            for (int k2 = 0; k2 < B; ++k2) s[(k1 * B + k2) * TK + l] = u[k2];
        }
    }
    if (!direct) {
        lds_barrier();
        store_col_tile_perm<B>(g, Hp, s, Ny, lnl, Py, ay.off, ay.foldn, nc);
    }
}

template <int B>
__device__ void pass_x_2s(const pvd_vol_d &v, int tile, float2 *tw, float2 *s) {
    const pvd_axis_d &ax = v.ax[0];
    const int Nx = ax.N, Px = ax.n, H = v.H, Hp = v.Hp;
    const int lnl = v.ltkx, TK = 1 << lnl;
    const int ntc = (H + TK - 1) >> lnl;
    const int u0 = tile / ntc;
    const int Py = v.ax[1].n;
    const int y = x_tile_row(u0, Py);
    const int c0 = (tile - u0 * ntc) << lnl;
    const int nc = min(TK, H - c0);
    const long long off0 = (long long)y * Hp + c0;
    float2 *g = v.work + off0;                                   // uniform bases of the tile
    const float2 *sp = v.spec + off0;
    const float *sr = reinterpret_cast<const float *>(v.spec) + (long long)min(y, Py - y) * Hp + c0;
    const uint32_t pitch_b = (uint32_t)v.wplane * 8u;             // x-plane step of the workspace
    const uint32_t kplane_b = (uint32_t)Py * (uint32_t)Hp * 8u;   // kx step of the general spectrum
    const uint32_t krow_b = (uint32_t)(Py / 2 + 1) * (uint32_t)Hp * 4u;   // kx step of the real octant
    const bool direct = ax.foldn == 0;                                  // block-uniform
    load_twiddles(tw, ax);
    lds_barrier();
    xy_step1_global<B, false>(s, lnl, tw, g, pitch_b, Nx, nc);
    lds_barrier();
    // forward DFT-B, times the spectrum, inverse DFT-B and twiddles, in registers
    for (int w = threadIdx.x; w < (16 << lnl); w += blockDim.x) {
        const int k1 = w >> lnl, l = w & (TK - 1);
        // spectrum of the frequencies k1 + 16 k2 (issued before the transform; masked
        // columns read the tile's first entry and are never stored)
        const bool ok = l < nc;
        float2 sv[B];
        if (v.sym) {
#pragma unroll
            for (int k2 = 0; k2 < B; ++k2) {
                const int k = k1 + 16 * k2;
                sv[k2] = make_float2(at(sr, ok ? (uint32_t)min(k, Px - k) * krow_b + (uint32_t)l * 4u : 0u), 0.f);
            }
        } else {
#pragma unroll
            for (int k2 = 0; k2 < B; ++k2)
                sv[k2] = at(sp, ok ? (uint32_t)(k1 + 16 * k2) * kplane_b + (uint32_t)l * 8u : 0u);
        }
        float2 u[B];
#pragma unroll
        for (int n2 = 0; n2 < B; ++n2) u[n2] = s[(k1 * B + n2) * TK + l];
        dft_any<B, false>(u);
#pragma unroll
        for (int k2 = 0; k2 < B; ++k2) u[k2] = v.sym ? cscale(u[k2], sv[k2].x) : cmul(u[k2], sv[k2]);
        dft_any<B, true>(u);
#pragma unroll
        for (int b = 1; b < B; ++b) {
            float2 t = tw[b * k1];
            t.y = -t.y;
            u[b] = cmul(u[b], t);
        }
#pragma unroll
        for (int b = 0; b < B; ++b) s[(k1 * B + b) * TK + l] = u[b];
    }
    lds_barrier();
    // inverse DFT-16 of item (b, l): x[b + B n1] in registers, stored row by row
    for (int w = threadIdx.x; w < (B << lnl); w += blockDim.x) {
        const int b = w >> lnl, l = w & (TK - 1);
        float2 u[16];
#pragma unroll
        for (int k1 = 0; k1 < 16; ++k1) u[k1] = s[(k1 * B + b) * TK + l];
        Dft<16, true>::run(u);
        if (direct) {
#pragma unroll
            for (int n1 = 0; n1 < 16; ++n1) xy_put(g, pitch_b, b + B * n1, l, nc, ax, u[n1]);
        } else {
            // rows b + B n1 are read only by this item: x[b + B n1] back to row b + B n1
#pragma unroll
            for (int n1 = 0; n1 < 16; ++n1) s[(n1 * B + b) * TK + l] = u[n1];
        }
    }
    if (!direct) {
        lds_barrier();
        store_col_tile(g, (long long)v.wplane, s, Nx, lnl, Px, ax.off, ax.foldn, nc);
    }
}

// LDS-resident form of the two-step, for B >= kXyLdsB (P = 560, 576): a 16-column tile
// of 35 or 36 B-rows has more step-1 items than the 512-thread block, so the fused
// form above would load in two rounds; here the tile is loaded into LDS by all
// threads, both steps run over LDS with one barrier each, and the result leaves
// through a row map.  xy_nat2perm: natural order in; step 1 (item (b, l)) a DFT-16
// over the rows b + B n1 and the twiddles W_P^(b k1), written back to the same rows;
// step 2 (item (k1, l)) a DFT-B over rows k1 B .. k1 B + B - 1.  X[k] is left in row
// xy_row<B>(k) = (k mod 16) B + k / 16.  xy_perm2nat: the same split run backwards,
// from that order to natural order.
constexpr int kXyLdsB = PVD_XY_LDS_B;
template <int B, bool INV>
__device__ __forceinline__ void xy_nat2perm(float2 *s, int lnl, const float2 *tw) {
    const int TK = 1 << lnl;
    for (int w = threadIdx.x; w < (B << lnl); w += blockDim.x) {
        const int b = w >> lnl, l = w & (TK - 1);
        float2 v[16];
#pragma unroll
        for (int n1 = 0; n1 < 16; ++n1) v[n1] = s[(b + B * n1) * TK + l];
        Dft<16, INV>::run(v);
#pragma unroll
        for (int k1 = 1; k1 < 16; ++k1) {
            float2 t = tw[b * k1];
            if (INV) t.y = -t.y;
            v[k1] = cmul(v[k1], t);
        }
#pragma unroll
        for (int k1 = 0; k1 < 16; ++k1) s[(b + B * k1) * TK + l] = v[k1];
    }
    lds_barrier();
    for (int w = threadIdx.x; w < (16 << lnl); w += blockDim.x) {
        const int k1 = w >> lnl, l = w & (TK - 1);
        float2 v[B];
#pragma unroll
        for (int n2 = 0; n2 < B; ++n2) v[n2] = s[(k1 * B + n2) * TK + l];
        dft_any<B, INV>(v);
#pragma unroll
        for (int k2 = 0; k2 < B; ++k2) s[(k1 * B + k2) * TK + l] = v[k2];
    }
    lds_barrier();
}

template <int B, bool INV>
__device__ __forceinline__ void xy_perm2nat(float2 *s, int lnl, const float2 *tw) {
    const int TK = 1 << lnl;
    for (int w = threadIdx.x; w < (16 << lnl); w += blockDim.x) {
        const int k1 = w >> lnl, l = w & (TK - 1);
        float2 v[B];
#pragma unroll
        for (int k2 = 0; k2 < B; ++k2) v[k2] = s[(k1 * B + k2) * TK + l];
        dft_any<B, INV>(v);
#pragma unroll
        for (int b = 1; b < B; ++b) {
            float2 t = tw[b * k1];
            if (INV) t.y = -t.y;
            v[b] = cmul(v[b], t);
        }
#pragma unroll
        for (int b = 0; b < B; ++b) s[(k1 * B + b) * TK + l] = v[b];
    }
    lds_barrier();
    for (int w = threadIdx.x; w < (B << lnl); w += blockDim.x) {
        const int b = w >> lnl, l = w & (TK - 1);
        float2 v[16];
#pragma unroll
        for (int k1 = 0; k1 < 16; ++k1) v[k1] = s[(k1 * B + b) * TK + l];
        Dft<16, INV>::run(v);
#pragma unroll
        for (int n1 = 0; n1 < 16; ++n1) s[(n1 * B + b) * TK + l] = v[n1];
    }
    lds_barrier();
}

template <int B, bool INV>
__device__ void pass_y_2s_lds(const pvd_vol_d &v, int tile, float2 *tw, float2 *s) {
    const pvd_axis_d &ay = v.ax[1];
    const int Ny = ay.N, Py = ay.n, H = v.H, Hp = v.Hp;
    const int lnl = v.ltky, TK = 1 << lnl;
    const int ntc = (H + TK - 1) >> lnl;
    const int x = tile / ntc;
    const int c0 = (tile - x * ntc) << lnl;
    const int nc = min(TK, H - c0);
    float2 *g = v.work + (long long)x * v.wplane + c0;
    load_twiddles(tw, ay);
    load_col_tile(s, g, Hp, Py, INV ? Py : Ny, lnl, nc);
    lds_barrier();
    xy_nat2perm<B, INV>(s, lnl, tw);
    if (!INV) store_col_tile_perm<B>(g, Hp, s, Py, lnl, Py, 0, 0, nc);
    else store_col_tile_perm<B>(g, Hp, s, Ny, lnl, Py, ay.off, ay.foldn, nc);
}

template <int B>
__device__ void pass_x_2s_lds(const pvd_vol_d &v, int tile, float2 *tw, float2 *s) {
    const pvd_axis_d &ax = v.ax[0];
    const int Nx = ax.N, Px = ax.n, H = v.H, Hp = v.Hp;
    const int lnl = v.ltkx, TK = 1 << lnl;
    const int ntc = (H + TK - 1) >> lnl;
    const int u0 = tile / ntc;
    const int y = x_tile_row(u0, v.ax[1].n);
    const int c0 = (tile - u0 * ntc) << lnl;
    const int nc = min(TK, H - c0);
    const long long plane = v.wplane;                 // work
    const long long splane = (long long)v.ax[1].n * Hp;   // spectrum
    const long long off0 = (long long)y * Hp + c0;
    float2 *g = v.work + off0;
    const PVD_GLOBAL float2 *sp = gmem(v.spec) + off0;
    // After the forward two-step, LDS row r = k1 B + k2 holds frequency kx = k1 + 16 k2.
    // Every spectrum load is issued at a clamped address (the tile's first entry where
    // masked) and masked afterwards.  (Loading the first batch before the forward
    // transform took the 35-point instance from 131 to 170 VGPRs and X from 471 to
    // 735 us on the 512x512x400 torso: it halved the workgroups per CU.)
    const int total = Px << lnl;
    const PVD_GLOBAL float *sr = (const PVD_GLOBAL float *)v.spec;
    const long long pyh = v.ax[1].n / 2 + 1;
    const long long row0 = (long long)min(y, v.ax[1].n - y) * Hp + c0;
    auto sym_index = [&](int e) -> long long {
        const int r = e >> lnl, l = e & (TK - 1);
        const int k = r / B + 16 * (r % B);
        return (e < total && l < nc) ? (long long)min(k, Px - k) * pyh * Hp + row0 + l : row0;
    };
    load_twiddles(tw, ax);
    load_col_tile(s, g, plane, Px, Nx, lnl, nc);
    lds_barrier();
    xy_nat2perm<B, false>(s, lnl, tw);
    if (v.sym) {
        for (int e0 = threadIdx.x; e0 < total; e0 += kBatch * blockDim.x) {
            float sv[kBatch];
#pragma unroll
            for (int u = 0; u < kBatch; ++u) sv[u] = sr[sym_index(e0 + u * blockDim.x)];
#pragma unroll
            for (int u = 0; u < kBatch; ++u) {
                const int e = e0 + u * blockDim.x;
                const bool ok = e < total && (e & (TK - 1)) < nc;
                if (e < total) s[e] = cscale(s[e], ok ? sv[u] : 0.f);
            }
        }
    } else {
        for (int e0 = threadIdx.x; e0 < total; e0 += kBatch * blockDim.x) {
            float2 sv[kBatch];
#pragma unroll
            for (int u = 0; u < kBatch; ++u) {
                const int e = e0 + u * blockDim.x;
                const int r = e >> lnl, l = e & (TK - 1);
                const int k = r / B + 16 * (r % B);
                sv[u] = sp[(e < total && l < nc) ? k * splane + l : 0];
            }
#pragma unroll
            for (int u = 0; u < kBatch; ++u) {
                const int e = e0 + u * blockDim.x;
                const bool ok = e < total && (e & (TK - 1)) < nc;
                if (e < total) s[e] = cmul(s[e], ok ? sv[u] : make_float2(0.f, 0.f));
            }
        }
    }
    lds_barrier();
    xy_perm2nat<B, true>(s, lnl, tw);
    store_col_tile(g, plane, s, Nx, lnl, Px, ax.off, ax.foldn, nc);
}

// ---------------------------------------------------------------------------------
// IZ: tile = 2*nl consecutive output lines (x, y); inverse two-for-one real FFT.
// IZ phases with TP threads per line pair: the half-spectrum load (both halves of
// each complex line) and the epilogue (roll / fold, scale * rho_w/rho, store).
template <int TP>
__device__ __forceinline__ void iz_load(const pvd_vol_d &v, float2 *s, const long long *lineoff, int nl, int nreal,
                                        int ld, int H, int Pz) {
    // Z_k = A_k + i B_k on [0, H); Z_{P-k} = conj(A_k) + i conj(B_k) for k in [1, P - H];
    // pairs past the last line get zeros
    constexpr int tpl = z_tpl<TP>();
    for (int w = threadIdx.x; w < nl * TP; w += blockDim.x) {
        const int q = w >> tpl, t = w & (TP - 1);
        const bool h0 = 2 * q < nreal, h1 = 2 * q + 1 < nreal;
        const PVD_GLOBAL float2 *g0 = gmem(v.work) + (h0 ? lineoff[2 * q] : 0);
        const PVD_GLOBAL float2 *g1 = gmem(v.work) + (h1 ? lineoff[2 * q + 1] : 0);
        for (int k0 = t; k0 < H; k0 += TP * kZBatch) {
            float2 A[kZBatch], B[kZBatch];
#pragma unroll
            for (int u = 0; u < kZBatch; ++u) {
                const int k = k0 + TP * u;
                A[u] = (k < H && h0) ? float2(g0[k]) : make_float2(0.f, 0.f);
                B[u] = (k < H && h1) ? float2(g1[k]) : make_float2(0.f, 0.f);
            }
#pragma unroll
            for (int u = 0; u < kZBatch; ++u) {
                const int k = k0 + TP * u;
                if (k < H) {
                    s[k * ld + q] = make_float2(A[u].x - B[u].y, A[u].y + B[u].x);
                    if (k >= 1 && k <= Pz - H) s[(Pz - k) * ld + q] = make_float2(A[u].x + B[u].y, B[u].x - A[u].y);
                }
            }
        }
    }
}

template <int TP>
__device__ __forceinline__ void iz_epilogue(const pvd_vol_d &v, const float2 *s, const pvd_axis_d &az, long long L0,
                                            int npair, int nreal, int ld, int Nz, int Pz) {
    constexpr int tpl = z_tpl<TP>();
    // epilogue: output o of real lines 2q and 2q+1 is the real and imaginary part of
    // index i = (o + off) mod P of complex line q (plus index (o + N + off) mod P when
    // o < foldn), times scale * rho_w/rho.  For o = t + TP j the row of i advances by
    // zrow_step per j and drops by zrow_wrap once i has wrapped past P.
    const int rs = zrow_step(az, TP) * ld, rw = zrow_wrap(az) * ld;
    const bool has_dens = v.dens != nullptr;
    const float sc = v.scale;
    for (int w = threadIdx.x; w < npair * TP; w += blockDim.x) {
        const int q = w >> tpl, t = w & (TP - 1);
        if (t >= Nz) continue;
        const bool h1 = 2 * q + 1 < nreal;
        const long long o0 = (L0 + 2 * q) * Nz;         // element 0 of real line 2q
        const int i0 = (t + az.off) % Pz;
        const float2 *p = s + zrow(az, i0) * ld + q;
        const int jw = (Pz - i0 + TP - 1) >> tpl;       // first j with i0 + TP j >= P
        for (int oo = t; oo < Nz; oo += TP * kZBatch) {
            const int jb = (oo - t) >> tpl;
            float2 d[kZBatch];
#pragma unroll
            for (int u = 0; u < kZBatch; ++u) {
                const int o = oo + TP * u;
                d[u].x = (has_dens && o < Nz) ? gmem(v.dens)[o0 + o] : 1.f;
                d[u].y = (has_dens && o < Nz && h1) ? gmem(v.dens)[o0 + Nz + o] : 1.f;
            }
            float2 vals[kZBatch];                   // all LDS reads of the batch before any store
#pragma unroll
            for (int u = 0; u < kZBatch; ++u) {
                const int j = jb + u;
                vals[u] = oo + TP * u < Nz ? p[j * rs - (j >= jw ? rw : 0)] : make_float2(0.f, 0.f);
            }
#pragma unroll
            for (int u = 0; u < kZBatch; ++u) {
                const int o = oo + TP * u;
                if (o < Nz) {
                    float2 val = vals[u];
                    if (o < az.foldn) {
                        const float2 f = s[zrow(az, (o + az.N + az.off) % Pz) * ld + q];
                        val.x += f.x;
                        val.y += f.y;
                    }
                    float f0 = sc, f1 = sc;
                    if (has_dens) {
                        f0 *= density_factor(d[u].x);
                        f1 *= density_factor(d[u].y);
                    }
                    gmem(v.out)[o0 + o] = val.x * f0;
                    if (h1) gmem(v.out)[o0 + Nz + o] = val.y * f1;
                }
            }
        }
    }
}

template <bool RZ, int TP>
__device__ void pass_iz(const pvd_vol_d &v, int tile, float2 *tw, float2 *s) {
    __shared__ long long lineoff[2 * PVD_TILE_MAX];
    const pvd_axis_d &az = v.ax[2];
    const int Nz = az.N, Pz = az.n, H = v.H, Hp = v.Hp;
    const int lnl = v.llz, nl = 1 << lnl, ld = nl + 1;
    const int Ny = v.ax[1].N, Py = v.ax[1].n;
    const long long nlines = (long long)v.ax[0].N * Ny;
    const long long L0 = (long long)tile * 2 * nl;
    const int nreal = (int)min((long long)(2 * nl), nlines - L0);
    const int npair = (nreal + 1) >> 1;
    load_twiddles(tw, az);
    tile_line_offsets(lineoff, L0, nreal, Ny, v.wplane, Hp);
    lds_barrier();
    iz_load<TP>(v, s, lineoff, nl, nreal, ld, H, Pz);
    lds_barrier();
    z_fft<true, RZ>(s, lnl, ld, az, tw);
    iz_epilogue<TP>(v, s, az, L0, npair, nreal, ld, Nz, Pz);
}

template <int PASS, bool RZ, int TP, int XB = 0>
__device__ __forceinline__ void run_pass(const pvd_vol_d &v, int tile) {
    extern __shared__ float2 smem[];
    const int axis = (PASS == PVD_PASS_X) ? 0 : (PASS == PVD_PASS_FY || PASS == PVD_PASS_IY) ? 1 : 2;
    float2 *tw = smem;
    float2 *s = smem + ((v.ax[axis].n + 7) & ~7);
    if constexpr (XB >= kXyLdsB) {    // register two-step over an LDS tile, P = 16 * XB
        if (PASS == PVD_PASS_FY) pass_y_2s_lds<XB, false>(v, tile, tw, s);
        else if (PASS == PVD_PASS_IY) pass_y_2s_lds<XB, true>(v, tile, tw, s);
        else if (PASS == PVD_PASS_X) pass_x_2s_lds<XB>(v, tile, tw, s);
        return;
    } else if constexpr (XB > 0) {    // fused register two-step, P = 16 * XB
        if (PASS == PVD_PASS_FY) pass_y_2s<XB, false>(v, tile, tw, s);
        else if (PASS == PVD_PASS_IY) pass_y_2s<XB, true>(v, tile, tw, s);
        else if (PASS == PVD_PASS_X) pass_x_2s<XB>(v, tile, tw, s);
        return;
    }
    if (PASS == PVD_PASS_FZ) pass_fz<RZ, TP>(v, tile, tw, s);
    else if (PASS == PVD_PASS_FY) {
        if (sq256_ok(v.ax[1], v.ltky)) pass_y_sq<false>(v, tile, tw, s);
        else pass_y<false>(v, tile, tw, s);
    } else if (PASS == PVD_PASS_X) {
        if (sq256_ok(v.ax[0], v.ltkx)) pass_x_sq(v, tile, tw, s);
        else pass_x(v, tile, tw, s);
    } else if (PASS == PVD_PASS_IY) {
        if (sq256_ok(v.ax[1], v.ltky)) pass_y_sq<true>(v, tile, tw, s);
        else pass_y<true>(v, tile, tw, s);
    } else pass_iz<RZ, TP>(v, tile, tw, s);
}

}  // namespace

// ---------------------------------------------------------------------------------
// Kernel templates.  Each specialization is instantiated (launched) in exactly one
// source file: the generic y/x kernels in pvd_fft.hip, the z kernels in
// pvd_fft_z.hip, the two-step kernels in pvd_fft_xy1/2/3.hip.

namespace {
// Workgroups go to the 8 XCDs round-robin and the X grid is a multiple of 8:
// block b takes tile (b % 8) * (grid / 8) + b / 8, so each XCD walks a contiguous
// range of tiles and the tiles sharing spectrum lines meet in its L2.
__device__ __forceinline__ int x_band_tile() {
    return (int)(blockIdx.x & 7) * (int)(gridDim.x >> 3) + (int)(blockIdx.x >> 3);
}

// FY and IZ walk their tiles last to first: FZ and IY write the workspace in x-plane
// order, so the planes they wrote last (those still in the last-level cache when the
// workspace is larger than it) are read first.  (512x512x400 torso: FY 178 -> 167 us,
// IZ 287 -> 272 us; the 256x256x200 CT dose 113.5 -> 112.3 us; same bits.)
template <int PASS> constexpr bool kReversed = PASS == PVD_PASS_FY || PASS == PVD_PASS_IZ;

// Batch launch: the member whose tile range holds grid tile t (base ascending),
// then that volume's own tile.
template <int PASS, bool RZ, int TP, int XB>
__device__ __forceinline__ void run_batch(const pvd_batch_args &a) {
    int t = blockIdx.x;
    if (PASS == PVD_PASS_X) {
        t = x_band_tile();
        if (t >= a.ntiles) return;
    }
    if (kReversed<PASS>) t = a.ntiles - 1 - t;
    int lo = 0, hi = a.nmem - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (a.base[mid] <= t) lo = mid; else hi = mid - 1;
    }
    const pvd_vol_d &v = a.vols[a.vol[lo]];
    const int tile = t - a.base[lo];
    if (tile < v.tiles[PASS]) run_pass<PASS, RZ, TP, XB>(v, tile);
}

// The y/x register two-step for P = 16 * B has its own kernels, so that their
// register budget does not touch the other instances.  Waves per SIMD (exactly; the
// VGPR cap is 512 / waves): FY/IY hold 16 loads or a B-point line per item and keep
// four up to B = 27, three up to 32 and two above; X also holds the B spectrum values
// of its line through the forward transform and keeps four up to B = 21, three up to
// 30 and two above.  A 512-thread block of a 512- or 560-point 16-column tile (over
// 64 KB of LDS) is one block per CU, two waves per SIMD, either way.
// The LDS-resident instances (B >= 35) hold only one B-point line per item and ask for
// at least four, the batch X at least three (its descriptor lookup would spill at 128
// VGPRs); with an exact count there the compiler spilled 2-10 VGPRs.
template <int PASS, int B, bool BATCH> constexpr int xy_waves() {
    if (B >= kXyLdsB) return (BATCH && PASS == PVD_PASS_X) ? 3 : 4;
    return PASS == PVD_PASS_X ? (B <= 21 ? 4 : B <= 30 ? 3 : 2) : (B <= 27 ? 4 : B <= 32 ? 3 : 2);
}
template <int PASS, int B, bool BATCH> constexpr int xy_waves_max() {
    return B >= kXyLdsB ? 8 : xy_waves<PASS, B, BATCH>();
}
}  // namespace

// single volume: descriptor by value (kernarg segment, scalar loads).  Only the z
// passes use the template arguments: RZ for a plan with a register z transform,
// TP threads per line pair in their I/O phases (the host picks both; one kernel
// per combination, as branches between them inflate the register allocation).
template <int PASS, bool RZ = false, int TP = 16>
__global__ __launch_bounds__(1024) PVD_FFT_OCC void pvd_fft_pass(pvd_vol_d v) {
    int tile = blockIdx.x;
    if (PASS == PVD_PASS_X) {
        tile = x_band_tile();
        if (tile >= v.tiles[PVD_PASS_X]) return;
    }
    if (kReversed<PASS>) tile = v.tiles[PASS] - 1 - tile;
    run_pass<PASS, RZ, TP>(v, tile);
}

// batch: the same instance for the members of one batch launch
template <int PASS, bool RZ = false, int TP = 16>
__global__ __launch_bounds__(1024) PVD_FFT_OCC void pvd_fft_pass_batch(pvd_batch_args a) {
    run_batch<PASS, RZ, TP, 0>(a);
}

template <int PASS, int B>
__global__ __launch_bounds__(kXyThreads)
__attribute__((amdgpu_waves_per_eu(xy_waves<PASS, B, false>(), xy_waves_max<PASS, B, false>())))
void pvd_fft_pass_xy(pvd_vol_d v) {
    int tile = blockIdx.x;
    if (PASS == PVD_PASS_X) {
        tile = x_band_tile();
        if (tile >= v.tiles[PVD_PASS_X]) return;
    }
    if (kReversed<PASS>) tile = v.tiles[PASS] - 1 - tile;
    run_pass<PASS, false, 16, B>(v, tile);
}

template <int PASS, int B>
__global__ __launch_bounds__(kXyThreads)
__attribute__((amdgpu_waves_per_eu(xy_waves<PASS, B, true>(), xy_waves_max<PASS, B, true>())))
void pvd_fft_pass_xy_batch(pvd_batch_args a) {
    run_batch<PASS, false, 16, B>(a);
}
