// 3-D residual CNN kernels on the gfx950 matrix cores (BF16 in, FP32 accumulate).
//
// Layout: activations channels-last BF16 [n0][n1][n2][C] (n2 % 32 == 0), weights
// BF16 [27][32 cout][cin] with tap = (dx+1)*9 + (dy+1)*3 + (dz+1).
//
// conv_fwd  implicit GEMM per output row of 32 voxels along z:
//           D[32 voxels][32 cout] += A[32 voxels][16 cin] * B[16 cin][32 cout]
//           with v_mfma_f32_32x32x16_bf16, K-loop over 27 taps x cin/16.
//           Persistent workgroups (8 waves) keep all weights in LDS and stage a
//           halo of 2x4 output rows; voxel/cout strides are padded by 16 bytes so
//           the 16-byte fragment reads of 8 consecutive lanes hit distinct banks.
// wgrad     D[32 cout][32 cin] per tap += A[32 cout][16 voxels] * B[16 voxels][32 cin]
//           from channel-major (transposed) LDS copies; the input halo is stored
//           three times, pre-shifted by dz = -1, 0, +1, so every fragment read is a
//           16-byte aligned load.  Each wave owns a fixed subset of taps; per-
//           workgroup partial sums are reduced in a fixed order (deterministic).
#include <algorithm>

#include "pvd_internal.h"

#include "../../include/pvd_learned.h"

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

namespace {

// row of accumulator i held by `lane` in the 32x32 MFMA output; its column is lane % 32
__device__ __forceinline__ int mfma_row(int i, int lane) { return ((i >> 2) << 3) + ((lane >> 5) << 2) + (i & 3); }

__device__ __forceinline__ float bf2f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

// round to nearest even; NaN stays NaN
__device__ __forceinline__ uint16_t f2bf(float f) {
    uint32_t u = __float_as_uint(f);
    if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((u >> 16) | 0x40);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}

__device__ __forceinline__ bf16x8 lds_frag(const char *p) {
    const uint4 v = *reinterpret_cast<const uint4 *>(p);
    return __builtin_bit_cast(bf16x8, v);
}

struct ConvArgs {
    const uint16_t *x;
    const uint16_t *w;
    const float *bias;
    const uint16_t *resid;
    uint16_t *y;
    uint16_t *act;
    float *y32;
    int n0, n1, n2;
    const int *list;          // blocks to compute (NULL: every block)
    const int *list_n;        // number of entries of list (device)
};

constexpr int FWD_BX = 2, FWD_BY = 4, FWD_THREADS = 512;
constexpr int HX = FWD_BX + 2, HY = FWD_BY + 2, HZ = 34;

// per-wave staging of one 32 x 32 BF16 output tile (32 voxels x 32 channels)
constexpr int FWD_STAGE = 32 * 32 * 2;

template <int CIN>
constexpr size_t fwd_lds_bytes() {
    return (size_t)27 * 32 * (CIN * 2 + 16) + (size_t)HX * HY * HZ * (CIN * 2 + 16) +
           (size_t)(FWD_THREADS / 64) * FWD_STAGE;
}

template <int CIN, int EPI>
__global__ __launch_bounds__(FWD_THREADS) void cnn_conv_fwd(ConvArgs a) {
    extern __shared__ uint4 lds_u4[];
    char *lds = reinterpret_cast<char *>(lds_u4);
    constexpr int VS = CIN * 2 + 16;   // bytes per halo voxel
    constexpr int WS = CIN * 2 + 16;   // bytes per (tap, cout) weight row
    constexpr int C8 = CIN / 8;        // 16-byte chunks per voxel
    char *wl = lds;
    char *hl = lds + 27 * 32 * WS;
    const int tid = threadIdx.x;
    char *stg = hl + HX * HY * HZ * VS + (tid >> 6) * FWD_STAGE;   // this wave's output staging
    for (int i = tid; i < 27 * 32 * C8; i += FWD_THREADS) {
        const int row = i / C8, c8 = i - row * C8;
        *reinterpret_cast<uint4 *>(wl + row * WS + c8 * 16) = reinterpret_cast<const uint4 *>(a.w)[i];
    }
    const int nbx = (a.n0 + FWD_BX - 1) / FWD_BX, nby = (a.n1 + FWD_BY - 1) / FWD_BY, nbz = a.n2 >> 5;
    // blocks to compute: the list's entries (a region of the volume) or all of them
    const int total = a.list ? *a.list_n : nbx * nby * nbz;
    auto block_of = [&](int i) { return a.list ? a.list[i] : i; };
    const int wave = tid >> 6, lane = tid & 63;
    const int wx = wave / FWD_BY, wy = wave % FWD_BY;
    const int m = lane & 31, kh = lane >> 5;
    // this lane's output-channel bias, loaded once: a global load in the epilogue
    // would wait (in-order counter) for the next block's halo prefetch
    const float bco = a.bias ? a.bias[m] : 0.0f;
    // The halo of the next block is fetched into registers while this block's
    // MFMA loop runs (all loads in flight at once), then written to LDS.
    constexpr int HALO = HX * HY * HZ * C8;
    constexpr int PER_T = (HALO + FWD_THREADS - 1) / FWD_THREADS;
    uint4 pf[PER_T];
    auto fetch = [&](int b) {
        const int bz = b % nbz, by = (b / nbz) % nby, bx = b / (nbz * nby);
        const int x0 = bx * FWD_BX, y0 = by * FWD_BY, z0 = bz * 32;
#pragma unroll
        for (int u = 0; u < PER_T; ++u) {
            const int i = tid + u * FWD_THREADS;
            const int c8 = i % C8, vox = i / C8;
            const int hz = vox % HZ, hy = (vox / HZ) % HY, hx = vox / (HZ * HY);
            const int gx = x0 + hx - 1, gy = y0 + hy - 1, gz = z0 + hz - 1;
            pf[u] = make_uint4(0u, 0u, 0u, 0u);
            if (i < HALO && gx >= 0 && gx < a.n0 && gy >= 0 && gy < a.n1 && gz >= 0 && gz < a.n2)
                pf[u] = reinterpret_cast<const uint4 *>(a.x + (((long long)gx * a.n1 + gy) * a.n2 + gz) * CIN)[c8];
        }
    };
    auto put_halo = [&]() {
#pragma unroll
        for (int u = 0; u < PER_T; ++u) {
            const int i = tid + u * FWD_THREADS;
            if (i < HALO) *reinterpret_cast<uint4 *>(hl + (i / C8) * VS + (i % C8) * 16) = pf[u];
        }
    };
    // Per block: MFMA loop, the skip values out of the halo, barrier, the next block's
    // halo into LDS, then the epilogue, whose output stores overlap the next block's
    // MFMAs.  (The wait for the prefetched halo registers counts every earlier vector
    // memory operation; the latest stores are then a block old.)
    if ((int)blockIdx.x < total) {
        fetch(block_of(blockIdx.x));
        put_halo();
    }
    for (int bi = blockIdx.x; bi < total; bi += gridDim.x) {
        const int b = block_of(bi);
        const int bz = b % nbz, by = (b / nbz) % nby, bx = b / (nbz * nby);
        const int x0 = bx * FWD_BX, y0 = by * FWD_BY, z0 = bz * 32;
        // LDS-only barriers: the previous block's output stores stay in flight (a
        // full __syncthreads would wait for them)
        lds_barrier();     // weights and this block's halo visible
        const int ox = x0 + wx, oy = y0 + wy;
        const bool live = ox < a.n0 && oy < a.n1;    // wave-uniform
        const long long vrow = ((long long)ox * a.n1 + oy) * a.n2 + z0;
        // skip term of a residual layer: a layer whose skip input is its own input
        // (resid == x, 32 channels) reads it from the centre of the halo in LDS;
        // any other tensor is loaded here, before the next halo, so the epilogue
        // waits for these loads only
        const bool resid_lds = EPI == PVD_CNN_EPI_RESID_RELU && CIN == 32 && a.resid == a.x;
        uint16_t rs[16];
        if (EPI == PVD_CNN_EPI_RESID_RELU && !resid_lds && live) {
#pragma unroll
            for (int i = 0; i < 16; ++i) rs[i] = a.resid[(vrow + mfma_row(i, lane)) * 32 + m];
        }
        if (bi + (int)gridDim.x < total) fetch(block_of(bi + gridDim.x));
        // one accumulation chain in tap order; the fragments of tap t + 1 are read
        // before the MFMAs of tap t (two register sets), so each LDS round trip
        // overlaps the previous tap's MFMAs
        f32x16 acc = {};
        if (live) {
            constexpr int KC = CIN / 16;
            bf16x8 fa[2][KC], fb[2][KC];
            auto read_tap = [&](int tap, bf16x8 *ta, bf16x8 *tb) {
                const int dx = tap / 9 - 1, dy = (tap / 3) % 3 - 1, dz = tap % 3 - 1;
                const char *hrow = hl + (((wx + dx + 1) * HY + (wy + dy + 1)) * HZ + (m + dz + 1)) * VS + kh * 16;
                const char *wrow = wl + (tap * 32 + m) * WS + kh * 16;
#pragma unroll
                for (int cc = 0; cc < KC; ++cc) {
                    ta[cc] = lds_frag(hrow + cc * 32);
                    tb[cc] = lds_frag(wrow + cc * 32);
                }
            };
            read_tap(0, fa[0], fb[0]);
#pragma unroll
            for (int tap = 0; tap < 27; ++tap) {
                const int cur = tap & 1;
                if (tap + 1 < 27) read_tap(tap + 1, fa[cur ^ 1], fb[cur ^ 1]);
                __builtin_amdgcn_sched_barrier(0);
#pragma unroll
                for (int cc = 0; cc < KC; ++cc)
                    acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(fa[cur][cc], fb[cur][cc], acc, 0, 0, 0);
                __builtin_amdgcn_sched_barrier(0);
            }
            if (resid_lds) {   // halo voxel (wx + 1, wy + 1, z + 1), channel m
#pragma unroll
                for (int i = 0; i < 16; ++i)
                    rs[i] = *reinterpret_cast<const uint16_t *>(
                        hl + (((wx + 1) * HY + (wy + 1)) * HZ + mfma_row(i, lane) + 1) * VS + m * 2);
            }
        }
        lds_barrier();     // every wave finished with this block's halo
        // unconditional, so that no path leaves the prefetch loads outstanding (they
        // would be waited for, with the epilogue's stores, at the next fetch); after
        // the last block it rewrites values that nothing reads
        put_halo();
        if (!live) continue;
        const int co = m;
        if (EPI == PVD_CNN_EPI_RELU || EPI == PVD_CNN_EPI_RESID_RELU) {
            // The wave's output tile (voxels z0..z0+31 of one column, 32 channels) is
            // 2 KB contiguous in y: stage it in the wave's LDS slot as [row][channel],
            // then write it as two 16-byte stores per lane (in place of 16 two-byte
            // stores).  One wave reads only what it wrote, in program order.
            uint16_t *sy = reinterpret_cast<uint16_t *>(stg);
            const uint4 *src = reinterpret_cast<const uint4 *>(stg);
            if (a.act) {                         // training: the activations first
#pragma unroll
                for (int i = 0; i < 16; ++i) sy[mfma_row(i, lane) * 32 + co] = f2bf(fmaxf(acc[i] + bco, 0.0f));
                uint4 *dst = reinterpret_cast<uint4 *>(a.act + vrow * 32);
                dst[lane] = src[lane];
                dst[lane + 64] = src[lane + 64];
            }
#pragma unroll
            for (int i = 0; i < 16; ++i) {
                const float av = fmaxf(acc[i] + bco, 0.0f);
                uint16_t h;
                if (EPI == PVD_CNN_EPI_RELU) {
                    h = f2bf(av);
                } else {
                    h = f2bf(bf2f(rs[i]) + av);
                }
                sy[mfma_row(i, lane) * 32 + co] = h;
            }
            uint4 *dst = reinterpret_cast<uint4 *>(a.y + vrow * 32);
            dst[lane] = src[lane];
            dst[lane + 64] = src[lane + 64];
        } else {
#pragma unroll
            for (int i = 0; i < 16; ++i) {
                const long long v = vrow + mfma_row(i, lane);
                if (EPI == PVD_CNN_EPI_LINEAR_F32) {
                    if (co == 0) a.y32[v] = acc[i] + bco;
                } else {
                    a.y32[v * 32 + co] += acc[i];
                }
            }
        }
    }
}

// ---------------------------------------------------------------------------------
// Input layer with four live channels (the packer's activity, density and plain dose,
// and a zero) in a 16-channel BF16 input whose channels 4..15 are zero, ReLU epilogue.
// K of each MFMA packs 4 taps x 4 channels: 7 MFMAs per 32-voxel row (taps 24..27, the
// last zero) in place of 27 with 13 of 16 channels zero.  The packed weights
// wp[c][co][k] = w[4c + k/4][co][k%4] are built in LDS from the [27][32][16] layout;
// the halo holds the four channels (8 bytes a voxel).  Blocks, region lists, the
// next-halo prefetch and the epilogue are cnn_conv_fwd's.
constexpr int IN_CHUNKS = 7;
constexpr int IN_WS = 16 * 2 + 16;                              // bytes per (chunk, cout) row
template <int BX, int BY>
constexpr int in_halo() { return (BX + 2) * (BY + 2) * HZ; }
template <int BX, int BY>
constexpr size_t in_lds_bytes() {
    return (size_t)IN_CHUNKS * 32 * IN_WS + (size_t)in_halo<BX, BY>() * 8 + (size_t)(FWD_THREADS / 64) * FWD_STAGE;
}

// BX x BY rows per block (2 x 4: the blocks of region lists; 4 x 4 for whole volumes,
// less halo per output); each of the 8 waves computes BX * BY / 8 rows.
template <int BX, int BY>
__global__ __launch_bounds__(FWD_THREADS) void cnn_conv_in(ConvArgs a) {
    constexpr int IN_HALO = in_halo<BX, BY>(), HYL = BY + 2, NR = BX * BY / 8;
    static_assert(NR * 8 == BX * BY, "whole rows per wave");
    extern __shared__ uint4 lds_u4[];
    char *lds = reinterpret_cast<char *>(lds_u4);
    char *wl = lds;
    char *hl = lds + IN_CHUNKS * 32 * IN_WS;
    const int tid = threadIdx.x;
    char *stg = hl + IN_HALO * 8 + (tid >> 6) * FWD_STAGE;
    for (int i = tid; i < IN_CHUNKS * 32 * 16; i += FWD_THREADS) {
        const int c = i >> 9, co = (i >> 4) & 31, k = i & 15;
        const int tap = 4 * c + (k >> 2), ch = k & 3;
        const uint16_t v = tap < 27 ? a.w[((size_t)tap * 32 + co) * 16 + ch] : (uint16_t)0;
        *reinterpret_cast<uint16_t *>(wl + (c * 32 + co) * IN_WS + k * 2) = v;
    }
    const int nbx = (a.n0 + BX - 1) / BX, nby = (a.n1 + BY - 1) / BY, nbz = a.n2 >> 5;
    const int total = a.list ? *a.list_n : nbx * nby * nbz;
    auto block_of = [&](int i) { return a.list ? a.list[i] : i; };
    const int wave = tid >> 6, lane = tid & 63;
    const int m = lane & 31, kh = lane >> 5;
    const float bco = a.bias ? a.bias[m] : 0.0f;
    constexpr int PER_T = (IN_HALO + FWD_THREADS - 1) / FWD_THREADS;
    uint2 pf[PER_T];
    auto fetch = [&](int b) {
        const int bz = b % nbz, by = (b / nbz) % nby, bx = b / (nbz * nby);
        const int x0 = bx * BX, y0 = by * BY, z0 = bz * 32;
#pragma unroll
        for (int u = 0; u < PER_T; ++u) {
            const int vox = tid + u * FWD_THREADS;
            const int hz = vox % HZ, hy = (vox / HZ) % HYL, hx = vox / (HZ * HYL);
            const int gx = x0 + hx - 1, gy = y0 + hy - 1, gz = z0 + hz - 1;
            pf[u] = make_uint2(0u, 0u);
            if (vox < IN_HALO && gx >= 0 && gx < a.n0 && gy >= 0 && gy < a.n1 && gz >= 0 && gz < a.n2)
                pf[u] = *reinterpret_cast<const uint2 *>(a.x + (((long long)gx * a.n1 + gy) * a.n2 + gz) * 16);
        }
    };
    if ((int)blockIdx.x < total) fetch(block_of(blockIdx.x));
    for (int bi = blockIdx.x; bi < total; bi += gridDim.x) {
        const int b = block_of(bi);
        const int bz = b % nbz, by = (b / nbz) % nby, bx = b / (nbz * nby);
        const int x0 = bx * BX, y0 = by * BY, z0 = bz * 32;
        lds_barrier();     // weights visible / previous block finished with the halo
#pragma unroll
        for (int u = 0; u < PER_T; ++u) {
            const int vox = tid + u * FWD_THREADS;
            if (vox < IN_HALO) *reinterpret_cast<uint2 *>(hl + vox * 8) = pf[u];
        }
        lds_barrier();
        if (bi + (int)gridDim.x < total) fetch(block_of(bi + gridDim.x));
#pragma unroll
        for (int r = 0; r < NR; ++r) {
        const int wx = (wave + 8 * r) / BY, wy = (wave + 8 * r) % BY;
        const int ox = x0 + wx, oy = y0 + wy;
        const bool live = ox < a.n0 && oy < a.n1;    // wave-uniform
        const long long vrow = ((long long)ox * a.n1 + oy) * a.n2 + z0;
        if (!live) continue;
        f32x16 acc0 = {}, acc1 = {};
#pragma unroll
        for (int c = 0; c < IN_CHUNKS; ++c) {
            uint2 h[2];
#pragma unroll
            for (int j = 0; j < 2; ++j) {
                const int tap = 4 * c + 2 * kh + j;
                const int dx = tap / 9 - 1, dy = (tap / 3) % 3 - 1, dz = tap % 3 - 1;
                h[j] = tap < 27 ? *reinterpret_cast<const uint2 *>(
                                      hl + (((wx + dx + 1) * HYL + (wy + dy + 1)) * HZ + (m + dz + 1)) * 8)
                                : make_uint2(0u, 0u);
            }
            const bf16x8 fa = __builtin_bit_cast(bf16x8, make_uint4(h[0].x, h[0].y, h[1].x, h[1].y));
            const bf16x8 fb = lds_frag(wl + (c * 32 + m) * IN_WS + kh * 16);
            if ((c & 1) == 0) acc0 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(fa, fb, acc0, 0, 0, 0);
            else acc1 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(fa, fb, acc1, 0, 0, 0);
        }
        const f32x16 acc = acc0 + acc1;
        // the output tile (32 voxels x 32 channels) staged in the wave's LDS slot as
        // [row][channel], then written as two 16-byte stores per lane
        uint16_t *sy = reinterpret_cast<uint16_t *>(stg);
        const uint4 *src = reinterpret_cast<const uint4 *>(stg);
#pragma unroll
        for (int i = 0; i < 16; ++i) sy[mfma_row(i, lane) * 32 + m] = f2bf(fmaxf(acc[i] + bco, 0.0f));
        uint4 *dst = reinterpret_cast<uint4 *>(a.y + vrow * 32);
        dst[lane] = src[lane];
        dst[lane + 64] = src[lane + 64];
        }
    }
}

// ---------------------------------------------------------------------------------
// Output layer (32 channels in, channel 0 out, FP32, + bias[0]), by taps:
//   Y[u][t] = sum_ci W[t][0][ci] x[u][ci]   for every voxel u of the block's halo, t < 27
//   y[v]    = sum_t Y[v + tap(t)][t] + bias[0]
// The first is a GEMM of the halo (32 voxels per MFMA row block, 32 channels) with the
// 32 x 27 weight columns (taps padded to 32), so the matrix cores do 27 useful columns of
// 32 instead of 1 of the 32 output channels a full layer computes; the second is a
// 27-term sum per output voxel from LDS, in tap order.  Blocks and the region list are
// those of cnn_conv_fwd (2 x 4 rows of 32 voxels along z).
constexpr int OUT_YS = 28;                                     // floats per halo voxel of Y
template <int BX, int BY>
constexpr size_t out_lds_bytes() { return (size_t)(BX + 2) * (BY + 2) * HZ * OUT_YS * 4; }

// BX x BY rows per block: 2 x 4 (816 halo voxels, 26 row blocks of 32) for region lists,
// 4 x 4 (1224, 39) for whole volumes; one thread per output voxel in the sums.
template <int BX, int BY>
__global__ __launch_bounds__(FWD_THREADS) void cnn_conv_out(ConvArgs a) {
    constexpr int HYL = BY + 2;
    constexpr int OUT_HALO = (BX + 2) * HYL * HZ;
    constexpr int OUT_GROUPS = (OUT_HALO + 31) / 32;
    static_assert(BX * BY * 32 <= FWD_THREADS, "one thread per output voxel");
    extern __shared__ uint4 lds_u4[];
    float *ys = reinterpret_cast<float *>(lds_u4);
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const int m = lane & 31, kh = lane >> 5;
    constexpr int NW = FWD_THREADS / 64;
    constexpr int GPW = (OUT_GROUPS + NW - 1) / NW;            // row blocks per wave (at most)
    // B: column n = tap (zero past 26), rows = input channels; both K halves of 16
    bf16x8 bw[2];
#pragma unroll
    for (int cc = 0; cc < 2; ++cc) {
        uint4 v = make_uint4(0u, 0u, 0u, 0u);
        if (m < 27) v = *reinterpret_cast<const uint4 *>(a.w + ((size_t)m * 32 * 32 + cc * 16 + kh * 8));
        bw[cc] = __builtin_bit_cast(bf16x8, v);
    }
    const float b0 = a.bias ? a.bias[0] : 0.0f;
    const int nbx = (a.n0 + BX - 1) / BX, nby = (a.n1 + BY - 1) / BY, nbz = a.n2 >> 5;
    const int total = a.list ? *a.list_n : nbx * nby * nbz;
    for (int bi = blockIdx.x; bi < total; bi += gridDim.x) {
        const int b = a.list ? a.list[bi] : bi;
        const int bz = b % nbz, by = (b / nbz) % nby, bx = b / (nbz * nby);
        const int x0 = bx * BX, y0 = by * BY, z0 = bz * 32;
        // A fragments of this wave's row blocks, all loads issued before the MFMAs
        uint4 av[GPW][2];
#pragma unroll
        for (int j = 0; j < GPW; ++j) {
            const int u = (wave + j * NW) * 32 + m;
            const int hz = u % HZ, hy = (u / HZ) % HYL, hx = u / (HZ * HYL);
            const int gx = x0 + hx - 1, gy = y0 + hy - 1, gz = z0 + hz - 1;
            const bool in = u < OUT_HALO && gx >= 0 && gx < a.n0 && gy >= 0 && gy < a.n1 && gz >= 0 && gz < a.n2;
            const uint4 *src = reinterpret_cast<const uint4 *>(
                a.x + (in ? (((long long)gx * a.n1 + gy) * a.n2 + gz) * 32 : 0));
#pragma unroll
            for (int cc = 0; cc < 2; ++cc) {
                const uint4 v = src[cc * 2 + kh];
                av[j][cc] = in ? v : make_uint4(0u, 0u, 0u, 0u);
            }
        }
        lds_barrier();      // the previous block's sums have read Y
#pragma unroll
        for (int j = 0; j < GPW; ++j) {
            const int g = wave + j * NW;
            if (g >= OUT_GROUPS) continue;           // wave-uniform
            f32x16 acc = {};
            acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(bf16x8, av[j][0]), bw[0], acc, 0, 0, 0);
            acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(bf16x8, av[j][1]), bw[1], acc, 0, 0, 0);
            if (m < 27) {
#pragma unroll
                for (int i = 0; i < 16; ++i) {
                    const int u = g * 32 + mfma_row(i, lane);
                    if (u < OUT_HALO) ys[u * OUT_YS + m] = acc[i];
                }
            }
        }
        lds_barrier();
        if (tid < BX * BY * 32) {
            const int ox = tid / (BY * 32), oy = (tid >> 5) % BY, z = tid & 31;
            if (x0 + ox < a.n0 && y0 + oy < a.n1) {
                float sum = 0.0f;
                for (int t = 0; t < 27; ++t) {
                    const int dx = t / 9 - 1, dy = (t / 3) % 3 - 1, dz = t % 3 - 1;
                    sum += ys[(((ox + dx + 1) * HYL + (oy + dy + 1)) * HZ + (z + dz + 1)) * OUT_YS + t];
                }
                a.y32[((long long)(x0 + ox) * a.n1 + (y0 + oy)) * a.n2 + z0 + z] = sum + b0;
            }
        }
    }
}

// ---------------------------------------------------------------------------------
struct WgradArgs {
    const uint16_t *x;
    const uint16_t *dz;
    float *work;
    int n0, n1, n2;
};

constexpr int WG_BX = 2, WG_BY = 2, WG_THREADS = 512;
constexpr int WHX = WG_BX + 2, WHY = WG_BY + 2, WROWS = WHX * WHY;   // 16 halo rows
constexpr int XCS = WROWS * 32 * 2 + 16;     // bytes per cin row of one shifted copy (1040)
constexpr int XCOPY = 32 * XCS;              // one dz-shifted copy
constexpr int DCS = WG_BX * WG_BY * 32 * 2 + 16;   // bytes per cout row of dz (272)
constexpr size_t WG_LDS = 3 * (size_t)XCOPY + 32 * (size_t)DCS;
constexpr int WG_MAX_BLOCKS = 256;

__device__ __forceinline__ void st_bf16(char *p, uint16_t v) { *reinterpret_cast<uint16_t *>(p) = v; }

template <int CIN>
__global__ __launch_bounds__(WG_THREADS) void cnn_wgrad(WgradArgs a) {
    extern __shared__ uint4 lds_u4[];
    char *lds = reinterpret_cast<char *>(lds_u4);
    char *xs = lds;                  // [3 dz][32 ci][16 rows][32 z]
    char *ds = lds + 3 * XCOPY;      // [32 co][4 rows][32 z]
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const int m = lane & 31, kh = lane >> 5;
    const int nbx = (a.n0 + WG_BX - 1) / WG_BX, nby = (a.n1 + WG_BY - 1) / WG_BY, nbz = a.n2 >> 5;
    const int total = nbx * nby * nbz;
    f32x16 acc[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) acc[j] = f32x16{};
    for (int b = blockIdx.x; b < total; b += gridDim.x) {
        const int bz = b % nbz, by = (b / nbz) % nby, bx = b / (nbz * nby);
        const int x0 = bx * WG_BX, y0 = by * WG_BY, z0 = bz * 32;
        __syncthreads();
        // input halo rows, z in [z0-1, z0+33), into the three shifted channel-major copies
        for (int i = tid; i < WROWS * 34 * 4; i += WG_THREADS) {
            const int c8 = i & 3, vox = i >> 2;           // 4 chunks of 8 channels (cin 32; 16 pads zeros)
            const int hz = vox % 34, hrow = vox / 34;
            const int hx = hrow / WHY, hy = hrow % WHY;
            const int gx = x0 + hx - 1, gy = y0 + hy - 1, gz = z0 + hz - 1;
            uint4 v = make_uint4(0u, 0u, 0u, 0u);
            if (c8 < CIN / 8 && gx >= 0 && gx < a.n0 && gy >= 0 && gy < a.n1 && gz >= 0 && gz < a.n2)
                v = reinterpret_cast<const uint4 *>(a.x + (((long long)gx * a.n1 + gy) * a.n2 + gz) * CIN)[c8];
            const uint16_t h[8] = {(uint16_t)(v.x & 0xffff), (uint16_t)(v.x >> 16), (uint16_t)(v.y & 0xffff),
                                   (uint16_t)(v.y >> 16),    (uint16_t)(v.z & 0xffff), (uint16_t)(v.z >> 16),
                                   (uint16_t)(v.w & 0xffff), (uint16_t)(v.w >> 16)};
#pragma unroll
            for (int s = 0; s < 3; ++s) {      // copy s holds x at z0 + z + (s - 1)
                const int z = hz - s;            // gz = z0 + hz - 1 = z0 + z + s - 1
                if (z < 0 || z >= 32) continue;
#pragma unroll
                for (int j = 0; j < 8; ++j)
                    st_bf16(xs + s * XCOPY + (c8 * 8 + j) * XCS + (hrow * 32 + z) * 2, h[j]);
            }
        }
        // output-gradient rows of the block, channel-major
        for (int i = tid; i < WG_BX * WG_BY * 32 * 4; i += WG_THREADS) {
            const int c8 = i & 3, vox = i >> 2;
            const int z = vox & 31, r = vox >> 5;
            const int ox = x0 + r / WG_BY, oy = y0 + r % WG_BY;
            uint4 v = make_uint4(0u, 0u, 0u, 0u);
            if (ox < a.n0 && oy < a.n1)
                v = reinterpret_cast<const uint4 *>(a.dz + (((long long)ox * a.n1 + oy) * a.n2 + z0 + z) * 32)[c8];
            const uint16_t h[8] = {(uint16_t)(v.x & 0xffff), (uint16_t)(v.x >> 16), (uint16_t)(v.y & 0xffff),
                                   (uint16_t)(v.y >> 16),    (uint16_t)(v.z & 0xffff), (uint16_t)(v.z >> 16),
                                   (uint16_t)(v.w & 0xffff), (uint16_t)(v.w >> 16)};
#pragma unroll
            for (int j = 0; j < 8; ++j) st_bf16(ds + (c8 * 8 + j) * DCS + (r * 32 + z) * 2, h[j]);
        }
        __syncthreads();
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int tap = wave + 8 * j;
            if (tap >= 27) break;
            const int dx = tap / 9 - 1, dy = (tap / 3) % 3 - 1, dzs = tap % 3 - 1;
            const char *xcopy = xs + (dzs + 1) * XCOPY + m * XCS;
            for (int r = 0; r < WG_BX * WG_BY; ++r) {
                const int hrow = (r / WG_BY + dx + 1) * WHY + (r % WG_BY + dy + 1);
#pragma unroll
                for (int c = 0; c < 2; ++c) {
                    const bf16x8 A = lds_frag(ds + m * DCS + (r * 32 + c * 16 + kh * 8) * 2);
                    const bf16x8 B = lds_frag(xcopy + (hrow * 32 + c * 16 + kh * 8) * 2);
                    acc[j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(A, B, acc[j], 0, 0, 0);
                }
            }
        }
    }
    // partial dW of this workgroup: work[blockIdx][tap][co][ci], ci < CIN
    float *out = a.work + (size_t)blockIdx.x * 27 * 32 * CIN;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const int tap = wave + 8 * j;
        if (tap >= 27) break;
        const int ci = m;
        if (ci >= CIN) continue;
#pragma unroll
        for (int i = 0; i < 16; ++i) out[((size_t)tap * 32 + mfma_row(i, lane)) * CIN + ci] = acc[j][i];
    }
}

// ---------------------------------------------------------------------------------
__global__ __launch_bounds__(256) void cnn_flip(const uint16_t *w, uint16_t *wt) {
    const int i = blockIdx.x * 256 + threadIdx.x;     // over 27*32*32
    if (i >= 27 * 32 * 32) return;
    const int ci = i & 31, co = (i >> 5) & 31, tap = i >> 10;
    // wt[tap][ci][co] = w[26 - tap][co][ci]  (stored with the second index as "cout")
    wt[((size_t)tap * 32 + ci) * 32 + co] = w[((size_t)(26 - tap) * 32 + co) * 32 + ci];
}

__global__ __launch_bounds__(256) void cnn_relu_grad(const float *__restrict__ dh, const uint16_t *__restrict__ act,
                                                     uint16_t *__restrict__ dz, float *__restrict__ part,
                                                     long long n) {
    __shared__ float red[8][32];
    const int c = threadIdx.x & 31, vo = threadIdx.x >> 5;
    float sum = 0.0f;
    const long long base = (long long)blockIdx.x * 1024;
    for (int k = 0; k < 128; ++k) {
        const long long v = base + k * 8 + vo;
        if (v < n) {
            const long long idx = v * 32 + c;
            const uint16_t av = act[idx];
            const float g = ((av & 0x8000u) == 0 && av != 0) ? dh[idx] : 0.0f;
            dz[idx] = f2bf(g);
            sum += g;
        }
    }
    red[vo][c] = sum;
    __syncthreads();
    if (threadIdx.x < 32) {
        float s = 0.0f;
        for (int j = 0; j < 8; ++j) s += red[j][threadIdx.x];
        part[(size_t)blockIdx.x * 32 + threadIdx.x] = s;
    }
}

__global__ __launch_bounds__(256) void cnn_loss(const float *__restrict__ r, const float *__restrict__ dp,
                                                const float *__restrict__ dt, const float *__restrict__ mk,
                                                long long n, float inv_count, uint16_t *__restrict__ dr,
                                                float *__restrict__ part) {
    __shared__ float red[3][256];
    float se = 0.0f, sw = 0.0f, sg = 0.0f;
    for (long long v = blockIdx.x * 256LL + threadIdx.x; v < n; v += (long long)gridDim.x * 256) {
        const float w = mk[v];
        const float e = dp[v] * r[v] - (dt[v] - dp[v]);
        se += w * e * e;
        sw += w;
        const uint16_t g = f2bf(2.0f * w * e * dp[v] * inv_count);
        sg += bf2f(g);
        uint4 *row = reinterpret_cast<uint4 *>(dr + v * 32);
        row[0] = make_uint4((uint32_t)g, 0u, 0u, 0u);
        row[1] = make_uint4(0u, 0u, 0u, 0u);
        row[2] = make_uint4(0u, 0u, 0u, 0u);
        row[3] = make_uint4(0u, 0u, 0u, 0u);
    }
    red[0][threadIdx.x] = se;
    red[1][threadIdx.x] = sw;
    red[2][threadIdx.x] = sg;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s) {
            red[0][threadIdx.x] += red[0][threadIdx.x + s];
            red[1][threadIdx.x] += red[1][threadIdx.x + s];
            red[2][threadIdx.x] += red[2][threadIdx.x + s];
        }
        __syncthreads();
    }
    if (threadIdx.x < 3) part[3 * blockIdx.x + threadIdx.x] = red[threadIdx.x][0];
}

__global__ __launch_bounds__(256) void cnn_adam(float *p, const float *g, float *m, float *v, uint16_t *pb, long long n,
                                                float lr, float b1, float b2, float eps, float c1, float c2) {
    for (long long i = blockIdx.x * 256LL + threadIdx.x; i < n; i += (long long)gridDim.x * 256) {
        const float gi = g[i];
        const float mi = b1 * m[i] + (1.0f - b1) * gi;
        const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
        m[i] = mi;
        v[i] = vi;
        const float pi = p[i] - lr * (mi * c1) / (sqrtf(vi * c2) + eps);
        p[i] = pi;
        if (pb) pb[i] = f2bf(pi);
    }
}

__global__ __launch_bounds__(256) void reduce_rows(const float *__restrict__ part, int rows, int cols,
                                                   float *__restrict__ out, int accumulate) {
    const int c = blockIdx.x * 256 + threadIdx.x;
    if (c >= cols) return;
    float s = 0.0f;
    for (int r = 0; r < rows; ++r) s += part[(size_t)r * cols + c];
    out[c] = accumulate ? out[c] + s : s;
}

__device__ __forceinline__ float inv_scale(const float *m) { return 1.0f / fmaxf(*m, 1e-30f); }

__global__ __launch_bounds__(256) void cnn_pack(const float *__restrict__ a, const float *__restrict__ d,
                                                const float *__restrict__ dp, const float *__restrict__ amax,
                                                const float *__restrict__ dmax, uint16_t *__restrict__ x,
                                                long long n) {
    const float sa = inv_scale(amax), sd = inv_scale(dmax);
    for (long long v = blockIdx.x * 256LL + threadIdx.x; v < n; v += (long long)gridDim.x * 256) {
        const uint32_t c0 = f2bf(a[v] * sa), c1 = f2bf(d[v]), c2 = f2bf(dp[v] * sd);
        uint4 *row = reinterpret_cast<uint4 *>(x + v * 16);
        row[0] = make_uint4(c0 | (c1 << 16), c2, 0u, 0u);
        row[1] = make_uint4(0u, 0u, 0u, 0u);
    }
}

__global__ __launch_bounds__(256) void cnn_apply(const float *__restrict__ dp, const float *__restrict__ r,
                                                 float bound, const float *__restrict__ dmax, float region_frac,
                                                 float apply_frac, float *__restrict__ out,
                                                 unsigned long long *__restrict__ counts, long long n) {
    __shared__ unsigned red[2][256];
    const float region_min = region_frac * *dmax;
    const float apply_min = apply_frac * *dmax;
    unsigned ex = 0, reg = 0;
    for (long long v = blockIdx.x * 256LL + threadIdx.x; v < n; v += (long long)gridDim.x * 256) {
        const float d = dp[v];
        const bool applied = d >= apply_min;     // r is read (and defined) only there
        const float rv = applied ? r[v] : 0.0f;
        const bool ok = fabsf(rv) <= bound;      // NaN fails too
        out[v] = (ok && applied) ? d * (1.0f + rv) : d;
        const bool in_region = d >= region_min;
        ex += (!ok && in_region) ? 1u : 0u;
        reg += in_region ? 1u : 0u;
    }
    red[0][threadIdx.x] = ex;
    red[1][threadIdx.x] = reg;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s) {
            red[0][threadIdx.x] += red[0][threadIdx.x + s];
            red[1][threadIdx.x] += red[1][threadIdx.x + s];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        if (red[0][0]) atomicAdd(&counts[0], (unsigned long long)red[0][0]);
        if (red[1][0]) atomicAdd(&counts[1], (unsigned long long)red[1][0]);
    }
}

// Maxima of the activity and of the plain dose (the network input's scales) in one pass:
// per-block maxima into part[2 * block + {0, 1}], then one block reduces them into
// out[0], out[1].  A maximum is exact, whatever the order.
constexpr int MAX2_BLOCKS = (PVD_CNN_MAXIMA_FLOATS - 2) / 2;   // partial maxima per call

__device__ __forceinline__ float wave_max(float v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = fmaxf(v, __shfl_xor(v, o));
    return v;
}

__global__ __launch_bounds__(256) void cnn_max2_part(const float *__restrict__ a, const float *__restrict__ b,
                                                     long long n, float *__restrict__ part) {
    __shared__ float red[2][4];
    float ma = -INFINITY, mb = -INFINITY;
    for (long long v = blockIdx.x * 256LL + threadIdx.x; v < n; v += (long long)gridDim.x * 256) {
        ma = fmaxf(ma, a[v]);
        mb = fmaxf(mb, b[v]);
    }
    ma = wave_max(ma);
    mb = wave_max(mb);
    if ((threadIdx.x & 63) == 0) {
        red[0][threadIdx.x >> 6] = ma;
        red[1][threadIdx.x >> 6] = mb;
    }
    __syncthreads();
    if (threadIdx.x < 2) {
        const float *r = red[threadIdx.x];
        part[2 * blockIdx.x + threadIdx.x] = fmaxf(fmaxf(r[0], r[1]), fmaxf(r[2], r[3]));
    }
}

__global__ __launch_bounds__(256) void cnn_max2_final(const float *__restrict__ part, int nb, float *__restrict__ out) {
    __shared__ float red[2][4];
    float ma = -INFINITY, mb = -INFINITY;
    for (int i = threadIdx.x; i < nb; i += 256) {
        ma = fmaxf(ma, part[2 * i]);
        mb = fmaxf(mb, part[2 * i + 1]);
    }
    ma = wave_max(ma);
    mb = wave_max(mb);
    if ((threadIdx.x & 63) == 0) {
        red[0][threadIdx.x >> 6] = ma;
        red[1][threadIdx.x >> 6] = mb;
    }
    __syncthreads();
    if (threadIdx.x < 2) {
        const float *r = red[threadIdx.x];
        out[threadIdx.x] = fmaxf(fmaxf(r[0], r[1]), fmaxf(r[2], r[3]));
    }
}

// Whole-volume fallback decided on the device from cnn_apply's counts: out = dp when
// counts[1] > 0 and counts[0] > max_exceed_frac * counts[1]; otherwise every block
// returns at once.
__global__ __launch_bounds__(256) void cnn_fallback(const float *__restrict__ dp,
                                                    const unsigned long long *__restrict__ counts,
                                                    double max_exceed_frac, float *__restrict__ out, long long n) {
    const unsigned long long ex = counts[0], reg = counts[1];
    if (!(reg > 0 && (double)ex > max_exceed_frac * (double)reg)) return;
    for (long long v = blockIdx.x * 256LL + threadIdx.x; v < n; v += (long long)gridDim.x * 256) out[v] = dp[v];
}

// Region blocks of the forward kernel (FWD_BX x FWD_BY x 32 output voxels):
// flag[b] = 1 if block b holds a voxel with dp >= frac * max; count reset.
__global__ __launch_bounds__(256) void cnn_block_flags(const float *__restrict__ dp, int n0, int n1, int n2,
                                                       const float *__restrict__ dmax, float frac,
                                                       int *__restrict__ flag, int *__restrict__ count) {
    const int nby = (n1 + FWD_BY - 1) / FWD_BY, nbz = n2 >> 5;
    const int b = blockIdx.x;
    const int bz = b % nbz, by = (b / nbz) % nby, bx = b / (nbz * nby);
    const float thr = frac * *dmax;
    const int t = threadIdx.x;                       // 256 = FWD_BX * FWD_BY * 32 voxels
    const int z = bz * 32 + (t & 31), y = by * FWD_BY + ((t >> 5) % FWD_BY), x = bx * FWD_BX + (t >> 5) / FWD_BY;
    const bool hit = x < n0 && y < n1 && dp[((long long)x * n1 + y) * n2 + z] >= thr;
    const int any = __syncthreads_or(hit ? 1 : 0);
    if (t == 0) flag[b] = any ? 1 : 0;
    if (b == 0 && t == 0) *count = 0;
}

// list the blocks within (3, 2, 1) blocks (6 voxels, the network's receptive field)
// of a flagged block; the order follows the atomic slots (the outputs do not depend on it)
__global__ __launch_bounds__(256) void cnn_block_list(const int *__restrict__ flag, int nbx, int nby, int nbz,
                                                      int *__restrict__ list, int *__restrict__ count) {
    const int total = nbx * nby * nbz;
    for (int b = blockIdx.x * 256 + threadIdx.x; b < total; b += gridDim.x * 256) {
        const int bz = b % nbz, by = (b / nbz) % nby, bx = b / (nbz * nby);
        bool near = false;
        for (int i = max(bx - 3, 0); i <= min(bx + 3, nbx - 1) && !near; ++i)
            for (int j = max(by - 2, 0); j <= min(by + 2, nby - 1) && !near; ++j)
                for (int k = max(bz - 1, 0); k <= min(bz + 1, nbz - 1); ++k)
                    if (flag[(i * nby + j) * nbz + k]) { near = true; break; }
        if (near) list[atomicAdd(count, 1)] = b;
    }
}

int g_num_cus = 0;

int num_cus() {
    if (g_num_cus == 0) {
        int dev = 0, v = 0;
        if (hipGetDevice(&dev) == hipSuccess &&
            hipDeviceGetAttribute(&v, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess && v > 0)
            g_num_cus = v;
        else
            g_num_cus = 256;
    }
    return g_num_cus;
}

template <int CIN, int EPI>
hipError_t launch_fwd(const ConvArgs &a, int blocks, hipStream_t st) {
    static bool attr = false;
    if (!attr) {
        hipError_t e = hipFuncSetAttribute((const void *)cnn_conv_fwd<CIN, EPI>,
                                           hipFuncAttributeMaxDynamicSharedMemorySize, (int)fwd_lds_bytes<CIN>());
        if (e != hipSuccess) return e;
        attr = true;
    }
    hipLaunchKernelGGL((cnn_conv_fwd<CIN, EPI>), dim3(blocks), dim3(FWD_THREADS), fwd_lds_bytes<CIN>(), st, a);
    return hipGetLastError();
}

template <int BX, int BY>
hipError_t launch_in(const ConvArgs &a, int blocks, hipStream_t st) {
    static bool attr = false;
    constexpr size_t lds = in_lds_bytes<BX, BY>();
    if (!attr) {
        hipError_t e = hipFuncSetAttribute((const void *)cnn_conv_in<BX, BY>,
                                           hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        if (e != hipSuccess) return e;
        attr = true;
    }
    hipLaunchKernelGGL((cnn_conv_in<BX, BY>), dim3(blocks), dim3(FWD_THREADS), lds, st, a);
    return hipGetLastError();
}

template <int BX, int BY>
hipError_t launch_out(const ConvArgs &a, int blocks, hipStream_t st) {
    static bool attr = false;
    constexpr size_t lds = out_lds_bytes<BX, BY>();
    if (!attr) {
        hipError_t e = hipFuncSetAttribute((const void *)cnn_conv_out<BX, BY>,
                                           hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        if (e != hipSuccess) return e;
        attr = true;
    }
    hipLaunchKernelGGL((cnn_conv_out<BX, BY>), dim3(blocks), dim3(FWD_THREADS), lds, st, a);
    return hipGetLastError();
}

// whole volumes (no region list): 4 x 4-row blocks for the input and output layers
#ifndef PVD_CNN_IO44
#define PVD_CNN_IO44 1
#endif
template <bool IN>
hipError_t launch_io(const ConvArgs &a, int grid, bool whole, hipStream_t st) {
    if (whole && PVD_CNN_IO44) {
        const long long t = (long long)((a.n0 + 3) / 4) * ((a.n1 + 3) / 4) * (a.n2 >> 5);
        const int g = (int)(t < num_cus() ? t : num_cus());
        return IN ? launch_in<4, 4>(a, g, st) : launch_out<4, 4>(a, g, st);
    }
    return IN ? launch_in<FWD_BX, FWD_BY>(a, grid, st) : launch_out<FWD_BX, FWD_BY>(a, grid, st);
}

template <int CIN>
hipError_t launch_wgrad(const WgradArgs &a, int blocks, hipStream_t st) {
    static bool attr = false;
    if (!attr) {
        hipError_t e = hipFuncSetAttribute((const void *)cnn_wgrad<CIN>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                           (int)WG_LDS);
        if (e != hipSuccess) return e;
        attr = true;
    }
    hipLaunchKernelGGL((cnn_wgrad<CIN>), dim3(blocks), dim3(WG_THREADS), WG_LDS, st, a);
    return hipGetLastError();
}

int check_grid(int n0, int n1, int n2) {
    if (n0 < 1 || n1 < 1 || n2 < 32 || (n2 & 31))
        return pvd_set_error(PVD_EINVAL, "CNN grid %dx%dx%d: need n2 a positive multiple of 32", n0, n1, n2);
    return PVD_OK;
}

int wgrad_blocks(int n0, int n1, int n2) {
    const long long total = (long long)((n0 + WG_BX - 1) / WG_BX) * ((n1 + WG_BY - 1) / WG_BY) * (n2 >> 5);
    return (int)(total < WG_MAX_BLOCKS ? total : WG_MAX_BLOCKS);
}

}  // namespace

// =================================================================================
extern "C" {

int pvd_cnn_conv_fwd(const void *x, int cin, const void *w, const float *bias, int epi, const void *resid, void *y,
                     void *act, float *y32, int n0, int n1, int n2, const int *blocks, const int *nblocks,
                     void *stream) {
    if ((blocks == nullptr) != (nblocks == nullptr)) return pvd_set_error(PVD_EINVAL, "blocks and nblocks go together");
    if (int rc = check_grid(n0, n1, n2)) return rc;
    if (!x || !w) return pvd_set_error(PVD_EINVAL, "null input or weights");
    if (cin != 4 && cin != 16 && cin != 32) return pvd_set_error(PVD_EINVAL, "cin must be 4, 16 or 32 (got %d)", cin);
    if (cin == 4 && (epi != PVD_CNN_EPI_RELU || act))
        return pvd_set_error(PVD_EINVAL, "cin 4 (the input layer) takes the ReLU epilogue without act");
    switch (epi) {
        case PVD_CNN_EPI_RELU: if (!y) return pvd_set_error(PVD_EINVAL, "null output"); break;
        case PVD_CNN_EPI_RESID_RELU:
            if (!y || !resid) return pvd_set_error(PVD_EINVAL, "residual layer needs y and resid");
            break;
        case PVD_CNN_EPI_LINEAR_F32:
        case PVD_CNN_EPI_ACCUM_F32: if (!y32) return pvd_set_error(PVD_EINVAL, "null FP32 output"); break;
        default: return pvd_set_error(PVD_EINVAL, "unknown epilogue %d", epi);
    }
    // BF16 outputs are written as 16-byte chunks of whole 32-voxel tiles
    if ((epi == PVD_CNN_EPI_RELU || epi == PVD_CNN_EPI_RESID_RELU) &&
        ((reinterpret_cast<uintptr_t>(y) | reinterpret_cast<uintptr_t>(act)) & 15))
        return pvd_set_error(PVD_EINVAL, "y and act must be 16-byte aligned");
    ConvArgs a{(const uint16_t *)x, (const uint16_t *)w, bias, (const uint16_t *)resid, (uint16_t *)y,
               (uint16_t *)act, y32, n0, n1, n2, blocks, nblocks};
    const long long total = (long long)((n0 + FWD_BX - 1) / FWD_BX) * ((n1 + FWD_BY - 1) / FWD_BY) * (n2 >> 5);
    const int grid = (int)(total < num_cus() ? total : num_cus());
    hipStream_t st = (hipStream_t)stream;
    hipError_t e;
    if (cin == 4) {
        e = launch_io<true>(a, grid, blocks == nullptr, st);
    } else if (cin == 16) {
        switch (epi) {
            case 0: e = launch_fwd<16, 0>(a, grid, st); break;
            case 1: e = launch_fwd<16, 1>(a, grid, st); break;
            case 2: e = launch_fwd<16, 2>(a, grid, st); break;
            default: e = launch_fwd<16, 3>(a, grid, st); break;
        }
    } else {
        switch (epi) {
            case 0: e = launch_fwd<32, 0>(a, grid, st); break;
            case 1: e = launch_fwd<32, 1>(a, grid, st); break;
            case 2: e = launch_io<false>(a, grid, blocks == nullptr, st); break;
            default: e = launch_fwd<32, 3>(a, grid, st); break;
        }
    }
    return e == hipSuccess ? PVD_OK : pvd_hip_error(e, "launch CNN convolution");
}

int pvd_cnn_flip_weights(const void *w, void *wt, void *stream) {
    if (!w || !wt) return pvd_set_error(PVD_EINVAL, "null argument");
    hipLaunchKernelGGL(cnn_flip, dim3((27 * 32 * 32 + 255) / 256), dim3(256), 0, (hipStream_t)stream,
                       (const uint16_t *)w, (uint16_t *)wt);
    const hipError_t e = hipGetLastError();
    return e == hipSuccess ? PVD_OK : pvd_hip_error(e, "launch weight flip");
}

int pvd_cnn_relu_grad(const float *dh, const void *act, void *dz, float *dbias_part, long long n, void *stream) {
    if (!dh || !act || !dz || !dbias_part || n < 1) return pvd_set_error(PVD_EINVAL, "bad argument");
    const long long blocks = (n + 1023) / 1024;
    if (blocks > 0x7fffffffLL) return pvd_set_error(PVD_ENOTSUP, "too many voxels");
    hipLaunchKernelGGL(cnn_relu_grad, dim3((unsigned)blocks), dim3(256), 0, (hipStream_t)stream, dh,
                       (const uint16_t *)act, (uint16_t *)dz, dbias_part, n);
    const hipError_t e = hipGetLastError();
    return e == hipSuccess ? PVD_OK : pvd_hip_error(e, "launch ReLU gradient");
}

size_t pvd_cnn_wgrad_work_floats(int n0, int n1, int n2) {
    if (n0 < 1 || n1 < 1 || n2 < 32) return 0;
    return (size_t)wgrad_blocks(n0, n1, n2) * 27 * 32 * 32;
}

int pvd_cnn_wgrad(const void *x, int cin, const void *dz, float *dw, float *work, int n0, int n1, int n2,
                  void *stream) {
    if (int rc = check_grid(n0, n1, n2)) return rc;
    if (!x || !dz || !dw || !work) return pvd_set_error(PVD_EINVAL, "null argument");
    if (cin != 16 && cin != 32) return pvd_set_error(PVD_EINVAL, "cin must be 16 or 32 (got %d)", cin);
    WgradArgs a{(const uint16_t *)x, (const uint16_t *)dz, work, n0, n1, n2};
    const int blocks = wgrad_blocks(n0, n1, n2);
    hipStream_t st = (hipStream_t)stream;
    hipError_t e = cin == 16 ? launch_wgrad<16>(a, blocks, st) : launch_wgrad<32>(a, blocks, st);
    if (e != hipSuccess) return pvd_hip_error(e, "launch weight gradient");
    const int cols = 27 * 32 * cin;
    hipLaunchKernelGGL(reduce_rows, dim3((cols + 255) / 256), dim3(256), 0, st, (const float *)work, blocks, cols, dw, 0);
    e = hipGetLastError();
    return e == hipSuccess ? PVD_OK : pvd_hip_error(e, "launch reduction");
}

int pvd_cnn_loss(const float *r, const float *dp, const float *dt, const float *m, long long n, float inv_count,
                 void *dr, float *loss_part, void *stream) {
    if (!r || !dp || !dt || !m || !dr || !loss_part || n < 1) return pvd_set_error(PVD_EINVAL, "bad argument");
    hipLaunchKernelGGL(cnn_loss, dim3(PVD_CNN_LOSS_BLOCKS), dim3(256), 0, (hipStream_t)stream, r, dp, dt, m, n, inv_count,
                       (uint16_t *)dr, loss_part);
    const hipError_t e = hipGetLastError();
    return e == hipSuccess ? PVD_OK : pvd_hip_error(e, "launch loss");
}

int pvd_cnn_adam(float *param, const float *grad, float *m, float *v, void *param_bf16, long long n, float lr,
                 float beta1, float beta2, float eps, int step, void *stream) {
    if (!param || !grad || !m || !v || n < 1 || step < 1) return pvd_set_error(PVD_EINVAL, "bad argument");
    const float c1 = 1.0f / (1.0f - powf(beta1, (float)step));
    const float c2 = 1.0f / (1.0f - powf(beta2, (float)step));
    long long blocks = (n + 255) / 256;
    if (blocks > 4096) blocks = 4096;
    hipLaunchKernelGGL(cnn_adam, dim3((unsigned)blocks), dim3(256), 0, (hipStream_t)stream, param, grad, m, v,
                       (uint16_t *)param_bf16, n, lr, beta1, beta2, eps, c1, c2);
    const hipError_t e = hipGetLastError();
    return e == hipSuccess ? PVD_OK : pvd_hip_error(e, "launch Adam");
}

int pvd_cnn_pack_inputs(const float *activity, const float *density, const float *dose_plain, const float *act_max,
                        const float *dose_max, void *x, long long n, void *stream) {
    if (!activity || !density || !dose_plain || !act_max || !dose_max || !x || n < 1)
        return pvd_set_error(PVD_EINVAL, "bad argument");
    long long blocks = (n + 255) / 256;
    if (blocks > 8192) blocks = 8192;
    hipLaunchKernelGGL(cnn_pack, dim3((unsigned)blocks), dim3(256), 0, (hipStream_t)stream, activity, density,
                       dose_plain, act_max, dose_max, (uint16_t *)x, n);
    const hipError_t e = hipGetLastError();
    return e == hipSuccess ? PVD_OK : pvd_hip_error(e, "launch input packing");
}

int pvd_cnn_apply_correction(const float *dp, const float *r, float bound, const float *dose_max, float region_frac,
                             float apply_frac, float *out, unsigned long long *counts, long long n, void *stream) {
    if (!dp || !r || !dose_max || !out || !counts || n < 1 || !(bound >= 0.0f) || !(region_frac >= 0.0f) ||
        !(apply_frac >= 0.0f))
        return pvd_set_error(PVD_EINVAL, "bad argument");
    // 8 or more voxels per thread: the block's reduction and atomics are paid once per
    // 2048 voxels rather than per 256
    long long blocks = (n + 2047) / 2048;
    if (blocks > 2048) blocks = 2048;
    hipLaunchKernelGGL(cnn_apply, dim3((unsigned)blocks), dim3(256), 0, (hipStream_t)stream, dp, r, bound, dose_max,
                       region_frac, apply_frac, out, counts, n);
    const hipError_t e = hipGetLastError();
    return e == hipSuccess ? PVD_OK : pvd_hip_error(e, "launch correction");
}

int pvd_cnn_input_maxima(const float *activity, const float *dose_plain, long long n, float *maxima, void *stream) {
    if (!activity || !dose_plain || !maxima || n < 1) return pvd_set_error(PVD_EINVAL, "bad argument");
    long long blocks = (n + 2047) / 2048;
    if (blocks > MAX2_BLOCKS) blocks = MAX2_BLOCKS;
    hipStream_t st = (hipStream_t)stream;
    hipLaunchKernelGGL(cnn_max2_part, dim3((unsigned)blocks), dim3(256), 0, st, activity, dose_plain, n, maxima + 2);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return pvd_hip_error(e, "launch input maxima");
    hipLaunchKernelGGL(cnn_max2_final, dim3(1), dim3(256), 0, st, (const float *)(maxima + 2), (int)blocks, maxima);
    e = hipGetLastError();
    return e == hipSuccess ? PVD_OK : pvd_hip_error(e, "launch input maxima");
}

int pvd_cnn_apply_fallback(const float *dp, const unsigned long long *counts, double max_exceed_frac, float *out,
                           long long n, void *stream) {
    if (!dp || !counts || !out || n < 1 || !(max_exceed_frac >= 0.0))
        return pvd_set_error(PVD_EINVAL, "bad argument");
    long long blocks = (n + 255) / 256;
    if (blocks > 1024) blocks = 1024;     // blocks that only read the counts, most often
    hipLaunchKernelGGL(cnn_fallback, dim3((unsigned)blocks), dim3(256), 0, (hipStream_t)stream, dp, counts,
                       max_exceed_frac, out, n);
    const hipError_t e = hipGetLastError();
    return e == hipSuccess ? PVD_OK : pvd_hip_error(e, "launch fallback");
}

int pvd_cnn_region_blocks(const float *dp, int n0, int n1, int n2, const float *dose_max, float frac, int *flags,
                          int *blocks, int *nblocks, void *stream) {
    if (int rc = check_grid(n0, n1, n2)) return rc;
    if (!dp || !dose_max || !flags || !blocks || !nblocks || !(frac >= 0.0f))
        return pvd_set_error(PVD_EINVAL, "bad argument");
    const int nbx = (n0 + FWD_BX - 1) / FWD_BX, nby = (n1 + FWD_BY - 1) / FWD_BY, nbz = n2 >> 5;
    const long long total = (long long)nbx * nby * nbz;
    if (total > 0x7fffffffLL) return pvd_set_error(PVD_ENOTSUP, "too many blocks");
    hipStream_t st = (hipStream_t)stream;
    hipLaunchKernelGGL(cnn_block_flags, dim3((unsigned)total), dim3(256), 0, st, dp, n0, n1, n2, dose_max, frac, flags,
                       nblocks);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return pvd_hip_error(e, "launch region flags");
    const unsigned g = (unsigned)std::min<long long>((total + 255) / 256, 1024);
    hipLaunchKernelGGL(cnn_block_list, dim3(g), dim3(256), 0, st, (const int *)flags, nbx, nby, nbz, blocks, nblocks);
    e = hipGetLastError();
    return e == hipSuccess ? PVD_OK : pvd_hip_error(e, "launch region list");
}

size_t pvd_cnn_region_block_count(int n0, int n1, int n2) {
    if (n0 < 1 || n1 < 1 || n2 < 32) return 0;
    return (size_t)((n0 + FWD_BX - 1) / FWD_BX) * ((n1 + FWD_BY - 1) / FWD_BY) * (n2 >> 5);
}

int pvd_reduce_rows(const float *part, int rows, int cols, float *out, int accumulate, void *stream) {
    if (!part || !out || rows < 1 || cols < 1) return pvd_set_error(PVD_EINVAL, "bad argument");
    hipLaunchKernelGGL(reduce_rows, dim3((cols + 255) / 256), dim3(256), 0, (hipStream_t)stream, part, rows, cols, out,
                       accumulate);
    const hipError_t e = hipGetLastError();
    return e == hipSuccess ? PVD_OK : pvd_hip_error(e, "launch reduction");
}

}  // extern "C"
