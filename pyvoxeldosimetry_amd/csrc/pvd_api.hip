// Host side of libpvd: the C ABI declared in include/pvd.h.
//
// Plans own their device buffers (twiddles, complex workspace, kernel
// spectrum or reversed small kernel).  Dose calls only launch kernels on the
// caller's stream; nothing in a dose call allocates or synchronises.
#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <map>
#include <mutex>
#include <new>
#include <string>
#include <vector>

#include "pvd_internal.h"

// kernels (pvd_fft*.hip launchers are declared in pvd_internal.h; pvd_spectrum.hip, pvd_direct.hip)
hipError_t pvd_launch_dpk(double v, const int g[3], const pvd_dpk_term *terms, int nterms, double factor,
                          double *out, hipStream_t st);
hipError_t pvd_launch_spectrum(const int P[3], const int lo[3], const int u[3], const int s[3], const int K[3],
                               int Hp, int sym, const double *kernel, double2 *T1, double2 *T2, float2 *S,
                               hipStream_t st);
size_t pvd_direct_lds_bytes(const int k[3]);
hipError_t pvd_direct_set_lds_limit(size_t bytes);
hipError_t pvd_launch_wsum(const float *const *maps, const float *w, int T, size_t n, float *out, hipStream_t st);
hipError_t pvd_launch_direct(const int N[3], const int K[3], const int sh[3], int wrap, const float *const *acts,
                             const float *w, int T, const float *kr, const float *dens, float *out, float scale,
                             hipStream_t st);

#define PVD_VERSION 100

namespace {

thread_local std::string g_err;

int fail(int code, const char *fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

int hip_fail(hipError_t e, const char *what) {
    return fail(e == hipErrorOutOfMemory ? PVD_ENOMEM : PVD_EHIP, "%s: %s (%d)", what, hipGetErrorString(e),
                (int)e);
}

}  // namespace

int pvd_set_error(int code, const char *fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

int pvd_hip_error(hipError_t e, const char *what) { return hip_fail(e, what); }

namespace {

#define PVD_HIP(call, what)                         \
    do {                                            \
        hipError_t e_ = (call);                     \
        if (e_ != hipSuccess) return hip_fail(e_, what); \
    } while (0)

constexpr size_t kFftLdsCap = 96 * 1024;   // per-workgroup LDS budget of the FFT passes
constexpr size_t kDescribeLds = 160 * 1024; // LDS per workgroup assumed by pvd_plan_describe (gfx950)
// y/x lengths 16 * b with a register two-step instance (the kernels' PVD_XY_BS list)
#define PVD_XY_LIST(B) B,
constexpr int kXyRegB[] = {PVD_XY_BS(PVD_XY_LIST)};
#undef PVD_XY_LIST

bool is_smooth(int n) {
    if (n < 1) return false;
    for (int p : {2, 3, 5, 7})
        while (n % p == 0) n /= p;
    return n == 1;
}

int next_smooth(int n) {
    if (n < 1) n = 1;
    while (!is_smooth(n)) ++n;
    return n;
}

// Stockham radices of n (largest power-of-two radix max_radix2 = 16 or 8)
int factor_radices(int n, unsigned char *radix, int max_radix2) {
    int ns = 0;
    int p2 = 0;
    while (n % 2 == 0) { n /= 2; ++p2; }
    while (max_radix2 >= 16 && p2 >= 4) { radix[ns++] = 16; p2 -= 4; }
    while (p2 >= 3) { radix[ns++] = 8; p2 -= 3; }
    if (p2 == 3) radix[ns++] = 8;
    else if (p2 == 2) radix[ns++] = 4;
    else if (p2 == 1) radix[ns++] = 2;
    for (int p : {3, 5, 7})
        while (n % p == 0) { n /= p; radix[ns++] = (unsigned char)p; }
    return n == 1 ? ns : -1;
}

int round_up(int a, int b) { return (a + b - 1) / b * b; }

struct device_limits {
    int dev = -1;
    size_t max_lds = 64 * 1024;
};

std::mutex g_mu;
std::map<int, bool> g_attr_done;

int ensure_kernel_attrs(int dev, size_t *max_lds) {
    int v = 0;
    PVD_HIP(hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev), "query LDS size");
    *max_lds = (size_t)v;
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_attr_done[dev]) {
        PVD_HIP(pvd_set_fft_lds_limits(std::min(kFftLdsCap, (size_t)v)), "set FFT LDS limit");
        PVD_HIP(pvd_direct_set_lds_limit((size_t)v), "set direct LDS limit");
        g_attr_done[dev] = true;
    }
    return PVD_OK;
}

// per-process cache of FP32 twiddle tables built in FP64 (one per length and device)
struct tw_key {
    int dev, n;
    bool operator<(const tw_key &o) const { return dev != o.dev ? dev < o.dev : n < o.n; }
};
std::map<tw_key, float2 *> g_tw;

int get_twiddles(int dev, int n, const float2 **out) {
    std::lock_guard<std::mutex> lk(g_mu);
    auto it = g_tw.find({dev, n});
    if (it != g_tw.end()) { *out = it->second; return PVD_OK; }
    std::vector<float2> h(n);
    const double two_pi = 6.283185307179586476925286766559;
    for (int m = 0; m < n; ++m) {
        // exact reduction: angle = 2 pi m / n with m < n
        const double a = two_pi * (double)m / (double)n;
        h[m] = make_float2((float)cos(a), (float)(-sin(a)));
    }
    float2 *d = nullptr;
    PVD_HIP(hipMalloc(&d, sizeof(float2) * n), "allocate twiddles");
    PVD_HIP(hipMemcpy(d, h.data(), sizeof(float2) * n, hipMemcpyHostToDevice), "upload twiddles");
    g_tw[{dev, n}] = d;
    *out = d;
    return PVD_OK;
}

struct launch_cfg {
    int grid = 0, block = 0;
    size_t lds = 0;
};

}  // namespace

struct pvd_plan {
    pvd_plan_params prm{};
    int dev = 0;
    int algo = PVD_ALGO_FFT;
    int N[3]{}, K[3]{}, Ku[3]{};   // grid, kernel, kernel extent used (crop)
    int P[3]{};
    int lo[3]{}, s[3]{};           // kernel placement (first used sample, shift)
    int tile[3]{};                 // tile targets of the x, y, z passes
    int max_radix = 16;            // largest power-of-two Stockham radix
    bool has_kernel = false;
    pvd_vol_d vol{};               // descriptor template
    launch_cfg cfg[PVD_NPASS_FFT];
    size_t work_elems = 0, spec_elems = 0;   // complex elements of the workspace and spectrum
    float2 *work = nullptr;
    float2 *spec = nullptr;
    float *kr = nullptr;           // direct path: reversed cropped kernel (FP32)
    size_t bytes = 0;
    // profiling
    hipEvent_t ev[PVD_NPASS_FFT + 1]{};
    int n_ev = 0;
    bool profiled_call = false;
};

// One launch of a batch pass: members that share a kernel instance and block.
struct batch_launch_d {
    int pass = 0;
    int exemplar = 0;                  // a member's volume index (selects the instance)
    launch_cfg cfg;                    // grid = tiles of all members
    int first = 0, nmem = 0;           // members in the base/vol pools
};

struct pvd_batch {
    std::vector<pvd_plan *> plans;
    int dev = 0;
    int nvol = 0;
    pvd_vol_d *h_tab = nullptr;        // pinned host copy of the descriptor table
    pvd_vol_d *d_tab = nullptr;        // device table read by the batch kernels
    int *d_pool = nullptr;             // device member pools: tile bases, then volume indices
    std::vector<batch_launch_d> launches;   // in pass order
    int pool_n = 0;
    std::vector<char> ready;           // volume inputs set
    hipEvent_t uploaded = nullptr;     // last table upload (host table reusable after it)
    bool upload_pending = false;
    hipStream_t cap = nullptr;         // private stream used only for graph capture
    hipStream_t fork[PVD_BATCH_BRANCHES]{};   // branch streams of the capture (default priority)
    hipEvent_t join[PVD_BATCH_BRANCHES + 1]{};
    hipGraphExec_t exec = nullptr;
};

namespace {

// 16 * B for the two-step instances (256 among them: it runs there when the 256-point
// path does not, with a reference-mode fold or other than 16-line tiles)
bool xy_two_step_len(int P) {
    if (P % 16) return false;
    for (int b : kXyRegB)
        if (P == 16 * b) return true;
    return false;
}

int z_reg_code(int P) {
    for (int c = 1; c < PVD_ZREG_CODES; ++c)
        if (kZregA[c] * kZregB[c] == P) return c;
    return 0;
}

// A length whose transform runs register-resident on axis a (0 x, 1 y, 2 z): 16 * B
// for the two-step instances on y/x (256 on the 256-point path with 16-line tiles and
// no reference-mode fold), and the z two-step table (16-line tiles).
bool register_len(int a, int P, bool fold, int tile) {
    (void)fold;
    if (a < 2) return xy_two_step_len(P);
    return tile == 16 && z_reg_code(P) != 0;
}

// Largest padding (relative to the 7-smooth length) accepted to reach a register
// length.  x: the padded x-planes are implicit zeros and never stored, so a longer
// x transform costs compute only; y: the y passes and X read and write Py rows;
// z: the half spectrum H = Pz/2 + 1 sizes every pass.
constexpr double kRegSlack[3] = {1.0, 0.3, 0.2};

int axis_setup(pvd_plan *p) {
    const pvd_plan_params &q = p->prm;
    const bool ref = q.mode == PVD_MODE_REFERENCE;
    const bool regs = !(q.flags & (PVD_FLAG_STOCKHAM | PVD_FLAG_SMOOTH_LENGTHS));
    for (int a = 0; a < 3; ++a) {
        const int N = q.n[a], K = q.k[a];
        p->N[a] = N;
        p->K[a] = K;
        int need;          // smallest FFT length free of aliasing when padded
        if (ref) {
            const int Ku = std::min(K, N);           // np.fft.fftn(kernel, s=N) crops to [0:N]
            p->Ku[a] = Ku;
            p->lo[a] = 0;
            p->s[a] = 0;
            need = N + Ku - 1;
        } else {
            const int c = K / 2;
            const int mlo = std::max(0, c - N + 1), mhi = std::min(K - 1, c + N - 1);
            p->lo[a] = mlo;
            p->Ku[a] = mhi - mlo + 1;
            p->s[a] = c;
            need = N + std::max(mhi - c, c - mlo);
        }
        int P = q.fft_n[a];
        if (P == 0) {
            P = (ref && is_smooth(N)) ? N : next_smooth(need);
            // pad toward a register-resident length when it costs little (a padded
            // reference-mode length folds the wrap back in)
            // (lengths up to 128 take at most two Stockham stages and are left alone)
            const int tile = p->tile[a];
            if (regs && P > 128 && !register_len(a, P, ref && P != N, tile)) {
                int best = 0;
                for (int L = need; !best && L <= std::min(PVD_MAX_FFT_LEN, 2 * need); ++L)
                    if (is_smooth(L) && register_len(a, L, ref && L != N, tile)) best = L;
                if (best && best <= P * (1.0 + kRegSlack[a])) P = best;
            }
        }
        if (!is_smooth(P)) return fail(PVD_EINVAL, "fft_n[%d]=%d is not 7-smooth", a, P);
        if (ref ? !(P == N || P >= need) : P < need)
            return fail(PVD_EINVAL, "fft_n[%d]=%d aliases: need %s%d", a, P, ref ? "N or >= " : ">= ", need);
        if (P > PVD_MAX_FFT_LEN) return fail(PVD_ENOTSUP, "FFT length %d on axis %d exceeds %d", P, a,
                                             PVD_MAX_FFT_LEN);
        p->P[a] = P;
        pvd_axis_d &ax = p->vol.ax[a];
        ax.n = P;
        ax.N = N;
        ax.nst = factor_radices(P, ax.radix, p->max_radix);
        if (ax.nst < 0 || ax.nst > PVD_MAXSTAGE) return fail(PVD_EINVAL, "cannot factor FFT length %d", P);
        // output o reads FFT index (o - s) mod P in reference mode, o in linear mode
        ax.off = ref ? ((P - (p->s[a] % P)) % P) : 0;
        ax.foldn = (ref && P != N) ? p->Ku[a] - 1 : 0;
    }
    return PVD_OK;
}

// power-of-two lines per tile and a block size so that lines*n <= 16*block and
// the tile plus twiddles fit the LDS budget
void pick_tile(int n, int target, int extra_ld, int *lines, int *block, size_t *lds) {
    const size_t cap_elems = kFftLdsCap / sizeof(float2);
    int l = 1;
    while (l * 2 <= target) l *= 2;
    const int twn = round_up(n, 8);
    while (l > 1 && ((size_t)twn + (size_t)n * (l + extra_ld) > cap_elems || l * n > 16 * 1024)) l /= 2;
    *lines = l;
    int b = round_up((l * n + 15) / 16, 64);
    b = std::max(64, std::min(1024, b));
    *block = b;
    *lds = sizeof(float2) * ((size_t)twn + (size_t)n * (l + extra_ld));
}

int ilog2(int v) {
    int r = 0;
    while ((1 << (r + 1)) <= v) ++r;
    return r;
}

// The register-resident X passes (256-point and fused two-step) address the workspace
// and the kernel spectrum with 32-bit byte offsets from a tile's base: the x-plane walk
// reaches (Nx - 1) * wplane * 8 bytes, the general spectrum (Px - 1) * Py * Hp * 8
// bytes (the real octant less), plus a tile's columns.  Larger volumes take the
// generic X pass (64-bit offsets).
bool x_offsets_fit(long long Nx, long long Px, long long Py, long long Hp, long long wplane) {
    const long long lim = 0xffffffffLL - 256 - 8 * Hp;
    return (Nx - 1) * wplane * 8 <= lim && (Px - 1) * Py * Hp * 8 <= lim;
}

// Host-only layout of an FFT plan: tiles, register paths, launch shapes and buffer
// sizes (no device call).
int fft_layout(pvd_plan *p) {
    pvd_vol_d &v = p->vol;
    const int Nx = p->N[0], Ny = p->N[1];
    const int Px = p->P[0], Py = p->P[1], Pz = p->P[2];
    v.H = Pz / 2 + 1;
    int lz, bz; size_t ldsz;
    pick_tile(Pz, p->tile[2], 1, &lz, &bz, &ldsz);
    // z blocks are a power-of-two multiple of the lines per tile, so the z passes'
    // I/O phases (blockDim / lines threads per line pair) keep every thread busy
    // (at most 1024: both factors are powers of two)
    { int per = 1; while (per * lz < bz) per *= 2; bz = per * lz; }
    int tky, by; size_t ldsy;
    pick_tile(Py, p->tile[1], 0, &tky, &by, &ldsy);
    const bool regs = !(p->prm.flags & PVD_FLAG_STOCKHAM);
    // x tile left to the plan: 8 columns for the fused two-step lengths 16 * B with
    // B >= 18 (the X pass holds a column's two transforms and its spectrum values in
    // registers there, and half-width tiles fit more workgroups per CU)
    int tile_x = p->tile[0];
    if (p->prm.tile[0] == 0 && regs && xy_two_step_len(Px) && Px >= 16 * 18 && Px < 16 * PVD_XY_LDS_B) tile_x = 8;
    int tkx, bx; size_t ldsx;
    pick_tile(Px, tile_x, 0, &tkx, &bx, &ldsx);
    // y/x register paths: 1 = the 256-point path (16-line tiles, no reference-mode fold;
    // checked again in the kernel); 2 = the two-step P = 16 * rb (rb in kXyRegB) in its
    // own kernel instances
    for (int a = 0; a < 2; ++a) {
        pvd_axis_d &ax = v.ax[a];
        const int tk = a ? tky : tkx;
        const bool sq = ax.n == 256 && ax.foldn == 0 && tk == 16;
        ax.reg = regs ? 1 : 0;
        ax.ra = ax.ra_log = ax.rb = 0;
        if (regs && !sq && xy_two_step_len(ax.n)) {
            ax.reg = 2;
            ax.ra = 16;
            ax.ra_log = 4;
            ax.rb = ax.n / 16;
        }
    }
    // room for the padded exchange array of the register-resident 256-point path
    if (Py == 256 && tky == 16) ldsy += PVD_SQ256_EXTRA_LDS;
    if (Px == 256 && tkx == 16) ldsx += PVD_SQ256_EXTRA_LDS;
    // the y and x passes move 16-byte column pairs and read whole tiles of columns
    if (tky < 2 || tkx < 2) return fail(PVD_ENOTSUP, "FFT length too large for a two-column tile");
    // z two-step P = A * B for 16-line tiles whose block covers 16*A threads in step 2
    // and 16*B items in two rounds in step 1
    pvd_axis_d &az = v.ax[2];
    az.reg = 0;
    az.ra = az.ra_log = az.rb = 0;
    for (int c = 1; regs && lz == 16 && c < PVD_ZREG_CODES; ++c) {
        const int A = kZregA[c], B = kZregB[c];
        if (A * B == Pz && 16 * A <= bz && 16 * B <= 2 * bz) {
            az.reg = c;
            az.ra = A;
            az.ra_log = ilog2(A);
            az.rb = B;
            break;
        }
    }
    // the two-step instances run blocks of at most PVD_XY_MAX_BLOCK threads
    if (v.ax[1].reg == 2) by = std::min(by, PVD_XY_MAX_BLOCK);
    if (v.ax[0].reg == 2) bx = std::min(bx, PVD_XY_MAX_BLOCK);
    // rows padded to 128 bytes and to whole column tiles
    v.Hp = round_up(v.H, std::max(16, std::max(tky, tkx)));
    v.llz = ilog2(lz);
    v.ltky = ilog2(tky);
    v.ltkx = ilog2(tkx);
    // x-planes of the workspace are Py * Hp complex plus PVD_PLANE_PAD (0: padding by one
    // or three 128-byte lines measured the same on the CT workload)
    v.wplane = Py * v.Hp + PVD_PLANE_PAD;
    // 32-bit offsets of the register X passes (the LDS-resident two-step and the generic
    // pass address with 64 bits)
    {
        pvd_axis_d &ax = v.ax[0];
        const bool reg32 = (ax.reg == 1 && Px == 256) || (ax.reg == 2 && ax.rb < PVD_XY_LDS_B);
        if (reg32 && !x_offsets_fit(Nx, Px, Py, v.Hp, v.wplane)) ax.reg = 0, ax.ra = ax.ra_log = ax.rb = 0;
    }
    const long long nlines = (long long)Nx * Ny;
    const long long tz = (nlines + 2 * lz - 1) / (2 * lz);
    const long long ty = (long long)Nx * ((v.H + tky - 1) / tky);
    const long long tx = (long long)Py * ((v.H + tkx - 1) / tkx);
    if (tz > 0x7fffffffLL || ty > 0x7fffffffLL || tx > 0x7ffffff0LL) return fail(PVD_ENOTSUP, "grid too large");
    p->cfg[PVD_PASS_FZ] = {(int)tz, bz, ldsz};
    p->cfg[PVD_PASS_IZ] = {(int)tz, bz, ldsz};
    p->cfg[PVD_PASS_FY] = {(int)ty, by, ldsy};
    p->cfg[PVD_PASS_IY] = {(int)ty, by, ldsy};
    p->cfg[PVD_PASS_X] = {(int)tx, bx, ldsx};
    for (int i = 0; i < PVD_NPASS_FFT; ++i) v.tiles[i] = p->cfg[i].grid;
    p->work_elems = (size_t)Nx * v.wplane;
    p->spec_elems = (size_t)Px * Py * v.Hp;
    p->bytes = sizeof(float2) * (p->work_elems + p->spec_elems);
    return PVD_OK;
}

int fft_alloc(pvd_plan *p) {
    pvd_vol_d &v = p->vol;
    for (int a = 0; a < 3; ++a) {
        const float2 *tw = nullptr;
        int rc = get_twiddles(p->dev, p->P[a], &tw);
        if (rc) return rc;
        v.ax[a].tw = tw;
    }
    PVD_HIP(hipMalloc(&p->work, sizeof(float2) * p->work_elems), "allocate FFT workspace");
    PVD_HIP(hipMalloc(&p->spec, sizeof(float2) * p->spec_elems), "allocate kernel spectrum");
    v.work = p->work;
    v.spec = p->spec;
    return PVD_OK;
}

bool aligned16(const void *ptr) { return (reinterpret_cast<uintptr_t>(ptr) & 15) == 0; }

double fft_cost_model(const pvd_plan *p) {
    double vol = (double)p->N[0] * p->N[1] * p->N[2];
    double padded = (double)p->N[0] * p->P[1] * (p->P[2] / 2 + 1) * 2.0;
    return (vol * 12.0 + padded * 4.0 * 7.0) / 5.0e12 + 5 * 3.0e-6;
}

double direct_cost_model(const pvd_plan *p) {
    double vol = (double)p->N[0] * p->N[1] * p->N[2];
    return vol * 2.0 * p->Ku[0] * p->Ku[1] * p->Ku[2] / 40.0e12 + 3.0e-6;
}

int launch_fft(pvd_plan *p, const pvd_vol_d &v, hipStream_t st) {
    const bool prof = (p->prm.flags & PVD_FLAG_PROFILE) != 0;
    if (prof) PVD_HIP(hipEventRecord(p->ev[0], st), "record event");
    for (int pass = 0; pass < PVD_NPASS_FFT; ++pass) {
        const launch_cfg &c = p->cfg[pass];
        PVD_HIP(pvd_launch_fft_pass(pass, v, c.grid, c.block, c.lds, st), "launch FFT pass");
        if (prof) PVD_HIP(hipEventRecord(p->ev[pass + 1], st), "record event");
    }
    p->profiled_call = prof;
    return PVD_OK;
}

// the plan's descriptor template completed with one call's inputs
void fill_desc(const pvd_plan *p, const float *const *acts, const float *w, int T, const float *dens, float *out,
               double scale, pvd_vol_d *vd) {
    pvd_vol_d &v = *vd;
    v = p->vol;
    v.T = T;
    for (int t = 0; t < PVD_MAX_TIMEPOINTS; ++t) {
        v.act[t] = t < T ? acts[t] : nullptr;
        v.w[t] = t < T ? w[t] : 0.f;
    }
    v.dens = dens;
    v.out = out;
    v.scale = (float)scale;
    // float4 access when every z-line starts 16-byte aligned (1), else float4 runs of a
    // tile's contiguous lines when the activity maps are 16-byte aligned and a tile's
    // first line (a multiple of 2 * pairs per tile) starts a multiple of 4 floats in (2)
    bool aligned = true;
    for (int t = 0; t < T; ++t) aligned = aligned && aligned16(acts[t]);
    const bool vec = aligned && p->N[2] % 4 == 0 && aligned16(out) && aligned16(dens);
    v.vec_io = vec ? 1 : (aligned && p->vol.llz >= 1) ? 2 : 0;
}

int check_inputs(const pvd_plan *p, const float *const *acts, int T, const float *out, double scale) {
    if (!p) return fail(PVD_EINVAL, "null plan");
    if (!p->has_kernel) return fail(PVD_EINVAL, "plan has no kernel: call pvd_plan_set_kernel first");
    if (T < 1 || T > PVD_MAX_TIMEPOINTS) return fail(PVD_EINVAL, "T=%d outside [1, %d]", T, PVD_MAX_TIMEPOINTS);
    if (!out) return fail(PVD_EINVAL, "null dose pointer");
    for (int t = 0; t < T; ++t)
        if (!acts[t]) return fail(PVD_EINVAL, "null activity pointer %d", t);
    if (!isfinite(scale)) return fail(PVD_EINVAL, "non-finite scale");
    return PVD_OK;
}

int run_dose(pvd_plan *p, const float *const *acts, const float *w, int T, const float *dens, float *out,
             double scale, hipStream_t st) {
    if (int rc = check_inputs(p, acts, T, out, scale)) return rc;
    PVD_HIP(hipSetDevice(p->dev), "set device");
    if (p->algo == PVD_ALGO_DIRECT) {
        const bool prof = (p->prm.flags & PVD_FLAG_PROFILE) != 0;
        const int sh[3] = {p->prm.mode == PVD_MODE_LINEAR ? p->K[0] / 2 - p->lo[0] : 0,
                           p->prm.mode == PVD_MODE_LINEAR ? p->K[1] / 2 - p->lo[1] : 0,
                           p->prm.mode == PVD_MODE_LINEAR ? p->K[2] / 2 - p->lo[2] : 0};
        if (prof) PVD_HIP(hipEventRecord(p->ev[0], st), "record event");
        PVD_HIP(pvd_launch_direct(p->N, p->Ku, sh, p->prm.mode == PVD_MODE_REFERENCE, acts, w, T, p->kr, dens,
                                  out, (float)scale, st),
                "launch direct convolution");
        if (prof) PVD_HIP(hipEventRecord(p->ev[1], st), "record event");
        p->profiled_call = prof;
        return PVD_OK;
    }
    pvd_vol_d v;
    fill_desc(p, acts, w, T, dens, out, scale, &v);
    return launch_fft(p, v, st);
}

// Kernel instance and block of one pass of a plan, as one number: volumes with
// equal keys share a batch launch.
long long batch_instance_key(const pvd_plan *p, int pass) {
    const pvd_vol_d &v = p->vol;
    long long inst;
    if (pass == PVD_PASS_FZ || pass == PVD_PASS_IZ) inst = (v.ax[2].reg != 0 ? 1 : 0) + 2 * v.llz;
    else {
        const pvd_axis_d &ax = v.ax[pass == PVD_PASS_X ? 0 : 1];
        inst = ax.reg == 2 ? ax.rb : 0;
    }
    return inst * 4096 + p->cfg[pass].block;
}

int check_params(const pvd_plan_params *q) {
    for (int a = 0; a < 3; ++a) {
        if (q->n[a] < 1) return fail(PVD_EINVAL, "activity grid extent %d on axis %d", q->n[a], a);
        if (q->k[a] < 1) return fail(PVD_EINVAL, "kernel grid extent %d on axis %d", q->k[a], a);
        if (q->fft_n[a] < 0) return fail(PVD_EINVAL, "negative fft_n");
    }
    if (q->mode != PVD_MODE_REFERENCE && q->mode != PVD_MODE_LINEAR) return fail(PVD_EINVAL, "bad mode %d", q->mode);
    if (q->algo < PVD_ALGO_AUTO || q->algo > PVD_ALGO_DIRECT) return fail(PVD_EINVAL, "bad algo %d", q->algo);
    for (int a = 0; a < 3; ++a) {
        const int t = q->tile[a];
        if (t < 0 || t > PVD_TILE_MAX || (t & (t - 1)))
            return fail(PVD_EINVAL, "tile[%d]=%d must be 0 or a power of two <= %d", a, t, PVD_TILE_MAX);
    }
    if (q->max_radix != 0 && q->max_radix != 8 && q->max_radix != 16)
        return fail(PVD_EINVAL, "max_radix must be 0, 8 or 16 (got %d)", q->max_radix);
    return PVD_OK;
}

void plan_init(pvd_plan *p, const pvd_plan_params *q) {
    p->prm = *q;
    for (int a = 0; a < 3; ++a) p->tile[a] = q->tile[a] ? q->tile[a] : PVD_TILE_DEFAULT;
    p->max_radix = q->max_radix ? q->max_radix : 16;
}

// Host-only: FFT lengths, algorithm, and (FFT) the layout and launch shapes.
int plan_layout(pvd_plan *p, size_t max_lds) {
    if (int rc = axis_setup(p)) return rc;
    const bool direct_ok = p->Ku[0] <= 16 && p->Ku[1] <= 16 && p->Ku[2] <= 16 &&
                           pvd_direct_lds_bytes(p->Ku) <= max_lds;
    int algo = p->prm.algo;
    if (algo == PVD_ALGO_DIRECT && !direct_ok)
        return fail(PVD_ENOTSUP, "direct path needs kernel extents <= 16 (got %dx%dx%d)", p->Ku[0], p->Ku[1],
                    p->Ku[2]);
    if (algo == PVD_ALGO_AUTO)
        algo = (direct_ok && direct_cost_model(p) < fft_cost_model(p)) ? PVD_ALGO_DIRECT : PVD_ALGO_FFT;
    p->algo = algo;
    if (algo == PVD_ALGO_FFT) return fft_layout(p);
    p->bytes = sizeof(float) * (size_t)p->Ku[0] * p->Ku[1] * p->Ku[2];
    return PVD_OK;
}

int axis_path(const pvd_plan *p, int a) {
    const pvd_axis_d &ax = p->vol.ax[a];
    if (a == 2) return ax.reg ? PVD_PATH_TWO_STEP : PVD_PATH_STOCKHAM;
    if (ax.reg == 2) return PVD_PATH_TWO_STEP;
    const int lines = 1 << (a == 0 ? p->vol.ltkx : p->vol.ltky);
    const int block = p->cfg[a == 0 ? PVD_PASS_X : PVD_PASS_FY].block;
    // the conditions the kernels check before taking the 256-point path (sq256_ok)
    if (ax.reg && ax.n == 256 && lines == 16 && block == 256 && ax.foldn == 0) return PVD_PATH_SQ256;
    return PVD_PATH_STOCKHAM;
}

void fill_info(const pvd_plan *p, pvd_plan_info *info) {
    memset(info, 0, sizeof *info);
    info->algo = p->algo;
    info->symmetric_spectrum = p->vol.sym;
    const bool fft = p->algo == PVD_ALGO_FFT;
    for (int a = 0; a < 3; ++a) {
        info->fft_n[a] = fft ? p->P[a] : 0;
        info->kernel_used[a] = p->Ku[a];
        info->fft_path[a] = fft ? axis_path(p, a) : 0;
    }
    info->n_passes = fft ? PVD_NPASS_FFT : 1;
    info->workspace_bytes = p->bytes;
    info->tile[0] = fft ? 1 << p->vol.ltkx : 0;
    info->tile[1] = fft ? 1 << p->vol.ltky : 0;
    info->tile[2] = fft ? 1 << p->vol.llz : 0;
    info->max_radix = fft ? p->max_radix : 0;
    for (int i = 0; fft && i < PVD_NPASS_FFT; ++i) {
        info->block[i] = p->cfg[i].block;
        info->grid[i] = p->cfg[i].grid;
    }
}

}  // namespace

// =================================================================================
extern "C" {

const char *pvd_last_error(void) { return g_err.c_str(); }

int pvd_version(void) { return PVD_VERSION; }

int pvd_dpk_generate(double voxel_mm, int gx, int gy, int gz, const pvd_dpk_term *terms, int nterms,
                     double factor, double *out, void *stream) {
    if (gx < 1 || gy < 1 || gz < 1) return fail(PVD_EINVAL, "kernel grid must be positive");
    if (nterms < 0 || nterms > 16 || (nterms > 0 && !terms)) return fail(PVD_EINVAL, "bad term list");
    if (!out) return fail(PVD_EINVAL, "null output");
    for (int i = 0; i < nterms; ++i)
        if (terms[i].kind != PVD_DPK_BETA && terms[i].kind != PVD_DPK_PHOTON)
            return fail(PVD_EINVAL, "unknown term kind %d", terms[i].kind);
    const int g[3] = {gx, gy, gz};
    PVD_HIP(pvd_launch_dpk(voxel_mm, g, terms, nterms, factor, out, (hipStream_t)stream), "launch DPK generator");
    return PVD_OK;
}

int pvd_plan_create(pvd_plan **out, const pvd_plan_params *q) {
    if (!out || !q) return fail(PVD_EINVAL, "null argument");
    *out = nullptr;
    if (int rc = check_params(q)) return rc;
    pvd_plan *p = new (std::nothrow) pvd_plan();
    if (!p) return fail(PVD_ENOMEM, "host allocation");
    plan_init(p, q);
    int rc = PVD_OK;
    do {
        if (hipGetDevice(&p->dev) != hipSuccess) { rc = fail(PVD_EHIP, "no HIP device"); break; }
        size_t max_lds = 0;
        if ((rc = ensure_kernel_attrs(p->dev, &max_lds))) break;
        if ((rc = plan_layout(p, max_lds))) break;
        if (p->algo == PVD_ALGO_FFT) {
            if ((rc = fft_alloc(p))) break;
        } else {
            if (hipMalloc(&p->kr, p->bytes) != hipSuccess) { rc = fail(PVD_ENOMEM, "allocate kernel"); break; }
        }
        if (q->flags & PVD_FLAG_PROFILE) {
            const int nev = (p->algo == PVD_ALGO_FFT) ? PVD_NPASS_FFT + 1 : 2;
            for (int i = 0; i < nev; ++i)
                if (hipEventCreate(&p->ev[i]) != hipSuccess) { rc = fail(PVD_EHIP, "create event"); break; }
            p->n_ev = nev;
        }
    } while (0);
    if (rc) {
        pvd_plan_destroy(p);
        return rc;
    }
    *out = p;
    return PVD_OK;
}

int pvd_plan_describe(const pvd_plan_params *q, pvd_plan_info *info) {
    if (!q || !info) return fail(PVD_EINVAL, "null argument");
    if (int rc = check_params(q)) return rc;
    pvd_plan *p = new (std::nothrow) pvd_plan();
    if (!p) return fail(PVD_ENOMEM, "host allocation");
    plan_init(p, q);
    int rc = plan_layout(p, kDescribeLds);
    if (!rc) fill_info(p, info);
    delete p;
    return rc;
}

int pvd_plan_info_get(const pvd_plan *p, pvd_plan_info *info) {
    if (!p || !info) return fail(PVD_EINVAL, "null argument");
    fill_info(p, info);
    return PVD_OK;
}

int pvd_plan_destroy(pvd_plan *p) {
    if (!p) return PVD_OK;
    if (p->work) (void)hipFree(p->work);
    if (p->spec) (void)hipFree(p->spec);
    if (p->kr) (void)hipFree(p->kr);
    for (int i = 0; i < p->n_ev; ++i)
        if (p->ev[i]) (void)hipEventDestroy(p->ev[i]);
    delete p;
    return PVD_OK;
}

int pvd_plan_set_kernel(pvd_plan *p, const double *kernel, void *stream) {
    if (!p || !kernel) return fail(PVD_EINVAL, "null argument");
    hipStream_t st = (hipStream_t)stream;
    PVD_HIP(hipSetDevice(p->dev), "set device");
    if (p->algo == PVD_ALGO_DIRECT) {
        const size_t nk = (size_t)p->K[0] * p->K[1] * p->K[2];
        std::vector<double> h(nk);
        PVD_HIP(hipMemcpyAsync(h.data(), kernel, sizeof(double) * nk, hipMemcpyDeviceToHost, st), "download kernel");
        PVD_HIP(hipStreamSynchronize(st), "synchronise");
        const int *Ku = p->Ku, *lo = p->lo;
        std::vector<float> r((size_t)Ku[0] * Ku[1] * Ku[2]);
        for (int i = 0; i < Ku[0]; ++i)
            for (int j = 0; j < Ku[1]; ++j)
                for (int l = 0; l < Ku[2]; ++l) {
                    const size_t src = ((size_t)(lo[0] + Ku[0] - 1 - i) * p->K[1] + (lo[1] + Ku[1] - 1 - j)) * p->K[2] +
                                       (lo[2] + Ku[2] - 1 - l);
                    r[((size_t)i * Ku[1] + j) * Ku[2] + l] = (float)h[src];
                }
        PVD_HIP(hipMemcpyAsync(p->kr, r.data(), sizeof(float) * r.size(), hipMemcpyHostToDevice, st), "upload kernel");
        PVD_HIP(hipStreamSynchronize(st), "synchronise");
        p->has_kernel = true;
        return PVD_OK;
    }
    // An odd, uncropped kernel that is mirror-symmetric on every axis, placed with its
    // centre at the FFT origin, has a real spectrum even in kx and ky: keep only that
    // octant (1/16 of the complex bytes) and recover the reference's origin-anchored
    // result by the output roll (off = -K//2 mod P).
    bool sym = !(p->prm.flags & PVD_FLAG_GENERAL_SPECTRUM);
    for (int a = 0; a < 3 && sym; ++a) sym = (p->K[a] & 1) && p->Ku[a] == p->K[a] && p->lo[a] == 0;
    if (sym) {
        const size_t nk = (size_t)p->K[0] * p->K[1] * p->K[2];
        std::vector<double> h(nk);
        PVD_HIP(hipMemcpyAsync(h.data(), kernel, sizeof(double) * nk, hipMemcpyDeviceToHost, st), "download kernel");
        PVD_HIP(hipStreamSynchronize(st), "synchronise");
        const int K0 = p->K[0], K1 = p->K[1], K2 = p->K[2];
        for (int i = 0; i < K0 && sym; ++i)
            for (int j = 0; j < K1 && sym; ++j)
                for (int l = 0; l < K2 && sym; ++l) {
                    const double v = h[((size_t)i * K1 + j) * K2 + l];
                    sym = v == h[((size_t)(K0 - 1 - i) * K1 + j) * K2 + l] &&
                          v == h[((size_t)i * K1 + (K1 - 1 - j)) * K2 + l] &&
                          v == h[((size_t)i * K1 + j) * K2 + (K2 - 1 - l)];
                }
    }
    const bool ref = p->prm.mode == PVD_MODE_REFERENCE;
    for (int a = 0; a < 3; ++a) {
        if (ref) p->s[a] = sym ? p->K[a] / 2 : 0;
        const int P = p->P[a];
        p->vol.ax[a].off = ref ? ((P - (p->s[a] % P)) % P) : 0;
    }
    p->vol.sym = sym ? 1 : 0;
    const int H = p->P[2] / 2 + 1;
    // FP64 scratch of the separable spectrum build: inside the plan's workspace when it
    // fits (the workspace holds nothing between dose calls, and this call is ordered on
    // `st` and synchronises before returning), else allocated for the call
    double2 *T1 = nullptr, *T2 = nullptr;
    const size_t n1 = (size_t)p->Ku[0] * p->Ku[1] * H, n2 = (size_t)p->Ku[0] * p->P[1] * H;
    const bool in_work = sizeof(double2) * (n1 + n2) <= sizeof(float2) * p->work_elems;
    hipError_t e = hipSuccess;
    if (in_work) {
        T1 = reinterpret_cast<double2 *>(p->vol.work);
        T2 = T1 + n1;
    } else {
        PVD_HIP(hipMalloc(&T1, sizeof(double2) * n1), "allocate spectrum scratch");
        e = hipMalloc(&T2, sizeof(double2) * n2);
        if (e != hipSuccess) { (void)hipFree(T1); return hip_fail(e, "allocate spectrum scratch"); }
    }
    e = pvd_launch_spectrum(p->P, p->lo, p->Ku, p->s, p->K, p->vol.Hp, sym ? 1 : 0, kernel, T1, T2, p->spec, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (!in_work) {
        (void)hipFree(T1);
        (void)hipFree(T2);
    }
    if (e != hipSuccess) return hip_fail(e, "kernel spectrum");
    p->has_kernel = true;
    return PVD_OK;
}

int pvd_conv_dose(pvd_plan *p, const float *act, const float *density, float *dose, double scale, void *stream) {
    const float *acts[1] = {act};
    const float w[1] = {1.0f};
    return run_dose(p, acts, w, 1, density, dose, scale, (hipStream_t)stream);
}

int pvd_conv_dose_weighted(pvd_plan *p, const float *const *acts, const double *weights, int T,
                           const float *density, float *dose, double scale, void *stream) {
    if (!acts || !weights) return fail(PVD_EINVAL, "null argument");
    if (T < 1 || T > PVD_MAX_TIMEPOINTS) return fail(PVD_EINVAL, "T=%d outside [1, %d]", T, PVD_MAX_TIMEPOINTS);
    float w[PVD_MAX_TIMEPOINTS];
    for (int t = 0; t < T; ++t) {
        if (!isfinite(weights[t])) return fail(PVD_EINVAL, "non-finite weight %d", t);
        w[t] = (float)weights[t];
    }
    return run_dose(p, acts, w, T, density, dose, scale, (hipStream_t)stream);
}

int pvd_conv_dose_timeseries(pvd_plan *p, const float *const *acts, const double *t_hours, int T,
                             const float *density, float *dose, double scale, void *stream) {
    if (!acts || !t_hours) return fail(PVD_EINVAL, "null argument");
    if (T < 1 || T > PVD_MAX_TIMEPOINTS) return fail(PVD_EINVAL, "T=%d outside [1, %d]", T, PVD_MAX_TIMEPOINTS);
    // trapezoid weights, kernel_convolution.py:101-104 (dt in hours * 3600)
    double w[PVD_MAX_TIMEPOINTS] = {0};
    for (int i = 0; i + 1 < T; ++i) {
        const double dt = (t_hours[i + 1] - t_hours[i]) * 3600.0;
        w[i] += dt / 2.0;
        w[i + 1] += dt / 2.0;
    }
    return pvd_conv_dose_weighted(p, acts, w, T, density, dose, scale, stream);
}

int pvd_weighted_sum(const float *const *maps, const double *weights, int T, size_t n, float *out,
                     void *stream) {
    if (!maps || !weights || !out) return fail(PVD_EINVAL, "null argument");
    if (T < 1 || T > PVD_MAX_TIMEPOINTS) return fail(PVD_EINVAL, "T=%d outside [1, %d]", T, PVD_MAX_TIMEPOINTS);
    float w[PVD_MAX_TIMEPOINTS];
    for (int t = 0; t < T; ++t) {
        if (!maps[t]) return fail(PVD_EINVAL, "null map %d", t);
        if (!isfinite(weights[t])) return fail(PVD_EINVAL, "non-finite weight %d", t);
        w[t] = (float)weights[t];
    }
    if (n == 0) return PVD_OK;
    PVD_HIP(pvd_launch_wsum(maps, w, T, n, out, (hipStream_t)stream), "launch weighted sum");
    return PVD_OK;
}

int pvd_plan_timings(pvd_plan *p, float *ms, int n) {
    if (!p) return fail(PVD_EINVAL, "null plan");
    if (!p->profiled_call) return fail(PVD_EINVAL, "no profiled call recorded (create the plan with PVD_FLAG_PROFILE)");
    const int np = p->n_ev - 1;
    PVD_HIP(hipEventSynchronize(p->ev[np]), "synchronise event");
    for (int i = 0; i < np && i < n; ++i)
        PVD_HIP(hipEventElapsedTime(&ms[i], p->ev[i], p->ev[i + 1]), "event time");
    return np;
}

// ---------------------------------------------------------------------------------
// batches
int pvd_batch_destroy(pvd_batch *b) {
    if (!b) return PVD_OK;
    if (b->uploaded) (void)hipEventSynchronize(b->uploaded);
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->cap) (void)hipStreamDestroy(b->cap);
    for (hipStream_t f : b->fork)
        if (f) (void)hipStreamDestroy(f);
    for (hipEvent_t j : b->join)
        if (j) (void)hipEventDestroy(j);
    if (b->uploaded) (void)hipEventDestroy(b->uploaded);
    if (b->d_tab) (void)hipFree(b->d_tab);
    if (b->d_pool) (void)hipFree(b->d_pool);
    if (b->h_tab) (void)hipHostFree(b->h_tab);
    delete b;
    return PVD_OK;
}

int pvd_batch_create(pvd_batch **out, pvd_plan *const *plans, int nvol) {
    if (!out || !plans || nvol < 1) return fail(PVD_EINVAL, "need at least one plan");
    *out = nullptr;
    for (int i = 0; i < nvol; ++i) {
        if (!plans[i]) return fail(PVD_EINVAL, "null plan %d", i);
        if (plans[i]->algo != PVD_ALGO_FFT) return fail(PVD_EINVAL, "plan %d is not an FFT plan", i);
        if (!plans[i]->has_kernel) return fail(PVD_EINVAL, "plan %d has no kernel", i);
        if (plans[i]->dev != plans[0]->dev) return fail(PVD_EINVAL, "plans on different devices");
        for (int j = 0; j < i; ++j)
            if (plans[j] == plans[i]) return fail(PVD_EINVAL, "plan %d repeated (volumes need their own workspace)", i);
    }
    pvd_batch *b = new (std::nothrow) pvd_batch();
    if (!b) return fail(PVD_ENOMEM, "host allocation");
    b->plans.assign(plans, plans + nvol);
    b->nvol = nvol;
    b->dev = plans[0]->dev;
    b->ready.assign(nvol, 0);
    int rc = PVD_OK;
    do {
        // Each pass launches once per group of volumes that run the same kernel instance
        // with the same block as their own plans (so a batched volume gives the bits of
        // its single-volume plan); a group's grid concatenates its members' tiles.
        std::vector<int> bases, vols;
        for (int pass = 0; pass < PVD_NPASS_FFT && !rc; ++pass) {
            std::vector<std::vector<int>> groups;
            std::vector<long long> keys;
            for (int i = 0; i < nvol; ++i) {
                const long long key = batch_instance_key(plans[i], pass);
                size_t g = 0;
                while (g < keys.size() && keys[g] != key) ++g;
                if (g == keys.size()) { keys.push_back(key); groups.emplace_back(); }
                groups[g].push_back(i);
            }
            for (const auto &grp : groups) {
                batch_launch_d L;
                L.pass = pass;
                L.exemplar = grp[0];
                L.first = (int)bases.size();
                L.nmem = (int)grp.size();
                long long total = 0;
                for (int i : grp) {
                    const launch_cfg &pc = plans[i]->cfg[pass];
                    L.cfg.block = pc.block;
                    L.cfg.lds = std::max(L.cfg.lds, pc.lds);
                    bases.push_back((int)total);
                    vols.push_back(i);
                    total += pc.grid;
                }
                if (total > 0x7ffffff0LL) { rc = fail(PVD_ENOTSUP, "batch grid too large"); break; }
                L.cfg.grid = (int)total;
                b->launches.push_back(L);
            }
        }
        if (rc) break;
        b->pool_n = (int)bases.size();
        if (hipSetDevice(b->dev) != hipSuccess) { rc = fail(PVD_EHIP, "set device"); break; }
        if (hipHostMalloc((void **)&b->h_tab, sizeof(pvd_vol_d) * nvol, hipHostMallocDefault) != hipSuccess) {
            rc = fail(PVD_ENOMEM, "pinned descriptor table");
            break;
        }
        if (hipMalloc(&b->d_tab, sizeof(pvd_vol_d) * nvol) != hipSuccess) {
            rc = fail(PVD_ENOMEM, "device descriptor table");
            break;
        }
        if (hipMalloc(&b->d_pool, sizeof(int) * 2 * b->pool_n) != hipSuccess) {
            rc = fail(PVD_ENOMEM, "device member tables");
            break;
        }
        std::vector<int> pool(bases);
        pool.insert(pool.end(), vols.begin(), vols.end());
        if (hipMemcpy(b->d_pool, pool.data(), sizeof(int) * pool.size(), hipMemcpyHostToDevice) != hipSuccess) {
            rc = fail(PVD_EHIP, "upload member tables");
            break;
        }
        if (hipEventCreateWithFlags(&b->uploaded, hipEventDisableTiming) != hipSuccess) {
            rc = fail(PVD_EHIP, "create event");
            break;
        }
        for (int i = 0; i < nvol; ++i) b->h_tab[i] = plans[i]->vol;
    } while (0);
    if (rc) {
        pvd_batch_destroy(b);
        return rc;
    }
    *out = b;
    return PVD_OK;
}

int pvd_batch_set_volume(pvd_batch *b, int i, const float *const *acts, const double *weights, int T,
                         const float *density, float *dose, double scale) {
    if (!b) return fail(PVD_EINVAL, "null batch");
    if (i < 0 || i >= b->nvol) return fail(PVD_EINVAL, "volume index %d outside [0, %d)", i, b->nvol);
    if (!acts || !weights) return fail(PVD_EINVAL, "null argument");
    if (T < 1 || T > PVD_MAX_TIMEPOINTS) return fail(PVD_EINVAL, "T=%d outside [1, %d]", T, PVD_MAX_TIMEPOINTS);
    float w[PVD_MAX_TIMEPOINTS];
    for (int t = 0; t < T; ++t) {
        if (!isfinite(weights[t])) return fail(PVD_EINVAL, "non-finite weight %d", t);
        w[t] = (float)weights[t];
    }
    pvd_plan *p = b->plans[i];
    if (int rc = check_inputs(p, acts, T, dose, scale)) return rc;
    if (b->upload_pending) {   // the previous upload still reads the host table
        PVD_HIP(hipEventSynchronize(b->uploaded), "wait for descriptor upload");
        b->upload_pending = false;
    }
    fill_desc(p, acts, w, T, density, dose, scale, &b->h_tab[i]);
    b->ready[i] = 1;
    return PVD_OK;
}

namespace {
int batch_upload(pvd_batch *b, hipStream_t st) {
    for (int i = 0; i < b->nvol; ++i)
        if (!b->ready[i]) return fail(PVD_EINVAL, "volume %d has no inputs (pvd_batch_set_volume)", i);
    PVD_HIP(hipSetDevice(b->dev), "set device");
    PVD_HIP(hipMemcpyAsync(b->d_tab, b->h_tab, sizeof(pvd_vol_d) * b->nvol, hipMemcpyHostToDevice, st),
            "upload descriptor table");
    PVD_HIP(hipEventRecord(b->uploaded, st), "record event");
    b->upload_pending = true;
    return PVD_OK;
}

// One pass's groups in the captured graph: branch k takes groups k, k + K, ...; the
// groups of a pass run concurrently (each group's tail overlaps the others), and the
// next pass starts when all of them are done.
int batch_capture(pvd_batch *b) {
    constexpr int K = PVD_BATCH_BRANCHES;
    size_t i = 0;
    while (i < b->launches.size()) {
        const int pass = b->launches[i].pass;
        PVD_HIP(hipEventRecord(b->join[K], b->cap), "record pass start");
        for (int k = 0; k < K; ++k) PVD_HIP(hipStreamWaitEvent(b->fork[k], b->join[K], 0), "fork");
        int k = 0;
        for (; i < b->launches.size() && b->launches[i].pass == pass; ++i, k = (k + 1) % K) {
            const batch_launch_d &L = b->launches[i];
            pvd_batch_args a;
            a.vols = b->d_tab;
            a.base = b->d_pool + L.first;
            a.vol = b->d_pool + b->pool_n + L.first;
            a.nmem = L.nmem;
            a.ntiles = L.cfg.grid;
            PVD_HIP(pvd_launch_fft_pass_batch(L.pass, b->plans[L.exemplar]->vol, a, L.cfg.grid, L.cfg.block,
                                              L.cfg.lds, b->fork[k]),
                    "launch batch FFT pass");
        }
        for (int j = 0; j < K; ++j) {
            PVD_HIP(hipEventRecord(b->join[j], b->fork[j]), "record branch end");
            PVD_HIP(hipStreamWaitEvent(b->cap, b->join[j], 0), "join");
        }
    }
    return PVD_OK;
}

int batch_launch(pvd_batch *b, hipStream_t st) {
    for (const batch_launch_d &L : b->launches) {
        pvd_batch_args a;
        a.vols = b->d_tab;
        a.base = b->d_pool + L.first;
        a.vol = b->d_pool + b->pool_n + L.first;
        a.nmem = L.nmem;
        a.ntiles = L.cfg.grid;
        PVD_HIP(pvd_launch_fft_pass_batch(L.pass, b->plans[L.exemplar]->vol, a, L.cfg.grid, L.cfg.block, L.cfg.lds,
                                          st),
                "launch batch FFT pass");
    }
    return PVD_OK;
}
}  // namespace

int pvd_batch_run(pvd_batch *b, void *stream) {
    if (!b) return fail(PVD_EINVAL, "null batch");
    hipStream_t st = (hipStream_t)stream;
    if (int rc = batch_upload(b, st)) return rc;
    return batch_launch(b, st);
}

int pvd_batch_run_graph(pvd_batch *b, void *stream) {
    if (!b) return fail(PVD_EINVAL, "null batch");
    hipStream_t st = (hipStream_t)stream;
    if (!b->exec) {
        PVD_HIP(hipSetDevice(b->dev), "set device");
        if (!b->cap) PVD_HIP(hipStreamCreateWithFlags(&b->cap, hipStreamNonBlocking), "create capture stream");
        for (int k = 0; k < PVD_BATCH_BRANCHES; ++k)
            if (!b->fork[k]) PVD_HIP(hipStreamCreateWithFlags(&b->fork[k], hipStreamNonBlocking), "create branch stream");
        for (int k = 0; k <= PVD_BATCH_BRANCHES; ++k)
            if (!b->join[k]) PVD_HIP(hipEventCreateWithFlags(&b->join[k], hipEventDisableTiming), "create event");
        PVD_HIP(hipStreamBeginCapture(b->cap, hipStreamCaptureModeThreadLocal), "begin capture");
        int rc = batch_capture(b);
        hipGraph_t g = nullptr;
        hipError_t e = hipStreamEndCapture(b->cap, &g);
        if (rc) {
            if (g) (void)hipGraphDestroy(g);
            return rc;
        }
        if (e != hipSuccess) return hip_fail(e, "end capture");
        e = hipGraphInstantiate(&b->exec, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        if (e != hipSuccess) {
            b->exec = nullptr;
            return hip_fail(e, "instantiate graph");
        }
    }
    if (int rc = batch_upload(b, st)) return rc;
    PVD_HIP(hipGraphLaunch(b->exec, st), "launch graph");
    return PVD_OK;
}

const char *pvd_pass_name(const pvd_plan *p, int i) {
    static const char *fft[PVD_NPASS_FFT] = {"fft_z_forward", "fft_y_forward", "fft_x_spectrum_multiply",
                                             "fft_y_inverse", "fft_z_inverse_epilogue"};
    if (!p) return "";
    if (p->algo == PVD_ALGO_DIRECT) return i == 0 ? "direct_conv" : "";
    return (i >= 0 && i < PVD_NPASS_FFT) ? fft[i] : "";
}

}  // extern "C"
