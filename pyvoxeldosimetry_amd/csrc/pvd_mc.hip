// Voxel Monte Carlo of beta-electron energy deposition (gfx950).
//
// This is the project's own transport model, used to generate heterogeneous-
// medium dose targets for the learned correction (the reference's GPU Monte
// Carlo, core/monte_carlo.py:10-173, is a non-functional CuPy placeholder):
//  * stratified sources: every voxel with activity A emits `histories` electrons
//    of weight A/histories from uniform positions inside the voxel;
//  * energy from the nuclide's beta spectrum (inverse-CDF table);
//  * continuous slowing down along straight segments: the residual range r is
//    tracked in water-equivalent mm (Katz-Penfold range-energy relation, tabulated
//    as E(r)), a step of water-equivalent length s covers s/rho mm of the local
//    medium (density scaling), so the voxel map of deposits is invariant under
//    (rho, voxel) -> (c rho, voxel / c);
//  * the energy lost over a step, E(r) - E(r - s), is deposited in the voxel where
//    the step starts; steps are at most half a voxel long;
//  * multiple scattering: after each step the direction turns by a Gaussian polar
//    angle of width scatter * sqrt(s / R(E0)) about the old direction;
//  * electrons that leave the grid carry their residual energy out ("escaped");
//  * photons (gamma lines per decay, bremsstrahlung packets per electron) as kerma:
//    isotropic straight rays from the decay point, traversed voxel by voxel
//    (Amanatides & Woo, Eurographics 1987); over a chord of L mm in a voxel of
//    density rho the packet deposits E (1 - exp(-mu_en rho L)) of its energy E,
//    mu_en the water energy-absorption coefficient at its energy.
// Energy bookkeeping (emitted, deposited, escaped, MeV x weight) is reduced per
// block in FP64 and added with FP64 atomics; the deposit map uses FP64 atomics.
// Random numbers: Philox4x32-10 (Salmon et al., SC'11), counter = (history, draw,
// stream) with separate streams for the electron and the photons.
#include "pvd_internal.h"

#include "../../include/pvd_learned.h"

namespace {

struct Philox {
    uint32_t k0, k1;
    __device__ __forceinline__ static void round(uint32_t &c0, uint32_t &c1, uint32_t &c2, uint32_t &c3, uint32_t k0,
                                                 uint32_t k1) {
        const uint32_t M0 = 0xD2511F53u, M1 = 0xCD9E8D57u;
        const uint32_t h0 = __umulhi(M0, c0), l0 = M0 * c0;
        const uint32_t h1 = __umulhi(M1, c2), l1 = M1 * c2;
        const uint32_t n0 = h1 ^ c1 ^ k0, n2 = h0 ^ c3 ^ k1;
        c0 = n0; c1 = l1; c2 = n2; c3 = l0;
    }
    // 4 x 32 random bits for counter (a, b, c, d)
    __device__ __forceinline__ uint4 operator()(uint32_t a, uint32_t b, uint32_t c, uint32_t d) const {
        uint32_t c0 = a, c1 = b, c2 = c, c3 = d, x0 = k0, x1 = k1;
#pragma unroll
        for (int i = 0; i < 10; ++i) {
            round(c0, c1, c2, c3, x0, x1);
            x0 += 0x9E3779B9u;
            x1 += 0xBB67AE85u;
        }
        return make_uint4(c0, c1, c2, c3);
    }
};

// uniform in (0, 1]
__device__ __forceinline__ float u01(uint32_t x) { return ((float)(x >> 8) + 1.0f) * (1.0f / 16777216.0f); }

__device__ __forceinline__ float table_lerp(const float *t, int n, float x) {   // x in [0, n-1]
    x = fminf(fmaxf(x, 0.0f), (float)(n - 1));
    const int i = min((int)x, n - 2);
    const float f = x - (float)i;
    return fmaf(f, t[i + 1] - t[i], t[i]);
}

struct MCArgs {
    int n[3];
    float inv_voxel;          // 1 / voxel_mm
    float half_voxel_mm;
    int histories;
    uint32_t seed_lo, seed_hi;
    const float *e_of_u;      // inverse CDF of the beta spectrum, n_u points on u in [0, 1]
    int n_u;
    const float *e_of_r;      // E(r), n_r points on r in [0, r_max] (water-equivalent mm)
    int n_r;
    float r_max;
    float ka, kb, kc;         // Katz-Penfold: R(E) = 10 * ka * E^(kb - kc ln E) mm (water)
    float step_mm;
    float scatter;
    float rho_min;
    const long long *active;
    long long n_src;          // number of source voxels (active list length or all voxels)
    int n_lines;
    float line_e[PVD_MC_MAX_LINES], line_y[PVD_MC_MAX_LINES];
    float brems_z;
    int photon_every;
    int electrons;
    const float *mu_en;       // water mu_en/rho (cm^2/g) at n_mu log-spaced energies
    int n_mu;
    float mu_lo, mu_inv_step; // ln E of entry 0, entries per unit ln E
    float rho_min_ph;
};

// mu_en/rho of water at energy e (cm^2/g), log-spaced table, clamped at its ends
__device__ __forceinline__ float mu_en_of(const MCArgs &a, float e) {
    return table_lerp(a.mu_en, a.n_mu, (__logf(fmaxf(e, 1e-6f)) - a.mu_lo) * a.mu_inv_step);
}

// One photon packet carrying energy e (absorbed with mu_en at energy e_mu) and weight w
// from position p (voxel units) along the unit direction d: exact voxel traversal,
// kerma deposits; adds to the deposited and escaped totals.
__device__ void photon_packet(const MCArgs &a, const float *__restrict__ dens, double *__restrict__ edep, float px,
                              float py, float pz, float dx, float dy, float dz, float e, float e_mu, float w,
                              double &dep_tot, double &esc_tot) {
    const float mu_mm = mu_en_of(a, e_mu) * 0.1f;       // 1/mm per g/cc
    const float voxel_mm = 2.0f * a.half_voxel_mm;
    int ix = (int)floorf(px), iy = (int)floorf(py), iz = (int)floorf(pz);
    const int sx = dx > 0.0f ? 1 : -1, sy = dy > 0.0f ? 1 : -1, sz = dz > 0.0f ? 1 : -1;
    const float big = 3.0e38f;
    const float tdx = dx != 0.0f ? fabsf(1.0f / dx) : big;
    const float tdy = dy != 0.0f ? fabsf(1.0f / dy) : big;
    const float tdz = dz != 0.0f ? fabsf(1.0f / dz) : big;
    float tmx = dx != 0.0f ? ((dx > 0.0f ? (float)(ix + 1) - px : px - (float)ix) * tdx) : big;
    float tmy = dy != 0.0f ? ((dy > 0.0f ? (float)(iy + 1) - py : py - (float)iy) * tdy) : big;
    float tmz = dz != 0.0f ? ((dz > 0.0f ? (float)(iz + 1) - pz : pz - (float)iz) * tdz) : big;
    float t = 0.0f, erem = e;
    const float e_stop = 1e-6f * e;
    const int max_steps = a.n[0] + a.n[1] + a.n[2] + 3;     // a straight ray crosses each plane once
    for (int step = 0; step < max_steps; ++step) {
        if (ix < 0 || iy < 0 || iz < 0 || ix >= a.n[0] || iy >= a.n[1] || iz >= a.n[2] || erem <= e_stop) break;
        const float tn = fminf(tmx, fminf(tmy, tmz));
        const long long vox = ((long long)ix * a.n[1] + iy) * a.n[2] + iz;
        const float rho = fmaxf(dens ? dens[vox] : 1.0f, a.rho_min_ph);
        const float dep = erem * (1.0f - __expf(-mu_mm * rho * (tn - t) * voxel_mm));
        if (dep > 0.0f) {
            atomicAdd(&edep[vox], (double)(w * dep));
            dep_tot += (double)w * dep;
            erem -= dep;
        }
        t = tn;
        if (tn == tmx) { ix += sx; tmx += tdx; }
        else if (tn == tmy) { iy += sy; tmy += tdy; }
        else { iz += sz; tmz += tdz; }
    }
    esc_tot += (double)w * erem;
}

// isotropic unit direction from two uniform draws
__device__ __forceinline__ void iso_dir(uint32_t u1, uint32_t u2, float &dx, float &dy, float &dz) {
    const float cth = 2.0f * u01(u1) - 1.0f;
    const float sth = sqrtf(fmaxf(0.0f, 1.0f - cth * cth));
    float sph, cph;
    __sincosf(6.28318530718f * u01(u2), &sph, &cph);
    dx = sth * cph; dy = sth * sph; dz = cth;
}

__device__ __forceinline__ float range_of(const MCArgs &a, float e) {
    e = fmaxf(e, 1e-4f);
    return 10.0f * a.ka * __powf(e, a.kb - a.kc * __logf(e));
}

__global__ __launch_bounds__(256) void pvd_mc_kernel(MCArgs a, const float *__restrict__ act,
                                                     const float *__restrict__ dens, double *__restrict__ edep,
                                                     double *__restrict__ totals) {
    __shared__ double red[3][256 / 64];
    const long long nvox = (long long)a.n[0] * a.n[1] * a.n[2];
    const long long tid = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    const long long src = tid / a.histories;
    long long v = -1;
    if (src < a.n_src) v = a.active ? a.active[src] : src;
    double emitted = 0.0, escaped = 0.0, deposited = 0.0;
    if (v >= 0 && v < nvox && act[v] > 0.0f) {
        const float w = act[v] / (float)a.histories;
        const Philox rng{a.seed_lo, a.seed_hi};
        const uint32_t hid_lo = (uint32_t)tid, hid_hi = (uint32_t)(tid >> 32);
        uint32_t draw = 0;
        uint4 r4 = rng(hid_lo, hid_hi, draw++, 0x5eedu);
        const int vz = (int)(v % a.n[2]);
        const int vy = (int)((v / a.n[2]) % a.n[1]);
        const int vx = (int)(v / ((long long)a.n[2] * a.n[1]));
        // position in voxel units, uniform inside the emitting voxel
        float px = (float)vx + u01(r4.x), py = (float)vy + u01(r4.y), pz = (float)vz + u01(r4.z);
        const float px0 = px, py0 = py, pz0 = pz;
        float e0 = table_lerp(a.e_of_u, a.n_u, u01(r4.w) * (float)(a.n_u - 1));
        r4 = rng(hid_lo, hid_hi, draw++, 0x5eedu);
        float dx, dy, dz;
        iso_dir(r4.x, r4.y, dx, dy, dz);
        float r = fminf(range_of(a, e0), a.r_max);
        const float r0 = fmaxf(r, 1e-6f);
        const float dr = a.r_max / (float)(a.n_r - 1);
        const float inv_dr = 1.0f / dr;
        float e_cur = table_lerp(a.e_of_r, a.n_r, r * inv_dr);
        const float e_start = e_cur;
        // thick-target bremsstrahlung: the electron keeps (1 - Y) of its losses
        float yb = 0.0f;
        if (a.brems_z > 0.0f) {
            const float x = 6.0e-4f * a.brems_z * e_start;
            yb = x / (1.0f + x);
        }
        const float keep = 1.0f - yb;
        if (a.electrons) emitted = (double)w * keep * e_cur;
        // photon packets of every photon_every-th history of this voxel, weight x photon_every
        const int hist = (int)(tid - src * a.histories);
        if (hist % a.photon_every == 0 && (a.n_lines > 0 || yb > 0.0f)) {
            const float wp = w * (float)a.photon_every;
            uint32_t pd = 0;
            for (int l = 0; l < a.n_lines; ++l) {
                const uint4 q = rng(hid_lo, hid_hi, pd++, 0xf07u);
                float qx, qy, qz;
                iso_dir(q.x, q.y, qx, qy, qz);
                const float wl = wp * a.line_y[l];
                emitted += (double)wl * a.line_e[l];
                photon_packet(a, dens, edep, px0, py0, pz0, qx, qy, qz, a.line_e[l], a.line_e[l], wl, deposited,
                              escaped);
            }
            if (yb > 0.0f) {
                const uint4 q = rng(hid_lo, hid_hi, pd++, 0xf07u);
                float qx, qy, qz;
                iso_dir(q.x, q.y, qx, qy, qz);
                const float k = fmaxf(e_start * (1.0f - sqrtf(u01(q.z))), 1e-3f);    // energy spectrum ~ (E0 - k)
                const float eb = yb * e_start;
                emitted += (double)wp * eb;
                // the packet carries Y E0 and is absorbed with mu_en at its sampled energy k
                photon_packet(a, dens, edep, px0, py0, pz0, qx, qy, qz, eb, k, wp, deposited, escaped);
            }
        }
        int guard = 0;
        while (a.electrons && r > 0.0f && guard++ < 100000) {
            const int ix = (int)floorf(px), iy = (int)floorf(py), iz = (int)floorf(pz);
            if (ix < 0 || iy < 0 || iz < 0 || ix >= a.n[0] || iy >= a.n[1] || iz >= a.n[2]) {
                escaped += (double)w * keep * e_cur;
                break;
            }
            const long long vox = ((long long)ix * a.n[1] + iy) * a.n[2] + iz;
            const float rho = fmaxf(dens ? dens[vox] : 1.0f, a.rho_min);
            const float s = fminf(fminf(a.step_mm, r), a.half_voxel_mm * rho);   // water-equivalent mm
            r = fmaxf(r - s, 0.0f);
            const float e_next = r > 0.0f ? table_lerp(a.e_of_r, a.n_r, r * inv_dr) : 0.0f;
            const float de = keep * (e_cur - e_next);
            atomicAdd(&edep[vox], (double)(w * de));
            deposited += (double)w * de;
            e_cur = e_next;
            const float len = (s / rho) * a.inv_voxel;   // geometric step in voxel units
            px = fmaf(dx, len, px);
            py = fmaf(dy, len, py);
            pz = fmaf(dz, len, pz);
            // Gaussian deflection about the current direction
            r4 = rng(hid_lo, hid_hi, draw++, 0x5eedu);
            const float g = sqrtf(-2.0f * __logf(u01(r4.x))) * __cosf(6.28318530718f * u01(r4.y));
            const float th = a.scatter * sqrtf(s / r0) * g;
            float st, ct, sp, cp;
            __sincosf(th, &st, &ct);
            __sincosf(6.28318530718f * u01(r4.z), &sp, &cp);
            // orthonormal frame (e1, e2) perpendicular to d
            float e1x, e1y, e1z;
            if (fabsf(dz) < 0.9f) {            // e1 = normalize(d x z)
                const float inv = rsqrtf(dx * dx + dy * dy);
                e1x = dy * inv; e1y = -dx * inv; e1z = 0.0f;
            } else {                            // e1 = normalize(d x x)
                const float inv = rsqrtf(dy * dy + dz * dz);
                e1x = 0.0f; e1y = dz * inv; e1z = -dy * inv;
            }
            const float e2x = dy * e1z - dz * e1y, e2y = dz * e1x - dx * e1z, e2z = dx * e1y - dy * e1x;
            const float nx = ct * dx + st * (cp * e1x + sp * e2x);
            const float ny = ct * dy + st * (cp * e1y + sp * e2y);
            const float nz = ct * dz + st * (cp * e1z + sp * e2z);
            const float inv = rsqrtf(nx * nx + ny * ny + nz * nz);
            dx = nx * inv; dy = ny * inv; dz = nz * inv;
        }
    }
    // block reduction of the energy bookkeeping, one FP64 atomic per block and quantity
    double vals[3] = {emitted, deposited, escaped};
#pragma unroll
    for (int q = 0; q < 3; ++q) {
        double x = vals[q];
        for (int off = 32; off > 0; off >>= 1) x += __shfl_down(x, off, 64);
        if ((threadIdx.x & 63) == 0) red[q][threadIdx.x >> 6] = x;
    }
    __syncthreads();
    if (threadIdx.x < 3) {
        double s = 0.0;
        for (int i = 0; i < (int)(blockDim.x >> 6); ++i) s += red[threadIdx.x][i];
        atomicAdd(&totals[threadIdx.x], s);
    }
}

}  // namespace

hipError_t pvd_launch_mc(const pvd_mc_params *p, const float *act, const float *dens, double *edep, double *totals,
                         hipStream_t st) {
    MCArgs a;
    for (int i = 0; i < 3; ++i) a.n[i] = p->n[i];
    a.inv_voxel = (float)(1.0 / p->voxel_mm);
    a.half_voxel_mm = (float)(0.5 * p->voxel_mm);
    a.histories = p->histories;
    a.seed_lo = (uint32_t)p->seed;
    a.seed_hi = (uint32_t)(p->seed >> 32);
    a.e_of_u = p->e_of_u;
    a.n_u = p->n_u;
    a.e_of_r = p->e_of_r;
    a.n_r = p->n_r;
    a.r_max = (float)p->r_max_mm;
    a.ka = (float)p->katz_a;
    a.kb = (float)p->katz_b;
    a.kc = (float)p->katz_c;
    a.step_mm = (float)p->step_mm;
    a.scatter = (float)p->scatter;
    a.rho_min = (float)p->rho_min;
    a.active = p->active;
    a.n_src = p->active ? p->n_active : (long long)p->n[0] * p->n[1] * p->n[2];
    a.n_lines = p->n_lines;
    for (int l = 0; l < PVD_MC_MAX_LINES; ++l) {
        a.line_e[l] = l < p->n_lines ? (float)p->line_energy[l] : 0.0f;
        a.line_y[l] = l < p->n_lines ? (float)p->line_yield[l] : 0.0f;
    }
    a.brems_z = (float)p->brems_z;
    a.photon_every = p->photon_every;
    a.electrons = p->electrons;
    a.mu_en = p->mu_en;
    a.n_mu = p->n_mu;
    a.mu_lo = (float)p->mu_log_e_lo;
    a.mu_inv_step = p->n_mu > 1 ? (float)((p->n_mu - 1) / (p->mu_log_e_hi - p->mu_log_e_lo)) : 0.0f;
    a.rho_min_ph = (float)p->rho_min;
    const long long threads = a.n_src * p->histories;
    if (threads == 0) return hipSuccess;
    const long long blocks = (threads + 255) / 256;
    if (blocks > 0x7fffffffLL) return hipErrorInvalidValue;
    hipLaunchKernelGGL(pvd_mc_kernel, dim3((unsigned)blocks), dim3(256), 0, st, a, act, dens, edep, totals);
    return hipGetLastError();
}

extern "C" int pvd_mc_dose(const pvd_mc_params *p, const float *activity, const float *density, double *edep,
                           double *totals, void *stream) {
    if (!p || !activity || !edep || !totals) return pvd_set_error(PVD_EINVAL, "null argument");
    for (int i = 0; i < 3; ++i)
        if (p->n[i] < 1) return pvd_set_error(PVD_EINVAL, "grid extent %d on axis %d", p->n[i], i);
    if (p->histories < 1) return pvd_set_error(PVD_EINVAL, "histories must be >= 1");
    if (!p->e_of_u || p->n_u < 2 || !p->e_of_r || p->n_r < 2 || !(p->r_max_mm > 0.0))
        return pvd_set_error(PVD_EINVAL, "bad spectrum or range table");
    if (!(p->voxel_mm > 0.0) || !(p->step_mm > 0.0) || !(p->rho_min > 0.0))
        return pvd_set_error(PVD_EINVAL, "voxel_mm, step_mm and rho_min must be positive");
    if (p->active && p->n_active < 0) return pvd_set_error(PVD_EINVAL, "negative n_active");
    if (p->n_lines < 0 || p->n_lines > PVD_MC_MAX_LINES)
        return pvd_set_error(PVD_EINVAL, "n_lines must be in [0, %d]", PVD_MC_MAX_LINES);
    for (int l = 0; l < p->n_lines; ++l)
        if (!(p->line_energy[l] > 0.0) || !(p->line_yield[l] >= 0.0))
            return pvd_set_error(PVD_EINVAL, "photon line %d needs energy > 0 and yield >= 0", l);
    if (p->photon_every < 1) return pvd_set_error(PVD_EINVAL, "photon_every must be >= 1");
    if (!(p->brems_z >= 0.0)) return pvd_set_error(PVD_EINVAL, "brems_z must be >= 0");
    if ((p->n_lines > 0 || p->brems_z > 0.0) &&
        (!p->mu_en || p->n_mu < 2 || !(p->mu_log_e_hi > p->mu_log_e_lo)))
        return pvd_set_error(PVD_EINVAL, "photons need the mu_en table");
    const hipError_t e = pvd_launch_mc(p, activity, density, edep, totals, (hipStream_t)stream);
    if (e != hipSuccess) return pvd_hip_error(e, "launch Monte Carlo");
    return PVD_OK;
}
