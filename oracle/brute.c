/* Independent plain-C direct 3-D convolution in double precision.
 *
 * Test infrastructure only: it pins the NumPy FFT restatement in
 * oracle/conv.py (ref_conv, linear_same) by brute force, with no FFT and no
 * shared code.
 *
 * mode 0 ("reference", pyvoxeldosimetry/core/kernel_convolution.py:71-74):
 *   D[i] = sum_{m in [0,K')} K[m] * A[(i - m) mod N],  K' = min(K, N) per axis
 *   (np.fft.fftn(kernel, s=N) crops the kernel to [0:N] or zero-pads it).
 * mode 1 ("linear", centred 'same'):
 *   D[i] = sum_{m in [0,K)} K[m] * A[i - m + K/2],  A = 0 outside [0, N).
 */
#include <stddef.h>

static inline long wrap(long i, long n) { i %= n; return i < 0 ? i + n : i; }

void pvd_oracle_brute_conv3d(const double *a, int nx, int ny, int nz,
                             const double *k, int kx, int ky, int kz,
                             int mode, double *out)
{
    const int cx = kx / 2, cy = ky / 2, cz = kz / 2;
    const int ex = (mode == 0 && kx > nx) ? nx : kx;
    const int ey = (mode == 0 && ky > ny) ? ny : ky;
    const int ez = (mode == 0 && kz > nz) ? nz : kz;
    for (long i = 0; i < nx; ++i)
    for (long j = 0; j < ny; ++j)
    for (long l = 0; l < nz; ++l) {
        double acc = 0.0;
        for (long p = 0; p < ex; ++p) {
            long si = (mode == 0) ? wrap(i - p, nx) : i - p + cx;
            if (si < 0 || si >= nx) continue;
            for (long q = 0; q < ey; ++q) {
                long sj = (mode == 0) ? wrap(j - q, ny) : j - q + cy;
                if (sj < 0 || sj >= ny) continue;
                const double *krow = k + ((size_t)p * ky + q) * kz;
                const double *arow = a + ((size_t)si * ny + sj) * nz;
                for (long r = 0; r < ez; ++r) {
                    long sl = (mode == 0) ? wrap(l - r, nz) : l - r + cz;
                    if (sl < 0 || sl >= nz) continue;
                    acc += krow[r] * arow[sl];
                }
            }
        }
        out[((size_t)i * ny + j) * nz + l] = acc;
    }
}
