// The misfit after the path sum: the R-factor of chi per segment against data, and its gradient in chi, in one
// launch (src/feffgpu/exafs/transform.py:104-145, src/feffgpu/compare.py:17-27). xraylarch's conventions: the
// Hanning k window (sills centered on kmin and kmax, dk wide), the transform of k^w chi(k) window(k) zero-padded to
// nfft points with the factor kstep / sqrt(pi), R_m = m pi / (nfft kstep), and
// R = sum_m |M_m - D_m|^2 / sum_m |D_m|^2 over the caller's range of m. Everything in FP64.
//
// One workgroup per segment. The transform's twiddles e^{-2 pi i t / nfft} are tabulated in LDS; the sums run over
// the k points in order (one R point per thread), and the gradient over the R points in order (one k point per
// thread), so the results do not depend on the launch.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdio.h>

#include "fm_device.h"

int fm_set_err(int code, const char *msg);

namespace {

constexpr int RF_THREADS = 256;

__device__ __forceinline__ double hann(double k, const fm_window &w) {
    if (w.dk <= 0.0) return (k >= w.kmin && k <= w.kmax) ? 1.0 : 0.0;
    const double a = fmin(fmax((k - (w.kmin - 0.5 * w.dk)) / w.dk, 0.0), 1.0);
    const double b = fmin(fmax((k - (w.kmax - 0.5 * w.dk)) / w.dk, 0.0), 1.0);
    return 0.5 * (1.0 - cospi(a)) * 0.5 * (1.0 + cospi(b));
}

// ordered tree over the workgroup (RF_THREADS values in red[])
__device__ __forceinline__ double block_sum(double v, double *red) {
    const int t = threadIdx.x;
    red[t] = v;
    __syncthreads();
    for (int h = RF_THREADS / 2; h > 0; h >>= 1) {
        if (t < h) red[t] += red[t + h];
        __syncthreads();
    }
    const double s = red[0];
    __syncthreads();
    return s;
}

__global__ __launch_bounds__(RF_THREADS) void k_r_factor(int nk, double kstep, int nfft, fm_window w, const double *chi,
                                                         double scale, const double *data, int dstride, int m0, int nm,
                                                         double *rfac, double *gchi) {
    extern __shared__ __align__(16) double sm[];
    double *twc = sm, *tws = twc + nfft;            // cos, sin of 2 pi t / nfft
    double *y = tws + nfft, *yd = y + nk, *f = yd + nk;
    double *Mr = f + nk, *Mi = Mr + nm, *red = Mi + nm;
    const int s = blockIdx.x, t = threadIdx.x;
    for (int i = t; i < nfft; i += RF_THREADS) {
        double sn, cs;
        sincospi(2.0 * (double)i / (double)nfft, &sn, &cs);
        twc[i] = cs;
        tws[i] = sn;
    }
    const double *cs = chi + (size_t)s * nk;
    const double *ds = data + (size_t)s * dstride;
    for (int j = t; j < nk; j += RF_THREADS) {
        const double kj = j * kstep;
        double p = 1.0;
        for (int e = 0; e < w.kweight; ++e) p *= kj;
        const double fj = p * hann(kj, w);
        f[j] = fj;
        y[j] = scale * cs[j] * fj - ds[j] * fj;   // the transform is linear: FT(model) - FT(data) = FT(difference)
        yd[j] = ds[j] * fj;
    }
    __syncthreads();
    const double c = kstep / sqrt(M_PI);
    const int mask = nfft - 1;
    double num = 0.0, den = 0.0;
    for (int q = t; q < nm; q += RF_THREADS) {
        const int m = m0 + q;
        double ar = 0.0, ai = 0.0, br = 0.0, bi = 0.0;
        int idx = 0;                                  // j m mod nfft
        for (int j = 0; j < nk; ++j) {
            const double co = twc[idx], si = tws[idx];
            ar += y[j] * co;
            ai -= y[j] * si;
            br += yd[j] * co;
            bi -= yd[j] * si;
            idx = (idx + m) & mask;
        }
        ar *= c; ai *= c; br *= c; bi *= c;
        Mr[q] = ar;
        Mi[q] = ai;
        num += ar * ar + ai * ai;
        den += br * br + bi * bi;
    }
    num = block_sum(num, red);
    den = block_sum(den, red);
    if (t == 0) rfac[s] = num / den;
    if (!gchi) return;
    // dR/dchi_j = 2 c scale f_j / den sum_m Re(conj(M_m) e^{-2 pi i j m / nfft})
    const double g0 = 2.0 * c * scale / den;
    for (int j = t; j < nk; j += RF_THREADS) {
        double g = 0.0;
        int idx = (int)(((long long)j * m0) & mask);
        for (int q = 0; q < nm; ++q) {
            g += Mr[q] * twc[idx] - Mi[q] * tws[idx];
            idx = (idx + j) & mask;
        }
        gchi[(size_t)s * nk + j] = g0 * f[j] * g;
    }
}

}  // namespace

extern "C" {

int fm_r_factor(int32_t nseg, int32_t nk, double kstep, int32_t nfft, const fm_window *w, int32_t m0, int32_t nm,
                const double *chi, double scale, const double *data, int32_t data_stride, double *rfac, double *gchi,
                void *stream) {
    if (nseg < 0 || nk < 2) return fm_set_err(FM_EINVAL, "fm_r_factor: nseg < 0 or nk < 2");
    if (nseg == 0) return FM_OK;
    if (!w || !chi || !data || !rfac) return fm_set_err(FM_EINVAL, "fm_r_factor: NULL buffer");
    if (nfft < 2 || nfft > FM_RF_NFFT_MAX || (nfft & (nfft - 1))) return fm_set_err(FM_EINVAL, "fm_r_factor: nfft must be a power of two in 2..4096");
    if (nk > nfft) return fm_set_err(FM_EINVAL, "fm_r_factor: more k points than nfft");
    if (!(kstep > 0.0)) return fm_set_err(FM_EINVAL, "fm_r_factor: kstep must be positive");
    if (w->kweight < 0 || w->kweight > 16) return fm_set_err(FM_EINVAL, "fm_r_factor: kweight out of range 0..16");
    if (m0 < 0 || nm < 1 || m0 + nm > nfft / 2) return fm_set_err(FM_EINVAL, "fm_r_factor: R range outside 0..nfft/2 or empty");
    if (data_stride != 0 && data_stride != nk) return fm_set_err(FM_EINVAL, "fm_r_factor: data_stride must be 0 or nk");
    const size_t shm = sizeof(double) * (2 * (size_t)nfft + 3 * (size_t)nk + 2 * (size_t)nm + RF_THREADS);
    if (shm > 160 * 1024) return fm_set_err(FM_EINVAL, "fm_r_factor: nfft, nk and the R range exceed the LDS");
    hipStream_t st = (hipStream_t)stream;
    if (shm > 64 * 1024) {
        hipError_t e = hipFuncSetAttribute((const void *)k_r_factor, hipFuncAttributeMaxDynamicSharedMemorySize, (int)shm);
        if (e != hipSuccess) return fm_set_err(FM_EHIP, hipGetErrorString(e));
    }
    hipLaunchKernelGGL(k_r_factor, dim3((unsigned)nseg), dim3(RF_THREADS), shm, st, nk, kstep, nfft, *w, chi, scale,
                       data, data_stride, m0, nm, rfac, gchi);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
        char buf[256];
        snprintf(buf, sizeof buf, "fm_r_factor launch: %s", hipGetErrorString(e));
        return fm_set_err(FM_EHIP, buf);
    }
    return FM_OK;
}

}  // extern "C"
