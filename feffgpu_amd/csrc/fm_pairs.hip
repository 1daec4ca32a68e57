// Single scattering from tabulated two-leg signals (PairTables, src/feffgpu/exafs/forward.py:178-320): the histogram
// of absorber-neighbor distances on the tables' grid r0 + i dr (linear-interpolation weights, minimum image of the
// cell, pairs within rmax, self excluded) and its adjoint in the positions.
//
// The histogram is accumulated in 64-bit fixed point (units of 2^-32 of a pair): integer atomics are associative, so
// the result does not depend on the order the pairs arrive in, and each pair adds exactly 2^32 over its two bins. A
// bin holds up to 2^31 pairs before the signed 64-bit value overflows (the host splits larger calls). The
// adjoint needs no atomics: one thread per atom sums its pairs, as absorber and as neighbor, in index order.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdio.h>

#include "fm_device.h"

int fm_set_err(int code, const char *msg);

namespace {

constexpr int PH_THREADS = 256;
constexpr double PH_ONE = 4294967296.0;   // 2^32

// minimum image of v (src/feffgpu/cell.py:26-35): cubic v - L rint(v / L), general cell by fractional rounding
__device__ __forceinline__ void min_image(double (&v)[3], const fm_pair_grid &g) {
    if (g.cell_kind == 1) {
        const double L = g.cell[0];
#pragma unroll
        for (int c = 0; c < 3; ++c) v[c] -= L * rint(v[c] / L);
    } else if (g.cell_kind == 2) {
        double f[3];
#pragma unroll
        for (int c = 0; c < 3; ++c) f[c] = v[0] * g.cellinv[0 * 3 + c] + v[1] * g.cellinv[1 * 3 + c] + v[2] * g.cellinv[2 * 3 + c];
#pragma unroll
        for (int c = 0; c < 3; ++c) f[c] -= rint(f[c]);
#pragma unroll
        for (int c = 0; c < 3; ++c) v[c] = f[0] * g.cell[0 * 3 + c] + f[1] * g.cell[1 * 3 + c] + f[2] * g.cell[2 * 3 + c];
    }
}

// Pair (a, j): displacement j - a (minimum image) and distance; false outside rmax or for coincident atoms
// (r^2 <= 1e-6, as forward.py:259, 292 exclude the atom itself).
__device__ __forceinline__ bool pair(const double *pa, const double *pj, const fm_pair_grid &g, double (&v)[3], double &r) {
#pragma unroll
    for (int c = 0; c < 3; ++c) v[c] = pj[c] - pa[c];
    min_image(v, g);
    const double r2 = v[0] * v[0] + v[1] * v[1] + v[2] * v[2];
    if (!(r2 <= g.rmax * g.rmax) || !(r2 > 1e-6)) return false;
    r = sqrt(r2);
    return true;
}

// lower bin and upper weight of distance r (forward.py:264-266, 296-298): u = (r - r0) / dr, i0 = floor(u) clamped
// to 0..nr-2, f = u - i0 clamped to [0, 1]; dfdr: d f / d r (1 / dr inside the clamp, else 0)
__device__ __forceinline__ void bin_of(double r, const fm_pair_grid &g, int &i0, double &f, double &dfdr) {
    const double u = (r - g.r0) / g.dr;
    double fl = floor(u);
    fl = fl < 0.0 ? 0.0 : (fl > g.nr - 2 ? (double)(g.nr - 2) : fl);
    i0 = (int)fl;
    f = u - fl;
    dfdr = (f >= 0.0 && f <= 1.0) ? 1.0 / g.dr : 0.0;
    f = f < 0.0 ? 0.0 : (f > 1.0 ? 1.0 : f);
}

__global__ __launch_bounds__(PH_THREADS) void k_pair_hist(int natom, const double *pos, int nabs, const int32_t *absorbers,
                                                           const int32_t *ipot, int npot, fm_pair_grid g,
                                                           unsigned long long *hist, int32_t *nshort) {
    __shared__ int cnt[PH_THREADS / FM_WAVE];
    const int c = blockIdx.y, ai = blockIdx.x;
    const int a = absorbers ? absorbers[ai] : ai;
    if (a < 0 || a >= natom) return;           // an absorber index outside the configuration adds nothing
    const double *P = pos + (size_t)c * natom * 3;
    const double pa[3] = {P[3 * a], P[3 * a + 1], P[3 * a + 2]};
    int shorts = 0;
    for (int j = threadIdx.x; j < natom; j += PH_THREADS) {
        double v[3], r;
        if (!pair(pa, P + 3 * (size_t)j, g, v, r)) continue;
        if (r < g.r0) ++shorts;
        const int p = ipot[j];
        if (p < 0 || p >= npot) continue;
        int i0;
        double f, dfdr;
        bin_of(r, g, i0, f, dfdr);
        const unsigned long long up = (unsigned long long)llrint(f * PH_ONE);
        unsigned long long *h = hist + (size_t)p * g.nr + i0;
        atomicAdd(h, (unsigned long long)PH_ONE - up);
        atomicAdd(h + 1, up);
    }
    if (nshort) {
        for (int off = FM_WAVE / 2; off > 0; off >>= 1) shorts += __shfl_xor(shorts, off, FM_WAVE);
        if ((threadIdx.x & (FM_WAVE - 1)) == 0) cnt[threadIdx.x / FM_WAVE] = shorts;
        __syncthreads();
        if (threadIdx.x == 0) {
            int s = 0;
            for (int w = 0; w < PH_THREADS / FM_WAVE; ++w) s += cnt[w];
            if (s) atomicAdd(nshort, s);
        }
    }
}

// The same histogram with the bins of PH_ABS absorbers first summed in LDS (64-bit fixed point, LDS atomics) and only
// the non-zero bins then added to hist: the pairs of a frame crowd onto the few bins of its coordination shells, so
// global atomics on them would contend; integer sums stay order-independent.
constexpr int PH_ABS = 64;

__global__ __launch_bounds__(PH_THREADS) void k_pair_hist_lds(int natom, const double *pos, int nabs, const int32_t *absorbers,
                                                               const int32_t *ipot, int npot, fm_pair_grid g,
                                                               unsigned long long *hist, int32_t *nshort) {
    extern __shared__ __align__(16) unsigned long long lh[];
    __shared__ int cnt[PH_THREADS / FM_WAVE];
    const int c = blockIdx.y, t = threadIdx.x;
    const int nb = npot * g.nr;
    for (int b = t; b < nb; b += PH_THREADS) lh[b] = 0ull;
    __syncthreads();
    const double *P = pos + (size_t)c * natom * 3;
    int shorts = 0;
    const int a_end = min(nabs, (int)(blockIdx.x + 1) * PH_ABS);
    for (int ai = blockIdx.x * PH_ABS; ai < a_end; ++ai) {
        const int a = absorbers ? absorbers[ai] : ai;
        if (a < 0 || a >= natom) continue;          // uniform over the workgroup
        const double pa[3] = {P[3 * a], P[3 * a + 1], P[3 * a + 2]};
        for (int j = t; j < natom; j += PH_THREADS) {
            double v[3], r;
            if (!pair(pa, P + 3 * (size_t)j, g, v, r)) continue;
            if (r < g.r0) ++shorts;
            const int p = ipot[j];
            if (p < 0 || p >= npot) continue;
            int i0;
            double f, dfdr;
            bin_of(r, g, i0, f, dfdr);
            const unsigned long long up = (unsigned long long)llrint(f * PH_ONE);
            atomicAdd(&lh[p * g.nr + i0], (unsigned long long)PH_ONE - up);
            atomicAdd(&lh[p * g.nr + i0 + 1], up);
        }
    }
    __syncthreads();
    for (int b = t; b < nb; b += PH_THREADS) {
        const unsigned long long v = lh[b];
        if (v) atomicAdd(hist + b, v);
    }
    if (nshort) {
        for (int off = FM_WAVE / 2; off > 0; off >>= 1) shorts += __shfl_xor(shorts, off, FM_WAVE);
        if ((t & (FM_WAVE - 1)) == 0) cnt[t / FM_WAVE] = shorts;
        __syncthreads();
        if (t == 0) {
            int sum = 0;
            for (int w = 0; w < PH_THREADS / FM_WAVE; ++w) sum += cnt[w];
            if (sum) atomicAdd(nshort, sum);
        }
    }
}

// d L / d pos for L = sum_p,i gbin[p][i] hist[p][i], every atom an absorber: pair (a, j) puts (1 - f, f) on the
// bins (i0, i0 + 1) of potential ipot[j]; its distance moves with pos_j - pos_a.
__global__ __launch_bounds__(PH_THREADS) void k_pair_grad(int natom, const double *pos, const int32_t *ipot, int npot,
                                                           fm_pair_grid g, const double *gbin, double *grad) {
    const int c = blockIdx.y, a = blockIdx.x * PH_THREADS + threadIdx.x;
    if (a >= natom) return;
    const double *P = pos + (size_t)c * natom * 3;
    const double pa[3] = {P[3 * a], P[3 * a + 1], P[3 * a + 2]};
    const int pa_pot = ipot[a];
    double ga[3] = {0.0, 0.0, 0.0};
    for (int j = 0; j < natom; ++j) {
        double v[3], r;
        if (!pair(pa, P + 3 * (size_t)j, g, v, r)) continue;
        int i0;
        double f, dfdr;
        bin_of(r, g, i0, f, dfdr);
        double dLdr = 0.0;
        const int pj = ipot[j];
        if (pj >= 0 && pj < npot) {       // a the absorber, j the neighbor
            const double *gb = gbin + (size_t)pj * g.nr + i0;
            dLdr += (gb[1] - gb[0]) * dfdr;
        }
        if (pa_pot >= 0 && pa_pot < npot) {   // j the absorber, a the neighbor (same distance)
            const double *gb = gbin + (size_t)pa_pot * g.nr + i0;
            dLdr += (gb[1] - gb[0]) * dfdr;
        }
        const double s = dLdr / r;          // d r / d pos_a = -(pos_j - pos_a) / r
#pragma unroll
        for (int k = 0; k < 3; ++k) ga[k] -= s * v[k];
    }
    double *o = grad + ((size_t)c * natom + a) * 3;
#pragma unroll
    for (int k = 0; k < 3; ++k) o[k] = ga[k];
}

const char *grid_problem(const fm_pair_grid *g) {
    if (!g) return "NULL grid";
    if (g->nr < 2) return "the distance grid needs nr >= 2";
    if (!(g->dr > 0.0)) return "dr must be positive";
    if (!(g->rmax > 0.0)) return "rmax must be positive";
    if (g->cell_kind < 0 || g->cell_kind > 2) return "cell_kind must be 0, 1 or 2";
    if (g->cell_kind == 1 && !(g->cell[0] > 0.0)) return "the cubic cell needs a positive length";
    return nullptr;
}

int launch_err(const char *what) {
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) return FM_OK;
    char buf[256];
    snprintf(buf, sizeof buf, "%s launch: %s", what, hipGetErrorString(e));
    return fm_set_err(FM_EHIP, buf);
}

}  // namespace

extern "C" {

int fm_pair_histogram(int32_t nconf, int32_t natom, const double *pos, int32_t nabs, const int32_t *absorbers,
                      const int32_t *ipot, int32_t npot, const fm_pair_grid *grid, int64_t *hist, int32_t *nshort,
                      void *stream) {
    const char *m = grid_problem(grid);
    if (m) return fm_set_err(FM_EINVAL, m);
    if (nconf < 0 || natom < 0 || npot < 1) return fm_set_err(FM_EINVAL, "fm_pair_histogram: bad sizes");
    const int na = absorbers ? nabs : natom;
    if (na < 0 || nconf > 65535) return fm_set_err(FM_EINVAL, "fm_pair_histogram: bad absorber count or more than 65535 configurations");
    if (nconf == 0 || natom == 0 || na == 0) return FM_OK;
    if (!pos || !ipot || !hist) return fm_set_err(FM_EINVAL, "fm_pair_histogram: NULL buffer");
    // every ordered pair adds one unit over two bins: nconf * na * (natom - 1) < 2^31 keeps every bin below overflow
    if ((double)nconf * na * (natom - 1) >= 2147483648.0)
        return fm_set_err(FM_EINVAL, "fm_pair_histogram: nconf * absorbers * (natom - 1) >= 2^31 pairs could overflow a "
                                     "bin; split the configurations over several calls");
    const size_t lds = sizeof(unsigned long long) * (size_t)npot * grid->nr;
    if (lds <= 144 * 1024) {   // the bins fit a workgroup's LDS: sum PH_ABS absorbers there first
        if (lds > 64 * 1024) {
            hipError_t e = hipFuncSetAttribute((const void *)k_pair_hist_lds, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
            if (e != hipSuccess) return fm_set_err(FM_EHIP, hipGetErrorString(e));
        }
        hipLaunchKernelGGL(k_pair_hist_lds, dim3((unsigned)((na + PH_ABS - 1) / PH_ABS), (unsigned)nconf), dim3(PH_THREADS),
                           lds, (hipStream_t)stream, natom, pos, na, absorbers, ipot, npot, *grid,
                           reinterpret_cast<unsigned long long *>(hist), nshort);
    } else {
        hipLaunchKernelGGL(k_pair_hist, dim3((unsigned)na, (unsigned)nconf), dim3(PH_THREADS), 0, (hipStream_t)stream, natom,
                           pos, na, absorbers, ipot, npot, *grid, reinterpret_cast<unsigned long long *>(hist), nshort);
    }
    return launch_err("fm_pair_histogram");
}

int fm_pair_histogram_grad(int32_t nconf, int32_t natom, const double *pos, const int32_t *ipot, int32_t npot,
                           const fm_pair_grid *grid, const double *gbin, double *grad_pos, void *stream) {
    const char *m = grid_problem(grid);
    if (m) return fm_set_err(FM_EINVAL, m);
    if (nconf < 0 || natom < 0 || npot < 1 || nconf > 65535) return fm_set_err(FM_EINVAL, "fm_pair_histogram_grad: bad sizes");
    if (nconf == 0 || natom == 0) return FM_OK;
    if (!pos || !ipot || !gbin || !grad_pos) return fm_set_err(FM_EINVAL, "fm_pair_histogram_grad: NULL buffer");
    hipLaunchKernelGGL(k_pair_grad, dim3((unsigned)((natom + PH_THREADS - 1) / PH_THREADS), (unsigned)nconf),
                       dim3(PH_THREADS), 0, (hipStream_t)stream, natom, pos, ipot, npot, *grid, gbin, grad_pos);
    return launch_err("fm_pair_histogram_grad");
}

}  // extern "C"
