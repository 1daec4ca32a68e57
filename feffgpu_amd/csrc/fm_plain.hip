// The plain FP64 path stage: the op sequence of feffgpu's GPUPathStage.cchi / chi_fine
// (src/feffgpu/stage.py:147-388) as one kernel per stage, with every intermediate in HBM, in FP64:
//   geometry (r, beta, eta per leg; stage.py:147-206)                       k_pl_geo     [C][n]
//   rotation matrices d^l_{m1 m2}(beta), |m| <= 2 (stage.py:41-93)          k_pl_wigner  [C][n][L1][5][5]
//   spherical-wave factors c[il][im] per leg and energy (_clm, :209-232)    k_pl_clm     [C][n][L1+2][mnp1+1][E]
//   gam and gamtl per leg (:267-285)                                        k_pl_gam     [C][n][L1+2][lambda][E] x 2
//   scattering matrices F_s (fmat, :287-305)                                k_pl_fmat    [C][n-1][lambda][lambda][E]
//   termination T (:318-334)                                                k_pl_term    [C][lambda0][lambda0][E]
//   matrix chain, trace and cfac (:309-317, 335-338)                        k_pl_chain   [C][E]
//   amplitude, unwrapped phase, interpolation to q (chi_fine, :365-388)     k_pl_fine    [P][NK]
//   degeneracy-weighted sum per segment (src/feffgpu/exafs/forward.py:162) k_segment_sum
// The energy index is the fastest of every intermediate (coalesced). A bucket's paths are processed in chunks of
// as many paths as the workspace holds. This is the comparator the fused kernels are measured against and a second,
// independent evaluation of the path stage for the tests: it shares only the leg geometry, the Wigner seeds and the
// fine-grid epilogue with them.
#include <hip/hip_runtime.h>
#include <stdio.h>
#include <string.h>

#include "fm_device.h"
#include "fm_path.h"

using namespace fm;

int fm_set_err(int code, const char *msg);

namespace {

using Cd = cx<double>;

struct LamRt {
    int L, L0, mnp1;
    int M[6], N[6];
};

struct PDims {
    int C;               // paths of this chunk
    int n, E, L1, LC, MU;  // legs, energies, L1, rows of c (L1 + 2), columns of c (mnp1 + 1)
};

// The chunk's arrays, carved from the workspace.
struct PBuf {
    int32_t *pidx, *ok, *ipa;     // [C], [C], [C][n]: ipot of path atom j+1 (atom n = the absorber)
    double *r, *beta, *eta, *reff;
    double *dl;
    Cd *clm, *gam, *gamtl, *F, *T, *cchi;
};

__host__ __device__ inline size_t al256(size_t x) { return (x + 255) & ~size_t(255); }

// Carves the arrays of a chunk of C paths from `base` (base == nullptr: only the size). Returns the bytes used.
__host__ size_t carve(const PDims &d, unsigned char *base, PBuf *b) {
    size_t o = 0;
    auto take = [&](size_t bytes) {
        unsigned char *p = base ? base + o : nullptr;
        o += al256(bytes);
        return p;
    };
    const size_t C = d.C, n = d.n, E = d.E;
    unsigned char *p;
    p = take(4 * C); if (b) b->pidx = (int32_t *)p;
    p = take(4 * C); if (b) b->ok = (int32_t *)p;
    p = take(4 * C * n); if (b) b->ipa = (int32_t *)p;
    p = take(8 * C * n); if (b) b->r = (double *)p;
    p = take(8 * C * n); if (b) b->beta = (double *)p;
    p = take(8 * C * n); if (b) b->eta = (double *)p;
    p = take(8 * C); if (b) b->reff = (double *)p;
    p = take(8 * C * n * d.L1 * 25); if (b) b->dl = (double *)p;
    p = take(16 * C * n * d.LC * d.MU * E); if (b) b->clm = (Cd *)p;
    p = take(16 * C * n * d.LC * 6 * E); if (b) b->gam = (Cd *)p;
    p = take(16 * C * n * d.LC * 6 * E); if (b) b->gamtl = (Cd *)p;
    p = take(16 * C * (n - 1) * 36 * E); if (b) b->F = (Cd *)p;
    p = take(16 * C * 36 * E); if (b) b->T = (Cd *)p;
    p = take(16 * C * E); if (b) b->cchi = (Cd *)p;
    return o;
}

__device__ __forceinline__ Cd cexp_i(double re_arg, double im_arg) {   // exp(i (re + i im))
    const double a = exp(-im_arg);
    double s, c;
    sincos(re_arg, &s, &c);
    return mk<double>(a * c, a * s);
}

// sqrt((2l+1)(l-m)!/(l+m)!), l = il - 1; 0 for m > l (snlm, src/feffgpu/genfmt.py:41-53, without FEFF's common
// scale, which cancels between gam and gamtl)
__device__ __forceinline__ double xnlm_of(int il, int m) {
    const int l = il - 1;
    if (m < 0 || m > l) return 0.0;
    return sqrt((2.0 * l + 1.0) * exp(lgamma((double)(l - m + 1)) - lgamma((double)(l + m + 1))));
}

// ---- geometry: one thread per path
__global__ __launch_bounds__(256) void k_pl_geo(PDims d, fm_paths P, const int32_t *order, int first, PBuf b) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= d.C) return;
    const int n = d.n;
    const int p = order ? order[first + c] : first + c;
    const bool ok = P.nleg[p] == n;
    double rat[FM_MAX_LEGS + 1][3];
    for (int j = 0; j < n; ++j) {
        double v[3];
        if (ok) {
            path_point(P, p, j, v);
        } else {   // placeholder geometry (NaN output)
            v[0] = 4.0 * j; v[1] = 1.5 * (j & 1); v[2] = 0.5 * j;
        }
        for (int k = 0; k < 3; ++k) rat[j][k] = v[k];
    }
    for (int k = 0; k < 3; ++k) rat[n][k] = rat[0][k];
    double alpha[FM_MAX_LEGS + 1], gam[FM_MAX_LEGS + 1];
    double R = 0.0;
    for (int J = 1; J <= n; ++J) {
        const LegGeo lg = leg_geometry(rat, J, n);
        b.r[(size_t)c * n + J - 1] = lg.r;
        b.beta[(size_t)c * n + J - 1] = lg.beta;
        alpha[J] = lg.alpha;
        gam[J] = lg.gamma;
        R += lg.r;
    }
    for (int J = 1; J <= n; ++J) b.eta[(size_t)c * n + J - 1] = (J == 1 ? alpha[n] : alpha[J - 1]) + gam[J];
    b.reff[c] = 0.5 * R;
    b.pidx[c] = p;
    b.ok[c] = ok ? 1 : 0;
    for (int j = 1; j <= n; ++j) b.ipa[(size_t)c * n + j - 1] = ok ? P.ipots[(size_t)p * P.nmax + (j == n ? 0 : j)] : 0;
}

// ---- d^l_{m1 m2}(beta) of every leg, |m| <= 2, l = 0..L1-1: one thread per (path, leg, m1, m2)
__global__ __launch_bounds__(256) void k_pl_wigner(PDims d, const double *rotc, PBuf b) {
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (size_t)d.C * d.n * 25) return;
    const int mm = (int)(t % 25);
    const size_t cl = t / 25;                 // c * n + leg
    const int m1 = mm / 5 - 2, m2 = mm % 5 - 2;
    const double be = b.beta[cl];
    const double cb = cos(be), sb = sin(be), c2 = cos(0.5 * be), s2 = sin(0.5 * be);
    const int k = (m1 + 3) * 7 + (m2 + 3);
    double dm1 = 0.0, d0 = 0.0;
    for (int l = 0; l < d.L1; ++l) {
        double v;
        if (l <= 2) {
            v = dseed(l, m1, m2, cb, sb, c2, s2);
        } else {
            const double *co = rotc + ((size_t)(l - 1) * 49 + k) * 3;
            v = (co[0] * cb - co[1]) * d0 - co[2] * dm1;
        }
        dm1 = d0;
        d0 = v;
        b.dl[(cl * d.L1 + l) * 25 + mm] = v;
    }
}

// ---- c[il][im] of leg j at energy e (FEFF's sclmz as the stage vectorises it): one thread per (path, leg, e)
__device__ __forceinline__ Cd &clm_at(const PDims &d, Cd *clm, size_t cl, int il, int im, int e) {
    return clm[((cl * d.LC + il) * d.MU + im) * d.E + e];
}
__global__ __launch_bounds__(256) void k_pl_clm(PDims d, const double *ck, const int32_t *lmax, int mnp1, PBuf b) {
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (size_t)d.C * d.n * d.E) return;
    const int e = (int)(t % d.E);
    const size_t cl = t / d.E;
    const int c = (int)(cl / d.n), j = (int)(cl % d.n);
    const int n = d.n, E = d.E, L1 = d.L1;
    const int ipj = b.ipa[(size_t)c * n + j], ipp = b.ipa[(size_t)c * n + (j + n - 1) % n];
    const int lxp1 = max(lmax[(size_t)ipj * E + e], lmax[(size_t)ipp * E + e]) + 1;
    const Cd rho = mk<double>(ck[2 * e], ck[2 * e + 1]) * b.r[cl];
    const Cd z = minus_i_over(rho);
    const Cd zero = mk<double>(0.0, 0.0), one = mk<double>(1.0, 0.0);
    auto alive = [&](int il) { return il <= lxp1 ? 1.0 : 0.0; };
    for (int il = 0; il < d.LC; ++il)
        for (int im = 0; im < d.MU; ++im) clm_at(d, b.clm, cl, il, im, e) = zero;
    const int lmaxv = L1 - 1;
    clm_at(d, b.clm, cl, 1, 1, e) = one;
    if (2 < d.LC) clm_at(d, b.clm, cl, 2, 1, e) = one - z;
    for (int il = 2; il <= lmaxv; ++il)
        clm_at(d, b.clm, cl, il + 1, 1, e) =
            (clm_at(d, b.clm, cl, il - 1, 1, e) - z * (double)(2 * il - 1) * clm_at(d, b.clm, cl, il, 1, e)) * alive(il + 1);
    Cd cmm = one;
    for (int im = 2; im <= mnp1; ++im) {
        const int m = im - 1;
        cmm = mk<double>(-cmm.r, -cmm.i) * (double)(2 * m - 1) * z;
        clm_at(d, b.clm, cl, im, im, e) = cmm * alive(im);
        if (im + 1 <= L1 + 1) clm_at(d, b.clm, cl, im + 1, im, e) = cmm * (double)(2 * m + 1) * (one - z * (double)im) * alive(im + 1);
        for (int il = im + 1; il <= lmaxv; ++il) {
            const int l = il - 1;
            clm_at(d, b.clm, cl, il + 1, im, e) =
                (clm_at(d, b.clm, cl, l, im, e) -
                 z * (double)(2 * l + 1) * (clm_at(d, b.clm, cl, il, im, e) + clm_at(d, b.clm, cl, il, m, e))) *
                alive(il + 1);
        }
    }
}

// ---- gam[il][lambda] and gamtl[il][lambda] of every leg: one thread per (path, leg, il, lambda, e). gamtl carries
// the t-matrix of the atom at the leg's end for legs 1..n-1 (the scattering matrices), 2il - 1 for leg n (the
// termination).
__global__ __launch_bounds__(256) void k_pl_gam(PDims d, LamRt lam, const double *tl64, PBuf b) {
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (size_t)d.C * d.n * d.LC * 6 * d.E) return;
    const int e = (int)(t % d.E);
    size_t r = t / d.E;
    const int lam_i = (int)(r % 6); r /= 6;
    const int il = (int)(r % d.LC);
    const size_t cl = r / d.LC;
    const int c = (int)(cl / d.n), j = (int)(cl % d.n);
    const size_t o = ((cl * d.LC + il) * 6 + lam_i) * d.E + e;
    if (lam_i >= lam.L) {
        b.gam[o] = mk<double>(0.0, 0.0);
        b.gamtl[o] = mk<double>(0.0, 0.0);
        return;
    }
    const int m = lam.M[lam_i], am = m < 0 ? -m : m, nn = lam.N[lam_i];
    const int idx = am + nn + 1;
    const double xn = xnlm_of(il, am);
    const double sg = (am & 1) ? -1.0 : 1.0;
    Cd g = mk<double>(0.0, 0.0);
    if (idx <= il && idx <= lam.mnp1) g = clm_at(d, b.clm, cl, il, idx, e) * (xn * sg);
    b.gam[o] = g;
    const Cd cn = nn + 1 < d.MU ? clm_at(d, b.clm, cl, il, nn + 1, e) * (1.0 / (xn == 0.0 ? 1.0 : xn)) : mk<double>(0.0, 0.0);
    Cd tv;
    if (j < d.n - 1) {
        const int ip = b.ipa[(size_t)c * d.n + j];
        tv = il <= d.L1 ? mk<double>(tl64[(((size_t)ip * (d.L1 + 1) + il) * d.E + e) * 2],
                                     tl64[(((size_t)ip * (d.L1 + 1) + il) * d.E + e) * 2 + 1])
                        : mk<double>(0.0, 0.0);
    } else {
        tv = mk<double>(2.0 * il - 1.0, 0.0);
    }
    b.gamtl[o] = cn * tv;
}

// ---- F_s[l1][l2] = exp(-i eta_{s+1} m1) sum_il gam_{leg s+1}[il][l1] d^{il-1}_{m1 m2}(beta_s) mask gamtl_{leg s}[il][l2]
// (fmat, src/feffgpu/stage.py:287-305): one thread per (path, s, l1, l2, e)
__global__ __launch_bounds__(256) void k_pl_fmat(PDims d, LamRt lam, const int32_t *lmax, PBuf b) {
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int ns = d.n - 1;
    if (t >= (size_t)d.C * ns * 36 * d.E) return;
    const int e = (int)(t % d.E);
    size_t r = t / d.E;
    const int l2 = (int)(r % 6); r /= 6;
    const int l1 = (int)(r % 6); r /= 6;
    const int s = (int)(r % ns) + 1;           // scatterer 1..n-1
    const int c = (int)(r / ns);
    const size_t o = ((((size_t)c * ns + s - 1) * 6 + l1) * 6 + l2) * d.E + e;
    if (l1 >= lam.L || l2 >= lam.L) {
        b.F[o] = mk<double>(0.0, 0.0);
        return;
    }
    const int m1 = lam.M[l1], m2 = lam.M[l2], n1 = lam.N[l1], n2 = lam.N[l2];
    const int am1 = m1 < 0 ? -m1 : m1, am2 = m2 < 0 ? -m2 : m2;
    const int ilmin = max(max(am1 + 1, am2 + 1), max(am1 + n1 + 1, n2 + 1));
    const int ip = b.ipa[(size_t)c * d.n + s - 1];
    const int ilx = lmax[(size_t)ip * d.E + e] + 1;
    const size_t cg = (size_t)c * d.n + s, ct = (size_t)c * d.n + s - 1;   // gam leg s+1, gamtl / d leg s
    Cd acc = mk<double>(0.0, 0.0);
    for (int il = 1; il <= d.L1; ++il) {
        if (il > ilx || il < ilmin || il <= am1 || il <= am2) continue;
        const double dv = b.dl[(ct * d.L1 + il - 1) * 25 + (m1 + 2) * 5 + (m2 + 2)];
        cfma(acc, b.gam[((cg * d.LC + il) * 6 + l1) * d.E + e] * dv, b.gamtl[((ct * d.LC + il) * 6 + l2) * d.E + e]);
    }
    const double et = b.eta[(size_t)c * d.n + s];
    b.F[o] = acc * cexp_i(-et * m1, 0.0);
}

// ---- the termination T[l1][l2] (src/feffgpu/stage.py:318-334): one thread per (path, l1, l2, e), l1, l2 < lambda0
__global__ __launch_bounds__(256) void k_pl_term(PDims d, LamRt lam, fm_edge ed, PBuf b) {
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (size_t)d.C * 36 * d.E) return;
    const int e = (int)(t % d.E);
    size_t r = t / d.E;
    const int l2 = (int)(r % 6); r /= 6;
    const int l1 = (int)(r % 6);
    const int c = (int)(r / 6);
    const size_t o = (((size_t)c * 6 + l1) * 6 + l2) * d.E + e;
    if (l1 >= lam.L0 || l2 >= lam.L0) {
        b.T[o] = mk<double>(0.0, 0.0);
        return;
    }
    const int m1 = lam.M[l1], m2 = lam.M[l2];
    const int am1 = m1 < 0 ? -m1 : m1, am2 = m2 < 0 ? -m2 : m2;
    const size_t c0 = (size_t)c * d.n, cn = (size_t)c * d.n + d.n - 1;    // gam leg 1, gamtl / d leg n
    Cd acc = mk<double>(0.0, 0.0);
    for (int q = 0; q < ed.nchan; ++q) {
        const int il = ed.chan_il[q];
        if (am1 + 1 > il || am2 + 1 > il) continue;
        const Cd w = mk<double>(ed.chw64[((size_t)q * d.E + e) * 2], ed.chw64[((size_t)q * d.E + e) * 2 + 1]);
        const double dv = b.dl[(cn * d.L1 + il - 1) * 25 + (m1 + 2) * 5 + (m2 + 2)];
        cfma(acc, w * b.gam[((c0 * d.LC + il) * 6 + l1) * d.E + e] * dv, b.gamtl[((cn * d.LC + il) * 6 + l2) * d.E + e]);
    }
    const double et = b.eta[c0];
    b.T[o] = acc * cexp_i(-et * m1, 0.0);
}

// ---- P = F_{n-1} ... F_1, ptrac = tr(T P), cchi = ptrac exp(i R (ck - xk)) / prod(ck r_j) (src/feffgpu/stage.py:
// 309-317, 335-338): one thread per (path, e)
__global__ __launch_bounds__(256) void k_pl_chain(PDims d, LamRt lam, const double *ck, const double *ckmxk, PBuf b) {
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (size_t)d.C * d.E) return;
    const int e = (int)(t % d.E);
    const int c = (int)(t / d.E);
    const int n = d.n, ns = n - 1, L = lam.L, L0 = lam.L0;
    auto Fv = [&](int s, int i, int j) { return b.F[((((size_t)c * ns + s - 1) * 6 + i) * 6 + j) * d.E + e]; };
    Cd P[6][6], Q[6][6];
    // P = F_1 (L x L0)
#pragma unroll
    for (int i = 0; i < 6; ++i)
#pragma unroll
        for (int j = 0; j < 6; ++j) P[i][j] = (i < L && j < L0) ? Fv(1, i, j) : mk<double>(0.0, 0.0);
    for (int s = 2; s <= ns; ++s) {
        const int rows = s == ns ? L0 : L;
#pragma unroll
        for (int i = 0; i < 6; ++i)
#pragma unroll
            for (int j = 0; j < 6; ++j) {
                Cd acc = mk<double>(0.0, 0.0);
                if (i < rows && j < L0)
#pragma unroll
                    for (int k = 0; k < 6; ++k)
                        if (k < L) cfma(acc, Fv(s, i, k), P[k][j]);
                Q[i][j] = acc;
            }
#pragma unroll
        for (int i = 0; i < 6; ++i)
#pragma unroll
            for (int j = 0; j < 6; ++j) P[i][j] = Q[i][j];
    }
    Cd ptrac = mk<double>(0.0, 0.0);
#pragma unroll
    for (int a = 0; a < 6; ++a)
#pragma unroll
        for (int bb = 0; bb < 6; ++bb)
            if (a < L0 && bb < L0) cfma(ptrac, b.T[(((size_t)c * 6 + a) * 6 + bb) * d.E + e], P[bb][a]);
    const Cd kk = mk<double>(ck[2 * e], ck[2 * e + 1]), dk = mk<double>(ckmxk[2 * e], ckmxk[2 * e + 1]);
    const double R = 2.0 * b.reff[c];
    Cd prho = mk<double>(1.0, 0.0);
    for (int j = 0; j < n; ++j) prho = prho * (kk * b.r[(size_t)c * n + j]);
    const Cd cf = cdiv(cexp_i(R * dk.r, R * dk.i), prho);
    b.cchi[(size_t)c * d.E + e] = ptrac * cf;
}

// ---- chi_fine: one wavefront per path (amplitude and phase per energy, unwrap, interpolation to q)
__global__ __launch_bounds__(FM_WAVE * 4) void k_pl_fine(PDims d, fm_edge ed, fm_grid G, int edge_id,
                                                         const double *sigma2, PBuf b, double *chi_path,
                                                         double *reff_out) {
    extern __shared__ __align__(16) unsigned char smem[];
    const int wid = threadIdx.x / FM_WAVE, lane = lane_id();
    const int c = blockIdx.x * 4 + wid;
    if (c >= d.C) return;
    double *amp = reinterpret_cast<double *>(smem) + (size_t)wid * 2 * FM_MAX_E;
    double *ang = amp + FM_MAX_E;
    const int p = b.pidx[c];
    const int nk = G.nk;
    double *out = chi_path + (size_t)p * nk;
    const double reff_A = b.ok[c] ? b.reff[c] * FM_BOHR : __builtin_nan("");
    if (lane == 0 && reff_out) reff_out[p] = reff_A;
    if (!b.ok[c]) {
        for (int i = lane; i < nk; i += FM_WAVE) out[i] = __builtin_nan("");
        return;
    }
    for (int e = lane; e < d.E; e += FM_WAVE) {
        const Cd v = b.cchi[(size_t)c * d.E + e];
        amp[e] = sqrt(v.r * v.r + v.i * v.i + 1e-30);
        ang[e] = atan2(v.i, v.r);
    }
    wave_sync();
    fine_grid<double>(ed, G.jq + (size_t)edge_id * nk, G.fq + (size_t)edge_id * nk, G.q, nk, reff_A,
                      sigma2 ? sigma2[p] : 0.0, amp, ang, out, G.S ? G.S + (size_t)edge_id * nk * G.sstride : nullptr,
                      G.sstride);
}

int set_err(int code, const char *msg) { return fm_set_err(code, msg); }
int hip_err(hipError_t e, const char *where) {
    char buf[256];
    snprintf(buf, sizeof buf, "%s: %s", where, hipGetErrorString(e));
    return fm_set_err(FM_EHIP, buf);
}

LamRt lam_of(int set) {
    const LamDef D = lamdef(set);
    LamRt r;
    r.L = D.L;
    r.L0 = D.L0;
    int mmax = 0, nmax = 0;
    for (int i = 0; i < 6; ++i) {
        r.M[i] = D.M[i];
        r.N[i] = D.N[i];
        if (i < D.L) {
            mmax = max(mmax, D.M[i] < 0 ? -D.M[i] : D.M[i]);
            nmax = max(nmax, D.N[i]);
        }
    }
    r.mnp1 = mmax + 1 + nmax;   // min(L1, mmaxp1 + nmax), clipped by the caller
    return r;
}

PDims dims_of(const fm_edge &ed, int nleg, int set, int C) {
    PDims d;
    d.C = C;
    d.n = nleg;
    d.E = ed.E;
    d.L1 = ed.L1;
    d.LC = ed.L1 + 2;
    const LamRt lam = lam_of(set);
    const int mnp1 = min(ed.L1, lam.mnp1);
    d.MU = mnp1 + 1;
    return d;
}

inline unsigned nblk(size_t threads) { return (unsigned)((threads + 255) / 256); }

}  // namespace

extern "C" {

size_t fm_plain_path_bytes(const fm_edge *edge, int32_t nleg) {
    if (edge_problem(edge) || nleg < 2 || nleg > FM_MAX_LEGS) return 0;
    // a chunk of C paths needs at most C times the arrays of one (each array is rounded up to 256 bytes)
    return carve(dims_of(*edge, nleg, nleg == 2 ? 2 : 1, 1), nullptr, nullptr);
}

int fm_chi_paths_plain(const fm_edge *edges, int32_t nedge, const fm_paths *paths, const fm_grid *grid, int32_t iorder,
                       const int32_t *order, int32_t nbucket, const int32_t *bucket_offset, const int32_t *bucket_kind,
                       const int32_t *bucket_edge, const int32_t *bucket_nleg, int32_t nseg, const int32_t *seg_ptr,
                       const int32_t *seg_paths, double *chi_seg, double *chi_path, double *reff, void *workspace,
                       size_t workspace_bytes, void *stream) {
    if (nedge < 1 || nedge > FM_MAX_EDGE) return set_err(FM_EINVAL, "nedge out of range 1..4");
    for (int i = 0; i < nedge; ++i) {
        const char *m = edge_problem(&edges[i]);
        if (m) return set_err(FM_EINVAL, m);
        if (edges[i].ilinit < 1 || edges[i].ilinit > 3) return set_err(FM_EUNSUPPORTED, "edge: only K, L and M edges");
        if (!edges[i].tl64 || !edges[i].chw64) return set_err(FM_EINVAL, "edge: NULL FP64 table");
    }
    if (iorder != 2) return set_err(FM_EUNSUPPORTED, "only iorder = 2 is implemented on the GPU");
    if (!paths || !grid || !chi_path || !workspace) return set_err(FM_EINVAL, "NULL paths, grid, chi_path or workspace");
    if (paths->nmax < 2 || paths->nmax > FM_MAX_LEGS) return set_err(FM_EINVAL, "nmax out of range 2..9");
    if (nbucket > 0 && (!bucket_offset || !bucket_kind || !bucket_nleg)) return set_err(FM_EINVAL, "NULL bucket arrays");
    hipStream_t st = (hipStream_t)stream;
    for (int bk = 0; bk < nbucket; ++bk) {
        const int cnt = bucket_offset[bk + 1] - bucket_offset[bk];
        if (cnt <= 0) continue;
        const int set = bucket_kind[bk];
        const int nleg = bucket_nleg[bk];
        const int eid = bucket_edge ? bucket_edge[bk] : 0;
        if (set < 0 || set > 2) return set_err(FM_EUNSUPPORTED, "bucket kind not implemented on the GPU");
        if (eid < 0 || eid >= nedge) return set_err(FM_EINVAL, "bucket edge out of range");
        if (nleg < 2 || nleg > paths->nmax || (set == 2) != (nleg == 2)) return set_err(FM_EINVAL, "bucket nleg out of range");
        const fm_edge &ed = edges[eid];
        if ((set == 0 || set == 2) != (ed.ilinit <= 1))
            return set_err(set == 2 ? FM_EUNSUPPORTED : FM_EINVAL,
                           set == 2 ? "the plain pipeline runs two-leg paths of K and L1 edges only"
                                    : "bucket kind does not match its edge");
        LamRt lam = lam_of(set);
        lam.mnp1 = min(ed.L1, lam.mnp1);
        // chunk: as many paths as the workspace holds
        int C = 1;
        {
            const size_t per = fm_plain_path_bytes(&ed, nleg);
            if (per == 0) return set_err(FM_EINVAL, "bad edge or nleg");
            size_t cfit = workspace_bytes / per;
            if (cfit < 1) return set_err(FM_EINVAL, "workspace smaller than fm_plain_path_bytes");
            C = (int)(cfit < (size_t)cnt ? cfit : (size_t)cnt);
        }
        for (int c0 = 0; c0 < cnt; c0 += C) {
            const int Cc = cnt - c0 < C ? cnt - c0 : C;
            const PDims d = dims_of(ed, nleg, set, Cc);
            PBuf b;
            carve(d, (unsigned char *)workspace, &b);
            const int first = bucket_offset[bk] + c0;
            hipLaunchKernelGGL(k_pl_geo, dim3(nblk(Cc)), dim3(256), 0, st, d, *paths, order, first, b);
            hipLaunchKernelGGL(k_pl_wigner, dim3(nblk((size_t)Cc * nleg * 25)), dim3(256), 0, st, d, ed.rotc, b);
            hipLaunchKernelGGL(k_pl_clm, dim3(nblk((size_t)Cc * nleg * d.E)), dim3(256), 0, st, d, ed.ck, ed.lmax, lam.mnp1, b);
            hipLaunchKernelGGL(k_pl_gam, dim3(nblk((size_t)Cc * nleg * d.LC * 6 * d.E)), dim3(256), 0, st, d, lam, ed.tl64, b);
            hipLaunchKernelGGL(k_pl_fmat, dim3(nblk((size_t)Cc * (nleg - 1) * 36 * d.E)), dim3(256), 0, st, d, lam, ed.lmax, b);
            hipLaunchKernelGGL(k_pl_term, dim3(nblk((size_t)Cc * 36 * d.E)), dim3(256), 0, st, d, lam, ed, b);
            hipLaunchKernelGGL(k_pl_chain, dim3(nblk((size_t)Cc * d.E)), dim3(256), 0, st, d, lam, ed.ck, ed.ckmxk, b);
            hipLaunchKernelGGL(k_pl_fine, dim3((Cc + 3) / 4), dim3(FM_WAVE * 4), 4 * 2 * FM_MAX_E * sizeof(double), st, d, ed,
                               *grid, eid, paths->sigma2, b, chi_path, reff);
            hipError_t e = hipGetLastError();
            if (e != hipSuccess) return hip_err(e, "fm_chi_paths_plain launch");
        }
    }
    if (chi_seg && nseg > 0) {
        if (!seg_ptr || !seg_paths) return set_err(FM_EINVAL, "NULL segment lists");
        dim3 grd((grid->nk + FM_WAVE - 1) / FM_WAVE, nseg);
        hipLaunchKernelGGL(k_segment_sum<double>, grd, dim3(FM_WAVE * SEG_WAVES), 0, st, nseg, grid->nk, seg_ptr, seg_paths,
                           paths->deg, (const double *)chi_path, chi_seg);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return hip_err(e, "segment sum launch");
    }
    return FM_OK;
}

}  // extern "C"
