// Fused path-stage forward kernels for gfx950 and the C ABI that launches them.
//
// One wavefront per path, one lane per phase-grid energy (energies beyond 64 loop). Per path the wave computes
// the leg geometry in FP64 (one lane per leg), the rotation tables d^l_{m1 m2}(beta_s) of every leg into LDS (one
// lane per (leg, pair)), then every lane evaluates the Rehr-Albers amplitude at its energy: the termination matrix
// T, then the scattering matrices F_{n-1} ... F_1 multiplied into U = T F_{n-1} ... F_s, each F streamed over the
// partial waves with the spherical-wave recursion of its two legs kept in registers; cchi = tr(U) cfac. The fine
// grid (amplitude and unwrapped phase interpolated to q, chi = A sin(Phi + 2 q reff)) is the wave's epilogue.
//
// Reference semantics: GPUPathStage.cchi / chi_fine (src/feffgpu/stage.py:235-388), the literal port
// genfmt.PathStage (src/feffgpu/genfmt.py:395-445) and the path-list gather/sum of
// src/feffgpu/exafs/forward.py:143-163.
#include <hip/hip_runtime.h>
#include <stdio.h>
#include <string.h>

#include <type_traits>

#include "fm_device.h"

using namespace fm;

thread_local char g_err[512] = "";
int fm_set_err(int code, const char *msg) {
    snprintf(g_err, sizeof g_err, "%s", msg);
    return code;
}

namespace {

int set_err(int code, const char *msg) {
    snprintf(g_err, sizeof g_err, "%s", msg);
    return code;
}
int hip_err(hipError_t e, const char *where) {
    snprintf(g_err, sizeof g_err, "%s: %s", where, hipGetErrorString(e));
    return FM_EHIP;
}

constexpr int FM_MAX_EDGE = 4;
constexpr int WPB_MAX = 4;        // wavefronts (paths) per workgroup at most
constexpr int FM_MAX_ECHUNK = 4;  // energies up to 4 x 64
#define FM_WPE(T) (sizeof(T) == 4 ? 2 : 1)   // wavefronts per SIMD the register budget is sized for

// ------------------------------------------------------------------------------------------------ per-wave LDS
struct PathLDS {
    double rat[FM_MAX_LEGS + 1][3];   // bohr, rat[0] = absorber (origin), rat[1..n-1] scatterers
    double r[FM_MAX_LEGS + 1], beta[FM_MAX_LEGS + 1], alpha[FM_MAX_LEGS + 1], gamma[FM_MAX_LEGS + 1];
    double eta[FM_MAX_LEGS + 1];
    double c1[FM_MAX_LEGS + 1], s1[FM_MAX_LEGS + 1], c2[FM_MAX_LEGS + 1], s2[FM_MAX_LEGS + 1];  // cos/sin eta, 2 eta
    int ipot[FM_MAX_LEGS + 1];       // potential of path atom j (0 and n: absorber)
    double reff, R, prod_r;           // bohr
    int n, pad_[3];
};

struct KArgs {
    fm_edge ed[FM_MAX_EDGE];
    int nedge;
    int mode;             // 0: coords given (cchi), 1: path list (gather + fine grid)
    int npath_bucket;     // paths in this launch
    int first;            // offset into order[]
    const int32_t *order;
    // mode 0
    int nleg;
    const double *coords;
    const int64_t *ipots64;
    void *cchi_out;
    // mode 1
    fm_paths P;
    fm_grid G;
    void *chi_path;
    double *reff_out;
    int ltab;             // L1 + 1 rows per rotation table
    int emax;             // largest E over the edges
    int wpb;              // wavefronts per workgroup
    size_t wave_bytes;
};

__host__ __device__ inline size_t align16(size_t x) { return (x + 15) & ~size_t(15); }

template <class T> __host__ __device__ inline size_t wave_lds_bytes(int nlegmax, int ltab, int emax) {
    return align16(sizeof(PathLDS)) + align16(sizeof(T) * (size_t)nlegmax * ltab * FM_NREP) +
           align16(sizeof(T) * (size_t)emax) + align16(sizeof(double) * (size_t)emax);
}

// ------------------------------------------------------------------------------------------------ setup of one path
// Fills g (geometry, potentials) for path slot `p`; all lanes call it. Returns n.
__device__ int load_path(const KArgs &a, int p, PathLDS &g, int &edge_idx) {
    const int lane = lane_id();
    int n;
    if (a.mode == 0) {
        n = a.nleg;
        edge_idx = 0;
        if (lane < n) {
            const double *c = a.coords + ((size_t)p * n + lane) * 3;
            double x = c[0] / FM_BOHR, y = c[1] / FM_BOHR, z = c[2] / FM_BOHR;
            // absorber (lane 0) is the path's first point; the stage places it at rat[0] and rat[n]
            g.rat[lane][0] = x; g.rat[lane][1] = y; g.rat[lane][2] = z;
            g.ipot[lane] = (int)a.ipots64[(size_t)p * n + lane];
        }
    } else {
        n = a.P.nleg[p];
        edge_idx = a.P.edge ? a.P.edge[p] : 0;
        const int nmax = a.P.nmax;
        if (lane < n) {
            double v[3] = {0.0, 0.0, 0.0};
            if (lane > 0) {
                const int at = a.P.atoms[(size_t)p * (nmax - 1) + lane - 1];
                const int ab = a.P.absorber[p];
#pragma unroll
                for (int c = 0; c < 3; ++c) v[c] = a.P.pos[3 * (size_t)at + c] - a.P.pos[3 * (size_t)ab + c];
                if (a.P.cell_kind == 1) {
                    const double L = a.P.cell[0];
#pragma unroll
                    for (int c = 0; c < 3; ++c) v[c] -= L * rint(v[c] / L);
                } else if (a.P.cell_kind == 2) {
                    double f[3];
#pragma unroll
                    for (int c = 0; c < 3; ++c)
                        f[c] = v[0] * a.P.cellinv[0 * 3 + c] + v[1] * a.P.cellinv[1 * 3 + c] + v[2] * a.P.cellinv[2 * 3 + c];
#pragma unroll
                    for (int c = 0; c < 3; ++c) f[c] -= rint(f[c]);
#pragma unroll
                    for (int c = 0; c < 3; ++c)
                        v[c] = f[0] * a.P.cell[0 * 3 + c] + f[1] * a.P.cell[1 * 3 + c] + f[2] * a.P.cell[2 * 3 + c];
                }
            }
#pragma unroll
            for (int c = 0; c < 3; ++c) g.rat[lane][c] = v[c] / FM_BOHR;
            g.ipot[lane] = a.P.ipots[(size_t)p * nmax + lane];
        }
    }
    wave_sync();
    if (lane < 3) g.rat[n][lane] = g.rat[0][lane];
    if (lane == 0) { g.ipot[n] = g.ipot[0]; g.n = n; }
    wave_sync();
    // one lane per leg
    if (lane < n) {
        LegGeo lg = leg_geometry(g.rat, lane + 1, n);
        g.r[lane + 1] = lg.r;
        g.beta[lane + 1] = lg.beta;
        g.alpha[lane + 1] = lg.alpha;
        g.gamma[lane + 1] = lg.gamma;
    }
    wave_sync();
    if (lane < n) {
        const int J = lane + 1;
        double eta = (J == 1 ? g.alpha[n] : g.alpha[J - 1]) + g.gamma[J];
        g.eta[J] = eta;
        double s, c;
        sincos(eta, &s, &c);
        g.c1[J] = c; g.s1[J] = s;
        sincos(2.0 * eta, &s, &c);
        g.c2[J] = c; g.s2[J] = s;
    }
    if (lane == 0) {
        double R = 0.0, pr = 1.0;
        for (int j = 1; j <= n; ++j) { R += g.r[j]; pr *= g.r[j]; }
        g.R = R; g.reff = 0.5 * R; g.prod_r = pr;
    }
    wave_sync();
    return n;
}

// Rotation tables: W[s][il][k] = d^{il-1}_{m1 m2}(beta_{s+1}) xnlm(il,|m1|)(-1)^|m1| / xnlm(il,|m2|) for legs
// s = 0..n-1 and the FM_NREP representatives k, il = 1..L1 (row 0 unused). Three-term recursion in l for fixed
// (m1, m2) (src/feffgpu/stage.py:77-92), seeded by the closed forms at l <= 2, in FP64.
template <class T, int SET>
__device__ void build_rotation(const fm_edge &ed, const PathLDS &g, int n, int ltab, T *W) {
    constexpr PairDef PD = pairdef(SET);
    const int lane = lane_id();
    const int ntask = n * FM_NREP;
    for (int t = lane; t < ntask; t += FM_WAVE) {
        const int s = t / FM_NREP, k = t % FM_NREP;
        int m1 = 0, m2 = 0;
#pragma unroll
        for (int q = 0; q < FM_NREP; ++q)
            if (q == k) { m1 = PD.m1[q]; m2 = PD.m2[q]; }
        const double be = g.beta[s + 1];
        double c = cos(be), sn = sin(be), c2 = cos(0.5 * be), s2 = sin(0.5 * be);
        T *Ws = W + (size_t)s * ltab * FM_NREP;
        const int am1 = m1 < 0 ? -m1 : m1, am2 = m2 < 0 ? -m2 : m2;
        const int mm = (m1 + 3) * 7 + (m2 + 3), nn = am1 * 4 + am2;
        double dlm1 = 0.0, dl = 0.0;
        for (int il = 1; il <= ed.L1; ++il) {
            const int l = il - 1;
            double d;
            if (l <= 2) {
                d = dseed(l, m1, m2, c, sn, c2, s2);
            } else {
                const double *co = ed.rotc + ((size_t)(l - 1) * 49 + mm) * 3;   // from l - 1 to l
                d = (co[0] * c - co[1]) * dl - co[2] * dlm1;
            }
            dlm1 = dl;
            dl = d;
            Ws[(size_t)il * FM_NREP + k] = (T)(d * ed.rotn[(size_t)il * 16 + nn]);
        }
    }
    wave_sync();
}

// One partial-wave row of a scattering matrix: Ft[k] += X_a(k) Y_b(k) w[k] for the representatives k of the
// NR x NC block (Y already carries the t-matrix).
template <class T, int SET, int NR, int NC>
__device__ __forceinline__ void accum_row(cx<T> (&Ft)[FM_NREP], const cx<T> (&xa)[3], cx<T> y0, cx<T> y1,
                                          const T *__restrict__ w) {
    constexpr PairDef PD = pairdef(SET);
    cx<T> Z[3][2];
#pragma unroll
    for (int aa = 0; aa < 3; ++aa) {
        Z[aa][0] = xa[aa] * y0;
        Z[aa][1] = xa[aa] * y1;
    }
#pragma unroll
    for (int k = 0; k < FM_NREP; ++k)
        if (PD.ri[k] < NR && PD.rj[k] < NC) rfma(Ft[k], Z[PD.a[k] - 1][PD.b[k] - 1], w[k]);
}

// F_s of the path at the lane's energy: X = Hankel factors of leg s+1 (gam), Y = of leg s (gamtl) times the
// scatterer's t-matrix, summed over il <= ilx with the rotation table of leg s (src/feffgpu/stage.py:287-305).
// Accumulates the representatives k with ri < NR, rj < NC into Ft (zeroed here). Rows up to ilmax (the wave's
// largest ilx) are generated; a lane adds only its rows il <= ilx.
template <class T, int SET, int NR, int NC>
__device__ __forceinline__ void fblock(cx<T> (&Ft)[FM_NREP], cx<T> ck, T rX, T rY, int ilx, int ilmax,
                                       const T *__restrict__ Wl, const T *__restrict__ tl, int tlstride) {
#pragma unroll
    for (int k = 0; k < FM_NREP; ++k) Ft[k] = mk<T>(T(0), T(0));
    Hank3<T> X;
    Hank2<T> Y;
    X.init(minus_i_over(ck * rX));
    Y.init(minus_i_over(ck * rY));
    cx<T> xa[3], yb[2];
    const int rseed = ilmax < 4 ? ilmax : 4;
    for (int r = 1; r <= rseed; ++r) {
        X.seed_row(r, xa);
        if (r < 4) Y.seed_row(r, yb); else Y.step(r, yb);
        const cx<T> t = mk<T>(tl[(size_t)r * tlstride], tl[(size_t)r * tlstride + 1]);
        if (r <= ilx) accum_row<T, SET, NR, NC>(Ft, xa, yb[0] * t, yb[1] * t, Wl + (size_t)r * FM_NREP);
    }
#pragma unroll 1
    for (int r = 5; r <= ilmax; ++r) {
        X.step(r, xa);
        Y.step(r, yb);
        const cx<T> t = mk<T>(tl[(size_t)r * tlstride], tl[(size_t)r * tlstride + 1]);
        const T *w = Wl + (size_t)r * FM_NREP;
        if (r <= ilx) accum_row<T, SET, NR, NC>(Ft, xa, yb[0] * t, yb[1] * t, w);
    }
}

// The row phases exp(-i eta m) for m = M[i].
template <class T, int SET>
__device__ __forceinline__ void row_phases(cx<T> (&ph)[6], const PathLDS &g, int J) {
    constexpr LamDef D = lamdef(SET);
    const T c1 = (T)g.c1[J], s1 = (T)g.s1[J], c2 = (T)g.c2[J], s2 = (T)g.s2[J];
#pragma unroll
    for (int i = 0; i < 6; ++i) {
        const int m = D.M[i];
        ph[i] = m == 0 ? mk<T>(T(1), T(0))
                : m == 1 ? mk<T>(c1, -s1)
                : m == -1 ? mk<T>(c1, s1)
                : m == 2 ? mk<T>(c2, -s2)
                          : mk<T>(c2, s2);
    }
}

template <class T> struct TL;
template <> struct TL<float> {
    static __device__ __forceinline__ const float *tl(const fm_edge &e) { return e.tl32; }
    static __device__ __forceinline__ const float *chw(const fm_edge &e) { return e.chw32; }
};
template <> struct TL<double> {
    static __device__ __forceinline__ const double *tl(const fm_edge &e) { return e.tl64; }
    static __device__ __forceinline__ const double *chw(const fm_edge &e) { return e.chw64; }
};

// cchi at energy e of the path held in g (tables in W). All lanes of the wave run it (e may be out of range for
// some; they compute garbage that the caller drops, which keeps the loops uniform).
template <class T, int SET>
__device__ cx<T> path_cchi(const fm_edge &ed, const PathLDS &g, const T *W, int ltab, int e, int n) {
    constexpr LamDef D = lamdef(SET);
    constexpr PairDef PD = pairdef(SET);
    constexpr int L = D.L, L0 = D.L0;
    const int E = ed.E;
    const int ee = e < E ? e : E - 1;
    const cx<T> ck = mk<T>((T)ed.ck[2 * ee], (T)ed.ck[2 * ee + 1]);
    const int tlstride = 2 * E;
    const T *tlbase = TL<T>::tl(ed);
    auto lmaxA = [&](int j) -> int { return ed.lmax[(size_t)g.ipot[j] * E + ee]; };

    // ---- termination T (mmtrxi + mmtr, src/feffgpu/stage.py:318-334): X = leg 1, Y = leg n, channel rows only
    cx<T> U[L0][L];
    {
        const int lx1 = max(lmaxA(0), lmaxA(1)) + 1;          // leg 1: absorber -> atom 1
        const int lxn = max(lmaxA(n - 1), lmaxA(n)) + 1;      // leg n: atom n-1 -> absorber
        Hank3<T> X;
        Hank2<T> Y;
        X.init(minus_i_over(ck * (T)g.r[1]));
        Y.init(minus_i_over(ck * (T)g.r[n]));
        cx<T> Tt[FM_NREP];
#pragma unroll
        for (int k = 0; k < FM_NREP; ++k) Tt[k] = mk<T>(T(0), T(0));
        const T *Wn = W + (size_t)(n - 1) * ltab * FM_NREP;
        const T *chw = TL<T>::chw(ed);
        cx<T> xa[3], yb[2];
        for (int r = 1; r <= 4; ++r) {
            X.seed_row(r, xa, r <= lx1);
            if (r < 4) Y.seed_row(r, yb, r <= lxn); else Y.step(r, yb, r <= lxn);
            int c = -1;
            for (int q = 0; q < ed.nchan; ++q)
                if (ed.chan_il[q] == r) c = q;
            if (c >= 0) {
                const cx<T> w = mk<T>(chw[((size_t)c * E + ee) * 2], chw[((size_t)c * E + ee) * 2 + 1]) * T(2 * r - 1);
                accum_row<T, SET, L0, L0>(Tt, xa, yb[0] * w, yb[1] * w, Wn + (size_t)r * FM_NREP);
            }
        }
        cx<T> ph[6];
        row_phases<T, SET>(ph, g, 1);
#pragma unroll
        for (int i = 0; i < L0; ++i)
#pragma unroll
            for (int j = 0; j < L; ++j)
                U[i][j] = (j < L0) ? ph[i] * Tt[PD.idx[i][j]] * T(PD.sg[i][j]) : mk<T>(T(0), T(0));
    }

    cx<T> ptrac = mk<T>(T(0), T(0));
    cx<T> Ft[FM_NREP];
    cx<T> ph[6];
    auto scat = [&](int s, int &ilx, int &ilmax, const T *&tl, const T *&Wl) {
        const int ip = g.ipot[s];
        ilx = ed.lmax[(size_t)ip * E + ee] + 1;
        ilmax = ed.lmax1[ip];
        tl = tlbase + ((size_t)ip * ltab) * tlstride + 2 * ee;
        Wl = W + (size_t)(s - 1) * ltab * FM_NREP;
    };
    // U <- U F with F = diag(ph) Ft (expanded), row by row of U; U has NCU valid columns (rows of F)
    auto chain = [&](auto ncu_tag) {
        constexpr int NCU = decltype(ncu_tag)::value;
#pragma unroll
        for (int a = 0; a < L0; ++a) {
            cx<T> row[L];
#pragma unroll
            for (int i = 0; i < L; ++i) row[i] = (i < NCU) ? U[a][i] * ph[i] : mk<T>(T(0), T(0));
#pragma unroll
            for (int j = 0; j < L; ++j) {
                cx<T> acc = mk<T>(T(0), T(0));
#pragma unroll
                for (int i = 0; i < NCU; ++i) {
                    const cx<T> f = Ft[PD.idx[i][j]];
                    if (PD.sg[i][j] > 0) cfma(acc, row[i], f);
                    else cfma(acc, row[i], mk<T>(-f.r, -f.i));
                }
                U[a][j] = acc;
            }
        }
    };
    if (n == 2) {
        int ilx, ilmax;
        const T *tl, *Wl;
        scat(1, ilx, ilmax, tl, Wl);
        fblock<T, SET, L0, L0>(Ft, ck, (T)g.r[2], (T)g.r[1], ilx, ilmax, Wl, tl, tlstride);
        row_phases<T, SET>(ph, g, 2);
#pragma unroll
        for (int a = 0; a < L0; ++a)
#pragma unroll
            for (int b = 0; b < L0; ++b) {
                const cx<T> f = Ft[PD.idx[b][a]] * T(PD.sg[b][a]);
                cfma(ptrac, U[a][b] * ph[b], f);
            }
    } else {
        // s = n - 1: F is L0 x L; U (L0 x L0) -> U F (L0 x L)
        {
            int ilx, ilmax;
            const T *tl, *Wl;
            scat(n - 1, ilx, ilmax, tl, Wl);
            fblock<T, SET, L0, L>(Ft, ck, (T)g.r[n], (T)g.r[n - 1], ilx, ilmax, Wl, tl, tlstride);
            row_phases<T, SET>(ph, g, n);
            chain(std::integral_constant<int, L0>{});
        }
#pragma unroll 1
        for (int s = n - 2; s >= 2; --s) {
            int ilx, ilmax;
            const T *tl, *Wl;
            scat(s, ilx, ilmax, tl, Wl);
            fblock<T, SET, L, L>(Ft, ck, (T)g.r[s + 1], (T)g.r[s], ilx, ilmax, Wl, tl, tlstride);
            row_phases<T, SET>(ph, g, s + 1);
            chain(std::integral_constant<int, L>{});
        }
        // s = 1: F is L x L0; trace of U (L0 x L) F (L x L0)
        {
            int ilx, ilmax;
            const T *tl, *Wl;
            scat(1, ilx, ilmax, tl, Wl);
            fblock<T, SET, L, L0>(Ft, ck, (T)g.r[2], (T)g.r[1], ilx, ilmax, Wl, tl, tlstride);
            row_phases<T, SET>(ph, g, 2);
#pragma unroll
            for (int a = 0; a < L0; ++a)
#pragma unroll
                for (int i = 0; i < L; ++i) {
                    const cx<T> f = Ft[PD.idx[i][a]] * T(PD.sg[i][a]);
                    cfma(ptrac, U[a][i] * ph[i], f);
                }
        }
    }
    // cfac = exp(i R (ck - xk)) / prod_j (ck r_j), R = sum r_j = 2 reff (src/feffgpu/stage.py:336-337); the phase
    // R Re(ck - xk) in FP64 reduced mod 2 pi
    const double dkr = ed.ckmxk[2 * ee], dki = ed.ckmxk[2 * ee + 1];
    double phs = g.R * dkr;
    phs -= FM_TWO_PI * rint(phs / FM_TWO_PI);
    const double damp = exp(-g.R * dki);
    T sp, cp;
    if constexpr (sizeof(T) == 4) {
        sincosf((float)phs, &sp, &cp);
    } else {
        sincos(phs, &sp, &cp);
    }
    cx<T> num = mk<T>((T)damp * cp, (T)damp * sp);
    cx<T> ckn = ck;
    for (int j = 1; j < n; ++j) ckn = ckn * ck;
    cx<T> den = ckn * (T)g.prod_r;
    return cdiv(ptrac * num, den);
}

// Fine-grid epilogue for the path held by the wave: amp/phase of cchi (lanes = energies, already in LDS as
// amp[] and ang[]), unwrap (FP64 scan), interpolation to q, chi = A sin(Phi + 2 q reff) exp(-2 sigma2 q^2).
template <class T>
__device__ void fine_grid(const fm_edge &ed, const int32_t *jq, const double *fq, const double *q, int nk, double reff_A,
                          double sigma2, T *amp, double *ang, T *out) {
    const int lane = lane_id();
    const int E = ed.E;
    // unwrap: pha[e] = ang[0] + sum_{i<=e} wrap(ang[i] - ang[i-1]) (src/feffgpu/stage.py:365-370)
    const double a0 = ang[0];
    double dl[FM_MAX_ECHUNK];
#pragma unroll
    for (int c = 0; c < FM_MAX_ECHUNK; ++c) {
        const int e = c * FM_WAVE + lane;
        double d = 0.0;
        if (e < E && e > 0) {
            d = ang[e] - ang[e - 1];
            d -= FM_TWO_PI * rint(d / FM_TWO_PI);
        }
        dl[c] = d;
    }
    wave_sync();
    double carry = 0.0;
#pragma unroll
    for (int c = 0; c < FM_MAX_ECHUNK; ++c) {
        if (c * FM_WAVE >= E) break;
        const int e = c * FM_WAVE + lane;
        const double s = wave_scan_add(dl[c]) + carry;
        carry = __shfl(s, FM_WAVE - 1, FM_WAVE);
        if (e < E) ang[e] = a0 + s;
    }
    wave_sync();
    for (int i = lane; i < nk; i += FM_WAVE) {
        const int j = jq[i];
        const double f = fq[i];
        const T A = (T)((1.0 - f) * (double)amp[j] + f * (double)amp[j + 1]);
        double ph = (1.0 - f) * ang[j] + f * ang[j + 1] + 2.0 * q[i] * reff_A;
        ph -= FM_TWO_PI * rint(ph / FM_TWO_PI);
        T v;
        if constexpr (sizeof(T) == 4) v = A * sinf((float)ph);
        else v = A * sin(ph);
        if (sigma2 != 0.0) v *= (T)exp(-2.0 * sigma2 * q[i] * q[i]);
        out[i] = v;
    }
}

template <class T, int SET>
__global__ __launch_bounds__(FM_WAVE *WPB_MAX) __attribute__((amdgpu_waves_per_eu(FM_WPE(T)))) void k_forward(KArgs a) {
    extern __shared__ __align__(16) unsigned char smem[];
    const int wid = threadIdx.x / FM_WAVE;
    const int lane = lane_id();
    unsigned char *base = smem + (size_t)wid * a.wave_bytes;
    PathLDS &g = *reinterpret_cast<PathLDS *>(base);
    T *W = reinterpret_cast<T *>(base + align16(sizeof(PathLDS)));
    for (int slot = blockIdx.x * a.wpb + wid; slot < a.npath_bucket; slot += gridDim.x * a.wpb) {
        const int p = a.order ? a.order[a.first + slot] : a.first + slot;
        int ei;
        const int n = load_path(a, p, g, ei);
        const fm_edge &ed = a.ed[ei];
        const int ltab = ed.L1 + 1;
        build_rotation<T, SET>(ed, g, n, ltab, W);
        T *amp = reinterpret_cast<T *>(base + a.wave_bytes - align16(sizeof(double) * (size_t)a.emax) -
                                       align16(sizeof(T) * (size_t)a.emax));
        double *ang = reinterpret_cast<double *>(base + a.wave_bytes - align16(sizeof(double) * (size_t)a.emax));
        for (int e0 = 0; e0 < ed.E; e0 += FM_WAVE) {
            const int e = e0 + lane;
            cx<T> c = path_cchi<T, SET>(ed, g, W, ltab, e, n);
            if (e < ed.E) {
                if (a.mode == 0) {
                    T *o = reinterpret_cast<T *>(a.cchi_out) + ((size_t)p * ed.E + e) * 2;
                    o[0] = c.r;
                    o[1] = c.i;
                } else {
                    amp[e] = (T)sqrt((double)c.r * c.r + (double)c.i * c.i + 1e-30);
                    ang[e] = (double)atan2(c.i, c.r);
                }
            }
        }
        const double reff_A = g.reff * FM_BOHR;
        if (a.mode == 0) {
            if (lane == 0 && a.reff_out) a.reff_out[p] = reff_A;
        } else {
            wave_sync();
            if (lane == 0 && a.reff_out) a.reff_out[p] = reff_A;
            const int nk = a.G.nk;
            T *out = reinterpret_cast<T *>(a.chi_path) + (size_t)p * nk;
            fine_grid<T>(ed, a.G.jq + (size_t)ei * nk, a.G.fq + (size_t)ei * nk, a.G.q, nk, reff_A,
                         a.P.sigma2 ? a.P.sigma2[p] : 0.0, amp, ang, out);
        }
        wave_sync();
    }
}

// GPUPathStage.chi_fine on its own: one wave per path reads cchi [npath][E] and runs the fine-grid epilogue.
template <class T>
__global__ __launch_bounds__(FM_WAVE *WPB_MAX) void k_chi_fine(fm_edge ed, int npath, const T *cchi, const double *reff,
                                                                fm_grid G, T *chi, size_t wave_bytes) {
    extern __shared__ __align__(16) unsigned char smem[];
    const int wid = threadIdx.x / FM_WAVE, lane = lane_id();
    const int wpb = blockDim.x / FM_WAVE;
    unsigned char *base = smem + (size_t)wid * wave_bytes;
    T *amp = reinterpret_cast<T *>(base);
    double *ang = reinterpret_cast<double *>(base + align16(sizeof(T) * (size_t)ed.E));
    for (int p = blockIdx.x * wpb + wid; p < npath; p += gridDim.x * wpb) {
        for (int e = lane; e < ed.E; e += FM_WAVE) {
            const T cr = cchi[((size_t)p * ed.E + e) * 2], ci = cchi[((size_t)p * ed.E + e) * 2 + 1];
            amp[e] = (T)sqrt((double)cr * cr + (double)ci * ci + 1e-30);
            ang[e] = (double)atan2(ci, cr);
        }
        wave_sync();
        fine_grid<T>(ed, G.jq, G.fq, G.q, G.nk, reff[p], 0.0, amp, ang, chi + (size_t)p * G.nk);
        wave_sync();
    }
}

// chi_seg[s][k] = sum over the CSR list of segment s of deg[p] chi_path[p][k], in list order (deterministic).
template <class T>
__global__ void k_segment_sum(int nseg, int nk, const int32_t *seg_ptr, const int32_t *seg_paths, const double *deg,
                              const T *chi_path, double *chi_seg) {
    const int s = blockIdx.y;
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nseg || k >= nk) return;
    double acc = 0.0;
    for (int t = seg_ptr[s]; t < seg_ptr[s + 1]; ++t) {
        const int p = seg_paths[t];
        acc += deg[p] * (double)chi_path[(size_t)p * nk + k];
    }
    chi_seg[(size_t)s * nk + k] = acc;
}

// ------------------------------------------------------------------------------------------------ launch helpers
int set_for(const fm_edge &ed, int nleg) {
    if (ed.ilinit <= 1) return nleg == 2 ? 2 : 0;
    return nleg == 2 ? -1 : 1;
}

template <class T, int SET> hipError_t launch_forward(KArgs &a, hipStream_t st, int nlegmax) {
    int ltab = 0, emax = 0;
    for (int i = 0; i < a.nedge; ++i) {
        ltab = ltab > a.ed[i].L1 + 1 ? ltab : a.ed[i].L1 + 1;
        emax = emax > a.ed[i].E ? emax : a.ed[i].E;
    }
    a.ltab = ltab;
    a.emax = emax;
    a.wave_bytes = wave_lds_bytes<T>(nlegmax, ltab, emax);
    if (a.wave_bytes > 160 * 1024) return hipErrorInvalidValue;
    int wpb = (int)((64 * 1024) / a.wave_bytes);
    wpb = wpb < 1 ? 1 : (wpb > WPB_MAX ? WPB_MAX : wpb);
    a.wpb = wpb;
    const size_t shm = a.wave_bytes * wpb;
    if (shm > 64 * 1024) {
        hipError_t e = hipFuncSetAttribute((const void *)k_forward<T, SET>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)shm);
        if (e != hipSuccess) return e;
    }
    const int blocks = (a.npath_bucket + wpb - 1) / wpb;
    if (blocks == 0) return hipSuccess;
    hipLaunchKernelGGL((k_forward<T, SET>), dim3(blocks), dim3(FM_WAVE * wpb), shm, st, a);
    return hipGetLastError();
}

template <class T> hipError_t dispatch(KArgs &a, int set, hipStream_t st, int nlegmax) {
    switch (set) {
        case 0: return launch_forward<T, 0>(a, st, nlegmax);
        case 1: return launch_forward<T, 1>(a, st, nlegmax);
        case 2: return launch_forward<T, 2>(a, st, nlegmax);
    }
    return hipErrorInvalidValue;
}

int check_edge(const fm_edge *e) {
    if (!e) return set_err(FM_EINVAL, "edge table is NULL");
    if (e->E < 2 || e->E > 64 * FM_MAX_ECHUNK) return set_err(FM_EINVAL, "edge: E out of range 2..256");
    if (e->L1 < 1 || e->L1 > 120) return set_err(FM_EINVAL, "edge: L1 out of range 1..120");
    if (e->nph1 < 1 || e->nph1 > FM_MAX_POT) return set_err(FM_EINVAL, "edge: nph1 out of range");
    if (e->nchan < 1 || e->nchan > FM_MAX_CHAN) return set_err(FM_EINVAL, "edge: nchan out of range");
    for (int c = 0; c < e->nchan; ++c)
        if (e->chan_il[c] < 1 || e->chan_il[c] > 4 || e->chan_il[c] > e->L1)
            return set_err(FM_EINVAL, "edge: channel il out of range 1..min(4, L1)");
    for (int i = 0; i < e->nph1; ++i)
        if (e->lmax1[i] < 1 || e->lmax1[i] > e->L1) return set_err(FM_EINVAL, "edge: lmax1 out of range");
    if (e->ilinit < 1 || e->ilinit > 3) return set_err(FM_EUNSUPPORTED, "edge: only K, L and M edges (ilinit 1..3)");
    if (!e->lmax || !e->ck || !e->ckmxk || !e->rotc || !e->rotn) return set_err(FM_EINVAL, "edge: NULL table");
    return FM_OK;
}

}  // namespace

extern "C" {

const char *fm_last_error(void) { return g_err; }
const char *fm_version(void) { return "feffgpu_amd 0.1 (gfx950)"; }

int fm_cchi(const fm_edge *edge, int32_t precision, int32_t iorder, int32_t npath, int32_t nleg, const double *coords,
            const int64_t *ipots, void *cchi, double *reff, void *stream) {
    int rc = check_edge(edge);
    if (rc) return rc;
    if (iorder != 2) return set_err(FM_EUNSUPPORTED, "only iorder = 2 is implemented on the GPU");
    if (nleg < 2 || nleg > FM_MAX_LEGS) return set_err(FM_EINVAL, "nleg out of range 2..9");
    if (npath < 0) return set_err(FM_EINVAL, "npath < 0");
    if (npath == 0) return FM_OK;
    if (!coords || !ipots || !cchi) return set_err(FM_EINVAL, "NULL buffer");
    if (precision != 32 && precision != 64) return set_err(FM_EINVAL, "precision must be 32 or 64");
    const int set = set_for(*edge, nleg);
    if (set < 0) return set_err(FM_EUNSUPPORTED, "two-leg paths of L and M edges are not implemented on the GPU");
    KArgs a;
    memset(&a, 0, sizeof a);
    a.ed[0] = *edge;
    a.nedge = 1;
    a.mode = 0;
    a.npath_bucket = npath;
    a.first = 0;
    a.order = nullptr;
    a.nleg = nleg;
    a.coords = coords;
    a.ipots64 = ipots;
    a.cchi_out = cchi;
    a.reff_out = reff;
    hipStream_t st = (hipStream_t)stream;
    hipError_t e = precision == 32 ? dispatch<float>(a, set, st, nleg) : dispatch<double>(a, set, st, nleg);
    return e == hipSuccess ? FM_OK : hip_err(e, "fm_cchi launch");
}

int fm_chi_fine(const fm_edge *edge, int32_t precision, int32_t npath, const void *cchi, const double *reff,
                const fm_grid *grid, void *chi, void *stream) {
    int rc = check_edge(edge);
    if (rc) return rc;
    if (npath < 0) return set_err(FM_EINVAL, "npath < 0");
    if (npath == 0) return FM_OK;
    if (!cchi || !reff || !grid || !chi || !grid->q || !grid->jq || !grid->fq) return set_err(FM_EINVAL, "NULL buffer");
    if (precision != 32 && precision != 64) return set_err(FM_EINVAL, "precision must be 32 or 64");
    const size_t wb = align16((precision == 32 ? 4 : 8) * (size_t)edge->E) + align16(8 * (size_t)edge->E);
    const int wpb = WPB_MAX;
    const int blocks = (npath + wpb - 1) / wpb;
    hipStream_t st = (hipStream_t)stream;
    if (precision == 32)
        hipLaunchKernelGGL(k_chi_fine<float>, dim3(blocks), dim3(FM_WAVE * wpb), wb * wpb, st, *edge, npath,
                           (const float *)cchi, reff, *grid, (float *)chi, wb);
    else
        hipLaunchKernelGGL(k_chi_fine<double>, dim3(blocks), dim3(FM_WAVE * wpb), wb * wpb, st, *edge, npath,
                           (const double *)cchi, reff, *grid, (double *)chi, wb);
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? FM_OK : hip_err(e, "fm_chi_fine launch");
}

int fm_chi_paths(const fm_edge *edges, int32_t nedge, const fm_paths *paths, const fm_grid *grid,
                         int32_t precision, int32_t iorder, const int32_t *order, int32_t nbucket,
                         const int32_t *bucket_offset, const int32_t *bucket_set, const int32_t *bucket_nlegmax,
                         int32_t nseg, const int32_t *seg_ptr, const int32_t *seg_paths, double *chi_seg,
                         void *chi_path, double *reff, void *stream) {
    if (nedge < 1 || nedge > FM_MAX_EDGE) return set_err(FM_EINVAL, "nedge out of range 1..4");
    for (int i = 0; i < nedge; ++i) {
        int rc = check_edge(&edges[i]);
        if (rc) return rc;
    }
    if (iorder != 2) return set_err(FM_EUNSUPPORTED, "only iorder = 2 is implemented on the GPU");
    if (!paths || !grid || !chi_path) return set_err(FM_EINVAL, "NULL paths, grid or chi_path");
    if (precision != 32 && precision != 64) return set_err(FM_EINVAL, "precision must be 32 or 64");
    if (paths->nmax < 2 || paths->nmax > FM_MAX_LEGS) return set_err(FM_EINVAL, "nmax out of range 2..9");
    if (grid->nk < 1) return set_err(FM_EINVAL, "empty grid");
    hipStream_t st = (hipStream_t)stream;
    for (int b = 0; b < nbucket; ++b) {
        const int cnt = bucket_offset[b + 1] - bucket_offset[b];
        if (cnt <= 0) continue;
        const int set = bucket_set[b];
        if (set < 0 || set > 2) return set_err(FM_EUNSUPPORTED, "bucket kind not implemented on the GPU");
        KArgs a;
        memset(&a, 0, sizeof a);
        for (int i = 0; i < nedge; ++i) a.ed[i] = edges[i];
        a.nedge = nedge;
        a.mode = 1;
        a.npath_bucket = cnt;
        a.first = bucket_offset[b];
        a.order = order;
        a.P = *paths;
        a.G = *grid;
        a.chi_path = chi_path;
        a.reff_out = reff;
        int nlm = bucket_nlegmax ? bucket_nlegmax[b] : paths->nmax;
        if (nlm < 2 || nlm > paths->nmax) nlm = paths->nmax;
        hipError_t e = precision == 32 ? dispatch<float>(a, set, st, nlm) : dispatch<double>(a, set, st, nlm);
        if (e != hipSuccess) return hip_err(e, "fm_chi_paths launch");
    }
    if (chi_seg && nseg > 0) {
        if (!seg_ptr || !seg_paths) return set_err(FM_EINVAL, "NULL segment lists");
        dim3 grd((grid->nk + 127) / 128, nseg);
        if (precision == 32)
            hipLaunchKernelGGL(k_segment_sum<float>, grd, dim3(128), 0, st, nseg, grid->nk, seg_ptr, seg_paths, paths->deg,
                               (const float *)chi_path, chi_seg);
        else
            hipLaunchKernelGGL(k_segment_sum<double>, grd, dim3(128), 0, st, nseg, grid->nk, seg_ptr, seg_paths, paths->deg,
                               (const double *)chi_path, chi_seg);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return hip_err(e, "segment sum launch");
    }
    return FM_OK;
}

}  // extern "C"
