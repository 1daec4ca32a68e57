// Adjoint (reverse-mode) kernels of the path stage: d loss / d coordinates (GPUPathStage.cchi), d loss / d (cchi,
// reff) (chi_fine) and d loss / d positions for a path list (the fused entry), the gradient that feffgpu obtains by
// autograd through src/feffgpu/stage.py:147-388 and the gather of src/feffgpu/exafs/forward.py:153-163.
//
// One wavefront per path, one lane per phase-grid energy, nothing kept from the forward call:
//   1. leg geometry (FP64), rotation tables d^l(beta) and their beta-derivatives into LDS;
//   2. pass A per energy: the amplitude with the prefix products Pre_{s+1} = F_s ... F_1 stored in LDS;
//   3. the adjoint of the fine grid (mode 1) or the given dL/dcchi (mode 0): dL/dcchi per energy;
//   4. pass B per energy, the chain walked backwards with U_s = T F_{n-1} ... F_{s+1}: every scattering matrix is
//      streamed over the partial waves again with the forward-mode tangents of its two leg lengths (the Hankel
//      recursions differentiated) and of its rotation angle (the d recursion differentiated), contracted on the fly
//      with its adjoint conj(dL/dptrac) (Pre_s U_s)^T; eta enters analytically through exp(-i eta m);
//   5. per-leg gradients reduced over the lanes (fixed order), then pushed through the FP64 geometry with dual
//      numbers (one lane per coordinate), and for a path list gathered per atom in a fixed order.
#include <hip/hip_runtime.h>
#include <stdio.h>
#include <string.h>

#include <type_traits>

#include "fm_device.h"
#include "fm_path.h"

using namespace fm;

int fm_set_err(int code, const char *msg);

namespace {

__host__ __device__ constexpr int pre_elems(int set) { return lamdef(set).L * lamdef(set).L0; }   // one stored prefix product

struct BArgs {
    fm_edge ed[FM_MAX_EDGE];
    int nedge, npath_bucket, first;
    const int32_t *order;
    PathSrc src;
    const void *gcchi;          // mode 0: [npath][E] complex
    const double *greff;        // mode 0: [npath] or NULL
    fm_grid G;                  // mode 1
    const int32_t *segment;     // mode 1
    const double *gseg;         // mode 1: [nseg][nk]
    const void *gpath;          // mode 1: [npath][nk] or NULL
    double *gcoord;             // [npath][nslot][3]
    int nslot;
    int ltab, emax, nk, nlegmax, wpb;
    size_t wave_bytes, off_W, off_Wd, off_cc, off_amp, off_ang, off_cb, off_gq, off_gacc, off_gpar, off_tmp;
    unsigned char *work;        // per wave slot: the prefix products (global memory, cache-resident)
    size_t work_per_wave, work_bytes;
};

int set_err(int code, const char *msg) { return fm_set_err(code, msg); }
int hip_err(hipError_t e, const char *where) {
    char buf[256];
    snprintf(buf, sizeof buf, "%s: %s", where, hipGetErrorString(e));
    return fm_set_err(FM_EHIP, buf);
}

template <class T, int SET> __host__ void layout(BArgs &a) {
    size_t o = align16(sizeof(PathLDS));
    const size_t tab = align16(sizeof(T) * (size_t)a.nlegmax * a.ltab * FM_NREP);
    a.off_W = o; o += tab;
    a.off_Wd = o; o += tab;
    a.off_cc = o; o += align16(sizeof(cx<T>) * a.emax);
    a.off_amp = o; o += align16(sizeof(T) * a.emax);
    a.off_ang = o; o += align16(sizeof(double) * a.emax);
    a.off_cb = o; o += align16(sizeof(cx<T>) * a.emax);
    a.off_gq = o; o += align16(sizeof(T) * 2 * (size_t)a.nk);
    const int nslots = a.nlegmax > 2 ? a.nlegmax - 1 : 1;
    a.work_per_wave = (sizeof(cx<T>) * (size_t)nslots * pre_elems(SET) * FM_WAVE + 255) & ~size_t(255);
    a.off_gacc = o; o += align16(sizeof(T) * 3 * (size_t)a.nlegmax * FM_WAVE);
    a.off_gpar = o; o += align16(sizeof(double) * (3 * (size_t)a.nlegmax + 1));
    a.off_tmp = o; o += align16(sizeof(double) * 3 * FM_MAX_LEGS);
    a.wave_bytes = o;
}

// Derivative tables Wd[s][il][k] = d/dbeta d^{il-1}_{m1 m2}(beta_{s+1}) times the same normalisation as W: the
// three-term recursion differentiated, d'^{l} = (P c - Q) d'^{l-1} - P sin(beta) d^{l-1} - R d'^{l-2}.
template <class T, int SET>
__device__ void build_rotation_d(const fm_edge &ed, const PathLDS &g, int n, int ltab, T *Wd) {
    constexpr PairDef PD = pairdef(SET);
    const int lane = lane_id();
    for (int t = lane; t < n * FM_NREP; t += FM_WAVE) {
        const int s = t / FM_NREP, k = t % FM_NREP;
        int m1 = 0, m2 = 0;
#pragma unroll
        for (int q = 0; q < FM_NREP; ++q)
            if (q == k) { m1 = PD.m1[q]; m2 = PD.m2[q]; }
        const double be = g.beta[s + 1];
        const double c = cos(be), sn = sin(be), c2 = cos(0.5 * be), s2 = sin(0.5 * be);
        T *Ws = Wd + (size_t)s * ltab * FM_NREP;
        const int am1 = m1 < 0 ? -m1 : m1, am2 = m2 < 0 ? -m2 : m2;
        const int mm = (m1 + 3) * 7 + (m2 + 3), nn = am1 * 4 + am2;
        double dlm1 = 0.0, dl = 0.0, tlm1 = 0.0, tl = 0.0;
        for (int il = 1; il <= ed.L1; ++il) {
            const int l = il - 1;
            double d, t1;
            if (l <= 2) {
                d = dseed(l, m1, m2, c, sn, c2, s2);
                t1 = dseed_db(l, m1, m2, c, sn, c2, s2);
            } else {
                const double *co = ed.rotc + ((size_t)(l - 1) * 49 + mm) * 3;
                d = (co[0] * c - co[1]) * dl - co[2] * dlm1;
                t1 = (co[0] * c - co[1]) * tl - co[0] * sn * dl - co[2] * tlm1;
            }
            dlm1 = dl; dl = d;
            tlm1 = tl; tl = t1;
            Ws[(size_t)il * FM_NREP + k] = (T)(t1 * ed.rotn[(size_t)il * 16 + nn]);
        }
    }
    wave_sync();
}

template <class T> struct TLB;
template <> struct TLB<float> {
    static __device__ __forceinline__ const float *tl(const fm_edge &e) { return e.tl32; }
    static __device__ __forceinline__ const float *chw(const fm_edge &e) { return e.chw32; }
};
template <> struct TLB<double> {
    static __device__ __forceinline__ const double *tl(const fm_edge &e) { return e.tl64; }
    static __device__ __forceinline__ const double *chw(const fm_edge &e) { return e.chw64; }
};

// One partial-wave row of the adjoint of a block: adds the primal row to Ft and the contractions of the tangents
// with K to gX (gam leg length), gY (gamtl leg length), gB (rotation angle).
template <class T, int SET, int NR, int NC>
__device__ __forceinline__ void bwd_row(cx<T> (&Ft)[FM_NREP], const cx<T> (&K)[FM_NREP], const cx<T> (&xa)[3],
                                        const cx<T> (&xd)[3], cx<T> y0, cx<T> y1, cx<T> y0d, cx<T> y1d,
                                        const T *__restrict__ w, const T *__restrict__ wd, T &gX, T &gY, T &gB) {
    constexpr PairDef PD = pairdef(SET);
    cx<T> Z[3][2], Q[3][2], Qd[3][2];
#pragma unroll
    for (int aa = 0; aa < 3; ++aa) {
        Z[aa][0] = xa[aa] * y0;
        Z[aa][1] = xa[aa] * y1;
        Q[aa][0] = Q[aa][1] = Qd[aa][0] = Qd[aa][1] = mk<T>(T(0), T(0));
    }
#pragma unroll
    for (int k = 0; k < FM_NREP; ++k)
        if (PD.ri[k] < NR && PD.rj[k] < NC) {
            const int aa = PD.a[k] - 1, bb = PD.b[k] - 1;
            rfma(Ft[k], Z[aa][bb], w[k]);
            rfma(Q[aa][bb], K[k], w[k]);
            rfma(Qd[aa][bb], K[k], wd[k]);
        }
#pragma unroll
    for (int aa = 0; aa < 3; ++aa) {
        // beta: Re sum Qd Z; gam leg: Re sum X'_a (sum_b Q_ab Y_b); gamtl leg: Re sum Y'_b (sum_a Q_ab X_a)
        gB += Qd[aa][0].r * Z[aa][0].r - Qd[aa][0].i * Z[aa][0].i + Qd[aa][1].r * Z[aa][1].r - Qd[aa][1].i * Z[aa][1].i;
        const cx<T> v = Q[aa][0] * y0 + Q[aa][1] * y1;
        gX += xd[aa].r * v.r - xd[aa].i * v.i;
    }
    const cx<T> v0 = Q[0][0] * xa[0] + Q[1][0] * xa[1] + Q[2][0] * xa[2];
    const cx<T> v1 = Q[0][1] * xa[0] + Q[1][1] * xa[1] + Q[2][1] * xa[2];
    gY += y0d.r * v0.r - y0d.i * v0.i + y1d.r * v1.r - y1d.i * v1.i;
}

template <class T, int SET, int NR, int NC>
__device__ __forceinline__ void fblock_bwd(cx<T> (&Ft)[FM_NREP], const cx<T> (&K)[FM_NREP], cx<T> ck, T rX, T rY,
                                           int ilx, int ilmax, const T *__restrict__ Wl, const T *__restrict__ Wdl,
                                           const T *__restrict__ tl, int tlstride, T &gX, T &gY, T &gB) {
#pragma unroll
    for (int k = 0; k < FM_NREP; ++k) Ft[k] = mk<T>(T(0), T(0));
    gX = gY = gB = T(0);
    Hank3T<T> X;
    Hank2T<T> Y;
    X.init(minus_i_over(ck * rX), rX);
    Y.init(minus_i_over(ck * rY), rY);
    cx<T> xa[3], xd[3], ya[2], yd[2];
    const int rseed = ilmax < 4 ? ilmax : 4;
    for (int r = 1; r <= rseed; ++r) {
        X.seed_row(r, xa, xd);
        if (r < 4) Y.seed_row(r, ya, yd); else Y.step(r, ya, yd);
        const cx<T> t = mk<T>(tl[(size_t)r * tlstride], tl[(size_t)r * tlstride + 1]);
        if (r <= ilx)
            bwd_row<T, SET, NR, NC>(Ft, K, xa, xd, ya[0] * t, ya[1] * t, yd[0] * t, yd[1] * t, Wl + (size_t)r * FM_NREP,
                                    Wdl + (size_t)r * FM_NREP, gX, gY, gB);
    }
#pragma unroll 1
    for (int r = 5; r <= ilmax; ++r) {
        X.step(r, xa, xd);
        Y.step(r, ya, yd);
        const cx<T> t = mk<T>(tl[(size_t)r * tlstride], tl[(size_t)r * tlstride + 1]);
        if (r <= ilx)
            bwd_row<T, SET, NR, NC>(Ft, K, xa, xd, ya[0] * t, ya[1] * t, yd[0] * t, yd[1] * t, Wl + (size_t)r * FM_NREP,
                                    Wdl + (size_t)r * FM_NREP, gX, gY, gB);
    }
}

template <class T, int SET>
__device__ __forceinline__ void phases_of(cx<T> (&ph)[6], const PathLDS &g, int J) {
    constexpr LamDef D = lamdef(SET);
    const T c1 = (T)g.c1[J], s1 = (T)g.s1[J], c2 = (T)g.c2[J], s2 = (T)g.s2[J];
#pragma unroll
    for (int i = 0; i < 6; ++i) {
        const int m = D.M[i];
        ph[i] = m == 0 ? mk<T>(T(1), T(0)) : m == 1 ? mk<T>(c1, -s1) : m == -1 ? mk<T>(c1, s1) : m == 2 ? mk<T>(c2, -s2) : mk<T>(c2, s2);
    }
}

// Per-lane accumulators of the per-leg gradients (kind 0: r, 1: beta, 2: eta; leg j = 1..n) in LDS, lane-private.
template <class T> struct GAcc {
    T *base;
    int nlegmax;
    __device__ __forceinline__ void add(int kind, int j, T v) {
        base[((size_t)kind * nlegmax + (j - 1)) * FM_WAVE + lane_id()] += v;
    }
};

template <class T, int SET> struct PreStore {
    cx<T> *base;
    __device__ __forceinline__ cx<T> &at(int slot, int elem) {
        return base[((size_t)slot * pre_elems(SET) + elem) * FM_WAVE + lane_id()];
    }
};

// Pass A: the amplitude at energy e with the prefix products stored: slot s-2 holds Pre_s = F_{s-1} ... F_1
// (L x L0) for s = 2..n-1, slot max(n-2, 0) holds Pre_n (L0 x L0). Returns ptrac.
template <class T, int SET>
__device__ cx<T> pass_A(const fm_edge &ed, const PathLDS &g, const T *W, int ltab, int ee, int n, PreStore<T, SET> pre) {
    constexpr LamDef D = lamdef(SET);
    constexpr PairDef PD = pairdef(SET);
    constexpr int L = D.L, L0 = D.L0;
    const int E = ed.E;
    const cx<T> ck = mk<T>((T)ed.ck[2 * ee], (T)ed.ck[2 * ee + 1]);
    const int tlstride = 2 * E;
    const T *tlbase = TLB<T>::tl(ed);
    cx<T> Ft[FM_NREP], ph[6];
    auto scat = [&](int s, int &ilx, int &ilmax, const T *&tl, const T *&Wl) {
        const int ip = g.ipot[s];
        ilx = ed.lmax[(size_t)ip * E + ee] + 1;
        ilmax = ed.lmax1[ip];
        tl = tlbase + ((size_t)ip * ltab) * tlstride + 2 * ee;
        Wl = W + (size_t)(s - 1) * ltab * FM_NREP;
    };
    const int pn_slot = n > 2 ? n - 2 : 0;
    if (n == 2) {
        int ilx, ilmax;
        const T *tl, *Wl;
        scat(1, ilx, ilmax, tl, Wl);
        fblock<T, SET, L0, L0>(Ft, ck, (T)g.r[2], (T)g.r[1], ilx, ilmax, Wl, tl, tlstride);
        phases_of<T, SET>(ph, g, 2);
#pragma unroll
        for (int a = 0; a < L0; ++a)
#pragma unroll
            for (int b = 0; b < L0; ++b) pre.at(pn_slot, a * L0 + b) = ph[a] * Ft[PD.idx[a][b]] * T(PD.sg[a][b]);
    } else {
        {   // Pre_2 = F_1 (L x L0)
            int ilx, ilmax;
            const T *tl, *Wl;
            scat(1, ilx, ilmax, tl, Wl);
            fblock<T, SET, L, L0>(Ft, ck, (T)g.r[2], (T)g.r[1], ilx, ilmax, Wl, tl, tlstride);
            phases_of<T, SET>(ph, g, 2);
#pragma unroll
            for (int i = 0; i < L; ++i)
#pragma unroll
                for (int b = 0; b < L0; ++b) pre.at(0, i * L0 + b) = ph[i] * Ft[PD.idx[i][b]] * T(PD.sg[i][b]);
        }
#pragma unroll 1
        for (int s = 2; s <= n - 2; ++s) {   // Pre_{s+1} = F_s Pre_s
            int ilx, ilmax;
            const T *tl, *Wl;
            scat(s, ilx, ilmax, tl, Wl);
            fblock<T, SET, L, L>(Ft, ck, (T)g.r[s + 1], (T)g.r[s], ilx, ilmax, Wl, tl, tlstride);
            phases_of<T, SET>(ph, g, s + 1);
#pragma unroll
            for (int b = 0; b < L0; ++b) {
                cx<T> col[L];
#pragma unroll
                for (int j = 0; j < L; ++j) col[j] = pre.at(s - 2, j * L0 + b);
#pragma unroll
                for (int i = 0; i < L; ++i) {
                    cx<T> acc = mk<T>(T(0), T(0));
#pragma unroll
                    for (int j = 0; j < L; ++j) cfma(acc, Ft[PD.idx[i][j]] * T(PD.sg[i][j]), col[j]);
                    pre.at(s - 1, i * L0 + b) = ph[i] * acc;
                }
            }
        }
        {   // Pre_n = F_{n-1} Pre_{n-1} (L0 x L0)
            int ilx, ilmax;
            const T *tl, *Wl;
            scat(n - 1, ilx, ilmax, tl, Wl);
            fblock<T, SET, L0, L>(Ft, ck, (T)g.r[n], (T)g.r[n - 1], ilx, ilmax, Wl, tl, tlstride);
            phases_of<T, SET>(ph, g, n);
            cx<T> pn[L0][L0];
#pragma unroll
            for (int b = 0; b < L0; ++b) {
                cx<T> col[L];
#pragma unroll
                for (int j = 0; j < L; ++j) col[j] = pre.at(n - 3, j * L0 + b);
#pragma unroll
                for (int i = 0; i < L0; ++i) {
                    cx<T> acc = mk<T>(T(0), T(0));
#pragma unroll
                    for (int j = 0; j < L; ++j) cfma(acc, Ft[PD.idx[i][j]] * T(PD.sg[i][j]), col[j]);
                    pn[i][b] = ph[i] * acc;
                }
            }
#pragma unroll
            for (int i = 0; i < L0; ++i)
#pragma unroll
                for (int b = 0; b < L0; ++b) pre.at(pn_slot, i * L0 + b) = pn[i][b];
        }
    }
    // termination (primal) and the trace
    const auto lmaxA = [&](int j) -> int { return ed.lmax[(size_t)g.ipot[j] * E + ee]; };
    const int lx1 = max(lmaxA(0), lmaxA(1)) + 1, lxn = max(lmaxA(n - 1), lmaxA(n)) + 1;
    Hank3<T> X;
    Hank2<T> Y;
    X.init(minus_i_over(ck * (T)g.r[1]));
    Y.init(minus_i_over(ck * (T)g.r[n]));
    cx<T> Tt[FM_NREP];
#pragma unroll
    for (int k = 0; k < FM_NREP; ++k) Tt[k] = mk<T>(T(0), T(0));
    const T *Wn = W + (size_t)(n - 1) * ltab * FM_NREP;
    const T *chw = TLB<T>::chw(ed);
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
    phases_of<T, SET>(ph, g, 1);
    cx<T> ptrac = mk<T>(T(0), T(0));
#pragma unroll
    for (int a = 0; a < L0; ++a)
#pragma unroll
        for (int b = 0; b < L0; ++b) cfma(ptrac, ph[a] * Tt[PD.idx[a][b]] * T(PD.sg[a][b]), pre.at(pn_slot, b * L0 + a));
    return ptrac;
}

// cfac = exp(i R (ck - xk)) / prod (ck r_j), as the forward computes it
template <class T>
__device__ __forceinline__ cx<T> cfac_of(const fm_edge &ed, const PathLDS &g, int ee, int n) {
    const cx<T> ck = mk<T>((T)ed.ck[2 * ee], (T)ed.ck[2 * ee + 1]);
    double phs = g.R * ed.ckmxk[2 * ee];
    phs -= FM_TWO_PI * rint(phs / FM_TWO_PI);
    const double damp = exp(-g.R * ed.ckmxk[2 * ee + 1]);
    T sp, cp;
    if constexpr (sizeof(T) == 4) sincosf((float)phs, &sp, &cp); else sincos(phs, &sp, &cp);
    cx<T> ckn = ck;
    for (int j = 1; j < n; ++j) ckn = ckn * ck;
    return cdiv(mk<T>((T)damp * cp, (T)damp * sp), ckn * (T)g.prod_r);
}

// One step of pass B for scatterer s with an NR x NC block: H = conj(pbar) (Pre_s U_s)^T, its contraction with
// the tangents of F_s, eta_{s+1}, and U <- U F_s (unless last).
template <class T, int SET, int NR, int NC>
__device__ __forceinline__ void chain_step(const fm_edge &ed, const PathLDS &g, const T *W, const T *Wd, int ltab,
                                           int ee, int s, cx<T> pbc, cx<T> (&U)[lamdef(SET).L0][lamdef(SET).L],
                                           PreStore<T, SET> pre, bool has_pre, bool last, GAcc<T> acc) {
    constexpr LamDef D = lamdef(SET);
    constexpr PairDef PD = pairdef(SET);
    constexpr int L0 = D.L0;
    const int E = ed.E;
    const cx<T> ck = mk<T>((T)ed.ck[2 * ee], (T)ed.ck[2 * ee + 1]);
    cx<T> ph[6];
    phases_of<T, SET>(ph, g, s + 1);
    // K[rep] = sum over the orbit of H[l1][l2] ph[l1] sg, H[l1][l2] = pbc * M[l2][l1], M = Pre_s U (NC x NR)
    cx<T> K[FM_NREP];
#pragma unroll
    for (int k = 0; k < FM_NREP; ++k) K[k] = mk<T>(T(0), T(0));
#pragma unroll
    for (int l2 = 0; l2 < NC; ++l2) {
        cx<T> prow[L0];
#pragma unroll
        for (int a = 0; a < L0; ++a)
            prow[a] = has_pre ? pre.at(s - 2, l2 * L0 + a) : (a == l2 ? mk<T>(T(1), T(0)) : mk<T>(T(0), T(0)));
#pragma unroll
        for (int l1 = 0; l1 < NR; ++l1) {
            cx<T> m = mk<T>(T(0), T(0));
#pragma unroll
            for (int a = 0; a < L0; ++a) cfma(m, prow[a], U[a][l1]);
            const cx<T> h = pbc * m * ph[l1];
            if (PD.sg[l1][l2] > 0) K[PD.idx[l1][l2]] = K[PD.idx[l1][l2]] + h;
            else K[PD.idx[l1][l2]] = K[PD.idx[l1][l2]] - h;
        }
    }
    const int ip = g.ipot[s];
    const int ilx = ed.lmax[(size_t)ip * E + ee] + 1, ilmax = ed.lmax1[ip];
    const T *tl = TLB<T>::tl(ed) + ((size_t)ip * ltab) * (2 * E) + 2 * ee;
    cx<T> Ft[FM_NREP];
    T gX, gY, gB;
    fblock_bwd<T, SET, NR, NC>(Ft, K, ck, (T)g.r[s + 1], (T)g.r[s], ilx, ilmax, W + (size_t)(s - 1) * ltab * FM_NREP,
                               Wd + (size_t)(s - 1) * ltab * FM_NREP, tl, 2 * E, gX, gY, gB);
    acc.add(0, s + 1, gX);
    acc.add(0, s, gY);
    acc.add(1, s, gB);
    // eta_{s+1}: sum m_{l1} Im(H F) with F = ph sg Ft; recompute H from K's ingredients is costlier than this
    // second pass over (l1, l2) using M again
    T geta = T(0);
#pragma unroll
    for (int l2 = 0; l2 < NC; ++l2) {
        cx<T> prow[L0];
#pragma unroll
        for (int a = 0; a < L0; ++a)
            prow[a] = has_pre ? pre.at(s - 2, l2 * L0 + a) : (a == l2 ? mk<T>(T(1), T(0)) : mk<T>(T(0), T(0)));
#pragma unroll
        for (int l1 = 0; l1 < NR; ++l1) {
            if (D.M[l1] == 0) continue;
            cx<T> m = mk<T>(T(0), T(0));
#pragma unroll
            for (int a = 0; a < L0; ++a) cfma(m, prow[a], U[a][l1]);
            const cx<T> hf = pbc * m * ph[l1] * Ft[PD.idx[l1][l2]] * T(PD.sg[l1][l2]);
            geta += T(D.M[l1]) * hf.i;
        }
    }
    acc.add(2, s + 1, geta);
    if (!last) {
#pragma unroll
        for (int a = 0; a < L0; ++a) {
            cx<T> row[6];
#pragma unroll
            for (int i = 0; i < 6; ++i) row[i] = (i < NR) ? U[a][i] * ph[i] : mk<T>(T(0), T(0));
#pragma unroll
            for (int j = 0; j < D.L; ++j) {
                cx<T> v = mk<T>(T(0), T(0));
                if (j < NC) {
#pragma unroll
                    for (int i = 0; i < NR; ++i) {
                        const cx<T> f = Ft[PD.idx[i][j]];
                        if (PD.sg[i][j] > 0) cfma(v, row[i], f); else cfma(v, row[i], mk<T>(-f.r, -f.i));
                    }
                }
                U[a][j] = v;
            }
        }
    }
}

// Pass B at energy ee: cbar = dL/dcchi, cc = cchi.
template <class T, int SET>
__device__ void pass_B(const fm_edge &ed, const PathLDS &g, const T *W, const T *Wd, int ltab, int ee, int n,
                       cx<T> cbar, cx<T> cc, PreStore<T, SET> pre, GAcc<T> acc) {
    constexpr LamDef D = lamdef(SET);
    constexpr PairDef PD = pairdef(SET);
    constexpr int L = D.L, L0 = D.L0;
    const int E = ed.E;
    // cfac's own dependence on the leg lengths: d cfac / d r_j = cfac (i (ck - xk) - 1 / r_j)
    {
        const cx<T> w = conj(cbar) * cc;
        const cx<T> dk = mk<T>((T)ed.ckmxk[2 * ee], (T)ed.ckmxk[2 * ee + 1]);
        const cx<T> wdk = w * dk;
        const T a1 = -wdk.i, a2 = w.r;
        for (int j = 1; j <= n; ++j) acc.add(0, j, a1 - a2 / (T)g.r[j]);
    }
    const cx<T> cf = cfac_of<T>(ed, g, ee, n);
    const cx<T> pbc = conj(cbar * conj(cf));          // conj(dL/dptrac)
    // ---- termination: H_T[a][b] = pbc Pre_n[b][a]; tangents of leg 1 (X), leg n (Y), beta_n; eta_1
    const int pn_slot = n > 2 ? n - 2 : 0;
    const cx<T> ck = mk<T>((T)ed.ck[2 * ee], (T)ed.ck[2 * ee + 1]);
    cx<T> U[L0][L];
    {
        cx<T> ph[6];
        phases_of<T, SET>(ph, g, 1);
        cx<T> K[FM_NREP];
#pragma unroll
        for (int k = 0; k < FM_NREP; ++k) K[k] = mk<T>(T(0), T(0));
#pragma unroll
        for (int a = 0; a < L0; ++a)
#pragma unroll
            for (int b = 0; b < L0; ++b) {
                const cx<T> h = pbc * pre.at(pn_slot, b * L0 + a) * ph[a];
                if (PD.sg[a][b] > 0) K[PD.idx[a][b]] = K[PD.idx[a][b]] + h;
                else K[PD.idx[a][b]] = K[PD.idx[a][b]] - h;
            }
        const auto lmaxA = [&](int j) -> int { return ed.lmax[(size_t)g.ipot[j] * E + ee]; };
        const int lx1 = max(lmaxA(0), lmaxA(1)) + 1, lxn = max(lmaxA(n - 1), lmaxA(n)) + 1;
        Hank3T<T> X;
        Hank2T<T> Y;
        X.init(minus_i_over(ck * (T)g.r[1]), (T)g.r[1]);
        Y.init(minus_i_over(ck * (T)g.r[n]), (T)g.r[n]);
        cx<T> Tt[FM_NREP];
#pragma unroll
        for (int k = 0; k < FM_NREP; ++k) Tt[k] = mk<T>(T(0), T(0));
        const T *Wn = W + (size_t)(n - 1) * ltab * FM_NREP;
        const T *Wdn = Wd + (size_t)(n - 1) * ltab * FM_NREP;
        const T *chw = TLB<T>::chw(ed);
        T gX = T(0), gY = T(0), gB = T(0);
        cx<T> xa[3], xd[3], ya[2], yd[2];
        for (int r = 1; r <= 4; ++r) {
            X.seed_row(r, xa, xd, r <= lx1);
            if (r < 4) Y.seed_row(r, ya, yd, r <= lxn); else Y.step(r, ya, yd, r <= lxn);
            int c = -1;
            for (int q = 0; q < ed.nchan; ++q)
                if (ed.chan_il[q] == r) c = q;
            if (c >= 0) {
                const cx<T> w = mk<T>(chw[((size_t)c * E + ee) * 2], chw[((size_t)c * E + ee) * 2 + 1]) * T(2 * r - 1);
                bwd_row<T, SET, L0, L0>(Tt, K, xa, xd, ya[0] * w, ya[1] * w, yd[0] * w, yd[1] * w,
                                        Wn + (size_t)r * FM_NREP, Wdn + (size_t)r * FM_NREP, gX, gY, gB);
            }
        }
        acc.add(0, 1, gX);
        acc.add(0, n, gY);
        acc.add(1, n, gB);
        T geta = T(0);
#pragma unroll
        for (int a = 0; a < L0; ++a)
#pragma unroll
            for (int b = 0; b < L; ++b) {
                const cx<T> t = (b < L0) ? ph[a] * Tt[PD.idx[a][b]] * T(PD.sg[a][b]) : mk<T>(T(0), T(0));
                U[a][b] = t;
                if (b < L0 && D.M[a] != 0) {
                    const cx<T> hf = pbc * pre.at(pn_slot, b * L0 + a) * t;
                    geta += T(D.M[a]) * hf.i;
                }
            }
        acc.add(2, 1, geta);
    }
    // ---- the chain, s = n-1 .. 1
    if (n == 2) {
        chain_step<T, SET, L0, L0>(ed, g, W, Wd, ltab, ee, 1, pbc, U, pre, false, true, acc);
    } else {
        chain_step<T, SET, L0, L>(ed, g, W, Wd, ltab, ee, n - 1, pbc, U, pre, true, false, acc);
#pragma unroll 1
        for (int s = n - 2; s >= 2; --s) chain_step<T, SET, L, L>(ed, g, W, Wd, ltab, ee, s, pbc, U, pre, true, false, acc);
        chain_step<T, SET, L, L0>(ed, g, W, Wd, ltab, ee, 1, pbc, U, pre, false, true, acc);
    }
}

// deterministic wave sum (butterfly)
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, FM_WAVE);
    return v;
}

// Adjoint of the fine grid for the wave's path: from gq (dL/dA(q), dL/dPhi(q) already in LDS) to dL/dcchi per
// energy, gathered over the q points that interpolate from each energy (qorder / qstart runs, fixed order).
template <class T>
__device__ void fine_grid_gather(int E, const int32_t *qorder, const int32_t *qstart, const double *fq, const T *gA,
                                 const T *gP, const cx<T> *cc, const T *amp, cx<T> *cb) {
    const int lane = lane_id();
    for (int e = lane; e < E; e += FM_WAVE) {
        double ab = 0.0, pb = 0.0;
        for (int t = qstart[e]; t < qstart[e + 1]; ++t) {
            const int i = qorder[t];
            ab += (1.0 - fq[i]) * (double)gA[i];
            pb += (1.0 - fq[i]) * (double)gP[i];
        }
        if (e > 0)
            for (int t = qstart[e - 1]; t < qstart[e]; ++t) {
                const int i = qorder[t];
                ab += fq[i] * (double)gA[i];
                pb += fq[i] * (double)gP[i];
            }
        const double cr = (double)cc[e].r, ci = (double)cc[e].i;
        const double c2 = cr * cr + ci * ci, a = (double)amp[e];
        cx<T> v = mk<T>(T(0), T(0));
        if (c2 > 0.0) v = mk<T>((T)(ab * cr / a - pb * ci / c2), (T)(ab * ci / a + pb * cr / c2));
        cb[e] = v;
    }
}

// amp / unwrapped phase of cc[0..E) into amp[], ang[] (the forward epilogue's first half)
template <class T>
__device__ void amp_phase(int E, const cx<T> *cc, T *amp, double *ang) {
    const int lane = lane_id();
    for (int e = lane; e < E; e += FM_WAVE) {
        amp[e] = (T)sqrt((double)cc[e].r * cc[e].r + (double)cc[e].i * cc[e].i + 1e-30);
        ang[e] = (double)atan2(cc[e].i, cc[e].r);
    }
    wave_sync();
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
}

// dL/dA(q) and dL/dPhi(q) from dL/dchi(q) (chi = A sin(Phi + 2 q reff) dw); returns dL/dreff (A), wave-summed.
template <class T>
__device__ double fine_grid_adjoint_q(const int32_t *jq, const double *fq, const double *q, int nk, double reff_A,
                                      double sigma2, const T *amp, const double *ang, const double *gup_d,
                                      double gup_scale, const T *gup_p, T *gA, T *gP) {
    const int lane = lane_id();
    double part = 0.0;
    for (int i = lane; i < nk; i += FM_WAVE) {
        const int j = jq[i];
        const double f = fq[i];
        const double A = (1.0 - f) * (double)amp[j] + f * (double)amp[j + 1];
        double ph = (1.0 - f) * ang[j] + f * ang[j + 1] + 2.0 * q[i] * reff_A;
        ph -= FM_TWO_PI * rint(ph / FM_TWO_PI);
        double gu = 0.0;
        if (gup_d) gu += gup_scale * gup_d[i];
        if (gup_p) gu += (double)gup_p[i];
        if (sigma2 != 0.0) gu *= exp(-2.0 * sigma2 * q[i] * q[i]);
        T sn, cs;
        if constexpr (sizeof(T) == 4) sincosf((float)ph, &sn, &cs); else sincos(ph, &sn, &cs);
        gA[i] = (T)(gu * (double)sn);
        const T gp = (T)(gu * A * (double)cs);
        gP[i] = gp;
        part += 2.0 * q[i] * (double)gp;
    }
    return wave_sum(part);
}

template <class T, int SET>
__global__ __launch_bounds__(FM_WAVE *WPB_MAX) __attribute__((amdgpu_waves_per_eu(1))) void k_backward(BArgs a) {
    extern __shared__ __align__(16) unsigned char smem[];
    const int wid = threadIdx.x / FM_WAVE;
    const int lane = lane_id();
    unsigned char *base = smem + (size_t)wid * a.wave_bytes;
    PathLDS &g = *reinterpret_cast<PathLDS *>(base);
    T *W = reinterpret_cast<T *>(base + a.off_W);
    T *Wd = reinterpret_cast<T *>(base + a.off_Wd);
    cx<T> *cc = reinterpret_cast<cx<T> *>(base + a.off_cc);
    T *amp = reinterpret_cast<T *>(base + a.off_amp);
    double *ang = reinterpret_cast<double *>(base + a.off_ang);
    cx<T> *cb = reinterpret_cast<cx<T> *>(base + a.off_cb);
    T *gA = reinterpret_cast<T *>(base + a.off_gq);
    T *gP = gA + a.nk;
    PreStore<T, SET> pre{reinterpret_cast<cx<T> *>(a.work + (size_t)(blockIdx.x * a.wpb + wid) * a.work_per_wave)};
    GAcc<T> acc{reinterpret_cast<T *>(base + a.off_gacc), a.nlegmax};
    double *gpar = reinterpret_cast<double *>(base + a.off_gpar);
    double *tmp = reinterpret_cast<double *>(base + a.off_tmp);
    for (int slot = blockIdx.x * a.wpb + wid; slot < a.npath_bucket; slot += gridDim.x * a.wpb) {
        const int p = a.order ? a.order[a.first + slot] : a.first + slot;
        int ei;
        const int n = load_path(a.src, p, g, ei);
        const fm_edge &ed = a.ed[ei];
        const int ltab = ed.L1 + 1;
        const int E = ed.E;
        build_rotation<T, SET>(ed, g, n, ltab, W);
        build_rotation_d<T, SET>(ed, g, n, ltab, Wd);
        for (int k = 0; k < 3 * a.nlegmax; ++k) acc.base[(size_t)k * FM_WAVE + lane] = T(0);
        const bool single = E <= FM_WAVE;
        // 1-2: amplitude per energy (prefix products kept when one chunk covers every energy)
        for (int e0 = 0; e0 < E; e0 += FM_WAVE) {
            const int e = e0 + lane, ee = e < E ? e : E - 1;
            const cx<T> pt = pass_A<T, SET>(ed, g, W, ltab, ee, n, pre);
            const cx<T> c = pt * cfac_of<T>(ed, g, ee, n);
            if (e < E) cc[e] = c;
        }
        wave_sync();
        // 3: dL/dcchi
        double reffbar = 0.0;
        if (a.src.mode == 0) {
            const T *gin = reinterpret_cast<const T *>(a.gcchi) + (size_t)p * E * 2;
            for (int e = lane; e < E; e += FM_WAVE) cb[e] = mk<T>(gin[2 * e], gin[2 * e + 1]);
            reffbar = a.greff ? a.greff[p] : 0.0;
        } else {
            amp_phase<T>(E, cc, amp, ang);
            const double reff_A = g.reff * FM_BOHR;
            const double scale = a.src.P.deg[p];
            const double *gs = a.gseg ? a.gseg + (size_t)a.segment[p] * a.nk : nullptr;
            const T *gpp = a.gpath ? reinterpret_cast<const T *>(a.gpath) + (size_t)p * a.nk : nullptr;
            reffbar = fine_grid_adjoint_q<T>(a.G.jq + (size_t)ei * a.nk, a.G.fq + (size_t)ei * a.nk, a.G.q, a.nk, reff_A,
                                             a.src.P.sigma2 ? a.src.P.sigma2[p] : 0.0, amp, ang, gs, scale, gpp, gA, gP);
            wave_sync();
            fine_grid_gather<T>(E, a.G.qorder, a.G.qstart + (size_t)ei * a.G.qstride, a.G.fq + (size_t)ei * a.nk, gA, gP, cc,
                                amp, cb);
        }
        wave_sync();
        // 4: pass B (pass A again per chunk when the energies did not fit one chunk)
        for (int e0 = 0; e0 < E; e0 += FM_WAVE) {
            const int e = e0 + lane, ee = e < E ? e : E - 1;
            if (!single) pass_A<T, SET>(ed, g, W, ltab, ee, n, pre);
            cx<T> cbar = cb[ee], cce = cc[ee];
            if (e >= E) cbar = mk<T>(T(0), T(0));
            pass_B<T, SET>(ed, g, W, Wd, ltab, ee, n, cbar, cce, pre, acc);
        }
        wave_sync();
        // 5: per-leg gradients summed over the lanes (fixed order), plus reff = sum r / 2
        if (lane < 3 * n) {
            const int kind = lane / n, j = lane % n;
            double s = 0.0;
            const T *row = acc.base + ((size_t)kind * a.nlegmax + j) * FM_WAVE;
            for (int l = 0; l < FM_WAVE; ++l) s += (double)row[l];
            if (kind == 0) s += 0.5 * FM_BOHR * reffbar;
            gpar[lane] = s;
        }
        wave_sync();
        // geometry adjoint: one lane per coordinate direction (bohr), dual numbers through every leg
        const int ndir = a.src.mode == 0 ? 3 * n : 3 * (n - 1);
        double gout = 0.0;
        if (lane < ndir) {
            const int atom = a.src.mode == 0 ? lane / 3 : lane / 3 + 1, comp = lane % 3;
            Dd rat[FM_MAX_LEGS + 1][3];
            for (int j = 0; j <= n; ++j)
                for (int c = 0; c < 3; ++c) rat[j][c] = dd(g.rat[j][c], 0.0);
            rat[atom][comp].d = 1.0;
            if (atom == 0) rat[n][comp].d = 1.0;
            double acc_g = 0.0;
            for (int J = 1; J <= n; ++J) {
                const LegGeoD lg = leg_geometry_dual(rat, J, n);
                const int next = J == n ? 1 : J + 1;
                acc_g += gpar[J - 1] * lg.r.d + gpar[n + J - 1] * lg.beta.d + gpar[2 * n + J - 1] * lg.gamma.d +
                         gpar[2 * n + next - 1] * lg.alpha.d;
            }
            gout = acc_g / FM_BOHR;
            tmp[lane] = gout;
        }
        wave_sync();
        double *out = a.gcoord + (size_t)p * a.nslot * 3;
        if (a.src.mode == 0) {
            if (lane < 3 * n) out[lane] = gout;
        } else {
            if (lane < 3 * (n - 1)) out[3 + lane] = gout;
            if (lane < 3) {
                double s = 0.0;
                for (int j = 0; j < n - 1; ++j) s += tmp[3 * j + lane];
                out[lane] = -s;
            }
            if (lane >= 3 * n && lane < 3 * a.nslot) out[lane] = 0.0;
        }
        wave_sync();
    }
}

// grad_pos[atom] = sum of gcoord over the atom's references, in list order
__global__ void k_atom_grad(int natom, const int32_t *ptr, const int32_t *refs, const double *gcoord, double *grad) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= 3 * natom) return;
    const int atom = t / 3, c = t % 3;
    double s = 0.0;
    for (int k = ptr[atom]; k < ptr[atom + 1]; ++k) s += gcoord[(size_t)refs[k] * 3 + c];
    grad[t] = s;
}

// GPUPathStage.chi_fine's adjoint on its own: one wave per path.
template <class T>
__global__ __launch_bounds__(FM_WAVE *WPB_MAX) void k_chi_fine_bwd(fm_edge ed, int npath, const T *cchi, const double *reff,
                                                                    fm_grid G, const T *gchi, T *gcchi, double *greff,
                                                                    size_t wave_bytes) {
    extern __shared__ __align__(16) unsigned char smem[];
    const int wid = threadIdx.x / FM_WAVE, lane = lane_id();
    const int wpb = blockDim.x / FM_WAVE;
    const int E = ed.E, nk = G.nk;
    unsigned char *base = smem + (size_t)wid * wave_bytes;
    cx<T> *cc = reinterpret_cast<cx<T> *>(base);
    T *amp = reinterpret_cast<T *>(base + align16(sizeof(cx<T>) * E));
    double *ang = reinterpret_cast<double *>(base + align16(sizeof(cx<T>) * E) + align16(sizeof(T) * E));
    cx<T> *cb = reinterpret_cast<cx<T> *>(reinterpret_cast<unsigned char *>(ang) + align16(sizeof(double) * E));
    T *gA = reinterpret_cast<T *>(reinterpret_cast<unsigned char *>(cb) + align16(sizeof(cx<T>) * E));
    T *gP = gA + nk;
    for (int p = blockIdx.x * wpb + wid; p < npath; p += gridDim.x * wpb) {
        for (int e = lane; e < E; e += FM_WAVE) cc[e] = mk<T>(cchi[((size_t)p * E + e) * 2], cchi[((size_t)p * E + e) * 2 + 1]);
        wave_sync();
        amp_phase<T>(E, cc, amp, ang);
        const double rb = fine_grid_adjoint_q<T>(G.jq, G.fq, G.q, nk, reff[p], 0.0, amp, ang, nullptr, 0.0,
                                                 gchi + (size_t)p * nk, gA, gP);
        wave_sync();
        fine_grid_gather<T>(E, G.qorder, G.qstart, G.fq, gA, gP, cc, amp, cb);
        wave_sync();
        for (int e = lane; e < E; e += FM_WAVE) {
            gcchi[((size_t)p * E + e) * 2] = cb[e].r;
            gcchi[((size_t)p * E + e) * 2 + 1] = cb[e].i;
        }
        if (lane == 0) greff[p] = rb;
        wave_sync();
    }
}

template <class T, int SET> hipError_t launch_backward(BArgs &a, hipStream_t st) {
    layout<T, SET>(a);
    if (a.wave_bytes > 160 * 1024) return hipErrorInvalidValue;
    int wpb = (int)((64 * 1024) / a.wave_bytes);
    wpb = wpb < 1 ? 1 : (wpb > WPB_MAX ? WPB_MAX : wpb);
    a.wpb = wpb;
    const size_t shm = a.wave_bytes * wpb;
    if (shm > 64 * 1024) {
        hipError_t e = hipFuncSetAttribute((const void *)k_backward<T, SET>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)shm);
        if (e != hipSuccess) return e;
    }
    const long slots = (long)(a.work_bytes / a.work_per_wave);
    if (slots < wpb) return hipErrorInvalidValue;
    long blocks = (a.npath_bucket + wpb - 1) / wpb;
    if (blocks > slots / wpb) blocks = slots / wpb;
    if (blocks == 0) return hipSuccess;
    hipLaunchKernelGGL((k_backward<T, SET>), dim3((unsigned)blocks), dim3(FM_WAVE * wpb), shm, st, a);
    return hipGetLastError();
}

template <class T> hipError_t dispatch_b(BArgs &a, int set, hipStream_t st) {
    switch (set) {
        case 0: return launch_backward<T, 0>(a, st);
        case 1: return launch_backward<T, 1>(a, st);
        case 2: return launch_backward<T, 2>(a, st);
    }
    return hipErrorInvalidValue;
}

int check_edge_b(const fm_edge *e) {
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

int set_of(const fm_edge &ed, int nleg) {
    if (ed.ilinit <= 1) return nleg == 2 ? 2 : 0;
    return nleg == 2 ? -1 : 1;
}

}  // namespace

extern "C" {

size_t fm_backward_workspace_bytes(int32_t precision, int32_t nlegmax, int32_t wave_slots) {
    if (nlegmax < 2 || nlegmax > FM_MAX_LEGS || wave_slots < 1) return 0;
    const size_t es = precision == 32 ? 8 : 16;
    const size_t nslots = nlegmax > 2 ? nlegmax - 1 : 1;
    const size_t per = (es * nslots * 36 * FM_WAVE + 255) & ~size_t(255);
    return per * (size_t)wave_slots;
}

int fm_cchi_backward(const fm_edge *edge, int32_t precision, int32_t iorder, int32_t npath, int32_t nleg,
                     const double *coords, const int64_t *ipots, const void *gcchi, const double *greff, double *gcoords,
                     void *workspace, size_t workspace_bytes, void *stream) {
    int rc = check_edge_b(edge);
    if (rc) return rc;
    if (iorder != 2) return set_err(FM_EUNSUPPORTED, "only iorder = 2 is implemented on the GPU");
    if (nleg < 2 || nleg > FM_MAX_LEGS) return set_err(FM_EINVAL, "nleg out of range 2..9");
    if (npath < 0) return set_err(FM_EINVAL, "npath < 0");
    if (npath == 0) return FM_OK;
    if (!coords || !ipots || !gcchi || !gcoords || !workspace) return set_err(FM_EINVAL, "NULL buffer");
    if (precision != 32 && precision != 64) return set_err(FM_EINVAL, "precision must be 32 or 64");
    if (workspace_bytes < fm_backward_workspace_bytes(precision, nleg, WPB_MAX))
        return set_err(FM_EINVAL, "workspace smaller than fm_backward_workspace_bytes(precision, nleg, 4)");
    const int set = set_of(*edge, nleg);
    if (set < 0) return set_err(FM_EUNSUPPORTED, "two-leg paths of L and M edges are not implemented on the GPU");
    BArgs a;
    memset(&a, 0, sizeof a);
    a.ed[0] = *edge;
    a.nedge = 1;
    a.npath_bucket = npath;
    a.src.mode = 0;
    a.src.nleg = nleg;
    a.src.coords = coords;
    a.src.ipots64 = ipots;
    a.gcchi = gcchi;
    a.greff = greff;
    a.gcoord = gcoords;
    a.nslot = nleg;
    a.ltab = edge->L1 + 1;
    a.emax = edge->E;
    a.nk = 0;
    a.nlegmax = nleg;
    a.work = (unsigned char *)workspace;
    a.work_bytes = workspace_bytes;
    hipStream_t st = (hipStream_t)stream;
    hipError_t e = precision == 32 ? dispatch_b<float>(a, set, st) : dispatch_b<double>(a, set, st);
    return e == hipSuccess ? FM_OK : hip_err(e, "fm_cchi_backward launch");
}

int fm_chi_fine_backward(const fm_edge *edge, int32_t precision, int32_t npath, const void *cchi, const double *reff,
                         const fm_grid *grid, const void *gchi, void *gcchi, double *greff, void *stream) {
    int rc = check_edge_b(edge);
    if (rc) return rc;
    if (npath < 0) return set_err(FM_EINVAL, "npath < 0");
    if (npath == 0) return FM_OK;
    if (!cchi || !reff || !grid || !gchi || !gcchi || !greff || !grid->q || !grid->jq || !grid->fq || !grid->qorder ||
        !grid->qstart)
        return set_err(FM_EINVAL, "NULL buffer");
    if (precision != 32 && precision != 64) return set_err(FM_EINVAL, "precision must be 32 or 64");
    const int E = edge->E, nk = grid->nk;
    const size_t es = precision == 32 ? 4 : 8;
    const size_t wb = align16(2 * es * E) + align16(es * E) + align16(8 * (size_t)E) + align16(2 * es * E) + align16(2 * es * nk);
    if (wb > 64 * 1024) return set_err(FM_EINVAL, "grid too large for one workgroup's LDS");
    const int wpb = (int)((64 * 1024) / wb) < WPB_MAX ? (int)((64 * 1024) / wb) : WPB_MAX;
    const int blocks = (npath + wpb - 1) / wpb;
    hipStream_t st = (hipStream_t)stream;
    if (precision == 32)
        hipLaunchKernelGGL(k_chi_fine_bwd<float>, dim3(blocks), dim3(FM_WAVE * wpb), wb * wpb, st, *edge, npath,
                           (const float *)cchi, reff, *grid, (const float *)gchi, (float *)gcchi, greff, wb);
    else
        hipLaunchKernelGGL(k_chi_fine_bwd<double>, dim3(blocks), dim3(FM_WAVE * wpb), wb * wpb, st, *edge, npath,
                           (const double *)cchi, reff, *grid, (const double *)gchi, (double *)gcchi, greff, wb);
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? FM_OK : hip_err(e, "fm_chi_fine_backward launch");
}

int fm_chi_paths_backward(const fm_edge *edges, int32_t nedge, const fm_paths *paths, const fm_grid *grid,
                          int32_t precision, int32_t iorder, const int32_t *order, int32_t nbucket,
                          const int32_t *bucket_offset, const int32_t *bucket_kind, const int32_t *bucket_edge,
                          const int32_t *bucket_nlegmax, const int32_t *segment, const double *gseg, const void *gpath, int32_t natom,
                          const int32_t *atom_ptr, const int32_t *atom_refs, double *gcoord, double *grad_pos,
                          void *workspace, size_t workspace_bytes, void *stream) {
    if (nedge < 1 || nedge > FM_MAX_EDGE) return set_err(FM_EINVAL, "nedge out of range 1..4");
    int emax = 0, ltab = 0;
    for (int i = 0; i < nedge; ++i) {
        int rc = check_edge_b(&edges[i]);
        if (rc) return rc;
        emax = emax > edges[i].E ? emax : edges[i].E;
        ltab = ltab > edges[i].L1 + 1 ? ltab : edges[i].L1 + 1;
    }
    if (iorder != 2) return set_err(FM_EUNSUPPORTED, "only iorder = 2 is implemented on the GPU");
    if (!paths || !grid || !gcoord || !grid->qorder || !grid->qstart) return set_err(FM_EINVAL, "NULL paths, grid or gcoord");
    if (!gseg && !gpath) return set_err(FM_EINVAL, "no upstream gradient");
    if (gseg && !segment) return set_err(FM_EINVAL, "gseg needs segment");
    if (precision != 32 && precision != 64) return set_err(FM_EINVAL, "precision must be 32 or 64");
    if (paths->nmax < 2 || paths->nmax > FM_MAX_LEGS) return set_err(FM_EINVAL, "nmax out of range 2..9");
    if (grid->qstride < emax + 1) return set_err(FM_EINVAL, "grid.qstride < E + 1");
    if (!workspace || workspace_bytes < fm_backward_workspace_bytes(precision, paths->nmax, WPB_MAX))
        return set_err(FM_EINVAL, "workspace smaller than fm_backward_workspace_bytes(precision, nmax, 4)");
    hipStream_t st = (hipStream_t)stream;
    for (int b = 0; b < nbucket; ++b) {
        const int cnt = bucket_offset[b + 1] - bucket_offset[b];
        if (cnt <= 0) continue;
        const int set = bucket_kind[b];
        if (set < 0 || set > 2) return set_err(FM_EUNSUPPORTED, "bucket kind not implemented on the GPU");
        if (bucket_edge && (bucket_edge[b] < 0 || bucket_edge[b] >= nedge)) return set_err(FM_EINVAL, "bucket edge out of range");
        BArgs a;
        memset(&a, 0, sizeof a);
        for (int i = 0; i < nedge; ++i) a.ed[i] = edges[i];
        a.nedge = nedge;
        a.npath_bucket = cnt;
        a.first = bucket_offset[b];
        a.order = order;
        a.src.mode = 1;
        a.src.P = *paths;
        a.G = *grid;
        a.segment = segment;
        a.gseg = gseg;
        a.gpath = gpath;
        a.gcoord = gcoord;
        a.nslot = paths->nmax;
        a.ltab = ltab;
        a.emax = emax;
        a.nk = grid->nk;
        int nlm = bucket_nlegmax ? bucket_nlegmax[b] : paths->nmax;
        a.nlegmax = (nlm < 2 || nlm > paths->nmax) ? paths->nmax : nlm;
        a.work = (unsigned char *)workspace;
        a.work_bytes = workspace_bytes;
        hipError_t e = precision == 32 ? dispatch_b<float>(a, set, st) : dispatch_b<double>(a, set, st);
        if (e != hipSuccess) return hip_err(e, "fm_chi_paths_backward launch");
    }
    if (grad_pos && natom > 0) {
        if (!atom_ptr || !atom_refs) return set_err(FM_EINVAL, "NULL atom lists");
        const int nt = 3 * natom;
        hipLaunchKernelGGL(k_atom_grad, dim3((nt + 255) / 256), dim3(256), 0, st, natom, atom_ptr, atom_refs, gcoord, grad_pos);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return hip_err(e, "atom gradient launch");
    }
    return FM_OK;
}

}  // extern "C"
