// Single-scattering (two-leg) paths of every edge: the path absorber -> scatterer -> absorber is collinear, so both
// scattering angles are pi and every rotation matrix is anti-diagonal, d^l_{m1 m2}(pi) = (-1)^(l + m1) delta(m1, -m2)
// (Edmonds' convention, as FEFF's rot3i and the stage's closed forms, src/feffgpu/stage.py:41-76); with the
// normalisation xnlm(il,|m1|)(-1)^|m1| / xnlm(il,|m2|) of gam / gamtl (src/feffgpu/stage.py:267-285) every weight is
// (-1)^(il-1) times the static mask. The lambda set is the nsc = 1 set of setlam (src/feffgpu/genfmt.py:115-153):
// 6 entries for K / L1 edges, 15 for L2 / L3, the first 15 of 22 for M4 / M5 (the LAMTOT truncation). Only the pairs
// (lambda1, lambda2) with m2 = -m1 are non-zero; of each pair and its mirror (-m1, -m2) one is computed (the two are
// equal before the row phase). Both legs have the same length, so one spherical-wave table serves gam and gamtl of
// both the scattering matrix and the termination.
//   cchi = tr(T F) cfac, T[a][b] = e^{-i eta_1 m_a} T_k, F[b][a] = e^{-i eta_2 m_b} F_k' (src/feffgpu/stage.py:287-338).
// The adjoint differentiates the same evaluation with dual numbers in the common leg length (the angle is pi for every
// position, so it contributes nothing).
#pragma once
#include "fm_device.h"

namespace fm {

constexpr int FM_SET_SS = 3;   // kernel sets FM_SET_SS + ilinit - 1: single scattering of K/L1, L2/L3, M4/M5 edges
constexpr int SS_MAXL = 15;    // lambda entries
constexpr int SS_MAXR = 27;    // representative anti-diagonal pairs
constexpr int SS_MAXP = 45;    // anti-diagonal pairs
constexpr int SS_COLS = 5;     // spherical-wave columns |m| + n + 1 read by gam (3 for K / L1 edges)
__host__ __device__ constexpr int ss_cols(int ilinit) { return ilinit == 1 ? 3 : SS_COLS; }

struct SSDef {
    int L;
    int M[SS_MAXL], N[SS_MAXL];
    int nrep;
    int ra[SS_MAXR], rb[SS_MAXR];      // lambda indices (a, b) of representative k
    int ca[SS_MAXR], cb[SS_MAXR];      // spherical-wave columns of gam (|m| + n_a + 1) and gamtl (n_b + 1)
    int ilmin[SS_MAXR];                // static mask: il >= max(|m| + n_a + 1, n_b + 1)
    int am[SS_MAXR];                   // |m_a|
    int npair;
    int pa[SS_MAXP], pb[SS_MAXP];      // every anti-diagonal (a, b)
    int kab[SS_MAXP], kba[SS_MAXP];    // representatives of (a, b) and (b, a)
};

__host__ __device__ constexpr SSDef ssdef(int ilinit) {
    SSDef d{};
    const int mmax = ilinit, nmax = ilinit, iord = 2 * nmax + mmax;
    int L = 0;
    for (int n = 0; n <= nmax; ++n)
        for (int m = 0; m <= mmax; ++m) {
            if (2 * n + m > iord) continue;
            const int cnt = m == 0 ? 1 : 2;
            for (int t = 0; t < cnt; ++t) {
                if (L < SS_MAXL) {
                    d.M[L] = m == 0 ? 0 : (t == 0 ? -m : m);
                    d.N[L] = n;
                    ++L;
                }
            }
        }
    d.L = L;
    int k = 0;
    for (int a = 0; a < L; ++a)
        for (int b = 0; b < L; ++b) {
            if (d.M[b] != -d.M[a] || d.M[a] < 0) continue;
            const int am = d.M[a] < 0 ? -d.M[a] : d.M[a];
            d.ra[k] = a;
            d.rb[k] = b;
            d.ca[k] = am + d.N[a] + 1;
            d.cb[k] = d.N[b] + 1;
            d.ilmin[k] = d.ca[k] > d.cb[k] ? d.ca[k] : d.cb[k];
            d.am[k] = am;
            ++k;
        }
    d.nrep = k;
    auto rep_of = [&](int a, int b) {
        int ma = d.M[a], na = d.N[a], nb = d.N[b];
        if (ma < 0) ma = -ma;   // the mirror pair has the same value before the row phase
        for (int q = 0; q < d.nrep; ++q)
            if (d.M[d.ra[q]] == ma && d.N[d.ra[q]] == na && d.N[d.rb[q]] == nb) return q;
        return -1;
    };
    int p = 0;
    for (int a = 0; a < L; ++a)
        for (int b = 0; b < L; ++b) {
            if (d.M[b] != -d.M[a]) continue;
            d.pa[p] = a;
            d.pb[p] = b;
            d.kab[p] = rep_of(a, b);
            d.kba[p] = rep_of(b, a);
            ++p;
        }
    d.npair = p;
    return d;
}

// ------------------------------------------------------------------------------------------------ dual complex
// value and derivative along one parameter (the common leg length)
template <class T> struct dcx { cx<T> v, d; };
template <class T> __device__ __forceinline__ dcx<T> operator+(dcx<T> a, dcx<T> b) { return {a.v + b.v, a.d + b.d}; }
template <class T> __device__ __forceinline__ dcx<T> operator-(dcx<T> a, dcx<T> b) { return {a.v - b.v, a.d - b.d}; }
template <class T> __device__ __forceinline__ dcx<T> operator*(dcx<T> a, dcx<T> b) { return {a.v * b.v, a.d * b.v + a.v * b.d}; }
template <class T> __device__ __forceinline__ dcx<T> operator*(dcx<T> a, T s) { return {a.v * s, a.d * s}; }
template <class T> __device__ __forceinline__ dcx<T> operator*(dcx<T> a, cx<T> s) { return {a.v * s, a.d * s}; }

// acc += a * b: complex values as the product, then the sum (the forward); dual values, the single-scattering
// adjoint's, as four fused multiply-adds per complex product (cfma4)
template <class T> __device__ __forceinline__ void vfma(cx<T> &acc, cx<T> a, cx<T> b) { acc = acc + a * b; }
template <class T> __device__ __forceinline__ void vfma(dcx<T> &acc, dcx<T> a, dcx<T> b) {
    cfma4(acc.v, a.v, b.v);
    cfma4(acc.d, a.d, b.v);
    cfma4(acc.d, a.v, b.d);
}

template <class V> struct VOps;
template <class T> struct VOps<cx<T>> {
    using R = T;
    static __device__ __forceinline__ cx<T> zero() { return mk<T>(T(0), T(0)); }
    static __device__ __forceinline__ cx<T> one() { return mk<T>(T(1), T(0)); }
    static __device__ __forceinline__ cx<T> val(cx<T> x) { return x; }
};
template <class T> struct VOps<dcx<T>> {
    using R = T;
    static __device__ __forceinline__ dcx<T> zero() { return {mk<T>(T(0), T(0)), mk<T>(T(0), T(0))}; }
    static __device__ __forceinline__ dcx<T> one() { return {mk<T>(T(1), T(0)), mk<T>(T(0), T(0))}; }
    static __device__ __forceinline__ cx<T> val(dcx<T> x) { return x.v; }
};

// ------------------------------------------------------------------------------------------------ spherical waves
// The c[il][im] table of one leg (FEFF's sclmz as src/feffgpu/stage.py:209-232 vectorises it), streamed row by row
// for columns 1..C: row il from rows il-1 and il-2; rows beyond the leg's limit (alive false) are zero.
template <class V, int C> struct HankRows {
    using T = typename VOps<V>::R;
    V z, p1[C], p2[C], cmm[C];
    __device__ __forceinline__ void init(V z_) {
        z = z_;
        cmm[0] = VOps<V>::one();
#pragma unroll
        for (int c = 1; c < C; ++c) cmm[c] = (cmm[c - 1] * z) * T(-(2 * c - 1));   // column im = c + 1
#pragma unroll
        for (int c = 0; c < C; ++c) p1[c] = p2[c] = VOps<V>::zero();
    }
    // row r >= 1 into o[0..C-1] (column im = c + 1)
    __device__ __forceinline__ void row(int r, V (&o)[C], bool alive) {
#pragma unroll
        for (int c = 0; c < C; ++c) {
            const int im = c + 1;
            V v;
            if (c == 0) {
                if (r == 1) v = VOps<V>::one();
                else if (r == 2) v = VOps<V>::one() - z;
                else v = (p2[0] - (z * p1[0]) * T(2 * r - 3)) * T(alive ? 1 : 0);
            } else {
                if (r < im) v = VOps<V>::zero();
                else if (r == im) v = cmm[c] * T(alive ? 1 : 0);
                else if (r == im + 1) v = (cmm[c] * (VOps<V>::one() - z * T(im))) * T((2 * im - 1) * (alive ? 1 : 0));
                else v = (p2[c] - (z * (p1[c] + p1[c - 1])) * T(2 * r - 3)) * T(alive ? 1 : 0);
            }
            o[c] = v;
        }
#pragma unroll
        for (int c = 0; c < C; ++c) { p2[c] = p1[c]; p1[c] = o[c]; }
    }
    // row r > C + 1 (past every column's explicit rows) of a live row: the recursion alone, the same arithmetic as
    // row()
    __device__ __forceinline__ void row_rec(int r, V (&o)[C]) {
#pragma unroll
        for (int c = 0; c < C; ++c) {
            if (c == 0) o[c] = p2[0] - (z * p1[0]) * T(2 * r - 3);
            else o[c] = p2[c] - (z * (p1[c] + p1[c - 1])) * T(2 * r - 3);
        }
#pragma unroll
        for (int c = 0; c < C; ++c) { p2[c] = p1[c]; p1[c] = o[c]; }
    }
};

// The representatives of T and F of a single-scattering path at energy ee (V = cx<T> or dcx<T> for the derivative
// along the leg length r, dz/dr = -z / r). lxa: rows alive on the legs (both legs join the same two atoms).
template <class T, int ILI, class V>
__device__ __forceinline__ void ss_reps(const fm_edge &ed, int ee, int ip_s, V z, int lxa, V (&Tk)[SS_MAXR],
                                        V (&Fk)[SS_MAXR]) {
    constexpr SSDef D = ssdef(ILI);
    constexpr int C = ss_cols(ILI);
    const int E = ed.E;
    const T *tlbase = sizeof(T) == 4 ? (const T *)(const void *)ed.tl32 : (const T *)(const void *)ed.tl64;
    const T *chw = sizeof(T) == 4 ? (const T *)(const void *)ed.chw32 : (const T *)(const void *)ed.chw64;
    const int ltab = ed.L1 + 1;
    const int ilx = ed.lmax[(size_t)ip_s * E + ee] + 1;     // the scatterer's partial waves in F
#pragma unroll
    for (int k = 0; k < D.nrep; ++k) Tk[k] = Fk[k] = VOps<V>::zero();
    HankRows<V, C> H;
    H.init(z);
    const T *tl = tlbase + ((size_t)ip_s * ltab) * (2 * E) + 2 * ee;
    // the t-matrix two rows ahead (rows past ilx read row ilx, inside the table, and are not used), and the
    // termination's channel weights of rows 1..4 read with the first rows (a row without a channel reads channel 0's
    // and does not use it)
    auto trow = [&](int r) {
        const size_t o = (size_t)(r < ilx ? r : ilx) * 2 * E;
        return mk<T>(tl[o], tl[o + 1]);
    };
    int chan[4];
    cx<T> wraw[4];
#pragma unroll
    for (int r = 1; r <= 4; ++r) {
        int c = -1;
        for (int q = 0; q < ed.nchan; ++q)
            if (ed.chan_il[q] == r) c = q;
        chan[r - 1] = c;
        const size_t o = ((size_t)(c >= 0 ? c : 0) * E + ee) * 2;
        wraw[r - 1] = mk<T>(chw[o], chw[o + 1]);
    }
    cx<T> t = trow(1), tn = trow(2);
    // rows 1 .. RP (the columns' explicit rows and the termination's) unrolled, so that their cases resolve at
    // compile time; the rest (r <= ilx) by the recursion alone, two rows per iteration
    constexpr int RP = C + 1 > 4 ? C + 1 : 4;
    auto scatter_row = [&](int r, const V (&h)[C], cx<T> traw) {
        const T sgn = (r & 1) ? T(1) : T(-1);          // (-1)^(il - 1), exact on the t-matrix factor
        const cx<T> tr = mk<T>(traw.r * sgn, traw.i * sgn);
        V yt[3];
#pragma unroll
        for (int c = 0; c < 3; ++c) yt[c] = h[c] * tr;
#pragma unroll
        for (int k = 0; k < D.nrep; ++k)
            if (r >= D.ilmin[k]) vfma(Fk[k], h[D.ca[k] - 1], yt[D.cb[k] - 1]);
    };
    // rows past RP up to ilx: ilx <= lxa (lxa is the larger lmax + 1 of the two atoms, one of them the
    // scatterer), so every such row is alive
    auto rec_row = [&](int r) {
        V h[C];
        H.row_rec(r, h);
        const cx<T> tnn = trow(r + 2);
        scatter_row(r, h, t);
        t = tn;
        tn = tnn;
    };
#pragma unroll
    for (int r = 1; r <= RP; ++r) {
        V h[C];
        H.row(r, h, r <= lxa);
        const T sgn = (r & 1) ? T(1) : T(-1);          // (-1)^(il - 1)
        const cx<T> tnn = trow(r + 2);
        if (r <= ilx) scatter_row(r, h, t);
        t = tn;
        tn = tnn;
        if (r <= 4) {
            if (chan[r - 1] >= 0) {
                const cx<T> w = wraw[r - 1] * (T(2 * r - 1) * sgn);
                V yw[3];
#pragma unroll
                for (int c = 0; c < 3; ++c) yw[c] = h[c] * w;
#pragma unroll
                for (int k = 0; k < D.nrep; ++k)
                    if (r >= D.ilmin[k] && D.am[k] + 1 <= r) Tk[k] = Tk[k] + h[D.ca[k] - 1] * yw[D.cb[k] - 1];
            }
        }
    }
    int r = RP + 1;
#pragma unroll 1
    for (; r < ilx; r += 2) {
        rec_row(r);
        rec_row(r + 1);
    }
    if (r == ilx) rec_row(r);
}

// e^{-i eta m} for m = -3..3 of leg J
template <class T> __device__ __forceinline__ cx<T> eta_phase(double eta, int m) {
    double s, c;
    sincos(eta * (double)m, &s, &c);
    return mk<T>((T)c, (T)-s);
}

// tr(T F) of the representatives with the row phases of legs 1 (T) and 2 (F)
template <class T, int ILI, class V>
__device__ __forceinline__ V ss_trace(const double *eta, const V (&Tk)[SS_MAXR], const V (&Fk)[SS_MAXR]) {
    constexpr SSDef D = ssdef(ILI);
    cx<T> p1[7], p2[7];
#pragma unroll
    for (int m = -3; m <= 3; ++m) { p1[m + 3] = eta_phase<T>(eta[1], m); p2[m + 3] = eta_phase<T>(eta[2], m); }
    V acc = VOps<V>::zero();
#pragma unroll
    for (int p = 0; p < D.npair; ++p) {
        const cx<T> ph = p1[D.M[D.pa[p]] + 3] * p2[D.M[D.pb[p]] + 3];
        acc = acc + (Tk[D.kab[p]] * Fk[D.kba[p]]) * ph;
    }
    return acc;
}

// The eta derivatives of tr: d tr / d eta_1 = -i h1, d tr / d eta_2 = -i h2 with h1 (h2) = the sum over the pairs of
// m_a (m_b) times the pair's term (the values of V).
template <class T, int ILI, class V>
__device__ __forceinline__ void ss_trace_eta(const double *eta, const V (&Tk)[SS_MAXR], const V (&Fk)[SS_MAXR],
                                             cx<T> &h1, cx<T> &h2) {
    constexpr SSDef D = ssdef(ILI);
    cx<T> p1[7], p2[7];
#pragma unroll
    for (int m = -3; m <= 3; ++m) { p1[m + 3] = eta_phase<T>(eta[1], m); p2[m + 3] = eta_phase<T>(eta[2], m); }
    h1 = h2 = mk<T>(T(0), T(0));
#pragma unroll
    for (int p = 0; p < D.npair; ++p) {
        const cx<T> term = (VOps<V>::val(Tk[D.kab[p]]) * VOps<V>::val(Fk[D.kba[p]])) *
                           (p1[D.M[D.pa[p]] + 3] * p2[D.M[D.pb[p]] + 3]);
        h1 = h1 + term * T(D.M[D.pa[p]]);
        h2 = h2 + term * T(D.M[D.pb[p]]);
    }
}

}  // namespace fm
