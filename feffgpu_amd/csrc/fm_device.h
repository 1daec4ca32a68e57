// Device-side building blocks of the path-stage kernels: complex arithmetic, the lambda index sets, the
// independent-pair tables of the scattering matrices, the spherical-wave (Hankel) recursion streamed over the
// partial waves, wavefront helpers and the FP64 leg geometry.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/feffgpu_amd.h"

#define FM_WAVE 64
#define FM_TWO_PI 6.283185307179586476925286766559
#define FM_BOHR 0.52917721067

namespace fm {

// ------------------------------------------------------------------------------------------------ complex
template <class T> struct cx { T r, i; };
template <class T> __device__ __forceinline__ cx<T> mk(T r, T i) { return cx<T>{r, i}; }
template <class T> __device__ __forceinline__ cx<T> operator+(cx<T> a, cx<T> b) { return {a.r + b.r, a.i + b.i}; }
template <class T> __device__ __forceinline__ cx<T> operator-(cx<T> a, cx<T> b) { return {a.r - b.r, a.i - b.i}; }
template <class T> __device__ __forceinline__ cx<T> operator*(cx<T> a, cx<T> b) {
    return {a.r * b.r - a.i * b.i, a.r * b.i + a.i * b.r};
}
template <class T> __device__ __forceinline__ cx<T> operator*(cx<T> a, T s) { return {a.r * s, a.i * s}; }
template <class T> __device__ __forceinline__ cx<T> conj(cx<T> a) { return {a.r, -a.i}; }
__device__ __forceinline__ float fmad(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
__device__ __forceinline__ double fmad(double a, double b, double c) { return __builtin_fma(a, b, c); }
// acc += a * b: the product's components first (a multiply and a fused multiply-add), then one rounding of the
// accumulator per term
template <class T> __device__ __forceinline__ void cfma(cx<T> &acc, cx<T> a, cx<T> b) {
    acc.r += a.r * b.r - a.i * b.i;
    acc.i += a.r * b.i + a.i * b.r;
}
// acc += a * b as four fused multiply-adds into the accumulator: two instructions per component instead of three,
// at the cost of two roundings of the accumulator per term (about 1.1-1.3 times the rounding error of cfma over a
// sum of 6-20 terms). Used in the adjoint's sweep only (its per-row contraction and the chain products of each
// energy), where it leaves the FP32 gradient's error against FP64 unchanged; the forward keeps cfma for its chain,
// and its Hankel recursions keep the product-then-subtract form, whose FP32 amplitudes measure more accurate.
template <class T> __device__ __forceinline__ void cfma4(cx<T> &acc, cx<T> a, cx<T> b) {
    acc.r = fmad(a.r, b.r, acc.r);
    acc.r = fmad(-a.i, b.i, acc.r);
    acc.i = fmad(a.r, b.i, acc.i);
    acc.i = fmad(a.i, b.r, acc.i);
}
// c - a * b as four fused multiply-adds
template <class T> __device__ __forceinline__ cx<T> cfms(cx<T> c, cx<T> a, cx<T> b) {
    cx<T> o;
    o.r = fmad(-a.r, b.r, c.r);
    o.r = fmad(a.i, b.i, o.r);
    o.i = fmad(-a.r, b.i, c.i);
    o.i = fmad(-a.i, b.r, o.i);
    return o;
}
// acc += a * s (s real)
template <class T> __device__ __forceinline__ void rfma(cx<T> &acc, cx<T> a, T s) {
    acc.r += a.r * s;
    acc.i += a.i * s;
}
// -i / rho
template <class T> __device__ __forceinline__ cx<T> minus_i_over(cx<T> rho) {
    T d = T(1) / (rho.r * rho.r + rho.i * rho.i);
    return {-rho.i * d, -rho.r * d};
}
template <class T> __device__ __forceinline__ cx<T> cdiv(cx<T> a, cx<T> b) {
    T d = T(1) / (b.r * b.r + b.i * b.i);
    return {(a.r * b.r + a.i * b.i) * d, (a.i * b.r - a.r * b.i) * d};
}

// ------------------------------------------------------------------------------------------------ lambda sets
// The lambda = (m, n) index sets of setlam (src/feffgpu/genfmt.py:115-153) for iorder = 2, in an internal order
// whose first L0 entries are the set's first laml0x (the order inside each part is free: only the trace is
// used). SET 0: three or more legs, K/L1 edge (ilinit 1): laml0x 4. SET 1: three or more legs, L2/L3/M edges:
// laml0x 6. SET 2: two legs, K/L1 edge: mmax = nmax = 1, iord 3.
struct LamDef {
    int L, L0, M[6], N[6];
};
__host__ __device__ constexpr LamDef lamdef(int set) {
    return set == 2   ? LamDef{6, 6, {0, -1, 1, 0, -1, 1}, {0, 0, 0, 1, 1, 1}}
           : set == 1 ? LamDef{6, 6, {0, -1, 1, 0, -2, 2}, {0, 0, 0, 1, 0, 0}}
                      : LamDef{6, 4, {0, -1, 1, 0, -2, 2}, {0, 0, 0, 1, 0, 0}};
}

// Independent pairs of a scattering matrix. The rotation matrices obey d_{-m1,-m2} = (-1)^(m1-m2) d_{m1,m2} and the
// spherical-wave factors depend on |m| only, so before the row phase exp(-i eta m1) the entry of (-m1, n1; -m2, n2)
// is (-1)^(m1-m2) times that of (m1, n1; m2, n2). A representative has m1 > 0, or m1 = 0 and m2 >= 0.
struct PairDef {
    int n;                 // representatives of the 6 x 6 matrix (20)
    int ri[FM_NREP], rj[FM_NREP];
    int idx[6][6];         // (i, j) -> representative
    int sg[6][6];          // sign relating (i, j) to its representative
    int a[FM_NREP], b[FM_NREP];   // Hankel columns: a = |m1| + n1 + 1 of the gam leg, b = n2 + 1 of the gamtl leg
    int m1[FM_NREP], m2[FM_NREP]; // signed m of the representative
};
__host__ __device__ constexpr int flip_of(const LamDef &D, int i) {
    int f = -1;
    for (int k = 0; k < D.L; ++k)
        if (D.M[k] == -D.M[i] && D.N[k] == D.N[i]) f = k;
    return f;
}
__host__ __device__ constexpr bool is_rep(const LamDef &D, int i, int j) {
    return D.M[i] > 0 || (D.M[i] == 0 && D.M[j] >= 0);
}
__host__ __device__ constexpr PairDef pairdef(int set) {
    LamDef D = lamdef(set);
    PairDef P{};
    int k = 0;
    for (int i = 0; i < D.L; ++i)
        for (int j = 0; j < D.L; ++j)
            if (is_rep(D, i, j)) {
                P.ri[k] = i;
                P.rj[k] = j;
                P.a[k] = (D.M[i] < 0 ? -D.M[i] : D.M[i]) + D.N[i] + 1;
                P.b[k] = D.N[j] + 1;
                P.m1[k] = D.M[i];
                P.m2[k] = D.M[j];
                P.idx[i][j] = k;
                P.sg[i][j] = 1;
                ++k;
            }
    P.n = k;
    for (int i = 0; i < D.L; ++i)
        for (int j = 0; j < D.L; ++j)
            if (!is_rep(D, i, j)) {
                int fi = flip_of(D, i), fj = flip_of(D, j);
                P.idx[i][j] = P.idx[fi][fj];
                int dm = D.M[i] - D.M[j];
                P.sg[i][j] = (dm & 1) ? -1 : 1;
            }
    return P;
}

// ------------------------------------------------------------------------------------------------ Hankel factors
// FEFF's spherical-wave factors c[il][im] (sclmz; src/feffgpu/stage.py:209-232) of one leg, produced row by row
// (il = 1, 2, ...) for columns im = 1..3 (the gam leg) or 1..2 (the gamtl leg), with z = -i / rho:
//   col 1: 1, 1 - z, then c[r] = c[r-2] - (2r-3) z c[r-1]
//   col 2: 0, -z, -3z (1 - 2z), then c[r] = c[r-2] - (2r-3) z (c[r-1] + c1[r-1])
//   col 3: 0, 0, 3z^2, 15 z^2 (1 - 3z), then c[r] = c[r-2] - (2r-3) z (c[r-1] + c2[r-1])
// (diagonal seeds cmm_m = -(2m - 1) z cmm_{m-1}, next row cmm_m (2m + 1)(1 - (m+1) z)). FEFF zeroes rows beyond
// the leg's partial-wave limit lxp1; the scattering matrices only read rows il <= the scatterer's lmax + 1 <= lxp1
// of both legs, where no zeroing applies, so it is applied only for the termination (rows 1..4, `mask`).
template <class T> struct Hank3 {
    cx<T> z, a2, a1, b2, b1, c2, c1;   // columns 1, 2, 3: rows r-2 and r-1
    __device__ __forceinline__ void init(cx<T> z_) {
        z = z_;
        a2 = a1 = b2 = b1 = c2 = c1 = mk<T>(T(0), T(0));
    }
    // rows 1..4 (explicit); alive[r] false zeroes row r as FEFF does beyond lxp1 (col 1 rows 1, 2 excepted)
    __device__ __forceinline__ void seed_row(int r, cx<T> (&o)[3], bool alive = true) {
        const cx<T> one = mk<T>(T(1), T(0)), zero = mk<T>(T(0), T(0));
        if (r == 1) { o[0] = one; o[1] = zero; o[2] = zero; }
        else if (r == 2) { o[0] = one - z; o[1] = mk<T>(-z.r, -z.i); o[2] = zero; }
        else if (r == 3) {
            const cx<T> zt = z * T(3);
            o[0] = a2 - zt * a1;
            o[1] = (z * T(-3)) * (one - z * T(2));
            o[2] = z * z * T(3);
        } else {
            const cx<T> zt = z * T(5);
            o[0] = a2 - zt * a1;
            o[1] = b2 - zt * (b1 + a1);
            o[2] = (z * z * T(15)) * (one - z * T(3));
        }
        if (!alive) {
            if (r > 2) o[0] = zero;
            o[1] = zero;
            o[2] = zero;
        }
        a2 = a1; a1 = o[0]; b2 = b1; b1 = o[1]; c2 = c1; c1 = o[2];
    }
    // rows r >= 5
    __device__ __forceinline__ void step(int r, cx<T> (&o)[3]) {
        const cx<T> zt = z * T(2 * r - 3);
        o[0] = a2 - zt * a1;
        o[1] = b2 - zt * (b1 + a1);
        o[2] = c2 - zt * (c1 + b1);
        a2 = a1; a1 = o[0]; b2 = b1; b1 = o[1]; c2 = c1; c1 = o[2];
    }
};
template <class T> struct Hank2 {
    cx<T> z, a2, a1, b2, b1;
    __device__ __forceinline__ void init(cx<T> z_) {
        z = z_;
        a2 = a1 = b2 = b1 = mk<T>(T(0), T(0));
    }
    // rows 1..3 (explicit); row 4 and beyond with step()
    __device__ __forceinline__ void seed_row(int r, cx<T> (&o)[2], bool alive = true) {
        const cx<T> one = mk<T>(T(1), T(0)), zero = mk<T>(T(0), T(0));
        if (r == 1) { o[0] = one; o[1] = zero; }
        else if (r == 2) { o[0] = one - z; o[1] = mk<T>(-z.r, -z.i); }
        else {
            o[0] = a2 - (z * T(3)) * a1;
            o[1] = (z * T(-3)) * (one - z * T(2));
        }
        if (!alive) {
            if (r > 2) o[0] = zero;
            o[1] = zero;
        }
        a2 = a1; a1 = o[0]; b2 = b1; b1 = o[1];
    }
    __device__ __forceinline__ void step(int r, cx<T> (&o)[2], bool alive = true) {
        const cx<T> zt = z * T(2 * r - 3);
        o[0] = a2 - zt * a1;
        o[1] = b2 - zt * (b1 + a1);
        if (!alive) { o[0] = mk<T>(T(0), T(0)); o[1] = o[0]; }
        a2 = a1; a1 = o[0]; b2 = b1; b1 = o[1];
    }
};

// ------------------------------------------------------------------------------------------------ wave helpers
__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}
__device__ __forceinline__ int lane_id() { return threadIdx.x & (FM_WAVE - 1); }

// inclusive prefix sum over the 64 lanes (double)
__device__ __forceinline__ double wave_scan_add(double v) {
    const int lane = lane_id();
#pragma unroll
    for (int off = 1; off < FM_WAVE; off <<= 1) {
        double t = __shfl_up(v, off, FM_WAVE);
        if (lane >= off) v += t;
    }
    return v;
}

// ------------------------------------------------------------------------------------------------ geometry
// pathgeom/trig/arg of FEFF as the batched stage regularises them (src/feffgpu/stage.py:147-206), FP64.
struct Trig { double ct, st, cp, sp; };
__device__ __forceinline__ Trig trig_of(double x, double y, double z) {
    Trig t;
    double r = sqrt(x * x + y * y + z * z), rxy = sqrt(x * x + y * y);
    if (r < 1e-6) { t.ct = 1.0; t.st = 0.0; } else { t.ct = z / r; t.st = rxy / r; }
    if (rxy < 1e-6) { t.cp = t.ct < 0 ? -1.0 : 1.0; t.sp = 0.0; } else { t.cp = x / rxy; t.sp = y / rxy; }
    return t;
}
__device__ __forceinline__ double arg_reg(double x, double y, double fi) {
    if (fabs(x) < 1e-6) x = 0.0;
    if (fabs(y) < 1e-6) y = 0.0;
    return (x == 0.0 && y == 0.0) ? fi : atan2(y, x);
}

struct LegGeo { double r, beta, alpha, gamma; };
// Leg J (1..n) of the closed path rat[0..n] (bohr), rat[0] = rat[n] = absorber: the leg runs from atom im1 to
// atom i, the next leg from i to ip1 (j = n wraps to i = 0, ip1 = 1, im1 = n - 1).
__device__ __forceinline__ LegGeo leg_geometry(const double (*rat)[3], int J, int n) {
    int i, ip1, im1;
    if (J == n) { i = 0; ip1 = 1; im1 = n - 1; } else { i = J; ip1 = J + 1; im1 = J - 1; }
    double vo[3], vi[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        vo[c] = rat[ip1][c] - rat[i][c];
        vi[c] = rat[i][c] - rat[im1][c];
    }
    Trig p = trig_of(vo[0], vo[1], vo[2]);
    Trig q = trig_of(vi[0], vi[1], vi[2]);
    double cppp = q.cp * p.cp + q.sp * p.sp, sppp = p.sp * q.cp - p.cp * q.sp;
    double phi = atan2(q.sp, q.cp), phip = atan2(p.sp, p.cp);
    double ar = -(q.st * p.ct - q.ct * p.st * cppp), ai = p.st * sppp;
    double gr = -(q.st * p.ct * cppp - q.ct * p.st), gi = -q.st * sppp;
    double b = q.ct * p.ct + q.st * p.st * cppp;
    b = fmin(1.0, fmax(-1.0, b));
    double u1x = q.st * q.cp, u1y = q.st * q.sp, u1z = q.ct;
    double u2x = p.st * p.cp, u2y = p.st * p.sp, u2z = p.ct;
    double cx_ = u1y * u2z - u1z * u2y, cy_ = u1z * u2x - u1x * u2z, cz_ = u1x * u2y - u1y * u2x;
    LegGeo g;
    g.beta = atan2(sqrt(cx_ * cx_ + cy_ * cy_ + cz_ * cz_ + 1e-30), b);
    double a_j = arg_reg(ar, ai, phip - phi), g_j = arg_reg(gr, gi, 0.0);
    g.alpha = M_PI - g_j;
    g.gamma = M_PI - a_j;
    g.r = sqrt(vi[0] * vi[0] + vi[1] * vi[1] + vi[2] * vi[2]);
    return g;
}

// Wigner small-d seeds d^l_{m1 m2}(beta), l <= 2, Edmonds' convention as FEFF's rot3i (src/feffgpu/stage.py:52-76).
__device__ __forceinline__ double dseed(int l, int m, int n, double c, double s, double c2, double s2) {
    if (l == 0) return (m == 0 && n == 0) ? 1.0 : 0.0;
    if (l == 1) {
        if (m < -1 || m > 1 || n < -1 || n > 1) return 0.0;
        const double r2 = 0.70710678118654752440;
        if (m == 1 && n == 1) return c2 * c2;
        if (m == 1 && n == 0) return -s * r2;
        if (m == 1 && n == -1) return s2 * s2;
        if (m == 0 && n == 1) return s * r2;
        if (m == 0 && n == 0) return c;
        if (m == 0 && n == -1) return -s * r2;
        if (m == -1 && n == 1) return s2 * s2;
        if (m == -1 && n == 0) return s * r2;
        return c2 * c2;   // (-1, -1)
    }
    // l = 2: base entries with m >= |n| ... reached through the symmetries
    //   d_{nm} = (-1)^{m-n} d_{mn},  d_{-m,-n} = (-1)^{m-n} d_{mn}
    if (m < -2 || m > 2 || n < -2 || n > 2) return 0.0;
    double sign = 1.0;
    int am = m, an = n;
    // bring to m >= 0 and m >= |n| with the symmetries
    if (abs(an) > abs(am)) {   // swap: d_{mn} = (-1)^{m-n} d_{nm}
        int t = am; am = an; an = t;
        if ((m - n) & 1) sign = -sign;
    }
    if (am < 0) {              // d_{mn} = (-1)^{m-n} d_{-m,-n}
        am = -am; an = -an;
        if ((m - n) & 1) sign = -sign;
    }
    double v;
    if (am == 2) {
        if (an == 2) v = c2 * c2 * c2 * c2;
        else if (an == 1) v = -(1 + c) / 2 * s;
        else if (an == 0) v = 0.61237243569579452455 * s * s;      // sqrt(3/8)
        else if (an == -1) v = -(1 - c) / 2 * s;
        else v = s2 * s2 * s2 * s2;
    } else if (am == 1) {
        if (an == 1) v = (1 + c) / 2 * (2 * c - 1);
        else if (an == 0) v = -1.2247448713915890491 * s * c;       // sqrt(3/2)
        else v = (1 - c) / 2 * (2 * c + 1);
    } else {
        v = (3 * c * c - 1) / 2;
    }
    return sign * v;
}

// d/dbeta of the seeds above (same symmetry relations, which are linear).
__device__ __forceinline__ double dseed_db(int l, int m, int n, double c, double s, double c2, double s2) {
    if (l == 0) return 0.0;
    if (l == 1) {
        if (m < -1 || m > 1 || n < -1 || n > 1) return 0.0;
        const double r2 = 0.70710678118654752440;
        if ((m == 1 && n == 1) || (m == -1 && n == -1)) return -0.5 * s;
        if ((m == 1 && n == -1) || (m == -1 && n == 1)) return 0.5 * s;
        if (m == 0 && n == 0) return -s;
        if ((m == 1 && n == 0) || (m == 0 && n == -1)) return -c * r2;
        return c * r2;   // (0, 1), (-1, 0)
    }
    if (m < -2 || m > 2 || n < -2 || n > 2) return 0.0;
    double sign = 1.0;
    int am = m, an = n;
    if (abs(an) > abs(am)) {
        int t = am; am = an; an = t;
        if ((m - n) & 1) sign = -sign;
    }
    if (am < 0) {
        am = -am; an = -an;
        if ((m - n) & 1) sign = -sign;
    }
    double v;
    if (am == 2) {
        if (an == 2) v = -2.0 * c2 * c2 * c2 * s2;
        else if (an == 1) v = 0.5 * s * s - 0.5 * (1 + c) * c;
        else if (an == 0) v = 2.0 * 0.61237243569579452455 * s * c;
        else if (an == -1) v = -0.5 * s * s - 0.5 * (1 - c) * c;
        else v = 2.0 * s2 * s2 * s2 * c2;
    } else if (am == 1) {
        if (an == 1) v = -0.5 * s * (4 * c + 1);
        else if (an == 0) v = -1.2247448713915890491 * (c * c - s * s);
        else v = 0.5 * s * (4 * c - 1);
    } else {
        v = -3.0 * c * s;
    }
    return sign * v;
}

// ------------------------------------------------------------------------------------------------ dual numbers
// Forward-mode derivative of the FP64 leg geometry along one coordinate direction, with the value branches of the
// regularised geometry (trig / arg / clamp) taken as in the forward: the derivative of a constant branch is zero.
struct Dd { double v, d; };
__device__ __forceinline__ Dd dd(double v, double d = 0.0) { return Dd{v, d}; }
__device__ __forceinline__ Dd operator+(Dd a, Dd b) { return {a.v + b.v, a.d + b.d}; }
__device__ __forceinline__ Dd operator-(Dd a, Dd b) { return {a.v - b.v, a.d - b.d}; }
__device__ __forceinline__ Dd operator-(Dd a) { return {-a.v, -a.d}; }
__device__ __forceinline__ Dd operator*(Dd a, Dd b) { return {a.v * b.v, a.d * b.v + a.v * b.d}; }
__device__ __forceinline__ Dd operator/(Dd a, Dd b) { return {a.v / b.v, (a.d * b.v - a.v * b.d) / (b.v * b.v)}; }
__device__ __forceinline__ Dd dsqrt(Dd a) {
    const double s = sqrt(a.v);
    return {s, s > 0 ? a.d / (2.0 * s) : 0.0};
}
__device__ __forceinline__ Dd datan2(Dd y, Dd x) {
    const double q = x.v * x.v + y.v * y.v;
    return {atan2(y.v, x.v), q > 0 ? (x.v * y.d - y.v * x.d) / q : 0.0};
}

struct TrigD { Dd ct, st, cp, sp; };
__device__ __forceinline__ TrigD trig_dual(Dd x, Dd y, Dd z) {
    TrigD t;
    Dd r = dsqrt(x * x + y * y + z * z), rxy = dsqrt(x * x + y * y);
    if (r.v < 1e-6) { t.ct = dd(1.0); t.st = dd(0.0); } else { t.ct = z / r; t.st = rxy / r; }
    if (rxy.v < 1e-6) { t.cp = dd(t.ct.v < 0 ? -1.0 : 1.0); t.sp = dd(0.0); } else { t.cp = x / rxy; t.sp = y / rxy; }
    return t;
}
__device__ __forceinline__ Dd arg_dual(Dd x, Dd y, Dd fi) {
    if (fabs(x.v) < 1e-6) x = dd(0.0);
    if (fabs(y.v) < 1e-6) y = dd(0.0);
    return (x.v == 0.0 && y.v == 0.0) ? fi : datan2(y, x);
}
struct LegGeoD { Dd r, beta, alpha, gamma; };
// The leg from atom `pm` to atom `pi`, the next leg from `pi` to `pp` (leg_geometry's im1, i, ip1).
__device__ __forceinline__ LegGeoD leg_geometry_dual(const Dd (&pm)[3], const Dd (&pi)[3], const Dd (&pp)[3]) {
    Dd vo[3], vi[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        vo[c] = pp[c] - pi[c];
        vi[c] = pi[c] - pm[c];
    }
    TrigD p = trig_dual(vo[0], vo[1], vo[2]);
    TrigD q = trig_dual(vi[0], vi[1], vi[2]);
    Dd cppp = q.cp * p.cp + q.sp * p.sp, sppp = p.sp * q.cp - p.cp * q.sp;
    Dd phi = datan2(q.sp, q.cp), phip = datan2(p.sp, p.cp);
    Dd ar = -(q.st * p.ct - q.ct * p.st * cppp), ai = p.st * sppp;
    Dd gr = -(q.st * p.ct * cppp - q.ct * p.st), gi = -(q.st * sppp);
    Dd b = q.ct * p.ct + q.st * p.st * cppp;
    if (b.v > 1.0) b = dd(1.0);
    if (b.v < -1.0) b = dd(-1.0);
    Dd u1x = q.st * q.cp, u1y = q.st * q.sp, u1z = q.ct;
    Dd u2x = p.st * p.cp, u2y = p.st * p.sp, u2z = p.ct;
    Dd cx_ = u1y * u2z - u1z * u2y, cy_ = u1z * u2x - u1x * u2z, cz_ = u1x * u2y - u1y * u2x;
    LegGeoD g;
    g.beta = datan2(dsqrt(cx_ * cx_ + cy_ * cy_ + cz_ * cz_ + dd(1e-30)), b);
    Dd a_j = arg_dual(ar, ai, phip - phi), g_j = arg_dual(gr, gi, dd(0.0));
    g.alpha = dd(M_PI) - g_j;
    g.gamma = dd(M_PI) - a_j;
    g.r = dsqrt(vi[0] * vi[0] + vi[1] * vi[1] + vi[2] * vi[2]);
    return g;
}

// ------------------------------------------------------------------------------------------------ Hankel tangents
// The spherical-wave factors and their derivative with respect to the leg length r (z = -i/(ck r), dz/dr = -z/r),
// the recursions of Hank3/Hank2 differentiated row by row.
template <class T> struct Hank3T {
    cx<T> z, zd, a2, a1, b2, b1, c2, c1, a2d, a1d, b2d, b1d, c2d, c1d;
    T mr;   // -1 / r
    __device__ __forceinline__ void init(cx<T> z_, T r) {
        z = z_;
        mr = T(-1) / r;
        zd = z * mr;
        a2 = a1 = b2 = b1 = c2 = c1 = a2d = a1d = b2d = b1d = c2d = c1d = mk<T>(T(0), T(0));
    }
    __device__ __forceinline__ void shift(const cx<T> (&o)[3], const cx<T> (&od)[3]) {
        a2 = a1; a1 = o[0]; b2 = b1; b1 = o[1]; c2 = c1; c1 = o[2];
        a2d = a1d; a1d = od[0]; b2d = b1d; b1d = od[1]; c2d = c1d; c1d = od[2];
    }
    __device__ __forceinline__ void seed_row(int r, cx<T> (&o)[3], cx<T> (&od)[3], bool alive = true) {
        const cx<T> one = mk<T>(T(1), T(0)), zero = mk<T>(T(0), T(0));
        if (r == 1) { o[0] = one; o[1] = zero; o[2] = zero; od[0] = od[1] = od[2] = zero; }
        else if (r == 2) {
            o[0] = one - z; o[1] = mk<T>(-z.r, -z.i); o[2] = zero;
            od[0] = mk<T>(-zd.r, -zd.i); od[1] = od[0]; od[2] = zero;
        } else if (r == 3) {
            o[0] = a2 - (z * T(3)) * a1;
            od[0] = a2d - (zd * a1 + z * a1d) * T(3);
            o[1] = (z * T(-3)) * (one - z * T(2));
            od[1] = zd * T(-3) + (z * zd) * T(12);
            o[2] = z * z * T(3);
            od[2] = (z * zd) * T(6);
        } else {
            o[0] = a2 - (z * T(5)) * a1;
            od[0] = a2d - (zd * a1 + z * a1d) * T(5);
            o[1] = b2 - (z * T(5)) * (b1 + a1);
            od[1] = b2d - (zd * (b1 + a1) + z * (b1d + a1d)) * T(5);
            o[2] = (z * z * T(15)) * (one - z * T(3));
            od[2] = (z * zd) * T(30) - (z * z * zd) * T(135);
        }
        if (!alive) {
            if (r > 2) { o[0] = zero; od[0] = zero; }
            o[1] = o[2] = od[1] = od[2] = zero;
        }
        shift(o, od);
    }
    // with zt = (2r - 3) z and d zt / dr = -zt / r: od = a2d - zt (s' - s / r)
    __device__ __forceinline__ void step(int r, cx<T> (&o)[3], cx<T> (&od)[3]) {
        const cx<T> zt = z * T(2 * r - 3);
        const cx<T> s0 = a1, s1 = b1 + a1, s2 = c1 + b1;
        const cx<T> s0d = a1d, s1d = b1d + a1d, s2d = c1d + b1d;
        o[0] = cfms(a2, zt, s0);
        o[1] = cfms(b2, zt, s1);
        o[2] = cfms(c2, zt, s2);
        od[0] = cfms(a2d, zt, mk<T>(s0d.r + s0.r * mr, s0d.i + s0.i * mr));
        od[1] = cfms(b2d, zt, mk<T>(s1d.r + s1.r * mr, s1d.i + s1.i * mr));
        od[2] = cfms(c2d, zt, mk<T>(s2d.r + s2.r * mr, s2d.i + s2.i * mr));
        shift(o, od);
    }
};
template <class T> struct Hank2T {
    cx<T> z, zd, a2, a1, b2, b1, a2d, a1d, b2d, b1d;
    T mr;   // -1 / r
    __device__ __forceinline__ void init(cx<T> z_, T r) {
        z = z_;
        mr = T(-1) / r;
        zd = z * mr;
        a2 = a1 = b2 = b1 = a2d = a1d = b2d = b1d = mk<T>(T(0), T(0));
    }
    __device__ __forceinline__ void shift(const cx<T> (&o)[2], const cx<T> (&od)[2]) {
        a2 = a1; a1 = o[0]; b2 = b1; b1 = o[1];
        a2d = a1d; a1d = od[0]; b2d = b1d; b1d = od[1];
    }
    __device__ __forceinline__ void seed_row(int r, cx<T> (&o)[2], cx<T> (&od)[2], bool alive = true) {
        const cx<T> one = mk<T>(T(1), T(0)), zero = mk<T>(T(0), T(0));
        if (r == 1) { o[0] = one; o[1] = zero; od[0] = od[1] = zero; }
        else if (r == 2) { o[0] = one - z; o[1] = mk<T>(-z.r, -z.i); od[0] = mk<T>(-zd.r, -zd.i); od[1] = od[0]; }
        else {
            o[0] = a2 - (z * T(3)) * a1;
            od[0] = a2d - (zd * a1 + z * a1d) * T(3);
            o[1] = (z * T(-3)) * (one - z * T(2));
            od[1] = zd * T(-3) + (z * zd) * T(12);
        }
        if (!alive) {
            if (r > 2) { o[0] = zero; od[0] = zero; }
            o[1] = od[1] = zero;
        }
        shift(o, od);
    }
    __device__ __forceinline__ void step(int r, cx<T> (&o)[2], cx<T> (&od)[2], bool alive = true) {
        const cx<T> zt = z * T(2 * r - 3);
        const cx<T> s0 = a1, s1 = b1 + a1, s0d = a1d, s1d = b1d + a1d;
        o[0] = cfms(a2, zt, s0);
        o[1] = cfms(b2, zt, s1);
        od[0] = cfms(a2d, zt, mk<T>(s0d.r + s0.r * mr, s0d.i + s0.i * mr));
        od[1] = cfms(b2d, zt, mk<T>(s1d.r + s1.r * mr, s1d.i + s1.i * mr));
        if (!alive) { o[0] = o[1] = od[0] = od[1] = mk<T>(T(0), T(0)); }
        shift(o, od);
    }
};

}  // namespace fm
