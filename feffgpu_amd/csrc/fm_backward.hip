// Adjoint (reverse-mode) kernels of the path stage: d loss / d coordinates (GPUPathStage.cchi), d loss / d (cchi,
// reff) (chi_fine) and d loss / d positions for a path list (the fused entry), the gradient that feffgpu obtains by
// autograd through src/feffgpu/stage.py:147-388 and the gather of src/feffgpu/exafs/forward.py:153-163.
//
// The workgroup layout is the forward's: G = ed.gpaths paths and 4 wavefronts, wavefront w takes the energy band
// [gstart[w], gstart[w+1]) of all G paths, lane (path j, energy i) = j * (64 / G) + i. Nothing is kept from the
// forward call:
//   1. per path (one wave): leg geometry (FP64), the rotation tables d^l(beta) and their beta-derivatives into LDS;
//   2. per (path, energy) lane, pass A: the amplitude, with the scattering matrices F_s and the prefix products
//      Pre_{s+1} = F_s ... F_1 stored in the lane's slice of the workspace;
//   3. per path (one wave): the adjoint of the fine grid (path lists) or the given dL/dcchi: dL/dcchi per energy;
//   4. per (path, energy) lane, pass B: the chain walked backwards with U_s = T F_{n-1} ... F_{s+1}; every scattering
//      matrix is streamed over the partial waves again for the forward-mode tangents of its two leg lengths (the
//      Hankel recursions differentiated) and of its rotation angle (the d recursion differentiated), contracted on
//      the fly with its adjoint conj(dL/dptrac) (Pre_s U_s)^T; eta enters analytically through exp(-i eta m). Each
//      lane writes its per-leg contributions once to its gradient slots (workspace);
//   5. per path (one wave): the slots summed over the path's lanes in a fixed order, then pushed through the FP64
//      geometry with dual numbers (one lane per coordinate); for a path list gathered per atom in a fixed order.
#include <hip/hip_runtime.h>
#include <stdio.h>
#include <string.h>

#include <mutex>
#include <type_traits>

#include "fm_device.h"
#include "fm_path.h"

using namespace fm;

// The workspace of fm_cchi_backward_noalloc: a static device array, one per GPU the library runs on, held until the
// process ends. Host code never passes `scratch` itself to a kernel (in host code the name is a stand-in, not a
// device address): it looks the address up per device with hipGetSymbolAddress (scratch_of_current_device).
constexpr size_t FM_SCRATCH_BYTES = size_t(64) << 20;
static __device__ __align__(256) unsigned char scratch[FM_SCRATCH_BYTES];

int fm_set_err(int code, const char *msg);
size_t fm_lds_budget(int kernel);

namespace {

__host__ __device__ constexpr int pre_elems(int set) { return lamdef(set).L * lamdef(set).L0; }   // one prefix product
__host__ __device__ inline int nacc(int nlegmax) { return 4 * nlegmax + 2; }   // gradient slots per lane

// Gradient slots of a lane: kind X (leg length through the gam leg), Y (leg length through the gamtl leg), B (beta),
// H (eta), each for legs 1..nlegmax, then the two sums of the cfac term.
enum { GK_X = 0, GK_Y = 1, GK_B = 2, GK_H = 3 };

struct BArgs {
    fm_edge ed;                 // the edge of every path of this launch
    fm_bands bp;                // its band plan for gp paths per workgroup
    int gp;
    int npath_bucket, first, edge_id;
    const int32_t *order;
    PathSrc src;
    const void *gcchi;          // mode 0: [npath][E] complex
    const double *greff;        // mode 0: [npath] or NULL
    fm_grid G;                  // mode 1
    const int32_t *segment;     // mode 1
    const double *gseg;         // mode 1: [nseg][nk]
    const void *gpath;          // mode 1: [npath][nk] or NULL
    double *gcoord;             // [npath][nslot][3]
    int nslot, ltab, nk;
    int nlegmax;                // the leg count of every path of this launch
    size_t path_bytes, off_W, off_Wd, off_cc, off_amp, off_ang, off_cb, off_red;
    size_t off_qs;              // 3 nk doubles of staging for the fine-grid adjoint, or 0 (none)
    size_t off_gs;              // 3 n^2 doubles of scratch for the geometry adjoint
    unsigned char *work;        // per band slot: prefix products and scattering matrices, gradient slots
    size_t work_per_wave, work_off_F, work_off_acc, work_bytes;
    size_t dstride;             // fused adjoint: complex elements of a path's workspace region, (4 nleg + 1) E
    size_t dbytes;              // fused adjoint: bytes of the launch's regions (its paths and one spare)
    unsigned *tickets;          // the path groups' counter pair of a persistent launch, or NULL (fixed assignment)
    const void *keep;           // fused adjoint: the launch's kept scattering matrices (its first path first), or NULL
};

int set_err(int code, const char *msg) { return fm_set_err(code, msg); }
int hip_err(hipError_t e, const char *where) {
    char buf[256];
    snprintf(buf, sizeof buf, "%s: %s", where, hipGetErrorString(e));
    return fm_set_err(FM_EHIP, buf);
}

constexpr size_t FM_QS_MAX = 32 * 1024;   // the fine-grid adjoint stages its q points in LDS up to this size

// LDS of one path. fused: the rotation tables are dead when the fine-grid adjoint runs, so its staging reuses them.
template <class T, int SET> __host__ void layout(BArgs &a, const fm_edge &ed, bool fused = false) {
    size_t o = align16(sizeof(PathLDS));
    const size_t tab = SET >= FM_SET_SS ? 0 : align16(sizeof(T) * rot_elems(a.nlegmax, a.ltab));
    a.off_W = o; o += tab;
    a.off_Wd = o; o += tab;
    a.off_cc = o; o += align16(sizeof(cx<T>) * ed.E);
    a.off_amp = o; o += align16(sizeof(T) * ed.E);
    a.off_ang = o; o += align16(sizeof(double) * ed.E);
    a.off_cb = o; o += align16(sizeof(cx<T>) * ed.E);
    a.off_red = o; o += align16(sizeof(double) * (nacc(a.nlegmax) + 6 * FM_MAX_LEGS));
    const size_t qs = sizeof(double) * 3 * (size_t)a.nk;
    a.off_qs = 0;
    if (a.src.mode == 1 && qs > 0 && qs <= FM_QS_MAX) {
        if (fused && 2 * tab >= qs) a.off_qs = a.off_W;
        else { a.off_qs = o; o += align16(qs); }
    }
    // the rotation tables are dead when the geometry adjoint runs
    const size_t gsb = sizeof(double) * 3 * (size_t)a.nlegmax * a.nlegmax;
    if (2 * tab >= gsb) a.off_gs = a.off_W;
    else { a.off_gs = o; o += align16(gsb); }
    a.path_bytes = path_stride(o);
    const int nslots = a.nlegmax > 2 ? a.nlegmax - 1 : 1;
    const bool store = SET < FM_SET_SS;   // single scattering keeps nothing but the gradient slots
    a.work_off_F = store ? sizeof(cx<T>) * (size_t)nslots * pre_elems(SET) * FM_WAVE : 0;
    a.work_off_acc = store ? a.work_off_F + sizeof(cx<T>) * (size_t)(a.nlegmax - 1) * FM_NREP * FM_WAVE : 0;
    a.work_per_wave = (a.work_off_acc + sizeof(T) * (size_t)nacc(a.nlegmax) * FM_WAVE + 255) & ~size_t(255);
}

// One partial-wave row of the adjoint of a block: the contractions of the row's tangents with K added to gX (gam
// leg length), gY (gamtl leg length), gB (rotation angle).
template <class T, int SET, int NR, int NC>
__device__ __forceinline__ void bwd_row(const cx<T> (&K)[FM_NREP], const cx<T> (&xa)[3],
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

// The adjoint of fblock (its tangents contracted with K): rows il = 1..ilx of the lane, t-matrix fetched one row
// ahead.
template <class T, int SET, int NR, int NC>
__device__ __forceinline__ void fblock_bwd(const cx<T> (&K)[FM_NREP], cx<T> ck, T rX, T rY, int ilx,
                                           const T *__restrict__ Wl, const T *__restrict__ Wdl,
                                           const T *__restrict__ tl, int tlstride, T &gX, T &gY, T &gB) {
    gX = gY = gB = T(0);
    Hank3T<T> X;
    Hank2T<T> Y;
    X.init(minus_i_over(ck * rX), rX);
    Y.init(minus_i_over(ck * rY), rY);
    cx<T> xa[3], xd[3], ya[2], yd[2];
    cx<T> t = mk<T>(tl[tlstride], tl[tlstride + 1]);
    const int rseed = ilx < 4 ? ilx : 4;
    for (int r = 1; r <= rseed; ++r) {
        X.seed_row(r, xa, xd);
        if (r < 4) Y.seed_row(r, ya, yd); else Y.step(r, ya, yd);
        const size_t rn = (size_t)(r < ilx ? r + 1 : r) * tlstride;
        const cx<T> tn = mk<T>(tl[rn], tl[rn + 1]);
        bwd_row<T, SET, NR, NC>(K, xa, xd, ya[0] * t, ya[1] * t, yd[0] * t, yd[1] * t, Wl + (size_t)r * FM_NREP,
                                Wdl + (size_t)r * FM_NREP, gX, gY, gB);
        t = tn;
    }
#pragma unroll 2
    for (int r = 5; r <= ilx; ++r) {
        X.step(r, xa, xd);
        Y.step(r, ya, yd);
        const size_t rn = (size_t)(r < ilx ? r + 1 : r) * tlstride;
        const cx<T> tn = mk<T>(tl[rn], tl[rn + 1]);
        bwd_row<T, SET, NR, NC>(K, xa, xd, ya[0] * t, ya[1] * t, yd[0] * t, yd[1] * t, Wl + (size_t)r * FM_NREP,
                                Wdl + (size_t)r * FM_NREP, gX, gY, gB);
        t = tn;
    }
}

// Lane-private arrays in the workspace slice of one wavefront: element-major, lane-minor (coalesced), accessed
// with buffer instructions. The slice's descriptor is wave-uniform (scalar registers), the lane enters as one
// 32-bit offset and the element as a scalar offset, so no 64-bit address is formed per element. Accesses outside
// the slice read 0 and write nothing.
constexpr int FM_RSRC_WORD3 = 0x00020000;   // buffer descriptor word 3 on gfx9: 32-bit data format, no swizzle
__device__ __forceinline__ __amdgpu_buffer_rsrc_t slice_rsrc(unsigned char *p, size_t bytes) {
    return __builtin_amdgcn_make_buffer_rsrc(p, (short)0, (int)bytes, FM_RSRC_WORD3);
}
template <class E>
__device__ __forceinline__ void buf_st(__amdgpu_buffer_rsrc_t rs, uint32_t vo, int so, const E &v) {
    if constexpr (sizeof(E) == 4) {
        unsigned x;
        __builtin_memcpy(&x, &v, 4);
        __builtin_amdgcn_raw_buffer_store_b32(x, rs, vo, so, 0);
    } else if constexpr (sizeof(E) == 8) {
        __attribute__((ext_vector_type(2))) unsigned x;
        __builtin_memcpy(&x, &v, 8);
        __builtin_amdgcn_raw_buffer_store_b64(x, rs, vo, so, 0);
    } else {
        static_assert(sizeof(E) == 16, "element of 4, 8 or 16 bytes");
        __attribute__((ext_vector_type(4))) unsigned x;
        __builtin_memcpy(&x, &v, 16);
        __builtin_amdgcn_raw_buffer_store_b128(x, rs, vo, so, 0);
    }
}
template <class E> struct LaneArray {
    __amdgpu_buffer_rsrc_t rs;
    uint32_t lo;           // lane * sizeof(E)
    __device__ __forceinline__ E ld(int elem) const {
        const int so = elem * (int)(FM_WAVE * sizeof(E));
        E v;
        if constexpr (sizeof(E) == 4) {
            const unsigned x = __builtin_amdgcn_raw_buffer_load_b32(rs, lo, so, 0);
            __builtin_memcpy(&v, &x, 4);
        } else if constexpr (sizeof(E) == 8) {
            const auto x = __builtin_amdgcn_raw_buffer_load_b64(rs, lo, so, 0);
            __builtin_memcpy(&v, &x, 8);
        } else {
            static_assert(sizeof(E) == 16, "element of 4, 8 or 16 bytes");
            const auto x = __builtin_amdgcn_raw_buffer_load_b128(rs, lo, so, 0);
            __builtin_memcpy(&v, &x, 16);
        }
        return v;
    }
    __device__ __forceinline__ void st(int elem, const E &v) const { buf_st(rs, lo, elem * (int)(FM_WAVE * sizeof(E)), v); }
};

// Per-lane gradient slots (kind, leg j), each written once per path.
template <class T> struct GSlots {
    LaneArray<T> v;
    int nlegmax;
    __device__ __forceinline__ void put(int kind, int j, T x) const { v.st(kind * nlegmax + j - 1, x); }
    __device__ __forceinline__ void put_slot(int slot, T x) const { v.st(slot, x); }
};

// Prefix products (slot, element) and the FM_NREP representatives of F_s (in slot s - 1).
template <class T, int SET> struct PreStore {
    LaneArray<cx<T>> pre, fs;
    __device__ __forceinline__ cx<T> get(int slot, int elem) const { return pre.ld(slot * pre_elems(SET) + elem); }
    __device__ __forceinline__ void set(int slot, int elem, cx<T> v) const { pre.st(slot * pre_elems(SET) + elem, v); }
    __device__ __forceinline__ void put_f(int s, const cx<T> (&Ft)[FM_NREP]) const {
#pragma unroll
        for (int k = 0; k < FM_NREP; ++k) fs.st((s - 1) * FM_NREP + k, Ft[k]);
    }
    __device__ __forceinline__ void get_f(int s, cx<T> (&Ft)[FM_NREP]) const {
#pragma unroll
        for (int k = 0; k < FM_NREP; ++k) Ft[k] = fs.ld((s - 1) * FM_NREP + k);
    }
    // Pre_s (rows L; L0 for s = n, whose slot holds Pre_n)
    __device__ __forceinline__ void mat(int s, const PathLDS &, cx<T> (&P)[lamdef(SET).L][lamdef(SET).L0]) const {
        constexpr int L = lamdef(SET).L, L0 = lamdef(SET).L0;
#pragma unroll
        for (int i = 0; i < L; ++i)
#pragma unroll
            for (int b = 0; b < L0; ++b) P[i][b] = get(s - 2, i * L0 + b);
    }
};

// Pass A: the amplitude at energy ee with the prefix products stored: slot s-2 holds Pre_s = F_{s-1} ... F_1
// (L x L0) for s = 2..n-1, slot max(n-2, 0) holds Pre_n (L0 x L0). Returns ptrac.
template <class T, int SET>
__device__ __forceinline__ cx<T> pass_A(const fm_edge &ed, const PathLDS &g, const T *W, int ltab, int ee, int n, PreStore<T, SET> pre) {
    constexpr LamDef D = lamdef(SET);
    constexpr PairDef PD = pairdef(SET);
    constexpr int L = D.L, L0 = D.L0;
    const int E = ed.E;
    const cx<T> ck = mk<T>((T)ed.ck[2 * ee], (T)ed.ck[2 * ee + 1]);
    const int tlstride = 2 * E;
    const T *tlbase = TL<T>::tl(ed);
    cx<T> Ft[FM_NREP], ph[6];
    auto scat = [&](int s, int &ilx, const T *&tl, const T *&Wl) {
        const int ip = g.ipot[s];
        ilx = ed.lmax[(size_t)ip * E + ee] + 1;
        tl = tlbase + ((size_t)ip * ltab) * tlstride + 2 * ee;
        Wl = W + (size_t)(s - 1) * ltab * FM_NREP;
    };
    const int pn_slot = n > 2 ? n - 2 : 0;
    if (n == 2) {
        int ilx;
        const T *tl, *Wl;
        scat(1, ilx, tl, Wl);
        fblock<T, SET, L0, L0>(Ft, ck, (T)g.r[2], (T)g.r[1], ilx, Wl, tl, tlstride);
        pre.put_f(1, Ft);
        row_phases<T, SET>(ph, g, 2);
#pragma unroll
        for (int a = 0; a < L0; ++a)
#pragma unroll
            for (int b = 0; b < L0; ++b) pre.set(pn_slot, a * L0 + b, ph[a] * Ft[PD.idx[a][b]] * T(PD.sg[a][b]));
    } else {
        {   // Pre_2 = F_1 (L x L0)
            int ilx;
            const T *tl, *Wl;
            scat(1, ilx, tl, Wl);
            fblock<T, SET, L, L0>(Ft, ck, (T)g.r[2], (T)g.r[1], ilx, Wl, tl, tlstride);
            pre.put_f(1, Ft);
            row_phases<T, SET>(ph, g, 2);
#pragma unroll
            for (int i = 0; i < L; ++i)
#pragma unroll
                for (int b = 0; b < L0; ++b) pre.set(0, i * L0 + b, ph[i] * Ft[PD.idx[i][b]] * T(PD.sg[i][b]));
        }
#pragma unroll 1
        for (int s = 2; s <= n - 2; ++s) {   // Pre_{s+1} = F_s Pre_s
            int ilx;
            const T *tl, *Wl;
            scat(s, ilx, tl, Wl);
            fblock<T, SET, L, L>(Ft, ck, (T)g.r[s + 1], (T)g.r[s], ilx, Wl, tl, tlstride);
            pre.put_f(s, Ft);
            row_phases<T, SET>(ph, g, s + 1);
#pragma unroll
            for (int b = 0; b < L0; ++b) {
                cx<T> col[L];
#pragma unroll
                for (int j = 0; j < L; ++j) col[j] = pre.get(s - 2, j * L0 + b);
#pragma unroll
                for (int i = 0; i < L; ++i) {
                    cx<T> acc = mk<T>(T(0), T(0));
#pragma unroll
                    for (int j = 0; j < L; ++j) cfma(acc, Ft[PD.idx[i][j]] * T(PD.sg[i][j]), col[j]);
                    pre.set(s - 1, i * L0 + b, ph[i] * acc);
                }
            }
        }
        {   // Pre_n = F_{n-1} Pre_{n-1} (L0 x L0)
            int ilx;
            const T *tl, *Wl;
            scat(n - 1, ilx, tl, Wl);
            fblock<T, SET, L0, L>(Ft, ck, (T)g.r[n], (T)g.r[n - 1], ilx, Wl, tl, tlstride);
            pre.put_f(n - 1, Ft);
            row_phases<T, SET>(ph, g, n);
            cx<T> pn[L0][L0];
#pragma unroll
            for (int b = 0; b < L0; ++b) {
                cx<T> col[L];
#pragma unroll
                for (int j = 0; j < L; ++j) col[j] = pre.get(n - 3, j * L0 + b);
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
                for (int b = 0; b < L0; ++b) pre.set(pn_slot, i * L0 + b, pn[i][b]);
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
    const T *chw = TL<T>::chw(ed);
    cx<T> xa[3], yb[2];
#pragma unroll
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
    row_phases<T, SET>(ph, g, 1);
    cx<T> ptrac = mk<T>(T(0), T(0));
#pragma unroll
    for (int a = 0; a < L0; ++a)
#pragma unroll
        for (int b = 0; b < L0; ++b) cfma(ptrac, ph[a] * Tt[PD.idx[a][b]] * T(PD.sg[a][b]), pre.get(pn_slot, b * L0 + a));
    return ptrac;
}

// One step of pass B for scatterer s with an NR x NC block, with U = U_s = T F_{n-1} ... F_{s+1}:
//   K[rep]  = sum over the orbit of H[l1][l2] ph[l1] sg, H[l1][l2] = conj(pbar) M[l2][l1], M = Pre_s U (NC x NR);
//   Kp[rep] = the same sum weighted by m_{l1}, so that dL/deta_{s+1} = Im sum_rep Kp[rep] Ft[rep];
//   U <- U diag(ph) F_s (unless last), with F_s as pass A stored it;
// then F_s is streamed over the partial waves for the tangents of its two leg lengths and its angle, contracted
// with K. Only K and U are live across the stream.
template <class T, int SET, int NR, int NC, int WPE, class Src>
__device__ __forceinline__ void chain_step(const fm_edge &ed, const PathLDS &g, const T *W, const T *Wd, int ltab,
                                           int ee, int s, cx<T> pbc, cx<T> (&U)[lamdef(SET).L0][lamdef(SET).L],
                                           const Src &pre, bool has_pre, bool last, GSlots<T> gs) {
    constexpr LamDef D = lamdef(SET);
    constexpr PairDef PD = pairdef(SET);
    constexpr int L0 = D.L0;
    const int E = ed.E;
    const cx<T> ck = mk<T>((T)ed.ck[2 * ee], (T)ed.ck[2 * ee + 1]);
    cx<T> ph[6];
    row_phases<T, SET>(ph, g, s + 1);
    cx<T> K[FM_NREP];
    {
        cx<T> Kp[FM_NREP];
        cx<T> P[D.L][L0];
        if (has_pre) pre.mat(s, g, P);
#pragma unroll
        for (int k = 0; k < FM_NREP; ++k) K[k] = Kp[k] = mk<T>(T(0), T(0));
#pragma unroll
        for (int l2 = 0; l2 < NC; ++l2) {
            cx<T> prow[L0];
#pragma unroll
            for (int a = 0; a < L0; ++a)
                prow[a] = has_pre ? P[l2][a] : (a == l2 ? mk<T>(T(1), T(0)) : mk<T>(T(0), T(0)));
#pragma unroll
            for (int l1 = 0; l1 < NR; ++l1) {
                cx<T> m = mk<T>(T(0), T(0));
#pragma unroll
                for (int a = 0; a < L0; ++a) cfma(m, prow[a], U[a][l1]);
                const cx<T> h = pbc * m * ph[l1];
                const T hs = T(PD.sg[l1][l2] * D.M[l1]);
                if (PD.sg[l1][l2] > 0) K[PD.idx[l1][l2]] = K[PD.idx[l1][l2]] + h;
                else K[PD.idx[l1][l2]] = K[PD.idx[l1][l2]] - h;
                if (D.M[l1] != 0) rfma(Kp[PD.idx[l1][l2]], h, hs);
            }
        }
        cx<T> Ft[FM_NREP];
        pre.get_f(s, Ft);
        // eta_{s+1}: sum over (l1, l2) of m_{l1} Im(H F), F = ph sg Ft
        T geta = T(0);
#pragma unroll
        for (int k = 0; k < FM_NREP; ++k)
            if (PD.ri[k] < NR && PD.rj[k] < NC) geta += Kp[k].r * Ft[k].i + Kp[k].i * Ft[k].r;
        gs.put(GK_H, s + 1, geta);
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
    const int ip = g.ipot[s];
    const int ilx = ed.lmax[(size_t)ip * E + ee] + 1;
    const T *tl = TL<T>::tl(ed) + ((size_t)ip * ltab) * (2 * E) + 2 * ee;
    T gX, gY, gB;
    fblock_bwd<T, SET, NR, NC>(K, ck, (T)g.r[s + 1], (T)g.r[s], ilx, W + (size_t)(s - 1) * ltab * FM_NREP,
                               Wd + (size_t)(s - 1) * ltab * FM_NREP, tl, 2 * E, gX, gY, gB);
    gs.put(GK_X, s + 1, gX);
    gs.put(GK_Y, s, gY);
    gs.put(GK_B, s, gB);
}

// Pass B at energy ee: cbar = dL/dcchi, cc = cchi.
template <class T, int SET, int WPE, class Src>
__device__ __forceinline__ void pass_B(const fm_edge &ed, const PathLDS &g, const T *W, const T *Wd, int ltab, int ee, int n,
                       cx<T> cbar, cx<T> cc, const Src &pre, GSlots<T> gs) {
    constexpr LamDef D = lamdef(SET);
    constexpr PairDef PD = pairdef(SET);
    constexpr int L = D.L, L0 = D.L0;
    const int E = ed.E;
    // cfac's own dependence on the leg lengths: d cfac / d r_j = cfac (i (ck - xk) - 1 / r_j), i.e. a1 - a2 / r_j
    {
        const cx<T> w = conj(cbar) * cc;
        const cx<T> dk = mk<T>((T)ed.ckmxk[2 * ee], (T)ed.ckmxk[2 * ee + 1]);
        const cx<T> wdk = w * dk;
        gs.put_slot(4 * gs.nlegmax, -wdk.i);
        gs.put_slot(4 * gs.nlegmax + 1, w.r);
    }
    const cx<T> cf = cfac_at<T>(ed, g, ee, n);
    const cx<T> pbc = conj(cbar * conj(cf));          // conj(dL/dptrac)
    // ---- termination: H_T[a][b] = pbc Pre_n[b][a]; tangents of leg 1 (X), leg n (Y), beta_n; eta_1
    const cx<T> ck = mk<T>((T)ed.ck[2 * ee], (T)ed.ck[2 * ee + 1]);
    cx<T> U[L0][L];
    {
        cx<T> Pn[L][L0];   // Pre_n (rows L0)
        pre.mat(n, g, Pn);
        cx<T> ph[6];
        row_phases<T, SET>(ph, g, 1);
        cx<T> K[FM_NREP];
#pragma unroll
        for (int k = 0; k < FM_NREP; ++k) K[k] = mk<T>(T(0), T(0));
#pragma unroll
        for (int a = 0; a < L0; ++a)
#pragma unroll
            for (int b = 0; b < L0; ++b) {
                const cx<T> h = pbc * Pn[b][a] * ph[a];
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
        const T *chw = TL<T>::chw(ed);
        T gX = T(0), gY = T(0), gB = T(0);
        cx<T> xa[3], xd[3], ya[2], yd[2];
#pragma unroll
        for (int r = 1; r <= 4; ++r) {
            X.seed_row(r, xa, xd, r <= lx1);
            if (r < 4) Y.seed_row(r, ya, yd, r <= lxn); else Y.step(r, ya, yd, r <= lxn);
            int c = -1;
            for (int q = 0; q < ed.nchan; ++q)
                if (ed.chan_il[q] == r) c = q;
            if (c >= 0) {
                const cx<T> w = mk<T>(chw[((size_t)c * E + ee) * 2], chw[((size_t)c * E + ee) * 2 + 1]) * T(2 * r - 1);
                accum_row<T, SET, L0, L0>(Tt, xa, ya[0] * w, ya[1] * w, Wn + (size_t)r * FM_NREP);
                bwd_row<T, SET, L0, L0>(K, xa, xd, ya[0] * w, ya[1] * w, yd[0] * w, yd[1] * w,
                                        Wn + (size_t)r * FM_NREP, Wdn + (size_t)r * FM_NREP, gX, gY, gB);
            }
        }
        gs.put(GK_X, 1, gX);
        gs.put(GK_Y, n, gY);
        gs.put(GK_B, n, gB);
        T geta = T(0);
#pragma unroll
        for (int a = 0; a < L0; ++a)
#pragma unroll
            for (int b = 0; b < L; ++b) {
                const cx<T> t = (b < L0) ? ph[a] * Tt[PD.idx[a][b]] * T(PD.sg[a][b]) : mk<T>(T(0), T(0));
                U[a][b] = t;
                if (b < L0 && D.M[a] != 0) {
                    const cx<T> hf = pbc * Pn[b][a] * t;
                    geta += T(D.M[a]) * hf.i;
                }
            }
        gs.put(GK_H, 1, geta);
    }
    // ---- the chain, s = n-1 .. 1
    if (n == 2) {
        chain_step<T, SET, L0, L0, WPE, Src>(ed, g, W, Wd, ltab, ee, 1, pbc, U, pre, false, true, gs);
    } else {
        chain_step<T, SET, L0, L, WPE, Src>(ed, g, W, Wd, ltab, ee, n - 1, pbc, U, pre, true, false, gs);
#pragma unroll 1
        for (int s = n - 2; s >= 2; --s)
            chain_step<T, SET, L, L, WPE, Src>(ed, g, W, Wd, ltab, ee, s, pbc, U, pre, true, false, gs);
        chain_step<T, SET, L, L0, WPE, Src>(ed, g, W, Wd, ltab, ee, 1, pbc, U, pre, false, true, gs);
    }
}

// ------------------------------------------------------------------------------------------------ fused adjoint
// The adjoint of paths of NL legs in one sweep per energy, with nothing of the amplitude leaving the registers.
// ptrac is a polynomial in the entries of the scattering matrices, so for every real parameter theta of the
// geometry (leg lengths, angles, eta) the sweep forms the complex derivative D_theta = d ptrac / d theta without
// the upstream gradient; the loss gradient is Re(pbc D_theta) with pbc = conj(dL/dptrac), applied once the
// fine-grid adjoint of the whole path is known. Per energy: the F_s representatives of every scatterer (kept),
// the termination with its tangents against Pre_n = F_{n-1} ... F_1, then s = n-1 .. 1: K_s from Pre_s (rebuilt
// from the kept F) and U_s = T F_{n-1} ... F_{s+1}, the eta derivative, U_{s-1}, and the tangent stream of F_s.

// Bit 2 aa + bb set: some representative of an NR x NC block has Hankel columns (aa + 1, bb + 1).
template <int SET, int NR, int NC>
__host__ __device__ constexpr unsigned group_mask() {
    constexpr PairDef PD = pairdef(SET);
    unsigned m = 0;
    for (int k = 0; k < FM_NREP; ++k)
        if (PD.ri[k] < NR && PD.rj[k] < NC) m |= 1u << (2 * (PD.a[k] - 1) + PD.b[k] - 1);
    return m;
}

// One partial-wave row of a block's tangents contracted with K, accumulated as complex derivatives. y, yd: the
// gamtl leg's factors and their tangents without the row's t-matrix factor t, which scales the row's three sums once.
template <class T, int SET, int NR, int NC, class WRow>
__device__ __forceinline__ void bwd_row_c(const cx<T> (&K)[FM_NREP], const cx<T> (&xa)[3], const cx<T> (&xd)[3],
                                          const cx<T> (&y)[2], const cx<T> (&yd)[2], cx<T> t, const WRow &w,
                                          const WRow &wd, cx<T> &dX, cx<T> &dY, cx<T> &dB) {
    constexpr PairDef PD = pairdef(SET);
    constexpr unsigned used = group_mask<SET, NR, NC>();
    const cx<T> zero = mk<T>(T(0), T(0));
    cx<T> Q[3][2], Qd[3][2];
#pragma unroll
    for (int aa = 0; aa < 3; ++aa) Q[aa][0] = Q[aa][1] = Qd[aa][0] = Qd[aa][1] = zero;
#pragma unroll
    for (int k = 0; k < FM_NREP; ++k)
        if (PD.ri[k] < NR && PD.rj[k] < NC) {
            const int aa = PD.a[k] - 1, bb = PD.b[k] - 1;
            rfma(Q[aa][bb], K[k], w[k]);
            rfma(Qd[aa][bb], K[k], wd[k]);
        }
    // dB: sum Qd[aa][bb] xa[aa] y[bb]; dY: sum Q[aa][bb] xa[aa] yd[bb]; dX: sum Q[aa][bb] xd[aa] y[bb]
    cx<T> sB = zero, sY = zero, sX = zero;
#pragma unroll
    for (int bb = 0; bb < 2; ++bb) {
        if (!((used >> bb) & 0x15u)) continue;
        cx<T> rd = zero, r = zero, v = zero;
#pragma unroll
        for (int aa = 0; aa < 3; ++aa)
            if ((used >> (2 * aa + bb)) & 1u) {
                cfma4(rd, Qd[aa][bb], xa[aa]);
                cfma4(r, Q[aa][bb], xa[aa]);
                cfma4(v, Q[aa][bb], xd[aa]);
            }
        cfma4(sB, y[bb], rd);
        cfma4(sY, yd[bb], r);
        cfma4(sX, y[bb], v);
    }
    cfma4(dB, t, sB);
    cfma4(dY, t, sY);
    cfma4(dX, t, sX);
}

// The tangent stream of one scattering matrix (rows il = 1..ilx, t-matrix two rows ahead), complex derivatives.
template <class T, int SET, int NR, int NC>
__device__ __forceinline__ void fblock_bwd_c(const cx<T> (&K)[FM_NREP], cx<T> ck, T rX, T rY, int ilx,
                                             const T *__restrict__ Wl, const T *__restrict__ Wdl,
                                             const T *__restrict__ tl, int tlstride, cx<T> t1, cx<T> t2, cx<T> &dX,
                                             cx<T> &dY, cx<T> &dB) {
    dX = dY = dB = mk<T>(T(0), T(0));
    Hank3T<T> X;
    Hank2T<T> Y;
    X.init(minus_i_over(ck * rX), rX);
    Y.init(minus_i_over(ck * rY), rY);
    cx<T> xa[3], xd[3], ya[2], yd[2];
    // the t-matrix two rows ahead (rows past ilx repeat row ilx, which stays in the table); rows 1 and min(2, ilx)
    // come from the caller
    cx<T> t = t1;
    cx<T> tn = t2;
    // rows 1..4 unrolled (their seeds are explicit per row); a lane whose energy has fewer rows runs them with the
    // t-matrix factor zero, which adds nothing
#pragma unroll
    for (int r = 1; r <= 4; ++r) {
        X.seed_row(r, xa, xd);
        if (r < 4) Y.seed_row(r, ya, yd); else Y.step(r, ya, yd);
        const size_t rn = (size_t)(r + 2 < ilx ? r + 2 : ilx) * tlstride;
        const cx<T> tnn = mk<T>(tl[rn], tl[rn + 1]);
        // rows past ilx read the table's row ilx (inside the table, finite) and add zero
        const cx<T> tr = r <= ilx ? t : mk<T>(T(0), T(0));
        const size_t rw = (size_t)(r <= ilx ? r : ilx) * FM_NREP;
        bwd_row_c<T, SET, NR, NC>(K, xa, xd, ya, yd, tr, Wl + rw, Wdl + rw, dX, dY, dB);
        t = tn;
        tn = tnn;
    }
    // as in fblock: the row's table values read before the recursion steps that hide them, two rows per iteration
    auto row = [&](int r) {
        T wv[FM_NREP], wdv[FM_NREP];
        load_wrow<T, SET, NR, NC>(wv, Wl + (size_t)r * FM_NREP);
        load_wrow<T, SET, NR, NC>(wdv, Wdl + (size_t)r * FM_NREP);
        __builtin_amdgcn_sched_barrier(0);
        X.step(r, xa, xd);
        Y.step(r, ya, yd);
        const size_t rn = (size_t)(r + 2 < ilx ? r + 2 : ilx) * tlstride;
        const cx<T> tnn = mk<T>(tl[rn], tl[rn + 1]);
        bwd_row_c<T, SET, NR, NC>(K, xa, xd, ya, yd, t, wv, wdv, dX, dY, dB);
        t = tn;
        tn = tnn;
    };
    int r = 5;
#pragma unroll 1
    for (; r < ilx; r += 2) {
        row(r);
        row(r + 1);
    }
    if (r == ilx) row(r);
}

// Complex derivative slots of a lane, (kind, leg j) -> slot kind nleg + j - 1, in the workspace region of the lane's
// path: slot s of the lane's energy at byte lo + s sbytes (lo: the lane's (path, energy) offset in slot 0, sbytes: E
// complex).
template <class T> struct DSlots {
    __amdgpu_buffer_rsrc_t rs;
    uint32_t lo;
    int sbytes, nleg;
    __device__ __forceinline__ void put(int kind, int j, cx<T> x) const { buf_st(rs, lo, (kind * nleg + j - 1) * sbytes, x); }
};

// F(i, j) of a block from its representatives, with the row phase: ph[i] sg Ft[idx[i][j]]
template <class T, int SET>
__device__ __forceinline__ cx<T> fent(const cx<T> (&Ft)[FM_NREP], const cx<T> (&ph)[6], int i, int j) {
    constexpr PairDef PD = pairdef(SET);
    return ph[i] * Ft[PD.idx[i][j]] * T(PD.sg[i][j]);
}

// Pre_S = F_{S-1} ... F_1 (L x L0, with the row phases) from the kept representatives; S >= 2.
template <class T, int SET, int NL, int S>
__device__ __forceinline__ void prefix(const PathLDS &g, const cx<T> (&F)[NL - 1][FM_NREP],
                                       cx<T> (&P)[lamdef(SET).L][lamdef(SET).L0]) {
    constexpr int L = lamdef(SET).L, L0 = lamdef(SET).L0;
    cx<T> ph[6];
    row_phases<T, SET>(ph, g, 2);
#pragma unroll
    for (int i = 0; i < L; ++i)
#pragma unroll
        for (int b = 0; b < L0; ++b) P[i][b] = fent<T, SET>(F[0], ph, i, b);
#pragma unroll
    for (int t = 2; t < S; ++t) {   // Pre_{t+1} = F_t Pre_t
        row_phases<T, SET>(ph, g, t + 1);
#pragma unroll
        for (int b = 0; b < L0; ++b) {
            cx<T> col[L];
#pragma unroll
            for (int j = 0; j < L; ++j) col[j] = P[j][b];
#pragma unroll
            for (int i = 0; i < L; ++i) {
                cx<T> acc = mk<T>(T(0), T(0));
#pragma unroll
                for (int j = 0; j < L; ++j) cfma4(acc, fent<T, SET>(F[t - 1], ph, i, j), col[j]);
                P[i][b] = acc;
            }
        }
    }
}

// What a chain step needs of its scatterer S at the lane's energy: the partial waves ilx, the scatterer's t-matrix
// column and its rows 1 and min(2, ilx) (row 2 read whole from min(2, L1), used only when ilx >= 2).
template <class T> struct ScatIn {
    int ilx;
    const T *tl;
    cx<T> t1, t2;
};
template <class T>
__device__ __forceinline__ ScatIn<T> scat_in(const fm_edge &ed, const PathLDS &g, int ltab, int ee, int S) {
    const int E = ed.E, ip = g.ipot[S];
    const T *tl = TL<T>::tl(ed) + ((size_t)ip * ltab) * (2 * E) + 2 * ee;
    const size_t r2 = (size_t)(ed.L1 < 2 ? ed.L1 : 2) * (2 * E);
    return ScatIn<T>{ed.lmax[(size_t)ip * E + ee] + 1, tl, mk<T>(tl[2 * E], tl[2 * E + 1]), mk<T>(tl[r2], tl[r2 + 1])};
}
// Whether the sweep reads every scatterer's ScatIn when the energy starts (their loads in flight together, hidden by
// Pre_n) or at its chain step: early where the registers hold them without more spills (measured: K edges up to 4
// legs, L/M edges 3 legs).
template <int SET, int NL> constexpr bool scat_early() { return NL <= (SET == 0 ? 4 : 3); }

// One chain step of the fused sweep (scatterer S; NR x NC block): K and the eta derivative from Pre_S and U, then
// U <- U F_S (unless S = 1), then the tangent stream of F_S.
template <class T, int SET, int NL, int S>
__device__ __forceinline__ void fused_step(const fm_edge &ed, const PathLDS &g, const T *W, const T *Wd, int ltab, int ee,
                                           cx<T> ck, const ScatIn<T> (&sc)[NL - 1], const cx<T> (&F)[NL - 1][FM_NREP],
                                           cx<T> (&U)[lamdef(SET).L0][lamdef(SET).L], DSlots<T> ds) {
    constexpr LamDef D = lamdef(SET);
    constexpr PairDef PD = pairdef(SET);
    constexpr int L = D.L, L0 = D.L0;
    constexpr int NR = S == NL - 1 ? L0 : L, NC = S == 1 ? L0 : L;
    const int E = ed.E;
    cx<T> ph[6];
    row_phases<T, SET>(ph, g, S + 1);
    cx<T> K[FM_NREP];
    {
        cx<T> P[L][L0];
        if constexpr (S >= 2) prefix<T, SET, NL, S>(g, F, P);
#pragma unroll
        for (int k = 0; k < FM_NREP; ++k) K[k] = mk<T>(T(0), T(0));
        // K from Pre_S and U; the eta_{S+1} derivative -i sum_{l1,l2} M[l1] h(l1, l2) F_S(l1, l2) on the way
        cx<T> de = mk<T>(T(0), T(0));
#pragma unroll
        for (int l2 = 0; l2 < NC; ++l2) {
#pragma unroll
            for (int l1 = 0; l1 < NR; ++l1) {
                cx<T> m = mk<T>(T(0), T(0));
                if constexpr (S >= 2) {
#pragma unroll
                    for (int a = 0; a < L0; ++a) cfma4(m, P[l2][a], U[a][l1]);
                } else {
                    m = U[l2][l1];
                }
                const cx<T> h = m * ph[l1];
                if (PD.sg[l1][l2] > 0) K[PD.idx[l1][l2]] = K[PD.idx[l1][l2]] + h;
                else K[PD.idx[l1][l2]] = K[PD.idx[l1][l2]] - h;
                if (D.M[l1] != 0) cfma4(de, h * T(PD.sg[l1][l2] * D.M[l1]), F[S - 1][PD.idx[l1][l2]]);
            }
        }
        ds.put(GK_H, S + 1, mk<T>(de.i, -de.r));
        if constexpr (S > 1) {
#pragma unroll
            for (int a = 0; a < L0; ++a) {
                cx<T> row[6];
#pragma unroll
                for (int i = 0; i < 6; ++i) row[i] = (i < NR) ? U[a][i] * ph[i] : mk<T>(T(0), T(0));
#pragma unroll
                for (int j = 0; j < L; ++j) {
                    cx<T> v = mk<T>(T(0), T(0));
                    if (j < NC) {
#pragma unroll
                        for (int i = 0; i < NR; ++i) {
                            const cx<T> f = F[S - 1][PD.idx[i][j]];
                            if (PD.sg[i][j] > 0) cfma4(v, row[i], f); else cfma4(v, row[i], mk<T>(-f.r, -f.i));
                        }
                    }
                    U[a][j] = v;
                }
            }
        }
    }
    ScatIn<T> si;
    if constexpr (scat_early<SET, NL>()) si = sc[S - 1];
    else si = scat_in<T>(ed, g, ltab, ee, S);
    cx<T> dX, dY, dB;
    fblock_bwd_c<T, SET, NR, NC>(K, ck, (T)g.r[S + 1], (T)g.r[S], si.ilx, W + (size_t)(S - 1) * ltab * FM_NREP,
                                 Wd + (size_t)(S - 1) * ltab * FM_NREP, si.tl, 2 * E, si.t1,
                                 si.ilx >= 2 ? si.t2 : si.t1, dX, dY, dB);
    ds.put(GK_X, S + 1, dX);
    ds.put(GK_Y, S, dY);
    ds.put(GK_B, S, dB);
}

template <class T, int SET, int NL, int S>
__device__ __forceinline__ void fused_chain(const fm_edge &ed, const PathLDS &g, const T *W, const T *Wd, int ltab, int ee,
                                            cx<T> ck, const ScatIn<T> (&sc)[NL - 1], const cx<T> (&F)[NL - 1][FM_NREP],
                                            cx<T> (&U)[lamdef(SET).L0][lamdef(SET).L], DSlots<T> ds) {
    fused_step<T, SET, NL, S>(ed, g, W, Wd, ltab, ee, ck, sc, F, U, ds);
    if constexpr (S > 1) fused_chain<T, SET, NL, S - 1>(ed, g, W, Wd, ltab, ee, ck, sc, F, U, ds);
}

template <class T, int SET, int NL, int S>
__device__ __forceinline__ void fused_fblocks(const fm_edge &ed, const PathLDS &g, const T *W, int ltab, int ee,
                                              cx<T> (&F)[NL - 1][FM_NREP]) {
    constexpr int L = lamdef(SET).L, L0 = lamdef(SET).L0;
    const int E = ed.E;
    const cx<T> ck = mk<T>((T)ed.ck[2 * ee], (T)ed.ck[2 * ee + 1]);
    const int ip = g.ipot[S];
    const int ilx = ed.lmax[(size_t)ip * E + ee] + 1;
    const T *tl = TL<T>::tl(ed) + ((size_t)ip * ltab) * (2 * E) + 2 * ee;
    const T *Wl = W + (size_t)(S - 1) * ltab * FM_NREP;
    if constexpr (S == 1) fblock<T, SET, L, L0, true>(F[0], ck, (T)g.r[2], (T)g.r[1], ilx, Wl, tl, 2 * E);
    else if constexpr (S == NL - 1) fblock<T, SET, L0, L, true>(F[S - 1], ck, (T)g.r[S + 1], (T)g.r[S], ilx, Wl, tl, 2 * E);
    else fblock<T, SET, L, L, true>(F[S - 1], ck, (T)g.r[S + 1], (T)g.r[S], ilx, Wl, tl, 2 * E);
    if constexpr (S < NL - 1) fused_fblocks<T, SET, NL, S + 1>(ed, g, W, ltab, ee, F);
}

// The kept matrices of scatterers S..NL-1 from the lane's rows (the first scatterer's block is L x L0, the last's
// L0 x L, the others L x L).
template <class T, int SET, int NL, int S>
__device__ __forceinline__ void keep_loads(const T *kp, int E, cx<T> (&F)[NL - 1][FM_NREP]) {
    constexpr int L = lamdef(SET).L, L0 = lamdef(SET).L0;
    const T *b = kp + (size_t)2 * E * keep_off<SET>(NL, S);
    if constexpr (S == 1) keep_load<T, SET, L, L0>(b, E, F[S - 1]);
    else if constexpr (S == NL - 1) keep_load<T, SET, L0, L>(b, E, F[S - 1]);
    else keep_load<T, SET, L, L>(b, E, F[S - 1]);
    if constexpr (S < NL - 1) keep_loads<T, SET, NL, S + 1>(kp, E, F);
}

// The fused sweep at energy ee from the scattering matrices F of every scatterer (kept by the forward or rebuilt):
// returns cchi and stores the complex derivatives of ptrac in ds.
template <class T, int SET, int NL>
__device__ __forceinline__ cx<T> fused_energy(const fm_edge &ed, const PathLDS &g, const T *W, const T *Wd, int ltab,
                                              int ee, DSlots<T> ds, const cx<T> (&F)[NL - 1][FM_NREP]) {
    constexpr LamDef D = lamdef(SET);
    constexpr PairDef PD = pairdef(SET);
    constexpr int L = D.L, L0 = D.L0;
    const int E = ed.E;
    // read ahead of Pre_n, whose work hides their latency: the termination's channel weights of rows 1..4 (a row
    // without a channel reads channel 0's and does not use it), ck, and (scat_early) every scatterer's ScatIn
    const cx<T> ck = mk<T>((T)ed.ck[2 * ee], (T)ed.ck[2 * ee + 1]);
    ScatIn<T> sc[NL - 1];
    if constexpr (scat_early<SET, NL>()) {
#pragma unroll
        for (int S = 1; S < NL; ++S) sc[S - 1] = scat_in<T>(ed, g, ltab, ee, S);
    }
    const T *chw = TL<T>::chw(ed);
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
    // Pre_n (L0 x L0)
    cx<T> Pn[L0][L0];
    {
        cx<T> P[L][L0];
        prefix<T, SET, NL, NL - 1>(g, F, P);
        cx<T> ph[6];
        row_phases<T, SET>(ph, g, NL);
#pragma unroll
        for (int i = 0; i < L0; ++i)
#pragma unroll
            for (int b = 0; b < L0; ++b) {
                cx<T> acc = mk<T>(T(0), T(0));
#pragma unroll
                for (int j = 0; j < L; ++j) cfma4(acc, fent<T, SET>(F[NL - 2], ph, i, j), P[j][b]);
                Pn[i][b] = acc;
            }
    }
    // termination: T with its tangents against K_T[a][b] = Pn[b][a] ph1[a]
    cx<T> U[L0][L];
    cx<T> ptrac = mk<T>(T(0), T(0));
    {
        cx<T> ph[6];
        row_phases<T, SET>(ph, g, 1);
        cx<T> K[FM_NREP];
#pragma unroll
        for (int k = 0; k < FM_NREP; ++k) K[k] = mk<T>(T(0), T(0));
#pragma unroll
        for (int a = 0; a < L0; ++a)
#pragma unroll
            for (int b = 0; b < L0; ++b) {
                const cx<T> h = Pn[b][a] * ph[a];
                if (PD.sg[a][b] > 0) K[PD.idx[a][b]] = K[PD.idx[a][b]] + h;
                else K[PD.idx[a][b]] = K[PD.idx[a][b]] - h;
            }
        const int *lm = ed.lmax + ee;
        const int lx1 = max(lm[(size_t)g.ipot[0] * E], lm[(size_t)g.ipot[1] * E]) + 1;
        const int lxn = max(lm[(size_t)g.ipot[NL - 1] * E], lm[(size_t)g.ipot[NL] * E]) + 1;
        Hank3T<T> X;
        Hank2T<T> Y;
        X.init(minus_i_over(ck * (T)g.r[1]), (T)g.r[1]);
        Y.init(minus_i_over(ck * (T)g.r[NL]), (T)g.r[NL]);
        cx<T> Tt[FM_NREP];
#pragma unroll
        for (int k = 0; k < FM_NREP; ++k) Tt[k] = mk<T>(T(0), T(0));
        const T *Wn = W + (size_t)(NL - 1) * ltab * FM_NREP;
        const T *Wdn = Wd + (size_t)(NL - 1) * ltab * FM_NREP;
        cx<T> dX = mk<T>(T(0), T(0)), dY = dX, dB = dX;
        cx<T> xa[3], xd[3], ya[2], yd[2];
#pragma unroll
        for (int r = 1; r <= 4; ++r) {
            X.seed_row(r, xa, xd, r <= lx1);
            if (r < 4) Y.seed_row(r, ya, yd, r <= lxn); else Y.step(r, ya, yd, r <= lxn);
            if (chan[r - 1] >= 0) {
                const cx<T> w = wraw[r - 1] * T(2 * r - 1);
                accum_row<T, SET, L0, L0>(Tt, xa, ya[0] * w, ya[1] * w, Wn + (size_t)r * FM_NREP);
                bwd_row_c<T, SET, L0, L0>(K, xa, xd, ya, yd, w, Wn + (size_t)r * FM_NREP, Wdn + (size_t)r * FM_NREP,
                                          dX, dY, dB);
            }
        }
        ds.put(GK_X, 1, dX);
        ds.put(GK_Y, NL, dY);
        ds.put(GK_B, NL, dB);
        cx<T> de = mk<T>(T(0), T(0));
#pragma unroll
        for (int a = 0; a < L0; ++a)
#pragma unroll
            for (int b = 0; b < L; ++b) {
                const cx<T> t = (b < L0) ? ph[a] * Tt[PD.idx[a][b]] * T(PD.sg[a][b]) : mk<T>(T(0), T(0));
                U[a][b] = t;
                if (b < L0) {
                    const cx<T> tp = t * Pn[b][a];
                    ptrac = ptrac + tp;
                    if (D.M[a] != 0) rfma(de, tp, T(D.M[a]));
                }
            }
        ds.put(GK_H, 1, mk<T>(de.i, -de.r));
    }
    fused_chain<T, SET, NL, NL - 1>(ed, g, W, Wd, ltab, ee, ck, sc, F, U, ds);
    return ptrac * cfac_at<T>(ed, g, ee, NL);
}

// The geometry adjoint of one path (one wavefront): for direction lane < ndir (atom lane / 3, plus 1 for a path list
// whose absorber is atom 0; component lane % 3), the sum over the legs J of gp . d(r_J, beta_J, gamma_J, alpha_{J+1})
// / d(direction) by dual numbers through the leg geometry, in 1/A. The (direction, leg) terms are spread over the
// lanes (legs that do not touch the atom contribute 0 and are skipped) and summed over J in order; scr holds 3 n^2
// doubles.
__device__ __forceinline__ double geometry_adjoint(const PathLDS &g, int n, int mode, const double *gp, double *scr) {
    const int lane = lane_id();
    const int ndir = mode == 0 ? 3 * n : 3 * (n - 1);
    for (int t = lane; t < ndir * n; t += FM_WAVE) {
        const int dir = t / n, J = t - dir * n + 1;
        const int atom = mode == 0 ? dir / 3 : dir / 3 + 1, comp = dir % 3;
        const int i = J == n ? 0 : J, ip1 = J == n ? 1 : J + 1, im1 = J - 1;
        const int ip1a = ip1 == n ? 0 : ip1;
        double v = 0.0;
        if (atom == im1 || atom == i || atom == ip1a) {
            Dd pm[3], pi[3], pp[3];
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                pm[c] = dd(g.rat[im1][c], (im1 == atom && c == comp) ? 1.0 : 0.0);
                pi[c] = dd(g.rat[i][c], (i == atom && c == comp) ? 1.0 : 0.0);
                pp[c] = dd(g.rat[ip1][c], (ip1a == atom && c == comp) ? 1.0 : 0.0);
            }
            const LegGeoD lg = leg_geometry_dual(pm, pi, pp);
            const int next = J == n ? 1 : J + 1;
            v = gp[J - 1] * lg.r.d + gp[n + J - 1] * lg.beta.d + gp[2 * n + J - 1] * lg.gamma.d +
                gp[2 * n + next - 1] * lg.alpha.d;
        }
        scr[t] = v;
    }
    wave_sync();
    double acc = 0.0;
    if (lane < ndir)
        for (int J = 0; J < n; ++J) acc += scr[lane * n + J];
    wave_sync();
    return acc / FM_BOHR;
}

// deterministic wave sum (butterfly)
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, FM_WAVE);
    return v;
}

// amp / unwrapped phase of cc[0..E) into amp[], ang[] (the forward epilogue's first half)
template <class T>
__device__ __forceinline__ void amp_phase(int E, const cx<T> *cc, T *amp, double *ang) {
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

// Adjoint of the fine grid of one path (one wave): chi(q) = A(q) sin(Phi(q) + 2 q reff) exp(-2 sigma2 q^2), A and
// Phi linear in amp[] and the unwrapped ang[]. From dL/dchi(q) = gup_scale gup_d[q] + gup_p[q] to dL/dcchi per
// energy: each energy gathers the q points that interpolate from it (qorder / qstart runs, in a fixed order).
// With qs (3 nk doubles of LDS), dL/dA(q), dL/dPhi(q) and the fraction are first evaluated once per q point with
// the lanes over the points (their loads independent), then gathered from LDS; without it, each energy evaluates
// them where it uses them. Returns dL/dreff (A), wave-summed.
template <class T>
__device__ __forceinline__ double fine_grid_adjoint(int E, const int32_t *jq, const double *fq, const double *q, const int32_t *qorder,
                                    const int32_t *qstart, double reff_A, double sigma2, const double *gup_d,
                                    double gup_scale, const T *gup_p, const cx<T> *cc, const T *amp, const double *ang,
                                    cx<T> *cb, int nk = 0, double *qs = nullptr, const double *S = nullptr,
                                    int sstride = 0) {
    const int lane = lane_id();
    // dL/dA and dL/dPhi at q point i from its interpolation (j, f), q and the upstream gradient gu
    auto eval = [&](int j, double f, double qi, double gu, double &gA, double &gP) {
        const double A = (1.0 - f) * (double)amp[j] + f * (double)amp[j + 1];
        double ph = (1.0 - f) * ang[j] + f * ang[j + 1] + 2.0 * qi * reff_A;
        ph -= FM_TWO_PI * rint(ph / FM_TWO_PI);
        if (sigma2 != 0.0) gu *= exp(-2.0 * sigma2 * qi * qi);
        T sn, cs;
        if constexpr (sizeof(T) == 4) sincosf((float)ph, &sn, &cs); else sincos(ph, &sn, &cs);
        gA = (double)(T)(gu * (double)sn);
        gP = (double)(T)(gu * A * (double)cs);
    };
    auto gup = [&](int i) -> double {
        double gu = 0.0;
        if (gup_d) gu += gup_scale * gup_d[i];
        if (gup_p) gu += (double)gup_p[i];
        return gu;
    };
    auto adj = [&](int i, double &gA, double &gP) { eval(jq[i], fq[i], q[i], gup(i), gA, gP); };
    double rb = 0.0;
    // dL/dcchi at energy e from dL/dA and dL/dPhi there (through amp = |cchi| and the phase)
    auto finish = [&](int e, double ab, double pb) {
        const double cr = (double)cc[e].r, ci = (double)cc[e].i;
        const double c2 = cr * cr + ci * ci, a = (double)amp[e];
        cx<T> v = mk<T>(T(0), T(0));
        if (c2 > 0.0) v = mk<T>((T)(ab * cr / a - pb * ci / c2), (T)(ab * ci / a + pb * cr / c2));
        cb[e] = v;
    };
    // staging: dL/dA and dL/dPhi per q point (values already rounded to T, so stored as T), then the fractions
    T *const sA = qs ? reinterpret_cast<T *>(qs) : nullptr;
    T *const sP = qs ? sA + nk : nullptr;
    double *const sF = qs ? reinterpret_cast<double *>(sP + nk) : nullptr;
    if (S) {   // dense interpolation (qs required): dL/dA, dL/dPhi per q point, then S^T per energy
        for (int i = lane; i < nk; i += FM_WAVE) {
            const double *Si = S + (size_t)i * sstride;
            double sa = 0.0, sp = 0.0;
            for (int e = 0; e < E; ++e) {
                sa += Si[e] * (double)amp[e];
                sp += Si[e] * ang[e];
            }
            double ph = sp + 2.0 * q[i] * reff_A;
            ph -= FM_TWO_PI * rint(ph / FM_TWO_PI);
            double gu = gup(i);
            if (sigma2 != 0.0) gu *= exp(-2.0 * sigma2 * q[i] * q[i]);
            T sn, cs;
            if constexpr (sizeof(T) == 4) sincosf((float)ph, &sn, &cs); else sincos(ph, &sn, &cs);
            const double gP = (double)(T)(gu * sa * (double)cs);
            sA[i] = (T)(gu * (double)sn);
            sP[i] = (T)gP;
            rb += 2.0 * q[i] * gP;
        }
        wave_sync();
        for (int e = lane; e < E; e += FM_WAVE) {
            double ab = 0.0, pb = 0.0;
            for (int i = 0; i < nk; ++i) {
                const double s = S[(size_t)i * sstride + e];
                ab += s * (double)sA[i];
                pb += s * (double)sP[i];
            }
            finish(e, ab, pb);
        }
        return wave_sum(rb);
    }
    if (qs) {
        const int t1 = qstart[E];
        for (int t = lane; t < t1; t += FM_WAVE) {
            const int i = qorder[t];
            double gA, gP;
            adj(i, gA, gP);
            sA[t] = (T)gA;
            sP[t] = (T)gP;
            sF[t] = fq[i];
            rb += 2.0 * q[i] * gP;
        }
        wave_sync();
        for (int e = lane; e < E; e += FM_WAVE) {
            double ab = 0.0, pb = 0.0;
            const int lo0 = qstart[e], lo1 = qstart[e + 1], hi0 = e > 0 ? qstart[e - 1] : lo0;
            for (int t = lo0; t < lo1; ++t) {
                ab += (1.0 - sF[t]) * (double)sA[t];
                pb += (1.0 - sF[t]) * (double)sP[t];
            }
            for (int t = hi0; t < lo0; ++t) {
                ab += sF[t] * (double)sA[t];
                pb += sF[t] * (double)sP[t];
            }
            finish(e, ab, pb);
        }
        return wave_sum(rb);
    }
    for (int e = lane; e < E; e += FM_WAVE) {
        double ab = 0.0, pb = 0.0;
        for (int t = qstart[e]; t < qstart[e + 1]; ++t) {    // q points with lower interpolation energy e
            const int i = qorder[t];
            double gA, gP;
            adj(i, gA, gP);
            ab += (1.0 - fq[i]) * gA;
            pb += (1.0 - fq[i]) * gP;
            rb += 2.0 * q[i] * gP;
        }
        if (e > 0)
            for (int t = qstart[e - 1]; t < qstart[e]; ++t) {  // ... and upper interpolation energy e
                const int i = qorder[t];
                double gA, gP;
                adj(i, gA, gP);
                ab += fq[i] * gA;
                pb += fq[i] * gP;
            }
        finish(e, ab, pb);
    }
    return wave_sum(rb);
}

template <class T, int SET, int WPE>
__global__ __launch_bounds__(FM_WAVE * 4) __attribute__((amdgpu_waves_per_eu(WPE))) void k_backward(BArgs a) {
    extern __shared__ __align__(16) unsigned char smem[];
    __shared__ int pidx[8], pval[8], tslot;
    __shared__ double rbar[8];
    const int wid = __builtin_amdgcn_readfirstlane(threadIdx.x / FM_WAVE);
    const int lane = lane_id();
    const fm_edge &ed = a.ed;
    const int G = a.gp, cap = FM_WAVE / G;
    const int ltab = ed.L1 + 1, E = ed.E;
    const int n = a.nlegmax;
    const int nband = a.bp.nband;
    const uint32_t loc = (uint32_t)lane * sizeof(cx<T>), lor = (uint32_t)lane * sizeof(T);
    // this lane's path and energy within a band, and that path's LDS
    const int jl = lane / cap, il = lane % cap;
    unsigned char *pbase = smem + (size_t)jl * a.path_bytes;
    const PathLDS &gj = *reinterpret_cast<const PathLDS *>(pbase);
    const T *Wj = reinterpret_cast<const T *>(pbase + a.off_W);
    const T *Wdj = reinterpret_cast<const T *>(pbase + a.off_Wd);
    cx<T> *ccj = reinterpret_cast<cx<T> *>(pbase + a.off_cc);
    const cx<T> *cbj = reinterpret_cast<const cx<T> *>(pbase + a.off_cb);
    unsigned char *const wg_work = a.work + (size_t)blockIdx.x * nband * a.work_per_wave;   // band b: + b * work_per_wave
    const size_t wpw = a.work_per_wave;

    for (int blk = first_group(a.tickets, tslot); blk * G < a.npath_bucket; blk = next_group(a.tickets, tslot, blk)) {
        // ---- 1. setup: one wavefront per path
        for (int j = wid; j < G; j += 4) {
            const int slot = blk * G + j;
            const bool valid = slot < a.npath_bucket;
            const int s0 = valid ? slot : blk * G;
            const int p = a.order ? a.order[a.first + s0] : a.first + s0;
            unsigned char *base = smem + (size_t)j * a.path_bytes;
            PathLDS &g = *reinterpret_cast<PathLDS *>(base);
            int ei;
            load_path(a.src, p, g, ei, n);
            if (lane == 0) {
                pidx[j] = p;
                pval[j] = valid ? 1 : 0;
                rbar[j] = (a.src.mode == 0 && a.greff) ? a.greff[p] : 0.0;
            }
        }
        build_tables<T, SET, true>(ed, smem, a.path_bytes, a.off_W, a.off_Wd, wid, 4, G, n, ltab);
        __syncthreads();
        // ---- 2. pass A over this wavefront's bands
        for (int b = 0; b < nband; ++b) {
            if (a.bp.wave[b] != wid) continue;
            const int e0 = a.bp.start[b], nb = a.bp.start[b + 1] - e0;
            const int e = e0 + (il < nb ? il : nb - 1);
            {
                unsigned char *wb = wg_work + (size_t)b * wpw;
                const PreStore<T, SET> pre{{slice_rsrc(wb, a.work_off_F), loc},
                                           {slice_rsrc(wb + a.work_off_F, a.work_off_acc - a.work_off_F), loc}};
                const cx<T> pt = pass_A<T, SET>(ed, gj, Wj, ltab, e, n, pre);
                const cx<T> c = pt * cfac_at<T>(ed, gj, e, n);
                if (il < nb) ccj[e] = c;
            }
        }
        __syncthreads();
        // ---- 3. dL/dcchi per energy (path lists: the fine grid's adjoint), one wavefront per path
        if (a.src.mode == 1)
            for (int j = wid; j < G; j += 4) {
                if (!pval[j]) continue;
                unsigned char *base = smem + (size_t)j * a.path_bytes;
                const PathLDS &g = *reinterpret_cast<const PathLDS *>(base);
                const cx<T> *cc = reinterpret_cast<const cx<T> *>(base + a.off_cc);
                T *amp = reinterpret_cast<T *>(base + a.off_amp);
                double *ang = reinterpret_cast<double *>(base + a.off_ang);
                cx<T> *cb = reinterpret_cast<cx<T> *>(base + a.off_cb);
                const int p = pidx[j];
                amp_phase<T>(E, cc, amp, ang);
                const double *gsr = a.gseg ? a.gseg + (size_t)a.segment[p] * a.nk : nullptr;
                const T *gpr = a.gpath ? reinterpret_cast<const T *>(a.gpath) + (size_t)p * a.nk : nullptr;
                const double rb = fine_grid_adjoint<T>(E, a.G.jq + (size_t)a.edge_id * a.nk, a.G.fq + (size_t)a.edge_id * a.nk,
                                                       a.G.q, a.G.qorder, a.G.qstart + (size_t)a.edge_id * a.G.qstride,
                                                       g.reff * FM_BOHR, a.src.P.sigma2 ? a.src.P.sigma2[p] : 0.0, gsr,
                                                       a.src.P.deg[p], gpr, cc, amp, ang, cb, a.nk,
                                                       a.off_qs ? reinterpret_cast<double *>(base + a.off_qs) : nullptr,
                                                       a.G.S ? a.G.S + (size_t)a.edge_id * a.nk * a.G.sstride : nullptr,
                                                       a.G.sstride);
                if (lane == 0) rbar[j] = rb;
            }
        __syncthreads();
        // ---- 4. pass B over this wavefront's bands
        for (int b = 0; b < nband; ++b) {
            if (a.bp.wave[b] != wid) continue;
            unsigned char *wb = wg_work + (size_t)b * wpw;
            const GSlots<T> gs{{slice_rsrc(wb + a.work_off_acc, a.work_per_wave - a.work_off_acc), lor}, n};
            const int e0 = a.bp.start[b], nb = a.bp.start[b + 1] - e0;
            const int e = e0 + (il < nb ? il : nb - 1);
            cx<T> cbar = mk<T>(T(0), T(0));
            const cx<T> cce = ccj[e];
            if (il < nb && pval[jl]) {
                if (a.src.mode == 0) {
                    const T *gin = reinterpret_cast<const T *>(a.gcchi) + ((size_t)pidx[jl] * E + e) * 2;
                    cbar = mk<T>(gin[0], gin[1]);
                } else {
                    cbar = cbj[e];
                }
            }
            {
                const PreStore<T, SET> pre{{slice_rsrc(wb, a.work_off_F), loc},
                                           {slice_rsrc(wb + a.work_off_F, a.work_off_acc - a.work_off_F), loc}};
                pass_B<T, SET, WPE>(ed, gj, Wj, Wdj, ltab, e, n, cbar, cce, pre, gs);
            }
        }
        __syncthreads();
        // ---- 5. per path: the gradient slots summed over the path's lanes of every band (fixed order), the geometry
        // adjoint, the output
        for (int j = wid; j < G; j += 4) {
            if (!pval[j]) continue;
            unsigned char *base = smem + (size_t)j * a.path_bytes;
            const PathLDS &g = *reinterpret_cast<const PathLDS *>(base);
            double *red = reinterpret_cast<double *>(base + a.off_red);
            double *gp = red + nacc(a.nlegmax);
            double *tmp = gp + 3 * FM_MAX_LEGS;
            const int nlm = n;
            const int nq = nacc(nlm);
            if (lane < nq) {
                double s = 0.0;
                for (int b = 0; b < nband; ++b) {
                    const int nb = a.bp.start[b + 1] - a.bp.start[b];
                    const T *src = reinterpret_cast<const T *>(wg_work + (size_t)b * wpw + a.work_off_acc) + (size_t)lane * FM_WAVE + j * cap;
                    for (int i0 = 0; i0 < nb; i0 += 8) {   // 8 loads in flight, summed in order
                        T d[8];
#pragma unroll
                        for (int u = 0; u < 8; ++u) d[u] = i0 + u < nb ? src[i0 + u] : T(0);
#pragma unroll
                        for (int u = 0; u < 8; ++u)
                            if (i0 + u < nb) s += (double)d[u];
                    }
                }
                red[lane] = s;
            }
            wave_sync();
            // per-leg gradients: r_j (bohr) from both leg factors, the cfac term and reff = sum r / 2; beta_j; eta_j
            if (lane < 3 * n) {
                const int kind = lane / n, jg = lane % n + 1;
                double v;
                if (kind == 0)
                    v = red[GK_X * nlm + jg - 1] + red[GK_Y * nlm + jg - 1] + red[4 * nlm] - red[4 * nlm + 1] / g.r[jg] +
                        0.5 * FM_BOHR * rbar[j];
                else if (kind == 1)
                    v = red[GK_B * nlm + jg - 1];
                else
                    v = red[GK_H * nlm + jg - 1];
                gp[lane] = v;
            }
            wave_sync();
            // geometry adjoint (dual numbers through the legs)
            const int ndir = a.src.mode == 0 ? 3 * n : 3 * (n - 1);
            double gout = geometry_adjoint(g, n, a.src.mode, gp, reinterpret_cast<double *>(base + a.off_gs));
            if (lane < ndir) {
                if (!g.ok) gout = __builtin_nan("");
                tmp[lane] = gout;
            }
            wave_sync();
            const int p = pidx[j];
            double *out = a.gcoord + (size_t)p * a.nslot * 3;
            if (a.src.mode == 0) {
                if (lane < 3 * n) out[lane] = gout;
            } else {
                if (lane < 3 * (n - 1)) out[3 + lane] = gout;
                if (lane < 3) {
                    double s = 0.0;
                    for (int k = 0; k < n - 1; ++k) s += tmp[3 * k + lane];
                    out[lane] = -s;
                }
                if (lane >= 3 * n && lane < 3 * a.nslot) out[lane] = 0.0;
            }
            wave_sync();
        }
        __syncthreads();
    }
    release_groups(a.tickets);
}

// The fused sweep of a single-scattering path (fm_ss.h) at energy ee: cchi, and the complex derivatives of ptrac
// along the common leg length (leg-1 slot; d r_1 = d r_2 for every displacement, the angles are pi for every
// position) and the two eta, with dual numbers in the leg length.
template <class T, int ILI>
__device__ __forceinline__ cx<T> ss_fused_energy(const fm_edge &ed, const PathLDS &g, int ee, DSlots<T> ds) {
    const int E = ed.E;
    const int lxa = max(ed.lmax[(size_t)g.ipot[0] * E + ee], ed.lmax[(size_t)g.ipot[1] * E + ee]) + 1;
    const cx<T> ck = mk<T>((T)ed.ck[2 * ee], (T)ed.ck[2 * ee + 1]);
    const T r = (T)g.r[1];
    const cx<T> z = minus_i_over(ck * r);
    dcx<T> Tk[SS_MAXR], Fk[SS_MAXR];
    ss_reps<T, ILI, dcx<T>>(ed, ee, g.ipot[1], dcx<T>{z, z * (T(-1) / r)}, lxa, Tk, Fk);
    const dcx<T> pt = ss_trace<T, ILI, dcx<T>>(g.eta, Tk, Fk);
    const cx<T> zero = mk<T>(T(0), T(0));
    ds.put(GK_X, 1, pt.d);
    ds.put(GK_X, 2, zero);
    ds.put(GK_Y, 1, zero);
    ds.put(GK_Y, 2, zero);
    ds.put(GK_B, 1, zero);
    ds.put(GK_B, 2, zero);
    cx<T> h1, h2;
    ss_trace_eta<T, ILI, dcx<T>>(g.eta, Tk, Fk, h1, h2);
    ds.put(GK_H, 1, mk<T>(h1.i, -h1.r));
    ds.put(GK_H, 2, mk<T>(h2.i, -h2.r));
    return pt.v * cfac_at<T>(ed, g, ee, 2);
}

// The fused adjoint of multiple-scattering paths of NL legs: per band of energies one sweep (fused_energy) that
// leaves cchi in LDS and the complex derivatives of ptrac in the band's workspace; then per path the fine-grid
// adjoint (or the given dL/dcchi), pbc and the cfac terms per energy; then the derivative slots contracted with pbc
// and summed over the energies in band order, and the geometry adjoint.
// The fused adjoint's sweep (one wavefront per SIMD): per block of G paths, the geometry and the rotation tables
// with their beta derivatives into LDS, then every energy of every path in one sweep. Each path leaves cchi and the
// 4 nleg complex derivative slots of every energy in its workspace region (path-major: cchi [E], slot s at
// [(1 + s) E]); k_adjoint_tail takes it from there. Lanes past their band's end and paths past the launch's end
// compute on a duplicate and write their slots to the spare region after the last path.
template <class T, int SET, int NL, int WPE>
__global__ __launch_bounds__(FM_WAVE * 4) __attribute__((amdgpu_waves_per_eu(WPE))) void k_adjoint(BArgs a) {
    extern __shared__ __align__(16) unsigned char smem[];
    __shared__ int tslot;
    const int wid = __builtin_amdgcn_readfirstlane(threadIdx.x / FM_WAVE);
    const int lane = lane_id();
    const fm_edge &ed = a.ed;
    const int G = a.gp, cap = FM_WAVE / G;
    const int ltab = ed.L1 + 1, E = ed.E;
    const int nband = a.bp.nband;
    const int jl = lane / cap, il = lane % cap;
    unsigned char *pbase = smem + (size_t)jl * a.path_bytes;
    const PathLDS &gj = *reinterpret_cast<const PathLDS *>(pbase);
    const T *Wj = reinterpret_cast<const T *>(pbase + a.off_W);
    const T *Wdj = reinterpret_cast<const T *>(pbase + a.off_Wd);
    const __amdgpu_buffer_rsrc_t drs = slice_rsrc(a.work, a.dbytes);
    cx<T> *const wcc = reinterpret_cast<cx<T> *>(a.work);
    const int sbytes = E * (int)sizeof(cx<T>);

    for (int blk = first_group(a.tickets, tslot); blk * G < a.npath_bucket; blk = next_group(a.tickets, tslot, blk)) {
        {   // ---- 1. setup: the wavefront's paths j = wid, wid + 4 at once, a 16-lane group each (G <= 8)
            const int j = wid + 4 * (lane >> 4);
            const bool act = j < G;
            const int slot = blk * G + j;
            const int s0 = act && slot < a.npath_bucket ? slot : blk * G;
            const int p = a.order ? a.order[a.first + s0] : a.first + s0;
            PathLDS &g = *reinterpret_cast<PathLDS *>(smem + (size_t)(act ? j : 0) * a.path_bytes);
            int ei;
            load_path_group(a.src, p, g, ei, NL, lane & 15, act, SET < FM_SET_SS);
        }
        if constexpr (SET < FM_SET_SS) build_tables<T, SET, true, 16>(ed, smem, a.path_bytes, a.off_W, a.off_Wd, wid, 4, G, NL, ltab);
        __syncthreads();
        // ---- 2. the fused sweep over this wavefront's bands
        const int sj = blk * G + jl;
        const bool vj = sj < a.npath_bucket;
        // this lane's energy in band b, and its path's (a duplicate in the launch for lanes past its end) kept rows
        auto energy_of = [&](int b) {
            const int e0 = a.bp.start[b], nb = a.bp.start[b + 1] - e0;
            return e0 + (il < nb ? il : nb - 1);
        };
        auto kept_at = [&](int e) {
            return reinterpret_cast<const T *>(a.keep) + ((size_t)(vj ? sj : blk * G) * keep_rows<SET>(NL) * E + e) * 2;
        };
        auto next_band = [&](int b) {
            for (++b; b < nband && a.bp.wave[b] != wid; ++b) {}
            return b;
        };
        for (int b = next_band(-1); b < nband; b = next_band(b)) {
            const int e = energy_of(b);
            const bool on = vj && il < a.bp.start[b + 1] - a.bp.start[b];
            const size_t reg = on ? (size_t)sj : (size_t)a.npath_bucket;
            const DSlots<T> ds{drs, (uint32_t)((reg * a.dstride + E + e) * sizeof(cx<T>)), sbytes, NL};
            cx<T> c;
            if constexpr (SET >= FM_SET_SS) {
                c = ss_fused_energy<T, SET - FM_SET_SS + 1>(ed, gj, e, ds);
            } else {
                cx<T> F[NL - 1][FM_NREP];
                if (a.keep) {
                    keep_loads<T, SET, NL, 1>(kept_at(e), E, F);
                } else {
                    fused_fblocks<T, SET, NL, 1>(ed, gj, Wj, ltab, e, F);
                }
                c = fused_energy<T, SET, NL>(ed, gj, Wj, Wdj, ltab, e, ds, F);
            }
            if (on) wcc[(size_t)sj * a.dstride + e] = c;
        }
        __syncthreads();
    }
    release_groups(a.tickets);
}

// The per-path tail of the fused adjoint, one wavefront per path, from the workspace regions k_adjoint left:
// dL/dcchi (the fine-grid adjoint, or the given gcchi), pbc = conj(dL/dptrac) and the cfac terms, Re(pbc D) per
// slot summed over the energies (FP64, a fixed butterfly), the geometry adjoint and the output. Few registers and a
// few KB of LDS per wavefront, so that several wavefronts share a SIMD and hide the latencies this per-path work
// consists of. No workgroup barrier: a wavefront without a path returns.
template <class T, int NL>
__global__ __launch_bounds__(FM_WAVE * 4) void k_adjoint_tail(BArgs a) {
    extern __shared__ __align__(16) unsigned char smem[];
    const int wid = threadIdx.x / FM_WAVE, lane = lane_id();
    const int s0 = blockIdx.x * 4 + wid;
    if (s0 >= a.npath_bucket) return;
    const fm_edge &ed = a.ed;
    const int E = ed.E;
    constexpr int n = NL, NS = 4 * NL;
    unsigned char *base = smem + (size_t)wid * a.path_bytes;
    PathLDS &g = *reinterpret_cast<PathLDS *>(base);
    cx<T> *cc = reinterpret_cast<cx<T> *>(base + a.off_cc);
    T *amp = reinterpret_cast<T *>(base + a.off_amp);
    double *ang = reinterpret_cast<double *>(base + a.off_ang);
    cx<T> *cb = reinterpret_cast<cx<T> *>(base + a.off_cb);
    double *red = reinterpret_cast<double *>(base + a.off_red);
    double *gp = red + nacc(NL);
    double *tmp = gp + 3 * FM_MAX_LEGS;
    const int p = a.order ? a.order[a.first + s0] : a.first + s0;
    int ei;
    load_path(a.src, p, g, ei, NL, false);
    const cx<T> *wp = reinterpret_cast<const cx<T> *>(a.work) + (size_t)s0 * a.dstride;
    for (int e = lane; e < E; e += FM_WAVE) cc[e] = wp[e];
    wave_sync();
    double rb = 0.0;
    if (a.src.mode == 1) {
        amp_phase<T>(E, cc, amp, ang);
        const double *gsr = a.gseg ? a.gseg + (size_t)a.segment[p] * a.nk : nullptr;
        const T *gpr = a.gpath ? reinterpret_cast<const T *>(a.gpath) + (size_t)p * a.nk : nullptr;
        rb = fine_grid_adjoint<T>(E, a.G.jq + (size_t)a.edge_id * a.nk, a.G.fq + (size_t)a.edge_id * a.nk, a.G.q,
                                  a.G.qorder, a.G.qstart + (size_t)a.edge_id * a.G.qstride, g.reff * FM_BOHR,
                                  a.src.P.sigma2 ? a.src.P.sigma2[p] : 0.0, gsr, a.src.P.deg[p], gpr, cc, amp, ang, cb,
                                  a.nk, a.off_qs ? reinterpret_cast<double *>(base + a.off_qs) : nullptr,
                                  a.G.S ? a.G.S + (size_t)a.edge_id * a.nk * a.G.sstride : nullptr, a.G.sstride);
    } else {
        const T *gin = reinterpret_cast<const T *>(a.gcchi) + (size_t)p * E * 2;
        for (int e = lane; e < E; e += FM_WAVE) cb[e] = mk<T>(gin[2 * e], gin[2 * e + 1]);
        rb = a.greff ? a.greff[p] : 0.0;
    }
    wave_sync();
    // pbc and the two cfac terms: d cfac / d r_j = a1 - a2 / r_j with a1 = sum -Im(w (ck - xk)), a2 = sum Re w
    double s1 = 0.0, s2 = 0.0;
    for (int e = lane; e < E; e += FM_WAVE) {
        const cx<T> cbar = cb[e];
        const cx<T> w = conj(cbar) * cc[e];
        const cx<T> wdk = w * mk<T>((T)ed.ckmxk[2 * e], (T)ed.ckmxk[2 * e + 1]);
        s1 += (double)(-wdk.i);
        s2 += (double)w.r;
        cb[e] = conj(cbar * conj(cfac_at<T>(ed, g, e, NL)));
    }
    // Re(pbc D) per slot, the lanes over the energies with all the slots' loads in flight
    double acc[NS];
#pragma unroll
    for (int k = 0; k < NS; ++k) acc[k] = 0.0;
    for (int e = lane; e < E; e += FM_WAVE) {
        const cx<T> q = cb[e];
        cx<T> d[NS];
#pragma unroll
        for (int k = 0; k < NS; ++k) d[k] = wp[(size_t)(1 + k) * E + e];
#pragma unroll
        for (int k = 0; k < NS; ++k) acc[k] += (double)(q.r * d[k].r - q.i * d[k].i);
    }
#pragma unroll
    for (int k = 0; k < NS; ++k) acc[k] = wave_sum(acc[k]);
    s1 = wave_sum(s1);
    s2 = wave_sum(s2);
    if (lane == 0) {
#pragma unroll
        for (int k = 0; k < NS; ++k) red[k] = acc[k];
        red[NS] = s1;
        red[NS + 1] = s2;
    }
    wave_sync();
    if (lane < 3 * n) {
        const int kind = lane / n, jg = lane % n + 1;
        double v;
        if (kind == 0)
            v = red[GK_X * n + jg - 1] + red[GK_Y * n + jg - 1] + red[NS] - red[NS + 1] / g.r[jg] + 0.5 * FM_BOHR * rb;
        else if (kind == 1)
            v = red[GK_B * n + jg - 1];
        else
            v = red[GK_H * n + jg - 1];
        gp[lane] = v;
    }
    wave_sync();
    const int ndir = a.src.mode == 0 ? 3 * n : 3 * (n - 1);
    double gout = geometry_adjoint(g, n, a.src.mode, gp, reinterpret_cast<double *>(base + a.off_gs));
    if (lane < ndir) {
        if (!g.ok) gout = __builtin_nan("");
        tmp[lane] = gout;
    }
    wave_sync();
    double *out = a.gcoord + (size_t)p * a.nslot * 3;
    if (a.src.mode == 0) {
        if (lane < 3 * n) out[lane] = gout;
    } else {
        if (lane < 3 * (n - 1)) out[3 + lane] = gout;
        if (lane < 3) {
            double sum = 0.0;
            for (int k = 0; k < n - 1; ++k) sum += tmp[3 * k + lane];
            out[lane] = -sum;
        }
        if (lane >= 3 * n && lane < 3 * a.nslot) out[lane] = 0.0;
    }
}

// grad_pos[atom] = sum of gcoord over the atom's references: one wavefront per atom (an absorber can appear in
// thousands of paths), lane l summing references l, l + 64, ... in list order, the 64 partial sums of each
// component then added in a fixed butterfly. Four atoms per 256-thread block; a wavefront past natom returns whole.
__global__ __launch_bounds__(FM_WAVE * 4) void k_atom_grad(int natom, const int32_t *ptr, const int32_t *refs,
                                                           const double *gcoord, double *grad) {
    const int atom = blockIdx.x * 4 + (int)(threadIdx.x / FM_WAVE);
    if (atom >= natom) return;
    const int lane = lane_id();
    double s0 = 0.0, s1 = 0.0, s2 = 0.0;
    for (int k = ptr[atom] + lane; k < ptr[atom + 1]; k += FM_WAVE) {
        const double *g = gcoord + (size_t)refs[k] * 3;
        s0 += g[0];
        s1 += g[1];
        s2 += g[2];
    }
    s0 = wave_sum(s0);
    s1 = wave_sum(s1);
    s2 = wave_sum(s2);
    if (lane == 0) {
        grad[(size_t)atom * 3] = s0;
        grad[(size_t)atom * 3 + 1] = s1;
        grad[(size_t)atom * 3 + 2] = s2;
    }
}

// GPUPathStage.chi_fine's adjoint on its own: one wave per path.
template <class T>
__global__ __launch_bounds__(FM_WAVE *WPB_MAX) void k_chi_fine_bwd(fm_edge ed, int npath, const T *cchi, const double *reff,
                                                                    fm_grid G, const T *gchi, T *gcchi, double *greff,
                                                                    size_t wave_bytes, int staged) {
    extern __shared__ __align__(16) unsigned char smem[];
    const int wid = threadIdx.x / FM_WAVE, lane = lane_id();
    const int wpb = blockDim.x / FM_WAVE;
    const int E = ed.E, nk = G.nk;
    unsigned char *base = smem + (size_t)wid * wave_bytes;
    cx<T> *cc = reinterpret_cast<cx<T> *>(base);
    T *amp = reinterpret_cast<T *>(base + align16(sizeof(cx<T>) * E));
    double *ang = reinterpret_cast<double *>(base + align16(sizeof(cx<T>) * E) + align16(sizeof(T) * E));
    cx<T> *cb = reinterpret_cast<cx<T> *>(reinterpret_cast<unsigned char *>(ang) + align16(sizeof(double) * E));
    double *qs = staged ? reinterpret_cast<double *>(reinterpret_cast<unsigned char *>(cb) + align16(sizeof(cx<T>) * E))
                        : nullptr;
    for (int p = blockIdx.x * wpb + wid; p < npath; p += gridDim.x * wpb) {
        for (int e = lane; e < E; e += FM_WAVE) cc[e] = mk<T>(cchi[((size_t)p * E + e) * 2], cchi[((size_t)p * E + e) * 2 + 1]);
        wave_sync();
        amp_phase<T>(E, cc, amp, ang);
        const double rb = fine_grid_adjoint<T>(E, G.jq, G.fq, G.q, G.qorder, G.qstart, reff[p], 0.0, nullptr, 0.0,
                                               gchi + (size_t)p * nk, cc, amp, ang, cb, nk, qs, G.S, G.sstride);
        wave_sync();
        for (int e = lane; e < E; e += FM_WAVE) {
            gcchi[((size_t)p * E + e) * 2] = cb[e].r;
            gcchi[((size_t)p * E + e) * 2 + 1] = cb[e].i;
        }
        if (lane == 0) greff[p] = rb;
        wave_sync();
    }
}

// Paths per workgroup: the largest plan within the LDS budget (by default one workgroup per CU, whose 4 wavefronts
// then have the whole register file: 1 wavefront per SIMD); a plan small enough for two workgroups per CU runs the
// variant sized for 2 wavefronts per SIMD (FP32).
// groups: the launch's path groups; grid, tickets: the bucket's policy; cap: the most workgroups the workspace holds.
template <class T, int SET, int WPE>
hipError_t launch_k(BArgs a, long groups, long cap, const fm_launch &pol, size_t shm, hipStream_t st) {
    const void *fn = (const void *)k_backward<T, SET, WPE>;
    if (shm > 64 * 1024) {
        hipError_t e = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)shm);
        if (e != hipSuccess) return e;
    }
    long blocks = 0;
    hipError_t e = launch_grid(fn, shm, groups, pol.grid, &blocks);
    if (e != hipSuccess) return e;
    if (blocks > cap) blocks = cap;
    if (blocks == 0) return hipSuccess;
    // a grid capped below the groups by the workspace loops over them as a persistent grid does
    a.tickets = pol.grid != 0 ? pol.tickets : nullptr;
    hipLaunchKernelGGL((k_backward<T, SET, WPE>), dim3((unsigned)blocks), dim3(FM_WAVE * 4), shm, st, a);
    return hipGetLastError();
}

constexpr int FM_FUSED_MAXLEG = 6;   // multiple-scattering paths up to this many legs run the fused adjoint

template <class T, int SET, int NL, int WPE = 1>
hipError_t launch_fused_nl(BArgs a, long groups, const fm_launch &pol, size_t shm, hipStream_t st) {
    const void *fn = (const void *)k_adjoint<T, SET, NL, WPE>;
    if (shm > 64 * 1024) {
        hipError_t e = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)shm);
        if (e != hipSuccess) return e;
    }
    long blocks = 0;
    hipError_t e = launch_grid(fn, shm, groups, pol.grid, &blocks);
    if (e != hipSuccess) return e;
    if (blocks == 0) return hipSuccess;
    a.tickets = pol.grid != 0 ? pol.tickets : nullptr;
    hipLaunchKernelGGL((k_adjoint<T, SET, NL, WPE>), dim3((unsigned)blocks), dim3(FM_WAVE * 4), shm, st, a);
    return hipGetLastError();
}

// LDS of one path in k_adjoint: the geometry and the rotation tables with their beta derivatives.
template <class T, int SET> __host__ void layout_sweep(BArgs &a) {
    size_t o = align16(sizeof(PathLDS));
    const size_t tab = SET >= FM_SET_SS ? 0 : align16(sizeof(T) * rot_elems(a.nlegmax, a.ltab));
    a.off_W = o; o += tab;
    a.off_Wd = o; o += tab;
    a.path_bytes = path_stride(o);
}

// LDS of one wavefront (one path) of k_adjoint_tail.
template <class T> __host__ void layout_tail(BArgs &a, const fm_edge &ed) {
    size_t o = align16(sizeof(PathLDS));
    a.off_cc = o; o += align16(sizeof(cx<T>) * ed.E);
    a.off_amp = o; o += align16(sizeof(T) * ed.E);
    a.off_ang = o; o += align16(sizeof(double) * ed.E);
    a.off_cb = o; o += align16(sizeof(cx<T>) * ed.E);
    a.off_red = o; o += align16(sizeof(double) * (nacc(a.nlegmax) + 6 * FM_MAX_LEGS));
    const size_t qs = sizeof(double) * 3 * (size_t)a.nk;
    a.off_qs = 0;
    if (a.src.mode == 1 && qs > 0 && qs <= FM_QS_MAX) { a.off_qs = o; o += align16((2 * sizeof(T) + sizeof(double)) * (size_t)a.nk); }
    a.off_gs = o; o += align16(sizeof(double) * 3 * (size_t)a.nlegmax * a.nlegmax);
    a.path_bytes = o;
}

template <class T, int NL> hipError_t launch_tail(const BArgs &a, hipStream_t st) {
    const size_t shm = a.path_bytes * 4;
    if (shm > 64 * 1024) {
        hipError_t e = hipFuncSetAttribute((const void *)k_adjoint_tail<T, NL>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)shm);
        if (e != hipSuccess) return e;
    }
    hipLaunchKernelGGL((k_adjoint_tail<T, NL>), dim3((unsigned)((a.npath_bucket + 3) / 4)), dim3(FM_WAVE * 4), shm, st, a);
    return hipGetLastError();
}

constexpr size_t FM_DREGION_MAX = (size_t(1) << 31) - 4096;   // the sweep addresses its workspace with 32-bit offsets

// The fused adjoint (multiple scattering of 3..FM_FUSED_MAXLEG legs, single scattering): k_adjoint's sweep, then
// k_adjoint_tail, over chunks of the bucket of as many paths as the workspace holds regions for (one spare). Single
// scattering of K edges keeps little in registers: two workgroups per CU (FP32).
template <class T, int SET> hipError_t launch_fused(BArgs &a, const fm_launch &pol, hipStream_t st) {
    const fm_edge &ed = a.ed;
    const int nl = SET >= FM_SET_SS ? 2 : a.nlegmax;
    a.dstride = (size_t)(4 * nl + 1) * ed.E;
    const size_t per = a.dstride * sizeof(cx<T>);
    const size_t room = a.work_bytes < FM_DREGION_MAX ? a.work_bytes : FM_DREGION_MAX;
    const long chunk = (long)(room / per) - 1;
    if (chunk < 1) return hipErrorOutOfMemory;
    BArgs A = a, B = a;
    layout_sweep<T, SET>(A);
    size_t budget = fm_lds_budget(sizeof(T) == 4 ? 1 : 3);
    if (SET >= FM_SET_SS && budget > FM_LDS_SHARE) budget = FM_LDS_SHARE;
    const int g = choose_gpaths(ed, A.path_bytes, budget, pol.gpaths);
    if (g == -2) return hipErrorInvalidConfiguration;
    if (g < 0) return hipErrorInvalidValue;
    A.gp = 1 << g;
    A.bp = ed.bands[g];
    const size_t shm = A.path_bytes * A.gp;
    layout_tail<T>(B, ed);
    if (B.path_bytes * 4 > FM_LDS_MAX) return hipErrorInvalidValue;
    // bytes per path of the kept matrices
    const size_t keep_path = SET >= FM_SET_SS ? 0 : (size_t)ed.E * keep_rows<SET>(nl) * 2 * sizeof(T);
    for (long c0 = 0; c0 < a.npath_bucket; c0 += chunk) {
        const int cnt = (int)(a.npath_bucket - c0 < chunk ? a.npath_bucket - c0 : chunk);
        A.keep = a.keep ? (const unsigned char *)a.keep + (size_t)c0 * keep_path : nullptr;
        A.first = B.first = a.first + (int)c0;
        A.npath_bucket = B.npath_bucket = cnt;
        A.dbytes = B.dbytes = (size_t)(cnt + 1) * per;
        const long groups = (cnt + A.gp - 1) / A.gp;
        hipError_t e;
        if constexpr (SET >= FM_SET_SS) {
            // K: two workgroups per CU; L and M (27 and 21 representatives) would spill at that register budget
            if (sizeof(T) == 4 && SET == FM_SET_SS && shm <= FM_LDS_SHARE - 256) e = launch_fused_nl<T, SET, 2, 2>(A, groups, pol, shm, st);
            else e = launch_fused_nl<T, SET, 2, 1>(A, groups, pol, shm, st);
            if (e == hipSuccess) e = launch_tail<T, 2>(B, st);
        } else {
            switch (a.nlegmax) {
                case 3: e = launch_fused_nl<T, SET, 3>(A, groups, pol, shm, st); if (e == hipSuccess) e = launch_tail<T, 3>(B, st); break;
                case 4: e = launch_fused_nl<T, SET, 4>(A, groups, pol, shm, st); if (e == hipSuccess) e = launch_tail<T, 4>(B, st); break;
                case 5: e = launch_fused_nl<T, SET, 5>(A, groups, pol, shm, st); if (e == hipSuccess) e = launch_tail<T, 5>(B, st); break;
                case 6: e = launch_fused_nl<T, SET, 6>(A, groups, pol, shm, st); if (e == hipSuccess) e = launch_tail<T, 6>(B, st); break;
                default: e = hipErrorInvalidValue;
            }
        }
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

// Multiple scattering of more than FM_FUSED_MAXLEG legs: pass A stores the prefix products in the workspace, pass B
// reads them (k_backward). Everything else: the fused adjoint.
template <class T, int SET> hipError_t launch_backward(BArgs &a, const fm_launch &pol, hipStream_t st) {
    const fm_edge &ed = a.ed;
    a.ltab = ed.L1 + 1;
    if constexpr (SET >= FM_SET_SS) {
        return launch_fused<T, SET>(a, pol, st);
    } else {
        if (a.nlegmax >= 3 && a.nlegmax <= FM_FUSED_MAXLEG) return launch_fused<T, SET>(a, pol, st);
        layout<T, SET>(a, ed);
        const int g = choose_gpaths(ed, a.path_bytes, fm_lds_budget(sizeof(T) == 4 ? 1 : 3), pol.gpaths);
        if (g == -2) return hipErrorInvalidConfiguration;
        if (g < 0) return hipErrorInvalidValue;
        a.gp = 1 << g;
        a.bp = ed.bands[g];
        const size_t shm = a.path_bytes * a.gp;
        const long slots = (long)(a.work_bytes / a.work_per_wave);
        if (slots < a.bp.nband) return hipErrorInvalidValue;
        const long groups = (a.npath_bucket + a.gp - 1) / a.gp, cap = slots / a.bp.nband;
        if constexpr (sizeof(T) == 4) {
            if (shm <= FM_LDS_SHARE - 256) return launch_k<T, SET, 2>(a, groups, cap, pol, shm, st);
        }
        return launch_k<T, SET, 1>(a, groups, cap, pol, shm, st);
    }
}

template <class T> hipError_t dispatch_b(BArgs &a, int set, const fm_launch &pol, hipStream_t st) {
    switch (set) {
        case 0: return launch_backward<T, 0>(a, pol, st);
        case 1: return launch_backward<T, 1>(a, pol, st);
        case FM_SET_SS: return launch_backward<T, FM_SET_SS>(a, pol, st);
        case FM_SET_SS + 1: return launch_backward<T, FM_SET_SS + 1>(a, pol, st);
        case FM_SET_SS + 2: return launch_backward<T, FM_SET_SS + 2>(a, pol, st);
    }
    return hipErrorInvalidValue;
}

fm_launch policy_of_b(const fm_launch *launch, int b) {
    fm_launch d;
    memset(&d, 0, sizeof d);
    return launch ? launch[b] : d;
}

int check_edge_b(const fm_edge *e) {
    const char *m = edge_problem(e);
    if (m) return set_err(FM_EINVAL, m);
    if (e->ilinit < 1 || e->ilinit > 3) return set_err(FM_EUNSUPPORTED, "edge: only K, L and M edges (ilinit 1..3)");
    return FM_OK;
}

int set_of(const fm_edge &ed, int nleg) {
    if (nleg == 2) return FM_SET_SS + ed.ilinit - 1;
    return ed.ilinit <= 1 ? 0 : 1;
}

constexpr int FM_SCRATCH_MAX_DEV = 64;
std::mutex g_scratch_mu;
void *g_scratch_ptr[FM_SCRATCH_MAX_DEV];

// The device address of the current device's `scratch`, looked up on the first call on that device.
int scratch_of_current_device(void **out) {
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return hip_err(e, "hipGetDevice");
    if (dev < 0 || dev >= FM_SCRATCH_MAX_DEV) return set_err(FM_EINVAL, "fm_cchi_backward_noalloc: device index above 63");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void *p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return hip_err(e, "hipGetSymbolAddress(scratch)");
        if (!p) return set_err(FM_EHIP, "hipGetSymbolAddress(scratch) gave a null address");
        g_scratch_ptr[dev] = p;
    }
    *out = g_scratch_ptr[dev];
    return FM_OK;
}

}  // namespace

extern "C" {

size_t fm_backward_workspace_bytes(int32_t precision, int32_t nlegmax, int32_t band_slots) {
    if (nlegmax < 2 || nlegmax > FM_MAX_LEGS || band_slots < 1) return 0;
    const size_t ec = precision == 32 ? 8 : 16, er = ec / 2;
    const size_t nslots = nlegmax > 2 ? nlegmax - 1 : 1;
    const size_t cxe = nslots * 36 + (size_t)(nlegmax - 1) * FM_NREP;
    const size_t per = (ec * cxe * FM_WAVE + er * (size_t)nacc(nlegmax) * FM_WAVE + 255) & ~size_t(255);
    return per * (size_t)band_slots;
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
    if (workspace_bytes < fm_backward_workspace_bytes(precision, nleg, 16))
        return set_err(FM_EINVAL, "workspace smaller than fm_backward_workspace_bytes(precision, nleg, 16)");
    const int set = set_of(*edge, nleg);
    BArgs a;
    memset(&a, 0, sizeof a);
    a.ed = *edge;
    a.edge_id = 0;
    a.npath_bucket = npath;
    a.src.mode = 0;
    a.src.nleg = nleg;
    a.src.coords = coords;
    a.src.ipots64 = ipots;
    a.src.nph1 = edge->nph1;
    a.gcchi = gcchi;
    a.greff = greff;
    a.gcoord = gcoords;
    a.nslot = nleg;
    a.nk = 0;
    a.nlegmax = nleg;
    a.work = (unsigned char *)workspace;
    a.work_bytes = workspace_bytes;
    hipStream_t st = (hipStream_t)stream;
    const fm_launch pol = policy_of_b(nullptr, 0);
    hipError_t e = precision == 32 ? dispatch_b<float>(a, set, pol, st) : dispatch_b<double>(a, set, pol, st);
    if (e == hipErrorInvalidValue) return set_err(FM_EINVAL, "fm_cchi_backward: a path's tables do not fit one workgroup's LDS");
    if (e == hipErrorOutOfMemory) return set_err(FM_EINVAL, "fm_cchi_backward: workspace too small for one path");
    return e == hipSuccess ? FM_OK : hip_err(e, "fm_cchi_backward launch");
}

// fm_cchi_backward with the static device array `scratch` (FM_SCRATCH_BYTES per GPU) as its workspace: no allocation
// by the library or the caller. Nothing in the workspace needs to start at zero (every slot the adjoint reads was
// written earlier in the same call). Calls that overlap on one GPU share its array and race.
int fm_cchi_backward_noalloc(const fm_edge *edge, int32_t precision, int32_t iorder, int32_t npath, int32_t nleg,
                             const double *coords, const int64_t *ipots, const void *gcchi, const double *greff,
                             double *gcoords, void *stream) {
    int rc = check_edge_b(edge);
    if (rc) return rc;
    if (nleg < 2 || nleg > FM_MAX_LEGS) return set_err(FM_EINVAL, "nleg out of range 2..9");
    if (precision != 32 && precision != 64) return set_err(FM_EINVAL, "precision must be 32 or 64");
    if (npath <= 0) return npath == 0 ? FM_OK : set_err(FM_EINVAL, "npath < 0");
    // the least the adjoint works with: 16 band slots (7-9 legs) or one path's region and the spare (2-6 legs)
    const size_t per_path = (size_t)(4 * (nleg < 6 ? nleg : 6) + 1) * edge->E * (precision == 32 ? 8 : 16);
    if (fm_backward_workspace_bytes(precision, nleg, 16) > FM_SCRATCH_BYTES || 2 * per_path > FM_SCRATCH_BYTES)
        return set_err(FM_EINVAL, "fm_cchi_backward_noalloc: one path of this edge needs more than the 64 MiB scratch");
    void *ws = nullptr;
    rc = scratch_of_current_device(&ws);
    if (rc) return rc;
    return fm_cchi_backward(edge, precision, iorder, npath, nleg, coords, ipots, gcchi, greff, gcoords, ws,
                            FM_SCRATCH_BYTES, stream);
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
    const int E = edge->E;
    const size_t es = precision == 32 ? 4 : 8;
    if (grid->S && grid->sstride < E) return set_err(FM_EINVAL, "grid.sstride < E");
    const size_t qsb = sizeof(double) * 3 * (size_t)grid->nk;
    const int staged = qsb <= FM_QS_MAX ? 1 : 0;
    if (grid->S && !staged) return set_err(FM_EUNSUPPORTED, "a dense interpolation matrix needs nk <= 1365 in the adjoint");
    const size_t wb = align16(2 * es * E) + align16(es * E) + align16(8 * (size_t)E) + align16(2 * es * E) +
                      (staged ? align16(qsb) : 0);
    const int wpb = wb * WPB_MAX <= 64 * 1024 ? WPB_MAX : 1;
    const int blocks = (npath + wpb - 1) / wpb;
    hipStream_t st = (hipStream_t)stream;
    if (precision == 32)
        hipLaunchKernelGGL(k_chi_fine_bwd<float>, dim3(blocks), dim3(FM_WAVE * wpb), wb * wpb, st, *edge, npath,
                           (const float *)cchi, reff, *grid, (const float *)gchi, (float *)gcchi, greff, wb, staged);
    else
        hipLaunchKernelGGL(k_chi_fine_bwd<double>, dim3(blocks), dim3(FM_WAVE * wpb), wb * wpb, st, *edge, npath,
                           (const double *)cchi, reff, *grid, (const double *)gchi, (double *)gcchi, greff, wb, staged);
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? FM_OK : hip_err(e, "fm_chi_fine_backward launch");
}

int fm_chi_paths_backward(const fm_edge *edges, int32_t nedge, const fm_paths *paths, const fm_grid *grid,
                          int32_t precision, int32_t iorder, const int32_t *order, int32_t nbucket,
                          const int32_t *bucket_offset, const int32_t *bucket_kind, const int32_t *bucket_edge,
                          const int32_t *bucket_nleg, const int32_t *segment, const double *gseg, const void *gpath, int32_t natom,
                          const int32_t *atom_ptr, const int32_t *atom_refs, double *gcoord, double *grad_pos,
                          void *workspace, size_t workspace_bytes, const fm_launch *launch, void *stream) {
    if (nedge < 1 || nedge > FM_MAX_EDGE) return set_err(FM_EINVAL, "nedge out of range 1..4");
    int emax = 0;
    for (int i = 0; i < nedge; ++i) {
        int rc = check_edge_b(&edges[i]);
        if (rc) return rc;
        emax = emax > edges[i].E ? emax : edges[i].E;
    }
    if (iorder != 2) return set_err(FM_EUNSUPPORTED, "only iorder = 2 is implemented on the GPU");
    if (!paths || !grid || !gcoord || !grid->qorder || !grid->qstart) return set_err(FM_EINVAL, "NULL paths, grid or gcoord");
    if (grid->S && grid->sstride < emax) return set_err(FM_EINVAL, "grid.sstride < E");
    if (grid->S && sizeof(double) * 3 * (size_t)grid->nk > FM_QS_MAX)
        return set_err(FM_EUNSUPPORTED, "a dense interpolation matrix needs nk <= 1365 in the adjoint");
    if (!gseg && !gpath) return set_err(FM_EINVAL, "no upstream gradient");
    if (gseg && !segment) return set_err(FM_EINVAL, "gseg needs segment");
    if (precision != 32 && precision != 64) return set_err(FM_EINVAL, "precision must be 32 or 64");
    if (paths->nmax < 2 || paths->nmax > FM_MAX_LEGS) return set_err(FM_EINVAL, "nmax out of range 2..9");
    if (grid->qstride < emax + 1) return set_err(FM_EINVAL, "grid.qstride < E + 1");
    if (!workspace || workspace_bytes < fm_backward_workspace_bytes(precision, paths->nmax, 16))
        return set_err(FM_EINVAL, "workspace smaller than fm_backward_workspace_bytes(precision, nmax, 16)");
    if (nbucket > 0 && (!bucket_offset || !bucket_kind || !bucket_nleg)) return set_err(FM_EINVAL, "NULL bucket arrays");
    hipStream_t st = (hipStream_t)stream;
    for (int b = 0; b < nbucket; ++b) {
        const int cnt = bucket_offset[b + 1] - bucket_offset[b];
        if (cnt <= 0) continue;
        const int kind = bucket_kind[b];
        if (kind < 0 || kind > 2) return set_err(FM_EINVAL, "bucket kind out of range 0..2");
        const int eid = bucket_edge ? bucket_edge[b] : 0;
        if (eid < 0 || eid >= nedge) return set_err(FM_EINVAL, "bucket edge out of range");
        BArgs a;
        memset(&a, 0, sizeof a);
        a.ed = edges[eid];
        a.edge_id = eid;
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
        a.nk = grid->nk;
        a.nlegmax = bucket_nleg[b];
        if (a.nlegmax < 2 || a.nlegmax > paths->nmax) return set_err(FM_EINVAL, "bucket nleg out of range 2..nmax");
        if ((kind == FM_BUCKET_SS) != (a.nlegmax == 2)) return set_err(FM_EINVAL, "bucket kind does not match its nleg");
        if ((kind == FM_BUCKET_K_MS) != (a.ed.ilinit <= 1) && kind != FM_BUCKET_SS)
            return set_err(FM_EINVAL, "bucket kind does not match its edge");
        const int set = set_of(a.ed, a.nlegmax);
        a.work = (unsigned char *)workspace;
        a.work_bytes = workspace_bytes;
        const fm_launch pol = policy_of_b(launch, b);
        if (pol.gpaths != 0 && pol.gpaths != 1 && pol.gpaths != 2 && pol.gpaths != 4 && pol.gpaths != 8)
            return set_err(FM_EINVAL, "launch policy: gpaths must be 0, 1, 2, 4 or 8");
        if (pol.grid < -1) return set_err(FM_EINVAL, "launch policy: grid must be -1, 0 or a number of workgroups");
        if (pol.keep && kind != FM_BUCKET_SS && a.nlegmax >= 3 && a.nlegmax <= 6) {
            if (pol.keep_bytes < fm_keep_bytes(precision, &a.ed, a.nlegmax, cnt))
                return set_err(FM_EINVAL, "fm_chi_paths_backward: keep_bytes smaller than fm_keep_bytes for the bucket");
            a.keep = pol.keep;
        }
        hipError_t e = precision == 32 ? dispatch_b<float>(a, set, pol, st) : dispatch_b<double>(a, set, pol, st);
        if (e == hipErrorInvalidValue)
            return set_err(FM_EINVAL, "fm_chi_paths_backward: the bucket's paths do not fit one workgroup's LDS");
        if (e == hipErrorInvalidConfiguration)
            return set_err(FM_EINVAL, "fm_chi_paths_backward: the launch policy's gpaths has no band plan or does not "
                                      "fit one CU's LDS");
        if (e == hipErrorOutOfMemory) return set_err(FM_EINVAL, "fm_chi_paths_backward: workspace too small for one path");
        if (e != hipSuccess) return hip_err(e, "fm_chi_paths_backward launch");
    }
    if (grad_pos && natom > 0) {
        if (!atom_ptr || !atom_refs) return set_err(FM_EINVAL, "NULL atom lists");
        hipLaunchKernelGGL(k_atom_grad, dim3((natom + 3) / 4), dim3(FM_WAVE * 4), 0, st, natom, atom_ptr, atom_refs,
                           gcoord, grad_pos);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return hip_err(e, "atom gradient launch");
    }
    return FM_OK;
}

}  // extern "C"
