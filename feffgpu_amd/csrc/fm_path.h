// Per-path machinery shared by the forward and adjoint kernels: per-wave LDS layout, path loading and FP64 leg
// geometry, rotation tables, the streamed scattering-matrix blocks, the amplitude at one energy and the fine-grid
// epilogue. Reference semantics: src/feffgpu/stage.py:147-388, src/feffgpu/exafs/forward.py:153-163.
#pragma once
#include <hip/hip_runtime.h>

#include <type_traits>

#include "fm_device.h"
#include "fm_ss.h"

namespace fm {

constexpr int FM_MAX_EDGE = 4;
constexpr int WPB_MAX = 4;        // wavefronts (paths) per workgroup at most
constexpr int FM_MAX_ECHUNK = 4;  // energies up to 4 x 64
#define FM_WPE(T) (sizeof(T) == 4 ? 2 : 1)   // wavefronts per SIMD the register budget is sized for

// Where a kernel reads its paths: mode 0 explicit coordinates (GPUPathStage.cchi), mode 1 a path list.
struct PathSrc {
    int mode;
    int nleg;                 // mode 0
    const double *coords;     // mode 0: [npath][nleg][3] A
    const int64_t *ipots64;   // mode 0: [npath][nleg]
    int nph1;                 // mode 0: potentials 0..nph1-1 exist; a path naming another is a placeholder (NaN)
    fm_paths P;               // mode 1
};

// ------------------------------------------------------------------------------------------------ per-wave LDS
struct PathLDS {
    double rat[FM_MAX_LEGS + 1][3];   // bohr, rat[0] = absorber (origin), rat[1..n-1] scatterers
    double r[FM_MAX_LEGS + 1], beta[FM_MAX_LEGS + 1], alpha[FM_MAX_LEGS + 1], gamma[FM_MAX_LEGS + 1];
    double eta[FM_MAX_LEGS + 1];
    double c1[FM_MAX_LEGS + 1], s1[FM_MAX_LEGS + 1], c2[FM_MAX_LEGS + 1], s2[FM_MAX_LEGS + 1];  // cos/sin eta, 2 eta
    double cb[FM_MAX_LEGS + 1], sb[FM_MAX_LEGS + 1], ch[FM_MAX_LEGS + 1], sh[FM_MAX_LEGS + 1];  // cos/sin beta, beta/2
    int ipot[FM_MAX_LEGS + 1];       // potential of path atom j (0 and n: absorber)
    double reff, R, prod_r;           // bohr
    int n, ok, pad_[2];               // ok = 0: placeholder for a path whose leg count is not the launch's
};


__host__ __device__ inline size_t align16(size_t x) { return (x + 15) & ~size_t(15); }

// The LDS stride between the paths of a workgroup: an odd multiple of 16 bytes, so that the up to 8 paths whose
// lanes share a wavefront read the same table offset from 8 different groups of banks.
__host__ __device__ inline size_t path_stride(size_t bytes) {
    const size_t s = align16(bytes);
    return (s / 16) % 2 ? s : s + 16;
}

// Elements of the rotation tables of a path of at most nl legs: legs 1..nl-1 hold ltab rows each; the last leg
// (the termination's) is read at rows 1..4 only, and is stored with 5 rows at leg offset n - 1 <= nl - 1.
__host__ __device__ inline size_t rot_elems(int nl, int ltab) { return ((size_t)(nl - 1) * ltab + 5) * FM_NREP; }

// ------------------------------------------------------------------------------------------------ host checks
constexpr int FM_MAX_E = 256;
constexpr size_t FM_LDS_MAX = 160 * 1024;       // LDS of one CU
constexpr size_t FM_LDS_SHARE = 80 * 1024;      // what lets two workgroups share a CU

// The edge's tables and band plans are consistent; returns the problem or nullptr.
inline const char *edge_problem(const fm_edge *e) {
    if (!e) return "edge table is NULL";
    if (e->E < 2 || e->E > FM_MAX_E) return "edge: E out of range 2..256";
    if (e->L1 < 1 || e->L1 > 120) return "edge: L1 out of range 1..120";
    if (e->nph1 < 1 || e->nph1 > FM_MAX_POT) return "edge: nph1 out of range";
    if (e->nchan < 1 || e->nchan > FM_MAX_CHAN) return "edge: nchan out of range";
    for (int c = 0; c < e->nchan; ++c)
        if (e->chan_il[c] < 1 || e->chan_il[c] > 4 || e->chan_il[c] > e->L1) return "edge: channel il out of range 1..min(4, L1)";
    for (int i = 0; i < e->nph1; ++i)
        if (e->lmax1[i] < 1 || e->lmax1[i] > e->L1) return "edge: lmax1 out of range";
    if (!e->lmax || !e->ck || !e->ckmxk || !e->rotc || !e->rotn) return "edge: NULL table";
    for (int g = 0; g < 4; ++g) {
        const fm_bands &b = e->bands[g];
        if (b.nband == 0) continue;
        if (b.nband < 0 || b.nband > 16) return "edge: band count out of range 0..16";
        if (b.start[0] != 0 || b.start[b.nband] != e->E) return "edge: bands must cover the energies 0..E";
        for (int k = 0; k < b.nband; ++k) {
            const int len = b.start[k + 1] - b.start[k];
            if (len < 1 || len > FM_WAVE >> g) return "edge: a band holds 0 or more than 64 / gpaths energies";
            if (b.wave[k] < 0 || b.wave[k] > 3) return "edge: band wavefront out of range 0..3";
        }
    }
    if (e->bands[0].nband == 0) return "edge: no band plan for one path per workgroup";
    return nullptr;
}

// log2 of the largest gpaths with a band plan whose paths (path_bytes of LDS each) fit `budget`; -1 if none.
inline int pick_gpaths(const fm_edge &e, size_t path_bytes, size_t budget) {
    for (int g = 3; g >= 0; --g)
        if (e.bands[g].nband > 0 && (path_bytes << g) <= budget) return g;
    return -1;
}

// log2 of the paths per workgroup of a launch: `want` (1, 2, 4 or 8) if it has a band plan and fits one CU's LDS
// (-2 if not), else (want 0) the largest within `budget`, else the largest that fits one CU (-1 if none).
inline int choose_gpaths(const fm_edge &e, size_t path_bytes, size_t budget, int want) {
    if (want > 0) {
        const int g = want == 1 ? 0 : want == 2 ? 1 : want == 4 ? 2 : want == 8 ? 3 : -1;
        if (g < 0 || e.bands[g].nband == 0 || (path_bytes << g) > FM_LDS_MAX - 256) return -2;
        return g;
    }
    const int g = pick_gpaths(e, path_bytes, budget - 256);
    return g >= 0 ? g : pick_gpaths(e, path_bytes, FM_LDS_MAX - 256);
}

// Workgroups of a launch over `groups` path groups: one per group (grid 0), or a persistent grid of `grid`
// workgroups (-1: as many as the device's CUs hold at once with this kernel's registers and `shm` bytes of LDS),
// never more than the groups. The kernel's dynamic LDS limit must already allow `shm`.
inline hipError_t launch_grid(const void *fn, size_t shm, long groups, int grid, long *out) {
    *out = groups;
    if (grid == 0 || groups == 0) return hipSuccess;
    long g = grid;
    if (grid < 0) {
        int dev = 0, cus = 0, per = 0;
        hipError_t e = hipGetDevice(&dev);
        if (e == hipSuccess) e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
        if (e == hipSuccess) e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, fn, FM_WAVE * 4, shm);
        if (e != hipSuccess) return e;
        g = (long)(cus > 0 ? cus : 1) * (per > 0 ? per : 1);
    }
    if (g < groups) *out = g;
    return hipSuccess;
}

// ------------------------------------------------------------------------------------------------ path groups
// Which group of paths a workgroup runs next. tickets NULL: workgroup b takes groups b, b + grid, b + 2 grid, ...
// Otherwise every group comes from the counter tickets[0], one atomic per group; tickets[1] counts the workgroups
// that have run out of groups, and the last of them puts both back to zero, so the pair (zero before the first
// launch) is ready for the next launch that uses it on the same stream, with no fill in between. Every thread of the
// workgroup calls these (they hold barriers when tickets is set); `slot` is a __shared__ int of the kernel.
__device__ inline int take_group(unsigned *tickets, int &slot) {
    __syncthreads();
    if (threadIdx.x == 0) slot = (int)atomicAdd(tickets, 1u);
    __syncthreads();
    return __builtin_amdgcn_readfirstlane(slot);
}
__device__ inline int first_group(unsigned *tickets, int &slot) {
    return tickets ? take_group(tickets, slot) : (int)blockIdx.x;
}
__device__ inline int next_group(unsigned *tickets, int &slot, int blk) {
    return tickets ? take_group(tickets, slot) : blk + (int)gridDim.x;
}
__device__ inline void release_groups(unsigned *tickets) {
    if (tickets && threadIdx.x == 0) {
        __threadfence();
        if (atomicAdd(tickets + 1, 1u) == gridDim.x - 1) {
            atomicExch(tickets, 0u);
            atomicExch(tickets + 1, 0u);
        }
    }
}

// ------------------------------------------------------------------------------------------------ setup of one path
// Point j of path p of a list, in bohr relative to its absorber (point 0 = the absorber itself), by the minimum
// image of the list's cell (src/feffgpu/exafs/forward.py:153-156, src/feffgpu/cell.py:26-35).
__device__ inline void path_point(const fm_paths &P, int p, int j, double (&v)[3]) {
    v[0] = v[1] = v[2] = 0.0;
    if (j > 0) {
        const int at = P.atoms[(size_t)p * (P.nmax - 1) + j - 1];
        const int ab = P.absorber[p];
#pragma unroll
        for (int c = 0; c < 3; ++c) v[c] = P.pos[3 * (size_t)at + c] - P.pos[3 * (size_t)ab + c];
        if (P.cell_kind == 1) {
            const double L = P.cell[0];
#pragma unroll
            for (int c = 0; c < 3; ++c) v[c] -= L * rint(v[c] / L);
        } else if (P.cell_kind == 2) {
            double f[3];
#pragma unroll
            for (int c = 0; c < 3; ++c) f[c] = v[0] * P.cellinv[0 * 3 + c] + v[1] * P.cellinv[1 * 3 + c] + v[2] * P.cellinv[2 * 3 + c];
#pragma unroll
            for (int c = 0; c < 3; ++c) f[c] -= rint(f[c]);
#pragma unroll
            for (int c = 0; c < 3; ++c) v[c] = f[0] * P.cell[0 * 3 + c] + f[1] * P.cell[1 * 3 + c] + f[2] * P.cell[2 * 3 + c];
        }
    }
#pragma unroll
    for (int c = 0; c < 3; ++c) v[c] /= FM_BOHR;
}

// Fills g (geometry, potentials) for path slot `p`; all lanes of the wavefront call it, those with act set work on
// this path as lanes ln = 0..15 of a group (several groups of a wavefront fill several paths at once). Every path
// of a launch has nuni legs (the kernels loop over the legs uniformly): a path-list entry with another leg count is
// replaced by a fixed placeholder path of nuni legs with g.ok = 0, so that every index stays in bounds; the kernels
// write NaN for it. In mode 0 a path that names a potential outside 0..nph1-1 reads the tables of potential 0 and
// has g.ok = 0 (NaN output), so the caller need not check its potentials on the host.
__device__ inline int load_path_group(const PathSrc &a, int p, PathLDS &g, int &edge_idx, int nuni, int lane, bool act,
                                      bool beta_trig = true) {
    int n;
    bool ok = true;
    if (a.mode == 0) {
        n = nuni;
        edge_idx = 0;
        if (act && lane < n) {
            const double *c = a.coords + ((size_t)p * n + lane) * 3;
            double x = c[0] / FM_BOHR, y = c[1] / FM_BOHR, z = c[2] / FM_BOHR;
            // absorber (lane 0) is the path's first point; the stage places it at rat[0] and rat[n]
            g.rat[lane][0] = x; g.rat[lane][1] = y; g.rat[lane][2] = z;
            const int64_t ip = a.ipots64[(size_t)p * n + lane];
            g.ipot[lane] = (ip >= 0 && ip < a.nph1) ? (int)ip : 0;    // tables are read at potential 0 instead
        }
        if (act && lane == 0)
            for (int j = 0; j < n; ++j) {
                const int64_t ip = a.ipots64[(size_t)p * n + j];
                if (ip < 0 || ip >= a.nph1) ok = false;
            }
    } else {
        ok = a.P.nleg[p] == nuni;
        n = nuni;
        edge_idx = a.P.edge ? a.P.edge[p] : 0;
        const int nmax = a.P.nmax;
        if (!ok) {
            if (act && lane < n) {
                g.rat[lane][0] = 4.0 * lane;
                g.rat[lane][1] = 1.5 * (lane & 1);
                g.rat[lane][2] = 0.5 * lane;
                g.ipot[lane] = 0;
            }
        } else if (act && lane < n) {
            double v[3];
            path_point(a.P, p, lane, v);
#pragma unroll
            for (int c = 0; c < 3; ++c) g.rat[lane][c] = v[c];
            g.ipot[lane] = a.P.ipots[(size_t)p * nmax + lane];
        }
    }
    wave_sync();
    if (act && lane < 3) g.rat[n][lane] = g.rat[0][lane];
    if (act && lane == 0) { g.ipot[n] = g.ipot[0]; g.n = n; g.ok = ok ? 1 : 0; }
    wave_sync();
    // one lane per leg
    if (act && lane < n) {
        LegGeo lg = leg_geometry(g.rat, lane + 1, n);
        g.r[lane + 1] = lg.r;
        g.beta[lane + 1] = lg.beta;
        g.alpha[lane + 1] = lg.alpha;
        g.gamma[lane + 1] = lg.gamma;
    }
    wave_sync();
    if (act && lane < n) {
        const int J = lane + 1;
        double eta = (J == 1 ? g.alpha[n] : g.alpha[J - 1]) + g.gamma[J];
        g.eta[J] = eta;
        double s, c;
        sincos(eta, &s, &c);
        g.c1[J] = c; g.s1[J] = s;
        sincos(2.0 * eta, &s, &c);
        g.c2[J] = c; g.s2[J] = s;
        if (beta_trig) {   // for the rotation tables only
            const double be = g.beta[J];
            sincos(be, &s, &c);
            g.cb[J] = c; g.sb[J] = s;
            sincos(0.5 * be, &s, &c);
            g.ch[J] = c; g.sh[J] = s;
        }
    }
    if (act && lane == 0) {
        double R = 0.0, pr = 1.0;
        for (int j = 1; j <= n; ++j) { R += g.r[j]; pr *= g.r[j]; }
        g.R = R; g.reff = 0.5 * R; g.prod_r = pr;
    }
    wave_sync();
    return n;
}

__device__ inline int load_path(const PathSrc &a, int p, PathLDS &g, int &edge_idx, int nuni, bool beta_trig = true) {
    return load_path_group(a, p, g, edge_idx, nuni, lane_id(), true, beta_trig);
}

// The recursion coefficients of B consecutive rows il0.. of one (m1, m2) (mm, nn as build_tables forms them),
// fetched together so that their load latencies overlap: from l - 1 to l for l >= 3, and the normalisation.
template <int B> struct RotCoef {
    typedef double vec __attribute__((ext_vector_type(B)));   // vector values, not arrays: kept in registers
    vec c0, c1, c2, nr;
    // B > 1: the loads are unconditional (a row past `rows`, or below l = 3, reads a row inside the tables, rotn
    // row min(il, rows) <= L1 and rotc row min(max(l - 1, 2), L1)) and the values are zeroed after, so that the
    // compiler issues all of them before the first wait instead of branching around each. B = 1 (the forward's
    // tables, one row at a time) loads only what it uses (measured faster there).
    __device__ __forceinline__ void fetch(const fm_edge &ed, int il0, int rows, int mm, int nn) {
        if constexpr (B == 1) {
            const int il = il0, l = il - 1;
            nr[0] = il <= rows ? ed.rotn[(size_t)il * 16 + nn] : 0.0;
            if (il <= rows && l >= 3) {
                const double *co = ed.rotc + ((size_t)(l - 1) * 49 + mm) * 3;
                c0[0] = co[0];
                c1[0] = co[1];
                c2[0] = co[2];
            } else {
                c0[0] = c1[0] = c2[0] = 0.0;
            }
            return;
        }
#pragma unroll
        for (int u = 0; u < B; ++u) {
            const int il = il0 + u, l = il - 1;
            const bool live = il <= rows, rec = live && l >= 3;
            const int ilc = live ? il : rows;
            const int lrow = min(ilc - 1 >= 3 ? ilc - 2 : 2, ed.L1);
            const double n0 = ed.rotn[(size_t)ilc * 16 + nn];
            const double *co = ed.rotc + ((size_t)lrow * 49 + mm) * 3;
            const double a0 = co[0], a1 = co[1], a2 = co[2];
            nr[u] = live ? n0 : 0.0;
            c0[u] = rec ? a0 : 0.0;
            c1[u] = rec ? a1 : 0.0;
            c2[u] = rec ? a2 : 0.0;
        }
    }
};

// Rotation tables of the paths j = j0, j0 + jstep, ... < G held by a wavefront (path j at smem + j path_bytes, its
// PathLDS first): W[s][il][k] = d^{il-1}_{m1 m2}(beta_{s+1}) xnlm(il,|m1|)(-1)^|m1| / xnlm(il,|m2|) for legs s = 0..n-1
// and the FM_NREP representatives k, il = 1..L1 (row 0 unused; the last leg, which only the termination reads, up to
// il = 4); with WITH_D also Wd = the same with d/dbeta d. Three-term recursion in l for fixed (m1, m2)
// (src/feffgpu/stage.py:77-92) seeded by the closed forms at l <= 2, in FP64, differentiated:
// d'^l = (P c - Q) d'^{l-1} - P sin(beta) d^{l-1} - R d'^{l-2}. The (path, leg, representative) tasks of all the
// wavefront's paths are spread over its lanes; cos/sin of beta and beta/2 come from the path's geometry. The
// coefficients of BATCH rows are fetched at a time, as vector values held in registers.
template <class T, int SET, bool WITH_D, int BATCH = 8, int ILP = 1>
__device__ void build_tables(const fm_edge &ed, unsigned char *smem, size_t path_bytes, size_t off_W, size_t off_Wd,
                             int j0, int jstep, int G, int n, int ltab) {
    constexpr PairDef PD = pairdef(SET);
    const int lane = lane_id();
    const int np = j0 < G ? (G - 1 - j0) / jstep + 1 : 0;
    const int per = n * FM_NREP, ntask = np * per;
    // ILP independent recursions per lane (tasks t, t + 64, ...), interleaved row by row
    for (int t0 = lane; t0 < ntask; t0 += FM_WAVE * ILP) {
        int m1[ILP], m2[ILP], mm[ILP], nn[ILP], rows[ILP], kk[ILP];
        double c[ILP], sn[ILP], c2[ILP], s2[ILP], dlm1[ILP], dl[ILP], tlm1[ILP], tl[ILP];
        T *Ws[ILP], *Wds[ILP];
        int rmax = 0;
#pragma unroll
        for (int q = 0; q < ILP; ++q) {
            const int t = t0 + q * FM_WAVE;
            const bool on = t < ntask;
            const int tt = on ? t : t0;
            const int jj = tt / per, rr = tt - jj * per;
            const int s = rr / FM_NREP, k = rr - s * FM_NREP;
            unsigned char *base = smem + (size_t)(j0 + jj * jstep) * path_bytes;
            const PathLDS &g = *reinterpret_cast<const PathLDS *>(base);
            m1[q] = m2[q] = 0;
#pragma unroll
            for (int r = 0; r < FM_NREP; ++r)
                if (r == k) { m1[q] = PD.m1[r]; m2[q] = PD.m2[r]; }
            c[q] = g.cb[s + 1]; sn[q] = g.sb[s + 1]; c2[q] = g.ch[s + 1]; s2[q] = g.sh[s + 1];
            Ws[q] = reinterpret_cast<T *>(base + off_W) + (size_t)s * ltab * FM_NREP;
            Wds[q] = reinterpret_cast<T *>(base + off_Wd) + (size_t)s * ltab * FM_NREP;
            kk[q] = k;
            const int am1 = m1[q] < 0 ? -m1[q] : m1[q], am2 = m2[q] < 0 ? -m2[q] : m2[q];
            mm[q] = (m1[q] + 3) * 7 + (m2[q] + 3);
            nn[q] = am1 * 4 + am2;
            rows[q] = on ? (s == n - 1 ? (ed.L1 < 4 ? ed.L1 : 4) : ed.L1) : 0;
            rmax = rows[q] > rmax ? rows[q] : rmax;
            dlm1[q] = dl[q] = tlm1[q] = tl[q] = 0.0;
        }
        // rows in batches of BATCH (their coefficients fetched together); the seeds l <= 2 fall in the first batch,
        // which is peeled so that the seed-or-recursion choice is made at compile time
        auto batch = [&](int il0, auto first) {
            RotCoef<BATCH> rc[ILP];
#pragma unroll
            for (int q = 0; q < ILP; ++q) rc[q].fetch(ed, il0, rows[q], mm[q], nn[q]);
#pragma unroll
            for (int u = 0; u < BATCH; ++u) {   // no early exit: the batch stays in registers
                const int il = il0 + u, l = il - 1;
#pragma unroll
                for (int q = 0; q < ILP; ++q) {
                    if (il > rows[q]) continue;
                    double d, t1 = 0.0;
                    // the first batch starts at row 1 (l = u); batches of fewer than 3 rows test l at run time
                    const bool seed = decltype(first)::value ? (u <= 2) : (BATCH < 3 && l <= 2);
                    if (seed) {
                        d = dseed(l, m1[q], m2[q], c[q], sn[q], c2[q], s2[q]);
                        if constexpr (WITH_D) t1 = dseed_db(l, m1[q], m2[q], c[q], sn[q], c2[q], s2[q]);
                    } else {
                        const double a1 = rc[q].c0[u] * c[q] - rc[q].c1[u];
                        d = a1 * dl[q] - rc[q].c2[u] * dlm1[q];
                        if constexpr (WITH_D) t1 = a1 * tl[q] - rc[q].c0[u] * sn[q] * dl[q] - rc[q].c2[u] * tlm1[q];
                    }
                    dlm1[q] = dl[q]; dl[q] = d;
                    Ws[q][(size_t)il * FM_NREP + kk[q]] = (T)(d * rc[q].nr[u]);
                    if constexpr (WITH_D) {
                        tlm1[q] = tl[q]; tl[q] = t1;
                        Wds[q][(size_t)il * FM_NREP + kk[q]] = (T)(t1 * rc[q].nr[u]);
                    }
                }
            }
        };
        if (rmax >= 1) batch(1, std::true_type{});
        for (int il0 = 1 + BATCH; il0 <= rmax; il0 += BATCH) batch(il0, std::false_type{});
    }
    wave_sync();
}

// One partial-wave row of a scattering matrix: Ft[k] += X_a(k) Y_b(k) w[k] for the representatives k of the
// NR x NC block (Y already carries the t-matrix).
// The rotation-table values of one row for the representatives of an NR x NC block, LDS -> registers.
template <class T, int SET, int NR, int NC>
__device__ __forceinline__ void load_wrow(T (&v)[FM_NREP], const T *__restrict__ w) {
    constexpr PairDef PD = pairdef(SET);
#pragma unroll
    for (int k = 0; k < FM_NREP; ++k) v[k] = (PD.ri[k] < NR && PD.rj[k] < NC) ? w[k] : T(0);
}

template <class T, int SET, int NR, int NC, class WRow>
__device__ __forceinline__ void accum_row(cx<T> (&Ft)[FM_NREP], const cx<T> (&xa)[3], cx<T> y0, cx<T> y1,
                                          const WRow &w) {
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
// Accumulates the representatives k with ri < NR, rj < NC into Ft (zeroed here). Each lane runs its own rows
// il = 1..ilx (the wave runs to the largest ilx of its lanes); the t-matrix is fetched two rows ahead.
// PREFETCH: each row's table values are read into registers before its recursion steps (the adjoint's sweep, at one
// wavefront per SIMD, has no second wavefront to hide the LDS latency; the forward measures faster with it too);
// otherwise where the compiler puts them.
template <class T, int SET, int NR, int NC, bool PREFETCH = false>
__device__ __forceinline__ void fblock(cx<T> (&Ft)[FM_NREP], cx<T> ck, T rX, T rY, int ilx,
                                       const T *__restrict__ Wl, const T *__restrict__ tl, int tlstride) {
#pragma unroll
    for (int k = 0; k < FM_NREP; ++k) Ft[k] = mk<T>(T(0), T(0));
    Hank3<T> X;
    Hank2<T> Y;
    X.init(minus_i_over(ck * rX));
    Y.init(minus_i_over(ck * rY));
    cx<T> xa[3], yb[2];
    // the t-matrix two rows ahead (rows past ilx repeat row ilx, which stays in the table)
    cx<T> t = mk<T>(tl[tlstride], tl[tlstride + 1]);
    const size_t r2 = (size_t)(ilx < 2 ? ilx : 2) * tlstride;
    cx<T> tn = mk<T>(tl[r2], tl[r2 + 1]);
    // rows 1..4 unrolled (their seeds are explicit per row); a lane whose energy has fewer rows runs them with the
    // t-matrix factor zero (and the table's row ilx, inside the table and finite), which adds nothing
#pragma unroll
    for (int r = 1; r <= 4; ++r) {
        X.seed_row(r, xa);
        if (r < 4) Y.seed_row(r, yb); else Y.step(r, yb);
        const size_t rn = (size_t)(r + 2 < ilx ? r + 2 : ilx) * tlstride;
        const cx<T> tnn = mk<T>(tl[rn], tl[rn + 1]);
        const cx<T> tr = r <= ilx ? t : mk<T>(T(0), T(0));
        accum_row<T, SET, NR, NC>(Ft, xa, yb[0] * tr, yb[1] * tr, Wl + (size_t)(r <= ilx ? r : ilx) * FM_NREP);
        t = tn;
        tn = tnn;
    }
    if constexpr (PREFETCH) {
        // two rows per iteration, so that the recursions' two-row state is renamed instead of copied every row
        auto row = [&](int r) {
            const size_t rn = (size_t)(r + 2 < ilx ? r + 2 : ilx) * tlstride;
            T wv[FM_NREP];
            load_wrow<T, SET, NR, NC>(wv, Wl + (size_t)r * FM_NREP);
            __builtin_amdgcn_sched_barrier(0);
            X.step(r, xa);
            Y.step(r, yb);
            const cx<T> tnn = mk<T>(tl[rn], tl[rn + 1]);
            accum_row<T, SET, NR, NC>(Ft, xa, yb[0] * t, yb[1] * t, wv);
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
    } else {
        // two rows per iteration as well (the state renamed rather than copied), the table values read where used
        auto row = [&](int r) {
            X.step(r, xa);
            Y.step(r, yb);
            const size_t rn = (size_t)(r + 2 < ilx ? r + 2 : ilx) * tlstride;
            const cx<T> tnn = mk<T>(tl[rn], tl[rn + 1]);
            accum_row<T, SET, NR, NC>(Ft, xa, yb[0] * t, yb[1] * t, Wl + (size_t)r * FM_NREP);
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

// cfac = exp(i R (ck - xk)) / prod_j (ck r_j), R = sum r_j = 2 reff (src/feffgpu/stage.py:336-337); the phase
// R Re(ck - xk) in FP64 reduced mod 2 pi
template <class T>
__device__ __forceinline__ cx<T> cfac_at(const fm_edge &ed, const PathLDS &g, int ee, int n) {
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

// cchi at energy e of the path held in g (tables in W). All lanes of the wave run it (e may be out of range for
// some; they compute garbage that the caller drops, which keeps the loops uniform).
// Kept scattering matrices: [path][row][energy] complex T, energy fastest, so that the lanes of a band (consecutive
// energies of one path) write and read consecutive elements of each row. The rows of a path hold the
// representatives of each scatterer's block, scatterer 1 first: its block is L x L0, the last one's L0 x L, the
// others L x L (only those representatives are ever used). A lane's pointer kp is at (its path, row 0, its
// energy); row r of the lane is at kp + 2 r E.
__host__ __device__ constexpr bool keep_used(int set, int nr, int nc, int k) {
    return pairdef(set).ri[k] < nr && pairdef(set).rj[k] < nc;
}
__host__ __device__ constexpr int keep_count(int set, int nr, int nc) {
    int c = 0;
    for (int k = 0; k < FM_NREP; ++k) c += keep_used(set, nr, nc, k) ? 1 : 0;
    return c;
}
// the representatives of an NR x NC block, in index order
struct KeepList {
    int n;
    int k[FM_NREP];
};
__host__ __device__ constexpr KeepList keep_list(int set, int nr, int nc) {
    KeepList L{};
    for (int k = 0; k < FM_NREP; ++k)
        if (keep_used(set, nr, nc, k)) L.k[L.n++] = k;
    return L;
}
// rows before scatterer s's block for a path of n legs (run-time n and s; the block sizes are compile-time
// constants), and the rows of a path
template <int SET> __host__ __device__ inline int keep_off(int n, int s) {
    constexpr int c1 = keep_count(SET, lamdef(SET).L, lamdef(SET).L0), cm = keep_count(SET, lamdef(SET).L, lamdef(SET).L);
    return s == 1 ? 0 : c1 + (s - 2) * cm;
}
template <int SET> __host__ __device__ inline int keep_rows(int n) {
    constexpr int cl = keep_count(SET, lamdef(SET).L0, lamdef(SET).L);
    return keep_off<SET>(n, n - 1) + cl;
}

template <class T> using cx2 = T __attribute__((ext_vector_type(2)));
// one scatterer's NR x NC block, its first row at b
template <class T, int SET, int NR, int NC>
__device__ __forceinline__ void keep_store(T *b, int E, const cx<T> (&Ft)[FM_NREP]) {
    constexpr KeepList KL = keep_list(SET, NR, NC);
#pragma unroll
    for (int q = 0; q < KL.n; ++q) {
        cx2<T> v;
        v.x = Ft[KL.k[q]].r;
        v.y = Ft[KL.k[q]].i;
        *reinterpret_cast<cx2<T> *>(b + (size_t)q * 2 * E) = v;
    }
}
template <class T, int SET, int NR, int NC>
__device__ __forceinline__ void keep_load(const T *b, int E, cx<T> (&Ft)[FM_NREP]) {
    constexpr KeepList KL = keep_list(SET, NR, NC);
#pragma unroll
    for (int k = 0; k < FM_NREP; ++k) Ft[k] = mk<T>(T(0), T(0));
#pragma unroll
    for (int q = 0; q < KL.n; ++q) {
        const cx2<T> v = *reinterpret_cast<const cx2<T> *>(b + (size_t)q * 2 * E);
        Ft[KL.k[q]] = mk<T>(v.x, v.y);
    }
}

// keep: the lane's pointer into the kept scattering matrices (written after each is built), or NULL.
template <class T, int SET>
__device__ cx<T> path_cchi(const fm_edge &ed, const PathLDS &g, const T *W, int ltab, int e, int n, T *keep = nullptr) {
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
    auto scat = [&](int s, int &ilx, const T *&tl, const T *&Wl) {
        const int ip = g.ipot[s];
        ilx = ed.lmax[(size_t)ip * E + ee] + 1;
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
        int ilx;
        const T *tl, *Wl;
        scat(1, ilx, tl, Wl);
        fblock<T, SET, L0, L0>(Ft, ck, (T)g.r[2], (T)g.r[1], ilx, Wl, tl, tlstride);
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
            int ilx;
            const T *tl, *Wl;
            scat(n - 1, ilx, tl, Wl);
            fblock<T, SET, L0, L, true>(Ft, ck, (T)g.r[n], (T)g.r[n - 1], ilx, Wl, tl, tlstride);
            if (keep) keep_store<T, SET, L0, L>(keep + (size_t)2 * E * keep_off<SET>(n, n - 1), E, Ft);
            row_phases<T, SET>(ph, g, n);
            chain(std::integral_constant<int, L0>{});
        }
#pragma unroll 1
        for (int s = n - 2; s >= 2; --s) {
            int ilx;
            const T *tl, *Wl;
            scat(s, ilx, tl, Wl);
            fblock<T, SET, L, L, true>(Ft, ck, (T)g.r[s + 1], (T)g.r[s], ilx, Wl, tl, tlstride);
            if (keep) keep_store<T, SET, L, L>(keep + (size_t)2 * E * keep_off<SET>(n, s), E, Ft);
            row_phases<T, SET>(ph, g, s + 1);
            chain(std::integral_constant<int, L>{});
        }
        // s = 1: F is L x L0; trace of U (L0 x L) F (L x L0)
        {
            int ilx;
            const T *tl, *Wl;
            scat(1, ilx, tl, Wl);
            fblock<T, SET, L, L0, true>(Ft, ck, (T)g.r[2], (T)g.r[1], ilx, Wl, tl, tlstride);
            if (keep) keep_store<T, SET, L, L0>(keep, E, Ft);
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
    return ptrac * cfac_at<T>(ed, g, ee, n);
}

// cchi at energy ee of a single-scattering path (fm_ss.h) of an edge with initial-state il = ILI
template <class T, int ILI>
__device__ cx<T> ss_cchi(const fm_edge &ed, const PathLDS &g, int ee) {
    const int E = ed.E;
    const int lxa = max(ed.lmax[(size_t)g.ipot[0] * E + ee], ed.lmax[(size_t)g.ipot[1] * E + ee]) + 1;
    const cx<T> ck = mk<T>((T)ed.ck[2 * ee], (T)ed.ck[2 * ee + 1]);
    const cx<T> z = minus_i_over(ck * (T)g.r[1]);
    cx<T> Tk[SS_MAXR], Fk[SS_MAXR];
    ss_reps<T, ILI, cx<T>>(ed, ee, g.ipot[1], z, lxa, Tk, Fk);
    const cx<T> ptrac = ss_trace<T, ILI, cx<T>>(g.eta, Tk, Fk);
    return ptrac * cfac_at<T>(ed, g, ee, 2);
}

// Fine-grid epilogue for the path held by the wave: amp/phase of cchi (lanes = energies, already in LDS as
// amp[] and ang[]), unwrap (FP64 scan), interpolation to q, chi = A sin(Phi + 2 q reff) exp(-2 sigma2 q^2).
template <class T>
__device__ void fine_grid(const fm_edge &ed, const int32_t *jq, const double *fq, const double *q, int nk, double reff_A,
                          double sigma2, T *amp, double *ang, T *out, const double *S = nullptr, int sstride = 0) {
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
        T A;
        double ph;
        if (S) {   // dense interpolation: row i of S (this edge's block)
            const double *Si = S + (size_t)i * sstride;
            double sa = 0.0, sp = 0.0;
            for (int e = 0; e < E; ++e) {
                sa += Si[e] * (double)amp[e];
                sp += Si[e] * ang[e];
            }
            A = (T)sa;
            ph = sp + 2.0 * q[i] * reff_A;
        } else {
            const int j = jq[i];
            const double f = fq[i];
            A = (T)((1.0 - f) * (double)amp[j] + f * (double)amp[j + 1]);
            ph = (1.0 - f) * ang[j] + f * ang[j + 1] + 2.0 * q[i] * reff_A;
        }
        ph -= FM_TWO_PI * rint(ph / FM_TWO_PI);
        T v;
        if constexpr (sizeof(T) == 4) v = A * sinf((float)ph);
        else v = A * sin(ph);
        if (sigma2 != 0.0) v *= (T)exp(-2.0 * sigma2 * q[i] * q[i]);
        out[i] = v;
    }
}


// chi_seg[s][k] = sum over the CSR list of segment s of deg[p] chi_path[p][k]: a workgroup per (64 k-points,
// segment); wave w sums the list entries w, w + 8, ... in order, then the 8 partial sums are added in wave order.
// Deterministic for a given list and grid.
constexpr int SEG_WAVES = 16;
template <class T>
__global__ __launch_bounds__(FM_WAVE *SEG_WAVES) void k_segment_sum(int nseg, int nk, const int32_t *seg_ptr,
                                                                     const int32_t *seg_paths, const double *deg,
                                                                     const T *chi_path, double *chi_seg) {
    __shared__ double part[SEG_WAVES][FM_WAVE];
    const int s = blockIdx.y;
    const int w = threadIdx.x / FM_WAVE, lane = threadIdx.x % FM_WAVE;
    const int k = blockIdx.x * FM_WAVE + lane;
    double acc = 0.0;
    if (k < nk) {
        const int t0 = seg_ptr[s] + w, t1 = seg_ptr[s + 1];
        constexpr int U = 16;   // paths whose loads are in flight together; summed in list order
        for (int tb = t0; tb < t1; tb += U * SEG_WAVES) {
            int pv[U];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const int t = tb + u * SEG_WAVES;
                pv[u] = t < t1 ? seg_paths[t] : -1;
            }
            // unconditional loads (past the list's end they read path 0, a valid row, and are not summed), so that
            // the compiler issues all U of them before the first wait instead of branching around each
            double dv[U], cv[U];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const int p = pv[u] < 0 ? 0 : pv[u];
                dv[u] = deg[p];
                cv[u] = (double)chi_path[(size_t)p * nk + k];
            }
#pragma unroll
            for (int u = 0; u < U; ++u)
                if (pv[u] >= 0) acc += dv[u] * cv[u];
        }
    }
    part[w][lane] = acc;
    __syncthreads();
    if (w == 0 && k < nk) {
        double tot = 0.0;
        for (int i = 0; i < SEG_WAVES; ++i) tot += part[i][lane];
        chi_seg[(size_t)s * nk + k] = tot;
    }
}

}  // namespace fm
