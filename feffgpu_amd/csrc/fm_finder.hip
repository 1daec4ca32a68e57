// FEFF's path finder on the GPU (src/feffgpu/pathfinder.py): for each absorber of a configuration, every closed path
// of 2..nlegmax legs whose total length is at most 2 rmax, its plane-wave importance, and the degeneracy totals of
// symmetry-equivalent multiple-scattering paths; the importance filter relative to the shortest single-scattering
// path. Everything is float32, as the reference computes it, with contraction of a * b + c into FMAs off so that the
// lengths, angles and importance round as the reference's separate operations do.
//
// Enumeration runs in two passes over work items: a count per item, an exclusive scan over all items of all
// absorbers, then a write at each item's offset, so that the paths come out in the reference's order (absorber, then
// the first atom, then the next ...). Paths of five and more legs grow from partial paths, one atom per level.
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stdio.h>
#include <string.h>

#include "../../include/feffgpu_amd.h"

#pragma clang fp contract(off)

int fm_set_err(int code, const char *msg);

namespace {

constexpr int FD_THREADS = 256;
constexpr int FD_NBETA = 40;
constexpr int FD_SCAN_ITEMS = 4096;     // items per block of the scan (1024 threads x 4)

__device__ __forceinline__ float norm3(float x, float y, float z) { return sqrtf(x * x + y * y + z * z); }

// minimum image of a float32 displacement (src/feffgpu/cell.py:26-35)
__device__ __forceinline__ void wrap(const fm_find_sys &s, float (&v)[3]) {
    if (s.cell_kind == 1) {
        const float L = s.cell[0];
#pragma unroll
        for (int c = 0; c < 3; ++c) v[c] = v[c] - L * rintf(v[c] / L);
    } else if (s.cell_kind == 2) {
        float f[3];
#pragma unroll
        for (int c = 0; c < 3; ++c) f[c] = v[0] * s.cellinv[0 * 3 + c] + v[1] * s.cellinv[1 * 3 + c] + v[2] * s.cellinv[2 * 3 + c];
#pragma unroll
        for (int c = 0; c < 3; ++c) f[c] = f[c] - rintf(f[c]);
#pragma unroll
        for (int c = 0; c < 3; ++c) v[c] = f[0] * s.cell[0 * 3 + c] + f[1] * s.cell[1 * 3 + c] + f[2] * s.cell[2 * 3 + c];
    }
}

// The candidates of one absorber: atoms of its configuration (natom atoms from a multiple of natom: positions hold
// configurations one after the other) other than the absorber within rmax, in index order (block-wide scan).
__global__ __launch_bounds__(FD_THREADS) void k_find_cand(fm_find_sys s, const int32_t *absorbers, int32_t maxc,
                                                          int32_t *cidx, float *crel, float *cd0, int32_t *ncand) {
    __shared__ int wsum[FD_THREADS / 64];
    __shared__ int base_count;
    const int b = blockIdx.x, tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int a = absorbers[b];
    const int first = (a / s.natom) * s.natom, last = first + s.natom;
    const float ax = s.pos[3 * a], ay = s.pos[3 * a + 1], az = s.pos[3 * a + 2];
    if (tid == 0) base_count = 0;
    __syncthreads();
    for (int c0 = first; c0 < last; c0 += FD_THREADS) {
        const int i = c0 + tid;
        bool keep = false;
        float v[3] = {0.f, 0.f, 0.f}, d = 0.f;
        if (i < last) {
            v[0] = s.pos[3 * i] - ax;
            v[1] = s.pos[3 * i + 1] - ay;
            v[2] = s.pos[3 * i + 2] - az;
            wrap(s, v);
            d = norm3(v[0], v[1], v[2]);
            keep = i != a && d <= s.rmax;
        }
        const unsigned long long m = __ballot(keep);
        const int before = __popcll(m & ((1ull << lane) - 1ull));
        if (lane == 0) wsum[wv] = __popcll(m);
        __syncthreads();
        int off = base_count;
        for (int w = 0; w < wv; ++w) off += wsum[w];
        if (keep && off + before < maxc) {
            const size_t k = (size_t)b * maxc + off + before;
            cidx[k] = i;
            crel[3 * k] = v[0];
            crel[3 * k + 1] = v[1];
            crel[3 * k + 2] = v[2];
            cd0[k] = d;
        }
        __syncthreads();
        if (tid == 0)
            for (int w = 0; w < FD_THREADS / 64; ++w) base_count += wsum[w];
        __syncthreads();
    }
    if (tid == 0) ncand[b] = base_count;
}

// Access to the intermediate atoms of absorber slot b: positions 0..nc-1 are the candidates, nc the absorber.
struct Inter {
    const float *rel;
    const float *d0;
    int nc;
    __device__ __forceinline__ void at(int j, float (&v)[3]) const {
        if (j < nc) { v[0] = rel[3 * j]; v[1] = rel[3 * j + 1]; v[2] = rel[3 * j + 2]; }
        else { v[0] = v[1] = v[2] = 0.f; }
    }
    __device__ __forceinline__ float back(int j) const { return j < nc ? d0[j] : 0.f; }
    __device__ __forceinline__ float dist(int i, int j) const {
        float p[3], q[3];
        at(i, p);
        at(j, q);
        return norm3(p[0] - q[0], p[1] - q[1], p[2] - q[2]);
    }
};

__device__ __forceinline__ Inter inter_of(const float *crel, const float *cd0, const int32_t *ncand, int maxc, int b) {
    Inter t;
    t.rel = crel + (size_t)b * maxc * 3;
    t.d0 = cd0 + (size_t)b * maxc;
    t.nc = ncand[b] < maxc ? ncand[b] : maxc;
    return t;
}

// Counts of 3- and 4-leg paths per work item: item (b, i) for three legs (partners j > i), item (b, i, k) for four
// legs (k an intermediate atom, the last atom l >= i); item offsets per absorber in off3 / off4 (items of absorber b
// start at off[b]).
__global__ __launch_bounds__(FD_THREADS) void k_find_count(int32_t nabs, const int64_t *off3, const int64_t *off4,
                                                           int32_t maxc, const float *crel, const float *cd0,
                                                           const int32_t *ncand, float two_r, int32_t nlegmax,
                                                           int32_t *cnt3, int32_t *cnt4) {
    const int b = blockIdx.y;
    const Inter I = inter_of(crel, cd0, ncand, maxc, b);
    const int nc = I.nc, ni = nc + 1;
    const int64_t t = (int64_t)blockIdx.x * FD_THREADS + threadIdx.x;
    if (nc < 2) return;    // one candidate: single scattering only (the absorber has no 3+-leg items)
    if (nlegmax >= 3 && t < nc) {
        const int i = (int)t;
        int c = 0;
        for (int j = i + 1; j < nc; ++j)
            if ((I.d0[i] + I.dist(i, j)) + I.d0[j] <= two_r) ++c;
        cnt3[off3[b] + i] = c;
    }
    if (nlegmax >= 4 && t < (int64_t)nc * ni) {
        const int i = (int)(t / ni), k = (int)(t % ni);
        int c = 0;
        if (k != i) {
            const float head = I.d0[i] + I.dist(i, k);
            for (int l = i; l < nc; ++l)
                if (l != k && ((head + I.dist(l, k)) + I.d0[l]) <= two_r) ++c;
        }
        cnt4[off4[b] + t] = c;
    }
}

struct PathOut {
    int64_t *atoms;     // [P][nleg - 1] global atom indices
    float *coords;      // [P][nleg][3] relative to the absorber, absorber first (0)
    int64_t *ipots;     // [P][nleg]
    float *deg;         // [P]
    int32_t *slot;      // [P] absorber slot
};

__device__ __forceinline__ void put_path(const PathOut &o, int64_t p, int nleg, int b, int absorber, const int32_t *cidx,
                                         const int32_t *ipot, const Inter &I, const int *loc, float deg) {
    float *c = o.coords + (size_t)p * nleg * 3;
    c[0] = c[1] = c[2] = 0.f;
    o.ipots[(size_t)p * nleg] = 0;
    for (int j = 1; j < nleg; ++j) {
        const int q = loc[j - 1];
        float v[3];
        I.at(q, v);
        c[3 * j] = v[0];
        c[3 * j + 1] = v[1];
        c[3 * j + 2] = v[2];
        const int atom = q < I.nc ? cidx[q] : absorber;
        o.atoms[(size_t)p * (nleg - 1) + j - 1] = atom;
        o.ipots[(size_t)p * nleg + j] = atom == absorber ? 0 : ipot[atom];
    }
    o.deg[p] = deg;
    o.slot[p] = b;
}

// The paths of two, three and four legs at the scanned offsets.
__global__ __launch_bounds__(FD_THREADS) void k_find_write(int32_t nabs, const int32_t *absorbers, const int64_t *off3,
                                                           const int64_t *off4, int32_t maxc, const int32_t *cidx,
                                                           const float *crel, const float *cd0, const int32_t *ncand,
                                                           const int32_t *ipot, float two_r, int32_t nlegmax,
                                                           const int64_t *off2, PathOut o2, const int64_t *pos3,
                                                           PathOut o3, const int64_t *pos4, PathOut o4) {
    const int b = blockIdx.y;
    const Inter I = inter_of(crel, cd0, ncand, maxc, b);
    const int nc = I.nc, ni = nc + 1, absorber = absorbers[b];
    const int32_t *ci = cidx + (size_t)b * maxc;
    const int64_t t = (int64_t)blockIdx.x * FD_THREADS + threadIdx.x;
    if (t < nc) {
        const int loc[1] = {(int)t};
        put_path(o2, off2[b] + t, 2, b, absorber, ci, ipot, I, loc, 1.f);
    }
    if (nc < 2) return;    // one candidate: single scattering only (no 3+-leg items)
    if (nlegmax >= 3 && t < nc) {
        const int i = (int)t;
        int64_t p = pos3[off3[b] + i];
        for (int j = i + 1; j < nc; ++j)
            if ((I.d0[i] + I.dist(i, j)) + I.d0[j] <= two_r) {
                const int loc[2] = {i, j};
                put_path(o3, p++, 3, b, absorber, ci, ipot, I, loc, 2.f);
            }
    }
    if (nlegmax >= 4 && t < (int64_t)nc * ni) {
        const int i = (int)(t / ni), k = (int)(t % ni);
        if (k != i) {
            int64_t p = pos4[off4[b] + t];
            const float head = I.d0[i] + I.dist(i, k);
            for (int l = i; l < nc; ++l)
                if (l != k && ((head + I.dist(l, k)) + I.d0[l]) <= two_r) {
                    const int loc[3] = {i, k, l};
                    put_path(o4, p++, 4, b, absorber, ci, ipot, I, loc, i == l ? 1.f : 2.f);
                }
        }
    }
}

// Partial paths of the 5+-leg extension: m atoms (positions in the absorber's intermediate list), their length
// from the absorber, the absorber slot.
struct Partial {
    int16_t *seq;      // [P][FM_MAX_LEGS - 1]
    float *len;
    int32_t *slot;
};

__global__ void k_find_seed(int32_t nabs, const int64_t *off2, int32_t maxc, const float *cd0, const int32_t *ncand,
                            Partial out) {
    const int b = blockIdx.y;
    const int nc = ncand[b] < maxc ? ncand[b] : maxc;
    const int i = blockIdx.x * FD_THREADS + threadIdx.x;
    if (i >= nc || nc < 2) return;    // one candidate: no 5+-leg paths (the host drops its seed slot)
    const int64_t p = off2[b] + i;
    out.seq[p * (FM_MAX_LEGS - 1)] = (int16_t)i;
    out.len[p] = cd0[(size_t)b * maxc + i];
    out.slot[p] = b;
}

// One more atom for every partial path of m atoms: counts (mode 0) or the new partials at the scanned offsets.
__global__ void k_find_extend(int64_t npart, int32_t m, Partial in, int32_t maxc, const float *crel, const float *cd0,
                              const int32_t *ncand, float two_r, int32_t *cnt, const int64_t *pos, Partial out) {
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= npart) return;
    const int b = in.slot[p];
    const Inter I = inter_of(crel, cd0, ncand, maxc, b);
    const int16_t *sq = in.seq + p * (FM_MAX_LEGS - 1);
    const int last = sq[m - 1];
    const float len = in.len[p];
    int c = 0;
    int64_t q = pos ? pos[p] : 0;
    for (int j = 0; j <= I.nc; ++j) {
        if (j == last) continue;
        const float tot = len + I.dist(last, j);
        if (tot + I.back(j) <= two_r) {
            if (pos) {
                int16_t *d = out.seq + q * (FM_MAX_LEGS - 1);
                for (int k = 0; k < m; ++k) d[k] = sq[k];
                d[m] = (int16_t)j;
                out.len[q] = tot;
                out.slot[q] = b;
                ++q;
            } else {
                ++c;
            }
        }
    }
    if (!pos) cnt[p] = c;
}

// Completed paths among partials of m atoms (the last atom a scatterer), the one of a path and its reverse whose
// sequence comes first (degeneracy 2; 1 for a palindrome): counts (pos NULL) or written at the scanned offsets.
__global__ void k_find_emit(int64_t npart, int32_t m, Partial in, const int32_t *absorbers, int32_t maxc,
                            const int32_t *cidx, const float *crel, const float *cd0, const int32_t *ncand,
                            const int32_t *ipot, int32_t *cnt, const int64_t *pos, PathOut o) {
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= npart) return;
    const int b = in.slot[p];
    const Inter I = inter_of(crel, cd0, ncand, maxc, b);
    const int16_t *sq = in.seq + p * (FM_MAX_LEGS - 1);
    bool keep = sq[m - 1] != I.nc;
    bool pal = true;
    if (keep) {
        for (int k = 0; k < m; ++k) {
            const int f = sq[k], r = sq[m - 1 - k];
            if (f != r) {
                pal = false;
                keep = f < r;
                break;
            }
        }
    }
    if (!pos) {
        cnt[p] = keep ? 1 : 0;
        return;
    }
    if (!keep) return;
    int loc[FM_MAX_LEGS - 1];
    for (int k = 0; k < m; ++k) loc[k] = sq[k];
    put_path(o, pos[p], m + 1, b, absorbers[b], cidx + (size_t)b * maxc, ipot, I, loc, pal ? 1.f : 2.f);
}

// Block scan of FD_SCAN_ITEMS counts per block (exclusive, int64), block totals in sums.
__global__ __launch_bounds__(1024) void k_scan_local(const int32_t *in, int64_t n, int64_t *out, int64_t *sums) {
    __shared__ int64_t ws[16];
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int64_t base = (int64_t)blockIdx.x * FD_SCAN_ITEMS + (int64_t)tid * 4;
    int64_t v[4], s = 0;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        v[k] = base + k < n ? in[base + k] : 0;
        s += v[k];
    }
    int64_t x = s;     // inclusive wave scan of the thread sums
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const int64_t y = __shfl_up(x, d, 64);
        if (lane >= d) x += y;
    }
    if (lane == 63) ws[wv] = x;
    __syncthreads();
    int64_t wave_off = 0;
    for (int w = 0; w < wv; ++w) wave_off += ws[w];
    int64_t run = wave_off + x - s;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        if (base + k < n) out[base + k] = run;
        run += v[k];
    }
    if (tid == 1023) sums[blockIdx.x] = run;
}

// Exclusive scan of the block totals (one block, up to 4096 totals) and their addition to every block.
__global__ __launch_bounds__(1024) void k_scan_sums(int64_t *sums, int64_t nb, int64_t *total) {
    __shared__ int64_t ws[16];
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    int64_t v[4], s = 0;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const int64_t i = (int64_t)tid * 4 + k;
        v[k] = i < nb ? sums[i] : 0;
        s += v[k];
    }
    int64_t x = s;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const int64_t y = __shfl_up(x, d, 64);
        if (lane >= d) x += y;
    }
    if (lane == 63) ws[wv] = x;
    __syncthreads();
    int64_t wave_off = 0;
    for (int w = 0; w < wv; ++w) wave_off += ws[w];
    int64_t run = wave_off + x - s;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const int64_t i = (int64_t)tid * 4 + k;
        if (i < nb) sums[i] = run;
        run += v[k];
    }
    if (tid == 1023) *total = run;
}

__global__ void k_scan_add(int64_t *out, int64_t n, const int64_t *sums) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] += sums[i / FD_SCAN_ITEMS];
}

// The plane-wave importance of each path (src/feffgpu/pathfinder.py:72-94) and its total length.
__global__ void k_find_importance(fm_find_tables t, int64_t npath, int32_t nleg, const float *coords,
                                  const int64_t *ipots, float *xport, float *rtot_out) {
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= npath) return;
    const int m = nleg - 1;
    const float *c = coords + (size_t)p * nleg * 3;
    float u[FM_MAX_LEGS][3], ri[FM_MAX_LEGS];
    float rtot = 0.f;
    for (int j = 0; j < nleg; ++j) {     // leg j: atom j -> atom j + 1 (atom nleg = atom 0)
        const int a1 = j + 1 < nleg ? j + 1 : 0;
        float l[3];
        for (int k = 0; k < 3; ++k) l[k] = c[3 * a1 + k] - c[3 * j + k];
        ri[j] = norm3(l[0], l[1], l[2]);
        rtot = rtot + ri[j];
        const float den = ri[j] > 1e-12f ? ri[j] : 1e-12f;
        for (int k = 0; k < 3; ++k) u[j][k] = l[k] / den;
    }
    float cosv[FM_MAX_LEGS];
    for (int j = 0; j < m; ++j) cosv[j] = u[j][0] * u[j + 1][0] + u[j][1] * u[j + 1][1] + u[j][2] * u[j + 1][2];
    const float cos_abs = u[m][0] * u[0][0] + u[m][1] * u[0][1] + u[m][2] * u[0][2];
    int bin[FM_MAX_LEGS];
    for (int j = 0; j < m; ++j) {
        const float a = fabsf(cosv[j]);
        float n = floorf(a / 0.025f);
        if (a - n * 0.025f > 0.0125f) n = n + 1.f;
        if (cosv[j] < 0.f) n = -n;
        int k = (int)n + FD_NBETA;
        bin[j] = k < 0 ? 0 : (k > 2 * FD_NBETA ? 2 * FD_NBETA : k);
    }
    const float cmax = fabsf(cos_abs) > 0.3f ? fabsf(cos_abs) : 0.3f;
    const int64_t *ip = ipots + (size_t)p * nleg;
    float acc = 0.f;
    for (int e = t.ik0; e < t.E; ++e) {
        const float ck = t.cksp[e];
        float crit = cmax / (ri[m] * ck);
        for (int j = 0; j < m; ++j) {
            const float f = t.fbeta[((size_t)e * t.nph1 + (int)ip[j + 1]) * (2 * FD_NBETA + 1) + bin[j]];
            crit = (crit * f) / (ri[j] * ck);
        }
        crit = crit * expf(-rtot / t.xlam[e]);
        acc = acc + fabsf(crit) * t.trap[e - t.ik0];
    }
    xport[p] = acc;
    if (rtot_out) rtot_out[p] = rtot;
}

// FNV-1a over 32-bit words
__device__ __forceinline__ uint64_t fnv(uint64_t h, uint32_t w) {
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        h ^= (w >> (8 * k)) & 0xffu;
        h *= 1099511628211ull;
    }
    return h;
}

// The group of each path among paths of equal leg lengths and angles (to 1e-3, in either direction) of the same
// absorber slot: a 64-bit hash of the canonical key, found or inserted in an open-addressing table; the path's
// degeneracy added to its group's total (sums of small integers: exact in any order).
__global__ void k_find_group(int64_t npath, int32_t nleg, const float *coords, const float *deg, const int32_t *slot,
                             unsigned long long *tags, float *sums, uint64_t mask, int64_t *where) {
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= npath) return;
    const float *c = coords + (size_t)p * nleg * 3;
    float u[FM_MAX_LEGS][3], ri[FM_MAX_LEGS];
    for (int j = 0; j < nleg; ++j) {
        const int a1 = j + 1 < nleg ? j + 1 : 0;
        float l[3];
        for (int k = 0; k < 3; ++k) l[k] = c[3 * a1 + k] - c[3 * j + k];
        ri[j] = norm3(l[0], l[1], l[2]);
        const float den = ri[j] > 1e-12f ? ri[j] : 1e-12f;
        for (int k = 0; k < 3; ++k) u[j][k] = l[k] / den;
    }
    int kf[2 * FM_MAX_LEGS], kr[2 * FM_MAX_LEGS];
    float cb[FM_MAX_LEGS];
    for (int j = 0; j < nleg; ++j) {
        const int j1 = j + 1 < nleg ? j + 1 : 0;
        cb[j] = u[j][0] * u[j1][0] + u[j][1] * u[j1][1] + u[j][2] * u[j1][2];
    }
    for (int j = 0; j < nleg; ++j) {
        kf[j] = (int)rintf(ri[j] * 1000.f);
        kf[nleg + j] = (int)rintf(cb[j] * 1000.f);
        kr[j] = (int)rintf(ri[nleg - 1 - j] * 1000.f);
    }
    for (int j = 0; j < nleg - 1; ++j) kr[nleg + j] = (int)rintf(cb[nleg - 2 - j] * 1000.f);
    kr[2 * nleg - 1] = kf[2 * nleg - 1];
    bool less = false;
    for (int j = 0; j < 2 * nleg; ++j)
        if (kr[j] != kf[j]) { less = kr[j] < kf[j]; break; }
    uint64_t h = 14695981039346656037ull;
    h = fnv(h, (uint32_t)slot[p]);
    h = fnv(h, (uint32_t)nleg);
    for (int j = 0; j < 2 * nleg; ++j) h = fnv(h, (uint32_t)(less ? kr[j] : kf[j]));
    if (h == 0) h = 1;
    uint64_t i = h & mask;
    while (true) {
        const unsigned long long prev = atomicCAS(&tags[i], 0ull, (unsigned long long)h);
        if (prev == 0ull || prev == h) break;
        i = (i + 1) & mask;
    }
    atomicAdd(&sums[i], deg[p]);
    where[p] = (int64_t)i;
}

__global__ void k_find_clear(unsigned long long *tags, float *sums, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) { tags[i] = 0ull; sums[i] = 0.f; }
}

__global__ void k_find_group_read(int64_t npath, const int64_t *where, const float *sums, float *total) {
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p < npath) total[p] = sums[where[p]];
}

// keep[p] = 100 dtot xp / ref[slot] >= critpw, in float32 as the reference evaluates it
__global__ void k_find_filter(int64_t npath, const float *dtot, const float *xp, const int32_t *slot, const float *ref,
                              float critpw, uint8_t *keep) {
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= npath) return;
    const float frac = ((100.f * dtot[p]) * xp[p]) / ref[slot[p]];
    keep[p] = frac >= critpw ? 1 : 0;
}

int launch_err(const char *what) {
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) return FM_OK;
    static thread_local char buf[256];
    snprintf(buf, sizeof buf, "%s: %s", what, hipGetErrorString(e));
    return fm_set_err(FM_EHIP, buf);
}

unsigned blocks_of(int64_t n, int t) { return (unsigned)((n + t - 1) / t); }

}  // namespace

extern "C" {

int fm_find_candidates(const fm_find_sys *sys, int32_t nabs, const int32_t *absorbers, int32_t maxc, int32_t *cidx,
                       float *crel, float *cd0, int32_t *ncand, void *stream) {
    if (!sys || nabs < 0 || maxc < 1) return fm_set_err(FM_EINVAL, "fm_find_candidates: bad arguments");
    if (nabs == 0) return FM_OK;
    if (!sys->pos || !absorbers || !cidx || !crel || !cd0 || !ncand) return fm_set_err(FM_EINVAL, "fm_find_candidates: NULL buffer");
    if (sys->cell_kind < 0 || sys->cell_kind > 2) return fm_set_err(FM_EINVAL, "fm_find_candidates: cell_kind 0..2");
    hipLaunchKernelGGL(k_find_cand, dim3(nabs), dim3(FD_THREADS), 0, (hipStream_t)stream, *sys, absorbers, maxc, cidx,
                       crel, cd0, ncand);
    return launch_err("fm_find_candidates");
}

int fm_find_count(int32_t nabs, const int64_t *off3, const int64_t *off4, int32_t maxc, const float *crel,
                  const float *cd0, const int32_t *ncand, int32_t ncmax, float rmax, int32_t nlegmax, int32_t *cnt3,
                  int32_t *cnt4, void *stream) {
    if (nabs <= 0 || ncmax <= 0 || nlegmax < 3) return FM_OK;
    if (ncmax > maxc) return fm_set_err(FM_EINVAL, "fm_find_count: ncmax > maxc");
    const int64_t items = (int64_t)ncmax * (nlegmax >= 4 ? ncmax + 1 : 1);
    hipLaunchKernelGGL(k_find_count, dim3(blocks_of(items, FD_THREADS), nabs), dim3(FD_THREADS), 0, (hipStream_t)stream,
                       nabs, off3, off4, maxc, crel, cd0, ncand, 2.f * rmax, nlegmax, cnt3, cnt4);
    return launch_err("fm_find_count");
}

size_t fm_scan_workspace_bytes(int64_t n) { return sizeof(int64_t) * ((n + FD_SCAN_ITEMS - 1) / FD_SCAN_ITEMS + 2); }

int fm_exclusive_scan(const int32_t *in, int64_t n, int64_t *out, int64_t *total, void *workspace, size_t bytes,
                      void *stream) {
    if (n < 0) return fm_set_err(FM_EINVAL, "fm_exclusive_scan: n < 0");
    const int64_t nb = (n + FD_SCAN_ITEMS - 1) / FD_SCAN_ITEMS;
    if (nb > FD_SCAN_ITEMS) return fm_set_err(FM_EINVAL, "fm_exclusive_scan: more than 2^24 items");
    if (bytes < fm_scan_workspace_bytes(n)) return fm_set_err(FM_EINVAL, "fm_exclusive_scan: workspace too small");
    int64_t *sums = (int64_t *)workspace;
    hipStream_t st = (hipStream_t)stream;
    if (nb > 0) hipLaunchKernelGGL(k_scan_local, dim3((unsigned)nb), dim3(1024), 0, st, in, n, out, sums);
    hipLaunchKernelGGL(k_scan_sums, dim3(1), dim3(1024), 0, st, sums, nb, total);
    if (n > 0) hipLaunchKernelGGL(k_scan_add, dim3(blocks_of(n, 256)), dim3(256), 0, st, out, n, sums);
    return launch_err("fm_exclusive_scan");
}

int fm_find_write(int32_t nabs, const int32_t *absorbers, const int64_t *off3, const int64_t *off4, int32_t maxc,
                  const int32_t *cidx, const float *crel, const float *cd0, const int32_t *ncand, int32_t ncmax,
                  const int32_t *ipot, float rmax, int32_t nlegmax, const int64_t *off2, const fm_find_out *o2,
                  const int64_t *pos3, const fm_find_out *o3, const int64_t *pos4, const fm_find_out *o4, void *stream) {
    if (nabs <= 0 || ncmax <= 0) return FM_OK;
    if (!o2 || (nlegmax >= 3 && (!o3 || !pos3)) || (nlegmax >= 4 && (!o4 || !pos4)))
        return fm_set_err(FM_EINVAL, "fm_find_write: NULL output");
    auto cv = [](const fm_find_out *o) {
        PathOut p;
        memset(&p, 0, sizeof p);
        if (o) { p.atoms = o->atoms; p.coords = o->coords; p.ipots = o->ipots; p.deg = o->deg; p.slot = o->slot; }
        return p;
    };
    const int64_t items = (int64_t)ncmax * (nlegmax >= 4 ? ncmax + 1 : 1);
    hipLaunchKernelGGL(k_find_write, dim3(blocks_of(items, FD_THREADS), nabs), dim3(FD_THREADS), 0, (hipStream_t)stream,
                       nabs, absorbers, off3, off4, maxc, cidx, crel, cd0, ncand, ipot, 2.f * rmax, nlegmax, off2, cv(o2),
                       pos3, cv(o3), pos4, cv(o4));
    return launch_err("fm_find_write");
}

int fm_find_seed(int32_t nabs, const int64_t *off2, int32_t maxc, const float *cd0, const int32_t *ncand,
                 int32_t ncmax, const fm_find_partial *out, void *stream) {
    if (nabs <= 0 || ncmax <= 0) return FM_OK;
    Partial p = {out->seq, out->len, out->slot};
    hipLaunchKernelGGL(k_find_seed, dim3(blocks_of(ncmax, FD_THREADS), nabs), dim3(FD_THREADS), 0, (hipStream_t)stream,
                       nabs, off2, maxc, cd0, ncand, p);
    return launch_err("fm_find_seed");
}

int fm_find_extend(int64_t npart, int32_t m, const fm_find_partial *in, int32_t maxc, const float *crel,
                   const float *cd0, const int32_t *ncand, float rmax, int32_t *cnt, const int64_t *pos,
                   const fm_find_partial *out, void *stream) {
    if (npart <= 0) return FM_OK;
    if (m < 1 || m >= FM_MAX_LEGS - 1) return fm_set_err(FM_EINVAL, "fm_find_extend: m out of range");
    if (!pos && !cnt) return fm_set_err(FM_EINVAL, "fm_find_extend: NULL counts");
    Partial a = {in->seq, in->len, in->slot};
    Partial b = {nullptr, nullptr, nullptr};
    if (pos) b = {out->seq, out->len, out->slot};
    hipLaunchKernelGGL(k_find_extend, dim3(blocks_of(npart, 256)), dim3(256), 0, (hipStream_t)stream, npart, m, a, maxc,
                       crel, cd0, ncand, 2.f * rmax, cnt, pos, b);
    return launch_err("fm_find_extend");
}

int fm_find_emit(int64_t npart, int32_t m, const fm_find_partial *in, const int32_t *absorbers, int32_t maxc,
                 const int32_t *cidx, const float *crel, const float *cd0, const int32_t *ncand, const int32_t *ipot,
                 int32_t *cnt, const int64_t *pos, const fm_find_out *o, void *stream) {
    if (npart <= 0) return FM_OK;
    if (m < 4 || m > FM_MAX_LEGS - 1) return fm_set_err(FM_EINVAL, "fm_find_emit: m out of range 4..8");
    Partial a = {in->seq, in->len, in->slot};
    PathOut p;
    memset(&p, 0, sizeof p);
    if (pos) { p.atoms = o->atoms; p.coords = o->coords; p.ipots = o->ipots; p.deg = o->deg; p.slot = o->slot; }
    hipLaunchKernelGGL(k_find_emit, dim3(blocks_of(npart, 256)), dim3(256), 0, (hipStream_t)stream, npart, m, a,
                       absorbers, maxc, cidx, crel, cd0, ncand, ipot, cnt, pos, p);
    return launch_err("fm_find_emit");
}

int fm_find_importance(const fm_find_tables *t, int64_t npath, int32_t nleg, const float *coords, const int64_t *ipots,
                       float *xport, float *rtot, void *stream) {
    if (npath <= 0) return FM_OK;
    if (!t || !t->fbeta || !t->cksp || !t->xlam || !t->trap || !coords || !ipots || !xport)
        return fm_set_err(FM_EINVAL, "fm_find_importance: NULL buffer");
    if (nleg < 2 || nleg > FM_MAX_LEGS) return fm_set_err(FM_EINVAL, "fm_find_importance: nleg out of range 2..9");
    hipLaunchKernelGGL(k_find_importance, dim3(blocks_of(npath, 256)), dim3(256), 0, (hipStream_t)stream, *t, npath,
                       nleg, coords, ipots, xport, rtot);
    return launch_err("fm_find_importance");
}

size_t fm_find_group_workspace_bytes(int64_t npath) {
    uint64_t size = 1024;
    while (size < 2 * (uint64_t)npath) size <<= 1;
    return size * (sizeof(unsigned long long) + sizeof(float)) + 8 * (size_t)npath + 256;
}

int fm_find_degenerate_total(int64_t npath, int32_t nleg, const float *coords, const float *deg, const int32_t *slot,
                             float *total, void *workspace, size_t bytes, void *stream) {
    if (npath <= 0) return FM_OK;
    if (nleg < 2 || nleg > FM_MAX_LEGS) return fm_set_err(FM_EINVAL, "fm_find_degenerate_total: nleg out of range");
    if (!coords || !deg || !slot || !total || !workspace) return fm_set_err(FM_EINVAL, "fm_find_degenerate_total: NULL buffer");
    if (bytes < fm_find_group_workspace_bytes(npath)) return fm_set_err(FM_EINVAL, "fm_find_degenerate_total: workspace too small");
    uint64_t size = 1024;
    while (size < 2 * (uint64_t)npath) size <<= 1;
    unsigned long long *tags = (unsigned long long *)workspace;
    float *sums = (float *)(tags + size);
    int64_t *where = (int64_t *)(((uintptr_t)(sums + size) + 7) & ~(uintptr_t)7);
    hipStream_t st = (hipStream_t)stream;
    hipLaunchKernelGGL(k_find_clear, dim3(blocks_of((int64_t)size, 256)), dim3(256), 0, st, tags, sums, (int64_t)size);
    hipLaunchKernelGGL(k_find_group, dim3(blocks_of(npath, 256)), dim3(256), 0, st, npath, nleg, coords, deg, slot, tags,
                       sums, (uint64_t)(size - 1), where);
    hipLaunchKernelGGL(k_find_group_read, dim3(blocks_of(npath, 256)), dim3(256), 0, st, npath, where, sums, total);
    return launch_err("fm_find_degenerate_total");
}

int fm_find_filter(int64_t npath, const float *dtot, const float *xport, const int32_t *slot, const float *ref,
                   float critpw, uint8_t *keep, void *stream) {
    if (npath <= 0) return FM_OK;
    if (!dtot || !xport || !slot || !ref || !keep) return fm_set_err(FM_EINVAL, "fm_find_filter: NULL buffer");
    hipLaunchKernelGGL(k_find_filter, dim3(blocks_of(npath, 256)), dim3(256), 0, (hipStream_t)stream, npath, dtot,
                       xport, slot, ref, critpw, keep);
    return launch_err("fm_find_filter");
}

}  // extern "C"
