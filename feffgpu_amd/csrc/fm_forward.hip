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
#include "fm_path.h"

using namespace fm;

thread_local char g_err[512] = "";
int fm_set_err(int code, const char *msg) {
    snprintf(g_err, sizeof g_err, "%s", msg);
    return code;
}

namespace {

struct KArgs {
    fm_edge ed[FM_MAX_EDGE];
    int nedge;
    int npath_bucket;     // paths in this launch
    int first;            // offset into order[]
    const int32_t *order;
    PathSrc src;
    void *cchi_out;       // mode 0
    fm_grid G;            // mode 1
    void *chi_path;
    double *reff_out;
    int ltab;             // L1 + 1 rows per rotation table
    int emax;             // largest E over the edges
    int edge_id;          // the edge of every path of this launch
    size_t path_bytes;    // LDS per path of a workgroup
};

int set_err(int code, const char *msg) {
    snprintf(g_err, sizeof g_err, "%s", msg);
    return code;
}
int hip_err(hipError_t e, const char *where) {
    snprintf(g_err, sizeof g_err, "%s: %s", where, hipGetErrorString(e));
    return FM_EHIP;
}


// A workgroup holds G = ed.gpaths paths and 4 wavefronts. Setup: wave j < G loads path j (geometry, rotation tables)
// into its LDS area. Energies: wave w takes the energy band [gstart[w], gstart[w+1]) of all G paths, lane
// (path j, energy i) = j * (64 / G) + i: the lanes of a wave share similar partial-wave counts. Epilogue: wave j
// interpolates path j to the output grid.
template <class T, int SET>
__global__ __launch_bounds__(FM_WAVE * 4) __attribute__((amdgpu_waves_per_eu(FM_WPE(T)))) void k_forward(KArgs a) {
    extern __shared__ __align__(16) unsigned char smem[];
    const int wid = threadIdx.x / FM_WAVE;
    const int lane = lane_id();
    const fm_edge &ed = a.ed[a.edge_id];
    const int G = ed.gpaths, cap = FM_WAVE / G;
    const int ltab = ed.L1 + 1, E = ed.E;
    const size_t off_W = align16(sizeof(PathLDS));
    const int nl = a.src.mode == 0 ? a.src.nleg : a.src.P.nmax;     // rotation tables sized for the longest path
    const size_t off_amp = off_W + align16(sizeof(T) * (size_t)nl * ltab * FM_NREP);
    const size_t off_ang = off_amp + align16(sizeof(T) * (size_t)E);
    __shared__ int pidx[4];
    __shared__ int pval[4];
    for (int blk = blockIdx.x; blk * G < a.npath_bucket; blk += gridDim.x) {
        // ---- setup: one wave per path
        if (wid < G) {
            const int slot = blk * G + wid;
            const bool valid = slot < a.npath_bucket;
            const int s0 = valid ? slot : blk * G;
            const int p = a.order ? a.order[a.first + s0] : a.first + s0;
            unsigned char *base = smem + (size_t)wid * a.path_bytes;
            PathLDS &g = *reinterpret_cast<PathLDS *>(base);
            int ei;
            const int n = load_path(a.src, p, g, ei);
            build_rotation<T, SET>(ed, g, n, ltab, reinterpret_cast<T *>(base + off_W));
            if (lane == 0) { pidx[wid] = p; pval[wid] = valid ? 1 : 0; }
        }
        __syncthreads();
        // ---- energies: band wid of every path
        {
            const int e0 = ed.gstart[wid], nb = ed.gstart[wid + 1] - e0;
            const int j = lane / cap, i = lane % cap;
            const bool active = nb > 0 && i < nb && j < G;
            const int jj = j < G ? j : G - 1;
            const int e = e0 + (i < nb ? i : (nb > 0 ? nb - 1 : 0));
            unsigned char *base = smem + (size_t)jj * a.path_bytes;
            const PathLDS &g = *reinterpret_cast<const PathLDS *>(base);
            const T *W = reinterpret_cast<const T *>(base + off_W);
            const int ee = e < E ? e : E - 1;
            const cx<T> c = path_cchi<T, SET>(ed, g, W, ltab, ee, g.n);
            if (active && pval[jj]) {
                if (a.src.mode == 0) {
                    T *o = reinterpret_cast<T *>(a.cchi_out) + ((size_t)pidx[jj] * E + e) * 2;
                    o[0] = c.r;
                    o[1] = c.i;
                } else {
                    T *amp = reinterpret_cast<T *>(base + off_amp);
                    double *ang = reinterpret_cast<double *>(base + off_ang);
                    amp[e] = (T)sqrt((double)c.r * c.r + (double)c.i * c.i + 1e-30);
                    ang[e] = (double)atan2(c.i, c.r);
                }
            }
        }
        __syncthreads();
        // ---- epilogue: one wave per path
        if (wid < G && pval[wid]) {
            unsigned char *base = smem + (size_t)wid * a.path_bytes;
            const PathLDS &g = *reinterpret_cast<const PathLDS *>(base);
            const int p = pidx[wid];
            const double reff_A = g.reff * FM_BOHR;
            if (lane == 0 && a.reff_out) a.reff_out[p] = reff_A;
            if (a.src.mode == 1) {
                const int nk = a.G.nk;
                T *out = reinterpret_cast<T *>(a.chi_path) + (size_t)p * nk;
                fine_grid<T>(ed, a.G.jq + (size_t)a.edge_id * nk, a.G.fq + (size_t)a.edge_id * nk, a.G.q, nk, reff_A,
                             a.src.P.sigma2 ? a.src.P.sigma2[p] : 0.0, reinterpret_cast<T *>(base + off_amp),
                             reinterpret_cast<double *>(base + off_ang), out);
            }
        }
        __syncthreads();
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

// chi_seg[s][k] = sum over the CSR list of segment s of deg[p] chi_path[p][k]: a workgroup per (64 k-points,
// segment); wave w sums the list entries w, w + 8, ... in order, then the 8 partial sums are added in wave order.
// Deterministic for a given list and grid.
constexpr int SEG_WAVES = 8;
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
        for (int t = seg_ptr[s] + w; t < seg_ptr[s + 1]; t += SEG_WAVES) {
            const int p = seg_paths[t];
            acc += deg[p] * (double)chi_path[(size_t)p * nk + k];
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

// ------------------------------------------------------------------------------------------------ launch helpers
int set_for(const fm_edge &ed, int nleg) {
    if (ed.ilinit <= 1) return nleg == 2 ? 2 : 0;
    return nleg == 2 ? -1 : 1;
}

template <class T, int SET> hipError_t launch_forward(KArgs &a, hipStream_t st, int nlegmax) {
    const fm_edge &ed = a.ed[a.edge_id];
    a.ltab = ed.L1 + 1;
    a.emax = ed.E;
    const int nl = a.src.mode == 0 ? a.src.nleg : a.src.P.nmax;
    (void)nlegmax;
    a.path_bytes = align16(sizeof(PathLDS)) + align16(sizeof(T) * (size_t)nl * a.ltab * FM_NREP) + align16(sizeof(T) * ed.E) +
                   align16(sizeof(double) * ed.E);
    const size_t shm = a.path_bytes * ed.gpaths;
    if (shm > 160 * 1024) return hipErrorInvalidValue;
    if (shm > 64 * 1024) {
        hipError_t e = hipFuncSetAttribute((const void *)k_forward<T, SET>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)shm);
        if (e != hipSuccess) return e;
    }
    const int blocks = (a.npath_bucket + ed.gpaths - 1) / ed.gpaths;
    if (blocks == 0) return hipSuccess;
    hipLaunchKernelGGL((k_forward<T, SET>), dim3(blocks), dim3(FM_WAVE * 4), shm, st, a);
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
    if (e->gpaths != 1 && e->gpaths != 2 && e->gpaths != 4) return set_err(FM_EINVAL, "edge: gpaths must be 1, 2 or 4");
    if (e->gstart[0] != 0 || e->gstart[4] != e->E) return set_err(FM_EINVAL, "edge: energy bands must cover 0..E");
    for (int w = 0; w < 4; ++w)
        if (e->gstart[w + 1] < e->gstart[w] || e->gstart[w + 1] - e->gstart[w] > FM_WAVE / e->gpaths)
            return set_err(FM_EINVAL, "edge: an energy band is longer than 64 / gpaths");
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
    a.src.mode = 0;
    a.npath_bucket = npath;
    a.first = 0;
    a.order = nullptr;
    a.src.nleg = nleg;
    a.src.coords = coords;
    a.src.ipots64 = ipots;
    a.cchi_out = cchi;
    a.reff_out = reff;
    a.edge_id = 0;
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
                         const int32_t *bucket_offset, const int32_t *bucket_set, const int32_t *bucket_edge,
                         const int32_t *bucket_nlegmax, int32_t nseg, const int32_t *seg_ptr, const int32_t *seg_paths, double *chi_seg,
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
        a.src.mode = 1;
        a.npath_bucket = cnt;
        a.first = bucket_offset[b];
        a.order = order;
        a.src.P = *paths;
        a.G = *grid;
        a.chi_path = chi_path;
        a.reff_out = reff;
        a.edge_id = bucket_edge ? bucket_edge[b] : 0;
        if (a.edge_id < 0 || a.edge_id >= nedge) return set_err(FM_EINVAL, "bucket edge out of range");
        int nlm = bucket_nlegmax ? bucket_nlegmax[b] : paths->nmax;
        if (nlm < 2 || nlm > paths->nmax) nlm = paths->nmax;
        hipError_t e = precision == 32 ? dispatch<float>(a, set, st, nlm) : dispatch<double>(a, set, st, nlm);
        if (e != hipSuccess) return hip_err(e, "fm_chi_paths launch");
    }
    if (chi_seg && nseg > 0) {
        if (!seg_ptr || !seg_paths) return set_err(FM_EINVAL, "NULL segment lists");
        dim3 grd((grid->nk + FM_WAVE - 1) / FM_WAVE, nseg);
        if (precision == 32)
            hipLaunchKernelGGL(k_segment_sum<float>, grd, dim3(FM_WAVE * SEG_WAVES), 0, st, nseg, grid->nk, seg_ptr, seg_paths,
                               paths->deg, (const float *)chi_path, chi_seg);
        else
            hipLaunchKernelGGL(k_segment_sum<double>, grd, dim3(FM_WAVE * SEG_WAVES), 0, st, nseg, grid->nk, seg_ptr, seg_paths,
                               paths->deg, (const double *)chi_path, chi_seg);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return hip_err(e, "segment sum launch");
    }
    return FM_OK;
}

}  // extern "C"
