// Fused path-stage forward kernels for gfx950 and the C ABI that launches them.
//
// A workgroup of 4 wavefronts holds G paths (1, 2, 4 or 8: as many as fit the LDS) of the same edge and leg count.
// Per path one wavefront computes the leg geometry in FP64 (one lane per leg) and the rotation tables
// d^l_{m1 m2}(beta_s) of every leg into LDS (one lane per (leg, pair)). The (path, energy) pairs are then cut into
// energy bands of the edge's plan: a wavefront runs whole bands, lane (path j, energy i) = j * (64 / G) + i, so its
// lanes loop over similar numbers of partial waves, and the bands are spread over the wavefronts so that their
// loads balance. Every lane evaluates the Rehr-Albers amplitude at its energy: the termination matrix T, then the
// scattering matrices F_{n-1} ... F_1 multiplied into U = T F_{n-1} ... F_s, each F streamed over the lane's partial
// waves with the spherical-wave recursion of its two legs kept in registers; cchi = tr(U) cfac. The fine grid
// (amplitude and unwrapped phase interpolated to q, chi = A sin(Phi + 2 q reff)) is one wavefront's epilogue per
// path.
//
// Reference semantics: GPUPathStage.cchi / chi_fine (src/feffgpu/stage.py:235-388), the literal port
// genfmt.PathStage (src/feffgpu/genfmt.py:395-445) and the path-list gather/sum of
// src/feffgpu/exafs/forward.py:143-163.
#include <hip/hip_runtime.h>
#include <stdio.h>
#include <string.h>

#include <atomic>
#include <type_traits>

#include "fm_device.h"
#include "fm_path.h"

using namespace fm;

thread_local char g_err[512] = "";
int fm_set_err(int code, const char *msg) {
    snprintf(g_err, sizeof g_err, "%s", msg);
    return code;
}

// LDS a workgroup may take when the paths per workgroup are chosen (bytes): FP32 forward, FP32 adjoint, FP64
// forward, FP64 adjoint.
static std::atomic<size_t> g_lds_budget[4] = {{80 * 1024}, {160 * 1024}, {160 * 1024}, {160 * 1024}};
size_t fm_lds_budget(int kernel) { return g_lds_budget[kernel & 3].load(std::memory_order_relaxed); }

namespace {

struct KArgs {
    fm_edge ed;           // the edge of every path of this launch
    fm_bands bp;          // its band plan for gp paths per workgroup
    int gp;
    int npath_bucket;     // paths in this launch
    int first;            // offset into order[]
    const int32_t *order;
    PathSrc src;
    void *cchi_out;       // mode 0
    fm_grid G;            // mode 1
    void *chi_path;
    double *reff_out;
    int edge_id;          // index of the edge (rows of the grid's jq / fq)
    int nleg;             // the leg count of every path of this launch
    size_t path_bytes;    // LDS per path of a workgroup
    unsigned *tickets;    // the path groups' counter pair of a persistent launch, or NULL (fixed assignment)
    void *keep;           // the bucket's kept scattering matrices (written), or NULL
};

int set_err(int code, const char *msg) {
    snprintf(g_err, sizeof g_err, "%s", msg);
    return code;
}
int hip_err(hipError_t e, const char *where) {
    snprintf(g_err, sizeof g_err, "%s: %s", where, hipGetErrorString(e));
    return FM_EHIP;
}


template <class T, int SET>
__global__ __launch_bounds__(FM_WAVE * 4) __attribute__((amdgpu_waves_per_eu(FM_WPE(T)))) void k_forward(KArgs a) {
    extern __shared__ __align__(16) unsigned char smem[];
    __shared__ int pidx[8], pval[8], tslot;
    const int wid = __builtin_amdgcn_readfirstlane(threadIdx.x / FM_WAVE);
    const int lane = lane_id();
    const fm_edge &ed = a.ed;
    const int G = a.gp, cap = FM_WAVE / G;
    const int ltab = ed.L1 + 1, E = ed.E;
    const int nl = a.nleg;
    const size_t off_W = align16(sizeof(PathLDS));
    const size_t off_amp = off_W + (SET < FM_SET_SS ? align16(sizeof(T) * rot_elems(nl, ltab)) : 0);
    const size_t off_ang = off_amp + align16(sizeof(T) * (size_t)E);
    const int jl = lane / cap, il = lane % cap;      // this lane's path and energy within a band
    unsigned char *lbase = smem + (size_t)jl * a.path_bytes;
    const PathLDS &gl = *reinterpret_cast<const PathLDS *>(lbase);
    const T *Wl = reinterpret_cast<const T *>(lbase + off_W);
    for (int blk = first_group(a.tickets, tslot); blk * G < a.npath_bucket; blk = next_group(a.tickets, tslot, blk)) {
        // ---- setup. Single scattering (no tables): the wavefront's paths j = wid, wid + 4 at once, a 16-lane group
        // each (G <= 8); multiple scattering: one path after the other (measured faster there, the two workgroups of
        // a CU overlapping their setups)
        if constexpr (SET >= FM_SET_SS) {
            const int j = wid + 4 * (lane >> 4);
            const bool act = j < G;
            const int slot = blk * G + j;
            const bool valid = act && slot < a.npath_bucket;
            const int s0 = valid ? slot : blk * G;
            const int p = a.order ? a.order[a.first + s0] : a.first + s0;
            PathLDS &g = *reinterpret_cast<PathLDS *>(smem + (size_t)(act ? j : 0) * a.path_bytes);
            int ei;
            load_path_group(a.src, p, g, ei, nl, lane & 15, act, false);   // single scattering: no tables
            if (act && (lane & 15) == 0) { pidx[j] = p; pval[j] = valid ? 1 : 0; }
        } else {
            for (int j = wid; j < G; j += 4) {
                const int slot = blk * G + j;
                const bool valid = slot < a.npath_bucket;
                const int s0 = valid ? slot : blk * G;
                const int p = a.order ? a.order[a.first + s0] : a.first + s0;
                PathLDS &g = *reinterpret_cast<PathLDS *>(smem + (size_t)j * a.path_bytes);
                int ei;
                load_path(a.src, p, g, ei, nl);
                if (lane == 0) { pidx[j] = p; pval[j] = valid ? 1 : 0; }
            }
        }
        if constexpr (SET < FM_SET_SS) build_tables<T, SET, false, 1>(ed, smem, a.path_bytes, off_W, 0, wid, 4, G, nl, ltab);
        __syncthreads();
        // ---- energies: the bands of this wavefront, each wavefront stepping from one of its bands to the next
        auto next_band = [&](int b) {
            for (++b; b < a.bp.nband && a.bp.wave[b] != wid; ++b) {}
            return b;
        };
        for (int b = next_band(-1); b < a.bp.nband; b = next_band(b)) {
            const int e0 = a.bp.start[b], nb = a.bp.start[b + 1] - e0;
            const bool active = il < nb && pval[jl];
            const int e = e0 + (il < nb ? il : nb - 1);
            cx<T> c;
            if constexpr (SET >= FM_SET_SS) {
                c = ss_cchi<T, SET - FM_SET_SS + 1>(ed, gl, e);
            } else {
                // this lane's (path, energy) in the kept rows, written by active lanes only
                T *kp = (a.keep && active)
                            ? reinterpret_cast<T *>(a.keep) + ((size_t)(blk * G + jl) * keep_rows<SET>(nl) * E + e) * 2
                            : nullptr;
                c = path_cchi<T, SET>(ed, gl, Wl, ltab, e, nl, kp);
            }
            if (active) {
                if (a.src.mode == 0) {
                    T *o = reinterpret_cast<T *>(a.cchi_out) + ((size_t)pidx[jl] * E + e) * 2;
                    o[0] = gl.ok ? c.r : (T)__builtin_nan("");
                    o[1] = gl.ok ? c.i : (T)__builtin_nan("");
                } else {
                    T *amp = reinterpret_cast<T *>(lbase + off_amp);
                    double *ang = reinterpret_cast<double *>(lbase + off_ang);
                    amp[e] = (T)sqrt((double)c.r * c.r + (double)c.i * c.i + 1e-30);
                    ang[e] = (double)atan2(c.i, c.r);
                }
            }
        }
        __syncthreads();
        // ---- epilogue: one wavefront per path
        for (int j = wid; j < G; j += 4) {
            if (!pval[j]) continue;
            unsigned char *base = smem + (size_t)j * a.path_bytes;
            const PathLDS &g = *reinterpret_cast<const PathLDS *>(base);
            const int p = pidx[j];
            const double reff_A = g.ok ? g.reff * FM_BOHR : __builtin_nan("");
            if (lane == 0 && a.reff_out) a.reff_out[p] = reff_A;
            if (a.src.mode == 1) {
                const int nk = a.G.nk;
                T *out = reinterpret_cast<T *>(a.chi_path) + (size_t)p * nk;
                if (g.ok)
                    fine_grid<T>(ed, a.G.jq + (size_t)a.edge_id * nk, a.G.fq + (size_t)a.edge_id * nk, a.G.q, nk, reff_A,
                                 a.src.P.sigma2 ? a.src.P.sigma2[p] : 0.0, reinterpret_cast<T *>(base + off_amp),
                                 reinterpret_cast<double *>(base + off_ang), out,
                                 a.G.S ? a.G.S + (size_t)a.edge_id * nk * a.G.sstride : nullptr, a.G.sstride);
                else
                    for (int i = lane; i < nk; i += FM_WAVE) out[i] = (T)__builtin_nan("");
            }
        }
        __syncthreads();
    }
    release_groups(a.tickets);
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
        fine_grid<T>(ed, G.jq, G.fq, G.q, G.nk, reff[p], 0.0, amp, ang, chi + (size_t)p * G.nk, G.S, G.sstride);
        wave_sync();
    }
}

// ------------------------------------------------------------------------------------------------ launch helpers
int set_for(const fm_edge &ed, int nleg) {
    if (nleg == 2) return FM_SET_SS + ed.ilinit - 1;
    return ed.ilinit <= 1 ? 0 : 1;
}

// Paths per workgroup: the policy's, else the largest plan within the LDS budget (by default room for two workgroups
// per CU), else the largest that fits one workgroup. Grid: one workgroup per group, or the policy's persistent grid.
template <class T, int SET> hipError_t launch_forward(KArgs &a, const fm_launch &pol, hipStream_t st) {
    const fm_edge &ed = a.ed;
    a.path_bytes = path_stride(align16(sizeof(PathLDS)) +
                               (SET < FM_SET_SS ? align16(sizeof(T) * rot_elems(a.nleg, ed.L1 + 1)) : 0) +
                               align16(sizeof(T) * ed.E) + align16(sizeof(double) * ed.E));
    const int g = choose_gpaths(ed, a.path_bytes, fm_lds_budget(sizeof(T) == 4 ? 0 : 2), pol.gpaths);
    if (g == -2) return hipErrorInvalidConfiguration;
    if (g < 0) return hipErrorInvalidValue;
    a.gp = 1 << g;
    a.bp = ed.bands[g];
    const size_t shm = a.path_bytes * a.gp;
    const void *fn = (const void *)k_forward<T, SET>;
    if (shm > 64 * 1024) {
        hipError_t e = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)shm);
        if (e != hipSuccess) return e;
    }
    long blocks = 0;
    hipError_t e = launch_grid(fn, shm, (a.npath_bucket + a.gp - 1) / a.gp, pol.grid, &blocks);
    if (e != hipSuccess) return e;
    if (blocks == 0) return hipSuccess;
    a.tickets = pol.grid != 0 ? pol.tickets : nullptr;
    hipLaunchKernelGGL((k_forward<T, SET>), dim3((unsigned)blocks), dim3(FM_WAVE * 4), shm, st, a);
    return hipGetLastError();
}

template <class T> hipError_t dispatch(KArgs &a, int set, const fm_launch &pol, hipStream_t st) {
    switch (set) {
        case 0: return launch_forward<T, 0>(a, pol, st);
        case 1: return launch_forward<T, 1>(a, pol, st);
        case FM_SET_SS: return launch_forward<T, FM_SET_SS>(a, pol, st);
        case FM_SET_SS + 1: return launch_forward<T, FM_SET_SS + 1>(a, pol, st);
        case FM_SET_SS + 2: return launch_forward<T, FM_SET_SS + 2>(a, pol, st);
    }
    return hipErrorInvalidValue;
}

// The policy of bucket b: the caller's, or the defaults (paths per workgroup by the LDS budget, one workgroup each).
fm_launch policy_of(const fm_launch *launch, int b) {
    fm_launch d;
    memset(&d, 0, sizeof d);
    return launch ? launch[b] : d;
}

const char *check_policy(const fm_launch &p) {
    if (p.gpaths != 0 && p.gpaths != 1 && p.gpaths != 2 && p.gpaths != 4 && p.gpaths != 8)
        return "launch policy: gpaths must be 0, 1, 2, 4 or 8";
    if (p.grid < -1) return "launch policy: grid must be -1, 0 or a number of workgroups";
    return nullptr;
}

int check_edge(const fm_edge *e) {
    const char *m = edge_problem(e);
    if (m) return set_err(FM_EINVAL, m);
    if (e->ilinit < 1 || e->ilinit > 3) return set_err(FM_EUNSUPPORTED, "edge: only K, L and M edges (ilinit 1..3)");
    return FM_OK;
}

}  // namespace

extern "C" {

const char *fm_last_error(void) { return g_err; }
const char *fm_version(void) { return "feffgpu_amd 0.1 (gfx950)"; }

size_t fm_keep_bytes(int32_t precision, const fm_edge *edge, int32_t nleg, int32_t npath) {
    if (!edge || nleg < 3 || nleg > 6 || npath <= 0 || (precision != 32 && precision != 64)) return 0;
    const int rows = edge->ilinit <= 1 ? keep_rows<0>(nleg) : keep_rows<1>(nleg);
    return (size_t)npath * (size_t)rows * (size_t)edge->E * 2 * (precision == 32 ? 4 : 8);
}

int32_t fm_set_lds_budget(int32_t kernel, int32_t kib) {
    if (kernel < 0 || kernel > 3 || kib < 16 || kib > 160) return -1;
    return (int32_t)(g_lds_budget[kernel].exchange((size_t)kib * 1024) / 1024);
}

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
    KArgs a;
    memset(&a, 0, sizeof a);
    a.ed = *edge;
    a.src.mode = 0;
    a.npath_bucket = npath;
    a.first = 0;
    a.order = nullptr;
    a.src.nleg = nleg;
    a.src.coords = coords;
    a.src.ipots64 = ipots;
    a.src.nph1 = edge->nph1;
    a.cchi_out = cchi;
    a.reff_out = reff;
    a.edge_id = 0;
    a.nleg = nleg;
    hipStream_t st = (hipStream_t)stream;
    const fm_launch pol = policy_of(nullptr, 0);
    hipError_t e = precision == 32 ? dispatch<float>(a, set, pol, st) : dispatch<double>(a, set, pol, st);
    if (e == hipErrorInvalidValue) return set_err(FM_EINVAL, "fm_cchi: a path's tables do not fit one workgroup's LDS");
    return e == hipSuccess ? FM_OK : hip_err(e, "fm_cchi launch");
}

int fm_chi_fine(const fm_edge *edge, int32_t precision, int32_t npath, const void *cchi, const double *reff,
                const fm_grid *grid, void *chi, void *stream) {
    int rc = check_edge(edge);
    if (rc) return rc;
    if (npath < 0) return set_err(FM_EINVAL, "npath < 0");
    if (npath == 0) return FM_OK;
    if (!cchi || !reff || !grid || !chi || !grid->q || !grid->jq || !grid->fq) return set_err(FM_EINVAL, "NULL buffer");
    if (grid->S && grid->sstride < edge->E) return set_err(FM_EINVAL, "grid.sstride < E");
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
                         const int32_t *bucket_nleg, int32_t nseg, const int32_t *seg_ptr, const int32_t *seg_paths, double *chi_seg,
                         void *chi_path, double *reff, const fm_launch *launch, void *stream) {
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
    for (int i = 0; i < nedge; ++i)
        if (grid->S && grid->sstride < edges[i].E) return set_err(FM_EINVAL, "grid.sstride < E");
    if (nbucket > 0 && (!bucket_offset || !bucket_set || !bucket_nleg)) return set_err(FM_EINVAL, "NULL bucket arrays");
    hipStream_t st = (hipStream_t)stream;
    for (int b = 0; b < nbucket; ++b) {
        const int cnt = bucket_offset[b + 1] - bucket_offset[b];
        if (cnt <= 0) continue;
        const int kind = bucket_set[b];
        if (kind < 0 || kind > 2) return set_err(FM_EINVAL, "bucket kind out of range 0..2");
        KArgs a;
        memset(&a, 0, sizeof a);
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
        a.ed = edges[a.edge_id];
        a.nleg = bucket_nleg[b];
        if (a.nleg < 2 || a.nleg > paths->nmax) return set_err(FM_EINVAL, "bucket nleg out of range 2..nmax");
        if ((kind == FM_BUCKET_SS) != (a.nleg == 2)) return set_err(FM_EINVAL, "bucket kind does not match its nleg");
        if ((kind == FM_BUCKET_K_MS) != (a.ed.ilinit <= 1) && kind != FM_BUCKET_SS)
            return set_err(FM_EINVAL, "bucket kind does not match its edge");
        const int set = set_for(a.ed, a.nleg);
        const fm_launch pol = policy_of(launch, b);
        if (const char *m = check_policy(pol)) return set_err(FM_EINVAL, m);
        if (pol.keep && kind != FM_BUCKET_SS && a.nleg >= 3 && a.nleg <= 6) {
            if (pol.keep_bytes < fm_keep_bytes(precision, &a.ed, a.nleg, cnt))
                return set_err(FM_EINVAL, "fm_chi_paths: keep_bytes smaller than fm_keep_bytes for the bucket");
            a.keep = pol.keep;
        }
        hipError_t e = precision == 32 ? dispatch<float>(a, set, pol, st) : dispatch<double>(a, set, pol, st);
        if (e == hipErrorInvalidValue) return set_err(FM_EINVAL, "fm_chi_paths: the bucket's paths do not fit one workgroup's LDS");
        if (e == hipErrorInvalidConfiguration)
            return set_err(FM_EINVAL, "fm_chi_paths: the launch policy's gpaths has no band plan or does not fit one CU's LDS");
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
