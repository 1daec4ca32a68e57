/*
 * feffgpu_amd: the path stage of feffgpu (Rehr-Albers multiple-scattering path amplitudes and their sum to
 * chi(k)) as hand-written HIP kernels for MI355X (gfx950), behind a plain C ABI.
 *
 * Every pointer in these structs and calls is a DEVICE pointer owned by the caller, unless marked "host". No call
 * below allocates device memory: workspaces are passed in, so every call can be captured in a hipGraph. All calls
 * are asynchronous on the given stream (a hipStream_t, NULL for the default stream) and return 0 on success or a
 * negative code; fm_last_error() gives the message of the last failure on the calling thread.
 *
 * Units as in FEFF and feffgpu: coordinates and positions in Angstrom, momenta in 1/bohr inside the tables, the
 * output grid q in 1/Angstrom. Precision 32 = FP32 amplitudes with FP64 leg geometry and FP64 path phase reduced
 * mod 2 pi before the FP32 sine; precision 64 = FP64 throughout.
 */
#ifndef FEFFGPU_AMD_H
#define FEFFGPU_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define FM_MAX_LEGS 9          /* FEFF's own limit (src/feffgpu/pathfinder.py:26) */
#define FM_MAX_CHAN 4          /* distinct final-state l of the termination */
#define FM_MAX_POT 16          /* potentials per phase file */
#define FM_NREP 20             /* independent (lambda1, lambda2) pairs of a 6 x 6 scattering matrix */

enum { FM_OK = 0, FM_EINVAL = -1, FM_EUNSUPPORTED = -2, FM_EHIP = -3 };

/* A split of the phase-grid energies into bands for a workgroup of 4 wavefronts holding gpaths paths: a band is
 * run by one wavefront, lane (path j, energy i) = j * (64 / gpaths) + i, so the lanes of a wavefront loop over
 * similar numbers of partial waves; the bands are spread over the wavefronts so that their totals are balanced. */
typedef struct fm_bands {
    int32_t nband;               /* 0 (no plan for this gpaths) or 1..16 */
    int32_t start[17];           /* band b = energies [start[b], start[b+1]); start[0] = 0, start[nband] = E, each band
                                    1..64 / gpaths energies */
    int32_t wave[16];            /* the wavefront (0..3) that runs band b; a wavefront runs its bands in increasing b */
} fm_bands;

/* Device tables of one absorber edge, built once from a phase.pad (GPUPathStage.__init__, src/feffgpu/stage.py:97-125,
 * and the lambda / rotation bookkeeping of src/feffgpu/stage.py:41-93, 128-144). */
typedef struct fm_edge {
    int32_t E;                   /* phase-grid energies evaluated (first E of ne1) */
    int32_t L1;                  /* largest il = lmax + 1 over those energies and potentials */
    int32_t nph1;                /* potentials 0..nph */
    int32_t ilinit;              /* initial-state l + 1 (1: K/L1, 2: L2/L3, 3: M4/M5) */
    int32_t nchan;               /* distinct final-state channels */
    int32_t chan_il[FM_MAX_CHAN];/* il = l + 1 of each channel */
    int32_t lmax1[FM_MAX_POT];   /* host: max over the E energies of lmax + 1 per potential */
    const int32_t *lmax;         /* [nph1][E] per-energy partial-wave limit */
    const double *ck;            /* [E][2] complex momentum (1/bohr) */
    const double *ckmxk;         /* [E][2] ck - xk */
    const double *kf;            /* [E] xk / BOHR (1/A) */
    const float *tl32;           /* [nph1][L1 + 1][E][2] j-averaged t-matrix (l + 1) t(-l) + l t(l), index il */
    const double *tl64;
    const float *chw32;          /* [nchan][E][2] termination weight sum_k bdiag_k rkk_k^2 of each channel */
    const double *chw64;
    const double *rotc;          /* [L1 + 1][49][3] d-recursion coefficients P, Q, R per l and (m1 + 3) * 7 + m2 + 3 */
    const double *rotn;          /* [L1 + 1][16] normalisation xnlm(il, |m1|) (-1)^|m1| / xnlm(il, |m2|), index |m1| * 4 + |m2| */
    fm_bands bands[4];           /* host: band plans for gpaths = 1, 2, 4, 8 (a launch takes the largest gpaths whose
                                    paths fit the workgroup's LDS) */
} fm_edge;

/* A path list in the atom-index form of feffgpu's forward model (src/feffgpu/exafs/forward.py:143-163, 664-685). */
typedef struct fm_paths {
    int32_t npath;
    int32_t nmax;                /* legs per row of ipots; atoms has nmax - 1 columns */
    const double *pos;           /* [N][3] atom positions (A) */
    const int32_t *atoms;        /* [npath][nmax - 1] scatterers in path order */
    const int32_t *nleg;         /* [npath] 2..9 */
    const int32_t *absorber;     /* [npath] absorbing atom */
    const int32_t *ipots;        /* [npath][nmax] potential of the absorber, then of each scatterer */
    const double *deg;           /* [npath] degeneracy */
    const int32_t *edge;         /* [npath] index into the edge array, or NULL (all 0) */
    const double *sigma2;        /* [npath] Debye-Waller sigma^2 (A^2) applied as exp(-2 sigma^2 q^2), or NULL */
    int32_t cell_kind;           /* 0: cluster, 1: cubic box cell[0], 2: general cell (rows = lattice vectors) */
    double cell[9];              /* host values */
    double cellinv[9];           /* host values: inverse of the general cell */
} fm_paths;

/* The output k grid and its linear-interpolation weights on every edge's phase grid (GPUPathStage.linear_matrix,
 * src/feffgpu/stage.py:353-363). */
typedef struct fm_grid {
    int32_t nk;
    const double *q;             /* [nk] 1/A */
    const int32_t *jq;           /* [nedge][nk] lower phase-grid index, clipped to 0..E-2 */
    const double *fq;            /* [nedge][nk] weight of the upper point */
    /* for the adjoints (may be NULL for the forward calls): q sorted ascending is also sorted by jq on every
     * edge, so the points interpolating from energy e form one run of qorder */
    const int32_t *qorder;       /* [nk] indices of q in ascending order */
    const int32_t *qstart;       /* [nedge][qstride] first position in qorder whose jq >= e, e = 0..E_edge */
    int32_t qstride;             /* row length of qstart (>= largest E + 1) */
    /* optional dense interpolation matrix (GPUPathStage.chi_fine's S, e.g. spline_matrix for the "larch"
     * convention, src/feffgpu/stage.py:341-351, 372-388): amplitude and unwrapped phase at q[i] of edge ed are
     * sum_e S[(ed * nk + i) * sstride + e] times their values at energy e; NULL: the linear jq / fq interpolation.
     * The adjoints need nk <= 1365 with S set. */
    const double *S;
    int32_t sstride;             /* row length of S (>= largest E) */
} fm_grid;

/* GPUPathStage.cchi (src/feffgpu/stage.py:235-338): npath paths of nleg legs each; coords [npath][nleg][3] (A,
 * absorber first at the origin), ipots [npath][nleg] (absorber first). Writes cchi [npath][E] complex (interleaved
 * float for precision 32, double for 64) and reff [npath] (A). */
int fm_cchi(const fm_edge *edge, int32_t precision, int32_t iorder, int32_t npath, int32_t nleg, const double *coords,
            const int64_t *ipots, void *cchi, double *reff, void *stream);

/* Adjoint of fm_cchi: gcchi [npath][E] complex is dL/dRe + i dL/dIm of cchi (same element type as cchi), greff
 * [npath] dL/dreff or NULL. Writes gcoords [npath][nleg][3] = dL/dcoords. */
int fm_cchi_backward(const fm_edge *edge, int32_t precision, int32_t iorder, int32_t npath, int32_t nleg,
                     const double *coords, const int64_t *ipots, const void *gcchi, const double *greff, double *gcoords,
                     void *workspace, size_t workspace_bytes, void *stream);

/* fm_cchi_backward without a workspace argument and without any allocation: no counterpart in the reference. For
 * callers that cannot allocate device memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc). Its workspace is a static device array of FM_NOALLOC_SCRATCH_BYTES (64 MiB) per GPU, held from
 * the library's load to the end of the process; larger batches run in chunks of it, and an edge whose single path
 * needs more than it (none of the shipped phase files) is refused with FM_EINVAL. Calls that overlap on one GPU share
 * that array and race: serialise them (one stream per GPU). Same arguments and results as fm_cchi_backward. */
#define FM_NOALLOC_SCRATCH_BYTES (64u << 20)
int fm_cchi_backward_noalloc(const fm_edge *edge, int32_t precision, int32_t iorder, int32_t npath, int32_t nleg,
                             const double *coords, const int64_t *ipots, const void *gcchi, const double *greff,
                             double *gcoords, void *stream);

/* Bytes of the workspace the adjoints need for `band_slots` energy bands of paths of at most nlegmax legs in flight
 * at once (paths of 7-9 legs: each lane keeps its scattering matrices, prefix products and gradient terms there; a
 * workgroup holds one slot per band of its plan, at most 16, and a launch uses as many workgroups as the workspace
 * holds). The adjoint of 2-6 legs keeps (4 nleg + 1) E complex per path (cchi and the derivative slots of every
 * energy) for as many paths of a bucket at a time as the workspace holds regions for, less one spare: larger buckets
 * run in chunks. Any workspace of at least fm_backward_workspace_bytes(precision, nlegmax, 16) bytes works. */
size_t fm_backward_workspace_bytes(int32_t precision, int32_t nlegmax, int32_t band_slots);

/* GPUPathStage.chi_fine with the linear ("feff") convention (src/feffgpu/stage.py:365-388): chi [npath][nk] =
 * A(q) sin(Phi(q) + 2 q reff). cchi as written by fm_cchi; jq/fq of grid are read for edge 0. */
int fm_chi_fine(const fm_edge *edge, int32_t precision, int32_t npath, const void *cchi, const double *reff,
                const fm_grid *grid, void *chi, void *stream);

/* Adjoint of fm_chi_fine: gchi [npath][nk] (element type of chi) -> gcchi [npath][E] complex and greff [npath]
 * (grid.qorder/qstart required). */
int fm_chi_fine_backward(const fm_edge *edge, int32_t precision, int32_t npath, const void *cchi, const double *reff,
                         const fm_grid *grid, const void *gchi, void *gcchi, double *greff, void *stream);

/* The fused path stage over a path list (MultipleScattering.signals / PathSpectra.spectra,
 * src/feffgpu/exafs/forward.py:143-163, 657-685; chain.run's sum, src/feffgpu/chain.py:45-49):
 *   chi_path[p][k] = chi of path p on the grid (without the degeneracy, with its Debye-Waller factor),
 *   chi_seg[s][k] = sum over the paths of segment s, in the order seg_paths lists them, of deg[p] chi_path[p][k].
 * order [npath] (device) is the processing order, a permutation of 0..npath-1 grouped into nbucket buckets:
 * bucket b holds order[bucket_offset[b] .. bucket_offset[b+1]) (host array), all of kind bucket_kind[b] (host):
 * FM_BUCKET_K_MS (K/L1 edge, 3..9 legs), FM_BUCKET_LM_MS (L2/L3/M edges, 3..9 legs), FM_BUCKET_SS (any
 * edge, 2 legs); every path of bucket b has edge bucket_edge[b] (host, may be NULL: edge 0) and bucket_nleg[b]
 * legs (host, 2..nmax; a path whose nleg differs gets NaN for its chi_path row and reff).
 * seg_ptr [nseg + 1] and seg_paths [npath] (device) list the paths of each segment (CSR). chi_path (float for
 * precision 32, double for 64) is required: it is also the workspace of the deterministic segment sum.
 * chi_seg (double) and reff (double, A) may be NULL. Nothing is kept for the adjoint, which recomputes the
 * amplitudes. */
enum { FM_BUCKET_K_MS = 0, FM_BUCKET_LM_MS = 1, FM_BUCKET_SS = 2 };

/* Launch options of one bucket (host array launch[nbucket], or NULL for the defaults; no counterpart in the
 * reference, whose stage has no launch configuration). gpaths: paths per workgroup, 1, 2, 4 or 8 (refused where the
 * edge has no band plan for it or its paths do not fit one CU's LDS), or 0: the largest within the LDS budget of
 * fm_set_lds_budget. grid: 0 launches one workgroup per group of gpaths paths; n > 0 a persistent grid of at most n
 * workgroups, -1 as many as the device's CUs hold at once, that loop over the groups. tickets (device, two uint32,
 * zero before first use; used only with grid != 0): the groups are then taken from this counter one at a time, and
 * the launch's last workgroup puts the pair back to zero, so it can serve every later launch on the same stream (not
 * two launches that overlap in time); NULL: workgroup i takes groups i, i + grid, ... The multiple-scattering
 * adjoint of 3-6 legs (and single scattering) applies the policy to its sweep; its per-path tail runs one wavefront
 * per path.
 * keep (device, keep_bytes >= fm_keep_bytes(precision, edge, nleg, paths of the bucket)), for buckets of multiple
 * scattering of 3-6 legs (ignored by the others): fm_chi_paths writes the scattering matrices of every path's
 * scatterers into it ([path][row][energy] complex, paths in bucket order; the rows of a path are the independent
 * entries of each scatterer's block, scatterer 1 first), and fm_chi_paths_backward reads them instead of
 * recomputing them. The caller guarantees that the
 * adjoint call sees the positions of the forward call that wrote the buffer; NULL: the forward keeps nothing and the
 * adjoint recomputes the amplitudes, as the reference's checkpointed autograd does. */
typedef struct fm_launch {
    int32_t gpaths;
    int32_t grid;
    uint32_t *tickets;
    void *keep;
    size_t keep_bytes;
} fm_launch;

/* Bytes of a bucket's kept scattering matrices: npath x rows x E complex elements of the precision, the rows being
 * 14 (K edges: the 6 x 4 and 4 x 6 blocks of the first and last scatterers) or 20 per scatterer (0 for buckets that
 * keep nothing: nleg outside 3..6). */
size_t fm_keep_bytes(int32_t precision, const fm_edge *edge, int32_t nleg, int32_t npath);

int fm_chi_paths(const fm_edge *edges, int32_t nedge, const fm_paths *paths, const fm_grid *grid, int32_t precision,
                 int32_t iorder, const int32_t *order, int32_t nbucket, const int32_t *bucket_offset,
                 const int32_t *bucket_kind, const int32_t *bucket_edge, const int32_t *bucket_nleg, int32_t nseg,
                 const int32_t *seg_ptr, const int32_t *seg_paths, double *chi_seg, void *chi_path, double *reff,
                 const fm_launch *launch, void *stream);

/* Adjoint of fm_chi_paths: the same path list, buckets and grid (grid.qorder/qstart required); gseg [nseg][nk] =
 * dL/dchi_seg (double) with segment [npath] the row of each path, gpath [npath][nk] = dL/dchi_path (element type by
 * precision) or NULL. Writes gcoord [npath][nmax][3] (double; dL/dcoords of each path's scatterers in slots
 * 1..nleg-1 and minus their sum, the absorber's share, in slot 0) and grad_pos [natom][3] = dL/dpos, the sum over
 * atom_ptr [natom + 1] / atom_refs (CSR of path * nmax + slot references, device), in list order (a path whose
 * nleg differs from its bucket's gets NaN). The amplitudes and scattering matrices are recomputed inside (nothing is
 * kept from the forward call), as the reference's checkpointed autograd does (src/feffgpu/exafs/forward.py:683). */
int fm_chi_paths_backward(const fm_edge *edges, int32_t nedge, const fm_paths *paths, const fm_grid *grid,
                          int32_t precision, int32_t iorder, const int32_t *order, int32_t nbucket,
                          const int32_t *bucket_offset, const int32_t *bucket_kind, const int32_t *bucket_edge,
                          const int32_t *bucket_nleg, const int32_t *segment, const double *gseg, const void *gpath, int32_t natom,
                          const int32_t *atom_ptr, const int32_t *atom_refs, double *gcoord, double *grad_pos,
                          void *workspace, size_t workspace_bytes, const fm_launch *launch, void *stream);

/* The plain FP64 path stage: the reference's op sequence (src/feffgpu/stage.py:147-388) with one kernel per stage
 * and every intermediate (rotation matrices, spherical-wave factors, gam / gamtl, scattering matrices, termination,
 * cchi) in HBM -- the comparator of the fused kernels on the same GPU, and an independent evaluation for tests.
 * Same path list, buckets, grid and outputs as fm_chi_paths at precision 64 (chi_path, reff, chi_seg double); each
 * bucket is processed in chunks of as many paths as `workspace` holds, fm_plain_path_bytes(edge, nleg) per path. */
size_t fm_plain_path_bytes(const fm_edge *edge, int32_t nleg);
int fm_chi_paths_plain(const fm_edge *edges, int32_t nedge, const fm_paths *paths, const fm_grid *grid, int32_t iorder,
                       const int32_t *order, int32_t nbucket, const int32_t *bucket_offset, const int32_t *bucket_kind,
                       const int32_t *bucket_edge, const int32_t *bucket_nleg, int32_t nseg, const int32_t *seg_ptr,
                       const int32_t *seg_paths, double *chi_seg, double *chi_path, double *reff, void *workspace,
                       size_t workspace_bytes, void *stream);

/* The k window and R range of a comparison (FitWindow, src/feffgpu/exafs/transform.py:30-62): Hanning sills
 * centered on kmin and kmax, dk wide (dk <= 0: the rectangular window kmin..kmax); k^kweight chi(k) is transformed. */
typedef struct fm_window {
    double kmin, kmax, dk;
    int32_t kweight;             /* 0..16 */
    double rmin, rmax;           /* A; the caller turns them into the range m0 .. m0 + nm - 1 of R_m = m pi / (nfft kstep) */
} fm_window;

#define FM_RF_NFFT_MAX 4096

/* The misfit after the path sum, r_factor_torch (src/feffgpu/exafs/transform.py:130-145) and compare.r_factor
 * (src/feffgpu/compare.py:17-27), per segment: chi [nseg][nk] on the grid k_j = j kstep (times `scale`, e.g. 1 / the
 * number of absorbers), data [nseg][nk] (data_stride nk) or one spectrum [nk] (data_stride 0). The transform of
 * k^w chi window, zero-padded to nfft (a power of two <= 4096, >= nk), times kstep / sqrt(pi), is taken at
 * m = m0 .. m0 + nm - 1 (< nfft / 2). rfac [nseg] = sum_m |M_m - D_m|^2 / sum_m |D_m|^2; gchi [nseg][nk] (or NULL) =
 * d rfac / d chi. FP64, one workgroup per segment, no allocation (capturable). */
int fm_r_factor(int32_t nseg, int32_t nk, double kstep, int32_t nfft, const fm_window *w, int32_t m0, int32_t nm,
                const double *chi, double scale, const double *data, int32_t data_stride, double *rfac, double *gchi,
                void *stream);

/* The distance grid and cell of single scattering from tables (PairTables, src/feffgpu/exafs/forward.py:178-320). */
typedef struct fm_pair_grid {
    double r0, dr;               /* grid r0 + i dr, i = 0 .. nr - 1 (A) */
    double rmax;                 /* pairs within rmax (A) count */
    int32_t nr;                  /* >= 2 */
    int32_t cell_kind;           /* 0: cluster, 1: cubic box cell[0], 2: general cell (rows = lattice vectors) */
    double cell[9], cellinv[9];  /* host values */
} fm_pair_grid;

/* PairTables.histogram (src/feffgpu/exafs/forward.py:233-280) for nconf configurations pos [nconf][natom][3] (A):
 * every pair (absorber a, neighbor j != a) within rmax, minimum image, adds 1 - f to bin i0 and f to bin i0 + 1 of
 * row ipot[j] (u = (r - r0) / dr, i0 = floor(u) clamped to 0..nr-2, f = u - i0 clamped to [0, 1]). absorbers
 * [nabs] (indices into each configuration) or NULL for every atom; ipot [natom] in 0..npot-1 (other values: never a
 * neighbor). hist [npot][nr] accumulates in fixed point, units of 2^-32 of a pair (the caller zeroes it):
 * order-independent. A bin holds up to 2^31 pairs, so a call with nconf * absorbers * (natom - 1) >= 2^31 is refused
 * (FM_EINVAL) and the caller splits its configurations. nshort (or NULL) += the pairs closer than r0 (placed on the first bin). */
int fm_pair_histogram(int32_t nconf, int32_t natom, const double *pos, int32_t nabs, const int32_t *absorbers,
                      const int32_t *ipot, int32_t npot, const fm_pair_grid *grid, int64_t *hist, int32_t *nshort,
                      void *stream);

/* Adjoint of the histogram with every atom an absorber (PairTables.single_scattering's autograd, forward.py:284-320):
 * gbin [npot][nr] = dL/dhist (FP64) -> grad_pos [nconf][natom][3] = dL/dpos, one thread per atom, no atomics. */
int fm_pair_histogram_grad(int32_t nconf, int32_t natom, const double *pos, const int32_t *ipot, int32_t npot,
                           const fm_pair_grid *grid, const double *gbin, double *grad_pos, void *stream);

/* ---------------------------------------------------------------- path finder (src/feffgpu/pathfinder.py) --------
 * FEFF's path finder for absorbers of one configuration, float32 as the reference: GPUPathFinder.enumerate
 * (pathfinder.py:211-279), importance (:72-94), degenerate_total (:135-158) and the importance filter (:266-279).
 * The enumeration runs in two passes: counts per work item, an exclusive scan (fm_exclusive_scan), then the paths
 * written at the scanned offsets; feffgpu_amd/pathfinder.py drives it. */
typedef struct fm_find_sys {
    int32_t natom;               /* atoms of one configuration */
    const float *pos;            /* [nconf][natom][3] A: configurations one after the other */
    int32_t cell_kind;           /* 0 cluster, 1 cubic (cell[0] the length), 2 general (rows the lattice vectors) */
    float cell[9];
    float cellinv[9];
    float rmax;                  /* A: half the largest total path length */
} fm_find_sys;

typedef struct fm_find_tables {
    int32_t E, nph1, ik0;        /* energies, potentials, first energy of the importance integral (0-based) */
    const float *fbeta;          /* [E][nph1][81] |f(beta)| on cos(beta) = -1, -0.975 .. 0.975, 1 */
    const float *cksp;           /* [E] Re ck (1/A) */
    const float *xlam;           /* [E] mean free path 1 / Im ck (A; 1e10 where Im ck vanishes) */
    const float *trap;           /* [E - ik0] trapezoid weights over cksp[ik0:] */
} fm_find_tables;

typedef struct fm_find_out {     /* paths of one number of legs n */
    int64_t *atoms;              /* [P][n - 1] global atom indices after the absorber */
    float *coords;               /* [P][n][3] relative to the absorber (minimum image), absorber first at 0 */
    int64_t *ipots;              /* [P][n] potential per atom, 0 for the absorber */
    float *deg;                  /* [P] 2 for a path and its reverse, 1 for a palindrome */
    int32_t *slot;               /* [P] which absorber of the call */
} fm_find_out;

typedef struct fm_find_partial { /* partial paths of the 5+-leg extension */
    int16_t *seq;                /* [P][FM_MAX_LEGS - 1] positions in the absorber's candidate list (nc = absorber) */
    float *len;                  /* [P] A, from the absorber */
    int32_t *slot;               /* [P] */
} fm_find_partial;

/* Candidates of each absorber (atoms of its own configuration other than it within rmax, minimum image, in index
 * order; absorbers and candidate indices count atoms over all configurations): cidx [nabs][maxc],
 * crel [nabs][maxc][3], cd0 [nabs][maxc], ncand [nabs] (a count above maxc means the list was cut: call again with a
 * larger maxc). The candidate selection of GPUPathFinder.enumerate (pathfinder.py:229-236). */
int fm_find_candidates(const fm_find_sys *sys, int32_t nabs, const int32_t *absorbers, int32_t maxc, int32_t *cidx,
                       float *crel, float *cd0, int32_t *ncand, void *stream);

/* Counts of 3-leg paths per (absorber b, first atom i) at cnt3[off3[b] + i], and of 4-leg paths per (b, i,
 * intermediate k) at cnt4[off4[b] + i (nc_b + 1) + k] (pathfinder.py:240-262); ncmax = the largest nc. */
int fm_find_count(int32_t nabs, const int64_t *off3, const int64_t *off4, int32_t maxc, const float *crel,
                  const float *cd0, const int32_t *ncand, int32_t ncmax, float rmax, int32_t nlegmax, int32_t *cnt3,
                  int32_t *cnt4, void *stream);

/* out[i] = sum of in[0..i) (int64) and *total = the sum of all n, for n <= 2^24; workspace of
 * fm_scan_workspace_bytes(n). */
size_t fm_scan_workspace_bytes(int64_t n);
int fm_exclusive_scan(const int32_t *in, int64_t n, int64_t *out, int64_t *total, void *workspace, size_t bytes,
                      void *stream);

/* The paths of 2, 3 and 4 legs: two-leg paths of absorber b from off2[b]; 3- and 4-leg paths at the scanned item
 * offsets pos3 / pos4, in the reference's order. */
int fm_find_write(int32_t nabs, const int32_t *absorbers, const int64_t *off3, const int64_t *off4, int32_t maxc,
                  const int32_t *cidx, const float *crel, const float *cd0, const int32_t *ncand, int32_t ncmax,
                  const int32_t *ipot, float rmax, int32_t nlegmax, const int64_t *off2, const fm_find_out *o2,
                  const int64_t *pos3, const fm_find_out *o3, const int64_t *pos4, const fm_find_out *o4, void *stream);

/* Paths of 5+ legs (pathfinder.py:160-209): the seed partials (one atom: every candidate, from off2[b]); one more
 * atom per partial (counts, or with pos the new partials); the completed paths of m + 1 legs among partials of m >= 4
 * atoms (counts, or with pos the paths). */
int fm_find_seed(int32_t nabs, const int64_t *off2, int32_t maxc, const float *cd0, const int32_t *ncand,
                 int32_t ncmax, const fm_find_partial *out, void *stream);
int fm_find_extend(int64_t npart, int32_t m, const fm_find_partial *in, int32_t maxc, const float *crel,
                   const float *cd0, const int32_t *ncand, float rmax, int32_t *cnt, const int64_t *pos,
                   const fm_find_partial *out, void *stream);
int fm_find_emit(int64_t npart, int32_t m, const fm_find_partial *in, const int32_t *absorbers, int32_t maxc,
                 const int32_t *cidx, const float *crel, const float *cd0, const int32_t *ncand, const int32_t *ipot,
                 int32_t *cnt, const int64_t *pos, const fm_find_out *o, void *stream);

/* GPUPathFinder.importance (pathfinder.py:72-94): xport [npath] of paths of nleg legs, coords [npath][nleg][3]
 * (absorber first), ipots [npath][nleg]; rtot [npath] (or NULL) the total length. */
int fm_find_importance(const fm_find_tables *t, int64_t npath, int32_t nleg, const float *coords, const int64_t *ipots,
                       float *xport, float *rtot, void *stream);

/* GPUPathFinder.degenerate_total (pathfinder.py:135-158) per absorber slot: total [npath] = the sum of deg over the
 * paths of the same slot whose leg lengths and angles round to the same 1e-3 in either direction (a 64-bit hash of
 * that key identifies the group). Workspace of fm_find_group_workspace_bytes(npath). */
size_t fm_find_group_workspace_bytes(int64_t npath);
int fm_find_degenerate_total(int64_t npath, int32_t nleg, const float *coords, const float *deg, const int32_t *slot,
                             float *total, void *workspace, size_t bytes, void *stream);

/* The importance filter (pathfinder.py:272-278): keep[p] = 100 dtot xp / ref[slot] >= critpw, in float32. */
int fm_find_filter(int64_t npath, const float *dtot, const float *xport, const int32_t *slot, const float *ref,
                   float critpw, uint8_t *keep, void *stream);

/* Launch policy (process-wide): the LDS in KiB (16..160) that a workgroup may take when the number of paths per
 * workgroup is chosen, for kernel 0 (FP32 forward calls), 1 (FP32 adjoints), 2 (FP64 forward), 3 (FP64 adjoints);
 * defaults 80 (two FP32 forward workgroups per CU), then 160. Set by feffgpu_amd.tune from timings. Returns the
 * previous value, or -1 for bad arguments. No counterpart in the reference (its stage has no launch configuration). */
int32_t fm_set_lds_budget(int32_t kernel, int32_t kib);

/* Message of the last failed call on this thread. */
const char *fm_last_error(void);

/* Library version string. */
const char *fm_version(void);

#ifdef __cplusplus
}
#endif
#endif
