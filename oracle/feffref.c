/*
 * FP64 CPU restatement of feffgpu's path stage: TEST INFRASTRUCTURE ONLY.
 *
 * This file is the parity oracle (and the CPU baseline of the benchmark). Only the test suite, the smoke check and
 * the benchmark's CPU-baseline leg load it, and only to check or time against; the product never calls it.
 *
 * It restates, literally and per energy, the FEFF GENFMT port of the reference:
 *   phase.pad decoding ............ src/feffgpu/pad.py:18-27 (unpad), :124-151 (record order), :93-121 (lmax, xk, ck)
 *   setkap / snlm / setlam ......... src/feffgpu/genfmt.py:22-53, :115-153
 *   rot3i (Edmonds recursion) ...... src/feffgpu/genfmt.py:156-204
 *   sclmz (spherical-wave factors)  src/feffgpu/genfmt.py:207-227
 *   bcoef (unpolarised diagonal) ... src/feffgpu/genfmt.py:230-264
 *   fmtrxi / mmtrxi / mmtr ......... src/feffgpu/genfmt.py:306-392
 *   path loop, trace, cfac ......... src/feffgpu/genfmt.py:395-445
 * with the leg geometry of the batched stage (beta = atan2(|u1 x u2|, u1.u2), regularised; src/feffgpu/stage.py:147-206)
 * and its fine-grid evaluation (linear interpolation of amplitude and unwrapped phase; src/feffgpu/stage.py:353-388),
 * the degeneracy-weighted sum (src/feffgpu/chain.py:45-49), the minimum image (src/feffgpu/cell.py:26-35) and the
 * coordinate gather of src/feffgpu/exafs/forward.py:153-163. The optional Debye-Waller factor exp(-2 sigma^2 q^2) is
 * not in the reference (docs/feff/limits.md:19-20); it is applied per path after the fine-grid evaluation.
 *
 * Build: make -C oracle  (gcc -O2 -fopenmp, no GPU).
 */
#include <complex.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define ORC_BOHR 0.52917721067
#define ORC_PHMIN 1.0e-7
#define MTOT 4
#define NTOT 2
#define LAMTOT 15
#define ORC_EPS 1.0e-16
#define NLEG_MAX 9

typedef double complex cplx;

static __thread char orc_err[256];
const char *orc_last_error(void) { return orc_err; }
static int fail(const char *msg) { snprintf(orc_err, sizeof orc_err, "%s", msg); return -1; }

/* ---------------------------------------------------------------- phase file -------------------------------- */
typedef struct {
    int nsp, ne, ne1, ne3, nph, ihole, ik0, npack, ixc;
    double rs, vint, rnrmav, xmu, edge;
    cplx *em, *eref;            /* em[ne], eref[ne*nsp] (ie fastest within a spin) */
    int *lmax0, *iz;
    int ltot;
    cplx *ph;                   /* ph[((ie*(2ltot+1) + l+ltot)*nsp + isp)*(nph+1) + iph] */
    cplx *rkk;                  /* rkk[(ie*8 + k)*nsp + isp] */
    /* derived */
    int *lmax;                  /* lmax[ie*(nph+1)+iph], ie < ne */
    double *xk;                 /* xk[ne] 1/bohr */
    cplx *ck;                   /* ck[ne] */
    int kinit, linit, ilinit, lmaxp1;
    int lind[9];
    double bdiag[9];
    double xnlm[128][MTOT + 2];
} orc_phase;

typedef struct { char **lines; int n, pos; char *buf; } lines_t;

static double unpad1(const char *s, int npack) {
    int iexp = (s[0] - 37) - 45;
    int c1 = s[1] - 37;
    int isgn = (c1 % 2) * 2 - 1;
    double total = (double)(c1 / 2) / (90.0 * 90.0);
    for (int i = npack; i >= 3; --i) total += (double)(s[i - 1] - 37) / pow(90.0, i);
    return 2.0 * isgn * 90.0 * total * pow(10.0, iexp);
}

static const char *next_line(lines_t *L) { return L->pos < L->n ? L->lines[L->pos++] : NULL; }

static int read_reals(lines_t *L, int npack, int count, double *out) {
    int have = 0;
    while (have < count) {
        const char *s = next_line(L);
        if (!s || s[0] != '!') return fail("phase file: expected a packed real line");
        int len = (int)strlen(s);
        while (len > 0 && (s[len - 1] == ' ' || s[len - 1] == '\r')) --len;
        int n = len / npack;
        for (int j = 0; j < n && have < count; ++j) out[have++] = unpad1(s + 1 + j * npack, npack);
    }
    return 0;
}

static int read_cplx(lines_t *L, int npack, int count, cplx *out) {
    int have = 0;
    while (have < count) {
        const char *s = next_line(L);
        if (!s || s[0] != '$') return fail("phase file: expected a packed complex line");
        int len = (int)strlen(s);
        while (len > 0 && (s[len - 1] == ' ' || s[len - 1] == '\r')) --len;
        int n = len / (2 * npack);
        for (int j = 0; j < n && have < count; ++j) {
            const char *p = s + 1 + j * 2 * npack;
            out[have++] = unpad1(p, npack) + I * unpad1(p + npack, npack);
        }
    }
    return 0;
}

static int setkap(int ihole, int *kinit, int *linit) {
    static const int a[] = {1, 2, 5, 10, 17, 24, 27}, b[] = {3, 6, 11, 18, 25, 30}, c[] = {4, 7, 12, 19, 26},
                     d[] = {8, 13, 20, 27}, e[] = {9, 14, 21, 28}, f[] = {15, 22}, g[] = {16, 23};
#define IN(arr) ({ int hit = 0; for (size_t i_ = 0; i_ < sizeof(arr) / sizeof(arr[0]); ++i_) hit |= (arr[i_] == ihole); hit; })
    if (IN(a)) { *kinit = -1; *linit = 0; return 0; }
    if (IN(b)) { *kinit = 1; *linit = 1; return 0; }
    if (IN(c)) { *kinit = -2; *linit = 1; return 0; }
    if (IN(d)) { *kinit = 2; *linit = 2; return 0; }
    if (IN(e)) { *kinit = -3; *linit = 2; return 0; }
    if (IN(f)) { *kinit = 3; *linit = 3; return 0; }
    if (IN(g)) { *kinit = -4; *linit = 3; return 0; }
#undef IN
    return fail("invalid hole number");
}

static void bcoef_unpol(int kinit, int lind[9], double bdiag[9]) {
    const int lx = 4, le2 = 0;
    for (int k = -1; k <= 1; ++k) {
        int kap = kinit + k;
        if (k == 0) kap = -kap;
        int lkap = kap > 0 ? kap : abs(kap) - 1;
        if (lkap > lx) lkap = -1;
        lind[k + 2] = lkap;
    }
    for (int k = -2; k <= 2; ++k) {
        int jkap = abs(kinit) + k;
        if (jkap <= 0) jkap = 0;
        int kap = jkap;
        if (kinit < 0 && abs(k) != 1) kap = -jkap;
        if (kinit > 0 && abs(k) == 1) kap = -jkap;
        int lkap = kap > 0 ? kap : -kap - 1;
        if (lkap > lx || le2 == 0 || (le2 == 1 && abs(k) == 2)) lkap = -1;
        lind[k + 6] = lkap;
    }
    lind[0] = -1;
    for (int k = 0; k < 9; ++k) bdiag[k] = 0.0;
    for (int k = 1; k <= 8; ++k) {
        if (lind[k] < 0) continue;
        int i2 = (le2 == 2 && k > 3) ? 5 : 3;
        double v = 0.5 / (2 * lind[k] + 1.0) / i2;
        bdiag[k] = 2.0 * (k <= 3 ? -v : v);   /* doubled: the spin trace of bcoef (ispin = 0) */
    }
}

static void snlm_tab(double xnlm[128][MTOT + 2], int lmaxp1, int mmaxp1) {
    const double afac = 1.0 / 64;
    double flg[211];
    flg[0] = 1.0; flg[1] = afac;
    for (int i = 2; i < 211; ++i) flg[i] = flg[i - 1] * i * afac;
    memset(xnlm, 0, sizeof(double) * 128 * (MTOT + 2));
    for (int il = 1; il <= lmaxp1 && il < 128; ++il)
        for (int im = 1; im <= (mmaxp1 < il ? mmaxp1 : il); ++im) {
            int l = il - 1, m = im - 1;
            xnlm[il][im] = sqrt((2 * l + 1) * flg[l - m] / flg[l + m]) * pow(afac, m);
        }
}

void orc_free(orc_phase *P) {
    if (!P) return;
    free(P->em); free(P->eref); free(P->lmax0); free(P->iz); free(P->ph); free(P->rkk);
    free(P->lmax); free(P->xk); free(P->ck);
    free(P);
}

int orc_load(const char *path, orc_phase **out) {
    *out = NULL;
    FILE *fp = fopen(path, "rb");
    if (!fp) return fail("cannot open phase file");
    fseek(fp, 0, SEEK_END);
    long size = ftell(fp);
    fseek(fp, 0, SEEK_SET);
    lines_t L = {0};
    L.buf = malloc(size + 1);
    if (fread(L.buf, 1, size, fp) != (size_t)size) { fclose(fp); free(L.buf); return fail("short read"); }
    fclose(fp);
    L.buf[size] = 0;
    int cap = 1024;
    L.lines = malloc(sizeof(char *) * cap);
    for (char *p = L.buf; *p;) {
        if (L.n == cap) { cap *= 2; L.lines = realloc(L.lines, sizeof(char *) * cap); }
        L.lines[L.n++] = p;
        char *nl = strchr(p, '\n');
        if (!nl) break;
        *nl = 0;
        p = nl + 1;
    }
    orc_phase *P = calloc(1, sizeof *P);
    int rc = -1;
    const char *h = next_line(&L);
    char tmp[16];
    int ints[9];
    for (int j = 0; j < 9; ++j) { memcpy(tmp, h + 1 + 5 * j, 4); tmp[4] = 0; ints[j] = atoi(tmp); }
    P->nsp = ints[0]; P->ne = ints[1]; P->ne1 = ints[2]; P->ne3 = ints[3]; P->nph = ints[4];
    P->ihole = ints[5]; P->ik0 = ints[6]; P->npack = ints[7]; P->ixc = ints[8];
    memcpy(tmp, h + 45, 11); tmp[11] = 0; P->rs = atof(tmp);
    memcpy(tmp, h + 56, 11); tmp[11] = 0; P->vint = atof(tmp);
    int ne = P->ne, nsp = P->nsp, nph = P->nph, np = P->npack;
    double three[3];
    if (read_reals(&L, np, 3, three)) goto done;
    P->rnrmav = three[0]; P->xmu = three[1]; P->edge = three[2];
    P->em = malloc(sizeof(cplx) * ne);
    P->eref = malloc(sizeof(cplx) * ne * nsp);
    if (read_cplx(&L, np, ne, P->em) || read_cplx(&L, np, ne * nsp, P->eref)) goto done;
    P->lmax0 = calloc(nph + 1, sizeof(int));
    P->iz = calloc(nph + 1, sizeof(int));
    cplx *blocks[16] = {0};
    if (nph + 1 > 16) { fail("too many potentials"); goto done; }
    for (int iph = 0; iph <= nph; ++iph) {
        const char *s = next_line(&L);
        if (!s) { fail("phase file ends early"); goto done; }
        memcpy(tmp, s + 1, 3); tmp[3] = 0; P->lmax0[iph] = atoi(tmp);
        memcpy(tmp, s + 5, 3); tmp[3] = 0; P->iz[iph] = atoi(tmp);
        int nl = 2 * P->lmax0[iph] + 1;
        blocks[iph] = malloc(sizeof(cplx) * ne * nl * nsp);
        for (int isp = 0; isp < nsp; ++isp)
            if (read_cplx(&L, np, ne * nl, blocks[iph] + (size_t)isp * ne * nl)) goto done;
    }
    P->ltot = 0;
    for (int iph = 0; iph <= nph; ++iph) if (P->lmax0[iph] > P->ltot) P->ltot = P->lmax0[iph];
    int nlt = 2 * P->ltot + 1;
    P->ph = calloc((size_t)ne * nlt * nsp * (nph + 1), sizeof(cplx));
    for (int iph = 0; iph <= nph; ++iph) {
        int l0 = P->lmax0[iph], nl = 2 * l0 + 1, lo = P->ltot - l0;
        for (int isp = 0; isp < nsp; ++isp)
            for (int ie = 0; ie < ne; ++ie)
                for (int j = 0; j < nl; ++j)     /* ll fastest within an energy */
                    P->ph[(((size_t)ie * nlt + lo + j) * nsp + isp) * (nph + 1) + iph] = blocks[iph][(size_t)isp * ne * nl + (size_t)ie * nl + j];
        free(blocks[iph]); blocks[iph] = NULL;
    }
    cplx *rk = malloc(sizeof(cplx) * ne * 8 * nsp);
    if (read_cplx(&L, np, ne * 8 * nsp, rk)) { free(rk); goto done; }
    P->rkk = malloc(sizeof(cplx) * ne * 8 * nsp);
    for (int isp = 0; isp < nsp; ++isp)          /* ie fastest, then kdif, then spin */
        for (int k = 0; k < 8; ++k)
            for (int ie = 0; ie < ne; ++ie) P->rkk[((size_t)ie * 8 + k) * nsp + isp] = rk[((size_t)isp * 8 + k) * ne + ie];
    free(rk);
    /* derived tables */
    P->lmax = calloc((size_t)ne * (nph + 1), sizeof(int));
    for (int iph = 0; iph <= nph; ++iph)
        for (int ie = 0; ie < ne; ++ie) {
            int lm = 0;
            for (int l = P->lmax0[iph]; l >= 0; --l) {
                lm = l;
                cplx a = P->ph[(((size_t)ie * nlt + l + P->ltot) * nsp + 0) * (nph + 1) + iph];
                cplx b = P->ph[(((size_t)ie * nlt + l + P->ltot) * nsp + nsp - 1) * (nph + 1) + iph];
                if (cabs(csin(a)) > ORC_PHMIN || cabs(csin(b)) > ORC_PHMIN) break;
            }
            P->lmax[ie * (nph + 1) + iph] = lm;
        }
    P->xk = malloc(sizeof(double) * ne);
    P->ck = malloc(sizeof(cplx) * ne);
    for (int ie = 0; ie < ne; ++ie) {
        double e = creal(P->em[ie]) - P->edge;
        P->xk[ie] = (e > 0) - (e < 0) ? ((e > 0) ? 1.0 : -1.0) * sqrt(fabs(2 * e)) : 0.0;
        cplx er = nsp == 1 ? P->eref[ie] : 0.5 * (P->eref[ie] + P->eref[(size_t)(nsp - 1) * ne + ie]);
        P->ck[ie] = csqrt(2.0 * (P->em[ie] - er));
    }
    if (setkap(P->ihole, &P->kinit, &P->linit)) goto done;
    P->ilinit = P->linit + 1;
    int lmx = 0;
    for (int ie = 0; ie < P->ne1; ++ie)
        for (int iph = 0; iph <= nph; ++iph) if (P->lmax[ie * (nph + 1) + iph] > lmx) lmx = P->lmax[ie * (nph + 1) + iph];
    P->lmaxp1 = lmx + 1;
    if (P->ltot + 2 > 126) { fail("ltot too large for the oracle tables"); goto done; }
    snlm_tab(P->xnlm, P->ltot + 1, MTOT + 1);
    bcoef_unpol(P->kinit, P->lind, P->bdiag);
    rc = 0;
done:
    free(L.lines); free(L.buf);
    if (rc) { orc_free(P); return -1; }
    *out = P;
    return 0;
}

/* Scalars and arrays a caller may want to inspect (tests compare them with the product's own reader). */
int orc_info(const orc_phase *P, int *ints /* ne, ne1, nph, ihole, ltot, lmaxp1, kinit, linit */) {
    ints[0] = P->ne; ints[1] = P->ne1; ints[2] = P->nph; ints[3] = P->ihole;
    ints[4] = P->ltot; ints[5] = P->lmaxp1; ints[6] = P->kinit; ints[7] = P->linit;
    return 0;
}
int orc_tables(const orc_phase *P, int *lmax /* ne1*(nph+1) */, double *xk /* ne1 */, double *ck /* 2*ne1 */) {
    for (int i = 0; i < P->ne1 * (P->nph + 1); ++i) lmax[i] = P->lmax[i];
    for (int i = 0; i < P->ne1; ++i) { xk[i] = P->xk[i]; ck[2 * i] = creal(P->ck[i]); ck[2 * i + 1] = cimag(P->ck[i]); }
    return 0;
}

/* The phase shifts of spin 0, ph_out[((ie * (ltot + 1) + l) * (nph + 1) + iph) * 2 + re/im] for l = -ltot side
 * only when sign < 0 (the j = l - 1/2 shifts FEFF's path finder tabulates, ph(ie, -l, iph)), else l >= 0; and ik0. */
int orc_phase_shifts(const orc_phase *P, int sign, double *ph_out, int *ik0) {
    const int nlt = 2 * P->ltot + 1;
    for (int ie = 0; ie < P->ne1; ++ie)
        for (int l = 0; l <= P->ltot; ++l)
            for (int iph = 0; iph <= P->nph; ++iph) {
                const int lo = sign < 0 ? -l : l;
                cplx v = P->ph[(((size_t)ie * nlt + lo + P->ltot) * P->nsp + 0) * (P->nph + 1) + iph];
                size_t o = (((size_t)ie * (P->ltot + 1) + l) * (P->nph + 1) + iph) * 2;
                ph_out[o] = creal(v);
                ph_out[o + 1] = cimag(v);
            }
    *ik0 = P->ik0;
    return 0;
}

/* ---------------------------------------------------------------- path geometry ------------------------------ */
static void trig(const double v[3], double *ct, double *st, double *cp, double *sp) {
    double x = v[0], y = v[1], z = v[2];
    double r = sqrt(x * x + y * y + z * z), rxy = sqrt(x * x + y * y);
    if (r < 1e-6) { *ct = 1.0; *st = 0.0; } else { *ct = z / r; *st = rxy / r; }
    if (rxy < 1e-6) { *cp = (*ct < 0) ? -1.0 : 1.0; *sp = 0.0; } else { *cp = x / rxy; *sp = y / rxy; }
}

static double argc(double x, double y, double fi) {
    if (fabs(x) < 1e-6) x = 0.0;
    if (fabs(y) < 1e-6) y = 0.0;
    return (x == 0.0 && y == 0.0) ? fi : atan2(y, x);
}

/* rat[j] j = 0..nleg (bohr), rat[0] = rat[nleg] = absorber. ri, beta, eta indexed 1..nleg. */
static void pathgeom(double rat[][3], int nleg, double *ri, double *beta, double *eta) {
    int nsc = nleg - 1;
    double alpha[NLEG_MAX + 2], gamma[NLEG_MAX + 2];
    for (int j = 1; j <= nleg; ++j) {
        int i, ip1, im1;
        if (j == nsc + 1) { i = 0; ip1 = 1; im1 = nsc; } else { i = j; ip1 = j + 1; im1 = j - 1; }
        double vo[3], vi[3];
        for (int c = 0; c < 3; ++c) { vo[c] = rat[ip1][c] - rat[i][c]; vi[c] = rat[i][c] - rat[im1][c]; }
        double ctp, stp, cpp, spp, ct, st, cp, sp;
        trig(vo, &ctp, &stp, &cpp, &spp);
        trig(vi, &ct, &st, &cp, &sp);
        double cppp = cp * cpp + sp * spp, sppp = spp * cp - cpp * sp;
        double phi = atan2(sp, cp), phip = atan2(spp, cpp);
        double ar = -(st * ctp - ct * stp * cppp), ai = stp * sppp;            /* alph */
        double gr = -(st * ctp * cppp - ct * stp), gi = -st * sppp;           /* gamm */
        double b = ct * ctp + st * stp * cppp;
        if (b > 1.0) b = 1.0;
        if (b < -1.0) b = -1.0;
        double u1[3] = {st * cp, st * sp, ct}, u2[3] = {stp * cpp, stp * spp, ctp};
        double cx = u1[1] * u2[2] - u1[2] * u2[1], cy = u1[2] * u2[0] - u1[0] * u2[2], cz = u1[0] * u2[1] - u1[1] * u2[0];
        beta[j] = atan2(sqrt(cx * cx + cy * cy + cz * cz + 1e-30), b);
        double a_j = argc(ar, ai, phip - phi), g_j = argc(gr, gi, 0.0);
        alpha[j] = M_PI - g_j;
        gamma[j] = M_PI - a_j;
        ri[j] = sqrt(vi[0] * vi[0] + vi[1] * vi[1] + vi[2] * vi[2]);
    }
    alpha[0] = alpha[nleg];
    for (int j = 1; j <= nleg; ++j) eta[j] = alpha[j - 1] + gamma[j];
    eta[0] = 0.0;
    eta[nleg + 1] = 0.0;
}

/* ---------------------------------------------------------------- lambda sets -------------------------------- */
typedef struct { int m[LAMTOT + 1], n[LAMTOT + 1], lamx, laml0x, mmaxp1, nmax; } lamset;

static int setlam(int iorder, int nsc, int ilinit, lamset *S) {
    int mmax, nmax, iord;
    if (nsc == 1) { mmax = nmax = ilinit; iord = 2 * nmax + mmax; }
    else if (iorder < 10) { iord = iorder; mmax = iorder; nmax = iorder / 2; }
    else return fail("icalc >= 10 not ported");
    int m0[LAMTOT], n0[LAMTOT], cnt = 0, filled = 0;
    for (int n = 0; n <= nmax && !filled; ++n)
        for (int m = 0; m <= mmax; ++m) {
            if (2 * n + m > iord) continue;
            if (cnt >= LAMTOT) { filled = 1; break; }
            m0[cnt] = -m; n0[cnt] = n; ++cnt;
            if (m == 0) continue;
            if (cnt >= LAMTOT) { filled = 1; break; }
            m0[cnt] = m; n0[cnt] = n; ++cnt;
        }
    int k = 1;
    for (int i = 0; i < cnt; ++i) if (n0[i] <= ilinit && abs(m0[i]) <= ilinit) { S->m[k] = m0[i]; S->n[k] = n0[i]; ++k; }
    S->laml0x = k - 1;
    for (int i = 0; i < cnt; ++i) if (!(n0[i] <= ilinit && abs(m0[i]) <= ilinit)) { S->m[k] = m0[i]; S->n[k] = n0[i]; ++k; }
    S->lamx = cnt;
    S->m[0] = S->n[0] = 0;
    S->mmaxp1 = -1000; S->nmax = 0;
    for (int i = 1; i <= cnt; ++i) {
        if (S->m[i] + 1 > S->mmaxp1) S->mmaxp1 = S->m[i] + 1;
        if (S->n[i] > S->nmax) S->nmax = S->n[i];
    }
    if (S->nmax > NTOT || S->mmaxp1 > MTOT + 1) return fail("lambda set exceeds mtot/ntot");
    return 0;
}

/* ---------------------------------------------------------------- rot3i ------------------------------------ */
#define DRI_L 128
#define DRI_M (2 * MTOT + 2)
typedef double dri_t[DRI_L][DRI_M][DRI_M];   /* dri[il][m1+MTOT+1][m2+MTOT+1] */

static void rot3i(int lxp1, int mxp1, double beta, dri_t dri, double *work /* 2 * size * size */) {
    int nm = mxp1, ndm = lxp1 + nm - 1;
    int size = (2 * lxp1 + 1 > ndm + 1 ? 2 * lxp1 + 1 : ndm + 1) + 1;
    double xc = cos(beta / 2), xs = sin(beta / 2), s = sin(beta);
    double *prev = work, *cur = work + (size_t)size * size;
#define D(a, n, m) a[(size_t)(n) * size + (m)]
    memset(dri, 0, sizeof(dri_t));
    memset(prev, 0, sizeof(double) * size * size);
    memset(cur, 0, sizeof(double) * size * size);
    /* l = 1 (il = 1): d = 1 */
    D(prev, 1, 1) = 1.0;
    dri[1][MTOT + 1][MTOT + 1] = 1.0;
    if (lxp1 < 2) return;
    double *d2 = cur;
    D(d2, 1, 1) = xc * xc;
    D(d2, 1, 2) = s / sqrt(2.0);
    D(d2, 1, 3) = xs * xs;
    D(d2, 2, 1) = -D(d2, 1, 2);
    D(d2, 2, 2) = cos(beta);
    D(d2, 2, 3) = D(d2, 1, 2);
    D(d2, 3, 1) = D(d2, 1, 3);
    D(d2, 3, 2) = -D(d2, 2, 3);
    D(d2, 3, 3) = D(d2, 1, 1);
    { double *t = prev; prev = cur; cur = t; }
    for (int il = 2; il <= lxp1; ++il) {
        if (il >= 3) {
            int l = il;
            memset(cur, 0, sizeof(double) * size * size);
            int ln = 2 * l - 1 < ndm ? 2 * l - 1 : ndm, lm = 2 * l - 3 < ndm ? 2 * l - 3 : ndm;
            for (int n = 1; n <= ln; ++n) {
                for (int m = 1; m <= lm; ++m) {
                    double t1 = (double)(2 * l - 1 - n) * (2 * l - 2 - n);
                    double t = (double)(2 * l - 1 - m) * (2 * l - 2 - m);
                    double f1 = sqrt(t1 / t), f2 = sqrt((double)(2 * l - 1 - n) * (n - 1) / t), f3 = sqrt((double)(n - 2) * (n - 1) / t);
                    double v = f1 * xc * xc * D(prev, n, m);
                    if (n - 1 > 0) v -= f2 * s * D(prev, n - 1, m);
                    if (n - 2 > 0) v += f3 * xs * xs * D(prev, n - 2, m);
                    D(cur, n, m) = v;
                    if (n > 2 * l - 3) D(cur, m, n) = (((n - m) & 1) ? -1.0 : 1.0) * v;
                }
                if (n > 2 * l - 3) {
                    D(cur, 2 * l - 2, 2 * l - 2) = D(cur, 2, 2);
                    D(cur, 2 * l - 1, 2 * l - 2) = -D(cur, 1, 2);
                    D(cur, 2 * l - 2, 2 * l - 1) = -D(cur, 2, 1);
                    D(cur, 2 * l - 1, 2 * l - 1) = D(cur, 1, 1);
                }
            }
            double *t = prev; prev = cur; cur = t;
        }
        int mx = il - 1 < mxp1 - 1 ? il - 1 : mxp1 - 1;
        for (int m1 = -mx; m1 <= mx; ++m1)
            for (int m2 = -mx; m2 <= mx; ++m2) dri[il][m1 + MTOT + 1][m2 + MTOT + 1] = D(prev, m1 + il, m2 + il);
    }
#undef D
}

/* Exact rotation matrices: Wigner's d^j_{m1 m2}(beta) through Jacobi polynomials,
 *   d = (-1)^lam sqrt(C(2j-k, k+a) / C(k+b, b)) sin(beta/2)^a cos(beta/2)^b P_k^(a,b)(cos beta),
 *   k = min(j+m2, j-m2, j+m1, j-m1), a, lam by which bound is reached, b = 2j - 2k - a,
 * in Edmonds' convention (d^1_{10} = -sin(beta)/sqrt 2), the convention of FEFF's rot3i and of the batched stage's
 * three-term recursion (src/feffgpu/stage.py:41-93). This is what the batched stage computes; FEFF's rot3i
 * (Edmonds' recursion on truncated matrices) drifts from it above l ~ 20 (rotmode 0 reproduces that). */
static double jacobi_p(int k, double a, double b, double x) {
    if (k == 0) return 1.0;
    double p0 = 1.0, p1 = 0.5 * (a - b + (a + b + 2.0) * x);
    for (int n = 2; n <= k; ++n) {
        double s = 2.0 * n + a + b;
        double c1 = 2.0 * n * (n + a + b) * (s - 2.0);
        double c2 = (s - 1.0) * (s * (s - 2.0) * x + a * a - b * b);
        double c3 = 2.0 * (n + a - 1.0) * (n + b - 1.0) * s;
        double p2 = (c2 * p1 - c3 * p0) / c1;
        p0 = p1;
        p1 = p2;
    }
    return p1;
}
static double lbinom(int n, int k) { return lgamma(n + 1.0) - lgamma(k + 1.0) - lgamma(n - k + 1.0); }
static double wigner_d(int j, int m1, int m2, double beta) {
    if (abs(m1) > j || abs(m2) > j) return 0.0;
    int k = j + m2, a = m1 - m2, lam = m1 - m2;
    if (j - m2 < k) { k = j - m2; a = m2 - m1; lam = 0; }
    if (j + m1 < k) { k = j + m1; a = m2 - m1; lam = 0; }
    if (j - m1 < k) { k = j - m1; a = m1 - m2; lam = m1 - m2; }
    int b = 2 * j - 2 * k - a;
    double pref = exp(0.5 * (lbinom(2 * j - k, k + a) - lbinom(k + b, b)));
    double v = pref * pow(sin(0.5 * beta), a) * pow(cos(0.5 * beta), b) * jacobi_p(k, a, b, cos(beta));
    return (lam & 1) ? -v : v;
}
static void rot_exact(int lxp1, int mxp1, double beta, dri_t dri) {
    memset(dri, 0, sizeof(dri_t));
    for (int il = 1; il <= lxp1; ++il) {
        int mx = il - 1 < mxp1 - 1 ? il - 1 : mxp1 - 1;
        for (int m1 = -mx; m1 <= mx; ++m1)
            for (int m2 = -mx; m2 <= mx; ++m2) dri[il][m1 + MTOT + 1][m2 + MTOT + 1] = wigner_d(il - 1, m1, m2, beta);
    }
}
/* for tests: d^l_{m1 m2}(beta) of either algorithm */
int orc_wigner(int rotmode, int l, int m1, int m2, double beta, double *out) {
    static __thread dri_t d;
    static __thread double work[2 * 260 * 260];
    if (l < 0 || l > 120 || abs(m1) > 3 || abs(m2) > 3) return fail("orc_wigner: l or m out of range");
    if (rotmode == 0) rot3i(l + 1, 4, beta, d, work); else rot_exact(l + 1, 4, beta, d);
    *out = d[l + 1][m1 + MTOT + 1][m2 + MTOT + 1];
    return 0;
}

/* ---------------------------------------------------------------- sclmz ------------------------------------ */
#define CLM_M (MTOT + NTOT + 2)
typedef cplx clm_t[DRI_L][CLM_M];

static void sclmz(cplx rho, int lmaxp1, int mmaxp1, clm_t c) {
    memset(c, 0, sizeof(cplx) * CLM_M * (size_t)(lmaxp1 + 2));
    cplx z = -I / rho, cmm = 1.0;
    c[1][1] = 1.0;
    c[2][1] = c[1][1] - z;
    int lmax = lmaxp1 - 1;
    for (int il = 2; il <= lmax; ++il) c[il + 1][1] = c[il - 1][1] - z * (2 * il - 1) * c[il][1];
    int mmxp1 = mmaxp1 < lmaxp1 ? mmaxp1 : lmaxp1;
    for (int im = 2; im <= mmxp1; ++im) {
        int m = im - 1, imp1 = im + 1;
        cmm = -cmm * (2 * m - 1) * z;
        c[im][im] = cmm;
        c[imp1][im] = cmm * (2 * m + 1) * (1.0 - im * z);
        for (int il = imp1; il <= lmax; ++il) {
            int l = il - 1;
            c[il + 1][im] = c[l][im] - (2 * l + 1) * z * (c[il][im] + c[il][m]);
        }
    }
}

/* ---------------------------------------------------------------- one path --------------------------------- */
typedef struct {
    dri_t dri[NLEG_MAX + 1];
    clm_t clm[NLEG_MAX + 1];
    double *rwork;
    cplx fmat[NLEG_MAX + 1][LAMTOT + 1][LAMTOT + 1];   /* fmat[ilegp][lam1][lam2] */
    cplx gam[DRI_L][MTOT + 2][NTOT + 2], gamtl[DRI_L][MTOT + 2][NTOT + 2];
} orc_work;

static cplx ph_at(const orc_phase *P, int ie, int l, int iph) {
    int nlt = 2 * P->ltot + 1;
    return P->ph[(((size_t)ie * nlt + l + P->ltot) * P->nsp + 0) * (P->nph + 1) + iph];
}

static void fmtrxi(const orc_phase *P, orc_work *W, int lam1x, int lam2x, int ie, int ileg, int ilegp, const lamset *S,
                   const double *eta, const int *ipot) {
    int ilx = P->lmax[ie * (P->nph + 1) + ipot[ilegp]] + 1;
    memset(W->gam, 0, sizeof(W->gam[0]) * (size_t)(ilx + 1));
    memset(W->gamtl, 0, sizeof(W->gamtl[0]) * (size_t)(ilx + 1));
    int lmx = lam1x > lam2x ? lam1x : lam2x;
    for (int il = 1; il <= ilx; ++il) {
        int ll = il - 1;
        cplx tm = (cexp(2 * I * ph_at(P, ie, -ll, ipot[ilegp])) - 1) / (2 * I);
        cplx tp = (cexp(2 * I * ph_at(P, ie, ll, ipot[ilegp])) - 1) / (2 * I);
        cplx tltl = tm * (ll + 1) + tp * ll;
        for (int lam = 1; lam <= lmx; ++lam) {
            int m = S->m[lam];
            if (m < 0) continue;
            int im = m + 1;
            if (im > il) continue;
            int in = S->n[lam] + 1, imn = in + m;
            if (lam <= lam1x) W->gam[il][im][in] = imn <= il ? P->xnlm[il][im] * ((m & 1) ? -1.0 : 1.0) * W->clm[ileg][il][imn] : 0.0;
            if (lam <= lam2x) W->gamtl[il][im][in] = tltl / P->xnlm[il][im] * W->clm[ilegp][il][in];
        }
    }
    for (int l1 = 1; l1 <= lam1x; ++l1) {
        int m1 = S->m[l1], in1 = S->n[l1] + 1, iam1 = abs(m1) + 1;
        for (int l2 = 1; l2 <= lam2x; ++l2) {
            int m2 = S->m[l2], in2 = S->n[l2] + 1, iam2 = abs(m2) + 1;
            int imn1 = iam1 + in1 - 1;
            int ilmin = iam1;
            if (iam2 > ilmin) ilmin = iam2;
            if (imn1 > ilmin) ilmin = imn1;
            if (in2 > ilmin) ilmin = in2;
            cplx ct = 0;
            for (int il = ilmin; il <= ilx; ++il) {
                if (abs(m1) >= il || abs(m2) >= il) continue;
                ct += W->gam[il][iam1][in1] * W->gamtl[il][iam2][in2] * W->dri[ilegp][il][m1 + MTOT + 1][m2 + MTOT + 1];
            }
            if (eta[ileg] != 0.0) ct *= cexp(-I * eta[ileg] * m1);
            W->fmat[ilegp][l1][l2] = ct;
        }
    }
}

static void mmtrxi(const orc_phase *P, orc_work *W, int lam1x, int ie, int ileg, int ilegp, const lamset *S,
                   const double *eta, cplx bmati[2 * MTOT + 1][9][2 * MTOT + 1][9]) {
    const int *lind = P->lind;
    int iln = 1000, ilx = -1000;
    for (int k = 1; k <= 8; ++k) if (lind[k] >= 0) { if (lind[k] + 1 < iln) iln = lind[k] + 1; if (lind[k] + 1 > ilx) ilx = lind[k] + 1; }
    memset(W->gam, 0, sizeof(W->gam[0]) * (size_t)(ilx + 1));
    memset(W->gamtl, 0, sizeof(W->gamtl[0]) * (size_t)(ilx + 1));
    for (int il = iln; il <= ilx; ++il) {
        double tltl = 2 * il - 1;
        for (int lam = 1; lam <= lam1x; ++lam) {
            int m = S->m[lam];
            if (m < 0) continue;
            int im = m + 1;
            if (im > il) continue;
            int in = S->n[lam] + 1, imn = in + m;
            W->gam[il][im][in] = imn <= il ? P->xnlm[il][im] * ((m & 1) ? -1.0 : 1.0) * W->clm[ileg][il][imn] : 0.0;
            W->gamtl[il][im][in] = tltl / P->xnlm[il][im] * W->clm[ilegp][il][in];
        }
    }
    const int ns = P->nsp;
    for (int l1 = 1; l1 <= lam1x; ++l1) {
        int m1 = S->m[l1], in1 = S->n[l1] + 1, iam1 = abs(m1) + 1;
        for (int l2 = 1; l2 <= lam1x; ++l2) {
            int m2 = S->m[l2], in2 = S->n[l2] + 1, iam2 = abs(m2) + 1;
            cplx f = 0;
            for (int k1 = 1; k1 <= 8; ++k1)
                for (int k2 = 1; k2 <= 8; ++k2) {
                    int a = lind[k1] + 1, b = lind[k2] + 1;
                    if (a > 0 && b > 0 && iam1 <= a && iam2 <= b)
                        f += bmati[m1 + MTOT][k1][m2 + MTOT][k2] * P->rkk[((size_t)ie * 8 + k1 - 1) * ns] * P->rkk[((size_t)ie * 8 + k2 - 1) * ns] *
                             W->gam[a][iam1][in1] * W->gamtl[b][iam2][in2];
                }
            W->fmat[ilegp][l1][l2] = f * cexp(-I * eta[ileg] * m1);
        }
    }
}

/* cchi on the first ne1 energies of the phase grid; coords (nleg x 3, A) absorber first at the origin; ipots (nleg)
 * with the absorber first. Returns reff in A. geom (optional, 3 x nleg): ri (A), beta, eta of legs 1..nleg. */
int orc_cchi(const orc_phase *P, int nleg, const double *coords, const long *ipots, int iorder, int rotmode,
             double *cchi_out, double *reff_out, double *geom) {
    if (nleg < 2 || nleg > NLEG_MAX) return fail("nleg out of range 2..9");
    int nsc = nleg - 1;
    double rat[NLEG_MAX + 2][3];
    for (int j = 1; j < nleg; ++j) for (int c = 0; c < 3; ++c) rat[j][c] = coords[3 * j + c] / ORC_BOHR;
    for (int c = 0; c < 3; ++c) rat[nleg][c] = rat[0][c] = coords[c] / ORC_BOHR;
    int ipot[NLEG_MAX + 2];
    for (int j = 1; j < nleg; ++j) ipot[j] = (int)ipots[j];
    ipot[nleg] = (int)ipots[0];
    ipot[0] = ipot[nleg];
    for (int j = 0; j <= nleg; ++j) if (ipot[j] < 0 || ipot[j] > P->nph) return fail("potential index out of range");
    double ri[NLEG_MAX + 2], beta[NLEG_MAX + 2], eta[NLEG_MAX + 2];
    pathgeom(rat, nleg, ri, beta, eta);
    double reff = 0;
    for (int j = 1; j <= nleg; ++j) reff += ri[j];
    reff /= 2;
    lamset S;
    if (setlam(iorder, nsc, P->ilinit, &S)) return -1;
    orc_work *W = malloc(sizeof *W);
    if (!W) return fail("out of memory");
    int size = (2 * P->lmaxp1 + 1 > P->lmaxp1 + S.mmaxp1 ? 2 * P->lmaxp1 + 1 : P->lmaxp1 + S.mmaxp1) + 1;
    W->rwork = malloc(sizeof(double) * 2 * size * size);
    for (int isc = 1; isc <= nleg; ++isc) {
        if (rotmode == 0) rot3i(P->lmaxp1, S.mmaxp1, beta[isc], W->dri[isc], W->rwork);
        else rot_exact(P->lmaxp1, S.mmaxp1, beta[isc], W->dri[isc]);
    }
    /* termination coefficients (mmtr, ipol = 0) */
    static __thread cplx bmati[2 * MTOT + 1][9][2 * MTOT + 1][9];
    memset(bmati, 0, sizeof bmati);
    int lxx = MTOT < P->ilinit ? MTOT : P->ilinit;
    for (int mu1 = -lxx; mu1 <= lxx; ++mu1)
        for (int mu2 = -lxx; mu2 <= lxx; ++mu2)
            for (int k1 = 1; k1 <= 8; ++k1) {
                int l1 = P->lind[k1] + 1;
                if (l1 > 0) bmati[mu1 + MTOT][k1][mu2 + MTOT][k1] += P->bdiag[k1] * W->dri[nleg][l1][mu1 + MTOT + 1][mu2 + MTOT + 1];
            }
    int mnmxp1 = S.mmaxp1 + S.nmax;
    const int nph1 = P->nph + 1;
    for (int ie = 0; ie < P->ne1; ++ie) {
        cplx ck = P->ck[ie];
        cplx rho[NLEG_MAX + 2];
        rho[0] = 0;
        for (int j = 1; j <= nleg; ++j) rho[j] = ck * ri[j];
        if (cabs(ck) <= ORC_EPS) { cchi_out[2 * ie] = cchi_out[2 * ie + 1] = 0; continue; }
        for (int ileg = 1; ileg <= nleg; ++ileg) {
            int isc0 = ileg > 1 ? ileg - 1 : nleg;
            int a = P->lmax[ie * nph1 + ipot[isc0]] + 1, b = P->lmax[ie * nph1 + ipot[ileg]] + 1;
            int lxp1 = a > b ? a : b;
            int mnp1 = lxp1 < mnmxp1 ? lxp1 : mnmxp1;
            sclmz(rho[ileg], lxp1, mnp1, W->clm[ileg]);
        }
        memset(W->fmat, 0, sizeof(W->fmat[0]) * (size_t)(nleg + 1));
        fmtrxi(P, W, S.lamx, S.laml0x, ie, 2, 1, &S, eta, ipot);
        if (nleg > 2) fmtrxi(P, W, S.laml0x, S.lamx, ie, nleg, nleg - 1, &S, eta, ipot);
        for (int ilegp = 2; ilegp < nsc; ++ilegp) fmtrxi(P, W, S.lamx, S.lamx, ie, ilegp + 1, ilegp, &S, eta, ipot);
        /* pmat = f(nsc) ... f(1), lamx x laml0x */
        cplx pm[LAMTOT + 1][LAMTOT + 1], tmp[LAMTOT + 1][LAMTOT + 1];
        for (int a = 1; a <= S.lamx; ++a) for (int b = 1; b <= S.laml0x; ++b) pm[a][b] = W->fmat[1][a][b];
        for (int isc = 2; isc < nleg; ++isc) {
            for (int a = 1; a <= S.lamx; ++a)
                for (int b = 1; b <= S.laml0x; ++b) {
                    cplx s = 0;
                    for (int c = 1; c <= S.lamx; ++c) s += W->fmat[isc][a][c] * pm[c][b];
                    tmp[a][b] = s;
                }
            for (int a = 1; a <= S.lamx; ++a) for (int b = 1; b <= S.laml0x; ++b) pm[a][b] = tmp[a][b];
        }
        mmtrxi(P, W, S.laml0x, ie, 1, nleg, &S, eta, bmati);
        cplx ptrac = 0;
        for (int a = 1; a <= S.laml0x; ++a) for (int b = 1; b <= S.laml0x; ++b) ptrac += W->fmat[nleg][a][b] * pm[b][a];
        cplx srho = 0, prho = 1;
        for (int j = 1; j <= nleg; ++j) { srho += rho[j]; prho *= rho[j]; }
        cplx cfac = cexp(I * (srho - 2 * P->xk[ie] * reff)) / prho;
        cplx v = ptrac * cfac;
        cchi_out[2 * ie] = creal(v);
        cchi_out[2 * ie + 1] = cimag(v);
    }
    free(W->rwork);
    free(W);
    *reff_out = reff * ORC_BOHR;
    if (geom) for (int j = 1; j <= nleg; ++j) { geom[j - 1] = ri[j] * ORC_BOHR; geom[nleg + j - 1] = beta[j]; geom[2 * nleg + j - 1] = eta[j]; }
    return 0;
}

/* ---------------------------------------------------------------- fine grid -------------------------------- */
/* chi(q) = A(q) sin(Phi(q) + 2 q reff) exp(-2 sigma2 q^2): A and Phi linear interpolations on xk/BOHR of
 * sqrt(|cchi|^2 + 1e-30) and of the unwrapped phase of cchi (src/feffgpu/stage.py:353-388). */
static void chi_fine_one(const orc_phase *P, const double *cchi, double reff, double sigma2, int nq, const double *q,
                         double *out, double *amp, double *pha) {
    int E = P->ne1;
    for (int e = 0; e < E; ++e) {
        double re = cchi[2 * e], im = cchi[2 * e + 1];
        amp[e] = sqrt(re * re + im * im + 1e-30);
        pha[e] = amp[e] >= 1e-16 ? atan2(im, re) : 0.0;
    }
    double acc = pha[0], prev = pha[0];
    for (int e = 1; e < E; ++e) {
        double d = pha[e] - prev;
        prev = pha[e];
        d -= 2 * M_PI * nearbyint(d / (2 * M_PI));
        acc += d;
        pha[e] = acc;
    }
    const double k0 = P->xk[0] / ORC_BOHR, k1 = P->xk[E - 1] / ORC_BOHR;
    for (int i = 0; i < nq; ++i) {
        double qn = q[i] < k0 ? k0 : (q[i] > k1 ? k1 : q[i]);
        /* searchsorted(kf, qn, 'right') - 1, clipped to 0..E-2 */
        int lo = 0, hi = E;
        while (lo < hi) { int mid = (lo + hi) / 2; if (P->xk[mid] / ORC_BOHR <= qn) lo = mid + 1; else hi = mid; }
        int j = lo - 1;
        if (j < 0) j = 0;
        if (j > E - 2) j = E - 2;
        double kj = P->xk[j] / ORC_BOHR, kj1 = P->xk[j + 1] / ORC_BOHR;
        double f = (qn - kj) / (kj1 - kj);
        double a = (1 - f) * amp[j] + f * amp[j + 1];
        double ph = (1 - f) * pha[j] + f * pha[j + 1];
        double v = a * sin(ph + 2 * q[i] * reff);
        if (sigma2 != 0.0) v *= exp(-2.0 * sigma2 * q[i] * q[i]);
        out[i] = v;
    }
}

int orc_chi_fine(const orc_phase *P, int npath, const double *cchi, const double *reff, const double *sigma2, int nq,
                 const double *q, double *chi) {
    int E = P->ne1;
    double *amp = malloc(sizeof(double) * E), *pha = malloc(sizeof(double) * E);
    for (int p = 0; p < npath; ++p)
        chi_fine_one(P, cchi + (size_t)2 * E * p, reff[p], sigma2 ? sigma2[p] : 0.0, nq, q, chi + (size_t)nq * p, amp, pha);
    free(amp); free(pha);
    return 0;
}

/* ---------------------------------------------------------------- batched path list ------------------------ */
/* The fused path-list evaluation: per path p, coords = [0, wrap(pos[atoms[p, :nleg-1]] - pos[absorber[p]])],
 * chi_p(q) on the grid q, chi_seg[segment[p]] += deg[p] chi_p summed in path order (deterministic).
 * phases[edge[p]] selects the phase file. cell: NULL (cluster), 1 value (cubic length) or 9 (rows are lattice
 * vectors). chi_path (optional) receives chi_p without the degeneracy. OpenMP over paths with nthreads threads. */
static double round_half_even(double x) { return nearbyint(x); }

static void wrap_vec(double v[3], const double *cell, int ncell, const double *inv) {
    if (ncell == 1) {
        for (int c = 0; c < 3; ++c) v[c] -= cell[0] * round_half_even(v[c] / cell[0]);
    } else if (ncell == 9) {
        double f[3];
        for (int c = 0; c < 3; ++c) f[c] = v[0] * inv[0 * 3 + c] + v[1] * inv[1 * 3 + c] + v[2] * inv[2 * 3 + c];
        for (int c = 0; c < 3; ++c) f[c] -= round_half_even(f[c]);
        for (int c = 0; c < 3; ++c) v[c] = f[0] * cell[0 * 3 + c] + f[1] * cell[1 * 3 + c] + f[2] * cell[2 * 3 + c];
    }
}

static void inv3(const double *a, double *o) {
    double det = a[0] * (a[4] * a[8] - a[5] * a[7]) - a[1] * (a[3] * a[8] - a[5] * a[6]) + a[2] * (a[3] * a[7] - a[4] * a[6]);
    o[0] = (a[4] * a[8] - a[5] * a[7]) / det; o[1] = (a[2] * a[7] - a[1] * a[8]) / det; o[2] = (a[1] * a[5] - a[2] * a[4]) / det;
    o[3] = (a[5] * a[6] - a[3] * a[8]) / det; o[4] = (a[0] * a[8] - a[2] * a[6]) / det; o[5] = (a[2] * a[3] - a[0] * a[5]) / det;
    o[6] = (a[3] * a[7] - a[4] * a[6]) / det; o[7] = (a[1] * a[6] - a[0] * a[7]) / det; o[8] = (a[0] * a[4] - a[1] * a[3]) / det;
}

int orc_chi_paths(const orc_phase *const *phases, int nphase, int npath, int nmax, const double *pos, const int *atoms,
                  const int *nleg, const int *absorber, const long *ipots, const double *deg, const int *segment,
                  const int *edge, const double *sigma2, int nseg, int nq, const double *q, const double *cell, int ncell,
                  int iorder, int rotmode, int nthreads, double *chi_seg, double *chi_path, double *reff_out) {
    double inv[9] = {0};
    if (ncell == 9) inv3(cell, inv);
    int Emax = 0;
    for (int i = 0; i < nphase; ++i) if (phases[i]->ne1 > Emax) Emax = phases[i]->ne1;
    double *chi_all = malloc(sizeof(double) * (size_t)npath * nq);
    int err = 0;
    if (nthreads <= 0) nthreads = 1;
#pragma omp parallel num_threads(nthreads)
    {
        double *cc = malloc(sizeof(double) * 2 * Emax), *amp = malloc(sizeof(double) * Emax), *pha = malloc(sizeof(double) * Emax);
#pragma omp for schedule(dynamic, 16)
        for (int p = 0; p < npath; ++p) {
            int n = nleg[p];
            int ed = edge ? edge[p] : 0;
            if (n < 2 || n > nmax || ed < 0 || ed >= nphase) { err = 1; continue; }
            double coords[3 * NLEG_MAX];
            long ip[NLEG_MAX];
            const double *pa = pos + 3 * (size_t)absorber[p];
            coords[0] = coords[1] = coords[2] = 0.0;
            ip[0] = ipots[(size_t)p * nmax];
            for (int j = 1; j < n; ++j) {
                const double *pj = pos + 3 * (size_t)atoms[(size_t)p * (nmax - 1) + j - 1];
                double v[3] = {pj[0] - pa[0], pj[1] - pa[1], pj[2] - pa[2]};
                wrap_vec(v, cell, ncell, inv);
                for (int c = 0; c < 3; ++c) coords[3 * j + c] = v[c];
                ip[j] = ipots[(size_t)p * nmax + j];
            }
            double reff;
            if (orc_cchi(phases[ed], n, coords, ip, iorder, rotmode, cc, &reff, NULL)) { err = 2; continue; }
            if (reff_out) reff_out[p] = reff;
            chi_fine_one(phases[ed], cc, reff, sigma2 ? sigma2[p] : 0.0, nq, q, chi_all + (size_t)nq * p, amp, pha);
        }
        free(cc); free(amp); free(pha);
    }
    if (err) { free(chi_all); return fail(err == 1 ? "bad path (nleg or edge index)" : orc_err); }
    if (chi_path) memcpy(chi_path, chi_all, sizeof(double) * (size_t)npath * nq);
    if (chi_seg) {
        memset(chi_seg, 0, sizeof(double) * (size_t)nseg * nq);
        for (int p = 0; p < npath; ++p) {
            int s = segment ? segment[p] : 0;
            if (s < 0 || s >= nseg) { free(chi_all); return fail("segment index out of range"); }
            for (int i = 0; i < nq; ++i) chi_seg[(size_t)s * nq + i] += deg[p] * chi_all[(size_t)nq * p + i];
        }
    }
    free(chi_all);
    return 0;
}
