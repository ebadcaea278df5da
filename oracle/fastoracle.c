/*
 * fastoracle.c -- CPU restatement of the FastS structured NS iteration (test oracle).
 *
 * TEST INFRASTRUCTURE ONLY (see fastoracle.h).  Each routine cites the reference kernel it
 * restates; loop nests, index ranges and the floating-point operation order of every cell
 * follow the reference for a single OpenMP thread and a single cache block.  Build with
 * -ffp-contract=off.
 *
 * Conventions: 1-based (i,j,k) as in the reference, 0-based linear offsets in C.
 */
#include "fastoracle.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <stdio.h>

enum { NIJK = 0, NIJK_MTR = 5, NIJK_XYZ = 10, IJKV = 20, ITYPZONE = 25, IFLOW = 27, ITYPCP = 29,
       SIZE_SSDOM = 30, KFLUDOM = 33, LALE = 34, NEQ = 36, NEQ_IJ = 37, NEQ_K = 38, NEQ_COE = 39,
       NEQ_VENT = 40,
       NDIMDX = 41, NDIMDX_XYZ = 42, NDIMDX_MTR = 43, DTLOC = 46, RK = 52, LEVEL = 53,
       NSSITER = 58, CACHEBLCKI = 59, CACHEBLCKJ = 60, CACHEBLCKK = 61, SLOPE = 64, PT_BC = 70, LU_MATCH = 75, SENSORTYPE = 85,
       BC_TYPE = 0, BC_IDIR = 1, BC_FEN = 2, BC_NBDATA = 8 };
enum { DTC = 0, GAMMA = 1, CVINF = 2, ROINF = 3, PINF = 4, VINF = 5, TINF = 6,
       RONUTILDEINF = 8, PRANDT = 10, XMUL0 = 11, TEMP0 = 12, CS = 13, CFL = 15, VXINF = 19,
       VYINF = 20, VZINF = 21, PSIROE = 34, PRANDT_TUR = 38, WIG_AMP = 72 };

/* ------------------------------------------------------------------------------------ */
/* index helpers: formule_param.h (inddm), formule_mtr_param.h (indmtr),                 */
/* formule_xyz_param.h (indcg)                                                          */
typedef struct {
    const int* p;
    long ni, nj;          /* NIJK(1), NIJK(2) */
    long nf, nfk;         /* ghost layers */
    long mi, mj, mk, mf, mfk;  /* metric strides */
    long xi, xj, xf, xfk; /* vertex grid */
    long nd;              /* NDIMDX */
    long ndm;             /* NDIMDX_MTR */
    long incj, inck;
} Ix;

static Ix mkix(const int* p)
{
    Ix z;
    z.p = p;
    z.ni = p[NIJK]; z.nj = p[NIJK + 1]; z.nf = p[NIJK + 3]; z.nfk = p[NIJK + 4];
    z.mi = p[NIJK_MTR]; z.mj = p[NIJK_MTR + 1]; z.mk = p[NIJK_MTR + 2];
    z.mf = p[NIJK_MTR + 3]; z.mfk = p[NIJK_MTR + 4];
    z.xi = p[NIJK_XYZ]; z.xj = p[NIJK_XYZ + 1]; z.xf = p[NIJK_XYZ + 3]; z.xfk = p[NIJK_XYZ + 4];
    z.nd = p[NDIMDX]; z.ndm = p[NDIMDX_MTR];
    z.incj = z.ni; z.inck = z.ni * z.nj;
    return z;
}
static inline long IDX(const Ix* z, long i, long j, long k)
{ return (i + z->nf - 1) + (j + z->nf - 1) * z->ni + (k + z->nfk - 1) * z->ni * z->nj; }
static inline long MTR(const Ix* z, long i, long j, long k)
{ return (i + z->mf - 1) * z->mi + (j + z->mf - 1) * z->mj + (k + z->mfk - 1) * z->mk; }
static inline long XYZ(const Ix* z, long i, long j, long k)
{ return (i + z->xf - 1) + (j + z->xf - 1) * z->xi + (k + z->xfk - 1) * z->xi * z->xj; }

static inline double fsign(double a, double b) { return copysign(fabs(a), b); }
static inline double fmax2(double a, double b) { return a > b ? a : b; }
static inline double fmin2(double a, double b) { return a < b ? a : b; }

/* ------------------------------------------------------------------------------------ */
/* Metric: cp_tijk.for (3dfull :76-124 and the imax/jmax/kmax faces), cp_vol.for          */
int fo_init_metric(int* p, const double* x, const double* y, const double* z, double* ti,
                   double* tj, double* tk, double* vol)
{
    Ix zz = mkix(p); Ix* Z = &zz;
    int typ = p[ITYPZONE];
    long ind[6] = {1 - p[NIJK + 3], p[IJKV] + p[NIJK + 3], 1 - p[NIJK + 3],
                   p[IJKV + 1] + p[NIJK + 3], 1 - p[NIJK + 4], p[IJKV + 2] + p[NIJK + 4]};
    long inci = 1, incj = Z->xi, inck = Z->xi * Z->xj;
    long ndm = Z->ndm;
    if (typ == 2) {
        /* 3dcart: one scalar per direction, computed at cell (1,1,1) (init_metric.cpp:312-316) */
        long l0 = XYZ(Z, 1, 1, 1);
        long l1 = l0 + incj, l2 = l0 + inck, l3 = l1 + inck, l4 = l0 + inci;
        long l5 = l1 + inci, l6 = l2 + inci;
        double tix10 = (y[l1] - y[l0]) * (z[l3] - z[l0]) - (z[l1] - z[l0]) * (y[l3] - y[l0]);
        double tix11 = (y[l2] - y[l3]) * (z[l0] - z[l2]) - (z[l2] - z[l3]) * (y[l0] - y[l2]);
        double tjy10 = (z[l0] - z[l4]) * (x[l2] - x[l0]) - (x[l0] - x[l4]) * (z[l2] - z[l0]);
        double tjy11 = (z[l6] - z[l2]) * (x[l4] - x[l6]) - (x[l6] - x[l2]) * (z[l4] - z[l6]);
        double tkz10 = (x[l4] - x[l0]) * (y[l5] - y[l4]) - (y[l4] - y[l0]) * (x[l5] - x[l4]);
        double tkz11 = (x[l1] - x[l5]) * (y[l0] - y[l1]) - (y[l1] - y[l5]) * (x[l0] - x[l1]);
        ti[0] = .5 * (tix10 + tix11);
        tj[0] = .5 * (tjy10 + tjy11);
        tk[0] = .5 * (tkz10 + tkz11);
        {
            long l7 = l3 + inci;
            double ax = .25 * ((x[l4] + x[l5] + x[l7] + x[l6]) * ti[0] - (x[l0] + x[l1] + x[l3] + x[l2]) * ti[0]);
            double ay = .25 * ((y[l1] + y[l3] + y[l5] + y[l7]) * tj[0] - (y[l0] + y[l2] + y[l4] + y[l6]) * tj[0]);
            double az = .25 * ((z[l2] + z[l3] + z[l6] + z[l7]) * tk[0] - (z[l0] + z[l1] + z[l4] + z[l5]) * tk[0]);
            vol[0] = fmax2((ax + ay + az) / 3., 1.e-30);
        }
        return 0;
    }
    if (typ != 0) return -1; /* 3dhomo and 2d zones are outside this restatement */
    double* tix = ti; double* tiy = ti + ndm; double* tiz = ti + 2 * ndm;
    double* tjx = tj; double* tjy = tj + ndm; double* tjz = tj + 2 * ndm;
    double* tkx = tk; double* tky = tk + ndm; double* tkz = tk + 2 * ndm;
    long iv = p[IJKV], jv = p[IJKV + 1], kv = p[IJKV + 2];
    int li = ind[1] >= iv + 1, lj = ind[3] >= jv + 1, lk = ind[5] >= kv + 1;
    for (long k = ind[4]; k <= ind[5]; k++)
        for (long j = ind[2]; j <= ind[3]; j++)
            for (long i = ind[0]; i <= ind[1]; i++) {
                long lt = MTR(Z, i, j, k), l0 = XYZ(Z, i, j, k);
                long l1 = l0 + incj, l2 = l0 + inck, l3 = l1 + inck, l4 = l0 + inci;
                long l5 = l1 + inci, l6 = l2 + inci;
                double a10, a11, b10, b11, c10, c11;
                a10 = (y[l1] - y[l0]) * (z[l3] - z[l0]) - (z[l1] - z[l0]) * (y[l3] - y[l0]);
                b10 = (z[l1] - z[l0]) * (x[l3] - x[l0]) - (x[l1] - x[l0]) * (z[l3] - z[l0]);
                c10 = (x[l1] - x[l0]) * (y[l3] - y[l0]) - (y[l1] - y[l0]) * (x[l3] - x[l0]);
                a11 = (y[l2] - y[l3]) * (z[l0] - z[l2]) - (z[l2] - z[l3]) * (y[l0] - y[l2]);
                b11 = (z[l2] - z[l3]) * (x[l0] - x[l2]) - (x[l2] - x[l3]) * (z[l0] - z[l2]);
                c11 = (x[l2] - x[l3]) * (y[l0] - y[l2]) - (y[l2] - y[l3]) * (x[l0] - x[l2]);
                tix[lt] = .5 * (a10 + a11); tiy[lt] = .5 * (b10 + b11); tiz[lt] = .5 * (c10 + c11);
                a10 = (y[l0] - y[l4]) * (z[l2] - z[l0]) - (z[l0] - z[l4]) * (y[l2] - y[l0]);
                b10 = (z[l0] - z[l4]) * (x[l2] - x[l0]) - (x[l0] - x[l4]) * (z[l2] - z[l0]);
                c10 = (x[l0] - x[l4]) * (y[l2] - y[l0]) - (y[l0] - y[l4]) * (x[l2] - x[l0]);
                a11 = (y[l6] - y[l2]) * (z[l4] - z[l6]) - (z[l6] - z[l2]) * (y[l4] - y[l6]);
                b11 = (z[l6] - z[l2]) * (x[l4] - x[l6]) - (x[l6] - x[l2]) * (z[l4] - z[l6]);
                c11 = (x[l6] - x[l2]) * (y[l4] - y[l6]) - (y[l6] - y[l2]) * (x[l4] - x[l6]);
                tjx[lt] = .5 * (a10 + a11); tjy[lt] = .5 * (b10 + b11); tjz[lt] = .5 * (c10 + c11);
                a10 = (y[l4] - y[l0]) * (z[l5] - z[l4]) - (z[l4] - z[l0]) * (y[l5] - y[l4]);
                b10 = (z[l4] - z[l0]) * (x[l5] - x[l4]) - (x[l4] - x[l0]) * (z[l5] - z[l4]);
                c10 = (x[l4] - x[l0]) * (y[l5] - y[l4]) - (y[l4] - y[l0]) * (x[l5] - x[l4]);
                a11 = (y[l1] - y[l5]) * (z[l0] - z[l1]) - (z[l1] - z[l5]) * (y[l0] - y[l1]);
                b11 = (z[l1] - z[l5]) * (x[l0] - x[l1]) - (x[l1] - x[l5]) * (z[l0] - z[l1]);
                c11 = (x[l1] - x[l5]) * (y[l0] - y[l1]) - (y[l1] - y[l5]) * (x[l0] - x[l1]);
                tkx[lt] = .5 * (a10 + a11); tky[lt] = .5 * (b10 + b11); tkz[lt] = .5 * (c10 + c11);
            }
    if (li) {
        long i = ind[1] + 1;
        for (long k = ind[4]; k <= ind[5]; k++)
            for (long j = ind[2]; j <= ind[3]; j++) {
                long lt = MTR(Z, i, j, k), l0 = XYZ(Z, i, j, k);
                long l1 = l0 + incj, l2 = l0 + inck, l3 = l1 + inck;
                double a10 = (y[l1] - y[l0]) * (z[l3] - z[l0]) - (z[l1] - z[l0]) * (y[l3] - y[l0]);
                double b10 = (z[l1] - z[l0]) * (x[l3] - x[l0]) - (x[l1] - x[l0]) * (z[l3] - z[l0]);
                double c10 = (x[l1] - x[l0]) * (y[l3] - y[l0]) - (y[l1] - y[l0]) * (x[l3] - x[l0]);
                double a11 = (y[l2] - y[l3]) * (z[l0] - z[l2]) - (z[l2] - z[l3]) * (y[l0] - y[l2]);
                double b11 = (z[l2] - z[l3]) * (x[l0] - x[l2]) - (x[l2] - x[l3]) * (z[l0] - z[l2]);
                double c11 = (x[l2] - x[l3]) * (y[l0] - y[l2]) - (y[l2] - y[l3]) * (x[l0] - x[l2]);
                tix[lt] = .5 * (a10 + a11); tiy[lt] = .5 * (b10 + b11); tiz[lt] = .5 * (c10 + c11);
            }
    }
    if (lj) {
        long j = ind[3] + 1;
        for (long k = ind[4]; k <= ind[5]; k++)
            for (long i = ind[0]; i <= ind[1]; i++) {
                long lt = MTR(Z, i, j, k), l0 = XYZ(Z, i, j, k);
                long l2 = l0 + inck, l4 = l0 + inci, l6 = l2 + inci;
                double a10 = (y[l0] - y[l4]) * (z[l2] - z[l0]) - (z[l0] - z[l4]) * (y[l2] - y[l0]);
                double b10 = (z[l0] - z[l4]) * (x[l2] - x[l0]) - (x[l0] - x[l4]) * (z[l2] - z[l0]);
                double c10 = (x[l0] - x[l4]) * (y[l2] - y[l0]) - (y[l0] - y[l4]) * (x[l2] - x[l0]);
                double a11 = (y[l6] - y[l2]) * (z[l4] - z[l6]) - (z[l6] - z[l2]) * (y[l4] - y[l6]);
                double b11 = (z[l6] - z[l2]) * (x[l4] - x[l6]) - (x[l6] - x[l2]) * (z[l4] - z[l6]);
                double c11 = (x[l6] - x[l2]) * (y[l4] - y[l6]) - (y[l6] - y[l2]) * (x[l4] - x[l6]);
                tjx[lt] = .5 * (a10 + a11); tjy[lt] = .5 * (b10 + b11); tjz[lt] = .5 * (c10 + c11);
            }
    }
    if (lk) {
        long k = ind[5] + 1;
        for (long j = ind[2]; j <= ind[3]; j++)
            for (long i = ind[0]; i <= ind[1]; i++) {
                long lt = MTR(Z, i, j, k), l0 = XYZ(Z, i, j, k);
                long l1 = l0 + incj, l4 = l0 + inci, l5 = l1 + inci;
                double a10 = (y[l4] - y[l0]) * (z[l5] - z[l4]) - (z[l4] - z[l0]) * (y[l5] - y[l4]);
                double b10 = (z[l4] - z[l0]) * (x[l5] - x[l4]) - (x[l4] - x[l0]) * (z[l5] - z[l4]);
                double c10 = (x[l4] - x[l0]) * (y[l5] - y[l4]) - (y[l4] - y[l0]) * (x[l5] - x[l4]);
                double a11 = (y[l1] - y[l5]) * (z[l0] - z[l1]) - (z[l1] - z[l5]) * (y[l0] - y[l1]);
                double b11 = (z[l1] - z[l5]) * (x[l0] - x[l1]) - (x[l1] - x[l5]) * (z[l0] - z[l1]);
                double c11 = (x[l1] - x[l5]) * (y[l0] - y[l1]) - (y[l1] - y[l5]) * (x[l0] - x[l1]);
                tkx[lt] = .5 * (a10 + a11); tky[lt] = .5 * (b10 + b11); tkz[lt] = .5 * (c10 + c11);
            }
    }
    /* cp_vol.for 3dfull */
    for (long k = ind[4]; k <= ind[5]; k++)
        for (long j = ind[2]; j <= ind[3]; j++)
            for (long i = ind[0]; i <= ind[1]; i++) {
                long lt = MTR(Z, i, j, k), l0 = XYZ(Z, i, j, k);
                long l1 = l0 + incj, l2 = l0 + inck, l3 = l1 + inck, l4 = l0 + inci;
                long l5 = l1 + inci, l6 = l2 + inci, l7 = l3 + inci;
                long lip = lt + Z->mi, ljp = lt + Z->mj, lkp = lt + Z->mk;
                double ax = .25 * ((x[l4] + x[l5] + x[l7] + x[l6]) * tix[lip]
                                  - (x[l0] + x[l1] + x[l3] + x[l2]) * tix[lt]
                                  + (x[l1] + x[l3] + x[l5] + x[l7]) * tjx[ljp]
                                  - (x[l0] + x[l2] + x[l4] + x[l6]) * tjx[lt]
                                  + (x[l2] + x[l3] + x[l6] + x[l7]) * tkx[lkp]
                                  - (x[l0] + x[l1] + x[l4] + x[l5]) * tkx[lt]);
                double ay = .25 * ((y[l4] + y[l5] + y[l7] + y[l6]) * tiy[lip]
                                  - (y[l0] + y[l1] + y[l3] + y[l2]) * tiy[lt]
                                  + (y[l1] + y[l3] + y[l5] + y[l7]) * tjy[ljp]
                                  - (y[l0] + y[l2] + y[l4] + y[l6]) * tjy[lt]
                                  + (y[l2] + y[l3] + y[l6] + y[l7]) * tky[lkp]
                                  - (y[l0] + y[l1] + y[l4] + y[l5]) * tky[lt]);
                double az = .25 * ((z[l4] + z[l5] + z[l7] + z[l6]) * tiz[lip]
                                  - (z[l0] + z[l1] + z[l3] + z[l2]) * tiz[lt]
                                  + (z[l1] + z[l3] + z[l5] + z[l7]) * tjz[ljp]
                                  - (z[l0] + z[l2] + z[l4] + z[l6]) * tjz[lt]
                                  + (z[l2] + z[l3] + z[l6] + z[l7]) * tkz[lkp]
                                  - (z[l0] + z[l1] + z[l4] + z[l5]) * tkz[lt]);
                vol[lt] = fmax2((ax + ay + az) / 3., 1.e-30);
            }
    /* tijk_extrap.for only rewrites metrics of degenerate (collapsed) ghost cells; the
       zone classifier rejects meshes with degenerate cells, so nothing to do here. */
    return 0;
}

/* ------------------------------------------------------------------------------------ */
/* invist.for: Sutherland viscosity over ind_loop                                        */
void fo_invist(const int* p, const double* pr, const int* ind, const double* rop, double* xmut)
{
    Ix zz = mkix(p); Ix* Z = &zz;
    double cmus1 = pr[CS], temp01 = 1. / pr[TEMP0];
    double coesut = pr[XMUL0] * (1. + cmus1 * temp01) * sqrt(temp01);
    const double* T = rop + 4 * Z->nd;
    for (long k = ind[4]; k <= ind[5]; k++)
        for (long j = ind[2]; j <= ind[3]; j++)
            for (long i = ind[0]; i <= ind[1]; i++) {
                long l = IDX(Z, i, j, k);
                double t1 = T[l];
                xmut[l] = coesut * sqrt(t1) * t1 / (t1 + cmus1);
            }
}

/* tstb3_global.for (non-chimera): coe1 = DTC/vol */
static void fo_tstb3_global(const int* p, const double* pr, const int* ind, double* coe,
                            const double* vol)
{
    Ix zz = mkix(p); Ix* Z = &zz;
    if (p[ITYPZONE] != 2) {
        for (long k = ind[4]; k <= ind[5]; k++)
            for (long j = ind[2]; j <= ind[3]; j++)
                for (long i = ind[0]; i <= ind[1]; i++)
                    coe[IDX(Z, i, j, k)] = pr[DTC] / vol[MTR(Z, i, j, k)];
    } else {
        double volinv = pr[DTC] / vol[0];
        for (long k = ind[4]; k <= ind[5]; k++)
            for (long j = ind[2]; j <= ind[3]; j++)
                for (long i = ind[0]; i <= ind[1]; i++) coe[IDX(Z, i, j, k)] = volinv;
    }
}

/* tstb3c.for (global dt, implicit LU coefficients) and tstb3_impli.for + dtloc.for (local dt).
   local = 0 -> tstb3c (xid = 1.5, or 1 at the first step; cut0x = 1e-20)
   local = 1 -> tstb3_impli (coe1 = local dt/vol; xid = 1; cut0x = 1e-30)                 */
static void fo_tstb3(const int* p, const double* pr, int first_it, int local, const int* ind,
                     const double* rop, double* coe, const double* xmut, const double* ti,
                     const double* tj, const double* tk, const double* vol)
{
    Ix zz = mkix(p); Ix* Z = &zz;
    long nd = Z->nd, ndm = Z->ndm;
    double cut0x = local ? 1e-30 : 1e-20;
    double gamma = pr[GAMMA];
    double rgp = pr[CVINF] * (gamma - 1.);
    double prandt = pr[PRANDT];
    double xid = local ? 1. : 1.5;
    if (!local && first_it == 0) xid = 1.;
    double gam1 = gamma / prandt;
    if (p[IFLOW] == 1) gam1 = local ? cut0x : 0.;
    double gam1_1 = 1. / gam1;
    double gam2 = gamma * rgp;
    double* c1 = coe; double* c2 = coe + nd; double* c3 = coe + 2 * nd; double* c4 = coe + 3 * nd;
    double* c5 = coe + 4 * nd;
    const double *r = rop, *u = rop + nd, *v = rop + 2 * nd, *w = rop + 3 * nd, *t = rop + 4 * nd;
    if (p[ITYPZONE] == 0) {
        for (long k = ind[4]; k <= ind[5]; k++)
            for (long j = ind[2]; j <= ind[3]; j++)
                for (long i = ind[0]; i <= ind[1]; i++) {
                    long l = IDX(Z, i, j, k), lt = MTR(Z, i, j, k);
                    long li = lt + Z->mi, lj = lt + Z->mj, lk = lt + Z->mk;
                    double rr = r[l];
                    double detj = fmax2(vol[lt], cut0x);
                    double uu = u[l], vv = v[l], ww = w[l];
                    double c = sqrt(gam2 * t[l]);
                    double vmu = gam1 * xmut[l];
                    double volu = detj * rr;
                    double xmvis = vmu / volu;
                    double tcx, tcy, tcz, tc2i, tc2j, tc2k, qn, vm;
                    tcx = .5 * (ti[lt] + ti[li]); tcy = .5 * (ti[lt + ndm] + ti[li + ndm]);
                    tcz = .5 * (ti[lt + 2 * ndm] + ti[li + 2 * ndm]);
                    tc2i = .5 * (sqrt(ti[lt] * ti[lt] + ti[lt + ndm] * ti[lt + ndm] + ti[lt + 2 * ndm] * ti[lt + 2 * ndm])
                               + sqrt(ti[li] * ti[li] + ti[li + ndm] * ti[li + ndm] + ti[li + 2 * ndm] * ti[li + 2 * ndm]));
                    qn = tcx * uu + tcy * vv + tcz * ww;
                    vm = fabs(qn) + c * tc2i;
                    c2[l] = vm * 2. + xmvis * tc2i * tc2i * 4.;
                    tcx = .5 * (tj[lt] + tj[lj]); tcy = .5 * (tj[lt + ndm] + tj[lj + ndm]);
                    tcz = .5 * (tj[lt + 2 * ndm] + tj[lj + 2 * ndm]);
                    tc2j = .5 * (sqrt(tj[lt] * tj[lt] + tj[lt + ndm] * tj[lt + ndm] + tj[lt + 2 * ndm] * tj[lt + 2 * ndm])
                               + sqrt(tj[lj] * tj[lj] + tj[lj + ndm] * tj[lj + ndm] + tj[lj + 2 * ndm] * tj[lj + 2 * ndm]));
                    qn = tcx * uu + tcy * vv + tcz * ww;
                    vm = fabs(qn) + c * tc2j;
                    c3[l] = vm * 2. + xmvis * tc2j * tc2j * 4.;
                    tcx = .5 * (tk[lt] + tk[lk]); tcy = .5 * (tk[lt + ndm] + tk[lk + ndm]);
                    tcz = .5 * (tk[lt + 2 * ndm] + tk[lk + 2 * ndm]);
                    tc2k = .5 * (sqrt(tk[lt] * tk[lt] + tk[lt + ndm] * tk[lt + ndm] + tk[lt + 2 * ndm] * tk[lt + 2 * ndm])
                               + sqrt(tk[lk] * tk[lk] + tk[lk + ndm] * tk[lk + ndm] + tk[lk + 2 * ndm] * tk[lk + 2 * ndm]));
                    qn = tcx * uu + tcy * vv + tcz * ww;
                    vm = fabs(qn) + c * tc2k;
                    c4[l] = vm * 2. + xmvis * tc2k * tc2k * 4.;
                    if (local) {
                        double d = 1. / (tc2i * tc2i + tc2j * tc2j + tc2k * tc2k);
                        double dtvis = 0.5 * detj * d * gam1_1 / xmut[l] * rr;
                        double sp = sqrt(uu * uu + vv * vv + ww * ww) + c;
                        double dtconv = sqrt(d) / sp;
                        c1[l] = pr[CFL] * fmin2(dtconv, dtvis);
                    }
                    c5[l] = xid + c1[l] * (c2[l] + c3[l] + c4[l]) * 0.5;
                }
    } else { /* 3dcart */
        double tcxi = ti[0], tcyj = tj[0], tczk = tk[0];
        double tc2i = sqrt(tcxi * tcxi), tc2j = sqrt(tcyj * tcyj), tc2k = sqrt(tczk * tczk);
        double detj = fmax2(vol[0], cut0x);
        for (long k = ind[4]; k <= ind[5]; k++)
            for (long j = ind[2]; j <= ind[3]; j++)
                for (long i = ind[0]; i <= ind[1]; i++) {
                    long l = IDX(Z, i, j, k);
                    double rr = r[l];
                    double uu = u[l], vv = v[l], ww = w[l];
                    double c = sqrt(gam2 * t[l]);
                    double vmu = gam1 * xmut[l];
                    double volu = detj * rr;
                    double xmvis = vmu / volu;
                    double vm;
                    vm = fabs(tcxi * uu) + c * tc2i;
                    c2[l] = vm * 2. + xmvis * tc2i * tc2i * 4.;
                    vm = fabs(tcyj * vv) + c * tc2j;
                    c3[l] = vm * 2. + xmvis * tc2j * tc2j * 4.;
                    vm = fabs(tczk * ww) + c * tc2k;
                    c4[l] = vm * 2. + xmvis * tc2k * tc2k * 4.;
                    if (local) {
                        double d = 1. / (tc2i * tc2i + tc2j * tc2j + tc2k * tc2k);
                        double dtvis = 0.5 * detj * d * gam1_1 / xmut[l] * rr;
                        double sp = sqrt(uu * uu + vv * vv + ww * ww) + c;
                        double dtconv = sqrt(d) / sp;
                        c1[l] = pr[CFL] * fmin2(dtconv, dtvis);
                    }
                    c5[l] = xid + c1[l] * (c2[l] + c3[l] + c4[l]) * 0.5;
                }
    }
}

/* init_rhs.for: drodm <- 0, or coefH*drodm for RK stages 2 and 3 */
static void fo_init_rhs(const int* p, int nitcfg, const int* ind, double* drodm)
{
    Ix zz = mkix(p); Ix* Z = &zz;
    long nd = Z->nd;
    int b = p[RK];
    double coefH = 0.;
    if (b == 3) { if (nitcfg == 2) coefH = -0.6830127018922193; else if (nitcfg == 3) coefH = -4. / 3.; }
    else if (b == 2) { if (nitcfg == 2) coefH = -1.; }
    int zero = (p[ITYPCP] <= 1 || nitcfg == 1);
    for (long k = ind[4]; k <= ind[5]; k++)
        for (long j = ind[2]; j <= ind[3]; j++)
            for (long i = ind[0]; i <= ind[1]; i++) {
                long l = IDX(Z, i, j, k);
                for (int e = 0; e < 5; e++)
                    drodm[l + e * nd] = zero ? 0. : drodm[l + e * nd] * coefH;
            }
}

/* SENSOR/wiggle2001.for + wiggle_{i,j,k}.for: wig = 1 where a primitive changes slope sign */
static void fo_wiggle_one(const Ix* Z, const double* rop, double* wig, long l, long inc, int comp)
{
    long nd = Z->nd;
    long nm = l - inc, nm2 = l - 2 * inc, np = l + inc;
    double f[5];
    for (int e = 0; e < 5; e++) {
        const double* q = rop + e * nd;
        double qm = (q[nm] - q[nm2]) * (q[l] - q[nm]);
        double qp = (q[l] - q[nm]) * (q[np] - q[l]);
        f[e] = fsign(1., qm) + fsign(1., qp);
    }
    double test = fmin2(fmin2(fmin2(fmin2(f[0], f[1]), f[2]), f[3]), f[4]);
    wig[l + comp * nd] = (test < -1e-12) ? 1. : 0.;
}
static void fo_wiggle(const int* p, const double* pr, const int* ind, const double* rop, double* wig)
{
    Ix zz = mkix(p); Ix* Z = &zz;
    (void)pr;
    int icorr = ind[1] == p[IJKV], jcorr = ind[3] == p[IJKV + 1], kcorr = ind[5] == p[IJKV + 2];
    for (long k = ind[4]; k <= ind[5]; k++)
        for (long j = ind[2]; j <= ind[3]; j++)
            for (long i = ind[0]; i <= ind[1]; i++) {
                long l = IDX(Z, i, j, k);
                fo_wiggle_one(Z, rop, wig, l, Z->inck, 2);
                fo_wiggle_one(Z, rop, wig, l, Z->incj, 1);
                fo_wiggle_one(Z, rop, wig, l, 1, 0);
            }
    if (icorr) {
        long i = ind[1] + 1;
        for (long k = ind[4]; k <= ind[5]; k++)
            for (long j = ind[2]; j <= ind[3]; j++) fo_wiggle_one(Z, rop, wig, IDX(Z, i, j, k), 1, 0);
    }
    if (jcorr) {
        long j = ind[3] + 1;
        for (long k = ind[4]; k <= ind[5]; k++)
            for (long i = ind[0]; i <= ind[1]; i++) fo_wiggle_one(Z, rop, wig, IDX(Z, i, j, k), Z->incj, 1);
    }
    if (kcorr) {
        long k = ind[5] + 1;
        for (long j = ind[2]; j <= ind[3]; j++)
            for (long i = ind[0]; i <= ind[1]; i++) fo_wiggle_one(Z, rop, wig, IDX(Z, i, j, k), Z->inck, 2);
    }
}

/* ------------------------------------------------------------------------------------ */
/* Face flux.  Generated kernels flu<scheme>_<model>_<slope>_<mesh> from
   template_FluxAndBalance.for with the fragments Slope/{o1,o3,minmod}_slope_var.for, etat_GD.for,
   etat_roe_GD.for, Vit_ent/{qn,fludiffer,fluvector}_3d{cart,full}_{i,j,k}.for,
   {ROE,AUSM,SENSOR}/fluFaceEuler.for, fluvisq_3d{cart,full}_{i,j,k}.for, mut_interface.for. */
typedef struct {
    const Ix* Z;
    const double *rop, *xmut, *wig, *ti, *tj, *tk, *vol;
    int kflu, slope, lamin, cart;
    /* scheme constants (flu template header) */
    double gam, rgp, gam1, gam3, psiroe, tmin_1, c1, c2, c3, opt0, c4, c5, c6, cvisq, wig_cte;
    /* 3dcart metric */
    double tc[3], s[3], volinv;
} FC;

static void slope_state(const FC* f, long l, long inc, double qm[5], double qp[5])
{
    long nd = f->Z->nd;
    long nm = l - inc, nm2 = l - 2 * inc, np = l + inc;
    const double* rop = f->rop;
    if (f->slope == 1) {
        for (int e = 0; e < 5; e++) { qm[e] = rop[l + e * nd]; qp[e] = rop[nm + e * nd]; }
    } else if (f->slope == 2) {
        for (int e = 0; e < 5; e++) {
            const double* q = rop + e * nd;
            qm[e] = f->c4 * q[l] + f->c5 * q[nm] + f->c6 * q[np];
            qp[e] = f->c4 * q[nm] + f->c6 * q[nm2] + f->c5 * q[l];
        }
    } else { /* minmod_slope_var.for */
        for (int e = 0; e < 5; e++) {
            const double* q = rop + e * nd;
            double roff = .5 * fmax2(0., fmin2(rop[l + 4 * nd] * f->tmin_1 - .5, 1.));
            double delm = q[l] - q[nm], delp = q[np] - q[l], delq = q[nm] - q[nm2];
            double sg, slp, slq;
            sg = fsign(1., delm); slp = roff * (sg * fmax2(0., fmin2(fabs(delm), sg * delp)));
            sg = fsign(1., delq); slq = roff * (sg * fmax2(0., fmin2(fabs(delq), sg * delm)));
            qm[e] = q[l] - slp;
            qp[e] = q[nm] + slq;
        }
    }
}

/* inviscid flux through face (l-inc, l) in direction d; tn = face vector; sn = |tn| */
static void inviscid_flux(const FC* f, long l, int d, const double tn[3], double sn, int cart,
                          const double qm[5], const double qp[5], double flu[5])
{
    long nd = f->Z->nd;
    if (f->kflu == 5) { /* Roe: etat_roe_GD.for + qn + fludiffer */
        double gam = f->gam, rgp = f->rgp, gam1 = f->gam1, psiroe = f->psiroe;
        double r1 = qp[0];
        double p1 = r1 * qp[4] * rgp;
        double h1 = gam1 * qp[4] * rgp + .5 * (qp[1] * qp[1] + qp[2] * qp[2] + qp[3] * qp[3]);
        double r2 = qm[0];
        double p2 = r2 * qm[4] * rgp;
        double h2 = gam1 * qm[4] * rgp + .5 * (qm[1] * qm[1] + qm[2] * qm[2] + qm[3] * qm[3]);
        double sp1 = sqrt(r1), sm1 = sqrt(r2);
        double r = sp1 * sm1;
        double r_1 = 1. / (sp1 + sm1);
        double a[3];
        a[0] = (sp1 * qp[1] + sm1 * qm[1]) * r_1;
        a[1] = (sp1 * qp[2] + sm1 * qm[2]) * r_1;
        a[2] = (sp1 * qp[3] + sm1 * qm[3]) * r_1;
        double h = (sp1 * h1 + sm1 * h2) * r_1;
        double q = .5 * (a[0] * a[0] + a[1] * a[1] + a[2] * a[2]);
        double c = sqrt((gam - 1.) * fabs(h - q));
        double qn1, qn2;
        if (cart) { qn1 = qp[1 + d] * tn[d]; qn2 = qm[1 + d] * tn[d]; }
        else {
            qn1 = qp[1] * tn[0] + qp[2] * tn[1] + qp[3] * tn[2];
            qn2 = qm[1] * tn[0] + qm[2] * tn[1] + qm[3] * tn[2];
        }
        double du[3] = {(qm[1] - qp[1]) * sn, (qm[2] - qp[2]) * sn, (qm[3] - qp[3]) * sn};
        double dp = p2 - p1;
        double dqn = qn2 - qn1;
        double s_1 = 1. / fmax2(sn, 1.e-30);
        double nn[3], qn;
        if (cart) { nn[d] = tn[d] * s_1; qn = a[d] * nn[d]; }
        else {
            nn[0] = tn[0] * s_1; nn[1] = tn[1] * s_1; nn[2] = tn[2] * s_1;
            qn = a[0] * nn[0] + a[1] * nn[1] + a[2] * nn[2];
        }
        double f1 = qn;
        double f4 = psiroe * (fabs(f1) + c);
        double f2 = .5 + fsign(.5, fabs(f1) - f4);
        double f3 = (f2 - 1.) * (f1 * f1 / f4 + f4) * 0.5 - f2 * fabs(f1);
        f1 = (r2 - r1 - dp / (c * c)) * sn;
        flu[0] = f3 * (f1);
        for (int m = 0; m < 3; m++) {
            int dir = cart ? (m == d) : 1;
            flu[1 + m] = f3 * (f1 * a[m] + r * (dir ? du[m] - nn[m] * dqn : du[m]));
        }
        flu[4] = f3 * (f1 * q + r * (a[0] * du[0] + a[1] * du[1] + a[2] * du[2] - qn * dqn));
        f1 = qn + c;
        f2 = .5 + fsign(.5, fabs(f1) - f4);
        f3 = (f2 - 1.) * (f1 * f1 / f4 + f4) * 0.5 - f2 * fabs(f1);
        f1 = (dp + r * c * dqn * s_1) / (2. * c * c) * sn * f3;
        flu[0] = flu[0] + f1;
        for (int m = 0; m < 3; m++) {
            int dir = cart ? (m == d) : 1;
            flu[1 + m] = flu[1 + m] + f1 * (dir ? a[m] + nn[m] * c : a[m]);
        }
        flu[4] = flu[4] + f1 * (h + qn * c);
        f1 = qn - c;
        f2 = .5 + fsign(.5, fabs(f1) - f4);
        f3 = (f2 - 1.) * (f1 * f1 / f4 + f4) * 0.5 - f2 * fabs(f1);
        f1 = (dp - r * c * dqn * s_1) / (2. * c * c) * f3 * sn;
        flu[0] = flu[0] + f1;
        for (int m = 0; m < 3; m++) {
            int dir = cart ? (m == d) : 1;
            flu[1 + m] = flu[1 + m] + f1 * (dir ? a[m] - nn[m] * c : a[m]);
        }
        flu[4] = flu[4] + f1 * (h - qn * c);
        double rou1 = r1 * qn1, rou2 = r2 * qn2, p1p2 = p1 + p2;
        flu[0] = (flu[0] + rou1 + rou2) * 0.5;
        for (int m = 0; m < 3; m++) {
            int dir = cart ? (m == d) : 1;
            flu[1 + m] = dir ? (flu[1 + m] + rou1 * qp[1 + m] + rou2 * qm[1 + m] + tn[m] * p1p2) * 0.5
                             : (flu[1 + m] + rou1 * qp[1 + m] + rou2 * qm[1 + m]) * 0.5;
        }
        flu[4] = (flu[4] + rou1 * h1 + rou2 * h2) * 0.5;
        return;
    }
    /* AUSM+(P) (kflu 1) and SENSEUR (kflu 2): etat_GD.for + fluvector */
    double rgp = f->rgp, gam1 = f->gam1;
    double r1 = qp[0];
    double rou1 = r1 * qp[1], rov1 = r1 * qp[2], row1 = r1 * qp[3];
    double p1 = r1 * qp[4] * rgp;
    double h1 = gam1 * p1 + .5 * (rou1 * qp[1] + rov1 * qp[2] + row1 * qp[3]);
    double r2 = qm[0];
    double rou2 = r2 * qm[1], rov2 = r2 * qm[2], row2 = r2 * qm[3];
    double p2 = r2 * qm[4] * rgp;
    double h2 = gam1 * p2 + .5 * (rou2 * qm[1] + rov2 * qm[2] + row2 * qm[3]);
    double qn1, qn2;
    if (cart) { qn1 = qp[1 + d] * tn[d]; qn2 = qm[1 + d] * tn[d]; }
    else {
        qn1 = qp[1] * tn[0] + qp[2] * tn[1] + qp[3] * tn[2];
        qn2 = qm[1] * tn[0] + qm[2] * tn[1] + qm[3] * tn[2];
    }
    double u, tdu;
    if (f->kflu == 1) {
        double c = rgp * f->gam * f->rop[l + 4 * nd];
        double son = sqrt(qn1 * qn1 / c);
        double tam = f->c3 * son + sn;
        double tam1 = fmax2(0., tam) * f->c2;
        u = 0.25 * (qn1 + qn2) - tam1 * (p2 - p1);
        tdu = fmax2(fabs(u), f->c1 * sn) * f->wig_cte;
    } else {
        double wg = f->wig[l + d * nd];
        u = 0.25 * (qn1 + qn2) - f->c2 * sn * (p2 - p1) * (f->opt0 * wg + 1. - f->opt0);
        tdu = fmax2(fabs(u), f->c1 * sn) * wg;
    }
    double p1p2 = (p1 + p2) * 0.5;
    double mom1[3] = {rou1, rov1, row1}, mom2[3] = {rou2, rov2, row2};
    flu[0] = u * (r1 + r2) - tdu * (r2 - r1);
    for (int m = 0; m < 3; m++) {
        int dir = cart ? (m == d) : 1;
        flu[1 + m] = dir ? u * (mom1[m] + mom2[m]) - tdu * (mom2[m] - mom1[m]) + tn[m] * p1p2
                         : u * (mom1[m] + mom2[m]) - tdu * (mom2[m] - mom1[m]);
    }
    flu[4] = u * (h1 + h2) - tdu * (h2 - h1);
}

/* centred viscous flux, 3dcart: fluvisq_3dcart_{i,j,k}.for */
static void visc_cart(const FC* f, long l, int d, double flu[5], double* gv)
{
    const Ix* Z = f->Z;
    long nd = Z->nd, inci = 1, incj = Z->incj, inck = Z->inck;
    const double* rop = f->rop;
    const double *U = rop + nd, *V = rop + 2 * nd, *W = rop + 3 * nd, *T = rop + 4 * nd;
    double tcx = f->tc[0], tcy = f->tc[1], tcz = f->tc[2];
    double xmutvol, xktvol, fv, fv5, f1, f2, f3, f4, f5, f6, div, fvs;
    if (d == 0) {
        long il = l - inci, ir = l;
        long l220 = l - inci - incj, l020 = l - incj, l210 = l - inci + incj, l010 = l + incj;
        long l202 = l - inci - inck, l002 = l - inck, l201 = l - inci + inck, l001 = l + inck;
        double tix = tcx, tjy = .25 * tcy, tkz = .25 * tcz;
        double gUx, gUy, gUz, gVx, gVy, gWx, gWz, gTx;
        f1 = U[il]; f2 = U[ir]; fvs = f1 + f2;
        f3 = fvs + U[l220] + U[l020]; f4 = fvs + U[l210] + U[l010];
        f5 = fvs + U[l202] + U[l002]; f6 = fvs + U[l201] + U[l001];
        gUx = (f2 - f1) * tix; gUy = (f4 - f3) * tjy; gUz = (f6 - f5) * tkz;
        f1 = V[il]; f2 = V[ir]; fvs = f1 + f2;
        f3 = fvs + V[l220] + V[l020]; f4 = fvs + V[l210] + V[l010];
        gVx = (f2 - f1) * tix; gVy = (f4 - f3) * tjy;
        f1 = W[il]; f2 = W[ir]; fvs = f1 + f2;
        f5 = fvs + W[l202] + W[l002]; f6 = fvs + W[l201] + W[l001];
        gWx = (f2 - f1) * tix; gWz = (f6 - f5) * tkz;
        f1 = T[il]; f2 = T[ir];
        gTx = (f2 - f1) * tix;
        div = (gUx + gVy + gWz) * f->cvisq;
        f1 = gUx - div; f2 = gUy + gVx; f3 = gUz + gWx;
        xmutvol = (f->xmut[ir] + f->xmut[il]) * f->volinv;
        xktvol = xmutvol * f->gam3;
        fv = -(2. * f1 * tcx) * xmutvol; fv5 = fv * (U[ir] + U[il]); flu[1] = flu[1] + fv;
        fv = -(f2 * tcx) * xmutvol; fv5 = fv5 + fv * (V[ir] + V[il]); flu[2] = flu[2] + fv;
        fv = -(f3 * tcx) * xmutvol; fv5 = fv5 + fv * (W[ir] + W[il]); flu[3] = flu[3] + fv;
        fv = fv5 * 0.5 - (gTx * tcx) * xktvol; flu[4] = flu[4] + fv;
        if (gv) {
            gv[0] = gUx; gv[1] = gUy; gv[2] = gUz; gv[3] = gVx; gv[4] = gVy;
            gv[6] = gWx; gv[8] = gWz; gv[9] = gTx;
        }
    } else if (d == 1) {
        long il = l - incj, ir = l;
        long l022 = l - incj - inck, l002 = l - inck, l021 = l - incj + inck, l001 = l + inck;
        long l220 = l - inci - incj, l200 = l - inci, l120 = l + inci - incj, l100 = l + inci;
        double tiy = tcy, tjz = .25 * tcz, tkx = .25 * tcx;
        double gUx, gUy, gVx, gVy, gVz, gWy, gWz, gTy;
        f1 = U[il]; f2 = U[ir]; fvs = f1 + f2;
        f5 = fvs + U[l220] + U[l200]; f6 = fvs + U[l120] + U[l100];
        gUx = (f6 - f5) * tkx; gUy = (f2 - f1) * tiy;
        f1 = V[il]; f2 = V[ir]; fvs = f1 + f2;
        f3 = fvs + V[l022] + V[l002]; f4 = fvs + V[l021] + V[l001];
        f5 = fvs + V[l220] + V[l200]; f6 = fvs + V[l120] + V[l100];
        gVx = (f6 - f5) * tkx; gVy = (f2 - f1) * tiy; gVz = (f4 - f3) * tjz;
        f1 = W[il]; f2 = W[ir]; fvs = f1 + f2;
        f3 = fvs + W[l022] + W[l002]; f4 = fvs + W[l021] + W[l001];
        gWy = (f2 - f1) * tiy; gWz = (f4 - f3) * tjz;
        f1 = T[il]; f2 = T[ir];
        gTy = (f2 - f1) * tiy;
        div = (gUx + gVy + gWz) * f->cvisq;
        f2 = gUy + gVx; f4 = gVy - div; f5 = gVz + gWy;
        xmutvol = (f->xmut[ir] + f->xmut[il]) * f->volinv;
        xktvol = xmutvol * f->gam3;
        fv = -(f2 * tcy) * xmutvol; fv5 = fv * (U[ir] + U[il]); flu[1] = flu[1] + fv;
        fv = -(2. * f4 * tcy) * xmutvol; fv5 = fv5 + fv * (V[ir] + V[il]); flu[2] = flu[2] + fv;
        fv = -(f5 * tcy) * xmutvol; fv5 = fv5 + fv * (W[ir] + W[il]); flu[3] = flu[3] + fv;
        fv = fv5 * 0.5 - (gTy * tcy) * xktvol; flu[4] = flu[4] + fv;
        if (gv) {
            gv[0] = gUx; gv[1] = gUy; gv[3] = gVx; gv[4] = gVy; gv[5] = gVz;
            gv[7] = gWy; gv[8] = gWz; gv[10] = gTy;
        }
    } else {
        long il = l - inck, ir = l;
        long l200 = l - inci, l202 = l - inci - inck, l102 = l + inci - inck, l100 = l + inci;
        long l022 = l - incj - inck, l020 = l - incj, l012 = l + incj - inck, l010 = l + incj;
        double tiz = tcz, tky = .25 * tcy, tjx = .25 * tcx;
        double gUx, gUz, gVy, gVz, gWx, gWy, gWz, gTz;
        f1 = U[il]; f2 = U[ir]; fvs = f1 + f2;
        f3 = fvs + U[l202] + U[l200]; f4 = fvs + U[l102] + U[l100];
        gUx = (f4 - f3) * tjx; gUz = (f2 - f1) * tiz;
        f1 = V[il]; f2 = V[ir]; fvs = f1 + f2;
        f5 = fvs + V[l022] + V[l020]; f6 = fvs + V[l012] + V[l010];
        gVy = (f6 - f5) * tky; gVz = (f2 - f1) * tiz;
        f1 = W[il]; f2 = W[l]; fvs = f1 + f2;
        f3 = fvs + W[l202] + W[l200]; f4 = fvs + W[l102] + W[l100];
        f5 = fvs + W[l022] + W[l020]; f6 = fvs + W[l012] + W[l010];
        gWx = (f4 - f3) * tjx; gWy = (f6 - f5) * tky; gWz = (f2 - f1) * tiz;
        f1 = T[il]; f2 = T[ir];
        gTz = (f2 - f1) * tiz;
        div = (gUx + gVy + gWz) * f->cvisq;
        f3 = gUz + gWx; f5 = gVz + gWy; f6 = gWz - div;
        xmutvol = (f->xmut[ir] + f->xmut[il]) * f->volinv;
        xktvol = xmutvol * f->gam3;
        fv = -f3 * tcz * xmutvol; fv5 = fv * (U[ir] + U[il]); flu[1] = flu[1] + fv;
        fv = -f5 * tcz * xmutvol; fv5 = fv5 + fv * (V[ir] + V[il]); flu[2] = flu[2] + fv;
        fv = -2. * f6 * tcz * xmutvol; fv5 = fv5 + fv * (W[ir] + W[il]); flu[3] = flu[3] + fv;
        fv = fv5 * 0.5 - gTz * tcz * xktvol; flu[4] = flu[4] + fv;
        if (gv) {
            gv[0] = gUx; gv[2] = gUz; gv[4] = gVy; gv[5] = gVz;
            gv[6] = gWx; gv[7] = gWy; gv[8] = gWz; gv[11] = gTz;
        }
    }
    if (gv) gv[12] = f->volinv;
}

/* centred viscous flux, 3dfull: fluvisq_3dfull_{i,j,k}.for */
static void visc_full(const FC* f, long l, long lt, int d, const double tn[3], double flu[5], double* gv)
{
    const Ix* Z = f->Z;
    long nd = Z->nd, ndm = Z->ndm;
    long inci = 1, incj = Z->incj, inck = Z->inck, mi = Z->mi, mj = Z->mj, mk = Z->mk;
    const double* rop = f->rop;
    const double *A, *B, *C;   /* metric arrays for the "i", "j", "k" roles of the dual cell */
    long il, ir, lvol;
    long a0, a1, b0, b1, b2, b3, c0, c1, c2, c3;     /* metric points */
    long n3a, n3b, n4a, n4b, n5a, n5b, n6a, n6b;      /* cell neighbours for f3..f6 */
    if (d == 0) {
        A = f->ti; B = f->tj; C = f->tk;
        il = l - inci; ir = l; lvol = lt - mi;
        a0 = lt - mi; a1 = lt + mi;
        b0 = lt; b1 = lt - mi; b2 = lt + mj; b3 = lt - mi + mj;
        c0 = lt; c1 = lt - mi; c2 = lt + mk; c3 = lt - mi + mk;
        n3a = l - inci - incj; n3b = l - incj; n4a = l - inci + incj; n4b = l + incj;
        n5a = l - inci - inck; n5b = l - inck; n6a = l - inci + inck; n6b = l + inck;
    } else if (d == 1) {
        A = f->tj; B = f->tk; C = f->ti;
        il = l - incj; ir = l; lvol = lt - mj;
        a0 = lt - mj; a1 = lt + mj;
        b0 = lt; b1 = lt - mj; b2 = lt + mk; b3 = lt - mj + mk;
        c0 = lt; c1 = lt - mj; c2 = lt + mi; c3 = lt + mi - mj;
        n3a = l - incj - inck; n3b = l - inck; n4a = l - incj + inck; n4b = l + inck;
        n5a = l - inci - incj; n5b = l - inci; n6a = l + inci - incj; n6b = l + inci;
    } else {
        A = f->tk; B = f->ti; C = f->tj;
        il = l - inck; ir = l; lvol = lt - mk;
        a0 = lt - mk; a1 = lt + mk;
        b0 = lt; b1 = lt - mk; b2 = lt + mi; b3 = lt + mi - mk;
        c0 = lt; c1 = lt - mk; c2 = lt + mj; c3 = lt + mj - mk;
        n3a = l - inci - inck; n3b = l - inci; n4a = l + inci - inck; n4b = l + inci;
        n5a = l - incj - inck; n5b = l - incj; n6a = l + incj - inck; n6b = l + incj;
    }
    double t0[3], t1[3], u0[3], u1[3], w0[3], w1[3];
    for (int m = 0; m < 3; m++) {
        t0[m] = .5 * (A[lt + m * ndm] + A[a0 + m * ndm]);
        t1[m] = .5 * (A[lt + m * ndm] + A[a1 + m * ndm]);
        u0[m] = .125 * (B[b0 + m * ndm] + B[b1 + m * ndm]);
        u1[m] = .125 * (B[b2 + m * ndm] + B[b3 + m * ndm]);
        w0[m] = .125 * (C[c0 + m * ndm] + C[c1 + m * ndm]);
        w1[m] = .125 * (C[c2 + m * ndm] + C[c3 + m * ndm]);
    }
    double g[4][3];
    for (int e = 0; e < 4; e++) {
        const double* q = rop + (e + 1) * nd;
        double f1 = q[il], f2 = q[ir];
        double f3 = q[ir] + q[il] + q[n3a] + q[n3b];
        double f4 = q[ir] + q[il] + q[n4a] + q[n4b];
        double f5 = q[ir] + q[il] + q[n5a] + q[n5b];
        double f6 = q[ir] + q[il] + q[n6a] + q[n6b];
        for (int m = 0; m < 3; m++)
            g[e][m] = (f2 * t1[m] - f1 * t0[m]) + (f4 * u1[m] - f3 * u0[m]) + (f6 * w1[m] - f5 * w0[m]);
    }
    double div = (g[0][0] + g[1][1] + g[2][2]) * f->cvisq;
    double f1 = g[0][0] - div, f2 = g[0][1] + g[1][0], f3 = g[0][2] + g[2][0];
    double f4 = g[1][1] - div, f5 = g[1][2] + g[2][1], f6 = g[2][2] - div;
    double volinv = 1. / (f->vol[lt] + f->vol[lvol]);
    double xmutvol = (f->xmut[ir] + f->xmut[il]) * volinv;
    double xktvol = xmutvol * f->gam3;
    const double *U = rop + nd, *V = rop + 2 * nd, *W = rop + 3 * nd;
    double tcx = tn[0], tcy = tn[1], tcz = tn[2], fv, fv5;
    fv = -(2. * f1 * tcx + f2 * tcy + f3 * tcz) * xmutvol;
    fv5 = fv * (U[ir] + U[il]); flu[1] = flu[1] + fv;
    fv = -(f2 * tcx + 2. * f4 * tcy + f5 * tcz) * xmutvol;
    fv5 = fv5 + fv * (V[ir] + V[il]); flu[2] = flu[2] + fv;
    fv = -(f3 * tcx + f5 * tcy + 2. * f6 * tcz) * xmutvol;
    fv5 = fv5 + fv * (W[ir] + W[il]); flu[3] = flu[3] + fv;
    fv = fv5 * 0.5 - (g[3][0] * tcx + g[3][1] * tcy + g[3][2] * tcz) * xktvol;
    flu[4] = flu[4] + fv;
    if (gv) {
        for (int e = 0; e < 4; e++)
            for (int m = 0; m < 3; m++) gv[3 * e + m] = g[e][m];
        gv[12] = volinv;
    }
}

static void face_flux(const FC* f, long i, long j, long k, int d, double flu[5])
{
    const Ix* Z = f->Z;
    long l = IDX(Z, i, j, k);
    long inc = d == 0 ? 1 : d == 1 ? Z->incj : Z->inck;
    double qm[5], qp[5], tn[3], sn;
    long lt = 0;
    if (f->cart) {
        tn[0] = f->tc[0]; tn[1] = f->tc[1]; tn[2] = f->tc[2]; sn = f->s[d];
    } else {
        const double* t = d == 0 ? f->ti : d == 1 ? f->tj : f->tk;
        lt = MTR(Z, i, j, k);
        long ndm = Z->ndm;
        tn[0] = t[lt]; tn[1] = t[lt + ndm]; tn[2] = t[lt + 2 * ndm];
        sn = sqrt(tn[0] * tn[0] + tn[1] * tn[1] + tn[2] * tn[2]);
    }
    slope_state(f, l, inc, qm, qp);
    inviscid_flux(f, l, d, tn, sn, f->cart, qm, qp, flu);
    if (f->lamin) {
        if (f->cart) visc_cart(f, l, d, flu, NULL);
        else visc_full(f, l, lt, d, tn, flu, NULL);
    }
}

static int cl_window(const int* p, int idir, const int* fen, int ind_cl[6]);
static int cl_window_sdm(const int* p, int idir, const int* fen, const int sdm[6], int ind_cl[6]);
static void shape_tab(int neq_mtr, int idir, int* ic, int* jc, int* kc, double* cim, double* cjm,
                      double* ckm);

/* correct_flux.for:44-260 for Roe (lapply = 4): at every wall face (BCWallInviscid /
   BCSymmetryPlane, BCWall / BCWallViscous and variants) of the zone, in BC order,
   1. the Euler face flux that the balance added is taken back out (template_correction.for,
      corr_fluroe_*, assemble_drodm_corrflu.for: df = flu*sens + flux(li)*c28, c28 = 0);
   2. the wall flux is put in: bflwall.for (pressure and the convective term of the face
      averaged state, no mass or energy flux) or bflwallslip.for (slip wall: convective fluxes
      with the normal velocity of the averaged state), assemble_drodm_corr.for.
   Both act on the interior cell next to the face; the window is indice_cl_sdm.for with
   npass = 0, the intersection of the BC window with the cache block sdm (navier_stokes_struct.for
   calls correct_flux once per cache block, after that block's fluxes), and its normal range
   set to the face index (correct_flux.for:116-133).                                       */
static void correct_flux_roe(const int* p, const double* pr, const FC* f, const int* sdm,
                             double* drodm)
{
    const Ix* Z = f->Z;
    long nd = Z->nd, ndm = Z->ndm;
    int pt_bcs = p[PT_BC], nb_bc = pt_bcs > 0 ? p[pt_bcs] : 0;
    double rgp = pr[CVINF] * (pr[GAMMA] - 1.);
    double gam2 = pr[GAMMA] / (pr[GAMMA] - 1.);
    double ck_vent = p[NEQ_VENT] != 3 ? 0. : 1.;
    double c_ale = 1. * (p[LALE] < 1 ? p[LALE] : 1);
    FC g = *f;
    g.lamin = 0;      /* iflow = 1: only the Euler flux is corrected */
    for (int n = 0; n < nb_bc; n++) {
        int pt = p[pt_bcs + 1 + n], idir = p[pt + BC_IDIR], t = p[pt + BC_TYPE];
        int slip = t == 3, wall = t == 4 || t == 6 || t == 7 || t == 12;
        if (!slip && !wall) continue;
        int cl[6];
        if (cl_window_sdm(p, idir, p + pt + BC_FEN, sdm, cl)) continue;
        switch (idir) {
        case 1: cl[0] = 1; cl[1] = 1; break;
        case 2: cl[1] = cl[0]; break;
        case 3: cl[2] = 1; cl[3] = 1; break;
        case 4: cl[3] = cl[2]; break;
        case 5: cl[4] = 1; cl[5] = 1; break;
        default: cl[5] = cl[4]; break;
        }
        int d = (idir - 1) / 2;
        long inc = d == 0 ? 1 : d == 1 ? Z->incj : Z->inck;
        int even = idir % 2 == 0;
        long shift = even ? inc : 0;
        double flu[5];
        /* 1. take the Euler face flux back out */
        double sens = even ? 1. : -1.;
        for (long k = cl[4]; k <= cl[5]; k++)
            for (long j = cl[2]; j <= cl[3]; j++)
                for (long i = cl[0]; i <= cl[1]; i++) {
                    long l = IDX(Z, i, j, k), l0 = l - shift;
                    face_flux(&g, i, j, k, d, flu);
                    for (int e = 0; e < 5; e++) {
                        double df = flu[e] * sens + 0.;
                        drodm[l0 + e * nd] = drodm[l0 + e * nd] + df;
                    }
                }
        /* 2. add the wall flux */
        const double* tijk = d == 0 ? f->ti : d == 1 ? f->tj : f->tk;
        int neq_mtr = idir <= 4 ? p[NEQ_IJ] : p[NEQ_K];
        int ic, jc, kc;
        double cim, cjm, ckm;
        shape_tab(neq_mtr, idir, &ic, &jc, &kc, &cim, &cjm, &ckm);
        double sw = even ? -1. : 1.;
        const double* rop = f->rop;
        for (long k = cl[4]; k <= cl[5]; k++)
            for (long j = cl[2]; j <= cl[3]; j++)
                for (long i = cl[0]; i <= cl[1]; i++) {
                    long l = IDX(Z, i, j, k), iadrf = l - inc, l0 = l - shift;
                    long lt = MTR(Z, i, j, k);
                    double tcx = tijk[lt + ndm * (ic - 1)] * cim;
                    double tcy = tijk[lt + ndm * (jc - 1)] * cjm;
                    double tcz = tijk[lt + ndm * (kc - 1)] * ckm;
                    double qen = (0. * tcx + 0. * tcy + 0. * tcz * ck_vent) * c_ale;
                    double r = 0.5 * (rop[l] + rop[iadrf]);
                    double u = 0.5 * (rop[l + nd] + rop[iadrf + nd]);
                    double v = 0.5 * (rop[l + 2 * nd] + rop[iadrf + 2 * nd]);
                    double w = 0.5 * (rop[l + 3 * nd] + rop[iadrf + 3 * nd]);
                    double u_int = tcx * u + tcy * v + tcz * w - qen;
                    if (wall) {
                        double pw = 0.5 * (rop[l + 4 * nd] * rop[l] + rop[iadrf + 4 * nd] * rop[iadrf]) * rgp;
                        flu[0] = 0.;
                        flu[1] = tcx * pw + u_int * u * r;
                        flu[2] = tcy * pw + u_int * v * r;
                        flu[3] = tcz * pw + u_int * w * r;
                        flu[4] = pw * qen;
                    } else {
                        double pi = rop[iadrf + 4 * nd] * rop[iadrf] * rgp;
                        double h = gam2 * pi + r * (u * u + v * v + w * w) * 0.5;
                        double pw = 0.5 * (rop[l + 4 * nd] * rop[l] + rop[iadrf + 4 * nd] * rop[iadrf]) * rgp;
                        flu[0] = u_int * r;
                        flu[1] = tcx * pw + u_int * r * u;
                        flu[2] = tcy * pw + u_int * r * v;
                        flu[3] = tcz * pw + u_int * r * w;
                        flu[4] = u_int * h + pw * qen;
                    }
                    for (int e = 0; e < 5; e++) drodm[l0 + e * nd] = drodm[l0 + e * nd] + flu[e] * sw;
                }
    }
}

/* template_FluxAndBalance.for:192-271 with assemble_drodm_plus_vec1.for /
   flu_send_scater_vec1.for: per (j,k) row, k-faces, j-faces, i-faces, then the block-edge
   max faces.  The per-cell accumulation order is therefore
   +Fk(k) +Fj(j) +Fi(i) -Fi(i+1) -Fj(j+1) -Fk(k+1).                                     */
/* the constants of the flux templates (template_FluxAndBalance.for, template_effort.for) */
static int fc_init(FC* fp, const Ix* Z, const int* p, const double* pr, const double* rop,
                   const double* wig, const double* ti, const double* tj, const double* tk,
                   const double* vol, const double* xmut)
{
    FC f;
    memset(&f, 0, sizeof f);
    f.Z = Z; f.rop = rop; f.xmut = xmut; f.wig = wig; f.ti = ti; f.tj = tj; f.tk = tk; f.vol = vol;
    f.kflu = p[KFLUDOM]; f.slope = p[SLOPE]; f.lamin = p[IFLOW] == 2;
    f.cart = p[ITYPZONE] == 2;
    if (p[ITYPZONE] != 0 && p[ITYPZONE] != 2) return -1;
    if (!(f.kflu == 1 || f.kflu == 2 || f.kflu == 5)) return -1;
    if ((f.kflu == 1 || f.kflu == 2) && f.slope != 2) return -1; /* AUSM/SENSEUR: o3 only */
    f.gam = pr[GAMMA];
    f.rgp = pr[CVINF] * (f.gam - 1.);
    f.gam1 = f.gam / (f.gam - 1.);
    f.gam3 = f.gam1 / pr[PRANDT] * f.rgp;
    f.psiroe = pr[PSIROE];
    f.tmin_1 = 100. / pr[TINF];
    f.c1 = 0.02 * pr[VINF];
    f.c2 = 0.02 / (pr[VINF] * pr[ROINF]);
    f.c3 = -2.;
    f.opt0 = (double)p[SENSORTYPE];
    f.wig_cte = pr[WIG_AMP];
    f.c6 = 1. / 6.; f.c4 = 5. * f.c6; f.c5 = 2. * f.c6; f.c6 = -1. * f.c6;
    f.cvisq = 1. / 3;
    if (f.cart) {
        f.tc[0] = ti[0]; f.tc[1] = tj[0]; f.tc[2] = tk[0];
        f.s[0] = fabs(f.tc[0]); f.s[1] = fabs(f.tc[1]); f.s[2] = fabs(f.tc[2]);
        f.volinv = 0.5 / vol[0];
    }
    *fp = f;
    return 0;
}

/* The flux pass of one cache block ind (template_FluxAndBalance.for:192-271): the faces of the
   block's cells, then the max faces that no later block computes (icorr / jcorr at the last block
   in i / j, the k complement at the last block in k). */
static void flux_block(const FC* f, const int* ind, int icorr, int jcorr, int kcorr, double* drodm)
{
    const Ix* Z = f->Z;
    long nd = Z->nd;
    double flu[5];
#define PLUS(L) for (int e = 0; e < 5; e++) { drodm[(L) + e * nd] = drodm[(L) + e * nd] + flu[e]; }
#define MINUS(L) for (int e = 0; e < 5; e++) { drodm[(L) + e * nd] = drodm[(L) + e * nd] - flu[e]; }
    for (long k = ind[4]; k <= ind[5]; k++) {
        for (long j = ind[2]; j <= ind[3]; j++) {
            for (long i = ind[0]; i <= ind[1]; i++) {
                long l = IDX(Z, i, j, k);
                face_flux(f, i, j, k, 2, flu); PLUS(l); MINUS(l - Z->inck);
            }
            for (long i = ind[0]; i <= ind[1]; i++) {
                long l = IDX(Z, i, j, k);
                face_flux(f, i, j, k, 1, flu); PLUS(l); MINUS(l - Z->incj);
            }
            for (long i = ind[0]; i <= ind[1]; i++) {
                long l = IDX(Z, i, j, k);
                face_flux(f, i, j, k, 0, flu); PLUS(l); MINUS(l - 1);
            }
            if (icorr) {
                long i = ind[1] + 1, l = IDX(Z, i, j, k);
                face_flux(f, i, j, k, 0, flu); MINUS(l - 1);
            }
        }
        if (jcorr) {
            long j = ind[3] + 1;
            for (long i = ind[0]; i <= ind[1]; i++) {
                long l = IDX(Z, i, j, k);
                face_flux(f, i, j, k, 1, flu); MINUS(l - Z->incj);
            }
        }
    }
    if (kcorr) {
        long k = ind[5] + 1;
        for (long j = ind[2]; j <= ind[3]; j++)
            for (long i = ind[0]; i <= ind[1]; i++) {
                long l = IDX(Z, i, j, k);
                face_flux(f, i, j, k, 2, flu); MINUS(l - Z->inck);
            }
    }
#undef PLUS
#undef MINUS
}

/* Number of cache blocks of a direction of n cells (crsdm_scater.for:42-44, one thread: the
   target size is the CACHEBLCK* slot unchanged, WORK_DISTRIBUTION_BEGIN.for:18-28). */
static int cache_nblk(int n, int size)
{
    int b = n / size;
    return b < 1 ? 1 : b;
}

/* The RHS flux pass of navier_stokes_struct.for:157-498 for one thread owning the box ind: the
   cache blocks (LOOP_CACHE_BEGIN.for: kcache, jcache, icache from the outermost) each get their
   fluxes and then, for Roe, their wall corrections.  Block c of a direction spans
   [lo + (c-1) size, lo + c size - 1], the last one up to the end of the box
   (indice_boucle_ssdom.for:40-50).  With one block per direction the per-cell accumulation order
   is +Fk(k) +Fj(j) +Fi(i) -Fi(i+1) -Fj(j+1) -Fk(k+1); a face between two blocks is computed by
   the later block, so block edges change the order in which a cell receives its max faces. */
static int fo_flux(const int* p, const double* pr, const int* ind, const double* rop,
                   double* drodm, const double* wig, const double* ti, const double* tj,
                   const double* tk, const double* vol, const double* xmut)
{
    Ix zz = mkix(p); Ix* Z = &zz;
    FC f;
    if (fc_init(&f, Z, p, pr, rop, wig, ti, tj, tk, vol, xmut)) return -1;
    int size[3] = {p[CACHEBLCKI], p[CACHEBLCKJ], p[CACHEBLCKK]}, nb[3];
    for (int d = 0; d < 3; d++) {
        if (size[d] < 1) return -1;
        nb[d] = cache_nblk(ind[2 * d + 1] - ind[2 * d] + 1, size[d]);
    }
    for (int kc = 1; kc <= nb[2]; kc++)
        for (int jc = 1; jc <= nb[1]; jc++)
            for (int ic = 1; ic <= nb[0]; ic++) {
                int c[3] = {ic, jc, kc}, sdm[6];
                for (int d = 0; d < 3; d++) {
                    sdm[2 * d] = ind[2 * d] + (c[d] - 1) * size[d];
                    sdm[2 * d + 1] = c[d] == nb[d] ? ind[2 * d + 1] : ind[2 * d] + c[d] * size[d] - 1;
                }
                flux_block(&f, sdm, ic == nb[0], jc == nb[1], kc == nb[2], drodm);
                if (f.kflu == 5) correct_flux_roe(p, pr, &f, sdm, drodm);
            }
    return 0;
}

/* ------------------------------------------------------------------------------------ */
/* Wall efforts: bceffort.for -> eff_flu<scheme>_<model>_<slope>_<mesh>, generated from
   template_effort.for:100-440.  On every face of the window ind (the face index in the
   normal direction idir, cells across it), in loop order k, j, i:
   the full face flux (inviscid + viscous, the balance's own fluxes) minus the free-stream
   pressure, normalised by 2 / (rho_inf V_inf^2) and signed outward (sens), stored with the
   face gradients, pressure coefficient, viscosity, density and pressure (21 fields, the
   teff variables of createStressNodes); the 12 sums are forces (1-3), moments about pos
   of the face centre (4-6), the wetted area (7) and the raw fluxes (8-12).
   flu: 21 x nfaces (face-major within each field) or NULL.                               */
int fo_effort(const int* p, const double* pr, int idir, const int* ind, const double* pos,
              const double* rop, const double* xmut, const double* wig, const double* x,
              const double* y, const double* z, const double* ti, const double* tj,
              const double* tk, const double* vol, double* eff, double* flu)
{
    Ix zz = mkix(p); Ix* Z = &zz;
    FC f;
    if (fc_init(&f, Z, p, pr, rop, wig, ti, tj, tk, vol, xmut)) return -1;
    int d = (idir - 1) / 2;
    long nd = Z->nd, ndm = Z->ndm;
    long n0 = ind[1] - ind[0] + 1, n1 = ind[3] - ind[2] + 1, n2 = ind[5] - ind[4] + 1;
    long nf = n0 * n1 * n2;
    double pinf = pr[PINF];
    double norm = 2. / (pr[ROINF] * (pr[VINF] * pr[VINF]));
    double sens = norm, sens1 = 1.;
    if (idir % 2 == 1) { sens = -sens; sens1 = -sens1; }
    double flagi = 1., flagj = 1., flagk = 1.;
    if (p[ITYPZONE] == 2) { flagi = 0.; flagj = 0.; flagk = 0.; }
    long inc = d == 0 ? 1 : d == 1 ? Z->incj : Z->inck;
    long x1, x2;
    if (d == 0) { x1 = Z->xi; x2 = Z->xi * Z->xj; }
    else if (d == 1) { x1 = 1; x2 = Z->xi * Z->xj; }
    else { x1 = 1; x2 = Z->xi; }
    long x3 = x1 + x2;
    for (int q = 0; q < 12; q++) eff[q] = 0.;
    for (long k = ind[4]; k <= ind[5]; k++)
        for (long j = ind[2]; j <= ind[3]; j++)
            for (long i = ind[0]; i <= ind[1]; i++) {
                long l = IDX(Z, i, j, k), lt = 0;
                double qm[5], qp[5], tn[3], sn, fl[5], gv[13];
                for (int q = 0; q < 12; q++) gv[q] = 0.;
                gv[12] = f.volinv;
                if (f.cart) {
                    tn[0] = f.tc[0]; tn[1] = f.tc[1]; tn[2] = f.tc[2]; sn = f.s[d];
                } else {
                    const double* t = d == 0 ? ti : d == 1 ? tj : tk;
                    lt = MTR(Z, i, j, k);
                    tn[0] = t[lt]; tn[1] = t[lt + ndm]; tn[2] = t[lt + 2 * ndm];
                    sn = sqrt(tn[0] * tn[0] + tn[1] * tn[1] + tn[2] * tn[2]);
                }
                slope_state(&f, l, inc, qm, qp);
                inviscid_flux(&f, l, d, tn, sn, f.cart, qm, qp, fl);
                if (f.lamin) {
                    if (f.cart) visc_cart(&f, l, d, fl, gv);
                    else visc_full(&f, l, lt, d, tn, fl, gv);
                }
                double p1 = qp[0] * qp[4] * f.rgp, p2 = qm[0] * qm[4] * f.rgp;
                double tcx = tn[0], tcy = tn[1], tcz = tn[2];
                double v[21];
                v[0] = fl[0] * sens1;
                v[1] = (fl[1] - pinf * tcx * (d == 0 ? 1. : flagi)) * sens;
                v[2] = (fl[2] - pinf * tcy * (d == 1 ? 1. : flagj)) * sens;
                v[3] = (fl[3] - pinf * tcz * (d == 2 ? 1. : flagk)) * sens;
                if (d == 0) v[1] = (fl[1] - pinf * tcx) * sens;
                if (d == 1) v[2] = (fl[2] - pinf * tcy) * sens;
                if (d == 2) v[3] = (fl[3] - pinf * tcz) * sens;
                v[4] = fl[4] * sens;
                for (int q = 0; q < 12; q++) v[5 + q] = gv[q] * 2. * gv[12];
                v[17] = (0.5 * (p1 + p2) - pinf) * norm;
                v[18] = f.lamin ? 0.5 * (xmut[l] + xmut[l - inc]) : 0.;
                /* Roe: etat_roe_GD.for:11-12 leaves qm1, qp1 holding the square roots */
                v[19] = f.kflu == 5 ? 0.5 * (sqrt(qm[0]) + sqrt(qp[0])) : 0.5 * (qm[0] + qp[0]);
                v[20] = 0.5 * (p1 + p2);
                long lx = XYZ(Z, i, j, k);
                double f4 = 0.25 * (x[lx] + x[lx + x1] + x[lx + x2] + x[lx + x3]);
                double f5 = 0.25 * (y[lx] + y[lx + x1] + y[lx + x2] + y[lx + x3]);
                double f6 = 0.25 * (z[lx] + z[lx + x1] + z[lx + x2] + z[lx + x3]);
                double m1 = v[3] * (f5 - pos[1]) - v[2] * (f6 - pos[2]);
                double m2 = v[1] * (f6 - pos[2]) - v[3] * (f4 - pos[0]);
                double m3 = v[2] * (f4 - pos[0]) - v[1] * (f5 - pos[1]);
                eff[0] = eff[0] + v[1]; eff[1] = eff[1] + v[2]; eff[2] = eff[2] + v[3];
                eff[3] = eff[3] + m1; eff[4] = eff[4] + m2; eff[5] = eff[5] + m3;
                eff[6] = eff[6] + sn;
                eff[7] = eff[7] + v[0];
                for (int e = 1; e < 5; e++) eff[7 + e] = eff[7 + e] + fl[e] * sens1;
                if (flu) {
                    long lf = (i - ind[0]) + (j - ind[2]) * n0 + (k - ind[4]) * n0 * n1;
                    for (int q = 0; q < 21; q++) flu[lf + q * nf] = v[q];
                }
            }
    return 0;
}

/* ------------------------------------------------------------------------------------ */
/* Gradients of the pressure drag and lift coefficients at a slip wall:
   compute_dpJ_dpW.cpp:33-261 -> dpj_dpw_calc_meth.for:55-70 -> dpJ_dpW_fluausm_select.for:60-170
   -> dpJ_dpW_fluausm_euler_o3_2d.for (option 213: AUSM+(P), o3, Euler; the reference codes
   the 2-D zone only and leaves every 3-D option empty, so this is its 3-D statement).
   Per face of the window, with F the AUSM+(P) o3 momentum flux of the face (AUSM
   fluFaceEuler.for, the flux inviscid_flux forms) and s = sens * surfinv, sens = -+2/(rho_inf
   V_inf^2) on min / max faces (:359-366):
       J_D = s (cos a F_x + sin a F_y),   J_L = s (cos a F_y - sin a F_x)      (:855-890)
   differentiated (reverse mode here, a forward chain of partials in the reference,
   :468-850) with respect to the primitive states of the four-cell stencil, the sound speed
   c^2 = rgp gam T_l included (:468, :587-614).  The two ghost cells of the stencil follow the
   slip-wall mirror of bvbs_wall_inviscid.for (density and temperature copied, velocity
   reflected about the face plane, :620-700), so their gradients are carried onto the interior
   cells they mirror, and the primitive gradients of the two interior cells are mapped to the
   conservative state (rho, rho u, rho v, rho w, rho E) through dpP_dpW_fluausm_euler_2d_
   idir_{min,max}.for.  The values are assigned, not accumulated, at the two interior cells
   (:958-978), every other cell is left as it was.                                        */
static void dpj_cons(const double q[5], double cvm, const double g[5], double out[5])
{
    double w1 = q[0], w2 = q[0] * q[1], w3 = q[0] * q[2], w4 = q[0] * q[3];
    double ke = 0.5 * (w2 * w2 + w3 * w3 + w4 * w4);
    double w5 = cvm * w1 * q[4] + ke / w1;
    double w11 = w1 * w1;
    double t1 = 1. / cvm * (-w5 / w11 + 2. * ke / (w11 * w1));
    out[0] = g[0] + g[1] * (-w2 / w11) + g[2] * (-w3 / w11) + g[3] * (-w4 / w11) + g[4] * t1;
    out[1] = g[1] * (1. / w1) + g[4] * (-1. / cvm * w2 / w11);
    out[2] = g[2] * (1. / w1) + g[4] * (-1. / cvm * w3 / w11);
    out[3] = g[3] * (1. / w1) + g[4] * (-1. / cvm * w4 / w11);
    out[4] = g[4] * (1. / (cvm * w1));
}

/* reverse pass of one face: q[s] = primitive states of nm2, nm, l, np; fb = weights of
   (F_x, F_y, F_z); wb[s] += d(fb . F)/d q[s] */
static void dpj_face(const FC* f, const double q[4][5], const double tn[3], double sn,
                     const double fb[3], double wb[4][5])
{
    double qp[5], qm[5];
    for (int e = 0; e < 5; e++) {
        qm[e] = f->c4 * q[2][e] + f->c5 * q[1][e] + f->c6 * q[3][e];
        qp[e] = f->c4 * q[1][e] + f->c6 * q[0][e] + f->c5 * q[2][e];
    }
    double rgp = f->rgp, r1 = qp[0], r2 = qm[0];
    double p1 = r1 * qp[4] * rgp, p2 = r2 * qm[4] * rgp;
    double qn1 = qp[1] * tn[0] + qp[2] * tn[1] + qp[3] * tn[2];
    double qn2 = qm[1] * tn[0] + qm[2] * tn[1] + qm[3] * tn[2];
    double c = rgp * f->gam * q[2][4];
    double son = sqrt(qn1 * qn1 / c);
    double tam = f->c3 * son + sn;
    double tam1 = fmax2(0., tam) * f->c2;
    double u = 0.25 * (qn1 + qn2) - tam1 * (p2 - p1);
    double tdu = fmax2(fabs(u), f->c1 * sn) * f->wig_cte;
    double ub = 0., tdub = 0., p1p2b = 0., m1b[3], m2b[3];
    for (int m = 0; m < 3; m++) {
        double mom1 = r1 * qp[1 + m], mom2 = r2 * qm[1 + m];
        ub = ub + fb[m] * (mom1 + mom2);
        tdub = tdub + fb[m] * (mom1 - mom2);
        m1b[m] = fb[m] * (u + tdu);
        m2b[m] = fb[m] * (u - tdu);
        p1p2b = p1p2b + fb[m] * tn[m];
    }
    double p1b = 0.5 * p1p2b, p2b = 0.5 * p1p2b;
    if (fabs(u) >= f->c1 * sn) ub = ub + tdub * f->wig_cte * (u >= 0. ? 1. : -1.);
    double qn1b = 0.25 * ub, qn2b = 0.25 * ub;
    double tam1b = -ub * (p2 - p1);
    p1b = p1b + ub * tam1;
    p2b = p2b - ub * tam1;
    double tamb = tam >= 0. ? tam1b * f->c2 : 0.;
    double sonb = f->c3 * tamb;
    double sc = sqrt(c);
    qn1b = qn1b + sonb * (qn1 >= 0. ? 1. / sc : -1. / sc);
    double cb = sonb * (-0.5 * sqrt(qn1 * qn1) / sqrt(c * c * c));
    double qpb[5], qmb[5];
    qpb[0] = p1b * qp[4] * rgp; qmb[0] = p2b * qm[4] * rgp;
    for (int m = 0; m < 3; m++) {
        qpb[0] = qpb[0] + m1b[m] * qp[1 + m];
        qmb[0] = qmb[0] + m2b[m] * qm[1 + m];
        qpb[1 + m] = qn1b * tn[m] + m1b[m] * r1;
        qmb[1 + m] = qn2b * tn[m] + m2b[m] * r2;
    }
    qpb[4] = p1b * r1 * rgp; qmb[4] = p2b * r2 * rgp;
    for (int e = 0; e < 5; e++) {
        wb[0][e] = wb[0][e] + f->c6 * qpb[e];
        wb[1][e] = wb[1][e] + (f->c4 * qpb[e] + f->c5 * qmb[e]);
        wb[2][e] = wb[2][e] + (f->c5 * qpb[e] + f->c4 * qmb[e]);
        wb[3][e] = wb[3][e] + f->c6 * qmb[e];
    }
    wb[2][4] = wb[2][4] + cb * (rgp * f->gam);
}

/* ghost gradient g carried onto the interior cell it mirrors: x += M^T g, M = the slip-wall
   mirror (velocity reflected about the plane of unit normal n) */
static void dpj_mirror(const double n[3], const double g[5], double x[5])
{
    x[0] = x[0] + g[0];
    for (int a = 0; a < 3; a++) {
        double s = 0.;
        for (int b = 0; b < 3; b++) s = s + g[1 + b] * ((a == b ? 1. : 0.) - 2. * n[b] * n[a]);
        x[1 + a] = x[1 + a] + s;
    }
    x[4] = x[4] + g[4];
}

int fo_dpj_dpw(const int* p, const double* pr, int idir, const int* ind, const double* rop,
               const double* ti, const double* tj, const double* tk, double cosa, double sina,
               double surfinv, double* dcd, double* dcl)
{
    Ix zz = mkix(p); Ix* Z = &zz;
    FC f;
    double vdum = 1.;                /* fc_init reads a Cartesian zone's volume */
    if (fc_init(&f, Z, p, pr, rop, NULL, ti, tj, tk, &vdum, NULL)) return -1;
    if (f.kflu != 1 || f.slope != 2 || p[IFLOW] != 1) return -1;    /* option 213 only */
    int d = (idir - 1) / 2, up = idir % 2 == 0;
    long nd = Z->nd, ndm = Z->ndm;
    long inc = d == 0 ? 1 : d == 1 ? Z->incj : Z->inck;
    double norm = 2. / (pr[ROINF] * (pr[VINF] * pr[VINF]));
    double sens = up ? norm : -norm;
    double cc = sens * surfinv * cosa, cs = sens * surfinv * sina;
    double fbd[3] = {cc, cs, 0.}, fbl[3] = {-cs, cc, 0.};
    double cvm = f.rgp / (f.gam - 1.);
    for (long k = ind[4]; k <= ind[5]; k++)
        for (long j = ind[2]; j <= ind[3]; j++)
            for (long i = ind[0]; i <= ind[1]; i++) {
                long l = IDX(Z, i, j, k);
                long cell[4] = {l - 2 * inc, l - inc, l, l + inc};
                double tn[3] = {0., 0., 0.}, sn;
                if (f.cart) { tn[d] = f.tc[d]; sn = f.s[d]; }
                else {
                    const double* t = d == 0 ? ti : d == 1 ? tj : tk;
                    long lt = MTR(Z, i, j, k);
                    tn[0] = t[lt]; tn[1] = t[lt + ndm]; tn[2] = t[lt + 2 * ndm];
                    sn = sqrt(tn[0] * tn[0] + tn[1] * tn[1] + tn[2] * tn[2]);
                }
                double nrm = fmax2(sqrt(tn[0] * tn[0] + tn[1] * tn[1] + tn[2] * tn[2]), 1e-30);
                double n[3] = {tn[0] / nrm, tn[1] / nrm, tn[2] / nrm};
                double q[4][5];
                for (int s = 0; s < 4; s++)
                    for (int e = 0; e < 5; e++) q[s][e] = rop[cell[s] + e * nd];
                /* interior cells (A nearest the wall) and the ghosts mirroring them */
                int A = up ? 1 : 2, B = up ? 0 : 3, gA = up ? 2 : 1, gB = up ? 3 : 0;
                for (int o = 0; o < 2; o++) {
                    double wb[4][5];
                    memset(wb, 0, sizeof wb);
                    dpj_face(&f, q, tn, sn, o == 0 ? fbd : fbl, wb);
                    dpj_mirror(n, wb[gA], wb[A]);
                    dpj_mirror(n, wb[gB], wb[B]);
                    double ga[5], gb[5];
                    dpj_cons(q[A], cvm, wb[A], ga);
                    dpj_cons(q[B], cvm, wb[B], gb);
                    double* out = o == 0 ? dcd : dcl;
                    for (int e = 0; e < 5; e++) {
                        out[cell[A] + e * nd] = ga[e];
                        out[cell[B] + e * nd] = gb[e];
                    }
                }
            }
    return 0;
}

/* ------------------------------------------------------------------------------------ */
/* Transposed Jacobian product of the AUSM+(P) o3 Euler face balance: the reverse mode the
   reference has for 2-D zones as Tapenade output (ADJOINT/AUSM/2d/fluausm_euler_o3_2d_b.for:
   the gradient of drodm with respect to rop), stated here for 3-D zones.  out = (dR/dW)^T lam,
   R_l = sum_d F_d(l) - F_d(l + inc_d) over the interior cells (the balance fo_flux forms),
   W the primitive state of every cell, ghost cells included.  Per direction i, j, k: every
   face (index 1..n_d + 1, interior across) is weighted by lam_l - lam_{l - inc} (each term only
   where its cell is interior) and its reverse pass gives the contributions to the four stencil
   cells; each cell then adds the contributions of the faces whose stencil holds it, in face
   order (faces c - 1, c, c + 1, c + 2 of cell c), and adds that sum to out.               */
static void ausm_face_rev(const FC* f, const double q[4][5], const double tn[3], double sn,
                          const double fb[5], double wb[4][5])
{
    double qp[5], qm[5];
    for (int e = 0; e < 5; e++) {
        qm[e] = f->c4 * q[2][e] + f->c5 * q[1][e] + f->c6 * q[3][e];
        qp[e] = f->c4 * q[1][e] + f->c6 * q[0][e] + f->c5 * q[2][e];
    }
    double rgp = f->rgp, gam1 = f->gam1, r1 = qp[0], r2 = qm[0];
    double mom1[3], mom2[3];
    for (int m = 0; m < 3; m++) { mom1[m] = r1 * qp[1 + m]; mom2[m] = r2 * qm[1 + m]; }
    double p1 = r1 * qp[4] * rgp, p2 = r2 * qm[4] * rgp;
    double h1 = gam1 * p1 + .5 * (mom1[0] * qp[1] + mom1[1] * qp[2] + mom1[2] * qp[3]);
    double h2 = gam1 * p2 + .5 * (mom2[0] * qm[1] + mom2[1] * qm[2] + mom2[2] * qm[3]);
    double qn1 = qp[1] * tn[0] + qp[2] * tn[1] + qp[3] * tn[2];
    double qn2 = qm[1] * tn[0] + qm[2] * tn[1] + qm[3] * tn[2];
    double c = rgp * f->gam * q[2][4];
    double son = sqrt(qn1 * qn1 / c);
    double tam = f->c3 * son + sn;
    double tam1 = fmax2(0., tam) * f->c2;
    double u = 0.25 * (qn1 + qn2) - tam1 * (p2 - p1);
    double tdu = fmax2(fabs(u), f->c1 * sn) * f->wig_cte;
    double ub = fb[0] * (r1 + r2), tdub = -(fb[0] * (r2 - r1)), p1p2b = 0.;
    double m1b[3], m2b[3];
    for (int m = 0; m < 3; m++) {
        ub = ub + fb[1 + m] * (mom1[m] + mom2[m]);
        tdub = tdub - fb[1 + m] * (mom2[m] - mom1[m]);
        m1b[m] = fb[1 + m] * (u + tdu);
        m2b[m] = fb[1 + m] * (u - tdu);
        p1p2b = p1p2b + fb[1 + m] * tn[m];
    }
    ub = ub + fb[4] * (h1 + h2);
    tdub = tdub - fb[4] * (h2 - h1);
    double r1b = fb[0] * (u + tdu), r2b = fb[0] * (u - tdu);
    double h1b = fb[4] * (u + tdu), h2b = fb[4] * (u - tdu);
    double qpb[5], qmb[5];
    double p1b = gam1 * h1b + 0.5 * p1p2b, p2b = gam1 * h2b + 0.5 * p1p2b;
    for (int m = 0; m < 3; m++) {
        m1b[m] = m1b[m] + .5 * h1b * qp[1 + m];
        m2b[m] = m2b[m] + .5 * h2b * qm[1 + m];
        qpb[1 + m] = .5 * h1b * mom1[m];
        qmb[1 + m] = .5 * h2b * mom2[m];
    }
    if (fabs(u) >= f->c1 * sn) ub = ub + tdub * f->wig_cte * (u >= 0. ? 1. : -1.);
    double qn1b = 0.25 * ub, qn2b = 0.25 * ub;
    double tam1b = -ub * (p2 - p1);
    p1b = p1b + ub * tam1;
    p2b = p2b - ub * tam1;
    double tamb = tam >= 0. ? tam1b * f->c2 : 0.;
    double sonb = f->c3 * tamb;
    double sc = sqrt(c);
    qn1b = qn1b + sonb * (qn1 >= 0. ? 1. / sc : -1. / sc);
    double cb = sonb * (-0.5 * sqrt(qn1 * qn1) / sqrt(c * c * c));
    r1b = r1b + p1b * qp[4] * rgp;
    r2b = r2b + p2b * qm[4] * rgp;
    for (int m = 0; m < 3; m++) {
        r1b = r1b + m1b[m] * qp[1 + m];
        r2b = r2b + m2b[m] * qm[1 + m];
        qpb[1 + m] = qpb[1 + m] + qn1b * tn[m] + m1b[m] * r1;
        qmb[1 + m] = qmb[1 + m] + qn2b * tn[m] + m2b[m] * r2;
    }
    qpb[0] = r1b; qmb[0] = r2b;
    qpb[4] = p1b * r1 * rgp; qmb[4] = p2b * r2 * rgp;
    for (int e = 0; e < 5; e++) {
        wb[0][e] = f->c6 * qpb[e];
        wb[1][e] = f->c4 * qpb[e] + f->c5 * qmb[e];
        wb[2][e] = f->c5 * qpb[e] + f->c4 * qmb[e];
        wb[3][e] = f->c6 * qmb[e];
    }
    wb[2][4] = wb[2][4] + cb * (rgp * f->gam);
}

int fo_residual_vjp(const int* p, const double* pr, const double* rop, const double* ti,
                    const double* tj, const double* tk, const double* lam, double* out)
{
    Ix zz = mkix(p); Ix* Z = &zz;
    FC f;
    double vdum = 1.;
    if (fc_init(&f, Z, p, pr, rop, NULL, ti, tj, tk, &vdum, NULL)) return -1;
    if (f.kflu != 1 || f.slope != 2 || p[IFLOW] != 1) return -1;
    long nd = Z->nd, ndm = Z->ndm;
    int n[3] = {p[IJKV], p[IJKV + 1], p[IJKV + 2]};
    for (long q = 0; q < 5 * nd; q++) out[q] = 0.;
    for (int d = 0; d < 3; d++) {
        long inc = d == 0 ? 1 : d == 1 ? Z->incj : Z->inck;
        int e0 = n[0] + (d == 0), e1 = n[1] + (d == 1), e2 = n[2] + (d == 2);
        long nf = (long)e0 * e1 * e2;
        double* buf = (double*)malloc(sizeof(double) * 20 * (size_t)nf);
        if (!buf) return -2;
        long t = 0;
        for (int k = 1; k <= e2; k++)
            for (int j = 1; j <= e1; j++)
                for (int i = 1; i <= e0; i++, t++) {
                    long l = IDX(Z, i, j, k);
                    int fi = d == 0 ? i : d == 1 ? j : k;
                    double fb[5];
                    for (int e = 0; e < 5; e++) {
                        double a = fi <= n[d] ? lam[l + e * nd] : 0.;
                        double b = fi >= 2 ? lam[l - inc + e * nd] : 0.;
                        fb[e] = a - b;
                    }
                    double tn[3] = {0., 0., 0.}, sn;
                    if (f.cart) { tn[d] = f.tc[d]; sn = f.s[d]; }
                    else {
                        const double* tm = d == 0 ? ti : d == 1 ? tj : tk;
                        long lt = MTR(Z, i, j, k);
                        tn[0] = tm[lt]; tn[1] = tm[lt + ndm]; tn[2] = tm[lt + 2 * ndm];
                        sn = sqrt(tn[0] * tn[0] + tn[1] * tn[1] + tn[2] * tn[2]);
                    }
                    double q[4][5], wb[4][5];
                    for (int s = 0; s < 4; s++)
                        for (int e = 0; e < 5; e++) q[s][e] = rop[l + (s - 2) * inc + e * nd];
                    ausm_face_rev(&f, q, tn, sn, fb, wb);
                    for (int s = 0; s < 4; s++)
                        for (int e = 0; e < 5; e++) buf[(s * 5 + e) * nf + t] = wb[s][e];
                }
        /* cells -1 .. n_d + 2 along d, interior across */
        int lo[3] = {1, 1, 1}, hi[3] = {n[0], n[1], n[2]};
        lo[d] = -1; hi[d] = n[d] + 2;
        for (int k = lo[2]; k <= hi[2]; k++)
            for (int j = lo[1]; j <= hi[1]; j++)
                for (int i = lo[0]; i <= hi[0]; i++) {
                    int c = d == 0 ? i : d == 1 ? j : k;
                    long l = IDX(Z, i, j, k);
                    double acc[5] = {0., 0., 0., 0., 0.};
                    for (int s = 3; s >= 0; s--) {       /* faces c - 1, c, c + 1, c + 2 */
                        int fidx = c + 2 - s;
                        if (fidx < 1 || fidx > n[d] + 1) continue;
                        int fi = d == 0 ? fidx : i, fj = d == 1 ? fidx : j, fk = d == 2 ? fidx : k;
                        long tf = (long)(fi - 1) + (long)(fj - 1) * e0 + (long)(fk - 1) * e0 * e1;
                        for (int e = 0; e < 5; e++) acc[e] = acc[e] + buf[(s * 5 + e) * nf + tf];
                    }
                    for (int e = 0; e < 5; e++) out[l + e * nd] = out[l + e * nd] + acc[e];
                }
        free(buf);
    }
    return 0;
}

/* The face balance alone over the interior cells (the part of the RHS the transposed product
   above differentiates): drodm = 0 then fo_flux. */
int fo_flux_balance(const int* p, const double* pr, const double* rop, const double* wig,
                    const double* ti, const double* tj, const double* tk, const double* vol,
                    const double* xmut, double* drodm)
{
    Ix zz = mkix(p);
    for (long q = 0; q < 5 * zz.nd; q++) drodm[q] = 0.;
    int ind[6] = {1, p[IJKV], 1, p[IJKV + 1], 1, p[IJKV + 2]};
    return fo_flux(p, pr, ind, rop, drodm, wig, ti, tj, tk, vol, xmut);
}

/* ------------------------------------------------------------------------------------ */
/* core3as2.for (neq = 5, non-2d): Newton/BDF residual                                   */
static void fo_core3as2(const int* p, const double* pr, int nitcfg, int first_it,
                        const int* ind, const double* rop, const double* rop_n,
                        const double* rop_n1, double* drodm, const double* coe)
{
    Ix zz = mkix(p); Ix* Z = &zz;
    long nd = Z->nd;
    double cv = pr[CVINF], c1, c2, c3;
    if (first_it == 0 || p[DTLOC] == 1) { c1 = -1.; c2 = 1.; c3 = 0.; }
    else { c1 = -1.5; c2 = 2.; c3 = -0.5; }
    for (long k = ind[4]; k <= ind[5]; k++)
        for (long j = ind[2]; j <= ind[3]; j++)
            for (long i = ind[0]; i <= ind[1]; i++) {
                long l = IDX(Z, i, j, k);
                const double *a = rop, *b = rop_n, *c = rop_n1;
                double* dr = drodm;
                double cf = coe[l];
                if (nitcfg == 1) {
                    double c4 = c1 + c2;
                    dr[l] = cf * dr[l] + c4 * a[l] + c3 * c[l];
                    for (int m = 1; m < 4; m++)
                        dr[l + m * nd] = cf * dr[l + m * nd] + c4 * a[l + m * nd] * a[l] + c3 * c[l + m * nd] * c[l];
                    double roe = a[l] * (cv * a[l + 4 * nd] + 0.5 * (a[l + nd] * a[l + nd] + a[l + 2 * nd] * a[l + 2 * nd] + a[l + 3 * nd] * a[l + 3 * nd]));
                    double roe_n1 = c[l] * (cv * c[l + 4 * nd] + 0.5 * (c[l + nd] * c[l + nd] + c[l + 2 * nd] * c[l + 2 * nd] + c[l + 3 * nd] * c[l + 3 * nd]));
                    dr[l + 4 * nd] = cf * dr[l + 4 * nd] + c4 * roe + c3 * roe_n1;
                } else {
                    dr[l] = cf * dr[l] + c1 * a[l] + c2 * b[l] + c3 * c[l];
                    for (int m = 1; m < 4; m++)
                        dr[l + m * nd] = cf * dr[l + m * nd] + c1 * a[l + m * nd] * a[l] + c2 * b[l + m * nd] * b[l] + c3 * c[l + m * nd] * c[l];
                    double roe = a[l] * (cv * a[l + 4 * nd] + 0.5 * (a[l + nd] * a[l + nd] + a[l + 2 * nd] * a[l + 2 * nd] + a[l + 3 * nd] * a[l + 3 * nd]));
                    double roe_n = b[l] * (cv * b[l + 4 * nd] + 0.5 * (b[l + nd] * b[l + nd] + b[l + 2 * nd] * b[l + 2 * nd] + b[l + 3 * nd] * b[l + 3 * nd]));
                    double roe_n1 = c[l] * (cv * c[l + 4 * nd] + 0.5 * (c[l + nd] * c[l + nd] + c[l + 2 * nd] * c[l + 2 * nd] + c[l + 3 * nd] * c[l + 3 * nd]));
                    dr[l + 4 * nd] = cf * dr[l + 4 * nd] + c1 * roe + c2 * roe_n + c3 * roe_n1;
                }
            }
}

/* core3ark3.for (explicit global, neq = 5): W_k = W_{k-1} + coefW*coe1*drodm */
static void fo_core3ark3(const int* p, const double* pr, int nitcfg, const int* ind,
                         double* rop_1, const double* rop, const double* drodm, const double* coe)
{
    Ix zz = mkix(p); Ix* Z = &zz;
    long nd = Z->nd;
    double cv = pr[CVINF], cvinv = 1. / cv, cvinv2 = 0.5 * cvinv;
    int b = p[RK];
    double coefW = 1.;
    if (b == 2 && nitcfg == 2) coefW = 0.5;
    if (b == 3) coefW = nitcfg == 1 ? 0.5 : nitcfg == 2 ? 0.9106836025229591 : 0.3660254037844387;
    for (long k = ind[4]; k <= ind[5]; k++)
        for (long j = ind[2]; j <= ind[3]; j++)
            for (long i = ind[0]; i <= ind[1]; i++) {
                long l = IDX(Z, i, j, k);
                double r_1 = coefW * coe[l];
                double t1 = drodm[l] * r_1, t2 = drodm[l + nd] * r_1, t3 = drodm[l + 2 * nd] * r_1;
                double t4 = drodm[l + 3 * nd] * r_1, t5 = drodm[l + 4 * nd] * r_1;
                double ro_old = rop[l], u_old = rop[l + nd], v_old = rop[l + 2 * nd];
                double w_old = rop[l + 3 * nd], t_old = rop[l + 4 * nd];
                rop_1[l] = rop[l] + t1;
                r_1 = 1. / rop_1[l];
                rop_1[l + nd] = (ro_old * u_old + t2) * r_1;
                rop_1[l + 2 * nd] = (ro_old * v_old + t3) * r_1;
                rop_1[l + 3 * nd] = (ro_old * w_old + t4) * r_1;
                double roe_old = ro_old * (cv * t_old + 0.5 * (u_old * u_old + v_old * v_old + w_old * w_old));
                rop_1[l + 4 * nd] = (roe_old + t5) * r_1 * cvinv
                    - cvinv2 * (rop_1[l + nd] * rop_1[l + nd] + rop_1[l + 2 * nd] * rop_1[l + 2 * nd] + rop_1[l + 3 * nd] * rop_1[l + 3 * nd]);
            }
}

/* ------------------------------------------------------------------------------------ */
/* LU-SGS: invlu_l.for (lower, signe = +1/2) and invlu_u.for (upper, signe = -1/2) with
   lu_{i,j,k}_3d{cart,full}.for, mjr_drodm*.for, lu_dinv*.for, lu_d.for, mjr_newton.for.
   The recurrence visits cells lexicographically; for each cell the neighbour terms are added
   in a fixed order (3dcart lower J,K,I; 3dfull lower K,J,I; upper D,K,J,I for both), each
   term being xal * B(W_nb) * dW_nb with xal = coe1(l)*signe.                             */
typedef struct {
    const Ix* Z;
    const double *rop, *coe, *ti, *tj, *tk;
    double gam1, gamm1, cp, gam2, signe;
    int cart;
    long dm;        /* metric direction sign: +1 lower, -1 upper */
} LU;

/* b = B(W_nb) dW_nb for neighbour l1 in direction d (0 i, 1 j, 2 k); lt = metric index of l */
static void lu_term(const LU* s, long l1, long lt, int d, const double* dw, long nd, double b[5])
{
    const Ix* Z = s->Z;
    const double* rop = s->rop;
    double u = rop[l1 + nd], v = rop[l1 + 2 * nd], w = rop[l1 + 3 * nd], t = rop[l1 + 4 * nd];
    double q2 = .5 * (u * u + v * v + w * w);
    double h = s->cp * t + q2;
    double ph2 = s->gamm1 * q2;
    double cs = s->coe[l1 + (1 + d) * nd] * s->signe;
    double gam1 = s->gam1, gamm1 = s->gamm1, gam2 = s->gam2;
    double d1 = dw[l1], d2 = dw[l1 + nd], d3 = dw[l1 + 2 * nd], d4 = dw[l1 + 3 * nd], d5 = dw[l1 + 4 * nd];
    if (s->cart) {
        double tc = d == 0 ? s->ti[0] : d == 1 ? s->tj[0] : s->tk[0];
        if (d == 0) {
            double qn = tc * u;
            double b11 = cs, b12 = tc;
            double b21 = (tc * ph2 - u * qn), b22 = (qn - tc * gam2 * u) + cs, b23 = (-tc * gamm1 * v);
            double b24 = (-tc * gamm1 * w), b25 = tc * gamm1;
            double b31 = (-v * qn), b32 = (tc * v), b33 = (qn) + cs;
            double b41 = (-w * qn), b42 = (tc * w), b44 = (qn) + cs;
            double b51 = qn * (ph2 - h), b52 = (tc * h - gamm1 * u * qn), b53 = (-gamm1 * v * qn);
            double b54 = (-gamm1 * w * qn), b55 = (gam1 * qn) + cs;
            b[0] = b11 * d1 + b12 * d2;
            b[1] = b21 * d1 + b22 * d2 + b23 * d3 + b24 * d4 + b25 * d5;
            b[2] = b31 * d1 + b32 * d2 + b33 * d3;
            b[3] = b41 * d1 + b42 * d2 + b44 * d4;
            b[4] = b51 * d1 + b52 * d2 + b53 * d3 + b54 * d4 + b55 * d5;
        } else if (d == 1) {
            double qn = tc * v;
            double b11 = cs, b13 = tc;
            double b21 = (-u * qn), b22 = qn + cs, b23 = tc * u;
            double b31 = (tc * ph2 - v * qn), b32 = (-tc * gamm1 * u), b33 = (qn - tc * gam2 * v) + cs;
            double b34 = (-tc * gamm1 * w), b35 = tc * gamm1;
            double b41 = (-w * qn), b43 = (tc * w), b44 = (qn) + cs;
            double b51 = qn * (ph2 - h), b52 = (-gamm1 * u * qn), b53 = (tc * h - gamm1 * v * qn);
            double b54 = (-gamm1 * w * qn), b55 = (gam1 * qn) + cs;
            b[0] = b11 * d1 + b13 * d3;
            b[1] = b21 * d1 + b22 * d2 + b23 * d3;
            b[2] = b31 * d1 + b32 * d2 + b33 * d3 + b34 * d4 + b35 * d5;
            b[3] = b41 * d1 + b43 * d3 + b44 * d4;
            b[4] = b51 * d1 + b52 * d2 + b53 * d3 + b54 * d4 + b55 * d5;
        } else {
            double qn = tc * w;
            double b11 = cs, b14 = tc;
            double b21 = -u * qn, b22 = qn + cs, b24 = tc * u;
            double b31 = -v * qn, b33 = qn + cs, b34 = tc * v;
            double b41 = tc * ph2 - w * qn, b42 = -tc * gamm1 * u, b43 = -tc * gamm1 * v;
            double b44 = (qn - tc * gam2 * w) + cs, b45 = tc * gamm1;
            double b51 = qn * (ph2 - h), b52 = -gamm1 * u * qn, b53 = -gamm1 * v * qn;
            double b54 = tc * h - gamm1 * w * qn, b55 = gam1 * qn + cs;
            b[0] = b11 * d1 + b14 * d4;
            b[1] = b21 * d1 + b22 * d2 + b24 * d4;
            b[2] = b31 * d1 + b33 * d3 + b34 * d4;
            b[3] = b41 * d1 + b42 * d2 + b43 * d3 + b44 * d4 + b45 * d5;
            b[4] = b51 * d1 + b52 * d2 + b53 * d3 + b54 * d4 + b55 * d5;
        }
        return;
    }
    /* 3dfull: metric at the neighbour centre = mean of its two faces */
    const double* tm = d == 0 ? s->ti : d == 1 ? s->tj : s->tk;
    long inc = d == 0 ? Z->mi : d == 1 ? Z->mj : Z->mk;
    long ndm = Z->ndm;
    long lt1, lt2;
    if (s->dm > 0) { lt1 = lt - inc; lt2 = lt; }          /* inc2_mtr = 0 in the lower sweep */
    else { lt1 = lt + inc; lt2 = lt + 2 * inc; }          /* inc_mtr < 0, inc2_mtr = 2*inc_mtr */
    double tcx = 0.5 * (tm[lt2] + tm[lt1]);
    double tcy = 0.5 * (tm[lt2 + ndm] + tm[lt1 + ndm]);
    double tcz = 0.5 * (tm[lt2 + 2 * ndm] + tm[lt1 + 2 * ndm]);
    double qn = tcx * u + tcy * v + tcz * w;
    double b11 = cs, b12 = tcx, b13 = tcy, b14 = tcz;
    double b21 = (tcx * ph2 - u * qn), b22 = (qn - tcx * gam2 * u) + cs;
    double b23 = (tcy * u - tcx * gamm1 * v), b24 = (tcz * u - tcx * gamm1 * w), b25 = tcx * gamm1;
    double b31 = (tcy * ph2 - v * qn), b32 = (tcx * v - tcy * gamm1 * u);
    double b33 = (qn - tcy * gam2 * v) + cs, b34 = (tcz * v - tcy * gamm1 * w), b35 = tcy * gamm1;
    double b41 = (tcz * ph2 - w * qn), b42 = (tcx * w - tcz * gamm1 * u);
    double b43 = (tcy * w - tcz * gamm1 * v), b44 = (qn - tcz * gam2 * w) + cs, b45 = tcz * gamm1;
    double b51 = qn * (ph2 - h), b52 = (tcx * h - gamm1 * u * qn), b53 = (tcy * h - gamm1 * v * qn);
    double b54 = (tcz * h - gamm1 * w * qn), b55 = (gam1 * qn) + cs;
    b[0] = b11 * d1 + b12 * d2 + b13 * d3 + b14 * d4;
    b[1] = b21 * d1 + b22 * d2 + b23 * d3 + b24 * d4 + b25 * d5;
    b[2] = b31 * d1 + b32 * d2 + b33 * d3 + b34 * d4 + b35 * d5;
    b[3] = b41 * d1 + b42 * d2 + b43 * d3 + b44 * d4 + b45 * d5;
    b[4] = b51 * d1 + b52 * d2 + b53 * d3 + b54 * d4 + b55 * d5;
}

static void mjr_newton(const Ix* Z, const double* pr, long l, const double* rop, double* rop_1,
                       const double* dw)
{
    long nd = Z->nd;
    double cvinv = 1. / pr[CVINF], cvinv2 = 0.5 * cvinv;
    double ro_old = rop[l], u_old = rop[l + nd], v_old = rop[l + 2 * nd];
    double w_old = rop[l + 3 * nd], t_old = rop[l + 4 * nd];
    rop_1[l] = rop[l] + dw[l];
    double r_1 = 1. / rop_1[l];
    rop_1[l + nd] = (ro_old * u_old + dw[l + nd]) * r_1;
    rop_1[l + 2 * nd] = (ro_old * v_old + dw[l + 2 * nd]) * r_1;
    rop_1[l + 3 * nd] = (ro_old * w_old + dw[l + 3 * nd]) * r_1;
    double roe_old = ro_old * (pr[CVINF] * t_old + 0.5 * (u_old * u_old + v_old * v_old + w_old * w_old));
    rop_1[l + 4 * nd] = (roe_old + dw[l + 4 * nd]) * r_1 * cvinv
        - cvinv2 * (rop_1[l + nd] * rop_1[l + nd] + rop_1[l + 2 * nd] * rop_1[l + 2 * nd] + rop_1[l + 3 * nd] * rop_1[l + 3 * nd]);
}

int fo_lu_sgs(int* p, double* pr, const int* ind, double* rotmp, const double* rop,
              double* dw, const double* ti, const double* tj, const double* tk, const double* coe,
              int mjrnewton)
{
    Ix zz = mkix(p); Ix* Z = &zz;
    long nd = Z->nd;
    if (ind[1] < ind[0]) return 0;
    if (p[ITYPZONE] != 0 && p[ITYPZONE] != 2) return -1;
    LU s;
    s.Z = Z; s.rop = rop; s.coe = coe; s.ti = ti; s.tj = tj; s.tk = tk;
    s.gam1 = pr[GAMMA]; s.gamm1 = s.gam1 - 1.; s.cp = pr[GAMMA] * pr[CVINF]; s.gam2 = s.gamm1 - 1.;
    s.cart = p[ITYPZONE] == 2;
    double b[5];
    /* lower sweep */
    s.signe = 0.5; s.dm = 1;
    for (long k = ind[4]; k <= ind[5]; k++)
        for (long j = ind[2]; j <= ind[3]; j++)
            for (long i = ind[0]; i <= ind[1]; i++) {
                long l = IDX(Z, i, j, k), lt = s.cart ? 0 : MTR(Z, i, j, k);
                double xal = coe[l] * s.signe;
                int hasI = i > ind[0], hasJ = j > ind[2], hasK = k > ind[4];
                int order[2] = {s.cart ? 1 : 2, s.cart ? 2 : 1};
                for (int o = 0; o < 2; o++) {
                    int d = order[o];
                    if ((d == 1 && !hasJ) || (d == 2 && !hasK)) continue;
                    long l1 = l - (d == 1 ? Z->incj : Z->inck);
                    lu_term(&s, l1, lt, d, dw, nd, b);
                    for (int e = 0; e < 5; e++) dw[l + e * nd] = dw[l + e * nd] + b[e] * xal;
                }
                if (hasI) {
                    lu_term(&s, l - 1, lt, 0, dw, nd, b);
                    for (int e = 0; e < 5; e++) dw[l + e * nd] = dw[l + e * nd] + b[e] * xal;
                }
                double diag = 1. / coe[l + 4 * nd];
                for (int e = 0; e < 5; e++) dw[l + e * nd] = dw[l + e * nd] * diag;
            }
    /* upper sweep */
    s.signe = -0.5; s.dm = -1;
    for (long k = ind[5]; k >= ind[4]; k--) {
        for (long j = ind[3]; j >= ind[2]; j--)
            for (long i = ind[1]; i >= ind[0]; i--) {
                long l = IDX(Z, i, j, k), lt = s.cart ? 0 : MTR(Z, i, j, k);
                double xal = coe[l] * s.signe;
                double dg = coe[l + 4 * nd];
                for (int e = 0; e < 5; e++) dw[l + e * nd] = dw[l + e * nd] * dg;
                if (k < ind[5]) {
                    lu_term(&s, l + Z->inck, lt, 2, dw, nd, b);
                    for (int e = 0; e < 5; e++) dw[l + e * nd] = dw[l + e * nd] + b[e] * xal;
                }
                if (j < ind[3]) {
                    lu_term(&s, l + Z->incj, lt, 1, dw, nd, b);
                    for (int e = 0; e < 5; e++) dw[l + e * nd] = dw[l + e * nd] + b[e] * xal;
                }
                if (i < ind[1]) {
                    lu_term(&s, l + 1, lt, 0, dw, nd, b);
                    for (int e = 0; e < 5; e++) dw[l + e * nd] = dw[l + e * nd] + b[e] * xal;
                }
                double diag = 1. / coe[l + 4 * nd];
                for (int e = 0; e < 5; e++) dw[l + e * nd] = dw[l + e * nd] * diag;
            }
        if (mjrnewton == 1)
            for (long j = ind[2]; j <= ind[3]; j++)
                for (long i = ind[0]; i <= ind[1]; i++) mjr_newton(Z, pr, IDX(Z, i, j, k), rop, rotmp, dw);
    }
    return 0;
}

/* ------------------------------------------------------------------------------------ */
/* cprdu3s1.for + cprdu3s1_incl.for: sum of squares and max|.| per equation and residual
   sub-block (SIZE_SSDOM), stored in rdm(nssiter, neq, 2, ndim_rdm) at it_bloc = 1.        */
void fo_residual_blocks(const int* p, int nitcfg, int nssiter, const double* drodm, double* rdm)
{
    Ix zz = mkix(p); Ix* Z = &zz;
    (void)nitcfg;
    long nd = Z->nd;
    int ijkv_lu[3];
    for (int i = 0; i < 3; i++) { ijkv_lu[i] = p[IJKV + i] / p[SIZE_SSDOM + i]; if (ijkv_lu[i] < 1) ijkv_lu[i] = 1; }
    int ndim = ijkv_lu[0] * ijkv_lu[1] * ijkv_lu[2];
    int neq = 5;
    for (int no = 1; no <= ndim; no++) {
        int k_lu = 1 + (no - 1) / (ijkv_lu[0] * ijkv_lu[1]);
        int j_lu = 1 + (no - 1 - ijkv_lu[0] * ijkv_lu[1] * (k_lu - 1)) / ijkv_lu[0];
        int i_lu = no - (j_lu - 1) * ijkv_lu[0] - (k_lu - 1) * ijkv_lu[0] * ijkv_lu[1];
        int lo[3] = {1 + (i_lu - 1) * p[SIZE_SSDOM], 1 + (j_lu - 1) * p[SIZE_SSDOM + 1], 1 + (k_lu - 1) * p[SIZE_SSDOM + 2]};
        int hi[3] = {i_lu * p[SIZE_SSDOM], j_lu * p[SIZE_SSDOM + 1], k_lu * p[SIZE_SSDOM + 2]};
        if (i_lu == ijkv_lu[0]) hi[0] = p[IJKV];
        if (j_lu == ijkv_lu[1]) hi[1] = p[IJKV + 1];
        if (k_lu == ijkv_lu[2]) hi[2] = p[IJKV + 2];
        double s2[5] = {0, 0, 0, 0, 0}, mx[5] = {0, 0, 0, 0, 0};
        for (long k = lo[2]; k <= hi[2]; k++)
            for (long j = lo[1]; j <= hi[1]; j++)
                for (long i = lo[0]; i <= hi[0]; i++) {
                    long l = IDX(Z, i, j, k);
                    for (int e = 0; e < 5; e++) {
                        double xp = fabs(drodm[l + e * nd]);
                        if (xp > mx[e]) mx[e] = xp;
                        s2[e] = s2[e] + drodm[l + e * nd] * drodm[l + e * nd];
                    }
                }
        for (int e = 0; e < neq; e++) {
            rdm[0 + (long)nssiter * (e + neq * (0 + 2 * (no - 1)))] = s2[e];
            rdm[0 + (long)nssiter * (e + neq * (1 + 2 * (no - 1)))] = mx[e];
        }
    }
}

/* ------------------------------------------------------------------------------------ */
/* Boundary conditions: bvbs_*.for + BC*_{i,j,k}.for bodies; indice_cl_sdm.for windows;
   BCzone.cpp dispatch; correct_coins.for corner/edge fill.                              */
typedef struct { long lo[3], hi[3]; } Box;

/* indice_cl_sdm.for for a thread box equal to the zone and npass = 0 */
static int cl_window(const int* p, int idir, const int* fen, int ind_cl[6])
{
    int sdm[6] = {1, p[IJKV], 1, p[IJKV + 1], 1, p[IJKV + 2]};
    return cl_window_sdm(p, idir, fen, sdm, ind_cl);
}

/* indice_cl_sdm.for with npass = 0: the BC window intersected with the box sdm, widened to
   the ghost layers in the normal direction; 1 if they do not intersect */
static int cl_window_sdm(const int* p, int idir, const int* fen, const int sdm[6], int ind_cl[6])
{
    int ific = p[NIJK + 3], kfic = p[NIJK + 4];
    if (!(sdm[0] <= fen[1] && sdm[1] >= fen[0] && sdm[2] <= fen[3] && sdm[3] >= fen[2] &&
          sdm[4] <= fen[5] && sdm[5] >= fen[4])) return 1;
    for (int a = 0; a < 3; a++) {
        ind_cl[2 * a] = fen[2 * a] > sdm[2 * a] ? fen[2 * a] : sdm[2 * a];
        ind_cl[2 * a + 1] = fen[2 * a + 1] < sdm[2 * a + 1] ? fen[2 * a + 1] : sdm[2 * a + 1];
    }
    switch (idir) {
    case 1: ind_cl[1] = ind_cl[0] - 1; ind_cl[0] = ind_cl[0] - ific; break;
    case 2: ind_cl[1] = ind_cl[0] + ific; ind_cl[0] = ind_cl[0] + 1; break;
    case 3: ind_cl[3] = ind_cl[2] - 1; ind_cl[2] = ind_cl[2] - ific; break;
    case 4: ind_cl[3] = ind_cl[2] + ific; ind_cl[2] = ind_cl[2] + 1; break;
    case 5: ind_cl[5] = ind_cl[4] - 1; ind_cl[4] = ind_cl[4] - kfic; break;
    default: ind_cl[5] = ind_cl[4] + kfic; ind_cl[4] = ind_cl[4] + 1; break;
    }
    return 0;
}

/* mirror / periodic partner of ghost (i,j,k) for face idir and window ind */
static long bc_partner(const Ix* Z, const int* p, int idir, const int* ind, long i, long j, long k,
                       int kind /* 0 extrapolate, 1 mirror, 2 periodic */)
{
    int d = (idir - 1) / 2;
    int up = (idir % 2) == 0;
    long c[3] = {i, j, k};
    long n = d == 0 ? p[IJKV] : d == 1 ? p[IJKV + 1] : p[IJKV + 2];
    if (kind == 0) c[d] = up ? ind[2 * d] - 1 : ind[2 * d + 1] + 1;
    else if (kind == 1) c[d] = up ? 2 * ind[2 * d] - 1 - c[d] : 2 * ind[2 * d + 1] + 1 - c[d];
    else c[d] = up ? c[d] - n : c[d] + n;
    return IDX(Z, c[0], c[1], c[2]);
}

static void bc_apply(const Ix* Z, const int* p, int bctype, int idir, int lrhs, const int* ind,
                     double* rop)
{
    long nd = Z->nd;
    int kind = bctype == 0 ? 0 : bctype == 11 ? 2 : 1;
    double c1 = lrhs == 1 ? 0. : 1.;
    for (long k = ind[4]; k <= ind[5]; k++)
        for (long j = ind[2]; j <= ind[3]; j++)
            for (long i = ind[0]; i <= ind[1]; i++) {
                long l = IDX(Z, i, j, k);
                long l1 = bc_partner(Z, p, idir, ind, i, j, k, kind);
                if (kind == 0) {          /* BCExtrapolate.for, vmax = -1e30 */
                    for (int e = 0; e < 5; e++) rop[l + e * nd] = fmax2(-1.e30, rop[l1 + e * nd]) * c1;
                } else if (kind == 2) {   /* BCPeriodic.for */
                    for (int e = 0; e < 5; e++) rop[l + e * nd] = rop[l1 + e * nd] * c1;
                } else {                  /* BCWallViscous_{i,j,k}.for, no mesh motion */
                    double u = rop[l1 + nd], v = rop[l1 + 2 * nd], w = rop[l1 + 3 * nd];
                    rop[l] = rop[l1];
                    rop[l + nd] = 0. - u;
                    rop[l + 2 * nd] = 0. - v;
                    rop[l + 3 * nd] = 0. - w;
                    rop[l + 4 * nd] = rop[l1 + 4 * nd];
                }
            }
}

/* shape_tab_mtr.for: which metric components a face BC reads */
static void shape_tab(int neq_mtr, int idir, int* ic, int* jc, int* kc, double* cim, double* cjm,
                      double* ckm)
{
    if (neq_mtr == 3) { *ic = 1; *jc = 2; *kc = 3; *cim = 1.; *cjm = 1.; *ckm = 1.; }
    else if (neq_mtr == 2) { *ic = 1; *jc = 2; *kc = 2; *cim = 1.; *cjm = 1.; *ckm = 0.; }
    else if (neq_mtr == 1) {
        *ic = 1; *jc = 1; *kc = 1;
        *cim = idir <= 2 ? 1. : 0.;
        *cjm = (idir == 3 || idir == 4) ? 1. : 0.;
        *ckm = idir >= 5 ? 1. : 0.;
    } else { *ic = 1; *jc = 1; *kc = 1; *cim = 0.; *cjm = 0.; *ckm = 0.; }
}

/* Characteristic (non-reflecting) state of the first ghost rank: FastS BC/non_reflection.for
   with the exterior state (roe..ete) and the interior cell l1; returns the primitive ghost
   state (BCFarfield.for / BCInflow_i.for epilogue) or, for BCOutflow, qn, sn and qvar5.      */
typedef struct { double svar1, svar2, svar3, svar4, qvar5, r, u, v, w, qn, sn; } NonRefl;

static NonRefl non_reflection(const double* rop, long nd, long l1, double roe, double rue,
                              double rve, double rwe, double ete, double tcx, double tcy,
                              double tcz, double gam, double gam1, double gam1_1, double cvinf)
{
    NonRefl o;
    double sn = sqrt(tcx * tcx + tcy * tcy + tcz * tcz);
    double roi = rop[l1];
    double rui = rop[l1 + nd] * roi, rvi = rop[l1 + 2 * nd] * roi, rwi = rop[l1 + 3 * nd] * roi;
    double eti = rop[l1 + 4 * nd] * cvinf * roi + .5 * (rui * rui + rvi * rvi + rwi * rwi) / roi;
    double qen = 0.;   /* entrainment velocity: no mesh motion */
    double qref1 = .5 * (roe + roi), qref2 = .5 * (rue + rui), qref3 = .5 * (rve + rvi);
    double qref4 = .5 * (rwe + rwi), qref5 = .5 * (ete + eti);
    double r = qref1, ri = 1. / r;
    double u = qref2 * ri, v = qref3 * ri, w = qref4 * ri;
    double p = gam1 * (qref5 - .5 * r * (u * u + v * v + w * w));
    double c = sqrt(gam * p * ri), ci = 1. / c, s_1 = 1. / sn;
    double qn = tcx * u + tcy * v + tcz * w - qen;
    double tnx, tny;
    if (sn == 0.) { tnx = 1.; tny = 0.; } else { tnx = tcx / sn; tny = tcy / sn; }
    double eigv1 = qn, eigv2 = qn, eigv3 = qn, eigv4 = qn + sn * c, eigv5 = qn - sn * c;
    double q[2][5] = {{roe, rue, rve, rwe, ete}, {roi, rui, rvi, rwi, eti}};
    double ch[2][5];
    for (int s = 0; s < 2; s++) {
        double q1 = q[s][0], q2 = q[s][1], q3 = q[s][2], q4 = q[s][3], q5 = q[s][4];
        double s1 = q1, s2 = (q2 - u * q1) * ri, s3 = (q3 - v * q1) * ri, s4 = (q4 - w * q1) * ri;
        double s5 = gam1 * (.5 * (u * u + v * v + w * w) * q1 - u * q2 - v * q3 - w * q4 + q5);
        ch[s][0] = s1 - ci * ci * s5;
        ch[s][1] = tcz * s_1 * (tnx * s2 + tny * s3) - s4;
        ch[s][2] = -tny * s2 + tnx * s3;
        ch[s][3] = .5 * s_1 * (tcx * s2 + tcy * s3 + tcz * s4) + .5 * ri * ci * s5;
        ch[s][4] = -.5 * s_1 * (tcx * s2 + tcy * s3 + tcz * s4) + .5 * ri * ci * s5;
    }
    double rv1 = ch[1][0], rv2 = ch[1][1], rv3 = ch[1][2], rv4 = ch[1][3], rv5 = ch[1][4];
    if (eigv1 < 0.) rv1 = ch[0][0];
    if (eigv2 < 0.) rv2 = ch[0][1];
    if (eigv3 < 0.) rv3 = ch[0][2];
    if (eigv4 < 0.) rv4 = ch[0][3];
    if (eigv5 < 0.) rv5 = ch[0][4];
    o.svar1 = rv1 + r * ci * (rv4 + rv5);
    o.svar2 = tcz * s_1 * tnx * rv2 - tny * rv3 + tcx * s_1 * (rv4 - rv5);
    o.svar3 = tcz * s_1 * tny * rv2 + tnx * rv3 + tcy * s_1 * (rv4 - rv5);
    o.svar4 = -rv2 + tcz * s_1 * (rv4 - rv5);
    double svar5 = r * c * (rv4 + rv5);
    o.qvar5 = .5 * (u * u + v * v + w * w) * o.svar1 + r * (u * o.svar2 + v * o.svar3 + w * o.svar4)
              + gam1_1 * svar5;
    o.r = r; o.u = u; o.v = v; o.w = w; o.qn = qn; o.sn = sn;
    return o;
}

/* Face BCs with a first-rank formula and an outer-rank rule, one ghost column at a time:
   bvbs_farfield.for / bvbs_inflow.for (BCFarfield.for, then BC_nextrank.for),
   bvbs_outflow.for (BCOutflow.for, then BC_extrap.for), bvbs_outpres.for
   (BCOutpres_{i,j,k}.for + linearization_state.for, then BC_extrap.for),
   bvbs_wall_inviscid.for (BCWallInviscid.for on every rank).
   state: the BC's conservative state (BCzone.cpp DEFAULT_STATE); pext: per-face-point
   pressure (outpres) indexed like indbci with inc_bc.                                      */
static void bc_column(const Ix* Z, const int* p, const double* pr, int t, int idir,
                      const int* ind, const double* tijk, int neq_mtr, const double* state,
                      const double* pext, const int* inc_bc, double* rop)
{
    long nd = Z->nd, ndm = Z->ndm;
    int d = (idir - 1) / 2, up = idir % 2 == 0;
    int ic, jc, kc;
    double cim, cjm, ckm;
    shape_tab(neq_mtr, idir, &ic, &jc, &kc, &cim, &cjm, &ckm);
    double snorm = up ? 1. : -1.;
    double gam = pr[GAMMA], gam1 = pr[GAMMA] - 1., gam1_1 = 1. / gam1;
    double rgp = pr[CVINF] * gam1, cv = pr[CVINF], cvinv = 1. / cv;
    double c4 = 5. / 6., c5 = 2. / 6., c6 = -1. / 6.;
    double c0 = 1. / c6, c1 = -(c4 + c5) * c0, c2 = -c6 * c0, c3 = (2. - c5 - c4) * c0;
    long inc = d == 0 ? 1 : d == 1 ? Z->incj : Z->inck;
    long sinc = up ? -inc : inc;                 /* towards the interior */
    int g0 = up ? ind[2 * d] : ind[2 * d + 1];   /* first ghost rank */
    int ta = d == 0 ? 1 : 0, tb = d == 2 ? 1 : 2;
    for (long b = ind[2 * tb]; b <= ind[2 * tb + 1]; b++)
        for (long a = ind[2 * ta]; a <= ind[2 * ta + 1]; a++) {
            long c[3];
            c[ta] = a; c[tb] = b;
            c[d] = up ? g0 : g0 + 1;
            long lmtr = MTR(Z, c[0], c[1], c[2]);
            double tx = tijk[lmtr + (ic - 1) * ndm], ty = tijk[lmtr + (jc - 1) * ndm];
            double tz = tijk[lmtr + (kc - 1) * ndm];
            if (t == 3) {
                /* BCWallInviscid.for: mirror of the interior, normal velocity reversed */
                double tcx = tx * cim, tcy = ty * cjm, tcz = tz * ckm;
                double surf = sqrt(tcx * tcx + tcy * tcy + tcz * tcz);
                surf = fmax2(surf, 1e-30);
                double s_1 = 1. / surf;
                tcx = tcx * s_1; tcy = tcy * s_1; tcz = tcz * s_1;
                for (long g = ind[2 * d]; g <= ind[2 * d + 1]; g++) {
                    c[d] = g;
                    long l = IDX(Z, c[0], c[1], c[2]);
                    long cm[3] = {c[0], c[1], c[2]};
                    cm[d] = up ? 2 * ind[2 * d] - 1 - g : 2 * ind[2 * d + 1] + 1 - g;
                    long ldjr = IDX(Z, cm[0], cm[1], cm[2]);
                    /* the mesh velocity term (vent x c_ale) is +0 without mesh motion */
                    double u = rop[ldjr + nd], v = rop[ldjr + 2 * nd], w = rop[ldjr + 3 * nd];
                    double qn = (u * tcx + v * tcy + w * tcz);
                    double ut = u - qn * tcx, vt = v - qn * tcy, wt = w - qn * tcz;
                    double ua = 2. * ut - u, va = 2. * vt - v, wa = 2. * wt - w;
                    rop[l] = rop[ldjr];
                    rop[l + nd] = 0. + ua;
                    rop[l + 2 * nd] = 0. + va;
                    rop[l + 3 * nd] = 0. + wa;
                    rop[l + 4 * nd] = rop[ldjr + 4 * nd];
                }
                continue;
            }
            double tcx = tx * cim * snorm, tcy = ty * cjm * snorm, tcz = tz * ckm * snorm;
            c[d] = g0;
            long l = IDX(Z, c[0], c[1], c[2]);
            long l1 = l + sinc;
            if (t == 1 || t == 13) {
                NonRefl o = non_reflection(rop, nd, l1, state[0], state[1], state[2], state[3],
                                           state[4], tcx, tcy, tcz, gam, gam1, gam1_1, pr[CVINF]);
                double roinv = 1. / o.svar1;
                double u = o.u + o.r * o.svar2 * roinv, v = o.v + o.r * o.svar3 * roinv;
                double w = o.w + o.r * o.svar4 * roinv;
                rop[l] = o.svar1; rop[l + nd] = u; rop[l + 2 * nd] = v; rop[l + 3 * nd] = w;
                rop[l + 4 * nd] = (o.qvar5 * roinv - .5 * (u * u + v * v + w * w)) * cvinv;
            } else if (t == 10) {
                double roe = rop[l1], rue = rop[l1 + nd] * roe, rve = rop[l1 + 2 * nd] * roe;
                double rwe = rop[l1 + 3 * nd] * roe;
                double roe_inv = 1. / state[0];
                double ete = state[4] - 0.5 * (state[1] * state[1] + state[2] * state[2] + state[3] * state[3]) * roe_inv
                             + 0.5 * (rue * rop[l1 + nd] + rve * rop[l1 + 2 * nd] + rwe * rop[l1 + 3 * nd]);
                NonRefl o = non_reflection(rop, nd, l1, roe, rue, rve, rwe, ete, tcx, tcy, tcz, gam,
                                           gam1, gam1_1, pr[CVINF]);
                rop[l] = rop[l1];
                double roinv = 1. / rop[l];
                if (o.qn < 0) {
                    rop[l + nd] = rop[l1 + nd] - o.qn / (o.sn * o.sn) * tcx;
                    rop[l + 2 * nd] = rop[l1 + 2 * nd] - o.qn / (o.sn * o.sn) * tcy;
                    rop[l + 3 * nd] = rop[l1 + 3 * nd] - o.qn / (o.sn * o.sn) * tcz;
                } else {
                    rop[l + nd] = rop[l1 + nd]; rop[l + 2 * nd] = rop[l1 + 2 * nd];
                    rop[l + 3 * nd] = rop[l1 + 3 * nd];
                }
                rop[l + 4 * nd] = (o.qvar5 * roinv - .5 * (rop[l + nd] * rop[l + nd] + rop[l + 2 * nd] * rop[l + 2 * nd]
                                                          + rop[l + 3 * nd] * rop[l + 3 * nd])) * cvinv;
            } else { /* 16: BCOutpres_{i,j,k}.for with linearization_state.for */
                long li = (a - inc_bc[1]) + (b - inc_bc[2]) * inc_bc[0];
                double sn = sqrt(tcx * tcx + tcy * tcy + tcz * tcz);
                double s_1 = 1. / sn, tnx, tny, tnz;
                if (sn == 0.) { tnx = 1.; tny = 0.; tnz = 0.; }
                else { tnx = tcx * s_1; tny = tcy * s_1; tnz = tcz * s_1; }
                double roi = rop[l1], ui = rop[l1 + nd], vi = rop[l1 + 2 * nd], wi = rop[l1 + 3 * nd];
                double Ti = rop[l1 + 4 * nd], pi = roi * rgp * Ti;
                double ro0 = rop[l1], T0 = rop[l1 + 4 * nd], p0 = ro0 * rgp * T0;
                double roc0 = sqrt(ro0 * gam * p0);
                double dqn = (pext[li] - pi) / roc0;
                double rog = roi + ro0 * ro0 * dqn / roc0;
                rop[l] = rog;
                rop[l + nd] = ui - dqn * tnx; rop[l + 2 * nd] = vi - dqn * tny; rop[l + 3 * nd] = wi - dqn * tnz;
                rop[l + 4 * nd] = pext[li] / (rog * rgp);
            }
            /* outer ranks in the reference loop order (ascending index) */
            int glo = up ? g0 + 1 : ind[2 * d], ghi = up ? ind[2 * d + 1] : g0 - 1;
            for (int g = glo; g <= ghi; g++) {
                c[d] = g;
                long lg = IDX(Z, c[0], c[1], c[2]);
                long q0 = lg + sinc, q1 = lg + 2 * sinc, q2 = lg + 3 * sinc;
                if (t == 1 || t == 13) {   /* BC_nextrank.for */
                    rop[lg] = rop[q0] * c3 + rop[q1] * c1 + rop[q2] * c2;
                    if (fabs(rop[lg] / rop[q0] - 1.) >= 0.7) {
                        for (int e = 0; e < 5; e++) rop[lg + e * nd] = rop[q0 + e * nd];
                    } else {
                        for (int e = 1; e < 5; e++)
                            rop[lg + e * nd] = rop[q0 + e * nd] * c3 + rop[q1 + e * nd] * c1 + rop[q2 + e * nd] * c2;
                    }
                } else {                   /* BC_extrap.for: copy of the first rank */
                    for (int e = 0; e < 5; e++) rop[lg + e * nd] = rop[l + e * nd];
                }
            }
        }
}

/* correct_coins.for */
static void fo_correct_coins(const int* p, const int* ind_loop, double* rop)
{
    Ix zz = mkix(p); Ix* Z = &zz;
    long nd = Z->nd;
    int nf = p[NIJK + 3], nfk = p[NIJK + 4];
    int L[6];
    L[0] = ind_loop[0] > 1 ? ind_loop[0] : 1;
    L[2] = ind_loop[2] > 1 ? ind_loop[2] : 1;
    L[4] = ind_loop[4] > 1 ? ind_loop[4] : 1;
    L[1] = ind_loop[1] < p[IJKV] ? ind_loop[1] : p[IJKV];
    L[3] = ind_loop[3] < p[IJKV + 1] ? ind_loop[3] : p[IJKV + 1];
    L[5] = ind_loop[5] < p[IJKV + 2] ? ind_loop[5] : p[IJKV + 2];
#define AVG(l1, l11, l12) rop[(l1) + e * nd] = (rop[(l11) + e * nd] + rop[(l12) + e * nd]) * 0.5
    int itest[2] = {1, p[IJKV]};
    for (int idir = 1; idir <= 2; idir++) {
        int is = idir == 1 ? -1 : 1;
        int ic = L[idir - 1];
        if (ic != itest[idir - 1]) continue;
        if (L[2] == 1)
            for (int e = 0; e < 5; e++) for (long k = L[4]; k <= L[5]; k++)
                for (long j = 1; j <= nf; j++) for (long i = 1; i <= nf; i++) {
                    long i1 = i + 1 < nf ? i + 1 : nf, j1 = j + 1 < nf ? j + 1 : nf;
                    AVG(IDX(Z, ic + i * is, L[2] - j, k), IDX(Z, ic + i1 * is, L[2], k), IDX(Z, ic, L[2] - j1, k));
                }
        if (L[3] == p[IJKV + 1])
            for (int e = 0; e < 5; e++) for (long k = L[4]; k <= L[5]; k++)
                for (long j = 1; j <= nf; j++) for (long i = 1; i <= nf; i++) {
                    long i1 = i + 1 < nf ? i + 1 : nf, j1 = j + 1 < nf ? j + 1 : nf;
                    AVG(IDX(Z, ic + i * is, L[3] + j, k), IDX(Z, ic + i1 * is, L[3], k), IDX(Z, ic, L[3] + j1, k));
                }
        if (L[4] == 1)
            for (int e = 0; e < 5; e++) for (long k = 1; k <= nfk; k++)
                for (long j = L[2]; j <= L[3]; j++) for (long i = 1; i <= nf; i++) {
                    long k1 = k + 1 < nfk ? k + 1 : nfk, i1 = i + 1 < nf ? i + 1 : nf;
                    AVG(IDX(Z, ic + i * is, j, L[4] - k), IDX(Z, ic, j, L[4] - k1), IDX(Z, ic + i1 * is, j, L[4]));
                }
        if (L[5] == p[IJKV + 2])
            for (int e = 0; e < 5; e++) for (long k = 1; k <= nfk; k++)
                for (long j = L[2]; j <= L[3]; j++) for (long i = 1; i <= nf; i++) {
                    long k1 = k + 1 < nfk ? k + 1 : nfk, i1 = i + 1 < nf ? i + 1 : nf;
                    AVG(IDX(Z, ic + i * is, j, L[5] + k), IDX(Z, ic, j, L[5] + k1), IDX(Z, ic + i1 * is, j, L[5]));
                }
    }
    int jtest[2] = {1, p[IJKV + 1]};
    for (int idir = 3; idir <= 4; idir++) {
        int is = idir == 3 ? -1 : 1;
        int jc = L[idir - 1];
        if (jc != jtest[idir - 3]) continue;
        if (L[4] == 1)
            for (int e = 0; e < 5; e++) for (long k = 1; k <= nfk; k++)
                for (long j = 1; j <= nf; j++) for (long i = L[0]; i <= L[1]; i++) {
                    long k1 = k + 1 < nfk ? k + 1 : nfk, j1 = j + 1 < nf ? j + 1 : nf;
                    AVG(IDX(Z, i, jc + j * is, L[4] - k), IDX(Z, i, jc, L[4] - k1), IDX(Z, i, jc + j1 * is, L[4]));
                }
        if (L[5] == p[IJKV + 2])
            for (int e = 0; e < 5; e++) for (long k = 1; k <= nfk; k++)
                for (long j = 1; j <= nf; j++) for (long i = L[0]; i <= L[1]; i++) {
                    long k1 = k + 1 < nfk ? k + 1 : nfk, j1 = j + 1 < nf ? j + 1 : nf;
                    AVG(IDX(Z, i, jc + j * is, L[5] + k), IDX(Z, i, jc, L[5] + k1), IDX(Z, i, jc + j1 * is, L[5]));
                }
    }
#undef AVG
}

/* correct_coins.for, the 8 corner blocks: mean of the three face-adjacent ghosts */
static void fo_correct_corners(const int* p, const int* ind_loop, double* rop)
{
    Ix zz = mkix(p); Ix* Z = &zz;
    long nd = Z->nd;
    int nf = p[NIJK + 3], nfk = p[NIJK + 4];
    int L[6];
    L[0] = ind_loop[0] > 1 ? ind_loop[0] : 1;
    L[2] = ind_loop[2] > 1 ? ind_loop[2] : 1;
    L[4] = ind_loop[4] > 1 ? ind_loop[4] : 1;
    L[1] = ind_loop[1] < p[IJKV] ? ind_loop[1] : p[IJKV];
    L[3] = ind_loop[3] < p[IJKV + 1] ? ind_loop[3] : p[IJKV + 1];
    L[5] = ind_loop[5] < p[IJKV + 2] ? ind_loop[5] : p[IJKV + 2];
    double c3 = 1. / 3.;
    /* order of the reference: (1,1,1) (iv,1,1) (1,jv,1) (iv,jv,1) then the same at kv */
    for (int kc = 0; kc < 2; kc++)
        for (int jc = 0; jc < 2; jc++)
            for (int ic = 0; ic < 2; ic++) {
                int i0 = ic ? L[1] : L[0], j0 = jc ? L[3] : L[2], k0 = kc ? L[5] : L[4];
                int ok = (ic ? i0 == p[IJKV] : i0 == 1) && (jc ? j0 == p[IJKV + 1] : j0 == 1) &&
                         (kc ? k0 == p[IJKV + 2] : k0 == 1);
                if (!ok) continue;
                int si = ic ? 1 : -1, sj = jc ? 1 : -1, sk = kc ? 1 : -1;
                for (int e = 0; e < 5; e++)
                    for (long k = 1; k <= nfk; k++)
                        for (long j = 1; j <= nf; j++)
                            for (long i = 1; i <= nf; i++) {
                                long i1 = i + 1 < nf ? i + 1 : nf, j1 = j + 1 < nf ? j + 1 : nf;
                                long k1 = k + 1 < nfk ? k + 1 : nfk;
                                long l = IDX(Z, i0 + si * i, j0 + sj * j, k0 + sk * k);
                                long l11 = IDX(Z, i0, j0, k0 + sk * k1);
                                long l12 = IDX(Z, i0 + si * i1, j0, k0);
                                long l13 = IDX(Z, i0, j0 + sj * j1, k0);
                                rop[l + e * nd] = (rop[l11 + e * nd] + rop[l12 + e * nd] + rop[l13 + e * nd]) * c3;
                            }
            }
}

static void correct_coins_all(const int* p, const int* ind_loop, double* rop)
{
    fo_correct_coins(p, ind_loop, rop);
    fo_correct_corners(p, ind_loop, rop);
}

/* BCzone.cpp:60-865 for one thread owning the zone (npass = 0).  Supported types:
   0 extrapolate, 1 farfield, 3/5 (and 4 in Euler) inviscid wall, 4/6 viscous adiabatic
   wall, 10 outflow, 11 autoperiod, 13 inflow (free-stream state), 16 outpres (pressure data).
   Returns -1 otherwise.                                                                    */

static int fo_bczone(int* p, double* pr, const double* ti, const double* tj, const double* tk,
                     int lrhs, int* shift_lu, double* rop, int* lcorner_out)
{
    Ix zz = mkix(p); Ix* Z = &zz;
    (void)pr;
    int ind_dm[6] = {1, p[IJKV], 1, p[IJKV + 1], 1, p[IJKV + 2]};
    int lskip[6] = {1, 1, 1, 1, 1, 1}, ificmax[6] = {0, 0, 0, 0, 0, 0};
    int lcorner = 1;
    int pt_bcs = p[PT_BC], nb_bc = pt_bcs > 0 ? p[pt_bcs] : 0;
    if (p[LU_MATCH] != 0) return -1;
    for (int n = 0; n < nb_bc; n++) {
        int pt = p[pt_bcs + 1 + n], idir = p[pt + BC_IDIR], t = p[pt + BC_TYPE];
        int ip = idir - 1;
        int itest = (ip % 2 == 0) ? 1 : ind_dm[ip];
        if (lrhs == 1 && ind_dm[ip] == itest && ((t >= 3 && t <= 7) || t == 12 || t == 18 || t == 30)) {
            lskip[ip] = 0; lcorner = 0;
            if ((t >= 3 && t <= 7) || t == 12) ificmax[ip] = 1;
        }
    }
    if (lrhs == 1) {
        for (int a = 0; a < 3; a++) {
            shift_lu[2 * a] = ind_dm[2 * a] - (1 - lskip[2 * a]) * ificmax[2 * a];
            shift_lu[2 * a + 1] = ind_dm[2 * a + 1] + (1 - lskip[2 * a + 1]) * ificmax[2 * a + 1];
        }
        for (int idir = 0; idir < 6; idir++) {
            if (lskip[idir] != 0) continue;
            int fen[6];
            for (int l = 0; l < 6; l++) fen[l] = ind_dm[l];
            if (idir % 2 == 0) { fen[idir] = 1; fen[idir + 1] = 1; } else fen[idir - 1] = fen[idir];
            int cl[6];
            if (cl_window(p, idir + 1, fen, cl)) continue;
            bc_apply(Z, p, 0, idir + 1, 1, cl, rop);
        }
    }
    for (int n = 0; n < nb_bc; n++) {
        int pt = p[pt_bcs + 1 + n], idir = p[pt + BC_IDIR], t = p[pt + BC_TYPE];
        if (!(lskip[idir - 1] == 0 || lrhs == 0)) continue;
        int cl[6];
        if (cl_window(p, idir, p + pt + BC_FEN, cl)) continue;
        if (lrhs == 1 && (t == 0 || t == 1 || t == 2 || t == 10 || t == 11 || t == 14 || t == 16 ||
                          t == 17 || t == 20 || t == 21)) t = 0;
        const double* tijk = idir <= 2 ? ti : idir <= 4 ? tj : tk;
        int neq_mtr = idir <= 4 ? p[NEQ_IJ] : p[NEQ_K];
        int nbdata = p[pt + BC_NBDATA];
        const double* data = nbdata ? pr + p[pt_bcs + 1 + n + nb_bc] : NULL;
        /* BCzone.cpp DEFAULT_STATE: free stream unless the BC carries data */
        double st[5];
        const double* state = st;
        if (nbdata == 0) {
            st[0] = pr[ROINF]; st[1] = pr[VXINF] * pr[ROINF]; st[2] = pr[VYINF] * pr[ROINF];
            st[3] = pr[VZINF] * pr[ROINF];
            st[4] = pr[PINF] / (pr[GAMMA] - 1.) + 0.5 * (pr[VXINF] * st[1] + pr[VYINF] * st[2] + pr[VZINF] * st[3]);
        } else state = data;
        if (t == 0 || t == 11) bc_apply(Z, p, t, idir, lrhs, cl, rop);
        else if (t == 1 || t == 10) bc_column(Z, p, pr, t, idir, cl, tijk, neq_mtr, state, NULL, NULL, rop);
        else if (t == 13) {
            if (nbdata != 0) return -1;           /* BCInflow from file data (bvbs_inflow_fich) */
            bc_column(Z, p, pr, 13, idir, cl, tijk, neq_mtr, state, NULL, NULL, rop);
        } else if (t == 3 || (t == 4 && p[IFLOW] == 1) || t == 5)
            bc_column(Z, p, pr, 3, idir, cl, tijk, neq_mtr, state, NULL, NULL, rop);
        else if ((t == 6 || t == 4) && p[IFLOW] > 1) bc_apply(Z, p, 6, idir, lrhs, cl, rop);
        else if (t == 6) { /* viscous wall in an Euler zone: BCzone.cpp has no branch for it */ }
        else if (t == 16) {
            /* BCOutpres: one pressure per face point (nbdata 1; a leading 1-value array is the
               mesh-motion coefficient and is skipped) */
            const int* fen = p + pt + BC_FEN;
            const int* size_data = p + pt + BC_NBDATA + 1;
            int inc_bc[3];
            if (idir <= 2) { inc_bc[0] = fen[3] - fen[2] + 1; inc_bc[1] = fen[2]; inc_bc[2] = fen[4]; }
            else if (idir <= 4) { inc_bc[0] = fen[1] - fen[0] + 1; inc_bc[1] = fen[0]; inc_bc[2] = fen[4]; }
            else { inc_bc[0] = fen[1] - fen[0] + 1; inc_bc[1] = fen[0]; inc_bc[2] = fen[2]; }
            if (nbdata == 0 || nbdata > 2) return -1;
            const double* pext = size_data[0] == 1 ? data + 1 : data;
            bc_column(Z, p, pr, 16, idir, cl, tijk, neq_mtr, state, pext, inc_bc, rop);
        } else return -1;
    }
    if (lcorner_out) *lcorner_out = lcorner;
    return 0;
}

int fo_apply_bc(int* p, double* pr, int nstep, double* rop, double* ti, double* tj, double* tk,
                const double* x, const double* y, const double* z, double* xmut)
{
    int shift[6];
    int ind_dm[6] = {1, p[IJKV], 1, p[IJKV + 1], 1, p[IJKV + 2]};
    (void)nstep; (void)x; (void)y; (void)z; (void)xmut;
    int rc = fo_bczone(p, pr, ti, tj, tk, 0, shift, rop, NULL);
    if (rc) return rc;
    correct_coins_all(p, ind_dm, rop);
    return 0;
}

/* ------------------------------------------------------------------------------------ */
/* One sub-iteration (gsdr3.cpp:152-444; navier_stokes_struct.for:216-496; lhs.cpp)     */
/* cpcfl3.for, no mesh motion (LALE = 0): local CFL number dt / min(viscous, convective time)
   of every interior cell; cfl[0] = max, cfl[1] = min, cfl[2] += sum.  cptst3.for:114-119 calls
   it once per cache block (over ind_sdm), so the sum runs block by block (kcache, jcache,
   icache from the outermost) and in the loop order k, j, i inside a block.
   3dcart uses the single face areas (ddi = 2|ti|), 3dfull the sum of the two face areas.     */
static void fo_cfl(const int* p, const double* pr, const double* rop, const double* xmut,
                   const double* ti, const double* tj, const double* tk, const double* vol,
                   double* cfl)
{
    Ix zz = mkix(p); Ix* Z = &zz;
    long nd = Z->nd, ndm = Z->ndm;
    double rgp = pr[CVINF] * (pr[GAMMA] - 1.);
    double gam1 = pr[GAMMA] / pr[PRANDT];
    double gam2 = pr[GAMMA] * rgp;
    double dt = pr[DTC];
    int lamin = p[IFLOW] >= 2, cart = p[ITYPZONE] == 2;
    double surf_c = 0.;
    if (cart) {
        double ddi = 2. * sqrt(ti[0] * ti[0]), ddj = 2. * sqrt(tj[0] * tj[0]), ddk = 2. * sqrt(tk[0] * tk[0]);
        surf_c = 1. / fmax2(fmax2(ddi, ddj), ddk);
    }
    int size[3] = {p[CACHEBLCKI], p[CACHEBLCKJ], p[CACHEBLCKK]}, nb[3];
    for (int d = 0; d < 3; d++) nb[d] = cache_nblk(p[IJKV + d], size[d]);
    for (int kc = 1; kc <= nb[2]; kc++)
    for (int jc = 1; jc <= nb[1]; jc++)
    for (int ic = 1; ic <= nb[0]; ic++) {
    int c[3] = {ic, jc, kc}, sdm[6];
    for (int d = 0; d < 3; d++) {
        sdm[2 * d] = 1 + (c[d] - 1) * size[d];
        sdm[2 * d + 1] = c[d] == nb[d] ? p[IJKV + d] : c[d] * size[d];
    }
    for (long k = sdm[4]; k <= sdm[5]; k++)
        for (long j = sdm[2]; j <= sdm[3]; j++)
            for (long i = sdm[0]; i <= sdm[1]; i++) {
                long l = IDX(Z, i, j, k);
                double surf, vo;
                if (cart) { surf = surf_c; vo = vol[0]; }
                else {
                    long lt = MTR(Z, i, j, k);
                    long li = lt + Z->mi, lj = lt + Z->mj, lk = lt + Z->mk;
                    double ddi = sqrt(ti[li] * ti[li] + ti[li + ndm] * ti[li + ndm] + ti[li + 2 * ndm] * ti[li + 2 * ndm]);
                    ddi = ddi + sqrt(ti[lt] * ti[lt] + ti[lt + ndm] * ti[lt + ndm] + ti[lt + 2 * ndm] * ti[lt + 2 * ndm]);
                    double ddj = sqrt(tj[lj] * tj[lj] + tj[lj + ndm] * tj[lj + ndm] + tj[lj + 2 * ndm] * tj[lj + 2 * ndm]);
                    ddj = ddj + sqrt(tj[lt] * tj[lt] + tj[lt + ndm] * tj[lt + ndm] + tj[lt + 2 * ndm] * tj[lt + 2 * ndm]);
                    double ddk = sqrt(tk[lk] * tk[lk] + tk[lk + ndm] * tk[lk + ndm] + tk[lk + 2 * ndm] * tk[lk + 2 * ndm]);
                    ddk = ddk + sqrt(tk[lt] * tk[lt] + tk[lt + ndm] * tk[lt + ndm] + tk[lt + 2 * ndm] * tk[lt + 2 * ndm]);
                    surf = 1. / fmax2(fmax2(ddi, ddj), ddk);
                    vo = vol[lt];
                }
                double u = rop[l + nd], v = rop[l + 2 * nd], w = rop[l + 3 * nd], r = rop[l];
                double c = sqrt(gam2 * rop[l + 4 * nd]);
                double ur2 = u * u + v * v + w * w;
                double sp = sqrt(ur2);
                double xinvspc = 1. / (sp + c);
                double conv = 2. * vo * surf * xinvspc;
                double cfloc;
                if (lamin) {
                    double xinvkt = gam1 / xmut[l];
                    double vis = 2. * r * vo * vo * surf * surf * xinvkt;
                    double comp = fmin2(vis, conv);
                    cfloc = dt / comp;
                } else cfloc = dt / conv;
                cfl[0] = fmax2(cfl[0], cfloc);
                cfl[1] = fmin2(cfl[1], cfloc);
                cfl[2] = cfl[2] + cfloc;
            }
    }
}

/* Kinetic energy and enstrophy sums (POST/post.for -> cp_gradu_3d{cart,full}.for with the
   4-point face values of grad_{i,j,k}_3d{cart,full}.for, c1 = 7/12, c2 = 1/12, then
   cptaylor.for).  Per gradient component the face contributions are accumulated in the order
   of the reference loops: -Fk(k), -Fj(j), -Fi(i), +Fi(i+1), +Fj(j+1), +Fk(k+1) (3dcart: only
   the face direction of the component).  out[0] = sum 1/2 rho |u|^2 vol, out[1] = sum
   1/2 rho |curl|^2 vol (Green-Gauss gradients divided by the volume, cp_gradu.for),
   out[2] = sum vol.  grad (optional, 9 x NDIMDX): the velocity gradients (du/dx, du/dy, ...).  */
static double face_u(const double* q, long l, long inc, double c1, double c2)
{
    return (q[l] + q[l - inc]) * c1 - (q[l + inc] + q[l - 2 * inc]) * c2;
}

void fo_enstrophy(const int* p, const double* pr, const double* rop, const double* ti,
                  const double* tj, const double* tk, const double* vol, double* out,
                  double* grad)
{
    (void)pr;
    Ix zz = mkix(p); Ix* Z = &zz;
    long nd = Z->nd, ndm = Z->ndm;
    double c1 = 7. / 12, c2 = 1. / 12;
    int cart = p[ITYPZONE] == 2;
    long inc[3] = {1, Z->incj, Z->inck};
    const double* tm[3] = {ti, tj, tk};
    long minc[3] = {Z->mi, Z->mj, Z->mk};
    double tke = 0., enst = 0., cpt = 0.;
    /* post.for:70-95 runs cp_gradu and cptaylor once per cache block (kcache, jcache, icache
       from the outermost; blocks as in fo_flux).  A face between two blocks belongs to the later
       block: a cell receives first the max faces its own block computes (i, j, k), then those of
       the next block in i, in j, in k (in the order the blocks run); the sums run block by
       block. */
    int size[3] = {p[CACHEBLCKI], p[CACHEBLCKJ], p[CACHEBLCKK]}, nb[3];
    for (int d = 0; d < 3; d++) nb[d] = cache_nblk(p[IJKV + d], size[d]);
    for (int kc = 1; kc <= nb[2]; kc++)
    for (int jc = 1; jc <= nb[1]; jc++)
    for (int ic = 1; ic <= nb[0]; ic++) {
    int cb[3] = {ic, jc, kc}, sdm[6];
    for (int d = 0; d < 3; d++) {
        sdm[2 * d] = 1 + (cb[d] - 1) * size[d];
        sdm[2 * d + 1] = cb[d] == nb[d] ? p[IJKV + d] : cb[d] * size[d];
    }
    for (long k = sdm[4]; k <= sdm[5]; k++)
        for (long j = sdm[2]; j <= sdm[3]; j++)
            for (long i = sdm[0]; i <= sdm[1]; i++) {
                long l = IDX(Z, i, j, k);
                long lt = cart ? 0 : MTR(Z, i, j, k);
                int nx[3] = {i == sdm[1] && ic < nb[0], j == sdm[3] && jc < nb[1], k == sdm[5] && kc < nb[2]};
                double g[3][3];      /* g[velocity][x,y,z] */
                for (int v = 0; v < 3; v++) {
                    const double* q = rop + (1 + v) * nd;
                    /* face values: low faces of this cell and high faces (low faces of the next) */
                    double flo[3], fhi[3];
                    for (int d = 0; d < 3; d++) {
                        flo[d] = face_u(q, l, inc[d], c1, c2);
                        fhi[d] = face_u(q, l + inc[d], inc[d], c1, c2);
                    }
                    for (int x = 0; x < 3; x++) {
                        if (cart) {
                            double t = tm[x][0];
                            g[v][x] = (0. - flo[x] * t) + fhi[x] * t;
                        } else {
                            double a = 0.;
                            a = a - flo[2] * tm[2][lt + x * ndm];
                            a = a - flo[1] * tm[1][lt + x * ndm];
                            a = a - flo[0] * tm[0][lt + x * ndm];
                            /* cp_gradu_3dfull.for:129-135: the i-face above the last
                               cell is added with the metric index of the previous loop
                               (the face below that cell), as the reference does */
                            long lhi[3] = {(i == p[IJKV]) ? lt : lt + minc[0], lt + minc[1], lt + minc[2]};
                            for (int pass = 0; pass < 2; pass++)
                                for (int d = 0; d < 3; d++)
                                    if (nx[d] == pass) a = a + fhi[d] * tm[d][lhi[d] + x * ndm];
                            g[v][x] = a;
                        }
                    }
                }
                double vo = cart ? vol[0] : vol[lt];
                /* cp_gradu.for: gradients divided by the cell volume */
                double volinv = 1. / vo;
                for (int v = 0; v < 3; v++)
                    for (int x = 0; x < 3; x++) {
                        g[v][x] = g[v][x] * volinv;
                        if (grad) grad[l + (3 * v + x) * nd] = g[v][x];
                    }
                cpt = cpt + vo;
                double rho = rop[l], u1 = rop[l + nd], u2 = rop[l + 2 * nd], u3 = rop[l + 3 * nd];
                tke = tke + 0.5 * rho * (u1 * u1 + u2 * u2 + u3 * u3) * vo;
                double w1 = (g[2][1] - g[1][2]), w2 = (g[0][2] - g[2][0]), w3 = (g[1][0] - g[0][1]);
                enst = enst + 0.5 * rho * (w1 * w1 + w2 * w2 + w3 * w3) * vo;
            }
    }
    out[0] = tke; out[1] = enst; out[2] = cpt;
}

int fo_substep(int* p, double* pr, int nitcfg, int nitrun, int first_it, int lssiter_verif,
               int lexit_lu, const double* x, const double* y, const double* z, double* ro,
               double* ro_m1, double* ro_p1, double* xmut, double* ti, double* tj, double* tk,
               double* vol, double* drodm, double* coe, double* wig, double* rdm, double* cfl,
               int stop)
{
    (void)nitrun; (void)x; (void)y; (void)z;
    int nf = p[NIJK + 3], nfk = p[NIJK + 4];
    int sdm[6] = {1, p[IJKV], 1, p[IJKV + 1], 1, p[IJKV + 2]};
    int ind_coe[6] = {1 - nf, p[IJKV] + nf, 1 - nf, p[IJKV + 1] + nf, 1 - nfk, p[IJKV + 2] + nfk};
    int ind_grad[6] = {0, p[IJKV] + 1, 0, p[IJKV + 1] + 1, 0, p[IJKV + 2] + 1};
    double *ro_ssiter, *ro_cl;
    if (nitcfg == 1) { ro_ssiter = ro; ro_cl = ro_p1; }
    else if (p[ITYPCP] < 2) { ro_ssiter = ro_p1; ro_cl = ro_p1; }
    else if (nitcfg % 2 == 0) { ro_ssiter = ro_p1; ro_cl = ro; }
    else { ro_ssiter = ro; ro_cl = ro_p1; }
    int implicit = p[ITYPCP] <= 1;
    if (nitcfg == 1) {
        if (p[IFLOW] == 2) fo_invist(p, pr, implicit ? ind_coe : ind_grad, ro_ssiter, xmut);
        if (p[DTLOC] == 1) {
            if (!implicit) return -1;  /* explicit local time step: outside the restatement */
            fo_tstb3(p, pr, first_it, 1, ind_coe, ro_ssiter, coe, xmut, ti, tj, tk, vol);
        } else if (implicit) {
            fo_tstb3_global(p, pr, ind_coe, coe, vol);
            fo_tstb3(p, pr, first_it, 0, ind_coe, ro_ssiter, coe, xmut, ti, tj, tk, vol);
        } else {
            fo_tstb3_global(p, pr, sdm, coe, vol);
        }
        /* cptst3.for:114-119 */
        if (lssiter_verif != 0 && cfl) fo_cfl(p, pr, ro_ssiter, xmut, ti, tj, tk, vol, cfl);
    }
    fo_init_rhs(p, nitcfg, sdm, drodm);
    if (p[KFLUDOM] == 2 && nitcfg == 1) fo_wiggle(p, pr, sdm, ro_ssiter, wig);
    if (fo_flux(p, pr, sdm, ro_ssiter, drodm, wig, ti, tj, tk, vol, xmut)) return -1;
    if (implicit) fo_core3as2(p, pr, nitcfg, first_it, sdm, ro_ssiter, ro, ro_m1, drodm, coe);
    else if (nitcfg % 2 == 1) fo_core3ark3(p, pr, nitcfg, sdm, ro_p1, ro, drodm, coe);
    else fo_core3ark3(p, pr, nitcfg, sdm, ro, ro_p1, drodm, coe);
    if (stop == 1) return 0;
    if (implicit) {
        int shift_lu[6], lcorner = 1;
        if (fo_bczone(p, pr, ti, tj, tk, 1, shift_lu, drodm, &lcorner)) return -1;
        if (lcorner == 0) correct_coins_all(p, shift_lu, drodm);
        if (lssiter_verif == 1 && rdm) fo_residual_blocks(p, nitcfg, p[NSSITER], drodm, rdm);
        if (stop == 2) return 0;
        if (lexit_lu == 0) fo_lu_sgs(p, pr, shift_lu, ro_p1, ro_ssiter, drodm, ti, tj, tk, coe, 1);
    }
    if (stop == 3) return 0;
    if (lexit_lu == 0) {
        int shift[6];
        if (fo_bczone(p, pr, ti, tj, tk, 0, shift, ro_cl, NULL)) return -1;
        correct_coins_all(p, sdm, ro_cl);
    }
    return 0;
}
