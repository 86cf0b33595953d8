/*
 * quick_oracle.c - TEST INFRASTRUCTURE ONLY.
 *
 * A plain-C, FP64, CPU restatement of QUICK's closed-shell direct-SCF
 * two-electron Fock build, written from QUICK-26.03:
 *
 *   Boys function FmT ............ src/subs/fmt.f90:13-85
 *   precomputables .............. src/modules/include/eri.fh:62-154
 *   Schwarz Ycutoff / cutprim ... src/modules/quick_cutoff_module.f90:54-588
 *   density cutoff .............. src/modules/include/cutoff.fh:18-92
 *   quartet loop + screening .... src/modules/include/eri.fh:161-202
 *   primitive loop (cshell) ..... src/modules/include/eri.fh:209-441
 *   contraction + HRR + digest .. src/modules/include/eri.fh:448-1057,
 *                                 src/subs/hrr.f90:7-857 (hrrwhole, lefthrr)
 *   VRR ......................... src/subs/hrr.f90:1051 (vertical) and
 *                                 src/subs/spdfgh.f90: generic Obara-Saika /
 *                                 Head-Gordon-Pople recursion restated here
 *   copySym ..................... src/subs/copySym.f90
 *
 * The product path never links this file; it is test infrastructure.
 *
 * Every function is pinned by a check in tests/test_oracle.py (closed forms,
 * an independent McMurchie-Davidson code, brute-force J/K, finite
 * differences, QUICK's saved SCF energies).
 */
#include "quick_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define QO_PI 3.14159265358979323846264338327950288
#define MAXAM 6            /* highest AM on one side: f+f */
#define NCUM(L) ((L) < 0 ? 0 : (((L) + 1) * ((L) + 2) * ((L) + 3) / 6))
#define NCMAX 84           /* NCUM(6) */
#define MMAX 13            /* m = 0..12 */

/* ------------------------------------------------------------------------- */
/* Cartesian component tables                                                */
/* ------------------------------------------------------------------------- */
static int g_comp[NCMAX][3];
static int g_idx[MAXAM + 1][MAXAM + 1][MAXAM + 1];
static int g_tables_ready = 0;

static void init_tables(void)
{
    if (g_tables_ready) return;
    int k = 0;
    for (int L = 0; L <= MAXAM; ++L)
        for (int l = L; l >= 0; --l)
            for (int m = L - l; m >= 0; --m) {
                int n = L - l - m;
                g_comp[k][0] = l; g_comp[k][1] = m; g_comp[k][2] = n;
                g_idx[l][m][n] = k;
                ++k;
            }
    g_tables_ready = 1;
}

/* QUICK's per-shell component order and cons (basis.f90:567-640). */
static const int q_comp_p[3][3] = {{1,0,0},{0,1,0},{0,0,1}};
static const int q_comp_d[6][3] = {{2,0,0},{1,1,0},{0,2,0},{1,0,1},{0,1,1},{0,0,2}};
static const int q_comp_f[10][3] = {{3,0,0},{2,1,0},{1,2,0},{0,3,0},{2,0,1},{1,1,1},{0,2,1},{1,0,2},{0,1,2},{0,0,3}};
/* cons of the f components (basis.f90:652-720) */
static double q_cons_f(int k)
{
    static const double c[10] = {1.0, 2.2360679774997896964, 2.2360679774997896964, 1.0, 2.2360679774997896964,
                                 3.8729833462074168852, 2.2360679774997896964, 2.2360679774997896964,
                                 2.2360679774997896964, 1.0};
    return c[k];
}

/* ------------------------------------------------------------------------- */
struct qo_ctx {
    int natom, nshell, nbasis, jbasis, maxcontract;
    double *xyz, *charge;
    int *katom, *qstart, *qfinal, *kprim, *kstart, *ksumtype;
    double *gcexpo, *gccoeff;
    /* per basis function */
    int *fn_lmn;        /* nbasis*3 */
    double *fn_cons;    /* nbasis */
    int *qsbasis;       /* nshell*4 : first function offset of sub-shell L */
    int *qfbasis;       /* nshell*4 : last function offset of sub-shell L */
    /* precomputables */
    double *Apri, *Kpri, *Ppri, *Xcoeff;
    double *Ycutoff, *cutprim;
    double X0, PITO3HALF;
};

/* ------------------------------------------------------------------------- */
/* Boys function: fmt.f90:13-85 (same piecewise fits, same branch limits).   */
/* The fit coefficients are those of QUICK's routine; parity with QUICK      */
/* requires them.                                                            */
/* ------------------------------------------------------------------------- */
/* Test switch: with it on, the small-X start value F_maxm comes from the
 * convergent series sum_k (-X)^k / (k! (2 maxm + 2k + 1)) instead of QUICK's
 * first-order (1 - X)/(2 maxm + 1).  Used only by the finite-difference
 * checks of the gradient, where energy and gradient evaluate FmT with
 * different maxm (NABCD and NABCD + 2) and QUICK's start value makes them
 * disagree by ~1e-6. */
static int g_fmt_series = 0;

void qo_set_fmt_series(int on) { g_fmt_series = on; }

void qo_fmt(int maxm, double X, double *vals)
{
    const double PIE4 = QO_PI / 4.0;
    const double XINV = 1.0 / X;
    const double E = exp(-X);
    double WW1, F1, Y;

    if (X > 5.0) {
        if (X > 15.0) {
            if (X > 33.0) {
                WW1 = sqrt(PIE4 * XINV);
            } else {
                WW1 = ((1.9623264149430e-01 * XINV - 4.9695241464490e-01) * XINV -
                       6.0156581186481e-05) * E + sqrt(PIE4 * XINV);
            }
        } else if (X > 10.0) {
            WW1 = (((-1.8784686463512e-01 * XINV + 2.2991849164985e-01) * XINV
                    - 4.9893752514047e-01) * XINV - 2.1916512131607e-05) * E
                  + sqrt(PIE4 * XINV);
        } else {
            WW1 = ((((((4.6897511375022e-01 * XINV - 6.9955602298985e-01) * XINV +
                       5.3689283271887e-01) * XINV - 3.2883030418398e-01) * XINV +
                     2.4645596956002e-01) * XINV - 4.9984072848436e-01) * XINV -
                   3.1501078774085e-06) * E + sqrt(PIE4 * XINV);
        }
    } else if (X > 1.0) {
        if (X > 3.0) {
            Y = X - 4.0;
            F1 = ((((((((((-2.62453564772299e-11 * Y + 3.24031041623823e-10) * Y -
                      3.614965656163e-09) * Y + 3.760256799971e-08) * Y -
                      3.553558319675e-07) * Y + 3.022556449731e-06) * Y -
                      2.290098979647e-05) * Y + 1.526537461148e-04) * Y -
                      8.81947375894379e-04) * Y + 4.33207949514611e-03) * Y -
                      1.75257821619926e-02) * Y + 5.28406320615584e-02;
            WW1 = (X + X) * F1 + E;
        } else {
            Y = X - 2.0;
            F1 = ((((((((((-1.61702782425558e-10 * Y + 1.96215250865776e-09) * Y -
                      2.14234468198419e-08) * Y + 2.17216556336318e-07) * Y -
                      1.98850171329371e-06) * Y + 1.62429321438911e-05) * Y -
                      1.16740298039895e-04) * Y + 7.24888732052332e-04) * Y -
                      3.79490003707156e-03) * Y + 1.61723488664661e-02) * Y -
                      5.29428148329736e-02) * Y + 1.15702180856167e-01;
            WW1 = (X + X) * F1 + E;
        }
    } else if (X > 1.0e-01 || (X > 1.0e-04 && maxm < 4)) {
        F1 = ((((((((-8.36313918003957e-08 * X + 1.21222603512827e-06) * X -
                 1.15662609053481e-05) * X + 9.25197374512647e-05) * X -
                 6.40994113129432e-04) * X + 3.78787044215009e-03) * X -
                 1.85185172458485e-02) * X + 7.14285713298222e-02) * X -
                 1.99999999997023e-01) * X + 3.33333333333318e-01;
        WW1 = (X + X) * F1 + E;
    } else {
        if (g_fmt_series) {
            double t = 1.0, acc = 0.0;
            for (int k = 0; k < 30; ++k) {
                acc += t / (double)(2 * maxm + 2 * k + 1);
                t *= -X / (double)(k + 1);
            }
            WW1 = acc;
        } else {
            WW1 = (1.0 - X) / (double)(2 * maxm + 1);
        }
    }

    if (X > 1.0e-1 || (X > 1.0e-04 && maxm < 4)) {
        vals[0] = WW1;
        for (int m = 1; m <= maxm; ++m)
            vals[m] = (((2 * m - 1) * vals[m - 1]) - E) * 0.5 * XINV;
    } else {
        vals[maxm] = WW1;
        const double twot = X * 2.0;
        for (int m = maxm - 1; m >= 0; --m)
            vals[m] = (twot * vals[m + 1] + E) / (double)(m * 2 + 1);
    }
}

/* ------------------------------------------------------------------------- */
/* Vertical recursion for one primitive quartet.                             */
/* Builds v[e][f] (m = 0) for every Cartesian e with AM <= E and f with      */
/* AM <= F, starting from [00|00]^(m) = FM(m)/sqrt(zeta+eta) as QUICK's      */
/* Yxiaotemp(1,1,m) (eri.fh:398-405).  Recurrences (HGP eqs. 6 and 16):      */
/*  [e+1i|0]^m = PA_i[e|0]^m + WP_i[e|0]^m+1 + e_i/2z([e-1i|0]^m - r/z[e-1i|0]^m+1) */
/*  [e|f+1i]^m = QC_i[e|f]^m + WQ_i[e|f]^m+1 + f_i/2n([e|f-1i]^m - r/n[e|f-1i]^m+1) */
/*               + e_i/2(z+n) [e-1i|f]^m+1                                    */
/* ------------------------------------------------------------------------- */
typedef struct {
    double PA[3], WP[3], QC[3], WQ[3];
    double oo2z, oo2n, oo2zn, rz, rn;  /* 1/2z, 1/2n, 1/2(z+n), rho/z, rho/n */
} vrr_par;

static void vrr_full(int E, int F, const double *fm, double prefac, const vrr_par *q,
                     double v[NCMAX][NCMAX][MMAX])
{
    const int L = E + F;
    for (int m = 0; m <= L; ++m) v[0][0][m] = fm[m] * prefac;
    /* bra build, f = 0 */
    for (int e = 1; e < NCUM(E); ++e) {
        const int *c = g_comp[e];
        int ae = c[0] + c[1] + c[2];
        int i = c[0] ? 0 : (c[1] ? 1 : 2);
        int lm[3] = {c[0], c[1], c[2]};
        lm[i] -= 1;
        int e1 = g_idx[lm[0]][lm[1]][lm[2]];
        int has2 = lm[i] > 0;
        int e2 = 0;
        if (has2) { lm[i] -= 1; e2 = g_idx[lm[0]][lm[1]][lm[2]]; lm[i] += 1; }
        double fac = (double)lm[i] * q->oo2z;
        for (int m = 0; m <= L - ae; ++m) {
            double t = q->PA[i] * v[e1][0][m] + q->WP[i] * v[e1][0][m + 1];
            if (has2) t += fac * (v[e2][0][m] - q->rz * v[e2][0][m + 1]);
            v[e][0][m] = t;
        }
    }
    /* ket build, every e */
    for (int f = 1; f < NCUM(F); ++f) {
        const int *c = g_comp[f];
        int af = c[0] + c[1] + c[2];
        int i = c[0] ? 0 : (c[1] ? 1 : 2);
        int lm[3] = {c[0], c[1], c[2]};
        lm[i] -= 1;
        int f1 = g_idx[lm[0]][lm[1]][lm[2]];
        int hasf2 = lm[i] > 0;
        int f2 = 0;
        if (hasf2) { lm[i] -= 1; f2 = g_idx[lm[0]][lm[1]][lm[2]]; lm[i] += 1; }
        double ffac = (double)lm[i] * q->oo2n;
        for (int e = 0; e < NCUM(E); ++e) {
            const int *ce = g_comp[e];
            int ae = ce[0] + ce[1] + ce[2];
            int ei = ce[i];
            int e1 = 0;
            if (ei > 0) {
                int le[3] = {ce[0], ce[1], ce[2]};
                le[i] -= 1;
                e1 = g_idx[le[0]][le[1]][le[2]];
            }
            double efac = (double)ei * q->oo2zn;
            for (int m = 0; m <= L - ae - af; ++m) {
                double t = q->QC[i] * v[e][f1][m] + q->WQ[i] * v[e][f1][m + 1];
                if (hasf2) t += ffac * (v[e][f2][m] - q->rn * v[e][f2][m + 1]);
                if (ei > 0) t += efac * v[e1][f1][m + 1];
                v[e][f][m] = t;
            }
        }
    }
}

/* Horizontal recursion in closed form, as lefthrr (hrr.f90:289-857):      */
/* (a b| = sum_{t<=b} prod_d C(b_d,t_d) (A_d-B_d)^(b_d-t_d) (a+t 0|.        */
static int hrr_terms(const int *a, const int *b, const double *AB, double *coef, int *ang)
{
    static const double binom[4][4] = {{1, 0, 0, 0}, {1, 1, 0, 0}, {1, 2, 1, 0}, {1, 3, 3, 1}};
    int n = 0;
    for (int tx = 0; tx <= b[0]; ++tx)
        for (int ty = 0; ty <= b[1]; ++ty)
            for (int tz = 0; tz <= b[2]; ++tz) {
                double c = binom[b[0]][tx] * binom[b[1]][ty] * binom[b[2]][tz];
                for (int k = 0; k < b[0] - tx; ++k) c *= AB[0];
                for (int k = 0; k < b[1] - ty; ++k) c *= AB[1];
                for (int k = 0; k < b[2] - tz; ++k) c *= AB[2];
                coef[n] = c;
                ang[n] = g_idx[a[0] + tx][a[1] + ty][a[2] + tz];
                ++n;
            }
    return n;
}

/* ------------------------------------------------------------------------- */
/* Context: basis copy, per-function tables, precomputables, Schwarz         */
/* ------------------------------------------------------------------------- */
#define XC(c, na, nb, I, J) ((c)->Xcoeff[((size_t)(na) * (c)->jbasis + (nb)) * 16 + (I) * 4 + (J)])

static void *dup(const void *p, size_t n)
{
    void *q = malloc(n ? n : 1);
    if (n) memcpy(q, p, n);
    return q;
}

/* get_eri_precomputables (eri.fh:62-154) */
static void precompute(qo_ctx *c)
{
    const int jb = c->jbasis;
    for (int ics = 0; ics < c->nshell; ++ics)
        for (int ips = 0; ips < c->kprim[ics]; ++ips)
            for (int jcs = 0; jcs < c->nshell; ++jcs)
                for (int jps = 0; jps < c->kprim[jcs]; ++jps) {
                    int NA = c->kstart[ics] + ips;
                    int NB = c->kstart[jcs] + jps;
                    double AA = c->gcexpo[NA], BB = c->gcexpo[NB];
                    c->Apri[NA * jb + NB] = AA + BB;
                    const double *A = c->xyz + 3 * c->katom[ics];
                    const double *B = c->xyz + 3 * c->katom[jcs];
                    double DAB = sqrt((A[0] - B[0]) * (A[0] - B[0]) + (A[1] - B[1]) * (A[1] - B[1]) +
                                      (A[2] - B[2]) * (A[2] - B[2]));
                    for (int i = 0; i < 3; ++i)
                        c->Ppri[(NA * jb + NB) * 3 + i] = (A[i] * AA + B[i] * BB) / (AA + BB);
                    c->Kpri[NA * jb + NB] = exp(-AA * BB / (AA + BB) * (DAB * DAB)) / (AA + BB);
                    for (int I = c->qstart[ics]; I <= c->qfinal[ics]; ++I)
                        for (int J = c->qstart[jcs]; J <= c->qfinal[jcs]; ++J)
                            XC(c, NA, NB, I, J) = c->Kpri[NA * jb + NB] *
                                                  c->gccoeff[NA * 4 + I] * c->gccoeff[NB * 4 + J];
                }
}

/* Primitive-quartet parameters of eri.fh:262-390 */
static void prim_par(const qo_ctx *c, int Nii, int Njj, int Nkk, int Nll,
                     const double *RA, const double *RC, vrr_par *q, double *T, double *prefac)
{
    const int jb = c->jbasis;
    double AB = c->Apri[Nii * jb + Njj];
    double CD = c->Apri[Nkk * jb + Nll];
    const double *P = c->Ppri + (Nii * jb + Njj) * 3;
    const double *Q = c->Ppri + (Nkk * jb + Nll) * 3;
    double ABCD = AB + CD;
    double ROU = AB * CD / ABCD;
    double RPQ = 0.0;
    for (int M = 0; M < 3; ++M) {
        double W = (P[M] * AB + Q[M] * CD) / ABCD;
        double d = P[M] - Q[M];
        RPQ += d * d;
        q->PA[M] = P[M] - RA[M];
        q->QC[M] = Q[M] - RC[M];
        q->WQ[M] = W - Q[M];
        q->WP[M] = W - P[M];
    }
    q->oo2z = 0.5 / AB;
    q->oo2n = 0.5 / CD;
    q->oo2zn = 0.5 / ABCD;
    q->rz = ROU / AB;     /* ABcom-style rho/zeta = CD/ABCD */
    q->rn = ROU / CD;
    *T = RPQ * ROU;
    *prefac = 1.0 / sqrt(ABCD);
}

/* Sub-shell function ranges (QUICK's Qsbasis/Qfbasis, basis.f90:400-560). */
static void build_function_tables(qo_ctx *c)
{
    c->fn_lmn = calloc((size_t)c->nbasis * 3, sizeof(int));
    c->fn_cons = calloc((size_t)c->nbasis, sizeof(double));
    c->qsbasis = calloc((size_t)c->nshell * 4, sizeof(int));
    c->qfbasis = calloc((size_t)c->nshell * 4, sizeof(int));
    for (int s = 0; s < c->nshell; ++s) {
        int off = 0;
        for (int L = c->qstart[s]; L <= c->qfinal[s]; ++L) {
            int nc = (L + 1) * (L + 2) / 2;
            c->qsbasis[s * 4 + L] = off;
            c->qfbasis[s * 4 + L] = off + nc - 1;
            for (int k = 0; k < nc; ++k) {
                int f = c->ksumtype[s] + off + k;
                const int *lmn = L == 0 ? (const int[3]){0, 0, 0}
                                 : (L == 1 ? q_comp_p[k] : (L == 2 ? q_comp_d[k] : q_comp_f[k]));
                c->fn_lmn[f * 3 + 0] = lmn[0];
                c->fn_lmn[f * 3 + 1] = lmn[1];
                c->fn_lmn[f * 3 + 2] = lmn[2];
                c->fn_cons[f] = L == 3 ? q_cons_f(k) : ((L == 2 && (k == 1 || k == 3 || k == 4)) ? sqrt(3.0) : 1.0);
            }
            off += nc;
        }
    }
}

/* Contracted shell-quartet store for one sub-class (I,J,K,L):              */
/* store[a][c] = sum_prim X0*Xcoeff(ij,I,J)*Xcoeff(kl,K,L) * Yxiao[prim][a][c] */
typedef struct {
    int nprim;             /* surviving primitive quartets */
    int E, F;
    int *pidx;             /* nprim*4: Nii, Njj, Nkk, Nll */
    double *Y;             /* nprim * NCUM(E) * NCUM(F) */
} prim_store;

static void shell_prims(const qo_ctx *c, int II, int JJ, int KK, int LL, double dnmax,
                        double prim_limit, int screen, prim_store *ps, double v[NCMAX][NCMAX][MMAX])
{
    const int jb = c->jbasis;
    const double *RA = c->xyz + 3 * c->katom[II];
    const double *RC = c->xyz + 3 * c->katom[KK];
    int E = c->qfinal[II] + c->qfinal[JJ];
    int F = c->qfinal[KK] + c->qfinal[LL];
    int ne = NCUM(E), nf = NCUM(F);
    double fm[MMAX + 5];
    ps->E = E;
    ps->F = F;
    ps->nprim = 0;
    /* loop order of cshell: JJJ, III, LLL, KKK (eri.fh:262-418) */
    for (int jjj = 0; jjj < c->kprim[JJ]; ++jjj) {
        int Njj = c->kstart[JJ] + jjj;
        for (int iii = 0; iii < c->kprim[II]; ++iii) {
            int Nii = c->kstart[II] + iii;
            double cutoffprim1 = dnmax * c->cutprim[Nii * jb + Njj];
            for (int lll = 0; lll < c->kprim[LL]; ++lll) {
                int Nll = c->kstart[LL] + lll;
                for (int kkk = 0; kkk < c->kprim[KK]; ++kkk) {
                    int Nkk = c->kstart[KK] + kkk;
                    double cutoffprim = cutoffprim1 * c->cutprim[Nkk * jb + Nll];
                    if (screen && !(cutoffprim > prim_limit)) continue;
                    vrr_par q;
                    double T, prefac;
                    prim_par(c, Nii, Njj, Nkk, Nll, RA, RC, &q, &T, &prefac);
                    qo_fmt(E + F, T, fm);
                    vrr_full(E, F, fm, prefac, &q, v);
                    int n = ps->nprim++;
                    ps->pidx[n * 4 + 0] = Nii;
                    ps->pidx[n * 4 + 1] = Njj;
                    ps->pidx[n * 4 + 2] = Nkk;
                    ps->pidx[n * 4 + 3] = Nll;
                    double *dst = ps->Y + (size_t)n * ne * nf;
                    for (int a = 0; a < ne; ++a)
                        for (int f = 0; f < nf; ++f) dst[a * nf + f] = v[a][f][0];
                }
            }
        }
    }
}

static void contract(const qo_ctx *c, const prim_store *ps, int I, int J, int K, int L,
                     double store[NCMAX][NCMAX])
{
    int ne = NCUM(ps->E), nf = NCUM(ps->F);
    int a0 = NCUM(I - 1), a1 = NCUM(I + J);
    int c0 = NCUM(K - 1), c1 = NCUM(K + L);
    for (int a = a0; a < a1; ++a)
        for (int f = c0; f < c1; ++f) store[a][f] = 0.0;
    for (int n = 0; n < ps->nprim; ++n) {
        const int *p = ps->pidx + n * 4;
        double X2 = c->X0 * XC(c, p[0], p[1], I, J);
        double X44 = X2 * XC(c, p[2], p[3], K, L);
        const double *src = ps->Y + (size_t)n * ne * nf;
        for (int a = a0; a < a1; ++a)
            for (int f = c0; f < c1; ++f) store[a][f] += X44 * src[a * nf + f];
    }
}

static double hrr_whole(const qo_ctx *c, int III, int JJJ, int KKK, int LLL, const double *AB,
                        const double *CD, double store[NCMAX][NCMAX])
{
    double cl[64], cr[64];
    int al[64], ar[64];
    int nl = hrr_terms(c->fn_lmn + 3 * III, c->fn_lmn + 3 * JJJ, AB, cl, al);
    int nr = hrr_terms(c->fn_lmn + 3 * KKK, c->fn_lmn + 3 * LLL, CD, cr, ar);
    double Y = 0.0;
    for (int i = 0; i < nl; ++i)
        for (int j = 0; j < nr; ++j) Y += cl[i] * cr[j] * store[al[i]][ar[j]];
    return Y * c->fn_cons[III] * c->fn_cons[JJJ] * c->fn_cons[KKK] * c->fn_cons[LLL];
}

/* shellcutoff / classcutoff / classprim (quick_cutoff_module.f90:91-588). */
static void schwarzoff(qo_ctx *c)
{
    const int ns = c->nshell, jb = c->jbasis;
    size_t ystride = 0;
    int mc = c->maxcontract;
    prim_store ps;
    ps.pidx = malloc(sizeof(int) * 4 * (size_t)mc * mc * mc * mc);
    ps.Y = malloc(sizeof(double) * (size_t)mc * mc * mc * mc * NCMAX * NCMAX);
    (void)ystride;
    double (*v)[NCMAX][MMAX] = malloc(sizeof(double) * NCMAX * NCMAX * MMAX);
    double (*store)[NCMAX] = malloc(sizeof(double) * NCMAX * NCMAX);
    for (int II = 0; II < ns; ++II)
        for (int JJ = II; JJ < ns; ++JJ) {
            const double *RA = c->xyz + 3 * c->katom[II];
            const double *RB = c->xyz + 3 * c->katom[JJ];
            double AB[3] = {RA[0] - RB[0], RA[1] - RB[1], RA[2] - RB[2]};
            /* all primitive quartets (ij|ij), KK=II, LL=JJ, no screening */
            shell_prims(c, II, JJ, II, JJ, 0.0, 0.0, 0, &ps, v);
            int ne = NCUM(ps.E), nf = NCUM(ps.F);
            double Ymax = 0.0;
            /* classcutoff: contracted (ij|ij) */
            for (int I = c->qstart[II]; I <= c->qfinal[II]; ++I)
                for (int J = c->qstart[JJ]; J <= c->qfinal[JJ]; ++J) {
                    contract(c, &ps, I, J, I, J, store);
                    for (int III = c->ksumtype[II] + c->qsbasis[II * 4 + I];
                         III <= c->ksumtype[II] + c->qfbasis[II * 4 + I]; ++III)
                        for (int JJJ = c->ksumtype[JJ] + c->qsbasis[JJ * 4 + J];
                             JJJ <= c->ksumtype[JJ] + c->qfbasis[JJ * 4 + J]; ++JJJ) {
                            double Y = hrr_whole(c, III, JJJ, III, JJJ, AB, AB, store);
                            if (fabs(Y) > Ymax) Ymax = fabs(Y);
                        }
                }
            c->Ycutoff[II * ns + JJ] = sqrt(Ymax);
            c->Ycutoff[JJ * ns + II] = sqrt(Ymax);
            /* classprim: diagonal primitive quartet (ij|ij) with the QUICK
             * weighting X4444^2 = (X0*Xcoeff)^2 (quick_cutoff_module.f90:456-467) */
            for (int iii = 0; iii < c->kprim[II]; ++iii)
                for (int jjj = 0; jjj < c->kprim[JJ]; ++jjj) {
                    int Nii = c->kstart[II] + iii, Njj = c->kstart[JJ] + jjj;
                    int n;
                    for (n = 0; n < ps.nprim; ++n) {
                        const int *p = ps.pidx + n * 4;
                        if (p[0] == Nii && p[1] == Njj && p[2] == Nii && p[3] == Njj) break;
                    }
                    const double *src = ps.Y + (size_t)n * ne * nf;
                    double Ymaxprim = 0.0;
                    for (int I = c->qstart[II]; I <= c->qfinal[II]; ++I)
                        for (int J = c->qstart[JJ]; J <= c->qfinal[JJ]; ++J) {
                            double X4444 = c->X0 * XC(c, Nii, Njj, I, J);
                            double w = pow(X4444, 2.0);
                            int a0 = NCUM(I - 1), a1 = NCUM(I + J);
                            for (int a = a0; a < a1; ++a)
                                for (int f = a0; f < a1; ++f) store[a][f] = src[a * nf + f] * w;
                            for (int III = c->ksumtype[II] + c->qsbasis[II * 4 + I];
                                 III <= c->ksumtype[II] + c->qfbasis[II * 4 + I]; ++III)
                                for (int JJJ = c->ksumtype[JJ] + c->qsbasis[JJ * 4 + J];
                                     JJJ <= c->ksumtype[JJ] + c->qfbasis[JJ * 4 + J]; ++JJJ) {
                                    double Y = hrr_whole(c, III, JJJ, III, JJJ, AB, AB, store);
                                    if (fabs(Y) > Ymaxprim) Ymaxprim = fabs(Y);
                                }
                        }
                    c->cutprim[Nii * jb + Njj] = sqrt(Ymaxprim);
                }
        }
    free(ps.pidx);
    free(ps.Y);
    free(v);
    free(store);
}

qo_ctx *qo_create(const qo_basis *b)
{
    init_tables();
    qo_ctx *c = calloc(1, sizeof(qo_ctx));
    c->natom = b->natom;
    c->nshell = b->nshell;
    c->nbasis = b->nbasis;
    c->jbasis = b->jbasis;
    c->xyz = dup(b->xyz, sizeof(double) * 3 * b->natom);
    c->charge = dup(b->charge, sizeof(double) * b->natom);
    c->katom = dup(b->katom, sizeof(int) * b->nshell);
    c->qstart = dup(b->qstart, sizeof(int) * b->nshell);
    c->qfinal = dup(b->qfinal, sizeof(int) * b->nshell);
    c->kprim = dup(b->kprim, sizeof(int) * b->nshell);
    c->kstart = dup(b->kstart, sizeof(int) * b->nshell);
    c->ksumtype = dup(b->ksumtype, sizeof(int) * b->nshell);
    c->gcexpo = dup(b->gcexpo, sizeof(double) * b->jbasis);
    c->gccoeff = dup(b->gccoeff, sizeof(double) * 4 * b->jbasis);
    c->maxcontract = 1;
    for (int s = 0; s < c->nshell; ++s)
        if (c->kprim[s] > c->maxcontract) c->maxcontract = c->kprim[s];
    c->X0 = 2.0 * pow(QO_PI, 2.5);
    c->PITO3HALF = pow(QO_PI, 1.5);
    build_function_tables(c);
    size_t jb2 = (size_t)c->jbasis * c->jbasis;
    c->Apri = calloc(jb2, sizeof(double));
    c->Kpri = calloc(jb2, sizeof(double));
    c->Ppri = calloc(jb2 * 3, sizeof(double));
    c->Xcoeff = calloc(jb2 * 16, sizeof(double));
    c->cutprim = calloc(jb2, sizeof(double));
    c->Ycutoff = calloc((size_t)c->nshell * c->nshell, sizeof(double));
    precompute(c);
    schwarzoff(c);
    return c;
}

void qo_destroy(qo_ctx *c)
{
    if (!c) return;
    free(c->xyz); free(c->charge); free(c->katom); free(c->qstart); free(c->qfinal);
    free(c->kprim); free(c->kstart); free(c->ksumtype); free(c->gcexpo); free(c->gccoeff);
    free(c->fn_lmn); free(c->fn_cons); free(c->qsbasis); free(c->qfbasis);
    free(c->Apri); free(c->Kpri); free(c->Ppri); free(c->Xcoeff);
    free(c->Ycutoff); free(c->cutprim);
    free(c);
}

void qo_get_schwarz(const qo_ctx *c, double *ycutoff, double *cutprim)
{
    if (ycutoff) memcpy(ycutoff, c->Ycutoff, sizeof(double) * c->nshell * c->nshell);
    if (cutprim) memcpy(cutprim, c->cutprim, sizeof(double) * c->jbasis * c->jbasis);
}

/* cshell_dnscreen / cshell_density_cutoff (cutoff.fh:18-92) */
void qo_density_cutoff(const qo_ctx *c, const double *P, double *cutmatrix)
{
    const int n = c->nbasis, ns = c->nshell;
    for (int II = 0; II < ns; ++II)
        for (int JJ = II; JJ < ns; ++JJ) {
            int i0 = c->ksumtype[II] + c->qsbasis[II * 4 + c->qstart[II]];
            int i1 = c->ksumtype[II] + c->qfbasis[II * 4 + c->qfinal[II]];
            int j0 = c->ksumtype[JJ] + c->qsbasis[JJ * 4 + c->qstart[JJ]];
            int j1 = c->ksumtype[JJ] + c->qfbasis[JJ * 4 + c->qfinal[JJ]];
            double dn = 0.0;
            for (int i = i0; i <= i1; ++i)
                for (int j = j0; j <= j1; ++j) {
                    double d = fabs(P[j * n + i]);
                    if (d > dn) dn = d;
                }
            cutmatrix[II * ns + JJ] = dn;
            cutmatrix[JJ * ns + II] = dn;
        }
}

/* ------------------------------------------------------------------------- */
/* Digestion of one sub-class: iclass_cshell / iclass_oshell                 */
/* (eri.fh:448-1057).  One target adds J(PJ) - hk/2 K(PK) into o:            */
/*   closed shell: PJ = PK = P,  hk = hyb                                     */
/*   open shell, alpha: PJ = Pa + Pb, PK = Pa, hk = 2 hyb; beta with Pb.      */
/* Every OSHELL branch of eri.fh is the closed-shell branch with the Coulomb */
/* densities replaced by dense + denseb and the exchange densities by dense  */
/* (or denseb) at twice the weight, so one restatement serves both.          */
/* ------------------------------------------------------------------------- */
typedef struct {
    const double *PJ, *PK;
    double hk;
    double *o;
} digest_t;

#define O_(r, cc) o[(size_t)(r) * n + (cc)]
#define PJ_(r, cc) PJ[(size_t)(r) * n + (cc)]
#define PK_(r, cc) PK[(size_t)(r) * n + (cc)]

/* III < JJJ, KKK < LLL, III < KKK (eri.fh:649-704 and 719-766) */
static void dig_general(const digest_t *d, int n, int I, int J, int K, int L, double Y)
{
    const double *PJ = d->PJ, *PK = d->PK, h = d->hk;
    double *o = d->o;
    O_(J, I) += 2.0 * PJ_(L, K) * Y;
    O_(L, K) += 2.0 * PJ_(J, I) * Y;
    O_(K, I) -= h * 0.5 * PK_(L, J) * Y;
    O_(L, I) -= h * 0.5 * PK_(K, J) * Y;
    O_(J, K) -= h * 0.5 * PK_(L, I) * Y;
    O_(J, L) -= h * 0.5 * PK_(K, I) * Y;
    O_(K, J) -= h * 0.5 * PK_(L, I) * Y;
    O_(L, J) -= h * 0.5 * PK_(K, I) * Y;
}

/* III < KKK, the special index patterns (eri.fh:716-904) */
static void dig_lt(const digest_t *d, int n, int I, int J, int K, int L, double Y)
{
    const double *PJ = d->PJ, *PK = d->PK, h = d->hk;
    double *o = d->o;
    if (I < J && K < L) {
        dig_general(d, n, I, J, K, L, Y);
    } else if (I == J && K == L) {
        O_(I, I) += PJ_(K, K) * Y;
        O_(K, K) += PJ_(I, I) * Y;
        O_(K, I) -= h * 0.5 * PK_(K, I) * Y;
    } else if (J == K && J == L) {
        O_(J, I) += PJ_(J, J) * Y - h * 0.5 * PK_(J, J) * Y;
        O_(J, J) += 2.0 * PJ_(J, I) * Y - h * PK_(J, I) * Y;
    } else if (K == L && I < J && J != K) {
        O_(J, I) += PJ_(K, K) * Y;
        O_(K, K) += 2.0 * PJ_(J, I) * Y;
        O_(K, I) -= h * 0.5 * PK_(K, J) * Y;
        O_(K, J) -= h * 0.5 * PK_(K, I) * Y;
        O_(J, K) -= h * 0.5 * PK_(K, I) * Y;
    } else if (I == J && K < L) {
        O_(L, K) += PJ_(I, I) * Y;
        O_(I, I) += 2.0 * PJ_(L, K) * Y;
        O_(K, I) -= h * 0.5 * PK_(L, I) * Y;
        O_(L, I) -= h * 0.5 * PK_(K, I) * Y;
    }
}

/* III == KKK, JJJ <= LLL (eri.fh:906-1046) */
static void dig_eq(const digest_t *d, int n, int I, int J, int K, int L, double Y)
{
    const double *PJ = d->PJ, *PK = d->PK, h = d->hk;
    double *o = d->o;
    if (I == J && I == K && I == L) {
        O_(I, I) += PJ_(I, I) * Y - h * 0.5 * PK_(I, I) * Y;
    } else if (I == J && I == K && I < L) {
        O_(L, I) += PJ_(I, I) * Y - h * 0.5 * PK_(I, I) * Y;
        O_(I, I) += 2.0 * PJ_(L, I) * Y - h * PK_(L, I) * Y;
    } else if (I == K && J == L && I < J) {
        O_(J, I) += 2.0 * PJ_(J, I) * Y - h * 0.5 * PK_(J, I) * Y;
        O_(J, J) -= h * 0.5 * PK_(I, I) * Y;
        O_(I, I) -= h * 0.5 * PK_(J, J) * Y;
    } else if (I == K && I < J && J < L) {
        O_(J, I) += 2.0 * PJ_(L, I) * Y - h * 0.5 * PK_(L, I) * Y;
        O_(L, I) += 2.0 * PJ_(J, I) * Y - h * 0.5 * PK_(J, I) * Y;
        O_(I, I) -= h * 1.0 * PK_(L, J) * Y;
        O_(L, J) -= h * 0.5 * PK_(I, I) * Y;
    }
}

static void iclass(const qo_ctx *c, int II, int JJ, int KK, int LL, int I, int J, int K, int L,
                   const prim_store *ps, const digest_t *dg, int ndg, double store[NCMAX][NCMAX])
{
    const int n = c->nbasis;
    contract(c, ps, I, J, K, L, store);
    const double *RA = c->xyz + 3 * c->katom[II], *RB = c->xyz + 3 * c->katom[JJ];
    const double *RC = c->xyz + 3 * c->katom[KK], *RD = c->xyz + 3 * c->katom[LL];
    double AB[3] = {RA[0] - RB[0], RA[1] - RB[1], RA[2] - RB[2]};
    double CD[3] = {RC[0] - RD[0], RC[1] - RD[1], RC[2] - RD[2]};
    int III1 = c->ksumtype[II] + c->qsbasis[II * 4 + I], III2 = c->ksumtype[II] + c->qfbasis[II * 4 + I];
    int JJJ1 = c->ksumtype[JJ] + c->qsbasis[JJ * 4 + J], JJJ2 = c->ksumtype[JJ] + c->qfbasis[JJ * 4 + J];
    int KKK1 = c->ksumtype[KK] + c->qsbasis[KK * 4 + K], KKK2 = c->ksumtype[KK] + c->qfbasis[KK * 4 + K];
    int LLL1 = c->ksumtype[LL] + c->qsbasis[LL * 4 + L], LLL2 = c->ksumtype[LL] + c->qfbasis[LL * 4 + L];

    if (II < JJ && II < KK && KK < LL) {
        for (int III = III1; III <= III2; ++III)
            for (int JJJ = JJJ1; JJJ <= JJJ2; ++JJJ)
                for (int KKK = KKK1; KKK <= KKK2; ++KKK)
                    for (int LLL = LLL1; LLL <= LLL2; ++LLL) {
                        double Y = hrr_whole(c, III, JJJ, KKK, LLL, AB, CD, store);
                        for (int t = 0; t < ndg; ++t) dig_general(&dg[t], n, III, JJJ, KKK, LLL, Y);
                    }
        return;
    }
    for (int III = III1; III <= III2; ++III)
        for (int JJJ = (III > JJJ1 ? III : JJJ1); JJJ <= JJJ2; ++JJJ)
            for (int KKK = (III > KKK1 ? III : KKK1); KKK <= KKK2; ++KKK)
                for (int LLL = (KKK > LLL1 ? KKK : LLL1); LLL <= LLL2; ++LLL) {
                    if (III < KKK) {
                        double Y = hrr_whole(c, III, JJJ, KKK, LLL, AB, CD, store);
                        for (int t = 0; t < ndg; ++t) dig_lt(&dg[t], n, III, JJJ, KKK, LLL, Y);
                    } else if (JJJ <= LLL) {
                        double Y = hrr_whole(c, III, JJJ, KKK, LLL, AB, CD, store);
                        for (int t = 0; t < ndg; ++t) dig_eq(&dg[t], n, III, JJJ, KKK, LLL, Y);
                    }
                }
}

typedef struct {
    prim_store ps;
    double (*v)[NCMAX][MMAX];
    double (*store)[NCMAX];
} work_t;

static void work_init(const qo_ctx *c, work_t *w)
{
    size_t mc4 = (size_t)c->maxcontract * c->maxcontract * c->maxcontract * c->maxcontract;
    w->ps.pidx = malloc(sizeof(int) * 4 * mc4);
    w->ps.Y = malloc(sizeof(double) * mc4 * NCMAX * NCMAX);
    w->v = malloc(sizeof(double) * NCMAX * NCMAX * MMAX);
    w->store = malloc(sizeof(double) * NCMAX * NCMAX);
}

static void work_free(work_t *w)
{
    free(w->ps.pidx);
    free(w->ps.Y);
    free(w->v);
    free(w->store);
}

/* get_cshell_eri / get_oshell_eri (eri.fh:161-202) for one II, then
 * cshell / oshell (eri.fh:209-441): screening and primitives are shared, the
 * digestion adds into every target. */
static long get_eri_ii(const qo_ctx *c, int II, const double *cm, double cutoff, double prim_limit,
                       const digest_t *dg, int ndg, work_t *w)
{
    const int ns = c->nshell;
    long nq = 0;
    for (int JJ = II; JJ < ns; ++JJ) {
        double testtmp = c->Ycutoff[II * ns + JJ];
        for (int KK = II; KK < ns; ++KK)
            for (int LL = KK; LL < ns; ++LL) {
                double cutoffTest = testtmp * c->Ycutoff[KK * ns + LL];
                if (!(cutoffTest > cutoff)) continue;
                double DNmax = 4.0 * cm[II * ns + JJ];
                double t;
                if ((t = 4.0 * cm[KK * ns + LL]) > DNmax) DNmax = t;
                if ((t = cm[II * ns + LL]) > DNmax) DNmax = t;
                if ((t = cm[II * ns + KK]) > DNmax) DNmax = t;
                if ((t = cm[JJ * ns + KK]) > DNmax) DNmax = t;
                if ((t = cm[JJ * ns + LL]) > DNmax) DNmax = t;
                if (!(cutoffTest * DNmax > cutoff)) continue;
                ++nq;
                shell_prims(c, II, JJ, KK, LL, DNmax, prim_limit, 1, &w->ps, w->v);
                for (int I = c->qstart[II]; I <= c->qfinal[II]; ++I)
                    for (int J = c->qstart[JJ]; J <= c->qfinal[JJ]; ++J)
                        for (int K = c->qstart[KK]; K <= c->qfinal[KK]; ++K)
                            for (int L = c->qstart[LL]; L <= c->qfinal[LL]; ++L)
                                iclass(c, II, JJ, KK, LL, I, J, K, L, &w->ps, dg, ndg, w->store);
            }
    }
    return nq;
}

static void copy_sym(double *o, int n)
{
    for (int i = 0; i < n; ++i)
        for (int j = i + 1; j < n; ++j) o[(size_t)i * n + j] = o[(size_t)j * n + i];
}

/* The build over II in [ii_begin, ii_end) on nthreads threads (QUICK's MPI
 * build weights II by jshell - II, mpi_setup.f90:317-322; a dynamic schedule
 * in that order gives the same balance on threads).  Each thread digests
 * into private copies of the targets, summed at the end. */
static long run_get2e(const qo_ctx *c, const double *cm, double cutoff, double prim_limit,
                      const digest_t *dg, int ndg, int ii_begin, int ii_end, int nthreads)
{
    const int n = c->nbasis;
    long nq = 0;
#ifdef _OPENMP
    if (nthreads <= 0) nthreads = omp_get_max_threads();
#else
    nthreads = 1;
#endif
    if (nthreads == 1) {
        work_t w;
        work_init(c, &w);
        for (int II = ii_begin; II < ii_end; ++II) nq += get_eri_ii(c, II, cm, cutoff, prim_limit, dg, ndg, &w);
        work_free(&w);
        return nq;
    }
    const size_t nn = (size_t)n * n;
    double *part = calloc((size_t)nthreads * ndg * nn, sizeof(double));
#pragma omp parallel num_threads(nthreads) reduction(+ : nq)
    {
        int tid = 0;
#ifdef _OPENMP
        tid = omp_get_thread_num();
#endif
        work_t w;
        work_init(c, &w);
        digest_t mine[2];
        for (int t = 0; t < ndg; ++t) {
            mine[t] = dg[t];
            mine[t].o = part + ((size_t)tid * ndg + t) * nn;
        }
#pragma omp for schedule(dynamic, 1)
        for (int II = ii_begin; II < ii_end; ++II) nq += get_eri_ii(c, II, cm, cutoff, prim_limit, mine, ndg, &w);
        work_free(&w);
    }
    for (int th = 0; th < nthreads; ++th)
        for (int t = 0; t < ndg; ++t)
            for (size_t k = 0; k < nn; ++k) dg[t].o[k] += part[((size_t)th * ndg + t) * nn + k];
    free(part);
    return nq;
}

long qo_get2e_range(const qo_ctx *c, const double *P, const double *cutmatrix, double cutoff,
                    double prim_limit, double hyb, int ii_begin, int ii_end, double *o)
{
    const int ns = c->nshell;
    double *cm = NULL;
    if (!cutmatrix) {
        cm = malloc(sizeof(double) * ns * ns);
        qo_density_cutoff(c, P, cm);
        cutmatrix = cm;
    }
    digest_t dg = {P, P, hyb, o};
    long nq = run_get2e(c, cutmatrix, cutoff, prim_limit, &dg, 1, ii_begin, ii_end, 1);
    free(cm);
    return nq;
}

long qo_get2e(const qo_ctx *c, const double *P, const double *cutmatrix, double cutoff,
              double prim_limit, double hyb, double *o)
{
    long nq = qo_get2e_range(c, P, cutmatrix, cutoff, prim_limit, hyb, 0, c->nshell, o);
    copy_sym(o, c->nbasis);
    return nq;
}

long qo_get2e_omp(const qo_ctx *c, const double *P, const double *cutmatrix, double cutoff,
                  double prim_limit, double hyb, int nthreads, double *o)
{
    const int ns = c->nshell;
    double *cm = NULL;
    if (!cutmatrix) {
        cm = malloc(sizeof(double) * ns * ns);
        qo_density_cutoff(c, P, cm);
        cutmatrix = cm;
    }
    digest_t dg = {P, P, hyb, o};
    long nq = run_get2e(c, cutmatrix, cutoff, prim_limit, &dg, 1, 0, ns, nthreads);
    free(cm);
    copy_sym(o, c->nbasis);
    return nq;
}

/* oshell_dnscreen / oshell_density_cutoff (cutoff.fh:16-89):
 * cutmatrix(II,JJ) = max over the block of max(|Pa+Pb|, |Pa|, |Pb|). */
void qo_density_cutoff_os(const qo_ctx *c, const double *Pa, const double *Pb, double *cutmatrix)
{
    const int n = c->nbasis, ns = c->nshell;
    for (int II = 0; II < ns; ++II)
        for (int JJ = II; JJ < ns; ++JJ) {
            int i0 = c->ksumtype[II] + c->qsbasis[II * 4 + c->qstart[II]];
            int i1 = c->ksumtype[II] + c->qfbasis[II * 4 + c->qfinal[II]];
            int j0 = c->ksumtype[JJ] + c->qsbasis[JJ * 4 + c->qstart[JJ]];
            int j1 = c->ksumtype[JJ] + c->qfbasis[JJ * 4 + c->qfinal[JJ]];
            double dn = 0.0;
            for (int i = i0; i <= i1; ++i)
                for (int j = j0; j <= j1; ++j) {
                    double a = Pa[j * n + i], b = Pb[j * n + i];
                    double d = fmax(fabs(a + b), fmax(fabs(a), fabs(b)));
                    if (d > dn) dn = d;
                }
            cutmatrix[II * ns + JJ] = dn;
            cutmatrix[JJ * ns + II] = dn;
        }
}

/* get_oshell_eri (eri.fh:159-200, iclass_oshell :446-1055) over all II, then
 * copySym of both matrices: oa += J(Pa+Pb) - hyb K(Pa), ob += J(Pa+Pb) - hyb K(Pb). */
long qo_get2e_os(const qo_ctx *c, const double *Pa, const double *Pb, const double *cutmatrix, double cutoff,
                 double prim_limit, double hyb, int nthreads, double *oa, double *ob)
{
    const int ns = c->nshell, n = c->nbasis;
    const size_t nn = (size_t)n * n;
    double *cm = NULL;
    if (!cutmatrix) {
        cm = malloc(sizeof(double) * ns * ns);
        qo_density_cutoff_os(c, Pa, Pb, cm);
        cutmatrix = cm;
    }
    double *Pt = malloc(sizeof(double) * nn);
    for (size_t k = 0; k < nn; ++k) Pt[k] = Pa[k] + Pb[k];
    digest_t dg[2] = {{Pt, Pa, 2.0 * hyb, oa}, {Pt, Pb, 2.0 * hyb, ob}};
    long nq = run_get2e(c, cutmatrix, cutoff, prim_limit, dg, 2, 0, ns, nthreads);
    free(Pt);
    free(cm);
    copy_sym(oa, n);
    copy_sym(ob, n);
    return nq;
}

/* ------------------------------------------------------------------------- */
/* Brute-force integral tensor (tiny systems; digestion checks)              */
/* ------------------------------------------------------------------------- */
/* ------------------------------------------------------------------------- */
/* Two-electron gradient: get_cshell_eri_grad (gradient.fh:826-930),         */
/* cshell_eri_grad (eri_grad.fh:43-199) and iclass_grad_cshell               */
/* (eri_grad.fh:207-768).  Shells up to d: the raised VRR reaches            */
/* [e|f] with |e|, |f| <= 5.                                                 */
/* ------------------------------------------------------------------------- */
#define GMAXL 2

typedef struct {
    int nprim;
    int E, F;        /* Qfinal sums of bra and ket; the VRR runs to E+1, F+1 */
    int *pidx;       /* nprim*4: Nii, Njj, Nkk, Nll */
    double *Y;       /* nprim * NCUM(E+1) * NCUM(F+1) */
} gprim_store;

/* cshell_eri_grad's primitive loops (eri_grad.fh:97-170): FmT with
 * NABCD = Qfinal sum + 2, VRR raised by one on both sides. */
static void grad_shell_prims(const qo_ctx *c, int II, int JJ, int KK, int LL, double dnmax, double prim_limit,
                             gprim_store *ps, double v[NCMAX][NCMAX][MMAX])
{
    const int jb = c->jbasis;
    const double *RA = c->xyz + 3 * c->katom[II];
    const double *RC = c->xyz + 3 * c->katom[KK];
    const int E = c->qfinal[II] + c->qfinal[JJ], F = c->qfinal[KK] + c->qfinal[LL];
    const int ne = NCUM(E + 1), nf = NCUM(F + 1);
    double fm[MMAX + 5];
    ps->E = E;
    ps->F = F;
    ps->nprim = 0;
    for (int jjj = 0; jjj < c->kprim[JJ]; ++jjj) {
        int Njj = c->kstart[JJ] + jjj;
        for (int iii = 0; iii < c->kprim[II]; ++iii) {
            int Nii = c->kstart[II] + iii;
            double cutoffprim1 = dnmax * c->cutprim[Nii * jb + Njj];
            for (int lll = 0; lll < c->kprim[LL]; ++lll) {
                int Nll = c->kstart[LL] + lll;
                for (int kkk = 0; kkk < c->kprim[KK]; ++kkk) {
                    int Nkk = c->kstart[KK] + kkk;
                    if (!(cutoffprim1 * c->cutprim[Nkk * jb + Nll] > prim_limit)) continue;
                    vrr_par q;
                    double T, prefac;
                    prim_par(c, Nii, Njj, Nkk, Nll, RA, RC, &q, &T, &prefac);
                    qo_fmt(E + F + 2, T, fm);
                    vrr_full(E + 1, F + 1, fm, prefac, &q, v);
                    int n = ps->nprim++;
                    ps->pidx[n * 4 + 0] = Nii;
                    ps->pidx[n * 4 + 1] = Njj;
                    ps->pidx[n * 4 + 2] = Nkk;
                    ps->pidx[n * 4 + 3] = Nll;
                    double *dst = ps->Y + (size_t)n * ne * nf;
                    for (int a = 0; a < ne; ++a)
                        for (int f = 0; f < nf; ++f) dst[a * nf + f] = v[a][f][0];
                }
            }
        }
    }
}

/* (a b|c d) from a contracted store by the closed-form HRR on both sides,
 * for explicit Cartesian exponents (hrrwholeopt). */
static double hrr_lmn(const int *a, const int *b, const int *cc, const int *d, const double *AB, const double *CD,
                      double store[NCMAX][NCMAX])
{
    double cl[64], cr[64];
    int al[64], ar[64];
    int nl = hrr_terms(a, b, AB, cl, al);
    int nr = hrr_terms(cc, d, CD, cr, ar);
    double Y = 0.0;
    for (int i = 0; i < nl; ++i)
        for (int j = 0; j < nr; ++j) Y += cl[i] * cr[j] * store[al[i]][ar[j]];
    return Y;
}

typedef struct {
    gprim_store ps;
    double (*v)[NCMAX][MMAX];
    double (*st)[NCMAX][NCMAX];   /* store, storeAA, storeBB, storeCC */
} gwork_t;

/* iclass_grad_cshell for one sub-class (I,J,K,L): the four contracted stores
 * (weights 1, 2 AA, 2 BB, 2 CC; eri_grad.fh:238-309), then per Cartesian
 * quartet of QUICK's visiting order (:357-766) the derivatives Yaa, Ybb, Ycc
 * (hrrwholeopt: 2a (a+1|..) - n (a-1|..), times the cons factors) weighted
 * with QUICK's density constant.  Every branch's constant equals
 * (f/8) [4 P_ij P_kl - hyb (P_ik P_jl + P_il P_jk)], f the number of distinct
 * index permutations of the quartet (8, or 4 / 2 for the coincidences QUICK
 * enumerates). */
static void iclass_grad(const qo_ctx *c, int II, int JJ, int KK, int LL, int I, int J, int K, int L,
                        const gprim_store *ps, const double *P, double hyb, double *grad, gwork_t *w)
{
    const int n = c->nbasis, jb = c->jbasis;
    const int ne = NCUM(ps->E + 1), nf = NCUM(ps->F + 1);
    for (int t = 0; t < 4; ++t)
        for (int a = 0; a < ne; ++a)
            for (int f = 0; f < nf; ++f) w->st[t][a][f] = 0.0;
    for (int k = 0; k < ps->nprim; ++k) {
        const int *p = ps->pidx + k * 4;
        double X44 = c->X0 * XC(c, p[0], p[1], I, J) * XC(c, p[2], p[3], K, L);
        double wt[4] = {X44, X44 * c->gcexpo[p[0]] * 2.0, X44 * c->gcexpo[p[1]] * 2.0, X44 * c->gcexpo[p[2]] * 2.0};
        const double *src = ps->Y + (size_t)k * ne * nf;
        for (int t = 0; t < 4; ++t)
            for (int a = 0; a < ne; ++a)
                for (int f = 0; f < nf; ++f) w->st[t][a][f] += wt[t] * src[a * nf + f];
    }
    (void)jb;
    const double *RA = c->xyz + 3 * c->katom[II], *RB = c->xyz + 3 * c->katom[JJ];
    const double *RC = c->xyz + 3 * c->katom[KK], *RD = c->xyz + 3 * c->katom[LL];
    double AB[3] = {RA[0] - RB[0], RA[1] - RB[1], RA[2] - RB[2]};
    double CD[3] = {RC[0] - RD[0], RC[1] - RD[1], RC[2] - RD[2]};
    int III1 = c->ksumtype[II] + c->qsbasis[II * 4 + I], III2 = c->ksumtype[II] + c->qfbasis[II * 4 + I];
    int JJJ1 = c->ksumtype[JJ] + c->qsbasis[JJ * 4 + J], JJJ2 = c->ksumtype[JJ] + c->qfbasis[JJ * 4 + J];
    int KKK1 = c->ksumtype[KK] + c->qsbasis[KK * 4 + K], KKK2 = c->ksumtype[KK] + c->qfbasis[KK * 4 + K];
    int LLL1 = c->ksumtype[LL] + c->qsbasis[LL * 4 + L], LLL2 = c->ksumtype[LL] + c->qfbasis[LL * 4 + L];
    const int general = II < JJ && II < KK && KK < LL;
    double g[9] = {0};
    for (int III = III1; III <= III2; ++III)
        for (int JJJ = general ? JJJ1 : (III > JJJ1 ? III : JJJ1); JJJ <= JJJ2; ++JJJ)
            for (int KKK = general ? KKK1 : (III > KKK1 ? III : KKK1); KKK <= KKK2; ++KKK)
                for (int LLL = general ? LLL1 : (KKK > LLL1 ? KKK : LLL1); LLL <= LLL2; ++LLL) {
                    if (!general && !(III < KKK) && !(JJJ <= LLL)) continue;
                    if (!general && III == JJJ && III == KKK && III == LLL) continue;   /* (ii|ii): nothing */
                    const double f = (III != JJJ ? 2.0 : 1.0) * (KKK != LLL ? 2.0 : 1.0) *
                                     ((III != KKK || JJJ != LLL) ? 2.0 : 1.0);
                    const double Pij = P[(size_t)JJJ * n + III], Pkl = P[(size_t)LLL * n + KKK];
                    const double Pik = P[(size_t)KKK * n + III], Pjl = P[(size_t)LLL * n + JJJ];
                    const double Pil = P[(size_t)LLL * n + III], Pjk = P[(size_t)KKK * n + JJJ];
                    const double constant = f / 8.0 * (4.0 * Pij * Pkl - hyb * (Pik * Pjl + Pil * Pjk));
                    const double cons = c->fn_cons[III] * c->fn_cons[JJJ] * c->fn_cons[KKK] * c->fn_cons[LLL];
                    const int *a = c->fn_lmn + 3 * III, *b = c->fn_lmn + 3 * JJJ;
                    const int *cc = c->fn_lmn + 3 * KKK, *d = c->fn_lmn + 3 * LLL;
                    for (int x = 0; x < 3; ++x) {
                        int ap[3] = {a[0], a[1], a[2]}, am_[3] = {a[0], a[1], a[2]};
                        int bp[3] = {b[0], b[1], b[2]}, bm[3] = {b[0], b[1], b[2]};
                        int cp[3] = {cc[0], cc[1], cc[2]}, cm_[3] = {cc[0], cc[1], cc[2]};
                        ap[x] += 1;
                        bp[x] += 1;
                        cp[x] += 1;
                        double ya = hrr_lmn(ap, b, cc, d, AB, CD, w->st[1]);
                        if (a[x] > 0) {
                            am_[x] -= 1;
                            ya -= a[x] * hrr_lmn(am_, b, cc, d, AB, CD, w->st[0]);
                        }
                        double yb = hrr_lmn(a, bp, cc, d, AB, CD, w->st[2]);
                        if (b[x] > 0) {
                            bm[x] -= 1;
                            yb -= b[x] * hrr_lmn(a, bm, cc, d, AB, CD, w->st[0]);
                        }
                        double yc = hrr_lmn(a, b, cp, d, AB, CD, w->st[3]);
                        if (cc[x] > 0) {
                            cm_[x] -= 1;
                            yc -= cc[x] * hrr_lmn(a, b, cm_, d, AB, CD, w->st[0]);
                        }
                        g[x] += constant * cons * ya;
                        g[3 + x] += constant * cons * yb;
                        g[6 + x] += constant * cons * yc;
                    }
                }
    const int iA = c->katom[II], iB = c->katom[JJ], iC = c->katom[KK], iD = c->katom[LL];
    for (int x = 0; x < 3; ++x) {
        grad[iA * 3 + x] += g[x];
        grad[iB * 3 + x] += g[3 + x];
        grad[iC * 3 + x] += g[6 + x];
        grad[iD * 3 + x] -= g[x] + g[3 + x] + g[6 + x];
    }
}

/* get_cshell_eri_grad's loop for one II (gradient.fh:896-925). */
static long grad_ii(const qo_ctx *c, int II, const double *cm, const double *P, double grad_cutoff,
                    double prim_limit, double hyb, double *grad, gwork_t *w)
{
    const int ns = c->nshell;
    long nq = 0;
    for (int JJ = II; JJ < ns; ++JJ) {
        double testtmp = c->Ycutoff[II * ns + JJ];
        for (int KK = II; KK < ns; ++KK)
            for (int LL = KK; LL < ns; ++LL) {
                if (c->katom[II] == c->katom[JJ] && c->katom[II] == c->katom[KK] && c->katom[II] == c->katom[LL])
                    continue;
                double testCutoff = testtmp * c->Ycutoff[KK * ns + LL];
                if (!(testCutoff > grad_cutoff)) continue;
                double DNmax = 4.0 * cm[II * ns + JJ];
                double t;
                if ((t = 4.0 * cm[KK * ns + LL]) > DNmax) DNmax = t;
                if ((t = cm[II * ns + LL]) > DNmax) DNmax = t;
                if ((t = cm[II * ns + KK]) > DNmax) DNmax = t;
                if ((t = cm[JJ * ns + KK]) > DNmax) DNmax = t;
                if ((t = cm[JJ * ns + LL]) > DNmax) DNmax = t;
                if (!(testCutoff * DNmax > grad_cutoff)) continue;
                ++nq;
                grad_shell_prims(c, II, JJ, KK, LL, DNmax, prim_limit, &w->ps, w->v);
                for (int I = c->qstart[II]; I <= c->qfinal[II]; ++I)
                    for (int J = c->qstart[JJ]; J <= c->qfinal[JJ]; ++J)
                        for (int K = c->qstart[KK]; K <= c->qfinal[KK]; ++K)
                            for (int L = c->qstart[LL]; L <= c->qfinal[LL]; ++L)
                                iclass_grad(c, II, JJ, KK, LL, I, J, K, L, &w->ps, P, hyb, grad, w);
            }
    }
    return nq;
}

long qo_eri_grad(const qo_ctx *c, const double *P, const double *cutmatrix, double grad_cutoff, double prim_limit,
                 double hyb, int nthreads, double *grad)
{
    const int ns = c->nshell;
    for (int s = 0; s < ns; ++s)
        if (c->qfinal[s] > GMAXL) return -1;
    double *cm = NULL;
    if (!cutmatrix) {
        cm = malloc(sizeof(double) * ns * ns);
        qo_density_cutoff(c, P, cm);
        cutmatrix = cm;
    }
#ifdef _OPENMP
    if (nthreads <= 0) nthreads = omp_get_max_threads();
#else
    nthreads = 1;
#endif
    const size_t ng = (size_t)3 * c->natom;
    double *part = calloc((size_t)nthreads * ng, sizeof(double));
    long nq = 0;
#pragma omp parallel num_threads(nthreads) reduction(+ : nq)
    {
        int tid = 0;
#ifdef _OPENMP
        tid = omp_get_thread_num();
#endif
        size_t mc4 = (size_t)c->maxcontract * c->maxcontract * c->maxcontract * c->maxcontract;
        gwork_t w;
        w.ps.pidx = malloc(sizeof(int) * 4 * mc4);
        w.ps.Y = malloc(sizeof(double) * mc4 * NCUM(2 * GMAXL + 1) * NCUM(2 * GMAXL + 1));
        w.v = malloc(sizeof(double) * NCMAX * NCMAX * MMAX);
        w.st = malloc(sizeof(double) * 4 * NCMAX * NCMAX);
#pragma omp for schedule(dynamic, 1)
        for (int II = 0; II < ns; ++II)
            nq += grad_ii(c, II, cutmatrix, P, grad_cutoff, prim_limit, hyb, part + (size_t)tid * ng, &w);
        free(w.ps.pidx);
        free(w.ps.Y);
        free(w.v);
        free(w.st);
    }
    for (int t = 0; t < nthreads; ++t)
        for (size_t k = 0; k < ng; ++k) grad[k] += part[(size_t)t * ng + k];
    free(part);
    free(cm);
    return nq;
}

void qo_eri_tensor(const qo_ctx *c, double *eri)
{
    const int ns = c->nshell;
    const size_t n = (size_t)c->nbasis;
    work_t w;
    work_init(c, &w);
    for (int II = 0; II < ns; ++II)
        for (int JJ = 0; JJ < ns; ++JJ)
            for (int KK = 0; KK < ns; ++KK)
                for (int LL = 0; LL < ns; ++LL) {
                    shell_prims(c, II, JJ, KK, LL, 0.0, 0.0, 0, &w.ps, w.v);
                    const double *RA = c->xyz + 3 * c->katom[II], *RB = c->xyz + 3 * c->katom[JJ];
                    const double *RC = c->xyz + 3 * c->katom[KK], *RD = c->xyz + 3 * c->katom[LL];
                    double AB[3] = {RA[0] - RB[0], RA[1] - RB[1], RA[2] - RB[2]};
                    double CD[3] = {RC[0] - RD[0], RC[1] - RD[1], RC[2] - RD[2]};
                    for (int I = c->qstart[II]; I <= c->qfinal[II]; ++I)
                        for (int J = c->qstart[JJ]; J <= c->qfinal[JJ]; ++J)
                            for (int K = c->qstart[KK]; K <= c->qfinal[KK]; ++K)
                                for (int L = c->qstart[LL]; L <= c->qfinal[LL]; ++L) {
                                    contract(c, &w.ps, I, J, K, L, w.store);
                                    for (int a = c->qsbasis[II * 4 + I]; a <= c->qfbasis[II * 4 + I]; ++a)
                                        for (int b = c->qsbasis[JJ * 4 + J]; b <= c->qfbasis[JJ * 4 + J]; ++b)
                                            for (int cc = c->qsbasis[KK * 4 + K]; cc <= c->qfbasis[KK * 4 + K]; ++cc)
                                                for (int d = c->qsbasis[LL * 4 + L]; d <= c->qfbasis[LL * 4 + L]; ++d) {
                                                    size_t i = c->ksumtype[II] + a, j = c->ksumtype[JJ] + b;
                                                    size_t k = c->ksumtype[KK] + cc, l = c->ksumtype[LL] + d;
                                                    eri[((i * n + j) * n + k) * n + l] =
                                                        hrr_whole(c, (int)i, (int)j, (int)k, (int)l, AB, CD, w.store);
                                                }
                                }
                }
    work_free(&w);
}

/* ------------------------------------------------------------------------- */
/* One-electron integrals (Obara-Saika) for the SCF harness                  */
/* QUICK: quick_overlap_module.f90 (overlap), quick_oei_module.f90:69-257    */
/* (kinetic, attrashell nuclear attraction with FmT).                        */
/* ------------------------------------------------------------------------- */
static void os_1d_overlap(int la, int lb, double PA, double PB, double oo2p, double S[8][8])
{
    S[0][0] = 1.0;
    for (int i = 1; i <= la; ++i) {
        S[i][0] = PA * S[i - 1][0] + (i > 1 ? (i - 1) * oo2p * S[i - 2][0] : 0.0);
    }
    for (int j = 1; j <= lb; ++j)
        for (int i = 0; i <= la; ++i) {
            double t = PB * S[i][j - 1];
            if (i > 0) t += i * oo2p * S[i - 1][j - 1];
            if (j > 1) t += (j - 1) * oo2p * S[i][j - 2];
            S[i][j] = t;
        }
}

void qo_one_electron(const qo_ctx *c, double *Smat, double *Tmat, double *Vmat)
{
    const int n = c->nbasis;
    int *fshell = malloc(sizeof(int) * n);
    int *fL = malloc(sizeof(int) * n);
    for (int s = 0; s < c->nshell; ++s)
        for (int L = c->qstart[s]; L <= c->qfinal[s]; ++L)
            for (int f = c->ksumtype[s] + c->qsbasis[s * 4 + L]; f <= c->ksumtype[s] + c->qfbasis[s * 4 + L]; ++f) {
                fshell[f] = s;
                fL[f] = L;
            }
    for (int mu = 0; mu < n; ++mu)
        for (int nu = 0; nu < n; ++nu) {
            int sa = fshell[mu], sb = fshell[nu];
            int La = fL[mu], Lb = fL[nu];
            const int *a = c->fn_lmn + 3 * mu, *b = c->fn_lmn + 3 * nu;
            const double *A = c->xyz + 3 * c->katom[sa], *B = c->xyz + 3 * c->katom[sb];
            double s_sum = 0.0, t_sum = 0.0, v_sum = 0.0;
            for (int ip = 0; ip < c->kprim[sa]; ++ip)
                for (int jp = 0; jp < c->kprim[sb]; ++jp) {
                    int pa = c->kstart[sa] + ip, pb = c->kstart[sb] + jp;
                    double al = c->gcexpo[pa], be = c->gcexpo[pb];
                    double ca = c->gccoeff[pa * 4 + La] * c->fn_cons[mu];
                    double cb = c->gccoeff[pb * 4 + Lb] * c->fn_cons[nu];
                    double p = al + be, oo2p = 0.5 / p;
                    double P[3], PA[3], PB[3], AB2 = 0.0;
                    for (int d = 0; d < 3; ++d) {
                        P[d] = (al * A[d] + be * B[d]) / p;
                        PA[d] = P[d] - A[d];
                        PB[d] = P[d] - B[d];
                        AB2 += (A[d] - B[d]) * (A[d] - B[d]);
                    }
                    double Kab = exp(-al * be / p * AB2);
                    double S1[3][8][8];
                    for (int d = 0; d < 3; ++d) os_1d_overlap(a[d] + 1, b[d] + 2, PA[d], PB[d], oo2p, S1[d]);
                    double s0 = pow(QO_PI / p, 1.5) * Kab;
                    double Sx = S1[0][a[0]][b[0]], Sy = S1[1][a[1]][b[1]], Sz = S1[2][a[2]][b[2]];
                    double ovl = s0 * Sx * Sy * Sz;
                    /* kinetic: T_d = -2 be^2 S(a,b+2) + be(2b+1) S(a,b) - b(b-1)/2 S(a,b-2) */
                    double Td[3];
                    for (int d = 0; d < 3; ++d) {
                        int bj = b[d];
                        double t = -2.0 * be * be * S1[d][a[d]][bj + 2] + be * (2 * bj + 1) * S1[d][a[d]][bj];
                        if (bj >= 2) t -= 0.5 * bj * (bj - 1) * S1[d][a[d]][bj - 2];
                        Td[d] = t;
                    }
                    double kin = s0 * (Td[0] * Sy * Sz + Sx * Td[1] * Sz + Sx * Sy * Td[2]);
                    s_sum += ca * cb * ovl;
                    t_sum += ca * cb * kin;
                    /* nuclear attraction: OS recursion on a then b with auxiliary m;
                     * FmT is called with Maxm = Qfinal(I)+Qfinal(J) as attrashell
                     * (attrashell.fh:46-48,89) */
                    const int L = La + Lb;
                    const int maxm = c->qfinal[sa] + c->qfinal[sb];
                    double vn = 0.0;
                    for (int at = 0; at < c->natom; ++at) {
                        const double *Cn = c->xyz + 3 * at;
                        double PC[3], PC2 = 0.0;
                        for (int d = 0; d < 3; ++d) {
                            PC[d] = P[d] - Cn[d];
                            PC2 += PC[d] * PC[d];
                        }
                        double fm[16];
                        qo_fmt(maxm, p * PC2, fm);
                        /* table over (a-comp, b-comp, m): use lmn up to la/lb */
                        static __thread double V[4][4][4][4][4][4][7];
                        /* V[ax][ay][az][bx][by][bz][m]; fill base */
                        for (int m = 0; m <= L; ++m) V[0][0][0][0][0][0][m] = 2.0 * QO_PI / p * Kab * fm[m];
                        /* build a (b = 0) */
                        for (int ax = 0; ax <= a[0]; ++ax)
                            for (int ay = 0; ay <= a[1]; ++ay)
                                for (int az = 0; az <= a[2]; ++az) {
                                    int aa[3] = {ax, ay, az};
                                    int asum = ax + ay + az;
                                    if (asum == 0) continue;
                                    int i = ax ? 0 : (ay ? 1 : 2);
                                    int a1[3] = {ax, ay, az};
                                    a1[i]--;
                                    for (int m = 0; m <= L - asum; ++m) {
                                        double t = PA[i] * V[a1[0]][a1[1]][a1[2]][0][0][0][m] -
                                                   PC[i] * V[a1[0]][a1[1]][a1[2]][0][0][0][m + 1];
                                        if (a1[i] > 0) {
                                            int a2[3] = {a1[0], a1[1], a1[2]};
                                            a2[i]--;
                                            t += a1[i] * oo2p *
                                                 (V[a2[0]][a2[1]][a2[2]][0][0][0][m] - V[a2[0]][a2[1]][a2[2]][0][0][0][m + 1]);
                                        }
                                        V[aa[0]][aa[1]][aa[2]][0][0][0][m] = t;
                                    }
                                }
                        /* build b for every a */
                        for (int bx = 0; bx <= b[0]; ++bx)
                            for (int by = 0; by <= b[1]; ++by)
                                for (int bz = 0; bz <= b[2]; ++bz) {
                                    int bb[3] = {bx, by, bz};
                                    int bsum = bx + by + bz;
                                    if (bsum == 0) continue;
                                    int i = bx ? 0 : (by ? 1 : 2);
                                    int b1[3] = {bx, by, bz};
                                    b1[i]--;
                                    for (int ax = 0; ax <= a[0]; ++ax)
                                        for (int ay = 0; ay <= a[1]; ++ay)
                                            for (int az = 0; az <= a[2]; ++az) {
                                                int aa[3] = {ax, ay, az};
                                                int asum = ax + ay + az;
                                                for (int m = 0; m <= L - asum - bsum; ++m) {
                                                    double t = PB[i] * V[ax][ay][az][b1[0]][b1[1]][b1[2]][m] -
                                                               PC[i] * V[ax][ay][az][b1[0]][b1[1]][b1[2]][m + 1];
                                                    if (aa[i] > 0) {
                                                        int a1[3] = {ax, ay, az};
                                                        a1[i]--;
                                                        t += aa[i] * oo2p *
                                                             (V[a1[0]][a1[1]][a1[2]][b1[0]][b1[1]][b1[2]][m] -
                                                              V[a1[0]][a1[1]][a1[2]][b1[0]][b1[1]][b1[2]][m + 1]);
                                                    }
                                                    if (b1[i] > 0) {
                                                        int b2[3] = {b1[0], b1[1], b1[2]};
                                                        b2[i]--;
                                                        t += b1[i] * oo2p *
                                                             (V[ax][ay][az][b2[0]][b2[1]][b2[2]][m] -
                                                              V[ax][ay][az][b2[0]][b2[1]][b2[2]][m + 1]);
                                                    }
                                                    V[ax][ay][az][bb[0]][bb[1]][bb[2]][m] = t;
                                                }
                                            }
                                }
                        vn -= c->charge[at] * V[a[0]][a[1]][a[2]][b[0]][b[1]][b[2]][0];
                    }
                    v_sum += ca * cb * vn;
                }
            Smat[mu * n + nu] = s_sum;
            Tmat[mu * n + nu] = t_sum;
            Vmat[mu * n + nu] = v_sum;
        }
    free(fshell);
    free(fL);
}
