/*
 * quick_oracle.h - CPU restatement of QUICK's closed-shell two-electron Fock
 * build (getCshellEri, and getOshellEri), used ONLY as test infrastructure: by tests/, by the
 * smoke check and by bench.py's cpu_baseline leg.  Nothing in the product
 * (quick_amd/, libquickfock) links or calls it.
 *
 * Every function cites the QUICK source it restates (paths relative to the
 * QUICK checkout).  All arithmetic is FP64; indices are 0-based; matrices are
 * n x n row-major (P and the Fock matrix are symmetric, so the layout only
 * fixes which triangle "lower" means: element (r, c) lives at [r*n + c]).
 */
#ifndef QUICK_ORACLE_H
#define QUICK_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
    int natom;
    const double *xyz;      /* natom*3, bohr */
    const double *charge;   /* natom nuclear charges */
    int nshell;
    const int *katom;       /* shell -> atom */
    const int *qstart;      /* lowest L of the shell (SP: 0) */
    const int *qfinal;      /* highest L of the shell (SP: 1) */
    const int *kprim;       /* primitives per shell */
    const int *kstart;      /* offset of the shell's first primitive */
    const int *ksumtype;    /* index of the shell's first basis function */
    int nbasis;
    int jbasis;             /* total primitives */
    const double *gcexpo;   /* jbasis exponents */
    const double *gccoeff;  /* jbasis*4: normalised coefficient [p*4 + L] */
} qo_basis;

typedef struct qo_ctx qo_ctx;

/* eri.fh:62-154 (get_eri_precomputables) + quick_cutoff_module.f90:54-588
 * (schwarzoff).  Copies the basis description. */
qo_ctx *qo_create(const qo_basis *b);
void qo_destroy(qo_ctx *c);

/* Ycutoff (nshell*nshell) and cutprim (jbasis*jbasis) as computed by schwarzoff. */
void qo_get_schwarz(const qo_ctx *c, double *ycutoff, double *cutprim);

/* cutoff.fh:65-92 (cshell_density_cutoff): cutmatrix(II,JJ) = max |P_ij|. */
void qo_density_cutoff(const qo_ctx *c, const double *P, double *cutmatrix);

/* quick_scf_operator_module.f90:146-157 + eri.fh:161-1057: o += G(P) over all
 * II (getCshellEri), then copySym.  hyb = x_hybrid_coeff (1 for HF).
 * If cutmatrix is NULL it is computed from P.  Returns the number of shell
 * quartets that passed screening. */
long qo_get2e(const qo_ctx *c, const double *P, const double *cutmatrix,
              double integral_cutoff, double prim_limit, double hyb, double *o);

/* Same, restricted to II in [ii_begin, ii_end) and without copySym: the unit
 * of work QUICK's MPI build hands to one rank (quick_scf_operator_module.f90:135-144). */
long qo_get2e_range(const qo_ctx *c, const double *P, const double *cutmatrix,
                    double integral_cutoff, double prim_limit, double hyb,
                    int ii_begin, int ii_end, double *o);

/* OpenMP-parallel build over II shells with QUICK's greedy weights
 * (mpi_setup.f90:317-322) and per-thread Fock reduction.  nthreads<=0: all. */
long qo_get2e_omp(const qo_ctx *c, const double *P, const double *cutmatrix,
                  double integral_cutoff, double prim_limit, double hyb,
                  int nthreads, double *o);

/* cutoff.fh:16-89 (oshell_density_cutoff): cutmatrix(II,JJ) = max over the
 * shell block of max(|Pa+Pb|, |Pa|, |Pb|). */
void qo_density_cutoff_os(const qo_ctx *c, const double *Pa, const double *Pb, double *cutmatrix);

/* Open-shell build (getOshellEri, eri.fh:159-200 with iclass_oshell
 * :446-1055): oa += J(Pa+Pb) - hyb K(Pa), ob += J(Pa+Pb) - hyb K(Pb), both
 * symmetric.  cutmatrix NULL: computed by qo_density_cutoff_os.  nthreads
 * as qo_get2e_omp (1: serial).  Returns the screened shell-quartet count. */
long qo_get2e_os(const qo_ctx *c, const double *Pa, const double *Pb, const double *cutmatrix, double cutoff,
                 double prim_limit, double hyb, int nthreads, double *oa, double *ob);

/* Two-electron gradient (get_cshell_eri_grad / cshell_eri_grad / iclass_grad_cshell,
 * gradient.fh:826-930, eri_grad.fh:43-768): grad[3*atom + xyz] += d/dR of
 * 1/2 sum P G(P) with P fixed, quartets screened with grad_cutoff (quartets on one
 * atom skipped).  cutmatrix NULL: qo_density_cutoff.  Shells up to d (returns -1
 * otherwise); nthreads as qo_get2e_omp.  Returns the screened shell-quartet count. */
long qo_eri_grad(const qo_ctx *c, const double *P, const double *cutmatrix, double grad_cutoff, double prim_limit,
                 double hyb, int nthreads, double *grad);

/* Test switch (default off = QUICK's FmT): the small-x start value from the
 * convergent series instead of QUICK's first-order form. */
void qo_set_fmt_series(int on);

/* fmt.f90:13-85: F_m(x) for m = 0..maxm into vals[0..maxm]. */
void qo_fmt(int maxm, double x, double *vals);

/* One-electron matrices S, T, V (nbasis*nbasis each) over normalised
 * Cartesian functions, nuclear attraction with the same FmT. */
void qo_one_electron(const qo_ctx *c, double *S, double *T, double *V);

/* Every contracted integral (ij|kl) over basis functions, nbasis^4 doubles
 * (tiny systems only; used to check digestion by brute force). */
void qo_eri_tensor(const qo_ctx *c, double *eri);

#ifdef __cplusplus
}
#endif
#endif
