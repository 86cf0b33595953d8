/*
 * quickfock.h - C ABI of libquickfock, the MI355X (gfx950) two-electron
 * Fock-build engine that replaces QUICK's gpu_get2e path.
 *
 * Two layers:
 *
 *  1. QUICK-compatible entry points (lower case, trailing underscore, every
 *     argument by reference, Fortran column-major arrays).  They keep QUICK's
 *     names, argument order and meaning so QUICK's Fortran scf_operator can
 *     call them unchanged.  Each cites the QUICK function it replaces.
 *
 *  2. The native API (qf_*): an explicit engine handle, device or host
 *     buffers, a HIP stream, precision mode, batched builds over many
 *     molecules, and error codes.  The QUICK layer is a thin shim over it.
 *
 * All matrices are nbasis x nbasis FP64.  Density and Fock matrices are
 * symmetric, so row- and column-major layouts coincide.
 */
#ifndef QUICKFOCK_H
#define QUICKFOCK_H

#include <stdbool.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------ */
/* Error codes                                                         */
/* ------------------------------------------------------------------ */
enum {
    QF_OK = 0,
    QF_ERR_ARG = 1,        /* bad argument / unsupported angular momentum */
    QF_ERR_STATE = 2,      /* call out of order (e.g. build before prepare) */
    QF_ERR_HIP = 3,        /* a HIP runtime call failed */
    QF_ERR_NODEVICE = 24,  /* QUICK's ierr 24: no GPU */
    QF_ERR_BADDEVICE = 25, /* QUICK's ierr 25: bad device id */
    QF_ERR_MEMORY = 30     /* device memory budget exceeded */
};

enum { QF_PREC_FP64 = 0, QF_PREC_MIXED = 1 };

typedef struct qf_engine qf_engine;

typedef struct {
    long long n_pairs;           /* shell pairs kept after Schwarz */
    long long n_tasks_fp64;      /* wave tasks (1 bra x <=64 kets) */
    long long n_tasks_fp32;
    long long n_quartets_bound;  /* quartets passing the Schwarz-only test */
    long long device_bytes;      /* device memory held by the engine */
} qf_stats;

/* Last error message of the calling thread (never NULL). */
const char *qf_last_error(void);

/* Engine lifetime.  device < 0 selects the current device. */
int qf_engine_create(int device, qf_engine **out);
void qf_engine_destroy(qf_engine *e);

/* Geometry in bohr, xyz[3*natom]. */
int qf_set_molecule(qf_engine *e, int natom, const double *xyz);

/* Basis in QUICK's shell layout (0-based indices):
 *   shell_atom[s], shell_lmin[s] (Qstart), shell_lmax[s] (Qfinal; SP = 0..1),
 *   shell_nprim[s] (kprim), shell_pstart[s] (kstart), shell_fstart[s]
 *   (ksumtype: first basis function), exps[nprim_total] (gcexpo) and
 *   coef[nprim_total*4] the normalised coefficient of primitive p for L
 *   (QUICK's gccoeff of function ksumtype+L), at coef[p*4+L].
 *   Angular momentum must be <= 3 (f); the ERI gradient covers shells up to d. */
int qf_set_basis(qf_engine *e, int nshell, const int *shell_atom, const int *shell_lmin,
                 const int *shell_lmax, const int *shell_nprim, const int *shell_pstart,
                 const int *shell_fstart, int nprim_total, const double *exps, const double *coef,
                 int nbasis);

/* Screening thresholds: integral cutoff, primitive limit (QUICK's
 * integralCutoff, primLimit) and exchange weight (x_hybrid_coeff, 1 = HF). */
int qf_set_cutoffs(qf_engine *e, double integral_cutoff, double prim_limit, double hyb);

/* Precision: QF_PREC_FP64, or QF_PREC_MIXED where shell quartets with
 * Ycut_ij*Ycut_kl < split*max(Ycut_ij*Ycut_kl) run in FP32: exponents, Boys
 * function (FP32 grid table), VRR and HRR (screening bounds, coordinate
 * differences and the Fock accumulation stay FP64). */
int qf_set_precision(qf_engine *e, int mode, double split);
/* Which classes a mixed build splits (native tuning, no QUICK counterpart):
 * those with La+Lb+Lc+Ld >= min_lsum that have per-lane kernels; -1 (the
 * default): 0 (all of them) above 256 basis functions, 5 up to 256.  The
 * others run every quartet in FP64. */
int qf_set_split_classes(qf_engine *e, int min_lsum);
/* Per-class override of that choice (native tuning, no QUICK counterpart; class
 * indices as qf_set_class_tuning): -1 follows qf_set_split_classes, 0 the class
 * runs every quartet in FP64, 1 it is split (where it has per-lane kernels). */
int qf_set_class_split(qf_engine *e, int cb, int ck, int mode);

/* Schwarz bounds.  Pass NULL to have them computed on the GPU, or QUICK's
 * Ycutoff[nshell*nshell] and cutprim[nprim_total*nprim_total]. */
int qf_set_schwarz(qf_engine *e, const double *ycutoff, const double *cutprim);
int qf_get_schwarz(qf_engine *e, double *ycutoff, double *cutprim);

/* Build pair lists, bounds and wave task lists (after the calls above). */
int qf_prepare(qf_engine *e);
int qf_get_stats(qf_engine *e, qf_stats *out);

/* F += G(P) for a device-resident density.  d_F may equal NULL, in which
 * case G is left in the engine and can be read with qf_copy_g.  d_cutmat
 * (nshell*nshell max |P| per shell pair) may be NULL: computed on device.
 * stream is a hipStream_t (NULL = default stream).  Asynchronous. */
int qf_fock_build_device(qf_engine *e, const double *d_P, double *d_F, const double *d_cutmat,
                         void *stream);

/* No counterpart in QUICK.  For callers that cannot allocate device memory
 * during a build (HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): the same build as qf_fock_build_device (same kernels,
 * same result), with its scratch (G partials, symmetrised G, the Cutmatrix
 * when d_cutmat is NULL) in a fixed device array of QF_NOALLOC_SCRATCH_BYTES
 * per GPU instead of engine buffers; no hipMalloc/hipFree, no HIP events.
 * Limits: 2 nbasis^2 (+ nshell^2 without d_cutmat) doubles must fit, and no
 * class may keep node-program slots in global memory ((ff|ff)-sized f
 * classes); QF_ERR_ARG otherwise.  Every call on one GPU shares that array:
 * calls that overlap on one GPU (other streams, other engines) race, and G
 * is only valid until the next no-alloc call (read it through d_F). */
#define QF_NOALLOC_SCRATCH_BYTES (32u << 20)
int qf_fock_build_device_noalloc(qf_engine *e, const double *d_P, double *d_F, const double *d_cutmat,
                                 void *stream);

/* Two-electron nuclear gradient (QUICK's gpu_upload_grad_ + gpu_get_cshell_eri_grad_,
 * gpu.cu:2136 and gpu_get2e_getxc_drivers.h:127; kernels getGrad_kernel_sp/_spd,
 * gpu_get2e_subs_grad.h:1211-1650; CPU cshell_eri_grad, eri_grad.fh:43-768):
 * grad[3*atom + xyz] += d/dR of E2 = 1/2 sum P G(P) with P held fixed, over the
 * quartets with Ycut Ycut > grad_cutoff and Ycut Ycut DNmax > grad_cutoff (quartets on
 * one atom skipped), primitive screen primLimit.  Shells up to d (QF_ERR_ARG with f).
 * qf_set_grad_cutoff: QUICK's gradCutoff (default: the integral cutoff); a value below
 * the integral cutoff makes the next qf_prepare extend the pair lists. */
int qf_set_grad_cutoff(qf_engine *e, double grad_cutoff);
/* lane = 1 (default): per-lane gradient kernels for the classes with La+Lb+Lc+Ld <= 5,
 * the wave-per-quartet kernel for the rest; lane = 0: the wave-per-quartet kernel for all */
int qf_set_grad_kernel(qf_engine *e, int lane);
/* Per-class gradient kernel (native tuning, no QUICK counterpart; class indices as
 * qf_set_class_tuning): -1 follows qf_set_grad_kernel, 0 the wave-per-quartet kernel,
 * 1 the per-lane kernel where the class has one.  With qf_set_timing on, a gradient
 * call records one event pair per class launch like a Fock build (qf_read_timing). */
int qf_set_grad_class(qf_engine *e, int cb, int ck, int mode);
/* device pointers, asynchronous on stream; d_cutmat may be NULL (computed) */
int qf_eri_grad_device(qf_engine *e, const double *d_P, double *d_grad, const double *d_cutmat, void *stream);
/* host arrays, synchronous */
int qf_eri_grad_host(qf_engine *e, const double *P, double *grad, const double *cutmat);
/* per Fock launch: algorithmic FLOPs, quartets and primitive quartets of the last
 * gradient (which must have computed its own Cutmatrix); any pointer may be NULL */
int qf_grad_launch_flops(qf_engine *e, double *flops, long long *quartets, long long *prim_quartets);

/* Host convenience: F[n*n] += G(P[n*n]); synchronous. */
int qf_fock_build_host(qf_engine *e, const double *P, double *F, const double *cutmat);

/* Open-shell build (QUICK's get_oshell_eri: gpu_get2e_getxc_drivers.h:25 and the
 * OSHELL branches of gpu_get2e_subs.h:391-575):
 *   F_a += J(P_a + P_b) - hyb K(P_a),   F_b += J(P_a + P_b) - hyb K(P_b).
 * Screening uses QUICK's open-shell density matrix (cutoff.fh:16-89, max of
 * |P_a + P_b|, |P_a|, |P_b| per shell pair) unless d_cutmat is given.
 * Device pointers, asynchronous on stream; needs qf_prepare. */
int qf_fock_build_os_device(qf_engine *e, const double *d_Pa, const double *d_Pb, double *d_Fa, double *d_Fb,
                            const double *d_cutmat, void *stream);

/* Host convenience of the open-shell build: Fa += G_a, Fb += G_b; synchronous. */
int qf_fock_build_os_host(qf_engine *e, const double *Pa, const double *Pb, double *Fa, double *Fb,
                          const double *cutmat);

/* Device pointer of the engine's G (= Gt + Gt^T after the last build). */
const double *qf_device_g(qf_engine *e);

/* One-shot batched build: F_i += G(P_i) for every engine of the list with
 * one launch per quartet class covering all of them (a temporary qf_batch).
 * d_P[i] / d_F[i] are device pointers per engine (d_F may be NULL); all
 * engines live on one device.  Synchronises the stream before returning. */
int qf_fock_build_batch(qf_engine *const *engines, int count, const double *const *d_P,
                        double *const *d_F, void *stream);

/* Persistent batches (native; no QUICK counterpart; used for many-molecule
 * SCF iterations and the serving loop).  A batch takes prepared engines
 * (it prepares them if needed) and owns pooled device buffers: P_i, F_i,
 * G_i and Cutmatrix_i per molecule.  qf_batch_build adds G(P_i) into F_i
 * for every molecule with one launch per quartet class; with
 * qf_batch_set_graph(b, 1) the first build is captured into a HIP graph
 * and later builds replay it.  The engines must stay alive and unchanged
 * while the batch exists.  G replicas: 1..64 copies of each G. */
typedef struct qf_batch qf_batch;
int qf_batch_create(qf_engine *const *engines, int count, int replicas, qf_batch **out);
/* Same, with caller-owned device buffers: molecule i reads P from d_P[i] and
 * adds G into d_F[i] (both nbasis_i^2, alive as long as the batch). */
int qf_batch_create_ext(qf_engine *const *engines, int count, int replicas, const double *const *d_P,
                        double *const *d_F, qf_batch **out);
void qf_batch_destroy(qf_batch *b);
int qf_batch_size(qf_batch *b);
double *qf_batch_density(qf_batch *b, int i);   /* device, nbasis_i^2, written by the caller */
double *qf_batch_fock(qf_batch *b, int i);      /* device, nbasis_i^2, F_i += G_i per build */
const double *qf_batch_g(qf_batch *b, int i);   /* device, nbasis_i^2, G_i of the last build */
int qf_batch_set_graph(qf_batch *b, int enable);
int qf_batch_set_density_host(qf_batch *b, int i, const double *P);   /* synchronous upload */
int qf_batch_get_host(qf_batch *b, int i, double *G, double *F);       /* synchronous; either may be NULL */
int qf_batch_clear_fock(qf_batch *b, void *stream);                     /* F_i = 0 for every molecule */
int qf_batch_build(qf_batch *b, void *stream);
int qf_batch_launch_count(qf_batch *b);
int qf_batch_launch_info(qf_batch *b, int idx, int *cb, int *ck, int *fp32, int *ntasks);
int qf_batch_set_timing(qf_batch *b, int enable);   /* per-launch events; builds run eagerly */
int qf_batch_read_timing(qf_batch *b, double *ms_per_launch, int *calls_per_launch);

/* One-electron matrices for the SCF harness: S, T, V (nbasis*nbasis, host)
 * with nuclear charges[natom]; computed on the GPU. */
int qf_one_electron(qf_engine *e, const double *charges, double *S, double *T, double *V);

/* Same with an explicit list of point charges (the atoms and any external
 * charges, QUICK's allxyz/allchg): V = -sum_C Z_C <mu|1/|r-C||nu>.
 * center_xyz[3*ncenter] in bohr. */
int qf_one_electron_ext(qf_engine *e, int ncenter, const double *center_xyz, const double *charges, double *S,
                        double *T, double *V);

/* Algorithmic FLOPs of a build with the density of the last build: the
 * quartets and primitive quartets that pass the screens, counted on the
 * device, priced with the generated per-class FLOP table. */
double qf_flops_estimate(qf_engine *e);

/* Per-launch introspection: a build issues one kernel launch per
 * (bra class, ket class, precision) that has work. */
int qf_launch_count(qf_engine *e);
int qf_launch_info(qf_engine *e, int idx, int *bra_class, int *ket_class, int *fp32, int *ntasks);
int qf_launch_flops(qf_engine *e, double *flops);
int qf_launch_counts(qf_engine *e, long long *quartets, long long *prim_quartets);

/* Digestion tuning: the FP64 atomics of concurrent waves are spread over
 * `replicas` copies of G (summed when the build finishes), and tasks inside
 * a launch are permuted so that concurrent waves touch different kets. */
int qf_set_tuning(qf_engine *e, int replicas, int shuffle);

/* Exchange digestion through LDS tiles (native; no QUICK counterpart):
 * enable = 1 accumulates J_ab and all exchange terms of a bra chunk in a
 * workgroup-shared tile and adds it to G once, instead of one FP64 atomic
 * per quartet.  waves: waves per workgroup (1..4); tasks_per_group: tasks
 * one workgroup takes (0, the default: 4 up to 256 basis functions, 16
 * above); lds_bytes: tile budget (1 KiB..64 KiB).  Classes
 * whose tile does not fit fall back to direct atomics.  With enable = 1 the
 * d-heavy classes (La+Lb+Lc+Ld >= 6; >= 5 up to 128 basis functions, >= 4 up to 48)
 * default to the wave-per-quartet kernel;
 * enable = 0 runs every class without a qf_set_class_tuning entry on the
 * per-lane direct-atomic kernels. */
int qf_set_tile(qf_engine *e, int enable, int waves, int tasks_per_group, int lds_bytes);

/* Per-class launch shape (native; used by the autotuner): for the quartet
 * class (cb|ck), pair classes 0..9 = ss, ps, pp, ds, dp, dd, fs, fp, fd, ff
 * with ck <= cb: tile (-1 engine default, 0 direct atomics, 1 LDS tile, 2
 * one quartet per wave with the recursions in LDS - d-heavy classes; others
 * use the tile), waves per tile workgroup, tasks per tile workgroup (with
 * tile = 2: kets per task, 1..64, default 1 up to 256 basis functions and 8
 * above) and bras per task (0: default). */
int qf_set_class_tuning(qf_engine *e, int cb, int ck, int tile, int waves, int tasks_per_group,
                        int bras_per_task);

/* Persistent launches: at most max_workgroups workgroups per class launch,
 * each striding over the tasks (0: one workgroup per task / tile group). */
int qf_set_persistent(qf_engine *e, int max_workgroups);
/* Primitive pieces (no QUICK counterpart: a launch-shape choice).  Each task
 * sums at most `piece` bra and `piece` ket primitive pairs of its quartets, so
 * the most contracted quartets of a small molecule are spread over several
 * waves; the pieces add up to the same Fock matrix.  piece = 0: automatic (1
 * up to 96 basis functions, 2 up to 128, 4 up to 256, 12 above); piece < 0:
 * never split. */
int qf_set_prim_split(qf_engine *e, int piece);
/* Per-class primitive piece size (class indices as qf_set_class_tuning):
 * 0 follows qf_set_prim_split, < 0 never splits the class. */
int qf_set_class_piece(qf_engine *e, int cb, int ck, int piece);
/* Open-shell form of a class (native tuning, no QUICK counterpart): 0 one
 * fused launch that digests each integral with both densities, the second
 * exchange's per-task terms held in registers; 2 (and -1, the default) the
 * same, lean: on the LDS-tile kernel (FP64) those terms are re-read and
 * added to the tile every quartet instead, so the integral loop keeps the
 * closed-shell register count (other kernels take form 0); 1 two
 * closed-shell launches, the total density into G and the spin density,
 * exchange terms only, into the second accumulator (each integral computed
 * twice, but no second set of exchange state per lane). */
int qf_set_class_os(qf_engine *e, int cb, int ck, int mode);

/* HIP-event timing of every class launch (on the build stream); read sums
 * the durations per launch index since the last read. */
int qf_set_timing(qf_engine *e, int enable);
int qf_read_timing(qf_engine *e, double *ms_per_launch, int *calls_per_launch);

/* ------------------------------------------------------------------ */
/* QUICK-compatible layer (src/gpu/hip/gpu.cu, src/gpu/gpu_get2e_getxc_drivers.h) */
/* ------------------------------------------------------------------ */
/* gpu.cu:197 gpu_new_ */
void gpu_new_(int *ierr);
/* gpu.cu:184 gpu_set_device_ */
void gpu_set_device_(int *gpu_dev_id, int *ierr);
/* gpu.cu:281 gpu_init_device_ */
void gpu_init_device_(int *ierr);
/* gpu.cu:437 gpu_allocate_scratch_ (nothing to allocate: kernels keep no global scratch) */
void gpu_allocate_scratch_(bool *allocate_gradient_scratch);
/* gpu.cu:496 gpu_deallocate_scratch_ */
void gpu_deallocate_scratch_(bool *deallocate_gradient_scratch);
/* gpu.cu:514 gpu_delete_ */
void gpu_delete_(int *ierr);
/* gpu.cu:722 gpu_upload_method_ (0 = HF, hyb forced to 1) */
void gpu_upload_method_(int *quick_method, bool *is_oshell, double *hyb_coeff);
/* gpu.cu:539 gpu_setup_ */
void gpu_setup_(int *natom, int *nbasis, int *nElec, int *imult, int *molchg, int *iAtomType);
/* gpu.cu:776 gpu_upload_xyz_ (xyz(3, natom), bohr) */
void gpu_upload_xyz_(double *atom_xyz);
/* gpu.cu:820 gpu_upload_atom_and_chg_ */
void gpu_upload_atom_and_chg_(int *atom, double *atom_chg);
/* gpu.cu:1782 gpu_upload_basis_ (Fortran 1-based indices; gccoeff/gcexpo (MAXPRIM=14, nbasis)) */
void gpu_upload_basis_(int *nshell, int *nprim, int *jshell, int *jbasis, int *maxcontract,
                       int *ncontract, int *itype, double *aexp, double *dcoeff,
                       int *first_basis_function, int *last_basis_function,
                       int *first_shell_basis_function, int *last_shell_basis_function, int *ncenter,
                       int *kstart, int *katom, int *ktype, int *kprim, int *kshell, int *Ksumtype,
                       int *Qnumber, int *Qstart, int *Qfinal, int *Qsbasis, int *Qfbasis,
                       double *gccoeff, double *cons, double *gcexpo, int *KLMN);
/* gpu.cu:879 gpu_upload_cutoff_matrix_ (YCutoff(nshell,nshell), cutPrim(jbasis,jbasis)) */
void gpu_upload_cutoff_matrix_(double *YCutoff, double *cutPrim);
/* gpu.cu:1763 gpu_upload_density_matrix_ (the density of the next build) */
void gpu_upload_density_matrix_(double *dense);
/* gpu.cu:1490 gpu_upload_oei_ (external point charges: extxyz(3, nextatom), extchg(nextatom)) */
void gpu_upload_oei_(int *nextatom, double *extxyz, double *extchg, int *ierr);
/* gpu_get2e_getxc_drivers.h:360 gpu_get_oei_: o += V, the nuclear attraction of the atoms and the
 * external charges (the kinetic part stays on the CPU, quick_oei_module.f90:95-125) */
void gpu_get_oei_(double *o);
/* gpu.cu:405 gpu_get_device_info_, with the argument list hip/gpu_write_info.f90:33-34 passes
 * (device name and architecture name are Fortran character(len=256)) */
void gpu_get_device_info_(int *gpu_dev_count, int *gpu_dev_id, int *gpu_dev_mem, int *gpu_num_proc,
                          double *gpu_core_freq, char *gpu_dev_name, int *name_len, char *gpu_arch_name,
                          int *arch_len, int *majorv, int *minorv, int *ierr);
/* gpu.cu:1690 gpu_upload_calculated_ (only dense is read) */
void gpu_upload_calculated_(double *o, double *co, double *vec, double *dense);
/* gpu.cu:840 gpu_upload_cutoff_ */
void gpu_upload_cutoff_(double *cutMatrix, double *integralCutoff, double *primLimit, double *DMCutoff,
                        double *coreIntegralCutoff);
/* gpu_get2e_getxc_drivers.h:27 gpu_get_cshell_eri_: o += G(dense), both triangles */
void gpu_get_cshell_eri_(bool *deltaO, double *o);
/* gpu.cu:2136 gpu_upload_grad_: QUICK's gradCutoff for the next gradient */
void gpu_upload_grad_(double *gradCutoff);
/* gpu_get2e_getxc_drivers.h:127 gpu_get_cshell_eri_grad_: grad(3, natom) += d/dR of the
 * two-electron energy of the density of gpu_upload_density_matrix_ (s, p, d shells) */
void gpu_get_cshell_eri_grad_(double *grad);
/* gpu.cu:1771 gpu_upload_beta_density_matrix_ (the beta density of the next open-shell build) */
void gpu_upload_beta_density_matrix_(double *denseb);
/* gpu.cu:1728 gpu_upload_calculated_beta_ (only denseb is read) */
void gpu_upload_calculated_beta_(double *ob, double *denseb);
/* gpu_get2e_getxc_drivers.h:25 gpu_get_oshell_eri_ (called from quick_uscf_operator_module.f90:128):
 * o += J(dense + denseb) - hyb K(dense), ob += J(dense + denseb) - hyb K(denseb), both triangles */
void gpu_get_oshell_eri_(bool *deltaO, double *o, double *ob);
/* gpu.cu:2882 gpu_cleanup_ */
void gpu_cleanup_(void);

#ifdef __cplusplus
}
#endif
#endif
