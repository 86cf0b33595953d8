// QUICK-compatible C ABI: the gpu_*_ entry points QUICK's Fortran calls for
// the closed-shell two-electron path, as a thin shim over the native qf_*
// API.  Names, argument order, by-reference passing and Fortran layouts are
// QUICK's (src/gpu/hip/gpu.cu, src/gpu/gpu_get2e_getxc_drivers.h); the state
// they share is a process-wide singleton, like QUICK's `gpu` object
// (src/gpu/hip/gpu_type.h:511-512).  Not re-entrant, one molecule at a time.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "../../include/quickfock.h"

namespace {

// QUICK's MAXPRIM for s-d builds (src/modules/quick_size_module.f90:39-43,
// src/gpu/gpu_common.h:37-40): leading dimension of gccoeff and gcexpo.
constexpr int kMaxPrim = 14;

struct QuickState {
    qf_engine* eng = nullptr;
    int dev = 0;
    int natom = 0, nbasis = 0;
    double hyb = 1.0;
    double cutoff = 1e-7, primlim = 1e-8;
    bool cut_set = false;
    bool oshell = false;
    int nshell = 0, jbasis = 0;
    std::vector<double> dense, denseb, cutmat;
    std::vector<double> xyz, chg;          // atoms (bohr) and nuclear charges
    std::vector<double> ext_xyz, ext_chg;  // external point charges (gpu_upload_oei_)
};

QuickState* g_state = nullptr;

void report(const char* where)
{
    std::fprintf(stderr, "quickfock: %s failed: %s\n", where, qf_last_error());
}

// The Fock build has no error argument in QUICK's interface; continuing
// would hand the SCF a Fock matrix without its two-electron part.
[[noreturn]] void fatal(const char* where)
{
    report(where);
    std::fflush(stderr);
    std::exit(EXIT_FAILURE);
}

}  // namespace

extern "C" {

void gpu_new_(int* ierr)
{
    if (!g_state) g_state = new QuickState();
    if (ierr) *ierr = 0;
}

void gpu_set_device_(int* gpu_dev_id, int* ierr)
{
    if (!g_state) g_state = new QuickState();
    g_state->dev = *gpu_dev_id;
    if (ierr) *ierr = 0;
}

void gpu_init_device_(int* ierr)
{
    if (!g_state) g_state = new QuickState();
    int rc = qf_engine_create(g_state->dev, &g_state->eng);
    // QUICK's codes: 24 no GPU, 25 bad device id (gpu.cu:297-339)
    if (ierr) *ierr = rc == QF_OK ? 0 : (rc == QF_ERR_NODEVICE ? 24 : 25);
}

void gpu_allocate_scratch_(bool*) {}
void gpu_deallocate_scratch_(bool*) {}

void gpu_upload_method_(int* quick_method, bool* is_oshell, double* hyb_coeff)
{
    if (!g_state) return;
    g_state->oshell = is_oshell && *is_oshell;
    g_state->hyb = (*quick_method == 0) ? 1.0 : *hyb_coeff;   // HF: full exchange (gpu.cu:724-727)
}

void gpu_setup_(int* natom, int* nbasis, int* /*nElec*/, int* /*imult*/, int* /*molchg*/, int* /*iAtomType*/)
{
    if (!g_state) return;
    g_state->natom = *natom;
    g_state->nbasis = *nbasis;
    g_state->dense.assign((size_t)*nbasis * *nbasis, 0.0);
}

void gpu_upload_xyz_(double* atom_xyz)
{
    if (!g_state || !g_state->eng) return;
    g_state->xyz.assign(atom_xyz, atom_xyz + 3 * (size_t)g_state->natom);
    if (qf_set_molecule(g_state->eng, g_state->natom, atom_xyz)) report("gpu_upload_xyz_");
}

void gpu_upload_atom_and_chg_(int* /*atom*/, double* atom_chg)
{
    if (!g_state) return;
    g_state->chg.assign(atom_chg, atom_chg + g_state->natom);
}

void gpu_upload_oei_(int* nextatom, double* extxyz, double* extchg, int* ierr)
{
    if (ierr) *ierr = 0;
    if (!g_state) {
        if (ierr) *ierr = 1;
        return;
    }
    const int m = nextatom ? *nextatom : 0;
    g_state->ext_xyz.assign(extxyz, extxyz + 3 * (size_t)std::max(m, 0));
    g_state->ext_chg.assign(extchg, extchg + std::max(m, 0));
}

void gpu_get_oei_(double* o)
{
    // QUICK adds the nuclear attraction of atoms and external charges into o
    // (gpu_get2e_getxc_drivers.h:360-392); the kinetic part is the CPU's.
    if (!g_state || !g_state->eng || g_state->chg.size() != (size_t)g_state->natom) {
        std::fprintf(stderr, "quickfock: gpu_get_oei_ called before gpu_upload_xyz_/gpu_upload_atom_and_chg_\n");
        std::exit(EXIT_FAILURE);
    }
    std::vector<double> cx(g_state->xyz), cq(g_state->chg);
    cx.insert(cx.end(), g_state->ext_xyz.begin(), g_state->ext_xyz.end());
    cq.insert(cq.end(), g_state->ext_chg.begin(), g_state->ext_chg.end());
    const size_t n2 = (size_t)g_state->nbasis * g_state->nbasis;
    std::vector<double> S(n2), T(n2), V(n2);
    if (qf_one_electron_ext(g_state->eng, (int)cq.size(), cx.data(), cq.data(), S.data(), T.data(), V.data()))
        fatal("gpu_get_oei_");
    for (size_t k = 0; k < n2; ++k) o[k] += V[k];
}

void gpu_upload_density_matrix_(double* dense)
{
    if (!g_state) return;
    g_state->dense.assign(dense, dense + (size_t)g_state->nbasis * g_state->nbasis);
}

void gpu_upload_beta_density_matrix_(double* denseb)
{
    if (!g_state) return;
    g_state->denseb.assign(denseb, denseb + (size_t)g_state->nbasis * g_state->nbasis);
}

void gpu_get_device_info_(int* gpu_dev_count, int* gpu_dev_id, int* gpu_dev_mem, int* gpu_num_proc,
                          double* gpu_core_freq, char* gpu_dev_name, int* name_len, char* gpu_arch_name,
                          int* arch_len, int* majorv, int* minorv, int* ierr)
{
    if (ierr) *ierr = 0;
    *gpu_dev_count = 0;
    if (hipGetDeviceCount(gpu_dev_count) != hipSuccess || *gpu_dev_count == 0) {
        *gpu_dev_count = 0;
        if (ierr) *ierr = 24;
        return;
    }
    const int dev = g_state ? g_state->dev : 0;
    *gpu_dev_id = dev;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, dev) != hipSuccess) {
        if (ierr) *ierr = 25;
        return;
    }
    *gpu_dev_mem = (int)(prop.totalGlobalMem / (1024 * 1024));
    *gpu_num_proc = prop.multiProcessorCount;
    int khz = 0;
    (void)hipDeviceGetAttribute(&khz, hipDeviceAttributeClockRate, dev);
    *gpu_core_freq = khz * 1e-6;
    // Fortran character(len=256) buffers: blank padded, no terminator needed
    const int cap = 256;
    const int nl = (int)strnlen(prop.name, cap), al = (int)strnlen(prop.gcnArchName, cap);
    std::memset(gpu_dev_name, ' ', cap);
    std::memcpy(gpu_dev_name, prop.name, nl);
    *name_len = nl;
    std::memset(gpu_arch_name, ' ', cap);
    std::memcpy(gpu_arch_name, prop.gcnArchName, al);
    *arch_len = al;
    *majorv = prop.major;
    *minorv = prop.minor;
}

void gpu_upload_basis_(int* nshell, int* /*nprim*/, int* /*jshell*/, int* jbasis, int* /*maxcontract*/,
                       int* /*ncontract*/, int* /*itype*/, double* /*aexp*/, double* /*dcoeff*/,
                       int* /*first_basis_function*/, int* /*last_basis_function*/,
                       int* /*first_shell_basis_function*/, int* /*last_shell_basis_function*/, int* /*ncenter*/,
                       int* kstart, int* katom, int* /*ktype*/, int* kprim, int* /*kshell*/, int* Ksumtype,
                       int* /*Qnumber*/, int* Qstart, int* Qfinal, int* /*Qsbasis*/, int* /*Qfbasis*/,
                       double* gccoeff, double* /*cons*/, double* gcexpo, int* /*KLMN*/)
{
    if (!g_state || !g_state->eng) return;
    const int ns = *nshell, jb = *jbasis;
    std::vector<int> atom(ns), lmin(ns), lmax(ns), np(ns), ps(ns), fs(ns);
    std::vector<double> ex(jb), co((size_t)jb * 4, 0.0);
    for (int s = 0; s < ns; ++s) {
        atom[s] = katom[s] - 1;
        lmin[s] = Qstart[s];
        lmax[s] = Qfinal[s];
        np[s] = kprim[s];
        ps[s] = kstart[s] - 1;
        fs[s] = Ksumtype[s] - 1;
        for (int k = 0; k < kprim[s]; ++k) {
            // gcexpo(k, Ksumtype(s)); gccoeff(k, Ksumtype(s)+L) as eri.fh:131-140
            ex[ps[s] + k] = gcexpo[(size_t)fs[s] * kMaxPrim + k];
            for (int L = Qstart[s]; L <= Qfinal[s]; ++L)
                co[(size_t)(ps[s] + k) * 4 + L] = gccoeff[(size_t)(fs[s] + L) * kMaxPrim + k];
        }
    }
    g_state->nshell = ns;
    g_state->jbasis = jb;
    if (qf_set_basis(g_state->eng, ns, atom.data(), lmin.data(), lmax.data(), np.data(), ps.data(), fs.data(), jb,
                     ex.data(), co.data(), g_state->nbasis))
        report("gpu_upload_basis_");
}

void gpu_upload_cutoff_matrix_(double* YCutoff, double* cutPrim)
{
    if (!g_state || !g_state->eng) return;
    if (qf_set_schwarz(g_state->eng, YCutoff, cutPrim)) report("gpu_upload_cutoff_matrix_");
}

void gpu_upload_calculated_(double* /*o*/, double* /*co*/, double* /*vec*/, double* dense)
{
    if (!g_state) return;
    std::memcpy(g_state->dense.data(), dense, g_state->dense.size() * sizeof(double));
}

void gpu_upload_calculated_beta_(double* /*ob*/, double* denseb)
{
    // only denseb is read: ob is added to by gpu_get_oshell_eri_
    gpu_upload_beta_density_matrix_(denseb);
}

void gpu_upload_cutoff_(double* cutMatrix, double* integralCutoff, double* primLimit, double* /*DMCutoff*/,
                        double* /*coreIntegralCutoff*/)
{
    if (!g_state || !g_state->eng) return;
    if (!g_state->cut_set || *integralCutoff != g_state->cutoff || *primLimit != g_state->primlim) {
        g_state->cutoff = *integralCutoff;
        g_state->primlim = *primLimit;
        g_state->cut_set = true;
        if (qf_set_cutoffs(g_state->eng, g_state->cutoff, g_state->primlim, g_state->hyb)) report("gpu_upload_cutoff_");
    }
    g_state->cutmat.assign(cutMatrix, cutMatrix + (size_t)g_state->nshell * g_state->nshell);
}

void gpu_get_cshell_eri_(bool* /*deltaO*/, double* o)
{
    // deltaO needs no handling: the incremental density already arrives in
    // dense (quick_scf_operator_module.f90:69-79); o += G(dense) as QUICK.
    if (!g_state || !g_state->eng) {
        std::fprintf(stderr, "quickfock: gpu_get_cshell_eri_ called before gpu_new_/gpu_upload_basis_\n");
        std::exit(EXIT_FAILURE);
    }
    if (g_state->oshell) {
        std::fprintf(stderr, "quickfock: gpu_get_cshell_eri_ called for an open-shell method (use gpu_get_oshell_eri_)\n");
        std::exit(EXIT_FAILURE);
    }
    if (!g_state->cut_set) {
        qf_set_cutoffs(g_state->eng, g_state->cutoff, g_state->primlim, g_state->hyb);
        g_state->cut_set = true;
    }
    // pair and task lists are rebuilt only after a cutoff or geometry change
    if (qf_prepare(g_state->eng) != QF_OK) fatal("gpu_get_cshell_eri_ (prepare)");
    const double* cm = g_state->cutmat.empty() ? nullptr : g_state->cutmat.data();
    if (qf_fock_build_host(g_state->eng, g_state->dense.data(), o, cm)) fatal("gpu_get_cshell_eri_");
}

void gpu_get_oshell_eri_(bool* /*deltaO*/, double* o, double* ob)
{
    // o += J(dense + denseb) - hyb K(dense), ob += J(dense + denseb) - hyb K(denseb)
    // (quick_uscf_operator_module.f90:96-128); with deltaO the incremental
    // densities arrive in dense and denseb, as for the closed shell.
    if (!g_state || !g_state->eng) {
        std::fprintf(stderr, "quickfock: gpu_get_oshell_eri_ called before gpu_new_/gpu_upload_basis_\n");
        std::exit(EXIT_FAILURE);
    }
    if (g_state->denseb.size() != g_state->dense.size()) {
        std::fprintf(stderr, "quickfock: gpu_get_oshell_eri_ called before gpu_upload_calculated_beta_\n");
        std::exit(EXIT_FAILURE);
    }
    if (!g_state->cut_set) {
        qf_set_cutoffs(g_state->eng, g_state->cutoff, g_state->primlim, g_state->hyb);
        g_state->cut_set = true;
    }
    if (qf_prepare(g_state->eng) != QF_OK) fatal("gpu_get_oshell_eri_ (prepare)");
    const double* cm = g_state->cutmat.empty() ? nullptr : g_state->cutmat.data();
    if (qf_fock_build_os_host(g_state->eng, g_state->dense.data(), g_state->denseb.data(), o, ob, cm))
        fatal("gpu_get_oshell_eri_");
}

void gpu_upload_grad_(double* gradCutoff)
{
    // QUICK allocates the device gradient here; the shim only records the
    // cutoff (a lower one reaches further down the pair lists at the next prepare)
    if (!g_state || !g_state->eng) return;
    if (qf_set_grad_cutoff(g_state->eng, *gradCutoff)) report("gpu_upload_grad_");
}

void gpu_get_cshell_eri_grad_(double* grad)
{
    // grad(3, natom) += d/dR of the two-electron energy of dense, with the
    // Cutmatrix of the last gpu_upload_cutoff_ (gradient.fh:575-584, 867)
    if (!g_state || !g_state->eng) {
        std::fprintf(stderr, "quickfock: gpu_get_cshell_eri_grad_ called before gpu_new_/gpu_upload_basis_\n");
        std::exit(EXIT_FAILURE);
    }
    if (g_state->oshell) {
        std::fprintf(stderr, "quickfock: gpu_get_cshell_eri_grad_ called for an open-shell method\n");
        std::exit(EXIT_FAILURE);
    }
    if (!g_state->cut_set) {
        qf_set_cutoffs(g_state->eng, g_state->cutoff, g_state->primlim, g_state->hyb);
        g_state->cut_set = true;
    }
    if (qf_prepare(g_state->eng) != QF_OK) fatal("gpu_get_cshell_eri_grad_ (prepare)");
    const double* cm = g_state->cutmat.empty() ? nullptr : g_state->cutmat.data();
    if (qf_eri_grad_host(g_state->eng, g_state->dense.data(), grad, cm)) fatal("gpu_get_cshell_eri_grad_");
}

void gpu_cleanup_(void) {}

void gpu_delete_(int* ierr)
{
    if (g_state) {
        qf_engine_destroy(g_state->eng);
        delete g_state;
        g_state = nullptr;
    }
    if (ierr) *ierr = 0;
}

}  // extern "C"
