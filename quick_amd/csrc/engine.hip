// Host runtime of the Fock-build engine (native C ABI, include/quickfock.h).
//
// Replaces QUICK's GPU host side for the two-electron path:
//   gpu_upload_basis_ Q-shell split and sort   (src/gpu/hip/gpu.cu:1782-2134)
//   gpu_upload_cutoff_matrix_ pair-list build  (src/gpu/hip/gpu.cu:879-1484)
//   get2e launches                             (src/gpu/hip/gpu_get2e.cu:595-638)
//   download / symmetrise / accumulate         (src/gpu/gpu_get2e_getxc_drivers.h:27-121)
// but keeps everything device-resident: pair data and task lists are built
// once per geometry (O(n log n) sorts), a build only uploads nothing, screens
// the density on the device, launches one kernel per class and symmetrises.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <numeric>
#include <string>
#include <vector>

#include "../../include/quickfock.h"
#include "gen/grad_launchers.h"
#include "gen/launchers.h"
#include "qf_types.h"

namespace {

thread_local std::string g_err;

int fail(int code, const std::string& msg)
{
    g_err = msg;
    return code;
}

#define QF_HIP(call)                                                                              \
    do {                                                                                          \
        hipError_t _e = (call);                                                                   \
        if (_e != hipSuccess)                                                                     \
            return fail(QF_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(_e));         \
    } while (0)

constexpr double kPi = 3.14159265358979323846264338327950288;
const double kX0 = 2.0 * std::pow(kPi, 2.5);   // QUICK's X0 (quick_constants_module.f90:61)

inline int pair_class(int La, int Lb) { return La * (La + 1) / 2 + Lb; }
const char* const kDirectFp32Refused =
    "the direct-atomic kernels run FP64 only (their FP32 form gave wrong Fock matrices on gfx950 and is not "
    "built); FP32 halves run on the LDS-tile kernel";
constexpr int kCoopWindow = 8;   // kets per task of the wave-per-quartet kernels
constexpr int kNPC = 10;   // pair classes (La >= Lb) for s, p, d, f
constexpr size_t kMaxTileLds = 64 * 1024;   // dynamic LDS of one tile workgroup
const int kClassL[kNPC][2] = {{0, 0}, {1, 0}, {1, 1}, {2, 0}, {2, 1}, {2, 2}, {3, 0}, {3, 1}, {3, 2}, {3, 3}};
inline int ncart(int L) { return (L + 1) * (L + 2) / 2; }

template <typename T>
struct DevBuf {
    T* p = nullptr;
    size_t n = 0;
    ~DevBuf() { release(); }
    void release()
    {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
    hipError_t resize(size_t count)
    {
        if (count <= n && p) return hipSuccess;
        release();
        if (count == 0) return hipSuccess;
        hipError_t e = hipMalloc(&p, count * sizeof(T));
        if (e == hipSuccess) n = count;
        return e;
    }
    hipError_t upload(const std::vector<T>& v)
    {
        hipError_t e = resize(v.size());
        if (e != hipSuccess || v.empty()) return e;
        return hipMemcpy(p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice);
    }
    size_t bytes() const { return n * sizeof(T); }
};

struct QShell {
    int L, atom, shell, fn, nprim, pstart, qfin;
};

struct HostBlock {
    int cls = 0, K = 0;
    std::vector<qf::PairInfo> info;
    std::vector<int> bperm;                      // pair indices grouped by first Q-shell
    std::vector<std::pair<int, int>> groups;     // [start, end) in bperm, one per first Q-shell
    std::vector<double> prim;   // PF_N * K * n, filled at finalisation
    std::vector<std::vector<double>> rows;   // per pair: PF_N*K values (field-major)
    std::vector<int> colf;                   // compact column -> basis function
};

// --------------------------------------------------------------------------
// Device helper kernels
// --------------------------------------------------------------------------
__global__ void k_cutmat(const double* __restrict__ P, int n, int nshell, const int* __restrict__ f0,
                         const int* __restrict__ f1, double* __restrict__ cm)
{
    // QUICK's cshell_density_cutoff (cutoff.fh:65-92): max |P_ij| over the
    // functions of the two original shells.
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= nshell * nshell) return;
    const int I = idx / nshell, J = idx % nshell;
    double m = 0.0;
    for (int i = f0[I]; i <= f1[I]; ++i)
        for (int j = f0[J]; j <= f1[J]; ++j) m = fmax(m, fabs(P[i * n + j]));
    cm[idx] = m;
}

// QUICK's oshell_density_cutoff (cutoff.fh:16-89): max over the block of
// max(|Pa+Pb|, |Pa|, |Pb|).
__global__ void k_cutmat_os(const double* __restrict__ Pa, const double* __restrict__ Pb, int n, int nshell,
                            const int* __restrict__ f0, const int* __restrict__ f1, double* __restrict__ cm)
{
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= nshell * nshell) return;
    const int I = idx / nshell, J = idx % nshell;
    double m = 0.0;
    for (int i = f0[I]; i <= f1[I]; ++i)
        for (int j = f0[J]; j <= f1[J]; ++j) {
            const double a = Pa[i * n + j], b = Pb[i * n + j];
            m = fmax(m, fmax(fabs(a + b), fmax(fabs(a), fabs(b))));
        }
    cm[idx] = m;
}

// Total and spin densities of the open-shell build: Pt = Pa + Pb, Ps = Pa - Pb.
__global__ void k_spin(const double* __restrict__ Pa, const double* __restrict__ Pb, double* __restrict__ Pt,
                       double* __restrict__ Ps, int n2)
{
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= n2) return;
    const double a = Pa[idx], b = Pb[idx];
    Pt[idx] = a + b;
    Ps[idx] = a - b;
}

__global__ void k_zero(double* __restrict__ p, size_t count)
{
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < count; i += (size_t)gridDim.x * blockDim.x)
        p[i] = 0.0;
}

__global__ void k_finalize(const double* __restrict__ Gt, int replicas, double* __restrict__ Gs,
                           double* __restrict__ F, int n)
{
    // G = sum over copies of (Gt + Gt^T); F += G
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= n * n) return;
    const int i = idx / n, j = idx % n;
    const size_t n2 = (size_t)n * n;
    double g = 0.0;
    for (int r = 0; r < replicas; ++r) g += Gt[r * n2 + i * n + j] + Gt[r * n2 + j * n + i];
    Gs[idx] = g;
    if (F) F[idx] += g;
}

// Open-shell pass: A = sum(Gt + Gt^T) (J and exchange of Pt), B = sum(Gt2 +
// Gt2^T) (exchange of Ps); Fa += A + B, Fb += A - B, Gs = A.
__global__ void k_finalize_os(const double* __restrict__ Gt, const double* __restrict__ Gt2, int replicas,
                              double* __restrict__ Gs, double* __restrict__ Fa, double* __restrict__ Fb, int n)
{
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= n * n) return;
    const int i = idx / n, j = idx % n;
    const size_t n2 = (size_t)n * n;
    double a = 0.0, b = 0.0;
    for (int r = 0; r < replicas; ++r) {
        a += Gt[r * n2 + i * n + j] + Gt[r * n2 + j * n + i];
        b += Gt2[r * n2 + i * n + j] + Gt2[r * n2 + j * n + i];
    }
    Gs[idx] = a;
    Fa[idx] += a + b;
    Fb[idx] += a - b;
}

// Counts, per quartet class, the quartets and primitive quartets that pass
// the screens (same tests as the Fock kernel); used for the FLOP model.
// grad = 1: the gradient's screen (md.cutoff is gradCutoff, quartets on one
// atom skipped).
__global__ void k_count(const qf::Task* __restrict__ tasks, int ntasks, const qf::PairBlock* __restrict__ blocks,
                        const qf::MolDesc* __restrict__ mols, unsigned long long* __restrict__ counts, int slot,
                        int fp32, int grad)
{
    const int tix = blockIdx.x;
    if (tix >= ntasks) return;
    const qf::Task t = tasks[tix];
    const qf::PairBlock bb = blocks[t.bra_block];
    const qf::PairBlock kb = blocks[t.ket_block];
    const qf::MolDesc md = mols[bb.mol];
    const int lane = threadIdx.x;
    if (lane >= t.ket_count) return;
    const int j = t.ket_start + lane;
    const qf::PairInfo ki = kb.info[j];
    const int ns = md.nshell;
    const double* cm = md.cutmat;
    unsigned long long nq = 0, np = 0;
    for (int kk = t.kb; kk < t.ke; ++kk) {
        const int i = bb.bperm[kk];
        const qf::PairInfo bi = bb.info[i];
        double dn = 4.0 * cm[bi.sa * ns + bi.sb];
        dn = fmax(dn, 4.0 * cm[ki.sa * ns + ki.sb]);
        dn = fmax(dn, cm[bi.sa * ns + ki.sb]);
        dn = fmax(dn, cm[bi.sa * ns + ki.sa]);
        dn = fmax(dn, cm[bi.sb * ns + ki.sa]);
        dn = fmax(dn, cm[bi.sb * ns + ki.sb]);
        const double y = bi.ycut * ki.ycut;
        if (!(y > md.cutoff && y * dn > md.cutoff)) continue;
        if (t.same_block && j > i) continue;
        if (fp32 ? !(y < md.split) : !(t.all_fp64 || y >= md.split)) continue;
        if (grad) {
            const int a = md.shell_atom[bi.sa];
            if (a == md.shell_atom[bi.sb] && a == md.shell_atom[ki.sa] && a == md.shell_atom[ki.sb]) continue;
        }
        if (t.pb == 0 && t.qb == 0) ++nq;   // a quartet spread over several tasks counts once
        for (int ip = t.pb; ip < t.pe; ++ip) {
            const double c1 = dn * bb.prim[(qf::PF_CUT * bb.K + ip) * bb.n + i];
            for (int jp = t.qb; jp < t.qe; ++jp)
                if (c1 * kb.prim[(qf::PF_CUT * kb.K + jp) * kb.n + j] > md.primlim) ++np;
        }
    }
    atomicAdd(&counts[2 * slot], nq);
    atomicAdd(&counts[2 * slot + 1], np);
}

}  // namespace

// ==========================================================================
struct qf_engine {
    int device = 0;
    // molecule + basis (host)
    int natom = 0;
    std::vector<double> xyz;
    int nshell = 0, nprim = 0, nbasis = 0;
    std::vector<int> s_atom, s_lmin, s_lmax, s_nprim, s_pstart, s_fstart;
    std::vector<double> exps, coef;
    std::vector<QShell> qs;
    // screening
    std::vector<double> ycut, cutprim;
    bool have_schwarz = false;
    double cutoff = 1e-12, primlim = 1e-13, hyb = 1.0;
    double grad_cutoff = -1.0;   // QUICK's gradCutoff (qf_set_grad_cutoff); <= 0: the integral cutoff
    bool grad_lane = true;       // per-lane gradient kernels where a class has one (qf_set_grad_kernel)
    // pair lists and task windows cover every quartet either screen can keep
    double list_cutoff() const { return grad_cutoff > 0.0 ? std::min(cutoff, grad_cutoff) : cutoff; }
    int prec = QF_PREC_FP64;
    double split = 1e-4;
    int split_lmin = -1;     // classes with La+Lb+Lc+Ld >= split_lmin are split in mixed builds (-1: by size)
    double split_thr = 0.0;  // Ycut_ij*Ycut_kl below which quartets run in FP32
    int replicas = 8;        // copies of G the atomics are spread over
    int shuffle = 1;         // permute tasks inside a launch (decorrelates atomic targets)
    int tile = 1;            // accumulate exchange terms in LDS tiles (fock_tile_kernel)
    int tile_waves = 4;      // max waves per tile workgroup
    int tile_tasks = 0;      // max tasks per tile workgroup (0: 4 up to 256 functions, 16 above)
    int lds_budget = 64 * 1024;
    int persist = 0;         // grid cap of the class launches (0: one workgroup per task / group)
    int prim_piece = 0;      // primitive pairs per task piece: 0 auto, < 0 never split
    struct ClassTune {
        int tile = -1, waves = 0, tpg = 0, mmax = 0;   // <0 / 0: engine default
        int piece = 0;                                  // primitive pairs per piece: 0 engine default, < 0 none
        int os = -1;   // open-shell form: -1 default (2), 0 fused, 1 two passes, 2 fused lean (tile kernels)
        int grad = -1; // gradient kernel: -1 engine default, 0 wave-per-quartet, 1 per-lane
        int split = -1; // mixed builds: -1 engine rule (split_lmin), 0 all FP64, 1 split
    };
    ClassTune ctune[kNPC][kNPC];
    // pair blocks
    std::vector<HostBlock> blocks;
    std::vector<qf::PairBlock> hblocks;
    std::vector<qf::Task> tasks;
    std::vector<qf::TileGroup> groups;
    struct Launch {
        int cb, ck, fp32, offset, count;
        int tile, goffset, gcount, waves;
        size_t lds;
        size_t lds_os;   // open-shell pass with a second exchange tile (0: it does not fit)
        int os_split;    // open-shell builds run this class as two closed-shell passes
        int os_lean;     // ... or, on the tile kernel, in the fused lean form
    };
    std::vector<Launch> launches;
    bool prepared = false;
    unsigned long long generation = 0;   // bumped by every prepare(): batches check it
    bool built = false;                  // a build has run since the last prepare()
    bool count_cm_own = false;           // ... and screened with the engine's own Cutmatrix (d_cm)
    qf_stats stats{};
    double flops_last = -1.0;
    // device
    DevBuf<qf::PairInfo> d_info;
    DevBuf<double> d_prim;
    DevBuf<int> d_perm;
    DevBuf<qf::PairBlock> d_blocks;
    DevBuf<qf::MolDesc> d_mol;
    DevBuf<qf::Task> d_tasks;
    DevBuf<qf::TileGroup> d_groups;
    DevBuf<int> d_colf;
    DevBuf<double> d_Gt, d_Gs, d_cm, d_P;
    DevBuf<double> d_Pb, d_Pt, d_Ps, d_Fa, d_Fb;   // open-shell builds (allocated on first use)
    DevBuf<double> d_Gt2;                           // their second exchange accumulators (replicas x n^2)
    DevBuf<double> d_zero;                          // nbasis^2 zeros: Coulomb density of an exchange-only pass
    DevBuf<int> d_f0, d_f1;
    DevBuf<unsigned long long> d_counts;
    DevBuf<double> d_cscratch;   // global slots of the wave-per-quartet classes that need them
    DevBuf<int> d_shell_atom;    // original shell -> atom (gradient)
    DevBuf<double> d_gscratch;   // global slots of the gradient classes that need them (on first use)
    DevBuf<double> d_grad;       // gradient of the host entry point (on first use)
    bool grad_built = false;     // a gradient has run since the last prepare() (counts need it) ...
    bool grad_cm_own = false;    // ... screened with the engine's own Cutmatrix
    qf::MolDesc mol[4]{};   // slot 0: the build (or open-shell pass 1), slot 1: open-shell pass 2, slot 2: no-alloc build, slot 3: gradient
    // per-launch timing (HIP events on the build stream)
    bool timing = false;
    std::vector<hipEvent_t> ev_pool;
    size_t ev_used = 0;
    std::vector<int> ev_launch;        // launch index of each recorded (start, stop) pair
    ~qf_engine()
    {
        for (hipEvent_t ev : ev_pool) (void)hipEventDestroy(ev);
    }

    int set_device() { return hipSetDevice(device) == hipSuccess ? QF_OK : fail(QF_ERR_BADDEVICE, "hipSetDevice"); }
};

namespace {

// Q-shells: an SP shell becomes an s and a p Q-shell (gpu.cu:1884-1887).
void build_qshells(qf_engine* e)
{
    e->qs.clear();
    for (int s = 0; s < e->nshell; ++s) {
        int off = 0;
        for (int L = e->s_lmin[s]; L <= e->s_lmax[s]; ++L) {
            e->qs.push_back({L, e->s_atom[s], s, e->s_fstart[s] + off, e->s_nprim[s], e->s_pstart[s], e->s_lmax[s]});
            off += ncart(L);
        }
    }
}

double cutprim_sym(const qf_engine* e, int pa, int pb)
{
    // QUICK fills cutprim for shell order II <= JJ only; the value is
    // symmetric, so take whichever triangle holds it.
    const size_t jb = e->nprim;
    return std::max(e->cutprim[pa * jb + pb], e->cutprim[pb * jb + pa]);
}

// One canonical pair (La >= Lb; equal L: qa >= qb).  If ip >= 0 the pair
// holds the single primitive pair (ip, jp) (cutprim evaluation).
void make_pair(const qf_engine* e, int qa, int qb, int ip, int jp, bool use_cut, qf::PairInfo& info,
               std::vector<double>& row, int& K)
{
    const QShell& a = e->qs[qa];
    const QShell& b = e->qs[qb];
    const double* A = &e->xyz[3 * a.atom];
    const double* B = &e->xyz[3 * b.atom];
    info.qa = qa;
    info.qb = qb;
    info.fa = a.fn;
    info.fb = b.fn;
    info.sa = a.shell;
    info.sb = b.shell;
    info.qfs = a.qfin + e->s_lmax[b.shell];
    info.id = 0;
    info.ycut = e->have_schwarz ? e->ycut[(size_t)a.shell * e->nshell + b.shell] : 1.0;
    double AB2 = 0.0;
    for (int d = 0; d < 3; ++d) {
        info.A[d] = A[d];
        info.AB[d] = A[d] - B[d];
        AB2 += info.AB[d] * info.AB[d];
    }
    info.cola = info.colb = 0;
    struct Prim {
        double v[qf::PF_N];
    };
    std::vector<Prim> prims;
    const double sx0 = std::sqrt(kX0);
    for (int i = 0; i < a.nprim; ++i)
        for (int j = 0; j < b.nprim; ++j) {
            if (ip >= 0 && (i != ip || j != jp)) continue;
            const int pa = a.pstart + i, pb = b.pstart + j;
            const double al = e->exps[pa], be = e->exps[pb];
            const double z = al + be;
            Prim p;
            p.v[qf::PF_ZETA] = z;
            p.v[qf::PF_OOZ2] = 0.5 / z;
            for (int d = 0; d < 3; ++d) p.v[qf::PF_PX + d] = (A[d] * al + B[d] * be) / z;
            // Kpri * c_a * c_b (eri.fh:120-140) with sqrt(X0) per pair, so a
            // quartet carries X0 once (eri.fh:514).
            const double kpri = std::exp(-al * be / z * AB2) / z;
            p.v[qf::PF_KC] = sx0 * kpri * e->coef[(size_t)pa * 4 + a.L] * e->coef[(size_t)pb * 4 + b.L];
            p.v[qf::PF_CUT] = use_cut ? cutprim_sym(e, pa, pb) : 1.0;
            p.v[qf::PF_ALPHA] = al;
            prims.push_back(p);
        }
    std::stable_sort(prims.begin(), prims.end(),
                     [](const Prim& x, const Prim& y) { return x.v[qf::PF_CUT] > y.v[qf::PF_CUT]; });
    K = (int)prims.size();
    row.assign((size_t)qf::PF_N * K, 0.0);
    for (int f = 0; f < qf::PF_N; ++f)
        for (int k = 0; k < K; ++k) row[(size_t)f * K + k] = prims[k].v[f];
}

// Assemble host blocks into flat device arrays; fills hblocks (device pointers).
int upload_blocks(qf_engine* e, std::vector<HostBlock>& blocks, DevBuf<qf::PairInfo>& d_info,
                  DevBuf<double>& d_prim, std::vector<qf::PairBlock>& out, int mol, DevBuf<int>* d_perm = nullptr,
                  DevBuf<int>* d_colf = nullptr)
{
    size_t ninfo = 0, nprimv = 0;
    for (auto& b : blocks) {
        ninfo += b.info.size();
        nprimv += (size_t)qf::PF_N * b.K * b.info.size();
    }
    std::vector<qf::PairInfo> info;
    info.reserve(ninfo);
    std::vector<double> prim(nprimv);
    std::vector<size_t> ioff, poff;
    size_t po = 0;
    for (auto& b : blocks) {
        ioff.push_back(info.size());
        poff.push_back(po);
        const size_t n = b.info.size();
        info.insert(info.end(), b.info.begin(), b.info.end());
        for (size_t j = 0; j < n; ++j)
            for (int f = 0; f < qf::PF_N; ++f)
                for (int k = 0; k < b.K; ++k) prim[po + ((size_t)f * b.K + k) * n + j] = b.rows[j][(size_t)f * b.K + k];
        po += (size_t)qf::PF_N * b.K * n;
    }
    QF_HIP(d_info.upload(info));
    QF_HIP(d_prim.upload(prim));
    std::vector<int> perm;
    std::vector<size_t> permoff;
    for (auto& b : blocks) {
        permoff.push_back(perm.size());
        perm.insert(perm.end(), b.bperm.begin(), b.bperm.end());
    }
    if (d_perm) QF_HIP(d_perm->upload(perm));
    std::vector<int> colf;
    std::vector<size_t> coloff;
    for (auto& b : blocks) {
        coloff.push_back(colf.size());
        colf.insert(colf.end(), b.colf.begin(), b.colf.end());
    }
    if (d_colf) QF_HIP(d_colf->upload(colf));
    out.clear();
    for (size_t k = 0; k < blocks.size(); ++k) {
        qf::PairBlock pb;
        pb.info = d_info.p + ioff[k];
        pb.prim = d_prim.p + poff[k];
        pb.bperm = (d_perm && !blocks[k].bperm.empty()) ? d_perm->p + permoff[k] : nullptr;
        pb.colf = (d_colf && !blocks[k].colf.empty()) ? d_colf->p + coloff[k] : nullptr;
        pb.ncol = (int)blocks[k].colf.size();
        pb.pad = 0;
        pb.n = (int)blocks[k].info.size();
        pb.K = blocks[k].K;
        pb.cls = blocks[k].cls;
        pb.mol = mol;
        out.push_back(pb);
    }
    return QF_OK;
}

// Group pairs by (class, K).
void add_to_blocks(std::vector<HostBlock>& blocks, int cls, int K, const qf::PairInfo& info, std::vector<double>&& row)
{
    for (auto& b : blocks)
        if (b.cls == cls && b.K == K) {
            b.info.push_back(info);
            b.rows.push_back(std::move(row));
            return;
        }
    HostBlock b;
    b.cls = cls;
    b.K = K;
    b.info.push_back(info);
    b.rows.push_back(std::move(row));
    blocks.push_back(std::move(b));
}

template <class F>
void for_each_canonical_pair(const qf_engine* e, F&& f)
{
    const int nq = (int)e->qs.size();
    for (int a = 0; a < nq; ++a)
        for (int b = 0; b < nq; ++b) {
            const int La = e->qs[a].L, Lb = e->qs[b].L;
            if (La < Lb || (La == Lb && a < b)) continue;
            f(a, b);
        }
}

// Schwarz bounds on the GPU with the diagonal-class kernels:
//   Ycutoff(I,J) = sqrt(max |(ij|ij)|) over the Q-shell pairs of (I,J)
//   cutprim(p,q) = sqrt(X0 * max |(pq|pq)|) for single primitive pairs,
// as QUICK's schwarzoff (quick_cutoff_module.f90:54-588).
int compute_schwarz(qf_engine* e)
{
    e->have_schwarz = false;
    e->ycut.assign((size_t)e->nshell * e->nshell, 0.0);
    e->cutprim.assign((size_t)e->nprim * e->nprim, 0.0);
    for (int pass = 0; pass < 2; ++pass) {
        std::vector<HostBlock> blocks;
        std::vector<int> owner_a, owner_b;   // per block, per pair: Q-shells / prims
        std::vector<std::vector<int>> keys;
        for_each_canonical_pair(e, [&](int a, int b) {
            if (pass == 0) {
                qf::PairInfo info;
                std::vector<double> row;
                int K;
                make_pair(e, a, b, -1, -1, false, info, row, K);
                info.id = a * (int)e->qs.size() + b;
                add_to_blocks(blocks, pair_class(e->qs[a].L, e->qs[b].L), K, info, std::move(row));
            } else {
                for (int i = 0; i < e->qs[a].nprim; ++i)
                    for (int j = 0; j < e->qs[b].nprim; ++j) {
                        qf::PairInfo info;
                        std::vector<double> row;
                        int K;
                        make_pair(e, a, b, i, j, false, info, row, K);
                        info.id = (e->qs[a].pstart + i) * e->nprim + (e->qs[b].pstart + j);
                        add_to_blocks(blocks, pair_class(e->qs[a].L, e->qs[b].L), K, info, std::move(row));
                    }
            }
        });
        DevBuf<qf::PairInfo> dinfo;
        DevBuf<double> dprim;
        std::vector<qf::PairBlock> hb;
        int rc = upload_blocks(e, blocks, dinfo, dprim, hb, 0);
        if (rc) return rc;
        size_t total = 0;
        for (auto& b : blocks) total += b.info.size();
        DevBuf<double> dout;
        QF_HIP(dout.resize(total));
        // global slot buffer of the wave-per-quartet diagonals that need one
        size_t gdoubles = 0;
        for (auto& b : blocks) gdoubles = std::max<size_t>(gdoubles, (size_t)qf::coop_global_doubles(b.cls, b.cls));
        DevBuf<double> dgs;
        if (gdoubles) QF_HIP(dgs.resize(gdoubles * qf::COOP_GLOBAL_SLICES));
        size_t off = 0;
        for (size_t k = 0; k < blocks.size(); ++k) {
            qf::SchwarzLaunch L = qf::schwarz_launcher(blocks[k].cls);
            L(hb[k], dout.p + off, dgs.p, 0);
            off += blocks[k].info.size();
        }
        QF_HIP(hipGetLastError());
        std::vector<double> out(total);
        QF_HIP(hipMemcpy(out.data(), dout.p, total * sizeof(double), hipMemcpyDeviceToHost));
        off = 0;
        for (auto& b : blocks)
            for (auto& info : b.info) {
                const double v = out[off++];
                if (pass == 0) {
                    double& y = e->ycut[(size_t)info.sa * e->nshell + info.sb];
                    y = std::max(y, v);
                    e->ycut[(size_t)info.sb * e->nshell + info.sa] = y;
                } else {
                    const int pa = info.id / e->nprim, pb = info.id % e->nprim;
                    double& c = e->cutprim[(size_t)pa * e->nprim + pb];
                    c = std::max(c, v);
                    e->cutprim[(size_t)pb * e->nprim + pa] = c;
                }
            }
    }
    for (auto& y : e->ycut) y = std::sqrt(y);
    for (auto& c : e->cutprim) c = std::sqrt(kX0 * c);
    e->have_schwarz = true;
    return QF_OK;
}

// Upper bound of kets j (sorted by ycut desc) with ycut_i*ycut_j > thr.
int prefix_count(const std::vector<qf::PairInfo>& kets, double yi, double thr, bool strict)
{
    int lo = 0, hi = (int)kets.size();
    while (lo < hi) {
        const int mid = (lo + hi) / 2;
        const double y = yi * kets[mid].ycut;
        if (strict ? (y > thr) : (y >= thr))
            lo = mid + 1;
        else
            hi = mid;
    }
    return lo;
}

int prepare(qf_engine* e)
{
    if (e->prepared) return QF_OK;   // idempotent: setters clear the flag
    if (e->nshell == 0 || e->natom == 0) return fail(QF_ERR_STATE, "molecule and basis must be set");
    int rc = e->set_device();
    if (rc) return rc;
    if (!e->have_schwarz) {
        rc = compute_schwarz(e);
        if (rc) return rc;
    }
    double ymax = 0.0;
    for (double y : e->ycut) ymax = std::max(ymax, y);

    // ---- pair blocks: canonical pairs that can meet some partner -----------
    e->blocks.clear();
    int next_id = 0;
    for_each_canonical_pair(e, [&](int a, int b) {
        const double y = e->ycut[(size_t)e->qs[a].shell * e->nshell + e->qs[b].shell];
        if (!(y * ymax > e->list_cutoff())) return;
        qf::PairInfo info;
        std::vector<double> row;
        int K;
        make_pair(e, a, b, -1, -1, true, info, row, K);
        info.id = next_id++;
        add_to_blocks(e->blocks, pair_class(e->qs[a].L, e->qs[b].L), K, info, std::move(row));
    });
    // order: by class, then K; pairs inside a block by ycut descending
    std::sort(e->blocks.begin(), e->blocks.end(),
              [](const HostBlock& x, const HostBlock& y) { return x.cls != y.cls ? x.cls < y.cls : x.K < y.K; });
    for (auto& b : e->blocks) {
        std::vector<size_t> ord(b.info.size());
        std::iota(ord.begin(), ord.end(), 0);
        std::stable_sort(ord.begin(), ord.end(), [&](size_t x, size_t y) { return b.info[x].ycut > b.info[y].ycut; });
        std::vector<qf::PairInfo> info2;
        std::vector<std::vector<double>> rows2;
        for (size_t k : ord) {
            info2.push_back(b.info[k]);
            rows2.push_back(std::move(b.rows[k]));
        }
        b.info.swap(info2);
        b.rows.swap(rows2);
    }
    for (auto& b : e->blocks) {
        // bra order: grouped by first Q-shell a, Schwarz-descending inside a group
        const int nb = (int)b.info.size();
        b.bperm.resize(nb);
        std::iota(b.bperm.begin(), b.bperm.end(), 0);
        std::stable_sort(b.bperm.begin(), b.bperm.end(),
                         [&](int x, int y) { return b.info[x].qa < b.info[y].qa; });
        b.groups.clear();
        for (int k = 0; k < nb;) {
            int e2 = k;
            while (e2 < nb && b.info[b.bperm[e2]].qa == b.info[b.bperm[k]].qa) ++e2;
            b.groups.push_back({k, e2});
            k = e2;
        }
    }
    for (auto& b : e->blocks) {
        // compact columns: the functions of every Q-shell of the block, by Q-shell id
        std::vector<int> qlist;
        for (const auto& p : b.info) {
            qlist.push_back(p.qa);
            qlist.push_back(p.qb);
        }
        std::sort(qlist.begin(), qlist.end());
        qlist.erase(std::unique(qlist.begin(), qlist.end()), qlist.end());
        std::vector<int> qcol(e->qs.size(), -1);
        b.colf.clear();
        for (int q : qlist) {
            qcol[q] = (int)b.colf.size();
            for (int f = 0; f < ncart(e->qs[q].L); ++f) b.colf.push_back(e->qs[q].fn + f);
        }
        for (auto& p : b.info) {
            p.cola = qcol[p.qa];
            p.colb = qcol[p.qb];
        }
    }
    rc = upload_blocks(e, e->blocks, e->d_info, e->d_prim, e->hblocks, 0, &e->d_perm, &e->d_colf);
    if (rc) return rc;
    for (auto& b : e->blocks) b.rows.clear();
    QF_HIP(e->d_blocks.upload(e->hblocks));

    // ---- wave tasks per quartet class and precision ------------------------
    const double thr_hi = e->split * ymax * ymax;
    e->split_thr = thr_hi;
    e->tasks.clear();
    e->groups.clear();
    e->launches.clear();
    e->stats = qf_stats{};
    for (auto& b : e->blocks) e->stats.n_pairs += (long long)b.info.size();
    // Primitive pieces: a wave's lanes each run the whole primitive loop of
    // one quartet, so the most contracted quartets set the tail of a launch
    // (and, for small molecules with few waves, its whole time).  The
    // primitive pairs of a task are cut into pieces of at most 1 up to 96
    // functions, 2 up to 128, 4 up to 256 and 12 above (e.g. 36 -> 3 x 12).
    // Measured on (H2O)32 6-31G(d): no split 17.0 ms, pieces of 6 / 12 / 18:
    // 15.8 / 15.4 / 15.9 ms; on QUICK's 69 regression inputs (12-210
    // functions), geometric mean of the build times up to 32 functions with
    // pieces of 1 / 2: 0.39 / 0.55 ms, 33-64: 0.93 / 1.23 ms, 79 functions
    // 2.20 / 2.49 ms, 114: 2.02 / 1.50 ms, 210 (pieces of 1 / 4): 5.53 /
    // 4.57 ms.  A small molecule has too few quartets to fill the GPU, and
    // one primitive quartet per lane gives it the most waves.
    const int piece = e->prim_piece > 0   ? e->prim_piece
                      : e->prim_piece < 0 ? 1 << 20
                      : e->nbasis <= 96   ? 1
                      : e->nbasis <= 128  ? 2
                      : e->nbasis <= 256  ? 4
                                          : 12;
    for (int cb = 0; cb < kNPC; ++cb)
        for (int ck = 0; ck <= cb; ++ck)
            for (int fp32 = 0; fp32 < 2; ++fp32) {
                const int offset = (int)e->tasks.size();
                const int goffset = (int)e->groups.size();
                const int na = ncart(kClassL[cb][0]), nb = ncart(kClassL[cb][1]);
                // bras per task: long loops for cheap classes (more register
                // accumulation, fewer atomics), short ones for expensive classes
                // (enough waves to fill 256 CUs)
                const int ni = na * nb * ncart(kClassL[ck][0]) * ncart(kClassL[ck][1]);
                const qf_engine::ClassTune& ct = e->ctune[cb][ck];
                int mmax = ni <= 10 ? 16 : ni <= 36 ? 8 : ni <= 108 ? 4 : ni <= 324 ? 2 : 1;
                // up to 32 functions at most 2 bras per task: more, shorter
                // tasks for launches that hold few (QUICK's regression inputs
                // up to 32 functions, geometric mean 16.3x -> 17.8x QUICK's
                // saved build times; no gain above)
                if (e->nbasis <= 32) mmax = std::min(mmax, 2);
                if (ct.mmax > 0) mmax = ct.mmax;
                // tasks per tile workgroup: few for molecules up to 256
                // functions, whose launches hold few tasks (QUICK's 69
                // regression inputs, light classes with 1 / 2 / 4 / 8 / 16
                // tasks per group: 15.7x / 16.5x / 16.5x / 13.9x / 10.4x
                // QUICK's saved build times, geometric mean)
                const int tile_tasks = ct.tpg > 0        ? ct.tpg
                                       : e->tile_tasks > 0 ? e->tile_tasks
                                       : e->nbasis <= 256  ? 4
                                                           : 16;
                const int tile_waves = ct.waves > 0 ? ct.waves : e->tile_waves;
                // LDS tile: (na + m*nb) x ncol exchange terms + m*na*nb Coulomb terms
                int maxcol = 0;
                for (const auto& Kb : e->blocks)
                    if (Kb.cls == ck) maxcol = std::max(maxcol, (int)Kb.colf.size());
                const long long room = e->lds_budget / 8 - (long long)na * maxcol;
                const long long mfit = room > 0 ? room / ((long long)nb * maxcol + na * nb) : 0;
                // Mixed precision splits a class only where its FP32 half can run
                // through the LDS-tile kernel: the direct-atomic FP32 kernels of
                // (dp|dp) gave wrong Fock matrices on gfx950 and the (dd|dp) one
                // non-finite values, while CPU builds of the same code are right.
                // (dd|dp) and (dd|dd) stay FP64 (their FP32 forms spill several KB
                // per lane: little to gain).
                // classes without per-lane kernels (La+Lb+Lc+Ld >= 6) run every
                // quartet in FP64 on the wave-per-quartet kernel
                const bool lane_class = qf::fock_launcher(cb, ck) != nullptr;
                // Classes with La+Lb+Lc+Ld >= split_lmin are split (qf_set_split_classes).
                // FP32 quartets compute their Boys function, exponents and
                // recursions in FP32, which shortens the primitive loop (the
                // FP64 Boys function is most of a light class's work) and halves
                // the registers of the spilling d classes.
                const int Lsum_s = kClassL[cb][0] + kClassL[cb][1] + kClassL[ck][0] + kClassL[ck][1];
                // By default every class above 256 functions, and only the
                // classes with La+Lb+Lc+Ld >= 5 below: a small molecule's
                // light classes hold too few quartets for the second launch
                // to pay (256 batched molecules: 1389 builds/s splitting all,
                // FP64 1432).
                const int lmin = e->split_lmin >= 0 ? e->split_lmin : e->nbasis <= 256 ? 5 : 0;
                const bool split_class = e->prec == QF_PREC_MIXED && mfit >= 1 && lane_class &&
                                         (ct.split >= 0 ? ct.split == 1 : Lsum_s >= lmin);
                if (fp32 && !split_class) continue;
                // wave-per-quartet kernel (FP64 launches of classes that have one)
                // wave-per-quartet kernel: where the class tuning asks for it, and by
                // default (unless qf_set_tile turned the shared-memory kernels off)
                // for the d-heavy classes (La+Lb+Lc+Ld >= 6; >= 5 up to 128 functions and >= 4 up to 48,
                // where few waves run and latency decides), whose per-lane
                // kernels spill several KB per lane: one such quartet alone on a
                // lane costs milliseconds, which dominates the build of a small
                // molecule
                const int Lsum = kClassL[cb][0] + kClassL[cb][1] + kClassL[ck][0] + kClassL[ck][1];
                // (QUICK's regression inputs of 49-128 functions run their
                // La+Lb+Lc+Ld = 4 classes faster on the tile kernel: PO4 6-31G*
                // 15.8x -> 18.0x, PSB5 6-31G 26x -> 34x QUICK's build times;
                // up to 48 functions and at 5 the wave-per-quartet kernel wins)
                const int coop_lmin = e->nbasis <= 48 ? 4 : e->nbasis <= 128 ? 5 : 6;
                const bool want_coop = !lane_class || ct.tile == 2 || (ct.tile < 0 && e->tile != 0 && Lsum >= coop_lmin);
                const bool coop = !fp32 && want_coop && qf::fock_coop_launcher(cb, ck) != nullptr;
                if (!coop && !lane_class) return fail(QF_ERR_STATE, "class has neither per-lane nor wave-per-quartet kernels");
                const bool tile = !coop && (fp32 || (ct.tile >= 0 ? ct.tile != 0 : e->tile != 0)) && mfit >= 1;
                if (tile) mmax = (int)std::min<long long>(mmax, mfit);
                // kets per task: a full wave for the per-lane kernels; for the
                // wave-per-quartet kernel a window (the class tuning's
                // tasks_per_group): small windows give more workgroups to
                // classes with few quartets
                const int coop_win = e->nbasis <= 256 ? 1 : kCoopWindow;
                const int win = coop ? (ct.tpg > 0 ? std::min(ct.tpg, qf::WAVE) : coop_win) : qf::WAVE;
                size_t lds = 0, lds_os = 0;
                long long group_tasks = 0;
                for (size_t bbk = 0; bbk < e->blocks.size(); ++bbk) {
                    const HostBlock& B = e->blocks[bbk];
                    if (B.cls != cb) continue;
                    for (size_t kbk = 0; kbk < e->blocks.size(); ++kbk) {
                        const HostBlock& Kb = e->blocks[kbk];
                        if (Kb.cls != ck) continue;
                        if (cb == ck && kbk > bbk) continue;
                        const bool same = (cb == ck && kbk == bbk);
                        std::vector<std::pair<int, int>> chunks;
                        for (const auto& g : B.groups)
                            for (int c0 = g.first; c0 < g.second; c0 += mmax)
                                chunks.push_back({c0, std::min(g.second, c0 + mmax)});
                        for (const auto& g : chunks) {
                            // per bra of the group: ket range [lo, hi) of this precision
                            const int M = g.second - g.first;
                            std::vector<int> lo(M), hi(M);
                            int nmax = 0;
                            for (int k = 0; k < M; ++k) {
                                const int i = B.bperm[g.first + k];
                                const double yi = B.info[i].ycut;
                                int n = prefix_count(Kb.info, yi, e->list_cutoff(), true);
                                if (same) n = std::min(n, i + 1);
                                if (!fp32) e->stats.n_quartets_bound += n;
                                int n64 = n;
                                if (split_class) n64 = std::min(n, prefix_count(Kb.info, yi, thr_hi, false));
                                lo[k] = fp32 ? n64 : 0;
                                hi[k] = fp32 ? n : n64;
                                nmax = std::max(nmax, hi[k]);
                            }
                            const int chunk_first = (int)e->tasks.size();
                            for (int s0 = 0; s0 < nmax; s0 += win) {
                                int kb = -1, ke = -1;
                                for (int k = 0; k < M; ++k)
                                    if (lo[k] < hi[k] && lo[k] < s0 + win && hi[k] > s0) {
                                        if (kb < 0) kb = k;
                                        ke = k + 1;
                                    }
                                if (kb < 0) continue;
                                qf::Task t{};
                                t.bra_block = (int)bbk;
                                t.kb = g.first + kb;
                                t.ke = g.first + ke;
                                t.ket_block = (int)kbk;
                                t.ket_start = s0;
                                t.ket_count = std::min(win, nmax - s0);
                                t.same_block = same ? 1 : 0;
                                t.all_fp64 = split_class ? 0 : 1;
                                // highly contracted pairs: spread the primitive
                                // pairs of the window over several tasks
                                const int cpiece = ct.piece > 0 ? ct.piece : ct.piece < 0 ? 1 << 20 : piece;
                                const int nbp = (B.K + cpiece - 1) / cpiece, nkp = (Kb.K + cpiece - 1) / cpiece;
                                const int sb = (B.K + nbp - 1) / nbp, sk = (Kb.K + nkp - 1) / nkp;
                                for (int pbi = 0; pbi < B.K; pbi += sb)
                                    for (int qbi = 0; qbi < Kb.K; qbi += sk) {
                                        t.pb = pbi;
                                        t.pe = std::min(B.K, pbi + sb);
                                        t.qb = qbi;
                                        t.qe = std::min(Kb.K, qbi + sk);
                                        e->tasks.push_back(t);
                                    }
                            }
                            if (!tile) continue;
                            // workgroups of the chunk: consecutive runs of its tasks
                            const int chunk_end = (int)e->tasks.size();
                            for (int t0 = chunk_first; t0 < chunk_end; t0 += tile_tasks) {
                                qf::TileGroup tg{};
                                tg.task_begin = t0;
                                tg.task_end = std::min(chunk_end, t0 + tile_tasks);
                                tg.chunk_kb = g.first;
                                tg.nbra = M;
                                e->groups.push_back(tg);
                                group_tasks += tg.task_end - tg.task_begin;
                            }
                            if (chunk_end > chunk_first) {
                                const size_t kt = (size_t)(na + M * nb) * Kb.colf.size();
                                lds = std::max(lds, sizeof(double) * (kt + (size_t)M * na * nb));
                                lds_os = std::max(lds_os, sizeof(double) * (2 * kt + (size_t)M * na * nb));
                            }
                        }
                    }
                }
                const int count = (int)e->tasks.size() - offset;
                const int gcount = (int)e->groups.size() - goffset;
                // work of a task: bras x kets x primitive pairs (bra and ket);
                // without the primitive factor the longest-first order misjudges
                // the contracted classes (13 % of the (H2O)32 build)
                auto cost = [](const qf::Task& x) {
                    return (long long)(x.ke - x.kb) * x.ket_count * (x.pe - x.pb) * (x.qe - x.qb);
                };
                if (tile && gcount > 0) {
                    // longest workgroups first so the launch ends evenly
                    std::vector<long long> gc(gcount, 0);
                    for (int q = 0; q < gcount; ++q) {
                        const auto& tg = e->groups[goffset + q];
                        for (int x = tg.task_begin; x < tg.task_end; ++x) gc[q] += cost(e->tasks[x]);
                    }
                    std::vector<int> ord(gcount);
                    std::iota(ord.begin(), ord.end(), 0);
                    std::stable_sort(ord.begin(), ord.end(), [&](int x, int y) { return gc[x] > gc[y]; });
                    std::vector<qf::TileGroup> sorted(gcount);
                    for (int q = 0; q < gcount; ++q) sorted[q] = e->groups[goffset + ord[q]];
                    std::copy(sorted.begin(), sorted.end(), e->groups.begin() + goffset);
                } else if (e->shuffle && count > 1) {
                    // deterministic permutation (concurrent waves then hit different
                    // kets), then longest tasks first so the launch ends evenly
                    uint64_t st = 0x9E3779B97F4A7C15ull ^ (uint64_t)(cb * 12 + ck * 2 + fp32);
                    for (int q = count - 1; q > 0; --q) {
                        st ^= st << 13;
                        st ^= st >> 7;
                        st ^= st << 17;
                        const int r = (int)(st % (uint64_t)(q + 1));
                        std::swap(e->tasks[offset + q], e->tasks[offset + r]);
                    }
                    std::stable_sort(e->tasks.begin() + offset, e->tasks.end(),
                                     [&](const qf::Task& x, const qf::Task& y) { return cost(x) > cost(y); });
                }
                if (count > 0) {
                    qf_engine::Launch L{};
                    L.cb = cb;
                    L.ck = ck;
                    L.fp32 = fp32;
                    L.offset = offset;
                    L.count = count;
                    L.tile = coop ? 2 : (tile && gcount > 0) ? 1 : 0;
                    L.goffset = goffset;
                    L.gcount = gcount;
                    L.waves = gcount > 0 ? (int)std::min<long long>(tile_waves,
                                                                  std::max<long long>(1, (group_tasks + gcount - 1) / gcount))
                                         : 1;
                    L.lds = lds;
                    L.lds_os = lds_os <= kMaxTileLds ? lds_os : 0;
                    L.os_split = ct.os == 1;
                    L.os_lean = ct.os == 2 || ct.os < 0;
                    e->launches.push_back(L);
                }
                if (fp32) e->stats.n_tasks_fp32 += count; else e->stats.n_tasks_fp64 += count;
            }
    // every tile workgroup must stay inside its chunk and its LDS allocation
    for (const auto& L : e->launches) {
        if (L.tile != 1) continue;
        const int na = ncart(kClassL[L.cb][0]), nb = ncart(kClassL[L.cb][1]);
        for (int q = L.goffset; q < L.goffset + L.gcount; ++q) {
            const qf::TileGroup& g = e->groups[q];
            if (g.task_begin < L.offset || g.task_end > L.offset + L.count || g.task_begin >= g.task_end)
                return fail(QF_ERR_STATE, "tile group outside its launch");
            const qf::Task& t0 = e->tasks[g.task_begin];
            const size_t need = sizeof(double) * ((size_t)(na + g.nbra * nb) * e->blocks[t0.ket_block].colf.size() +
                                                  (size_t)g.nbra * na * nb);
            if (need > L.lds) return fail(QF_ERR_STATE, "tile group exceeds its LDS allocation");
            if (L.lds_os && 2 * need - sizeof(double) * (size_t)g.nbra * na * nb > L.lds_os)
                return fail(QF_ERR_STATE, "tile group exceeds its open-shell LDS allocation");
            for (int x = g.task_begin; x < g.task_end; ++x) {
                const qf::Task& t = e->tasks[x];
                if (t.bra_block != t0.bra_block || t.ket_block != t0.ket_block || t.kb < g.chunk_kb ||
                    t.ke > g.chunk_kb + g.nbra || t.kb >= t.ke)
                    return fail(QF_ERR_STATE, "tile group mixes chunks");
            }
        }
    }
    QF_HIP(e->d_tasks.upload(e->tasks));
    QF_HIP(e->d_groups.upload(e->groups));
    // global slot buffer of the wave-per-quartet classes that keep their
    // node-program slots in global memory
    size_t gdoubles = 0;
    for (const auto& L : e->launches)
        if (L.tile == 2) gdoubles = std::max<size_t>(gdoubles, (size_t)qf::coop_global_doubles(L.cb, L.ck));
    if (gdoubles) QF_HIP(e->d_cscratch.resize(gdoubles * qf::COOP_GLOBAL_SLICES));

    // ---- per-build buffers -------------------------------------------------
    const size_t n2 = (size_t)e->nbasis * e->nbasis;
    QF_HIP(e->d_Gt.resize(n2 * e->replicas));
    QF_HIP(e->d_Gs.resize(n2));
    QF_HIP(e->d_P.resize(n2));
    QF_HIP(e->d_cm.resize((size_t)e->nshell * e->nshell));
    std::vector<int> f0(e->nshell), f1(e->nshell);
    for (int s = 0; s < e->nshell; ++s) {
        int nf = 0;
        for (int L = e->s_lmin[s]; L <= e->s_lmax[s]; ++L) nf += ncart(L);
        f0[s] = e->s_fstart[s];
        f1[s] = e->s_fstart[s] + nf - 1;
    }
    QF_HIP(e->d_f0.upload(f0));
    QF_HIP(e->d_f1.upload(f1));
    QF_HIP(e->d_mol.resize(4));   // slots: build / open-shell pass 1, open-shell pass 2, no-alloc build, gradient
    QF_HIP(e->d_shell_atom.upload(e->s_atom));
    QF_HIP(e->d_counts.resize(2 * std::max<size_t>(1, e->launches.size())));
    e->stats.device_bytes = (long long)(e->d_info.bytes() + e->d_prim.bytes() + e->d_tasks.bytes() +
                                        e->d_Gt.bytes() * 3 + e->d_cm.bytes() + e->d_blocks.bytes());
    // launch indices recorded for timing refer to the old launch list
    e->ev_used = 0;
    e->ev_launch.clear();
    e->built = false;
    e->count_cm_own = false;
    e->grad_built = false;
    e->grad_cm_own = false;
    ++e->generation;
    e->prepared = true;
    return QF_OK;
}

// Fixed scratch of qf_fock_build_device_noalloc: G partials, symmetrised G
// and the Cutmatrix of one build live here instead of in buffers the
// library allocates.  Each device holds its own copy for the life of the
// process.  Host code never passes `scratch` itself to a kernel (on the host
// it names a stand-in, not device memory): it uses the address
// hipGetSymbolAddress returns, looked up once per device in
// qf_engine_create with that device current.
constexpr size_t kNoAllocDoubles = QF_NOALLOC_SCRATCH_BYTES / sizeof(double);
__device__ __align__(256) double scratch[kNoAllocDoubles];

constexpr int kMaxDevices = 64;
double* g_scratch_ptr[kMaxDevices] = {};
std::mutex g_scratch_mu;

int noalloc_scratch_init(int device)
{
    if (device < 0 || device >= kMaxDevices) return fail(QF_ERR_BADDEVICE, "device id beyond the no-alloc table");
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (g_scratch_ptr[device]) return QF_OK;
    int prev = 0;
    if (hipGetDevice(&prev) != hipSuccess) return fail(QF_ERR_NODEVICE, "hipGetDevice");
    if (hipSetDevice(device) != hipSuccess) return fail(QF_ERR_BADDEVICE, "hipSetDevice");
    void* p = nullptr;
    const hipError_t err = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    (void)hipSetDevice(prev);
    if (err != hipSuccess || !p) return fail(QF_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(err));
    g_scratch_ptr[device] = static_cast<double*>(p);
    return QF_OK;
}

double* noalloc_scratch_ptr(int device)
{
    if (device < 0 || device >= kMaxDevices) return nullptr;
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    return g_scratch_ptr[device];
}

// One pass over every class launch: G(P) with the density screen dcm, then
// F += G.  `slot` picks the
// device MolDesc, so two passes can be queued back to back on one stream.
// Where a build keeps G: `replicas` copies of nbasis^2 at Gt (the atomics'
// targets) and the symmetrised G at Gs.  The engine's own buffers, or the
// fixed scratch array of the no-alloc entry point.
struct GBufs {
    double* Gt;
    int replicas;
    double* Gs;
    bool timing;   // per-launch HIP events (never in the no-alloc path: it creates events)
    double* Gt2 = nullptr;   // open-shell pass: accumulators of the second exchange density
};

// dP2 != nullptr: the open-shell pass.  Every integral is digested with dP
// (Coulomb and exchange, into Gt) and with dP2 (exchange only, into gb.Gt2);
// then dF += A + B and dF2 += A - B (k_finalize_os).  A class with os_split
// runs as two closed-shell launches instead of the fused one: dP into Gt,
// then dP2 with a zero Coulomb density (e->d_zero) into Gt2 (descriptor slot
// 1).
int build_pass(qf_engine* e, const double* dP, const double* dcm, double* dF, int slot, hipStream_t s,
               const GBufs& gb, const double* dP2 = nullptr, double* dF2 = nullptr)
{
    const int os = dP2 != nullptr;
    if (os && (!gb.Gt2 || !dF || !dF2 || !e->d_zero.p)) return fail(QF_ERR_ARG, "open-shell pass without its buffers");
    const int n = e->nbasis, ns = e->nshell;
    qf::MolDesc md{};
    md.P = dP;
    md.PJ = dP;
    md.G = gb.Gt;
    md.cutmat = dcm;
    md.nbasis = n;
    md.nshell = ns;
    md.cutoff = e->cutoff;
    md.primlim = e->primlim;
    md.hyb = e->hyb;
    md.replicas = gb.replicas;
    md.split = (e->prec == QF_PREC_MIXED) ? e->split_thr : 0.0;
    md.P2 = dP2;
    md.G2 = gb.Gt2;
    e->mol[slot] = md;
    qf::MolDesc* dmol = e->d_mol.p + slot;
    QF_HIP(hipMemcpyAsync(dmol, &e->mol[slot], sizeof(md), hipMemcpyHostToDevice, s));
    qf::MolDesc* dmol_x = nullptr;   // exchange-only pass of the two-pass open-shell classes
    if (os) {
        qf::MolDesc mx = md;
        mx.P = dP2;
        mx.PJ = e->d_zero.p;
        mx.G = gb.Gt2;
        mx.P2 = nullptr;
        mx.G2 = nullptr;
        e->mol[1] = mx;
        dmol_x = e->d_mol.p + 1;
        QF_HIP(hipMemcpyAsync(dmol_x, &e->mol[1], sizeof(mx), hipMemcpyHostToDevice, s));
    }
    hipLaunchKernelGGL(k_zero, dim3(1024), dim3(256), 0, s, gb.Gt, (size_t)n * n * gb.replicas);
    if (os) hipLaunchKernelGGL(k_zero, dim3(1024), dim3(256), 0, s, gb.Gt2, (size_t)n * n * gb.replicas);
    for (size_t li = 0; li < e->launches.size(); ++li) {
        const auto& L = e->launches[li];
        hipEvent_t ea = nullptr, eb = nullptr;
        if (gb.timing) {
            while (e->ev_pool.size() < e->ev_used + 2) {
                hipEvent_t ev;
                QF_HIP(hipEventCreate(&ev));
                e->ev_pool.push_back(ev);
            }
            ea = e->ev_pool[e->ev_used];
            eb = e->ev_pool[e->ev_used + 1];
            e->ev_used += 2;
            e->ev_launch.push_back((int)li);
            QF_HIP(hipEventRecord(ea, s));
        }
        // one class launch with descriptor m; los: its open-shell form
        auto launch = [&](qf::MolDesc* m, int los) -> int {
            if (L.tile == 2) {
                qf::FockCoopLaunch fc = qf::fock_coop_launcher(L.cb, L.ck);
                if (!fc) return fail(QF_ERR_ARG, "no wave-per-quartet kernel for class");
                fc(e->d_tasks.p + L.offset, L.count, e->d_blocks.p, m, los, e->persist, e->d_cscratch.p, s);
            } else if (L.tile) {
                qf::FockTileLaunch ft = qf::fock_tile_launcher(L.cb, L.ck);
                if (!ft) return fail(QF_ERR_ARG, "no tile kernel for class");
                // open shell: a second exchange tile where two fit, else its terms go to Gt2
                const int tos = !los ? 0 : L.lds_os ? 1 : 2;
                ft(e->d_groups.p + L.goffset, L.gcount, e->d_tasks.p, e->d_blocks.p, m, L.fp32,
                   tos | (los && L.os_lean ? 4 : 0), L.waves, tos == 1 ? L.lds_os : L.lds, e->persist, s);
            } else {
                qf::FockLaunch f = qf::fock_launcher(L.cb, L.ck);
                if (!f) return fail(QF_ERR_ARG, "no per-lane kernel for class");
                if (f(e->d_tasks.p + L.offset, L.count, e->d_blocks.p, m, L.fp32, los, e->persist, s))
                    return fail(QF_ERR_ARG, kDirectFp32Refused);
            }
            return QF_OK;
        };
        int rc = QF_OK;
        if (os && L.os_split) {
            rc = launch(dmol, 0);
            if (rc == QF_OK) rc = launch(dmol_x, 0);
        } else {
            rc = launch(dmol, os);
        }
        if (rc) return rc;
        if (gb.timing) QF_HIP(hipEventRecord(eb, s));
    }
    if (os)
        hipLaunchKernelGGL(k_finalize_os, dim3((n * n + 255) / 256), dim3(256), 0, s, gb.Gt, gb.Gt2, gb.replicas,
                           gb.Gs, dF, dF2, n);
    else
        hipLaunchKernelGGL(k_finalize, dim3((n * n + 255) / 256), dim3(256), 0, s, gb.Gt, gb.replicas, gb.Gs,
                           dF, n);
    QF_HIP(hipGetLastError());
    return QF_OK;
}

int build(qf_engine* e, const double* dP, double* dF, const double* dcm, hipStream_t s)
{
    if (!e->prepared) return fail(QF_ERR_STATE, "qf_prepare has not been called");
    const int n = e->nbasis, ns = e->nshell;
    if (!dcm) {
        const int nthr = 256, tot = ns * ns;
        hipLaunchKernelGGL(k_cutmat, dim3((tot + nthr - 1) / nthr), dim3(nthr), 0, s, dP, n, ns, e->d_f0.p,
                           e->d_f1.p, e->d_cm.p);
        dcm = e->d_cm.p;
    }
    const int rc = build_pass(e, dP, dcm, dF, 0, s, GBufs{e->d_Gt.p, e->replicas, e->d_Gs.p, e->timing});
    if (rc == QF_OK) {
        e->built = true;
        e->count_cm_own = (dcm == e->d_cm.p);
    }
    return rc;
}

// Open-shell build (QUICK's get_oshell_eri, gpu_get2e_subs.h OSHELL
// branches): Fa += J(Pt) - hyb K(Pa), Fb += J(Pt) - hyb K(Pb).  With the
// spin density Ps = Pa - Pb:
//   A = J(Pt) - hyb/2 K(Pt)   (the closed-shell G of Pt)
//   B = -hyb/2 K(Ps)          (exchange terms only)
// and Fa += A + B, Fb += A - B.  One pass computes each integral once and
// digests it with both densities (the OS kernel instantiations); it screens
// with QUICK's open-shell density matrix, so it sees exactly the quartets
// QUICK's kernel sees.
int build_os(qf_engine* e, const double* dPa, const double* dPb, double* dFa, double* dFb, const double* dcm,
             hipStream_t s)
{
    if (!e->prepared) return fail(QF_ERR_STATE, "qf_prepare has not been called");
    const int n = e->nbasis, ns = e->nshell, n2 = n * n;
    QF_HIP(e->d_Pt.resize(n2));
    QF_HIP(e->d_Ps.resize(n2));
    hipLaunchKernelGGL(k_spin, dim3((n2 + 255) / 256), dim3(256), 0, s, dPa, dPb, e->d_Pt.p, e->d_Ps.p, n2);
    if (!dcm) {
        const int nthr = 256, tot = ns * ns;
        hipLaunchKernelGGL(k_cutmat_os, dim3((tot + nthr - 1) / nthr), dim3(nthr), 0, s, dPa, dPb, n, ns, e->d_f0.p,
                           e->d_f1.p, e->d_cm.p);
        dcm = e->d_cm.p;
    }
    QF_HIP(e->d_Gt2.resize((size_t)n2 * e->replicas));
    if (e->d_zero.n < (size_t)n2) {
        QF_HIP(e->d_zero.resize(n2));
        hipLaunchKernelGGL(k_zero, dim3(256), dim3(256), 0, s, e->d_zero.p, (size_t)n2);
    }
    GBufs gb{e->d_Gt.p, e->replicas, e->d_Gs.p, e->timing};
    gb.Gt2 = e->d_Gt2.p;
    const int rc = build_pass(e, e->d_Pt.p, dcm, dFa, 0, s, gb, e->d_Ps.p, dFb);
    if (rc == QF_OK) {
        e->built = true;
        e->count_cm_own = (dcm == e->d_cm.p);
    }
    return rc;
}

}  // namespace

// ==========================================================================
// C ABI
// ==========================================================================
extern "C" {

const char* qf_last_error(void) { return g_err.c_str(); }

int qf_engine_create(int device, qf_engine** out)
{
    if (!out) return fail(QF_ERR_ARG, "out is NULL");
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count == 0) return fail(QF_ERR_NODEVICE, "no GPU visible");
    if (device < 0) {
        if (hipGetDevice(&device) != hipSuccess) return fail(QF_ERR_NODEVICE, "hipGetDevice");
    }
    if (device >= count) return fail(QF_ERR_BADDEVICE, "device id out of range");
    const int rc = noalloc_scratch_init(device);
    if (rc) return rc;
    qf_engine* e = new qf_engine();
    e->device = device;
    *out = e;
    return QF_OK;
}

void qf_engine_destroy(qf_engine* e)
{
    if (!e) return;
    (void)hipSetDevice(e->device);
    delete e;
}

int qf_set_molecule(qf_engine* e, int natom, const double* xyz)
{
    if (!e || natom <= 0 || !xyz) return fail(QF_ERR_ARG, "bad molecule");
    e->natom = natom;
    e->xyz.assign(xyz, xyz + 3 * natom);
    e->prepared = false;
    e->have_schwarz = false;
    return QF_OK;
}

int qf_set_basis(qf_engine* e, int nshell, const int* shell_atom, const int* shell_lmin, const int* shell_lmax,
                 const int* shell_nprim, const int* shell_pstart, const int* shell_fstart, int nprim_total,
                 const double* exps, const double* coef, int nbasis)
{
    if (!e || nshell <= 0 || nprim_total <= 0 || nbasis <= 0) return fail(QF_ERR_ARG, "bad basis sizes");
    for (int s = 0; s < nshell; ++s) {
        if (shell_lmin[s] < 0 || shell_lmax[s] > 3 || shell_lmin[s] > shell_lmax[s])
            return fail(QF_ERR_ARG, "angular momentum above f is not supported");
        if (shell_atom[s] < 0 || shell_atom[s] >= e->natom) return fail(QF_ERR_ARG, "shell atom out of range");
        if (shell_pstart[s] < 0 || shell_pstart[s] + shell_nprim[s] > nprim_total || shell_nprim[s] <= 0)
            return fail(QF_ERR_ARG, "shell primitive range out of range");
        int nf = 0;
        for (int L = shell_lmin[s]; L <= shell_lmax[s]; ++L) nf += ncart(L);
        if (shell_fstart[s] < 0 || shell_fstart[s] + nf > nbasis) return fail(QF_ERR_ARG, "function range");
    }
    e->nshell = nshell;
    e->nprim = nprim_total;
    e->nbasis = nbasis;
    e->s_atom.assign(shell_atom, shell_atom + nshell);
    e->s_lmin.assign(shell_lmin, shell_lmin + nshell);
    e->s_lmax.assign(shell_lmax, shell_lmax + nshell);
    e->s_nprim.assign(shell_nprim, shell_nprim + nshell);
    e->s_pstart.assign(shell_pstart, shell_pstart + nshell);
    e->s_fstart.assign(shell_fstart, shell_fstart + nshell);
    e->exps.assign(exps, exps + nprim_total);
    e->coef.assign(coef, coef + 4 * (size_t)nprim_total);
    build_qshells(e);
    e->prepared = false;
    e->have_schwarz = false;
    return QF_OK;
}

int qf_set_cutoffs(qf_engine* e, double integral_cutoff, double prim_limit, double hyb)
{
    if (!e || !(integral_cutoff > 0) || !(prim_limit > 0)) return fail(QF_ERR_ARG, "cutoffs must be > 0");
    e->cutoff = integral_cutoff;
    e->primlim = prim_limit;
    e->hyb = hyb;
    e->prepared = false;
    return QF_OK;
}

int qf_set_precision(qf_engine* e, int mode, double split)
{
    if (!e || (mode != QF_PREC_FP64 && mode != QF_PREC_MIXED) || !(split > 0)) return fail(QF_ERR_ARG, "precision");
    e->prec = mode;
    e->split = split;
    e->prepared = false;
    return QF_OK;
}

int qf_set_split_classes(qf_engine* e, int min_lsum)
{
    if (!e || min_lsum < -1) return fail(QF_ERR_ARG, "min_lsum must be >= -1");
    e->split_lmin = min_lsum;
    e->prepared = false;
    return QF_OK;
}

int qf_set_class_split(qf_engine* e, int cb, int ck, int mode)
{
    if (!e) return fail(QF_ERR_ARG, "NULL");
    if (cb < 0 || cb >= kNPC || ck < 0 || ck > cb) return fail(QF_ERR_ARG, "class index out of range");
    if (mode < -1 || mode > 1) return fail(QF_ERR_ARG, "mode must be -1, 0 (FP64) or 1 (split)");
    e->ctune[cb][ck].split = mode;
    e->prepared = false;
    return QF_OK;
}

int qf_set_schwarz(qf_engine* e, const double* ycutoff, const double* cutprim)
{
    if (!e || e->nshell == 0) return fail(QF_ERR_STATE, "set the basis first");
    e->prepared = false;
    if (!ycutoff || !cutprim) {
        e->have_schwarz = false;
        int rc = e->set_device();
        return rc ? rc : compute_schwarz(e);
    }
    e->ycut.assign(ycutoff, ycutoff + (size_t)e->nshell * e->nshell);
    e->cutprim.assign(cutprim, cutprim + (size_t)e->nprim * e->nprim);
    e->have_schwarz = true;
    return QF_OK;
}

int qf_get_schwarz(qf_engine* e, double* ycutoff, double* cutprim)
{
    if (!e || !e->have_schwarz) return fail(QF_ERR_STATE, "no Schwarz bounds yet");
    if (ycutoff) std::memcpy(ycutoff, e->ycut.data(), e->ycut.size() * sizeof(double));
    if (cutprim) std::memcpy(cutprim, e->cutprim.data(), e->cutprim.size() * sizeof(double));
    return QF_OK;
}

int qf_prepare(qf_engine* e)
{
    if (!e) return fail(QF_ERR_ARG, "engine is NULL");
    return prepare(e);
}

int qf_get_stats(qf_engine* e, qf_stats* out)
{
    if (!e || !out) return fail(QF_ERR_ARG, "NULL");
    *out = e->stats;
    return QF_OK;
}

int qf_fock_build_device(qf_engine* e, const double* d_P, double* d_F, const double* d_cutmat, void* stream)
{
    if (!e || !d_P) return fail(QF_ERR_ARG, "NULL");
    int rc = e->set_device();
    if (rc) return rc;
    return build(e, d_P, d_F, d_cutmat, (hipStream_t)stream);
}

int qf_fock_build_device_noalloc(qf_engine* e, const double* d_P, double* d_F, const double* d_cutmat, void* stream)
{
    if (!e || !d_P) return fail(QF_ERR_ARG, "NULL");
    int rc = e->set_device();
    if (rc) return rc;
    if (!e->prepared) return fail(QF_ERR_STATE, "qf_prepare has not been called");
    double* base = noalloc_scratch_ptr(e->device);
    if (!base) return fail(QF_ERR_STATE, "no-alloc scratch not set up on this device");
    const size_t n2 = (size_t)e->nbasis * e->nbasis, ns2 = (size_t)e->nshell * e->nshell;
    if (2 * n2 + (d_cutmat ? 0 : ns2) > (size_t)kNoAllocDoubles)
        return fail(QF_ERR_ARG, "no-alloc build: nbasis too large for the fixed scratch (2 nbasis^2 + nshell^2 "
                                "doubles must fit in QF_NOALLOC_SCRATCH_BYTES)");
    for (const auto& L : e->launches)
        if (L.tile == 2 && qf::coop_global_doubles(L.cb, L.ck) > 0)
            return fail(QF_ERR_ARG, "no-alloc build: a class keeps its node-program slots in global memory");
    const hipStream_t s = (hipStream_t)stream;
    const int n = e->nbasis, ns = e->nshell;
    // scratch layout: [G partials n^2 | symmetrised G n^2 | Cutmatrix nshell^2]
    double* Gt = base;
    double* Gs = base + n2;
    if (!d_cutmat) {
        double* cm = base + 2 * n2;
        const int nthr = 256, tot = ns * ns;
        hipLaunchKernelGGL(k_cutmat, dim3((tot + nthr - 1) / nthr), dim3(nthr), 0, s, d_P, n, ns, e->d_f0.p,
                           e->d_f1.p, cm);
        d_cutmat = cm;
    }
    return build_pass(e, d_P, d_cutmat, d_F, 2, s, GBufs{Gt, 1, Gs, false});
}

// Two-electron gradient (QUICK's gpu_get_cshell_eri_grad_ -> getGrad,
// gpu_get2e.cu:695-725): one wave-per-quartet gradient launch per class
// over the tasks of the Fock build's launches (their union covers every
// quartet in both precision halves), screened with gradCutoff.
namespace {
constexpr size_t kGradLdsMax = 64 * 1024;   // dynamic LDS of one gradient workgroup

int grad_pass(qf_engine* e, const double* dP, const double* dcm, double* dgrad, hipStream_t s)
{
    // every class needs a gradient kernel: check before launching anything
    size_t gdoubles = 0;
    for (const auto& L : e->launches) {
        if (!qf::grad_launcher(L.cb, L.ck))
            return fail(QF_ERR_ARG, "ERI gradients cover shells up to d; this basis has f shells");
        gdoubles = std::max<size_t>(gdoubles, (size_t)qf::grad_global_doubles(L.cb, L.ck));
    }
    if (gdoubles && e->d_gscratch.n < gdoubles * qf::COOP_GLOBAL_SLICES)
        QF_HIP(e->d_gscratch.resize(gdoubles * qf::COOP_GLOBAL_SLICES));
    qf::MolDesc md{};
    md.P = dP;
    md.PJ = dP;
    md.G = nullptr;
    md.cutmat = dcm;
    md.nbasis = e->nbasis;
    md.nshell = e->nshell;
    md.cutoff = e->grad_cutoff > 0.0 ? e->grad_cutoff : e->cutoff;
    md.primlim = e->primlim;
    md.hyb = e->hyb;
    md.split = (e->prec == QF_PREC_MIXED) ? e->split_thr : 0.0;   // only separates the launches' halves
    md.replicas = 1;
    md.natom = e->natom;
    md.shell_atom = e->d_shell_atom.p;
    md.grad = dgrad;
    e->mol[3] = md;
    qf::MolDesc* dmol = e->d_mol.p + 3;
    QF_HIP(hipMemcpyAsync(dmol, &e->mol[3], sizeof(md), hipMemcpyHostToDevice, s));
    const size_t tile_bytes = (size_t)3 * e->natom * sizeof(double);
    for (size_t li = 0; li < e->launches.size(); ++li) {
        const auto& L = e->launches[li];
        const int gm = e->ctune[L.cb][L.ck].grad;
        const bool want_lane = gm >= 0 ? gm == 1 : e->grad_lane;
        const qf::GradLaunch lane = want_lane ? qf::grad_lane_launcher(L.cb, L.ck) : nullptr;
        const size_t slots = lane ? 0 : (size_t)qf::grad_lds_doubles(L.cb, L.ck) * sizeof(double);
        const size_t tile = slots + tile_bytes <= kGradLdsMax ? tile_bytes : 0;
        // one workgroup per task unless the engine caps the grid (qf_set_persistent):
        // the hardware then hands out tasks as workgroups finish
        const int grid = e->persist > 0 ? std::min(L.count, e->persist) : L.count;
        hipEvent_t eb = nullptr;
        if (e->timing) {
            while (e->ev_pool.size() < e->ev_used + 2) {
                hipEvent_t ev;
                QF_HIP(hipEventCreate(&ev));
                e->ev_pool.push_back(ev);
            }
            eb = e->ev_pool[e->ev_used + 1];
            QF_HIP(hipEventRecord(e->ev_pool[e->ev_used], s));
            e->ev_used += 2;
            e->ev_launch.push_back((int)li);
        }
        (lane ? lane : qf::grad_launcher(L.cb, L.ck))(e->d_tasks.p + L.offset, L.count, e->d_blocks.p, dmol, grid,
                                                      tile, e->d_gscratch.p, L.fp32, s);
        if (eb) QF_HIP(hipEventRecord(eb, s));
    }
    QF_HIP(hipGetLastError());
    return QF_OK;
}
}  // namespace

int qf_set_grad_cutoff(qf_engine* e, double grad_cutoff)
{
    if (!e || !(grad_cutoff > 0.0)) return fail(QF_ERR_ARG, "gradient cutoff must be positive");
    const double before = e->list_cutoff();
    e->grad_cutoff = grad_cutoff;
    if (e->list_cutoff() != before) e->prepared = false;   // the lists must reach further down
    return QF_OK;
}

int qf_set_grad_kernel(qf_engine* e, int lane)
{
    if (!e) return fail(QF_ERR_ARG, "engine is NULL");
    e->grad_lane = lane != 0;
    return QF_OK;
}

int qf_set_grad_class(qf_engine* e, int cb, int ck, int mode)
{
    if (!e) return fail(QF_ERR_ARG, "engine is NULL");
    if (cb < 0 || cb >= kNPC || ck < 0 || ck > cb) return fail(QF_ERR_ARG, "class index out of range");
    if (mode < -1 || mode > 1) return fail(QF_ERR_ARG, "mode must be -1, 0 (wave-per-quartet) or 1 (per-lane)");
    e->ctune[cb][ck].grad = mode;
    return QF_OK;
}

int qf_eri_grad_device(qf_engine* e, const double* d_P, double* d_grad, const double* d_cutmat, void* stream)
{
    if (!e || !d_P || !d_grad) return fail(QF_ERR_ARG, "NULL");
    int rc = e->set_device();
    if (rc) return rc;
    if (!e->prepared) return fail(QF_ERR_STATE, "qf_prepare has not been called");
    const hipStream_t s = (hipStream_t)stream;
    const double* dcm = d_cutmat;
    if (!dcm) {
        const int nthr = 256, tot = e->nshell * e->nshell;
        hipLaunchKernelGGL(k_cutmat, dim3((tot + nthr - 1) / nthr), dim3(nthr), 0, s, d_P, e->nbasis, e->nshell,
                           e->d_f0.p, e->d_f1.p, e->d_cm.p);
        dcm = e->d_cm.p;
    }
    rc = grad_pass(e, d_P, dcm, d_grad, s);
    if (rc == QF_OK) {
        e->grad_built = true;
        e->grad_cm_own = (dcm == e->d_cm.p);
    }
    return rc;
}

int qf_eri_grad_host(qf_engine* e, const double* P, double* grad, const double* cutmat)
{
    if (!e || !P || !grad) return fail(QF_ERR_ARG, "NULL");
    int rc = e->set_device();
    if (rc) return rc;
    if (!e->prepared) return fail(QF_ERR_STATE, "qf_prepare has not been called");
    const size_t n2 = (size_t)e->nbasis * e->nbasis, ng = (size_t)3 * e->natom;
    QF_HIP(hipMemcpy(e->d_P.p, P, n2 * sizeof(double), hipMemcpyHostToDevice));
    const double* dcm = nullptr;
    if (cutmat) {
        QF_HIP(hipMemcpy(e->d_cm.p, cutmat, (size_t)e->nshell * e->nshell * sizeof(double), hipMemcpyHostToDevice));
        dcm = e->d_cm.p;
    }
    if (e->d_grad.n < ng) QF_HIP(e->d_grad.resize(ng));
    QF_HIP(hipMemset(e->d_grad.p, 0, ng * sizeof(double)));
    rc = qf_eri_grad_device(e, e->d_P.p, e->d_grad.p, dcm, nullptr);
    if (rc) return rc;
    std::vector<double> g(ng);
    QF_HIP(hipMemcpy(g.data(), e->d_grad.p, ng * sizeof(double), hipMemcpyDeviceToHost));
    for (size_t k = 0; k < ng; ++k) grad[k] += g[k];
    return QF_OK;
}

int qf_grad_launch_flops(qf_engine* e, double* flops, long long* quartets, long long* prims)
{
    // counts, with the density screen of the last gradient, the quartets and
    // primitive quartets the gradient launches evaluate, priced with the
    // generated per-class gradient FLOP table
    if (!e || !e->prepared) return fail(QF_ERR_STATE, "not prepared");
    if (!e->grad_built) return fail(QF_ERR_STATE, "no gradient since the last qf_prepare");
    if (!e->grad_cm_own)
        return fail(QF_ERR_STATE, "the last gradient screened with a caller-owned Cutmatrix; counts need one that "
                                  "computes its own (d_cutmat = NULL)");
    int rc = e->set_device();
    if (rc) return rc;
    const size_t nl = e->launches.size();
    QF_HIP(hipMemset(e->d_counts.p, 0, 2 * std::max<size_t>(nl, 1) * sizeof(unsigned long long)));
    for (size_t k = 0; k < nl; ++k) {
        const auto& L = e->launches[k];
        hipLaunchKernelGGL(k_count, dim3(L.count), dim3(qf::WAVE), 0, 0, e->d_tasks.p + L.offset, L.count,
                           e->d_blocks.p, e->d_mol.p + 3, e->d_counts.p, (int)k, L.fp32, 1);
    }
    QF_HIP(hipGetLastError());
    std::vector<unsigned long long> c(2 * std::max<size_t>(nl, 1));
    QF_HIP(hipMemcpy(c.data(), e->d_counts.p, c.size() * sizeof(unsigned long long), hipMemcpyDeviceToHost));
    for (size_t k = 0; k < nl; ++k) {
        if (flops) flops[k] = qf::grad_class_flops(e->launches[k].cb, e->launches[k].ck, c[2 * k], c[2 * k + 1]);
        if (quartets) quartets[k] = (long long)c[2 * k];
        if (prims) prims[k] = (long long)c[2 * k + 1];
    }
    return QF_OK;
}

int qf_fock_build_host(qf_engine* e, const double* P, double* F, const double* cutmat)
{
    if (!e || !P || !F) return fail(QF_ERR_ARG, "NULL");
    int rc = e->set_device();
    if (rc) return rc;
    if (!e->prepared) return fail(QF_ERR_STATE, "qf_prepare has not been called");
    const size_t n2 = (size_t)e->nbasis * e->nbasis;
    QF_HIP(hipMemcpy(e->d_P.p, P, n2 * sizeof(double), hipMemcpyHostToDevice));
    const double* dcm = nullptr;
    if (cutmat) {
        QF_HIP(hipMemcpy(e->d_cm.p, cutmat, (size_t)e->nshell * e->nshell * sizeof(double), hipMemcpyHostToDevice));
        dcm = e->d_cm.p;
    }
    rc = build(e, e->d_P.p, nullptr, dcm, 0);
    if (rc) return rc;
    std::vector<double> G(n2);
    QF_HIP(hipMemcpy(G.data(), e->d_Gs.p, n2 * sizeof(double), hipMemcpyDeviceToHost));
    for (size_t k = 0; k < n2; ++k) F[k] += G[k];
    return QF_OK;
}

int qf_fock_build_os_device(qf_engine* e, const double* d_Pa, const double* d_Pb, double* d_Fa, double* d_Fb,
                            const double* d_cutmat, void* stream)
{
    if (!e || !d_Pa || !d_Pb || !d_Fa || !d_Fb) return fail(QF_ERR_ARG, "NULL");
    int rc = e->set_device();
    if (rc) return rc;
    return build_os(e, d_Pa, d_Pb, d_Fa, d_Fb, d_cutmat, (hipStream_t)stream);
}

int qf_fock_build_os_host(qf_engine* e, const double* Pa, const double* Pb, double* Fa, double* Fb,
                          const double* cutmat)
{
    if (!e || !Pa || !Pb || !Fa || !Fb) return fail(QF_ERR_ARG, "NULL");
    int rc = e->set_device();
    if (rc) return rc;
    if (!e->prepared) return fail(QF_ERR_STATE, "qf_prepare has not been called");
    const size_t n2 = (size_t)e->nbasis * e->nbasis;
    QF_HIP(e->d_Pb.resize(n2));
    QF_HIP(e->d_Fa.resize(n2));
    QF_HIP(e->d_Fb.resize(n2));
    QF_HIP(hipMemcpy(e->d_P.p, Pa, n2 * sizeof(double), hipMemcpyHostToDevice));
    QF_HIP(hipMemcpy(e->d_Pb.p, Pb, n2 * sizeof(double), hipMemcpyHostToDevice));
    QF_HIP(hipMemcpy(e->d_Fa.p, Fa, n2 * sizeof(double), hipMemcpyHostToDevice));
    QF_HIP(hipMemcpy(e->d_Fb.p, Fb, n2 * sizeof(double), hipMemcpyHostToDevice));
    const double* dcm = nullptr;
    if (cutmat) {
        QF_HIP(hipMemcpy(e->d_cm.p, cutmat, (size_t)e->nshell * e->nshell * sizeof(double), hipMemcpyHostToDevice));
        dcm = e->d_cm.p;
    }
    rc = build_os(e, e->d_P.p, e->d_Pb.p, e->d_Fa.p, e->d_Fb.p, dcm, 0);
    if (rc) return rc;
    QF_HIP(hipMemcpy(Fa, e->d_Fa.p, n2 * sizeof(double), hipMemcpyDeviceToHost));
    QF_HIP(hipMemcpy(Fb, e->d_Fb.p, n2 * sizeof(double), hipMemcpyDeviceToHost));
    return QF_OK;
}

const double* qf_device_g(qf_engine* e) { return e ? e->d_Gs.p : nullptr; }

int qf_internal_basis(qf_engine* e, int* natom, const double** xyz, int* nshell, const int** s_atom,
                      const int** s_lmin, const int** s_lmax, const int** s_nprim, const int** s_pstart,
                      const int** s_fstart, int* nprim, const double** exps, const double** coef, int* nbasis)
{
    if (!e || e->nshell == 0 || e->natom == 0) return fail(QF_ERR_STATE, "molecule and basis must be set");
    int rc = e->set_device();
    if (rc) return rc;
    *natom = e->natom;
    *xyz = e->xyz.data();
    *nshell = e->nshell;
    *s_atom = e->s_atom.data();
    *s_lmin = e->s_lmin.data();
    *s_lmax = e->s_lmax.data();
    *s_nprim = e->s_nprim.data();
    *s_pstart = e->s_pstart.data();
    *s_fstart = e->s_fstart.data();
    *nprim = e->nprim;
    *exps = e->exps.data();
    *coef = e->coef.data();
    *nbasis = e->nbasis;
    return QF_OK;
}

double qf_flops_estimate(qf_engine* e)
{
    const int n = qf_launch_count(e);
    if (n < 0) return -1.0;
    std::vector<double> f(std::max(n, 1));
    if (qf_launch_flops(e, f.data()) != QF_OK) return -1.0;
    double tot = 0.0;
    for (int k = 0; k < n; ++k) tot += f[k];
    e->flops_last = tot;
    return tot;
}

int qf_launch_count(qf_engine* e)
{
    if (!e || !e->prepared) return -1;
    return (int)e->launches.size();
}

int qf_launch_info(qf_engine* e, int idx, int* cb, int* ck, int* fp32, int* ntasks)
{
    if (!e || !e->prepared || idx < 0 || idx >= (int)e->launches.size()) return fail(QF_ERR_ARG, "launch index");
    const auto& L = e->launches[idx];
    *cb = L.cb;
    *ck = L.ck;
    *fp32 = L.fp32;
    *ntasks = L.count;
    return QF_OK;
}

int qf_launch_flops(qf_engine* e, double* flops)
{
    // Counts, with the density of the last build, the quartets and primitive
    // quartets that pass the screens in every launch, and prices them with
    // the generated per-class FLOP table.
    // The kernel reads the screen settings and the Cutmatrix of the last
    // build (device MolDesc slot 0), so it needs a build since the last
    // prepare, screened with the engine's own Cutmatrix (a caller-owned one
    // may be gone by now).
    if (!e || !e->prepared) return fail(QF_ERR_STATE, "not prepared");
    if (!e->built) return fail(QF_ERR_STATE, "no Fock build since the last qf_prepare: counts use its density screen");
    if (!e->count_cm_own)
        return fail(QF_ERR_STATE, "the last build screened with a caller-owned Cutmatrix; counts need a build "
                                  "that computes its own (d_cutmat = NULL)");
    int rc = e->set_device();
    if (rc) return rc;
    const size_t nl = e->launches.size();
    QF_HIP(hipMemset(e->d_counts.p, 0, 2 * std::max<size_t>(nl, 1) * sizeof(unsigned long long)));
    for (size_t k = 0; k < nl; ++k) {
        const auto& L = e->launches[k];
        hipLaunchKernelGGL(k_count, dim3(L.count), dim3(qf::WAVE), 0, 0, e->d_tasks.p + L.offset, L.count,
                           e->d_blocks.p, e->d_mol.p, e->d_counts.p, (int)k, L.fp32, 0);
    }
    QF_HIP(hipGetLastError());
    std::vector<unsigned long long> c(2 * std::max<size_t>(nl, 1));
    QF_HIP(hipMemcpy(c.data(), e->d_counts.p, c.size() * sizeof(unsigned long long), hipMemcpyDeviceToHost));
    for (size_t k = 0; k < nl; ++k) flops[k] = qf::class_flops(e->launches[k].cb, e->launches[k].ck, c[2 * k], c[2 * k + 1]);
    return QF_OK;
}

int qf_launch_counts(qf_engine* e, long long* quartets, long long* prims)
{
    if (!e || !e->prepared) return fail(QF_ERR_STATE, "not prepared");
    const size_t nl = e->launches.size();
    std::vector<double> tmp(std::max<size_t>(nl, 1));
    int rc = qf_launch_flops(e, tmp.data());
    if (rc) return rc;
    std::vector<unsigned long long> c(2 * std::max<size_t>(nl, 1));
    QF_HIP(hipMemcpy(c.data(), e->d_counts.p, c.size() * sizeof(unsigned long long), hipMemcpyDeviceToHost));
    for (size_t k = 0; k < nl; ++k) {
        quartets[k] = (long long)c[2 * k];
        prims[k] = (long long)c[2 * k + 1];
    }
    return QF_OK;
}

int qf_set_tuning(qf_engine* e, int replicas, int shuffle)
{
    if (!e || replicas < 1 || replicas > 256) return fail(QF_ERR_ARG, "replicas must be in 1..256");
    e->replicas = replicas;
    e->shuffle = shuffle != 0;
    e->prepared = false;
    return QF_OK;
}

int qf_set_tile(qf_engine* e, int enable, int waves, int tasks_per_group, int lds_bytes)
{
    if (!e) return fail(QF_ERR_ARG, "NULL");
    if (waves < 1 || waves > 4) return fail(QF_ERR_ARG, "waves must be in 1..4");
    if (tasks_per_group < 0) return fail(QF_ERR_ARG, "tasks_per_group must be >= 0 (0: automatic)");
    if (lds_bytes < 1024 || lds_bytes > 64 * 1024) return fail(QF_ERR_ARG, "lds_bytes must be in 1 KiB..64 KiB");
    e->tile = enable != 0;
    e->tile_waves = waves;
    e->tile_tasks = tasks_per_group;
    e->lds_budget = lds_bytes;
    e->prepared = false;
    return QF_OK;
}

int qf_set_class_tuning(qf_engine* e, int cb, int ck, int tile, int waves, int tasks_per_group, int bras_per_task)
{
    if (!e) return fail(QF_ERR_ARG, "NULL");
    if (cb < 0 || cb >= kNPC || ck < 0 || ck > cb) return fail(QF_ERR_ARG, "class index out of range");
    if (waves > 4 || bras_per_task > 64) return fail(QF_ERR_ARG, "waves <= 4 and bras_per_task <= 64");
    qf_engine::ClassTune& t = e->ctune[cb][ck];
    if (tile > 2) return fail(QF_ERR_ARG, "tile must be -1, 0 (direct), 1 (LDS tile) or 2 (wave per quartet)");
    t.tile = tile < 0 ? -1 : tile;
    t.waves = std::max(0, waves);
    t.tpg = std::max(0, tasks_per_group);
    t.mmax = std::max(0, bras_per_task);
    e->prepared = false;
    return QF_OK;
}

int qf_set_class_os(qf_engine* e, int cb, int ck, int mode)
{
    if (!e) return fail(QF_ERR_ARG, "NULL");
    if (cb < 0 || cb >= kNPC || ck < 0 || ck > cb) return fail(QF_ERR_ARG, "class index out of range");
    if (mode < -1 || mode > 2) return fail(QF_ERR_ARG, "mode must be -1, 0 (fused), 1 (two passes) or 2 (fused lean)");
    e->ctune[cb][ck].os = mode;
    e->prepared = false;
    return QF_OK;
}

int qf_set_class_piece(qf_engine* e, int cb, int ck, int piece)
{
    if (!e) return fail(QF_ERR_ARG, "NULL");
    if (cb < 0 || cb >= kNPC || ck < 0 || ck > cb) return fail(QF_ERR_ARG, "class index out of range");
    e->ctune[cb][ck].piece = piece;
    e->prepared = false;
    return QF_OK;
}

int qf_set_prim_split(qf_engine* e, int piece)
{
    if (!e) return fail(QF_ERR_ARG, "NULL");
    e->prim_piece = piece;
    e->prepared = false;
    return QF_OK;
}

int qf_set_persistent(qf_engine* e, int max_workgroups)
{
    if (!e || max_workgroups < 0) return fail(QF_ERR_ARG, "max_workgroups must be >= 0");
    e->persist = max_workgroups;
    return QF_OK;
}

int qf_set_timing(qf_engine* e, int enable)
{
    if (!e) return fail(QF_ERR_ARG, "NULL");
    e->timing = enable != 0;
    e->ev_used = 0;
    e->ev_launch.clear();
    return QF_OK;
}

int qf_read_timing(qf_engine* e, double* ms_per_launch, int* calls_per_launch)
{
    // Sums the event-measured duration of every recorded class launch since
    // the last read, per launch index; waits for the last recorded event.
    if (!e || !e->prepared) return fail(QF_ERR_STATE, "not prepared");
    const size_t nl = e->launches.size();
    for (size_t k = 0; k < nl; ++k) {
        ms_per_launch[k] = 0.0;
        calls_per_launch[k] = 0;
    }
    if (e->ev_used) QF_HIP(hipEventSynchronize(e->ev_pool[e->ev_used - 1]));
    for (size_t r = 0; r < e->ev_launch.size(); ++r) {
        if (e->ev_launch[r] < 0 || (size_t)e->ev_launch[r] >= nl) continue;
        float ms = 0.0f;
        QF_HIP(hipEventElapsedTime(&ms, e->ev_pool[2 * r], e->ev_pool[2 * r + 1]));
        ms_per_launch[e->ev_launch[r]] += ms;
        calls_per_launch[e->ev_launch[r]] += 1;
    }
    e->ev_used = 0;
    e->ev_launch.clear();
    return QF_OK;
}

}  // extern "C"

// ==========================================================================
// Batches: many prepared engines (molecules) built together, one launch per
// quartet class for all of them, optionally replayed from a HIP graph.
// Each molecule keeps its own pair data (owned by its engine); the batch
// owns the combined block/task/group tables and pooled P, F, G and Cutmatrix
// buffers, so every pointer a launch uses is fixed and the whole build can be
// captured once and replayed.
// ==========================================================================
namespace {

struct BatchMol {
    const double* P;
    double* F;
    double* Gt;      // replicas * n * n
    double* Gs;      // n * n
    double* cm;      // ns * ns
    const int* f0;
    const int* f1;
    int n, ns;
};

__global__ void k_cutmat_batch(const BatchMol* __restrict__ bm)
{
    const BatchMol m = bm[blockIdx.y];
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= m.ns * m.ns) return;
    const int I = idx / m.ns, J = idx % m.ns;
    double v = 0.0;
    for (int i = m.f0[I]; i <= m.f1[I]; ++i)
        for (int j = m.f0[J]; j <= m.f1[J]; ++j) v = fmax(v, fabs(m.P[i * m.n + j]));
    m.cm[idx] = v;
}

__global__ void k_add(double* __restrict__ y, const double* __restrict__ x, int count)
{
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx < count) y[idx] += x[idx];
}

__global__ void k_finalize_batch(const BatchMol* __restrict__ bm, int replicas)
{
    const BatchMol m = bm[blockIdx.y];
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= m.n * m.n) return;
    const int i = idx / m.n, j = idx % m.n;
    const size_t n2 = (size_t)m.n * m.n;
    double g = 0.0;
    for (int r = 0; r < replicas; ++r) g += m.Gt[r * n2 + i * m.n + j] + m.Gt[r * n2 + j * m.n + i];
    m.Gs[idx] = g;
    m.F[idx] += g;
}

}  // namespace

struct qf_batch {
    int device = 0;
    std::vector<qf_engine*> engs;
    std::vector<unsigned long long> gens;   // each engine's generation when the batch copied its tables
    int replicas = 2;
    int persist = 0;
    std::vector<qf_engine::Launch> launches;
    std::vector<BatchMol> mols;
    DevBuf<BatchMol> d_bm;
    DevBuf<qf::PairBlock> d_blocks;
    DevBuf<qf::Task> d_tasks;
    DevBuf<qf::TileGroup> d_groups;
    DevBuf<qf::MolDesc> d_moldesc;
    DevBuf<double> d_P, d_F, d_Gt, d_Gs, d_cm;
    DevBuf<double> d_cscratch;   // global slots of the wave-per-quartet classes that need them
    size_t gt_total = 0;
    int max_n2 = 0, max_ns2 = 0;
    // graph replay
    bool use_graph = false;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    hipStream_t cap_stream = nullptr;
    // per-launch timing (eager builds only)
    bool timing = false;
    std::vector<hipEvent_t> ev_pool;
    size_t ev_used = 0;
    std::vector<int> ev_launch;
    ~qf_batch()
    {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        if (cap_stream) (void)hipStreamDestroy(cap_stream);
        for (hipEvent_t ev : ev_pool) (void)hipEventDestroy(ev);
    }
};

namespace {

int batch_issue(qf_batch* b, hipStream_t s)
{
    const int nm = (int)b->mols.size();
    hipLaunchKernelGGL(k_cutmat_batch, dim3((b->max_ns2 + 255) / 256, nm), dim3(256), 0, s, b->d_bm.p);
    hipLaunchKernelGGL(k_zero, dim3(1024), dim3(256), 0, s, b->d_Gt.p, b->gt_total);
    for (size_t li = 0; li < b->launches.size(); ++li) {
        const auto& L = b->launches[li];
        hipEvent_t ea = nullptr, eb = nullptr;
        if (b->timing) {
            while (b->ev_pool.size() < b->ev_used + 2) {
                hipEvent_t ev;
                QF_HIP(hipEventCreate(&ev));
                b->ev_pool.push_back(ev);
            }
            ea = b->ev_pool[b->ev_used];
            eb = b->ev_pool[b->ev_used + 1];
            b->ev_used += 2;
            b->ev_launch.push_back((int)li);
            QF_HIP(hipEventRecord(ea, s));
        }
        if (L.tile == 2) {
            qf::FockCoopLaunch fc = qf::fock_coop_launcher(L.cb, L.ck);
            if (!fc) return fail(QF_ERR_ARG, "no wave-per-quartet kernel for class");
            fc(b->d_tasks.p + L.offset, L.count, b->d_blocks.p, b->d_moldesc.p, 0, b->persist, b->d_cscratch.p, s);
        } else if (L.tile) {
            qf::FockTileLaunch ft = qf::fock_tile_launcher(L.cb, L.ck);
            if (!ft) return fail(QF_ERR_ARG, "no tile kernel for class");
            ft(b->d_groups.p + L.goffset, L.gcount, b->d_tasks.p, b->d_blocks.p, b->d_moldesc.p, L.fp32, 0, L.waves,
               L.lds, b->persist, s);
        } else {
            qf::FockLaunch f = qf::fock_launcher(L.cb, L.ck);
            if (!f) return fail(QF_ERR_ARG, "no per-lane kernel for class");
            if (f(b->d_tasks.p + L.offset, L.count, b->d_blocks.p, b->d_moldesc.p, L.fp32, 0, b->persist, s))
                return fail(QF_ERR_ARG, kDirectFp32Refused);
        }
        if (b->timing) QF_HIP(hipEventRecord(eb, s));
    }
    hipLaunchKernelGGL(k_finalize_batch, dim3((b->max_n2 + 255) / 256, nm), dim3(256), 0, s, b->d_bm.p,
                       b->replicas);
    QF_HIP(hipGetLastError());
    return QF_OK;
}

}  // namespace

extern "C" {

int qf_batch_create_ext(qf_engine* const* engines, int count, int replicas, const double* const* d_P,
                        double* const* d_F, qf_batch** out)
{
    if (!engines || count <= 0 || !out) return fail(QF_ERR_ARG, "engines/count/out");
    if ((d_P == nullptr) != (d_F == nullptr)) return fail(QF_ERR_ARG, "give both d_P and d_F or neither");
    if (replicas < 1 || replicas > 64) return fail(QF_ERR_ARG, "replicas must be in 1..64");
    *out = nullptr;
    auto* b = new qf_batch();
    b->device = engines[0]->device;
    b->replicas = replicas;
    b->persist = engines[0]->persist;
    int rc = QF_OK;
    auto bail = [&](int code) {
        delete b;
        return code;
    };
    if (hipSetDevice(b->device) != hipSuccess) return bail(fail(QF_ERR_BADDEVICE, "hipSetDevice"));
    for (int i = 0; i < count; ++i) {
        qf_engine* e = engines[i];
        if (!e) return bail(fail(QF_ERR_ARG, "NULL engine"));
        if (e->device != b->device) return bail(fail(QF_ERR_ARG, "engines of a batch must share one device"));
        rc = prepare(e);
        if (rc) return bail(rc);
        b->engs.push_back(e);
        b->gens.push_back(e->generation);
    }
    // pooled buffers
    size_t np = 0, ncm = 0;
    std::vector<size_t> poff, cmoff, gtoff;
    for (qf_engine* e : b->engs) {
        const size_t n2 = (size_t)e->nbasis * e->nbasis;
        poff.push_back(np);
        cmoff.push_back(ncm);
        gtoff.push_back(b->gt_total);
        np += n2;
        ncm += (size_t)e->nshell * e->nshell;
        b->gt_total += n2 * replicas;
        b->max_n2 = std::max<int>(b->max_n2, (int)n2);
        b->max_ns2 = std::max<int>(b->max_ns2, e->nshell * e->nshell);
    }
    if ((!d_P && (b->d_P.resize(np) != hipSuccess || b->d_F.resize(np) != hipSuccess)) || b->d_Gs.resize(np) != hipSuccess || b->d_cm.resize(ncm) != hipSuccess ||
        b->d_Gt.resize(b->gt_total) != hipSuccess)
        return bail(fail(QF_ERR_MEMORY, "batch buffers"));
    if (!d_P && (hipMemset(b->d_P.p, 0, np * sizeof(double)) != hipSuccess ||
                 hipMemset(b->d_F.p, 0, np * sizeof(double)) != hipSuccess))
        return bail(fail(QF_ERR_HIP, "hipMemset"));
    // combined tables
    std::vector<qf::PairBlock> blocks;
    std::vector<qf::MolDesc> md;
    std::vector<int> boff;
    for (int i = 0; i < count; ++i) {
        qf_engine* e = b->engs[i];
        boff.push_back((int)blocks.size());
        for (qf::PairBlock pb : e->hblocks) {
            pb.mol = i;
            blocks.push_back(pb);
        }
        if (d_P && (!d_P[i] || !d_F[i])) return bail(fail(QF_ERR_ARG, "NULL density/Fock pointer"));
        qf::MolDesc m{};
        m.P = d_P ? d_P[i] : b->d_P.p + poff[i];
        m.PJ = m.P;
        m.G = b->d_Gt.p + gtoff[i];
        m.cutmat = b->d_cm.p + cmoff[i];
        m.nbasis = e->nbasis;
        m.nshell = e->nshell;
        m.cutoff = e->cutoff;
        m.primlim = e->primlim;
        m.hyb = e->hyb;
        m.replicas = replicas;
        m.split = (e->prec == QF_PREC_MIXED) ? e->split_thr : 0.0;
        md.push_back(m);
        BatchMol bm{};
        bm.P = m.P;
        bm.F = d_F ? d_F[i] : b->d_F.p + poff[i];
        bm.Gt = m.G;
        bm.Gs = b->d_Gs.p + poff[i];
        bm.cm = b->d_cm.p + cmoff[i];
        bm.f0 = e->d_f0.p;
        bm.f1 = e->d_f1.p;
        bm.n = e->nbasis;
        bm.ns = e->nshell;
        b->mols.push_back(bm);
    }
    std::vector<qf::Task> tasks;
    std::vector<qf::TileGroup> groups;
    // work of a task: bras x kets x primitive pairs (bra and ket)
    auto cost = [](const qf::Task& x) {
        return (long long)(x.ke - x.kb) * x.ket_count * (x.pe - x.pb) * (x.qe - x.qb);
    };
    for (int cb = 0; cb < kNPC; ++cb)
        for (int ck = 0; ck <= cb; ++ck)
            for (int fp32 = 0; fp32 < 2; ++fp32)
                for (int tile = 0; tile < 3; ++tile) {
                    qf_engine::Launch C{};
                    C.cb = cb;
                    C.ck = ck;
                    C.fp32 = fp32;
                    C.tile = tile;
                    C.offset = (int)tasks.size();
                    C.goffset = (int)groups.size();
                    C.waves = 1;
                    for (int i = 0; i < count; ++i) {
                        const qf_engine* e = b->engs[i];
                        for (const auto& L : e->launches) {
                            if (L.cb != cb || L.ck != ck || L.fp32 != fp32 || L.tile != tile) continue;
                            const int tbase = (int)tasks.size();
                            for (int x = L.offset; x < L.offset + L.count; ++x) {
                                qf::Task t = e->tasks[x];
                                t.bra_block += boff[i];
                                t.ket_block += boff[i];
                                tasks.push_back(t);
                            }
                            for (int q = L.goffset; q < L.goffset + L.gcount; ++q) {
                                qf::TileGroup g = e->groups[q];
                                g.task_begin += tbase - L.offset;
                                g.task_end += tbase - L.offset;
                                groups.push_back(g);
                            }
                            C.lds = std::max(C.lds, L.lds);
                            C.waves = std::max(C.waves, L.waves);
                        }
                    }
                    C.count = (int)tasks.size() - C.offset;
                    C.gcount = (int)groups.size() - C.goffset;
                    if (C.count == 0) continue;
                    if (tile == 1) {
                        std::vector<long long> gc(C.gcount, 0);
                        for (int q = 0; q < C.gcount; ++q)
                            for (int x = groups[C.goffset + q].task_begin; x < groups[C.goffset + q].task_end; ++x)
                                gc[q] += cost(tasks[x]);
                        std::vector<int> ord(C.gcount);
                        std::iota(ord.begin(), ord.end(), 0);
                        std::stable_sort(ord.begin(), ord.end(), [&](int x, int y) { return gc[x] > gc[y]; });
                        std::vector<qf::TileGroup> sorted(C.gcount);
                        for (int q = 0; q < C.gcount; ++q) sorted[q] = groups[C.goffset + ord[q]];
                        std::copy(sorted.begin(), sorted.end(), groups.begin() + C.goffset);
                    } else {
                        std::stable_sort(tasks.begin() + C.offset, tasks.end(),
                                         [&](const qf::Task& x, const qf::Task& y) { return cost(x) > cost(y); });
                    }
                    b->launches.push_back(C);
                }
    size_t gdoubles = 0;
    for (const auto& L : b->launches)
        if (L.tile == 2) gdoubles = std::max<size_t>(gdoubles, (size_t)qf::coop_global_doubles(L.cb, L.ck));
    if (gdoubles && b->d_cscratch.resize(gdoubles * qf::COOP_GLOBAL_SLICES) != hipSuccess)
        return bail(fail(QF_ERR_MEMORY, "batch global slot buffer"));
    if (b->d_blocks.upload(blocks) != hipSuccess || b->d_tasks.upload(tasks) != hipSuccess ||
        b->d_groups.upload(groups) != hipSuccess || b->d_moldesc.upload(md) != hipSuccess ||
        b->d_bm.upload(b->mols) != hipSuccess)
        return bail(fail(QF_ERR_HIP, "batch table upload"));
    *out = b;
    return QF_OK;
}

int qf_batch_create(qf_engine* const* engines, int count, int replicas, qf_batch** out)
{
    return qf_batch_create_ext(engines, count, replicas, nullptr, nullptr, out);
}

void qf_batch_destroy(qf_batch* b)
{
    if (!b) return;
    (void)hipSetDevice(b->device);
    delete b;
}

int qf_batch_size(qf_batch* b) { return b ? (int)b->mols.size() : -1; }

double* qf_batch_density(qf_batch* b, int i)
{
    if (!b || i < 0 || i >= (int)b->mols.size()) return nullptr;
    return const_cast<double*>(b->mols[i].P);
}

double* qf_batch_fock(qf_batch* b, int i)
{
    if (!b || i < 0 || i >= (int)b->mols.size()) return nullptr;
    return b->mols[i].F;
}

const double* qf_batch_g(qf_batch* b, int i)
{
    if (!b || i < 0 || i >= (int)b->mols.size()) return nullptr;
    return b->mols[i].Gs;
}

int qf_batch_set_graph(qf_batch* b, int enable)
{
    if (!b) return fail(QF_ERR_ARG, "NULL");
    b->use_graph = enable != 0;
    return QF_OK;
}

int qf_batch_build(qf_batch* b, void* stream)
{
    if (!b) return fail(QF_ERR_ARG, "NULL");
    // the batch (and its graph) holds pointers into every engine's pair and
    // task tables: a re-prepared engine has replaced them
    for (size_t i = 0; i < b->engs.size(); ++i)
        if (b->engs[i]->generation != b->gens[i] || !b->engs[i]->prepared)
            return fail(QF_ERR_STATE, "engine " + std::to_string(i) +
                                          " was changed or re-prepared after the batch was created; create the batch again");
    QF_HIP(hipSetDevice(b->device));
    hipStream_t s = (hipStream_t)stream;
    if (!b->use_graph || b->timing) return batch_issue(b, s);
    if (!b->exec) {
        // capture on a private stream (capture needs a non-legacy stream)
        if (!b->cap_stream) QF_HIP(hipStreamCreateWithFlags(&b->cap_stream, hipStreamNonBlocking));
        QF_HIP(hipStreamBeginCapture(b->cap_stream, hipStreamCaptureModeThreadLocal));
        int rc = batch_issue(b, b->cap_stream);
        hipGraph_t g = nullptr;
        hipError_t ce = hipStreamEndCapture(b->cap_stream, &g);
        if (rc) {
            if (g) (void)hipGraphDestroy(g);
            return rc;
        }
        if (ce != hipSuccess) return fail(QF_ERR_HIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(ce));
        b->graph = g;
        QF_HIP(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
    }
    QF_HIP(hipGraphLaunch(b->exec, s));
    return QF_OK;
}

int qf_fock_build_batch(qf_engine* const* engines, int count, const double* const* d_P, double* const* d_F,
                        void* stream)
{
    // One-shot form of the batch API: a temporary batch, one launch per
    // class for all engines.  Repeated builds should keep a qf_batch.
    if (!engines || !d_P || count <= 0) return fail(QF_ERR_ARG, "NULL");
    qf_batch* b = nullptr;
    int rc = qf_batch_create(engines, count, 2, &b);
    if (rc) return rc;
    hipStream_t s = (hipStream_t)stream;
    for (int i = 0; i < count && !rc; ++i) {
        const size_t n2 = (size_t)b->mols[i].n * b->mols[i].n;
        if (hipMemcpyAsync(const_cast<double*>(b->mols[i].P), d_P[i], n2 * sizeof(double), hipMemcpyDeviceToDevice,
                           s) != hipSuccess)
            rc = fail(QF_ERR_HIP, "hipMemcpyAsync (batch density)");
    }
    if (!rc) rc = batch_issue(b, s);
    for (int i = 0; i < count && !rc; ++i) {
        const int n2 = b->mols[i].n * b->mols[i].n;
        if (d_F && d_F[i])
            hipLaunchKernelGGL(k_add, dim3((n2 + 255) / 256), dim3(256), 0, s, d_F[i], b->mols[i].Gs, n2);
        if (hipMemcpyAsync(engines[i]->d_Gs.p, b->mols[i].Gs, (size_t)n2 * sizeof(double), hipMemcpyDeviceToDevice,
                           s) != hipSuccess)
            rc = fail(QF_ERR_HIP, "hipMemcpyAsync (batch G)");
    }
    if (!rc && hipStreamSynchronize(s) != hipSuccess) rc = fail(QF_ERR_HIP, "hipStreamSynchronize");
    qf_batch_destroy(b);
    return rc;
}

int qf_batch_set_density_host(qf_batch* b, int i, const double* P)
{
    if (!b || !P || i < 0 || i >= (int)b->mols.size()) return fail(QF_ERR_ARG, "batch density");
    QF_HIP(hipSetDevice(b->device));
    const size_t n2 = (size_t)b->mols[i].n * b->mols[i].n;
    QF_HIP(hipMemcpy(const_cast<double*>(b->mols[i].P), P, n2 * sizeof(double), hipMemcpyHostToDevice));
    return QF_OK;
}

int qf_batch_get_host(qf_batch* b, int i, double* G, double* F)
{
    if (!b || i < 0 || i >= (int)b->mols.size()) return fail(QF_ERR_ARG, "batch index");
    QF_HIP(hipSetDevice(b->device));
    const size_t n2 = (size_t)b->mols[i].n * b->mols[i].n;
    if (G) QF_HIP(hipMemcpy(G, b->mols[i].Gs, n2 * sizeof(double), hipMemcpyDeviceToHost));
    if (F) QF_HIP(hipMemcpy(F, b->mols[i].F, n2 * sizeof(double), hipMemcpyDeviceToHost));
    return QF_OK;
}

int qf_batch_clear_fock(qf_batch* b, void* stream)
{
    if (!b) return fail(QF_ERR_ARG, "NULL");
    QF_HIP(hipSetDevice(b->device));
    for (const auto& m : b->mols)
        hipLaunchKernelGGL(k_zero, dim3(64), dim3(256), 0, (hipStream_t)stream, m.F, (size_t)m.n * m.n);
    QF_HIP(hipGetLastError());
    return QF_OK;
}

int qf_batch_launch_count(qf_batch* b) { return b ? (int)b->launches.size() : -1; }

int qf_batch_launch_info(qf_batch* b, int idx, int* cb, int* ck, int* fp32, int* ntasks)
{
    if (!b || idx < 0 || idx >= (int)b->launches.size()) return fail(QF_ERR_ARG, "launch index");
    const auto& L = b->launches[idx];
    *cb = L.cb;
    *ck = L.ck;
    *fp32 = L.fp32;
    *ntasks = L.count;
    return QF_OK;
}

int qf_batch_set_timing(qf_batch* b, int enable)
{
    if (!b) return fail(QF_ERR_ARG, "NULL");
    b->timing = enable != 0;
    b->ev_used = 0;
    b->ev_launch.clear();
    return QF_OK;
}

int qf_batch_read_timing(qf_batch* b, double* ms_per_launch, int* calls_per_launch)
{
    if (!b) return fail(QF_ERR_ARG, "NULL");
    const size_t nl = b->launches.size();
    for (size_t k = 0; k < nl; ++k) {
        ms_per_launch[k] = 0.0;
        calls_per_launch[k] = 0;
    }
    if (b->ev_used) QF_HIP(hipEventSynchronize(b->ev_pool[b->ev_used - 1]));
    for (size_t r = 0; r < b->ev_launch.size(); ++r) {
        if (b->ev_launch[r] < 0 || (size_t)b->ev_launch[r] >= nl) continue;
        float ms = 0.0f;
        QF_HIP(hipEventElapsedTime(&ms, b->ev_pool[2 * r], b->ev_pool[2 * r + 1]));
        ms_per_launch[b->ev_launch[r]] += ms;
        calls_per_launch[b->ev_launch[r]] += 1;
    }
    b->ev_used = 0;
    b->ev_launch.clear();
    return QF_OK;
}

}  // extern "C"
