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
#include <numeric>
#include <string>
#include <vector>

#include "../../include/quickfock.h"
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
const int kClassL[6][2] = {{0, 0}, {1, 0}, {1, 1}, {2, 0}, {2, 1}, {2, 2}};
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

// Counts, per quartet class, the quartets and primitive quartets that pass
// the screens (same tests as the Fock kernel); used for the FLOP model.
__global__ void k_count(const qf::Task* __restrict__ tasks, int ntasks, const qf::PairBlock* __restrict__ blocks,
                        const qf::MolDesc* __restrict__ mols, unsigned long long* __restrict__ counts, int slot,
                        int fp32)
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
        ++nq;
        for (int ip = 0; ip < bb.K; ++ip) {
            const double c1 = dn * bb.prim[(qf::PF_CUT * bb.K + ip) * bb.n + i];
            for (int jp = 0; jp < kb.K; ++jp)
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
    int prec = QF_PREC_FP64;
    double split = 1e-4;
    double split_thr = 0.0;  // Ycut_ij*Ycut_kl below which quartets run in FP32
    int replicas = 8;        // copies of G the atomics are spread over
    int shuffle = 1;         // permute tasks inside a launch (decorrelates atomic targets)
    // pair blocks
    std::vector<HostBlock> blocks;
    std::vector<qf::PairBlock> hblocks;
    std::vector<qf::Task> tasks;
    struct Launch {
        int cb, ck, fp32, offset, count;
    };
    std::vector<Launch> launches;
    bool prepared = false;
    qf_stats stats{};
    double flops_last = -1.0;
    // device
    DevBuf<qf::PairInfo> d_info;
    DevBuf<double> d_prim;
    DevBuf<int> d_perm;
    DevBuf<qf::PairBlock> d_blocks;
    DevBuf<qf::MolDesc> d_mol;
    DevBuf<qf::Task> d_tasks;
    DevBuf<double> d_Gt, d_Gs, d_cm, d_P;
    DevBuf<int> d_f0, d_f1;
    DevBuf<unsigned long long> d_counts;
    qf::MolDesc mol{};
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
    info.pad = 0.0;
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
                  DevBuf<double>& d_prim, std::vector<qf::PairBlock>& out, int mol, DevBuf<int>* d_perm = nullptr)
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
    out.clear();
    for (size_t k = 0; k < blocks.size(); ++k) {
        qf::PairBlock pb;
        pb.info = d_info.p + ioff[k];
        pb.prim = d_prim.p + poff[k];
        pb.bperm = (d_perm && !blocks[k].bperm.empty()) ? d_perm->p + permoff[k] : nullptr;
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
        size_t off = 0;
        for (size_t k = 0; k < blocks.size(); ++k) {
            qf::SchwarzLaunch L = qf::schwarz_launcher(blocks[k].cls);
            L(hb[k], dout.p + off, 0);
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
        if (!(y * ymax > e->cutoff)) return;
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
    rc = upload_blocks(e, e->blocks, e->d_info, e->d_prim, e->hblocks, 0, &e->d_perm);
    if (rc) return rc;
    for (auto& b : e->blocks) b.rows.clear();
    QF_HIP(e->d_blocks.upload(e->hblocks));

    // ---- wave tasks per quartet class and precision ------------------------
    const double thr_hi = e->split * ymax * ymax;
    e->split_thr = thr_hi;
    e->tasks.clear();
    e->launches.clear();
    e->stats = qf_stats{};
    for (auto& b : e->blocks) e->stats.n_pairs += (long long)b.info.size();
    for (int cb = 0; cb < 6; ++cb)
        for (int ck = 0; ck <= cb; ++ck)
            for (int fp32 = 0; fp32 < 2; ++fp32) {
                // (dd|dp) and (dd|dd) stay FP64 in mixed mode: their float
                // kernels spill several KB per lane (little to gain) and the
                // (dd|dp) one returned non-finite values on gfx950.
                const bool split_class = e->prec == QF_PREC_MIXED &&
                                         kClassL[cb][0] + kClassL[cb][1] + kClassL[ck][0] + kClassL[ck][1] < 7;
                if (fp32 && !split_class) continue;
                const int offset = (int)e->tasks.size();
                for (size_t bbk = 0; bbk < e->blocks.size(); ++bbk) {
                    const HostBlock& B = e->blocks[bbk];
                    if (B.cls != cb) continue;
                    for (size_t kbk = 0; kbk < e->blocks.size(); ++kbk) {
                        const HostBlock& Kb = e->blocks[kbk];
                        if (Kb.cls != ck) continue;
                        if (cb == ck && kbk > bbk) continue;
                        const bool same = (cb == ck && kbk == bbk);
                        // bras per task: long loops for cheap classes (more register
                        // accumulation, fewer atomics), short ones for expensive classes
                        // (enough waves to fill 256 CUs)
                        const int ni = ncart(kClassL[cb][0]) * ncart(kClassL[cb][1]) * ncart(kClassL[ck][0]) *
                                       ncart(kClassL[ck][1]);
                        const int mmax = ni <= 10 ? 16 : ni <= 36 ? 8 : ni <= 108 ? 4 : ni <= 324 ? 2 : 1;
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
                                int n = prefix_count(Kb.info, yi, e->cutoff, true);
                                if (same) n = std::min(n, i + 1);
                                if (!fp32) e->stats.n_quartets_bound += n;
                                int n64 = n;
                                if (split_class) n64 = std::min(n, prefix_count(Kb.info, yi, thr_hi, false));
                                lo[k] = fp32 ? n64 : 0;
                                hi[k] = fp32 ? n : n64;
                                nmax = std::max(nmax, hi[k]);
                            }
                            for (int s0 = 0; s0 < nmax; s0 += qf::WAVE) {
                                int kb = -1, ke = -1;
                                for (int k = 0; k < M; ++k)
                                    if (lo[k] < hi[k] && lo[k] < s0 + qf::WAVE && hi[k] > s0) {
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
                                t.ket_count = std::min(qf::WAVE, nmax - s0);
                                t.same_block = same ? 1 : 0;
                                t.all_fp64 = split_class ? 0 : 1;
                                e->tasks.push_back(t);
                            }
                        }
                    }
                }
                const int count = (int)e->tasks.size() - offset;
                if (e->shuffle && count > 1) {
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
                    std::stable_sort(e->tasks.begin() + offset, e->tasks.end(), [](const qf::Task& x, const qf::Task& y) {
                        return (x.ke - x.kb) * x.ket_count > (y.ke - y.kb) * y.ket_count;
                    });
                }
                if (count > 0) e->launches.push_back({cb, ck, fp32, offset, count});
                if (fp32) e->stats.n_tasks_fp32 += count; else e->stats.n_tasks_fp64 += count;
            }
    QF_HIP(e->d_tasks.upload(e->tasks));

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
    QF_HIP(e->d_mol.resize(1));
    QF_HIP(e->d_counts.resize(2 * std::max<size_t>(1, e->launches.size())));
    e->stats.device_bytes = (long long)(e->d_info.bytes() + e->d_prim.bytes() + e->d_tasks.bytes() +
                                        e->d_Gt.bytes() * 3 + e->d_cm.bytes() + e->d_blocks.bytes());
    e->prepared = true;
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
    qf::MolDesc md{};
    md.P = dP;
    md.G = e->d_Gt.p;
    md.cutmat = dcm;
    md.nbasis = n;
    md.nshell = ns;
    md.cutoff = e->cutoff;
    md.primlim = e->primlim;
    md.hyb = e->hyb;
    md.replicas = e->replicas;
    md.split = (e->prec == QF_PREC_MIXED) ? e->split_thr : 0.0;
    e->mol = md;
    QF_HIP(hipMemcpyAsync(e->d_mol.p, &e->mol, sizeof(md), hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(k_zero, dim3(1024), dim3(256), 0, s, e->d_Gt.p, (size_t)n * n * e->replicas);
    for (size_t li = 0; li < e->launches.size(); ++li) {
        const auto& L = e->launches[li];
        qf::FockLaunch f = qf::fock_launcher(L.cb, L.ck);
        if (!f) return fail(QF_ERR_ARG, "no kernel for class");
        hipEvent_t ea = nullptr, eb = nullptr;
        if (e->timing) {
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
        f(e->d_tasks.p + L.offset, L.count, e->d_blocks.p, e->d_mol.p, L.fp32, s);
        if (e->timing) QF_HIP(hipEventRecord(eb, s));
    }
    hipLaunchKernelGGL(k_finalize, dim3((n * n + 255) / 256), dim3(256), 0, s, e->d_Gt.p, e->replicas, e->d_Gs.p,
                       dF, n);
    QF_HIP(hipGetLastError());
    return QF_OK;
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
        if (shell_lmin[s] < 0 || shell_lmax[s] > 2 || shell_lmin[s] > shell_lmax[s])
            return fail(QF_ERR_ARG, "angular momentum above d is not supported");
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

int qf_fock_build_batch(qf_engine* const* engines, int count, const double* const* d_P, double* const* d_F,
                        void* stream)
{
    for (int k = 0; k < count; ++k) {
        int rc = qf_fock_build_device(engines[k], d_P[k], d_F ? d_F[k] : nullptr, nullptr, stream);
        if (rc) return rc;
    }
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
    if (!e || !e->prepared) return fail(QF_ERR_STATE, "not prepared");
    int rc = e->set_device();
    if (rc) return rc;
    const size_t nl = e->launches.size();
    QF_HIP(hipMemset(e->d_counts.p, 0, 2 * std::max<size_t>(nl, 1) * sizeof(unsigned long long)));
    for (size_t k = 0; k < nl; ++k) {
        const auto& L = e->launches[k];
        hipLaunchKernelGGL(k_count, dim3(L.count), dim3(qf::WAVE), 0, 0, e->d_tasks.p + L.offset, L.count,
                           e->d_blocks.p, e->d_mol.p, e->d_counts.p, (int)k, L.fp32);
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
