// Two-electron gradient on the wave-per-quartet kernel (QUICK's
// getGrad_kernel_sp / _spd, gpu_get2e_subs_grad.h:1211-1650 and 2484-2620;
// CPU get_cshell_eri_grad / cshell_eri_grad, gradient.fh:826-930 and
// eri_grad.fh:43-768).
//
// For every canonical quartet that passes QUICK's gradient screen
// (Ycut_ab Ycut_cd > gradCutoff and x DNmax > gradCutoff, quartets with all
// four shells on one atom skipped) the 64 lanes of one wave evaluate, per
// primitive quartet, the VRR of the class raised by one in bra and ket and
// add it into four contracted sets with weights 1, 2 alpha, 2 beta and
// 2 gamma (the generated node programs of gen_grad.py); then six HRR
// blocks give (a+-1 b|cd), (a b+-1|cd) and (ab|c+-1 d), and each lane
// contracts its integrals with the two-particle density
//   Gamma_abcd = s [2 P_ab P_cd - hyb/2 (P_ac P_bd + P_ad P_bc)],
// s the quartet's degeneracy weight f/4 (as in the Fock digestion), into
//   dE/dA_x = sum Gamma (2a (a+1_x b|cd) - n_x(a) (a-1_x b|cd)), likewise B
// and C, dE/dD = -(dE/dA + dE/dB + dE/dC).  E here is QUICK's two-electron
// energy 1/2 sum P G(P) with P held fixed.
//
// The gradient of a workgroup collects in LDS (3 natom doubles) when it
// fits and goes to md.grad with FP64 atomics at the end.
#pragma once

#include "gen/grad_maps.h"
#include "qf_coop.h"

namespace qf {

struct GradAcc {
    // s[dst] += w[set] * s[src] after the VRR of one primitive quartet;
    // w = 1, 2 alpha, 2 beta, 2 gamma
    unsigned short src, dst, set;
};

template <class D, typename SP>
__device__ __forceinline__ void grad_contract(int pb, int pe, int qb, int qe, const PairBlock& bb, int i,
                                              const PairInfo& bi, const PairBlock& kb, int j, const PairInfo& ki,
                                              double dn, double primlim, SP s, const QF_AS_LDS double* boys_tab)
{
    const int lane = threadIdx.x;
    SP cf = s + D::CF;
    for (int x = lane; x < D::NACC; x += D::NT) s[x] = 0.0;

    const auto bp = cptr(bb.prim);
    const auto kp = cptr(kb.prim);
    const int bn = bb.n, bK = bb.K, kn = kb.n, kK = kb.K;
    // QUICK's CPU gradient evaluates FmT with NABCD + 2 (eri_grad.fh:104-106)
    const int maxm = bi.qfs + ki.qfs + 2;
    const double c0 = dn * bp[(PF_CUT * bK) * bn + i];
    const auto vrr = gptr(D::vrr());
    const auto vlev = cptr(D::vlev());
    const auto acc = gptr(D::acc());
    for (int jp = qb; jp < qe; ++jp) {
        const double cutk = kp[(PF_CUT * kK + jp) * kn + j];
        if (!(c0 * cutk > primlim)) break;
        const double zk = kp[(PF_ZETA * kK + jp) * kn + j];
        const double ozk = kp[(PF_OOZ2 * kK + jp) * kn + j];
        const double gk = kp[(PF_ALPHA * kK + jp) * kn + j];
        const double Qk[3] = {kp[(PF_PX * kK + jp) * kn + j], kp[(PF_PY * kK + jp) * kn + j],
                              kp[(PF_PZ * kK + jp) * kn + j]};
        const double Kk = kp[(PF_KC * kK + jp) * kn + j];
        const double QC[3] = {Qk[0] - ki.A[0], Qk[1] - ki.A[1], Qk[2] - ki.A[2]};
        for (int ip = pb; ip < pe; ++ip) {
            if (!((dn * bp[(PF_CUT * bK + ip) * bn + i]) * cutk > primlim)) break;
            const double zb = bp[(PF_ZETA * bK + ip) * bn + i];
            const double ozb = bp[(PF_OOZ2 * bK + ip) * bn + i];
            const double ab = bp[(PF_ALPHA * bK + ip) * bn + i];
            const double Pb[3] = {bp[(PF_PX * bK + ip) * bn + i], bp[(PF_PY * bK + ip) * bn + i],
                                  bp[(PF_PZ * bK + ip) * bn + i]};
            const double Kb = bp[(PF_KC * bK + ip) * bn + i];
            const double rs = qf_rsqrt(zb + zk);
            const double inv = rs * rs;
            const double PQ[3] = {Pb[0] - Qk[0], Pb[1] - Qk[1], Pb[2] - Qk[2]};
            const double rpq = PQ[0] * PQ[0] + PQ[1] * PQ[1] + PQ[2] * PQ[2];
            const double T = rpq * (zb * zk * inv);
            double F[D::L + 1];
            boys_quick<D::L, const QF_AS_LDS double*, 4>(T, maxm, F, boys_tab);
            const double pref = Kb * Kk * rs;
            const double wp = zk * inv, wq = zb * inv;
            __syncthreads();   // the previous primitive's accumulation pass has read its sources
#pragma unroll
            for (int m = 0; m <= D::L; ++m)
                if (lane == m) s[D::FSLOT[m]] = F[m] * pref;
            if (lane == 0) {
#pragma unroll
                for (int d = 0; d < 3; ++d) {
                    cf[4 * d + 0] = Pb[d] - bi.A[d];
                    cf[4 * d + 1] = -wp * PQ[d];
                    cf[4 * d + 2] = ozb;
                    cf[4 * d + 3] = wp;
                    cf[12 + 4 * d + 0] = QC[d];
                    cf[12 + 4 * d + 1] = wq * PQ[d];
                    cf[12 + 4 * d + 2] = ozk;
                    cf[12 + 4 * d + 3] = wq;
                }
                cf[24] = 0.5 * inv;
                cf[32] = 1.0;
                cf[33] = 2.0 * ab;
                cf[34] = 2.0 * (zb - ab);
                cf[35] = 2.0 * gk;
                s[D::ZS] = 0.0;
            }
            __syncthreads();
            run_levels<kVU, D::NT>(vrr, vlev, D::NLV, [&](const CoopVNode (&q)[kVU], int x0, int e1) {
                double v[kVU];
#pragma unroll
                for (int u = 0; u < kVU; ++u) {
                    const SP c4 = cf + 4 * (q[u].meta & 15);
                    const double n1 = double((q[u].meta >> 4) & 15), n2 = double(q[u].meta >> 8);
                    v[u] = c4[0] * s[q[u].a] + c4[1] * s[q[u].b] + n1 * c4[2] * (s[q[u].c] - c4[3] * s[q[u].d]) +
                           n2 * cf[24] * s[q[u].e];
                }
#pragma unroll
                for (int u = 0; u < kVU; ++u)
                    if (x0 + u * D::NT < e1) s[q[u].dst] = v[u];
            });
            // the four weighted sets: every entry has its own accumulator
            for (int x = lane; x < D::NX; x += D::NT) {
                const GradAcc g = acc[x];
                s[g.dst] += cf[32 + g.set] * s[g.src];
            }
        }
    }
}

// One quartet: contract, HRR, then the lane-parallel contraction of the
// derivative integrals with Gamma; returns (in every lane) the wave's nine
// sums dE/dA, dE/dB, dE/dC.
template <class D, typename SP>
__device__ __forceinline__ void grad_quartet(const Task& t, const PairBlock& bb, int i, const PairInfo& bi,
                                             const PairBlock& kb, int j, const PairInfo& ki, double dn,
                                             const MolDesc& md, SP s, const QF_AS_LDS double* boys_tab,
                                             double (&g)[9])
{
    constexpr int na = ncart(D::La), nb = ncart(D::Lb), nc = ncart(D::Lc), nd = ncart(D::Ld);
    constexpr int nap = ncart(D::La + 1), nam = D::La > 0 ? ncart(D::La - 1) : 1;
    constexpr int nbp = ncart(D::Lb + 1), nbm = D::Lb > 0 ? ncart(D::Lb - 1) : 1;
    constexpr int ncp = ncart(D::Lc + 1), ncm = D::Lc > 0 ? ncart(D::Lc - 1) : 1;
    const int lane = threadIdx.x;
    grad_contract<D>(t.pb, t.pe, t.qb, t.qe, bb, i, bi, kb, j, ki, dn, md.primlim, s, boys_tab);
    coop_hrr<D>(bi, ki, s);

    const int n = md.nbasis;
    const auto P = gptr(md.P);
    SP Dab = s + D::DM;
    SP Dcd = Dab + na * nb;
    SP Dac = Dcd + nc * nd;
    SP Dad = Dac + na * nc;
    SP Dbc = Dad + na * nd;
    SP Dbd = Dbc + nb * nc;
    for (int x = lane; x < na * nb; x += D::NT) Dab[x] = P[(bi.fa + x / nb) * n + bi.fb + x % nb];
    for (int x = lane; x < nc * nd; x += D::NT) Dcd[x] = P[(ki.fa + x / nd) * n + ki.fb + x % nd];
    for (int x = lane; x < na * nc; x += D::NT) Dac[x] = P[(bi.fa + x / nc) * n + ki.fa + x % nc];
    for (int x = lane; x < na * nd; x += D::NT) Dad[x] = P[(bi.fa + x / nd) * n + ki.fb + x % nd];
    for (int x = lane; x < nb * nc; x += D::NT) Dbc[x] = P[(bi.fb + x / nc) * n + ki.fa + x % nc];
    for (int x = lane; x < nb * nd; x += D::NT) Dbd[x] = P[(bi.fb + x / nd) * n + ki.fb + x % nd];
    __syncthreads();

    const double sdeg = 0.25 * (bi.qa != bi.qb ? 2.0 : 1.0) * (ki.qa != ki.qb ? 2.0 : 1.0) * (bi.id != ki.id ? 2.0 : 1.0);
    const double hh = 0.5 * md.hyb;
    const auto gplus = gptr(&kGradPlus[0][0]);
    const auto gminus = gptr(&kGradMinus[0][0]);
    const auto gpow = gptr(&kGradPow[0][0]);
#pragma unroll
    for (int k = 0; k < 9; ++k) g[k] = 0.0;
    for (int x = lane; x < D::NI; x += D::NT) {
        const int d = x % nd, c = (x / nd) % nc, b = (x / (nd * nc)) % nb, a = x / (nd * nc * nb);
        const double cons = kCons[D::La][a] * kCons[D::Lb][b] * kCons[D::Lc][c] * kCons[D::Ld][d];
        const double gam = sdeg * cons *
                           (2.0 * Dab[a * nb + b] * Dcd[c * nd + d] -
                            hh * (Dac[a * nc + c] * Dbd[b * nd + d] + Dad[a * nd + d] * Dbc[b * nc + c]));
#pragma unroll
        for (int dir = 0; dir < 3; ++dir) {
            const int ia = D::La * 18 + a * 3 + dir, ib = D::Lb * 18 + b * 3 + dir, ic = D::Lc * 18 + c * 3 + dir;
            double vA = s[D::OAP + ((gplus[ia] * nb + b) * nc + c) * nd + d];
            if (D::OAM >= 0 && gpow[ia]) vA -= double(gpow[ia]) * s[D::OAM + ((gminus[ia] * nb + b) * nc + c) * nd + d];
            double vB = s[D::OBP + ((a * nbp + gplus[ib]) * nc + c) * nd + d];
            if (D::OBM >= 0 && gpow[ib]) vB -= double(gpow[ib]) * s[D::OBM + ((a * nbm + gminus[ib]) * nc + c) * nd + d];
            double vC = s[D::OCP + ((a * nb + b) * ncp + gplus[ic]) * nd + d];
            if (D::OCM >= 0 && gpow[ic]) vC -= double(gpow[ic]) * s[D::OCM + ((a * nb + b) * ncm + gminus[ic]) * nd + d];
            g[dir] += gam * vA;
            g[3 + dir] += gam * vB;
            g[6 + dir] += gam * vC;
        }
    }
    (void)nap;
    (void)nam;
#pragma unroll
    for (int k = 0; k < 9; ++k) g[k] = wave_sum_all_active(g[k]);
    __syncthreads();   // the next quartet reuses the slots
}

// One workgroup (one wave) per task, as fock_coop_kernel, with QUICK's
// gradient screen in md.cutoff; fp32 = 1 for the tasks of a Fock launch that
// holds the FP32 half of a split class (the gradient itself is FP64).  Dynamic LDS: the slots (unless they are
// global), then `tile` doubles for the workgroup's gradient (0: none, the
// sums go straight to md.grad).
template <class D>
__global__ __launch_bounds__(D::NT) void grad_coop_kernel(const Task* __restrict__ tasks, int ntasks,
                                                         const PairBlock* __restrict__ blocks,
                                                         const MolDesc* __restrict__ mols, double* gscratch, int tile,
                                                         int fp32)
{
    static_assert(D::NT == WAVE, "one wave per workgroup");
    QF_BOYS_LDS_DECL;
    QF_BOYS_LDS_FILL();
    extern __shared__ double qf_lds[];
    const auto s = coop_slots<D>(gscratch);
    QF_AS_LDS double* gt = (QF_AS_LDS double*)qf_lds + (D::GLOBAL_SLOTS ? 0 : D::LDS);
    const int lane = threadIdx.x % WAVE;
    for (int x = lane; x < tile; x += WAVE) gt[x] = 0.0;
    __syncthreads();
    const MolDesc* md_last = nullptr;
    for (int ti = blockIdx.x; ti < ntasks; ti += gridDim.x) {
        const Task t = cptr(tasks)[ti];
        const PairBlock bb = cptr(blocks)[t.bra_block];
        const PairBlock kb = cptr(blocks)[t.ket_block];
        const MolDesc md = cptr(mols)[bb.mol];
        md_last = mols + bb.mol;
        const int ns = md.nshell;
        const auto cm = gptr(md.cutmat);
        const auto sat = gptr(md.shell_atom);
        const auto bperm = cptr(bb.bperm);
        const bool in_range = lane < t.ket_count;
        const int jl = t.ket_start + (in_range ? lane : 0);
        const PairInfo kl = gptr(kb.info)[jl];
        const int atC = sat[kl.sa], atD = sat[kl.sb];
        for (int kk = t.kb; kk < t.ke; ++kk) {
            const int i = bperm[kk];
            const PairInfo bi = cptr(bb.info)[i];
            const int atA = sat[bi.sa], atB = sat[bi.sb];
            const double ytest = bi.ycut * kl.ycut;
            double dn = fmax(4.0 * cm[bi.sa * ns + bi.sb], 4.0 * cm[kl.sa * ns + kl.sb]);
            dn = fmax(dn, cm[bi.sa * ns + kl.sb]);
            dn = fmax(dn, cm[bi.sa * ns + kl.sa]);
            dn = fmax(dn, cm[bi.sb * ns + kl.sa]);
            dn = fmax(dn, cm[bi.sb * ns + kl.sb]);
            const bool one_atom = atA == atB && atA == atC && atA == atD;
            // the two precision halves of a split class share boundary windows: each
            // launch keeps its own side of md.split, as the Fock kernels do
            const bool half = fp32 ? (ytest < md.split) : (t.all_fp64 || ytest >= md.split);
            const bool valid = in_range && !one_atom && half && (ytest > md.cutoff) && (ytest * dn > md.cutoff) &&
                               (!t.same_block || jl <= i);
            unsigned long long mask = __ballot(valid);
            while (mask) {
                const int v = __builtin_ctzll(mask);
                mask &= mask - 1;
                const int j = t.ket_start + v;
                const PairInfo ki = cptr(kb.info)[j];
                const double dnv = __shfl(dn, v, WAVE);
                const int atCv = __shfl(atC, v, WAVE), atDv = __shfl(atD, v, WAVE);
                double g[9];
                grad_quartet<D>(t, bb, i, bi, kb, j, ki, dnv, md, s, QF_BOYS_LDS_PTR, g);
                // lanes 0-11: A, B, C, D x y z
                if (lane < 12) {
                    const int which = lane / 3, dir = lane % 3;
                    double val;
                    if (which < 3) {
                        val = g[3 * which + dir];
                    } else {
                        val = -(g[dir] + g[3 + dir] + g[6 + dir]);
                    }
                    const int at = which == 0 ? atA : which == 1 ? atB : which == 2 ? atCv : atDv;
                    if (tile)
                        lds_add(gt + 3 * at + dir, val);
                    else
                        atomic_add(gptr_rw(md.grad) + 3 * at + dir, val);
                }
            }
        }
    }
    __syncthreads();
    if (tile && md_last) {
        const MolDesc md = *md_last;
        for (int x = lane; x < tile; x += WAVE)
            if (gt[x] != 0.0) atomic_add(gptr_rw(md.grad) + x, gt[x]);
    }
}

// ---------------------------------------------------------------------------
// Per-lane gradient kernel for the light classes: the layout of the per-lane
// Fock kernels (one ket pair per lane, the wave's bras sharing the first
// Q-shell a), with the raised VRR, the four weighted sets, the six HRR
// blocks and the contraction with the density in registers (generated
// C::vrr and C::digest).  Atom A is the same for the whole task and atoms
// C, D for a lane, so dE/dA, dE/dC and dE/dD are summed in registers over
// the bra loop; dE/dB is summed over the wave per bra.  Sums go to the
// workgroup's LDS gradient (or straight to md.grad when it does not fit).
// ---------------------------------------------------------------------------
template <class C, typename Tab>
__device__ __forceinline__ void grad_lane_contract(const PairBlock& bb, int i, const PairInfo& bi, const PairBlock& kb,
                                                   int j, const PairInfo& ki, double dn, double primlim, int maxm,
                                                   double (&acc)[C::NACC], Tab boys_tab, int pb, int pe, int qb,
                                                   int qe)
{
    const auto bpc = cptr(bb.prim);
    const auto kp = gptr(kb.prim);
    const int bn = bb.n, bK = bb.K, kn = kb.n, kK = kb.K;
    const size_t bfs = (size_t)bK * bn, kfs = (size_t)kK * kn;
#define BP(f, q) (*(const QF_AS_CONST double*)((const QF_AS_CONST char*)(bpc + (f) * bfs) + (uint32_t)((q) * bn + i) * 8u))
#define KP(f, off) (*(const QF_AS_GLOBAL double*)((const QF_AS_GLOBAL char*)(kp + (f) * kfs) + (off)))
    const double c0 = dn * BP(PF_CUT, 0);
    for (int jp = qb; jp < qe; ++jp) {
        const uint32_t ko = (uint32_t)(jp * kn + j) * 8u;
        const double cutk = KP(PF_CUT, ko);
        if (!(c0 * cutk > primlim)) break;
        const double zk = KP(PF_ZETA, ko);
        const double ozk = KP(PF_OOZ2, ko);
        const double Kk = KP(PF_KC, ko);
        const double gk = KP(PF_ALPHA, ko);
        const double Qk[3] = {KP(PF_PX, ko), KP(PF_PY, ko), KP(PF_PZ, ko)};
        const double QC[3] = {Qk[0] - ki.A[0], Qk[1] - ki.A[1], Qk[2] - ki.A[2]};
        for (int ip0 = pb; ip0 < pe; ++ip0) {
            const int ip = wave_uniform(ip0);    // the active lanes agree: scalar loads of the bra data
            if (!((dn * BP(PF_CUT, ip)) * cutk > primlim)) break;
            const double zb = BP(PF_ZETA, ip);
            const double ozb = BP(PF_OOZ2, ip);
            const double ab = BP(PF_ALPHA, ip);
            const double Pb[3] = {BP(PF_PX, ip), BP(PF_PY, ip), BP(PF_PZ, ip)};
            const double Kb = BP(PF_KC, ip);
            const double rs = qf_rsqrt(zb + zk);
            const double inv = rs * rs;
            const double PQ[3] = {Pb[0] - Qk[0], Pb[1] - Qk[1], Pb[2] - Qk[2]};
            const double rpq = PQ[0] * PQ[0] + PQ[1] * PQ[1] + PQ[2] * PQ[2];
            const double T = rpq * (zb * zk * inv);
            double F[C::L + 1];
            boys_quick<C::L, Tab, 4>(T, maxm, F, boys_tab);
            const double pref = Kb * Kk * rs;
#pragma unroll
            for (int m = 0; m <= C::L; ++m) F[m] *= pref;
            VrrPar<double> p;
            const double wp = zk * inv, wq = zb * inv;
#pragma unroll
            for (int d = 0; d < 3; ++d) {
                p.PA[d] = Pb[d] - bi.A[d];
                p.WP[d] = -wp * PQ[d];
                p.QC[d] = QC[d];
                p.WQ[d] = wq * PQ[d];
            }
            p.oo2z = ozb;
            p.oo2n = ozk;
            p.oo2zn = 0.5 * inv;
            p.rz = wp;
            p.rn = wq;
            const double w[4] = {1.0, 2.0 * ab, 2.0 * (zb - ab), 2.0 * gk};
            C::vrr(F, p, w, acc);
        }
    }
#undef BP
#undef KP
}

template <class C>
__global__ __launch_bounds__(WAVE) void grad_lane_kernel(const Task* __restrict__ tasks, int ntasks,
                                                        const PairBlock* __restrict__ blocks,
                                                        const MolDesc* __restrict__ mols, int tile, int fp32)
{
    constexpr int na = ncart(C::La), nb = ncart(C::Lb), nc = ncart(C::Lc), nd = ncart(C::Ld);
    QF_BOYS_LDS_DECL;
    QF_BOYS_LDS_FILL();
    extern __shared__ double qf_lds[];
    QF_AS_LDS double* gt = (QF_AS_LDS double*)qf_lds;
    const int lane = threadIdx.x;
    for (int x = lane; x < tile; x += WAVE) gt[x] = 0.0;
    __syncthreads();
    const MolDesc* md_last = nullptr;
    for (int ti = blockIdx.x; ti < ntasks; ti += gridDim.x) {
        const Task t = cptr(tasks)[ti];
        const PairBlock bb = cptr(blocks)[t.bra_block];
        const PairBlock kb = cptr(blocks)[t.ket_block];
        const MolDesc md = cptr(mols)[bb.mol];
        md_last = mols + bb.mol;
        const auto grad = gptr_rw(md.grad);
        const int ns = md.nshell, n = md.nbasis;
        const auto cm = gptr(md.cutmat);
        const auto sat = gptr(md.shell_atom);
        const auto P = gptr(md.P);
        const auto bperm = cptr(bb.bperm);
        const bool in_range = lane < t.ket_count;
        const int j = t.ket_start + (in_range ? lane : 0);
        const PairInfo ki = gptr(kb.info)[j];
        const int atC = sat[ki.sa], atD = sat[ki.sb];
        const double hh = 0.5 * md.hyb;
        double gA[3] = {0.0, 0.0, 0.0}, gC[3] = {0.0, 0.0, 0.0}, gD[3] = {0.0, 0.0, 0.0};
        int atA = 0;
        // density blocks of the ket pair (the same for every bra)
        double Dcd[nc * nd];
#pragma unroll
        for (int x = 0; x < nc * nd; ++x) Dcd[x] = P[(ki.fa + x / nd) * n + ki.fb + x % nd];
        for (int kk0 = t.kb; kk0 < t.ke; ++kk0) {
            const int kk = wave_uniform(kk0);
            const int i = bperm[kk];
            const PairInfo bi = cptr(bb.info)[i];
            atA = sat[bi.sa];
            const int atB = sat[bi.sb];
            const double ytest = bi.ycut * ki.ycut;
            double dn = fmax(4.0 * cm[bi.sa * ns + bi.sb], 4.0 * cm[ki.sa * ns + ki.sb]);
            dn = fmax(dn, cm[bi.sa * ns + ki.sb]);
            dn = fmax(dn, cm[bi.sa * ns + ki.sa]);
            dn = fmax(dn, cm[bi.sb * ns + ki.sa]);
            dn = fmax(dn, cm[bi.sb * ns + ki.sb]);
            const bool one_atom = atA == atB && atA == atC && atA == atD;
            const bool half = fp32 ? (ytest < md.split) : (t.all_fp64 || ytest >= md.split);
            const bool valid = in_range && !one_atom && half && (ytest > md.cutoff) && (ytest * dn > md.cutoff) &&
                               (!t.same_block || j <= i);
            double g[9];
#pragma unroll
            for (int k = 0; k < 9; ++k) g[k] = 0.0;
            if (valid) {
                double acc[C::NACC];
#pragma unroll
                for (int x = 0; x < C::NACC; ++x) acc[x] = 0.0;
                grad_lane_contract<C>(bb, i, bi, kb, j, ki, dn, md.primlim, bi.qfs + ki.qfs + 2, acc,
                                      QF_BOYS_LDS_PTR, t.pb, t.pe, t.qb, t.qe);
                double Dab[na * nb], Dac[na * nc], Dad[na * nd], Dbc[nb * nc], Dbd[nb * nd];
#pragma unroll
                for (int x = 0; x < na * nb; ++x) Dab[x] = P[(bi.fa + x / nb) * n + bi.fb + x % nb];
#pragma unroll
                for (int x = 0; x < na * nc; ++x) Dac[x] = P[(bi.fa + x / nc) * n + ki.fa + x % nc];
#pragma unroll
                for (int x = 0; x < na * nd; ++x) Dad[x] = P[(bi.fa + x / nd) * n + ki.fb + x % nd];
#pragma unroll
                for (int x = 0; x < nb * nc; ++x) Dbc[x] = P[(bi.fb + x / nc) * n + ki.fa + x % nc];
#pragma unroll
                for (int x = 0; x < nb * nd; ++x) Dbd[x] = P[(bi.fb + x / nd) * n + ki.fb + x % nd];
                C::digest(acc, bi.AB, ki.AB, Dab, Dcd, Dac, Dad, Dbc, Dbd, hh, g);
                const double sdeg = 0.25 * (bi.qa != bi.qb ? 2.0 : 1.0) * (ki.qa != ki.qb ? 2.0 : 1.0) *
                                    (bi.id != ki.id ? 2.0 : 1.0);
#pragma unroll
                for (int k = 0; k < 9; ++k) g[k] *= sdeg;
            }
#pragma unroll
            for (int x = 0; x < 3; ++x) {
                gA[x] += g[x];
                gC[x] += g[6 + x];
                gD[x] -= g[x] + g[3 + x] + g[6 + x];
                const double b = wave_sum_all_active(g[3 + x]);
                if (lane == 0) {
                    if (tile)
                        lds_add(gt + 3 * atB + x, b);
                    else if (b != 0.0)
                        atomic_add(grad + 3 * atB + x, b);
                }
            }
        }
#pragma unroll
        for (int x = 0; x < 3; ++x) {
            const double a = wave_sum_all_active(gA[x]);
            if (tile) {
                if (lane == 0) lds_add(gt + 3 * atA + x, a);
                if (in_range) {
                    lds_add(gt + 3 * atC + x, gC[x]);
                    lds_add(gt + 3 * atD + x, gD[x]);
                }
            } else {
                if (lane == 0 && a != 0.0) atomic_add(grad + 3 * atA + x, a);
                if (in_range && gC[x] != 0.0) atomic_add(grad + 3 * atC + x, gC[x]);
                if (in_range && gD[x] != 0.0) atomic_add(grad + 3 * atD + x, gD[x]);
            }
        }
    }
    __syncthreads();
    if (tile && md_last) {
        const MolDesc md = *md_last;
        for (int x = lane; x < tile; x += WAVE)
            if (gt[x] != 0.0) atomic_add(gptr_rw(md.grad) + x, gt[x]);
    }
}

}  // namespace qf
