// Wave-per-quartet Fock kernel for the d-heavy classes.
//
// The per-lane kernels give every lane its own quartet, so a (dd|dd)
// quartet's 961 accumulators and ~2,200 VRR intermediates live in scratch
// memory and the launch is latency-bound at one wave per SIMD.  Here the 64
// lanes of one wave (the whole workgroup) share ONE quartet: the generated
// node programs (gen_eri.coop_tables) evaluate the VRR level by level - every
// node of a level is independent, one node per lane - and then the HRR the
// same way, with all values in LDS slots that the generator reuses by
// liveness, so several workgroups fit on one CU.  A VRR node whose value is a
// target also adds itself into its contracted accumulator.  Digestion is
// lane-parallel too: each lane owns output elements of J_ab, J_cd, K_ac,
// K_ad, K_bc, K_bd and sums its contraction over the integral block.
//
// Classes whose slots exceed the LDS cap keep them in a global scratch
// buffer instead (one slice per workgroup, GLOBAL_SLOTS): the same code with
// global instead of LDS addresses.
//
// The arithmetic is the same as in fock_task (same screening, same Boys
// function, same primitive loops and prefactors, same digestion weights);
// only the evaluation order of the recursions differs.
#pragma once
#include <type_traits>

#include "qf_device.h"

namespace qf {

struct CoopVNode {
    // value = C1 s[a] + C2 s[b] + n1 H (s[c] - Rr s[d]) + n2 (1/2(z+n)) s[e], written to s[dst]
    // (and added to s[acc] unless acc == 0xFFFF); meta = sel | n1 << 4 | n2 << 8;
    // (C1, C2, H, Rr) = (PA, WP, 1/2z, rho/z) for sel = dir (bra step) and
    // (QC, WQ, 1/2n, rho/n) for sel = 3 + dir (ket step)
    unsigned short a, b, c, d, e, dst, acc, meta;
};

struct CoopHNode {
    // s[dst] = s[s1] + X s[s2], X = CD[dir] (sel = dir) or AB[dir] (sel = 3 + dir)
    unsigned short dst, s1, s2, sel;
};

constexpr int kVU = 2;   // VRR nodes per lane in flight
constexpr int kHU = 4;   // HRR nodes per lane in flight

// Runs a level-scheduled node program: levels [lev[l], lev[l+1]) of nodes
// `prog`, U nodes per lane per round, a barrier after each level.  The
// rounds are one uniform sequence over all levels, and the descriptors of
// the next round are loaded before the current round reads and writes its
// slots, so their global-memory latency overlaps the LDS work and the
// barriers.  exec(node, x_valid) evaluates one node.
template <int U, int NT, typename Node, typename Exec>
__device__ __forceinline__ void run_levels(const QF_AS_GLOBAL Node* prog, const QF_AS_CONST int* lev, int nlev,
                                           Exec exec)
{
    const int lane = threadIdx.x;
    if (nlev <= 0) return;
    int lv = 0, e1 = lev[1], x0 = lev[0];
    Node q[U], qn[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
        const int x = x0 + u * NT + lane;
        q[u] = prog[x < e1 ? x : e1 - 1];
    }
    for (;;) {
        int nx0 = x0 + U * NT, nlv = lv, ne1 = e1;
        const bool level_end = nx0 >= e1;
        if (level_end) {
            nlv = lv + 1;
            if (nlv < nlev) {
                nx0 = lev[nlv];
                ne1 = lev[nlv + 1];
            }
        }
        if (nlv < nlev) {
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const int x = nx0 + u * NT + lane;
                qn[u] = prog[x < ne1 ? x : ne1 - 1];
            }
        }
        exec(q, x0 + lane, e1);
        if (level_end) {
            __syncthreads();
            if (nlv >= nlev) break;
        }
        lv = nlv;
        x0 = nx0;
        e1 = ne1;
#pragma unroll
        for (int u = 0; u < U; ++u) q[u] = qn[u];
    }
}

// Slot memory of a class: LDS, or a global scratch slice.
template <class D>
using CoopSlots = typename std::conditional<D::GLOBAL_SLOTS, QF_AS_GLOBAL double*, QF_AS_LDS double*>::type;

// Contracted [e0|f0] accumulators of one quartet into s[0, NACC): the
// primitive loops with QUICK's screen and order, as in contract().  The
// bra/ket primitive-pair ranges [pb, pe), [qb, qe) select a piece of the
// quartet.  With primlim < 0 nothing is screened (Schwarz bounds).
template <class D, typename SP>
__device__ __forceinline__ void coop_contract(int pb, int pe, int qb, int qe, const PairBlock& bb, int i,
                                              const PairInfo& bi, const PairBlock& kb, int j, const PairInfo& ki,
                                              double dn, double primlim, SP s, const QF_AS_LDS double* boys_tab)
{
    const int lane = threadIdx.x;
    SP cf = s + D::CF;   // [0,24) VRR coefficients, 24: 1/2(z+n), [25,31): CD, AB
    for (int x = lane; x < D::NACC; x += D::NT) s[x] = 0.0;

    const auto bp = cptr(bb.prim);
    const auto kp = cptr(kb.prim);
    const int bn = bb.n, bK = bb.K, kn = kb.n, kK = kb.K;
    const int maxm = bi.qfs + ki.qfs;
    const double c0 = dn * bp[(PF_CUT * bK) * bn + i];
    const auto vrr = gptr(D::vrr());
    const auto vlev = cptr(D::vlev());
    for (int jp = qb; jp < qe; ++jp) {
        const double cutk = kp[(PF_CUT * kK + jp) * kn + j];
        if (!(c0 * cutk > primlim)) break;
        const double zk = kp[(PF_ZETA * kK + jp) * kn + j];
        const double ozk = kp[(PF_OOZ2 * kK + jp) * kn + j];
        const double Qk[3] = {kp[(PF_PX * kK + jp) * kn + j], kp[(PF_PY * kK + jp) * kn + j],
                              kp[(PF_PZ * kK + jp) * kn + j]};
        const double Kk = kp[(PF_KC * kK + jp) * kn + j];
        const double QC[3] = {Qk[0] - ki.A[0], Qk[1] - ki.A[1], Qk[2] - ki.A[2]};
        for (int ip = pb; ip < pe; ++ip) {
            if (!((dn * bp[(PF_CUT * bK + ip) * bn + i]) * cutk > primlim)) break;
            const double zb = bp[(PF_ZETA * bK + ip) * bn + i];
            const double ozb = bp[(PF_OOZ2 * bK + ip) * bn + i];
            const double Pb[3] = {bp[(PF_PX * bK + ip) * bn + i], bp[(PF_PY * bK + ip) * bn + i],
                                  bp[(PF_PZ * bK + ip) * bn + i]};
            const double Kb = bp[(PF_KC * bK + ip) * bn + i];
            const double rs = qf_rsqrt(zb + zk);
            const double inv = rs * rs;
            const double PQ[3] = {Pb[0] - Qk[0], Pb[1] - Qk[1], Pb[2] - Qk[2]};
            const double rpq = PQ[0] * PQ[0] + PQ[1] * PQ[1] + PQ[2] * PQ[2];
            const double T = rpq * (zb * zk * inv);
            double F[D::L + 1];
            boys_quick<D::L>(T, maxm, F, boys_tab);
            const double pref = Kb * Kk * rs;
            const double wp = zk * inv, wq = zb * inv;
            __syncthreads();   // the previous primitive's last level has read its sources
#pragma unroll
            for (int m = 0; m <= D::L; ++m)
                if (lane == m) s[D::FSLOT[m]] = F[m] * pref;
            if (lane == 0) {
                // the values are the same for the whole wave (one quartet): one
                // lane stores them with compile-time indices (a lane-indexed read
                // would place the pair records and these arrays in scratch)
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
                s[D::ZS] = 0.0;
            }
            __syncthreads();
            // nodes of one level are independent: each lane reads the sources
            // of its kVU nodes, then computes and stores, so the LDS latencies
            // overlap instead of chaining node by node
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
                    if (x0 + u * D::NT < e1) {
                        s[q[u].dst] = v[u];
                        if (q[u].acc != 0xFFFF) s[q[u].acc] += v[u];
                    }
            });
        }
    }
}

// HRR on the contracted block in s (the VRR values are dead now); the
// integral of output index x then sits at s[D::OUT + x].
template <class D, typename SP>
__device__ __forceinline__ void coop_hrr(const PairInfo& bi, const PairInfo& ki, SP s)
{
    const int lane = threadIdx.x;
    SP cf = s + D::CF;
    __syncthreads();
    if (lane == 0) {
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            cf[25 + d] = ki.AB[d];
            cf[28 + d] = bi.AB[d];
        }
        s[D::ZS] = 0.0;
    }
    __syncthreads();
    const auto hrr = gptr(D::hrr());
    const auto hlev = cptr(D::hlev());
    run_levels<kHU, D::NT>(hrr, hlev, D::NLH, [&](const CoopHNode (&q)[kHU], int x0, int e1) {
        double v[kHU];
#pragma unroll
        for (int u = 0; u < kHU; ++u) v[u] = s[q[u].s1] + cf[25 + q[u].sel] * s[q[u].s2];
#pragma unroll
        for (int u = 0; u < kHU; ++u)
            if (x0 + u * D::NT < e1) s[q[u].dst] = v[u];
    });
}

// OS (the fused open-shell pass): after the six blocks of md.P the
// four exchange blocks are reloaded from md.P2 and the exchange terms are
// contracted a second time into G2.
template <class D, bool OS, typename SP>
__device__ __forceinline__ void coop_quartet(const Task& t, const PairBlock& bb, int i, const PairInfo& bi,
                                             const PairBlock& kb, int j, const PairInfo& ki, double dn,
                                             const MolDesc& md, QF_AS_GLOBAL double* G, QF_AS_GLOBAL double* G2,
                                             SP s, const QF_AS_LDS double* boys_tab)
{
    constexpr int na = ncart(D::La), nb = ncart(D::Lb), nc = ncart(D::Lc), nd = ncart(D::Ld);
    const int lane = threadIdx.x;
    coop_contract<D>(t.pb, t.pe, t.qb, t.qe, bb, i, bi, kb, j, ki, dn, md.primlim, s, boys_tab);
    coop_hrr<D>(bi, ki, s);

    // integrals with cons and degeneracy weight, in place; density blocks
    const double sdeg = 0.25 * (bi.qa != bi.qb ? 2.0 : 1.0) * (ki.qa != ki.qb ? 2.0 : 1.0) * (bi.id != ki.id ? 2.0 : 1.0);
    SP Ic = s + D::OUT;
    for (int x = lane; x < D::NI; x += D::NT) {
        const int d = x % nd, c = (x / nd) % nc, b = (x / (nd * nc)) % nb, a = x / (nd * nc * nb);
        const double cons = kCons[D::La][a] * kCons[D::Lb][b] * kCons[D::Lc][c] * kCons[D::Ld][d];
        Ic[x] = Ic[x] * (cons * sdeg);
    }
    const int n = md.nbasis;
    const auto P = gptr(md.P);
    SP Dab = s + D::DM;
    SP Dcd = Dab + na * nb;
    SP Dac = Dcd + nc * nd;
    SP Dad = Dac + na * nc;
    SP Dbc = Dad + na * nd;
    SP Dbd = Dbc + nb * nc;
    const auto PJ = gptr(md.PJ);              // Coulomb density (zeros in an exchange-only pass)
    for (int x = lane; x < na * nb; x += D::NT) Dab[x] = PJ[(bi.fa + x / nb) * n + bi.fb + x % nb];
    for (int x = lane; x < nc * nd; x += D::NT) Dcd[x] = PJ[(ki.fa + x / nd) * n + ki.fb + x % nd];
    for (int x = lane; x < na * nc; x += D::NT) Dac[x] = P[(bi.fa + x / nc) * n + ki.fa + x % nc];
    for (int x = lane; x < na * nd; x += D::NT) Dad[x] = P[(bi.fa + x / nd) * n + ki.fb + x % nd];
    for (int x = lane; x < nb * nc; x += D::NT) Dbc[x] = P[(bi.fb + x / nc) * n + ki.fa + x % nc];
    for (int x = lane; x < nb * nd; x += D::NT) Dbd[x] = P[(bi.fb + x / nd) * n + ki.fb + x % nd];
    __syncthreads();

    // lane-owned contractions, one G element each
    const double kscale = -0.25 * md.hyb;
    constexpr int o1 = na * nb, o2 = o1 + nc * nd, o3 = o2 + na * nc, o4 = o3 + na * nd, o5 = o4 + nb * nc,
                  o6 = o5 + nb * nd;
#define IX(a, b, c, d) Ic[(((a) * nb + (b)) * nc + (c)) * nd + (d)]
    for (int o = lane; o < o6; o += D::NT) {
        double v = 0.0;
        int row, col;
        if (o < o1) {                              // J_ab = sum_cd D_cd I
            const int a = o / nb, b = o % nb;
            for (int c = 0; c < nc; ++c)
                for (int d = 0; d < nd; ++d) v += Dcd[c * nd + d] * IX(a, b, c, d);
            row = bi.fa + a;
            col = bi.fb + b;
        } else if (o < o2) {                       // J_cd = sum_ab D_ab I
            const int c = (o - o1) / nd, d = (o - o1) % nd;
            for (int a = 0; a < na; ++a)
                for (int b = 0; b < nb; ++b) v += Dab[a * nb + b] * IX(a, b, c, d);
            row = ki.fa + c;
            col = ki.fb + d;
        } else if (o < o3) {                       // K_ac = sum_bd D_bd I
            const int a = (o - o2) / nc, c = (o - o2) % nc;
            for (int b = 0; b < nb; ++b)
                for (int d = 0; d < nd; ++d) v += Dbd[b * nd + d] * IX(a, b, c, d);
            v *= kscale;
            row = bi.fa + a;
            col = ki.fa + c;
        } else if (o < o4) {                       // K_ad = sum_bc D_bc I
            const int a = (o - o3) / nd, d = (o - o3) % nd;
            for (int b = 0; b < nb; ++b)
                for (int c = 0; c < nc; ++c) v += Dbc[b * nc + c] * IX(a, b, c, d);
            v *= kscale;
            row = bi.fa + a;
            col = ki.fb + d;
        } else if (o < o5) {                       // K_bc = sum_ad D_ad I
            const int b = (o - o4) / nc, c = (o - o4) % nc;
            for (int a = 0; a < na; ++a)
                for (int d = 0; d < nd; ++d) v += Dad[a * nd + d] * IX(a, b, c, d);
            v *= kscale;
            row = bi.fb + b;
            col = ki.fa + c;
        } else {                                   // K_bd = sum_ac D_ac I
            const int b = (o - o5) / nd, d = (o - o5) % nd;
            for (int a = 0; a < na; ++a)
                for (int c = 0; c < nc; ++c) v += Dac[a * nc + c] * IX(a, b, c, d);
            v *= kscale;
            row = bi.fb + b;
            col = ki.fb + d;
        }
        if (v != 0.0) atomic_add(&G[row * n + col], v);
    }
    if constexpr (OS) {
        __syncthreads();
        const auto P2 = gptr(md.P2);
        for (int x = lane; x < na * nc; x += D::NT) Dac[x] = P2[(bi.fa + x / nc) * n + ki.fa + x % nc];
        for (int x = lane; x < na * nd; x += D::NT) Dad[x] = P2[(bi.fa + x / nd) * n + ki.fb + x % nd];
        for (int x = lane; x < nb * nc; x += D::NT) Dbc[x] = P2[(bi.fb + x / nc) * n + ki.fa + x % nc];
        for (int x = lane; x < nb * nd; x += D::NT) Dbd[x] = P2[(bi.fb + x / nd) * n + ki.fb + x % nd];
        __syncthreads();
        for (int o = o2 + lane; o < o6; o += D::NT) {
            double v = 0.0;
            int row, col;
            if (o < o3) {
                const int a = (o - o2) / nc, c = (o - o2) % nc;
                for (int b = 0; b < nb; ++b)
                    for (int d = 0; d < nd; ++d) v += Dbd[b * nd + d] * IX(a, b, c, d);
                row = bi.fa + a;
                col = ki.fa + c;
            } else if (o < o4) {
                const int a = (o - o3) / nd, d = (o - o3) % nd;
                for (int b = 0; b < nb; ++b)
                    for (int c = 0; c < nc; ++c) v += Dbc[b * nc + c] * IX(a, b, c, d);
                row = bi.fa + a;
                col = ki.fb + d;
            } else if (o < o5) {
                const int b = (o - o4) / nc, c = (o - o4) % nc;
                for (int a = 0; a < na; ++a)
                    for (int d = 0; d < nd; ++d) v += Dad[a * nd + d] * IX(a, b, c, d);
                row = bi.fb + b;
                col = ki.fa + c;
            } else {
                const int b = (o - o5) / nd, d = (o - o5) % nd;
                for (int a = 0; a < na; ++a)
                    for (int c = 0; c < nc; ++c) v += Dac[a * nc + c] * IX(a, b, c, d);
                row = bi.fb + b;
                col = ki.fb + d;
            }
            v *= kscale;
            if (v != 0.0) atomic_add(&G2[row * n + col], v);
        }
    }
#undef IX
    __syncthreads();   // the next quartet reuses the slots
}

// Slot memory of this workgroup: dynamic LDS, or its slice of the global
// scratch buffer (gscratch holds gridDim.x slices of D::LDS doubles).
template <class D>
__device__ __forceinline__ CoopSlots<D> coop_slots(double* gscratch)
{
    if constexpr (D::GLOBAL_SLOTS) {
        return (QF_AS_GLOBAL double*)(gscratch + (size_t)blockIdx.x * D::LDS);
    } else {
        extern __shared__ double qf_lds[];
        return (QF_AS_LDS double*)qf_lds;
    }
}

// One workgroup (one wave) per task of a narrow ket window: the wave
// screens the task's quartets lane-parallel, then processes the surviving
// quartets one after another.  FP64 throughout.  Dynamic LDS: D::LDS doubles
// (none with GLOBAL_SLOTS).
template <class D, bool OS = false>
__global__ __launch_bounds__(D::NT) void fock_coop_kernel(const Task* __restrict__ tasks, int ntasks,
                                                         const PairBlock* __restrict__ blocks,
                                                         const MolDesc* __restrict__ mols, double* gscratch)
{
    static_assert(D::NT == WAVE, "one wave per workgroup");
    QF_BOYS_LDS_DECL;
    QF_BOYS_LDS_FILL();
    const auto s = coop_slots<D>(gscratch);
    const int lane = threadIdx.x % WAVE;
    for (int ti = blockIdx.x; ti < ntasks; ti += gridDim.x) {
        const Task t = cptr(tasks)[ti];
        const PairBlock bb = cptr(blocks)[t.bra_block];
        const PairBlock kb = cptr(blocks)[t.ket_block];
        const MolDesc md = cptr(mols)[bb.mol];
        const int ns = md.nshell;
        const size_t roff = (size_t)(blockIdx.x % md.replicas) * md.nbasis * md.nbasis;
        const auto G = gptr_rw(md.G + roff);
        const auto G2 = gptr_rw(OS ? md.G2 + roff : md.G);
        const auto cm = gptr(md.cutmat);
        const auto bperm = cptr(bb.bperm);
        const bool in_range = lane < t.ket_count;
        const int jl = t.ket_start + (in_range ? lane : 0);
        const PairInfo kl = gptr(kb.info)[jl];
        for (int kk = t.kb; kk < t.ke; ++kk) {
            const int i = bperm[kk];
            const PairInfo bi = cptr(bb.info)[i];
            const double ytest = bi.ycut * kl.ycut;
            double dn = fmax(4.0 * cm[bi.sa * ns + bi.sb], 4.0 * cm[kl.sa * ns + kl.sb]);
            dn = fmax(dn, cm[bi.sa * ns + kl.sb]);
            dn = fmax(dn, cm[bi.sa * ns + kl.sa]);
            dn = fmax(dn, cm[bi.sb * ns + kl.sa]);
            dn = fmax(dn, cm[bi.sb * ns + kl.sb]);
            const bool valid = in_range && (ytest > md.cutoff) && (ytest * dn > md.cutoff) &&
                               (!t.same_block || jl <= i) && (t.all_fp64 || ytest >= md.split);
            unsigned long long mask = __ballot(valid);
            while (mask) {
                const int v = __builtin_ctzll(mask);
                mask &= mask - 1;
                const int j = t.ket_start + v;
                const PairInfo ki = cptr(kb.info)[j];
                const double dnv = wave_read_lane(dn, v);
                coop_quartet<D, OS>(t, bb, i, bi, kb, j, ki, dnv, md, G, G2, s, QF_BOYS_LDS_PTR);
            }
        }
    }
}

// Schwarz diagonal of a class (La Lb | La Lb) on the wave-per-quartet node
// programs: one workgroup per pair of blk (striding when the grid is capped),
// out[j] = max over components |(ab|ab)| with no screening (shellcutoff /
// classcutoff, quick_cutoff_module.f90:91-424), like schwarz_kernel for the
// classes that have per-lane kernels.
template <class D>
__global__ __launch_bounds__(D::NT) void schwarz_coop_kernel(PairBlock blk, double* __restrict__ out,
                                                            double* gscratch)
{
    constexpr int na = ncart(D::La), nb = ncart(D::Lb);
    static_assert(D::La == D::Lc && D::Lb == D::Ld, "diagonal class");
    QF_BOYS_LDS_DECL;
    QF_BOYS_LDS_FILL();
    const auto s = coop_slots<D>(gscratch);
    for (int j = blockIdx.x; j < blk.n; j += gridDim.x) {
        const PairInfo pi = cptr(blk.info)[j];
        coop_contract<D>(0, blk.K, 0, blk.K, blk, j, pi, blk, j, pi, 1.0, -1.0, s, QF_BOYS_LDS_PTR);
        coop_hrr<D>(pi, pi, s);
        double m = 0.0;
        for (int x = threadIdx.x; x < na * nb; x += D::NT) {
            const int a = x / nb, b = x % nb;
            const double cons = kCons[D::La][a] * kCons[D::Lb][b];
            m = fmax(m, fabs(s[D::OUT + ((a * nb + b) * na + a) * nb + b] * cons * cons));
        }
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) m = fmax(m, __shfl_xor(m, o, WAVE));
        if (threadIdx.x == 0) out[j] = m;
        __syncthreads();   // the next pair reuses the slots
    }
}

}  // namespace qf
