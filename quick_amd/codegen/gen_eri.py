"""Generate the per-class ERI recursions for the HIP kernels.

For every canonical class (La Lb | Lc Ld) with La >= Lb, Lc >= Ld, and bra
pair class >= ket pair class (21 classes for s, p, d) this emits a struct
with two straight-line device functions:

* ``vrr`` - Obara-Saika / Head-Gordon-Pople vertical recursion for one
  primitive quartet: builds [e0|f0]^(0) for every Cartesian e with
  La <= |e| <= La+Lb and f with Lc <= |f| <= Lc+Ld from [00|00]^(m) and adds
  them to the contracted accumulators (HGP 1988, eqs. 6 and 16 - the same
  recurrences QUICK's ``vertical`` / ``ERint_vertical_spd`` evaluate).
* ``hrr`` - horizontal recursion on the contracted accumulators:
  ket (e0|c d+1_i) = (e0|c+1_i d) + CD_i (e0|c d), then bra
  (a b+1_i|cd) = (a+1_i b|cd) + AB_i (a b|cd), producing the (ab|cd) block
  in QUICK's component order (s; x y z; xx xy yy xz yz zz).

The generator also counts the floating-point operations of each recursion
(FMA = 2), which is the per-class FLOP table used for the roofline.

    python -m quick_amd.codegen.gen_eri   # writes quick_amd/csrc/gen/*.h
"""
from __future__ import annotations

import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
OUT = os.path.join(os.path.dirname(HERE), "csrc", "gen")

QCOMP = {
    0: [(0, 0, 0)],
    1: [(1, 0, 0), (0, 1, 0), (0, 0, 1)],
    2: [(2, 0, 0), (1, 1, 0), (0, 2, 0), (1, 0, 1), (0, 1, 1), (0, 0, 2)],
}


def comps(L):
    if L in QCOMP:
        return QCOMP[L]
    out = []
    for l in range(L, -1, -1):
        for m in range(L - l, -1, -1):
            out.append((l, m, L - l - m))
    return out


def am(t):
    return t[0] + t[1] + t[2]


def sub(t, i, k=1):
    u = list(t)
    u[i] -= k
    return tuple(u)


def add(t, i, k=1):
    u = list(t)
    u[i] += k
    return tuple(u)


def tag(t):
    return "%d%d%d" % t


Z = (0, 0, 0)


class Vrr:
    """Dependency graph of the vertical recursion for one class."""

    def __init__(self, La, Lb, Lc, Ld):
        self.La, self.Lb, self.Lc, self.Ld = La, Lb, Lc, Ld
        self.L = La + Lb + Lc + Ld
        self.nodes = {}   # (e, f, m) -> (kind, deps)
        self.targets = []
        for E in range(La, La + Lb + 1):
            for e in comps(E):
                for F in range(Lc, Lc + Ld + 1):
                    for f in comps(F):
                        self.targets.append((e, f))
        for e, f in self.targets:
            self.need(e, f, 0)

    def choose_ket(self, e, f):
        best = None
        for i in range(3):
            if f[i] == 0:
                continue
            score = (e[i] > 0, f[i] > 1, -f[i])
            if best is None or score < best[0]:
                best = (score, i)
        return best[1]

    def choose_bra(self, e):
        best = None
        for i in range(3):
            if e[i] == 0:
                continue
            score = (e[i] > 1, -e[i])
            if best is None or score < best[0]:
                best = (score, i)
        return best[1]

    def need(self, e, f, m):
        key = (e, f, m)
        if key in self.nodes:
            return
        if f == Z and e == Z:
            self.nodes[key] = ("base", ())
            return
        if f == Z:
            i = self.choose_bra(e)
            e1 = sub(e, i)
            deps = [("PA", i, (e1, Z, m)), ("WP", i, (e1, Z, m + 1))]
            if e1[i] > 0:
                e2 = sub(e1, i)
                deps.append(("Z2", e1[i], (e2, Z, m), (e2, Z, m + 1)))
            self.nodes[key] = ("bra", (i, deps))
        else:
            i = self.choose_ket(e, f)
            f1 = sub(f, i)
            deps = [("QC", i, (e, f1, m)), ("WQ", i, (e, f1, m + 1))]
            if f1[i] > 0:
                f2 = sub(f1, i)
                deps.append(("N2", f1[i], (e, f2, m), (e, f2, m + 1)))
            if e[i] > 0:
                deps.append(("ZN", e[i], (sub(e, i), f1, m + 1)))
            self.nodes[key] = ("ket", (i, deps))
        for d in self.nodes[key][1][1] if self.nodes[key][0] != "base" else ():
            for x in d[2:]:
                self.need(*x)

    def order(self):
        return sorted(self.nodes, key=lambda k: (am(k[0]) + am(k[1]), am(k[1]), -k[2], k))

    @staticmethod
    def var(k):
        e, f, m = k
        return "v%s_%s_%d" % (tag(e), tag(f), m)

    def emit(self, acc_index):
        lines = []
        flops = 0
        X = "xyz"
        for k in self.order():
            kind, info = self.nodes[k]
            v = self.var(k)
            if kind == "base":
                lines.append("const R %s = F[%d];" % (v, k[2]))
                continue
            i, deps = info
            terms = []
            if kind == "bra":
                terms.append("p.PA[%d] * %s" % (i, self.var(deps[0][2])))
                terms.append("p.WP[%d] * %s" % (i, self.var(deps[1][2])))
                flops += 3
                if len(deps) > 2:
                    n, a, b = deps[2][1], deps[2][2], deps[2][3]
                    terms.append("R(%d) * p.oo2z * (%s - p.rz * %s)" % (n, self.var(a), self.var(b)))
                    flops += 4 if n > 1 else 3
            else:
                terms.append("p.QC[%d] * %s" % (i, self.var(deps[0][2])))
                terms.append("p.WQ[%d] * %s" % (i, self.var(deps[1][2])))
                flops += 3
                for d in deps[2:]:
                    if d[0] == "N2":
                        terms.append("R(%d) * p.oo2n * (%s - p.rn * %s)" % (d[1], self.var(d[2]), self.var(d[3])))
                        flops += 4 if d[1] > 1 else 3
                    else:
                        terms.append("R(%d) * p.oo2zn * %s" % (d[1], self.var(d[2])))
                        flops += 3 if d[1] > 1 else 2
            lines.append("const R %s = %s;" % (v, " + ".join(terms)))
        for (e, f) in self.targets:
            lines.append("acc[%d] += %s;" % (acc_index[(e, f)], self.var((e, f, 0))))
            flops += 1
        return lines, flops


class Hrr:
    def __init__(self, La, Lb, Lc, Ld):
        self.La, self.Lb, self.Lc, self.Ld = La, Lb, Lc, Ld
        self.ket = {}    # (e, c, d) -> expr
        self.bra = {}    # (a, b, c, d) -> expr
        self.order_k = []
        self.order_b = []

    def need_ket(self, e, c, d):
        k = (e, c, d)
        if k in self.ket:
            return
        if d == Z:
            self.ket[k] = None
            return
        i = next(j for j in range(3) if d[j] > 0)
        self.need_ket(e, add(c, i), sub(d, i))
        self.need_ket(e, c, sub(d, i))
        self.ket[k] = i
        self.order_k.append(k)

    def need_bra(self, a, b, c, d):
        k = (a, b, c, d)
        if k in self.bra:
            return
        if b == Z:
            self.need_ket(a, c, d)
            self.bra[k] = None
            return
        i = next(j for j in range(3) if b[j] > 0)
        self.need_bra(add(a, i), sub(b, i), c, d)
        self.need_bra(a, sub(b, i), c, d)
        self.bra[k] = i
        self.order_b.append(k)

    def emit(self, acc_index):
        outs = []
        for a in comps(self.La):
            for b in comps(self.Lb):
                for c in comps(self.Lc):
                    for d in comps(self.Ld):
                        self.need_bra(a, b, c, d)
                        outs.append((a, b, c, d))
        lines = []
        flops = 0

        def kv(e, c, d):
            if d == Z:
                return "acc[%d]" % acc_index[(e, c)]
            return "k%s_%s_%s" % (tag(e), tag(c), tag(d))

        def bv(a, b, c, d):
            if b == Z:
                return kv(a, c, d)
            return "b%s_%s_%s_%s" % (tag(a), tag(b), tag(c), tag(d))

        for (e, c, d) in self.order_k:
            i = self.ket[(e, c, d)]
            lines.append("const R %s = %s + CD[%d] * %s;" % (kv(e, c, d), kv(e, add(c, i), sub(d, i)), i,
                                                            kv(e, c, sub(d, i))))
            flops += 2
        for (a, b, c, d) in self.order_b:
            i = self.bra[(a, b, c, d)]
            lines.append("const R %s = %s + AB[%d] * %s;" % (bv(a, b, c, d), bv(add(a, i), sub(b, i), c, d), i,
                                                            bv(a, sub(b, i), c, d)))
            flops += 2
        for n, (a, b, c, d) in enumerate(outs):
            lines.append("I[%d] = %s;" % (n, bv(a, b, c, d)))
        return lines, flops


# Classes that also get a wave-per-quartet kernel (fock_coop_kernel): the
# VRR and HRR run as level-scheduled node programs over LDS, so a quartet's
# intermediates never leave the chip.
COOP_CLASSES = [(1, 1, 1, 1), (2, 0, 1, 1), (2, 0, 2, 0), (2, 1, 1, 1), (2, 1, 2, 0), (2, 1, 2, 1), (2, 2, 0, 0),
                (2, 2, 1, 0), (2, 2, 1, 1), (2, 2, 2, 0), (2, 2, 2, 1), (2, 2, 2, 2)]


# Classes with La+Lb+Lc+Ld above this get no per-lane kernels: their per-lane
# forms need several KB of scratch per lane (a (dd|dd) quartet holds 961
# contracted accumulators), and provisioning that private segment for a
# full-occupancy launch is what the runtime cannot always do.  Every launch of
# such a class runs on the wave-per-quartet kernel.
LANE_LSUM_MAX = 5


def has_lane_kernels(La, Lb, Lc, Ld):
    return La + Lb + Lc + Ld <= LANE_LSUM_MAX


def coop_tables(La, Lb, Lc, Ld):
    """Node programs of the wave-per-quartet kernel for one class.

    LDS slots (doubles): [0, NACC) contracted [e0|f0] accumulators; from NACC
    a region shared by the VRR (base [00|00]^(m), m = 0..L, the VRR nodes in
    level order, one zero slot) and afterwards the HRR (ket nodes, then bra
    nodes, each in level order); then the coefficient table (per (bra|ket,
    direction): PA|QC, WP|WQ, 1/2z|1/2n, rho/z|rho/n; 1/2(z+n); CD and AB),
    the integrals with cons and degeneracy, and the six density blocks."""
    v = Vrr(La, Lb, Lc, Ld)
    L = v.L
    acc = {t: i for i, t in enumerate(v.targets)}
    nacc = len(v.targets)
    order = [k for k in v.order() if v.nodes[k][0] != "base"]
    nv = len(order)
    slot = {}
    for m in range(L + 1):
        slot[(Z, Z, m)] = nacc + m
    for idx, k in enumerate(order):
        slot[k] = nacc + L + 1 + idx
    zs = nacc + L + 1 + nv
    vnodes, vlev = [], [0]
    cur = None
    for idx, k in enumerate(order):
        lv = am(k[0]) + am(k[1])
        if cur is not None and lv != cur:
            vlev.append(idx)
        cur = lv
        kind, (i, deps) = v.nodes[k]
        a, b = slot[deps[0][2]], slot[deps[1][2]]
        c = d = e = zs
        n1 = n2 = 0
        if kind == "bra":
            sel = i
            if len(deps) > 2:
                n1, c, d = deps[2][1], slot[deps[2][2]], slot[deps[2][3]]
        else:
            sel = 3 + i
            for dd in deps[2:]:
                if dd[0] == "N2":
                    n1, c, d = dd[1], slot[dd[2]], slot[dd[3]]
                else:
                    n2, e = dd[1], slot[dd[2]]
        vnodes.append((a, b, c, d, e, slot[k], sel, n1, n2))
    vlev.append(nv)
    tgt = [slot[(e, f, 0)] for (e, f) in v.targets]
    # HRR node programs
    h = Hrr(La, Lb, Lc, Ld)
    outs = []
    for a in comps(La):
        for b in comps(Lb):
            for c in comps(Lc):
                for d in comps(Ld):
                    h.need_bra(a, b, c, d)
                    outs.append((a, b, c, d))
    korder = sorted(h.order_k, key=lambda k: am(k[2]))
    border = sorted(h.order_b, key=lambda k: am(k[1]))
    nk, nb_ = len(korder), len(border)
    kslot = {k: nacc + i for i, k in enumerate(korder)}
    bslot = {k: nacc + nk + i for i, k in enumerate(border)}

    def ks(e, c, d):
        return acc[(e, c)] if d == Z else kslot[(e, c, d)]

    def bs(a, b, c, d):
        return ks(a, c, d) if b == Z else bslot[(a, b, c, d)]

    hnodes, hlev = [], [0]
    cur = None
    for idx, (e, c, d) in enumerate(korder):
        lv = ("k", am(d))
        if cur is not None and lv != cur:
            hlev.append(idx)
        cur = lv
        i = h.ket[(e, c, d)]
        hnodes.append((kslot[(e, c, d)], ks(e, add(c, i), sub(d, i)), ks(e, c, sub(d, i)), i))
    for idx, (a, b, c, d) in enumerate(border):
        lv = ("b", am(b))
        if cur is not None and lv != cur:
            hlev.append(nk + idx)
        cur = lv
        i = h.bra[(a, b, c, d)]
        hnodes.append((bslot[(a, b, c, d)], bs(add(a, i), sub(b, i), c, d), bs(a, sub(b, i), c, d), 3 + i))
    hlev.append(nk + nb_)
    omap = [bs(*o) for o in outs]
    region = max(L + 2 + nv, nk + nb_)
    na, nb, nc, nd = (len(comps(x)) for x in (La, Lb, Lc, Ld))
    ni = na * nb * nc * nd
    cf = nacc + region
    ic = cf + 32
    dm = ic + ni
    total = dm + na * nb + nc * nd + na * nc + na * nd + nb * nc + nb * nd
    assert total * 8 <= 64 * 1024, (La, Lb, Lc, Ld, total)
    assert all(max(x[:6]) < 65536 for x in vnodes) and all(max(x[:3]) < 65536 for x in hnodes)
    return dict(L=L, nacc=nacc, nv=nv, vlev=vlev, vnodes=vnodes, tgt=tgt, zs=zs, nk=nk, nb=nb_, hnodes=hnodes,
                hlev=hlev, omap=omap, ni=ni, cf=cf, ic=ic, dm=dm, total=total, outs=outs, targets=v.targets)


def coop_source(name, La, Lb, Lc, Ld):
    t = coop_tables(La, Lb, Lc, Ld)
    L, nacc, nv, vlev, vnodes, tgt, zs = t["L"], t["nacc"], t["nv"], t["vlev"], t["vnodes"], t["tgt"], t["zs"]
    nk, nb_, hnodes, hlev, omap, ni = t["nk"], t["nb"], t["hnodes"], t["hlev"], t["omap"], t["ni"]
    cf, ic, dm, total = t["cf"], t["ic"], t["dm"], t["total"]
    cname = name + "Coop"
    src = ["struct %s {" % cname,
           "  static constexpr int La = %d, Lb = %d, Lc = %d, Ld = %d, L = %d;" % (La, Lb, Lc, Ld, L),
           "  static constexpr int NACC = %d, NV = %d, NLV = %d, NH = %d, NLH = %d, NI = %d;"
           % (nacc, nv, len(vlev) - 1, nk + nb_, len(hlev) - 1, ni),
           "  static constexpr int BASE = %d, ZS = %d, CF = %d, IC = %d, DM = %d, LDS = %d;"
           % (nacc, zs, cf, ic, dm, total),
           "  static constexpr int NT = %d;   // threads per workgroup" % (256 if nv >= 1000 else 64),
           "  __device__ static const CoopVNode* vrr();",
           "  __device__ static const int* vlev();",
           "  __device__ static const unsigned short* tgt();",
           "  __device__ static const CoopHNode* hrr();",
           "  __device__ static const int* hlev();",
           "  __device__ static const unsigned short* out();",
           "};"]
    fmt_v = ",".join("{%d,%d,%d,%d,%d,%d,%d,%d,%d,0}" % x for x in vnodes)
    fmt_h = ",".join("{%d,%d,%d,%d,0}" % x for x in hnodes)
    src += ["__device__ const CoopVNode k%s_vrr[%d] = {%s};" % (cname, nv, fmt_v),
            "__device__ const int k%s_vlev[%d] = {%s};" % (cname, len(vlev), ",".join(map(str, vlev))),
            "__device__ const unsigned short k%s_tgt[%d] = {%s};" % (cname, nacc, ",".join(map(str, tgt))),
            "__device__ const CoopHNode k%s_hrr[%d] = {%s};" % (cname, max(1, nk + nb_), fmt_h or "{0,0,0,0,0}"),
            "__device__ const int k%s_hlev[%d] = {%s};" % (cname, len(hlev), ",".join(map(str, hlev))),
            "__device__ const unsigned short k%s_out[%d] = {%s};" % (cname, ni, ",".join(map(str, omap))),
            "__device__ inline const CoopVNode* %s::vrr() { return k%s_vrr; }" % (cname, cname),
            "__device__ inline const int* %s::vlev() { return k%s_vlev; }" % (cname, cname),
            "__device__ inline const unsigned short* %s::tgt() { return k%s_tgt; }" % (cname, cname),
            "__device__ inline const CoopHNode* %s::hrr() { return k%s_hrr; }" % (cname, cname),
            "__device__ inline const int* %s::hlev() { return k%s_hlev; }" % (cname, cname),
            "__device__ inline const unsigned short* %s::out() { return k%s_out; }" % (cname, cname)]
    return cname, "\n".join(src) + "\n", total


PAIR_CLASSES = [(0, 0), (1, 0), (1, 1), (2, 0), (2, 1), (2, 2)]


def quartet_classes():
    out = []
    for kb, (La, Lb) in enumerate(PAIR_CLASSES):
        for kk, (Lc, Ld) in enumerate(PAIR_CLASSES):
            if kk <= kb:
                out.append((La, Lb, Lc, Ld))
    return out


def min_waves(nacc, ni):
    """Occupancy target per class (caps VGPRs: 4 waves -> 128, 3 -> 168, 2 -> 256).
    (ss|ss) fits 128 VGPRs without spilling and gets 4 waves; every other
    class needs more than 128 for the Boys function and the digestion state
    (at 128 its spills cost more than the fourth wave gains), so it is
    capped at 168."""
    size = nacc + ni
    if size <= 2:
        return 4
    if size <= 40:
        return 3
    if size <= 80:
        return 2
    return 1


def boys_setup_flop(L):
    """Boys function (fit polynomial, exp, sqrt, recursion) plus the per
    primitive-quartet setup (1/(z+n), PQ, T, prefactor, PA/WP/QC/WQ):
    a fixed convention, 58 + 4 L."""
    return 58 + 4 * L


def ncart(L):
    return (L + 1) * (L + 2) // 2


def gen_class(La, Lb, Lc, Ld):
    name = "C%d%d%d%d" % (La, Lb, Lc, Ld)
    acc_index = {}
    for E in range(La, La + Lb + 1):
        for e in comps(E):
            for F in range(Lc, Lc + Ld + 1):
                for f in comps(F):
                    acc_index[(e, f)] = len(acc_index)
    vrr = Vrr(La, Lb, Lc, Ld)
    vlines, vflops = vrr.emit(acc_index)
    hrr = Hrr(La, Lb, Lc, Ld)
    hlines, hflops = hrr.emit(acc_index)
    nacc = len(acc_index)
    ni = ncart(La) * ncart(Lb) * ncart(Lc) * ncart(Ld)
    L = La + Lb + Lc + Ld
    src = []
    src.append("// (%s%s|%s%s): %d accumulators, %d integrals, VRR %d FLOP/prim, HRR %d FLOP"
               % ("spd"[La], "spd"[Lb], "spd"[Lc], "spd"[Ld], nacc, ni, vflops, hflops))
    src.append("struct %s {" % name)
    src.append("  static constexpr int La = %d, Lb = %d, Lc = %d, Ld = %d, L = %d;" % (La, Lb, Lc, Ld, L))
    src.append("  static constexpr int NACC = %d, NI = %d;" % (nacc, ni))
    src.append("  static constexpr int VRR_FLOP = %d, HRR_FLOP = %d;" % (vflops, hflops))
    src.append("  static constexpr int MIN_WAVES = %d;   // waves per SIMD the register budget targets"
               % min_waves(nacc, ni))
    src.append("  template <typename R>")
    src.append("  __host__ __device__ static __forceinline__ void vrr(const R (&F)[L + 1], const VrrPar<R>& p, R (&acc)[NACC]) {")
    src += ["    " + l for l in vlines]
    src.append("  }")
    src.append("  template <typename R>")
    src.append("  __host__ __device__ static __forceinline__ void hrr(const R (&acc)[NACC], const R (&AB)[3], const R (&CD)[3], R (&I)[NI]) {")
    src += ["    " + l for l in hlines]
    src.append("  }")
    src.append("};")
    return name, "\n".join(src) + "\n", dict(La=La, Lb=Lb, Lc=Lc, Ld=Ld, nacc=nacc, ni=ni,
                                             vrr_flop=vflops, hrr_flop=hflops, nodes=len(vrr.nodes))


def launcher_source(name, diag, coop=False, lane=True):
    """Launchers of one class.  ``lane``: the class has per-lane kernels
    (direct FP64 and LDS-tile FP64/FP32); without them every launch of the
    class goes to the wave-per-quartet kernel and its Schwarz diagonal to the
    wave-per-quartet Schwarz kernel."""
    out = ["// Generated by quick_amd/codegen/gen_eri.py - do not edit."]
    out += (['#include "%s.h"' % name] if lane else []) + (['#include "coop_%s.h"' % name] if coop else [])
    out += ["", "namespace qf {", ""]
    if lane:
        out += ["// Direct-atomic launch; FP64 only.  The FP32 instantiation of the direct",
                "// kernel is not built (it gave wrong Fock matrices on gfx950, cause not",
                "// found): a request for it is refused with a nonzero return.",
                "int launch_fock_%s(const Task* t, int nt, const PairBlock* b, const MolDesc* m, int fp32,"
                " int grid, hipStream_t s)" % name, "{",
                "    if (fp32) return 1;",
                "    if (nt <= 0) return 0;",
                "    if (grid <= 0 || grid > nt) grid = nt;",
                "    hipLaunchKernelGGL((fock_kernel<%s, double>), dim3(grid), dim3(WAVE), 0, s, t, nt, b, m);" % name,
                "    return 0;",
                "}", "",
                "void launch_fock_tile_%s(const TileGroup* g, int ng, const Task* t, const PairBlock* b,"
                " const MolDesc* m, int fp32, int waves, size_t lds, int grid, hipStream_t s)" % name, "{",
                "    if (ng <= 0) return;",
                "    if (grid <= 0 || grid > ng) grid = ng;",
                "    if (fp32)",
                "        hipLaunchKernelGGL((fock_tile_kernel<%s, float>), dim3(grid), dim3(waves * WAVE), lds, s,"
                " g, ng, t, b, m);" % name,
                "    else",
                "        hipLaunchKernelGGL((fock_tile_kernel<%s, double>), dim3(grid), dim3(waves * WAVE), lds, s,"
                " g, ng, t, b, m);" % name,
                "}", ""]
    if coop:
        out += ["void launch_fock_coop_%s(const Task* t, int nt, const PairBlock* b, const MolDesc* m, int grid,"
                " hipStream_t s)" % name, "{",
                "    if (nt <= 0) return;",
                "    if (grid <= 0 || grid > nt) grid = nt;",
                "    hipLaunchKernelGGL((fock_coop_kernel<%sCoop>), dim3(grid), dim3(%sCoop::NT), %sCoop::LDS * sizeof(double),"
                " s, t, nt, b, m);" % (name, name, name),
                "}", ""]
    if diag:
        out += ["void launch_schwarz_%s(PairBlock blk, double* out, hipStream_t s)" % name, "{",
                "    if (blk.n <= 0) return;"]
        if lane:
            out += ["    hipLaunchKernelGGL((schwarz_kernel<%s>), dim3((blk.n + WAVE - 1) / WAVE), dim3(WAVE), 0, s,"
                    " blk, out);" % name]
        else:
            out += ["    hipLaunchKernelGGL((schwarz_coop_kernel<%sCoop>), dim3(blk.n), dim3(%sCoop::NT),"
                    " %sCoop::LDS * sizeof(double), s, blk, out);" % (name, name, name)]
        out += ["}", ""]
    out += ["}  // namespace qf", ""]
    return "\n".join(out)


def _write(path, text):
    """Write only when the content changes, so incremental builds stay incremental."""
    if os.path.exists(path):
        with open(path) as fh:
            if fh.read() == text:
                return
    with open(path, "w") as fh:
        fh.write(text)


def main():
    os.makedirs(OUT, exist_ok=True)
    table = {}
    names = []
    for (La, Lb, Lc, Ld) in quartet_classes():
        name, src, info = gen_class(La, Lb, Lc, Ld)
        hdr = ("// Generated by quick_amd/codegen/gen_eri.py - do not edit.\n"
               "#pragma once\n#include \"../qf_device.h\"\nnamespace qf {\n" + src + "}  // namespace qf\n")
        _write(os.path.join(OUT, name + ".h"), hdr)
        diag = (La, Lb) == (Lc, Ld)
        coop = (La, Lb, Lc, Ld) in COOP_CLASSES
        if coop:
            cname, csrc, ctotal = coop_source(name, La, Lb, Lc, Ld)
            _write(os.path.join(OUT, "coop_" + name + ".h"),
                   "// Generated by quick_amd/codegen/gen_eri.py - do not edit.\n#pragma once\n"
                   "#include \"../qf_coop.h\"\nnamespace qf {\n" + csrc + "}  // namespace qf\n")
        lane = has_lane_kernels(La, Lb, Lc, Ld)
        assert lane or coop, "a class needs per-lane or wave-per-quartet kernels"
        _write(os.path.join(OUT, "eri_" + name + ".hip"), launcher_source(name, diag, coop, lane))
        names.append((name, La, Lb, Lc, Ld, diag))
        info["boys_setup_flop"] = boys_setup_flop(La + Lb + Lc + Ld)
        info["prim_flop"] = info["vrr_flop"] + info["boys_setup_flop"]
        info["quartet_flop"] = info["hrr_flop"] + 12 * info["ni"]
        table[name] = info
    # launch table
    pc = {pcl: k for k, pcl in enumerate(PAIR_CLASSES)}
    lines = ["// Generated by quick_amd/codegen/gen_eri.py - do not edit.", "#pragma once",
             "#include <hip/hip_runtime.h>", '#include "../qf_types.h"', "", "namespace qf {", "",
             "// returns nonzero when the launch is refused (FP32 on the direct kernel)",
             "using FockLaunch = int (*)(const Task*, int, const PairBlock*, const MolDesc*, int, int, hipStream_t);",
             "using FockTileLaunch = void (*)(const TileGroup*, int, const Task*, const PairBlock*, const MolDesc*,"
             " int, int, size_t, int, hipStream_t);",
             "using FockCoopLaunch = void (*)(const Task*, int, const PairBlock*, const MolDesc*, int, hipStream_t);",
             "using SchwarzLaunch = void (*)(PairBlock, double*, hipStream_t);", ""]
    for (name, La, Lb, Lc, Ld, diag) in names:
        if has_lane_kernels(La, Lb, Lc, Ld):
            lines.append("int launch_fock_%s(const Task*, int, const PairBlock*, const MolDesc*, int, int, hipStream_t);"
                         % name)
            lines.append("void launch_fock_tile_%s(const TileGroup*, int, const Task*, const PairBlock*, const MolDesc*,"
                         " int, int, size_t, int, hipStream_t);" % name)
        if (La, Lb, Lc, Ld) in COOP_CLASSES:
            lines.append("void launch_fock_coop_%s(const Task*, int, const PairBlock*, const MolDesc*, int, hipStream_t);"
                         % name)
        if diag:
            lines.append("void launch_schwarz_%s(PairBlock, double*, hipStream_t);" % name)
    # FLOP model (FMA = 2): per surviving primitive quartet the VRR plus
    # BOYS_SETUP_FLOP(L) for the Boys function and the quartet setup, per
    # surviving contracted quartet the HRR plus 12 per Cartesian integral
    # (six Fock updates).
    lines += ["", "// Algorithmic FLOPs of nq quartets with np primitive quartets in class (cb|ck).",
              "inline double class_flops(int cb, int ck, unsigned long long nq, unsigned long long np)", "{"]
    for (name, La, Lb, Lc, Ld, diag) in names:
        info = table[name]
        L = La + Lb + Lc + Ld
        lines.append("    if (cb == %d && ck == %d) return double(np) * %d.0 + double(nq) * %d.0;"
                     % (pc[(La, Lb)], pc[(Lc, Ld)], info["vrr_flop"] + boys_setup_flop(L),
                        info["hrr_flop"] + 12 * info["ni"]))
    lines += ["    return 0.0;", "}", ""]
    lines += ["", "// indexed [bra pair class][ket pair class], pair class = La*(La+1)/2 + Lb;",
              "// nullptr: the class has no per-lane kernels (wave-per-quartet only)",
              "inline FockLaunch fock_launcher(int cb, int ck)", "{"]
    for (name, La, Lb, Lc, Ld, diag) in names:
        if has_lane_kernels(La, Lb, Lc, Ld):
            lines.append("    if (cb == %d && ck == %d) return launch_fock_%s;" % (pc[(La, Lb)], pc[(Lc, Ld)], name))
    lines += ["    return nullptr;", "}", "", "inline FockTileLaunch fock_tile_launcher(int cb, int ck)", "{"]
    for (name, La, Lb, Lc, Ld, diag) in names:
        if has_lane_kernels(La, Lb, Lc, Ld):
            lines.append("    if (cb == %d && ck == %d) return launch_fock_tile_%s;" % (pc[(La, Lb)], pc[(Lc, Ld)], name))
    lines += ["    return nullptr;", "}", "", "// wave-per-quartet kernels (nullptr: the class has none)",
              "inline FockCoopLaunch fock_coop_launcher(int cb, int ck)", "{"]
    for (name, La, Lb, Lc, Ld, diag) in names:
        if (La, Lb, Lc, Ld) in COOP_CLASSES:
            lines.append("    if (cb == %d && ck == %d) return launch_fock_coop_%s;" % (pc[(La, Lb)], pc[(Lc, Ld)], name))
    lines += ["    return nullptr;", "}", "", "inline SchwarzLaunch schwarz_launcher(int c)", "{"]
    for (name, La, Lb, Lc, Ld, diag) in names:
        if diag:
            lines.append("    if (c == %d) return launch_schwarz_%s;" % (pc[(La, Lb)], name))
    lines += ["    return nullptr;", "}", "", "}  // namespace qf", ""]
    _write(os.path.join(OUT, "launchers.h"), "\n".join(lines))
    _write(os.path.join(OUT, "flop_table.json"), json.dumps(table, indent=1))
    for k, v in table.items():
        print(k, v)


if __name__ == "__main__":
    sys.exit(main())
