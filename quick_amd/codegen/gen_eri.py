"""Generate the per-class ERI recursions for the HIP kernels.

For every canonical class (La Lb | Lc Ld) with La >= Lb, Lc >= Ld, and bra
pair class >= ket pair class (55 classes for s, p, d, f) this emits, for the
classes that have per-lane kernels, a struct with two straight-line device
functions:

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

# QUICK's Cartesian component order per shell (basis.f90:567-720 KLMN):
# by ascending z power, then descending x power.
QCOMP = {
    0: [(0, 0, 0)],
    1: [(1, 0, 0), (0, 1, 0), (0, 0, 1)],
    2: [(2, 0, 0), (1, 1, 0), (0, 2, 0), (1, 0, 1), (0, 1, 1), (0, 0, 2)],
    3: [(3, 0, 0), (2, 1, 0), (1, 2, 0), (0, 3, 0), (2, 0, 1), (1, 1, 1), (0, 2, 1), (1, 0, 2), (0, 1, 2),
        (0, 0, 3)],
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


PAIR_CLASSES = [(0, 0), (1, 0), (1, 1), (2, 0), (2, 1), (2, 2), (3, 0), (3, 1), (3, 2), (3, 3)]


def quartet_classes():
    out = []
    for kb, (La, Lb) in enumerate(PAIR_CLASSES):
        for kk, (Lc, Ld) in enumerate(PAIR_CLASSES):
            if kk <= kb:
                out.append((La, Lb, Lc, Ld))
    return out


# Classes that get a wave-per-quartet kernel (fock_coop_kernel): the VRR and
# HRR run as level-scheduled node programs over LDS, so a quartet's
# intermediates never leave the chip.  Every class with an f shell has only
# this kernel.
COOP_CLASSES = [(1, 1, 1, 1), (2, 0, 1, 1), (2, 0, 2, 0), (2, 1, 1, 1), (2, 1, 2, 0), (2, 1, 2, 1), (2, 2, 0, 0),
                (2, 2, 1, 0), (2, 2, 1, 1), (2, 2, 2, 0), (2, 2, 2, 1), (2, 2, 2, 2)]
COOP_CLASSES += [c for c in quartet_classes() if max(c) == 3]


# Classes with La+Lb+Lc+Ld above this get no per-lane kernels: their per-lane
# forms need several KB of scratch per lane (a (dd|dd) quartet holds 961
# contracted accumulators), and provisioning that private segment for a
# full-occupancy launch is what the runtime cannot always do.  Every launch of
# such a class runs on the wave-per-quartet kernel.
LANE_LSUM_MAX = 5


def has_lane_kernels(La, Lb, Lc, Ld):
    return La + Lb + Lc + Ld <= LANE_LSUM_MAX and max(La, Lb, Lc, Ld) <= 2


# Per-workgroup LDS the wave-per-quartet kernel may use; classes whose node
# programs need more keep their slots in a global scratch buffer instead.
COOP_LDS_CAP = 64 * 1024
COOP_NT = 64          # one wave per quartet
COOP_GLOBAL_SLICES = 1024   # workgroups of a launch with global slots (qf_types.h)


class SlotPool:
    """Lowest-free-slot allocator for node-program values."""

    def __init__(self, first):
        self.next = first
        self.free = []

    def take(self):
        import heapq
        if self.free:
            return heapq.heappop(self.free)
        self.next += 1
        return self.next - 1

    def give(self, slot):
        import heapq
        heapq.heappush(self.free, slot)


def _allocate(values, levels, consumers_last, pool, pinned=()):
    """Assigns slots level by level: before the values of level l are
    placed, every value whose last reader ran at a level < l gives its slot
    back.  A slot is therefore never written in the level that reads its old
    value.  Values without readers go to no slot (None)."""
    slot = dict(pinned)
    by_level = {}
    for v in values:
        by_level.setdefault(levels[v], []).append(v)
    live = [(consumers_last[v], v) for v in slot if v in consumers_last]
    for lv in sorted(by_level):
        still = []
        for last, v in live:
            if last < lv:
                pool.give(slot[v])
            else:
                still.append((last, v))
        live = still
        for v in by_level[lv]:
            if v in slot:
                continue
            if consumers_last.get(v, -1) < 0:
                slot[v] = None
                continue
            slot[v] = pool.take()
            live.append((consumers_last[v], v))
    return slot


def coop_tables(La, Lb, Lc, Ld):
    """Node programs of the wave-per-quartet kernel for one class.

    Slots (doubles): [0, NACC) contracted [e0|f0] accumulators; ZS a zero,
    SINK a write-only slot; [OUT, OUT + NI) the integrals (ab|cd) in output
    order; from POOL the VRR values of one primitive quartet ([00|00]^(m)
    and the recursion nodes) and afterwards the HRR intermediates, both
    placed by liveness over the levels; then the coefficient table CF
    (per (bra|ket, direction): PA|QC, WP|WQ, 1/2z|1/2n, rho/z|rho/n;
    1/2(z+n); CD and AB) and the six density blocks DM.

    A VRR node of level l (= |e| + |f|) reads values of levels l-1 and l-2
    only and, when its (e, f, m = 0) is a target, also adds itself into that
    accumulator.  An HRR node of level l reads level l-1 (ket levels |d|,
    then bra levels Ld + |b|); the last level writes the integrals."""
    v = Vrr(La, Lb, Lc, Ld)
    L = v.L
    acc = {t: i for i, t in enumerate(v.targets)}
    nacc = len(v.targets)
    ZS, SINK = nacc, nacc + 1
    na, nb, nc, nd = (len(comps(x)) for x in (La, Lb, Lc, Ld))
    ni = na * nb * nc * nd
    OUT = nacc + 2
    POOL = OUT + ni
    # ---- VRR ---------------------------------------------------------------
    order = [k for k in v.order() if v.nodes[k][0] != "base"]
    vlevel = {k: am(k[0]) + am(k[1]) for k in v.nodes}
    last = {}
    for k in order:
        kind, (i, deps) = v.nodes[k]
        for d in deps:
            for x in d[2:]:
                last[x] = max(last.get(x, -1), vlevel[k])
    base = [(Z, Z, m) for m in range(L + 1)]
    for b in base:                      # the setup writes every F_m(T): give each a slot
        last.setdefault(b, 0)
    pool = SlotPool(POOL)
    vslot = _allocate(base + order, vlevel, last, pool)
    vmax = pool.next
    vnodes, vlev = [], [0]
    cur = None
    sl = lambda k: vslot[k] if vslot[k] is not None else ZS
    for idx, k in enumerate(order):
        lv = vlevel[k]
        if cur is not None and lv != cur:
            vlev.append(idx)
        cur = lv
        kind, (i, deps) = v.nodes[k]
        a, b = sl(deps[0][2]), sl(deps[1][2])
        c = d = e = ZS
        n1 = n2 = 0
        if kind == "bra":
            sel = i
            if len(deps) > 2:
                n1, c, d = deps[2][1], sl(deps[2][2]), sl(deps[2][3])
        else:
            sel = 3 + i
            for dd in deps[2:]:
                if dd[0] == "N2":
                    n1, c, d = dd[1], sl(dd[2]), sl(dd[3])
                else:
                    n2, e = dd[1], sl(dd[2])
        dst = vslot[k] if vslot[k] is not None else SINK
        ac = acc[(k[0], k[1])] if (k[2] == 0 and (k[0], k[1]) in acc) else 0xFFFF
        assert n1 < 16 and n2 < 16
        vnodes.append((a, b, c, d, e, dst, ac, sel | (n1 << 4) | (n2 << 8)))
    vlev.append(len(order))
    fslot = [vslot[b] for b in base]
    # every target must be produced by exactly one node
    assert sorted(x[6] for x in vnodes if x[6] != 0xFFFF) == list(range(nacc))
    # ---- HRR ---------------------------------------------------------------
    h = Hrr(La, Lb, Lc, Ld)
    outs = []
    for a in comps(La):
        for b in comps(Lb):
            for c in comps(Lc):
                for d in comps(Ld):
                    h.need_bra(a, b, c, d)
                    outs.append((a, b, c, d))
    oidx = {o: x for x, o in enumerate(outs)}

    def kkey(e, c, d):
        return ("acc", e, c) if d == Z else ("k", e, c, d)

    def bkey(a, b, c, d):
        return kkey(a, c, d) if b == Z else ("b", a, b, c, d)

    hn = {}        # key -> (level, sel, src1 key, src2 key)
    for (e, c, d) in h.order_k:
        i = h.ket[(e, c, d)]
        hn[("k", e, c, d)] = (am(d), i, kkey(e, add(c, i), sub(d, i)), kkey(e, c, sub(d, i)))
    for (a, b, c, d) in h.order_b:
        i = h.bra[(a, b, c, d)]
        hn[("b", a, b, c, d)] = (Ld + am(b), 3 + i, bkey(add(a, i), sub(b, i), c, d), bkey(a, sub(b, i), c, d))
    out_keys = {bkey(*o): oidx[o] for o in outs}
    copies = [k for k in out_keys if k[0] == "acc"]       # Lb = Ld = 0: the accumulators are the integrals
    for k in copies:
        hn[("copy",) + k[1:]] = (1, 0, k, "ZERO")
    hlevel = {k: lv for k, (lv, _, _, _) in hn.items()}
    hlast = {}
    for k, (lv, _, s1, s2) in hn.items():
        for x in (s1, s2):
            hlast[x] = max(hlast.get(x, -1), lv)
    accs = [("acc", e, f) for (e, f) in v.targets]
    for a_ in accs:
        hlevel[a_] = 0
    pinned = {("acc", e, f): acc[(e, f)] for (e, f) in v.targets}
    pinned["ZERO"] = ZS
    hlevel["ZERO"] = 0
    for k in hn:
        if k in out_keys:
            pinned[k] = OUT + out_keys[k]
    for k in copies:
        pinned[("copy",) + k[1:]] = OUT + out_keys[k]
    hpool = SlotPool(POOL)
    inner = [k for k in hn if k not in pinned]
    for k in inner:
        hlast.setdefault(k, -1)
    # accumulator slots go back to the pool once their last reader has run
    hslot = _allocate(accs + ["ZERO"] + sorted(inner, key=lambda k: hlevel[k]) + [k for k in hn if k in pinned],
                      hlevel, hlast, hpool, pinned=pinned)
    hmax = hpool.next
    hkeys = sorted(hn, key=lambda k: (hlevel[k], str(k)))
    hnodes, hlev = [], [0]
    cur = None
    for idx, k in enumerate(hkeys):
        lv = hlevel[k]
        if cur is not None and lv != cur:
            hlev.append(idx)
        cur = lv
        _, sel, s1, s2 = hn[k]
        assert hslot[k] is not None
        hnodes.append((hslot[k], hslot[s1], hslot[s2], sel))
    hlev.append(len(hkeys))
    foot = max(vmax, hmax, POOL)
    cf = foot
    dm = cf + 32
    total = dm + na * nb + nc * nd + na * nc + na * nd + nb * nc + nb * nd
    assert total < 65536
    return dict(L=L, nacc=nacc, zs=ZS, sink=SINK, out=OUT, pool=POOL, nv=len(order), vlev=vlev, vnodes=vnodes,
                fslot=fslot, nh=len(hnodes), hnodes=hnodes, hlev=hlev, ni=ni, cf=cf, dm=dm, total=total, outs=outs,
                targets=v.targets, glob=total * 8 > COOP_LDS_CAP)


def coop_source(name, La, Lb, Lc, Ld):
    t = coop_tables(La, Lb, Lc, Ld)
    cname = name + "Coop"
    src = ["struct %s {" % cname,
           "  static constexpr int La = %d, Lb = %d, Lc = %d, Ld = %d, L = %d;" % (La, Lb, Lc, Ld, t["L"]),
           "  static constexpr int NACC = %d, NV = %d, NLV = %d, NH = %d, NLH = %d, NI = %d;"
           % (t["nacc"], t["nv"], len(t["vlev"]) - 1, t["nh"], len(t["hlev"]) - 1, t["ni"]),
           "  static constexpr int ZS = %d, SINK = %d, OUT = %d, CF = %d, DM = %d, LDS = %d;"
           % (t["zs"], t["sink"], t["out"], t["cf"], t["dm"], t["total"]),
           "  static constexpr bool GLOBAL_SLOTS = %s;   // slots in a global scratch buffer (over %d KiB)"
           % ("true" if t["glob"] else "false", COOP_LDS_CAP // 1024),
           "  static constexpr int NT = %d;   // threads per workgroup: one wave" % COOP_NT,
           "  static constexpr unsigned short FSLOT[%d] = {%s};   // slots of [00|00]^(m)"
           % (t["L"] + 1, ",".join(map(str, t["fslot"]))),
           "  __device__ static const CoopVNode* vrr();",
           "  __device__ static const int* vlev();",
           "  __device__ static const CoopHNode* hrr();",
           "  __device__ static const int* hlev();",
           "};"]
    fmt_v = ",".join("{%d,%d,%d,%d,%d,%d,%d,%d}" % x for x in t["vnodes"])
    fmt_h = ",".join("{%d,%d,%d,%d}" % x for x in t["hnodes"])
    src += ["__device__ const CoopVNode k%s_vrr[%d] = {%s};" % (cname, t["nv"], fmt_v),
            "__device__ const int k%s_vlev[%d] = {%s};" % (cname, len(t["vlev"]), ",".join(map(str, t["vlev"]))),
            "__device__ const CoopHNode k%s_hrr[%d] = {%s};" % (cname, max(1, t["nh"]), fmt_h or "{0,0,0,0}"),
            "__device__ const int k%s_hlev[%d] = {%s};" % (cname, len(t["hlev"]), ",".join(map(str, t["hlev"]))),
            "__device__ inline const CoopVNode* %s::vrr() { return k%s_vrr; }" % (cname, cname),
            "__device__ inline const int* %s::vlev() { return k%s_vlev; }" % (cname, cname),
            "__device__ inline const CoopHNode* %s::hrr() { return k%s_hrr; }" % (cname, cname),
            "__device__ inline const int* %s::hlev() { return k%s_hlev; }" % (cname, cname)]
    return cname, "\n".join(src) + "\n", t["total"]


# per-class occupancy targets that differ from the size rule, measured on
# (H2O)32 6-31G(d) (FP64 ms per build before -> after): (ds|ps) 2 waves
# 0.87 -> 0.62 (no spills), (ds|ds) 1 wave 0.45 -> 0.23 (no spills),
# (pp|ss) 4 waves 0.35 -> 0.32, (ss|ss) 5 waves 1.31 -> 1.23.  Tried and
# slower: (ps|ps), (dp|ss), (dp|ds) at 2; (pp|ps) at 3; (ds|pp) at 2;
# (ss|ss) at 6, (ps|ss) at 5, (ds|ss) at 4 (they spill).
MIN_WAVES_OVERRIDE = {(2, 0, 1, 0): 2, (2, 0, 2, 0): 1, (1, 1, 0, 0): 4, (0, 0, 0, 0): 5, (1, 0, 0, 0): 4}
# the open-shell kernels carry a second set of exchange accumulators and
# density blocks: one wave fewer than the closed-shell kernel by default
MIN_WAVES_OS_OVERRIDE = {(1, 0, 0, 0): 3}

# Classes whose FP64 tile kernels take the first exchange lean (LEAN_K: the
# a-row density blocks re-read and the a-row exchange terms added to the LDS
# tile every quartet instead of held in registers over the task).  Chosen
# from per-class times of both forms with the persisted launch shapes, ms
# per build (H2O)32 6-31G(d) / (Gly)10 6-31G(d,p): (ps|ps) 0.574 -> 0.538 /
# 1.43 -> 1.36, (pp|ss) 0.321 -> 0.307 / 0.773 -> 0.753, (pp|pp) 0.206 ->
# 0.198 / 0.83 -> 0.81, (ds|ss) 0.550 -> 0.547 / 0.697 -> 0.648, (dp|ss)
# 0.321 -> 0.281 / 0.655 -> 0.630, (dp|ds) 0.647 -> 0.557 / 2.00 -> 1.77,
# (dd|ps) 0.243 -> 0.210 / 0.70 -> 0.57.  Slower in that form: (ss|ss) 1.10
# -> 1.14, (ps|ss) 1.29 -> 1.33, (ds|ps) 0.54 -> 0.58, (dp|pp) 0.73 -> 0.90.
LEAN_K_CLASSES = {(1, 0, 1, 0), (1, 1, 0, 0), (1, 1, 1, 1), (2, 0, 0, 0), (2, 1, 0, 0), (2, 1, 2, 0), (2, 2, 1, 0)}


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
               % ("spdf"[La], "spdf"[Lb], "spdf"[Lc], "spdf"[Ld], nacc, ni, vflops, hflops))
    src.append("struct %s {" % name)
    src.append("  static constexpr int La = %d, Lb = %d, Lc = %d, Ld = %d, L = %d;" % (La, Lb, Lc, Ld, L))
    src.append("  static constexpr int NACC = %d, NI = %d;" % (nacc, ni))
    src.append("  static constexpr int VRR_FLOP = %d, HRR_FLOP = %d;" % (vflops, hflops))
    mw = MIN_WAVES_OVERRIDE.get((La, Lb, Lc, Ld), min_waves(nacc, ni))
    src.append("  static constexpr int MIN_WAVES = %d;   // waves per SIMD the register budget targets" % mw)
    src.append("  static constexpr int MIN_WAVES_OS = %d;   // the same for the open-shell kernels"
               % MIN_WAVES_OS_OVERRIDE.get((La, Lb, Lc, Ld), max(1, mw - 1)))
    src.append("  static constexpr bool LEAN_K = %s;   // FP64 tile kernels: first exchange re-read per quartet"
               % ("true" if (La, Lb, Lc, Ld) in LEAN_K_CLASSES else "false"))
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
                "// os: 1 = the fused open-shell pass (second exchange density)",
                "int launch_fock_%s(const Task* t, int nt, const PairBlock* b, const MolDesc* m, int fp32,"
                " int os, int grid, hipStream_t s)" % name, "{",
                "    if (fp32) return 1;",
                "    if (nt <= 0) return 0;",
                "    if (grid <= 0 || grid > nt) grid = nt;",
                "    if (os)",
                "        hipLaunchKernelGGL((fock_kernel<%s, double, true>), dim3(grid), dim3(WAVE), 0, s, t, nt, b, m);"
                % name,
                "    else",
                "        hipLaunchKernelGGL((fock_kernel<%s, double>), dim3(grid), dim3(WAVE), 0, s, t, nt, b, m);" % name,
                "    return 0;",
                "}", "",
                "// os & 3: 0 closed shell, 1 open shell with a second LDS tile, 2 open",
                "// shell with the second exchange terms straight to G2; os & 4: the lean",
                "// open-shell form (FP64; FP32 launches take the register form)",
                "void launch_fock_tile_%s(const TileGroup* g, int ng, const Task* t, const PairBlock* b,"
                " const MolDesc* m, int fp32, int os, int waves, size_t lds, int grid, hipStream_t s)" % name, "{",
                "    if (ng <= 0) return;",
                "    if (grid <= 0 || grid > ng) grid = ng;",
                "    const bool lean = (os & 4) != 0;",
                "    os &= 3;",
                "    const int k2 = os == 1;",
                "    if (fp32 && os)",
                "        hipLaunchKernelGGL((fock_tile_kernel<%s, float, true>), dim3(grid), dim3(waves * WAVE), lds, s,"
                " g, ng, t, b, m, k2);" % name,
                "    else if (fp32)",
                "        hipLaunchKernelGGL((fock_tile_kernel<%s, float>), dim3(grid), dim3(waves * WAVE), lds, s,"
                " g, ng, t, b, m, 0);" % name,
                "    else if (os && lean)",
                "        hipLaunchKernelGGL((fock_tile_kernel<%s, double, true, true>), dim3(grid), dim3(waves * WAVE), lds,"
                " s, g, ng, t, b, m, k2);" % name,
                "    else if (os)",
                "        hipLaunchKernelGGL((fock_tile_kernel<%s, double, true>), dim3(grid), dim3(waves * WAVE), lds, s,"
                " g, ng, t, b, m, k2);" % name,
                "    else",
                "        hipLaunchKernelGGL((fock_tile_kernel<%s, double>), dim3(grid), dim3(waves * WAVE), lds, s,"
                " g, ng, t, b, m, 0);" % name,
                "}", ""]
    if coop:
        out += ["// gscratch: COOP_GLOBAL_SLICES slices of LDS doubles when the class keeps its",
                "// slots in global memory (the grid is then capped to that many workgroups)",
                "void launch_fock_coop_%s(const Task* t, int nt, const PairBlock* b, const MolDesc* m, int os,"
                " int grid, double* gscratch, hipStream_t s)" % name, "{",
                "    using D = %sCoop;" % name,
                "    if (nt <= 0) return;",
                "    if (grid <= 0 || grid > nt) grid = nt;",
                "    if (D::GLOBAL_SLOTS && grid > COOP_GLOBAL_SLICES) grid = COOP_GLOBAL_SLICES;",
                "    const size_t lds = D::GLOBAL_SLOTS ? 0 : D::LDS * sizeof(double);",
                "    if (os)",
                "        hipLaunchKernelGGL((fock_coop_kernel<D, true>), dim3(grid), dim3(D::NT), lds, s, t, nt, b, m,"
                " gscratch);",
                "    else",
                "        hipLaunchKernelGGL((fock_coop_kernel<D>), dim3(grid), dim3(D::NT), lds, s, t, nt, b, m, gscratch);",
                "}", ""]
    if diag:
        out += ["void launch_schwarz_%s(PairBlock blk, double* out, double* gscratch, hipStream_t s)" % name, "{",
                "    if (blk.n <= 0) return;"]
        if lane:
            out += ["    hipLaunchKernelGGL((schwarz_kernel<%s>), dim3((blk.n + WAVE - 1) / WAVE), dim3(WAVE), 0, s,"
                    " blk, out);" % name]
        else:
            out += ["    using D = %sCoop;" % name,
                    "    const int grid = D::GLOBAL_SLOTS && blk.n > COOP_GLOBAL_SLICES ? COOP_GLOBAL_SLICES : blk.n;",
                    "    hipLaunchKernelGGL((schwarz_coop_kernel<D>), dim3(grid), dim3(D::NT),"
                    " D::GLOBAL_SLOTS ? 0 : D::LDS * sizeof(double), s, blk, out, gscratch);"]
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
        if has_lane_kernels(La, Lb, Lc, Ld):       # the per-lane recursions of the other classes are not built
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
             "// (tasks, ntasks, blocks, mols, fp32, os, grid, stream)",
             "using FockLaunch = int (*)(const Task*, int, const PairBlock*, const MolDesc*, int, int, int, hipStream_t);",
             "// (groups, ngroups, tasks, blocks, mols, fp32, os, waves, lds bytes, grid, stream)",
             "using FockTileLaunch = void (*)(const TileGroup*, int, const Task*, const PairBlock*, const MolDesc*,"
             " int, int, int, size_t, int, hipStream_t);",
             "// (tasks, ntasks, blocks, mols, os, grid, global slots, stream)",
             "using FockCoopLaunch = void (*)(const Task*, int, const PairBlock*, const MolDesc*, int, int, double*,"
             " hipStream_t);",
             "using SchwarzLaunch = void (*)(PairBlock, double*, double*, hipStream_t);", ""]
    for (name, La, Lb, Lc, Ld, diag) in names:
        if has_lane_kernels(La, Lb, Lc, Ld):
            lines.append("int launch_fock_%s(const Task*, int, const PairBlock*, const MolDesc*, int, int, int,"
                         " hipStream_t);" % name)
            lines.append("void launch_fock_tile_%s(const TileGroup*, int, const Task*, const PairBlock*, const MolDesc*,"
                         " int, int, int, size_t, int, hipStream_t);" % name)
        if (La, Lb, Lc, Ld) in COOP_CLASSES:
            lines.append("void launch_fock_coop_%s(const Task*, int, const PairBlock*, const MolDesc*, int, int, double*,"
                         " hipStream_t);" % name)
        if diag:
            lines.append("void launch_schwarz_%s(PairBlock, double*, double*, hipStream_t);" % name)
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
    lines += ["    return nullptr;", "}", "",
              "// doubles per workgroup slice of the global slot buffer (0: the class uses LDS)",
              "inline int coop_global_doubles(int cb, int ck)", "{"]
    for (name, La, Lb, Lc, Ld, diag) in names:
        if (La, Lb, Lc, Ld) in COOP_CLASSES:
            t = coop_tables(La, Lb, Lc, Ld)
            if t["glob"]:
                lines.append("    if (cb == %d && ck == %d) return %d;" % (pc[(La, Lb)], pc[(Lc, Ld)], t["total"]))
    lines += ["    return 0;", "}", "", "inline SchwarzLaunch schwarz_launcher(int c)", "{"]
    for (name, La, Lb, Lc, Ld, diag) in names:
        if diag:
            lines.append("    if (c == %d) return launch_schwarz_%s;" % (pc[(La, Lb)], name))
    lines += ["    return nullptr;", "}", "", "}  // namespace qf", ""]
    _write(os.path.join(OUT, "launchers.h"), "\n".join(lines))
    from . import gen_boys, gen_grad
    _write(os.path.join(OUT, "boys_grid.h"), gen_boys.header())
    table["gradient"] = gen_grad.main(OUT, lambda name, text: _write(os.path.join(OUT, name), text))
    _write(os.path.join(OUT, "flop_table.json"), json.dumps(table, indent=1))
    for k, v in table.items():
        print(k, v)


if __name__ == "__main__":
    sys.exit(main())
