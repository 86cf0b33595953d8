"""Node programs of the two-electron gradient kernel (QUICK's getGrad_kernel_*,
gpu_get2e_subs_grad.h; CPU cshell_eri_grad, eri_grad.fh:43-768).

For a contracted quartet (ab|cd) the derivative with respect to the centre
A of shell a is, per primitive,

    d/dA_x (ab|cd) = 2 alpha (a+1_x b|cd) - n_x(a) (a-1_x b|cd),

and likewise for B and C; D follows from translational invariance,
d/dD = -(d/dA + d/dB + d/dC).  Summed over primitives this needs four
contracted sets of [e0|f0] built by one VRR per primitive quartet:

    S0  weight 1      e in [La-1, La+Lb], f in [Lc-1, Lc+Ld] (as needed by the -1 terms)
    SA  weight 2 alpha  e in [La+1, La+Lb+1], f in [Lc, Lc+Ld]
    SB  weight 2 beta   e in [La, La+Lb+1],   f in [Lc, Lc+Ld]
    SC  weight 2 gamma  e in [La, La+Lb],     f in [Lc+1, Lc+Ld+1]

(QUICK's store, storeAA, storeBB and storeCC), and six HRR blocks:
(a+1 b|cd) from SA, (a-1 b|cd) from S0, (a b+1|cd) from SB, (a b-1|cd)
from S0, (ab|c+1 d) from SC, (ab|c-1 d) from S0.

The program runs on the wave-per-quartet kernel (qf_grad.h): the VRR level
by level with one node per lane, then one accumulation pass (one entry per
lane: acc[dst] += w[set] * value), then the six HRR blocks level by level,
all in LDS slots placed by liveness (gen_eri._allocate)."""
from __future__ import annotations

from .gen_eri import (COOP_LDS_CAP, COOP_NT, PAIR_CLASSES, Hrr, SlotPool, Vrr, Z, _allocate, add, am,
                      boys_setup_flop, comps, ncart, sub)

# derivative blocks: (name, accumulator set, (dLa, dLb, dLc))
BLOCKS = [("AP", 1, (1, 0, 0)), ("AM", 0, (-1, 0, 0)), ("BP", 2, (0, 1, 0)), ("BM", 0, (0, -1, 0)),
          ("CP", 3, (0, 0, 1)), ("CM", 0, (0, 0, -1))]

GRAD_LMAX = 2   # shells up to d (QUICK's getGrad_kernel_sp / _spd)


def grad_classes():
    out = []
    pcs = [pc for pc in PAIR_CLASSES if max(pc) <= GRAD_LMAX]
    for kb, (La, Lb) in enumerate(pcs):
        for kk, (Lc, Ld) in enumerate(pcs):
            if kk <= kb:
                out.append((La, Lb, Lc, Ld))
    return out


class VrrTargets(Vrr):
    """The VRR dependency graph for an explicit target list."""

    def __init__(self, targets, L):
        self.L = L
        self.nodes = {}
        self.targets = list(targets)
        for e, f in self.targets:
            self.need(e, f, 0)


def _rect(e0, e1, f0, f1):
    return [(e, f) for E in range(e0, e1 + 1) for e in comps(E) for F in range(f0, f1 + 1) for f in comps(F)]


def grad_sets(La, Lb, Lc, Ld):
    s0 = []
    if La > 0:
        s0 += _rect(La - 1, La + Lb - 1, Lc, Lc + Ld)
    if Lb > 0:
        s0 += _rect(La, La + Lb - 1, Lc, Lc + Ld)
    if Lc > 0:
        s0 += _rect(La, La + Lb, Lc - 1, Lc + Ld - 1)
    seen, u = set(), []
    for t in s0:
        if t not in seen:
            seen.add(t)
            u.append(t)
    return [u, _rect(La + 1, La + Lb + 1, Lc, Lc + Ld), _rect(La, La + Lb + 1, Lc, Lc + Ld),
            _rect(La, La + Lb, Lc + 1, Lc + Ld + 1)]


def grad_tables(La, Lb, Lc, Ld):
    """Slots (doubles): [0, NACC) the four accumulator sets; ZS zero; SINK;
    then the six output blocks (OUT[k] = -1 when absent); the pool of VRR
    and HRR values; CF (40: VRR coefficients as the Fock node programs at
    [0, 25), CD and AB at [25, 31), weights 1, 2a, 2b, 2c at [32, 36)); DM
    the six density blocks."""
    sets = grad_sets(La, Lb, Lc, Ld)
    L = La + Lb + Lc + Ld + 1
    acc = {}
    for si, st in enumerate(sets):
        for t in st:
            acc[(si, t)] = len(acc)
    nacc = len(acc)
    targets = []
    seen = set()
    for st in sets:
        for t in st:
            if t not in seen:
                seen.add(t)
                targets.append(t)
    v = VrrTargets(targets, L)
    ZS, SINK = nacc, nacc + 1
    na, nb, nc, nd = (ncart(x) for x in (La, Lb, Lc, Ld))
    ni = na * nb * nc * nd
    blocks = []
    out = {}
    o = nacc + 2
    for name, si, (da, db, dc) in BLOCKS:
        cls = (La + da, Lb + db, Lc + dc, Ld)
        if min(cls) < 0:
            out[name] = -1
            continue
        n = ncart(cls[0]) * ncart(cls[1]) * ncart(cls[2]) * ncart(cls[3])
        out[name] = o
        blocks.append((name, si, cls, o))
        o += n
    POOL = o
    # ---- VRR: every target stays live until the accumulation pass ----------
    order = [k for k in v.order() if v.nodes[k][0] != "base"]
    vlevel = {k: am(k[0]) + am(k[1]) for k in v.nodes}
    top = max(vlevel.values()) + 1
    last = {}
    for k in order:
        kind, (i, deps) = v.nodes[k]
        for d in deps:
            for x in d[2:]:
                last[x] = max(last.get(x, -1), vlevel[k])
    base = [(Z, Z, m) for m in range(L + 1)]
    for b in base:
        last.setdefault(b, 0)
    for (e, f) in targets:
        last[(e, f, 0)] = top
    pool = SlotPool(POOL)
    vslot = _allocate(base + order, vlevel, last, pool)
    vmax = pool.next
    sl = lambda k: vslot[k] if vslot[k] is not None else ZS
    vnodes, vlev = [], [0]
    cur = None
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
        assert n1 < 16 and n2 < 16
        vnodes.append((a, b, c, d, e, dst, 0xFFFF, sel | (n1 << 4) | (n2 << 8)))
    vlev.append(len(order))
    fslot = [vslot[b] for b in base]
    accs = []
    for (si, t), x in sorted(acc.items(), key=lambda kv: kv[1]):
        src = vslot[(t[0], t[1], 0)]
        assert src is not None
        accs.append((src, x, si))
    # ---- HRR: the six blocks, merged level by level ------------------------
    hn = {}
    out_pin = {}
    for name, si, (A_, B_, C_, D_), ob in blocks:
        h = Hrr(A_, B_, C_, D_)
        outs = []
        for a in comps(A_):
            for b in comps(B_):
                for c in comps(C_):
                    for d in comps(D_):
                        h.need_bra(a, b, c, d)
                        outs.append((a, b, c, d))

        def kkey(e, c, d, _n=name, _s=si):
            return ("acc", _s, e, c) if d == Z else ("k", _n, e, c, d)

        def bkey(a, b, c, d, _n=name, _k=kkey):
            return _k(a, c, d) if b == Z else ("b", _n, a, b, c, d)

        for (e, c, d) in h.order_k:
            i = h.ket[(e, c, d)]
            hn[kkey(e, c, d)] = (am(d), i, kkey(e, add(c, i), sub(d, i)), kkey(e, c, sub(d, i)))
        for (a, b, c, d) in h.order_b:
            i = h.bra[(a, b, c, d)]
            hn[bkey(a, b, c, d)] = (D_ + am(b), 3 + i, bkey(add(a, i), sub(b, i), c, d), bkey(a, sub(b, i), c, d))
        for x, o4 in enumerate(outs):
            k = bkey(*o4)
            if k[0] == "acc":      # the accumulator is the integral: copy it out
                ck = ("copy", name, x)
                hn[ck] = (1, 0, k, "ZERO")
                out_pin[ck] = ob + x
            else:
                assert k not in out_pin
                out_pin[k] = ob + x
    hlevel = {k: lv for k, (lv, _, _, _) in hn.items()}
    hlast = {}
    for k, (lv, _, s1, s2) in hn.items():
        for x in (s1, s2):
            hlast[x] = max(hlast.get(x, -1), lv)
    acc_keys = [("acc", si, t[0], t[1]) for (si, t) in acc]
    for a_ in acc_keys:
        hlevel[a_] = 0
    pinned = {("acc", si, t[0], t[1]): x for (si, t), x in acc.items()}
    pinned["ZERO"] = ZS
    hlevel["ZERO"] = 0
    pinned.update(out_pin)
    for k in hn:
        assert k[0] != "acc"
    hpool = SlotPool(POOL)
    inner = [k for k in hn if k not in pinned]
    for k in inner:
        hlast.setdefault(k, -1)
    hslot = _allocate(acc_keys + ["ZERO"] + sorted(inner, key=lambda k: (hlevel[k], str(k))) +
                      [k for k in hn if k in pinned], hlevel, hlast, hpool, pinned=pinned)
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
        assert hslot[k] is not None and hslot[s1] is not None and hslot[s2] is not None
        hnodes.append((hslot[k], hslot[s1], hslot[s2], sel))
    hlev.append(len(hkeys))
    cf = max(vmax, hmax, POOL)
    dm = cf + 40
    total = dm + na * nb + nc * nd + na * nc + na * nd + nb * nc + nb * nd
    assert total < 65536
    # FLOPs (FMA = 2): VRR as in the Fock tables; 2 per accumulation entry;
    # 2 per HRR node; per integral of the quartet the weight (10) and the
    # nine derivative terms (4 each)
    vflops = 0
    for k in order:
        kind, (i, deps) = v.nodes[k]
        vflops += 3
        for d in deps[2:]:
            if d[0] in ("Z2", "N2"):
                vflops += 4 if d[1] > 1 else 3
            else:
                vflops += 3 if d[1] > 1 else 2
    hflops = 2 * sum(1 for k in hn if k[0] != "copy")
    return dict(L=L, nacc=nacc, zs=ZS, sink=SINK, out=out, pool=POOL, nv=len(order), vlev=vlev, vnodes=vnodes,
                fslot=fslot, nx=len(accs), accs=accs, nh=len(hnodes), hnodes=hnodes, hlev=hlev, ni=ni, cf=cf, dm=dm,
                total=total, glob=total * 8 > COOP_LDS_CAP,
                prim_flop=boys_setup_flop(L) + vflops + 2 * len(accs), quartet_flop=hflops + ni * (10 + 9 * 4))


# Classes with La+Lb+Lc+Ld up to this also get a per-lane gradient kernel
# (one quartet per lane, recursions in registers); the wave-per-quartet
# kernel serves the rest.  At 4 and 5 the per-lane kernels spill 1.3-10 KB
# per lane and are still faster than the wave-per-quartet ones ((H2O)32
# 6-31G(d) gradient: 192 ms with 3, 134 ms with 4, 123 ms with 5).
LANE_GRAD_LSUM_MAX = 5


def has_lane_grad(La, Lb, Lc, Ld):
    return La + Lb + Lc + Ld <= LANE_GRAD_LSUM_MAX


def _hrr_lines(prefix, cls, accexpr):
    """Straight-line HRR of one block; returns (lines, {output (a,b,c,d): expression})."""
    A_, B_, C_, D_ = cls
    h = Hrr(A_, B_, C_, D_)
    outs = []
    for a in comps(A_):
        for b in comps(B_):
            for c in comps(C_):
                for d in comps(D_):
                    h.need_bra(a, b, c, d)
                    outs.append((a, b, c, d))
    tg = lambda t: "%d%d%d" % t

    def kv(e, c, d):
        return accexpr(e, c) if d == Z else "%sk%s_%s_%s" % (prefix, tg(e), tg(c), tg(d))

    def bv(a, b, c, d):
        return kv(a, c, d) if b == Z else "%sb%s_%s_%s_%s" % (prefix, tg(a), tg(b), tg(c), tg(d))

    lines = []
    for (e, c, d) in h.order_k:
        i = h.ket[(e, c, d)]
        lines.append("const double %s = %s + CD[%d] * %s;" % (kv(e, c, d), kv(e, add(c, i), sub(d, i)), i,
                                                           kv(e, c, sub(d, i))))
    for (a, b, c, d) in h.order_b:
        i = h.bra[(a, b, c, d)]
        lines.append("const double %s = %s + AB[%d] * %s;" % (bv(a, b, c, d), bv(add(a, i), sub(b, i), c, d), i,
                                                           bv(a, sub(b, i), c, d)))
    return lines, {o: bv(*o) for o in outs}


def lane_grad_source(name, La, Lb, Lc, Ld):
    """Per-lane gradient recursions of one class: ``vrr`` adds one primitive
    quartet into the four weighted sets, ``digest`` runs the six HRR blocks
    and contracts the derivative integrals with the density blocks into
    g[9] = dE/dA, dE/dB, dE/dC (without the degeneracy weight)."""
    sets = grad_sets(La, Lb, Lc, Ld)
    L = La + Lb + Lc + Ld + 1
    acc = {}
    for si, st in enumerate(sets):
        for t in st:
            acc[(si, t)] = len(acc)
    targets, seen = [], set()
    for st in sets:
        for t in st:
            if t not in seen:
                seen.add(t)
                targets.append(t)
    v = VrrTargets(targets, L)
    vlines, _ = v.emit({t: 0 for t in targets})
    vlines = vlines[:-len(targets)]
    for (si, t), x in sorted(acc.items(), key=lambda kv: kv[1]):
        vlines.append("acc[%d] += w[%d] * %s;" % (x, si, v.var((t[0], t[1], 0))))
    vlines = [l.replace("const R ", "const double ").replace("R(", "double(") for l in vlines]
    na, nb, nc, nd = (ncart(x) for x in (La, Lb, Lc, Ld))
    dlines, outs = [], {}
    for bname, si, (da, db, dc) in BLOCKS:
        cls = (La + da, Lb + db, Lc + dc, Ld)
        if min(cls) < 0:
            continue
        lines, o = _hrr_lines(bname.lower() + "_", cls, lambda e, c, _s=si: "acc[%d]" % acc[(_s, (e, c))])
        dlines += lines
        outs[bname] = o
    ca, cb_, cc, cd = comps(La), comps(Lb), comps(Lc), comps(Ld)
    for ia, a in enumerate(ca):
        for ib, b in enumerate(cb_):
            for ic, c in enumerate(cc):
                for id_, d in enumerate(cd):
                    dlines.append("{")
                    dlines.append("  const double gm = kCons[%d][%d] * kCons[%d][%d] * kCons[%d][%d] * kCons[%d][%d] * "
                                  "(2.0 * Dab[%d] * Dcd[%d] - hh * (Dac[%d] * Dbd[%d] + Dad[%d] * Dbc[%d]));"
                                  % (La, ia, Lb, ib, Lc, ic, Ld, id_, ia * nb + ib, ic * nd + id_, ia * nc + ic,
                                     ib * nd + id_, ia * nd + id_, ib * nc + ic))
                    for x in range(3):
                        tA = outs["AP"][(add(a, x), b, c, d)]
                        if a[x] > 0:
                            tA = "(%s - %d.0 * %s)" % (tA, a[x], outs["AM"][(sub(a, x), b, c, d)])
                        tB = outs["BP"][(a, add(b, x), c, d)]
                        if b[x] > 0:
                            tB = "(%s - %d.0 * %s)" % (tB, b[x], outs["BM"][(a, sub(b, x), c, d)])
                        tC = outs["CP"][(a, b, add(c, x), d)]
                        if c[x] > 0:
                            tC = "(%s - %d.0 * %s)" % (tC, c[x], outs["CM"][(a, b, sub(c, x), d)])
                        dlines.append("  g[%d] += gm * %s;" % (x, tA))
                        dlines.append("  g[%d] += gm * %s;" % (3 + x, tB))
                        dlines.append("  g[%d] += gm * %s;" % (6 + x, tC))
                    dlines.append("}")
    g = name + "GradLane"
    src = ["struct %s {" % g,
           "  static constexpr int La = %d, Lb = %d, Lc = %d, Ld = %d, L = %d;" % (La, Lb, Lc, Ld, L),
           "  static constexpr int NACC = %d;" % len(acc),
           "  // w = 1, 2 alpha, 2 beta, 2 gamma: QUICK's store, storeAA, storeBB, storeCC",
           "  __device__ static __forceinline__ void vrr(const double (&F)[L + 1], const VrrPar<double>& p,"
           " const double (&w)[4], double (&acc)[NACC]) {"]
    src += ["    " + l for l in vlines]
    src += ["  }",
            "  __device__ static __forceinline__ void digest(const double (&acc)[NACC], const double (&AB)[3],"
            " const double (&CD)[3], const double (&Dab)[%d], const double (&Dcd)[%d], const double (&Dac)[%d],"
            " const double (&Dad)[%d], const double (&Dbc)[%d], const double (&Dbd)[%d], double hh, double (&g)[9]) {"
            % (na * nb, nc * nd, na * nc, na * nd, nb * nc, nb * nd)]
    src += ["    " + l for l in dlines]
    src += ["  }", "};"]
    return g, "\n".join(src) + "\n"


def grad_source(name, La, Lb, Lc, Ld):
    t = grad_tables(La, Lb, Lc, Ld)
    g = name + "Grad"
    o = t["out"]
    src = ["struct %s {" % g,
           "  static constexpr int La = %d, Lb = %d, Lc = %d, Ld = %d, L = %d;" % (La, Lb, Lc, Ld, t["L"]),
           "  static constexpr int NACC = %d, NV = %d, NLV = %d, NX = %d, NH = %d, NLH = %d, NI = %d;"
           % (t["nacc"], t["nv"], len(t["vlev"]) - 1, t["nx"], t["nh"], len(t["hlev"]) - 1, t["ni"]),
           "  static constexpr int ZS = %d, SINK = %d, CF = %d, DM = %d, LDS = %d;"
           % (t["zs"], t["sink"], t["cf"], t["dm"], t["total"]),
           "  // output blocks (a+1 b|cd), (a-1 b|cd), (a b+1|cd), (a b-1|cd), (ab|c+1 d), (ab|c-1 d); -1: absent",
           "  static constexpr int OAP = %d, OAM = %d, OBP = %d, OBM = %d, OCP = %d, OCM = %d;"
           % (o["AP"], o["AM"], o["BP"], o["BM"], o["CP"], o["CM"]),
           "  static constexpr bool GLOBAL_SLOTS = %s;   // slots in a global scratch buffer (over %d KiB)"
           % ("true" if t["glob"] else "false", COOP_LDS_CAP // 1024),
           "  static constexpr int NT = %d;   // threads per workgroup: one wave" % COOP_NT,
           "  static constexpr int PRIM_FLOP = %d, QUARTET_FLOP = %d;" % (t["prim_flop"], t["quartet_flop"]),
           "  static constexpr unsigned short FSLOT[%d] = {%s};   // slots of [00|00]^(m)"
           % (t["L"] + 1, ",".join(map(str, t["fslot"]))),
           "  __device__ static const CoopVNode* vrr();",
           "  __device__ static const int* vlev();",
           "  __device__ static const GradAcc* acc();",
           "  __device__ static const CoopHNode* hrr();",
           "  __device__ static const int* hlev();",
           "};"]
    fv = ",".join("{%d,%d,%d,%d,%d,%d,%d,%d}" % x for x in t["vnodes"])
    fa = ",".join("{%d,%d,%d}" % x for x in t["accs"])
    fh = ",".join("{%d,%d,%d,%d}" % x for x in t["hnodes"])
    src += ["__device__ const CoopVNode k%s_vrr[%d] = {%s};" % (g, max(1, t["nv"]), fv or "{0,0,0,0,0,0,0,0}"),
            "__device__ const int k%s_vlev[%d] = {%s};" % (g, len(t["vlev"]), ",".join(map(str, t["vlev"]))),
            "__device__ const GradAcc k%s_acc[%d] = {%s};" % (g, t["nx"], fa),
            "__device__ const CoopHNode k%s_hrr[%d] = {%s};" % (g, max(1, t["nh"]), fh or "{0,0,0,0}"),
            "__device__ const int k%s_hlev[%d] = {%s};" % (g, len(t["hlev"]), ",".join(map(str, t["hlev"]))),
            "__device__ inline const CoopVNode* %s::vrr() { return k%s_vrr; }" % (g, g),
            "__device__ inline const int* %s::vlev() { return k%s_vlev; }" % (g, g),
            "__device__ inline const GradAcc* %s::acc() { return k%s_acc; }" % (g, g),
            "__device__ inline const CoopHNode* %s::hrr() { return k%s_hrr; }" % (g, g),
            "__device__ inline const int* %s::hlev() { return k%s_hlev; }" % (g, g)]
    return g, "\n".join(src) + "\n", t


def grad_index_tables():
    """Component maps of the derivative terms for L <= GRAD_LMAX: index of
    comp + 1_x in comps(L + 1), of comp - 1_x in comps(L - 1) (0 when the
    power is 0) and the power n_x itself, [L][comp * 3 + x]."""
    plus, minus, power = [], [], []
    for L in range(GRAD_LMAX + 1):
        cs = comps(L)
        p, m, w = [0] * 18, [0] * 18, [0] * 18
        for a, t in enumerate(cs):
            for x in range(3):
                p[a * 3 + x] = comps(L + 1).index(add(t, x))
                w[a * 3 + x] = t[x]
                m[a * 3 + x] = comps(L - 1).index(sub(t, x)) if t[x] > 0 else 0
        plus.append(p)
        minus.append(m)
        power.append(w)
    return plus, minus, power


def launcher_source(name, g, lane):
    extra = []
    if lane:
        extra = [
            "// per-lane kernel: grad_tile as above; no global slots",
            "void launch_grad_lane_%s(const Task* t, int nt, const PairBlock* b, const MolDesc* m, int grid,"
            " size_t grad_tile, double* gscratch, int fp32, hipStream_t s)" % name, "{",
            "    (void)gscratch;",
            "    if (nt <= 0) return;",
            "    if (grid <= 0 || grid > nt) grid = nt;",
            "    hipLaunchKernelGGL((grad_lane_kernel<%sGradLane>), dim3(grid), dim3(WAVE), grad_tile, s, t, nt, b, m,"
            " (int)(grad_tile / sizeof(double)), fp32);" % name,
            "}", ""]
    return "\n".join([
        "// Generated by quick_amd/codegen/gen_grad.py - do not edit.",
        '#include "grad_%s.h"' % name, "", "namespace qf {", "",
        "// grad_tile: 3 natom doubles of LDS for the workgroup's gradient (0: atomics",
        "// straight to md.grad); gscratch: COOP_GLOBAL_SLICES slices when the class",
        "// keeps its slots in global memory",
        "void launch_grad_%s(const Task* t, int nt, const PairBlock* b, const MolDesc* m, int grid,"
        " size_t grad_tile, double* gscratch, int fp32, hipStream_t s)" % name, "{",
        "    using D = %s;" % g,
        "    if (nt <= 0) return;",
        "    if (grid <= 0 || grid > nt) grid = nt;",
        "    if (D::GLOBAL_SLOTS && grid > COOP_GLOBAL_SLICES) grid = COOP_GLOBAL_SLICES;",
        "    const size_t lds = (D::GLOBAL_SLOTS ? 0 : D::LDS * sizeof(double)) + grad_tile;",
        "    hipLaunchKernelGGL((grad_coop_kernel<D>), dim3(grid), dim3(D::NT), lds, s, t, nt, b, m, gscratch,"
        " (int)(grad_tile / sizeof(double)), fp32);",
        "}", ""] + extra + ["}  // namespace qf", ""])


def main(outdir, write):
    pc = {pcl: k for k, pcl in enumerate(PAIR_CLASSES)}
    names = []
    table = {}
    for (La, Lb, Lc, Ld) in grad_classes():
        name = "C%d%d%d%d" % (La, Lb, Lc, Ld)
        g, src, t = grad_source(name, La, Lb, Lc, Ld)
        lane = has_lane_grad(La, Lb, Lc, Ld)
        if lane:
            src += lane_grad_source(name, La, Lb, Lc, Ld)[1]
        write("grad_%s.h" % name, "// Generated by quick_amd/codegen/gen_grad.py - do not edit.\n#pragma once\n"
              "#include \"../qf_grad.h\"\nnamespace qf {\n" + src + "}  // namespace qf\n")
        write("egrad_%s.hip" % name, launcher_source(name, g, lane))
        names.append((name, La, Lb, Lc, Ld, t))
        table[name] = dict(prim_flop=t["prim_flop"], quartet_flop=t["quartet_flop"], lds=t["total"], glob=t["glob"])
    plus, minus, power = grad_index_tables()
    fmt = lambda rows: ",".join("{%s}" % ",".join(map(str, r)) for r in rows)
    lines = ["// Generated by quick_amd/codegen/gen_grad.py - do not edit.", "#pragma once",
             "#include <hip/hip_runtime.h>", '#include "../qf_types.h"', "", "namespace qf {", "",
             "using GradLaunch = void (*)(const Task*, int, const PairBlock*, const MolDesc*, int, size_t, double*,"
             " int, hipStream_t);", ""]
    for (name, La, Lb, Lc, Ld, t) in names:
        lines.append("void launch_grad_%s(const Task*, int, const PairBlock*, const MolDesc*, int, size_t, double*,"
                     " int, hipStream_t);" % name)
        if has_lane_grad(La, Lb, Lc, Ld):
            lines.append("void launch_grad_lane_%s(const Task*, int, const PairBlock*, const MolDesc*, int, size_t,"
                         " double*, int, hipStream_t);" % name)
    lines += ["", "// nullptr: no gradient kernel for the class (f shells)",
              "inline GradLaunch grad_launcher(int cb, int ck)", "{"]
    for (name, La, Lb, Lc, Ld, t) in names:
        lines.append("    if (cb == %d && ck == %d) return launch_grad_%s;" % (pc[(La, Lb)], pc[(Lc, Ld)], name))
    lines += ["    return nullptr;", "}", "", "// per-lane gradient kernels (nullptr: the class has none)",
              "inline GradLaunch grad_lane_launcher(int cb, int ck)", "{"]
    for (name, La, Lb, Lc, Ld, t) in names:
        if has_lane_grad(La, Lb, Lc, Ld):
            lines.append("    if (cb == %d && ck == %d) return launch_grad_lane_%s;" % (pc[(La, Lb)], pc[(Lc, Ld)], name))
    lines += ["    return nullptr;", "}", "", "// doubles per workgroup slice of the global slot buffer (0: LDS slots)",
              "inline int grad_global_doubles(int cb, int ck)", "{"]
    for (name, La, Lb, Lc, Ld, t) in names:
        if t["glob"]:
            lines.append("    if (cb == %d && ck == %d) return %d;" % (pc[(La, Lb)], pc[(Lc, Ld)], t["total"]))
    lines += ["    return 0;", "}", "", "// LDS doubles of the node-program slots (0 when they are global)",
              "inline int grad_lds_doubles(int cb, int ck)", "{"]
    for (name, La, Lb, Lc, Ld, t) in names:
        if not t["glob"]:
            lines.append("    if (cb == %d && ck == %d) return %d;" % (pc[(La, Lb)], pc[(Lc, Ld)], t["total"]))
    lines += ["    return 0;", "}", "",
              "// Algorithmic FLOPs of the gradient: per primitive quartet (Boys, VRR,",
              "// accumulation) and per contracted quartet (HRR blocks, weights, nine derivatives).",
              "inline double grad_class_flops(int cb, int ck, unsigned long long nq, unsigned long long np)", "{"]
    for (name, La, Lb, Lc, Ld, t) in names:
        lines.append("    if (cb == %d && ck == %d) return double(np) * %d.0 + double(nq) * %d.0;"
                     % (pc[(La, Lb)], pc[(Lc, Ld)], t["prim_flop"], t["quartet_flop"]))
    lines += ["    return 0.0;", "}", "", "}  // namespace qf", ""]
    write("grad_launchers.h", "\n".join(lines))
    write("grad_maps.h", "\n".join([
        "// Generated by quick_amd/codegen/gen_grad.py - do not edit.", "#pragma once", "", "namespace qf {", "",
        "// derivative component maps, [L][component * 3 + direction]: index of",
        "// component + 1_x in comps(L + 1), of component - 1_x in comps(L - 1), and n_x",
        "__device__ const unsigned char kGradPlus[%d][18] = {%s};" % (len(plus), fmt(plus)),
        "__device__ const unsigned char kGradMinus[%d][18] = {%s};" % (len(minus), fmt(minus)),
        "__device__ const unsigned char kGradPow[%d][18] = {%s};" % (len(power), fmt(power)),
        "", "}  // namespace qf", ""]))
    return table
