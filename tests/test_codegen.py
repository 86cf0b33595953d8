"""The node programs of the wave-per-quartet kernels (level-scheduled VRR
and HRR over LDS slots) compute exactly the recursions of the class: run
them with a small interpreter on random recursion parameters and compare
with a direct evaluation of the same Head-Gordon-Pople recurrences."""
import numpy as np
import pytest

from quick_amd.codegen import gen_eri as G


def reference(La, Lb, Lc, Ld, F, p):
    """Direct recursive evaluation of the VRR (HGP eqs. 6, 16) and HRR."""
    memo = {}

    def vrr(e, f, m):
        k = (e, f, m)
        if k in memo:
            return memo[k]
        if e == G.Z and f == G.Z:
            r = F[m]
        elif f == G.Z:
            i = next(j for j in range(3) if e[j] > 0)
            i = min((j for j in range(3) if e[j] > 0), key=lambda j: (e[j] > 1, -e[j]))
            e1 = G.sub(e, i)
            r = p["PA"][i] * vrr(e1, f, m) + p["WP"][i] * vrr(e1, f, m + 1)
            if e1[i] > 0:
                e2 = G.sub(e1, i)
                r += e1[i] * p["oo2z"] * (vrr(e2, f, m) - p["rz"] * vrr(e2, f, m + 1))
        else:
            i = min((j for j in range(3) if f[j] > 0), key=lambda j: (e[j] > 0, f[j] > 1, -f[j]))
            f1 = G.sub(f, i)
            r = p["QC"][i] * vrr(e, f1, m) + p["WQ"][i] * vrr(e, f1, m + 1)
            if f1[i] > 0:
                f2 = G.sub(f1, i)
                r += f1[i] * p["oo2n"] * (vrr(e, f2, m) - p["rn"] * vrr(e, f2, m + 1))
            if e[i] > 0:
                r += e[i] * p["oo2zn"] * vrr(G.sub(e, i), f1, m + 1)
        memo[k] = r
        return r

    def ket(e, c, d):
        if d == G.Z:
            return vrr(e, c, 0)
        i = next(j for j in range(3) if d[j] > 0)
        return ket(e, G.add(c, i), G.sub(d, i)) + p["CD"][i] * ket(e, c, G.sub(d, i))

    def bra(a, b, c, d):
        if b == G.Z:
            return ket(a, c, d)
        i = next(j for j in range(3) if b[j] > 0)
        return bra(G.add(a, i), G.sub(b, i), c, d) + p["AB"][i] * bra(a, G.sub(b, i), c, d)

    return np.array([bra(*o) for o in G.coop_tables(La, Lb, Lc, Ld)["outs"]])


def interpret(t, prims, AB, CD):
    """Execute the node programs the way fock_coop_kernel does, for a
    quartet of several primitive quartets (F, p) sharing the pair geometry
    (AB, CD).  Checks that no slot is both read and written in one level."""
    s = np.full(t["total"] + 8, np.nan)
    nacc = t["nacc"]
    s[:nacc] = 0.0
    for F, p in prims:
        s[t["zs"]] = 0.0
        for m in range(t["L"] + 1):
            s[t["fslot"][m]] = F[m]
        cf = np.zeros(6 * 4)
        for sel in range(6):
            op, d = sel // 3, sel % 3
            cf[sel * 4:sel * 4 + 4] = ([p["PA"][d], p["WP"][d], p["oo2z"], p["rz"]] if op == 0 else
                                       [p["QC"][d], p["WQ"][d], p["oo2n"], p["rn"]])
        lev = t["vlev"]
        for lv in range(len(lev) - 1):
            nodes = t["vnodes"][lev[lv]:lev[lv + 1]]
            reads = {x for nd in nodes for x in nd[:5]}
            writes = {nd[5] for nd in nodes} | {nd[6] for nd in nodes if nd[6] != 0xFFFF}
            assert not (reads & writes), "a level reads a slot it writes"
            vals = []
            for (a, b, c, d, e, dst, acc, meta) in nodes:
                sel, n1, n2 = meta & 15, (meta >> 4) & 15, meta >> 8
                c4 = cf[sel * 4:sel * 4 + 4]
                vals.append(c4[0] * s[a] + c4[1] * s[b] + n1 * c4[2] * (s[c] - c4[3] * s[d]) + n2 * p["oo2zn"] * s[e])
            for (a, b, c, d, e, dst, acc, meta), val in zip(nodes, vals):
                s[dst] = val
                if acc != 0xFFFF:
                    s[acc] += val
    s[nacc + 1:t["out"]] = np.nan          # the sink
    s[t["zs"]] = 0.0
    hcf = list(CD) + list(AB)
    lev = t["hlev"]
    for lv in range(len(lev) - 1):
        nodes = t["hnodes"][lev[lv]:lev[lv + 1]]
        reads = {x for nd in nodes for x in nd[1:3]}
        writes = {nd[0] for nd in nodes}
        assert not (reads & writes), "an HRR level reads a slot it writes"
        vals = [s[s1] + hcf[sel] * s[s2] for (dst, s1, s2, sel) in nodes]
        for (dst, s1, s2, sel), val in zip(nodes, vals):
            s[dst] = val
    return s[t["out"]:t["out"] + t["ni"]]


def _params(rng):
    p = {k: rng.uniform(-1, 1, 3) for k in ("PA", "WP", "QC", "WQ")}
    for k in ("oo2z", "oo2n", "oo2zn", "rz", "rn"):
        p[k] = float(rng.uniform(0.1, 0.9))
    return p


@pytest.mark.parametrize("cls", G.COOP_CLASSES, ids=lambda c: "%d%d%d%d" % c)
def test_coop_node_programs_match_recursion(cls):
    """Two primitive quartets: the accumulators must hold their sum after
    the slots of the first one were reused by the second."""
    rng = np.random.default_rng(sum(cls) * 7 + 1)
    t = G.coop_tables(*cls)
    L = t["L"]
    AB, CD = rng.uniform(-1, 1, 3), rng.uniform(-1, 1, 3)
    prims = []
    ref = 0.0
    for _ in range(2):
        F = rng.uniform(0.1, 1.0, L + 1)
        p = _params(rng)
        prims.append((F, p))
        ref = ref + reference(*cls, F, dict(p, AB=AB, CD=CD))
    got = interpret(t, prims, AB, CD)
    assert got.shape == ref.shape
    assert np.all(np.isfinite(got))
    assert np.allclose(got, ref, rtol=1e-12, atol=1e-14)
    assert t["total"] * 8 <= G.COOP_LDS_CAP or t["glob"]


def test_fp32_boys_grid_table():
    """The FP32 Boys grid (gen_boys.py, read by boys_f32 for the FP32 quartets
    of split-precision builds): the committed header equals the generator,
    the series values equal F_m(x) = Gamma(m+1/2) P(m+1/2, x) / (2 x^(m+1/2))
    (scipy's regularized incomplete gamma) to 1e-13, and the 5-term Taylor
    step the kernel takes from the nearest grid point stays below 1e-7
    relative over the grid range for every L the per-lane classes use."""
    import math
    import os
    from scipy.special import gamma, gammainc
    from quick_amd.codegen import gen_boys as GB
    hdr = os.path.join(os.path.dirname(G.__file__), "..", "csrc", "gen", "boys_grid.h")
    assert open(hdr).read() == GB.header()
    tab = GB.table()
    worst = 0.0
    for i in range(1, GB.NG, 7):
        x = i * GB.H
        for m in range(GB.MG):
            ref = gamma(m + 0.5) * gammainc(m + 0.5, x) / (2.0 * x ** (m + 0.5))
            worst = max(worst, abs(tab[i][m] - ref) / ref)
    assert worst < 1e-13
    assert all(abs(tab[0][m] - 1.0 / (2 * m + 1)) < 1e-16 for m in range(GB.MG))
    rng = np.random.default_rng(5)
    worst = 0.0
    for L in range(6):
        for x in rng.uniform(0.0, (GB.NG - 1) * GB.H - 1e-9, 400):
            i = int(x / GB.H + 0.5)
            d = i * GB.H - x
            s = sum(tab[i][L + k] * d ** k / math.factorial(k) for k in range(5))
            worst = max(worst, abs(s - GB.boys_series(L, x)) / GB.boys_series(L, x))
    assert worst < 1e-7
