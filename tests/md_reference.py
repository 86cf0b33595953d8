"""Independent two-electron integrals by the McMurchie-Davidson scheme.

A different algorithm from the Obara-Saika/HGP recursions used by the oracle
and by the HIP kernels: Hermite expansion coefficients E^{ij}_t per Cartesian
direction and Hermite Coulomb integrals R_{tuv}.  Used only by tests, on tiny
shell quartets, to pin the VRR/HRR/normalisation of the oracle.  The Boys
function is passed in, so the comparison can either share QUICK's FmT (to
isolate the recursions) or use an exact one.
"""
import math

import numpy as np

CART = {
    0: [(0, 0, 0)],
    1: [(1, 0, 0), (0, 1, 0), (0, 0, 1)],
    2: [(2, 0, 0), (1, 1, 0), (0, 2, 0), (1, 0, 1), (0, 1, 1), (0, 0, 2)],
    3: [(3, 0, 0), (2, 1, 0), (1, 2, 0), (0, 3, 0), (2, 0, 1), (1, 1, 1), (0, 2, 1), (1, 0, 2), (0, 1, 2),
        (0, 0, 3)],
}
_S5 = math.sqrt(5)
CONS = {0: [1.0], 1: [1.0] * 3, 2: [1.0, math.sqrt(3), 1.0, math.sqrt(3), math.sqrt(3), 1.0],
        3: [1.0, _S5, _S5, 1.0, _S5, math.sqrt(15), _S5, _S5, _S5, 1.0]}
NH = 7  # Hermite index range per direction (0..6)
_IDX = np.array([(t, u, v) for t in range(NH) for u in range(NH) for v in range(NH)])


def hermite_E(la, lb, a, b, XA, XB):
    """E[i][j][t] for i<=la, j<=lb in one direction."""
    p = a + b
    mu = a * b / p
    XAB = XA - XB
    XP = (a * XA + b * XB) / p
    XPA, XPB = XP - XA, XP - XB
    E = np.zeros((la + 1, lb + 1, la + lb + 2))
    E[0, 0, 0] = math.exp(-mu * XAB * XAB)
    for i in range(la + 1):
        for j in range(lb + 1):
            if i == 0 and j == 0:
                continue
            for t in range(i + j + 1):
                if i > 0:
                    v = XPA * E[i - 1, j, t]
                    if t > 0:
                        v += E[i - 1, j, t - 1] / (2 * p)
                    v += (t + 1) * E[i - 1, j, t + 1]
                else:
                    v = XPB * E[i, j - 1, t]
                    if t > 0:
                        v += E[i, j - 1, t - 1] / (2 * p)
                    v += (t + 1) * E[i, j - 1, t + 1]
                E[i, j, t] = v
    return E


def hermite_R(L, alpha, PQ, boys):
    """R[t][u][v] (n = 0) for t+u+v <= L."""
    T = alpha * float(PQ @ PQ)
    F = boys(L, T)
    Rn = {}
    for n in range(L + 1):
        Rn[(0, 0, 0, n)] = (-2 * alpha) ** n * F[n]
    X, Y, Z = PQ

    def R(t, u, v, n):
        key = (t, u, v, n)
        if key in Rn:
            return Rn[key]
        if t > 0:
            val = X * R(t - 1, u, v, n + 1)
            if t > 1:
                val += (t - 1) * R(t - 2, u, v, n + 1)
        elif u > 0:
            val = Y * R(t, u - 1, v, n + 1)
            if u > 1:
                val += (u - 1) * R(t, u - 2, v, n + 1)
        else:
            val = Z * R(t, u, v - 1, n + 1)
            if v > 1:
                val += (v - 1) * R(t, u, v - 2, n + 1)
        Rn[key] = val
        return val

    out = np.zeros((2 * NH - 1,) * 3)
    for t in range(L + 1):
        for u in range(L + 1 - t):
            for v in range(L + 1 - t - u):
                out[t, u, v] = R(t, u, v, 0)
    return out


def shell_block(sh, boys):
    """(ab|cd) block for four shells; each shell = (L, center[3], exps, coefs)
    with coefs the normalised x^L contraction coefficients; the per-component
    cons factor is applied here.  Returns array [na, nb, nc, nd]."""
    (La, A, ea, ca), (Lb, B, eb, cb), (Lc, C, ec, cc), (Ld, D, ed, cd) = sh
    A, B, C, D = (np.asarray(x, dtype=float) for x in (A, B, C, D))
    ca_, cb_, cc_, cd_ = CART[La], CART[Lb], CART[Lc], CART[Ld]
    out = np.zeros((len(ca_), len(cb_), len(cc_), len(cd_)))
    L = La + Lb + Lc + Ld
    idx = [(t, u, v) for t in range(NH) for u in range(NH) for v in range(NH)]
    for a_e, a_c in zip(ea, ca):
        for b_e, b_c in zip(eb, cb):
            p = a_e + b_e
            P = (a_e * A + b_e * B) / p
            Ex = [hermite_E(La, Lb, a_e, b_e, A[d], B[d]) for d in range(3)]
            Eab = np.zeros((len(ca_) * len(cb_), NH ** 3))
            for i, ai in enumerate(ca_):
                for j, bj in enumerate(cb_):
                    for k, (t, u, v) in enumerate(idx):
                        if t <= ai[0] + bj[0] and u <= ai[1] + bj[1] and v <= ai[2] + bj[2]:
                            Eab[i * len(cb_) + j, k] = Ex[0][ai[0], bj[0], t] * \
                                Ex[1][ai[1], bj[1], u] * Ex[2][ai[2], bj[2], v]
            for c_e, c_c in zip(ec, cc):
                for d_e, d_c in zip(ed, cd):
                    q = c_e + d_e
                    Q = (c_e * C + d_e * D) / q
                    Fx = [hermite_E(Lc, Ld, c_e, d_e, C[d], D[d]) for d in range(3)]
                    Ecd = np.zeros((len(cc_) * len(cd_), NH ** 3))
                    for i, ci in enumerate(cc_):
                        for j, dj in enumerate(cd_):
                            for k, (t, u, v) in enumerate(idx):
                                if t <= ci[0] + dj[0] and u <= ci[1] + dj[1] and v <= ci[2] + dj[2]:
                                    Ecd[i * len(cd_) + j, k] = (-1) ** (t + u + v) * \
                                        Fx[0][ci[0], dj[0], t] * Fx[1][ci[1], dj[1], u] * Fx[2][ci[2], dj[2], v]
                    alpha = p * q / (p + q)
                    R = hermite_R(L, alpha, P - Q, boys)
                    Rm = R[_IDX[:, None, 0] + _IDX[None, :, 0], _IDX[:, None, 1] + _IDX[None, :, 1],
                           _IDX[:, None, 2] + _IDX[None, :, 2]]
                    pref = 2 * math.pi ** 2.5 / (p * q * math.sqrt(p + q)) * a_c * b_c * c_c * d_c
                    blk = pref * (Eab @ Rm @ Ecd.T)
                    out += blk.reshape(out.shape)
    cons = [np.array(CONS[x]) for x in (La, Lb, Lc, Ld)]
    return out * np.einsum("i,j,k,l->ijkl", *cons)


def boys_exact(maxm, T):
    """F_m(T) to double precision with mpmath."""
    import mpmath as mp
    mp.mp.dps = 40
    out = np.zeros(maxm + 1)
    for m in range(maxm + 1):
        if T == 0:
            out[m] = 1.0 / (2 * m + 1)
        else:
            out[m] = float(mp.gammainc(m + 0.5, 0, T) / (2 * mp.mpf(T) ** (m + 0.5)))
    return out
