"""Seeded geometry generators for the Fock-build workloads.

All generators return :class:`quick_amd.basis.Molecule` (coordinates in bohr)
and are deterministic for a given seed.

* ``water()``                 - O at the origin, O-H 0.9572 A, HOH 104.52 deg.
* ``water_cluster(n, seed)``  - n rigid waters, uniform random orientations,
                                O positions uniform in a cube sized for liquid
                                density (29.9 A^3 per molecule), min O-O 2.7 A.
* ``alkane(n)``               - all-trans CnH2n+2, C-C 1.54 A, C-H 1.09 A,
                                tetrahedral angles.
* ``polyglycine(n)``          - H-(Gly)n-OH extended beta strand
                                (phi -139, psi 135, omega 180).
"""
from __future__ import annotations

import math

import numpy as np

from .basis import Molecule

SEED = 20261202
OH = 0.9572
HOH = 104.52


def water_local():
    h = math.radians(HOH / 2.0)
    return np.array([[0.0, 0.0, 0.0],
                     [OH * math.sin(h), 0.0, OH * math.cos(h)],
                     [-OH * math.sin(h), 0.0, OH * math.cos(h)]])


def water():
    return Molecule.from_angstrom(["O", "H", "H"], water_local())


def _random_rotation(rng):
    q = rng.normal(size=4)
    q /= np.linalg.norm(q)
    w, x, y, z = q
    return np.array([
        [1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
        [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
        [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)],
    ])


def water_cluster(n, seed=SEED, min_oo=2.7, volume_per_water=29.9):
    rng = np.random.default_rng(seed)
    side = (n * volume_per_water) ** (1.0 / 3.0)
    opos = []
    tries = 0
    while len(opos) < n:
        tries += 1
        if tries > 200000:
            raise RuntimeError("could not place waters")
        p = rng.uniform(0.0, side, size=3)
        if all(np.linalg.norm(p - q) >= min_oo for q in opos):
            opos.append(p)
    loc = water_local()
    syms, xyz = [], []
    for p in opos:
        R = _random_rotation(rng)
        for s, r in zip(["O", "H", "H"], loc):
            syms.append(s)
            xyz.append(p + R @ r)
    return Molecule.from_angstrom(syms, np.array(xyz))


def alkane(n):
    """All-trans zig-zag CnH2n+2 in the xz plane."""
    cc, ch = 1.54, 1.09
    tet = math.acos(-1.0 / 3.0)
    half = tet / 2.0
    dx = cc * math.sin(half)
    dz = cc * math.cos(half)
    carbons = [np.array([i * dx, 0.0, (i % 2) * dz]) for i in range(n)]
    syms, xyz = [], []
    for i, c in enumerate(carbons):
        syms.append("C")
        xyz.append(c)
    for i, c in enumerate(carbons):
        up = 1.0 if i % 2 == 0 else -1.0     # point away from the chain
        # two out-of-plane hydrogens
        bis = np.array([0.0, 0.0, -up])
        for sgn in (1.0, -1.0):
            v = math.cos(tet / 2) * bis + math.sin(tet / 2) * np.array([0.0, sgn, 0.0])
            syms.append("H")
            xyz.append(c + ch * v / np.linalg.norm(v))
        ends = []
        if i == 0:
            ends.append(-1.0)
        if i == n - 1:
            ends.append(1.0 if n > 1 else 1.0)
        if n == 1:
            # methane: complete the tetrahedron
            d = [np.array([1, 1, 1]), np.array([-1, -1, 1]),
                 np.array([-1, 1, -1]), np.array([1, -1, -1])]
            syms = ["C"] + ["H"] * 4
            xyz = [c] + [c + ch * v / np.linalg.norm(v) for v in d]
            break
        for sgn in ends:
            v = np.array([sgn * math.sin(half), 0.0, up * math.cos(half)])
            syms.append("H")
            xyz.append(c + ch * v / np.linalg.norm(v))
    return Molecule.from_angstrom(syms, np.array(xyz))


def _place(a, b, c, bond, angle, torsion):
    """Natural-extension reference frame: position of d given a, b, c."""
    angle = math.radians(angle)
    torsion = math.radians(torsion)
    bc = c - b
    bc /= np.linalg.norm(bc)
    n = np.cross(b - a, bc)
    n /= np.linalg.norm(n)
    m = np.cross(n, bc)
    d2 = np.array([-bond * math.cos(angle),
                   bond * math.sin(angle) * math.cos(torsion),
                   bond * math.sin(angle) * math.sin(torsion)])
    return c + d2[0] * bc + d2[1] * m + d2[2] * n


def polyglycine(n=10, phi=-139.0, psi=135.0, omega=180.0):
    """H-(Gly)n-OH beta strand: N-CA 1.458, CA-C 1.525, C-N 1.329, C=O 1.231,
    N-H 1.010, CA-H 1.090 A; backbone angles N-CA-C 111.2, CA-C-N 116.2,
    C-N-CA 121.7 deg."""
    nca, cac, cn, co, nh, cah, oh = 1.458, 1.525, 1.329, 1.231, 1.010, 1.090, 0.97
    ang_ncac, ang_cacn, ang_cnca = 111.2, 116.2, 121.7
    backbone = []  # (N, CA, C) per residue
    N = np.array([0.0, 0.0, 0.0])
    CA = np.array([nca, 0.0, 0.0])
    t = math.radians(180.0 - ang_ncac)
    C = CA + cac * np.array([math.cos(t), math.sin(t), 0.0])
    backbone.append((N, CA, C))
    for i in range(1, n):
        pN, pCA, pC = backbone[-1]
        Nn = _place(pN, pCA, pC, cn, ang_cacn, psi)
        CAn = _place(pCA, pC, Nn, nca, ang_cnca, omega)
        Cn = _place(pC, Nn, CAn, cac, ang_ncac, phi)
        backbone.append((Nn, CAn, Cn))
    syms, xyz = [], []
    for i, (N, CA, C) in enumerate(backbone):
        # amide H (trans to carbonyl O of previous residue); N-terminus gets two H
        if i == 0:
            prevC = _place(C, CA, N, cn, ang_cnca, phi)   # virtual previous C
        else:
            prevC = backbone[i - 1][2]
        H = _place(prevC if i > 0 else C, CA, N, nh, 119.0, 180.0 if i > 0 else 60.0)
        syms += ["N", "H"]
        xyz += [N, H]
        if i == 0:
            syms.append("H")
            xyz.append(_place(C, CA, N, nh, 109.5, -60.0))
        syms.append("C")
        xyz.append(CA)
        syms += ["H", "H"]
        xyz.append(_place(N, C, CA, cah, 109.5, 120.0))
        xyz.append(_place(N, C, CA, cah, 109.5, -120.0))
        syms.append("C")
        xyz.append(C)
        if i < n - 1:
            nextN = backbone[i + 1][0]
            syms.append("O")
            xyz.append(_place(nextN, CA, C, co, 121.0, 180.0))
        else:
            # C-terminal carboxylic acid
            psi_t = psi
            O1 = _place(N, CA, C, co, 121.0, psi_t + 180.0)
            O2 = _place(N, CA, C, 1.34, 112.0, psi_t)
            Hc = _place(CA, C, O2, oh, 107.0, 180.0)
            syms += ["O", "O", "H"]
            xyz += [O1, O2, Hc]
    return Molecule.from_angstrom(syms, np.array(xyz))


def mixed_set(count=256, seed=SEED):
    """Molecules of uneven size: water clusters of 1..16 waters and alkanes
    C1..C12, interleaved, cycling until ``count`` molecules exist."""
    out = []
    w, a = 1, 1
    k = 0
    while len(out) < count:
        if k % 2 == 0:
            out.append(water_cluster(w, seed=seed + len(out)))
            w = w % 16 + 1
        else:
            out.append(alkane(a))
            a = a % 12 + 1
        k += 1
    return out
