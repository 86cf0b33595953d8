"""Cartesian Gaussian basis sets laid out the way QUICK lays them out.

Follows QUICK's reader and normalisation (``src/basis.f90:370-720``,
``src/subs/xnorm.f90``, ``src/modules/quick_basis_module.f90:385-425``):

* shells are kept per atom in file order; an ``SP`` shell is *one* shell with
  ``Qstart=0, Qfinal=1`` that shares its exponents between an s and a p part
  (``basis.f90:476-540``);
* Cartesian components are ordered s; x, y, z; xx, xy, yy, xz, yz, zz;
  xxx, xxy, xyy, yyy, xxz, xyz, yyz, xzz, yzz, zzz (``basis.f90:567-720``
  KLMN);
* every Cartesian function is normalised to one.  The contraction
  coefficient stored per (shell, L) is the one of the axis-aligned function
  x^L (QUICK's ``gccoeff`` of function ``ksumtype+L``); mixed components carry
  the extra factor ``cons`` (sqrt(3) for xy, xz, yz; sqrt(5) for the f
  components with one exponent 1 and one 2, sqrt(15) for xyz), as QUICK's
  ``cons``.  f contractions are not renormalised (QUICK's reader does not).

Coordinates are converted from Angstrom with QUICK's BOHR = 0.52917720859
(``src/modules/quick_constants_module.f90:65``).
"""
from __future__ import annotations

import json
import math
import os
from dataclasses import dataclass, field

import numpy as np

BOHR = 0.52917720859
_DATA = os.path.join(os.path.dirname(os.path.abspath(__file__)), "basis_data")

SYMBOLS = ["X", "H", "He", "Li", "Be", "B", "C", "N", "O", "F", "Ne",
           "Na", "Mg", "Al", "Si", "P", "S", "Cl", "Ar"]
Z_OF = {s: z for z, s in enumerate(SYMBOLS)}

# keyword -> json file (aliases follow QUICK's basis/basis_link table)
_ALIASES = {
    "sto-3g": "STO-3G", "3-21g": "3-21G", "6-31g": "6-31G",
    "6-31g*": "6-31Gs", "6-31g(d)": "6-31Gs",
    "6-31g**": "6-31Gss", "6-31g(d,p)": "6-31Gss",
    "cc-pvdz": "cc-pVDZ", "def2-svp": "def2-SVP", "cc-pvtz": "cc-pVTZ", "def2-tzvp": "def2-TZVP",
}

CART = {
    0: [(0, 0, 0)],
    1: [(1, 0, 0), (0, 1, 0), (0, 0, 1)],
    2: [(2, 0, 0), (1, 1, 0), (0, 2, 0), (1, 0, 1), (0, 1, 1), (0, 0, 2)],
    3: [(3, 0, 0), (2, 1, 0), (1, 2, 0), (0, 3, 0), (2, 0, 1), (1, 1, 1), (0, 2, 1), (1, 0, 2), (0, 1, 2),
        (0, 0, 3)],
}
_S3, _S5 = math.sqrt(3.0), math.sqrt(5.0)
CONS = {0: [1.0], 1: [1.0, 1.0, 1.0],
        2: [1.0, _S3, 1.0, _S3, _S3, 1.0],
        3: [1.0, _S5, _S5, 1.0, _S5, _S5 * _S3, _S5, _S5, _S5, 1.0]}
_SHELL_L = {"S": (0, 0), "P": (1, 1), "SP": (0, 1), "D": (2, 2), "F": (3, 3)}


def _dfact(n):
    r = 1.0
    while n > 1:
        r *= n
        n -= 2
    return r


def normalized_coefficients(exps, coefs, L):
    """Coefficients of the normalised x^L contracted function.

    Primitive factor (2a)^(L/2+3/4) pi^(-3/4) ((2L-1)!!/2^L)^(-1/2), then,
    for s, p and d, the contraction is rescaled to unit self-overlap
    (``xnorm`` and ``xnewnorm``).  QUICK's reader applies no contraction
    rescale to f shells (basis.f90:652-720), so neither does this.
    """
    a = np.asarray(exps, dtype=np.float64)
    d = np.asarray(coefs, dtype=np.float64)
    g = _dfact(2 * L - 1) / (2.0 ** L)
    c = d * (2.0 * a) ** (L / 2.0 + 0.75) * math.pi ** (-0.75) * g ** (-0.5)
    if L >= 3:
        return c
    s = np.sum(np.outer(c, c) * (a[:, None] + a[None, :]) ** (-1.5 - L))
    s = (s * g * math.pi ** 1.5) ** (-0.5)
    return c * s


@dataclass
class Shell:
    atom: int            # 0-based atom index
    lmin: int            # QUICK Qstart
    lmax: int            # QUICK Qfinal
    exps: np.ndarray     # (nprim,)
    coef: dict           # L -> normalised coefficients (nprim,)
    first_bf: int        # 0-based index of the first basis function
    prim_start: int      # 0-based index into the global primitive list

    @property
    def nprim(self):
        return len(self.exps)

    def nfunc(self):
        return sum(len(CART[L]) for L in range(self.lmin, self.lmax + 1))

    def bf_offset(self, L):
        """Offset of the first function of angular momentum L inside the shell."""
        return sum(len(CART[l]) for l in range(self.lmin, L))


@dataclass
class Molecule:
    symbols: list
    xyz: np.ndarray          # (natom, 3) bohr
    charge: int = 0

    @property
    def Z(self):
        return np.array([Z_OF[s] for s in self.symbols], dtype=np.float64)

    @property
    def natom(self):
        return len(self.symbols)

    def nelec(self):
        return int(round(self.Z.sum())) - self.charge

    def nuclear_repulsion(self):
        Z = self.Z
        e = 0.0
        for i in range(self.natom):
            d = np.linalg.norm(self.xyz[i + 1:] - self.xyz[i], axis=1)
            e += float(np.sum(Z[i] * Z[i + 1:] / d))
        return e

    @staticmethod
    def from_angstrom(symbols, xyz_ang, charge=0):
        return Molecule(list(symbols), np.asarray(xyz_ang, dtype=np.float64) / BOHR, charge)


@dataclass
class BasisSet:
    name: str
    mol: Molecule
    shells: list = field(default_factory=list)
    nbasis: int = 0
    nprim: int = 0

    # --- QUICK-style flat arrays -------------------------------------------------
    def shell_arrays(self):
        """Flat arrays describing shells and primitives (QUICK's kstart, kprim,
        katom, Qstart, Qfinal, ksumtype, gcexpo and gccoeff per L)."""
        ns = len(self.shells)
        atom = np.array([s.atom for s in self.shells], dtype=np.int32)
        lmin = np.array([s.lmin for s in self.shells], dtype=np.int32)
        lmax = np.array([s.lmax for s in self.shells], dtype=np.int32)
        nprim = np.array([s.nprim for s in self.shells], dtype=np.int32)
        pstart = np.array([s.prim_start for s in self.shells], dtype=np.int32)
        fstart = np.array([s.first_bf for s in self.shells], dtype=np.int32)
        exps = np.zeros(self.nprim, dtype=np.float64)
        coef = np.zeros((self.nprim, 4), dtype=np.float64)   # [prim][L]
        for s in self.shells:
            exps[s.prim_start:s.prim_start + s.nprim] = s.exps
            for L, c in s.coef.items():
                coef[s.prim_start:s.prim_start + s.nprim, L] = c
        assert ns == len(atom)
        return dict(atom=atom, lmin=lmin, lmax=lmax, nprim=nprim, pstart=pstart,
                    fstart=fstart, exps=exps, coef=coef)

    def function_shell(self):
        """Shell index of every basis function."""
        out = np.zeros(self.nbasis, dtype=np.int32)
        for i, s in enumerate(self.shells):
            out[s.first_bf:s.first_bf + s.nfunc()] = i
        return out


def _load_json(name):
    key = _ALIASES.get(name.lower(), name)
    path = os.path.join(_DATA, key + ".json")
    if not os.path.exists(path):
        raise ValueError(f"unknown basis set {name!r}")
    with open(path) as fh:
        return json.load(fh)


def build_basis(mol: Molecule, name: str) -> BasisSet:
    doc = _load_json(name)
    bs = BasisSet(name=doc["name"], mol=mol)
    nbf = 0
    npr = 0
    for ia, sym in enumerate(mol.symbols):
        shells = doc["elements"].get(sym)
        if shells is None:
            raise ValueError(f"basis {name} has no entry for {sym}")
        for sh in shells:
            if sh["type"] not in _SHELL_L:
                raise ValueError(f"shell type {sh['type']} is beyond f (not supported)")
            lmin, lmax = _SHELL_L[sh["type"]]
            coef = {}
            for k, L in enumerate(range(lmin, lmax + 1)):
                coef[L] = normalized_coefficients(sh["exps"], sh["coefs"][k], L)
            s = Shell(atom=ia, lmin=lmin, lmax=lmax,
                      exps=np.asarray(sh["exps"], dtype=np.float64),
                      coef=coef, first_bf=nbf, prim_start=npr)
            bs.shells.append(s)
            nbf += s.nfunc()
            npr += s.nprim
    bs.nbasis = nbf
    bs.nprim = npr
    return bs
