"""Shared fixtures for the parity tests: one geometry + basis seen both by the
product (quick_amd.basis) and by the oracle (QUICK's layout restated)."""
import sys

import numpy as np

from oracle import oracle as O
from quick_amd import basis as B
from quick_amd import molecules as M


def announce(name):
    print(f"[case] {name}", flush=True)
    sys.stdout.flush()


def oracle_for(mol, basis_name):
    qb = O.QuickBasisRef(mol.symbols, mol.xyz, basis_name)
    return O.Oracle(qb)


def product_basis(mol, basis_name):
    return B.build_basis(mol, basis_name)


def molecule(name):
    if name == "water":
        return M.water()
    if name.startswith("water"):
        return M.water_cluster(int(name[5:]))
    if name.startswith("alkane"):
        return M.alkane(int(name[6:]))
    if name.startswith("gly"):
        return M.polyglycine(int(name[3:]))
    if name == "LiF":
        return B.Molecule.from_angstrom(["Li", "F"], [[0, 0, 0], [0, 0, 1.5639]])
    if name == "NH4+":
        d = 1.03 / np.sqrt(3)
        return B.Molecule.from_angstrom(["N", "H", "H", "H", "H"],
                                        [[0, 0, 0], [d, d, d], [-d, -d, d], [-d, d, -d], [d, -d, -d]], charge=1)
    raise KeyError(name)


def seeded_density_for(orc, mol):
    S, _, _ = orc.one_electron()
    return O.seeded_density(S, mol.nelec() // 2), S
