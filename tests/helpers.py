"""Shared fixtures for the parity tests: one geometry + basis seen both by the
product (quick_amd.basis) and by the oracle (QUICK's layout restated)."""
import ctypes
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
    if name.endswith("+") and name != "NH4+":
        m = molecule(name[:-1])
        return B.Molecule(m.symbols, m.xyz, charge=m.charge + 1)
    if name == "water":
        return M.water()
    if name.startswith("water"):
        return M.water_cluster(int(name[5:]))
    if name.startswith("alkane"):
        return M.alkane(int(name[6:]))
    if name.startswith("gly"):
        return M.polyglycine(int(name[3:]))
    if name == "c2h5":
        # ethyl radical at the input geometry of QUICK's opt_c2h5rad_uhf_631gss test
        return B.Molecule.from_angstrom(
            ["C", "C", "H", "H", "H", "H", "H"],
            [[0.0, 0.0, 0.769841], [0.0, 0.0, -0.769841], [0.889465, 0.513533, 1.134849],
             [0.889465, -0.513533, -1.134849], [-0.889465, 0.513533, 1.134849],
             [-0.889465, -0.513533, -1.134849], [0.0, -1.027065, 1.134849]])
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


def seeded_spin_densities_for(orc, mol, mult):
    """Pa, Pb: half of seeded_density with n_alpha / n_beta occupied
    orbitals (different seeds), so tr(Pa S) = n_alpha, tr(Pb S) = n_beta."""
    S, _, _ = orc.one_electron()
    nel = mol.nelec()
    na = (nel + mult - 1) // 2
    return 0.5 * O.seeded_density(S, na), 0.5 * O.seeded_density(S, nel - na, seed=20261203)


def quick_abi_setup(lib, mol, orc, oshell=False, mult=1):
    """The per-geometry gpu_*_ calls as QUICK's Fortran makes them (main.f90:
    99-204): 1-based shell arrays, gccoeff/gcexpo(MAXPRIM=14, nbasis)
    column-major, YCutoff/cutPrim from QUICK's schwarzoff.  Returns the
    (pointer, double-by-reference) helpers for the per-iteration calls."""
    qb = orc.qb
    n, ns, jb = qb.nbasis, qb.nshell, qb.jbasis
    I = lambda v: ctypes.byref(ctypes.c_int(int(v)))
    D = lambda v: ctypes.byref(ctypes.c_double(float(v)))
    vp = lambda a: a.ctypes.data_as(ctypes.c_void_p)
    ierr = ctypes.c_int(0)
    lib.gpu_new_(ctypes.byref(ierr))
    lib.gpu_set_device_(I(0), ctypes.byref(ierr))
    lib.gpu_init_device_(ctypes.byref(ierr))
    assert ierr.value == 0
    f = ctypes.c_bool(bool(oshell))
    lib.gpu_upload_method_(I(0), ctypes.byref(f), D(1.0))
    lib.gpu_setup_(I(len(mol.symbols)), I(n), I(mol.nelec()), I(mult), I(mol.charge), I(2))
    xyz = np.ascontiguousarray(qb.xyz)      # xyz(3, natom) column-major == [atom][3]
    lib.gpu_upload_xyz_(vp(xyz))
    MAXPRIM = 14
    gccoeff = np.zeros((n, MAXPRIM))        # Fortran (MAXPRIM, nbasis) == C [nbasis][MAXPRIM]
    gcexpo = np.zeros((n, MAXPRIM))
    for s in range(ns):
        f0 = qb.ksumtype[s]
        nf = sum((L + 1) * (L + 2) // 2 for L in range(qb.qstart[s], qb.qfinal[s] + 1))
        p0, k = qb.kstart[s], qb.kprim[s]
        for fi in range(f0, f0 + nf):
            gcexpo[fi, :k] = qb.gcexpo[p0:p0 + k]
        for L in range(qb.qstart[s], qb.qfinal[s] + 1):
            gccoeff[f0 + L, :k] = qb.gccoeff[p0:p0 + k, L]
    i32 = lambda a: np.ascontiguousarray(a, dtype=np.int32)
    kstart, katom = i32(qb.kstart + 1), i32(qb.katom + 1)
    kprim, ksum = i32(qb.kprim), i32(qb.ksumtype + 1)
    qs, qf = i32(qb.qstart), i32(qb.qfinal)
    dummy_i = np.zeros(max(n, ns) * 4 + 8, dtype=np.int32)
    dummy_d = np.zeros(max(n, 1) * 4, dtype=np.float64)
    lib.gpu_upload_basis_.restype = None
    lib.gpu_upload_basis_(I(ns), I(jb), I(ns), I(jb), I(int(qb.kprim.max())), vp(dummy_i), vp(dummy_i),
                          vp(dummy_d), vp(dummy_d), vp(dummy_i), vp(dummy_i), vp(dummy_i), vp(dummy_i),
                          vp(dummy_i), vp(kstart), vp(katom), vp(dummy_i), vp(kprim), vp(dummy_i), vp(ksum),
                          vp(dummy_i), vp(qs), vp(qf), vp(dummy_i), vp(dummy_i), vp(gccoeff), vp(dummy_d),
                          vp(gcexpo), vp(dummy_i))
    Y, CP = orc.schwarz()
    Yf, CPf = np.asfortranarray(Y), np.asfortranarray(CP)
    lib.gpu_upload_cutoff_matrix_(vp(Yf), vp(CPf))
    return vp, D


def md_schwarz(qb, a, b):
    """sqrt(max_components |(ab|ab)|) of oracle shells a, b from the
    McMurchie-Davidson code with QUICK's FmT (maxm = 2 (La + Lb), as the
    Schwarz diagonals use): a value independent of both HRR formulations."""
    from tests import md_reference as MD

    def desc(s):
        out = []
        for L in range(qb.qstart[s], qb.qfinal[s] + 1):
            p0, n = qb.kstart[s], qb.kprim[s]
            out.append((L, qb.xyz[qb.katom[s]], qb.gcexpo[p0:p0 + n], qb.gccoeff[p0:p0 + n, L]))
        return out

    m = 0.0
    for sa in desc(a):
        for sb in desc(b):
            maxm = 2 * (qb.qfinal[a] + qb.qfinal[b])
            blk = MD.shell_block((sa, sb, sa, sb), lambda L, T, _m=maxm: O.fmt(_m, T)[:L + 1])
            for i in range(blk.shape[0]):
                for j in range(blk.shape[1]):
                    m = max(m, abs(blk[i, j, i, j]))
    return float(np.sqrt(m))
