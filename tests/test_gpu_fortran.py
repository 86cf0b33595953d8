"""A Fortran program (tests/fortran/quick_driver.f90, built with amdflang
against libquickfock.so) makes QUICK's calls with Fortran's own argument
passing: implicit interfaces, everything by reference, LOGICAL flags,
column-major arrays.  Its Fock matrix and two-electron gradient must equal
the oracle's."""
import os
import subprocess

import numpy as np
import pytest

from tests.helpers import announce, molecule, oracle_for, seeded_density_for

pytestmark = pytest.mark.gpu

DRIVER = os.path.join(os.path.dirname(os.path.abspath(__file__)), "fortran", "quick_driver")


def _write_input(path, mol, orc, P, cutoff, prim_limit, grad_cutoff):
    qb = orc.qb
    n, ns, jb = qb.nbasis, qb.nshell, qb.jbasis
    MAXPRIM = 14
    gccoeff = np.zeros((n, MAXPRIM))        # Fortran (MAXPRIM, nbasis): row-major [nbasis][MAXPRIM]
    gcexpo = np.zeros((n, MAXPRIM))
    for s in range(ns):
        f0 = qb.ksumtype[s]
        nf = sum((L + 1) * (L + 2) // 2 for L in range(qb.qstart[s], qb.qfinal[s] + 1))
        p0, k = qb.kstart[s], qb.kprim[s]
        for fi in range(f0, f0 + nf):
            gcexpo[fi, :k] = qb.gcexpo[p0:p0 + k]
        for L in range(qb.qstart[s], qb.qfinal[s] + 1):
            gccoeff[f0 + L, :k] = qb.gccoeff[p0:p0 + k, L]
    Y, CP = orc.schwarz()
    cm = orc.density_cutoff(P)
    fmt = lambda a: "\n".join("%.17e" % v for v in np.ravel(a))
    with open(path, "w") as fh:
        fh.write("%d %d %d %d %d\n" % (len(mol.symbols), n, ns, jb, mol.nelec()))
        fh.write(fmt(qb.xyz) + "\n")                      # xyz(3, natom): row-major [atom][3]
        fh.write(fmt(qb.charge) + "\n")
        for s in range(ns):
            fh.write("%d %d %d %d %d %d\n" % (qb.kstart[s] + 1, qb.katom[s] + 1, qb.kprim[s], qb.ksumtype[s] + 1,
                                              qb.qstart[s], qb.qfinal[s]))
        fh.write(fmt(gccoeff) + "\n")
        fh.write(fmt(gcexpo) + "\n")
        fh.write(fmt(Y.T) + "\n")                         # column-major: Fortran reads the transpose
        fh.write(fmt(CP.T) + "\n")
        fh.write(fmt(P.T) + "\n")
        fh.write(fmt(cm.T) + "\n")
        fh.write("%.17e %.17e %.17e\n" % (cutoff, prim_limit, grad_cutoff))


@pytest.mark.skipif(not os.path.exists(DRIVER), reason="amdflang driver not built (no amdflang in the image)")
def test_fortran_caller_fock_and_gradient(tmp_path):
    announce("Fortran driver through QUICK's entry points: water3 6-31G*")
    mol = molecule("water3")
    orc = oracle_for(mol, "6-31G*")
    P, _ = seeded_density_for(orc, mol)
    inp, out = str(tmp_path / "in.txt"), str(tmp_path / "out.txt")
    _write_input(inp, mol, orc, P, 1e-11, 1e-12, 1e-10)
    r = subprocess.run([DRIVER, inp, out], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    vals = np.loadtxt(out)
    n = orc.n
    o = vals[:n * n].reshape(n, n).T - 0.25            # column order
    grad = vals[n * n:].reshape(-1, 3) - 0.125
    G_ref, _ = orc.fock2e(P, cutoff=1e-11, prim_limit=1e-12, threads=8)
    g_ref, _ = orc.eri_grad(P, grad_cutoff=1e-10, prim_limit=1e-12, threads=8)
    dF = float(np.max(np.abs(o - G_ref)))
    dg = float(np.max(np.abs(grad - g_ref))) / max(1.0, float(np.max(np.abs(g_ref))))
    print(f"  max|dF| {dF:.2e}, max|dg|/max(1,|g|) {dg:.2e}", flush=True)
    assert dF <= 1e-10
    assert dg <= 1e-10
