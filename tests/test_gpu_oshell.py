"""Open-shell (UHF) Fock build on the GPU against the oracle's getOshellEri.

Bar: Ga = J(Pa+Pb) - K(Pa) and Gb = J(Pa+Pb) - K(Pb) within max |dF| <=
1e-10 Eh of QUICK's CPU open-shell build restated (same screening: the
open-shell density matrix max(|Pa+Pb|, |Pa|, |Pb|)); UHF energies within
1e-8 Eh of the oracle's UHF and 3e-7 Eh of QUICK's saved output.
"""
import ctypes

import numpy as np
import pytest

from tests.helpers import (announce, molecule, oracle_for, product_basis, quick_abi_setup, seeded_density_for,
                           seeded_spin_densities_for)

pytestmark = pytest.mark.gpu

FOCK_TOL = 1e-10

CASES = [
    ("c2h5", "6-31G**", 2, 1e-10),
    ("water+", "6-31G*", 2, 1e-12),
    ("water3+", "6-31G*", 2, 1e-12),
    ("gly1", "6-31G*", 3, 1e-11),
]


def _engine(bs, **kw):
    from quick_amd.engine import FockEngine
    return FockEngine(bs, **kw)


def _set_mode(eng, mode):
    if mode == "direct":
        eng.set_tile(False)
    elif mode == "coop":
        for cb in range(6):
            for ck in range(cb + 1):
                eng.set_class_tuning(cb, ck, 2, 0, 0, 0)
    elif mode == "register":
        # the fused form with the second exchange's per-task terms in registers
        for cb in range(6):
            for ck in range(cb + 1):
                eng.set_class_os(cb, ck, 0)
    elif mode.startswith("two-pass"):
        # every class as two closed-shell launches (exchange-only second pass)
        for cb in range(6):
            for ck in range(cb + 1):
                eng.set_class_os(cb, ck, 1)
                if mode == "two-pass-direct":
                    eng.set_class_tuning(cb, ck, 0, 0, 0, 0)


@pytest.mark.parametrize("mode", ["default", "register", "direct", "coop", "two-pass", "two-pass-direct"])
@pytest.mark.parametrize("name,basis,mult,cutoff", CASES)
def test_open_shell_fock_matches_oracle(name, basis, mult, cutoff, mode):
    announce(f"open-shell fock {name} {basis} mult={mult} {mode}")
    mol = molecule(name)
    orc = oracle_for(mol, basis)
    Pa, Pb = seeded_spin_densities_for(orc, mol, mult)
    Ga_ref, Gb_ref, nq = orc.fock2e_os(Pa, Pb, cutoff=cutoff, threads=4)
    eng = _engine(product_basis(mol, basis), cutoff=cutoff)
    _set_mode(eng, mode)
    Ga, Gb = eng.fock2e_os(Pa, Pb)
    ea, eb = float(np.max(np.abs(Ga - Ga_ref))), float(np.max(np.abs(Gb - Gb_ref)))
    print(f"  max|dFa| = {ea:.3e}  max|dFb| = {eb:.3e}  (quartets {nq})", flush=True)
    assert ea <= FOCK_TOL and eb <= FOCK_TOL
    assert np.max(np.abs(Ga - Ga.T)) <= 1e-14 * max(1.0, np.max(np.abs(Ga)))


@pytest.mark.parametrize("name,basis", [("water+", "cc-pVTZ"), ("water+", "def2-TZVP")])
def test_open_shell_f_shells_match_oracle(name, basis):
    """f classes in the open-shell pass: the wave-per-quartet kernel reloads
    the exchange density blocks from P_a - P_b (the (ff|..) classes keep
    their node-program slots in global memory), fused, fused lean and
    two-pass forms."""
    announce(f"open-shell f shells {name} {basis}")
    mol = molecule(name)
    orc = oracle_for(mol, basis)
    Pa, Pb = seeded_spin_densities_for(orc, mol, 2)
    Ga_ref, Gb_ref, nq = orc.fock2e_os(Pa, Pb, cutoff=1e-11, threads=4)
    for form in (2, 0, 1):   # fused lean (default), fused, two passes
        eng = _engine(product_basis(mol, basis), cutoff=1e-11)
        for cb in range(10):
            for ck in range(cb + 1):
                eng.set_class_os(cb, ck, form)
        Ga, Gb = eng.fock2e_os(Pa, Pb)
        ea, eb = float(np.max(np.abs(Ga - Ga_ref))), float(np.max(np.abs(Gb - Gb_ref)))
        print(f"  form {form}: max|dFa| = {ea:.3e}  max|dFb| = {eb:.3e}  (quartets {nq})", flush=True)
        assert ea <= FOCK_TOL and eb <= FOCK_TOL


def test_open_shell_closed_shell_limit():
    """Pa = Pb = P/2 gives Ga = Gb = the closed-shell G(P)."""
    announce("open-shell closed-shell limit water3 6-31G*")
    mol = molecule("water3")
    orc = oracle_for(mol, "6-31G*")
    P, _ = seeded_density_for(orc, mol)
    eng = _engine(product_basis(mol, "6-31G*"), cutoff=1e-12)
    G = eng.fock2e(P)
    Ga, Gb = eng.fock2e_os(0.5 * P, 0.5 * P)
    assert np.max(np.abs(Ga - G)) <= 1e-12 and np.max(np.abs(Gb - G)) <= 1e-12


def test_open_shell_mixed_precision_error_is_bounded():
    announce("open-shell mixed precision water8+ 6-31G*")
    mol = molecule("water8+")
    orc = oracle_for(mol, "6-31G*")
    Pa, Pb = seeded_spin_densities_for(orc, mol, 2)
    bs = product_basis(mol, "6-31G*")
    Ga64, Gb64 = _engine(bs, cutoff=1e-12).fock2e_os(Pa, Pb)
    emx = _engine(bs, cutoff=1e-12, precision="mixed", split=1e-4)
    Gamx, Gbmx = emx.fock2e_os(Pa, Pb)
    assert emx.stats()["n_tasks_fp32"] > 0
    dF = max(float(np.max(np.abs(Gamx - Ga64))), float(np.max(np.abs(Gbmx - Gb64))))
    dE = 0.5 * abs(float(np.sum(Pa * (Gamx - Ga64)) + np.sum(Pb * (Gbmx - Gb64))))
    print(f"  mixed: max|dF| {dF:.2e}  |dE2| {dE:.2e} Eh", flush=True)
    assert dF < 1e-6 and dE < 1e-6


def test_uhf_energy_matches_oracle_and_quick_saved():
    """UHF of the ethyl radical (6-31G**, doublet) at the geometry of QUICK's
    opt_c2h5rad_uhf_631gss test: the GPU build in the UHF driver against the
    oracle's UHF (1e-8 Eh) and QUICK's printed first-step energy
    -78.593961730 Eh (3e-7 Eh).  The incremental (deltaO) SCF lands on the
    same energy."""
    announce("UHF c2h5 6-31G**")
    from oracle import oracle as O
    from quick_amd.scf import uhf
    mol = molecule("c2h5")
    orc = oracle_for(mol, "6-31G**")
    nel = mol.nelec()
    E_ref, _, _ = O.uhf(orc, (nel + 1) // 2, (nel - 1) // 2, cutoff=1e-10, prim_limit=1e-10)
    eng = _engine(product_basis(mol, "6-31G**"), cutoff=1e-10, prim_limit=1e-10)
    res = uhf(eng, mult=2, etol=1e-10, dtol=1e-8)
    print(f"  E_gpu {res.energy:.10f}  E_oracle {E_ref:.10f}  it {res.iterations}", flush=True)
    assert res.converged
    assert abs(res.energy - E_ref) < 1e-8
    assert abs(res.energy - (-78.593961730)) < 3e-7
    res2 = uhf(eng, mult=2, etol=1e-10, dtol=1e-8, incremental=True)
    assert res2.converged and abs(res2.energy - res.energy) < 1e-8


def test_quick_fortran_abi_open_shell():
    """gpu_upload_method_(is_oshell), gpu_upload_calculated_ (alpha density),
    gpu_upload_calculated_beta_ and gpu_get_oshell_eri_ as QUICK's
    uscf_operator calls them (quick_uscf_operator_module.f90:96-128): o and
    ob are added to."""
    announce("QUICK Fortran ABI open shell water+ 6-31G*")
    from quick_amd import _lib
    lib = _lib.lib()
    mol = molecule("water+")
    orc = oracle_for(mol, "6-31G*")
    n = orc.qb.nbasis
    Pa, Pb = seeded_spin_densities_for(orc, mol, 2)
    Ga_ref, Gb_ref, _ = orc.fock2e_os(Pa, Pb, cutoff=1e-12)
    vp, D = quick_abi_setup(lib, mol, orc, oshell=True, mult=2)
    dummy = np.zeros(4 * n)
    o = np.asfortranarray(np.full((n, n), 0.25))
    ob = np.asfortranarray(np.full((n, n), -0.5))
    dense, denseb = np.asfortranarray(Pa), np.asfortranarray(Pb)
    lib.gpu_upload_calculated_(vp(o), vp(dummy), vp(dummy), vp(dense))
    cm = np.asfortranarray(orc.density_cutoff_os(Pa, Pb))
    lib.gpu_upload_cutoff_(vp(cm), D(1e-12), D(1e-13), D(1e-10), D(1e-12))
    lib.gpu_upload_calculated_beta_(vp(ob), vp(denseb))
    f = ctypes.c_bool(False)
    lib.gpu_get_oshell_eri_(ctypes.byref(f), vp(o), vp(ob))
    ierr = ctypes.c_int(0)
    lib.gpu_delete_(ctypes.byref(ierr))
    assert np.max(np.abs((o - 0.25) - Ga_ref)) <= FOCK_TOL
    assert np.max(np.abs((ob + 0.5) - Gb_ref)) <= FOCK_TOL
