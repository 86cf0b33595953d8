"""Edge cases of the Fock build: empty work, a single atom, only s shells,
rejected angular momentum (above f), and a cluster large enough that some LDS tiles
do not fit (size-independent properties there: the digestion paths agree,
G is linear in P and symmetric)."""
import ctypes

import numpy as np
import pytest

from tests.helpers import announce, molecule, oracle_for, product_basis, seeded_density_for

pytestmark = pytest.mark.gpu


def _engine(bs, **kw):
    from quick_amd.engine import FockEngine
    return FockEngine(bs, **kw)


def test_zero_density_and_everything_screened_give_zero():
    announce("edge: zero density, cutoff 1")
    mol = molecule("water3")
    bs = product_basis(mol, "6-31G*")
    e = _engine(bs, cutoff=1e-12)
    n = bs.nbasis
    assert np.all(e.fock2e(np.zeros((n, n))) == 0.0)
    e2 = _engine(bs, cutoff=1.0e3)           # no pair survives
    S, _, _ = e2.one_electron()
    from quick_amd.scf import seeded_density
    P = seeded_density(S, mol.nelec() // 2)
    assert np.all(e2.fock2e(P) == 0.0)
    assert e2.stats()["n_tasks_fp64"] == 0


@pytest.mark.parametrize("name,symbols,xyz,basis", [
    ("He atom", ["He"], [[0.0, 0.0, 0.0]], "6-31G"),
    ("H2 STO-3G", ["H", "H"], [[0.0, 0.0, 0.0], [0.0, 0.0, 0.7408481]], "STO-3G"),
    ("Ne atom cc-pVDZ", ["Ne"], [[0.0, 0.0, 0.0]], "cc-pVDZ"),
])
def test_small_systems_match_oracle(name, symbols, xyz, basis):
    announce(f"edge: {name}")
    from quick_amd import basis as B
    mol = B.Molecule.from_angstrom(symbols, xyz)
    orc = oracle_for(mol, basis)
    P, _ = seeded_density_for(orc, mol)
    G_ref, _ = orc.fock2e(P, cutoff=1e-12)
    G = _engine(product_basis(mol, basis), cutoff=1e-12).fock2e(P)
    assert np.max(np.abs(G - G_ref)) <= 1e-10


def test_g_shells_are_rejected():
    """QUICK stops at f (its reader sets ierr 37 for G shells, basis.f90);
    the engine refuses L = 4 with QF_ERR_ARG and accepts L = 3."""
    announce("edge: g shell rejected, f accepted")
    from quick_amd import _lib
    lib = _lib.lib()
    h = ctypes.c_void_p()
    _lib.check(lib.qf_engine_create(0, ctypes.byref(h)), "create")
    xyz = np.zeros(3)
    _lib.check(lib.qf_set_molecule(h, 1, _lib.dptr(xyz)), "molecule")
    i32 = lambda v: np.ascontiguousarray(v, dtype=np.int32)
    rc = lib.qf_set_basis(h, 1, _lib.iptr(i32([0])), _lib.iptr(i32([4])), _lib.iptr(i32([4])), _lib.iptr(i32([1])),
                          _lib.iptr(i32([0])), _lib.iptr(i32([0])), 1, _lib.dptr(np.ones(1)),
                          _lib.dptr(np.ones(4)), 15)
    assert rc == _lib.QF_OK + 1      # QF_ERR_ARG
    assert b"above f" in lib.qf_last_error()
    rc = lib.qf_set_basis(h, 1, _lib.iptr(i32([0])), _lib.iptr(i32([3])), _lib.iptr(i32([3])), _lib.iptr(i32([1])),
                          _lib.iptr(i32([0])), _lib.iptr(i32([0])), 1, _lib.dptr(np.ones(1)),
                          _lib.dptr(np.ones(4)), 10)
    assert rc == _lib.QF_OK
    lib.qf_engine_destroy(h)


def test_large_cluster_paths_agree_and_properties_hold():
    """(H2O)48 6-31G(d), 912 functions: the ss/sp blocks' compact column
    spaces outgrow a small LDS budget, so tile and direct kernels mix; all
    digestion paths give the same G, which is linear in P and symmetric."""
    announce("edge: (H2O)48 6-31G* paths and linearity")
    mol = molecule("water48")
    bs = product_basis(mol, "6-31G*")
    e = _engine(bs, cutoff=1e-11)
    S, _, _ = e.one_electron()
    from quick_amd.scf import seeded_density
    P1 = seeded_density(S, mol.nelec() // 2, seed=3)
    P2 = seeded_density(S, mol.nelec() // 2, seed=4)
    G_tile = e.fock2e(P1)
    e.set_tile(True, waves=4, tasks_per_group=16, lds_bytes=8192)
    G_small = e.fock2e(P1)
    e.set_tile(False)
    G_direct = e.fock2e(P1)
    for cb in range(6):
        for ck in range(cb + 1):
            e.set_class_tuning(cb, ck, 2)
    G_coop = e.fock2e(P1)
    scale = np.max(np.abs(G_tile))
    for G in (G_small, G_direct, G_coop):
        assert np.max(np.abs(G - G_tile)) <= 1e-12 * scale
    lin = _engine(bs, cutoff=1e-30, prim_limit=1e-31)
    Ga, Gb, Gab = lin.fock2e(P1), lin.fock2e(P2), lin.fock2e(P1 + P2)
    assert np.max(np.abs(Ga + Gb - Gab)) <= 1e-12 * np.max(np.abs(Gab))
    assert np.max(np.abs(Gab - Gab.T)) <= 1e-14 * np.max(np.abs(Gab))
