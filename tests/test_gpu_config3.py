"""Split-precision Fock builds on the polyglycine workload (beta strand,
6-31G(d,p) Cartesian): quartets whose Schwarz product is below 1e-4 of the
largest run their VRR/HRR in FP32, the rest in FP64; Boys function,
prefactors, digestion and accumulation stay FP64.  Every result is stated
in Eh against the FP64 oracle: H-(Gly)n-OH at sizes the oracle builds in
seconds, and the full (Gly)10 strand by properties against the FP64 GPU
build (itself held to the oracle at 1e-10 on the small strands)."""
import numpy as np
import pytest

from tests.helpers import announce, molecule, oracle_for, product_basis, seeded_density_for

pytestmark = pytest.mark.gpu

BASIS = "6-31G**"


def _engine(bs, **kw):
    from quick_amd.engine import FockEngine
    return FockEngine(bs, **kw)


@pytest.mark.parametrize("n", [2, 3])
def test_polyglycine_mixed_fock_vs_oracle(n):
    announce(f"config3 (Gly){n} 6-31G** mixed and FP64 vs oracle")
    mol = molecule(f"gly{n}")
    orc = oracle_for(mol, BASIS)
    P, _ = seeded_density_for(orc, mol)
    G_ref, nq = orc.fock2e(P, cutoff=1e-11, threads=8)
    bs = product_basis(mol, BASIS)
    e64 = _engine(bs, cutoff=1e-11)
    G64 = e64.fock2e(P)
    emx = _engine(bs, cutoff=1e-11, precision="mixed", split=1e-4)
    Gmx = emx.fock2e(P)
    st = emx.stats()
    d64 = float(np.max(np.abs(G64 - G_ref)))
    dF = float(np.max(np.abs(Gmx - G_ref)))
    dE = 0.5 * abs(float(np.sum(P * (Gmx - G_ref))))
    print(f"  nbasis {bs.nbasis}, {nq} quartets, FP32 tasks {st['n_tasks_fp32']}: FP64 max|dF| {d64:.2e}; "
          f"mixed max|dF| {dF:.2e} Eh, |dE2| {dE:.2e} Eh", flush=True)
    assert st["n_tasks_fp32"] > 0
    assert d64 <= 1e-10
    assert dF <= 1e-6 and dE <= 1e-6


def test_per_class_split_choice():
    """qf_set_class_split: with every class forced to FP64 a mixed engine
    builds the FP64 G exactly (no FP32 tasks); with the split alternating over
    the classes it stays within the mixed bar of the oracle."""
    from quick_amd.autotune import _class_index
    announce("config3 (Gly)2 6-31G** per-class split choice")
    mol = molecule("gly2")
    orc = oracle_for(mol, BASIS)
    P, _ = seeded_density_for(orc, mol)
    G_ref, _ = orc.fock2e(P, cutoff=1e-11, threads=8)
    bs = product_basis(mol, BASIS)
    G64 = _engine(bs, cutoff=1e-11).fock2e(P)
    emx = _engine(bs, cutoff=1e-11, precision="mixed", split=1e-4)
    emx.set_split_classes(0)
    names = sorted({nm for nm, _, _ in emx.launches()})
    for nm in names:
        emx.set_class_split(*_class_index(nm), 0)
    G0 = emx.fock2e(P)
    assert emx.stats()["n_tasks_fp32"] == 0
    assert float(np.max(np.abs(G0 - G64))) <= 1e-13
    for k, nm in enumerate(names):
        emx.set_class_split(*_class_index(nm), k % 2)
    Gm = emx.fock2e(P)
    st = emx.stats()
    dF = float(np.max(np.abs(Gm - G_ref)))
    print(f"  FP32 tasks {st['n_tasks_fp32']}, max|dF| {dF:.2e} Eh", flush=True)
    assert st["n_tasks_fp32"] > 0 and dF <= 1e-6
    assert {f for _, f, _ in emx.launches()} == {False, True}


def test_polyglycine_mixed_scf_vs_oracle_scf():
    """H-(Gly)2-OH: SCF with mixed-precision GPU builds against the oracle's
    FP64 SCF, |dE_SCF| <= 1e-6 Eh (and the FP64 GPU SCF to 1e-8)."""
    announce("config3 (Gly)2 6-31G** SCF mixed/FP64 vs oracle SCF")
    from oracle import oracle as O
    from quick_amd.scf import rhf
    mol = molecule("gly2")
    orc = oracle_for(mol, BASIS)
    E_ref, _, _ = O.rhf(orc, mol.nelec(), cutoff=1e-11, prim_limit=1e-12, etol=1e-10, dtol=1e-8, threads=8)
    bs = product_basis(mol, BASIS)
    out = {}
    for prec in ("fp64", "mixed"):
        e = _engine(bs, cutoff=1e-11, prim_limit=1e-12, precision=prec)
        r = rhf(e, etol=1e-10, dtol=1e-8)
        assert r.converged
        out[prec] = r.energy
    print(f"  E_oracle {E_ref:.10f}  E_fp64 {out['fp64']:.10f} (dE {abs(out['fp64'] - E_ref):.1e})  "
          f"E_mixed {out['mixed']:.10f} (dE {abs(out['mixed'] - E_ref):.1e})", flush=True)
    assert abs(out["fp64"] - E_ref) <= 1e-8
    assert abs(out["mixed"] - E_ref) <= 1e-6


def test_gly10_full_size_mixed_vs_fp64_properties():
    """(Gly)10 6-31G(d,p), 73 atoms: mixed against FP64 on the GPU (max|dF|,
    |dE2| <= 1e-6 Eh), both symmetric, FP32 work present, and both linear
    in the density (G(P1) + G(P2) = G(P1 + P2) with screening off)."""
    announce("config3 (Gly)10 6-31G** full size, mixed vs FP64")
    mol = molecule("gly10")
    bs = product_basis(mol, BASIS)
    e64 = _engine(bs, cutoff=1e-11)
    S, _, _ = e64.one_electron()
    from quick_amd.scf import seeded_density
    P = seeded_density(S, mol.nelec() // 2)
    G64 = e64.fock2e(P)
    emx = _engine(bs, cutoff=1e-11, precision="mixed", split=1e-4)
    Gmx = emx.fock2e(P)
    st = emx.stats()
    dF = float(np.max(np.abs(Gmx - G64)))
    dE = 0.5 * abs(float(np.sum(P * (Gmx - G64))))
    print(f"  nbasis {bs.nbasis}: FP32 tasks {st['n_tasks_fp32']} of {st['n_tasks_fp32'] + st['n_tasks_fp64']}; "
          f"max|dF| {dF:.2e}  |dE2| {dE:.2e} Eh", flush=True)
    assert bs.nbasis == 775
    assert st["n_tasks_fp32"] > 0
    assert dF <= 1e-6 and dE <= 1e-6
    scale = float(np.max(np.abs(G64)))
    assert np.max(np.abs(G64 - G64.T)) <= 1e-14 * scale
    assert np.max(np.abs(Gmx - Gmx.T)) <= 1e-14 * scale
    lin = _engine(bs, cutoff=1e-30, prim_limit=1e-31, precision="mixed")
    P2 = seeded_density(S, mol.nelec() // 2, seed=11)
    Ga, Gb, Gab = lin.fock2e(P), lin.fock2e(P2), lin.fock2e(P + P2)
    assert np.max(np.abs(Ga + Gb - Gab)) <= 1e-12 * float(np.max(np.abs(Gab)))
