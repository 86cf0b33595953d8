"""f shells (QUICK's spdf kernels, gpu_get2e_subs.h:579-686): every class
with an f shell runs on the wave-per-quartet kernel (node programs generated
for La, Lb, Lc, Ld <= 3; the largest keep their slots in global memory).
Parity with the oracle, which is pinned for f by the McMurchie-Davidson
reference and by QUICK's saved def2-TZVP energy."""
import numpy as np
import pytest

from tests.helpers import announce, md_schwarz, molecule, oracle_for, product_basis, seeded_density_for

pytestmark = pytest.mark.gpu

FOCK_TOL = 1e-10


def _engine(bs, **kw):
    from quick_amd.engine import FockEngine
    return FockEngine(bs, **kw)


CASES = [("water", "cc-pVTZ"), ("water", "def2-TZVP"), ("LiF", "cc-pVTZ"), ("alkane2", "def2-TZVP"),
         ("gly1", "cc-pVTZ")]


@pytest.mark.parametrize("name,basis", CASES)
def test_f_shell_schwarz_and_fock_match_oracle(name, basis):
    announce(f"f shells: {name} {basis}")
    mol = molecule(name)
    orc = oracle_for(mol, basis)
    bs = product_basis(mol, basis)
    eng = _engine(bs, cutoff=1e-11)
    y_gpu, c_gpu = eng.schwarz()
    y_ref, c_ref = orc.schwarz()
    # The oracle's HRR is QUICK's binomial expansion (hrrwhole, (a|b) as a
    # sum over [e| with AB powers): for an s/p/d shell and an f shell on
    # different atoms it cancels and loses up to ~4e-10 relative, where the
    # GPU's stepwise HRR keeps ~1e-15.  Pairs outside 1e-11 are therefore
    # settled against the McMurchie-Davidson values (same FmT).
    assert np.allclose(y_gpu, y_ref, rtol=1e-9, atol=1e-14)
    far = np.argwhere(~np.isclose(y_gpu, y_ref, rtol=1e-11, atol=1e-14))
    for a, b in far:
        if a <= b:
            y_md = md_schwarz(orc.qb, a, b)
            assert abs(y_gpu[a, b] - y_md) <= 1e-13 * y_md, (a, b, y_gpu[a, b], y_ref[a, b], y_md)
    mask = c_ref > 0
    assert np.allclose(c_gpu[mask], c_ref[mask], rtol=1e-9, atol=1e-14)
    P, _ = seeded_density_for(orc, mol)
    G_ref, nq = orc.fock2e(P, cutoff=1e-11, threads=8)
    G = eng.fock2e(P)
    err = float(np.max(np.abs(G - G_ref)))
    names = sorted({nm for nm, _, _ in eng.launches()})
    print(f"  nbasis {bs.nbasis}, {nq} quartets, classes {names}: max|dF| = {err:.3e}", flush=True)
    assert any("f" in nm for nm in names)
    assert err <= FOCK_TOL
    assert np.max(np.abs(G - G.T)) <= 1e-14 * max(1.0, float(np.max(np.abs(G))))


def test_f_shell_one_electron_matches_oracle():
    announce("f shells: one-electron S T V, water cc-pVTZ and glycine def2-TZVP")
    for name, basis in [("water", "cc-pVTZ"), ("gly1", "def2-TZVP")]:
        mol = molecule(name)
        orc = oracle_for(mol, basis)
        eng = _engine(product_basis(mol, basis))
        for a, b in zip(eng.one_electron(), orc.one_electron()):
            assert np.max(np.abs(a - b)) < 1e-10


def test_f_shell_scf_matches_oracle_scf():
    """Water cc-pVTZ RHF with GPU Fock builds against the oracle's SCF."""
    announce("f shells: SCF water cc-pVTZ GPU vs oracle")
    from oracle import oracle as O
    from quick_amd.scf import rhf
    mol = molecule("water")
    orc = oracle_for(mol, "cc-pVTZ")
    E_ref, _, _ = O.rhf(orc, mol.nelec(), cutoff=1e-11, prim_limit=1e-12, etol=1e-11, dtol=1e-9, threads=8)
    eng = _engine(product_basis(mol, "cc-pVTZ"), cutoff=1e-11, prim_limit=1e-12)
    res = rhf(eng, etol=1e-11, dtol=1e-9)
    print(f"  E_gpu {res.energy:.12f}  E_oracle {E_ref:.12f}  dE {abs(res.energy - E_ref):.1e}", flush=True)
    assert res.converged
    assert abs(res.energy - E_ref) <= 1e-8


def test_f_shell_paths_agree():
    """The f classes on every launch setting: persistent grid caps and
    primitive pieces give the same Fock matrix."""
    announce("f shells: launch settings, glycine cc-pVTZ")
    mol = molecule("gly1")
    orc = oracle_for(mol, "cc-pVTZ")
    P, _ = seeded_density_for(orc, mol)
    bs = product_basis(mol, "cc-pVTZ")
    base = _engine(bs, cutoff=1e-11).fock2e(P)
    e = _engine(bs, cutoff=1e-11)
    e.set_persistent(64)
    e.set_prim_split(1)
    G = e.fock2e(P)
    assert np.max(np.abs(G - base)) <= 1e-12 * max(1.0, float(np.max(np.abs(base))))
