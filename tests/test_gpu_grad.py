"""Two-electron nuclear gradient on the GPU (QUICK's gpu_get_cshell_eri_grad_
-> getGrad_kernel_sp/_spd, gpu_get2e_subs_grad.h) against the oracle's
restatement of QUICK's CPU cshell_eri_grad (eri_grad.fh:43-768), which the
CPU suite pins to finite differences of the two-electron energy."""
import ctypes

import numpy as np
import pytest

from tests.helpers import announce, molecule, oracle_for, product_basis, quick_abi_setup, seeded_density_for

pytestmark = pytest.mark.gpu

GRAD_TOL = 1e-10   # Eh/bohr, relative to max(1, max |g|)


def _err(g, g_ref):
    return float(np.max(np.abs(g - g_ref))) / max(1.0, float(np.max(np.abs(g_ref))))


def _engine(bs, **kw):
    from quick_amd.engine import FockEngine
    return FockEngine(bs, **kw)


CASES = [("water", "6-31G*", 1e-12), ("water3", "6-31G*", 1e-12), ("c2h5", "6-31G**", 1e-11),
         ("gly1", "6-31G*", 1e-11), ("alkane3", "cc-pVDZ", 1e-11), ("NH4+", "def2-SVP", 1e-12),
         ("LiF", "6-31G*", 1e-12)]


@pytest.mark.parametrize("lane", [True, False], ids=["lane", "wave"])
@pytest.mark.parametrize("name,basis,cutoff", CASES)
def test_eri_gradient_matches_oracle(name, basis, cutoff, lane):
    """Both kernels: per-lane for the light classes (default) and the
    wave-per-quartet kernel for every class."""
    announce(f"ERI gradient: {name} {basis} cutoff {cutoff} {'lane' if lane else 'wave'}")
    mol = molecule(name)
    orc = oracle_for(mol, basis)
    P, _ = seeded_density_for(orc, mol)
    g_ref, nq = orc.eri_grad(P, grad_cutoff=cutoff, prim_limit=0.1 * cutoff, threads=8)
    eng = _engine(product_basis(mol, basis), cutoff=cutoff)
    eng.set_grad_kernel(lane)
    g = eng.eri_grad(P)
    err = _err(g, g_ref)
    print(f"  {nq} shell quartets, max|g| {np.max(np.abs(g_ref)):.3f}, max|dg|/max(1,|g|) = {err:.2e}", flush=True)
    assert err <= GRAD_TOL
    assert np.max(np.abs(g.sum(axis=0))) <= 1e-11 * max(1.0, float(np.max(np.abs(g))))


def test_eri_gradient_per_class_kernels_and_timing():
    """qf_set_grad_class: a mix of per-lane and wave-per-quartet classes gives
    the oracle's gradient, and a timed gradient call records every class
    launch (the event pairs qf_read_timing reads)."""
    from quick_amd.autotune import _class_index
    announce("ERI gradient: gly1 6-31G* per-class kernel mix, timed")
    mol = molecule("gly1")
    orc = oracle_for(mol, "6-31G*")
    P, _ = seeded_density_for(orc, mol)
    g_ref, _ = orc.eri_grad(P, grad_cutoff=1e-11, prim_limit=1e-12, threads=8)
    eng = _engine(product_basis(mol, "6-31G*"), cutoff=1e-11)
    names = sorted({nm for nm, _, _ in eng.launches()})
    for k, nm in enumerate(names):   # alternate the kernels over the classes
        eng.set_grad_class(*_class_index(nm), k % 2)
    eng.set_timing(True)
    g = eng.eri_grad(P)
    ms, calls = eng.read_timing()
    eng.set_timing(False)
    assert _err(g, g_ref) <= GRAD_TOL
    assert len(ms) == len(eng.launches()) and np.all(calls == 1) and np.all(ms >= 0.0) and ms.sum() > 0.0
    with pytest.raises(RuntimeError):
        eng.set_grad_class(0, 1, 0)   # ck > cb


def test_eri_gradient_quick_cutoffs_and_lists():
    """QUICK's production settings (integral cutoff 1e-9, gradCutoff 1e-7,
    primLimit 1e-10) and a gradCutoff below the integral cutoff, which
    extends the pair lists at the next prepare; the Fock build is unchanged."""
    announce("ERI gradient: water3 with QUICK's cutoffs and a looser integral cutoff")
    mol = molecule("water3")
    orc = oracle_for(mol, "6-31G*")
    P, _ = seeded_density_for(orc, mol)
    eng = _engine(product_basis(mol, "6-31G*"), cutoff=1e-9, prim_limit=1e-10)
    G0 = eng.fock2e(P)
    eng.set_grad_cutoff(1e-7)
    g_ref, _ = orc.eri_grad(P, grad_cutoff=1e-7, prim_limit=1e-10, threads=8)
    assert _err(eng.eri_grad(P), g_ref) <= GRAD_TOL
    eng.set_grad_cutoff(1e-12)
    g_ref, _ = orc.eri_grad(P, grad_cutoff=1e-12, prim_limit=1e-10, threads=8)
    assert _err(eng.eri_grad(P), g_ref) <= GRAD_TOL
    assert np.max(np.abs(eng.fock2e(P) - G0)) <= 1e-12


def test_eri_gradient_device_entry_and_mixed_engine():
    """qf_eri_grad_device adds into a device array on a stream; an engine in
    mixed precision (its Fock tasks split in two launches) gives the FP64
    gradient over the union of both halves."""
    announce("ERI gradient: device entry, mixed-precision engine, alkane3 6-31G*")
    import torch
    mol = molecule("alkane3")
    orc = oracle_for(mol, "6-31G*")
    P, _ = seeded_density_for(orc, mol)
    g_ref, _ = orc.eri_grad(P, grad_cutoff=1e-11, prim_limit=1e-12, threads=8)
    eng = _engine(product_basis(mol, "6-31G*"), cutoff=1e-11, precision="mixed")
    Pd = torch.from_numpy(P).to("cuda:0")
    gd = torch.full((mol.natom, 3), 0.5, dtype=torch.float64, device="cuda:0")
    s = torch.cuda.current_stream()
    eng.eri_grad_device(Pd.data_ptr(), gd.data_ptr(), s.cuda_stream)
    torch.cuda.synchronize()
    g = gd.cpu().numpy() - 0.5
    assert _err(g, g_ref) <= GRAD_TOL
    assert eng.stats()["n_tasks_fp32"] > 0
    f, q, p = eng.grad_launch_counts()
    assert q.sum() > 0 and p.sum() >= q.sum() and np.all(f >= 0)


def test_eri_gradient_refuses_f_shells():
    announce("ERI gradient: f shells refused")
    from quick_amd import _lib as L
    mol = molecule("water")
    orc = oracle_for(mol, "cc-pVTZ")
    eng = _engine(product_basis(mol, "cc-pVTZ"), cutoff=1e-10)
    with pytest.raises(L.QuickFockError, match="up to d"):
        eng.eri_grad(np.eye(orc.n))


def test_quick_fortran_abi_gradient():
    """gpu_upload_density_matrix_, gpu_upload_cutoff_, gpu_upload_grad_ and
    gpu_get_cshell_eri_grad_ as QUICK's gradient driver calls them
    (gradient.fh:575-584, 867): grad(3, natom) += the two-electron term."""
    announce("QUICK Fortran ABI gradient, water3 6-31G*")
    from quick_amd import _lib
    lib = _lib.lib()
    mol = molecule("water3")
    orc = oracle_for(mol, "6-31G*")
    P, _ = seeded_density_for(orc, mol)
    g_ref, _ = orc.eri_grad(P, grad_cutoff=1e-8, prim_limit=1e-10, threads=8)
    vp, D = quick_abi_setup(lib, mol, orc)
    ierr = ctypes.c_int(0)
    dense = np.asfortranarray(P)
    lib.gpu_upload_density_matrix_(dense.ctypes.data_as(ctypes.c_void_p))
    cm = np.asfortranarray(orc.density_cutoff(P))
    lib.gpu_upload_cutoff_(cm.ctypes.data_as(ctypes.c_void_p), D(1e-9), D(1e-10), D(1e-10), D(1e-12))
    lib.gpu_upload_grad_(D(1e-8))
    grad = np.full(3 * mol.natom, 0.125)
    lib.gpu_get_cshell_eri_grad_(grad.ctypes.data_as(ctypes.c_void_p))
    lib.gpu_delete_(ctypes.byref(ierr))
    assert _err((grad - 0.125).reshape(-1, 3), g_ref) <= GRAD_TOL


def test_eri_gradient_large_cluster_properties():
    """(H2O)16 6-31G(d) against the oracle (16 host threads), and the
    translational invariance of the full gradient."""
    announce("ERI gradient: (H2O)16 6-31G* vs oracle")
    mol = molecule("water16")
    orc = oracle_for(mol, "6-31G*")
    P, _ = seeded_density_for(orc, mol)
    g_ref, nq = orc.eri_grad(P, grad_cutoff=1e-10, prim_limit=1e-11, threads=16)
    eng = _engine(product_basis(mol, "6-31G*"), cutoff=1e-10)
    g = eng.eri_grad(P)
    err = _err(g, g_ref)
    print(f"  {nq} shell quartets: max|dg|/max(1,|g|) = {err:.2e}", flush=True)
    assert err <= GRAD_TOL
    assert np.max(np.abs(g.sum(axis=0))) <= 1e-10 * max(1.0, float(np.max(np.abs(g))))
