"""qf_fock_build_device_noalloc: the build with its scratch in the library's
fixed device array gives the same F as qf_fock_build_device, on two calls in
a row on one GPU (the second with a smaller molecule, so stale contents of
the larger call would show), and matches the oracle."""
import ctypes

import numpy as np
import pytest

from tests.helpers import announce, molecule, oracle_for, product_basis, seeded_density_for

pytestmark = pytest.mark.gpu


def _case(name, basis):
    import torch
    from quick_amd.engine import FockEngine
    mol = molecule(name)
    orc = oracle_for(mol, basis)
    P, _ = seeded_density_for(orc, mol)
    eng = FockEngine(product_basis(mol, basis), device=0, cutoff=1e-11)
    eng.prepare()
    Pd = torch.from_numpy(P).to("cuda:0")
    return mol, orc, P, eng, Pd


def _build(eng, Pd, noalloc):
    import torch
    F = torch.zeros_like(Pd)
    fn = eng.fock2e_device_noalloc if noalloc else eng.fock2e_device
    fn(Pd.data_ptr(), F.data_ptr(), torch.cuda.current_stream().cuda_stream)
    torch.cuda.synchronize()
    return F.cpu().numpy()


def test_noalloc_build_equals_device_build_two_calls():
    announce("no-alloc build: water3 6-31G* then water 6-31G*")
    cases = [_case("water3", "6-31G*"), _case("water", "6-31G*")]
    ref = [_build(c[3], c[4], False) for c in cases]
    for (mol, orc, P, eng, Pd), F_dev in zip(cases, ref):
        F_na = _build(eng, Pd, True)
        scale = max(1.0, float(np.max(np.abs(F_dev))))
        d_dev = float(np.max(np.abs(F_na - F_dev)))
        G_ref, _ = orc.fock2e(P, cutoff=1e-11, threads=8)
        d_orc = float(np.max(np.abs(F_na - G_ref)))
        print(f"  nbasis {eng.n}: max|F_noalloc - F_device| {d_dev:.2e}, vs oracle {d_orc:.2e}", flush=True)
        assert d_dev <= 1e-12 * scale
        assert d_orc <= 1e-10
    # the second call's result does not depend on what the first left in the array
    F_again = _build(cases[0][3], cases[0][4], True)
    assert float(np.max(np.abs(F_again - ref[0]))) <= 1e-12 * max(1.0, float(np.max(np.abs(ref[0]))))


def test_noalloc_build_refuses_what_does_not_fit():
    """A basis whose 2 nbasis^2 + nshell^2 doubles exceed the fixed array is
    refused with QF_ERR_ARG before anything is launched."""
    announce("no-alloc build: size limit")
    from quick_amd import _lib as L
    from quick_amd.engine import FockEngine
    cap = 32 << 20
    mol = molecule("water80")
    bs = product_basis(mol, "6-31G*")
    ns = len(bs.shells)
    need = 8 * (2 * bs.nbasis ** 2 + ns ** 2)
    assert need > cap, need
    eng = FockEngine(bs, device=0, cutoff=1e-11)
    eng.prepare()
    import torch
    Pd = torch.zeros((bs.nbasis, bs.nbasis), dtype=torch.float64, device="cuda:0")
    with pytest.raises(L.QuickFockError, match="fixed scratch"):
        eng.fock2e_device_noalloc(Pd.data_ptr(), 0, 0)
