"""Batched builds: many molecules, one launch per quartet class, optional
HIP-graph replay, per-class autotuning.  Each molecule's G must equal its
own single-engine build (and the oracle) to the FP64 bar."""
import numpy as np
import pytest

from tests.helpers import announce, molecule, oracle_for, product_basis, seeded_density_for

pytestmark = pytest.mark.gpu

FOCK_TOL = 1e-10
MIX = [("water", "6-31G*"), ("water3", "6-31G*"), ("alkane2", "6-31G*"), ("gly1", "6-31G*"), ("water", "STO-3G"),
       ("NH4+", "def2-SVP")]


def _setup(cases, **kw):
    from quick_amd.engine import FockEngine
    engs, dens, orcs = [], [], []
    for name, basis in cases:
        mol = molecule(name)
        orc = oracle_for(mol, basis)
        P, _ = seeded_density_for(orc, mol)
        engs.append(FockEngine(product_basis(mol, basis), cutoff=1e-12, **kw))
        dens.append(P)
        orcs.append(orc)
    return engs, dens, orcs


def test_batch_matches_single_engine_and_oracle():
    announce("batch of 6 molecules vs single builds")
    from quick_amd.engine import FockBatch
    engs, dens, orcs = _setup(MIX)
    singles = [e.fock2e(P) for e, P in zip(engs, dens)]
    b = FockBatch(engs, replicas=2)
    for i, P in enumerate(dens):
        b.set_density(i, P)
    b.build()
    for i, (G1, orc, P) in enumerate(zip(singles, orcs, dens)):
        G, F = b.get(i, g=True, f=True)
        assert np.max(np.abs(G - G1)) <= 1e-12, MIX[i]
        assert np.max(np.abs(F - G)) == 0.0
        if i < 3:
            G_ref, _ = orc.fock2e(P, cutoff=1e-12, threads=4)
            err = float(np.max(np.abs(G - G_ref)))
            print(f"  {MIX[i]} max|dF| vs oracle {err:.2e}", flush=True)
            assert err <= FOCK_TOL


def test_batch_graph_replays_equal_eager():
    """Every replay of the captured build equals the eager build; host work
    and synchronisation happen between replays."""
    announce("batch graph replay x4")
    import torch
    from quick_amd.engine import FockBatch
    engs, dens, _ = _setup(MIX[:4])
    ref = FockBatch(engs, replicas=2)
    for i, P in enumerate(dens):
        ref.set_density(i, P)
    ref.build()
    G_eager = [ref.get(i)[0] for i in range(len(engs))]
    ref.close()
    b = FockBatch(engs, replicas=2, graph=True)
    for i, P in enumerate(dens):
        b.set_density(i, P)
    s = torch.cuda.Stream()
    for rep in range(4):
        b.clear_fock(s.cuda_stream)
        b.build(s.cuda_stream)
        s.synchronize()
        for i in range(len(engs)):
            G, F = b.get(i, g=True, f=True)
            assert np.max(np.abs(G - G_eager[i])) <= 1e-12, (rep, i)
            assert np.max(np.abs(F - G)) == 0.0, (rep, i)
        _ = np.linalg.norm(dens[0])     # host work between replays


def test_one_shot_batch_api_and_autotune_keep_results():
    announce("qf_fock_build_batch + autotune")
    import ctypes

    import torch
    from quick_amd import _lib
    from quick_amd.autotune import autotune
    engs, dens, _ = _setup(MIX[:3])
    singles = [e.fock2e(P) for e, P in zip(engs, dens)]
    dev = torch.device("cuda:0")
    Pd = [torch.from_numpy(P).to(dev) for P in dens]
    Fd = [torch.zeros_like(p) for p in Pd]
    lib = _lib.lib()
    arrE = (ctypes.c_void_p * 3)(*[e.h for e in engs])
    arrP = (ctypes.c_void_p * 3)(*[p.data_ptr() for p in Pd])
    arrF = (ctypes.c_void_p * 3)(*[f.data_ptr() for f in Fd])
    _lib.check(lib.qf_fock_build_batch(arrE, 3, arrP, arrF, ctypes.c_void_p(0)), "qf_fock_build_batch")
    torch.cuda.synchronize()
    for i in range(3):
        assert np.max(np.abs(Fd[i].cpu().numpy() - singles[i])) <= 1e-12
    # autotune over the batch, then the single builds with the tuned shapes
    s = torch.cuda.current_stream().cuda_stream
    chosen = autotune(engs, s, builds=1, densities=dens)
    assert chosen
    for e, P, G1 in zip(engs, dens, singles):
        assert np.max(np.abs(e.fock2e(P) - G1)) <= 1e-12
